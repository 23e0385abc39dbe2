// Warps, elementwise operators and the fused motion-sweep kernels for gfx950.
//
// Volume kernels use 256-lane workgroups shaped 64 x 4 over (dim 1, dim 2) with one z-plane per
// grid row, so every wavefront reads and writes 64 contiguous floats.  Kernels that read a field
// at the xy-transposed index (src/cu_add.cu:58) stage 32 x 32 tiles through LDS so that both the
// transposed read and the natural write are coalesced.
#include "internal.hpp"

namespace mcsart {
namespace {

constexpr int EX = 64, EY = 4;

__device__ inline int clampi_d(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Clamp-to-edge trilinear sample of img at (i0 + m0, i1 + m1, i2 + m2), the value the
// reference's texture fetch returns at unnormalised coordinate (i + 1 + m - 0.5)
// (src/cu_deform.cu:86-89), with exact FP32 weights.  The integer part of each displacement is
// split off first, so the fraction keeps full precision for any index size.
__device__ inline float trilinear(const float *__restrict__ img, int n0, int n1, int n2, int i0,
                                  float m0, int i1, float m1, int i2, float m2) {
    // A displacement larger than the volume lands outside it on the same side, where clamp-to-edge
    // addressing returns the edge voxel whatever its size: clamping it to +-(n+1) changes no
    // result and keeps the float->int conversions in range (a NaN displacement becomes -(n+1)).
    m0 = fminf(fmaxf(m0, -(float)(n0 + 1)), (float)(n0 + 1));
    m1 = fminf(fmaxf(m1, -(float)(n1 + 1)), (float)(n1 + 1));
    m2 = fminf(fmaxf(m2, -(float)(n2 + 1)), (float)(n2 + 1));
    const float f0 = floorf(m0), f1 = floorf(m1), f2 = floorf(m2);
    const float t0 = m0 - f0, t1 = m1 - f1, t2 = m2 - f2;
    const int b0 = i0 + (int)f0, b1 = i1 + (int)f1, b2 = i2 + (int)f2;
    const int a0 = clampi_d(b0, 0, n0 - 1), c0 = clampi_d(b0 + 1, 0, n0 - 1);
    const int a1 = clampi_d(b1, 0, n1 - 1), c1 = clampi_d(b1 + 1, 0, n1 - 1);
    const int a2 = clampi_d(b2, 0, n2 - 1), c2 = clampi_d(b2 + 1, 0, n2 - 1);
    const long long s1 = n0, s2 = (long long)n0 * n1;
    const long long r00 = a1 * s1 + a2 * s2, r10 = c1 * s1 + a2 * s2;
    const long long r01 = a1 * s1 + c2 * s2, r11 = c1 * s1 + c2 * s2;
    const float u0 = 1.f - t0, u1 = 1.f - t1, u2 = 1.f - t2;
    const float c00 = img[r00 + a0] * u0 + img[r00 + c0] * t0;
    const float c10 = img[r10 + a0] * u0 + img[r10 + c0] * t0;
    const float c01 = img[r01 + a0] * u0 + img[r01 + c0] * t0;
    const float c11 = img[r11 + a0] * u0 + img[r11 + c0] * t0;
    const float e0 = c00 * u1 + c10 * t1;
    const float e1 = c01 * u1 + c11 * t1;
    return e0 * u2 + e1 * t2;
}

// out[id] = W(src)[id]: DVF m_c = alpha_c*vol + beta_c*flow formed on the fly
// (kernel_forwardDVF, src/cu_deform.cu:66-77), sample at (i0 + my, i1 + mx, i2 + mz) (:85-89).
template <bool ADD>
__global__ void __launch_bounds__(EX *EY) k_warp(float *__restrict__ out,
                                                const float *__restrict__ src, int nx, int ny,
                                                int nz, float vol, float flow, Fields f,
                                                float lambda, int nonneg) {
    const int i0 = blockIdx.x * EX + threadIdx.x;
    const int i1 = blockIdx.y * EY + threadIdx.y;
    const int i2 = blockIdx.z;
    if (i0 >= nx || i1 >= ny) return;
    const long long id = i0 + (long long)i1 * nx + (long long)i2 * nx * ny;
    const float mx = f.ax[id] * vol + f.bx[id] * flow;
    const float my = f.ay[id] * vol + f.by[id] * flow;
    const float mz = f.az[id] * vol + f.bz[id] * flow;
    const float v = trilinear(src, nx, ny, nz, i0, my, i1, mx, i2, mz);
    if (ADD) {   // x += lambda * W(u)  (src/AUMISART.cu:222-223), optional clamp
        float x = out[id] + v * lambda;
        if (nonneg) x = fmaxf(x, 0.f);
        out[id] = x;
    } else {
        out[id] = v;
    }
}

// kernel_deformation2 (src/cu_deform.cu:92-113) with its swapped boundary checks; the linear
// index is id = iy + ix*ny + iz*nx*ny, i.e. iy is the dim-1 index and ix the dim-2 index.
__global__ void __launch_bounds__(EX *EY) k_deform2(float *__restrict__ out,
                                                   const float *__restrict__ src, int nx, int ny,
                                                   int nz, float vol, float flow, Fields f) {
    const int iy = blockIdx.x * EX + threadIdx.x;   // dim 1
    const int ix = blockIdx.y * EY + threadIdx.y;   // dim 2
    const int iz = blockIdx.z;
    if (iy >= nx || ix >= ny) return;
    const long long N = (long long)nx * ny * nz;
    const long long id = iy + (long long)ix * nx + (long long)iz * nx * ny;
    const float c = src[id];
    const long long j1 = id + 1, j2 = id + nx, j3 = id + (long long)nx * ny;
    const float dx = (ix == nx - 1) ? 0.f : ((j1 < N) ? src[j1] : c) - c;   // :100-103
    const float dy = (iy == ny - 1) ? 0.f : ((j2 < N) ? src[j2] : c) - c;   // :104-107
    const float dz = (iz == nz - 1) ? 0.f : ((j3 < N) ? src[j3] : c) - c;   // :108-111
    const float mx = f.ax[id] * vol + f.bx[id] * flow;
    const float my = f.ay[id] * vol + f.by[id] * flow;
    const float mz = f.az[id] * vol + f.bz[id] * flow;
    out[id] = c + dy * mx + dx * my + dz * mz;   // :112
}

// Forward difference of src at voxel (i0, i1, i2) along ind, zero at the last index
// (src/cu_diff.cu:2-33).
__device__ inline float dfwd(const float *__restrict__ src, long long id, int i0, int i1, int i2,
                             int nx, int ny, int nz, int ind) {
    if (ind == 1) return (i0 == nx - 1) ? 0.f : src[id + 1] - src[id];
    if (ind == 2) return (i1 == ny - 1) ? 0.f : src[id + nx] - src[id];
    return (i2 == nz - 1) ? 0.f : src[id + (long long)nx * ny] - src[id];
}

#define VOXEL_PROLOGUE                                                             \
    const int i0 = blockIdx.x * EX + threadIdx.x;                                  \
    const int i1 = blockIdx.y * EY + threadIdx.y;                                  \
    const int i2 = blockIdx.z;                                                     \
    if (i0 >= nx || i1 >= ny) return;                                              \
    const long long id = i0 + (long long)i1 * nx + (long long)i2 * nx * ny;

__global__ void k_diff(float *__restrict__ out, const float *__restrict__ src, int nx, int ny,
                       int nz, int ind, int backward) {
    VOXEL_PROLOGUE
    float v;
    if (!backward) {
        v = dfwd(src, id, i0, i1, i2, nx, ny, nz, ind);
    } else {   // src/mex_diff.cu:86-109: backward difference, zero at both ends
        if (ind == 1) v = (i0 == nx - 1 || i0 == 0) ? 0.f : src[id] - src[id - 1];
        else if (ind == 2) v = (i1 == ny - 1 || i1 == 0) ? 0.f : src[id] - src[id - nx];
        else if (ind == 3)
            v = (i2 == nz - 1 || i2 == 0) ? 0.f : src[id] - src[id - (long long)nx * ny];
        else v = src[id];
    }
    out[id] = v;
}

// kernel_initial2 (src/cu_initial.cu:27-49)
__global__ void k_initial2(float *__restrict__ out, const float *__restrict__ src, int nx, int ny,
                           int nz, float vol, float flow) {
    VOXEL_PROLOGUE
    const float s = dfwd(src, id, i0, i1, i2, nx, ny, nz, 1) +
                    dfwd(src, id, i0, i1, i2, nx, ny, nz, 2) +
                    dfwd(src, id, i0, i1, i2, nx, ny, nz, 3);
    out[id] = s * (vol + flow);
}

// kernel_add2 (src/cu_add.cu:28-58): a += b^T * w * D_ind(c)
__global__ void k_add2(float *__restrict__ a, const float *__restrict__ b, int nx, int ny, int nz,
                       const float *__restrict__ c, float w, int ind) {
    VOXEL_PROLOGUE
    const float df = (ind >= 1 && ind <= 3) ? dfwd(c, id, i0, i1, i2, nx, ny, nz, ind) : 0.f;
    a[id] += b[i1 + (long long)i0 * ny + (long long)i2 * nx * ny] * w * df;
}

// kernel_update_udvf (src/cu_update_udvf.cu:11-62), x/y write-back swapped as written.
__global__ void k_update_udvf(float *ax, float *ay, float *az, float *bx, float *by, float *bz,
                              const float *__restrict__ img, const float *__restrict__ img0,
                              float vd, float fd, int nx, int ny, int nz) {
    VOXEL_PROLOGUE
    const float dfI = img0[id] - img[id];
    const float dfx = dfwd(img, id, i0, i1, i2, nx, ny, nz, 1);
    const float dfy = dfwd(img, id, i0, i1, i2, nx, ny, nz, 2);
    const float dfz = dfwd(img, id, i0, i1, i2, nx, ny, nz, 3);
    const float cax = dfx * vd, cay = dfy * vd, caz = dfz * vd;
    const float cbx = dfx * fd, cby = dfy * fd, cbz = dfz * fd;
    const float w = cax * cax + cay * cay + caz * caz + cbx * cbx + cby * cby + cbz * cbz;
    const float a_x = ay[id], a_y = ax[id], a_z = az[id];   // :45-47
    const float b_x = by[id], b_y = bx[id], b_z = bz[id];   // :48-50
    float b = dfI - cax * a_x - cay * a_y - caz * a_z - cbx * b_x - cby * b_y - cbz * b_z;
    if (w == 0.f) return;   // :52-53
    b /= w;
    const float r = 1.0f / 6;
    ax[id] += r * b * cax;
    ay[id] += r * b * cay;
    az[id] += r * b * caz;
    bx[id] += r * b * cbx;
    by[id] += r * b * cby;
    bz[id] += r * b * cbz;
}

__global__ void k_fill(float *__restrict__ a, long long n, float v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_add(float *__restrict__ a, const float *__restrict__ b, long long n, float w,
                      int nonneg) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x = a[i] + b[i] * w;   // src/cu_add.cu:17
    if (nonneg) x = fmaxf(x, 0.f);
    a[i] = x;
}

__global__ void k_division(float *a, const float *b, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = a[i] / b[i];   // src/cu_division.cu:20-24
    if (isnan(v) || isinf(v)) v = 0.f;
    a[i] = v;
}

// ---- fused motion sweep ----------------------------------------------------------------------
constexpr int TT = 32, TR = 8;   // 32x32 tile, 32x8 lanes, 4 rows per lane

// Load the xy-transposed 32x32 tile of f at plane i2 into t: afterwards t[j][i] holds
// f[dim1 = Y0 + i, dim2 = X0 + j], so lane (i0 = X0+tx, i1 = Y0+r) finds f^T at t[tx][r].
__device__ inline void load_transposed(float (*t)[TT + 1], const float *__restrict__ f, int nx,
                                       int ny, int X0, int Y0, int i2) {
    const int tx = threadIdx.x;
    for (int r = threadIdx.y; r < TT; r += TR) {
        const int d0 = Y0 + tx, d1 = X0 + r;
        t[r][tx] = (d0 < nx && d1 < ny) ? f[d0 + (long long)d1 * nx + (long long)i2 * nx * ny] : 0.f;
    }
}

// g = x - img0, img0 = sum over (alpha_y,D1),(alpha_x,D2),(alpha_z,D3) of alpha^T*vd*D x plus the
// same with beta and fd (src/AUMISART.cu:252-258); q = (D1+D2+D3)x * (-vd - fd) (:267).  The pair
// (g, q) is written interleaved in the natural layout (pair) and/or, through the LDS tile, in the
// xy-transposed layout (pairT), whichever the next projection reads.
__global__ void __launch_bounds__(TT *TR) k_motion_prep(float2 *__restrict__ pair,
                                                       float2 *__restrict__ pairT,
                                                       const float *__restrict__ x, Fields f,
                                                       int nx, int ny, int nz, float vd, float fd) {
    __shared__ float t[6][TT][TT + 1];
    const int X0 = blockIdx.x * TT, Y0 = blockIdx.y * TT, i2 = blockIdx.z;
    load_transposed(t[0], f.ax, nx, ny, X0, Y0, i2);
    load_transposed(t[1], f.ay, nx, ny, X0, Y0, i2);
    load_transposed(t[2], f.az, nx, ny, X0, Y0, i2);
    load_transposed(t[3], f.bx, nx, ny, X0, Y0, i2);
    load_transposed(t[4], f.by, nx, ny, X0, Y0, i2);
    load_transposed(t[5], f.bz, nx, ny, X0, Y0, i2);
    __syncthreads();
    const int tx = threadIdx.x;
    const int i0 = X0 + tx;
    const long long plane = (long long)nx * ny;
    float gv[TT / TR], qv[TT / TR];
#pragma unroll
    for (int k = 0; k < TT / TR; ++k) {
        const int r = threadIdx.y + k * TR;
        const int i1 = Y0 + r;
        gv[k] = qv[k] = 0.f;
        if (i0 >= nx || i1 >= ny) continue;
        const long long id = i0 + (long long)i1 * nx + (long long)i2 * plane;
        const float d1 = dfwd(x, id, i0, i1, i2, nx, ny, nz, 1);
        const float d2 = dfwd(x, id, i0, i1, i2, nx, ny, nz, 2);
        const float d3 = dfwd(x, id, i0, i1, i2, nx, ny, nz, 3);
        float img0 = 0.f;
        img0 += t[1][tx][r] * vd * d1;   // alpha_y, ind 1
        img0 += t[0][tx][r] * vd * d2;   // alpha_x, ind 2
        img0 += t[2][tx][r] * vd * d3;   // alpha_z, ind 3
        img0 += t[4][tx][r] * fd * d1;   // beta_y
        img0 += t[3][tx][r] * fd * d2;   // beta_x
        img0 += t[5][tx][r] * fd * d3;   // beta_z
        gv[k] = x[id] - img0;
        qv[k] = (d1 + d2 + d3) * (-vd + -fd);
        if (pair) pair[id] = make_float2(gv[k], qv[k]);
    }
    if (pairT) {   // uniform over the workgroup
        __syncthreads();   // every lane has finished reading the field tiles
#pragma unroll
        for (int k = 0; k < TT / TR; ++k) {
            const int r = threadIdx.y + k * TR;
            t[0][r][tx] = gv[k];
            t[1][r][tx] = qv[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < TT / TR; ++k) {
            const int r = threadIdx.y + k * TR;
            const int j0 = Y0 + tx, j1 = X0 + r;   // transposed element (iy = j0, ix = j1)
            if (j0 < ny && j1 < nx)
                pairT[j0 + (long long)j1 * ny + (long long)i2 * plane] = make_float2(t[0][tx][r], t[1][tx][r]);
        }
    }
}

// Store a 32x32 tile held in registers (lane (tx, r) owns voxel (X0 + tx, Y0 + r)) in the
// xy-transposed layout through LDS; lane (tx, r) then writes element (iy = Y0 + tx, ix = X0 + r).
__device__ inline void store_transposed(float (*t)[TT + 1], float *__restrict__ outT,
                                        const float *v, int nx, int ny, int X0, int Y0, int i2) {
    const int tx = threadIdx.x;
#pragma unroll
    for (int k = 0; k < TT / TR; ++k) t[threadIdx.y + k * TR][tx] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TT / TR; ++k) {
        const int r = threadIdx.y + k * TR;
        const int j0 = Y0 + tx, j1 = X0 + r;
        if (j0 < ny && j1 < nx)
            outT[j0 + (long long)j1 * ny + (long long)i2 * nx * ny] = t[tx][r];
    }
}

// Forward warp over 32x32 tiles: natural output (out) and/or transposed output (outT).
__global__ void __launch_bounds__(TT *TR) k_warp_tile(float *__restrict__ out,
                                                     float *__restrict__ outT,
                                                     const float *__restrict__ src, int nx, int ny,
                                                     int nz, float vol, float flow, Fields f) {
    __shared__ float t[TT][TT + 1];
    const int X0 = blockIdx.x * TT, Y0 = blockIdx.y * TT, i2 = blockIdx.z;
    const int i0 = X0 + threadIdx.x;
    float v[TT / TR];
#pragma unroll
    for (int k = 0; k < TT / TR; ++k) {
        const int i1 = Y0 + threadIdx.y + k * TR;
        v[k] = 0.f;
        if (i0 >= nx || i1 >= ny) continue;
        const long long id = i0 + (long long)i1 * nx + (long long)i2 * nx * ny;
        const float mx = f.ax[id] * vol + f.bx[id] * flow;   // src/cu_deform.cu:74-76
        const float my = f.ay[id] * vol + f.by[id] * flow;
        const float mz = f.az[id] * vol + f.bz[id] * flow;
        v[k] = trilinear(src, nx, ny, nz, i0, my, i1, mx, i2, mz);   // :85-89
        if (out) out[id] = v[k];
    }
    if (outT) store_transposed(t, outT, v, nx, ny, X0, Y0, i2);
}

// x += lambda * u (optional clamp at zero) over 32x32 tiles, with the transposed copy of the
// result when xT is given.
__global__ void __launch_bounds__(TT *TR) k_add_tile(float *__restrict__ x, float *__restrict__ xT,
                                                    const float *__restrict__ u, int nx, int ny,
                                                    int nz, float lambda, int nonneg) {
    __shared__ float t[TT][TT + 1];
    const int X0 = blockIdx.x * TT, Y0 = blockIdx.y * TT, i2 = blockIdx.z;
    const int i0 = X0 + threadIdx.x;
    float v[TT / TR];
#pragma unroll
    for (int k = 0; k < TT / TR; ++k) {
        const int i1 = Y0 + threadIdx.y + k * TR;
        v[k] = 0.f;
        if (i0 >= nx || i1 >= ny) continue;
        const long long id = i0 + (long long)i1 * nx + (long long)i2 * nx * ny;
        float y = x[id] + u[id] * lambda;   // src/cu_add.cu:17
        if (nonneg) y = fmaxf(y, 0.f);
        x[id] = y;
        v[k] = y;
    }
    if (xT) store_transposed(t, xT, v, nx, ny, X0, Y0, i2);
}

__global__ void __launch_bounds__(TT *TR) k_transpose(float *__restrict__ dst,
                                                     const float *__restrict__ src, int nx, int ny,
                                                     int nz) {
    __shared__ float t[TT][TT + 1];
    const int X0 = blockIdx.x * TT, Y0 = blockIdx.y * TT, i2 = blockIdx.z;
    const int i0 = X0 + threadIdx.x;
    float v[TT / TR];
#pragma unroll
    for (int k = 0; k < TT / TR; ++k) {
        const int i1 = Y0 + threadIdx.y + k * TR;
        v[k] = (i0 < nx && i1 < ny) ? src[i0 + (long long)i1 * nx + (long long)i2 * nx * ny] : 0.f;
    }
    store_transposed(t, dst, v, nx, ny, X0, Y0, i2);
}

// alpha_c += u^T * wv * D_c x, beta_c += u^T * wf * D_c x with the pairs (alpha_y, D1),
// (alpha_x, D2), (alpha_z, D3) (src/AUMISART.cu:272-277).
__global__ void __launch_bounds__(TT *TR) k_motion_update(float *ax, float *ay, float *az,
                                                         float *bx, float *by, float *bz,
                                                         const float *__restrict__ u,
                                                         const float *__restrict__ x, int nx,
                                                         int ny, int nz, float wv, float wf) {
    __shared__ float t[TT][TT + 1];
    const int X0 = blockIdx.x * TT, Y0 = blockIdx.y * TT, i2 = blockIdx.z;
    load_transposed(t, u, nx, ny, X0, Y0, i2);
    __syncthreads();
    const int tx = threadIdx.x;
    const int i0 = X0 + tx;
    if (i0 >= nx) return;
    for (int r = threadIdx.y; r < TT; r += TR) {
        const int i1 = Y0 + r;
        if (i1 >= ny) break;
        const long long id = i0 + (long long)i1 * nx + (long long)i2 * nx * ny;
        const float d1 = dfwd(x, id, i0, i1, i2, nx, ny, nz, 1);
        const float d2 = dfwd(x, id, i0, i1, i2, nx, ny, nz, 2);
        const float d3 = dfwd(x, id, i0, i1, i2, nx, ny, nz, 3);
        const float ut = t[tx][r];
        ay[id] += ut * wv * d1;
        ax[id] += ut * wv * d2;
        az[id] += ut * wv * d3;
        by[id] += ut * wf * d1;
        bx[id] += ut * wf * d2;
        bz[id] += ut * wf * d3;
    }
}

inline dim3 vgrid(int nx, int ny, int nz) {
    return dim3((nx + EX - 1) / EX, (ny + EY - 1) / EY, nz);
}
inline dim3 lin_grid(long long n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

inline dim3 tile_grid(int nx, int ny, int nz) {
    return dim3((nx + TT - 1) / TT, (ny + TT - 1) / TT, nz);
}

int launch_warp(float *out, float *outT, const float *src, int nx, int ny, int nz, float vol,
                float flow, const Fields &f, hipStream_t st) {
    hipLaunchKernelGGL(k_warp_tile, tile_grid(nx, ny, nz), dim3(TT, TR), 0, st, out, outT, src,
                       nx, ny, nz, vol, flow, f);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

int launch_transpose(float *dst, const float *src, int nx, int ny, int nz, hipStream_t st) {
    hipLaunchKernelGGL(k_transpose, tile_grid(nx, ny, nz), dim3(TT, TR), 0, st, dst, src, nx, ny,
                       nz);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

int launch_add_tile(float *x, float *xT, const float *u, int nx, int ny, int nz, float lambda,
                    int nonneg, hipStream_t st) {
    hipLaunchKernelGGL(k_add_tile, tile_grid(nx, ny, nz), dim3(TT, TR), 0, st, x, xT, u, nx, ny,
                       nz, lambda, nonneg);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

int launch_warp_add(float *x, const float *u, int nx, int ny, int nz, float vol, float flow,
                    const Fields &f, float lambda, int nonneg, hipStream_t st) {
    hipLaunchKernelGGL(k_warp<true>, vgrid(nx, ny, nz), dim3(EX, EY), 0, st, x, u, nx, ny, nz,
                       vol, flow, f, lambda, nonneg);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

int launch_deform2(float *out, const float *src, int nx, int ny, int nz, float vol, float flow,
                   const Fields &f, hipStream_t st) {
    hipLaunchKernelGGL(k_deform2, vgrid(nx, ny, nz), dim3(EX, EY), 0, st, out, src, nx, ny, nz,
                       vol, flow, f);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

int launch_add_clamp(float *x, const float *u, long long n, float lambda, int nonneg,
                     hipStream_t st) {
    hipLaunchKernelGGL(k_add, lin_grid(n), dim3(256), 0, st, x, u, n, lambda, nonneg);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

int launch_add(float *a, const float *b, long long n, float w, hipStream_t st) {
    return launch_add_clamp(a, b, n, w, 0, st);
}

int launch_add2(float *a, const float *b, int nx, int ny, int nz, const float *c, float w, int ind,
                hipStream_t st) {
    hipLaunchKernelGGL(k_add2, vgrid(nx, ny, nz), dim3(EX, EY), 0, st, a, b, nx, ny, nz, c, w,
                       ind);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

int launch_fill(float *a, long long n, float v, hipStream_t st) {
    hipLaunchKernelGGL(k_fill, lin_grid(n), dim3(256), 0, st, a, n, v);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

int launch_initial2(float *img, int nx, int ny, int nz, const float *img0, float vol, float flow,
                    hipStream_t st) {
    hipLaunchKernelGGL(k_initial2, vgrid(nx, ny, nz), dim3(EX, EY), 0, st, img, img0, nx, ny, nz,
                       vol, flow);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

int launch_division(float *a, const float *b, long long n, hipStream_t st) {
    hipLaunchKernelGGL(k_division, lin_grid(n), dim3(256), 0, st, a, b, n);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

int launch_diff(float *out, const float *img, int nx, int ny, int nz, int ind, int backward,
                hipStream_t st) {
    hipLaunchKernelGGL(k_diff, vgrid(nx, ny, nz), dim3(EX, EY), 0, st, out, img, nx, ny, nz, ind,
                       backward);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

int launch_update_udvf(float *ax, float *ay, float *az, float *bx, float *by, float *bz,
                       const float *img, const float *img0, float vd, float fd, int nx, int ny,
                       int nz, hipStream_t st) {
    hipLaunchKernelGGL(k_update_udvf, vgrid(nx, ny, nz), dim3(EX, EY), 0, st, ax, ay, az, bx, by,
                       bz, img, img0, vd, fd, nx, ny, nz);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

int launch_motion_prep(float2 *pair, float2 *pairT, const float *x, const float *const fields[6],
                       int nx, int ny, int nz, float vd, float fd, hipStream_t st) {
    const Fields f{fields[0], fields[1], fields[2], fields[3], fields[4], fields[5]};
    hipLaunchKernelGGL(k_motion_prep, tile_grid(nx, ny, nz), dim3(TT, TR), 0, st, pair, pairT, x,
                       f, nx, ny, nz, vd, fd);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

int launch_motion_update(float *const fields[6], const float *u, const float *x, int nx, int ny,
                         int nz, float wv, float wf, hipStream_t st) {
    const dim3 grid((nx + TT - 1) / TT, (ny + TT - 1) / TT, nz);
    hipLaunchKernelGGL(k_motion_update, grid, dim3(TT, TR), 0, st, fields[0], fields[1],
                       fields[2], fields[3], fields[4], fields[5], u, x, nx, ny, nz, wv, wf);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

}  // namespace mcsart

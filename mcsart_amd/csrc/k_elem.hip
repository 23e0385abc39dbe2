// Warps, elementwise operators and the fused motion-sweep kernels for gfx950.
//
// Simple volume kernels use 256-lane workgroups shaped 64 x 4 over (dim 1, dim 2) with one z-plane
// per grid row, so every wavefront reads and writes 64 contiguous floats.  The warps and the
// motion-sweep kernels work on 32 x 32 tiles (below); kernels that read a field at the
// xy-transposed index (src/cu_add.cu:58) stage the tile through LDS so that both the transposed
// read and the natural write are coalesced.
#include "internal.hpp"

namespace mcsart {
namespace {

constexpr int EX = 64, EY = 4;

__device__ inline int clampi_d(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

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

// kernel_multiple (src/cu_multiple.cu:11-43): img1 *= weight * D_ind(img0), forward difference,
// zero at the last index.
__global__ void k_multiple(float *__restrict__ a, const float *__restrict__ b, int nx, int ny,
                           int nz, float w, int ind) {
    VOXEL_PROLOGUE
    a[id] *= w * dfwd(b, id, i0, i1, i2, nx, ny, nz, ind);
}

// kernel_forwardDVF (src/cu_deform.cu:66-77, waste/kernel_forwardDVF.cu): m_c = alpha_c * vol +
// beta_c * flow.
__global__ void k_forward_dvf(float *__restrict__ mx, float *__restrict__ my,
                              float *__restrict__ mz, Fields f, int nx, int ny, int nz, float vol,
                              float flow) {
    VOXEL_PROLOGUE
    mx[id] = f.ax[id] * vol + f.bx[id] * flow;
    my[id] = f.ay[id] * vol + f.by[id] * flow;
    mz[id] = f.az[id] * vol + f.bz[id] * flow;
}

// ---- tile kernels: warps and the fused motion sweep ------------------------------------------
//
// A workgroup of 32 x 8 lanes owns a 32 x 32 tile of one z-plane; lane (tx, ty) owns the four
// voxels (X0 + tx, Y0 + ty + 8k), k = 0..3.  Each kernel first issues every load its lane needs
// (all four voxels), then computes, then stores: nothing it writes can alias what it still has to
// read, so the loads of all four voxels are in flight together.  Volume indices are 32-bit
// (volumes are limited to < 2^31 voxels at the C ABI).  Transposed reads and writes go through an
// LDS tile (+1 padding: conflict-free columns) so that global accesses stay coalesced.
constexpr int TT = 32, TR = 8, TK = TT / TR;

// Index arithmetic of the warps.  The fast form (FAST = true) uses 24-bit multiplies and unsigned
// 32-bit byte offsets from a volume's base address, so every load is one instruction on a uniform
// base; it needs n0 < 2^24, n1 * n2 < 2^24 and fewer than 2^30 voxels (fast_index below).  The
// general form uses 32-bit signed element indices (volumes of < 2^31 voxels).
inline bool fast_index(int n0, int n1, int n2) {
    return n0 >= 2 && n0 < (1 << 24) && (long long)n1 * n2 < (1 << 24) &&
           (long long)n0 * n1 * n2 < (1LL << 30);
}

__device__ inline float ldb(const float *__restrict__ base, unsigned off) {
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + off);
}
// Two consecutive floats at a 4-byte aligned byte offset, as one 8-byte load.
__device__ inline float2 ldb2(const float *__restrict__ base, unsigned off) {
    float2 v;
    __builtin_memcpy(&v, reinterpret_cast<const char *>(base) + off, sizeof(v));
    return v;
}
__device__ inline void stb(float *__restrict__ base, unsigned off, float v) {
    *reinterpret_cast<float *>(reinterpret_cast<char *>(base) + off) = v;
}

// Element index i0 + i1*n0 + i2*n0*n1 of a voxel: a byte offset in the fast form.
template <bool FAST>
__device__ inline unsigned voxel_off(int i0, int i1, int i2, int n0, int n1) {
    if constexpr (FAST)
        return (__umul24(__umul24((unsigned)i2, (unsigned)n1) + (unsigned)i1, (unsigned)n0) +
                (unsigned)i0) << 2;
    else
        return (unsigned)(i0 + i1 * n0 + i2 * n0 * n1);
}
template <bool FAST>
__device__ inline float vload(const float *__restrict__ p, unsigned o) {
    if constexpr (FAST)
        return ldb(p, o);
    else
        return p[(int)o];
}
template <bool FAST>
__device__ inline void vstore(float *__restrict__ p, unsigned o, float v) {
    if constexpr (FAST)
        stb(p, o, v);
    else
        p[(int)o] = v;
}

// Clamp-to-edge trilinear sample of img at (i0 + m0, i1 + m1, i2 + m2), the value the
// reference's texture fetch returns at unnormalised coordinate (i + 1 + m - 0.5)
// (src/cu_deform.cu:86-89), with exact FP32 weights.  The integer part of each displacement is
// split off first, so the fraction keeps full precision for any index size.
template <bool FAST>
__device__ inline float trilinear(const float *__restrict__ img, int n0, int n1, int n2, int i0,
                                  float m0, int i1, float m1, int i2, float m2) {
    // A displacement larger than the volume lands outside it on the same side, where clamp-to-edge
    // addressing returns the edge voxel whatever its size: clamping it to +-(n+1) changes no
    // result and keeps the float->int conversions in range.  v_med3_f32 returns the minimum of
    // its operands when one is NaN, so a NaN displacement (an arithmetic result, hence a quiet
    // NaN) becomes -(n+1).
    m0 = __builtin_amdgcn_fmed3f(m0, -(float)(n0 + 1), (float)(n0 + 1));
    m1 = __builtin_amdgcn_fmed3f(m1, -(float)(n1 + 1), (float)(n1 + 1));
    m2 = __builtin_amdgcn_fmed3f(m2, -(float)(n2 + 1), (float)(n2 + 1));
    const float f0 = floorf(m0), f1 = floorf(m1), f2 = floorf(m2);
    const float t0 = m0 - f0, t1 = m1 - f1, t2 = m2 - f2;
    const int b0 = i0 + (int)f0, b1 = i1 + (int)f1, b2 = i2 + (int)f2;
    const float u0 = 1.f - t0, u1 = 1.f - t1, u2 = 1.f - t2;
    float v[8];
    if constexpr (FAST) {
        // lower corner clamped into the volume; the upper corner is one step further along an
        // axis exactly when 0 <= b < n - 1, and the same voxel otherwise
        const int a0 = min(max(b0, 0), n0 - 1);
        const int a1 = min(max(b1, 0), n1 - 1);
        const int a2 = min(max(b2, 0), n2 - 1);
        const unsigned s0 = (unsigned)b0 < (unsigned)(n0 - 1) ? 4u : 0u;
        const unsigned s1 = (unsigned)b1 < (unsigned)(n1 - 1) ? 4u * (unsigned)n0 : 0u;
        const unsigned s2 =
            (unsigned)b2 < (unsigned)(n2 - 1) ? 4u * (unsigned)n0 * (unsigned)n1 : 0u;
        // Each row's pair (a0, c0) comes from one 8-byte load of elements (e, e + 1) of that
        // row: e = a0, except at the upper edge (c0 == a0 == n0 - 1), where e = a0 - 1 and both
        // values are the second element.  With n0 >= 2 both elements lie in the row.
        const bool hi = s0 == 0u && a0 > 0;
        const bool second = s0 != 0u || hi;   // c0 is element e + 1
        const unsigned o = voxel_off<true>(a0, a1, a2, n0, n1) - (hi ? 4u : 0u);
        const unsigned o1 = o + s1, o2 = o + s2, o3 = o2 + s1;
        const float2 p0 = ldb2(img, o), p1 = ldb2(img, o1), p2 = ldb2(img, o2),
                     p3 = ldb2(img, o3);
        v[0] = hi ? p0.y : p0.x;
        v[1] = second ? p0.y : p0.x;
        v[2] = hi ? p1.y : p1.x;
        v[3] = second ? p1.y : p1.x;
        v[4] = hi ? p2.y : p2.x;
        v[5] = second ? p2.y : p2.x;
        v[6] = hi ? p3.y : p3.x;
        v[7] = second ? p3.y : p3.x;
    } else {
        const int a0 = clampi_d(b0, 0, n0 - 1), c0 = clampi_d(b0 + 1, 0, n0 - 1);
        const int a1 = clampi_d(b1, 0, n1 - 1), c1 = clampi_d(b1 + 1, 0, n1 - 1);
        const int a2 = clampi_d(b2, 0, n2 - 1), c2 = clampi_d(b2 + 1, 0, n2 - 1);
        const int s2 = n0 * n1;
        const int r00 = a1 * n0 + a2 * s2, r10 = c1 * n0 + a2 * s2;
        const int r01 = a1 * n0 + c2 * s2, r11 = c1 * n0 + c2 * s2;
        v[0] = img[r00 + a0];
        v[1] = img[r00 + c0];
        v[2] = img[r10 + a0];
        v[3] = img[r10 + c0];
        v[4] = img[r01 + a0];
        v[5] = img[r01 + c0];
        v[6] = img[r11 + a0];
        v[7] = img[r11 + c0];
    }
    const float c00 = v[0] * u0 + v[1] * t0;
    const float c10 = v[2] * u0 + v[3] * t0;
    const float c01 = v[4] * u0 + v[5] * t0;
    const float c11 = v[6] * u0 + v[7] * t0;
    const float e0 = c00 * u1 + c10 * t1;
    const float e1 = c01 * u1 + c11 * t1;
    return e0 * u2 + e1 * t2;
}

// DVF of voxel o (kernel_forwardDVF, src/cu_deform.cu:74-76): (mx, my, mz).
struct Dvf {
    float x, y, z;
};
template <bool FAST>
__device__ inline Dvf dvf_at(const Fields &f, unsigned o, float vol, float flow) {
    return Dvf{vload<FAST>(f.ax, o) * vol + vload<FAST>(f.bx, o) * flow,
               vload<FAST>(f.ay, o) * vol + vload<FAST>(f.by, o) * flow,
               vload<FAST>(f.az, o) * vol + vload<FAST>(f.bz, o) * flow};
}

// Forward differences of x at voxel id, zero at the last index (src/cu_diff.cu:2-33).
struct Diffs {
    float x, d1, d2, d3;
};
__device__ inline Diffs diffs_at(const float *__restrict__ x, int id, int i0, int i1, int i2,
                                 int nx, int ny, int nz) {
    Diffs d;
    d.x = x[id];
    d.d1 = (i0 == nx - 1) ? 0.f : x[id + 1] - d.x;
    d.d2 = (i1 == ny - 1) ? 0.f : x[id + nx] - d.x;
    d.d3 = (i2 == nz - 1) ? 0.f : x[id + nx * ny] - d.x;
    return d;
}

// Load the tile of f whose natural coordinates are (dim0 = Y0 + i, dim1 = X0 + j) into t[j][i]
// (coalesced row reads); lane (tx, r) owning voxel (X0 + tx, Y0 + r) then finds
// f^T(voxel) = f(dim0 = Y0 + r, dim1 = X0 + tx) at t[tx][r].
__device__ inline void load_transposed(float (*t)[TT + 1], const float *__restrict__ f, int nx,
                                       int ny, int X0, int Y0, int i2) {
    const int tx = threadIdx.x;
#pragma unroll
    for (int k = 0; k < TK; ++k) {
        const int r = threadIdx.y + k * TR;
        const int d0 = Y0 + tx, d1 = X0 + r;
        t[r][tx] = (d0 < nx && d1 < ny) ? f[d0 + d1 * nx + i2 * nx * ny] : 0.f;
    }
}

// Store values v[k] of voxels (X0 + tx, Y0 + ty + 8k) in the xy-transposed layout
// outT[iy + ix*ny + iz*nx*ny] through LDS; every global store is a coalesced row.
template <class T, int ROWS = TR>
__device__ inline void store_transposed(T (*t)[TT + 1], T *__restrict__ outT, const T *v,
                                        int nx, int ny, int X0, int Y0, int i2) {
    constexpr int K = TT / ROWS;
    const int tx = threadIdx.x;
#pragma unroll
    for (int k = 0; k < K; ++k) t[threadIdx.y + k * ROWS][tx] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int r = threadIdx.y + k * ROWS;
        const int j0 = Y0 + tx, j1 = X0 + r;
        if (j0 < ny && j1 < nx) outT[j0 + j1 * ny + i2 * nx * ny] = t[tx][r];
    }
}

// The warps run 32 x ROWS lanes per 32 x 32 tile, i.e. 32 / ROWS voxels per lane (WARP_ROWS).
//
// Forward warp (src/cu_deform.cu:79-90): natural output (out) and/or transposed output (outT).
// Only the voxels the view's projector can read are warped (cc: its read set, ConeCull); the
// others are neither read nor written (outside the edge shell, written as NaN when cc.mode == 2),
// so the projector's nonzero weights meet exactly the values an uncut warp gives it.
template <bool FAST, int ROWS>
__global__ void __launch_bounds__(TT *ROWS) k_warp_tile(float *__restrict__ out,
                                                     float *__restrict__ outT,
                                                     const float *__restrict__ src, int nx, int ny,
                                                     int nz, float vol, float flow, Fields f,
                                                     ConeCull cc) {
    __shared__ float t[TT][TT + 1];
    const int X0 = spread_x(blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x) * TT;
    const int Y0 = blockIdx.y * TT, i2 = blockIdx.z;
    const int i0 = X0 + threadIdx.x;
    constexpr int K = TT / ROWS;
    Dvf m[K];
    unsigned id[K];
    bool ok[K], need[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i1 = Y0 + threadIdx.y + k * ROWS;
        ok[k] = i0 < nx && i1 < ny;
        need[k] = ok[k] && cone_keep(cc, i0, i1, i2);
        id[k] = voxel_off<FAST>(i0, i1, i2, nx, ny);
        m[k] = need[k] ? dvf_at<FAST>(f, id[k], vol, flow) : Dvf{0.f, 0.f, 0.f};
    }
    float v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i1 = Y0 + threadIdx.y + k * ROWS;
        // sample at (i0 + my, i1 + mx, i2 + mz) (:85-89)
        v[k] = need[k] ? trilinear<FAST>(src, nx, ny, nz, i0, m[k].y, i1, m[k].x, i2, m[k].z)
                       : __builtin_nanf("");
    }
    if (out) {
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (need[k] || (ok[k] && cone_poison(cc, i0, Y0 + threadIdx.y + k * ROWS, i2)))
                vstore<FAST>(out, id[k], v[k]);
    }
    if (outT) {   // through LDS, as store_transposed; lane (tx, r) writes voxel (X0 + r, Y0 + tx)
#pragma unroll
        for (int k = 0; k < K; ++k) t[threadIdx.y + k * ROWS][threadIdx.x] = v[k];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int r = threadIdx.y + k * ROWS;
            const int j0 = Y0 + threadIdx.x, j1 = X0 + r;
            if (j0 < ny && j1 < nx && (cone_keep(cc, j1, j0, i2) || cone_poison(cc, j1, j0, i2)))
                outT[j0 + j1 * ny + i2 * nx * ny] = t[threadIdx.x][r];
        }
    }
}

// The eight voxels a clamp-to-edge trilinear sample at (i0 + m0, i1 + m1, i2 + m2) reads (the
// same clamping as trilinear()) meet the support of u (sup, ConeCull): otherwise all eight are 0
// and so is the sample.
__device__ inline bool sample_meets(const ConeCull &sup, int n0, int n1, int n2, int i0, float m0,
                                    int i1, float m1, int i2, float m2) {
    if (sup.mode == 0) return true;
    m0 = __builtin_amdgcn_fmed3f(m0, -(float)(n0 + 1), (float)(n0 + 1));
    m1 = __builtin_amdgcn_fmed3f(m1, -(float)(n1 + 1), (float)(n1 + 1));
    m2 = __builtin_amdgcn_fmed3f(m2, -(float)(n2 + 1), (float)(n2 + 1));
    const int b0 = i0 + (int)floorf(m0), b1 = i1 + (int)floorf(m1), b2 = i2 + (int)floorf(m2);
    const int a0 = min(max(b0, 0), n0 - 1), c0 = min(max(b0 + 1, 0), n0 - 1);
    const int a1 = min(max(b1, 0), n1 - 1), c1 = min(max(b1 + 1, 0), n1 - 1);
    const int a2 = min(max(b2, 0), n2 - 1), c2 = min(max(b2 + 1, 0), n2 - 1);
    return cone_box_meets(sup, 0.5f * (float)(a0 + c0), 0.5f * (float)(a1 + c1),
                          0.5f * (float)(a2 + c2), 0.5f * (float)(c0 - a0),
                          0.5f * (float)(c1 - a1), 0.5f * (float)(c2 - a2));
}

// x += lambda * W(u; vol, flow) (src/AUMISART.cu:222-223), then the optional clamp at zero.
// Where the sample's voxels all lie outside the support of u (sup), W(u) is 0: x + 0 is x, so
// that voxel is neither read nor written.  With the clamp, that skip is exact once x has been
// clamped everywhere: the sweep's first view runs with sup.mode == 0.
template <bool FAST, int ROWS>
__global__ void __launch_bounds__(TT *ROWS) k_warp_add_tile(float *__restrict__ x,
                                                         const float *__restrict__ u, int nx,
                                                         int ny, int nz, float vol, float flow,
                                                         Fields f, float lambda, int nonneg,
                                                         ConeCull sup) {
    const int X0 = spread_x(blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x) * TT;
    const int Y0 = blockIdx.y * TT, i2 = blockIdx.z;
    const int i0 = X0 + threadIdx.x;
    constexpr int K = TT / ROWS;
    Dvf m[K];
    unsigned id[K];
    bool ok[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i1 = Y0 + threadIdx.y + k * ROWS;
        ok[k] = i0 < nx && i1 < ny;
        id[k] = voxel_off<FAST>(i0, i1, i2, nx, ny);
        m[k] = ok[k] ? dvf_at<FAST>(f, id[k], vol, flow) : Dvf{0.f, 0.f, 0.f};
    }
    float v[K];
    bool hit[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i1 = Y0 + threadIdx.y + k * ROWS;
        hit[k] = ok[k] && sample_meets(sup, nx, ny, nz, i0, m[k].y, i1, m[k].x, i2, m[k].z);
        v[k] = hit[k] ? trilinear<FAST>(u, nx, ny, nz, i0, m[k].y, i1, m[k].x, i2, m[k].z) : 0.f;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (!hit[k]) continue;
        float y = vload<FAST>(x, id[k]) + v[k] * lambda;
        if (nonneg) y = fmaxf(y, 0.f);
        vstore<FAST>(x, id[k], y);
    }
}

// One step of the motion (udvf) sweep, fused over a tile pair.  With A the workgroup's tile and
// A^T its transpose (nx == ny):
//   UPD : fields on A += u^T * (wv | wf) * D_c x      (view i,   src/AUMISART.cu:272-277)
//   PREP: on A^T, g = x - sum_c field_c^T (vd | fd) D_c x, q = (D1+D2+D3)x (-vd - fd)
//                                                       (view i+1, src/AUMISART.cu:252-258,267)
// PREP reads the fields at the transposed index, i.e. on A, exactly the values UPD has just
// written, so one workgroup needs nothing from another and the fields are read once per step.
// x is constant over a motion sweep (it projects the unwarped image, :246).  The pair (g, q)
// goes interleaved to `pair` (natural layout) and/or `pairT` (transposed layout).
template <bool UPD, bool PREP>
__global__ void __launch_bounds__(TT *TR) k_motion_step(float *__restrict__ ax,
                                                       float *__restrict__ ay,
                                                       float *__restrict__ az,
                                                       float *__restrict__ bx,
                                                       float *__restrict__ by,
                                                       float *__restrict__ bz,
                                                       const float *__restrict__ u,
                                                       const float *__restrict__ x,
                                                       float2 *__restrict__ pair,
                                                       float2 *__restrict__ pairT, int nx,
                                                       int ny, int nz, float wv, float wf,
                                                       float vd, float fd) {
    __shared__ float t[6][TT][TT + 1];
    const int X0 = blockIdx.x * TT, Y0 = blockIdx.y * TT, i2 = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int plane = nx * ny;
    float *const F[6] = {ax, ay, az, bx, by, bz};
    float nf[6][TK];   // fields on A after the update, voxel (X0 + tx, Y0 + ty + 8k)
    int id[TK];
    bool ok[TK];
    if (UPD) load_transposed(t[0], u, nx, ny, X0, Y0, i2);
#pragma unroll
    for (int k = 0; k < TK; ++k) {
        const int i1 = Y0 + ty + k * TR;
        ok[k] = X0 + tx < nx && i1 < ny;
        id[k] = X0 + tx + i1 * nx + i2 * plane;
#pragma unroll
        for (int c = 0; c < 6; ++c) nf[c][k] = ok[k] ? F[c][id[k]] : 0.f;
    }
    if (UPD) {
        Diffs d[TK];
#pragma unroll
        for (int k = 0; k < TK; ++k)
            d[k] = ok[k] ? diffs_at(x, id[k], X0 + tx, Y0 + ty + k * TR, i2, nx, ny, nz)
                         : Diffs{0.f, 0.f, 0.f, 0.f};
        __syncthreads();   // the u tile is complete
#pragma unroll
        for (int k = 0; k < TK; ++k) {
            const float ut = t[0][tx][ty + k * TR];
            nf[1][k] += ut * wv * d[k].d1;   // alpha_y with D1
            nf[0][k] += ut * wv * d[k].d2;   // alpha_x with D2
            nf[2][k] += ut * wv * d[k].d3;   // alpha_z with D3
            nf[4][k] += ut * wf * d[k].d1;   // beta_y
            nf[3][k] += ut * wf * d[k].d2;   // beta_x
            nf[5][k] += ut * wf * d[k].d3;   // beta_z
        }
#pragma unroll
        for (int k = 0; k < TK; ++k)
            if (ok[k])
#pragma unroll
                for (int c = 0; c < 6; ++c) F[c][id[k]] = nf[c][k];
    }
    if (!PREP) return;
    if (UPD) __syncthreads();   // every lane is done with the u tile
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int k = 0; k < TK; ++k) t[c][ty + k * TR][tx] = nf[c][k];
    __syncthreads();
    // lane (tx, r) now owns voxel p = (Y0 + tx, X0 + r) of A^T; field_c^T(p) is t[c][tx][r]
    float gv[TK], qv[TK];
    Diffs d[TK];
    int jd[TK];
    bool jok[TK];
#pragma unroll
    for (int k = 0; k < TK; ++k) {
        const int j0 = Y0 + tx, j1 = X0 + ty + k * TR;
        jok[k] = j0 < nx && j1 < ny;
        jd[k] = j0 + j1 * nx + i2 * plane;
        d[k] = jok[k] ? diffs_at(x, jd[k], j0, j1, i2, nx, ny, nz) : Diffs{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int k = 0; k < TK; ++k) {
        const int r = ty + k * TR;
        float img0 = 0.f;
        img0 += t[1][tx][r] * vd * d[k].d1;   // alpha_y, ind 1
        img0 += t[0][tx][r] * vd * d[k].d2;   // alpha_x, ind 2
        img0 += t[2][tx][r] * vd * d[k].d3;   // alpha_z, ind 3
        img0 += t[4][tx][r] * fd * d[k].d1;   // beta_y
        img0 += t[3][tx][r] * fd * d[k].d2;   // beta_x
        img0 += t[5][tx][r] * fd * d[k].d3;   // beta_z
        gv[k] = d[k].x - img0;
        qv[k] = (d[k].d1 + d[k].d2 + d[k].d3) * (-vd + -fd);
    }
    if (pair) {
#pragma unroll
        for (int k = 0; k < TK; ++k)
            if (jok[k]) pair[jd[k]] = make_float2(gv[k], qv[k]);
    }
    if (pairT) {
        // pairT at natural position P of A holds the pair of P^T in A^T; lane (tx, r) writes
        // P = (X0 + tx, Y0 + r), computed by lane (r, tx)
        __syncthreads();   // all field reads from t are done
#pragma unroll
        for (int k = 0; k < TK; ++k) {
            t[0][ty + k * TR][tx] = gv[k];
            t[1][ty + k * TR][tx] = qv[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < TK; ++k) {
            const int r = ty + k * TR;
            if (ok[k]) pairT[id[k]] = make_float2(t[0][tx][r], t[1][tx][r]);
        }
    }
}

// The motion sweep in accumulator form.  x is constant over a motion sweep, so every update of
// the sweep adds a multiple of the same D_c x to a field (src/AUMISART.cu:272-277):
//   alpha_c(P) = alpha_c0(P) + D_c x(P) * SV(P),  SV(P) = sum over updated views of wv_j u_j(P^T)
//   beta_c(P)  = beta_c0(P)  + D_c x(P) * SF(P),  SF(P) = sum of wf_j u_j(P^T)
// with the fields' pairing (alpha_y, D1), (alpha_x, D2), (alpha_z, D3).  The prep of a view
// (src/AUMISART.cu:252-258,267) then needs, at voxel p, only
//   img0(p) = vd Ga(p) + fd Gb(p) + H(p) (vd SV(p^T) + fd SF(p^T)),
//   Ga(p) = sum_c alpha_c0(p^T) D_c x(p),  Gb(p) = sum_c beta_c0(p^T) D_c x(p),
//   H(p) = sum_c D_c x(p^T) D_c x(p),  Q(p) = (D1 + D2 + D3) x(p),
// so a view step reads and writes two accumulators instead of six fields, and the view's
// q = Q (-vd - fd) is the sweep's Q times a scalar (the projector scales P(Q); q is never
// stored).  k_motion_begin makes (Ga, Gb, H) and Q once per sweep; each view's update and the
// next view's prep g run in the ratio backprojector's epilogue (k_backproject.hip), which keeps
// the accumulators transposed and interleaved (acc(p) = (SV(p^T), SF(p^T))); k_motion_finish
// writes the fields at the end.

// Per-sweep constants (Ga, Gb, H) at every voxel from the sweep's starting fields, and the pairs
// (g, Q) in both layouts: Q = (D1 + D2 + D3) x for the whole sweep, g the prep of its first
// view (both accumulators still zero).  The natural layout is written by the A^T lanes at jd[k],
// the transposed one through LDS at the A lanes' indices id[k].
__global__ void __launch_bounds__(TT *TR) k_motion_begin(Fields f, const float *__restrict__ x,
                                                        float *__restrict__ cst,
                                                        float2 *__restrict__ pair,
                                                        float2 *__restrict__ pairT, int nx,
                                                        int ny, int nz, float vd, float fd) {
    __shared__ float t[9][TT][TT + 1];
    const int X0 = blockIdx.x * TT, Y0 = blockIdx.y * TT, i2 = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int plane = nx * ny;
    const float *const F[6] = {f.ax, f.ay, f.az, f.bx, f.by, f.bz};
    int id[TK], jd[TK];
    bool ok[TK], jok[TK];
    // A side: starting fields and D x at P, staged so that the A^T lane of p finds them at p^T
#pragma unroll
    for (int k = 0; k < TK; ++k) {
        const int i1 = Y0 + ty + k * TR;
        ok[k] = X0 + tx < nx && i1 < ny;
        id[k] = X0 + tx + i1 * nx + i2 * plane;
        const Diffs d =
            ok[k] ? diffs_at(x, id[k], X0 + tx, i1, i2, nx, ny, nz) : Diffs{0.f, 0.f, 0.f, 0.f};
        const int r = ty + k * TR;
#pragma unroll
        for (int c = 0; c < 6; ++c) t[c][r][tx] = ok[k] ? F[c][id[k]] : 0.f;
        t[6][r][tx] = d.d1;
        t[7][r][tx] = d.d2;
        t[8][r][tx] = d.d3;
    }
    __syncthreads();
    float gv[TK], qv[TK];
#pragma unroll
    for (int k = 0; k < TK; ++k) {
        const int r = ty + k * TR;
        const int j0 = Y0 + tx, j1 = X0 + r;
        jok[k] = j0 < nx && j1 < ny;
        jd[k] = j0 + j1 * nx + i2 * plane;
        const Diffs d =
            jok[k] ? diffs_at(x, jd[k], j0, j1, i2, nx, ny, nz) : Diffs{0.f, 0.f, 0.f, 0.f};
        // fields index order: ax ay az bx by bz; pairing (y, D1), (x, D2), (z, D3)
        const float ga = t[1][tx][r] * d.d1 + t[0][tx][r] * d.d2 + t[2][tx][r] * d.d3;
        const float gb = t[4][tx][r] * d.d1 + t[3][tx][r] * d.d2 + t[5][tx][r] * d.d3;
        const float h = t[6][tx][r] * d.d1 + t[7][tx][r] * d.d2 + t[8][tx][r] * d.d3;
        if (jok[k]) {
            reinterpret_cast<float4 *>(cst)[jd[k]] = make_float4(ga, gb, h, d.x);
            pair[jd[k]] = make_float2(d.x - (vd * ga + fd * gb), d.d1 + d.d2 + d.d3);
        }
        gv[k] = d.x - (vd * ga + fd * gb);
        qv[k] = d.d1 + d.d2 + d.d3;
    }
    __syncthreads();   // every read of t[0], t[1] above is done
#pragma unroll
    for (int k = 0; k < TK; ++k) {
        t[0][ty + k * TR][tx] = gv[k];
        t[1][ty + k * TR][tx] = qv[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TK; ++k) {
        const int r = ty + k * TR;
        if (ok[k]) pairT[id[k]] = make_float2(t[0][tx][r], t[1][tx][r]);
    }
}

// End of a motion sweep in accumulator form: alpha_c = alpha_c0 + D_c x SV and
// beta_c = beta_c0 + D_c x SF in place, with (SV(P), SF(P)) = acc[P^T] (the ratio backprojector
// keeps the accumulators transposed and interleaved).
__global__ void __launch_bounds__(TT *TR) k_motion_finish(const float2 *__restrict__ acc,
                                                         const float *__restrict__ x, FieldsW f,
                                                         int nx, int ny, int nz) {
    __shared__ float t[2][TT][TT + 1];
    const int X0 = blockIdx.x * TT, Y0 = blockIdx.y * TT, i2 = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int plane = nx * ny;
    // the tile of acc whose natural coordinates are (dim0 = Y0 + i, dim1 = X0 + j) into
    // t[.][j][i], as load_transposed does for one volume
#pragma unroll
    for (int k = 0; k < TK; ++k) {
        const int r = ty + k * TR;
        const int d0 = Y0 + tx, d1 = X0 + r;
        const float2 v =
            (d0 < nx && d1 < ny) ? acc[d0 + d1 * nx + i2 * plane] : make_float2(0.f, 0.f);
        t[0][r][tx] = v.x;
        t[1][r][tx] = v.y;
    }
    float *const F[6] = {f.ax, f.ay, f.az, f.bx, f.by, f.bz};
    float a[TK][6];
    Diffs d[TK];
    int id[TK];
    bool ok[TK];
#pragma unroll
    for (int k = 0; k < TK; ++k) {
        const int i1 = Y0 + ty + k * TR;
        ok[k] = X0 + tx < nx && i1 < ny;
        id[k] = X0 + tx + i1 * nx + i2 * plane;
        d[k] = ok[k] ? diffs_at(x, id[k], X0 + tx, i1, i2, nx, ny, nz) : Diffs{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < 6; ++c) a[k][c] = ok[k] ? F[c][id[k]] : 0.f;
    }
    __syncthreads();   // both accumulator tiles are complete
#pragma unroll
    for (int k = 0; k < TK; ++k) {
        if (!ok[k]) continue;
        const float sv = t[0][tx][ty + k * TR], sf = t[1][tx][ty + k * TR];
        F[1][id[k]] = a[k][1] + d[k].d1 * sv;   // alpha_y with D1
        F[0][id[k]] = a[k][0] + d[k].d2 * sv;   // alpha_x with D2
        F[2][id[k]] = a[k][2] + d[k].d3 * sv;   // alpha_z with D3
        F[4][id[k]] = a[k][4] + d[k].d1 * sf;   // beta_y
        F[3][id[k]] = a[k][3] + d[k].d2 * sf;   // beta_x
        F[5][id[k]] = a[k][5] + d[k].d3 * sf;   // beta_z
    }
}

__global__ void __launch_bounds__(TT *TR) k_transpose(float *__restrict__ dst,
                                                     const float *__restrict__ src, int nx, int ny,
                                                     int nz) {
    __shared__ float t[TT][TT + 1];
    const int X0 = blockIdx.x * TT, Y0 = blockIdx.y * TT, i2 = blockIdx.z;
    const int i0 = X0 + threadIdx.x;
    float v[TK];
#pragma unroll
    for (int k = 0; k < TK; ++k) {
        const int i1 = Y0 + threadIdx.y + k * TR;
        v[k] = (i0 < nx && i1 < ny) ? src[i0 + i1 * nx + i2 * nx * ny] : 0.f;
    }
    store_transposed(t, dst, v, nx, ny, X0, Y0, i2);
}

// Fixed-point inverse of a DVF (waste/kernel_invertDVF.cu:1-18): starting from 0, niter
// Gauss-Seidel sweeps x <- -m_x(p + d), y <- -m_y(p + d), z <- -m_z(p + d), each component
// sampled (clamp-to-edge trilinear, index space, dim 0 fastest) at the current estimate
// d = (x, y, z) with the components updated so far.
template <bool FAST>
__global__ void k_invert_dvf(float *__restrict__ mx2, float *__restrict__ my2,
                             float *__restrict__ mz2, const float *__restrict__ mx,
                             const float *__restrict__ my, const float *__restrict__ mz, int nx,
                             int ny, int nz, int niter) {
    const int i0 = blockIdx.x * EX + threadIdx.x;
    const int i1 = blockIdx.y * EY + threadIdx.y;
    const int i2 = blockIdx.z;
    if (i0 >= nx || i1 >= ny) return;
    float x = 0.f, y = 0.f, z = 0.f;
    for (int it = 0; it < niter; ++it) {
        x = -trilinear<FAST>(mx, nx, ny, nz, i0, x, i1, y, i2, z);
        y = -trilinear<FAST>(my, nx, ny, nz, i0, x, i1, y, i2, z);
        z = -trilinear<FAST>(mz, nx, ny, nz, i0, x, i1, y, i2, z);
    }
    const unsigned id = voxel_off<FAST>(i0, i1, i2, nx, ny);
    vstore<FAST>(mx2, id, x);
    vstore<FAST>(my2, id, y);
    vstore<FAST>(mz2, id, z);
}

inline dim3 vgrid(int nx, int ny, int nz) {
    return dim3((nx + EX - 1) / EX, (ny + EY - 1) / EY, nz);
}
inline dim3 lin_grid(long long n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

inline dim3 tile_grid(int nx, int ny, int nz) {
    return dim3((nx + TT - 1) / TT, (ny + TT - 1) / TT, nz);
}

// Lanes per tile for the warps: 32 x 32, one voxel per lane (more wavefronts in flight to hide
// the field-load -> gather -> store chain).  Measured in one session on warps after a motion
// sweep: four voxels per lane 66 us, two 62 us, one 59.6 us at 256x256x128; two 28.4 us, one
// 26.1 us at 128^3.
constexpr int WARP_ROWS = 32;

int launch_warp(float *out, float *outT, const float *src, int nx, int ny, int nz, float vol,
                float flow, const Fields &f, hipStream_t st, const ConeCull *read_set) {
    const dim3 grid = tile_grid(nx, ny, nz);
    const ConeCull cc = read_set ? *read_set : ConeCull{};
#define MCSART_WARP(F, RW)                                                                      \
    hipLaunchKernelGGL((k_warp_tile<F, RW>), grid, dim3(TT, RW), 0, st, out, outT, src, nx, ny, \
                       nz, vol, flow, f, cc)
    if (fast_index(nx, ny, nz)) MCSART_WARP(true, WARP_ROWS);
    else MCSART_WARP(false, WARP_ROWS);
#undef MCSART_WARP
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

int launch_transpose(float *dst, const float *src, int nx, int ny, int nz, hipStream_t st) {
    hipLaunchKernelGGL(k_transpose, tile_grid(nx, ny, nz), dim3(TT, TR), 0, st, dst, src, nx, ny,
                       nz);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

int launch_warp_add(float *x, const float *u, int nx, int ny, int nz, float vol, float flow,
                    const Fields &f, float lambda, int nonneg, hipStream_t st,
                    const ConeCull *support) {
    const dim3 grid = tile_grid(nx, ny, nz);
    const ConeCull sup = support ? *support : ConeCull{};
#define MCSART_WARP_ADD(F, RW)                                                                  \
    hipLaunchKernelGGL((k_warp_add_tile<F, RW>), grid, dim3(TT, RW), 0, st, x, u, nx, ny, nz,    \
                       vol, flow, f, lambda, nonneg, sup)
    if (fast_index(nx, ny, nz)) MCSART_WARP_ADD(true, WARP_ROWS);
    else MCSART_WARP_ADD(false, WARP_ROWS);
#undef MCSART_WARP_ADD
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
    // the PREP half reads the fields only; the kernel's field pointers are writable for UPD
    float *const f[6] = {(float *)fields[0], (float *)fields[1], (float *)fields[2],
                         (float *)fields[3], (float *)fields[4], (float *)fields[5]};
    return launch_motion_step(f, nullptr, x, pair, pairT, nx, ny, nz, 0.f, 0.f, vd, fd, false,
                              true, st);
}

int launch_motion_update(float *const fields[6], const float *u, const float *x, int nx, int ny,
                         int nz, float wv, float wf, hipStream_t st) {
    return launch_motion_step(fields, u, x, nullptr, nullptr, nx, ny, nz, wv, wf, 0.f, 0.f, true,
                              false, st);
}

int launch_motion_begin(const Fields &f, const float *x, float *cst, float2 *pair,
                        float2 *pairT, int nx, int ny, int nz, float vd, float fd,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_motion_begin, tile_grid(nx, ny, nz), dim3(TT, TR), 0, st, f, x, cst,
                       pair, pairT, nx, ny, nz, vd, fd);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

int launch_motion_finish(const float2 *acc, const float *x, const FieldsW &f, int nx, int ny,
                         int nz, hipStream_t st) {
    hipLaunchKernelGGL(k_motion_finish, tile_grid(nx, ny, nz), dim3(TT, TR), 0, st, acc, x, f, nx,
                       ny, nz);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

int launch_motion_step(float *const fields[6], const float *u, const float *x, float2 *pair,
                       float2 *pairT, int nx, int ny, int nz, float wv, float wf, float vd,
                       float fd, bool update, bool prep, hipStream_t st) {
    const dim3 grid = tile_grid(nx, ny, nz), block(TT, TR);
#define MCSART_MOTION_STEP(U, P)                                                                 \
    hipLaunchKernelGGL((k_motion_step<U, P>), grid, block, 0, st, fields[0], fields[1],          \
                       fields[2], fields[3], fields[4], fields[5], u, x, pair, pairT, nx, ny, nz, \
                       wv, wf, vd, fd)
    if (update && prep) MCSART_MOTION_STEP(true, true);
    else if (update) MCSART_MOTION_STEP(true, false);
    else if (prep) MCSART_MOTION_STEP(false, true);
#undef MCSART_MOTION_STEP
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

int launch_multiple(float *a, const float *b, int nx, int ny, int nz, float w, int ind,
                    hipStream_t st) {
    hipLaunchKernelGGL(k_multiple, vgrid(nx, ny, nz), dim3(EX, EY), 0, st, a, b, nx, ny, nz, w,
                       ind);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

int launch_forward_dvf(float *mx, float *my, float *mz, const Fields &f, int nx, int ny, int nz,
                       float vol, float flow, hipStream_t st) {
    hipLaunchKernelGGL(k_forward_dvf, vgrid(nx, ny, nz), dim3(EX, EY), 0, st, mx, my, mz, f, nx,
                       ny, nz, vol, flow);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

int launch_invert_dvf(float *mx2, float *my2, float *mz2, const float *mx, const float *my,
                      const float *mz, int nx, int ny, int nz, int niter, hipStream_t st) {
    if (fast_index(nx, ny, nz))
        hipLaunchKernelGGL(k_invert_dvf<true>, vgrid(nx, ny, nz), dim3(EX, EY), 0, st, mx2, my2,
                           mz2, mx, my, mz, nx, ny, nz, niter);
    else
        hipLaunchKernelGGL(k_invert_dvf<false>, vgrid(nx, ny, nz), dim3(EX, EY), 0, st, mx2, my2,
                           mz2, mx, my, mz, nx, ny, nz, niter);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

}  // namespace mcsart

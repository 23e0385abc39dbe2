// CPU oracle for the McSART hot path -- TEST INFRASTRUCTURE ONLY.
//
// A restatement of the reference CUDA kernels and host driver of bill52547/McSART as plain C++,
// templated on the real type: the FP64 instance (suffix _f64) is the parity oracle, the FP32
// instance (suffix _f32) follows the reference's own single-precision arithmetic so that rounding
// effects can be separated from logic errors.  Parallelised with OpenMP over independent outputs
// (rays, voxels); every output element is computed by exactly one thread in a fixed order, so
// results do not depend on the thread count.
//
// Only the test-suite, the smoke check and the CPU-baseline leg of the benchmark may load this
// library.  The product (libmcsart_hip) never calls it.
//
// Reference quirks are reproduced as written (integer ny/2 in footprints, transposed reads in
// add2, swapped write-back in update_udvf, ...).  Deliberate deviations:
//   * cos/sin are evaluated exactly in T (the reference uses cosf/sinf in the projector and
//     __cosf/__sinf in the backprojector);
//   * trilinear interpolation uses exact weights (CUDA texture filtering uses 8-bit fixed-point
//     fractions);
//   * deform2 reads that fall past the end of the volume (reachable as written, undefined in the
//     reference) contribute a zero difference;
//   * warp displacements beyond +-(n+1) voxels are clamped to +-(n+1) (the same edge voxel under
//     clamp-to-edge addressing) and NaN displacements to -(n+1), so the float->int conversions of
//     the trilinear sample are always defined (the HIP warp does the same);
//   * the driver's host image h_outimg starts equal to h_img (uninitialised in the reference).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <omp.h>

namespace {

template <typename T> inline T tmin(T a, T b) { return (a < b) ? a : b; }  // universal.h:16
template <typename T> inline T tmax(T a, T b) { return (a > b) ? a : b; }  // universal.h:12
template <typename T> inline T tabs(T x) { return x > 0 ? x : -x; }        // universal.h:19

inline long long lin(int ix, int iy, int iz, int nx, int ny) {
    return (long long)ix + (long long)iy * nx + (long long)iz * nx * ny;
}

// ---------------------------------------------------------------------------------------------
// Forward projector: src/cu_projection.cu:15-145 (kernel_projection), one view.
// proj[ia + ib*na] is overwritten.  Footprint weights are normalised per plane (:83,:86).
// ---------------------------------------------------------------------------------------------
template <typename T>
void projection(T* proj, const T* img, T angle, T SO, T SD, T da, int na, T ai, T db, int nb,
                T bi, int nx, int ny, int nz) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int ib = 0; ib < nb; ++ib) {
        for (int ia = 0; ia < na; ++ia) {
            const long long id = ia + (long long)ib * na;
            T acc = 0;
            const T cphi = std::cos(angle), sphi = std::sin(angle);   // :23-24
            const T x1 = -SO * cphi, y1 = -SO * sphi, z1 = 0;         // :25-27
            const T x20 = SD - SO;                                     // :28
            const T y20 = ((T)ia + ai) * da;                           // :29
            const T x2 = x20 * cphi - y20 * sphi;                      // :30
            const T y2 = x20 * sphi + y20 * cphi;                      // :31
            const T z2 = ((T)ib + bi) * db;                            // :32
            const T x21 = x2 - x1, y21 = y2 - y1, z21 = z2 - z1;       // :34-36
            if (tabs(x21) > tabs(y21)) {                               // :39
                for (int ix = 0; ix < nx; ++ix) {                      // :44
                    const T xl = x21 - da / 2 * sphi, xr = x21 + da / 2 * sphi;   // :48-49
                    const T yl = y21 - da / 2 * cphi, yr = y21 + da / 2 * cphi;   // :50-51
                    const T xc = (T)ix + (T)0.5 - (T)nx / 2 - x1;                  // :52
                    // (ratio * xc + y1) + ny / 2, with integer ny/2 (:54-61)
                    const T hy = (T)(ny / 2);
                    const T cyll = (yl / xl * xc + y1) + hy, cylr = (yl / xr * xc + y1) + hy;
                    const T cyrl = (yr / xl * xc + y1) + hy, cyrr = (yr / xr * xc + y1) + hy;
                    const T yi1 = tmin(tmin(cyll, cylr), tmin(cyrl, cyrr));         // :63
                    const T yi2 = tmax(tmax(cyll, cylr), tmax(cyrl, cyrr));         // :64
                    const int Yi1 = (int)std::floor(yi1), Yi2 = (int)std::floor(yi2);
                    const T zl = z21 - db / 2, zr = z21 + db / 2;                  // :67-68
                    const T czl = zl / x21 * xc + z1 + (T)(nz / 2);                // :71-72
                    const T czr = zr / x21 * xc + z1 + (T)(nz / 2);                // :73-74
                    const T zi1 = tmin(czl, czr), zi2 = tmax(czl, czr);             // :76-77
                    const int Zi1 = (int)std::floor(zi1), Zi2 = (int)std::floor(zi2);
                    for (int iy = std::max(0, Yi1); iy <= std::min(ny - 1, Yi2); ++iy) {   // :81
                        T wy = tmin((T)iy + 1, yi2) - tmax((T)iy, yi1);
                        wy /= (yi2 - yi1);                                          // :83
                        for (int iz = std::max(0, Zi1); iz <= std::min(nz - 1, Zi2); ++iz) {
                            T wz = tmin((T)iz + 1, zi2) - tmax((T)iz, zi1);
                            wz /= (zi2 - zi1);                                      // :86
                            acc += img[lin(ix, iy, iz, nx, ny)] * wy * wz;          // :87
                        }
                    }
                }
            } else {
                for (int iy = 0; iy < ny; ++iy) {                      // :97
                    const T yl = y21 - da / 2 * cphi, yr = y21 + da / 2 * cphi;   // :101-102
                    const T xl = x21 - da / 2 * sphi, xr = x21 + da / 2 * sphi;   // :103-104
                    const T yc = (T)iy + (T)0.5 - (T)ny / 2 - y1;                  // :105
                    // (ratio * yc + x1) + nx / 2, with integer nx/2 (:107-114)
                    const T hx = (T)(nx / 2);
                    const T cxll = (xl / yl * yc + x1) + hx, cxlr = (xl / yr * yc + x1) + hx;
                    const T cxrl = (xr / yl * yc + x1) + hx, cxrr = (xr / yr * yc + x1) + hx;
                    const T xi1 = tmin(tmin(cxll, cxlr), tmin(cxrl, cxrr));         // :116
                    const T xi2 = tmax(tmax(cxll, cxlr), tmax(cxrl, cxrr));         // :117
                    const int Xi1 = (int)std::floor(xi1), Xi2 = (int)std::floor(xi2);
                    const T zl = z21 - db / 2, zr = z21 + db / 2;                  // :120-121
                    const T czl = zl / y21 * yc + z1 + (T)(nz / 2);                // :124-125
                    const T czr = zr / y21 * yc + z1 + (T)(nz / 2);                // :126-127
                    const T zi1 = tmin(czl, czr), zi2 = tmax(czl, czr);             // :129-130
                    const int Zi1 = (int)std::floor(zi1), Zi2 = (int)std::floor(zi2);
                    for (int ix = std::max(0, Xi1); ix <= std::min(nx - 1, Xi2); ++ix) {   // :134
                        T wx = tmin((T)ix + 1, xi2) - tmax((T)ix, xi1);
                        wx /= (xi2 - xi1);                                          // :136
                        for (int iz = std::max(0, Zi1); iz <= std::min(nz - 1, Zi2); ++iz) {
                            T wz = tmin((T)iz + 1, zi2) - tmax((T)iz, zi1);
                            wz /= (zi2 - zi1);                                      // :139
                            acc += img[lin(ix, iy, iz, nx, ny)] * wx * wz;          // :140
                        }
                    }
                }
            }
            proj[id] = acc;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Backprojector: src/cu_backprojection.cu:49-170 (kernel_backprojection), one view, overwrites
// img.  The texture fetch at texel centres (:167) is an exact read of proj[ia + ib*na].
// ---------------------------------------------------------------------------------------------
template <typename T>
void backprojection(T* img, const T* proj, T angle, T SO, T SD, T da, int na, T ai, T db, int nb,
                    T bi, int nx, int ny, int nz) {
    const T sphi = std::sin(angle), cphi = std::cos(angle);   // :62-63
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = 0; iz < nz; ++iz) {
        for (int iy = 0; iy < ny; ++iy) {
            for (int ix = 0; ix < nx; ++ix) {
                const T xc = (T)ix - (T)(nx / 2) + (T)0.5;   // :66 (integer nx/2)
                const T yc = (T)iy - (T)(ny / 2) + (T)0.5;   // :67
                const T zc = (T)iz - (T)(nz / 2) + (T)0.5;   // :68
                const T xr0 = xc * cphi + yc * sphi;          // rotated centre
                const T yr0 = -xc * sphi + yc * cphi;
                const T xll = xr0 - (T)0.5, yll = yr0 - (T)0.5;   // :82-83
                const T xrr = xr0 + (T)0.5, yrr = yr0 + (T)0.5;   // :84-85
                const T xrl = xr0 + (T)0.5, yrl = yr0 - (T)0.5;   // :87-88
                const T xlr = xr0 - (T)0.5, ylr = yr0 + (T)0.5;   // :89-90
                const T zll = zc, zrr = zc, zrl = zc, zlr = zc;   // :86,:91
                const T xt = xr0, yt = yr0, zt = zc + (T)0.5;     // :92-94
                const T xb = xr0, yb = yr0, zb = zc - (T)0.5;     // :95-97
                T ratio;
                ratio = SD / (xll + SO); const T all = ratio * yll, bll = ratio * zll;   // :105-107
                ratio = SD / (xrr + SO); const T arr = ratio * yrr, brr = ratio * zrr;   // :108-110
                ratio = SD / (xlr + SO); const T alr = ratio * ylr, blr = ratio * zlr;   // :111-113
                ratio = SD / (xrl + SO); const T arl = ratio * yrl, brl = ratio * zrl;   // :114-116
                ratio = SD / (xt + SO);  const T at = ratio * yt, bt = ratio * zt;       // :117-119
                ratio = SD / (xb + SO);  const T ab = ratio * yb, bb = ratio * zb;       // :120-122
                T a_max = tmax(tmax(tmax(all, arr), tmax(alr, arl)), tmax(at, ab));     // :129
                T a_min = tmin(tmin(tmin(all, arr), tmin(alr, arl)), tmin(at, ab));     // :130
                T b_max = tmax(tmax(tmax(bll, brr), tmax(blr, brl)), tmax(bt, bb));     // :131
                T b_min = tmin(tmin(tmin(bll, brr), tmin(blr, brl)), tmin(bt, bb));     // :132
                a_max = a_max / da - ai + (T)0.5;   // :135-138
                a_min = a_min / da - ai + (T)0.5;
                b_max = b_max / db - bi + (T)0.5;
                b_min = b_min / db - bi + (T)0.5;
                const int a_ind_min = (int)std::floor(a_min);   // :140
                const int b_ind_min = (int)std::floor(b_min);   // :142
                const T alim = ((T)na < a_max) ? (T)na : a_max; // MIN(na, a_max) is a float (:150)
                const T blim = ((T)nb < b_max) ? (T)nb : b_max; // :158
                T acc = 0;
                for (int ia = std::max(0, a_ind_min); (T)ia < alim; ++ia) {
                    const T wa = tmin((T)ia + 1, a_max) - tmax((T)ia, a_min);   // :156
                    for (int ib = std::max(0, b_ind_min); (T)ib < blim; ++ib) {
                        const T wb = tmin((T)ib + 1, b_max) - tmax((T)ib, b_min);   // :164
                        acc += wa * wb * proj[ia + (long long)ib * na];             // :167
                    }
                }
                img[lin(ix, iy, iz, nx, ny)] = acc;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Warp.  kernel_forwardDVF (src/cu_deform.cu:66-77) forms m_c = alpha_c*volume + beta_c*flow at
// each linear index; kernel_deformation (:79-90) writes out[id], id = i1 + i2*nx + i3*nx*ny, as
// the clamp-to-edge trilinear sample of img at (i1 + my[id], i2 + mx[id], i3 + mz[id]) (the
// texture is addressed with unnormalised coordinates and the -0.5 texel convention, :86-89).
// ---------------------------------------------------------------------------------------------
// A displacement beyond +-(n+1) lands outside the volume on the same side, where clamp-to-edge
// addressing returns the edge voxel whatever its size, so it is clamped to +-(n+1) before the
// floor and the integer conversion (which are then always in range).  A NaN displacement maps to
// -(n+1), the lower edge, as the HIP warp does (v_med3_f32 returns its smallest operand on NaN).
template <typename T>
inline T clamp_disp(T m, int n) {
    const T lim = (T)(n + 1);
    if (std::isnan(m)) return -lim;
    return m < -lim ? -lim : (m > lim ? lim : m);
}

// Clamp-to-edge trilinear sample of img (dim 0 fastest) at (i0 + m0, i1 + m1, i2 + m2).
template <typename T>
inline T trilinear(const T* img, int n0, int n1, int n2, int i0, T m0, int i1, T m1, int i2,
                   T m2) {
    const T p0 = (T)i0 + clamp_disp(m0, n0), p1 = (T)i1 + clamp_disp(m1, n1);
    const T p2 = (T)i2 + clamp_disp(m2, n2);
    const T f0 = std::floor(p0), f1 = std::floor(p1), f2 = std::floor(p2);
    const T t0 = p0 - f0, t1 = p1 - f1, t2 = p2 - f2;
    const long long i0f = (long long)f0, i1f = (long long)f1, i2f = (long long)f2;
    auto cl = [](long long i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); };
    const long long a0 = cl(i0f, n0), b0 = cl(i0f + 1, n0);
    const long long a1 = cl(i1f, n1), b1 = cl(i1f + 1, n1);
    const long long a2 = cl(i2f, n2), b2 = cl(i2f + 1, n2);
    auto at = [&](long long x, long long y, long long z) {
        return img[x + y * (long long)n0 + z * (long long)n0 * n1];
    };
    const T c00 = at(a0, a1, a2) * (1 - t0) + at(b0, a1, a2) * t0;
    const T c10 = at(a0, b1, a2) * (1 - t0) + at(b0, b1, a2) * t0;
    const T c01 = at(a0, a1, b2) * (1 - t0) + at(b0, a1, b2) * t0;
    const T c11 = at(a0, b1, b2) * (1 - t0) + at(b0, b1, b2) * t0;
    const T c0 = c00 * (1 - t1) + c10 * t1;
    const T c1 = c01 * (1 - t1) + c11 * t1;
    return c0 * (1 - t2) + c1 * t2;
}

template <typename T>
void deform(T* out, const T* img, int nx, int ny, int nz, T volume, T flow, const T* ax,
            const T* ay, const T* az, const T* bx, const T* by, const T* bz) {
    const long long N = (long long)nx * ny * nz;
    std::vector<T> src(img, img + N);   // the reference copies img into a CUDA array (:23)
#pragma omp parallel for schedule(static)
    for (long long id = 0; id < N; ++id) {
        const int i1 = (int)(id % nx);
        const int i2 = (int)((id / nx) % ny);
        const int i3 = (int)(id / ((long long)nx * ny));
        const T mx = ax[id] * volume + bx[id] * flow;   // :74
        const T my = ay[id] * volume + by[id] * flow;   // :75
        const T mz = az[id] * volume + bz[id] * flow;   // :76
        out[id] = trilinear(src.data(), nx, ny, nz, i1, my, i2, mx, i3, mz);
    }
}

// kernel_deformation2 (src/cu_deform.cu:92-113): first-order Taylor warp with the dim-1/dim-2
// boundary checks swapped as written (:100-107).
template <typename T>
void deform2(T* out, const T* img, int nx, int ny, int nz, T volume, T flow, const T* ax,
             const T* ay, const T* az, const T* bx, const T* by, const T* bz) {
    const long long N = (long long)nx * ny * nz;
    std::vector<T> src(img, img + N);
#pragma omp parallel for schedule(static)
    for (long long id = 0; id < N; ++id) {
        // thread coordinates: id = iy + ix*ny + iz*nx*ny (:98)
        const int iy = (int)(id % ny);
        const int ix = (int)((id / ny) % nx);
        const int iz = (int)(id / ((long long)nx * ny));
        auto rd = [&](long long j) { return (j < N) ? src[j] : src[id]; };
        const T dx = (ix == nx - 1) ? (T)0 : rd(id + 1) - src[id];                 // :100-103
        const T dy = (iy == ny - 1) ? (T)0 : rd(id + nx) - src[id];                // :104-107
        const T dz = (iz == nz - 1) ? (T)0 : rd(id + (long long)nx * ny) - src[id];   // :108-111
        const T mx = ax[id] * volume + bx[id] * flow;
        const T my = ay[id] * volume + by[id] * flow;
        const T mz = az[id] * volume + bz[id] * flow;
        out[id] = src[id] + dy * mx + dx * my + dz * mz;   // :112
    }
}

// Forward differences, zero at the last index: src/cu_diff.cu:2-33 (and inlined in
// cu_add.cu:37-57, cu_initial.cu:33-46, cu_update_udvf.cu:22-33).  ind: 1=x, 2=y, 3=z.
template <typename T>
inline T dfwd(const T* img, int ix, int iy, int iz, int nx, int ny, int nz, int ind) {
    const long long id = lin(ix, iy, iz, nx, ny);
    switch (ind) {
        case 1: return (ix == nx - 1) ? (T)0 : img[id + 1] - img[id];
        case 2: return (iy == ny - 1) ? (T)0 : img[id + nx] - img[id];
        default: return (iz == nz - 1) ? (T)0 : img[id + (long long)nx * ny] - img[id];
    }
}

#define FOR_VOXELS(nx, ny, nz)                                   \
    _Pragma("omp parallel for collapse(2) schedule(static)")     \
    for (int iz = 0; iz < (nz); ++iz)                            \
        for (int iy = 0; iy < (ny); ++iy)                        \
            for (int ix = 0; ix < (nx); ++ix)

template <typename T>
void diff_forward(T* out, const T* img, int nx, int ny, int nz, int ind) {
    FOR_VOXELS(nx, ny, nz) { out[lin(ix, iy, iz, nx, ny)] = dfwd(img, ix, iy, iz, nx, ny, nz, ind); }
}

// Backward difference, zero at the first and last index: src/mex_diff.cu:75-112.
template <typename T>
void diff_backward(T* out, const T* img, int nx, int ny, int nz, int ind) {
    FOR_VOXELS(nx, ny, nz) {
        const long long id = lin(ix, iy, iz, nx, ny);
        T v = 0;
        if (ind == 1) v = (ix == nx - 1 || ix == 0) ? (T)0 : img[id] - img[id - 1];
        else if (ind == 2) v = (iy == ny - 1 || iy == 0) ? (T)0 : img[id] - img[id - nx];
        else if (ind == 3) v = (iz == nz - 1 || iz == 0) ? (T)0 : img[id] - img[id - (long long)nx * ny];
        else v = img[id];   // the reference leaves img1 untouched for other ind; defined here as a copy
        out[id] = v;
    }
}

// kernel_update_udvf: src/cu_update_udvf.cu:11-62 (x/y write-back swapped as written).
template <typename T>
void update_udvf(T* ax, T* ay, T* az, T* bx, T* by, T* bz, const T* img, const T* img0,
                 T volume_diff, T flow_diff, int nx, int ny, int nz) {
    FOR_VOXELS(nx, ny, nz) {
        const long long id = lin(ix, iy, iz, nx, ny);
        const T dfI = img0[id] - img[id];                                   // :20
        const T dfx = dfwd(img, ix, iy, iz, nx, ny, nz, 1);                 // :22-25
        const T dfy = dfwd(img, ix, iy, iz, nx, ny, nz, 2);                 // :26-29
        const T dfz = dfwd(img, ix, iy, iz, nx, ny, nz, 3);                 // :30-33
        const T cax = dfx * volume_diff, cay = dfy * volume_diff, caz = dfz * volume_diff;
        const T cbx = dfx * flow_diff, cby = dfy * flow_diff, cbz = dfz * flow_diff;
        const T w = cax * cax + cay * cay + caz * caz + cbx * cbx + cby * cby + cbz * cbz;   // :43
        const T a_x = ay[id], a_y = ax[id], a_z = az[id];   // :45-47 (alpha_y read as "x")
        const T b_x = by[id], b_y = bx[id], b_z = bz[id];   // :48-50
        T b = dfI - cax * a_x - cay * a_y - caz * a_z - cbx * b_x - cby * b_y - cbz * b_z;
        if (w == 0) continue;                                                // :52-53
        b /= w;
        const T r = (T)1 / (T)6;
        ax[id] += r * b * cax;   // :55
        ay[id] += r * b * cay;   // :56
        az[id] += r * b * caz;
        bx[id] += r * b * cbx;
        by[id] += r * b * cby;
        bz[id] += r * b * cbz;   // :60
    }
}

// kernel_add (src/cu_add.cu:9-18): a += w*b over n elements.
template <typename T>
void add(T* a, const T* b, long long n, T w) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) a[i] += b[i] * w;
}

// kernel_add2 (src/cu_add.cu:28-58): a[ix,iy,iz] += b[iy,ix,iz] * w * D_ind(c)[ix,iy,iz].
template <typename T>
void add2(T* a, const T* b, int nx, int ny, int nz, const T* c, T w, int ind) {
    FOR_VOXELS(nx, ny, nz) {
        const long long id = lin(ix, iy, iz, nx, ny);
        const T df = dfwd(c, ix, iy, iz, nx, ny, nz, ind);
        a[id] += b[iy + (long long)ix * ny + (long long)iz * nx * ny] * w * df;   // :58
    }
}

// kernel_initial2 (src/cu_initial.cu:27-49): img = (D1+D2+D3)(img0) * (volume + flow).
template <typename T>
void initial2(T* img, int nx, int ny, int nz, const T* img0, T volume, T flow) {
    FOR_VOXELS(nx, ny, nz) {
        const T s = dfwd(img0, ix, iy, iz, nx, ny, nz, 1) + dfwd(img0, ix, iy, iz, nx, ny, nz, 2) +
                    dfwd(img0, ix, iy, iz, nx, ny, nz, 3);
        img[lin(ix, iy, iz, nx, ny)] = s * (volume + flow);
    }
}

// kernel_division (src/cu_division.cu:11-25): a /= b, NaN and +-Inf mapped to 0.
template <typename T>
void division(T* a, const T* b, long long n) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) {
        T v = a[i] / b[i];
        if (std::isnan(v) || std::isinf(v)) v = 0;
        a[i] = v;
    }
}

template <typename T>
T norm2(const T* x, long long n) {   // cublasSnrm2 at AUMISART.cu:48,51 (textbook 2-norm)
    T s = 0;
    for (long long i = 0; i < n; ++i) s += x[i] * x[i];
    return std::sqrt(s);
}

// ---------------------------------------------------------------------------------------------
// Host driver: both overloads of host_AUMISART (src/AUMISART.cu:3-141 static, :143-313 motion).
// motion == false: overload 1 (fields start at zero, no warps, motion sweep views step by 10).
// motion == true:  overload 2 (fields given, image sweep warps with v_i - v_0, P = 1).
// nonneg: clamp the image at zero after each view's update (absent in the reference; off = as
// written).  Outputs follow AUMISART.cu:52-53,64,121,209-211,234,291.
// ---------------------------------------------------------------------------------------------
template <typename T>
struct Geo {
    int nx, ny, nz, na, nb;
    T da, db, ai, bi, SO, SD;
};

template <typename T>
void sart_image_sweep(const Geo<T>& g, bool motion, int n_views, int i_iter, const T* angles,
                      const T* volumes, const T* flows, T lambda, bool nonneg, const T* h_proj,
                      T* img, const T* fields[6], T* outproj, T* outnorm, std::vector<T>& w) {
    const long long N = (long long)g.nx * g.ny * g.nz, D = (long long)g.na * g.nb;
    T* img1 = w.data();
    T* img_temp = img1 + N;
    T* img_ones = img_temp + N;
    T* proj = img_ones + N;
    T* proj_temp = proj + D;
    T* proj_ones = proj_temp + D;
    for (int iv = 0; iv < n_views; ++iv) {
        const T ang = angles[iv];
        const T* src = img;
        if (motion) {   // AUMISART.cu:193-194
            deform(img1, img, g.nx, g.ny, g.nz, volumes[iv] - volumes[0], flows[iv] - flows[0],
                   fields[0], fields[1], fields[2], fields[3], fields[4], fields[5]);
            src = img1;
        }
        projection(proj_temp, src, ang, g.SO, g.SD, g.da, g.na, g.ai, g.db, g.nb, g.bi, g.nx, g.ny, g.nz);
        std::memcpy(proj, h_proj + D * iv, sizeof(T) * D);                  // :203
        const T n0 = norm2(proj, D);                                        // :204
        add(proj, proj_temp, D, (T)-1);                                     // :206
        const T n1 = norm2(proj, D);                                        // :207
        outnorm[(long long)i_iter * n_views + iv] = n1 / n0;                // :209
        std::memcpy(outproj + D * iv, proj, sizeof(T) * D);                 // :211
        backprojection(img_temp, proj, ang, g.SO, g.SD, g.da, g.na, g.ai, g.db, g.nb, g.bi, g.nx, g.ny, g.nz);
        std::fill(img_ones, img_ones + N, (T)1);                            // :214
        projection(proj_ones, img_ones, ang, g.SO, g.SD, g.da, g.na, g.ai, g.db, g.nb, g.bi, g.nx, g.ny, g.nz);
        backprojection(img_ones, proj_ones, ang, g.SO, g.SD, g.da, g.na, g.ai, g.db, g.nb, g.bi, g.nx, g.ny, g.nz);
        division(img_temp, img_ones, N);                                    // :218
        if (motion) {   // :222-223
            deform(img1, img_temp, g.nx, g.ny, g.nz, volumes[0] - volumes[iv], flows[0] - flows[iv],
                   fields[0], fields[1], fields[2], fields[3], fields[4], fields[5]);
            add(img, img1, N, lambda);
        } else {
            add(img, img_temp, N, lambda);                                  // :62
        }
        if (nonneg) {
            for (long long i = 0; i < N; ++i) img[i] = img[i] < 0 ? (T)0 : img[i];
        }
    }
}

template <typename T>
void sart_motion_sweep(const Geo<T>& g, int n_views, int stride, int i_iter, const T* angles,
                       const T* volumes, const T* flows, const T* h_proj, const T* img,
                       T* fields[6], T* outnorm, std::vector<T>& w) {
    const long long N = (long long)g.nx * g.ny * g.nz, D = (long long)g.na * g.nb;
    T* img0 = w.data();
    T* img_temp = img0 + N;
    T* img_ones = img_temp + N;
    T* proj = img_ones + N;
    T* proj_temp = proj + D;
    T* proj_ones = proj_temp + D;
    T *ax = fields[0], *ay = fields[1], *az = fields[2], *bx = fields[3], *by = fields[4], *bz = fields[5];
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    for (int iv = 1; iv < n_views; iv += stride) {   // :242 (overload 1: :72)
        const T ang = angles[iv];
        projection(proj_temp, img, ang, g.SO, g.SD, g.da, g.na, g.ai, g.db, g.nb, g.bi, nx, ny, nz);
        std::memcpy(proj, h_proj + D * iv, sizeof(T) * D);
        const T n0 = norm2(proj, D);
        const T vd = volumes[iv] - volumes[iv - 1];   // :249
        const T fd = flows[iv] - flows[iv - 1];       // :250
        add(proj, proj_temp, D, (T)-1);               // :251
        std::fill(img0, img0 + N, (T)0);              // :252
        add2(img0, ay, nx, ny, nz, img, vd, 1);       // :253-258
        add2(img0, ax, nx, ny, nz, img, vd, 2);
        add2(img0, az, nx, ny, nz, img, vd, 3);
        add2(img0, by, nx, ny, nz, img, fd, 1);
        add2(img0, bx, nx, ny, nz, img, fd, 2);
        add2(img0, bz, nx, ny, nz, img, fd, 3);
        projection(proj_temp, img0, ang, g.SO, g.SD, g.da, g.na, g.ai, g.db, g.nb, g.bi, nx, ny, nz);
        add(proj, proj_temp, D, (T)1);                // :260
        const T n1 = norm2(proj, D);
        outnorm[(long long)i_iter * n_views + iv] = n1 / n0;   // :263
        backprojection(img_temp, proj, ang, g.SO, g.SD, g.da, g.na, g.ai, g.db, g.nb, g.bi, nx, ny, nz);
        initial2(img_ones, nx, ny, nz, img, -vd, -fd);   // :267
        projection(proj_ones, img_ones, ang, g.SO, g.SD, g.da, g.na, g.ai, g.db, g.nb, g.bi, nx, ny, nz);
        backprojection(img_ones, proj_ones, ang, g.SO, g.SD, g.da, g.na, g.ai, g.db, g.nb, g.bi, nx, ny, nz);
        division(img_temp, img_ones, N);                 // :270
        const T wv = volumes[iv - 1] - volumes[iv], wf = flows[iv - 1] - flows[iv];
        add2(ay, img_temp, nx, ny, nz, img, wv, 1);      // :272-277
        add2(ax, img_temp, nx, ny, nz, img, wv, 2);
        add2(az, img_temp, nx, ny, nz, img, wv, 3);
        add2(by, img_temp, nx, ny, nz, img, wf, 1);
        add2(bx, img_temp, nx, ny, nz, img, wf, 2);
        add2(bz, img_temp, nx, ny, nz, img, wf, 3);
    }
}

template <typename T>
void aumisart(T* outimg, T* outproj, T* outnorm, T* outalphax, const T* h_img, const T* h_proj,
              const Geo<T>& g, int n_views, int n_iter, const int* op_iter, T lambda,
              const T* volumes, const T* flows, const T* angles, const T* const in_fields[6],
              bool motion, bool nonneg) {
    const long long N = (long long)g.nx * g.ny * g.nz, D = (long long)g.na * g.nb;
    std::vector<T> img(h_img, h_img + N);
    std::vector<T> f[6];
    T* fp[6];
    const T* cfp[6];
    for (int c = 0; c < 6; ++c) {
        if (motion) f[c].assign(in_fields[c], in_fields[c] + N);
        else f[c].assign(N, (T)0);   // AUMISART.cu:25-30
        fp[c] = f[c].data();
        cfp[c] = fp[c];
    }
    std::vector<T> work(3 * N + 3 * D);
    std::memcpy(outimg, h_img, sizeof(T) * N);
    std::memset(outnorm, 0, sizeof(T) * (size_t)n_iter * n_views);
    std::memset(outproj, 0, sizeof(T) * (size_t)D * n_views);
    for (int it = 0; it < n_iter; ++it) {
        if (op_iter[it] == 1) {
            sart_image_sweep(g, motion, n_views, it, angles, volumes, flows, lambda, nonneg, h_proj,
                             img.data(), cfp, outproj, outnorm, work);
            std::memcpy(outimg, img.data(), sizeof(T) * N);   // :64 / :234
        } else {
            // :68-71 / :238-241 reload the image from the host copy
            if (it == 0) img.assign(h_img, h_img + N);
            else img.assign(outimg, outimg + N);
            sart_motion_sweep(g, n_views, motion ? 1 : 10, it, angles, volumes, flows, h_proj,
                              img.data(), fp, outnorm, work);
        }
    }
    std::memcpy(outalphax, fp[0], sizeof(T) * N);   // :121 / :291
    if (!motion) std::memcpy(outimg, img.data(), sizeof(T) * N);   // :123 (overload 1 only)
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// extern "C" surface used by the Python test harness (ctypes).  All arrays are T.
// ---------------------------------------------------------------------------------------------
// kernel_multiple (src/cu_multiple.cu:11-43): img1[id] *= weight * D_ind(img0)[id].
template <typename T>
void multiple(T* a, int nx, int ny, int nz, const T* b, T w, int ind) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const long long id = lin(ix, iy, iz, nx, ny);
                a[id] *= w * dfwd(b, ix, iy, iz, nx, ny, nz, ind);
            }
}

// kernel_invertDVF (waste/kernel_invertDVF.cu:1-18): the texture fetch tex3D(m, c + 0.5) with
// linear filtering and clamp addressing samples m at index-space point c, dim 0 = x fastest.
// Each component is updated in turn with the components already updated (Gauss-Seidel).
template <typename T>
void invert_dvf(T* mx2, T* my2, T* mz2, const T* mx, const T* my, const T* mz, int nx, int ny,
                int nz, int niter) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                T x = 0, y = 0, z = 0;
                for (int it = 0; it < niter; ++it) {
                    x = -trilinear(mx, nx, ny, nz, ix, x, iy, y, iz, z);   // :11
                    y = -trilinear(my, nx, ny, nz, ix, x, iy, y, iz, z);   // :12
                    z = -trilinear(mz, nx, ny, nz, ix, x, iy, y, iz, z);   // :13
                }
                const long long id = lin(ix, iy, iz, nx, ny);
                mx2[id] = x;
                my2[id] = y;
                mz2[id] = z;
            }
}

// Forward simulator in the order include/generateProj.h:21-24 lists its kernels: forwardDVF ->
// (optional) invertDVF -> deformation -> projection, per view.
template <typename T>
void generate_proj(T* proj, const T* img, const T* const fl[6], const T* volumes, const T* flows,
                   const T* angles, int n_views, int niter, T SO, T SD, T da, int na, T ai, T db,
                   int nb, T bi, int nx, int ny, int nz) {
    const long long N = (long long)nx * ny * nz, D = (long long)na * nb;
    std::vector<T> m(3 * N), mi(3 * N), w(N);
    const std::vector<T> zero(N, T(0));
    for (int iv = 0; iv < n_views; ++iv) {
        for (long long id = 0; id < N; ++id)
            for (int c = 0; c < 3; ++c)   // kernel_forwardDVF (src/cu_deform.cu:73-76)
                m[c * N + id] = fl[c][id] * volumes[iv] + fl[c + 3][id] * flows[iv];
        const T* u = m.data();
        if (niter > 0) {
            invert_dvf(mi.data(), mi.data() + N, mi.data() + 2 * N, m.data(), m.data() + N,
                       m.data() + 2 * N, nx, ny, nz, niter);
            u = mi.data();
        }
        deform(w.data(), img, nx, ny, nz, T(1), T(0), u, u + N, u + 2 * N, zero.data(),
               zero.data(), zero.data());
        projection(proj + iv * D, w.data(), angles[iv], SO, SD, da, na, ai, db, nb, bi, nx, ny,
                   nz);
    }
}

#define ORACLE_EXPORTS(T, SFX)                                                                     \
    extern "C" void oracle_projection_##SFX(T* proj, const T* img, T angle, T SO, T SD, T da,      \
                                            int na, T ai, T db, int nb, T bi, int nx, int ny,      \
                                            int nz) {                                              \
        projection<T>(proj, img, angle, SO, SD, da, na, ai, db, nb, bi, nx, ny, nz);               \
    }                                                                                              \
    extern "C" void oracle_backprojection_##SFX(T* img, const T* proj, T angle, T SO, T SD, T da,  \
                                                int na, T ai, T db, int nb, T bi, int nx, int ny,  \
                                                int nz) {                                          \
        backprojection<T>(img, proj, angle, SO, SD, da, na, ai, db, nb, bi, nx, ny, nz);           \
    }                                                                                              \
    extern "C" void oracle_deform_##SFX(T* out, const T* img, int nx, int ny, int nz, T volume,    \
                                        T flow, const T* ax, const T* ay, const T* az,             \
                                        const T* bx, const T* by, const T* bz) {                   \
        deform<T>(out, img, nx, ny, nz, volume, flow, ax, ay, az, bx, by, bz);                     \
    }                                                                                              \
    extern "C" void oracle_deform2_##SFX(T* out, const T* img, int nx, int ny, int nz, T volume,   \
                                         T flow, const T* ax, const T* ay, const T* az,            \
                                         const T* bx, const T* by, const T* bz) {                  \
        deform2<T>(out, img, nx, ny, nz, volume, flow, ax, ay, az, bx, by, bz);                    \
    }                                                                                              \
    extern "C" void oracle_update_udvf_##SFX(T* ax, T* ay, T* az, T* bx, T* by, T* bz,             \
                                             const T* img, const T* img0, T vd, T fd, int nx,      \
                                             int ny, int nz) {                                     \
        update_udvf<T>(ax, ay, az, bx, by, bz, img, img0, vd, fd, nx, ny, nz);                     \
    }                                                                                              \
    extern "C" void oracle_add_##SFX(T* a, const T* b, long long n, T w) { add<T>(a, b, n, w); }   \
    extern "C" void oracle_add2_##SFX(T* a, const T* b, int nx, int ny, int nz, const T* c, T w,   \
                                      int ind) {                                                   \
        add2<T>(a, b, nx, ny, nz, c, w, ind);                                                      \
    }                                                                                              \
    extern "C" void oracle_initial2_##SFX(T* img, int nx, int ny, int nz, const T* img0, T vol,    \
                                          T flow) {                                                \
        initial2<T>(img, nx, ny, nz, img0, vol, flow);                                             \
    }                                                                                              \
    extern "C" void oracle_division_##SFX(T* a, const T* b, long long n) { division<T>(a, b, n); } \
    extern "C" void oracle_diff_forward_##SFX(T* o, const T* i, int nx, int ny, int nz, int ind) { \
        diff_forward<T>(o, i, nx, ny, nz, ind);                                                    \
    }                                                                                              \
    extern "C" void oracle_diff_backward_##SFX(T* o, const T* i, int nx, int ny, int nz,           \
                                               int ind) {                                          \
        diff_backward<T>(o, i, nx, ny, nz, ind);                                                   \
    }                                                                                              \
    extern "C" T oracle_norm2_##SFX(const T* x, long long n) { return norm2<T>(x, n); }            \
    extern "C" void oracle_multiple_##SFX(T* a, int nx, int ny, int nz, const T* b, T w, int ind) { \
        multiple<T>(a, nx, ny, nz, b, w, ind);                                                     \
    }                                                                                              \
    extern "C" void oracle_invert_dvf_##SFX(T* mx2, T* my2, T* mz2, const T* mx, const T* my,      \
                                            const T* mz, int nx, int ny, int nz, int niter) {      \
        invert_dvf<T>(mx2, my2, mz2, mx, my, mz, nx, ny, nz, niter);                               \
    }                                                                                              \
    extern "C" void oracle_generate_proj_##SFX(                                                    \
        T* proj, const T* img, const T* ax, const T* ay, const T* az, const T* bx, const T* by,    \
        const T* bz, const T* volumes, const T* flows, const T* angles, int n_views, int niter,    \
        T SO, T SD, T da, int na, T ai, T db, int nb, T bi, int nx, int ny, int nz) {              \
        const T* const fl[6] = {ax, ay, az, bx, by, bz};                                           \
        generate_proj<T>(proj, img, fl, volumes, flows, angles, n_views, niter, SO, SD, da, na,    \
                         ai, db, nb, bi, nx, ny, nz);                                              \
    }                                                                                              \
    extern "C" void oracle_aumisart_##SFX(                                                         \
        T* outimg, T* outproj, T* outnorm, T* outalphax, const T* h_img, const T* h_proj, int nx,  \
        int ny, int nz, int na, int nb, int n_views, int n_iter, const int* op_iter, T da, T db,   \
        T ai, T bi, T SO, T SD, T lambda, const T* volumes, const T* flows, const T* angles,       \
        const T* ax, const T* ay, const T* az, const T* bx, const T* by, const T* bz, int motion,  \
        int nonneg) {                                                                              \
        Geo<T> g{nx, ny, nz, na, nb, da, db, ai, bi, SO, SD};                                      \
        const T* const fl[6] = {ax, ay, az, bx, by, bz};                                           \
        aumisart<T>(outimg, outproj, outnorm, outalphax, h_img, h_proj, g, n_views, n_iter,        \
                    op_iter, lambda, volumes, flows, angles, fl, motion != 0, nonneg != 0);        \
    }

// One motion (udvf) sweep of host_AUMISART on its own (src/AUMISART.cu:242-288, overload 1's
// stride-10 form :72-118): the six fields are updated in place and the per-view residual ratios
// go to outnorm[n_views] (view 0 stays 0).  Tests start it from any state (teacher forcing).
#define ORACLE_SWEEP_EXPORT(T, SFX)                                                                \
    extern "C" void oracle_motion_sweep_##SFX(                                                     \
        T* ax, T* ay, T* az, T* bx, T* by, T* bz, T* outnorm, const T* img, const T* h_proj,       \
        int nx, int ny, int nz, int na, int nb, int n_views, int stride, const T* angles,          \
        const T* volumes, const T* flows, T da, T db, T ai, T bi, T SO, T SD) {                    \
        Geo<T> g{nx, ny, nz, na, nb, da, db, ai, bi, SO, SD};                                      \
        T* fp[6] = {ax, ay, az, bx, by, bz};                                                       \
        const long long N = (long long)nx * ny * nz, D = (long long)na * nb;                       \
        std::vector<T> work(3 * N + 3 * D);                                                        \
        std::memset(outnorm, 0, sizeof(T) * (size_t)n_views);                                      \
        sart_motion_sweep<T>(g, n_views, stride, 0, angles, volumes, flows, h_proj, img, fp,       \
                             outnorm, work);                                                       \
    }

ORACLE_EXPORTS(double, f64)
ORACLE_EXPORTS(float, f32)
ORACLE_SWEEP_EXPORT(double, f64)
ORACLE_SWEEP_EXPORT(float, f32)

extern "C" int oracle_num_threads(void) { return omp_get_max_threads(); }
extern "C" void oracle_set_num_threads(int n) { omp_set_num_threads(n); }

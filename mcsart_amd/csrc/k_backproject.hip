// Voxel-driven footprint backprojector for gfx950.
//
// Reference: kernel_backprojection, src/cu_backprojection.cu:49-170 (one thread per voxel, every
// call allocates a CUDA array, copies the projection into it and builds a texture, :10-45).
//
// Here one lane owns an (ix, iy) column and walks a z-slab of voxels: the transverse footprint
// (a bounds, weights) of a column does not depend on z, so it is formed once per column in FP64
// (footprint.hpp) and kept in registers; each voxel only forms its axial bounds.  A wavefront is
// 64 neighbouring ix of one iy, so its detector reads fall on neighbouring cells of the same
// detector rows, and its volume writes are contiguous.  The projection is read directly from
// global memory (it is at most a few MB and stays in L2).  The SART mode reads (r, P1) interleaved
// as float2 and writes B(r)/B(P1) with NaN/Inf -> 0, fusing src/AUMISART.cu:212-218 (two
// backprojections and a division) into one pass.
#include "footprint.hpp"
#include "internal.hpp"

namespace mcsart {
namespace {

constexpr int BX = 64, BY = 4;

template <int MODE>
__global__ void __launch_bounds__(BX *BY) k_backproject(GeomArgs g, double c, double s,
                                                       const float *__restrict__ proj,
                                                       const float2 *__restrict__ rp,
                                                       float *__restrict__ out, int zslab,
                                                       NormFinish nf) {
    // the preceding projector's per-view norm (its partials are complete: kernel boundary)
    if (nf.out && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.y == 0)
        finish_norm_wave(nf);
    const int ix = blockIdx.x * BX + threadIdx.x;
    const int iy = blockIdx.y * BY + threadIdx.y;
    if (ix >= g.nx || iy >= g.ny) return;
    const int iz0 = blockIdx.z * zslab;
    const int iz1 = min(g.nz, iz0 + zslab);

    Column C;
    column_setup(g, c, s, ix, iy, C);
    // a range (src/cu_backprojection.cu:140,150): ia from max(0, floor(Amin)) while
    // ia < MIN(na, Amax); weights in coordinates local to the first cell.
    // clamped in FP64 before the conversion so that it is always in range (NaN -> 0)
    const int ia_lo = (int)fmin(fmax(floor(C.Amin), 0.0), (double)g.na);
    const float aminL = (float)(C.Amin - ia_lo);
    const float amaxL = (float)(C.Amax - ia_lo);
    const float alim = ((double)g.na < C.Amax) ? (float)(g.na - ia_lo) : amaxL;
    const long long plane = (long long)g.nx * g.ny;

    for (int iz = iz0; iz < iz1; ++iz) {
        double Bmin, Bmax;
        column_b(C, g.nz, iz, Bmin, Bmax);
        const int ib_lo = (int)fmin(fmax(floor(Bmin), 0.0), (double)g.nb);
        const float bminL = (float)(Bmin - ib_lo);
        const float bmaxL = (float)(Bmax - ib_lo);
        const float blim = ((double)g.nb < Bmax) ? (float)(g.nb - ib_lo) : bmaxL;
        float acc = 0.f, acc1 = 0.f;
        for (int lb = 0; (float)lb < blim; ++lb) {
            const float wb = fminf(lb + 1.f, bmaxL) - fmaxf((float)lb, bminL);   // :164
            const int row = (ib_lo + lb) * g.na + ia_lo;
            for (int la = 0; (float)la < alim; ++la) {
                const float wa = fminf(la + 1.f, amaxL) - fmaxf((float)la, aminL);   // :156
                const float w = wa * wb;
                if (MODE == BACK_PLAIN) {
                    acc += w * proj[row + la];
                } else {
                    const float2 v = rp[row + la];
                    acc += w * v.x;
                    acc1 += w * v.y;
                }
            }
        }
        const long long id = ix + (long long)iy * g.nx + iz * plane;
        if (MODE == BACK_PLAIN) {
            out[id] = acc;
        } else {
            float u = acc / acc1;   // src/cu_division.cu:20-24
            if (isnan(u) || isinf(u)) u = 0.f;
            out[id] = u;
        }
    }
}

int pick_zslab(const Geom &g) {
    const long long cols = (long long)((g.nx + BX - 1) / BX) * ((g.ny + BY - 1) / BY);
    int zs = 32;
    while (zs > 1 && cols * ((g.nz + zs - 1) / zs) < 4096) zs /= 2;
    return zs;
}

}  // namespace

int launch_backprojector(const Geom &g, const ViewTrig &t, BackMode mode, const float *proj,
                         const float2 *rp, float *out, hipStream_t st, const NormFinish *nfp) {
    const GeomArgs ga{g.nx, g.ny, g.nz, g.na, g.nb, g.da, g.db, g.ai, g.bi, g.SO, g.SD};
    const int zs = pick_zslab(g);
    const dim3 block(BX, BY);
    const dim3 grid((g.nx + BX - 1) / BX, (g.ny + BY - 1) / BY, (g.nz + zs - 1) / zs);
    const NormFinish nf = nfp ? *nfp : NormFinish{nullptr, 0, nullptr};
    if (mode == BACK_PLAIN)
        hipLaunchKernelGGL(k_backproject<BACK_PLAIN>, grid, block, 0, st, ga, t.c, t.s, proj,
                           rp, out, zs, nf);
    else
        hipLaunchKernelGGL(k_backproject<BACK_RATIO>, grid, block, 0, st, ga, t.c, t.s, proj,
                           rp, out, zs, nf);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

}  // namespace mcsart

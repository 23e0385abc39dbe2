// Cone-beam footprint forward projector for gfx950.
//
// Reference: kernel_projection, src/cu_projection.cu:15-145 (one thread per detector cell, a
// loop over every plane, accumulation straight into global memory, 4 redundant z-threads).
//
// Here a workgroup owns a 64-cell detector-row segment; its 64*S lanes are arranged as 64 rays x
// S plane-chunks, so that one wavefront is 64 neighbouring rays walking the same plane (their
// footprints are neighbouring voxels, so loads coalesce along the transverse axis in the y-branch
// and hit the same cache lines plane after plane in the x-branch).  Per-ray constants are formed
// in FP64 once (footprint.hpp); each plane costs four FMAs for its footprint and, in the common
// case of a footprint at most two voxels wide in each direction, four gathers with no
// data-dependent loop.  The S partial sums meet in LDS.  SART modes fuse the residual
// r = p - P(x), the normaliser projection P(1) (image sweep) or the second projection P(q)
// (motion sweep), and the per-view norms ||p||, ||r|| (reduced deterministically by the last
// workgroup to finish), replacing src/AUMISART.cu:46-52 / :194-209 / :246-263.
#include "footprint.hpp"
#include "internal.hpp"

namespace mcsart {
namespace {

constexpr int RAYS = 64;

template <int MODE>
__global__ void __launch_bounds__(512) k_project(GeomArgs g, double c, double s, ProjArgs a,
                                                 int nsplit) {
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int ia = blockIdx.x * RAYS + tx;
    const int ib = blockIdx.y;
    const bool live = ia < g.na;

    float acc0 = 0.f, acc1 = 0.f;
    if (live) {
        Ray R;
        ray_setup(g, c, s, ia, ib, R);
        const int np = R.xbranch ? g.nx : g.ny;
        const int nt = R.xbranch ? g.ny : g.nx;
        const int sp = R.xbranch ? 1 : g.nx;        // stride between planes
        const int st = R.xbranch ? g.nx : 1;        // stride along the transverse axis
        const int sz = g.nx * g.ny;                 // stride along z
        const float p0 = 0.5f - 0.5f * (float)np;   // (float)ip + 0.5 - n/2 (:52, :105)
        const int chunk = (np + nsplit - 1) / nsplit;
        const int ip_lo = ty * chunk;
        const int ip_hi = min(np, ip_lo + chunk);
        const float *__restrict__ v0 = a.vol0;
        const float *__restrict__ v1 = a.vol1;
        for (int ip = ip_lo; ip < ip_hi; ++ip) {
            PlaneFP F;
            if (!plane_fp(R, (float)ip + p0, nt, g.nz, F)) continue;
            const int pbase = ip * sp;
            const float sc = F.it * F.iz;
            if (F.fast) {
                const int t0 = clampi(F.T0, 0, nt - 1), t1 = clampi(F.T0 + 1, 0, nt - 1);
                const int z0 = clampi(F.Z0, 0, g.nz - 1), z1 = clampi(F.Z0 + 1, 0, g.nz - 1);
                const int i00 = pbase + t0 * st + z0 * sz, i01 = pbase + t0 * st + z1 * sz;
                const int i10 = pbase + t1 * st + z0 * sz, i11 = pbase + t1 * st + z1 * sz;
                if (MODE == PROJ_IMAGE) {
                    acc1 += sc * (F.wt0 + F.wt1) * (F.wz0 + F.wz1);   // P(1)
                }
                {
                    const float x00 = v0[i00], x01 = v0[i01], x10 = v0[i10], x11 = v0[i11];
                    acc0 += sc * (F.wt0 * (F.wz0 * x00 + F.wz1 * x01) +
                                  F.wt1 * (F.wz0 * x10 + F.wz1 * x11));
                }
                if (MODE == PROJ_MOTION) {
                    const float q00 = v1[i00], q01 = v1[i01], q10 = v1[i10], q11 = v1[i11];
                    acc1 += sc * (F.wt0 * (F.wz0 * q00 + F.wz1 * q01) +
                                  F.wt1 * (F.wz0 * q10 + F.wz1 * q11));
                }
            } else {
                const int Tl = max(F.T0, 0), Th = min(nt - 1, (int)fminf(floorf(F.t2), (float)nt));
                const int Zl = max(F.Z0, 0);
                const int Zh = min(g.nz - 1, (int)fminf(floorf(F.z2), (float)g.nz));
                for (int it = Tl; it <= Th; ++it) {
                    const float wt = (fminf(it + 1.f, F.t2) - fmaxf((float)it, F.t1)) * F.it;
                    for (int iz = Zl; iz <= Zh; ++iz) {
                        const float w =
                            wt * (fminf(iz + 1.f, F.z2) - fmaxf((float)iz, F.z1)) * F.iz;
                        const int idx = pbase + it * st + iz * sz;
                        acc0 += w * v0[idx];
                        if (MODE == PROJ_IMAGE) acc1 += w;
                        if (MODE == PROJ_MOTION) acc1 += w * v1[idx];
                    }
                }
            }
        }
    }

    // Combine the plane chunks (fixed order: deterministic).
    __shared__ float red0[8][RAYS];
    __shared__ float red1[8][RAYS];
    if (nsplit > 1) {
        red0[ty][tx] = acc0;
        if (MODE != PROJ_PLAIN) red1[ty][tx] = acc1;
        __syncthreads();
        if (ty == 0) {
            acc0 = red0[0][tx];
            if (MODE != PROJ_PLAIN) acc1 = red1[0][tx];
            for (int k = 1; k < nsplit; ++k) {
                acc0 += red0[k][tx];
                if (MODE != PROJ_PLAIN) acc1 += red1[k][tx];
            }
        }
    }
    if (ty != 0) return;

    const long long id = (long long)ib * g.na + ia;
    if (MODE == PROJ_PLAIN) {
        if (live) a.out[id] = acc0;
        return;
    }

    // SART epilogue: residual, normaliser, norms.
    double pp = 0.0, rr = 0.0;
    if (live) {
        const float p = a.meas[id];
        const float r = p - acc0;
        a.rp[id] = make_float2(r, acc1);
        if (a.resid_copy) a.resid_copy[id] = r;
        pp = (double)p * (double)p;
        rr = (double)r * (double)r;
    }
    // wavefront (64 lanes) reduction
    for (int off = 32; off > 0; off >>= 1) {
        pp += __shfl_xor(pp, off, 64);
        rr += __shfl_xor(rr, off, 64);
    }
    const int nblk = gridDim.x * gridDim.y;
    const int bid = blockIdx.y * gridDim.x + blockIdx.x;
    int last = 0;
    if (tx == 0) {
        a.partials[2 * bid] = pp;
        a.partials[2 * bid + 1] = rr;
        __threadfence();
        const unsigned prev = atomicAdd(a.counter, 1u);
        last = (prev == (unsigned)(nblk - 1)) ? 1 : 0;
    }
    last = __shfl(last, 0, 64);   // only wavefront 0 of the workgroup is still running
    if (!last) return;
    __threadfence();
    double sp2 = 0.0, sr2 = 0.0;
    const volatile double *part = a.partials;
    for (int k = tx; k < nblk; k += 64) {
        sp2 += part[2 * k];
        sr2 += part[2 * k + 1];
    }
    for (int off = 32; off > 0; off >>= 1) {
        sp2 += __shfl_xor(sp2, off, 64);
        sr2 += __shfl_xor(sr2, off, 64);
    }
    if (tx == 0) {
        const float np2 = (float)sqrt(sp2), nr2 = (float)sqrt(sr2);
        *a.norm_out = nr2 / np2;   // src/AUMISART.cu:209
        atomicExch(a.counter, 0u);
    }
}

int pick_split(const Geom &g) {
    const long long tiles = (long long)((g.na + RAYS - 1) / RAYS) * g.nb;
    int s = 1;
    while (s < 8 && tiles * s < 4096) s *= 2;
    const int np = g.nx > g.ny ? g.nx : g.ny;
    while (s > 1 && np / s < 8) s /= 2;
    return s;
}

}  // namespace

int projector_partials(const Geom &g) { return ((g.na + RAYS - 1) / RAYS) * g.nb; }

int launch_projector(const Geom &g, const ViewTrig &t, ProjMode mode, const ProjArgs &a,
                     hipStream_t st) {
    const GeomArgs ga{g.nx, g.ny, g.nz, g.na, g.nb, g.da, g.db, g.ai, g.bi, g.SO, g.SD};
    const int s = pick_split(g);
    const dim3 block(RAYS, s);
    const dim3 grid((g.na + RAYS - 1) / RAYS, g.nb);
    switch (mode) {
        case PROJ_PLAIN:
            hipLaunchKernelGGL(k_project<PROJ_PLAIN>, grid, block, 0, st, ga, t.c, t.s, a, s);
            break;
        case PROJ_IMAGE:
            hipLaunchKernelGGL(k_project<PROJ_IMAGE>, grid, block, 0, st, ga, t.c, t.s, a, s);
            break;
        case PROJ_MOTION:
            hipLaunchKernelGGL(k_project<PROJ_MOTION>, grid, block, 0, st, ga, t.c, t.s, a, s);
            break;
    }
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

}  // namespace mcsart

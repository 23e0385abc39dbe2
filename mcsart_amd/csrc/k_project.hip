// Cone-beam footprint forward projector for gfx950.
//
// Reference: kernel_projection, src/cu_projection.cu:15-145 (one thread per detector cell, a
// loop over every plane with data-dependent footprint loops, accumulation straight into global
// memory, 4 redundant z-threads per cell).
//
// Layout of the work: a workgroup owns 64 neighbouring detector cells of one row and splits their
// planes into S chunks, one wavefront per chunk, so a wavefront is 64 neighbouring rays walking
// the same plane.  The per-ray constants (footprint.hpp, FP64) are formed once per ray by the
// first wavefront and shared through LDS.
//
// Per plane the body is branch-free: the footprint's bounds are four FMAs, and the overlap
// weights of a fixed WT x WZ window of voxels (WT, WZ in {2, 3}, chosen per view on the host as
// the largest span any footprint can have) are formed with masks instead of loops, so the
// compiler can unroll and keep several planes' gathers in flight.  A plane outside the volume
// simply gets weight zero.  Gathers use clamped 32-bit indices.
//
// Memory layout: x-branch rays step through x-planes, so with the natural x-fastest layout
// neighbouring rays would read one float from 64 different cache lines.  They read instead an
// xy-transposed copy (written by the kernel that produced the volume), in which their footprints
// are contiguous; y-branch rays read the natural layout.  In the motion sweep the two volumes
// (g, q) are interleaved as float2, so each gather is one 8-byte load for both projections.
//
// SART modes fuse the residual r = p - P(x), the normaliser projection P(1) (image sweep) or the
// second projection P(q) (motion sweep) and the per-view norms ||p||, ||r|| (per-workgroup FP64
// partials written with plain stores; the next kernel on the stream sums them in a fixed order),
// replacing src/AUMISART.cu:46-52, :194-209 and :246-263.  No fence or atomic is used: on a
// multi-XCD part a device-scope release in every workgroup forces an L2 write-back each time.
#include "footprint.hpp"
#include "internal.hpp"

namespace mcsart {
namespace {

constexpr int RAYS = 64;
constexpr int MAXSPLIT = 8;

// Footprint of one plane over a fixed WT x WZ window of voxels: the overlap weights of
// src/cu_projection.cu:81-89 (masked to zero outside the volume), the normalisation
// 1/((t2-t1)(z2-z1)) (:83,:86) and the gather indices (clamped into the volume).
template <int WT, int WZ>
struct PlaneW {
    bool in;
    float sc;
    float wt[WT], wz[WZ];
    unsigned idx[WT][WZ];
};

template <int WT, int WZ>
__device__ inline void plane_weights(const Ray &R, int ip, float p0, int nt, int nz, float ntf,
                                     float nzf, unsigned sp, unsigned st, unsigned sz,
                                     PlaneW<WT, WZ> &P) {
    const float p = (float)ip + p0;
    const float ta = fmaf(R.rA, p, R.KA), tb = fmaf(R.rB, p, R.KB);
    const float za = fmaf(R.rzA, p, R.KzA), zb = fmaf(R.rzB, p, R.KzB);
    const float t1 = fminf(ta, tb), t2 = fmaxf(ta, tb);
    const float z1 = fminf(za, zb), z2 = fmaxf(za, zb);
    P.in = (t2 > 0.f) & (t1 < ntf) & (z2 > 0.f) & (z1 < nzf);
    P.sc = P.in ? __builtin_amdgcn_rcpf((t2 - t1) * (z2 - z1)) : 0.f;
    const float ft = fminf(fmaxf(floorf(t1), -4.f), ntf);
    const float fz = fminf(fmaxf(floorf(z1), -4.f), nzf);
    const int T0 = (int)ft, Z0 = (int)fz;
    unsigned it[WT], iz[WZ];
#pragma unroll
    for (int k = 0; k < WT; ++k) {
        // overlap of voxel T0+k with [t1, t2]: min(i+1, t2) - max(i, t1), >= 0
        const float lo = (k == 0) ? t1 : ft + (float)k;
        const float w = fmaxf(fminf(ft + (float)(k + 1), t2) - lo, 0.f);
        const int i = T0 + k;
        P.wt[k] = (i >= 0 && i < nt) ? w : 0.f;
        it[k] = (unsigned)clampi(i, 0, nt - 1) * st;
    }
#pragma unroll
    for (int k = 0; k < WZ; ++k) {
        const float lo = (k == 0) ? z1 : fz + (float)k;
        const float w = fmaxf(fminf(fz + (float)(k + 1), z2) - lo, 0.f);
        const int i = Z0 + k;
        P.wz[k] = (i >= 0 && i < nz) ? w : 0.f;
        iz[k] = (unsigned)clampi(i, 0, nz - 1) * sz;
    }
    const unsigned pb = (unsigned)ip * sp;
#pragma unroll
    for (int k = 0; k < WT; ++k)
#pragma unroll
        for (int j = 0; j < WZ; ++j) P.idx[k][j] = pb + it[k] + iz[j];
}

template <typename T, int WT, int WZ>
__device__ inline void gather(const T *__restrict__ v, const PlaneW<WT, WZ> &P, T (&out)[WT][WZ]) {
#pragma unroll
    for (int k = 0; k < WT; ++k)
#pragma unroll
        for (int j = 0; j < WZ; ++j) out[k][j] = v[P.idx[k][j]];
}

// select, not multiply: a plane outside the volume adds exactly nothing even when the clamped
// gathers hit non-finite values
template <bool ONES, int WT, int WZ>
__device__ inline void accumulate1(const PlaneW<WT, WZ> &P, const float (&val)[WT][WZ],
                                   float &acc0, float &acc1) {
    float s0 = 0.f;
#pragma unroll
    for (int k = 0; k < WT; ++k) {
        float r0 = 0.f;
#pragma unroll
        for (int j = 0; j < WZ; ++j) r0 = fmaf(P.wz[j], val[k][j], r0);
        s0 = fmaf(P.wt[k], r0, s0);
    }
    acc0 += P.in ? P.sc * s0 : 0.f;
    if (ONES) {   // P(1): the same weights on a volume of ones
        float swt = 0.f, swz = 0.f;
#pragma unroll
        for (int k = 0; k < WT; ++k) swt += P.wt[k];
#pragma unroll
        for (int k = 0; k < WZ; ++k) swz += P.wz[k];
        acc1 = fmaf(P.sc, swt * swz, acc1);
    }
}

template <int WT, int WZ>
__device__ inline void accumulate2(const PlaneW<WT, WZ> &P, const float2 (&val)[WT][WZ],
                                   float &acc0, float &acc1) {
    float s0 = 0.f, s1 = 0.f;
#pragma unroll
    for (int k = 0; k < WT; ++k) {
        float r0 = 0.f, r1 = 0.f;
#pragma unroll
        for (int j = 0; j < WZ; ++j) {
            r0 = fmaf(P.wz[j], val[k][j].x, r0);
            r1 = fmaf(P.wz[j], val[k][j].y, r1);
        }
        s0 = fmaf(P.wt[k], r0, s0);
        s1 = fmaf(P.wt[k], r1, s1);
    }
    acc0 += P.in ? P.sc * s0 : 0.f;
    acc1 += P.in ? P.sc * s1 : 0.f;
}

template <int MODE, int WT, int WZ>
__global__ void __launch_bounds__(RAYS *MAXSPLIT) k_project(GeomArgs g, double c, double s,
                                                            ProjArgs a, int nsplit) {
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int ia = blockIdx.x * RAYS + tx;
    const int ib = blockIdx.y;
    const bool live = ia < g.na;

    __shared__ Ray rays[RAYS];
    if (ty == 0 && live) ray_setup(g, c, s, ia, ib, rays[tx]);
    __syncthreads();

    float acc0 = 0.f, acc1 = 0.f;
    if (live) {
        const Ray R = rays[tx];
        const bool xb = R.xbranch != 0;
        const int np = xb ? g.nx : g.ny;
        const int nt = xb ? g.ny : g.nx;
        const int nz = g.nz;
        const unsigned sz = (unsigned)g.nx * (unsigned)g.ny;
        // strides of the layout this ray reads (see the header comment)
        const bool useT = xb && (MODE == PROJ_MOTION ? a.pairT != nullptr : a.vol0T != nullptr);
        const unsigned sp = xb ? (useT ? (unsigned)g.ny : 1u) : (unsigned)g.nx;
        const unsigned st = xb ? (useT ? 1u : (unsigned)g.nx) : 1u;
        const float *__restrict__ v = useT ? a.vol0T : a.vol0;
        const float2 *__restrict__ v2 = useT ? a.pairT : a.pair;
        const float p0 = 0.5f - 0.5f * (float)np;   // (float)ip + 0.5 - n/2 (:52, :105)
        const float ntf = (float)nt, nzf = (float)nz;
        const int chunk = (np + nsplit - 1) / nsplit;
        const int ip_lo = ty * chunk;
        const int ip_hi = min(np, ip_lo + chunk);
        // Two planes per iteration: both footprints first, then all their gathers (into distinct
        // registers, so they are in flight together), then the weighted sums.
        int ip = ip_lo;
        for (; ip + 1 < ip_hi; ip += 2) {
            PlaneW<WT, WZ> A, B;
            plane_weights(R, ip, p0, nt, nz, ntf, nzf, sp, st, sz, A);
            plane_weights(R, ip + 1, p0, nt, nz, ntf, nzf, sp, st, sz, B);
            if (MODE == PROJ_MOTION) {
                float2 va[WT][WZ], vb[WT][WZ];
                gather(v2, A, va);
                gather(v2, B, vb);
                accumulate2(A, va, acc0, acc1);
                accumulate2(B, vb, acc0, acc1);
            } else {
                float va[WT][WZ], vb[WT][WZ];
                gather(v, A, va);
                gather(v, B, vb);
                accumulate1<MODE == PROJ_IMAGE>(A, va, acc0, acc1);
                accumulate1<MODE == PROJ_IMAGE>(B, vb, acc0, acc1);
            }
        }
        if (ip < ip_hi) {
            PlaneW<WT, WZ> A;
            plane_weights(R, ip, p0, nt, nz, ntf, nzf, sp, st, sz, A);
            if (MODE == PROJ_MOTION) {
                float2 va[WT][WZ];
                gather(v2, A, va);
                accumulate2(A, va, acc0, acc1);
            } else {
                float va[WT][WZ];
                gather(v, A, va);
                accumulate1<MODE == PROJ_IMAGE>(A, va, acc0, acc1);
            }
        }
    }

    // Combine the plane chunks (fixed order: deterministic).
    __shared__ float red0[MAXSPLIT][RAYS];
    __shared__ float red1[MAXSPLIT][RAYS];
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
        const float pm = a.meas[id];
        const float r = pm - acc0;
        a.rp[id] = make_float2(r, acc1);
        if (a.resid_copy) a.resid_copy[id] = r;
        pp = (double)pm * (double)pm;
        rr = (double)r * (double)r;
    }
    for (int off = 32; off > 0; off >>= 1) {
        pp += __shfl_xor(pp, off, 64);
        rr += __shfl_xor(rr, off, 64);
    }
    if (tx == 0) {   // finished by the next kernel on the stream (NormFinish)
        const int bid = blockIdx.y * gridDim.x + blockIdx.x;
        a.partials[2 * bid] = pp;
        a.partials[2 * bid + 1] = rr;
    }
}

// Footprints wider than the fixed windows (coarse voxels against a fine detector): the same
// weights with data-dependent loops, as in the reference (src/cu_projection.cu:81-89).
template <int MODE>
__global__ void __launch_bounds__(RAYS *MAXSPLIT) k_project_wide(GeomArgs g, double c, double s,
                                                                 ProjArgs a, int nsplit);

int pick_split(const Geom &g) {
    const long long tiles = (long long)((g.na + RAYS - 1) / RAYS) * g.nb;
    int s = 1;
    while (s < MAXSPLIT && tiles * s < 4096) s *= 2;
    const int np = g.nx > g.ny ? g.nx : g.ny;
    while (s > 1 && np / s < 8) s /= 2;
    return s;
}

template <int MODE>
int launch_mode(const Geom &g, const ViewTrig &t, const ViewInfo &vi, const ProjArgs &a,
                hipStream_t st, int split) {
    const GeomArgs ga{g.nx, g.ny, g.nz, g.na, g.nb, g.da, g.db, g.ai, g.bi, g.SO, g.SD};
    const int s = split > 0 ? split : pick_split(g);
    const dim3 block(RAYS, s);
    const dim3 grid((g.na + RAYS - 1) / RAYS, g.nb);
    const int code = (vi.wt <= 2 ? 0 : vi.wt == 3 ? 1 : 2) * 3 + (vi.wz <= 2 ? 0 : vi.wz == 3 ? 1 : 2);
    switch (code) {
        case 0: hipLaunchKernelGGL((k_project<MODE, 2, 2>), grid, block, 0, st, ga, t.c, t.s, a, s); break;
        case 1: hipLaunchKernelGGL((k_project<MODE, 2, 3>), grid, block, 0, st, ga, t.c, t.s, a, s); break;
        case 3: hipLaunchKernelGGL((k_project<MODE, 3, 2>), grid, block, 0, st, ga, t.c, t.s, a, s); break;
        case 4: hipLaunchKernelGGL((k_project<MODE, 3, 3>), grid, block, 0, st, ga, t.c, t.s, a, s); break;
        default: hipLaunchKernelGGL((k_project_wide<MODE>), grid, block, 0, st, ga, t.c, t.s, a, s); break;
    }
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

template <int MODE>
__global__ void __launch_bounds__(RAYS *MAXSPLIT) k_project_wide(GeomArgs g, double c, double s,
                                                                 ProjArgs a, int nsplit) {
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int ia = blockIdx.x * RAYS + tx;
    const int ib = blockIdx.y;
    const bool live = ia < g.na;
    float acc0 = 0.f, acc1 = 0.f;
    if (live) {
        Ray R;
        ray_setup(g, c, s, ia, ib, R);
        const bool xb = R.xbranch != 0;
        const int np = xb ? g.nx : g.ny;
        const int nt = xb ? g.ny : g.nx;
        const bool useT = xb && (MODE == PROJ_MOTION ? a.pairT != nullptr : a.vol0T != nullptr);
        const long long sp = xb ? (useT ? g.ny : 1) : g.nx;
        const long long st = xb ? (useT ? 1 : g.nx) : 1;
        const long long sz = (long long)g.nx * g.ny;
        const float *v = useT ? a.vol0T : a.vol0;
        const float2 *v2 = useT ? a.pairT : a.pair;
        const float p0 = 0.5f - 0.5f * (float)np;
        const int chunk = (np + nsplit - 1) / nsplit;
        const int ip_lo = ty * chunk, ip_hi = min(np, ip_lo + chunk);
        for (int ip = ip_lo; ip < ip_hi; ++ip) {
            PlaneFP F;
            if (!plane_fp(R, (float)ip + p0, nt, g.nz, F)) continue;
            const int Tl = max(F.T0, 0), Th = min(nt - 1, (int)fminf(floorf(F.t2), (float)nt));
            const int Zl = max(F.Z0, 0);
            const int Zh = min(g.nz - 1, (int)fminf(floorf(F.z2), (float)g.nz));
            for (int i = Tl; i <= Th; ++i) {
                const float w1 = (fminf(i + 1.f, F.t2) - fmaxf((float)i, F.t1)) * F.it;
                for (int k = Zl; k <= Zh; ++k) {
                    const float w = w1 * (fminf(k + 1.f, F.z2) - fmaxf((float)k, F.z1)) * F.iz;
                    const long long idx = ip * sp + i * st + k * sz;
                    if (MODE == PROJ_MOTION) {
                        const float2 q = v2[idx];
                        acc0 += w * q.x;
                        acc1 += w * q.y;
                    } else {
                        acc0 += w * v[idx];
                        if (MODE == PROJ_IMAGE) acc1 += w;
                    }
                }
            }
        }
    }
    __shared__ float red0[MAXSPLIT][RAYS];
    __shared__ float red1[MAXSPLIT][RAYS];
    red0[ty][tx] = acc0;
    red1[ty][tx] = acc1;
    __syncthreads();
    if (ty != 0) return;
    for (int k = 1; k < nsplit; ++k) {
        acc0 += red0[k][tx];
        acc1 += red1[k][tx];
    }
    // hand the sums to the same epilogue through a tiny second launch-free path: write them
    // where the SART epilogue expects its inputs (the projection itself and the second sum)
    const long long id = (long long)ib * g.na + ia;
    if (!live) return;
    if (MODE == PROJ_PLAIN) {
        a.out[id] = acc0;
    } else {
        a.rp[id] = make_float2(acc0, acc1);   // raw sums; finished by k_sart_finish
    }
}

// Residual, normaliser and norms from raw (P x, n) sums, for the wide-footprint path.
__global__ void __launch_bounds__(256) k_sart_finish(ProjArgs a, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    double pp = 0.0, rr = 0.0;
    if (i < n) {
        const float2 v = a.rp[i];
        const float pm = a.meas[i];
        const float r = pm - v.x;
        a.rp[i] = make_float2(r, v.y);
        if (a.resid_copy) a.resid_copy[i] = r;
        pp = (double)pm * pm;
        rr = (double)r * r;
    }
    for (int off = 32; off > 0; off >>= 1) {
        pp += __shfl_xor(pp, off, 64);
        rr += __shfl_xor(rr, off, 64);
    }
    __shared__ double sp[4], sr[4];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) {
        sp[w] = pp;
        sr[w] = rr;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    a.partials[2 * blockIdx.x] = sp[0] + sp[1] + sp[2] + sp[3];
    a.partials[2 * blockIdx.x + 1] = sr[0] + sr[1] + sr[2] + sr[3];
}

__global__ void __launch_bounds__(64) k_norm_finish(NormFinish nf) {
    finish_norm_wave(nf, threadIdx.x);
}

}  // namespace

int launch_norm_finish(const NormFinish &nf, hipStream_t st) {
    hipLaunchKernelGGL(k_norm_finish, dim3(1), dim3(64), 0, st, nf);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

ViewInfo view_info_of(const Geom &g, const ViewTrig &t) {
    const GeomArgs ga{g.nx, g.ny, g.nz, g.na, g.nb, g.da, g.db, g.ai, g.bi, g.SO, g.SD};
    return view_info(ga, t.c, t.s);
}

int projector_partials(const Geom &g) {
    const int a = ((g.na + RAYS - 1) / RAYS) * g.nb;
    const int b = (int)((g.ndet() + 255) / 256);
    return a > b ? a : b;
}

int launch_projector(const Geom &g, const ViewTrig &t, const ViewInfo &vi, ProjMode mode,
                     const ProjArgs &a, hipStream_t st, int *nparts, int split) {
    const bool wide = !(vi.wt <= 3 && vi.wz <= 3);
    if (split < 0 || split > MAXSPLIT || (split & (split - 1)))
        return fail(MCSART_EINVAL, "projector split must be 0 (automatic) or a power of two <= " +
                                       std::to_string(MAXSPLIT));
    int rc;
    switch (mode) {
        case PROJ_PLAIN: rc = launch_mode<PROJ_PLAIN>(g, t, vi, a, st, split); break;
        case PROJ_IMAGE: rc = launch_mode<PROJ_IMAGE>(g, t, vi, a, st, split); break;
        default: rc = launch_mode<PROJ_MOTION>(g, t, vi, a, st, split); break;
    }
    if (nparts) *nparts = ((g.na + RAYS - 1) / RAYS) * g.nb;
    if (rc || !wide || mode == PROJ_PLAIN) return rc;
    const int n = (int)g.ndet();
    hipLaunchKernelGGL(k_sart_finish, dim3((n + 255) / 256), dim3(256), 0, st, a, n);
    MCSART_CHECK_LAUNCH();
    if (nparts) *nparts = (n + 255) / 256;
    return MCSART_OK;
}

}  // namespace mcsart

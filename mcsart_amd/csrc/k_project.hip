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
// (g, Q) are interleaved as float2, so each gather is one 8-byte load for both projections, and
// P(q) = qscale P(Q) (q = Q (-vd - fd): Q is written once per sweep, g per view).
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
        a.rp[id] = make_float2(r, MODE == PROJ_MOTION ? acc1 * a.qscale : acc1);
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

// ---- rows kernel ------------------------------------------------------------------------------
//
// A lane walks R rays of one detector column (rows ib0 .. ib0+R-1).  Their xy directions are the
// same, so the transverse half of every plane's work (bounds, window, weights) is done once and
// shared by the R rays; only the axial half is per ray.  The window of WT (WZ) voxels starts at
// floor(lo) clamped into [0, n - WT]: every voxel it names exists, so no per-voxel clamp or mask
// is needed, and a voxel of the window outside the footprint gets weight zero from the overlap
// formula itself.  Overlaps are differences of the clamped boundary F(b) = min(max(b, lo), hi):
// w_k = F(T0+k+1) - F(T0+k) = min(T0+k+1, hi) - max(T0+k, lo) whenever the cell meets the
// footprint (src/cu_projection.cu:83,86), and 0 otherwise.  Plane offsets use 24-bit multiplies
// and gathers 32-bit byte offsets from one base (launch requires nx*ny < 2^24 and both layouts
// within 4 GiB of the lower one).  Requires the xy-transposed layout for x-branch rays.  Where
// adjacent detector rows' footprints lie less than a voxel apart (rows_share), the R rays also
// share their z-rows: WZ + 1 gathers per plane for the lane (lane_values).
// med3(x, lo, hi) = min(max(x, lo), hi) when lo <= hi, in one instruction (v_med3_f32).
__device__ inline float clampf(float x, float lo, float hi) {
    return __builtin_amdgcn_fmed3f(x, lo, hi);
}

// Weights of the W-voxel window over [lo, hi], 0 <= lo <= hi <= n (bounds already clamped into
// the volume; lo == hi, all weights 0, for a plane outside it).  The window covers the
// footprint, hi <= f0 + W (W is the most cells any footprint of the view spans, view_info; at
// the volume's top edge f0 + W = n >= hi), so the last boundary min(f0 + W, hi) is hi itself.
template <int W>
__device__ inline void window(float lo, float hi, float top, float (&w)[W], float &sw, float &f0) {
    f0 = clampf(floorf(lo), 0.f, top);   // window start, inside [0, n - W]
    float F = lo;   // F(f0) = max(f0, lo) = lo: 0 <= lo, so f0 <= floor(lo) <= lo
#pragma unroll
    for (int k = 0; k < W; ++k) {
        // interior boundaries clamp to [lo, hi]; the last one is hi
        const float Fn = (k + 1 < W) ? clampf(f0 + (float)(k + 1), lo, hi) : hi;
        w[k] = Fn - F;
        F = Fn;
    }
    sw = hi - lo;
}

// Footprint of one plane for the R rays of a lane.  The bounds are clamped into the volume
// first, so a plane outside it gets weights of exactly zero from the window itself (no mask and
// no select: every gather of its window is in the volume, and zero times a finite value adds
// exactly nothing); inside, clamping changes no weight, because the window never reaches past
// [0, n].  The scale uses the unclamped widths (src/cu_projection.cu:83,86).
template <int WT, int WZ, int R>
struct PlaneFoot {
    float wt[WT], sw;
    float wz[R][WZ], swz[R], sc[R];
    unsigned off[R];   // byte offset of ray r's window corner
    unsigned g[R];     // ray r's first z-row
    unsigned base;     // byte offset of the window's first cell in z-row 0
};

template <int WT, int WZ, int R, int LOGES>
__device__ inline void plane_foot(const RayT &T, const RayZ (&Z)[R], float p, unsigned pofs,
                                  unsigned sz, float ntf, float nzf, float ttop, float ztop,
                                  PlaneFoot<WT, WZ, R> &P) {
    const float t1 = fmaf(T.rlo, p, T.Klo), t2 = fmaf(T.rhi, p, T.Khi);
    float f0;
    window<WT>(clampf(t1, 0.f, ntf), clampf(t2, 0.f, ntf), ttop, P.wt, P.sw, f0);
    const float width = t2 - t1;
    const unsigned t0 = (unsigned)(int)f0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const float z1 = fmaf(Z[r].rlo, p, Z[r].Klo), z2 = fmaf(Z[r].rhi, p, Z[r].Khi);
        float g0;
        window<WZ>(clampf(z1, 0.f, nzf), clampf(z2, 0.f, nzf), ztop, P.wz[r], P.swz[r], g0);
        P.sc[r] = __builtin_amdgcn_rcpf(width * (z2 - z1));
        P.g[r] = (unsigned)(int)g0;
        P.off[r] = ((__umul24(P.g[r], sz) + t0) << LOGES) + pofs;
    }
    P.base = (t0 << LOGES) + pofs;
}

// The window's values: WT consecutive elements on each of WZ z-rows, from a wave-uniform base
// and 32-bit byte offsets (one scalar base, one vector offset per row).
template <int WT, int WZ, typename E>
__device__ inline void window_values(const char *__restrict__ vb, unsigned off, unsigned szB,
                                     E (&val)[WT][WZ]) {
#pragma unroll
    for (int j = 0; j < WZ; ++j) {
        const char *row = vb + (off + (unsigned)j * szB);
#pragma unroll
        for (int k = 0; k < WT; ++k) val[k][j] = *(const E *)(row + k * sizeof(E));
    }
}

// The values of the R rays' windows, val[r][k][j] = cell k of z-row j of ray r's window.
// Without SH every ray gathers its own window.  With SH (rows_share holds for the view) the
// rays' windows start at most one z-row apart, so rows gmin .. gmin + WZ - 1 and gmax + WZ - 1
// hold all of them: the lane gathers those WZ + 1 rows once (instead of WZ per ray) and each ray
// selects its WZ rows, shifted by one where its window starts above gmin.  The values are the
// same either way, so are the sums, bit for bit, whatever the values are.
template <int WT, int WZ, int R, bool SH, typename E>
__device__ inline void lane_values(const char *__restrict__ vb, const PlaneFoot<WT, WZ, R> &P,
                                   unsigned sz, unsigned szB, E (&val)[R][WT][WZ]) {
    if (!SH || R == 1) {
#pragma unroll
        for (int r = 0; r < R; ++r) window_values(vb, P.off[r], szB, val[r]);
        return;
    }
    constexpr int LOGES = sizeof(E) == 8 ? 3 : 2;
    // g_r is monotone in r (adjacent rows' footprints are one row spacing apart)
    const unsigned gmin = min(P.g[0], P.g[R - 1]), gmax = max(P.g[0], P.g[R - 1]);
    E rows[WT][WZ + 1];
    const unsigned o0 = (__umul24(gmin, sz) << LOGES) + P.base;   // row gmin
#pragma unroll
    for (int j = 0; j <= WZ; ++j) {
        // rows gmin + j, then gmax + WZ - 1 (gmax is gmin or gmin + 1)
        const unsigned o = j < WZ ? o0 + (unsigned)j * szB
                                  : o0 + (unsigned)(WZ - 1) * szB + (gmax != gmin ? szB : 0u);
#pragma unroll
        for (int k = 0; k < WT; ++k)
            rows[k][j] = *(const E *)(vb + o + (unsigned)(k * sizeof(E)));
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const bool up = P.g[r] != gmin;
#pragma unroll
        for (int k = 0; k < WT; ++k)
#pragma unroll
            for (int j = 0; j < WZ; ++j) val[r][k][j] = up ? rows[k][j + 1] : rows[k][j];
    }
}

// PL planes from plane coordinate p: all footprints, then all gathers (in flight together),
// then the weighted sums.  SH: the lane's rays share their z-rows (lane_values).
template <int MODE, int WT, int WZ, int R, int PL, bool SH = false>
__device__ inline void rows_step(const char *__restrict__ vb, const RayT &T, const RayZ (&Z)[R],
                                 float p, unsigned pofs, unsigned spB, unsigned sz, float ntf,
                                 float nzf, float ttop, float ztop, float (&acc0)[R],
                                 float (&acc1)[R]) {
    constexpr int LOGES = MODE == PROJ_MOTION ? 3 : 2;
    const unsigned szB = sz << LOGES;
    PlaneFoot<WT, WZ, R> P[PL];
#pragma unroll
    for (int q = 0; q < PL; ++q)
        plane_foot<WT, WZ, R, LOGES>(T, Z, p + (float)q, pofs + (unsigned)q * spB, sz, ntf, nzf,
                                     ttop, ztop, P[q]);
    if (MODE == PROJ_MOTION) {
        float2 val[PL][R][WT][WZ];
#pragma unroll
        for (int q = 0; q < PL; ++q) lane_values<WT, WZ, R, SH>(vb, P[q], sz, szB, val[q]);
#pragma unroll
        for (int q = 0; q < PL; ++q)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                float s0 = 0.f, s1 = 0.f;
#pragma unroll
                for (int k = 0; k < WT; ++k) {
                    float u0 = 0.f, u1 = 0.f;
#pragma unroll
                    for (int j = 0; j < WZ; ++j) {
                        u0 = fmaf(P[q].wz[r][j], val[q][r][k][j].x, u0);
                        u1 = fmaf(P[q].wz[r][j], val[q][r][k][j].y, u1);
                    }
                    s0 = fmaf(P[q].wt[k], u0, s0);
                    s1 = fmaf(P[q].wt[k], u1, s1);
                }
                acc0[r] = fmaf(P[q].sc[r], s0, acc0[r]);
                acc1[r] = fmaf(P[q].sc[r], s1, acc1[r]);
            }
    } else {
        float val[PL][R][WT][WZ];
#pragma unroll
        for (int q = 0; q < PL; ++q) lane_values<WT, WZ, R, SH>(vb, P[q], sz, szB, val[q]);
#pragma unroll
        for (int q = 0; q < PL; ++q)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                float s0 = 0.f;
#pragma unroll
                for (int k = 0; k < WT; ++k) {
                    float u0 = 0.f;
#pragma unroll
                    for (int j = 0; j < WZ; ++j) u0 = fmaf(P[q].wz[r][j], val[q][r][k][j], u0);
                    s0 = fmaf(P[q].wt[k], u0, s0);
                }
                acc0[r] = fmaf(P[q].sc[r], s0, acc0[r]);
                if (MODE == PROJ_IMAGE)   // P(1): the same weights on a volume of ones
                    acc1[r] = fmaf(P[q].sc[r], P[q].sw * P[q].swz[r], acc1[r]);
            }
    }
}

template <int MODE, int WT, int WZ, int R, bool SQ, bool SH>
__global__ void __launch_bounds__(RAYS * 16) k_project_rows(GeomArgs g, double c, double s,
                                                           ProjArgs a, int nsplit) {
    const int tx = threadIdx.x, ty = threadIdx.y;
    // logical tile: detector-row bands stay on one XCD (their rays share the same z-slab)
    const int tile = xcd_order(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int bxl = tile % gridDim.x, byl = tile / gridDim.x;
    const int ia = bxl * RAYS + tx;
    const int ib0 = byl * R;
    const bool live = ia < g.na;

    __shared__ RayT rt[RAYS];
    __shared__ RayZ rz[R][RAYS];
    if (ty == 0 && live) {
        const RayDir d = ray_dir(g, c, s, ia);
        ray_t_setup(g, d, rt[tx]);
#pragma unroll
        for (int r = 0; r < R; ++r) ray_z_setup(g, d, min(ib0 + r, g.nb - 1), rz[r][tx]);
    }
    __syncthreads();

    float acc0[R], acc1[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc0[r] = acc1[r] = 0.f;
    if (live) {
        const RayT T = rt[tx];
        RayZ Z[R];
#pragma unroll
        for (int r = 0; r < R; ++r) Z[r] = rz[r][tx];
        // x-branch rays read the transposed layout, y-branch rays the natural one; both have
        // the same index form (plane stride nt) and are addressed from one base, so every
        // gather is that wave-uniform base plus a 32-bit byte offset
        constexpr int LOGES = MODE == PROJ_MOTION ? 3 : 2;
        const bool xb = T.xbranch != 0;
        // SQ (nx == ny): both branches walk the same number of planes, a kernel argument, so the
        // plane loop's bounds and counter stay scalar
        const int np = SQ ? g.nx : (xb ? g.nx : g.ny);
        const int nt = SQ ? g.nx : (xb ? g.ny : g.nx);
        const unsigned sz = (unsigned)g.nx * (unsigned)g.ny;
        const unsigned spB = (unsigned)nt << LOGES;   // plane stride in bytes
        const float p0 = 0.5f - 0.5f * (float)np;   // (float)ip + 0.5 - n/2 (:52, :105)
        const float ntf = (float)nt, nzf = (float)g.nz;
        const float ttop = (float)(nt - WT), ztop = (float)(g.nz - WZ);
        const int chunk = (np + nsplit - 1) / nsplit;
        // a wavefront is one row of the block (64 lanes), so ty is uniform across it
        const int ip_lo = __builtin_amdgcn_readfirstlane(ty) * chunk;
        const int ip_hi = min(np, ip_lo + chunk);
        // PL planes per step, all their gathers in flight together, the last ones one at a time.
        // One plane when a lane walks two or more rays: their gathers already overlap, and a
        // second plane's registers cost occupancy (R = 2 image mode: 74 VGPRs, 6 waves per SIMD,
        // against 53 and 8 with one plane; measured 52.0 -> 50.6 us at 256x256x128) -- except
        // for two rays sharing their z-rows, which hold fewer values: 68 VGPRs and 7 waves with
        // two planes, 46.2 -> 44.1 us.  p and the plane's byte offset are carried: p is a small
        // integer + 0.5, exact in float.
        constexpr int PL = (R == 1 || (SH && R == 2)) ? 2 : 1;
        int ip = ip_lo;
        float p = (float)ip_lo + p0;
        unsigned pofs = (xb ? a.roffT : a.roffN) + (unsigned)ip_lo * spB;
        for (; ip + PL <= ip_hi; ip += PL, p += (float)PL, pofs += PL * spB)
            rows_step<MODE, WT, WZ, R, PL, SH>(a.rbase, T, Z, p, pofs, spB, sz, ntf, nzf, ttop, ztop,
                                          acc0, acc1);
        for (; PL > 1 && ip < ip_hi; ++ip, p += 1.f, pofs += spB)
            rows_step<MODE, WT, WZ, R, 1, SH>(a.rbase, T, Z, p, pofs, spB, sz, ntf, nzf, ttop, ztop,
                                          acc0, acc1);
    }

    // Combine the plane chunks (fixed order: deterministic).
    extern __shared__ float red[];   // [2][nsplit][R][RAYS]
    if (nsplit > 1) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            red[((0 * nsplit + ty) * R + r) * RAYS + tx] = acc0[r];
            red[((1 * nsplit + ty) * R + r) * RAYS + tx] = acc1[r];
        }
        __syncthreads();
        if (ty == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                acc0[r] = red[(0 * R + r) * RAYS + tx];
                acc1[r] = red[((1 * nsplit) * R + r) * RAYS + tx];
                for (int k = 1; k < nsplit; ++k) {
                    acc0[r] += red[((0 * nsplit + k) * R + r) * RAYS + tx];
                    acc1[r] += red[((1 * nsplit + k) * R + r) * RAYS + tx];
                }
            }
        }
    }
    if (ty != 0) return;

    if (MODE == PROJ_PLAIN) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (live && ib0 + r < g.nb) a.out[(long long)(ib0 + r) * g.na + ia] = acc0[r];
        return;
    }
    // SART epilogue: residual, normaliser, norms.
    double pp = 0.0, rr = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (!live || ib0 + r >= g.nb) continue;
        const long long id = (long long)(ib0 + r) * g.na + ia;
        const float pm = a.meas[id];
        const float res = pm - acc0[r];
        a.rp[id] = make_float2(res, MODE == PROJ_MOTION ? acc1[r] * a.qscale : acc1[r]);
        if (a.resid_copy) a.resid_copy[id] = res;
        pp += (double)pm * (double)pm;
        rr += (double)res * (double)res;
    }
    for (int o = 32; o > 0; o >>= 1) {
        pp += __shfl_xor(pp, o, 64);
        rr += __shfl_xor(rr, o, 64);
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

// Rows kernel shape: R rays per lane and S plane chunks per workgroup.  Automatic choice: the
// most rays per lane for which chunks of at least 16 planes still give >= 4096 wavefronts (the
// shapes the timing autotuner picks on 128^3 / 256x256x128 scans: R = 1, S = 8 / R = 4, S = 8).
void pick_rows(const Geom &g, int rows, int split, int &R, int &S) {
    const int np = g.nx > g.ny ? g.nx : g.ny;
    const long long cols = (g.na + RAYS - 1) / RAYS;
    auto split_for = [&](int r) {
        const long long tiles = cols * ((g.nb + r - 1) / r);
        int sp = 1;
        while (sp < 16 && tiles * sp < 4096 && np / (2 * sp) >= 16) sp *= 2;
        return sp;
    };
    auto waves = [&](int r) { return cols * ((g.nb + r - 1) / r) * split_for(r); };
    R = rows;
    if (!R) {
        R = 1;
        for (int r = 4; r >= 2; r /= 2)
            if (waves(r) >= 4096) {
                R = r;
                break;
            }
    }
    S = split > 0 ? split : split_for(R);
}

template <int MODE, int WT, int WZ, int R, bool SH>
void launch_rows(const GeomArgs &ga, const ViewTrig &t, const ProjArgs &a, hipStream_t st, int S) {
    const dim3 block(RAYS, S);
    const dim3 grid((ga.na + RAYS - 1) / RAYS, (ga.nb + R - 1) / R);
    const size_t lds = S > 1 ? sizeof(float) * 2 * S * R * RAYS : 0;
    if (ga.nx == ga.ny)
        hipLaunchKernelGGL((k_project_rows<MODE, WT, WZ, R, true, SH>), grid, block, lds, st, ga,
                           t.c, t.s, a, S);
    else
        hipLaunchKernelGGL((k_project_rows<MODE, WT, WZ, R, false, SH>), grid, block, lds, st, ga,
                           t.c, t.s, a, S);
}

// R rays per lane, sharing their z-rows where rows_share holds for the view (vi)
template <int MODE, int WT, int WZ>
void launch_rows_r(const GeomArgs &ga, const ViewTrig &t, const ViewInfo &vi, const ProjArgs &a,
                   hipStream_t st, int R, int S) {
    if constexpr (MODE == PROJ_MOTION) {
        // the motion mode gathers (g, Q) pairs, twice the registers per footprint: one ray per
        // lane (two sharing their z-rows measured the same at 256x256x128, 51.5 us)
        launch_rows<MODE, WT, WZ, 1, false>(ga, t, a, st, S);
    } else {
        if (R == 1) launch_rows<MODE, WT, WZ, 1, false>(ga, t, a, st, S);
        else if (R == 2 && rows_share(vi, 2)) launch_rows<MODE, WT, WZ, 2, true>(ga, t, a, st, S);
        else if (R == 2) launch_rows<MODE, WT, WZ, 2, false>(ga, t, a, st, S);
        else if (rows_share(vi, 4)) launch_rows<MODE, WT, WZ, 4, true>(ga, t, a, st, S);
        else launch_rows<MODE, WT, WZ, 4, false>(ga, t, a, st, S);
    }
}

template <int MODE>
int launch_mode_rows(const Geom &g, const ViewTrig &t, const ViewInfo &vi, const ProjArgs &a,
                     hipStream_t st, int R, int S) {
    const GeomArgs ga{g.nx, g.ny, g.nz, g.na, g.nb, g.da, g.db, g.ai, g.bi, g.SO, g.SD};
    const int code = (vi.wt <= 2 ? 0 : 1) * 2 + (vi.wz <= 2 ? 0 : 1);
    switch (code) {
        case 0: launch_rows_r<MODE, 2, 2>(ga, t, vi, a, st, R, S); break;
        case 1: launch_rows_r<MODE, 2, 3>(ga, t, vi, a, st, R, S); break;
        case 2: launch_rows_r<MODE, 3, 2>(ga, t, vi, a, st, R, S); break;
        default: launch_rows_r<MODE, 3, 3>(ga, t, vi, a, st, R, S); break;
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
    } else {   // raw sums, finished by k_sart_finish
        a.rp[id] = make_float2(acc0, MODE == PROJ_MOTION ? acc1 * a.qscale : acc1);
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
                     const ProjArgs &a, hipStream_t st, int *nparts, int split, int rows) {
    const bool wide = !(vi.wt <= 3 && vi.wz <= 3);
    if (split < 0 || split > 16 || (split & (split - 1)))
        return fail(MCSART_EINVAL, "projector split must be 0 (automatic) or a power of two <= 16");
    if (rows != 0 && rows != 1 && rows != 2 && rows != 4)
        return fail(MCSART_EINVAL, "projector rows must be 0 (automatic), 1, 2 or 4");
    // The rows kernel needs contiguous footprints for every ray (the transposed layout when the
    // view has x-branch rays), windows that fit in the volume and 24-bit plane offsets.
    const uintptr_t un = (uintptr_t)(mode == PROJ_MOTION ? (const void *)a.pair : a.vol0);
    const uintptr_t ut = (uintptr_t)(mode == PROJ_MOTION ? (const void *)a.pairT : a.vol0T);
    const bool have_n = un != 0, have_t = ut != 0;
    // one base for both layouts, each within 32-bit byte offsets of it
    const uintptr_t base = !have_n ? ut : !have_t ? un : (un < ut ? un : ut);
    const unsigned long long offN = have_n ? un - base : 0ull, offT = have_t ? ut - base : 0ull;
    const unsigned long long vbytes = (unsigned long long)g.nvox() * (mode == PROJ_MOTION ? 8 : 4);
    const bool rows_ok = !wide && (!vi.has_x || have_t) && (!vi.has_y || have_n) &&
                         g.nx >= vi.wt && g.ny >= vi.wt && g.nz >= vi.wz &&
                         (long long)g.nx * g.ny < (1LL << 24) &&
                         (offN > offT ? offN : offT) + vbytes <= (1ull << 32);
    int rc;
    if (rows_ok) {
        ProjArgs ar = a;
        ar.rbase = (const char *)base;
        ar.roffN = (unsigned)offN;
        ar.roffT = (unsigned)offT;
        int R, S;
        pick_rows(g, rows, split, R, S);
        // The motion mode gathers (g, Q) pairs: with more rays per lane its footprints' values
        // hold twice the registers and the lanes wait on them.  One ray per lane is its best
        // shape at 128^3 and 256x256x128 (11.3 / 57.0 us against 12.3 / 57.8 us with two rays
        // and one plane per step, 15.0 / 61.4 us with two planes), while the residual mode is
        // fastest with two or four at 256x256x128.
        if (mode == PROJ_MOTION) R = 1;
        switch (mode) {
            case PROJ_PLAIN: rc = launch_mode_rows<PROJ_PLAIN>(g, t, vi, ar, st, R, S); break;
            case PROJ_IMAGE: rc = launch_mode_rows<PROJ_IMAGE>(g, t, vi, ar, st, R, S); break;
            default: rc = launch_mode_rows<PROJ_MOTION>(g, t, vi, ar, st, R, S); break;
        }
        if (nparts) *nparts = ((g.na + RAYS - 1) / RAYS) * ((g.nb + R - 1) / R);
        return rc;
    }
    const int s8 = split > MAXSPLIT ? MAXSPLIT : split;
    switch (mode) {
        case PROJ_PLAIN: rc = launch_mode<PROJ_PLAIN>(g, t, vi, a, st, s8); break;
        case PROJ_IMAGE: rc = launch_mode<PROJ_IMAGE>(g, t, vi, a, st, s8); break;
        default: rc = launch_mode<PROJ_MOTION>(g, t, vi, a, st, s8); break;
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

// Voxel-driven footprint backprojector for gfx950.
//
// Reference: kernel_backprojection, src/cu_backprojection.cu:49-170 (one thread per voxel, every
// call allocates a CUDA array, copies the projection into it and builds a texture, :10-45).
//
// Work layout: a workgroup owns 16 x 16 (ix, iy) columns and a slab of ZS voxels along z; one
// lane owns a column.  A square tile's footprint on the detector is the narrowest, so its staged
// rectangle (below) is the smallest (measured: 256x256x128 image sweep 61 -> 51 us, motion sweep
// 87 -> 77 us against 32 x 8; 64 x 4 was slower than both).  The transverse footprint (a bounds, weights) of a column does not depend on z,
// so it is formed once per column in FP64 (footprint.hpp); each voxel only forms its axial bounds.
//
// LDS staging of the detector: the workgroup first reduces, in LDS, the rectangle of detector
// cells its voxels touch (per column the a range is exact; along z the b range is monotone, so
// the slab's first and last voxels bound it), then copies that rectangle of the projection (or of
// the interleaved (r, n) pairs) into LDS with coalesced row reads, and every voxel gathers its
// footprint cells from LDS instead of issuing scattered global loads.  A cell outside the staged
// rectangle (or a rectangle larger than the LDS budget) is read from global memory, so the
// staging changes only speed, never results.
//
// The SART mode reads (r, P1) interleaved as float2 and writes B(r)/B(P1) with NaN/Inf -> 0,
// fusing src/AUMISART.cu:212-218 (two backprojections and a division) into one pass.
#include <cmath>
#include <type_traits>

#include "footprint.hpp"
#include "internal.hpp"

namespace mcsart {
namespace {

constexpr int BX = 16, BY = 16;
// Staged cells per workgroup: 4094 float2 + the four rectangle bounds = 32 KiB exactly, so five
// workgroups share a CU's 160 KiB (five waves per SIMD; the copy below keeps the kernel under
// 96 VGPRs, which five waves also need).
constexpr int LDS_CELLS = 4094;

// Axial bounds of a column's voxels relative to an integer anchor: the four lines
// q*zc + c of src/cu_backprojection.cu:107-132 (slopes Rm/db, Rp/db, Rc/db twice) evaluated in
// FP64 at the slab's first voxel, then stepped in FP32 with values of a few cells at most.
struct BLines {
    int anchor;
    float l0, l1, l2, l3;   // line values at the slab's first voxel, minus anchor
    float q0, q1, q2;       // slopes per voxel (l2 and l3 share q2)
};

__device__ inline BLines blines_at(const Column &C, int nz, int iz0) {
    const double zc = (double)iz0 - (double)(nz / 2) + 0.5;   // :68
    const double L0 = C.qm * zc + C.C0, L1 = C.qp * zc + C.C0;
    const double L2 = C.qc * (zc + 0.5) + C.C0, L3 = C.qc * (zc - 0.5) + C.C0;
    const double lo = fmin(fmin(L0, L1), L3);
    BLines b;
    b.anchor = (int)fmin(fmax(floor(lo), -1048576.0), 1048576.0);
    const double a = (double)b.anchor;
    b.l0 = (float)(L0 - a);
    b.l1 = (float)(L1 - a);
    b.l2 = (float)(L2 - a);
    b.l3 = (float)(L3 - a);
    b.q0 = (float)C.qm;
    b.q1 = (float)C.qp;
    b.q2 = (float)C.qc;
    return b;
}

// Detector value at byte offset `off` from `base`: (r, P1) pairs in the SART mode, (proj, 0) in
// the plain mode.  32-bit byte offsets (detectors are limited to < 2^29 cells at launch) let the
// compiler address every load from one scalar base plus one vector register.
template <int MODE>
__device__ inline float2 det_at(const void *base, unsigned off) {
    if (MODE == BACK_PLAIN)
        return make_float2(*(const float *)((const char *)base + off), 0.f);
    return *(const float2 *)((const char *)base + off);
}

// The motion-sweep step fused into the ratio mode (MotionStep, internal.hpp): the update of the
// transposed accumulators and the next view's prep at the voxel itself, instead of storing u.
// The voxel's inputs are loaded before its backprojection (MotionIn), so their latency hides
// behind the row loop.
// Per voxel: `hit` when its footprint has cells and rows (otherwise u = 0 exactly), `need` when
// the next view's projector can read its g (MotionStep::next).  The accumulators are read for
// either, the constants for `need` only.
struct MotionIn {
    float sv, sf;
    float4 c;   // (Ga, Gb, H, x): one 16-byte load
    bool hit, need, poison;
};
// nc: the voxel's column's share of the next view's read-set test (cone_column).
__device__ inline MotionIn motion_load(const MotionStep &m, const ConeColumn &nc, int id, int iz,
                                       bool hit) {
    MotionIn in;
    in.hit = hit || !m.skip_zero;
    // the shell too: an earlier view's g there may have overflowed (the next projector reads
    // shell cells with zero weights, and 0 times Inf is NaN)
    in.need = !m.last && (cone_keep_at(m.next, nc, iz) || cone_shell_at(m.next, nc, iz));
    in.poison = !m.last && cone_poison_at(m.next, nc, iz);   // never read: NaN in mode 2
    const float2 a = (m.first || !(in.hit || in.need)) ? make_float2(0.f, 0.f) : m.acc[id];
    in.sv = a.x;
    in.sf = a.y;
    in.c = in.need ? m.cst[id] : make_float4(0.f, 0.f, 0.f, 0.f);
    return in;
}
__device__ inline void motion_step_at(const MotionStep &m, const MotionIn &in, int id, int idT,
                                      float u) {
    // u = 0 without a hit: sv + wv * 0 = sv, so the stored sums stand (the first view writes
    // them everywhere: it starts them)
    const float sv = in.sv + m.wv * u;
    const float sf = in.sf + m.wf * u;
    if (in.hit || m.first) m.acc[id] = make_float2(sv, sf);
    if (m.last) return;
    // only the g half of the pair: the next view's q = Q (-vd - fd) is the sweep's Q, scaled
    // inside the projector
    float gv = __builtin_nanf("");
    if (in.need) {
        const float4 c = in.c;   // (Ga, Gb, H, x)
        const float img0 = m.vd * c.x + m.fd * c.y + c.z * (m.vd * sv + m.fd * sf);
        gv = c.w - img0;
    } else if (!in.poison) {
        return;   // read by the next view with zero weights at most
    }
    if (m.pair) reinterpret_cast<float *>(m.pair)[2u * (unsigned)id] = gv;
    if (m.pairT) reinterpret_cast<float *>(m.pairT)[2u * (unsigned)idT] = gv;
}

// Epilogue of the ratio mode: store u, or the motion-sweep step, or the static sweep's add.
enum BackEpi { EPI_STORE = 0, EPI_MOTION = 1, EPI_ADD = 2 };

template <int MODE, int WA, bool STAGE, int EPI>
__global__ void __launch_bounds__(BX *BY) k_backproject(GeomArgs g, double c, double s,
                                                       const float *__restrict__ proj,
                                                       const float2 *__restrict__ rp,
                                                       float *__restrict__ out, int zslab,
                                                       double ida, double idb, NormFinish nf,
                                                       MotionStep mot, AddStep add, int xbands) {
    const int tid = threadIdx.y * BX + threadIdx.x;
    // the preceding projector's per-view norm (its partials are complete: kernel boundary); the
    // whole first wavefront (threads 0..63) of the first workgroup does it
    if (nf.out && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && tid < 64)
        finish_norm_wave(nf, tid);

    // logical tile, z fastest: each XCD walks whole columns of tiles (a band of detector
    // columns, every z), which keeps its share of the work even (work grows with |z|).  Handing
    // out tiles in hardware order instead (every XCD every column) measured slower at 256x256x128
    // (image sweep 47.8 -> 49.1 us, motion sweep 59.8 -> 60.3 us)
    const int tile = xcd_order(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z),
                               gridDim.x * gridDim.y * gridDim.z);
    // xbands (rays mostly along x: the view's fan is a band across y, and out-of-fan voxels cost
    // little): each XCD takes a band of tile columns instead of tile rows, so every XCD has its
    // share of the fan
    const int bzl = tile % gridDim.z, rest = tile / gridDim.z;
    const int bxl = xbands ? rest / gridDim.y : rest % gridDim.x;
    const int byl = xbands ? rest % gridDim.y : rest / gridDim.x;
    const int ix = bxl * BX + threadIdx.x;
    const int iy = byl * BY + threadIdx.y;
    const bool live = ix < g.nx && iy < g.ny;
    const int iz0 = bzl * zslab;
    const int iz1 = min(g.nz, iz0 + zslab);

    // ---- column: transverse footprint (a) once, in FP64, clamped so conversions stay in range
    Column C;
    int ia_lo = 0, acount = 0;
    float aminL = 0.f, amaxL = 0.f;
    BLines bl{};
    if (live) {
        column_setup_dev(g, c, s, ix, iy, ida, idb, C);
        ia_lo = (int)fmin(fmax(floor(C.Amin), 0.0), (double)g.na);   // :140,:150
        aminL = (float)(C.Amin - ia_lo);
        amaxL = (float)(C.Amax - ia_lo);
        const float alim = ((double)g.na < C.Amax) ? (float)(g.na - ia_lo) : amaxL;   // :150
        acount = alim > 0.f ? (int)ceilf(alim) : 0;   // cells ia_lo .. ia_lo + acount - 1
        bl = blines_at(C, g.nz, iz0);
    }
    // weights of the first WA cells (zero beyond acount); :156.  Window slots past the footprint
    // read its last cell again (weight zero): every value read is one the footprint uses.
    float wa[WA];
    int col[WA];
    const int last = max(ia_lo + acount - 1, ia_lo);
#pragma unroll
    for (int k = 0; k < WA; ++k) {
        wa[k] = (k < acount) ? fminf(k + 1.f, amaxL) - fmaxf((float)k, aminL) : 0.f;
        col[k] = min(ia_lo + k, last);
    }
    const bool wide = acount > WA;   // rare: handled by the generic loop below
    // the motion step's test of the next view's read set: its column part once per column
    ConeColumn ncol{};
    if (EPI == EPI_MOTION && live) ncol = cone_column(mot.next, ix, iy);

    // ---- optional LDS staging of the workgroup's detector rectangle
    __shared__ float2 stage[STAGE ? LDS_CELLS : 1];
    __shared__ int rA0, rA1, rB0, rB1;
    int A0 = 0, B0 = 0, RA = 0, RB = 0;
    bool staged = false;
    if (STAGE) {
        if (tid == 0) {
            rA0 = g.na;
            rA1 = 0;
            rB0 = g.nb;
            rB1 = 0;
        }
        __syncthreads();
        int a0 = g.na, a1 = 0, b0 = g.nb, b1 = 0;
        if (live && acount > 0 && iz0 < iz1) {
            // the slab's rows: its first voxel's lower bound to its last voxel's upper bound
            // (both bounds grow with z), with the same arithmetic as the voxel loop below
            const float d = (float)(iz1 - 1 - iz0);
            const float lo = fminf(fminf(bl.l0, bl.l1), bl.l3);
            const float hi = fmaxf(fmaxf(fmaf(bl.q0, d, bl.l0), fmaf(bl.q1, d, bl.l1)),
                                   fmaf(bl.q2, d, bl.l2));
            a0 = ia_lo;
            a1 = min(g.na, ia_lo + acount);
            b0 = clampi(bl.anchor + (int)floorf(lo), 0, g.nb);
            b1 = clampi(bl.anchor + (int)ceilf(fminf(hi, (float)(g.nb - bl.anchor))), 0, g.nb);
        }
        for (int off = 32; off > 0; off >>= 1) {
            a0 = min(a0, __shfl_xor(a0, off, 64));
            a1 = max(a1, __shfl_xor(a1, off, 64));
            b0 = min(b0, __shfl_xor(b0, off, 64));
            b1 = max(b1, __shfl_xor(b1, off, 64));
        }
        if ((tid & 63) == 0) {
            atomicMin(&rA0, a0);
            atomicMax(&rA1, a1);
            atomicMin(&rB0, b0);
            atomicMax(&rB1, b1);
        }
        __syncthreads();
        A0 = rA0;
        B0 = rB0;
        RA = max(rA1 - A0, 0);
        RB = max(rB1 - B0, 0);
        staged = RA > 0 && RB > 0 && RA * RB <= LDS_CELLS;   // uniform over the workgroup
        if (staged) {   // coalesced copy of the rectangle, 4 loads in flight per lane
            const int total = RA * RB;
            const float inv = 1.0f / (float)RA;
            for (int base = tid; base < total; base += BX * BY * 4) {
                float2 v[4];
                int e[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    // past the end: element total-1 again (the same value to the same slot)
                    e[k] = min(base + k * BX * BY, total - 1);
                    int row = (int)((float)e[k] * inv);   // e < 4096: off by at most one
                    row -= (row * RA > e[k]) ? 1 : 0;
                    row += ((row + 1) * RA <= e[k]) ? 1 : 0;
                    const int cl = e[k] - row * RA;
                    const unsigned src = (unsigned)((B0 + row) * g.na + A0 + cl);
                    v[k] = MODE == BACK_PLAIN ? det_at<MODE>(proj, src * 4u)
                                              : det_at<MODE>(rp, src * 8u);
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) stage[e[k]] = v[k];
            }
        }
        __syncthreads();
    }
    // the wavefront's widest footprint, while every lane is active (lanes past the volume and
    // columns without cells count 0)
    int am = acount;
    if (STAGE) {
        for (int off = 32; off > 0; off >>= 1) am = max(am, __shfl_xor(am, off, 64));
        am = __builtin_amdgcn_readfirstlane(am);
    }
    if (!live) return;

    // Every value the voxels read lies in the staged rectangle when `staged` (uniform over the
    // workgroup), so the source is chosen once, outside the loops: element
    // e(ib, ia) = (ib - rb0) * rs + (ia - ra0) of the LDS rectangle or of the detector.
    const void *gsrc = MODE == BACK_PLAIN ? (const void *)proj : (const void *)rp;
    constexpr unsigned ES = MODE == BACK_PLAIN ? 4u : 8u;   // bytes per detector element
    const int plane = g.nx * g.ny;
    // W: the cells per row the loops read, WA or (from the LDS stage) the widest footprint of
    // the wavefront when that is narrower; the slots past it carry weight 0 on a cell the
    // footprint already counts, so leaving them out adds nothing
    auto run = [&](auto from_lds, auto width) {
        constexpr bool LDSRC = decltype(from_lds)::value;
        constexpr int W = decltype(width)::value;
        const int rs = LDSRC ? RA : g.na, rb0 = LDSRC ? B0 : 0, ra0 = LDSRC ? A0 : 0;
        int cof[WA];
#pragma unroll
        for (int k = 0; k < WA; ++k) cof[k] = col[k] - ra0;
        auto value = [&](int e) -> float2 {
            if (LDSRC) return stage[e];
            return det_at<MODE>(gsrc, (unsigned)e * ES);
        };
        // the same values by byte offset: EB bytes per element of the source, rows rsB apart
        constexpr unsigned EB = LDSRC ? 8u : ES;
        const unsigned rsB = (unsigned)rs * EB;
        unsigned cofB[WA];
#pragma unroll
        for (int k = 0; k < WA; ++k) cofB[k] = (unsigned)cof[k] * EB;
        auto value_b = [&](unsigned b) -> float2 {
            if (LDSRC) return *(const float2 *)((const char *)stage + b);
            return det_at<MODE>(gsrc, b);
        };
        // sum_k wa[k] * value(row, cell col[k]) for the row that starts at byte offset rowb
        auto row_sum = [&](unsigned rowb) -> float2 {
            float2 v[W];
#pragma unroll
            for (int k = 0; k < W; ++k) v[k] = value_b(rowb + cofB[k]);
            float s0 = 0.f, s1 = 0.f;
#pragma unroll
            for (int k = 0; k < W; ++k) {
                s0 = fmaf(wa[k], v[k].x, s0);
                s1 = fmaf(wa[k], v[k].y, s1);
            }
            return make_float2(s0, s1);
        };
        // The row sums depend on the column and the row only, and a voxel's first row is often
        // the previous voxel's last one: that row's sum is carried (prow, psum) instead of read
        // again.  The accumulators see the same products in the same row order.
        int prow = -2;
        float2 psum = make_float2(0.f, 0.f);
        // voxel index and slab-relative z carried from voxel to voxel (no multiply per voxel)
        int id = ix + iy * g.nx + iz0 * plane;
        float d = 0.f;
        for (int iz = iz0; iz < iz1; ++iz, id += plane, d += 1.f) {
            // axial bounds (:131-132,:137-138) in anchor-relative FP32
            const float L0 = fmaf(bl.q0, d, bl.l0), L1 = fmaf(bl.q1, d, bl.l1);
            const float L2 = fmaf(bl.q2, d, bl.l2), L3 = fmaf(bl.q2, d, bl.l3);
            const float lBmin = fminf(fminf(L0, L1), L3), lBmax = fmaxf(fmaxf(L0, L1), L2);
            const int ib_lo = max(0, bl.anchor + (int)floorf(lBmin));   // :142,:158
            const float off = (float)(bl.anchor - ib_lo);               // small integer
            const float bminL = off + lBmin, bmaxL = off + lBmax;
            // rows ib < MIN(nb, b_max) (:158); an empty a range contributes nothing
            const float top = (float)(g.nb - ib_lo);
            const float blim = (top < bmaxL) ? top : bmaxL;
            const int n = (acount > 0 && blim > 0.f) ? (int)ceilf(blim) : 0;
            MotionIn min_{};
            float x_old = 0.f;
            if (EPI == EPI_MOTION) min_ = motion_load(mot, ncol, id, iz, n > 0);
            if (EPI == EPI_ADD) x_old = add.x[id];
            // row weight (:164) as a difference of the clamped boundary F(b) = min(max(b, lo), hi),
            // one median: with rows (n > 0) lo <= hi (bminL < 1 <= top and bminL <= bmaxL), where
            // the median of (b, lo, hi) is that clamp exactly; F is used only then
            const float bhi = fminf(bmaxL, top);
            auto Ff = [&](float b) { return __builtin_amdgcn_fmed3f(b, bminL, bhi); };
            auto F = [&](int b) { return Ff((float)b); };
            float acc = 0.f, acc1 = 0.f;
            if (!wide) {
                // row 0 from the carry when it is the previous voxel's last row (n > 0 only:
                // the carried sum may be non-finite and must not meet a zero weight)
                const bool carry = n > 0 && ib_lo == prow;
                const float F0 = F(0), F1 = F(1);
                const float cx = fmaf(F1 - F0, psum.x, acc), cy = fmaf(F1 - F0, psum.y, acc1);
                acc = carry ? cx : acc;
                acc1 = carry ? cy : acc1;
                // two rows per step, all their loads in flight; a second row past the end reads
                // the first again with weight F(lb+2) - F(lb+1) = 0.  F at the step's upper
                // boundary is the next step's lower one (lbf counts rows in float: exact).
                // a step reads rows lb and min(lb + 1, n - 1): byte offsets e0 and
                // min(e0 + rsB, el), el the voxel's last row (rows at or above the stage's first)
                float Fa = carry ? F1 : F0;
                const int lb0 = carry ? 1 : 0;
                // (from the stage rows and strides are below 2^12 and 2^15: 24-bit products)
                auto mulB = [&](int r) {
                    return LDSRC ? __umul24((unsigned)r, rsB) : (unsigned)r * rsB;
                };
                unsigned rowB = mulB(ib_lo - rb0);
                asm("" : "+v"(rowB));   // keeps el one 24-bit multiply-add off rowB
                unsigned e0 = rowB + (carry ? rsB : 0u);
                const unsigned el = rowB + mulB(n - 1);
                const float nf = (float)n;
#pragma unroll 1   // one or two steps per voxel: unrolling only adds remainder code
                for (float lbf = (float)lb0; lbf < nf; lbf += 2.f, e0 += 2u * rsB) {
                    // e0 opaque to loop strength reduction (an empty asm statement, no
                    // instruction): one offset per row plus a cell offset per load, instead of
                    // one induction variable per cell
                    asm("" : "+v"(e0));
                    const unsigned e1 = min(e0 + rsB, el);
                    const float Fb = Ff(lbf + 1.f), Fc = Ff(lbf + 2.f);
                    const float2 s0 = row_sum(e0), s1 = row_sum(e1);
                    const float w0 = Fb - Fa, w1 = Fc - Fb;
                    acc = fmaf(w0, s0.x, acc);
                    acc1 = fmaf(w0, s0.y, acc1);
                    acc = fmaf(w1, s1.x, acc);
                    acc1 = fmaf(w1, s1.y, acc1);
                    Fa = Fc;
                    psum = s1;   // row min(lb + 1, n - 1): the voxel's last row after its last step
                }
                if (n > 0) prow = ib_lo + n - 1;
            } else {
                for (int lb = 0; lb < n; ++lb) {
                    const float wb = F(lb + 1) - F(lb);
                    const int rowe = (ib_lo + lb - rb0) * rs;
                    for (int la = 0; la < acount; ++la) {
                        const float w = fminf(la + 1.f, amaxL) - fmaxf((float)la, aminL);
                        const float2 v = value(rowe + ia_lo + la - ra0);
                        acc = fmaf(wb, w * v.x, acc);
                        acc1 = fmaf(wb, w * v.y, acc1);
                    }
                }
            }
            if (MODE == BACK_PLAIN) {
                out[id] = acc;
            } else {
                // src/cu_division.cu:20-24, with the hardware reciprocal (<= 2 ulp from the
                // IEEE quotient); 0/0 and x/0 still give NaN / Inf, mapped to 0
                float u = acc * __builtin_amdgcn_rcpf(acc1);
                if (isnan(u) || isinf(u)) u = 0.f;
                if (EPI == EPI_MOTION) {   // square slices: the transposed voxel is (iy, ix, iz)
                    motion_step_at(mot, min_, id, iy + ix * g.nx + iz * plane, u);
                } else if (EPI == EPI_ADD) {   // src/cu_add.cu:17, then the optional clamp
                    float y = x_old + u * add.lambda;
                    if (add.nonneg) y = fmaxf(y, 0.f);
                    add.x[id] = y;
                    if (add.xT) add.xT[iy + ix * g.ny + iz * plane] = y;
                } else {
                    out[id] = u;
                }
            }
        }
    };
    using Full = std::integral_constant<int, WA>;
    if (STAGE && __builtin_amdgcn_readfirstlane((int)staged)) {
        constexpr int W2 = WA - 2 > 2 ? WA - 2 : 2, W1 = WA - 1 > 2 ? WA - 1 : 2;
        if (am <= W2)
            run(std::true_type{}, std::integral_constant<int, W2>{});
        else if (am <= W1)
            run(std::true_type{}, std::integral_constant<int, W1>{});
        else
            run(std::true_type{}, Full{});
    } else {
        run(std::false_type{}, Full{});
    }
}

// Voxels per lane along z: the longest slab (up to 8) that still leaves >= 1024 workgroups.
int pick_zslab(const Geom &g) {
    const long long cols = (long long)((g.nx + BX - 1) / BX) * ((g.ny + BY - 1) / BY);
    int zs = 8;
    while (zs > 1 && cols * ((g.nz + zs - 1) / zs) < 1024) zs /= 2;
    return zs;
}

template <int MODE, int WA>
void launch_wa(dim3 grid, dim3 block, hipStream_t st, const GeomArgs &ga, const ViewTrig &t,
               const float *proj, const float2 *rp, float *out, int zs, const NormFinish &nf,
               bool stage, const MotionStep *mot, const AddStep *add) {
    const double ida = 1.0 / (double)ga.da, idb = 1.0 / (double)ga.db;
    const MotionStep m = mot ? *mot : MotionStep{};
    const AddStep ad = add ? *add : AddStep{};
    // bands of tile columns for views whose rays run mostly along x, when the fan at the
    // isocentre covers less than 85 % of the volume's width (a fan as wide as the volume leaves
    // no band empty, and column bands cost L2 locality: every XCD then reads the whole detector)
    const double mag = ga.SO / ga.SD, half = 0.5 * (double)ga.ny;
    const double lat0 = ((double)ga.ai - 0.5) * ga.da * mag;
    const double lat1 = ((double)ga.na + (double)ga.ai - 0.5) * ga.da * mag;
    const double inside = std::fmin(std::fmax(lat0, lat1), half) -
                          std::fmax(std::fmin(lat0, lat1), -half);
    const double covered = std::fmax(0.0, inside) / (double)ga.ny;
    const int xb = std::fabs(t.c) > std::fabs(t.s) && covered < 0.85 ? 1 : 0;
#define MCSART_BACK(ST, EP)                                                                      \
    hipLaunchKernelGGL((k_backproject<MODE, WA, ST, EP>), grid, block, 0, st, ga, t.c, t.s, proj, \
                       rp, out, zs, ida, idb, nf, m, ad, xb)
    constexpr int MOTION = MODE == BACK_RATIO ? EPI_MOTION : EPI_STORE;
    constexpr int ADD = MODE == BACK_RATIO ? EPI_ADD : EPI_STORE;
    if (MODE == BACK_RATIO && mot) {
        if (stage) MCSART_BACK(true, MOTION);
        else MCSART_BACK(false, MOTION);
    } else if (MODE == BACK_RATIO && add) {
        if (stage) MCSART_BACK(true, ADD);
        else MCSART_BACK(false, ADD);
    } else {
        if (stage) MCSART_BACK(true, EPI_STORE);
        else MCSART_BACK(false, EPI_STORE);
    }
#undef MCSART_BACK
}

template <int MODE>
void launch_mode(dim3 grid, dim3 block, hipStream_t st, const GeomArgs &ga, const ViewTrig &t,
                 int wa, const float *proj, const float2 *rp, float *out, int zs,
                 const NormFinish &nf, bool stage, const MotionStep *mot, const AddStep *add) {
#define MCSART_WA(W)                                                                             \
    launch_wa<MODE, W>(grid, block, st, ga, t, proj, rp, out, zs, nf, stage, mot, add)
    switch (wa <= 3 ? 3 : wa >= 8 ? 8 : wa) {
        case 3: MCSART_WA(3); break;
        case 4: MCSART_WA(4); break;
        case 5: MCSART_WA(5); break;
        case 6: MCSART_WA(6); break;
        case 7: MCSART_WA(7); break;
        default: MCSART_WA(8); break;
    }
#undef MCSART_WA
}

}  // namespace

int launch_backprojector(const Geom &g, const ViewTrig &t, const ViewInfo &vi, BackMode mode,
                         const float *proj, const float2 *rp, float *out, hipStream_t st,
                         const NormFinish *nfp, const BackTune *tune, const MotionStep *mot,
                         const AddStep *add) {
    const GeomArgs ga{g.nx, g.ny, g.nz, g.na, g.nb, g.da, g.db, g.ai, g.bi, g.SO, g.SD};
    if (g.ndet() >= (1LL << 29))
        return fail(MCSART_ENOTSUP, "backprojector: detectors of 2^29 cells or more");
    if (mot && (mode != BACK_RATIO || g.nx != g.ny))
        return fail(MCSART_EINVAL, "backprojector: the motion step needs the ratio mode and nx == ny");
    if (add && (mode != BACK_RATIO || mot))
        return fail(MCSART_EINVAL, "backprojector: the add step needs the ratio mode alone");
    const int zs = (tune && tune->zslab > 0) ? tune->zslab : pick_zslab(g);
    const bool stage = !tune || tune->stage != 2;   // 0 automatic (staged), 1 staged, 2 global
    const dim3 block(BX, BY);
    const dim3 grid((g.nx + BX - 1) / BX, (g.ny + BY - 1) / BY, (g.nz + zs - 1) / zs);
    const NormFinish nf = nfp ? *nfp : NormFinish{nullptr, 0, nullptr};
    if (mode == BACK_PLAIN)
        launch_mode<BACK_PLAIN>(grid, block, st, ga, t, vi.wa, proj, rp, out, zs, nf, stage,
                                nullptr, nullptr);
    else
        launch_mode<BACK_RATIO>(grid, block, st, ga, t, vi.wa, proj, rp, out, zs, nf, stage, mot,
                                add);
    MCSART_CHECK_LAUNCH();
    return MCSART_OK;
}

}  // namespace mcsart

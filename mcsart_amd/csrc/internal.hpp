// Internal declarations shared by the libmcsart_hip translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "footprint.hpp"
#include "mcsart_hip.h"

namespace mcsart {

// ---- error handling -------------------------------------------------------------------------
void set_error(const std::string &msg);
int fail(int status, const std::string &msg);

#define MCSART_HIP_TRY(expr)                                                              \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return ::mcsart::fail(MCSART_EHIP, std::string(#expr) + ": " +                \
                                                   hipGetErrorString(e_));                \
    } while (0)

#define MCSART_CHECK_LAUNCH()                                                             \
    do {                                                                                  \
        hipError_t e_ = hipGetLastError();                                                \
        if (e_ != hipSuccess)                                                             \
            return ::mcsart::fail(MCSART_EHIP, std::string("kernel launch: ") +           \
                                                   hipGetErrorString(e_));                \
    } while (0)

// ---- geometry -------------------------------------------------------------------------------
struct Geom {
    int nx, ny, nz, na, nb;
    float da, db, ai, bi, SO, SD;
    long long nvox() const { return (long long)nx * ny * nz; }
    long long ndet() const { return (long long)na * nb; }
};

int validate_geom(const Geom &g);
int validate_volume(int nx, int ny, int nz, bool need_square);

// Per-view trigonometry, evaluated once on the host in FP64 from the float angle the reference
// passes (src/cu_projection.cu:23-24, src/cu_backprojection.cu:62-63).
struct ViewTrig {
    double c, s;
};
ViewTrig view_trig(float angle);

// ---- XCD-aware workgroup order ----------------------------------------------------------------
// The device hands consecutive workgroups of a launch to its 8 XCDs in turn, and each XCD has its
// own L2.  xcd_order maps the hardware workgroup index L (0 <= L < T) to a logical tile index so
// that every XCD receives one contiguous range of logical tiles (a bijection of [0, T)); kernels
// lay tiles out so that neighbouring tiles share data.
constexpr int kXcds = 8;
__device__ inline int xcd_order(int L, int T) {
    const int q = T / kXcds, r = T % kXcds, x = L % kXcds;
    return x * q + min(x, r) + L / kXcds;
}
// The opposite need: a kernel whose work per tile depends on where the tile lies (the culled
// warps) must not give an XCD a region of the volume, which a view can leave empty.  For a
// (gx, gy, gz) grid the hardware hands XCD k the workgroups with blockIdx.x = k (mod 8) when gx
// is a multiple of 8, i.e. whole columns of tiles; tile column (hx + hy + hz) mod gx instead
// gives every XCD every column in turn (culled warp at 256x256x128: 46.4 -> 42.7 us).
__device__ inline int spread_x(int hx, int hy, int hz, int gx) { return (hx + hy + hz) % gx; }

// ---- projector ------------------------------------------------------------------------------
enum ProjMode {
    PROJ_PLAIN = 0,   // out[id] = P(x)
    PROJ_IMAGE = 1,   // rp[id] = (p[id] - P(x), P(1)), norms of p and r
    PROJ_MOTION = 2,  // rp[id] = (p[id] - P(g), qscale P(Q)), norms of p and r
};

// Volumes the projector reads.  y-branch rays read the natural layout img[ix + iy*nx + iz*nx*ny];
// x-branch rays read the xy-transposed layout imgT[iy + ix*ny + iz*nx*ny] when it is given (it
// makes their gathers contiguous) and the natural layout otherwise.  PLAIN/IMAGE project vol0;
// MOTION projects interleaved pairs (g, Q) and scales P(Q) by qscale: the motion sweep's
// q = Q (-vd - fd) is a per-sweep volume times a per-view scalar, so Q is written once per sweep
// and only g per view.
struct ProjArgs {
    const float *vol0;     // natural layout (PLAIN/IMAGE)
    const float *vol0T;    // transposed layout or null
    const float2 *pair;    // (g, Q) natural layout (MOTION)
    const float2 *pairT;   // (g, Q) transposed layout or null (MOTION)
    float qscale;          // MOTION: P(q) = qscale P(Q)
    const float *meas;     // measured projection of this view (IMAGE/MOTION)
    float *out;            // PLAIN: D floats
    float2 *rp;            // IMAGE/MOTION: D float2 (residual, normaliser projection)
    float *resid_copy;     // optional: residual copy (reference h_outproj slice)
    double *partials;      // per-block (sum p^2, sum r^2), plain stores
    // set by launch_projector for the rows kernel: the lower of the two layouts' addresses and
    // each layout's byte offset from it (both layouts within 4 GiB of it)
    const char *rbase;
    unsigned roffN, roffT;
};
// vi: the view's branches and footprint spans (footprint.hpp view_info).  SART modes leave one
// (sum p^2, sum r^2) pair per workgroup in a.partials and report their count in *nparts; a later
// kernel on the same stream turns them into ||r||/||p|| (NormFinish).  No fence or atomic is
// needed: the kernel boundary orders the stores.
// split: workgroup slices along the ray (power of two up to 16; 0 = automatic); rows: detector
// rows walked by one lane (1, 2 or 4; 0 = automatic).
int launch_projector(const Geom &g, const ViewTrig &t, const ViewInfo &vi, ProjMode mode,
                     const ProjArgs &a, hipStream_t st, int *nparts = nullptr, int split = 0,
                     int rows = 0);
ViewInfo view_info_of(const Geom &g, const ViewTrig &t);
int projector_partials(const Geom &g);   // number of per-block partial slots needed

// Per-view residual ratio ||r|| / ||p|| (src/AUMISART.cu:204-209) from projector partials,
// summed in a fixed order by one wavefront.
struct NormFinish {
    const double *partials;
    int nparts;
    float *out;
};
// Must be called by all 64 lanes of one wavefront; `lane` is the caller's lane index.
__device__ inline void finish_norm_wave(const NormFinish &nf, int lane) {
    double sp = 0.0, sr = 0.0;
    for (int k = lane; k < nf.nparts; k += 64) {
        sp += nf.partials[2 * k];
        sr += nf.partials[2 * k + 1];
    }
    for (int off = 32; off > 0; off >>= 1) {
        sp += __shfl_xor(sp, off, 64);
        sr += __shfl_xor(sr, off, 64);
    }
    if (lane == 0) *nf.out = (float)sqrt(sr) / (float)sqrt(sp);
}
int launch_norm_finish(const NormFinish &nf, hipStream_t st);

// ---- backprojector --------------------------------------------------------------------------
enum BackMode {
    BACK_PLAIN = 0,  // out = B(proj)
    BACK_RATIO = 1,  // out = B(rp.x) / B(rp.y), NaN/Inf -> 0
};
// Launch shape of the backprojector: voxels per lane along z (0 = automatic) and whether the
// workgroup stages its detector rectangle in LDS (0 = automatic: staged, 1 = staged, 2 = read
// the detector from global memory).
struct BackTune {
    int zslab;
    int stage;
};
// vi: the view's transverse window (view_info).  nf (optional): the first wavefront of the first
// workgroup also finishes the preceding projector's norm reduction.
// A motion-sweep step fused into the ratio backprojector (x is constant over the sweep; the
// accumulators are kept transposed, SVt(p) = SV(p^T), so every access is at the voxel itself,
// and interleaved, acc[p] = (SVt(p), SFt(p)), so a voxel reads and writes both with one 8-byte
// access):
//   SVt(p) += wv u(p), SFt(p) += wf u(p)                (update of view i, from zero if `first`)
//   img0(p) = vd Ga(p) + fd Gb(p) + H(p) (vd SVt(p) + fd SFt(p))
//   g(p) = x(p) - img0(p)                               (prep of view i+1, unless `last`)
// with cst = (Ga, Gb, H, x) per voxel (one 16-byte record: the constants and the sweep's image)
// from launch_motion_begin, which also wrote the sweep's Q (the projector scales P(Q) by
// -vd - fd); u itself is not stored.
// Work the next kernels never see is skipped: g only where the next view's projector can read it
// (`next`, its read set and edge shell: ConeCull), and, when `skip_zero` (wv and wf finite), the
// accumulators only where u can be nonzero (the voxel's footprint meets the detector; elsewhere
// u = 0 and the sums do not change).  A voxel with neither is not touched at all.
struct MotionStep {
    float2 *acc;         // (SVt, SFt) per voxel
    const float4 *cst;   // (Ga, Gb, H, x) per voxel
    float2 *pair, *pairT;   // the g halves of the (g, Q) pairs the next view reads (either
                            // may be null); Q stays as launch_motion_begin wrote it
    float wv, wf, vd, fd;
    int first, last;
    ConeCull next;       // read set of the next view's projector (mode 0: every voxel)
    int skip_zero;
};
// The static image sweep's update fused into the ratio backprojector (src/AUMISART.cu:62):
//   x(p) = max(x(p) + lambda u(p), 0 if nonneg), and the same value at p^T in xT when given (the
//   xy-transposed copy the next view's x-branch rays read); u itself is not stored.
struct AddStep {
    float *x, *xT;
    float lambda;
    int nonneg;
};
// At most one of mot and add: the ratio mode's u then goes to that step instead of out.
int launch_backprojector(const Geom &g, const ViewTrig &t, const ViewInfo &vi, BackMode mode,
                         const float *proj, const float2 *rp, float *out, hipStream_t st,
                         const NormFinish *nf = nullptr, const BackTune *tune = nullptr,
                         const MotionStep *mot = nullptr, const AddStep *add = nullptr);

// ---- warps and elementwise ------------------------------------------------------------------
struct Fields {
    const float *ax, *ay, *az, *bx, *by, *bz;
};
struct FieldsW {   // the same six fields, writable
    float *ax, *ay, *az, *bx, *by, *bz;
};
// out = W(src; vol, flow) in the natural layout and/or (outT) the xy-transposed layout; either
// output may be null; neither may alias src.
// read_set (optional): only the voxels the view's projector can read are written (ConeCull).
int launch_warp(float *out, float *outT, const float *src, int nx, int ny, int nz, float vol,
                float flow, const Fields &f, hipStream_t st, const ConeCull *read_set = nullptr);
// dst = xy-transpose of src (dst[iy + ix*ny + iz*nx*ny] = src[ix + iy*nx + iz*nx*ny])
int launch_transpose(float *dst, const float *src, int nx, int ny, int nz, hipStream_t st);
// x += lambda * W(u; vol, flow), then optional clamp at zero.  support (optional): the voxels
// where u can be nonzero (ConeCull); x is left as it is where W(u) is 0 for that reason, which
// with the clamp needs x already clamped everywhere.
int launch_warp_add(float *x, const float *u, int nx, int ny, int nz, float vol, float flow,
                    const Fields &f, float lambda, int nonneg, hipStream_t st,
                    const ConeCull *support = nullptr);
int launch_deform2(float *out, const float *src, int nx, int ny, int nz, float vol, float flow,
                   const Fields &f, hipStream_t st);
// x += lambda * u, then optional clamp at zero.
int launch_add_clamp(float *x, const float *u, long long n, float lambda, int nonneg,
                     hipStream_t st);
int launch_add(float *a, const float *b, long long n, float w, hipStream_t st);
int launch_add2(float *a, const float *b, int nx, int ny, int nz, const float *c, float w,
                int ind, hipStream_t st);
int launch_fill(float *a, long long n, float v, hipStream_t st);
int launch_initial2(float *img, int nx, int ny, int nz, const float *img0, float vol, float flow,
                    hipStream_t st);
int launch_division(float *a, const float *b, long long n, hipStream_t st);
int launch_diff(float *out, const float *img, int nx, int ny, int nz, int ind, int backward,
                hipStream_t st);
int launch_update_udvf(float *ax, float *ay, float *az, float *bx, float *by, float *bz,
                       const float *img, const float *img0, float vd, float fd, int nx, int ny,
                       int nz, hipStream_t st);
// Motion sweep, fused (src/AUMISART.cu:246-277):
//   g = x - sum_c (alpha_c^T vd + beta_c^T fd) * D_c x,  q = (D1+D2+D3)x * (-vd - fd)
// Pairs (g, q) go to `pair` (natural layout) and/or `pairT` (transposed); either may be null.
int launch_motion_prep(float2 *pair, float2 *pairT, const float *x, const float *const fields[6],
                       int nx, int ny, int nz, float vd, float fd, hipStream_t st);
//   alpha_c += u^T * wv * D_c x,  beta_c += u^T * wf * D_c x
int launch_motion_update(float *const fields[6], const float *u, const float *x, int nx, int ny,
                         int nz, float wv, float wf, hipStream_t st);
// Both in one pass: the update of one view (u, wv, wf) followed by the prep of the next view
// (vd, fd) on the updated fields; either half may be switched off.
int launch_motion_step(float *const fields[6], const float *u, const float *x, float2 *pair,
                       float2 *pairT, int nx, int ny, int nz, float wv, float wf, float vd,
                       float fd, bool update, bool prep, hipStream_t st);

// The motion sweep in accumulator form (x constant over the sweep): per-sweep constants
// cst = (Ga, Gb, H, x) (4 floats per voxel) from the starting fields, and pairs (g, Q) in both
// layouts: Q = (D1 + D2 + D3) x and the prep g of the first view ...
int launch_motion_begin(const Fields &f, const float *x, float *cst, float2 *pair,
                        float2 *pairT, int nx, int ny, int nz, float vd, float fd,
                        hipStream_t st);
// ... then per updated view the ratio backprojector with a MotionStep, and at the end of the
// sweep the fields alpha_c = alpha_c0 + D_c x SV, beta_c = beta_c0 + D_c x SF in place (f),
// with (SV(P), SF(P)) = acc[P^T].
int launch_motion_finish(const float2 *acc, const float *x, const FieldsW &f, int nx, int ny,
                         int nz, hipStream_t st);

// img1 *= w * D_ind(img0) (src/cu_multiple.cu:11-43)
int launch_multiple(float *a, const float *b, int nx, int ny, int nz, float w, int ind,
                    hipStream_t st);
// m_c = alpha_c * vol + beta_c * flow into three volumes (src/cu_deform.cu:66-77)
int launch_forward_dvf(float *mx, float *my, float *mz, const Fields &f, int nx, int ny, int nz,
                       float vol, float flow, hipStream_t st);
// fixed-point DVF inverse (waste/kernel_invertDVF.cu:1-18); outputs must not alias inputs
int launch_invert_dvf(float *mx2, float *my2, float *mz2, const float *mx, const float *my,
                      const float *mz, int nx, int ny, int nz, int niter, hipStream_t st);

// ---- contexts -------------------------------------------------------------------------------
enum TimingKind { TK_PROJ = 0, TK_BACK = 1, TK_WARP = 2, TK_MOTION = 3, TK_COUNT = 4 };

}  // namespace mcsart

struct mcsart_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // timing (events recorded around kernels on `stream`)
    bool timing = false;
    struct Pending {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    double total_ms[mcsart::TK_COUNT] = {0, 0, 0, 0};
    long long launches[mcsart::TK_COUNT] = {0, 0, 0, 0};
    // scratch for the synchronous operator shims
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    // norm-reduction scratch (per-block partials)
    void *red = nullptr;
    size_t red_bytes = 0;
};

namespace mcsart {
// mcsart_job_create in a caller-provided device block of block_bytes (at least
// mcsart_job_bytes): the job uses it and never frees it.  block == nullptr: the job allocates its
// own, as mcsart_job_create does.
int job_create_in(mcsart_ctx *ctx, const mcsart_geom *g, int n_views, const float *angles,
                  const float *volumes, const float *flows, const mcsart_opts *opts, int n_iter,
                  const int *op_iter, void *block, size_t block_bytes, mcsart_job **out);
mcsart_ctx *default_ctx();   // per-device context behind the synchronous operators
int ctx_scratch(mcsart_ctx *ctx, size_t bytes, void **ptr);
// Reduction scratch for the projector's SART modes: 2*nblocks doubles of partials.
int ctx_reduction(mcsart_ctx *ctx, int nblocks, double **partials);
// Timing helpers: bracket a launch with events when ctx->timing is on.
struct TimedRegion {
    mcsart_ctx *ctx;
    int kind;
    hipEvent_t a = nullptr, b = nullptr;
    TimedRegion(mcsart_ctx *c, int k);
    ~TimedRegion();
};
int ctx_collect_timing(mcsart_ctx *ctx);
}  // namespace mcsart

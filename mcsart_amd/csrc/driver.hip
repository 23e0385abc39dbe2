// Host driver: device-resident McSART jobs and the host_AUMISART drop-ins.
//
// Reference: both overloads of host_AUMISART (src/AUMISART.cu:3-141 static, :143-313 motion).
// The reference allocates per call, uploads each measured view and downloads each residual every
// view, synchronises the device after every kernel and reads two cuBLAS norms back to the host
// per view.  A job here uploads its measured projections, image and fields once, keeps them in
// HBM for the whole schedule, and enqueues each view as four kernels on the context's stream
// with no host round trip: the norms are reduced on the device into a per-view array.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "internal.hpp"

using namespace mcsart;

struct mcsart_job {
    mcsart_ctx *ctx = nullptr;
    Geom g{};
    int V = 0;
    int n_iter = 0;
    std::vector<int> op_iter;
    mcsart_opts opts{};
    std::vector<float> angles, volumes, flows;
    std::vector<ViewTrig> trig;
    std::vector<ViewInfo> vinfo;   // per view: branches present, footprint spans
    float *d_proj = nullptr;     // V*D measured projections
    float *d_img = nullptr;      // N
    float *d_fields = nullptr;   // 6N: ax ay az bx by bz
    float *d_init = nullptr;     // 7N: initial image and fields (mcsart_job_reset)
    float *d_w1 = nullptr;       // 2N: (g, Q) pairs (motion sweep) / scratch
    float *d_w1T = nullptr;      // 2N: the same in the xy-transposed layout / transposed image
    float *d_wi = nullptr;       // N: warped image (image sweep)
    float *d_wiT = nullptr;      // N: the same in the xy-transposed layout
    float *d_w2 = nullptr;       // N: u
    float *d_acc = nullptr;      // 2N: motion-sweep accumulators, transposed, as (SVt, SFt) pairs
    float *d_cst = nullptr;      // 4N: motion-sweep constants and image (Ga, Gb, H, x) per voxel
    float2 *d_rp = nullptr;      // D
    float *d_norm = nullptr;     // n_iter*V
    float *d_resid = nullptr;    // V*D (keep_residual)
    double *d_partials = nullptr;
    void *block = nullptr;       // single allocation backing all of the above
    size_t bytes = 0;
    bool owns_block = true;      // false: the block belongs to the caller (a server's pool)
    // One captured graph per sweep of the schedule: the kernels of a sweep and all their
    // arguments are fixed for the life of the job, so a sweep is captured on first use and
    // replayed afterwards (one launch instead of ~4 per view).  Sweeps run eagerly while the
    // context times kernels (HIP events recorded inside a graph carry no timestamps).
    struct SweepGraph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<SweepGraph> graphs;
    bool use_graphs = true;
    int cull = 1;   // 0: every voxel every view; 1: skip what the views never see; 2: as 1, with
                    // the skipped cells written as NaN (mcsart_job_set_culling)
    int proj_split = 0, proj_rows = 0;   // projector launch shape (0 = automatic)
    BackTune back_tune{0, 0};   // backprojector launch shape (0 = automatic)
    int back_zslab_motion = 0;  // z-slab of the motion sweep's backprojector (0 = automatic)
};

namespace {

size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

struct Layout {
    size_t proj, img, fields, init, w1, w1T, wi, wiT, w2, acc, cst, rp, norm, resid, partials,
        total;
};

Layout layout(const Geom &g, int V, int n_iter, int keep_residual) {
    const size_t N = (size_t)g.nvox(), D = (size_t)g.ndet();
    Layout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += align_up(bytes);
        return o;
    };
    L.proj = take(sizeof(float) * D * V);
    L.fields = take(sizeof(float) * N * 6);
    L.init = take(sizeof(float) * N * 7);
    // the volumes a projector reads next to each other: its gathers address both layouts from
    // one base with 32-bit byte offsets
    L.img = take(sizeof(float) * N);
    L.w1 = take(sizeof(float) * N * 2);
    L.w1T = take(sizeof(float) * N * 2);
    // the image sweep's warped image in its own volumes: with culling a view's warp leaves the
    // cells its projector reads with zero weights at most as an earlier view wrote them, so they
    // only ever hold warps of the image (or the zeros the job starts with)
    L.wi = take(sizeof(float) * N);
    L.wiT = take(sizeof(float) * N);
    L.w2 = take(sizeof(float) * N);
    L.acc = take(sizeof(float) * N * 2);
    L.cst = take(sizeof(float) * N * 4);
    L.rp = take(sizeof(float2) * D);
    L.norm = take(sizeof(float) * (size_t)n_iter * V);
    L.resid = take(keep_residual ? sizeof(float) * D * V : 0);
    L.partials = take(sizeof(double) * 2 * (size_t)projector_partials(g));
    L.total = off;
    return L;
}

int check_job(mcsart_job *job) {
    if (!job) return fail(MCSART_EINVAL, "null job");
    return MCSART_OK;
}

FieldsW fields_w(mcsart_job *j) {
    const size_t N = (size_t)j->g.nvox();
    float *f = j->d_fields;
    return FieldsW{f, f + N, f + 2 * N, f + 3 * N, f + 4 * N, f + 5 * N};
}

Fields fields_of(const mcsart_job *j) {
    const size_t N = (size_t)j->g.nvox();
    const float *f = j->d_fields;
    return Fields{f, f + N, f + 2 * N, f + 3 * N, f + 4 * N, f + 5 * N};
}

GeomArgs geom_args(const Geom &g) {
    return GeomArgs{g.nx, g.ny, g.nz, g.na, g.nb, g.da, g.db, g.ai, g.bi, g.SO, g.SD};
}

// Voxels the projector of view iv can read (ConeCull), under the job's cull mode.
ConeCull read_set(const mcsart_job *j, int iv) {
    return projector_read_set(geom_args(j->g), j->trig[iv].c, j->trig[iv].s, j->vinfo[iv],
                              j->cull);
}

// Voxels where view iv's ratio backprojection u can be nonzero (ConeCull); mode 0 on the first
// view of an image sweep, whose warp-add then clamps every voxel (the skip needs x clamped).
ConeCull u_support(const mcsart_job *j, int iv) {
    return backprojection_support(geom_args(j->g), j->trig[iv].c, j->trig[iv].s,
                                  (j->cull && iv > 0) ? 1 : 0);
}

// The motion step of view iv, preparing view nv (nv >= V: the sweep's last view).
MotionStep motion_step(const mcsart_job *j, int iv, int nv, float2 *pair, float2 *pairT,
                       int first) {
    const bool more = nv < j->V;
    const ViewInfo &vn = j->vinfo[more ? nv : iv];
    const float wv = j->volumes[iv - 1] - j->volumes[iv], wf = j->flows[iv - 1] - j->flows[iv];
    MotionStep m{reinterpret_cast<float2 *>(j->d_acc),
                 reinterpret_cast<const float4 *>(j->d_cst),
                 more && vn.has_y ? pair : nullptr,
                 more && vn.has_x ? pairT : nullptr,
                 wv,
                 wf,
                 more ? j->volumes[nv] - j->volumes[nv - 1] : 0.f,
                 more ? j->flows[nv] - j->flows[nv - 1] : 0.f,
                 first,
                 more ? 0 : 1,
                 more ? read_set(j, nv) : ConeCull{},
                 (j->cull && std::isfinite(wv) && std::isfinite(wf)) ? 1 : 0};
    return m;
}

// Image sweep (src/AUMISART.cu:182-235; static form :40-65).
int image_sweep(mcsart_job *j, int it) {
    const Geom &g = j->g;
    const long long D = g.ndet();
    hipStream_t st = j->ctx->stream;
    const Fields f = fields_of(j);
    const bool motion = j->opts.motion != 0;
    int rc, nparts = 0;
    // Static sweeps project x itself: keep its transposed copy (in d_w1T) current for views with
    // x-branch rays; the update kernel of view i writes it when view i+1 needs it.
    if (!motion && j->vinfo[0].has_x) {
        TimedRegion tr(j->ctx, TK_WARP);
        rc = launch_transpose(j->d_w1T, j->d_img, g.nx, g.ny, g.nz, st);
        if (rc) return rc;
    }
    for (int iv = 0; iv < j->V; ++iv) {
        const ViewInfo &vi = j->vinfo[iv];
        ProjArgs a{};
        if (motion) {   // x_i = W(x; v_i - v_0, f_i - f_0)  (:193), in the layouts P reads
            TimedRegion tr(j->ctx, TK_WARP);
            const ConeCull rs = read_set(j, iv);
            rc = launch_warp(vi.has_y ? j->d_wi : nullptr, vi.has_x ? j->d_wiT : nullptr,
                             j->d_img, g.nx, g.ny, g.nz, j->volumes[iv] - j->volumes[0],
                             j->flows[iv] - j->flows[0], f, st, &rs);
            if (rc) return rc;
            a.vol0 = j->d_wi;
            a.vol0T = vi.has_x ? j->d_wiT : nullptr;
        } else {
            a.vol0 = j->d_img;
            a.vol0T = vi.has_x ? j->d_w1T : nullptr;
        }
        {   // r = p_i - P x_i, P(1), rho_i  (:194-211)
            a.meas = j->d_proj + D * iv;
            a.rp = j->d_rp;
            a.resid_copy = j->d_resid ? j->d_resid + D * iv : nullptr;
            a.partials = j->d_partials;
            TimedRegion tr(j->ctx, TK_PROJ);
            rc = launch_projector(g, j->trig[iv], vi, PROJ_IMAGE, a, st, &nparts, j->proj_split,
                                  j->proj_rows);
            if (rc) return rc;
        }
        {   // u = B(r) / B(P1)  (:212-218); also rho_i from the projector's partials.  Static
            // sweeps add lambda u to x in the same pass (:62), with the transposed copy the next
            // view's x-branch rays read.
            const NormFinish nf{j->d_partials, nparts, j->d_norm + (long long)it * j->V + iv};
            const AddStep add{j->d_img,
                              (iv + 1 < j->V && j->vinfo[iv + 1].has_x) ? j->d_w1T : nullptr,
                              j->opts.lambda, j->opts.nonneg};
            TimedRegion tr(j->ctx, TK_BACK);
            rc = launch_backprojector(g, j->trig[iv], vi, BACK_RATIO, nullptr, j->d_rp,
                                      motion ? j->d_w2 : nullptr, st, &nf, &j->back_tune, nullptr,
                                      motion ? nullptr : &add);
            if (rc) return rc;
        }
        if (motion) {   // x += lambda * W(u; v_0 - v_i, f_0 - f_i)  (:222-223)
            TimedRegion tr(j->ctx, TK_WARP);
            const ConeCull sup = u_support(j, iv);
            rc = launch_warp_add(j->d_img, j->d_w2, g.nx, g.ny, g.nz,
                                 j->volumes[0] - j->volumes[iv], j->flows[0] - j->flows[iv], f,
                                 j->opts.lambda, j->opts.nonneg, st, &sup);
            if (rc) return rc;
        }
    }
    return MCSART_OK;
}

// Motion (udvf) sweep (src/AUMISART.cu:236-288; static overload :66-118 steps views by 10).
// Per view: prep (g, q) -> projector -> backprojector -> field update.  x is constant over the
// sweep, so the updates are carried as two accumulators and q = Q (-vd - fd) is a per-sweep
// volume times a scalar: launch_motion_begin makes the per-sweep constants, Q and the first
// view's prep g; each view's update and the next view's g run in the ratio backprojector's
// epilogue (MotionStep); the projector scales P(Q); launch_motion_finish writes the fields.
int motion_sweep(mcsart_job *j, int it) {
    const Geom &g = j->g;
    const long long D = g.ndet(), N = g.nvox();
    hipStream_t st = j->ctx->stream;
    const int stride = j->opts.motion ? 1 : 10;
    // pairs (g, Q) in each layout: the view's prep g, the sweep's Q (q = Q (-vd - fd))
    float2 *pair = (float2 *)j->d_w1, *pairT = (float2 *)j->d_w1T;
    auto vdiff = [&](int iv) { return j->volumes[iv] - j->volumes[iv - 1]; };   // :249
    auto fdiff = [&](int iv) { return j->flows[iv] - j->flows[iv - 1]; };       // :250
    int rc, nparts = 0;
    if (j->V < 2) return MCSART_OK;   // no view to update
    {   // constants, Q and the prep of the first view
        TimedRegion tr(j->ctx, TK_MOTION);
        rc = launch_motion_begin(fields_of(j), j->d_img, j->d_cst, pair, pairT, g.nx, g.ny, g.nz,
                                 vdiff(1), fdiff(1), st);
        if (rc) return rc;
    }
    for (int iv = 1; iv < j->V; iv += stride) {
        const ViewInfo &vi = j->vinfo[iv];
        {   // r = p_i - P(x - img0), P(q) = (-vd - fd) P(Q), rho_i  (:246-263)
            ProjArgs a{};
            a.pair = pair;
            a.pairT = vi.has_x ? pairT : nullptr;
            a.qscale = -vdiff(iv) + -fdiff(iv);
            a.meas = j->d_proj + D * iv;
            a.rp = j->d_rp;
            a.partials = j->d_partials;
            TimedRegion tr(j->ctx, TK_PROJ);
            rc = launch_projector(g, j->trig[iv], vi, PROJ_MOTION, a, st, &nparts, j->proj_split,
                                  j->proj_rows);
            if (rc) return rc;
        }
        {   // u = B(r) / B(P q) (:265-270), rho_i; fields += u^T * (-vd | -fd) * D_c x
            // (:272-277) as accumulators; then the next view's prep
            const MotionStep m = motion_step(j, iv, iv + stride, pair, pairT, iv == 1 ? 1 : 0);
            const NormFinish nf{j->d_partials, nparts, j->d_norm + (long long)it * j->V + iv};
            const BackTune bt{j->back_zslab_motion, j->back_tune.stage};
            TimedRegion tr(j->ctx, TK_BACK);
            rc = launch_backprojector(g, j->trig[iv], vi, BACK_RATIO, nullptr, j->d_rp, nullptr,
                                      st, &nf, &bt, &m);
            if (rc) return rc;
        }
    }
    {   // fields = fields_0 + D_c x * accumulators
        TimedRegion tr(j->ctx, TK_MOTION);
        rc = launch_motion_finish(reinterpret_cast<const float2 *>(j->d_acc), j->d_img,
                                  fields_w(j), g.nx, g.ny, g.nz, st);
        if (rc) return rc;
    }
    return MCSART_OK;
}

Geom geom_of(const mcsart_geom *g) {
    return Geom{g->nx, g->ny, g->nz, g->na, g->nb, g->da, g->db, g->ai, g->bi, g->SO, g->SD};
}

}  // namespace

extern "C" size_t mcsart_job_bytes(const mcsart_geom *g, int n_views, int n_iter,
                                   int keep_residual) {
    if (!g || n_views <= 0 || n_iter < 0) return 0;
    return layout(geom_of(g), n_views, n_iter, keep_residual).total;
}

int mcsart::job_create_in(mcsart_ctx *ctx, const mcsart_geom *gp, int n_views,
                          const float *angles, const float *volumes, const float *flows,
                          const mcsart_opts *opts, int n_iter, const int *op_iter, void *block,
                          size_t block_bytes, mcsart_job **out) {
    if (!out) return fail(MCSART_EINVAL, "null job pointer");
    *out = nullptr;
    if (!ctx || !gp || !angles || !volumes || !flows || !opts)
        return fail(MCSART_EINVAL, "null argument");
    if (n_views <= 0) return fail(MCSART_EINVAL, "n_views must be positive");
    if (n_iter < 0 || (n_iter > 0 && !op_iter)) return fail(MCSART_EINVAL, "bad op_iter");
    const Geom g = geom_of(gp);
    int rc = validate_geom(g);
    if (rc) return rc;
    rc = validate_volume(g.nx, g.ny, g.nz, true);
    if (rc) return rc;
    MCSART_HIP_TRY(hipSetDevice(ctx->device));
    const Layout L = layout(g, n_views, n_iter, opts->keep_residual);
    auto *j = new mcsart_job();
    j->ctx = ctx;
    j->g = g;
    j->V = n_views;
    j->n_iter = n_iter;
    j->op_iter.assign(op_iter, op_iter + n_iter);
    j->opts = *opts;
    j->angles.assign(angles, angles + n_views);
    j->volumes.assign(volumes, volumes + n_views);
    j->flows.assign(flows, flows + n_views);
    for (int i = 0; i < n_views; ++i) {
        j->trig.push_back(view_trig(angles[i]));
        j->vinfo.push_back(view_info_of(g, j->trig.back()));
    }
    j->graphs.resize(n_iter);
    if (const char *env = getenv("MCSART_GRAPHS")) j->use_graphs = atoi(env) != 0;
    if (block) {   // the caller's memory: used, never freed here
        if (block_bytes < L.total) {
            delete j;
            return fail(MCSART_EINVAL, "job block of " + std::to_string(block_bytes) +
                                           " bytes is smaller than the " +
                                           std::to_string(L.total) + " the job needs");
        }
        j->block = block;
        j->owns_block = false;
    } else if (hipMalloc(&j->block, L.total) != hipSuccess) {
        (void)hipGetLastError();
        delete j;
        return fail(MCSART_ENOMEM, "job allocation of " + std::to_string(L.total) + " bytes failed");
    }
    j->bytes = L.total;
    char *b = (char *)j->block;
    j->d_proj = (float *)(b + L.proj);
    j->d_img = (float *)(b + L.img);
    j->d_fields = (float *)(b + L.fields);
    j->d_init = (float *)(b + L.init);
    j->d_w1 = (float *)(b + L.w1);
    j->d_w1T = (float *)(b + L.w1T);
    j->d_wi = (float *)(b + L.wi);
    j->d_wiT = (float *)(b + L.wiT);
    j->d_w2 = (float *)(b + L.w2);
    j->d_acc = (float *)(b + L.acc);
    j->d_cst = (float *)(b + L.cst);
    j->d_rp = (float2 *)(b + L.rp);
    j->d_norm = (float *)(b + L.norm);
    j->d_resid = opts->keep_residual ? (float *)(b + L.resid) : nullptr;
    j->d_partials = (double *)(b + L.partials);
    // everything starts at zero: image, fields, norms, residuals
    if (hipMemsetAsync(j->block, 0, L.total, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        if (j->owns_block) (void)hipFree(j->block);
        delete j;
        return fail(MCSART_EHIP, "job initialisation failed");
    }
    *out = j;
    return MCSART_OK;
}

extern "C" int mcsart_job_create(mcsart_ctx *ctx, const mcsart_geom *gp, int n_views,
                                 const float *angles, const float *volumes, const float *flows,
                                 const mcsart_opts *opts, int n_iter, const int *op_iter,
                                 mcsart_job **out) {
    return job_create_in(ctx, gp, n_views, angles, volumes, flows, opts, n_iter, op_iter, nullptr,
                         0, out);
}

extern "C" int mcsart_job_set_projections(mcsart_job *j, const float *h_proj) {
    int rc = check_job(j);
    if (rc) return rc;
    if (!h_proj) return fail(MCSART_EINVAL, "null projections");
    MCSART_HIP_TRY(hipMemcpyAsync(j->d_proj, h_proj, sizeof(float) * j->g.ndet() * j->V,
                                  hipMemcpyHostToDevice, j->ctx->stream));
    MCSART_HIP_TRY(hipStreamSynchronize(j->ctx->stream));
    return MCSART_OK;
}

extern "C" int mcsart_job_set_projections_device(mcsart_job *j, const float *d_proj) {
    int rc = check_job(j);
    if (rc) return rc;
    if (!d_proj) return fail(MCSART_EINVAL, "null projections");
    MCSART_HIP_TRY(hipMemcpyAsync(j->d_proj, d_proj, sizeof(float) * j->g.ndet() * j->V,
                                  hipMemcpyDeviceToDevice, j->ctx->stream));
    return MCSART_OK;
}

extern "C" int mcsart_job_set_image(mcsart_job *j, const float *h_img) {
    int rc = check_job(j);
    if (rc) return rc;
    const size_t bytes = sizeof(float) * j->g.nvox();
    if (h_img)
        MCSART_HIP_TRY(hipMemcpyAsync(j->d_init, h_img, bytes, hipMemcpyHostToDevice, j->ctx->stream));
    else
        MCSART_HIP_TRY(hipMemsetAsync(j->d_init, 0, bytes, j->ctx->stream));
    MCSART_HIP_TRY(hipMemcpyAsync(j->d_img, j->d_init, bytes, hipMemcpyDeviceToDevice, j->ctx->stream));
    MCSART_HIP_TRY(hipStreamSynchronize(j->ctx->stream));
    return MCSART_OK;
}

extern "C" int mcsart_job_set_fields(mcsart_job *j, const float *const h_fields[6]) {
    int rc = check_job(j);
    if (rc) return rc;
    const size_t N = (size_t)j->g.nvox();
    float *init = j->d_init + N;
    for (int c = 0; c < 6; ++c) {
        if (h_fields && h_fields[c])
            MCSART_HIP_TRY(hipMemcpyAsync(init + c * N, h_fields[c], sizeof(float) * N,
                                          hipMemcpyHostToDevice, j->ctx->stream));
        else
            MCSART_HIP_TRY(hipMemsetAsync(init + c * N, 0, sizeof(float) * N, j->ctx->stream));
    }
    MCSART_HIP_TRY(hipMemcpyAsync(j->d_fields, init, sizeof(float) * N * 6,
                                  hipMemcpyDeviceToDevice, j->ctx->stream));
    MCSART_HIP_TRY(hipStreamSynchronize(j->ctx->stream));
    return MCSART_OK;
}

extern "C" int mcsart_job_reset(mcsart_job *j) {
    int rc = check_job(j);
    if (rc) return rc;
    const size_t N = (size_t)j->g.nvox();
    hipStream_t st = j->ctx->stream;
    MCSART_HIP_TRY(hipMemcpyAsync(j->d_img, j->d_init, sizeof(float) * N, hipMemcpyDeviceToDevice, st));
    MCSART_HIP_TRY(hipMemcpyAsync(j->d_fields, j->d_init + N, sizeof(float) * N * 6,
                                  hipMemcpyDeviceToDevice, st));
    rc = launch_fill(j->d_norm, (long long)j->n_iter * j->V, 0.f, st);
    if (rc) return rc;
    if (j->d_resid) rc = launch_fill(j->d_resid, j->g.ndet() * j->V, 0.f, st);
    return rc;
}

namespace {

void drop_graph(mcsart_job::SweepGraph &sg) {
    if (sg.exec) (void)hipGraphExecDestroy(sg.exec);
    if (sg.graph) (void)hipGraphDestroy(sg.graph);
    sg = mcsart_job::SweepGraph{};
}

int run_sweep_eager(mcsart_job *j, int it) {
    return (j->op_iter[it] == 1) ? image_sweep(j, it) : motion_sweep(j, it);
}

int capture_sweep(mcsart_job *j, int it) {
    mcsart_ctx *ctx = j->ctx;
    auto &sg = j->graphs[it];
    drop_graph(sg);
    MCSART_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    const int rc = run_sweep_eager(j, it);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    sg.graph = graph;
    if (rc || e != hipSuccess) {
        drop_graph(sg);
        return rc ? rc : fail(MCSART_EHIP, std::string("stream capture: ") + hipGetErrorString(e));
    }
    if (hipGraphInstantiate(&sg.exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        drop_graph(sg);
        return fail(MCSART_EHIP, "hipGraphInstantiate failed");
    }
    return MCSART_OK;
}

}  // namespace

namespace {

// First index in [0, n) where a predicate that is false..false true..true holds (n if none).
template <class P>
int first_true(int n, P pred) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (pred(mid)) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

struct ZRange {
    int lo, hi;   // [lo, hi)
};

// Number of z indices covered by the union of k ranges.
int union_length(ZRange *r, int k) {
    std::sort(r, r + k, [](const ZRange &a, const ZRange &b) { return a.lo < b.lo; });
    int total = 0, end = 0;
    for (int i = 0; i < k; ++i) {
        const int lo = std::max(r[i].lo, end), hi = r[i].hi;
        if (hi > lo) {
            total += hi - lo;
            end = hi;
        }
    }
    return total;
}

// The z ranges of column (ix, iy) in a read set (cone_keep), with the edge shell (cone_shell)
// when `shell`: up to three.
int read_set_ranges(const ConeCull &C, int ix, int iy, ZRange *out, bool shell) {
    const int nz = C.nz;
    if (shell && (ix < C.sx || ix >= C.nx - C.sx || iy < C.sx || iy >= C.ny - C.sx)) {
        out[0] = ZRange{0, nz};
        return 1;
    }
    const float x = (float)ix, y = (float)iy;
    if (!(cone_form(C, 0, x, y, 0.f) >= 0.f && cone_form(C, 1, x, y, 0.f) >= 0.f)) return 0;
    int k = 0;
    if (shell) {   // z edges, in the fan's columns
        out[k++] = ZRange{0, std::min(C.sz, nz)};
        out[k++] = ZRange{std::max(nz - C.sz, 0), nz};
    }
    // form 2 grows with z, form 3 falls
    const int lo = first_true(nz, [&](int iz) { return cone_form(C, 2, x, y, (float)iz) >= 0.f; });
    const int hi = first_true(nz, [&](int iz) { return cone_form(C, 3, x, y, (float)iz) < 0.f; });
    out[k++] = ZRange{lo, std::max(lo, hi)};
    return k;
}

// The z range of column (ix, iy) whose ratio backprojection has cells and rows (u can be
// nonzero): src/cu_backprojection.cu:150-169 with the bounds of column_setup / column_b.
ZRange hit_range(const GeomArgs &ga, double c, double s, int ix, int iy) {
    Column C;
    column_setup(ga, c, s, ix, iy, C);
    const double ia_lo = std::fmin(std::fmax(std::floor(C.Amin), 0.0), (double)ga.na);
    if (!(ia_lo < ga.na && C.Amax > ia_lo)) return ZRange{0, 0};
    auto bounds = [&](int iz, double &lo, double &hi) { column_b(C, ga.nz, iz, lo, hi); };
    // both bounds grow with z
    const int lo = first_true(ga.nz, [&](int iz) {
        double b0, b1;
        bounds(iz, b0, b1);
        return b1 > 0.0;
    });
    const int hi = first_true(ga.nz, [&](int iz) {
        double b0, b1;
        bounds(iz, b0, b1);
        return b0 >= (double)ga.nb;
    });
    return ZRange{lo, std::max(lo, hi)};
}

}  // namespace

extern "C" int mcsart_view_voxels(const mcsart_geom *gp, int n_views, const float *angles,
                                  long long *counts) {
    if (!gp || !angles || !counts) return fail(MCSART_EINVAL, "null argument");
    if (n_views <= 0) return fail(MCSART_EINVAL, "n_views must be positive");
    const Geom g = geom_of(gp);
    int rc = validate_geom(g);
    if (rc) return rc;
    const GeomArgs ga = geom_args(g);
    std::vector<ViewTrig> trig(n_views);
    std::vector<ConeCull> rs(n_views);
    std::vector<int> layouts(n_views);
    for (int v = 0; v < n_views; ++v) {
        trig[v] = view_trig(angles[v]);
        const ViewInfo vi = view_info_of(g, trig[v]);
        rs[v] = projector_read_set(ga, trig[v].c, trig[v].s, vi, 1);
        layouts[v] = (vi.has_y ? 1 : 0) | (vi.has_x ? 2 : 0);
    }
    auto count_view = [&](int v) {
        long long nread = 0, nprep = 0, nhit = 0, nunion = 0;
        for (int iy = 0; iy < g.ny; ++iy)
            for (int ix = 0; ix < g.nx; ++ix) {
                ZRange r[4];
                int k = read_set_ranges(rs[v], ix, iy, r, false);
                nread += union_length(r, k);
                k = read_set_ranges(rs[v], ix, iy, r, true);
                nprep += union_length(r, k);
                ZRange u[4];
                u[0] = hit_range(ga, trig[v].c, trig[v].s, ix, iy);
                nhit += u[0].hi - u[0].lo;
                const int kn =
                    v + 1 < n_views ? read_set_ranges(rs[v + 1], ix, iy, u + 1, true) : 0;
                nunion += union_length(u, 1 + kn);
            }
        long long *c = counts + 5 * (long long)v;
        c[0] = nread;
        c[1] = nprep;
        c[2] = nhit;
        c[3] = nunion;
        c[4] = layouts[v];
    };
    // views on up to 16 host threads (each writes its own entries)
    const int nt = std::max(1, std::min({16, n_views, (int)std::thread::hardware_concurrency()}));
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&, t] {
            for (int v = t; v < n_views; v += nt) count_view(v);
        });
    for (auto &th : pool) th.join();
    return MCSART_OK;
}

extern "C" int mcsart_job_set_culling(mcsart_job *j, int mode) {
    int rc = check_job(j);
    if (rc) return rc;
    if (mode < 0 || mode > 2) return fail(MCSART_EINVAL, "culling mode must be 0, 1 or 2");
    if (mode != j->cull)
        for (auto &sg : j->graphs) drop_graph(sg);   // captured launches carry the old mode
    j->cull = mode;
    return MCSART_OK;
}

extern "C" int mcsart_job_set_graphs(mcsart_job *j, int enable) {
    int rc = check_job(j);
    if (rc) return rc;
    j->use_graphs = enable != 0;
    return MCSART_OK;
}

extern "C" int mcsart_job_run_sweep(mcsart_job *j, int it) {
    int rc = check_job(j);
    if (rc) return rc;
    if (it < 0 || it >= j->n_iter) return fail(MCSART_EINVAL, "sweep index out of range");
    mcsart_ctx *ctx = j->ctx;
    if (!j->use_graphs || ctx->timing) return run_sweep_eager(j, it);
    auto &sg = j->graphs[it];
    if (!sg.exec) {
        rc = capture_sweep(j, it);
        if (rc) return rc;
    }
    MCSART_HIP_TRY(hipGraphLaunch(sg.exec, ctx->stream));
    return MCSART_OK;
}

extern "C" int mcsart_job_run(mcsart_job *j) {
    int rc = check_job(j);
    if (rc) return rc;
    for (int it = 0; it < j->n_iter; ++it) {
        rc = mcsart_job_run_sweep(j, it);
        if (rc) return rc;
    }
    return MCSART_OK;
}

extern "C" int mcsart_job_get_image(mcsart_job *j, float *h_img) {
    int rc = check_job(j);
    if (rc) return rc;
    MCSART_HIP_TRY(hipStreamSynchronize(j->ctx->stream));
    MCSART_HIP_TRY(hipMemcpy(h_img, j->d_img, sizeof(float) * j->g.nvox(), hipMemcpyDeviceToHost));
    return MCSART_OK;
}

extern "C" int mcsart_job_get_norms(mcsart_job *j, float *h_norm) {
    int rc = check_job(j);
    if (rc) return rc;
    MCSART_HIP_TRY(hipStreamSynchronize(j->ctx->stream));
    MCSART_HIP_TRY(hipMemcpy(h_norm, j->d_norm, sizeof(float) * (size_t)j->n_iter * j->V,
                             hipMemcpyDeviceToHost));
    return MCSART_OK;
}

extern "C" int mcsart_job_get_residual(mcsart_job *j, float *h_proj) {
    int rc = check_job(j);
    if (rc) return rc;
    if (!j->d_resid) return fail(MCSART_EINVAL, "job was created without keep_residual");
    MCSART_HIP_TRY(hipStreamSynchronize(j->ctx->stream));
    MCSART_HIP_TRY(hipMemcpy(h_proj, j->d_resid, sizeof(float) * j->g.ndet() * j->V,
                             hipMemcpyDeviceToHost));
    return MCSART_OK;
}

extern "C" int mcsart_job_get_fields(mcsart_job *j, float *const h_fields[6]) {
    int rc = check_job(j);
    if (rc) return rc;
    MCSART_HIP_TRY(hipStreamSynchronize(j->ctx->stream));
    const size_t N = (size_t)j->g.nvox();
    for (int c = 0; c < 6; ++c)
        if (h_fields[c])
            MCSART_HIP_TRY(hipMemcpy(h_fields[c], j->d_fields + c * N, sizeof(float) * N,
                                     hipMemcpyDeviceToHost));
    return MCSART_OK;
}

extern "C" int mcsart_job_device_image(mcsart_job *j, float **d_img) {
    int rc = check_job(j);
    if (rc) return rc;
    if (!d_img) return fail(MCSART_EINVAL, "null pointer");
    *d_img = j->d_img;
    return MCSART_OK;
}

namespace {

int check_tuning(const mcsart_tuning &t) {
    if (t.proj_split < 0 || t.proj_split > 16 || (t.proj_split & (t.proj_split - 1)))
        return fail(MCSART_EINVAL, "proj_split must be 0, 1, 2, 4, 8 or 16");
    if (t.proj_rows != 0 && t.proj_rows != 1 && t.proj_rows != 2 && t.proj_rows != 4)
        return fail(MCSART_EINVAL, "proj_rows must be 0, 1, 2 or 4");
    if (t.back_zslab < 0 || t.back_zslab > 64)
        return fail(MCSART_EINVAL, "back_zslab must be in 0..64");
    if (t.back_stage < 0 || t.back_stage > 2)
        return fail(MCSART_EINVAL, "back_stage must be 0, 1 or 2");
    if (t.back_zslab_motion < 0 || t.back_zslab_motion > 64)
        return fail(MCSART_EINVAL, "back_zslab_motion must be in 0..64");
    return MCSART_OK;
}

void apply_tuning(mcsart_job *j, const mcsart_tuning &t) {
    if (j->proj_split != t.proj_split || j->proj_rows != t.proj_rows ||
        j->back_tune.zslab != t.back_zslab || j->back_tune.stage != t.back_stage ||
        j->back_zslab_motion != t.back_zslab_motion)
        for (auto &sg : j->graphs) drop_graph(sg);   // captured launches carry the old shapes
    j->proj_split = t.proj_split;
    j->proj_rows = t.proj_rows;
    j->back_tune = BackTune{t.back_zslab, t.back_stage};
    j->back_zslab_motion = t.back_zslab_motion;
}

// Elapsed milliseconds of `fn` enqueued `reps` times on the job's stream.
template <class F>
int time_on_stream(mcsart_job *j, int reps, F &&fn, float *ms) {
    hipStream_t st = j->ctx->stream;
    hipEvent_t a = nullptr, b = nullptr;
    MCSART_HIP_TRY(hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return fail(MCSART_EHIP, "hipEventCreate failed");
    }
    int rc = MCSART_OK;
    if (hipEventRecord(a, st) != hipSuccess) rc = fail(MCSART_EHIP, "hipEventRecord failed");
    for (int r = 0; r < reps && !rc; ++r) rc = fn();
    if (!rc && (hipEventRecord(b, st) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                hipEventElapsedTime(ms, a, b) != hipSuccess))
        rc = fail(MCSART_EHIP, "event timing failed");
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}

}  // namespace

extern "C" int mcsart_job_set_tuning(mcsart_job *j, const mcsart_tuning *t) {
    int rc = check_job(j);
    if (rc) return rc;
    if (!t) return fail(MCSART_EINVAL, "null tuning");
    rc = check_tuning(*t);
    if (rc) return rc;
    apply_tuning(j, *t);
    return MCSART_OK;
}

extern "C" int mcsart_job_get_tuning(mcsart_job *j, mcsart_tuning *t) {
    int rc = check_job(j);
    if (rc) return rc;
    if (!t) return fail(MCSART_EINVAL, "null tuning");
    *t = mcsart_tuning{j->proj_split, j->back_tune.zslab, j->back_tune.stage, j->proj_rows,
                       j->back_zslab_motion};
    return MCSART_OK;
}

extern "C" int mcsart_job_view_layouts(mcsart_job *j, int *layouts) {
    int rc = check_job(j);
    if (rc) return rc;
    if (!layouts) return fail(MCSART_EINVAL, "null layouts");
    for (int iv = 0; iv < j->V; ++iv)
        layouts[iv] = (j->vinfo[iv].has_y ? 1 : 0) | (j->vinfo[iv].has_x ? 2 : 0);
    return MCSART_OK;
}

extern "C" int mcsart_job_kernel_time(mcsart_job *j, int kind, int reps, double *avg_us) {
    int rc = check_job(j);
    if (rc) return rc;
    if (!avg_us || reps < 1 || kind < 0 || kind > 5)
        return fail(MCSART_EINVAL, "kind must be 0..5 (projector, backprojector, warp, warp-add, "
                                   "motion step, motion-sweep projector), reps >= 1");
    mcsart_ctx *ctx = j->ctx;
    if (ctx->timing) return fail(MCSART_EINVAL, "kernel timing while per-kernel events are on");
    MCSART_HIP_TRY(hipSetDevice(ctx->device));
    const Geom &g = j->g;
    const long long D = g.ndet(), N = g.nvox();
    hipStream_t st = ctx->stream;
    MCSART_HIP_TRY(hipStreamSynchronize(st));
    rc = launch_transpose(j->d_w1T, j->d_img, g.nx, g.ny, g.nz, st);   // scratch copy
    if (rc) return rc;
    // warp-add updates a scratch copy of the image; the motion step works on the sweep's scratch
    // accumulators and constants (made here as a motion sweep makes them)
    if (kind == 3)
        MCSART_HIP_TRY(hipMemcpyAsync(j->d_w1, j->d_img, sizeof(float) * N,
                                      hipMemcpyDeviceToDevice, st));
    const Fields f = fields_of(j);
    float2 *pair = (float2 *)j->d_w1, *pairT = (float2 *)j->d_w1T;
    if (kind == 4) {
        if (j->V < 3) return fail(MCSART_EINVAL, "the motion step needs at least three views");
        rc = launch_motion_begin(f, j->d_img, j->d_cst, pair, pairT, g.nx, g.ny, g.nz, 0.f, 0.f,
                                 st);
        if (rc) return rc;
        MCSART_HIP_TRY(hipMemsetAsync(j->d_acc, 0, sizeof(float) * 2 * N, st));
    }
    // every view of the scan once per rep, back to back, as the sweeps launch them
    auto all_views = [&]() {
        for (int iv = 0; iv < j->V; ++iv) {
            const ViewInfo &vi = j->vinfo[iv];
            int r;
            if (kind == 2) {   // x_i = W(x; v_i - v_0, f_i - f_0) in the layouts P reads
                const ConeCull rs = read_set(j, iv);
                r = launch_warp(vi.has_y ? j->d_wi : nullptr, vi.has_x ? j->d_wiT : nullptr,
                                j->d_img, g.nx, g.ny, g.nz, j->volumes[iv] - j->volumes[0],
                                j->flows[iv] - j->flows[0], f, st, &rs);
            } else if (kind == 3) {   // x += lambda W(u; v_0 - v_i, f_0 - f_i), on the copy
                const ConeCull sup = u_support(j, iv);
                r = launch_warp_add(j->d_w1, j->d_w2, g.nx, g.ny, g.nz,
                                    j->volumes[0] - j->volumes[iv], j->flows[0] - j->flows[iv],
                                    f, j->opts.lambda, j->opts.nonneg, st, &sup);
            } else if (kind == 4) {   // ratio backprojection of view iv with its update and
                                      // the prep of view iv + 1 (views 1..V-2)
                if (iv == 0 || iv + 1 >= j->V) continue;
                const MotionStep m = motion_step(j, iv, iv + 1, pair, pairT, 0);
                const BackTune bt{j->back_zslab_motion, j->back_tune.stage};
                r = launch_backprojector(g, j->trig[iv], vi, BACK_RATIO, nullptr, j->d_rp,
                                         nullptr, st, nullptr, &bt, &m);
            } else if (kind == 5) {   // projector of the motion sweep: reads (g, Q) pairs
                ProjArgs a{};
                a.pair = pair;
                a.pairT = vi.has_x ? pairT : nullptr;
                a.qscale = -(j->volumes[iv] - j->volumes[iv > 0 ? iv - 1 : 0]) +
                           -(j->flows[iv] - j->flows[iv > 0 ? iv - 1 : 0]);
                a.meas = j->d_proj + D * iv;
                a.rp = j->d_rp;
                a.partials = j->d_partials;
                int np = 0;
                r = launch_projector(g, j->trig[iv], vi, PROJ_MOTION, a, st, &np, j->proj_split,
                                     j->proj_rows);
            } else if (kind == 0) {
                ProjArgs a{};
                a.vol0 = j->d_img;
                a.vol0T = j->vinfo[iv].has_x ? j->d_w1T : nullptr;
                a.meas = j->d_proj + D * iv;
                a.rp = j->d_rp;
                a.partials = j->d_partials;
                int np = 0;
                r = launch_projector(g, j->trig[iv], j->vinfo[iv], PROJ_IMAGE, a, st, &np,
                                     j->proj_split, j->proj_rows);
            } else {
                r = launch_backprojector(g, j->trig[iv], j->vinfo[iv], BACK_RATIO, nullptr,
                                         j->d_rp, j->d_w2, st, nullptr, &j->back_tune);
            }
            if (r) return r;
        }
        return (int)MCSART_OK;
    };
    rc = all_views();   // warm-up
    float ms = 0.f;
    if (!rc) rc = time_on_stream(j, reps, all_views, &ms);
    if (rc) return rc;
    const int launches = kind == 4 ? j->V - 2 : j->V;
    *avg_us = 1e3 * (double)ms / ((double)reps * launches);
    return MCSART_OK;
}

extern "C" int mcsart_job_autotune(mcsart_job *j, mcsart_tuning *chosen) {
    int rc = check_job(j);
    if (rc) return rc;
    mcsart_ctx *ctx = j->ctx;
    if (ctx->timing) return fail(MCSART_EINVAL, "autotune while kernel timing is enabled");
    MCSART_HIP_TRY(hipSetDevice(ctx->device));
    const Geom &g = j->g;
    const long long D = g.ndet();
    hipStream_t st = ctx->stream;
    // up to sixteen views spread over the scan at the midpoints of equal shares, so that a full
    // circle is not sampled at the axis-aligned angles alone: footprint shapes, window sizes and
    // the rays' branches change with the angle
    std::vector<int> views;
    const int nv = std::min(j->V, 16);
    for (int k = 0; k < nv; ++k) views.push_back((int)((2LL * k + 1) * j->V / (2LL * nv)));
    // transposed image for x-branch rays, into the job's scratch
    MCSART_HIP_TRY(hipStreamSynchronize(st));
    rc = launch_transpose(j->d_w1T, j->d_img, g.nx, g.ny, g.nz, st);
    if (rc) return rc;
    auto project = [&](int split, int rows) {
        for (int iv : views) {
            ProjArgs a{};
            a.vol0 = j->d_img;
            a.vol0T = j->vinfo[iv].has_x ? j->d_w1T : nullptr;
            a.meas = j->d_proj + D * iv;
            a.rp = j->d_rp;
            a.partials = j->d_partials;
            int np = 0;
            const int r = launch_projector(g, j->trig[iv], j->vinfo[iv], PROJ_IMAGE, a, st, &np,
                                           split, rows);
            if (r) return r;
        }
        return (int)MCSART_OK;
    };
    auto backproject = [&](const BackTune &bt) {
        for (int iv : views) {
            const int r = launch_backprojector(g, j->trig[iv], j->vinfo[iv], BACK_RATIO, nullptr,
                                               j->d_rp, j->d_w2, st, nullptr, &bt);
            if (r) return r;
        }
        return (int)MCSART_OK;
    };
    const int reps = 3;
    mcsart_tuning best{0, 0, 0, 0, 0};
    float best_ms = 3.4e38f;
    const int np = std::max(g.nx, g.ny);
    for (int rows = 1; rows <= 4; rows *= 2)
        for (int split = 1; split <= 16; split *= 2) {
            if (split > 1 && np / split < 4) break;
            rc = project(split, rows);   // warm-up (code object load, caches)
            if (rc) return rc;
            float ms = 0.f;
            rc = time_on_stream(j, reps, [&] { return project(split, rows); }, &ms);
            if (rc) return rc;
            if (ms < best_ms) {
                best_ms = ms;
                best.proj_split = split;
                best.proj_rows = rows;
            }
        }
    best_ms = 3.4e38f;
    for (int stage = 1; stage <= 2; ++stage)
        for (int zs = 1; zs <= 16 && zs <= g.nz; zs *= 2) {
            const BackTune bt{zs, stage};
            rc = backproject(bt);
            if (rc) return rc;
            float ms = 0.f;
            rc = time_on_stream(j, reps, [&] { return backproject(bt); }, &ms);
            if (rc) return rc;
            if (ms < best_ms) {
                best_ms = ms;
                best.back_zslab = zs;
                best.back_stage = stage;
            }
        }
    // the motion sweep's backprojector (fused update and next prep) on the same views (view 1
    // at the least: it needs a previous view), into the sweep's scratch accumulators and preps
    if (j->V >= 2) {
        const long long N = g.nvox();
        float2 *pair = (float2 *)j->d_w1, *pairT = (float2 *)j->d_w1T;
        rc = launch_motion_begin(fields_of(j), j->d_img, j->d_cst, pair, pairT, g.nx, g.ny, g.nz,
                                 0.f, 0.f, st);
        if (rc) return rc;
        MCSART_HIP_TRY(hipMemsetAsync(j->d_acc, 0, sizeof(float) * 2 * N, st));
        auto motion = [&](const BackTune &bt) {
            for (int iv0 : views) {
                // as the sweep launches it (culled), preparing the next view (or the view
                // itself at the end of the scan)
                const int iv = std::max(iv0, 1), nv = std::min(iv + 1, j->V - 1);
                const MotionStep m = motion_step(j, iv, nv, pair, pairT, 0);
                const int r = launch_backprojector(g, j->trig[iv], j->vinfo[iv], BACK_RATIO,
                                                   nullptr, j->d_rp, nullptr, st, nullptr, &bt,
                                                   &m);
                if (r) return r;
            }
            return (int)MCSART_OK;
        };
        best_ms = 3.4e38f;
        for (int zs = 1; zs <= 16 && zs <= g.nz; zs *= 2) {
            const BackTune bt{zs, best.back_stage};
            rc = motion(bt);
            if (rc) return rc;
            float ms = 0.f;
            rc = time_on_stream(j, reps, [&] { return motion(bt); }, &ms);
            if (rc) return rc;
            if (ms < best_ms) {
                best_ms = ms;
                best.back_zslab_motion = zs;
            }
        }
    }
    apply_tuning(j, best);
    if (chosen) *chosen = best;
    return MCSART_OK;
}

extern "C" int mcsart_job_destroy(mcsart_job *j) {
    if (!j) return MCSART_OK;
    (void)hipSetDevice(j->ctx->device);
    (void)hipStreamSynchronize(j->ctx->stream);
    for (auto &sg : j->graphs) drop_graph(sg);
    if (j->owns_block) (void)hipFree(j->block);
    delete j;
    return MCSART_OK;
}

// ---------------------------------------------------------------------------------------------
// host_AUMISART drop-ins (include/AUMISART.h:17,19)
// ---------------------------------------------------------------------------------------------
namespace {

int run_aumisart(float *h_outimg, float *h_outproj, float *h_outnorm, float *h_outalphax,
                 const float *h_img, const float *h_proj, int nx, int ny, int nz, int na, int nb,
                 int n_views, int n_iter, const int *op_iter, float da, float db, float ai,
                 float bi, float SO, float SD, float lambda, const float *volumes,
                 const float *flows, const float *angles, const float *const fields[6],
                 int motion) {
    if (!h_outimg || !h_outproj || !h_outnorm || !h_outalphax || !h_img || !h_proj || !volumes ||
        !flows || !angles)
        return fail(MCSART_EINVAL, "null host pointer");
    mcsart_ctx *ctx = default_ctx();
    if (!ctx) return fail(MCSART_EHIP, "no HIP device / context");
    const mcsart_geom g{nx, ny, nz, na, nb, da, db, ai, bi, SO, SD};
    mcsart_opts o{};
    o.motion = motion;
    o.nonneg = 0;   // the reference applies no clamp (src/AUMISART.cu:62,223)
    o.keep_residual = 1;
    o.lambda = lambda;
    mcsart_job *j = nullptr;
    int rc = mcsart_job_create(ctx, &g, n_views, angles, volumes, flows, &o, n_iter, op_iter, &j);
    if (rc) return rc;
    if (!(rc = mcsart_job_set_projections(j, h_proj)) && !(rc = mcsart_job_set_image(j, h_img)) &&
        !(rc = mcsart_job_set_fields(j, fields)) && !(rc = mcsart_job_run(j)) &&
        !(rc = mcsart_job_get_image(j, h_outimg)) && !(rc = mcsart_job_get_norms(j, h_outnorm)) &&
        !(rc = mcsart_job_get_residual(j, h_outproj))) {
        float *fo[6] = {h_outalphax, nullptr, nullptr, nullptr, nullptr, nullptr};
        rc = mcsart_job_get_fields(j, fo);
    }
    mcsart_job_destroy(j);
    return rc;
}

}  // namespace

extern "C" int mcsart_aumisart(float *h_outimg, float *h_outproj, float *h_outnorm,
                               float *h_outalphax, const float *h_img, const float *h_proj, int nx,
                               int ny, int nz, int na, int nb, int outIter, int n_views,
                               int n_iter, const int *op_iter, float da, float db, float ai,
                               float bi, float SO, float SD, float dx, float lambda,
                               const float *volumes, const float *flows, const float *err_weights,
                               const float *angles) {
    (void)outIter;
    (void)dx;
    (void)err_weights;   // parsed but unused by the reference
    return run_aumisart(h_outimg, h_outproj, h_outnorm, h_outalphax, h_img, h_proj, nx, ny, nz,
                        na, nb, n_views, n_iter, op_iter, da, db, ai, bi, SO, SD, lambda, volumes,
                        flows, angles, nullptr, 0);
}

extern "C" int mcsart_aumisart_motion(float *h_outimg, float *h_outproj, float *h_outnorm,
                                      float *h_outalphax, const float *h_img, const float *h_proj,
                                      int nx, int ny, int nz, int na, int nb, int outIter,
                                      int n_views, int n_iter, const int *op_iter, float da,
                                      float db, float ai, float bi, float SO, float SD, float dx,
                                      float lambda, const float *volumes, const float *flows,
                                      const float *err_weights, const float *angles,
                                      const float *ax, const float *ay, const float *az,
                                      const float *bx, const float *by, const float *bz) {
    (void)outIter;
    (void)dx;
    (void)err_weights;
    if (!ax || !ay || !az || !bx || !by || !bz) return fail(MCSART_EINVAL, "null field");
    const float *const f[6] = {ax, ay, az, bx, by, bz};
    return run_aumisart(h_outimg, h_outproj, h_outnorm, h_outalphax, h_img, h_proj, nx, ny, nz,
                        na, nb, n_views, n_iter, op_iter, da, db, ai, bi, SO, SD, lambda, volumes,
                        flows, angles, f, 1);
}

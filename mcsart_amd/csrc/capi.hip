// C ABI of libmcsart_hip: status handling, device memory, contexts, timing and the synchronous
// operator entry points that replace the reference's host_* functions (include/cu_*.h).
#include <cstring>
#include <mutex>
#include <string>

#include "internal.hpp"

namespace mcsart {

namespace {
thread_local std::string g_last_error;
std::mutex g_ctx_mutex;
mcsart_ctx *g_default_ctx[64] = {nullptr};
}  // namespace

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int status, const std::string &msg) {
    g_last_error = msg;
    return status;
}

int validate_volume(int nx, int ny, int nz, bool need_square) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return fail(MCSART_EINVAL, "volume sizes must be positive");
    if ((long long)nx * ny * nz >= (1LL << 31))
        return fail(MCSART_EINVAL, "volume larger than 2^31 voxels");
    if (need_square && nx != ny)
        return fail(MCSART_EINVAL,
                    "nx == ny is required (the reference indexes dim 1 with the dim-2 thread "
                    "coordinate: src/cu_deform.cu:85, src/cu_add.cu:58)");
    return MCSART_OK;
}

int validate_geom(const Geom &g) {
    int rc = validate_volume(g.nx, g.ny, g.nz, false);
    if (rc) return rc;
    if (g.na <= 0 || g.nb <= 0) return fail(MCSART_EINVAL, "detector sizes must be positive");
    if ((long long)g.na * g.nb >= (1LL << 31)) return fail(MCSART_EINVAL, "detector too large");
    if (!(g.SO > 0.f) || !(g.SD > g.SO))
        return fail(MCSART_EINVAL, "geometry needs 0 < SO < SD");
    if (!(g.da > 0.f) || !(g.db > 0.f)) return fail(MCSART_EINVAL, "detector pitch must be > 0");
    const double half_diag = 0.5 * sqrt((double)g.nx * g.nx + (double)g.ny * g.ny);
    if (g.SO <= half_diag)
        return fail(MCSART_EINVAL, "source orbit must lie outside the volume (SO > half diagonal)");
    return MCSART_OK;
}

ViewTrig view_trig(float angle) {
    const double a = (double)angle;
    return ViewTrig{cos(a), sin(a)};
}

mcsart_ctx *default_ctx() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    if (dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lk(g_ctx_mutex);
    if (!g_default_ctx[dev]) {
        mcsart_ctx *c = nullptr;
        if (mcsart_ctx_create(&c, dev) != MCSART_OK) return nullptr;
        g_default_ctx[dev] = c;
    }
    return g_default_ctx[dev];
}

int ctx_scratch(mcsart_ctx *ctx, size_t bytes, void **ptr) {
    if (ctx->scratch_bytes < bytes) {
        if (ctx->scratch) {
            MCSART_HIP_TRY(hipStreamSynchronize(ctx->stream));
            MCSART_HIP_TRY(hipFree(ctx->scratch));
            ctx->scratch = nullptr;
            ctx->scratch_bytes = 0;
        }
        if (hipMalloc(&ctx->scratch, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return fail(MCSART_ENOMEM, "scratch allocation failed");
        }
        ctx->scratch_bytes = bytes;
    }
    *ptr = ctx->scratch;
    return MCSART_OK;
}

// The transposed copy of a volume for the projector's x-branch rays, in the context's scratch.
// The rows projector addresses both layouts from one base with 32-bit byte offsets, so when the
// caller's volume lies too far from the scratch, a copy of it goes beside the transposed one.
static int stage_transposed(mcsart_ctx *ctx, const float *vol, int nx, int ny, int nz,
                            const float **nat, const float **tr) {
    const size_t bytes = sizeof(float) * (size_t)nx * ny * nz;
    void *tmp = nullptr;
    int rc = ctx_scratch(ctx, 2 * bytes, &tmp);
    if (rc) return rc;
    float *T = (float *)tmp;
    rc = launch_transpose(T, vol, nx, ny, nz, ctx->stream);
    if (rc) return rc;
    const uintptr_t a = (uintptr_t)vol, b = (uintptr_t)T;
    *tr = T;
    *nat = vol;
    if ((a > b ? a - b : b - a) + bytes > (1ull << 32)) {
        float *copy = (float *)((char *)tmp + bytes);
        MCSART_HIP_TRY(hipMemcpyAsync(copy, vol, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        *nat = copy;
    }
    return MCSART_OK;
}

int ctx_reduction(mcsart_ctx *ctx, int nblocks, double **partials) {
    const size_t need = sizeof(double) * 2 * (size_t)nblocks;
    if (ctx->red_bytes < need) {
        if (ctx->red) {
            MCSART_HIP_TRY(hipStreamSynchronize(ctx->stream));
            MCSART_HIP_TRY(hipFree(ctx->red));
            ctx->red = nullptr;
            ctx->red_bytes = 0;
        }
        if (hipMalloc(&ctx->red, need) != hipSuccess) {
            (void)hipGetLastError();
            return fail(MCSART_ENOMEM, "reduction scratch allocation failed");
        }
        ctx->red_bytes = need;
    }
    *partials = (double *)ctx->red;
    return MCSART_OK;
}

TimedRegion::TimedRegion(mcsart_ctx *c, int k) : ctx(c), kind(k) {
    if (!ctx || !ctx->timing) return;
    auto take = [&]() {
        hipEvent_t e = nullptr;
        if (!ctx->event_pool.empty()) {
            e = ctx->event_pool.back();
            ctx->event_pool.pop_back();
        } else if (hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
        }
        return e;
    };
    a = take();
    b = take();
    if (a && b) (void)hipEventRecord(a, ctx->stream);
}

TimedRegion::~TimedRegion() {
    if (!ctx || !ctx->timing || !a || !b) return;
    (void)hipEventRecord(b, ctx->stream);
    ctx->pending.push_back({kind, a, b});
    if (ctx->pending.size() >= 4096) ctx_collect_timing(ctx);
}

int ctx_collect_timing(mcsart_ctx *ctx) {
    for (auto &p : ctx->pending) {
        MCSART_HIP_TRY(hipEventSynchronize(p.b));
        float ms = 0.f;
        MCSART_HIP_TRY(hipEventElapsedTime(&ms, p.a, p.b));
        ctx->total_ms[p.kind] += ms;
        ctx->launches[p.kind] += 1;
        ctx->event_pool.push_back(p.a);
        ctx->event_pool.push_back(p.b);
    }
    ctx->pending.clear();
    return MCSART_OK;
}

}  // namespace mcsart

using namespace mcsart;

// ---------------------------------------------------------------------------------------------
// 1. status, memory, contexts
// ---------------------------------------------------------------------------------------------
extern "C" const char *mcsart_status_string(int status) {
    switch (status) {
        case MCSART_OK: return "ok";
        case MCSART_EINVAL: return "invalid argument";
        case MCSART_ENOMEM: return "out of device memory";
        case MCSART_EHIP: return "HIP runtime error";
        case MCSART_ENOTSUP: return "not supported";
        case MCSART_EBUSY: return "busy";
        case MCSART_ECANCELED: return "cancelled";
        default: return "unknown status";
    }
}

extern "C" const char *mcsart_last_error(void) { return g_last_error.c_str(); }

extern "C" int mcsart_version(void) { return 10000; }   // 1.0.0

extern "C" int mcsart_device_count(int *count) {
    if (!count) return fail(MCSART_EINVAL, "null count");
    MCSART_HIP_TRY(hipGetDeviceCount(count));
    return MCSART_OK;
}

extern "C" int mcsart_set_device(int device) {
    MCSART_HIP_TRY(hipSetDevice(device));
    return MCSART_OK;
}

extern "C" int mcsart_malloc(void **dptr, size_t bytes) {
    if (!dptr) return fail(MCSART_EINVAL, "null pointer");
    *dptr = nullptr;
    if (bytes == 0) return MCSART_OK;
    if (hipMalloc(dptr, bytes) != hipSuccess) {
        (void)hipGetLastError();
        *dptr = nullptr;
        return fail(MCSART_ENOMEM, "hipMalloc failed for " + std::to_string(bytes) + " bytes");
    }
    return MCSART_OK;
}

extern "C" int mcsart_free(void *dptr) {
    if (!dptr) return MCSART_OK;
    MCSART_HIP_TRY(hipFree(dptr));
    return MCSART_OK;
}

extern "C" int mcsart_memcpy_htod(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return MCSART_OK;
    if (!dst || !src) return fail(MCSART_EINVAL, "null pointer");
    MCSART_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return MCSART_OK;
}

extern "C" int mcsart_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return MCSART_OK;
    if (!dst || !src) return fail(MCSART_EINVAL, "null pointer");
    MCSART_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return MCSART_OK;
}

extern "C" int mcsart_memcpy_dtod(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return MCSART_OK;
    if (!dst || !src) return fail(MCSART_EINVAL, "null pointer");
    MCSART_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return MCSART_OK;
}

extern "C" int mcsart_memset(void *dptr, int value, size_t bytes) {
    if (bytes == 0) return MCSART_OK;
    if (!dptr) return fail(MCSART_EINVAL, "null pointer");
    MCSART_HIP_TRY(hipMemset(dptr, value, bytes));
    return MCSART_OK;
}

extern "C" int mcsart_device_synchronize(void) {
    MCSART_HIP_TRY(hipDeviceSynchronize());
    return MCSART_OK;
}

extern "C" int mcsart_mem_info(size_t *free_bytes, size_t *total_bytes) {
    size_t f = 0, t = 0;
    MCSART_HIP_TRY(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return MCSART_OK;
}

extern "C" int mcsart_ctx_create(mcsart_ctx **ctx, int device) {
    if (!ctx) return fail(MCSART_EINVAL, "null ctx");
    *ctx = nullptr;
    MCSART_HIP_TRY(hipSetDevice(device));
    auto *c = new mcsart_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(MCSART_EHIP, "hipStreamCreate failed");
    }
    *ctx = c;
    return MCSART_OK;
}

extern "C" int mcsart_ctx_destroy(mcsart_ctx *ctx) {
    if (!ctx) return MCSART_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &p : ctx->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->red) (void)hipFree(ctx->red);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return MCSART_OK;
}

extern "C" int mcsart_ctx_synchronize(mcsart_ctx *ctx) {
    if (!ctx) return fail(MCSART_EINVAL, "null ctx");
    MCSART_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MCSART_OK;
}

extern "C" int mcsart_ctx_enable_timing(mcsart_ctx *ctx, int enable) {
    if (!ctx) return fail(MCSART_EINVAL, "null ctx");
    ctx->timing = enable != 0;
    return MCSART_OK;
}

extern "C" int mcsart_ctx_kernel_time(mcsart_ctx *ctx, int kind, double *total_ms,
                                      long long *launches) {
    if (!ctx || kind < 0 || kind >= TK_COUNT) return fail(MCSART_EINVAL, "bad ctx or kind");
    int rc = ctx_collect_timing(ctx);
    if (rc) return rc;
    if (total_ms) *total_ms = ctx->total_ms[kind];
    if (launches) *launches = ctx->launches[kind];
    return MCSART_OK;
}

extern "C" int mcsart_ctx_reset_timing(mcsart_ctx *ctx) {
    if (!ctx) return fail(MCSART_EINVAL, "null ctx");
    int rc = ctx_collect_timing(ctx);
    if (rc) return rc;
    for (int k = 0; k < TK_COUNT; ++k) {
        ctx->total_ms[k] = 0;
        ctx->launches[k] = 0;
    }
    return MCSART_OK;
}

// ---------------------------------------------------------------------------------------------
// 2. synchronous operators (reference host_* argument order)
// ---------------------------------------------------------------------------------------------
namespace {

int finish(mcsart_ctx *ctx, int rc) {
    if (rc) return rc;
    MCSART_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MCSART_OK;
}

#define GET_CTX()                                                                    \
    mcsart_ctx *ctx = default_ctx();                                                 \
    if (!ctx) return fail(MCSART_EHIP, "no HIP device / context")

// Asynchronous operators run on the given context's stream (its device becomes current).
#define CHECK_CTX()                                                                  \
    if (!ctx) return fail(MCSART_EINVAL, "null context");                           \
    MCSART_HIP_TRY(hipSetDevice(ctx->device))

}  // namespace

extern "C" int mcsart_projection_async(mcsart_ctx *ctx, float *d_proj, const float *d_img,
                                       float angle, float SO, float SD, float da, int na, float ai,
                                       float db, int nb, float bi, int nx, int ny, int nz) {
    const Geom g{nx, ny, nz, na, nb, da, db, ai, bi, SO, SD};
    int rc = validate_geom(g);
    if (rc) return rc;
    if (!d_proj || !d_img) return fail(MCSART_EINVAL, "null device pointer");
    CHECK_CTX();
    const ViewTrig t = view_trig(angle);
    const ViewInfo vi = view_info_of(g, t);
    ProjArgs a{};
    a.vol0 = d_img;
    a.out = d_proj;
    if (vi.has_x) {   // x-branch rays read a transposed copy (contiguous footprints)
        rc = stage_transposed(ctx, d_img, nx, ny, nz, &a.vol0, &a.vol0T);
        if (rc) return rc;
    }
    TimedRegion tr(ctx, TK_PROJ);
    return launch_projector(g, t, vi, PROJ_PLAIN, a, ctx->stream);
}

extern "C" int mcsart_projection(float *d_proj, const float *d_img, float angle, float SO, float SD,
                                 float da, int na, float ai, float db, int nb, float bi, int nx,
                                 int ny, int nz) {
    GET_CTX();
    return finish(ctx, mcsart_projection_async(ctx, d_proj, d_img, angle, SO, SD, da, na, ai, db,
                                               nb, bi, nx, ny, nz));
}

extern "C" int mcsart_backprojection_async(mcsart_ctx *ctx, float *d_img, const float *d_proj,
                                           float angle, float SO, float SD, float da, int na,
                                           float ai, float db, int nb, float bi, int nx, int ny,
                                           int nz) {
    const Geom g{nx, ny, nz, na, nb, da, db, ai, bi, SO, SD};
    int rc = validate_geom(g);
    if (rc) return rc;
    if (!d_proj || !d_img) return fail(MCSART_EINVAL, "null device pointer");
    CHECK_CTX();
    const ViewTrig t = view_trig(angle);
    const ViewInfo vi = view_info_of(g, t);
    TimedRegion tr(ctx, TK_BACK);
    return launch_backprojector(g, t, vi, BACK_PLAIN, d_proj, nullptr, d_img,
                                            ctx->stream);
}

extern "C" int mcsart_backprojection(float *d_img, const float *d_proj, float angle, float SO,
                                     float SD, float da, int na, float ai, float db, int nb,
                                     float bi, int nx, int ny, int nz) {
    GET_CTX();
    return finish(ctx, mcsart_backprojection_async(ctx, d_img, d_proj, angle, SO, SD, da, na, ai,
                                                   db, nb, bi, nx, ny, nz));
}

extern "C" int mcsart_deform_async(mcsart_ctx *ctx, float *d_img1, const float *d_img, int nx,
                                   int ny, int nz, float volume, float flow, const float *alpha_x,
                                   const float *alpha_y, const float *alpha_z, const float *beta_x,
                                   const float *beta_y, const float *beta_z) {
    int rc = validate_volume(nx, ny, nz, true);
    if (rc) return rc;
    if (!d_img1 || !d_img || !alpha_x || !alpha_y || !alpha_z || !beta_x || !beta_y || !beta_z)
        return fail(MCSART_EINVAL, "null device pointer");
    CHECK_CTX();
    const Fields f{alpha_x, alpha_y, alpha_z, beta_x, beta_y, beta_z};
    const float *src = d_img;
    const long long n = (long long)nx * ny * nz;
    if (d_img1 == d_img) {   // in place: the reference samples a copy (src/cu_deform.cu:23)
        void *tmp = nullptr;
        rc = ctx_scratch(ctx, sizeof(float) * n, &tmp);
        if (rc) return rc;
        MCSART_HIP_TRY(hipMemcpyAsync(tmp, d_img, sizeof(float) * n, hipMemcpyDeviceToDevice,
                                      ctx->stream));
        src = (const float *)tmp;
    }
    TimedRegion tr(ctx, TK_WARP);
    return launch_warp(d_img1, nullptr, src, nx, ny, nz, volume, flow, f, ctx->stream);
}

extern "C" int mcsart_deform(float *d_img1, const float *d_img, int nx, int ny, int nz,
                             float volume, float flow, const float *alpha_x, const float *alpha_y,
                             const float *alpha_z, const float *beta_x, const float *beta_y,
                             const float *beta_z) {
    GET_CTX();
    return finish(ctx, mcsart_deform_async(ctx, d_img1, d_img, nx, ny, nz, volume, flow, alpha_x,
                                           alpha_y, alpha_z, beta_x, beta_y, beta_z));
}

extern "C" int mcsart_deform2_async(mcsart_ctx *ctx, float *d_img1, const float *d_img, int nx,
                                    int ny, int nz, float volume, float flow, const float *alpha_x,
                                    const float *alpha_y, const float *alpha_z, const float *beta_x,
                                    const float *beta_y, const float *beta_z) {
    int rc = validate_volume(nx, ny, nz, true);
    if (rc) return rc;
    if (!d_img1 || !d_img || !alpha_x || !alpha_y || !alpha_z || !beta_x || !beta_y || !beta_z)
        return fail(MCSART_EINVAL, "null device pointer");
    CHECK_CTX();
    const Fields f{alpha_x, alpha_y, alpha_z, beta_x, beta_y, beta_z};
    const float *src = d_img;
    const long long n = (long long)nx * ny * nz;
    if (d_img1 == d_img) {
        void *tmp = nullptr;
        rc = ctx_scratch(ctx, sizeof(float) * n, &tmp);
        if (rc) return rc;
        MCSART_HIP_TRY(hipMemcpyAsync(tmp, d_img, sizeof(float) * n, hipMemcpyDeviceToDevice,
                                      ctx->stream));
        src = (const float *)tmp;
    }
    return launch_deform2(d_img1, src, nx, ny, nz, volume, flow, f, ctx->stream);
}

extern "C" int mcsart_deform2(float *d_img1, const float *d_img, int nx, int ny, int nz,
                              float volume, float flow, const float *alpha_x, const float *alpha_y,
                              const float *alpha_z, const float *beta_x, const float *beta_y,
                              const float *beta_z) {
    GET_CTX();
    return finish(ctx, mcsart_deform2_async(ctx, d_img1, d_img, nx, ny, nz, volume, flow, alpha_x,
                                            alpha_y, alpha_z, beta_x, beta_y, beta_z));
}

extern "C" int mcsart_update_udvf_async(mcsart_ctx *ctx, float *alpha_x, float *alpha_y,
                                        float *alpha_z, float *beta_x, float *beta_y, float *beta_z,
                                        const float *img, const float *img0, float volume_diff,
                                        float flow_diff, int nx, int ny, int nz, int iter) {
    (void)iter;   // unused in the reference (src/cu_update_udvf.cu:11-62)
    int rc = validate_volume(nx, ny, nz, false);
    if (rc) return rc;
    if (!alpha_x || !alpha_y || !alpha_z || !beta_x || !beta_y || !beta_z || !img || !img0)
        return fail(MCSART_EINVAL, "null device pointer");
    CHECK_CTX();
    return launch_update_udvf(alpha_x, alpha_y, alpha_z, beta_x, beta_y, beta_z, img,
                                          img0, volume_diff, flow_diff, nx, ny, nz, ctx->stream);
}

extern "C" int mcsart_update_udvf(float *alpha_x, float *alpha_y, float *alpha_z, float *beta_x,
                                  float *beta_y, float *beta_z, const float *img, const float *img0,
                                  float volume_diff, float flow_diff, int nx, int ny, int nz,
                                  int iter) {
    GET_CTX();
    return finish(ctx, mcsart_update_udvf_async(ctx, alpha_x, alpha_y, alpha_z, beta_x, beta_y,
                                                beta_z, img, img0, volume_diff, flow_diff, nx, ny,
                                                nz, iter));
}

extern "C" int mcsart_add_async(mcsart_ctx *ctx, float *img1, const float *img, int nx, int ny,
                                int nz, float weight) {
    int rc = validate_volume(nx, ny, nz, false);
    if (rc) return rc;
    if (!img1 || !img) return fail(MCSART_EINVAL, "null device pointer");
    CHECK_CTX();
    return launch_add(img1, img, (long long)nx * ny * nz, weight, ctx->stream);
}

extern "C" int mcsart_add(float *img1, const float *img, int nx, int ny, int nz, float weight) {
    GET_CTX();
    return finish(ctx, mcsart_add_async(ctx, img1, img, nx, ny, nz, weight));
}

extern "C" int mcsart_add2_async(mcsart_ctx *ctx, float *img1, const float *img, int nx, int ny,
                                 int nz, const float *img0, float weight, int ind) {
    int rc = validate_volume(nx, ny, nz, true);
    if (rc) return rc;
    if (!img1 || !img || !img0) return fail(MCSART_EINVAL, "null device pointer");
    if (ind < 1 || ind > 3) return fail(MCSART_EINVAL, "ind must be 1, 2 or 3");
    if (img1 == img || img1 == img0)
        return fail(MCSART_EINVAL, "add2 output must not alias its inputs");
    CHECK_CTX();
    return launch_add2(img1, img, nx, ny, nz, img0, weight, ind, ctx->stream);
}

extern "C" int mcsart_add2(float *img1, const float *img, int nx, int ny, int nz, const float *img0,
                           float weight, int ind) {
    GET_CTX();
    return finish(ctx, mcsart_add2_async(ctx, img1, img, nx, ny, nz, img0, weight, ind));
}

extern "C" int mcsart_initial_async(mcsart_ctx *ctx, float *img, int nx, int ny, int nz,
                                    float value) {
    int rc = validate_volume(nx, ny, nz, false);
    if (rc) return rc;
    if (!img) return fail(MCSART_EINVAL, "null device pointer");
    CHECK_CTX();
    return launch_fill(img, (long long)nx * ny * nz, value, ctx->stream);
}

extern "C" int mcsart_initial(float *img, int nx, int ny, int nz, float value) {
    GET_CTX();
    return finish(ctx, mcsart_initial_async(ctx, img, nx, ny, nz, value));
}

extern "C" int mcsart_initial2_async(mcsart_ctx *ctx, float *img, int nx, int ny, int nz,
                                     const float *img0, float volume, float flow) {
    int rc = validate_volume(nx, ny, nz, false);
    if (rc) return rc;
    if (!img || !img0) return fail(MCSART_EINVAL, "null device pointer");
    if (img == img0) return fail(MCSART_EINVAL, "initial2 output must not alias its input");
    CHECK_CTX();
    return launch_initial2(img, nx, ny, nz, img0, volume, flow, ctx->stream);
}

extern "C" int mcsart_initial2(float *img, int nx, int ny, int nz, const float *img0, float volume,
                               float flow) {
    GET_CTX();
    return finish(ctx, mcsart_initial2_async(ctx, img, nx, ny, nz, img0, volume, flow));
}

extern "C" int mcsart_division_async(mcsart_ctx *ctx, float *img1, const float *img, int nx, int ny,
                                     int nz) {
    int rc = validate_volume(nx, ny, nz, false);
    if (rc) return rc;
    if (!img1 || !img) return fail(MCSART_EINVAL, "null device pointer");
    CHECK_CTX();
    return launch_division(img1, img, (long long)nx * ny * nz, ctx->stream);
}

extern "C" int mcsart_division(float *img1, const float *img, int nx, int ny, int nz) {
    GET_CTX();
    return finish(ctx, mcsart_division_async(ctx, img1, img, nx, ny, nz));
}

extern "C" int mcsart_diff_async(mcsart_ctx *ctx, float *img1, const float *img, int nx, int ny,
                                 int nz, int ind) {
    int rc = validate_volume(nx, ny, nz, false);
    if (rc) return rc;
    if (!img1 || !img) return fail(MCSART_EINVAL, "null device pointer");
    if (ind < 1 || ind > 3) return fail(MCSART_EINVAL, "ind must be 1, 2 or 3");
    if (img1 == img) return fail(MCSART_EINVAL, "diff output must not alias its input");
    CHECK_CTX();
    return launch_diff(img1, img, nx, ny, nz, ind, 1, ctx->stream);
}

extern "C" int mcsart_diff(float *img1, const float *img, int nx, int ny, int nz, int ind) {
    GET_CTX();
    return finish(ctx, mcsart_diff_async(ctx, img1, img, nx, ny, nz, ind));
}

extern "C" int mcsart_diff_forward_async(mcsart_ctx *ctx, float *img1, const float *img, int nx,
                                         int ny, int nz, int ind) {
    int rc = validate_volume(nx, ny, nz, false);
    if (rc) return rc;
    if (!img1 || !img) return fail(MCSART_EINVAL, "null device pointer");
    if (ind < 1 || ind > 3) return fail(MCSART_EINVAL, "ind must be 1, 2 or 3");
    if (img1 == img) return fail(MCSART_EINVAL, "diff output must not alias its input");
    CHECK_CTX();
    return launch_diff(img1, img, nx, ny, nz, ind, 0, ctx->stream);
}

extern "C" int mcsart_diff_forward(float *img1, const float *img, int nx, int ny, int nz, int ind) {
    GET_CTX();
    return finish(ctx, mcsart_diff_forward_async(ctx, img1, img, nx, ny, nz, ind));
}

// ---------------------------------------------------------------------------------------------
// fused SART steps
// ---------------------------------------------------------------------------------------------
extern "C" int mcsart_multiple_async(mcsart_ctx *ctx, float *img1, int nx, int ny, int nz,
                                     const float *img0, float weight, int ind) {
    int rc = validate_volume(nx, ny, nz, false);
    if (rc) return rc;
    if (!img1 || !img0) return fail(MCSART_EINVAL, "null device pointer");
    if (ind < 1 || ind > 3) return fail(MCSART_EINVAL, "ind must be 1, 2 or 3");
    CHECK_CTX();
    return launch_multiple(img1, img0, nx, ny, nz, weight, ind, ctx->stream);
}

extern "C" int mcsart_multiple(float *img1, int nx, int ny, int nz, const float *img0,
                               float weight, int ind) {
    GET_CTX();
    return finish(ctx, mcsart_multiple_async(ctx, img1, nx, ny, nz, img0, weight, ind));
}

extern "C" int mcsart_invert_dvf_async(mcsart_ctx *ctx, float *mx2, float *my2, float *mz2,
                                       const float *mx, const float *my, const float *mz, int nx,
                                       int ny, int nz, int niter) {
    int rc = validate_volume(nx, ny, nz, false);
    if (rc) return rc;
    if (!mx2 || !my2 || !mz2 || !mx || !my || !mz) return fail(MCSART_EINVAL, "null device pointer");
    if (niter < 0) return fail(MCSART_EINVAL, "niter must be >= 0");
    const float *out[3] = {mx2, my2, mz2}, *in[3] = {mx, my, mz};
    for (const float *o : out)
        for (const float *i : in)
            if (o == i) return fail(MCSART_EINVAL, "outputs must not alias inputs");
    CHECK_CTX();
    return launch_invert_dvf(mx2, my2, mz2, mx, my, mz, nx, ny, nz, niter, ctx->stream);
}

extern "C" int mcsart_invert_dvf(float *mx2, float *my2, float *mz2, const float *mx,
                                 const float *my, const float *mz, int nx, int ny, int nz,
                                 int niter) {
    GET_CTX();
    return finish(ctx, mcsart_invert_dvf_async(ctx, mx2, my2, mz2, mx, my, mz, nx, ny, nz, niter));
}

extern "C" int mcsart_generate_proj_async(mcsart_ctx *ctx, float *d_proj, const float *d_img,
                                          const float *alpha_x, const float *alpha_y,
                                          const float *alpha_z, const float *beta_x,
                                          const float *beta_y, const float *beta_z,
                                          const float *volumes, const float *flows,
                                          const float *angles, int n_views, int niter_inverse,
                                          float SO, float SD, float da, int na, float ai,
                                          float db, int nb, float bi, int nx, int ny, int nz) {
    const Geom g{nx, ny, nz, na, nb, da, db, ai, bi, SO, SD};
    int rc = validate_geom(g);
    if (rc) return rc;
    rc = validate_volume(nx, ny, nz, true);
    if (rc) return rc;
    if (!d_proj || !d_img || !alpha_x || !alpha_y || !alpha_z || !beta_x || !beta_y || !beta_z)
        return fail(MCSART_EINVAL, "null device pointer");
    if (!volumes || !flows || !angles || n_views <= 0)
        return fail(MCSART_EINVAL, "n_views and the per-view arrays are required");
    if (niter_inverse < 0) return fail(MCSART_EINVAL, "niter_inverse must be >= 0");
    CHECK_CTX();
    // scratch: forward DVF (3N), inverse DVF (3N), warped image (N), its transpose (N)
    const size_t N = (size_t)g.nvox();
    void *tmp = nullptr;
    rc = ctx_scratch(ctx, sizeof(float) * 8 * N, &tmp);
    if (rc) return rc;
    float *m = (float *)tmp, *mi = m + 3 * N, *w = m + 6 * N, *wT = m + 7 * N;
    const Fields f{alpha_x, alpha_y, alpha_z, beta_x, beta_y, beta_z};
    for (int iv = 0; iv < n_views; ++iv) {
        // DVF of the view's breathing state (kernel_forwardDVF)
        rc = launch_forward_dvf(m, m + N, m + 2 * N, f, nx, ny, nz, volumes[iv], flows[iv],
                                ctx->stream);
        if (rc) return rc;
        const float *u = m;
        if (niter_inverse > 0) {   // its fixed-point inverse (kernel_invertDVF)
            rc = launch_invert_dvf(mi, mi + N, mi + 2 * N, m, m + N, m + 2 * N, nx, ny, nz,
                                   niter_inverse, ctx->stream);
            if (rc) return rc;
            u = mi;
        }
        // warp by that DVF (kernel_deformation: components by name, dim 1 by my, dim 2 by mx)
        const Fields uf{u, u + N, u + 2 * N, u, u + N, u + 2 * N};
        const ViewTrig t = view_trig(angles[iv]);
        const ViewInfo vi = view_info_of(g, t);
        rc = launch_warp(vi.has_y ? w : nullptr, vi.has_x ? wT : nullptr, d_img, nx, ny, nz, 1.f,
                         0.f, uf, ctx->stream);
        if (rc) return rc;
        ProjArgs a{};
        a.vol0 = w;
        a.vol0T = vi.has_x ? wT : nullptr;
        a.out = d_proj + (size_t)iv * g.ndet();
        TimedRegion tr(ctx, TK_PROJ);
        rc = launch_projector(g, t, vi, PROJ_PLAIN, a, ctx->stream);
        if (rc) return rc;
    }
    return MCSART_OK;
}

extern "C" int mcsart_generate_proj(float *d_proj, const float *d_img, const float *alpha_x,
                                    const float *alpha_y, const float *alpha_z,
                                    const float *beta_x, const float *beta_y, const float *beta_z,
                                    const float *volumes, const float *flows, const float *angles,
                                    int n_views, int niter_inverse, float SO, float SD, float da,
                                    int na, float ai, float db, int nb, float bi, int nx, int ny,
                                    int nz) {
    GET_CTX();
    return finish(ctx, mcsart_generate_proj_async(ctx, d_proj, d_img, alpha_x, alpha_y, alpha_z,
                                                  beta_x, beta_y, beta_z, volumes, flows, angles,
                                                  n_views, niter_inverse, SO, SD, da, na, ai, db,
                                                  nb, bi, nx, ny, nz));
}

extern "C" int mcsart_sart_residual_async(mcsart_ctx *ctx, float *d_rp, float *d_norm,
                                          const float *d_vol, int motion_pairs, const float *d_meas,
                                          float angle, float SO, float SD, float da, int na,
                                          float ai, float db, int nb, float bi, int nx, int ny,
                                          int nz) {
    const Geom g{nx, ny, nz, na, nb, da, db, ai, bi, SO, SD};
    int rc = validate_geom(g);
    if (rc) return rc;
    if (!d_rp || !d_norm || !d_vol || !d_meas) return fail(MCSART_EINVAL, "null device pointer");
    CHECK_CTX();
    const ViewTrig t = view_trig(angle);
    const ViewInfo vi = view_info_of(g, t);
    ProjArgs a{};
    if (motion_pairs) {
        a.pair = (const float2 *)d_vol;
        a.qscale = 1.f;   // the pairs hold q itself
    } else {
        a.vol0 = d_vol;
        if (vi.has_x) {
            rc = stage_transposed(ctx, d_vol, nx, ny, nz, &a.vol0, &a.vol0T);
            if (rc) return rc;
        }
    }
    a.meas = d_meas;
    a.rp = (float2 *)d_rp;
    rc = ctx_reduction(ctx, projector_partials(g), &a.partials);
    if (rc) return rc;
    int nparts = 0;
    {
        TimedRegion tr(ctx, TK_PROJ);
        rc = launch_projector(g, t, vi, motion_pairs ? PROJ_MOTION : PROJ_IMAGE, a, ctx->stream,
                              &nparts);
        if (rc) return rc;
    }
    return launch_norm_finish(NormFinish{a.partials, nparts, d_norm}, ctx->stream);
}

extern "C" int mcsart_sart_residual(float *d_rp, float *d_norm, const float *d_vol,
                                    int motion_pairs, const float *d_meas, float angle, float SO,
                                    float SD, float da, int na, float ai, float db, int nb,
                                    float bi, int nx, int ny, int nz) {
    GET_CTX();
    return finish(ctx, mcsart_sart_residual_async(ctx, d_rp, d_norm, d_vol, motion_pairs, d_meas,
                                                  angle, SO, SD, da, na, ai, db, nb, bi, nx, ny,
                                                  nz));
}

extern "C" int mcsart_sart_backprojection_async(mcsart_ctx *ctx, float *d_u, const float *d_rp,
                                                float angle, float SO, float SD, float da, int na,
                                                float ai, float db, int nb, float bi, int nx,
                                                int ny, int nz) {
    const Geom g{nx, ny, nz, na, nb, da, db, ai, bi, SO, SD};
    int rc = validate_geom(g);
    if (rc) return rc;
    if (!d_u || !d_rp) return fail(MCSART_EINVAL, "null device pointer");
    CHECK_CTX();
    const ViewTrig t = view_trig(angle);
    const ViewInfo vi = view_info_of(g, t);
    TimedRegion tr(ctx, TK_BACK);
    return launch_backprojector(g, t, vi, BACK_RATIO, nullptr, (const float2 *)d_rp,
                                            d_u, ctx->stream);
}

extern "C" int mcsart_sart_backprojection(float *d_u, const float *d_rp, float angle, float SO,
                                          float SD, float da, int na, float ai, float db, int nb,
                                          float bi, int nx, int ny, int nz) {
    GET_CTX();
    return finish(ctx, mcsart_sart_backprojection_async(ctx, d_u, d_rp, angle, SO, SD, da, na, ai,
                                                        db, nb, bi, nx, ny, nz));
}

extern "C" int mcsart_motion_prep_async(mcsart_ctx *ctx, float *d_pairs, const float *d_img,
                                        const float *alpha_x, const float *alpha_y,
                                        const float *alpha_z, const float *beta_x,
                                        const float *beta_y, const float *beta_z, float vd,
                                        float fd, int nx, int ny, int nz) {
    int rc = validate_volume(nx, ny, nz, true);
    if (rc) return rc;
    if (!d_pairs || !d_img || !alpha_x || !alpha_y || !alpha_z || !beta_x || !beta_y || !beta_z)
        return fail(MCSART_EINVAL, "null device pointer");
    if (d_pairs == d_img) return fail(MCSART_EINVAL, "motion_prep output must not alias the image");
    CHECK_CTX();
    const float *f[6] = {alpha_x, alpha_y, alpha_z, beta_x, beta_y, beta_z};
    TimedRegion tr(ctx, TK_MOTION);
    return launch_motion_prep((float2 *)d_pairs, nullptr, d_img, f, nx, ny, nz, vd, fd,
                                          ctx->stream);
}

extern "C" int mcsart_motion_prep(float *d_pairs, const float *d_img, const float *alpha_x,
                                  const float *alpha_y, const float *alpha_z, const float *beta_x,
                                  const float *beta_y, const float *beta_z, float vd, float fd,
                                  int nx, int ny, int nz) {
    GET_CTX();
    return finish(ctx, mcsart_motion_prep_async(ctx, d_pairs, d_img, alpha_x, alpha_y, alpha_z,
                                                beta_x, beta_y, beta_z, vd, fd, nx, ny, nz));
}

extern "C" int mcsart_motion_update_async(mcsart_ctx *ctx, float *alpha_x, float *alpha_y,
                                          float *alpha_z, float *beta_x, float *beta_y,
                                          float *beta_z, const float *d_u, const float *d_img,
                                          float wv, float wf, int nx, int ny, int nz) {
    int rc = validate_volume(nx, ny, nz, true);
    if (rc) return rc;
    if (!alpha_x || !alpha_y || !alpha_z || !beta_x || !beta_y || !beta_z || !d_u || !d_img)
        return fail(MCSART_EINVAL, "null device pointer");
    float *f[6] = {alpha_x, alpha_y, alpha_z, beta_x, beta_y, beta_z};
    for (int c = 0; c < 6; ++c)
        if (f[c] == d_u || f[c] == d_img)
            return fail(MCSART_EINVAL, "motion_update fields must not alias u or the image");
    CHECK_CTX();
    TimedRegion tr(ctx, TK_MOTION);
    return launch_motion_update(f, d_u, d_img, nx, ny, nz, wv, wf, ctx->stream);
}

extern "C" int mcsart_motion_update(float *alpha_x, float *alpha_y, float *alpha_z, float *beta_x,
                                    float *beta_y, float *beta_z, const float *d_u,
                                    const float *d_img, float wv, float wf, int nx, int ny,
                                    int nz) {
    GET_CTX();
    return finish(ctx, mcsart_motion_update_async(ctx, alpha_x, alpha_y, alpha_z, beta_x, beta_y,
                                                  beta_z, d_u, d_img, wv, wf, nx, ny, nz));
}

/*
 * libmcsart_hip -- MI355X (gfx950) implementation of the McSART motion-compensated SART loop.
 *
 * C ABI.  Plain pointers and sizes only.  Every function returns an int status (MCSART_OK == 0,
 * negative on error); mcsart_last_error() gives a thread-local message for the last failure.
 *
 * Array layout is the reference's MATLAB column-major FP32 layout:
 *   volume      img[ix + iy*nx + iz*nx*ny]                (reference src/cu_projection.cu:87)
 *   projection  proj[ia + ib*na]                          (src/cu_projection.cu:20)
 *   stack       proj[ia + ib*na + iv*na*nb]               (src/AUMISART.cu:203)
 * Geometry is in voxel units (SO, SD, da, db pre-divided by the voxel size) and ai/bi are the
 * reference's internal detector offsets, user offset - (n/2 - 0.5) (src/entry_AUMISART.cu:76-94).
 *
 * Operators (section 2) keep the argument order of the reference's host_* functions and take
 * DEVICE pointers; they are synchronous: they run on a library-owned stream of the current
 * device and wait for it (the reference ends every call with cudaDeviceSynchronize()).  Outputs of
 * the projector and backprojector are overwritten, not accumulated.  The drivers (section 3) take
 * HOST pointers like the reference's host_AUMISART.  The job API (section 4) keeps a scan's
 * state resident in device memory and runs asynchronously on a context's stream.
 */
#ifndef MCSART_HIP_H
#define MCSART_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------------------------
 * 1. Status, device memory and contexts (no counterpart in the reference, which reports no errors
 *    and uses the CUDA runtime directly).
 * ------------------------------------------------------------------------------------------- */
enum {
    MCSART_OK = 0,
    MCSART_EINVAL = -1,  /* invalid argument (bad size, nx != ny where required, null pointer) */
    MCSART_ENOMEM = -2,  /* device allocation failed or HBM budget exceeded */
    MCSART_EHIP = -3,    /* HIP runtime error */
    MCSART_ENOTSUP = -4, /* unsupported configuration */
    MCSART_EBUSY = -5,   /* resource in use */
    MCSART_ECANCELED = -6 /* the request was cancelled (mcsart_server_cancel) */
};

const char *mcsart_status_string(int status);
const char *mcsart_last_error(void);
int mcsart_version(void);

int mcsart_device_count(int *count);
int mcsart_set_device(int device);
int mcsart_malloc(void **dptr, size_t bytes);
int mcsart_free(void *dptr);
int mcsart_memcpy_htod(void *dst, const void *src, size_t bytes);
int mcsart_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int mcsart_memcpy_dtod(void *dst, const void *src, size_t bytes);
int mcsart_memset(void *dptr, int value, size_t bytes);
int mcsart_device_synchronize(void);
int mcsart_mem_info(size_t *free_bytes, size_t *total_bytes);

typedef struct mcsart_ctx mcsart_ctx;
int mcsart_ctx_create(mcsart_ctx **ctx, int device);
int mcsart_ctx_destroy(mcsart_ctx *ctx);
int mcsart_ctx_synchronize(mcsart_ctx *ctx);

/* ---------------------------------------------------------------------------------------------
 * 2. Operators: the reference's host_* functions (include/cu_*.h), same argument order.
 * ------------------------------------------------------------------------------------------- */

/* Replaces host_projection (include/cu_projection.h:7, src/cu_projection.cu:7-13).
 * Cone-beam footprint-weighted forward projection of one view; d_proj (na*nb) is overwritten. */
int mcsart_projection(float *d_proj, const float *d_img, float angle, float SO, float SD, float da,
                      int na, float ai, float db, int nb, float bi, int nx, int ny, int nz);

/* Replaces host_backprojection (include/cu_backprojection.h:7, src/cu_backprojection.cu:8-46).
 * Voxel-driven footprint backprojection of one view; d_img (nx*ny*nz) is overwritten. */
int mcsart_backprojection(float *d_img, const float *d_proj, float angle, float SO, float SD,
                          float da, int na, float ai, float db, int nb, float bi, int nx, int ny,
                          int nz);

/* Replaces host_deform (include/cu_deform.h:8, src/cu_deform.cu:3-47): d_img1 = trilinear
 * clamp-to-edge warp of d_img by the DVF m_c = alpha_c*volume + beta_c*flow.  nx == ny.
 * In-place (d_img1 == d_img) is allowed, as in the reference (src/AUMISART.cu:228). */
int mcsart_deform(float *d_img1, const float *d_img, int nx, int ny, int nz, float volume,
                  float flow, const float *alpha_x, const float *alpha_y, const float *alpha_z,
                  const float *beta_x, const float *beta_y, const float *beta_z);

/* Replaces host_deform2 (include/cu_deform.h:9, src/cu_deform.cu:49-64,92-113): first-order
 * Taylor warp with the reference's swapped boundary checks; reads past the end of the volume
 * contribute zero.  In-place allowed. */
int mcsart_deform2(float *d_img1, const float *d_img, int nx, int ny, int nz, float volume,
                   float flow, const float *alpha_x, const float *alpha_y, const float *alpha_z,
                   const float *beta_x, const float *beta_y, const float *beta_z);

/* Replaces host_update_udvf (include/cu_update_udvf.h:8, src/cu_update_udvf.cu:3-62). */
int mcsart_update_udvf(float *alpha_x, float *alpha_y, float *alpha_z, float *beta_x,
                       float *beta_y, float *beta_z, const float *img, const float *img0,
                       float volume_diff, float flow_diff, int nx, int ny, int nz, int iter);

/* Replaces host_add (include/cu_add.h:4, src/cu_add.cu:2-18): img1 += weight * img. */
int mcsart_add(float *img1, const float *img, int nx, int ny, int nz, float weight);

/* Replaces host_add2 (include/cu_add.h:8, src/cu_add.cu:20-58):
 * img1[ix,iy,iz] += img[iy,ix,iz] * weight * D_ind(img0)[ix,iy,iz]; nx == ny. */
int mcsart_add2(float *img1, const float *img, int nx, int ny, int nz, const float *img0,
                float weight, int ind);

/* Replaces host_initial (include/cu_initial.h:5, src/cu_initial.cu:2-17). */
int mcsart_initial(float *img, int nx, int ny, int nz, float value);

/* Replaces host_initial2 (include/cu_initial.h:8, src/cu_initial.cu:20-49):
 * img = (D1 + D2 + D3)(img0) * (volume + flow). */
int mcsart_initial2(float *img, int nx, int ny, int nz, const float *img0, float volume,
                    float flow);

/* Replaces host_division (include/cu_division.h:5, src/cu_division.cu:2-25): img1 /= img with
 * NaN and +-Inf mapped to 0.  In-place (img1 == img) allowed. */
int mcsart_division(float *img1, const float *img, int nx, int ny, int nz);

/* Replaces host_diff (src/mex_diff.cu:5,67-112): backward difference along ind (1, 2, 3), zero
 * at the first and last index. */
int mcsart_diff(float *img1, const float *img, int nx, int ny, int nz, int ind);

/* Volume form of the __device__ diff_x/diff_y/diff_z (include/cu_diff.h:4-6,
 * src/cu_diff.cu:2-33): forward difference along ind, zero at the last index. */
int mcsart_diff_forward(float *img1, const float *img, int nx, int ny, int nz, int ind);

/* Fused SART steps (device pointers, synchronous).  Each is one kernel that replaces a sequence
 * of reference operator calls; the drivers are built from them.
 *
 * mcsart_sart_residual: fuses src/AUMISART.cu:194-209 (image sweep) and :246-263 (motion sweep).
 *   d_rp (2*na*nb floats) receives interleaved pairs (r, n).
 *   motion_pairs == 0 (image sweep): d_vol is a volume x (N floats); r = meas - P(x), n = P(1).
 *   motion_pairs != 0 (motion sweep): d_vol holds N interleaved pairs (g, q) as written by
 *   mcsart_motion_prep; r = meas - P(g) = meas - P(x) + P(img0) and n = P(q).
 *   *d_norm (one device float) receives ||r|| / ||meas||, the reference's per-view residual
 *   ratio (:209, :263). */
int mcsart_sart_residual(float *d_rp, float *d_norm, const float *d_vol, int motion_pairs,
                         const float *d_meas, float angle, float SO, float SD, float da, int na,
                         float ai, float db, int nb, float bi, int nx, int ny, int nz);

/* mcsart_sart_backprojection: u = B(r) / B(n) from the pairs of mcsart_sart_residual, NaN/Inf
 * mapped to 0; fuses src/AUMISART.cu:212-218 and :265-270 (two backprojections, one division). */
int mcsart_sart_backprojection(float *d_u, const float *d_rp, float angle, float SO, float SD,
                               float da, int na, float ai, float db, int nb, float bi, int nx,
                               int ny, int nz);

/* mcsart_motion_prep: fuses src/AUMISART.cu:252-258 and :267.  With
 * img0 = sum over (alpha_y,D1),(alpha_x,D2),(alpha_z,D3) of add2(alpha_c, vd) plus the same with
 * beta and fd, d_pairs (2*nx*ny*nz floats) receives per voxel the interleaved pair
 * (g, q) = (img - img0, (D1+D2+D3)(img) * (-vd - fd)).  nx == ny. */
int mcsart_motion_prep(float *d_pairs, const float *d_img, const float *alpha_x,
                       const float *alpha_y, const float *alpha_z, const float *beta_x,
                       const float *beta_y, const float *beta_z, float vd, float fd, int nx,
                       int ny, int nz);

/* mcsart_motion_update: fuses src/AUMISART.cu:272-277: alpha_c += add2(u, wv, D_c img) and
 * beta_c += add2(u, wf, D_c img) with the pairs (y,1), (x,2), (z,3).  nx == ny. */
int mcsart_motion_update(float *alpha_x, float *alpha_y, float *alpha_z, float *beta_x,
                         float *beta_y, float *beta_z, const float *d_u, const float *d_img,
                         float wv, float wf, int nx, int ny, int nz);

/* Replaces host_multiple (src/cu_multiple.cu:2-9, kernel :11-43): img1 *= weight * D_ind(img0),
 * D_ind the forward difference along dimension ind (1, 2, 3), zero at the last index. */
int mcsart_multiple(float *img1, int nx, int ny, int nz, const float *img0, float weight, int ind);

/* The DVF inverse of waste/kernel_invertDVF.cu:1-18 (no host code for it ships with the
 * reference): per voxel, starting from d = 0, niter Gauss-Seidel sweeps
 * x = -mx(p + d), y = -my(p + d), z = -mz(p + d), each a clamp-to-edge trilinear sample in index
 * space (dim 0 fastest) at the current d.  Outputs must not alias inputs. */
int mcsart_invert_dvf(float *mx2, float *my2, float *mz2, const float *mx, const float *my,
                      const float *mz, int nx, int ny, int nz, int niter);

/* Forward simulator after include/generateProj.h:21-24 (its MEX source is not in the reference):
 * for each view i, the DVF alpha*volumes[i] + beta*flows[i] (kernel_forwardDVF), replaced by its
 * inverse (mcsart_invert_dvf, niter_inverse steps) when niter_inverse > 0, warps d_img
 * (kernel_deformation) and the warped image is projected at angles[i]; d_proj receives
 * n_views*na*nb floats.  volumes, flows, angles are host arrays; the rest device pointers. */
int mcsart_generate_proj(float *d_proj, const float *d_img, const float *alpha_x,
                         const float *alpha_y, const float *alpha_z, const float *beta_x,
                         const float *beta_y, const float *beta_z, const float *volumes,
                         const float *flows, const float *angles, int n_views, int niter_inverse,
                         float SO, float SD, float da, int na, float ai, float db, int nb,
                         float bi, int nx, int ny, int nz);

/* ---------------------------------------------------------------------------------------------
 * 2b. Asynchronous operators: the same operators enqueued on a context's stream (the context owns
 *     the stream, the scratch buffers and the kernel timers; its device becomes current).  They
 *     return once the work is enqueued; mcsart_ctx_synchronize() waits for it.  Scratch buffers are
 *     reused in stream order, so calls on one context never race one another; calls on different
 *     contexts share nothing.  No counterpart in the reference (its calls all end with
 *     cudaDeviceSynchronize()).
 * ------------------------------------------------------------------------------------------- */
int mcsart_projection_async(mcsart_ctx *ctx, float *d_proj, const float *d_img, float angle,
                            float SO, float SD, float da, int na, float ai, float db, int nb,
                            float bi, int nx, int ny, int nz);
int mcsart_backprojection_async(mcsart_ctx *ctx, float *d_img, const float *d_proj, float angle,
                                float SO, float SD, float da, int na, float ai, float db, int nb,
                                float bi, int nx, int ny, int nz);
int mcsart_deform_async(mcsart_ctx *ctx, float *d_img1, const float *d_img, int nx, int ny, int nz,
                        float volume, float flow, const float *alpha_x, const float *alpha_y,
                        const float *alpha_z, const float *beta_x, const float *beta_y,
                        const float *beta_z);
int mcsart_deform2_async(mcsart_ctx *ctx, float *d_img1, const float *d_img, int nx, int ny, int nz,
                         float volume, float flow, const float *alpha_x, const float *alpha_y,
                         const float *alpha_z, const float *beta_x, const float *beta_y,
                         const float *beta_z);
int mcsart_update_udvf_async(mcsart_ctx *ctx, float *alpha_x, float *alpha_y, float *alpha_z,
                             float *beta_x, float *beta_y, float *beta_z, const float *img,
                             const float *img0, float volume_diff, float flow_diff, int nx, int ny,
                             int nz, int iter);
int mcsart_add_async(mcsart_ctx *ctx, float *img1, const float *img, int nx, int ny, int nz,
                     float weight);
int mcsart_add2_async(mcsart_ctx *ctx, float *img1, const float *img, int nx, int ny, int nz,
                      const float *img0, float weight, int ind);
int mcsart_initial_async(mcsart_ctx *ctx, float *img, int nx, int ny, int nz, float value);
int mcsart_initial2_async(mcsart_ctx *ctx, float *img, int nx, int ny, int nz, const float *img0,
                          float volume, float flow);
int mcsart_division_async(mcsart_ctx *ctx, float *img1, const float *img, int nx, int ny, int nz);
int mcsart_diff_async(mcsart_ctx *ctx, float *img1, const float *img, int nx, int ny, int nz,
                      int ind);
int mcsart_diff_forward_async(mcsart_ctx *ctx, float *img1, const float *img, int nx, int ny,
                              int nz, int ind);
int mcsart_sart_residual_async(mcsart_ctx *ctx, float *d_rp, float *d_norm, const float *d_vol,
                               int motion_pairs, const float *d_meas, float angle, float SO,
                               float SD, float da, int na, float ai, float db, int nb, float bi,
                               int nx, int ny, int nz);
int mcsart_sart_backprojection_async(mcsart_ctx *ctx, float *d_u, const float *d_rp, float angle,
                                     float SO, float SD, float da, int na, float ai, float db,
                                     int nb, float bi, int nx, int ny, int nz);
int mcsart_motion_prep_async(mcsart_ctx *ctx, float *d_pairs, const float *d_img,
                             const float *alpha_x, const float *alpha_y, const float *alpha_z,
                             const float *beta_x, const float *beta_y, const float *beta_z,
                             float vd, float fd, int nx, int ny, int nz);
int mcsart_motion_update_async(mcsart_ctx *ctx, float *alpha_x, float *alpha_y, float *alpha_z,
                               float *beta_x, float *beta_y, float *beta_z, const float *d_u,
                               const float *d_img, float wv, float wf, int nx, int ny, int nz);

int mcsart_multiple_async(mcsart_ctx *ctx, float *img1, int nx, int ny, int nz,
                          const float *img0, float weight, int ind);
int mcsart_invert_dvf_async(mcsart_ctx *ctx, float *mx2, float *my2, float *mz2, const float *mx,
                            const float *my, const float *mz, int nx, int ny, int nz, int niter);
int mcsart_generate_proj_async(mcsart_ctx *ctx, float *d_proj, const float *d_img,
                               const float *alpha_x, const float *alpha_y, const float *alpha_z,
                               const float *beta_x, const float *beta_y, const float *beta_z,
                               const float *volumes, const float *flows, const float *angles,
                               int n_views, int niter_inverse, float SO, float SD, float da,
                               int na, float ai, float db, int nb, float bi, int nx, int ny,
                               int nz);

/* ---------------------------------------------------------------------------------------------
 * 2c. Allocation-free projection.  No counterpart in the reference: it is for callers that
 *     cannot allocate device memory (code under HIP graph capture, tight serving loops,
 *     harnesses that forbid hipMalloc).
 * ------------------------------------------------------------------------------------------- */
/* mcsart_projection (same arguments after `stream`, same result to FP32 rounding) without any
 * device allocation: the xy-transposed copy of the volume that x-branch rays read goes into a
 * fixed device array inside the library (mcsart_projection_noalloc_max_voxels() floats, 32 MiB,
 * held by every GPU the library runs on until the process ends) instead of an allocated scratch.
 * Asynchronous: enqueued on `stream` (a hipStream_t of the current device; NULL = the null
 * stream); the caller synchronises.  Volumes of more voxels than the array holds are rejected
 * with MCSART_EINVAL, whatever the angle.  Every call on one GPU shares that GPU's array, so calls
 * that overlap on one GPU (on different streams, or from different threads) race: serialise them,
 * e.g. on one stream. */
int mcsart_projection_noalloc(void *stream, float *d_proj, const float *d_img, float angle,
                              float SO, float SD, float da, int na, float ai, float db, int nb,
                              float bi, int nx, int ny, int nz);
size_t mcsart_projection_noalloc_max_voxels(void);

/* ---------------------------------------------------------------------------------------------
 * 3. Drivers: host_AUMISART (include/AUMISART.h:17,19), host pointers, same argument order.
 * ------------------------------------------------------------------------------------------- */

/* Overload 1, static SART (src/AUMISART.cu:3-141).  Outputs: h_outimg (nx*ny*nz),
 * h_outproj (na*nb*n_views, last image sweep's residuals), h_outnorm (n_iter*n_views residual
 * ratios), h_outalphax (nx*ny*nz).  outIter, dx and err_weights are accepted and unused, as in
 * the reference. */
int mcsart_aumisart(float *h_outimg, float *h_outproj, float *h_outnorm, float *h_outalphax,
                    const float *h_img, const float *h_proj, int nx, int ny, int nz, int na, int nb,
                    int outIter, int n_views, int n_iter, const int *op_iter, float da, float db,
                    float ai, float bi, float SO, float SD, float dx, float lambda,
                    const float *volumes, const float *flows, const float *err_weights,
                    const float *angles);

/* Overload 2, motion-compensated SART (src/AUMISART.cu:143-313), with the six 5DCT fields. */
int mcsart_aumisart_motion(float *h_outimg, float *h_outproj, float *h_outnorm,
                           float *h_outalphax, const float *h_img, const float *h_proj, int nx,
                           int ny, int nz, int na, int nb, int outIter, int n_views, int n_iter,
                           const int *op_iter, float da, float db, float ai, float bi, float SO,
                           float SD, float dx, float lambda, const float *volumes,
                           const float *flows, const float *err_weights, const float *angles,
                           const float *ax, const float *ay, const float *az, const float *bx,
                           const float *by, const float *bz);

/* ---------------------------------------------------------------------------------------------
 * 4. Device-resident jobs (no counterpart in the reference: its driver re-uploads each measured
 *    view and downloads each residual every view, src/AUMISART.cu:203,211).
 * ------------------------------------------------------------------------------------------- */
typedef struct {
    int nx, ny, nz; /* volume; nx == ny */
    int na, nb;     /* detector */
    float da, db;   /* detector pitch, voxel units */
    float ai, bi;   /* internal detector offsets */
    float SO, SD;   /* source-isocentre and source-detector distances, voxel units */
} mcsart_geom;

typedef struct {
    int motion;          /* 1: overload 2 (warped image sweep, view stride 1 in motion sweeps);
                            0: overload 1 (no warps, motion sweeps step views by 10) */
    int nonneg;          /* clamp the image at zero after every view update (not in reference) */
    int keep_residual;   /* store each image-sweep view's residual (h_outproj) */
    float lambda;        /* SART relaxation */
} mcsart_opts;

typedef struct mcsart_job mcsart_job;

/* Bytes of device memory a job of this shape holds (admission control). */
size_t mcsart_job_bytes(const mcsart_geom *g, int n_views, int n_iter, int keep_residual);

/* Allocate a job and upload its inputs.  fields may be NULL (all six start at zero). */
int mcsart_job_create(mcsart_ctx *ctx, const mcsart_geom *g, int n_views, const float *angles,
                      const float *volumes, const float *flows, const mcsart_opts *opts,
                      int n_iter, const int *op_iter, mcsart_job **job);
int mcsart_job_set_projections(mcsart_job *job, const float *h_proj);   /* host (V*D) */
int mcsart_job_set_projections_device(mcsart_job *job, const float *d_proj);
int mcsart_job_set_image(mcsart_job *job, const float *h_img);          /* host, NULL -> zero */
int mcsart_job_set_fields(mcsart_job *job, const float *const h_fields[6]); /* ax ay az bx by bz */
/* Restore the image and fields last given to set_image / set_fields (kept resident in device
 * memory) and clear the norms, so the same scan can be reconstructed again; enqueued on the
 * context stream. */
int mcsart_job_reset(mcsart_job *job);
/* Enqueue the whole schedule (n_iter sweeps) on the context stream; returns immediately. */
int mcsart_job_run(mcsart_job *job);
/* Enqueue a single sweep it (0 <= it < n_iter).  By default each sweep is captured into a HIP
 * graph on its first run and replayed afterwards; enable = 0 launches kernel by kernel instead
 * (environment: MCSART_GRAPHS=0 sets the default for new jobs). */
int mcsart_job_run_sweep(mcsart_job *job, int it);
int mcsart_job_set_graphs(mcsart_job *job, int enable);
/* Work the sweeps skip because no later kernel can see it (no counterpart in the reference, which
 * processes every voxel every view): a view's warp writes only the voxels its projector can read
 * (those whose box, grown by the projector's window, meets the cone from the source to the
 * detector's edges, and the volume's edge shell that clamped windows read); the motion step
 * prepares g only there and updates the accumulators only where u can be nonzero; the warp-add
 * leaves x alone where W(u) is 0 because u is 0 at every voxel its sample reads.  Results are the
 * same bit for bit.  mode 0: off; 1: on (the default); 2: on, with every skipped write made as a
 * NaN instead (a check that no kernel reads a skipped cell: any such read would show in the
 * results).  Drops the captured sweeps. */
int mcsart_job_set_culling(mcsart_job *job, int mode);
/* Sizes of the culled sets of a scan, per view v (counts[5*n_views]), for the algorithmic bytes
 * of the culled kernels: counts[5v] voxels in view v's read set (its warp's writes),
 * counts[5v+1] the read set with the volume's edge shell (where a motion step prepares g for
 * view v), counts[5v+2] voxels whose ratio backprojection of view v has cells and rows (u can be
 * nonzero: the motion step's accumulator writes), counts[5v+3] the union of that set and view
 * v+1's read set with its shell (the motion step's accumulator reads; for the last view,
 * counts[5v+2]), counts[5v+4] the volume layouts view v's rays read (as
 * mcsart_job_view_layouts).  Host arithmetic only (no device needed). */
int mcsart_view_voxels(const mcsart_geom *g, int n_views, const float *angles,
                       long long *counts);
int mcsart_job_get_image(mcsart_job *job, float *h_img);
int mcsart_job_get_norms(mcsart_job *job, float *h_norm);                /* n_iter*n_views */
int mcsart_job_get_residual(mcsart_job *job, float *h_proj);             /* n_views*na*nb */
int mcsart_job_get_fields(mcsart_job *job, float *const h_fields[6]);
/* Device pointers of the job's resident state (for zero-copy callers and tests). */
int mcsart_job_device_image(mcsart_job *job, float **d_img);
int mcsart_job_destroy(mcsart_job *job);

/* Launch shapes of the two footprint kernels (the reference has fixed 16x16 blocks,
 * src/cu_projection.cu:140, src/cu_backprojection.cu:186).  Zero fields mean automatic: shapes
 * chosen from the geometry alone, so a given scan always runs the same kernels.  The shapes
 * change floating-point rounding (plane-chunk sums, slab anchors) but not the mathematics. */
typedef struct {
    int proj_split;   /* projector workgroup slices along each ray: 0, 1, 2, 4, 8 or 16 */
    int back_zslab;   /* backprojector voxels per lane along z: 0 or 1..64 */
    int back_stage;   /* backprojector detector reads: 0 automatic (LDS-staged rectangle),
                         1 staged in LDS, 2 straight from global memory */
    int proj_rows;    /* detector rows (rays of one column) per projector lane: 0, 1, 2 or 4;
                         the rays share their z-row gathers where adjacent rows' footprints lie
                         less than a voxel apart (bit-identical either way) */
    int back_zslab_motion; /* voxels per lane along z of the motion sweep's backprojector (with
                              the fused field update): 0 (automatic) or 1..64 */
} mcsart_tuning;
int mcsart_job_set_tuning(mcsart_job *job, const mcsart_tuning *t);
int mcsart_job_get_tuning(mcsart_job *job, mcsart_tuning *t);
/* Time the candidate launch shapes on up to sixteen of the job's views spread over the scan
 * (projector in residual mode on the current image, backprojector in ratio mode into scratch;
 * the job's image, fields and norms are untouched), keep the fastest and report it in *chosen
 * (may be NULL).  Synchronous. */
int mcsart_job_autotune(mcsart_job *job, mcsart_tuning *chosen);
/* Average duration (microseconds) of one launch of a kernel of the sweeps, with the job's launch
 * shapes: 0 projector (residual mode), 1 backprojector (ratio mode), 2 warp of the image into the
 * layouts the view's rays read, 3 warp-add (x += lambda W(u)), 4 motion step (the motion
 * sweep's ratio backprojection with the update of view i, as accumulators, and the prep of view
 * i+1 fused), 5 the motion sweep's projector (reads the (g, Q) pairs).  Every view of the scan
 * (views 1..V-2 for kind 4) is launched back to back `reps` times between two HIP events on the
 * job's stream.  Outputs go to the job's scratch: kind 3 updates a copy of the image, kinds 4
 * and 5 the motion sweep's accumulators and pairs; the job's image, fields and norms are
 * untouched.  Synchronous. */
int mcsart_job_kernel_time(mcsart_job *job, int kind, int reps, double *avg_us);
/* Detector layouts each view's rays read: bit 0 the natural volume layout (y-branch rays), bit 1
 * the xy-transposed layout (x-branch rays); layouts[n_views]. */
int mcsart_job_view_layouts(mcsart_job *job, int *layouts);

/* ---------------------------------------------------------------------------------------------
 * 5. Reconstruction server: continuous batching of jobs on one device (no counterpart in the
 *    reference, which reconstructs one scan per MATLAB call, src/entry_AUMISART.cu:193-207).
 *    Requests queue in arrival order; one is admitted whenever a slot (a stream driven by its own
 *    host thread) is free and its device footprint (mcsart_job_bytes) fits in the HBM budget.
 *    It runs as a device-resident job in a block of the server's memory pool, its sweeps
 *    enqueued with at most two in flight; on completion the results are copied into the
 *    request's output buffers and the block returns to the pool for a later request of the same
 *    or a somewhat smaller footprint (idle blocks are freed only when the budget needs them).
 *    Jobs on different slots overlap on the device.  A request's results do not depend on what
 *    else runs (same kernels, shapes and order).
 * ------------------------------------------------------------------------------------------- */
typedef struct mcsart_server mcsart_server;

typedef struct {
    int slots;           /* jobs in flight at once, one stream each: 1..16 */
    size_t hbm_budget;   /* device bytes admitted jobs may hold together; 0: 90% of free memory */
    int autotune;        /* 1: autotune each job's launch shapes after admission */
    int proj_split, back_zslab, back_stage, proj_rows;   /* fixed shapes when autotune == 0 */
    int back_zslab_motion;                               /* (mcsart_tuning's fields) */
} mcsart_server_opts;

/* All host buffers must stay valid until the request has completed (mcsart_server_wait). */
typedef struct {
    mcsart_geom geom;
    int n_views;
    const float *angles, *volumes, *flows;   /* n_views each; copied at submit */
    mcsart_opts opts;
    int n_iter;
    const int *op_iter;                      /* copied at submit */
    const float *h_proj;                     /* n_views*na*nb measured projections */
    const float *h_img;                      /* initial image or NULL (zero) */
    const float *h_fields[6];                /* ax ay az bx by bz, each N or NULL (zero) */
    float *out_img;                          /* N, or NULL */
    float *out_norm;                         /* n_iter*n_views, or NULL */
    float *out_fields[6];                    /* each N or NULL */
} mcsart_request;

enum {
    MCSART_REQ_QUEUED = 0,
    MCSART_REQ_RUNNING = 1,
    MCSART_REQ_DONE = 2,
    MCSART_REQ_FAILED = 3,
    MCSART_REQ_CANCELLED = 4
};

typedef struct {
    int state;          /* MCSART_REQ_* */
    int status;         /* MCSART_OK or the error that failed the request */
    int slot;           /* slot it ran on, -1 while queued */
    size_t bytes;       /* device footprint */
    double queue_ms;    /* submit -> admission */
    double run_ms;      /* admission -> results in the output buffers */
    double total_ms;    /* submit -> results in the output buffers */
    mcsart_tuning tuning;   /* launch shapes it ran with */
    double admit_at_ms; /* admission and completion on one clock for all requests: milliseconds */
    double done_at_ms;  /* since the server was created (0 until the event happened) */
} mcsart_request_info;

int mcsart_server_create(int device, const mcsart_server_opts *opts, mcsart_server **server);
/* MCSART_ENOMEM if the request could never fit the budget. */
int mcsart_server_submit(mcsart_server *server, const mcsart_request *req, long long *id);
/* Block until request id has ended: done, failed or cancelled (timeout_s < 0: no limit;
 * MCSART_EBUSY on timeout).  Returns the request's own status, with its message in
 * mcsart_last_error() on the calling thread. */
int mcsart_server_wait(mcsart_server *server, long long id, double timeout_s);
/* Cancel request id.  Queued: it leaves the queue at once.  Running: no further sweep is
 * enqueued; its slot (and its memory) is free once the sweeps already on the device complete (at
 * most two), without copying results.  Either way it ends as MCSART_REQ_CANCELLED and waiting
 * on it returns MCSART_ECANCELED.  MCSART_EINVAL for unknown or already ended requests. */
int mcsart_server_cancel(mcsart_server *server, long long id);
int mcsart_server_drain(mcsart_server *server);   /* wait for every submitted request */
int mcsart_server_info(mcsart_server *server, long long id, mcsart_request_info *info);
int mcsart_server_release(mcsart_server *server, long long id);   /* forget a completed request */
/* used: device bytes held by admitted jobs; with the pool's idle blocks it stays within budget. */
int mcsart_server_memory(mcsart_server *server, size_t *used, size_t *budget);
/* Runs every queued request to completion, then stops the scheduler and frees the slots. */
int mcsart_server_destroy(mcsart_server *server);

/* ---------------------------------------------------------------------------------------------
 * 6. Gateway parameter parsing: the MATLAB gateways' struct fields, aliases, defaults and
 *    "MATLAB:badInput" errors (src/entry_AUMISART.cu:18-160, src/mex_backprojection.cu:13-59,
 *    src/mex_diff.cu:17-37, waste/8888mex_projection.cu:13-126), over a caller-supplied lookup so
 *    any host (MEX wrapper, file driver, Python) shares one implementation.  A missing required
 *    field returns MCSART_EINVAL with the reference's message in mcsart_last_error().
 * ------------------------------------------------------------------------------------------- */
/* Return 1 and set *value (may be NULL: presence check) if the struct has field `name` (its
 * first element for arrays, as mxGetScalar), 0 otherwise. */
typedef int (*mcsart_scalar_fn)(void *user, const char *name, double *value);

enum {
    MCSART_GW_AUMISART = 0,        /* entry_AUMISART: volume, detector, dx, da, db, ai, bi, SO, SD */
    MCSART_GW_BACKPROJECTION = 1,  /* mex_backprojection: no dx, plus angle */
    MCSART_GW_PROJECTION = 2,      /* standalone projection: plus angle; bi shifted only if > -1 */
    MCSART_GW_DIFF = 3             /* mex_diff: nx, ny, nz, ind */
};
enum {   /* bits of `defaulted`: fields absent from the struct that took their default */
    MCSART_DEF_DX = 1, MCSART_DEF_DA = 2, MCSART_DEF_DB = 4, MCSART_DEF_AI = 8, MCSART_DEF_BI = 16,
    MCSART_DEF_NITER = 32
};

typedef struct {
    int nx, ny, nz, na, nb;
    float dx, da, db;
    float ai, bi;      /* internal offsets: user offset - (n/2 - 0.5), in float */
    float SO, SD;
    float angle;       /* projection / backprojection gateways */
    int ind;           /* diff gateway */
    unsigned defaulted;
} mcsart_geo_params;

typedef struct {
    int n_iter, n_views, outIter;
    float lambda;
    int motion;        /* 1 if alpha_x is present: the motion overload of host_AUMISART */
    unsigned defaulted;
} mcsart_iter_params;

int mcsart_parse_geo(int gateway, mcsart_scalar_fn get, void *user, mcsart_geo_params *out);
/* Scalars of ITER_PARA; presence of op_iter, volumes, flows, err_weights and angles. */
int mcsart_parse_iter(mcsart_scalar_fn get, void *user, mcsart_iter_params *out);

/* Per-kernel timing of the last run(s) on this context, from HIP events recorded on the stream
 * the kernels run on.  kind: 0 projector, 1 backprojector, 2 warp, 3 motion elementwise. */
int mcsart_ctx_enable_timing(mcsart_ctx *ctx, int enable);
int mcsart_ctx_kernel_time(mcsart_ctx *ctx, int kind, double *total_ms, long long *launches);
int mcsart_ctx_reset_timing(mcsart_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* MCSART_HIP_H */

// C++ compatibility layer: the reference's exact host_* names and overloads (include/cu_*.h,
// include/AUMISART.h), implemented inline on top of the C ABI in mcsart_hip.h.  Code written
// against the reference headers (for instance its MEX gateways) compiles unchanged against this
// header and links libmcsart_hip instead of the CUDA sources.
//
// The reference functions return void and report nothing; these shims keep the void signatures
// and route a failed status to mcsart_compat_on_error(), which by default prints the library's
// message to stderr.  Define MCSART_COMPAT_ON_ERROR before including this header to replace it
// (a MEX gateway would call mexErrMsgIdAndTxt there).
#ifndef MCSART_COMPAT_H
#define MCSART_COMPAT_H

#include <cstdio>

#include "mcsart_hip.h"

#ifndef MCSART_COMPAT_ON_ERROR
inline void mcsart_compat_on_error(int status, const char *where) {
    std::fprintf(stderr, "%s: %s: %s\n", where, mcsart_status_string(status), mcsart_last_error());
}
#define MCSART_COMPAT_ON_ERROR(status, where) mcsart_compat_on_error(status, where)
#endif

#define MCSART_COMPAT_CALL(expr, where)                      \
    do {                                                     \
        int s_ = (expr);                                     \
        if (s_ != MCSART_OK) MCSART_COMPAT_ON_ERROR(s_, where); \
    } while (0)

// include/cu_projection.h:7
inline void host_projection(float *d_proj, float *d_img, float angle, float SO, float SD, float da,
                            int na, float ai, float db, int nb, float bi, int nx, int ny, int nz) {
    MCSART_COMPAT_CALL(mcsart_projection(d_proj, d_img, angle, SO, SD, da, na, ai, db, nb, bi, nx,
                                         ny, nz), "host_projection");
}

// include/cu_backprojection.h:7
inline void host_backprojection(float *d_img, float *d_proj, float angle, float SO, float SD,
                                float da, int na, float ai, float db, int nb, float bi, int nx,
                                int ny, int nz) {
    MCSART_COMPAT_CALL(mcsart_backprojection(d_img, d_proj, angle, SO, SD, da, na, ai, db, nb, bi,
                                             nx, ny, nz), "host_backprojection");
}

// include/cu_deform.h:8-9
inline void host_deform(float *d_img1, float *d_img, int nx, int ny, int nz, float volume_diff,
                        float flow_diff, float *alpha_x, float *alpha_y, float *alpha_z,
                        float *beta_x, float *beta_y, float *beta_z) {
    MCSART_COMPAT_CALL(mcsart_deform(d_img1, d_img, nx, ny, nz, volume_diff, flow_diff, alpha_x,
                                     alpha_y, alpha_z, beta_x, beta_y, beta_z), "host_deform");
}
inline void host_deform2(float *d_img1, float *d_img, int nx, int ny, int nz, float volume_diff,
                         float flow_diff, float *alpha_x, float *alpha_y, float *alpha_z,
                         float *beta_x, float *beta_y, float *beta_z) {
    MCSART_COMPAT_CALL(mcsart_deform2(d_img1, d_img, nx, ny, nz, volume_diff, flow_diff, alpha_x,
                                      alpha_y, alpha_z, beta_x, beta_y, beta_z), "host_deform2");
}

// include/cu_update_udvf.h:8
inline void host_update_udvf(float *alpha_x, float *alpha_y, float *alpha_z, float *beta_x,
                             float *beta_y, float *beta_z, float *img, float *img0,
                             float volume_diff, float flow_diff, int nx, int ny, int nz,
                             int iter) {
    MCSART_COMPAT_CALL(mcsart_update_udvf(alpha_x, alpha_y, alpha_z, beta_x, beta_y, beta_z, img,
                                          img0, volume_diff, flow_diff, nx, ny, nz, iter),
                       "host_update_udvf");
}

// include/cu_add.h:4,8
inline void host_add(float *img1, float *img, int nx, int ny, int nz, float weight) {
    MCSART_COMPAT_CALL(mcsart_add(img1, img, nx, ny, nz, weight), "host_add");
}
inline void host_add2(float *img1, float *img, int nx, int ny, int nz, float *img0, float weight,
                      int ind) {
    MCSART_COMPAT_CALL(mcsart_add2(img1, img, nx, ny, nz, img0, weight, ind), "host_add2");
}

// include/cu_initial.h:5,8
inline void host_initial(float *img, int nx, int ny, int nz, float value) {
    MCSART_COMPAT_CALL(mcsart_initial(img, nx, ny, nz, value), "host_initial");
}
inline void host_initial2(float *img, int nx, int ny, int nz, float *img0, float volume,
                          float flow) {
    MCSART_COMPAT_CALL(mcsart_initial2(img, nx, ny, nz, img0, volume, flow), "host_initial2");
}

// include/cu_division.h:5
inline void host_division(float *img1, float *img, int nx, int ny, int nz) {
    MCSART_COMPAT_CALL(mcsart_division(img1, img, nx, ny, nz), "host_division");
}

// src/mex_diff.cu:5
inline void host_diff(float *img1, float *img, int nx, int ny, int nz, int ind) {
    MCSART_COMPAT_CALL(mcsart_diff(img1, img, nx, ny, nz, ind), "host_diff");
}

// include/cu_multiple.h (src/cu_multiple.cu:2)
inline void host_multiple(float *img1, int nx, int ny, int nz, float *img0, float weight,
                          int ind) {
    MCSART_COMPAT_CALL(mcsart_multiple(img1, nx, ny, nz, img0, weight, ind), "host_multiple");
}

// include/AUMISART.h:17 (overload 1, static) and :19 (overload 2, motion-compensated)
inline void host_AUMISART(float *h_outimg, float *h_outproj, float *h_outnorm, float *h_outalphax,
                          float *h_img, float *h_proj, int nx, int ny, int nz, int na, int nb,
                          int outIter, int n_views, int n_iter, int *op_iter, float da, float db,
                          float ai, float bi, float SO, float SD, float dx, float lambda,
                          float *volumes, float *flows, float *err_weights, float *angles) {
    MCSART_COMPAT_CALL(mcsart_aumisart(h_outimg, h_outproj, h_outnorm, h_outalphax, h_img, h_proj,
                                       nx, ny, nz, na, nb, outIter, n_views, n_iter, op_iter, da,
                                       db, ai, bi, SO, SD, dx, lambda, volumes, flows, err_weights,
                                       angles), "host_AUMISART");
}
inline void host_AUMISART(float *h_outimg, float *h_outproj, float *h_outnorm, float *h_outalphax,
                          float *h_img, float *h_proj, int nx, int ny, int nz, int na, int nb,
                          int outIter, int n_views, int n_iter, int *op_iter, float da, float db,
                          float ai, float bi, float SO, float SD, float dx, float lambda,
                          float *volumes, float *flows, float *err_weights, float *angles,
                          float *ax, float *ay, float *az, float *bx, float *by, float *bz) {
    MCSART_COMPAT_CALL(mcsart_aumisart_motion(h_outimg, h_outproj, h_outnorm, h_outalphax, h_img,
                                              h_proj, nx, ny, nz, na, nb, outIter, n_views, n_iter,
                                              op_iter, da, db, ai, bi, SO, SD, dx, lambda, volumes,
                                              flows, err_weights, angles, ax, ay, az, bx, by, bz),
                       "host_AUMISART");
}

#endif  // MCSART_COMPAT_H

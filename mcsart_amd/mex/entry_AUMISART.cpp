// [img, nrm, alpha_x, residual] = entry_AUMISART(img0, proj, geo_para, iter_para)
//
// Same inputs, outputs and overload choice as src/entry_AUMISART.cu:3-210 (parameters parsed by
// mcsart_parse_geo / mcsart_parse_iter with the reference's aliases, defaults and messages); the
// reconstruction runs in libmcsart_hip (mcsart_aumisart / mcsart_aumisart_motion, host pointers
// in, as host_AUMISART, src/AUMISART.cu:3,143).
#include "mex_common.hpp"

using namespace mcsart_mex;

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    if (nrhs < 4) mexErrMsgIdAndTxt("MATLAB:badInput", "entry_AUMISART(img, proj, geo, iter)\n");
    const mxArray *IN_IMG = prhs[0], *PROJ = prhs[1], *GEO = prhs[2], *ITER = prhs[3];
    mcsart_geo_params g;
    raise(mcsart_parse_geo(MCSART_GW_AUMISART, struct_scalar, (void *)GEO, &g));
    mcsart_iter_params it;
    raise(mcsart_parse_iter(struct_scalar, (void *)ITER, &it));
    const size_t N = (size_t)g.nx * g.ny * g.nz, D = (size_t)g.na * g.nb;
    const size_t V = (size_t)it.n_views;
    const float *h_img = single_data(IN_IMG, N, "img");
    const float *h_proj = single_data(PROJ, D * V, "proj");
    const mxArray *op = mxGetField(ITER, 0, "op_iter");
    if (!mxIsInt32(op) || mxGetNumberOfElements(op) < (size_t)it.n_iter)
        mexErrMsgIdAndTxt("MCSART:badInput", "iter_para.op_iter must be int32 with n_iter entries\n");
    const int *op_iter = static_cast<const int *>(mxGetData(op));
    const float *vol = single_data(mxGetField(ITER, 0, "volumes"), V, "iter_para.volumes");
    const float *flw = single_data(mxGetField(ITER, 0, "flows"), V, "iter_para.flows");
    const float *ew = single_data(mxGetField(ITER, 0, "err_weights"), V, "iter_para.err_weights");
    const float *ang = single_data(mxGetField(ITER, 0, "angles"), V, "iter_para.angles");

    float *out_img = new_single(&plhs[0], g.nx, g.ny, g.nz);
    float *out_nrm = new_single(&plhs[1], (mwSize)it.n_iter * V, 1, 0);
    mxArray *ax_arr = nullptr, *res_arr = nullptr;
    float *out_ax = new_single(&ax_arr, g.nx, g.ny, g.nz);
    float *out_res = new_single(&res_arr, g.na, g.nb, V);
    if (!it.motion) {
        raise(mcsart_aumisart(out_img, out_res, out_nrm, out_ax, h_img, h_proj, g.nx, g.ny, g.nz,
                              g.na, g.nb, it.outIter, it.n_views, it.n_iter, op_iter, g.da, g.db,
                              g.ai, g.bi, g.SO, g.SD, g.dx, it.lambda, vol, flw, ew, ang));
    } else {
        const char *names[6] = {"alpha_x", "alpha_y", "alpha_z", "beta_x", "beta_y", "beta_z"};
        const float *f[6];
        for (int c = 0; c < 6; ++c) f[c] = single_data(mxGetField(ITER, 0, names[c]), N, names[c]);
        raise(mcsart_aumisart_motion(out_img, out_res, out_nrm, out_ax, h_img, h_proj, g.nx,
                                     g.ny, g.nz, g.na, g.nb, it.outIter, it.n_views, it.n_iter,
                                     op_iter, g.da, g.db, g.ai, g.bi, g.SO, g.SD, g.dx,
                                     it.lambda, vol, flw, ew, ang, f[0], f[1], f[2], f[3], f[4],
                                     f[5]));
    }
    if (nlhs > 2) plhs[2] = ax_arr; else mxDestroyArray(ax_arr);
    if (nlhs > 3) plhs[3] = res_arr; else mxDestroyArray(res_arr);
}

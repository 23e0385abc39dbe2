// img = mex_backprojection(proj, geo_para): one view (src/mex_backprojection.cu:3-75).
#include "mex_common.hpp"

using namespace mcsart_mex;

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    if (nrhs < 2) mexErrMsgIdAndTxt("MATLAB:badInput", "mex_backprojection(proj, geo)\n");
    mcsart_geo_params g;
    raise(mcsart_parse_geo(MCSART_GW_BACKPROJECTION, struct_scalar, (void *)prhs[1], &g));
    const size_t N = (size_t)g.nx * g.ny * g.nz, D = (size_t)g.na * g.nb;
    const float *h_proj = single_data(prhs[0], D, "proj");
    DeviceBuffer d_proj(D * sizeof(float)), d_img(N * sizeof(float));
    raise(mcsart_memcpy_htod(d_proj.p, h_proj, D * sizeof(float)));
    raise(mcsart_backprojection(d_img.f(), d_proj.f(), g.angle, g.SO, g.SD, g.da, g.na, g.ai,
                                g.db, g.nb, g.bi, g.nx, g.ny, g.nz));
    float *out = new_single(&plhs[0], g.nx, g.ny, g.nz);
    raise(mcsart_memcpy_dtoh(out, d_img.p, N * sizeof(float)));
}

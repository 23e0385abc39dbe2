// proj = mex_projection(img, geo_para): one view, the authors' standalone projection gateway
// (waste/8888mex_projection.cu:3-139, including its "bi > -1" offset rule).
#include "mex_common.hpp"

using namespace mcsart_mex;

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    if (nrhs < 2) mexErrMsgIdAndTxt("MATLAB:badInput", "mex_projection(img, geo)\n");
    mcsart_geo_params g;
    raise(mcsart_parse_geo(MCSART_GW_PROJECTION, struct_scalar, (void *)prhs[1], &g));
    const size_t N = (size_t)g.nx * g.ny * g.nz, D = (size_t)g.na * g.nb;
    const float *h_img = single_data(prhs[0], N, "img");
    DeviceBuffer d_img(N * sizeof(float)), d_proj(D * sizeof(float));
    raise(mcsart_memcpy_htod(d_img.p, h_img, N * sizeof(float)));
    raise(mcsart_projection(d_proj.f(), d_img.f(), g.angle, g.SO, g.SD, g.da, g.na, g.ai, g.db,
                            g.nb, g.bi, g.nx, g.ny, g.nz));
    float *out = new_single(&plhs[0], g.na, g.nb, 0);
    raise(mcsart_memcpy_dtoh(out, d_proj.p, D * sizeof(float)));
}

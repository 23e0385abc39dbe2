// out = mex_diff(img, geo_para): backward difference along dimension geo_para.ind, zero at the
// first and last index (src/mex_diff.cu:5-112).
#include "mex_common.hpp"

using namespace mcsart_mex;

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    if (nrhs < 2) mexErrMsgIdAndTxt("MATLAB:badInput", "mex_diff(img, geo)\n");
    mcsart_geo_params g;
    raise(mcsart_parse_geo(MCSART_GW_DIFF, struct_scalar, (void *)prhs[1], &g));
    const size_t N = (size_t)g.nx * g.ny * g.nz;
    const float *h_img = single_data(prhs[0], N, "img");
    DeviceBuffer d_img(N * sizeof(float)), d_out(N * sizeof(float));
    raise(mcsart_memcpy_htod(d_img.p, h_img, N * sizeof(float)));
    raise(mcsart_diff(d_out.f(), d_img.f(), g.nx, g.ny, g.nz, g.ind));
    float *out = new_single(&plhs[0], g.nx, g.ny, g.nz);
    raise(mcsart_memcpy_dtoh(out, d_out.p, N * sizeof(float)));
}

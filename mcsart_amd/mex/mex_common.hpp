// Shared helpers of the MATLAB gateways over libmcsart_hip (built only where MATLAB exists).
#pragma once

#include <cstring>
#include <string>

#include "mex.h"
#include "mcsart_hip.h"

namespace mcsart_mex {

// mcsart_scalar_fn over a MATLAB struct: field present -> mxGetScalar of it.
inline int struct_scalar(void *user, const char *name, double *value) {
    const mxArray *f = mxGetField(static_cast<const mxArray *>(user), 0, name);
    if (!f) return 0;
    if (value) *value = mxGetScalar(f);
    return 1;
}

// Raise the library's last error: parse errors keep the reference's "MATLAB:badInput" id.
inline void raise(int status) {
    if (status == MCSART_OK) return;
    std::string msg = mcsart_last_error();
    const std::string tag = "MATLAB:badInput: ";
    if (msg.compare(0, tag.size(), tag) == 0)
        mexErrMsgIdAndTxt("MATLAB:badInput", "%s\n", msg.substr(tag.size()).c_str());
    mexErrMsgIdAndTxt("MCSART:error", "%s (%s)\n", msg.c_str(), mcsart_status_string(status));
}

inline const float *single_data(const mxArray *a, size_t n, const char *what) {
    if (!a || !mxIsSingle(a) || mxIsComplex(a) || mxGetNumberOfElements(a) < n)
        mexErrMsgIdAndTxt("MCSART:badInput", "%s must be a real single array of %zu elements\n",
                          what, n);
    return static_cast<const float *>(mxGetData(a));
}

inline float *new_single(mxArray **out, mwSize d0, mwSize d1, mwSize d2) {
    const mwSize dims[3] = {d0, d1, d2};
    *out = mxCreateNumericArray(d2 ? 3 : 2, dims, mxSINGLE_CLASS, mxREAL);
    return static_cast<float *>(mxGetData(*out));
}

// Device buffer released on scope exit (also when a MATLAB error unwinds the gateway).
struct DeviceBuffer {
    void *p = nullptr;
    explicit DeviceBuffer(size_t bytes) { raise(mcsart_malloc(&p, bytes)); }
    ~DeviceBuffer() { mcsart_free(p); }
    float *f() const { return static_cast<float *>(p); }
};

}  // namespace mcsart_mex

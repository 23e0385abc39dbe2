// Allocation-free forward projection for callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// mcsart_projection's x-branch rays read an xy-transposed copy of the volume, which it writes into
// a context scratch buffer allocated on demand.  This variant writes that copy into `scratch`, a
// file-scope device array of a fixed size that every GPU the library runs on holds until the
// process ends; it calls no hipMalloc, hipFree or allocating library routine.
//
// In host code `scratch` names a host-side stand-in, not device memory: its device address is
// looked up once per device (hipGetSymbolAddress, with that device current) and only that address
// is passed to kernels.
#include <mutex>

#include "internal.hpp"

namespace mcsart {
namespace {

// One transposed volume of up to 2^23 voxels (32 MiB): 128^3, 192x192x128 and 256x256x128 fit.
constexpr size_t kScratchFloats = size_t(1) << 23;
__device__ __align__(256) float scratch[kScratchFloats];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float *g_scratch_addr[kMaxDevices] = {nullptr};

// Device address of `scratch` on the current device, looked up on the first call per device.
int scratch_address(float **out) {
    int dev = 0;
    MCSART_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(MCSART_ENOTSUP, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (!g_scratch_addr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(MCSART_EHIP, std::string("hipGetSymbolAddress(scratch): ") +
                                         hipGetErrorString(e));
        g_scratch_addr[dev] = (float *)p;
    }
    *out = g_scratch_addr[dev];
    return MCSART_OK;
}

}  // namespace
}  // namespace mcsart

using namespace mcsart;

extern "C" size_t mcsart_projection_noalloc_max_voxels(void) { return kScratchFloats; }

extern "C" int mcsart_projection_noalloc(void *stream, float *d_proj, const float *d_img,
                                         float angle, float SO, float SD, float da, int na,
                                         float ai, float db, int nb, float bi, int nx, int ny,
                                         int nz) {
    const Geom g{nx, ny, nz, na, nb, da, db, ai, bi, SO, SD};
    int rc = validate_geom(g);
    if (rc) return rc;
    if (!d_proj || !d_img) return fail(MCSART_EINVAL, "null device pointer");
    const ViewTrig t = view_trig(angle);
    const ViewInfo vi = view_info_of(g, t);
    hipStream_t st = (hipStream_t)stream;
    ProjArgs a{};
    a.vol0 = d_img;
    a.out = d_proj;
    // The limit holds for every view, also those whose rays all take the y-branch (they read no
    // transposed copy), so whether a call is accepted does not depend on the angle.
    if ((size_t)g.nvox() > kScratchFloats)
        return fail(MCSART_EINVAL, "mcsart_projection_noalloc: volume of " +
                                       std::to_string(g.nvox()) + " voxels exceeds the scratch of " +
                                       std::to_string(kScratchFloats) +
                                       " voxels (use mcsart_projection)");
    if (vi.has_x) {   // x-branch rays read the transposed copy, here in the static scratch
        float *T = nullptr;
        rc = scratch_address(&T);
        if (rc) return rc;
        rc = launch_transpose(T, d_img, nx, ny, nz, st);
        if (rc) return rc;
        a.vol0T = T;
    }
    return launch_projector(g, t, vi, PROJ_PLAIN, a, st);
}

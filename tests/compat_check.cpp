// The reference-side binding, compiled: code written against the reference's host_* functions
// (include/cu_*.h, include/AUMISART.h:17,19) built against include/mcsart_compat.h and linked with
// libmcsart_hip.  Every shim is called with the reference's argument types (float* for inputs
// the reference does not mark const, int* op_iter), in the reference's argument order.
//
//     compat_check <dir>
//
// reads <dir>/params.txt ("nx ny nz na nb V n_iter da db ai bi SO SD lambda" then n_iter op_iter
// entries) and raw float32 files img.bin (N), proj.bin (V*D), angles.bin, volumes.bin,
// flows.bin (V each), fields.bin (6N: ax ay az bx by bz); writes out_*.bin for the test to check
// against the FP64 oracle.  Exit status 0 when every call succeeded.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

static int g_failures = 0;
#define MCSART_COMPAT_ON_ERROR(status, where)                                                  \
    do {                                                                                       \
        std::fprintf(stderr, "%s failed: %s (%s)\n", where, mcsart_status_string(status),      \
                     mcsart_last_error());                                                     \
        ++g_failures;                                                                          \
    } while (0)
#include "mcsart_compat.h"

namespace {

std::vector<float> read_floats(const std::string &path, size_t n) {
    std::vector<float> v(n);
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f || std::fread(v.data(), sizeof(float), n, f) != n) {
        std::fprintf(stderr, "cannot read %zu floats from %s\n", n, path.c_str());
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

void write_floats(const std::string &path, const float *p, size_t n) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(p, sizeof(float), n, f) != n) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(2);
    }
    std::fclose(f);
}

float *device_copy(const std::vector<float> &h) {
    void *d = nullptr;
    if (mcsart_malloc(&d, h.size() * sizeof(float)) != MCSART_OK ||
        mcsart_memcpy_htod(d, h.data(), h.size() * sizeof(float)) != MCSART_OK) {
        std::fprintf(stderr, "device allocation failed: %s\n", mcsart_last_error());
        std::exit(2);
    }
    return (float *)d;
}

std::vector<float> host_copy(const float *d, size_t n) {
    std::vector<float> h(n);
    if (mcsart_memcpy_dtoh(h.data(), d, n * sizeof(float)) != MCSART_OK) {
        std::fprintf(stderr, "download failed: %s\n", mcsart_last_error());
        std::exit(2);
    }
    return h;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: compat_check <dir>\n");
        return 2;
    }
    const std::string dir = argv[1];
    int nx, ny, nz, na, nb, V, n_iter;
    float da, db, ai, bi, SO, SD, lambda;
    FILE *pf = std::fopen((dir + "/params.txt").c_str(), "r");
    if (!pf || std::fscanf(pf, "%d %d %d %d %d %d %d %f %f %f %f %f %f %f", &nx, &ny, &nz, &na,
                           &nb, &V, &n_iter, &da, &db, &ai, &bi, &SO, &SD, &lambda) != 14) {
        std::fprintf(stderr, "bad params.txt\n");
        return 2;
    }
    std::vector<int> op_iter(n_iter);
    for (int i = 0; i < n_iter; ++i)
        if (std::fscanf(pf, "%d", &op_iter[i]) != 1) return 2;
    std::fclose(pf);
    const size_t N = (size_t)nx * ny * nz, D = (size_t)na * nb;
    std::vector<float> img = read_floats(dir + "/img.bin", N);
    std::vector<float> proj = read_floats(dir + "/proj.bin", D * V);
    std::vector<float> angles = read_floats(dir + "/angles.bin", V);
    std::vector<float> volumes = read_floats(dir + "/volumes.bin", V);
    std::vector<float> flows = read_floats(dir + "/flows.bin", V);
    std::vector<float> fields = read_floats(dir + "/fields.bin", 6 * N);

    // ---- operators on device pointers (include/cu_*.h) ----------------------------------------
    float *d_img = device_copy(img);
    float *d_proj = device_copy(std::vector<float>(proj.begin(), proj.begin() + D));
    float *d_f[6];
    for (int c = 0; c < 6; ++c)
        d_f[c] = device_copy(std::vector<float>(fields.begin() + c * N, fields.begin() + (c + 1) * N));
    float *d_out = device_copy(std::vector<float>(D, 0.f));
    float *d_vol = device_copy(std::vector<float>(N, 0.f));
    float *d_tmp = device_copy(std::vector<float>(N, 0.f));

    host_projection(d_out, d_img, angles[0], SO, SD, da, na, ai, db, nb, bi, nx, ny, nz);
    write_floats(dir + "/out_projection.bin", host_copy(d_out, D).data(), D);
    host_backprojection(d_vol, d_proj, angles[0], SO, SD, da, na, ai, db, nb, bi, nx, ny, nz);
    write_floats(dir + "/out_backprojection.bin", host_copy(d_vol, N).data(), N);
    // in place, as src/AUMISART.cu:228 calls it
    host_initial(d_tmp, nx, ny, nz, 0.f);
    host_add(d_tmp, d_img, nx, ny, nz, 1.f);
    host_deform(d_tmp, d_tmp, nx, ny, nz, volumes[1] - volumes[0], flows[1] - flows[0], d_f[0],
                d_f[1], d_f[2], d_f[3], d_f[4], d_f[5]);
    write_floats(dir + "/out_deform.bin", host_copy(d_tmp, N).data(), N);
    host_deform2(d_vol, d_img, nx, ny, nz, 0.5f, 0.25f, d_f[0], d_f[1], d_f[2], d_f[3], d_f[4],
                 d_f[5]);
    write_floats(dir + "/out_deform2.bin", host_copy(d_vol, N).data(), N);
    host_initial2(d_vol, nx, ny, nz, d_img, -0.3f, 0.1f);
    write_floats(dir + "/out_initial2.bin", host_copy(d_vol, N).data(), N);
    host_add2(d_tmp, d_vol, nx, ny, nz, d_img, 0.7f, 2);
    write_floats(dir + "/out_add2.bin", host_copy(d_tmp, N).data(), N);
    host_division(d_tmp, d_vol, nx, ny, nz);
    write_floats(dir + "/out_division.bin", host_copy(d_tmp, N).data(), N);
    host_diff(d_vol, d_img, nx, ny, nz, 3);
    write_floats(dir + "/out_diff.bin", host_copy(d_vol, N).data(), N);
    host_multiple(d_vol, nx, ny, nz, d_img, 0.5f, 1);
    write_floats(dir + "/out_multiple.bin", host_copy(d_vol, N).data(), N);
    host_update_udvf(d_f[0], d_f[1], d_f[2], d_f[3], d_f[4], d_f[5], d_img, d_tmp, 0.4f, -0.2f,
                     nx, ny, nz, 0);
    std::vector<float> upd(6 * N);
    for (int c = 0; c < 6; ++c) {
        std::vector<float> h = host_copy(d_f[c], N);
        std::copy(h.begin(), h.end(), upd.begin() + c * N);
    }
    write_floats(dir + "/out_update_udvf.bin", upd.data(), 6 * N);
    for (float *p : {d_img, d_proj, d_out, d_vol, d_tmp, d_f[0], d_f[1], d_f[2], d_f[3], d_f[4],
                     d_f[5]})
        mcsart_free(p);

    // ---- both host_AUMISART overloads on host pointers (include/AUMISART.h:17,19) -------------
    std::vector<float> out_img(N), out_proj(D * V), out_norm((size_t)n_iter * V), out_ax(N);
    std::vector<float> err_weights(V, 1.f);
    std::vector<float> zero(N, 0.f);
    host_AUMISART(out_img.data(), out_proj.data(), out_norm.data(), out_ax.data(), zero.data(),
                  proj.data(), nx, ny, nz, na, nb, 1, V, n_iter, op_iter.data(), da, db, ai, bi,
                  SO, SD, 1.f, lambda, volumes.data(), flows.data(), err_weights.data(),
                  angles.data());
    write_floats(dir + "/out_static_img.bin", out_img.data(), N);
    write_floats(dir + "/out_static_norm.bin", out_norm.data(), out_norm.size());
    float *f[6];
    for (int c = 0; c < 6; ++c) f[c] = fields.data() + c * N;
    host_AUMISART(out_img.data(), out_proj.data(), out_norm.data(), out_ax.data(), zero.data(),
                  proj.data(), nx, ny, nz, na, nb, 1, V, n_iter, op_iter.data(), da, db, ai, bi,
                  SO, SD, 1.f, lambda, volumes.data(), flows.data(), err_weights.data(),
                  angles.data(), f[0], f[1], f[2], f[3], f[4], f[5]);
    write_floats(dir + "/out_motion_img.bin", out_img.data(), N);
    write_floats(dir + "/out_motion_norm.bin", out_norm.data(), out_norm.size());
    write_floats(dir + "/out_motion_ax.bin", out_ax.data(), N);
    write_floats(dir + "/out_motion_res.bin", out_proj.data(), D * V);
    std::printf("compat_check: %d failed calls\n", g_failures);
    return g_failures == 0 ? 0 : 1;
}

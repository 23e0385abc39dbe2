// Sanitizer run of the CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// Built by `make sanitize` in this directory with AddressSanitizer and UndefinedBehaviorSanitizer
// (every report aborts), this program drives each oracle function once on small inputs,
// including the inputs that used to reach undefined behaviour: warp displacements of +-1e20,
// +-Inf and NaN (float->int conversion out of range before the displacement clamp), detector
// footprints that leave the volume, and a motion sweep whose B(P q) has zeros (0/0 -> 0).  It
// prints one checksum per function and "SANITIZE_OK" at the end; tests/test_oracle.py compiles
// and runs it.
#include <cstdio>
#include <limits>

#include "mcsart_oracle.cpp"

namespace {

template <typename T>
std::vector<T> ramp(long long n, T scale, T offset) {
    std::vector<T> v((size_t)n);
    for (long long i = 0; i < n; ++i) v[(size_t)i] = offset + scale * (T)((i * 7919) % 97);
    return v;
}

template <typename T>
double checksum(const std::vector<T>& v) {
    double s = 0;
    for (const T& x : v) s += std::isfinite((double)x) ? (double)x : 0.0;
    return s;
}

template <typename T>
int run(const char* name) {
    const int nx = 12, ny = 12, nz = 9, na = 20, nb = 14, V = 4;
    const long long N = (long long)nx * ny * nz, D = (long long)na * nb;
    const T SO = 40, SD = 64, da = 1.1, db = 1.2;
    const T ai = -(T)(na / 2.0 - 0.5), bi = -(T)(nb / 2.0 - 0.5) + (T)3;   // offset detector
    std::vector<T> img = ramp<T>(N, (T)0.01, (T)0.2);
    std::vector<T> proj((size_t)D), back((size_t)N), out((size_t)N);
    const T angles[V] = {(T)0.0, (T)0.7853981633974483, (T)2.1, (T)4.0};
    projection<T>(proj.data(), img.data(), angles[1], SO, SD, da, na, ai, db, nb, bi, nx, ny, nz);
    std::printf("%s projection %.9g\n", name, checksum(proj));
    backprojection<T>(back.data(), proj.data(), angles[2], SO, SD, da, na, ai, db, nb, bi, nx, ny,
                      nz);
    std::printf("%s backprojection %.9g\n", name, checksum(back));

    // warps: fields with extreme and non-finite displacements
    std::vector<std::vector<T>> f(6);
    for (int c = 0; c < 6; ++c) f[c] = ramp<T>(N, (T)0.05, (T)(-2 + c));
    const T inf = std::numeric_limits<T>::infinity(), nan = std::numeric_limits<T>::quiet_NaN();
    const T extreme[5] = {(T)1e20, (T)-1e20, inf, -inf, nan};
    for (long long i = 0; i < N; ++i) {
        f[0][(size_t)i] = extreme[i % 5];
        f[2][(size_t)i] = extreme[(i + 2) % 5];
    }
    deform<T>(out.data(), img.data(), nx, ny, nz, (T)1.5, (T)-0.5, f[0].data(), f[1].data(),
              f[2].data(), f[3].data(), f[4].data(), f[5].data());
    std::printf("%s deform %.9g\n", name, checksum(out));
    deform2<T>(out.data(), img.data(), nx, ny, nz, (T)0.5, (T)0.25, f[1].data(), f[3].data(),
               f[4].data(), f[5].data(), f[3].data(), f[4].data());
    std::printf("%s deform2 %.9g\n", name, checksum(out));
    std::vector<T> m2(3 * (size_t)N);
    invert_dvf<T>(m2.data(), m2.data() + N, m2.data() + 2 * N, f[0].data(), f[1].data(),
                  f[2].data(), nx, ny, nz, 3);
    std::printf("%s invert_dvf %.9g\n", name, checksum(m2));

    // elementwise operators
    std::vector<T> a = ramp<T>(N, (T)0.3, (T)-1), b = ramp<T>(N, (T)0.2, (T)0);
    add2<T>(a.data(), b.data(), nx, ny, nz, img.data(), (T)0.7, 2);
    initial2<T>(b.data(), nx, ny, nz, img.data(), (T)0.3, (T)-0.1);
    division<T>(a.data(), b.data(), N);
    multiple<T>(a.data(), nx, ny, nz, img.data(), (T)0.5, 3);
    diff_backward<T>(out.data(), img.data(), nx, ny, nz, 1);
    std::vector<T> g[6];
    for (int c = 0; c < 6; ++c) g[c] = ramp<T>(N, (T)0.01, (T)(c - 3) * (T)0.1);
    update_udvf<T>(g[0].data(), g[1].data(), g[2].data(), g[3].data(), g[4].data(), g[5].data(),
                   img.data(), b.data(), (T)0.4, (T)-0.9, nx, ny, nz);
    std::printf("%s elementwise %.9g %.9g %.9g\n", name, checksum(a), checksum(out),
                checksum(g[0]));

    // whole drivers: both overloads, image and motion sweeps
    std::vector<T> hp((size_t)(D * V));
    for (int v = 0; v < V; ++v)
        projection<T>(hp.data() + v * D, img.data(), angles[v], SO, SD, da, na, ai, db, nb, bi,
                      nx, ny, nz);
    const T vol[V] = {0, (T)0.3, (T)-0.2, (T)0.5}, flo[V] = {(T)0.1, 0, (T)0.2, (T)-0.3};
    const int op[3] = {1, 0, 1};
    std::vector<T> oimg((size_t)N), oproj((size_t)(D * V)), onorm(3 * V), oax((size_t)N);
    std::vector<T> zero((size_t)N, (T)0);
    const Geo<T> geo{nx, ny, nz, na, nb, da, db, ai, bi, SO, SD};
    for (int motion = 0; motion < 2; ++motion) {
        const T* fl[6] = {g[0].data(), g[1].data(), g[2].data(), g[3].data(), g[4].data(),
                          g[5].data()};
        aumisart<T>(oimg.data(), oproj.data(), onorm.data(), oax.data(), zero.data(), hp.data(),
                    geo, V, 3, op, (T)1, vol, flo, angles, fl, motion != 0, motion != 0);
        std::printf("%s aumisart(motion=%d) %.9g %.9g\n", name, motion, checksum(oimg),
                    checksum(onorm));
    }
    return 0;
}

}  // namespace

int main() {
    omp_set_num_threads(2);
    run<double>("f64");
    run<float>("f32");
    std::printf("SANITIZE_OK\n");
    return 0;
}

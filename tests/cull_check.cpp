// Host check of the culling geometry (ConeCull in mcsart_amd/csrc/footprint.hpp): every cell a
// projector kernel reads with a nonzero weight lies in the view's read set, every cell it reads
// with zero weights only lies in the read set or the edge shell, and every voxel whose
// backprojection can be nonzero lies in the view's u support.
//
// The projector's reads are replayed here on the CPU from the same per-ray constants the kernels
// use (ray_dir / ray_t_setup / ray_z_setup / ray_setup, FP32 plane bounds):
//   rows kernel (k_project_rows): window start floor(clamp(t1, 0, nt)) clamped into [0, nt - W],
//     W = 2 or 3, the same along z;
//   masked kernel (k_project): start floor(t1) clamped into [-4, nt], cells clamped into the
//     volume;
//   wide kernel (k_project_wide): the footprint's own cells (plane_fp).
// The backprojector's u is nonzero only where the voxel's transverse cells and axial rows are both
// non-empty (src/cu_backprojection.cu:150-169, the kernel's acount > 0 and n > 0), evaluated here
// in FP64 from column_setup / column_b.
//
// Usage: cull_check nx ny nz na nb da db ai bi SO SD n_angles shrink [angle ...]
// (environment: CULL_CHECK_FORCE_SHARE=1, see share_check)
// shrink > 0 narrows both sets' margins by that many voxels (a check that the replay catches a
// set that is too small).  Prints the kept fractions and the number of violations; exit status
// 1 on any violation.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <vector>

#include "footprint.hpp"

using namespace mcsart;

namespace {

// Per voxel: bit 0 read at all, bit 1 read with a nonzero weight.
struct Marks {
    int nx, ny, nz;
    std::vector<unsigned char> m;
    void set(int ix, int iy, int iz, bool nonzero) {
        m[(size_t)ix + (size_t)nx * (iy + (size_t)ny * iz)] |= nonzero ? 3 : 1;
    }
};

float clampf(float x, float lo, float hi) { return std::fmin(std::fmax(x, lo), hi); }

// The rows kernel's window weights (k_project.hip, window()): F(b) = clamp(b, lo, hi) with the
// window's interior boundaries, the last one hi: the window covers the footprint, which is
// counted here where it does not (a window whose last boundary f0 + W falls below hi).
long long window_misses = 0;
void window_weights(float lo, float hi, float top, int W, float *w, float &f0) {
    f0 = clampf(std::floor(lo), 0.f, top);
    if (f0 + (float)W < hi) ++window_misses;
    float F = lo;
    for (int k = 0; k < W; ++k) {
        const float Fn = (k + 1 < W) ? clampf(f0 + (float)(k + 1), lo, hi) : hi;
        w[k] = Fn - F;
        F = Fn;
    }
}

// The image projector's shared z-rows (rows_share): a lane's R rays (detector rows ib0 ..
// ib0 + R - 1, the last ones repeated past the detector) read rows gmin .. gmin + WZ - 1 and
// gmax + WZ - 1 with gmin / gmax the smaller / larger window start of its first and last ray.
// Counted here: planes where that misses a ray's window (starts more than one row apart, or the
// end rays not holding the extremes) for R = 2 and 4 wherever rows_share allows them.
long long share_misses = 0, shared_views = 0;
// CULL_CHECK_FORCE_SHARE=1 checks R = 2 and 4 on every view, allowed or not (a check that the
// replay catches rays whose windows lie too far apart to share)
const bool force_share = std::getenv("CULL_CHECK_FORCE_SHARE") != nullptr;
void share_check(const std::vector<int> &g0, int nb, int np, int R) {
    for (int ib0 = 0; ib0 < nb; ib0 += R)
        for (int ip = 0; ip < np; ++ip) {
            int lo = 1 << 30, hi = -(1 << 30);
            for (int r = 0; r < R; ++r) {
                const int v = g0[(size_t)std::min(ib0 + r, nb - 1) * np + ip];
                lo = std::min(lo, v);
                hi = std::max(hi, v);
            }
            const int a = g0[(size_t)ib0 * np + ip];
            const int b = g0[(size_t)std::min(ib0 + R - 1, nb - 1) * np + ip];
            if (std::min(a, b) != lo || std::max(a, b) != hi || hi - lo > 1) ++share_misses;
        }
}

// k_project_rows: windows W x WZ from the clamped lower bounds
void rows_reads(const GeomArgs &g, const ViewInfo &vi, double c, double s, int W, int WZ,
                Marks &M) {
    for (int ia = 0; ia < g.na; ++ia) {
        const RayDir d = ray_dir(g, c, s, ia);
        RayT T;
        ray_t_setup(g, d, T);
        const bool xb = T.xbranch != 0;
        const int np = xb ? g.nx : g.ny, nt = xb ? g.ny : g.nx;
        const float p0 = 0.5f - 0.5f * (float)np;
        const float ntf = (float)nt, nzf = (float)g.nz;
        const float ttop = (float)(nt - W), ztop = (float)(g.nz - WZ);
        std::vector<int> gz((size_t)g.nb * np);   // window start row per (ib, ip)
        for (int ib = 0; ib < g.nb; ++ib) {
            RayZ Z;
            ray_z_setup(g, d, ib, Z);
            for (int ip = 0; ip < np; ++ip) {
                const float p = (float)ip + p0;
                const float t1 = std::fma(T.rlo, p, T.Klo), t2 = std::fma(T.rhi, p, T.Khi);
                const float z1 = std::fma(Z.rlo, p, Z.Klo), z2 = std::fma(Z.rhi, p, Z.Khi);
                float wt[3], wz[3], f0, g0;
                window_weights(clampf(t1, 0.f, ntf), clampf(t2, 0.f, ntf), ttop, W, wt, f0);
                window_weights(clampf(z1, 0.f, nzf), clampf(z2, 0.f, nzf), ztop, WZ, wz, g0);
                gz[(size_t)ib * np + ip] = (int)g0;
                for (int k = 0; k < W; ++k)
                    for (int j = 0; j < WZ; ++j) {
                        const int it = (int)f0 + k, iz = (int)g0 + j;
                        const bool nz = wt[k] != 0.f && wz[j] != 0.f;
                        if (xb) M.set(ip, it, iz, nz);
                        else M.set(it, ip, iz, nz);
                    }
            }
        }
        for (int R = 2; R <= 4; R *= 2)
            if (force_share || rows_share(vi, R)) share_check(gz, g.nb, np, R);
        if (ia == 0 && rows_share(vi, 2)) ++shared_views;
    }
}

// k_project: windows WT x WZ from floor(t1) clamped into [-4, nt], cells clamped into the volume
void masked_reads(const GeomArgs &g, double c, double s, int WT, int WZ, Marks &M) {
    for (int ia = 0; ia < g.na; ++ia)
        for (int ib = 0; ib < g.nb; ++ib) {
            Ray R;
            ray_setup(g, c, s, ia, ib, R);
            const bool xb = R.xbranch != 0;
            const int np = xb ? g.nx : g.ny, nt = xb ? g.ny : g.nx;
            const float p0 = 0.5f - 0.5f * (float)np;
            for (int ip = 0; ip < np; ++ip) {
                const float p = (float)ip + p0;
                const float ta = std::fma(R.rA, p, R.KA), tb = std::fma(R.rB, p, R.KB);
                const float za = std::fma(R.rzA, p, R.KzA), zb = std::fma(R.rzB, p, R.KzB);
                const float t1 = std::fmin(ta, tb), z1 = std::fmin(za, zb);
                const float t2 = std::fmax(ta, tb), z2 = std::fmax(za, zb);
                const bool in = (t2 > 0.f) & (t1 < (float)nt) & (z2 > 0.f) & (z1 < (float)g.nz);
                const float ft = clampf(std::floor(t1), -4.f, (float)nt);
                const float fz = clampf(std::floor(z1), -4.f, (float)g.nz);
                const int T0 = (int)ft, Z0 = (int)fz;
                for (int k = 0; k < WT; ++k)
                    for (int j = 0; j < WZ; ++j) {
                        // plane_weights(): masked to zero outside the plane or the volume
                        const float lk = (k == 0) ? t1 : ft + (float)k;
                        const float lj = (j == 0) ? z1 : fz + (float)j;
                        const float wk = std::fmax(std::fmin(ft + (float)(k + 1), t2) - lk, 0.f);
                        const float wj = std::fmax(std::fmin(fz + (float)(j + 1), z2) - lj, 0.f);
                        const bool vk = T0 + k >= 0 && T0 + k < nt, vj = Z0 + j >= 0 && Z0 + j < g.nz;
                        const bool nz = in && vk && vj && wk != 0.f && wj != 0.f;
                        const int it = clampi(T0 + k, 0, nt - 1), iz = clampi(Z0 + j, 0, g.nz - 1);
                        if (xb) M.set(ip, it, iz, nz);
                        else M.set(it, ip, iz, nz);
                    }
            }
        }
}

// k_project_wide: the footprint's cells inside the volume
void wide_reads(const GeomArgs &g, double c, double s, Marks &M) {
    for (int ia = 0; ia < g.na; ++ia)
        for (int ib = 0; ib < g.nb; ++ib) {
            Ray R;
            ray_setup(g, c, s, ia, ib, R);
            const bool xb = R.xbranch != 0;
            const int np = xb ? g.nx : g.ny, nt = xb ? g.ny : g.nx;
            const float p0 = 0.5f - 0.5f * (float)np;
            for (int ip = 0; ip < np; ++ip) {
                PlaneFP F;
                if (!plane_fp(R, (float)ip + p0, nt, g.nz, F)) continue;
                const int Tl = std::max(F.T0, 0);
                const int Th = std::min(nt - 1, (int)std::fmin(std::floor(F.t2), (float)nt));
                const int Zl = std::max(F.Z0, 0);
                const int Zh = std::min(g.nz - 1, (int)std::fmin(std::floor(F.z2), (float)g.nz));
                for (int i = Tl; i <= Th; ++i)
                    for (int k = Zl; k <= Zh; ++k) {
                        if (xb) M.set(ip, i, k, true);
                        else M.set(i, ip, k, true);
                    }
            }
        }
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 14) {
        std::fprintf(stderr, "usage: cull_check nx ny nz na nb da db ai bi SO SD n_angles shrink "
                             "[angle ...]\n");
        return 2;
    }
    GeomArgs g{atoi(argv[1]), atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atoi(argv[5]),
               (float)atof(argv[6]), (float)atof(argv[7]), (float)atof(argv[8]),
               (float)atof(argv[9]), (float)atof(argv[10]), (float)atof(argv[11])};
    const int n_uniform = atoi(argv[12]);
    std::vector<float> angles;
    for (int i = 0; i < n_uniform; ++i) angles.push_back((float)(2.0 * M_PI * i / n_uniform));
    const float shrink = (float)atof(argv[13]);
    for (int i = 14; i < argc; ++i) angles.push_back((float)atof(argv[i]));
    const size_t N = (size_t)g.nx * g.ny * g.nz;
    long long bad_read = 0, bad_support = 0;
    double kept = 0.0, support = 0.0;
    // per view, brute force: voxels in the read set, voxels with cells and rows, and the union
    // of the latter with the next view's read set (what mcsart_view_voxels counts)
    std::vector<unsigned char> prev_hit;
    long long prev_nhit = 0;
    int view = 0;
    auto print_counts = [&](int v, long long nread, long long nprep, long long nhit,
                            long long nunion) {
        std::printf("count %d %lld %lld %lld %lld\n", v, nread, nprep, nhit, nunion);
    };
    std::vector<long long> reads, preps;
    for (float ang : angles) {
        const double c = std::cos((double)ang), s = std::sin((double)ang);
        const ViewInfo vi = view_info(g, c, s);
        ConeCull rs = projector_read_set(g, c, s, vi, 1);
        ConeCull sup = backprojection_support(g, c, s, 1);
        for (int j = 0; j < 4; ++j) {   // a margin of h fewer voxels: k lower by h |n|_1
            const float l1 = std::fabs(rs.n[j][0]) + std::fabs(rs.n[j][1]) + std::fabs(rs.n[j][2]);
            rs.k[j] -= shrink * l1;
            sup.k[j] -= shrink * l1;
        }
        Marks M{g.nx, g.ny, g.nz, std::vector<unsigned char>(N, 0)};
        if (vi.wt <= 3 && vi.wz <= 3) {
            rows_reads(g, vi, c, s, vi.wt <= 2 ? 2 : 3, vi.wz <= 2 ? 2 : 3, M);
            masked_reads(g, c, s, vi.wt <= 2 ? 2 : 3, vi.wz <= 2 ? 2 : 3, M);
        } else {
            wide_reads(g, c, s, M);
        }
        long long nkeep = 0, nprep = 0, nsup = 0, viol = 0, vsup = 0, nhit = 0;
        std::vector<unsigned char> keepm(N, 0), hitm(N, 0);   // keepm: read set and shell
        for (int iz = 0; iz < g.nz; ++iz)
            for (int iy = 0; iy < g.ny; ++iy)
                for (int ix = 0; ix < g.nx; ++ix) {
                    const bool keep = cone_keep(rs, ix, iy, iz);
                    const bool shell = cone_shell(rs, ix, iy, iz);
                    nkeep += keep;
                    nprep += keep || shell;
                    keepm[(size_t)ix + (size_t)g.nx * (iy + (size_t)g.ny * iz)] = keep || shell;
                    const unsigned char mk = M.m[(size_t)ix + (size_t)g.nx * (iy + (size_t)g.ny * iz)];
                    // read with a nonzero weight: in the read set; read at all: or in the shell
                    if (((mk & 2) && !keep) || ((mk & 1) && !keep && !shell)) ++viol;
                }
        for (int iy = 0; iy < g.ny; ++iy)
            for (int ix = 0; ix < g.nx; ++ix) {
                Column C;
                column_setup(g, c, s, ix, iy, C);
                // acount > 0: Amin < na and Amax > max(floor(Amin), 0)
                const double ia_lo = std::fmin(std::fmax(std::floor(C.Amin), 0.0), (double)g.na);
                const bool cells = ia_lo < g.na && C.Amax > ia_lo;
                for (int iz = 0; iz < g.nz; ++iz) {
                    double Bmin, Bmax;
                    column_b(C, g.nz, iz, Bmin, Bmax);
                    const double ib_lo = std::fmax(0.0, std::floor(Bmin));
                    const bool rows = ib_lo < g.nb && Bmax > ib_lo;
                    const bool in = cone_box_meets(sup, (float)ix, (float)iy, (float)iz, 0.f, 0.f, 0.f);
                    nsup += in;
                    if (cells && rows && !in) ++vsup;
                    if (cells && rows) {
                        hitm[(size_t)ix + (size_t)g.nx * (iy + (size_t)g.ny * iz)] = 1;
                        ++nhit;
                    }
                }
            }
        if (view > 0) {   // the previous view's union with this view's read set
            long long nu = 0;
            for (size_t i = 0; i < N; ++i) nu += prev_hit[i] || keepm[i];
            print_counts(view - 1, reads.back(), preps.back(), prev_nhit, nu);
        }
        reads.push_back(nkeep);
        preps.push_back(nprep);
        prev_hit.swap(hitm);
        prev_nhit = nhit;
        ++view;
        bad_read += viol;
        bad_support += vsup;
        kept += (double)nkeep / N;
        support += (double)nsup / N;
        if (viol || vsup)
            std::printf("angle %.9g: %lld read cells outside the read set, %lld support voxels "
                        "outside the support (windows %d x %d)\n",
                        (double)ang, viol, vsup, vi.wt, vi.wz);
    }
    if (view > 0) print_counts(view - 1, reads.back(), preps.back(), prev_nhit, prev_nhit);
    std::printf("views %zu kept %.4f support %.4f bad_read %lld bad_support %lld "
                "window_misses %lld share_misses %lld shared_views %lld\n", angles.size(),
                kept / angles.size(), support / angles.size(), bad_read, bad_support,
                window_misses, share_misses, shared_views);
    return (bad_read || bad_support || window_misses || share_misses) ? 1 : 0;
}

// Footprint geometry of the projector and backprojector, shared by the device kernels.
//
// The reference forms footprint bounds as ratio*xc + y1 + ny/2 with xc ~ SO (src/cu_projection.cu:
// 52-61) and a/da - ai + 0.5 with a ~ na*da/2 (src/cu_backprojection.cu:135-138), in FP32.  Both are
// badly conditioned: the terms reach ~10^3 and cancel.  Here the same quantities are rearranged
// into per-ray / per-column constants evaluated in FP64, so that the per-plane / per-voxel work in
// FP32 only touches well-scaled numbers.  The mathematics is identical to the reference's.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace mcsart {

struct GeomArgs {
    int nx, ny, nz, na, nb;
    float da, db, ai, bi, SO, SD;
};

// ------------------------------------------------------------------------------------------
// Projector ray (src/cu_projection.cu:22-77).
//
// A ray from source S=(-SO c, -SO s, 0) to the detector cell centre.  If |x21| > |y21| (:39) the
// ray steps through x-planes ix and its footprint in each plane spans, along y, the min..max of
// four lines through the source with slopes yl/xl, yl/xr, yr/xl, yr/xr (:48-64), and along z
// the two lines with slopes zl/x21, zr/x21 (:67-77); otherwise the roles of x and y swap
// (:93-130).  With p = ip + 0.5 - n_plane/2 the plane centre offset, the plane's distance from
// the source is xc = p + SO*c (x-branch) or p + SO*s (y-branch), and a line of slope r gives the
// bound r*xc + (y1 + ny/2) = r*p + K with K = r*SO*c + y1 + ny/2.  The min/max over four slopes
// of r*xc equals the min/max over the two extreme slopes, for either sign of xc.
// ------------------------------------------------------------------------------------------
struct Ray {
    int xbranch;       // 1: planes are x-planes (transverse axis y)
    float rA, KA;      // transverse bound A: t = rA*p + KA
    float rB, KB;      // transverse bound B
    float rzA, KzA;    // z bound A
    float rzB, KzB;    // z bound B
};

__host__ __device__ inline void ray_setup(const GeomArgs &g, double c, double s, int ia, int ib,
                                          Ray &R) {
    const double SO = g.SO, SD = g.SD, da = g.da, db = g.db;
    const double x1 = -SO * c, y1 = -SO * s;                       // :25-26
    const double x20 = SD - SO;                                     // :28
    const double y20 = ((double)ia + (double)g.ai) * da;            // :29
    const double x2 = x20 * c - y20 * s, y2 = x20 * s + y20 * c;    // :30-31
    const double z2 = ((double)ib + (double)g.bi) * db;             // :32
    const double x21 = x2 - x1, y21 = y2 - y1, z21 = z2;            // :34-36
    const double hda = da / 2, hdb = db / 2;
    const double xl = x21 - hda * s, xr = x21 + hda * s;            // :48-49
    const double yl = y21 - hda * c, yr = y21 + hda * c;            // :50-51
    const double zl = z21 - hdb, zr = z21 + hdb;                    // :67-68
    double r0, r1, r2, r3, rz0, rz1, dist, base, zbase;
    R.xbranch = fabs(x21) > fabs(y21) ? 1 : 0;                      // :39
    if (R.xbranch) {
        r0 = yl / xl; r1 = yl / xr; r2 = yr / xl; r3 = yr / xr;     // :54-61
        rz0 = zl / x21; rz1 = zr / x21;                              // :71-74
        dist = SO * c;                                               // xc = p - x1
        base = y1 + (double)(g.ny / 2);                              // integer ny/2
    } else {
        r0 = xl / yl; r1 = xl / yr; r2 = xr / yl; r3 = xr / yr;     // :107-114
        rz0 = zl / y21; rz1 = zr / y21;                              // :124-127
        dist = SO * s;                                               // yc = p - y1
        base = x1 + (double)(g.nx / 2);                              // integer nx/2
    }
    const double rmin = fmin(fmin(r0, r1), fmin(r2, r3));
    const double rmax = fmax(fmax(r0, r1), fmax(r2, r3));
    zbase = (double)(g.nz / 2);                                      // integer nz/2
    R.rA = (float)rmin;
    R.KA = (float)(rmin * dist + base);
    R.rB = (float)rmax;
    R.KB = (float)(rmax * dist + base);
    R.rzA = (float)rz0;
    R.KzA = (float)(rz0 * dist + zbase);
    R.rzB = (float)rz1;
    R.KzB = (float)(rz1 * dist + zbase);
}

// The same constants split by what they depend on: the transverse lines depend on the detector
// column ia only (the ray's xy direction), the axial lines on (ia, ib).  Lines are ordered for the
// planes in front of the source (xc of the sign of x21 / y21): there t_lo = rlo*p + Klo is the
// lower bound and t_hi the upper one; z lines zl/x21 < zr/x21 keep their order there.  Planes
// behind the source (the reference evaluates every plane) lie outside the volume, where the
// bounds' order does not matter.
struct RayT {
    int xbranch;
    float rlo, Klo, rhi, Khi;
};
struct RayZ {
    float rlo, Klo, rhi, Khi;
};

struct RayDir {
    double x21, y21, dist, s_, c_;
    int xbranch;
};

__host__ __device__ inline RayDir ray_dir(const GeomArgs &g, double c, double s, int ia) {
    const double SO = g.SO, SD = g.SD, da = g.da;
    const double x1 = -SO * c, y1 = -SO * s;                       // :25-26
    const double x20 = SD - SO;                                     // :28
    const double y20 = ((double)ia + (double)g.ai) * da;            // :29
    const double x2 = x20 * c - y20 * s, y2 = x20 * s + y20 * c;    // :30-31
    RayDir d;
    d.x21 = x2 - x1;                                                // :34
    d.y21 = y2 - y1;                                                // :35
    d.xbranch = fabs(d.x21) > fabs(d.y21) ? 1 : 0;                  // :39
    d.dist = d.xbranch ? SO * c : SO * s;
    d.c_ = c;
    d.s_ = s;
    return d;
}

__host__ __device__ inline void ray_t_setup(const GeomArgs &g, const RayDir &d, RayT &R) {
    const double SO = g.SO, hda = g.da / 2, c = d.c_, s = d.s_;
    const double xl = d.x21 - hda * s, xr = d.x21 + hda * s;        // :48-49
    const double yl = d.y21 - hda * c, yr = d.y21 + hda * c;        // :50-51
    double r0, r1, r2, r3, base, fwd;
    if (d.xbranch) {
        r0 = yl / xl; r1 = yl / xr; r2 = yr / xl; r3 = yr / xr;     // :54-61
        base = -SO * s + (double)(g.ny / 2);                         // y1 + integer ny/2
        fwd = d.x21;
    } else {
        r0 = xl / yl; r1 = xl / yr; r2 = xr / yl; r3 = xr / yr;     // :107-114
        base = -SO * c + (double)(g.nx / 2);                         // x1 + integer nx/2
        fwd = d.y21;
    }
    double rmin = fmin(fmin(r0, r1), fmin(r2, r3));
    double rmax = fmax(fmax(r0, r1), fmax(r2, r3));
    if (fwd < 0) {   // xc < 0 in front of the source: the larger slope gives the lower bound
        const double t = rmin;
        rmin = rmax;
        rmax = t;
    }
    R.xbranch = d.xbranch;
    R.rlo = (float)rmin;
    R.Klo = (float)(rmin * d.dist + base);
    R.rhi = (float)rmax;
    R.Khi = (float)(rmax * d.dist + base);
}

__host__ __device__ inline void ray_z_setup(const GeomArgs &g, const RayDir &d, int ib, RayZ &Z) {
    const double z21 = ((double)ib + (double)g.bi) * (double)g.db;   // :32,:36
    const double hdb = (double)g.db / 2;
    const double den = d.xbranch ? d.x21 : d.y21;
    const double rz0 = (z21 - hdb) / den, rz1 = (z21 + hdb) / den;   // :67-74, :120-127
    const double zbase = (double)(g.nz / 2);                          // integer nz/2
    Z.rlo = (float)rz0;
    Z.Klo = (float)(rz0 * d.dist + zbase);
    Z.rhi = (float)rz1;
    Z.Khi = (float)(rz1 * d.dist + zbase);
}

// ------------------------------------------------------------------------------------------
// Backprojector column (src/cu_backprojection.cu:62-138).
//
// A voxel at centre (xc, yc, zc) (integer n/2, :66-68) is rotated by -phi (:82-97) into
// (xr0, yr0); six points (xr0 +-0.5, yr0 +-0.5, zc) and (xr0, yr0, zc +- 0.5) are mapped to the
// detector with ratio = SD/(x + SO) (:105-122).  The transverse (a) bounds depend only on the
// (ix, iy) column; the axial (b) bounds are the extremes of four lines in zc with slopes
// Rm = SD/(xr0-0.5+SO), Rp = SD/(xr0+0.5+SO) and Rc = SD/(xr0+SO) (the latter at zc +- 0.5).
// ------------------------------------------------------------------------------------------
struct Column {
    double Amin, Amax;   // a bounds in cell coordinates: a/da - ai + 0.5 (:135-136)
    double qm, qp, qc;   // b slopes divided by db
    double C0;           // 0.5 - bi
};

__host__ __device__ inline void column_setup(const GeomArgs &g, double c, double s, int ix, int iy,
                                             Column &C) {
    const double SO = g.SO, SD = g.SD;
    const double xc = (double)ix - (double)(g.nx / 2) + 0.5;   // :66
    const double yc = (double)iy - (double)(g.ny / 2) + 0.5;   // :67
    const double xr0 = xc * c + yc * s;                        // :92
    const double yr0 = -xc * s + yc * c;                       // :93
    const double Rm = SD / (xr0 - 0.5 + SO);                   // ll, lr (:105,:111)
    const double Rp = SD / (xr0 + 0.5 + SO);                   // rr, rl (:108,:114)
    const double Rc = SD / (xr0 + SO);                         // t, b (:117,:120)
    const double a_ll = Rm * (yr0 - 0.5), a_lr = Rm * (yr0 + 0.5);
    const double a_rr = Rp * (yr0 + 0.5), a_rl = Rp * (yr0 - 0.5);
    const double a_tb = Rc * yr0;
    const double amax = fmax(fmax(fmax(a_ll, a_lr), fmax(a_rr, a_rl)), a_tb);   // :129
    const double amin = fmin(fmin(fmin(a_ll, a_lr), fmin(a_rr, a_rl)), a_tb);   // :130
    const double da = g.da, db = g.db;
    C.Amax = amax / da - (double)g.ai + 0.5;
    C.Amin = amin / da - (double)g.ai + 0.5;
    C.qm = Rm / db;
    C.qp = Rp / db;
    C.qc = Rc / db;
    C.C0 = 0.5 - (double)g.bi;
}

// 1/x for x > 0 from the hardware estimate and two Newton steps (relative error far below FP64
// rounding of the quotients it replaces).
__device__ inline double rcp64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}

// column_setup for the device: the same constants, with the eight FP64 divisions replaced by
// the host's 1/da, 1/db and rcp64 (x + SO > 0: the source lies outside the volume).
__device__ inline void column_setup_dev(const GeomArgs &g, double c, double s, int ix, int iy,
                                        double ida, double idb, Column &C) {
    const double SO = g.SO, SD = g.SD;
    const double xc = (double)ix - (double)(g.nx / 2) + 0.5;   // :66
    const double yc = (double)iy - (double)(g.ny / 2) + 0.5;   // :67
    const double xr0 = xc * c + yc * s;                        // :92
    const double yr0 = -xc * s + yc * c;                       // :93
    const double Rm = SD * rcp64(xr0 - 0.5 + SO);              // ll, lr (:105,:111)
    const double Rp = SD * rcp64(xr0 + 0.5 + SO);              // rr, rl (:108,:114)
    const double Rc = SD * rcp64(xr0 + SO);                    // t, b (:117,:120)
    const double a_ll = Rm * (yr0 - 0.5), a_lr = Rm * (yr0 + 0.5);
    const double a_rr = Rp * (yr0 + 0.5), a_rl = Rp * (yr0 - 0.5);
    const double a_tb = Rc * yr0;
    const double amax = fmax(fmax(fmax(a_ll, a_lr), fmax(a_rr, a_rl)), a_tb);   // :129
    const double amin = fmin(fmin(fmin(a_ll, a_lr), fmin(a_rr, a_rl)), a_tb);   // :130
    C.Amax = amax * ida - (double)g.ai + 0.5;
    C.Amin = amin * ida - (double)g.ai + 0.5;
    C.qm = Rm * idb;
    C.qp = Rp * idb;
    C.qc = Rc * idb;
    C.C0 = 0.5 - (double)g.bi;
}

// b bounds of voxel iz of a column: b/db - bi + 0.5 over the six points (:131-132,:137-138).
__host__ __device__ inline void column_b(const Column &C, int nz, int iz, double &Bmin,
                                         double &Bmax) {
    const double zc = (double)iz - (double)(nz / 2) + 0.5;   // :68
    const double bm = C.qm * zc, bp = C.qp * zc;
    const double bt = C.qc * (zc + 0.5), bb = C.qc * (zc - 0.5);
    Bmax = fmax(fmax(bm, bp), fmax(bt, bb)) + C.C0;
    Bmin = fmin(fmin(bm, bp), fmin(bt, bb)) + C.C0;
}

}  // namespace mcsart

namespace mcsart {

// ------------------------------------------------------------------------------------------
// One plane of one ray: the footprint [t1, t2] x [z1, z2] in voxel-index units and the
// normalised overlap weights of src/cu_projection.cu:81-89 (x-branch) / :134-142 (y-branch):
// w(i) = (min(i+1, hi) - max(i, lo)) / (hi - lo) for i in [max(0, floor(lo)), min(n-1, floor(hi))].
// When the footprint spans at most two voxels along both axes (the common case), the weights of
// the two candidate voxels are returned directly (wt*, wz* already masked to zero for voxels
// outside the volume) and `fast` is set; otherwise the caller loops over the ranges.
// ------------------------------------------------------------------------------------------
struct PlaneFP {
    float t1, t2, z1, z2;   // bounds
    float it, iz;           // 1/(t2-t1), 1/(z2-z1)
    int T0, Z0;             // floor(t1), floor(z1)
    float wt0, wt1, wz0, wz1;
    bool fast;
};

__host__ __device__ inline bool plane_fp(const Ray &R, float p, int nt, int nz, PlaneFP &F) {
    const float ta = fmaf(R.rA, p, R.KA), tb = fmaf(R.rB, p, R.KB);
    const float za = fmaf(R.rzA, p, R.KzA), zb = fmaf(R.rzB, p, R.KzB);
    F.t1 = fminf(ta, tb);
    F.t2 = fmaxf(ta, tb);
    F.z1 = fminf(za, zb);
    F.z2 = fmaxf(za, zb);
    if (!(F.t2 > 0.f && F.t1 < (float)nt && F.z2 > 0.f && F.z1 < (float)nz)) return false;
    const float ft = floorf(F.t1), fz = floorf(F.z1);
    // t1 < nt and z1 < nz here; clamping the lower side keeps every float->int conversion in
    // range.  The fast path needs ft >= -1 (t2 > 0 and t2 < ft + 2), where T0 == ft exactly.
    F.T0 = (int)fmaxf(ft, -2.f);
    F.Z0 = (int)fmaxf(fz, -2.f);
    F.it = 1.0f / (F.t2 - F.t1);
    F.iz = 1.0f / (F.z2 - F.z1);
    const float wt0 = fminf(ft + 1.f, F.t2) - F.t1;
    const float wt1 = F.t2 - (ft + 1.f);
    const float wz0 = fminf(fz + 1.f, F.z2) - F.z1;
    const float wz1 = F.z2 - (fz + 1.f);
    F.fast = (wt1 < 1.f) && (wz1 < 1.f);
    const bool vt0 = (F.T0 >= 0) && (F.T0 < nt), vt1 = (F.T0 >= -1) && (F.T0 + 1 < nt);
    const bool vz0 = (F.Z0 >= 0) && (F.Z0 < nz), vz1 = (F.Z0 >= -1) && (F.Z0 + 1 < nz);
    F.wt0 = vt0 ? wt0 : 0.f;
    F.wt1 = vt1 ? fmaxf(wt1, 0.f) : 0.f;
    F.wz0 = vz0 ? wz0 : 0.f;
    F.wz1 = vz1 ? fmaxf(wz1, 0.f) : 0.f;
    return true;
}

__host__ __device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ------------------------------------------------------------------------------------------
// Per-view facts the host needs before launching the projector:
//   has_x / has_y : whether any ray takes the x-branch / y-branch (the branch depends on the
//                   detector column ia only: x21 and y21 do not involve ib);
//   wt / wz       : the most voxels any footprint can cover along the transverse axis / z, i.e.
//                   floor(max width) + 2.  The transverse width of ray ia in the plane at
//                   distance xc from the source is (rmax - rmin)*|xc|, the z width db*|xc|/|x21|
//                   (x-branch; mirrored for the y-branch); |xc| is largest at an end plane.
// ------------------------------------------------------------------------------------------
struct ViewInfo {
    int has_x, has_y;
    int wt, wz;      // projector windows
    int wa;          // backprojector transverse window (cells)
    double t_box;    // how far (voxels) a projector footprint's transverse bounds reach past the
                     // ray's own pyramid: the four-ratio box against the two edge ratios
    double zrow;     // the largest z extent of a projector footprint in any plane, which is also
                     // how far apart the footprints of adjacent detector rows lie (z voxels)
};

// The rays of R adjacent detector rows (one detector column) can read one window of WZ + 1
// z-rows together when their footprints' lower bounds lie less than one voxel apart in every
// plane, so that their windows start at most one row apart: (R - 1) zrow below 1, with a margin
// for the FP32 rounding of the per-plane bounds.
inline bool rows_share(const ViewInfo &v, int R) {
    return R >= 2 && (double)(R - 1) * v.zrow < 1.0 - 1e-3;
}

inline ViewInfo view_info(const GeomArgs &g, double c, double s) {
    ViewInfo v{0, 0, 0, 0, 0, 0.0, 0.0};
    double wt_max = 0.0, wz_max = 0.0;
    const double SO = g.SO, SD = g.SD, da = g.da, db = g.db;
    for (int ia = 0; ia < g.na; ++ia) {
        const double y20 = ((double)ia + (double)g.ai) * da;
        const double x21 = (SD - SO) * c - y20 * s + SO * c;
        const double y21 = (SD - SO) * s + y20 * c + SO * s;
        const double h = da / 2;
        const double xl = x21 - h * s, xr = x21 + h * s, yl = y21 - h * c, yr = y21 + h * c;
        double r0, r1, r2, r3, dist, zslope;
        int np;
        const bool xb = fabs(x21) > fabs(y21);
        if (xb) {
            r0 = yl / xl; r1 = yl / xr; r2 = yr / xl; r3 = yr / xr;
            dist = SO * c;
            zslope = db / fabs(x21);
            np = g.nx;
            v.has_x = 1;
        } else {
            r0 = xl / yl; r1 = xl / yr; r2 = xr / yl; r3 = xr / yr;
            dist = SO * s;
            zslope = db / fabs(y21);
            np = g.ny;
            v.has_y = 1;
        }
        const double rmin = fmin(fmin(r0, r1), fmin(r2, r3));
        const double rmax = fmax(fmax(r0, r1), fmax(r2, r3));
        const double dr = rmax - rmin;
        const double p0 = 0.5 - 0.5 * np, p1 = p0 + (np - 1);
        const double xc = fmax(fabs(p0 + dist), fabs(p1 + dist));
        wt_max = fmax(wt_max, dr * xc);
        wz_max = fmax(wz_max, zslope * xc);
        // the cell's edge rays run through (xr, yl) and (xl, yr): their ratios bound the ray's
        // pyramid; the box of four ratios reaches past them by at most this much in any plane
        const double e0 = xb ? yl / xr : xr / yl, e1 = xb ? yr / xl : xl / yr;
        const double excess = fmax(fmin(e0, e1) - rmin, rmax - fmax(e0, e1));
        v.t_box = fmax(v.t_box, excess * xc);
    }
    // a small margin absorbs the FP32 rounding of the per-plane bounds
    v.wt = (int)floor(wt_max * (1.0 + 1e-5) + 1e-5) + 2;
    v.wz = (int)floor(wz_max * (1.0 + 1e-5) + 1e-5) + 2;
    v.zrow = wz_max;
    // backprojector: the widest transverse footprint of a voxel column, in cells.  The width
    // grows toward the source and with |y'|, so over the square of columns it is largest on the
    // boundary, and the boundary columns are scanned; a column that still exceeds the window is
    // handled by a generic loop in the kernel.
    double wa_max = 0.0;
    auto scan = [&](int ix, int iy) {
        Column C;
        column_setup(g, c, s, ix, iy, C);
        wa_max = fmax(wa_max, C.Amax - C.Amin);
    };
    for (int ix = 0; ix < g.nx; ++ix) {
        scan(ix, 0);
        scan(ix, g.ny - 1);
    }
    for (int iy = 0; iy < g.ny; ++iy) {
        scan(0, iy);
        scan(g.nx - 1, iy);
    }
    // a column covers cells floor(Amin) .. ceil(Amax) - 1: at most floor(width) + 2 of them (the
    // guard absorbs the device's own FP64 rounding of the same bounds)
    v.wa = (int)floor(wa_max * (1.0 + 1e-9) + 1e-9) + 2;
    return v;
}

// ------------------------------------------------------------------------------------------
// The cone of a view: the points whose projection lands on the detector, i.e. detector
// coordinates a in [(ai - 1/2) da, (na + ai - 1/2) da] and b in [(bi - 1/2) db, (nb + bi - 1/2) db]
// at depth SD (the outer edges of the first and last cells, src/cu_projection.cu:28-36).  With
// depth t = (P - S).e and lateral l = (P - S).u (e = (c, s), u = (-s, c), S the source), a point
// is inside when the four linear forms SD l - a_lo t, a_hi t - SD l, SD z - b_lo t, b_hi t - SD z
// are all >= 0.  ConeCull holds them normalised (voxel units) as n.(ix, iy, iz) + k with voxel
// centres at i + 1/2 - floor(n/2), and k already raised by a box margin: a voxel passes when a
// box around its centre meets the cone.
//
// Two uses:
//  * the read set of a view's projector.  A window cell lies within W - 1/2 voxels (W = window
//    size) of a point of the ray's footprint box in its plane; the box reaches t_box past the
//    ray's pyramid transversely and, axially, b_max * (box width) / SD past it; the planes of odd
//    sizes sit 1/2 voxel off the integer grid.  The margin covers all of it.  A window clamped at
//    the volume's edge (a plane the ray misses, or a footprint outside [0, nz]) reads the edge
//    shell instead (cells within W of a transverse edge, any z; in columns inside the fan, within
//    W of a z edge), and only with weights of exactly zero: there any finite value, such as an
//    earlier view's, adds exactly nothing, so the shell is not part of the read set.  Other voxels
//    outside the read set are never read by that view's projector.  (A volume whose earlier
//    values may not be finite, as the motion sweep's g once its fields overflow, has its shell
//    written as well: 0 times Inf is NaN.)
//  * the support of a ratio backprojection u = B(r)/B(P q): u is 0 exactly at voxels whose
//    six-point footprint misses the detector (no cells or no rows, src/cu_backprojection.cu:
//    150-169); the six points lie within 1/sqrt(2) (xy) and 1/2 (z) of the centre.
// ------------------------------------------------------------------------------------------
struct ConeCull {
    float n[4][3];   // unit normals (x, y, z) of the four forms in voxel-index coordinates
    float k[4];      // offsets, margin included
    int sx, sz;      // edge shell: transverse cells < sx or >= n - sx; z cells < sz or >= nz - sz
    int nx, ny, nz;
    int mode;        // 0: every voxel passes; 1: cull; 2: cull, and write NaN where a write is
                     // culled outside the shell (a check that nothing reads those cells)
};

// Margin-raised form j at voxel index (x, y, z); the z term last, so that a column's part
// (cone_column) is the same partial sum.
__host__ __device__ inline float cone_form_xy(const ConeCull &C, int j, float x, float y) {
    return fmaf(C.n[j][0], x, fmaf(C.n[j][1], y, C.k[j]));
}
__host__ __device__ inline float cone_form(const ConeCull &C, int j, float x, float y, float z) {
    return fmaf(C.n[j][2], z, cone_form_xy(C, j, x, y));
}

// A column's share of the read-set test: the fan forms (0, 1: no z term) and the xy parts of the
// axial forms (2, 3), so that each voxel of the column costs two FMAs and compares.
struct ConeColumn {
    float p2, p3;
    bool fan, shell_xy;
};
__host__ __device__ inline ConeColumn cone_column(const ConeCull &C, int ix, int iy) {
    const float x = (float)ix, y = (float)iy;
    ConeColumn c;
    c.fan = cone_form_xy(C, 0, x, y) >= 0.f && cone_form_xy(C, 1, x, y) >= 0.f;
    c.p2 = cone_form_xy(C, 2, x, y);
    c.p3 = cone_form_xy(C, 3, x, y);
    c.shell_xy = ix < C.sx || ix >= C.nx - C.sx || iy < C.sx || iy >= C.ny - C.sx;
    return c;
}
__host__ __device__ inline bool cone_keep_at(const ConeCull &C, const ConeColumn &c, int iz) {
    if (C.mode == 0) return true;
    const float z = (float)iz;
    return c.fan && fmaf(C.n[2][2], z, c.p2) >= 0.f && fmaf(C.n[3][2], z, c.p3) >= 0.f;
}
__host__ __device__ inline bool cone_shell_at(const ConeCull &C, const ConeColumn &c, int iz) {
    return c.shell_xy || (c.fan && (iz < C.sz || iz >= C.nz - C.sz));
}

// Voxel (ix, iy, iz) is in the read set (or the cull is off).
__host__ __device__ inline bool cone_keep(const ConeCull &C, int ix, int iy, int iz) {
    return cone_keep_at(C, cone_column(C, ix, iy), iz);
}

// Voxel (ix, iy, iz) lies in the edge shell: within sx of a transverse edge, or within sz of a
// z edge in a column that passes the fan forms (a superset of the cells clamped windows read: a
// window clamped along z sits beside its ray's footprint transversely).
__host__ __device__ inline bool cone_shell(const ConeCull &C, int ix, int iy, int iz) {
    return cone_shell_at(C, cone_column(C, ix, iy), iz);
}

// A culled voxel whose write is made as a NaN in mode 2: one outside the read set and outside
// the edge shell (shell cells are read with zero weights, so they keep their earlier values).
__host__ __device__ inline bool cone_poison_at(const ConeCull &C, const ConeColumn &c, int iz) {
    return C.mode == 2 && !cone_shell_at(C, c, iz) && !cone_keep_at(C, c, iz);
}
__host__ __device__ inline bool cone_poison(const ConeCull &C, int ix, int iy, int iz) {
    return cone_poison_at(C, cone_column(C, ix, iy), iz);
}

// A box of centre (x, y, z) and half-widths (hx, hy, hz), on top of the built-in margin, meets
// the cone (or the cull is off).
__host__ __device__ inline bool cone_box_meets(const ConeCull &C, float x, float y, float z,
                                               float hx, float hy, float hz) {
    if (C.mode == 0) return true;
    bool in = true;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float m = fmaf(fabsf(C.n[j][0]), hx, fmaf(fabsf(C.n[j][1]), hy, fabsf(C.n[j][2]) * hz));
        in = in && cone_form(C, j, x, y, z) + m >= 0.f;
    }
    return in;
}

// The cone's forms for a view, with box half-widths (ht, ht, hz) around each voxel centre built
// into the offsets.  FP64 here; the device evaluates FP32 forms of magnitude ~n, far inside
// the 0.05-voxel guard added to every offset.
inline ConeCull cone_forms(const GeomArgs &g, double c, double s, double ht, double hz,
                           int mode) {
    ConeCull C{};
    C.nx = g.nx;
    C.ny = g.ny;
    C.nz = g.nz;
    C.mode = mode;
    const double SO = g.SO, SD = g.SD;
    const double alo = ((double)g.ai - 0.5) * g.da, ahi = ((double)g.na + g.ai - 0.5) * g.da;
    const double blo = ((double)g.bi - 0.5) * g.db, bhi = ((double)g.nb + g.bi - 0.5) * g.db;
    // t = c X + s Y + SO, l = -s X + c Y (S = -SO e); each form a.(X, Y, Z) + a0
    const double a[4][3] = {{-SD * s - alo * c, SD * c - alo * s, 0.0},
                            {ahi * c + SD * s, ahi * s - SD * c, 0.0},
                            {-blo * c, -blo * s, SD},
                            {bhi * c, bhi * s, -SD}};
    const double a0[4] = {-alo * SO, ahi * SO, -blo * SO, bhi * SO};
    const double o[3] = {0.5 - (double)(g.nx / 2), 0.5 - (double)(g.ny / 2),
                         0.5 - (double)(g.nz / 2)};
    const double h[3] = {ht, ht, hz};
    for (int j = 0; j < 4; ++j) {
        const double len = sqrt(a[j][0] * a[j][0] + a[j][1] * a[j][1] + a[j][2] * a[j][2]);
        double k = a0[j];
        for (int d = 0; d < 3; ++d) k += a[j][d] * o[d] + fabs(a[j][d]) * h[d];
        for (int d = 0; d < 3; ++d) C.n[j][d] = (float)(a[j][d] / len);
        C.k[j] = (float)(k / len + 0.05);
    }
    return C;
}

// Read set of the view's projector (first use above).
inline ConeCull projector_read_set(const GeomArgs &g, double c, double s, const ViewInfo &vi,
                                   int mode) {
    const double bmax = fmax(fabs(((double)g.bi - 0.5) * g.db),
                             fabs(((double)g.nb + g.bi - 0.5) * g.db));
    const double ht = (vi.wt - 0.5) + vi.t_box + 0.5;
    const double hz = (vi.wz - 0.5) + bmax * (vi.wt + vi.t_box + 0.5) / g.SD;
    ConeCull C = cone_forms(g, c, s, ht, hz, mode);
    C.sx = vi.wt;
    C.sz = vi.wz;
    return C;
}

// Support of the view's ratio backprojection (second use above); no shell.
inline ConeCull backprojection_support(const GeomArgs &g, double c, double s, int mode) {
    ConeCull C = cone_forms(g, c, s, 0.70710678118654752, 0.5, mode);
    C.sx = 0;
    C.sz = 0;
    return C;
}

}  // namespace mcsart

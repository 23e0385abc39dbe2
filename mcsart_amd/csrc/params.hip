// Parameter parsing of the MATLAB gateways, as a C ABI usable from any host.
//
// The reference parses MATLAB structs inline in each gateway (src/entry_AUMISART.cu:18-160,
// src/mex_backprojection.cu:13-59, src/mex_diff.cu:17-37 and the standalone projection gateway
// waste/8888mex_projection.cu:13-126), with field aliases, defaults and "MATLAB:badInput" errors.
// Here the same rules are written once.  The caller supplies a scalar lookup over its own struct
// type (a MATLAB struct in the MEX wrappers, a dict from Python, a JSON object in the file
// driver); array fields are only checked for presence here, in the reference's order.
#include <cstdio>
#include <cstring>
#include <string>

#include "internal.hpp"

using namespace mcsart;

namespace {

struct Lookup {
    mcsart_scalar_fn fn;
    void *user;
    bool has(const char *name, double *v) const {
        double t = 0.0;
        const int r = fn(user, name, &t);
        if (r && v) *v = t;
        return r != 0;
    }
};

int bad_input(const std::string &msg) { return fail(MCSART_EINVAL, "MATLAB:badInput: " + msg); }

int need_int(const Lookup &L, const char *name, const char *msg, int *out) {
    double v;
    if (!L.has(name, &v)) return bad_input(msg);
    *out = (int)v;
    return MCSART_OK;
}

int need_int2(const Lookup &L, const char *a, const char *b, const char *msg, int *out) {
    double v;
    if (L.has(a, &v) || L.has(b, &v)) {
        *out = (int)v;
        return MCSART_OK;
    }
    return bad_input(msg);
}

float or_default(const Lookup &L, const char *name, float dflt, unsigned bit, unsigned *mask) {
    double v;
    if (L.has(name, &v)) return (float)v;
    *mask |= bit;
    return dflt;
}

}  // namespace

extern "C" int mcsart_parse_geo(int gateway, mcsart_scalar_fn get, void *user,
                                mcsart_geo_params *out) {
    if (!get || !out) return fail(MCSART_EINVAL, "null argument");
    if (gateway < MCSART_GW_AUMISART || gateway > MCSART_GW_DIFF)
        return fail(MCSART_EINVAL, "unknown gateway");
    const Lookup L{get, user};
    mcsart_geo_params p;
    std::memset(&p, 0, sizeof(p));
    int rc;
    // volume (all gateways)
    if ((rc = need_int(L, "nx", "Can't found valid volume resolution nx.", &p.nx))) return rc;
    if ((rc = need_int(L, "ny", "Can't found valid volume resolution ny.", &p.ny))) return rc;
    if ((rc = need_int(L, "nz", "Can't found valid volume resolution nz.", &p.nz))) return rc;
    if (gateway == MCSART_GW_DIFF) {   // src/mex_diff.cu:34-37
        if ((rc = need_int(L, "ind", "Can't found ind.", &p.ind))) return rc;
        *out = p;
        return MCSART_OK;
    }
    // detector: na | nv, nb | nu (src/entry_AUMISART.cu:35-45)
    if ((rc = need_int2(L, "na", "nv",
                        "Can't found valid number of detector in plane, which is denoted as na "
                        "or nu.",
                        &p.na)))
        return rc;
    if ((rc = need_int2(L, "nb", "nu",
                        "Can't found valid number of detector across plane, which is denoted as "
                        "nb or nv.",
                        &p.nb)))
        return rc;
    if (gateway == MCSART_GW_AUMISART)   // voxel size (:48-55)
        p.dx = or_default(L, "dx", 1.0f, MCSART_DEF_DX, &p.defaulted);
    else
        p.dx = 1.0f;
    p.da = or_default(L, "da", 1.0f, MCSART_DEF_DA, &p.defaulted);   // :57-63
    p.db = or_default(L, "db", 1.0f, MCSART_DEF_DB, &p.defaulted);   // :66-71
    // offsets: user value minus (n/2 - 0.5) in float (:76-94)
    double v;
    if (L.has("ai", &v)) {
        p.ai = (float)v;
        p.ai -= ((float)p.na / 2 - 0.5f);
    } else {
        p.ai = -((float)p.na / 2 - 0.5f);
        p.defaulted |= MCSART_DEF_AI;
    }
    if (L.has("bi", &v)) {
        p.bi = (float)v;
        // the standalone projection gateway shifts only offsets above -1
        // (waste/8888mex_projection.cu:81-82)
        if (gateway != MCSART_GW_PROJECTION || p.bi > -1) p.bi -= ((float)p.nb / 2 - 0.5f);
    } else {
        p.bi = -((float)p.nb / 2 - 0.5f);
        p.defaulted |= MCSART_DEF_BI;
    }
    // SO | SI, SD | DI + SO (:96-108)
    if (L.has("SO", &v) || L.has("SI", &v))
        p.SO = (float)v;
    else
        return bad_input("Can't found valid distance between source and isocenter, which is "
                         "denoted with para.SO or para.DI.");
    if (L.has("SD", &v))
        p.SD = (float)v;
    else if (L.has("DI", &v))
        p.SD = (float)v + p.SO;
    else
        return bad_input("Can't found valid distance between source and detector plane, which "
                         "is denoted with para.SD or para.SI + para.DI.");
    if (gateway == MCSART_GW_BACKPROJECTION || gateway == MCSART_GW_PROJECTION) {
        // one view (src/mex_backprojection.cu:56-59)
        if (!L.has("angle", &v))
            return bad_input("Can't found valid projection angle, which is denoted with "
                             "para.angle.");
        p.angle = (float)v;
    }
    *out = p;
    return MCSART_OK;
}

extern "C" int mcsart_parse_iter(mcsart_scalar_fn get, void *user, mcsart_iter_params *out) {
    if (!get || !out) return fail(MCSART_EINVAL, "null argument");
    const Lookup L{get, user};
    mcsart_iter_params p;
    std::memset(&p, 0, sizeof(p));
    double v;
    int rc;
    if (L.has("n_iter", &v)) {   // src/entry_AUMISART.cu:111-117
        p.n_iter = (int)v;
    } else {
        p.n_iter = 1;
        p.defaulted |= MCSART_DEF_NITER;
    }
    if (!L.has("op_iter", nullptr))
        return bad_input("Can't found valid iteration distribution as iter_para.op_iter.");
    if ((rc = need_int(L, "n_views",
                       "Can't found valid number bins, which is denoted as iter_para.n_views.",
                       &p.n_views)))
        return rc;
    if (!L.has("lambda", &v)) return bad_input("Can't found valid coefficience iter_para.lambda.");
    p.lambda = (float)v;
    if ((rc = need_int(L, "outIter", "Can't found valid coefficience iter_para.outIter.",
                       &p.outIter)))
        return rc;
    static const char *const arrays[4][2] = {
        {"volumes", "Can't found valid volume in iter_para.volumes."},
        {"flows", "Can't found valid flow in iter_para.flows."},
        {"err_weights", "Can't found valid flow in iter_para.err_weights."},
        {"angles", "Can't found valid angles iter_para.angles."}};
    for (const auto &a : arrays)
        if (!L.has(a[0], nullptr)) return bad_input(a[1]);
    p.motion = L.has("alpha_x", nullptr) ? 1 : 0;   // overload choice (:193-207)
    *out = p;
    return MCSART_OK;
}

// mcsart_recon: file-driven McSART reconstruction without MATLAB.
//
//   mcsart_recon SCAN.json OUTDIR [--dry-run]
//
// SCAN.json holds the gateway's two structs and the names of raw FP32 files (MATLAB column-major
// order, the layouts of src/entry_AUMISART.cu:163-190: volumes (nx, ny, nz), projection stacks
// (na, nb, n_views)); file names are relative to the JSON file:
//
//   {"geo":  {"nx": 64, "ny": 64, "nz": 64, "na": 96, "nb": 96, "da": 1, "db": 1,
//             "SO": 250, "SD": 375, ...},                       // GEO_PARA fields and aliases
//    "iter": {"n_iter": 5, "op_iter": [1, 1, 1, 1, 1], "n_views": 60, "lambda": 1,
//             "outIter": 1, "volumes": [...], "flows": [...], "err_weights": [...],
//             "angles": [...]},                                 // ITER_PARA fields
//    "proj": "proj.raw",                                        // required
//    "img0": "img0.raw",                                        // optional (zero)
//    "fields": {"alpha_x": "ax.raw", "alpha_y": ..., "beta_z": ...}}   // optional: motion overload
//
// Writes OUTDIR/img.raw, nrm.raw (n_iter*n_views), alpha_x.raw, residual.raw and result.json: the
// four outputs of entry_AUMISART.  --dry-run parses and checks everything (parameters with the
// gateway's rules, file sizes), prints the parsed parameters as one JSON line and exits without
// touching the GPU.
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "mcsart_hip.h"

namespace {

// ---- a small JSON reader (objects, arrays, numbers, strings, true/false/null) -------------------
struct Json {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    double num = 0;
    bool b = false;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;
    const Json *get(const std::string &k) const {
        if (kind != Obj) return nullptr;
        auto it = obj.find(k);
        return it == obj.end() ? nullptr : &it->second;
    }
};

struct Parser {
    const std::string &s;
    size_t i = 0;
    std::string err;
    explicit Parser(const std::string &t) : s(t) {}
    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) ++i;
    }
    bool fail(const std::string &m) {
        if (err.empty()) err = m + " at offset " + std::to_string(i);
        return false;
    }
    bool lit(const char *w) {
        const size_t n = std::strlen(w);
        if (s.compare(i, n, w) != 0) return false;
        i += n;
        return true;
    }
    bool string(std::string &out) {
        if (i >= s.size() || s[i] != '"') return fail("expected string");
        ++i;
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c == '\\') {
                if (i >= s.size()) return fail("bad escape");
                const char e = s[i++];
                c = e == 'n' ? '\n' : e == 't' ? '\t' : e == 'r' ? '\r' : e;   // \" \\ \/
            }
            out.push_back(c);
        }
        if (i >= s.size()) return fail("unterminated string");
        ++i;
        return true;
    }
    bool value(Json &v) {
        ws();
        if (i >= s.size()) return fail("unexpected end");
        const char c = s[i];
        if (c == '{') {
            v.kind = Json::Obj;
            ++i;
            ws();
            if (i < s.size() && s[i] == '}') return ++i, true;
            while (true) {
                ws();
                std::string k;
                if (!string(k)) return false;
                ws();
                if (i >= s.size() || s[i] != ':') return fail("expected ':'");
                ++i;
                if (!value(v.obj[k])) return false;
                ws();
                if (i < s.size() && s[i] == ',') { ++i; continue; }
                if (i < s.size() && s[i] == '}') return ++i, true;
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = Json::Arr;
            ++i;
            ws();
            if (i < s.size() && s[i] == ']') return ++i, true;
            while (true) {
                v.arr.emplace_back();
                if (!value(v.arr.back())) return false;
                ws();
                if (i < s.size() && s[i] == ',') { ++i; continue; }
                if (i < s.size() && s[i] == ']') return ++i, true;
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = Json::Str;
            return string(v.str);
        }
        if (lit("true")) return v.kind = Json::Bool, v.b = true, true;
        if (lit("false")) return v.kind = Json::Bool, v.b = false, true;
        if (lit("null")) return v.kind = Json::Null, true;
        char *end = nullptr;
        errno = 0;
        const double d = std::strtod(s.c_str() + i, &end);
        if (end == s.c_str() + i || errno == ERANGE) return fail("bad value");
        i = (size_t)(end - s.c_str());
        v.kind = Json::Num;
        v.num = d;
        return true;
    }
};

// mcsart_scalar_fn over a JSON object: numbers, booleans, and the first element of arrays
int json_scalar(void *user, const char *name, double *value) {
    const Json *o = static_cast<const Json *>(user);
    const Json *f = o->get(name);
    if (!f || f->kind == Json::Null) return 0;
    double v = 0;
    if (f->kind == Json::Num) v = f->num;
    else if (f->kind == Json::Bool) v = f->b ? 1 : 0;
    else if (f->kind == Json::Arr && !f->arr.empty() && f->arr[0].kind == Json::Num) v = f->arr[0].num;
    if (value) *value = v;
    return 1;
}

template <class T>
bool numbers(const Json *f, size_t n, std::vector<T> &out, std::string &err, const char *name) {
    if (!f || f->kind != Json::Arr || f->arr.size() < n) {
        err = std::string("iter.") + name + " must be an array of at least " + std::to_string(n) +
              " numbers";
        return false;
    }
    out.resize(n);
    for (size_t k = 0; k < n; ++k) {
        if (f->arr[k].kind != Json::Num) {
            err = std::string("iter.") + name + " holds a non-number";
            return false;
        }
        out[k] = (T)f->arr[k].num;
    }
    return true;
}

std::string dir_of(const std::string &p) {
    const size_t k = p.find_last_of('/');
    return k == std::string::npos ? std::string(".") : p.substr(0, k);
}

bool read_raw(const std::string &path, size_t n, std::vector<float> &out, std::string &err) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return err = "cannot open " + path, false;
    const std::streamoff bytes = f.tellg();
    if (bytes != (std::streamoff)(n * sizeof(float))) {
        err = path + " holds " + std::to_string(bytes) + " bytes, expected " +
              std::to_string(n * sizeof(float)) + " (float32)";
        return false;
    }
    out.resize(n);
    f.seekg(0);
    f.read(reinterpret_cast<char *>(out.data()), (std::streamsize)(n * sizeof(float)));
    return (bool)f || (err = "short read from " + path, false);
}

bool write_raw(const std::string &path, const float *p, size_t n, std::string &err) {
    std::ofstream f(path, std::ios::binary);
    if (!f) return err = "cannot write " + path, false;
    f.write(reinterpret_cast<const char *>(p), (std::streamsize)(n * sizeof(float)));
    return (bool)f || (err = "short write to " + path, false);
}

int die(const std::string &msg, int code = 2) {
    std::fprintf(stderr, "mcsart_recon: %s\n", msg.c_str());
    return code;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) return die("usage: mcsart_recon SCAN.json OUTDIR [--dry-run]");
    const std::string json_path = argv[1], outdir = argv[2];
    const bool dry = argc > 3 && std::strcmp(argv[3], "--dry-run") == 0;
    std::ifstream jf(json_path);
    if (!jf) return die("cannot open " + json_path);
    std::stringstream ss;
    ss << jf.rdbuf();
    const std::string text = ss.str();
    Json root;
    Parser P(text);
    if (!P.value(root) || root.kind != Json::Obj) return die("bad JSON: " + P.err);
    const Json *geo = root.get("geo"), *iter = root.get("iter");
    if (!geo || geo->kind != Json::Obj || !iter || iter->kind != Json::Obj)
        return die("the JSON needs objects \"geo\" and \"iter\"");

    mcsart_geo_params g;
    if (mcsart_parse_geo(MCSART_GW_AUMISART, json_scalar, (void *)geo, &g) != MCSART_OK)
        return die(mcsart_last_error());
    mcsart_iter_params it;
    if (mcsart_parse_iter(json_scalar, (void *)iter, &it) != MCSART_OK)
        return die(mcsart_last_error());
    const size_t V = (size_t)it.n_views, N = (size_t)g.nx * g.ny * g.nz,
                 D = (size_t)g.na * g.nb;
    std::string err;
    std::vector<int> op_iter;
    std::vector<float> volumes, flows, err_w, angles;
    if (!numbers(iter->get("op_iter"), (size_t)it.n_iter, op_iter, err, "op_iter") ||
        !numbers(iter->get("volumes"), V, volumes, err, "volumes") ||
        !numbers(iter->get("flows"), V, flows, err, "flows") ||
        !numbers(iter->get("err_weights"), V, err_w, err, "err_weights") ||
        !numbers(iter->get("angles"), V, angles, err, "angles"))
        return die(err);

    const std::string base = dir_of(json_path) + "/";
    const Json *pj = root.get("proj");
    if (!pj || pj->kind != Json::Str) return die("\"proj\" must name the projection file");
    std::vector<float> proj, img0(N, 0.f);
    if (!read_raw(base + pj->str, D * V, proj, err)) return die(err);
    const Json *ij = root.get("img0");
    if (ij && ij->kind == Json::Str && !read_raw(base + ij->str, N, img0, err)) return die(err);
    const Json *fj = root.get("fields");
    const bool motion = fj && fj->kind == Json::Obj;
    static const char *names[6] = {"alpha_x", "alpha_y", "alpha_z", "beta_x", "beta_y", "beta_z"};
    std::vector<float> fields[6];
    if (motion)
        for (int c = 0; c < 6; ++c) {
            const Json *f = fj->get(names[c]);
            if (!f || f->kind != Json::Str)
                return die(std::string("fields.") + names[c] + " must name a file");
            if (!read_raw(base + f->str, N, fields[c], err)) return die(err);
        }

    if (dry) {
        std::printf("{\"nx\": %d, \"ny\": %d, \"nz\": %d, \"na\": %d, \"nb\": %d, \"dx\": %.9g, "
                    "\"da\": %.9g, \"db\": %.9g, \"ai\": %.9g, \"bi\": %.9g, \"SO\": %.9g, "
                    "\"SD\": %.9g, \"defaulted\": %u, \"n_iter\": %d, \"n_views\": %d, "
                    "\"outIter\": %d, \"lambda\": %.9g, \"motion\": %d, \"op_iter\": [",
                    g.nx, g.ny, g.nz, g.na, g.nb, g.dx, g.da, g.db, g.ai, g.bi, g.SO, g.SD,
                    g.defaulted, it.n_iter, it.n_views, it.outIter, it.lambda, motion ? 1 : 0);
        for (size_t k = 0; k < op_iter.size(); ++k) std::printf("%s%d", k ? ", " : "", op_iter[k]);
        std::printf("]}\n");
        return 0;
    }

    std::vector<float> out_img(N), out_nrm((size_t)it.n_iter * V), out_ax(N), out_res(D * V);
    const auto t0 = std::chrono::steady_clock::now();
    int rc;
    if (!motion)
        rc = mcsart_aumisart(out_img.data(), out_res.data(), out_nrm.data(), out_ax.data(),
                             img0.data(), proj.data(), g.nx, g.ny, g.nz, g.na, g.nb, it.outIter,
                             it.n_views, it.n_iter, op_iter.data(), g.da, g.db, g.ai, g.bi, g.SO,
                             g.SD, g.dx, it.lambda, volumes.data(), flows.data(), err_w.data(),
                             angles.data());
    else
        rc = mcsart_aumisart_motion(
            out_img.data(), out_res.data(), out_nrm.data(), out_ax.data(), img0.data(),
            proj.data(), g.nx, g.ny, g.nz, g.na, g.nb, it.outIter, it.n_views, it.n_iter,
            op_iter.data(), g.da, g.db, g.ai, g.bi, g.SO, g.SD, g.dx, it.lambda, volumes.data(),
            flows.data(), err_w.data(), angles.data(), fields[0].data(), fields[1].data(),
            fields[2].data(), fields[3].data(), fields[4].data(), fields[5].data());
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc != MCSART_OK)
        return die(std::string("reconstruction failed: ") + mcsart_last_error(), 3);
    const std::string o = outdir + "/";
    if (!write_raw(o + "img.raw", out_img.data(), N, err) ||
        !write_raw(o + "nrm.raw", out_nrm.data(), out_nrm.size(), err) ||
        !write_raw(o + "alpha_x.raw", out_ax.data(), N, err) ||
        !write_raw(o + "residual.raw", out_res.data(), D * V, err))
        return die(err);
    std::ofstream r(o + "result.json");
    r << "{\"img\": [" << g.nx << ", " << g.ny << ", " << g.nz << "], \"nrm\": ["
      << out_nrm.size() << ", 1], \"alpha_x\": [" << g.nx << ", " << g.ny << ", " << g.nz
      << "], \"residual\": [" << g.na << ", " << g.nb << ", " << V << "], \"motion\": "
      << (motion ? "true" : "false") << ", \"seconds\": " << secs << "}\n";
    return r ? 0 : die("cannot write result.json");
}

// extern "C" surface of the oracle for the Python test harness (ctypes).
// Test infrastructure only (see oracle.h).
#include <cstring>

#include "oracle.h"

using namespace oracle;

namespace {
bool parse_opt(const char* opt, size_t len, Value* v) {
    if (!opt || !len) return false;
    std::string err;
    return parse((const uint8_t*)opt, len, v, &err);
}
void put_err(char* err, size_t n, const std::string& m) {
    if (err && n) {
        std::strncpy(err, m.c_str(), n - 1);
        err[n - 1] = 0;
    }
}
}  // namespace

extern "C" {

// Option-string probe: returns 1 and fills the typed readings when the path exists.
// kind: 0 int, 1 float, 2 bool.  path: keys separated by '.'.
int oracle_opt_lookup(const char* opt, size_t len, const char* path, int kind, int* parsed, int64_t* iv,
                      double* fv) {
    Value v;
    std::string err;
    *parsed = parse((const uint8_t*)opt, len, &v, &err) ? 1 : 0;
    if (!*parsed) return 0;
    std::vector<std::string> keys;
    std::string cur;
    for (const char* p = path; *p; p++) {
        if (*p == '.') { keys.push_back(cur); cur.clear(); } else cur += *p;
    }
    keys.push_back(cur);
    const Value* at = &v;
    for (auto& k : keys) {
        at = at->get(k);
        if (!at) return 0;
    }
    Value root;
    root.kind = Value::Object;
    root.object.emplace_back("k", *at);
    if (kind == 0) return root.int_at({"k"}, iv) ? 1 : 0;
    if (kind == 1) return root.float_at({"k"}, fv) ? 1 : 0;
    bool b;
    if (!root.bool_at({"k"}, &b)) return 0;
    *iv = b;
    return 1;
}

// SuperOpts::from_opt; out = {width,height,pel,gpu,scale_down,scale_up,full,rc,levels}.
int oracle_super_opts(const char* opt, size_t len, int w, int h, int32_t* out, int64_t* sdata, int32_t* super_h,
                      char* err, size_t errlen) {
    Value v;
    bool have = parse_opt(opt, len, &v);
    SuperOpts o;
    std::string e = super_opts_from(have ? &v : nullptr, w, h, &o);
    if (!e.empty()) {
        put_err(err, errlen, e);
        return -3;
    }
    out[0] = o.width; out[1] = o.height; out[2] = o.pel; out[3] = o.gpu; out[4] = o.scale_down;
    out[5] = o.scale_up; out[6] = o.full; out[7] = o.rc; out[8] = o.levels;
    *sdata = super_pack_data(o);
    *super_h = super_height(o);
    return 0;
}

// get_frame_super (super_filter.rs:143-268): zero the three planes, build_plane each.
int oracle_super_build(const char* opt, size_t len, int w, int h, const uint8_t* sy, int ssy, const uint8_t* su,
                       int ssu, const uint8_t* sv, int ssv, uint8_t* dy, int dsy, uint8_t* du, int dsu, uint8_t* dv,
                       int dsv) {
    Value v;
    bool have = parse_opt(opt, len, &v);
    SuperOpts o;
    if (!super_opts_from(have ? &v : nullptr, w, h, &o).empty()) return -3;
    int sh = super_height(o);
    const uint8_t* src[3] = {sy, su, sv};
    int ss[3] = {ssy, ssu, ssv};
    uint8_t* dst[3] = {dy, du, dv};
    int ds[3] = {dsy, dsu, dsv};
    for (int p = 0; p < 3; p++) {
        size_t rows = p ? (size_t)sh / 2 : (size_t)sh;
        std::memset(dst[p], 0, (size_t)ds[p] * rows);
    }
    for (int p = 0; p < 3; p++) {
        size_t bw = p ? (size_t)w / 2 : (size_t)w, bh = p ? (size_t)h / 2 : (size_t)h;
        size_t rows = p ? (size_t)sh / 2 : (size_t)sh;
        build_plane(dst[p], (size_t)ds[p] * rows, (size_t)ds[p], src[p], (size_t)ss[p] * bh, (size_t)ss[p], bw, bh,
                    (size_t)w, (size_t)h, p != 0, o);
    }
    return 0;
}

}  // extern "C"

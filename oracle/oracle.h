// CPU oracle for the open-svpflow Super -> Analyse -> SmoothFps hot path.
//
// TEST INFRASTRUCTURE ONLY.  This is a plain C++ restatement of the
// reference's CPU path (Rust, crates/svpflow1, crates/svpflow-core,
// crates/svpflow2), written function by function with the reference
// file:line each part follows.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it, and only as the checker or as the
// timed CPU baseline; the product library never links or calls it.
//
// Build: oracle/Makefile (g++ -O3 -ffp-contract=off) -> oracle/_build/liboracle.so
//
// Pinning.  The reference holds no golden vectors for this path and cannot be built
// here, so each part is pinned by checks of its own in tests/ that restate the
// reference's rules independently (closed forms, textbook restatements, brute force,
// invariants):
//   - option parser, Super/Analyse/SmoothFps defaults: test_super_oracle.py,
//     test_lib_exports.py;
//   - Super build: constant/impulse closed forms and a naive per-pixel restatement;
//   - block costs: textbook SAD and Hadamard SATD (test_oracle_pins.py);
//   - the search hierarchy, block by block from the oracle's trace (BlockTrace):
//     visit orders, predictors, projection, global predictor, lambdas, seeds, the
//     predictor stage, exhaustive refinement (brute force), the bad-block count,
//     luma DC, rescale_scores and the refine pass (test_oracle_search_rules.py);
//   - hex2/UMH/exhaustive over whole frames: pure integer translation gives the exact
//     shift and fwd = -bwd (test_oracle_pins.py);
//   - TruncDiv, global predictor, rescale_scores, lambdas, luma_byte, scene classes,
//     coverage and magnitude masks as closed forms (test_oracle_closed_forms.py);
//   - the render: identical sources render the source for every algorithm;
//   - the light border and padding (light.cpp): the published srand/rand sequence,
//     the textbook Catmull-Rom curve, brute-force window means, the corner closed
//     forms, worked padding sizes and a numpy restatement of the strip and corner
//     layout (test_oracle_light.py).
// One rule of the reference has no observable effect and so no check: the lsad/3 of
// the coarsest level, where the lambda it scales is always 0.
#pragma once
#include <cstdint>
#include <cstddef>
#include <string>
#include <utility>
#include <vector>
#include <memory>

namespace oracle {

// ---------------------------------------------------------------- params.rs
// JSON5-ish option parser: crates/svpflow-core/src/params.rs:1-450.
struct Value {
    enum Kind { Object, Array, String, Number, Bool, Null } kind = Null;
    std::vector<std::pair<std::string, Value>> object;
    std::vector<Value> array;
    std::string str;
    double number = 0.0;
    bool boolean = false;

    const Value* get(const std::string& key) const;                   // params.rs:77-85 (last wins)
    const Value* at(std::initializer_list<const char*> path) const;   // params.rs:69-75
    bool bool_at(std::initializer_list<const char*> path, bool* out) const;     // :22-32
    bool int_at(std::initializer_list<const char*> path, int64_t* out) const;   // :34-43
    bool float_at(std::initializer_list<const char*> path, double* out) const;  // :45-52
    const std::vector<Value>* array_at(std::initializer_list<const char*> path) const;  // :61-66
};

// Returns false on a syntax error (the reference's Err), message in *err.
bool parse(const uint8_t* bytes, size_t len, Value* out, std::string* err);

// ---------------------------------------------------------------- super_opts.rs
struct SuperOpts {
    int pel = 2, gpu = 1, scale_down = 4, scale_up = 2;
    bool full = true, rc = false;
    int levels = 0, width = 0, height = 0;
};
int reduce_dim(int dim, int n);                                   // super_opts.rs:70-75
int auto_levels(int width, int height);                           // :77-87
int super_plane_height(int base_h, int pel, int levels, bool full);  // :89-95
// super_opts.rs:17-49; returns "" on success, else the reference's error string.
std::string super_opts_from(const Value* opt, int width, int height, SuperOpts* out);
int64_t super_pack_data(const SuperOpts& o);                      // :51-63
inline int super_height(const SuperOpts& o) { return super_plane_height(o.height, o.pel, o.levels, o.full); }

// ---------------------------------------------------------------- super_build.rs
// super_build.rs:8-93.  `dst` is the whole plane (dst_len bytes), pre-zeroed.
void build_plane(uint8_t* dst, size_t dst_len, size_t dst_stride, const uint8_t* src, size_t src_len,
                 size_t src_stride, size_t src_w, size_t src_h, size_t luma_w, size_t luma_h, bool chroma,
                 const SuperOpts& opts);

}  // namespace oracle

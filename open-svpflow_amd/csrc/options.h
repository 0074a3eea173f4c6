// Option strings of the svpflow filters (the `opt` argument of Super, Analyse and
// SmoothFps).  Same syntax as crates/svpflow-core/src/params.rs:1-450: JSON with
// unquoted identifier keys, // and /* */ comments between values, a lone '-' read as
// 0, and the LAST duplicate key winning on lookup.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace svpf {

class Opt {
public:
    enum Kind : uint8_t { kObject, kArray, kString, kNumber, kBool, kNull };

    // Parses `text`; returns false with a message on a syntax error (the caller then
    // treats the option string as absent, as the reference's `.ok()` does).
    static bool parse(const char* text, size_t len, Opt* out, std::string* error);

    Kind kind() const { return kind_; }
    // Typed lookups along a key path (params.rs:22-66).  Return false when the path
    // is missing or the value has no reading of that type.
    bool get_int(std::initializer_list<const char*> path, int64_t* v) const;
    bool get_float(std::initializer_list<const char*> path, double* v) const;
    bool get_bool(std::initializer_list<const char*> path, bool* v) const;
    bool get_string(std::initializer_list<const char*> path, std::string* v) const;
    const Opt* get_array_first(std::initializer_list<const char*> path) const;  // first element, if any
    const Opt* find(std::initializer_list<const char*> path) const;

private:
    friend struct OptReader;
    Kind kind_ = kNull;
    double num_ = 0;
    bool flag_ = false;
    std::string text_;
    std::vector<std::string> keys_;
    std::vector<Opt> items_;  // object members (parallel to keys_) or array elements
};

}  // namespace svpf

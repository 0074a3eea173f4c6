// Option-string reader; behaviour of crates/svpflow-core/src/params.rs:1-450.
#include "options.h"

#include <cmath>
#include <cstdlib>
#include <cstring>

namespace svpf {

struct OptReader {
    const unsigned char* s;
    size_t n, i = 0;
    std::string err;

    int cur() const { return i < n ? s[i] : -1; }
    bool at2(char a, char b) const { return i + 1 < n && s[i] == (unsigned char)a && s[i + 1] == (unsigned char)b; }
    bool eat(int c) { if (cur() == c) { ++i; return true; } return false; }
    bool bad(const char* m) { err = m; return false; }
    static bool space(int c) { return c == ' ' || c == '\n' || c == '\r' || c == '\t'; }
    static bool id0(int c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || c == '_'; }
    static bool digit(int c) { return c >= '0' && c <= '9'; }

    void line_comment() { i += 2; while (cur() != -1 && cur() != '\n' && cur() != '\r') ++i; }
    bool block_comment() {  // false when the input ends before "*/"
        i += 2;
        while (!at2('*', '/')) { if (cur() < 0) return false; ++i; }
        i += 2;
        return true;
    }
    // Whitespace and comments between tokens (params.rs:117-138).
    void gap() {
        for (;;) {
            while (space(cur())) ++i;
            if (at2('/', '/')) line_comment();
            else if (at2('/', '*')) { if (!block_comment()) return; }
            else return;
        }
    }
    // Before an object member name a comment is a syntax error (params.rs:146-168).
    bool member_gap() {
        while (space(cur())) ++i;
        if (at2('/', '/')) { line_comment(); return bad("Missing '}' or object member name"); }
        if (at2('/', '*')) { block_comment(); return bad("Missing '}' or object member name"); }
        return true;
    }

    static void utf8(std::string& o, uint32_t c) {
        if (c < 0x80) { o += (char)c; return; }
        if (c < 0x800) { o += (char)(0xC0 | c >> 6); o += (char)(0x80 | (c & 63)); return; }
        if (c < 0x10000) { o += (char)(0xE0 | c >> 12); o += (char)(0x80 | ((c >> 6) & 63)); o += (char)(0x80 | (c & 63)); return; }
        o += (char)(0xF0 | c >> 18); o += (char)(0x80 | ((c >> 12) & 63));
        o += (char)(0x80 | ((c >> 6) & 63)); o += (char)(0x80 | (c & 63));
    }
    bool hex(uint32_t* out) {
        uint32_t v = 0;
        for (int k = 0; k < 4; k++) {
            int c = cur();
            if (c < 0) return bad("Bad unicode escape sequence in string: hexadecimal digit expected.");
            ++i;
            int d = digit(c) ? c - '0' : (c >= 'a' && c <= 'f') ? c - 'a' + 10 : (c >= 'A' && c <= 'F') ? c - 'A' + 10 : -1;
            if (d < 0) return bad("Bad unicode escape sequence in string: hexadecimal digit expected.");
            v = v * 16 + (uint32_t)d;
        }
        *out = v;
        return true;
    }
    bool str(std::string* o) {
        ++i;  // opening quote
        for (;;) {
            int c = cur();
            if (c < 0) return bad("Syntax error: value, object or array expected.");
            ++i;
            if (c == '"') return true;
            if (c != '\\') { *o += (char)c; continue; }
            int e = cur();
            if (e >= 0) ++i;
            switch (e) {
                case '"': case '\\': case '/': *o += (char)e; break;
                case 'b': *o += '\b'; break;
                case 'f': *o += '\f'; break;
                case 'n': *o += '\n'; break;
                case 'r': *o += '\r'; break;
                case 't': *o += '\t'; break;
                case 'u': {
                    uint32_t hi;
                    if (!hex(&hi)) return false;
                    if (hi >= 0xd800 && hi <= 0xdbff) {
                        uint32_t lo;
                        if (!eat('\\') || !eat('u')) return bad("additional six characters expected to parse unicode surrogate pair.");
                        if (!hex(&lo)) return false;
                        if (lo < 0xdc00 || lo > 0xdfff) return bad("additional six characters expected to parse unicode surrogate pair.");
                        utf8(*o, 0x10000 + ((hi - 0xd800) << 10) + (lo - 0xdc00));
                    } else {
                        utf8(*o, (hi >= 0xdc00 && hi <= 0xdfff) ? 0xfffd : hi);
                    }
                    break;
                }
                default: return bad("Bad escape sequence in string");
            }
        }
    }
    bool num(Opt* o) {
        size_t b = i;
        bool neg = eat('-');
        bool dot_first = false;
        o->kind_ = Opt::kNumber;
        if (digit(cur())) {
            while (digit(cur())) ++i;
        } else if (neg && eat('.')) {
            if (!digit(cur())) return bad("is not a number.");
            while (digit(cur())) ++i;
            dot_first = true;
        } else if (neg) {
            o->num_ = 0.0;
            return true;
        } else {
            return bad("invalid number");
        }
        if (!dot_first && eat('.')) while (digit(cur())) ++i;
        if (eat('e') || eat('E')) {
            if (!eat('+')) eat('-');
            if (!digit(cur())) return bad("is not a number.");
            while (digit(cur())) ++i;
        }
        std::string t((const char*)s + b, i - b);
        if (dot_first) t = "-0" + t.substr(1);
        o->num_ = std::strtod(t.c_str(), nullptr);
        return true;
    }
    bool word(const char* w) {
        size_t l = std::strlen(w);
        if (i + l <= n && std::memcmp(s + i, w, l) == 0) { i += l; return true; }
        return bad("invalid literal");
    }
    bool value(Opt* o) {
        gap();
        int c = cur();
        if (c == '{') {
            ++i;
            o->kind_ = Opt::kObject;
            gap();
            if (eat('}')) return true;
            for (;;) {
                if (!member_gap()) return false;
                std::string key;
                if (cur() == '"') {
                    if (!str(&key)) return false;
                } else {
                    if (!id0(cur())) return bad("Missing '}' or object member name");
                    size_t b = i;
                    while (id0(cur()) || digit(cur())) ++i;
                    key.assign((const char*)s + b, i - b);
                }
                gap();
                if (!eat(':')) return bad("Missing ':' after object member name");
                Opt v;
                if (!value(&v)) return false;
                o->keys_.push_back(std::move(key));
                o->items_.push_back(std::move(v));
                gap();
                if (eat('}')) return true;
                if (!eat(',')) return bad("Missing ',' or '}' in object declaration");
            }
        }
        if (c == '[') {
            ++i;
            o->kind_ = Opt::kArray;
            gap();
            if (eat(']')) return true;
            for (;;) {
                Opt v;
                if (!value(&v)) return false;
                o->items_.push_back(std::move(v));
                gap();
                if (eat(']')) return true;
                if (!eat(',')) return bad("unexpected token");
            }
        }
        if (c == '"') { o->kind_ = Opt::kString; return str(&o->text_); }
        if (c == '-' || digit(c)) return num(o);
        if (c == 't') { o->kind_ = Opt::kBool; o->flag_ = true; return word("true"); }
        if (c == 'f') { o->kind_ = Opt::kBool; o->flag_ = false; return word("false"); }
        if (c == 'n') { o->kind_ = Opt::kNull; return word("null"); }
        return bad("Syntax error: value, object or array expected.");
    }
};

bool Opt::parse(const char* text, size_t len, Opt* out, std::string* error) {
    OptReader r{(const unsigned char*)text, len};
    r.gap();
    Opt v;
    if (!r.value(&v)) {
        if (error) *error = r.err;
        return false;
    }
    r.gap();
    if (r.i != r.n && v.kind_ != kObject) {  // trailing data is only tolerated after an object
        if (error) *error = "extra data after JSON value";
        return false;
    }
    *out = std::move(v);
    return true;
}

const Opt* Opt::find(std::initializer_list<const char*> path) const {
    const Opt* o = this;
    for (const char* key : path) {
        if (o->kind_ != kObject) return nullptr;
        const Opt* hit = nullptr;
        for (size_t k = o->keys_.size(); k-- > 0;)
            if (o->keys_[k] == key) { hit = &o->items_[k]; break; }
        if (!hit) return nullptr;
        o = hit;
    }
    return o;
}

bool Opt::get_int(std::initializer_list<const char*> path, int64_t* v) const {
    const Opt* o = find(path);
    if (!o) return false;
    switch (o->kind_) {
        case kNumber: {
            if (!std::isfinite(o->num_)) return false;
            double t = std::trunc(o->num_);
            if (!(t >= -9223372036854775808.0 && t < 9223372036854775808.0)) return false;
            *v = (int64_t)t;
            return true;
        }
        case kBool: *v = o->flag_; return true;
        case kNull: *v = 0; return true;
        default: return false;
    }
}

bool Opt::get_float(std::initializer_list<const char*> path, double* v) const {
    const Opt* o = find(path);
    if (!o) return false;
    switch (o->kind_) {
        case kNumber: *v = o->num_; return true;
        case kBool: *v = o->flag_ ? 1.0 : 0.0; return true;
        case kNull: *v = 0.0; return true;
        default: return false;
    }
}

bool Opt::get_bool(std::initializer_list<const char*> path, bool* v) const {
    const Opt* o = find(path);
    if (!o) return false;
    switch (o->kind_) {
        case kNumber: *v = o->num_ != 0.0; return true;
        case kString: *v = !o->text_.empty(); return true;
        case kObject: case kArray: *v = !o->items_.empty(); return true;
        case kBool: *v = o->flag_; return true;
        case kNull: *v = false; return true;
    }
    return false;
}

bool Opt::get_string(std::initializer_list<const char*> path, std::string* v) const {
    const Opt* o = find(path);
    if (!o || o->kind_ != kString) return false;
    *v = o->text_;
    return true;
}

const Opt* Opt::get_array_first(std::initializer_list<const char*> path) const {
    const Opt* o = find(path);
    if (!o || o->kind_ != kArray || o->items_.empty()) return nullptr;
    return &o->items_[0];
}

}  // namespace svpf

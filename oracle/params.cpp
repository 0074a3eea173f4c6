// Oracle restatement of crates/svpflow-core/src/params.rs (JSON5-ish parser).
// Test infrastructure only (see oracle.h).
#include "oracle.h"

#include <cmath>
#include <cstdlib>
#include <cstring>

namespace oracle {

const Value* Value::get(const std::string& key) const {  // params.rs:77-85: reverse search, last wins
    if (kind != Object) return nullptr;
    for (auto it = object.rbegin(); it != object.rend(); ++it)
        if (it->first == key) return &it->second;
    return nullptr;
}

const Value* Value::at(std::initializer_list<const char*> path) const {  // params.rs:69-75
    const Value* v = this;
    for (const char* k : path) {
        v = v->get(k);
        if (!v) return nullptr;
    }
    return v;
}

bool Value::bool_at(std::initializer_list<const char*> path, bool* out) const {  // params.rs:22-32
    const Value* v = at(path);
    if (!v) return false;
    switch (v->kind) {
        case Number: *out = v->number != 0.0; return true;
        case String: *out = !v->str.empty(); return true;
        case Object: *out = !v->object.empty(); return true;
        case Array: *out = !v->array.empty(); return true;
        case Bool: *out = v->boolean; return true;
        case Null: *out = false; return true;
    }
    return false;
}

bool Value::int_at(std::initializer_list<const char*> path, int64_t* out) const {  // params.rs:34-43
    const Value* v = at(path);
    if (!v) return false;
    if (v->kind == Number) {
        if (!std::isfinite(v->number)) return false;
        // value.trunc().to_string().parse::<i64>(): succeeds iff the truncated value is in i64 range.
        double t = std::trunc(v->number);
        if (t >= 9223372036854775808.0 || t < -9223372036854775808.0) return false;
        *out = (int64_t)t;
        return true;
    }
    if (v->kind == Bool) { *out = v->boolean ? 1 : 0; return true; }
    if (v->kind == Null) { *out = 0; return true; }
    return false;
}

bool Value::float_at(std::initializer_list<const char*> path, double* out) const {  // params.rs:45-52
    const Value* v = at(path);
    if (!v) return false;
    if (v->kind == Number) { *out = v->number; return true; }
    if (v->kind == Bool) { *out = v->boolean ? 1.0 : 0.0; return true; }
    if (v->kind == Null) { *out = 0.0; return true; }
    return false;
}

const std::vector<Value>* Value::array_at(std::initializer_list<const char*> path) const {  // :61-66
    const Value* v = at(path);
    if (!v || v->kind != Array) return nullptr;
    return &v->array;
}

namespace {

struct Parser {  // params.rs:88-449
    const uint8_t* b;
    size_t n;
    size_t pos = 0;
    std::string err;

    int peek() const { return pos < n ? b[pos] : -1; }
    int next() { int c = peek(); if (c >= 0) pos++; return c; }
    bool two(const char* s) const { return pos + 2 <= n && b[pos] == (uint8_t)s[0] && b[pos + 1] == (uint8_t)s[1]; }
    bool consume(int c) { if (peek() == c) { pos++; return true; } return false; }

    bool fail(const char* msg) { err = msg; return false; }

    void skip_space() {
        while (peek() == ' ' || peek() == '\n' || peek() == '\r' || peek() == '\t') pos++;
    }
    void skip_ws() {  // params.rs:117-138
        for (;;) {
            skip_space();
            if (two("//")) {
                pos += 2;
                while (peek() != -1 && peek() != '\n' && peek() != '\r') pos++;
            } else if (two("/*")) {
                pos += 2;
                while (!two("*/")) {
                    if (next() < 0) return;
                }
                pos += 2;
            } else {
                return;
            }
        }
    }
    bool skip_member_space() {  // params.rs:146-168: a comment before a member name is an error
        skip_space();
        if (two("//")) {
            pos += 2;
            while (peek() != -1 && peek() != '\n' && peek() != '\r') pos++;
            skip_space();
            return fail("Missing '}' or object member name");
        }
        if (two("/*")) {
            pos += 2;
            while (!two("*/")) {
                if (next() < 0) return fail("Missing '}' or object member name");
            }
            pos += 2;
            skip_space();
            return fail("Missing '}' or object member name");
        }
        return true;
    }

    bool value(Value* out) {  // params.rs:170-182
        skip_ws();
        int c = peek();
        if (c == '{') return object(out);
        if (c == '[') return array(out);
        if (c == '"') { out->kind = Value::String; return string(&out->str); }
        if (c == '-' || (c >= '0' && c <= '9')) return number(out);
        if (c == 't') { out->kind = Value::Bool; out->boolean = true; return literal("true"); }
        if (c == 'f') { out->kind = Value::Bool; out->boolean = false; return literal("false"); }
        if (c == 'n') { out->kind = Value::Null; return literal("null"); }
        return fail("Syntax error: value, object or array expected.");
    }

    bool object(Value* out) {  // params.rs:184-212
        out->kind = Value::Object;
        if (!consume('{')) return fail("unexpected token");
        skip_ws();
        if (consume('}')) return true;
        for (;;) {
            if (!skip_member_space()) return false;
            std::string key;
            if (peek() == '"') {
                if (!string(&key)) return false;
            } else if (!identifier(&key)) {
                return false;
            }
            skip_ws();
            if (!consume(':')) return fail("Missing ':' after object member name");
            Value v;
            if (!value(&v)) return false;
            out->object.emplace_back(std::move(key), std::move(v));
            skip_ws();
            if (consume('}')) return true;
            if (!consume(',')) return fail("Missing ',' or '}' in object declaration");
        }
    }

    bool array(Value* out) {  // params.rs:214-229
        out->kind = Value::Array;
        if (!consume('[')) return fail("unexpected token");
        skip_ws();
        if (consume(']')) return true;
        for (;;) {
            Value v;
            if (!value(&v)) return false;
            out->array.push_back(std::move(v));
            skip_ws();
            if (consume(']')) return true;
            if (!consume(',')) return fail("unexpected token");
        }
    }

    bool identifier(std::string* out) {  // params.rs:231-245
        size_t start = pos;
        int c = peek();
        if ((c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || c == '_') pos++;
        else return fail("Missing '}' or object member name");
        for (;;) {
            c = peek();
            if ((c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9') || c == '_') pos++;
            else break;
        }
        out->assign((const char*)b + start, pos - start);
        return true;
    }

    static void put_utf8(std::string* s, uint32_t cp) {
        if (cp < 0x80) s->push_back((char)cp);
        else if (cp < 0x800) { s->push_back((char)(0xC0 | (cp >> 6))); s->push_back((char)(0x80 | (cp & 63))); }
        else if (cp < 0x10000) {
            s->push_back((char)(0xE0 | (cp >> 12)));
            s->push_back((char)(0x80 | ((cp >> 6) & 63)));
            s->push_back((char)(0x80 | (cp & 63)));
        } else {
            s->push_back((char)(0xF0 | (cp >> 18)));
            s->push_back((char)(0x80 | ((cp >> 12) & 63)));
            s->push_back((char)(0x80 | ((cp >> 6) & 63)));
            s->push_back((char)(0x80 | (cp & 63)));
        }
    }

    bool hex4(uint32_t* code) {  // params.rs:285-306
        uint32_t c = 0;
        for (int i = 0; i < 4; i++) {
            int d = next();
            if (d >= '0' && d <= '9') c = c * 16 + (uint32_t)(d - '0');
            else if (d >= 'a' && d <= 'f') c = c * 16 + (uint32_t)(d - 'a' + 10);
            else if (d >= 'A' && d <= 'F') c = c * 16 + (uint32_t)(d - 'A' + 10);
            else return fail("Bad unicode escape sequence in string: hexadecimal digit expected.");
        }
        *code = c;
        return true;
    }

    bool string(std::string* out) {  // params.rs:247-283 and 308-335
        if (!consume('"')) return fail("unexpected token");
        std::string bytes;
        for (;;) {
            int c = next();
            if (c < 0) return fail("Syntax error: value, object or array expected.");
            if (c == '"') { *out = bytes; return true; }
            if (c != '\\') { bytes.push_back((char)c); continue; }
            int e = next();
            switch (e) {
                case '"': bytes.push_back('"'); break;
                case '\\': bytes.push_back('\\'); break;
                case '/': bytes.push_back('/'); break;
                case 'b': bytes.push_back('\b'); break;
                case 'f': bytes.push_back('\f'); break;
                case 'n': bytes.push_back('\n'); break;
                case 'r': bytes.push_back('\r'); break;
                case 't': bytes.push_back('\t'); break;
                case 'u': {
                    uint32_t code;
                    if (!hex4(&code)) return false;
                    if (code >= 0xd800 && code <= 0xdbff) {
                        if (next() != '\\' || next() != 'u')
                            return fail("additional six characters expected to parse unicode surrogate pair.");
                        uint32_t low;
                        if (!hex4(&low)) return false;
                        if (low < 0xdc00 || low > 0xdfff)
                            return fail("additional six characters expected to parse unicode surrogate pair.");
                        put_utf8(&bytes, 0x10000 + ((code - 0xd800) << 10) + (low - 0xdc00));
                    } else if (code >= 0xdc00 && code <= 0xdfff) {
                        put_utf8(&bytes, 0xfffd);
                    } else {
                        put_utf8(&bytes, code);
                    }
                    break;
                }
                default: return fail("Bad escape sequence in string");
            }
        }
    }

    bool digits() {
        size_t s = pos;
        while (peek() >= '0' && peek() <= '9') pos++;
        return pos != s;
    }

    bool number(Value* out) {  // params.rs:356-394
        size_t start = pos;
        bool negative = consume('-');
        bool leading_dot = false;
        if (peek() >= '0' && peek() <= '9') {
            pos++;
            digits();
        } else if (negative && consume('.')) {
            if (!digits()) return fail("is not a number.");
            leading_dot = true;
        } else if (negative) {
            out->kind = Value::Number;  // a lone '-' reads as 0.0 (params.rs:368-369)
            out->number = 0.0;
            return true;
        } else {
            return fail("invalid number");
        }
        if (!leading_dot && consume('.')) digits();
        if (consume('e') || consume('E')) {
            if (!consume('+')) consume('-');
            if (!digits()) return fail("is not a number.");
        }
        std::string text((const char*)b + start, pos - start);
        if (leading_dot) text = "-0" + text.substr(1);
        // Rust's f64::from_str accepts "1." and "-0.5"; strtod does too and rounds correctly.
        out->kind = Value::Number;
        out->number = std::strtod(text.c_str(), nullptr);
        return true;
    }

    bool literal(const char* lit) {  // params.rs:404-411
        size_t l = std::strlen(lit);
        if (pos + l <= n && std::memcmp(b + pos, lit, l) == 0) { pos += l; return true; }
        return fail("invalid literal");
    }
};

}  // namespace

bool parse(const uint8_t* bytes, size_t len, Value* out, std::string* err) {  // params.rs:1-11
    Parser p{bytes, len};
    p.skip_ws();
    Value v;
    if (!p.value(&v)) {
        if (err) *err = p.err;
        return false;
    }
    p.skip_ws();
    if (p.pos == p.n || v.kind == Value::Object) {
        *out = std::move(v);
        return true;
    }
    if (err) *err = "extra data after JSON value";
    return false;
}

}  // namespace oracle

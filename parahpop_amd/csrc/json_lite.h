// Minimal JSON reader/writer for CONFIG.json, TARGETS.json, constants.json and
// summary.json. Objects keep their key order (body order in the force sum
// follows the order written in the configuration). Numbers are parsed with
// strtod, which is correctly rounded; the writer prints the shortest digits
// that read back exactly.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace phx {
namespace json {

struct Value;
using Object = std::vector<std::pair<std::string, Value>>;
using Array = std::vector<Value>;

struct Value {
    enum Kind { Null, Bool, Number, String, Arr, Obj } kind = Null;
    bool b = false;
    double num = 0.0;
    bool integral = false;  // number text had no '.', 'e' or 'E'
    std::string str;
    std::shared_ptr<Array> arr;
    std::shared_ptr<Object> obj;

    static Value makeObject() { Value v; v.kind = Obj; v.obj = std::make_shared<Object>(); return v; }
    static Value makeArray() { Value v; v.kind = Arr; v.arr = std::make_shared<Array>(); return v; }
    static Value of(double d) { Value v; v.kind = Number; v.num = d; return v; }
    static Value ofInt(long long i) { Value v; v.kind = Number; v.num = double(i); v.integral = true; return v; }
    static Value of(bool x) { Value v; v.kind = Bool; v.b = x; return v; }
    static Value of(const std::string& s) { Value v; v.kind = String; v.str = s; return v; }

    bool isNull() const { return kind == Null; }
    bool isBool() const { return kind == Bool; }
    bool isNumber() const { return kind == Number; }
    bool isString() const { return kind == String; }
    bool isArray() const { return kind == Arr; }
    bool isObject() const { return kind == Obj; }

    const Value* find(const std::string& key) const {
        if (kind != Obj) return nullptr;
        for (const auto& kv : *obj) if (kv.first == key) return &kv.second;
        return nullptr;
    }
    bool has(const std::string& key) const { return find(key) != nullptr; }
    const Value& at(const std::string& key) const {
        const Value* v = find(key);
        if (!v) throw std::runtime_error("Missing key: " + key);
        return *v;
    }
    const Value& at(size_t i) const {
        if (kind != Arr || i >= arr->size()) throw std::runtime_error("Array index out of range");
        return (*arr)[i];
    }
    size_t size() const {
        if (kind == Arr) return arr->size();
        if (kind == Obj) return obj->size();
        return 0;
    }
    double asDouble() const {
        if (kind != Number) throw std::runtime_error("Expected a number");
        return num;
    }
    long long asInt() const {
        if (kind != Number || !integral) throw std::runtime_error("Expected an integer");
        return (long long)num;
    }
    bool asBool() const {
        if (kind != Bool) throw std::runtime_error("Expected a boolean");
        return b;
    }
    const std::string& asString() const {
        if (kind != String) throw std::runtime_error("Expected a string");
        return str;
    }
    void set(const std::string& key, const Value& v) {
        if (kind != Obj) throw std::runtime_error("set() on a non-object");
        for (auto& kv : *obj) if (kv.first == key) { kv.second = v; return; }
        obj->emplace_back(key, v);
    }
    void erase(const std::string& key) {
        if (kind != Obj) return;
        for (auto it = obj->begin(); it != obj->end(); ++it)
            if (it->first == key) { obj->erase(it); return; }
    }
    void push(const Value& v) {
        if (kind != Arr) throw std::runtime_error("push() on a non-array");
        arr->push_back(v);
    }
};

class Parser {
public:
    explicit Parser(const std::string& text) : s_(text) {}
    Value parse() {
        Value v = value_();
        ws_();
        if (p_ != s_.size()) fail_("trailing characters");
        return v;
    }

private:
    const std::string& s_;
    size_t p_ = 0;

    [[noreturn]] void fail_(const char* what) const {
        throw std::runtime_error(std::string("JSON parse error at offset ") + std::to_string(p_) + ": " + what);
    }
    void ws_() {
        while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\t' || s_[p_] == '\n' || s_[p_] == '\r')) ++p_;
    }
    bool lit_(const char* w) {
        size_t n = 0;
        while (w[n]) ++n;
        if (s_.compare(p_, n, w) == 0) { p_ += n; return true; }
        return false;
    }
    Value value_() {
        ws_();
        if (p_ >= s_.size()) fail_("unexpected end");
        char c = s_[p_];
        if (c == '{') return object_();
        if (c == '[') return array_();
        if (c == '"') { Value v; v.kind = Value::String; v.str = string_(); return v; }
        if (lit_("true")) return Value::of(true);
        if (lit_("false")) return Value::of(false);
        if (lit_("null")) return Value();
        return number_();
    }
    Value object_() {
        Value v = Value::makeObject();
        ++p_;
        ws_();
        if (p_ < s_.size() && s_[p_] == '}') { ++p_; return v; }
        for (;;) {
            ws_();
            if (p_ >= s_.size() || s_[p_] != '"') fail_("expected key");
            std::string k = string_();
            ws_();
            if (p_ >= s_.size() || s_[p_] != ':') fail_("expected ':'");
            ++p_;
            Value item = value_();
            for (const auto& kv : *v.obj)
                if (kv.first == k) fail_("duplicate key");
            v.obj->emplace_back(std::move(k), std::move(item));
            ws_();
            if (p_ < s_.size() && s_[p_] == ',') { ++p_; continue; }
            if (p_ < s_.size() && s_[p_] == '}') { ++p_; return v; }
            fail_("expected ',' or '}'");
        }
    }
    Value array_() {
        Value v = Value::makeArray();
        ++p_;
        ws_();
        if (p_ < s_.size() && s_[p_] == ']') { ++p_; return v; }
        for (;;) {
            v.arr->push_back(value_());
            ws_();
            if (p_ < s_.size() && s_[p_] == ',') { ++p_; continue; }
            if (p_ < s_.size() && s_[p_] == ']') { ++p_; return v; }
            fail_("expected ',' or ']'");
        }
    }
    static void utf8_(std::string& out, unsigned cp) {
        if (cp < 0x80) out += char(cp);
        else if (cp < 0x800) { out += char(0xC0 | (cp >> 6)); out += char(0x80 | (cp & 0x3F)); }
        else if (cp < 0x10000) { out += char(0xE0 | (cp >> 12)); out += char(0x80 | ((cp >> 6) & 0x3F)); out += char(0x80 | (cp & 0x3F)); }
        else { out += char(0xF0 | (cp >> 18)); out += char(0x80 | ((cp >> 12) & 0x3F)); out += char(0x80 | ((cp >> 6) & 0x3F)); out += char(0x80 | (cp & 0x3F)); }
    }
    unsigned hex4_() {
        if (p_ + 4 > s_.size()) fail_("bad \\u escape");
        unsigned cp = 0;
        for (int i = 0; i < 4; ++i) {
            char h = s_[p_++];
            cp <<= 4;
            if (h >= '0' && h <= '9') cp |= unsigned(h - '0');
            else if (h >= 'a' && h <= 'f') cp |= unsigned(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') cp |= unsigned(h - 'A' + 10);
            else fail_("bad hex digit");
        }
        return cp;
    }
    std::string string_() {
        ++p_;  // opening quote
        std::string out;
        while (p_ < s_.size()) {
            char c = s_[p_++];
            if (c == '"') return out;
            if (c != '\\') { out += c; continue; }
            if (p_ >= s_.size()) break;
            char e = s_[p_++];
            switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    unsigned cp = hex4_();
                    if (cp >= 0xD800 && cp < 0xDC00 && p_ + 6 <= s_.size() && s_[p_] == '\\' && s_[p_ + 1] == 'u') {
                        p_ += 2;
                        unsigned lo = hex4_();
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    utf8_(out, cp);
                    break;
                }
                default: fail_("bad escape");
            }
        }
        fail_("unterminated string");
    }
    Value number_() {
        size_t start = p_;
        bool integral = true;
        if (p_ < s_.size() && (s_[p_] == '-' || s_[p_] == '+')) ++p_;
        while (p_ < s_.size()) {
            char c = s_[p_];
            if (c >= '0' && c <= '9') { ++p_; continue; }
            if (c == '.' || c == 'e' || c == 'E' || c == '-' || c == '+') { integral = false; ++p_; continue; }
            break;
        }
        if (p_ == start) fail_("unexpected character");
        std::string tok = s_.substr(start, p_ - start);
        char* end = nullptr;
        double d = std::strtod(tok.c_str(), &end);
        if (end == tok.c_str() || *end != '\0') fail_("bad number");
        Value v = Value::of(d);
        v.integral = integral;
        return v;
    }
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

inline Value readFile(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot read file: " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    return parse(ss.str());
}

inline std::string quote(const std::string& s) {
    std::string out = "\"";
    for (char c : s) {
        switch (c) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\n': out += "\\n"; break;
            case '\r': out += "\\r"; break;
            case '\t': out += "\\t"; break;
            default:
                if ((unsigned char)c < 0x20) {
                    char buf[8];
                    std::snprintf(buf, sizeof buf, "\\u%04x", c);
                    out += buf;
                } else {
                    out += c;
                }
        }
    }
    return out + "\"";
}

// Shortest decimal digits that read back as v (v finite, nonzero, positive):
// digits "d1..dk" and n with v = 0.d1..dk x 10^n.
inline void shortestDigits(double v, std::string& digits, int& n) {
    char buf[40];
    for (int p = 1; p <= 17; ++p) {
        std::snprintf(buf, sizeof buf, "%.*e", p - 1, v);
        if (std::strtod(buf, nullptr) == v || p == 17) break;
    }
    // buf = "d.ddde[+-]XX"
    const char* e = std::strchr(buf, 'e');
    digits.clear();
    for (const char* c = buf; c < e; ++c)
        if (*c != '.') digits += *c;
    while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
    n = std::atoi(e + 1) + 1;
}

// Number text as nlohmann::json writes it: integers as parsed, doubles in
// their shortest round-trip digits, decimal notation for decimal exponents in
// (-4, 15] (".0" after an integral value), else d.ddde+XX with at least two
// exponent digits (nlohmann/detail/conversions/to_chars.hpp format_buffer).
inline std::string numberText(const Value& v) {
    if (v.integral && std::fabs(v.num) < 9.007199254740992e15) {
        char buf[32];
        std::snprintf(buf, sizeof buf, "%lld", (long long)v.num);
        return buf;
    }
    if (!std::isfinite(v.num)) return "null";
    std::string out = std::signbit(v.num) ? "-" : "";
    const double a = std::fabs(v.num);
    if (a == 0.0) return out + "0.0";
    std::string d;
    int n = 0;
    shortestDigits(a, d, n);
    const int k = int(d.size());
    constexpr int kMinExp = -4, kMaxExp = 15;
    if (k <= n && n <= kMaxExp) return out + d + std::string(size_t(n - k), '0') + ".0";
    if (0 < n && n <= kMaxExp) return out + d.substr(0, size_t(n)) + "." + d.substr(size_t(n));
    if (kMinExp < n && n <= 0) return out + "0." + std::string(size_t(-n), '0') + d;
    out += d.substr(0, 1);
    if (k > 1) out += "." + d.substr(1);
    const int e = n - 1;
    char buf[16];
    std::snprintf(buf, sizeof buf, "e%c%02d", e < 0 ? '-' : '+', e < 0 ? -e : e);
    return out + buf;
}

// Copy of v with every object's keys in ascending order (nlohmann::json keeps
// objects in a std::map, so its dumps list keys sorted).
inline Value sortedKeys(const Value& v) {
    if (v.kind == Value::Arr) {
        Value r = v;
        r.arr = std::make_shared<std::vector<Value>>();
        for (const Value& x : *v.arr) r.arr->push_back(sortedKeys(x));
        return r;
    }
    if (v.kind != Value::Obj) return v;
    Value r = Value::makeObject();
    std::vector<std::pair<std::string, Value>> items(v.obj->begin(), v.obj->end());
    std::stable_sort(items.begin(), items.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (auto& kv : items) r.obj->emplace_back(kv.first, sortedKeys(kv.second));
    return r;
}

inline void dump(const Value& v, std::string& out, int indent, int depth) {
    const std::string pad(size_t(indent * (depth + 1)), ' ');
    const std::string padEnd(size_t(indent * depth), ' ');
    switch (v.kind) {
        case Value::Null: out += "null"; break;
        case Value::Bool: out += v.b ? "true" : "false"; break;
        case Value::Number: out += numberText(v); break;
        case Value::String: out += quote(v.str); break;
        case Value::Arr:
            if (v.arr->empty()) { out += "[]"; break; }
            out += "[\n";
            for (size_t i = 0; i < v.arr->size(); ++i) {
                out += pad;
                dump((*v.arr)[i], out, indent, depth + 1);
                out += (i + 1 < v.arr->size()) ? ",\n" : "\n";
            }
            out += padEnd + "]";
            break;
        case Value::Obj:
            if (v.obj->empty()) { out += "{}"; break; }
            out += "{\n";
            for (size_t i = 0; i < v.obj->size(); ++i) {
                out += pad + quote((*v.obj)[i].first) + ": ";
                dump((*v.obj)[i].second, out, indent, depth + 1);
                out += (i + 1 < v.obj->size()) ? ",\n" : "\n";
            }
            out += padEnd + "}";
            break;
    }
}

inline std::string dump(const Value& v, int indent = 2) {
    std::string out;
    dump(v, out, indent, 0);
    return out;
}

}  // namespace json
}  // namespace phx

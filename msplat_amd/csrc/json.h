// json.h — a small JSON reader for the dataset loaders (transforms.json, Polycam camera
// files): objects, arrays, strings (with \uXXXX escapes, encoded as UTF-8), numbers,
// booleans and null. Numbers are kept as double. Parse errors throw std::runtime_error
// with the byte offset.
#pragma once
#include <cstdint>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace msplat_hip {

class Json {
public:
    enum Kind { Null, Bool, Number, String, Array, Object };

    Kind kind() const { return kind_; }
    bool is_null() const { return kind_ == Null; }
    bool is_object() const { return kind_ == Object; }
    bool is_array() const { return kind_ == Array; }
    bool is_number() const { return kind_ == Number; }
    bool is_string() const { return kind_ == String; }

    double number() const {
        if (kind_ == Number) return num_;
        if (kind_ == Bool) return b_ ? 1.0 : 0.0;
        throw std::runtime_error("JSON value is not a number");
    }
    const std::string& str() const {
        if (kind_ != String) throw std::runtime_error("JSON value is not a string");
        return s_;
    }
    size_t size() const { return kind_ == Array ? arr_.size() : kind_ == Object ? keys_.size() : 0; }
    const Json& operator[](size_t i) const {
        if (kind_ != Array || i >= arr_.size()) throw std::runtime_error("JSON array index out of range");
        return arr_[i];
    }
    bool contains(const std::string& k) const { return kind_ == Object && obj_.count(k) != 0; }
    const Json& at(const std::string& k) const {
        if (!contains(k)) throw std::runtime_error("JSON object has no key '" + k + "'");
        return arr_[obj_.at(k)];
    }
    // The value of key k as a number, or `def` when absent or null (nlohmann's value()).
    double value(const std::string& k, double def) const {
        if (!contains(k)) return def;
        const Json& v = at(k);
        return v.is_null() ? def : v.number();
    }
    std::string value(const std::string& k, const std::string& def) const {
        if (!contains(k)) return def;
        const Json& v = at(k);
        return v.is_string() ? v.s_ : def;
    }
    // Members of an object in document order, or elements of an array.
    const std::vector<std::string>& keys() const { return keys_; }
    const std::vector<Json>& items() const { return arr_; }

    static Json parse(const std::string& text) {
        size_t p = 0;
        Json v = parse_value(text, p, 0);
        skip_ws(text, p);
        if (p != text.size()) fail(p, "trailing characters");
        return v;
    }

private:
    Kind kind_ = Null;
    bool b_ = false;
    double num_ = 0.0;
    std::string s_;
    std::vector<Json> arr_;                  // array elements, or object values
    std::vector<std::string> keys_;          // object keys in order
    std::map<std::string, size_t> obj_;      // key -> index into arr_ (last one wins)

    [[noreturn]] static void fail(size_t p, const char* what) {
        throw std::runtime_error(std::string("JSON parse error at byte ") + std::to_string(p) + ": " + what);
    }
    static void skip_ws(const std::string& t, size_t& p) {
        while (p < t.size() && (t[p] == ' ' || t[p] == '\t' || t[p] == '\n' || t[p] == '\r')) p++;
    }
    static void put_utf8(std::string& out, uint32_t c) {
        if (c < 0x80) {
            out += (char)c;
        } else if (c < 0x800) {
            out += (char)(0xC0 | (c >> 6));
            out += (char)(0x80 | (c & 0x3F));
        } else if (c < 0x10000) {
            out += (char)(0xE0 | (c >> 12));
            out += (char)(0x80 | ((c >> 6) & 0x3F));
            out += (char)(0x80 | (c & 0x3F));
        } else {
            out += (char)(0xF0 | (c >> 18));
            out += (char)(0x80 | ((c >> 12) & 0x3F));
            out += (char)(0x80 | ((c >> 6) & 0x3F));
            out += (char)(0x80 | (c & 0x3F));
        }
    }
    static uint32_t hex4(const std::string& t, size_t& p) {
        if (p + 4 > t.size()) fail(p, "short \\u escape");
        uint32_t v = 0;
        for (int k = 0; k < 4; k++) {
            const char c = t[p++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
            else fail(p - 1, "bad hex digit");
        }
        return v;
    }
    static std::string parse_string(const std::string& t, size_t& p) {
        if (t[p] != '"') fail(p, "expected a string");
        p++;
        std::string out;
        while (true) {
            if (p >= t.size()) fail(p, "unterminated string");
            const char c = t[p++];
            if (c == '"') break;
            if (c != '\\') {
                out += c;
                continue;
            }
            if (p >= t.size()) fail(p, "unterminated escape");
            const char e = t[p++];
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
                    uint32_t c1 = hex4(t, p);
                    if (c1 >= 0xD800 && c1 < 0xDC00 && p + 1 < t.size() && t[p] == '\\' && t[p + 1] == 'u') {
                        p += 2;
                        const uint32_t c2 = hex4(t, p);
                        if (c2 < 0xDC00 || c2 >= 0xE000) fail(p, "bad surrogate pair");
                        c1 = 0x10000 + ((c1 - 0xD800) << 10) + (c2 - 0xDC00);
                    }
                    put_utf8(out, c1);
                    break;
                }
                default: fail(p - 1, "bad escape");
            }
        }
        return out;
    }
    static Json parse_value(const std::string& t, size_t& p, int depth) {
        if (depth > 512) fail(p, "nesting too deep");
        skip_ws(t, p);
        if (p >= t.size()) fail(p, "unexpected end");
        Json v;
        const char c = t[p];
        if (c == '{') {
            v.kind_ = Object;
            p++;
            skip_ws(t, p);
            if (p < t.size() && t[p] == '}') {
                p++;
                return v;
            }
            while (true) {
                skip_ws(t, p);
                if (p >= t.size()) fail(p, "unterminated object");
                std::string k = parse_string(t, p);
                skip_ws(t, p);
                if (p >= t.size() || t[p] != ':') fail(p, "expected ':'");
                p++;
                v.arr_.push_back(parse_value(t, p, depth + 1));
                if (!v.obj_.count(k)) v.keys_.push_back(k);
                v.obj_[k] = v.arr_.size() - 1;
                skip_ws(t, p);
                if (p >= t.size()) fail(p, "unterminated object");
                if (t[p] == ',') { p++; continue; }
                if (t[p] == '}') { p++; break; }
                fail(p, "expected ',' or '}'");
            }
        } else if (c == '[') {
            v.kind_ = Array;
            p++;
            skip_ws(t, p);
            if (p < t.size() && t[p] == ']') {
                p++;
                return v;
            }
            while (true) {
                v.arr_.push_back(parse_value(t, p, depth + 1));
                skip_ws(t, p);
                if (p >= t.size()) fail(p, "unterminated array");
                if (t[p] == ',') { p++; continue; }
                if (t[p] == ']') { p++; break; }
                fail(p, "expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind_ = String;
            v.s_ = parse_string(t, p);
        } else if (t.compare(p, 4, "true") == 0) {
            v.kind_ = Bool; v.b_ = true; p += 4;
        } else if (t.compare(p, 5, "false") == 0) {
            v.kind_ = Bool; v.b_ = false; p += 5;
        } else if (t.compare(p, 4, "null") == 0) {
            p += 4;
        } else if (c == '-' || (c >= '0' && c <= '9')) {
            const size_t s = p;
            if (t[p] == '-') p++;
            while (p < t.size() && ((t[p] >= '0' && t[p] <= '9') || t[p] == '.' || t[p] == 'e' || t[p] == 'E' ||
                                    t[p] == '+' || t[p] == '-'))
                p++;
            const std::string num = t.substr(s, p - s);
            char* end = nullptr;
            v.kind_ = Number;
            v.num_ = std::strtod(num.c_str(), &end);
            if (!end || *end != '\0') fail(s, "bad number");
        } else {
            fail(p, "unexpected character");
        }
        return v;
    }
};

}  // namespace msplat_hip

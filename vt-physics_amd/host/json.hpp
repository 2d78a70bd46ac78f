// json.hpp - minimal JSON value, parser and writer for the solver configs.
//
// The reference reads its solver templates and scene files with nlohmann::json
// (Simulator/Library/Source/JSON/JSONHandler.cpp:8-95).  That library is not in
// this image, so the host side carries this small self-contained value type:
// null / bool / number / string / array / object, insertion-ordered objects,
// strict RFC 8259 parsing (plus tolerance of a UTF-8 BOM), and a writer.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace vtp {

class Json {
public:
    enum class Type { Null, Bool, Number, String, Array, Object };

    Json() = default;
    Json(std::nullptr_t) {}
    Json(bool b) : type_(Type::Bool), b_(b) {}
    Json(int v) : type_(Type::Number), n_(v) {}
    Json(unsigned v) : type_(Type::Number), n_(v) {}
    Json(long v) : type_(Type::Number), n_(double(v)) {}
    Json(unsigned long v) : type_(Type::Number), n_(double(v)) {}
    Json(long long v) : type_(Type::Number), n_(double(v)) {}
    Json(unsigned long long v) : type_(Type::Number), n_(double(v)) {}
    Json(float v) : type_(Type::Number), n_(v) {}
    Json(double v) : type_(Type::Number), n_(v) {}
    Json(const char *s) : type_(Type::String), s_(s) {}
    Json(std::string s) : type_(Type::String), s_(std::move(s)) {}
    template <typename T>
    Json(const std::vector<T> &v) : type_(Type::Array) {
        for (const auto &e : v) a_.emplace_back(e);
    }
    Json(std::initializer_list<double> v) : type_(Type::Array) {
        for (double e : v) a_.emplace_back(e);
    }

    static Json array() { Json j; j.type_ = Type::Array; return j; }
    static Json object() { Json j; j.type_ = Type::Object; return j; }

    Type type() const { return type_; }
    bool is_null() const { return type_ == Type::Null; }
    bool is_bool() const { return type_ == Type::Bool; }
    bool is_number() const { return type_ == Type::Number; }
    bool is_string() const { return type_ == Type::String; }
    bool is_array() const { return type_ == Type::Array; }
    bool is_object() const { return type_ == Type::Object; }
    bool empty() const {
        if (type_ == Type::Null) return true;
        if (type_ == Type::Array) return a_.empty();
        if (type_ == Type::Object) return keys_.empty();
        return false;
    }
    size_t size() const {
        if (type_ == Type::Array) return a_.size();
        if (type_ == Type::Object) return keys_.size();
        return type_ == Type::Null ? 0 : 1;
    }

    bool contains(const std::string &k) const {
        return type_ == Type::Object && index_.count(k) != 0;
    }

    // Object access; a null value becomes an object on first write, as in
    // nlohmann::json.
    Json &operator[](const std::string &k) {
        if (type_ == Type::Null) type_ = Type::Object;
        if (type_ != Type::Object) throw std::runtime_error("json: not an object");
        auto it = index_.find(k);
        if (it != index_.end()) return vals_[it->second];
        index_[k] = vals_.size();
        keys_.push_back(k);
        vals_.emplace_back();
        return vals_.back();
    }
    Json &operator[](const char *k) { return (*this)[std::string(k)]; }
    const Json &at(const std::string &k) const {
        auto it = index_.find(k);
        if (type_ != Type::Object || it == index_.end())
            throw std::runtime_error("json: missing key '" + k + "'");
        return vals_[it->second];
    }
    const Json &operator[](const std::string &k) const { return at(k); }
    const Json &operator[](const char *k) const { return at(std::string(k)); }

    Json &operator[](size_t i) {
        if (type_ != Type::Array || i >= a_.size()) throw std::runtime_error("json: bad index");
        return a_[i];
    }
    const Json &operator[](size_t i) const {
        if (type_ != Type::Array || i >= a_.size()) throw std::runtime_error("json: bad index");
        return a_[i];
    }
    Json &operator[](int i) { return (*this)[size_t(i)]; }
    const Json &operator[](int i) const { return (*this)[size_t(i)]; }

    void push_back(Json v) {
        if (type_ == Type::Null) type_ = Type::Array;
        if (type_ != Type::Array) throw std::runtime_error("json: not an array");
        a_.push_back(std::move(v));
    }

    const std::vector<std::string> &keys() const { return keys_; }
    const std::vector<Json> &items() const { return a_; }

    double number() const {
        if (type_ == Type::Number) return n_;
        if (type_ == Type::Bool) return b_ ? 1.0 : 0.0;
        throw std::runtime_error("json: not a number");
    }
    bool boolean() const {
        if (type_ == Type::Bool) return b_;
        if (type_ == Type::Number) return n_ != 0.0;
        throw std::runtime_error("json: not a bool");
    }
    const std::string &str() const {
        if (type_ != Type::String) throw std::runtime_error("json: not a string");
        return s_;
    }
    std::vector<float> floats() const {
        if (type_ != Type::Array) throw std::runtime_error("json: not an array");
        std::vector<float> v;
        for (const auto &e : a_) v.push_back(float(e.number()));
        return v;
    }
    std::vector<int> ints() const {
        if (type_ != Type::Array) throw std::runtime_error("json: not an array");
        std::vector<int> v;
        for (const auto &e : a_) v.push_back(int(e.number()));
        return v;
    }
    std::vector<std::vector<float>> float_rows() const {
        if (type_ != Type::Array) throw std::runtime_error("json: not an array");
        std::vector<std::vector<float>> v;
        for (const auto &e : a_) v.push_back(e.floats());
        return v;
    }

    bool operator==(const Json &o) const {
        if (type_ != o.type_) return false;
        switch (type_) {
            case Type::Null: return true;
            case Type::Bool: return b_ == o.b_;
            case Type::Number: return n_ == o.n_;
            case Type::String: return s_ == o.s_;
            case Type::Array: return a_ == o.a_;
            case Type::Object:
                if (keys_.size() != o.keys_.size()) return false;
                for (const auto &k : keys_)
                    if (!o.contains(k) || !(at(k) == o.at(k))) return false;
                return true;
        }
        return false;
    }
    bool operator!=(const Json &o) const { return !(*this == o); }

    static Json parse(const std::string &text) {
        Parser p{text, 0};
        if (text.size() >= 3 && (unsigned char)text[0] == 0xEF &&
            (unsigned char)text[1] == 0xBB && (unsigned char)text[2] == 0xBF)
            p.i = 3;
        p.ws();
        Json v = p.value();
        p.ws();
        if (p.i != text.size()) p.fail("trailing characters");
        return v;
    }

    std::string dump(int indent = -1) const {
        std::string out;
        write(out, indent, 0);
        return out;
    }

private:
    struct Parser {
        const std::string &t;
        size_t i;
        [[noreturn]] void fail(const char *what) const {
            throw std::runtime_error(std::string("json parse error at offset ") +
                                     std::to_string(i) + ": " + what);
        }
        void ws() {
            while (i < t.size() && (t[i] == ' ' || t[i] == '\t' || t[i] == '\n' || t[i] == '\r')) ++i;
        }
        bool lit(const char *s) {
            size_t n = std::char_traits<char>::length(s);
            if (t.compare(i, n, s) == 0) { i += n; return true; }
            return false;
        }
        Json value() {
            if (i >= t.size()) fail("unexpected end");
            char c = t[i];
            if (c == '{') return obj();
            if (c == '[') return arr();
            if (c == '"') return Json(string());
            if (lit("true")) return Json(true);
            if (lit("false")) return Json(false);
            if (lit("null")) return Json();
            if (c == '-' || (c >= '0' && c <= '9')) return num();
            fail("unexpected character");
        }
        Json num() {
            size_t s = i;
            if (t[i] == '-') ++i;
            if (i >= t.size() || !(t[i] >= '0' && t[i] <= '9')) fail("bad number");
            if (t[i] == '0') ++i; else while (i < t.size() && t[i] >= '0' && t[i] <= '9') ++i;
            if (i < t.size() && t[i] == '.') {
                ++i;
                if (i >= t.size() || !(t[i] >= '0' && t[i] <= '9')) fail("bad fraction");
                while (i < t.size() && t[i] >= '0' && t[i] <= '9') ++i;
            }
            if (i < t.size() && (t[i] == 'e' || t[i] == 'E')) {
                ++i;
                if (i < t.size() && (t[i] == '+' || t[i] == '-')) ++i;
                if (i >= t.size() || !(t[i] >= '0' && t[i] <= '9')) fail("bad exponent");
                while (i < t.size() && t[i] >= '0' && t[i] <= '9') ++i;
            }
            return Json(std::strtod(t.c_str() + s, nullptr));
        }
        static void utf8(std::string &o, uint32_t cp) {
            if (cp < 0x80) o += char(cp);
            else if (cp < 0x800) { o += char(0xC0 | (cp >> 6)); o += char(0x80 | (cp & 0x3F)); }
            else if (cp < 0x10000) {
                o += char(0xE0 | (cp >> 12)); o += char(0x80 | ((cp >> 6) & 0x3F));
                o += char(0x80 | (cp & 0x3F));
            } else {
                o += char(0xF0 | (cp >> 18)); o += char(0x80 | ((cp >> 12) & 0x3F));
                o += char(0x80 | ((cp >> 6) & 0x3F)); o += char(0x80 | (cp & 0x3F));
            }
        }
        uint32_t hex4() {
            if (i + 4 > t.size()) fail("bad \\u escape");
            uint32_t v = 0;
            for (int k = 0; k < 4; ++k) {
                char c = t[i++];
                v <<= 4;
                if (c >= '0' && c <= '9') v |= uint32_t(c - '0');
                else if (c >= 'a' && c <= 'f') v |= uint32_t(c - 'a' + 10);
                else if (c >= 'A' && c <= 'F') v |= uint32_t(c - 'A' + 10);
                else fail("bad hex digit");
            }
            return v;
        }
        std::string string() {
            ++i;  // opening quote
            std::string o;
            while (true) {
                if (i >= t.size()) fail("unterminated string");
                char c = t[i++];
                if (c == '"') break;
                if ((unsigned char)c < 0x20) fail("control character in string");
                if (c != '\\') { o += c; continue; }
                if (i >= t.size()) fail("bad escape");
                char e = t[i++];
                switch (e) {
                    case '"': o += '"'; break;
                    case '\\': o += '\\'; break;
                    case '/': o += '/'; break;
                    case 'b': o += '\b'; break;
                    case 'f': o += '\f'; break;
                    case 'n': o += '\n'; break;
                    case 'r': o += '\r'; break;
                    case 't': o += '\t'; break;
                    case 'u': {
                        uint32_t cp = hex4();
                        if (cp >= 0xD800 && cp < 0xDC00 && i + 1 < t.size() && t[i] == '\\' && t[i + 1] == 'u') {
                            i += 2;
                            uint32_t lo = hex4();
                            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                        }
                        utf8(o, cp);
                        break;
                    }
                    default: fail("bad escape");
                }
            }
            return o;
        }
        Json arr() {
            ++i;
            Json a = Json::array();
            ws();
            if (i < t.size() && t[i] == ']') { ++i; return a; }
            while (true) {
                ws();
                a.push_back(value());
                ws();
                if (i < t.size() && t[i] == ',') { ++i; continue; }
                if (i < t.size() && t[i] == ']') { ++i; return a; }
                fail("expected ',' or ']'");
            }
        }
        Json obj() {
            ++i;
            Json o = Json::object();
            ws();
            if (i < t.size() && t[i] == '}') { ++i; return o; }
            while (true) {
                ws();
                if (i >= t.size() || t[i] != '"') fail("expected key");
                std::string k = string();
                ws();
                if (i >= t.size() || t[i] != ':') fail("expected ':'");
                ++i;
                ws();
                o[k] = value();
                ws();
                if (i < t.size() && t[i] == ',') { ++i; continue; }
                if (i < t.size() && t[i] == '}') { ++i; return o; }
                fail("expected ',' or '}'");
            }
        }
    };

    static void escape(std::string &out, const std::string &s) {
        out += '"';
        for (unsigned char c : s) {
            switch (c) {
                case '"': out += "\\\""; break;
                case '\\': out += "\\\\"; break;
                case '\n': out += "\\n"; break;
                case '\r': out += "\\r"; break;
                case '\t': out += "\\t"; break;
                default:
                    if (c < 0x20) {
                        char buf[8];
                        std::snprintf(buf, sizeof buf, "\\u%04x", c);
                        out += buf;
                    } else {
                        out += char(c);
                    }
            }
        }
        out += '"';
    }
    void write(std::string &out, int indent, int depth) const {
        auto nl = [&](int d) {
            if (indent < 0) return;
            out += '\n';
            out.append(size_t(indent * d), ' ');
        };
        switch (type_) {
            case Type::Null: out += "null"; break;
            case Type::Bool: out += b_ ? "true" : "false"; break;
            case Type::Number: {
                if (std::isfinite(n_) && n_ == std::floor(n_) && std::fabs(n_) < 1e15) {
                    out += std::to_string((long long)n_);
                } else {
                    char buf[32];
                    std::snprintf(buf, sizeof buf, "%.9g", n_);
                    out += buf;
                }
                break;
            }
            case Type::String: escape(out, s_); break;
            case Type::Array:
                out += '[';
                for (size_t k = 0; k < a_.size(); ++k) {
                    if (k) out += ',';
                    nl(depth + 1);
                    a_[k].write(out, indent, depth + 1);
                }
                if (!a_.empty()) nl(depth);
                out += ']';
                break;
            case Type::Object:
                out += '{';
                for (size_t k = 0; k < keys_.size(); ++k) {
                    if (k) out += ',';
                    nl(depth + 1);
                    escape(out, keys_[k]);
                    out += indent < 0 ? ":" : ": ";
                    vals_[k].write(out, indent, depth + 1);
                }
                if (!keys_.empty()) nl(depth);
                out += '}';
                break;
        }
    }

    Type type_{Type::Null};
    bool b_{false};
    double n_{0};
    std::string s_;
    std::vector<Json> a_;
    std::vector<std::string> keys_;
    std::vector<Json> vals_;
    std::map<std::string, size_t> index_;
};

}  // namespace vtp

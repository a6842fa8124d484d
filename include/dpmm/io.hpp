// File formats of the reference's command-line path (header only, C++17): a JSON reader for the params
// file (global_params::init, global_params.cpp:8-70, read with jsoncpp's comment-tolerant reader, so `//`
// and `/* */` comments are accepted), a JSON writer for result.json (main.cpp:59-117) and a reader of
// .npy arrays (utils::load_data_model, utils.cpp:12-57, which the reference reads with cnpy).
#pragma once
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace dpmm {
namespace json {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::map<std::string, Value> obj;

    bool is_null() const { return kind == Null; }
    bool has(const std::string& k) const { return kind == Object && obj.count(k) && !obj.at(k).is_null(); }
    const Value& operator[](const std::string& k) const {
        static const Value none;
        if (kind != Object) return none;
        auto it = obj.find(k);
        return it == obj.end() ? none : it->second;
    }
    double as_double() const {
        if (kind == Number) return num;
        if (kind == Bool) return b ? 1.0 : 0.0;
        throw std::runtime_error("json: number expected");
    }
    bool as_bool() const {
        if (kind == Bool) return b;
        if (kind == Number) return num != 0.0;
        throw std::runtime_error("json: boolean expected");
    }
    // jsoncpp's get(key, default) for the scalar fields of the params file
    double get(const std::string& k, double def) const { return has(k) ? (*this)[k].as_double() : def; }
    bool get(const std::string& k, bool def) const { return has(k) ? (*this)[k].as_bool() : def; }
    std::vector<double> numbers() const {
        if (kind != Array) throw std::runtime_error("json: array expected");
        std::vector<double> v;
        for (const auto& e : arr) v.push_back(e.as_double());
        return v;
    }
};

class Parser {
  public:
    explicit Parser(const std::string& text) : s_(text) {}
    Value parse() {
        Value v = value();
        skip();
        if (p_ != s_.size()) fail("trailing characters");
        return v;
    }

  private:
    const std::string& s_;
    size_t p_ = 0;

    [[noreturn]] void fail(const char* what) const {
        throw std::runtime_error(std::string("json: ") + what + " at offset " + std::to_string(p_));
    }
    void skip() {
        for (;;) {
            while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) ++p_;
            if (p_ + 1 < s_.size() && s_[p_] == '/' && s_[p_ + 1] == '/') {
                while (p_ < s_.size() && s_[p_] != '\n') ++p_;
            } else if (p_ + 1 < s_.size() && s_[p_] == '/' && s_[p_ + 1] == '*') {
                const size_t e = s_.find("*/", p_ + 2);
                if (e == std::string::npos) fail("unterminated comment");
                p_ = e + 2;
            } else {
                return;
            }
        }
    }
    bool literal(const char* w) {
        const size_t n = std::strlen(w);
        if (s_.compare(p_, n, w) == 0) { p_ += n; return true; }
        return false;
    }
    std::string string_lit() {
        if (s_[p_] != '"') fail("string expected");
        ++p_;
        std::string out;
        while (p_ < s_.size() && s_[p_] != '"') {
            char c = s_[p_++];
            if (c == '\\') {
                if (p_ >= s_.size()) fail("bad escape");
                const char e = s_[p_++];
                switch (e) {
                    case 'n': c = '\n'; break;
                    case 't': c = '\t'; break;
                    case 'r': c = '\r'; break;
                    case 'b': c = '\b'; break;
                    case 'f': c = '\f'; break;
                    case 'u': {   // code points below 128 only (file paths and keys)
                        if (p_ + 4 > s_.size()) fail("bad \\u escape");
                        c = (char)std::stoi(s_.substr(p_, 4), nullptr, 16);
                        p_ += 4;
                        break;
                    }
                    default: c = e;
                }
            }
            out.push_back(c);
        }
        if (p_ >= s_.size()) fail("unterminated string");
        ++p_;
        return out;
    }
    Value value() {
        skip();
        if (p_ >= s_.size()) fail("value expected");
        Value v;
        const char c = s_[p_];
        if (c == '{') {
            ++p_;
            v.kind = Value::Object;
            skip();
            if (s_[p_] == '}') { ++p_; return v; }
            for (;;) {
                skip();
                std::string k = string_lit();
                skip();
                if (s_[p_] != ':') fail("':' expected");
                ++p_;
                v.obj[k] = value();
                skip();
                if (s_[p_] == ',') { ++p_; continue; }
                if (s_[p_] == '}') { ++p_; return v; }
                fail("',' or '}' expected");
            }
        }
        if (c == '[') {
            ++p_;
            v.kind = Value::Array;
            skip();
            if (s_[p_] == ']') { ++p_; return v; }
            for (;;) {
                v.arr.push_back(value());
                skip();
                if (s_[p_] == ',') { ++p_; continue; }
                if (s_[p_] == ']') { ++p_; return v; }
                fail("',' or ']' expected");
            }
        }
        if (c == '"') {
            v.kind = Value::String;
            v.str = string_lit();
            return v;
        }
        if (literal("true")) { v.kind = Value::Bool; v.b = true; return v; }
        if (literal("false")) { v.kind = Value::Bool; v.b = false; return v; }
        if (literal("null")) return v;
        const char* begin = s_.c_str() + p_;
        char* end = nullptr;
        v.num = std::strtod(begin, &end);
        if (end == begin) fail("unexpected character");
        p_ += (size_t)(end - begin);
        v.kind = Value::Number;
        return v;
    }
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

inline Value parse_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
}

// Numbers as the shortest text that reads back to the same double (17 significant digits at most).
inline std::string number(double v) {
    if (std::isnan(v) || std::isinf(v)) return "null";
    char buf[32];
    for (int prec = 1; prec <= 17; ++prec) {
        std::snprintf(buf, sizeof buf, "%.*g", prec, v);
        if (std::strtod(buf, nullptr) == v) break;
    }
    return buf;
}

inline std::string quote(const std::string& s) {
    std::string o = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') { o.push_back('\\'); o.push_back(c); }
        else if (c == '\n') o += "\\n";
        else if ((unsigned char)c < 0x20) { char b[8]; std::snprintf(b, sizeof b, "\\u%04x", c); o += b; }
        else o.push_back(c);
    }
    return o + "\"";
}

}  // namespace json

// A 2-D array read from a .npy file (format versions 1-3): shape[0] x shape[1] values as double, in C
// order whatever the file's order.
struct NpyArray {
    size_t rows = 0, cols = 0;
    std::string dtype;   // the file's descr, e.g. "<f8"
    std::vector<double> data;
};

inline NpyArray load_npy(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    char magic[8];
    f.read(magic, 8);
    if (!f || std::memcmp(magic, "\x93NUMPY", 6) != 0) throw std::runtime_error(path + ": not a .npy file");
    uint32_t hlen = 0;
    if (magic[6] == 1) {
        unsigned char b[2];
        f.read((char*)b, 2);
        hlen = b[0] | (b[1] << 8);
    } else {
        unsigned char b[4];
        f.read((char*)b, 4);
        hlen = b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24);
    }
    std::string h(hlen, ' ');
    f.read(&h[0], hlen);
    if (!f) throw std::runtime_error(path + ": truncated header");
    auto field = [&](const char* key) {
        const size_t k = h.find(key);
        if (k == std::string::npos) throw std::runtime_error(path + ": header lacks " + key);
        return h.substr(h.find(':', k) + 1);
    };
    std::string descr = field("'descr'");
    descr = descr.substr(descr.find('\'') + 1);
    descr = descr.substr(0, descr.find('\''));
    const std::string fo = field("'fortran_order'");
    const bool fortran = fo.compare(fo.find_first_not_of(' '), 4, "True") == 0;
    std::string shape = field("'shape'");
    shape = shape.substr(shape.find('(') + 1, shape.find(')') - shape.find('(') - 1);
    std::vector<size_t> dims;
    std::stringstream ss(shape);
    std::string tok;
    while (std::getline(ss, tok, ','))
        if (tok.find_first_not_of(" ") != std::string::npos) dims.push_back(std::stoull(tok));
    if (dims.size() == 1) dims.push_back(1);
    if (dims.size() != 2) throw std::runtime_error(path + ": a 2-D array is expected");
    if (descr.size() < 3 || descr[0] == '>') throw std::runtime_error(path + ": big-endian data is not supported");
    const char type = descr[1];
    const int size = std::stoi(descr.substr(2));
    NpyArray a;
    a.rows = dims[0];
    a.cols = dims[1];
    a.dtype = descr;
    const size_t n = a.rows * a.cols;
    std::vector<char> raw(n * (size_t)size);
    f.read(raw.data(), (std::streamsize)raw.size());
    if (!f) throw std::runtime_error(path + ": truncated data");
    a.data.resize(n);
    for (size_t e = 0; e < n; ++e) {
        const char* p = raw.data() + e * size;
        double v;
        if (type == 'f' && size == 8) { double x; std::memcpy(&x, p, 8); v = x; }
        else if (type == 'f' && size == 4) { float x; std::memcpy(&x, p, 4); v = x; }
        else if (type == 'i' && size == 8) { int64_t x; std::memcpy(&x, p, 8); v = (double)x; }
        else if (type == 'i' && size == 4) { int32_t x; std::memcpy(&x, p, 4); v = x; }
        else if (type == 'i' && size == 2) { int16_t x; std::memcpy(&x, p, 2); v = x; }
        else if ((type == 'u' || type == 'b') && size == 1) v = (unsigned char)*p;
        else if (type == 'u' && size == 2) { uint16_t x; std::memcpy(&x, p, 2); v = x; }
        else if (type == 'u' && size == 4) { uint32_t x; std::memcpy(&x, p, 4); v = x; }
        else throw std::runtime_error(path + ": unsupported dtype " + descr);
        // element e of the file is (row, col) = (e / cols, e % cols) in C order, (e % rows, e / rows) in F order
        const size_t r = fortran ? e % a.rows : e / a.cols, c = fortran ? e / a.rows : e % a.cols;
        a.data[r * a.cols + c] = v;
    }
    return a;
}

}  // namespace dpmm

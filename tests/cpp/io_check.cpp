// Prints what include/dpmm/io.hpp reads, for tests/test_io_formats.py:
//   io_check npy <file>    -> "rows cols dtype" then every value (C order), one per line
//   io_check json <file>   -> the parsed value written back as compact JSON
#include <cstdio>
#include <iostream>
#include <string>

#include "dpmm/io.hpp"

static std::string dump(const dpmm::json::Value& v) {
    using V = dpmm::json::Value;
    switch (v.kind) {
        case V::Null: return "null";
        case V::Bool: return v.b ? "true" : "false";
        case V::Number: return dpmm::json::number(v.num);
        case V::String: return dpmm::json::quote(v.str);
        case V::Array: {
            std::string s = "[";
            for (size_t i = 0; i < v.arr.size(); ++i) s += (i ? "," : "") + dump(v.arr[i]);
            return s + "]";
        }
        default: {
            std::string s = "{";
            bool first = true;
            for (const auto& kv : v.obj) {
                s += (first ? "" : ",") + dpmm::json::quote(kv.first) + ":" + dump(kv.second);
                first = false;
            }
            return s + "}";
        }
    }
}

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    try {
        const std::string mode = argv[1];
        if (mode == "npy") {
            const dpmm::NpyArray a = dpmm::load_npy(argv[2]);
            std::printf("%zu %zu %s\n", a.rows, a.cols, a.dtype.c_str());
            for (double x : a.data) std::printf("%s\n", dpmm::json::number(x).c_str());
        } else {
            std::printf("%s\n", dump(dpmm::json::parse_file(argv[2])).c_str());
        }
    } catch (const std::exception& e) {
        std::printf("error: %s\n", e.what());
        return 1;
    }
    return 0;
}

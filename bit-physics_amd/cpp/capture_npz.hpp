// Capture-file plumbing for the C++ clients of libd3q19: SHA-256, an .npz
// writer (a stored zip of .npy members) and a small JSON value.  Host code
// only; no compression and no zip64, so an archive is limited to 65535
// members and 4 GiB (larger ones throw).
#pragma once

#include <cstdint>
#include <filesystem>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace d3q19_capture {

// Lower-case hex SHA-256 (FIPS 180-4) of len bytes.
std::string sha256_hex(const void* data, size_t len);
std::string sha256_file_hex(const std::filesystem::path& p);

// One .npy member: little-endian array of float64 ("<f8"), int64 ("<i8") or
// a unicode string ("<U{n}", UTF-32LE), C order.
struct NpyArray {
    std::string descr;
    std::vector<int64_t> shape;
    std::vector<unsigned char> bytes;
};
NpyArray npy_f64(const std::vector<double>& v, std::vector<int64_t> shape);
NpyArray npy_f64_scalar(double v);
NpyArray npy_i64_scalar(int64_t v);
NpyArray npy_str_scalar(const std::string& ascii);

// Writes members in the given order with a fixed timestamp (1980-01-01), so
// equal content gives equal bytes.
void write_npz(const std::filesystem::path& p, const std::vector<std::pair<std::string, NpyArray>>& members);

// JSON values for the manifest: objects keep their keys sorted.
struct Json {
    enum Kind { Null, Bool, Int, Num, Str, Arr, Obj } kind = Null;
    bool b = false;
    int64_t i = 0;
    double d = 0.0;
    std::string s;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;

    static Json boolean(bool v) { Json j; j.kind = Bool; j.b = v; return j; }
    static Json integer(int64_t v) { Json j; j.kind = Int; j.i = v; return j; }
    static Json number(double v) { Json j; j.kind = Num; j.d = v; return j; }
    static Json string(std::string v) { Json j; j.kind = Str; j.s = std::move(v); return j; }
    static Json array() { Json j; j.kind = Arr; return j; }
    static Json object() { Json j; j.kind = Obj; return j; }
    Json& operator[](const std::string& k) { kind = Obj; return obj[k]; }
    void push(Json v) { kind = Arr; arr.push_back(std::move(v)); }
    // Two-space indentation, like Python's json.dump(indent=2, sort_keys=True).
    std::string dump(int indent = 0) const;
};

}  // namespace d3q19_capture

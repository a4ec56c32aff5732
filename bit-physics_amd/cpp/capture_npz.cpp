// SHA-256, .npz and JSON writers for the C++ capture path (see capture_npz.hpp).
#include "capture_npz.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>

namespace d3q19_capture {

namespace {

// ---- SHA-256, FIPS 180-4 section 6.2 ---------------------------------------
// Round constants: the first 32 bits of the fractional parts of the cube
// roots of the first 64 primes (FIPS 180-4 section 4.2.2).
const uint32_t kRound[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2,
};

inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

struct Sha256 {
    // Initial hash value: fractional parts of the square roots of the first 8 primes.
    uint32_t h[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                     0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    unsigned char buf[64];
    size_t fill = 0;
    uint64_t total = 0;

    void block(const unsigned char* p) {
        uint32_t w[64];
        for (int t = 0; t < 16; ++t)
            w[t] = (uint32_t)p[4 * t] << 24 | (uint32_t)p[4 * t + 1] << 16 | (uint32_t)p[4 * t + 2] << 8 | p[4 * t + 3];
        for (int t = 16; t < 64; ++t) {
            const uint32_t s0 = rotr(w[t - 15], 7) ^ rotr(w[t - 15], 18) ^ (w[t - 15] >> 3);
            const uint32_t s1 = rotr(w[t - 2], 17) ^ rotr(w[t - 2], 19) ^ (w[t - 2] >> 10);
            w[t] = w[t - 16] + s0 + w[t - 7] + s1;
        }
        uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], k = h[7];
        for (int t = 0; t < 64; ++t) {
            const uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
            const uint32_t ch = (e & f) ^ (~e & g);
            const uint32_t t1 = k + S1 + ch + kRound[t] + w[t];
            const uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
            const uint32_t maj = (a & b) ^ (a & c) ^ (b & c);
            const uint32_t t2 = S0 + maj;
            k = g;
            g = f;
            f = e;
            e = d + t1;
            d = c;
            c = b;
            b = a;
            a = t1 + t2;
        }
        h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += k;
    }
    void update(const void* data, size_t len) {
        const unsigned char* p = (const unsigned char*)data;
        total += len;
        while (len) {
            const size_t take = std::min(len, (size_t)64 - fill);
            std::memcpy(buf + fill, p, take);
            fill += take;
            p += take;
            len -= take;
            if (fill == 64) {
                block(buf);
                fill = 0;
            }
        }
    }
    std::string hex() {
        const uint64_t bits = total * 8;
        const unsigned char one = 0x80, zero = 0;
        update(&one, 1);
        while (fill != 56) update(&zero, 1);
        unsigned char len[8];
        for (int i = 0; i < 8; ++i) len[i] = (unsigned char)(bits >> (56 - 8 * i));
        update(len, 8);
        char out[65];
        for (int i = 0; i < 8; ++i) std::snprintf(out + 8 * i, 9, "%08x", h[i]);
        return std::string(out, 64);
    }
};

// ---- CRC-32 (ISO-HDLC, reflected polynomial 0xEDB88320) used by zip --------
struct CrcTable {
    uint32_t t[256];
    CrcTable() {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            t[i] = c;
        }
    }
};

uint32_t crc32(const unsigned char* p, size_t n) {
    static const CrcTable tab;  // initialised once, thread-safe
    const uint32_t* table = tab.t;
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

void put16(std::vector<unsigned char>& o, uint32_t v) {
    o.push_back((unsigned char)v);
    o.push_back((unsigned char)(v >> 8));
}
void put32(std::vector<unsigned char>& o, uint32_t v) {
    put16(o, v & 0xFFFF);
    put16(o, v >> 16);
}

// The .npy v1.0 encoding: magic, version, header length, a Python-literal
// dict padded with spaces to a 64-byte boundary and ended by '\n', then data.
std::vector<unsigned char> npy_bytes(const NpyArray& a) {
    std::string shape = "(";
    for (size_t k = 0; k < a.shape.size(); ++k) {
        shape += std::to_string(a.shape[k]);
        if (a.shape.size() == 1 || k + 1 < a.shape.size()) shape += ",";
        if (k + 1 < a.shape.size()) shape += " ";
    }
    shape += ")";
    std::string hdr = "{'descr': '" + a.descr + "', 'fortran_order': False, 'shape': " + shape + ", }";
    const size_t pre = 10;  // magic (6) + version (2) + header length (2)
    size_t total = pre + hdr.size() + 1;
    const size_t padded = (total + 63) / 64 * 64;
    hdr.append(padded - total, ' ');
    hdr.push_back('\n');
    std::vector<unsigned char> out;
    out.reserve(pre + hdr.size() + a.bytes.size());
    const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    out.insert(out.end(), magic, magic + 8);
    put16(out, (uint32_t)hdr.size());
    out.insert(out.end(), hdr.begin(), hdr.end());
    out.insert(out.end(), a.bytes.begin(), a.bytes.end());
    return out;
}

template <class T>
void append_le(std::vector<unsigned char>& o, T v) {
    unsigned char b[sizeof(T)];
    std::memcpy(b, &v, sizeof(T));  // x86-64 and gfx hosts are little-endian
    o.insert(o.end(), b, b + sizeof(T));
}

std::string json_escape(const std::string& s) {
    std::string o = "\"";
    for (unsigned char c : s) {
        switch (c) {
            case '"': o += "\\\""; break;
            case '\\': o += "\\\\"; break;
            case '\n': o += "\\n"; break;
            case '\t': o += "\\t"; break;
            case '\r': o += "\\r"; break;
            default:
                if (c < 0x20) {
                    char u[8];
                    std::snprintf(u, sizeof u, "\\u%04x", c);
                    o += u;
                } else {
                    o.push_back((char)c);
                }
        }
    }
    return o + "\"";
}

}  // namespace

std::string sha256_hex(const void* data, size_t len) {
    Sha256 s;
    s.update(data, len);
    return s.hex();
}

std::string sha256_file_hex(const std::filesystem::path& p) {
    std::ifstream in(p, std::ios::binary);
    if (!in) throw std::runtime_error("cannot read " + p.string());
    Sha256 s;
    std::vector<char> chunk(1 << 20);
    while (in) {
        in.read(chunk.data(), (std::streamsize)chunk.size());
        s.update(chunk.data(), (size_t)in.gcount());
    }
    return s.hex();
}

NpyArray npy_f64(const std::vector<double>& v, std::vector<int64_t> shape) {
    NpyArray a;
    a.descr = "<f8";
    a.shape = std::move(shape);
    a.bytes.resize(v.size() * sizeof(double));
    if (!v.empty()) std::memcpy(a.bytes.data(), v.data(), a.bytes.size());
    return a;
}

NpyArray npy_f64_scalar(double v) {
    NpyArray a;
    a.descr = "<f8";
    append_le(a.bytes, v);
    return a;
}

NpyArray npy_i64_scalar(int64_t v) {
    NpyArray a;
    a.descr = "<i8";
    append_le(a.bytes, v);
    return a;
}

NpyArray npy_str_scalar(const std::string& ascii) {
    NpyArray a;
    const size_t n = std::max<size_t>(ascii.size(), 1);
    a.descr = "<U" + std::to_string(n);
    for (size_t k = 0; k < n; ++k) append_le<uint32_t>(a.bytes, k < ascii.size() ? (unsigned char)ascii[k] : 0u);
    return a;
}

void write_npz(const std::filesystem::path& p, const std::vector<std::pair<std::string, NpyArray>>& members) {
    if (members.size() > 0xFFFF) throw std::runtime_error("capture has more than 65535 arrays (no zip64 here)");
    std::ofstream out(p, std::ios::binary | std::ios::trunc);
    if (!out) throw std::runtime_error("cannot write " + p.string());
    const uint32_t kDate1980 = (0u << 9) | (1u << 5) | 1u;  // DOS date: 1980-01-01
    std::vector<unsigned char> central;
    uint64_t offset = 0;
    for (const auto& [key, arr] : members) {
        const std::string name = key + ".npy";
        const std::vector<unsigned char> data = npy_bytes(arr);
        if (offset + data.size() + 30 + name.size() > 0xFFFFFFFFull)
            throw std::runtime_error("capture payload exceeds 4 GiB (no zip64 here)");
        const uint32_t crc = crc32(data.data(), data.size());
        std::vector<unsigned char> local;
        put32(local, 0x04034b50);  // local file header
        put16(local, 20);          // version needed: 2.0
        put16(local, 0);           // flags
        put16(local, 0);           // stored
        put16(local, 0);           // time 00:00:00
        put16(local, kDate1980);
        put32(local, crc);
        put32(local, (uint32_t)data.size());
        put32(local, (uint32_t)data.size());
        put16(local, (uint32_t)name.size());
        put16(local, 0);           // no extra field
        local.insert(local.end(), name.begin(), name.end());
        out.write((const char*)local.data(), (std::streamsize)local.size());
        out.write((const char*)data.data(), (std::streamsize)data.size());

        put32(central, 0x02014b50);  // central directory header
        put16(central, 20);          // made by: 2.0
        put16(central, 20);
        put16(central, 0);
        put16(central, 0);
        put16(central, 0);
        put16(central, kDate1980);
        put32(central, crc);
        put32(central, (uint32_t)data.size());
        put32(central, (uint32_t)data.size());
        put16(central, (uint32_t)name.size());
        put16(central, 0);           // extra
        put16(central, 0);           // comment
        put16(central, 0);           // disk
        put16(central, 0);           // internal attributes
        put32(central, 0);           // external attributes
        put32(central, (uint32_t)offset);
        central.insert(central.end(), name.begin(), name.end());
        offset += local.size() + data.size();
    }
    if (offset + central.size() > 0xFFFFFFFFull) throw std::runtime_error("capture payload exceeds 4 GiB");
    std::vector<unsigned char> end;
    put32(end, 0x06054b50);  // end of central directory
    put16(end, 0);
    put16(end, 0);
    put16(end, (uint32_t)members.size());
    put16(end, (uint32_t)members.size());
    put32(end, (uint32_t)central.size());
    put32(end, (uint32_t)offset);
    put16(end, 0);
    out.write((const char*)central.data(), (std::streamsize)central.size());
    out.write((const char*)end.data(), (std::streamsize)end.size());
    if (!out) throw std::runtime_error("write failed: " + p.string());
}

std::string Json::dump(int indent) const {
    const std::string pad(indent + 2, ' '), close(indent, ' ');
    switch (kind) {
        case Null: return "null";
        case Bool: return b ? "true" : "false";
        case Int: return std::to_string(i);
        case Num: {
            if (!std::isfinite(d)) throw std::runtime_error("JSON cannot hold a non-finite number");
            // the shortest of 15, 16, 17 significant digits that reads back exactly
            char t[32];
            for (int prec = 15; prec <= 17; ++prec) {
                std::snprintf(t, sizeof t, "%.*g", prec, d);
                if (std::strtod(t, nullptr) == d) break;
            }
            std::string s = t;
            if (s.find_first_of(".eE") == std::string::npos) s += ".0";
            return s;
        }
        case Str: return json_escape(s);
        case Arr: {
            if (arr.empty()) return "[]";
            std::string o = "[\n";
            for (size_t k = 0; k < arr.size(); ++k)
                o += pad + arr[k].dump(indent + 2) + (k + 1 < arr.size() ? ",\n" : "\n");
            return o + close + "]";
        }
        case Obj: {
            if (obj.empty()) return "{}";
            std::string o = "{\n";
            size_t k = 0;
            for (const auto& [key, v] : obj)
                o += pad + json_escape(key) + ": " + v.dump(indent + 2) + (++k < obj.size() ? ",\n" : "\n");
            return o + close + "}";
        }
    }
    return "null";
}

}  // namespace d3q19_capture

// Test infrastructure only: minimal .npy (v1.0) writer shared by the
// reference drivers (little-endian, C order).
#pragma once

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace npy {

inline void write_npy(const std::string& path, const char* descr, size_t elsize,
               const std::vector<size_t>& shape, const void* data) {
    std::string shp = "(";
    size_t n = 1;
    for (size_t i = 0; i < shape.size(); ++i) {
        shp += std::to_string(shape[i]);
        shp += (shape.size() == 1) ? "," : (i + 1 < shape.size() ? ", " : "");
        n *= shape[i];
    }
    shp += ")";
    std::string hdr = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': " + shp + ", }";
    size_t total = 10 + hdr.size() + 1;
    size_t pad = (64 - total % 64) % 64;
    hdr += std::string(pad, ' ');
    hdr += '\n';
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    std::fwrite(magic, 1, 8, f);
    uint16_t hl = (uint16_t)hdr.size();
    std::fwrite(&hl, 2, 1, f);
    std::fwrite(hdr.data(), 1, hdr.size(), f);
    if (n) std::fwrite(data, elsize, n, f);
    std::fclose(f);
}

inline void npy_f64(const std::string& p, const std::vector<size_t>& s, const double* d) { write_npy(p, "<f8", 8, s, d); }
inline void npy_u32(const std::string& p, const std::vector<size_t>& s, const uint32_t* d) { write_npy(p, "<u4", 4, s, d); }

}  // namespace npy

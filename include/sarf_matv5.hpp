// sarf_matv5.hpp -- reader for MATLAB Level 5 MAT-files (the -v5/-v6/-v7
// formats: uncompressed and zlib-compressed elements, little-endian), for
// the phase-history records the reference reads through matio
// (src/core/uncc_sar_matio.cpp:23-52, include/uncc_sar_matio.hpp:132-285).
//
// Numeric arrays, structs (nested, 1x1 or arrays) and cells are read; char,
// sparse, logical-as-numeric and object arrays are skipped.  Struct fields
// are flattened into dotted names ("sph_MATData.Data.SampleData"); element
// e > 0 of a struct array gets the name "root(e).field" (e 0-based,
// column-major).  Values are converted to double; complex arrays keep an
// imaginary part.  Implemented in libsarf (sarf_matv5.cpp), no matio.
#ifndef SARF_MATV5_HPP
#define SARF_MATV5_HPP

#include <cstddef>
#include <map>
#include <string>
#include <vector>

namespace sarf {
namespace matv5 {

struct Array {
    int mx_class = 0;          // MATLAB class id (6 double, 7 single, 8.. integers)
    std::vector<int> dims;     // MATLAB dimensions (column-major, dims[0] fastest)
    std::vector<double> re;    // prod(dims) values
    std::vector<double> im;    // empty unless complex
    bool complex() const { return !im.empty(); }
    size_t numel() const {
        size_t n = 1;
        for (int d : dims) n *= (size_t)d;
        return dims.empty() ? 0 : n;
    }
};

// Reads every numeric variable and struct field of `path` into `out`.
// Returns 0 on success; on failure returns -1 and sets `err`.
int read_file(const std::string &path, std::map<std::string, Array> &out, std::string &err);

}  // namespace matv5
}  // namespace sarf

#endif  // SARF_MATV5_HPP

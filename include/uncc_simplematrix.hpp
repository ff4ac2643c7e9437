// uncc_simplematrix.hpp -- the column-major container of the SAR focusing
// C++ interface (shape + flat data), with the import helpers callers use to
// wrap raw arrays.  Same names and semantics as the reference
// (include/uncc_simplematrix.hpp:37-293): shape[0] varies fastest, complex
// sources are interleaved re/im.
#ifndef SARF_UNCC_SIMPLEMATRIX_HPP
#define SARF_UNCC_SIMPLEMATRIX_HPP

#include <cstdlib>
#include <functional>
#include <iostream>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

// Half-open index range [start, end) with a stride.
class Range {
   public:
    int start, end, stride;
    Range(int end_) : Range(0, end_, 1) {}
    Range(int start_, int end_) : Range(start_, end_, 1) {}
    Range(int start_, int end_, int stride_) : start(start_), end(end_), stride(stride_) {
        if (start < 0) start = 0;
        if (end < 0) end = 0;
        if (stride == 0) stride = (start <= end) ? 1 : -1;
        if ((start < end && stride < 0) || (start > end && stride > 0)) stride = -stride;
    }
    std::vector<int> values() const {
        std::vector<int> v;
        if (stride > 0)
            for (int i = start; i < end; i += stride) v.push_back(i);
        else
            for (int i = start; i > end; i += stride) v.push_back(i);
        return v;
    }
};

template <typename T>
struct SimpleMatrix {
    typedef std::shared_ptr<SimpleMatrix> Ptr;
    std::vector<int> shape;
    std::vector<T> data;

    int nelem() const {
        return shape.empty() ? 0
                             : std::accumulate(shape.begin(), shape.end(), 1, std::multiplies<int>());
    }
    bool isEmpty() const { return shape.empty(); }

    // column-major linear index of a multi-index
    int sub2ind(const std::vector<int> &idx) const {
        int r = 0;
        for (int d = (int)idx.size() - 1; d >= 0; d--) r = r * (d < (int)shape.size() ? shape[d] : 1) + idx[d];
        return r;
    }
    T at(const std::vector<int> &idx) const { return data[sub2ind(idx)]; }

    // Sub-matrix over one Range per dimension; dimensions of length 1 drop.
    SimpleMatrix getRange(const std::vector<Range> &sel) const {
        SimpleMatrix out;
        std::vector<std::vector<int>> idx;
        for (const Range &r : sel) {
            idx.push_back(r.values());
            if (idx.back().size() > 1) out.shape.push_back((int)idx.back().size());
        }
        std::vector<int> cur(sel.size(), 0);
        collect(idx, cur, (int)sel.size() - 1, out.data);
        return out;
    }

    template <typename D>
    std::vector<D> toVector() const {
        return std::vector<D>(data.begin(), data.end());
    }

   private:
    void collect(const std::vector<std::vector<int>> &idx, std::vector<int> &cur, int dim,
                 std::vector<T> &out) const {
        for (int v : idx[dim]) {
            cur[dim] = v;
            if (dim > 0)
                collect(idx, cur, dim - 1, out);
            else
                out.push_back(at(cur));
        }
    }
};

template <typename T>
inline std::ostream &operator<<(std::ostream &os, const SimpleMatrix<T> &m) {
    os << "[";
    for (size_t i = 0; i < m.shape.size(); i++) os << (i ? "," : "") << m.shape[i];
    os << "] = {";
    for (size_t i = 0; i < m.data.size() && i < 10; i++) os << m.data[i] << ", ";
    return os << (m.data.size() > 10 ? " ..." : "") << " }";
}

// Copy a real array of the given dims into sMat (dims appended to shape).
template <typename S, typename D>
int import_MatrixReal(S *data, int *dims, int ndims, SimpleMatrix<D> &sMat) {
    size_t total = 1;
    for (int d = 0; d < ndims; d++) {
        sMat.shape.push_back(dims[d]);
        total *= (size_t)dims[d];
    }
    sMat.data.insert(sMat.data.end(), data, data + total);
    return EXIT_SUCCESS;
}

template <typename S, typename D>
int import_Vector(S *data, int *dims, int ndims, SimpleMatrix<D> &sMat) {
    return import_MatrixReal<S, D>(data, dims, ndims, sMat);
}

// Copy an interleaved (re, im, re, im, ...) array into complex elements.
template <typename S, typename DC>
int import_MatrixComplex(S *data, int *dims, int ndims, SimpleMatrix<DC> &sMat) {
    size_t total = 1;
    for (int d = 0; d < ndims; d++) {
        sMat.shape.push_back(dims[d]);
        total *= (size_t)dims[d];
    }
    sMat.data.reserve(sMat.data.size() + total);
    for (size_t i = 0; i < total; i++) sMat.data.push_back(DC(data[2 * i], data[2 * i + 1]));
    return EXIT_SUCCESS;
}

#endif  // SARF_UNCC_SIMPLEMATRIX_HPP

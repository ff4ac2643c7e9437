// sarf_matv5.cpp -- MATLAB Level 5 MAT-file reader (see sarf_matv5.hpp).
//
// Layout followed (MathWorks "MAT-File Format", Level 5): a 128-byte header
// ending in version 0x0100 and the endian indicator "IM" (little-endian);
// then data elements, each an 8-byte tag {u32 type, u32 bytes} followed by
// the data padded to 8 bytes, or a "small" element whose tag packs
// {u16 type, u16 bytes} and up to 4 data bytes.  miCOMPRESSED (15) holds a
// zlib stream that inflates to one element; miMATRIX (14) holds sub-
// elements: array flags, dimensions, name, then the class's payload.
#include <zlib.h>

#include <cstdint>
#include <cstring>
#include <fstream>
#include <iterator>
#include <sstream>

#include "sarf_matv5.hpp"

namespace sarf {
namespace matv5 {
namespace {

enum : uint32_t {
    miINT8 = 1, miUINT8 = 2, miINT16 = 3, miUINT16 = 4, miINT32 = 5, miUINT32 = 6,
    miSINGLE = 7, miDOUBLE = 9, miINT64 = 12, miUINT64 = 13, miMATRIX = 14,
    miCOMPRESSED = 15, miUTF8 = 16, miUTF16 = 17, miUTF32 = 18
};
enum : int { mxCELL = 1, mxSTRUCT = 2, mxOBJECT = 3, mxCHAR = 4, mxSPARSE = 5 };

struct Error {
    std::string msg;
};

[[noreturn]] void bad(const std::string &m) { throw Error{m}; }

struct Elem {
    uint32_t type = 0;
    const uint8_t *data = nullptr;
    size_t bytes = 0;
};

class Reader {
   public:
    Reader(const uint8_t *p, size_t n) : p_(p), n_(n) {}
    bool done() const { return pos_ >= n_; }
    // Next element; `storage` keeps inflated bytes of compressed elements.
    Elem next(std::vector<uint8_t> &storage) {
        if (pos_ + 8 > n_) bad("truncated element tag");
        uint32_t t0, t1;
        std::memcpy(&t0, p_ + pos_, 4);
        std::memcpy(&t1, p_ + pos_ + 4, 4);
        Elem e;
        if (t0 >> 16) {  // small element
            e.type = t0 & 0xffff;
            e.bytes = t0 >> 16;
            if (e.bytes > 4) bad("bad small element");
            e.data = p_ + pos_ + 4;
            pos_ += 8;
            return e;
        }
        e.type = t0;
        e.bytes = t1;
        if (pos_ + 8 + e.bytes > n_) bad("truncated element data");
        e.data = p_ + pos_ + 8;
        if (e.type == miCOMPRESSED) {
            pos_ += 8 + e.bytes;  // compressed elements are not padded
            storage = inflate_all(e.data, e.bytes);
            Reader inner(storage.data(), storage.size());
            std::vector<uint8_t> unused;
            Elem in = inner.next(unused);
            if (in.type == miCOMPRESSED) bad("nested compression");
            return in;
        }
        pos_ += 8 + ((e.bytes + 7) & ~size_t(7));
        return e;
    }

   private:
    static std::vector<uint8_t> inflate_all(const uint8_t *src, size_t n) {
        z_stream zs;
        std::memset(&zs, 0, sizeof zs);
        if (inflateInit(&zs) != Z_OK) bad("zlib init failed");
        std::vector<uint8_t> out(n * 4 + 1024);
        zs.next_in = const_cast<Bytef *>(src);
        zs.avail_in = (uInt)n;
        int rc;
        for (;;) {
            if (zs.total_out == out.size()) out.resize(out.size() * 2);
            zs.next_out = out.data() + zs.total_out;
            zs.avail_out = (uInt)(out.size() - zs.total_out);
            rc = inflate(&zs, Z_NO_FLUSH);
            if (rc == Z_STREAM_END) break;
            if (rc == Z_OK) continue;
            if (rc == Z_BUF_ERROR && zs.avail_out == 0) continue;  // grow the output
            break;  // truncated input or a corrupt stream
        }
        const size_t got = zs.total_out;
        inflateEnd(&zs);
        if (rc != Z_STREAM_END) bad("bad or truncated zlib stream");
        out.resize(got);
        return out;
    }

    const uint8_t *p_;
    size_t n_;
    size_t pos_ = 0;
};

size_t type_size(uint32_t t) {
    switch (t) {
        case miINT8: case miUINT8: case miUTF8: return 1;
        case miINT16: case miUINT16: case miUTF16: return 2;
        case miINT32: case miUINT32: case miSINGLE: case miUTF32: return 4;
        case miDOUBLE: case miINT64: case miUINT64: return 8;
        default: return 0;
    }
}

template <class T>
T load(const uint8_t *p) {
    T v;
    std::memcpy(&v, p, sizeof v);
    return v;
}

std::vector<double> numeric(const Elem &e) {
    const size_t sz = type_size(e.type);
    if (sz == 0 || e.type == miUTF8 || e.type == miUTF16 || e.type == miUTF32)
        bad("non-numeric data element");
    const size_t n = e.bytes / sz;
    std::vector<double> v(n);
    for (size_t i = 0; i < n; i++) {
        const uint8_t *q = e.data + i * sz;
        switch (e.type) {
            case miINT8: v[i] = load<int8_t>(q); break;
            case miUINT8: v[i] = load<uint8_t>(q); break;
            case miINT16: v[i] = load<int16_t>(q); break;
            case miUINT16: v[i] = load<uint16_t>(q); break;
            case miINT32: v[i] = load<int32_t>(q); break;
            case miUINT32: v[i] = load<uint32_t>(q); break;
            case miSINGLE: v[i] = load<float>(q); break;
            case miDOUBLE: v[i] = load<double>(q); break;
            case miINT64: v[i] = (double)load<int64_t>(q); break;
            case miUINT64: v[i] = (double)load<uint64_t>(q); break;
        }
    }
    return v;
}

std::string text(const Elem &e) {
    if (e.type != miINT8 && e.type != miUINT8 && e.type != miUTF8) bad("name is not text");
    std::string s((const char *)e.data, e.bytes);
    const size_t z = s.find('\0');
    return z == std::string::npos ? s : s.substr(0, z);
}

void parse_matrix(const Elem &m, const std::string &prefix, bool named,
                  std::map<std::string, Array> &out);

// Parses the sub-elements of one miMATRIX element.  `prefix` is the dotted
// path of a struct field (its own name element is empty) or "" at top level.
void parse_matrix(const Elem &m, const std::string &prefix, bool named,
                  std::map<std::string, Array> &out) {
    if (m.bytes == 0) return;  // empty array
    Reader r(m.data, m.bytes);
    std::vector<uint8_t> st;
    const Elem flags = r.next(st);
    if (flags.type != miUINT32 || flags.bytes < 8) bad("bad array flags");
    const uint32_t f0 = load<uint32_t>(flags.data);
    const int cls = (int)(f0 & 0xff);
    const bool is_complex = (f0 & 0x800) != 0;
    const Elem de = r.next(st);
    if (de.type != miINT32) bad("bad dimensions");
    std::vector<int> dims;
    for (size_t i = 0; i < de.bytes / 4; i++) dims.push_back(load<int32_t>(de.data + 4 * i));
    const Elem ne = r.next(st);
    std::string name = text(ne);
    std::string full = named ? (prefix.empty() ? name : prefix + "." + name) : prefix;
    size_t numel = 1;
    for (int d : dims) numel *= (size_t)(d < 0 ? 0 : d);

    if (cls == mxSTRUCT) {
        const Elem fl = r.next(st);
        if (fl.type != miINT32 || fl.bytes < 4) bad("bad field name length");
        const int flen = load<int32_t>(fl.data);
        const Elem fn = r.next(st);
        if (flen <= 0) return;
        const int nf = (int)(fn.bytes / (size_t)flen);
        std::vector<std::string> fields(nf);
        for (int i = 0; i < nf; i++) {
            std::string s((const char *)fn.data + (size_t)i * flen, (size_t)flen);
            const size_t z = s.find('\0');
            fields[i] = z == std::string::npos ? s : s.substr(0, z);
        }
        for (size_t e = 0; e < numel; e++) {
            const std::string base = e == 0 ? full : full + "(" + std::to_string(e) + ")";
            for (int i = 0; i < nf; i++) {
                std::vector<uint8_t> st2;
                const Elem sub = r.next(st2);
                if (sub.type != miMATRIX) bad("struct field is not a matrix");
                parse_matrix(sub, base + "." + fields[i], false, out);
            }
        }
        return;
    }
    if (cls == mxCELL) {
        for (size_t e = 0; e < numel; e++) {
            std::vector<uint8_t> st2;
            const Elem sub = r.next(st2);
            if (sub.type != miMATRIX) bad("cell element is not a matrix");
            parse_matrix(sub, full + "{" + std::to_string(e) + "}", false, out);
        }
        return;
    }
    if (cls == mxCHAR || cls == mxSPARSE || cls == mxOBJECT || cls < 6 || cls > 15) return;
    Array a;
    a.mx_class = cls;
    a.dims = dims;
    const Elem re = r.next(st);
    a.re = numeric(re);
    if (is_complex) {
        std::vector<uint8_t> st2;
        const Elem im = r.next(st2);
        a.im = numeric(im);
        if (a.im.size() != a.re.size()) bad("real and imaginary parts differ in size");
    }
    if (a.re.size() != numel) bad("data size does not match dimensions for " + full);
    out[full] = std::move(a);
}

}  // namespace

int read_file(const std::string &path, std::map<std::string, Array> &out, std::string &err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open " + path;
        return -1;
    }
    std::vector<uint8_t> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    try {
        if (buf.size() < 128) bad("file shorter than the 128-byte header");
        if (buf[126] == 'M' && buf[127] == 'I') bad("big-endian MAT-files are not supported");
        if (buf[126] != 'I' || buf[127] != 'M') bad("not a Level 5 MAT-file (v7.3/HDF5?)");
        Reader r(buf.data() + 128, buf.size() - 128);
        while (!r.done()) {
            std::vector<uint8_t> st;
            const Elem e = r.next(st);
            if (e.type == miMATRIX) parse_matrix(e, "", true, out);
        }
    } catch (const Error &e) {
        err = path + ": " + e.msg;
        return -1;
    }
    return 0;
}

}  // namespace matv5
}  // namespace sarf

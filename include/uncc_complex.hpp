// uncc_complex.hpp -- complex type of the SAR focusing C++ interface.
//
// Same name, layout (two consecutive T: real then imaginary) and member
// names as the reference's unccComplex (include/uncc_complex.hpp:32-214),
// so caller code that reads _M_real/_M_imag or uses polar/conj/norm/abs
// compiles unchanged.  Host-only arithmetic; the GPU path never uses it.
#ifndef SARF_UNCC_COMPLEX_HPP
#define SARF_UNCC_COMPLEX_HPP

#include <algorithm>
#include <cmath>
#include <iostream>

template <typename T>
class unccComplex {
   public:
    T _M_real;
    T _M_imag;

    constexpr unccComplex(const T &re = T(), const T &im = T()) : _M_real(re), _M_imag(im) {}
    template <typename U>
    constexpr unccComplex(const unccComplex<U> &c) : _M_real(T(c.real())), _M_imag(T(c.imag())) {}

    constexpr T real() const { return _M_real; }
    constexpr T imag() const { return _M_imag; }

    static unccComplex conj(const unccComplex &x) { return unccComplex(x._M_real, -x._M_imag); }

    // r e^{j phi}
    template <typename R, typename P>
    static unccComplex polar(const R &r, const P &phi) {
        return unccComplex(T(r * std::cos(phi)), T(r * std::sin(phi)));
    }

    // squared magnitude
    static T norm(const unccComplex &z) { return z._M_real * z._M_real + z._M_imag * z._M_imag; }

    // magnitude, scaled to avoid overflow
    static T abs(const unccComplex &z) {
        const T s = std::max(std::abs(z._M_real), std::abs(z._M_imag));
        if (s == T(0)) return s;
        const T a = z._M_real / s, b = z._M_imag / s;
        return s * std::sqrt(a * a + b * b);
    }

    unccComplex &operator+=(const unccComplex &o) {
        _M_real += o._M_real;
        _M_imag += o._M_imag;
        return *this;
    }
    unccComplex &operator-=(const unccComplex &o) {
        _M_real -= o._M_real;
        _M_imag -= o._M_imag;
        return *this;
    }
    unccComplex &operator*=(const unccComplex &o) {
        const T re = _M_real * o._M_real - _M_imag * o._M_imag;
        _M_imag = _M_real * o._M_imag + _M_imag * o._M_real;
        _M_real = re;
        return *this;
    }
    unccComplex &operator/=(const unccComplex &o) {
        const T n = norm(o);
        const T re = (_M_real * o._M_real + _M_imag * o._M_imag) / n;
        _M_imag = (_M_imag * o._M_real - _M_real * o._M_imag) / n;
        _M_real = re;
        return *this;
    }
    template <typename S>
    unccComplex &operator*=(const S &s) {
        _M_real *= s;
        _M_imag *= s;
        return *this;
    }
    template <typename S>
    unccComplex &operator/=(const S &s) {
        _M_real /= s;
        _M_imag /= s;
        return *this;
    }
};

template <typename T>
inline unccComplex<T> operator+(unccComplex<T> a, const unccComplex<T> &b) { return a += b; }
template <typename T>
inline unccComplex<T> operator-(unccComplex<T> a, const unccComplex<T> &b) { return a -= b; }
template <typename T>
inline unccComplex<T> operator*(unccComplex<T> a, const unccComplex<T> &b) { return a *= b; }
template <typename T>
inline unccComplex<T> operator/(unccComplex<T> a, const unccComplex<T> &b) { return a /= b; }
template <typename T, typename S>
inline unccComplex<T> operator*(unccComplex<T> a, const S &s) { return a *= s; }
template <typename T, typename S>
inline unccComplex<T> operator*(const S &s, unccComplex<T> a) { return a *= s; }
template <typename T, typename S>
inline unccComplex<T> operator/(unccComplex<T> a, const S &s) { return a /= s; }
template <typename T>
inline bool operator==(const unccComplex<T> &a, const unccComplex<T> &b) {
    return a._M_real == b._M_real && a._M_imag == b._M_imag;
}

template <typename T>
inline std::ostream &operator<<(std::ostream &os, const unccComplex<T> &c) {
    return os << "(" << c._M_real << "," << c._M_imag << ")";
}

#endif  // SARF_UNCC_COMPLEX_HPP

// complex<T> for host code and HIP device code (user potentials receive complex<float>*).
//
// Same class name, member functions and free functions as the reference's
// modules/complex_numbers.h (re, im, norm, cabs, arg, conj, cexp, csin, ccos, clog, icpow,
// cpow, isnan), so code written against it compiles unchanged.  Layout is two packed T
// (real, imaginary), so a complex<float> array aliases the engine's float2 positions.
// icpow(z, n) is the integer power z^n by binary exponentiation in complex arithmetic
// (the reference goes through |z|^n e^{i n arg z}; both are z^n).
#ifndef LE_COMPLEX_NUMBERS_H
#define LE_COMPLEX_NUMBERS_H

#include <cmath>
#include <complex>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define LE_HD __host__ __device__
#else
#define LE_HD
#endif

template <typename T>
class complex {
    T x_;
    T y_;

public:
    LE_HD complex() {}
    LE_HD complex(const T& x, const T& y) : x_(x), y_(y) {}
    LE_HD complex(const T& x) : x_(x), y_(0) {}

    LE_HD T real() const { return x_; }
    LE_HD T imag() const { return y_; }
    LE_HD T* pointer_to_real() { return &x_; }
    LE_HD T* pointer_to_imag() { return &y_; }
    LE_HD T norm() const { return x_ * x_ + y_ * y_; }
    LE_HD T cabs() const { return sqrt(x_ * x_ + y_ * y_); }
    LE_HD T arg() const { return atan2(y_, x_); }
    LE_HD complex<T> conj() const { return complex<T>(x_, -y_); }

    template <typename T2>
    LE_HD operator complex<T2>() const { return complex<T2>((T2)x_, (T2)y_); }
    template <typename T2>
    operator std::complex<T2>() const { return std::complex<T2>((T2)x_, (T2)y_); }

    LE_HD complex<T> operator+(const complex<T>& z) const { return complex<T>(x_ + z.x_, y_ + z.y_); }
    LE_HD complex<T> operator-(const complex<T>& z) const { return complex<T>(x_ - z.x_, y_ - z.y_); }
    LE_HD complex<T> operator*(const complex<T>& z) const
    {
        return complex<T>(x_ * z.x_ - y_ * z.y_, x_ * z.y_ + y_ * z.x_);
    }
    LE_HD complex<T> operator/(const complex<T>& z) const
    {
        const T d = z.x_ * z.x_ + z.y_ * z.y_;
        return complex<T>((x_ * z.x_ + y_ * z.y_) / d, (y_ * z.x_ - x_ * z.y_) / d);
    }
    LE_HD complex<T> operator+(const T& a) const { return complex<T>(x_ + a, y_); }
    LE_HD complex<T> operator-(const T& a) const { return complex<T>(x_ - a, y_); }
    LE_HD complex<T> operator*(const T& a) const { return complex<T>(x_ * a, y_ * a); }
    LE_HD complex<T> operator/(const T& a) const { return complex<T>(x_ / a, y_ / a); }
    LE_HD complex<T> operator-() const { return complex<T>(-x_, -y_); }
    LE_HD friend complex<T> operator+(const T& a, const complex<T>& z) { return complex<T>(a + z.x_, z.y_); }
    LE_HD friend complex<T> operator-(const T& a, const complex<T>& z) { return complex<T>(a - z.x_, -z.y_); }
    LE_HD friend complex<T> operator*(const T& a, const complex<T>& z) { return complex<T>(a * z.x_, a * z.y_); }
    LE_HD friend complex<T> operator/(const T& a, const complex<T>& z) { return complex<T>(a) / z; }
    LE_HD void operator+=(const complex<T>& z) { *this = *this + z; }
    LE_HD void operator-=(const complex<T>& z) { *this = *this - z; }
    LE_HD void operator*=(const complex<T>& z) { *this = *this * z; }
    LE_HD void operator/=(const complex<T>& z) { *this = *this / z; }
    LE_HD bool operator==(const complex<T>& z) const { return x_ == z.x_ && y_ == z.y_; }

    LE_HD friend T re(const complex<T>& z) { return z.x_; }
    LE_HD friend T im(const complex<T>& z) { return z.y_; }
    LE_HD friend T norm(const complex<T>& z) { return z.norm(); }
    LE_HD friend T cabs(const complex<T>& z) { return z.cabs(); }
    LE_HD friend T arg(const complex<T>& z) { return z.arg(); }
    LE_HD friend complex<T> conj(const complex<T>& z) { return z.conj(); }
    LE_HD friend complex<T> cexp(const complex<T>& z)
    {
        const T r = exp(z.x_);
        return complex<T>(r * cos(z.y_), r * sin(z.y_));
    }
    LE_HD friend complex<T> csin(const complex<T>& z)
    {
        return complex<T>(sin(z.x_) * cosh(z.y_), cos(z.x_) * sinh(z.y_));
    }
    LE_HD friend complex<T> ccos(const complex<T>& z)
    {
        return complex<T>(cos(z.x_) * cosh(z.y_), -sin(z.x_) * sinh(z.y_));
    }
    LE_HD friend complex<T> clog(const complex<T>& z) { return complex<T>(log(z.cabs()), z.arg()); }
    LE_HD friend complex<T> icpow(const complex<T>& z, int n)
    {
        complex<T> base = z, acc((T)1);
        bool inv = n < 0;
        unsigned e = inv ? (unsigned)(-n) : (unsigned)n;
        while (e) {
            if (e & 1u) acc = acc * base;
            e >>= 1;
            if (e) base = base * base;
        }
        return inv ? complex<T>((T)1) / acc : acc;
    }
    LE_HD friend complex<T> cpow(const complex<T>& z, const T& a)
    {
        const T th = z.arg() * a, r = pow(z.cabs(), a);
        return complex<T>(r * cos(th), r * sin(th));
    }
    LE_HD friend bool isnan(const complex<T>& z) { return std::isnan(z.x_) || std::isnan(z.y_); }
};

#endif

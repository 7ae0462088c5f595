// Boundary containers of the reference's public signatures (modules/tensors.h):
// vector<T> (with the inclusive extract(v, b, e)), matrix<T>, cmatrix<T> (row-major
// complex, adjoint()) and rank3tensor<T> (index i1 + n1 (i2 + n2 i3)).  Host-only,
// backed by std::vector.
#ifndef LE_TENSORS_H
#define LE_TENSORS_H

#include <cstddef>
#include <initializer_list>
#include <vector>

#include "complex_numbers.h"

template <typename T>
class vector {
    std::vector<T> v_;

public:
    vector() {}
    explicit vector(int n) : v_((size_t)(n > 0 ? n : 0)) {}
    vector(int n, const T& x) : v_((size_t)(n > 0 ? n : 0), x) {}
    vector(std::initializer_list<T> l) : v_(l) {}
    vector(const std::vector<T>& v) : v_(v) {}
    T& operator[](const int& i) { return v_[(size_t)i]; }
    const T& operator[](const int& i) const { return v_[(size_t)i]; }
    int size() const { return (int)v_.size(); }
    void push_back(const T& x) { v_.push_back(x); }
    void reserve(int n) { v_.reserve((size_t)n); }
    void resize(int n) { v_.resize((size_t)n); }
    T* data() { return v_.data(); }
    const T* data() const { return v_.data(); }
    template <typename T2>
    operator std::vector<T2>() const { return std::vector<T2>(v_.begin(), v_.end()); }
    // elements b..e, both ends included
    friend vector<T> extract(vector<T>& v, int b, int e)
    {
        vector<T> out;
        for (int i = b; i <= e; ++i) out.push_back(v[i]);
        return out;
    }
};

template <typename T>
class matrix {
    int r_ = 0, c_ = 0;
    std::vector<T> v_;

public:
    matrix() {}
    matrix(int rows, int cols) : r_(rows), c_(cols), v_((size_t)rows * cols, T(0)) {}
    T& operator()(const int& i, const int& j) { return v_[(size_t)i * c_ + j]; }
    T operator()(const int& i, const int& j) const { return v_[(size_t)i * c_ + j]; }
    int rows() const { return r_; }
    int cols() const { return c_; }
    T* data() { return v_.data(); }
    matrix<T> operator+(const matrix<T>& m) const
    {
        matrix<T> u(r_, c_);
        for (size_t q = 0; q < v_.size(); ++q) u.v_[q] = v_[q] + m.v_[q];
        return u;
    }
    matrix<T> operator-(const matrix<T>& m) const
    {
        matrix<T> u(r_, c_);
        for (size_t q = 0; q < v_.size(); ++q) u.v_[q] = v_[q] - m.v_[q];
        return u;
    }
    matrix<T> operator*(const matrix<T>& m) const
    {
        matrix<T> u(r_, m.c_);
        for (int i = 0; i < r_; ++i)
            for (int k = 0; k < c_; ++k)
                for (int j = 0; j < m.c_; ++j) u(i, j) += (*this)(i, k) * m(k, j);
        return u;
    }
};

template <typename T>
class cmatrix {
    int r_ = 0, c_ = 0;
    std::vector<complex<T>> v_;

public:
    cmatrix() {}
    cmatrix(int rows, int cols) : r_(rows), c_(cols), v_((size_t)rows * cols, complex<T>(0)) {}
    complex<T>& operator()(const int& i, const int& j) { return v_[(size_t)i * c_ + j]; }
    complex<T> operator()(const int& i, const int& j) const { return v_[(size_t)i * c_ + j]; }
    int rows() const { return r_; }
    int cols() const { return c_; }
    complex<T>* data() { return v_.data(); }
    const complex<T>* data() const { return v_.data(); }
    cmatrix<T> adjoint() const
    {
        cmatrix<T> u(c_, r_);
        for (int i = 0; i < r_; ++i)
            for (int j = 0; j < c_; ++j) u(j, i) = conj((*this)(i, j));
        return u;
    }
    cmatrix<T> operator+(const cmatrix<T>& m) const
    {
        cmatrix<T> u(r_, c_);
        for (size_t q = 0; q < v_.size(); ++q) u.v_[q] = v_[q] + m.v_[q];
        return u;
    }
    cmatrix<T> operator-(const cmatrix<T>& m) const
    {
        cmatrix<T> u(r_, c_);
        for (size_t q = 0; q < v_.size(); ++q) u.v_[q] = v_[q] - m.v_[q];
        return u;
    }
    cmatrix<T> operator*(const cmatrix<T>& m) const
    {
        cmatrix<T> u(r_, m.c_);
        for (int i = 0; i < r_; ++i)
            for (int k = 0; k < c_; ++k)
                for (int j = 0; j < m.c_; ++j) u(i, j) += (*this)(i, k) * m(k, j);
        return u;
    }
    cmatrix<T> operator*(const T& a) const
    {
        cmatrix<T> u(r_, c_);
        for (size_t q = 0; q < v_.size(); ++q) u.v_[q] = v_[q] * a;
        return u;
    }
    cmatrix<T> operator*(const complex<T>& a) const
    {
        cmatrix<T> u(r_, c_);
        for (size_t q = 0; q < v_.size(); ++q) u.v_[q] = v_[q] * a;
        return u;
    }
    cmatrix<T> operator/(const T& a) const
    {
        cmatrix<T> u(r_, c_);
        for (size_t q = 0; q < v_.size(); ++q) u.v_[q] = v_[q] / a;
        return u;
    }
};

template <typename T>
class rank3tensor {
    int n1_ = 0, n2_ = 0, n3_ = 0;
    std::vector<T> v_;

public:
    rank3tensor() {}
    rank3tensor(int n1, int n2, int n3) : n1_(n1), n2_(n2), n3_(n3), v_((size_t)n1 * n2 * n3, T(0)) {}
    T& operator()(const int& i1, const int& i2, const int& i3) { return v_[(size_t)i1 + (size_t)n1_ * (i2 + (size_t)n2_ * i3)]; }
    T operator()(const int& i1, const int& i2, const int& i3) const
    {
        return v_[(size_t)i1 + (size_t)n1_ * (i2 + (size_t)n2_ * i3)];
    }
    int dim1() const { return n1_; }
    int dim2() const { return n2_; }
    int dim3() const { return n3_; }
    T* data() { return v_.data(); }
};

#endif

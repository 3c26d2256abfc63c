// wenofv_array.h -- the containers of the CWENOFV / TwoPhaseEquation interface.
//
// The reference's solver class and physics take and hold Array1D / Array2D objects
// (third_party/scTools/Array.h of the reference: construction with a size and optional
// value, Resize, dim / row / col, setZero, element access by () and [], Array2D rows by
// [i] as pointers, printing with <<), so code written against them indexes a state as
// m_Uh[ix][iy].vector[r].  These are this library's own containers with that interface:
// contiguous storage, and vectors of up to four elements (a cell's conserved variables)
// stored inline, so a grid of cells costs no allocation per cell.  Elements are
// value-initialised (the reference leaves them uninitialised).
//
// The facade headers are compiled in the caller's translation unit, so a caller that keeps
// the reference's own container header (third_party/scTools/Array.h, include guard
// __ARRAY_H_) and includes it first gets the facade over those containers; the facade uses
// only the interface both share.  Included first, this header takes that guard, so the
// reference's header is not read a second time.
#pragma once
#ifndef __ARRAY_H_
#define __ARRAY_H_
#include <algorithm>
#include <cmath>
#include <fstream>
#include <initializer_list>
#include <ostream>
#include <string>
#include <vector>

template <typename T>
class Array1D {
public:
    Array1D() = default;
    Array1D(unsigned int n) { Resize(n); }
    Array1D(unsigned int n, const T &val) { Resize(n, val); }
    Array1D(std::initializer_list<T> v) {
        Resize((unsigned int)v.size());
        std::copy(v.begin(), v.end(), data());
    }
    Array1D(const Array1D &A) { *this = A; }
    Array1D &operator=(const Array1D &A) {
        if (this == &A) return *this;
        Resize(A.n_);
        std::copy(A.data(), A.data() + A.n_, data());
        return *this;
    }

    // a size of 0 keeps the array as it is, as the reference's Resize does
    void Resize(unsigned int n) {
        if (n == 0) return;
        n_ = n;
        if (n > INLINE) heap_.assign(n, T());
        else {
            heap_.clear();
            std::fill(inl_, inl_ + INLINE, T());
        }
    }
    void Resize(unsigned int n, const T &val) {
        if (n == 0) return;
        Resize(n);
        std::fill(data(), data() + n_, val);
    }
    unsigned int dim() const { return n_; }
    void setZero() { std::fill(data(), data() + n_, T(0)); }
    void abs() {
        for (unsigned int i = 0; i < n_; ++i) data()[i] = std::fabs(data()[i]);
    }
    double sum() {
        double s = 0;
        for (unsigned int i = 0; i < n_; ++i) s += data()[i];
        return s;
    }
    double max() { return n_ ? *std::max_element(data(), data() + n_) : 0.0; }
    double min() { return n_ ? *std::min_element(data(), data() + n_) : 0.0; }
    // "l1", "l2" or "inf"
    double norm(std::string type = "l2") {
        double s = 0;
        for (unsigned int i = 0; i < n_; ++i) {
            const double a = std::fabs(data()[i]);
            if (type == "l1") s += a;
            else if (type == "inf") s = a > s ? a : s;
            else s += a * a;
        }
        return type == "l2" ? std::sqrt(s) : s;
    }
    void output(std::string filename = "output.txt") {
        std::ofstream f(filename);
        for (unsigned int i = 0; i < n_; ++i) f << data()[i] << "\n";
    }

    T &operator()(unsigned int i) { return data()[i]; }
    const T &operator()(unsigned int i) const { return data()[i]; }
    T &operator[](unsigned int i) { return data()[i]; }
    const T &operator[](unsigned int i) const { return data()[i]; }
    T *data() { return n_ > INLINE ? heap_.data() : inl_; }
    const T *data() const { return n_ > INLINE ? heap_.data() : inl_; }

private:
    static constexpr unsigned int INLINE = 4;
    T inl_[INLINE]{};
    std::vector<T> heap_;
    unsigned int n_ = 0;
};

template <typename T>
std::ostream &operator<<(std::ostream &os, const Array1D<T> &A) {
    for (unsigned int i = 0; i < A.dim(); ++i) os << A[i] << (i + 1 < A.dim() ? " " : "");
    return os;
}

// row-major: A[i] is row i, A[i][j] = A(i, j)
template <typename T>
class Array2D {
public:
    Array2D() = default;
    Array2D(unsigned int rows, unsigned int cols) { Resize(rows, cols); }
    Array2D(unsigned int rows, unsigned int cols, const T &val) { Resize(rows, cols, val); }

    void Resize(unsigned int rows, unsigned int cols) {
        if (rows == 0 || cols == 0) return;
        r_ = rows;
        c_ = cols;
        a_.assign((size_t)rows * cols, T());
    }
    void Resize(unsigned int rows, unsigned int cols, const T &val) {
        if (rows == 0 || cols == 0) return;
        r_ = rows;
        c_ = cols;
        a_.assign((size_t)rows * cols, val);
    }
    unsigned int row() const { return r_; }
    unsigned int col() const { return c_; }
    void setZero() { std::fill(a_.begin(), a_.end(), T(0)); }
    void transpose(Array2D<T> &At) {
        At.Resize(c_, r_);
        for (unsigned int i = 0; i < r_; ++i)
            for (unsigned int j = 0; j < c_; ++j) At(j, i) = (*this)(i, j);
    }

    T &operator()(unsigned int i, unsigned int j) { return a_[(size_t)i * c_ + j]; }
    const T &operator()(unsigned int i, unsigned int j) const { return a_[(size_t)i * c_ + j]; }
    T *operator[](unsigned int i) { return a_.data() + (size_t)i * c_; }
    const T *operator[](unsigned int i) const { return a_.data() + (size_t)i * c_; }

private:
    std::vector<T> a_;
    unsigned int r_ = 0, c_ = 0;
};
#endif

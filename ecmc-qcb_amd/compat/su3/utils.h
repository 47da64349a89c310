// utils.h -- the parts of src/su3/utils.h that callers of the flow path use.
//
// SU3 is a row-major 3x3 complex value type standing in for the reference's
// Eigen matrix (utils.h:12) with the few members callers apply to flow
// results (element access, adjoint, trace, squaredNorm, sums and products).
// exp_su3_luscher (utils.cpp:145-193) runs on the device (qcbf_su3_exp).
#pragma once
#include <array>
#include <complex>

using Complex = std::complex<double>;

class SU3 {
    std::array<Complex, 9> a_{};

public:
    static SU3 Zero() { return SU3{}; }
    static SU3 Identity() {
        SU3 m;
        for (int i = 0; i < 3; i++) m(i, i) = 1.0;
        return m;
    }
    Complex& operator()(int i, int j) { return a_[3 * i + j]; }
    const Complex& operator()(int i, int j) const { return a_[3 * i + j]; }
    Complex* data() { return a_.data(); }
    const Complex* data() const { return a_.data(); }
    [[nodiscard]] SU3 adjoint() const {
        SU3 r;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) r(i, j) = std::conj((*this)(j, i));
        return r;
    }
    [[nodiscard]] Complex trace() const { return a_[0] + a_[4] + a_[8]; }
    [[nodiscard]] double squaredNorm() const {
        double s = 0.0;
        for (const Complex& z : a_) s += std::norm(z);
        return s;
    }
    SU3& operator+=(const SU3& b) {
        for (int k = 0; k < 9; k++) a_[k] += b.a_[k];
        return *this;
    }
    SU3& operator-=(const SU3& b) {
        for (int k = 0; k < 9; k++) a_[k] -= b.a_[k];
        return *this;
    }
    SU3& operator*=(double s) {
        for (Complex& z : a_) z *= s;
        return *this;
    }
    friend SU3 operator+(SU3 a, const SU3& b) { return a += b; }
    friend SU3 operator-(SU3 a, const SU3& b) { return a -= b; }
    friend SU3 operator*(SU3 a, double s) { return a *= s; }
    friend SU3 operator*(double s, SU3 a) { return a *= s; }
    friend SU3 operator*(const SU3& a, const SU3& b) {
        SU3 r;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                r(i, j) = a(i, 0) * b(0, j) + a(i, 1) * b(1, j) + a(i, 2) * b(2, j);
        return r;
    }
};

// exp_su3_luscher(Z, coeff) (src/su3/utils.cpp:145-193) on the device: the
// exponential of coeff * Z for a traceless anti-hermitian Z.  Prints the
// reference's warning when t = -1/2 tr X^2 > 1.
SU3 exp_su3_luscher(const SU3& Z, double coeff);

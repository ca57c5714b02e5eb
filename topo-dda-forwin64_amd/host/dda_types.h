// dda_types.h -- small dense vector types used by the host classes.
//
// The reference's classes take and return Eigen types (VectorXcd, VectorXi,
// VectorXd, Vector3d, Matrix3cd).  Eigen is not a dependency of this library;
// these types keep the same names and the operations the reference's callers
// use: element access with (), size(), Zero(n), +, -, scalar *, and Eigen's
// sesquilinear a.dot(b) = sum conj(a_i) b_i and norm()
// (Eigen/src/Core/Dot.h:61-72 of the vendored Eigen 3.3.90).
#pragma once
#include <cmath>
#include <complex>
#include <cstddef>
#include <vector>

namespace topodda {

using cd = std::complex<double>;

template <typename S>
class DenseVector {
public:
    DenseVector() = default;
    explicit DenseVector(std::size_t n) : v_(n, S(0)) {}
    static DenseVector Zero(std::size_t n) { return DenseVector(n); }
    std::size_t size() const { return v_.size(); }
    S& operator()(std::size_t i) { return v_[i]; }
    const S& operator()(std::size_t i) const { return v_[i]; }
    S* data() { return v_.data(); }
    const S* data() const { return v_.data(); }
    void resize(std::size_t n) { v_.assign(n, S(0)); }
    std::vector<S>& raw() { return v_; }
    const std::vector<S>& raw() const { return v_; }

    DenseVector operator+(const DenseVector& o) const
    {
        DenseVector r(*this);
        for (std::size_t i = 0; i < v_.size(); i++) r.v_[i] += o.v_[i];
        return r;
    }
    DenseVector operator-(const DenseVector& o) const
    {
        DenseVector r(*this);
        for (std::size_t i = 0; i < v_.size(); i++) r.v_[i] -= o.v_[i];
        return r;
    }
    DenseVector operator*(S a) const
    {
        DenseVector r(*this);
        for (auto& x : r.v_) x *= a;
        return r;
    }
    DenseVector operator/(S a) const
    {
        DenseVector r(*this);
        for (auto& x : r.v_) x /= a;
        return r;
    }
    bool operator==(const DenseVector& o) const { return v_ == o.v_; }
    bool operator!=(const DenseVector& o) const { return v_ != o.v_; }

    // sum conj(a_i) b_i for complex, sum a_i b_i for real
    S dot(const DenseVector& o) const
    {
        S s(0);
        for (std::size_t i = 0; i < v_.size(); i++) s += conj_(v_[i]) * o.v_[i];
        return s;
    }
    double norm() const
    {
        double s = 0;
        for (const auto& x : v_) s += std::norm(x);
        return std::sqrt(s);
    }
    S sum() const
    {
        S s(0);
        for (const auto& x : v_) s += x;
        return s;
    }

private:
    static cd conj_(cd a) { return std::conj(a); }
    static double conj_(double a) { return a; }
    static int conj_(int a) { return a; }
    std::vector<S> v_;
};

template <typename S>
DenseVector<S> operator*(S a, const DenseVector<S>& v)
{
    return v * a;
}

using VectorXcd = DenseVector<cd>;
using VectorXd = DenseVector<double>;
using VectorXi = DenseVector<int>;

struct Vector3d {
    double v[3] = {0, 0, 0};
    Vector3d() = default;
    Vector3d(double a, double b, double c) : v{a, b, c} {}
    double& operator()(int i) { return v[i]; }
    double operator()(int i) const { return v[i]; }
    double dot(const Vector3d& o) const { return v[0] * o.v[0] + v[1] * o.v[1] + v[2] * o.v[2]; }
};

struct Matrix3cd {
    cd a[9];
    cd& operator()(int i, int j) { return a[3 * i + j]; }
    const cd& operator()(int i, int j) const { return a[3 * i + j]; }
};

}  // namespace topodda

// The reference's code names these classes without a namespace; define
// TOPODDA_NO_GLOBAL_NAMES to keep them in namespace topodda only.
#ifndef TOPODDA_NO_GLOBAL_NAMES
using topodda::cd;
using topodda::VectorXcd;
using topodda::VectorXd;
using topodda::VectorXi;
using topodda::Vector3d;
using topodda::Matrix3cd;
#endif

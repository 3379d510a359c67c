// A minimal stand-in for the Eigen types of the reference's interface, for builds without Eigen: the
// members the swept-CCD interface and the reference's example programs use (Vector3d, Vector3i,
// VectorXd, Matrix3Xi, Matrix3Xd with construction, element access, arithmetic, dot / cross / norm,
// segment<3> and col).  Arithmetic follows Eigen's for 3-vectors: dot is ((x0 y0 + x1 y1) + x2 y2),
// cross is (a1 b2 - a2 b1, a2 b0 - a0 b2, a0 b1 - a1 b0).  Include it as <Eigen/Core> through
// include/eigen_subset, or through EigenTypes.h when Eigen is not installed.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <vector>

namespace Eigen {

class Vector3d {
public:
    Vector3d() : v_{0, 0, 0} {}
    Vector3d(double x, double y, double z) : v_{x, y, z} {}
    double &operator[](std::ptrdiff_t i) { return v_[i]; }
    double operator[](std::ptrdiff_t i) const { return v_[i]; }
    double &operator()(std::ptrdiff_t i) { return v_[i]; }
    double operator()(std::ptrdiff_t i) const { return v_[i]; }
    double coeff(std::ptrdiff_t i) const { return v_[i]; }
    double &coeffRef(std::ptrdiff_t i) { return v_[i]; }
    double *data() { return v_; }
    const double *data() const { return v_; }
    static constexpr std::ptrdiff_t size() { return 3; }
    void setZero() { v_[0] = v_[1] = v_[2] = 0.0; }
    static Vector3d Zero() { return Vector3d(); }
    Vector3d operator+(const Vector3d &o) const { return {v_[0] + o.v_[0], v_[1] + o.v_[1], v_[2] + o.v_[2]}; }
    Vector3d operator-(const Vector3d &o) const { return {v_[0] - o.v_[0], v_[1] - o.v_[1], v_[2] - o.v_[2]}; }
    Vector3d operator-() const { return {-v_[0], -v_[1], -v_[2]}; }
    Vector3d operator*(double s) const { return {v_[0] * s, v_[1] * s, v_[2] * s}; }
    Vector3d operator/(double s) const { return {v_[0] / s, v_[1] / s, v_[2] / s}; }
    Vector3d &operator*=(double s) { return *this = *this * s; }
    Vector3d &operator/=(double s) { return *this = *this / s; }
    Vector3d &operator+=(const Vector3d &o) { return *this = *this + o; }
    Vector3d &operator-=(const Vector3d &o) { return *this = *this - o; }
    bool operator==(const Vector3d &o) const { return v_[0] == o.v_[0] && v_[1] == o.v_[1] && v_[2] == o.v_[2]; }
    bool operator!=(const Vector3d &o) const { return !(*this == o); }
    double dot(const Vector3d &o) const { return (v_[0] * o.v_[0] + v_[1] * o.v_[1]) + v_[2] * o.v_[2]; }
    Vector3d cross(const Vector3d &o) const
    {
        return {v_[1] * o.v_[2] - v_[2] * o.v_[1], v_[2] * o.v_[0] - v_[0] * o.v_[2], v_[0] * o.v_[1] - v_[1] * o.v_[0]};
    }
    double squaredNorm() const { return dot(*this); }
    double norm() const { return std::sqrt(squaredNorm()); }
    Vector3d normalized() const { return *this / norm(); }

private:
    double v_[3];
};
inline Vector3d operator*(double s, const Vector3d &v) { return {s * v[0], s * v[1], s * v[2]}; }

class Vector3i {
public:
    Vector3i() : v_{0, 0, 0} {}
    Vector3i(int a, int b, int c) : v_{a, b, c} {}
    int &operator[](std::ptrdiff_t i) { return v_[i]; }
    int operator[](std::ptrdiff_t i) const { return v_[i]; }
    int &operator()(std::ptrdiff_t i) { return v_[i]; }
    int operator()(std::ptrdiff_t i) const { return v_[i]; }
    int coeff(std::ptrdiff_t i) const { return v_[i]; }
    static constexpr std::ptrdiff_t size() { return 3; }

private:
    int v_[3];
};

// A 3-element window of a VectorXd (VectorXd::segment<3>): reads as a Vector3d, writes through.
class Segment3 {
public:
    explicit Segment3(double *p) : p_(p) {}
    Segment3 &operator=(const Vector3d &v)
    {
        p_[0] = v[0];
        p_[1] = v[1];
        p_[2] = v[2];
        return *this;
    }
    Segment3 &operator=(const Segment3 &o) { return *this = Vector3d(o); }
    Segment3 &operator+=(const Vector3d &v) { return *this = Vector3d(*this) + v; }
    Segment3 &operator-=(const Vector3d &v) { return *this = Vector3d(*this) - v; }
    operator Vector3d() const { return Vector3d(p_[0], p_[1], p_[2]); }
    double &operator[](std::ptrdiff_t i) { return p_[i]; }
    double operator[](std::ptrdiff_t i) const { return p_[i]; }
    Vector3d operator+(const Vector3d &o) const { return Vector3d(*this) + o; }
    Vector3d operator-(const Vector3d &o) const { return Vector3d(*this) - o; }
    Vector3d operator*(double s) const { return Vector3d(*this) * s; }
    Vector3d operator/(double s) const { return Vector3d(*this) / s; }
    double dot(const Vector3d &o) const { return Vector3d(*this).dot(o); }
    Vector3d cross(const Vector3d &o) const { return Vector3d(*this).cross(o); }
    double norm() const { return Vector3d(*this).norm(); }
    double squaredNorm() const { return Vector3d(*this).squaredNorm(); }

private:
    double *p_;
};

class VectorXd {
public:
    VectorXd() {}
    explicit VectorXd(std::ptrdiff_t n) : v_(n) {}
    std::ptrdiff_t size() const { return (std::ptrdiff_t)v_.size(); }
    void resize(std::ptrdiff_t n) { v_.resize(n); }
    void setZero() { std::fill(v_.begin(), v_.end(), 0.0); }
    static VectorXd Zero(std::ptrdiff_t n) { return VectorXd(n); }
    double &operator[](std::ptrdiff_t i) { return v_[i]; }
    double operator[](std::ptrdiff_t i) const { return v_[i]; }
    double &operator()(std::ptrdiff_t i) { return v_[i]; }
    double operator()(std::ptrdiff_t i) const { return v_[i]; }
    double coeff(std::ptrdiff_t i) const { return v_[i]; }
    double &coeffRef(std::ptrdiff_t i) { return v_[i]; }
    double *data() { return v_.data(); }
    const double *data() const { return v_.data(); }
    template <int N>
    Segment3 segment(std::ptrdiff_t i)
    {
        static_assert(N == 3, "only 3-segments");
        return Segment3(v_.data() + i);
    }
    template <int N>
    Vector3d segment(std::ptrdiff_t i) const
    {
        static_assert(N == 3, "only 3-segments");
        return Vector3d(v_[i], v_[i + 1], v_[i + 2]);
    }
    VectorXd &operator+=(const VectorXd &o)
    {
        for (size_t i = 0; i < v_.size(); i++) v_[i] += o.v_[i];
        return *this;
    }
    VectorXd &operator-=(const VectorXd &o)
    {
        for (size_t i = 0; i < v_.size(); i++) v_[i] -= o.v_[i];
        return *this;
    }
    VectorXd operator+(const VectorXd &o) const { return VectorXd(*this) += o; }
    VectorXd operator-(const VectorXd &o) const { return VectorXd(*this) -= o; }
    VectorXd operator*(double s) const
    {
        VectorXd r(*this);
        for (double &x : r.v_) x *= s;
        return r;
    }

private:
    std::vector<double> v_;
};
inline VectorXd operator*(double s, const VectorXd &v) { return v * s; }

// 3 x n column-major matrices (one column per triangle / per vertex).
template <class T, class Col>
class Matrix3X {
public:
    Matrix3X() : cols_(0) {}
    Matrix3X(std::ptrdiff_t rows, std::ptrdiff_t cols) : cols_(cols), v_(3 * cols) { (void)rows; }
    void resize(std::ptrdiff_t rows, std::ptrdiff_t cols)
    {
        (void)rows;
        cols_ = cols;
        v_.resize(3 * cols);
    }
    std::ptrdiff_t cols() const { return cols_; }
    static constexpr std::ptrdiff_t rows() { return 3; }
    std::ptrdiff_t size() const { return 3 * cols_; }
    void setZero() { std::fill(v_.begin(), v_.end(), T(0)); }
    T coeff(std::ptrdiff_t r, std::ptrdiff_t c) const { return v_[3 * c + r]; }
    T &coeffRef(std::ptrdiff_t r, std::ptrdiff_t c) { return v_[3 * c + r]; }
    T &operator()(std::ptrdiff_t r, std::ptrdiff_t c) { return v_[3 * c + r]; }
    T operator()(std::ptrdiff_t r, std::ptrdiff_t c) const { return v_[3 * c + r]; }
    Col col(std::ptrdiff_t c) const { return Col(v_[3 * c], v_[3 * c + 1], v_[3 * c + 2]); }
    T *data() { return v_.data(); }
    const T *data() const { return v_.data(); }

private:
    std::ptrdiff_t cols_;
    std::vector<T> v_;
};
using Matrix3Xi = Matrix3X<int, Vector3i>;
using Matrix3Xd = Matrix3X<double, Vector3d>;

}   // namespace Eigen

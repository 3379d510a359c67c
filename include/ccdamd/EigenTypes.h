// Vector and matrix types of the reference's interface (Eigen::Vector3d, VectorXd, Matrix3Xi).
// With Eigen on the include path these are Eigen's own; without it, a minimal set of value types
// with the members the swept-CCD interface touches (construction, element access, size, data(),
// segment<3>, col, dot/cross/norm) stands in so that this interface compiles and links on its own.
#pragma once

#if __has_include(<Eigen/Core>)
#include <Eigen/Core>
#include <Eigen/Geometry>
#define CCDAMD_HAVE_EIGEN 1
#else
#include <cmath>
#include <cstddef>
#include <vector>

namespace Eigen {

class Vector3d {
public:
    Vector3d() : v_{0, 0, 0} {}
    Vector3d(double x, double y, double z) : v_{x, y, z} {}
    double &operator[](int i) { return v_[i]; }
    double operator[](int i) const { return v_[i]; }
    double &operator()(int i) { return v_[i]; }
    double operator()(int i) const { return v_[i]; }
    double coeff(int i) const { return v_[i]; }
    double *data() { return v_; }
    const double *data() const { return v_; }
    static constexpr int size() { return 3; }
    Vector3d operator+(const Vector3d &o) const { return {v_[0] + o.v_[0], v_[1] + o.v_[1], v_[2] + o.v_[2]}; }
    Vector3d operator-(const Vector3d &o) const { return {v_[0] - o.v_[0], v_[1] - o.v_[1], v_[2] - o.v_[2]}; }
    Vector3d operator-() const { return {-v_[0], -v_[1], -v_[2]}; }
    Vector3d operator*(double s) const { return {v_[0] * s, v_[1] * s, v_[2] * s}; }
    Vector3d operator/(double s) const { return {v_[0] / s, v_[1] / s, v_[2] / s}; }
    Vector3d &operator/=(double s) { return *this = *this / s; }
    Vector3d &operator+=(const Vector3d &o) { return *this = *this + o; }
    Vector3d &operator-=(const Vector3d &o) { return *this = *this - o; }
    bool operator==(const Vector3d &o) const { return v_[0] == o.v_[0] && v_[1] == o.v_[1] && v_[2] == o.v_[2]; }
    double dot(const Vector3d &o) const { return (v_[0] * o.v_[0] + v_[1] * o.v_[1]) + v_[2] * o.v_[2]; }
    Vector3d cross(const Vector3d &o) const
    {
        return {v_[1] * o.v_[2] - v_[2] * o.v_[1], v_[2] * o.v_[0] - v_[0] * o.v_[2], v_[0] * o.v_[1] - v_[1] * o.v_[0]};
    }
    double squaredNorm() const { return dot(*this); }
    double norm() const { return std::sqrt(squaredNorm()); }

private:
    double v_[3];
};
inline Vector3d operator*(double s, const Vector3d &v) { return v * s; }

class Vector3i {
public:
    Vector3i() : v_{0, 0, 0} {}
    Vector3i(int a, int b, int c) : v_{a, b, c} {}
    int &operator[](int i) { return v_[i]; }
    int operator[](int i) const { return v_[i]; }

private:
    int v_[3];
};

class VectorXd {
public:
    VectorXd() {}
    explicit VectorXd(std::ptrdiff_t n) : v_(n) {}
    std::ptrdiff_t size() const { return (std::ptrdiff_t)v_.size(); }
    void resize(std::ptrdiff_t n) { v_.resize(n); }
    void setZero() { std::fill(v_.begin(), v_.end(), 0.0); }
    double &operator[](std::ptrdiff_t i) { return v_[i]; }
    double operator[](std::ptrdiff_t i) const { return v_[i]; }
    double &operator()(std::ptrdiff_t i) { return v_[i]; }
    double operator()(std::ptrdiff_t i) const { return v_[i]; }
    double *data() { return v_.data(); }
    const double *data() const { return v_.data(); }
    // segment<3>(i): a copy for reading; use setSegment to write.
    template <int N>
    Vector3d segment(std::ptrdiff_t i) const
    {
        static_assert(N == 3, "only 3-segments");
        return Vector3d(v_[i], v_[i + 1], v_[i + 2]);
    }
    void setSegment(std::ptrdiff_t i, const Vector3d &p)
    {
        v_[i] = p[0];
        v_[i + 1] = p[1];
        v_[i + 2] = p[2];
    }

private:
    std::vector<double> v_;
};

class Matrix3Xi {
public:
    Matrix3Xi() : cols_(0) {}
    Matrix3Xi(std::ptrdiff_t rows, std::ptrdiff_t cols) : cols_(cols), v_(3 * cols) { (void)rows; }
    void resize(std::ptrdiff_t rows, std::ptrdiff_t cols)
    {
        (void)rows;
        cols_ = cols;
        v_.resize(3 * cols);
    }
    std::ptrdiff_t cols() const { return cols_; }
    static constexpr std::ptrdiff_t rows() { return 3; }
    std::ptrdiff_t size() const { return 3 * cols_; }
    int coeff(std::ptrdiff_t r, std::ptrdiff_t c) const { return v_[3 * c + r]; }
    int &coeffRef(std::ptrdiff_t r, std::ptrdiff_t c) { return v_[3 * c + r]; }
    int &operator()(std::ptrdiff_t r, std::ptrdiff_t c) { return v_[3 * c + r]; }
    int operator()(std::ptrdiff_t r, std::ptrdiff_t c) const { return v_[3 * c + r]; }
    Vector3i col(std::ptrdiff_t c) const { return Vector3i(v_[3 * c], v_[3 * c + 1], v_[3 * c + 2]); }
    int *data() { return v_.data(); }
    const int *data() const { return v_.data(); }

private:
    std::ptrdiff_t cols_;
    std::vector<int> v_;
};

}   // namespace Eigen
#endif

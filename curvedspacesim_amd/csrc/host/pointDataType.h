// Value types of the reference-API host layer, without CGAL: point3 / vector3 behave like the CGAL Point_3 /
// Vector_3 of the reference (inc/cgalIncludesAndTypedefs.h) as far as the simulation classes use them, and
// meshPosition is the reference's {barycentric point, face index} pair (inc/pointDataType.h:10-28).
#pragma once

#include <cmath>

struct double3 {
    double x = 0, y = 0, z = 0;
};

class vector3 {
public:
    vector3() = default;
    vector3(double a, double b, double c) : v{a, b, c} {}
    double x() const { return v[0]; }
    double y() const { return v[1]; }
    double z() const { return v[2]; }
    double operator[](int i) const { return v[i]; }
    double& operator[](int i) { return v[i]; }
    double squared_length() const { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }
    vector3 operator+(const vector3& o) const { return {v[0] + o.v[0], v[1] + o.v[1], v[2] + o.v[2]}; }
    vector3 operator-(const vector3& o) const { return {v[0] - o.v[0], v[1] - o.v[1], v[2] - o.v[2]}; }
    vector3 operator-() const { return {-v[0], -v[1], -v[2]}; }
    // CGAL's Vector_3 * Vector_3 is the scalar product
    double operator*(const vector3& o) const { return v[0] * o.v[0] + v[1] * o.v[1] + v[2] * o.v[2]; }
    vector3 operator*(double s) const { return {v[0] * s, v[1] * s, v[2] * s}; }
    vector3 operator/(double s) const { return {v[0] / s, v[1] / s, v[2] / s}; }
    vector3& operator+=(const vector3& o) {
        for (int i = 0; i < 3; ++i) v[i] += o.v[i];
        return *this;
    }
    vector3& operator-=(const vector3& o) {
        for (int i = 0; i < 3; ++i) v[i] -= o.v[i];
        return *this;
    }
    vector3& operator*=(double s) {
        for (double& a : v) a *= s;
        return *this;
    }
    vector3& operator/=(double s) {
        for (double& a : v) a /= s;
        return *this;
    }

private:
    double v[3] = {0, 0, 0};
};

inline vector3 operator*(double s, const vector3& a) { return a * s; }

inline vector3 cross_product(const vector3& a, const vector3& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}

class point3 {
public:
    point3() = default;
    point3(double a, double b, double c) : p{a, b, c} {}
    double x() const { return p[0]; }
    double y() const { return p[1]; }
    double z() const { return p[2]; }
    double operator[](int i) const { return p[i]; }
    double& operator[](int i) { return p[i]; }
    vector3 operator-(const point3& o) const { return {p[0] - o.p[0], p[1] - o.p[1], p[2] - o.p[2]}; }
    point3 operator+(const vector3& d) const { return {p[0] + d[0], p[1] + d[1], p[2] + d[2]}; }

private:
    double p[3] = {0, 0, 0};
};

inline double squared_distance(const point3& a, const point3& b) { return (a - b).squared_length(); }

// a particle location: barycentric weights in the order of the face's corners, and the face index (-1: unset)
class meshPosition {
public:
    meshPosition() = default;
    meshPosition(point3 _x, int _fIdx) : x(_x), faceIndex(_fIdx) {}
    point3 x;
    int faceIndex = -1;
};

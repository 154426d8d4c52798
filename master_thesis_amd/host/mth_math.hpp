// mth_math.hpp — host-side vector / matrix arithmetic with the reference's
// conventions and rounding: float components, dot products summed from 0 left
// to right, normalisation in double, row-vector products v * M
// (software_renderer/math/vector.h, math/matrix.h, math/matrix.cpp,
// math/rotation.cpp).  Compiled with -ffp-contract=off.
#pragma once

#include <cmath>
#include <cstdint>

namespace mth {

struct Vec3 {
    float v[3] = {0.0f, 0.0f, 0.0f};
    Vec3() = default;
    Vec3(float x, float y, float z) { v[0] = x; v[1] = y; v[2] = z; }
    float& operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
};

inline Vec3 operator+(const Vec3& a, const Vec3& b) { return Vec3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
inline Vec3 operator-(const Vec3& a, const Vec3& b) { return Vec3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
inline float dot(const Vec3& a, const Vec3& b) {
    float s = 0.0f;
    for (int i = 0; i < 3; ++i) s += a[i] * b[i];
    return s;
}
inline Vec3 cross(const Vec3& a, const Vec3& b) {
    return Vec3(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
// |v|^2 in float, magnitude and factor in double, products rounded to float
inline Vec3 normalized(Vec3 a) {
    double m = std::sqrt((double)dot(a, a));
    if (m != 0) {
        double f = 1.0 / m;
        for (int i = 0; i < 3; ++i) a[i] = (float)((double)a[i] * f);
    }
    return a;
}
// double scalar times float vector, each product rounded to float
inline Vec3 scaled(double s, const Vec3& a) { return Vec3((float)(s * a[0]), (float)(s * a[1]), (float)(s * a[2])); }

// 4x4 float matrix, m[i][j], acting on row vectors.
struct Mat4 {
    float m[4][4];
    Mat4() {
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) m[i][j] = (i == j) ? 1.0f : 0.0f;
    }
};

// Matrix product, each element summed from 0 over n (matrix.h:216-241)
inline Mat4 mul(const Mat4& a, const Mat4& b) {
    Mat4 r;
    for (int i = 0; i < 4; ++i) {
        r.m[i][i] = 0.0f;
        for (int j = 0; j < 4; ++j)
            for (int n = 0; n < 4; ++n) r.m[i][j] += a.m[i][n] * b.m[n][j];
    }
    return r;
}

// v * M (matrix.h:245-253)
inline void vec4_mul(const Mat4& M, const float v[4], float out[4]) {
    for (int j = 0; j < 4; ++j) out[j] = 0.0f;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) out[j] += M.m[i][j] * v[i];
}

// Determinant by expansion along the first row, recursively
// (matrix.h:128-145,191-210); inverse = (1/det) * cofactor^T (matrix.h:176-189,212-214).
template <int S>
struct SqM {
    float m[S][S];
};

template <int S>
inline SqM<S - 1> minor_of(const SqM<S>& a, int ri, int rj) {
    SqM<S - 1> r;
    for (int i = 0; i < S - 1; ++i)
        for (int j = 0; j < S - 1; ++j) r.m[i][j] = a.m[i < ri ? i : i + 1][j < rj ? j : j + 1];
    return r;
}

inline float det(const SqM<2>& a) { return a.m[0][0] * a.m[1][1] - a.m[0][1] * a.m[1][0]; }

template <int S>
inline float det(const SqM<S>& a) {
    float d = 0;
    for (int i = 0; i < S; ++i) {
        float sub = det(minor_of<S>(a, 0, i));
        if (i % 2 == 0) d += a.m[0][i] * sub;
        else d -= a.m[0][i] * sub;
    }
    return d;
}

inline Mat4 inverse(const Mat4& M) {
    SqM<4> a;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) a.m[i][j] = M.m[i][j];
    float inv_det = 1.0f / det(a);
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            float c = det(minor_of<4>(a, j, i));
            if ((i + j) % 2 == 1) c *= -1;
            r.m[i][j] = c * inv_det;
        }
    return r;
}

inline Mat4 scale_matrix(const Vec3& s) {
    Mat4 r;
    r.m[0][0] = s[0]; r.m[1][1] = s[1]; r.m[2][2] = s[2];
    return r;
}

inline Mat4 translation_matrix(const Vec3& p) {
    Mat4 r;
    r.m[3][0] = p[0]; r.m[3][1] = p[1]; r.m[3][2] = p[2];
    return r;
}

// matrix.cpp:23-44
inline Mat4 look_at(const Vec3& position, const Vec3& target, const Vec3& up) {
    Mat4 r;
    const Vec3 f = normalized(target - position);
    const Vec3 s = normalized(cross(f, up));
    const Vec3 u = cross(s, f);
    r.m[0][0] = s[0]; r.m[1][0] = s[1]; r.m[2][0] = s[2];
    r.m[0][1] = u[0]; r.m[1][1] = u[1]; r.m[2][1] = u[2];
    r.m[0][2] = -f[0]; r.m[1][2] = -f[1]; r.m[2][2] = -f[2];
    r.m[3][0] = -dot(s, position);
    r.m[3][1] = -dot(u, position);
    r.m[3][2] = dot(f, position);
    return r;
}

// matrix.cpp:46-61
inline Mat4 perspective(float aspect, float fov, float near_z, float far_z, bool invert_y) {
    Mat4 r;
    const float s = (float)std::tan((double)(fov / 2.0f));
    r.m[0][0] = 1.0f / (aspect * s);
    r.m[1][1] = 1.0f / s;
    r.m[2][2] = -(far_z + near_z) / (far_z - near_z);
    r.m[2][3] = -1.0f;
    r.m[3][2] = -(2.0f * far_z * near_z) / (far_z - near_z);
    r.m[3][3] = 0.0f;
    if (invert_y) r.m[1][1] *= -1;
    return r;
}

// Unit quaternion (s, v) from axis and angle, rotation.cpp:6-8, and its
// matrix, rotation.cpp:35-56.
struct Rotation {
    float s = 1.0f;
    Vec3 v;
    bool set = false;
    static Rotation from_axis_angle(Vec3 axis, float angle) {
        Rotation r;
        r.s = (float)std::cos((double)(angle / 2.0f));
        r.v = scaled(std::sin((double)(angle / 2.0f)), normalized(axis));
        r.set = true;
        return r;
    }
    Mat4 matrix() const {
        Mat4 r;
        if (std::sqrt((double)dot(v, v)) > 0) {
            float a2 = s * s, x2 = v[0] * v[0], y2 = v[1] * v[1], z2 = v[2] * v[2];
            r.m[0][0] = a2 + x2 - y2 - z2;
            r.m[1][1] = a2 - x2 + y2 - z2;
            r.m[2][2] = a2 - x2 - y2 + z2;
            r.m[3][3] = a2 + x2 + y2 + z2;
            r.m[0][1] = 2.0f * (v[0] * v[1] + s * v[2]);
            r.m[0][2] = 2.0f * (v[0] * v[2] - s * v[1]);
            r.m[1][0] = 2.0f * (v[0] * v[1] - s * v[2]);
            r.m[1][2] = 2.0f * (v[1] * v[2] + s * v[0]);
            r.m[2][0] = 2.0f * (v[0] * v[2] + s * v[1]);
            r.m[2][1] = 2.0f * (v[1] * v[2] - s * v[0]);
        }
        return r;
    }
};

}  // namespace mth

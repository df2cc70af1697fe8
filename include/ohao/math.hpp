// Small vector / quaternion / matrix types of the C++ façade. The reference
// uses glm 0.9.9.8 for these (external/cmake/glm.cmake:7-8); the façade keeps
// the same conventions (column-major matrices, radians, Euler XYZ
// quaternion constructor) so caller code ports by swapping the type names.
#pragma once

#include <cmath>

namespace ohao {

struct vec2 {
    float x = 0, y = 0;
    constexpr vec2() = default;
    constexpr vec2(float a, float b) : x(a), y(b) {}
};

struct vec3 {
    float x = 0, y = 0, z = 0;
    constexpr vec3() = default;
    constexpr explicit vec3(float s) : x(s), y(s), z(s) {}
    constexpr vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

struct vec4 {
    float x = 0, y = 0, z = 0, w = 0;
    constexpr vec4() = default;
    constexpr vec4(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
    constexpr vec4(const vec3& v, float d) : x(v.x), y(v.y), z(v.z), w(d) {}
};

inline vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec3 operator-(vec3 a) { return {-a.x, -a.y, -a.z}; }
inline vec3 operator*(vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline vec3 operator*(float s, vec3 a) { return a * s; }
inline vec3 operator*(vec3 a, vec3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline vec3 operator/(vec3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline vec3 cross(vec3 a, vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(vec3 a) { return std::sqrt(dot(a, a)); }
inline vec3 normalize(vec3 a) { return a / length(a); }
inline vec3 min(vec3 a, vec3 b) { return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
inline vec3 max(vec3 a, vec3 b) { return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }
inline float radians(float deg) { return deg * 0.017453292519943295f; }
inline float degrees(float rad) { return rad * 57.29577951308232f; }
inline vec3 radians(vec3 d) { return {radians(d.x), radians(d.y), radians(d.z)}; }

// Unit quaternion (w, x, y, z).
struct quat {
    float w = 1, x = 0, y = 0, z = 0;
    constexpr quat() = default;
    constexpr quat(float w_, float x_, float y_, float z_) : w(w_), x(x_), y(y_), z(z_) {}
    // Rotation from Euler angles in radians (pitch about X, yaw about Y, roll
    // about Z), composed as R = Rz * Ry * Rx.
    explicit quat(vec3 euler) {
        const float cx = std::cos(euler.x * 0.5f), sx = std::sin(euler.x * 0.5f);
        const float cy = std::cos(euler.y * 0.5f), sy = std::sin(euler.y * 0.5f);
        const float cz = std::cos(euler.z * 0.5f), sz = std::sin(euler.z * 0.5f);
        w = cx * cy * cz + sx * sy * sz;
        x = sx * cy * cz - cx * sy * sz;
        y = cx * sy * cz + sx * cy * sz;
        z = cx * cy * sz - sx * sy * cz;
    }
};

inline vec3 rotate(const quat& q, vec3 v) {
    const vec3 u{q.x, q.y, q.z};
    const vec3 t = 2.0f * cross(u, v);
    return v + q.w * t + cross(u, t);
}

// Column-major 4x4 (m[col][row]).
struct mat4 {
    float m[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    const float* data() const { return &m[0][0]; }
    vec3 point(vec3 p) const {
        return {m[0][0] * p.x + m[1][0] * p.y + m[2][0] * p.z + m[3][0],
                m[0][1] * p.x + m[1][1] * p.y + m[2][1] * p.z + m[3][1],
                m[0][2] * p.x + m[1][2] * p.y + m[2][2] * p.z + m[3][2]};
    }
};

// translate * rotate * scale
inline mat4 trs(vec3 t, const quat& r, vec3 s) {
    mat4 M;
    const vec3 cx = rotate(r, {s.x, 0, 0}), cy = rotate(r, {0, s.y, 0}), cz = rotate(r, {0, 0, s.z});
    M.m[0][0] = cx.x; M.m[0][1] = cx.y; M.m[0][2] = cx.z; M.m[0][3] = 0;
    M.m[1][0] = cy.x; M.m[1][1] = cy.y; M.m[1][2] = cy.z; M.m[1][3] = 0;
    M.m[2][0] = cz.x; M.m[2][1] = cz.y; M.m[2][2] = cz.z; M.m[2][3] = 0;
    M.m[3][0] = t.x;  M.m[3][1] = t.y;  M.m[3][2] = t.z;  M.m[3][3] = 1;
    return M;
}

}  // namespace ohao

// Oracle vector/quaternion helpers, FP32, in the reference's operation
// order (scene.cpp:52-70 rotate/cross, 101-110 multiply). Templated on the
// scalar so the FP64 twin of each numeric stage shares the same sequence.
// Test infrastructure only.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

template <class T>
struct V3 {
    T x, y, z;
};
template <class T>
struct Q4 {
    T x, y, z, w;
};

template <class T>
inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T>
inline T dot(V3<T> a, V3<T> b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class T>
inline V3<T> add(V3<T> a, V3<T> b) {
    return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class T>
inline V3<T> sub(V3<T> a, V3<T> b) {
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class T>
inline V3<T> scale(V3<T> a, T s) {
    return {a.x * s, a.y * s, a.z * s};
}
// PxQuat::rotate as restated at scene.cpp:57-66: t = 2(u x v),
// v + w t + u x t.
template <class T>
inline V3<T> rotate(Q4<T> q, V3<T> v) {
    const V3<T> u{q.x, q.y, q.z};
    const V3<T> c = cross(u, v);
    const V3<T> t{c.x * T(2), c.y * T(2), c.z * T(2)};
    const V3<T> d = cross(u, t);
    return {v.x + q.w * t.x + d.x, v.y + q.w * t.y + d.y, v.z + q.w * t.z + d.z};
}
template <class T>
inline V3<T> rotate_inverse(Q4<T> q, V3<T> v) {
    return rotate(Q4<T>{-q.x, -q.y, -q.z, q.w}, v);
}
// Hamilton product, scene.cpp:101-110.
template <class T>
inline Q4<T> multiply(Q4<T> a, Q4<T> b) {
    return {a.w * b.x + b.w * a.x + a.y * b.z - a.z * b.y,
            a.w * b.y + b.w * a.y + a.z * b.x - a.x * b.z,
            a.w * b.z + b.w * a.z + a.x * b.y - a.y * b.x,
            a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}

inline std::uint32_t f2u(float f) {
    std::uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
// Order-preserving float -> u32 (scene.cpp:25-28) and the 4-bit
// conservative min/max expansion (scene.cpp:29-34).
inline std::uint32_t encode_float(float v) {
    const std::uint32_t bits = f2u(v);
    return (bits & 0x80000000u) ? ~bits : (bits | 0x80000000u);
}
inline std::uint32_t encode_min(float v) { return ((encode_float(v) >> 4) - 1u) << 4; }
inline std::uint32_t encode_max(float v) {
    return (((encode_float(v) >> 4) + 1u) << 4) | 4u;
}

}  // namespace orc

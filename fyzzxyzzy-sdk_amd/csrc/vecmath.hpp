// Float vector/quaternion helpers shared by the PCM and Dy kernels, in the
// reference's FP32 operation order (no contraction: the library builds with
// -ffp-contract=off).
#pragma once
#include "common.hpp"

namespace fyzz_hip::vm {

struct V {
    float x, y, z;
};
struct Q {
    float x, y, z, w;
};
__device__ inline V vadd(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline V vsub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline V vmul(V a, float f) { return {a.x * f, a.y * f, a.z * f}; }
__device__ inline V vneg(V a) { return {-a.x, -a.y, -a.z}; }
__device__ inline float vdot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline V vcross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ inline float comp(V v, unsigned i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
// std::min / std::max / std::clamp semantics (first argument on ties).
__device__ inline float smin(float a, float b) { return b < a ? b : a; }
__device__ inline float smax(float a, float b) { return a < b ? b : a; }
__device__ inline float sclamp(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }
constexpr float kFltMax = 3.402823466e+38f;
constexpr float kFltEps = 1.1920928955078125e-7f;

// scene.cpp:57-66
__device__ inline V srot(Q q, V v) {
    const V u{q.x, q.y, q.z};
    const V c = vcross(u, v);
    const V t{c.x * 2.0f, c.y * 2.0f, c.z * 2.0f};
    const V d = vcross(u, t);
    return {v.x + q.w * t.x + d.x, v.y + q.w * t.y + d.y, v.z + q.w * t.z + d.z};
}
__device__ inline V srot_inv(Q q, V v) { return srot(Q{-q.x, -q.y, -q.z, q.w}, v); }
// scene.cpp:101-110
__device__ inline Q qmul(Q a, Q b) {
    return {a.w * b.x + b.w * a.x + a.y * b.z - a.z * b.y, a.w * b.y + b.w * a.y + a.z * b.x - a.x * b.z,
            a.w * b.z + b.w * a.z + a.x * b.y - a.y * b.x, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
// GuPCMContactBoxBox.cpp:29-37 / GuPCMGjkEpa.cpp:34-40
__device__ inline V grot(Q q, V v) {
    const V u{q.x, q.y, q.z};
    const V t = vmul(vcross(u, v), 2.0f);
    return vadd(v, vadd(vmul(t, q.w), vcross(u, t)));
}
__device__ inline V grot_inv(Q q, V v) {
    const V u{-q.x, -q.y, -q.z};
    const V t = vmul(vcross(u, v), 2.0f);
    return vadd(v, vadd(vmul(t, q.w), vcross(u, t)));
}

}  // namespace fyzz_hip::vm

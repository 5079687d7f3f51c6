// World AABB keys of boxes with the reference's 32-bit endpoint encoding,
// shared by the broadphase objects and the per-scene broadphase of the
// step pipeline.
#pragma once
#include "common.hpp"

namespace fyzz_hip::keys {

__device__ inline u32 encode_float(float v) {
    const u32 bits = __float_as_uint(v);
    return (bits & 0x80000000u) ? ~bits : (bits | 0x80000000u);
}
__device__ inline u32 encode_min(float v) { return ((encode_float(v) >> 4) - 1u) << 4; }
__device__ inline u32 encode_max(float v) { return (((encode_float(v) >> 4) + 1u) << 4) | 4u; }

// World AABB keys of a box (reference scene.cpp:1233-1316, box branch):
// extent = sum |R e_i| h_i unless the orientation is exactly identity,
// plus the contact offset. Same operation order; contraction is off.
__device__ inline bool box_keys(const float* pos, const float* quat, const float* half,
                                const float* offset, u32 i, u32 out[6], const uint8_t* sphere = nullptr) {
    const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
    const float hx = half[3 * i], hy = half[3 * i + 1], hz = half[3 * i + 2];
    const float qx = quat[4 * i], qy = quat[4 * i + 1], qz = quat[4 * i + 2], qw = quat[4 * i + 3];
    float ex = hx, ey = hy, ez = hz;
    // a sphere (radius in all three half extents) ignores its orientation
    // (scene.cpp:1243-1245)
    const bool round = sphere && sphere[i];
    if (!round && (qx != 0.0f || qy != 0.0f || qz != 0.0f || qw != 1.0f)) {
        // rotate(q, e_k): t = 2 (u x e_k); r = e_k + w t + u x t.
        auto rot = [&](float vx, float vy, float vz, float& rx, float& ry, float& rz) {
            const float cx = qy * vz - qz * vy, cy = qz * vx - qx * vz, cz = qx * vy - qy * vx;
            const float tx = cx * 2.0f, ty = cy * 2.0f, tz = cz * 2.0f;
            const float dx = qy * tz - qz * ty, dy = qz * tx - qx * tz, dz = qx * ty - qy * tx;
            rx = vx + qw * tx + dx;
            ry = vy + qw * ty + dy;
            rz = vz + qw * tz + dz;
        };
        float xx, xy, xz, yx, yy, yz, zx, zy, zz;
        rot(1.0f, 0.0f, 0.0f, xx, xy, xz);
        rot(0.0f, 1.0f, 0.0f, yx, yy, yz);
        rot(0.0f, 0.0f, 1.0f, zx, zy, zz);
        ex = fabsf(xx) * hx + fabsf(yx) * hy + fabsf(zx) * hz;
        ey = fabsf(xy) * hx + fabsf(yy) * hy + fabsf(zy) * hz;
        ez = fabsf(xz) * hx + fabsf(yz) * hy + fabsf(zz) * hz;
    }
    const float c = offset[i];
    ex += c;
    ey += c;
    ez += c;
    const float b[6] = {px - ex, py - ey, pz - ez, px + ex, py + ey, pz + ez};
    bool finite = true;
#pragma unroll
    for (int k = 0; k < 6; ++k) finite &= isfinite(b[k]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        out[k] = encode_min(b[k]);
        out[k + 3] = encode_max(b[k + 3]);
    }
    return finite;
}

}  // namespace fyzz_hip::keys

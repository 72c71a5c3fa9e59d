// Types.h — the plain vector types of the reference's include/core/Types.h:40-77
// (particle and GPU-buffer data), with the helpers its callers use.
#pragma once

#include <cmath>
#include <cstdint>

using F32 = float;
using I32 = std::int32_t;
using U32 = std::uint32_t;

struct PVec3 {
    F32 x, y, z;
};
struct PVec4 {
    F32 x, y, z, w;
};
struct UVec2 {
    U32 x, y;
};
struct IVec2 {
    I32 x, y;
};

inline PVec3 make_pvec3(F32 x, F32 y, F32 z) { return {x, y, z}; }
inline PVec3 operator+(PVec3 a, PVec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline PVec3 operator-(PVec3 a, PVec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline PVec3 operator*(PVec3 a, F32 s) { return {a.x * s, a.y * s, a.z * s}; }
inline PVec3 operator*(F32 s, PVec3 a) { return a * s; }
inline PVec3& operator+=(PVec3& a, PVec3 b) { return a = a + b; }
inline PVec3& operator-=(PVec3& a, PVec3 b) { return a = a - b; }
inline PVec3& operator*=(PVec3& a, F32 s) { return a = a * s; }
inline F32 dot(PVec3 a, PVec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F32 norm2(PVec3 a) { return dot(a, a); }
inline F32 norm(PVec3 a) { return std::sqrt(norm2(a)); }

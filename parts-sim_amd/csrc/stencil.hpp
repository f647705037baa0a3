// Range-checked buffer accesses shared by the z-marching operator kernels (fine.hip, fine_el.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace psx {

// Raw buffer accesses over one plane: the descriptor's byte count is the plane's, so a lane whose
// offset is OUT_OF_RANGE (a position outside the part) reads 0 and its store is dropped, and a plane
// outside the grid (count 0) reads 0 everywhere, without predicated branches. Plane byte counts stay
// below 2^31 (checkGridLimits bounds a node plane at 2^31 / 24 nodes).
constexpr int32_t OUT_OF_RANGE = int32_t(0x80000000u);
constexpr int RSRC_WORD3 = 0x00020000;  // raw buffer, 32-bit data (gfx9 family)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t planeRsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, int(bytes), RSRC_WORD3);
}
// a node's three components at byte offset off
__device__ __forceinline__ void bload3(__amdgpu_buffer_rsrc_t rs, int32_t off, float (&d)[3]) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b96(rs, off, 0, 0);
    d[0] = __uint_as_float(v[0]);
    d[1] = __uint_as_float(v[1]);
    d[2] = __uint_as_float(v[2]);
}
__device__ __forceinline__ void bload3(__amdgpu_buffer_rsrc_t rs, int32_t off, double (&d)[3]) {
    const auto a = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
    const auto b = __builtin_amdgcn_raw_buffer_load_b64(rs, off + 16, 0, 0);
    d[0] = __hiloint2double(int(a[1]), int(a[0]));
    d[1] = __hiloint2double(int(a[3]), int(a[2]));
    d[2] = __hiloint2double(int(b[1]), int(b[0]));
}
__device__ __forceinline__ void bstore3(__amdgpu_buffer_rsrc_t rs, int32_t off, const float (&d)[3]) {
    typedef unsigned u3 __attribute__((ext_vector_type(3)));
    const u3 v = {__float_as_uint(d[0]), __float_as_uint(d[1]), __float_as_uint(d[2])};
    __builtin_amdgcn_raw_buffer_store_b96(v, rs, off, 0, 0);
}
__device__ __forceinline__ void bstore3(__amdgpu_buffer_rsrc_t rs, int32_t off, const double (&d)[3]) {
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    const u4 a = {unsigned(__double2loint(d[0])), unsigned(__double2hiint(d[0])), unsigned(__double2loint(d[1])), unsigned(__double2hiint(d[1]))};
    const u2 b = {unsigned(__double2loint(d[2])), unsigned(__double2hiint(d[2]))};
    __builtin_amdgcn_raw_buffer_store_b128(a, rs, off, 0, 0);
    __builtin_amdgcn_raw_buffer_store_b64(b, rs, off + 16, 0, 0);
}

}  // namespace psx

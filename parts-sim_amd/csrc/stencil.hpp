// Range-checked buffer accesses shared by the z-marching operator kernels (fine.hip, fine_el.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <utility>

namespace psx {

// The reference element's corners (x, y, z offsets of corner b, numbered 0..3 counter-clockwise in
// the lower face, 4..7 above) and its symmetry classes
__device__ constexpr int cx(int b) { return (b == 1 || b == 2 || b == 5 || b == 6) ? 1 : 0; }
__device__ constexpr int cy(int b) { return (b == 2 || b == 3 || b == 6 || b == 7) ? 1 : 0; }
__device__ constexpr int cz(int b) { return b >= 4 ? 1 : 0; }
__device__ constexpr int cornerOf(int x, int y, int z) { return (y ? (x ? 2 : 3) : (x ? 1 : 0)) + 4 * z; }
// Symmetry classes (the host's elementClass, hierarchy.cpp): every element coefficient is
// +-one of 22 values, at an index known at compile time once the loops are unrolled, so the values
// stay in scalar registers for the whole kernel.
__device__ constexpr int elementClass(int a, int d, int b, int c) {
    const int ca[3] = {cx(a), cy(a), cz(a)};
    int r[3] = {cx(a) ^ cx(b), cy(a) ^ cy(b), cz(a) ^ cz(b)};
    const bool neg = ((ca[d] + ca[c]) & 1) != 0;
    int t = r[0];
    r[0] = r[d];
    r[d] = t;
    int c2 = c == 0 ? d : (c == d ? 0 : c);
    if (c2 == 2) {
        t = r[1];
        r[1] = r[2];
        r[2] = t;
        c2 = 1;
    } else if (c2 == 0 && r[1] < r[2]) {
        t = r[1];
        r[1] = r[2];
        r[2] = t;
    }
    const int idx = c2 == 0 ? r[0] * 3 + r[1] + r[2] : 6 + r[0] + 2 * r[1] + 4 * r[2];
    return idx | (neg ? 0x100 : 0);
}
template <class T> struct Sym {
    T v[22];
};
// coefficient times x, class and sign fixed at compile time
template <int CLS, class T> __device__ __forceinline__ T coefMul(const Sym<T>& C, T x) {
    if constexpr ((CLS & 0x100) != 0) return -C.v[CLS & 0xff] * x;
    else return C.v[CLS & 0xff] * x;
}
// loop with a compile-time index: f(std::integral_constant<int, i>) for i in [0, N)
template <class F, int... I> __device__ __forceinline__ void staticForImpl(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F> __device__ __forceinline__ void staticFor(F&& f) { staticForImpl(f, std::make_integer_sequence<int, N>{}); }

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

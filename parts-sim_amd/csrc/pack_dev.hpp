// Packed launches, device side (included by the kernel sources): a kernel body is a struct with
// NTH (threads per block), WPE (waves per SIMD to hold registers for, 0 = the compiler's choice) and
// a device function run(VBlk, parameters...) that reads its block position only from the VBlk. The
// named kernel calls run with its own block index; kPacked calls it for a block of a packed launch
// with the block index and arguments of the launch that block belongs to.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <utility>

#include "device.hpp"
#include "pack.hpp"

namespace psx {

// a block's position in its own launch: x, y and the grid's x extent
struct VBlk {
    uint32_t x, y, gx;
};
__device__ __forceinline__ VBlk ownBlock() { return VBlk{blockIdx.x, blockIdx.y, gridDim.x}; }

// a kernel's parameters as one plain-data block (the same layout on host and device)
template <class... A> struct Args {};
template <class H, class... R> struct Args<H, R...> {
    H h;
    Args<R...> r;
};
template <size_t I, class H, class... R> __host__ __device__ __forceinline__ constexpr const auto& argAt(const Args<H, R...>& a) {
    if constexpr (I == 0) return a.h;
    else return argAt<I - 1>(a.r);
}
template <class... A> struct MakeArgs;
template <> struct MakeArgs<> {
    static Args<> make() { return {}; }
};
template <class H, class... R> struct MakeArgs<H, R...> {
    template <class V, class... W> static Args<H, R...> make(V&& v, W&&... w) {
        return Args<H, R...>{H(static_cast<V&&>(v)), MakeArgs<R...>::make(static_cast<W&&>(w)...)};
    }
};

// Pointers read from the argument table carry no address space, where a kernel's pointer parameters
// are known to point to global memory and its const __restrict__ ones to memory the kernel does not
// write; without that the compiler issues flat loads and stores, and vector loads for the uniform
// table reads that the named kernel makes scalar. Re-deriving each pointer from its bits restores
// both: outputs in the global address space, read-only inputs (pointers to const: every one of
// them is read-only in the solve kernels) in the constant address space.
template <class T> __device__ __forceinline__ T* asGlobal(T* p) {
    if constexpr (std::is_const_v<T>) return (T*)(__attribute__((address_space(4))) T*)(__builtin_bit_cast(uint64_t, p));
    else return (T*)(__attribute__((address_space(1))) T*)(__builtin_bit_cast(uint64_t, p));
}
template <class T> __device__ __forceinline__ T globalArg(T v) {
    if constexpr (std::is_pointer_v<T>) return asGlobal(v);
    else {
        // the reductions' finishers hold their conjugate-gradient state by pointer
        if constexpr (requires { v.st0; }) v.st0 = asGlobal(v.st0);
        return v;
    }
}
__device__ __forceinline__ Reduce globalArg(Reduce v) {
    v.partials = asGlobal(v.partials);
    v.counter = asGlobal(v.counter);
    return v;
}
__device__ __forceinline__ DevLevel globalArg(DevLevel v) {
    v.words = asGlobal(v.words);
    v.Kown = asGlobal(v.Kown);
    v.KownH = asGlobal(v.KownH);
    v.ownScale = asGlobal(v.ownScale);
    v.Kb64 = asGlobal(v.Kb64);
    v.Kb32 = asGlobal(v.Kb32);
    v.St64 = asGlobal(v.St64);
    v.St32 = asGlobal(v.St32);
    v.elA = asGlobal(v.elA);
    v.nodeScale = asGlobal(v.nodeScale);
    v.nodeInfo = asGlobal(v.nodeInfo);
    v.childS = asGlobal(v.childS);
    v.crow64 = asGlobal(v.crow64);
    v.crow32 = asGlobal(v.crow32);
    v.childTwo = asGlobal(v.childTwo);
    v.ctab64 = asGlobal(v.ctab64);
    v.ctab32 = asGlobal(v.ctab32);
    v.ctab16 = asGlobal(v.ctab16);
    v.ctabRowScale = asGlobal(v.ctabRowScale);
    v.invD64 = asGlobal(v.invD64);
    v.invD32 = asGlobal(v.invD32);
    v.diagAdd = asGlobal(v.diagAdd);
    v.ownVox = asGlobal(v.ownVox);
    v.childA = asGlobal(v.childA);
    v.ownT = asGlobal(v.ownT);
    return v;
}

// the block's launch: the one whose block range holds it (binary search over the head, uniform)
template <class B, class... A, size_t... I>
__device__ __forceinline__ void runPacked(const PackHead& h, const Args<A...>* __restrict__ args, std::index_sequence<I...>) {
    const uint32_t bx = blockIdx.x;
    int lo = 0, hi = h.n;  // first[lo] <= bx < first[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (h.first[mid] <= bx) lo = mid;
        else hi = mid;
    }
    const uint32_t local = bx - h.first[lo], gx = h.gx[lo];
    const Args<A...>& a = args[lo];
    B::run(VBlk{local % gx, local / gx, gx}, globalArg(argAt<I>(a))...);
}
template <class B, class... A> __global__ __launch_bounds__(B::NTH) void kPacked(PackHead h, const Args<A...>* args) {
    runPacked<B, A...>(h, args, std::index_sequence_for<A...>{});
}
template <class B, class... A>
__global__ __launch_bounds__(B::NTH) __attribute__((amdgpu_waves_per_eu(3))) void kPacked3(PackHead h, const Args<A...>* args) {
    runPacked<B, A...>(h, args, std::index_sequence_for<A...>{});
}
template <class B, class... A>
__global__ __launch_bounds__(B::NTH) __attribute__((amdgpu_waves_per_eu(4))) void kPacked4(PackHead h, const Args<A...>* args) {
    runPacked<B, A...>(h, args, std::index_sequence_for<A...>{});
}
template <class B, class... A>
__global__ __launch_bounds__(B::NTH) __attribute__((amdgpu_waves_per_eu(5))) void kPacked5(PackHead h, const Args<A...>* args) {
    runPacked<B, A...>(h, args, std::index_sequence_for<A...>{});
}

template <class B, class... A> void launchPacked(const PackHead& h, const void* args, uint32_t blocks, hipStream_t s) {
    const auto* a = static_cast<const Args<A...>*>(args);
    if constexpr (B::WPE == 5) kPacked5<B, A...><<<blocks, B::NTH, 0, s>>>(h, a);
    else if constexpr (B::WPE == 4) kPacked4<B, A...><<<blocks, B::NTH, 0, s>>>(h, a);
    else if constexpr (B::WPE == 3) kPacked3<B, A...><<<blocks, B::NTH, 0, s>>>(h, a);
    else kPacked<B, A...><<<blocks, B::NTH, 0, s>>>(h, a);
}

template <class... A, size_t... I>
void launchOwn(void (*k)(A...), dim3 g, int nth, hipStream_t s, const Args<A...>& a, std::index_sequence<I...>) {
    void* p[sizeof...(A) + 1] = {const_cast<void*>(static_cast<const void*>(&argAt<I>(a)))..., nullptr};
    (void)hipLaunchKernel(reinterpret_cast<const void*>(k), g, dim3(unsigned(nth)), p, 0, s);
}

// launches kernel k (whose body is B) on grid g, or records the launch when this thread records
template <class B, class... A, class... V> void go(void (*k)(A...), dim3 g, hipStream_t s, V&&... v) {
    static_assert(sizeof...(A) == sizeof...(V), "argument count");
    const Args<A...> a = MakeArgs<A...>::make(static_cast<V&&>(v)...);
    if (PackRecorder* rec = packRecorder()) {
        rec->add(reinterpret_cast<const void*>(k), &launchPacked<B, A...>, g.x, g.y, &a, uint32_t(sizeof(a)));
        return;
    }
    launchOwn(k, g, B::NTH, s, a, std::index_sequence_for<A...>{});
}

}  // namespace psx

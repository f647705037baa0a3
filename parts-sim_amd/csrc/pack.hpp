// Packed launches, host side: while a thread has a recorder set, the solve path's launchers
// (kernels.hip, fine.hip) record their launches instead of issuing them. A packed solve records one
// batch of conjugate-gradient iterations per model and merges the records: the same kernel at the
// same point of several models' sequences becomes one launch whose grid is the concatenation of their
// grids, each block finding its model's arguments and its own block index from a small device table.
// Every block then runs exactly the code and arguments it would have run in its own launch.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

namespace psx {

// launches one packed launch can hold
constexpr int PACK_MAX = 64;

// The launches inside a packed launch, passed as a kernel argument (read through the scalar cache):
// launch s owns blocks [first[s], first[s + 1]) of the packed grid, numbered x fastest over its own
// grid's x extent gx[s], then y; first[n] = the packed grid's size.
struct PackHead {
    int n;
    uint32_t first[PACK_MAX + 1];
    uint32_t gx[PACK_MAX];
};

using PackLaunchFn = void (*)(const PackHead& h, const void* args, uint32_t blocks, hipStream_t s);

// a recorded launch
struct PackRec {
    const void* kernel;   // the kernel it stands for: records of the same kernel can share a launch
    PackLaunchFn launch;  // issues the packed form of that kernel
    uint32_t gx, gy;      // its own grid
    uint32_t argBytes;    // its argument block (kernel parameters in order, natural alignment)
    size_t argOff;        // into the recorder's pool, 16-byte aligned
};

struct PackRecorder {
    std::vector<PackRec> recs;
    std::vector<unsigned char> pool;
    void add(const void* kernel, PackLaunchFn fn, uint32_t gx, uint32_t gy, const void* args, uint32_t bytes) {
        const size_t off = (pool.size() + 15) / 16 * 16;
        pool.resize(off + bytes);
        std::memcpy(pool.data() + off, args, bytes);
        recs.push_back(PackRec{kernel, fn, gx, gy, bytes, off});
    }
};

// this thread's recorder (null: launches are issued)
inline PackRecorder*& packRecorder() {
    static thread_local PackRecorder* r = nullptr;
    return r;
}

}  // namespace psx

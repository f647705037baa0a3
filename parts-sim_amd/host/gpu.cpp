// Everything the reference's native/src/gpu/gpu.hpp declares outside its WebGPU section (gpu.hpp:22-31,
// gpu.cpp:250-285), for a build whose GPU solver is the psx library: the device queries over the HIP
// runtime, and the two WebGPU-only queries answered as a build without WebGPU answers them (no 16-bit
// shader floats, no embedded WGSL), so callers such as the airflow solver (airflow.cpp:25) still link.
// The PARTS_SIM_NO_GPU switch keeps its meaning (gpu.cpp:252): set, it reports no GPU.
#include "../gpu/gpu.hpp"

#include <cstdlib>

#include "psx.h"

namespace ps {

bool gpuAvailable() { return !std::getenv("PARTS_SIM_NO_GPU") && psx_available(); }

std::string gpuUnavailableReason() {
    if (gpuAvailable()) return "";
    if (std::getenv("PARTS_SIM_NO_GPU")) return "PARTS_SIM_NO_GPU is set";
    return "no HIP device";
}

std::string gpuName() {
    char buf[256] = {0};
    if (psx_device_name(buf, sizeof buf) != PSX_OK) return "";
    return buf;
}

bool gpuShaderF16() { return false; }

const char* wgslSource(const std::string&) { return ""; }

}  // namespace ps

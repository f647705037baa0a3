// The device-availability queries of the reference's native/src/gpu/gpu.hpp (gpu.hpp:22-26,
// gpu.cpp:250-285) over the HIP runtime, for a build whose GPU solver is the psx library. The
// PARTS_SIM_NO_GPU switch keeps its meaning (gpu.cpp:252): set, it reports no GPU.
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

}  // namespace ps

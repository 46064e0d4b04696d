// Streaming kernels (x-marching with a register queue along the slow axis).
#include "fw_internal.h"

namespace fw {

hipError_t launch_velocity_simple(const DevSim& s, hipStream_t st);
hipError_t launch_pressure_simple(const DevSim& s, hipStream_t st);

bool streaming_supported(const DevSim& s) { (void)s; return false; }

hipError_t launch_velocity(const DevSim& s, hipStream_t st) { return launch_velocity_simple(s, st); }
hipError_t launch_pressure(const DevSim& s, hipStream_t st) { return launch_pressure_simple(s, st); }

}  // namespace fw

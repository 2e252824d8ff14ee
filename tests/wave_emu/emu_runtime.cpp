// State of the CPU execution shim (tests/wave_emu/hip_cpu.hpp).  TEST INFRASTRUCTURE ONLY.
#include "hip_cpu.hpp"

#include <cstdlib>

// launches run at once on the CPU: no graph capture (hip_cpu.hpp)
static const int g_no_graphs = setenv("MI_GRAPHS", "0", 1);

// the device's fast log2 (declared by the C library's math header, not
// exported by it); only the scorer features use it
#include <cmath>
extern "C" float __log2f(float x) noexcept { return log2f(x); }

namespace emu {
thread_local Ctx* g = nullptr;
thread_local int lane = 0;
thread_local const void* t_kernarg = nullptr;
thread_local dim3 t_block, t_thread;
dim3 g_bdim, g_gdim;
YieldBarrier* g_blockbar = nullptr;
}  // namespace emu

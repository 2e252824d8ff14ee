// State of the CPU execution shim (tests/wave_emu/hip_cpu.hpp).  TEST INFRASTRUCTURE ONLY.
#include "hip_cpu.hpp"

namespace emu {
thread_local Ctx* g = nullptr;
thread_local int lane = 0;
thread_local dim3 t_block, t_thread;
dim3 g_bdim, g_gdim;
YieldBarrier* g_blockbar = nullptr;
}  // namespace emu

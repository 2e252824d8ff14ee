// Stand-in for <hip/hip_runtime.h> when the device code is compiled for the
// CPU wave emulator (tests only): maps the intrinsics to wave_emu.hpp.
#pragma once
#include "../../wave_emu.hpp"

// Stand-in for <hip/hip_runtime.h> when the library's HIP sources are
// compiled for the CPU (tests only): see ../../hip_cpu.hpp.
#pragma once
#include "../../hip_cpu.hpp"

// Host-thread test build of the liborx_hip kernels: stands in for the HIP
// runtime header so the kernel sources compile with g++ and run on the CPU
// under AddressSanitizer (tests/cpu_build/hip_cpu.hpp).
#pragma once
#include "../../hip_cpu.hpp"

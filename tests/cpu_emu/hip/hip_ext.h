// CPU stand-in of hip/hip_ext.h for the kernel emulation: hipExtLaunchKernelGGL
// is a no-op macro in hip_runtime.h
#pragma once
#include "hip_runtime.h"

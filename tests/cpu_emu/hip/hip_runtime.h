// CPU stand-in for the HIP runtime header, for the kernel-equivalence tests:
// a kernel body runs as plain C++, one host thread per work-item of a block
// (with a real barrier for __syncthreads) or sequentially for kernels that
// have no barrier.  Only what the library's kernels use is provided.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstddef>
#include <pthread.h>
using std::isfinite;
#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __shared__ static
#define __launch_bounds__(...)
#define __align__(n) alignas(n)
#define __attribute__(x)
struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
extern thread_local dim3 threadIdx, blockIdx;
extern dim3 gridDim, blockDim;
extern pthread_barrier_t emu_barrier;
extern double emu_shfl[1024];
inline void __syncthreads() { pthread_barrier_wait(&emu_barrier); }
// lane exchange within a 64-wide wave through a shared table
template <typename T>
inline T __shfl_down(T v, unsigned d, int width = 64) {
    (void)width;
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    emu_shfl[tid] = (double)v;
    __syncthreads();
    const int lane = tid % 64;
    T r = (lane + (int)d < 64) ? (T)emu_shfl[tid + d] : v;
    __syncthreads();
    return r;
}
typedef int hipError_t;
typedef void *hipStream_t;
typedef void *hipEvent_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1 };
enum { hipFuncAttributeMaxDynamicSharedMemorySize = 8 };
inline hipError_t hipGetLastError() { return 0; }
inline hipError_t hipFuncSetAttribute(const void *, int, int) { return 0; }
#define hipLaunchKernelGGL(...) ((void)0)
#define hipExtLaunchKernelGGL(...) ((void)0)
struct float4 { float x, y, z, w; };
inline float4 make_float4(float a, float b, float c, float d) { return {a, b, c, d}; }
struct double2 { double x, y; };
inline double2 make_double2(double a, double b) { return {a, b}; }
inline void __threadfence_system() {}
#define __HIP_MEMORY_SCOPE_SYSTEM 0
template <typename T, typename U>
inline void __hip_atomic_store(T *p, U v, int, int) { *p = v; }
inline float __builtin_amdgcn_rcpf(float x) { return 1.0f / x; }
// a cache hint only: a plain store on the CPU
#define __builtin_nontemporal_store(v, p) (*(p) = (v))

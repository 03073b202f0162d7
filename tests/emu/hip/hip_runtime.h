// CPU stand-in for the HIP runtime and device built-ins that libfsa's kernels
// use, so the kernel source can be compiled for the host and run under
// AddressSanitizer (test infrastructure only; never part of libfsa.so).
//
// Execution model: a launch runs its workgroups one after another; a
// workgroup runs one OS thread per lane.  __syncthreads is a barrier over the
// workgroup's threads, __shfl_xor and the 16x16x4 MFMA built-ins exchange
// values through a per-wave buffer and barrier, so they must be reached by all
// 64 lanes of a wave together (as on the GPU).  __shared__ arrays are static,
// which is correct because workgroups never run at the same time.  Graph
// capture executes the captured launches immediately and replays do nothing.
#pragma once
#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __builtin_amdgcn_readfirstlane(x) (x)
#define __builtin_amdgcn_s_waitcnt(x) ((void)(x))   // memory operations complete at once here
#define __builtin_amdgcn_sched_barrier(x) ((void)(x))   // instruction scheduling only
#define HIP_SYMBOL(x) (&x)

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct EmuIdx { unsigned x = 0, y = 0, z = 0; };
extern thread_local EmuIdx threadIdx;
extern EmuIdx blockIdx, gridDim, blockDim;

void emu_syncthreads();
double emu_shfl_xor_d(double v, int mask);
double emu_shfl_down_d(double v, unsigned off);
void emu_wave_gather(double a, double b, double* A, double* B);   // all 64 lanes' a and b
inline void __syncthreads() { emu_syncthreads(); }
template <class T>
inline T __shfl_xor(T v, int mask, int = 64) { return (T)emu_shfl_xor_d((double)v, mask); }
template <class T>
inline T __shfl_down(T v, unsigned off, int = 64) { return (T)emu_shfl_down_d((double)v, off); }
inline unsigned atomicAdd(unsigned* p, unsigned v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }

// v_mfma_*_16x16x4 (layouts measured on gfx950): lane l holds A[l%16][l/16] and
// B[l/16][l%16]; accumulator element j of lane l is column l%16 and row
// 4j + l/16 (f64) or 4(l/16) + j (f32)
typedef double emu_d4 __attribute__((ext_vector_type(4)));
typedef float emu_f4 __attribute__((ext_vector_type(4)));
inline emu_d4 __builtin_amdgcn_mfma_f64_16x16x4f64(double a, double b, emu_d4 c, int, int, int) {
    double A[64], B[64];
    emu_wave_gather(a, b, A, B);
    const unsigned l = threadIdx.x % 64;
    for (int j = 0; j < 4; j++) {
        const unsigned row = 4 * j + l / 16, col = l % 16;
        double sum = 0;
        for (int k = 0; k < 4; k++) sum += A[row + 16 * k] * B[col + 16 * k];
        c[j] += sum;
    }
    return c;
}
inline emu_f4 __builtin_amdgcn_mfma_f32_16x16x4f32(float a, float b, emu_f4 c, int, int, int) {
    double A[64], B[64];
    emu_wave_gather(a, b, A, B);
    const unsigned l = threadIdx.x % 64;
    for (int j = 0; j < 4; j++) {
        const unsigned row = 4 * (l / 16) + j, col = l % 16;
        double sum = 0;
        for (int k = 0; k < 4; k++) sum += A[row + 16 * k] * B[col + 16 * k];
        c[j] += (float)sum;
    }
    return c;
}

void emu_launch(dim3 grid, dim3 block, std::function<void()> body);
#define hipLaunchKernelGGL(K, G, B, S, ST, ...) emu_launch(dim3(G), dim3(B), [=]() { K(__VA_ARGS__); })

typedef int hipError_t;
typedef void* hipStream_t;
typedef void* hipEvent_t;
typedef void* hipGraph_t;
typedef void* hipGraphExec_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2 };
enum { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3, hipMemcpyDefault = 4 };
enum { hipHostMallocDefault = 0, hipStreamNonBlocking = 1, hipStreamCaptureModeThreadLocal = 1 };
enum { hipDeviceAttributeMultiprocessorCount = 1, hipFuncAttributeMaxDynamicSharedMemorySize = 8 };
inline hipError_t hipMalloc(void** p, size_t n) { *p = calloc(1, n ? n : 1); return hipSuccess; }
template <class T> inline hipError_t hipMalloc(T** p, size_t n) { return hipMalloc((void**)p, n); }
inline hipError_t hipFree(void* p) { free(p); return hipSuccess; }
inline hipError_t hipHostMalloc(void** p, size_t n, int) { *p = calloc(1, n ? n : 1); return hipSuccess; }
inline hipError_t hipHostFree(void* p) { free(p); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, int, hipStream_t) { memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char* hipGetErrorString(hipError_t) { return "cpu emulation"; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
// one CU: persistent grids hold two workgroups, so each walks several tiles
inline hipError_t hipDeviceGetAttribute(int* v, int attr, int) {
    *v = attr == hipDeviceAttributeMultiprocessorCount ? 1 : 4;
    return hipSuccess;
}
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* b, const void*, int, size_t) { *b = 1; return hipSuccess; }
inline hipError_t hipFuncSetAttribute(const void*, int, int) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = nullptr; return hipSuccess; }
enum { hipEventDisableTiming = 2 };
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { *e = nullptr; return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }   // launches run in order
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 0; return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, int) { *s = nullptr; return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamBeginCapture(hipStream_t, int) { return hipSuccess; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) { *g = (void*)1; return hipSuccess; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t, void*, void*, size_t) { *e = (void*)1; return hipSuccess; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }

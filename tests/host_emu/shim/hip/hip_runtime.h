// Test-only host shim of the HIP names the tracer's device code uses, so that g++ can compile
// that code (ptir_amd/csrc/*.hpp) for the CPU and run it one lane at a time: a wave of one
// lane (ballot = the lane's own predicate), plain loads/stores for atomics, LDS as an array.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __builtin_amdgcn_readfirstlane(x) (x)
#define __builtin_amdgcn_readlane(x, lane) (x)
#define __builtin_amdgcn_fence(order, scope) ((void)0)
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))
#define __align__(n) alignas(n)
struct float4 { float x, y, z, w; };
struct int4 { int x, y, z, w; };
struct uint2 { unsigned x, y; };
inline float4 make_float4(float a, float b, float c, float d) { return {a, b, c, d}; }
inline uint2 make_uint2(unsigned a, unsigned b) { return {a, b}; }
struct Dim3 { unsigned x, y, z; };
extern Dim3 threadIdx, blockIdx, blockDim;
typedef int hipError_t;
enum { hipSuccess = 0 };
typedef void* hipStream_t;
inline hipError_t hipGetLastError() { return 0; }
inline unsigned long long __ballot(int p) { return p ? 1ull : 0ull; }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __popc(unsigned x) { return __builtin_popcount(x); }
inline unsigned __lane_id() { return 0; }
template <class T> inline T __shfl(T v, int) { return v; }
// one-lane waves: no other lane holds a value to exchange
template <class T> inline T __shfl_xor(T, int) { return T(0); }
inline float __uint_as_float(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline float __int_as_float(int32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline uint32_t __float_as_uint(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline int32_t __float_as_int(float f) { int32_t u; std::memcpy(&u, &f, 4); return u; }
inline float atomicAdd(float* p, float v) { float o = *p; *p = o + v; return o; }
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) { auto o = *p; *p = o + v; return o; }
inline uint32_t atomicAdd(uint32_t* p, uint32_t v) { auto o = *p; *p = o + v; return o; }
inline int atomicOr(int* p, int v) { int o = *p; *p = o | v; return o; }
inline uint32_t atomicCAS(uint32_t* p, uint32_t cmp, uint32_t v) { uint32_t o = *p; if (o == cmp) *p = v; return o; }
inline void __syncthreads() {}
inline float rsqrtf(float x) { return 1.0f / std::sqrt(x); }
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }
inline unsigned max(unsigned a, unsigned b) { return a > b ? a : b; }

// CPU stand-ins for the HIP names that the stage-kernel headers
// (sddc_amd/csrc/sd2d_device.hpp, sd2d_stage.hpp, sd2d_stage_mx.hpp) use,
// so that the very same kernel source compiles with g++ and runs on host
// threads under AddressSanitizer (tests/emu/stage_emu.cpp,
// tests/test_emu_stage.py).  Test infrastructure: the product library is
// built by hipcc against the real HIP headers and never sees this file.
//
// Execution model of a launch (tests/emu/emu_runtime.hpp): workgroups run
// one after another; the 128 work-items of a workgroup are 128 host
// threads.  __syncthreads is a barrier over the workgroup's live threads,
// __shfl_xor an exchange among the 64 lanes of a wavefront (and
// __builtin_amdgcn_wave_barrier a barrier over them, since host threads do
// not run in lockstep as a wavefront's lanes do), atomicMax a
// locked read-modify-write.  Dynamic LDS is one buffer shared by the
// running workgroup, poisoned with NaN before every workgroup so that a
// read of LDS nothing wrote shows up in the results.
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <cmath>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __shared__
#define __launch_bounds__(...)
#define __builtin_amdgcn_sched_barrier(x) ((void)0)
#define __builtin_amdgcn_fence(order, scope) ((void)0)
#define __builtin_amdgcn_wave_barrier() emu::wave_sync()
#define __builtin_amdgcn_rcp(x) (1.0 / (x))
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))

struct emu_dim3 {
  unsigned x, y, z;
};

namespace emu {
extern thread_local emu_dim3 tid3, bid3;
void syncthreads();
void wave_sync();
uint64_t shfl_xor_bits(uint64_t bits, int mask);
unsigned long long atomic_max(unsigned long long *p, unsigned long long v);
}  // namespace emu

#define threadIdx (emu::tid3)
#define blockIdx (emu::bid3)

inline void __syncthreads() { emu::syncthreads(); }

template <class T>
inline T __shfl_xor(T v, int mask, int width = 64) {
  static_assert(sizeof(T) <= 8, "exchange of at most 8 bytes");
  (void)width;
  uint64_t b = 0;
  memcpy(&b, &v, sizeof(T));
  b = emu::shfl_xor_bits(b, mask);
  T r;
  memcpy(&r, &b, sizeof(T));
  return r;
}

inline unsigned long long atomicMax(unsigned long long *p, unsigned long long v) {
  return emu::atomic_max(p, v);
}

inline long long __double_as_longlong(double x) {
  long long r;
  memcpy(&r, &x, 8);
  return r;
}
inline double __longlong_as_double(long long x) {
  double r;
  memcpy(&r, &x, 8);
  return r;
}

struct double2 {
  double x, y;
};
struct float2 {
  float x, y;
};
inline float2 make_float2(float a, float b) { return float2{a, b}; }
inline double2 make_double2(double a, double b) { return double2{a, b}; }

// Small device helpers shared by the kernels: wave64 / workgroup reductions.
#pragma once

#include <hip/hip_runtime.h>

#include "common.h"

#define TA_WAVE 64

// Address-space-qualified element types for device functions that are not
// inlined: a callee that only sees generic pointers issues flat memory
// instructions, which wait on both the LDS and the vector memory counters.
typedef __attribute__((address_space(3))) double lds_dbl;
typedef __attribute__((address_space(3))) int lds_int;
typedef __attribute__((address_space(1))) double gbl_dbl;
#define TA_LDS(p) ((lds_dbl*)(p))
#define TA_LDSI(p) ((lds_int*)(p))
#define TA_GBL(p) ((gbl_dbl*)(p))

// Wave index within the workgroup as a wave-uniform (scalar) value, so that
// loops and addresses derived from it stay in scalar registers.
__device__ __forceinline__ int wave_id()
{
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, TA_WAVE);
  return v;
}

// Sum over the 16-lane segment of a wave the lane belongs to (lanes 0-15,
// 16-31, 32-47, 48-63); every lane of the segment gets the result.
__device__ __forceinline__ double seg16_sum(double v)
{
#pragma unroll
  for (int off = 8; off > 0; off >>= 1) v += __shfl_xor(v, off, TA_WAVE);
  return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, TA_WAVE));
  return v;
}

// Workgroup sum; red must hold blockDim.x / 64 doubles.  Every thread gets
// the result.  Contains two barriers.
template <class P>
__device__ __forceinline__ double block_sum(double v, P red)
{
  const int lane = threadIdx.x & (TA_WAVE - 1);
  const int wid = threadIdx.x / TA_WAVE;
  const int nw = blockDim.x / TA_WAVE;
  v = wave_sum(v);
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double r = 0.0;
  for (int w = 0; w < nw; ++w) r += red[w];
  __syncthreads();
  return r;
}

template <class P>
__device__ __forceinline__ double block_max(double v, P red)
{
  const int lane = threadIdx.x & (TA_WAVE - 1);
  const int wid = threadIdx.x / TA_WAVE;
  const int nw = blockDim.x / TA_WAVE;
  v = wave_max(v);
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double r = red[0];
  for (int w = 1; w < nw; ++w) r = fmax(r, red[w]);
  __syncthreads();
  return r;
}

// Exclusive prefix sum of ints over the workgroup (red: blockDim.x/64 ints);
// returns the thread's exclusive prefix and writes the total to *total.
__device__ __forceinline__ int block_exclusive_scan(int v, int* red, int* total)
{
  const int lane = threadIdx.x & (TA_WAVE - 1);
  const int wid = threadIdx.x / TA_WAVE;
  const int nw = blockDim.x / TA_WAVE;
  int incl = v;
#pragma unroll
  for (int off = 1; off < TA_WAVE; off <<= 1) {
    int t = __shfl_up(incl, off, TA_WAVE);
    if (lane >= off) incl += t;
  }
  if (lane == TA_WAVE - 1) red[wid] = incl;
  __syncthreads();
  int base = 0, tot = 0;
  for (int w = 0; w < nw; ++w) {
    if (w < wid) base += red[w];
    tot += red[w];
  }
  __syncthreads();
  *total = tot;
  return base + incl - v;
}

// Per-phase wall-clock accumulator of one workgroup (thread 0 only); used by
// the SCF kernel when profiling is enabled.
struct PhaseClock {
  long long acc[TA_PROF_SLOTS];
  long long last;
  bool on;
  __device__ void start(bool enable)
  {
    on = enable && threadIdx.x == 0;
    for (int i = 0; i < TA_PROF_SLOTS; ++i) acc[i] = 0;
    if (on) last = wall_clock64();
  }
  __device__ void lap(int slot)
  {
    if (on) {
      const long long t = wall_clock64();
      acc[slot] += t - last;
      last = t;
    }
  }
};

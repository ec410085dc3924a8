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

// A wave-uniform pointer or integer moved into scalar registers.  Arguments
// of non-inlined device functions arrive in vector registers, so the compiler
// cannot see that they are uniform; re-reading them through readfirstlane
// lets memory accesses use a scalar base plus a 32-bit vector offset instead
// of a 64-bit address per lane.
template <class T>
__device__ __forceinline__ T* sgpr_ptr(T* p)
{
  const unsigned long long b = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
  return (T*)(((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ int sgpr_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Element idx of a global array addressed as base + 32-bit byte offset (the
// scalar-base addressing mode when base is uniform); idx < 2^29.
typedef __attribute__((address_space(1))) char gbl_char;
__device__ __forceinline__ double gld(const gbl_dbl* base, unsigned idx)
{
  return *(const gbl_dbl*)((const gbl_char*)base + (idx << 3));
}
__device__ __forceinline__ void gst(gbl_dbl* base, unsigned idx, double v)
{
  *(gbl_dbl*)((gbl_char*)base + (idx << 3)) = v;
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

// Lane exchanges on the VALU (no LDS crossbar): DPP within a 16-lane row,
// v_permlane16/32_swap across rows.  Every lane of a reduction gets a
// bit-identical sum (each step adds the same two operands in every lane).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, true);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// sum over the lane's 16-lane row (quad_perm xor 1, xor 2, row_half_mirror,
// row_mirror)
__device__ __forceinline__ double row16_sum(double v)
{
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x141>(v);
  v += dpp_f64<0x140>(v);
  return v;
}

// v(lane) + v(lane ^ 16)
__device__ __forceinline__ double xor16_sum(double v)
{
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)b, (unsigned)b, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false,
                                                   false);
  const double a = __builtin_bit_cast(double, ((unsigned long long)hi[0] << 32) | lo[0]);
  const double c = __builtin_bit_cast(double, ((unsigned long long)hi[1] << 32) | lo[1]);
  return a + c;
}

// v(lane) + v(lane ^ 32)
__device__ __forceinline__ double xor32_sum(double v)
{
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)b, (unsigned)b, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false,
                                                   false);
  const double a = __builtin_bit_cast(double, ((unsigned long long)hi[0] << 32) | lo[0]);
  const double c = __builtin_bit_cast(double, ((unsigned long long)hi[1] << 32) | lo[1]);
  return a + c;
}

// Orders the LDS accesses of one wave around this point (the hardware runs a
// wave's LDS instructions in order; this keeps the compiler from moving them
// across), so that lanes read what other lanes of the wave wrote before it.
__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// lane 0's value in every lane (scalar broadcast)
__device__ __forceinline__ double bcast_lane0(double v)
{
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// wave64 sum on the VALU
__device__ __forceinline__ double wave_sum_dpp(double v) { return xor32_sum(xor16_sum(row16_sum(v))); }

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

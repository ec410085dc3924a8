// Two-stage reduction of a symmetric matrix to tridiagonal form, and the
// matching back-transformations, for the global-memory eigensolver (one
// workgroup of NT = 512 threads per molecule).  It replaces the one-stage
// tridiagonalize_blocked() of eig_big.h in the reduction step of sygvd
// (src/lapack/sygvd.cu:223-313) with far less memory traffic:
//
//   stage 1   dense -> band of half-width BW = 16, in panels of BW columns.
//             The strip of a panel is held in registers; a Householder QR of
//             its rows below the band gives Q = I - V T V^T, and the trailing
//             matrix takes Q^T A2 Q = A2 - V W^T - W V^T with X = A2 V T,
//             W = X - V (T^T V^T X) / 2.  Panel p's update is applied while
//             the trailing lower triangle is read for panel p+1's product
//             X = A2 V, so each panel reads and writes that triangle once
//             (the one-stage reduction reads it once per column).  The
//             products run on the FP64 matrix cores.
//   stage 2   band -> tridiagonal by bulge chasing: sweep s annihilates
//             column s with a reflector of length BW and chases the bulge it
//             creates down the band, one BW x BW block step (task) at a time.
//             A wave runs the tasks of one sweep (three blocks in registers);
//             sweep s's task t waits for sweep s - 1's task t + 2 through a
//             progress counter in LDS, which keeps every pair of tasks that
//             share an element in sweep order and tasks running at the same
//             time on disjoint elements.  The band stays in A's lower
//             triangle, with room for the bulges (2 BW - 1 below the
//             diagonal; stage 1 leaves zeros there).
//   vectors   Z = Q1 Q2 Y: the stage-2 reflectors in blocks of 16 sweeps at
//             one task (block reflectors of 31 rows on the matrix cores, each
//             wave carrying its column strips of Y through the blocks in
//             order), then the stage-1 reflectors one panel of 16 at a time
//             in compact WY form with the T factor stage 1 saved.
//
// Global layouts (ld = leading dimension, rows r global):
//   V1[k * ld + r]           stage-1 reflector k = c0 + j (panel c0, column
//                            j): zero above row k + BW, 1 there; a reflector
//                            with tau1[k] = 0 is the identity (zero column)
//   V2[(first(s) + t) * 16]  stage-2 reflector of sweep s, task t: entry 0 is
//                            tau, entries 1..15 the vector below its leading 1
//                            (first(s) = chase_first: records of consecutive
//                            sweeps are contiguous)
//   W[k * ld + r], X[k * ld + r]   panel work vectors (16 columns each)
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "device_util.h"
#include "eig_big.h"

namespace band {

constexpr int BW = 16;  // half-bandwidth after stage 1
// The panel strip of stage 1 is held in registers, QB blocks of 16 rows per
// wave (8 QB BW rows): QB = 4 up to order 512, QB = 8 up to 1024.
constexpr int band_qb(int n) { return n <= 8 * 4 * BW ? 4 : 8; }

using big::ta_double4;

// stage-2 tasks of sweep s (s <= n - 3): task 0, then one per further block
// that still holds at least two rows
__device__ __forceinline__ int chase_tasks(int n, int s) { return 1 + (n - 3 - s) / BW; }

// sum_{u=0}^{x} floor(u / BW) (0 for x < 0)
__device__ __forceinline__ int floor_sum(int x)
{
  if (x < 0) return 0;
  const int q = x / BW, r = x - q * BW;
  return (BW / 2) * q * (q - 1) + q * (r + 1);
}

// index of the first record of sweep s in the compact stage-2 record array
// (records of consecutive sweeps are contiguous; record (s, t) = first + t)
__device__ __forceinline__ int chase_first(int n, int s)
{
  return s + floor_sum(n - 3) - floor_sum(n - 3 - s);
}

// number of stage-1 reflectors (16 per panel; panels while n - c0 - BW >= 2)
__device__ __forceinline__ int band_reflectors(int n)
{
  return n - BW >= 2 ? ((n - BW - 2) / BW + 1) * BW : 0;
}

__device__ __forceinline__ ta_double4 mfma(double a, double b, ta_double4 c)
{
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ double pick4(ta_double4 v, int i)
{
  return i == 0 ? v[0] : (i == 1 ? v[1] : (i == 2 ? v[2] : v[3]));
}

// 64 x 64 LDS tile with rows padded to 65 doubles: row and column fragments
// of 16 lanes spread over the banks, and every access of a lane is one base
// address plus a compile-time offset
constexpr int TS = 65;
constexpr int TILE_DOUBLES = 64 * TS;
__device__ __forceinline__ int tix(int r, int c) { return r * TS + c; }

// ------------------------------------------------------ stage 1: one pass
// Trailing matrix A2 = A[r2:, r2:] (m = n - r2), lower 64 x 64 tiles in row
// order.  With upd: A2 -= Vp Wp^T + Wp Vp^T on the lower triangle (stored).
// With want_x: X = A2 Vc (A2 after the update, 16 columns), from the lower
// tiles only: tile (I, J) adds A_IJ V_J to X_I and, below the diagonal,
// A_IJ^T V_I to X_J.  Waves 0-3 keep X_I in registers over a tile row and
// store it at its end; waves 4-7 add the transposed contributions to rows
// stored earlier.  tile: TILE_DOUBLES of LDS.
template <int NT>
__device__ __noinline__ void band_pass(gbl_dbl* A, int n, int ld, int r2, const gbl_dbl* Vp,
                                       const gbl_dbl* Wp, bool upd, const gbl_dbl* Vc,
                                       gbl_dbl* X, bool want_x, lds_dbl* tile)
{
  static_assert(NT == 512, "tiling assumes 8 waves");
  A = sgpr_ptr(A);
  Vp = sgpr_ptr(Vp);
  Wp = sgpr_ptr(Wp);
  Vc = sgpr_ptr(Vc);
  X = sgpr_ptr(X);
  n = sgpr_int(n);
  ld = sgpr_int(ld);
  r2 = sgpr_int(r2);
  const int lane = threadIdx.x & 63, w = wave_id();
  const int lr = lane & 15, lk = lane >> 4;
  const int bi = w >> 1, bj = (w & 1) * 2;
  const int m = n - r2, nt = (m + 63) >> 6;
  // loads use clamped (in-range) indices and zero the value afterwards, so
  // they are branch-free; row-major offsets fit in 32 bits (ld <= 1024)
  gbl_dbl* A2 = sgpr_ptr(A + (int64_t)r2 * ld + r2);
  const gbl_dbl* Vr = want_x ? Vc + r2 : nullptr;  // V[row][lr] = Vr[lr * ld + row]
  for (int I = 0; I < nt; ++I) {
    const int I0 = I * 64;
    double au[8], vb[16];
    {
      const int row = I0 + 16 * bi + lr, rc = min(row, m - 1);
#pragma unroll
      for (int kk = 0; kk < 8; ++kk) {
        const int k = lk + 4 * kk;
        double a = 0.0;
        if (upd) a = k < BW ? gld(Vp, k * ld + r2 + rc) : gld(Wp, (k - BW) * ld + r2 + rc);
        au[kk] = row < m ? -a : 0.0;
      }
    }
    if (want_x && w >= 4) {
#pragma unroll
      for (int kk = 0; kk < 16; ++kk) {
        const int row = I0 + 4 * kk + lk;
        const double v = gld(Vr, lr * ld + min(row, m - 1));
        vb[kk] = row < m ? v : 0.0;
      }
    }
    ta_double4 xi = {0.0, 0.0, 0.0, 0.0};
    for (int J = 0; J <= I; ++J) {
      const int J0 = J * 64;
      // each wave loads (and later stores) only the lower-triangle elements
      // of its own blocks: the upper half of a diagonal tile is another
      // wave's lower half, which that wave may already have updated and
      // stored, so it is taken from the tile in LDS instead (below)
      ta_double4 acc[2];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int row = I0 + 16 * bi + 4 * i + lk, col = J0 + 16 * (bj + h) + lr;
          const int rc = min(row, m - 1), cc = min(min(col, row), m - 1);
          const double v = gld(A2, rc * ld + cc);
          acc[h][i] = (row < m && col <= row) ? v : 0.0;
        }
      if (upd) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int col = J0 + 16 * (bj + h) + lr, cc = min(col, m - 1);
          double bo[8];
#pragma unroll
          for (int kk = 0; kk < 8; ++kk) {
            const int k = lk + 4 * kk;
            const double b = k < BW ? gld(Wp, k * ld + r2 + cc) : gld(Vp, (k - BW) * ld + r2 + cc);
            bo[kk] = col < m ? b : 0.0;
          }
#pragma unroll
          for (int kk = 0; kk < 8; ++kk) acc[h] = mfma(au[kk], bo[kk], acc[h]);
        }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int row = I0 + 16 * bi + 4 * i + lk, col = J0 + 16 * (bj + h) + lr;
            if (row < m && col <= row) gst(A2, row * ld + col, acc[h][i]);
          }
      }
      if (want_x) {
        // the full symmetric tile: lower elements at their place and, in a
        // diagonal tile, mirrored into the upper half
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int tr = 16 * bi + 4 * i + lk, tc = 16 * (bj + h) + lr;
            if (I != J || tr >= tc) tile[tix(tr, tc)] = acc[h][i];
            if (I == J && tr > tc) tile[tix(tc, tr)] = acc[h][i];
          }
        __syncthreads();
        if (w < 4) {
          // V rows of the tile column
#pragma unroll
          for (int kk = 0; kk < 16; ++kk) {
            const int col = J0 + 4 * kk + lk;
            const double v = gld(Vr, lr * ld + min(col, m - 1));
            vb[kk] = col < m ? v : 0.0;
          }
#pragma unroll
          for (int kk = 0; kk < 16; ++kk)
            xi = mfma(tile[tix(16 * w + lr, 4 * kk + lk)], vb[kk], xi);
        } else if (I != J) {
          const int wb = w - 4;
          ta_double4 xj = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int kk = 0; kk < 16; ++kk)
            xj = mfma(tile[tix(4 * kk + lk, 16 * wb + lr)], vb[kk], xj);
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int row = J0 + 16 * wb + 4 * i + lk;
            if (row < m) gst(X, lr * ld + r2 + row, gld(X, lr * ld + r2 + row) + xj[i]);
          }
        }
        __syncthreads();
      }
    }
    if (want_x && w < 4) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = I0 + 16 * w + 4 * i + lk;
        if (row < m) gst(X, lr * ld + r2 + row, xi[i]);
      }
    }
  }
  __syncthreads();
}

// Sum of the eight waves' 16 x 16 partials (D layout) into out[16 * 16];
// part: 2048 doubles of LDS.  Two barriers.
__device__ __forceinline__ void reduce_partials(ta_double4 g, lds_dbl* part, lds_dbl* out)
{
  const int lane = threadIdx.x & 63, w = wave_id();
  const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int i = 0; i < 4; ++i) part[w * 256 + (4 * i + lk) * 16 + lr] = g[i];
  __syncthreads();
  if (threadIdx.x < 256) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) s += part[q * 256 + threadIdx.x];
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

// ------------------------------------------------------ stage 1: panel
// Panel at column c0 (m = n - c0 - BW >= 2 rows below the band): the strip
// (rows c0.., columns c0..c0+15) is loaded into registers (MFMA D layout of
// 16-row blocks, block b on wave b % 8) and brought up to date with the
// previous panel's update (Vp, Wp; upd); its diagonal block is stored as band
// data; a Householder QR (dlarfg per column) of its rows below the band gives
// the reflectors (V1 at row c0, tau1) and R, which goes into A with zeros in
// the bulge region below it; T of the block reflector I - V T V^T ends in
// sm[0..256).  tile: TILE_DOUBLES of LDS (reductions); sm: 560 doubles.
template <int NT, int QB>
__device__ __noinline__ void band_panel_qr(gbl_dbl* A, int n, int ld, int c0, gbl_dbl* V1,
                                           gbl_dbl* tau1, gbl_dbl* Tp, const gbl_dbl* Vp,
                                           const gbl_dbl* Wp, bool upd, lds_dbl* tile,
                                           lds_dbl* sm, lds_dbl* tprof)
{
  const int tid = threadIdx.x, lane = tid & 63, w = wave_id();
  const int lr = lane & 15, lk = lane >> 4;
  const bool clk_on = tprof != nullptr && tid == 0;
  long long clk_last = clk_on ? wall_clock64() : 0;
  auto lap = [&](int slot) {
    if (clk_on) {
      const long long t = wall_clock64();
      tprof[slot] += (double)(t - clk_last);
      clk_last = t;
    }
  };
  lds_dbl* Tm = sm;          // T of the panel, [k][c]
  lds_dbl* Gs = sm + 256;    // reduced V^T V
  lds_dbl* sc = sm + 512;    // pivot (0), tau of the panel's reflectors (1..16)
  lds_dbl* red2 = tile;  // QR: partial Gram rows [2][8][16], pivot rows [2][16]
  const int ms = n - c0, m = ms - BW, r2 = c0 + BW;
  gbl_dbl* Vc = V1 + (int64_t)c0 * ld;
  // (a) strip into registers with the previous panel's update
  ta_double4 x[QB];
  double bop[8];
#pragma unroll
  for (int kk = 0; kk < 8; ++kk) {
    const int k = lk + 4 * kk;
    double b = 0.0;
    if (upd) b = k < BW ? Wp[(int64_t)k * ld + c0 + lr] : Vp[(int64_t)(k - BW) * ld + c0 + lr];
    bop[kk] = b;
  }
#pragma unroll
  for (int q = 0; q < QB; ++q) {
    const int rb = 16 * (w + 8 * q);
    ta_double4 acc;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int rl = rb + 4 * i + lk;
      double v = 0.0;
      if (rl < ms) {
        const int gr = c0 + rl, gc = c0 + lr;
        v = rl >= lr ? A[(int64_t)gr * ld + gc] : A[(int64_t)gc * ld + gr];
      }
      acc[i] = v;
    }
    if (upd && rb < ms) {
      const int gr = c0 + rb + lr;
#pragma unroll
      for (int kk = 0; kk < 8; ++kk) {
        const int k = lk + 4 * kk;
        double a = 0.0;
        if (rb + lr < ms) a = k < BW ? Vp[(int64_t)k * ld + gr] : Wp[(int64_t)(k - BW) * ld + gr];
        acc = mfma(-a, bop[kk], acc);
      }
    }
    x[q] = acc;
  }
  if (w == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int rl = 4 * i + lk;
      if (rl >= lr) A[(int64_t)(c0 + rl) * ld + c0 + lr] = x[0][i];
    }
  }
  lap(0);
  // (b) QR.  The lane's rows are rb(q) + 4 i + lk with rb(q) = 16 (w + 8 q)
  // wave-uniform; column j lives in the lanes lr == j of each 16-lane row.
  const int kref = m < BW ? m : BW;
  const int src_row = lane & 48;
  // One barrier per column: the waves' partial Gram row g_c = sum over the
  // rows below the pivot of x_j x_c (g_j = the squared norm below the pivot)
  // and the pivot row go through LDS, double-buffered by column parity; the
  // reflector v (1 at the pivot, x_j * scale below) then has
  // v . x_c = x_c[piv] + scale g_c.  Column j's broadcast along each 16-lane
  // row (all shuffles issued before the first use) serves both the Gram row
  // and the update.
  // The broadcast is issued QH blocks at a time; with QB > QH (the deep
  // strip) it is issued again for the update instead of being held across
  // the barrier, which keeps the strip and the broadcast within the
  // register budget.
  constexpr int QH = QB > 4 ? 4 : QB;
  for (int j = 0; j < kref; ++j) {
    const int piv = BW + j;
    lds_dbl* R = red2 + (j & 1) * 8 * 16;
    lds_dbl* P = red2 + 2 * 8 * 16 + (j & 1) * 16;
    double xs[QH][4];
    double g = 0.0;
#pragma unroll
    for (int h = 0; h < QB; h += QH) {
#pragma unroll
      for (int q = 0; q < QH; ++q)
#pragma unroll
        for (int i = 0; i < 4; ++i) xs[q][i] = __shfl(x[h + q][i], src_row | j, 64);
#pragma unroll
      for (int q = 0; q < QH; ++q)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int rl = 16 * (w + 8 * (h + q)) + 4 * i + lk;
          g += (rl > piv && rl < ms) ? xs[q][i] * x[h + q][i] : 0.0;
        }
    }
    g = xor32_sum(xor16_sum(g));
    if (lk == 0) R[w * 16 + lr] = g;
    if (w == 1 && lk == (j & 3)) P[lr] = pick4(x[0], j >> 2);
    __syncthreads();
    double gc = 0.0, gj = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      gc += R[q * 16 + lr];
      gj += R[q * 16 + j];
    }
    const double alpha = P[j];
    double beta, tau, scale;
    if (gj == 0.0) {
      beta = alpha;
      tau = 0.0;
      scale = 0.0;
    } else {
      beta = -copysign(sqrt(alpha * alpha + gj), alpha);
      tau = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    const double f = lr > j ? tau * (P[lr] + scale * gc) : 0.0;
    if (tid == 0) sc[1 + j] = tau;
    // column j <- (beta at the pivot, reflector below); columns c > j take
    // x_c -= f v
#pragma unroll
    for (int h = 0; h < QB; h += QH) {
      if (QB > QH) {
#pragma unroll
        for (int q = 0; q < QH; ++q)
#pragma unroll
          for (int i = 0; i < 4; ++i) xs[q][i] = __shfl(x[h + q][i], src_row | j, 64);
      }
#pragma unroll
      for (int q = 0; q < QH; ++q)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int rl = 16 * (w + 8 * (h + q)) + 4 * i + lk;
          const bool below = rl > piv && rl < ms;
          const double v = rl == piv ? 1.0 : (below ? xs[q][i] * scale : 0.0);
          const double xo = x[h + q][i];
          x[h + q][i] = lr == j ? (rl == piv ? beta : (below ? v : xo)) : xo - f * v;
        }
    }
  }
  __syncthreads();
  lap(1);
  // (c) R and the zeroed bulge region into A, the reflectors into V1 (zero
  // columns for identities); x keeps V for T
  const double tcol = lr < kref ? sc[1 + lr] : 0.0;
  if (tid < BW) tau1[c0 + tid] = tid < kref ? sc[1 + tid] : 0.0;
#pragma unroll
  for (int q = 0; q < QB; ++q)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int rl = 16 * (w + 8 * q) + 4 * i + lk;
      double vv = 0.0;
      if (rl >= BW && rl < ms) {
        const int pr = rl - BW;
        const double xv = x[q][i];
        if (pr <= lr) A[(int64_t)(r2 + pr) * ld + c0 + lr] = xv;
        else if (pr - lr < BW) A[(int64_t)(r2 + pr) * ld + c0 + lr] = 0.0;
        if (tcol != 0.0 && pr >= lr) vv = pr == lr ? 1.0 : xv;
        Vc[(int64_t)lr * ld + r2 + pr] = vv;
      }
      x[q][i] = vv;
    }
  // (d) T from T^-1 = diag(1/tau) + strict_upper(V^T V)
  ta_double4 g = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < QB; ++q)
#pragma unroll
    for (int i = 0; i < 4; ++i) g = mfma(x[q][i], x[q][i], g);
  reduce_partials(g, tile, Gs);
  if (tid < BW) {
    const int c = tid;
    for (int i = BW - 1; i >= 0; --i) {
      const double ti = sc[1 + i];
      const double dii = (i < kref && ti != 0.0) ? 1.0 / ti : 1.0;
      double t = 0.0;
      if (i == c) {
        t = 1.0 / dii;
      } else if (i < c) {
        double acc = 0.0;
        for (int k = i + 1; k <= c; ++k) acc += Gs[i * BW + k] * Tm[k * BW + c];
        t = -acc / dii;
      }
      Tm[i * BW + c] = t;
    }
  }
  __syncthreads();
  if (tid < BW * BW) Tp[tid] = Tm[tid];  // kept for the back-transformation
  lap(2);
}

// W = X T - V (T^T V^T X T) / 2 = X T + V M, M = -(T^T Z T) / 2, Z = V^T X,
// for the panel at c0 (rows c0 + BW .. n-1 of V, X and W).  T in sm[0..256);
// tile: 4096 doubles of LDS; sm[256..512) scratch.
template <int NT, int QB>
__device__ __noinline__ void band_panel_w(int n, int ld, int c0, const gbl_dbl* Vc,
                                          const gbl_dbl* X, gbl_dbl* Wc, lds_dbl* tile,
                                          lds_dbl* sm)
{
  const int tid = threadIdx.x, lane = tid & 63, w = wave_id();
  const int lr = lane & 15, lk = lane >> 4;
  const int ms = n - c0;
  lds_dbl* Tm = sm;
  lds_dbl* Zs = sm + 256;
  lds_dbl* Ps = tile + 2048;
  lds_dbl* Ms = tile + 2304;
  // Z = V^T X from D-layout blocks
  ta_double4 z = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < QB; ++q) {
    const int rb = 16 * (w + 8 * q);
    if (rb < BW || rb >= ms) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int rl = rb + 4 * i + lk;
      const int off = lr * ld + c0 + rl;
      const double v = rl < ms ? Vc[off] : 0.0;
      const double xv = rl < ms ? X[off] : 0.0;
      z = mfma(v, xv, z);
    }
  }
  reduce_partials(z, tile, Zs);
  if (tid < 256) {
    const int a = tid >> 4, c = tid & 15;
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < BW; ++k) t += Zs[a * BW + k] * Tm[k * BW + c];
    Ps[tid] = t;
  }
  __syncthreads();
  if (tid < 256) {
    const int a = tid >> 4, c = tid & 15;
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < BW; ++k) t += Tm[k * BW + a] * Ps[k * BW + c];
    Ms[tid] = -0.5 * t;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < QB; ++q) {
    const int rb = 16 * (w + 8 * q);
    if (rb < BW || rb >= ms) continue;
    ta_double4 acc = {0.0, 0.0, 0.0, 0.0};
    const bool in = rb + lr < ms;
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const int k = 4 * kk + lk;
      const int off = k * ld + c0 + rb + lr;
      const double xa = in ? X[off] : 0.0;
      const double va = in ? Vc[off] : 0.0;
      acc = mfma(xa, Tm[k * BW + lr], acc);
      acc = mfma(va, Ms[k * BW + lr], acc);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int rl = rb + 4 * i + lk;
      if (rl < ms) Wc[lr * ld + c0 + rl] = acc[i];
    }
  }
  __syncthreads();
}

// ------------------------------------------------------ stage 1: driver
// A (n x n, lower triangle valid, n <= 8 QB BW) -> band of half-width BW in its
// lower triangle, zeros 17..31 below the diagonal.  V1, tau1: stage-1
// reflectors (layout above); T1: the T factor of each panel (16 x 16, T[i][c]
// at 256 p + 16 i + c); Wb: 2 x 16 x ld, X: 16 x ld doubles of global
// scratch; tile: TILE_DOUBLES of LDS; sm: 560 doubles of LDS; tprof: optional
// step clocks (strip, QR, store + T, trailing pass, W).
template <int NT, int QB>
__device__ __noinline__ void band_reduce(gbl_dbl* A, int n, int ld, gbl_dbl* V1, gbl_dbl* tau1,
                                         gbl_dbl* T1, gbl_dbl* Wb, gbl_dbl* X, lds_dbl* tile,
                                         lds_dbl* sm, lds_dbl* tprof)
{
  static_assert(NT == 512, "strip layout assumes 8 waves");
  const bool clk_on = tprof != nullptr && threadIdx.x == 0;
  long long clk_last = 0;
  auto lap = [&](int slot) {
    if (clk_on) {
      const long long t = wall_clock64();
      tprof[slot] += (double)(t - clk_last);
      clk_last = t;
    }
  };
  int p = 0;
  for (int c0 = 0; n - c0 - BW >= 2; c0 += BW, ++p) {
    const int r2 = c0 + BW;
    const bool upd = c0 > 0;
    const gbl_dbl* Vp = V1 + (int64_t)(upd ? c0 - BW : 0) * ld;
    const gbl_dbl* Wp = Wb + (int64_t)((p + 1) & 1) * BW * ld;
    gbl_dbl* Wc = Wb + (int64_t)(p & 1) * BW * ld;
    const gbl_dbl* Vc = V1 + (int64_t)c0 * ld;
    band_panel_qr<NT, QB>(A, n, ld, c0, V1, tau1, T1 + BW * BW * p, Vp, Wp, upd, tile, sm, tprof);
    if (clk_on) clk_last = wall_clock64();
    band_pass<NT>(A, n, ld, r2, Vp, Wp, upd, Vc, X, true, tile);
    lap(3);
    band_panel_w<NT, QB>(n, ld, c0, Vc, X, Wc, tile, sm);
    lap(4);
  }
  // the last panel's update of the remaining trailing block
  if (p > 0) {
    const int c0 = (p - 1) * BW, r2 = c0 + BW;
    if (clk_on) clk_last = wall_clock64();
    band_pass<NT>(A, n, ld, r2, V1 + (int64_t)c0 * ld, Wb + (int64_t)((p - 1) & 1) * BW * ld,
                  true, nullptr, nullptr, false, tile);
    lap(3);
  }
}

// ------------------------------------------------------ stage 2
// Task (sweep s, task t) on the band in A, by one wave.  Rows of the task
// lo .. hi-1 (L of them), the annihilated column c (the left block's Lc
// columns start there), the block below hi .. bh-1 (Lr rows).  Lane (i, g):
// row i = lane % 16 of each 16 x 16 block, columns 4 g .. 4 g + 3.
struct ChaseGeom {
  int lo, hi, L, c, Lc, Lr;
};

__device__ __forceinline__ ChaseGeom chase_geom(int n, int s, int t)
{
  ChaseGeom q;
  q.lo = s + 1 + BW * t;
  q.hi = min(s + BW * (t + 1), n - 1) + 1;
  q.L = q.hi - q.lo;
  q.c = t == 0 ? s : q.lo - BW;
  q.Lc = t == 0 ? 1 : BW;
  q.Lr = min(q.hi + BW, n) - q.hi;
  return q;
}

// Storage of the band during the chase: A's lower triangle in global memory
// (row r, column c at r * ld + c), or an LDS copy of its 32 lower diagonals
// (the band of half-width 16 and the bulge room below it; element (r, c),
// 0 <= r - c < 32, at r * 32 + r - c).
struct BandGbl {
  gbl_dbl* A;
  int ld;
  __device__ __forceinline__ double get(int r, int c) const { return gld(A, r * ld + c); }
  __device__ __forceinline__ void put(int r, int c, double v) const { gst(A, r * ld + c, v); }
};
struct BandLds {
  lds_dbl* B;
  __device__ __forceinline__ double get(int r, int c) const { return B[r * 32 + (r - c)]; }
  __device__ __forceinline__ void put(int r, int c, double v) const { B[r * 32 + (r - c)] = v; }
};

// the task's diagonal block and block below (and, with left, its left block)
template <class BS>
__device__ __forceinline__ void chase_load(const BS& A, int n, int s, int t, bool left,
                                           double* lb, double* dd, double* rr)
{
  const int lane = threadIdx.x & 63;
  const int i = lane & 15, g = lane >> 4;
  const ChaseGeom q = chase_geom(n, s, t);
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) {
    const int j = 4 * g + jj;
    if (left) lb[jj] = (i < q.L && j < q.Lc) ? A.get(q.lo + i, q.c + j) : 0.0;
    const int a = i > j ? i : j, b = i > j ? j : i;
    dd[jj] = (i < q.L && j < q.L) ? A.get(q.lo + a, q.lo + b) : 0.0;
    rr[jj] = (i < q.Lr && j < q.L) ? A.get(q.hi + i, q.lo + j) : 0.0;
  }
}

// The task on its three blocks in registers: the reflector from column 0 of
// the left block, applied from the left to the left block, two-sided to the
// diagonal block and from the right to the block below; the blocks and the
// reflector record are stored, rr keeps the updated block below (the next
// task's left block).
template <class BS>
__device__ __forceinline__ void chase_apply(const BS& A, int n, int s, int t, double* lb,
                                            double* dd, double* rr, gbl_dbl* V2)
{
  const int lane = threadIdx.x & 63;
  const int i = lane & 15, g = lane >> 4;
  const ChaseGeom q = chase_geom(n, s, t);
  const double x = __shfl(lb[0], i, 64);
  const double ss = row16_sum(i >= 1 ? x * x : 0.0);
  const double alpha = bcast_lane0(x);
  double beta, tau, scale;
  if (ss == 0.0) {
    beta = alpha;
    tau = 0.0;
    scale = 0.0;
  } else {
    beta = -copysign(sqrt(alpha * alpha + ss), alpha);
    tau = (beta - alpha) / beta;
    scale = 1.0 / (alpha - beta);
  }
  const double vi = i == 0 ? 1.0 : (i < q.L ? x * scale : 0.0);
  double vj[4];
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) vj[jj] = __shfl(vi, 4 * g + jj, 64);
  if (t > 0) {
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      const double pcol = row16_sum(vi * lb[jj]);
      lb[jj] -= tau * vi * pcol;
    }
  }
  if (g == 0) lb[0] = i == 0 ? beta : 0.0;
  // H D H = D - v w^T - w v^T, w = y - (tau/2)(y.v) v, y = tau D v
  double part = 0.0;
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) part += dd[jj] * vj[jj];
  part = xor32_sum(xor16_sum(part));
  const double y = tau * part;
  const double yv = row16_sum(y * vi);
  const double wv = y - 0.5 * tau * yv * vi;
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) {
    const double wj = __shfl(wv, 4 * g + jj, 64);
    dd[jj] -= vi * wj + wv * vj[jj];
  }
  double qr = 0.0;
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) qr += rr[jj] * vj[jj];
  qr = xor32_sum(xor16_sum(qr));
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) rr[jj] -= tau * qr * vj[jj];
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) {
    const int j = 4 * g + jj;
    if (i < q.L && j < q.Lc) A.put(q.lo + i, q.c + j, lb[jj]);
    if (i < q.L && j <= i) A.put(q.lo + i, q.lo + j, dd[jj]);
    if (i < q.Lr && j < q.L) A.put(q.hi + i, q.lo + j, rr[jj]);
  }
  if (g == 0) gst(V2, (chase_first(n, s) + t) * BW + i, i == 0 ? tau : vi);
}

// Band (A's lower triangle, half-width BW, zeros below) -> tridiagonal:
// d[k] = T[k][k], e[k] = T[k+1][k] in LDS.  Requires n >= 3.
//
// Wave w runs sweeps w, w + 8, ... in order, its tasks one after another.
// Sweep s's task t starts once sweep s - 1 has finished task t + 2: tasks
// that share an element then keep sweep order, and those running at the
// same time touch disjoint elements (the model of the pipelined order with a
// lag of three tasks, checked element by element).  Progress is published
// per wave in LDS (stat: NW ints, "sweep s has finished task t" as
// s * 65536 + t + 1, after the task's stores have completed); sweeps only
// wait for lower sweeps, and all waves of the workgroup are resident, so the
// waits end; each is still bounded (a wait that runs out marks the result
// non-finite).  The next task's diagonal block and block below are loaded
// before the progress store when sweep s - 1 is already past them (its task
// t + 3 finished; task t + 4 and later touch rows below them), so that load
// overlaps the store's completion; its left block is the block below of the
// task just done, kept in registers.
//
// band (optional, LDS, 32 n doubles): the chase runs on an LDS copy of the
// band's 32 lower diagonals (copied in first; stage 1 leaves zeros in the
// bulge room), so each task's blocks cost LDS round trips instead of memory
// round trips; A is then only read.
template <int NT, class BS>
__device__ __forceinline__ void band_chase_on(const BS& A, int n, gbl_dbl* V2, lds_dbl* d,
                                              lds_dbl* e, lds_int* stat)
{
  constexpr int NW = NT / 64;
  constexpr int SPIN_MAX = 1 << 26;
  V2 = sgpr_ptr(V2);
  n = sgpr_int(n);
  const int lane = threadIdx.x & 63, w = wave_id();
  if (threadIdx.x < NW + 1) stat[threadIdx.x] = threadIdx.x < NW ? -1 : 0;
  __syncthreads();
  bool ok = true;
  auto progress = [&](int sw) {
    return __hip_atomic_load(stat + (sw % NW), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
  };
  auto passed = [&](int sw, int t) {  // sweep sw has finished task t (or all of its tasks)
    return sw < 0 || progress(sw) >= sw * 65536 + min(t, chase_tasks(n, sw) - 1) + 1;
  };
  auto wait_passed = [&](int sw, int t) {
    int spins = 0;
    while (ok && !passed(sw, t)) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > SPIN_MAX) ok = false;
    }
  };
  for (int s = w; s <= n - 3 && ok; s += NW) {
    const int ns = chase_tasks(n, s);
    double lb[4], dd[4], rr[4];
    wait_passed(s - 1, 2);
    chase_load(A, n, s, 0, true, lb, dd, rr);
    for (int t = 0; t < ns && ok; ++t) {
      chase_apply(A, n, s, t, lb, dd, rr, V2);
      bool pre = false;
      if (t + 1 < ns) {
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) lb[jj] = rr[jj];
        if (passed(s - 1, t + 3)) {
          chase_load(A, n, s, t + 1, false, lb, dd, rr);
          pre = true;
        }
      }
      if (lane == 0)
        __hip_atomic_store(stat + w, s * 65536 + t + 1, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
      if (t + 1 < ns && !pre) {
        wait_passed(s - 1, t + 3);
        chase_load(A, n, s, t + 1, false, lb, dd, rr);
      }
    }
  }
  if (!ok && lane == 0) __hip_atomic_store(stat + NW, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  __syncthreads();
  const bool failed = stat[NW] != 0;
  for (int k = threadIdx.x; k < n; k += NT) {
    d[k] = failed ? NAN : A.get(k, k);
    if (k < n - 1) e[k] = failed ? NAN : A.get(k + 1, k);
  }
  __syncthreads();
}

template <int NT>
__device__ __noinline__ void band_chase(gbl_dbl* A, int n, int ld, gbl_dbl* V2, lds_dbl* d,
                                        lds_dbl* e, lds_int* stat, lds_dbl* band = nullptr)
{
  if (band) {
    for (int x = threadIdx.x; x < 32 * n; x += NT) {
      const int r = x >> 5, o = x & 31;
      band[x] = r - o >= 0 ? gld(A, (int64_t)r * ld + (r - o)) : 0.0;
    }
    __syncthreads();
    band_chase_on<NT>(BandLds{band}, n, V2, d, e, stat);
  } else {
    band_chase_on<NT>(BandGbl{sgpr_ptr(A), sgpr_int(ld)}, n, V2, d, e, stat);
  }
}

// ------------------------------------------------------ back-transforms
// Stage-2 reflectors in blocks: group g holds sweeps 16 g .. 16 g + 15, and
// block (g, t) their reflectors of task t, on rows rb .. rb + 30 (rb =
// 16 g + 1 + 16 t); reflector j of the block (sweep 16 g + j) acts on rows
// rb + j .. rb + j + 15, and a sweep without a task t adds the identity.
// Group g has F + 1 - g blocks (F = (n - 3) / 16), numbered from
// q2_first_block(F, g).  Y <- Q2 Y takes the groups last first and the
// blocks of a group in task order: two reflectors that share a row are
// sweeps s > s' with tasks (t, t) or (t, t + 1), so the one of the later
// sweep always comes first, as in the reverse of the chase.
__device__ __forceinline__ int q2_first_block(int F, int g) { return g * (F + 1) - g * (g - 1) / 2; }

// Matrix-core operands of every block into Op (1024 doubles per block; value
// q of lane l at q * 64 + l).  With V (32 x 16, row 31 zero) the block's
// reflectors and T its upper triangular factor (forward recurrence T[j][j] =
// tau_j, T[0:j, j] = -tau_j T[0:j, 0:j] V[:, 0:j]^T v_j; the block's product
// H_0 H_1 ... H_15 = I - V T V^T):
//   q = 0..7          U = V T^T, rows 4q .. 4q + 3: U[4q + l/16][l%16], the A
//                     operand of W' = T V^T Y = U^T Y
//   q = 8 + 4h + k    V[16h + l%16][4k + l/16], the A operand of
//                     Y_h -= V_h W' (rows 16h .. 16h + 15 of the block)
// V^T V on the matrix cores (the A and B operands of a k-step coincide), T
// row by row (row a depends only on its own earlier entries), U from T in
// LDS.  One block per wave at a time; lds: 512 doubles per wave.
template <int NT>
__device__ __noinline__ void q2_operands(const gbl_dbl* V2, int n, gbl_dbl* Op, lds_dbl* lds)
{
  constexpr int NW = NT / 64;
  V2 = sgpr_ptr(V2);
  Op = sgpr_ptr(Op);
  n = sgpr_int(n);
  const int lane = threadIdx.x & 63, w = wave_id();
  const int lr = lane & 15, lk = lane >> 4;
  const int F = (n - 3) / BW, G = F + 1;
  const int nb = q2_first_block(F, G);
  lds_dbl* Gs = lds + w * 512;  // V^T V, tau on the diagonal
  lds_dbl* Ts = Gs + 256;       // T
  for (int blk = w; blk < nb; blk += NW) {
    int g = 0;
    while (q2_first_block(F, g + 1) <= blk) ++g;
    const int t = blk - q2_first_block(F, g), s0 = BW * g;
    // V[r][j] (0 outside reflector j's rows, or for a sweep without task t)
    auto vel = [&](int r, int j, bool tau) -> double {
      const int s = s0 + j, e = tau ? 0 : r - j;
      const bool in = e >= 0 && e < BW && s <= n - 3 && t < chase_tasks(n, s);
      const unsigned rec = in ? chase_first(n, s) + t : 0;
      const double x = gld(V2, rec * BW + (in ? e : 0));
      return in ? ((e == 0 && !tau) ? 1.0 : x) : 0.0;
    };
    double x[8], vo[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = vel(4 * k + lk, lr, false);
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int k = 0; k < 4; ++k) vo[4 * h + k] = vel(16 * h + lr, 4 * k + lk, false);
    const double tau = vel(0, lr, true);
    ta_double4 gm = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 8; ++k) gm = mfma(x[k], x[k], gm);
#pragma unroll
    for (int q = 0; q < 4; ++q) Gs[(4 * q + lk) * BW + lr] = gm[q];
    if (lk == 0) Gs[lr * BW + lr] = tau;
    wave_lds_sync();
    // row a = lr of T
    {
      const int a = lr;
      double tr[BW];
#pragma unroll
      for (int j = 0; j < BW; ++j) {
        double v = 0.0;
        if (j == a) v = Gs[j * BW + j];
        if (j > a) {
          double acc = 0.0;
#pragma unroll
          for (int k = 0; k < j; ++k)
            if (k >= a) acc += tr[k] * Gs[k * BW + j];
          v = -Gs[j * BW + j] * acc;
        }
        tr[j] = v;
      }
      if (lk == 0) {
#pragma unroll
        for (int j = 0; j < BW; ++j) Ts[a * BW + j] = tr[j];
      }
    }
    wave_lds_sync();
    const unsigned ob = (unsigned)blk * 1024 + lane;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      ta_double4 du = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k = 0; k < 4; ++k) du = mfma(vo[4 * h + k], Ts[lr * BW + 4 * k + lk], du);
#pragma unroll
      for (int q = 0; q < 4; ++q) gst(Op, ob + (4 * h + q) * 64, du[q]);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) gst(Op, ob + (8 + i) * 64, vo[i]);
    wave_lds_sync();
  }
}

// Y <- Q2 Y on the first ncol columns of Y (Y[k * ld + col]) with the block
// operands of q2_operands.  Each wave carries one or two 16-column strips of
// Y (strips w + 16 i and w + 16 i + 8) through every block in order, with no
// barrier: per block W' = U^T Y_rows (eight matrix-core steps over the
// block's 32 rows, the D layout of the rows being the B-operand layout),
// then Y_rows -= V W' (the D layout of W' is again the B-operand layout).
// The block's 32 rows stay in registers as two 16-row halves: after block t
// the first half is final for the group and stored, the second becomes the
// first half of block t + 1, and the next 16 rows are loaded while the
// current block is multiplied (as are the next block's U operands).
template <int NT>
__device__ __noinline__ void apply_q2(const gbl_dbl* Op, int n, int ld, int ncol, gbl_dbl* Y)
{
  constexpr int NW = NT / 64;
  Op = sgpr_ptr(Op);
  Y = sgpr_ptr(Y);
  n = sgpr_int(n);
  ld = sgpr_int(ld);
  ncol = sgpr_int(ncol);
  const int lane = threadIdx.x & 63, w = wave_id();
  const int lr = lane & 15, lk = lane >> 4;
  const int F = (n - 3) / BW, G = F + 1;
  const int nst = (ncol + 15) >> 4;
  auto yload = [&](int r0, unsigned col, ta_double4& y) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int row = r0 + 4 * q + lk;
      y[q] = gld(Y, (row < n ? row : n - 1) * ld + col);
    }
  };
  auto ystore = [&](int r0, int col, const ta_double4& y) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int row = r0 + 4 * q + lk;
      if (row < n && col < ncol) gst(Y, row * ld + col, y[q]);
    }
  };
  for (int sa = w; sa < nst; sa += 2 * NW) {
    const bool hb = sa + NW < nst;
    const int ca = 16 * sa + lr, cb = 16 * (sa + NW) + lr;
    const unsigned cla = ca < ncol ? ca : 0, clb = (hb && cb < ncol) ? cb : 0;
    for (int g = G - 1; g >= 0; --g) {
      const int nt = F + 1 - g;
      const unsigned ob0 = (unsigned)q2_first_block(F, g) * 1024 + lane;
      int rb = BW * g + 1;
      ta_double4 a0, a1, b0, b1, na, nb;
      yload(rb, cla, a0);
      yload(rb + 16, cla, a1);
      if (hb) {
        yload(rb, clb, b0);
        yload(rb + 16, clb, b1);
      }
      double u[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) u[q] = gld(Op, ob0 + q * 64);
      for (int t = 0; t < nt; ++t) {
        const unsigned ob = ob0 + t * 1024;
        const bool more = t + 1 < nt;
        double v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = gld(Op, ob + (8 + i) * 64);
        if (more) {
          yload(rb + 32, cla, na);
          if (hb) yload(rb + 32, clb, nb);
        }
        ta_double4 wa = {0.0, 0.0, 0.0, 0.0}, wb = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < 4; ++k) wa = mfma(u[k], a0[k], wa);
#pragma unroll
        for (int k = 0; k < 4; ++k) wa = mfma(u[4 + k], a1[k], wa);
        if (hb) {
#pragma unroll
          for (int k = 0; k < 4; ++k) wb = mfma(u[k], b0[k], wb);
#pragma unroll
          for (int k = 0; k < 4; ++k) wb = mfma(u[4 + k], b1[k], wb);
        }
        if (more) {
#pragma unroll
          for (int q = 0; q < 8; ++q) u[q] = gld(Op, ob + 1024 + q * 64);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          a0 = mfma(v[k], -wa[k], a0);
          a1 = mfma(v[4 + k], -wa[k], a1);
        }
        ystore(rb, ca, a0);
        if (hb) {
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            b0 = mfma(v[k], -wb[k], b0);
            b1 = mfma(v[4 + k], -wb[k], b1);
          }
          ystore(rb, cb, b0);
        }
        if (more) {
          a0 = a1;
          a1 = na;
          b0 = b1;
          b1 = nb;
          rb += 16;
        } else {
          ystore(rb + 16, ca, a1);
          if (hb) ystore(rb + 16, cb, b1);
        }
      }
    }
  }
}

// Y <- Q1 Y on the first ncol columns of Y, Q1 = H_0 H_1 ... (stage-1
// reflectors, reflector k on rows k + BW..), one panel of 16 at a time, last
// panel first: Y -= V (T (V^T Y)) with the panel's T saved by stage 1.
//   W' = T V^T Y   one 16-column strip per wave: W = V^T Y on the matrix
//                  cores over the panel's rows (two accumulators, the next
//                  16 rows loaded while the current ones are multiplied),
//                  then W' = T W from the accumulator registers (the D layout
//                  of W is the B-operand layout of T W); W' into LDS.
//   Y -= V W'      16 x 16 blocks of Y (row block, strip) spread over the
//                  waves, each four matrix-core steps on the block loaded as
//                  the accumulator; the next block's operands are loaded
//                  before the current block is multiplied and stored.
// ws: nws doubles of LDS (T, then W' rows for as many columns as fit).
template <int NT>
__device__ __noinline__ void back_transform_q1(const gbl_dbl* V1, const gbl_dbl* T1, int n,
                                               int ld, int ncol, gbl_dbl* Y, lds_dbl* ws, int nws)
{
  constexpr int NW = NT / 64;
  V1 = sgpr_ptr(V1);
  T1 = sgpr_ptr(T1);
  Y = sgpr_ptr(Y);
  n = sgpr_int(n);
  ld = sgpr_int(ld);
  ncol = sgpr_int(ncol);
  const int tid = threadIdx.x, lane = tid & 63, w = wave_id();
  const int lr = lane & 15, lk = lane >> 4;
  const int np = band_reflectors(n) / BW;
  lds_dbl* Ts = ws;
  lds_dbl* Wq = ws + BW * BW;
  int cc = (((nws - BW * BW) / BW - 4) / 16) * 16;  // columns per chunk
  if (cc > ((ncol + 15) & ~15)) cc = (ncol + 15) & ~15;
  const int CSW = cc + 4;  // W' row stride
  for (int c0 = 0; c0 < ncol; c0 += cc) {
    const int cw = min(cc, ncol - c0);
    const int nst = (cw + 15) >> 4;
    for (int p = np - 1; p >= 0; --p) {
      const int r0 = BW * p + BW, m = n - r0;
      const gbl_dbl* V = V1 + (int64_t)(BW * p) * ld;  // reflector j: V[j * ld + r], r >= r0
      __syncthreads();  // the previous panel's blocks of Y are written, T and W' are free
      if (tid < BW * BW) Ts[tid] = T1[BW * BW * p + tid];
      __syncthreads();
      // ---- W' = T V^T Y, strip st = 16 columns
      for (int st = w; st < nst; st += NW) {
        const int cl = 16 * st + lr;
        const unsigned col = c0 + (cl < cw ? cl : 0);
        ta_double4 a0 = {0.0, 0.0, 0.0, 0.0}, a1 = {0.0, 0.0, 0.0, 0.0};
        double va[4], ya[4];
        auto fetch = [&](int k) {
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int r = k + 4 * u + lk;
            const bool ok = r < m;
            const unsigned rr = r0 + (ok ? r : 0);
            const double v = gld(V, lr * ld + rr);
            va[u] = ok ? v : 0.0;
            ya[u] = gld(Y, rr * ld + col);
          }
        };
        fetch(0);
        for (int k = 0; k < m; k += 16) {
          const double v0 = va[0], v1 = va[1], v2 = va[2], v3 = va[3];
          const double y0 = ya[0], y1 = ya[1], y2 = ya[2], y3 = ya[3];
          if (k + 16 < m) fetch(k + 16);
          a0 = mfma(v0, y0, a0);
          a1 = mfma(v1, y1, a1);
          a0 = mfma(v2, y2, a0);
          a1 = mfma(v3, y3, a1);
        }
        ta_double4 wq = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
          wq = mfma(Ts[lr * BW + 4 * kk + lk], a0[kk] + a1[kk], wq);
#pragma unroll
        for (int q = 0; q < 4; ++q) Wq[(4 * q + lk) * CSW + 16 * st + lr] = wq[q];
      }
      __syncthreads();
      // ---- Y -= V W' over (row block, strip) blocks
      const int nrb = (m + 15) >> 4;
      const int ntask = nrb * nst;
      ta_double4 cy;
      double av[4], bw[4];
      auto load = [&](int tk) {
        const int rb = tk / nst, st = tk - rb * nst;
        const int cl = 16 * st + lr;
        const unsigned col = c0 + (cl < cw ? cl : 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = r0 + 16 * rb + 4 * q + lk;
          cy[q] = gld(Y, (row < n ? row : r0) * ld + col);
        }
        const int ra = r0 + 16 * rb + lr;
        const unsigned rac = ra < n ? ra : r0;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
          av[ks] = gld(V, (4 * ks + lk) * ld + rac);
          bw[ks] = -Wq[(4 * ks + lk) * CSW + 16 * st + lr];
        }
      };
      if (w < ntask) load(w);
      for (int tk = w; tk < ntask; tk += NW) {
        ta_double4 c = cy;
        const double a0 = av[0], a1 = av[1], a2 = av[2], a3 = av[3];
        const double b0 = bw[0], b1 = bw[1], b2 = bw[2], b3 = bw[3];
        if (tk + NW < ntask) load(tk + NW);
        c = mfma(a0, b0, c);
        c = mfma(a1, b1, c);
        c = mfma(a2, b2, c);
        c = mfma(a3, b3, c);
        const int rb = tk / nst, st = tk - rb * nst;
        const int cl = 16 * st + lr;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = r0 + 16 * rb + 4 * q + lk;
          if (row < n && cl < cw) gst(Y, row * ld + c0 + cl, c[q]);
        }
      }
    }
  }
  __syncthreads();
}

}  // namespace band

// Dense symmetric eigensolver and GEMMs for molecules whose matrices live in
// global memory (one workgroup per molecule, NT = 512 threads).  This is the
// sygvd replacement (src/lapack/sygvd.cu:223-313) for AO counts above the
// LDS-resident Jacobi limit:
//
//   reduction      A = X F X^T with X = L^-1 (Cholesky factor of S)
//   tridiagonal    A = Q T Q^T, Householder reflectors, lower triangle only;
//                  the symmetric matrix-vector product of the next column is
//                  fused into the rank-2 update of the current one, so every
//                  column costs one read and one write of the trailing
//                  triangle
//   eigenvalues    bisection on Sturm counts, four-way multisection (three
//                  independent recurrences per step), one thread per level
//   eigenvectors   inverse iteration with a partial-pivoting tridiagonal LU,
//                  one thread per level, for the levels that carry occupation
//                  only; levels closer than ORTOL |T| are re-orthogonalised
//                  (modified Gram-Schmidt, two passes)
//   back-transform Z = Q Y in blocks of reflectors (compact WY form,
//                  I - V T V^T), one thread per column of Y
//   MO basis       C^T = Z^T X
//
// All matrices are row-major with leading dimension ld; the eigenvalues come
// out in ascending order.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "device_util.h"

namespace big {

constexpr double EPS = 2.220446049250313e-16;
constexpr double ORTOL = 1.0e-5;  // relative level gap below which vectors are re-orthogonalised

// ------------------------------------------------------------------ GEMM
// out(i, j) = sum_{k in [k0, k1)} a(i, k) b(k, j) over 64 x 64 output tiles;
// a and b are loader functors, kr(ti, tj, &k0, &k1) restricts the k range of
// a tile (triangular operands), skip(ti, tj) drops tiles, st(i, j, value)
// stores.  AK: a is read k-fastest (row-major a), else i-fastest.  BJ: b is
// read j-fastest (row-major b), else k-fastest.  tile: 128 KS doubles of LDS
// (KS = 16 by default: 2048).
//
// The product runs on the FP64 matrix cores (v_mfma_f64_16x16x4f64: lane l
// supplies a[l%16][l/16] and b[l/16][l%16] and holds d[4i + l/16][l%16] of
// the 16 x 16 result).  Each of the 8 waves owns a 16 x 32 strip of the tile;
// k slices are staged through LDS.  The global loads run GQ slices ahead of
// the multiplication, so that the memory latency of a slice is covered by
// the products of the slices before it rather than by one.
typedef double ta_double4 __attribute__((ext_vector_type(4)));

// One 64 x 64 output tile (rows i0.., columns j0..) over k in [k0, k1): the
// body shared by the single-workgroup product below and the grid-wide one of
// the multi-workgroup path.  All NT = 512 threads of the workgroup take part.
// KS: depth of the k slices staged through LDS (16 or 32; tile holds
// 128 KS doubles), two barriers per slice.
template <int NT, bool AK, bool BJ, class LA, class LB, class ST, int KS = 16>
__device__ __noinline__ void gemm_tile(int M, int N, int i0, int j0, int k0, int k1, LA la,
                                          LB lb, ST st, lds_dbl* tile)
{
  static_assert(NT == 512, "gemm tiling assumes 512 threads");
  static_assert(KS == 16 || KS == 32, "k slices of 16 or 32");
  constexpr int R = 64 * KS / NT;        // staged elements per thread and operand
  constexpr int GQ = KS == 16 ? 4 : 2;   // slices in flight
  lds_dbl* as = tile;            // [KS][64]: as[k * 64 + i]
  lds_dbl* bs = tile + 64 * KS;  // [KS][64]: bs[k * 64 + j]
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = wave_id();
  const int bi = w >> 1, bj = (w & 1) * 2;
  const int lr = lane & 15, lk = lane >> 4;
  // staging coordinates of this thread's elements of each operand slice
  int ai[R], ak[R], bjj[R], bk[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int idx = tid + r * NT;
    ai[r] = AK ? idx / KS : idx & 63;
    ak[r] = AK ? idx % KS : idx >> 6;
    bjj[r] = BJ ? idx & 63 : idx / KS;
    bk[r] = BJ ? idx >> 6 : idx % KS;
  }
  ta_double4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
  double ra[GQ][R], rb[GQ][R];
  auto fetch = [&](int q, int kb) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int gi = i0 + ai[r], gk = kb + ak[r];
      ra[q][r] = (gi < M && gk < k1) ? la(gi, gk) : 0.0;
      const int gj = j0 + bjj[r], gk2 = kb + bk[r];
      rb[q][r] = (gj < N && gk2 < k1) ? lb(gk2, gj) : 0.0;
    }
  };
#pragma unroll
  for (int q = 0; q < GQ; ++q)
    if (k0 + KS * q < k1) fetch(q, k0 + KS * q);
  for (int kq = k0; kq < k1; kq += KS * GQ) {
#pragma unroll
    for (int q = 0; q < GQ; ++q) {
      const int kb = kq + KS * q;
      if (kb < k1) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          as[ak[r] * 64 + ai[r]] = ra[q][r];
          bs[bk[r] * 64 + bjj[r]] = rb[q][r];
        }
        __syncthreads();
        if (kb + KS * GQ < k1) fetch(q, kb + KS * GQ);
#pragma unroll
        for (int kk = 0; kk < KS / 4; ++kk) {
          const int kr_ = (4 * kk + lk) * 64;
          const double a = as[kr_ + 16 * bi + lr];
          const double b0 = bs[kr_ + 16 * bj + lr];
          const double b1 = bs[kr_ + 16 * bj + 16 + lr];
          acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b0, acc0, 0, 0, 0);
          acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b1, acc1, 0, 0, 0);
        }
        __syncthreads();
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gi = i0 + 16 * bi + 4 * i + lk;
    const int gj = j0 + 16 * bj + lr;
    if (gi < M) {
      if (gj < N) st(gi, gj, acc0[i]);
      if (gj + 16 < N) st(gi, gj + 16, acc1[i]);
    }
  }
}

template <int NT, bool AK, bool BJ, int KS = 16, class LA, class LB, class KR, class SK, class ST>
__device__ __noinline__ void gemm(int M, int N, LA la, LB lb, KR kr, SK skip, ST st, lds_dbl* tile)
{
  const int mt = (M + 63) >> 6, nt = (N + 63) >> 6;
  for (int ti = 0; ti < mt; ++ti)
    for (int tj = 0; tj < nt; ++tj) {
      if (skip(ti, tj)) continue;
      int k0, k1;
      kr(ti, tj, &k0, &k1);
      gemm_tile<NT, AK, BJ, LA, LB, ST, KS>(M, N, ti * 64, tj * 64, k0, k1, la, lb, st, tile);
    }
  __syncthreads();
}

// out(i, j) = sum_{k in [k0, k1)} a(i, k) b(k, j) over 32 x 32 output blocks
// without LDS or barriers: each wave takes blocks w, w + NW, .. (skip(bi,
// bj) drops a block, kr(bi, bj, &k0, &k1) gives its k range) and holds the
// block as four 16 x 16 accumulators on the FP64 matrix cores, its operands
// read straight from memory in the operand layout (lane l supplies
// a[i0 + l%16][k + l/16] and b[k + l/16][j0 + l%16] per k-step of 4) and the
// next 16 k loaded while the current 16 are multiplied.  Out-of-range indices
// are clamped for the loads (the loaders see only i < M, j < N and k in
// [k0, k1)); k past k1 contributes zero.  st(i, j, value) for i < M, j < N.
template <int NT, class LA, class LB, class KR, class SK, class ST>
__device__ __noinline__ void gemm_reg(int M, int N, LA la, LB lb, KR kr, SK skip, ST st)
{
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, w = wave_id();
  const int lr = lane & 15, lk = lane >> 4;
  const int nbi = (M + 31) >> 5, nbj = (N + 31) >> 5;
  for (int b = w; b < nbi * nbj; b += NW) {
    const int bi = b / nbj, bj = b - bi * nbj;
    if (skip(bi, bj)) continue;
    int k0, k1;
    kr(bi, bj, &k0, &k1);
    const int i0 = 32 * bi, j0 = 32 * bj;
    const int ia = min(i0 + lr, M - 1), ib = min(i0 + 16 + lr, M - 1);
    const int ja = min(j0 + lr, N - 1), jb = min(j0 + 16 + lr, N - 1);
    ta_double4 c00 = {0.0, 0.0, 0.0, 0.0}, c01 = c00, c10 = c00, c11 = c00;
    double pa[4], pb[4], qa[4], qb[4];
    auto fetch = [&](int k) {
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int kk = k + 4 * s + lk;
        const bool in = kk < k1;
        const int kc = in ? kk : k1 - 1;
        const double x0 = la(ia, kc), x1 = la(ib, kc);
        pa[s] = in ? x0 : 0.0;
        pb[s] = in ? x1 : 0.0;
        qa[s] = lb(kc, ja);
        qb[s] = lb(kc, jb);
      }
    };
    if (k0 < k1) fetch(k0);
    for (int k = k0; k < k1; k += 16) {
      double ca[4], cb[4], da[4], db[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        ca[s] = pa[s];
        cb[s] = pb[s];
        da[s] = qa[s];
        db[s] = qb[s];
      }
      if (k + 16 < k1) fetch(k + 16);
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        c00 = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[s], da[s], c00, 0, 0, 0);
        c01 = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[s], db[s], c01, 0, 0, 0);
        c10 = __builtin_amdgcn_mfma_f64_16x16x4f64(cb[s], da[s], c10, 0, 0, 0);
        c11 = __builtin_amdgcn_mfma_f64_16x16x4f64(cb[s], db[s], c11, 0, 0, 0);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r0 = i0 + 4 * q + lk, r1 = r0 + 16;
      const int s0 = j0 + lr, s1 = s0 + 16;
      if (r0 < M && s0 < N) st(r0, s0, c00[q]);
      if (r0 < M && s1 < N) st(r0, s1, c01[q]);
      if (r1 < M && s0 < N) st(r1, s0, c10[q]);
      if (r1 < M && s1 < N) st(r1, s1, c11[q]);
    }
  }
  __syncthreads();
}

// ------------------------------------------------------ Householder vector
// x[lo..hi) (LDS) -> v with v[lo] = 1, H = I - tau v v^T, H x = beta e_lo
// (dlarfg).  Wave 0 computes the scalars; returns beta, tau via scal[0..1].
template <int NT>
__device__ void householder(lds_dbl* x, int lo, int hi, lds_dbl* scal)
{
  if (threadIdx.x < 64) {
    double s = 0.0;
    for (int r = lo + 1 + threadIdx.x; r < hi; r += 64) s += x[r] * x[r];
    s = wave_sum(s);
    if (threadIdx.x == 0) {
      const double alpha = x[lo];
      double beta, tau, scale;
      if (s == 0.0) {
        beta = alpha;
        tau = 0.0;
        scale = 0.0;
      } else {
        beta = -copysign(sqrt(alpha * alpha + s), alpha);
        tau = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
      }
      scal[0] = beta;
      scal[1] = tau;
      scal[2] = scale;
    }
  }
  __syncthreads();
  const double scale = scal[2];
  for (int r = lo + 1 + threadIdx.x; r < hi; r += NT) x[r] *= scale;
  __syncthreads();
  if (threadIdx.x == 0) x[lo] = 1.0;
  __syncthreads();
}

// ------------------------------------------------- fused trailing update
// Rows r in [lo, n), columns c in [lo, r]:
//   a' = a - v_r w_c - w_r v_c   (when UPDATE)
//   y_r += sum_c a' u_c (c <= r),  y_c += a' u_r (c < r)
// A wave takes RQ rows at a time (rows r0 + q NW) and walks their column
// blocks of 64 together, so RQ loads per lane are in flight; row sums go to
// yrow (LDS, one writer per row), column sums to row (wave mod npart) of ycol
// (npart x ld), combined afterwards.  With npart == NW every wave owns its
// row (deterministic sums); with fewer rows (large ld, less LDS) waves that
// share a row add atomically.  ycol must be zero on entry.
template <int NT, bool UPDATE, int RQ = 4>
__device__ __noinline__ void trail_pass(gbl_dbl* A, int n, int ld, int lo, const lds_dbl* v,
                                        const lds_dbl* w, const lds_dbl* u, lds_dbl* yrow,
                                        lds_dbl* ycol, int npart)
{
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wid = wave_id();
  lds_dbl* yc = ycol + (wid % npart) * ld;
  const bool shared_row = npart < NW;
  for (int r0 = lo + wid; r0 < n; r0 += RQ * NW) {
    int r[RQ];
    bool h[RQ];
    gbl_dbl* row[RQ];
    double vr[RQ], wr[RQ], ur[RQ], rs[RQ];
    int rmax = r0;
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
      r[q] = r0 + q * NW;
      h[q] = r[q] < n;
      if (h[q]) rmax = r[q];
      row[q] = A + (int64_t)(h[q] ? r[q] : r0) * ld;
      vr[q] = (UPDATE && h[q]) ? v[r[q]] : 0.0;
      wr[q] = (UPDATE && h[q]) ? w[r[q]] : 0.0;
      ur[q] = h[q] ? u[r[q]] : 0.0;
      rs[q] = 0.0;
    }
#pragma unroll 2
    for (int cb = lo; cb <= rmax; cb += 64) {
      const int c = cb + lane;
      double x[RQ];
      bool in[RQ];
#pragma unroll
      for (int q = 0; q < RQ; ++q) {
        in[q] = h[q] && c <= r[q];
        x[q] = in[q] ? row[q][c] : 0.0;
      }
      const bool kc = c < n;
      const double uc = kc ? u[c] : 0.0;
      if constexpr (UPDATE) {
        const double wc = kc ? w[c] : 0.0, vc = kc ? v[c] : 0.0;
#pragma unroll
        for (int q = 0; q < RQ; ++q) {
          x[q] -= vr[q] * wc + wr[q] * vc;
          if (in[q]) row[q][c] = x[q];
          if (!in[q]) x[q] = 0.0;
        }
      }
      double cs = 0.0;
#pragma unroll
      for (int q = 0; q < RQ; ++q) {
        rs[q] += x[q] * uc;
        if (c < r[q]) cs += x[q] * ur[q];
      }
      if (kc) {
        if (shared_row)
          __hip_atomic_fetch_add(yc + c, cs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else
          yc[c] += cs;
      }
    }
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
      const double t = wave_sum(rs[q]);
      if (lane == 0 && h[q]) yrow[r[q]] = t;
    }
  }
  __syncthreads();
}

// rows per wave in flight in the read-only symmetric product
#ifndef TA_SYMV_RQ
#define TA_SYMV_RQ 4
#endif

// --------------------------------------------- blocked tridiagonalisation
// Same result layout as tridiagonalize(), in panels of NP columns (dlatrd
// scheme): within a panel each column j is first brought up to date with
// the panel's earlier reflectors,
//   a_j -= V W_j^T + W V_j^T,
// its reflector is formed, and y = A_trail v is a read-only pass over the
// trailing triangle as it was at the start of the panel, corrected by
//   y -= V (W^T v) + W (V^T v),
//   w = tau y - (tau^2/2)(y.v) v.
// After the panel the trailing triangle takes the rank-2 NP update
// A -= V W^T + W V^T on the matrix cores.  Per column this reads the trailing
// triangle once (an unblocked two-sided update reads and writes it).  vt, wt:
// NP x ld doubles of global scratch (reflectors and w vectors, one per row);
// dots: 2 NP doubles of LDS; part: npart x ld doubles of LDS (also the GEMM
// tile region, free between passes).  Result: d[0..n), e[0..n-1)
// (T[k+1][k] = e[k]); reflector k (k = 0..n-3) stored as v_k[r] = A[r][k],
// r >= k+1 (v_k[k+1] = 1), tau[k].  Requires n >= 3.
template <int NT, int NP>
__device__ __noinline__ void tridiagonalize_blocked(gbl_dbl* A, int n, int ld, lds_dbl* d,
                                                   lds_dbl* e, lds_dbl* tau, lds_dbl* v,
                                                   lds_dbl* x, lds_dbl* y, lds_dbl* yrow,
                                                   lds_dbl* dots, lds_dbl* part, gbl_dbl* vt,
                                                   gbl_dbl* wt, int npart, lds_dbl* red,
                                                   lds_dbl* scal, lds_dbl* tprof)
{
  constexpr int NW = NT / 64;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wid = wave_id();
  const int nr = n - 2;  // reflectors 0 .. n-3
  for (int r = tid; r < ld; r += NT)
    for (int w = 0; w < npart; ++w) part[w * ld + r] = 0.0;
  // optional step clocks (thread 0, wall-clock ticks added to tprof[0..4])
  const bool clk_on = tprof != nullptr && tid == 0;
  long long clk_last = clk_on ? wall_clock64() : 0;
  auto lap = [&](int slot) {
    if (clk_on) {
      const long long t = wall_clock64();
      tprof[slot] += (double)(t - clk_last);
      clk_last = t;
    }
  };
  for (int k0 = 0; k0 < nr; k0 += NP) {
    const int np = (k0 + NP <= nr) ? NP : nr - k0;
    for (int i = 0; i < np; ++i) {
      const int j = k0 + i;
      // (a) column j, rows j..n-1, brought up to date with the panel's reflectors 0..i-1
      for (int r = j + tid; r < n; r += NT) {
        double a = A[(int64_t)r * ld + j];
        for (int q = 0; q < i; ++q)
          a -= vt[(int64_t)q * ld + r] * wt[(int64_t)q * ld + j] +
               wt[(int64_t)q * ld + r] * vt[(int64_t)q * ld + j];
        x[r] = a;
      }
      __syncthreads();
      lap(0);
      if (tid == 0) d[j] = x[j];
      // (b) reflector from x[j+1..n)
      householder<NT>(x, j + 1, n, scal);
      const double tj = scal[1];
      if (tid == 0) {
        e[j] = scal[0];
        tau[j] = tj;
      }
      for (int r = j + 1 + tid; r < n; r += NT) {
        A[(int64_t)r * ld + j] = x[r];
        vt[(int64_t)i * ld + r] = x[r];
        v[r] = x[r];
      }
      __syncthreads();
      lap(1);
      // (c) panel dot products dots[q] = w_q . v, dots[NP + q] = v_q . v (waves
      // 0 .. 2i-1 first), then y = A_trail v over rows/cols j+1.. (read-only)
      for (int pq = wid; pq < 2 * i; pq += NW) {
        const int q = pq < i ? pq : pq - i;
        const gbl_dbl* col = (pq < i ? wt : vt) + (int64_t)q * ld;
        double sacc = 0.0;
        for (int r = j + 1 + lane; r < n; r += 64) sacc += col[r] * v[r];
        sacc = wave_sum(sacc);
        if (lane == 0) dots[(pq < i ? 0 : NP) + q] = sacc;
      }
      trail_pass<NT, false, TA_SYMV_RQ>(A, n, ld, j + 1, v, v, v, yrow, part, npart);
      lap(2);
      // (d) y = tau (A_trail v - V (W^T v) - W (V^T v)), combined with the
      // per-wave partial sums in the same pass
      double sv = 0.0;
      for (int r = j + 1 + tid; r < n; r += NT) {
        double yr = yrow[r];
        for (int w = 0; w < npart; ++w) {
          yr += part[w * ld + r];
          part[w * ld + r] = 0.0;  // ready for the next column
        }
        for (int q = 0; q < i; ++q)
          yr -= vt[(int64_t)q * ld + r] * dots[q] + wt[(int64_t)q * ld + r] * dots[NP + q];
        yr *= tj;
        y[r] = yr;
        sv += yr * v[r];
      }
      sv = block_sum(sv, red);
      // (e) w = tau y - (tau/2)(tau y . v) v
      const double alpha = -0.5 * tj * sv;
      for (int r = j + 1 + tid; r < n; r += NT) wt[(int64_t)i * ld + r] = y[r] + alpha * v[r];
      // rows <= j of the panel vectors are zero (keeps the rank-2 update exact)
      for (int r = tid; r <= j; r += NT) {
        wt[(int64_t)i * ld + r] = 0.0;
        vt[(int64_t)i * ld + r] = 0.0;
      }
      __syncthreads();
      lap(3);
    }
    const int r0 = k0 + np;
    if (r0 >= nr) {
      // last two rows/columns: bring the 2 x 2 block up to date, no reflector
      if (tid < 3) {
        const int rr = (tid == 0) ? n - 2 : n - 1, cc = (tid == 2) ? n - 1 : n - 2;
        double a = A[(int64_t)rr * ld + cc];
        for (int q = 0; q < np; ++q)
          a -= vt[(int64_t)q * ld + rr] * wt[(int64_t)q * ld + cc] +
               wt[(int64_t)q * ld + rr] * vt[(int64_t)q * ld + cc];
        if (tid == 0) d[n - 2] = a;
        if (tid == 1) e[n - 2] = a;
        if (tid == 2) d[n - 1] = a;
      }
      __syncthreads();
      break;
    }
    // trailing update A[r0:, r0:] -= V W^T + W V^T (lower tiles; element-wise lower)
    const int m = n - r0;
    gemm<NT, false, true>(
        m, m,
        [=](int ii, int k) {
          return k < NP ? vt[(int64_t)k * ld + r0 + ii] : wt[(int64_t)(k - NP) * ld + r0 + ii];
        },
        [=](int k, int jj) {
          return k < NP ? wt[(int64_t)k * ld + r0 + jj] : vt[(int64_t)(k - NP) * ld + r0 + jj];
        },
        [](int, int, int* ka, int* kb) {
          *ka = 0;
          *kb = 2 * NP;
        },
        [](int ti, int tj) { return tj > ti; },
        [=](int ii, int jj, double val) {
          if (ii >= jj) A[(int64_t)(r0 + ii) * ld + r0 + jj] -= val;
        },
        part);
    // the product used the partial-sum rows as its tile: clear the columns
    // the next panel accumulates into
    for (int r = r0 + tid; r < n; r += NT)
      for (int w = 0; w < npart; ++w) part[w * ld + r] = 0.0;
    __syncthreads();
    lap(4);
  }
}

// ---------------------------------------------------------- eigenvalues
// Number of eigenvalues below each of three points (Sturm sequences run in
// lockstep for instruction-level parallelism).
// 1 / x to full double precision without the IEEE division sequence: the
// hardware reciprocal estimate refined by one Newton step (the Sturm
// recurrence only needs the sign of its pivots, and a relative error near
// the rounding unit perturbs T by that much).
__device__ __forceinline__ double rcp_nr(double x)
{
  const double r = __builtin_amdgcn_rcp(x);
  return fma(r, fma(-x, r, 1.0), r);
}

template <class PD, class PE>
__device__ __forceinline__ void sturm3(PD d, PE e2, int n, double x1,
                                       double x2, double x3, double pivmin, int& c1, int& c2,
                                       int& c3)
{
  double q1 = d[0] - x1, q2 = d[0] - x2, q3 = d[0] - x3;
  if (fabs(q1) < pivmin) q1 = -pivmin;
  if (fabs(q2) < pivmin) q2 = -pivmin;
  if (fabs(q3) < pivmin) q3 = -pivmin;
  int a1 = q1 < 0.0, a2 = q2 < 0.0, a3 = q3 < 0.0;
  for (int k = 1; k < n; ++k) {
    const double dk = d[k], ek = e2[k - 1];
    q1 = (dk - x1) - ek * rcp_nr(q1);
    q2 = (dk - x2) - ek * rcp_nr(q2);
    q3 = (dk - x3) - ek * rcp_nr(q3);
    if (fabs(q1) < pivmin) q1 = -pivmin;
    if (fabs(q2) < pivmin) q2 = -pivmin;
    if (fabs(q3) < pivmin) q3 = -pivmin;
    a1 += q1 < 0.0;
    a2 += q2 < 0.0;
    a3 += q3 < 0.0;
  }
  c1 = a1;
  c2 = a2;
  c3 = a3;
}

// All eigenvalues of T (d, e) in ascending order into lam, to an absolute
// accuracy of 1e-12 |T| (far below what the SCF observables resolve); e2 =
// e^2 scratch (LDS).  BTPL adjacent lanes share a level (4 by default): each
// evaluates three Sturm counts per step, so a step splits the bracket into
// 3 BTPL + 1 parts (13 for four lanes: log13 instead of log4 steps) and a
// workgroup holds NT / BTPL levels at once.  Fewer lanes per level do fewer
// Sturm counts in total (one lane: 3 log4 against 12 log13 per level), which
// is what counts once the levels fill the workgroup and every CU is busy.
// Warm start: with *warm > 0 each level is first bracketed in prev[i] -/+
// *warm (the previous SCF step's levels), checked by Sturm counts, and falls
// back to the Gershgorin interval when the check fails.  On return *warm
// holds the bracket for the next call (twice the largest level shift seen,
// plus a margin) and scal[0] = |T| (Gershgorin).
template <int NT, int BTPL = 4>
__device__ __noinline__ void bisect_all(const lds_dbl* d, const lds_dbl* e, lds_dbl* e2, int n,
                                       lds_dbl* lam, const lds_dbl* prev, lds_dbl* warm,
                                       lds_dbl* red, lds_dbl* scal)
{
  static_assert(NT % BTPL == 0, "lane groups of BTPL");
  constexpr int NPT = 3 * BTPL;  // points per step
  const int tid = threadIdx.x;
  double gl = INFINITY, gu = -INFINITY, emax = 0.0;
  for (int k = tid; k < n; k += NT) {
    const double el = k > 0 ? fabs(e[k - 1]) : 0.0, er = k < n - 1 ? fabs(e[k]) : 0.0;
    gl = fmin(gl, d[k] - el - er);
    gu = fmax(gu, d[k] + el + er);
    if (k < n - 1) {
      e2[k] = e[k] * e[k];
      emax = fmax(emax, e2[k]);
    }
  }
  gl = -block_max(-gl, red);
  gu = block_max(gu, red);
  emax = block_max(emax, red);
  const double wd = warm[0];
  const double tnorm = fmax(fabs(gl), fabs(gu));
  const double pivmin = 2.2250738585072014e-308 * fmax(1.0, emax);
  const double tol = 1.0e-12 * tnorm + 4.0 * EPS * tnorm + pivmin;
  gl -= 2.0 * tol;
  gu += 2.0 * tol;
  double shift = 0.0;
  const int sub = tid % BTPL;
  for (int i = tid / BTPL; i < n; i += NT / BTPL) {
    double lo = gl, hi = gu;
    if (wd > 0.0) {
      const double a = prev[i] - wd, b = prev[i] + wd;
      int ca, cb, cc;
      sturm3(d, e2, n, a, b, b, pivmin, ca, cb, cc);
      if (ca <= i && cb > i) {
        lo = a;
        hi = b;
      }
    }
    for (int it = 0; it < 40 && hi - lo > tol; ++it) {
      const double h = (hi - lo) / (double)(NPT + 1);
      const double x1 = lo + (3 * sub + 1) * h, x2 = lo + (3 * sub + 2) * h,
                   x3 = lo + (3 * sub + 3) * h;
      int c1, c2, c3;
      sturm3(d, e2, n, x1, x2, x3, pivmin, c1, c2, c3);
      // first point of this lane whose count exceeds i (NPT + 1: none)
      int j = (c1 > i) ? 3 * sub : ((c2 > i) ? 3 * sub + 1 : ((c3 > i) ? 3 * sub + 2 : NPT));
#pragma unroll
      for (int off = 1; off < BTPL; off <<= 1) j = min(j, __shfl_xor(j, off, BTPL));
      // the level lies between points j-1 and j (points -1 and NPT: lo, hi)
      const double nlo = j == 0 ? lo : lo + j * h;
      const double nhi = j == NPT ? hi : lo + (j + 1) * h;
      lo = nlo;
      hi = nhi;
    }
    const double l = 0.5 * (lo + hi);
    if (wd > 0.0) shift = fmax(shift, fabs(l - prev[i]));
    if (sub == 0) lam[i] = l;
  }
  shift = block_max(shift, red);
  if (tid == 0) {
    scal[0] = tnorm;
    warm[0] = wd > 0.0 ? 2.0 * shift + 1.0e-10 * tnorm : 1.0e-2 * tnorm;
  }
  __syncthreads();
}

// ---------------------------------------------------------- eigenvectors
__device__ __forceinline__ double start_vector(int i, int k)
{
  // deterministic pseudo-random entries in (-1, 1)
  unsigned x = (unsigned)i * 2654435761u ^ ((unsigned)k * 2246822519u + 0x9e3779b9u);
  x ^= x >> 15;
  x *= 2246822519u;
  x ^= x >> 13;
  x *= 3266489917u;
  x ^= x >> 16;
  return (double)(x >> 8) * (2.0 / 16777216.0) - 1.0;
}

// Inverse iteration: column i of Y (Y[k * ld + i]) receives the normalised
// eigenvector of T for lam[i].  scr: 3 * ld * NT doubles of global scratch,
// slot-fastest: per elimination step k the pivot reciprocal 1 / U_kk (and
// 1 / U_{n-1,n-1} at k = n - 1), the first superdiagonal U_k,k+1, and the
// row interchange: 2 when rows k and k+1 were not exchanged (the multiplier
// is then e_k / U_kk, and U_k,k+2 = 0), else the multiplier itself (|m| < 1;
// U_k,k+2 = e_k+1).  The factorisation of T - lam is formed in the first
// sweep only; the later sweeps reuse it, so their forward chain is one FMA
// per step instead of a division.
//
// Each sweep is a dependency chain along k whose loads do not depend on the
// chain, but follow the stores of earlier steps (which the compiler must
// assume may alias them); they are issued IB steps at a time ahead of those
// steps' stores, so that one memory latency is paid per IB steps.
template <int NT>
__device__ __noinline__ void inverse_iteration(const lds_dbl* d, const lds_dbl* e,
                                              const lds_dbl* lam, int n, int nvec, double tnorm,
                                              gbl_dbl* Y, int ld, gbl_dbl* scr)
{
  const double pert = EPS * tnorm + 1e-300;
  constexpr int NIT = 3;
  constexpr int IB = 8;
  Y = sgpr_ptr(Y);
  scr = sgpr_ptr(scr);
  n = sgpr_int(n);
  ld = sgpr_int(ld);
  for (int i0 = 0; i0 < nvec; i0 += NT) {
    const int i = i0 + threadIdx.x;
    if (i < nvec) {
      const int s = threadIdx.x;
      gbl_dbl* u0 = scr;                         // [k][NT]
      gbl_dbl* u1 = scr + (int64_t)ld * NT;
      gbl_dbl* u2 = scr + 2 * (int64_t)ld * NT;
      const double l = lam[i];
      double scale = 1.0;
      for (int it = 0; it < NIT; ++it) {
        // forward elimination with partial pivoting, rhs transformed in place
        double p = d[0] - l, q = n > 1 ? e[0] : 0.0;
        double bp = it == 0 ? start_vector(i, 0) : gld(Y, i) * scale;
        if (it > 0) {
          // stored factors: the rhs only
          for (int k0 = 0; k0 < n - 1; k0 += IB) {
            double bq[IB], a0[IB], a2[IB];
#pragma unroll
            for (int u = 0; u < IB; ++u) {
              const int k = k0 + u;
              const unsigned kc = k + 1 < n ? k + 1 : n - 1;
              const unsigned kf = k < n - 1 ? k : n - 2;
              bq[u] = gld(Y, kc * ld + i) * scale;
              a0[u] = gld(u0, kf * NT + s);
              a2[u] = gld(u2, kf * NT + s);
            }
#pragma unroll
            for (int u = 0; u < IB; ++u) {
              const int k = k0 + u;
              if (k < n - 1) {
                double rk;
                if (a2[u] == 2.0) {
                  const double m = e[k] * a0[u];
                  rk = bp;
                  bp = bq[u] - m * bp;
                } else {
                  rk = bq[u];
                  bp = bp - a2[u] * bq[u];
                }
                gst(Y, k * ld + i, rk);
              }
            }
          }
        } else
        for (int k0 = 0; k0 < n - 1; k0 += IB) {
          double bq[IB];
#pragma unroll
          for (int u = 0; u < IB; ++u) {
            const int k = k0 + u;
            const unsigned kc = k + 1 < n ? k + 1 : n - 1;
            bq[u] = it == 0 ? start_vector(i, k + 1) : gld(Y, kc * ld + i) * scale;
          }
#pragma unroll
          for (int u = 0; u < IB; ++u) {
            const int k = k0 + u;
            if (k < n - 1) {
              const double sk = e[k], dk1 = d[k + 1] - l, ek1 = (k + 1 < n - 1) ? e[k + 1] : 0.0;
              const double bk1 = bq[u];
              double U0, U1, X, rk;
              if (fabs(p) >= fabs(sk)) {
                if (fabs(p) < pert) p = p < 0.0 ? -pert : pert;
                const double m = sk / p;
                U0 = p;
                U1 = q;
                X = 2.0;
                rk = bp;
                p = dk1 - m * q;
                q = ek1;
                bp = bk1 - m * bp;
              } else {
                const double m = p / sk;
                U0 = sk;
                U1 = dk1;
                X = m;
                rk = bk1;
                p = q - m * dk1;
                q = -m * ek1;
                bp = bp - m * bk1;
              }
              gst(u0, k * NT + s, 1.0 / U0);
              gst(u1, k * NT + s, U1);
              gst(u2, k * NT + s, X);
              gst(Y, k * ld + i, rk);
            }
          }
        }
        // last pivot: formed in the first sweep, reused after
        double ip;
        if (it == 0) {
          if (fabs(p) < pert) p = p < 0.0 ? -pert : pert;
          ip = 1.0 / p;
          gst(u0, (n - 1) * NT + s, ip);
        } else {
          ip = gld(u0, (n - 1) * NT + s);
        }
        // back substitution
        double x2 = 0.0, x1 = bp * ip, nrm = x1 * x1;
        gst(Y, (n - 1) * ld + i, x1);
        for (int k1 = n - 2; k1 >= 0; k1 -= IB) {
          double yq[IB], a0[IB], a1[IB], a2[IB];
#pragma unroll
          for (int u = 0; u < IB; ++u) {
            const int k = k1 - u;
            const unsigned kc = k >= 0 ? k : 0;
            yq[u] = gld(Y, kc * ld + i);
            a0[u] = gld(u0, kc * NT + s);
            a1[u] = gld(u1, kc * NT + s);
            a2[u] = gld(u2, kc * NT + s);
          }
#pragma unroll
          for (int u = 0; u < IB; ++u) {
            const int k = k1 - u;
            if (k >= 0) {
              // U_k,k+2: e_k+1 after a row interchange, else 0
              const double u2k = a2[u] == 2.0 ? 0.0 : (k + 1 < n - 1 ? e[k + 1] : 0.0);
              const double x = (yq[u] - a1[u] * x1 - u2k * x2) * a0[u];
              gst(Y, k * ld + i, x);
              nrm += x * x;
              x2 = x1;
              x1 = x;
            }
          }
        }
        scale = 1.0 / sqrt(nrm);
      }
      for (int k0 = 0; k0 < n; k0 += IB) {
        double yq[IB];
#pragma unroll
        for (int u = 0; u < IB; ++u) {
          const unsigned kc = k0 + u < n ? k0 + u : n - 1;
          yq[u] = gld(Y, kc * ld + i);
        }
#pragma unroll
        for (int u = 0; u < IB; ++u)
          if (k0 + u < n) gst(Y, (k0 + u) * ld + i, yq[u] * scale);
      }
    }
  }
  __syncthreads();
}

// Re-orthogonalise eigenvectors of close levels among the first nvec (runs of
// lam with gaps below ORTOL |T|): one wave per run, modified Gram-Schmidt
// twice, then normalise.
// runs: int scratch of ld entries (LDS); nruns via scal.
template <int NT>
__device__ __noinline__ void orthogonalize_clusters(const lds_dbl* lam, int n, int nvec,
                                                   double tnorm, gbl_dbl* Y, int ld,
                                                   lds_int* runs, lds_dbl* scal)
{
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wid = wave_id();
  if (threadIdx.x == 0) {
    int nr = 0;
    const double gap = ORTOL * tnorm;
    int i = 0;
    while (i < nvec) {
      int j = i + 1;
      while (j < nvec && lam[j] - lam[j - 1] <= gap) ++j;
      if (j - i > 1) {
        runs[2 * nr] = i;
        runs[2 * nr + 1] = j;
        ++nr;
      }
      i = j;
    }
    scal[5] = (double)nr;
  }
  __syncthreads();
  const int nr = (int)scal[5];
  for (int c = wid; c < nr; c += NW) {
    const int a = runs[2 * c], b = runs[2 * c + 1];
    for (int j = a; j < b; ++j) {
      for (int pass = 0; pass < 2; ++pass)
        for (int l = a; l < j; ++l) {
          double s = 0.0;
          for (int k = lane; k < n; k += 64) s += Y[(int64_t)k * ld + l] * Y[(int64_t)k * ld + j];
          s = wave_sum(s);
          for (int k = lane; k < n; k += 64) Y[(int64_t)k * ld + j] -= s * Y[(int64_t)k * ld + l];
        }
      double s = 0.0;
      for (int k = lane; k < n; k += 64) {
        const double y = Y[(int64_t)k * ld + j];
        s += y * y;
      }
      s = wave_sum(s);
      const double inv = 1.0 / sqrt(s);
      for (int k = lane; k < n; k += 64) Y[(int64_t)k * ld + j] *= inv;
    }
  }
  __syncthreads();
}

// -------------------------------------------------------- back-transform
// Y <- Q Y on the first ncol columns of Y, Q = H_0 H_1 ... H_{n-3},
// reflectors from tridiagonalize().  Blocks
// of NB reflectors, last block first.  A block is I - V T V^T with T^-1 =
// diag(1/tau) + strict_upper(V^T V) (the reflectors satisfy tau = 2 / v.v), so
//   W = V^T Y,  solve T^-1 W' = W,  Y -= V W'
// three matrix-core products and a small triangular inversion per block.
// W: NB x ld doubles of global scratch; sinv, tm: NB x NB doubles of LDS.
template <int NT, int NB>
__device__ __noinline__ void back_transform(const gbl_dbl* A, const lds_dbl* tau, int n, int ld,
                                            int ncol, gbl_dbl* Y, gbl_dbl* W, lds_dbl* sinv,
                                            lds_dbl* tm, lds_dbl* tile)
{
  const int nr = n - 2;
  const int nblk = (nr + NB - 1) / NB;
  for (int bi = nblk - 1; bi >= 0; --bi) {
    const int k0 = bi * NB;
    const int nb = (k0 + NB <= nr) ? NB : nr - k0;
    const int r0 = k0 + 1;
    const int m = n - r0;
    // V(rr, j) = v_{k0+j}[r0 + rr]: nonzero for rr >= j; a reflector with
    // tau = 0 is the identity and gets a zero column
    auto vel = [=](int rr, int j) -> double {
      return (j < nb && rr >= j && tau[k0 + j] != 0.0) ? A[(int64_t)(r0 + rr) * ld + k0 + j]
                                                       : 0.0;
    };
    auto all_k = [m](int, int, int* ka, int* kb) {
      *ka = 0;
      *kb = m;
    };
    auto no_skip = [](int, int) { return false; };
    // T^-1 = diag(1/tau) + strict_upper(V^T V)
    gemm<NT, false, true>(
        NB, NB, [=](int i, int k) { return vel(k, i); }, [=](int k, int j) { return vel(k, j); },
        all_k, no_skip,
        [=](int i, int j, double val) {
          double s = 0.0;
          if (i < j) s = val;
          if (i == j) s = (i < nb && tau[k0 + i] != 0.0) ? 1.0 / tau[k0 + i] : 1.0;
          sinv[i * NB + j] = s;
        },
        tile);
    // W = V^T Y[r0:, :ncol]
    gemm<NT, false, true>(
        NB, ncol, [=](int i, int k) { return vel(k, i); },
        [=](int k, int j) { return Y[(int64_t)(r0 + k) * ld + j]; }, all_k, no_skip,
        [=](int i, int j, double val) { W[(int64_t)i * ld + j] = val; }, tile);
    // T = (T^-1)^-1, upper triangular: one thread per column, back substitution
    if (threadIdx.x < NB) {
      const int j = threadIdx.x;
      for (int i = NB - 1; i >= 0; --i) {
        double t = 0.0;
        if (i == j) {
          t = 1.0 / sinv[i * NB + i];
        } else if (i < j) {
          double acc = 0.0;
          for (int k = i + 1; k <= j; ++k) acc += sinv[i * NB + k] * tm[k * NB + j];
          t = -acc / sinv[i * NB + i];
        }
        tm[i * NB + j] = t;
      }
    }
    __syncthreads();
    // W' = T W (in place: a tile reads its own columns of W before storing them)
    gemm<NT, true, true>(
        NB, ncol, [=](int i, int k) { return tm[i * NB + k]; },
        [=](int k, int j) { return W[(int64_t)k * ld + j]; },
        [](int, int, int* ka, int* kb) {
          *ka = 0;
          *kb = NB;
        },
        no_skip, [=](int i, int j, double val) { W[(int64_t)i * ld + j] = val; }, tile);
    // Y[r0:, :ncol] -= V W'
    gemm<NT, true, true>(
        m, ncol, [=](int i, int k) { return vel(i, k); },
        [=](int k, int j) { return W[(int64_t)k * ld + j]; },
        [](int, int, int* ka, int* kb) {
          *ka = 0;
          *kb = NB;
        },
        no_skip, [=](int i, int j, double val) { Y[(int64_t)(r0 + i) * ld + j] -= val; }, tile);
  }
}

}  // namespace big

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
//                  one thread per level; levels closer than ORTOL |T| are
//                  re-orthogonalised (modified Gram-Schmidt, two passes)
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
// read j-fastest (row-major b), else k-fastest.  tile: 2048 doubles of LDS.
// Each thread holds a 4 x 2 micro tile.
template <int NT, bool AK, bool BJ, class LA, class LB, class KR, class SK, class ST>
__device__ __noinline__ void gemm(int M, int N, LA la, LB lb, KR kr, SK skip, ST st, double* tile)
{
  static_assert(NT == 512, "gemm tiling assumes 512 threads");
  double* as = tile;         // [16][64]: as[k * 64 + i]
  double* bs = tile + 1024;  // [16][64]: bs[k * 64 + j]
  const int tid = threadIdx.x;
  const int tx = tid & 31, ty = tid >> 5;
  const int mt = (M + 63) >> 6, nt = (N + 63) >> 6;
  for (int ti = 0; ti < mt; ++ti)
    for (int tj = 0; tj < nt; ++tj) {
      if (skip(ti, tj)) continue;
      int k0, k1;
      kr(ti, tj, &k0, &k1);
      const int i0 = ti * 64, j0 = tj * 64;
      double acc[4][2];
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u][0] = acc[u][1] = 0.0;
      for (int kb = k0; kb < k1; kb += 16) {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          const int idx = tid + r * NT;
          int i, k;
          if constexpr (AK) {
            i = idx >> 4;
            k = idx & 15;
          } else {
            i = idx & 63;
            k = idx >> 6;
          }
          const int gi = i0 + i, gk = kb + k;
          as[k * 64 + i] = (gi < M && gk < k1) ? la(gi, gk) : 0.0;
          int j;
          if constexpr (BJ) {
            j = idx & 63;
            k = idx >> 6;
          } else {
            j = idx >> 4;
            k = idx & 15;
          }
          const int gj = j0 + j, gk2 = kb + k;
          bs[k * 64 + j] = (gj < N && gk2 < k1) ? lb(gk2, gj) : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          double a[4], bv[2];
#pragma unroll
          for (int u = 0; u < 4; ++u) a[u] = as[k * 64 + 4 * ty + u];
          bv[0] = bs[k * 64 + 2 * tx];
          bv[1] = bs[k * 64 + 2 * tx + 1];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            acc[u][0] += a[u] * bv[0];
            acc[u][1] += a[u] * bv[1];
          }
        }
        __syncthreads();
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const int gi = i0 + 4 * ty + u, gj = j0 + 2 * tx + c;
          if (gi < M && gj < N) st(gi, gj, acc[u][c]);
        }
    }
  __syncthreads();
}

// ------------------------------------------------------ Householder vector
// x[lo..hi) (LDS) -> v with v[lo] = 1, H = I - tau v v^T, H x = beta e_lo
// (dlarfg).  Wave 0 computes the scalars; returns beta, tau via scal[0..1].
template <int NT>
__device__ void householder(double* x, int lo, int hi, double* scal)
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
// Row sums go to yrow (LDS, one writer per row), column sums to the
// per-wave partials part[w * ld + c] (global), combined afterwards.
template <int NT, int MAXT, bool UPDATE>
__device__ void trail_pass(double* A, int n, int ld, int lo, const double* v, const double* w,
                           const double* u, double* yrow, double* part)
{
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double ycol[MAXT];
#pragma unroll
  for (int t = 0; t < MAXT; ++t) ycol[t] = 0.0;
  for (int r = lo + wid; r < n; r += NW) {
    double* row = A + (int64_t)r * ld;
    const double vr = UPDATE ? v[r] : 0.0, wr = UPDATE ? w[r] : 0.0, ur = u[r];
    const int len = r - lo + 1;
    double rs = 0.0;
#pragma unroll
    for (int t = 0; t < MAXT; ++t) {
      // wave-uniform guard first: the blocks stay separate basic blocks, so
      // the loads of later column blocks are not hoisted (register pressure)
      if (64 * t >= len) break;
      const int c = lo + lane + 64 * t;
      if (c <= r) {
        double a = row[c];
        if constexpr (UPDATE) {
          a -= vr * w[c] + wr * v[c];
          row[c] = a;
        }
        rs += a * u[c];
        if (c < r) ycol[t] += a * ur;
      }
    }
    rs = wave_sum(rs);
    if (lane == 0) yrow[r] = rs;
  }
#pragma unroll
  for (int t = 0; t < MAXT; ++t) {
    const int c = lo + lane + 64 * t;
    if (c < n) part[(int64_t)wid * ld + c] = ycol[t];
  }
  __syncthreads();
}

template <int NT>
__device__ void combine_partials(int lo, int n, int ld, const double* yrow, const double* part,
                                 double* y)
{
  constexpr int NW = NT / 64;
  for (int c = lo + threadIdx.x; c < n; c += NT) {
    double s = yrow[c];
#pragma unroll
    for (int w = 0; w < NW; ++w) s += part[(int64_t)w * ld + c];
    y[c] = s;
  }
  __syncthreads();
}

// ------------------------------------------------------ tridiagonalisation
// A (global, lower triangle) -> d[0..n), e[0..n-1) (T[k+1][k] = e[k]);
// reflector k (k = 0..n-3) stored as v_k[r] = A[r][k], r >= k+1 (v_k[k+1] = 1),
// tau[k].  Work vectors v, w, vn, y, yrow: ld doubles of LDS each; part:
// NW * ld doubles of global scratch.  Requires n >= 3.
template <int NT, int MAXT>
__device__ __noinline__ void tridiagonalize(double* A, int n, int ld, double* d, double* e, double* tau,
                               double* v, double* w, double* vn, double* y, double* yrow,
                               double* part, double* red, double* scal)
{
  const int tid = threadIdx.x;
  // column 0
  if (tid == 0) d[0] = A[0];
  for (int r = 1 + tid; r < n; r += NT) v[r] = A[(int64_t)r * ld];
  __syncthreads();
  householder<NT>(v, 1, n, scal);
  if (tid == 0) {
    e[0] = scal[0];
    tau[0] = scal[1];
  }
  for (int r = 1 + tid; r < n; r += NT) A[(int64_t)r * ld] = v[r];
  // y = B_0 v_0, B_0 = A[1:, 1:]
  trail_pass<NT, MAXT, false>(A, n, ld, 1, v, v, v, yrow, part);
  combine_partials<NT>(1, n, ld, yrow, part, y);

  for (int k = 0; k <= n - 3; ++k) {
    const int lo = k + 1;
    const double tk = tau[k];
    // w = tau y - (tau/2)(tau y . v) v
    double s = 0.0;
    for (int r = lo + tid; r < n; r += NT) s += y[r] * v[r];
    s = block_sum(s, red);
    const double K = 0.5 * tk * tk * s;
    for (int r = lo + tid; r < n; r += NT) w[r] = tk * y[r] - K * v[r];
    __syncthreads();
    // updated column k+1 of the trailing matrix
    const int c1 = k + 1;
    if (tid == 0) d[c1] = A[(int64_t)c1 * ld + c1] - 2.0 * v[c1] * w[c1];
    for (int r = c1 + 1 + tid; r < n; r += NT)
      vn[r] = A[(int64_t)r * ld + c1] - v[r] * w[c1] - w[r] * v[c1];
    __syncthreads();
    if (k == n - 3) {
      // last 2 x 2 block: one off-diagonal element, last diagonal element
      if (tid == 0) {
        e[n - 2] = vn[n - 1];
        d[n - 1] = A[(int64_t)(n - 1) * ld + (n - 1)] - 2.0 * v[n - 1] * w[n - 1];
      }
      __syncthreads();
      break;
    }
    householder<NT>(vn, c1 + 1, n, scal);
    if (tid == 0) {
      e[c1] = scal[0];
      tau[c1] = scal[1];
    }
    for (int r = c1 + 1 + tid; r < n; r += NT) A[(int64_t)r * ld + c1] = vn[r];
    // rows/cols >= k+2: rank-2 update fused with y = B_{k+1} v_{k+1}
    trail_pass<NT, MAXT, true>(A, n, ld, c1 + 1, v, w, vn, yrow, part);
    combine_partials<NT>(c1 + 1, n, ld, yrow, part, y);
    for (int r = c1 + 1 + tid; r < n; r += NT) v[r] = vn[r];
    __syncthreads();
  }
}

// ---------------------------------------------------------- eigenvalues
// Number of eigenvalues below each of three points (Sturm sequences run in
// lockstep for instruction-level parallelism).
__device__ __forceinline__ void sturm3(const double* d, const double* e2, int n, double x1,
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
    q1 = (dk - x1) - ek / q1;
    q2 = (dk - x2) - ek / q2;
    q3 = (dk - x3) - ek / q3;
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

// All eigenvalues of T (d, e) in ascending order into lam; e2 = e^2 scratch
// (LDS).  Returns |T| (Gershgorin bound) through scal[0].
template <int NT>
__device__ __noinline__ void bisect_all(const double* d, const double* e, double* e2, int n, double* lam,
                           double* red, double* scal)
{
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
  const double tnorm = fmax(fabs(gl), fabs(gu));
  const double pivmin = 2.2250738585072014e-308 * fmax(1.0, emax);
  const double tol = 4.0 * EPS * tnorm + pivmin;
  gl -= 2.0 * tol;
  gu += 2.0 * tol;
  for (int i = tid; i < n; i += NT) {
    double lo = gl, hi = gu;
    for (int it = 0; it < 64 && hi - lo > tol; ++it) {
      const double h = hi - lo;
      const double x1 = lo + 0.25 * h, x2 = lo + 0.5 * h, x3 = lo + 0.75 * h;
      int c1, c2, c3;
      sturm3(d, e2, n, x1, x2, x3, pivmin, c1, c2, c3);
      if (c1 > i) {
        hi = x1;
      } else if (c2 > i) {
        lo = x1;
        hi = x2;
      } else if (c3 > i) {
        lo = x2;
        hi = x3;
      } else {
        lo = x3;
      }
    }
    lam[i] = 0.5 * (lo + hi);
  }
  if (tid == 0) scal[0] = tnorm;
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
// eigenvector of T for lam[i].  scr: 3 * ld * NT doubles of global scratch
// (pivots^-1, first and second superdiagonals of U, slot-fastest).
template <int NT>
__device__ __noinline__ void inverse_iteration(const double* d, const double* e, const double* lam, int n,
                                  double tnorm, double* Y, int ld, double* scr)
{
  const double pert = EPS * tnorm + 1e-300;
  constexpr int NIT = 3;
  for (int i0 = 0; i0 < n; i0 += NT) {
    const int i = i0 + threadIdx.x;
    if (i < n) {
      const int s = threadIdx.x;
      double* u0 = scr;                         // [k][NT]
      double* u1 = scr + (int64_t)ld * NT;
      double* u2 = scr + 2 * (int64_t)ld * NT;
      const double l = lam[i];
      double scale = 1.0;
      for (int it = 0; it < NIT; ++it) {
        // forward elimination with partial pivoting, rhs transformed in place
        double p = d[0] - l, q = n > 1 ? e[0] : 0.0;
        double bp = it == 0 ? start_vector(i, 0) : Y[i] * scale;
        for (int k = 0; k < n - 1; ++k) {
          const double sk = e[k], dk1 = d[k + 1] - l, ek1 = (k + 1 < n - 1) ? e[k + 1] : 0.0;
          const double bk1 =
              it == 0 ? start_vector(i, k + 1) : Y[(int64_t)(k + 1) * ld + i] * scale;
          double U0, U1, U2, rk;
          if (fabs(p) >= fabs(sk)) {
            if (fabs(p) < pert) p = p < 0.0 ? -pert : pert;
            const double m = sk / p;
            U0 = p;
            U1 = q;
            U2 = 0.0;
            rk = bp;
            p = dk1 - m * q;
            q = ek1;
            bp = bk1 - m * bp;
          } else {
            const double m = p / sk;
            U0 = sk;
            U1 = dk1;
            U2 = ek1;
            rk = bk1;
            p = q - m * dk1;
            q = -m * ek1;
            bp = bp - m * bk1;
          }
          u0[(int64_t)k * NT + s] = 1.0 / U0;
          u1[(int64_t)k * NT + s] = U1;
          u2[(int64_t)k * NT + s] = U2;
          Y[(int64_t)k * ld + i] = rk;
        }
        if (fabs(p) < pert) p = p < 0.0 ? -pert : pert;
        // back substitution
        double x2 = 0.0, x1 = bp / p, nrm = x1 * x1;
        Y[(int64_t)(n - 1) * ld + i] = x1;
        for (int k = n - 2; k >= 0; --k) {
          const double x = (Y[(int64_t)k * ld + i] - u1[(int64_t)k * NT + s] * x1 -
                            u2[(int64_t)k * NT + s] * x2) *
                           u0[(int64_t)k * NT + s];
          Y[(int64_t)k * ld + i] = x;
          nrm += x * x;
          x2 = x1;
          x1 = x;
        }
        scale = 1.0 / sqrt(nrm);
      }
      for (int k = 0; k < n; ++k) Y[(int64_t)k * ld + i] *= scale;
    }
  }
  __syncthreads();
}

// Re-orthogonalise eigenvectors of close levels (runs of lam with gaps below
// ORTOL |T|): one wave per run, modified Gram-Schmidt twice, then normalise.
// runs: int scratch of ld entries (LDS); nruns via scal.
template <int NT>
__device__ __noinline__ void orthogonalize_clusters(const double* lam, int n, double tnorm, double* Y, int ld,
                                       int* runs, double* scal)
{
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (threadIdx.x == 0) {
    int nr = 0;
    const double gap = ORTOL * tnorm;
    int i = 0;
    while (i < n) {
      int j = i + 1;
      while (j < n && lam[j] - lam[j - 1] <= gap) ++j;
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
// Y <- Q Y, Q = H_0 H_1 ... H_{n-3}, reflectors from tridiagonalize().  Blocks
// of NB reflectors, last block first; vb: NB * ld doubles of LDS, tm: NB*NB.
template <int NT, int NB>
__device__ __noinline__ void back_transform(const double* A, const double* tau, int n, int ld, double* Y,
                               double* vb, double* tm, double* gm)
{
  const int tid = threadIdx.x;
  const int lane = tid & 63, wid = tid >> 6;
  constexpr int NW = NT / 64;
  const int nr = n - 2;
  const int nblk = (nr + NB - 1) / NB;
  for (int bi = nblk - 1; bi >= 0; --bi) {
    const int k0 = bi * NB;
    const int nb = (k0 + NB <= nr) ? NB : nr - k0;
    const int r0 = k0 + 1;
    const int m = n - r0;
    // V block: vb[(r - r0) * NB + j] = v_{k0+j}[r]
    for (int idx = tid; idx < m * NB; idx += NT) {
      const int rr = idx / NB, j = idx - (idx / NB) * NB;
      const int r = r0 + rr;
      vb[idx] = (j < nb && r >= k0 + j + 1) ? A[(int64_t)r * ld + k0 + j] : 0.0;
    }
    __syncthreads();
    // Gram entries G[l][j] = v_l . v_j (l < j), one wave per pair
    for (int pr = wid; pr < NB * NB; pr += NW) {
      const int l = pr / NB, j = pr - (pr / NB) * NB;
      if (l < j && j < nb) {
        double s = 0.0;
        for (int rr = lane; rr < m; rr += 64) s += vb[rr * NB + l] * vb[rr * NB + j];
        s = wave_sum(s);
        if (lane == 0) gm[l * NB + j] = s;
      }
    }
    __syncthreads();
    // compact WY factor (dlarft, forward, columnwise)
    if (tid == 0) {
      for (int j = 0; j < NB; ++j)
        for (int l = 0; l < NB; ++l) tm[l * NB + j] = 0.0;
      for (int j = 0; j < nb; ++j) {
        const double tj = tau[k0 + j];
        tm[j * NB + j] = tj;
        for (int l = 0; l < j; ++l) {
          double s = 0.0;
          for (int q = l; q < j; ++q) s += tm[l * NB + q] * gm[q * NB + j];
          tm[l * NB + j] = -tj * s;
        }
      }
    }
    __syncthreads();
    for (int c = tid; c < n; c += NT) {
      double wv[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) wv[j] = 0.0;
      for (int rr = 0; rr < m; ++rr) {
        const double y = Y[(int64_t)(r0 + rr) * ld + c];
#pragma unroll
        for (int j = 0; j < NB; ++j) wv[j] += vb[rr * NB + j] * y;
      }
      double w2[NB];
#pragma unroll
      for (int l = 0; l < NB; ++l) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < NB; ++j)
          if (j >= l) s += tm[l * NB + j] * wv[j];
        w2[l] = s;
      }
      for (int rr = 0; rr < m; ++rr) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < NB; ++j) s += vb[rr * NB + j] * w2[j];
        Y[(int64_t)(r0 + rr) * ld + c] -= s;
      }
    }
    __syncthreads();
  }
}

}  // namespace big

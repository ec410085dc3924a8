// Symmetric eigensolver for the SCF kernel whose matrices live in LDS
// (padded AO dimension ld <= 96, one workgroup per molecule): the sygvd
// replacement (src/lapack/sygvd.cu:223-313) for the small sizes.
//
//   basis          A = CT F CT^T in the orthonormal MO basis of the previous
//                  iteration (rows of CT; L^-T from the Cholesky factor of S
//                  at the start), on the matrix cores (scf.hip)
//   tridiagonal    A = Q T Q^T by Householder reflectors, the full square of
//                  A kept (exactly symmetric); each reflector is applied to
//                  the rows of CT as soon as it is formed (CT <- H CT), so no
//                  reflector is stored.  Two barriers per column: (A) the
//                  symmetric product p = tau A v and the column sums v^T CT,
//                  (B) the rank-2 update of A and the rank-1 update of CT,
//                  while wave 0 updates the next column itself and forms the
//                  next reflector (double-buffered)
//   eigenvalues    bisection on Sturm counts (eig_big.h), warm-started from
//                  the previous iteration's levels
//   eigenvectors   one thread per level: twisted factorisation of T - lam
//                  (the eigenvector of the twist with the smallest pivot,
//                  z_r = 1), repeated at the Rayleigh quotient of that
//                  vector; levels closer than ORTOL |T| to a neighbour (rare)
//                  by inverse iteration from distinct start vectors and
//                  modified Gram-Schmidt within their cluster (two passes)
//   new basis      CT <- Z^T CT on the matrix cores: its rows are the MO
//                  coefficients, in ascending level order
//
// Work region (doubles): vb [2 ld] (reflectors), p [ld], cpart [5 ld], d, e,
// e2 [ld each]: 11 ld, the size of the Mulliken partial sums.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "device_util.h"
#include "eig_big.h"

namespace ldseig {

constexpr double ORTOL = 1.0e-5;  // relative level gap below which levels form a cluster
constexpr int kMaxG = 5;          // row groups of the column sums v^T CT
constexpr int kPick = 4;          // twist candidates kept per level

__device__ __forceinline__ int skx(int i, int j, int ld)
{
  int c = i + j;
  c = c >= ld ? c - ld : c;
  return i * ld + c;
}

struct Work {
  lds_dbl *vb, *p, *cpart, *d, *e, *e2;
};

__device__ __forceinline__ Work work_layout(double* wk_, int ld)
{
  lds_dbl* wk = TA_LDS(wk_);
  Work W;
  W.vb = wk;
  W.p = wk + 2 * ld;
  W.cpart = wk + 3 * ld;
  W.d = wk + 8 * ld;
  W.e = wk + 9 * ld;
  W.e2 = wk + 10 * ld;
  return W;
}

// Reflector of x[0..len) (dlarfg): H = I - tau v v^T, v[0] = 1, H x = beta e_0.
// Called by all 64 lanes of one wave; x(i) reads element i; v receives the
// vector, *tau_out and *beta_out (lane 0) the scalars.
template <class X>
__device__ __forceinline__ void make_reflector(X x, int len, lds_dbl* v, lds_dbl* tau_out,
                                               lds_dbl* beta_out)
{
  const int lane = threadIdx.x & 63;
  const double alpha = x(0);
  double s = 0.0;
  for (int i = 1 + lane; i < len; i += 64) {
    const double xi = x(i);
    v[i] = xi;
    s += xi * xi;
  }
  s = wave_sum(s);
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
  for (int i = 1 + lane; i < len; i += 64) v[i] *= scale;  // each lane rescales its own writes
  if (lane == 0) {
    v[0] = 1.0;
    *tau_out = tau;
    *beta_out = beta;
  }
}

// Householder tridiagonalisation of the symmetric n x n matrix A (skewed
// layout, both triangles), reflectors applied to the rows of CT (plain,
// stride ld).  Result d[0..n), e[0..n-1) (T[k+1][k] = e[k]) in the work
// region; A is left unspecified.  red: 64 doubles (entries 32..39 used),
// scal: 16 doubles (entries 8..11 used).
template <int NT>
__device__ __noinline__ void tridiagonalize(lds_dbl* A, lds_dbl* CT, int n, int ld, Work W,
                                            lds_dbl* red, lds_dbl* scal)
{
  constexpr int NW = NT / 64;
  static_assert(NW >= 2, "needs a wave for the next column and at least one worker wave");
  const int tid = threadIdx.x, lane = tid & 63, wid = wave_id();
  lds_dbl* d = W.d;
  lds_dbl* e = W.e;
  if (n <= 2) {
    if (tid == 0) {
      d[0] = A[skx(0, 0, ld)];
      if (n == 2) {
        d[1] = A[skx(1, 1, ld)];
        e[0] = A[skx(1, 0, ld)];
      }
    }
    __syncthreads();
    return;
  }
  if (wid == 0) {
    if (lane == 0) d[0] = A[skx(0, 0, ld)];
    make_reflector([&](int i) { return A[skx(1 + i, 0, ld)]; }, n - 1, W.vb, scal + 8, e);
  }
  __syncthreads();
  const int G = min(kMaxG, NT / n);
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;  // trailing rows/columns k+1 .. n-1
    const int bsel = k & 1;
    const lds_dbl* v = W.vb + bsel * ld;
    const double tau = scal[8 + bsel];
    // (A) p_i = tau sum_j A(k+1+i, k+1+j) v_j (16-lane segments over rows),
    // partial sums of p.v, and the column sums of the CT rows k+1..
    {
      constexpr int NSEG = NT / 16;
      const int seg = tid >> 4, sl = tid & 15;
      double pv = 0.0;
      for (int i = seg; i < m; i += NSEG) {
        const int row = k + 1 + i;
        double s = 0.0;
        for (int j = sl; j < m; j += 16) s += A[skx(row, k + 1 + j, ld)] * v[j];
        s = seg16_sum(s);
        const double pi = tau * s;
        if (sl == 0) {
          W.p[i] = pi;
          pv += pi * v[i];
        }
      }
      pv = wave_sum(pv);
      if (lane == 0) red[32 + wid] = pv;
      if (tid < G * n) {
        // four independent partial sums keep four LDS loads in flight
        const int g = tid / n, c = tid - g * n;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int r = g;
        for (; r + 3 * G < m; r += 4 * G) {
          a0 += v[r] * CT[(k + 1 + r) * ld + c];
          a1 += v[r + G] * CT[(k + 1 + r + G) * ld + c];
          a2 += v[r + 2 * G] * CT[(k + 1 + r + 2 * G) * ld + c];
          a3 += v[r + 3 * G] * CT[(k + 1 + r + 3 * G) * ld + c];
        }
        for (; r < m; r += G) a0 += v[r] * CT[(k + 1 + r) * ld + c];
        W.cpart[g * ld + c] = (a0 + a1) + (a2 + a3);
      }
    }
    __syncthreads();
    // (B) w = p - (tau/2)(p.v) v; A -= v w^T + w v^T; CT rows -= tau v_r (v^T CT)
    double K = 0.0;
    for (int q = 0; q < NW; ++q) K += red[32 + q];
    const double alpha = -0.5 * tau * K;
    if (wid == 0) {
      // column k+1 and its mirror row, the diagonal (k+1, k+1), then the
      // reflector of column k+1 from the values the lanes hold in registers
      // (or the last off-diagonal)
      const int bn = bsel ^ 1;
      lds_dbl* vn = W.vb + bn * ld;
      const double v0 = v[0], w0 = W.p[0] + alpha * v0;
      double xr[2] = {0.0, 0.0}, s2 = 0.0;
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int i = 1 + lane + 64 * t;
        if (i < m) {
          const double vi = v[i], wi = W.p[i] + alpha * vi;
          const double x = A[skx(k + 1 + i, k + 1, ld)] - (vi * w0 + wi * v0);
          A[skx(k + 1 + i, k + 1, ld)] = x;
          A[skx(k + 1, k + 1 + i, ld)] = x;
          xr[t] = x;
          if (i >= 2) s2 += x * x;
        }
      }
      const double x1 = __shfl(xr[0], 0, 64);  // trailing index 1 (lane 0)
      s2 = wave_sum(s2);
      if (lane == 0) {
        const double a = A[skx(k + 1, k + 1, ld)] - (v0 * w0 + w0 * v0);
        A[skx(k + 1, k + 1, ld)] = a;
        d[k + 1] = a;
      }
      if (m - 1 >= 2) {
        double beta, taun, scale;
        if (s2 == 0.0) {
          beta = x1;
          taun = 0.0;
          scale = 0.0;
        } else {
          beta = -copysign(sqrt(x1 * x1 + s2), x1);
          taun = (beta - x1) / beta;
          scale = 1.0 / (x1 - beta);
        }
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          const int i = 1 + lane + 64 * t;
          if (i < m) vn[i - 1] = (i == 1) ? 1.0 : xr[t] * scale;
        }
        if (lane == 0) {
          scal[8 + bn] = taun;
          e[k + 1] = beta;
        }
      } else if (lane == 0) {
        e[k + 1] = x1;
      }
    } else {
      // loads of a batch (two rows x two column blocks) are issued before its
      // stores, so four LDS round trips overlap
      constexpr int NWW = NW - 1;
      const int wv = wid - 1;
      for (int i0 = 1 + wv; i0 < m; i0 += 2 * NWW) {
        int ad[4];
        bool ok[4];
        double x[4], vi[2], wi[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int i = i0 + q * NWW;
          const bool oi = i < m;
          vi[q] = oi ? v[i] : 0.0;
          wi[q] = oi ? W.p[i] + alpha * vi[q] : 0.0;
#pragma unroll
          for (int t = 0; t < 2; ++t) {
            const int j = 1 + lane + 64 * t;
            ok[2 * q + t] = oi && j < m;
            ad[2 * q + t] = ok[2 * q + t] ? skx(k + 1 + i, k + 1 + j, ld) : 0;
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = ok[u] ? A[ad[u]] : 0.0;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          const int j = 1 + lane + 64 * t;
          const double vj = j < m ? v[j] : 0.0;
          const double wj = j < m ? W.p[j] + alpha * vj : 0.0;
#pragma unroll
          for (int q = 0; q < 2; ++q) x[2 * q + t] -= vi[q] * wj + wi[q] * vj;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (ok[u]) A[ad[u]] = x[u];
      }
      double sc0 = 0.0, sc1 = 0.0;
      for (int g = 0; g < G; ++g) {
        if (lane < n) sc0 += W.cpart[g * ld + lane];
        if (lane + 64 < n) sc1 += W.cpart[g * ld + lane + 64];
      }
      for (int r0 = wv; r0 < m; r0 += 2 * NWW) {
        const int r1 = r0 + NWW;
        const bool o1 = r1 < m;
        lds_dbl* row0 = CT + (k + 1 + r0) * ld;
        lds_dbl* row1 = CT + (k + 1 + (o1 ? r1 : r0)) * ld;
        const double t0 = tau * v[r0], t1 = o1 ? tau * v[r1] : 0.0;
        const bool c0 = lane < n, c1 = lane + 64 < n;
        double y00 = c0 ? row0[lane] : 0.0, y01 = c1 ? row0[lane + 64] : 0.0;
        double y10 = (o1 && c0) ? row1[lane] : 0.0, y11 = (o1 && c1) ? row1[lane + 64] : 0.0;
        if (c0) row0[lane] = y00 - t0 * sc0;
        if (c1) row0[lane + 64] = y01 - t0 * sc1;
        if (o1 && c0) row1[lane] = y10 - t1 * sc0;
        if (o1 && c1) row1[lane + 64] = y11 - t1 * sc1;
      }
    }
    __syncthreads();
  }
  if (tid == 0) d[n - 1] = A[skx(n - 1, n - 1, ld)];
  __syncthreads();
}

// Eigenvector of T (d, e; n) for the level lam by the twisted
// factorisation T - lam = N_r Delta N_r^T: z_r = 1, z_k = -L_k z_{k+1}
// (k < r), z_{k+1} = -U_k z_k (k >= r), with r the twist of the pick-th
// smallest |gamma_k| (pick 0: the smallest).  z(k) addresses element k of the
// output row, which also holds the intermediate pivots.  Returns the
// Euclidean norm of the unnormalised vector.
template <class Z>
__device__ __forceinline__ double twisted_vector(const lds_dbl* __restrict__ d,
                                                 const lds_dbl* __restrict__ e, int n,
                                                 double lam, double pivmin, int pick, Z z)
{
  auto clampp = [pivmin](double x) { return fabs(x) < pivmin ? -pivmin : x; };
  // forward: D+_k into z(k)
  double dp = clampp(d[0] - lam);
  z(0) = dp;
  for (int k = 0; k + 1 < n; ++k) {
    const double ek = e[k];
    dp = clampp((d[k + 1] - lam) - (ek * big::rcp_nr(dp)) * ek);
    z(k + 1) = dp;
  }
  // backward: D-_k, gamma_k = D+_k + D-_k - (d_k - lam); U_k into z(k+1)
  double gbest[kPick];
  int rbest[kPick];
#pragma unroll
  for (int t = 0; t < kPick; ++t) {
    gbest[t] = INFINITY;
    rbest[t] = 0;
  }
  auto track = [&](double g, int k) {
    double ag = fabs(g);
    int rk = k;
#pragma unroll
    for (int t = 0; t < kPick; ++t) {
      if (ag < gbest[t]) {
        const double tg = gbest[t];
        const int tr = rbest[t];
        gbest[t] = ag;
        rbest[t] = rk;
        ag = tg;
        rk = tr;
      }
    }
  };
  double dm = clampp(d[n - 1] - lam);
  track(z(n - 1) + dm - (d[n - 1] - lam), n - 1);
  // D+_k is read one step ahead, before the store of U_{k+1} that follows it
  double dplus_next = n >= 2 ? z(n - 2) : 0.0;
  for (int k = n - 2; k >= 0; --k) {
    const double dplus = dplus_next;
    if (k > 0) dplus_next = z(k - 1);
    const double ek = e[k];
    const double u = ek * big::rcp_nr(dm);
    dm = clampp((d[k] - lam) - u * ek);
    z(k + 1) = u;
    track(dplus + dm - (d[k] - lam), k);
  }
  int r = rbest[0];
#pragma unroll
  for (int t = 1; t < kPick; ++t)
    if (pick >= t && gbest[t] < INFINITY) r = rbest[t];
  // L_k (k < r) again, into z(k)
  dp = clampp(d[0] - lam);
  for (int k = 0; k < r; ++k) {
    const double ek = e[k];
    const double l = ek * big::rcp_nr(dp);
    z(k) = l;
    dp = clampp((d[k + 1] - lam) - l * ek);
  }
  z(r) = 1.0;
  double nrm = 1.0;
  for (int k = r - 1; k >= 0; --k) {
    const double zk = -z(k) * z(k + 1);
    z(k) = zk;
    nrm += zk * zk;
  }
  for (int k = r; k + 1 < n; ++k) {
    const double zk = -z(k + 1) * z(k);
    z(k + 1) = zk;
    nrm += zk * zk;
  }
  return sqrt(nrm);
}

// Inverse iteration for one level by one thread (levels in a cluster): three
// solves of (T - lam) y = x with partial pivoting from a pseudo-random start
// (the published dstein scheme, as in eig_big.h), the factors in scr (3 n
// doubles).  z(k) addresses element k of the result (normalised).
template <class Z>
__device__ void inverse_iteration_one(const lds_dbl* d, const lds_dbl* e, int n, double lam,
                                      double pert, int seed, lds_dbl* scr, Z z)
{
  lds_dbl* u0 = scr;
  lds_dbl* u1 = scr + n;
  lds_dbl* u2 = scr + 2 * n;
  double scale = 1.0;
  for (int it = 0; it < 3; ++it) {
    double p = d[0] - lam, q = n > 1 ? e[0] : 0.0;
    double bp = it == 0 ? big::start_vector(seed, 0) : z(0) * scale;
    for (int k = 0; k < n - 1; ++k) {
      const double sk = e[k], dk1 = d[k + 1] - lam, ek1 = (k + 1 < n - 1) ? e[k + 1] : 0.0;
      const double bk1 = it == 0 ? big::start_vector(seed, k + 1) : z(k + 1) * scale;
      double U0, U1, U2, rk;
      if (fabs(p) >= fabs(sk)) {
        if (fabs(p) < pert) p = p < 0.0 ? -pert : pert;
        const double mlt = sk / p;
        U0 = p;
        U1 = q;
        U2 = 0.0;
        rk = bp;
        p = dk1 - mlt * q;
        q = ek1;
        bp = bk1 - mlt * bp;
      } else {
        const double mlt = p / sk;
        U0 = sk;
        U1 = dk1;
        U2 = ek1;
        rk = bk1;
        p = q - mlt * dk1;
        q = -mlt * ek1;
        bp = bp - mlt * bk1;
      }
      u0[k] = 1.0 / U0;
      u1[k] = U1;
      u2[k] = U2;
      z(k) = rk;
    }
    if (fabs(p) < pert) p = p < 0.0 ? -pert : pert;
    double x2 = 0.0, x1 = bp / p, nrm = x1 * x1;
    z(n - 1) = x1;
    for (int k = n - 2; k >= 0; --k) {
      const double x = (z(k) - u1[k] * x1 - u2[k] * x2) * u0[k];
      z(k) = x;
      nrm += x * x;
      x2 = x1;
      x1 = x;
    }
    scale = 1.0 / sqrt(nrm);
  }
  for (int k = 0; k < n; ++k) z(k) *= scale;
}

// Eigenvectors of T for the ascending levels lam[0..n) into the rows of Zt
// (skewed, ld x ld; entries beyond n zero).  tnorm = |T|.  Isolated levels:
// one thread each, twisted factorisation at the bisection level, then at
// the Rayleigh quotient of that vector.  Levels closer than ORTOL |T| to a
// neighbour (rare): thread 0, inverse iteration from distinct start
// vectors, then modified Gram-Schmidt within the cluster (two passes).
// scr: 3 n doubles.
template <int NT>
__device__ __noinline__ void eigenvectors(const lds_dbl* __restrict__ d,
                                          const lds_dbl* __restrict__ e,
                                          const lds_dbl* __restrict__ lam, int n, int ld,
                                          lds_dbl* __restrict__ Zt, double tnorm,
                                          lds_dbl* __restrict__ scr)
{
  const int tid = threadIdx.x;
  for (int idx = tid; idx < ld * ld; idx += NT) Zt[idx] = 0.0;
  __syncthreads();
  const double ctol = ORTOL * tnorm;
  double emax = 0.0;
  for (int k = 0; k + 1 < n; ++k) emax = fmax(emax, e[k] * e[k]);
  const double pivmin = 2.2250738585072014e-308 * fmax(1.0, emax);
  auto in_cluster = [&](int i) {
    return (i > 0 && lam[i] - lam[i - 1] < ctol) || (i + 1 < n && lam[i + 1] - lam[i] < ctol);
  };
  for (int i = tid; i < n; i += NT) {
    if (in_cluster(i)) continue;
    auto z = [&](int k) -> lds_dbl& { return Zt[skx(i, k, ld)]; };
    double l = lam[i];
    for (int round = 0; round < 2; ++round) {
      const double nrm = twisted_vector(d, e, n, l, pivmin, 0, z);
      if (!(nrm > 0.0 && isfinite(nrm))) {
        for (int k = 0; k < n; ++k) z(k) = 0.0;
        z(i) = 1.0;  // not reached for an unreduced T; keeps the basis finite
        break;
      }
      const double inv = 1.0 / nrm;
      double rq = 0.0, zprev = 0.0;
      for (int k = 0; k < n; ++k) {
        const double zk = z(k) * inv;
        z(k) = zk;
        rq += d[k] * zk * zk;
        if (k > 0) rq += 2.0 * e[k - 1] * zprev * zk;
        zprev = zk;
      }
      l = rq;
    }
  }
  if (tid == 0) {
    const double pert = 2.220446049250313e-16 * tnorm + 1e-300;
    int i = 0;
    while (i < n) {
      int end = i + 1;
      while (end < n && lam[end] - lam[end - 1] < ctol) ++end;
      if (end - i > 1) {
        for (int a = i; a < end; ++a) {
          auto za = [&](int k) -> lds_dbl& { return Zt[skx(a, k, ld)]; };
          inverse_iteration_one(d, e, n, lam[a], pert, a, scr, za);
          for (int pass = 0; pass < 2; ++pass)
            for (int bq = i; bq < a; ++bq) {
              double s = 0.0;
              for (int k = 0; k < n; ++k) s += Zt[skx(bq, k, ld)] * za(k);
              for (int k = 0; k < n; ++k) za(k) -= s * Zt[skx(bq, k, ld)];
            }
          double s = 0.0;
          for (int k = 0; k < n; ++k) s += za(k) * za(k);
          const double inv = s > 0.0 ? 1.0 / sqrt(s) : 0.0;
          for (int k = 0; k < n; ++k) za(k) *= inv;
        }
      }
      i = end;
    }
  }
  __syncthreads();
}

// CT <- Zt * CT on the matrix cores (Zt skewed, CT plain, both ld x ld; the
// result replaces CT: every wave's tiles stay in registers until all waves
// have read).  Requires mfma-sized tiles to cover ld (MT tiles per wave).
template <int NT, int MT>
__device__ __noinline__ void rotate_basis(const lds_dbl* Zt, lds_dbl* CT, int ld)
{
  typedef double d4 __attribute__((ext_vector_type(4)));
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, w = wave_id();
  const int lr = lane & 15, lk = lane >> 4;
  const int nb = (ld + 15) >> 4, ntile = nb * nb;
  d4 acc[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    acc[t] = d4{0.0, 0.0, 0.0, 0.0};
    const int tile = w + t * NW;
    if (tile < ntile) {
      const int i0 = (tile / nb) * 16, j0 = (tile % nb) * 16;
      const bool oka = i0 + lr < ld, okb = j0 + lr < ld;
      const int ra = oka ? i0 + lr : 0, cb = okb ? j0 + lr : 0;
      for (int k = 0; k < ld; k += 4) {
        const double a = Zt[skx(ra, k + lk, ld)];
        const double bv = CT[(k + lk) * ld + cb];
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(oka ? a : 0.0, okb ? bv : 0.0, acc[t], 0, 0, 0);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    const int tile = w + t * NW;
    if (tile < ntile) {
      const int i0 = (tile / nb) * 16, j0 = (tile % nb) * 16;
      const int c = j0 + lr;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = i0 + 4 * i + lk;
        if (r < ld && c < ld) CT[r * ld + c] = acc[t][i];
      }
    }
  }
  __syncthreads();
}

}  // namespace ldseig

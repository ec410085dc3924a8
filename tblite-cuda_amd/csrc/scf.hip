// Self-consistent field loop, one workgroup per molecule.
//
// The workgroup runs the whole SCF of its molecule (next_scf,
// src/scf/iterators.cu:317-424; driver loop src/xtb.cu:413-433) without
// returning to the host: molecules that converge retire and free their CU
// slot while the rest of the batch keeps iterating.
//
// Per iteration:
//   potential     ES2 / AES2 / ES3 from the current charges
//   Fock          F = H0 - S.(v_i+v_j)/2 - (D.vdp + Q.vqp terms)/2   -> A (LDS)
//   transform     A <- C^T F C, C the S-orthonormal MO basis of the last step
//   eigensolve    parallel cyclic Jacobi on A, rotations applied to C
//                 (matrices in global memory: A <- X F X^T with X = L^-1,
//                 Householder tridiagonalisation, bisection, inverse
//                 iteration, back-transformation; see eig_big.h)
//   occupation    Fermi smearing on the sorted levels
//   density       P = C f C^T                                          -> A
//   Mulliken      shell charges, atomic dipoles / quadrupoles, band energy
//   energy        E = band + TS + ES2 + AES2 + ES3
//   mixing        modified Broyden on [qsh, dpat, qpat]
// C starts as L^-T from the Cholesky factor of S.  Because C carries the
// previous eigenvectors, A is nearly diagonal after the first iterations
// and the Jacobi solver needs only a few sweeps.
//
// A and C are ld x ld (ld = nao rounded up to a multiple of 4, padding
// zero); they sit in LDS for ld <= 96 and in global memory above that,
// where the matrices are plain row-major instead of skewed.
#include <hip/hip_runtime.h>
#include <math.h>

#include "aes2.h"
#include "common.h"
#include "d4.h"
#include "device_util.h"
#include "eig_band.h"
#include "eig_big.h"
#include "eig_lds.h"
#include "fermi.h"

#define JACOBI_MAX_SWEEPS 40

// A (Fock / working / density matrix) is stored skewed: element (i, j) at
// i*ld + (i + j) mod ld.  Rows and columns then both spread over the LDS
// banks (a plain row stride of ld = 0 mod 64 banks would serialise column
// accesses).  CT (MO coefficients, one MO per row) is stored plainly.
__device__ __forceinline__ int sk(int i, int j, int ld)
{
  int c = i + j;
  c = c >= ld ? c - ld : c;
  return i * ld + c;
}

struct MolView {
  int m, n, ld, nat, nsh, a0, s0, o0, ndim, mem;
  const int* el;
  const int* sh_at;
  const int* sh_ish;
  const int* sh_iao;
  const int* ao_at;
  const int* ao_sh;
  const int* at_ish;
  const int* at_iao;
  const double* S;
  const double* H;
  const double* D;
  const double* Q;
  const double* gam;
  const double* xyz;   // atom positions and multipole radii (AES2 pair kernels)
  const double* mrad;
  // periodic cell: the Ewald-summed AES2 planes (row a, column i), nat^2 each
  const double* aes_sd;
  const double* aes_dd;
  const double* aes_sq;
  double* qsh;
  double* qat;
  double* dpat;
  double* qpat;
};

// ------------------------------------------------------------- GEMM tiles
// out = X * Y (TRANS: out = (X * Y)^T) on ld x ld matrices with 4x4 register
// tiles; X plain row-major, Y and out skewed.  A thread's tile is rows
// r0..r0+3 and the columns c, c + nb, c + 2 nb, c + 3 nb (nb = ld / 4), so
// neighbouring threads read neighbouring elements of Y (no LDS bank
// conflicts) and share the rows of X (a broadcast).  When HOLD is set the
// output aliases an input: a thread's tiles stay in registers until every
// thread has read.
template <int NT, int MAXT, bool HOLD, bool TRANS>
__device__ void mm_nn(const double* X, const double* Y, double* out, int ld)
{
  const int nb = ld >> 2;
  const int ntile = nb * nb;
  auto tile_mm = [&](int tile, double* acc) {
    const int r0 = (tile / nb) * 4, c0 = tile % nb;
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.0;
    for (int k = 0; k < ld; ++k) {
      double x[4], y[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) x[u] = X[(r0 + u) * ld + k];
#pragma unroll
      for (int v = 0; v < 4; ++v) y[v] = Y[sk(k, c0 + v * nb, ld)];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u * 4 + v] += x[u] * y[v];
    }
  };
  auto tile_store = [&](int tile, const double* acc) {
    const int r0 = (tile / nb) * 4, c0 = tile % nb;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        if constexpr (TRANS) out[sk(c0 + v * nb, r0 + u, ld)] = acc[u * 4 + v];
        else out[sk(r0 + u, c0 + v * nb, ld)] = acc[u * 4 + v];
      }
  };
  if constexpr (HOLD) {
    double acc[MAXT][16];
#pragma unroll
    for (int t = 0; t < MAXT; ++t) {
      const int tile = threadIdx.x + t * NT;
      if (tile < ntile) tile_mm(tile, acc[t]);
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < MAXT; ++t) {
      const int tile = threadIdx.x + t * NT;
      if (tile < ntile) tile_store(tile, acc[t]);
    }
  } else {
    for (int tile = threadIdx.x; tile < ntile; tile += NT) {
      double acc[16];
      tile_mm(tile, acc);
      tile_store(tile, acc);
    }
  }
  __syncthreads();
}

// The same products on the FP64 matrix cores (v_mfma_f64_16x16x4f64: lane l
// supplies a[l%16][l/16] and b[l/16][l%16] and holds d[4i + l/16][l%16] of
// the 16 x 16 result), operands read straight from LDS.  Each wave owns
// 16 x 16 output tiles (tile w, w + NW, ...); at most MFMA_TILES(NT) per wave,
// enough for the largest ld of each instantiation (128 threads: ld <= 32,
// 256: ld <= 64, 512: ld <= 96).  Rows and columns beyond ld read index 0
// and are masked to zero.
#define MFMA_TILES(NT) ((NT) >= 512 ? 5 : ((NT) >= 256 ? 4 : 2))
typedef double scf_double4 __attribute__((ext_vector_type(4)));

template <int NT>
__device__ __forceinline__ bool mfma_fits(int ld)
{
  const int nb = (ld + 15) >> 4;
  return nb * nb <= MFMA_TILES(NT) * (NT / 64);
}

// out = X * Y (TRANS: out = (X * Y)^T); X plain row-major, Y and out skewed.
// With HOLD the output aliases an input: every wave's tiles stay in
// registers until all waves have read.
template <int NT, bool HOLD, bool TRANS>
__device__ void mm_mfma(const double* X, const double* Y, double* out, int ld)
{
  constexpr int NW = NT / 64;
  constexpr int MT = MFMA_TILES(NT);
  const int lane = threadIdx.x & 63, w = wave_id();
  const int lr = lane & 15, lk = lane >> 4;
  const int nb = (ld + 15) >> 4, ntile = nb * nb;
  scf_double4 acc[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    acc[t] = scf_double4{0.0, 0.0, 0.0, 0.0};
    const int tile = w + t * NW;
    if (tile < ntile) {
      const int i0 = (tile / nb) * 16, j0 = (tile % nb) * 16;
      const bool oka = i0 + lr < ld, okb = j0 + lr < ld;
      const int ra = oka ? i0 + lr : 0, cb = okb ? j0 + lr : 0;
      for (int k = 0; k < ld; k += 4) {
        const double a = X[ra * ld + k + lk];
        const double bv = Y[sk(k + lk, cb, ld)];
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(oka ? a : 0.0, okb ? bv : 0.0, acc[t], 0, 0, 0);
      }
    }
  }
  if constexpr (HOLD) __syncthreads();
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    const int tile = w + t * NW;
    if (tile < ntile) {
      const int i0 = (tile / nb) * 16, j0 = (tile % nb) * 16;
      const int c = j0 + lr;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = i0 + 4 * i + lk;
        if (r < ld && c < ld) {
          if constexpr (TRANS) out[sk(c, r, ld)] = acc[t][i];
          else out[sk(r, c, ld)] = acc[t][i];
        }
      }
    }
  }
  __syncthreads();
}

// P = sum_t f_t CT[occ_t, :]^T CT[occ_t, :] on the matrix cores; the operands
// are rows of CT (consecutive lanes read consecutive elements)
template <int NT>
__device__ void density_mfma(const double* CT, const double* f, const int* occ, int nocc,
                             double* P, int ld)
{
  constexpr int NW = NT / 64;
  constexpr int MT = MFMA_TILES(NT);
  const int lane = threadIdx.x & 63, w = wave_id();
  const int lr = lane & 15, lk = lane >> 4;
  const int nb = (ld + 15) >> 4, ntile = nb * nb;
  // k outer, tiles inner: the MO row and its occupation are read once per
  // k step for all of the wave's tiles, whose accumulation chains overlap
  scf_double4 acc[MT];
  int ci[MT], cj[MT];
  bool oki[MT], okj[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    acc[t] = scf_double4{0.0, 0.0, 0.0, 0.0};
    const int tile = w + t * NW;
    const int i0 = (tile / nb) * 16, j0 = (tile % nb) * 16;
    oki[t] = i0 + lr < ld;
    okj[t] = j0 + lr < ld;
    ci[t] = oki[t] ? i0 + lr : 0;
    cj[t] = okj[t] ? j0 + lr : 0;
  }
  for (int k = 0; k < nocc; k += 4) {
    const bool okt = k + lk < nocc;
    const int row = occ[okt ? k + lk : 0];
    const double* cr = CT + row * ld;
    const double fr = f[row];
    double a[MT], bv[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t)
      if (w + t * NW < ntile) {
        a[t] = cr[ci[t]] * fr;
        bv[t] = cr[cj[t]];
      }
#pragma unroll
    for (int t = 0; t < MT; ++t)
      if (w + t * NW < ntile)
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64((okt && oki[t]) ? a[t] : 0.0,
                                                      (okt && okj[t]) ? bv[t] : 0.0, acc[t], 0, 0, 0);
  }
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    const int tile = w + t * NW;
    if (tile >= ntile) continue;
    const int i0 = (tile / nb) * 16, j0 = (tile % nb) * 16;
    const int c = j0 + lr;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = i0 + 4 * i + lk;
      if (r < ld && c < ld) P[sk(r, c, ld)] = acc[t][i];
    }
  }
  __syncthreads();
}

// P = sum_t f_t CT[occ_t, :]^T CT[occ_t, :] over the occupied MO rows
// (get_density_matrix, src/wavefunction/type.cu:17-35 text)
template <int NT>
__device__ void density(const double* CT, const double* f, const int* occ, int nocc, double* P,
                        int ld)
{
  // tiles as in mm_nn: rows r0..r0+3, columns c, c + nb, c + 2 nb, c + 3 nb
  const int nb = ld >> 2;
  for (int tile = threadIdx.x; tile < nb * nb; tile += NT) {
    const int r0 = (tile / nb) * 4, c0 = tile % nb;
    double acc[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.0;
    for (int t = 0; t < nocc; ++t) {
      const int k = occ[t];
      const double fk = f[k];
      const double* row = CT + k * ld;
      double x[4], y[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) x[u] = row[r0 + u] * fk;
#pragma unroll
      for (int v = 0; v < 4; ++v) y[v] = row[c0 + v * nb];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u * 4 + v] += x[u] * y[v];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) P[sk(r0 + u, c0 + v * nb, ld)] = acc[u * 4 + v];
  }
  __syncthreads();
}

// ------------------------------------------------------------ Cholesky
// S = L L^T in place (lower), then C = L^-T (stored as its transpose L^-1,
// one start MO per row).  Returns false if S is not positive definite.  col:
// LDS scratch, n doubles (2 ld with the matrices in LDS).  Right-looking,
// the trailing lower triangle updated one wave per row with lanes over its
// columns (coalesced along the row).  L^-1 is formed by forward
// substitution: with the matrices in LDS by columns of L^-1, four lanes per
// column (no barriers); in global memory row by row (row i of L in LDS, all
// columns j <= i in parallel, reads of earlier rows coalesced).
template <int NT, bool MATS_LDS>
__device__ bool orthonormalizer(double* A, double* C, int n, int ld, double* col)
{
  constexpr int NW = NT / 64;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  // Right-looking L D L^T elimination without square roots, one barrier per
  // column: column j and the pivot d_j are final when step j starts and step
  // j writes only columns c > j, A[r][c] -= A[r][j] A[c][j] / d_j
  // (j < c <= r).  Then L = (unit lower part) D^1/2: L[i][j] = A[i][j] /
  // sqrt(d_j), L[j][j] = sqrt(d_j).
  for (int j = 0; j < n; ++j) {
    const double dj = A[sk(j, j, ld)];
    if (!(dj > 0.0)) return false;  // the same value in every thread
    const double rdj = 1.0 / dj;
#pragma unroll 2
    for (int r = j + 1 + wid; r < n; r += NW) {
      const double lr = A[sk(r, j, ld)] * rdj;
      for (int c = j + 1 + lane; c <= r; c += 64) A[sk(r, c, ld)] -= lr * A[sk(c, j, ld)];
    }
    __syncthreads();
  }
  for (int j = tid; j < n; j += NT) col[j] = sqrt(A[sk(j, j, ld)]);
  __syncthreads();
  for (int idx = tid; idx < n * n; idx += NT) {
    const int i = idx / n, j = idx - (idx / n) * n;
    if (i > j) A[sk(i, j, ld)] /= col[j];
    else if (i == j) A[sk(i, i, ld)] = col[i];
  }
  __syncthreads();
  for (int idx = tid; idx < ld * ld; idx += NT) C[idx] = 0.0;
  __syncthreads();
  if constexpr (MATS_LDS) {
    // C = L^-1 by columns, four adjacent lanes per column: lane t sums the
    // terms k = j + t (mod 4) of C[i][j] = (delta_ij - sum_k L[i][k] C[k][j])
    // / L[i][i], two butterfly steps give every lane the same total, and lane
    // (i - j) % 4 stores C[i][j], so a lane reads only the elements it stored
    double* rdg = col + ld;
    for (int i = tid; i < n; i += NT) rdg[i] = 1.0 / col[i];
    __syncthreads();
    for (int j = tid >> 2; j < n; j += NT >> 2) {
      const int t = tid & 3;
      for (int i = j; i < n; ++i) {
        double s = 0.0;
        for (int k = j + t; k < i; k += 4) s -= A[sk(i, k, ld)] * C[k * ld + j];
        s += __shfl_xor(s, 1, 4);
        s += __shfl_xor(s, 2, 4);
        if (((i - j) & 3) == t) C[i * ld + j] = ((i == j ? 1.0 : 0.0) + s) * rdg[i];
      }
    }
  } else {
    for (int i = 0; i < n; ++i) {
      for (int k = tid; k <= i; k += NT) col[k] = A[sk(i, k, ld)];
      __syncthreads();
      const double dii = col[i];
      for (int j = tid; j <= i; j += NT) {
        double s = (i == j) ? 1.0 : 0.0;
        for (int k = j; k < i; ++k) s -= col[k] * C[k * ld + j];
        C[i * ld + j] = s / dii;
      }
      __syncthreads();
    }
  }
  __syncthreads();
  return true;
}

// ------------------------------------------------ blocked orthonormalizer
// Global-memory path: S (in A, row-major, lower triangle read) = L L^T with
// the blocks of L below the diagonal left in A, and C = L^-1 (row-major, zero
// above the diagonal), in panels of 64 columns:
//   diagonal block  Cholesky in LDS (one barrier per column: the block's
//                   trailing triangle takes a[i][c] -= a[i][j] a[c][j] / d_j,
//                   then L = (unit lower part) D^1/2), then its inverse in
//                   place, columns last to first (x[j][j] = 1 / l[j][j],
//                   x[i][j] = -x[j][j] sum_{k=j+1}^{i} x[i][k] l[k][j], eight
//                   lanes per row), into the diagonal block of C;
//   below           L21 = A21 X11^T, then the trailing update
//                   A22 -= L21 L21^T (lower tiles), on the matrix cores.
// Then the blocks of L^-1 below the diagonal, subdiagonal d = 1, 2, .. at a
// time: X_IJ = -X_II (sum_{K=J}^{I-1} L_IK X_KJ), I = J + d (the sum into
// C's block first), on the matrix cores.  The GEMM tiles are the 64 x 64
// blocks.  Returns false if S is not positive definite.  lds: 4224 doubles
// (diagonal block and its pivots; the products' tile overlays them).
template <int NT>
__device__ __noinline__ bool orthonormalizer_blocked(gbl_dbl* A, gbl_dbl* C, int n, int ld,
                                                     lds_dbl* lds)
{
  constexpr int NB = 64, TS = 65;
  const int tid = threadIdx.x;
  const int nb = (n + NB - 1) / NB;
  lds_dbl* T = lds;
  lds_dbl* dg = lds + NB * TS;
  lds_dbl* tile = lds;
  auto no_skip = [](int, int) { return false; };
  for (int idx = tid; idx < ld * ld; idx += NT) C[idx] = 0.0;
  for (int p = 0; p < nb; ++p) {
    const int c0 = NB * p, w = min(NB, n - c0), r0 = c0 + w, m = n - r0;
    __syncthreads();  // the previous panel's products are done with the tile
    for (int idx = tid; idx < NB * NB; idx += NT) {
      const int i = idx >> 6, j = idx & 63;
      T[i * TS + j] = (i < w && j <= i) ? A[(int64_t)(c0 + i) * ld + c0 + j] : 0.0;
    }
    __syncthreads();
    for (int j = 0; j < w; ++j) {
      const double dj = T[j * TS + j];
      if (!(dj > 0.0)) return false;  // the same value in every thread
      const double rdj = 1.0 / dj;
      for (int idx = tid; idx < (w - j - 1) * NB; idx += NT) {
        const int i = j + 1 + (idx >> 6), c = idx & 63;
        if (c > j && c <= i) T[i * TS + c] -= T[i * TS + j] * T[c * TS + j] * rdj;
      }
      __syncthreads();
    }
    if (tid < w) dg[tid] = sqrt(T[tid * TS + tid]);
    __syncthreads();
    for (int idx = tid; idx < NB * NB; idx += NT) {
      const int i = idx >> 6, j = idx & 63;
      if (i < w && j < i) T[i * TS + j] /= dg[j];
      else if (i < w && j == i) T[i * TS + i] = dg[i];
    }
    __syncthreads();
    {
      const int g = tid >> 3, q = tid & 7;
      for (int j = w - 1; j >= 0; --j) {
        const int i = j + 1 + g;
        double s = 0.0;
        if (i < w)
          for (int k = j + 1 + q; k <= i; k += 8) s += T[i * TS + k] * T[k * TS + j];
        s += dpp_f64<0xB1>(s);
        s += dpp_f64<0x4E>(s);
        s += dpp_f64<0x141>(s);
        const double xjj = 1.0 / T[j * TS + j];
        __syncthreads();  // column j (still L) read by every row
        if (q == 0 && i < w) T[i * TS + j] = -xjj * s;
        if (tid == 0) T[j * TS + j] = xjj;
        __syncthreads();
      }
    }
    for (int idx = tid; idx < NB * NB; idx += NT) {
      const int i = idx >> 6, j = idx & 63;
      if (i < w && j <= i) C[(int64_t)(c0 + i) * ld + c0 + j] = T[i * TS + j];
    }
    __syncthreads();
    if (m <= 0) continue;
    auto kpanel = [w](int, int, int* ka, int* kb) {
      *ka = 0;
      *kb = w;
    };
    // L21 = A21 X11^T in place (a tile reads all of its rows before it stores)
    big::gemm<NT, true, false, 32>(
        m, w, [=](int i, int k) { return A[(int64_t)(r0 + i) * ld + c0 + k]; },
        [=](int k, int j) { return C[(int64_t)(c0 + j) * ld + c0 + k]; }, kpanel, no_skip,
        [=](int i, int j, double v) { A[(int64_t)(r0 + i) * ld + c0 + j] = v; }, tile);
    // A22 -= L21 L21^T (tiles on and below the diagonal)
    big::gemm<NT, true, false, 32>(
        m, m, [=](int i, int k) { return A[(int64_t)(r0 + i) * ld + c0 + k]; },
        [=](int k, int j) { return A[(int64_t)(r0 + j) * ld + c0 + k]; }, kpanel,
        [](int ti, int tj) { return tj > ti; },
        [=](int i, int j, double v) { A[(int64_t)(r0 + i) * ld + r0 + j] -= v; }, tile);
  }
  for (int d = 1; d < nb; ++d) {
    auto sub = [d](int ti, int tj) { return ti - tj != d; };
    big::gemm<NT, true, true, 32>(
        n, n, [=](int i, int k) { return A[(int64_t)i * ld + k]; },
        [=](int k, int j) { return C[(int64_t)k * ld + j]; },
        [](int ti, int tj, int* ka, int* kb) {
          *ka = NB * tj;
          *kb = NB * ti;
        },
        sub, [=](int i, int j, double v) { C[(int64_t)i * ld + j] = v; }, tile);
    big::gemm<NT, true, true, 32>(
        n, n, [=](int i, int k) { return C[(int64_t)i * ld + k]; },
        [=](int k, int j) { return C[(int64_t)k * ld + j]; },
        [n](int ti, int, int* ka, int* kb) {
          *ka = NB * ti;
          *kb = min(NB * ti + NB, n);
        },
        sub, [=](int i, int j, double v) { C[(int64_t)i * ld + j] = -v; }, tile);
  }
  __syncthreads();
  return true;
}

// ---------------------------------------------------------------- Jacobi
// Cyclic parallel Jacobi with round-robin pairing on the symmetric matrix A
// (upper triangle kept during the sweeps), rotations accumulated into the
// rows of CT (the MO coefficients, one MO per row).  Each round applies ld/2
// disjoint plane rotations, one per pair slot k = (position k, position
// ld-1-k).
//
// One barrier per round.  Under round-robin pairing every pair of the next
// round joins one element of slot k-1 with one of slot k+1 (slot 0 with slot
// 1, the last slot with the one before), so the 2x2 block of A holding that
// pair's new off-diagonal element is known in advance.  Wave 0 owns these
// "owner" blocks and the diagonal blocks: it updates them and computes the
// next round's rotations from the rotation records of the current round
// (which carry the pre-rotation diagonal elements).  The other waves update
// the remaining blocks of A and the rows of CT.  Records are double buffered.
//
// Work items of the worker waves are fixed for the whole SCF (slots do not
// move between rounds, only the indices they hold) and precomputed once.
// With MAXB == 0 they are decoded on the fly (global-memory path for large
// molecules).
//
// Convergence: a sweep whose largest rotated off-diagonal element is below
// a tolerance tol leaves off-diagonals of order tol^2 / gap (quadratic
// convergence), so the solve stops after that sweep.
struct JRot {
  int2* pq;      // [2][h] indices (p, q) held by the slot
  double2* cs;   // [2][h] cos, sin
  double2* ta;   // [2][h] tan, a_pq before the rotation
  double2* dg;   // [2][h] a_pp, a_qq before the rotation
};

// CT rows: JACOBI_TPS(NT) worker threads share one slot's row pair, thread t
// of the group taking the two-AO pieces t, t + TPS, ... (up to MAXC of them),
// so a thread reads its slot's rotation record once per round and a group's
// lanes read contiguous 16-byte pieces.  Coverage: TPS * MAXC >= ld / 2 and
// TPS * ld / 2 <= NT - 64 for the largest ld of each instantiation
// (128 threads: ld <= 32, 256: ld <= 64, 512: ld <= 96).
#define JACOBI_TPS(NT) ((NT) >= 512 ? 8 : ((NT) >= 256 ? 6 : 4))

template <int MAXB, int MAXC>
struct JPlan {
  int bkl[MAXB > 0 ? MAXB : 1];  // worker block items (k << 16 | l), -1 for none
  int cslot;                     // worker CT slot, -1 for none
  int cpc;                       // first two-AO piece of the slot's rows
};

__device__ __forceinline__ bool jacobi_owner_block(int k, int l, int h)
{
  return (k == 0 && l == 1) || (l == k + 2 && k + 1 <= h - 2) || (k == h - 2 && l == h - 1);
}

template <int NT, int MAXB, int MAXC>
__device__ void jacobi_plan(JPlan<MAXB, MAXC>& pl, int ld)
{
  if constexpr (MAXB > 0) {
    const int h = ld >> 1;
    const int nw = NT - 64, w = (int)threadIdx.x - 64;
#pragma unroll
    for (int j = 0; j < MAXB; ++j) pl.bkl[j] = -1;
    pl.cslot = -1;
    pl.cpc = 0;
    if (w < 0) return;
    // consecutive workers take consecutive blocks (neighbouring columns of A)
    const int nblk = h * (h - 1) / 2;
#pragma unroll
    for (int j = 0; j < MAXB; ++j) {
      const int bi = w + j * nw;
      if (bi >= nblk) continue;
      int k = 0, rem = bi;
      while (rem >= h - 1 - k) {
        rem -= h - 1 - k;
        ++k;
      }
      const int l = k + 1 + rem;
      if (!jacobi_owner_block(k, l, h)) pl.bkl[j] = (k << 16) | l;
    }
    // CT pieces cover two adjacent AOs (one 16-byte LDS access per row)
    constexpr int TPS = JACOBI_TPS(NT);
    static_assert(MAXC == 0 || TPS * MAXC >= (NT >= 512 ? 48 : (NT >= 256 ? 32 : 16)),
                  "CT pieces per slot not covered");
    if (w / TPS < h) {
      pl.cslot = w / TPS;
      pl.cpc = w % TPS;
    }
  }
}

__device__ __forceinline__ int rr_player(int pos, int r, int ld)
{
  return pos == 0 ? 0 : 1 + ((pos - 1 + r) % (ld - 1));
}

__device__ __forceinline__ void make_rotation(const JRot& R, int buf, int h, int slot, int p,
                                              int q, double app, double aqq, double apq,
                                              double skip)
{
  double c = 1.0, sn = 0.0, t = 0.0;
  if (fabs(apq) > skip) {
    // t = 2 apq sgn(d) / (|d| + sqrt(d^2 + 4 apq^2)), the smaller rotation angle
    const double d = aqq - app;
    const double a2 = 2.0 * apq;
    t = a2 / (fabs(d) + sqrt(d * d + a2 * a2));
    if (d < 0.0) t = -t;
    c = rsqrt(t * t + 1.0);
    sn = t * c;
  }
  const int o = buf * h + slot;
  R.pq[o] = make_int2(p, q);
  R.cs[o] = make_double2(c, sn);
  R.ta[o] = make_double2(t, apq);
  R.dg[o] = make_double2(app, aqq);
}

// upper-triangle element (i != j)
__device__ __forceinline__ double& aup(double* A, int ld, int i, int j)
{
  return i < j ? A[sk(i, j, ld)] : A[sk(j, i, ld)];
}

template <int NT, int MAXB, int MAXC>
__device__ bool jacobi(double* A, double* CT, int ld, const JRot& R, const JPlan<MAXB, MAXC>& pl,
                       double* red, double tol, double skip, int* sweeps_out)
{
  const int h = ld >> 1;
  const int lane = threadIdx.x & 63;
  const bool owner_wave = threadIdx.x < 64;
  // rotations of the first round
  if (owner_wave)
    for (int k = lane; k < h; k += 64) {
      const int p = rr_player(k, 0, ld), q = rr_player(ld - 1 - k, 0, ld);
      make_rotation(R, 0, h, k, p, q, A[sk(p, p, ld)], A[sk(q, q, ld)], aup(A, ld, p, q), skip);
    }
  __syncthreads();

  int sweep = 0, round = 0;
  bool ok = false;
  for (; sweep < JACOBI_MAX_SWEEPS; ++sweep) {
    double offmax = 0.0;
    for (int r = 0; r < ld - 1; ++r, ++round) {
      const int co = (round & 1) * h, nxt = (round & 1) ^ 1;
      // 2x2 block (rows of slot k, columns of slot l) <- G_k^T block G_l
      auto block_rot = [&](const int2 pk, const double2 ck, const int2 pl_, const double2 cl,
                           double& z00, double& z01, double& z10, double& z11) {
        const double x00 = aup(A, ld, pk.x, pl_.x), x01 = aup(A, ld, pk.x, pl_.y);
        const double x10 = aup(A, ld, pk.y, pl_.x), x11 = aup(A, ld, pk.y, pl_.y);
        const double y00 = ck.x * x00 - ck.y * x10, y01 = ck.x * x01 - ck.y * x11;
        const double y10 = ck.y * x00 + ck.x * x10, y11 = ck.y * x01 + ck.x * x11;
        z00 = cl.x * y00 - cl.y * y01;
        z01 = cl.y * y00 + cl.x * y01;
        z10 = cl.x * y10 - cl.y * y11;
        z11 = cl.y * y10 + cl.x * y11;
        aup(A, ld, pk.x, pl_.x) = z00;
        aup(A, ld, pk.x, pl_.y) = z01;
        aup(A, ld, pk.y, pl_.x) = z10;
        aup(A, ld, pk.y, pl_.y) = z11;
      };
      auto block = [&](int k, int l, double& z00, double& z01, double& z10, double& z11,
                       int2& pk, int2& pl_) {
        pk = R.pq[co + k];
        pl_ = R.pq[co + l];
        block_rot(pk, R.cs[co + k], pl_, R.cs[co + l], z00, z01, z10, z11);
      };
      if (owner_wave) {
        // diagonal element after this round of the first (e = 0) or second index of slot k
        auto diag_after = [&](int k, int e) {
          const double2 ta = R.ta[co + k], dg = R.dg[co + k];
          return e == 0 ? dg.x - ta.x * ta.y : dg.y + ta.x * ta.y;
        };
        for (int kn = lane; kn < h; kn += 64) {
          {
            const int2 pq = R.pq[co + kn];
            const double2 ta = R.ta[co + kn], dg = R.dg[co + kn];
            offmax = fmax(offmax, fabs(ta.y));
            A[sk(pq.x, pq.x, ld)] = dg.x - ta.x * ta.y;
            A[sk(pq.y, pq.y, ld)] = dg.y + ta.x * ta.y;
            aup(A, ld, pq.x, pq.y) = 0.0;
          }
          // The block whose new off-diagonal element is slot kn's pivot next
          // round: slot 0 takes element 00 of block (0, 1), a middle slot
          // element 10 of block (kn-1, kn+1), the last slot element 11 of
          // block (h-2, h-1).  One code path for every lane (no divergence:
          // the owner wave is the critical path of a round).
          const int e = kn == 0 ? 0 : (kn <= h - 2 ? 1 : 2);
          const int bk = e == 0 ? 0 : (e == 1 ? kn - 1 : h - 2);
          const int bl = e == 0 ? 1 : (e == 1 ? kn + 1 : h - 1);
          double z00, z01, z10, z11;
          int2 pk, pl_;
          block(bk, bl, z00, z01, z10, z11, pk, pl_);
          const int p = e == 0 ? pk.x : (e == 1 ? pl_.x : pl_.y);
          const int q = e == 0 ? pl_.x : pk.y;
          const double apq = e == 0 ? z00 : (e == 1 ? z10 : z11);
          // diagonal elements after this round: slot/index holding p and q
          const int sp = e == 0 ? 0 : (e == 1 ? kn + 1 : h - 1), ip = e == 2 ? 1 : 0;
          const int sq = e == 0 ? 1 : (e == 1 ? kn - 1 : h - 2), iq = e == 0 ? 0 : 1;
          make_rotation(R, nxt, h, kn, p, q, diag_after(sp, ip), diag_after(sq, iq), apq, skip);
        }
      } else {
        auto crow = [&](int k, int ao) {
          const int2 pq = R.pq[co + k];
          const double2 cs = R.cs[co + k];
          double* rp = CT + pq.x * ld + ao;
          double* rq = CT + pq.y * ld + ao;
          const double u = *rp, v = *rq;
          *rp = cs.x * u - cs.y * v;
          *rq = cs.y * u + cs.x * v;
        };
        double z00, z01, z10, z11;
        int2 pk, pl_;
        if constexpr (MAXB > 0) {
#pragma unroll
          for (int j = 0; j < MAXB; ++j)
            if (pl.bkl[j] >= 0) {
              const int k = pl.bkl[j] >> 16, l = pl.bkl[j] & 0xffff;
              // both rotations identities: the block is unchanged
              if (R.cs[co + k].y == 0.0 && R.cs[co + l].y == 0.0) continue;
              block(k, l, z00, z01, z10, z11, pk, pl_);
            }
          // the slot's row pair: one record read, then every piece's loads
          // before any store
          if (pl.cslot >= 0) {
            constexpr int TPS = JACOBI_TPS(NT);
            const double2 cs = R.cs[co + pl.cslot];
            if (cs.y != 0.0) {  // identity rotation: rows unchanged
              const int2 pq = R.pq[co + pl.cslot];
              double2* rp = reinterpret_cast<double2*>(CT + pq.x * ld);
              double2* rq = reinterpret_cast<double2*>(CT + pq.y * ld);
              double2 u[MAXC], v[MAXC];
#pragma unroll
              for (int j = 0; j < MAXC; ++j) {
                const int c2 = pl.cpc + j * TPS;
                if (c2 < h) {
                  u[j] = rp[c2];
                  v[j] = rq[c2];
                }
              }
#pragma unroll
              for (int j = 0; j < MAXC; ++j) {
                const int c2 = pl.cpc + j * TPS;
                if (c2 < h) {
                  rp[c2] = make_double2(cs.x * u[j].x - cs.y * v[j].x, cs.x * u[j].y - cs.y * v[j].y);
                  rq[c2] = make_double2(cs.y * u[j].x + cs.x * v[j].x, cs.y * u[j].y + cs.x * v[j].y);
                }
              }
            }
          }
        } else {
          const int nw = NT - 64, w = (int)threadIdx.x - 64;
          const int nblk = h * (h - 1) / 2;
          for (int bi = w; bi < nblk; bi += nw) {
            int k = 0, rem = bi;
            while (rem >= h - 1 - k) {
              rem -= h - 1 - k;
              ++k;
            }
            const int l = k + 1 + rem;
            if (!jacobi_owner_block(k, l, h)) block(k, l, z00, z01, z10, z11, pk, pl_);
          }
          for (int ci = w; ci < h * ld; ci += nw) crow(ci / ld, ci % ld);
        }
      }
      __syncthreads();
    }
    offmax = block_max(offmax, red);
    if (offmax < tol) {
      ok = true;
      ++sweep;
      break;
    }
  }
  *sweeps_out = sweep;
  return ok;
}


// ------------------------------------------- two-stage reduction
// A (lower triangle) -> tridiagonal d, e (LDS) through a band of half-width
// 16; the stage-1 and stage-2 reflectors stay in the molecule's scratch for
// big_eigenvectors.  tile: band::TILE_DOUBLES of LDS, sm: 560 doubles of LDS.
template <int NT>
__device__ __noinline__ void big_band_tridiagonalize(gbl_dbl* A, gbl_dbl* scr, int n, int ld,
                                                     lds_dbl* d, lds_dbl* e, lds_dbl* tile,
                                                     lds_dbl* sm, lds_dbl* red, int region,
                                                     PhaseClock& clk)
{
  // stage-1 step clocks in red[16..20] -> the tri_* slots
  if (threadIdx.x == 0)
    for (int q = 0; q < 5; ++q) red[16 + q] = 0.0;
  __syncthreads();
  if (band::band_qb(n) == 4)
    band::band_reduce<NT, 4>(A, n, ld, scr + band_v1_offset(ld), scr + band_tau_offset(ld),
                             scr + band_t1_offset(ld), scr + band_w_offset(ld),
                             scr + band_x_offset(ld), tile, sm, clk.on ? red + 16 : nullptr);
  else
    band::band_reduce<NT, 8>(A, n, ld, scr + band_v1_offset(ld), scr + band_tau_offset(ld),
                             scr + band_t1_offset(ld), scr + band_w_offset(ld),
                             scr + band_x_offset(ld), tile, sm, clk.on ? red + 16 : nullptr);
  if (clk.on) {
    for (int q = 0; q < 5; ++q) clk.acc[TA_PROF_TRI_COLUMN + q] += (long long)red[16 + q];
  }
  clk.lap(TA_PROF_BAND1);
  // the chase on an LDS copy of the band (32 n doubles after d and e, then
  // its progress counters) when it fits the eigensolver region (region
  // doubles from d), else on A
  if (2 * ld + 32 * n + 8 <= region) {
    lds_dbl* band = d + 2 * ld;
    band::band_chase<NT>(A, n, ld, scr + band_v2_offset(ld), d, e, (lds_int*)(band + 32 * n), band);
  } else {
    band::band_chase<NT>(A, n, ld, scr + band_v2_offset(ld), d, e, (lds_int*)(sm + 532));
  }
  clk.lap(TA_PROF_BAND2);
}

// ------------------------------------------- global-memory eigensolve
// F (in A, full) -> eigenvalues ev (ascending) of the generalized problem
// F C = S C e, with X = L^-1 (lower triangular): A = X F X^T, reduced to
// tridiagonal form in place (reflectors in the lower triangle of A).  The
// eigensolver region er (LDS) holds d, e, tau, e^2 and the work vectors of
// the tridiagonalisation, followed by the GEMM tiles / reflector block; d, e
// and tau stay there for big_eigenvectors.  |T| ends in scal[0].  npart:
// per-wave column-sum partials of the tridiagonalisation, npart * ld doubles
// after the work vectors; the LDS layout reserves big_npart(ldmax) * ldmax,
// so npart follows the bucket's largest molecule, not this one.
template <int NT>
__device__ __noinline__ void big_eigenvalues(double* A_, const double* X_, double* Y_,
                                             double* scr_, int n, int ld, int npart, double* er_,
                                             double* ev_, double* red_, double* scal_,
                                             const double* prev_, double* warm_, bool two_stage,
                                             int eig_region, PhaseClock& clk)
{
  gbl_dbl* A = TA_GBL(A_);
  const gbl_dbl* X = TA_GBL(X_);
  gbl_dbl* Y = TA_GBL(Y_);
  gbl_dbl* scr = TA_GBL(scr_);
  lds_dbl* er = TA_LDS(er_);
  lds_dbl* ev = TA_LDS(ev_);
  lds_dbl* red = TA_LDS(red_);
  lds_dbl* scal = TA_LDS(scal_);
  lds_dbl* d = er;
  lds_dbl* e = er + ld;
  lds_dbl* tau = er + 2 * ld;
  lds_dbl* e2 = er + 3 * ld;
  lds_dbl* vv = er + 4 * ld;
  lds_dbl* ww = er + 5 * ld;
  lds_dbl* vn = er + 6 * ld;
  lds_dbl* yy = er + 7 * ld;
  lds_dbl* yrow = er + 8 * ld;
  lds_dbl* tmp = er + 9 * ld;
  auto k_lower = [n](int ti, int, int* k0, int* k1) {
    *k0 = 0;
    *k1 = min(n, ti * 64 + 64);
  };
  auto no_skip = [](int, int) { return false; };
  auto upper_tile = [](int ti, int tj) { return tj > ti; };
  // U = X F, stored transposed in Y; A = X U^T = X F X^T (lower tiles); 32-deep
  // k slices (tmp holds 4096 doubles)
  big::gemm<NT, true, true, 32>(
      n, n, [=](int i, int k) { return X[(int64_t)i * ld + k]; },
      [=](int k, int j) { return A[(int64_t)k * ld + j]; }, k_lower, no_skip,
      [=](int i, int j, double val) { Y[(int64_t)j * ld + i] = val; }, tmp);
  big::gemm<NT, true, true, 32>(
      n, n, [=](int i, int k) { return X[(int64_t)i * ld + k]; },
      [=](int k, int j) { return Y[(int64_t)k * ld + j]; }, k_lower, upper_tile,
      [=](int i, int j, double val) { A[(int64_t)i * ld + j] = val; }, tmp);
  clk.lap(TA_PROF_TRANSFORM);
  if (two_stage) {
    // dense -> band (panel QR, one pass over the trailing triangle per panel)
    // -> tridiagonal (bulge chasing); reflectors kept in the scratch
    // LDS: the products' tile from er + 8 ld (the row-sum vector of the
    // one-stage path and the tile region after it: ld + 4096 doubles), the
    // panel scalars at er + 2 ld; d and e are written last
    // panel scalars (560 doubles) at er + 2 ld clear the tile at er + 8 ld
    // from ld = 96 on; below, they follow the tile (the LDS layout reserves
    // the room, scf_lds_layout)
    lds_dbl* band_sm = ld >= 96 ? er + 2 * ld : er + 8 * ld + band::TILE_DOUBLES;
    big_band_tridiagonalize<NT>(A, scr, n, ld, d, e, er + 8 * ld, band_sm, red, eig_region, clk);
  } else {
    // per-wave column partials of the tridiagonalisation share the GEMM tile region
    gbl_dbl* wbt = scr + 3 * (int64_t)ld * TA_BIG_NT;
    gbl_dbl* pvt = wbt + (int64_t)TA_BIG_NB * ld;
    if (threadIdx.x == 0)
      for (int q = 0; q < 5; ++q) red[16 + q] = 0.0;
    big::tridiagonalize_blocked<NT, TA_BIG_NP>(A, n, ld, d, e, tau, vv, vn, yy, yrow, ww, tmp, pvt,
                                               pvt + (int64_t)TA_BIG_NP * ld, npart, red,
                                               scal, clk.on ? red + 16 : nullptr);
    if (clk.on) {
      for (int q = 0; q < 5; ++q) clk.acc[TA_PROF_TRI_COLUMN + q] += (long long)red[16 + q];
      clk.last = wall_clock64();
    }
  }
  clk.lap(TA_PROF_TRIDIAG);
  // lanes per level: as few as still fill the workgroup (see bisect_all)
  if (n > NT / 2)
    big::bisect_all<NT, 1>(d, e, e2, n, ev, TA_LDS(prev_), TA_LDS(warm_), red, scal);
  else if (n > NT / 4)
    big::bisect_all<NT, 2>(d, e, e2, n, ev, TA_LDS(prev_), TA_LDS(warm_), red, scal);
  else
    big::bisect_all<NT, 4>(d, e, e2, n, ev, TA_LDS(prev_), TA_LDS(warm_), red, scal);
  __syncthreads();
  clk.lap(TA_PROF_BISECT);
}

// Eigenvectors of the first nvec levels (those that carry occupation) from
// the tridiagonal form left by big_eigenvalues: inverse iteration, clusters,
// back-transform, then the MO rows CT[0:nvec] = Z^T X.  runs_: ld ints of
// LDS scratch.
template <int NT>
__device__ __noinline__ void big_eigenvectors(const double* A_, const double* X_, double* Y_,
                                              double* CT_, double* scr_, int n, int ld, int nvec,
                                              double* er_, const double* ev_, double* scal_,
                                              int* runs_, bool two_stage, int eig_region,
                                              PhaseClock& clk)
{
  const gbl_dbl* A = TA_GBL(A_);
  const gbl_dbl* X = TA_GBL(X_);
  gbl_dbl* Y = TA_GBL(Y_);
  gbl_dbl* CT = TA_GBL(CT_);
  gbl_dbl* scr = TA_GBL(scr_);
  lds_dbl* er = TA_LDS(er_);
  const lds_dbl* ev = TA_LDS(ev_);
  lds_dbl* scal = TA_LDS(scal_);
  lds_int* runs = TA_LDSI(runs_);
  const lds_dbl* d = er;
  const lds_dbl* e = er + ld;
  const lds_dbl* tau = er + 2 * ld;
  lds_dbl* tmp = er + 9 * ld;
  const double tnorm = scal[0];
  __syncthreads();
  gbl_dbl* wbt = scr + 3 * (int64_t)ld * TA_BIG_NT;
  big::inverse_iteration<NT>(d, e, ev, n, nvec, tnorm, Y, ld, scr);
  clk.lap(TA_PROF_INVIT);
  big::orthogonalize_clusters<NT>(ev, n, nvec, tnorm, Y, ld, runs, scal);
  clk.lap(TA_PROF_CLUSTERS);
  if (two_stage) {
    // Z = Q1 Q2 Y: the bulge-chasing reflectors in blocks (operands formed
    // first, the eigensolver region of LDS is free here), then the panel
    // reflectors
    band::q2_operands<NT>(scr + band_v2_offset(ld), n, scr + band_q2op_offset(ld), er);
    __syncthreads();
    band::apply_q2<NT>(scr + band_q2op_offset(ld), n, ld, nvec, Y);
    __syncthreads();
    clk.lap(TA_PROF_Q2);
    band::back_transform_q1<NT>(scr + band_v1_offset(ld), scr + band_t1_offset(ld), n, ld, nvec,
                                Y, er, eig_region);
  } else {
    big::back_transform<NT, TA_BIG_NB>(A, tau, n, ld, nvec, Y, wbt, tmp + 2048,
                                       tmp + 2048 + TA_BIG_NB * TA_BIG_NB, tmp);
  }
  clk.lap(TA_PROF_BACKTR);
  // CT = Z^T X; X[k][a] = 0 for a > k
  big::gemm<NT, false, true, 32>(
      nvec, n, [=](int i, int k) { return Y[(int64_t)k * ld + i]; },
      [=](int k, int j) { return X[(int64_t)k * ld + j]; },
      [n](int, int tj, int* k0, int* k1) {
        *k0 = tj * 64;
        *k1 = n;
      },
      [](int, int) { return false; },
      [=](int i, int j, double val) { CT[(int64_t)i * ld + j] = val; }, tmp);
  clk.lap(TA_PROF_JACOBI);
}

// P = sum_t f_t CT[occ_t]^T CT[occ_t] (get_density_matrix,
// src/wavefunction/type.cu:17-35 text), lower tiles mirrored; P plain n x n.
template <int NT>
__device__ __noinline__ void big_density(const double* CT_, const double* f_, const int* occ_,
                                         int nocc, double* P_, int n, int ld, double* er_)
{
  const gbl_dbl* CT = TA_GBL(CT_);
  gbl_dbl* P = TA_GBL(P_);
  const lds_dbl* f = TA_LDS(f_);
  const lds_int* occ = TA_LDSI(occ_);
  big::gemm_reg<NT>(
      n, n,
      [=](int i, int t) {
        const int k = occ[t];
        return CT[(int64_t)k * ld + i] * f[k];
      },
      [=](int t, int j) { return CT[(int64_t)occ[t] * ld + j]; },
      [nocc](int, int, int* k0, int* k1) {
        *k0 = 0;
        *k1 = nocc;
      },
      [](int bi, int bj) { return bj > bi; },
      [=](int i, int j, double val) {
        if (i >= j) {
          P[(int64_t)i * ld + j] = val;
          P[(int64_t)j * ld + i] = val;
        }
      });
}

// ------------------------------------------------------------- D4 phases
// Kept out of line so that their reference-weight accumulators do not add to
// the register pressure of the SCF kernel's other phases.
//
// Potential dE2/dq_a added to vat[a]: four atoms per wave (16-lane segments),
// lanes over partners i; the lane with sl == 0 of atom a's segment writes.
template <int NW>
__device__ __noinline__ void d4_potential(const DevD4& d4, const int* __restrict__ el, int nat,
                                          const double* __restrict__ disp,
                                          const double* __restrict__ gw,
                                          const double* __restrict__ dgw, lds_dbl* vat)
{
  const int lane = threadIdx.x & 63, wid = wave_id();
  const int seg = lane >> 4, sl = lane & 15;
  for (int a0 = 4 * wid; a0 < nat; a0 += 4 * NW) {
    const int a = a0 + seg;
    double vv[TA_D4_MAXREF] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int ea = a < nat ? el[a] : 0;
    for (int i = sl; i < nat && a < nat; i += 16) {
      const double dai = disp[(int64_t)a * nat + i];
      if (dai != 0.0) d4_scf_accumulate(d4, ea, el[i], dai, gw + TA_D4_MAXREF * i, vv);
    }
    double va = 0.0;
#pragma unroll
    for (int r = 0; r < TA_D4_MAXREF; ++r) {
      vv[r] = seg16_sum(vv[r]);
      if (a < nat && r < d4_nw(d4, ea)) va += vv[r] * dgw[TA_D4_MAXREF * a + r];
    }
    if (sl == 0 && a < nat) vat[a] += va;
  }
}

// This thread's share of E2 = 1/2 sum_{a != i} sum_rs g_ar d_ai C6ref g_is,
// threads over all atom pairs.
template <int NT>
__device__ __noinline__ double d4_energy_part(const DevD4& d4, const int* __restrict__ el, int nat,
                                              const double* __restrict__ disp,
                                              const double* __restrict__ gw)
{
  double ed = 0.0;
  for (int ai = threadIdx.x; ai < nat * nat; ai += NT) {
    const double dai = disp[ai];
    if (dai == 0.0) continue;
    const int a = ai / nat, i = ai - a * nat;
    const int ea = el[a];
    double vv[TA_D4_MAXREF] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    d4_scf_accumulate(d4, ea, el[i], 0.5 * dai, gw + TA_D4_MAXREF * i, vv);
#pragma unroll
    for (int r = 0; r < TA_D4_MAXREF; ++r)
      if (r < d4_nw(d4, ea)) ed += vv[r] * gw[TA_D4_MAXREF * a + r];
  }
  return ed;
}

// ---------------------------------------------------------------- kernel
// The global-memory variant runs two workgroups per CU (128 VGPRs): its phases
// are bound by memory latency, not by arithmetic.  CELL: periodic cells (LDS
// instantiations only), whose AES2 interactions are read from the
// Ewald-summed planes of k_pbc_pairs, diagonal included (an atom interacts
// with its own images), instead of the molecular pair kernels evaluated in
// place.
template <int NT, int MAXT, int MAXB, int MAXC, bool MATS_LDS, bool CELL>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(MATS_LDS ? (NT >= 512 ? 1 : 2) : 4))) void k_scf(BatchDev b, const int* __restrict__ mols, int ldmax,
                                            int natmax, int nshmax)
{
  extern __shared__ double smem[];
  const int tid = threadIdx.x;
  // the wave index as a scalar: loops over a wave's rows, shells or atoms
  // then run under scalar control instead of per-lane exit masks
  const int lane = tid & 63, wid = wave_id();
  constexpr int NW = NT / 64;
  const ScfLds L = scf_lds_layout(ldmax, natmax, nshmax, MATS_LDS);
  int* ismem = reinterpret_cast<int*>(smem);

  MolView v;
  v.m = mols[blockIdx.x];
  const int m = v.m;
  v.n = b.mol_nao[m];
  v.ld = (v.n + 3) & ~3;
  v.nat = b.mol_nat[m];
  v.nsh = b.mol_nsh[m];
  v.a0 = b.mol_atom_off[m];
  v.s0 = b.mol_sh_off[m];
  v.o0 = b.mol_ao_off[m];
  v.ndim = v.nsh + 9 * v.nat;
  v.mem = b.mix_memory;
  const int n = v.n, ld = v.ld, nat = v.nat, nsh = v.nsh;
  v.el = b.at_el + v.a0;
  v.sh_at = b.sh_at + v.s0;
  v.sh_ish = b.sh_ish + v.s0;
  v.sh_iao = b.sh_iao + v.s0;
  v.ao_at = b.ao_at + v.o0;
  v.ao_sh = b.ao_sh + v.o0;
  v.at_ish = b.at_ish + v.a0;
  v.at_iao = b.at_iao + v.a0;
  const int64_t moff = b.mol_mat_off[m];
  const int64_t n2 = (int64_t)n * n;
  v.S = b.S + moff;
  v.H = b.H + moff;
  v.D = b.D + 3 * moff;
  v.Q = b.Q + 6 * moff;
  v.gam = b.gam + b.mol_gam_off[m];
  v.xyz = b.xyz + 3 * (int64_t)v.a0;
  v.mrad = b.mrad + v.a0;
  if constexpr (CELL) {
    v.aes_sd = b.aes_sd + 3 * b.mol_pair_off[m];
    v.aes_dd = b.aes_dd + 9 * b.mol_pair_off[m];
    v.aes_sq = b.aes_sq + 6 * b.mol_pair_off[m];
  }
  v.qsh = b.qsh + v.s0;
  v.qat = b.qat + v.a0;
  v.dpat = b.dpat + 3 * (int64_t)v.a0;
  v.qpat = b.qpat + 6 * (int64_t)v.a0;
  const DevParams& P = *b.par;

  // LDS path: A (skewed) and C = CT in LDS.  Global path: A, X = L^-1 (in C),
  // Y (eigenvectors, then the density) and CT, plain row-major.
  double* A;
  double* C;
  double* Y = nullptr;
  double* CT;
  double* scr = nullptr;
  if constexpr (MATS_LDS) {
    A = smem + L.a;
    C = smem + L.c;
    CT = C;
  } else {
    A = b.big_a + b.mol_big_off[m];
    C = b.big_c + b.mol_big_off[m];
    Y = b.big_t + b.mol_big_off[m];
    CT = b.big_ct + b.mol_big_off[m];
    scr = b.big_s + b.mol_scr_off[m];
  }
  // two-stage (band) reduction in the global-memory eigensolver
  const bool two_stage = !MATS_LDS && b.big_eig != 0 && ld <= TA_BAND_MAXLD;
  // element (i, j) of the working matrix A / the density
  auto aix = [ld](int i, int j) { return MATS_LDS ? sk(i, j, ld) : i * ld + j; };
  double* Pm = MATS_LDS ? A : Y;
  double* red = smem + L.red;
  double* scal = smem + L.scal;
  double* ev = smem + L.ev;
  double* focc_u = smem + L.focc;
  double* emo_s = smem + L.emo;
  double* fs = smem + L.fs;
  JPlan<MAXB, MAXC> jplan;
  jacobi_plan<NT, MAXB, MAXC>(jplan, ld);
  // Jacobi rotation records live in the work region (free between the Fock
  // build and the Mulliken analysis): 7 ld doubles
  JRot jr;
  jr.cs = reinterpret_cast<double2*>(smem + L.work);
  jr.ta = jr.cs + ld;
  jr.dg = jr.ta + ld;
  jr.pq = reinterpret_cast<int2*>(jr.dg + ld);
  double* work = smem + L.work;
  int* rank = ismem + L.rank;
  int* inv = ismem + L.inv;
  int* occl = ismem + L.occ;


  double* qin = b.mix_qin + b.mol_mix_off[m];
  double* dq = b.mix_dq + b.mol_mix_off[m];
  double* dqlast = b.mix_dqlast + b.mol_mix_off[m];
  double* qlast = b.mix_qlast + b.mol_mix_off[m];
  double* dfm = b.mix_df + b.mol_mixmem_off[m];
  double* um = b.mix_u + b.mol_mixmem_off[m];
  const int mem = v.mem;
  double* amix = b.mix_a + (int64_t)m * mem * mem;
  double* beta = b.mix_beta + (int64_t)m * mem * (mem + 1);
  double* omega = b.mix_omega + (int64_t)m * mem;
  double* cvec = b.mix_c + (int64_t)m * mem;
  const int ndim = v.ndim;
  const double damp = b.damping;

  int status = 0;
  PhaseClock clk;
  clk.start(b.prof != nullptr);

  // --- occupation numbers, src/xtb.cu:290-313 and src/h0.cu:61-91 (text)
  double nref = 0.0;
  for (int s = tid; s < nsh; s += NT) nref += b.n0sh[v.s0 + s];
  nref = block_sum(nref, red);
  const double charge = b.mol_charge[m];
  const double nocc = nref - charge;
  const double nel = nref - charge;
  const int nint_nel = (int)lround(nel);
  const int uhf = b.mol_uhf[m];
  const double nuhf = ((uhf % 2) == (nint_nel % 2)) ? (double)uhf : (double)(nint_nel % 2);
  const double diff = fmin(nuhf, nocc);
  const double nalp = (nocc - diff) / 2.0 + diff, nbet = (nocc - diff) / 2.0;

  const DevD4& d4 = P.d4;
  double* d4gw = b.d4_gw + TA_D4_MAXREF * (int64_t)v.a0;
  double* d4dgw = b.d4_dgw + TA_D4_MAXREF * (int64_t)v.a0;
  const double* d4cn = b.d4_cn + v.a0;
  const double* d4disp = b.d4_disp + b.mol_pair_off[m];
  double ed2 = 0.0;

  // --- initial guess: zero charges and multipoles (superposition of neutral atoms)
  for (int i = tid; i < nsh; i += NT) v.qsh[i] = 0.0;
  for (int i = tid; i < nat; i += NT) v.qat[i] = 0.0;
  for (int i = tid; i < 3 * nat; i += NT) v.dpat[i] = 0.0;
  for (int i = tid; i < 6 * nat; i += NT) v.qpat[i] = 0.0;
  for (int i = tid; i < ndim; i += NT) qin[i] = 0.0;

  // --- orthonormal start basis C = L^-T
  {
    // six loads per thread in flight; the trip count is the same for every
    // thread (a loop whose threads leave it at different times had the
    // values live across it spilled under the exit mask, i.e. not at all)
    constexpr int SB = 6;
    for (int base0 = 0; base0 < ld * ld; base0 += SB * NT) {
      const int base = base0 + tid;
      double sv[SB];
#pragma unroll
      for (int u = 0; u < SB; ++u) {
        const int idx = base + u * NT;
        const int i = idx / ld, j = idx - (idx / ld) * ld;
        sv[u] = (idx < ld * ld && i < n && j < n) ? v.S[(int64_t)i * n + j] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < SB; ++u) {
        const int idx = base + u * NT;
        if (idx < ld * ld) A[aix(idx / ld, idx - (idx / ld) * ld)] = sv[u];
      }
    }
  }
  __syncthreads();
  bool chol_ok;
  if constexpr (MATS_LDS)
    chol_ok = orthonormalizer<NT, MATS_LDS>(A, C, n, ld, work);
  else
    chol_ok = orthonormalizer_blocked<NT>(TA_GBL(A), TA_GBL(C), n, ld, TA_LDS(work));
  if (!chol_ok) status |= TA_STATUS_CHOLESKY;
  clk.lap(TA_PROF_INIT);

  if (tid == 0) red[24] = 0.0;  // bisection warm-start bracket (global-memory path)
  __syncthreads();
  const double econv = 1.0e-6 * b.accuracy, pconv = 2.0e-5 * b.accuracy;
  const double kt = b.kt;
  double esum = 0.0, elast = 0.0;
  int iter = 0, mixiter = 0;
  bool converged = false;
  int last_nocc = 0;
  double perr_last = 1.0;  // charge residual of the previous iteration
  // levels, occupations and their order (emo_s, focc_u, rank, inv) hold the
  // values of a completed eigensolve (false if the first one failed)
  bool have_levels = false;

  while (!status && !converged && iter < b.max_iter) {
    elast = esum;
    // ------------------------------------------------ Broyden step
    if (iter > 0) {
      ++mixiter;
      if (mixiter == 1) {
        for (int i = tid; i < ndim; i += NT) {
          dqlast[i] = dq[i];
          qlast[i] = qin[i];
          qin[i] += damp * dq[i];
        }
        __syncthreads();
      } else {
        const int it1 = (mixiter - 2) % mem;
        const int ldim = (mixiter - 1) < mem ? (mixiter - 1) : mem;
        double* dfc = dfm + (int64_t)it1 * ndim;
        double* uc = um + (int64_t)it1 * ndim;
        double a1 = 0.0, a2 = 0.0;
        for (int i = tid; i < ndim; i += NT) {
          const double t = dq[i] - dqlast[i];
          dfc[i] = t;
          a1 += dq[i] * dq[i];
          a2 += t * t;
        }
        a1 = block_sum(a1, red);
        a2 = block_sum(a2, red);
        double om = sqrt(a1);
        om = (om > 0.01 / 100000.0) ? 0.01 / om : 100000.0;
        if (om < 1.0) om = 1.0;
        const double inv_ = 1.0 / fmax(sqrt(a2), 2.220446049250313e-16);
        if (tid == 0) omega[it1] = om;
        for (int i = tid; i < ndim; i += NT) {
          const double d = dfc[i] * inv_;
          dfc[i] = d;
          uc[i] = damp * d + inv_ * (qin[i] - qlast[i]);
        }
        __syncthreads();
        // a(j, it1) = <df_j, df_it1>, c(j) = omega_j <df_j, dq>
        for (int j = wid; j < ldim; j += NW) {
          const double* dfj = dfm + (int64_t)j * ndim;
          double s1 = 0.0, s2 = 0.0;
          for (int i = lane; i < ndim; i += 64) {
            s1 += dfj[i] * dfc[i];
            s2 += dfj[i] * dq[i];
          }
          s1 = wave_sum(s1);
          s2 = wave_sum(s2);
          if (lane == 0) {
            amix[j + (int64_t)mem * it1] = s1;
            amix[it1 + (int64_t)mem * j] = s1;
            cvec[j] = (j == it1 ? om : omega[j]) * s2;
          }
        }
        __syncthreads();
        // beta = omega omega a + omega0^2 I, augmented with c: solve beta gamma = c
        const int lda = ldim + 1;
        for (int idx = tid; idx < ldim * ldim; idx += NT) {
          const int i = idx / ldim, j = idx - (idx / ldim) * ldim;
          double val = omega[i] * omega[j] * amix[i + (int64_t)mem * j];
          if (i == j) val += 0.01 * 0.01;
          beta[i * lda + j] = val;
        }
        for (int i = tid; i < ldim; i += NT) beta[i * lda + ldim] = cvec[i];
        __syncthreads();
        // Gaussian elimination with partial pivoting (rows i, columns j)
        for (int col = 0; col < ldim; ++col) {
          if (wid == 0) {
            double best = -1.0;
            int bi = col;
            for (int r = col + lane; r < ldim; r += 64) {
              const double a = fabs(beta[r * lda + col]);
              if (a > best) {
                best = a;
                bi = r;
              }
            }
            for (int off = 32; off > 0; off >>= 1) {
              const double ob = __shfl_xor(best, off, 64);
              const int oi = __shfl_xor(bi, off, 64);
              if (ob > best || (ob == best && oi < bi)) {
                best = ob;
                bi = oi;
              }
            }
            if (lane == 0) {
              scal[1] = (double)bi;
              scal[2] = best;
            }
          }
          __syncthreads();
          const int piv = (int)scal[1];
          if (!(scal[2] > 0.0)) {
            status |= TA_STATUS_MIXER;
            break;
          }
          if (piv != col)
            for (int k = tid; k <= ldim; k += NT) {
              const double t = beta[col * lda + k];
              beta[col * lda + k] = beta[piv * lda + k];
              beta[piv * lda + k] = t;
            }
          __syncthreads();
          const double pv = beta[col * lda + col];
          const int rows = ldim - col - 1, cols = ldim - col;
          for (int idx = tid; idx < rows * cols; idx += NT) {
            const int r = col + 1 + idx / cols, k = col + 1 + idx % cols;
            beta[r * lda + k] -= beta[r * lda + col] / pv * beta[col * lda + k];
          }
          __syncthreads();
          for (int r = col + 1 + tid; r < ldim; r += NT) beta[r * lda + col] = 0.0;
          __syncthreads();
        }
        if (status) break;
        // back substitution into cvec (gamma), one wave
        if (wid == 0) {
          for (int r = ldim - 1; r >= 0; --r) {
            double s = 0.0;
            for (int k = r + 1 + lane; k < ldim; k += 64) s += beta[r * lda + k] * cvec[k];
            s = wave_sum(s);
            if (lane == 0) cvec[r] = (beta[r * lda + ldim] - s) / beta[r * lda + r];
          }
        }
        __syncthreads();
        for (int i = tid; i < ndim; i += NT) {
          double q = qin[i] + damp * dq[i];
          for (int j = 0; j < ldim; ++j) q -= omega[j] * cvec[j] * um[(int64_t)j * ndim + i];
          dqlast[i] = dq[i];
          qlast[i] = qin[i];
          qin[i] = q;
        }
        __syncthreads();
      }
      // get_mixer: unpack into the wavefunction
      for (int i = tid; i < nsh; i += NT) v.qsh[i] = qin[i];
      for (int i = tid; i < 3 * nat; i += NT) v.dpat[i] = qin[nsh + i];
      for (int i = tid; i < 6 * nat; i += NT) v.qpat[i] = qin[nsh + 3 * nat + i];
      __syncthreads();
      for (int a = tid; a < nat; a += NT) {
        const int s0a = v.at_ish[a];
        const int ns = P.el[v.el[a]].nsh;
        double q = 0.0;
        for (int s = 0; s < ns; ++s) q += v.qsh[s0a + s];
        v.qat[a] = q;
      }
      __syncthreads();
    }
    ++iter;
    clk.lap(TA_PROF_MIX);

    // ------------------------------------------------ potential
    // D4 reference weights at the charges of this iteration
    if (d4.enabled) {
      for (int a = tid; a < nat; a += NT)
        d4_atom_weights(d4, v.el[a], d4cn[a], v.qat[a], d4gw + TA_D4_MAXREF * a, d4dgw + TA_D4_MAXREF * a);
      __syncthreads();
    }
    double* vao = work;
    double* vsh = work + ld;
    double* vat = vsh + nshmax;
    double* vdp = vat + natmax;
    double* vqp = vdp + 3 * natmax;
    // ES2 + ES3 shell potential: four shells per wave (16-lane segments),
    // lanes over partner shells; four rows share one memory round trip
    const int seg = lane >> 4, sl = lane & 15;
    for (int sh0 = 4 * wid; sh0 < nsh; sh0 += 4 * NW) {
      const int sh = sh0 + seg;
      double acc = 0.0;
      if (sh < nsh) {
        const double* g = v.gam + (int64_t)sh * nsh;
        for (int j = sl; j < nsh; j += 16) acc += g[j] * v.qsh[j];
      }
      acc = seg16_sum(acc);
      if (sl == 0 && sh < nsh) {
        const double q = v.qsh[sh];
        vsh[sh] = acc + q * q * P.el[v.el[v.sh_at[sh]]].hubbard_derivs[v.sh_ish[sh]];
      }
    }
    // AES2 atomic potentials: four atoms a per wave (16-lane segments), lanes
    // over partners i, each lane evaluating the pair kernels of (a, i)
    // (amat_sd is antisymmetric in the atom pair, amat_dd and amat_sq
    // symmetric)
    for (int a0 = 4 * wid; a0 < nat; a0 += 4 * NW) {
      const int a = a0 + seg;
      double va = 0.0, d[3] = {0.0, 0.0, 0.0}, q[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      if constexpr (CELL) {
        const int64_t p2 = (int64_t)nat * nat;
        for (int i = sl; i < nat && a < nat; i += 16) {
          const int64_t ai = (int64_t)a * nat + i;
          const double qi = v.qat[i];
          const double dpi[3] = {v.dpat[3 * i], v.dpat[3 * i + 1], v.dpat[3 * i + 2]};
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const double sdk = v.aes_sd[k * p2 + ai];
            va += sdk * dpi[k];
            d[k] -= sdk * qi;
#pragma unroll
            for (int l = 0; l < 3; ++l) d[k] += v.aes_dd[(3 * l + k) * p2 + ai] * dpi[l];
          }
#pragma unroll
          for (int k = 0; k < 6; ++k) {
            const double sqk = v.aes_sq[k * p2 + ai];
            va += sqk * v.qpat[6 * i + k];
            q[k] += sqk * qi;
          }
        }
      } else {
      for (int i = sl; i < nat && a < nat; i += 16) {
        const AesPair pr = aes_pair(v.xyz, v.mrad, a, i, P.dmp3, P.dmp5);
        const double qi = v.qat[i];
        const double dpi[3] = {v.dpat[3 * i], v.dpat[3 * i + 1], v.dpat[3 * i + 2]};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double sdk = aes_sd(pr, k);  // amat_sd(k, i, a) = -amat_sd(k, a, i)
          va += sdk * dpi[k];
          d[k] -= sdk * qi;
#pragma unroll
          for (int l = 0; l < 3; ++l) d[k] += aes_dd(pr, l, k) * dpi[l];
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          const double sqk = aes_sq(pr, k);
          va += sqk * v.qpat[6 * i + k];
          q[k] += sqk * qi;
        }
      }
      }
      va = seg16_sum(va);
#pragma unroll
      for (int k = 0; k < 3; ++k) d[k] = seg16_sum(d[k]);
#pragma unroll
      for (int k = 0; k < 6; ++k) q[k] = seg16_sum(q[k]);
      if (sl == 0 && a < nat) {
        const DevElem& e = P.el[v.el[a]];
        vat[a] = va;
#pragma unroll
        for (int k = 0; k < 3; ++k) vdp[3 * a + k] = d[k] + 2.0 * e.dkernel * v.dpat[3 * a + k];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          const double sc = (k == 1 || k == 3 || k == 4) ? 2.0 : 1.0;
          vqp[6 * a + k] = q[k] + 2.0 * e.qkernel * v.qpat[6 * a + k] * sc;
        }
      }
    }
    // D4 potential dE2/dq_a, same atom-to-lane mapping as above (the lane
    // that wrote vat[a] adds to it)
    if (d4.enabled) d4_potential<NW>(d4, v.el, nat, d4disp, d4gw, d4dgw, TA_LDS(vat));
    __syncthreads();
    for (int i = tid; i < n; i += NT) vao[i] = vsh[v.ao_sh[i]] + vat[v.ao_at[i]];
    // set_mixer: q_in = current charges
    if (iter == 1)
      for (int i = tid; i < ndim; i += NT) qin[i] = 0.0;
    __syncthreads();
    clk.lap(TA_PROF_POT);

    // ------------------------------------------------ Fock matrix
    // A1[i][j] = (H_ij - S_ij (v_i + v_j)/2)/2 - (D(.,j,i).vdp(at i) + Q(.,j,i).vqp(at i))/2
    // LDS path: four rows per wave (16-lane segments, lanes over j), the row's
    // potentials read once per row, the integral loads of CH column steps
    // issued together (as in the Mulliken loop), then F = A1 + A1^T.
    // Global-memory path: F directly from the lower triangle (below).
    if constexpr (MATS_LDS) {
      constexpr int JS = NT >= 512 ? 6 : (NT >= 256 ? 4 : 2), CH = NT >= 512 ? 3 : 2;
      for (int i0 = 4 * wid; i0 < ld; i0 += 4 * NW) {
        const int i = i0 + seg;
        const bool rin = i < n;
        const int ai = rin ? v.ao_at[i] : 0;
        const double vi = rin ? vao[i] : 0.0;
        const double* vd = vdp + 3 * ai;
        const double* vq = vqp + 6 * ai;
        const double vd0 = vd[0], vd1 = vd[1], vd2 = vd[2];
        const double vq0 = vq[0], vq1 = vq[1], vq2 = vq[2], vq3 = vq[3], vq4 = vq[4], vq5 = vq[5];
        const int64_t row = (int64_t)(rin ? i : 0) * n;
#pragma unroll 1
        for (int u0 = 0; u0 < JS; u0 += CH) {
          double fv[CH];
#pragma unroll
          for (int u = 0; u < CH; ++u) {
            const int j = sl + 16 * (u0 + u);
            const bool in = rin && j < n;
            const int jc = in ? j : 0;
            const int64_t e = row + jc;
            double f = 0.5 * (v.H[e] - v.S[e] * 0.5 * (vi + vao[jc]));
            f -= 0.5 * (v.D[e] * vd0 + v.D[n2 + e] * vd1 + v.D[2 * n2 + e] * vd2);
            f -= 0.5 * (v.Q[e] * vq0 + v.Q[n2 + e] * vq1 + v.Q[2 * n2 + e] * vq2 +
                        v.Q[3 * n2 + e] * vq3 + v.Q[4 * n2 + e] * vq4 + v.Q[5 * n2 + e] * vq5);
            fv[u] = in ? f : 0.0;
          }
#pragma unroll
          for (int u = 0; u < CH; ++u) {
            const int j = sl + 16 * (u0 + u);
            if (j < ld) A[aix(i, j)] = fv[u];
          }
        }
      }
    } else {
      // Global-memory path: the lower triangle only.  The integral kernel
      // stores the dipole and quadrupole integrals of an atom pair on the
      // row atom's center below the diagonal (the half neighbour list has
      // jat <= iat) and writes the element above the diagonal as the same
      // operator moved to the other center (r - R_j = (r - R_i) + vec); that
      // element is formed here from the one below in the same way, so each
      // integral is read once and F_ij = f(i,j) + f(j,i) is written to both
      // triangles (f(i,i) twice on the diagonal).  Two 64-column chunks of a
      // row per wave and step.
      const gbl_dbl* gH = TA_GBL(v.H);
      const gbl_dbl* gS = TA_GBL(v.S);
      const gbl_dbl* gD = TA_GBL(v.D);
      const gbl_dbl* gQ = TA_GBL(v.Q);
      const double* xyz = v.xyz;
      auto felem = [&](int i, int j, int ai, const double* xi, double vi, const double* vdi,
                       const double* vqi) {
        const int64_t e = (int64_t)i * n + j;
        const double sv = gS[e], hv = gH[e];
        double di[3], qi[6];
#pragma unroll
        for (int k = 0; k < 3; ++k) di[k] = gD[k * n2 + e];
#pragma unroll
        for (int k = 0; k < 6; ++k) qi[k] = gQ[k * n2 + e];
        const double base = 0.5 * (hv - sv * 0.5 * (vi + vao[j]));
        const double fij = base - 0.5 * (di[0] * vdi[0] + di[1] * vdi[1] + di[2] * vdi[2]) -
                           0.5 * (qi[0] * vqi[0] + qi[1] * vqi[1] + qi[2] * vqi[2] +
                                  qi[3] * vqi[3] + qi[4] * vqi[4] + qi[5] * vqi[5]);
        if (j == i) return 2.0 * fij;
        const int aj = v.ao_at[j];
        const double* vdj = vdp + 3 * aj;
        const double* vqj = vqp + 6 * aj;
        double dj[3], qj[6];
        if constexpr (CELL) {
          // a cell's elements are sums over lattice images, each with its own
          // shift: the element above the diagonal is read as stored
          const int64_t et = (int64_t)j * n + i;
#pragma unroll
          for (int k = 0; k < 3; ++k) dj[k] = gD[k * n2 + et];
#pragma unroll
          for (int k = 0; k < 6; ++k) qj[k] = gQ[k * n2 + et];
        } else {
        double vec[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) vec[k] = xi[k] - xyz[3 * aj + k];
        dj[0] = di[0] + vec[0] * sv;
        dj[1] = di[1] + vec[1] * sv;
        dj[2] = di[2] + vec[2] * sv;
        const double t0 = 2 * vec[0] * di[0] + vec[0] * vec[0] * sv;
        const double t2 = 2 * vec[1] * di[1] + vec[1] * vec[1] * sv;
        const double t5 = 2 * vec[2] * di[2] + vec[2] * vec[2] * sv;
        const double t1 = vec[0] * di[1] + vec[1] * di[0] + vec[0] * vec[1] * sv;
        const double t3 = vec[0] * di[2] + vec[2] * di[0] + vec[0] * vec[2] * sv;
        const double t4 = vec[1] * di[2] + vec[2] * di[1] + vec[1] * vec[2] * sv;
        const double tr = 0.5 * (t0 + t2 + t5);
        qj[0] = qi[0] + 1.5 * t0 - tr;
        qj[1] = qi[1] + 1.5 * t1;
        qj[2] = qi[2] + 1.5 * t2 - tr;
        qj[3] = qi[3] + 1.5 * t3;
        qj[4] = qi[4] + 1.5 * t4;
        qj[5] = qi[5] + 1.5 * t5 - tr;
        }
        const double fji = base - 0.5 * (dj[0] * vdj[0] + dj[1] * vdj[1] + dj[2] * vdj[2]) -
                           0.5 * (qj[0] * vqj[0] + qj[1] * vqj[1] + qj[2] * vqj[2] +
                                  qj[3] * vqj[3] + qj[4] * vqj[4] + qj[5] * vqj[5]);
        return fij + fji;
      };
      for (int i = wid; i < n; i += NW) {
        const int ai = v.ao_at[i];
        const double xi[3] = {xyz[3 * ai], xyz[3 * ai + 1], xyz[3 * ai + 2]};
        const double vi = vao[i];
        const double vdi[3] = {vdp[3 * ai], vdp[3 * ai + 1], vdp[3 * ai + 2]};
        const double vqi[6] = {vqp[6 * ai],     vqp[6 * ai + 1], vqp[6 * ai + 2],
                               vqp[6 * ai + 3], vqp[6 * ai + 4], vqp[6 * ai + 5]};
        for (int j0 = 0; j0 <= i; j0 += 128) {
          const int ja = j0 + lane, jb = ja + 64;
          const bool ina = ja <= i, inb = jb <= i;
          const double fa = felem(i, ina ? ja : i, ai, xi, vi, vdi, vqi);
          const double fb = felem(i, inb ? jb : i, ai, xi, vi, vdi, vqi);
          if (ina) {
            A[(int64_t)i * ld + ja] = fa;
            A[(int64_t)ja * ld + i] = fa;
          }
          if (inb) {
            A[(int64_t)i * ld + jb] = fb;
            A[(int64_t)jb * ld + i] = fb;
          }
        }
      }
      __syncthreads();
    }
    if constexpr (MATS_LDS) {
      __syncthreads();
      for (int idx = tid; idx < ld * ld; idx += NT) {
        const int i = idx / ld, j = idx - (idx / ld) * ld;
        if (i < j) {
          const double t = A[aix(i, j)] + A[aix(j, i)];
          A[aix(i, j)] = t;
          A[aix(j, i)] = t;
        } else if (i == j) {
          A[aix(i, i)] *= 2.0;
        }
      }
      __syncthreads();
    }

    clk.lap(TA_PROF_FOCK);
    // ------------------------------------------------ eigensolve
    if constexpr (MATS_LDS) {
      // A <- C^T F C = CT F CT^T: U = CT F stored transposed, then CT U^T
      if (mfma_fits<NT>(ld)) {
        mm_mfma<NT, true, true>(C, A, A, ld);
        mm_mfma<NT, true, false>(C, A, A, ld);
      } else {
        mm_nn<NT, MAXT, true, true>(C, A, A, ld);
        mm_nn<NT, MAXT, true, false>(C, A, A, ld);
      }
      clk.lap(TA_PROF_TRANSFORM);
      if (b.lds_eig) {
        // Householder in LDS: reflectors applied to the basis rows as they are
        // formed, levels by bisection, eigenvectors of T into A, then the new
        // MO rows CT <- Z^T CT (ascending levels)
        const ldseig::Work W = ldseig::work_layout(work, ld);
        ldseig::tridiagonalize<NT>(TA_LDS(A), TA_LDS(C), n, ld, W, TA_LDS(red), TA_LDS(scal));
        clk.lap(TA_PROF_TRIDIAG);
        big::bisect_all<NT>(W.d, W.e, W.e2, n, TA_LDS(ev), TA_LDS(emo_s),
                            TA_LDS(red + 24), TA_LDS(red), TA_LDS(scal));
        clk.lap(TA_PROF_BISECT);
        ldseig::eigenvectors<NT>(W.d, W.e, TA_LDS(ev), n, ld, TA_LDS(A), scal[0], W.vb);
        clk.lap(TA_PROF_INVIT);
        ldseig::rotate_basis<NT, MFMA_TILES(NT)>(TA_LDS(A), TA_LDS(C), ld);
        clk.lap(TA_PROF_BACKTR);
        for (int k = tid; k < n; k += NT) {
          rank[k] = k;
          inv[k] = k;
          emo_s[k] = ev[k];
        }
        __syncthreads();
      } else {
      int sweeps = 0;
      // looser stopping rule while the charges are still far from
      // self-consistency (jacobi_adapt > 0): the next iterations start from
      // these vectors and refine them.  Rotations below jtol^2 (at most
      // jacobi_skip) are skipped: off-diagonals of that size remain anyway.
      const double jtol =
          b.jacobi_adapt > 0.0 ? fmax(b.jacobi_tol, fmin(1.0e-2, b.jacobi_adapt * perr_last))
                               : b.jacobi_tol;
      const double jskip = fmin(b.jacobi_skip, jtol * jtol);
      const bool jok = jacobi<NT, MAXB, MAXC>(A, C, ld, jr, jplan, red, jtol, jskip, &sweeps);
      clk.lap(TA_PROF_JACOBI);
      clk.acc[TA_PROF_SWEEPS] += sweeps;
      if (!jok) {
        status |= TA_STATUS_EIGEN;
        break;
      }
      for (int k = tid; k < n; k += NT) ev[k] = A[sk(k, k, ld)];
      __syncthreads();
      for (int k = tid; k < n; k += NT) {
        const double e = ev[k];
        int r = 0;
        for (int j = 0; j < n; ++j) {
          const double o = ev[j];
          r += (o < e) || (o == e && j < k);
        }
        rank[k] = r;
        inv[r] = k;
        emo_s[r] = e;
      }
      __syncthreads();
      }
    } else {
      // levels of the previous step (emo_s) warm-start the bisection; the
      // eigenvectors follow once the occupations are known
      big_eigenvalues<NT>(A, C, Y, scr, n, ld, big_npart(ldmax), smem + L.eig, ev, red, scal,
                          emo_s, red + 24, two_stage, L.iint / 2 - L.eig, clk);
      if (b.coeff != nullptr) {
        // keep d, e, tau and |T| for the eigenvectors of a requested wavefunction
        double* sv = scr + big_tri_save_offset(ld);
        const double* er = smem + L.eig;
        for (int k = tid; k < 3 * ld; k += NT) sv[k] = er[k];
        if (tid == 0) sv[3 * ld] = scal[0];
      }
      for (int k = tid; k < n; k += NT) {
        rank[k] = k;
        inv[k] = k;
        emo_s[k] = ev[k];
      }
      __syncthreads();
    }

    // ------------------------------------------------ occupations
    double ts = 0.0;
    if (wid == 0) {
      double* fsp = fs;
      ts = fermi_wave(nalp, kt, n, emo_s, fsp);
      for (int r = lane; r < n; r += 64) focc_u[inv[r]] = fsp[r];
      ts += fermi_wave(nbet, kt, n, emo_s, fsp);
      for (int r = lane; r < n; r += 64) focc_u[inv[r]] += fsp[r];
      // levels with an occupation above TA_OCC_MIN enter the density, listed
      // in level order (ballot compaction; each lane reads the occupations
      // it wrote); the occupations fall with the level, so on the
      // global-memory path (levels in ascending order) these are a leading
      // range
      int c = 0;
      for (int r0 = 0; r0 < n; r0 += 64) {
        const int r = r0 + lane;
        const int k = r < n ? inv[r] : 0;
        const bool take = r < n && focc_u[k] > TA_OCC_MIN;
        const unsigned long long m = __ballot(take);
        if (take) occl[c + __popcll(m & ((1ull << lane) - 1ull))] = k;
        c += __popcll(m);
      }
      if (lane == 0) {
        scal[3] = ts;
        scal[4] = (double)c;
      }
    }
    __syncthreads();
    ts = scal[3];
    last_nocc = (int)scal[4];
    have_levels = true;
    clk.lap(TA_PROF_OCC);
    if constexpr (!MATS_LDS) {
      int nvec = 0;
      if (last_nocc > 0) nvec = occl[last_nocc - 1] + 1;
      big_eigenvectors<NT>(A, C, Y, CT, scr, n, ld, nvec, smem + L.eig, ev, scal, rank, two_stage,
                           L.iint / 2 - L.eig, clk);
    }

    // ------------------------------------------------ density -> A
    if constexpr (MATS_LDS) {
      if (mfma_fits<NT>(ld)) density_mfma<NT>(C, focc_u, occl, last_nocc, A, ld);
      else density<NT>(C, focc_u, occl, last_nocc, A, ld);
    }
    else big_density<NT>(CT, focc_u, occl, last_nocc, Y, n, ld, smem + L.eig);
    clk.lap(TA_PROF_DENSITY);

    // ------------------------------------------------ Mulliken + band energy
    double* part = work; // part[c * ld + i], c = 0 (PS), 1..3 (PD), 4..9 (PQ), 10 (PH)
    // four rows per wave (16-lane segments, lanes over j): four rows share
    // each memory round trip.  With the matrices in LDS (n <= 16 JS) every
    // lane issues the loads of CH column steps at once: indices past n are
    // clamped to column n - 1 and weighted by zero, so no load waits on a
    // branch and the sums keep the order of the general loop.
    for (int i0 = 4 * wid; i0 < n; i0 += 4 * NW) {
      const int i = i0 + seg;
      double acc[11];
#pragma unroll
      for (int c = 0; c < 11; ++c) acc[c] = 0.0;
      const int64_t row = (int64_t)i * n;
      auto mul_step = [&](int j, bool in) {
        const int jc = in ? j : n - 1;
        const double p = in ? Pm[aix(i, jc)] : 0.0;
        const int64_t e = row + jc;
        acc[0] += p * v.S[e];
        acc[1] += p * v.D[e];
        acc[2] += p * v.D[n2 + e];
        acc[3] += p * v.D[2 * n2 + e];
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[4 + c] += p * v.Q[c * n2 + e];
        acc[10] += p * v.H[e];
      };
      if constexpr (MATS_LDS) {
        constexpr int JS = NT >= 512 ? 6 : (NT >= 256 ? 4 : 2), CH = NT >= 512 ? 3 : 2;
        if (i < n)
#pragma unroll 1
          for (int u0 = 0; u0 < JS; u0 += CH)
#pragma unroll
            for (int u = u0; u < u0 + CH; ++u) mul_step(sl + 16 * u, sl + 16 * u < n);
      } else {
#pragma unroll 2
        for (int j = sl; j < n && i < n; j += 16) mul_step(j, true);
      }
#pragma unroll
      for (int c = 0; c < 11; ++c) acc[c] = seg16_sum(acc[c]);
      if (sl == 0 && i < n)
#pragma unroll
        for (int c = 0; c < 11; ++c) part[c * ld + i] = acc[c];
    }
    __syncthreads();
    double err = 0.0;
    for (int s = tid; s < nsh; s += NT) {
      const int i0 = v.sh_iao[s];
      const int l = P.el[v.el[v.sh_at[s]]].ang[v.sh_ish[s]];
      double q = b.n0sh[v.s0 + s];
      for (int i = i0; i < i0 + 2 * l + 1; ++i) q -= part[i];
      v.qsh[s] = q;
      const double d = q - qin[s];
      dq[s] = d;
      err += d * d;
    }
    for (int a = tid; a < nat; a += NT) {
      const DevElem& e = P.el[v.el[a]];
      const int i0 = v.at_iao[a];
      int nao_a = 0;
      for (int s = 0; s < e.nsh; ++s) nao_a += 2 * e.ang[s] + 1;
      double dp[3] = {0.0, 0.0, 0.0}, qp[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      for (int i = i0; i < i0 + nao_a; ++i) {
        for (int k = 0; k < 3; ++k) dp[k] -= part[(1 + k) * ld + i];
        for (int k = 0; k < 6; ++k) qp[k] -= part[(4 + k) * ld + i];
      }
      for (int k = 0; k < 3; ++k) {
        v.dpat[3 * a + k] = dp[k];
        const double d = dp[k] - qin[nsh + 3 * a + k];
        dq[nsh + 3 * a + k] = d;
        err += d * d;
      }
      for (int k = 0; k < 6; ++k) {
        v.qpat[6 * a + k] = qp[k];
        const double d = qp[k] - qin[nsh + 3 * nat + 6 * a + k];
        dq[nsh + 3 * nat + 6 * a + k] = d;
        err += d * d;
      }
    }
    double eband = 0.0;
    for (int i = tid; i < n; i += NT) eband += part[10 * ld + i];
    __syncthreads();
    for (int a = tid; a < nat; a += NT) {
      const int s0a = v.at_ish[a];
      const int ns = P.el[v.el[a]].nsh;
      double q = 0.0;
      for (int s = 0; s < ns; ++s) q += v.qsh[s0a + s];
      v.qat[a] = q;
      // D4 weights at the output charges (dispersion energy of this iteration)
      if (d4.enabled)
        d4_atom_weights(d4, v.el[a], d4cn[a], q, d4gw + TA_D4_MAXREF * a, d4dgw + TA_D4_MAXREF * a);
    }
    __syncthreads();
    clk.lap(TA_PROF_MULLIKEN);

    // ------------------------------------------------ electrostatic energy
    // E_ES2 = q^T gamma q / 2, E_ES3 = sum q^3 Gamma / 3, E_AES2 over atom
    // pairs (a, i): dp_a . (amat_sd(a,i) q_i + amat_dd(a,i) dp_i / 2)
    // + qp_a . amat_sq(a,i) q_i, plus the on-site kernels; all threads over
    // all shell and atom pairs (no per-row reduction needed), one reduction
    double ec = 0.0;
    for (int idx = tid; idx < nsh * nsh; idx += NT) {
      const int sh = idx / nsh, j = idx - sh * nsh;
      ec += 0.5 * v.qsh[sh] * v.gam[idx] * v.qsh[j];
    }
    for (int sh = tid; sh < nsh; sh += NT) {
      const double qs = v.qsh[sh];
      ec += qs * qs * qs * P.el[v.el[v.sh_at[sh]]].hubbard_derivs[v.sh_ish[sh]] / 3.0;
    }
    if constexpr (CELL) {
      const int64_t p2 = (int64_t)nat * nat;
      for (int ai = tid; ai < nat * nat; ai += NT) {
        const int a = ai / nat, i = ai - a * nat;
        const double qi = v.qat[i];
        const double dpi[3] = {v.dpat[3 * i], v.dpat[3 * i + 1], v.dpat[3 * i + 2]};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          double t = -v.aes_sd[k * p2 + ai] * qi;
#pragma unroll
          for (int l = 0; l < 3; ++l) t += 0.5 * v.aes_dd[(3 * l + k) * p2 + ai] * dpi[l];
          ec += v.dpat[3 * a + k] * t;
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) ec += v.qpat[6 * a + k] * v.aes_sq[k * p2 + ai] * qi;
      }
    }
    for (int ai = tid; ai < nat * nat && !CELL; ai += NT) {
      const int a = ai / nat, i = ai - a * nat;
      if (a == i) continue;  // the pair kernels vanish on the diagonal
      const AesPair pr = aes_pair(v.xyz, v.mrad, a, i, P.dmp3, P.dmp5);
      const double qi = v.qat[i];
      const double dpi[3] = {v.dpat[3 * i], v.dpat[3 * i + 1], v.dpat[3 * i + 2]};
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double t = -aes_sd(pr, k) * qi;
#pragma unroll
        for (int l = 0; l < 3; ++l) t += 0.5 * aes_dd(pr, l, k) * dpi[l];
        ec += v.dpat[3 * a + k] * t;
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) ec += v.qpat[6 * a + k] * aes_sq(pr, k) * qi;
    }
    for (int a = tid; a < nat; a += NT) {
      const DevElem& e = P.el[v.el[a]];
      const double* dpa = v.dpat + 3 * a;
      const double* qpa = v.qpat + 6 * a;
      ec += e.dkernel * (dpa[0] * dpa[0] + dpa[1] * dpa[1] + dpa[2] * dpa[2]);
      ec += e.qkernel * (qpa[0] * qpa[0] + 2.0 * qpa[1] * qpa[1] + qpa[2] * qpa[2] +
                         2.0 * qpa[3] * qpa[3] + 2.0 * qpa[4] * qpa[4] + qpa[5] * qpa[5]);
    }
    // D4 two-body energy E2 = 1/2 sum_{a != i} sum_rs g_ar d_ai C6ref g_is
    const double ed = d4.enabled ? d4_energy_part<NT>(d4, v.el, nat, d4disp, d4gw) : 0.0;
    ec = block_sum(ec, red);
    eband = block_sum(eband, red);
    err = block_sum(err, red);
    ed2 = d4.enabled ? block_sum(ed, red) : 0.0;
    esum = eband + ts + ec + ed2;
    const double perr = sqrt(err / ndim);
    perr_last = perr;
    converged = (fabs(esum - elast) < econv) && (perr < pconv);
    if (!isfinite(esum)) status |= TA_STATUS_NAN;
    clk.lap(TA_PROF_ENERGY);
  }
  if (!converged && status == 0) status |= TA_STATUS_NOT_CONVERGED;

  // ------------------------------------------------ results
  // repulsion total; three-body dispersion total (zero charges, k_d4_atm),
  // summed here where they are used, so that nothing holds them across the
  // SCF
  double erep = 0.0, e3 = 0.0;
  for (int i = tid; i < nat; i += NT) {
    erep += b.erep[v.a0 + i];
    if (d4.enabled) e3 += b.d4_e3[v.a0 + i];
  }
  erep = block_sum(erep, red);
  e3 = block_sum(e3, red);
  for (int r = tid; r < n; r += NT) {
    b.emo[v.o0 + r] = have_levels ? emo_s[r] : 0.0;
    b.focc[v.o0 + r] = have_levels ? focc_u[inv[r]] : 0.0;
  }
  // wavefunction arrays in tblite's column-major layout: density(i, j) of the
  // last iteration (still in A / Y: the energy phase does not touch it) and
  // coeff(i, r) = coefficient of AO i in level r (levels ascending)
  const bool wfn_ok = have_levels && !(status & (TA_STATUS_EIGEN | TA_STATUS_NAN));
  if (b.density != nullptr) {
    double* dm = b.density + moff;
    for (int idx = tid; idx < n * n; idx += NT) {
      const int j = idx / n, i = idx - j * n;
      dm[idx] = wfn_ok ? Pm[aix(i, j)] : 0.0;
    }
  }
  if (b.coeff != nullptr) {
    double* co = b.coeff + moff;
    if constexpr (!MATS_LDS) {
      // the SCF formed the occupied levels' vectors only: all n from the
      // tridiagonal form of the last iteration (its reflectors stay in A,
      // d, e, tau and |T| in LDS); overwrites Y, read out above
      __syncthreads();
      if (wfn_ok) {
        const double* sv = scr + big_tri_save_offset(ld);
        double* er = smem + L.eig;
        for (int k = tid; k < 3 * ld; k += NT) er[k] = sv[k];
        if (tid == 0) scal[0] = sv[3 * ld];
        __syncthreads();
        big_eigenvectors<NT>(A, C, Y, CT, scr, n, ld, n, smem + L.eig, ev, scal, rank, two_stage,
                             L.iint / 2 - L.eig, clk);
      }
      __syncthreads();
    }
    for (int idx = tid; idx < n * n; idx += NT) {
      const int r = idx / n, i = idx - r * n;
      co[idx] = wfn_ok ? CT[(int64_t)inv[r] * ld + i] : 0.0;
    }
  }
  if (clk.on) {
    clk.acc[TA_PROF_ITERS] = iter;
    for (int i = 0; i < TA_PROF_SLOTS; ++i) b.prof[(int64_t)m * TA_PROF_SLOTS + i] = clk.acc[i];
  }
  if (tid == 0) {
    b.energy[m] = erep + e3 + esum;
    b.edisp2[m] = ed2;
    b.edisp3[m] = e3;
    b.eelec[m] = esum;
    b.erep_tot[m] = erep;
    b.iterations[m] = iter;
    b.status[m] = status;
    if (b.done != nullptr) {
      TaDoneRecord* r = b.done + m;
      r->energy = erep + e3 + esum;
      r->iterations = iter;
      r->status = status;
      __hip_atomic_store(&r->flag, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

// explicit instantiations, launched by the host per size bucket
template __global__ void k_scf<128, 1, 2, 4, true, false>(BatchDev, const int*, int, int, int);
template __global__ void k_scf<256, 1, 3, 6, true, false>(BatchDev, const int*, int, int, int);
template __global__ void k_scf<512, 2, 3, 6, true, false>(BatchDev, const int*, int, int, int);
template __global__ void k_scf<512, 1, 0, 0, false, false>(BatchDev, const int*, int, int, int);
template __global__ void k_scf<128, 1, 2, 4, true, true>(BatchDev, const int*, int, int, int);
template __global__ void k_scf<256, 1, 3, 6, true, true>(BatchDev, const int*, int, int, int);
template __global__ void k_scf<512, 2, 3, 6, true, true>(BatchDev, const int*, int, int, int);
template __global__ void k_scf<512, 1, 0, 0, false, true>(BatchDev, const int*, int, int, int);

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
// zero); they sit in LDS for ld <= 96 and in global memory above that.
#include <hip/hip_runtime.h>
#include <math.h>

#include "common.h"
#include "device_util.h"

#define JACOBI_MAX_SWEEPS 40
#define JACOBI_TOL 1.0e-11

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
  const double* sd;
  const double* dd;
  const double* sq;
  double* qsh;
  double* qat;
  double* dpat;
  double* qpat;
};

// ------------------------------------------------------------- GEMM tiles
// out = X * Y on ld x ld row-major matrices, 4x4 register tiles.  When HOLD
// is set the output aliases an input: all tiles of a thread are kept in
// registers until every thread has finished reading.
template <int NT, int MAXT, bool HOLD>
__device__ void mm_nn(const double* X, const double* Y, double* out, int ld)
{
  const int nb = ld >> 2;
  const int ntile = nb * nb;
  if constexpr (HOLD) {
    double acc[MAXT][16];
#pragma unroll
    for (int t = 0; t < MAXT; ++t) {
      const int tile = threadIdx.x + t * NT;
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[t][e] = 0.0;
      if (tile < ntile) {
        const int r0 = (tile / nb) * 4, c0 = (tile % nb) * 4;
        for (int k = 0; k < ld; ++k) {
          double x[4], y[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) x[u] = X[(r0 + u) * ld + k];
#pragma unroll
          for (int v = 0; v < 4; ++v) y[v] = Y[k * ld + c0 + v];
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[t][u * 4 + v] += x[u] * y[v];
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < MAXT; ++t) {
      const int tile = threadIdx.x + t * NT;
      if (tile < ntile) {
        const int r0 = (tile / nb) * 4, c0 = (tile % nb) * 4;
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) out[(r0 + u) * ld + c0 + v] = acc[t][u * 4 + v];
      }
    }
  } else {
    for (int tile = threadIdx.x; tile < ntile; tile += NT) {
      const int r0 = (tile / nb) * 4, c0 = (tile % nb) * 4;
      double acc[16];
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[e] = 0.0;
      for (int k = 0; k < ld; ++k) {
        double x[4], y[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = X[(r0 + u) * ld + k];
#pragma unroll
        for (int v = 0; v < 4; ++v) y[v] = Y[k * ld + c0 + v];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[u * 4 + v] += x[u] * y[v];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) out[(r0 + u) * ld + c0 + v] = acc[u * 4 + v];
    }
  }
  __syncthreads();
}

// out = X^T * Y
template <int NT, int MAXT, bool HOLD>
__device__ void mm_tn(const double* X, const double* Y, double* out, int ld)
{
  const int nb = ld >> 2;
  const int ntile = nb * nb;
  if constexpr (HOLD) {
    double acc[MAXT][16];
#pragma unroll
    for (int t = 0; t < MAXT; ++t) {
      const int tile = threadIdx.x + t * NT;
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[t][e] = 0.0;
      if (tile < ntile) {
        const int r0 = (tile / nb) * 4, c0 = (tile % nb) * 4;
        for (int k = 0; k < ld; ++k) {
          double x[4], y[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) x[u] = X[k * ld + r0 + u];
#pragma unroll
          for (int v = 0; v < 4; ++v) y[v] = Y[k * ld + c0 + v];
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[t][u * 4 + v] += x[u] * y[v];
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < MAXT; ++t) {
      const int tile = threadIdx.x + t * NT;
      if (tile < ntile) {
        const int r0 = (tile / nb) * 4, c0 = (tile % nb) * 4;
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) out[(r0 + u) * ld + c0 + v] = acc[t][u * 4 + v];
      }
    }
  } else {
    for (int tile = threadIdx.x; tile < ntile; tile += NT) {
      const int r0 = (tile / nb) * 4, c0 = (tile % nb) * 4;
      double acc[16];
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[e] = 0.0;
      for (int k = 0; k < ld; ++k) {
        double x[4], y[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = X[k * ld + r0 + u];
#pragma unroll
        for (int v = 0; v < 4; ++v) y[v] = Y[k * ld + c0 + v];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[u * 4 + v] += x[u] * y[v];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) out[(r0 + u) * ld + c0 + v] = acc[u * 4 + v];
    }
  }
  __syncthreads();
}

// P = sum_t C[:, occ_t] f_t C[:, occ_t]^T over the nocc occupied MO slots
// (get_density_matrix, src/wavefunction/type.cu:17-35 text)
template <int NT>
__device__ void density(const double* C, const double* f, const int* occ, int nocc, double* P,
                        int ld)
{
  const int nb = ld >> 2;
  for (int tile = threadIdx.x; tile < nb * nb; tile += NT) {
    const int r0 = (tile / nb) * 4, c0 = (tile % nb) * 4;
    double acc[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.0;
    for (int t = 0; t < nocc; ++t) {
      const int k = occ[t];
      const double fk = f[k];
      double x[4], y[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) x[u] = C[(r0 + u) * ld + k] * fk;
#pragma unroll
      for (int v = 0; v < 4; ++v) y[v] = C[(c0 + v) * ld + k];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u * 4 + v] += x[u] * y[v];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) P[(r0 + u) * ld + c0 + v] = acc[u * 4 + v];
  }
  __syncthreads();
}

// ------------------------------------------------------------ Cholesky
// S = L L^T in place (lower), then C = L^-T.  Returns false if S is not
// positive definite.
template <int NT>
__device__ bool orthonormalizer(double* A, double* C, int n, int ld, double* scal)
{
  for (int j = 0; j < n; ++j) {
    if (threadIdx.x == 0) {
      const double d = A[j * ld + j];
      scal[0] = (d > 0.0) ? sqrt(d) : -1.0;
    }
    __syncthreads();
    const double d = scal[0];
    if (!(d > 0.0)) return false;
    for (int i = j + 1 + threadIdx.x; i < n; i += NT) A[i * ld + j] /= d;
    if (threadIdx.x == 0) A[j * ld + j] = d;
    __syncthreads();
    // trailing update of the lower triangle
    const int m = n - j - 1;
    for (int idx = threadIdx.x; idx < m * m; idx += NT) {
      const int r = j + 1 + idx / m, c = j + 1 + idx % m;
      if (c <= r) A[r * ld + c] -= A[r * ld + j] * A[c * ld + j];
    }
    __syncthreads();
  }
  // row j of C = column j of L^-1 (forward substitution, one thread per column)
  for (int j = threadIdx.x; j < ld; j += NT) {
    double* y = C + j * ld;
    for (int i = 0; i < ld; ++i) y[i] = 0.0;
    if (j < n) {
      for (int i = j; i < n; ++i) {
        double s = (i == j) ? 1.0 : 0.0;
        for (int k = j; k < i; ++k) s -= A[i * ld + k] * y[k];
        y[i] = s / A[i * ld + i];
      }
    }
  }
  __syncthreads();
  return true;
}

// ---------------------------------------------------------------- Jacobi
// Cyclic parallel Jacobi with round-robin pairing: each round applies ld/2
// disjoint plane rotations; A is updated in 2x2 blocks (row pair, column
// pair) and the rotations are accumulated into the columns of C.
__device__ __forceinline__ int rr_player(int pos, int r, int ld)
{
  return pos == 0 ? 0 : 1 + ((pos - 1 + r) % (ld - 1));
}

template <int NT>
__device__ bool jacobi(double* A, double* C, int ld, double* rc, double* rs, double* rt, int* rp,
                       int* rq, double* red, int* sweeps_out)
{
  const int h = ld >> 1;
  int sweep = 0;
  bool ok = false;
  for (; sweep < JACOBI_MAX_SWEEPS; ++sweep) {
    double offmax = 0.0;
    for (int r = 0; r < ld - 1; ++r) {
      for (int k = threadIdx.x; k < h; k += NT) {
        int p = rr_player(k, r, ld), q = rr_player(ld - 1 - k, r, ld);
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        const double apq = A[p * ld + q];
        double c = 1.0, s = 0.0, t = 0.0;
        offmax = fmax(offmax, fabs(apq));
        if (apq != 0.0) {
          const double app = A[p * ld + p], aqq = A[q * ld + q];
          const double theta = (aqq - app) / (2.0 * apq);
          if (fabs(theta) > 1.0e150) {
            t = 0.5 / theta;
          } else {
            t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
            if (theta < 0.0) t = -t;
          }
          c = 1.0 / sqrt(t * t + 1.0);
          s = t * c;
        }
        rc[k] = c;
        rs[k] = s;
        rt[k] = t;
        rp[k] = p;
        rq[k] = q;
      }
      __syncthreads();
      const int nblk = h * h;
      const int total = nblk + ld * h;
      for (int it = threadIdx.x; it < total; it += NT) {
        if (it < nblk) {
          const int k = it / h, l = it - (it / h) * h;
          const int pk = rp[k], qk = rq[k];
          if (k == l) {
            const double t = rt[k];
            const double apq = A[pk * ld + qk];
            A[pk * ld + pk] -= t * apq;
            A[qk * ld + qk] += t * apq;
            A[pk * ld + qk] = 0.0;
            A[qk * ld + pk] = 0.0;
          } else {
            const int pl = rp[l], ql = rq[l];
            const double ck = rc[k], sk = rs[k], cl = rc[l], sl = rs[l];
            const double x00 = A[pk * ld + pl], x01 = A[pk * ld + ql];
            const double x10 = A[qk * ld + pl], x11 = A[qk * ld + ql];
            const double y00 = ck * x00 - sk * x10, y01 = ck * x01 - sk * x11;
            const double y10 = sk * x00 + ck * x10, y11 = sk * x01 + ck * x11;
            A[pk * ld + pl] = cl * y00 - sl * y01;
            A[pk * ld + ql] = sl * y00 + cl * y01;
            A[qk * ld + pl] = cl * y10 - sl * y11;
            A[qk * ld + ql] = sl * y10 + cl * y11;
          }
        } else {
          const int it2 = it - nblk;
          const int row = it2 / h, k = it2 - (it2 / h) * h;
          const int pk = rp[k], qk = rq[k];
          const double ck = rc[k], sk = rs[k];
          const double u = C[row * ld + pk], v = C[row * ld + qk];
          C[row * ld + pk] = ck * u - sk * v;
          C[row * ld + qk] = sk * u + ck * v;
        }
      }
      __syncthreads();
    }
    offmax = block_max(offmax, red);
    if (offmax < JACOBI_TOL) {
      ok = true;
      ++sweep;
      break;
    }
  }
  *sweeps_out = sweep;
  return ok;
}

// ----------------------------------------------------------------- Fermi
// get_fermi_filling (src/wavefunction/fermi.cu:100-124) on sorted levels,
// run by one wave; occ receives this spin channel's occupations, returns
// the electronic entropy term kT * sum(f ln f + (1-f) ln(1-f)).
__device__ double fermi_wave(double nel, double kt, int n, const double* emo, double* occ)
{
  const int lane = threadIdx.x & 63;
  const double thr = 1.4901161193847656e-08; // sqrt(epsilon(1.0_wp))
  int homo = (int)floor(nel);
  const double frac = fmod(nel, 1.0);
  for (int i = lane; i < n; i += 64) occ[i] = (i < homo) ? 1.0 : ((i == homo) ? frac : 0.0);
  int homo_idx = homo;
  if (frac > 0.5) homo_idx += 1;
  if (homo_idx > 0) {
    const double occt = (double)homo_idx;
    const int lo = homo_idx - 1 < 0 ? 0 : homo_idx - 1;
    const int hi = homo_idx < n - 1 ? homo_idx : n - 1;
    double ef = 0.5 * (emo[lo] + emo[hi]);
    for (int cyc = 0; cyc < 200; ++cyc) {
      double tn = 0.0, td = 0.0;
      for (int i = lane; i < n; i += 64) {
        const double x = (emo[i] - ef) / kt;
        double f = 0.0, dfn = 0.0;
        if (x < 50.0) {
          const double ex = exp(x);
          f = 1.0 / (ex + 1.0);
          dfn = ex / (kt * (ex + 1.0) * (ex + 1.0));
        }
        occ[i] = f;
        tn += f;
        td += dfn;
      }
      tn = wave_sum(tn);
      td = wave_sum(td);
      ef += (occt - tn) / td;
      if (fabs(occt - tn) <= thr) break;
    }
  }
  double s = 0.0;
  for (int i = lane; i < n; i += 64) {
    const double f = occ[i];
    if (f > thr && 1.0 - f > thr) s += f * log(f) + (1.0 - f) * log(1.0 - f);
  }
  return wave_sum(s) * kt;
}

// ---------------------------------------------------------------- kernel
template <int NT, int MAXT, bool MATS_LDS>
__global__ __launch_bounds__(NT) void k_scf(BatchDev b, const int* __restrict__ mols, int ldmax,
                                            int natmax, int nshmax)
{
  extern __shared__ double smem[];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wid = tid >> 6;
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
  const int64_t p2 = (int64_t)nat * nat;
  v.sd = b.aes_sd + 3 * b.mol_pair_off[m];
  v.dd = b.aes_dd + 9 * b.mol_pair_off[m];
  v.sq = b.aes_sq + 6 * b.mol_pair_off[m];
  v.qsh = b.qsh + v.s0;
  v.qat = b.qat + v.a0;
  v.dpat = b.dpat + 3 * (int64_t)v.a0;
  v.qpat = b.qpat + 6 * (int64_t)v.a0;
  const DevParams& P = *b.par;

  double* A;
  double* C;
  double* T;
  if constexpr (MATS_LDS) {
    A = smem + L.a;
    C = smem + L.c;
    T = A;
  } else {
    A = b.big_a + b.mol_big_off[m];
    C = b.big_c + b.mol_big_off[m];
    T = b.big_t + b.mol_big_off[m];
  }
  double* red = smem + L.red;
  double* scal = smem + L.scal;
  double* ev = smem + L.ev;
  double* focc_u = smem + L.focc;
  double* emo_s = smem + L.emo;
  double* fs = smem + L.fs;
  double* rc = smem + L.rc;
  double* rs = smem + L.rs;
  double* rt = smem + L.rt;
  double* work = smem + L.work;
  int* rank = ismem + L.rank;
  int* inv = ismem + L.inv;
  int* occl = ismem + L.occ;
  int* rp = ismem + L.rp;
  int* rq = ismem + L.rq;

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

  // repulsion total
  double erep = 0.0;
  for (int i = tid; i < nat; i += NT) erep += b.erep[v.a0 + i];
  erep = block_sum(erep, red);

  // --- initial guess: zero charges and multipoles (superposition of neutral atoms)
  for (int i = tid; i < nsh; i += NT) v.qsh[i] = 0.0;
  for (int i = tid; i < nat; i += NT) v.qat[i] = 0.0;
  for (int i = tid; i < 3 * nat; i += NT) v.dpat[i] = 0.0;
  for (int i = tid; i < 6 * nat; i += NT) v.qpat[i] = 0.0;
  for (int i = tid; i < ndim; i += NT) qin[i] = 0.0;

  // --- orthonormal start basis C = L^-T
  for (int idx = tid; idx < ld * ld; idx += NT) {
    const int i = idx / ld, j = idx - (idx / ld) * ld;
    A[idx] = (i < n && j < n) ? v.S[(int64_t)i * n + j] : 0.0;
  }
  __syncthreads();
  if (!orthonormalizer<NT>(A, C, n, ld, scal)) status |= TA_STATUS_CHOLESKY;

  const double econv = 1.0e-6 * b.accuracy, pconv = 2.0e-5 * b.accuracy;
  const double kt = b.kt;
  double esum = 0.0, elast = 0.0;
  int iter = 0, mixiter = 0;
  bool converged = false;
  int last_nocc = 0;

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

    // ------------------------------------------------ potential
    double* vao = work;
    double* vsh = work + ld;
    double* vat = vsh + nshmax;
    double* vdp = vat + natmax;
    double* vqp = vdp + 3 * natmax;
    for (int s = tid; s < nsh; s += NT) {
      double acc = 0.0;
      for (int j = 0; j < nsh; ++j) acc += v.gam[(int64_t)j * nsh + s] * v.qsh[j];
      const int at = v.sh_at[s];
      const double q = v.qsh[s];
      acc += q * q * P.el[v.el[at]].hubbard_derivs[v.sh_ish[s]];
      vsh[s] = acc;
    }
    for (int a = tid; a < nat; a += NT) {
      const DevElem& e = P.el[v.el[a]];
      double va = 0.0, d0 = 0.0, d1 = 0.0, d2 = 0.0;
      double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0, q4 = 0.0, q5 = 0.0;
      for (int i = 0; i < nat; ++i) {
        const int64_t ia = (int64_t)i * nat + a;
        const double qi = v.qat[i];
        const double s0_ = v.sd[ia], s1_ = v.sd[p2 + ia], s2_ = v.sd[2 * p2 + ia];
        // amat_sd(k, i, a) = -amat_sd(k, a, i)
        va -= s0_ * v.dpat[3 * i] + s1_ * v.dpat[3 * i + 1] + s2_ * v.dpat[3 * i + 2];
        d0 += s0_ * qi;
        d1 += s1_ * qi;
        d2 += s2_ * qi;
        const double dpi0 = v.dpat[3 * i], dpi1 = v.dpat[3 * i + 1], dpi2 = v.dpat[3 * i + 2];
        d0 += v.dd[0 * p2 + ia] * dpi0 + v.dd[1 * p2 + ia] * dpi1 + v.dd[2 * p2 + ia] * dpi2;
        d1 += v.dd[3 * p2 + ia] * dpi0 + v.dd[4 * p2 + ia] * dpi1 + v.dd[5 * p2 + ia] * dpi2;
        d2 += v.dd[6 * p2 + ia] * dpi0 + v.dd[7 * p2 + ia] * dpi1 + v.dd[8 * p2 + ia] * dpi2;
        const double t0 = v.sq[ia], t1 = v.sq[p2 + ia], t2 = v.sq[2 * p2 + ia];
        const double t3 = v.sq[3 * p2 + ia], t4 = v.sq[4 * p2 + ia], t5 = v.sq[5 * p2 + ia];
        va += t0 * v.qpat[6 * i] + t1 * v.qpat[6 * i + 1] + t2 * v.qpat[6 * i + 2] +
              t3 * v.qpat[6 * i + 3] + t4 * v.qpat[6 * i + 4] + t5 * v.qpat[6 * i + 5];
        q0 += t0 * qi;
        q1 += t1 * qi;
        q2 += t2 * qi;
        q3 += t3 * qi;
        q4 += t4 * qi;
        q5 += t5 * qi;
      }
      vat[a] = va;
      vdp[3 * a + 0] = d0 + 2.0 * e.dkernel * v.dpat[3 * a + 0];
      vdp[3 * a + 1] = d1 + 2.0 * e.dkernel * v.dpat[3 * a + 1];
      vdp[3 * a + 2] = d2 + 2.0 * e.dkernel * v.dpat[3 * a + 2];
      vqp[6 * a + 0] = q0 + 2.0 * e.qkernel * v.qpat[6 * a + 0];
      vqp[6 * a + 1] = q1 + 4.0 * e.qkernel * v.qpat[6 * a + 1];
      vqp[6 * a + 2] = q2 + 2.0 * e.qkernel * v.qpat[6 * a + 2];
      vqp[6 * a + 3] = q3 + 4.0 * e.qkernel * v.qpat[6 * a + 3];
      vqp[6 * a + 4] = q4 + 4.0 * e.qkernel * v.qpat[6 * a + 4];
      vqp[6 * a + 5] = q5 + 2.0 * e.qkernel * v.qpat[6 * a + 5];
    }
    __syncthreads();
    for (int i = tid; i < n; i += NT) vao[i] = vsh[v.ao_sh[i]] + vat[v.ao_at[i]];
    // set_mixer: q_in = current charges
    if (iter == 1)
      for (int i = tid; i < ndim; i += NT) qin[i] = 0.0;
    __syncthreads();

    // ------------------------------------------------ Fock matrix
    // A1[i][j] = (H_ij - S_ij (v_i + v_j)/2)/2 - (D(.,j,i).vdp(at i) + Q(.,j,i).vqp(at i))/2
    for (int idx = tid; idx < ld * ld; idx += NT) {
      const int i = idx / ld, j = idx - (idx / ld) * ld;
      double f = 0.0;
      if (i < n && j < n) {
        const int64_t e = (int64_t)i * n + j;
        const int ai = v.ao_at[i];
        const double* vd = vdp + 3 * ai;
        const double* vq = vqp + 6 * ai;
        f = 0.5 * (v.H[e] - v.S[e] * 0.5 * (vao[i] + vao[j]));
        f -= 0.5 * (v.D[e] * vd[0] + v.D[n2 + e] * vd[1] + v.D[2 * n2 + e] * vd[2]);
        f -= 0.5 * (v.Q[e] * vq[0] + v.Q[n2 + e] * vq[1] + v.Q[2 * n2 + e] * vq[2] +
                    v.Q[3 * n2 + e] * vq[3] + v.Q[4 * n2 + e] * vq[4] + v.Q[5 * n2 + e] * vq[5]);
      }
      A[idx] = f;
    }
    __syncthreads();
    for (int idx = tid; idx < ld * ld; idx += NT) {
      const int i = idx / ld, j = idx - (idx / ld) * ld;
      if (i < j) {
        const double t = A[i * ld + j] + A[j * ld + i];
        A[i * ld + j] = t;
        A[j * ld + i] = t;
      } else if (i == j) {
        A[idx] *= 2.0;
      }
    }
    __syncthreads();

    // ------------------------------------------------ eigensolve
    if constexpr (MATS_LDS) {
      mm_nn<NT, MAXT, true>(A, C, A, ld);   // A <- F C
      mm_tn<NT, MAXT, true>(C, A, A, ld);   // A <- C^T F C
    } else {
      mm_nn<NT, 1, false>(A, C, T, ld);
      mm_tn<NT, 1, false>(C, T, A, ld);
    }
    int sweeps = 0;
    if (!jacobi<NT>(A, C, ld, rc, rs, rt, rp, rq, red, &sweeps)) {
      status |= TA_STATUS_EIGEN;
      break;
    }
    for (int k = tid; k < n; k += NT) ev[k] = A[k * ld + k];
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

    // ------------------------------------------------ occupations
    double ts = 0.0;
    if (wid == 0) {
      double* fsp = fs;
      ts = fermi_wave(nalp, kt, n, emo_s, fsp);
      for (int r = lane; r < n; r += 64) focc_u[inv[r]] = fsp[r];
      ts += fermi_wave(nbet, kt, n, emo_s, fsp);
      for (int r = lane; r < n; r += 64) focc_u[inv[r]] += fsp[r];
      if (lane == 0) scal[3] = ts;
    }
    __syncthreads();
    ts = scal[3];
    if (tid == 0) {
      int c = 0;
      for (int r = 0; r < n; ++r) {
        const int k = inv[r];
        if (focc_u[k] != 0.0) occl[c++] = k;
      }
      scal[4] = (double)c;
    }
    __syncthreads();
    last_nocc = (int)scal[4];

    // ------------------------------------------------ density -> A
    density<NT>(C, focc_u, occl, last_nocc, A, ld);

    // ------------------------------------------------ Mulliken + band energy
    double* part = work; // part[c * ld + i], c = 0 (PS), 1..3 (PD), 4..9 (PQ), 10 (PH)
    for (int i = wid; i < n; i += NW) {
      double acc[11];
#pragma unroll
      for (int c = 0; c < 11; ++c) acc[c] = 0.0;
      const int64_t row = (int64_t)i * n;
      for (int j = lane; j < n; j += 64) {
        const double p = A[i * ld + j];
        const int64_t e = row + j;
        acc[0] += p * v.S[e];
        acc[1] += p * v.D[e];
        acc[2] += p * v.D[n2 + e];
        acc[3] += p * v.D[2 * n2 + e];
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[4 + c] += p * v.Q[c * n2 + e];
        acc[10] += p * v.H[e];
      }
#pragma unroll
      for (int c = 0; c < 11; ++c) acc[c] = wave_sum(acc[c]);
      if (lane == 0)
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
    }
    __syncthreads();

    // ------------------------------------------------ electrostatic energy
    double ec = 0.0;
    for (int s = tid; s < nsh; s += NT) {
      double acc = 0.0;
      for (int j = 0; j < nsh; ++j) acc += v.gam[(int64_t)j * nsh + s] * v.qsh[j];
      const double q = v.qsh[s];
      ec += 0.5 * acc * q;
      ec += q * q * q * P.el[v.el[v.sh_at[s]]].hubbard_derivs[v.sh_ish[s]] / 3.0;
    }
    for (int a = tid; a < nat; a += NT) {
      const DevElem& e = P.el[v.el[a]];
      double d[3] = {0.0, 0.0, 0.0}, q[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      for (int i = 0; i < nat; ++i) {
        const int64_t ia = (int64_t)i * nat + a;
        const double qi = v.qat[i];
        const double dpi0 = v.dpat[3 * i], dpi1 = v.dpat[3 * i + 1], dpi2 = v.dpat[3 * i + 2];
        for (int k = 0; k < 3; ++k)
          d[k] += v.sd[k * p2 + ia] * qi +
                  0.5 * (v.dd[(3 * k + 0) * p2 + ia] * dpi0 + v.dd[(3 * k + 1) * p2 + ia] * dpi1 +
                         v.dd[(3 * k + 2) * p2 + ia] * dpi2);
        for (int k = 0; k < 6; ++k) q[k] += v.sq[k * p2 + ia] * qi;
      }
      const double* dpa = v.dpat + 3 * a;
      const double* qpa = v.qpat + 6 * a;
      ec += dpa[0] * d[0] + dpa[1] * d[1] + dpa[2] * d[2];
      for (int k = 0; k < 6; ++k) ec += qpa[k] * q[k];
      ec += e.dkernel * (dpa[0] * dpa[0] + dpa[1] * dpa[1] + dpa[2] * dpa[2]);
      ec += e.qkernel * (qpa[0] * qpa[0] + 2.0 * qpa[1] * qpa[1] + qpa[2] * qpa[2] +
                         2.0 * qpa[3] * qpa[3] + 2.0 * qpa[4] * qpa[4] + qpa[5] * qpa[5]);
    }
    ec = block_sum(ec, red);
    eband = block_sum(eband, red);
    err = block_sum(err, red);
    esum = eband + ts + ec;
    const double perr = sqrt(err / ndim);
    converged = (fabs(esum - elast) < econv) && (perr < pconv);
    if (!isfinite(esum)) status |= TA_STATUS_NAN;
  }
  if (!converged && status == 0) status |= TA_STATUS_NOT_CONVERGED;

  // ------------------------------------------------ results
  for (int r = tid; r < n; r += NT) {
    b.emo[v.o0 + r] = iter > 0 ? emo_s[r] : 0.0;
    b.focc[v.o0 + r] = iter > 0 ? focc_u[inv[r]] : 0.0;
  }
  if (b.coeff != nullptr && iter > 0) {
    double* co = b.coeff + moff;
    for (int idx = tid; idx < n * n; idx += NT) {
      const int i = idx / n, r = idx - (idx / n) * n;
      co[idx] = C[i * ld + inv[r]];
    }
  }
  if (tid == 0) {
    b.energy[m] = erep + esum;
    b.eelec[m] = esum;
    b.erep_tot[m] = erep;
    b.iterations[m] = iter;
    b.status[m] = status;
  }
}

// explicit instantiations, launched by the host per size bucket
template __global__ void k_scf<128, 1, true>(BatchDev, const int*, int, int, int);
template __global__ void k_scf<256, 1, true>(BatchDev, const int*, int, int, int);
template __global__ void k_scf<512, 2, true>(BatchDev, const int*, int, int, int);
template __global__ void k_scf<512, 1, false>(BatchDev, const int*, int, int, int);

// Multi-workgroup SCF for one molecule above the single-workgroup size
// (padded AO dimension > TA_BIG_MAXLD, e.g. a (H2O)700 cluster with 4200
// AOs).  The iteration of next_scf (src/scf/iterators.cu:317-424) runs as a
// host loop over kernels, each phase spread over the whole GPU:
//
//   mixing        modified Broyden (one workgroup; O(ndim x memory))
//   potential     one workgroup per shell (ES2 + ES3) and per atom (AES2)
//   Fock          grid-stride over the n^2 elements, then symmetrised
//   reduction     A = X F X on the matrix cores, X = S^-1/2 (Loewdin; its
//                 S-orthonormal columns serve exactly as the Cholesky basis)
//   tridiagonal   blocked Householder: per column one workgroup forms the
//                 reflector, the symmetric product is tiled over 64 x 64
//                 blocks of the trailing triangle (one workgroup each), one
//                 workgroup folds the partial sums and the panel corrections;
//                 after NP columns a rank-2 NP matrix-core update
//   eigenvalues   bisection, one thread per level, warm-started
//   eigenvectors  inverse iteration, one thread per level; runs of close
//                 levels orthonormalised in panels (block Gram-Schmidt on
//                 the matrix cores, Cholesky QR per panel)
//   back-transf.  compact WY blocks, three matrix-core products per block
//   occupation    Fermi smearing (one wave), density on the matrix cores
//   Mulliken      one wave per AO row, then one workgroup for shells / atoms
//   energy        one wave per shell / atom, then one workgroup
// The host reads one status word per iteration.
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <vector>

#include "common.h"
#include "d4.h"
#include "device_util.h"
#include "eig_big.h"
#include "fermi.h"
#include "scf_grid.h"

namespace {

constexpr int GT = 512;  // threads of the one-workgroup kernels and GEMM tiles

// ---------------------------------------------------------------- GEMM
template <bool AK, bool BJ, class LA, class LB, class KR, class SK, class ST>
__global__ __launch_bounds__(512) void k_g_gemm(int M, int N, LA la, LB lb, KR kr, SK skip, ST st)
{
  __shared__ double tile_[2048];
  lds_dbl* tile = TA_LDS(tile_);
  const int mt = (M + 63) >> 6, nt = (N + 63) >> 6;
  for (int t = blockIdx.x; t < mt * nt; t += gridDim.x) {
    const int ti = t / nt, tj = t - (t / nt) * nt;
    if (skip(ti, tj)) continue;
    int k0, k1;
    kr(ti, tj, &k0, &k1);
    big::gemm_tile<512, AK, BJ>(M, N, ti * 64, tj * 64, k0, k1, la, lb, st, tile);
  }
}

template <bool AK, bool BJ, class LA, class LB, class KR, class SK, class ST>
void g_gemm(hipStream_t s, int M, int N, LA la, LB lb, KR kr, SK skip, ST st)
{
  const int tiles = ((M + 63) >> 6) * ((N + 63) >> 6);
  const int grid = std::max(1, std::min(tiles, 2048));
  hipLaunchKernelGGL((k_g_gemm<AK, BJ, LA, LB, KR, SK, ST>), dim3(grid), dim3(512), 0, s, M, N,
                     la, lb, kr, skip, st);
}

struct AllK {
  int k;
  __device__ void operator()(int, int, int* a, int* b) const
  {
    *a = 0;
    *b = k;
  }
};
struct NoSkip {
  __device__ bool operator()(int, int) const { return false; }
};
struct UpperSkip {
  __device__ bool operator()(int ti, int tj) const { return tj > ti; }
};

// ---------------------------------------------------------------- init
// zero charges and multipoles; repulsion total, occupation numbers
// (src/xtb.cu:290-313 text) in the device scalars
__global__ __launch_bounds__(256) void k_g_init(GridMol g)
{
  __shared__ double red_[8];
  lds_dbl* red = TA_LDS(red_);
  double nref = 0.0, erep = 0.0, e3 = 0.0;
  const bool d4on = g.par->d4.enabled;
  for (int s = threadIdx.x; s < g.nsh; s += blockDim.x) nref += g.n0sh[s];
  for (int a = threadIdx.x; a < g.nat; a += blockDim.x) {
    erep += g.erep_at[a];
    if (d4on) e3 += g.d4_e3[a];
  }
  nref = block_sum(nref, red);
  erep = block_sum(erep, red);
  e3 = block_sum(e3, red);
  if (threadIdx.x == 0) {
    const double nel = nref - g.charge;
    const int nint_nel = (int)lround(nel);
    const double nuhf =
        ((g.uhf % 2) == (nint_nel % 2)) ? (double)g.uhf : (double)(nint_nel % 2);
    const double diff = fmin(nuhf, nel);
    g.sc[GS_NALP] = (nel - diff) / 2.0 + diff;
    g.sc[GS_NBET] = (nel - diff) / 2.0;
    g.sc[GS_EREP] = erep;
    g.sc[GS_E3] = e3;
    g.sc[GS_ED2] = 0.0;
  }
  for (int i = threadIdx.x; i < g.nsh; i += blockDim.x) g.qsh[i] = 0.0;
  for (int i = threadIdx.x; i < g.nat; i += blockDim.x) g.qat[i] = 0.0;
  for (int i = threadIdx.x; i < 3 * g.nat; i += blockDim.x) g.dpat[i] = 0.0;
  for (int i = threadIdx.x; i < 6 * g.nat; i += blockDim.x) g.qpat[i] = 0.0;
  for (int i = threadIdx.x; i < g.ndim; i += blockDim.x) g.qin[i] = 0.0;
  if (threadIdx.x == 0) {
    g.sc[GS_WARM] = 0.0;
    g.sc[GS_ESUM] = 0.0;
    g.ic[GI_STATUS] = 0;
    g.ic[GI_CONVERGED] = 0;
  }
}

// S (n-strided) -> A (ld-strided, full)
__global__ void k_g_copy_s(GridMol g)
{
  const int64_t tot = (int64_t)g.n * g.n;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < tot;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx / g.n), j = (int)(idx - (idx / g.n) * g.n);
    g.A[(int64_t)i * g.ld + j] = g.S[idx];
  }
}

// ---------------------------------------------------------------- potential
// blocks [0, nsh): ES2 + ES3 shell potential; [nsh, nsh + nat): AES2 atom potentials
__global__ __launch_bounds__(256) void k_g_potential(GridMol g)
{
  __shared__ double red_[8];
  lds_dbl* red = TA_LDS(red_);
  const DevParams& P = *g.par;
  const int nat = g.nat, nsh = g.nsh;
  const int64_t p2 = (int64_t)nat * nat;
  if ((int)blockIdx.x < nsh) {
    const int sh = blockIdx.x;
    const double* gr = g.gam + (int64_t)sh * nsh;
    double acc = 0.0;
    for (int j = threadIdx.x; j < nsh; j += blockDim.x) acc += gr[j] * g.qsh[j];
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) {
      const double q = g.qsh[sh];
      g.vsh[sh] = acc + q * q * P.el[g.el[g.sh_at[sh]]].hubbard_derivs[g.sh_ish[sh]];
    }
    return;
  }
  const int a = blockIdx.x - nsh;
  double va = 0.0, d[3] = {0.0, 0.0, 0.0}, q[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int64_t row = (int64_t)a * nat;
  for (int i = threadIdx.x; i < nat; i += blockDim.x) {
    const int64_t ai = row + i;
    const double qi = g.qat[i];
    const double dpi[3] = {g.dpat[3 * i], g.dpat[3 * i + 1], g.dpat[3 * i + 2]};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double sdk = g.sd[k * p2 + ai];
      va += sdk * dpi[k];
      d[k] -= sdk * qi;
#pragma unroll
      for (int l = 0; l < 3; ++l) d[k] += g.dd[(3 * l + k) * p2 + ai] * dpi[l];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double sqk = g.sq[k * p2 + ai];
      va += sqk * g.qpat[6 * i + k];
      q[k] += sqk * qi;
    }
  }
  va = block_sum(va, red);
#pragma unroll
  for (int k = 0; k < 3; ++k) d[k] = block_sum(d[k], red);
#pragma unroll
  for (int k = 0; k < 6; ++k) q[k] = block_sum(q[k], red);
  if (threadIdx.x == 0) {
    const DevElem& e = P.el[g.el[a]];
    g.vat[a] = va;
#pragma unroll
    for (int k = 0; k < 3; ++k) g.vdp[3 * a + k] = d[k] + 2.0 * e.dkernel * g.dpat[3 * a + k];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double sc = (k == 1 || k == 3 || k == 4) ? 2.0 : 1.0;
      g.vqp[6 * a + k] = q[k] + 2.0 * e.qkernel * g.qpat[6 * a + k] * sc;
    }
  }
}

// ---------------------------------------------------------------- D4
// reference weights at the current atomic charges, one thread per atom
__global__ void k_g_d4_weights(GridMol g)
{
  const DevD4& d4 = g.par->d4;
  for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < g.nat; a += gridDim.x * blockDim.x)
    d4_atom_weights(d4, g.el[a], g.d4_cn[a], g.qat[a], g.d4_gw + TA_D4_MAXREF * a,
               g.d4_dgw + TA_D4_MAXREF * a);
}

// D4 potential dE2/dq_a added to vat (after k_g_potential), one wave per atom
__global__ __launch_bounds__(64) void k_g_d4_potential(GridMol g)
{
  const DevD4& d4 = g.par->d4;
  const int a = blockIdx.x, nat = g.nat, lane = threadIdx.x;
  const int ea = g.el[a];
  double vv[TA_D4_MAXREF] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = lane; i < nat; i += 64) {
    const double dai = g.d4_disp[(int64_t)a * nat + i];
    if (dai != 0.0) d4_scf_accumulate(d4, ea, g.el[i], dai, g.d4_gw + TA_D4_MAXREF * i, vv);
  }
  double va = 0.0;
#pragma unroll
  for (int r = 0; r < TA_D4_MAXREF; ++r) {
    vv[r] = wave_sum(vv[r]);
    if (r < d4_nw(d4, ea)) va += vv[r] * g.d4_dgw[TA_D4_MAXREF * a + r];
  }
  if (lane == 0) g.vat[a] += va;
}

// ---------------------------------------------------------------- Fock
__global__ void k_g_fock(GridMol g)
{
  const int n = g.n;
  const int64_t n2 = (int64_t)n * n;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n2;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx / n), j = (int)(idx - (idx / n) * n);
    const int ai = g.ao_at[i], aj = g.ao_at[j];
    const double vi = g.vsh[g.ao_sh[i]] + g.vat[ai], vj = g.vsh[g.ao_sh[j]] + g.vat[aj];
    const double* vd = g.vdp + 3 * ai;
    const double* vq = g.vqp + 6 * ai;
    double f = 0.5 * (g.H[idx] - g.S[idx] * 0.5 * (vi + vj));
    f -= 0.5 * (g.D[idx] * vd[0] + g.D[n2 + idx] * vd[1] + g.D[2 * n2 + idx] * vd[2]);
    f -= 0.5 * (g.Q[idx] * vq[0] + g.Q[n2 + idx] * vq[1] + g.Q[2 * n2 + idx] * vq[2] +
                g.Q[3 * n2 + idx] * vq[3] + g.Q[4 * n2 + idx] * vq[4] + g.Q[5 * n2 + idx] * vq[5]);
    g.Y[(int64_t)i * g.ld + j] = f;
  }
}

// F = F1 + F1^T from Y into A
__global__ void k_g_symmetrize(GridMol g)
{
  const int n = g.n;
  const int64_t n2 = (int64_t)n * n;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n2;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx / n), j = (int)(idx - (idx / n) * n);
    g.A[(int64_t)i * g.ld + j] = g.Y[(int64_t)i * g.ld + j] + g.Y[(int64_t)j * g.ld + i];
  }
}

// ------------------------------------------------------ tridiagonalisation
// Per column j (panel position i) four kernels: the column brought up to
// date across the device (k_g_tri_colupd), its reflector (k_g_tri_refl, one
// workgroup), the symmetric product with the panel dot products
// (k_g_tri_symv) and y (k_g_tri_rows); w = y - (tau/2)(y.v) v of a column is
// formed by the next column's k_g_tri_colupd, or by k_g_tri_fin at the end
// of a panel.
#define GC_T 256

// Rows [j, n) over all workgroups.  With i > 0 first w of column j - 1
// (rows j..), then x_r = a_rj - sum_{q<i}
// (v_q[r] w_q[j] + w_q[r] v_q[j]) into g.e2 (free until the eigenvalues),
// d[j] = x_j, and each workgroup's partial sum of x_r^2 over r >= j + 2 into
// part[blockIdx.x].
__global__ __launch_bounds__(GC_T) void k_g_tri_colupd(GridMol g, int j, int i)
{
  __shared__ double red_[GC_T / 64];
  lds_dbl* red = TA_LDS(red_);
  const int n = g.n, ld = g.ld, tid = threadIdx.x;
  const int r = j + blockIdx.x * GC_T + tid;
  const gbl_dbl* vt = TA_GBL(g.vt);
  double* wt = g.wt;
  double alpha = 0.0, wj = 0.0;
  if (i > 0) {
    // y.v of column j - 1 over its row blocks (rows j..), in block order
    const int nbp = (n - j + 63) >> 6;
    double sv = 0.0;
    if (tid < 64)
      for (int R = tid; R < nbp; R += 64) sv += g.svp[R];
    sv = block_sum(sv, red);
    alpha = -0.5 * g.tau[j - 1] * sv;
    wj = g.y[j] + alpha * g.v[j];
  }
  double x2 = 0.0;
  if (r < n) {
    double a = g.A[(int64_t)j * ld + r];  // the trailing square is symmetric: row j
    for (int q = 0; q < i; ++q) {
      const double wr = (q == i - 1) ? g.y[r] + alpha * g.v[r] : wt[(int64_t)q * ld + r];
      const double wjq = (q == i - 1) ? wj : wt[(int64_t)q * ld + j];
      a -= vt[(int64_t)q * ld + r] * wjq + wr * vt[(int64_t)q * ld + j];
    }
    if (i > 0) wt[(int64_t)(i - 1) * ld + r] = g.y[r] + alpha * g.v[r];
    g.e2[r] = a;
    if (r == j) g.d[j] = a;
    if (r >= j + 2) x2 = a * a;
  }
  x2 = block_sum(x2, red);
  if (tid == 0) g.part[blockIdx.x] = x2;
}

// Reflector of column j from x = g.e2[j+1..n) (dlarfg, as big::householder):
// stored as A column j, panel row i of vt and g.v; e[j], tau[j].  Rows <= j
// of the panel vectors are never read on this path (the column update, the
// products and the panel GEMM read rows > j of vector q >= the panel's first
// column), so they are not cleared.
__global__ __launch_bounds__(512) void k_g_tri_refl(GridMol g, int j, int i)
{
  __shared__ double red_[16];
  __shared__ double scal_[2];
  lds_dbl* red = TA_LDS(red_);
  lds_dbl* scal = TA_LDS(scal_);
  const int n = g.n, ld = g.ld, tid = threadIdx.x;
  const int nwg = (n - j + GC_T - 1) / GC_T;
  double s = 0.0;
  if (tid < 64)
    for (int b = tid; b < nwg; b += 64) s += g.part[b];
  s = block_sum(s, red);
  if (tid == 0) {
    const double alpha = g.e2[j + 1];
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
    g.e[j] = beta;
    g.tau[j] = tau;
    scal[0] = scale;
  }
  __syncthreads();
  const double scale = scal[0];
  for (int r = j + 1 + tid; r < n; r += GT) {
    const double v = (r == j + 1) ? 1.0 : g.e2[r] * scale;
    g.A[(int64_t)r * ld + j] = v;
    g.vt[(int64_t)i * ld + r] = v;
    g.v[r] = v;
  }
}

// Symmetric product over 64 x 64 tiles (bi >= bj) of the trailing triangle
// [lo, n): tile (bi, bj) writes its row sums to part[bi][bj] and, off the
// diagonal, its column sums to part[bj][bi] (64 doubles each).
__global__ __launch_bounds__(256) void k_g_tri_symv(GridMol g, int lo, int i)
{
  __shared__ double t_[64 * 65];
  __shared__ double vr_[64], vc_[64];
  lds_dbl* T = TA_LDS(t_);
  lds_dbl* vr = TA_LDS(vr_);
  lds_dbl* vc = TA_LDS(vc_);
  const int n = g.n, ld = g.ld, m = n - lo;
  const int nb = (m + 63) >> 6;
  const int t = blockIdx.x;
  if (t >= nb * (nb + 1) / 2) return;
  int bi = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while (bi * (bi + 1) / 2 > t) --bi;
  while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
  const int bj = t - bi * (bi + 1) / 2;
  const int r0 = lo + 64 * bi, c0 = lo + 64 * bj;
  const int tid = threadIdx.x;
  for (int idx = tid; idx < 4096; idx += 256) {
    const int r = idx >> 6, c = idx & 63;
    const int gr = r0 + r, gc = c0 + c;
    double a = 0.0;
    if (gr < n && gc < n) {
      if (gc <= gr) a = g.A[(int64_t)gr * ld + gc];
      else a = g.A[(int64_t)gc * ld + gr];  // diagonal tile: mirror of the lower triangle
    }
    T[r * 65 + c] = a;
  }
  if (tid < 64) {
    vr[tid] = (r0 + tid < n) ? g.v[r0 + tid] : 0.0;
    vc[tid] = (c0 + tid < n) ? g.v[c0 + tid] : 0.0;
  }
  __syncthreads();
  const int rr = tid >> 2, qq = tid & 3;
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < 16; ++c) s += T[rr * 65 + 16 * qq + c] * vc[16 * qq + c];
  s += __shfl_xor(s, 1, 4);
  s += __shfl_xor(s, 2, 4);
  const int64_t nbm = g.nbmax;
  if (qq == 0) g.part[((int64_t)bi * nbm + bj) * 64 + rr] = s;
  if (bi != bj) {
    double cs = 0.0;
#pragma unroll
    for (int r = 0; r < 16; ++r) cs += T[(16 * qq + r) * 65 + rr] * vr[16 * qq + r];
    cs += __shfl_xor(cs, 1, 4);
    cs += __shfl_xor(cs, 2, 4);
    if (qq == 0) g.part[((int64_t)bj * nbm + bi) * 64 + rr] = cs;
  }
  // diagonal tiles: the block's share of the panel dot products
  // dpart[bi][q] = w_q . v, dpart[bi][NP + q] = v_q . v over its rows
  if (bi == bj && i > 0) {
    const int lane = tid & 63, w = tid >> 6;
    const int gr = r0 + lane;
    const double vl = vr[lane];
    for (int pq = w; pq < 2 * i; pq += 4) {
      const int q = pq < i ? pq : pq - i;
      const double* col = (pq < i ? g.wt : g.vt) + (int64_t)q * ld;
      double t = gr < n ? col[gr] * vl : 0.0;
      t = wave_sum(t);
      if (lane == 0) g.dpart[(int64_t)bi * 2 * TA_BIG_NP + (pq < i ? 0 : TA_BIG_NP) + q] = t;
    }
  }
}

// y = tau (sum of partials - V (W^T v) - W (V^T v)) over one 64-row block
// (blockIdx.x; four waves split the partial columns), the panel dot
// products summed from the diagonal tiles' shares, and the block's share of
// y.v in svp
__global__ __launch_bounds__(256) void k_g_tri_rows(GridMol g, int j, int i)
{
  __shared__ double acc_[4 * 64];
  __shared__ double dacc_[8 * 2 * TA_BIG_NP];
  __shared__ double dots_[2 * TA_BIG_NP];
  lds_dbl* acc = TA_LDS(acc_);
  lds_dbl* dacc = TA_LDS(dacc_);
  lds_dbl* dots = TA_LDS(dots_);
  const int n = g.n, ld = g.ld, lo = j + 1, m = n - lo;
  const int nb = (m + 63) >> 6;
  const int64_t nbm = g.nbmax;
  const int R = blockIdx.x, tid = threadIdx.x, l = tid & 63, w = tid >> 6;
  // panel dots: 32 slots x 8 groups of row blocks, then the 8 partials in order
  static_assert(2 * TA_BIG_NP == 32, "slot layout assumes 16 panel columns");
  {
    const int slot = tid & 31, grp = tid >> 5;
    double s = 0.0;
    if ((slot & 15) < i)
      for (int b = grp; b < nb; b += 8) s += g.dpart[(int64_t)b * 2 * TA_BIG_NP + slot];
    dacc[grp * 32 + slot] = s;
  }
  double s = 0.0;
  for (int q = w; q < nb; q += 4) s += g.part[((int64_t)R * nbm + q) * 64 + l];
  acc[w * 64 + l] = s;
  __syncthreads();
  if (tid < 32) {
    double t = 0.0;
    for (int gi = 0; gi < 8; ++gi) t += dacc[gi * 32 + tid];
    dots[tid] = t;
  }
  __syncthreads();
  if (w == 0) {
    const int r = lo + 64 * R + l;
    const double tj = g.tau[j];
    double sv = 0.0;
    if (r < n) {
      double y = ((acc[l] + acc[64 + l]) + acc[128 + l]) + acc[192 + l];
      for (int q = 0; q < i; ++q)
        y -= g.vt[(int64_t)q * ld + r] * dots[q] + g.wt[(int64_t)q * ld + r] * dots[TA_BIG_NP + q];
      y *= tj;
      g.y[r] = y;
      sv = y * g.v[r];
    }
    sv = wave_sum(sv);
    if (l == 0) g.svp[R] = sv;
  }
}

// w = y - (tau/2)(y.v) v, y.v summed over the row blocks in order
__global__ __launch_bounds__(512) void k_g_tri_fin(GridMol g, int j, int i)
{
  __shared__ double red_[16];
  lds_dbl* red = TA_LDS(red_);
  const int n = g.n, ld = g.ld, lo = j + 1, m = n - lo;
  const int nb = (m + 63) >> 6;
  const double tj = g.tau[j];
  double sv = 0.0;
  if (threadIdx.x < 64)
    for (int R = threadIdx.x; R < nb; R += 64) sv += g.svp[R];
  sv = block_sum(sv, red);
  const double alpha = -0.5 * tj * sv;
  for (int r = lo + threadIdx.x; r < n; r += GT) g.wt[(int64_t)i * ld + r] = g.y[r] + alpha * g.v[r];
  for (int r = threadIdx.x; r <= j; r += GT) {
    g.wt[(int64_t)i * ld + r] = 0.0;
    g.vt[(int64_t)i * ld + r] = 0.0;
  }
}

// last 2 x 2 block after the final panel (np reflectors in the panel)
__global__ void k_g_tri_last(GridMol g, int np)
{
  const int n = g.n, ld = g.ld, tid = threadIdx.x;
  if (tid < 3) {
    const int rr = (tid == 0) ? n - 2 : n - 1, cc = (tid == 2) ? n - 1 : n - 2;
    double a = g.A[(int64_t)rr * ld + cc];
    for (int q = 0; q < np; ++q)
      a -= g.vt[(int64_t)q * ld + rr] * g.wt[(int64_t)q * ld + cc] +
           g.wt[(int64_t)q * ld + rr] * g.vt[(int64_t)q * ld + cc];
    if (tid == 0) g.d[n - 2] = a;
    if (tid == 1) g.e[n - 2] = a;
    if (tid == 2) g.d[n - 1] = a;
  }
}

// ------------------------------------------------------------ eigenvalues
__global__ __launch_bounds__(512) void k_g_gersh(GridMol g)
{
  __shared__ double red_[16];
  lds_dbl* red = TA_LDS(red_);
  const int n = g.n;
  double gl = INFINITY, gu = -INFINITY, emax = 0.0;
  for (int k = threadIdx.x; k < n; k += GT) {
    const double el = k > 0 ? fabs(g.e[k - 1]) : 0.0, er = k < n - 1 ? fabs(g.e[k]) : 0.0;
    gl = fmin(gl, g.d[k] - el - er);
    gu = fmax(gu, g.d[k] + el + er);
    if (k < n - 1) {
      g.e2[k] = g.e[k] * g.e[k];
      emax = fmax(emax, g.e2[k]);
    }
  }
  gl = -block_max(-gl, red);
  gu = block_max(gu, red);
  emax = block_max(emax, red);
  if (threadIdx.x == 0) {
    const double tnorm = fmax(fabs(gl), fabs(gu));
    const double pivmin = 2.2250738585072014e-308 * fmax(1.0, emax);
    const double tol = 1.0e-12 * tnorm + 4.0 * big::EPS * tnorm + pivmin;
    g.sc[GS_GL] = gl - 2.0 * tol;
    g.sc[GS_GU] = gu + 2.0 * tol;
    g.sc[GS_TNORM] = tnorm;
    g.sc[GS_PIVMIN] = pivmin;
    g.sc[GS_TOL] = tol;
    g.ic[GI_SHIFT] = 0;
  }
}

// one thread per level: bisection (warm start from the previous levels)
__global__ __launch_bounds__(256) void k_g_bisect(GridMol g, int warm_ok)
{
  const int n = g.n;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double gl = g.sc[GS_GL], gu = g.sc[GS_GU], pivmin = g.sc[GS_PIVMIN], tol = g.sc[GS_TOL];
  const double wd = warm_ok ? g.sc[GS_WARM] : 0.0;
  const gbl_dbl* d = TA_GBL(g.d);
  const gbl_dbl* e2 = TA_GBL(g.e2);
  double lo = gl, hi = gu;
  if (wd > 0.0) {
    const double a = g.prevlam[i] - wd, b = g.prevlam[i] + wd;
    int ca, cb, cc;
    big::sturm3(d, e2, n, a, b, b, pivmin, ca, cb, cc);
    if (ca <= i && cb > i) {
      lo = a;
      hi = b;
    }
  }
  for (int it = 0; it < 64 && hi - lo > tol; ++it) {
    const double h = hi - lo;
    const double x1 = lo + 0.25 * h, x2 = lo + 0.5 * h, x3 = lo + 0.75 * h;
    int c1, c2, c3;
    big::sturm3(d, e2, n, x1, x2, x3, pivmin, c1, c2, c3);
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
  const double l = 0.5 * (lo + hi);
  g.lam[i] = l;
  if (wd > 0.0) {
    // non-negative doubles order like their bit patterns
    const double sh = fabs(l - g.prevlam[i]);
    atomicMax((unsigned long long*)(g.ic + GI_SHIFT), (unsigned long long)__double_as_longlong(sh));
  }
}

__global__ void k_g_warm(GridMol g, int warm_ok)
{
  const double tnorm = g.sc[GS_TNORM];
  if (threadIdx.x == 0) {
    const double sh = __longlong_as_double(*(const long long*)(g.ic + GI_SHIFT));
    g.sc[GS_WARM] = (warm_ok && g.sc[GS_WARM] > 0.0) ? 2.0 * sh + 1.0e-10 * tnorm : 1.0e-2 * tnorm;
  }
  for (int i = threadIdx.x; i < g.n; i += blockDim.x) g.prevlam[i] = g.lam[i];
}

// ---------------------------------------------------------- eigenvectors
// one thread per level; factors in scr ([k][i], three planes of n x n)
__global__ __launch_bounds__(256) void k_g_invit(GridMol g, int nvec)
{
  const int n = g.n, ld = g.ld;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nvec) return;
  const double tnorm = g.sc[GS_TNORM];
  const double pert = big::EPS * tnorm + 1e-300;
  const int64_t pl = (int64_t)n * n;
  double* u0 = g.scr;
  double* u1 = g.scr + pl;
  double* u2 = g.scr + 2 * pl;
  double* Y = g.Y;
  const double l = g.lam[i];
  double scale = 1.0;
  for (int it = 0; it < 3; ++it) {
    double p = g.d[0] - l, q = n > 1 ? g.e[0] : 0.0;
    double bp = it == 0 ? big::start_vector(i, 0) : Y[i] * scale;
    for (int k = 0; k < n - 1; ++k) {
      const double sk = g.e[k], dk1 = g.d[k + 1] - l, ek1 = (k + 1 < n - 1) ? g.e[k + 1] : 0.0;
      const double bk1 = it == 0 ? big::start_vector(i, k + 1) : Y[(int64_t)(k + 1) * ld + i] * scale;
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
      u0[(int64_t)k * n + i] = 1.0 / U0;
      u1[(int64_t)k * n + i] = U1;
      u2[(int64_t)k * n + i] = U2;
      Y[(int64_t)k * ld + i] = rk;
    }
    if (fabs(p) < pert) p = p < 0.0 ? -pert : pert;
    double x2 = 0.0, x1 = bp / p, nrm = x1 * x1;
    Y[(int64_t)(n - 1) * ld + i] = x1;
    for (int k = n - 2; k >= 0; --k) {
      const double x = (Y[(int64_t)k * ld + i] - u1[(int64_t)k * n + i] * x1 -
                        u2[(int64_t)k * n + i] * x2) *
                       u0[(int64_t)k * n + i];
      Y[(int64_t)k * ld + i] = x;
      nrm += x * x;
      x2 = x1;
      x1 = x;
    }
    scale = 1.0 / sqrt(nrm);
  }
  for (int k = 0; k < n; ++k) Y[(int64_t)k * ld + i] *= scale;
}

// runs of close levels (gap <= ORTOL |T|) with more than one member among
// the first nvec
__global__ void k_g_runs(GridMol g, int nvec)
{
  if (threadIdx.x != 0) return;
  const double gap = big::ORTOL * g.sc[GS_TNORM];
  int nr = 0, i = 0;
  while (i < nvec) {
    int j = i + 1;
    while (j < nvec && g.lam[j] - g.lam[j - 1] <= gap) ++j;
    if (j - i > 1) {
      g.runs[2 * nr] = i;
      g.runs[2 * nr + 1] = j;
      ++nr;
    }
    i = j;
  }
  g.ic[GI_NRUNS] = nr;
}

// ------------------------------------------------ cluster orthogonalisation
// The vectors of a run of close levels span the run's invariant subspace but
// are only approximately orthogonal after inverse iteration.  A run gets an
// orthonormal basis panel by panel (OW columns): the run's earlier columns
// are projected out of the panel twice (block classical Gram-Schmidt; the
// products run on the matrix cores, split over row chunks), then the panel
// is orthonormalised by Cholesky QR, twice, in one workgroup.  A panel whose
// Gram matrix is numerically singular (a pivot below OPIV of its diagonal)
// is orthonormalised column by column instead (classical Gram-Schmidt, two
// passes).
constexpr int OW = 64;     // panel width
constexpr int OT = 256;    // threads of the panel kernel
constexpr int OR = 16;     // rows staged per step in the panel kernel
constexpr int OCH = 256;   // rows per chunk of the split projection products
constexpr double OPIV = 1.0e-14;

// LDS of the panel kernel: Gram / Cholesky factor, its inverse, the Gram
// diagonal, a reduction area and a flag word.
constexpr int orth_panel_lds_doubles() { return 2 * OW * OW + OW + 4 * OW + 2; }

// Column-by-column classical Gram-Schmidt, two passes, of the w columns at
// c0 (the fallback for a numerically dependent panel).
__device__ void orth_panel_cgs(gbl_dbl* Y, int n, int ld, int c0, int w, lds_dbl* part,
                               lds_dbl* red)
{
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  constexpr int NWV = OT / 64;
  for (int j = 0; j < w; ++j) {
    for (int pass = 0; pass < 2 && j > 0; ++pass) {
      double s = 0.0;
      if (lane < j)
        for (int k = wv; k < n; k += NWV)
          s += Y[(int64_t)k * ld + c0 + lane] * Y[(int64_t)k * ld + c0 + j];
      part[wv * OW + lane] = s;
      __syncthreads();
      if (t < OW) {
        double c = 0.0;
        for (int q = 0; q < NWV; ++q) c += part[q * OW + t];
        red[t] = c;
      }
      __syncthreads();
      for (int k = t; k < n; k += OT) {
        double y = Y[(int64_t)k * ld + c0 + j];
        for (int l = 0; l < j; ++l) y -= red[l] * Y[(int64_t)k * ld + c0 + l];
        Y[(int64_t)k * ld + c0 + j] = y;
      }
      __syncthreads();
    }
    double s = 0.0;
    for (int k = t; k < n; k += OT) {
      const double y = Y[(int64_t)k * ld + c0 + j];
      s += y * y;
    }
    s = block_sum(s, red + OW);
    const double inv = s > 0.0 ? 1.0 / sqrt(s) : 0.0;
    for (int k = t; k < n; k += OT) Y[(int64_t)k * ld + c0 + j] *= inv;
    __syncthreads();
  }
}

// Orthonormalises the w <= OW columns at c0 of Y (n rows).  col0 < 0: the
// first panel of run blockIdx.x (runs of any length), else the panel
// (col0, w_in).  Dynamic LDS: orth_panel_lds_doubles().
__global__ __launch_bounds__(OT) void k_g_orth_panel(GridMol g, int col0, int w_in)
{
  extern __shared__ double smem[];
  lds_dbl* G = TA_LDS(smem);   // OW x OW: Gram matrix, then R (upper), staging rows
  lds_dbl* X = G + OW * OW;    // OW x OW: R^-1 (upper)
  lds_dbl* dg = X + OW * OW;   // Gram diagonal
  lds_dbl* red = dg + OW;      // 4 OW
  lds_dbl* flag = red + 4 * OW;
  int c0 = col0, w = w_in;
  if (col0 < 0) {
    if ((int)blockIdx.x >= g.ic[GI_NRUNS]) return;
    c0 = g.runs[2 * blockIdx.x];
    w = min(OW, g.runs[2 * blockIdx.x + 1] - c0);
  }
  const int n = g.n, ld = g.ld, t = threadIdx.x;
  gbl_dbl* Y = TA_GBL(g.Y);
  // Gram blocks: thread (bi, bj) owns G[4bi..4bi+3][4bj..4bj+3]
  const int bi = t >> 4, bj = t & 15;
  const bool gown = 4 * bi < w && 4 * bj < w;
  // product blocks: thread owns rows 2rq, 2rq+1 of a staged chunk, columns 4cb..4cb+3
  const int rq = t >> 4, cb = t & 15;
  for (int pass = 0; pass < 2; ++pass) {
    double acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.0;
    for (int k0 = 0; k0 < n; k0 += OR) {
      const int rows = min(OR, n - k0);
      for (int idx = t; idx < OR * OW; idx += OT) {
        const int r = idx >> 6, l = idx & 63;
        G[idx] = (r < rows && l < w) ? Y[(int64_t)(k0 + r) * ld + c0 + l] : 0.0;
      }
      __syncthreads();
      if (gown) {
        for (int r = 0; r < OR; ++r) {
          double a[4], b[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            a[u] = G[r * OW + 4 * bi + u];
            b[u] = G[r * OW + 4 * bj + u];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[4 * u + v] += a[u] * b[v];
        }
      }
      __syncthreads();
    }
    if (gown) {
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) G[(4 * bi + u) * OW + 4 * bj + v] = acc[4 * u + v];
    }
    if (t == 0) flag[0] = 0.0;
    __syncthreads();
    if (t < w) dg[t] = G[t * OW + t];
    __syncthreads();
    // Cholesky G = R^T R, R in the upper triangle of G
    for (int p = 0; p < w; ++p) {
      if (t == 0) {
        const double dpp = G[p * OW + p];
        if (!(dpp > OPIV * dg[p])) flag[0] = 1.0;
        G[p * OW + p] = sqrt(fmax(dpp, 1e-300));
      }
      __syncthreads();
      if (flag[0] != 0.0) break;
      const double rpp = G[p * OW + p];
      for (int j = p + 1 + t; j < w; j += OT) G[p * OW + j] /= rpp;
      __syncthreads();
      const int m = w - p - 1;
      for (int e = t; e < m * m; e += OT) {
        const int i = p + 1 + e / m, j = p + 1 + e % m;
        if (j >= i) G[i * OW + j] -= G[p * OW + i] * G[p * OW + j];
      }
      __syncthreads();
    }
    if (flag[0] != 0.0) {
      orth_panel_cgs(Y, n, ld, c0, w, X, red);
      return;
    }
    // X = R^-1, one column per thread
    if (t < w) {
      const int j = t;
      for (int i = j; i >= 0; --i) {
        double s = (i == j) ? 1.0 : 0.0;
        for (int l = i + 1; l <= j; ++l) s -= G[i * OW + l] * X[l * OW + j];
        X[i * OW + j] = s / G[i * OW + i];
      }
      for (int i = j + 1; i < OW; ++i) X[i * OW + j] = 0.0;
    }
    for (int idx = t; idx < OW * (OW - w); idx += OT) {
      const int i = idx / (OW - w), j = w + idx % (OW - w);
      X[i * OW + j] = 0.0;
    }
    __syncthreads();
    // panel <- panel R^-1, OR rows at a time
    for (int k0 = 0; k0 < n; k0 += OR) {
      const int rows = min(OR, n - k0);
      for (int idx = t; idx < OR * OW; idx += OT) {
        const int r = idx >> 6, l = idx & 63;
        G[idx] = (r < rows && l < w) ? Y[(int64_t)(k0 + r) * ld + c0 + l] : 0.0;
      }
      __syncthreads();
      if (2 * rq < rows && 4 * cb < w) {
        double o0[4] = {0.0, 0.0, 0.0, 0.0}, o1[4] = {0.0, 0.0, 0.0, 0.0};
        const int lmax = min(w, 4 * cb + 4);
        for (int l = 0; l < lmax; ++l) {
          const double s0 = G[(2 * rq) * OW + l], s1 = G[(2 * rq + 1) * OW + l];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const double x = X[l * OW + 4 * cb + v];
            o0[v] += s0 * x;
            o1[v] += s1 * x;
          }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int j = 4 * cb + v;
          if (j < w) {
            Y[(int64_t)(k0 + 2 * rq) * ld + c0 + j] = o0[v];
            if (2 * rq + 1 < rows) Y[(int64_t)(k0 + 2 * rq + 1) * ld + c0 + j] = o1[v];
          }
        }
      }
      __syncthreads();
    }
  }
}

// Split product C_c = Q^T P over row chunk c = blockIdx.y (OCH rows), Q the
// q columns at qa, P the w columns at p0: part[(c * q + i) * OW + j].
__global__ __launch_bounds__(512) void k_g_orth_part(GridMol g, int qa, int q, int p0, int w,
                                                     double* part)
{
  __shared__ double tile_[2048];
  const int ld = g.ld;
  const double* Y = g.Y;
  const int c = blockIdx.y;
  const int k0 = c * OCH, k1 = min(g.n, k0 + OCH);
  big::gemm_tile<512, false, true>(
      q, w, blockIdx.x * 64, 0, k0, k1,
      [=] __device__(int i, int k) { return Y[(int64_t)k * ld + qa + i]; },
      [=] __device__(int k, int j) { return Y[(int64_t)k * ld + p0 + j]; },
      [=] __device__(int i, int j, double v) { part[((int64_t)c * q + i) * OW + j] = v; },
      TA_LDS(tile_));
}

// C[i * OW + j] = sum over the chunks of the split product
__global__ void k_g_orth_sum(int q, int w, int nch, const double* part, double* C)
{
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < q * OW; idx += gridDim.x * blockDim.x) {
    const int i = idx / OW, j = idx % OW;
    double s = 0.0;
    if (j < w)
      for (int c = 0; c < nch; ++c) s += part[((int64_t)c * q + i) * OW + j];
    C[idx] = s;
  }
}

// -------------------------------------------------------- back-transform
// T of reflector block k0 .. k0 + nb - 1 (T^-1 = diag(1/tau) + strict_upper(V^T V))
// V^T V of reflector block k0 .. k0 + nb - 1 over row chunk blockIdx.x
// (OCH rows): part[c * NB * NB + i * NB + j]
__global__ __launch_bounds__(512) void k_g_bt_part(GridMol g, int k0, int nb, double* part)
{
  __shared__ double tile_[2048];
  const int n = g.n, ld = g.ld, r0 = k0 + 1, m = n - r0;
  const double* A = g.A;
  const double* tau = g.tau;
  const int c = blockIdx.x;
  const int ka = c * OCH, kb = min(m, ka + OCH);
  auto vel = [=] __device__(int rr, int j) -> double {
    return (j < nb && rr >= j && tau[k0 + j] != 0.0) ? A[(int64_t)(r0 + rr) * ld + k0 + j] : 0.0;
  };
  big::gemm_tile<512, false, true>(
      TA_BIG_NB, TA_BIG_NB, 0, 0, ka, kb, [=] __device__(int i, int k) { return vel(k, i); },
      [=] __device__(int k, int j) { return vel(k, j); },
      [=] __device__(int i, int j, double v) { part[(int64_t)c * TA_BIG_NB * TA_BIG_NB + i * TA_BIG_NB + j] = v; },
      TA_LDS(tile_));
}

// T of reflector block k0 .. k0 + nb - 1 from the chunk sums of V^T V
// (T^-1 = diag(1/tau) + strict_upper(V^T V))
__global__ __launch_bounds__(512) void k_g_bt_t(GridMol g, int k0, int nb, int nch,
                                                const double* part)
{
  __shared__ double sinv_[TA_BIG_NB * TA_BIG_NB];
  lds_dbl* sinv = TA_LDS(sinv_);
  const double* tau = g.tau;
  for (int idx = threadIdx.x; idx < TA_BIG_NB * TA_BIG_NB; idx += GT) {
    const int i = idx / TA_BIG_NB, j = idx % TA_BIG_NB;
    double s = 0.0;
    if (i < j)
      for (int c = 0; c < nch; ++c) s += part[(int64_t)c * TA_BIG_NB * TA_BIG_NB + idx];
    if (i == j) s = (i < nb && tau[k0 + i] != 0.0) ? 1.0 / tau[k0 + i] : 1.0;
    sinv[idx] = s;
  }
  __syncthreads();
  if (threadIdx.x < TA_BIG_NB) {
    const int j = threadIdx.x;
    for (int i = TA_BIG_NB - 1; i >= 0; --i) {
      double t = 0.0;
      if (i == j) {
        t = 1.0 / sinv[i * TA_BIG_NB + i];
      } else if (i < j) {
        double acc = 0.0;
        for (int k = i + 1; k <= j; ++k) acc += sinv[i * TA_BIG_NB + k] * g.tm[k * TA_BIG_NB + j];
        t = -acc / sinv[i * TA_BIG_NB + i];
      }
      g.tm[i * TA_BIG_NB + j] = t;
    }
  }
}

// ---------------------------------------------------------- occupation
__global__ __launch_bounds__(64) void k_g_fermi(GridMol g)
{
  const int n = g.n;
  double ts = fermi_wave(g.sc[GS_NALP], g.kt, n, g.lam, g.fs);
  for (int r = threadIdx.x; r < n; r += 64) g.focc[r] = g.fs[r];
  ts += fermi_wave(g.sc[GS_NBET], g.kt, n, g.lam, g.fs);
  for (int r = threadIdx.x; r < n; r += 64) g.focc[r] += g.fs[r];
  if (threadIdx.x == 0) {
    int last = 0;
    for (int r = 0; r < n; ++r)
      if (g.focc[r] > TA_OCC_MIN) last = r + 1;
    g.ic[GI_NOCC] = last;
    g.sc[GS_TS] = ts;
  }
}

// ---------------------------------------------------------- Mulliken
// one wave per AO row: part[c][i], c = 0 PS, 1..3 PD, 4..9 PQ, 10 PH
__global__ __launch_bounds__(64) void k_g_mull_rows(GridMol g)
{
  const int i = blockIdx.x, n = g.n, lane = threadIdx.x;
  const int64_t n2 = (int64_t)n * n, row = (int64_t)i * n;
  double acc[11];
#pragma unroll
  for (int c = 0; c < 11; ++c) acc[c] = 0.0;
  for (int j = lane; j < n; j += 64) {
    const double p = g.Y[(int64_t)i * g.ld + j];
    const int64_t e = row + j;
    acc[0] += p * g.S[e];
    acc[1] += p * g.D[e];
    acc[2] += p * g.D[n2 + e];
    acc[3] += p * g.D[2 * n2 + e];
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[4 + c] += p * g.Q[c * n2 + e];
    acc[10] += p * g.H[e];
  }
#pragma unroll
  for (int c = 0; c < 11; ++c) acc[c] = wave_sum(acc[c]);
  if (lane == 0)
#pragma unroll
    for (int c = 0; c < 11; ++c) g.part11[(int64_t)c * n + i] = acc[c];
}

__global__ __launch_bounds__(512) void k_g_mull_fin(GridMol g)
{
  __shared__ double red_[16];
  lds_dbl* red = TA_LDS(red_);
  const DevParams& P = *g.par;
  const int n = g.n, nsh = g.nsh, nat = g.nat, tid = threadIdx.x;
  const double* part = g.part11;
  double err = 0.0;
  for (int s = tid; s < nsh; s += GT) {
    const int i0 = g.sh_iao[s];
    const int l = P.el[g.el[g.sh_at[s]]].ang[g.sh_ish[s]];
    double q = g.n0sh[s];
    for (int i = i0; i < i0 + 2 * l + 1; ++i) q -= part[i];
    g.qsh[s] = q;
    const double d = q - g.qin[s];
    g.dq[s] = d;
    err += d * d;
  }
  for (int a = tid; a < nat; a += GT) {
    const DevElem& e = P.el[g.el[a]];
    const int i0 = g.at_iao[a];
    int nao_a = 0;
    for (int s = 0; s < e.nsh; ++s) nao_a += 2 * e.ang[s] + 1;
    double dp[3] = {0.0, 0.0, 0.0}, qp[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = i0; i < i0 + nao_a; ++i) {
      for (int k = 0; k < 3; ++k) dp[k] -= part[(int64_t)(1 + k) * n + i];
      for (int k = 0; k < 6; ++k) qp[k] -= part[(int64_t)(4 + k) * n + i];
    }
    for (int k = 0; k < 3; ++k) {
      g.dpat[3 * a + k] = dp[k];
      const double d = dp[k] - g.qin[nsh + 3 * a + k];
      g.dq[nsh + 3 * a + k] = d;
      err += d * d;
    }
    for (int k = 0; k < 6; ++k) {
      g.qpat[6 * a + k] = qp[k];
      const double d = qp[k] - g.qin[nsh + 3 * nat + 6 * a + k];
      g.dq[nsh + 3 * nat + 6 * a + k] = d;
      err += d * d;
    }
  }
  double eband = 0.0;
  for (int i = tid; i < n; i += GT) eband += part[(int64_t)10 * n + i];
  err = block_sum(err, red);
  eband = block_sum(eband, red);
  __syncthreads();
  for (int a = tid; a < nat; a += GT) {
    const int s0a = g.at_ish[a];
    const int ns = P.el[g.el[a]].nsh;
    double q = 0.0;
    for (int s = 0; s < ns; ++s) q += g.qsh[s0a + s];
    g.qat[a] = q;
  }
  if (tid == 0) {
    g.sc[GS_ERR] = err;
    g.sc[GS_EBAND] = eband;
  }
}

// ---------------------------------------------------------- energy
// blocks [0, nsh): ES2 + ES3 of one shell; [nsh, nsh + nat): AES2 of one atom
__global__ __launch_bounds__(64) void k_g_energy(GridMol g)
{
  const DevParams& P = *g.par;
  const int nat = g.nat, nsh = g.nsh, lane = threadIdx.x;
  const int64_t p2 = (int64_t)nat * nat;
  double ec = 0.0;
  if ((int)blockIdx.x < nsh) {
    const int sh = blockIdx.x;
    const double* gr = g.gam + (int64_t)sh * nsh;
    const double qs = g.qsh[sh];
    for (int j = lane; j < nsh; j += 64) ec += 0.5 * qs * gr[j] * g.qsh[j];
    ec = wave_sum(ec);
    if (lane == 0)
      g.epart[sh] = ec + qs * qs * qs * P.el[g.el[g.sh_at[sh]]].hubbard_derivs[g.sh_ish[sh]] / 3.0;
    return;
  }
  const int a = blockIdx.x - nsh;
  const int64_t row = (int64_t)a * nat;
  const double dpa[3] = {g.dpat[3 * a], g.dpat[3 * a + 1], g.dpat[3 * a + 2]};
  double qpa[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) qpa[k] = g.qpat[6 * a + k];
  for (int i = lane; i < nat; i += 64) {
    const int64_t ai = row + i;
    const double qi = g.qat[i];
    const double dpi[3] = {g.dpat[3 * i], g.dpat[3 * i + 1], g.dpat[3 * i + 2]};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double t = -g.sd[k * p2 + ai] * qi;
#pragma unroll
      for (int l = 0; l < 3; ++l) t += 0.5 * g.dd[(3 * l + k) * p2 + ai] * dpi[l];
      ec += dpa[k] * t;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) ec += qpa[k] * g.sq[k * p2 + ai] * qi;
  }
  ec = wave_sum(ec);
  // D4 two-body energy of atom a: 1/2 sum_i sum_rs g_ar d_ai C6ref g_is
  double ed = 0.0;
  if (P.d4.enabled) {
    const int ea = g.el[a];
    double vv[TA_D4_MAXREF] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = lane; i < nat; i += 64) {
      const double dai = g.d4_disp[row + i];
      if (dai != 0.0) d4_scf_accumulate(P.d4, ea, g.el[i], 0.5 * dai, g.d4_gw + TA_D4_MAXREF * i, vv);
    }
#pragma unroll
    for (int r = 0; r < TA_D4_MAXREF; ++r) {
      vv[r] = wave_sum(vv[r]);
      if (r < d4_nw(P.d4, ea)) ed += vv[r] * g.d4_gw[TA_D4_MAXREF * a + r];
    }
  }
  if (lane == 0) {
    const DevElem& e = P.el[g.el[a]];
    g.ed4[a] = ed;
    ec += ed;
    ec += e.dkernel * (dpa[0] * dpa[0] + dpa[1] * dpa[1] + dpa[2] * dpa[2]);
    ec += e.qkernel * (qpa[0] * qpa[0] + 2.0 * qpa[1] * qpa[1] + qpa[2] * qpa[2] +
                       2.0 * qpa[3] * qpa[3] + 2.0 * qpa[4] * qpa[4] + qpa[5] * qpa[5]);
    g.epart[nsh + a] = ec;
  }
}

__global__ __launch_bounds__(512) void k_g_energy_fin(GridMol g, int iter)
{
  __shared__ double red_[16];
  lds_dbl* red = TA_LDS(red_);
  double ec = 0.0, ed = 0.0;
  for (int k = threadIdx.x; k < g.nsh + g.nat; k += GT) ec += g.epart[k];
  for (int a = threadIdx.x; a < g.nat; a += GT) ed += g.ed4[a];
  ec = block_sum(ec, red);
  ed = block_sum(ed, red);
  if (threadIdx.x == 0) {
    g.sc[GS_ED2] = ed;
    const double elast = g.sc[GS_ESUM];
    const double esum = g.sc[GS_EBAND] + g.sc[GS_TS] + ec;
    const double perr = sqrt(g.sc[GS_ERR] / g.ndim);
    const int conv = (fabs(esum - elast) < g.econv) && (perr < g.pconv);
    g.sc[GS_ESUM] = esum;
    g.ic[GI_CONVERGED] = conv;
    if (!isfinite(esum)) g.ic[GI_STATUS] |= TA_STATUS_NAN;
    (void)iter;
  }
}

// ---------------------------------------------------------- mixing
// modified Broyden step (same scheme as the single-workgroup kernel), then
// unpack into the wavefunction
__global__ __launch_bounds__(512) void k_g_broyden(GridMol g, int mixiter)
{
  __shared__ double red_[16];
  __shared__ double scal_[4];
  lds_dbl* red = TA_LDS(red_);
  const int tid = threadIdx.x, lane = tid & 63, wid = wave_id();
  constexpr int NW = GT / 64;
  const int ndim = g.ndim, mem = g.mem, nsh = g.nsh, nat = g.nat;
  const double damp = g.damping;
  double* qin = g.qin;
  double* dq = g.dq;
  double* dqlast = g.dqlast;
  double* qlast = g.qlast;
  if (mixiter == 1) {
    for (int i = tid; i < ndim; i += GT) {
      dqlast[i] = dq[i];
      qlast[i] = qin[i];
      qin[i] += damp * dq[i];
    }
    __syncthreads();
  } else {
    const int it1 = (mixiter - 2) % mem;
    const int ldim = (mixiter - 1) < mem ? (mixiter - 1) : mem;
    double* dfc = g.dfm + (int64_t)it1 * ndim;
    double* uc = g.um + (int64_t)it1 * ndim;
    double a1 = 0.0, a2 = 0.0;
    for (int i = tid; i < ndim; i += GT) {
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
    if (tid == 0) g.omega[it1] = om;
    for (int i = tid; i < ndim; i += GT) {
      const double d = dfc[i] * inv_;
      dfc[i] = d;
      uc[i] = damp * d + inv_ * (qin[i] - qlast[i]);
    }
    __syncthreads();
    for (int j = wid; j < ldim; j += NW) {
      const double* dfj = g.dfm + (int64_t)j * ndim;
      double s1 = 0.0, s2 = 0.0;
      for (int i = lane; i < ndim; i += 64) {
        s1 += dfj[i] * dfc[i];
        s2 += dfj[i] * dq[i];
      }
      s1 = wave_sum(s1);
      s2 = wave_sum(s2);
      if (lane == 0) {
        g.amix[j + (int64_t)mem * it1] = s1;
        g.amix[it1 + (int64_t)mem * j] = s1;
        g.cvec[j] = (j == it1 ? om : g.omega[j]) * s2;
      }
    }
    __syncthreads();
    const int lda = ldim + 1;
    double* beta = g.beta;
    for (int idx = tid; idx < ldim * ldim; idx += GT) {
      const int i = idx / ldim, j = idx - (idx / ldim) * ldim;
      double val = g.omega[i] * g.omega[j] * g.amix[i + (int64_t)mem * j];
      if (i == j) val += 0.01 * 0.01;
      beta[i * lda + j] = val;
    }
    for (int i = tid; i < ldim; i += GT) beta[i * lda + ldim] = g.cvec[i];
    __syncthreads();
    bool fail = false;
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
          scal_[1] = (double)bi;
          scal_[2] = best;
        }
      }
      __syncthreads();
      const int piv = (int)scal_[1];
      if (!(scal_[2] > 0.0)) {
        fail = true;
        break;
      }
      if (piv != col)
        for (int k = tid; k <= ldim; k += GT) {
          const double t = beta[col * lda + k];
          beta[col * lda + k] = beta[piv * lda + k];
          beta[piv * lda + k] = t;
        }
      __syncthreads();
      const double pv = beta[col * lda + col];
      const int rows = ldim - col - 1, cols = ldim - col;
      for (int idx = tid; idx < rows * cols; idx += GT) {
        const int r = col + 1 + idx / cols, k = col + 1 + idx % cols;
        beta[r * lda + k] -= beta[r * lda + col] / pv * beta[col * lda + k];
      }
      __syncthreads();
      for (int r = col + 1 + tid; r < ldim; r += GT) beta[r * lda + col] = 0.0;
      __syncthreads();
    }
    if (fail) {
      if (tid == 0) g.ic[GI_STATUS] |= TA_STATUS_MIXER;
      return;
    }
    if (wid == 0) {
      for (int r = ldim - 1; r >= 0; --r) {
        double s = 0.0;
        for (int k = r + 1 + lane; k < ldim; k += 64) s += beta[r * lda + k] * g.cvec[k];
        s = wave_sum(s);
        if (lane == 0) g.cvec[r] = (beta[r * lda + ldim] - s) / beta[r * lda + r];
      }
    }
    __syncthreads();
    for (int i = tid; i < ndim; i += GT) {
      double q = qin[i] + damp * dq[i];
      for (int j = 0; j < ldim; ++j) q -= g.omega[j] * g.cvec[j] * g.um[(int64_t)j * ndim + i];
      dqlast[i] = dq[i];
      qlast[i] = qin[i];
      qin[i] = q;
    }
    __syncthreads();
  }
  for (int i = tid; i < nsh; i += GT) g.qsh[i] = qin[i];
  for (int i = tid; i < 3 * nat; i += GT) g.dpat[i] = qin[nsh + i];
  for (int i = tid; i < 6 * nat; i += GT) g.qpat[i] = qin[nsh + 3 * nat + i];
  __syncthreads();
  const DevParams& P = *g.par;
  for (int a = tid; a < nat; a += GT) {
    const int s0a = g.at_ish[a];
    const int ns = P.el[g.el[a]].nsh;
    double q = 0.0;
    for (int s = 0; s < ns; ++s) q += g.qsh[s0a + s];
    g.qat[a] = q;
  }
}

// q_in = current charges (set_mixer) before the first mixing step
__global__ void k_g_set_mixer(GridMol g)
{
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < g.ndim; i += gridDim.x * blockDim.x)
    g.qin[i] = 0.0;
}

// X = U s^-1/2 U^T needs s > 0: flag a singular overlap
__global__ void k_g_check_overlap(GridMol g)
{
  if (threadIdx.x == 0 && !(g.lam[0] > 1.0e-10)) g.ic[GI_STATUS] |= TA_STATUS_CHOLESKY;
}

// energy, status and iteration count of the molecule
__global__ void k_g_final(GridMol g, int iter, int status, double* energy, double* eelec,
                          double* erep, int* iterations, int* st)
{
  if (threadIdx.x == 0) {
    const double esum = g.sc[GS_ESUM], er = g.sc[GS_EREP];
    *energy = er + g.sc[GS_E3] + esum;
    *g.edisp2 = g.sc[GS_ED2];
    *g.edisp3 = g.sc[GS_E3];
    *eelec = esum;
    *erep = er;
    *iterations = iter;
    *st = status;
  }
}

// results: sorted levels and occupations
__global__ void k_g_results(GridMol g, double* emo, double* focc)
{
  const int n = g.n;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    emo[r] = g.lam[r];
    focc[r] = g.focc[r];
  }
}

// n x n matrix out(i, j) = M[j * ld + i] (column-major, tblite's layout) from
// a row-major ld-strided matrix; rows past nrow read as 0
__global__ void k_g_unpad(const double* M, int n, int ld, int nrow, double* out)
{
  const int64_t tot = (int64_t)n * n;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < tot;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(idx / n), i = (int)(idx - (idx / n) * n);
    out[idx] = j < nrow ? M[(int64_t)j * ld + i] : 0.0;
  }
}

int grid_blocks(int64_t work, int per = 256) { return (int)std::min<int64_t>((work + per - 1) / per, 8192); }

// Eigenvalues (g.lam, ascending) of the symmetric matrix in g.A (lower
// triangle used): reduced to tridiagonal form in place (d, e, tau and the
// reflectors in A), then bisection.
hipError_t grid_eigenvalues(GridMol& g, hipStream_t s, bool warm)
{
  const int n = g.n, ld = g.ld;
  const int nr = n - 2;
  constexpr int NP = TA_BIG_NP;
  for (int k0 = 0; k0 < nr; k0 += NP) {
    const int np = (k0 + NP <= nr) ? NP : nr - k0;
    for (int i = 0; i < np; ++i) {
      const int j = k0 + i;
      hipLaunchKernelGGL(k_g_tri_colupd, dim3((n - j + GC_T - 1) / GC_T), dim3(GC_T), 0, s, g, j, i);
      hipLaunchKernelGGL(k_g_tri_refl, dim3(1), dim3(GT), 0, s, g, j, i);
      const int m = n - (j + 1), nb = (m + 63) >> 6;
      hipLaunchKernelGGL(k_g_tri_symv, dim3(nb * (nb + 1) / 2), dim3(256), 0, s, g, j + 1, i);
      hipLaunchKernelGGL(k_g_tri_rows, dim3(nb), dim3(256), 0, s, g, j, i);
    }
    // w of the panel's last column (the next column's update forms the others)
    hipLaunchKernelGGL(k_g_tri_fin, dim3(1), dim3(GT), 0, s, g, k0 + np - 1, np - 1);
    const int r0 = k0 + np;
    if (r0 >= nr) {
      hipLaunchKernelGGL(k_g_tri_last, dim3(1), dim3(64), 0, s, g, np);
      break;
    }
    const int m = n - r0;
    double* A = g.A;
    const double* vt = g.vt;
    const double* wt = g.wt;
    g_gemm<false, true>(
        s, m, m,
        [=] __device__(int ii, int k) {
          return k < NP ? vt[(int64_t)k * ld + r0 + ii] : wt[(int64_t)(k - NP) * ld + r0 + ii];
        },
        [=] __device__(int k, int jj) {
          return k < NP ? wt[(int64_t)k * ld + r0 + jj] : vt[(int64_t)(k - NP) * ld + r0 + jj];
        },
        AllK{2 * NP}, UpperSkip{},
        [=] __device__(int ii, int jj, double val) {
          // lower tiles, mirrored: the trailing square stays symmetric, so a
          // column of it can be read as a (contiguous) row
          if (ii > jj) {
            A[(int64_t)(r0 + ii) * ld + r0 + jj] -= val;
            A[(int64_t)(r0 + jj) * ld + r0 + ii] -= val;
          } else if (ii == jj) {
            A[(int64_t)(r0 + ii) * ld + r0 + jj] -= val;
          }
        });
  }
  hipLaunchKernelGGL(k_g_gersh, dim3(1), dim3(GT), 0, s, g);
  hipLaunchKernelGGL(k_g_bisect, dim3(grid_blocks(n)), dim3(256), 0, s, g, warm ? 1 : 0);
  hipLaunchKernelGGL(k_g_warm, dim3(1), dim3(256), 0, s, g, warm ? 1 : 0);
  return hipGetLastError();
}

// Eigenvectors of the first nvec levels (g.Y columns 0..nvec-1) from the
// tridiagonal form left by grid_eigenvalues.
hipError_t grid_eigenvectors(GridMol& g, hipStream_t s, int nvec)
{
  const int n = g.n, ld = g.ld;
  const int nr = n - 2;
  if (nvec <= 0) return hipGetLastError();
  hipLaunchKernelGGL(k_g_invit, dim3(grid_blocks(nvec)), dim3(256), 0, s, g, nvec);
  hipLaunchKernelGGL(k_g_runs, dim3(1), dim3(64), 0, s, g, nvec);
  {
    // runs of close levels (the host needs their lengths for the panel loop)
    std::vector<int> hr(GI_COUNT + n + 2);
    hipError_t e = hipMemcpyAsync(hr.data(), g.ic, sizeof(int) * (GI_COUNT + n),
                                  hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    const int nruns = hr[GI_NRUNS];
    const int* runs = hr.data() + GI_COUNT;
    const size_t plds = sizeof(double) * orth_panel_lds_doubles();
    if (nruns > 0) hipLaunchKernelGGL(k_g_orth_panel, dim3(nruns), dim3(OT), plds, s, g, -1, 0);
    double* part = g.scr;
    double* C = g.scr + 2 * (int64_t)n * n;
    const int nch = (n + OCH - 1) / OCH;
    for (int c = 0; c < nruns; ++c) {
      const int a = runs[2 * c], b = runs[2 * c + 1];
      for (int p0 = a + OW; p0 < b; p0 += OW) {
        const int w = std::min(OW, b - p0), q = p0 - a;
        for (int pass = 0; pass < 2; ++pass) {
          hipLaunchKernelGGL(k_g_orth_part, dim3((q + 63) / 64, nch), dim3(512), 0, s, g, a, q, p0,
                             w, part);
          hipLaunchKernelGGL(k_g_orth_sum, dim3(grid_blocks((int64_t)q * OW)), dim3(256), 0, s, q,
                             w, nch, (const double*)part, C);
          double* Y = g.Y;
          const double* Cc = C;
          g_gemm<true, true>(
              s, n, w, [=] __device__(int i, int k) { return Y[(int64_t)i * ld + a + k]; },
              [=] __device__(int k, int j) { return Cc[k * OW + j]; }, AllK{q}, NoSkip{},
              [=] __device__(int i, int j, double val) { Y[(int64_t)i * ld + p0 + j] -= val; });
        }
        hipLaunchKernelGGL(k_g_orth_panel, dim3(1), dim3(OT), plds, s, g, p0, w);
      }
    }
  }
  // back-transformation, last block of reflectors first
  constexpr int NB = TA_BIG_NB;
  const int nblk = (nr + NB - 1) / NB;
  for (int bi = nblk - 1; bi >= 0; --bi) {
    const int k0 = bi * NB;
    const int nb = (k0 + NB <= nr) ? NB : nr - k0;
    const int r0 = k0 + 1;
    const int m = n - r0;
    {
      const int nch_bt = (m + OCH - 1) / OCH;
      hipLaunchKernelGGL(k_g_bt_part, dim3(nch_bt), dim3(512), 0, s, g, k0, nb, g.scr);
      hipLaunchKernelGGL(k_g_bt_t, dim3(1), dim3(GT), 0, s, g, k0, nb, nch_bt,
                         (const double*)g.scr);
    }
    const double* A = g.A;
    const double* tau = g.tau;
    double* Y = g.Y;
    double* W = g.wbt;
    const double* tm = g.tm;
    auto vel = [=] __device__(int rr, int j) -> double {
      return (j < nb && rr >= j && tau[k0 + j] != 0.0) ? A[(int64_t)(r0 + rr) * ld + k0 + j] : 0.0;
    };
    g_gemm<false, true>(
        s, NB, nvec, [=] __device__(int i, int k) { return vel(k, i); },
        [=] __device__(int k, int j) { return Y[(int64_t)(r0 + k) * ld + j]; }, AllK{m}, NoSkip{},
        [=] __device__(int i, int j, double val) { W[(int64_t)i * ld + j] = val; });
    g_gemm<true, true>(
        s, NB, nvec, [=] __device__(int i, int k) { return tm[i * NB + k]; },
        [=] __device__(int k, int j) { return W[(int64_t)k * ld + j]; }, AllK{NB}, NoSkip{},
        [=] __device__(int i, int j, double val) { W[(int64_t)i * ld + j] = val; });
    g_gemm<true, true>(
        s, m, nvec, [=] __device__(int i, int k) { return vel(i, k); },
        [=] __device__(int k, int j) { return W[(int64_t)k * ld + j]; }, AllK{NB}, NoSkip{},
        [=] __device__(int i, int j, double val) { Y[(int64_t)(r0 + i) * ld + j] -= val; });
  }
  return hipGetLastError();
}

}  // namespace

// Full SCF of one molecule on the grid path; writes energy/status/iterations
// of molecule m into the batch arrays.
namespace {
// The eigenvalue phase of an SCF iteration (tridiagonalisation, about four
// kernels per column, and bisection) is the same kernel sequence with the
// same arguments in every iteration after the first: captured once into a
// HIP graph and replayed.
struct EigGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~EigGraph()
  {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
  hipError_t run(GridMol& g, hipStream_t s)
  {
    if (!exec) {
      hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) return e;
      const hipError_t ek = grid_eigenvalues(g, s, true);
      e = hipStreamEndCapture(s, &graph);
      if (ek != hipSuccess) return ek;
      if (e != hipSuccess) return e;
      e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (e != hipSuccess) return e;
    }
    return hipGraphLaunch(exec, s);
  }
};
}  // namespace

int grid_scf(GridMol g, hipStream_t s, int max_iter, double* d_energy, double* d_eelec,
             double* d_erep, int* d_iter, int* d_status, double* emo, double* focc, double* coeff,
             double* density, GridHostStats* hs, int verbose, int use_graph)
{
  const int n = g.n, ld = g.ld;
  hipError_t err;
  EigGraph eg;
  if (hipFuncSetAttribute((const void*)k_g_orth_panel, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)(sizeof(double) * orth_panel_lds_doubles())) != hipSuccess)
    return -2;
  hipLaunchKernelGGL(k_g_init, dim3(1), dim3(256), 0, s, g);
  // X = S^-1/2 from the eigenpairs of S
  hipLaunchKernelGGL(k_g_copy_s, dim3(grid_blocks((int64_t)n * n)), dim3(256), 0, s, g);
  if ((err = grid_eigenvalues(g, s, false)) != hipSuccess) return -2;
  if ((err = grid_eigenvectors(g, s, n)) != hipSuccess) return -2;
  hipLaunchKernelGGL(k_g_check_overlap, dim3(1), dim3(64), 0, s, g);
  {
    const double* Y = g.Y;
    const double* lam = g.lam;
    double* X = g.X;
    g_gemm<true, false>(
        s, n, n,
        [=] __device__(int i, int k) { return Y[(int64_t)i * ld + k] * rsqrt(sqrt(fmax(lam[k], 1e-300))); },
        [=] __device__(int k, int j) { return Y[(int64_t)j * ld + k] * rsqrt(sqrt(fmax(lam[k], 1e-300))); },
        AllK{n}, NoSkip{}, [=] __device__(int i, int j, double val) { X[(int64_t)i * ld + j] = val; });
  }
  int iter = 0, mixiter = 0, converged = 0, status = 0;
  int hbuf[4];
  const auto t_start = std::chrono::steady_clock::now();
  if (hipMemcpyAsync(hbuf, g.ic, sizeof(int) * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return -2;
  status = hbuf[GI_STATUS];
  while (!converged && !status && iter < max_iter) {
    if (iter > 0) {
      ++mixiter;
      hipLaunchKernelGGL(k_g_broyden, dim3(1), dim3(GT), 0, s, g, mixiter);
    }
    ++iter;
    const bool d4on = g.d4_on != 0;
    if (d4on) hipLaunchKernelGGL(k_g_d4_weights, dim3(grid_blocks(g.nat)), dim3(256), 0, s, g);
    hipLaunchKernelGGL(k_g_potential, dim3(g.nsh + g.nat), dim3(256), 0, s, g);
    if (d4on) hipLaunchKernelGGL(k_g_d4_potential, dim3(g.nat), dim3(64), 0, s, g);
    if (iter == 1) hipLaunchKernelGGL(k_g_set_mixer, dim3(grid_blocks(g.ndim)), dim3(256), 0, s, g);
    hipLaunchKernelGGL(k_g_fock, dim3(grid_blocks((int64_t)n * n)), dim3(256), 0, s, g);
    hipLaunchKernelGGL(k_g_symmetrize, dim3(grid_blocks((int64_t)n * n)), dim3(256), 0, s, g);
    // A = X F X: U = X F into Y, A = U X (lower tiles)
    {
      const double* X = g.X;
      const double* F = g.A;
      double* U = g.Y;
      g_gemm<true, true>(
          s, n, n, [=] __device__(int i, int k) { return X[(int64_t)i * ld + k]; },
          [=] __device__(int k, int j) { return F[(int64_t)k * ld + j]; }, AllK{n}, NoSkip{},
          [=] __device__(int i, int j, double val) { U[(int64_t)i * ld + j] = val; });
      double* A = g.A;
      const double* Uc = g.Y;
      g_gemm<true, true>(
          s, n, n, [=] __device__(int i, int k) { return Uc[(int64_t)i * ld + k]; },
          [=] __device__(int k, int j) { return X[(int64_t)k * ld + j]; }, AllK{n}, UpperSkip{},
          [=] __device__(int i, int j, double val) {
            // lower tiles, mirrored (the tridiagonalisation reads rows)
            if (i >= j) {
              A[(int64_t)i * ld + j] = val;
              A[(int64_t)j * ld + i] = val;
            }
          });
    }
    if (use_graph && iter > 1) {
      if ((err = eg.run(g, s)) != hipSuccess) return -2;
    } else if ((err = grid_eigenvalues(g, s, iter > 1)) != hipSuccess) {
      return -2;
    }
    // occupations from the levels; only occupied levels need eigenvectors
    hipLaunchKernelGGL(k_g_fermi, dim3(1), dim3(64), 0, s, g);
    if (hipMemcpyAsync(hbuf, g.ic, sizeof(int) * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return -2;
    const int nocc = hbuf[GI_NOCC];
    if ((err = grid_eigenvectors(g, s, nocc)) != hipSuccess) return -2;
    // MO rows CT = Z^T X of the occupied levels
    {
      const double* Z = g.Y;
      const double* X = g.X;
      double* CT = g.CT;
      g_gemm<false, true>(
          s, nocc, n, [=] __device__(int i, int k) { return Z[(int64_t)k * ld + i]; },
          [=] __device__(int k, int j) { return X[(int64_t)k * ld + j]; }, AllK{n}, NoSkip{},
          [=] __device__(int i, int j, double val) { CT[(int64_t)i * ld + j] = val; });
    }
    // density P = sum_k f_k c_k c_k^T into Y (lower tiles mirrored)
    {
      const double* CT = g.CT;
      const double* f = g.focc;
      double* P = g.Y;
      g_gemm<false, true>(
          s, n, n, [=] __device__(int i, int k) { return CT[(int64_t)k * ld + i] * f[k]; },
          [=] __device__(int k, int j) { return CT[(int64_t)k * ld + j]; }, AllK{nocc}, UpperSkip{},
          [=] __device__(int i, int j, double val) {
            if (i >= j) {
              P[(int64_t)i * ld + j] = val;
              P[(int64_t)j * ld + i] = val;
            }
          });
    }
    hipLaunchKernelGGL(k_g_mull_rows, dim3(n), dim3(64), 0, s, g);
    hipLaunchKernelGGL(k_g_mull_fin, dim3(1), dim3(GT), 0, s, g);
    if (d4on) hipLaunchKernelGGL(k_g_d4_weights, dim3(grid_blocks(g.nat)), dim3(256), 0, s, g);
    hipLaunchKernelGGL(k_g_energy, dim3(g.nsh + g.nat), dim3(64), 0, s, g);
    hipLaunchKernelGGL(k_g_energy_fin, dim3(1), dim3(GT), 0, s, g, iter);
    if ((err = hipGetLastError()) != hipSuccess) return -2;
    if (hipMemcpyAsync(hbuf, g.ic, sizeof(int) * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return -2;
    converged = hbuf[GI_CONVERGED];
    status = hbuf[GI_STATUS];
    if (verbose) {
      const double ms = std::chrono::duration<double, std::milli>(
                            std::chrono::steady_clock::now() - t_start).count();
      std::fprintf(stderr, "[tblite_amd] grid SCF n=%d iteration %d: converged %d status %d (%.0f ms)\n",
                   n, iter, converged, status, ms);
    }
  }
  if (!converged && status == 0) status |= TA_STATUS_NOT_CONVERGED;
  hipLaunchKernelGGL(k_g_final, dim3(1), dim3(64), 0, s, g, iter, status, d_energy, d_eelec,
                     d_erep, d_iter, d_status);
  hipLaunchKernelGGL(k_g_results, dim3(grid_blocks(n)), dim3(256), 0, s, g, emo, focc);
  // wavefunction arrays: the density of the last iteration (in Y), then the
  // coefficients of all n levels from the last tridiagonal form (the SCF
  // formed the occupied ones only), MO rows CT = Z^T X
  const bool wfn_ok = iter > 0 && !(status & (TA_STATUS_EIGEN | TA_STATUS_NAN | TA_STATUS_CHOLESKY));
  if (density)
    hipLaunchKernelGGL(k_g_unpad, dim3(grid_blocks((int64_t)n * n)), dim3(256), 0, s, g.Y, n, ld,
                       wfn_ok ? n : 0, density);
  if (coeff) {
    if (wfn_ok) {
      if (grid_eigenvectors(g, s, n) != hipSuccess) return -2;
      const double* Z = g.Y;
      const double* X = g.X;
      double* CT = g.CT;
      g_gemm<false, true>(
          s, n, n, [=] __device__(int i, int k) { return Z[(int64_t)k * ld + i]; },
          [=] __device__(int k, int j) { return X[(int64_t)k * ld + j]; }, AllK{n}, NoSkip{},
          [=] __device__(int i, int j, double val) { CT[(int64_t)i * ld + j] = val; });
    }
    hipLaunchKernelGGL(k_g_unpad, dim3(grid_blocks((int64_t)n * n)), dim3(256), 0, s, g.CT, n, ld,
                       wfn_ok ? n : 0, coeff);
  }
  if (hipStreamSynchronize(s) != hipSuccess) return -2;
  if (hs) {
    hs->iterations = iter;
    hs->status = status;
  }
  return 0;
}

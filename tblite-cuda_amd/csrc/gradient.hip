// Analytic nuclear gradient of the converged SCF energy (dE/dR, Eh/bohr).
//
// The free energy is stationary in the orbitals and the Fermi occupations, so
// only explicit geometry derivatives and one overlap term remain:
//   dE/dR = sum P dH0/dR - sum W dS/dR + sum_s v_s dq_s/dR|_P
//         + sum_a vdp_a . ddpat_a/dR|_P + sum_a vqp_a . dqpat_a/dR|_P
//         + repulsion, ES2, AES2, D4 two- and three-body at fixed moments
//         + the CN chains of the self energies, the multipole radii and the
//           D4 weights,
// with W = sum_k f_k e_k c_k c_k^T.  This is the gradient the reference
// sketches in FP64 (multipole_grad_3d / multipole_grad_cgto and the gradient
// shift_operator, src/multipole.cu:290-323, :422-519, :846-981 text; the
// repulsion dG, src/xtb.cu:118-121 text; the CN derivative dcndr,
// src/coulomb.cu:289-292 text) and fills the gradient argument of
// xtb_singlepoint (src/xtb.cu:188-197).
//
// Kernels, in stream order after the SCF (the wavefunction arrays coeff and
// density of the batch are required):
//   k_grad_pot    one workgroup per molecule: D4 weights, potentials of the
//                 final moments, on-site CN term, zeroed accumulators
//   k_grad_wdens  64 x 64 tiles of W over the occupied levels
//   k_grad_h0     one thread per neighbour pair: integral derivatives
//                 contracted on the fly (grad_pair.h), atomics per atom
//   k_grad_pairs  one wave per atom: repulsion, ES2, AES2, D4 two-body
//   k_grad_c6d    pair planes: CN derivatives of the three-body C6
//   k_grad_atm    three-body dispersion triples: geometry and dE/dCN(D4),
//                 accumulated in LDS per workgroup
//   k_grad_cn     one wave per atom: both CN chains
#include <hip/hip_runtime.h>
#include <math.h>

#include "aes2.h"
#include "common.h"
#include "d4.h"
#include "device_util.h"
#include "grad_pair.h"

#define GRAD_THREADS 256

__device__ __forceinline__ double exp_count_g(double k, double r, double r0)
{
  return 1.0 / (1.0 + exp(-k * (r0 / r - 1.0)));
}

// d/dr of exp_count: -k r0 / r^2 f (1 - f)
__device__ __forceinline__ double dexp_count_g(double k, double r, double r0)
{
  const double f = exp_count_g(k, r, r0);
  return -k * r0 / (r * r) * f * (1.0 - f);
}

// ------------------------------------------------------------ k_grad_pot
__global__ __launch_bounds__(GRAD_THREADS) void k_grad_pot(BatchDev b)
{
  const int m = blockIdx.x, tid = threadIdx.x;
  const int a0 = b.mol_atom_off[m], nat = b.mol_nat[m];
  const int s0 = b.mol_sh_off[m], nsh = b.mol_nsh[m];
  const int n = b.mol_nao[m];
  const double* __restrict__ xyz = b.xyz + 3 * (int64_t)a0;
  const int* __restrict__ el = b.at_el + a0;
  const DevParams& P = *b.par;
  const DevD4& d4 = P.d4;
  const double* __restrict__ qat = b.qat + a0;
  const double* __restrict__ qsh = b.qsh + s0;
  const double* __restrict__ dpat = b.dpat + 3 * (int64_t)a0;
  const double* __restrict__ qpat = b.qpat + 6 * (int64_t)a0;
  const double* __restrict__ mrad = b.mrad + a0;
  double* __restrict__ gd4 = b.g_d4 + (int64_t)TA_GD4 * TA_D4_MAXREF * a0;
  constexpr int K = TA_D4_MAXREF;

  // D4 weights at the final charges and at zero charge, with CN derivatives
  for (int i = tid; i < nat; i += blockDim.x) {
    for (int k = 0; k < 3; ++k) b.grad[3 * (int64_t)(a0 + i) + k] = 0.0;
    b.g_dcn4[a0 + i] = 0.0;
    if (d4.enabled) {
      double* w = gd4 + (int64_t)TA_GD4 * K * i;
      const double cn4 = b.d4_cn[a0 + i];
      double tmp[K];
      d4_weights(d4, el[i], cn4, qat[i], w, w + K);
      d4_weights_dcn(d4, el[i], cn4, qat[i], w + 2 * K);
      d4_weights(d4, el[i], cn4, 0.0, w + 3 * K, tmp);
      d4_weights_dcn(d4, el[i], cn4, 0.0, w + 4 * K);
    }
  }
  __syncthreads();

  // per atom: AES2 and D4 potentials (coulomb_potential and the D4
  // get_potential hook), on-site CN term of the self energies
  const double* __restrict__ Pm = b.density + b.mol_mat_off[m];
  const double d2cut = d4.disp2_cutoff * d4.disp2_cutoff;
  for (int j = tid; j < nat; j += blockDim.x) {
    const DevElem& ej = P.el[el[j]];
    double vd[3] = {0.0, 0.0, 0.0}, vq[6] = {0, 0, 0, 0, 0, 0}, va = 0.0;
    double vv[K] = {0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < nat; ++i) {
      if (i == j) continue;
      // pair (i, j): amat_sd(:, j, i) etc.
      const AesPair pij = aes_pair(xyz, mrad, i, j, P.dmp3, P.dmp5);
      for (int k = 0; k < 3; ++k) vd[k] += aes_sd(pij, k) * qat[i];
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) vd[k] += aes_dd(pij, k, l) * dpat[3 * i + l];
      for (int c = 0; c < 6; ++c) vq[c] += aes_sq(pij, c) * qat[i];
      // pair (j, i): vat_j += amat_sd(:, i, j) . dpat_i + amat_sq(:, i, j) . qpat_i
      const AesPair pji = aes_pair(xyz, mrad, j, i, P.dmp3, P.dmp5);
      for (int k = 0; k < 3; ++k) va += aes_sd(pji, k) * dpat[3 * i + k];
      for (int c = 0; c < 6; ++c) va += aes_sq(pji, c) * qpat[6 * i + c];
      if (d4.enabled) {
        const double dx = xyz[3 * j] - xyz[3 * i], dy = xyz[3 * j + 1] - xyz[3 * i + 1];
        const double dz = xyz[3 * j + 2] - xyz[3 * i + 2];
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 <= d2cut) {
          double dmp, ddr;
          d4_damp(d4, el[j], el[i], r2, dmp, ddr);
          d4_accumulate(d4, el[j], el[i], dmp, gd4 + (int64_t)TA_GD4 * K * i, vv);
        }
      }
    }
    for (int k = 0; k < 3; ++k) vd[k] += 2.0 * ej.dkernel * dpat[3 * j + k];
    for (int c = 0; c < 6; ++c) {
      const double sc = (c == 1 || c == 3 || c == 4) ? 2.0 : 1.0;
      vq[c] += 2.0 * ej.qkernel * qpat[6 * j + c] * sc;
    }
    if (d4.enabled) {
      const double* dgq = gd4 + (int64_t)TA_GD4 * K * j + K;
      for (int r = 0; r < d4.nref[el[j]]; ++r) va += dgq[r] * vv[r];
    }
    for (int k = 0; k < 3; ++k) b.g_vdp[3 * (int64_t)(a0 + j) + k] = vd[k];
    for (int c = 0; c < 6; ++c) b.g_vqp[6 * (int64_t)(a0 + j) + c] = vq[c];
    b.g_vat[a0 + j] = va;
    // on-site self energies: dE/dCN_j = -sum_{mu on j} P_mumu kcn
    const int is = b.at_ish[a0 + j];
    double dcn = 0.0;
    for (int s = 0; s < ej.nsh; ++s) {
      const int i0 = b.sh_iao[s0 + is + s];
      for (int a = 0; a < 2 * ej.ang[s] + 1; ++a)
        dcn -= Pm[(int64_t)(i0 + a) * n + (i0 + a)] * ej.kcn[s];
    }
    b.g_dcn[a0 + j] = dcn;
  }
  __syncthreads();
  // shell potentials: ES2 (gam qsh), ES3, atomic potential
  const double* __restrict__ gam = b.gam + b.mol_gam_off[m];
  for (int s = tid; s < nsh; s += blockDim.x) {
    double v = 0.0;
    for (int t = 0; t < nsh; ++t) v += gam[(int64_t)s * nsh + t] * qsh[t];
    const int at = b.sh_at[s0 + s];
    const DevElem& e = P.el[el[at]];
    v += qsh[s] * qsh[s] * e.hubbard_derivs[b.sh_ish[s0 + s]];
    b.g_vsh[s0 + s] = v + b.g_vat[a0 + at];
  }
}

// ---------------------------------------------------------- k_grad_wdens
// W(i, j) = sum_k C(i, k) f_k e_k C(j, k) over the levels with f_k above
// TA_OCC_MIN (a prefix of the ascending levels).  Work item w: molecule
// tile_mol[w], 64 x 64 tile (ti, tj), tj <= ti, written to both triangles.
#define WD_T 64
#define WD_K 16
__global__ __launch_bounds__(256) void k_grad_wdens(BatchDev b, const int* __restrict__ tiles)
{
  __shared__ double ci[WD_K][WD_T + 1], cj[WD_K][WD_T + 1], fe[WD_K];
  __shared__ int kmax_s;
  const int w = blockIdx.x, tid = threadIdx.x;
  const int m = tiles[3 * w], ti = tiles[3 * w + 1], tj = tiles[3 * w + 2];
  const int n = b.mol_nao[m], o0 = b.mol_ao_off[m];
  const double* __restrict__ C = b.coeff + b.mol_mat_off[m];
  const double* __restrict__ focc = b.focc + o0;
  const double* __restrict__ emo = b.emo + o0;
  double* __restrict__ W = b.g_w + b.mol_mat_off[m];
  if (tid == 0) kmax_s = 0;
  __syncthreads();
  {
    int last = 0;
    for (int k = tid; k < n; k += blockDim.x)
      if (focc[k] > TA_OCC_MIN) last = k + 1;
    atomicMax(&kmax_s, last);
  }
  __syncthreads();
  const int kmax = kmax_s;
  const int tx = tid & 15, ty = tid >> 4;  // 16 x 16 threads, 4 x 4 outputs each
  const int i0 = ti * WD_T, j0 = tj * WD_T;
  double acc[4][4];
  for (int a = 0; a < 4; ++a)
    for (int c = 0; c < 4; ++c) acc[a][c] = 0.0;
  for (int k0 = 0; k0 < kmax; k0 += WD_K) {
    // load C(i0 + r, k0 + kk) and C(j0 + r, k0 + kk): rows fastest in memory
    for (int x = tid; x < WD_K * WD_T; x += blockDim.x) {
      const int kk = x / WD_T, r = x - kk * WD_T;
      const int k = k0 + kk;
      const bool kin = k < kmax;
      ci[kk][r] = (kin && i0 + r < n) ? C[(int64_t)k * n + i0 + r] : 0.0;
      cj[kk][r] = (kin && j0 + r < n) ? C[(int64_t)k * n + j0 + r] : 0.0;
    }
    if (tid < WD_K) fe[tid] = (k0 + tid < kmax) ? focc[k0 + tid] * emo[k0 + tid] : 0.0;
    __syncthreads();
    for (int kk = 0; kk < WD_K; ++kk) {
      const double f = fe[kk];
      double a[4], c[4];
      for (int q = 0; q < 4; ++q) {
        a[q] = ci[kk][ty * 4 + q] * f;
        c[q] = cj[kk][tx * 4 + q];
      }
      for (int p = 0; p < 4; ++p)
        for (int q = 0; q < 4; ++q) acc[p][q] += a[p] * c[q];
    }
    __syncthreads();
  }
  for (int p = 0; p < 4; ++p) {
    const int i = i0 + ty * 4 + p;
    if (i >= n) continue;
    for (int q = 0; q < 4; ++q) {
      const int j = j0 + tx * 4 + q;
      if (j >= n) continue;
      W[(int64_t)j * n + i] = acc[p][q];
      W[(int64_t)i * n + j] = acc[p][q];
    }
  }
}

// ------------------------------------------------------------ k_grad_h0
// Spherical weights of one shell block -> Cartesian (both indices), then the
// contraction of grad_pair.h.
template <int LI, int LJ>
__device__ void h0_block_grad(const DevElem& ej, int jsh, const DevElem& ei, int ish, double r2,
                              const double vec[3], double intcut, const double* __restrict__ Pm,
                              const double* __restrict__ Wm, int n, int ii, int jj, double hfac,
                              const GradPairConst& k, double g[3], double& psh)
{
  constexpr int SI = GCart<LI>::s, SJ = GCart<LJ>::s, MI = GCart<LI>::n, MJ = GCart<LJ>::n;
  double ac[MJ * MI], pc[MJ * MI];
  {
    // half transform over j for each spherical i
    double ha[SI][MJ], hp[SI][MJ];
    for (int bb = 0; bb < SI; ++bb) {
      double sa[SJ], sp[SJ];
      for (int a = 0; a < SJ; ++a) {
        // density(jj + a, ii + bb), column-major
        const int64_t x = (int64_t)(ii + bb) * n + (jj + a);
        const double p = Pm[x];
        sp[a] = p;
        sa[a] = p * hfac - 2.0 * Wm[x];
      }
      sph_to_cart<LJ>(sa, ha[bb]);
      sph_to_cart<LJ>(sp, hp[bb]);
    }
    for (int mj = 0; mj < MJ; ++mj) {
      double ca[SI], cp[SI], wa[MI], wp[MI];
      for (int bb = 0; bb < SI; ++bb) {
        ca[bb] = ha[bb][mj];
        cp[bb] = hp[bb][mj];
      }
      sph_to_cart<LI>(ca, wa);
      sph_to_cart<LI>(cp, wp);
      for (int mi = 0; mi < MI; ++mi) {
        ac[mj + MJ * mi] = wa[mi];
        pc[mj + MJ * mi] = wp[mi];
      }
    }
  }
  shell_pair_grad<LI, LJ>(ej.alpha[jsh], ej.coeff[jsh], ej.nprim[jsh], ei.alpha[ish], ei.coeff[ish],
                          ei.nprim[ish], r2, vec, intcut, ac, pc, k, g, psh);
}

template <int LMAX>
__device__ void h0_shell_grad(int li, int lj, const DevElem& ej, int jsh, const DevElem& ei,
                              int ish, double r2, const double vec[3], double intcut,
                              const double* Pm, const double* Wm, int n, int ii, int jj,
                              double hfac, const GradPairConst& k, double g[3], double& psh)
{
  if constexpr (LMAX < 2) {
    switch (li * 2 + lj) {
      case 0: h0_block_grad<0, 0>(ej, jsh, ei, ish, r2, vec, intcut, Pm, Wm, n, ii, jj, hfac, k, g, psh); break;
      case 1: h0_block_grad<0, 1>(ej, jsh, ei, ish, r2, vec, intcut, Pm, Wm, n, ii, jj, hfac, k, g, psh); break;
      case 2: h0_block_grad<1, 0>(ej, jsh, ei, ish, r2, vec, intcut, Pm, Wm, n, ii, jj, hfac, k, g, psh); break;
      default: h0_block_grad<1, 1>(ej, jsh, ei, ish, r2, vec, intcut, Pm, Wm, n, ii, jj, hfac, k, g, psh); break;
    }
  } else {
    switch (li * 3 + lj) {
      case 0: h0_block_grad<0, 0>(ej, jsh, ei, ish, r2, vec, intcut, Pm, Wm, n, ii, jj, hfac, k, g, psh); break;
      case 1: h0_block_grad<0, 1>(ej, jsh, ei, ish, r2, vec, intcut, Pm, Wm, n, ii, jj, hfac, k, g, psh); break;
      case 2: h0_block_grad<0, 2>(ej, jsh, ei, ish, r2, vec, intcut, Pm, Wm, n, ii, jj, hfac, k, g, psh); break;
      case 3: h0_block_grad<1, 0>(ej, jsh, ei, ish, r2, vec, intcut, Pm, Wm, n, ii, jj, hfac, k, g, psh); break;
      case 4: h0_block_grad<1, 1>(ej, jsh, ei, ish, r2, vec, intcut, Pm, Wm, n, ii, jj, hfac, k, g, psh); break;
      case 5: h0_block_grad<1, 2>(ej, jsh, ei, ish, r2, vec, intcut, Pm, Wm, n, ii, jj, hfac, k, g, psh); break;
      case 6: h0_block_grad<2, 0>(ej, jsh, ei, ish, r2, vec, intcut, Pm, Wm, n, ii, jj, hfac, k, g, psh); break;
      case 7: h0_block_grad<2, 1>(ej, jsh, ei, ish, r2, vec, intcut, Pm, Wm, n, ii, jj, hfac, k, g, psh); break;
      default: h0_block_grad<2, 2>(ej, jsh, ei, ish, r2, vec, intcut, Pm, Wm, n, ii, jj, hfac, k, g, psh); break;
    }
  }
}

// Work list of k_integrals; one thread per neighbour-list entry (on-site
// items have no geometry derivative and return).
template <int LMAX>
__global__ __launch_bounds__(128) void k_grad_h0(BatchDev b, const int* __restrict__ work_mol,
                                                 const int* __restrict__ work_kind,
                                                 const int* __restrict__ work_start)
{
  const int w = blockIdx.x;
  if (work_kind[w] != 0) return;
  const int m = work_mol[w];
  const int item = work_start[w] + threadIdx.x;
  if (item >= b.mol_nnz[m]) return;
  const int a0 = b.mol_atom_off[m], nat = b.mol_nat[m];
  const int s0 = b.mol_sh_off[m];
  const int n = b.mol_nao[m];
  const double* __restrict__ xyz = b.xyz + 3 * (int64_t)a0;
  const int* __restrict__ el = b.at_el + a0;
  const DevParams& P = *b.par;
  const double* __restrict__ Pm = b.density + b.mol_mat_off[m];
  const double* __restrict__ Wm = b.g_w + b.mol_mat_off[m];
  const double* __restrict__ se = b.selfenergy + s0;
  const double* __restrict__ vsh = b.g_vsh + s0;

  // owning atom of the list entry: largest iat with inl(iat) <= item
  const int* inl = b.nl_inl + a0;
  int lo = 0, hi = nat - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (inl[mid] <= item) lo = mid;
    else hi = mid - 1;
  }
  while (lo < nat - 1 && inl[lo + 1] <= item) ++lo;
  const int iat = lo;
  const int jat = b.nl_jat[b.mol_nl_off[m] + item];
  if (jat == iat) return;
  const int ei_ = el[iat], ej_ = el[jat];
  const DevElem& ei = P.el[ei_];
  const DevElem& ej = P.el[ej_];
  const int is = b.at_ish[a0 + iat], js = b.at_ish[a0 + jat];
  double vec[3];
  for (int k = 0; k < 3; ++k) vec[k] = xyz[3 * iat + k] - xyz[3 * jat + k];
  const double r2 = vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2];
  const double rr = sqrt(sqrt(r2) / (ej.atomic_rad + ei.atomic_rad));
  const double drr = rr / (2.0 * r2);  // d rr / d vec = drr * vec

  GradPairConst kc;
  double kappa;
  grad_pair_const(b.g_vdp + 3 * (int64_t)(a0 + iat), b.g_vqp + 6 * (int64_t)(a0 + iat),
                  b.g_vdp + 3 * (int64_t)(a0 + jat), b.g_vqp + 6 * (int64_t)(a0 + jat), vec, kc,
                  kappa);
  double g[3] = {0.0, 0.0, 0.0}, dcn_i = 0.0, dcn_j = 0.0, gsh = 0.0;
  for (int ish = 0; ish < ei.nsh; ++ish) {
    const int li = ei.ang[ish];
    const int ii = b.sh_iao[s0 + is + ish];
    for (int jsh = 0; jsh < ej.nsh; ++jsh) {
      const int lj = ej.ang[jsh];
      const int jj = b.sh_iao[s0 + js + jsh];
      const double ki = ei.shpoly[ish], kj = ej.shpoly[jsh];
      const double shp = (1.0 + ki * rr) * (1.0 + kj * rr);
      const double hs = P.hscale[ei_][ej_][ish][jsh];
      const double hav = 0.5 * (se[is + ish] + se[js + jsh]);
      const double hij = hav * hs * shp;
      // alpha = P (2 H0 - v_a - v_b - kappa) - 2 W
      const double hfac = 2.0 * hij - vsh[js + jsh] - vsh[is + ish] - kappa;
      double psh = 0.0;
      h0_shell_grad<LMAX>(li, lj, ej, jsh, ei, ish, r2, vec, b.intcut, Pm, Wm, n, ii, jj, hfac, kc,
                          g, psh);
      // shell polynomial of H0: 2 sum P S dh/dvec
      gsh += 2.0 * psh * hav * hs * (ki * (1.0 + kj * rr) + kj * (1.0 + ki * rr)) * drr;
      dcn_i -= psh * ei.kcn[ish] * hs * shp;
      dcn_j -= psh * ej.kcn[jsh] * hs * shp;
    }
  }
  for (int k = 0; k < 3; ++k) {
    const double gk = g[k] + gsh * vec[k];
    atomicAdd(&b.grad[3 * (int64_t)(a0 + iat) + k], gk);
    atomicAdd(&b.grad[3 * (int64_t)(a0 + jat) + k], -gk);
  }
  atomicAdd(&b.g_dcn[a0 + iat], dcn_i);
  atomicAdd(&b.g_dcn[a0 + jat], dcn_j);
}

template __global__ void k_grad_h0<1>(BatchDev, const int*, const int*, const int*);
template __global__ void k_grad_h0<2>(BatchDev, const int*, const int*, const int*);

// ---------------------------------------------------------- k_grad_pairs
// AES2 energy of the ordered pair (i, j), u = R_i - R_j:
//   e = q_i dp_j . u g3 f3 + 1/2 dp_j^T (g3 f5 I - 3 u u^T g5 f5) dp_i + q_i g5 f5 qp_j . t(u)
// out: de/du (3) and de/d mrad (either atom)
__device__ __forceinline__ void aes_pair_grad(const double u[3], double mi, double mj, double dmp3,
                                              double dmp5, double qi, const double* di,
                                              const double* dj, const double* qj, double gv[3],
                                              double& dm)
{
  const double r1 = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  const double g1 = 1.0 / r1, g3 = g1 * g1 * g1, g5 = g3 * g1 * g1;
  const double rr = 0.5 * (mi + mj) * g1;
  const double p3 = aes_damp_pow(rr, dmp3), p5 = aes_damp_pow(rr, dmp5);
  const double f3 = 1.0 / (1.0 + 6.0 * p3), f5 = 1.0 / (1.0 + 6.0 * p5);
  // d f / d rr = -6 k rr^(k-1) f^2
  const double df3 = -6.0 * dmp3 * p3 / rr * f3 * f3;
  const double df5 = -6.0 * dmp5 * p5 / rr * f5 * f5;
  const double d3f3 = -3.0 * g3 * f3 * g1 - g3 * df3 * rr * g1;
  const double d3f5 = -3.0 * g3 * f5 * g1 - g3 * df5 * rr * g1;
  const double d5f5 = -5.0 * g5 * f5 * g1 - g5 * df5 * rr * g1;
  const double dju = dj[0] * u[0] + dj[1] * u[1] + dj[2] * u[2];
  const double diu = di[0] * u[0] + di[1] * u[1] + di[2] * u[2];
  const double djdi = dj[0] * di[0] + dj[1] * di[1] + dj[2] * di[2];
  const double t[6] = {u[0] * u[0], 2 * u[0] * u[1], u[1] * u[1], 2 * u[0] * u[2], 2 * u[1] * u[2],
                       u[2] * u[2]};
  double qt = 0.0;
  for (int k = 0; k < 6; ++k) qt += qj[k] * t[k];
  const double dqt[3] = {2.0 * (qj[0] * u[0] + qj[1] * u[1] + qj[3] * u[2]),
                         2.0 * (qj[1] * u[0] + qj[2] * u[1] + qj[4] * u[2]),
                         2.0 * (qj[3] * u[0] + qj[4] * u[1] + qj[5] * u[2])};
  const double radial = qi * dju * d3f3 + 0.5 * (djdi * d3f5 - 3.0 * d5f5 * dju * diu) + qi * d5f5 * qt;
  for (int k = 0; k < 3; ++k)
    gv[k] = radial * u[k] * g1 + qi * dj[k] * g3 * f3 - 1.5 * g5 * f5 * (dj[k] * diu + di[k] * dju) +
            qi * g5 * f5 * dqt[k];
  dm = (qi * dju * g3 * df3 + 0.5 * (djdi * g3 - 3.0 * g5 * dju * diu) * df5 + qi * g5 * qt * df5) *
       0.5 * g1;
}

// Work items of the pair planes (rows [a_begin, a_end) of one molecule); one
// wave per atom of the rows, lanes over the partner atoms.
__global__ __launch_bounds__(GRAD_THREADS) void k_grad_pairs(BatchDev b,
                                                             const int* __restrict__ work_mol,
                                                             const int* __restrict__ work_a0,
                                                             const int* __restrict__ work_a1)
{
  const int w = blockIdx.x;
  const int m = work_mol[w], ab = work_a0[w], ae = work_a1[w];
  const int lane = threadIdx.x & (TA_WAVE - 1), wv = wave_id();
  const int nwv = GRAD_THREADS / TA_WAVE;
  const int a0 = b.mol_atom_off[m], nat = b.mol_nat[m];
  const int s0 = b.mol_sh_off[m];
  const double* __restrict__ xyz = b.xyz + 3 * (int64_t)a0;
  const int* __restrict__ el = b.at_el + a0;
  const DevParams& P = *b.par;
  const DevD4& d4 = P.d4;
  const double* __restrict__ qat = b.qat + a0;
  const double* __restrict__ qsh = b.qsh + s0;
  const double* __restrict__ dpat = b.dpat + 3 * (int64_t)a0;
  const double* __restrict__ qpat = b.qpat + 6 * (int64_t)a0;
  const double* __restrict__ mrad = b.mrad + a0;
  const double* __restrict__ gd4 = b.g_d4 + (int64_t)TA_GD4 * TA_D4_MAXREF * a0;
  constexpr int K = TA_D4_MAXREF;
  const double rep_cut2 = P.rep_cutoff * P.rep_cutoff;
  const double d2cut = d4.disp2_cutoff * d4.disp2_cutoff;
  const double gexp = P.gexp;
  for (int a = ab + wv; a < ae; a += nwv) {
    const int ea_ = el[a];
    const DevElem& ea = P.el[ea_];
    const int isa = b.at_ish[a0 + a];
    double g[3] = {0.0, 0.0, 0.0}, dm = 0.0;
    double vv[K] = {0, 0, 0, 0, 0, 0, 0};
    for (int j = lane; j < nat; j += TA_WAVE) {
      if (j == a) continue;
      const int ej_ = el[j];
      const DevElem& ej = P.el[ej_];
      double u[3];
      for (int k = 0; k < 3; ++k) u[k] = xyz[3 * a + k] - xyz[3 * j + k];
      const double r2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
      if (r2 < 1.0e-12) continue;
      const double r1 = sqrt(r2);
      double gr = 0.0;  // radial part: dE/dr / r
      // repulsion
      if (r2 <= rep_cut2) {
        const double alpha = sqrt(ea.arep * ej.arep);
        const double zeff = ea.zeff * ej.zeff;
        const double kexp = (ea.z <= 2 && ej.z <= 2) ? P.rep_kexp_light : P.rep_kexp;
        const double r1k = pow(r1, kexp);
        const double de = zeff * exp(-alpha * r1k) / pow(r1, P.rep_rexp);
        gr += -(alpha * r1k * kexp + P.rep_rexp) * de / r2;
      }
      // ES2: d gamma / dr = -r^(g-1) gamma^(1+g)
      {
        const int isj = b.at_ish[a0 + j];
        const double r1g = pow(r1, gexp), rg1 = pow(r1, gexp - 1.0);
        double de = 0.0;
        for (int s = 0; s < ea.nsh; ++s)
          for (int t = 0; t < ej.nsh; ++t) {
            const double eta = 0.5 * (ea.hardness[s] + ej.hardness[t]);
            const double gm = 1.0 / pow(r1g + pow(eta, -gexp), 1.0 / gexp);
            de -= qsh[isa + s] * qsh[isj + t] * rg1 * pow(gm, 1.0 + gexp);
          }
        gr += de / r1;
      }
      // D4 two-body at the final charges
      if (d4.enabled && r2 <= d2cut) {
        double dmp, ddr;
        d4_damp(d4, ea_, ej_, r2, dmp, ddr);
        double acc[K] = {0, 0, 0, 0, 0, 0, 0};
        d4_accumulate(d4, ea_, ej_, 1.0, gd4 + (int64_t)TA_GD4 * K * j, acc);
        const double* ga = gd4 + (int64_t)TA_GD4 * K * a;
        double c6 = 0.0;
        for (int r = 0; r < d4.nref[ea_]; ++r) {
          c6 += ga[r] * acc[r];
          vv[r] += dmp * acc[r];
        }
        gr += c6 * ddr;
      }
      for (int k = 0; k < 3; ++k) g[k] += gr * u[k];
      // AES2: e(a, j) with u, e(j, a) with -u
      {
        double gv[3], d1, d2;
        aes_pair_grad(u, mrad[a], mrad[j], P.dmp3, P.dmp5, qat[a], dpat + 3 * a, dpat + 3 * j,
                      qpat + 6 * j, gv, d1);
        for (int k = 0; k < 3; ++k) g[k] += gv[k];
        const double un[3] = {-u[0], -u[1], -u[2]};
        aes_pair_grad(un, mrad[j], mrad[a], P.dmp3, P.dmp5, qat[j], dpat + 3 * j, dpat + 3 * a,
                      qpat + 6 * a, gv, d2);
        for (int k = 0; k < 3; ++k) g[k] -= gv[k];
        dm += d1 + d2;
      }
    }
    for (int k = 0; k < 3; ++k) g[k] = wave_sum(g[k]);
    dm = wave_sum(dm);
    double dcn4 = 0.0;
    if (d4.enabled) {
      const double* gc = gd4 + (int64_t)TA_GD4 * K * a + 2 * K;
      for (int r = 0; r < K; ++r) {
        const double s = wave_sum(vv[r]);
        if (r < d4.nref[ea_]) dcn4 += gc[r] * s;
      }
    }
    if (lane == 0) {
      for (int k = 0; k < 3; ++k) b.grad[3 * (int64_t)(a0 + a) + k] += g[k];
      // d mrad / d CN
      const double t1 = exp(-P.mp_kexp * (b.cn[a0 + a] - ea.mpvcn - P.mp_shift));
      const double dmr = (P.mp_rmax - ea.mprad) * P.mp_kexp * t1 / ((1.0 + t1) * (1.0 + t1));
      b.g_dcn[a0 + a] += dm * dmr;
      b.g_dcn4[a0 + a] += dcn4;
    }
  }
}

// ------------------------------------------------------------ k_grad_c6d
// Per atom pair the CN derivatives of the zero-charge C6 of the three-body
// term: c6d(i, k) = dC6_ik(0)/dCN_i = sum_rs dg0_ir/dCN C6ref(ir, ks) g0_ks
// (row i of the pair plane; column k's own derivative is c6d(k, i)).  Work
// items of the pair planes, threads over the pairs of their rows.
__global__ __launch_bounds__(GRAD_THREADS) void k_grad_c6d(BatchDev b, const int* __restrict__ work_mol,
                                                           const int* __restrict__ work_a0,
                                                           const int* __restrict__ work_a1)
{
  const int w = blockIdx.x, tid = threadIdx.x;
  const int m = work_mol[w], ab = work_a0[w], ae = work_a1[w];
  const int a0 = b.mol_atom_off[m], nat = b.mol_nat[m];
  const int* __restrict__ el = b.at_el + a0;
  const DevD4& d4 = b.par->d4;
  const double* __restrict__ gd4 = b.g_d4 + (int64_t)TA_GD4 * TA_D4_MAXREF * a0;
  double* __restrict__ c6d = b.g_c6d + b.mol_pair_off[m];
  constexpr int K = TA_D4_MAXREF;
  const int64_t p0 = (int64_t)ab * nat, p1 = (int64_t)ae * nat;
  for (int64_t idx = p0 + tid; idx < p1; idx += GRAD_THREADS) {
    const int i = (int)(idx / nat), k = (int)(idx - (int64_t)i * nat);
    double v = 0.0;
    if (i != k) {
      double acc[K] = {0, 0, 0, 0, 0, 0, 0};
      d4_accumulate(d4, el[i], el[k], 1.0, gd4 + (int64_t)TA_GD4 * K * k + 3 * K, acc);
      const double* gc0 = gd4 + (int64_t)TA_GD4 * K * i + 4 * K;
      for (int r = 0; r < d4.nref[el[i]]; ++r) v += gc0[r] * acc[r];
    }
    c6d[idx] = v;
  }
}

// ------------------------------------------------------------ k_grad_atm
// Three-body dispersion (zero-charge C6 of k_setup_pairs): for the triple
// with squared distances x (ij), y (ik), z (jk) and
//   E = c9 ang fdmp / (x y z)^(3/2),  ang = 3/8 A B C / (x y z) + 1,
//   A = x + z - y, B = x - z + y, C = -x + z + y,
//   fdmp = 1 / (1 + 6 (R0 / sqrt(x y z))^(alp/3)),
// the gradient follows from dE/dx, dE/dy, dE/dz, and dE/dC6 = E / (2 C6)
// for each of the three pairs goes straight into dE/dCN(D4) through the
// pair plane c6d of k_grad_c6d.  Work items and pair numbering as
// k_d4_atm.  With lds != 0 the workgroup accumulates per-atom gradient and
// CN terms in LDS (4 nat doubles of dynamic shared memory, LDS atomics) and
// adds them to the global arrays once at the end; otherwise every term is a
// global atomic.
__device__ __forceinline__ int64_t tri_count_g(int64_t i) { return i * (i - 1) / 2; }

__global__ __launch_bounds__(GRAD_THREADS) void k_grad_atm(BatchDev b, const int* __restrict__ work_mol,
                                                           const int* __restrict__ work_i0,
                                                           const int* __restrict__ work_i1, int lds)
{
  extern __shared__ double smem[];
  const int w = blockIdx.x, tid = threadIdx.x;
  const int m = work_mol[w], i0 = work_i0[w], i1 = work_i1[w];
  const int a0 = b.mol_atom_off[m], nat = b.mol_nat[m];
  const double* __restrict__ xyz = b.xyz + 3 * (int64_t)a0;
  const int* __restrict__ el = b.at_el + a0;
  const double* __restrict__ c60 = b.d4_c60 + b.mol_pair_off[m];
  const double* __restrict__ c6d = b.g_c6d + b.mol_pair_off[m];
  double* __restrict__ grad = b.grad + 3 * (int64_t)a0;
  double* __restrict__ dcn4 = b.g_dcn4 + a0;
  lds_dbl* lg = TA_LDS(smem);          // 3 nat gradient terms
  lds_dbl* lc = lg + 3 * nat;          // nat CN terms
  if (lds) {
    for (int x = tid; x < 4 * nat; x += GRAD_THREADS) lg[x] = 0.0;
    __syncthreads();
  }
  auto add_g = [&](int a, int x, double v) {
    if (lds) __hip_atomic_fetch_add(lg + 3 * a + x, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    else atomicAdd(&grad[3 * a + x], v);
  };
  auto add_c = [&](int a, double v) {
    if (lds) __hip_atomic_fetch_add(lc + a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    else atomicAdd(&dcn4[a], v);
  };
  const DevD4& d = b.par->d4;
  const double cut2 = d.disp3_cutoff * d.disp3_cutoff;
  const double p3 = d.alp / 3.0;
  const int64_t t0 = tri_count_g(i0), npair = tri_count_g(i1) - t0;
  for (int64_t p = tid; p < npair; p += GRAD_THREADS) {
    const int64_t q = t0 + p;
    int i = (int)((1.0 + sqrt(1.0 + 8.0 * (double)q)) * 0.5);
    while (tri_count_g(i) > q) --i;
    while (tri_count_g(i + 1) <= q) ++i;
    const int j = (int)(q - tri_count_g(i));
    double rij[3];
    for (int x = 0; x < 3; ++x) rij[x] = xyz[3 * j + x] - xyz[3 * i + x];
    const double x2 = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
    if (x2 > cut2) continue;
    const double ri = d.r4r2[el[i]], rj = d.r4r2[el[j]];
    const double r0ij = d.a1 * sqrt(3.0 * ri * rj) + d.a2;
    const double cij = c60[(int64_t)i * nat + j];
    double gi[3] = {0, 0, 0}, gj[3] = {0, 0, 0}, fij = 0.0, ci = 0.0, cj = 0.0;
    for (int k = 0; k < j; ++k) {
      double rik[3], rjk[3];
      for (int x = 0; x < 3; ++x) {
        rik[x] = xyz[3 * k + x] - xyz[3 * i + x];
        rjk[x] = xyz[3 * k + x] - xyz[3 * j + x];
      }
      const double y2 = rik[0] * rik[0] + rik[1] * rik[1] + rik[2] * rik[2];
      const double z2 = rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2];
      if (y2 > cut2 || z2 > cut2) continue;
      const double rk = d.r4r2[el[k]];
      const double r0p = r0ij * (d.a1 * sqrt(3.0 * ri * rk) + d.a2) * (d.a1 * sqrt(3.0 * rj * rk) + d.a2);
      const double cik = c60[(int64_t)i * nat + k], cjk = c60[(int64_t)j * nat + k];
      const double c9 = d.s9 * sqrt(fabs(cij * cik * cjk));
      const double pp = x2 * y2 * z2, rp = sqrt(pp);
      const double t = 6.0 * pow(r0p / rp, p3);
      const double fdmp = 1.0 / (1.0 + t);
      const double A = x2 + z2 - y2, B = x2 - z2 + y2, Cc = -x2 + z2 + y2;
      const double abc = A * B * Cc;
      const double ang = 0.375 * abc / pp + 1.0;
      const double dn = 1.0 / (rp * pp);
      const double et = c9 * ang * fdmp * dn;
      const double dabc[3] = {B * Cc + A * Cc - A * B, -B * Cc + A * Cc + A * B, B * Cc - A * Cc + A * B};
      const double sq[3] = {x2, y2, z2};
      double de[3];
      for (int s = 0; s < 3; ++s) {
        const double dang = 0.375 * (dabc[s] / pp - abc / (pp * sq[s]));
        const double dfd = fdmp * fdmp * t * (d.alp / 6.0) / sq[s];
        const double ddn = -1.5 * dn / sq[s];
        de[s] = c9 * (dang * fdmp * dn + ang * dfd * dn + ang * fdmp * ddn);
      }
      for (int x = 0; x < 3; ++x) {
        gi[x] -= 2.0 * (de[0] * rij[x] + de[1] * rik[x]);
        gj[x] += 2.0 * (de[0] * rij[x] - de[2] * rjk[x]);
        add_g(k, x, 2.0 * (de[1] * rik[x] + de[2] * rjk[x]));
      }
      // dE/dC6 = E / (2 C6) of each pair, into the CN terms of its two atoms
      if (cij != 0.0) fij += 0.5 * et / cij;
      double ck = 0.0;
      if (cik != 0.0) {
        const double f = 0.5 * et / cik;
        ci += f * c6d[(int64_t)i * nat + k];
        ck += f * c6d[(int64_t)k * nat + i];
      }
      if (cjk != 0.0) {
        const double f = 0.5 * et / cjk;
        cj += f * c6d[(int64_t)j * nat + k];
        ck += f * c6d[(int64_t)k * nat + j];
      }
      if (ck != 0.0) add_c(k, ck);
    }
    ci += fij * c6d[(int64_t)i * nat + j];
    cj += fij * c6d[(int64_t)j * nat + i];
    for (int x = 0; x < 3; ++x) {
      add_g(i, x, gi[x]);
      add_g(j, x, gj[x]);
    }
    if (ci != 0.0) add_c(i, ci);
    if (cj != 0.0) add_c(j, cj);
  }
  if (lds) {
    __syncthreads();
    for (int x = tid; x < 3 * nat; x += GRAD_THREADS)
      if (lg[x] != 0.0) atomicAdd(&grad[x], (double)lg[x]);
    for (int a = tid; a < nat; a += GRAD_THREADS)
      if (lc[a] != 0.0) atomicAdd(&dcn4[a], (double)lc[a]);
  }
}

// ------------------------------------------------------------- k_grad_cn
// grad_a += sum_j (dE/dCN_a + dE/dCN_j) dcount_aj/dr (R_a - R_j)/r for the
// GFN CN (double logistic count) and the D4 CN (error-function count)
__global__ __launch_bounds__(GRAD_THREADS) void k_grad_cn(BatchDev b, const int* __restrict__ work_mol,
                                                          const int* __restrict__ work_a0,
                                                          const int* __restrict__ work_a1)
{
  const int w = blockIdx.x;
  const int m = work_mol[w], ab = work_a0[w], ae = work_a1[w];
  const int lane = threadIdx.x & (TA_WAVE - 1), wv = wave_id();
  const int nwv = GRAD_THREADS / TA_WAVE;
  const int a0 = b.mol_atom_off[m], nat = b.mol_nat[m];
  const double* __restrict__ xyz = b.xyz + 3 * (int64_t)a0;
  const int* __restrict__ el = b.at_el + a0;
  const DevParams& P = *b.par;
  const DevD4& d4 = P.d4;
  const double* __restrict__ dcn = b.g_dcn + a0;
  const double* __restrict__ dcn4 = b.g_dcn4 + a0;
  const double cn_cut2 = P.cn_cutoff * P.cn_cutoff;
  const double d4_cut2 = d4.cn_cutoff * d4.cn_cutoff;
  const double sqrtpi = 1.7724538509055160273;
  for (int a = ab + wv; a < ae; a += nwv) {
    const DevElem& ea = P.el[el[a]];
    double g[3] = {0.0, 0.0, 0.0};
    for (int j = lane; j < nat; j += TA_WAVE) {
      if (j == a) continue;
      const DevElem& ej = P.el[el[j]];
      double u[3];
      for (int k = 0; k < 3; ++k) u[k] = xyz[3 * a + k] - xyz[3 * j + k];
      const double r2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
      if (r2 < 1.0e-12) continue;
      const double r1 = sqrt(r2);
      double gr = 0.0;
      if (r2 <= cn_cut2) {
        const double rc = ea.rcov + ej.rcov;
        const double fa = exp_count_g(P.cn_ka, r1, rc), fb = exp_count_g(P.cn_kb, r1, rc + P.cn_rshift);
        const double dc = dexp_count_g(P.cn_ka, r1, rc) * fb + fa * dexp_count_g(P.cn_kb, r1, rc + P.cn_rshift);
        gr += (dcn[a] + dcn[j]) * dc;
      }
      if (d4.enabled && r2 <= d4_cut2) {
        const double rc = ea.rcov + ej.rcov;
        const double den = fabs(ea.en - ej.en) + d4.k5;
        const double pre = d4.k4 * exp(-den * den / d4.k6);
        const double t = d4.kn * (r1 - rc) / rc;
        const double dc = -pre * d4.kn / (rc * sqrtpi) * exp(-t * t);
        gr += (dcn4[a] + dcn4[j]) * dc;
      }
      for (int k = 0; k < 3; ++k) g[k] += gr * u[k] / r1;
    }
    for (int k = 0; k < 3; ++k) g[k] = wave_sum(g[k]);
    if (lane == 0)
      for (int k = 0; k < 3; ++k) b.grad[3 * (int64_t)(a0 + a) + k] += g[k];
  }
}

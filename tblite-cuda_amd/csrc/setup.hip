// Per-molecule setup kernels: coordination numbers, repulsion, self energies,
// Coulomb interaction matrices, the neighbour list, and the overlap /
// multipole / H0 integrals.
//
// Launch shapes: k_setup and k_adjlist run one workgroup per molecule (the
// work is O(nat^2) per molecule); k_integrals runs over a host-built work
// list of (molecule, chunk) items so that large molecules spread over many
// workgroups.
#include <hip/hip_runtime.h>
#include <math.h>

#include "aes2.h"
#include "common.h"
#include "d4.h"
#include "device_util.h"

#define SETUP_THREADS 256

// ---------------------------------------------------------------- k_fill
__global__ void k_fill(double* __restrict__ p, int64_t n, double v)
{
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// --------------------------------------------------------------- k_setup
// Per atom and shell: ncoord_dexp (src/coulomb.cu:173-245 text),
// get_repulsion_derivs energy (src/xtb.cu:62-91 text), get_mrad
// (src/coulomb.cu:387-397 text), the D4 CN and zero-charge weights,
// get_selfenergy (src/h0.cu:137-170 text), get_occupation (src/h0.cu:61-91).
// The pair planes (get_amat_0d, src/coulomb.cu:15-47 text;
// get_multipole_matrix_0d, src/coulomb.cu:439-467 text; D4) follow in
// k_setup_pairs.
__device__ __forceinline__ double exp_count(double k, double r, double r0)
{
  return 1.0 / (1.0 + exp(-k * (r0 / r - 1.0)));
}

// CN, D4 CN and repulsion contributions of one partner at squared distance r2
__device__ __forceinline__ void setup_pair_terms(const DevParams& P, const DevElem& ei,
                                                 const DevElem& ej, double r2, double cn_cut2,
                                                 double d4_cut2, double rep_cut2, double& cn,
                                                 double& cn4, double& er)
{
  const DevD4& d4 = P.d4;
  double r1 = sqrt(r2);
  if (r2 <= cn_cut2) {
    double rc = ei.rcov + ej.rcov;
    cn += exp_count(P.cn_ka, r1, rc) * exp_count(P.cn_kb, r1, rc + P.cn_rshift);
  }
  if (d4.enabled && r2 <= d4_cut2) {
    // D4 CN: electronegativity-weighted error-function count
    const double rc = ei.rcov + ej.rcov;
    const double den = fabs(ei.en - ej.en) + d4.k5;
    cn4 += d4.k4 * exp(-den * den / d4.k6) * 0.5 * (1.0 + erf(-d4.kn * (r1 - rc) / rc));
  }
  if (r2 <= rep_cut2) {
    double alpha = sqrt(ei.arep * ej.arep);
    double zeff = ei.zeff * ej.zeff;
    double kexp = (ei.z <= 2 && ej.z <= 2) ? P.rep_kexp_light : P.rep_kexp;
    double r1k = pow(r1, kexp);
    double r1r = pow(r1, P.rep_rexp);
    er += 0.5 * zeff * exp(-alpha * r1k) / r1r;
  }
}

// For a periodic cell (mol_ntr > 0) the partners are every image (j, T) other
// than the atom itself within each cutoff (ncoord_dexp's trans(:, itr) loop,
// src/coulomb.cu:212-213 text); its three-body energies are summed later by
// k_pbc_atm and start from zero here.
__global__ __launch_bounds__(SETUP_THREADS) void k_setup(BatchDev b)
{
  const int m = blockIdx.x;
  const int tid = threadIdx.x;
  const int a0 = b.mol_atom_off[m], nat = b.mol_nat[m];
  const int s0 = b.mol_sh_off[m], nsh = b.mol_nsh[m];
  const double* __restrict__ xyz = b.xyz + 3 * (int64_t)a0;
  const int* __restrict__ el = b.at_el + a0;
  const DevParams& P = *b.par;

  const double cn_cut2 = P.cn_cutoff * P.cn_cutoff;
  const double rep_cut2 = P.rep_cutoff * P.rep_cutoff;
  const DevD4& d4 = P.d4;
  const double d4_cut2 = d4.cn_cutoff * d4.cn_cutoff;
  for (int i = tid; i < nat; i += blockDim.x) {
    const DevElem& ei = P.el[el[i]];
    const double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
    double cn = 0.0, er = 0.0, cn4 = 0.0;
    const int ntr = b.mol_ntr ? b.mol_ntr[m] : 0;
    if (ntr == 0) {
      for (int j = 0; j < nat; ++j) {
        if (j == i) continue;
        const DevElem& ej = P.el[el[j]];
        double dx = xi - xyz[3 * j], dy = yi - xyz[3 * j + 1], dz = zi - xyz[3 * j + 2];
        double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 < 1.0e-12) continue;
        setup_pair_terms(P, ei, ej, r2, cn_cut2, d4_cut2, rep_cut2, cn, cn4, er);
      }
    } else {
      const double* __restrict__ tr = b.trans + 3 * b.mol_tr_off[m];
      const double* cst = b.mol_ewald + TA_PBC_EWALD * m;
      const double cut2 = fmax(cn_cut2, fmax(rep_cut2, d4.enabled ? d4_cut2 : 0.0));
      for (int j = 0; j < nat; ++j) {
        const DevElem& ej = P.el[el[j]];
        const double d0[3] = {xi - xyz[3 * j], yi - xyz[3 * j + 1], zi - xyz[3 * j + 2]};
        int lo[3], hi[3];
        cell_box(cst, d0, sqrt(cut2), lo, hi);
        for (int n1 = lo[0]; n1 <= hi[0]; ++n1)
          for (int n2 = lo[1]; n2 <= hi[1]; ++n2)
            for (int n3 = lo[2]; n3 <= hi[2]; ++n3) {
              const int t = cell_tindex(cst, n1, n2, n3);
              const double dx = d0[0] - tr[3 * t], dy = d0[1] - tr[3 * t + 1], dz = d0[2] - tr[3 * t + 2];
              const double r2 = dx * dx + dy * dy + dz * dz;
              if (r2 < 1.0e-12 || r2 > cut2) continue;
              setup_pair_terms(P, ei, ej, r2, cn_cut2, d4_cut2, rep_cut2, cn, cn4, er);
            }
      }
      b.d4_e3[a0 + i] = 0.0;
    }
    b.cn[a0 + i] = cn;
    b.erep[a0 + i] = er;
    if (d4.enabled) {
      // zero-charge reference weights (C6 of the three-body term)
      b.d4_cn[a0 + i] = cn4;
      d4_weights(d4, el[i], cn4, 0.0, b.d4_gw + TA_D4_MAXREF * (int64_t)(a0 + i),
                 b.d4_dgw + TA_D4_MAXREF * (int64_t)(a0 + i));
    }
    double arg = cn - ei.mpvcn - P.mp_shift;
    double t1 = exp(-P.mp_kexp * arg);
    b.mrad[a0 + i] = ei.mprad + (P.mp_rmax - ei.mprad) / (1.0 + t1);
  }
  __syncthreads();

  for (int s = tid; s < nsh; s += blockDim.x) {
    int at = b.sh_at[s0 + s];
    const DevElem& e = P.el[el[at]];
    int ish = b.sh_ish[s0 + s];
    b.selfenergy[s0 + s] = e.selfenergy[ish] - e.kcn[ish] * b.cn[a0 + at];
    b.n0sh[s0 + s] = e.refocc[ish];
  }
}

// ---------------------------------------------------------- k_setup_pairs
// Atom-pair and shell-pair interaction planes of rows [a_begin, a_end) of one
// molecule (work item w; k_setup has written the CN-dependent radii and the
// zero-charge D4 weights): the ES2 matrix rows of the shells of those atoms,
// the AES2 planes, the D4 damping factors and zero-charge C6.  Large
// molecules spread over many workgroups.
__global__ __launch_bounds__(SETUP_THREADS) void k_setup_pairs(BatchDev b, const int* __restrict__ work_mol,
                                                               const int* __restrict__ work_a0,
                                                               const int* __restrict__ work_a1)
{
  const int w = blockIdx.x, tid = threadIdx.x;
  const int m = work_mol[w], ab = work_a0[w], ae = work_a1[w];
  const int a0 = b.mol_atom_off[m], nat = b.mol_nat[m];
  const int s0 = b.mol_sh_off[m], nsh = b.mol_nsh[m];
  const double* __restrict__ xyz = b.xyz + 3 * (int64_t)a0;
  const int* __restrict__ el = b.at_el + a0;
  const DevParams& P = *b.par;
  const DevD4& d4 = P.d4;

  // ES2 interaction matrix rows of the shells of atoms [ab, ae), arithmetic
  // mean of shell hardness
  double* __restrict__ gam = b.gam + b.mol_gam_off[m];
  const double g = P.gexp;
  const int sb = b.at_ish[a0 + ab];
  const int se = ae < nat ? b.at_ish[a0 + ae] : nsh;
  for (int idx = tid; idx < (se - sb) * nsh; idx += blockDim.x) {
    int i = sb + idx / nsh, j = idx - (idx / nsh) * nsh;
    int ai = b.sh_at[s0 + i], aj = b.sh_at[s0 + j];
    const DevElem& ei = P.el[el[ai]];
    const DevElem& ej = P.el[el[aj]];
    double eta = 0.5 * (ei.hardness[b.sh_ish[s0 + i]] + ej.hardness[b.sh_ish[s0 + j]]);
    double v;
    if (ai == aj) {
      v = eta;
    } else {
      double dx = xyz[3 * aj] - xyz[3 * ai], dy = xyz[3 * aj + 1] - xyz[3 * ai + 1];
      double dz = xyz[3 * aj + 2] - xyz[3 * ai + 2];
      double r1 = sqrt(dx * dx + dy * dy + dz * dz);
      v = 1.0 / pow(pow(r1, g) + pow(eta, -g), 1.0 / g);
    }
    gam[(int64_t)i * nsh + j] = v;
  }

  // AES2 matrices, planar: sd[k][i][j] = amat_sd(k, j, i), etc.
  const int64_t n2 = (int64_t)nat * nat;
  // (only the multi-workgroup SCF reads the planes: null when it has no molecule)
  double* __restrict__ sd = b.aes_sd ? b.aes_sd + 3 * b.mol_pair_off[m] : nullptr;
  double* __restrict__ dd = b.aes_dd ? b.aes_dd + 9 * b.mol_pair_off[m] : nullptr;
  double* __restrict__ sq = b.aes_sq ? b.aes_sq + 6 * b.mol_pair_off[m] : nullptr;
  const double* __restrict__ mrad = b.mrad + a0;
  const int64_t p0 = (int64_t)ab * nat, p1 = (int64_t)ae * nat;
  if (sd)
    for (int64_t idx = p0 + tid; idx < p1; idx += blockDim.x) {
      const int i = (int)(idx / nat), j = (int)(idx - (int64_t)i * nat);
      const AesPair pr = aes_pair(xyz, mrad, i, j, P.dmp3, P.dmp5);
      for (int k = 0; k < 3; ++k) sd[k * n2 + idx] = aes_sd(pr, k);
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) dd[(k * 3 + l) * n2 + idx] = (i != j) ? aes_dd(pr, k, l) : 0.0;
      for (int c = 0; c < 6; ++c) sq[c * n2 + idx] = aes_sq(pr, c);
    }

  // D4 two-body damping factors d_ij and zero-charge C6_ij (three-body term)
  if (d4.enabled) {
    double* __restrict__ disp = b.d4_disp + b.mol_pair_off[m];
    double* __restrict__ c60 = b.d4_c60 + b.mol_pair_off[m];
    const double* __restrict__ gw0 = b.d4_gw + TA_D4_MAXREF * (int64_t)a0;
    const double cut2 = d4.disp2_cutoff * d4.disp2_cutoff;
    for (int64_t idx = p0 + tid; idx < p1; idx += blockDim.x) {
      const int i = (int)(idx / nat), j = (int)(idx - (int64_t)i * nat);
      double dv = 0.0, cv = 0.0;
      if (i != j) {
        const int ei = el[i], ej = el[j];
        const double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1];
        const double dz = xyz[3 * i + 2] - xyz[3 * j + 2];
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 <= cut2) {
          const double rr = 3.0 * d4.r4r2[ei] * d4.r4r2[ej];
          const double r0 = d4.a1 * sqrt(rr) + d4.a2;
          const double r02 = r0 * r0, r06 = r02 * r02 * r02;
          const double r6 = r2 * r2 * r2;
          dv = -(d4.s6 / (r6 + r06) + d4.s8 * rr / (r6 * r2 + r06 * r02));
        }
        double acc[TA_D4_MAXREF] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        d4_accumulate(d4, ei, ej, 1.0, gw0 + TA_D4_MAXREF * j, acc);
        for (int r = 0; r < d4.nref[ei]; ++r) cv += gw0[TA_D4_MAXREF * i + r] * acc[r];
      }
      disp[idx] = d4.qref_uniform ? dv * cv : dv;
      c60[idx] = cv;
    }
  }
}

// ------------------------------------------------------------- k_d4_atm
// Three-body (Axilrod-Teller-Muto) dispersion with zero-charge C6 and
// rational damping of the product of the pair radii:
//   E3 = sum_{k<j<i} s9 sqrt(C6_ij C6_ik C6_jk) (3 cos a cos b cos c + 1)
//        / (r_ij r_ik r_jk)^3 / (1 + 6 (R0_ij R0_ik R0_jk / r_ij r_ik r_jk)^(alp/3))
// Work item w: anchors i in [i0, i1) of one molecule.  Its atom pairs
// (i, j < i) are numbered p = T(i) - T(i0) + j with T(i) = i (i - 1) / 2; the
// threads stride over p and loop over k < j.  The item's energy goes to
// d4_e3 of its first anchor, zero to the others (a deterministic
// per-molecule sum).
__device__ __forceinline__ int64_t tri_count(int64_t i) { return i * (i - 1) / 2; }

__global__ __launch_bounds__(SETUP_THREADS) void k_d4_atm(BatchDev b, const int* __restrict__ work_mol,
                                                          const int* __restrict__ work_i0,
                                                          const int* __restrict__ work_i1)
{
  __shared__ double red[SETUP_THREADS / 64];
  const int w = blockIdx.x, tid = threadIdx.x;
  const int m = work_mol[w], i0 = work_i0[w], i1 = work_i1[w];
  const int a0 = b.mol_atom_off[m], nat = b.mol_nat[m];
  const double* __restrict__ xyz = b.xyz + 3 * (int64_t)a0;
  const int* __restrict__ el = b.at_el + a0;
  const double* __restrict__ c60 = b.d4_c60 + b.mol_pair_off[m];
  const DevD4& d = b.par->d4;
  const double cut2 = d.disp3_cutoff * d.disp3_cutoff;
  const double p3 = d.alp / 3.0;
  const int64_t t0 = tri_count(i0), npair = tri_count(i1) - t0;
  double e = 0.0;
  for (int64_t p = tid; p < npair; p += SETUP_THREADS) {
    // anchor i: T(i) <= t0 + p < T(i + 1)
    const int64_t q = t0 + p;
    int i = (int)((1.0 + sqrt(1.0 + 8.0 * (double)q)) * 0.5);
    while (tri_count(i) > q) --i;
    while (tri_count(i + 1) <= q) ++i;
    const int j = (int)(q - tri_count(i));
    const double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
    const double xj = xyz[3 * j], yj = xyz[3 * j + 1], zj = xyz[3 * j + 2];
    const double r2ij = (xj - xi) * (xj - xi) + (yj - yi) * (yj - yi) + (zj - zi) * (zj - zi);
    if (r2ij > cut2) continue;
    const double ri = d.r4r2[el[i]], rj = d.r4r2[el[j]];
    const double r0ij = d.a1 * sqrt(3.0 * ri * rj) + d.a2;
    const double c6ij = c60[(int64_t)i * nat + j];
    for (int k = 0; k < j; ++k) {
      const double xk = xyz[3 * k], yk = xyz[3 * k + 1], zk = xyz[3 * k + 2];
      const double r2ik = (xk - xi) * (xk - xi) + (yk - yi) * (yk - yi) + (zk - zi) * (zk - zi);
      const double r2jk = (xk - xj) * (xk - xj) + (yk - yj) * (yk - yj) + (zk - zj) * (zk - zj);
      if (r2ik > cut2 || r2jk > cut2) continue;
      const double rk = d.r4r2[el[k]];
      const double r0ik = d.a1 * sqrt(3.0 * ri * rk) + d.a2;
      const double r0jk = d.a1 * sqrt(3.0 * rj * rk) + d.a2;
      const double c9 = d.s9 * sqrt(fabs(c6ij * c60[(int64_t)i * nat + k] * c60[(int64_t)j * nat + k]));
      const double r2ijk = r2ij * r2ik * r2jk, rijk = sqrt(r2ijk);
      const double fdmp = 1.0 / (1.0 + 6.0 * pow(r0ij * r0ik * r0jk / rijk, p3));
      const double ang = 0.375 * (r2ij + r2jk - r2ik) * (r2ij - r2jk + r2ik) *
                             (-r2ij + r2jk + r2ik) / r2ijk + 1.0;
      e += c9 * ang * fdmp / (rijk * r2ijk);
    }
  }
  e = block_sum(e, red);
  for (int i = i0 + tid; i < i1; i += SETUP_THREADS) b.d4_e3[a0 + i] = (i == i0) ? e : 0.0;
}

// ------------------------------------------------------------- k_adjlist
// generate (src/adjlist.cu:15-44 text): half list, jat <= iat, skip
// r2 < epsilon or r2 > cutoff^2.  Integer output, bit-exact: the squared
// distance is formed as ((dx*dx + dy*dy) + dz*dz) with no contraction, the
// same operation sequence as the CPU restatement.
__device__ __forceinline__ double pair_r2_exact(const double* __restrict__ xyz, int i, int j)
{
#pragma clang fp contract(off)
  double dx = xyz[3 * i] - xyz[3 * j];
  double dy = xyz[3 * i + 1] - xyz[3 * j + 1];
  double dz = xyz[3 * i + 2] - xyz[3 * j + 2];
  double xx = dx * dx;
  double yy = dy * dy;
  double zz = dz * dz;
  double s = xx + yy;
  return s + zz;
}

// (R_i - R_j) - T with the same operation sequence as pair_r2_exact
__device__ __forceinline__ double pair_r2_image_exact(const double* __restrict__ xyz, int i, int j,
                                                      const double* __restrict__ t)
{
#pragma clang fp contract(off)
  double dx = (xyz[3 * i] - xyz[3 * j]) - t[0];
  double dy = (xyz[3 * i + 1] - xyz[3 * j + 1]) - t[1];
  double dz = (xyz[3 * i + 2] - xyz[3 * j + 2]) - t[2];
  double xx = dx * dx;
  double yy = dy * dy;
  double zz = dz * dz;
  double s = xx + yy;
  return s + zz;
}

// Neighbour list of a periodic cell: entries (jat, itr), jat <= iat, every
// translation itr in table order (generate, src/adjlist.cu:26-30 text).  One
// thread per atom, block scan per chunk of atoms as for molecules.
__device__ void adjlist_cell(const BatchDev& b, int m, int* red)
{
  const int tid = threadIdx.x;
  const int a0 = b.mol_atom_off[m], nat = b.mol_nat[m];
  const double* __restrict__ xyz = b.xyz + 3 * (int64_t)a0;
  const double* __restrict__ tr = b.trans + 3 * b.mol_tr_off[m];
  const double* cst = b.mol_ewald + TA_PBC_EWALD * m;
  const double cutoff = b.mol_cutoff[m];
  const double cutoff2 = cutoff * cutoff;
  const double eps = 2.220446049250313e-16;
  int* __restrict__ jat_out = b.nl_jat + b.mol_nl_off[m];
  int* __restrict__ itr_out = b.nl_itr + b.mol_nl_off[m];
  int base = 0;
  for (int chunk = 0; chunk < nat; chunk += blockDim.x) {
    const int i = chunk + tid;
    // translations of each pair's box in table order (the rest lie beyond
    // the cutoff)
    auto walk = [&](int j, auto&& visit) {
      const double d[3] = {xyz[3 * i] - xyz[3 * j], xyz[3 * i + 1] - xyz[3 * j + 1],
                           xyz[3 * i + 2] - xyz[3 * j + 2]};
      int lo[3], hi[3];
      cell_box(cst, d, cutoff, lo, hi);
      for (int n1 = lo[0]; n1 <= hi[0]; ++n1)
        for (int n2 = lo[1]; n2 <= hi[1]; ++n2)
          for (int n3 = lo[2]; n3 <= hi[2]; ++n3) {
            const int t = cell_tindex(cst, n1, n2, n3);
            double r2 = pair_r2_image_exact(xyz, i, j, tr + 3 * t);
            if (!(r2 < eps || r2 > cutoff2)) visit(t);
          }
    };
    int cnt = 0;
    if (i < nat)
      for (int j = 0; j <= i; ++j) walk(j, [&](int) { ++cnt; });
    int tot;
    int excl = block_exclusive_scan(cnt, red, &tot);
    if (i < nat) {
      b.nl_inl[a0 + i] = base + excl;
      b.nl_nnl[a0 + i] = cnt;
      int pos = base + excl;
      for (int j = 0; j <= i; ++j)
        walk(j, [&](int t) {
          jat_out[pos] = j;
          itr_out[pos] = t;
          ++pos;
        });
    }
    base += tot;
  }
  if (tid == 0) b.mol_nnz[m] = base;
}

__global__ __launch_bounds__(SETUP_THREADS) void k_adjlist(BatchDev b)
{
  __shared__ int red[SETUP_THREADS / TA_WAVE];
  const int m = blockIdx.x;
  if (b.mol_ntr && b.mol_ntr[m] > 0) {
    adjlist_cell(b, m, red);
    return;
  }
  const int tid = threadIdx.x;
  const int a0 = b.mol_atom_off[m], nat = b.mol_nat[m];
  const double* __restrict__ xyz = b.xyz + 3 * (int64_t)a0;
  const double cutoff = b.mol_cutoff[m];
  const double cutoff2 = cutoff * cutoff;
  const double eps = 2.220446049250313e-16;
  int* __restrict__ jat_out = b.nl_jat + b.mol_nl_off[m];
  int base = 0;
  for (int chunk = 0; chunk < nat; chunk += blockDim.x) {
    const int i = chunk + tid;
    int cnt = 0;
    if (i < nat)
      for (int j = 0; j <= i; ++j) {
        double r2 = pair_r2_exact(xyz, i, j);
        if (!(r2 < eps || r2 > cutoff2)) ++cnt;
      }
    int tot;
    int excl = block_exclusive_scan(cnt, red, &tot);
    if (i < nat) {
      b.nl_inl[a0 + i] = base + excl;
      b.nl_nnl[a0 + i] = cnt;
      int pos = base + excl;
      for (int j = 0; j <= i; ++j) {
        double r2 = pair_r2_exact(xyz, i, j);
        if (!(r2 < eps || r2 > cutoff2)) jat_out[pos++] = j;
      }
    }
    base += tot;
  }
  if (tid == 0) b.mol_nnz[m] = base;
}

// ----------------------------------------------------------- integrals
// Overlap, dipole and quadrupole integrals over contracted Cartesian
// Gaussians (multipole_cgto, src/integral/multipole.h:34-120 text; 1D moment
// recursion multipole_3d src/integral/multipole.cu:411-457 text), the
// Cartesian -> spherical transform (src/integral/trafo.h:6-108) and the
// operator shift to the other centre (shift_operator, src/h0.cu:251-289).
//
// For each primitive pair the 1D factors are tabulated once per axis and
// power pair, then combined for every Cartesian component pair.

template <int L>
struct Cart {
  static constexpr int n = (L + 1) * (L + 2) / 2;
  static constexpr int s = 2 * L + 1;
};

// Cartesian exponents (x, y, z), tblite order: s; p = y, z, x;
// d = xx, yy, zz, xy, xz, yz
__device__ __forceinline__ void cart_pow(int l, int c, int& px, int& py, int& pz)
{
  if (l == 0) {
    px = py = pz = 0;
  } else if (l == 1) {
    px = (c == 2);
    py = (c == 0);
    pz = (c == 1);
  } else {
    const int tx[6] = {2, 0, 0, 1, 1, 0}, ty[6] = {0, 2, 0, 1, 0, 1}, tz[6] = {0, 0, 2, 0, 1, 1};
    px = tx[c];
    py = ty[c];
    pz = tz[c];
  }
}

// spherical component a of a Cartesian vector v (d order z2, xz, yz, x2-y2, xy)
template <int L>
__device__ __forceinline__ double to_sph(const double* v, int a)
{
  if constexpr (L < 2) {
    return v[a];
  } else {
    const double s3 = 1.7320508075688772;
    switch (a) {
      case 0: return v[2] - 0.5 * (v[0] + v[1]);
      case 1: return s3 * v[4];
      case 2: return s3 * v[5];
      case 3: return 0.5 * s3 * (v[0] - v[1]);
      default: return s3 * v[3];
    }
  }
}

// Block for shell ish of element ei (center i) and shell jsh of element ej
// (center j), vec = R_i - R_j.  Outputs in tblite order ij = jao + SJ*iao:
// s[ij], d[3*ij + k], q[6*ij + k] (operator on center i, traceless).
template <int LI, int LJ>
__device__ void multipole_block(const DevElem& ej, int jsh, const DevElem& ei, int ish,
                                double r2, const double vec[3], double intcut, double* s,
                                double* d, double* q)
{
  constexpr int MI = Cart<LI>::n, MJ = Cart<LJ>::n, SI = Cart<LI>::s, SJ = Cart<LJ>::s;
  constexpr int LL = LI + LJ;
  double s3[MI * MJ], d3[MI * MJ * 3], q3[MI * MJ * 6];
#pragma unroll
  for (int x = 0; x < MI * MJ; ++x) s3[x] = 0.0;
#pragma unroll
  for (int x = 0; x < MI * MJ * 3; ++x) d3[x] = 0.0;
#pragma unroll
  for (int x = 0; x < MI * MJ * 6; ++x) q3[x] = 0.0;
  const double sqrtpi3 = 5.568327996831707845; // pi^(3/2)
  const double dfac[5] = {1.0, 1.0, 3.0, 15.0, 105.0};
  const int npi = ei.nprim[ish], npj = ej.nprim[jsh];
  for (int ip = 0; ip < npi; ++ip) {
    const double ai = ei.alpha[ish][ip], ci = ei.coeff[ish][ip];
    for (int jp = 0; jp < npj; ++jp) {
      const double aj = ej.alpha[jsh][jp];
      const double eab = ai + aj, oab = 1.0 / eab;
      const double est = ai * aj * r2 * oab;
      if (est > intcut) continue;
      const double cc = ci * ej.coeff[jsh][jp] * exp(-est) * sqrtpi3 * oab * sqrt(oab);
      double s1d[LL + 3];
      {
        double h = 0.5 * oab, pw = 1.0;
#pragma unroll
        for (int l = 0; l <= LL + 2; ++l) {
          if (l & 1) {
            s1d[l] = 0.0;
          } else {
            s1d[l] = pw * dfac[l / 2];
            pw *= h;
          }
        }
      }
      // v1[axis][pi][pj][moment]
      double v1[3][LI + 1][LJ + 1][3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double rpi = -vec[k] * aj * oab;
        const double rpj = vec[k] * ai * oab;
#pragma unroll
        for (int pi = 0; pi <= LI; ++pi)
#pragma unroll
          for (int pj = 0; pj <= LJ; ++pj) {
            // (t + rpi)^pi (t + rpj)^pj expanded in powers of t = x - P
            double vv[LL + 1];
#pragma unroll
            for (int l = 0; l <= LL; ++l) vv[l] = 0.0;
            double ci_[LI + 1], cj_[LJ + 1];
            {
              const double bin[3][3] = {{1, 0, 0}, {1, 1, 0}, {1, 2, 1}};
#pragma unroll
              for (int a = 0; a <= LI; ++a) {
                double pw = 1.0;
                for (int e = 0; e < pi - a; ++e) pw *= rpi;
                ci_[a] = (a <= pi) ? bin[pi][a] * pw : 0.0;
              }
#pragma unroll
              for (int a = 0; a <= LJ; ++a) {
                double pw = 1.0;
                for (int e = 0; e < pj - a; ++e) pw *= rpj;
                cj_[a] = (a <= pj) ? bin[pj][a] * pw : 0.0;
              }
            }
#pragma unroll
            for (int a = 0; a <= LI; ++a)
#pragma unroll
              for (int c = 0; c <= LJ; ++c) vv[a + c] += ci_[a] * cj_[c];
            double e0 = 0.0, e1 = 0.0, e2 = 0.0;
#pragma unroll
            for (int l = 0; l <= LL; ++l) {
              e0 += s1d[l] * vv[l];
              e1 += (s1d[l + 1] + rpi * s1d[l]) * vv[l];
              e2 += (s1d[l + 2] + 2.0 * rpi * s1d[l + 1] + rpi * rpi * s1d[l]) * vv[l];
            }
            v1[k][pi][pj][0] = e0;
            v1[k][pi][pj][1] = e1;
            v1[k][pi][pj][2] = e2;
          }
      }
#pragma unroll
      for (int mi = 0; mi < MI; ++mi) {
        int ix, iy, iz;
        cart_pow(LI, mi, ix, iy, iz);
#pragma unroll
        for (int mj = 0; mj < MJ; ++mj) {
          int jx, jy, jz;
          cart_pow(LJ, mj, jx, jy, jz);
          const double* X = v1[0][ix][jx];
          const double* Y = v1[1][iy][jy];
          const double* Z = v1[2][iz][jz];
          const int ij = mj + MJ * mi;
          s3[ij] += cc * X[0] * Y[0] * Z[0];
          d3[3 * ij + 0] += cc * X[1] * Y[0] * Z[0];
          d3[3 * ij + 1] += cc * X[0] * Y[1] * Z[0];
          d3[3 * ij + 2] += cc * X[0] * Y[0] * Z[1];
          q3[6 * ij + 0] += cc * X[2] * Y[0] * Z[0];
          q3[6 * ij + 1] += cc * X[1] * Y[1] * Z[0];
          q3[6 * ij + 2] += cc * X[0] * Y[2] * Z[0];
          q3[6 * ij + 3] += cc * X[1] * Y[0] * Z[1];
          q3[6 * ij + 4] += cc * X[0] * Y[1] * Z[1];
          q3[6 * ij + 5] += cc * X[0] * Y[0] * Z[2];
        }
      }
    }
  }
  // Cartesian -> spherical on both indices, then remove the quadrupole trace
#pragma unroll
  for (int comp = 0; comp < 10; ++comp) {
    // half transform over j for every Cartesian i
    double half[MI][SJ];
#pragma unroll
    for (int mi = 0; mi < MI; ++mi) {
      double row[MJ];
#pragma unroll
      for (int mj = 0; mj < MJ; ++mj) {
        const int ij = mj + MJ * mi;
        row[mj] = comp == 0 ? s3[ij] : (comp < 4 ? d3[3 * ij + comp - 1] : q3[6 * ij + comp - 4]);
      }
#pragma unroll
      for (int a = 0; a < SJ; ++a) half[mi][a] = to_sph<LJ>(row, a);
    }
#pragma unroll
    for (int a = 0; a < SJ; ++a) {
      double col[MI];
#pragma unroll
      for (int mi = 0; mi < MI; ++mi) col[mi] = half[mi][a];
#pragma unroll
      for (int bb = 0; bb < SI; ++bb) {
        const double val = to_sph<LI>(col, bb);
        const int ij = a + SJ * bb;
        if (comp == 0) s[ij] = val;
        else if (comp < 4) d[3 * ij + comp - 1] = val;
        else q[6 * ij + comp - 4] = val;
      }
    }
  }
#pragma unroll
  for (int ij = 0; ij < SI * SJ; ++ij) {
    double* qq = q + 6 * ij;
    const double tr = 0.5 * (qq[0] + qq[2] + qq[5]);
    qq[0] = 1.5 * qq[0] - tr;
    qq[1] = 1.5 * qq[1];
    qq[2] = 1.5 * qq[2] - tr;
    qq[3] = 1.5 * qq[3];
    qq[4] = 1.5 * qq[4];
    qq[5] = 1.5 * qq[5] - tr;
  }
}

// LMAX: the largest angular momentum the kernel instantiation handles (1:
// s and p shells only, so the d-shell blocks and their registers are not
// compiled in; 2: all)
template <int LMAX>
__device__ void shell_block(int li, int lj, const DevElem& ej, int jsh, const DevElem& ei, int ish,
                            double r2, const double vec[3], double intcut, double* s, double* d,
                            double* q)
{
  if constexpr (LMAX < 2) {
    switch (li * 2 + lj) {
      case 0: multipole_block<0, 0>(ej, jsh, ei, ish, r2, vec, intcut, s, d, q); break;
      case 1: multipole_block<0, 1>(ej, jsh, ei, ish, r2, vec, intcut, s, d, q); break;
      case 2: multipole_block<1, 0>(ej, jsh, ei, ish, r2, vec, intcut, s, d, q); break;
      default: multipole_block<1, 1>(ej, jsh, ei, ish, r2, vec, intcut, s, d, q); break;
    }
  } else {
    switch (li * 3 + lj) {
      case 0: multipole_block<0, 0>(ej, jsh, ei, ish, r2, vec, intcut, s, d, q); break;
      case 1: multipole_block<0, 1>(ej, jsh, ei, ish, r2, vec, intcut, s, d, q); break;
      case 2: multipole_block<0, 2>(ej, jsh, ei, ish, r2, vec, intcut, s, d, q); break;
      case 3: multipole_block<1, 0>(ej, jsh, ei, ish, r2, vec, intcut, s, d, q); break;
      case 4: multipole_block<1, 1>(ej, jsh, ei, ish, r2, vec, intcut, s, d, q); break;
      case 5: multipole_block<1, 2>(ej, jsh, ei, ish, r2, vec, intcut, s, d, q); break;
      case 6: multipole_block<2, 0>(ej, jsh, ei, ish, r2, vec, intcut, s, d, q); break;
      case 7: multipole_block<2, 1>(ej, jsh, ei, ish, r2, vec, intcut, s, d, q); break;
      default: multipole_block<2, 2>(ej, jsh, ei, ish, r2, vec, intcut, s, d, q); break;
    }
  }
}

// Work item kinds for k_integrals
#define WORK_PAIRS 0
#define WORK_ONSITE 1

// get_hamiltonian (src/h0.cu:306-477 text) plus the on-site blocks.  One
// thread per atom pair (neighbour-list entry) or per atom (on-site); each
// thread writes a disjoint set of matrix elements, so no atomics.  LMAX = 1
// serves molecules without d shells (the host splits the work list by it).
template <int LMAX>
__global__ __launch_bounds__(128) void k_integrals(BatchDev b, const int* __restrict__ work_mol,
                                                 const int* __restrict__ work_kind,
                                                 const int* __restrict__ work_start)
{
  const int w = blockIdx.x;
  const int m = work_mol[w];
  const int kind = work_kind[w];
  const int item = work_start[w] + threadIdx.x;
  const int a0 = b.mol_atom_off[m], nat = b.mol_nat[m];
  const int s0 = b.mol_sh_off[m];
  const int n = b.mol_nao[m];
  const int64_t n2 = (int64_t)n * n;
  const double* __restrict__ xyz = b.xyz + 3 * (int64_t)a0;
  const int* __restrict__ el = b.at_el + a0;
  const DevParams& P = *b.par;
  double* __restrict__ S = b.S + b.mol_mat_off[m];
  double* __restrict__ H = b.H + b.mol_mat_off[m];
  double* __restrict__ D = b.D + 3 * b.mol_mat_off[m];
  double* __restrict__ Q = b.Q + 6 * b.mol_mat_off[m];
  const double* __restrict__ se = b.selfenergy + s0;
  const double intcut = b.intcut;

  int iat, jat;
  if (kind == WORK_PAIRS) {
    if (item >= b.mol_nnz[m]) return;
    // locate the owning atom: largest iat with inl(iat) <= item
    const int* inl = b.nl_inl + a0;
    int lo = 0, hi = nat - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (inl[mid] <= item) lo = mid;
      else hi = mid - 1;
    }
    // skip atoms with empty lists that share the same offset
    while (lo < nat - 1 && inl[lo + 1] <= item) ++lo;
    iat = lo;
    jat = b.nl_jat[b.mol_nl_off[m] + item];
  } else {
    if (item >= nat) return;
    iat = jat = item;
  }
  const int ei_ = el[iat], ej_ = el[jat];
  const DevElem& ei = P.el[ei_];
  const DevElem& ej = P.el[ej_];
  // first shell / AO of each atom
  int is = 0, ii0 = 0;
  {
    // shells of this molecule are ordered by atom; scan the atom's first shell
    // through the shell->atom map (nsh per atom is at most 3)
    const int* sh_at = b.sh_at + s0;
    int lo = 0, hi = b.mol_nsh[m] - 1;
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (sh_at[mid] < iat) lo = mid + 1;
      else hi = mid;
    }
    is = lo;
    ii0 = b.sh_iao[s0 + is];
  }
  int js = is, jj0 = ii0;
  if (jat != iat) {
    const int* sh_at = b.sh_at + s0;
    int lo = 0, hi = b.mol_nsh[m] - 1;
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (sh_at[mid] < jat) lo = mid + 1;
      else hi = mid;
    }
    js = lo;
    jj0 = b.sh_iao[s0 + js];
  }
  double vec[3];
  for (int k = 0; k < 3; ++k) vec[k] = xyz[3 * iat + k] - xyz[3 * jat + k];
  const double r2 = vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2];
  const double rr = sqrt(sqrt(r2) / (ej.atomic_rad + ei.atomic_rad));

  constexpr int NB = (2 * LMAX + 1) * (2 * LMAX + 1);  // largest shell block
  double st[NB], dt[3 * NB], qt[6 * NB];
  for (int ish = 0; ish < ei.nsh; ++ish) {
    const int li = ei.ang[ish];
    const int ii = b.sh_iao[s0 + is + ish];
    for (int jsh = 0; jsh < ej.nsh; ++jsh) {
      const int lj = ej.ang[jsh];
      const int jj = b.sh_iao[s0 + js + jsh];
      shell_block<LMAX>(li, lj, ej, jsh, ei, ish, r2, vec, intcut, st, dt, qt);
      const int si = 2 * li + 1, sj = 2 * lj + 1;
      if (iat != jat) {
        const double shp = (1.0 + ei.shpoly[ish] * rr) * (1.0 + ej.shpoly[jsh] * rr);
        const double hij =
            0.5 * (se[is + ish] + se[js + jsh]) * P.hscale[ei_][ej_][ish][jsh] * shp;
        for (int iao = 0; iao < si; ++iao)
          for (int jao = 0; jao < sj; ++jao) {
            const int ij = jao + sj * iao;
            const double sv = st[ij];
            const double* di = dt + 3 * ij;
            const double* qi = qt + 6 * ij;
            // operator moved to center j: r - R_j = (r - R_i) + vec
            double dj[3], qj[6];
            dj[0] = di[0] + vec[0] * sv;
            dj[1] = di[1] + vec[1] * sv;
            dj[2] = di[2] + vec[2] * sv;
            double t0 = 2 * vec[0] * di[0] + vec[0] * vec[0] * sv;
            double t2 = 2 * vec[1] * di[1] + vec[1] * vec[1] * sv;
            double t5 = 2 * vec[2] * di[2] + vec[2] * vec[2] * sv;
            double t1 = vec[0] * di[1] + vec[1] * di[0] + vec[0] * vec[1] * sv;
            double t3 = vec[0] * di[2] + vec[2] * di[0] + vec[0] * vec[2] * sv;
            double t4 = vec[1] * di[2] + vec[2] * di[1] + vec[1] * vec[2] * sv;
            double tr = 0.5 * (t0 + t2 + t5);
            qj[0] = qi[0] + 1.5 * t0 - tr;
            qj[1] = qi[1] + 1.5 * t1;
            qj[2] = qi[2] + 1.5 * t2 - tr;
            qj[3] = qi[3] + 1.5 * t3;
            qj[4] = qi[4] + 1.5 * t4;
            qj[5] = qi[5] + 1.5 * t5 - tr;
            const int64_t a_ij = (int64_t)(ii + iao) * n + (jj + jao); // fortran (jj+jao, ii+iao)
            const int64_t a_ji = (int64_t)(jj + jao) * n + (ii + iao); // fortran (ii+iao, jj+jao)
            S[a_ij] = sv;
            S[a_ji] = sv;
            H[a_ij] = sv * hij;
            H[a_ji] = sv * hij;
            for (int k = 0; k < 3; ++k) {
              D[k * n2 + a_ij] = di[k];
              D[k * n2 + a_ji] = dj[k];
            }
            for (int k = 0; k < 6; ++k) {
              Q[k * n2 + a_ij] = qi[k];
              Q[k * n2 + a_ji] = qj[k];
            }
          }
      } else {
        for (int iao = 0; iao < si; ++iao)
          for (int jao = 0; jao < sj; ++jao) {
            const int ij = jao + sj * iao;
            const int64_t a_ij = (int64_t)(ii + iao) * n + (jj + jao);
            S[a_ij] = st[ij];
            H[a_ij] = (ish == jsh && iao == jao) ? se[is + ish] : 0.0;
            for (int k = 0; k < 3; ++k) D[k * n2 + a_ij] = dt[3 * ij + k];
            for (int k = 0; k < 6; ++k) Q[k * n2 + a_ij] = qt[6 * ij + k];
          }
      }
    }
  }
}

// Gamma-point integrals of periodic cells, one workgroup per atom pair
// (i, j <= i) of a cell (work item w): its neighbour-list entries (all images
// of the pair, contiguous in i's list) are spread over the threads, each
// image's shell blocks (vec = R_i - R_j - T, src/h0.cu:359 text) are summed
// into an LDS copy of the pair block with LDS atomics, and the block is
// stored once with plain stores (no global atomics, each element written by
// one workgroup).  The pair (i, i) adds the on-site blocks and the atom's
// images without a mirrored copy; every other pair also stores the mirrored
// block, whose dipole and quadrupole elements are summed per image with the
// operator moved to the other centre.
#define PBC_INT_THREADS 64
template <int LMAX>
__global__ __launch_bounds__(PBC_INT_THREADS) void k_pbc_integrals(BatchDev b, const int* __restrict__ work_mol,
                                                                  const int* __restrict__ work_i,
                                                                  const int* __restrict__ work_j)
{
  constexpr int MA = TA_MAX_SHELL * (2 * LMAX + 1);  // AOs per atom, at most
  constexpr int NE = MA * MA;
  // block: c = 0 S, 1 H, 2..4 D, 5..10 Q (row i, column j); 11..13 D and
  // 14..19 Q of the mirrored block
  __shared__ double blk[20 * NE];
  const int w = blockIdx.x, tid = threadIdx.x;
  const int m = work_mol[w], iat = work_i[w], jat = work_j[w];
  const int a0 = b.mol_atom_off[m], nat = b.mol_nat[m];
  const int s0 = b.mol_sh_off[m];
  const int n = b.mol_nao[m];
  const int64_t n2 = (int64_t)n * n;
  const double* __restrict__ xyz = b.xyz + 3 * (int64_t)a0;
  const int* __restrict__ el = b.at_el + a0;
  const DevParams& P = *b.par;
  const double* __restrict__ se = b.selfenergy + s0;
  const double intcut = b.intcut;
  const int ei_ = el[iat], ej_ = el[jat];
  const DevElem& ei = P.el[ei_];
  const DevElem& ej = P.el[ej_];
  const int is = b.at_ish[a0 + iat], js = b.at_ish[a0 + jat];
  const int ii0 = b.at_iao[a0 + iat], jj0 = b.at_iao[a0 + jat];
  int nai = 0, naj = 0;
  for (int s = 0; s < ei.nsh; ++s) nai += 2 * ei.ang[s] + 1;
  for (int s = 0; s < ej.nsh; ++s) naj += 2 * ej.ang[s] + 1;
  const int ne = nai * naj;
  for (int x = tid; x < 20 * ne; x += PBC_INT_THREADS) blk[x] = 0.0;
  // entries of the pair: j's run inside i's list (ascending j, then T)
  const int64_t nl0 = b.mol_nl_off[m];
  const int* __restrict__ jl = b.nl_jat + nl0;
  const int* __restrict__ tl = b.nl_itr + nl0;
  const double* __restrict__ tr = b.trans + 3 * b.mol_tr_off[m];
  const int l0 = b.nl_inl[a0 + iat], l1 = l0 + b.nl_nnl[a0 + iat];
  int lo = l0, hi = l1;  // first entry with jat >= j
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (jl[mid] < jat) lo = mid + 1;
    else hi = mid;
  }
  int e1 = lo, e2 = l1;  // first entry with jat > j
  while (e1 < e2) {
    const int mid = (e1 + e2) >> 1;
    if (jl[mid] <= jat) e1 = mid + 1;
    else e2 = mid;
  }
  __syncthreads();
  constexpr int NB = (2 * LMAX + 1) * (2 * LMAX + 1);
  double st[NB], dt[3 * NB], qt[6 * NB];
  const bool mirror = iat != jat;
  // one item per image, and for the pair (i, i) one more for the on-site block
  const int nitem = (e1 - lo) + (mirror ? 0 : 1);
  for (int it = tid; it < nitem; it += PBC_INT_THREADS) {
    const bool onsite = it == e1 - lo;
    double vec[3];
    for (int k = 0; k < 3; ++k) vec[k] = xyz[3 * iat + k] - xyz[3 * jat + k];
    if (!onsite) {
      const double* t = tr + 3 * tl[lo + it];
      for (int k = 0; k < 3; ++k) vec[k] -= t[k];
    }
    const double r2 = vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2];
    const double rr = sqrt(sqrt(r2) / (ej.atomic_rad + ei.atomic_rad));
    int ioff = 0;
    for (int ish = 0; ish < ei.nsh; ++ish) {
      const int li = ei.ang[ish], si = 2 * li + 1;
      int joff = 0;
      for (int jsh = 0; jsh < ej.nsh; ++jsh) {
        const int lj = ej.ang[jsh], sj = 2 * lj + 1;
        shell_block<LMAX>(li, lj, ej, jsh, ei, ish, r2, vec, intcut, st, dt, qt);
        double hij = 0.0;
        if (!onsite) {
          const double shp = (1.0 + ei.shpoly[ish] * rr) * (1.0 + ej.shpoly[jsh] * rr);
          hij = 0.5 * (se[is + ish] + se[js + jsh]) * P.hscale[ei_][ej_][ish][jsh] * shp;
        }
        for (int iao = 0; iao < si; ++iao)
          for (int jao = 0; jao < sj; ++jao) {
            const int ij = jao + sj * iao;
            const double sv = st[ij];
            const double* di = dt + 3 * ij;
            const double* qi = qt + 6 * ij;
            const int e = (ioff + iao) * naj + (joff + jao);
            if (sv == 0.0 && di[0] == 0.0 && di[1] == 0.0 && di[2] == 0.0 && qi[0] == 0.0 &&
                qi[1] == 0.0 && qi[2] == 0.0 && qi[3] == 0.0 && qi[4] == 0.0 && qi[5] == 0.0)
              continue;  // every primitive pair screened
            atomicAdd(blk + e, sv);
            if (!onsite) atomicAdd(blk + ne + e, sv * hij);
            for (int k = 0; k < 3; ++k) atomicAdd(blk + (2 + k) * ne + e, di[k]);
            for (int k = 0; k < 6; ++k) atomicAdd(blk + (5 + k) * ne + e, qi[k]);
            if (mirror) {
              double dj[3], qj[6];
              dj[0] = di[0] + vec[0] * sv;
              dj[1] = di[1] + vec[1] * sv;
              dj[2] = di[2] + vec[2] * sv;
              double t0 = 2 * vec[0] * di[0] + vec[0] * vec[0] * sv;
              double t2 = 2 * vec[1] * di[1] + vec[1] * vec[1] * sv;
              double t5 = 2 * vec[2] * di[2] + vec[2] * vec[2] * sv;
              double t1 = vec[0] * di[1] + vec[1] * di[0] + vec[0] * vec[1] * sv;
              double t3 = vec[0] * di[2] + vec[2] * di[0] + vec[0] * vec[2] * sv;
              double t4 = vec[1] * di[2] + vec[2] * di[1] + vec[1] * vec[2] * sv;
              double trc = 0.5 * (t0 + t2 + t5);
              qj[0] = qi[0] + 1.5 * t0 - trc;
              qj[1] = qi[1] + 1.5 * t1;
              qj[2] = qi[2] + 1.5 * t2 - trc;
              qj[3] = qi[3] + 1.5 * t3;
              qj[4] = qi[4] + 1.5 * t4;
              qj[5] = qi[5] + 1.5 * t5 - trc;
              for (int k = 0; k < 3; ++k) atomicAdd(blk + (11 + k) * ne + e, dj[k]);
              for (int k = 0; k < 6; ++k) atomicAdd(blk + (14 + k) * ne + e, qj[k]);
            }
          }
        joff += sj;
      }
      ioff += si;
    }
  }
  __syncthreads();
  double* __restrict__ S = b.S + b.mol_mat_off[m];
  double* __restrict__ H = b.H + b.mol_mat_off[m];
  double* __restrict__ D = b.D + 3 * b.mol_mat_off[m];
  double* __restrict__ Q = b.Q + 6 * b.mol_mat_off[m];
  for (int e = tid; e < ne; e += PBC_INT_THREADS) {
    const int a = e / naj, c = e - (e / naj) * naj;
    const int64_t a_ij = (int64_t)(ii0 + a) * n + (jj0 + c);  // fortran (jj+c, ii+a)
    double h = blk[ne + e];
    if (!mirror && a == c) {
      // self energy on the diagonal of the on-site block
      int off = 0, sh = 0;
      while (off + 2 * ei.ang[sh] + 1 <= a) off += 2 * ei.ang[sh++] + 1;
      h += se[is + sh];
    }
    S[a_ij] = blk[e];
    H[a_ij] = h;
    for (int k = 0; k < 3; ++k) D[k * n2 + a_ij] = blk[(2 + k) * ne + e];
    for (int k = 0; k < 6; ++k) Q[k * n2 + a_ij] = blk[(5 + k) * ne + e];
    if (mirror) {
      const int64_t a_ji = (int64_t)(jj0 + c) * n + (ii0 + a);
      S[a_ji] = blk[e];
      H[a_ji] = h;
      for (int k = 0; k < 3; ++k) D[k * n2 + a_ji] = blk[(11 + k) * ne + e];
      for (int k = 0; k < 6; ++k) Q[k * n2 + a_ji] = blk[(14 + k) * ne + e];
    }
  }
}

template __global__ void k_pbc_integrals<1>(BatchDev, const int*, const int*, const int*);
template __global__ void k_pbc_integrals<2>(BatchDev, const int*, const int*, const int*);
template __global__ void k_integrals<1>(BatchDev, const int*, const int*, const int*);
template __global__ void k_integrals<2>(BatchDev, const int*, const int*, const int*);

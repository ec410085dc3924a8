// Interaction planes and three-body dispersion of periodic cells (Gamma
// point).  The reference carries the periodic branches only as comments: the
// Ewald paths of effective_coulomb::update (src/coulomb.cu:94-96) and of the
// multipole update (src/coulomb.cu:554-557, get_multipole_matrix_3d), and the
// lattice translations of the short-range sums (src/coulomb.cu:212-213).
//
//   ES2   A_ab = sum'_T [f(r) (gamma_ab(r) - 1/r) + erfc(alpha r) / r]
//              + (4 pi / V) sum_G exp(-G^2 / 4 alpha^2) / G^2 cos(G . R)
//              - pi / (V alpha^2) + [A == B] (eta_ab - 2 alpha / sqrt(pi))
//         r = |R - T|, R = R_A - R_B, f the switch of the Klopman-Ohno kernel
//         to 1/r between TA_PBC_ES2_RON and TA_PBC_ES2_ROFF (without it the
//         lattice sum of gamma - 1/r ~ -1/(2 eta^2 r^3) diverges)
//   AES2  sd = -grad phi, dd = -hess phi, sq = hess phi / 3 of the Ewald sum
//         phi of 1/r (tin-foil boundary), plus the damping corrections
//         (f3 - 1), (f5 - 1) of the molecular kernels over the images within
//         TA_PBC_AES_CUT; an atom and its own images interact (self term
//         4 alpha^3 / (3 sqrt(pi)) of the Hessian)
//   D4    two-body damping factors summed over the images; three-body term
//         over the triangles of sites (atom images) within the three-body
//         cutoff of each other, each class of translated triangles counted
//         once: anchored at a site of its lowest atom index i (sites of atoms
//         >= i), weighted by 1 / (number of its sites with atom index i)
//
// Launch shapes: k_pbc_pairs one thread per atom pair, work items of atom
// rows of one cell; k_pbc_atm_sites one workgroup per atom of a cell (block
// scan keeps the sites in (atom, translation) order); k_pbc_atm work items of
// site rows of one anchor atom, threads over the second site.
#include <hip/hip_runtime.h>
#include <math.h>

#include "common.h"
#include "d4.h"
#include "device_util.h"

#define PBC_THREADS 256

__device__ __forceinline__ double es2_switch(double r)
{
  if (r <= TA_PBC_ES2_RON) return 1.0;
  if (r >= TA_PBC_ES2_ROFF) return 0.0;
  const double x = (r - TA_PBC_ES2_RON) / (TA_PBC_ES2_ROFF - TA_PBC_ES2_RON);
  return 1.0 - x * x * x * (10.0 + x * (-15.0 + 6.0 * x));
}

// ES2 row elements, AES2 planes, D4 damping factors and sqrt(|C6_ij(q = 0)|)
// (in d4_c60: the only reader for a cell is k_pbc_atm) of the atom pairs
// (i, j) with i in [a_begin, a_end) of cell m (work item w).
__global__ __launch_bounds__(PBC_THREADS) void k_pbc_pairs(BatchDev b, const int* __restrict__ work_mol,
                                                           const int* __restrict__ work_a0,
                                                           const int* __restrict__ work_a1)
{
  const int w = blockIdx.x;
  const int m = work_mol[w], ab = work_a0[w], ae = work_a1[w];
  const int a0 = b.mol_atom_off[m], nat = b.mol_nat[m];
  const int s0 = b.mol_sh_off[m], nsh = b.mol_nsh[m];
  const double* __restrict__ xyz = b.xyz + 3 * (int64_t)a0;
  const int* __restrict__ el = b.at_el + a0;
  const DevParams& P = *b.par;
  const DevD4& d4 = P.d4;
  const int ntr = b.mol_ntr[m], ng = b.mol_ng[m];
  const double* __restrict__ tr = b.trans + 3 * b.mol_tr_off[m];
  const double* __restrict__ gv = b.recip + 3 * b.mol_g_off[m];
  const double* cst = b.mol_ewald + TA_PBC_EWALD * m;
  const double vol = cst[TA_PBC_VOL], al = cst[TA_PBC_ALPHA], rreal = cst[TA_PBC_RREAL];
  const double fac = 4.0 * M_PI / vol, a4 = 4.0 * al * al;
  const double tpi = 2.0 * al / sqrt(M_PI);
  const double rcut_es = fmax(TA_PBC_ES2_ROFF, rreal);
  const double es_cut2 = rcut_es * rcut_es, rr2 = rreal * rreal;
  const double ra2 = TA_PBC_AES_CUT * TA_PBC_AES_CUT;
  const double d2_cut2 = d4.disp2_cutoff * d4.disp2_cutoff;
  // every term vanishes beyond the largest of its cutoffs
  const double rmax = fmax(fmax(rcut_es, rreal), fmax(TA_PBC_AES_CUT, d4.enabled ? d4.disp2_cutoff : 0.0));
  const double gexp = P.gexp;
  const double* __restrict__ mrad = b.mrad + a0;
  const int64_t n2 = (int64_t)nat * nat;
  double* __restrict__ gam = b.gam + b.mol_gam_off[m];
  double* __restrict__ sd = b.aes_sd + 3 * b.mol_pair_off[m];
  double* __restrict__ dd = b.aes_dd + 9 * b.mol_pair_off[m];
  double* __restrict__ sq = b.aes_sq + 6 * b.mol_pair_off[m];

  for (int64_t idx = (int64_t)ab * nat + threadIdx.x; idx < (int64_t)ae * nat; idx += blockDim.x) {
    const int i = (int)(idx / nat), j = (int)(idx - (int64_t)i * nat);
    const DevElem& ei = P.el[el[i]];
    const DevElem& ej = P.el[el[j]];
    const int ii = b.at_ish[a0 + i], jj = b.at_ish[a0 + j];
    const int nsi = ei.nsh, nsj = ej.nsh;
    double R[3];
    for (int x = 0; x < 3; ++x) R[x] = xyz[3 * i + x] - xyz[3 * j + x];
    // reciprocal space: phi, grad, Hessian (xx xy xz yx yy yz zx zy zz)
    double phi = 0.0, g[3] = {0.0, 0.0, 0.0};
    double h[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < ng; ++k) {
      const double gx = gv[3 * k], gy = gv[3 * k + 1], gz = gv[3 * k + 2];
      const double g2 = gx * gx + gy * gy + gz * gz;
      const double wk = fac * exp(-g2 / a4) / g2;
      double sn, cs;
      sincos(gx * R[0] + gy * R[1] + gz * R[2], &sn, &cs);
      phi += wk * cs;
      const double G[3] = {gx, gy, gz};
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        g[x] -= wk * sn * G[x];
#pragma unroll
        for (int y = 0; y < 3; ++y) h[3 * x + y] -= wk * cs * G[x] * G[y];
      }
    }
    phi -= M_PI / (vol * al * al);
    // real space over the images
    double es[TA_MAX_SHELL * TA_MAX_SHELL], etag[TA_MAX_SHELL * TA_MAX_SHELL];
    for (int p = 0; p < TA_MAX_SHELL * TA_MAX_SHELL; ++p) {
      es[p] = 0.0;
      etag[p] = 0.0;
    }
    for (int ish = 0; ish < nsi; ++ish)
      for (int jsh = 0; jsh < nsj; ++jsh) {
        const double eta = 0.5 * (ei.hardness[ish] + ej.hardness[jsh]);
        etag[ish * TA_MAX_SHELL + jsh] = pow(eta, -gexp);
      }
    double csd[3] = {0.0, 0.0, 0.0};
    double cdd[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double csq[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double dv = 0.0;
    const double mr = 0.5 * (mrad[j] + mrad[i]);
    int blo[3], bhi[3];
    cell_box(cst, R, rmax, blo, bhi);
    for (int n1 = blo[0]; n1 <= bhi[0]; ++n1)
    for (int n2 = blo[1]; n2 <= bhi[1]; ++n2)
    for (int n3 = blo[2]; n3 <= bhi[2]; ++n3) {
      const int t = cell_tindex(cst, n1, n2, n3);
      double v[3];
      for (int x = 0; x < 3; ++x) v[x] = R[x] - tr[3 * t + x];
      const double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
      if (r2 < 1.0e-12) continue;
      const double r1 = sqrt(r2);
      if (r2 <= es_cut2) {
        const double ec = erfc(al * r1) / r1;
        const double f = es2_switch(r1);
        for (int ish = 0; ish < nsi; ++ish)
          for (int jsh = 0; jsh < nsj; ++jsh) {
            const int p = ish * TA_MAX_SHELL + jsh;
            double term = ec;
            if (f > 0.0) term += f * (1.0 / pow(pow(r1, gexp) + etag[p], 1.0 / gexp) - 1.0 / r1);
            es[p] += term;
          }
      }
      if (r2 <= rr2) {
        const double ex = exp(-al * al * r2), ec = erfc(al * r1) / r1;
        const double B = (ec + tpi * ex) / r2;
        const double C = (3.0 * ec + tpi * (3.0 + 2.0 * al * al * r2) * ex) / (r2 * r2);
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          g[x] -= v[x] * B;
#pragma unroll
          for (int y = 0; y < 3; ++y) h[3 * x + y] += v[x] * v[y] * C - (x == y ? B : 0.0);
        }
      }
      if (r2 <= ra2) {
        const double g1 = 1.0 / r1, g3 = g1 * g1 * g1, g5 = g3 * g1 * g1;
        const double rr = mr * g1;
        const double f3 = 1.0 / (1.0 + 6.0 * pow(rr, P.dmp3)) - 1.0;
        const double f5 = 1.0 / (1.0 + 6.0 * pow(rr, P.dmp5)) - 1.0;
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          csd[x] += v[x] * g3 * f3;
#pragma unroll
          for (int y = 0; y < 3; ++y) {
            cdd[3 * x + y] += ((x == y ? g3 : 0.0) - 3.0 * v[x] * v[y] * g5) * f5;
            csq[3 * x + y] += v[x] * v[y] * g5 * f5;
          }
        }
      }
      if (d4.enabled && r2 <= d2_cut2) {
        double dmp, ddr;
        d4_damp(d4, el[i], el[j], r2, dmp, ddr);
        dv += dmp;
      }
    }
    for (int ish = 0; ish < nsi; ++ish)
      for (int jsh = 0; jsh < nsj; ++jsh) {
        double a = es[ish * TA_MAX_SHELL + jsh] + phi;
        if (i == j) a += 0.5 * (ei.hardness[ish] + ej.hardness[jsh]) - tpi;
        gam[(int64_t)(ii + ish) * nsh + (jj + jsh)] = a;
      }
    if (i == j)
      for (int x = 0; x < 3; ++x) h[4 * x] += 2.0 * al * al * tpi / 3.0;
    for (int k = 0; k < 3; ++k) sd[k * n2 + idx] = -g[k] + csd[k];
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l < 3; ++l) dd[(k * 3 + l) * n2 + idx] = -h[3 * k + l] + cdd[3 * k + l];
    double mq[9];
    for (int k = 0; k < 9; ++k) mq[k] = h[k] / 3.0 + csq[k];
    sq[0 * n2 + idx] = mq[0];
    sq[1 * n2 + idx] = 2.0 * mq[1];
    sq[2 * n2 + idx] = mq[4];
    sq[3 * n2 + idx] = 2.0 * mq[2];
    sq[4 * n2 + idx] = 2.0 * mq[5];
    sq[5 * n2 + idx] = mq[8];
    if (d4.enabled) {
      const double* __restrict__ gw0 = b.d4_gw + TA_D4_MAXREF * (int64_t)a0;
      double acc[TA_D4_MAXREF] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      d4_accumulate(d4, el[i], el[j], 1.0, gw0 + TA_D4_MAXREF * j, acc);
      double cv = 0.0;
      for (int r = 0; r < d4.nref[el[i]]; ++r) cv += gw0[TA_D4_MAXREF * i + r] * acc[r];
      b.d4_disp[b.mol_pair_off[m] + idx] = d4.qref_uniform ? dv * cv : dv;
      // a cell's three-body term takes the product of three square roots
      b.d4_c60[b.mol_pair_off[m] + idx] = sqrt(fabs(cv));
    }
  }
}

// Three-body sites of atom i of cell m (work item w): every image (j, T) of
// an atom j >= i with 1e-12 < |R_j - T - R_i|^2 <= cutoff^2, in (j, T)
// order, stored as (R_j - T - R_i, j) up to the host-computed capacity.
__global__ __launch_bounds__(PBC_THREADS) void k_pbc_atm_sites(BatchDev b, const int* __restrict__ work_mol,
                                                               const int* __restrict__ work_atom)
{
  __shared__ int red[PBC_THREADS / TA_WAVE];
  const int w = blockIdx.x, tid = threadIdx.x;
  const int m = work_mol[w], i = work_atom[w];
  const int a0 = b.mol_atom_off[m], nat = b.mol_nat[m];
  const double* __restrict__ xyz = b.xyz + 3 * (int64_t)a0;
  const double* __restrict__ tr = b.trans + 3 * b.mol_tr_off[m];
  const double* cst = b.mol_ewald + TA_PBC_EWALD * m;
  const double cut = b.par->d4.disp3_cutoff, cut2 = cut * cut;
  const int64_t off = b.atm_site_off[a0 + i];
  const int cap = b.atm_site_cap[a0 + i];
  double* __restrict__ site = b.atm_site + 4 * off;
  int base = 0;
  for (int j = i; j < nat; ++j) {
    // the box of translations around R_j - R_i, walked in table order
    const double d[3] = {xyz[3 * j] - xyz[3 * i], xyz[3 * j + 1] - xyz[3 * i + 1],
                         xyz[3 * j + 2] - xyz[3 * i + 2]};
    int lo[3], hi[3];
    cell_box(cst, d, cut, lo, hi);
    const int w1 = hi[1] - lo[1] + 1, w2 = hi[2] - lo[2] + 1;
    const int nb = (hi[0] - lo[0] + 1) * w1 * w2;
    for (int c0 = 0; c0 < nb; c0 += PBC_THREADS) {
      const int c = c0 + tid;
      double v[3] = {0.0, 0.0, 0.0};
      int keep = 0;
      if (c < nb) {
        const int q = c / w2;
        const int n1 = lo[0] + q / w1, n2 = lo[1] + q % w1, n3 = lo[2] + c % w2;
        const int t = cell_tindex(cst, n1, n2, n3);
        for (int x = 0; x < 3; ++x) v[x] = d[x] - tr[3 * t + x];
        const double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        keep = (r2 >= 1.0e-12 && r2 <= cut2) ? 1 : 0;
      }
      int tot;
      const int pos = base + block_exclusive_scan(keep, red, &tot);
      if (keep && pos < cap) {
        site[4 * pos] = v[0];
        site[4 * pos + 1] = v[1];
        site[4 * pos + 2] = v[2];
        site[4 * pos + 3] = (double)j;
      }
      base += tot;
    }
  }
  if (tid == 0) b.atm_site_n[a0 + i] = base < cap ? base : cap;
}

// Three-body energy of anchor i of cell m over the site pairs (a, b > a) with
// a in [r0, r1) (work item w), each triangle weighted by one over its number
// of sites of atom i (the anchor and images of it), added to d4_e3 of atom i.
// The damping exponent alp / 3 = 16 / 3 of GFN2's D4 is formed as x^5 cbrt(x).
__device__ __forceinline__ double atm_damp_pow(double x, double p3, bool p16)
{
  if (p16) {
    const double x2 = x * x;
    return x2 * x2 * x * cbrt(x);
  }
  return pow(x, p3);
}

__global__ __launch_bounds__(PBC_THREADS) void k_pbc_atm(BatchDev b, const int* __restrict__ work_mol,
                                                         const int* __restrict__ work_atom,
                                                         const int* __restrict__ work_r0,
                                                         const int* __restrict__ work_r1)
{
  __shared__ double red[PBC_THREADS / TA_WAVE];
  __shared__ double r0t[TA_MAX_ELEM * TA_MAX_ELEM];  // R0 of every element pair
  const int w = blockIdx.x, tid = threadIdx.x;
  const int m = work_mol[w], i = work_atom[w];
  const int a0 = b.mol_atom_off[m], nat = b.mol_nat[m];
  const int n = b.atm_site_n[a0 + i];
  const int r0 = work_r0[w], r1 = min(work_r1[w], n);
  const double* __restrict__ site = b.atm_site + 4 * b.atm_site_off[a0 + i];
  const int* __restrict__ el = b.at_el + a0;
  const double* __restrict__ sc6 = b.d4_c60 + b.mol_pair_off[m];  // sqrt(|C6(0)|)
  const DevD4& d = b.par->d4;
  const int ne = b.par->nelem;
  for (int x = tid; x < ne * ne; x += PBC_THREADS) {
    const int e1 = x / ne, e2 = x - (x / ne) * ne;
    r0t[e1 * TA_MAX_ELEM + e2] = d.a1 * sqrt(3.0 * d.r4r2[e1] * d.r4r2[e2]) + d.a2;
  }
  __syncthreads();
  const double cut2 = d.disp3_cutoff * d.disp3_cutoff;
  const double p3 = d.alp / 3.0;
  const bool p16 = d.alp == 16.0;
  const int ei = el[i];
  const double* __restrict__ sci = sc6 + (int64_t)i * nat;
  double e = 0.0;
  for (int a = r0; a < r1; ++a) {
    const double* sa = site + 4 * a;
    const int j = (int)sa[3], ej = el[j];
    const double r2ij = sa[0] * sa[0] + sa[1] * sa[1] + sa[2] * sa[2];
    const double r0ij = r0t[ei * TA_MAX_ELEM + ej];
    const double cij = d.s9 * sci[j];
    const double* __restrict__ scj = sc6 + (int64_t)j * nat;
    const double wa = (j == i) ? 1.0 : 0.0;
    for (int bb = a + 1 + tid; bb < n; bb += PBC_THREADS) {
      const double* sb = site + 4 * bb;
      const double dx = sb[0] - sa[0], dy = sb[1] - sa[1], dz = sb[2] - sa[2];
      const double r2jk = dx * dx + dy * dy + dz * dz;
      if (r2jk < 1.0e-12 || r2jk > cut2) continue;
      const int k = (int)sb[3], ek = el[k];
      const double r2ik = sb[0] * sb[0] + sb[1] * sb[1] + sb[2] * sb[2];
      const double c9 = cij * sci[k] * scj[k];
      const double r2ijk = r2ij * r2ik * r2jk;
      const double irijk = 1.0 / sqrt(r2ijk), ir2 = irijk * irijk;
      const double x = r0ij * r0t[ei * TA_MAX_ELEM + ek] * r0t[ej * TA_MAX_ELEM + ek] * irijk;
      const double ang = 0.375 * (r2ij + r2jk - r2ik) * (r2ij - r2jk + r2ik) *
                             (-r2ij + r2jk + r2ik) * ir2 + 1.0;
      const double wt = (k == i ? 2.0 : 1.0) + wa;  // sites of atom i in the triangle
      e += c9 * ang * irijk * ir2 / (wt * (1.0 + 6.0 * atm_damp_pow(x, p3, p16)));
    }
  }
  e = block_sum(e, red);
  if (tid == 0 && e != 0.0) atomicAdd(b.d4_e3 + a0 + i, e);
}

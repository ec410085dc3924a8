// Host side of the engine: calculator construction from parameters, batch
// planning (per-molecule offsets into flat device buffers), the device
// workspace, kernel launches, and the C ABI declared in include/tblite_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <atomic>
#include <thread>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../params/gfn2_default.h"
#include "../params/sto_ng_table.h"
#include "common.h"
#include "scf_grid.h"
#include "tblite_amd.h"

// kernels (setup.hip / scf.hip)
__global__ void k_fill(double* p, int64_t n, double v);
__global__ void k_setup(BatchDev b);
__global__ void k_adjlist(BatchDev b);
__global__ void k_d4_atm(BatchDev b, const int* work_mol, const int* work_i0, const int* work_i1);
__global__ void k_setup_pairs(BatchDev b, const int* work_mol, const int* work_a0, const int* work_a1);
template <int LMAX>
__global__ void k_integrals(BatchDev b, const int* work_mol, const int* work_kind,
                            const int* work_start);
template <int LMAX>
__global__ void k_pbc_integrals(BatchDev b, const int* work_mol, const int* work_i, const int* work_j);
// periodic cells (periodic.hip)
__global__ void k_pbc_pairs(BatchDev b, const int* work_mol, const int* work_a0, const int* work_a1);
__global__ void k_pbc_atm_sites(BatchDev b, const int* work_mol, const int* work_atom);
__global__ void k_pbc_atm(BatchDev b, const int* work_mol, const int* work_atom, const int* work_r0,
                          const int* work_r1);
template <int NT, int MAXT, int MAXB, int MAXC, bool MATS_LDS, bool CELL>
__global__ void k_scf(BatchDev b, const int* mols, int ldmax, int natmax, int nshmax);
// analytic gradient (gradient.hip)
__global__ void k_grad_pot(BatchDev b);
__global__ void k_grad_wdens(BatchDev b, const int* tiles);
template <int LMAX>
__global__ void k_grad_h0(BatchDev b, const int* work_mol, const int* work_kind,
                          const int* work_start);
__global__ void k_grad_pairs(BatchDev b, const int* work_mol, const int* work_a0, const int* work_a1);
__global__ void k_grad_c6d(BatchDev b, const int* work_mol, const int* work_a0, const int* work_a1);
__global__ void k_grad_atm(BatchDev b, const int* work_mol, const int* work_i0, const int* work_i1,
                           int lds);
__global__ void k_grad_cn(BatchDev b, const int* work_mol, const int* work_a0, const int* work_a1);

namespace {

// work-tile defaults (tblite_amd_tuning: integral_chunk, pair_chunk_log2,
// atm_chunk_log2): neighbour-list entries per integral work item (its
// workgroup size, at most the kernels' launch bound of 128), atom pairs per
// pair-plane work item, triples per three-body work item
constexpr int kIntegralChunk = 128;
constexpr int kAtmChunkLog2 = 18;
constexpr int kPairChunkLog2 = 15;
constexpr int kLdsMaxLd = 96;
constexpr int kNumStreams = 4;
constexpr int kMaxBuckets = 16;  // size buckets of molecules plus those of periodic cells
constexpr int kGridMaxLd = 12288;  // largest padded AO dimension of the multi-workgroup path

// ------------------------------------------------------------ calculator
struct Calculator {
  tblite_amd_param param;
  DevParams dev;
  double min_alpha_el[TA_MAX_ELEM];
  int zmap[120];
  double accuracy = 1.0;
  int max_iter = 250;
  double damping = 0.4;
  double etemp = 300.0;
  DevParams* d_par = nullptr;
};

int sto_shell_index(int n, int l)
{
  if (n == 1 && l == 0) return 0;
  if (n == 2 && l == 0) return 1;
  if (n == 2 && l == 1) return 2;
  if (n == 3 && l == 0) return 3;
  if (n == 3 && l == 1) return 4;
  if (n == 3 && l == 2) return 5;
  return -1;
}

// Build the device parameter block: STO-nG contraction of every shell
// (exponents alpha * zeta * zeta, primitive norms (2a/pi)^(3/4) (4a)^(l/2) /
// sqrt((2l-1)!!)), levels and CN shifts in Eh, shell hardness, third-order
// derivatives, and the H0 scaling table hscale (zeta term, EN term,
// shell-pair factor) of tblite's GFN2 specification.
bool build_calculator(Calculator& c, std::string& err)
{
  const tblite_amd_param& p = c.param;
  if (p.nelem < 1 || p.nelem > TA_MAX_ELEM) {
    err = "parameter set has an invalid element count";
    return false;
  }
  std::memset(&c.dev, 0, sizeof(c.dev));
  for (int z = 0; z < 120; ++z) c.zmap[z] = -1;
  c.dev.nelem = p.nelem;
  const double dfact[3] = {1.0, 1.0, 3.0};
  for (int e = 0; e < p.nelem; ++e) {
    const tblite_amd_element& el = p.elem[e];
    DevElem& d = c.dev.el[e];
    if (el.z <= 0 || el.z >= 120 || el.nsh < 1 || el.nsh > TA_MAX_SHELL) {
      err = "invalid element record in parameter set";
      return false;
    }
    c.zmap[el.z] = e;
    d.z = el.z;
    d.nsh = el.nsh;
    c.min_alpha_el[e] = INFINITY;
    for (int s = 0; s < el.nsh; ++s) {
      const int l = el.ang[s];
      const int ng = el.ngauss[s];
      const int sh = sto_shell_index(el.pqn[s], l);
      if (sh < 0 || ng < 1 || ng > STO_NG_MAXG || l < 0 || l > 2) {
        err = "unsupported shell (n, l, ngauss) in parameter set";
        return false;
      }
      d.ang[s] = l;
      d.nprim[s] = ng;
      const double zeta = el.slater[s];
      for (int i = 0; i < ng; ++i) {
        const double a = sto_ng_alpha[sh][ng - 1][i] * zeta * zeta;
        if (!(a > 0.0)) {
          err = "STO-nG table has no fit for this shell";
          return false;
        }
        d.alpha[s][i] = a;
        d.coeff[s][i] = sto_ng_coeff[sh][ng - 1][i] * std::pow(2.0 * a / M_PI, 0.75) *
                        std::pow(4.0 * a, 0.5 * l) / std::sqrt(dfact[l]);
        c.min_alpha_el[e] = std::min(c.min_alpha_el[e], a);
      }
      d.selfenergy[s] = el.levels[s] / TBLITE_AMD_AUTOEV;
      d.kcn[s] = el.kcn[s] / TBLITE_AMD_AUTOEV;
      d.hardness[s] = el.gam * el.lgam[s];
      d.hubbard_derivs[s] = el.gam3 * p.thirdorder_shell[l];
      d.refocc[s] = el.refocc[s];
      d.shpoly[s] = el.shpoly[s];
    }
    d.rcov = el.rcov;
    d.atomic_rad = el.atomic_rad;
    d.zeff = el.zeff;
    d.arep = el.arep;
    d.dkernel = el.dkernel;
    d.qkernel = el.qkernel;
    d.mprad = el.mprad;
    d.mpvcn = el.mpvcn;
    d.en = el.en;
  }
  for (int ei = 0; ei < p.nelem; ++ei)
    for (int ej = 0; ej < p.nelem; ++ej) {
      const tblite_amd_element& a = p.elem[ei];
      const tblite_amd_element& b = p.elem[ej];
      const double den = (a.en - b.en) * (a.en - b.en);
      for (int is = 0; is < a.nsh; ++is)
        for (int js = 0; js < b.nsh; ++js) {
          const double zi = a.slater[is], zj = b.slater[js];
          const double zij = std::pow(2.0 * std::sqrt(zi * zj) / (zi + zj), p.wexp);
          c.dev.hscale[ei][ej][is][js] =
              zij * (1.0 + p.enscale * den) * p.kshell[b.ang[js]][a.ang[is]];
        }
    }
  c.dev.rep_kexp = p.rep_kexp;
  c.dev.rep_kexp_light = p.rep_kexp_light;
  c.dev.rep_rexp = p.rep_rexp;
  c.dev.rep_cutoff = p.rep_cutoff;
  c.dev.gexp = p.gexp;
  c.dev.dmp3 = p.dmp3;
  c.dev.dmp5 = p.dmp5;
  c.dev.mp_shift = p.mp_shift;
  c.dev.mp_kexp = p.mp_kexp;
  c.dev.mp_rmax = p.mp_rmax;
  c.dev.cn_ka = p.cn_ka;
  c.dev.cn_kb = p.cn_kb;
  c.dev.cn_rshift = p.cn_rshift;
  c.dev.cn_cutoff = p.cn_cutoff;

  // D4: device copy of the damping / weighting constants and the reference
  // C6 of every element and reference pair, C6 = 3/pi int alpha_A alpha_B dw
  // (trapezoid rule on the model's imaginary-frequency grid)
  const tblite_amd_d4_param& d4 = p.d4;
  DevD4& dd = c.dev.d4;
  dd.enabled = d4.enabled ? 1 : 0;
  if (dd.enabled) {
    dd.s6 = d4.s6; dd.s8 = d4.s8; dd.a1 = d4.a1; dd.a2 = d4.a2;
    dd.s9 = d4.s9; dd.alp = d4.alp; dd.ga = d4.ga; dd.wf = d4.wf;
    dd.k4 = d4.cn_k4; dd.k5 = d4.cn_k5; dd.k6 = d4.cn_k6; dd.kn = d4.cn_kn;
    dd.cn_cutoff = d4.cn_cutoff;
    dd.disp2_cutoff = d4.disp2_cutoff;
    dd.disp3_cutoff = d4.disp3_cutoff;
    double w[TBLITE_AMD_D4_NFREQ];
    for (int k = 0; k < TBLITE_AMD_D4_NFREQ; ++k) {
      const double* f = tblite_amd_d4_freq;
      const double left = k > 0 ? f[k] - f[k - 1] : 0.0;
      const double right = k + 1 < TBLITE_AMD_D4_NFREQ ? f[k + 1] - f[k] : 0.0;
      w[k] = 0.5 * (left + right);
    }
    for (int e = 0; e < p.nelem; ++e) {
      const tblite_amd_d4_element& de = d4.elem[e];
      if (de.nref < 1 || de.nref > TA_D4_MAXREF || !(de.zeff > 0.0)) {
        err = "D4 data missing or invalid for an element (set d4.enabled = 0 to run without dispersion)";
        return false;
      }
      dd.nref[e] = de.nref;
      for (int r = 0; r < de.nref; ++r) {
        dd.cnref[e][r] = de.cn[r];
        dd.qref[e][r] = de.q[r];
      }
      dd.zeff[e] = de.zeff;
      dd.gi[e] = d4.gc * de.eta;
      dd.r4r2[e] = de.r4r2;
    }
    dd.qref_uniform = 1;
    for (int e = 0; e < p.nelem; ++e)
      for (int r = 1; r < d4.elem[e].nref; ++r)
        if (d4.elem[e].q[r] != d4.elem[e].q[0]) dd.qref_uniform = 0;
    for (int ei = 0; ei < p.nelem; ++ei)
      for (int ej = 0; ej < p.nelem; ++ej)
        for (int r = 0; r < d4.elem[ei].nref; ++r)
          for (int s = 0; s < d4.elem[ej].nref; ++s) {
            double acc = 0.0;
            for (int k = 0; k < TBLITE_AMD_D4_NFREQ; ++k)
              acc += w[k] * d4.elem[ei].alpha[r][k] * d4.elem[ej].alpha[s][k];
            dd.c6[ei][ej][r][s] = 3.0 / M_PI * acc;
          }
  }
  return true;
}

// integral_cutoff / get_cutoff, src/basis/type.cu:4-64 (text)
double integral_cutoff(double acc)
{
  auto clip = [](double v, double lo, double hi) { return std::fmin(std::fmax(v, lo), hi); };
  const double min_intcut = 5.0, max_intcut = 25.0, max_acc = 1.0e-4, min_acc = 1.0e+3;
  return clip(max_intcut - 10.0 * std::log10(clip(acc, min_acc, max_acc)), min_intcut, max_intcut);
}

double get_cutoff(double min_alpha, double acc)
{
  return std::fmin(std::sqrt(2.0 * integral_cutoff(acc) / min_alpha), 40.0);
}

// ------------------------------------------------------- periodic cells
// Lattice tables of one cell (a_k = lat[3k .. 3k+2], bohr): the translations
// of every real-space sum and the reciprocal vectors of the Ewald sums, with
// the constants listed at TA_PBC_RMIN in common.h.  The translations are
// formed with a fixed operation order, T_x = (n1 a1x + n2 a2x) + n3 a3x, so
// the neighbour list's translation indices and distances are reproducible
// on any host.  Returns false for a degenerate lattice.
struct CellTables {
  double vol = 0.0, alpha = 0.0, rreal = 0.0;
  int nrep[3] = {0, 0, 0};          // translation index ranges [-N_k, N_k]
  double rec[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, dk[3] = {0, 0, 0};  // b_k, plane spacings
  std::vector<double> tr, g;
  // the per-cell constant block of the device (TA_PBC_EWALD doubles)
  void constants(double* cst) const
  {
    for (int k = 0; k < TA_PBC_EWALD; ++k) cst[k] = 0.0;
    cst[TA_PBC_VOL] = vol;
    cst[TA_PBC_ALPHA] = alpha;
    cst[TA_PBC_RREAL] = rreal;
    for (int k = 0; k < 9; ++k) cst[TA_PBC_REC + k] = rec[k];
    for (int k = 0; k < 3; ++k) {
      cst[TA_PBC_DK + k] = dk[k];
      cst[TA_PBC_NREP + k] = nrep[k];
    }
  }
};

bool cell_tables(const double* lat, int nat, const double* xyz, CellTables& ct)
{
#pragma clang fp contract(off)
  const double *a1 = lat, *a2 = lat + 3, *a3 = lat + 6;
  auto cross = [](const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
  };
  double c[9], rec[9];
  cross(a2, a3, c);
  cross(a3, a1, c + 3);
  cross(a1, a2, c + 6);
  const double det = a1[0] * c[0] + a1[1] * c[1] + a1[2] * c[2];
  if (!(std::fabs(det) > 1.0e-6)) return false;
  const double pi = 3.14159265358979323846;
  ct.vol = std::fabs(det);
  for (int k = 0; k < 9; ++k) rec[k] = 2.0 * pi * c[k] / det;
  ct.alpha = std::sqrt(pi) / std::cbrt(ct.vol);
  ct.rreal = TA_PBC_EWALD_X / ct.alpha;
  const double gmax = 2.0 * TA_PBC_EWALD_X * ct.alpha;
  const double rmax = std::fmax(TA_PBC_RMIN, ct.rreal);
  int* nrep = ct.nrep;
  for (int k = 0; k < 9; ++k) ct.rec[k] = rec[k];
  for (int k = 0; k < 3; ++k) {
    const double* ck = c + 3 * k;
    const double dk = ct.vol / std::sqrt(ck[0] * ck[0] + ck[1] * ck[1] + ck[2] * ck[2]);
    ct.dk[k] = dk;
    double fmn = 0.0, fmx = 0.0;
    for (int i = 0; i < nat; ++i) {
      const double f = (xyz[3 * i] * rec[3 * k] + xyz[3 * i + 1] * rec[3 * k + 1] +
                        xyz[3 * i + 2] * rec[3 * k + 2]) / (2.0 * pi);
      if (i == 0 || f < fmn) fmn = f;
      if (i == 0 || f > fmx) fmx = f;
    }
    nrep[k] = (int)std::ceil(rmax / dk + (fmx - fmn));
  }
  ct.tr.clear();
  for (int n1 = -nrep[0]; n1 <= nrep[0]; ++n1)
    for (int n2 = -nrep[1]; n2 <= nrep[1]; ++n2)
      for (int n3 = -nrep[2]; n3 <= nrep[2]; ++n3)
        for (int x = 0; x < 3; ++x)
          ct.tr.push_back(((double)n1 * a1[x] + (double)n2 * a2[x]) + (double)n3 * a3[x]);
  int mrep[3];
  for (int k = 0; k < 3; ++k) {
    const double* ak = lat + 3 * k;
    mrep[k] = (int)std::ceil(gmax * std::sqrt(ak[0] * ak[0] + ak[1] * ak[1] + ak[2] * ak[2]) /
                             (2.0 * pi));
  }
  ct.g.clear();
  for (int m1 = -mrep[0]; m1 <= mrep[0]; ++m1)
    for (int m2 = -mrep[1]; m2 <= mrep[1]; ++m2)
      for (int m3 = -mrep[2]; m3 <= mrep[2]; ++m3) {
        if (m1 == 0 && m2 == 0 && m3 == 0) continue;
        double gv[3];
        for (int x = 0; x < 3; ++x)
          gv[x] = ((double)m1 * rec[x] + (double)m2 * rec[3 + x]) + (double)m3 * rec[6 + x];
        if (gv[0] * gv[0] + gv[1] * gv[1] + gv[2] * gv[2] > gmax * gmax) continue;
        ct.g.insert(ct.g.end(), gv, gv + 3);
      }
  return true;
}

// Upper bounds of a cell's neighbour-list length (pairs j <= i with an image
// within cutoff) and of each atom's three-body site count (images of atoms
// j >= i within cut3), with a relative slack of 1e-9 so that rounding on the
// device can only give fewer entries.
// Only the translations of the box n_k in [f_k - c / d_k, f_k + c / d_k]
// around a pair's fractional separation f can lie within c of it
// (|R - T| >= d_k |f_k - n_k|), so each pair scans that box, not the table.
void cell_counts(int nat, const double* xyz, const CellTables& ct, double cutoff, double cut3,
                 int64_t& nl, std::vector<int>& sites)
{
  const std::vector<double>& tr = ct.tr;
  const double c2 = cutoff * cutoff * (1.0 + 1.0e-9) + 1.0e-9;
  const double s2 = cut3 * cut3 * (1.0 + 1.0e-9) + 1.0e-9;
  const double cmax = std::max(cutoff, cut3) * (1.0 + 1.0e-9) + 1.0e-9;
  double cst[TA_PBC_EWALD];
  ct.constants(cst);
  nl = 0;
  sites.assign(nat, 0);
  // pairs j <= i once: the list entries of i, and the sites of anchor j
  // (atom i >= j; the table holds -T with every T, so the count is symmetric)
  for (int i = 0; i < nat; ++i)
    for (int j = 0; j <= i; ++j) {
      const double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1];
      const double dz = xyz[3 * i + 2] - xyz[3 * j + 2];
      const double d[3] = {dx, dy, dz};
      int lo[3], hi[3];
      cell_box(cst, d, cmax, lo, hi);
      for (int n1 = lo[0]; n1 <= hi[0]; ++n1)
        for (int n2 = lo[1]; n2 <= hi[1]; ++n2)
          for (int n3 = lo[2]; n3 <= hi[2]; ++n3) {
            const int t = cell_tindex(cst, n1, n2, n3);
            const double x = dx - tr[3 * t], y = dy - tr[3 * t + 1], z = dz - tr[3 * t + 2];
            const double r2 = x * x + y * y + z * z;
            if (r2 <= c2) ++nl;
            if (r2 <= s2) ++sites[j];
          }
    }
}

// -------------------------------------------------------------- workspace
struct Arena {
  char* base = nullptr;
  size_t cap = 0;
  size_t used = 0;
};

template <class T>
T* carve(Arena& a, size_t count)
{
  size_t bytes = ((count * sizeof(T)) + 255) & ~size_t(255);
  T* p = reinterpret_cast<T*>(a.base + a.used);
  a.used += bytes;
  return p;
}

// Host-side plan of one batch
struct Plan {
  int nmol = 0;
  int mem = 0;
  std::vector<int> mol_atom_off, mol_nat, mol_sh_off, mol_nsh, mol_ao_off, mol_nao, mol_uhf;
  std::vector<int64_t> mol_mat_off, mol_nl_off, mol_gam_off, mol_pair_off, mol_mix_off,
      mol_mixmem_off, mol_mixa_off, mol_big_off, mol_scr_off;
  std::vector<double> mol_charge, mol_cutoff;
  std::vector<int> at_el, at_ish, at_iao;
  std::vector<int> sh_at, sh_ish, sh_iao;
  std::vector<int> ao_at, ao_sh;
  // integral work items: those of molecules without d shells first (nwork_sp
  // of them, the s/p-only kernel), then the rest
  std::vector<int> work_mol, work_kind, work_start;
  int nwork_sp = 0;
  // three-body dispersion work: chunks [atm_i0, atm_i1) of anchor atoms
  std::vector<int> atm_mol, atm_i0, atm_i1;
  // pair-plane work: atom rows [pair_a0, pair_a1) of one molecule
  std::vector<int> pair_mol, pair_a0, pair_a1;
  std::vector<int> grid_mols;  // molecules run by the multi-workgroup SCF path
  // periodic cells (empty for a batch of molecules): per molecule the
  // translation and reciprocal-vector counts and offsets and the Ewald
  // constants; the tables; integral work of cells (one item per atom pair
  // (i, j <= i), s/p-only cells first, nwork_pbc_sp of them); pair-plane
  // items (atom rows); three-body site
  // items (one per atom) and pair items (site rows of one anchor atom); per
  // atom the site capacity and offset
  bool any_pbc = false;
  std::vector<int> mol_ntr, mol_ng;
  std::vector<int64_t> mol_tr_off, mol_g_off;
  std::vector<double> trans, recip, mol_ewald;
  std::vector<int> pw_mol, pw_i, pw_j;
  int nwork_pbc_sp = 0;
  std::vector<int> pp_mol, pp_a0, pp_a1;
  std::vector<int> ps_mol, ps_atom;
  std::vector<int> pa_mol, pa_atom, pa_r0, pa_r1;
  std::vector<int64_t> site_off;
  std::vector<int> site_cap;
  int64_t tot_site = 0;
  struct Bucket {
    int nt, maxt;
    bool lds;
    bool cell = false;  // periodic cells (AES2 from the Ewald-summed planes)
    int ldmax, natmax, nshmax;
    std::vector<int> mols;
  };
  std::vector<Bucket> buckets;
  int64_t tot_atoms = 0, tot_sh = 0, tot_ao = 0, tot_mat = 0, tot_nl = 0, tot_gam = 0,
          tot_pair = 0, tot_mix = 0, tot_mixmem = 0, tot_big = 0, tot_scr = 0;
  std::vector<int> too_large;
  std::vector<char> mol_path;  // 0 LDS, 1 one workgroup in global memory, 2 multi-workgroup
  int int_chunk = kIntegralChunk;  // integral work-item size the work list was cut with
};

struct DeviceBatch {
  BatchDev b;
  int* d_work_mol = nullptr;
  int* d_work_kind = nullptr;
  int* d_work_start = nullptr;
  int* d_pair_mol = nullptr;
  int* d_pair_a0 = nullptr;
  int* d_pair_a1 = nullptr;
  int* d_atm_mol = nullptr;
  int* d_atm_i0 = nullptr;
  int* d_atm_i1 = nullptr;
  int *d_pw_mol = nullptr, *d_pw_i = nullptr, *d_pw_j = nullptr;
  int *d_pp_mol = nullptr, *d_pp_a0 = nullptr, *d_pp_a1 = nullptr;
  int *d_ps_mol = nullptr, *d_ps_atom = nullptr;
  int *d_pa_mol = nullptr, *d_pa_atom = nullptr, *d_pa_r0 = nullptr, *d_pa_r1 = nullptr;
  std::vector<int*> d_bucket_mols;
  // energy-weighted density tiles of the gradient (molecule, row tile,
  // column tile); the host copy stays alive while the upload is in flight
  std::vector<int> h_wtiles;
  int* d_wtiles = nullptr;
  bool grad = false;
  size_t bytes = 0;
};

// 64 x 64 tiles of the lower triangle of each molecule's W
void gradient_tiles(const Plan& pl, std::vector<int>& t)
{
  t.clear();
  for (int m = 0; m < pl.nmol; ++m) {
    const int nt = (pl.mol_nao[m] + 63) / 64;
    for (int i = 0; i < nt; ++i)
      for (int j = 0; j <= i; ++j) {
        t.push_back(m);
        t.push_back(i);
        t.push_back(j);
      }
  }
}

}  // namespace

// ------------------------------------------------- fixed device workspace
// Workspace of tblite_amd_singlepoint_flat_noalloc: a static array that every
// device the library runs on holds (its own copy) until the process ends.  In
// host code `scratch` names a host-side stand-in, never the device memory:
// the device address comes from hipGetSymbolAddress, resolved once per device
// (scratch_address) and used for every kernel argument and copy.
static __device__ __align__(256) char scratch[TBLITE_AMD_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
bool g_scratch_resolved[kMaxDevices] = {};
char* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

// Device address of `scratch` on `device` (which must be current); resolved
// on the first call for that device, then cached.
char* scratch_address(int device, hipError_t* err)
{
  if (device < 0 || device >= kMaxDevices) {
    *err = hipErrorInvalidDevice;
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_resolved[device]) {
    void* ptr = nullptr;
    g_scratch_err[device] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_scratch_ptr[device] = g_scratch_err[device] == hipSuccess ? static_cast<char*>(ptr) : nullptr;
    g_scratch_resolved[device] = true;
  }
  *err = g_scratch_err[device];
  return g_scratch_ptr[device];
}
}  // namespace

// ------------------------------------------------------------- handles
struct tblite_amd_context_s {
  int device = 0;
  hipStream_t stream = nullptr;
  // device address of this device's fixed workspace (the static array of the
  // no-allocation entry point), or null if it could not be resolved
  char* fixed_ws = nullptr;
  hipStream_t aux[kNumStreams] = {};
  hipEvent_t ev[8] = {};
  hipEvent_t aux_done[kNumStreams] = {};
  hipEvent_t bev[2 * kMaxBuckets] = {};  // around each SCF bucket launch, on its stream
  double bucket_ms[kMaxBuckets] = {};
  std::string error;
  Arena arena;
  Plan plan;
  DeviceBatch batch;
  bool have_batch = false;
  double timings[5] = {0, 0, 0, 0, 0};
  size_t lds_limit = 0;
  tblite_amd_calculator prepared = nullptr;
  double prepared_acc = 1.0;
  bool profile = false;
  // padded AO dimension above which a molecule runs on the multi-workgroup
  // path (TBLITE_AMD_GRID_MIN_LD may lower it, e.g. to test that path on
  // molecules the oracle handles quickly)
  int grid_min_ld = TA_BIG_MAXLD;
  // padded AO dimension above which a molecule's matrices leave LDS for the
  // global-memory single-workgroup path (TBLITE_AMD_LDS_MAX_LD may lower it)
  int lds_max_ld = kLdsMaxLd;
  // size buckets in steps of 16 AOs instead of one per kernel instantiation
  bool fine_buckets = false;
  // work tiles of the setup kernels (see kIntegralChunk)
  int int_chunk = kIntegralChunk;
  int pair_chunk_log2 = kPairChunkLog2;
  int atm_chunk_log2 = kAtmChunkLog2;
  // replay the multi-workgroup path's eigenvalue phase from a HIP graph
  // (TBLITE_AMD_GRID_GRAPH=0 launches the kernels one by one instead)
  bool grid_graph = true;
  // Jacobi stopping threshold on the largest rotated element of a sweep at
  // self-consistency (TBLITE_AMD_JACOBI_TOL).  After such a sweep the
  // off-diagonal elements are of order tol^2 / gap.  1e-6 keeps slowly
  // converging SCFs (cations of water clusters, 29-49 iterations) within
  // 2e-8 e of the CPU oracle's charges; 1e-4 throughout moved them by up to
  // 3e-5 e, beyond the 1e-5 parity bound.
  double jacobi_tol = 1.0e-6;
  // Jacobi rotations with |a_pq| at or below min(this, threshold^2) are
  // skipped (TBLITE_AMD_JACOBI_SKIP): off-diagonals of that size are left by
  // the stopping rule anyway; the worker waves then skip the blocks and MO
  // rows of identity rotations, most of them in the last sweeps.
  double jacobi_skip = 1.0e-8;
  // > 0: the Jacobi threshold follows the SCF charge residual,
  // max(jacobi_tol, min(1e-2, jacobi_adapt * residual of the last iteration))
  // (TBLITE_AMD_JACOBI_ADAPT; 0, the default, keeps jacobi_tol throughout).
  // The loose early thresholds this allows perturb the SCF path: a (H2O)5
  // cation (uhf 1) took 34 iterations with factor 0.001 against the CPU
  // oracle's 32 (32 with 0, energy within 1e-11 Eh), beyond the +-1 parity
  // bound, so the faster schedule (about 7 % on 4096 small water clusters) is
  // off by default.
  double jacobi_adapt = 0.0;
  // eigensolver of the LDS-resident SCF kernels: 0 warm-started Jacobi (the
  // default), 1 Householder tridiagonalisation + bisection + twisted
  // factorisation (TBLITE_AMD_LDS_EIG=householder).  Both pass the parity
  // tests; on 4096 (H2O)n clusters the Jacobi takes 518 us per SCF iteration
  // at ld 65-96 against 368 + 156 + 134 + 12 us for tridiagonalisation,
  // bisection, eigenvectors and basis product (47.8 vs 60.8 ms per step)
  int lds_eig = 0;
  // reduction to tridiagonal form of the global-memory SCF kernel (ld <= 512):
  // 1 two-stage dense -> band -> tridiagonal (the default), 0 one-stage
  // Householder (TBLITE_AMD_BIG_EIG=band | householder)
  int big_eig = 1;
  // minimum LDS request of the global-memory SCF kernel (bytes); above half a
  // CU's LDS it runs one molecule per CU (TBLITE_AMD_BIG_LDS_MIN_KB)
  size_t big_lds_min = 0;
  // cost-model routing between one workgroup per molecule and the
  // multi-workgroup path (seconds per SCF iteration, measured on MI355X by
  // autotune.py: one workgroup alone wg_a * ld^wg_p, across the device
  // grid_b * ld + grid_c)
  bool auto_grid = true;
  double wg_a = 9.30e-8, wg_p = 2.085, grid_b = 4.16e-5, grid_c = -2.53e-3;
  int wg_slots = 512;  // one-workgroup molecules resident at once (two per CU)
  int verbose = 0;  // TBLITE_AMD_VERBOSE=1: progress lines of long runs on stderr
  // every kernel on the main stream (the serving slots: one stream per slot,
  // so concurrent slots do not queue behind each other's kernels)
  bool single_stream = false;
};

struct tblite_amd_structure_s {
  tblite_amd_context ctx;
  int nat;
  std::vector<int> num;
  std::vector<double> xyz;
  double charge;
  int uhf;
  bool periodic = false;  // three-dimensional cell with lattice vectors lat
  double lat[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
};

struct tblite_amd_calculator_s {
  tblite_amd_context ctx;
  Calculator c;
};

struct tblite_amd_wavefunction_s {
  tblite_amd_context ctx;
  int nat = 0, nsh = 0, nao = 0;
  std::vector<double> qsh, qat, dpat, qpat, emo, focc, n0sh, n0at, coeff, density;
};

namespace {

void set_error(tblite_amd_context ctx, const char* fmt, ...)
{
  if (!ctx) return;
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  ctx->error = buf;
}

#define HIP_TRY(ctx, call)                                                               \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) {                                                              \
      set_error(ctx, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__,    \
                __LINE__);                                                               \
      return TBLITE_AMD_ERR_HIP;                                                         \
    }                                                                                    \
  } while (0)

int ensure_params_on_device(tblite_amd_context ctx, Calculator& c)
{
  if (c.d_par) return TBLITE_AMD_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipMalloc(&c.d_par, sizeof(DevParams)));
  HIP_TRY(ctx, hipMemcpy(c.d_par, &c.dev, sizeof(DevParams), hipMemcpyHostToDevice));
  return TBLITE_AMD_OK;
}

// Build the host plan for a batch.  Returns OK or ERR_ARGUMENT.
// SCF path of every molecule from its padded AO dimension (see the tuning
// block of the header).  With auto_grid, a batch whose one-workgroup
// molecules would fill less than a quarter of the device's workgroup slots
// moves its largest molecules, one at a time, to the multi-workgroup path
// while the estimated finish time (the longer of the multi-workgroup chain
// on the main stream and the one-workgroup bucket: its slowest molecule or
// its total work over the slots) goes down.
void route_molecules(const tblite_amd_context ctx, const std::vector<int>& ld,
                     std::vector<char>& path)
{
  const int nmol = (int)ld.size();
  path.assign(nmol, 0);
  std::vector<int> wg;
  for (int m = 0; m < nmol; ++m) {
    if (ld[m] > ctx->grid_min_ld) {
      path[m] = 2;
    } else if (ld[m] > ctx->lds_max_ld) {
      path[m] = 1;
      wg.push_back(m);
    }
  }
  if (!ctx->auto_grid || wg.empty() || 4 * (int64_t)wg.size() > ctx->wg_slots) return;
  auto t_wg = [&](int m) { return ctx->wg_a * std::pow((double)ld[m], ctx->wg_p); };
  auto t_grid = [&](int m) { return std::max(0.0, ctx->grid_b * ld[m] + ctx->grid_c); };
  std::stable_sort(wg.begin(), wg.end(), [&](int a, int b) { return ld[a] > ld[b]; });
  double total = 0.0, chain = 0.0;
  for (int m : wg) total += t_wg(m);
  for (int m = 0; m < nmol; ++m)
    if (path[m] == 2) chain += t_grid(m);
  const double slots = (double)ctx->wg_slots;
  for (size_t k = 0; k < wg.size(); ++k) {
    const int m = wg[k];
    const double before = std::max(chain, std::max(t_wg(m), total / slots));
    const double rest = k + 1 < wg.size() ? t_wg(wg[k + 1]) : 0.0;
    const double after = std::max(chain + t_grid(m), std::max(rest, (total - t_wg(m)) / slots));
    if (!(after < before)) break;
    path[m] = 2;
    chain += t_grid(m);
    total -= t_wg(m);
  }
}

// lattice (9 per molecule) with xyz: periodic cells; is_cell (optional,
// one flag per molecule) marks which molecules of the batch are cells (all
// of them when null).  Cells run on the multi-workgroup SCF path, which
// reads the interaction planes their Ewald sums fill.
int make_plan(tblite_amd_context ctx, const Calculator& c, int nmol, const int* mol_nat,
              const int* num, const double* charge, const int* uhf, double acc, Plan& pl,
              const double* xyz = nullptr, const double* lattice = nullptr,
              const char* is_cell = nullptr)
{
  pl = Plan();
  pl.nmol = nmol;
  pl.mem = std::max(1, c.max_iter);
  pl.int_chunk = ctx->int_chunk;
  pl.mol_atom_off.resize(nmol);
  pl.mol_nat.resize(nmol);
  pl.mol_sh_off.resize(nmol);
  pl.mol_nsh.resize(nmol);
  pl.mol_ao_off.resize(nmol);
  pl.mol_nao.resize(nmol);
  pl.mol_uhf.resize(nmol);
  pl.mol_mat_off.resize(nmol);
  pl.mol_nl_off.resize(nmol);
  pl.mol_gam_off.resize(nmol);
  pl.mol_pair_off.resize(nmol);
  pl.mol_mix_off.resize(nmol);
  pl.mol_mixmem_off.resize(nmol);
  pl.mol_mixa_off.resize(nmol);
  pl.mol_big_off.resize(nmol);
  pl.mol_scr_off.resize(nmol);
  pl.mol_charge.resize(nmol);
  pl.mol_cutoff.resize(nmol);
  {
    // padded AO dimensions first: they decide the SCF path of each molecule
    std::vector<int> lds(nmol);
    int64_t a0 = 0;
    for (int m = 0; m < nmol; ++m) {
      int nao = 0;
      for (int i = 0; i < mol_nat[m] && mol_nat[m] > 0; ++i) {
        const int z = num[a0 + i];
        const int e = (z > 0 && z < 120) ? c.zmap[z] : -1;
        if (e < 0) break;  // reported by the pass below
        const DevElem& el = c.dev.el[e];
        for (int s = 0; s < el.nsh; ++s) nao += 2 * el.ang[s] + 1;
      }
      a0 += std::max(0, mol_nat[m]);
      lds[m] = (nao + 3) & ~3;
    }
    route_molecules(ctx, lds, pl.mol_path);
    // cells by size, like molecules (no cost-model routing): the CELL
    // instantiations of the one-workgroup kernels read the Ewald-summed AES2
    // planes and, in the global-memory kernel's Fock build, the stored
    // transposed multipole elements (a cell's image sums do not obey the
    // molecular origin shift it derives them by)
    if (lattice)
      for (int m = 0; m < nmol; ++m)
        if (!is_cell || is_cell[m])
          pl.mol_path[m] = lds[m] <= ctx->lds_max_ld ? 0 : (lds[m] <= ctx->grid_min_ld ? 1 : 2);
  }
  if (lattice && !xyz) {
    set_error(ctx, "internal: periodic plan without coordinates");
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  pl.any_pbc = lattice != nullptr;
  if (pl.any_pbc) {
    pl.mol_ntr.assign(nmol, 0);
    pl.mol_ng.assign(nmol, 0);
    pl.mol_tr_off.assign(nmol, 0);
    pl.mol_g_off.assign(nmol, 0);
    pl.mol_ewald.assign((size_t)TA_PBC_EWALD * nmol, 0.0);
  }
  // cells: lattice tables and list capacities of every cell first, spread
  // over host threads (the capacity counts are O(nat^2) box walks per cell)
  struct CellPre {
    bool ok = false;
    CellTables ct;
    int64_t nl = 0;
    std::vector<int> sites;
  };
  std::vector<CellPre> pre;
  if (lattice) {
    pre.resize(nmol);
    std::vector<int64_t> aoff(nmol + 1, 0);
    for (int m = 0; m < nmol; ++m) aoff[m + 1] = aoff[m] + std::max(0, mol_nat[m]);
    auto work = [&](int m) {
      const int nat = mol_nat[m];
      if ((is_cell && !is_cell[m]) || nat < 1) return;
      double min_alpha = INFINITY;
      for (int i = 0; i < nat; ++i) {
        const int z = num[aoff[m] + i];
        const int e = (z > 0 && z < 120) ? c.zmap[z] : -1;
        if (e < 0) return;  // reported by the molecule loop
        min_alpha = std::min(min_alpha, c.min_alpha_el[e]);
      }
      CellPre& p = pre[m];
      if (!cell_tables(lattice + 9 * (size_t)m, nat, xyz + 3 * aoff[m], p.ct)) return;
      p.ok = true;
      if ((int64_t)p.ct.tr.size() / 3 > (1 << 20)) return;  // rejected by the molecule loop
      cell_counts(nat, xyz + 3 * aoff[m], p.ct, get_cutoff(min_alpha, acc),
                  c.dev.d4.enabled ? c.dev.d4.disp3_cutoff : 0.0, p.nl, p.sites);
    };
    const unsigned hc = std::thread::hardware_concurrency();
    const int nth = std::max(1, std::min(nmol, std::min(16, (int)(hc ? hc : 1))));
    std::atomic<int> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < nth; ++t)
      th.emplace_back([&]() {
        for (int m = next++; m < nmol; m = next++) work(m);
      });
    for (auto& t : th) t.join();
  }
  int64_t atom = 0, shell = 0, ao = 0, mat = 0, nl = 0, gam = 0, pair = 0, mix = 0, mixmem = 0,
          big = 0, scr = 0;
  std::vector<int> wd_mol, wd_kind, wd_start;  // integral work of molecules with d shells
  std::vector<int> pwd_mol, pwd_i, pwd_j;  // periodic cells with d shells
  for (int m = 0; m < nmol; ++m) {
    const int nat = mol_nat[m];
    bool has_d = false;
    if (nat < 1) {
      set_error(ctx, "molecule %d has %d atoms", m, nat);
      return TBLITE_AMD_ERR_ARGUMENT;
    }
    pl.mol_atom_off[m] = (int)atom;
    pl.mol_nat[m] = nat;
    pl.mol_sh_off[m] = (int)shell;
    pl.mol_ao_off[m] = (int)ao;
    int nsh = 0, nao = 0;
    double min_alpha = INFINITY;
    for (int i = 0; i < nat; ++i) {
      const int z = num[atom + i];
      const int e = (z > 0 && z < 120) ? c.zmap[z] : -1;
      if (e < 0) {
        set_error(ctx, "molecule %d atom %d: element Z=%d has no parameters", m, i, z);
        return TBLITE_AMD_ERR_ARGUMENT;
      }
      const DevElem& el = c.dev.el[e];
      pl.at_el.push_back(e);
      pl.at_ish.push_back(nsh);
      pl.at_iao.push_back(nao);
      min_alpha = std::min(min_alpha, c.min_alpha_el[e]);
      for (int s = 0; s < el.nsh; ++s) {
        if (el.ang[s] >= 2) has_d = true;
        pl.sh_at.push_back(i);
        pl.sh_ish.push_back(s);
        pl.sh_iao.push_back(nao);
        const int na = 2 * el.ang[s] + 1;
        for (int k = 0; k < na; ++k) {
          pl.ao_at.push_back(i);
          pl.ao_sh.push_back(nsh);
        }
        nao += na;
        ++nsh;
      }
    }
    pl.mol_nsh[m] = nsh;
    pl.mol_nao[m] = nao;
    pl.mol_uhf[m] = uhf ? uhf[m] : 0;
    pl.mol_charge[m] = charge ? charge[m] : 0.0;
    pl.mol_cutoff[m] = get_cutoff(min_alpha, acc);
    const bool cell = lattice && (!is_cell || is_cell[m]);
    int64_t nl_cap = (int64_t)nat * (nat + 1) / 2;
    std::vector<int> site_cnt;
    if (cell) {
      if (!pre[m].ok) {
        set_error(ctx, "molecule %d: degenerate lattice", m);
        return TBLITE_AMD_ERR_ARGUMENT;
      }
      const CellTables& ct = pre[m].ct;
      const int64_t ntr = (int64_t)ct.tr.size() / 3;
      if (ntr > (1 << 20)) {
        set_error(ctx, "molecule %d: %lld lattice translations (cell too thin for the cutoffs)", m,
                  (long long)ntr);
        return TBLITE_AMD_ERR_TOO_LARGE;
      }
      pl.mol_ntr[m] = (int)ntr;
      pl.mol_tr_off[m] = (int64_t)pl.trans.size() / 3;
      pl.trans.insert(pl.trans.end(), ct.tr.begin(), ct.tr.end());
      pl.mol_ng[m] = (int)(ct.g.size() / 3);
      pl.mol_g_off[m] = (int64_t)pl.recip.size() / 3;
      pl.recip.insert(pl.recip.end(), ct.g.begin(), ct.g.end());
      ct.constants(pl.mol_ewald.data() + TA_PBC_EWALD * (size_t)m);
      nl_cap = pre[m].nl;
      site_cnt.swap(pre[m].sites);
      if (nl_cap > (int64_t)1 << 30) {
        set_error(ctx, "molecule %d: neighbour list of %lld entries", m, (long long)nl_cap);
        return TBLITE_AMD_ERR_TOO_LARGE;
      }
    }
    if (pl.any_pbc)
      for (int i = 0; i < nat; ++i) {
        pl.site_off.push_back(pl.tot_site);
        const int cap = cell && c.dev.d4.enabled ? site_cnt[i] : 0;
        pl.site_cap.push_back(cap);
        pl.tot_site += cap;
      }
    pl.mol_mat_off[m] = mat;
    pl.mol_nl_off[m] = nl;
    pl.mol_gam_off[m] = gam;
    pl.mol_pair_off[m] = pair;
    pl.mol_mix_off[m] = mix;
    pl.mol_mixmem_off[m] = mixmem;
    pl.mol_mixa_off[m] = (int64_t)m * pl.mem * pl.mem;
    const int ld = (nao + 3) & ~3;
    const int ndim = nsh + 9 * nat;
    if (ld > kGridMaxLd) {
      set_error(ctx, "molecule %d has %d atomic orbitals; the engine serves up to %d", m, nao,
                kGridMaxLd);
      return TBLITE_AMD_ERR_TOO_LARGE;
    }
    if (pl.mol_path[m] == 2) {
      pl.grid_mols.push_back(m);
      pl.mol_big_off[m] = big;
      big += (int64_t)ld * ld;
      pl.mol_scr_off[m] = scr;
      scr += grid_scratch_doubles(nao, ld, nat, nsh);
    } else if (pl.mol_path[m] == 1) {
      pl.mol_big_off[m] = big;
      big += (int64_t)ld * ld;
      pl.mol_scr_off[m] = scr;
      scr += big_scratch_doubles(ld);
    } else {
      pl.mol_big_off[m] = -1;
      pl.mol_scr_off[m] = -1;
    }
    atom += nat;
    shell += nsh;
    ao += nao;
    mat += (int64_t)nao * nao;
    nl += nl_cap;
    gam += (int64_t)nsh * nsh;
    pair += (int64_t)nat * nat;
    mix += ndim;
    mixmem += (int64_t)ndim * pl.mem;
    // integral work: neighbour chunks over the list capacity, on-site chunks
    // (cells: their own lists, s/p-only molecules first)
    if (cell) {
      std::vector<int>& wm = has_d ? pwd_mol : pl.pw_mol;
      std::vector<int>& wi = has_d ? pwd_i : pl.pw_i;
      std::vector<int>& wj = has_d ? pwd_j : pl.pw_j;
      for (int i = 0; i < nat; ++i)
        for (int j = 0; j <= i; ++j) {
          wm.push_back(m);
          wi.push_back(i);
          wj.push_back(j);
        }
      // pair planes: one thread per atom pair, rows of up to 256 pairs
      const int rows = std::max(1, 256 / nat);
      for (int a = 0; a < nat; a += rows) {
        pl.pp_mol.push_back(m);
        pl.pp_a0.push_back(a);
        pl.pp_a1.push_back(std::min(nat, a + rows));
      }
      // three-body: site lists per atom, then pair items of about 2^16
      // site pairs (rows of one anchor's sites)
      if (c.dev.d4.enabled)
        for (int i = 0; i < nat; ++i) {
          pl.ps_mol.push_back(m);
          pl.ps_atom.push_back(i);
          const int n = site_cnt[i];
          int r0 = 0;
          int64_t accp = 0;
          for (int a = 0; a < n; ++a) {
            accp += n - a - 1;
            if (accp >= (1 << 16) || a == n - 1) {
              pl.pa_mol.push_back(m);
              pl.pa_atom.push_back(i);
              pl.pa_r0.push_back(r0);
              pl.pa_r1.push_back(a + 1);
              r0 = a + 1;
              accp = 0;
            }
          }
        }
    } else {
      std::vector<int>& wm = has_d ? wd_mol : pl.work_mol;
      std::vector<int>& wk = has_d ? wd_kind : pl.work_kind;
      std::vector<int>& ws = has_d ? wd_start : pl.work_start;
      const int64_t cap = (int64_t)nat * (nat + 1) / 2;
      for (int64_t s = 0; s < cap; s += ctx->int_chunk) {
        wm.push_back(m);
        wk.push_back(0);
        ws.push_back((int)s);
      }
      for (int s = 0; s < nat; s += ctx->int_chunk) {
        wm.push_back(m);
        wk.push_back(1);
        ws.push_back(s);
      }
    }
    // pair planes: rows of about kPairChunk pairs per work item
    if (!cell) {
      const int rows = (int)std::max<int64_t>(1, ((int64_t)1 << ctx->pair_chunk_log2) / nat);
      for (int a = 0; a < nat; a += rows) {
        pl.pair_mol.push_back(m);
        pl.pair_a0.push_back(a);
        pl.pair_a1.push_back(std::min(nat, a + rows));
      }
    }
    // three-body work: anchor atom i owns the triples k < j < i; chunks of
    // consecutive anchors hold about kAtmChunk triples each
    if (c.dev.d4.enabled && !cell) {
      int i0 = 0;
      int64_t acc = 0;
      for (int i = 0; i < nat; ++i) {
        acc += (int64_t)i * (i - 1) / 2;
        if (acc >= ((int64_t)1 << ctx->atm_chunk_log2) || i == nat - 1) {
          pl.atm_mol.push_back(m);
          pl.atm_i0.push_back(i0);
          pl.atm_i1.push_back(i + 1);
          i0 = i + 1;
          acc = 0;
        }
      }
    }
  }
  pl.tot_atoms = atom;
  pl.tot_sh = shell;
  pl.tot_ao = ao;
  pl.tot_mat = mat;
  pl.tot_nl = nl;
  pl.nwork_sp = (int)pl.work_mol.size();
  pl.work_mol.insert(pl.work_mol.end(), wd_mol.begin(), wd_mol.end());
  pl.work_kind.insert(pl.work_kind.end(), wd_kind.begin(), wd_kind.end());
  pl.work_start.insert(pl.work_start.end(), wd_start.begin(), wd_start.end());
  pl.nwork_pbc_sp = (int)pl.pw_mol.size();
  pl.pw_mol.insert(pl.pw_mol.end(), pwd_mol.begin(), pwd_mol.end());
  pl.pw_i.insert(pl.pw_i.end(), pwd_i.begin(), pwd_i.end());
  pl.pw_j.insert(pl.pw_j.end(), pwd_j.begin(), pwd_j.end());
  pl.tot_gam = gam;
  pl.tot_pair = pair;
  pl.tot_mix = mix;
  pl.tot_mixmem = mixmem;
  pl.tot_big = big;
  pl.tot_scr = scr;

  // size buckets over the padded AO dimension; largest molecules first
  // One bucket per kernel instantiation (ld <= 32, 33..64, 65..96) plus the
  // LDS limit of the tuning as an extra bound; a bucket's ldmax sets its LDS
  // request, and within each instantiation the occupancy is the same for
  // every ldmax (the 512-thread kernel runs one molecule per CU, the 256- and
  // 128-thread ones are held by waves), so wider buckets cost nothing and
  // let the largest-first order cover all sizes (shorter tail).
  // TBLITE_AMD_FINE_BUCKETS=1 restores steps of 16.
  std::vector<int> bounds;
  if (ctx->fine_buckets)
    bounds = {16, 32, 48, 64, 80, 96};
  else
    bounds = {32, 64, 96};
  if (std::find(bounds.begin(), bounds.end(), ctx->lds_max_ld) == bounds.end())
    bounds.push_back(ctx->lds_max_ld);
  std::sort(bounds.begin(), bounds.end());
  bounds.push_back(1 << 30);
  std::vector<int> order(nmol);
  for (int m = 0; m < nmol; ++m) order[m] = m;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return pl.mol_nao[a] > pl.mol_nao[b]; });
  std::vector<char> cellf(nmol, 0);
  for (int m = 0; m < nmol && pl.any_pbc; ++m) cellf[m] = pl.mol_ntr[m] > 0;
  for (int k = (int)bounds.size() - 1; k >= 0; --k)
  for (int want_cell = 0; want_cell < (pl.any_pbc ? 2 : 1); ++want_cell) {
    Plan::Bucket bk;
    const int lo = (k == 0) ? 0 : bounds[k - 1];
    const int hi = bounds[k];
    bk.ldmax = 0;
    bk.natmax = 0;
    bk.nshmax = 0;
    bk.cell = want_cell != 0;
    for (int m : order) {
      const int ld = (pl.mol_nao[m] + 3) & ~3;
      if (ld > lo && ld <= hi && pl.mol_path[m] != 2 && (cellf[m] != 0) == bk.cell) {
        bk.mols.push_back(m);
        bk.ldmax = std::max(bk.ldmax, ld);
        bk.natmax = std::max(bk.natmax, pl.mol_nat[m]);
        bk.nshmax = std::max(bk.nshmax, pl.mol_nsh[m]);
      }
    }
    if (bk.mols.empty()) continue;
    if (hi > ctx->lds_max_ld) {
      bk.nt = TA_BIG_NT;
      bk.maxt = 1;
      bk.lds = false;
    } else if (hi <= 32) {
      bk.nt = 128;
      bk.maxt = 1;
      bk.lds = true;
    } else if (hi <= 64) {
      bk.nt = 256;
      bk.maxt = 1;
      bk.lds = true;
    } else if (hi <= kLdsMaxLd) {
      bk.nt = 512;
      bk.maxt = 2;
      bk.lds = true;
    } else {
      bk.nt = TA_BIG_NT;
      bk.maxt = 1;
      bk.lds = false;
    }
    pl.buckets.push_back(bk);
  }
  return TBLITE_AMD_OK;
}

template <class T>
void host_pack(std::vector<char>& buf, std::vector<std::pair<size_t, size_t>>& segs,
               const std::vector<T>& v)
{
  size_t off = (buf.size() + 255) & ~size_t(255);
  buf.resize(off + v.size() * sizeof(T) + 1);
  if (!v.empty()) std::memcpy(buf.data() + off, v.data(), v.size() * sizeof(T));
  segs.push_back({off, v.size() * sizeof(T)});
}

// Packed host image of a plan's integer / offset arrays: segment k of seg
// holds array k below (fixed ones first, then one per bucket).
void pack_plan(const Plan& pl, std::vector<char>& img, std::vector<std::pair<size_t, size_t>>& seg)
{
  host_pack(img, seg, pl.mol_atom_off);   // 0
  host_pack(img, seg, pl.mol_nat);        // 1
  host_pack(img, seg, pl.mol_sh_off);     // 2
  host_pack(img, seg, pl.mol_nsh);        // 3
  host_pack(img, seg, pl.mol_ao_off);     // 4
  host_pack(img, seg, pl.mol_nao);        // 5
  host_pack(img, seg, pl.mol_mat_off);    // 6
  host_pack(img, seg, pl.mol_nl_off);     // 7
  host_pack(img, seg, pl.mol_gam_off);    // 8
  host_pack(img, seg, pl.mol_pair_off);   // 9
  host_pack(img, seg, pl.mol_mix_off);    // 10
  host_pack(img, seg, pl.mol_mixmem_off); // 11
  host_pack(img, seg, pl.mol_mixa_off);   // 12
  host_pack(img, seg, pl.mol_charge);     // 13
  host_pack(img, seg, pl.mol_uhf);        // 14
  host_pack(img, seg, pl.mol_cutoff);     // 15
  host_pack(img, seg, pl.at_el);          // 16
  host_pack(img, seg, pl.at_ish);         // 17
  host_pack(img, seg, pl.at_iao);         // 18
  host_pack(img, seg, pl.sh_at);          // 19
  host_pack(img, seg, pl.sh_ish);         // 20
  host_pack(img, seg, pl.sh_iao);         // 21
  host_pack(img, seg, pl.ao_at);          // 22
  host_pack(img, seg, pl.ao_sh);          // 23
  host_pack(img, seg, pl.work_mol);       // 24
  host_pack(img, seg, pl.work_kind);      // 25
  host_pack(img, seg, pl.work_start);     // 26
  host_pack(img, seg, pl.mol_big_off);    // 27
  host_pack(img, seg, pl.mol_scr_off);    // 28
  host_pack(img, seg, pl.atm_mol);        // 29
  host_pack(img, seg, pl.atm_i0);         // 30
  host_pack(img, seg, pl.atm_i1);         // 31
  host_pack(img, seg, pl.pair_mol);       // 32
  host_pack(img, seg, pl.pair_a0);        // 33
  host_pack(img, seg, pl.pair_a1);        // 34
  host_pack(img, seg, pl.mol_ntr);        // 35
  host_pack(img, seg, pl.mol_tr_off);     // 36
  host_pack(img, seg, pl.trans);          // 37
  host_pack(img, seg, pl.mol_ng);         // 38
  host_pack(img, seg, pl.mol_g_off);      // 39
  host_pack(img, seg, pl.recip);          // 40
  host_pack(img, seg, pl.mol_ewald);      // 41
  host_pack(img, seg, pl.pw_mol);         // 42
  host_pack(img, seg, pl.pw_i);           // 43
  host_pack(img, seg, pl.pw_j);           // 44
  host_pack(img, seg, pl.pp_mol);         // 45
  host_pack(img, seg, pl.pp_a0);          // 46
  host_pack(img, seg, pl.pp_a1);          // 47
  host_pack(img, seg, pl.ps_mol);         // 48
  host_pack(img, seg, pl.ps_atom);        // 49
  host_pack(img, seg, pl.pa_mol);         // 50
  host_pack(img, seg, pl.pa_atom);        // 51
  host_pack(img, seg, pl.pa_r0);          // 52
  host_pack(img, seg, pl.pa_r1);          // 53
  host_pack(img, seg, pl.site_off);       // 54
  host_pack(img, seg, pl.site_cap);       // 55
  for (const auto& bk : pl.buckets) host_pack(img, seg, bk.mols);

}

// Device workspace of a plan in bytes (every array upload_plan carves, each
// rounded to 256 bytes).  img_bytes: size of the packed plan image.
size_t workspace_need(const Plan& pl, size_t img_bytes, bool wfn_arrays, bool grad = false)
{
  const int nmol = pl.nmol, mem = pl.mem;
  const int64_t A = pl.tot_atoms, Sh = pl.tot_sh, AO = pl.tot_ao;
  auto dbytes = [](int64_t n) { return ((size_t)n * 8 + 255) & ~size_t(255); };
  auto ibytes = [](int64_t n) { return ((size_t)n * 4 + 255) & ~size_t(255); };
  size_t need = ((img_bytes + 255) & ~size_t(255));
  need += dbytes(3 * A) + 3 * dbytes(A);                 // xyz, cn, mrad, erep
  need += 2 * dbytes(Sh);                                // selfenergy, n0sh
  need += 3 * ibytes(A) + ibytes(pl.tot_nl) + ibytes(nmol); // inl, nnl, jat, nnz (+pad)
  need += 11 * dbytes(pl.tot_mat);                       // S, H, D, Q
  if (wfn_arrays) need += 2 * dbytes(pl.tot_mat);        // coeff, density
  // gam; sd, dd, sq planes only for the multi-workgroup SCF (the
  // one-workgroup kernels evaluate the AES2 pair kernels where they use them)
  const bool aes_planes = !pl.grid_mols.empty() || pl.any_pbc;
  need += dbytes(pl.tot_gam) + (aes_planes ? 18 * dbytes(pl.tot_pair) : 0);
  need += dbytes(Sh) + dbytes(A) + dbytes(3 * A) + dbytes(6 * A) + 2 * dbytes(AO); // wfn
  need += 4 * dbytes(pl.tot_mix) + 2 * dbytes(pl.tot_mixmem);
  need += dbytes((int64_t)nmol * mem * mem) + dbytes((int64_t)nmol * mem * (mem + 1)) +
          2 * dbytes((int64_t)nmol * mem);
  need += 3 * dbytes(nmol) + 2 * ibytes(nmol);
  need += 4 * dbytes(std::max<int64_t>(pl.tot_big, 1));
  need += dbytes(std::max<int64_t>(pl.tot_scr, 1));
  need += dbytes((int64_t)nmol * TA_PROF_SLOTS);
  need += 2 * dbytes(A) + 2 * dbytes(TA_D4_MAXREF * A) + 2 * dbytes(pl.tot_pair) + 2 * dbytes(nmol); // D4
  if (pl.any_pbc) need += ibytes(pl.tot_nl) + ibytes(A) + dbytes(4 * std::max<int64_t>(pl.tot_site, 1));
  if (grad) {
    int64_t ntiles = 0;
    for (int m = 0; m < nmol; ++m) {
      const int64_t nt = (pl.mol_nao[m] + 63) / 64;
      ntiles += nt * (nt + 1) / 2;
    }
    if (!wfn_arrays) need += 2 * dbytes(pl.tot_mat);     // coeff, density
    need += dbytes(3 * A) + 3 * dbytes(A) + dbytes(3 * A) + dbytes(6 * A);  // grad, dcn, dcn4, vat, vdp, vqp
    need += dbytes((int64_t)TA_GD4 * TA_D4_MAXREF * A) + dbytes(Sh);          // D4 weights, vsh
    need += dbytes(pl.tot_mat) + dbytes(pl.tot_pair) + ibytes(3 * ntiles);  // W, dE3/dC6, tiles
  }
  need += 64 * 256;

  return need;
}

// Allocate (or reuse) the workspace for a plan and upload the plan arrays.
// wfn_arrays: also the per-molecule coefficient and density matrices of a
// requested wavefunction (2 n^2 doubles per molecule).
// fixed: carve from this caller-provided arena instead of the context's and
// never allocate (a plan that does not fit is an error); par: device address
// of the parameter block when it is not the calculator's own copy.
int upload_plan(tblite_amd_context ctx, Calculator& c, const Plan& pl, DeviceBatch& db,
                bool wfn_arrays = false, Arena* fixed = nullptr, const DevParams* par = nullptr,
                bool grad = false)
{
  // the gradient reads the density and MO coefficients of the SCF
  wfn_arrays = wfn_arrays || grad;
  const int nmol = pl.nmol, mem = pl.mem;
  std::vector<char> img;
  std::vector<std::pair<size_t, size_t>> seg;
  pack_plan(pl, img, seg);
  const size_t nfixed = seg.size() - pl.buckets.size();  // fixed arrays before the bucket lists
  const int64_t A = pl.tot_atoms, Sh = pl.tot_sh, AO = pl.tot_ao;
  const size_t need = workspace_need(pl, img.size(), wfn_arrays, grad);
  const bool aes_planes = !pl.grid_mols.empty() || pl.any_pbc;

  Arena& ar = fixed ? *fixed : ctx->arena;
  if (fixed && ar.cap < need) {
    set_error(ctx, "batch needs %zu bytes of device workspace; the fixed workspace holds %zu",
              need, ar.cap);
    return TBLITE_AMD_ERR_TOO_LARGE;
  }
  if (ar.cap < need) {
    if (ar.base) {
      HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
      HIP_TRY(ctx, hipFree(ar.base));
      ar.base = nullptr;
      ar.cap = 0;
    }
    size_t cap = need + need / 4;
    HIP_TRY(ctx, hipMalloc(&ar.base, cap));
    ar.cap = cap;
  }
  ar.used = 0;
  char* d_img = carve<char>(ar, img.size());
  HIP_TRY(ctx, hipMemcpyAsync(d_img, img.data(), img.size(), hipMemcpyHostToDevice, ctx->stream));
  auto I = [&](size_t k) { return reinterpret_cast<int*>(d_img + seg[k].first); };
  auto L = [&](size_t k) { return reinterpret_cast<int64_t*>(d_img + seg[k].first); };
  auto Dd = [&](size_t k) { return reinterpret_cast<double*>(d_img + seg[k].first); };

  BatchDev& b = db.b;
  std::memset(&b, 0, sizeof(b));
  b.nmol = nmol;
  b.max_iter = c.max_iter;
  b.mix_memory = mem;
  b.damping = c.damping;
  b.kt = c.etemp * TBLITE_AMD_KB_EH;
  b.accuracy = 0.0; // set by the caller
  b.par = par ? par : c.d_par;
  b.mol_atom_off = I(0);
  b.mol_nat = I(1);
  b.mol_sh_off = I(2);
  b.mol_nsh = I(3);
  b.mol_ao_off = I(4);
  b.mol_nao = I(5);
  b.mol_mat_off = L(6);
  b.mol_nl_off = L(7);
  b.mol_gam_off = L(8);
  b.mol_pair_off = L(9);
  b.mol_mix_off = L(10);
  b.mol_mixmem_off = L(11);
  b.mol_mixa_off = L(12);
  b.mol_charge = Dd(13);
  b.mol_uhf = I(14);
  b.mol_cutoff = Dd(15);
  b.at_el = I(16);
  b.at_ish = I(17);
  b.at_iao = I(18);
  b.sh_at = I(19);
  b.sh_ish = I(20);
  b.sh_iao = I(21);
  b.ao_at = I(22);
  b.ao_sh = I(23);
  db.d_work_mol = I(24);
  db.d_work_kind = I(25);
  db.d_work_start = I(26);
  b.mol_big_off = L(27);
  b.mol_scr_off = L(28);
  db.d_atm_mol = I(29);
  db.d_atm_i0 = I(30);
  db.d_atm_i1 = I(31);
  db.d_pair_mol = I(32);
  db.d_pair_a0 = I(33);
  db.d_pair_a1 = I(34);
  if (pl.any_pbc) {
    b.mol_ntr = I(35);
    b.mol_tr_off = L(36);
    b.trans = Dd(37);
    b.mol_ng = I(38);
    b.mol_g_off = L(39);
    b.recip = Dd(40);
    b.mol_ewald = Dd(41);
    b.atm_site_off = L(54);
    b.atm_site_cap = I(55);
  }
  db.d_pw_mol = I(42);
  db.d_pw_i = I(43);
  db.d_pw_j = I(44);
  db.d_pp_mol = I(45);
  db.d_pp_a0 = I(46);
  db.d_pp_a1 = I(47);
  db.d_ps_mol = I(48);
  db.d_ps_atom = I(49);
  db.d_pa_mol = I(50);
  db.d_pa_atom = I(51);
  db.d_pa_r0 = I(52);
  db.d_pa_r1 = I(53);
  db.d_bucket_mols.clear();
  for (size_t k = 0; k < pl.buckets.size(); ++k) db.d_bucket_mols.push_back(I(nfixed + k));

  b.xyz = carve<double>(ar, 3 * A);
  b.cn = carve<double>(ar, A);
  b.mrad = carve<double>(ar, A);
  b.erep = carve<double>(ar, A);
  b.selfenergy = carve<double>(ar, Sh);
  b.n0sh = carve<double>(ar, Sh);
  b.nl_inl = carve<int>(ar, A);
  b.nl_nnl = carve<int>(ar, A);
  b.nl_jat = carve<int>(ar, std::max<int64_t>(pl.tot_nl, 1));
  b.mol_nnz = carve<int>(ar, nmol);
  b.S = carve<double>(ar, pl.tot_mat);
  b.H = carve<double>(ar, pl.tot_mat);
  b.D = carve<double>(ar, 3 * pl.tot_mat);
  b.Q = carve<double>(ar, 6 * pl.tot_mat);
  b.gam = carve<double>(ar, pl.tot_gam);
  b.aes_sd = aes_planes ? carve<double>(ar, 3 * pl.tot_pair) : nullptr;
  b.aes_dd = aes_planes ? carve<double>(ar, 9 * pl.tot_pair) : nullptr;
  b.aes_sq = aes_planes ? carve<double>(ar, 6 * pl.tot_pair) : nullptr;
  b.qsh = carve<double>(ar, Sh);
  b.qat = carve<double>(ar, A);
  b.dpat = carve<double>(ar, 3 * A);
  b.qpat = carve<double>(ar, 6 * A);
  b.emo = carve<double>(ar, AO);
  b.focc = carve<double>(ar, AO);
  b.coeff = wfn_arrays ? carve<double>(ar, pl.tot_mat) : nullptr;
  b.density = wfn_arrays ? carve<double>(ar, pl.tot_mat) : nullptr;
  b.mix_qin = carve<double>(ar, pl.tot_mix);
  b.mix_dq = carve<double>(ar, pl.tot_mix);
  b.mix_dqlast = carve<double>(ar, pl.tot_mix);
  b.mix_qlast = carve<double>(ar, pl.tot_mix);
  b.mix_df = carve<double>(ar, pl.tot_mixmem);
  b.mix_u = carve<double>(ar, pl.tot_mixmem);
  b.mix_a = carve<double>(ar, (int64_t)nmol * mem * mem);
  b.mix_beta = carve<double>(ar, (int64_t)nmol * mem * (mem + 1));
  b.mix_omega = carve<double>(ar, (int64_t)nmol * mem);
  b.mix_c = carve<double>(ar, (int64_t)nmol * mem);
  b.energy = carve<double>(ar, nmol);
  b.eelec = carve<double>(ar, nmol);
  b.erep_tot = carve<double>(ar, nmol);
  b.iterations = carve<int>(ar, nmol);
  b.status = carve<int>(ar, nmol);
  b.big_a = carve<double>(ar, std::max<int64_t>(pl.tot_big, 1));
  b.big_c = carve<double>(ar, std::max<int64_t>(pl.tot_big, 1));
  b.big_t = carve<double>(ar, std::max<int64_t>(pl.tot_big, 1));
  b.big_ct = carve<double>(ar, std::max<int64_t>(pl.tot_big, 1));
  b.big_s = carve<double>(ar, std::max<int64_t>(pl.tot_scr, 1));
  b.d4_cn = carve<double>(ar, A);
  b.d4_e3 = carve<double>(ar, A);
  b.d4_gw = carve<double>(ar, TA_D4_MAXREF * A);
  b.d4_dgw = carve<double>(ar, TA_D4_MAXREF * A);
  b.d4_disp = carve<double>(ar, pl.tot_pair);
  b.d4_c60 = carve<double>(ar, pl.tot_pair);
  b.edisp2 = carve<double>(ar, nmol);
  b.edisp3 = carve<double>(ar, nmol);
  if (pl.any_pbc) {
    b.nl_itr = carve<int>(ar, std::max<int64_t>(pl.tot_nl, 1));
    b.atm_site_n = carve<int>(ar, A);
    b.atm_site = carve<double>(ar, 4 * std::max<int64_t>(pl.tot_site, 1));
  }
  b.intcut = 0.0;
  b.jacobi_tol = ctx->jacobi_tol;
  b.jacobi_skip = ctx->jacobi_skip;
  b.jacobi_adapt = ctx->jacobi_adapt;
  b.lds_eig = ctx->lds_eig;
  b.big_eig = ctx->big_eig;
  b.prof = ctx->profile ? carve<long long>(ar, (int64_t)nmol * TA_PROF_SLOTS) : nullptr;
  db.grad = grad;
  db.d_wtiles = nullptr;
  db.h_wtiles.clear();
  if (grad) {
    b.grad = carve<double>(ar, 3 * A);
    b.g_dcn = carve<double>(ar, A);
    b.g_dcn4 = carve<double>(ar, A);
    b.g_vat = carve<double>(ar, A);
    b.g_vdp = carve<double>(ar, 3 * A);
    b.g_vqp = carve<double>(ar, 6 * A);
    b.g_d4 = carve<double>(ar, (int64_t)TA_GD4 * TA_D4_MAXREF * A);
    b.g_vsh = carve<double>(ar, Sh);
    b.g_w = carve<double>(ar, pl.tot_mat);
    b.g_c6d = carve<double>(ar, pl.tot_pair);
    gradient_tiles(pl, db.h_wtiles);
    db.d_wtiles = carve<int>(ar, (int64_t)db.h_wtiles.size());
    if (ar.used <= ar.cap && !db.h_wtiles.empty())
      HIP_TRY(ctx, hipMemcpyAsync(db.d_wtiles, db.h_wtiles.data(), sizeof(int) * db.h_wtiles.size(),
                                  hipMemcpyHostToDevice, ctx->stream));
  }
  db.bytes = ar.used;
  if (ar.used > ar.cap) {
    set_error(ctx, "internal: workspace estimate too small (%zu > %zu)", ar.used, ar.cap);
    return TBLITE_AMD_ERR_HIP;
  }
  return TBLITE_AMD_OK;
}

template <int NT, int MAXT, int MAXB, int MAXC, bool LDS, bool CELL = false>
int launch_scf(tblite_amd_context ctx, hipStream_t st, const BatchDev& b, const int* mols,
               int count, int ldmax, int natmax, int nshmax)
{
  const ScfLds L = scf_lds_layout(ldmax, natmax, nshmax, LDS);
  size_t bytes = (size_t)L.total * sizeof(double);
  // a larger request than the layout needs limits how many workgroups share a CU
  if (!LDS) bytes = std::max(bytes, ctx->big_lds_min);
  if (bytes > ctx->lds_limit) {
    set_error(ctx, "SCF bucket needs %zu bytes of LDS (limit %zu)", bytes, ctx->lds_limit);
    return TBLITE_AMD_ERR_TOO_LARGE;
  }
  HIP_TRY(ctx, hipFuncSetAttribute((const void*)k_scf<NT, MAXT, MAXB, MAXC, LDS, CELL>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  hipLaunchKernelGGL((k_scf<NT, MAXT, MAXB, MAXC, LDS, CELL>), dim3(count), dim3(NT), bytes, st, b,
                     mols, ldmax, natmax, nshmax);
  HIP_TRY(ctx, hipGetLastError());
  return TBLITE_AMD_OK;
}

int finish_batch(tblite_amd_context ctx);

// Device view of grid-path molecule m of the planned batch.
int run_grid_molecule(tblite_amd_context ctx, Calculator& c, int m)
{
  const Plan& pl = ctx->plan;
  const BatchDev& b = ctx->batch.b;
  GridMol g;
  std::memset(&g, 0, sizeof(g));
  const int n = pl.mol_nao[m], nat = pl.mol_nat[m], nsh = pl.mol_nsh[m];
  const int ld = (n + 3) & ~3;
  const int a0 = pl.mol_atom_off[m], s0 = pl.mol_sh_off[m], o0 = pl.mol_ao_off[m];
  const int64_t moff = pl.mol_mat_off[m], poff = pl.mol_pair_off[m];
  const int mem = b.mix_memory;
  g.n = n;
  g.ld = ld;
  g.nat = nat;
  g.nsh = nsh;
  g.ndim = nsh + 9 * nat;
  g.mem = mem;
  g.nbmax = (n + 63) / 64;
  g.kt = b.kt;
  g.damping = b.damping;
  g.econv = 1.0e-6 * b.accuracy;
  g.pconv = 2.0e-5 * b.accuracy;
  g.charge = pl.mol_charge[m];
  g.uhf = pl.mol_uhf[m];
  g.par = b.par;
  g.el = b.at_el + a0;
  g.sh_at = b.sh_at + s0;
  g.sh_ish = b.sh_ish + s0;
  g.sh_iao = b.sh_iao + s0;
  g.ao_at = b.ao_at + o0;
  g.ao_sh = b.ao_sh + o0;
  g.at_ish = b.at_ish + a0;
  g.at_iao = b.at_iao + a0;
  g.S = b.S + moff;
  g.H = b.H + moff;
  g.D = b.D + 3 * moff;
  g.Q = b.Q + 6 * moff;
  g.gam = b.gam + pl.mol_gam_off[m];
  g.sd = b.aes_sd + 3 * poff;
  g.dd = b.aes_dd + 9 * poff;
  g.sq = b.aes_sq + 6 * poff;
  g.n0sh = b.n0sh + s0;
  g.erep_at = b.erep + a0;
  g.d4_on = c.dev.d4.enabled;
  g.d4_cn = b.d4_cn + a0;
  g.d4_disp = b.d4_disp + poff;
  g.d4_e3 = b.d4_e3 + a0;
  g.d4_gw = b.d4_gw + TA_D4_MAXREF * (int64_t)a0;
  g.d4_dgw = b.d4_dgw + TA_D4_MAXREF * (int64_t)a0;
  g.edisp2 = b.edisp2 + m;
  g.edisp3 = b.edisp3 + m;
  g.qsh = b.qsh + s0;
  g.qat = b.qat + a0;
  g.dpat = b.dpat + 3 * (int64_t)a0;
  g.qpat = b.qpat + 6 * (int64_t)a0;
  g.qin = b.mix_qin + pl.mol_mix_off[m];
  g.dq = b.mix_dq + pl.mol_mix_off[m];
  g.dqlast = b.mix_dqlast + pl.mol_mix_off[m];
  g.qlast = b.mix_qlast + pl.mol_mix_off[m];
  g.dfm = b.mix_df + pl.mol_mixmem_off[m];
  g.um = b.mix_u + pl.mol_mixmem_off[m];
  g.amix = b.mix_a + (int64_t)m * mem * mem;
  g.beta = b.mix_beta + (int64_t)m * mem * (mem + 1);
  g.omega = b.mix_omega + (int64_t)m * mem;
  g.cvec = b.mix_c + (int64_t)m * mem;
  const int64_t boff = pl.mol_big_off[m];
  g.A = b.big_a + boff;
  g.X = b.big_c + boff;
  g.Y = b.big_t + boff;
  g.CT = b.big_ct + boff;
  double* p = b.big_s + pl.mol_scr_off[m];
  auto take = [&p](int64_t cnt) {
    double* q = p;
    p += cnt;
    return q;
  };
  const int64_t nbm = g.nbmax;
  g.scr = take(3 * (int64_t)n * n);
  g.vt = take((int64_t)TA_BIG_NP * ld);
  g.wt = take((int64_t)TA_BIG_NP * ld);
  g.wbt = take((int64_t)TA_BIG_NB * ld);
  g.part = take(nbm * nbm * 64);
  g.svp = take(nbm);
  g.dpart = take(nbm * 2 * TA_BIG_NP);
  g.part11 = take(11 * (int64_t)n);
  g.epart = take(nsh + nat);
  g.d = take(ld);
  g.e = take(ld);
  g.tau = take(ld);
  g.e2 = take(ld);
  g.lam = take(ld);
  g.prevlam = take(ld);
  g.v = take(ld);
  g.y = take(ld);
  g.focc = take(ld);
  g.fs = take(ld);
  g.vsh = take(nsh);
  g.vat = take(nat);
  g.vdp = take(3 * (int64_t)nat);
  g.vqp = take(6 * (int64_t)nat);
  g.ed4 = take(nat);
  g.dots = take(2 * TA_BIG_NP);
  g.tm = take(TA_BIG_NB * TA_BIG_NB);
  g.sc = take(GS_COUNT);
  g.ic = reinterpret_cast<int*>(p);
  g.runs = g.ic + GI_COUNT;
  GridHostStats hs;
  int rc = grid_scf(g, ctx->stream, c.max_iter, b.energy + m, b.eelec + m, b.erep_tot + m,
                    b.iterations + m, b.status + m, b.emo + o0, b.focc + o0,
                    b.coeff ? b.coeff + moff : nullptr, b.density ? b.density + moff : nullptr,
                    &hs, ctx->verbose,
                    ctx->grid_graph ? 1 : 0);
  if (rc) {
    set_error(ctx, "multi-workgroup SCF of molecule %d failed: %s", m,
              hipGetErrorString(hipGetLastError()));
    return TBLITE_AMD_ERR_HIP;
  }
  return TBLITE_AMD_OK;
}

// Run the whole pipeline for the planned batch.  xyz (host) is uploaded
// first when given; with xyz == nullptr the coordinates already resident in
// the batch buffer are used.  With wait == false the work is only enqueued:
// ev[4] on the context stream marks its end and finish_batch() collects the
// phase timings.
int run_batch(tblite_amd_context ctx, Calculator& c, const double* xyz, double acc,
              bool wait = true)
{
  Plan& pl = ctx->plan;
  DeviceBatch& db = ctx->batch;
  BatchDev& b = db.b;
  b.accuracy = acc;
  b.intcut = integral_cutoff(acc);
  hipStream_t st = ctx->stream;
  if (xyz)
    HIP_TRY(ctx, hipMemcpyAsync((void*)b.xyz, xyz, sizeof(double) * 3 * pl.tot_atoms,
                                hipMemcpyHostToDevice, st));
  HIP_TRY(ctx, hipEventRecord(ctx->ev[0], st));
  // zero the integral matrices: pairs beyond the cutoff keep zero blocks
  {
    const int64_t n = 11 * pl.tot_mat;
    int grid = (int)std::min<int64_t>((n + 255) / 256, 65536);
    if (n > 0) hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, st, b.S, pl.tot_mat, 0.0);
    if (n > 0) hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, st, b.H, pl.tot_mat, 0.0);
    if (n > 0) hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, st, b.D, 3 * pl.tot_mat, 0.0);
    if (n > 0) hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, st, b.Q, 6 * pl.tot_mat, 0.0);
  }
  hipLaunchKernelGGL(k_setup, dim3(pl.nmol), dim3(256), 0, st, b);
  HIP_TRY(ctx, hipGetLastError());
  // The pair planes and the three-body dispersion need only k_setup's radii,
  // CN and weights: on a multi-stream context they run on the first
  // auxiliary stream beside the neighbour list and the integrals (whose
  // kernels leave CUs room), and the SCF waits for both.
  const bool one = ctx->single_stream;
  hipStream_t ps = one ? st : ctx->aux[0];
  if (!one) {
    HIP_TRY(ctx, hipEventRecord(ctx->ev[5], st));
    HIP_TRY(ctx, hipStreamWaitEvent(ps, ctx->ev[5], 0));
  }
  if (!pl.pair_mol.empty())
    hipLaunchKernelGGL(k_setup_pairs, dim3((int)pl.pair_mol.size()), dim3(256), 0, ps, b,
                       db.d_pair_mol, db.d_pair_a0, db.d_pair_a1);
  HIP_TRY(ctx, hipGetLastError());
  // periodic cells: Ewald-summed ES2/AES2 and image-summed D4 planes, then
  // the three-body sites and energies (these need the zero-charge C6)
  if (!pl.pp_mol.empty())
    hipLaunchKernelGGL(k_pbc_pairs, dim3((int)pl.pp_mol.size()), dim3(256), 0, ps, b,
                       db.d_pp_mol, db.d_pp_a0, db.d_pp_a1);
  if (!pl.ps_mol.empty())
    hipLaunchKernelGGL(k_pbc_atm_sites, dim3((int)pl.ps_mol.size()), dim3(256), 0, ps, b,
                       db.d_ps_mol, db.d_ps_atom);
  if (!pl.pa_mol.empty())
    hipLaunchKernelGGL(k_pbc_atm, dim3((int)pl.pa_mol.size()), dim3(256), 0, ps, b, db.d_pa_mol,
                       db.d_pa_atom, db.d_pa_r0, db.d_pa_r1);
  HIP_TRY(ctx, hipGetLastError());
  // three-body dispersion (writes every atom's d4_e3 entry; needs the D4 CN
  // of k_setup and the zero-charge C6 planes of k_setup_pairs)
  const int natm = (int)pl.atm_mol.size();
  if (natm > 0) {
    hipLaunchKernelGGL(k_d4_atm, dim3(natm), dim3(256), 0, ps, b, db.d_atm_mol, db.d_atm_i0,
                       db.d_atm_i1);
    HIP_TRY(ctx, hipGetLastError());
  }
  if (!one) HIP_TRY(ctx, hipEventRecord(ctx->ev[6], ps));
  HIP_TRY(ctx, hipEventRecord(ctx->ev[1], st));
  hipLaunchKernelGGL(k_adjlist, dim3(pl.nmol), dim3(256), 0, st, b);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipEventRecord(ctx->ev[2], st));
  const int nwork = (int)pl.work_mol.size(), nsp = pl.nwork_sp;
  if (nsp > 0)
    hipLaunchKernelGGL(k_integrals<1>, dim3(nsp), dim3(pl.int_chunk), 0, st, b, db.d_work_mol,
                       db.d_work_kind, db.d_work_start);
  if (nwork > nsp)
    hipLaunchKernelGGL(k_integrals<2>, dim3(nwork - nsp), dim3(pl.int_chunk), 0, st, b,
                       db.d_work_mol + nsp, db.d_work_kind + nsp, db.d_work_start + nsp);
  // periodic cells: one workgroup per atom pair sums the pair's images in LDS
  const int npw = (int)pl.pw_mol.size(), npsp = pl.nwork_pbc_sp;
  if (npsp > 0)
    hipLaunchKernelGGL(k_pbc_integrals<1>, dim3(npsp), dim3(64), 0, st, b, db.d_pw_mol,
                       db.d_pw_i, db.d_pw_j);
  if (npw > npsp)
    hipLaunchKernelGGL(k_pbc_integrals<2>, dim3(npw - npsp), dim3(64), 0, st, b,
                       db.d_pw_mol + npsp, db.d_pw_i + npsp, db.d_pw_j + npsp);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipEventRecord(ctx->ev[3], st));
  // SCF buckets on the auxiliary streams, largest bucket first; on a
  // single-stream context one after the other, smallest first (they finish
  // in milliseconds and report early).  Every stream waits for the
  // integrals and for the pair planes / three-body term.
  if (!one) {
    for (int s = 0; s < kNumStreams; ++s) {
      HIP_TRY(ctx, hipStreamWaitEvent(ctx->aux[s], ctx->ev[3], 0));
      HIP_TRY(ctx, hipStreamWaitEvent(ctx->aux[s], ctx->ev[6], 0));
    }
    HIP_TRY(ctx, hipStreamWaitEvent(st, ctx->ev[6], 0));
  }
  const int nbk = (int)pl.buckets.size();
  for (int q = 0; q < nbk; ++q) {
    const int k = one ? nbk - 1 - q : q;
    const Plan::Bucket& bk = pl.buckets[k];
    hipStream_t as = one ? st : ctx->aux[k % kNumStreams];
    const int* mols = db.d_bucket_mols[k];
    const int cnt = (int)bk.mols.size();
    int rc;
    HIP_TRY(ctx, hipEventRecord(ctx->bev[2 * k], as));
    // per-worker-thread Jacobi items over the NT - 64 worker threads: MAXB
    // blocks of A (ld <= 32: 120 over 64; ld <= 64: 496 over 192; ld <= 96:
    // 1128 over 448) and up to MAXC two-AO pieces of one slot's CT rows (4, 6
    // and 8 threads per slot)
    if (bk.cell && !bk.lds)
      rc = launch_scf<TA_BIG_NT, 1, 0, 0, false, true>(ctx, as, b, mols, cnt, bk.ldmax, bk.natmax,
                                                       bk.nshmax);
    else if (bk.cell && bk.nt == 128)
      rc = launch_scf<128, 1, 2, 4, true, true>(ctx, as, b, mols, cnt, bk.ldmax, bk.natmax,
                                                bk.nshmax);
    else if (bk.cell && bk.nt == 256)
      rc = launch_scf<256, 1, 3, 6, true, true>(ctx, as, b, mols, cnt, bk.ldmax, bk.natmax,
                                                bk.nshmax);
    else if (bk.cell)
      rc = launch_scf<512, 2, 3, 6, true, true>(ctx, as, b, mols, cnt, bk.ldmax, bk.natmax,
                                                bk.nshmax);
    else if (!bk.lds)
      rc = launch_scf<TA_BIG_NT, 1, 0, 0, false>(ctx, as, b, mols, cnt, bk.ldmax, bk.natmax,
                                                 bk.nshmax);
    else if (bk.nt == 128)
      rc = launch_scf<128, 1, 2, 4, true>(ctx, as, b, mols, cnt, bk.ldmax, bk.natmax, bk.nshmax);
    else if (bk.nt == 256)
      rc = launch_scf<256, 1, 3, 6, true>(ctx, as, b, mols, cnt, bk.ldmax, bk.natmax, bk.nshmax);
    else
      rc = launch_scf<512, 2, 3, 6, true>(ctx, as, b, mols, cnt, bk.ldmax, bk.natmax, bk.nshmax);
    if (rc) return rc;
    HIP_TRY(ctx, hipEventRecord(ctx->bev[2 * k + 1], as));
  }
  // molecules above the single-workgroup size: one at a time across the GPU,
  // on the main stream while the buckets run on the auxiliary streams
  for (int m : pl.grid_mols) {
    int rc = run_grid_molecule(ctx, c, m);
    if (rc) return rc;
  }
  if (!one)
    for (int s = 0; s < kNumStreams; ++s) {
      HIP_TRY(ctx, hipEventRecord(ctx->aux_done[s], ctx->aux[s]));
      HIP_TRY(ctx, hipStreamWaitEvent(st, ctx->aux_done[s], 0));
    }
  HIP_TRY(ctx, hipEventRecord(ctx->ev[4], st));
  if (!wait) return TBLITE_AMD_OK;
  return finish_batch(ctx);
}

int finish_batch(tblite_amd_context ctx)
{
  Plan& pl = ctx->plan;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  float t01, t12, t23, t34, t04;
  HIP_TRY(ctx, hipEventElapsedTime(&t01, ctx->ev[0], ctx->ev[1]));
  HIP_TRY(ctx, hipEventElapsedTime(&t12, ctx->ev[1], ctx->ev[2]));
  HIP_TRY(ctx, hipEventElapsedTime(&t23, ctx->ev[2], ctx->ev[3]));
  HIP_TRY(ctx, hipEventElapsedTime(&t34, ctx->ev[3], ctx->ev[4]));
  HIP_TRY(ctx, hipEventElapsedTime(&t04, ctx->ev[0], ctx->ev[4]));
  ctx->timings[0] = t01;
  ctx->timings[1] = t12;
  ctx->timings[2] = t23;
  ctx->timings[3] = t34;
  ctx->timings[4] = t04;
  for (size_t k = 0; k < pl.buckets.size(); ++k) {
    float tb = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&tb, ctx->bev[2 * k], ctx->bev[2 * k + 1]));
    ctx->bucket_ms[k] = tb;
  }
  return TBLITE_AMD_OK;
}

int copy_results(tblite_amd_context ctx, int nmol, double* energies, int* iterations,
                 int* status);

// Analytic gradient of the batch's SCF result (after run_batch): the kernels
// of gradient.hip in stream order on the context stream.
int run_gradient(tblite_amd_context ctx)
{
  Plan& pl = ctx->plan;
  DeviceBatch& db = ctx->batch;
  const BatchDev& b = db.b;
  if (!db.grad) {
    set_error(ctx, "internal: the batch holds no gradient buffers");
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  hipStream_t st = ctx->stream;
  hipLaunchKernelGGL(k_grad_pot, dim3(pl.nmol), dim3(256), 0, st, b);
  const int ntiles = (int)(db.h_wtiles.size() / 3);
  if (ntiles > 0)
    hipLaunchKernelGGL(k_grad_wdens, dim3(ntiles), dim3(256), 0, st, b, db.d_wtiles);
  const int nwork = (int)pl.work_mol.size(), nsp = pl.nwork_sp;
  if (nsp > 0)
    hipLaunchKernelGGL(k_grad_h0<1>, dim3(nsp), dim3(pl.int_chunk), 0, st, b, db.d_work_mol,
                       db.d_work_kind, db.d_work_start);
  if (nwork > nsp)
    hipLaunchKernelGGL(k_grad_h0<2>, dim3(nwork - nsp), dim3(pl.int_chunk), 0, st, b,
                       db.d_work_mol + nsp, db.d_work_kind + nsp, db.d_work_start + nsp);
  const int npw = (int)pl.pair_mol.size();
  hipLaunchKernelGGL(k_grad_pairs, dim3(npw), dim3(256), 0, st, b, db.d_pair_mol, db.d_pair_a0,
                     db.d_pair_a1);
  const int natm = (int)pl.atm_mol.size();
  if (natm > 0) {
    hipLaunchKernelGGL(k_grad_c6d, dim3(npw), dim3(256), 0, st, b, db.d_pair_mol, db.d_pair_a0,
                       db.d_pair_a1);
    // per-workgroup LDS accumulation of the three-body terms when 4 natmax
    // doubles fit (LDS per workgroup of the device)
    int natmax = 0;
    for (int m = 0; m < pl.nmol; ++m) natmax = std::max(natmax, pl.mol_nat[m]);
    const size_t lds_bytes = sizeof(double) * 4 * (size_t)natmax;
    const int use_lds = lds_bytes <= std::min<size_t>(ctx->lds_limit, 64 * 1024) ? 1 : 0;
    hipLaunchKernelGGL(k_grad_atm, dim3(natm), dim3(256), use_lds ? lds_bytes : 0, st, b,
                       db.d_atm_mol, db.d_atm_i0, db.d_atm_i1, use_lds);
  }
  hipLaunchKernelGGL(k_grad_cn, dim3(npw), dim3(256), 0, st, b, db.d_pair_mol, db.d_pair_a0,
                     db.d_pair_a1);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipStreamSynchronize(st));
  return TBLITE_AMD_OK;
}

// Gradients (3, total atoms) of the last gradient batch and, when sigmas is
// given, each molecule's strain derivative sum_a g_a (x) R_a (column-major
// 3 x 3 per molecule).  imol >= 0: that molecule only.
int copy_gradients(tblite_amd_context ctx, int imol, double* gradients, double* sigmas)
{
  const Plan& pl = ctx->plan;
  const BatchDev& b = ctx->batch.b;
  if (!ctx->batch.grad) {
    set_error(ctx, "the last batch was run without gradients");
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  const int m0 = imol < 0 ? 0 : imol, m1 = imol < 0 ? pl.nmol : imol + 1;
  const int64_t a_lo = pl.mol_atom_off[m0];
  const int64_t a_hi = pl.mol_atom_off[m1 - 1] + pl.mol_nat[m1 - 1];
  const int64_t na = a_hi - a_lo;
  std::vector<double> g(3 * na), x(3 * na);
  HIP_TRY(ctx, hipMemcpy(g.data(), b.grad + 3 * a_lo, sizeof(double) * 3 * na, hipMemcpyDeviceToHost));
  HIP_TRY(ctx, hipMemcpy(x.data(), b.xyz + 3 * a_lo, sizeof(double) * 3 * na, hipMemcpyDeviceToHost));
  if (gradients) std::memcpy(gradients, g.data(), sizeof(double) * 3 * na);
  if (sigmas)
    for (int m = m0; m < m1; ++m) {
      const int64_t o = pl.mol_atom_off[m] - a_lo;
      double* sg = sigmas + 9 * (int64_t)(m - m0);
      for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) {
          double acc = 0.0;
          for (int i = 0; i < pl.mol_nat[m]; ++i) acc += g[3 * (o + i) + p] * x[3 * (o + i) + q];
          sg[p + 3 * q] = acc;
        }
    }
  return TBLITE_AMD_OK;
}

int flat_singlepoint(tblite_amd_context ctx, Calculator& c, int nmol, const int* mol_nat,
                     const int* num, const double* xyz, const double* charge, const int* uhf,
                     double acc, double* energies, int* iterations, int* status,
                     bool wfn_arrays = false, bool grad = false, const double* lattice = nullptr,
                     const char* is_cell = nullptr)
{
  if (!ctx) return TBLITE_AMD_ERR_ARGUMENT;
  if (nmol < 1 || !mol_nat || !num || !xyz || !energies) {
    set_error(ctx, "invalid arguments to batched singlepoint");
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  if (lattice && grad) {
    set_error(ctx, "analytic gradients of periodic cells are not available");
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  int rc = ensure_params_on_device(ctx, c);
  if (rc) return rc;
  ctx->have_batch = false;
  ctx->prepared = nullptr;
  rc = make_plan(ctx, c, nmol, mol_nat, num, charge, uhf, acc, ctx->plan, xyz, lattice, is_cell);
  if (rc) return rc;
  rc = upload_plan(ctx, c, ctx->plan, ctx->batch, wfn_arrays, nullptr, nullptr, grad);
  if (rc) return rc;
  rc = run_batch(ctx, c, xyz, acc);
  if (rc) return rc;
  if (grad) {
    rc = run_gradient(ctx);
    if (rc) return rc;
  }
  ctx->have_batch = true;
  return copy_results(ctx, nmol, energies, iterations, status);
}

int copy_results(tblite_amd_context ctx, int nmol, double* energies, int* iterations,
                 int* status)
{
  const BatchDev& b = ctx->batch.b;
  std::vector<int> st(nmol), it(nmol);
  if (energies)
    HIP_TRY(ctx, hipMemcpy(energies, b.energy, sizeof(double) * nmol, hipMemcpyDeviceToHost));
  HIP_TRY(ctx, hipMemcpy(st.data(), b.status, sizeof(int) * nmol, hipMemcpyDeviceToHost));
  HIP_TRY(ctx, hipMemcpy(it.data(), b.iterations, sizeof(int) * nmol, hipMemcpyDeviceToHost));
  if (status) std::memcpy(status, st.data(), sizeof(int) * nmol);
  if (iterations) std::memcpy(iterations, it.data(), sizeof(int) * nmol);
  for (int m = 0; m < nmol; ++m)
    if (st[m] != 0) {
      set_error(ctx, "molecule %d finished with status %d after %d iterations", m, st[m], it[m]);
      return TBLITE_AMD_ERR_NOT_CONVERGED;
    }
  return TBLITE_AMD_OK;
}

int check_batch(tblite_amd_context ctx, int imol)
{
  if (!ctx) return TBLITE_AMD_ERR_ARGUMENT;
  if (!ctx->have_batch) {
    set_error(ctx, "no batch has been run on this context");
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  if (imol < 0 || imol >= ctx->plan.nmol) {
    set_error(ctx, "molecule index %d out of range", imol);
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  return TBLITE_AMD_OK;
}

}  // namespace

// ================================================================= C ABI
static tblite_amd_context make_context(int device, bool single_stream);

extern "C" {

int tblite_amd_param_default(tblite_amd_param* param)
{
  if (!param) return TBLITE_AMD_ERR_ARGUMENT;
  tblite_amd_fill_gfn2_default(param);
  return TBLITE_AMD_OK;
}

tblite_amd_context tblite_amd_new_context(int device) { return make_context(device, false); }

}  // extern "C"

// A context on one device; single_stream: no auxiliary streams, every kernel
// on the main stream.
static tblite_amd_context make_context(int device, bool single_stream)
{
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1) return nullptr;
  if (device < 0 && hipGetDevice(&device) != hipSuccess) return nullptr;
  if (device >= count) return nullptr;
  if (hipSetDevice(device) != hipSuccess) return nullptr;
  tblite_amd_context ctx = new (std::nothrow) tblite_amd_context_s();
  if (!ctx) return nullptr;
  ctx->device = device;
  ctx->single_stream = single_stream;
  {
    // the fixed workspace's device address, once per device (a failure only
    // disables the no-allocation entry point, which reports it)
    hipError_t se = hipSuccess;
    ctx->fixed_ws = scratch_address(device, &se);
  }
  bool ok = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) == hipSuccess;
  for (int s = 0; s < kNumStreams && ok && !single_stream; ++s)
    ok = hipStreamCreateWithFlags(&ctx->aux[s], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&ctx->aux_done[s], hipEventDisableTiming) == hipSuccess;
  for (int e = 0; e < 8 && ok; ++e) ok = hipEventCreate(&ctx->ev[e]) == hipSuccess;
  for (int e = 0; e < 2 * kMaxBuckets && ok; ++e) ok = hipEventCreate(&ctx->bev[e]) == hipSuccess;
  int lds = 0, ncu = 0;
  if (ok) ok = hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess;
  if (ok) ok = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess;
  ctx->wg_slots = std::max(1, 2 * ncu);
  ctx->lds_limit = (size_t)lds;
  const char* pe = std::getenv("TBLITE_AMD_PROFILE");
  ctx->profile = pe && pe[0] == '1';
  const char* pv = std::getenv("TBLITE_AMD_VERBOSE");
  ctx->verbose = pv && pv[0] == '1';
  const char* pg = std::getenv("TBLITE_AMD_GRID_MIN_LD");
  if (pg) ctx->grid_min_ld = std::max(16, std::min(TA_BIG_MAXLD, std::atoi(pg)));
  const char* pgg = std::getenv("TBLITE_AMD_GRID_GRAPH");
  ctx->grid_graph = !(pgg && pgg[0] == '0');
  const char* pb = std::getenv("TBLITE_AMD_BIG_LDS_MIN_KB");
  if (pb) ctx->big_lds_min = (size_t)std::max(0, std::min(160, std::atoi(pb))) * 1024;
  const char* pfb = std::getenv("TBLITE_AMD_FINE_BUCKETS");
  if (pfb) ctx->fine_buckets = pfb[0] == '1';
  const char* pa = std::getenv("TBLITE_AMD_JACOBI_ADAPT");
  if (pa) ctx->jacobi_adapt = std::max(0.0, std::min(10.0, std::atof(pa)));
  const char* ps = std::getenv("TBLITE_AMD_JACOBI_SKIP");
  if (ps) ctx->jacobi_skip = std::max(0.0, std::min(1e-6, std::atof(ps)));
  const char* pj = std::getenv("TBLITE_AMD_JACOBI_TOL");
  if (pj) ctx->jacobi_tol = std::max(1e-12, std::min(1e-2, std::atof(pj)));
  const char* pe2 = std::getenv("TBLITE_AMD_LDS_EIG");
  if (pe2) ctx->lds_eig = std::strcmp(pe2, "householder") == 0 ? 1 : 0;
  const char* pbe = std::getenv("TBLITE_AMD_BIG_EIG");
  if (pbe) ctx->big_eig = std::strcmp(pbe, "band") == 0 ? 1 : 0;
  const char* pl = std::getenv("TBLITE_AMD_LDS_MAX_LD");
  if (pl) ctx->lds_max_ld = std::max(16, std::min(kLdsMaxLd, std::atoi(pl) / 16 * 16));
  if (!ok) {
    tblite_amd_delete_context(&ctx);
    return nullptr;
  }
  return ctx;
}

extern "C" {

void tblite_amd_delete_context(tblite_amd_context* pctx)
{
  if (!pctx || !*pctx) return;
  tblite_amd_context ctx = *pctx;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (int s = 0; s < kNumStreams; ++s) {
    if (ctx->aux[s]) {
      (void)hipStreamSynchronize(ctx->aux[s]);
      (void)hipStreamDestroy(ctx->aux[s]);
    }
    if (ctx->aux_done[s]) (void)hipEventDestroy(ctx->aux_done[s]);
  }
  for (int e = 0; e < 8; ++e)
    if (ctx->ev[e]) (void)hipEventDestroy(ctx->ev[e]);
  for (int e = 0; e < 2 * kMaxBuckets; ++e)
    if (ctx->bev[e]) (void)hipEventDestroy(ctx->bev[e]);
  if (ctx->arena.base) (void)hipFree(ctx->arena.base);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  *pctx = nullptr;
}

int tblite_amd_get_error(tblite_amd_context ctx, char* buffer, int buffer_size)
{
  if (!ctx || !buffer || buffer_size < 1) return TBLITE_AMD_ERR_ARGUMENT;
  std::snprintf(buffer, (size_t)buffer_size, "%s", ctx->error.c_str());
  return TBLITE_AMD_OK;
}

tblite_amd_structure tblite_amd_new_structure(tblite_amd_context ctx, int nat, const int* num,
                                              const double* xyz, double charge, int uhf)
{
  if (!ctx) return nullptr;
  if (nat < 1 || !num || !xyz) {
    set_error(ctx, "invalid structure arguments");
    return nullptr;
  }
  for (int i = 0; i < 3 * nat; ++i)
    if (!std::isfinite(xyz[i])) {
      set_error(ctx, "non-finite coordinate in structure");
      return nullptr;
    }
  tblite_amd_structure s = new (std::nothrow) tblite_amd_structure_s();
  if (!s) return nullptr;
  s->ctx = ctx;
  s->nat = nat;
  s->num.assign(num, num + nat);
  s->xyz.assign(xyz, xyz + 3 * nat);
  s->charge = charge;
  s->uhf = uhf;
  return s;
}

tblite_amd_structure tblite_amd_new_structure_periodic(tblite_amd_context ctx, int nat,
                                                       const int* num, const double* xyz,
                                                       double charge, int uhf,
                                                       const double* lattice,
                                                       const int* periodic)
{
  if (!ctx) return nullptr;
  int np = 0;
  for (int k = 0; periodic && k < 3; ++k) np += periodic[k] ? 1 : 0;
  if (np != 0 && np != 3) {
    set_error(ctx, "only three-dimensional periodicity is supported (periodic = %d %d %d)",
              periodic[0], periodic[1], periodic[2]);
    return nullptr;
  }
  if (np == 3 && !lattice) {
    set_error(ctx, "periodic structure without lattice vectors");
    return nullptr;
  }
  if (np == 3) {
    for (int k = 0; k < 9; ++k)
      if (!std::isfinite(lattice[k])) {
        set_error(ctx, "non-finite lattice vector");
        return nullptr;
      }
    CellTables ct;
    std::vector<double> probe(3, 0.0);
    if (!cell_tables(lattice, 1, probe.data(), ct)) {
      set_error(ctx, "degenerate lattice (zero cell volume)");
      return nullptr;
    }
  }
  tblite_amd_structure s = tblite_amd_new_structure(ctx, nat, num, xyz, charge, uhf);
  if (!s) return nullptr;
  if (np == 3) {
    s->periodic = true;
    std::memcpy(s->lat, lattice, sizeof(double) * 9);
  }
  return s;
}

void tblite_amd_delete_structure(tblite_amd_structure* mol)
{
  if (!mol || !*mol) return;
  delete *mol;
  *mol = nullptr;
}

int tblite_amd_get_structure_dims(tblite_amd_structure mol, tblite_amd_calculator calc, int* nat,
                                  int* nsh, int* nao)
{
  if (!mol) return TBLITE_AMD_ERR_ARGUMENT;
  // the calculator's element table (shell layout per element), else the
  // built-in GFN2 data
  tblite_amd_param def;
  const tblite_amd_param* p = &def;
  if (calc) p = &calc->c.param;
  else tblite_amd_fill_gfn2_default(&def);
  int ns = 0, na = 0;
  for (int i = 0; i < mol->nat; ++i) {
    int e = -1;
    for (int k = 0; k < p->nelem; ++k)
      if (p->elem[k].z == mol->num[i]) e = k;
    if (e < 0) {
      set_error(mol->ctx, "atom %d: element Z=%d has no parameters", i, mol->num[i]);
      return TBLITE_AMD_ERR_ARGUMENT;
    }
    ns += p->elem[e].nsh;
    for (int s = 0; s < p->elem[e].nsh; ++s) na += 2 * p->elem[e].ang[s] + 1;
  }
  if (nat) *nat = mol->nat;
  if (nsh) *nsh = ns;
  if (nao) *nao = na;
  return TBLITE_AMD_OK;
}

tblite_amd_calculator tblite_amd_new_gfn2_calculator(tblite_amd_context ctx,
                                                     const tblite_amd_param* param)
{
  if (!ctx) return nullptr;
  tblite_amd_calculator c = new (std::nothrow) tblite_amd_calculator_s();
  if (!c) return nullptr;
  c->ctx = ctx;
  if (param) c->c.param = *param;
  else tblite_amd_fill_gfn2_default(&c->c.param);
  std::string err;
  if (!build_calculator(c->c, err)) {
    set_error(ctx, "%s", err.c_str());
    delete c;
    return nullptr;
  }
  return c;
}

void tblite_amd_delete_calculator(tblite_amd_calculator* calc)
{
  if (!calc || !*calc) return;
  if ((*calc)->c.d_par) (void)hipFree((*calc)->c.d_par);
  delete *calc;
  *calc = nullptr;
}

int tblite_amd_set_calculator_accuracy(tblite_amd_calculator calc, double accuracy)
{
  if (!calc || !(accuracy > 0.0)) return TBLITE_AMD_ERR_ARGUMENT;
  calc->c.accuracy = accuracy;
  return TBLITE_AMD_OK;
}

int tblite_amd_set_calculator_max_iter(tblite_amd_calculator calc, int max_iter)
{
  if (!calc || max_iter < 1) return TBLITE_AMD_ERR_ARGUMENT;
  calc->c.max_iter = max_iter;
  return TBLITE_AMD_OK;
}

int tblite_amd_set_calculator_mixer_damping(tblite_amd_calculator calc, double damping)
{
  if (!calc || !(damping > 0.0) || damping > 1.0) return TBLITE_AMD_ERR_ARGUMENT;
  calc->c.damping = damping;
  return TBLITE_AMD_OK;
}

int tblite_amd_set_calculator_temperature(tblite_amd_calculator calc, double kelvin)
{
  if (!calc || !(kelvin > 0.0)) return TBLITE_AMD_ERR_ARGUMENT;
  calc->c.etemp = kelvin;
  return TBLITE_AMD_OK;
}

tblite_amd_wavefunction tblite_amd_new_wavefunction(tblite_amd_context ctx)
{
  if (!ctx) return nullptr;
  tblite_amd_wavefunction w = new (std::nothrow) tblite_amd_wavefunction_s();
  if (w) w->ctx = ctx;
  return w;
}

void tblite_amd_delete_wavefunction(tblite_amd_wavefunction* wfn)
{
  if (!wfn || !*wfn) return;
  delete *wfn;
  *wfn = nullptr;
}

int tblite_amd_get_wavefunction_array(tblite_amd_wavefunction wfn, const char* name, double* out,
                                      int* count)
{
  if (!wfn || !name || !count) return TBLITE_AMD_ERR_ARGUMENT;
  const std::vector<double>* v = nullptr;
  const std::string n(name);
  if (n == "qsh") v = &wfn->qsh;
  else if (n == "qat") v = &wfn->qat;
  else if (n == "dpat") v = &wfn->dpat;
  else if (n == "qpat") v = &wfn->qpat;
  else if (n == "emo") v = &wfn->emo;
  else if (n == "focc") v = &wfn->focc;
  else if (n == "n0sh") v = &wfn->n0sh;
  else if (n == "n0at") v = &wfn->n0at;
  else if (n == "coeff") v = &wfn->coeff;
  else if (n == "density") v = &wfn->density;
  else {
    set_error(wfn->ctx, "unknown wavefunction array '%s'", name);
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  if (v->empty()) {
    set_error(wfn->ctx, "wavefunction array '%s' is empty: no singlepoint has filled this "
              "wavefunction", name);
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  *count = (int)v->size();
  if (out && !v->empty()) std::memcpy(out, v->data(), sizeof(double) * v->size());
  return TBLITE_AMD_OK;
}

// Fill a wavefunction handle from molecule imol of the last batch.
static int fill_wavefunction(tblite_amd_context ctx, int imol, tblite_amd_wavefunction w)
{
  const Plan& pl = ctx->plan;
  const BatchDev& b = ctx->batch.b;
  const int nat = pl.mol_nat[imol], nsh = pl.mol_nsh[imol], nao = pl.mol_nao[imol];
  const int a0 = pl.mol_atom_off[imol], s0 = pl.mol_sh_off[imol], o0 = pl.mol_ao_off[imol];
  w->nat = nat;
  w->nsh = nsh;
  w->nao = nao;
  w->qsh.resize(nsh);
  w->qat.resize(nat);
  w->dpat.resize(3 * nat);
  w->qpat.resize(6 * nat);
  w->emo.resize(nao);
  w->focc.resize(nao);
  w->n0sh.resize(nsh);
  w->n0at.assign(nat, 0.0);
  HIP_TRY(ctx, hipMemcpy(w->qsh.data(), b.qsh + s0, 8 * nsh, hipMemcpyDeviceToHost));
  HIP_TRY(ctx, hipMemcpy(w->qat.data(), b.qat + a0, 8 * nat, hipMemcpyDeviceToHost));
  HIP_TRY(ctx, hipMemcpy(w->dpat.data(), b.dpat + 3 * a0, 8 * 3 * nat, hipMemcpyDeviceToHost));
  HIP_TRY(ctx, hipMemcpy(w->qpat.data(), b.qpat + 6 * a0, 8 * 6 * nat, hipMemcpyDeviceToHost));
  HIP_TRY(ctx, hipMemcpy(w->emo.data(), b.emo + o0, 8 * nao, hipMemcpyDeviceToHost));
  HIP_TRY(ctx, hipMemcpy(w->focc.data(), b.focc + o0, 8 * nao, hipMemcpyDeviceToHost));
  HIP_TRY(ctx, hipMemcpy(w->n0sh.data(), b.n0sh + s0, 8 * nsh, hipMemcpyDeviceToHost));
  for (int s = 0; s < nsh; ++s) w->n0at[pl.sh_at[s0 + s]] += w->n0sh[s];
  const int64_t moff = pl.mol_mat_off[imol];
  const size_t n2 = (size_t)nao * nao;
  w->coeff.clear();
  w->density.clear();
  if (b.coeff) {
    w->coeff.resize(n2);
    HIP_TRY(ctx, hipMemcpy(w->coeff.data(), b.coeff + moff, 8 * n2, hipMemcpyDeviceToHost));
  }
  if (b.density) {
    w->density.resize(n2);
    HIP_TRY(ctx, hipMemcpy(w->density.data(), b.density + moff, 8 * n2, hipMemcpyDeviceToHost));
  }
  return TBLITE_AMD_OK;
}

int tblite_amd_xtb_singlepoint(tblite_amd_context ctx, tblite_amd_structure mol,
                               tblite_amd_calculator calc, tblite_amd_wavefunction wfn,
                               double accuracy, double* energy, int* iterations, int verbosity)
{
  if (!ctx || !mol || !calc || !energy) return TBLITE_AMD_ERR_ARGUMENT;
  const double acc = accuracy > 0.0 ? accuracy : calc->c.accuracy;
  int st = 0, it = 0;
  int rc = flat_singlepoint(ctx, calc->c, 1, &mol->nat, mol->num.data(), mol->xyz.data(),
                            &mol->charge, &mol->uhf, acc, energy, &it, &st, wfn != nullptr, false,
                            mol->periodic ? mol->lat : nullptr);
  if (iterations) *iterations = it;
  if (ctx->have_batch && wfn) {
    int rc2 = fill_wavefunction(ctx, 0, wfn);
    if (rc2) return rc2;
  }
  if (verbosity > 0 && ctx->have_batch)
    std::printf("tblite_amd: E = %.12f Eh after %d iterations (status %d)\n", *energy, it, st);
  if (rc == TBLITE_AMD_ERR_NOT_CONVERGED) {
    if (st & TA_STATUS_CHOLESKY) return TBLITE_AMD_ERR_CHOLESKY;
    if (st & TA_STATUS_EIGEN) return TBLITE_AMD_ERR_EIGEN;
    if (st & TA_STATUS_NAN) return TBLITE_AMD_ERR_NAN;
    if (st & TA_STATUS_MIXER) return TBLITE_AMD_ERR_MIXER;
  }
  return rc;
}

int tblite_amd_xtb_singlepoint_batch(tblite_amd_context ctx, int nmol,
                                     const tblite_amd_structure* mols,
                                     tblite_amd_calculator calc, tblite_amd_wavefunction* wfns,
                                     double accuracy, double* energies, int* iterations,
                                     int* status)
{
  if (!ctx || nmol < 1 || !mols || !calc || !energies) return TBLITE_AMD_ERR_ARGUMENT;
  std::vector<int> mol_nat(nmol), num, uhf(nmol);
  std::vector<double> xyz, charge(nmol), lat(9 * (size_t)nmol, 0.0);
  std::vector<char> is_cell(nmol, 0);
  bool any_cell = false;
  for (int m = 0; m < nmol; ++m) {
    if (!mols[m]) {
      set_error(ctx, "structure %d is NULL", m);
      return TBLITE_AMD_ERR_ARGUMENT;
    }
    if (mols[m]->periodic) {
      any_cell = true;
      is_cell[m] = 1;
      std::memcpy(lat.data() + 9 * (size_t)m, mols[m]->lat, sizeof(double) * 9);
    }
    mol_nat[m] = mols[m]->nat;
    num.insert(num.end(), mols[m]->num.begin(), mols[m]->num.end());
    xyz.insert(xyz.end(), mols[m]->xyz.begin(), mols[m]->xyz.end());
    charge[m] = mols[m]->charge;
    uhf[m] = mols[m]->uhf;
  }
  const double acc = accuracy > 0.0 ? accuracy : calc->c.accuracy;
  bool any_wfn = false;
  for (int m = 0; wfns && m < nmol; ++m) any_wfn = any_wfn || wfns[m] != nullptr;
  int rc = flat_singlepoint(ctx, calc->c, nmol, mol_nat.data(), num.data(), xyz.data(),
                            charge.data(), uhf.data(), acc, energies, iterations, status, any_wfn,
                            false, any_cell ? lat.data() : nullptr,
                            any_cell ? is_cell.data() : nullptr);
  if (ctx->have_batch && wfns)
    for (int m = 0; m < nmol; ++m)
      if (wfns[m]) {
        int rc2 = fill_wavefunction(ctx, m, wfns[m]);
        if (rc2) return rc2;
      }
  return rc;
}

int tblite_amd_singlepoint_flat(tblite_amd_context ctx, tblite_amd_calculator calc, int nmol,
                                const int* mol_nat, const int* num, const double* xyz,
                                const double* charge, const int* uhf, double accuracy,
                                double* energies, int* iterations, int* status)
{
  if (!ctx || !calc) return TBLITE_AMD_ERR_ARGUMENT;
  const double acc = accuracy > 0.0 ? accuracy : calc->c.accuracy;
  return flat_singlepoint(ctx, calc->c, nmol, mol_nat, num, xyz, charge, uhf, acc, energies,
                          iterations, status);
}

int tblite_amd_singlepoint_flat_periodic(tblite_amd_context ctx, tblite_amd_calculator calc,
                                         int nmol, const int* mol_nat, const int* num,
                                         const double* xyz, const double* lattice,
                                         const double* charge, const int* uhf, double accuracy,
                                         double* energies, int* iterations, int* status)
{
  if (!ctx || !calc) return TBLITE_AMD_ERR_ARGUMENT;
  if (!lattice) {
    set_error(ctx, "periodic batch without lattice vectors");
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  for (int k = 0; k < 9 * nmol; ++k)
    if (!std::isfinite(lattice[k])) {
      set_error(ctx, "non-finite lattice vector");
      return TBLITE_AMD_ERR_ARGUMENT;
    }
  const double acc = accuracy > 0.0 ? accuracy : calc->c.accuracy;
  return flat_singlepoint(ctx, calc->c, nmol, mol_nat, num, xyz, charge, uhf, acc, energies,
                          iterations, status, false, false, lattice);
}

int tblite_amd_singlepoint_flat_gradient(tblite_amd_context ctx, tblite_amd_calculator calc,
                                         int nmol, const int* mol_nat, const int* num,
                                         const double* xyz, const double* charge, const int* uhf,
                                         double accuracy, double* energies, double* gradients,
                                         double* sigmas, int* iterations, int* status)
{
  if (!ctx || !calc) return TBLITE_AMD_ERR_ARGUMENT;
  if (!gradients) {
    set_error(ctx, "singlepoint_flat_gradient: gradients must not be null");
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  const double acc = accuracy > 0.0 ? accuracy : calc->c.accuracy;
  int rc = flat_singlepoint(ctx, calc->c, nmol, mol_nat, num, xyz, charge, uhf, acc, energies,
                            iterations, status, false, true);
  if (rc && rc != TBLITE_AMD_ERR_NOT_CONVERGED) return rc;
  const int rc2 = copy_gradients(ctx, -1, gradients, sigmas);
  return rc2 ? rc2 : rc;
}

int tblite_amd_xtb_singlepoint_gradient(tblite_amd_context ctx, tblite_amd_structure mol,
                                        tblite_amd_calculator calc, tblite_amd_wavefunction wfn,
                                        double accuracy, double* energy, double* gradient,
                                        double* sigma, int* iterations, int verbosity)
{
  if (!ctx || !mol || !calc || !energy || !gradient) return TBLITE_AMD_ERR_ARGUMENT;
  if (mol->periodic) {
    set_error(ctx, "analytic gradients of periodic cells are not available");
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  const double acc = accuracy > 0.0 ? accuracy : calc->c.accuracy;
  int st = 0, it = 0;
  int rc = flat_singlepoint(ctx, calc->c, 1, &mol->nat, mol->num.data(), mol->xyz.data(),
                            &mol->charge, &mol->uhf, acc, energy, &it, &st, wfn != nullptr, true);
  if (iterations) *iterations = it;
  if (rc && rc != TBLITE_AMD_ERR_NOT_CONVERGED) return rc;
  int rc2 = copy_gradients(ctx, 0, gradient, sigma);
  if (rc2) return rc2;
  if (wfn) {
    rc2 = fill_wavefunction(ctx, 0, wfn);
    if (rc2) return rc2;
  }
  if (verbosity > 0) {
    double gmax = 0.0;
    for (int i = 0; i < 3 * mol->nat; ++i) gmax = std::max(gmax, std::fabs(gradient[i]));
    std::printf("tblite_amd: E = %.12f Eh after %d iterations (status %d), max |dE/dR| = %.3e Eh/bohr\n",
                *energy, it, st, gmax);
  }
  if (rc == TBLITE_AMD_ERR_NOT_CONVERGED) {
    if (st & TA_STATUS_CHOLESKY) return TBLITE_AMD_ERR_CHOLESKY;
    if (st & TA_STATUS_EIGEN) return TBLITE_AMD_ERR_EIGEN;
    if (st & TA_STATUS_NAN) return TBLITE_AMD_ERR_NAN;
    if (st & TA_STATUS_MIXER) return TBLITE_AMD_ERR_MIXER;
  }
  return rc;
}

int tblite_amd_batch_get_gradient(tblite_amd_context ctx, int imol, double* gradient,
                                  double* sigma)
{
  int rc = check_batch(ctx, imol);
  if (rc) return rc;
  return copy_gradients(ctx, imol, gradient, sigma);
}

int tblite_amd_singlepoint_flat_noalloc(tblite_amd_context ctx, tblite_amd_calculator calc,
                                        int nmol, const int* mol_nat, const int* num,
                                        const double* xyz, const double* charge, const int* uhf,
                                        double accuracy, double* energies, int* iterations,
                                        int* status)
{
  if (!ctx || !calc) return TBLITE_AMD_ERR_ARGUMENT;
  if (nmol < 1 || !mol_nat || !num || !xyz || !energies) {
    set_error(ctx, "invalid arguments to the no-allocation singlepoint");
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (!ctx->fixed_ws) {
    hipError_t e = hipSuccess;
    (void)scratch_address(ctx->device, &e);
    set_error(ctx, "fixed device workspace unavailable on device %d: %s", ctx->device,
              hipGetErrorString(e));
    return TBLITE_AMD_ERR_HIP;
  }
  Calculator& c = calc->c;
  const double acc = accuracy > 0.0 ? accuracy : c.accuracy;
  // the parameter block first, then the batch; nothing else is allocated
  Arena ws;
  ws.base = ctx->fixed_ws;
  ws.cap = TBLITE_AMD_NOALLOC_SCRATCH_BYTES;
  ws.used = 0;
  DevParams* dpar = carve<DevParams>(ws, 1);
  Arena rest;
  rest.base = ws.base + ws.used;
  rest.cap = ws.cap - ws.used;
  ctx->have_batch = false;
  ctx->prepared = nullptr;
  // fixed routing by size: the multi-workgroup path (host loop with graph
  // capture) is not taken here
  const bool ag = ctx->auto_grid;
  ctx->auto_grid = false;
  int rc = make_plan(ctx, c, nmol, mol_nat, num, charge, uhf, acc, ctx->plan);
  ctx->auto_grid = ag;
  if (rc) return rc;
  if (!ctx->plan.grid_mols.empty()) {
    const int m = ctx->plan.grid_mols[0];
    set_error(ctx, "molecule %d (%d AOs) needs the multi-workgroup path, which the "
              "no-allocation entry point does not run", m, ctx->plan.mol_nao[m]);
    return TBLITE_AMD_ERR_TOO_LARGE;
  }
  HIP_TRY(ctx, hipMemcpyAsync(dpar, &c.dev, sizeof(DevParams), hipMemcpyHostToDevice, ctx->stream));
  rc = upload_plan(ctx, c, ctx->plan, ctx->batch, false, &rest, dpar);
  if (rc) return rc;
  rc = run_batch(ctx, c, xyz, acc);
  if (rc) return rc;
  ctx->have_batch = true;
  return copy_results(ctx, nmol, energies, iterations, status);
}

}  // extern "C"

namespace {
int prepare_batch(tblite_amd_context ctx, tblite_amd_calculator calc, int nmol, const int* mol_nat,
                  const int* num, const double* charge, const int* uhf, double accuracy, bool grad)
{
  if (!ctx || !calc) return TBLITE_AMD_ERR_ARGUMENT;
  if (nmol < 1 || !mol_nat || !num) {
    set_error(ctx, "invalid arguments to batch_prepare");
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  int rc = ensure_params_on_device(ctx, calc->c);
  if (rc) return rc;
  ctx->have_batch = false;
  ctx->prepared = nullptr;
  const double acc = accuracy > 0.0 ? accuracy : calc->c.accuracy;
  rc = make_plan(ctx, calc->c, nmol, mol_nat, num, charge, uhf, acc, ctx->plan);
  if (rc) return rc;
  rc = upload_plan(ctx, calc->c, ctx->plan, ctx->batch, false, nullptr, nullptr, grad);
  if (rc) return rc;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  ctx->prepared = calc;
  ctx->prepared_acc = acc;
  return TBLITE_AMD_OK;
}
}  // namespace

extern "C" {

int tblite_amd_batch_prepare(tblite_amd_context ctx, tblite_amd_calculator calc, int nmol,
                             const int* mol_nat, const int* num, const double* charge,
                             const int* uhf, double accuracy)
{
  return prepare_batch(ctx, calc, nmol, mol_nat, num, charge, uhf, accuracy, false);
}

int tblite_amd_batch_prepare_gradient(tblite_amd_context ctx, tblite_amd_calculator calc,
                                      int nmol, const int* mol_nat, const int* num,
                                      const double* charge, const int* uhf, double accuracy)
{
  return prepare_batch(ctx, calc, nmol, mol_nat, num, charge, uhf, accuracy, true);
}

int tblite_amd_batch_run_gradient(tblite_amd_context ctx)
{
  if (!ctx) return TBLITE_AMD_ERR_ARGUMENT;
  if (!ctx->prepared || !ctx->batch.grad) {
    set_error(ctx, "no batch prepared with tblite_amd_batch_prepare_gradient on this context");
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  int rc = run_batch(ctx, ctx->prepared->c, nullptr, ctx->prepared_acc);
  if (rc) return rc;
  rc = run_gradient(ctx);
  if (rc) return rc;
  ctx->have_batch = true;
  return TBLITE_AMD_OK;
}

int tblite_amd_batch_gradients(tblite_amd_context ctx, double* gradients, double* sigmas)
{
  if (!ctx) return TBLITE_AMD_ERR_ARGUMENT;
  if (!ctx->have_batch) {
    set_error(ctx, "no batch has been run on this context");
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  return copy_gradients(ctx, -1, gradients, sigmas);
}

const double* tblite_amd_batch_gradient_device(tblite_amd_context ctx)
{
  if (!ctx || !ctx->prepared || !ctx->batch.grad) return nullptr;
  return ctx->batch.b.grad;
}

int tblite_amd_batch_set_geometry(tblite_amd_context ctx, const double* xyz, int on_device)
{
  if (!ctx || !xyz) return TBLITE_AMD_ERR_ARGUMENT;
  if (!ctx->prepared) {
    set_error(ctx, "no prepared batch on this context");
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  HIP_TRY(ctx, hipMemcpy((void*)ctx->batch.b.xyz, xyz, sizeof(double) * 3 * ctx->plan.tot_atoms,
                         on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
  return TBLITE_AMD_OK;
}

const double* tblite_amd_batch_geometry_device(tblite_amd_context ctx)
{
  if (!ctx || !ctx->prepared) return nullptr;
  return ctx->batch.b.xyz;
}

int tblite_amd_batch_run(tblite_amd_context ctx)
{
  if (!ctx) return TBLITE_AMD_ERR_ARGUMENT;
  if (!ctx->prepared) {
    set_error(ctx, "no prepared batch on this context");
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  int rc = run_batch(ctx, ctx->prepared->c, nullptr, ctx->prepared_acc);
  if (rc) return rc;
  ctx->have_batch = true;
  return TBLITE_AMD_OK;
}

int tblite_amd_batch_results(tblite_amd_context ctx, double* energies, int* iterations,
                             int* status)
{
  if (!ctx) return TBLITE_AMD_ERR_ARGUMENT;
  if (!ctx->have_batch) {
    set_error(ctx, "no batch has been run on this context");
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  return copy_results(ctx, ctx->plan.nmol, energies, iterations, status);
}

int tblite_amd_batch_dims(tblite_amd_context ctx, int imol, int* nat, int* nsh, int* nao)
{
  int rc = check_batch(ctx, imol);
  if (rc) return rc;
  if (nat) *nat = ctx->plan.mol_nat[imol];
  if (nsh) *nsh = ctx->plan.mol_nsh[imol];
  if (nao) *nao = ctx->plan.mol_nao[imol];
  return TBLITE_AMD_OK;
}

int tblite_amd_batch_get_cn(tblite_amd_context ctx, int imol, double* cn)
{
  int rc = check_batch(ctx, imol);
  if (rc) return rc;
  if (!cn) return TBLITE_AMD_ERR_ARGUMENT;
  HIP_TRY(ctx, hipMemcpy(cn, ctx->batch.b.cn + ctx->plan.mol_atom_off[imol],
                         8 * ctx->plan.mol_nat[imol], hipMemcpyDeviceToHost));
  return TBLITE_AMD_OK;
}

int tblite_amd_batch_get_selfenergy(tblite_amd_context ctx, int imol, double* se)
{
  int rc = check_batch(ctx, imol);
  if (rc) return rc;
  if (!se) return TBLITE_AMD_ERR_ARGUMENT;
  HIP_TRY(ctx, hipMemcpy(se, ctx->batch.b.selfenergy + ctx->plan.mol_sh_off[imol],
                         8 * ctx->plan.mol_nsh[imol], hipMemcpyDeviceToHost));
  return TBLITE_AMD_OK;
}

int tblite_amd_batch_get_adjlist(tblite_amd_context ctx, int imol, int* inl, int* nnl, int* nlat,
                                 int capacity)
{
  int rc = check_batch(ctx, imol);
  if (rc) return rc;
  const BatchDev& b = ctx->batch.b;
  const int nat = ctx->plan.mol_nat[imol], a0 = ctx->plan.mol_atom_off[imol];
  int nnz = 0;
  HIP_TRY(ctx, hipMemcpy(&nnz, b.mol_nnz + imol, sizeof(int), hipMemcpyDeviceToHost));
  if (inl) HIP_TRY(ctx, hipMemcpy(inl, b.nl_inl + a0, 4 * nat, hipMemcpyDeviceToHost));
  if (nnl) HIP_TRY(ctx, hipMemcpy(nnl, b.nl_nnl + a0, 4 * nat, hipMemcpyDeviceToHost));
  if (nlat) {
    if (capacity < nnz) {
      set_error(ctx, "neighbour list needs %d entries, capacity %d", nnz, capacity);
      return TBLITE_AMD_ERR_ARGUMENT;
    }
    if (nnz > 0)
      HIP_TRY(ctx, hipMemcpy(nlat, b.nl_jat + ctx->plan.mol_nl_off[imol], 4 * (size_t)nnz,
                             hipMemcpyDeviceToHost));
  }
  return nnz;
}

int tblite_amd_batch_get_images(tblite_amd_context ctx, int imol, int* nltr, int nltr_capacity,
                                double* trans, int trans_capacity)
{
  int rc = check_batch(ctx, imol);
  if (rc) return rc;
  const Plan& pl = ctx->plan;
  if (!pl.any_pbc || pl.mol_ntr[imol] == 0) return 0;
  const BatchDev& b = ctx->batch.b;
  const int ntr = pl.mol_ntr[imol];
  if (trans) {
    if (trans_capacity < ntr) {
      set_error(ctx, "translation capacity %d < %d", trans_capacity, ntr);
      return TBLITE_AMD_ERR_ARGUMENT;
    }
    std::memcpy(trans, pl.trans.data() + 3 * pl.mol_tr_off[imol], sizeof(double) * 3 * ntr);
  }
  if (nltr) {
    int nnz = 0;
    HIP_TRY(ctx, hipMemcpy(&nnz, b.mol_nnz + imol, sizeof(int), hipMemcpyDeviceToHost));
    if (nltr_capacity < nnz) {
      set_error(ctx, "neighbour-list capacity %d < %d", nltr_capacity, nnz);
      return TBLITE_AMD_ERR_ARGUMENT;
    }
    if (nnz > 0)
      HIP_TRY(ctx, hipMemcpy(nltr, b.nl_itr + pl.mol_nl_off[imol], 4 * (size_t)nnz,
                             hipMemcpyDeviceToHost));
  }
  return ntr;
}

int tblite_amd_batch_get_integrals(tblite_amd_context ctx, int imol, double* overlap,
                                   double* dpint, double* qpint, double* hamiltonian)
{
  int rc = check_batch(ctx, imol);
  if (rc) return rc;
  const BatchDev& b = ctx->batch.b;
  const int n = ctx->plan.mol_nao[imol];
  const int64_t off = ctx->plan.mol_mat_off[imol];
  const size_t n2 = (size_t)n * n;
  if (overlap) HIP_TRY(ctx, hipMemcpy(overlap, b.S + off, 8 * n2, hipMemcpyDeviceToHost));
  if (hamiltonian) HIP_TRY(ctx, hipMemcpy(hamiltonian, b.H + off, 8 * n2, hipMemcpyDeviceToHost));
  std::vector<double> tmp;
  if (dpint) {
    tmp.resize(3 * n2);
    HIP_TRY(ctx, hipMemcpy(tmp.data(), b.D + 3 * off, 8 * 3 * n2, hipMemcpyDeviceToHost));
    for (size_t e = 0; e < n2; ++e)
      for (int k = 0; k < 3; ++k) dpint[3 * e + k] = tmp[k * n2 + e];
  }
  if (qpint) {
    tmp.resize(6 * n2);
    HIP_TRY(ctx, hipMemcpy(tmp.data(), b.Q + 6 * off, 8 * 6 * n2, hipMemcpyDeviceToHost));
    for (size_t e = 0; e < n2; ++e)
      for (int k = 0; k < 6; ++k) qpint[6 * e + k] = tmp[k * n2 + e];
  }
  return TBLITE_AMD_OK;
}

int tblite_amd_batch_get_charges(tblite_amd_context ctx, int imol, double* qsh, double* qat,
                                 double* dpat, double* qpat, double* emo, double* focc)
{
  int rc = check_batch(ctx, imol);
  if (rc) return rc;
  tblite_amd_wavefunction_s w;
  w.ctx = ctx;
  rc = fill_wavefunction(ctx, imol, &w);
  if (rc) return rc;
  if (qsh) std::memcpy(qsh, w.qsh.data(), 8 * w.qsh.size());
  if (qat) std::memcpy(qat, w.qat.data(), 8 * w.qat.size());
  if (dpat) std::memcpy(dpat, w.dpat.data(), 8 * w.dpat.size());
  if (qpat) std::memcpy(qpat, w.qpat.data(), 8 * w.qpat.size());
  if (emo) std::memcpy(emo, w.emo.data(), 8 * w.emo.size());
  if (focc) std::memcpy(focc, w.focc.data(), 8 * w.focc.size());
  return TBLITE_AMD_OK;
}

int tblite_amd_batch_get_energy_parts(tblite_amd_context ctx, int imol, double* erep,
                                      double* eelec)
{
  int rc = check_batch(ctx, imol);
  if (rc) return rc;
  const BatchDev& b = ctx->batch.b;
  if (erep) HIP_TRY(ctx, hipMemcpy(erep, b.erep_tot + imol, 8, hipMemcpyDeviceToHost));
  if (eelec) HIP_TRY(ctx, hipMemcpy(eelec, b.eelec + imol, 8, hipMemcpyDeviceToHost));
  return TBLITE_AMD_OK;
}

int tblite_amd_batch_get_dispersion(tblite_amd_context ctx, int imol, double* edisp2,
                                    double* edisp3)
{
  int rc = check_batch(ctx, imol);
  if (rc) return rc;
  const BatchDev& b = ctx->batch.b;
  // the SCF kernels write both slots of every molecule (zero without D4)
  double e[2] = {0.0, 0.0};
  HIP_TRY(ctx, hipMemcpy(&e[0], b.edisp2 + imol, 8, hipMemcpyDeviceToHost));
  HIP_TRY(ctx, hipMemcpy(&e[1], b.edisp3 + imol, 8, hipMemcpyDeviceToHost));
  if (edisp2) *edisp2 = e[0];
  if (edisp3) *edisp3 = e[1];
  return TBLITE_AMD_OK;
}

int tblite_amd_last_profile(tblite_amd_context ctx, long long* out)
{
  if (!ctx || !out) return TBLITE_AMD_ERR_ARGUMENT;
  if (!ctx->have_batch || !ctx->batch.b.prof) {
    set_error(ctx, "no profile: set TBLITE_AMD_PROFILE=1 before creating the context");
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  HIP_TRY(ctx, hipMemcpy(out, ctx->batch.b.prof,
                         sizeof(long long) * ctx->plan.nmol * TA_PROF_SLOTS, hipMemcpyDeviceToHost));
  return TA_PROF_SLOTS;
}

int tblite_amd_last_bucket_timings(tblite_amd_context ctx, double* ms, int* ldmax, int* nmol,
                                   int* nthreads, int capacity)
{
  if (!ctx || capacity < 0) return TBLITE_AMD_ERR_ARGUMENT;
  if (!ctx->have_batch) {
    set_error(ctx, "no batch has been run on this context");
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  const int nb = (int)ctx->plan.buckets.size();
  for (int k = 0; k < nb && k < capacity; ++k) {
    const Plan::Bucket& bk = ctx->plan.buckets[k];
    if (ms) ms[k] = ctx->bucket_ms[k];
    if (ldmax) ldmax[k] = bk.ldmax;
    if (nmol) nmol[k] = (int)bk.mols.size();
    if (nthreads) nthreads[k] = bk.nt;
  }
  return nb;
}

int tblite_amd_get_tuning(tblite_amd_context ctx, tblite_amd_tuning* out)
{
  if (!ctx || !out) return TBLITE_AMD_ERR_ARGUMENT;
  out->lds_max_ld = ctx->lds_max_ld;
  out->grid_min_ld = ctx->grid_min_ld;
  out->auto_grid = ctx->auto_grid ? 1 : 0;
  out->grid_graph = ctx->grid_graph ? 1 : 0;
  out->wg_a = ctx->wg_a;
  out->wg_p = ctx->wg_p;
  out->grid_b = ctx->grid_b;
  out->grid_c = ctx->grid_c;
  out->integral_chunk = ctx->int_chunk;
  out->pair_chunk_log2 = ctx->pair_chunk_log2;
  out->atm_chunk_log2 = ctx->atm_chunk_log2;
  out->fine_buckets = ctx->fine_buckets ? 1 : 0;
  out->big_lds_min_kb = (int)(ctx->big_lds_min / 1024);
  return TBLITE_AMD_OK;
}

int tblite_amd_set_tuning(tblite_amd_context ctx, const tblite_amd_tuning* in)
{
  if (!ctx || !in) return TBLITE_AMD_ERR_ARGUMENT;
  if (in->lds_max_ld < 16 || in->lds_max_ld > kLdsMaxLd || in->lds_max_ld % 16 != 0 ||
      in->grid_min_ld < 16 || in->grid_min_ld > TA_BIG_MAXLD || !(in->wg_a > 0.0) ||
      !(in->wg_p > 0.0) || !(in->grid_b >= 0.0) || !std::isfinite(in->grid_c)) {
    set_error(ctx, "set_tuning: lds_max_ld must be 16..%d in steps of 16, grid_min_ld 16..%d, "
              "wg_a, wg_p > 0, grid_b >= 0", kLdsMaxLd, TA_BIG_MAXLD);
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  if ((in->integral_chunk != 32 && in->integral_chunk != 64 && in->integral_chunk != 128) ||
      in->pair_chunk_log2 < 8 || in->pair_chunk_log2 > 22 || in->atm_chunk_log2 < 10 ||
      in->atm_chunk_log2 > 26 || in->big_lds_min_kb < 0 || in->big_lds_min_kb > 160) {
    set_error(ctx, "set_tuning: integral_chunk must be 32, 64 or 128, pair_chunk_log2 8..22, "
              "atm_chunk_log2 10..26, big_lds_min_kb 0..160");
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  ctx->lds_max_ld = in->lds_max_ld;
  ctx->grid_min_ld = in->grid_min_ld;
  ctx->auto_grid = in->auto_grid != 0;
  ctx->grid_graph = in->grid_graph != 0;
  ctx->wg_a = in->wg_a;
  ctx->wg_p = in->wg_p;
  ctx->grid_b = in->grid_b;
  ctx->grid_c = in->grid_c;
  ctx->int_chunk = in->integral_chunk;
  ctx->pair_chunk_log2 = in->pair_chunk_log2;
  ctx->atm_chunk_log2 = in->atm_chunk_log2;
  ctx->fine_buckets = in->fine_buckets != 0;
  ctx->big_lds_min = (size_t)in->big_lds_min_kb * 1024;
  ctx->prepared = nullptr;  // a prepared batch was planned with the old routing
  return TBLITE_AMD_OK;
}

int tblite_amd_molecule_path(tblite_amd_context ctx, int imol)
{
  if (!ctx || imol < 0 || imol >= (int)ctx->plan.mol_path.size()) return TBLITE_AMD_ERR_ARGUMENT;
  return ctx->plan.mol_path[imol];
}

int tblite_amd_last_timings(tblite_amd_context ctx, double* times, int n)
{
  if (!ctx || !times) return TBLITE_AMD_ERR_ARGUMENT;
  for (int i = 0; i < n && i < 5; ++i) times[i] = ctx->timings[i];
  return TBLITE_AMD_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ serving
// Continuous batching over independent singlepoint requests.  The server
// owns `slots` engine contexts on the same device, each with its own streams
// and a device workspace reserved once (budget / slots bytes).  A slot takes
// requests from the head of the FIFO queue while their estimated footprint
// fits its workspace, runs them as one batch, and is refilled as soon as its
// batch has finished; batches of different slots run concurrently.  Requests
// that arrive while others are on the device wait only for the next free
// slot.  Wavefunction and Broyden state stay on the device for the whole SCF
// of a batch.  No counterpart in the reference (one molecule per GPU thread,
// src/xtb.cu:479-497).
namespace {

struct ServeRequest {
  long long ticket;
  int nat;
  std::vector<int> num;
  std::vector<double> xyz;
  double charge;
  int uhf;
  double t_submit;
  size_t bytes;
  size_t plane_bytes;  // AES2 pair planes, allocated for the whole batch once it has a
                       // multi-workgroup molecule
  bool big;   // above the LDS-resident size
  bool grid;  // multi-workgroup path
  // light-load routing: the multi-workgroup path is at least kGridServeGain
  // times faster per SCF iteration than one workgroup (cost model), and the
  // request's footprint alone on that path
  bool grid_ok = false;
  size_t grid_bytes = 0;
};

struct ServeResult {
  long long ticket;
  double energy;
  int iterations;
  int status;
  double latency_ms;
};

struct ServeSlot {
  tblite_amd_context ctx = nullptr;
  // a batch of one request on the multi-workgroup path runs its host loop on
  // a worker thread (state 1 running, 2 finished), so poll() never blocks
  std::thread worker;
  std::atomic<int> async_state{0};
  int async_rc = 0;
  bool async = false;
  std::vector<ServeRequest> reqs;
  std::vector<double> xyz;  // host coordinates of the running batch (kept until it ends)
  std::vector<char> reported;       // per request of the running batch
  TaDoneRecord* rec = nullptr;      // completion records, pinned host memory (max_batch)
  TaDoneRecord* rec_dev = nullptr;  // the same memory as the device sees it
  bool busy = false;
  bool small_class = false;  // takes only molecules of the LDS-resident sizes
};

double host_ms()
{
  using clk = std::chrono::steady_clock;
  return std::chrono::duration<double, std::milli>(clk::now().time_since_epoch()).count();
}

// Upper estimate of the device workspace a molecule adds to a batch (all
// per-molecule arrays of upload_plan, alignment slack included).  grid_min_ld:
// padded AO dimension above which the molecule takes the multi-workgroup path
// (its scratch differs); *plane_bytes receives the molecule's share of the
// AES2 pair planes that a batch holding a multi-workgroup molecule allocates
// for all its molecules.
size_t request_bytes(const Calculator& c, int nat, const int* num, int grid_min_ld, int* nao_out,
                     size_t* plane_bytes, bool* grid)
{
  int nsh = 0, nao = 0;
  for (int i = 0; i < nat; ++i) {
    const int z = num[i];
    const int e = (z > 0 && z < 120) ? c.zmap[z] : -1;
    if (e < 0) return 0;
    const DevElem& el = c.dev.el[e];
    nsh += el.nsh;
    for (int s = 0; s < el.nsh; ++s) nao += 2 * el.ang[s] + 1;
  }
  const int64_t mem = std::max(1, c.max_iter);
  const int64_t ndim = nsh + 9 * (int64_t)nat;
  const int64_t ld = (nao + 3) & ~3;
  const bool on_grid = ld > grid_min_ld;
  int64_t d = 11 * (int64_t)nao * nao + (int64_t)nsh * nsh + 20 * (int64_t)nat * nat;
  d += (int64_t)nat * (nat + 1) / 2 + (24 + 2 + 2 * TA_D4_MAXREF) * (int64_t)nat + 6 * nsh + 4 * nao + 2;
  d += 4 * ndim + 2 * ndim * mem + mem * mem + mem * (mem + 1) + 2 * mem + 8 + TA_PROF_SLOTS;
  if (on_grid)
    d += 4 * ld * ld + grid_scratch_doubles(nao, (int)ld, nat, nsh);
  else if (ld > kLdsMaxLd)
    d += 4 * ld * ld + big_scratch_doubles((int)ld);
  if (nao_out) *nao_out = nao;
  if (plane_bytes) *plane_bytes = (size_t)18 * 8 * nat * nat + 3 * 256;
  if (grid) *grid = on_grid;
  return (size_t)d * 8 + 64 * 256;
}

// Requests for which the cost model puts the multi-workgroup path at least
// this many times ahead of one workgroup wait in their own queue, served one
// at a time by the server's multi-workgroup lane; beyond kGridServeQueue of
// them waiting, the regular slots take the oldest (one workgroup each), so
// under heavy load they batch like other large requests.
constexpr double kGridServeGain = 2.0;
constexpr size_t kGridServeQueue = 2;

// Workspace of a batch of requests: their own arrays, the AES2 pair planes of
// all of them once one takes the multi-workgroup path, and a fixed reserve.
size_t batch_bytes(const std::vector<ServeRequest>& reqs)
{
  size_t used = 64 * 1024 * 8, planes = 0;
  bool grid = false;
  for (const ServeRequest& r : reqs) {
    used += r.bytes;
    planes += r.plane_bytes;
    grid = grid || r.grid;
  }
  return used + (grid ? planes : 0);
}

}  // namespace

struct tblite_amd_server_s {
  tblite_amd_context owner = nullptr;
  tblite_amd_calculator calc = nullptr;
  std::vector<std::unique_ptr<ServeSlot>> slots;
  // multi-workgroup lane (two or more slots): one request at a time on the
  // multi-workgroup path, its SCF host loop on a worker thread
  std::unique_ptr<ServeSlot> lane;
  // waiting requests by size class: LDS-resident sizes, larger ones, and
  // those the lane serves faster
  std::deque<ServeRequest> queue_small, queue_big, queue_grid;
  std::deque<ServeResult> done;
  size_t slot_bytes = 0, lane_bytes = 0;
  int max_batch = 4096;
  long long next_ticket = 1, submitted = 0, completed = 0, batches = 0, grid_requests = 0;
};

namespace {
// context, workspace reservation and completion records of one slot
bool make_slot(tblite_amd_server srv, ServeSlot& sl, size_t bytes)
{
  tblite_amd_context ctx = srv->owner;
  sl.ctx = make_context(ctx->device, true);
  bool ok = sl.ctx != nullptr;
  // a multi-workgroup molecule runs synchronously on the host thread, which
  // would stall poll(): slots keep one workgroup per molecule (the lane
  // routes its requests itself)
  if (ok) sl.ctx->auto_grid = false;
  if (ok) {
    // reserve the workspace once: growing it later would free device
    // memory, which synchronises the whole device
    Arena& ar = sl.ctx->arena;
    ok = hipMalloc(&ar.base, bytes) == hipSuccess;
    if (ok) ar.cap = bytes;
  }
  if (ok)
    ok = hipHostMalloc((void**)&sl.rec, sizeof(TaDoneRecord) * srv->max_batch,
                       hipHostMallocCoherent) == hipSuccess &&
         hipHostGetDevicePointer((void**)&sl.rec_dev, sl.rec, 0) == hipSuccess;
  if (ok) std::memset(sl.rec, 0, sizeof(TaDoneRecord) * srv->max_batch);
  return ok;
}

void free_slot(ServeSlot& sl)
{
  if (sl.worker.joinable()) sl.worker.join();
  if (sl.ctx) tblite_amd_delete_context(&sl.ctx);
  if (sl.rec) (void)hipHostFree(sl.rec);
  sl.rec = nullptr;
}
}  // namespace

extern "C" {

tblite_amd_server tblite_amd_new_server(tblite_amd_context ctx, tblite_amd_calculator calc,
                                        int slots, double workspace_bytes)
{
  if (!ctx || !calc || slots < 1 || slots > 16 || !(workspace_bytes > 0.0)) {
    if (ctx) set_error(ctx, "new_server: need 1..16 slots and a positive workspace");
    return nullptr;
  }
  if (hipSetDevice(ctx->device) != hipSuccess) return nullptr;
  if (ensure_params_on_device(ctx, calc->c)) return nullptr;
  tblite_amd_server srv = new (std::nothrow) tblite_amd_server_s();
  if (!srv) return nullptr;
  srv->owner = ctx;
  srv->calc = calc;
  // The multi-workgroup lane (two or more slots) issues thousands of short
  // kernels per SCF iteration; a stream that shares a hardware queue with it
  // waits behind them, and the process has four hardware queues.  With four
  // or more slots the last slot is the lane (one stream per slot, a slot's
  // share of the workspace); with two or three it is an extra context with
  // half a slot's share.
  const int nreg = slots >= 4 ? slots - 1 : slots;
  const double shares = slots >= 4 ? (double)slots : slots >= 2 ? slots + 0.5 : (double)slots;
  srv->slot_bytes = (size_t)(workspace_bytes / shares);
  srv->lane_bytes = slots >= 4 ? srv->slot_bytes : slots >= 2 ? srv->slot_bytes / 2 : 0;
  for (int k = 0; k < nreg; ++k) srv->slots.emplace_back(new ServeSlot());
  for (size_t k = 0; k < srv->slots.size(); ++k) {
    ServeSlot& sl = *srv->slots[k];
    // with two or more slots the first one serves small molecules only, so
    // they never wait behind a batch of large ones
    sl.small_class = nreg >= 2 && k == 0;
    if (!make_slot(srv, sl, srv->slot_bytes)) {
      set_error(ctx, "new_server: cannot create slot context / reserve %zu bytes", srv->slot_bytes);
      tblite_amd_delete_server(&srv);
      return nullptr;
    }
  }
  if (slots >= 2) {
    srv->lane.reset(new ServeSlot());
    if (!make_slot(srv, *srv->lane, srv->lane_bytes)) {
      set_error(ctx, "new_server: cannot create the multi-workgroup lane / reserve %zu bytes",
                srv->lane_bytes);
      tblite_amd_delete_server(&srv);
      return nullptr;
    }
  }
  return srv;
}

void tblite_amd_delete_server(tblite_amd_server* psrv)
{
  if (!psrv || !*psrv) return;
  for (auto& up : (*psrv)->slots) free_slot(*up);
  if ((*psrv)->lane) free_slot(*(*psrv)->lane);
  delete *psrv;
  *psrv = nullptr;
}

int tblite_amd_server_submit(tblite_amd_server srv, int nat, const int* num, const double* xyz,
                             double charge, int uhf, long long* ticket)
{
  if (!srv) return TBLITE_AMD_ERR_ARGUMENT;
  tblite_amd_context ctx = srv->owner;
  if (nat < 1 || !num || !xyz) {
    set_error(ctx, "server_submit: invalid molecule");
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  int nao = 0;
  size_t plane_bytes = 0;
  bool grid = false;
  const size_t bytes = request_bytes(srv->calc->c, nat, num, srv->slots[0]->ctx->grid_min_ld, &nao,
                                     &plane_bytes, &grid);
  if (bytes == 0) {
    set_error(ctx, "server_submit: element without parameters");
    return TBLITE_AMD_ERR_ARGUMENT;
  }
  if (bytes + 64 * 1024 * 8 + (grid ? plane_bytes : 0) > srv->slot_bytes) {
    set_error(ctx, "server_submit: molecule needs %zu bytes, a slot holds %zu", bytes,
              srv->slot_bytes);
    return TBLITE_AMD_ERR_TOO_LARGE;
  }
  ServeRequest r;
  r.ticket = srv->next_ticket++;
  r.nat = nat;
  r.num.assign(num, num + nat);
  r.xyz.assign(xyz, xyz + 3 * (size_t)nat);
  r.charge = charge;
  r.uhf = uhf;
  r.t_submit = host_ms();
  r.bytes = bytes;
  r.plane_bytes = plane_bytes;
  r.grid = grid;
  r.big = ((nao + 3) & ~3) > kLdsMaxLd;
  {
    // light-load routing to the multi-workgroup path: the cost model of the
    // owner context (one SCF iteration: wg_a ld^wg_p on one workgroup,
    // grid_b ld + grid_c across the device)
    const tblite_amd_context o = srv->owner;
    const double ld = (double)((nao + 3) & ~3);
    const double t_wg = o->wg_a * std::pow(ld, o->wg_p);
    const double t_grid = std::max(1e-9, o->grid_b * ld + o->grid_c);
    if (srv->lane && r.big && !grid && t_wg >= kGridServeGain * t_grid) {
      size_t pb = 0;
      bool g2 = false;
      const size_t gb = request_bytes(srv->calc->c, nat, num, 16, nullptr, &pb, &g2);
      r.grid_bytes = gb + pb + 64 * 1024 * 8;
      r.grid_ok = r.grid_bytes <= srv->lane_bytes;
    }
  }
  if (ticket) *ticket = r.ticket;
  (r.grid_ok ? srv->queue_grid : r.big ? srv->queue_big : srv->queue_small).push_back(std::move(r));
  ++srv->submitted;
  return TBLITE_AMD_OK;
}

namespace {

// Report the molecules of a running batch whose completion records have
// appeared since the last call.
void serve_scan(tblite_amd_server srv, ServeSlot& sl)
{
  const int nmol = (int)sl.reqs.size();
  double now = -1.0;
  for (int m = 0; m < nmol; ++m) {
    if (sl.reported[m]) continue;
    const TaDoneRecord* r = sl.rec + m;
    if (__atomic_load_n(&r->flag, __ATOMIC_ACQUIRE) == 0) continue;
    if (now < 0.0) now = host_ms();
    ServeResult res;
    res.ticket = sl.reqs[m].ticket;
    res.energy = r->energy;
    res.iterations = r->iterations;
    res.status = r->status;
    res.latency_ms = now - sl.reqs[m].t_submit;
    srv->done.push_back(res);
    sl.reported[m] = 1;
    ++srv->completed;
  }
}

// Collect a finished slot: the molecules without a completion record (the
// multi-workgroup path) come from the batch arrays; returns OK or an error.
int serve_collect(tblite_amd_server srv, ServeSlot& sl)
{
  tblite_amd_context sc = sl.ctx;
  int rc = finish_batch(sc);
  if (rc) return rc;
  serve_scan(srv, sl);
  const int nmol = (int)sl.reqs.size();
  int left = 0;
  for (int m = 0; m < nmol; ++m) left += sl.reported[m] ? 0 : 1;
  std::vector<double> e(nmol);
  std::vector<int> it(nmol), st(nmol);
  const BatchDev& b = sc->batch.b;
  if (left > 0) {
    // on the slot's own stream: a copy on another stream could wait behind
    // another slot's kernels
    HIP_TRY(srv->owner, hipMemcpyAsync(e.data(), b.energy, 8 * nmol, hipMemcpyDeviceToHost, sc->stream));
    HIP_TRY(srv->owner, hipMemcpyAsync(it.data(), b.iterations, 4 * nmol, hipMemcpyDeviceToHost, sc->stream));
    HIP_TRY(srv->owner, hipMemcpyAsync(st.data(), b.status, 4 * nmol, hipMemcpyDeviceToHost, sc->stream));
    HIP_TRY(srv->owner, hipStreamSynchronize(sc->stream));
  }
  const double now = host_ms();
  for (int m = 0; m < nmol; ++m) {
    if (sl.reported[m]) continue;
    ServeResult res;
    res.ticket = sl.reqs[m].ticket;
    res.energy = e[m];
    res.iterations = it[m];
    res.status = st[m];
    res.latency_ms = now - sl.reqs[m].t_submit;
    srv->done.push_back(res);
    ++srv->completed;
  }
  sc->have_batch = true;
  sl.reqs.clear();
  sl.busy = false;
  return TBLITE_AMD_OK;
}

// Report every request of a slot's batch as failed and empty the slot.
void serve_fail(tblite_amd_server srv, ServeSlot& sl)
{
  const double now = host_ms();
  for (const ServeRequest& r : sl.reqs) {
    ServeResult res;
    res.ticket = r.ticket;
    res.energy = 0.0;
    res.iterations = 0;
    res.status = TA_STATUS_FAILED;
    res.latency_ms = now - r.t_submit;
    srv->done.push_back(res);
    ++srv->completed;
  }
  sl.reqs.clear();
}

// Launch the slot's single request on the multi-workgroup path: plan and
// upload on this thread, then the SCF host loop on a worker thread.
int serve_launch_grid(tblite_amd_server srv, ServeSlot& sl)
{
  ServeRequest& r = sl.reqs[0];
  tblite_amd_context sc = sl.ctx;
  Calculator& c = srv->calc->c;
  sl.xyz = r.xyz;
  const int mol_nat = r.nat;
  sc->have_batch = false;
  sc->prepared = nullptr;
  const int gmin = sc->grid_min_ld;
  sc->grid_min_ld = 16;  // this batch: the multi-workgroup path
  int rc = make_plan(sc, c, 1, &mol_nat, r.num.data(), &r.charge, &r.uhf, c.accuracy, sc->plan);
  sc->grid_min_ld = gmin;
  if (!rc) rc = upload_plan(sc, c, sc->plan, sc->batch);
  if (rc) {
    set_error(srv->owner, "server: batch launch failed: %s", sc->error.c_str());
    (void)hipStreamSynchronize(sc->stream);
    serve_fail(srv, sl);
    return rc;
  }
  // no completion record on this path: the result comes from the batch
  // arrays when the worker has finished.  The slot's record memory still
  // holds the last batch's records, which serve_scan would report.
  sc->batch.b.done = nullptr;
  std::memset(sl.rec, 0, sizeof(TaDoneRecord) * srv->max_batch);
  sl.reported.assign(1, 0);
  sl.async = true;
  sl.async_rc = 0;
  sl.async_state.store(1, std::memory_order_release);
  sl.busy = true;
  ++srv->batches;
  ++srv->grid_requests;
  const int dev = sc->device;
  ServeSlot* ps = &sl;
  sl.worker = std::thread([ps, sc, &c, dev]() {
    int e = hipSetDevice(dev) == hipSuccess ? TBLITE_AMD_OK : TBLITE_AMD_ERR_HIP;
    if (!e) e = run_batch(sc, c, ps->xyz.data(), c.accuracy, false);
    ps->async_rc = e;
    ps->async_state.store(2, std::memory_order_release);
  });
  return TBLITE_AMD_OK;
}

// Fill a free slot from the queue head and launch its batch.
// The lane takes the largest waiting request of its queue.
int serve_launch_lane(tblite_amd_server srv, ServeSlot& sl)
{
  std::deque<ServeRequest>& q = srv->queue_grid;
  size_t best = 0;
  for (size_t k = 1; k < q.size(); ++k)
    if (q[k].bytes > q[best].bytes) best = k;
  sl.reqs.push_back(std::move(q[best]));
  q.erase(q.begin() + (std::ptrdiff_t)best);
  return serve_launch_grid(srv, sl);
}

int serve_launch(tblite_amd_server srv, ServeSlot& sl)
{
  // a small-class slot takes small molecules; the others take the lane's
  // overflow (beyond kGridServeQueue waiting, oldest first), then large
  // ones, then fill up with small ones (those run first and report early).
  // The footprint counts the requests already in the slot (none unless a
  // caller changes that) and the pair planes a multi-workgroup molecule
  // brings.
  for (int pass = sl.small_class ? 2 : 0; pass < 3; ++pass) {
    std::deque<ServeRequest>& q = pass == 0 ? srv->queue_grid : pass == 1 ? srv->queue_big : srv->queue_small;
    const size_t keep = pass == 0 ? kGridServeQueue : 0;
    while (q.size() > keep && (int)sl.reqs.size() < srv->max_batch) {
      sl.reqs.push_back(std::move(q.front()));
      q.pop_front();
      if (batch_bytes(sl.reqs) > srv->slot_bytes && sl.reqs.size() > 1) {
        q.push_front(std::move(sl.reqs.back()));
        sl.reqs.pop_back();
        break;
      }
    }
  }
  if (sl.reqs.empty()) return TBLITE_AMD_OK;
  const int nmol = (int)sl.reqs.size();
  std::vector<int> mol_nat(nmol), num, uhf(nmol);
  std::vector<double>& xyz = sl.xyz;
  std::vector<double> charge(nmol);
  xyz.clear();
  for (int m = 0; m < nmol; ++m) {
    const ServeRequest& r = sl.reqs[m];
    mol_nat[m] = r.nat;
    num.insert(num.end(), r.num.begin(), r.num.end());
    xyz.insert(xyz.end(), r.xyz.begin(), r.xyz.end());
    charge[m] = r.charge;
    uhf[m] = r.uhf;
  }
  tblite_amd_context sc = sl.ctx;
  Calculator& c = srv->calc->c;
  sc->have_batch = false;
  sc->prepared = nullptr;
  int rc = make_plan(sc, c, nmol, mol_nat.data(), num.data(), charge.data(), uhf.data(),
                     c.accuracy, sc->plan);
  if (!rc) rc = upload_plan(sc, c, sc->plan, sc->batch);
  if (!rc) {
    std::memset(sl.rec, 0, sizeof(TaDoneRecord) * nmol);
    sl.reported.assign(nmol, 0);
    sc->batch.b.done = sl.rec_dev;
    rc = run_batch(sc, c, xyz.data(), c.accuracy, false);
  }
  if (rc) {
    // the batch's requests end here, each reported with TA_STATUS_FAILED, so
    // no ticket is lost and the next poll starts from an empty slot
    set_error(srv->owner, "server: batch launch failed: %s", sc->error.c_str());
    (void)hipStreamSynchronize(sc->stream);
    serve_fail(srv, sl);
    return rc;
  }
  sl.busy = true;
  ++srv->batches;
  return TBLITE_AMD_OK;
}

}  // namespace

int tblite_amd_server_poll(tblite_amd_server srv, int capacity, long long* tickets,
                           double* energies, int* iterations, int* status, double* latency_ms)
{
  if (!srv || capacity < 0) return TBLITE_AMD_ERR_ARGUMENT;
  HIP_TRY(srv->owner, hipSetDevice(srv->owner->device));
  const size_t nslot = srv->slots.size();
  for (size_t k = 0; k <= nslot; ++k) {
    if (k == nslot && !srv->lane) break;
    ServeSlot& sl = k < nslot ? *srv->slots[k] : *srv->lane;
    const bool is_lane = k == nslot;
    if (sl.busy && sl.async) {
      // the worker owns the slot's context until it has finished
      if (sl.async_state.load(std::memory_order_acquire) != 2) continue;
      sl.worker.join();
      sl.async = false;
      sl.async_state.store(0, std::memory_order_relaxed);
      if (sl.async_rc) {
        set_error(srv->owner, "server: multi-workgroup request failed: %s", sl.ctx->error.c_str());
        (void)hipStreamSynchronize(sl.ctx->stream);
        serve_fail(srv, sl);
        sl.busy = false;
        return sl.async_rc;
      }
    }
    if (sl.busy) {
      const hipError_t q = hipEventQuery(sl.ctx->ev[4]);
      if (q == hipErrorNotReady) {
        serve_scan(srv, sl);
        continue;
      }
      if (q != hipSuccess) {
        set_error(srv->owner, "server: batch failed: %s", hipGetErrorString(q));
        return TBLITE_AMD_ERR_HIP;
      }
      int rc = serve_collect(srv, sl);
      if (rc) return rc;
    }
    if (is_lane) {
      if (!sl.busy && !srv->queue_grid.empty()) {
        int rc = serve_launch_lane(srv, sl);
        if (rc) return rc;
      }
    } else if (!sl.busy && !(srv->queue_small.empty() && srv->queue_big.empty() &&
                             srv->queue_grid.size() <= kGridServeQueue)) {
      int rc = serve_launch(srv, sl);
      if (rc) return rc;
    }
  }
  int k = 0;
  while (k < capacity && !srv->done.empty()) {
    const ServeResult& r = srv->done.front();
    if (tickets) tickets[k] = r.ticket;
    if (energies) energies[k] = r.energy;
    if (iterations) iterations[k] = r.iterations;
    if (status) status[k] = r.status;
    if (latency_ms) latency_ms[k] = r.latency_ms;
    srv->done.pop_front();
    ++k;
  }
  return k;
}

int tblite_amd_workspace_estimate(const tblite_amd_param* param, int nmol, const int* mol_nat,
                                  const int* num, int grid_min_ld, double* need_bytes,
                                  double* admit_bytes)
{
  if (nmol < 1 || !mol_nat || !num) return TBLITE_AMD_ERR_ARGUMENT;
  Calculator c;
  if (param) c.param = *param;
  else tblite_amd_fill_gfn2_default(&c.param);
  std::string err;
  if (!build_calculator(c, err)) return TBLITE_AMD_ERR_ARGUMENT;
  // a host-only context: the planner reads only its routing settings
  tblite_amd_context_s cx;
  cx.grid_min_ld = std::max(16, std::min(TA_BIG_MAXLD, grid_min_ld));
  cx.auto_grid = false;
  Plan pl;
  int rc = make_plan(&cx, c, nmol, mol_nat, num, nullptr, nullptr, c.accuracy, pl);
  if (rc) return rc;
  std::vector<char> img;
  std::vector<std::pair<size_t, size_t>> seg;
  pack_plan(pl, img, seg);
  std::vector<ServeRequest> reqs(nmol);
  int64_t a0 = 0;
  for (int m = 0; m < nmol; ++m) {
    ServeRequest& r = reqs[m];
    r.bytes = request_bytes(c, mol_nat[m], num + a0, cx.grid_min_ld, nullptr, &r.plane_bytes, &r.grid);
    a0 += mol_nat[m];
  }
  if (need_bytes) *need_bytes = (double)workspace_need(pl, img.size(), false);
  if (admit_bytes) *admit_bytes = (double)batch_bytes(reqs);
  return TBLITE_AMD_OK;
}

int tblite_amd_server_stats(tblite_amd_server srv, long long* submitted, long long* completed,
                            int* queued, int* busy_slots, long long* batches,
                            long long* grid_requests)
{
  if (!srv) return TBLITE_AMD_ERR_ARGUMENT;
  int busy = 0;
  for (const auto& up : srv->slots) busy += up->busy ? 1 : 0;
  if (srv->lane) busy += srv->lane->busy ? 1 : 0;
  if (submitted) *submitted = srv->submitted;
  if (completed) *completed = srv->completed;
  if (queued) *queued = (int)(srv->queue_small.size() + srv->queue_big.size() + srv->queue_grid.size());
  if (busy_slots) *busy_slots = busy;
  if (batches) *batches = srv->batches;
  if (grid_requests) *grid_requests = srv->grid_requests;
  return TBLITE_AMD_OK;
}

}  // extern "C"

// Device-side data layout shared by the engine's kernels and its host code.
//
// A batch is a set of independent molecules.  Everything per atom, shell and
// atomic orbital lives in flat arrays indexed by a global position; each
// molecule knows its offsets.  Matrices are stored per molecule, n x n row
// blocks, component-planar for the multipole integrals:
//   S[moff + i*n + j]            overlap(j, i)        (tblite layout, j fastest)
//   H[moff + i*n + j]            hamiltonian(j, i)
//   D[3*moff + k*n*n + i*n + j]  dpint(k, j, i)  operator centred on atom(i)
//   Q[6*moff + k*n*n + i*n + j]  qpint(k, j, i)
// with quadrupole components ordered xx, xy, yy, xz, yz, zz (tblite order).
#pragma once

#include <stdint.h>

#define TA_MAX_ELEM 16
#define TA_MAX_SHELL 3
#define TA_MAXG 6

struct DevElem {
  int z, nsh;
  int ang[TA_MAX_SHELL];
  int nprim[TA_MAX_SHELL];
  double alpha[TA_MAX_SHELL][TA_MAXG];
  double coeff[TA_MAX_SHELL][TA_MAXG];
  double selfenergy[TA_MAX_SHELL];  // Eh
  double kcn[TA_MAX_SHELL];         // Eh
  double hardness[TA_MAX_SHELL];    // Eh, gam * lgam
  double hubbard_derivs[TA_MAX_SHELL];
  double refocc[TA_MAX_SHELL];
  double shpoly[TA_MAX_SHELL];
  double rcov, atomic_rad, zeff, arep, dkernel, qkernel, mprad, mpvcn, en;
};

#define TA_D4_MAXREF 7

// D4 dispersion data in device form (zeff, gi = gc * eta, reference C6 of
// every element/reference pair from the Casimir-Polder integral)
struct DevD4 {
  int enabled;
  // every element's references share one reference charge: the charge scale
  // factors out, g_r(q) = w_r(CN) zeta(q), and the SCF works with
  // s = zeta(q) / zeta(0) per atom and d_ij C6_ij(q = 0) per pair
  int qref_uniform;
  double s6, s8, a1, a2, s9, alp, ga, wf;
  double k4, k5, k6, kn, cn_cutoff, disp2_cutoff, disp3_cutoff;
  int nref[TA_MAX_ELEM];
  double cnref[TA_MAX_ELEM][TA_D4_MAXREF];
  double qref[TA_MAX_ELEM][TA_D4_MAXREF];
  double zeff[TA_MAX_ELEM], gi[TA_MAX_ELEM], r4r2[TA_MAX_ELEM];
  double c6[TA_MAX_ELEM][TA_MAX_ELEM][TA_D4_MAXREF][TA_D4_MAXREF];
};

struct DevParams {
  int nelem;
  DevElem el[TA_MAX_ELEM];
  // hscale[ei][ej][ish][jsh]
  double hscale[TA_MAX_ELEM][TA_MAX_ELEM][TA_MAX_SHELL][TA_MAX_SHELL];
  double rep_kexp, rep_kexp_light, rep_rexp, rep_cutoff;
  double gexp;
  double dmp3, dmp5, mp_shift, mp_kexp, mp_rmax;
  double cn_ka, cn_kb, cn_rshift, cn_cutoff;
  DevD4 d4;
};

// Pointers into one batch's device buffers (passed by value to kernels).
// Completion record of one molecule, written by the SCF kernel into pinned
// host memory when the molecule ends (flag last, with a system-scope release)
// so a host loop can report results before the whole batch has finished.
struct TaDoneRecord {
  double energy;
  int iterations;
  int status;
  int flag;
  int pad;
};

struct BatchDev {
  int nmol;
  int max_iter;
  int mix_memory;   // Broyden memory (= max_iter)
  double damping;
  double kt;
  double accuracy;
  double intcut;    // primitive-pair screening exponent
  const DevParams* par;

  // per molecule
  const int* mol_atom_off;
  const int* mol_nat;
  const int* mol_sh_off;
  const int* mol_nsh;
  const int* mol_ao_off;
  const int* mol_nao;
  const int64_t* mol_mat_off;     // sum of previous n^2
  const int64_t* mol_nl_off;      // sum of previous nat(nat+1)/2
  const int64_t* mol_gam_off;     // sum of previous nsh^2
  const int64_t* mol_pair_off;    // sum of previous nat^2
  const int64_t* mol_mix_off;     // sum of previous ndim
  const int64_t* mol_mixmem_off;  // sum of previous ndim * mix_memory
  const int64_t* mol_mixa_off;    // sum of previous mix_memory^2 (per molecule constant)
  const double* mol_charge;
  const int* mol_uhf;
  const double* mol_cutoff;       // neighbour-list cutoff (depends on species present)

  // per atom
  const int* at_el;               // element slot
  const int* at_ish;              // local offset of the atom's first shell
  const int* at_iao;              // local offset of the atom's first AO
  const double* xyz;              // 3 per atom, bohr
  double* cn;
  double* mrad;
  double* erep;                   // atom-resolved repulsion energy

  // per shell
  const int* sh_at;               // local atom index
  const int* sh_ish;              // shell index within the element
  const int* sh_iao;              // local AO offset
  double* selfenergy;
  double* n0sh;

  // per AO
  const int* ao_at;               // local atom index
  const int* ao_sh;               // local shell index

  // neighbour list (half list, jat <= iat, 0-based local atom index)
  int* nl_inl;                    // per atom, local offset within the molecule's list
  int* nl_nnl;                    // per atom
  int* nl_jat;                    // per list entry
  int* mol_nnz;                   // per molecule

  // integrals
  double* S;
  double* H;
  double* D;
  double* Q;

  // Coulomb matrices
  double* gam;                    // nsh x nsh (ES2)
  double* aes_sd;                 // [3][nat][nat]: sd_k(i,j) = amat_sd(k, j, i)
  double* aes_dd;                 // [3][3][nat][nat]: dd_kl(i,j) = amat_dd(k, j, l, i)
  double* aes_sq;                 // [6][nat][nat]

  // SCF results / wavefunction
  double* qsh;
  double* qat;
  double* dpat;                   // 3 per atom
  double* qpat;                   // 6 per atom
  double* emo;                    // per AO (sorted)
  double* focc;                   // per AO
  // optional (both null unless a wavefunction is requested): per molecule at
  // mol_mat_off, in tblite's column-major layout, the MO coefficients
  // coeff(nao, nao) (column r = level r, ascending) and the density matrix
  // density(nao, nao) of the last SCF iteration
  double* coeff;
  double* density;

  // Broyden mixer state
  double* mix_qin;
  double* mix_dq;
  double* mix_dqlast;
  double* mix_qlast;
  double* mix_df;
  double* mix_u;
  double* mix_a;                  // mix_memory^2 per molecule: <df_i, df_j>
  double* mix_beta;               // mix_memory * (mix_memory + 1) per molecule: solve scratch
  double* mix_omega;              // mix_memory per molecule
  double* mix_c;                  // mix_memory per molecule

  // per-molecule scalars
  double* energy;
  double* eelec;
  double* erep_tot;
  int* iterations;
  int* status;

  // global-memory matrices for molecules too large for LDS (ld x ld each):
  // big_a Fock / reduced matrix / reflectors, big_c X = L^-1, big_t
  // eigenvectors / density, big_ct MO coefficients (one MO per row)
  double* big_a;
  double* big_c;
  double* big_t;
  double* big_ct;
  const int64_t* mol_big_off;     // offset into the big_* matrices, -1 when LDS-resident
  // eigensolver scratch of those molecules (big_scratch_doubles(ld) each)
  double* big_s;
  const int64_t* mol_scr_off;

  // D4 dispersion: per atom the D4 CN, the reference weights g_r(q) and
  // dg_r/dq of the current charges (TA_D4_MAXREF each, SCF scratch; with
  // DevD4::qref_uniform only the first entry, the charge scale s and ds/dq), the
  // three-body energy of a chunk of triples (e3[a]: triples whose largest atom
  // index lies in the chunk starting at atom a, zero for other atoms); per
  // atom pair (nat x nat per molecule, mol_pair_off) the two-body damping
  // factors d_ij (d_ij C6_ij(0) with DevD4::qref_uniform) and the zero-charge
  // C6 of the three-body term; per molecule
  // the final two-body and the three-body energy
  double* d4_cn;
  double* d4_gw;
  double* d4_dgw;
  double* d4_e3;
  double* d4_disp;
  double* d4_c60;
  double* edisp2;
  double* edisp3;
  // Jacobi stopping rule: the solve ends after the sweep whose largest rotated
  // off-diagonal element is below this
  double jacobi_tol;
  // rotations whose off-diagonal element is at most this (absolute) are
  // identities: the element is set to zero and the workers skip the blocks
  // and MO rows they would have touched (0: every nonzero element rotates)
  double jacobi_skip;
  // > 0: per-iteration stopping threshold max(jacobi_tol, min(1e-2,
  // jacobi_adapt * previous charge residual))
  double jacobi_adapt;
  // eigensolver of the LDS-resident kernels: 1 Householder tridiagonalisation
  // + bisection + twisted-factorisation eigenvectors, 0 warm-started Jacobi
  int lds_eig;
  // reduction to tridiagonal form of the global-memory kernel (ld <=
  // TA_BAND_MAXLD): 1 two-stage (dense -> band -> tridiagonal), 0 one-stage
  // Householder
  int big_eig;
  // analytic gradient (all null unless a gradient is requested): per atom
  // dE/dR (3 each), dE/dCN of the GFN and of the D4 coordination number, the
  // atomic, dipole and quadrupole potentials of the final moments (1, 3, 6
  // each) and the D4 weights (TA_GD4 blocks of TA_D4_MAXREF: g_r(q),
  // dg_r/dq, dg_r/dCN at the final charge, g_r(0), dg_r(0)/dCN); per shell
  // the shell potential including the atomic one; per molecule (mol_mat_off,
  // column-major) the energy-weighted density W = sum_k f_k e_k c_k c_k^T;
  // per atom pair (mol_pair_off, i*nat + k) dC6_ik(0)/dCN_i of the
  // three-body term
  double* grad;
  double* g_dcn;
  double* g_dcn4;
  double* g_vat;
  double* g_vdp;
  double* g_vqp;
  double* g_d4;
  double* g_vsh;
  double* g_w;
  double* g_c6d;
  // periodic cells (all null for a batch of molecules): per molecule the
  // number of lattice translations (0 for a molecule), their offset into
  // trans (3 doubles each, in the order of the neighbour list's translation
  // index), the reciprocal vectors G != 0 of the Ewald sums (mol_ng,
  // mol_g_off into recip) and the cell constants (TA_PBC_EWALD doubles, see
  // TA_PBC_VOL ...); per neighbour-list
  // entry its translation index; per atom of a cell the three-body sites,
  // the images of atoms within the three-body cutoff (capacity-based offset
  // atm_site_off, count atm_site_n, 4 doubles each: R_site - R_atom and the
  // atom index; capacity atm_site_cap)
  const int* mol_ntr;
  const int64_t* mol_tr_off;
  const double* trans;
  const int* mol_ng;
  const int64_t* mol_g_off;
  const double* recip;
  const double* mol_ewald;
  int* nl_itr;
  const int64_t* atm_site_off;
  const int* atm_site_cap;
  int* atm_site_n;
  double* atm_site;
  // optional per-molecule phase clocks of the SCF kernel (TA_PROF_SLOTS each), may be null
  long long* prof;
  // optional completion records (one per molecule, pinned host memory), may be null
  TaDoneRecord* done;
};

// Periodic cells at the Gamma point.  Every real-space lattice sum runs over
// one translation table per cell, n_k in [-N_k, N_k] with
// N_k = ceil(max(TA_PBC_RMIN, Ewald cutoff) / d_k + fractional spread),
// n_1 outermost; the Ewald splitting parameter is sqrt(pi) / V^(1/3), its
// real-space cutoff TA_PBC_EWALD_X / alpha and its reciprocal cutoff
// 2 TA_PBC_EWALD_X alpha; the ES2 kernel is switched to 1/r between
// TA_PBC_ES2_RON and TA_PBC_ES2_ROFF (quintic), the AES2 damping corrections
// are summed out to TA_PBC_AES_CUT.
#define TA_PBC_RMIN 60.0
#define TA_PBC_ES2_RON 20.0
#define TA_PBC_ES2_ROFF 25.0
#define TA_PBC_AES_CUT 40.0
#define TA_PBC_EWALD_X 6.0
// per-cell constants: volume, splitting parameter, Ewald real-space cutoff,
// the reciprocal vectors b_k (a_j . b_k = 2 pi delta_jk), the spacings d_k of
// the lattice planes and the translation index ranges N_k
#define TA_PBC_VOL 0
#define TA_PBC_ALPHA 1
#define TA_PBC_RREAL 2
#define TA_PBC_REC 4
#define TA_PBC_DK 13
#define TA_PBC_NREP 16
#define TA_PBC_EWALD 20

// Translations of a cell's table that can lie within c of the separation d
// (R_i - R_j): the box n_k in [f_k - c / d_k, f_k + c / d_k] around the
// fractional separation f (|d - T| >= d_k |f_k - n_k|), widened by one and
// clipped to the table.  Table index of (n1, n2, n3):
// ((n1 + N1)(2 N2 + 1) + n2 + N2)(2 N3 + 1) + n3 + N3, ascending in the
// nested order n1, n2, n3, so a box walk visits translations in table order.
__host__ __device__ inline void cell_box(const double* cst, const double d[3], double c, int lo[3],
                                         int hi[3])
{
  for (int k = 0; k < 3; ++k) {
    const double* bk = cst + TA_PBC_REC + 3 * k;
    const double f = (d[0] * bk[0] + d[1] * bk[1] + d[2] * bk[2]) * (1.0 / (2.0 * M_PI));
    const double w = c / cst[TA_PBC_DK + k];
    const int nk = (int)cst[TA_PBC_NREP + k];
    const int l = (int)floor(f - w) - 1, h = (int)ceil(f + w) + 1;
    lo[k] = l > -nk ? l : -nk;
    hi[k] = h < nk ? h : nk;
  }
}

__host__ __device__ inline int cell_tindex(const double* cst, int n1, int n2, int n3)
{
  const int N1 = (int)cst[TA_PBC_NREP], N2 = (int)cst[TA_PBC_NREP + 1], N3 = (int)cst[TA_PBC_NREP + 2];
  return ((n1 + N1) * (2 * N2 + 1) + (n2 + N2)) * (2 * N3 + 1) + (n3 + N3);
}

// D4 weight blocks per atom of the gradient (BatchDev::g_d4)
#define TA_GD4 5

// SCF phase clock slots (wall clock ticks at 100 MHz; last two are counts)
#define TA_PROF_INIT 0
#define TA_PROF_MIX 1
#define TA_PROF_POT 2
#define TA_PROF_FOCK 3
#define TA_PROF_TRANSFORM 4
#define TA_PROF_JACOBI 5
#define TA_PROF_OCC 6
#define TA_PROF_DENSITY 7
#define TA_PROF_MULLIKEN 8
#define TA_PROF_ENERGY 9
#define TA_PROF_SWEEPS 10
#define TA_PROF_ITERS 11
// global-memory eigensolver phases (the Jacobi slot then holds the MO basis GEMM)
#define TA_PROF_TRIDIAG 12
#define TA_PROF_BISECT 13
#define TA_PROF_INVIT 14
#define TA_PROF_BACKTR 15
// steps of the blocked tridiagonalisation (column update, reflector, symmetric
// product, panel corrections, panel update) and the cluster re-orthogonalisation
#define TA_PROF_TRI_COLUMN 16
#define TA_PROF_TRI_REFLECTOR 17
#define TA_PROF_TRI_SYMV 18
#define TA_PROF_TRI_CORRECT 19
#define TA_PROF_TRI_PANEL 20
#define TA_PROF_CLUSTERS 21
// two-stage reduction: dense -> band, band -> tridiagonal, stage-2 back-transform
#define TA_PROF_BAND1 22
#define TA_PROF_BAND2 23
#define TA_PROF_Q2 24
#define TA_PROF_SLOTS 25

// Global-memory eigensolver constants: threads per workgroup, reflectors per
// back-transformation block, columns per tridiagonalisation panel, largest
// padded AO dimension served.
#define TA_BIG_NT 512
#define TA_BIG_NB 32
#define TA_BIG_NP 16
#define TA_BIG_MAXLD 1024

// rows of per-wave column sums kept in LDS by the tridiagonalisation: one per
// wave, or one per two waves above ld 768 (LDS budget).  The LDS layout of a
// size bucket reserves big_npart(ldmax) * ldmax doubles, so every molecule of
// the bucket uses the count of the bucket's largest ld.
__host__ __device__ inline int big_npart(int ld) { return ld <= 768 ? TA_BIG_NT / 64 : TA_BIG_NT / 128; }

// scratch per molecule of the global-memory path: inverse-iteration factors
// (3 ld x NT), back-transformation products (NB x ld), tridiagonalisation
// panel vectors (2 NP x ld)
// Offset, in a global-memory-path molecule's scratch, of the saved
// tridiagonal form of its last eigensolve (d, e, tau: 3 ld doubles, then |T|):
// the LDS copy is overwritten by the Mulliken phase, and the full eigenvector
// set of a requested wavefunction is formed from it after the SCF.
__host__ __device__ inline int64_t big_tri_save_offset(int ld)
{
  return 3 * (int64_t)ld * TA_BIG_NT + (int64_t)(TA_BIG_NB + 2 * TA_BIG_NP) * ld;
}

// Two-stage (band) reduction of the global-memory eigensolver, after the
// saved tridiagonal form: stage-1 reflectors (ld x ld), stage-2 reflectors
// (ld x (ld / 16 + 2) records of 16), two panel W blocks and the panel X
// block (16 x ld each), stage-1 tau (ld), the T factor of each stage-1 panel
// (16 x 16 per panel of 16 reflectors: 16 ld).
__host__ __device__ inline int64_t band_v1_offset(int ld)
{
  return big_tri_save_offset(ld) + 3 * (int64_t)ld + 8;
}
__host__ __device__ inline int64_t band_v2_offset(int ld)
{
  return band_v1_offset(ld) + (int64_t)ld * ld;
}
__host__ __device__ inline int64_t band_w_offset(int ld)
{
  return band_v2_offset(ld) + (int64_t)ld * (ld / 16 + 2) * 16;
}
__host__ __device__ inline int64_t band_x_offset(int ld) { return band_w_offset(ld) + 32 * (int64_t)ld; }
__host__ __device__ inline int64_t band_tau_offset(int ld) { return band_x_offset(ld) + 16 * (int64_t)ld; }

__host__ __device__ inline int64_t band_t1_offset(int ld) { return band_tau_offset(ld) + ld; }

// blocks of the stage-2 back-transformation of an order-ld band (group g of
// 16 sweeps has F + 1 - g blocks, F = (ld - 3) / 16), and their matrix-core
// operands (1024 doubles per block) after the stage-1 T factors
__host__ __device__ inline int64_t band_q2_blocks(int ld)
{
  if (ld < 3) return 0;
  const int64_t f = (ld - 3) / 16;
  return (f + 1) * (f + 2) / 2;
}
__host__ __device__ inline int64_t band_q2op_offset(int ld) { return band_t1_offset(ld) + 16 * (int64_t)ld; }

__host__ __device__ inline int64_t big_scratch_doubles(int ld)
{
  return band_q2op_offset(ld) + 1024 * band_q2_blocks(ld);
}

// largest padded AO dimension of the two-stage reduction (its panel strip
// is held in registers: up to 8 blocks of 16 rows per wave)
#define TA_BAND_MAXLD 1024

// LDS layout of the SCF kernel, in doubles (integer arrays in units of ints
// from the start of the region).  ld is the padded AO dimension of the
// bucket (multiple of 4); the two ld x ld matrices are only in LDS when
// mats is true.
struct ScfLds {
  int a, c;           // matrix offsets (doubles)
  int red, scal;      // reduction scratch (64) and broadcast scalars (16)
  int ev, focc, emo, fs;
  int rc, rs, rt;     // rotation cos, sin, tan (ld/2 each)
  int work;           // phase-shared vectors
  int eig;            // global-memory eigensolver region (overlays work)
  int iint;           // start of the int region, in ints
  int rank, inv, occ, rp, rq;  // int arrays (ints, relative to the LDS base)
  int total;          // doubles
};

__host__ __device__ inline ScfLds scf_lds_layout(int ld, int natmax, int nshmax, bool mats)
{
  ScfLds L;
  int off = 0;
  L.a = off;
  L.c = off + ld * ld;
  if (mats) off += 2 * ld * ld;
  L.red = off; off += 64;
  L.scal = off; off += 16;
  L.ev = off; off += ld;
  L.focc = off; off += ld;
  L.emo = off; off += ld;
  L.fs = off; off += ld;
  L.rc = L.rs = L.rt = off;
  // work: potentials (ld + nsh + 10 nat), Jacobi records (7 ld, double2
  // aligned), Mulliken partial sums (11 ld)
  // eigensolver region of the global-memory path: 9 ld persistent vectors
  // (d, e, tau, e^2, v, w, v', y, row sums), then either the GEMM tiles and
  // T^-1, T of a reflector block or the per-wave column sums of the
  // tridiagonalisation
  int w1 = ld + nshmax + 10 * natmax, w2 = 11 * ld;
  int t3 = 2048 + 2 * TA_BIG_NB * TA_BIG_NB, t4 = big_npart(ld) * ld;
  int w3 = mats ? 0 : 9 * ld + (t3 > t4 ? t3 : t4);
  // two-stage reduction of molecules below 96 AOs: the 64 x 65 product tile
  // at 8 ld, then the 560 panel scalars (above, they sit at 2 ld)
  const int wb = 8 * (ld < 92 ? ld : 92) + 64 * 65 + 560;
  if (!mats && w3 < wb) w3 = wb;
  L.work = off;
  L.eig = off;
  int wmax = w1 > w2 ? w1 : w2;
  off += (wmax > w3 ? wmax : w3);
  L.iint = 2 * off;
  L.rank = L.iint;
  L.inv = L.rank + ld;
  L.occ = L.inv + ld;
  L.rp = L.rq = L.occ + ld;
  off += (3 * ld + 1) / 2 + 1;
  L.total = off;
  return L;
}

// Levels whose occupation (electrons, both spins) is at or below this do not
// enter the density; on the global-memory paths their eigenvectors are not
// formed.  Their total weight is below n * 1e-15 electrons.
#define TA_OCC_MIN 1.0e-15

#define TA_STATUS_NOT_CONVERGED 1
#define TA_STATUS_CHOLESKY 2
#define TA_STATUS_EIGEN 4
#define TA_STATUS_NAN 8
#define TA_STATUS_MIXER 16
#define TA_STATUS_TOO_LARGE 32
#define TA_STATUS_FAILED 64

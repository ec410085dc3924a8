/*
 * TEST INFRASTRUCTURE -- not part of the engine.
 *
 * CPU FP64 restatement of tblite's GFN2-xTB energy singlepoint, following the
 * upstream Fortran that the reference embeds in its comments (file:line
 * citations on each function in gfn2_oracle.c).  Only tests/, the smoke check
 * and the CPU-baseline leg of bench.py load it, and only as the checker.
 */
#ifndef GFN2_ORACLE_H
#define GFN2_ORACLE_H

#include "tblite_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_calc_s or_calc;

/* Build the per-element basis and parameter tables from raw parameters. */
or_calc* or_new_calc(const tblite_amd_param* par);
void or_free_calc(or_calc* calc);

/* nsh/nao of a molecule; returns 0 or -1 for an unknown element. */
int or_dims(const or_calc* calc, int nat, const int* num, int* nsh, int* nao);

typedef struct {
  double energy;  /* total energy */
  double erep;    /* repulsion */
  double eelec;   /* electronic: band + entropy + ES2 + AES2 + ES3 + D4 two-body */
  double edisp2;  /* D4 two-body part of eelec (self-consistent charges) */
  double edisp3;  /* D4 three-body energy (zero charges); energy = erep + edisp3 + eelec */
  int iterations;
  int converged;
  int status;     /* TBLITE_AMD_STATUS_* bits */
  /* optional caller-owned outputs (may be NULL) */
  double* qsh;    /* nsh */
  double* qat;    /* nat */
  double* dpat;   /* 3*nat */
  double* qpat;   /* 6*nat */
  double* emo;    /* nao */
  double* focc;   /* nao */
  double* density; /* nao*nao, column-major: P of the last SCF iteration */
  double* coeff;   /* nao*nao, column-major: coeff(:, k) = level k (ascending) */
  double* gradient; /* 3*nat: dE/dR in Eh/bohr of the final energy */
  double* sigma;    /* 9: strain derivative sum_a g_a (x) R_a, column-major */
} or_result;

/* Full singlepoint (xtb_singlepoint, src/xtb.cu:188-434). kt in Eh. */
int or_singlepoint(const or_calc* calc, int nat, const int* num, const double* xyz, double charge,
                   int uhf, double accuracy, double kt, int max_iter, double damping,
                   or_result* res);

/* Singlepoint of a three-dimensional periodic cell at the Gamma point.
 * lat(9): lattice vectors a_k = lat[3k .. 3k+2] in bohr.  Gradients are not
 * built for cells (res->gradient / sigma must be NULL). */
int or_singlepoint_pbc(const or_calc* calc, int nat, const int* num, const double* xyz,
                       const double* lat, double charge, int uhf, double accuracy, double kt,
                       int max_iter, double damping, or_result* res);
/* Lattice tables of a cell: info = (V, alpha, real-space cutoff, |G|max,
 * number of translations, number of reciprocal vectors); the translations
 * tr(3, ntr) and reciprocal vectors g(3, ng) when the capacities suffice. */
int or_pbc_tables(const double* lat, int nat, const double* xyz, double* info, double* tr,
                  int tr_cap, double* g, int g_cap);
/* Neighbour list of a cell: entries (nlat, nltr) = (atom, translation index). */
int or_adjlist_pbc(int nat, const double* xyz, const double* lat, double cutoff, int* inl,
                   int* nnl, int* nlat, int* nltr, int capacity);
/* Classical terms of a cell (any output may be NULL): GFN CN, atom-resolved
 * repulsion, D4 CN, ES2 matrix (nsh, nsh), AES2 matrices (layout of
 * or_coulomb_matrices), multipole radii, D4 two-body damping matrix
 * dmat[i*nat + j] (image sums) and atom-resolved three-body energy. */
int or_pbc_terms(const or_calc* calc, int nat, const int* num, const double* xyz,
                 const double* lat, double* cn, double* erep, double* cn4, double* amat,
                 double* sd, double* dd, double* sq, double* mrad, double* dmat, double* e3);
/* Gamma-point S, D, Q, H of a cell (layouts of or_hamiltonian). */
int or_hamiltonian_pbc(const or_calc* calc, int nat, const int* num, const double* xyz,
                       const double* lat, double accuracy, double* S, double* D, double* Q,
                       double* H);
/* Test hook: scale factor of the Ewald splitting parameter (results must not
 * depend on it). */
void or_set_ewald_alpha_scale(double s);

/* Independent molecules over an OpenMP pool of nthreads (0: all cores).
 * Molecule m has mol_nat[m] atoms starting at atom offset sum(mol_nat[:m]). */
int or_singlepoint_many(const or_calc* calc, int nmol, const int* mol_nat, const int* num,
                        const double* xyz, double accuracy, double kt, int max_iter,
                        double damping, int nthreads, double* energies, int* iterations,
                        int* status);

/* ---- stage functions, exported for stage-level checks ---- */
/* D4: coordination numbers; reference weights of one atom (returns nref);
 * reference C6 (Casimir-Polder) of two element references (-1 if invalid);
 * per-atom two-body energy at charges qat, three-body energy (zero charges)
 * and the two-body potential dE2/dq. */
int or_d4_ncoord(const or_calc* calc, int nat, const int* num, const double* xyz, double* cn);
int or_d4_weights(const or_calc* calc, int z, double cn, double q, double* gw, double* dgw);
double or_d4_c6ref(const or_calc* calc, int zi, int ri, int zj, int rj);
int or_d4_eval(const or_calc* calc, int nat, const int* num, const double* xyz, const double* qat,
               double* e2, double* e3, double* vat);
void or_multipole_3d(const double rpj[3], const double rpi[3], double aj, double ai,
                     const int lj[3], const int li[3], const double* s1d, double* s3d,
                     double d3d[3], double q3d[6]);
void or_overlap_1d_table(double eab, int lmax, double* s1d);
/* cart -> spherical for one shell pair; cart(mlao_j, mlao_i), sphr(msao_j, msao_i) */
void or_transform0(int lj, int li, const double* cart, double* sphr);
/* Contracted multipole block: S(msao_j, msao_i), D(3, msao_j, msao_i),
 * Q(6, msao_j, msao_i) (Fortran order). Species index into the calculator,
 * shell index within the species. */
int or_multipole_cgto(const or_calc* calc, int zj, int jsh, int zi, int ish, double r2,
                      const double vec[3], double intcut, double* S, double* D, double* Q);
/* Contracted basis of shell ish of element z: ang, nprim, alpha(nprim),
 * coeff(nprim) (coefficients include the primitive norms). */
int or_get_cgto(const or_calc* calc, int z, int ish, int* ang, int* nprim, double* alpha,
                double* coeff);
int or_ncoord(const or_calc* calc, int nat, const int* num, const double* xyz, double* cn);
int or_repulsion(const or_calc* calc, int nat, const int* num, const double* xyz, double* erep);
double or_get_cutoff(const or_calc* calc, int nat, const int* num, double accuracy);
/* Half neighbour list; returns nnz (or -(needed) when capacity is too small). */
int or_adjlist(int nat, const double* xyz, double cutoff, int* inl, int* nnl, int* nlat,
               int capacity);
int or_selfenergy(const or_calc* calc, int nat, const int* num, const double* cn, double* se);
int or_hamiltonian(const or_calc* calc, int nat, const int* num, const double* xyz,
                   double accuracy, double* S, double* D, double* Q, double* H);
int or_coulomb_matrices(const or_calc* calc, int nat, const int* num, const double* xyz,
                        double* amat, double* amat_sd, double* amat_dd, double* amat_sq,
                        double* mrad);
/* Generalized eigenproblem F C = S C e (dsygvd itype 1, ascending e). */
int or_sygvd(int n, const double* F, const double* S, double* e, double* C);
void or_fermi_filling(double nel, double kt, int n, const double* emo, int* homo, double* focc,
                      double* e_fermi);
double or_entropy(int n, const double* focc, double kt);
/* Electrostatic energy and potential for given charges / multipoles:
 * e(nat) atom-resolved energies; vsh(nsh), vat(nat), vdp(3,nat), vqp(6,nat). */
int or_coulomb_eval(const or_calc* calc, int nat, const int* num, const double* xyz,
                    const double* qsh, const double* dpat, const double* qpat, double* e,
                    double* vsh, double* vat, double* vdp, double* vqp);

/* ---- gradient stages ----
 * Contracted multipole block and its derivatives by vec = R_i - R_j:
 * dS(3, sj, si), dD(3, 3, sj, si), dQ(3, 6, sj, si) (derivative index first). */
int or_multipole_grad_cgto(const or_calc* calc, int zj, int jsh, int zi, int ish, double r2,
                           const double vec[3], double intcut, double* S, double* D, double* Q,
                           double* dS, double* dD, double* dQ);
/* Classical term energy and gradient at fixed moments: which = 0 repulsion,
 * 1 electrostatics, 2 D4 two-body, 3 D4 three-body. */
int or_classical_gradient(const or_calc* calc, int nat, const int* num, const double* xyz, int which,
                          const double* qsh, const double* dpat, const double* qpat, double* e,
                          double* grad);
/* Integral terms at fixed P, W and potentials: e = sum P H0 - sum W S
 * + sum vsh q + sum vdp dpat + sum vqp qpat (Mulliken moments of P without
 * reference populations) and its gradient. */
int or_integral_gradient(const or_calc* calc, int nat, const int* num, const double* xyz,
                         double accuracy, const double* P, const double* W, const double* vsh,
                         const double* vdp, const double* vqp, double* e, double* grad);

/* Mixer on an affine map F(q) = A q + b (test hook for broyden_next). */
int or_broyden_linear(int n, const double* A, const double* b, const double* q0, double damp,
                      double omega0, int memory, int nstep, double* qs, double* res);
/* H0 scaling of a shell pair (element indices into the parameter table). */
double or_hscale(const or_calc* calc, int ei, int ish, int ej, int jsh);

#ifdef __cplusplus
}
#endif

#endif

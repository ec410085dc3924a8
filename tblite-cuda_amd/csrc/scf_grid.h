// Multi-workgroup SCF path (one molecule across the whole GPU): device view
// of the molecule and its work arrays, shared by scf_grid.hip and the
// engine's host code.
#pragma once

#include <stdint.h>

#include "common.h"

// device scalars (GridMol::sc) and integer words (GridMol::ic)
enum {
  GS_WARM = 0, GS_ESUM, GS_GL, GS_GU, GS_TNORM, GS_PIVMIN, GS_TOL, GS_TS, GS_ERR, GS_EBAND,
  GS_EREP, GS_NALP, GS_NBET, GS_E3, GS_ED2, GS_COUNT = 16
};
enum { GI_STATUS = 0, GI_CONVERGED, GI_NOCC, GI_NRUNS, GI_SHIFT = 4, GI_COUNT = 8 };

struct GridMol {
  int n, ld, nat, nsh, ndim, mem, nbmax;
  double kt, damping, econv, pconv, charge;
  int uhf;
  const DevParams* par;
  // molecule (local indices)
  const int *el, *sh_at, *sh_ish, *sh_iao, *ao_at, *ao_sh, *at_ish, *at_iao;
  const double *S, *H, *D, *Q, *gam, *sd, *dd, *sq, *n0sh, *erep_at;
  double *qsh, *qat, *dpat, *qpat;
  // mixer
  double *qin, *dq, *dqlast, *qlast, *dfm, *um, *amix, *beta, *omega, *cvec;
  // matrices (ld x ld)
  double *A, *X, *Y, *CT;
  // scratch
  double *scr;                    // inverse iteration factors, 3 n^2
  double *vt, *wt, *wbt, *part;   // panel vectors, W of the back-transform, symv partials
  double *svp;                    // per 64-row block partial sums of y.v (nbmax)
  double *dpart;                  // per 64-row block partial panel dot products (nbmax x 2 NP)
  double *part11, *epart;
  double *d, *e, *tau, *e2, *lam, *prevlam, *v, *y, *focc, *fs, *vsh, *vat, *vdp, *vqp, *dots, *tm;
  double *sc;
  int *ic, *runs;
  // D4 dispersion: per atom CN, weights g_r(q) and dg_r/dq (TA_D4_MAXREF each),
  // three-body chunk energies, per-atom two-body energies of the iteration;
  // per atom pair the two-body damping factors; per molecule result slots
  int d4_on;  // host copy of the parameter set's D4 switch
  const double *d4_cn, *d4_disp, *d4_e3;
  double *d4_gw, *d4_dgw, *ed4;
  double *edisp2, *edisp3;
};

// doubles of scratch a molecule of the grid path needs beyond its four ld x ld matrices
inline int64_t grid_scratch_doubles(int n, int ld, int nat, int nsh)
{
  const int64_t nbm = (n + 63) / 64;
  int64_t s = 3 * (int64_t)n * n;                          // scr
  s += (int64_t)(2 * TA_BIG_NP + TA_BIG_NB) * ld;          // vt, wt, wbt
  s += nbm * nbm * 64 + nbm;                               // symv partials, y.v partials
  s += nbm * 2 * TA_BIG_NP;                                // panel dot partials
  s += 11 * (int64_t)n + (nsh + nat);                      // Mulliken partials, energy parts
  s += 10 * (int64_t)ld + nsh + 11 * (int64_t)nat;         // vectors
  s += 2 * TA_BIG_NP + TA_BIG_NB * TA_BIG_NB + GS_COUNT;   // dots, tm, scalars
  s += (GI_COUNT + ld + 4 + 1) / 2 + 16;                   // int words and runs (as doubles)
  return s + 64;
}

struct GridHostStats {
  int iterations, status;
};

// Runs the SCF of one molecule on the grid path (stream s); writes the
// molecule's energy, electronic energy, repulsion, iterations and status
// into the given device slots and its levels / occupations (and optionally
// the coefficients and the density matrix, n x n column-major each).  verbose != 0 prints one line per SCF iteration to stderr.
// use_graph != 0 replays the eigenvalue phase of iterations 2.. from a HIP
// graph captured in iteration 2.  Returns 0 or < 0 on a HIP failure.
int grid_scf(GridMol g, hipStream_t s, int max_iter, double* d_energy, double* d_eelec,
             double* d_erep, int* d_iter, int* d_status, double* emo, double* focc, double* coeff,
             double* density, GridHostStats* hs, int verbose, int use_graph);

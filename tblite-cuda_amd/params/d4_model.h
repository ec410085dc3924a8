/*
 * Built-in D4 dispersion data for H, C, N, O and S.
 *
 * Pure data.  The damping parameters are GFN2's published D4 values
 * (s6 = 1, s8 = 2.7, a1 = 0.52, a2 = 5.0, s9 = 5.0); the CN constants,
 * charge-scaling constants, cutoffs, chemical hardnesses and
 * sqrt(0.5 <r^4>/<r^2> sqrt(Z)) expectation values are the D4 model's
 * published ones as far as they could be restated without the upstream
 * files.
 *
 * The reference polarisabilities are NOT dftd4's tables (neither dftd4 nor
 * its data is in the reference or anywhere in this build environment).  Each
 * reference system is a single-pole (London) model,
 *     alpha(i w) = alpha0 / (1 + (w / w0)^2),
 * with alpha0 and w0 chosen so that the same-reference C6 = 3/4 alpha0^2 w0
 * lands near literature values for the free atom and for typical bonding
 * situations (free H 6.5, H in X-H 3.0, sp3 C 18, water O 12 Eh bohr^6,
 * ...); reference charges are zero.  Dispersion energies therefore have the
 * right magnitude but are not tblite's.  A caller holding dftd4's tables fills
 * tblite_amd_param.d4 directly: the engine only ever reads alpha(i w).
 */
#ifndef TBLITE_AMD_D4_MODEL_H
#define TBLITE_AMD_D4_MODEL_H

#include "tblite_amd.h"

/* imaginary frequencies of the alpha(i w) tables / Eh */
static const double tblite_amd_d4_freq[TBLITE_AMD_D4_NFREQ] = {
    0.000001, 0.050000, 0.100000, 0.200000, 0.300000, 0.400000, 0.500000, 0.600000,
    0.700000, 0.800000, 0.900000, 1.000000, 1.200000, 1.400000, 1.600000, 1.800000,
    2.000000, 2.500000, 3.000000, 4.000000, 5.000000, 7.500000, 10.00000};

static inline void tblite_amd_d4_single_pole(tblite_amd_d4_element* e, int r, double cn,
                                             double alpha0, double w0)
{
  e->cn[r] = cn;
  e->q[r] = 0.0;
  for (int k = 0; k < TBLITE_AMD_D4_NFREQ; ++k) {
    const double x = tblite_amd_d4_freq[k] / w0;
    e->alpha[r][k] = alpha0 / (1.0 + x * x);
  }
}

/* Fill p->d4 for the elements already in p->elem (H, C, N, O, S are known;
 * any other element gets no references and makes the calculator fail). */
static inline void tblite_amd_fill_d4_model(tblite_amd_param* p)
{
  tblite_amd_d4_param* d = &p->d4;
  d->enabled = 1;
  d->s6 = 1.0;
  d->s8 = 2.7;
  d->a1 = 0.52;
  d->a2 = 5.0;
  d->s9 = 5.0;
  d->alp = 16.0;
  d->ga = 3.0;
  d->gc = 2.0;
  d->wf = 6.0;
  d->cn_k4 = 4.10451;
  d->cn_k5 = 19.08857;
  d->cn_k6 = 2.0 * 11.28174 * 11.28174;
  d->cn_kn = 7.5;
  d->cn_cutoff = 30.0;
  d->disp2_cutoff = 60.0;
  d->disp3_cutoff = 40.0;
  for (int i = 0; i < p->nelem; ++i) {
    tblite_amd_d4_element* e = &d->elem[i];
    switch (p->elem[i].z) {
      case 1:
        e->nref = 2; e->zeff = 1.0; e->eta = 0.47259288; e->r4r2 = 2.00734898;
        tblite_amd_d4_single_pole(e, 0, 0.0000, 4.50, 0.428);
        tblite_amd_d4_single_pole(e, 1, 0.9118, 2.75, 0.530);
        break;
      case 6:
        e->nref = 5; e->zeff = 6.0; e->eta = 0.42195412; e->r4r2 = 3.10492822;
        tblite_amd_d4_single_pole(e, 0, 0.0000, 12.0, 0.431);
        tblite_amd_d4_single_pole(e, 1, 0.9868, 9.60, 0.434);
        tblite_amd_d4_single_pole(e, 2, 1.9985, 8.90, 0.433);
        tblite_amd_d4_single_pole(e, 3, 2.9987, 8.00, 0.458);
        tblite_amd_d4_single_pole(e, 4, 3.9844, 7.20, 0.466);
        break;
      case 7:
        e->nref = 4; e->zeff = 7.0; e->eta = 0.50438193; e->r4r2 = 2.71175247;
        tblite_amd_d4_single_pole(e, 0, 0.0000, 7.40, 0.589);
        tblite_amd_d4_single_pole(e, 1, 0.9987, 6.60, 0.581);
        tblite_amd_d4_single_pole(e, 2, 2.0052, 6.20, 0.572);
        tblite_amd_d4_single_pole(e, 3, 2.9917, 6.00, 0.574);
        break;
      case 8:
        e->nref = 3; e->zeff = 8.0; e->eta = 0.58691863; e->r4r2 = 2.59361680;
        tblite_amd_d4_single_pole(e, 0, 0.0000, 5.30, 0.740);
        tblite_amd_d4_single_pole(e, 1, 0.9926, 5.00, 0.683);
        tblite_amd_d4_single_pole(e, 2, 1.9887, 4.70, 0.724);
        break;
      case 16:
        e->nref = 4; e->zeff = 16.0; e->eta = 0.38924725; e->r4r2 = 4.04108902;
        tblite_amd_d4_single_pole(e, 0, 0.0000, 19.6, 0.465);
        tblite_amd_d4_single_pole(e, 1, 0.9988, 17.5, 0.479);
        tblite_amd_d4_single_pole(e, 2, 1.9984, 16.5, 0.490);
        tblite_amd_d4_single_pole(e, 3, 2.9886, 16.0, 0.495);
        break;
      default:
        e->nref = 0;
        break;
    }
  }
}

#endif

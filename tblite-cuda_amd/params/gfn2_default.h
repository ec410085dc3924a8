/*
 * Built-in GFN2-xTB parameter data for H, C, N, O and S.
 *
 * Pure data.  Upstream tblite keeps these values in its GFN2 parameter module
 * / parameter file; neither is present in the reference (the calculator there
 * is zero-initialised, src/xtb.cu:439, and new_hamiltonian is declared but
 * never defined, src/h0.h:80-86) nor anywhere else in this build environment.
 * The numbers below are the published GFN2-xTB values as far as they could be
 * restated without the upstream file; entries marked [est] are estimates of
 * the right magnitude and sign.  Absolute energies are therefore not
 * comparable with tblite; every algorithm is parameter-agnostic and takes
 * this table (or a caller-supplied one) as input.
 *
 * Units: levels and kcn in eV (converted with 1 Eh = 27.21138602 eV);
 * hardness in Eh; radii in bohr (converted from angstrom with
 * 1 bohr = 0.52917721090380 angstrom).
 */
#ifndef TBLITE_AMD_GFN2_DEFAULT_H
#define TBLITE_AMD_GFN2_DEFAULT_H

#include <string.h>

#include "tblite_amd.h"
#include "d4_model.h"

#define TBLITE_AMD_AUTOAA 0.52917721090380
#define TBLITE_AMD_AUTOEV 27.21138602
#define TBLITE_AMD_KB_EH 3.166808578545117e-06 /* Boltzmann constant, Eh/K */

static inline void tblite_amd_fill_gfn2_default(tblite_amd_param* p)
{
  memset(p, 0, sizeof(*p));
  /* shell-pair scaling: ss 1.85, pp 2.23, dd 2.23, sp = (ss+pp)/2, sd = pd = 2.00 */
  const double ks[3][3] = {{1.85, 2.04, 2.00}, {2.04, 2.23, 2.00}, {2.00, 2.00, 2.23}};
  memcpy(p->kshell, ks, sizeof(ks));
  p->enscale = 2.0e-2;
  p->wexp = 0.5;
  p->kpol = 2.0;
  p->rep_kexp = 1.5;
  p->rep_kexp_light = 1.0;
  p->rep_rexp = 1.0;
  p->rep_cutoff = 25.0;
  p->gexp = 2.0;
  p->thirdorder_shell[0] = 1.0;
  p->thirdorder_shell[1] = 0.5;
  p->thirdorder_shell[2] = 0.25;
  p->dmp3 = 3.0;
  p->dmp5 = 4.0;
  p->mp_shift = 1.2;
  p->mp_kexp = 4.0;
  p->mp_rmax = 5.0;
  p->cn_ka = 10.0;
  p->cn_kb = 20.0;
  p->cn_rshift = 2.0;
  p->cn_cutoff = 25.0;

  const double aa = 1.0 / TBLITE_AMD_AUTOAA;
  /* D3 covalent radii (Pyykko & Atsumi 2009, scaled by 4/3) and xtb atomic radii, angstrom */
  p->nelem = 5;
  tblite_amd_element* e;

  e = &p->elem[0]; /* H */
  e->z = 1; e->nsh = 1;
  e->ang[0] = 0; e->pqn[0] = 1; e->ngauss[0] = 3;
  e->levels[0] = -10.707211; e->slater[0] = 1.230000; e->refocc[0] = 1.0;
  e->shpoly[0] = -0.953618e-2; e->kcn[0] = -0.05; e->lgam[0] = 1.0;
  e->gam = 0.405771; e->gam3 = 0.080000;
  e->zeff = 1.105388; e->arep = 2.213717; e->en = 2.20;
  e->dkernel = 5.563889e-2; e->qkernel = 0.0;
  e->mprad = 1.4; e->mpvcn = 1.0;
  e->rcov = 0.32 * 4.0 / 3.0 * aa; e->atomic_rad = 0.32 * aa;

  e = &p->elem[1]; /* C */
  e->z = 6; e->nsh = 2;
  e->ang[0] = 0; e->ang[1] = 1; e->pqn[0] = 2; e->pqn[1] = 2; e->ngauss[0] = 4; e->ngauss[1] = 4;
  e->levels[0] = -13.970922; e->levels[1] = -10.063292;
  e->slater[0] = 2.096432; e->slater[1] = 1.800000;
  e->refocc[0] = 1.0; e->refocc[1] = 3.0;
  e->shpoly[0] = -2.294321e-2; e->shpoly[1] = -0.271102e-2;
  e->kcn[0] = -1.02144e-2; e->kcn[1] = 1.61657e-2; /* [est] */
  e->lgam[0] = 1.0; e->lgam[1] = 1.1056358;
  e->gam = 0.538015; e->gam3 = 0.150000;
  e->zeff = 4.231078; e->arep = 1.247655; e->en = 2.55;
  e->dkernel = -0.411674e-2; e->qkernel = 2.13583e-2;
  e->mprad = 3.0; e->mpvcn = 3.0;
  e->rcov = 0.75 * 4.0 / 3.0 * aa; e->atomic_rad = 0.75 * aa;

  e = &p->elem[2]; /* N */
  e->z = 7; e->nsh = 2;
  e->ang[0] = 0; e->ang[1] = 1; e->pqn[0] = 2; e->pqn[1] = 2; e->ngauss[0] = 4; e->ngauss[1] = 4;
  e->levels[0] = -16.686243; e->levels[1] = -12.523956;
  e->slater[0] = 2.339881; e->slater[1] = 2.014332;
  e->refocc[0] = 1.5; e->refocc[1] = 3.5;
  e->shpoly[0] = -8.506003e-2; e->shpoly[1] = -2.504201e-2;
  e->kcn[0] = -1.53e-2; e->kcn[1] = 1.52e-2; /* [est] */
  e->lgam[0] = 1.0; e->lgam[1] = 1.1164892;
  e->gam = 0.461493; e->gam3 = 0.130000;
  e->zeff = 5.242592; e->arep = 1.682689; e->en = 3.04;
  e->dkernel = -0.331949e-2; e->qkernel = 1.50e-2; /* qkernel [est] */
  e->mprad = 1.9; e->mpvcn = 2.0;
  e->rcov = 0.71 * 4.0 / 3.0 * aa; e->atomic_rad = 0.71 * aa;

  e = &p->elem[3]; /* O */
  e->z = 8; e->nsh = 2;
  e->ang[0] = 0; e->ang[1] = 1; e->pqn[0] = 2; e->pqn[1] = 2; e->ngauss[0] = 4; e->ngauss[1] = 4;
  e->levels[0] = -20.229985; e->levels[1] = -15.503117;
  e->slater[0] = 2.439742; e->slater[1] = 2.137023;
  e->refocc[0] = 2.0; e->refocc[1] = 4.0;
  e->shpoly[0] = -14.955291e-2; e->shpoly[1] = -3.350819e-2;
  e->kcn[0] = -1.85e-2; e->kcn[1] = 1.18e-2; /* [est] */
  e->lgam[0] = 1.0; e->lgam[1] = 1.1426302;
  e->gam = 0.451896; e->gam3 = 0.1053856;
  e->zeff = 5.784415; e->arep = 2.165712; e->en = 3.44;
  e->dkernel = -0.435e-2; e->qkernel = 0.085e-2; /* [est] */
  e->mprad = 1.8; e->mpvcn = 2.0;
  e->rcov = 0.63 * 4.0 / 3.0 * aa; e->atomic_rad = 0.64 * aa;

  e = &p->elem[4]; /* S */
  e->z = 16; e->nsh = 3;
  e->ang[0] = 0; e->ang[1] = 1; e->ang[2] = 2;
  e->pqn[0] = 3; e->pqn[1] = 3; e->pqn[2] = 3;
  e->ngauss[0] = 4; e->ngauss[1] = 4; e->ngauss[2] = 3;
  e->levels[0] = -20.029654; e->levels[1] = -11.377694; e->levels[2] = -0.420282;
  e->slater[0] = 2.506934; e->slater[1] = 1.992775; e->slater[2] = 1.964867;
  e->refocc[0] = 2.0; e->refocc[1] = 4.0; e->refocc[2] = 0.0;
  e->shpoly[0] = -7.455e-2; e->shpoly[1] = -1.862e-2; e->shpoly[2] = 4.000e-2; /* [est] */
  e->kcn[0] = 0.50e-2; e->kcn[1] = 1.00e-2; e->kcn[2] = 1.00e-2;               /* [est] */
  e->lgam[0] = 1.0; e->lgam[1] = 0.9; e->lgam[2] = 1.0;                          /* [est] */
  e->gam = 0.338; e->gam3 = 0.050;                                               /* [est] */
  e->zeff = 15.49; e->arep = 1.19; e->en = 2.58;                                 /* [est] */
  e->dkernel = -0.181e-2; e->qkernel = 0.655e-2;                                 /* [est] */
  e->mprad = 3.4; e->mpvcn = 2.0;                                                /* [est] */
  e->rcov = 1.02 * 4.0 / 3.0 * aa; e->atomic_rad = 1.04 * aa;

  tblite_amd_fill_d4_model(p);
}

#endif

// Fermi smearing of one spin channel on one wave (get_fermi_filling,
// src/wavefunction/fermi.cu:6-124 text), shared by the SCF kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "device_util.h"

// ----------------------------------------------------------------- Fermi
// get_fermi_filling (src/wavefunction/fermi.cu:100-124) on sorted levels,
// run by one wave; occ receives this spin channel's occupations, returns
// the electronic entropy term kT * sum(f ln f + (1-f) ln(1-f)).
__device__ inline double fermi_wave(double nel, double kt, int n, const double* emo, double* occ)
{
  const int lane = threadIdx.x & 63;
  const double thr = 1.4901161193847656e-08; // sqrt(epsilon(1.0_wp))
  int homo = (int)floor(nel);
  const double frac = fmod(nel, 1.0);
  for (int i = lane; i < n; i += 64) occ[i] = (i < homo) ? 1.0 : ((i == homo) ? frac : 0.0);
  int homo_idx = homo;
  if (frac > 0.5) homo_idx += 1;
  if (homo_idx > 0) {
    const double occt = (double)homo_idx;
    const int lo = homo_idx - 1 < 0 ? 0 : homo_idx - 1;
    const int hi = homo_idx < n - 1 ? homo_idx : n - 1;
    double ef = 0.5 * (emo[lo] + emo[hi]);
    for (int cyc = 0; cyc < 200; ++cyc) {
      double tn = 0.0, td = 0.0;
      for (int i = lane; i < n; i += 64) {
        const double x = (emo[i] - ef) / kt;
        double f = 0.0, dfn = 0.0;
        if (x < 50.0) {
          const double ex = exp(x);
          f = 1.0 / (ex + 1.0);
          dfn = ex / (kt * (ex + 1.0) * (ex + 1.0));
        }
        occ[i] = f;
        tn += f;
        td += dfn;
      }
      tn = wave_sum(tn);
      td = wave_sum(td);
      ef += (occt - tn) / td;
      if (fabs(occt - tn) <= thr) break;
    }
  }
  double s = 0.0;
  for (int i = lane; i < n; i += 64) {
    const double f = occ[i];
    if (f > thr && 1.0 - f > thr) s += f * log(f) + (1.0 - f) * log(1.0 - f);
  }
  return wave_sum(s) * kt;
}


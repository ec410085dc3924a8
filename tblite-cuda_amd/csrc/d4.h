// D4 dispersion pieces shared by the setup and SCF kernels (self-consistent
// D4 of GFN2: tblite's d4_dispersion, whose get_potential / get_energy hooks
// the reference leaves as comments in next_scf, src/scf/iterators.cu:353-354
// and :415-417, and whose non-self-consistent part is dispersion_get_engrad,
// src/xtb.cu:274-277).
//
//   CN_i    = sum_j k4 exp(-(|EN_i - EN_j| + k5)^2 / k6) (1 + erf(-kn (r - rc) / rc)) / 2
//   g_ir(q) = w_ir(CN_i) zeta_ir(q_i)
//   E2      = 1/2 sum_{i != j} sum_rs g_ir d_ij C6ref(ir, js) g_js
//   v_i     = dE2/dq_i = sum_r dg_ir/dq_i sum_j d_ij sum_s C6ref(ir, js) g_js
// with d_ij = -(s6 / (r^6 + R0^6) + s8 3 rr / (r^8 + R0^8)) stored per atom
// pair (DevD4::c6 holds C6ref from the Casimir-Polder integral, built on the
// host).  The three-body term uses the weights at zero charge.  When every
// element's references share one reference charge, zeta does not depend on
// r, g_ir(q) = w_ir s_i(q) zeta_i(0) with s = zeta(q)/zeta(0), and
//   E2 = 1/2 sum_{i != j} s_i s_j [d_ij C6_ij(0)],  v_i = ds_i sum_j [d_ij C6_ij(0)] s_j
// with the bracket stored per pair: one load per pair and iteration.
#pragma once

#include "common.h"

// charge scale zeta(q) / zeta(0) and its q derivative for an element whose
// references share the reference charge qref[e][0]
__device__ __forceinline__ void d4_scale(const DevD4& d, int e, double q, double& s, double& ds)
{
  const double gi = d.gi[e], zi = d.zeff[e], qref = d.qref[e][0] + zi;
  const double z0 = exp(d.ga * (1.0 - exp(gi * (1.0 - qref / zi))));
  const double qmod = q + zi;
  if (qmod < 0.0) {
    s = exp(d.ga) / z0;
    ds = 0.0;
  } else {
    const double x = exp(gi * (1.0 - qref / qmod));
    s = exp(d.ga * (1.0 - x)) / z0;
    ds = -d.ga * gi * x * s * qref / (qmod * qmod);
  }
}

// number of per-atom weight entries the SCF kernels carry
__device__ __forceinline__ int d4_nw(const DevD4& d, int e) { return d.qref_uniform ? 1 : d.nref[e]; }

// reference weights g_r and dg_r/dq of one atom of element e
__device__ __forceinline__ void d4_weights(const DevD4& d, int e, double cn, double q, double* gw,
                                           double* dgw)
{
  const int nr = d.nref[e];
  double norm = 0.0, cnmax = -1.0e300;
  for (int r = 0; r < nr; ++r) {
    const double t = cn - d.cnref[e][r];
    norm += exp(-d.wf * t * t);
    cnmax = fmax(cnmax, d.cnref[e][r]);
  }
  const double gi = d.gi[e], zi = d.zeff[e];
  const double qmod = q + zi;
  for (int r = 0; r < nr; ++r) {
    const double t = cn - d.cnref[e][r];
    double w = exp(-d.wf * t * t) / norm;
    // every Gaussian underflowed: the reference with the largest CN takes it all
    if (!isfinite(w)) w = d.cnref[e][r] == cnmax ? 1.0 : 0.0;
    double z, dz;
    if (qmod < 0.0) {
      z = exp(d.ga);
      dz = 0.0;
    } else {
      const double qref = d.qref[e][r] + zi;
      const double x = exp(gi * (1.0 - qref / qmod));
      z = exp(d.ga * (1.0 - x));
      dz = -d.ga * gi * x * z * qref / (qmod * qmod);
    }
    gw[r] = w * z;
    dgw[r] = w * dz;
  }
}

// CN derivative of the reference weights of d4_weights, dg_r/dCN = dw_r/dCN
// zeta_r(q) (zero where the weights fall back to the largest-CN reference)
__device__ __forceinline__ void d4_weights_dcn(const DevD4& d, int e, double cn, double q,
                                               double* dgdcn)
{
  const int nr = d.nref[e];
  double norm = 0.0, dnorm = 0.0;
  for (int r = 0; r < nr; ++r) {
    const double t = cn - d.cnref[e][r];
    const double x = exp(-d.wf * t * t);
    norm += x;
    dnorm += -2.0 * d.wf * t * x;
  }
  const double gi = d.gi[e], zi = d.zeff[e];
  const double qmod = q + zi;
  for (int r = 0; r < nr; ++r) {
    const double t = cn - d.cnref[e][r];
    const double x = exp(-d.wf * t * t);
    const double w = x / norm;
    const double dw = (-2.0 * d.wf * t * x * norm - x * dnorm) / (norm * norm);
    if (!isfinite(w) || !isfinite(dw)) {
      dgdcn[r] = 0.0;
      continue;
    }
    double z;
    if (qmod < 0.0) {
      z = exp(d.ga);
    } else {
      const double qref = d.qref[e][r] + zi;
      z = exp(d.ga * (1.0 - exp(gi * (1.0 - qref / qmod))));
    }
    dgdcn[r] = dw * z;
  }
}

// two-body damping factor d(r) = -(s6 / (r^6 + R0^6) + s8 rr / (r^8 + R0^8))
// of elements ei, ej at squared distance r2, and (dd/dr) / r
__device__ __forceinline__ void d4_damp(const DevD4& d, int ei, int ej, double r2, double& dmp,
                                        double& ddr)
{
  const double rr = 3.0 * d.r4r2[ei] * d.r4r2[ej];
  const double r0 = d.a1 * sqrt(rr) + d.a2;
  const double r02 = r0 * r0, r06 = r02 * r02 * r02, r08 = r06 * r02;
  const double r6 = r2 * r2 * r2, r8 = r6 * r2;
  const double t6 = 1.0 / (r6 + r06), t8 = 1.0 / (r8 + r08);
  dmp = -(d.s6 * t6 + d.s8 * rr * t8);
  ddr = d.s6 * 6.0 * r2 * r2 * t6 * t6 + d.s8 * rr * 8.0 * r6 * t8 * t8;
}

// per-atom SCF weights: g_r and dg_r/dq, or (qref_uniform) s and ds/dq
__device__ __forceinline__ void d4_atom_weights(const DevD4& d, int e, double cn, double q,
                                                double* gw, double* dgw)
{
  if (d.qref_uniform)
    d4_scale(d, e, q, gw[0], dgw[0]);
  else
    d4_weights(d, e, cn, q, gw, dgw);
}

// sum_s C6ref(ea r, ei s) g_is for r < nref(ea), accumulated with factor f
__device__ __forceinline__ void d4_accumulate(const DevD4& d, int ea, int ei, double f,
                                              const double* __restrict__ gi,
                                              double (&acc)[TA_D4_MAXREF])
{
  const int na = d.nref[ea], ni = d.nref[ei];
  const double* c6 = &d.c6[ea][ei][0][0];
  for (int s = 0; s < ni; ++s) {
    const double g = f * gi[s];
#pragma unroll
    for (int r = 0; r < TA_D4_MAXREF; ++r)
      if (r < na) acc[r] += c6[r * TA_D4_MAXREF + s] * g;
  }
}

// SCF pair term: d4_accumulate on the weights of d4_atom_weights, or
// (qref_uniform) acc[0] += f s_i, the pair factor f already holding C6_ai(0)
__device__ __forceinline__ void d4_scf_accumulate(const DevD4& d, int ea, int ei, double f,
                                                  const double* __restrict__ gi,
                                                  double (&acc)[TA_D4_MAXREF])
{
  if (d.qref_uniform)
    acc[0] += f * gi[0];
  else
    d4_accumulate(d, ea, ei, f, gi, acc);
}

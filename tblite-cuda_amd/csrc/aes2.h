// Anisotropic electrostatics (AES2) atom-pair kernels of GFN2: the
// charge-dipole, dipole-dipole and charge-quadrupole interaction matrices
// amat_sd, amat_dd, amat_sq of get_multipole_matrix_0d (src/coulomb.cu:439-467
// text, :429-510 code).  For the pair (row i,
// column j) with v = R_i - R_j, r = |v| and the damping radius
// rr = (mrad_j + mrad_i) / (2 r):
//   sd_k  = v_k g3 f3
//   dd_kl = delta_kl g3 f5 - 3 v_k v_l g5 f5
//   sq    = (v0 v0, 2 v0 v1, v1 v1, 2 v0 v2, 2 v1 v2, v2 v2) g5 f5
// with g3 = 1/r^3, g5 = 1/r^5, f3 = 1/(1 + 6 rr^dmp3), f5 = 1/(1 + 6 rr^dmp5),
// and zero on the diagonal.  The setup kernel stores these as planes for the
// multi-workgroup SCF; the one-workgroup SCF kernels evaluate them where
// they are used (a square root, a division and two powers per pair instead
// of eighteen loads), with the same expressions, so both agree bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

// rr^e of the damping functions; GFN2's exponents 3 and 4 as products
// (within two ulp of pow, at a fraction of its cost)
__device__ __forceinline__ double aes_damp_pow(double rr, double e)
{
  if (e == 3.0) return rr * rr * rr;
  if (e == 4.0) {
    const double r2 = rr * rr;
    return r2 * r2;
  }
  return pow(rr, e);
}

struct AesPair {
  double v[3];
  double g3, g5, f3, f5;
};

__device__ __forceinline__ AesPair aes_pair(const double* __restrict__ xyz,
                                            const double* __restrict__ mrad, int i, int j,
                                            double dmp3, double dmp5)
{
  AesPair p;
  p.v[0] = p.v[1] = p.v[2] = 0.0;
  p.g3 = p.g5 = p.f3 = p.f5 = 0.0;
  if (i != j) {
    for (int k = 0; k < 3; ++k) p.v[k] = xyz[3 * i + k] - xyz[3 * j + k];
    const double r1 = sqrt(p.v[0] * p.v[0] + p.v[1] * p.v[1] + p.v[2] * p.v[2]);
    const double g1 = 1.0 / r1;
    p.g3 = g1 * g1 * g1;
    p.g5 = p.g3 * g1 * g1;
    const double rr = 0.5 * (mrad[j] + mrad[i]) * g1;
    p.f3 = 1.0 / (1.0 + 6.0 * aes_damp_pow(rr, dmp3));
    p.f5 = 1.0 / (1.0 + 6.0 * aes_damp_pow(rr, dmp5));
  }
  return p;
}

__device__ __forceinline__ double aes_sd(const AesPair& p, int k) { return p.v[k] * p.g3 * p.f3; }

__device__ __forceinline__ double aes_dd(const AesPair& p, int k, int l)
{
  return (k == l ? p.g3 * p.f5 : 0.0) - p.v[k] * p.v[l] * 3.0 * p.g5 * p.f5;
}

// component c of the six stored quadrupole components
__device__ __forceinline__ double aes_sq(const AesPair& p, int c)
{
  const double gf = p.g5 * p.f5;
  switch (c) {
    case 0: return p.v[0] * p.v[0] * gf;
    case 1: return 2.0 * p.v[0] * p.v[1] * gf;
    case 2: return p.v[1] * p.v[1] * gf;
    case 3: return 2.0 * p.v[0] * p.v[2] * gf;
    case 4: return 2.0 * p.v[1] * p.v[2] * gf;
    default: return p.v[2] * p.v[2] * gf;
  }
}

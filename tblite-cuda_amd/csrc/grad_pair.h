// Shell-pair part of the analytic gradient: the derivative of the overlap,
// dipole and quadrupole integrals of one shell pair by the separation vector
// vec = R_i - R_j, contracted on the fly with the density-dependent weights
// of the energy (the upstream multipole_grad_cgto / multipole_grad_3d,
// src/multipole.cu:422-519 and :846-981 text, used the way tblite's H0
// gradient uses them).
//
// The pair energy of a shell block, with P, W and the potentials held fixed,
//   E = sum_ab [ A_ab S_ab - P_ab (vdp_i . Di_ab + vqp_i . Qi_ab
//                                  + vdp_j . Dj_ab + vqp_j . Qj_ab) ]
// (A = 2 P H0 - 2 W - P (v_a + v_b); Di/Qi with the operator on center i,
// Dj/Qj shifted to center j), differentiates through the operator shift to
//   dE/dvec_m = sum_ab [ alpha_ab dS_m + P_ab ( beta . dDi_m + gamma . dQi_m
//                        + zeta_m S + (G Di)_m ) ]
// with per-atom-pair constants
//   alpha = A - P (vdp_j . vec + vec^T Gj vec / 2),  beta = -(vdp_i + vdp_j + Gj vec),
//   gamma = -(vqp_i + vqp_j) (through the trace removal),  zeta = -(vdp_j + Gj vec),
//   G = -Gj,  Gj_aa = 2 U_aa, Gj_ab = U_ab, U = 3/2 vqp_j - tr(vqp_j)/2 on the diagonal.
// Everything is linear in the integrals, so alpha and P are moved into the
// Cartesian basis once per shell pair and each primitive pair adds four
// scalars (three gradient components and sum P S) instead of forming the
// 10 x 4 derivative blocks.  The integral derivative follows from the
// derivative of a Cartesian Gaussian on center j:
//   dI/dvec_x = l_x I(l_x - 1) - 2 a_j I(l_x + 1).
#pragma once

#include <math.h>

#if !defined(__HIPCC__)
#ifndef __host__
#define __host__
#endif
#ifndef __device__
#define __device__
#endif
#endif

template <int L>
struct GCart {
  static constexpr int n = (L + 1) * (L + 2) / 2;
  static constexpr int s = 2 * L + 1;
};

// Cartesian exponents (x, y, z), tblite order: s; p = y, z, x;
// d = xx, yy, zz, xy, xz, yz
__host__ __device__ inline void gcart_pow(int l, int c, int& px, int& py, int& pz)
{
  if (l == 0) {
    px = py = pz = 0;
  } else if (l == 1) {
    px = (c == 2);
    py = (c == 0);
    pz = (c == 1);
  } else {
    const int tx[6] = {2, 0, 0, 1, 1, 0}, ty[6] = {0, 2, 0, 1, 0, 1}, tz[6] = {0, 0, 2, 0, 1, 1};
    px = tx[c];
    py = ty[c];
    pz = tz[c];
  }
}

// Cartesian weights of spherical weights: wc[p] = sum_a T[a][p] ws[a], T the
// Cartesian -> spherical map of tblite (d order z2, xz, yz, x2-y2, xy)
template <int L>
__host__ __device__ inline void sph_to_cart(const double* ws, double* wc)
{
  if constexpr (L < 2) {
    for (int a = 0; a < 2 * L + 1; ++a) wc[a] = ws[a];
  } else {
    const double s3 = 1.7320508075688772;
    wc[0] = -0.5 * ws[0] + 0.5 * s3 * ws[3];
    wc[1] = -0.5 * ws[0] - 0.5 * s3 * ws[3];
    wc[2] = ws[0];
    wc[3] = s3 * ws[4];
    wc[4] = s3 * ws[1];
    wc[5] = s3 * ws[2];
  }
}

// Per atom pair constants of the contraction (see the file head)
struct GradPairConst {
  double beta[3];
  double gam[6];   // trace removal already applied
  double zeta[3];
  double G[9];     // G[3 m + k]
};

// Build the constants from the potentials of the two atoms and vec.
// kappa receives vdp_j . vec + vec^T Gj vec / 2.
__host__ __device__ inline void grad_pair_const(const double* vdpi, const double* vqpi,
                                                const double* vdpj, const double* vqpj,
                                                const double vec[3], GradPairConst& k,
                                                double& kappa)
{
  const double trj = vqpj[0] + vqpj[2] + vqpj[5];
  double u[6];
  for (int c = 0; c < 6; ++c) u[c] = 1.5 * vqpj[c];
  u[0] -= 0.5 * trj;
  u[2] -= 0.5 * trj;
  u[5] -= 0.5 * trj;
  // Gj (symmetric), components xx xy yy xz yz zz -> 3 x 3
  const double gj[9] = {2.0 * u[0], u[1], u[3], u[1], 2.0 * u[2], u[4], u[3], u[4], 2.0 * u[5]};
  double gv[3];
  for (int m = 0; m < 3; ++m) gv[m] = gj[3 * m] * vec[0] + gj[3 * m + 1] * vec[1] + gj[3 * m + 2] * vec[2];
  kappa = vdpj[0] * vec[0] + vdpj[1] * vec[1] + vdpj[2] * vec[2] +
          0.5 * (vec[0] * gv[0] + vec[1] * gv[1] + vec[2] * gv[2]);
  for (int m = 0; m < 3; ++m) {
    k.beta[m] = -(vdpi[m] + vdpj[m] + gv[m]);
    k.zeta[m] = -(vdpj[m] + gv[m]);
  }
  for (int x = 0; x < 9; ++x) k.G[x] = -gj[x];
  // trace removal of the quadrupole integrals moved onto the coefficients
  double c[6];
  for (int x = 0; x < 6; ++x) c[x] = -(vqpi[x] + vqpj[x]);
  k.gam[0] = c[0] - 0.5 * c[2] - 0.5 * c[5];
  k.gam[1] = 1.5 * c[1];
  k.gam[2] = c[2] - 0.5 * c[0] - 0.5 * c[5];
  k.gam[3] = 1.5 * c[3];
  k.gam[4] = 1.5 * c[4];
  k.gam[5] = c[5] - 0.5 * c[0] - 0.5 * c[2];
}

// g[m] += dE/dvec_m of one shell pair, psh += sum_ab P_ab S_ab.
// Shell j (center j): exponents aj[npj], coefficients cj[npj] (with norms);
// shell i likewise.  alpha_c, p_c: Cartesian weights, index mj + MJ * mi.
template <int LI, int LJ>
__host__ __device__ void shell_pair_grad(const double* aj, const double* cj, int npj,
                                         const double* ai, const double* ci, int npi, double r2,
                                         const double vec[3], double intcut,
                                         const double* alpha_c, const double* p_c,
                                         const GradPairConst& k, double g[3], double& psh)
{
  constexpr int MI = GCart<LI>::n, MJ = GCart<LJ>::n;
  constexpr int LL = LI + LJ + 1;  // highest combined power (raised j)
  const double sqrtpi3 = 5.568327996831707845;  // pi^(3/2)
  const double dfac[5] = {1.0, 1.0, 3.0, 15.0, 105.0};
  const double bin[4][4] = {{1, 0, 0, 0}, {1, 1, 0, 0}, {1, 2, 1, 0}, {1, 3, 3, 1}};
  for (int ip = 0; ip < npi; ++ip) {
    const double ea = ai[ip];
    for (int jp = 0; jp < npj; ++jp) {
      const double eb = aj[jp];
      const double eab = ea + eb, oab = 1.0 / eab;
      const double est = ea * eb * r2 * oab;
      if (est > intcut) continue;
      const double cc = ci[ip] * cj[jp] * exp(-est) * sqrtpi3 * oab * sqrt(oab);
      double s1d[LL + 3];
      {
        double h = 0.5 * oab, pw = 1.0;
        for (int l = 0; l <= LL + 2; ++l) {
          if (l & 1) {
            s1d[l] = 0.0;
          } else {
            s1d[l] = pw * dfac[l / 2];
            pw *= h;
          }
        }
      }
      // v1[axis][pi][pj][moment], pj up to LJ + 1
      double v1[3][LI + 1][LJ + 2][3];
#pragma unroll
      for (int ax = 0; ax < 3; ++ax) {
        const double rpi = -vec[ax] * eb * oab;
        const double rpj = vec[ax] * ea * oab;
#pragma unroll
        for (int pi = 0; pi <= LI; ++pi)
#pragma unroll
          for (int pj = 0; pj <= LJ + 1; ++pj) {
            double vv[LL + 1];
#pragma unroll
            for (int l = 0; l <= LL; ++l) vv[l] = 0.0;
#pragma unroll
            for (int a = 0; a <= pi; ++a) {
              double pa = bin[pi][a];
#pragma unroll
              for (int e = 0; e < pi - a; ++e) pa *= rpi;
#pragma unroll
              for (int c = 0; c <= pj; ++c) {
                double pc = bin[pj][c];
#pragma unroll
                for (int e = 0; e < pj - c; ++e) pc *= rpj;
                vv[a + c] += pa * pc;
              }
            }
            double e0 = 0.0, e1 = 0.0, e2 = 0.0;
#pragma unroll
            for (int l = 0; l <= pi + pj; ++l) {
              e0 += s1d[l] * vv[l];
              e1 += (s1d[l + 1] + rpi * s1d[l]) * vv[l];
              e2 += (s1d[l + 2] + 2.0 * rpi * s1d[l + 1] + rpi * rpi * s1d[l]) * vv[l];
            }
            v1[ax][pi][pj][0] = e0;
            v1[ax][pi][pj][1] = e1;
            v1[ax][pi][pj][2] = e2;
          }
      }
#pragma unroll
      for (int mi = 0; mi < MI; ++mi) {
        int ip3[3];
        gcart_pow(LI, mi, ip3[0], ip3[1], ip3[2]);
#pragma unroll
        for (int mj = 0; mj < MJ; ++mj) {
          int jp3[3];
          gcart_pow(LJ, mj, jp3[0], jp3[1], jp3[2]);
          const double wa = cc * alpha_c[mj + MJ * mi];
          const double wp = cc * p_c[mj + MJ * mi];
          if (wa == 0.0 && wp == 0.0) continue;
          // value and derivative factors per axis and moment
          double A[3][3], dA[3][3];
#pragma unroll
          for (int ax = 0; ax < 3; ++ax) {
            const int pi = ip3[ax], pj = jp3[ax];
#pragma unroll
            for (int mo = 0; mo < 3; ++mo) {
              A[ax][mo] = v1[ax][pi][pj][mo];
              const double dn = pj > 0 ? v1[ax][pi][pj - 1][mo] : 0.0;
              dA[ax][mo] = pj * dn - 2.0 * eb * v1[ax][pi][pj + 1][mo];
            }
          }
          const double S = A[0][0] * A[1][0] * A[2][0];
          const double Di[3] = {A[0][1] * A[1][0] * A[2][0], A[0][0] * A[1][1] * A[2][0],
                                A[0][0] * A[1][0] * A[2][1]};
          psh += wp * S;
#pragma unroll
          for (int m = 0; m < 3; ++m) {
            // factors with the derivative on axis m
            const double* X = m == 0 ? dA[0] : A[0];
            const double* Y = m == 1 ? dA[1] : A[1];
            const double* Z = m == 2 ? dA[2] : A[2];
            const double dS = X[0] * Y[0] * Z[0];
            const double dD0 = X[1] * Y[0] * Z[0], dD1 = X[0] * Y[1] * Z[0], dD2 = X[0] * Y[0] * Z[1];
            const double dQ0 = X[2] * Y[0] * Z[0], dQ1 = X[1] * Y[1] * Z[0], dQ2 = X[0] * Y[2] * Z[0];
            const double dQ3 = X[1] * Y[0] * Z[1], dQ4 = X[0] * Y[1] * Z[1], dQ5 = X[0] * Y[0] * Z[2];
            const double t = k.beta[0] * dD0 + k.beta[1] * dD1 + k.beta[2] * dD2 +
                             k.gam[0] * dQ0 + k.gam[1] * dQ1 + k.gam[2] * dQ2 + k.gam[3] * dQ3 +
                             k.gam[4] * dQ4 + k.gam[5] * dQ5 + k.zeta[m] * S +
                             k.G[3 * m] * Di[0] + k.G[3 * m + 1] * Di[1] + k.G[3 * m + 2] * Di[2];
            g[m] += wa * dS + wp * t;
          }
        }
      }
    }
  }
}

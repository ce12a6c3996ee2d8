// device_math.hpp -- per-face and per-block arithmetic of the hot path for
// gfx950.  Compiled with -ffp-contract=off: every function keeps the
// reference's floating-point operation order, so the FP64 instantiations give
// bit-identical results to the reference on x86-64 (no FMA) for nonzero
// intermediate sums.  Where the reference adds a term that is an exact zero
// (the w/n_z components of its 3D embedding, jacobianRoe.cpp:46-56, and the
// zero entries of its unit derivative vectors) the term is left out; that can
// only change the sign of an exactly-zero result.
#pragma once
#include <hip/hip_runtime.h>

namespace hamstr {

constexpr double kThird = 0.3333333333333333e+00;  // utils.h:52

// --------------------------------------------------------------------------
// Roe flux, interfaceFlux.cpp:24-169.  Returns the spectral radius times the
// face length in *sr; flux already multiplied by the face length.
// --------------------------------------------------------------------------
__device__ __forceinline__ void flux_roe2d(double nxd, double nyd, const double* ql, const double* qr,
                                           double gamma, double* flux, double* sr) {
  const double gm1 = gamma - 1.0;
  const double area = sqrt(nxd * nxd + nyd * nyd);
  const double invArea = 1. / area;
  const double nx = nxd * invArea, ny = nyd * invArea;

  const double rol = ql[0];
  const double ul = ql[1] / ql[0];
  const double vl = ql[2] / ql[0];
  const double pl = gm1 * (ql[3] - 0.5 * rol * (ul * ul + vl * vl));
  const double hl = (ql[3] + pl) / rol;
  const double ror = qr[0];
  const double ur = qr[1] / qr[0];
  const double vr = qr[2] / qr[0];
  const double pr = gm1 * (qr[3] - 0.5 * ror * (ur * ur + vr * vr));
  const double hr = (qr[3] + pr) / ror;

  const double dp = pr - pl;
  const double uconr = ur * nx + vr * ny;
  const double uconl = ul * nx + vl * ny;

  const double fact = sqrt(ror / rol);
  const double A = 1.0 / (1.0 + fact);
  const double B = fact / (1.0 + fact);
  const double robar = rol * fact;
  const double ubar = ul * A + ur * B;
  const double vbar = vl * A + vr * B;
  const double hbar = hl * A + hr * B;
  double cbar = gm1 * (hbar - 0.5 * (ubar * ubar + vbar * vbar));
  cbar = sqrt(cbar);
  const double uconbar = ubar * nx + vbar * ny;

  const double eig1 = fabs(uconbar);
  const double eig2 = fabs(uconbar + cbar);
  const double eig3 = fabs(uconbar - cbar);
  const double srad = eig1 + cbar;

  const double term1 = -eig1 + 0.5 * (eig2 + eig3);
  const double term2 = 0.5 * (eig2 - eig3);
  const double dun = uconr - uconl;
  const double del1 = term1 * dp / cbar / cbar + term2 * robar * dun / cbar;
  const double del2 = term1 * dun * robar + term2 * dp / cbar;

  double fm0 = eig1 * (qr[0] - ql[0]);
  double fm1 = eig1 * (qr[1] - ql[1]);
  double fm2 = eig1 * (qr[2] - ql[2]);
  double fm3 = eig1 * (qr[3] - ql[3]);
  fm0 = fm0 + del1;
  fm1 = fm1 + del1 * ubar + del2 * nx;
  fm2 = fm2 + del1 * vbar + del2 * ny;
  fm3 = fm3 + del1 * hbar + del2 * uconbar;

  const double fl0 = rol * uconl, fr0 = ror * uconr;
  const double fl1 = rol * uconl * ul + nx * pl, fr1 = ror * uconr * ur + nx * pr;
  const double fl2 = rol * uconl * vl + ny * pl, fr2 = ror * uconr * vr + ny * pr;
  const double fl3 = (ql[3] + pl) * uconl, fr3 = (qr[3] + pr) * uconr;

  flux[0] = 0.5 * (fl0 + fr0 - fm0) * area;
  flux[1] = 0.5 * (fl1 + fr1 - fm1) * area;
  flux[2] = 0.5 * (fl2 + fr2 - fm2) * area;
  flux[3] = 0.5 * (fl3 + fr3 - fm3) * area;
  *sr = srad * area;
}

// --------------------------------------------------------------------------
// Frozen-eigenvalue Roe Jacobians (jacobianRoe.cpp:9-70 wrapper, IMODE 1 core
// :92-336, :493-620), 4x4 row-major, already multiplied by the face length.
// T = double reproduces the reference; T = float is the FP32-LHS mode.
// --------------------------------------------------------------------------
template <typename T>
struct SideLin {
  T ro, u, v, p, E, ucon;
  T du0, du1, dv0, dv2;     // du = [du0, du1, 0, 0], dv = [dv0, 0, dv2, 0]
  T dp[4];                  // (0.5 gm1 (u^2+v^2), -gm1 u, -gm1 v, gm1)
  T dc[3];                  // ducon = [dc0, dc1, dc2, 0]
};

template <typename T>
__device__ __forceinline__ SideLin<T> side_lin(const T* q, T nx, T ny, T gm1) {
  SideLin<T> s;
  s.ro = q[0];
  s.u = q[1] / q[0];
  s.v = q[2] / q[0];
  s.E = q[3];
  s.p = gm1 * (q[3] - T(0.5) * s.ro * (s.u * s.u + s.v * s.v));
  s.ucon = s.u * nx + s.v * ny;
  s.du0 = -s.u / s.ro;
  s.du1 = T(1.0) / s.ro;
  s.dv0 = -s.v / s.ro;
  s.dv2 = T(1.0) / s.ro;
  s.dp[0] = T(0.5) * gm1 * (s.u * s.u + s.v * s.v);
  s.dp[1] = -gm1 * s.u;
  s.dp[2] = -gm1 * s.v;
  s.dp[3] = gm1;
  s.dc[0] = -s.ucon / s.ro;
  s.dc[1] = nx / s.ro;
  s.dc[2] = ny / s.ro;
  return s;
}

// native flux Jacobian N (jacobianRoe.cpp:573-605) minus D, times 0.5*area
template <typename T, bool RIGHT>
__device__ __forceinline__ void jac_side(const SideLin<T>& s, T nx, T ny, T gm1, T eig1, T term1, T term2,
                                         T cbar, T robar, T ubar, T vbar, T hbar, T uconbar, T area, T* M) {
  // ddel1, ddel2 (jacobianRoe.cpp:328-335); the left side carries the minus signs
  T dd1[4], dd2[4];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    if (RIGHT) {
      dd1[j] = term1 * s.dp[j] / cbar / cbar + term2 * robar * s.dc[j] / cbar;
      dd2[j] = term1 * s.dc[j] * robar + term2 * s.dp[j] / cbar;
    } else {
      dd1[j] = -term1 * s.dp[j] / cbar / cbar - term2 * robar * s.dc[j] / cbar;
      dd2[j] = -term1 * s.dc[j] * robar - term2 * s.dp[j] / cbar;
    }
  }
  if (RIGHT) {
    dd1[3] = term1 * s.dp[3] / cbar / cbar;
    dd2[3] = term2 * s.dp[3] / cbar;
  } else {
    dd1[3] = -term1 * s.dp[3] / cbar / cbar;
    dd2[3] = -(term2 * s.dp[3] / cbar);
  }
  const T e = RIGHT ? eig1 : -eig1;
  // D rows 0,1,2,4 of the 5x5 (jacobianRoe.cpp:506-531)
  T D[16];
  D[0] = e + dd1[0];
  D[1] = dd1[1];
  D[2] = dd1[2];
  D[3] = dd1[3];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    D[4 + j] = (j == 1) ? ((e + dd1[j] * ubar) + dd2[j] * nx) : (dd1[j] * ubar + dd2[j] * nx);
    D[8 + j] = (j == 2) ? ((e + dd1[j] * vbar) + dd2[j] * ny) : (dd1[j] * vbar + dd2[j] * ny);
    D[12 + j] = (j == 3) ? ((e + dd1[j] * hbar) + dd2[j] * uconbar) : (dd1[j] * hbar + dd2[j] * uconbar);
  }
  // native Jacobian
  const T rodc0 = s.ro * s.dc[0], rodc1 = s.ro * s.dc[1], rodc2 = s.ro * s.dc[2];
  const T rouc = s.ro * s.ucon;
  const T Ep = s.E + s.p;
  T N[16];
  N[0] = s.ucon + rodc0;
  N[1] = rodc1;
  N[2] = rodc2;
  N[3] = T(0.0);
  N[4] = (((s.ucon * s.u) + rodc0 * s.u) + rouc * s.du0) + nx * s.dp[0];
  N[5] = (rodc1 * s.u + rouc * s.du1) + nx * s.dp[1];
  N[6] = (rodc2 * s.u) + nx * s.dp[2];
  N[7] = nx * s.dp[3];
  N[8] = (((s.ucon * s.v) + rodc0 * s.v) + rouc * s.dv0) + ny * s.dp[0];
  N[9] = (rodc1 * s.v) + ny * s.dp[1];
  N[10] = (rodc2 * s.v + rouc * s.dv2) + ny * s.dp[2];
  N[11] = ny * s.dp[3];
  N[12] = s.dp[0] * s.ucon + Ep * s.dc[0];
  N[13] = s.dp[1] * s.ucon + Ep * s.dc[1];
  N[14] = s.dp[2] * s.ucon + Ep * s.dc[2];
  N[15] = (T(1.0) + gm1) * s.ucon;
#pragma unroll
  for (int k = 0; k < 16; k++) M[k] = T(0.5) * (N[k] - D[k]) * area;
}

template <typename T>
__device__ __forceinline__ void jacobian_roe2d(double nxd_, double nyd_, const double* ql_, const double* qr_,
                                               double gamma, T* lmat, T* rmat) {
  const T gam = T(gamma);
  const T gm1 = gam - T(1.0);
  const T nxd = T(nxd_), nyd = T(nyd_);
  const T area = sqrt(nxd * nxd + nyd * nyd);
  const T nx = nxd / area, ny = nyd / area;
  T ql[4] = {T(ql_[0]), T(ql_[1]), T(ql_[2]), T(ql_[3])};
  T qr[4] = {T(qr_[0]), T(qr_[1]), T(qr_[2]), T(qr_[3])};
  const SideLin<T> L = side_lin<T>(ql, nx, ny, gm1);
  const SideLin<T> R = side_lin<T>(qr, nx, ny, gm1);
  const T hl = (L.E + L.p) / L.ro;
  const T hr = (R.E + R.p) / R.ro;
  const T fact = sqrt(R.ro / L.ro);
  const T A = T(1.0) / (T(1.0) + fact);
  const T B = fact / (T(1.0) + fact);
  const T robar = L.ro * fact;
  const T ubar = L.u * A + R.u * B;
  const T vbar = L.v * A + R.v * B;
  const T hbar = hl * A + hr * B;
  T cbar = gm1 * (hbar - T(0.5) * (ubar * ubar + vbar * vbar));
  cbar = sqrt(cbar);
  const T uconbar = ubar * nx + vbar * ny;
  const T eig1 = fabs(uconbar);
  const T eig2 = fabs(uconbar + cbar);
  const T eig3 = fabs(uconbar - cbar);
  const T term1 = -eig1 + T(0.5) * (eig2 + eig3);
  const T term2 = T(0.5) * (eig2 - eig3);
  jac_side<T, false>(L, nx, ny, gm1, eig1, term1, term2, cbar, robar, ubar, vbar, hbar, uconbar, area, lmat);
  jac_side<T, true>(R, nx, ny, gm1, eig1, term1, term2, cbar, robar, ubar, vbar, hbar, uconbar, area, rmat);
}

// --------------------------------------------------------------------------
// Reconstruction, pointwise (reconstruction.cpp).
// MUSCL at row i from f[i-1], f[i], f[i+1] (reconstruction.cpp:53-69).
// --------------------------------------------------------------------------
__device__ __forceinline__ void muscl_pt(double fm, double f0, double fp, double qt, double eps, double* ql,
                                         double* qr) {
  const double thm = 1.0 - kThird, thp = 1.0 + kThird;
  const double f2i = fp - f0;
  const double f2i1 = f0 - fm;
  const double a1 = 3.0 * (f2i * f2i1 + eps);
  const double a2 = 2.0 * (f2i - f2i1) * (f2i - f2i1) + a1;
  const double f3i = a1 / a2;
  const double f3qt = qt * f3i;
  *ql = f0 + f3qt * (thm * f2i1 + thp * f2i);
  *qr = f0 - f3qt * (thp * f2i1 + thm * f2i);
}

// reconstruction.cpp:216-234
__device__ __forceinline__ void muscl_deriv_pt(double fm, double f0, double fp, double dm, double d0, double dp,
                                               double qt, double eps, double* ql, double* qr) {
  const double thm = 1.0 - kThird, thp = 1.0 + kThird;
  const double f2i = fp - f0;
  const double f2i1 = f0 - fm;
  const double df2i = dp - d0;
  const double df2i1 = d0 - dm;
  const double a1 = 3.0 * (f2i * f2i1 + eps);
  const double da1 = 3.0 * (df2i * f2i1 + f2i * df2i1);
  const double a2 = 2.0 * (f2i - f2i1) * (f2i - f2i1) + a1;
  const double da2 = 2.0 * 2.0 * (f2i - f2i1) * (df2i - df2i1) + da1;
  const double f3i = a1 / a2;
  const double df3i = da1 / a2 - a1 * da2 / (a2 * a2);
  const double f3qt = qt * f3i;
  const double df3qt = qt * df3i;
  *ql = d0 + df3qt * (thm * f2i1 + thp * f2i) + f3qt * (thm * df2i1 + thp * df2i);
  *qr = d0 - df3qt * (thp * f2i1 + thm * f2i) - f3qt * (thp * df2i1 + thm * df2i);
}

// reconstruction.cpp:78-124 (p = 1)
__device__ __forceinline__ double weno5(double a, double b, double c, double d, double e, double epsw) {
  const double b1 = 13. / 12.;
  const double b2 = 1. / 6.;
  const double djm1 = a - 2. * b + c;
  const double ejm1 = a - 4. * b + 3. * c;
  const double dj = b - 2. * c + d;
  const double ej = b - d;
  const double djp1 = c - 2. * d + e;
  const double ejp1 = 3. * c - 4. * d + e;
  const double dis0 = b1 * djm1 * djm1 + 0.25 * ejm1 * ejm1 + epsw;
  const double dis1 = b1 * dj * dj + 0.25 * ej * ej + epsw;
  const double dis2 = b1 * djp1 * djp1 + 0.25 * ejp1 * ejp1 + epsw;
  const double q30 = 2. * a - 7. * b + 11. * c;
  const double q31 = -b + 5. * c + 2. * d;
  const double q32 = 2. * c + 5. * d - e;
  const double d01 = dis0 / dis1;
  const double d02 = dis0 / dis2;
  const double a1ba0 = 6. * d01;
  const double a2ba0 = 3. * d02;
  const double w0 = 1.0 / (1.0 + a1ba0 + a2ba0);
  const double w1 = a1ba0 * w0;
  const double w2 = 1. - w0 - w1;
  return b2 * (w0 * q30 + w1 * q31 + w2 * q32);
}

// reconstruction.cpp:243-347
__device__ __forceinline__ double weno5_deriv(double a, double b, double c, double d, double e, double a_d,
                                              double b_d, double c_d, double d_d, double e_d, double epsw) {
  const double b1 = 13. / 12.;
  const double b2 = 1. / 6.;
  const double djm1 = a - 2. * b + c;
  const double ejm1 = a - 4. * b + 3. * c;
  const double dj = b - 2. * c + d;
  const double ej = b - d;
  const double djp1 = c - 2. * d + e;
  const double ejp1 = 3. * c - 4. * d + e;
  const double dis0 = b1 * djm1 * djm1 + 0.25 * ejm1 * ejm1 + epsw;
  const double dis1 = b1 * dj * dj + 0.25 * ej * ej + epsw;
  const double dis2 = b1 * djp1 * djp1 + 0.25 * ejp1 * ejp1 + epsw;
  const double q30 = 2. * a - 7. * b + 11. * c;
  const double q31 = -b + 5. * c + 2. * d;
  const double q32 = 2. * c + 5. * d - e;
  const double d01 = dis0 / dis1;
  const double d02 = dis0 / dis2;
  const double a1ba0 = 6. * d01;
  const double a2ba0 = 3. * d02;
  const double w0 = 1.0 / (1.0 + a1ba0 + a2ba0);
  const double w1 = a1ba0 * w0;
  const double w2 = 1. - w0 - w1;

  const double djm1_d = a_d - 2. * b_d + c_d;
  const double ejm1_d = a_d - 4. * b_d + 3. * c_d;
  const double dj_d = b_d - 2. * c_d + d_d;
  const double ej_d = b_d - d_d;
  const double djp1_d = c_d - 2. * d_d + e_d;
  const double ejp1_d = 3. * c_d - 4. * d_d + e_d;
  const double q30_d = 2. * a_d - 7. * b_d + 11. * c_d;
  const double q31_d = -b_d + 5. * c_d + 2. * d_d;
  const double q32_d = 2. * c_d + 5. * d_d - e_d;

  const double term1 = b2 * w0 * q30_d;
  const double term2 = b2 * w1 * q31_d;
  const double term3 = b2 * w2 * q32_d;

  const double dbeta0 = b1 * 2 * djm1 * djm1_d + 0.25 * 2 * ejm1 * ejm1_d;
  const double dbeta1 = b1 * 2 * dj * dj_d + 0.25 * 2 * ej * ej_d;
  const double dbeta2 = b1 * 2 * djp1 * djp1_d + 0.25 * 2 * ejp1 * ejp1_d;
  const double beta0 = dis0, beta1 = dis1, beta2 = dis2;

  const double dw0_down = beta1 * beta2 + 6 * beta0 * beta2 + 3 * beta0 * beta1;
  const double dw0_up_1 = (dbeta1 * beta2 + beta1 * dbeta2) * dw0_down;
  const double dw0_up_2 = beta1 * beta2 *
                          (dbeta1 * beta2 + beta1 * dbeta2 + 6 * (dbeta0 * beta2 + beta0 * dbeta2) +
                           3 * (dbeta0 * beta1 + beta0 * dbeta1));
  const double dw0 = (dw0_up_1 - dw0_up_2) / (dw0_down * dw0_down);
  const double dw1_up_1 = 6 * (dbeta0 * beta2 + beta0 * dbeta2) * dw0_down;
  const double dw1_up_2 = 6 * dw0_up_2 / beta1 * beta0;
  const double dw1 = (dw1_up_1 - dw1_up_2) / (dw0_down * dw0_down);
  const double dw2_up_1 = 3 * (dbeta0 * beta1 + beta0 * dbeta1) * dw0_down;
  const double dw2_up_2 = 3 * dw0_up_2 / beta2 * beta0;
  const double dw2 = (dw2_up_1 - dw2_up_2) / (dw0_down * dw0_down);
  return term1 + term2 + term3 + b2 * dw0 * q30 + b2 * dw1 * q31 + b2 * dw2 * q32;
}

// --------------------------------------------------------------------------
// Unpivoted 4x4 LU with reciprocal pivots and its solve
// (triSolvers.cpp:805-831, repeated in blockTridagPeriodic4).
// --------------------------------------------------------------------------
template <typename T>
struct LU4 {
  T b11, b21, b22, b31, b32, b33, b41, b42, b43, b44;
  T u12, u13, u14, u23, u24, u34;
};

template <typename T>
__device__ __forceinline__ LU4<T> lu4(const T* b) {
  LU4<T> L;
  L.b11 = T(1.) / b[0];
  L.u12 = b[1] * L.b11;
  L.u13 = b[2] * L.b11;
  L.u14 = b[3] * L.b11;
  L.b21 = b[4];
  L.b22 = T(1.) / (b[5] - L.b21 * L.u12);
  L.u23 = (b[6] - L.b21 * L.u13) * L.b22;
  L.u24 = (b[7] - L.b21 * L.u14) * L.b22;
  L.b31 = b[8];
  L.b32 = b[9] - L.b31 * L.u12;
  L.b33 = T(1.) / (b[10] - L.b31 * L.u13 - L.b32 * L.u23);
  L.u34 = (b[11] - L.b31 * L.u14 - L.b32 * L.u24) * L.b33;
  L.b41 = b[12];
  L.b42 = b[13] - L.b41 * L.u12;
  L.b43 = b[14] - L.b41 * L.u13 - L.b42 * L.u23;
  L.b44 = T(1.) / (b[15] - L.b41 * L.u14 - L.b42 * L.u24 - L.b43 * L.u34);
  return L;
}

// v[0], v[s], v[2s], v[3s] := LU^{-1} v
template <typename T>
__device__ __forceinline__ void lu4_solve(const LU4<T>& L, T* v, int s) {
  const T d1 = v[0] * L.b11;
  const T d2 = (v[s] - L.b21 * d1) * L.b22;
  const T d3 = (v[2 * s] - L.b31 * d1 - L.b32 * d2) * L.b33;
  const T d4 = (v[3 * s] - L.b41 * d1 - L.b42 * d2 - L.b43 * d3) * L.b44;
  v[3 * s] = d4;
  v[2 * s] = d3 - L.u34 * v[3 * s];
  v[s] = d2 - L.u23 * v[2 * s] - L.u24 * v[3 * s];
  v[0] = d1 - L.u12 * v[s] - L.u13 * v[2 * s] - L.u14 * v[3 * s];
}

// X[k][j] -= sum_l A[k][l] Y[l][j], l sequential (triSolvers.cpp:790-797)
template <typename T>
__device__ __forceinline__ void mat_sub_mul(T* X, const T* A, const T* Y) {
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      T x = X[4 * k + j];
#pragma unroll
      for (int l = 0; l < 4; l++) x -= A[4 * k + l] * Y[4 * l + j];
      X[4 * k + j] = x;
    }
}

// x[k] -= sum_j A[k][j] y[j], j sequential (triSolvers.cpp:799-803)
template <typename T>
__device__ __forceinline__ void vec_sub_mul(T* x, const T* A, const T* y) {
#pragma unroll
  for (int k = 0; k < 4; k++) {
    T v = x[k];
#pragma unroll
    for (int j = 0; j < 4; j++) v -= A[4 * k + j] * y[j];
    x[k] = v;
  }
}

// --------------------------------------------------------------------------
// The same 4x4 LU for the FAST (reordered) solves: contracted FMAs and the
// hardware reciprocal refined by Newton steps (two in FP64, to full precision;
// the FP32 reciprocal is used as is).
// --------------------------------------------------------------------------
__device__ __forceinline__ double fast_rcp(double x) {
#pragma clang fp contract(fast)
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

template <typename T>
__device__ __forceinline__ LU4<T> lu4_fast(const T* b) {
#pragma clang fp contract(fast)
  LU4<T> L;
  L.b11 = fast_rcp(b[0]);
  L.u12 = b[1] * L.b11;
  L.u13 = b[2] * L.b11;
  L.u14 = b[3] * L.b11;
  L.b21 = b[4];
  L.b22 = fast_rcp(b[5] - L.b21 * L.u12);
  L.u23 = (b[6] - L.b21 * L.u13) * L.b22;
  L.u24 = (b[7] - L.b21 * L.u14) * L.b22;
  L.b31 = b[8];
  L.b32 = b[9] - L.b31 * L.u12;
  L.b33 = fast_rcp(b[10] - L.b31 * L.u13 - L.b32 * L.u23);
  L.u34 = (b[11] - L.b31 * L.u14 - L.b32 * L.u24) * L.b33;
  L.b41 = b[12];
  L.b42 = b[13] - L.b41 * L.u12;
  L.b43 = b[14] - L.b41 * L.u13 - L.b42 * L.u23;
  L.b44 = fast_rcp(b[15] - L.b41 * L.u14 - L.b42 * L.u24 - L.b43 * L.u34);
  return L;
}

template <typename T>
__device__ __forceinline__ void lu4_solve_fast(const LU4<T>& L, T* v, int s) {
#pragma clang fp contract(fast)
  const T d1 = v[0] * L.b11;
  const T d2 = (v[s] - L.b21 * d1) * L.b22;
  const T d3 = (v[2 * s] - L.b31 * d1 - L.b32 * d2) * L.b33;
  const T d4 = (v[3 * s] - L.b41 * d1 - L.b42 * d2 - L.b43 * d3) * L.b44;
  v[3 * s] = d4;
  v[2 * s] = d3 - L.u34 * d4;
  v[s] = d2 - L.u23 * v[2 * s] - L.u24 * d4;
  v[0] = d1 - L.u12 * v[s] - L.u13 * v[2 * s] - L.u14 * d4;
}

// --------------------------------------------------------------------------
// FAST-mode Jacobians: the same frozen-eigenvalue Roe Jacobians as
// jacobian_roe2d, with every division by rho, area and cbar replaced by a
// multiplication with one reciprocal each (hardware reciprocal refined by
// Newton steps).  Equal to jacobian_roe2d up to rounding; only the ADI's
// left-hand side uses them, never the residual.
// --------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ SideLin<T> side_lin_fast(const T* q, T nx, T ny, T gm1) {
#pragma clang fp contract(fast)
  SideLin<T> s;
  const T ir = fast_rcp(q[0]);
  s.ro = q[0];
  s.u = q[1] * ir;
  s.v = q[2] * ir;
  s.E = q[3];
  s.p = gm1 * (q[3] - T(0.5) * s.ro * (s.u * s.u + s.v * s.v));
  s.ucon = s.u * nx + s.v * ny;
  s.du0 = -s.u * ir;
  s.du1 = ir;
  s.dv0 = -s.v * ir;
  s.dv2 = ir;
  s.dp[0] = T(0.5) * gm1 * (s.u * s.u + s.v * s.v);
  s.dp[1] = -gm1 * s.u;
  s.dp[2] = -gm1 * s.v;
  s.dp[3] = gm1;
  s.dc[0] = -s.ucon * ir;
  s.dc[1] = nx * ir;
  s.dc[2] = ny * ir;
  return s;
}

template <typename T, bool RIGHT>
__device__ __forceinline__ void jac_side_fast(const SideLin<T>& s, T nx, T ny, T gm1, T eig1, T term1, T term2,
                                              T ic, T ic2, T robar, T ubar, T vbar, T hbar, T uconbar, T harea,
                                              T* M) {
#pragma clang fp contract(fast)
  const T sg = RIGHT ? T(1) : T(-1);
  const T t1c2 = sg * term1 * ic2, t2c = sg * term2 * ic, t2rc = sg * term2 * robar * ic, t1r = sg * term1 * robar;
  T dd1[4], dd2[4];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    dd1[j] = t1c2 * s.dp[j] + t2rc * s.dc[j];
    dd2[j] = t1r * s.dc[j] + t2c * s.dp[j];
  }
  dd1[3] = t1c2 * s.dp[3];
  dd2[3] = t2c * s.dp[3];
  const T e = sg * eig1;
  T D[16];
  D[0] = e + dd1[0];
  D[1] = dd1[1];
  D[2] = dd1[2];
  D[3] = dd1[3];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    D[4 + j] = dd1[j] * ubar + dd2[j] * nx + ((j == 1) ? e : T(0));
    D[8 + j] = dd1[j] * vbar + dd2[j] * ny + ((j == 2) ? e : T(0));
    D[12 + j] = dd1[j] * hbar + dd2[j] * uconbar + ((j == 3) ? e : T(0));
  }
  const T rodc0 = s.ro * s.dc[0], rodc1 = s.ro * s.dc[1], rodc2 = s.ro * s.dc[2];
  const T rouc = s.ro * s.ucon;
  const T Ep = s.E + s.p;
  T N[16];
  N[0] = s.ucon + rodc0;
  N[1] = rodc1;
  N[2] = rodc2;
  N[3] = T(0.0);
  N[4] = s.ucon * s.u + rodc0 * s.u + rouc * s.du0 + nx * s.dp[0];
  N[5] = rodc1 * s.u + rouc * s.du1 + nx * s.dp[1];
  N[6] = rodc2 * s.u + nx * s.dp[2];
  N[7] = nx * s.dp[3];
  N[8] = s.ucon * s.v + rodc0 * s.v + rouc * s.dv0 + ny * s.dp[0];
  N[9] = rodc1 * s.v + ny * s.dp[1];
  N[10] = rodc2 * s.v + rouc * s.dv2 + ny * s.dp[2];
  N[11] = ny * s.dp[3];
  N[12] = s.dp[0] * s.ucon + Ep * s.dc[0];
  N[13] = s.dp[1] * s.ucon + Ep * s.dc[1];
  N[14] = s.dp[2] * s.ucon + Ep * s.dc[2];
  N[15] = (T(1.0) + gm1) * s.ucon;
#pragma unroll
  for (int k = 0; k < 16; k++) M[k] = (N[k] - D[k]) * harea;
}

template <typename T>
__device__ __forceinline__ void jacobian_roe2d_fast(double nxd_, double nyd_, const double* ql_, const double* qr_,
                                                    double gamma, T* lmat, T* rmat) {
#pragma clang fp contract(fast)
  const T gam = T(gamma);
  const T gm1 = gam - T(1.0);
  const T nxd = T(nxd_), nyd = T(nyd_);
  const T area = sqrt(nxd * nxd + nyd * nyd);
  const T ia = fast_rcp(area);
  const T nx = nxd * ia, ny = nyd * ia;
  T ql[4] = {T(ql_[0]), T(ql_[1]), T(ql_[2]), T(ql_[3])};
  T qr[4] = {T(qr_[0]), T(qr_[1]), T(qr_[2]), T(qr_[3])};
  const SideLin<T> L = side_lin_fast<T>(ql, nx, ny, gm1);
  const SideLin<T> R = side_lin_fast<T>(qr, nx, ny, gm1);
  const T hl = (L.E + L.p) * L.du1;  // du1 = 1/rho
  const T hr = (R.E + R.p) * R.du1;
  const T fact = sqrt(R.ro * L.du1);
  const T A = fast_rcp(T(1.0) + fact);
  const T B = fact * A;
  const T robar = L.ro * fact;
  const T ubar = L.u * A + R.u * B;
  const T vbar = L.v * A + R.v * B;
  const T hbar = hl * A + hr * B;
  const T cbar = sqrt(gm1 * (hbar - T(0.5) * (ubar * ubar + vbar * vbar)));
  const T ic = fast_rcp(cbar);
  const T uconbar = ubar * nx + vbar * ny;
  const T eig1 = fabs(uconbar);
  const T eig2 = fabs(uconbar + cbar);
  const T eig3 = fabs(uconbar - cbar);
  const T term1 = -eig1 + T(0.5) * (eig2 + eig3);
  const T term2 = T(0.5) * (eig2 - eig3);
  const T harea = T(0.5) * area;
  jac_side_fast<T, false>(L, nx, ny, gm1, eig1, term1, term2, ic, ic * ic, robar, ubar, vbar, hbar, uconbar, harea,
                          lmat);
  jac_side_fast<T, true>(R, nx, ny, gm1, eig1, term1, term2, ic, ic * ic, robar, ubar, vbar, hbar, uconbar, harea,
                         rmat);
}

// FAST-mode MUSCL derivative (the linear residual inside the ADI): one
// reciprocal of a2 instead of three divisions; equal to muscl_deriv_pt up to
// rounding
__device__ __forceinline__ void muscl_deriv_pt_fast(double fm, double f0, double fp, double dm, double d0,
                                                    double dp, double qt, double eps, double* ql, double* qr) {
#pragma clang fp contract(fast)
  const double thm = 1.0 - kThird, thp = 1.0 + kThird;
  const double f2i = fp - f0;
  const double f2i1 = f0 - fm;
  const double df2i = dp - d0;
  const double df2i1 = d0 - dm;
  const double a1 = 3.0 * (f2i * f2i1 + eps);
  const double da1 = 3.0 * (df2i * f2i1 + f2i * df2i1);
  const double a2 = 2.0 * (f2i - f2i1) * (f2i - f2i1) + a1;
  const double da2 = 4.0 * (f2i - f2i1) * (df2i - df2i1) + da1;
  const double ia2 = fast_rcp(a2);
  const double f3i = a1 * ia2;
  const double df3i = (da1 - f3i * da2) * ia2;
  const double f3qt = qt * f3i;
  const double df3qt = qt * df3i;
  *ql = d0 + df3qt * (thm * f2i1 + thp * f2i) + f3qt * (thm * df2i1 + thp * df2i);
  *qr = d0 - df3qt * (thp * f2i1 + thm * f2i) - f3qt * (thp * df2i1 + thm * df2i);
}

}  // namespace hamstr

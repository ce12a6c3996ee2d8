// strand_math.hpp -- device functions of the 3D / strand building blocks (5
// conserved variables rho, rho u, rho v, rho w, E): the Roe flux, its
// frozen-Roe Jacobians and the 5x5 block-tridiagonal solves.  The reference's
// 3D build does not compile (io.cpp:306 uses an undeclared BC_WALL_INVIS), but
// these three functions are compiled in its 2D build; they are reproduced bit
// for bit: -ffp-contract=off, IEEE division and square root, and every
// expression evaluated in the reference's order (the reference is x86-64
// SSE2 code without FMA).
#pragma once

#include <hip/hip_runtime.h>

namespace hamstr {

// Primitive state of one side of a 3D face and the derivatives of its
// velocity, pressure and normal velocity with respect to the conserved state
// (jacobianRoe.cpp:196-272 for the left side, the same for the right).
struct Side5 {
  double ro, u, v, w, p, h, ucon;
  double dvel[3][5];  // d(u, v, w) / dq
  double dp[5];       // d p / dq
  double ducon[5];    // d (u.n) / dq
};

__device__ __forceinline__ void prim5(const double* q, double gm1, double nx, double ny, double nz, double& ro,
                                      double& u, double& v, double& w, double& p, double& h, double& ucon) {
  ro = q[0];
  u = q[1] / q[0];
  v = q[2] / q[0];
  w = q[3] / q[0];
  p = gm1 * (q[4] - 0.5 * ro * (u * u + v * v + w * w));
  h = (q[4] + p) / ro;
  ucon = u * nx + v * ny + w * nz;
}

__device__ __forceinline__ Side5 side5(const double* q, double gm1, double nx, double ny, double nz) {
  Side5 s;
  prim5(q, gm1, nx, ny, nz, s.ro, s.u, s.v, s.w, s.p, s.h, s.ucon);
  const double vel[3] = {s.u, s.v, s.w};
#pragma unroll
  for (int d = 0; d < 3; d++) {
#pragma unroll
    for (int k = 0; k < 5; k++) s.dvel[d][k] = 0.0;
    s.dvel[d][0] = -vel[d] / s.ro;
    s.dvel[d][d + 1] = 1.0 / s.ro;
  }
  s.dp[0] = 0.5 * gm1 * (s.u * s.u + s.v * s.v + s.w * s.w);
  s.dp[1] = -gm1 * s.u;
  s.dp[2] = -gm1 * s.v;
  s.dp[3] = -gm1 * s.w;
  s.dp[4] = gm1;
  s.ducon[0] = -s.ucon / s.ro;
  s.ducon[1] = nx / s.ro;
  s.ducon[2] = ny / s.ro;
  s.ducon[3] = nz / s.ro;
  s.ducon[4] = 0.0;
  return s;
}

// fluxRoe3D (interfaceFlux.cpp:176-319): the flux through the unit normal and
// the spectral radius, neither scaled by the face area.
__device__ __forceinline__ void flux_roe3d(const double* nxyz, const double* ql, const double* qr, double gamma,
                                           double* flux, double* specRadius) {
  const double gm1 = gamma - 1.0;
  const double area = sqrt(nxyz[0] * nxyz[0] + nxyz[1] * nxyz[1] + nxyz[2] * nxyz[2]);
  const double invArea = 1. / area;
  const double nx = nxyz[0] * invArea, ny = nxyz[1] * invArea, nz = nxyz[2] * invArea;
  double rol, ul, vl, wl, pl, hl, uconl, ror, ur, vr, wr, pr, hr, uconr;
  prim5(ql, gm1, nx, ny, nz, rol, ul, vl, wl, pl, hl, uconl);
  prim5(qr, gm1, nx, ny, nz, ror, ur, vr, wr, pr, hr, uconr);
  const double dp = pr - pl;
  const double fact = sqrt(ror / rol);
  const double A = 1.0 / (1.0 + fact);
  const double B = fact / (1.0 + fact);
  const double robar = rol * fact;
  const double ubar = ul * A + ur * B;
  const double vbar = vl * A + vr * B;
  const double wbar = wl * A + wr * B;
  const double hbar = hl * A + hr * B;
  const double cbar = sqrt(gm1 * (hbar - 0.5 * (ubar * ubar + vbar * vbar + wbar * wbar)));
  const double uconbar = ubar * nx + vbar * ny + wbar * nz;
  const double eig1 = fabs(uconbar);
  const double eig2 = fabs(uconbar + cbar);
  const double eig3 = fabs(uconbar - cbar);
  *specRadius = eig1 + cbar;
  const double term1 = -eig1 + 0.5 * (eig2 + eig3);
  const double term2 = 0.5 * (eig2 - eig3);
  const double dun = uconr - uconl;
  const double del1 = term1 * dp / cbar / cbar + term2 * robar * dun / cbar;
  const double del2 = term1 * dun * robar + term2 * dp / cbar;
  // dissipation |A| dq, per component: eig1 dq + del1 (1, ubar, vbar, wbar,
  // hbar) + del2 (0, n, uconbar)
  const double avg[5] = {1.0, ubar, vbar, wbar, hbar};
  const double nrm[5] = {0.0, nx, ny, nz, uconbar};
  const double fl[5] = {rol * uconl, rol * uconl * ul + nx * pl, rol * uconl * vl + ny * pl,
                        rol * uconl * wl + nz * pl, (ql[4] + pl) * uconl};
  const double fr[5] = {ror * uconr, ror * uconr * ur + nx * pr, ror * uconr * vr + ny * pr,
                        ror * uconr * wr + nz * pr, (qr[4] + pr) * uconr};
#pragma unroll
  for (int i = 0; i < 5; i++) {
    double fm = eig1 * (qr[i] - ql[i]);
    fm = (i == 0) ? fm + del1 : fm + del1 * avg[i] + del2 * nrm[i];
    flux[i] = 0.5 * (fl[i] + fr[i] - fm);
  }
}

// jacobianRoe3D with IMODE 1 (jacobianRoe.cpp:92-336 and 493-620): the
// frozen-Roe-average derivatives of fluxRoe3D, lmat = dF/dqL, rmat = dF/dqR,
// row-major 5x5, not scaled by the face area.
__device__ __forceinline__ void jacobian_roe3d(const double* nxyz, const double* ql, const double* qr, double gam,
                                               double* lmat, double* rmat) {
  const double gm1 = gam - 1.0;
  const double area = sqrt(nxyz[0] * nxyz[0] + nxyz[1] * nxyz[1] + nxyz[2] * nxyz[2]);
  const double nx = nxyz[0] / area, ny = nxyz[1] / area, nz = nxyz[2] / area;
  const Side5 L = side5(ql, gm1, nx, ny, nz);
  const Side5 R = side5(qr, gm1, nx, ny, nz);
  const double fact = sqrt(R.ro / L.ro);
  const double A = 1.0 / (1.0 + fact);
  const double B = fact / (1.0 + fact);
  const double robar = L.ro * fact;
  const double ubar = L.u * A + R.u * B;
  const double vbar = L.v * A + R.v * B;
  const double wbar = L.w * A + R.w * B;
  const double hbar = L.h * A + R.h * B;
  const double cbar = sqrt(gm1 * (hbar - 0.5 * (ubar * ubar + vbar * vbar + wbar * wbar)));
  const double uconbar = ubar * nx + vbar * ny + wbar * nz;
  const double eig1 = fabs(uconbar);
  const double eig2 = fabs(uconbar + cbar);
  const double eig3 = fabs(uconbar - cbar);
  const double term1 = -eig1 + 0.5 * (eig2 + eig3);
  const double term2 = 0.5 * (eig2 - eig3);
  const double avg[5] = {1.0, ubar, vbar, wbar, hbar};
  const double nrm[5] = {0.0, nx, ny, nz, uconbar};
#pragma unroll
  for (int side = 0; side < 2; side++) {
    const Side5& S = side ? R : L;
    const double* q = side ? qr : ql;
    double* out = side ? rmat : lmat;
    // derivatives of del1 / del2: + for the right state, - for the left
    // (jacobianRoe.cpp:327-333), with the sign on term1 as written there
    const double t1 = side ? term1 : -term1;
    double dd1[5], dd2[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
      if (side) {
        dd1[i] = t1 * S.dp[i] / cbar / cbar + term2 * robar * S.ducon[i] / cbar;
        dd2[i] = t1 * S.ducon[i] * robar + term2 * S.dp[i] / cbar;
      } else {
        dd1[i] = t1 * S.dp[i] / cbar / cbar - term2 * robar * S.ducon[i] / cbar;
        dd2[i] = t1 * S.ducon[i] * robar - term2 * S.dp[i] / cbar;
      }
    }
    const double vel[3] = {S.u, S.v, S.w};
    const double nv[3] = {nx, ny, nz};
#pragma unroll
    for (int i = 0; i < 5; i++) {        // column: d/dq_i
      const double dro = (i == 0) ? 1.0 : 0.0;
      const double dq5 = (i == 4) ? 1.0 : 0.0;
#pragma unroll
      for (int r = 0; r < 5; r++) {      // row: flux component
        // dissipation part: (0 -/+ eig1 I) + dd1 avg_r (+ dd2 nrm_r)
        double dis = (r == i) ? (side ? 0.0 + eig1 : 0.0 - eig1) : 0.0;
        dis = (r == 0) ? dis + dd1[i] : dis + dd1[i] * avg[r] + dd2[i] * nrm[r];
        // native flux Jacobian
        double nat;
        if (r == 0) {
          nat = dro * S.ucon + S.ro * S.ducon[i];
        } else if (r < 4) {
          nat = dro * S.ucon * vel[r - 1] + S.ro * S.ducon[i] * vel[r - 1] + S.ro * S.ucon * S.dvel[r - 1][i];
          nat = nat + nv[r - 1] * S.dp[i];
        } else {
          nat = (dq5 + S.dp[i]) * S.ucon + (q[4] + S.p) * S.ducon[i];
        }
        out[5 * r + i] = 0.5 * (nat - dis);
      }
    }
  }
}

// Unpivoted 5x5 LU with reciprocal pivots in the reference's unrolled order
// (triSolvers.cpp:480-521): L[r][k] = b[r][k] - sum_{m<k} L[r][m] U[m][k],
// pivot_r = b[r][r] - sum_{m<r} L[r][m] U[m][r], U[r][k] = (b[r][k] - sum_{m<r}
// L[r][m] U[m][k]) / pivot_r, every sum left to right; then forward and back
// substitution.
struct Lu5 {
  double L[5][5], inv[5], U[5][5];
};

__device__ __forceinline__ Lu5 lu5(const double* b) {
  Lu5 F;
#pragma unroll
  for (int r = 0; r < 5; r++) {
#pragma unroll
    for (int k = 0; k < r; k++) {
      double v = b[5 * r + k];
#pragma unroll
      for (int m = 0; m < k; m++) v = v - F.L[r][m] * F.U[m][k];
      F.L[r][k] = v;
    }
    double piv = b[5 * r + r];
#pragma unroll
    for (int m = 0; m < r; m++) piv = piv - F.L[r][m] * F.U[m][r];
    F.inv[r] = 1. / piv;
#pragma unroll
    for (int k = r + 1; k < 5; k++) {
      double v = b[5 * r + k];
#pragma unroll
      for (int m = 0; m < r; m++) v = v - F.L[r][m] * F.U[m][k];
      F.U[r][k] = v * F.inv[r];
    }
  }
  return F;
}

// solves in place for the vector v[0], v[s], .., v[4 s]
__device__ __forceinline__ void lu5_solve(const Lu5& F, double* v, int s) {
  double d[5];
#pragma unroll
  for (int r = 0; r < 5; r++) {
    double x = v[r * s];
#pragma unroll
    for (int m = 0; m < r; m++) x = x - F.L[r][m] * d[m];
    d[r] = x * F.inv[r];
  }
  v[4 * s] = d[4];
#pragma unroll
  for (int r = 3; r >= 0; r--) {
    double x = d[r];
#pragma unroll
    for (int k = r + 1; k < 5; k++) x = x - F.U[r][k] * v[k * s];
    v[r * s] = x;
  }
}

// dst -= P Q (5x5, row-major), each entry summed over l left to right
__device__ __forceinline__ void sub_prod5(double* dst, const double* P, const double* Q) {
#pragma unroll
  for (int k = 0; k < 5; k++)
#pragma unroll
    for (int j = 0; j < 5; j++)
#pragma unroll
      for (int l = 0; l < 5; l++) dst[5 * k + j] -= P[5 * k + l] * Q[5 * l + j];
}

__device__ __forceinline__ void sub_prodv5(double* dst, const double* P, const double* x) {
#pragma unroll
  for (int k = 0; k < 5; k++)
#pragma unroll
    for (int j = 0; j < 5; j++) dst[k] -= P[5 * k + j] * x[j];
}

// blockTridag (triSolvers.cpp:443-536) and blockTridagPeriodic (:540-766) on
// one system, in place: a, b, c [N][5][5], f [N][5]; f becomes the solution.
// The periodic solve carries the spike column in a (a_i <- -a_i a_{i-1},
// normalised by B_i), folds it into c at row N-2, eliminates the bordered
// last row against every earlier row and back-substitutes with the spike.
__device__ inline void block_tridag5(double* a, double* b, double* c, double* f, int N, bool periodic) {
  for (int i = 0; i < N; i++) {
    double* ai = a + 25 * i;
    double* bi = b + 25 * i;
    double* ci = c + 25 * i;
    double* fi = f + 5 * i;
    if (i != 0) {
      sub_prod5(bi, ai, c + 25 * (i - 1));
      sub_prodv5(fi, ai, f + 5 * (i - 1));
      if (periodic) {
        double t[25];
#pragma unroll
        for (int e = 0; e < 25; e++) t[e] = 0;
        sub_prod5(t, ai, a + 25 * (i - 1));
#pragma unroll
        for (int e = 0; e < 25; e++) ai[e] = t[e];
      }
    }
    const Lu5 F = lu5(bi);
    lu5_solve(F, fi, 1);
    if (!periodic) {
      if (i != N - 1)
        for (int j = 0; j < 5; j++) lu5_solve(F, ci + j, 5);
    } else if (i != N - 2) {
      for (int j = 0; j < 5; j++) {
        lu5_solve(F, ci + j, 5);
        lu5_solve(F, ai + j, 5);
      }
    } else {
#pragma unroll
      for (int e = 0; e < 25; e++) ci[e] += ai[e];
      for (int j = 0; j < 5; j++) lu5_solve(F, ci + j, 5);
#pragma unroll
      for (int e = 0; e < 25; e++) ai[e] = ci[e];
    }
  }
  if (!periodic) {
    for (int i = N - 2; i >= 0; i--) sub_prodv5(f + 5 * i, c + 25 * i, f + 5 * (i + 1));
    return;
  }
  double* bl = b + 25 * (N - 1);
  double* cl = c + 25 * (N - 1);
  double* fl = f + 5 * (N - 1);
#pragma unroll
  for (int e = 0; e < 25; e++) bl[e] = (e % 6 == 0) ? 1.0 : 0.0;
  for (int i = 0; i < N - 1; i++) {
    double t[25];
#pragma unroll
    for (int e = 0; e < 25; e++) t[e] = 0;
    // one pass over (k, j, l), as the reference interleaves the two products
#pragma unroll
    for (int k = 0; k < 5; k++)
#pragma unroll
      for (int j = 0; j < 5; j++)
#pragma unroll
        for (int l = 0; l < 5; l++) {
          t[5 * k + j] -= cl[5 * k + l] * c[25 * i + 5 * l + j];
          bl[5 * k + j] -= cl[5 * k + l] * a[25 * i + 5 * l + j];
        }
    sub_prodv5(fl, cl, f + 5 * i);
#pragma unroll
    for (int e = 0; e < 25; e++) cl[e] = t[e];
  }
  const Lu5 F = lu5(bl);
  lu5_solve(F, fl, 1);
  sub_prodv5(f + 5 * (N - 2), c + 25 * (N - 2), fl);
  for (int i = N - 3; i >= 0; i--) {
#pragma unroll
    for (int k = 0; k < 5; k++)
#pragma unroll
      for (int j = 0; j < 5; j++)
        f[5 * i + k] -= (c[25 * i + 5 * k + j] * f[5 * (i + 1) + j] + a[25 * i + 5 * k + j] * fl[j]);
  }
}

}  // namespace hamstr

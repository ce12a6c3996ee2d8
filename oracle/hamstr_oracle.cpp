// ============================================================================
// hamstr_oracle.cpp -- CPU restatement of HAMSTR_OOP's line-implicit steady
// Euler iteration.  TEST INFRASTRUCTURE ONLY: the parity tests, the smoke
// check and the CPU-baseline leg of the benchmark load this library as the
// oracle.  The GPU product never links or calls it.
//
// Every function follows the reference file:line named in its comment, with
// the same floating-point operation order (build with -ffp-contract=off), so
// that its results are bit-identical to the reference built with g++ on
// x86-64 (SSE2, no FMA).  Pins:
//   * golden vectors written by the reference itself (tests/golden), compared
//     bit for bit in tests/test_oracle_golden.py;
//   * closed-form and invariant checks in tests/test_oracle_math.py.
// Differences from the reference, all outside the arithmetic:
//   * leak-free (the reference allocates 30 blocks per face Jacobian,
//     jacobianRoe.cpp:28-34,141-150, and never frees them);
//   * fluxRoe2D's out-of-bounds fm[4] write/qr[4] read (interfaceFlux.cpp:138)
//     is not reproduced (it does not reach any output);
//   * a NaN spectral radius sets a flag instead of exit(1)
//     (interfaceFlux.cpp:117-124).
// ============================================================================
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <chrono>

namespace {

constexpr int NQ = 4;
constexpr double THIRD = 0.3333333333333333e+00;   // utils.h:52
constexpr double PI = 3.1415926535897932e+00;      // utils.h:55

// --------------------------------------------------------------------------
// fluxRoe2D  (interfaceFlux.cpp:24-169)
// --------------------------------------------------------------------------
int flux_roe2d(const double* nxyz, const double* ql, const double* qr,
               double* flux, double* specRadius, double gamma) {
  double gm1 = gamma - 1.0;
  double nxd = nxyz[0], nyd = nxyz[1];
  double area = std::sqrt(nxd * nxd + nyd * nyd);
  double invArea = 1. / area;
  double nx = nxd * invArea, ny = nyd * invArea;

  double rol = ql[0];
  double ul = ql[1] / ql[0];
  double vl = ql[2] / ql[0];
  double pl = gm1 * (ql[3] - 0.5 * rol * (ul * ul + vl * vl));
  double hl = (ql[3] + pl) / rol;

  double ror = qr[0];
  double ur = qr[1] / qr[0];
  double vr = qr[2] / qr[0];
  double pr = gm1 * (qr[3] - 0.5 * ror * (ur * ur + vr * vr));
  double hr = (qr[3] + pr) / ror;

  double dp = pr - pl;
  double uconr = ur * nx + vr * ny;
  double uconl = ul * nx + vl * ny;

  double fact = std::sqrt(ror / rol);
  double A = 1.0 / (1.0 + fact);
  double B = fact / (1.0 + fact);
  double robar = rol * fact;
  double ubar = ul * A + ur * B;
  double vbar = vl * A + vr * B;
  double hbar = hl * A + hr * B;
  double cbar = gm1 * (hbar - 0.5 * (ubar * ubar + vbar * vbar));
  cbar = std::sqrt(cbar);
  double uconbar = ubar * nx + vbar * ny;

  double eig1 = std::fabs(uconbar);
  double eig2 = std::fabs(uconbar + cbar);
  double eig3 = std::fabs(uconbar - cbar);
  double sr = eig1 + cbar;
  int bad = (sr != sr);

  double term1 = -eig1 + 0.5 * (eig2 + eig3);
  double term2 = 0.5 * (eig2 - eig3);
  double del1 = term1 * dp / cbar / cbar + term2 * robar * (uconr - uconl) / cbar;
  double del2 = term1 * (uconr - uconl) * robar + term2 * dp / cbar;

  double fm[4];
  for (int i = 0; i < 4; i++) fm[i] = eig1 * (qr[i] - ql[i]);
  fm[0] = fm[0] + del1;
  fm[1] = fm[1] + del1 * ubar + del2 * nx;
  fm[2] = fm[2] + del1 * vbar + del2 * ny;
  fm[3] = fm[3] + del1 * hbar + del2 * uconbar;

  double fl[4], fr[4];
  fl[0] = rol * uconl;
  fr[0] = ror * uconr;
  fl[1] = rol * uconl * ul + nx * pl;
  fr[1] = ror * uconr * ur + nx * pr;
  fl[2] = rol * uconl * vl + ny * pl;
  fr[2] = ror * uconr * vr + ny * pr;
  fl[3] = (ql[3] + pl) * uconl;
  fr[3] = (qr[3] + pr) * uconr;

  for (int i = 0; i < 4; i++) {
    flux[i] = 0.5 * (fl[i] + fr[i] - fm[i]);
    flux[i] *= area;
  }
  *specRadius = sr * area;
  return bad;
}

// --------------------------------------------------------------------------
// jacobianRoe2D -> jacobianRoe3D with IMODE == 1
// (jacobianRoe.cpp:9-70 wrapper, :92-336 and :493-620 core).  The 2D state
// is embedded as a 5-vector with w = 0 and n_z = 0 exactly as the wrapper
// does, rows/columns {0,1,2,4} are extracted and scaled by the face length.
// --------------------------------------------------------------------------
void jacobian_roe3d(const double* nxyz, const double* ql, const double* qr,
                    double lmat[5][5], double rmat[5][5], double gam) {
  double gm1 = gam - 1.0;
  double nxd = nxyz[0], nyd = nxyz[1], nzd = nxyz[2];
  double area = std::sqrt(nxd * nxd + nyd * nyd + nzd * nzd);
  double nx = nxd / area, ny = nyd / area, nz = nzd / area;

  double rol = ql[0];
  double ul = ql[1] / ql[0];
  double vl = ql[2] / ql[0];
  double wl = ql[3] / ql[0];
  double pl = gm1 * (ql[4] - 0.5 * rol * (ul * ul + vl * vl + wl * wl));
  double hl = (ql[4] + pl) / rol;

  double ror = qr[0];
  double ur = qr[1] / qr[0];
  double vr = qr[2] / qr[0];
  double wr = qr[3] / qr[0];
  double pr = gm1 * (qr[4] - 0.5 * ror * (ur * ur + vr * vr + wr * wr));
  double hr = (qr[4] + pr) / ror;

  double dp = pr - pl;
  double uconr = ur * nx + vr * ny + wr * nz;
  double uconl = ul * nx + vl * ny + wl * nz;

  double dro_dql[5] = {1.0, 0, 0, 0, 0};
  double du_dql[5] = {0, 0, 0, 0, 0}, dv_dql[5] = {0, 0, 0, 0, 0}, dw_dql[5] = {0, 0, 0, 0, 0};
  du_dql[0] = -ul / rol;  du_dql[1] = 1.0 / rol;
  dv_dql[0] = -vl / rol;  dv_dql[2] = 1.0 / rol;
  dw_dql[0] = -wl / rol;  dw_dql[3] = 1.0 / rol;
  double dp_dql[5];
  dp_dql[0] = 0.5 * gm1 * (ul * ul + vl * vl + wl * wl);
  dp_dql[1] = -gm1 * ul;
  dp_dql[2] = -gm1 * vl;
  dp_dql[3] = -gm1 * wl;
  dp_dql[4] = gm1;
  double dq5_dql[5] = {0, 0, 0, 0, 1.0};
  double ducon_dql[5];
  ducon_dql[0] = -uconl / rol;
  ducon_dql[1] = nx / rol;
  ducon_dql[2] = ny / rol;
  ducon_dql[3] = nz / rol;
  ducon_dql[4] = 0.0;

  double dro_dqr[5] = {1.0, 0, 0, 0, 0};
  double du_dqr[5] = {0, 0, 0, 0, 0}, dv_dqr[5] = {0, 0, 0, 0, 0}, dw_dqr[5] = {0, 0, 0, 0, 0};
  du_dqr[0] = -ur / ror;  du_dqr[1] = 1.0 / ror;
  dv_dqr[0] = -vr / ror;  dv_dqr[2] = 1.0 / ror;
  dw_dqr[0] = -wr / ror;  dw_dqr[3] = 1.0 / ror;
  double dp_dqr[5];
  dp_dqr[0] = 0.5 * gm1 * (ur * ur + vr * vr + wr * wr);
  dp_dqr[1] = -gm1 * ur;
  dp_dqr[2] = -gm1 * vr;
  dp_dqr[3] = -gm1 * wr;
  dp_dqr[4] = gm1;
  double dq5_dqr[5] = {0, 0, 0, 0, 1.0};
  double ducon_dqr[5];
  ducon_dqr[0] = -uconr / ror;
  ducon_dqr[1] = nx / ror;
  ducon_dqr[2] = ny / ror;
  ducon_dqr[3] = nz / ror;
  ducon_dqr[4] = 0.0;

  double fact = std::sqrt(ror / rol);
  double A = 1.0 / (1.0 + fact);
  double B = fact / (1.0 + fact);
  double robar = rol * fact;
  double ubar = ul * A + ur * B;
  double vbar = vl * A + vr * B;
  double wbar = wl * A + wr * B;
  double hbar = hl * A + hr * B;
  double cbar = gm1 * (hbar - 0.5 * (ubar * ubar + vbar * vbar + wbar * wbar));
  cbar = std::sqrt(cbar);
  double uconbar = ubar * nx + vbar * ny + wbar * nz;

  double eig1 = std::fabs(uconbar);
  double eig2 = std::fabs(uconbar + cbar);
  double eig3 = std::fabs(uconbar - cbar);

  // IMODE == 1 branch (jacobianRoe.cpp:321-336)
  double term1 = -eig1 + 0.5 * (eig2 + eig3);
  double term2 = 0.5 * (eig2 - eig3);
  double ddel1_dql[5], ddel1_dqr[5], ddel2_dql[5], ddel2_dqr[5];
  for (int i = 0; i < 5; i++) {
    ddel1_dql[i] = -term1 * dp_dql[i] / cbar / cbar - term2 * robar * ducon_dql[i] / cbar;
    ddel1_dqr[i] = +term1 * dp_dqr[i] / cbar / cbar + term2 * robar * ducon_dqr[i] / cbar;
    ddel2_dql[i] = -term1 * ducon_dql[i] * robar - term2 * dp_dql[i] / cbar;
    ddel2_dqr[i] = +term1 * ducon_dqr[i] * robar + term2 * dp_dqr[i] / cbar;
  }

  // common linearisation terms (jacobianRoe.cpp:500-531)
  double imat[5][5];
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) lmat[i][j] = rmat[i][j] = imat[i][j] = 0.0;
  for (int i = 0; i < 5; i++) imat[i][i] = 1.0;
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) {
      lmat[i][j] = lmat[i][j] - eig1 * imat[i][j];
      rmat[i][j] = rmat[i][j] + eig1 * imat[i][j];
    }
  for (int i = 0; i < 5; i++) {
    lmat[0][i] = lmat[0][i] + ddel1_dql[i];
    rmat[0][i] = rmat[0][i] + ddel1_dqr[i];
    lmat[1][i] = lmat[1][i] + ddel1_dql[i] * ubar + ddel2_dql[i] * nx;
    rmat[1][i] = rmat[1][i] + ddel1_dqr[i] * ubar + ddel2_dqr[i] * nx;
    lmat[2][i] = lmat[2][i] + ddel1_dql[i] * vbar + ddel2_dql[i] * ny;
    rmat[2][i] = rmat[2][i] + ddel1_dqr[i] * vbar + ddel2_dqr[i] * ny;
    lmat[3][i] = lmat[3][i] + ddel1_dql[i] * wbar + ddel2_dql[i] * nz;
    rmat[3][i] = rmat[3][i] + ddel1_dqr[i] * wbar + ddel2_dqr[i] * nz;
    lmat[4][i] = lmat[4][i] + ddel1_dql[i] * hbar + ddel2_dql[i] * uconbar;
    rmat[4][i] = rmat[4][i] + ddel1_dqr[i] * hbar + ddel2_dqr[i] * uconbar;
  }

  // native flux Jacobian (jacobianRoe.cpp:573-605)
  double lmat1[5][5], rmat1[5][5];
  for (int i = 0; i < 5; i++) {
    lmat1[0][i] = dro_dql[i] * uconl + rol * ducon_dql[i];
    lmat1[1][i] = dro_dql[i] * uconl * ul + rol * ducon_dql[i] * ul + rol * uconl * du_dql[i];
    lmat1[2][i] = dro_dql[i] * uconl * vl + rol * ducon_dql[i] * vl + rol * uconl * dv_dql[i];
    lmat1[3][i] = dro_dql[i] * uconl * wl + rol * ducon_dql[i] * wl + rol * uconl * dw_dql[i];
    lmat1[2][i] = lmat1[2][i] + ny * dp_dql[i];
    lmat1[1][i] = lmat1[1][i] + nx * dp_dql[i];
    lmat1[3][i] = lmat1[3][i] + nz * dp_dql[i];
    lmat1[4][i] = (dq5_dql[i] + dp_dql[i]) * uconl + (ql[4] + pl) * ducon_dql[i];
  }
  for (int i = 0; i < 5; i++) {
    rmat1[0][i] = dro_dqr[i] * uconr + ror * ducon_dqr[i];
    rmat1[1][i] = dro_dqr[i] * uconr * ur + ror * ducon_dqr[i] * ur + ror * uconr * du_dqr[i];
    rmat1[2][i] = dro_dqr[i] * uconr * vr + ror * ducon_dqr[i] * vr + ror * uconr * dv_dqr[i];
    rmat1[3][i] = dro_dqr[i] * uconr * wr + ror * ducon_dqr[i] * wr + ror * uconr * dw_dqr[i];
    rmat1[2][i] = rmat1[2][i] + ny * dp_dqr[i];
    rmat1[1][i] = rmat1[1][i] + nx * dp_dqr[i];
    rmat1[3][i] = rmat1[3][i] + nz * dp_dqr[i];
    rmat1[4][i] = (dq5_dqr[i] + dp_dqr[i]) * uconr + (qr[4] + pr) * ducon_dqr[i];
  }
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) {
      lmat[i][j] = 0.5 * (lmat1[i][j] - lmat[i][j]);
      rmat[i][j] = 0.5 * (rmat1[i][j] - rmat[i][j]);
    }
}

void jacobian_roe2d(const double* nxyz2, const double* ql2, const double* qr2,
                    double* lmat2, double* rmat2, double gam) {
  static const int idiv[4] = {0, 1, 2, 4};
  double nxyz[3] = {nxyz2[0], nxyz2[1], 0};
  double area = std::sqrt(nxyz2[0] * nxyz2[0] + nxyz2[1] * nxyz2[1]);
  double ql[5] = {ql2[0], ql2[1], ql2[2], 0, ql2[3]};
  double qr[5] = {qr2[0], qr2[1], qr2[2], 0, qr2[3]};
  double lmat[5][5], rmat[5][5];
  jacobian_roe3d(nxyz, ql, qr, lmat, rmat, gam);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      lmat2[4 * i + j] = lmat[idiv[i]][idiv[j]] * area;
      rmat2[4 * i + j] = rmat[idiv[i]][idiv[j]] * area;
    }
}

// --------------------------------------------------------------------------
// Reconstruction (reconstruction.cpp).  Arrays are row-major [imax][NQ].
// --------------------------------------------------------------------------
void muscl(const double* f, double* ql, double* qr, double* f2, int is, int ie,
           double th, double qt, double eps, int imax, int nq) {
  // reconstruction.cpp:16-70
  int im = imax - 1;
  if (qt == 0.0) {
    for (int n = 0; n < nq; n++)
      for (int i = is; i <= ie; i++) {
        ql[i * NQ + n] = f[i * NQ + n];
        qr[i * NQ + n] = f[i * NQ + n];
      }
    return;
  }
  double thm = 1.0 - th, thp = 1.0 + th;
  for (int n = 0; n < nq; n++) {
    for (int i = 0; i < im; i++) f2[i * NQ + n] = f[(i + 1) * NQ + n] - f[i * NQ + n];
    for (int i = is; i <= ie; i++) {
      double f2i = f2[i * NQ + n];
      double f2i1 = f2[(i - 1) * NQ + n];
      double a1 = 3.0 * (f2i * f2i1 + eps);
      double a2 = 2.0 * (f2i - f2i1) * (f2i - f2i1) + a1;
      double f3i = a1 / a2;
      double f3qt = qt * f3i;
      ql[i * NQ + n] = f[i * NQ + n] + f3qt * (thm * f2i1 + thp * f2i);
      qr[i * NQ + n] = f[i * NQ + n] - f3qt * (thp * f2i1 + thm * f2i);
    }
  }
}

double weno5(double a, double b, double c, double d, double e, double epsw) {
  // reconstruction.cpp:78-124 (p = 1, so pow(x, p) == x)
  double b1 = 13. / 12.;
  double b2 = 1. / 6.;
  double djm1 = a - 2. * b + c;
  double ejm1 = a - 4. * b + 3. * c;
  double dj = b - 2. * c + d;
  double ej = b - d;
  double djp1 = c - 2. * d + e;
  double ejp1 = 3. * c - 4. * d + e;
  double dis0 = b1 * djm1 * djm1 + 0.25 * ejm1 * ejm1 + epsw;
  double dis1 = b1 * dj * dj + 0.25 * ej * ej + epsw;
  double dis2 = b1 * djp1 * djp1 + 0.25 * ejp1 * ejp1 + epsw;
  double q30 = 2. * a - 7. * b + 11. * c;
  double q31 = -b + 5. * c + 2. * d;
  double q32 = 2. * c + 5. * d - e;
  double d01 = dis0 / dis1;
  double d02 = dis0 / dis2;
  double a1ba0 = 6. * d01;
  double a2ba0 = 3. * d02;
  double w0 = 1.0 / (1.0 + a1ba0 + a2ba0);
  double w1 = a1ba0 * w0;
  double w2 = 1. - w0 - w1;
  return b2 * (w0 * q30 + w1 * q31 + w2 * q32);
}

void weno(const double* f, double* ql, double* qr, int is, int ie, double epsw, int imax, int nq) {
  // reconstruction.cpp:133-160: epsilon is forced to 1e-15
  (void)ie;
  epsw = 1e-15;
  int im = imax - 2;
  for (int n = 0; n < nq; n++)
    for (int i = is; i < im; i++) {
      ql[i * NQ + n] = weno5(f[(i - 2) * NQ + n], f[(i - 1) * NQ + n], f[i * NQ + n],
                             f[(i + 1) * NQ + n], f[(i + 2) * NQ + n], epsw);
      qr[i * NQ + n] = weno5(f[(i + 2) * NQ + n], f[(i + 1) * NQ + n], f[i * NQ + n],
                             f[(i - 1) * NQ + n], f[(i - 2) * NQ + n], epsw);
    }
}

void muscl_deriv(const double* f, double* ql, double* qr, double* f2, const double* dq,
                 int is, int ie, double th, double qt, double eps, int imax, int nq) {
  // reconstruction.cpp:168-236
  int im = imax - 1;
  if (qt == 0.0) {
    for (int n = 0; n < nq; n++)
      for (int i = is; i <= ie; i++) {
        ql[i * NQ + n] = dq[i * NQ + n];
        qr[i * NQ + n] = dq[i * NQ + n];
      }
    return;
  }
  double thm = 1.0 - th, thp = 1.0 + th;
  for (int n = 0; n < nq; n++) {
    for (int i = 0; i < im; i++) f2[i * NQ + n] = f[(i + 1) * NQ + n] - f[i * NQ + n];
    for (int i = is; i <= ie; i++) {
      double f2i = f2[i * NQ + n];
      double f2i1 = f2[(i - 1) * NQ + n];
      // dq[i+1] at i == ie reads one row past the chain; that row is never used
      double dqp = (i + 1 < imax) ? dq[(i + 1) * NQ + n] : 0.0;
      double df2i = dqp - dq[i * NQ + n];
      double df2i1 = dq[i * NQ + n] - dq[(i - 1) * NQ + n];
      double a1 = 3.0 * (f2i * f2i1 + eps);
      double da1 = 3.0 * (df2i * f2i1 + f2i * df2i1);
      double a2 = 2.0 * (f2i - f2i1) * (f2i - f2i1) + a1;
      double da2 = 2.0 * 2.0 * (f2i - f2i1) * (df2i - df2i1) + da1;
      double f3i = a1 / a2;
      double df3i = da1 / a2 - a1 * da2 / (a2 * a2);
      double f3qt = qt * f3i;
      double df3qt = qt * df3i;
      ql[i * NQ + n] = dq[i * NQ + n] + df3qt * (thm * f2i1 + thp * f2i) + f3qt * (thm * df2i1 + thp * df2i);
      qr[i * NQ + n] = dq[i * NQ + n] - df3qt * (thp * f2i1 + thm * f2i) - f3qt * (thp * df2i1 + thm * df2i);
    }
  }
}

double weno5_deriv(double a, double b, double c, double d, double e,
                   double a_d, double b_d, double c_d, double d_d, double e_d, double epsw) {
  // reconstruction.cpp:243-347 (WENO_5)
  double b1 = 13. / 12.;
  double b2 = 1. / 6.;
  double djm1 = a - 2. * b + c;
  double ejm1 = a - 4. * b + 3. * c;
  double dj = b - 2. * c + d;
  double ej = b - d;
  double djp1 = c - 2. * d + e;
  double ejp1 = 3. * c - 4. * d + e;
  double dis0 = b1 * djm1 * djm1 + 0.25 * ejm1 * ejm1 + epsw;
  double dis1 = b1 * dj * dj + 0.25 * ej * ej + epsw;
  double dis2 = b1 * djp1 * djp1 + 0.25 * ejp1 * ejp1 + epsw;
  double q30 = 2. * a - 7. * b + 11. * c;
  double q31 = -b + 5. * c + 2. * d;
  double q32 = 2. * c + 5. * d - e;
  double d01 = dis0 / dis1;
  double d02 = dis0 / dis2;
  double a1ba0 = 6. * d01;
  double a2ba0 = 3. * d02;
  double w0 = 1.0 / (1.0 + a1ba0 + a2ba0);
  double w1 = a1ba0 * w0;
  double w2 = 1. - w0 - w1;

  double djm1_d = a_d - 2. * b_d + c_d;
  double ejm1_d = a_d - 4. * b_d + 3. * c_d;
  double dj_d = b_d - 2. * c_d + d_d;
  double ej_d = b_d - d_d;
  double djp1_d = c_d - 2. * d_d + e_d;
  double ejp1_d = 3. * c_d - 4. * d_d + e_d;
  double q30_d = 2. * a_d - 7. * b_d + 11. * c_d;
  double q31_d = -b_d + 5. * c_d + 2. * d_d;
  double q32_d = 2. * c_d + 5. * d_d - e_d;

  double term1 = b2 * w0 * q30_d;
  double term2 = b2 * w1 * q31_d;
  double term3 = b2 * w2 * q32_d;

  double dbeta0 = b1 * 2 * djm1 * djm1_d + 0.25 * 2 * ejm1 * ejm1_d;
  double dbeta1 = b1 * 2 * dj * dj_d + 0.25 * 2 * ej * ej_d;
  double dbeta2 = b1 * 2 * djp1 * djp1_d + 0.25 * 2 * ejp1 * ejp1_d;
  double beta0 = dis0, beta1 = dis1, beta2 = dis2;

  double dw0_down = beta1 * beta2 + 6 * beta0 * beta2 + 3 * beta0 * beta1;
  double dw0_up_1 = (dbeta1 * beta2 + beta1 * dbeta2) * dw0_down;
  double dw0_up_2 = beta1 * beta2 * (dbeta1 * beta2 + beta1 * dbeta2 + 6 * (dbeta0 * beta2 + beta0 * dbeta2) +
                                     3 * (dbeta0 * beta1 + beta0 * dbeta1));
  double dw0 = (dw0_up_1 - dw0_up_2) / (dw0_down * dw0_down);

  double dw1_down = dw0_down;
  double dw1_up_1 = 6 * (dbeta0 * beta2 + beta0 * dbeta2) * dw1_down;
  double dw1_up_2 = 6 * dw0_up_2 / beta1 * beta0;
  double dw1 = (dw1_up_1 - dw1_up_2) / (dw1_down * dw1_down);

  double dw2_down = dw0_down;
  double dw2_up_1 = 3 * (dbeta0 * beta1 + beta0 * dbeta1) * dw2_down;
  double dw2_up_2 = 3 * dw0_up_2 / beta2 * beta0;
  double dw2 = (dw2_up_1 - dw2_up_2) / (dw2_down * dw2_down);

  return term1 + term2 + term3 + b2 * dw0 * q30 + b2 * dw1 * q31 + b2 * dw2 * q32;
}

void weno_deriv(const double* f, double* ql, double* qr, const double* dq, int is, int ie,
                double eps, int imax, int nq) {
  // reconstruction.cpp:353-386
  (void)ie;
  int im = imax - 2;
  for (int n = 0; n < nq; n++)
    for (int i = is; i < im; i++) {
      ql[i * NQ + n] = weno5_deriv(f[(i - 2) * NQ + n], f[(i - 1) * NQ + n], f[i * NQ + n],
                                   f[(i + 1) * NQ + n], f[(i + 2) * NQ + n],
                                   dq[(i - 2) * NQ + n], dq[(i - 1) * NQ + n], dq[i * NQ + n],
                                   dq[(i + 1) * NQ + n], dq[(i + 2) * NQ + n], eps);
      qr[i * NQ + n] = weno5_deriv(f[(i + 2) * NQ + n], f[(i + 1) * NQ + n], f[i * NQ + n],
                                   f[(i - 1) * NQ + n], f[(i - 2) * NQ + n],
                                   dq[(i + 2) * NQ + n], dq[(i + 1) * NQ + n], dq[i * NQ + n],
                                   dq[(i - 1) * NQ + n], dq[(i - 2) * NQ + n], eps);
    }
}

// --------------------------------------------------------------------------
// Block tridiagonal solvers, 4x4 blocks (triSolvers.cpp:769-847, 850-1046).
// a, b, c: [N][4][4] row-major; f: [N][4]; all overwritten in place.
// --------------------------------------------------------------------------
#define M3(p, i, k, j) p[((i) * 4 + (k)) * 4 + (j)]
#define V2(p, i, k) p[(i) * 4 + (k)]

struct LU4 {
  double b11, b21, b22, b31, b32, b33, b41, b42, b43, b44;
  double u12, u13, u14, u23, u24, u34;
};

inline LU4 lu4(const double* b) {
  // unpivoted LU with reciprocal pivots (triSolvers.cpp:805-820)
  LU4 L;
  L.b11 = 1. / b[0];
  L.u12 = b[1] * L.b11;
  L.u13 = b[2] * L.b11;
  L.u14 = b[3] * L.b11;
  L.b21 = b[4];
  L.b22 = 1. / (b[5] - L.b21 * L.u12);
  L.u23 = (b[6] - L.b21 * L.u13) * L.b22;
  L.u24 = (b[7] - L.b21 * L.u14) * L.b22;
  L.b31 = b[8];
  L.b32 = b[9] - L.b31 * L.u12;
  L.b33 = 1. / (b[10] - L.b31 * L.u13 - L.b32 * L.u23);
  L.u34 = (b[11] - L.b31 * L.u14 - L.b32 * L.u24) * L.b33;
  L.b41 = b[12];
  L.b42 = b[13] - L.b41 * L.u12;
  L.b43 = b[14] - L.b41 * L.u13 - L.b42 * L.u23;
  L.b44 = 1. / (b[15] - L.b41 * L.u14 - L.b42 * L.u24 - L.b43 * L.u34);
  return L;
}

// solve in place for a vector stored with stride s: v[0], v[s], v[2s], v[3s]
inline void lu4_solve(const LU4& L, double* v, int s) {
  double d1 = v[0] * L.b11;
  double d2 = (v[s] - L.b21 * d1) * L.b22;
  double d3 = (v[2 * s] - L.b31 * d1 - L.b32 * d2) * L.b33;
  double d4 = (v[3 * s] - L.b41 * d1 - L.b42 * d2 - L.b43 * d3) * L.b44;
  v[3 * s] = d4;
  v[2 * s] = d3 - L.u34 * v[3 * s];
  v[s] = d2 - L.u23 * v[2 * s] - L.u24 * v[3 * s];
  v[0] = d1 - L.u12 * v[s] - L.u13 * v[2 * s] - L.u14 * v[3 * s];
}

void block_tridag4(double* a, double* b, double* c, double* f, int N) {
  for (int i = 0; i < N; i++) {
    if (i != 0) {
      for (int k = 0; k < 4; k++)
        for (int j = 0; j < 4; j++)
          for (int l = 0; l < 4; l++) M3(b, i, k, j) -= M3(a, i, k, l) * M3(c, i - 1, l, j);
      for (int k = 0; k < 4; k++)
        for (int j = 0; j < 4; j++) V2(f, i, k) -= M3(a, i, k, j) * V2(f, i - 1, j);
    }
    LU4 L = lu4(&M3(b, i, 0, 0));
    lu4_solve(L, &V2(f, i, 0), 1);
    if (i != N - 1)
      for (int j = 0; j < 4; j++) lu4_solve(L, &M3(c, i, 0, j), 4);
  }
  for (int i = N - 2; i >= 0; i--)
    for (int k = 0; k < 4; k++)
      for (int j = 0; j < 4; j++) V2(f, i, k) -= M3(c, i, k, j) * V2(f, i + 1, j);
}

void block_tridag_periodic4(double* a, double* b, double* c, double* f, int N) {
  double t[4][4];
  for (int i = 0; i < N; i++) {
    if (i != 0) {
      for (int k = 0; k < 4; k++)
        for (int j = 0; j < 4; j++)
          for (int l = 0; l < 4; l++) M3(b, i, k, j) -= M3(a, i, k, l) * M3(c, i - 1, l, j);
      for (int k = 0; k < 4; k++)
        for (int j = 0; j < 4; j++) V2(f, i, k) -= M3(a, i, k, j) * V2(f, i - 1, j);
      for (int k = 0; k < 4; k++)
        for (int j = 0; j < 4; j++) t[k][j] = 0;
      for (int k = 0; k < 4; k++)
        for (int j = 0; j < 4; j++)
          for (int l = 0; l < 4; l++) t[k][j] -= M3(a, i, k, l) * M3(a, i - 1, l, j);
      for (int k = 0; k < 4; k++)
        for (int j = 0; j < 4; j++) M3(a, i, k, j) = t[k][j];
    }
    LU4 L = lu4(&M3(b, i, 0, 0));
    lu4_solve(L, &V2(f, i, 0), 1);
    if (i != N - 2) {
      for (int j = 0; j < 4; j++) {
        lu4_solve(L, &M3(c, i, 0, j), 4);
        lu4_solve(L, &M3(a, i, 0, j), 4);
      }
    }
    if (i == N - 2) {
      for (int k = 0; k < 4; k++)
        for (int j = 0; j < 4; j++) M3(c, i, j, k) += M3(a, i, j, k);
      for (int j = 0; j < 4; j++) lu4_solve(L, &M3(c, i, 0, j), 4);
      for (int k = 0; k < 4; k++)
        for (int j = 0; j < 4; j++) M3(a, i, k, j) = M3(c, i, k, j);
    }
  }
  int i = N - 1;
  for (int k = 0; k < 4; k++)
    for (int j = 0; j < 4; j++) M3(b, i, j, k) = (j == k) ? 1.0 : 0.0;
  for (i = 0; i < N - 1; i++) {
    for (int k = 0; k < 4; k++)
      for (int j = 0; j < 4; j++) t[k][j] = 0;
    for (int k = 0; k < 4; k++)
      for (int j = 0; j < 4; j++)
        for (int l = 0; l < 4; l++) {
          t[k][j] -= M3(c, N - 1, k, l) * M3(c, i, l, j);
          M3(b, N - 1, k, j) -= M3(c, N - 1, k, l) * M3(a, i, l, j);
        }
    for (int k = 0; k < 4; k++)
      for (int j = 0; j < 4; j++) V2(f, N - 1, k) -= M3(c, N - 1, k, j) * V2(f, i, j);
    for (int k = 0; k < 4; k++)
      for (int j = 0; j < 4; j++) M3(c, N - 1, k, j) = t[k][j];
  }
  i = N - 1;
  LU4 L = lu4(&M3(b, i, 0, 0));
  lu4_solve(L, &V2(f, i, 0), 1);
  i = N - 2;
  for (int k = 0; k < 4; k++)
    for (int j = 0; j < 4; j++) V2(f, i, k) -= (M3(c, i, k, j) * V2(f, i + 1, j));
  for (i = N - 3; i >= 0; i--)
    for (int k = 0; k < 4; k++)
      for (int j = 0; j < 4; j++)
        V2(f, i, k) -= (M3(c, i, k, j) * V2(f, i + 1, j) + M3(a, i, k, j) * V2(f, N - 1, j));
}

// --------------------------------------------------------------------------
// The 3D / strand building blocks (5 conserved variables).  The reference's
// 3D build does not compile (io.cpp:306), but these functions are compiled in
// its 2D build and are pinned here against it.
// --------------------------------------------------------------------------

// fluxRoe3D (interfaceFlux.cpp:176-319): Roe flux with the unit normal; unlike
// fluxRoe2D neither the flux nor the spectral radius is scaled by the area.
void flux_roe3d(const double* nxyz, const double* ql, const double* qr, double* flux, double* specRadius,
                double gamma) {
  double gm1 = gamma - 1.0;
  double nxd = nxyz[0], nyd = nxyz[1], nzd = nxyz[2];
  double area = std::sqrt(nxd * nxd + nyd * nyd + nzd * nzd);
  double invArea = 1. / area;
  double nx = nxd * invArea, ny = nyd * invArea, nz = nzd * invArea;

  double rol = ql[0];
  double ul = ql[1] / ql[0];
  double vl = ql[2] / ql[0];
  double wl = ql[3] / ql[0];
  double pl = gm1 * (ql[4] - 0.5 * rol * (ul * ul + vl * vl + wl * wl));
  double hl = (ql[4] + pl) / rol;

  double ror = qr[0];
  double ur = qr[1] / qr[0];
  double vr = qr[2] / qr[0];
  double wr = qr[3] / qr[0];
  double pr = gm1 * (qr[4] - 0.5 * ror * (ur * ur + vr * vr + wr * wr));
  double hr = (qr[4] + pr) / ror;

  double dp = pr - pl;
  double uconr = ur * nx + vr * ny + wr * nz;
  double uconl = ul * nx + vl * ny + wl * nz;

  double fact = std::sqrt(ror / rol);
  double A = 1.0 / (1.0 + fact);
  double B = fact / (1.0 + fact);
  double robar = rol * fact;
  double ubar = ul * A + ur * B;
  double vbar = vl * A + vr * B;
  double wbar = wl * A + wr * B;
  double hbar = hl * A + hr * B;
  double cbar = gm1 * (hbar - 0.5 * (ubar * ubar + vbar * vbar + wbar * wbar));
  cbar = std::sqrt(cbar);
  double uconbar = ubar * nx + vbar * ny + wbar * nz;

  double eig1 = std::fabs(uconbar);
  double eig2 = std::fabs(uconbar + cbar);
  double eig3 = std::fabs(uconbar - cbar);
  *specRadius = eig1 + cbar;

  double term1 = -eig1 + 0.5 * (eig2 + eig3);
  double term2 = 0.5 * (eig2 - eig3);
  double del1 = term1 * dp / cbar / cbar + term2 * robar * (uconr - uconl) / cbar;
  double del2 = term1 * (uconr - uconl) * robar + term2 * dp / cbar;

  double fm[5], fl[5], fr[5];
  for (int i = 0; i < 5; i++) fm[i] = eig1 * (qr[i] - ql[i]);
  fm[0] = fm[0] + del1;
  fm[1] = fm[1] + del1 * ubar + del2 * nx;
  fm[2] = fm[2] + del1 * vbar + del2 * ny;
  fm[3] = fm[3] + del1 * wbar + del2 * nz;
  fm[4] = fm[4] + del1 * hbar + del2 * uconbar;

  fl[0] = rol * uconl;
  fr[0] = ror * uconr;
  fl[1] = rol * uconl * ul + nx * pl;
  fr[1] = ror * uconr * ur + nx * pr;
  fl[2] = rol * uconl * vl + ny * pl;
  fr[2] = ror * uconr * vr + ny * pr;
  fl[3] = rol * uconl * wl + nz * pl;
  fr[3] = ror * uconr * wr + nz * pr;
  fl[4] = (ql[4] + pl) * uconl;
  fr[4] = (qr[4] + pr) * uconr;
  for (int i = 0; i < 5; i++) flux[i] = 0.5 * (fl[i] + fr[i] - fm[i]);
}

// 5x5 block tridiagonal solvers (triSolvers.cpp:443-536 blockTridag,
// :540-766 blockTridagPeriodic).  a, b, c: [N][5][5] row-major, f: [N][5], all
// overwritten in place as the reference does.  The unpivoted LU keeps the
// reference's unrolled formulas (:480-509): L[r][k] = b[r][k] - sum_{m<k}
// L[r][m] U[m][k] and U[r][k] = (b[r][k] - sum_{m<r} L[r][m] U[m][k]) / pivot,
// with the pivot's reciprocal, every sum taken left to right.
#define M5(p, i, k, j) p[((i) * 5 + (k)) * 5 + (j)]
#define V5(p, i, k) p[(i) * 5 + (k)]

struct LU5 {
  double L[5][5];   // strictly lower part, the original column-0 entries on k = 0
  double inv[5];    // reciprocal pivots
  double U[5][5];   // strictly upper part
};

inline LU5 lu5(const double* b) {
  LU5 F;
  for (int r = 0; r < 5; r++) {
    for (int k = 0; k < r; k++) {
      double v = b[5 * r + k];
      for (int m = 0; m < k; m++) v = v - F.L[r][m] * F.U[m][k];
      F.L[r][k] = v;
    }
    double piv = b[5 * r + r];
    for (int m = 0; m < r; m++) piv = piv - F.L[r][m] * F.U[m][r];
    F.inv[r] = 1. / piv;
    for (int k = r + 1; k < 5; k++) {
      double v = b[5 * r + k];
      for (int m = 0; m < r; m++) v = v - F.L[r][m] * F.U[m][k];
      F.U[r][k] = v * F.inv[r];
    }
  }
  return F;
}

// solve in place for a vector stored with stride s (:511-521)
inline void lu5_solve(const LU5& F, double* v, int s) {
  double d[5];
  for (int r = 0; r < 5; r++) {
    double x = v[r * s];
    for (int m = 0; m < r; m++) x = x - F.L[r][m] * d[m];
    d[r] = x * F.inv[r];
  }
  v[4 * s] = d[4];
  for (int r = 3; r >= 0; r--) {
    double x = d[r];
    for (int k = r + 1; k < 5; k++) x = x - F.U[r][k] * v[k * s];
    v[r * s] = x;
  }
}

void block_tridag5(double* a, double* b, double* c, double* f, int N) {
  for (int i = 0; i < N; i++) {
    if (i != 0) {
      for (int k = 0; k < 5; k++)
        for (int j = 0; j < 5; j++)
          for (int l = 0; l < 5; l++) M5(b, i, k, j) -= M5(a, i, k, l) * M5(c, i - 1, l, j);
      for (int k = 0; k < 5; k++)
        for (int j = 0; j < 5; j++) V5(f, i, k) -= M5(a, i, k, j) * V5(f, i - 1, j);
    }
    LU5 F = lu5(&M5(b, i, 0, 0));
    lu5_solve(F, &V5(f, i, 0), 1);
    if (i != N - 1)
      for (int j = 0; j < 5; j++) lu5_solve(F, &M5(c, i, 0, j), 5);
  }
  for (int i = N - 2; i >= 0; i--)
    for (int k = 0; k < 5; k++)
      for (int j = 0; j < 5; j++) V5(f, i, k) -= M5(c, i, k, j) * V5(f, i + 1, j);
}

void block_tridag_periodic5(double* a, double* b, double* c, double* f, int N) {
  double t[5][5];
  for (int i = 0; i < N; i++) {
    if (i != 0) {
      for (int k = 0; k < 5; k++)
        for (int j = 0; j < 5; j++)
          for (int l = 0; l < 5; l++) M5(b, i, k, j) -= M5(a, i, k, l) * M5(c, i - 1, l, j);
      for (int k = 0; k < 5; k++)
        for (int j = 0; j < 5; j++) V5(f, i, k) -= M5(a, i, k, j) * V5(f, i - 1, j);
      for (int k = 0; k < 5; k++)
        for (int j = 0; j < 5; j++) t[k][j] = 0;
      for (int k = 0; k < 5; k++)
        for (int j = 0; j < 5; j++)
          for (int l = 0; l < 5; l++) t[k][j] -= M5(a, i, k, l) * M5(a, i - 1, l, j);
      for (int k = 0; k < 5; k++)
        for (int j = 0; j < 5; j++) M5(a, i, k, j) = t[k][j];
    }
    LU5 F = lu5(&M5(b, i, 0, 0));
    lu5_solve(F, &V5(f, i, 0), 1);
    if (i != N - 2) {
      for (int j = 0; j < 5; j++) {
        lu5_solve(F, &M5(c, i, 0, j), 5);
        lu5_solve(F, &M5(a, i, 0, j), 5);
      }
    }
    if (i == N - 2) {
      for (int k = 0; k < 5; k++)
        for (int j = 0; j < 5; j++) M5(c, i, j, k) += M5(a, i, j, k);
      for (int j = 0; j < 5; j++) lu5_solve(F, &M5(c, i, 0, j), 5);
      for (int k = 0; k < 5; k++)
        for (int j = 0; j < 5; j++) M5(a, i, k, j) = M5(c, i, k, j);
    }
  }
  int i = N - 1;
  for (int k = 0; k < 5; k++)
    for (int j = 0; j < 5; j++) M5(b, i, j, k) = (j == k) ? 1.0 : 0.0;
  for (i = 0; i < N - 1; i++) {
    for (int k = 0; k < 5; k++)
      for (int j = 0; j < 5; j++) t[k][j] = 0;
    for (int k = 0; k < 5; k++)
      for (int j = 0; j < 5; j++)
        for (int l = 0; l < 5; l++) {
          t[k][j] -= M5(c, N - 1, k, l) * M5(c, i, l, j);
          M5(b, N - 1, k, j) -= M5(c, N - 1, k, l) * M5(a, i, l, j);
        }
    for (int k = 0; k < 5; k++)
      for (int j = 0; j < 5; j++) V5(f, N - 1, k) -= M5(c, N - 1, k, j) * V5(f, i, j);
    for (int k = 0; k < 5; k++)
      for (int j = 0; j < 5; j++) M5(c, N - 1, k, j) = t[k][j];
  }
  i = N - 1;
  LU5 F = lu5(&M5(b, i, 0, 0));
  lu5_solve(F, &V5(f, i, 0), 1);
  i = N - 2;
  for (int k = 0; k < 5; k++)
    for (int j = 0; j < 5; j++) V5(f, i, k) -= (M5(c, i, k, j) * V5(f, i + 1, j));
  for (i = N - 3; i >= 0; i--)
    for (int k = 0; k < 5; k++)
      for (int j = 0; j < 5; j++)
        V5(f, i, k) -= (M5(c, i, k, j) * V5(f, i + 1, j) + M5(a, i, k, j) * V5(f, N - 1, j));
}

// --------------------------------------------------------------------------
// Solver state: SOLVER + MESHBLOCK + SOLNBLOCK fields on the hot path.
// --------------------------------------------------------------------------
struct Oracle {
  int nCell = 0, nFace = 0, nChain = 0, nMaxChain = 0;
  std::vector<int> faces;       // 6*nFace [n1,n2,cL,eL,cR,eR]
  std::vector<int> fstart;      // nChain+1
  std::vector<int> chainConn;
  std::vector<double> normVec;  // 2*nFace
  std::vector<double> refMtx;   // 4*nFace
  int order = 1, msweep = 1, nGhost = 2;
  double CFL = 1.0, gamma = 1.4, gm1 = 0.4;
  double rinf = 1, uinf = 0, vinf = 0, einf = 0, pinf = 0;
  std::vector<double> q, r, r0, dq, sigma, ff;  // ff: 32 per face (lmat then rmat)
  std::vector<double> qt;                       // rk3 stage-start state (SOLNBLOCK::qt)
  // Orphan faces (an extension for meshes with triangles, which chains cannot
  // cover face by face): interior faces in no chain, in list order.  od_chain
  // [cell]: the chain whose ADI rows take the cell's orphan diagonal blocks
  // (the last chain listing the cell as a row; -1 none); od_off/od_ent: per
  // cell, its orphan faces in list order, code = face << 1 | (cell is left).
  std::vector<int> orphans, od_chain, od_off, od_ent;
  double L2 = 0, Linf = 0;       // printed norm (overwritten by computeLinearRHS)
  double L2rhs = 0, Linfrhs = 0; // nonlinear norm of the last computeRHS
  double L2lin = 0, Linflin = 0; // linear norm of the last computeLinearRHS
  int nanFlag = 0;
  // per-chain work rows (solver.cpp:41-79), nMaxChain+5 rows
  std::vector<int> cindx;
  std::vector<double> fw, df, ql, qr, dql, dqr, f2, A, B, C, F;
};

inline int FACE(const Oracle& o, int iface, int k) { return o.faces[6 * iface + k]; }

// Chain gather with periodic wrap or mirrored ghosts (computeRHS.cpp:57-149,
// identical in computeLinearRHS.cpp:55-143).  Returns chainSize.
int gather_cindx(Oracle& o, int i) {
  int f1 = o.fstart[i], f2 = o.fstart[i + 1];
  int m = o.nGhost;
  int* cindx = o.cindx.data();
  for (int f = f1; f < f2; f++) {
    int iface = o.chainConn[f];
    cindx[m] = FACE(o, iface, 2);
    m++;
  }
  int chainSize;
  if (o.chainConn[f1] == o.chainConn[f2 - 1]) {
    int f = f1 + 1;
    int iface = o.chainConn[f];
    cindx[m] = FACE(o, iface, 2);
    m++;
    chainSize = m;
    m = 0;
    for (f = f2 - o.nGhost - 1; f < f2 - 1; f++) {
      iface = o.chainConn[f];
      cindx[m] = FACE(o, iface, 2);
      m++;
    }
  } else {
    if (o.order == 5) {
      m--;
      cindx[m] = -cindx[m];
      m++;
      cindx[m] = -cindx[m - 3];
      m++;
      cindx[m] = -cindx[m - 5];
      chainSize = m + 1;
      cindx[0] = -cindx[5];
      cindx[1] = -cindx[4];
      cindx[2] = -cindx[3];
    } else {
      m--;
      cindx[m] = -cindx[m];
      m++;
      cindx[m] = -cindx[m - 3];
      chainSize = m + 1;
      cindx[0] = -cindx[3];
      cindx[1] = -cindx[2];
    }
  }
  return chainSize;
}

// Fill the chain rows f (and df when dqsrc != nullptr) from q / dq with the
// wall-reflected or freestream ghosts (computeRHS.cpp:154-260,
// computeLinearRHS.cpp:148-276), including the cell-0 special cases.
void fill_rows(Oracle& o, int i, int chainSize, const double* dqsrc) {
  int f1 = o.fstart[i], f2 = o.fstart[i + 1];
  for (int j = 0; j < chainSize; j++) {
    int icell = o.cindx[j];
    if (icell >= 0 || (icell == 0 && j == o.nGhost) || (icell == 0 && j == chainSize - o.nGhost - 1)) {
      for (int k = 0; k < NQ; k++) {
        o.fw[j * NQ + k] = o.q[NQ * icell + k];
        if (dqsrc) o.df[j * NQ + k] = dqsrc[NQ * icell + k];
      }
    }
    if (icell < 0 || (icell == 0 && j == o.nGhost - 1) || (icell == 0 && j == chainSize - o.nGhost)) {
      int iface = (j < o.nGhost) ? o.chainConn[f1] : o.chainConn[f2 - 1];
      int rightCell = FACE(o, iface, 4);
      const double* R = &o.refMtx[4 * iface];
      if (rightCell == -2) {
        icell = -icell;
        const double* cv = &o.q[NQ * icell];
        o.fw[j * NQ + 0] = cv[0];
        o.fw[j * NQ + 1] = (cv[1] * R[0] + cv[2] * R[1]);
        o.fw[j * NQ + 2] = (cv[1] * R[2] + cv[2] * R[3]);
        o.fw[j * NQ + 3] = cv[3];
        if (dqsrc) {
          const double* dv = &dqsrc[NQ * icell];
          o.df[j * NQ + 0] = dv[0];
          o.df[j * NQ + 1] = (dv[1] * R[0] + dv[2] * R[1]);
          o.df[j * NQ + 2] = (dv[1] * R[2] + dv[2] * R[3]);
          o.df[j * NQ + 3] = dv[3];
        }
      } else if (rightCell == -5) {
        // interdomain halo: empty in the reference (computeRHS.cpp:233-238)
      } else {
        o.fw[j * NQ + 0] = o.rinf;
        o.fw[j * NQ + 1] = o.rinf * o.uinf;
        o.fw[j * NQ + 2] = o.rinf * o.vinf;
        o.fw[j * NQ + 3] = o.einf;
        if (dqsrc)
          for (int k = 0; k < NQ; k++) o.df[j * NQ + k] = 0.0;
      }
    }
  }
}

void norms(const Oracle& o, double* L2, double* Linf) {
  // computeRHS.cpp:415-425 / computeLinearRHS.cpp:354-368
  double li = 0, l2 = 0;
  for (int i = 0; i < o.nCell; i++) {
    double t = std::fabs(o.r[NQ * i]);
    if (li < t) li = t;
    l2 += o.r[NQ * i] * o.r[NQ * i];
  }
  *L2 = std::sqrt(l2) / o.nCell;
  *Linf = li;
}

void compute_rhs(Oracle& o, bool implicit) {
  // computeRHS.cpp:17-427
  for (int i = 0; i < NQ * o.nCell; i++) o.r[i] = 0.0;
  for (int i = 0; i < o.nCell; i++) o.sigma[i] = 0.0;
  for (int i = 0; i < o.nChain; i++) {
    int f1 = o.fstart[i], f2 = o.fstart[i + 1];
    int chainSize = gather_cindx(o, i);
    fill_rows(o, i, chainSize, nullptr);
    int is = o.nGhost - 1, ie = chainSize - 1;
    double th = THIRD, qt = 0.25, eps = 1e-10;
    if (o.order == 1) qt = 0.0;
    if (o.order == 1 || o.order == 3)
      muscl(o.fw.data(), o.ql.data(), o.qr.data(), o.f2.data(), is, ie, th, qt, eps, chainSize, NQ);
    if (o.order == 5) weno(o.fw.data(), o.ql.data(), o.qr.data(), is, ie, eps, chainSize, NQ);
    int n = is;
    int iPeriodic = (o.chainConn[f1] == o.chainConn[f2 - 1]);
    for (int f = f1; f < f2 - iPeriodic; f++) {
      int iface = o.chainConn[f];
      int leftCell = FACE(o, iface, 2), rightCell = FACE(o, iface, 4);
      double ls[NQ], rs[NQ], flux[NQ], sr;
      for (int m = 0; m < NQ; m++) {
        if (f == f2 - iPeriodic - 1 && iPeriodic == 0) {
          ls[m] = o.ql[n * NQ + m];
          rs[m] = o.qr[(n + 1) * NQ + m];
        } else {
          ls[m] = o.qr[(n + 1) * NQ + m];
          rs[m] = o.ql[n * NQ + m];
        }
      }
      double ds[2] = {o.normVec[2 * iface], o.normVec[2 * iface + 1]};
      o.nanFlag |= flux_roe2d(ds, ls, rs, flux, &sr, o.gamma);
      if (implicit) jacobian_roe2d(ds, ls, rs, &o.ff[32 * iface], &o.ff[32 * iface + 16], o.gamma);
      int m = NQ * leftCell;
      for (int j = 0; j < NQ; j++) o.r[m + j] -= flux[j];
      o.sigma[leftCell] += sr;
      if (rightCell > -1) {
        m = NQ * rightCell;
        for (int j = 0; j < NQ; j++) o.r[m + j] += flux[j];
        o.sigma[rightCell] += sr;
      }
      n++;
    }
  }
  // orphan faces: first-order Roe flux of the two cell states, summed after
  // every chain's contributions (extension; no orphans = the reference)
  for (int iface : o.orphans) {
    int leftCell = FACE(o, iface, 2), rightCell = FACE(o, iface, 4);
    double flux[NQ], sr;
    double ds[2] = {o.normVec[2 * iface], o.normVec[2 * iface + 1]};
    o.nanFlag |= flux_roe2d(ds, &o.q[NQ * leftCell], &o.q[NQ * rightCell], flux, &sr, o.gamma);
    if (implicit)
      jacobian_roe2d(ds, &o.q[NQ * leftCell], &o.q[NQ * rightCell], &o.ff[32 * iface], &o.ff[32 * iface + 16],
                     o.gamma);
    for (int j = 0; j < NQ; j++) o.r[NQ * leftCell + j] -= flux[j];
    o.sigma[leftCell] += sr;
    for (int j = 0; j < NQ; j++) o.r[NQ * rightCell + j] += flux[j];
    o.sigma[rightCell] += sr;
  }
  norms(o, &o.L2, &o.Linf);
  o.L2rhs = o.L2;
  o.Linfrhs = o.Linf;
}

void compute_linear_rhs(Oracle& o) {
  // computeLinearRHS.cpp:13-371
  for (int i = 0; i < NQ * o.nCell; i++) o.r[i] = (o.r0[i] - o.dq[i]);
  for (int i = 0; i < o.nChain; i++) {
    int f1 = o.fstart[i], f2 = o.fstart[i + 1];
    int chainSize = gather_cindx(o, i);
    fill_rows(o, i, chainSize, o.dq.data());
    int is = o.nGhost - 1, ie = chainSize - 1;
    double th = THIRD, qt = 0.25, eps = 1e-10;
    if (o.order == 1) qt = 0.0;
    if (o.order == 1 || o.order == 3)
      muscl_deriv(o.fw.data(), o.dql.data(), o.dqr.data(), o.f2.data(), o.df.data(), is, ie, th, qt, eps,
                  chainSize, NQ);
    if (o.order == 5) weno_deriv(o.fw.data(), o.dql.data(), o.dqr.data(), o.df.data(), is, ie, eps, chainSize, NQ);
    int n = is;
    int iPeriodic = (o.chainConn[f1] == o.chainConn[f2 - 1]);
    for (int f = f1; f < f2 - iPeriodic; f++) {
      int iface = o.chainConn[f];
      int leftCell = FACE(o, iface, 2), rightCell = FACE(o, iface, 4);
      double dl[NQ], dr[NQ], flux[NQ];
      for (int m = 0; m < NQ; m++) {
        if (f == f2 - iPeriodic - 1 && iPeriodic == 0) {
          dl[m] = o.dql[n * NQ + m];
          dr[m] = o.dqr[(n + 1) * NQ + m];
        } else {
          dl[m] = o.dqr[(n + 1) * NQ + m];
          dr[m] = o.dql[n * NQ + m];
        }
      }
      const double* L = &o.ff[32 * iface];
      const double* R = &o.ff[32 * iface + 16];
      for (int j = 0; j < NQ; j++) {
        flux[j] = 0;
        for (int k = 0; k < NQ; k++) flux[j] += ((L[4 * j + k] * dl[k]) + (R[4 * j + k] * dr[k]));
      }
      int m = NQ * leftCell;
      double dtfac = o.CFL / o.sigma[leftCell];
      for (int j = 0; j < NQ; j++) o.r[m + j] -= (flux[j] * dtfac);
      if (rightCell > -1) {
        m = NQ * rightCell;
        dtfac = o.CFL / o.sigma[rightCell];
        for (int j = 0; j < NQ; j++) o.r[m + j] += (flux[j] * dtfac);
      }
      n++;
    }
  }
  // orphan faces: the full linearised flux lmat dq_L + rmat dq_R (extension)
  for (int iface : o.orphans) {
    int leftCell = FACE(o, iface, 2), rightCell = FACE(o, iface, 4);
    const double* L = &o.ff[32 * iface];
    const double* R = &o.ff[32 * iface + 16];
    const double* dl = &o.dq[NQ * leftCell];
    const double* dr = &o.dq[NQ * rightCell];
    double flux[NQ];
    for (int j = 0; j < NQ; j++) {
      flux[j] = 0;
      for (int k = 0; k < NQ; k++) flux[j] += ((L[4 * j + k] * dl[k]) + (R[4 * j + k] * dr[k]));
    }
    double dtfac = o.CFL / o.sigma[leftCell];
    for (int j = 0; j < NQ; j++) o.r[NQ * leftCell + j] -= (flux[j] * dtfac);
    dtfac = o.CFL / o.sigma[rightCell];
    for (int j = 0; j < NQ; j++) o.r[NQ * rightCell + j] += (flux[j] * dtfac);
  }
  norms(o, &o.L2, &o.Linf);
  o.L2lin = o.L2;
  o.Linflin = o.Linf;
}

void adi(Oracle& o) {
  // ADI.cpp:15-266
  for (int i = 0; i < o.nCell; i++) {
    double dtfac = o.CFL / o.sigma[i];
    for (int m = 0; m < NQ; m++) {
      o.r[NQ * i + m] *= dtfac;
      o.r0[NQ * i + m] = o.r[NQ * i + m];
      o.dq[NQ * i + m] = 0.0;
    }
  }
  int worksize = o.nMaxChain + 5;
  double* A = o.A.data();
  double* B = o.B.data();
  double* C = o.C.data();
  double* F = o.F.data();
  for (int isweep = 0; isweep < o.msweep; isweep++) {
    for (int i = 0; i < o.nChain; i++) {
      std::memset(A, 0, sizeof(double) * 16 * worksize);
      std::memset(B, 0, sizeof(double) * 16 * worksize);
      std::memset(C, 0, sizeof(double) * 16 * worksize);
      std::memset(F, 0, sizeof(double) * 4 * worksize);
      int f1 = o.fstart[i], f2 = o.fstart[i + 1];
      int iPeriodic = (o.chainConn[f1] == o.chainConn[f2 - 1]);
      int m = 0;
      int chainSize = (f2 - iPeriodic - f1);
      for (int f = f1; f < f2 - iPeriodic; f++) {
        int iface = o.chainConn[f];
        int leftCell = FACE(o, iface, 2), rightCell = FACE(o, iface, 4);
        double lmat[4][4], rmat[4][4];
        for (int j = 0; j < 4; j++)
          for (int k = 0; k < 4; k++) {
            lmat[j][k] = o.ff[32 * iface + 4 * j + k];
            rmat[j][k] = o.ff[32 * iface + 16 + 4 * j + k];
          }
        if (iPeriodic == 1) {
          int mm1 = (m == 0) ? chainSize - 1 : m - 1;
          for (int j = 0; j < 4; j++) {
            for (int k = 0; k < 4; k++) {
              M3(B, m, j, k) += (lmat[j][k]);
              M3(B, mm1, j, k) -= (rmat[j][k]);
              M3(A, m, j, k) += (rmat[j][k]);
              M3(C, mm1, j, k) -= (lmat[j][k]);
            }
            V2(F, m, j) = o.r[NQ * leftCell + j];
          }
        } else {
          int mm1 = (m - 1);
          if (rightCell == -2) {
            const double* R = &o.refMtx[4 * iface];
            double a1 = R[0], b1 = R[1], a2 = R[2], b2 = R[3];
            for (int j = 0; j < 4; j++) {
              double r01 = rmat[j][1], r02 = rmat[j][2];
              rmat[j][1] = a1 * r01 + b1 * r02;
              rmat[j][2] = a2 * r01 + b2 * r02;
            }
          }
          if (rightCell < 0 && iPeriodic == 0 && f == f2 - 1) m--;
          for (int j = 0; j < 4; j++) {
            for (int k = 0; k < 4; k++) {
              M3(B, m, j, k) += (lmat[j][k]);
              if (mm1 > -1 && rightCell > -1) {
                M3(A, m, j, k) += (rmat[j][k]);
                M3(B, mm1, j, k) -= (rmat[j][k]);
                M3(C, mm1, j, k) -= (lmat[j][k]);
              } else {
                if (rightCell == -2) M3(B, m, j, k) += (rmat[j][k]);
              }
            }
            V2(F, m, j) = o.r[NQ * leftCell + j];
          }
        }
        m++;
      }
      m = 0;
      chainSize = chainSize - (iPeriodic == 0);
      // orphan diagonal blocks of the rows whose cell designates this chain
      // (extension): + lmat where the cell is the face's left, - rmat where right
      if (!o.orphans.empty()) {
        for (int f = f1; f < f2 - 1; f++, m++) {
          int cell = FACE(o, o.chainConn[f], 2);
          if (o.od_chain[cell] != i) continue;
          for (int e = o.od_off[cell]; e < o.od_off[cell + 1]; e++) {
            int iface = o.od_ent[e] >> 1;
            bool left = o.od_ent[e] & 1;
            const double* M = &o.ff[32 * iface + (left ? 0 : 16)];
            for (int j = 0; j < 4; j++)
              for (int k = 0; k < 4; k++) {
                if (left)
                  M3(B, m, j, k) += M[4 * j + k];
                else
                  M3(B, m, j, k) -= M[4 * j + k];
              }
          }
        }
        m = 0;
      }
      for (int f = f1; f < f2 - 1; f++) {
        int iface = o.chainConn[f];
        int leftCell = FACE(o, iface, 2);
        double dtfac = o.CFL / o.sigma[leftCell];
        for (int j = 0; j < 4; j++)
          for (int k = 0; k < 4; k++) {
            M3(A, m, j, k) *= dtfac;
            M3(B, m, j, k) *= dtfac;
            M3(C, m, j, k) *= dtfac;
          }
        for (int j = 0; j < 4; j++) M3(B, m, j, j) += 1.0;
        m++;
      }
      if (iPeriodic == 1) block_tridag_periodic4(A, B, C, F, chainSize);
      if (iPeriodic == 0) block_tridag4(A, B, C, F, chainSize);
      m = 0;
      for (int f = f1; f < f2 - 1; f++) {
        int iface = o.chainConn[f];
        int leftCell = FACE(o, iface, 2);
        for (int j = 0; j < 4; j++) o.r[NQ * leftCell + j] = V2(F, m, j);
        m++;
      }
    }
    for (int i = 0; i < o.nCell * NQ; i++) o.dq[i] += o.r[i];
    compute_linear_rhs(o);
  }
  for (int i = 0; i < o.nCell * NQ; i++) o.q[i] += o.dq[i];
}

void init_flow(Oracle& o, double mach, double alpha) {
  // initFlow.cpp:12-93 (test == 0, 2D)
  o.gamma = 1.4;
  o.gm1 = o.gamma - 1.;
  o.rinf = 1.0;
  o.pinf = 1.0 / o.gamma;
  o.nGhost = (o.order == 5) ? 3 : 2;
  o.uinf = mach * std::cos(alpha * (PI / 180.0));
  o.vinf = mach * std::sin(alpha * (PI / 180.0));
  o.einf = o.pinf / (o.gm1) + 0.5 * o.rinf * (o.uinf * o.uinf + o.vinf * o.vinf);
  for (int i = 0; i < o.nCell; i++) {
    o.q[NQ * i + 0] = o.rinf;
    o.q[NQ * i + 1] = o.rinf * o.uinf;
    o.q[NQ * i + 2] = o.rinf * o.vinf;
    o.q[NQ * i + 3] = o.einf;
  }
  o.qt = o.q;  // initFlow.cpp:71-88 sets qt alongside q
}

// SOLVER::updateSolution (march.cpp:142-164): qnew = qold + (coef*CFL/sigma) r.
// qold and qnew may be the same array (each element is read before it is written).
void update_solution(Oracle& o, const double* qold, double* qnew, double coef) {
  int m = 0;
  for (int i = 0; i < o.nCell; i++) {
    for (int j = 0; j < NQ; j++) {
      double dtfac = coef * o.CFL / o.sigma[i];
      qnew[m] = qold[m] + dtfac * o.r[m];
      m++;
    }
  }
}

// SOLVER::stepSolution (march.cpp:64-136) for stepType "ADI" (0), "euler" (1)
// and "rk3" (2).  *first gets the nonlinear norm of the step's first
// computeRHS (the residual of the state the step starts from).  The rk3 branch
// keeps the reference's stage coefficients and its qt handling: qt is written
// in stage 1 only, stages 2 and 3 both restart from it.
void step_solution(Oracle& o, int scheme, double* first, double* first_inf) {
  if (scheme == 0) {
    compute_rhs(o, true);
    *first = o.L2rhs;
    *first_inf = o.Linfrhs;
    adi(o);
    return;
  }
  compute_rhs(o, false);
  *first = o.L2rhs;
  *first_inf = o.Linfrhs;
  if (scheme == 1) {
    update_solution(o, o.q.data(), o.q.data(), 1.);
    return;
  }
  update_solution(o, o.q.data(), o.qt.data(), 0.25);
  update_solution(o, o.q.data(), o.q.data(), 8.0 / 15);
  compute_rhs(o, false);
  update_solution(o, o.qt.data(), o.q.data(), 5.0 / 12);
  compute_rhs(o, false);
  update_solution(o, o.qt.data(), o.q.data(), 3.0 / 4.0);
}

}  // namespace

extern "C" {

// ---- geometry set-up (preprocess.cpp:56-193) -------------------------------
void orc_preprocess(int nNode, const double* x, int nFace, int* faces, double* normVec, double* refMtx,
                    int sol_visc) {
  (void)nNode;
  for (int i = 0; i < nFace; i++) {
    int i1 = faces[6 * i], i2 = faces[6 * i + 1];
    double x1 = x[2 * i1], y1 = x[2 * i1 + 1], x2 = x[2 * i2], y2 = x[2 * i2 + 1];
    normVec[2 * i] = (y2 - y1);
    normVec[2 * i + 1] = -(x2 - x1);
    double dsnorm = normVec[2 * i] * normVec[2 * i] + normVec[2 * i + 1] * normVec[2 * i + 1];
    dsnorm = 1. / dsnorm;
    refMtx[4 * i + 0] = 1. - 2. * normVec[2 * i] * normVec[2 * i] * dsnorm;
    refMtx[4 * i + 1] = -2. * normVec[2 * i] * normVec[2 * i + 1] * dsnorm;
    refMtx[4 * i + 2] = -2. * normVec[2 * i + 1] * normVec[2 * i] * dsnorm;
    refMtx[4 * i + 3] = 1. - 2. * normVec[2 * i + 1] * normVec[2 * i + 1] * dsnorm;
    if (faces[6 * i + 4] == -1) {
      double r1 = std::sqrt(x[2 * i1] * x[2 * i1] + x[2 * i1 + 1] * x[2 * i1 + 1]);
      double r2 = std::sqrt(x[2 * i2] * x[2 * i2] + x[2 * i2 + 1] * x[2 * i2 + 1]);
      if (0.5 * (r1 + r2) < 3.0) faces[6 * i + 4] = -(sol_visc + 2);
    }
  }
}

// ---- unit functions ---------------------------------------------------------
int orc_flux_roe2d(const double* n, const double* ql, const double* qr, double* flux, double* sr,
                   double gamma) {
  return flux_roe2d(n, ql, qr, flux, sr, gamma);
}
void orc_jacobian_roe2d(const double* n, const double* ql, const double* qr, double* lmat, double* rmat,
                        double gamma) {
  jacobian_roe2d(n, ql, qr, lmat, rmat, gamma);
}
void orc_muscl(const double* f, double* ql, double* qr, int is, int ie, double th, double qt, double eps,
               int imax) {
  std::vector<double> f2(4 * (imax + 1), 0.0);
  muscl(f, ql, qr, f2.data(), is, ie, th, qt, eps, imax, NQ);
}
void orc_weno(const double* f, double* ql, double* qr, int is, int ie, double eps, int imax) {
  weno(f, ql, qr, is, ie, eps, imax, NQ);
}
void orc_muscl_deriv(const double* f, const double* dq, double* ql, double* qr, int is, int ie, double th,
                     double qt, double eps, int imax) {
  std::vector<double> f2(4 * (imax + 1), 0.0);
  muscl_deriv(f, ql, qr, f2.data(), dq, is, ie, th, qt, eps, imax, NQ);
}
void orc_weno_deriv(const double* f, const double* dq, double* ql, double* qr, int is, int ie, double eps,
                    int imax) {
  weno_deriv(f, ql, qr, dq, is, ie, eps, imax, NQ);
}
void orc_block_tridag4(double* a, double* b, double* c, double* f, int N) { block_tridag4(a, b, c, f, N); }
void orc_flux_roe3d(const double* n, const double* ql, const double* qr, double* flux, double* sr, double gamma) {
  flux_roe3d(n, ql, qr, flux, sr, gamma);
}
void orc_jacobian_roe3d(const double* n, const double* ql, const double* qr, double* lmat, double* rmat,
                        double gamma) {
  double L[5][5], R[5][5];
  jacobian_roe3d(n, ql, qr, L, R, gamma);
  std::memcpy(lmat, L, sizeof L);
  std::memcpy(rmat, R, sizeof R);
}
void orc_block_tridag5(double* a, double* b, double* c, double* f, int N) { block_tridag5(a, b, c, f, N); }
void orc_block_tridag_periodic5(double* a, double* b, double* c, double* f, int N) {
  block_tridag_periodic5(a, b, c, f, N);
}
void orc_block_tridag_periodic4(double* a, double* b, double* c, double* f, int N) {
  block_tridag_periodic4(a, b, c, f, N);
}

// ---- solver -----------------------------------------------------------------
void* orc_create(int nCell, int nFace, int nChain, const int* faces, const int* fstart, const int* chainConn,
                 const double* normVec, const double* refMtx, double mach, double alpha, int order, double cfl,
                 int msweep, int nOrphan, const int* orphans) {
  Oracle* o = new Oracle();
  o->nCell = nCell;
  o->nFace = nFace;
  o->nChain = nChain;
  o->faces.assign(faces, faces + 6 * nFace);
  o->fstart.assign(fstart, fstart + nChain + 1);
  o->chainConn.assign(chainConn, chainConn + fstart[nChain]);
  o->normVec.assign(normVec, normVec + 2 * nFace);
  o->refMtx.assign(refMtx, refMtx + 4 * nFace);
  o->order = order;
  o->CFL = cfl;
  o->msweep = msweep;
  o->nMaxChain = 0;
  for (int i = 0; i < nChain; i++) {
    int d = fstart[i + 1] - fstart[i];
    if (d > o->nMaxChain) o->nMaxChain = d;
  }
  int ws = o->nMaxChain + 5;
  o->cindx.assign(ws, 0);
  for (auto* v : {&o->fw, &o->df, &o->ql, &o->qr, &o->dql, &o->dqr, &o->f2, &o->F}) v->assign(4 * ws, 0.0);
  for (auto* v : {&o->A, &o->B, &o->C}) v->assign(16 * ws, 0.0);
  o->q.assign(4 * nCell, 0.0);
  o->r.assign(4 * nCell, 0.0);
  o->r0.assign(4 * nCell, 0.0);
  o->dq.assign(4 * nCell, 0.0);
  o->sigma.assign(nCell, 0.0);
  o->ff.assign(32 * (size_t)nFace, 0.0);
  o->qt.assign(4 * nCell, 0.0);
  if (nOrphan > 0) {
    o->orphans.assign(orphans, orphans + nOrphan);
    o->od_chain.assign(nCell, -1);
    for (int i = 0; i < nChain; i++)
      for (int f = fstart[i]; f < fstart[i + 1] - 1; f++) o->od_chain[faces[6 * chainConn[f] + 2]] = i;
    std::vector<int> cnt(nCell + 1, 0);
    for (int f : o->orphans) {
      cnt[faces[6 * f + 2] + 1]++;
      cnt[faces[6 * f + 4] + 1]++;
    }
    for (int c = 0; c < nCell; c++) cnt[c + 1] += cnt[c];
    o->od_off = cnt;
    o->od_ent.assign(cnt[nCell], 0);
    for (int f : o->orphans) {
      o->od_ent[cnt[faces[6 * f + 2]]++] = (f << 1) | 1;
      o->od_ent[cnt[faces[6 * f + 4]]++] = (f << 1);
    }
  }
  init_flow(*o, mach, alpha);
  return o;
}
void orc_destroy(void* h) { delete static_cast<Oracle*>(h); }

// field ids: 0 q, 1 r, 2 r0, 3 dq, 4 sigma, 5 ff, 6 qt
static std::vector<double>* field(Oracle* o, int id) {
  switch (id) {
    case 0: return &o->q;
    case 1: return &o->r;
    case 2: return &o->r0;
    case 3: return &o->dq;
    case 4: return &o->sigma;
    case 5: return &o->ff;
    case 6: return &o->qt;
    default: return nullptr;
  }
}
int orc_get(void* h, int id, double* out) {
  auto* v = field(static_cast<Oracle*>(h), id);
  if (!v) return -1;
  std::memcpy(out, v->data(), sizeof(double) * v->size());
  return 0;
}
int orc_set(void* h, int id, const double* in) {
  auto* v = field(static_cast<Oracle*>(h), id);
  if (!v) return -1;
  std::memcpy(v->data(), in, sizeof(double) * v->size());
  return 0;
}
void orc_freestream(void* h, double* out4) {
  Oracle* o = static_cast<Oracle*>(h);
  out4[0] = o->rinf; out4[1] = o->uinf; out4[2] = o->vinf; out4[3] = o->einf;
}
int orc_compute_rhs(void* h, int implicit, double* L2, double* Linf) {
  Oracle* o = static_cast<Oracle*>(h);
  compute_rhs(*o, implicit != 0);
  if (L2) *L2 = o->L2rhs;
  if (Linf) *Linf = o->Linfrhs;
  return o->nanFlag;
}
int orc_adi(void* h, double* L2lin, double* Linflin) {
  Oracle* o = static_cast<Oracle*>(h);
  adi(*o);
  if (L2lin) *L2lin = o->L2lin;
  if (Linflin) *Linflin = o->Linflin;
  return 0;
}
int orc_compute_linear_rhs(void* h, double* L2lin, double* Linflin) {
  Oracle* o = static_cast<Oracle*>(h);
  compute_linear_rhs(*o);
  if (L2lin) *L2lin = o->L2lin;
  if (Linflin) *Linflin = o->Linflin;
  return 0;
}
// n steps of stepSolution's ADI branch (march.cpp:112-117); per-step norms out
// (nonlinear L2 of computeRHS, printed linear L2) when the arrays are given.
// Returns the wall time of the loop in seconds through *seconds.
int orc_step(void* h, int nsteps, double* l2rhs, double* l2lin, double* seconds) {
  Oracle* o = static_cast<Oracle*>(h);
  auto t0 = std::chrono::steady_clock::now();
  for (int n = 0; n < nsteps; n++) {
    compute_rhs(*o, true);
    adi(*o);
    if (l2rhs) l2rhs[n] = o->L2rhs;
    if (l2lin) l2lin[n] = o->L2lin;
  }
  auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  return o->nanFlag;
}

// SOLVER::updateSolution between two per-cell fields (0 q, 6 qt) with the
// current r and sigma
int orc_update_solution(void* h, int src, int dst, double coef) {
  Oracle* o = static_cast<Oracle*>(h);
  if ((src != 0 && src != 6) || (dst != 0 && dst != 6)) return -1;
  update_solution(*o, field(o, src)->data(), field(o, dst)->data(), coef);
  return 0;
}

// n x SOLVER::stepSolution with stepType scheme (0 ADI, 1 euler, 2 rk3).
// first[n]: nonlinear L2 of the step's first computeRHS; printed[n]: the norm
// marchInTime prints after the step (sb->L2Norm: the last computeLinearRHS for
// ADI, the last computeRHS for euler and rk3).
int orc_step_scheme(void* h, int scheme, int nsteps, double* first, double* printed, double* seconds) {
  Oracle* o = static_cast<Oracle*>(h);
  if (scheme < 0 || scheme > 2) return -1;
  auto t0 = std::chrono::steady_clock::now();
  for (int n = 0; n < nsteps; n++) {
    double f = 0, fi = 0;
    step_solution(*o, scheme, &f, &fi);
    if (first) first[n] = f;
    if (printed) printed[n] = (scheme == 0) ? o->L2lin : o->L2rhs;
  }
  auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  return o->nanFlag;
}

}  // extern "C"

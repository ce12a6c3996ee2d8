// ============================================================================
// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A C-ABI driver around the reference's own classes and free functions.  The
// Makefile in this directory compiles it together with the reference's C++
// sources (read where they lie, never copied) into oracle/_ref/, with
// -Dprivate=public so the harness can read MESHBLOCK/SOLNBLOCK fields.  It is
// used only to write golden vectors (tests/golden) and to cross-check the
// oracle; nothing in the GPU product links it.
//
// Call surface used: MESHBLOCK::readMeshData / setViscosity / preprocess,
// SOLNBLOCK::readSolutionInputs, SOLVER::{SOLVER, initFlow, computeRHS, ADI,
// computeLinearRHS, stepSolution, updateSolution, fluxRoe2D, jacobianRoe2D, fluxRoe3D, jacobianRoe3D},
// RECONSTRUCTION::{MUSCL, WENO, MUSCL_DERIV, WENO_DERIV}, TRISOLVER::{blockTridag4, blockTridagPeriodic4,
// blockTridag, blockTridagPeriodic}.
// ============================================================================
#include <unistd.h>
#include <chrono>
#include <cstring>
#include <vector>

#include "hamstr.h"
#include "reconstruction.h"
#include "solver.h"
#include "triSolvers.h"

namespace {
struct RefCase {
  MESHBLOCK* mb;
  SOLNBLOCK* sb;
  SOLVER* solv;
  double L2rhs, Linfrhs, L2lin, Linflin;
};

// A SOLVER whose methods fluxRoe2D/jacobianRoe2D read no members; built on a
// one-cell block so its constructor's allocations are tiny.
SOLVER* unit_solver() {
  static SOLVER* s = nullptr;
  if (!s) {
    MESHBLOCK* mb = new MESHBLOCK[1];
    SOLNBLOCK* sb = new SOLNBLOCK[1];
    mb->nCell = 1;
    mb->nFace = 1;
    mb->nMaxChain = 1;
    std::strcpy(sb->scheme, "ADI");
    s = new SOLVER(mb, sb);
  }
  return s;
}

std::vector<double*> rows(double* base, int n) {
  std::vector<double*> p(n);
  for (int i = 0; i < n; i++) p[i] = base + 4 * i;
  return p;
}
}  // namespace

extern "C" {

void* ref_load(const char* dir) {
  char cwd[4096];
  if (!getcwd(cwd, sizeof(cwd))) return nullptr;
  if (chdir(dir) != 0) return nullptr;
  RefCase* c = new RefCase();
  c->mb = new MESHBLOCK[1];
  c->sb = new SOLNBLOCK[1];
  c->mb->readMeshData(0);
  c->sb->readSolutionInputs();
  c->solv = new SOLVER(c->mb, c->sb);
  c->mb->setViscosity(0);   // sol_visc is otherwise never set (wall tag -(visc+2))
  c->mb->preprocess();
  c->solv->initFlow();
  if (chdir(cwd) != 0) return nullptr;
  return c;
}

void ref_counts(void* h, int* out) {
  RefCase* c = static_cast<RefCase*>(h);
  out[0] = c->mb->nCell;
  out[1] = c->mb->nFace;
  out[2] = c->mb->nChain;
  out[3] = c->mb->faceStartPerChain[c->mb->nChain];
  out[4] = c->mb->nMaxChain;
  out[5] = c->mb->nNode;
  out[6] = c->sb->order;
  out[7] = c->sb->msweep;
  out[8] = c->sb->nsteps;
}

void ref_params(void* h, double* out) {
  RefCase* c = static_cast<RefCase*>(h);
  SOLNBLOCK* s = c->sb;
  out[0] = s->Mach; out[1] = s->alpha; out[2] = s->CFL; out[3] = s->gamma;
  out[4] = s->rinf; out[5] = s->uinf; out[6] = s->vinf; out[7] = s->einf;
}

// 0 q, 1 r, 2 r0, 3 dq, 4 sigma, 5 ff (32/face), 6 normVec (2/face), 7 refMtx (4/face), 8 qt
int ref_get(void* h, int id, double* out) {
  RefCase* c = static_cast<RefCase*>(h);
  MESHBLOCK* mb = c->mb;
  SOLNBLOCK* sb = c->sb;
  int n = mb->nCell;
  switch (id) {
    case 0: std::memcpy(out, sb->q, sizeof(double) * 4 * n); return 0;
    case 1: std::memcpy(out, sb->r, sizeof(double) * 4 * n); return 0;
    case 2: std::memcpy(out, sb->r0, sizeof(double) * 4 * n); return 0;
    case 3: std::memcpy(out, sb->dq, sizeof(double) * 4 * n); return 0;
    case 4: std::memcpy(out, sb->sigma, sizeof(double) * n); return 0;
    case 5:
      for (int f = 0; f < mb->nFace; f++) {
        for (int j = 0; j < 4; j++)
          for (int k = 0; k < 4; k++) {
            out[32 * f + 4 * j + k] = sb->ff[f].lmat[j][k];
            out[32 * f + 16 + 4 * j + k] = sb->ff[f].rmat[j][k];
          }
      }
      return 0;
    case 6:
      for (int f = 0; f < mb->nFace; f++) {
        out[2 * f] = mb->normVec[f][0];
        out[2 * f + 1] = mb->normVec[f][1];
      }
      return 0;
    case 7:
      for (int f = 0; f < mb->nFace; f++)
        for (int j = 0; j < 2; j++)
          for (int k = 0; k < 2; k++) out[4 * f + 2 * j + k] = mb->refMtx[f][j][k];
      return 0;
    case 8: std::memcpy(out, sb->qt, sizeof(double) * 4 * n); return 0;
  }
  return -1;
}

// 0 faces (6/face, after reading + preprocess), 1 faceStartPerChain (nChain+1), 2 chainConn
int ref_get_ints(void* h, int id, int* out) {
  RefCase* c = static_cast<RefCase*>(h);
  MESHBLOCK* mb = c->mb;
  switch (id) {
    case 0: std::memcpy(out, mb->faces, sizeof(int) * 6 * mb->nFace); return 0;
    case 1: std::memcpy(out, mb->faceStartPerChain, sizeof(int) * (mb->nChain + 1)); return 0;
    case 2: std::memcpy(out, mb->chainConn, sizeof(int) * mb->faceStartPerChain[mb->nChain]); return 0;
  }
  return -1;
}

int ref_set_q(void* h, const double* q) {
  RefCase* c = static_cast<RefCase*>(h);
  std::memcpy(c->sb->q, q, sizeof(double) * 4 * c->mb->nCell);
  return 0;
}

void ref_compute_rhs(void* h, int implicit, double* L2, double* Linf) {
  RefCase* c = static_cast<RefCase*>(h);
  char imp[] = "implicit", exp_[] = "explicit";
  c->solv->computeRHS(implicit ? imp : exp_);
  *L2 = c->sb->L2Norm;
  *Linf = c->sb->LInfNorm;
}

void ref_adi(void* h, double* L2, double* Linf) {
  RefCase* c = static_cast<RefCase*>(h);
  c->solv->ADI();
  *L2 = c->sb->L2Norm;
  *Linf = c->sb->LInfNorm;
}

void ref_compute_linear_rhs(void* h, double* L2, double* Linf) {
  RefCase* c = static_cast<RefCase*>(h);
  c->solv->computeLinearRHS();
  *L2 = c->sb->L2Norm;
  *Linf = c->sb->LInfNorm;
}

// n x stepSolution's ADI branch (march.cpp:112-117), keeping the nonlinear
// norm of computeRHS that the reference overwrites before printing.
void ref_step(void* h, int nsteps, double* l2rhs, double* l2lin, double* seconds) {
  RefCase* c = static_cast<RefCase*>(h);
  char imp[] = "implicit";
  auto t0 = std::chrono::steady_clock::now();
  for (int n = 0; n < nsteps; n++) {
    c->solv->computeRHS(imp);
    if (l2rhs) l2rhs[n] = c->sb->L2Norm;
    c->solv->ADI();
    if (l2lin) l2lin[n] = c->sb->L2Norm;
  }
  auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
}

// SOLVER::stepType points at SOLNBLOCK::scheme (solver.cpp:22): "ADI",
// "euler" or "rk3" as the input deck's scheme= line would set it
int ref_set_scheme(void* h, const char* name) {
  RefCase* c = static_cast<RefCase*>(h);
  if (std::strlen(name) >= sizeof(c->sb->scheme)) return -1;
  std::strcpy(c->sb->scheme, name);
  return 0;
}

// n x SOLVER::stepSolution (march.cpp:64-136) with the current stepType;
// printed[n] / printed_inf[n]: sb->L2Norm / LInfNorm after the step, the
// values marchInTime prints (march.cpp:45-46)
void ref_step_solution(void* h, int nsteps, double* printed, double* printed_inf, double* seconds) {
  RefCase* c = static_cast<RefCase*>(h);
  auto t0 = std::chrono::steady_clock::now();
  for (int n = 0; n < nsteps; n++) {
    c->solv->stepSolution();
    if (printed) printed[n] = c->sb->L2Norm;
    if (printed_inf) printed_inf[n] = c->sb->LInfNorm;
  }
  auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
}

// SOLVER::updateSolution(qold, qnew, coef) between q (0) and qt (8)
int ref_update_solution(void* h, int src, int dst, double coef) {
  RefCase* c = static_cast<RefCase*>(h);
  auto pick = [&](int id) -> double* { return id == 0 ? c->sb->q : (id == 8 ? c->sb->qt : nullptr); };
  double *a = pick(src), *b = pick(dst);
  if (!a || !b) return -1;
  c->solv->updateSolution(a, b, coef);
  return 0;
}

// ---- unit functions ---------------------------------------------------------
void ref_flux_roe2d(const double* n, const double* ql, const double* qr, double* flux, double* sr,
                    double gamma) {
  double nn[2] = {n[0], n[1]};
  double l[8] = {ql[0], ql[1], ql[2], ql[3], 0, 0, 0, 0};
  double r[8] = {qr[0], qr[1], qr[2], qr[3], 0, 0, 0, 0};
  unit_solver()->fluxRoe2D(nn, l, r, flux, sr, gamma);
}

void ref_jacobian_roe2d(const double* n, const double* ql, const double* qr, double* lmat, double* rmat,
                        double gamma) {
  double nn[2] = {n[0], n[1]};
  double l[4] = {ql[0], ql[1], ql[2], ql[3]};
  double r[4] = {qr[0], qr[1], qr[2], qr[3]};
  double* lp[4] = {lmat, lmat + 4, lmat + 8, lmat + 12};
  double* rp[4] = {rmat, rmat + 4, rmat + 8, rmat + 12};
  unit_solver()->jacobianRoe2D(nn, l, r, lp, rp, gamma);
}

void ref_muscl(double* f, double* ql, double* qr, int is, int ie, double th, double qt, double eps,
               int imax) {
  std::vector<double> f2(4 * (imax + 1), 0.0);
  auto F = rows(f, imax), L = rows(ql, imax), R = rows(qr, imax), F2 = rows(f2.data(), imax + 1);
  RECONSTRUCTION::MUSCL(F.data(), L.data(), R.data(), F2.data(), is, ie, th, qt, eps, imax, 4);
}

void ref_weno(double* f, double* ql, double* qr, int is, int ie, double eps, int imax) {
  auto F = rows(f, imax), L = rows(ql, imax), R = rows(qr, imax);
  RECONSTRUCTION::WENO(F.data(), L.data(), R.data(), is, ie, eps, imax, 4);
}

void ref_muscl_deriv(double* f, double* dq, double* ql, double* qr, int is, int ie, double th, double qt,
                     double eps, int imax) {
  std::vector<double> f2(4 * (imax + 1), 0.0);
  std::vector<double> dqx(4 * (imax + 1), 0.0);
  std::memcpy(dqx.data(), dq, sizeof(double) * 4 * imax);
  auto F = rows(f, imax), L = rows(ql, imax), R = rows(qr, imax), F2 = rows(f2.data(), imax + 1),
       D = rows(dqx.data(), imax + 1);
  RECONSTRUCTION::MUSCL_DERIV(F.data(), L.data(), R.data(), F2.data(), D.data(), is, ie, th, qt, eps, imax, 4);
}

void ref_weno_deriv(double* f, double* dq, double* ql, double* qr, int is, int ie, double eps, int imax) {
  auto F = rows(f, imax), L = rows(ql, imax), R = rows(qr, imax), D = rows(dq, imax);
  RECONSTRUCTION::WENO_DERIV(F.data(), L.data(), R.data(), D.data(), is, ie, eps, imax, 4);
}

static void blocks(double* base, int N, std::vector<double*>& r2, std::vector<double**>& r3) {
  r2.resize(4 * N);
  r3.resize(N);
  for (int i = 0; i < N; i++) {
    for (int k = 0; k < 4; k++) r2[4 * i + k] = base + 16 * i + 4 * k;
    r3[i] = &r2[4 * i];
  }
}

void ref_block_tridag4(double* a, double* b, double* c, double* f, int N) {
  std::vector<double*> a2, b2, c2;
  std::vector<double**> a3, b3, c3;
  blocks(a, N, a2, a3);
  blocks(b, N, b2, b3);
  blocks(c, N, c2, c3);
  auto F = rows(f, N);
  TRISOLVER::blockTridag4(a3.data(), b3.data(), c3.data(), F.data(), N, 4);
}

void ref_block_tridag_periodic4(double* a, double* b, double* c, double* f, int N) {
  std::vector<double*> a2, b2, c2;
  std::vector<double**> a3, b3, c3;
  blocks(a, N, a2, a3);
  blocks(b, N, b2, b3);
  blocks(c, N, c2, c3);
  auto F = rows(f, N);
  TRISOLVER::blockTridagPeriodic4(a3.data(), b3.data(), c3.data(), F.data(), N, 4);
}

// ---- the 3D / strand building blocks (compiled in the reference's 2D build) ----
void ref_flux_roe3d(const double* n, const double* ql, const double* qr, double* flux, double* sr,
                    double gamma) {
  double nn[3] = {n[0], n[1], n[2]};
  double l[5] = {ql[0], ql[1], ql[2], ql[3], ql[4]};
  double r[5] = {qr[0], qr[1], qr[2], qr[3], qr[4]};
  unit_solver()->fluxRoe3D(nn, l, r, flux, sr, gamma);
}

void ref_jacobian_roe3d(const double* n, const double* ql, const double* qr, double* lmat, double* rmat,
                        double gamma) {
  double nn[3] = {n[0], n[1], n[2]};
  double l[5] = {ql[0], ql[1], ql[2], ql[3], ql[4]};
  double r[5] = {qr[0], qr[1], qr[2], qr[3], qr[4]};
  double* lp[5] = {lmat, lmat + 5, lmat + 10, lmat + 15, lmat + 20};
  double* rp[5] = {rmat, rmat + 5, rmat + 10, rmat + 15, rmat + 20};
  unit_solver()->jacobianRoe3D(nn, l, r, lp, rp, gamma);
}

static void blocks5(double* base, int N, std::vector<double*>& r2, std::vector<double**>& r3) {
  r2.resize(5 * N);
  r3.resize(N);
  for (int i = 0; i < N; i++) {
    for (int k = 0; k < 5; k++) r2[5 * i + k] = base + 25 * i + 5 * k;
    r3[i] = &r2[5 * i];
  }
}

static std::vector<double*> rows5(double* base, int n) {
  std::vector<double*> p(n);
  for (int i = 0; i < n; i++) p[i] = base + 5 * i;
  return p;
}

void ref_block_tridag5(double* a, double* b, double* c, double* f, int N) {
  std::vector<double*> a2, b2, c2;
  std::vector<double**> a3, b3, c3;
  blocks5(a, N, a2, a3);
  blocks5(b, N, b2, b3);
  blocks5(c, N, c2, c3);
  auto F = rows5(f, N);
  TRISOLVER::blockTridag(a3.data(), b3.data(), c3.data(), F.data(), N, 5);
}

void ref_block_tridag_periodic5(double* a, double* b, double* c, double* f, int N) {
  std::vector<double*> a2, b2, c2;
  std::vector<double**> a3, b3, c3;
  blocks5(a, N, a2, a3);
  blocks5(b, N, b2, b3);
  blocks5(c, N, c2, c3);
  auto F = rows5(f, N);
  TRISOLVER::blockTridagPeriodic(a3.data(), b3.data(), c3.data(), F.data(), N, 5);
}

}  // extern "C"

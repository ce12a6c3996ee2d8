// hamstr_main.cpp -- drop-in driver: the reference's main/HAMSTR facade call
// sequence (main.cpp:12-55, hamstr.cpp:33-77) on the MI355X solver.  Reads
// input.hamstr and ./input/domain000/*.dat from the working directory, then
// readInputs -> preprocessor -> marchInTime with the reference's iteration
// table.  Options: --lhs32 (FP32 LHS), --exact (the reference's sequential
// block-tridiagonal order), --dump <file> (final q, binary doubles),
// --restart <file> (start from a solution file instead of freestream),
// --solution <file> (where nwrite > 0 in input.hamstr writes the solution
// file every nwrite steps; default solution.bin).
#include <cstdio>
#include <cstring>
#include <vector>

#include "solver_hip.hpp"

int main(int argc, char** argv) {
  int lhs = HAMSTR_LHS_FP64, adi = HAMSTR_ADI_FAST;
  const char *dump = nullptr, *restart = nullptr, *solution = nullptr;
  for (int i = 1; i < argc; i++) {
    if (!std::strcmp(argv[i], "--lhs32")) lhs = HAMSTR_LHS_FP32;
    else if (!std::strcmp(argv[i], "--exact")) adi = HAMSTR_ADI_EXACT;
    else if (!std::strcmp(argv[i], "--dump") && i + 1 < argc) dump = argv[++i];
    else if (!std::strcmp(argv[i], "--restart") && i + 1 < argc) restart = argv[++i];
    else if (!std::strcmp(argv[i], "--solution") && i + 1 < argc) solution = argv[++i];
  }
  std::printf("\n HAMSTR (MI355X): Read input files ... ");
  hamstr::MeshBlock mb;
  hamstr::SolnBlock sb;
  if (!mb.readMeshData(0) || !sb.readSolutionInputs()) return 1;
  hamstr::SOLVER solv(&mb, &sb, lhs, adi);
  std::printf(" [done]\n HAMSTR (MI355X): Preprocess data  ... ");
  mb.setViscosity(0);
  mb.preprocess();
  std::printf(" [done]\n HAMSTR (MI355X): Initialize flow  ... ");
  solv.initFlow();
  if (restart) solv.restart(restart);
  if (solution) solv.setSolutionFile(solution);
  std::printf(" [done]\n cells %d faces %d chains %d max chain %d order %d CFL %g Mach %g alpha %g\n", mb.nCell,
              mb.nFace, mb.nChain, mb.nMaxChain, sb.order, sb.CFL, sb.Mach, sb.alpha);
  std::printf(" HAMSTR (MI355X): Marching in time ... \n");
  solv.marchInTime();
  std::printf(" final nonlinear L2 %.16e\n", solv.nonlinearL2());
  if (dump) {
    std::vector<double> q(4 * (size_t)mb.nCell);
    solv.get(HAMSTR_FIELD_Q, q.data());
    FILE* fp = std::fopen(dump, "wb");
    if (!fp || std::fwrite(q.data(), sizeof(double), q.size(), fp) != q.size()) return 1;
    std::fclose(fp);
  }
  return 0;
}

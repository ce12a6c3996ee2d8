// hamstr_hip.hip -- MI355X (gfx950) implementation of HAMSTR_OOP's
// line-implicit steady Euler iteration behind the C ABI of include/hamstr_hip.h.
//
// Device data layout (per context):
//   q, r, r0, dq      double [nCell][4]   AoS, the reference's layout
//   sigma             double [nCell]
//   per processed face p (chain order, closed chains without their repeated
//   first face): flux [p][4], sr [p], lflux [p][4], normal [p][2],
//   ff  T [32][nProc]  element-major (lmat row-major then rmat), T = double or
//                      float (FP32-LHS)
//   chain slot table  int2 [nSlot]: the source of every row the reference
//                     gathers into f[] (computeRHS.cpp:57-260): a cell, a cell
//                     reflected at a wall face, or the freestream state.
//   cell CSR          the faces of every cell in the reference's
//                     accumulation order, so residual sums are bit-identical.
// Kernels per iteration: rhs_faces (reconstruction + Roe flux + Jacobians),
// rhs_cells (residual/sigma sums, norms, ADI scaling), one ADI kernel per
// chain phase (block-tridiagonal solves), accumulate dq, lin_faces
// (linearised fluxes), lin_cells (linear residual, norms, q += dq).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hamstr_hip.h"
#include "device_math.hpp"
#include "strand_math.hpp"

using namespace hamstr;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                           \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess)                                                                   \
      return fail(HAMSTR_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));         \
  } while (0)

constexpr int kBlock = 256;
constexpr int kRedBlock = 256;
constexpr int kGraphMaxIters = 64;  // iterations per captured graph
constexpr int kMaxCR = 3;           // cyclic-reduction levels the autotuner tries before PCR

inline int nblk(long n, int b = kBlock) { return (int)((n + b - 1) / b); }

// The hardware deals the workgroups of a launch to the 8 XCDs round-robin
// (workgroup b to XCD b % 8), and each XCD has its own L2.  Kernels work on
// virtual block v = (b % 8) * (nb / 8) + b / 8 instead, so every XCD covers a
// contiguous range: neighbouring radial lines, whose cells share cache lines,
// then meet in the same L2.  A bijection on [0, nb); the tail beyond the last
// multiple of 8 keeps its order.
constexpr int kXcd = 8;
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int per = nb / kXcd;
  return b < per * kXcd ? (b % kXcd) * per + b / kXcd : b;
}

// --------------------------------------------------------------------------
// Device tables
// --------------------------------------------------------------------------
struct Consts {
  double cfl, gamma, qt, eps;
  int nCell, nProc, order;
  int nChainProc;  // processed faces [0, nChainProc) lie on chains, [nChainProc, nProc) are orphan faces
};

// pslot codes: the slot index in bits 0-28; bit 30: last face of an open
// chain; bit 29: an orphan face, whose two cells sit in slots s and s+1
constexpr int kSlotMask = (1 << 29) - 1;
constexpr int kOrphanSlot = 1 << 29;

struct FaceTab {
  const int* pslot;      // slot code of the face (see kSlotMask)
  const int2* src;       // per slot: (cell, wall face); cell < 0: freestream of case .y; face < 0: copy
  const double* nvec;    // [nProc][2]
  const double* refmtx;  // [nFace][4]
  const double4* qinf;   // [nCase] freestream state per flow case
  int nslot;             // slots in src
};

// Retired cases of a batch: one byte per processed face, cell and chain, zero
// where the case no longer iterates.  Null pointers: every case is live.
struct Live {
  const unsigned char* face;
  const unsigned char* cell;
  const unsigned char* chain;
};

struct ChainTab {
  const int* pbase;   // first processed face of the chain
  const int* nf;      // processed faces
  const int* ncell;   // rows of the block system
  const int* periodic;
  const int* pleft;   // per processed face: left cell
  const int* pright;  // per processed face: right cell (or boundary id)
  const int* pface;   // per processed face: face id
  const double* refmtx;
  // orphan faces (null when there are none): od_chain[cell] is the pbase of
  // the chain whose rows take the cell's orphan diagonal blocks; od_off /
  // od_ent list the cell's orphan processed faces, code = p << 1 | (cell is left)
  const int* od_chain;
  const int* od_off;
  const int* od_ent;
};

__device__ __forceinline__ double4 load_state(const int2 e, const double* __restrict__ q,
                                              const double* __restrict__ refmtx,
                                              const double4* __restrict__ qinf) {
  if (e.x < 0) return qinf[e.y];
  double4 v = *reinterpret_cast<const double4*>(q + 4 * (size_t)e.x);
  if (e.y >= 0) {
    const double4 R = *reinterpret_cast<const double4*>(refmtx + 4 * (size_t)e.y);
    const double a = v.y, b = v.z;
    v.y = (a * R.x + b * R.y);
    v.z = (a * R.z + b * R.w);
  }
  return v;
}

// same rows for dq: the freestream ghost has dq = 0 (computeLinearRHS.cpp:262-270)
__device__ __forceinline__ double4 load_dstate(const int2 e, const double* __restrict__ dq,
                                               const double* __restrict__ refmtx) {
  if (e.x < 0) return make_double4(0.0, 0.0, 0.0, 0.0);
  double4 v = *reinterpret_cast<const double4*>(dq + 4 * (size_t)e.x);
  if (e.y >= 0) {
    const double4 R = *reinterpret_cast<const double4*>(refmtx + 4 * (size_t)e.y);
    const double a = v.y, b = v.z;
    v.y = (a * R.x + b * R.y);
    v.z = (a * R.z + b * R.w);
  }
  return v;
}

__device__ __forceinline__ void put4(double* d, const double4 v) {
  d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
}

// The face Jacobians from (n, left, right), behind an empty asm barrier on
// the inputs: the compiler can share no subexpression with the code around
// the call (k_rhs_faces' flux), so its multiply-add fusion under FASTJ is the
// same wherever it is called, and k_lin_faces can compute again, bit for bit,
// the Jacobians k_rhs_faces stored.
template <typename TL, bool FASTJ>
__device__ __forceinline__ void face_jacobians(double nx, double ny, const double* left, const double* right,
                                               double gamma, TL* lm, TL* rm) {
  double l[4], r[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    l[k] = left[k];
    r[k] = right[k];
    asm volatile("" : "+v"(l[k]), "+v"(r[k]));
  }
  asm volatile("" : "+v"(nx), "+v"(ny));
  if (FASTJ)
    jacobian_roe2d_fast<TL>(nx, ny, l, r, gamma, lm, rm);
  else
    jacobian_roe2d<TL>(nx, ny, l, r, gamma, lm, rm);
}

// ql[n] and qr[n+1] along the chain for one face (reconstruction.cpp)
template <int ORDER>
__device__ __forceinline__ void reconstruct_face(const double (*f)[4], double eps, double qt, double* qln,
                                                 double* qrn1) {
  // f[] holds slots n-2 .. n+3 at indices 0..5
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (ORDER == 1) {
      qln[k] = f[2][k];
      qrn1[k] = f[3][k];
    } else if (ORDER == 3) {
      double a, b;
      muscl_pt(f[1][k], f[2][k], f[3][k], qt, eps, &qln[k], &a);
      muscl_pt(f[2][k], f[3][k], f[4][k], qt, eps, &b, &qrn1[k]);
    } else {
      qln[k] = weno5(f[0][k], f[1][k], f[2][k], f[3][k], f[4][k], 1e-15);
      qrn1[k] = weno5(f[5][k], f[4][k], f[3][k], f[2][k], f[1][k], 1e-15);
    }
  }
}

template <int ORDER, bool FASTJ>
__device__ __forceinline__ void reconstruct_dface(const double (*f)[4], const double (*d)[4], double eps,
                                                  double qt, double* dln, double* drn1) {
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (ORDER == 1) {
      dln[k] = d[2][k];
      drn1[k] = d[3][k];
    } else if (ORDER == 3 && FASTJ) {
      double a, b;
      muscl_deriv_pt_fast(f[1][k], f[2][k], f[3][k], d[1][k], d[2][k], d[3][k], qt, eps, &dln[k], &a);
      muscl_deriv_pt_fast(f[2][k], f[3][k], f[4][k], d[2][k], d[3][k], d[4][k], qt, eps, &b, &drn1[k]);
    } else if (ORDER == 3) {
      double a, b;
      muscl_deriv_pt(f[1][k], f[2][k], f[3][k], d[1][k], d[2][k], d[3][k], qt, eps, &dln[k], &a);
      muscl_deriv_pt(f[2][k], f[3][k], f[4][k], d[2][k], d[3][k], d[4][k], qt, eps, &b, &drn1[k]);
    } else {
      dln[k] = weno5_deriv(f[0][k], f[1][k], f[2][k], f[3][k], f[4][k], d[0][k], d[1][k], d[2][k], d[3][k],
                           d[4][k], eps);
      drn1[k] = weno5_deriv(f[5][k], f[4][k], f[3][k], f[2][k], f[1][k], d[5][k], d[4][k], d[3][k], d[2][k],
                            d[1][k], eps);
    }
  }
}

// slots a face's stencil reads, relative to its own slot s: s-2+j for j in
// [lo, hi] of reconstruct_face's table (orphan faces: s and s+1)
constexpr __host__ __device__ int stencil_lo(int order) { return order == 5 ? -2 : (order == 3 ? -1 : 0); }
constexpr __host__ __device__ int stencil_hi(int order) { return order == 5 ? 3 : (order == 3 ? 2 : 1); }
// LDS capacity, in slots, of the staged face kernels: 256 orphan faces (two
// slots each) and their halo take 515; 256 faces on chains of a few dozen
// cells or more take fewer.  The fused kernel holds 12 FP64 rows per slot
// (49,920 B), as many workgroups per CU as its 154 VGPRs allow.
constexpr int kStageSlots = 520;

// Staged face kernels: slots grow with the processed face (chains in order,
// orphan faces after them), so the stencils of a workgroup's faces lie in one
// range of slots, from its first face's stencil to its last one's.  Each
// slot's rows (load_state, load_dstate, ...) are formed once into LDS and the
// faces read them there instead of each gathering them from global memory.
// The host launches them only where every workgroup's range fits kStageSlots.
template <int ORDER>
__device__ __forceinline__ int2 stage_range(const FaceTab& ft, int nProc, int pf) {
  const int pl = min(pf + kBlock - 1, nProc - 1);
  return make_int2(max((ft.pslot[pf] & kSlotMask) + stencil_lo(ORDER), 0),
                   min((ft.pslot[pl] & kSlotMask) + stencil_hi(ORDER), ft.nslot - 1));
}

// The two states of processed face `code` (its pslot code) that its flux and
// Jacobians take (computeRHS.cpp:53-260): for an orphan face the two cell
// states (first order), else the states reconstructed along the chain,
// oriented as the face is; get(slot, out4) gives a slot's state
template <int ORDER, typename Get>
__device__ __forceinline__ void face_states_from(Get get, double eps, double qt, int code, double* left,
                                                 double* right) {
  const int s = code & kSlotMask;
  if (code & kOrphanSlot) {
    get(s, left);
    get(s + 1, right);
  } else {
    const bool flip = (code >> 30) & 1;
    constexpr int lo = (ORDER == 5) ? 0 : (ORDER == 3 ? 1 : 2);
    constexpr int hi = (ORDER == 5) ? 5 : (ORDER == 3 ? 4 : 3);
    double f[6][4];
#pragma unroll
    for (int j = lo; j <= hi; j++) get(s - 2 + j, f[j]);
    double qln[4], qrn1[4];
    reconstruct_face<ORDER>(f, eps, qt, qln, qrn1);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      left[k] = flip ? qln[k] : qrn1[k];
      right[k] = flip ? qrn1[k] : qln[k];
    }
  }
}

// face_states_from with the states gathered from q
template <int ORDER>
__device__ __forceinline__ void face_states(const FaceTab& ft, const double* __restrict__ q, double eps, double qt,
                                            int code, double* left, double* right) {
  face_states_from<ORDER>(
      [&](int slot, double* o) __attribute__((always_inline)) {
        put4(o, load_state(ft.src[slot], q, ft.refmtx, ft.qinf));
      },
      eps, qt, code, left, right);
}

// Where a FAST ADI kernel takes the face Jacobians from.  order 0: the array
// ff k_rhs_faces stored.  Otherwise computed again (face_jac_q), which the
// fused FAST iteration can do for the chain faces instead of storing and
// reading back 32 values per face: from the face states k_rhs_faces stored
// (fst non-null: [8][nProc], left then right), or from q at `order` (1, 3
// or 5: the stencil gathered and reconstructed again).
struct JacSrc {
  FaceTab ft;
  const double* q;
  double gamma, qt, eps;
  int order;
  const double* fst;
  int nProc;
};

// the FAST-mode Jacobians of processed face p: the states and the code
// k_rhs_faces uses (face_jacobians is opaque to its surroundings), so bit
// for bit what k_rhs_faces stores
template <typename TL>
__device__ __forceinline__ void face_jac_q(const JacSrc& js, int p, TL* lm, TL* rm) {
  double left[4], right[4];
  if (js.fst) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      left[k] = js.fst[(size_t)k * js.nProc + p];
      right[k] = js.fst[(size_t)(4 + k) * js.nProc + p];
    }
  } else {
    const int code = js.ft.pslot[p];
    if (js.order == 1)
      face_states<1>(js.ft, js.q, js.eps, js.qt, code, left, right);
    else if (js.order == 3)
      face_states<3>(js.ft, js.q, js.eps, js.qt, code, left, right);
    else
      face_states<5>(js.ft, js.q, js.eps, js.qt, code, left, right);
  }
  const double2 n = *reinterpret_cast<const double2*>(js.ft.nvec + 2 * (size_t)p);
  face_jacobians<TL, true>(n.x, n.y, left, right, js.gamma, lm, rm);
}

// --------------------------------------------------------------------------
// rhs_faces: chain gather + reconstruction + Roe flux (+ Jacobians)
// computeRHS.cpp:53-408
// --------------------------------------------------------------------------
// FASTJ: the FAST-mode Jacobians (reciprocals instead of divisions); the flux
// and spectral radius are the reference's arithmetic either way
// ff_from: the Jacobians are computed and stored for processed faces p >=
// ff_from or p < jac_to and the last faces of open chains only (the fused
// FAST iteration can leave the other chain faces' to its ADI kernels, which
// compute them again); fst non-null: those other faces' two states are
// stored instead
// ([8][nProc], left then right) for the ADI kernels to compute them from
// STAGED: the workgroup's load_state rows staged in LDS (stage_range)
template <int ORDER, typename TL, bool IMPL, bool FASTJ, bool STAGED>
__global__ __launch_bounds__(kBlock) void k_rhs_faces(FaceTab ft, Consts c, Live lv, const double* __restrict__ q,
                                                      double* __restrict__ flux, double* __restrict__ sr,
                                                      TL* __restrict__ ff, int ff_from, double* __restrict__ fst,
                                                      int* __restrict__ nanflag, int jac_to = 0) {
  const int pf = xcd_block(blockIdx.x, gridDim.x) * blockDim.x;
  const int p = pf + threadIdx.x;
  __shared__ double stg[STAGED ? 4 : 1][STAGED ? kStageSlots : 1];
  int S0 = 0;
  if constexpr (STAGED) {
    // a workgroup whose faces are all retired (converged cases of a batch)
    // stages nothing
    if (!__syncthreads_or(p < c.nProc && (!lv.face || lv.face[p]))) return;
    const int2 R = stage_range<ORDER>(ft, c.nProc, pf);
    S0 = R.x;
    for (int i = threadIdx.x; i <= R.y - R.x; i += blockDim.x) {
      const double4 a = load_state(ft.src[S0 + i], q, ft.refmtx, ft.qinf);
      stg[0][i] = a.x; stg[1][i] = a.y; stg[2][i] = a.z; stg[3][i] = a.w;
    }
    __syncthreads();
  }
  if (p >= c.nProc || (lv.face && !lv.face[p])) return;
  const int code = ft.pslot[p];
  double left[4], right[4];
  if constexpr (STAGED)
    face_states_from<ORDER>(
        [&](int slot, double* o) __attribute__((always_inline)) {
#pragma unroll
          for (int k = 0; k < 4; k++) o[k] = stg[k][slot - S0];
        },
        c.eps, c.qt, code, left, right);
  else
    face_states<ORDER>(ft, q, c.eps, c.qt, code, left, right);
  const double2 n = *reinterpret_cast<const double2*>(ft.nvec + 2 * (size_t)p);
  double fl[4], s_r;
  flux_roe2d(n.x, n.y, left, right, c.gamma, fl, &s_r);
  *reinterpret_cast<double4*>(flux + 4 * (size_t)p) = make_double4(fl[0], fl[1], fl[2], fl[3]);
  sr[p] = s_r;
  if (s_r != s_r) atomicOr(nanflag, 1);
  const bool store_jac = p >= ff_from || p < jac_to || ((code >> 30) & 1);
  if (IMPL && fst && !store_jac) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      fst[(size_t)k * c.nProc + p] = left[k];
      fst[(size_t)(4 + k) * c.nProc + p] = right[k];
    }
  }
  if (IMPL && store_jac) {
    TL lm[16], rm[16];
    face_jacobians<TL, FASTJ>(n.x, n.y, left, right, c.gamma, lm, rm);
#pragma unroll
    for (int e = 0; e < 16; e++) {
      ff[(size_t)e * c.nProc + p] = lm[e];
      ff[(size_t)(16 + e) * c.nProc + p] = rm[e];
    }
  }
}

// --------------------------------------------------------------------------
// block reductions for the density-residual norms
// --------------------------------------------------------------------------
__device__ __forceinline__ void block_norm_partial(double sq, double mx, double* __restrict__ partial, int blk,
                                                   double* __restrict__ partial2 = nullptr) {
  // per-block partial (sum of r_rho^2, max |r_rho|); k_norm_final reduces them.
  // Written to partial and, when given, to partial2 as well.  A butterfly
  // over each wave's 64 lanes (lane 0's order is fixed), then the waves'
  // results in wave order: deterministic, with one barrier.
  static_assert(kBlock % 64 == 0, "whole waves");
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    sq += __shfl_xor(sq, o);
    mx = fmax(mx, __shfl_xor(mx, o));
  }
  __shared__ double wsq[kBlock / 64];
  __shared__ double wmx[kBlock / 64];
  const int t = threadIdx.x;
  if ((t & 63) == 0) {
    wsq[t >> 6] = sq;
    wmx[t >> 6] = mx;
  }
  __syncthreads();
  if (t == 0) {
    double a = wsq[0], b = wmx[0];
#pragma unroll
    for (int w = 1; w < kBlock / 64; w++) {
      a += wsq[w];
      b = fmax(b, wmx[w]);
    }
    partial[2 * blk] = a;
    partial[2 * blk + 1] = b;
    if (partial2) {
      partial2[2 * blk] = a;
      partial2[2 * blk + 1] = b;
    }
  }
}

// L2 = sqrt(sum r_rho^2) / nCell, Linf = max |r_rho| (computeRHS.cpp:415-425),
// partials reduced in block order so the result does not depend on scheduling
__global__ __launch_bounds__(kRedBlock) void k_norm_final(const double* __restrict__ partial, int nPart,
                                                          int nCell, double* __restrict__ out) {
  __shared__ double ssq[kRedBlock];
  __shared__ double smx[kRedBlock];
  const int t = threadIdx.x;
  double a = 0.0, m = 0.0;
  for (int i = t; i < nPart; i += kRedBlock) {
    a += partial[2 * i];
    m = fmax(m, partial[2 * i + 1]);
  }
  ssq[t] = a;
  smx[t] = m;
  __syncthreads();
  for (int w = kRedBlock / 2; w > 0; w >>= 1) {
    if (t < w) {
      ssq[t] += ssq[t + w];
      smx[t] = fmax(smx[t], smx[t + w]);
    }
    __syncthreads();
  }
  if (t == 0) {
    out[0] = sqrt(ssq[0]) / nCell;
    out[1] = smx[0];
  }
}

// the same reduction for a run of iterations at once: block b reduces the
// partials of iteration b/2, stage b%2 (0 nonlinear, 1 linear) into
// out[4*(b/2) + 2*(b%2)], out[... + 1]
__global__ __launch_bounds__(kRedBlock) void k_norm_slots(const double* __restrict__ part, int nPart, int nCell,
                                                          double* __restrict__ out) {
  __shared__ double ssq[kRedBlock];
  __shared__ double smx[kRedBlock];
  const int t = threadIdx.x, b = blockIdx.x;
  const double* partial = part + (size_t)b * nPart * 2;
  double a = 0.0, m = 0.0;
  for (int i = t; i < nPart; i += kRedBlock) {
    a += partial[2 * i];
    m = fmax(m, partial[2 * i + 1]);
  }
  ssq[t] = a;
  smx[t] = m;
  __syncthreads();
  for (int w = kRedBlock / 2; w > 0; w >>= 1) {
    if (t < w) {
      ssq[t] += ssq[t + w];
      smx[t] = fmax(smx[t], smx[t + w]);
    }
    __syncthreads();
  }
  if (t == 0) {
    out[2 * b] = sqrt(ssq[0]) / nCell;
    out[2 * b + 1] = smx[0];
  }
}

// --------------------------------------------------------------------------
// per-cell sums over the cell's face list in list order (the reference's
// order).  The first kGather entries and their face values are all loaded
// before any is added, so their loads are in flight together; cells of
// triangles and quadrilaterals have at most 4 entries.
// --------------------------------------------------------------------------
constexpr int kGather = 4;

__device__ __forceinline__ void cell_flux_sum(const int* __restrict__ ent, int e0, int e1,
                                              const double* __restrict__ flux, const double* __restrict__ sr,
                                              double& a0, double& a1, double& a2, double& a3, double& sg) {
  const int n = min(e1 - e0, kGather);
  int code[kGather];
  double4 fx[kGather];
  double s4[kGather];
#pragma unroll
  for (int k = 0; k < kGather; k++) code[k] = (k < n) ? ent[e0 + k] : 0;
#pragma unroll
  for (int k = 0; k < kGather; k++) {
    if (k < n) {
      fx[k] = *reinterpret_cast<const double4*>(flux + 4 * (size_t)(code[k] >> 1));
      s4[k] = sr[code[k] >> 1];
    }
  }
#pragma unroll
  for (int k = 0; k < kGather; k++) {
    if (k < n) {
      if (code[k] & 1) {
        a0 += fx[k].x; a1 += fx[k].y; a2 += fx[k].z; a3 += fx[k].w;
      } else {
        a0 -= fx[k].x; a1 -= fx[k].y; a2 -= fx[k].z; a3 -= fx[k].w;
      }
      sg += s4[k];
    }
  }
  for (int e = e0 + kGather; e < e1; e++) {
    const int cd = ent[e];
    const int p = cd >> 1;
    const double4 f = *reinterpret_cast<const double4*>(flux + 4 * (size_t)p);
    if (cd & 1) {
      a0 += f.x; a1 += f.y; a2 += f.z; a3 += f.w;
    } else {
      a0 -= f.x; a1 -= f.y; a2 -= f.z; a3 -= f.w;
    }
    sg += sr[p];
  }
}

__device__ __forceinline__ void cell_lin_sum(const int* __restrict__ ent, int e0, int e1,
                                             const double* __restrict__ lflux, double dtfac, double& a0, double& a1,
                                             double& a2, double& a3) {
  const int n = min(e1 - e0, kGather);
  int code[kGather];
  double4 fx[kGather];
#pragma unroll
  for (int k = 0; k < kGather; k++) code[k] = (k < n) ? ent[e0 + k] : 0;
#pragma unroll
  for (int k = 0; k < kGather; k++)
    if (k < n) fx[k] = *reinterpret_cast<const double4*>(lflux + 4 * (size_t)(code[k] >> 1));
#pragma unroll
  for (int k = 0; k < kGather; k++) {
    if (k < n) {
      if (code[k] & 1) {
        a0 += (fx[k].x * dtfac); a1 += (fx[k].y * dtfac); a2 += (fx[k].z * dtfac); a3 += (fx[k].w * dtfac);
      } else {
        a0 -= (fx[k].x * dtfac); a1 -= (fx[k].y * dtfac); a2 -= (fx[k].z * dtfac); a3 -= (fx[k].w * dtfac);
      }
    }
  }
  for (int e = e0 + kGather; e < e1; e++) {
    const int cd = ent[e];
    const double4 f = *reinterpret_cast<const double4*>(lflux + 4 * (size_t)(cd >> 1));
    if (cd & 1) {
      a0 += (f.x * dtfac); a1 += (f.y * dtfac); a2 += (f.z * dtfac); a3 += (f.w * dtfac);
    } else {
      a0 -= (f.x * dtfac); a1 -= (f.y * dtfac); a2 -= (f.z * dtfac); a3 -= (f.w * dtfac);
    }
  }
}

// --------------------------------------------------------------------------
// rhs_cells: r and sigma summed in the reference's order, norms, and (in the
// fused iteration) ADI's scaling r *= CFL/sigma, r0 = r (ADI.cpp:40-49)
// --------------------------------------------------------------------------
// Cells [c0, c1) (c0 a multiple of kBlock), block partial c0/kBlock + vb
// for the cells of virtual block vb: a launch split at c0 leaves every
// partial as one launch over all cells writes it.
template <bool ADI_INIT>
__global__ __launch_bounds__(kBlock) void k_rhs_cells(Consts c, Live lv, int c0, const int* __restrict__ off,
                                                      const int* __restrict__ ent,
                                                      const double* __restrict__ flux,
                                                      const double* __restrict__ sr, double* __restrict__ r,
                                                      double* __restrict__ r0, double* __restrict__ sigma,
                                                      double* __restrict__ partial, double* __restrict__ rrho,
                                                      int c1) {
  const int vb = xcd_block(blockIdx.x, gridDim.x);
  const int i = c0 + vb * blockDim.x + threadIdx.x;
  double sq = 0.0, mx = 0.0;
  if (i < c1 && (!lv.cell || lv.cell[i])) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, sg = 0.0;
    cell_flux_sum(ent, off[i], off[i + 1], flux, sr, a0, a1, a2, a3, sg);
    sq = a0 * a0;
    mx = fabs(a0);
    if (rrho) rrho[i] = a0;
    sigma[i] = sg;
    if (ADI_INIT) {
      const double dtfac = c.cfl / sg;
      const double4 v = make_double4(a0 * dtfac, a1 * dtfac, a2 * dtfac, a3 * dtfac);
      *reinterpret_cast<double4*>(r + 4 * (size_t)i) = v;
      *reinterpret_cast<double4*>(r0 + 4 * (size_t)i) = v;
    } else {
      *reinterpret_cast<double4*>(r + 4 * (size_t)i) = make_double4(a0, a1, a2, a3);
    }
  }
  block_norm_partial(sq, mx, partial, c0 / kBlock + vb);
}

// --------------------------------------------------------------------------
// One explicit stage (march.cpp:73-108): computeRHS("explicit")'s cell sums,
// summed as k_rhs_cells sums them, fused with up to two updateSolution calls
// on this stage's r and sigma: dstA = srcA + (coefA*CFL/sigma) r, then, when
// dstB is set, dstB = srcB + (coefB*CFL/sigma) r (rk3's first stage writes qt
// and q from the same q).  r and sigma are stored as computeRHS leaves them.
// Block norm partials go to partA and partB (either may be null; both null:
// no norms), the raw density residual to rrhoA / rrhoB (batches).
// --------------------------------------------------------------------------
struct ExplicitStage {
  const double* srcA;
  double* dstA;
  double coefA;
  const double* srcB;
  double* dstB;
  double coefB;
  double* partA;
  double* partB;
  double* rrhoA;
  double* rrhoB;
};

__device__ __forceinline__ void explicit_update(const double* src, double* dst, double coef, double cfl, double sg,
                                                double a0, double a1, double a2, double a3, int i) {
  const double dtfac = coef * cfl / sg;
  const double4 s = *reinterpret_cast<const double4*>(src + 4 * (size_t)i);
  *reinterpret_cast<double4*>(dst + 4 * (size_t)i) =
      make_double4(s.x + dtfac * a0, s.y + dtfac * a1, s.z + dtfac * a2, s.w + dtfac * a3);
}

__global__ __launch_bounds__(kBlock) void k_rhs_cells_explicit(Consts c, Live lv, const int* __restrict__ off,
                                                               const int* __restrict__ ent,
                                                               const double* __restrict__ flux,
                                                               const double* __restrict__ sr, double* __restrict__ r,
                                                               double* __restrict__ sigma, ExplicitStage st) {
  const int vb = xcd_block(blockIdx.x, gridDim.x);
  const int i = vb * blockDim.x + threadIdx.x;
  double sq = 0.0, mx = 0.0;
  if (i < c.nCell && (!lv.cell || lv.cell[i])) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, sg = 0.0;
    cell_flux_sum(ent, off[i], off[i + 1], flux, sr, a0, a1, a2, a3, sg);
    sq = a0 * a0;
    mx = fabs(a0);
    if (st.rrhoA) st.rrhoA[i] = a0;
    if (st.rrhoB) st.rrhoB[i] = a0;
    sigma[i] = sg;
    *reinterpret_cast<double4*>(r + 4 * (size_t)i) = make_double4(a0, a1, a2, a3);
    explicit_update(st.srcA, st.dstA, st.coefA, c.cfl, sg, a0, a1, a2, a3, i);
    if (st.dstB) explicit_update(st.srcB, st.dstB, st.coefB, c.cfl, sg, a0, a1, a2, a3, i);
  }
  // the condition is the same for every thread of the block
  if (st.partA || st.partB) block_norm_partial(sq, mx, st.partA ? st.partA : st.partB, vb, st.partA ? st.partB : nullptr);
}

// ADI.cpp:40-49 on its own (API path)
__global__ void k_adi_init(Consts c, Live lv, double* __restrict__ r, double* __restrict__ r0,
                           double* __restrict__ dq, const double* __restrict__ sigma) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c.nCell || (lv.cell && !lv.cell[i])) return;
  const double dtfac = c.cfl / sigma[i];
  double4 v = *reinterpret_cast<double4*>(r + 4 * (size_t)i);
  v.x *= dtfac; v.y *= dtfac; v.z *= dtfac; v.w *= dtfac;
  *reinterpret_cast<double4*>(r + 4 * (size_t)i) = v;
  *reinterpret_cast<double4*>(r0 + 4 * (size_t)i) = v;
  *reinterpret_cast<double4*>(dq + 4 * (size_t)i) = make_double4(0.0, 0.0, 0.0, 0.0);
}

// dq += r (ADI.cpp:247-248); with FIRST the old dq is zero
template <bool FIRST>
__global__ void k_accum_dq(int n4, Live lv, const double* __restrict__ r, double* __restrict__ dq) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n4 || (lv.cell && !lv.cell[i >> 2])) return;
  if (FIRST)
    dq[i] = 0.0 + r[i];
  else
    dq[i] += r[i];
}

// --------------------------------------------------------------------------
// ADI block-tridiagonal row assembly (ADI.cpp:57-217) for row m of a chain
// --------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void load_fmat(const T* __restrict__ ff, int nProc, int p, T* L, T* R) {
#pragma unroll
  for (int e = 0; e < 16; e++) {
    L[e] = ff[(size_t)e * nProc + p];
    R[e] = ff[(size_t)(16 + e) * nProc + p];
  }
}

template <typename T>
__device__ __forceinline__ void reflect_cols(T* R, const double* rm) {
  const T a1 = T(rm[0]), b1 = T(rm[1]), a2 = T(rm[2]), b2 = T(rm[3]);
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const T r01 = R[4 * j + 1], r02 = R[4 * j + 2];
    R[4 * j + 1] = a1 * r01 + b1 * r02;
    R[4 * j + 2] = a2 * r01 + b2 * r02;
  }
}

template <typename T>
__device__ __forceinline__ void assemble_row(const ChainTab& ct, const T* __restrict__ ff, int nProc, int pb,
                                             int nf, int N, bool periodic, int m, double cfl,
                                             const double* __restrict__ sigma, const double* __restrict__ r,
                                             T* A, T* B, T* C, T* F, int* cell_out) {
  const int pm = pb + m;
  T L[16], R[16];
  load_fmat(ff, nProc, pm, L, R);
  if (periodic) {
    const int pn = pb + ((m + 1 == N) ? 0 : m + 1);
    T Ln[16], Rn[16];
    load_fmat(ff, nProc, pn, Ln, Rn);
#pragma unroll
    for (int e = 0; e < 16; e++) {
      B[e] = L[e] - Rn[e];
      A[e] = R[e];
      C[e] = T(0) - Ln[e];
    }
  } else {
    const int rc = ct.pright[pm];
    if (rc == -2) reflect_cols(R, ct.refmtx + 4 * (size_t)ct.pface[pm]);
#pragma unroll
    for (int e = 0; e < 16; e++) {
      B[e] = L[e];
      A[e] = T(0);
      C[e] = T(0);
    }
    if (m > 0 && rc > -1) {
#pragma unroll
      for (int e = 0; e < 16; e++) A[e] = R[e];
    } else if (rc == -2) {
#pragma unroll
      for (int e = 0; e < 16; e++) B[e] += R[e];
    }
    if (m + 1 < nf) {
      const int pn = pm + 1;
      const int rn = ct.pright[pn];
      load_fmat(ff, nProc, pn, L, R);
      if (rn == -2) reflect_cols(R, ct.refmtx + 4 * (size_t)ct.pface[pn]);
      if (m + 1 == nf - 1 && rn < 0) {
#pragma unroll
        for (int e = 0; e < 16; e++) B[e] += L[e];
        if (rn == -2) {
#pragma unroll
          for (int e = 0; e < 16; e++) B[e] += R[e];
        }
      } else if (rn > -1) {
#pragma unroll
        for (int e = 0; e < 16; e++) {
          B[e] -= R[e];
          C[e] = T(0) - L[e];
        }
      }
    }
  }
  const int cell = ct.pleft[pm];
  *cell_out = cell;
  if (ct.od_chain && ct.od_chain[cell] == pb) {
    // the cell's orphan faces: + lmat where it is the left cell, - rmat where right
    const int e1 = ct.od_off[cell + 1];
    for (int e = ct.od_off[cell]; e < e1; e++) {
      const int code = ct.od_ent[e];
      const int po = code >> 1;
      if (code & 1) {
#pragma unroll
        for (int k = 0; k < 16; k++) B[k] += ff[(size_t)k * nProc + po];
      } else {
#pragma unroll
        for (int k = 0; k < 16; k++) B[k] -= ff[(size_t)(16 + k) * nProc + po];
      }
    }
  }
  const T dtfac = T(cfl / sigma[cell]);
#pragma unroll
  for (int e = 0; e < 16; e++) {
    A[e] *= dtfac;
    B[e] *= dtfac;
    C[e] *= dtfac;
  }
  B[0] += T(1.0);
  B[5] += T(1.0);
  B[10] += T(1.0);
  B[15] += T(1.0);
  const double4 rv = *reinterpret_cast<const double4*>(r + 4 * (size_t)cell);
  F[0] = T(rv.x);
  F[1] = T(rv.y);
  F[2] = T(rv.z);
  F[3] = T(rv.w);
}

// r[cell] = x; with dqmode 1 also dq[cell] = 0 + x, with 2 dq[cell] += x
// (ADI.cpp:240, 247-248, fused into the last phase of a sweep when that
// phase covers every cell)
template <typename T>
__device__ __forceinline__ void store_solution(double* __restrict__ r, double* __restrict__ dq, int dqmode,
                                               int cell, const T* x) {
  const double4 v = make_double4(double(x[0]), double(x[1]), double(x[2]), double(x[3]));
  *reinterpret_cast<double4*>(r + 4 * (size_t)cell) = v;
  if (dqmode == 1) {
    *reinterpret_cast<double4*>(dq + 4 * (size_t)cell) = make_double4(0.0 + v.x, 0.0 + v.y, 0.0 + v.z, 0.0 + v.w);
  } else if (dqmode == 2) {
    double4 d = *reinterpret_cast<double4*>(dq + 4 * (size_t)cell);
    d.x += v.x; d.y += v.y; d.z += v.z; d.w += v.w;
    *reinterpret_cast<double4*>(dq + 4 * (size_t)cell) = d;
  }
}

// --------------------------------------------------------------------------
// ADI phase, exact order (ADI.cpp:57-244 with blockTridag4 /
// blockTridagPeriodic4, triSolvers.cpp:769-1046), in two launches.
//
// k_exact_assemble: every row of the group's chains assembled at once
// (assemble_row, a thread per row, a workgroup per chain) into rows[p*52 + e]
// (p the row's processed-face slot; e < 52: A, B, C row-major, then F), so
// that the sequential walk reads each row from one 416-byte block.
//
// k_adi_exact16: 16 lanes per chain, four chains per 64-lane wavefront.  Lane
// (ri, ci) owns entry (ri, ci) of the 4x4 blocks.  Per row it updates its B
// entry (B -= A C'), its F row (F -= A F') and its spike entry (a = -A a');
// one LDS exchange gives every lane of the chain the whole B, F and its
// column of the spike; every lane factors B and solves F (the same
// operations in each lane), and the lanes of column ci solve column ci of C
// (and of the spike) -- each entry with the reference's own operation
// sequence, so the result is the reference's bit for bit.  The rows C', F', a'
// go to scratch scr[p*36 + e] (C' 0-15, F' 16-19, a' 20-35, row-major) for the
// back sweep, which every lane of the chain runs on whole rows; one lane per
// chain writes the solution into r (and dq).
// --------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_exact_assemble(ChainTab ct, Live lv, const int* __restrict__ chains, int nch,
                                                        const T* __restrict__ ff, int nProc, double cfl,
                                                        const double* __restrict__ sigma,
                                                        const double* __restrict__ r, T* __restrict__ rows) {
  const int w = blockIdx.x;
  if (w >= nch) return;
  const int ch = chains[w];
  if (lv.chain && !lv.chain[ch]) return;  // the same for the whole workgroup
  const int pb = ct.pbase[ch], nf = ct.nf[ch], N = ct.ncell[ch];
  const bool periodic = ct.periodic[ch] != 0;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    T A[16], B[16], C[16], F[4];
    int cell;
    assemble_row<T>(ct, ff, nProc, pb, nf, N, periodic, i, cfl, sigma, r, A, B, C, F, &cell);
    T* o = rows + (size_t)(pb + i) * 52;
#pragma unroll
    for (int e = 0; e < 16; e++) {
      o[e] = A[e];
      o[16 + e] = B[e];
      o[32 + e] = C[e];
    }
#pragma unroll
    for (int e = 0; e < 4; e++) o[48 + e] = F[e];
  }
}

// one lane's share of an assembled row: row ri of A, B[ri][ci], column ci of
// C, and all of F
template <typename T>
struct ExactRow {
  T a[4], b, c[4], f[4];
};

template <typename T>
__device__ __forceinline__ void load_exact_row(const T* __restrict__ rows, size_t p, int ri, int ci, ExactRow<T>& o) {
  const T* base = rows + p * 52;
#pragma unroll
  for (int l = 0; l < 4; l++) {
    o.a[l] = base[4 * ri + l];
    o.c[l] = base[32 + 4 * l + ci];
    o.f[l] = base[48 + l];
  }
  o.b = base[16 + 4 * ri + ci];
}

// one lane's share of a back-sweep row: row ri of C' and of a', and F'[ri]
template <typename T>
struct BackRow {
  T c[4], a[4], f;
};

template <typename T>
__device__ __forceinline__ void load_back_row(const T* __restrict__ scr, size_t p, int ri, bool per, BackRow<T>& o) {
  const T* base = scr + p * 36;
#pragma unroll
  for (int l = 0; l < 4; l++) {
    o.c[l] = base[4 * ri + l];
    o.a[l] = per ? base[20 + 4 * ri + l] : T(0);
  }
  o.f = base[16 + ri];
}

// one lane's share of a row of the bordered elimination: column ci of C' and
// of a', and all of F'
template <typename T>
struct BorderRow {
  T c[4], a[4], f[4];
};

template <typename T>
__device__ __forceinline__ void load_border_row(const T* __restrict__ scr, size_t p, int ci, BorderRow<T>& o) {
  const T* base = scr + p * 36;
#pragma unroll
  for (int l = 0; l < 4; l++) {
    o.c[l] = base[4 * l + ci];
    o.a[l] = base[20 + 4 * l + ci];
    o.f[l] = base[16 + l];
  }
}

// Orders the LDS accesses of one wavefront around an exchange.  A wavefront's
// LDS operations are performed in order, so no wait is needed; the fence and
// the barrier only keep the compiler from moving them.  Unlike
// __syncthreads(), it does not wait for outstanding global loads (the rows
// prefetched ahead), so it is only for workgroups of a single wavefront.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// v[k] for a lane-dependent k, as selects (no indexed register access)
template <typename T>
__device__ __forceinline__ T pick4(const T* v, int k) {
  return k == 0 ? v[0] : (k == 1 ? v[1] : (k == 2 ? v[2] : v[3]));
}

// The three sequential sweeps below each keep kRing rows in flight: the loop
// is unrolled kRing times over kRing buffers, and a buffer is refilled with
// the row kRing ahead right after its row is used (no register copies, which
// would wait for the newest loads).  The refills are unconditional, with the
// row clamped into the chain: with a fixed number of loads per step the
// compiler waits only for the row a step uses, not for every load in flight.
constexpr int kRing = 4;

template <typename T>
__global__ __launch_bounds__(64) void k_adi_exact16(ChainTab ct, Live lv, const int* __restrict__ chains, int nch,
                                                    const T* __restrict__ rows, double* __restrict__ r,
                                                    T* __restrict__ scr, double* __restrict__ dq, int dqmode) {
  __shared__ T xb[4][36];  // per chain: B (16), F (4), spike a (16), row-major
  __shared__ int xn[4];
  const int lane = threadIdx.x, slot = lane >> 4, q = lane & 15, ri = q & 3, ci = q >> 2;
  const int w = xcd_block(blockIdx.x, gridDim.x) * 4 + slot;
  int pb = 0, N = 0;
  bool per = false;
  if (w < nch) {
    const int ch = chains[w];
    if (!lv.chain || lv.chain[ch]) {
      pb = ct.pbase[ch];
      N = ct.ncell[ch];
      per = ct.periodic[ch] != 0;
    }
  }
  if (q == 0) xn[slot] = N;
  wave_lds_sync();
  // every lane runs the longest chain's number of steps (the exchanges below
  // are wavefront-wide)
  const int Nmax = max(max(xn[0], xn[1]), max(xn[2], xn[3]));
  T* X = xb[slot];

  // ---- forward elimination ----
  T Ccol[4] = {T(0), T(0), T(0), T(0)}, acol[4] = {T(0), T(0), T(0), T(0)}, Fp[4] = {T(0), T(0), T(0), T(0)};
  auto forward = [&](int i, const ExactRow<T>& cur) __attribute__((always_inline)) {
    const bool act = i < N;
    T bv = cur.b, fr = pick4(cur.f, ri), av = T(0);
    if (i != 0) {
      // triSolvers.cpp:790-803 (and the spike a = -A a', :877-891)
#pragma unroll
      for (int l = 0; l < 4; l++) bv -= cur.a[l] * Ccol[l];
#pragma unroll
      for (int j = 0; j < 4; j++) fr -= cur.a[j] * Fp[j];
      if (per) {
#pragma unroll
        for (int l = 0; l < 4; l++) av -= cur.a[l] * acol[l];
      }
    } else if (per) {
      av = pick4(cur.a, ci);
    }
    X[4 * ri + ci] = bv;
    if (ci == 0) X[16 + ri] = fr;
    X[20 + 4 * ri + ci] = av;
    wave_lds_sync();
    T B[16], F[4], ac[4];
#pragma unroll
    for (int e = 0; e < 16; e++) B[e] = X[e];
#pragma unroll
    for (int l = 0; l < 4; l++) {
      F[l] = X[16 + l];
      ac[l] = X[20 + 4 * l + ci];
    }
    wave_lds_sync();
    if (act) {
      const LU4<T> L = lu4(B);
      lu4_solve(L, F, 1);
      T cc[4] = {cur.c[0], cur.c[1], cur.c[2], cur.c[3]};
      if (!per) {
        if (i != N - 1) lu4_solve(L, cc, 1);
      } else if (i != N - 2) {
        lu4_solve(L, cc, 1);
        lu4_solve(L, ac, 1);
      } else {
#pragma unroll
        for (int l = 0; l < 4; l++) cc[l] += ac[l];
        lu4_solve(L, cc, 1);
#pragma unroll
        for (int l = 0; l < 4; l++) ac[l] = cc[l];
      }
      if (i < N - 1) {
        T* o = scr + (size_t)(pb + i) * 36;
        o[4 * ri + ci] = pick4(cc, ri);
        if (ci == 0) o[16 + ri] = pick4(F, ri);
        if (per) o[20 + 4 * ri + ci] = pick4(ac, ri);
      }
#pragma unroll
      for (int l = 0; l < 4; l++) {
        Ccol[l] = cc[l];
        acol[l] = ac[l];
        Fp[l] = F[l];
      }
    }
  };
  {
    const int last = max(N - 1, 0);
    ExactRow<T> ring[kRing];
#pragma unroll
    for (int k = 0; k < kRing; k++) load_exact_row(rows, (size_t)pb + min(k, last), ri, ci, ring[k]);
    // whole rounds of kRing steps: the steps past a chain's end do nothing
    // but the exchange, and the unrolled body stays free of branches around
    // its loads
    for (int i0 = 0; i0 < Nmax; i0 += kRing) {
#pragma unroll
      for (int k = 0; k < kRing; k++) {
        const int i = i0 + k;
        forward(i, ring[k]);
        load_exact_row(rows, (size_t)pb + min(i + kRing, last), ri, ci, ring[k]);
      }
    }
  }

  // ---- bordered last row of the cyclic solve (triSolvers.cpp:966-1029) ----
  // Lane (ri, ci) carries Bl[ri][ci], Fl[ri] and row ri of Cl; Cl starts as
  // the normalised C of the last row, whose column ci this lane holds.  Only
  // cyclic chains take part: their lengths are exchanged like N.
  if (q == 0) xn[slot] = per ? N : 0;
  wave_lds_sync();
  int Mmax = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) Mmax = max(Mmax, xn[k] - 1);
  T blv = (ri == ci) ? T(1.0) : T(0.0), flv = pick4(Fp, ri), clr[4];
  X[4 * ri + ci] = pick4(Ccol, ri);
  wave_lds_sync();
#pragma unroll
  for (int l = 0; l < 4; l++) clr[l] = X[4 * ri + l];
  wave_lds_sync();
  T xl[4] = {Fp[0], Fp[1], Fp[2], Fp[3]};
  if (Mmax > 0) {
    const int M = per ? N - 1 : 0;
    auto border = [&](int i, const BorderRow<T>& o) __attribute__((always_inline)) {
      const bool act = i < M;
      T t = T(0);
      if (act) {
#pragma unroll
        for (int l = 0; l < 4; l++) {
          t -= clr[l] * o.c[l];
          blv -= clr[l] * o.a[l];
        }
#pragma unroll
        for (int j = 0; j < 4; j++) flv -= clr[j] * o.f[j];
      }
      X[4 * ri + ci] = t;
      wave_lds_sync();
      if (act) {
#pragma unroll
        for (int l = 0; l < 4; l++) clr[l] = X[4 * ri + l];
      }
      wave_lds_sync();
    };
    const int lastm = max(M - 1, 0);
    BorderRow<T> ring[kRing];
#pragma unroll
    for (int k = 0; k < kRing; k++) load_border_row(scr, (size_t)pb + min(k, lastm), ci, ring[k]);
    for (int i0 = 0; i0 < Mmax; i0 += kRing) {
#pragma unroll
      for (int k = 0; k < kRing; k++) {
        const int i = i0 + k;
        border(i, ring[k]);
        load_border_row(scr, (size_t)pb + min(i + kRing, lastm), ci, ring[k]);
      }
    }
    X[4 * ri + ci] = blv;
    if (ci == 0) X[16 + ri] = flv;
    wave_lds_sync();
    T Bl[16], Fl[4];
#pragma unroll
    for (int e = 0; e < 16; e++) Bl[e] = X[e];
#pragma unroll
    for (int l = 0; l < 4; l++) Fl[l] = X[16 + l];
    wave_lds_sync();
    if (per) {
      const LU4<T> L = lu4(Bl);
      lu4_solve(L, Fl, 1);
#pragma unroll
      for (int l = 0; l < 4; l++) xl[l] = Fl[l];
    }
  }

  // ---- back sweep: lane (ri, ci) computes x[ri] of each row, every lane
  // of the chain then reads the whole x ----
  auto put = [&](int i, const T* x) __attribute__((always_inline)) {
    if (q == 0) store_solution(r, dq, dqmode, ct.pleft[pb + i], x);
  };
  if (N > 0) put(N - 1, xl);
  T x[4] = {xl[0], xl[1], xl[2], xl[3]};
  // row i = N - 2 - s at step s
  const int S = max(N - 1, 0);
  int Smax = 0;
  if (q == 0) xn[slot] = S;
  wave_lds_sync();
#pragma unroll
  for (int k = 0; k < 4; k++) Smax = max(Smax, xn[k]);
  auto back = [&](int s, const BackRow<T>& o) __attribute__((always_inline)) {
    const bool act = s < S;
    const int i = N - 2 - s;
    T v = o.f;
    if (!per || i == N - 2) {
#pragma unroll
      for (int j = 0; j < 4; j++) v -= o.c[j] * x[j];
    } else {
#pragma unroll
      for (int j = 0; j < 4; j++) v -= (o.c[j] * x[j] + o.a[j] * xl[j]);
    }
    if (ci == 0) X[ri] = v;
    wave_lds_sync();
    if (act) {
#pragma unroll
      for (int l = 0; l < 4; l++) x[l] = X[l];
      put(i, x);
    }
    wave_lds_sync();
  };
  {
    // row N-2-s, clamped into [0, N-2] (row 0 for chains without a back sweep)
    auto brow = [&](int s) { return (size_t)pb + max(N - 2 - min(s, max(S - 1, 0)), 0); };
    BackRow<T> ring[kRing];
#pragma unroll
    for (int k = 0; k < kRing; k++) load_back_row(scr, brow(k), ri, per, ring[k]);
    for (int s0 = 0; s0 < Smax; s0 += kRing) {
#pragma unroll
      for (int k = 0; k < kRing; k++) {
        const int s = s0 + k;
        back(s, ring[k]);
        load_back_row(scr, brow(s + kRing), ri, per, ring[k]);
      }
    }
  }
}

// --------------------------------------------------------------------------
// ADI phase, FAST order: one workgroup per chain.  The rows are assembled
// exactly as in the reference (assemble_row), one thread per row, and
// normalised by their diagonal block (A^ = B^-1 A, C^ = B^-1 C, F^ = B^-1 F).
// Then `ncr` levels of block cyclic reduction (compacted: thread t updates
// surviving row t*2d) halve the system, parallel cyclic reduction eliminates
// the remaining rows' neighbours at strides d, 2d, 4d, ... until every row is
// decoupled (open chains), the cyclic subsystems have two rows (solved
// directly) or an odd number of rows (solved with the cyclic block Thomas of
// triSolvers.cpp:850-1046 on the normalised rows), and back substitution
// walks the cyclic-reduction levels in reverse.  The solution equals the
// EXACT ordering up to rounding.  LDS holds the 36 normalised entries of every
// row, element-major with a pad slot per 32 rows: lds[e*S + lpad(i)].
// --------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void mm4_acc(T* X, const T* A, const T* B) {
  // X -= A*B
#pragma clang fp contract(fast)
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      T x = X[4 * k + j];
#pragma unroll
      for (int l = 0; l < 4; l++) x -= A[4 * k + l] * B[4 * l + j];
      X[4 * k + j] = x;
    }
}

template <typename T>
__device__ __forceinline__ void mv4_acc(T* x, const T* A, const T* y) {
#pragma clang fp contract(fast)
#pragma unroll
  for (int k = 0; k < 4; k++) {
    T v = x[k];
#pragma unroll
    for (int j = 0; j < 4; j++) v -= A[4 * k + j] * y[j];
    x[k] = v;
  }
}

// normalise a row in registers: a := B^-1 a, c := B^-1 c, f := B^-1 f
template <typename T>
__device__ __forceinline__ void normalise_row(const T* B, T* a, T* c, T* f) {
#pragma clang fp contract(fast)
  const LU4<T> L = lu4_fast(B);
#pragma unroll
  for (int j = 0; j < 4; j++) {
    lu4_solve_fast(L, a + j, 4);
    lu4_solve_fast(L, c + j, 4);
  }
  lu4_solve_fast(L, f, 1);
}

// LDS row layout of the FAST solve: element e of row i at lds[e*S + lpad(i)],
// S = lds_stride(N).  One pad slot per 32 rows makes the row strides of the
// reduction levels (powers of two up to 32) free of bank conflicts.
constexpr __host__ __device__ __forceinline__ int lpad(int i) { return i + (i >> 5); }
constexpr __host__ __device__ __forceinline__ int lds_stride(int N) { return N + (N >> 5) + 1; }

// LDS of a gfx950 workgroup (160 KiB)
constexpr int kLdsBytes = 160 * 1024;
// How many odd rows per chain k_adi_fast_half2 keeps in the LDS its two
// slices leave free (36 entries each, stride one more): FP64 84 of at most
// 192, FP32 all of them.  The rest go through global scratch.
template <typename T, int MAXH>
constexpr __host__ __device__ int half2_odd_lds() {
  constexpr int free_ent = kLdsBytes / (int)sizeof(T) - 2 * 36 * lds_stride((MAXH + 1) / 2);
  constexpr int ks = free_ent / (2 * 36);
  return ks - 1 < MAXH / 2 ? ks - 1 : MAXH / 2;
}

template <typename T>
__device__ __forceinline__ void lds_load_row(const T* lds, int N, int i, T* a, T* c, T* f) {
#pragma unroll
  for (int e = 0; e < 16; e++) {
    a[e] = lds[e * N + i];
    c[e] = lds[(16 + e) * N + i];
  }
#pragma unroll
  for (int e = 0; e < 4; e++) f[e] = lds[(32 + e) * N + i];
}

template <typename T>
__device__ __forceinline__ void lds_store_row(T* lds, int N, int i, const T* a, const T* c, const T* f) {
#pragma unroll
  for (int e = 0; e < 16; e++) {
    lds[e * N + i] = a[e];
    lds[(16 + e) * N + i] = c[e];
  }
#pragma unroll
  for (int e = 0; e < 4; e++) lds[(32 + e) * N + i] = f[e];
}

// X -= A * (LDS matrix at element offset e0, row i); A in registers
template <typename T>
__device__ __forceinline__ void mm4_acc_lds(T* X, const T* A, const T* lds, int N, int e0, int i) {
#pragma clang fp contract(fast)
#pragma unroll
  for (int l = 0; l < 4; l++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const T b = lds[(e0 + 4 * l + j) * N + i];
#pragma unroll
      for (int k = 0; k < 4; k++) X[4 * k + j] -= A[4 * k + l] * b;
    }
}

template <typename T>
__device__ __forceinline__ void mv4_acc_lds(T* x, const T* A, const T* lds, int N, int i) {
#pragma clang fp contract(fast)
#pragma unroll
  for (int l = 0; l < 4; l++) {
    const T b = lds[(32 + l) * N + i];
#pragma unroll
    for (int k = 0; k < 4; k++) x[k] -= A[4 * k + l] * b;
  }
}

enum { kFinalOpen = 0, kFinalPair = 1, kFinalOdd = 2 };

// lane J's value of the calling lane's quad (lanes 4g..4g+3), by DPP
// quad_perm; the whole quad must be active
template <int J>
__device__ __forceinline__ float quad_get(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), J * 0x55, 0xF, 0xF, false));
}
template <int J>
__device__ __forceinline__ double quad_get(double v) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(u & 0xffffffffLL), J * 0x55, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), J * 0x55, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// One parallel-cyclic-reduction step of the normalised row in LDS slot sO,
// spread over a quad of lanes: lane q computes column q of the new blocks.
// M = I - A C_L - C A_R (column q here, the full M gathered from the quad),
// F' = M^-1 (F - A F_L - C F_R) (every lane), A'[:, q] = -M^-1 A A_L[:, q],
// C'[:, q] = -M^-1 C C_R[:, q]: every entry takes the operations
// row_reduce_lds gives it.  A missing neighbour reads slot 0 as zeros.  The
// caller writes (Ac, Cc, F[q]) back after a barrier.
template <typename T>
__device__ __forceinline__ void pcr_quad_row(const T* lds, int S, int sO, bool hasL, int sL, bool hasR, int sR,
                                             int q, T* Ac, T* Cc, T* F) {
#pragma clang fp contract(fast)
  sL = hasL ? sL : 0;
  sR = hasR ? sR : 0;
  T A[16], C[16];
  lds_load_row(lds, S, sO, A, C, F);
  // the identity column from an opaque copy of q: hoisted out of the levels'
  // loop, the four selects would hold 8 VGPRs for the whole kernel
  int qo = q;
  asm volatile("" : "+v"(qo));
  T m[4];
#pragma unroll
  for (int k = 0; k < 4; k++) m[k] = (k == qo) ? T(1) : T(0);
#pragma unroll
  for (int l = 0; l < 4; l++) {
    const T v = lds[(16 + 4 * l + q) * S + sL];  // C_L[l][q]
    const T b = hasL ? v : T(0);
#pragma unroll
    for (int k = 0; k < 4; k++) m[k] -= A[4 * k + l] * b;
  }
#pragma unroll
  for (int l = 0; l < 4; l++) {
    const T v = lds[(32 + l) * S + sL];
    const T b = hasL ? v : T(0);
#pragma unroll
    for (int k = 0; k < 4; k++) F[k] -= A[4 * k + l] * b;
  }
#pragma unroll
  for (int l = 0; l < 4; l++) {
    const T v = lds[(4 * l + q) * S + sR];  // A_R[l][q]
    const T b = hasR ? v : T(0);
#pragma unroll
    for (int k = 0; k < 4; k++) m[k] -= C[4 * k + l] * b;
  }
#pragma unroll
  for (int l = 0; l < 4; l++) {
    const T v = lds[(32 + l) * S + sR];
    const T b = hasR ? v : T(0);
#pragma unroll
    for (int k = 0; k < 4; k++) F[k] -= C[4 * k + l] * b;
  }
  __builtin_amdgcn_sched_barrier(0);
  T M[16];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    M[4 * k + 0] = quad_get<0>(m[k]);
    M[4 * k + 1] = quad_get<1>(m[k]);
    M[4 * k + 2] = quad_get<2>(m[k]);
    M[4 * k + 3] = quad_get<3>(m[k]);
  }
  const LU4<T> L = lu4_fast(M);
  lu4_solve_fast(L, F, 1);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int k = 0; k < 4; k++) Ac[k] = T(0);
#pragma unroll
  for (int l = 0; l < 4; l++) {
    const T v = lds[(4 * l + q) * S + sL];  // A_L[l][q]
    const T b = hasL ? v : T(0);
#pragma unroll
    for (int k = 0; k < 4; k++) Ac[k] -= A[4 * k + l] * b;
  }
  lu4_solve_fast(L, Ac, 1);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int k = 0; k < 4; k++) Cc[k] = T(0);
#pragma unroll
  for (int l = 0; l < 4; l++) {
    const T v = lds[(16 + 4 * l + q) * S + sR];  // C_R[l][q]
    const T b = hasR ? v : T(0);
#pragma unroll
    for (int k = 0; k < 4; k++) Cc[k] -= C[4 * k + l] * b;
  }
  lu4_solve_fast(L, Cc, 1);
}

// Reduction depths of a FAST launch in one int: the cyclic-reduction levels in
// the low byte; above it, one more than the parallel-cyclic-reduction levels
// an open chain runs before its Thomas finish (0: none, PCR to the end).
constexpr __host__ __device__ __forceinline__ int depth_code(int ncr, int tpcr) { return ncr | ((tpcr + 1) << 8); }
constexpr __host__ __device__ __forceinline__ int cr_levels(int code) { return code & 0xff; }
constexpr __host__ __device__ __forceinline__ int pcr_cap(int code) { return (code >> 8) - 1; }
constexpr int kMaxThomasPcr = 4;  // at most 16 subsystems: one wave's quads hold them all

// Thomas finish of an open chain after parallel cyclic reduction has run to
// stride dd over the members at multiples of d: the rows r_j + k*dd
// (r_j = j*d, j < dd/d) form dd/d independent block-tridiagonal systems of
// normalised rows x_k + a_k x_(k-1) + c_k x_(k+1) = f_k, with a = 0 on the
// first row of each and c = 0 on its last (their couplings fell past the
// chain's ends).  Quad j (lane q) runs the block Thomas algorithm on system j:
// lane q carries column q of gamma_k = D_k^-1 c_k, D_k = I - a_k gamma_(k-1),
// and all of phi_k = D_k^-1 (f_k - a_k phi_(k-1)) (D_0 = I), written over
// column q of c_k and f_k[q]; then x_k = phi_k - gamma_k x_(k+1), lane q
// computing x_k[q] into the F slot.  The whole quad must call it.
template <typename T>
__device__ __forceinline__ void thomas_quad_open(T* lds, int S, int N, int d, int dd, int j, int q) {
#pragma clang fp contract(fast)
  const int r0 = j * d;
  if (r0 >= N) return;
  const int n = (N - 1 - r0) / dd + 1;
  T G[4], P[4];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    G[m] = lds[(16 + 4 * m + q) * S + lpad(r0)];
    P[m] = lds[(32 + m) * S + lpad(r0)];
  }
  for (int k = 1; k < n; k++) {
    const int s = lpad(r0 + k * dd);
    T a[16], cq[4], f[4], Dq[4];
#pragma unroll
    for (int e = 0; e < 16; e++) a[e] = lds[e * S + s];
    int qo = q;  // opaque, as in pcr_quad_row
    asm volatile("" : "+v"(qo));
#pragma unroll
    for (int m = 0; m < 4; m++) {
      cq[m] = lds[(16 + 4 * m + q) * S + s];
      f[m] = lds[(32 + m) * S + s];
      Dq[m] = (m == qo) ? T(1) : T(0);
    }
#pragma unroll
    for (int l = 0; l < 4; l++)
#pragma unroll
      for (int m = 0; m < 4; m++) {
        Dq[m] -= a[4 * m + l] * G[l];  // D[:, q]
        f[m] -= a[4 * m + l] * P[l];
      }
    T D[16];
#pragma unroll
    for (int m = 0; m < 4; m++) {
      D[4 * m + 0] = quad_get<0>(Dq[m]);
      D[4 * m + 1] = quad_get<1>(Dq[m]);
      D[4 * m + 2] = quad_get<2>(Dq[m]);
      D[4 * m + 3] = quad_get<3>(Dq[m]);
    }
    const LU4<T> L = lu4_fast(D);
    lu4_solve_fast(L, f, 1);
    lu4_solve_fast(L, cq, 1);
#pragma unroll
    for (int m = 0; m < 4; m++) {
      lds[(16 + 4 * m + q) * S + s] = cq[m];
      G[m] = cq[m];
      P[m] = f[m];
    }
#pragma unroll
    for (int m = 0; m < 4; m++)
      if (q == m) lds[(32 + m) * S + s] = f[m];
  }
  // the quad's lanes wrote the gamma columns the back sweep reads by rows: a
  // wavefront's LDS operations complete in order; this keeps the compiler
  // from moving them
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  T x[4] = {P[0], P[1], P[2], P[3]};  // x_(n-1) = phi_(n-1), already in its F slot
  for (int k = n - 2; k >= 0; k--) {
    const int s = lpad(r0 + k * dd);
    T xq = lds[(32 + q) * S + s];
#pragma unroll
    for (int l = 0; l < 4; l++) xq -= lds[(16 + 4 * q + l) * S + s] * x[l];
    lds[(32 + q) * S + s] = xq;
    x[0] = quad_get<0>(xq);
    x[1] = quad_get<1>(xq);
    x[2] = quad_get<2>(xq);
    x[3] = quad_get<3>(xq);
  }
}

// The Thomas finish by the quads of the first tw waves of a chain's group of
// threads (i: the thread's index in the group): system j on wave j % tw,
// lanes 4*(j / tw) to 4*(j / tw) + 3, so that tw waves (on tw SIMDs) share the
// sequential work.  Needs dd/d <= 16*tw.
template <typename T>
__device__ __forceinline__ void thomas_finish_open(T* lds, int S, int N, int d, int dd, int i, int tw) {
  const int w = i >> 6, lane = i & 63;
  if (w >= tw) return;
  const int j = (lane >> 2) * tw + w;
  if (j < dd / d) thomas_quad_open(lds, S, N, d, dd, j, lane & 3);
}

// the quad's results into row slot sO: column q of A' and C', and F'[q]
template <typename T>
__device__ __forceinline__ void pcr_quad_store(T* lds, int S, int sO, int q, const T* Ac, const T* Cc, const T* F) {
#pragma unroll
  for (int k = 0; k < 4; k++) {
    lds[(4 * k + q) * S + sO] = Ac[k];
    lds[(16 + 4 * k + q) * S + sO] = Cc[k];
  }
  // one predicated store per entry (a select chain on q turns into a
  // dynamically indexed private array)
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (q == k) lds[(32 + k) * S + sO] = F[k];
}

#ifdef HAMSTR_LONG_TU
// The long-chain path (chains longer than a workgroup) is compiled in a
// translation unit of its own, adi_long.hip, which includes this file with
// HAMSTR_LONG_TU defined: in one module with k_adi_fast it costs that kernel
// 17 VGPRs and 2-4% of its time.
// row i of chain ch assembled (assemble_row) and normalised by its diagonal
// block; an open chain's end rows lose the coupling past the chain's ends
template <typename T>
__device__ __forceinline__ void fast_row(const ChainTab& ct, const T* __restrict__ ff, int nProc, int pb, int nf,
                                         int N, bool periodic, int i, double cfl, const double* __restrict__ sigma,
                                         const double* __restrict__ r, T* a, T* c, T* f) {
  T B[16];
  int cell;
  assemble_row<T>(ct, ff, nProc, pb, nf, N, periodic, i, cfl, sigma, r, a, B, c, f, &cell);
  if (!periodic) {
    if (i == 0) {
#pragma unroll
      for (int e = 0; e < 16; e++) a[e] = T(0);
    }
    if (i == N - 1) {
#pragma unroll
      for (int e = 0; e < 16; e++) c[e] = T(0);
    }
  }
  normalise_row(B, a, c, f);
}

// One cyclic-reduction elimination, the row updated in place (A, C, F in
// registers): it absorbs the rows at slots sL and sR of `src` (element-major,
// stride SS): M = I - A C_L - C A_R, then A <- -M^-1 A A_L, C <- -M^-1 C C_R
// and F <- M^-1 (F - A F_L - C F_R), every element computed by the
// operations (and in the order) the one-row kernel's reduction levels use.  The order keeps M's factors, F, one new block and at
// most both old blocks live (68 values), and the scheduling barriers keep the
// neighbours' loads of one stage from going in flight during another, so the
// kernels that use it fit in 168 VGPRs.
// X[:, j] -= A * (column j of the LDS matrix at element offset e0, row s);
// with has == false the column reads as zeros (row s must still be in range)
template <typename T>
__device__ __forceinline__ void mcol_acc_lds(T* X, const T* A, const T* src, int SS, int e0, int s, int j, bool has) {
#pragma clang fp contract(fast)
#pragma unroll
  for (int l = 0; l < 4; l++) {
    const T v = src[(e0 + 4 * l + j) * SS + s];
    const T b = has ? v : T(0);
#pragma unroll
    for (int k = 0; k < 4; k++) X[4 * k + j] -= A[4 * k + l] * b;
  }
}

// x -= A * (the F slots of LDS row s), zeros with has == false
template <typename T>
__device__ __forceinline__ void mv4_acc_lds_m(T* x, const T* A, const T* src, int SS, int s, bool has) {
#pragma clang fp contract(fast)
#pragma unroll
  for (int l = 0; l < 4; l++) {
    const T v = src[(32 + l) * SS + s];
    const T b = has ? v : T(0);
#pragma unroll
    for (int k = 0; k < 4; k++) x[k] -= A[4 * k + l] * b;
  }
}

// A missing neighbour (hasL / hasR false) reads row 0 of src and counts as a
// zero block: the products then add exact zeros, as skipping them would (up
// to the sign of a zero).  With no branch around the products the compiler
// keeps no second copy of the row: in a branch the same code needs 30 more
// VGPRs.
template <typename T>
__device__ __forceinline__ void row_reduce_lds(T* A, T* C, T* F, const T* src, int SS, bool hasL, int sL, bool hasR,
                                               int sR) {
#pragma clang fp contract(fast)
  sL = hasL ? sL : 0;
  sR = hasR ? sR : 0;
  T M[16];
#pragma unroll
  for (int e = 0; e < 16; e++) M[e] = (e % 5 == 0) ? T(1) : T(0);
#pragma unroll
  for (int j = 0; j < 4; j += 2) {
    mcol_acc_lds(M, A, src, SS, 16, sL, j, hasL);  // M -= A C_L
    mcol_acc_lds(M, A, src, SS, 16, sL, j + 1, hasL);
    __builtin_amdgcn_sched_barrier(0);
  }
  mv4_acc_lds_m(F, A, src, SS, sL, hasL);  // F -= A F_L
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int j = 0; j < 4; j += 2) {
    mcol_acc_lds(M, C, src, SS, 0, sR, j, hasR);  // M -= C A_R
    mcol_acc_lds(M, C, src, SS, 0, sR, j + 1, hasR);
    __builtin_amdgcn_sched_barrier(0);
  }
  mv4_acc_lds_m(F, C, src, SS, sR, hasR);  // F -= C F_R
  __builtin_amdgcn_sched_barrier(0);
  const LU4<T> L = lu4_fast(M);
  lu4_solve_fast(L, F, 1);
  T X[16];
#pragma unroll
  for (int e = 0; e < 16; e++) X[e] = T(0);
#pragma unroll
  for (int j = 0; j < 4; j++) {
    __builtin_amdgcn_sched_barrier(0);
    mcol_acc_lds(X, A, src, SS, 0, sL, j, hasL);  // -A A_L
    lu4_solve_fast(L, X + j, 4);
  }
#pragma unroll
  for (int e = 0; e < 16; e++) {
    A[e] = X[e];
    X[e] = T(0);
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    __builtin_amdgcn_sched_barrier(0);
    mcol_acc_lds(X, C, src, SS, 16, sR, j, hasR);  // -C C_R
    lu4_solve_fast(L, X + j, 4);
  }
#pragma unroll
  for (int e = 0; e < 16; e++) C[e] = X[e];
}

// Solves the N-row normalised system held in LDS (row i at lds[e*S + lpad(i)])
// in place; the solution ends in the F slots.  Every thread of the workgroup
// calls it (it holds barriers); thread i works on the rows each step assigns.
// S = SFIX when SFIX > 0 (a stride fixed at compile time, so that every
// element of a row is one immediate offset from one address), else
// lds_stride(N).  The reduction levels run as for an NL-row system (NL >= N;
// NL == N for closed chains), so that groups of threads solving open chains
// of different lengths side by side meet the same barriers: the levels an
// N-row system does not need leave its rows without neighbours, which only
// renormalises them by the identity.  QUAD: run the parallel cyclic
// reduction by quads of lanes where the threads allow it (the half-LDS
// kernels: their 96-member PCR levels otherwise keep 1.5 waves of a chain
// busy; the two-row long-chain kernel has no threads to spare).
template <typename T, int FINAL, int SFIX = 0, bool QUAD = true>
__device__ __forceinline__ void fast_solve_lds(T* lds, const int N, const int NL, const int ncr, const int tpcr,
                                               const int i, const int NT, const int tw) {
#pragma clang fp contract(fast)
  const bool periodic = FINAL != kFinalOpen;
  const int S = SFIX > 0 ? SFIX : lds_stride(N);

  // (1) ncr levels of cyclic reduction: at stride d the survivors (rows at
  // multiples of 2d) absorb rows i - d and i + d, which drop out.  The work is
  // compacted: thread t updates survivor row t*2d, so only the first
  // ceil(survivors/64) waves issue instructions.  A row is read only at the
  // level that eliminates it, so one barrier per level suffices.
  int d = 1;  // stride between the rows still in the system
  int n = N;  // cycle length of the remaining system (closed chains)
  int ncr_done = 0;
  while (ncr_done < ncr && (periodic ? (n % 2 == 0 && n >= 4) : (d < NL))) {
    const int nsurv = (N + 2 * d - 1) / (2 * d);
    if (i < nsurv) {
      const int row = i * 2 * d;
      const int iL = periodic ? (row - d + N) % N : row - d;
      const int iR = periodic ? (row + d) % N : row + d;
      T ra[16], rc[16], rf[4];
      lds_load_row(lds, S, lpad(row), ra, rc, rf);
      row_reduce_lds(ra, rc, rf, lds, S, periodic || iL >= 0, lpad(iL), periodic || iR < N, lpad(iR));
      lds_store_row(lds, S, lpad(row), ra, rc, rf);
    }
    __syncthreads();
    d *= 2;
    n /= 2;
    ncr_done++;
  }

  // (2) parallel cyclic reduction of the rows at multiples of d (members):
  // by quads of lanes (pcr_quad_row) when QUAD and the NT threads give four
  // per member (a uniform choice: nmem follows from NL), else thread t
  // holding member row t*d in registers
  const int nmem = (N + d - 1) / d;
  const bool member = i < nmem;
  const int mrow = i * d;
  T a[16], c[16], f[4];
  int dd = d;
  // open chains with a Thomas finish stop after tpcr levels
  const int pcap = (FINAL == kFinalOpen && tpcr >= 0) ? tpcr : (1 << 30);
  int npl = 0;
  if (QUAD && 4 * ((NL + d - 1) / d) <= NT) {
    const int g = i >> 2, q = i & 3, grow = g * d;
    while ((periodic ? (n % 2 == 0 && n >= 4) : (dd < NL)) && npl < pcap) {
      T Ac[4], Cc[4], F[4];
      if (g < nmem) {
        const int iL = periodic ? (grow - dd + N) % N : grow - dd;
        const int iR = periodic ? (grow + dd) % N : grow + dd;
        pcr_quad_row(lds, S, lpad(grow), periodic || iL >= 0, lpad(iL), periodic || iR < N, lpad(iR), q, Ac, Cc, F);
      }
      __syncthreads();
      if (g < nmem) pcr_quad_store(lds, S, lpad(grow), q, Ac, Cc, F);
      __syncthreads();
      dd *= 2;
      n /= 2;
      npl++;
    }
  }
  if (member) lds_load_row(lds, S, lpad(mrow), a, c, f);
  while ((periodic ? (n % 2 == 0 && n >= 4) : (dd < NL)) && npl < pcap) {
    if (member) {
      const int iL = periodic ? (mrow - dd + N) % N : mrow - dd;
      const int iR = periodic ? (mrow + dd) % N : mrow + dd;
      row_reduce_lds(a, c, f, lds, S, periodic || iL >= 0, lpad(iL), periodic || iR < N, lpad(iR));
    }
    __syncthreads();
    if (member) lds_store_row(lds, S, lpad(mrow), a, c, f);
    __syncthreads();
    dd *= 2;
    n /= 2;
    npl++;
  }

  // (3) the members' final systems; their solutions go to their F slots
  if (FINAL == kFinalOpen) {
    // decoupled (x = f, already in the F slot), or after tpcr levels the
    // subsystems at stride dd, by the Thomas finish
    if (dd < NL) {
      thomas_finish_open(lds, S, N, d, dd, i, tw);
      __syncthreads();
    }
  } else if (FINAL == kFinalPair) {
    // rows i and j = i + N/2 couple only to each other:
    // x_i = (I - P Q)^-1 (f_i - P f_j), P = A^_i + C^_i, Q = A^_j + C^_j
    T x[4];
    if (member) {
      const int j = (mrow + N / 2) % N;
      T P[16], Q[16], M[16];
#pragma unroll
      for (int e = 0; e < 16; e++) {
        P[e] = a[e] + c[e];
        Q[e] = lds[e * S + lpad(j)] + lds[(16 + e) * S + lpad(j)];
        M[e] = (e % 5 == 0) ? T(1) : T(0);
      }
      mm4_acc(M, P, Q);
#pragma unroll
      for (int e = 0; e < 4; e++) x[e] = f[e];
      mv4_acc_lds(x, P, lds, S, lpad(j));
      const LU4<T> L = lu4_fast(M);
      lu4_solve_fast(L, x, 1);
    }
    __syncthreads();
    if (member) {
#pragma unroll
      for (int e = 0; e < 4; e++) lds[(32 + e) * S + lpad(mrow)] = x[e];
    }
    __syncthreads();
  } else {
    // odd cycles {r, r+dd, r+2dd, ...} of n rows: the leader r < dd runs the
    // cyclic block Thomas on the normalised rows (B = I) in place
    if (member && mrow < dd) {
      T Cp[16], Fp[4], ap[16];
      for (int k = 0; k < n; k++) {
        const int rk = mrow + k * dd;
        T A[16], Bm[16], C[16], F[4], sp[16];
        lds_load_row(lds, S, lpad(rk), A, C, F);
#pragma unroll
        for (int e = 0; e < 16; e++) Bm[e] = (e % 5 == 0) ? T(1) : T(0);
        if (k != 0) {
          mm4_acc(Bm, A, Cp);
          mv4_acc(F, A, Fp);
#pragma unroll
          for (int e = 0; e < 16; e++) sp[e] = T(0);
          mm4_acc(sp, A, ap);
        } else {
#pragma unroll
          for (int e = 0; e < 16; e++) sp[e] = A[e];
        }
        const LU4<T> L = lu4_fast(Bm);
        lu4_solve_fast(L, F, 1);
        if (k != n - 2) {
#pragma unroll
          for (int j = 0; j < 4; j++) {
            lu4_solve_fast(L, C + j, 4);
            lu4_solve_fast(L, sp + j, 4);
          }
        } else {
#pragma unroll
          for (int e = 0; e < 16; e++) C[e] += sp[e];
#pragma unroll
          for (int j = 0; j < 4; j++) lu4_solve_fast(L, C + j, 4);
#pragma unroll
          for (int e = 0; e < 16; e++) sp[e] = C[e];
        }
        lds_store_row(lds, S, lpad(rk), sp, C, F);  // row k holds (spike a', C', F')
#pragma unroll
        for (int e = 0; e < 16; e++) {
          Cp[e] = C[e];
          ap[e] = sp[e];
        }
#pragma unroll
        for (int e = 0; e < 4; e++) Fp[e] = F[e];
      }
      T Bl[16], Cl[16], Fl[4];
#pragma unroll
      for (int e = 0; e < 16; e++) {
        Bl[e] = (e % 5 == 0) ? T(1) : T(0);
        Cl[e] = Cp[e];
      }
#pragma unroll
      for (int e = 0; e < 4; e++) Fl[e] = Fp[e];
      for (int k = 0; k < n - 1; k++) {
        const int rk = mrow + k * dd;
        T tt[16];
#pragma unroll
        for (int e = 0; e < 16; e++) tt[e] = T(0);
        mm4_acc_lds(tt, Cl, lds, S, 16, lpad(rk));
        mm4_acc_lds(Bl, Cl, lds, S, 0, lpad(rk));
        mv4_acc_lds(Fl, Cl, lds, S, lpad(rk));
#pragma unroll
        for (int e = 0; e < 16; e++) Cl[e] = tt[e];
      }
      const LU4<T> L = lu4_fast(Bl);
      lu4_solve_fast(L, Fl, 1);
      T xn[4] = {Fl[0], Fl[1], Fl[2], Fl[3]};
      // back sweep; the solution overwrites the F slots
      for (int k = n - 1; k >= 0; k--) {
        const int rk = mrow + k * dd;
        T sf[4];
        if (k == n - 1) {
#pragma unroll
          for (int e = 0; e < 4; e++) sf[e] = Fl[e];
        } else {
          T sc[16];
#pragma unroll
          for (int e = 0; e < 4; e++) sf[e] = lds[(32 + e) * S + lpad(rk)];
#pragma unroll
          for (int e = 0; e < 16; e++) sc[e] = lds[(16 + e) * S + lpad(rk)];
          mv4_acc(sf, sc, xn);
          if (k != n - 2) {
#pragma unroll
            for (int e = 0; e < 16; e++) sc[e] = lds[e * S + lpad(rk)];
            mv4_acc(sf, sc, Fl);
          }
        }
#pragma unroll
        for (int e = 0; e < 4; e++) {
          xn[e] = sf[e];
          lds[(32 + e) * S + lpad(rk)] = sf[e];
        }
      }
    }
    __syncthreads();
  }

  // (4) back substitution through the cyclic-reduction levels, compacted:
  // thread t takes row (2t+1)*dl, eliminated at stride dl, whose LDS row
  // still holds its coefficients as eliminated: x = F^ - A^ x_(i-dl) - C^ x_(i+dl)
  for (int l = ncr_done - 1; l >= 0; l--) {
    const int dl = 1 << l;
    const int nel = ((N + dl - 1) / dl) / 2;  // odd multiples of dl below N
    if (i < nel) {
      const int row = (2 * i + 1) * dl;
      const int iL = periodic ? (row - dl + N) % N : row - dl;
      const int iR = periodic ? (row + dl) % N : row + dl;
      T ra[16], rc[16], rf[4];
      lds_load_row(lds, S, lpad(row), ra, rc, rf);
      if (periodic || iL >= 0) mv4_acc_lds(rf, ra, lds, S, lpad(iL));
      if (periodic || iR < N) mv4_acc_lds(rf, rc, lds, S, lpad(iR));
#pragma unroll
      for (int e = 0; e < 4; e++) lds[(32 + e) * S + lpad(row)] = rf[e];
    }
    __syncthreads();
  }
}

#endif  // HAMSTR_LONG_TU

// Row m assembled as assemble_row does and normalised by its diagonal block
// (an open chain's end rows without the coupling past the chain's ends), for
// a chain whose rows are all assembled at once by one group of threads, each
// row taking only its own face m (L_m, R_m) from js: face m+1, which row m
// also needs, comes through LDS (`xl`, element e of face j at xl[e * XS + j],
// 16 entries, XS > N; L first, then R), so that no face is read from memory
// or computed twice.  An open chain's last face (no row owns it) is read from
// ff by the last row (k_rhs_faces stores it in either case: a second face to
// compute would hold every row's barrier for the last row).  Row m holds A (= R_m scaled) and C
// (= 0 - L_(m+1) scaled) before the diagonal block is factored.  Orphan
// diagonal blocks are read from ff.  Every thread of the group calls it (it
// holds barriers); inactive threads pass active = false.  The entries take
// the operations assemble_row and normalise_row give them.
constexpr int kXStride = 385;  // the 384-thread half-LDS kernels' XS (MAXT + 1)
// JQ: the Jacobians computed from q (js), else read from ff (a compile-time
// choice: in one function the two paths cost the two-line kernel 160 spilled
// VGPRs)
template <typename T, bool JQ>
__device__ __forceinline__ void fast_row_xchg(const ChainTab& ct, const JacSrc& js, const T* __restrict__ ff,
                                              int nProc, int pb, int nf, int N, bool periodic, int m, bool active,
                                              double cfl, const double* __restrict__ sigma,
                                              const double* __restrict__ r, T* xl, const int XS, T* A, T* C,
                                              T* F) {
  const int pm = pb + m;
  const bool has_next = active && (periodic || m + 1 < nf);
  const int mn = periodic ? ((m + 1 == N) ? 0 : m + 1) : m + 1;  // the next face, as a row index
  const bool next_owned = periodic || m + 1 < N;                 // else the chain's last face
  const int pn = pb + mn;
  const int rc = (periodic || !active) ? 0 : ct.pright[pm];
  const int rn = (periodic || !has_next) ? 0 : ct.pright[pn];
  const bool end = !periodic && has_next && m + 1 == nf - 1 && rn < 0;
  const bool hasA = periodic || (m > 0 && rc > -1);
  const bool hasC = periodic || (has_next && !end && rn > -1 && m != N - 1);
  constexpr bool comp = JQ;
  T B[16];
  // L: own face into LDS (computed: R_m too, kept in A), the next face's back
  if (active) {
    if (comp) {
      face_jac_q<T>(js, pm, B, A);
    } else {
#pragma unroll
      for (int e = 0; e < 16; e++) B[e] = ff[(size_t)e * nProc + pm];
    }
#pragma unroll
    for (int e = 0; e < 16; e++) xl[e * XS + m] = B[e];
  }
  __syncthreads();
  if (has_next) {
#pragma unroll
    for (int e = 0; e < 16; e++) {
      const T v = next_owned ? xl[e * XS + mn] : ff[(size_t)e * nProc + pn];
      C[e] = v;  // L_(m+1)
    }
  }
  __syncthreads();
  // R: own face into LDS (A, and the wall case of B), the next face's back
  if (active) {
    if (!comp) {
#pragma unroll
      for (int e = 0; e < 16; e++) A[e] = ff[(size_t)(16 + e) * nProc + pm];
    }
#pragma unroll
    for (int e = 0; e < 16; e++) xl[e * XS + m] = A[e];
    if (!periodic && rc == -2) {
      T R[16];
#pragma unroll
      for (int e = 0; e < 16; e++) R[e] = A[e];
      reflect_cols(R, ct.refmtx + 4 * (size_t)ct.pface[pm]);
#pragma unroll
      for (int e = 0; e < 16; e++) B[e] += R[e];
    }
  }
  __syncthreads();
  if (has_next) {
    T R[16];
#pragma unroll
    for (int e = 0; e < 16; e++)
      R[e] = next_owned ? xl[e * XS + mn] : ff[(size_t)(16 + e) * nProc + pn];
    if (periodic) {
#pragma unroll
      for (int e = 0; e < 16; e++) B[e] -= R[e];
    } else if (end) {
      if (rn == -2) reflect_cols(R, ct.refmtx + 4 * (size_t)ct.pface[pn]);
#pragma unroll
      for (int e = 0; e < 16; e++) B[e] += C[e];  // L_(m+1)
      if (rn == -2) {
#pragma unroll
        for (int e = 0; e < 16; e++) B[e] += R[e];
      }
    } else if (rn > -1) {
#pragma unroll
      for (int e = 0; e < 16; e++) B[e] -= R[e];
    }
  }
  __syncthreads();  // the LDS slice is free again
  if (!active) return;
  const int cell = ct.pleft[pm];
  if (ct.od_chain && ct.od_chain[cell] == pb) {
    const int e1 = ct.od_off[cell + 1];
    for (int e = ct.od_off[cell]; e < e1; e++) {
      const int code = ct.od_ent[e];
      const int po = code >> 1;
      if (code & 1) {
#pragma unroll
        for (int k = 0; k < 16; k++) B[k] += ff[(size_t)k * nProc + po];
      } else {
#pragma unroll
        for (int k = 0; k < 16; k++) B[k] -= ff[(size_t)(16 + k) * nProc + po];
      }
    }
  }
  const T dtfac = T(cfl / sigma[cell]);
#pragma unroll
  for (int e = 0; e < 16; e++) B[e] *= dtfac;
  B[0] += T(1.0);
  B[5] += T(1.0);
  B[10] += T(1.0);
  B[15] += T(1.0);
  const LU4<T> LU = lu4_fast(B);
  const double4 rv = *reinterpret_cast<const double4*>(r + 4 * (size_t)cell);
  F[0] = T(rv.x);
  F[1] = T(rv.y);
  F[2] = T(rv.z);
  F[3] = T(rv.w);
  lu4_solve_fast(LU, F, 1);
#pragma unroll
  for (int e = 0; e < 16; e++) {
    A[e] = (hasA ? A[e] : T(0)) * dtfac;
    C[e] = (hasC ? T(0) - C[e] : T(0)) * dtfac;
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    lu4_solve_fast(LU, A + j, 4);
    lu4_solve_fast(LU, C + j, 4);
  }
}

// ADI phase, FAST order, one thread and one LDS row per block row (steps as
// in fast_solve_lds, written out in place: the shared function costs this
// kernel about 20 VGPRs and 4% of its time)
template <typename T, int FINAL, bool JQ>
__device__ __forceinline__ void adi_fast_chain(ChainTab ct, const int ch, const JacSrc& js, const T* __restrict__ ff,
                                               int nProc, double cfl, const double* __restrict__ sigma,
                                               double* __restrict__ r, double* __restrict__ dq, int dqmode,
                                               int ncr) {
#pragma clang fp contract(fast)
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  const int pb = ct.pbase[ch], nf = ct.nf[ch], N = ct.ncell[ch];
  const bool periodic = FINAL != kFinalOpen;
  const int i = threadIdx.x;
  const bool active = i < N;
  const int S = lds_stride(N);

  // every row, assembled and normalised, goes to LDS.  Jacobians computed
  // from q pass each face to the neighbouring row through LDS, so that none
  // is computed twice; read from ff, both of a row's faces are read directly
  // (passing them through LDS made this kernel slower: here the second read
  // of a face hits in cache).
  if constexpr (JQ) {
    T a[16], c[16], f[4];
    fast_row_xchg<T, true>(ct, js, ff, nProc, pb, nf, N, periodic, i, active, cfl, sigma, r, lds, S, a, c, f);
    if (active) lds_store_row(lds, S, lpad(i), a, c, f);
  } else if (active) {
    T a[16], c[16], f[4], B[16];
    int cell;
    assemble_row<T>(ct, ff, nProc, pb, nf, N, periodic, i, cfl, sigma, r, a, B, c, f, &cell);
    if (!periodic) {
      if (i == 0) {
#pragma unroll
        for (int e = 0; e < 16; e++) a[e] = T(0);
      }
      if (i == N - 1) {
#pragma unroll
        for (int e = 0; e < 16; e++) c[e] = T(0);
      }
    }
    normalise_row(B, a, c, f);
    lds_store_row(lds, S, lpad(i), a, c, f);
  }
  __syncthreads();

  // (1) ncr levels of cyclic reduction: at stride d the survivors (rows at
  // multiples of 2d) absorb rows i - d and i + d, which drop out.  The work is
  // compacted: thread t updates survivor row t*2d, so only the first
  // ceil(survivors/64) waves issue instructions.  A row is read only at the
  // level that eliminates it, so one barrier per level suffices.
  int d = 1;  // stride between the rows still in the system
  int n = N;  // cycle length of the remaining system (closed chains)
  int ncr_done = 0;
  const int ncrl = cr_levels(ncr);
  while (ncr_done < ncrl && (periodic ? (n % 2 == 0 && n >= 4) : (d < N))) {
    const int nsurv = (N + 2 * d - 1) / (2 * d);
    if (i < nsurv) {
      const int row = i * 2 * d;
      const int iL = periodic ? (row - d + N) % N : row - d;
      const int iR = periodic ? (row + d) % N : row + d;
      T ra[16], rc[16], rf[4], M[16], na[16], nc[16];
      lds_load_row(lds, S, lpad(row), ra, rc, rf);
#pragma unroll
      for (int e = 0; e < 16; e++) {
        M[e] = (e % 5 == 0) ? T(1) : T(0);
        na[e] = T(0);
        nc[e] = T(0);
      }
      if (periodic || iL >= 0) {
        mm4_acc_lds(M, ra, lds, S, 16, lpad(iL));  // M -= A^_i C^_L
        mm4_acc_lds(na, ra, lds, S, 0, lpad(iL));  // A^new = -A^_i A^_L
        mv4_acc_lds(rf, ra, lds, S, lpad(iL));     // F -= A^_i F^_L
      }
      if (periodic || iR < N) {
        mm4_acc_lds(M, rc, lds, S, 0, lpad(iR));   // M -= C^_i A^_R
        mm4_acc_lds(nc, rc, lds, S, 16, lpad(iR)); // C^new = -C^_i C^_R
        mv4_acc_lds(rf, rc, lds, S, lpad(iR));     // F -= C^_i F^_R
      }
      normalise_row(M, na, nc, rf);
      lds_store_row(lds, S, lpad(row), na, nc, rf);
    }
    __syncthreads();
    d *= 2;
    n /= 2;
    ncr_done++;
  }

  // (2) parallel cyclic reduction of the rows at multiples of d (members):
  // by quads of lanes (pcr_quad_row) when the workgroup has four threads per
  // member, else thread t holding member row t*d in registers.  Not for FP64
  // odd cycles, whose final solve leaves no registers for it (in FP32 it took
  // a 256-case batch's 384-row loops from 0.95 to 0.87 ms).
  const int nmem = (N + d - 1) / d;
  const bool member = i < nmem;
  const int mrow = i * d;
  T a[16], c[16], f[4];
  int dd = d;
  // open chains with a Thomas finish stop after pcr_cap(ncr) levels
  const int pcap = (FINAL == kFinalOpen && pcr_cap(ncr) >= 0) ? pcr_cap(ncr) : (1 << 30);
  int npl = 0;
  if ((FINAL != kFinalOdd || sizeof(T) == sizeof(float)) && 4 * nmem <= (int)blockDim.x) {
    const int g = i >> 2, q = i & 3, grow = g * d;
    while ((periodic ? (n % 2 == 0 && n >= 4) : (dd < N)) && npl < pcap) {
      T Ac[4], Cc[4], F[4];
      if (g < nmem) {
        const int iL = periodic ? (grow - dd + N) % N : grow - dd;
        const int iR = periodic ? (grow + dd) % N : grow + dd;
        pcr_quad_row(lds, S, lpad(grow), periodic || iL >= 0, lpad(iL), periodic || iR < N, lpad(iR), q, Ac, Cc, F);
      }
      __syncthreads();
      if (g < nmem) pcr_quad_store(lds, S, lpad(grow), q, Ac, Cc, F);
      __syncthreads();
      dd *= 2;
      n /= 2;
      npl++;
    }
  }
  if (member) lds_load_row(lds, S, lpad(mrow), a, c, f);
  while ((periodic ? (n % 2 == 0 && n >= 4) : (dd < N)) && npl < pcap) {
    if (member) {
      const int iL = periodic ? (mrow - dd + N) % N : mrow - dd;
      const int iR = periodic ? (mrow + dd) % N : mrow + dd;
      T M[16], na[16], nc[16];
#pragma unroll
      for (int e = 0; e < 16; e++) {
        M[e] = (e % 5 == 0) ? T(1) : T(0);
        na[e] = T(0);
        nc[e] = T(0);
      }
      if (periodic || iL >= 0) {
        mm4_acc_lds(M, a, lds, S, 16, lpad(iL));
        mm4_acc_lds(na, a, lds, S, 0, lpad(iL));
        mv4_acc_lds(f, a, lds, S, lpad(iL));
      }
      if (periodic || iR < N) {
        mm4_acc_lds(M, c, lds, S, 0, lpad(iR));
        mm4_acc_lds(nc, c, lds, S, 16, lpad(iR));
        mv4_acc_lds(f, c, lds, S, lpad(iR));
      }
      normalise_row(M, na, nc, f);
#pragma unroll
      for (int e = 0; e < 16; e++) {
        a[e] = na[e];
        c[e] = nc[e];
      }
    }
    __syncthreads();
    if (member) lds_store_row(lds, S, lpad(mrow), a, c, f);
    __syncthreads();
    dd *= 2;
    n /= 2;
    npl++;
  }

  // (3) the members' final systems; their solutions go to their F slots
  if (FINAL == kFinalOpen) {
    // decoupled (x = f, already in the F slot), or after pcr_cap(ncr) levels
    // the subsystems at stride dd, by the Thomas finish on up to four waves
    if (dd < N) {
      thomas_finish_open(lds, S, N, d, dd, i, max(1, min(4, (int)blockDim.x >> 6)));
      __syncthreads();
    }
  } else if (FINAL == kFinalPair) {
    // rows i and j = i + N/2 couple only to each other:
    // x_i = (I - P Q)^-1 (f_i - P f_j), P = A^_i + C^_i, Q = A^_j + C^_j
    T x[4];
    if (member) {
      const int j = (mrow + N / 2) % N;
      T P[16], Q[16], M[16];
#pragma unroll
      for (int e = 0; e < 16; e++) {
        P[e] = a[e] + c[e];
        Q[e] = lds[e * S + lpad(j)] + lds[(16 + e) * S + lpad(j)];
        M[e] = (e % 5 == 0) ? T(1) : T(0);
      }
      mm4_acc(M, P, Q);
#pragma unroll
      for (int e = 0; e < 4; e++) x[e] = f[e];
      mv4_acc_lds(x, P, lds, S, lpad(j));
      const LU4<T> L = lu4_fast(M);
      lu4_solve_fast(L, x, 1);
    }
    __syncthreads();
    if (member) {
#pragma unroll
      for (int e = 0; e < 4; e++) lds[(32 + e) * S + lpad(mrow)] = x[e];
    }
    __syncthreads();
  } else {
    // odd cycles {r, r+dd, r+2dd, ...} of n rows: the leader r < dd runs the
    // cyclic block Thomas on the normalised rows (B = I) in place
    if (member && mrow < dd) {
      T Cp[16], Fp[4], ap[16];
      for (int k = 0; k < n; k++) {
        const int rk = mrow + k * dd;
        T A[16], Bm[16], C[16], F[4], sp[16];
        lds_load_row(lds, S, lpad(rk), A, C, F);
#pragma unroll
        for (int e = 0; e < 16; e++) Bm[e] = (e % 5 == 0) ? T(1) : T(0);
        if (k != 0) {
          mm4_acc(Bm, A, Cp);
          mv4_acc(F, A, Fp);
#pragma unroll
          for (int e = 0; e < 16; e++) sp[e] = T(0);
          mm4_acc(sp, A, ap);
        } else {
#pragma unroll
          for (int e = 0; e < 16; e++) sp[e] = A[e];
        }
        const LU4<T> L = lu4_fast(Bm);
        lu4_solve_fast(L, F, 1);
        if (k != n - 2) {
#pragma unroll
          for (int j = 0; j < 4; j++) {
            lu4_solve_fast(L, C + j, 4);
            lu4_solve_fast(L, sp + j, 4);
          }
        } else {
#pragma unroll
          for (int e = 0; e < 16; e++) C[e] += sp[e];
#pragma unroll
          for (int j = 0; j < 4; j++) lu4_solve_fast(L, C + j, 4);
#pragma unroll
          for (int e = 0; e < 16; e++) sp[e] = C[e];
        }
        lds_store_row(lds, S, lpad(rk), sp, C, F);  // row k holds (spike a', C', F')
#pragma unroll
        for (int e = 0; e < 16; e++) {
          Cp[e] = C[e];
          ap[e] = sp[e];
        }
#pragma unroll
        for (int e = 0; e < 4; e++) Fp[e] = F[e];
      }
      T Bl[16], Cl[16], Fl[4];
#pragma unroll
      for (int e = 0; e < 16; e++) {
        Bl[e] = (e % 5 == 0) ? T(1) : T(0);
        Cl[e] = Cp[e];
      }
#pragma unroll
      for (int e = 0; e < 4; e++) Fl[e] = Fp[e];
      for (int k = 0; k < n - 1; k++) {
        const int rk = mrow + k * dd;
        T tt[16];
#pragma unroll
        for (int e = 0; e < 16; e++) tt[e] = T(0);
        mm4_acc_lds(tt, Cl, lds, S, 16, lpad(rk));
        mm4_acc_lds(Bl, Cl, lds, S, 0, lpad(rk));
        mv4_acc_lds(Fl, Cl, lds, S, lpad(rk));
#pragma unroll
        for (int e = 0; e < 16; e++) Cl[e] = tt[e];
      }
      const LU4<T> L = lu4_fast(Bl);
      lu4_solve_fast(L, Fl, 1);
      T xn[4] = {Fl[0], Fl[1], Fl[2], Fl[3]};
      // back sweep; the solution overwrites the F slots
      for (int k = n - 1; k >= 0; k--) {
        const int rk = mrow + k * dd;
        T sf[4];
        if (k == n - 1) {
#pragma unroll
          for (int e = 0; e < 4; e++) sf[e] = Fl[e];
        } else {
          T sc[16];
#pragma unroll
          for (int e = 0; e < 4; e++) sf[e] = lds[(32 + e) * S + lpad(rk)];
#pragma unroll
          for (int e = 0; e < 16; e++) sc[e] = lds[(16 + e) * S + lpad(rk)];
          mv4_acc(sf, sc, xn);
          if (k != n - 2) {
#pragma unroll
            for (int e = 0; e < 16; e++) sc[e] = lds[e * S + lpad(rk)];
            mv4_acc(sf, sc, Fl);
          }
        }
#pragma unroll
        for (int e = 0; e < 4; e++) {
          xn[e] = sf[e];
          lds[(32 + e) * S + lpad(rk)] = sf[e];
        }
      }
    }
    __syncthreads();
  }

  // (4) back substitution through the cyclic-reduction levels, compacted:
  // thread t takes row (2t+1)*dl, eliminated at stride dl, whose LDS row
  // still holds its coefficients as eliminated: x = F^ - A^ x_(i-dl) - C^ x_(i+dl)
  for (int l = ncr_done - 1; l >= 0; l--) {
    const int dl = 1 << l;
    const int nel = ((N + dl - 1) / dl) / 2;  // odd multiples of dl below N
    if (i < nel) {
      const int row = (2 * i + 1) * dl;
      const int iL = periodic ? (row - dl + N) % N : row - dl;
      const int iR = periodic ? (row + dl) % N : row + dl;
      T ra[16], rc[16], rf[4];
      lds_load_row(lds, S, lpad(row), ra, rc, rf);
      if (periodic || iL >= 0) mv4_acc_lds(rf, ra, lds, S, lpad(iL));
      if (periodic || iR < N) mv4_acc_lds(rf, rc, lds, S, lpad(iR));
#pragma unroll
      for (int e = 0; e < 4; e++) lds[(32 + e) * S + lpad(row)] = rf[e];
    }
    __syncthreads();
  }
  if (active) {
    T x[4];
#pragma unroll
    for (int e = 0; e < 4; e++) x[e] = lds[(32 + e) * S + lpad(i)];
    store_solution(r, dq, dqmode, ct.pleft[pb + i], x);
  }
}

#ifdef HAMSTR_LONG_TU
// ADI phase, FAST order, for chains longer than a workgroup: thread t takes
// rows 2t and 2t+1.  The first cyclic-reduction level runs before anything
// enters LDS: the odd rows go to global scratch `hs` (element-major, stride
// hstride, the chain's rows from ct.hbase[ch]), each even row absorbs its two
// odd neighbours in registers, and only the reduced even rows enter LDS, as a
// system of ceil(N/2) rows that fast_solve_lds finishes with ncr-1 more
// levels.  The odd rows then back-substitute from scratch.  The arithmetic is
// that of adi_fast_chain with ncr levels.  Needs open chains of N >= 2 rows or
// closed chains of an even N >= 4, and ncr >= 1.  hbase[ch]: the chain's first
// scratch row.
template <typename T, int FINAL>
__device__ __forceinline__ void adi_fast_chain_long(ChainTab ct, const int ch, const T* __restrict__ ff, int nProc,
                                                    double cfl, const double* __restrict__ sigma,
                                                    double* __restrict__ r, double* __restrict__ dq, int dqmode,
                                                    int ncr, T* __restrict__ hs, int hstride,
                                                    const int* __restrict__ hbase) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  const int pb = ct.pbase[ch], nf = ct.nf[ch], N = ct.ncell[ch];
  const bool periodic = FINAL != kFinalOpen;
  const int t = threadIdx.x;
  const int M = (N + 1) / 2;
  const int SM = lds_stride(M);
  T* hrow = hs + hbase[ch];
  if (2 * t + 1 < N) {
    T a[16], c[16], f[4];
    fast_row<T>(ct, ff, nProc, pb, nf, N, periodic, 2 * t + 1, cfl, sigma, r, a, c, f);
    lds_store_row(hrow, hstride, t, a, c, f);
  }
  __syncthreads();
  if (t < M) {
    const int i = 2 * t;
    const int iL = periodic ? (i - 1 + N) % N : i - 1;
    const int iR = periodic ? (i + 1) % N : i + 1;
    T a[16], c[16], f[4];
    fast_row<T>(ct, ff, nProc, pb, nf, N, periodic, i, cfl, sigma, r, a, c, f);
    row_reduce_lds(a, c, f, hrow, hstride, periodic || iL >= 0, iL >> 1, periodic || iR < N, iR >> 1);
    lds_store_row(lds, SM, lpad(t), a, c, f);
  }
  __syncthreads();
  fast_solve_lds<T, FINAL, 0, false>(lds, M, M, cr_levels(ncr) - 1, -1, t, (int)blockDim.x, 1);
  if (2 * t + 1 < N) {
    // x = F^ - A^ x_2t - C^ x_2t+2, with the even rows' x in the F slots
    const int row = 2 * t + 1;
    T ra[16], rc[16], x[4];
    lds_load_row(hrow, hstride, t, ra, rc, x);
    mv4_acc_lds(x, ra, lds, SM, lpad(t));
    if (periodic || row + 1 < N) mv4_acc_lds(x, rc, lds, SM, lpad(((row + 1) % N) >> 1));
    store_solution(r, dq, dqmode, ct.pleft[pb + row], x);
  }
  if (t < M) {
    T x[4];
#pragma unroll
    for (int e = 0; e < 4; e++) x[e] = lds[(32 + e) * SM + lpad(t)];
    store_solution(r, dq, dqmode, ct.pleft[pb + 2 * t], x);
  }
}

// ADI phase, FAST order, one row per thread with half the rows in LDS: thread
// t < M = ceil(N/2) holds even row 2t, thread t >= M odd row 2(t-M)+1.  The
// odd rows, which the first cyclic-reduction level eliminates, go to LDS and
// to global scratch `hs` (element-major, stride hstride, the chain's rows from
// hbase[ch]); each even row absorbs its two odd neighbours in registers and
// replaces them in LDS, as a system of M rows that fast_solve_lds finishes
// with ncr-1 more levels; the odd rows then back-substitute from scratch.
// Half the LDS of adi_fast_chain, at most 168 VGPRs.  The parallel cyclic
// reduction runs by quads of lanes.  The arithmetic is that of
// adi_fast_chain with ncr levels.
// Arguments: t, the thread's index in the group of NT threads solving chain
// ch (ch < 0: none; that group only meets the barriers), its LDS slice `lds`,
// and NL >= N, the row count the reduction levels run for (NL == N for
// closed chains).  Needs N <= the group's threads, open chains or closed
// chains of an even N >= 4, ncr >= 1, and ceil(NL/2) rows within the fixed
// LDS stride SM (SM >= lds_stride(ceil(NL/2))).  With KL > 0 the first KL
// odd rows stay in LDS instead, in `ol` (element-major, stride KS >= KL),
// and only the others go through global scratch.
template <typename T, int FINAL, int SM, int XS, bool JQ, int KL = 0, int KS = 1>
__device__ __forceinline__ void adi_fast_chain_half(ChainTab ct, const JacSrc& js, const int ch, const int NL, T* lds,
                                                    const int t, const int NT, const int tw,
                                                    const T* __restrict__ ff, int nProc, double cfl,
                                                    const double* __restrict__ sigma, double* __restrict__ r,
                                                    double* __restrict__ dq, int dqmode, int ncr,
                                                    T* __restrict__ hs, int hstride, const int* __restrict__ hbase,
                                                    T* ol = nullptr) {
  const int pb = ch >= 0 ? ct.pbase[ch] : 0, nf = ch >= 0 ? ct.nf[ch] : 0, N = ch >= 0 ? ct.ncell[ch] : 0;
  const bool periodic = FINAL != kFinalOpen;
  const int M = (N + 1) / 2;
  T* hrow = hs + (ch >= 0 ? hbase[ch] : 0);
  const bool even = t < M;
  const int row = even ? 2 * t : 2 * (t - M) + 1;
  const bool active = row < N;
  {
    T a[16], c[16], f[4];
    fast_row_xchg<T, JQ>(ct, js, ff, nProc, pb, nf, N, periodic, row, active, cfl, sigma, r, lds, XS, a, c, f);
    if (active && !even) {
      lds_store_row(lds, SM, lpad(t - M), a, c, f);
      if (KL > 0 && t - M < KL)
        lds_store_row(ol, KS, t - M, a, c, f);
      else
        lds_store_row(hrow, hstride, t - M, a, c, f);
    }
    __syncthreads();
    if (active && even) {
      // odd row 2k+1 sits in slot k
      const int iL = periodic ? (row - 1 + N) % N : row - 1;
      const int iR = periodic ? (row + 1) % N : row + 1;
      row_reduce_lds(a, c, f, lds, SM, periodic || iL >= 0, lpad(iL >> 1), periodic || iR < N, lpad(iR >> 1));
    }
    __syncthreads();
    if (active && even) lds_store_row(lds, SM, lpad(t), a, c, f);
  }
  __syncthreads();
  fast_solve_lds<T, FINAL, SM, true>(lds, M, (NL + 1) / 2, cr_levels(ncr) - 1, pcr_cap(ncr), t, NT, tw);
  if (active) {
    T x[4];
    if (even) {
#pragma unroll
      for (int e = 0; e < 4; e++) x[e] = lds[(32 + e) * SM + lpad(t)];
    } else {
      // x = F^ - A^ x_(row-1) - C^ x_(row+1), the even rows' x in the F slots
      const int k = t - M;
      T ra[16], rc[16];
      if (KL > 0 && k < KL)
        lds_load_row(ol, KS, k, ra, rc, x);
      else
        lds_load_row(hrow, hstride, k, ra, rc, x);
      mv4_acc_lds(x, ra, lds, SM, lpad(k));
      if (periodic || row + 1 < N) mv4_acc_lds(x, rc, lds, SM, lpad(((row + 1) % N) >> 1));
    }
    store_solution(r, dq, dqmode, ct.pleft[pb + row], x);
  }
}

#endif  // HAMSTR_LONG_TU

// one chain per workgroup pass; the workgroup walks the group's chains with a
// stride of the grid size (a persistent launch: chains are sorted longest
// first, so every pass of the grid does similar work).  FP32 workgroups of up
// to 512 threads are held to 128 VGPRs (4 waves per SIMD), so that two
// 6-wave workgroups (384-row chains) pack onto a CU: for odd cycles that costs
// 22 spilled VGPRs and still took a 256-case batch's 384-row loops from 1.34
// to 0.95 ms.
// Work for workgroups beyond the chains' in the same launch: the residual
// sums of cells [c0, c1) (k_rhs_cells<true>'s, in blocks of kBlock cells with
// the same block partials), done by workgroups gl.. of a launch whose first gl
// workgroups walk the chains.  c1 <= c0: none.
struct CellWork {
  const int *off, *ent;
  const double *flux, *sr;
  double *r0, *partial;
  int c0, c1, gl;
};

// the blocks of cw's cells, workgroup wi of nw taking blockDim/kBlock
// blocks per pass (blockDim a multiple of kBlock); every thread calls it
__device__ __forceinline__ void rhs_cells_by_blocks(const CellWork& cw, const Consts& c, Live lv, double* __restrict__ r,
                                                 double* __restrict__ sigma, int wi, int nw) {
  __shared__ double wsq[1024 / 64], wmx[1024 / 64];
  const int t = threadIdx.x;
  const int nsub = blockDim.x / kBlock;
  const int sub = t / kBlock;
  const int nb = (cw.c1 - cw.c0 + kBlock - 1) / kBlock;
  for (int b0 = wi * nsub; b0 < nb; b0 += nw * nsub) {
    const int blk = b0 + sub;
    const int i = cw.c0 + blk * kBlock + (t % kBlock);
    double sq = 0.0, mx = 0.0;
    if (blk < nb && i < cw.c1 && (!lv.cell || lv.cell[i])) {
      double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, sg = 0.0;
      cell_flux_sum(cw.ent, cw.off[i], cw.off[i + 1], cw.flux, cw.sr, a0, a1, a2, a3, sg);
      sq = a0 * a0;
      mx = fabs(a0);
      sigma[i] = sg;
      const double dtfac = c.cfl / sg;
      const double4 v = make_double4(a0 * dtfac, a1 * dtfac, a2 * dtfac, a3 * dtfac);
      *reinterpret_cast<double4*>(r + 4 * (size_t)i) = v;
      *reinterpret_cast<double4*>(cw.r0 + 4 * (size_t)i) = v;
    }
    // block_norm_partial's reduction, per block of kBlock threads
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      sq += __shfl_xor(sq, o);
      mx = fmax(mx, __shfl_xor(mx, o));
    }
    if ((t & 63) == 0) {
      wsq[t >> 6] = sq;
      wmx[t >> 6] = mx;
    }
    __syncthreads();
    if (t % kBlock == 0 && blk < nb) {
      const int w0 = t >> 6;
      double a = wsq[w0], b = wmx[w0];
#pragma unroll
      for (int w = 1; w < kBlock / 64; w++) {
        a += wsq[w0 + w];
        b = fmax(b, wmx[w0 + w]);
      }
      const int pb = cw.c0 / kBlock + blk;
      cw.partial[2 * pb] = a;
      cw.partial[2 * pb + 1] = b;
    }
    __syncthreads();
  }
}

template <typename T, int FINAL, int MAXT, bool JQ>
__global__ __launch_bounds__(MAXT)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && MAXT == 512 ? 4 : 1)))
void k_adi_fast(ChainTab ct, Live lv, JacSrc js, const int* __restrict__ chains, int nch,
                                                   const T* __restrict__ ff, int nProc, double cfl,
                                                   const double* __restrict__ sigma, double* __restrict__ r,
                                                   double* __restrict__ dq, int dqmode, int ncr, CellWork cw,
                                                   Consts cc) {
  // the chains' workgroups come first in the launch, so that the hardware
  // deals them to CUs of their own before any cell workgroup
  const int gl = cw.c1 > cw.c0 ? cw.gl : (int)gridDim.x;
  if ((int)blockIdx.x >= gl) {
    rhs_cells_by_blocks(cw, cc, lv, r, const_cast<double*>(sigma), blockIdx.x - gl, gridDim.x - gl);
    return;
  }
  const int vb = xcd_block(blockIdx.x, gl);
  for (int w = vb; w < nch; w += gl) {
    const int ch = chains[w];
    if (lv.chain && !lv.chain[ch]) continue;  // the same for the whole workgroup
    adi_fast_chain<T, FINAL, JQ>(ct, ch, js, ff, nProc, cfl, sigma, r, dq, dqmode, ncr);
    __syncthreads();
  }
}

#ifdef HAMSTR_LONG_TU
// the same walk over chains longer than a workgroup (a kernel of its own: a
// shared one costs the common case registers and time)
template <typename T, int FINAL, int MAXT>
__global__ __launch_bounds__(MAXT) void k_adi_fast_long(ChainTab ct, Live lv, const int* __restrict__ chains,
                                                        int nch, const T* __restrict__ ff, int nProc, double cfl,
                                                        const double* __restrict__ sigma, double* __restrict__ r,
                                                        double* __restrict__ dq, int dqmode, int ncr,
                                                        T* __restrict__ hs, int hstride,
                                                        const int* __restrict__ hbase) {
  const int vb = xcd_block(blockIdx.x, gridDim.x);
  for (int w = vb; w < nch; w += gridDim.x) {
    const int ch = chains[w];
    if (lv.chain && !lv.chain[ch]) continue;  // the same for the whole workgroup
    adi_fast_chain_long<T, FINAL>(ct, ch, ff, nProc, cfl, sigma, r, dq, dqmode, ncr, hs, hstride, hbase);
    __syncthreads();
  }
}

// the same walk with adi_fast_chain_half, one chain per workgroup; WPE:
// minimum waves per SIMD the register allocation must leave room for (3: at
// most 168 VGPRs).  The LDS stride is fixed at lds_stride(ceil(MAXT/2)); the
// launch provides 36 such element rows.
template <typename T, int FINAL, int MAXT, int WPE, bool JQ>
__global__ __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(WPE))) void k_adi_fast_half(
    ChainTab ct, Live lv, JacSrc js, const int* __restrict__ chains, int nch, const T* __restrict__ ff, int nProc,
    double cfl, const double* __restrict__ sigma, double* __restrict__ r, double* __restrict__ dq, int dqmode,
    int ncr, T* __restrict__ hs, int hstride, const int* __restrict__ hbase) {
  static_assert(16 * (MAXT + 1) <= 36 * lds_stride((MAXT + 1) / 2), "face exchange fits the slice");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  const int vb = xcd_block(blockIdx.x, gridDim.x);
  for (int w = vb; w < nch; w += gridDim.x) {
    const int ch = chains[w];
    if (lv.chain && !lv.chain[ch]) continue;  // the same for the whole workgroup
    adi_fast_chain_half<T, FINAL, lds_stride((MAXT + 1) / 2), MAXT + 1, JQ>(ct, js, ch, ct.ncell[ch], lds, threadIdx.x,
                                                                        (int)blockDim.x, min(4, (int)blockDim.x >> 6),
                                                                        ff, nProc, cfl, sigma, r,
                                                                        dq, dqmode, ncr, hs, hstride, hbase);
    __syncthreads();
  }
}

// Two open chains per workgroup, each on its own MAXH threads and its own
// LDS slice (2 x 36 element rows of stride lds_stride(ceil(MAXH/2))): chains
// 2w and 2w+1 of the group.  With MAXH = 384 a workgroup is 12 waves, 3 per
// SIMD, which 168 VGPRs allow; two workgroups of 6 waves each do not share a
// CU (their waves do not pack evenly onto the 4 SIMDs), so this is how two
// 384-row chains run on one CU.  Both halves run the reduction levels of the
// longer chain.
template <typename T, int MAXH, bool JQ>
__global__ __launch_bounds__(2 * MAXH) __attribute__((amdgpu_waves_per_eu(3))) void k_adi_fast_half2(
    ChainTab ct, Live lv, JacSrc js, const int* __restrict__ chains, int nch, const T* __restrict__ ff, int nProc,
    double cfl, const double* __restrict__ sigma, double* __restrict__ r, double* __restrict__ dq, int dqmode,
    int ncr, T* __restrict__ hs, int hstride, const int* __restrict__ hbase) {
  constexpr int SM = lds_stride((MAXH + 1) / 2);
  constexpr int KL = half2_odd_lds<T, MAXH>(), KS = KL + 1;
  static_assert(16 * (MAXH + 1) <= 36 * SM, "face exchange fits the slice");
  static_assert(36 * (2 * SM + 2 * KS) * (int)sizeof(T) <= kLdsBytes, "slices and odd rows fit the LDS");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // which chain of the pair: the same for a whole wavefront (MAXH is a multiple
  // of 64), so kept in scalar registers, as is everything derived from it
  const int sub = __builtin_amdgcn_readfirstlane(threadIdx.x >= MAXH ? 1 : 0);
  T* lds = reinterpret_cast<T*>(smem_raw) + sub * 36 * SM;
  // the pair's odd rows kept in LDS, after both slices
  T* ol = reinterpret_cast<T*>(smem_raw) + 2 * 36 * SM + sub * 36 * KS;
  const int t = threadIdx.x - sub * MAXH;
  const int npair = (nch + 1) / 2;
  const int vb = xcd_block(blockIdx.x, gridDim.x);
  for (int w = vb; w < npair; w += gridDim.x) {
    int c0 = chains[2 * w];
    int c1 = 2 * w + 1 < nch ? chains[2 * w + 1] : -1;
    if (lv.chain && !lv.chain[c0]) c0 = -1;
    if (c1 >= 0 && lv.chain && !lv.chain[c1]) c1 = -1;
    const int NL = max(c0 >= 0 ? ct.ncell[c0] : 0, c1 >= 0 ? ct.ncell[c1] : 0);
    if (NL == 0) continue;  // the same for the whole workgroup
    // the Thomas finish on two waves per chain: waves 0-1 and 6-7, four SIMDs
    adi_fast_chain_half<T, kFinalOpen, SM, MAXH + 1, JQ, KL, KS>(ct, js, __builtin_amdgcn_readfirstlane(sub ? c1 : c0),
                                                                 NL, lds, t, MAXH, 2, ff, nProc, cfl, sigma, r, dq,
                                                                 dqmode, ncr, hs, hstride, hbase, ol);
    __syncthreads();
  }
}
#endif  // HAMSTR_LONG_TU

// --------------------------------------------------------------------------
// lin_faces: reconstruction derivatives + lmat*dqL + rmat*dqR per face
// (computeLinearRHS.cpp:52-325)
// --------------------------------------------------------------------------
// RECOMP: the face's Jacobians are computed again here, by the code and from
// the states k_rhs_faces used (the same q), instead of read back from ff: 1.1
// kflop per face in place of 32 loads.  Only for the fused iteration, where ff
// is k_rhs_faces' output for the current q.
// STAGED: the workgroup's load_state (where read) and load_dstate rows staged
// in LDS (stage_range)
template <int ORDER, typename TL, bool FASTJ, bool RECOMP, bool STAGED>
__global__ __launch_bounds__(kBlock) void k_lin_faces(FaceTab ft, Consts c, Live lv, const double* __restrict__ q,
                                                      const double* __restrict__ dq, const TL* __restrict__ ff,
                                                      double* __restrict__ lflux) {
  constexpr bool QROWS = ORDER != 1 || RECOMP;  // the stencil's states are read
  const int pf = xcd_block(blockIdx.x, gridDim.x) * blockDim.x;
  const int p = pf + threadIdx.x;
  // rows 0-3: load_dstate; 4-7: load_state
  __shared__ double stg[STAGED ? (QROWS ? 8 : 4) : 1][STAGED ? kStageSlots : 1];
  int S0 = 0;
  if constexpr (STAGED) {
    // a workgroup whose faces are all retired (converged cases of a batch)
    // stages nothing
    if (!__syncthreads_or(p < c.nProc && (!lv.face || lv.face[p]))) return;
    const int2 R = stage_range<ORDER>(ft, c.nProc, pf);
    S0 = R.x;
    for (int i = threadIdx.x; i <= R.y - R.x; i += blockDim.x) {
      const int2 e = ft.src[S0 + i];
      const double4 b = load_dstate(e, dq, ft.refmtx);
      stg[0][i] = b.x; stg[1][i] = b.y; stg[2][i] = b.z; stg[3][i] = b.w;
      if constexpr (QROWS) {
        const double4 a = load_state(e, q, ft.refmtx, ft.qinf);
        stg[QROWS ? 4 : 0][i] = a.x; stg[QROWS ? 5 : 1][i] = a.y;
        stg[QROWS ? 6 : 2][i] = a.z; stg[QROWS ? 7 : 3][i] = a.w;
      }
    }
    __syncthreads();
  }
  if (p >= c.nProc || (lv.face && !lv.face[p])) return;
  const int code = ft.pslot[p];
  const int s = code & kSlotMask;
  // a slot's load_dstate (w 0) or load_state (w 1), staged or gathered
  auto row = [&](int slot, int w, double* o) __attribute__((always_inline)) {
    if constexpr (STAGED) {
#pragma unroll
      for (int k = 0; k < 4; k++) o[k] = stg[(QROWS && w) ? 4 + k : k][slot - S0];
    } else {
      const int2 e = ft.src[slot];
      put4(o, w ? load_state(e, q, ft.refmtx, ft.qinf) : load_dstate(e, dq, ft.refmtx));
    }
  };
  double dl[4], dr[4];
  double left[4], right[4];  // the face states (RECOMP)
  if (code & kOrphanSlot) {
    // orphan face: the full product of the two cells' dq
    row(s, 0, dl);
    row(s + 1, 0, dr);
    if (RECOMP) {
      row(s, 1, left);
      row(s + 1, 1, right);
    }
  } else {
    const bool flip = (code >> 30) & 1;
    constexpr int lo = (ORDER == 5) ? 0 : (ORDER == 3 ? 1 : 2);
    constexpr int hi = (ORDER == 5) ? 5 : (ORDER == 3 ? 4 : 3);
    double f[6][4], d[6][4];
#pragma unroll
    for (int j = lo; j <= hi; j++) {
      if (QROWS) row(s - 2 + j, 1, f[j]);
      row(s - 2 + j, 0, d[j]);
    }
    double dln[4], drn1[4];
    reconstruct_dface<ORDER, FASTJ>(f, d, c.eps, c.qt, dln, drn1);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      dl[k] = flip ? dln[k] : drn1[k];
      dr[k] = flip ? drn1[k] : dln[k];
    }
    if (RECOMP) {
      // the states from opaque copies of the stencil: shared with the
      // derivative above, common subexpressions would change how the
      // derivative's multiply-adds fuse (FASTJ) from k_lin_faces without RECOMP
      double fb[6][4], qln[4], qrn1[4];
#pragma unroll
      for (int j = lo; j <= hi; j++)
#pragma unroll
        for (int k = 0; k < 4; k++) {
          fb[j][k] = f[j][k];
          asm volatile("" : "+v"(fb[j][k]));
        }
      reconstruct_face<ORDER>(fb, c.eps, c.qt, qln, qrn1);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        left[k] = flip ? qln[k] : qrn1[k];
        right[k] = flip ? qrn1[k] : qln[k];
      }
    }
  }
  TL lmat[16], rmat[16];
  if (RECOMP) {
    const double2 n = *reinterpret_cast<const double2*>(ft.nvec + 2 * (size_t)p);
    face_jacobians<TL, FASTJ>(n.x, n.y, left, right, c.gamma, lmat, rmat);
  } else {
#pragma unroll
    for (int e = 0; e < 16; e++) {
      lmat[e] = ff[(size_t)e * c.nProc + p];
      rmat[e] = ff[(size_t)(16 + e) * c.nProc + p];
    }
  }
  double out[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    double acc = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const double lm = double(lmat[4 * j + k]);
      const double rm = double(rmat[4 * j + k]);
      acc += ((lm * dl[k]) + (rm * dr[k]));
    }
    out[j] = acc;
  }
  *reinterpret_cast<double4*>(lflux + 4 * (size_t)p) = make_double4(out[0], out[1], out[2], out[3]);
}

// Orphan faces (interior faces in no chain; meshes with triangles) run in
// k_rhs_faces and k_lin_faces with the chain faces: an extension with no
// counterpart in the reference, the first-order Roe flux of the two cell
// states (and its frozen-Roe Jacobians), and in the linear residual the full
// product lmat dq_L + rmat dq_R.  Their entries come after the chain entries
// in each cell's list, as the oracle sums them.

// lin_cells: r = r0 - dq - CFL/sigma * sum(+-lflux), norms, optional q += dq
// (computeLinearRHS.cpp:45, 326-368; ADI.cpp:258-264)
template <bool UPDATE>
__global__ __launch_bounds__(kBlock) void k_lin_cells(Consts c, Live lv, const int* __restrict__ off,
                                                      const int* __restrict__ ent,
                                                      const double* __restrict__ lflux,
                                                      const double* __restrict__ r0, const double* __restrict__ dq,
                                                      const double* __restrict__ sigma, double* __restrict__ r,
                                                      double* __restrict__ q, double* __restrict__ partial,
                                                      double* __restrict__ rrho) {
  const int vb = xcd_block(blockIdx.x, gridDim.x);
  const int i = vb * blockDim.x + threadIdx.x;
  double sq = 0.0, mx = 0.0;
  if (i < c.nCell && (!lv.cell || lv.cell[i])) {
    const double4 a = *reinterpret_cast<const double4*>(r0 + 4 * (size_t)i);
    const double4 b = *reinterpret_cast<const double4*>(dq + 4 * (size_t)i);
    double a0 = (a.x - b.x), a1 = (a.y - b.y), a2 = (a.z - b.z), a3 = (a.w - b.w);
    const double dtfac = c.cfl / sigma[i];
    cell_lin_sum(ent, off[i], off[i + 1], lflux, dtfac, a0, a1, a2, a3);
    *reinterpret_cast<double4*>(r + 4 * (size_t)i) = make_double4(a0, a1, a2, a3);
    sq = a0 * a0;
    mx = fabs(a0);
    if (rrho) rrho[i] = a0;
    if (UPDATE) {
      double4 qv = *reinterpret_cast<double4*>(q + 4 * (size_t)i);
      qv.x += b.x; qv.y += b.y; qv.z += b.z; qv.w += b.w;
      *reinterpret_cast<double4*>(q + 4 * (size_t)i) = qv;
    }
  }
  block_norm_partial(sq, mx, partial, vb);
}

// --------------------------------------------------------------------------
// The end of one ADI iteration fused with the start of the next (a single
// case, msweep = 1): computeLinearRHS of iteration n shares its face gather
// with computeRHS("implicit") of iteration n + 1, which reads the state
// q + dq that iteration n's update produces.  Each value is computed by the
// operations of the kernels it replaces (k_lin_faces with the Jacobians
// computed again, k_rhs_faces; k_lin_cells with the update, k_rhs_cells), so
// the iterations are bit-identical to the unfused ones.
// --------------------------------------------------------------------------

// load_state of the updated state q + dq (the sum formed per raw cell, as
// k_lin_cells<true> forms it, before any wall reflection)
__device__ __forceinline__ double4 load_state_sum(const int2 e, const double* __restrict__ q,
                                                  const double* __restrict__ dq,
                                                  const double* __restrict__ refmtx,
                                                  const double4* __restrict__ qinf) {
  if (e.x < 0) return qinf[e.y];
  const double4 a = *reinterpret_cast<const double4*>(q + 4 * (size_t)e.x);
  const double4 b = *reinterpret_cast<const double4*>(dq + 4 * (size_t)e.x);
  double4 v = make_double4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
  if (e.y >= 0) {
    const double4 R = *reinterpret_cast<const double4*>(refmtx + 4 * (size_t)e.y);
    const double u = v.y, w = v.z;
    v.y = (u * R.x + w * R.y);
    v.z = (u * R.z + w * R.w);
  }
  return v;
}

// face_states of the updated state q + dq
template <int ORDER>
__device__ __forceinline__ void face_states_sum(const FaceTab& ft, const double* __restrict__ q,
                                                const double* __restrict__ dq, double eps, double qt, int code,
                                                double* left, double* right) {
  const int s = code & kSlotMask;
  if (code & kOrphanSlot) {
    put4(left, load_state_sum(ft.src[s], q, dq, ft.refmtx, ft.qinf));
    put4(right, load_state_sum(ft.src[s + 1], q, dq, ft.refmtx, ft.qinf));
  } else {
    const bool flip = (code >> 30) & 1;
    constexpr int lo = (ORDER == 5) ? 0 : (ORDER == 3 ? 1 : 2);
    constexpr int hi = (ORDER == 5) ? 5 : (ORDER == 3 ? 4 : 3);
    double f[6][4];
#pragma unroll
    for (int j = lo; j <= hi; j++) put4(f[j], load_state_sum(ft.src[s - 2 + j], q, dq, ft.refmtx, ft.qinf));
    double qln[4], qrn1[4];
    reconstruct_face<ORDER>(f, eps, qt, qln, qrn1);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      left[k] = flip ? qln[k] : qrn1[k];
      right[k] = flip ? qrn1[k] : qln[k];
    }
  }
}

// k_lin_faces<ORDER, TL, FASTJ, true> on (q, dq), then k_rhs_faces<ORDER, TL,
// true, FASTJ> on q + dq (ff_from, fst as there).
// STAGED: the workgroup's load_state, load_dstate and load_state_sum rows
// staged in LDS (stage_range)

template <int ORDER, typename TL, bool FASTJ, bool STAGED>
__global__ __launch_bounds__(kBlock) void k_lin_rhs_faces(FaceTab ft, Consts c, Live lv, const double* __restrict__ q,
                                                          const double* __restrict__ dq,
                                                          double* __restrict__ lflux, double* __restrict__ flux,
                                                          double* __restrict__ sr, TL* __restrict__ ff, int ff_from,
                                                          double* __restrict__ fst, int* __restrict__ nanflag, int jac_to) {
  const int pf = xcd_block(blockIdx.x, gridDim.x) * blockDim.x;
  const int p = pf + threadIdx.x;
  // per staged slot: load_state (0-3), load_dstate (4-7), load_state_sum (8-11)
  __shared__ double stg[STAGED ? 12 : 1][STAGED ? kStageSlots : 1];
  int S0 = 0;
  if constexpr (STAGED) {
    // a workgroup whose faces are all retired (converged cases of a batch)
    // stages nothing
    if (!__syncthreads_or(p < c.nProc && (!lv.face || lv.face[p]))) return;
    const int2 R = stage_range<ORDER>(ft, c.nProc, pf);
    S0 = R.x;
    for (int i = threadIdx.x; i <= R.y - R.x; i += blockDim.x) {
      const int2 e = ft.src[S0 + i];
      const double4 a = load_state(e, q, ft.refmtx, ft.qinf);
      const double4 b = load_dstate(e, dq, ft.refmtx);
      const double4 u = load_state_sum(e, q, dq, ft.refmtx, ft.qinf);
      stg[0][i] = a.x; stg[1][i] = a.y; stg[2][i] = a.z; stg[3][i] = a.w;
      stg[4][i] = b.x; stg[5][i] = b.y; stg[6][i] = b.z; stg[7][i] = b.w;
      stg[8][i] = u.x; stg[9][i] = u.y; stg[10][i] = u.z; stg[11][i] = u.w;
    }
    __syncthreads();
  }
  if (p >= c.nProc || (lv.face && !lv.face[p])) return;
  const int code = ft.pslot[p];
  const int s = code & kSlotMask;
  // slot row k of quantity w (0 state, 1 dstate, 2 state sum), staged or gathered
  auto row = [&](int slot, int w, double* o) __attribute__((always_inline)) {
    if constexpr (STAGED) {
#pragma unroll
      for (int k = 0; k < 4; k++) o[k] = stg[4 * w + k][slot - S0];
    } else {
      const int2 e = ft.src[slot];
      put4(o, w == 0 ? load_state(e, q, ft.refmtx, ft.qinf)
                     : (w == 1 ? load_dstate(e, dq, ft.refmtx) : load_state_sum(e, q, dq, ft.refmtx, ft.qinf)));
    }
  };
  const double2 n = *reinterpret_cast<const double2*>(ft.nvec + 2 * (size_t)p);
  {
    double dl[4], dr[4], left[4], right[4];
    if (code & kOrphanSlot) {
      row(s, 1, dl);
      row(s + 1, 1, dr);
      row(s, 0, left);
      row(s + 1, 0, right);
    } else {
      const bool flip = (code >> 30) & 1;
      constexpr int lo = (ORDER == 5) ? 0 : (ORDER == 3 ? 1 : 2);
      constexpr int hi = (ORDER == 5) ? 5 : (ORDER == 3 ? 4 : 3);
      double f[6][4], d[6][4];
#pragma unroll
      for (int j = lo; j <= hi; j++) {
        row(s - 2 + j, 0, f[j]);
        row(s - 2 + j, 1, d[j]);
      }
      double dln[4], drn1[4];
      reconstruct_dface<ORDER, FASTJ>(f, d, c.eps, c.qt, dln, drn1);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        dl[k] = flip ? dln[k] : drn1[k];
        dr[k] = flip ? drn1[k] : dln[k];
      }
      // opaque copies, as in k_lin_faces
      double fb[6][4], qln[4], qrn1[4];
#pragma unroll
      for (int j = lo; j <= hi; j++)
#pragma unroll
        for (int k = 0; k < 4; k++) {
          fb[j][k] = f[j][k];
          asm volatile("" : "+v"(fb[j][k]));
        }
      reconstruct_face<ORDER>(fb, c.eps, c.qt, qln, qrn1);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        left[k] = flip ? qln[k] : qrn1[k];
        right[k] = flip ? qrn1[k] : qln[k];
      }
    }
    TL lmat[16], rmat[16];
    face_jacobians<TL, FASTJ>(n.x, n.y, left, right, c.gamma, lmat, rmat);
    double out[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      double acc = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const double lm = double(lmat[4 * j + k]);
        const double rm = double(rmat[4 * j + k]);
        acc += ((lm * dl[k]) + (rm * dr[k]));
      }
      out[j] = acc;
    }
    *reinterpret_cast<double4*>(lflux + 4 * (size_t)p) = make_double4(out[0], out[1], out[2], out[3]);
  }
  double left[4], right[4];
  if constexpr (STAGED) {
    // face_states_sum on the staged rows
    face_states_from<ORDER>([&](int slot, double* o) __attribute__((always_inline)) { row(slot, 2, o); }, c.eps,
                            c.qt, code, left, right);
  } else {
    face_states_sum<ORDER>(ft, q, dq, c.eps, c.qt, code, left, right);
  }
  double fl[4], s_r;
  flux_roe2d(n.x, n.y, left, right, c.gamma, fl, &s_r);
  *reinterpret_cast<double4*>(flux + 4 * (size_t)p) = make_double4(fl[0], fl[1], fl[2], fl[3]);
  sr[p] = s_r;
  if (s_r != s_r) atomicOr(nanflag, 1);
  const bool store_jac = p >= ff_from || p < jac_to || ((code >> 30) & 1);
  if (fst && !store_jac) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      fst[(size_t)k * c.nProc + p] = left[k];
      fst[(size_t)(4 + k) * c.nProc + p] = right[k];
    }
  }
  if (store_jac) {
    TL lm[16], rm[16];
    face_jacobians<TL, FASTJ>(n.x, n.y, left, right, c.gamma, lm, rm);
#pragma unroll
    for (int e = 0; e < 16; e++) {
      ff[(size_t)e * c.nProc + p] = lm[e];
      ff[(size_t)(16 + e) * c.nProc + p] = rm[e];
    }
  }
}

// k_lin_cells<true> over every live cell (its block partials to p_lin, and in
// a batch its density residual to rrho[nCell + i]; the linear residual itself
// is not kept: the next computeRHS overwrites r), then k_rhs_cells<true> over
// the cells of the first nb_rhs blocks (block partials to p_nl, rrho[i]).  The launch's first nb_rhs workgroups take those blocks, spread
// over the XCDs like a launch of their own (nb_rhs a multiple of 8), so that
// every XCD gets its share of the blocks with both sums (all of them on three
// XCDs took the kernel from 12.9 to 16.4 us on the bench mesh).
__global__ __launch_bounds__(kBlock) void k_lin_rhs_cells(Consts c, Live lv, const int* __restrict__ off,
                                                          const int* __restrict__ ent,
                                                          const double* __restrict__ lflux,
                                                          const double* __restrict__ flux,
                                                          const double* __restrict__ sr, double* __restrict__ r0,
                                                          const double* __restrict__ dq, double* __restrict__ sigma,
                                                          double* __restrict__ r, double* __restrict__ q,
                                                          double* __restrict__ p_lin, double* __restrict__ p_nl,
                                                          double* __restrict__ rrho, int nb_rhs) {
  const int bx = blockIdx.x;
  const int vb = bx < nb_rhs ? xcd_block(bx, nb_rhs) : nb_rhs + xcd_block(bx - nb_rhs, gridDim.x - nb_rhs);
  const int i = vb * blockDim.x + threadIdx.x;
  const bool in = i < c.nCell && (!lv.cell || lv.cell[i]);
  double sq = 0.0, mx = 0.0;
  if (in) {
    const double4 a = *reinterpret_cast<const double4*>(r0 + 4 * (size_t)i);
    const double4 b = *reinterpret_cast<const double4*>(dq + 4 * (size_t)i);
    double a0 = (a.x - b.x), a1 = (a.y - b.y), a2 = (a.z - b.z), a3 = (a.w - b.w);
    const double dtfac = c.cfl / sigma[i];
    cell_lin_sum(ent, off[i], off[i + 1], lflux, dtfac, a0, a1, a2, a3);
    sq = a0 * a0;
    mx = fabs(a0);
    if (rrho) rrho[c.nCell + i] = a0;
    double4 qv = *reinterpret_cast<double4*>(q + 4 * (size_t)i);
    qv.x += b.x; qv.y += b.y; qv.z += b.z; qv.w += b.w;
    *reinterpret_cast<double4*>(q + 4 * (size_t)i) = qv;
  }
  block_norm_partial(sq, mx, p_lin, vb);
  if (vb >= nb_rhs) return;  // the same for the whole block
  sq = 0.0;
  mx = 0.0;
  if (in) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, sg = 0.0;
    cell_flux_sum(ent, off[i], off[i + 1], flux, sr, a0, a1, a2, a3, sg);
    sq = a0 * a0;
    mx = fabs(a0);
    if (rrho) rrho[i] = a0;
    sigma[i] = sg;
    const double dtfac = c.cfl / sg;
    const double4 v = make_double4(a0 * dtfac, a1 * dtfac, a2 * dtfac, a3 * dtfac);
    *reinterpret_cast<double4*>(r + 4 * (size_t)i) = v;
    *reinterpret_cast<double4*>(r0 + 4 * (size_t)i) = v;
  }
  __syncthreads();  // block_norm_partial's LDS is reused
  block_norm_partial(sq, mx, p_nl, vb);
}

__global__ void k_update(int n4, Live lv, double* __restrict__ q, const double* __restrict__ dq) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n4 && (!lv.cell || lv.cell[i >> 2])) q[i] += dq[i];
}

__global__ void k_fill_u8(unsigned char* __restrict__ p, long n, unsigned char v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// initFlow (initFlow.cpp:152-176): q := freestream of the cell's case, and
// qt := q (initFlow.cpp:71-88)
__global__ void k_init_flow(int nCell, const int* __restrict__ cell_case, const double4* __restrict__ qinf,
                            double* __restrict__ q, double* __restrict__ qt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nCell) {
    const double4 v = qinf[cell_case[i]];
    *reinterpret_cast<double4*>(q + 4 * (size_t)i) = v;
    *reinterpret_cast<double4*>(qt + 4 * (size_t)i) = v;
  }
}

// SOLVER::updateSolution (march.cpp:142-164): dst = src + (coef*CFL/sigma) r,
// per element in the reference's operation order.  src and dst may be the
// same field (each element is read before it is written).
__global__ void k_update_solution(int nCell, Live lv, const double* src, double* dst, const double* __restrict__ r,
                                  const double* __restrict__ sigma, double coef, double cfl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nCell || (lv.cell && !lv.cell[i])) return;
  const double dtfac = coef * cfl / sigma[i];
  const double4 a = *reinterpret_cast<const double4*>(src + 4 * (size_t)i);
  const double4 rv = *reinterpret_cast<const double4*>(r + 4 * (size_t)i);
  *reinterpret_cast<double4*>(dst + 4 * (size_t)i) =
      make_double4(a.x + dtfac * rv.x, a.y + dtfac * rv.y, a.z + dtfac * rv.z, a.w + dtfac * rv.w);
}

// per-case density-residual norms, one block per case and stage, cells in
// order: stage s = slot/2 + blockIdx.y reads rrho + s*nCell and writes
// out[4*k + 2*s] = L2, out[4*k + 2*s + 1] = Linf.  keep[4*k + ...] holds each
// case's last norms; a retired case repeats them instead of reducing again.
__global__ __launch_bounds__(kRedBlock) void k_case_norms(const double* __restrict__ rrho_all, int nCell,
                                                          const int* __restrict__ case_off, int slot,
                                                          double* __restrict__ out, Live lv,
                                                          double* __restrict__ keep) {
  __shared__ double ssq[kRedBlock];
  __shared__ double smx[kRedBlock];
  const int k = blockIdx.x, t = threadIdx.x;
  const int stage = slot / 2 + blockIdx.y;
  const double* rrho = rrho_all + (size_t)stage * nCell;
  const int c0 = case_off[k], c1 = case_off[k + 1];
  if (lv.cell && !lv.cell[c0]) {  // the same for the whole block
    if (t < 2) out[4 * k + 2 * stage + t] = keep[4 * k + 2 * stage + t];
    return;
  }
  double a = 0.0, m = 0.0;
  for (int i = c0 + t; i < c1; i += kRedBlock) {
    const double v = rrho[i];
    a += v * v;
    m = fmax(m, fabs(v));
  }
  ssq[t] = a;
  smx[t] = m;
  __syncthreads();
  for (int w = kRedBlock / 2; w > 0; w >>= 1) {
    if (t < w) {
      ssq[t] += ssq[t + w];
      smx[t] = fmax(smx[t], smx[t + w]);
    }
    __syncthreads();
  }
  if (t == 0) {
    const double l2 = sqrt(ssq[0]) / (c1 - c0);
    out[4 * k + 2 * stage] = l2;
    out[4 * k + 2 * stage + 1] = smx[0];
    keep[4 * k + 2 * stage] = l2;
    keep[4 * k + 2 * stage + 1] = smx[0];
  }
}

// --------------------------------------------------------------------------
// Unit kernels (per-function parity)
// --------------------------------------------------------------------------
__global__ void k_unit_flux(int n, const double* nv, const double* ql, const double* qr, double gamma,
                            double* flux, double* sr, int* nanflag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double f[4], s;
  flux_roe2d(nv[2 * i], nv[2 * i + 1], ql + 4 * i, qr + 4 * i, gamma, f, &s);
  for (int k = 0; k < 4; k++) flux[4 * i + k] = f[k];
  sr[i] = s;
  if (s != s) atomicOr(nanflag, 1);
}

template <typename T>
__global__ void k_unit_jac(int n, const double* nv, const double* ql, const double* qr, double gamma,
                           double* lm, double* rm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T L[16], R[16];
  jacobian_roe2d<T>(nv[2 * i], nv[2 * i + 1], ql + 4 * i, qr + 4 * i, gamma, L, R);
  for (int k = 0; k < 16; k++) {
    lm[16 * i + k] = double(L[k]);
    rm[16 * i + k] = double(R[k]);
  }
}

// kind: 0 MUSCL, 1 WENO, 2 MUSCL_DERIV, 3 WENO_DERIV.  One thread per (chain, row).
__global__ void k_unit_recon(int kind, int nchain, int imax, int is, int iend, double qt, double eps,
                             const double* f, const double* dq, double* ql, double* qr) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  const int nrow = iend - is;
  if (id >= nchain * nrow * 4) return;
  const int k = id % 4;
  const int i = is + (id / 4) % nrow;
  const int ch = id / (4 * nrow);
  const double* F = f + (size_t)ch * imax * 4;
  auto at = [&](const double* a, int row) { return a[(size_t)row * 4 + k]; };
  double a = 0, b = 0;
  if (kind == 0) {
    if (qt == 0.0) {
      a = b = at(F, i);
    } else {
      muscl_pt(at(F, i - 1), at(F, i), at(F, i + 1), qt, eps, &a, &b);
    }
  } else if (kind == 1) {
    a = weno5(at(F, i - 2), at(F, i - 1), at(F, i), at(F, i + 1), at(F, i + 2), 1e-15);
    b = weno5(at(F, i + 2), at(F, i + 1), at(F, i), at(F, i - 1), at(F, i - 2), 1e-15);
  } else {
    const double* D = dq + (size_t)ch * imax * 4;
    if (kind == 2) {
      if (qt == 0.0) {
        a = b = at(D, i);
      } else {
        muscl_deriv_pt(at(F, i - 1), at(F, i), at(F, i + 1), at(D, i - 1), at(D, i), at(D, i + 1), qt, eps, &a,
                       &b);
      }
    } else {
      a = weno5_deriv(at(F, i - 2), at(F, i - 1), at(F, i), at(F, i + 1), at(F, i + 2), at(D, i - 2),
                      at(D, i - 1), at(D, i), at(D, i + 1), at(D, i + 2), eps);
      b = weno5_deriv(at(F, i + 2), at(F, i + 1), at(F, i), at(F, i - 1), at(F, i - 2), at(D, i + 2),
                      at(D, i + 1), at(D, i), at(D, i - 1), at(D, i - 2), eps);
    }
  }
  ql[((size_t)ch * imax + i) * 4 + k] = a;
  qr[((size_t)ch * imax + i) * 4 + k] = b;
}

// the 3D / strand building blocks, thread per face or per system
__global__ void k_unit_flux3d(int n, const double* nv, const double* ql, const double* qr, double gamma,
                              double* flux, double* sr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double f[5], s;
  hamstr::flux_roe3d(nv + 3 * (size_t)i, ql + 5 * (size_t)i, qr + 5 * (size_t)i, gamma, f, &s);
  for (int k = 0; k < 5; k++) flux[5 * (size_t)i + k] = f[k];
  sr[i] = s;
}

__global__ void k_unit_jac3d(int n, const double* nv, const double* ql, const double* qr, double gamma, double* lm,
                             double* rm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double L[25], R[25];
  hamstr::jacobian_roe3d(nv + 3 * (size_t)i, ql + 5 * (size_t)i, qr + 5 * (size_t)i, gamma, L, R);
  for (int k = 0; k < 25; k++) {
    lm[25 * (size_t)i + k] = L[k];
    rm[25 * (size_t)i + k] = R[k];
  }
}

// in place on the device copies of a, b, c, f
__global__ void k_unit_tridag5(int nsys, int N, int periodic, double* a, double* b, double* c, double* f) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsys) return;
  const size_t m = (size_t)t * N;
  hamstr::block_tridag5(a + 25 * m, b + 25 * m, c + 25 * m, f + 5 * m, N, periodic != 0);
}

// blockTridag4 / blockTridagPeriodic4 on stored systems, thread per system
__global__ void k_unit_tridag(int nsys, int N, int periodic, const double* a_, const double* b_,
                              const double* c_, double* f_, double* scratch) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsys) return;
  const double* a0 = a_ + (size_t)t * N * 16;
  const double* b0 = b_ + (size_t)t * N * 16;
  const double* c0 = c_ + (size_t)t * N * 16;
  double* f0 = f_ + (size_t)t * N * 4;
  double* S = scratch + (size_t)t * N * 36;
  double Cp[16], Fp[4], ap[16];
  for (int i = 0; i < N; i++) {
    double A[16], B[16], C[16], F[4], a[16];
    for (int e = 0; e < 16; e++) {
      A[e] = a0[16 * i + e];
      B[e] = b0[16 * i + e];
      C[e] = c0[16 * i + e];
    }
    for (int e = 0; e < 4; e++) F[e] = f0[4 * i + e];
    if (i != 0) {
      mat_sub_mul(B, A, Cp);
      vec_sub_mul(F, A, Fp);
      if (periodic) {
        for (int k = 0; k < 4; k++)
          for (int j = 0; j < 4; j++) {
            double x = 0.0;
            for (int l = 0; l < 4; l++) x -= A[4 * k + l] * ap[4 * l + j];
            a[4 * k + j] = x;
          }
      }
    } else if (periodic) {
      for (int e = 0; e < 16; e++) a[e] = A[e];
    }
    const LU4<double> L = lu4(B);
    lu4_solve(L, F, 1);
    if (!periodic) {
      if (i != N - 1)
        for (int j = 0; j < 4; j++) lu4_solve(L, C + j, 4);
    } else if (i != N - 2) {
      for (int j = 0; j < 4; j++) {
        lu4_solve(L, C + j, 4);
        lu4_solve(L, a + j, 4);
      }
    } else {
      for (int e = 0; e < 16; e++) C[e] += a[e];
      for (int j = 0; j < 4; j++) lu4_solve(L, C + j, 4);
      for (int e = 0; e < 16; e++) a[e] = C[e];
    }
    for (int e = 0; e < 16; e++) S[36 * i + e] = C[e];
    for (int e = 0; e < 4; e++) S[36 * i + 16 + e] = F[e];
    for (int e = 0; e < 16; e++) S[36 * i + 20 + e] = a[e];
    for (int e = 0; e < 16; e++) {
      Cp[e] = C[e];
      ap[e] = a[e];
    }
    for (int e = 0; e < 4; e++) Fp[e] = F[e];
  }
  if (!periodic) {
    double x[4] = {Fp[0], Fp[1], Fp[2], Fp[3]};
    for (int e = 0; e < 4; e++) f0[4 * (N - 1) + e] = x[e];
    for (int i = N - 2; i >= 0; i--) {
      double fv[4];
      for (int e = 0; e < 4; e++) fv[e] = S[36 * i + 16 + e];
      vec_sub_mul(fv, &S[36 * i], x);
      for (int e = 0; e < 4; e++) x[e] = f0[4 * i + e] = fv[e];
    }
    return;
  }
  double Bl[16], Cl[16], Fl[4];
  for (int e = 0; e < 16; e++) {
    Bl[e] = (e % 5 == 0) ? 1.0 : 0.0;
    Cl[e] = Cp[e];
  }
  for (int e = 0; e < 4; e++) Fl[e] = Fp[e];
  for (int i = 0; i < N - 1; i++) {
    const double* c = &S[36 * i];
    const double* a = &S[36 * i + 20];
    double tt[16];
    for (int k = 0; k < 4; k++)
      for (int j = 0; j < 4; j++) {
        double x = 0.0, b = Bl[4 * k + j];
        for (int l = 0; l < 4; l++) {
          x -= Cl[4 * k + l] * c[4 * l + j];
          b -= Cl[4 * k + l] * a[4 * l + j];
        }
        tt[4 * k + j] = x;
        Bl[4 * k + j] = b;
      }
    vec_sub_mul(Fl, Cl, &S[36 * i + 16]);
    for (int e = 0; e < 16; e++) Cl[e] = tt[e];
  }
  const LU4<double> L = lu4(Bl);
  lu4_solve(L, Fl, 1);
  for (int e = 0; e < 4; e++) f0[4 * (N - 1) + e] = Fl[e];
  double x[4];
  {
    double fv[4];
    for (int e = 0; e < 4; e++) fv[e] = S[36 * (N - 2) + 16 + e];
    vec_sub_mul(fv, &S[36 * (N - 2)], Fl);
    for (int e = 0; e < 4; e++) x[e] = f0[4 * (N - 2) + e] = fv[e];
  }
  for (int i = N - 3; i >= 0; i--) {
    const double* c = &S[36 * i];
    const double* a = &S[36 * i + 20];
    double fv[4];
    for (int k = 0; k < 4; k++) {
      double v = S[36 * i + 16 + k];
      for (int j = 0; j < 4; j++) v -= (c[4 * k + j] * x[j] + a[4 * k + j] * Fl[j]);
      fv[k] = v;
    }
    for (int e = 0; e < 4; e++) x[e] = f0[4 * i + e] = fv[e];
  }
}

// --------------------------------------------------------------------------
// host helpers
// --------------------------------------------------------------------------
template <typename V>
int upload(V** dst, const std::vector<V>& src) {
  *dst = nullptr;
  if (src.empty()) return HAMSTR_OK;
  HIPCHK(hipMalloc(dst, sizeof(V) * src.size()));
  HIPCHK(hipMemcpy(*dst, src.data(), sizeof(V) * src.size(), hipMemcpyHostToDevice));
  return HAMSTR_OK;
}

struct DevBuf {
  void* p = nullptr;
};

}  // namespace

#ifndef HAMSTR_LONG_TU
// launcher of k_adi_fast_long, defined in adi_long.hip (ct, lv: a ChainTab and
// a Live, whose layouts both translation units share)
hipError_t hamstr_adi_long_launch(bool fp32, int final_type, unsigned grid, unsigned block, size_t lds,
                                  hipStream_t stream, const void* ct, const void* lv, const int* chains, int nch,
                                  const void* ff, int nProc, double cfl, const double* sigma, double* r, double* dq,
                                  int dqmode, int ncr, void* hs, int hstride, const int* hbase);
// launcher of k_adi_fast_half / k_adi_fast_half2 (FP64, chains of at most
// kHalfMaxRows rows), in adi_long.hip
hipError_t hamstr_adi_half_launch(bool fp32, int final_type, int pairs, unsigned grid, unsigned block, size_t lds,
                                  hipStream_t stream, const void* ct, const void* lv, const void* js,
                                  const int* chains, int nch,
                                  const void* ff, int nProc, double cfl, const double* sigma, double* r, double* dq,
                                  int dqmode, int ncr, void* hs, int hstride, const int* hbase);
constexpr int kHalfMaxRows = 384;  // the workgroup bound of k_adi_fast_half, per chain of k_adi_fast_half2
// the k_adi_fast_half instance for a workgroup of `block` threads: 192 (a
// 28.5 KB LDS slice, four workgroups per CU) or 384 (57 KB)
inline int half_maxt(int block) { return block <= 192 ? 192 : kHalfMaxRows; }
// ADI kernels of a launch group (hamstr_launch_groups' second field)
enum { kKernExact = 0, kKernRow = 1, kKernHalf = 2, kKernTwo = 3, kKernHalf2 = 4 };

// ==========================================================================
// context
// ==========================================================================
struct Group {
  int phase = 0;       // ADI phase
  int fast = 0;        // LDS-resident FAST kernel (else the exact per-chain kernel)
  int final = 0;       // FAST: kFinalOpen / kFinalPair / kFinalOdd
  int off = 0, n = 0;  // slice of group_chains
  int maxN = 0;        // rows of the longest chain
  int grid = 0;        // workgroups of the launch (autotuned for FAST)
  int block = 64;      // threads per workgroup
  int ncr = 0;         // FAST: cyclic-reduction levels before the parallel cyclic reduction (autotuned)
  int two = 0;         // FAST: chains longer than a workgroup, two rows per thread (adi_fast_chain_long)
  int half_ok = 0;     // FAST, FP64: every chain fits adi_fast_chain_half (scratch rows reserved)
  int half2_ok = 0;    // FAST, open chains that fit k_adi_fast_half2 (FP64 or FP32; scratch rows reserved)
  int kern = 0;        // FAST one-row kernel: kKernRow, kKernHalf or (open chains) kKernHalf2
  int tpcr = -1;       // FAST open chains (not kKernTwo): PCR levels before the Thomas finish (-1: none; autotuned)
};

struct hamstr_ctx {
  int device = 0;
  int ncu = 1;  // compute units of the device
  hipStream_t stream = nullptr;
  hamstr_params prm{};
  int nCell = 0, nFace = 0, nChain = 0, nProc = 0, nSlot = 0, nGhost = 2, nCase = 1;
  Consts c{};
  double L2rhs = 0, Linfrhs = 0, L2lin = 0, Linflin = 0;
  std::vector<double> case_norm;  // [nCase][4] of the last read-back step
  // device arrays
  double *q = nullptr, *r = nullptr, *r0 = nullptr, *dq = nullptr, *sigma = nullptr;
  double* qt = nullptr;  // rk3 stage-start state (SOLNBLOCK::qt)
  int scheme = HAMSTR_SCHEME_ADI;
  double *flux = nullptr, *sr = nullptr, *lflux = nullptr;
  void* ff = nullptr;
  int* pslot = nullptr;
  int2* src = nullptr;
  // whether every workgroup's stencil slots fit kStageSlots, and whether the
  // fused face kernel stages them in LDS (by default where they fit)
  bool stage_fits = false, stage_faces = false;
  double *nvec = nullptr, *refmtx = nullptr;
  double4* qinf = nullptr;
  std::vector<double> qinf_h;  // [nCase][4]
  std::vector<double> case_mach, case_alpha;
  uint64_t mesh_hash = 0;      // faces, chainStart and chainConn of the (union) mesh
  int* cell_case = nullptr;
  // raw density residual per cell for the per-case norms of batches:
  // [0, nCell) nonlinear, [nCell, 2 nCell) linear (a retired case keeps both)
  double* rrho = nullptr;
  double* keep_norms = nullptr;  // batches: [nCase][4] last norms per case
  int* case_off = nullptr;
  std::vector<int> case_off_h;
  int *pleft = nullptr, *pright = nullptr, *pface = nullptr;
  int nChainProc = 0;                                       // processed faces on chains; orphans follow
  int *od_chain = nullptr, *od_off = nullptr, *od_ent = nullptr;  // orphan diagonal tables (or null)
  std::vector<int> case_orph_off;                           // [nCase+1] orphan processed-face ranges
  int *cell_off = nullptr, *cell_ent = nullptr;
  int *ch_pbase = nullptr, *ch_nf = nullptr, *ch_ncell = nullptr, *ch_periodic = nullptr;
  int* ch_hbase = nullptr;      // per chain: first row of its odd rows in long_scratch (long chains)
  void* long_scratch = nullptr; // FAST long chains: [36][nLong] odd rows eliminated by the first level
  int nLong = 0;
  int* group_chains = nullptr;
  std::vector<Group> groups;   // in phase order
  // where the fused FAST iteration's ADI kernels take the chain faces'
  // Jacobians from (HAMSTR_JAC_*; chosen by hamstr_autotune)
  int jac_src = HAMSTR_JAC_STORED;
  double* fst = nullptr;  // HAMSTR_JAC_FROM_STATES / _STATES_LINES: [8][nProc] face states
  // HAMSTR_JAC_STATES_LINES: the processed faces below jac_to are phase 0's
  // (its chains come first); 0: that source is not available
  int jac_to = 0;
  int nPhase = 0;
  bool last_phase_all_cells = false;
  // single case: ADI phase 0 covers exactly the cells [0, split_cells), whole
  // kBlock blocks, and later phases follow (else 0)
  int split_cells = 0;
  void* scratch = nullptr;     // EXACT groups: T[nChainProc][36] rows C', F', a' of the back sweep
  void* exact_rows = nullptr;  // EXACT groups: T[nChainProc][52] assembled rows (k_exact_assemble)
  double* partial = nullptr;
  int nPart = 0;
  double* npart = nullptr;  // single case: [kGraphMaxIters][2 stages][nPart][2] block partials
  double* norms = nullptr;  // [nCase][4]: L2rhs, Linfrhs, L2lin, Linflin
  double* hist = nullptr;   // [histCap][nCase][4]
  int histCap = 0;
  int* nanflag = nullptr;
  std::vector<int> proc_face;  // host: face id per processed face
  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev;
  std::vector<double> last_ms;
  std::vector<std::string> stage_names;
  // retirement of converged or stalled cases (batches)
  unsigned char *face_on = nullptr, *cell_on = nullptr, *chain_on = nullptr;
  bool any_retired = false;
  bool retire = false;
  // graph of K iterations (launch shapes and live masks are baked in)
  bool use_graph = true;
  hipGraphExec_t gexec = nullptr;
  int gK = 0;
  double* ghist = nullptr;
  int ghistCap = 0;
  int stall_window = 0;
  long it_total = 0;                   // iterations run since the last hamstr_set_retire
  std::vector<int> case_status;        // 0 live, 1 converged, 2 stalled
  std::vector<long> case_iters;        // iteration of convergence / stall detection
  std::vector<double> stall_ref;       // best norm that started the current window
  std::vector<long> stall_ref_it;
  std::vector<int> case_proc_off, case_chain_off;  // [nCase+1] processed-face and chain ranges

  Live live() const {
    return (any_retired || retire) ? Live{face_on, cell_on, chain_on} : Live{nullptr, nullptr, nullptr};
  }
  FaceTab ftab() const { return FaceTab{pslot, src, nvec, refmtx, qinf, nSlot}; }
  ChainTab ctab() const {
    return ChainTab{ch_pbase, ch_nf, ch_ncell, ch_periodic, pleft, pright, pface, refmtx,
                    od_chain, od_off, od_ent};
  }
  bool lhs32() const { return prm.lhs == HAMSTR_LHS_FP32; }
};

namespace {

// ff_from: the first processed face whose Jacobians are stored (0: all;
// nChainProc: the orphan faces only, for ADI kernels that compute the chain
// faces' Jacobians again); the faces below jac_to are stored as well
int launch_rhs_faces(hamstr_ctx* x, bool implicit, int ff_from = 0, double* fst = nullptr, int jac_to = 0) {
  const FaceTab ft = x->ftab();
  dim3 g(nblk(x->nProc)), b(kBlock);
#define RHSF(O, T, I, F)                                                                                         \
  if (x->stage_faces)                                                                                          \
    hipLaunchKernelGGL((k_rhs_faces<O, T, I, F, true>), g, b, 0, x->stream, ft, x->c, x->live(), x->q, x->flux, \
                       x->sr, (T*)x->ff, ff_from, fst, x->nanflag, jac_to);                                     \
  else                                                                                                         \
    hipLaunchKernelGGL((k_rhs_faces<O, T, I, F, false>), g, b, 0, x->stream, ft, x->c, x->live(), x->q, x->flux,\
                       x->sr, (T*)x->ff, ff_from, fst, x->nanflag, jac_to)
#define RHSO(T, I, F) \
  if (o == 1) { RHSF(1, T, I, F); } else if (o == 3) { RHSF(3, T, I, F); } else { RHSF(5, T, I, F); }
  const int o = x->prm.order;
  const bool fj = x->prm.adi_mode == HAMSTR_ADI_FAST;
  if (x->lhs32()) {
    if (!implicit) { RHSO(float, false, false); }
    else if (fj) { RHSO(float, true, true); }
    else { RHSO(float, true, false); }
  } else {
    if (!implicit) { RHSO(double, false, false); }
    else if (fj) { RHSO(double, true, true); }
    else { RHSO(double, true, false); }
  }
#undef RHSO
#undef RHSF
  HIPCHK(hipGetLastError());
  return HAMSTR_OK;
}

// recompute: the Jacobians from q (the fused iteration, where ff is
// k_rhs_faces' output for the current q), else read from ff
int launch_lin_faces(hamstr_ctx* x, bool recompute = false) {
  const FaceTab ft = x->ftab();
  dim3 g(nblk(x->nProc)), b(kBlock);
#define LINF2(O, T, F, S)                                                                                        \
  if (recompute)                                                                                                 \
    hipLaunchKernelGGL((k_lin_faces<O, T, F, true, S>), g, b, 0, x->stream, ft, x->c, x->live(), x->q, x->dq,    \
                       (const T*)x->ff, x->lflux);                                                               \
  else                                                                                                           \
    hipLaunchKernelGGL((k_lin_faces<O, T, F, false, S>), g, b, 0, x->stream, ft, x->c, x->live(), x->q, x->dq,   \
                       (const T*)x->ff, x->lflux)
#define LINF(O, T, F)          \
  if (x->stage_faces) {        \
    LINF2(O, T, F, true);      \
  } else {                     \
    LINF2(O, T, F, false);     \
  }
#define LINO(T, F) \
  if (o == 1) { LINF(1, T, F); } else if (o == 3) { LINF(3, T, F); } else { LINF(5, T, F); }
  const int o = x->prm.order;
  const bool fj = x->prm.adi_mode == HAMSTR_ADI_FAST;
  if (x->lhs32()) {
    if (fj) { LINO(float, true); } else { LINO(float, false); }
  } else {
    if (fj) { LINO(double, true); } else { LINO(double, false); }
  }
#undef LINO
#undef LINF
#undef LINF2
  HIPCHK(hipGetLastError());
  return HAMSTR_OK;
}

// norms of r into base[4*k + slot], base[4*k + slot + 1] for every case k
int launch_norm(hamstr_ctx* x, double* base, int slot) {
  if (x->nCase == 1)
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(kRedBlock), 0, x->stream, x->partial, x->nPart, x->nCell,
                       base + slot);
  else
    hipLaunchKernelGGL(k_case_norms, dim3(x->nCase, 1), dim3(kRedBlock), 0, x->stream, x->rrho, x->nCell, x->case_off,
                       slot, base, x->live(), x->keep_norms);
  HIPCHK(hipGetLastError());
  return HAMSTR_OK;
}

// part == nullptr: block partials to x->partial and the norms at once into
// normbase; otherwise block partials to part, reduced later (k_norm_slots /
// k_case_norms) by the caller
// k_rhs_cells over cells [c0, c1) (c0 a multiple of kBlock), block partials to pp
int launch_rhs_cells_range(hamstr_ctx* x, bool adi_init, double* pp, int c0, int c1) {
  if (c1 <= c0) return HAMSTR_OK;
  const dim3 g(nblk(c1 - c0)), b(kBlock);
  if (adi_init)
    hipLaunchKernelGGL((k_rhs_cells<true>), g, b, 0, x->stream, x->c, x->live(), c0, x->cell_off, x->cell_ent,
                       x->flux, x->sr, x->r, x->r0, x->sigma, pp, x->rrho, c1);
  else
    hipLaunchKernelGGL((k_rhs_cells<false>), g, b, 0, x->stream, x->c, x->live(), c0, x->cell_off, x->cell_ent,
                       x->flux, x->sr, x->r, x->r0, x->sigma, pp, x->rrho, c1);
  HIPCHK(hipGetLastError());
  return HAMSTR_OK;
}

int launch_rhs_cells(hamstr_ctx* x, bool adi_init, double* normbase, double* part = nullptr) {
  double* pp = part ? part : x->partial;
  const int rc = launch_rhs_cells_range(x, adi_init, pp, 0, x->nCell);
  if (rc) return rc;
  return part ? HAMSTR_OK : launch_norm(x, normbase, 0);
}

// Workgroups the fused iteration adds to ADI phase 0's launch for the residual
// sums of the cells phase 0 does not cover (0: none).  Only for a single case
// whose phase 0 is one group of the one-row kernel over the leading cells
// [0, split_cells) on fewer workgroups than the device has CUs: the loops of
// an O-grid then leave CUs idle (one 512-row loop per CU, 128 loops).
int cells_overlap_extra(const hamstr_ctx* x) {
  if (x->split_cells <= 0 || x->nCase != 1) return 0;
  const Group* g0 = nullptr;
  int n0 = 0;
  for (const Group& g : x->groups)
    if (g.phase == 0) {
      g0 = &g;
      n0++;
    }
  if (n0 != 1 || !g0->fast || g0->two || g0->kern == kKernHalf || g0->kern == kKernHalf2) return 0;
  if (g0->block % kBlock != 0 || g0->block > 1024) return 0;
  const int used = std::min(g0->grid, g0->n);
  return used < x->ncu ? x->ncu - used : 0;
}

// Whether every ADI launch group can compute its chain faces' Jacobians from
// q: FAST kernels other than the long-chain one, and not FP32 odd cycles,
// whose one-row kernel, held to 128 VGPRs, would spill some 500 registers
// computing them.  The fused iteration does so (and stores the orphan faces'
// Jacobians only) when this holds and x->jac_src asks for it.
bool jac_recompute_possible(const hamstr_ctx* x);
// the source the fused iteration uses now: the chosen one where possible
int jac_source(const hamstr_ctx* x) {
  return jac_recompute_possible(x) ? x->jac_src : HAMSTR_JAC_STORED;
}

// whether the face kernels store face states for source rj
bool jac_states(int rj) { return rj == HAMSTR_JAC_FROM_STATES || rj == HAMSTR_JAC_STATES_LINES; }

bool jac_recompute_possible(const hamstr_ctx* x) {
  if (x->prm.adi_mode != HAMSTR_ADI_FAST || x->groups.empty()) return false;
  for (const Group& g : x->groups)
    if (!g.fast || g.two || (x->lhs32() && g.final == kFinalOdd)) return false;
  return true;
}

// jsrc: where the group's chain-face Jacobians come from (HAMSTR_JAC_*; the
// fused iteration's choice, else read from ff)
// cells: k_rhs_cells<true>'s work on cells [c0, c1), done in the same launch
// by `extra` workgroups after the chains' (the one-row kernel only; null or
// extra == 0: none)
int launch_group(hamstr_ctx* x, const Group& g, int grid, int dqmode, int ncr, int jsrc = HAMSTR_JAC_STORED,
                 const CellWork* cells = nullptr, int extra = 0) {
  // HAMSTR_JAC_STATES_LINES: phase 0 reads its stored Jacobians, the other
  // phases compute theirs from the stored face states
  if (jsrc == HAMSTR_JAC_STATES_LINES) jsrc = g.phase == 0 ? HAMSTR_JAC_STORED : HAMSTR_JAC_FROM_STATES;
  const bool recompute = jsrc != HAMSTR_JAC_STORED;
  const ChainTab ct = x->ctab();
  const int* chains = x->group_chains + g.off;
  CellWork cw{};
  if (cells && extra > 0) {
    cw = *cells;
    cw.gl = grid;
  } else {
    extra = 0;
  }
  if (g.fast) {
    const JacSrc js{x->ftab(), x->q, x->c.gamma, x->c.qt, x->c.eps, recompute ? x->prm.order : 0,
                    jsrc == HAMSTR_JAC_FROM_STATES ? x->fst : nullptr, x->nProc};
    const dim3 gr(grid), b(g.block);
    // long chains always run the first reduction level (it is what halves
    // them), and so does the half-LDS kernel
    const bool half = g.kern == kKernHalf || g.kern == kKernHalf2;
    if (extra > 0 && (half || g.two)) return fail(HAMSTR_ERR_ARG, "cell work rides in the one-row kernel only");
    const int nc = depth_code((g.two || half) ? std::max(ncr, 1) : ncr,
                              (g.final == kFinalOpen && !g.two) ? g.tpcr : -1);
    // half2: two slices, then each chain's odd rows kept in LDS
    const size_t kodd = x->lhs32() ? half2_odd_lds<float, kHalfMaxRows>() : half2_odd_lds<double, kHalfMaxRows>();
    const size_t lds = (x->lhs32() ? sizeof(float) : sizeof(double)) * 36 *
                       (half ? (g.kern == kKernHalf2 ? 2 * (size_t)lds_stride((kHalfMaxRows + 1) / 2) + 2 * (kodd + 1)
                                                     : (size_t)lds_stride((half_maxt(g.block) + 1) / 2))
                             : (size_t)lds_stride(g.two ? (g.maxN + 1) / 2 : g.maxN));
#define FASTJ(T, F, M, J)                                                                                          \
  hipLaunchKernelGGL((k_adi_fast<T, F, M, J>), dim3(grid + extra), b, lds, x->stream, ct, x->live(), js, chains,    \
                     g.n, (const T*)x->ff, x->nProc, x->c.cfl, x->sigma, x->r, x->dq, dqmode, nc, cw, x->c)
#define FASTT(T, F, M)             \
  do {                             \
    if (recompute)                 \
      FASTJ(T, F, M, true);        \
    else                           \
      FASTJ(T, F, M, false);       \
  } while (0)
    if (half) {
      const Live lv = x->live();
      const bool pairs = g.kern == kKernHalf2;
      const int gw = pairs ? std::min(grid, (g.n + 1) / 2) : grid;  // workgroups of two chains
      HIPCHK(hamstr_adi_half_launch(x->lhs32(), g.final, pairs, (unsigned)gw, pairs ? 2u * kHalfMaxRows : (unsigned)g.block, lds,
                                    x->stream, &ct, &lv, &js, chains, g.n, x->ff, x->nProc, x->c.cfl, x->sigma, x->r,
                                    x->dq, dqmode, nc, x->long_scratch, x->nLong, x->ch_hbase));
    } else if (g.two) {
      const Live lv = x->live();
      HIPCHK(hamstr_adi_long_launch(x->lhs32(), g.final, (unsigned)grid, (unsigned)g.block, lds, x->stream, &ct, &lv,
                                    chains, g.n, x->ff, x->nProc, x->c.cfl, x->sigma, x->r, x->dq, dqmode, nc,
                                    x->long_scratch, x->nLong, x->ch_hbase));
    } else if (x->lhs32() && g.final == kFinalOdd && recompute) {
      return fail(HAMSTR_ERR_ARG, "FP32 odd cycles read their Jacobians from ff");
    } else if (x->lhs32() && g.final == kFinalOdd && g.block <= 384 && grid <= x->ncu) {
      // odd cycles in one round of workgroups (one per CU): without the
      // 128-VGPR bound, which would only add spills here
      FASTJ(float, kFinalOdd, 384, false);
    } else if (x->lhs32() && g.block <= 512) {
      // up to 512 threads: held to 128 VGPRs, so that two 384-row
      // workgroups share a CU
      if (g.final == kFinalOpen) FASTT(float, kFinalOpen, 512);
      else if (g.final == kFinalPair) FASTT(float, kFinalPair, 512);
      else FASTJ(float, kFinalOdd, 512, false);
    } else if (x->lhs32()) {
      if (g.final == kFinalOpen) FASTT(float, kFinalOpen, 1024);
      else if (g.final == kFinalPair) FASTT(float, kFinalPair, 1024);
      else FASTJ(float, kFinalOdd, 1024, false);
    } else {
      if (g.final == kFinalOpen) FASTT(double, kFinalOpen, 512);
      else if (g.final == kFinalPair) FASTT(double, kFinalPair, 512);
      else FASTT(double, kFinalOdd, 512);
    }
#undef FASTT
#undef FASTJ
  } else if (x->lhs32()) {
    hipLaunchKernelGGL((k_exact_assemble<float>), dim3(g.n), dim3(256), 0, x->stream, ct, x->live(), chains, g.n,
                       (const float*)x->ff, x->nProc, x->c.cfl, x->sigma, x->r, (float*)x->exact_rows);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL((k_adi_exact16<float>), dim3(nblk(g.n, 4)), dim3(64), 0, x->stream, ct, x->live(), chains,
                       g.n, (const float*)x->exact_rows, x->r, (float*)x->scratch, x->dq, dqmode);
  } else {
    hipLaunchKernelGGL((k_exact_assemble<double>), dim3(g.n), dim3(256), 0, x->stream, ct, x->live(), chains, g.n,
                       (const double*)x->ff, x->nProc, x->c.cfl, x->sigma, x->r, (double*)x->exact_rows);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL((k_adi_exact16<double>), dim3(nblk(g.n, 4)), dim3(64), 0, x->stream, ct, x->live(), chains,
                       g.n, (const double*)x->exact_rows, x->r, (double*)x->scratch, x->dq, dqmode);
  }
  HIPCHK(hipGetLastError());
  return HAMSTR_OK;
}

// all phases of one sweep; the last phase also accumulates dq when it covers
// every cell, otherwise a separate kernel does (ADI.cpp:57-248)
// cells / extra: residual sums added to phase 0's launch (launch_group)
int launch_sweep_phases(hamstr_ctx* x, int sweep, const std::function<int(const char*)>& mark,
                        int jsrc = HAMSTR_JAC_STORED, const CellWork* cells = nullptr, int extra = 0) {
  for (size_t k = 0; k < x->groups.size(); k++) {
    const Group& g = x->groups[k];
    const bool fuse = (g.phase == x->nPhase - 1) && x->last_phase_all_cells;
    const int rc = launch_group(x, g, g.grid, fuse ? (sweep == 0 ? 1 : 2) : 0, g.ncr, jsrc,
                                g.phase == 0 ? cells : nullptr, g.phase == 0 ? extra : 0);
    if (rc) return rc;
    const bool phase_end = (k + 1 == x->groups.size()) || x->groups[k + 1].phase != g.phase;
    if (mark && phase_end &&
        (g.phase == 0 ? mark("adi_phase0") : (g.phase == 1 ? mark("adi_phase1") : mark("adi_phase"))))
      return HAMSTR_ERR_HIP;
  }
  if (!x->last_phase_all_cells) {
    const int n4 = 4 * x->nCell;
    if (sweep == 0)
      hipLaunchKernelGGL((k_accum_dq<true>), dim3(nblk(n4)), dim3(kBlock), 0, x->stream, n4, x->live(), x->r, x->dq);
    else
      hipLaunchKernelGGL((k_accum_dq<false>), dim3(nblk(n4)), dim3(kBlock), 0, x->stream, n4, x->live(), x->r, x->dq);
    HIPCHK(hipGetLastError());
    if (mark && mark("accum_dq")) return HAMSTR_ERR_HIP;
  }
  return HAMSTR_OK;
}

int launch_lin_cells(hamstr_ctx* x, bool update, double* normbase, double* part = nullptr) {
  dim3 g(nblk(x->nCell)), b(kBlock);
  double* pp = part ? part : x->partial;
  if (update)
    hipLaunchKernelGGL((k_lin_cells<true>), g, b, 0, x->stream, x->c, x->live(), x->cell_off, x->cell_ent, x->lflux, x->r0,
                       x->dq, x->sigma, x->r, x->q, pp, x->rrho ? x->rrho + x->nCell : nullptr);
  else
    hipLaunchKernelGGL((k_lin_cells<false>), g, b, 0, x->stream, x->c, x->live(), x->cell_off, x->cell_ent, x->lflux,
                       x->r0, x->dq, x->sigma, x->r, x->q, pp, x->rrho ? x->rrho + x->nCell : nullptr);
  HIPCHK(hipGetLastError());
  return part ? HAMSTR_OK : launch_norm(x, normbase, 2);
}

int launch_linear(hamstr_ctx* x, bool update, double* normbase) {
  int rc = launch_lin_faces(x);
  if (rc) return rc;
  return launch_lin_cells(x, update, normbase);
}

// the ADI sweeps after r/r0 have been scaled (ADI.cpp:55-264)
int launch_sweeps(hamstr_ctx* x, double* normbase) {
  for (int s = 0; s < x->prm.msweep; s++) {
    int rc = launch_sweep_phases(x, s, nullptr);
    if (rc) return rc;
    rc = launch_linear(x, s == x->prm.msweep - 1, normbase);
    if (rc) return rc;
  }
  return HAMSTR_OK;
}

int read_norms(hamstr_ctx* x) {
  std::vector<double> h(4 * (size_t)x->nCase);
  HIPCHK(hipMemcpyAsync(h.data(), x->norms, sizeof(double) * h.size(), hipMemcpyDeviceToHost, x->stream));
  int flag = 0;
  HIPCHK(hipMemcpyAsync(&flag, x->nanflag, sizeof(int), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  x->case_norm = h;
  x->L2rhs = h[0];
  x->Linfrhs = h[1];
  x->L2lin = h[2];
  x->Linflin = h[3];
  if (flag) return fail(HAMSTR_ERR_NAN, "NaN spectral radius in the Roe flux");
  return HAMSTR_OK;
}

void free_ctx(hamstr_ctx* x) {
  void* ptrs[] = {x->q, x->qt, x->r, x->r0, x->dq, x->sigma, x->flux, x->sr, x->lflux, x->ff, x->pslot, x->src,
                  x->nvec, x->refmtx, x->qinf, x->cell_case, x->case_off, x->pleft, x->pright, x->pface,
                  x->cell_off, x->cell_ent, x->ch_pbase, x->ch_nf, x->ch_ncell, x->ch_periodic, x->group_chains,
                  x->scratch, x->partial, x->norms, x->hist, x->nanflag, x->rrho, x->face_on, x->cell_on,
                  x->chain_on, x->ghist, x->npart, x->keep_norms, x->od_chain, x->od_off, x->od_ent,
                  x->ch_hbase, x->long_scratch, x->exact_rows, x->fst};
  if (x->gexec) (void)hipGraphExecDestroy(x->gexec);
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto e : x->ev) (void)hipEventDestroy(e);
  if (x->stream) (void)hipStreamDestroy(x->stream);
}

// FNV-1a, 64 bit, over a byte range
constexpr uint64_t kFnvBasis = 1469598103934665603ULL;
uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; i++) {
    h ^= b[i];
    h *= 1099511628211ULL;
  }
  return h;
}

struct CaseInfo {
  int cellBase, cellCount;
  int chainBase, chainCount;
  double mach, alpha;
};

// freestream (initFlow.cpp:107-142)
void freestream_state(double mach, double alpha, double* q4) {
  const double gamma = 1.4, gm1 = gamma - 1., rinf = 1.0, pinf = 1.0 / gamma;
  const double PI = 3.1415926535897932e+00;
  const double uinf = mach * std::cos(alpha * (PI / 180.0));
  const double vinf = mach * std::sin(alpha * (PI / 180.0));
  const double einf = pinf / (gm1) + 0.5 * rinf * (uinf * uinf + vinf * vinf);
  q4[0] = rinf;
  q4[1] = rinf * uinf;
  q4[2] = rinf * vinf;
  q4[3] = einf;
}

// Builds a context for the disjoint union of one or more cases' meshes.
int create_internal(const hamstr_mesh_desc* mesh, const hamstr_params& P, const std::vector<CaseInfo>& cases,
                    hamstr_ctx** out) {
  *out = nullptr;
  if (P.order != 1 && P.order != 3 && P.order != 5) return fail(HAMSTR_ERR_ARG, "order must be 1, 3 or 5");
  if (P.msweep < 1) return fail(HAMSTR_ERR_ARG, "msweep must be >= 1");
  if (P.lhs != HAMSTR_LHS_FP64 && P.lhs != HAMSTR_LHS_FP32) return fail(HAMSTR_ERR_ARG, "bad lhs precision");
  if (P.adi_mode != HAMSTR_ADI_EXACT && P.adi_mode != HAMSTR_ADI_FAST) return fail(HAMSTR_ERR_ARG, "bad adi mode");
  const int nCell = mesh->nCell, nFace = mesh->nFace, nChain = mesh->nChain;
  if (nCell <= 0 || nFace <= 0 || nChain <= 0 || !mesh->faces || !mesh->chainStart || !mesh->chainConn ||
      !mesh->normVec || !mesh->refMtx)
    return fail(HAMSTR_ERR_ARG, "hamstr_create: bad mesh description");
  const int* faces = mesh->faces;
  const int* fs = mesh->chainStart;
  const int* cc = mesh->chainConn;
  const int nGhost = (P.order == 5) ? 3 : 2;
  const int nChainFace = fs[nChain];
  const int nCase = (int)cases.size();
  if (fs[0] != 0 || nChainFace <= 0) return fail(HAMSTR_ERR_MESH, "chainStart must start at 0");
  for (int i = 0; i < nChain; i++)
    if (fs[i + 1] - fs[i] < 2) return fail(HAMSTR_ERR_MESH, "chain with fewer than 2 faces");
  for (int f = 0; f < nChainFace; f++)
    if (cc[f] < 0 || cc[f] >= nFace) return fail(HAMSTR_ERR_MESH, "chainConn entry out of range");
  for (int i = 0; i < nFace; i++) {
    const int L = faces[6 * i + 2], R = faces[6 * i + 4];
    if (L < 0 || L >= nCell || R >= nCell) return fail(HAMSTR_ERR_MESH, "face cell out of range");
    if (R == -5) return fail(HAMSTR_ERR_MESH, "inter-domain faces (-5) are not supported");
  }
  std::vector<int> chain_case(nChain, 0), cell_case(nCell, 0), case_off(nCase + 1, 0);
  for (int k = 0; k < nCase; k++) {
    for (int i = cases[k].chainBase; i < cases[k].chainBase + cases[k].chainCount; i++) chain_case[i] = k;
    for (int c = cases[k].cellBase; c < cases[k].cellBase + cases[k].cellCount; c++) cell_case[c] = k;
    case_off[k] = cases[k].cellBase;
    case_off[k + 1] = cases[k].cellBase + cases[k].cellCount;
  }

  // ---- chains, slots and processed faces (computeRHS.cpp:53-298) ----
  std::vector<int> ch_pbase(nChain), ch_nf(nChain), ch_ncell(nChain), ch_periodic(nChain);
  std::vector<int2> src;
  std::vector<int> pslot, pleft, pright, pface;
  std::vector<double> nvec;
  std::vector<int> proc_count(nFace, 0);
  std::vector<int> cindx;
  for (int i = 0; i < nChain; i++) {
    const int f1 = fs[i], f2 = fs[i + 1];
    const int periodic = (cc[f1] == cc[f2 - 1]);
    const int nf = f2 - f1 - periodic;
    const int ncell = f2 - f1 - 1;
    if (periodic && ncell < 3) return fail(HAMSTR_ERR_MESH, "closed chain with fewer than 3 cells");
    if (!periodic && ncell < 1) return fail(HAMSTR_ERR_MESH, "open chain without cells");
    // cindx exactly as the reference builds it (cell 0 quirks included)
    cindx.assign((f2 - f1) + 8, 0);
    int m = nGhost;
    for (int f = f1; f < f2; f++) cindx[m++] = faces[6 * cc[f] + 2];
    int chainSize;
    if (periodic) {
      cindx[m++] = faces[6 * cc[f1 + 1] + 2];
      chainSize = m;
      m = 0;
      for (int f = f2 - nGhost - 1; f < f2 - 1; f++) cindx[m++] = faces[6 * cc[f] + 2];
    } else if (P.order == 5) {
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
    // the reference's cell-0 tests apply to global cell 0 only; in a batch that
    // is the first case's cell 0 (meshgen never puts cell 0 where they differ)
    const int slotBase = (int)src.size();
    for (int j = 0; j < chainSize; j++) {
      int icell = cindx[j];
      int2 e = make_int2(icell, -1);
      const bool copy = icell >= 0 || (icell == 0 && j == nGhost) || (icell == 0 && j == chainSize - nGhost - 1);
      const bool ghost = icell < 0 || (icell == 0 && j == nGhost - 1) || (icell == 0 && j == chainSize - nGhost);
      if (!copy && !ghost) return fail(HAMSTR_ERR_MESH, "internal: slot without source");
      if (ghost) {
        const int iface = (j < nGhost) ? cc[f1] : cc[f2 - 1];
        const int rightCell = faces[6 * iface + 4];
        if (rightCell == -2)
          e = make_int2(-icell, iface);
        else
          e = make_int2(-1, chain_case[i]);  // freestream of this chain's case
      }
      src.push_back(e);
    }
    ch_pbase[i] = (int)pslot.size();
    ch_nf[i] = nf;
    ch_ncell[i] = ncell;
    ch_periodic[i] = periodic;
    const int is = nGhost - 1;
    for (int k = 0; k < nf; k++) {
      const int f = f1 + k;
      const int iface = cc[f];
      const int flip = (!periodic && f == f2 - 1) ? 1 : 0;
      const int n = slotBase + is + k;
      pslot.push_back(n | (flip << 30));
      pleft.push_back(faces[6 * iface + 2]);
      pright.push_back(faces[6 * iface + 4]);
      pface.push_back(iface);
      nvec.push_back(mesh->normVec[2 * iface]);
      nvec.push_back(mesh->normVec[2 * iface + 1]);
      if (++proc_count[iface] > 1) return fail(HAMSTR_ERR_MESH, "face processed by more than one chain");
    }
  }
  const int nChainProc = (int)pslot.size();
  // ---- orphan faces (extension for meshes with triangles): interior faces in
  // no chain, processed after every chain face.  Their processed-face slots
  // follow the ADI rows that take their diagonal blocks (case, then chain,
  // then row), so that a chain's rows read them from consecutive addresses;
  // every per-cell sum still takes them in list order (cell_ent, od_ent). ----
  if (mesh->nOrphan < 0 || (mesh->nOrphan > 0 && !mesh->orphans))
    return fail(HAMSTR_ERR_ARG, "bad orphan face list");
  std::vector<int> orph_order(mesh->nOrphan), orph_p(mesh->nOrphan);
  {
    std::vector<long> cell_rank(nCell, std::numeric_limits<long>::max());
    long rank = 0;
    for (int i = 0; i < nChain; i++)  // the last chain listing a cell takes its blocks
      for (int k = 0; k < ch_ncell[i]; k++) cell_rank[pleft[ch_pbase[i] + k]] = rank++;
    std::vector<long> key(mesh->nOrphan);
    for (int o = 0; o < mesh->nOrphan; o++) {
      const int iface = mesh->orphans[o];
      if (iface < 0 || iface >= nFace) return fail(HAMSTR_ERR_MESH, "orphan face out of range");
      if (faces[6 * iface + 4] < 0) return fail(HAMSTR_ERR_MESH, "orphan faces must be interior faces");
      key[o] = std::min(cell_rank[faces[6 * iface + 2]], cell_rank[faces[6 * iface + 4]]);
      orph_order[o] = o;
    }
    std::stable_sort(orph_order.begin(), orph_order.end(), [&](int a, int b) {
      const int ka = cell_case[faces[6 * mesh->orphans[a] + 2]], kb = cell_case[faces[6 * mesh->orphans[b] + 2]];
      return ka != kb ? ka < kb : key[a] < key[b];
    });
  }
  for (int n = 0; n < mesh->nOrphan; n++) {
    const int o = orph_order[n];
    const int iface = mesh->orphans[o];
    if (++proc_count[iface] > 1) return fail(HAMSTR_ERR_MESH, "orphan face also on a chain or listed twice");
    orph_p[o] = (int)pslot.size();
    pslot.push_back((int)src.size() | kOrphanSlot);  // the two cells, as plain copies
    src.push_back(make_int2(faces[6 * iface + 2], -1));
    src.push_back(make_int2(faces[6 * iface + 4], -1));
    pleft.push_back(faces[6 * iface + 2]);
    pright.push_back(faces[6 * iface + 4]);
    pface.push_back(iface);
    nvec.push_back(mesh->normVec[2 * iface]);
    nvec.push_back(mesh->normVec[2 * iface + 1]);
  }
  const int nProc = (int)pslot.size();
  const int nSlot = (int)src.size();
  if (nSlot > kSlotMask) return fail(HAMSTR_ERR_MESH, "too many chain slots");

  // ---- per-cell accumulation lists in processing order ----
  std::vector<int> cnt(nCell + 1, 0);
  for (int p = 0; p < nProc; p++) {
    cnt[pleft[p] + 1]++;
    if (pright[p] > -1) cnt[pright[p] + 1]++;
  }
  for (int i = 0; i < nCell; i++) cnt[i + 1] += cnt[i];
  std::vector<int> cell_ent(cnt[nCell]);
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (int n = 0; n < nProc; n++) {  // chain faces, then the orphans in list order
    const int p = n < nChainProc ? n : orph_p[n - nChainProc];
    cell_ent[fill[pleft[p]]++] = (p << 1) | 0;
    if (pright[p] > -1) cell_ent[fill[pright[p]]++] = (p << 1) | 1;
  }
  // orphan diagonal blocks go to the rows of the last chain listing the cell
  std::vector<int> od_chain, od_off, od_ent;
  if (nProc > nChainProc) {
    od_chain.assign(nCell, -1);
    for (int i = 0; i < nChain; i++)
      for (int k = 0; k < ch_ncell[i]; k++) od_chain[pleft[ch_pbase[i] + k]] = ch_pbase[i];
    od_off.assign(nCell + 1, 0);
    for (int p = nChainProc; p < nProc; p++) {
      od_off[pleft[p] + 1]++;
      od_off[pright[p] + 1]++;
    }
    for (int i = 0; i < nCell; i++) od_off[i + 1] += od_off[i];
    od_ent.assign(od_off[nCell], 0);
    std::vector<int> at(od_off.begin(), od_off.end() - 1);
    for (int o = 0; o < nProc - nChainProc; o++) {  // list order
      const int p = orph_p[o];
      od_ent[at[pleft[p]]++] = (p << 1) | 1;
      od_ent[at[pright[p]]++] = (p << 1);
    }
  }

  // ---- ADI phases: chain i runs after every earlier chain sharing a cell ----
  std::vector<int> last(nCell, -1), phase(nChain, 0);
  int nPhase = 0;
  for (int i = 0; i < nChain; i++) {
    int ph = 0;
    for (int k = 0; k < ch_ncell[i]; k++) ph = std::max(ph, last[pleft[ch_pbase[i] + k]] + 1);
    for (int k = 0; k < ch_ncell[i]; k++) last[pleft[ch_pbase[i] + k]] = ph;
    phase[i] = ph;
    nPhase = std::max(nPhase, ph + 1);
  }
  // ---- launch groups: per phase; FAST groups by final-solve type, chains
  // sorted longest first for the persistent launches ----
  const size_t tl = (P.lhs == HAMSTR_LHS_FP32) ? sizeof(float) : sizeof(double);
  size_t lds_max = 64 * 1024;
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) == hipSuccess && prop.sharedMemPerBlock > 0) lds_max = prop.sharedMemPerBlock;
  }
  const int maxt = (P.lhs == HAMSTR_LHS_FP32) ? 1024 : 512;
  std::vector<Group> groups;
  std::vector<int> group_chains;
  size_t maxRows = 0;
  std::vector<int> ch_hbase(nChain, 0);
  long nLong = 0;  // odd rows of the long chains, the rows their scratch holds
  for (int ph = 0; ph < nPhase; ph++) {
    // buckets: 0 exact, 1 fast open, 2 fast pair, 3 fast odd; 4-6 the same
    // three for chains longer than a workgroup (two rows per thread, half in LDS)
    std::vector<int> bucket[7];
    for (int i = 0; i < nChain; i++) {
      if (phase[i] != ph) continue;
      const int N = ch_ncell[i];
      const bool fits = tl * 36 * (size_t)lds_stride(N) <= lds_max;
      const bool fast_ok = P.adi_mode == HAMSTR_ADI_FAST && N <= maxt && fits;
      const bool long_ok = P.adi_mode == HAMSTR_ADI_FAST && !fast_ok && N <= 2 * maxt &&
                           tl * 36 * (size_t)lds_stride((N + 1) / 2) <= lds_max &&
                           (ch_periodic[i] ? (N % 2 == 0 && N >= 4) : N >= 2);
      int kind;
      if (!fast_ok && !long_ok) {
        kind = 0;
      } else if (!ch_periodic[i]) {
        kind = 1;
      } else {
        int nn = N;
        while (nn % 2 == 0 && nn >= 4) nn /= 2;
        kind = nn == 2 ? 2 : 3;
      }
      if (long_ok) {
        kind += 3;
        ch_hbase[i] = (int)nLong;
        nLong += N / 2;
      }
      bucket[kind].push_back(i);
    }
    for (int b = 0; b < 7; b++) {
      if (bucket[b].empty()) continue;
      std::stable_sort(bucket[b].begin(), bucket[b].end(), [&](int a, int c) { return ch_ncell[a] > ch_ncell[c]; });
      Group g;
      g.phase = ph;
      g.fast = b > 0;
      g.two = b > 3;
      const int kind = g.two ? b - 3 : b;
      g.final = kind == 1 ? kFinalOpen : (kind == 2 ? kFinalPair : kFinalOdd);
      g.off = (int)group_chains.size();
      g.n = (int)bucket[b].size();
      g.maxN = ch_ncell[bucket[b][0]];
      const int threads = g.two ? (g.maxN + 1) / 2 : g.maxN;
      g.block = g.fast ? ((threads + 63) / 64) * 64 : 64;
      g.grid = g.fast ? g.n : nblk(g.n, 64);
      // two levels of cyclic reduction before the PCR (one before an odd
      // cycle); measured best or close to it on single meshes and batches
      g.ncr = g.fast ? (kind == 3 ? 1 : 2) : 0;
      if (g.fast && !g.two && g.maxN <= kHalfMaxRows && kind != 3) {
        // the half-LDS kernels: FP64 for one chain per workgroup (closed
        // chains of an even length of at least 4; odd cycles stay with the
        // one-row kernel, whose final cyclic Thomas solve would spill at 168
        // VGPRs), FP64 or FP32 open chains for two per workgroup
        bool ok = true;
        for (int i : bucket[b]) ok = ok && (!ch_periodic[i] || (ch_ncell[i] % 2 == 0 && ch_ncell[i] >= 4));
        const size_t slice = 36 * tl * (size_t)lds_stride((kHalfMaxRows + 1) / 2);  // one chain's LDS
        g.half_ok = ok && tl == sizeof(double) && slice <= lds_max;
        // (two per workgroup only pays for lines of more than 256 rows; it was
        // slower on every shorter length measured, and the autotuner's time
        // counts in batched runs)
        g.half2_ok = kind == 1 && g.maxN > 256 && 2 * slice <= lds_max;
        if (g.half_ok || g.half2_ok) {
          // FP64: two open chains per workgroup where one chain's waves
          // would not pack onto the SIMDs twice (over 256 rows); elsewhere,
          // and for FP32 (whose one-row workgroups already pair up on a CU),
          // the one-row kernel with its quad-lane PCR is faster
          g.kern = (g.half2_ok && g.maxN > 256 && tl == sizeof(double)) ? kKernHalf2 : kKernRow;
          for (int i : bucket[b]) {
            ch_hbase[i] = (int)nLong;
            nLong += ch_ncell[i] / 2;
          }
        }
      }
      for (int i : bucket[b]) group_chains.push_back(i);
      groups.push_back(g);
      if (!g.fast) maxRows = std::max(maxRows, (size_t)g.maxN);
    }
  }

  // scratch rows (long and half-LDS chains) are indexed with int, 36 per row
  if (nLong >= (1L << 31) / 36) return fail(HAMSTR_ERR_MESH, "too many rows in long chains");
  // every host table above is built before the device is looked for, so that
  // a machine without one (and its sanitizer builds) runs all of this code
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(HAMSTR_ERR_NODEV, "no HIP device");
  hamstr_ctx* x = new hamstr_ctx();
  x->prm = P;
  x->nCell = nCell;
  x->nFace = nFace;
  x->nChain = nChain;
  x->nProc = nProc;
  x->nChainProc = nChainProc;
  x->nLong = (int)std::max<long>(nLong, 1);
  x->nSlot = nSlot;
  {
    // the staged fused face kernel's LDS range per workgroup, as that kernel
    // forms it (first face's slot + lo .. last face's slot + hi), must hold
    // every live stencil of its faces
    const int lo = stencil_lo(P.order), hi = stencil_hi(P.order);
    bool fits = nProc > 0;
    for (int b0 = 0; fits && b0 < nProc; b0 += kBlock) {
      const int b1 = std::min(b0 + kBlock, nProc) - 1;
      const int S0 = std::max((pslot[b0] & kSlotMask) + lo, 0);
      const int S1 = std::min((pslot[b1] & kSlotMask) + hi, nSlot - 1);
      if (S1 - S0 + 1 > kStageSlots) fits = false;
      for (int p = b0; fits && p <= b1; p++) {
        const int sl = pslot[p] & kSlotMask;
        const bool orphan = pslot[p] & kOrphanSlot;
        if ((orphan ? sl : sl + lo) < S0 || (orphan ? sl + 1 : sl + hi) > S1) fits = false;
      }
    }
    x->stage_fits = x->stage_faces = fits;
  }
  x->nGhost = nGhost;
  x->nCase = nCase;
  x->nPhase = nPhase;
  {
    // HAMSTR_JAC_STATES_LINES needs phase 0's chains first: their processed
    // faces are then [0, jac_to)
    int p0_end = 0, rest_start = nChainProc;
    for (int i = 0; i < nChain; i++) {
      if (phase[i] == 0)
        p0_end = std::max(p0_end, ch_pbase[i] + ch_nf[i]);
      else
        rest_start = std::min(rest_start, ch_pbase[i]);
    }
    x->jac_to = (nPhase >= 2 && p0_end <= rest_start) ? p0_end : 0;
  }
  x->groups = groups;
  x->proc_face = pface;
  x->case_off_h = case_off;
  {
    long cover = 0;
    for (int i = 0; i < nChain; i++)
      if (phase[i] == nPhase - 1) cover += ch_ncell[i];
    x->last_phase_all_cells = (cover == nCell);
  }
  if (nCase == 1 && nPhase >= 2) {
    // are phase 0's cells exactly [0, P) with P a multiple of kBlock?
    std::vector<char> in0(nCell, 0);
    long n0 = 0;
    int top = -1;
    for (int i = 0; i < nChain; i++) {
      if (phase[i] != 0) continue;
      for (int k = 0; k < ch_ncell[i]; k++) {
        const int cl = pleft[ch_pbase[i] + k];
        if (!in0[cl]) {
          in0[cl] = 1;
          n0++;
          top = std::max(top, cl);
        }
      }
    }
    if (n0 > 0 && top + 1 == n0 && n0 % kBlock == 0 && n0 < nCell) x->split_cells = (int)n0;
  }
  x->qinf_h.assign(4 * (size_t)nCase, 0.0);
  for (int k = 0; k < nCase; k++) freestream_state(cases[k].mach, cases[k].alpha, &x->qinf_h[4 * k]);
  for (int k = 0; k < nCase; k++) {
    x->case_mach.push_back(cases[k].mach);
    x->case_alpha.push_back(cases[k].alpha);
  }
  {
    // fingerprint of the mesh the state belongs to (solution files carry it)
    uint64_t h = fnv1a(kFnvBasis, &nCell, sizeof(int));
    h = fnv1a(h, faces, sizeof(int) * 6 * (size_t)nFace);
    h = fnv1a(h, fs, sizeof(int) * (size_t)(nChain + 1));
    x->mesh_hash = fnv1a(h, cc, sizeof(int) * (size_t)nChainFace);
  }
  // chains, and so processed faces, are stored case by case
  x->case_chain_off.assign(nCase + 1, nChain);
  x->case_proc_off.assign(nCase + 1, nChainProc);
  for (int k = 0; k < nCase; k++) {
    x->case_chain_off[k] = cases[k].chainBase;
    x->case_proc_off[k] = cases[k].chainCount > 0 ? ch_pbase[cases[k].chainBase] : nChainProc;
  }
  for (int k = nCase - 1; k >= 0; k--)
    if (cases[k].chainCount == 0) x->case_proc_off[k] = x->case_proc_off[k + 1];
  // orphan faces, listed case by case: [case_orph_off[k], case_orph_off[k+1])
  x->case_orph_off.assign(nCase + 1, nProc);
  for (int p = nProc - 1; p >= nChainProc; p--) x->case_orph_off[cell_case[pleft[p]]] = p;
  for (int k = nCase - 1; k >= 0; k--) x->case_orph_off[k] = std::min(x->case_orph_off[k], x->case_orph_off[k + 1]);
  for (int p = nChainProc; p < nProc; p++) {
    const int k = cell_case[pleft[p]];
    if (p < x->case_orph_off[k] || p >= x->case_orph_off[k + 1] || cell_case[pright[p]] != k) {
      free_ctx(x);
      delete x;
      return fail(HAMSTR_ERR_MESH, "orphan faces must be listed case by case");
    }
  }
  x->case_status.assign(nCase, 0);
  x->case_iters.assign(nCase, -1);
  x->stall_ref.assign(nCase, HUGE_VAL);
  x->stall_ref_it.assign(nCase, 0);
  std::vector<double4> qinf4(nCase);
  for (int k = 0; k < nCase; k++)
    qinf4[k] = make_double4(x->qinf_h[4 * k], x->qinf_h[4 * k + 1], x->qinf_h[4 * k + 2], x->qinf_h[4 * k + 3]);
  x->c.cfl = P.cfl;
  x->c.gamma = 1.4;
  x->c.qt = (P.order == 1) ? 0.0 : 0.25;
  x->c.eps = 1e-10;
  x->c.nCell = nCell;
  x->c.nProc = nProc;
  x->c.nChainProc = nChainProc;
  x->c.order = P.order;
  x->case_norm.assign(4 * (size_t)nCase, 0.0);

  int rc = HAMSTR_OK;
  auto chk = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && rc == HAMSTR_OK) rc = fail(HAMSTR_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  chk(hipGetDevice(&x->device), "hipGetDevice");
  {
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, x->device) == hipSuccess && cu > 0)
      x->ncu = cu;
  }
  chk(hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking), "hipStreamCreate");
  const size_t c4 = sizeof(double) * 4 * nCell;
  chk(hipMalloc(&x->q, c4), "alloc q");
  chk(hipMalloc(&x->qt, c4), "alloc qt");
  chk(hipMalloc(&x->r, c4), "alloc r");
  chk(hipMalloc(&x->r0, c4), "alloc r0");
  chk(hipMalloc(&x->dq, c4), "alloc dq");
  chk(hipMalloc(&x->sigma, sizeof(double) * nCell), "alloc sigma");
  chk(hipMalloc(&x->flux, sizeof(double) * 4 * nProc), "alloc flux");
  chk(hipMalloc(&x->sr, sizeof(double) * nProc), "alloc sr");
  chk(hipMalloc(&x->lflux, sizeof(double) * 4 * nProc), "alloc lflux");
  chk(hipMalloc(&x->ff, tl * 32 * nProc), "alloc ff");
  // EXACT groups: the assembled rows (52 per row) and the rows C', F', a' of
  // the back sweep (36 per row), both indexed by the row's processed-face slot
  chk(hipMalloc(&x->scratch, tl * 36 * (maxRows > 0 ? (size_t)std::max(nChainProc, 1) : 1)), "alloc scratch");
  if (maxRows > 0) chk(hipMalloc(&x->exact_rows, tl * 52 * (size_t)std::max(nChainProc, 1)), "alloc exact rows");
  chk(hipMalloc(&x->long_scratch, tl * 36 * (size_t)x->nLong), "alloc long-chain scratch");
  x->nPart = nblk(nCell);
  chk(hipMalloc(&x->partial, sizeof(double) * 2 * x->nPart), "alloc partial");
  if (nCase == 1)
    chk(hipMalloc(&x->npart, sizeof(double) * 2 * 2 * (size_t)kGraphMaxIters * x->nPart), "alloc norm partials");
  chk(hipMalloc(&x->norms, sizeof(double) * 4 * nCase), "alloc norms");
  if (nCase > 1) chk(hipMalloc(&x->rrho, sizeof(double) * 2 * (size_t)nCell), "alloc rrho");
  if (nCase > 1) chk(hipMalloc(&x->keep_norms, sizeof(double) * 4 * (size_t)nCase), "alloc case norms");
  chk(hipMalloc(&x->face_on, nProc), "alloc face mask");
  chk(hipMalloc(&x->cell_on, nCell), "alloc cell mask");
  chk(hipMalloc(&x->chain_on, nChain), "alloc chain mask");
  chk(hipMalloc(&x->nanflag, sizeof(int)), "alloc nanflag");
  if (rc) {
    free_ctx(x);
    delete x;
    return rc;
  }
  std::vector<double> rm(mesh->refMtx, mesh->refMtx + 4 * (size_t)nFace);
  if ((rc = upload(&x->pslot, pslot)) || (rc = upload(&x->src, src)) || (rc = upload(&x->nvec, nvec)) ||
      (rc = upload(&x->refmtx, rm)) || (rc = upload(&x->qinf, qinf4)) || (rc = upload(&x->cell_case, cell_case)) ||
      (rc = upload(&x->case_off, case_off)) || (rc = upload(&x->pleft, pleft)) ||
      (rc = upload(&x->pright, pright)) || (rc = upload(&x->pface, pface)) || (rc = upload(&x->cell_off, cnt)) ||
      (rc = upload(&x->cell_ent, cell_ent)) || (rc = upload(&x->ch_pbase, ch_pbase)) ||
      (rc = upload(&x->ch_nf, ch_nf)) || (rc = upload(&x->ch_ncell, ch_ncell)) ||
      (rc = upload(&x->ch_periodic, ch_periodic)) || (rc = upload(&x->group_chains, group_chains)) ||
      (rc = upload(&x->ch_hbase, ch_hbase)) ||
      (!od_chain.empty() && ((rc = upload(&x->od_chain, od_chain)) || (rc = upload(&x->od_off, od_off)) ||
                             (rc = upload(&x->od_ent, od_ent))))) {
    free_ctx(x);
    delete x;
    return rc;
  }
  chk(hipMemset(x->nanflag, 0, sizeof(int)), "memset nanflag");
  chk(hipMemset(x->norms, 0, sizeof(double) * 4 * nCase), "memset norms");
  chk(hipMemset(x->dq, 0, c4), "memset dq");
  chk(hipMemset(x->r, 0, c4), "memset r");
  chk(hipMemset(x->r0, 0, c4), "memset r0");
  chk(hipMemset(x->sigma, 0, sizeof(double) * nCell), "memset sigma");
  chk(hipMemset(x->ff, 0, tl * 32 * nProc), "memset ff");
  chk(hipMemset(x->face_on, 1, nProc), "memset face mask");
  chk(hipMemset(x->cell_on, 1, nCell), "memset cell mask");
  chk(hipMemset(x->chain_on, 1, nChain), "memset chain mask");
  if (x->keep_norms) chk(hipMemset(x->keep_norms, 0, sizeof(double) * 4 * (size_t)nCase), "memset case norms");
  if (rc) {
    free_ctx(x);
    delete x;
    return rc;
  }
  *out = x;
  return HAMSTR_OK;
}

}  // namespace

// ==========================================================================
// C ABI
// ==========================================================================
extern "C" {

const char* hamstr_last_error(void) { return g_err.c_str(); }

int hamstr_preprocess(int nNode, const double* x, int nFace, int* faces, double* normVec, double* refMtx,
                      int sol_visc) {
  if (nNode <= 0 || nFace <= 0 || !x || !faces || !normVec || !refMtx)
    return fail(HAMSTR_ERR_ARG, "hamstr_preprocess: bad arguments");
  for (int i = 0; i < nFace; i++) {
    const int i1 = faces[6 * i], i2 = faces[6 * i + 1];
    if (i1 < 0 || i1 >= nNode || i2 < 0 || i2 >= nNode) return fail(HAMSTR_ERR_ARG, "face node out of range");
    const double x1 = x[2 * i1], y1 = x[2 * i1 + 1], x2 = x[2 * i2], y2 = x[2 * i2 + 1];
    const double n0 = (y2 - y1), n1 = -(x2 - x1);
    normVec[2 * i] = n0;
    normVec[2 * i + 1] = n1;
    double dsnorm = n0 * n0 + n1 * n1;
    dsnorm = 1. / dsnorm;
    refMtx[4 * i + 0] = 1. - 2. * n0 * n0 * dsnorm;
    refMtx[4 * i + 1] = -2. * n0 * n1 * dsnorm;
    refMtx[4 * i + 2] = -2. * n1 * n0 * dsnorm;
    refMtx[4 * i + 3] = 1. - 2. * n1 * n1 * dsnorm;
    if (faces[6 * i + 4] == -1) {
      const double r1 = std::sqrt(x[2 * i1] * x[2 * i1] + x[2 * i1 + 1] * x[2 * i1 + 1]);
      const double r2 = std::sqrt(x[2 * i2] * x[2 * i2] + x[2 * i2 + 1] * x[2 * i2 + 1]);
      if (0.5 * (r1 + r2) < 3.0) faces[6 * i + 4] = -(sol_visc + 2);
    }
  }
  return HAMSTR_OK;
}

int hamstr_create(const hamstr_mesh_desc* mesh, const hamstr_params* params, hamstr_ctx** out) {
  if (!mesh || !params || !out) return fail(HAMSTR_ERR_ARG, "hamstr_create: null argument");
  *out = nullptr;
  std::vector<CaseInfo> cases(1, CaseInfo{0, mesh->nCell, 0, mesh->nChain, params->mach, params->alpha});
  const int rc = create_internal(mesh, *params, cases, out);
  if (rc) return rc;
  return hamstr_init_flow(*out);
}

int hamstr_batch_create(int ncases, const hamstr_mesh_desc* meshes, const hamstr_params* params, hamstr_ctx** out) {
  if (ncases <= 0 || !meshes || !params || !out) return fail(HAMSTR_ERR_ARG, "hamstr_batch_create: bad arguments");
  *out = nullptr;
  const hamstr_params& P0 = params[0];
  for (int k = 1; k < ncases; k++)
    if (params[k].cfl != P0.cfl || params[k].order != P0.order || params[k].msweep != P0.msweep ||
        params[k].lhs != P0.lhs || params[k].adi_mode != P0.adi_mode)
      return fail(HAMSTR_ERR_ARG, "batched cases must share cfl, order, msweep, lhs and adi_mode");
  // disjoint union of the meshes
  long nc = 0, nf = 0, nch = 0, ncf = 0, nor = 0;
  for (int k = 0; k < ncases; k++) {
    const hamstr_mesh_desc& m = meshes[k];
    if (m.nCell <= 0 || m.nFace <= 0 || m.nChain <= 0 || !m.faces || !m.chainStart || !m.chainConn ||
        !m.normVec || !m.refMtx || m.nOrphan < 0 || (m.nOrphan > 0 && !m.orphans))
      return fail(HAMSTR_ERR_ARG, "hamstr_batch_create: bad mesh description");
    nc += m.nCell;
    nf += m.nFace;
    nch += m.nChain;
    ncf += m.chainStart[m.nChain];
    nor += m.nOrphan;
  }
  if (nc >= (1L << 30) || nf >= (1L << 29)) return fail(HAMSTR_ERR_ARG, "batch too large");
  std::vector<int> faces(6 * nf), fs(nch + 1), conn(ncf), orph;
  orph.reserve(nor);
  std::vector<double> nv(2 * nf), rm(4 * nf);
  std::vector<CaseInfo> cases(ncases);
  long c0 = 0, f0 = 0, h0 = 0, cf0 = 0;
  fs[0] = 0;
  for (int k = 0; k < ncases; k++) {
    const hamstr_mesh_desc& m = meshes[k];
    for (long i = 0; i < m.nFace; i++) {
      const int* s = m.faces + 6 * i;
      int* d = faces.data() + 6 * (f0 + i);
      d[0] = s[0];
      d[1] = s[1];
      d[2] = s[2] + (int)c0;
      d[3] = s[3];
      d[4] = s[4] >= 0 ? s[4] + (int)c0 : s[4];
      d[5] = s[5];
    }
    std::memcpy(nv.data() + 2 * f0, m.normVec, sizeof(double) * 2 * m.nFace);
    std::memcpy(rm.data() + 4 * f0, m.refMtx, sizeof(double) * 4 * m.nFace);
    for (int i = 0; i < m.nChain; i++) fs[h0 + i + 1] = (int)cf0 + m.chainStart[i + 1];
    for (int f = 0; f < m.chainStart[m.nChain]; f++) conn[cf0 + f] = m.chainConn[f] + (int)f0;
    for (int o = 0; o < m.nOrphan; o++) orph.push_back(m.orphans[o] + (int)f0);
    cases[k] = CaseInfo{(int)c0, m.nCell, (int)h0, m.nChain, params[k].mach, params[k].alpha};
    c0 += m.nCell;
    f0 += m.nFace;
    h0 += m.nChain;
    cf0 += m.chainStart[m.nChain];
  }
  hamstr_mesh_desc u{(int)nc,     (int)nf,   (int)nch,  faces.data(),        fs.data(), conn.data(),
                     nv.data(),    rm.data(), (int)nor,  orph.empty() ? nullptr : orph.data()};
  const int rc = create_internal(&u, P0, cases, out);
  if (rc) return rc;
  return hamstr_init_flow(*out);
}

int hamstr_batch_info(const hamstr_ctx* x, int* ncases, int* cell_off) {
  if (!x || !ncases) return fail(HAMSTR_ERR_ARG, "null argument");
  *ncases = x->nCase;
  if (cell_off)
    for (int k = 0; k <= x->nCase; k++) cell_off[k] = x->case_off_h[k];
  return HAMSTR_OK;
}

int hamstr_destroy(hamstr_ctx* x) {
  if (!x) return HAMSTR_OK;
  (void)hipStreamSynchronize(x->stream);
  free_ctx(x);
  delete x;
  return HAMSTR_OK;
}

int hamstr_case_freestream(const hamstr_ctx* x, int k, double* out4) {
  if (!x || !out4 || k < 0 || k >= x->nCase) return fail(HAMSTR_ERR_ARG, "bad arguments");
  for (int j = 0; j < 4; j++) out4[j] = x->qinf_h[4 * k + j];
  return HAMSTR_OK;
}

int hamstr_freestream(const hamstr_ctx* x, double* out4) {
  if (!x || !out4) return fail(HAMSTR_ERR_ARG, "null argument");
  for (int k = 0; k < 4; k++) out4[k] = x->qinf_h[k];
  return HAMSTR_OK;
}

int hamstr_init_flow(hamstr_ctx* x) {
  if (!x) return fail(HAMSTR_ERR_ARG, "null context");
  hipLaunchKernelGGL(k_init_flow, dim3(nblk(x->nCell)), dim3(kBlock), 0, x->stream, x->nCell, x->cell_case, x->qinf,
                     x->q, x->qt);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(x->stream));
  return HAMSTR_OK;
}

int hamstr_set_field(hamstr_ctx* x, int field, const double* host) {
  if (!x || !host) return fail(HAMSTR_ERR_ARG, "null argument");
  double* dst = nullptr;
  size_t n = 4 * (size_t)x->nCell;
  switch (field) {
    case HAMSTR_FIELD_Q: dst = x->q; break;
    case HAMSTR_FIELD_QT: dst = x->qt; break;
    case HAMSTR_FIELD_R: dst = x->r; break;
    case HAMSTR_FIELD_R0: dst = x->r0; break;
    case HAMSTR_FIELD_DQ: dst = x->dq; break;
    case HAMSTR_FIELD_SIGMA: dst = x->sigma; n = x->nCell; break;
    case HAMSTR_FIELD_FF: {
      const size_t tl = x->lhs32() ? sizeof(float) : sizeof(double);
      std::vector<char> buf(tl * 32 * (size_t)x->nProc);
      for (int p = 0; p < x->nProc; p++)
        for (int e = 0; e < 32; e++) {
          const double v = host[32 * (size_t)x->proc_face[p] + e];
          if (x->lhs32())
            reinterpret_cast<float*>(buf.data())[(size_t)e * x->nProc + p] = float(v);
          else
            reinterpret_cast<double*>(buf.data())[(size_t)e * x->nProc + p] = v;
        }
      HIPCHK(hipMemcpyAsync(x->ff, buf.data(), buf.size(), hipMemcpyHostToDevice, x->stream));
      HIPCHK(hipStreamSynchronize(x->stream));
      return HAMSTR_OK;
    }
    default: return fail(HAMSTR_ERR_ARG, "unknown field id");
  }
  HIPCHK(hipMemcpyAsync(dst, host, sizeof(double) * n, hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return HAMSTR_OK;
}

int hamstr_get_field(hamstr_ctx* x, int field, double* host) {
  if (!x || !host) return fail(HAMSTR_ERR_ARG, "null argument");
  const double* srcp = nullptr;
  size_t n = 4 * (size_t)x->nCell;
  switch (field) {
    case HAMSTR_FIELD_Q: srcp = x->q; break;
    case HAMSTR_FIELD_QT: srcp = x->qt; break;
    case HAMSTR_FIELD_R: srcp = x->r; break;
    case HAMSTR_FIELD_R0: srcp = x->r0; break;
    case HAMSTR_FIELD_DQ: srcp = x->dq; break;
    case HAMSTR_FIELD_SIGMA: srcp = x->sigma; n = x->nCell; break;
    case HAMSTR_FIELD_FF: {
      const size_t tl = x->lhs32() ? sizeof(float) : sizeof(double);
      std::vector<char> buf(tl * 32 * (size_t)x->nProc);
      HIPCHK(hipMemcpyAsync(buf.data(), x->ff, buf.size(), hipMemcpyDeviceToHost, x->stream));
      HIPCHK(hipStreamSynchronize(x->stream));
      std::memset(host, 0, sizeof(double) * 32 * (size_t)x->nFace);
      for (int p = 0; p < x->nProc; p++)
        for (int e = 0; e < 32; e++) {
          const size_t k = (size_t)e * x->nProc + p;
          host[32 * (size_t)x->proc_face[p] + e] =
              x->lhs32() ? double(reinterpret_cast<float*>(buf.data())[k]) : reinterpret_cast<double*>(buf.data())[k];
        }
      return HAMSTR_OK;
    }
    default: return fail(HAMSTR_ERR_ARG, "unknown field id");
  }
  HIPCHK(hipMemcpyAsync(host, srcp, sizeof(double) * n, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return HAMSTR_OK;
}

// cells [first, first+count) of a per-cell field (q, qt, r, r0, dq: 4 values
// per cell; sigma: 1)
static int cell_range(hamstr_ctx* x, int field, int first, int count, double** dev, size_t* per) {
  if (!x || first < 0 || count < 0 || (long)first + count > x->nCell) return fail(HAMSTR_ERR_ARG, "bad cell range");
  *per = 4;
  switch (field) {
    case HAMSTR_FIELD_Q: *dev = x->q; break;
    case HAMSTR_FIELD_QT: *dev = x->qt; break;
    case HAMSTR_FIELD_R: *dev = x->r; break;
    case HAMSTR_FIELD_R0: *dev = x->r0; break;
    case HAMSTR_FIELD_DQ: *dev = x->dq; break;
    case HAMSTR_FIELD_SIGMA: *dev = x->sigma; *per = 1; break;
    default: return fail(HAMSTR_ERR_ARG, "not a per-cell field");
  }
  *dev += *per * (size_t)first;
  return HAMSTR_OK;
}

int hamstr_get_cells(hamstr_ctx* x, int field, int first, int count, double* host) {
  double* d = nullptr;
  size_t per = 0;
  const int rc = cell_range(x, field, first, count, &d, &per);
  if (rc) return rc;
  if (!host && count) return fail(HAMSTR_ERR_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(host, d, sizeof(double) * per * count, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return HAMSTR_OK;
}

int hamstr_set_cells(hamstr_ctx* x, int field, int first, int count, const double* host) {
  double* d = nullptr;
  size_t per = 0;
  const int rc = cell_range(x, field, first, count, &d, &per);
  if (rc) return rc;
  if (!host && count) return fail(HAMSTR_ERR_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(d, host, sizeof(double) * per * count, hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return HAMSTR_OK;
}

// ---- solution files ---------------------------------------------------------
// Little-endian binary: a 64-byte header, one 32-byte record per case, then
// q[4*nCell] as doubles.
//   0 char[8] "HMSTRSOL"   8 int32 version (1)   12 int32 nCase   16 int64 nCell
//  24 int64 step          32 uint64 mesh hash   40 int32 order   44 int32 scheme
//  48 double CFL          56 8 bytes zero
//  per case: double Mach, double alpha, int64 first cell, int64 cell count
namespace {
constexpr char kSolMagic[8] = {'H', 'M', 'S', 'T', 'R', 'S', 'O', 'L'};
struct SolHeader {
  char magic[8];
  int32_t version, ncase;
  int64_t ncell, step;
  uint64_t mesh_hash;
  int32_t order, scheme;
  double cfl;
  int64_t reserved;
};
struct SolCase {
  double mach, alpha;
  int64_t first, count;
};
static_assert(sizeof(SolHeader) == 64 && sizeof(SolCase) == 32, "solution file layout");
}  // namespace

int hamstr_write_solution(hamstr_ctx* x, const char* path, long step) {
  if (!x || !path) return fail(HAMSTR_ERR_ARG, "null argument");
  std::vector<double> q(4 * (size_t)x->nCell);
  HIPCHK(hipMemcpyAsync(q.data(), x->q, sizeof(double) * q.size(), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  SolHeader h{};
  std::memcpy(h.magic, kSolMagic, 8);
  h.version = 1;
  h.ncase = x->nCase;
  h.ncell = x->nCell;
  h.step = step;
  h.mesh_hash = x->mesh_hash;
  h.order = x->prm.order;
  h.scheme = x->scheme;
  h.cfl = x->prm.cfl;
  std::vector<SolCase> cs(x->nCase);
  for (int k = 0; k < x->nCase; k++)
    cs[k] = SolCase{x->case_mach[k], x->case_alpha[k], x->case_off_h[k], x->case_off_h[k + 1] - x->case_off_h[k]};
  // written next to the target and renamed over it, so a reader never sees a partial file
  const std::string tmp = std::string(path) + ".tmp";
  FILE* fp = std::fopen(tmp.c_str(), "wb");
  if (!fp) return fail(HAMSTR_ERR_ARG, "cannot open " + tmp + " for writing");
  bool ok = std::fwrite(&h, sizeof(h), 1, fp) == 1 &&
            std::fwrite(cs.data(), sizeof(SolCase), cs.size(), fp) == cs.size() &&
            std::fwrite(q.data(), sizeof(double), q.size(), fp) == q.size();
  ok = (std::fclose(fp) == 0) && ok;
  if (!ok || std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    return fail(HAMSTR_ERR_ARG, std::string("writing ") + path + " failed");
  }
  return HAMSTR_OK;
}

int hamstr_read_solution(hamstr_ctx* x, const char* path, long* step) {
  if (!x || !path) return fail(HAMSTR_ERR_ARG, "null argument");
  FILE* fp = std::fopen(path, "rb");
  if (!fp) return fail(HAMSTR_ERR_ARG, std::string("cannot open ") + path);
  SolHeader h{};
  std::string err;
  std::vector<double> q(4 * (size_t)x->nCell);
  if (std::fread(&h, sizeof(h), 1, fp) != 1 || std::memcmp(h.magic, kSolMagic, 8) != 0 || h.version != 1)
    err = "not a solution file (version 1)";
  else if (h.ncase != x->nCase || h.ncell != x->nCell || h.mesh_hash != x->mesh_hash)
    err = "solution file is for another mesh or batch";
  else if (std::fseek(fp, (long)(sizeof(SolHeader) + sizeof(SolCase) * (size_t)h.ncase), SEEK_SET) != 0 ||
           std::fread(q.data(), sizeof(double), q.size(), fp) != q.size())
    err = "solution file is truncated";
  std::fclose(fp);
  if (!err.empty()) return fail(HAMSTR_ERR_ARG, std::string(path) + ": " + err);
  HIPCHK(hipMemcpyAsync(x->q, q.data(), sizeof(double) * q.size(), hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  if (step) *step = (long)h.step;
  return HAMSTR_OK;
}

int hamstr_compute_rhs(hamstr_ctx* x, int implicit, double* L2, double* Linf) {
  if (!x) return fail(HAMSTR_ERR_ARG, "null context");
  int rc = launch_rhs_faces(x, implicit != 0);
  if (rc) return rc;
  if ((rc = launch_rhs_cells(x, false, x->norms))) return rc;
  rc = read_norms(x);
  if (L2) *L2 = x->L2rhs;
  if (Linf) *Linf = x->Linfrhs;
  return rc;
}

int hamstr_adi(hamstr_ctx* x, double* L2lin, double* Linflin) {
  if (!x) return fail(HAMSTR_ERR_ARG, "null context");
  hipLaunchKernelGGL(k_adi_init, dim3(nblk(x->nCell)), dim3(kBlock), 0, x->stream, x->c, x->live(), x->r, x->r0, x->dq,
                     x->sigma);
  HIPCHK(hipGetLastError());
  int rc = launch_sweeps(x, x->norms);
  if (rc) return rc;
  rc = read_norms(x);
  if (L2lin) *L2lin = x->L2lin;
  if (Linflin) *Linflin = x->Linflin;
  return rc;
}

int hamstr_update_solution(hamstr_ctx* x, int src_field, int dst_field, double coef) {
  if (!x) return fail(HAMSTR_ERR_ARG, "null context");
  auto pick = [&](int f) -> double* {
    return f == HAMSTR_FIELD_Q ? x->q : (f == HAMSTR_FIELD_QT ? x->qt : nullptr);
  };
  const double* src = pick(src_field);
  double* dst = pick(dst_field);
  if (!src || !dst) return fail(HAMSTR_ERR_ARG, "updateSolution works on q and qt");
  hipLaunchKernelGGL(k_update_solution, dim3(nblk(x->nCell)), dim3(kBlock), 0, x->stream, x->nCell, x->live(), src,
                     dst, x->r, x->sigma, coef, x->c.cfl);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(x->stream));
  return HAMSTR_OK;
}

int hamstr_compute_linear_rhs(hamstr_ctx* x, double* L2lin, double* Linflin) {
  if (!x) return fail(HAMSTR_ERR_ARG, "null context");
  int rc = launch_linear(x, false, x->norms);
  if (rc) return rc;
  rc = read_norms(x);
  if (L2lin) *L2lin = x->L2lin;
  if (Linflin) *Linflin = x->Linflin;
  return rc;
}

}  // extern "C"

namespace {

// Stops case k from iterating: its faces, cells and chains drop out of every
// launch (a fill kernel on the context's stream, ordered with the iterations).
int retire_case(hamstr_ctx* x, int k, int status, long it) {
  x->case_status[k] = status;
  x->case_iters[k] = it;
  auto fill = [&](unsigned char* p, int a, int b) {
    if (b > a)
      hipLaunchKernelGGL(k_fill_u8, dim3(nblk(b - a)), dim3(kBlock), 0, x->stream, p + a, (long)(b - a),
                         (unsigned char)0);
  };
  fill(x->face_on, x->case_proc_off[k], x->case_proc_off[k + 1]);
  fill(x->face_on, x->case_orph_off[k], x->case_orph_off[k + 1]);
  fill(x->cell_on, x->case_off_h[k], x->case_off_h[k + 1]);
  fill(x->chain_on, x->case_chain_off[k], x->case_chain_off[k + 1]);
  HIPCHK(hipGetLastError());
  x->any_retired = true;
  return HAMSTR_OK;
}

void drop_graph(hamstr_ctx* x) {
  if (x->gexec) (void)hipGraphExecDestroy(x->gexec);
  x->gexec = nullptr;
  x->gK = 0;
}

int enqueue_adi_iteration(hamstr_ctx* x, double* nb, const std::function<int(const char*)>& mark, double* p_nl,
                          double* p_lin, bool fuse_in, double* p_nl_next, double* nl_norm_now);
bool can_fuse(const hamstr_ctx* x);

// stepSolution's euler and rk3 branches (march.cpp:73-108): per stage one
// face launch and one cell launch that also applies updateSolution.  The
// first stage's density-residual norms (of the state the step starts from)
// take the nonlinear slot, the last stage's (the norm marchInTime prints) the
// second slot: single case as block partials p_first / p_last, batches as the
// raw residual in rrho.
int enqueue_explicit(hamstr_ctx* x, const std::function<int(const char*)>& mark, double* p_first, double* p_last) {
  const bool one = x->nCase == 1;
  double* rr_first = one ? nullptr : x->rrho;
  double* rr_last = one ? nullptr : x->rrho + x->nCell;
  const int nstage = x->scheme == HAMSTR_SCHEME_RK3 ? 3 : 1;
  for (int s = 0; s < nstage; s++) {
    ExplicitStage st{};
    if (x->scheme == HAMSTR_SCHEME_EULER) {
      st = ExplicitStage{x->q, x->q, 1., nullptr, nullptr, 0., p_first, p_last, rr_first, rr_last};
    } else if (s == 0) {  // qt = q + 1/4 dt r, q = q + 8/15 dt r
      st = ExplicitStage{x->q, x->qt, 0.25, x->q, x->q, 8.0 / 15, p_first, nullptr, rr_first, nullptr};
    } else {  // q = qt + 5/12 dt r, then q = qt + 3/4 dt r
      const bool last = s == 2;
      st = ExplicitStage{x->qt, x->q, last ? 3.0 / 4.0 : 5.0 / 12, nullptr, nullptr, 0., last ? p_last : nullptr,
                         nullptr, last ? rr_last : nullptr, nullptr};
    }
    int rc;
    if ((rc = launch_rhs_faces(x, false))) return rc;
    if (mark && mark("rhs_faces")) return HAMSTR_ERR_HIP;
    hipLaunchKernelGGL(k_rhs_cells_explicit, dim3(nblk(x->nCell)), dim3(kBlock), 0, x->stream, x->c, x->live(),
                       x->cell_off, x->cell_ent, x->flux, x->sr, x->r, x->sigma, st);
    HIPCHK(hipGetLastError());
    if (mark && mark("rhs_cells_update")) return HAMSTR_ERR_HIP;
  }
  return HAMSTR_OK;
}

// one fused iteration (march.cpp:64-136, the scheme's branch).  Single case:
// the norms' block partials go to iteration slot `slot` of x->npart (reduced
// per run of iterations by enqueue_block); batches: per-case norms of both
// stages to nb[nCase][4] at the end of the iteration.
// fuse_in / fuse_out: the iteration's computeRHS is fused into the previous
// iteration's tail / its tail into the next iteration's computeRHS (can_fuse;
// nb_next: the next iteration's norms).  A batch's per-case norms are reduced
// from rrho, which the fused tail overwrites with the next iteration's
// residuals: each stage's norms are taken as soon as it is complete.
int enqueue_iteration(hamstr_ctx* x, double* nb, const std::function<int(const char*)>& mark, int slot,
                      bool fuse_in = false, bool fuse_out = false, double* nb_next = nullptr) {
  int rc;
  const bool one = x->nCase == 1;
  double* p_nl = one ? x->npart + (size_t)(2 * slot) * x->nPart * 2 : x->partial;
  double* p_lin = one ? x->npart + (size_t)(2 * slot + 1) * x->nPart * 2 : x->partial;
  double* p_next = fuse_out ? (one ? x->npart + (size_t)(2 * (slot + 1)) * x->nPart * 2 : x->partial) : nullptr;
  double* nl_now = (!one && fuse_out && !fuse_in) ? nb : nullptr;
  if (x->scheme != HAMSTR_SCHEME_ADI) {
    if ((rc = enqueue_explicit(x, mark, one ? p_nl : nullptr, one ? p_lin : nullptr))) return rc;
  } else if ((rc = enqueue_adi_iteration(x, nb, mark, p_nl, p_lin, fuse_in, p_next, nl_now))) {
    return rc;
  }
  if (!one) {
    if (fuse_out) {
      if ((rc = launch_norm(x, nb, 2)) || (rc = launch_norm(x, nb_next, 0))) return rc;
    } else if (fuse_in) {
      if ((rc = launch_norm(x, nb, 2))) return rc;
    } else {
      hipLaunchKernelGGL(k_case_norms, dim3(x->nCase, 2), dim3(kRedBlock), 0, x->stream, x->rrho, x->nCell,
                         x->case_off, 0, nb, x->live(), x->keep_norms);
      HIPCHK(hipGetLastError());
    }
    if (mark && mark("norms")) return HAMSTR_ERR_HIP;
  }
  return HAMSTR_OK;
}

// stepSolution's ADI branch (march.cpp:112-117) without the batch norms
// computeRHS("implicit") and the ADI phases of the first sweep, the chain
// faces' Jacobians taken from source rj (HAMSTR_JAC_*): the part of the
// fused iteration the Jacobian source changes (the autotuner times it)
// rhs_done: the fused tail of the previous iteration (launch_lin_rhs) has
// done computeRHS already, except the residual sums phase 0's launch carries
// nl_norm_now (a batch whose residual norm is read before the fused tail
// overwrites it): the per-case norms of this computeRHS go there at once
int enqueue_rhs_sweep0(hamstr_ctx* x, double* nb, const std::function<int(const char*)>& mark, double* p_nl,
                       int rj, bool rhs_done = false, double* nl_norm_now = nullptr) {
  int rc;
  // FAST: the ADI kernels may compute the chain faces' Jacobians again, from
  // the face states k_rhs_faces stores instead (64 B per face, FP64 256 B)
  // or from q; k_rhs_faces then stores the orphan faces' Jacobians only
  if (!rhs_done) {
    if ((rc = launch_rhs_faces(x, true, rj != HAMSTR_JAC_STORED ? x->nChainProc : 0,
                               jac_states(rj) ? x->fst : nullptr,
                               rj == HAMSTR_JAC_STATES_LINES ? x->jac_to : 0)))
      return rc;
    if (mark && mark("rhs_faces")) return HAMSTR_ERR_HIP;
  }
  // phase 0 over the leading cells on part of the device: the other cells'
  // residual sums ride in phase 0's launch, on the CUs its chains leave idle
  // (phase 0's rows read only their own cells' r and sigma; the block
  // partials are those of one k_rhs_cells launch)
  const int extra = p_nl ? cells_overlap_extra(x) : 0;
  CellWork cw{};
  if (extra > 0) {
    if (!rhs_done && (rc = launch_rhs_cells_range(x, true, p_nl, 0, x->split_cells))) return rc;
    cw = CellWork{x->cell_off, x->cell_ent, x->flux, x->sr, x->r0, p_nl, x->split_cells, x->nCell, 0};
  } else if (!rhs_done && (rc = launch_rhs_cells(x, true, nb, p_nl))) {
    return rc;
  }
  if (!rhs_done && mark && mark("rhs_cells")) return HAMSTR_ERR_HIP;
  if (!rhs_done && nl_norm_now && (rc = launch_norm(x, nl_norm_now, 0))) return rc;
  return launch_sweep_phases(x, 0, mark, rj, &cw, extra);
}

// Iterations of a captured block run fused across their boundary (launch_lin_rhs)
// for a single live case with one sweep: every value the same, bit for bit.
bool can_fuse(const hamstr_ctx* x) {
  if (x->prm.msweep != 1 || x->scheme != HAMSTR_SCHEME_ADI) return false;
  // a single case reduces block partials per iteration slot, a batch its
  // per-cell density residuals per case
  return x->nCase == 1 ? (x->npart != nullptr && x->rrho == nullptr) : x->rrho != nullptr;
}

// computeLinearRHS of this iteration with the update, fused with computeRHS
// ("implicit") of the next: block partials of this iteration's linear norm to
// p_lin, of the next one's residual norm to p_nl_next (for the cells phase 0's
// launch does not sum)
int launch_lin_rhs(hamstr_ctx* x, int rj, double* p_lin, double* p_nl_next,
                   const std::function<int(const char*)>& mark) {
  const FaceTab ft = x->ftab();
  const Live lv = x->live();
  const int ff_from = rj != HAMSTR_JAC_STORED ? x->nChainProc : 0;
  double* fst = jac_states(rj) ? x->fst : nullptr;
  const int jac_to = rj == HAMSTR_JAC_STATES_LINES ? x->jac_to : 0;
  const dim3 g(nblk(x->nProc)), b(kBlock);
#define LRF(O, T, F)                                                                                         \
  if (x->stage_faces)                                                                                          \
    hipLaunchKernelGGL((k_lin_rhs_faces<O, T, F, true>), g, b, 0, x->stream, ft, x->c, lv, x->q, x->dq,        \
                       x->lflux, x->flux, x->sr, (T*)x->ff, ff_from, fst, x->nanflag, jac_to);                 \
  else                                                                                                         \
    hipLaunchKernelGGL((k_lin_rhs_faces<O, T, F, false>), g, b, 0, x->stream, ft, x->c, lv, x->q, x->dq,       \
                       x->lflux, x->flux, x->sr, (T*)x->ff, ff_from, fst, x->nanflag, jac_to)
#define LRO(T, F) \
  if (o == 1) { LRF(1, T, F); } else if (o == 3) { LRF(3, T, F); } else { LRF(5, T, F); }
  const int o = x->prm.order;
  const bool fj = x->prm.adi_mode == HAMSTR_ADI_FAST;
  if (x->lhs32()) {
    if (fj) { LRO(float, true); } else { LRO(float, false); }
  } else {
    if (fj) { LRO(double, true); } else { LRO(double, false); }
  }
#undef LRO
#undef LRF
  HIPCHK(hipGetLastError());
  if (mark && mark("lin_rhs_faces")) return HAMSTR_ERR_HIP;
  // the blocks of k_rhs_cells' cells come first (a multiple of 8 of them,
  // else all of them)
  int nb_rhs = cells_overlap_extra(x) > 0 ? x->split_cells / kBlock : nblk(x->nCell);
  if (nb_rhs % kXcd != 0) nb_rhs = nblk(x->nCell);
  hipLaunchKernelGGL(k_lin_rhs_cells, dim3(nblk(x->nCell)), dim3(kBlock), 0, x->stream, x->c, lv, x->cell_off,
                     x->cell_ent, x->lflux, x->flux, x->sr, x->r0, x->dq, x->sigma, x->r, x->q, p_lin, p_nl_next,
                     x->rrho, nb_rhs);
  HIPCHK(hipGetLastError());
  if (mark && mark("lin_rhs_cells")) return HAMSTR_ERR_HIP;
  return HAMSTR_OK;
}

// fuse_in: this iteration's computeRHS was fused into the previous one's tail;
// p_nl_next non-null: fuse this iteration's tail with the next's computeRHS
// (can_fuse contexts only); nl_norm_now: as in enqueue_rhs_sweep0
int enqueue_adi_iteration(hamstr_ctx* x, double* nb, const std::function<int(const char*)>& mark, double* p_nl,
                          double* p_lin, bool fuse_in, double* p_nl_next, double* nl_norm_now) {
  int rc;
  const int rj = jac_source(x);
  for (int sw = 0; sw < x->prm.msweep; sw++) {
    if (sw == 0) {
      if ((rc = enqueue_rhs_sweep0(x, nb, mark, p_nl, rj, fuse_in, nl_norm_now))) return rc;
    } else if ((rc = launch_sweep_phases(x, sw, mark, rj))) {
      return rc;
    }
    if (p_nl_next && sw == x->prm.msweep - 1) {
      if ((rc = launch_lin_rhs(x, rj, p_lin, p_nl_next, mark))) return rc;
      continue;
    }
    // the Jacobians computed again beat reading them back (FP64: 26.9 ->
    // 24.1 us on the 512x256 mixed mesh; FP32: the same there, 1.14 -> 0.84
    // ms in a 256-case batch)
    if ((rc = launch_lin_faces(x, true))) return rc;
    if (mark && mark("lin_faces")) return HAMSTR_ERR_HIP;
    if ((rc = launch_lin_cells(x, sw == x->prm.msweep - 1, nb, p_lin))) return rc;
    if (mark && mark("lin_cells")) return HAMSTR_ERR_HIP;
  }
  return HAMSTR_OK;
}

// `count` iterations whose norms go to base[count][nCase][4]
int enqueue_block(hamstr_ctx* x, int count, double* base, const std::function<int(const char*)>& mark) {
  const int nc = x->nCase;
  const bool fuse = count > 1 && can_fuse(x);
  for (int s = 0; s < count; s++) {
    const int rc = enqueue_iteration(x, base + 4 * (size_t)nc * s, mark, s, fuse && s > 0, fuse && s + 1 < count,
                                     base + 4 * (size_t)nc * (s + 1));
    if (rc) return rc;
  }
  if (nc == 1) {
    hipLaunchKernelGGL(k_norm_slots, dim3(2 * count), dim3(kRedBlock), 0, x->stream, x->npart, x->nPart, x->nCell,
                       base);
    HIPCHK(hipGetLastError());
    if (mark && mark("norms")) return HAMSTR_ERR_HIP;
  }
  return HAMSTR_OK;
}

// iterations per captured graph for a read-back interval: the largest divisor
// of check_every up to kGraphMaxIters (0: no graph)
int graph_block(int check_every) {
  for (int d = std::min(check_every, kGraphMaxIters); d >= 2; d--)
    if (check_every % d == 0) return d;
  return 0;
}

// Captures K iterations into a graph whose norms go to ghist[K][nCase][4].
int build_graph(hamstr_ctx* x, int K) {
  drop_graph(x);
  const int nc = x->nCase;
  if (x->ghistCap < K) {
    if (x->ghist) (void)hipFree(x->ghist);
    x->ghist = nullptr;
    x->ghistCap = 0;
    HIPCHK(hipMalloc(&x->ghist, sizeof(double) * 4 * nc * K));
    x->ghistCap = K;
  }
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(x->stream, hipStreamCaptureModeThreadLocal));
  int rc = HAMSTR_OK;
  rc = enqueue_block(x, K, x->ghist, nullptr);
  const hipError_t e = hipStreamEndCapture(x->stream, &g);
  if (rc) {
    if (e == hipSuccess && g) (void)hipGraphDestroy(g);
    return rc;
  }
  HIPCHK(e);
  const hipError_t ei = hipGraphInstantiate(&x->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    x->gexec = nullptr;
    return fail(HAMSTR_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  }
  x->gK = K;
  // the executable graph goes to the device now, not on its first replay
  HIPCHK(hipGraphUpload(x->gexec, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return HAMSTR_OK;
}

// Runs up to nsteps iterations, reading the norm history back every
// check_every steps.  iters[k] gets the first step (1-based) at which case k's
// nonlinear L2 <= tol (-1 if never); the loop stops at the first read-back
// after every case has converged (tol > 0).  *done = steps executed.
int run_steps(hamstr_ctx* x, int nsteps, double tol, int check_every, std::vector<int>& iters,
              std::vector<double>& series, int* done) {
  const int nc = x->nCase;
  // the outputs are sized before any early return, so callers may index them
  iters.assign(nc, -1);
  if (done) *done = 0;
  if (nsteps < 0) return fail(HAMSTR_ERR_ARG, "bad arguments");
  series.assign(4 * (size_t)nc * nsteps, 0.0);
  if (check_every < 1) check_every = 1;
  if (x->histCap < nsteps) {
    if (x->hist) (void)hipFree(x->hist);
    x->hist = nullptr;
    x->histCap = 0;  // a failed allocation leaves no capacity behind
    HIPCHK(hipMalloc(&x->hist, sizeof(double) * 4 * nc * std::max(nsteps, 1)));
    x->histCap = nsteps;
  }
  int rc = HAMSTR_OK;
  std::vector<hipEvent_t> evs;
  int nstage = 0;
  if (x->timing) x->stage_names.clear();
  int cur = 0;
  auto mark = [&](const char* name) -> int {
    if (cur == 0) x->stage_names.push_back(name);
    const size_t need = evs.size() + 1;
    if (x->ev.size() < need) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return 1;
      x->ev.push_back(e);
    }
    hipEvent_t e = x->ev[need - 1];
    evs.push_back(e);
    return hipEventRecord(e, x->stream) != hipSuccess;
  };
  std::function<int(const char*)> mk;
  if (x->timing) mk = mark;
  // blocks of K iterations between read-backs replay a captured graph
  int K = 0;
  if (x->use_graph && !x->timing && nsteps >= 2) {
    K = graph_block(check_every);
    if (nsteps < K) K = 0;
    if (K && (!x->gexec || x->gK != K) && (rc = build_graph(x, K))) return rc;
  }
  int ran = 0, lastRead = 0;
  for (int n = 0; n < nsteps; n = ran) {
    cur = n;
    if (K && nsteps - n >= K && (n % check_every) + K <= check_every) {
      HIPCHK(hipGraphLaunch(x->gexec, x->stream));
      HIPCHK(hipMemcpyAsync(x->hist + 4 * (size_t)nc * n, x->ghist, sizeof(double) * 4 * nc * K,
                            hipMemcpyDeviceToDevice, x->stream));
      ran = n + K;
    } else {
      if (x->timing && mark("start")) return fail(HAMSTR_ERR_HIP, "hipEventRecord");
      if ((rc = enqueue_block(x, 1, x->hist + 4 * (size_t)nc * n, mk))) return rc;
      if (n == 0) nstage = (int)evs.size() - 1;
      ran = n + 1;
    }
    if (ran % check_every == 0 || ran == nsteps) {
      const size_t cnt = 4 * (size_t)nc * (ran - lastRead);
      HIPCHK(hipMemcpyAsync(series.data() + 4 * (size_t)nc * lastRead, x->hist + 4 * (size_t)nc * lastRead,
                            sizeof(double) * cnt, hipMemcpyDeviceToHost, x->stream));
      int flag = 0;
      HIPCHK(hipMemcpyAsync(&flag, x->nanflag, sizeof(int), hipMemcpyDeviceToHost, x->stream));
      HIPCHK(hipStreamSynchronize(x->stream));
      const double* last = series.data() + 4 * (size_t)nc * (ran - 1);
      x->case_norm.assign(last, last + 4 * nc);
      x->L2rhs = last[0];
      x->Linfrhs = last[1];
      x->L2lin = last[2];
      x->Linflin = last[3];
      if (flag && x->retire) {
        // with retirement on the run keeps going: cases whose norms are no
        // longer finite retire as diverged (their cells are disjoint from every
        // other case's)
        int found = 0;
        for (int j = 0; j < nc; j++) {
          if (x->case_status[j] != 0) continue;
          bool bad = false;
          for (int k = lastRead; k < ran && !bad; k++)
            for (int e = 0; e < 4; e++) bad = bad || !std::isfinite(series[4 * ((size_t)nc * k + j) + e]);
          if (bad) {
            if ((rc = retire_case(x, j, 3, x->it_total + ran))) return rc;
            found++;
          }
        }
        if (found) {
          HIPCHK(hipMemsetAsync(x->nanflag, 0, sizeof(int), x->stream));  // eager, between graph replays
          flag = 0;
        }
      }
      if (flag) {
        if (done) *done = ran;
        return fail(HAMSTR_ERR_NAN, "NaN spectral radius in the Roe flux");
      }
      bool all = tol > 0;
      for (int j = 0; j < nc; j++) {
        for (int k = lastRead; k < ran && iters[j] < 0; k++)
          if (series[4 * ((size_t)nc * k + j)] <= tol) iters[j] = k + 1;
        if (iters[j] < 0) all = false;
      }
      if (x->retire) {
        // retire cases that converged; a case whose norm has not halved within
        // stall_window iterations is stalled (a limit cycle) and retires too
        all = true;
        for (int j = 0; j < nc; j++) {
          if (x->case_status[j] != 0) continue;
          if (tol > 0 && iters[j] > 0) {
            if ((rc = retire_case(x, j, 1, x->it_total + iters[j]))) return rc;
            continue;
          }
          for (int k = lastRead; k < ran; k++) {
            const double v = series[4 * ((size_t)nc * k + j)];
            if (v < 0.5 * x->stall_ref[j]) {
              x->stall_ref[j] = v;
              x->stall_ref_it[j] = x->it_total + k + 1;
            }
          }
          if (x->stall_window > 0 && x->it_total + ran - x->stall_ref_it[j] > x->stall_window) {
            if ((rc = retire_case(x, j, 2, x->it_total + ran))) return rc;
            continue;
          }
          all = false;
        }
      }
      lastRead = ran;
      if (all) break;
    }
  }
  if (x->timing && nstage > 0) {
    HIPCHK(hipStreamSynchronize(x->stream));
    x->last_ms.assign(nstage, 0.0);
    const int per = nstage + 1;
    const int steps = (int)evs.size() / per;
    for (int k = 0; k < steps; k++)
      for (int st = 0; st < nstage; st++) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, evs[k * per + st], evs[k * per + st + 1]));
        x->last_ms[st] += ms / steps;
      }
    x->stage_names.erase(x->stage_names.begin());  // "start"
  }
  x->it_total += ran;
  if (done) *done = ran;
  return rc;
}

}  // namespace

extern "C" {

int hamstr_step(hamstr_ctx* x, int nsteps, double tol, int check_every, double* l2rhs, double* l2lin, int* done) {
  if (!x) return fail(HAMSTR_ERR_ARG, "null context");
  std::vector<int> iters;
  std::vector<double> series;
  int ran = 0;
  const int rc = run_steps(x, nsteps, tol, check_every, iters, series, &ran);
  for (int k = 0; k < ran; k++) {
    if (l2rhs) l2rhs[k] = series[4 * (size_t)x->nCase * k];
    if (l2lin) l2lin[k] = series[4 * (size_t)x->nCase * k + 2];
  }
  if (done) *done = ran;
  return rc;
}

int hamstr_batch_step(hamstr_ctx* x, int nsteps, double tol, int check_every, int* iters, double* norms,
                      int* done) {
  if (!x) return fail(HAMSTR_ERR_ARG, "null context");
  std::vector<int> it;
  std::vector<double> series;
  int ran = 0;
  const int rc = run_steps(x, nsteps, tol, check_every, it, series, &ran);
  if (iters && (int)it.size() == x->nCase)
    for (int k = 0; k < x->nCase; k++) iters[k] = it[k];
  if (norms && ran > 0) std::memcpy(norms, x->case_norm.data(), sizeof(double) * 4 * x->nCase);
  if (done) *done = ran;
  return rc;
}

int hamstr_set_launch(hamstr_ctx* x, int group, int grid, int cr_levels) {
  if (!x || group < 0 || group >= (int)x->groups.size()) return fail(HAMSTR_ERR_ARG, "bad launch group");
  Group& g = x->groups[group];
  if (!g.fast) return fail(HAMSTR_ERR_ARG, "launch shape of a non-FAST group");
  if (grid < 1 || cr_levels < 0 || cr_levels > kMaxCR) return fail(HAMSTR_ERR_ARG, "bad grid or cyclic-reduction levels");
  g.grid = grid;
  g.ncr = cr_levels;
  drop_graph(x);
  return HAMSTR_OK;
}

int hamstr_set_thomas(hamstr_ctx* x, int group, int pcr_levels) {
  if (!x || group < 0 || group >= (int)x->groups.size()) return fail(HAMSTR_ERR_ARG, "bad launch group");
  Group& g = x->groups[group];
  if (!g.fast || g.two || g.final != kFinalOpen)
    return fail(HAMSTR_ERR_ARG, "a Thomas finish needs a FAST group of open chains that fit a workgroup");
  if (pcr_levels < -1 || pcr_levels > kMaxThomasPcr) return fail(HAMSTR_ERR_ARG, "bad PCR levels");
  g.tpcr = pcr_levels;
  drop_graph(x);
  return HAMSTR_OK;
}

int hamstr_thomas(const hamstr_ctx* x, int group, int* pcr_levels) {
  if (!x || !pcr_levels || group < 0 || group >= (int)x->groups.size()) return fail(HAMSTR_ERR_ARG, "bad launch group");
  *pcr_levels = x->groups[group].tpcr;
  return HAMSTR_OK;
}

// the face-state array of HAMSTR_JAC_FROM_STATES, allocated on first use
int ensure_face_states(hamstr_ctx* x) {
  if (!x->fst) HIPCHK(hipMalloc(&x->fst, sizeof(double) * 8 * (size_t)x->nProc));
  return HAMSTR_OK;
}

int hamstr_set_jacobian_source(hamstr_ctx* x, int src) {
  if (!x) return fail(HAMSTR_ERR_ARG, "null context");
  if (src != HAMSTR_JAC_STORED && src != HAMSTR_JAC_FROM_Q && src != HAMSTR_JAC_FROM_STATES &&
      src != HAMSTR_JAC_STATES_LINES)
    return fail(HAMSTR_ERR_ARG, "unknown Jacobian source");
  if (src == HAMSTR_JAC_STATES_LINES && x->jac_to <= 0)
    return fail(HAMSTR_ERR_ARG, "phase 0's chains do not come first, or there is one phase only");
  if (src != HAMSTR_JAC_STORED && !jac_recompute_possible(x))
    return fail(HAMSTR_ERR_ARG, "the Jacobians can be computed again only by FAST kernels other than the "
                                "long-chain one, and not for FP32 odd cycles");
  if (jac_states(src)) {
    const int rc = ensure_face_states(x);
    if (rc) return rc;
  }
  x->jac_src = src;
  drop_graph(x);
  return HAMSTR_OK;
}

int hamstr_jacobian_source(const hamstr_ctx* x, int* src) {
  if (!x || !src) return fail(HAMSTR_ERR_ARG, "null argument");
  *src = jac_source(x);
  return HAMSTR_OK;
}

int hamstr_set_face_staging(hamstr_ctx* x, int on) {
  if (!x) return fail(HAMSTR_ERR_ARG, "null context");
  if (on && !x->stage_fits)
    return fail(HAMSTR_ERR_ARG, "a workgroup's stencil slots exceed the staging capacity on this mesh");
  x->stage_faces = on != 0;
  drop_graph(x);
  return HAMSTR_OK;
}

int hamstr_face_staging(const hamstr_ctx* x, int* on) {
  if (!x || !on) return fail(HAMSTR_ERR_ARG, "null argument");
  *on = x->stage_faces ? 1 : 0;
  return HAMSTR_OK;
}

int hamstr_set_adi_kernel(hamstr_ctx* x, int group, int kernel) {
  if (!x || group < 0 || group >= (int)x->groups.size()) return fail(HAMSTR_ERR_ARG, "bad launch group");
  Group& g = x->groups[group];
  if (!g.fast || g.two) return fail(HAMSTR_ERR_ARG, "the group has one ADI kernel only");
  if (kernel != kKernRow && kernel != kKernHalf && kernel != kKernHalf2) return fail(HAMSTR_ERR_ARG, "unknown ADI kernel");
  if ((kernel == kKernHalf && !g.half_ok) || (kernel == kKernHalf2 && !g.half2_ok))
    return fail(HAMSTR_ERR_ARG, "the group's chains do not fit that half-LDS kernel (FP64 only for one chain per "
                                "workgroup, open chains only for two)");
  g.kern = kernel;
  drop_graph(x);
  return HAMSTR_OK;
}

int hamstr_prepare_graph(hamstr_ctx* x, int check_every) {
  if (!x || check_every < 1) return fail(HAMSTR_ERR_ARG, "bad arguments");
  const int K = graph_block(check_every);
  if (!x->use_graph || K == 0 || (x->gexec && x->gK == K)) return HAMSTR_OK;
  const int rc = build_graph(x, K);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(x->stream));
  return HAMSTR_OK;
}

int hamstr_set_scheme(hamstr_ctx* x, int scheme) {
  if (!x) return fail(HAMSTR_ERR_ARG, "null context");
  if (scheme != HAMSTR_SCHEME_ADI && scheme != HAMSTR_SCHEME_EULER && scheme != HAMSTR_SCHEME_RK3)
    return fail(HAMSTR_ERR_ARG, "unknown scheme");
  if (scheme != x->scheme) drop_graph(x);  // the captured iteration is the scheme's
  x->scheme = scheme;
  return HAMSTR_OK;
}

int hamstr_get_scheme(const hamstr_ctx* x) { return x ? x->scheme : -1; }

int hamstr_set_graph(hamstr_ctx* x, int on) {
  if (!x) return fail(HAMSTR_ERR_ARG, "null context");
  x->use_graph = on != 0;
  drop_graph(x);
  return HAMSTR_OK;
}

int hamstr_set_retire(hamstr_ctx* x, int on, int stall_window) {
  if (!x || stall_window < 0) return fail(HAMSTR_ERR_ARG, "bad arguments");
  drop_graph(x);  // the live masks enter the launches only while retirement is on
  x->retire = on != 0;
  x->stall_window = stall_window;
  x->it_total = 0;
  x->case_status.assign(x->nCase, 0);
  x->case_iters.assign(x->nCase, -1);
  x->stall_ref.assign(x->nCase, HUGE_VAL);
  x->stall_ref_it.assign(x->nCase, 0);
  if (x->any_retired) {
    auto fill = [&](unsigned char* p, long n) {
      hipLaunchKernelGGL(k_fill_u8, dim3(nblk(n)), dim3(kBlock), 0, x->stream, p, n, (unsigned char)1);
    };
    fill(x->face_on, x->nProc);
    fill(x->cell_on, x->nCell);
    fill(x->chain_on, x->nChain);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(x->stream));
    x->any_retired = false;
  }
  return HAMSTR_OK;
}

int hamstr_set_case_live(hamstr_ctx* x, int k, int live) {
  if (!x || k < 0 || k >= x->nCase) return fail(HAMSTR_ERR_ARG, "bad arguments");
  if (!x->retire) return fail(HAMSTR_ERR_ARG, "hamstr_set_case_live needs retirement on (hamstr_set_retire)");
  if (!live) {
    if (x->case_status[k] == 0) return retire_case(x, k, 4, x->it_total);
    return HAMSTR_OK;
  }
  auto fill = [&](unsigned char* p, int a, int b) {
    if (b > a)
      hipLaunchKernelGGL(k_fill_u8, dim3(nblk(b - a)), dim3(kBlock), 0, x->stream, p + a, (long)(b - a),
                         (unsigned char)1);
  };
  fill(x->face_on, x->case_proc_off[k], x->case_proc_off[k + 1]);
  fill(x->face_on, x->case_orph_off[k], x->case_orph_off[k + 1]);
  fill(x->cell_on, x->case_off_h[k], x->case_off_h[k + 1]);
  fill(x->chain_on, x->case_chain_off[k], x->case_chain_off[k + 1]);
  HIPCHK(hipGetLastError());
  x->case_status[k] = 0;
  x->case_iters[k] = -1;
  x->stall_ref[k] = HUGE_VAL;
  x->stall_ref_it[k] = x->it_total;
  return HAMSTR_OK;
}

long hamstr_iterations(const hamstr_ctx* x) { return x ? x->it_total : -1; }

int hamstr_case_status(const hamstr_ctx* x, int* status, long* iters) {
  if (!x) return fail(HAMSTR_ERR_ARG, "null context");
  for (int k = 0; k < x->nCase; k++) {
    if (status) status[k] = x->case_status[k];
    if (iters) iters[k] = x->case_iters[k];
  }
  return HAMSTR_OK;
}

int hamstr_last_norms(const hamstr_ctx* x, double* out4) {
  if (!x || !out4) return fail(HAMSTR_ERR_ARG, "null argument");
  out4[0] = x->L2rhs;
  out4[1] = x->Linfrhs;
  out4[2] = x->L2lin;
  out4[3] = x->Linflin;
  return HAMSTR_OK;
}

// Times each FAST launch group's kernels at several persistent grid sizes (all
// chains, 8, 4, 2, 1 workgroups per CU) and cyclic-reduction depths, and keeps
// the fastest.  Per kernel the search is by coordinates: every grid at the
// group's current depth, then every other depth at the best grid (8 shapes
// instead of 20).  r and dq are restored.
int hamstr_autotune(hamstr_ctx* x, int reps, double* best_ms, int cap, int* ngroups) {
  if (!x) return fail(HAMSTR_ERR_ARG, "null context");
  if (reps < 1) reps = 3;
  drop_graph(x);  // launch shapes change
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, x->device));
  const int cu = std::max(prop.multiProcessorCount, 1);
  const size_t c4 = sizeof(double) * 4 * x->nCell;
  double *r_save = nullptr, *dq_save = nullptr, *r0_save = nullptr, *q_save = nullptr, *sg_save = nullptr;
  int* nan_save = nullptr;
  // r, r0, dq, q, sigma and the NaN flag copied out: restore them (the
  // trials of the fused iteration step the state, and a case that blows up
  // within them must not leave the flag set)
  bool saved = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // the trials; a failure stops them, and r, dq and the buffers are still
  // restored and released below
  auto search = [&]() -> int {
    HIPCHK(hipMalloc(&r_save, c4));
    HIPCHK(hipMalloc(&dq_save, c4));
    HIPCHK(hipMalloc(&r0_save, c4));
    HIPCHK(hipMalloc(&q_save, c4));
    HIPCHK(hipMalloc(&sg_save, c4 / 4));
    HIPCHK(hipMalloc(&nan_save, sizeof(int)));
    HIPCHK(hipMemcpyAsync(r_save, x->r, c4, hipMemcpyDeviceToDevice, x->stream));
    HIPCHK(hipMemcpyAsync(dq_save, x->dq, c4, hipMemcpyDeviceToDevice, x->stream));
    HIPCHK(hipMemcpyAsync(r0_save, x->r0, c4, hipMemcpyDeviceToDevice, x->stream));
    HIPCHK(hipMemcpyAsync(q_save, x->q, c4, hipMemcpyDeviceToDevice, x->stream));
    HIPCHK(hipMemcpyAsync(sg_save, x->sigma, c4 / 4, hipMemcpyDeviceToDevice, x->stream));
    HIPCHK(hipMemcpyAsync(nan_save, x->nanflag, sizeof(int), hipMemcpyDeviceToDevice, x->stream));
    saved = true;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    // events around `reps` launches of `launch` after one warm-up: ms each
    auto timed = [&](const std::function<int()>& launch, double* ms_out) -> int {
      int rc = launch();
      if (rc) return rc;
      HIPCHK(hipEventRecord(e0, x->stream));
      for (int i = 0; i < reps; i++)
        if ((rc = launch())) return rc;
      HIPCHK(hipEventRecord(e1, x->stream));
      HIPCHK(hipEventSynchronize(e1));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      *ms_out = ms / reps;
      return HAMSTR_OK;
    };
    // the group search with the Jacobians read from ff (rj false) or
    // computed again from q (rj true): gs gets every group's best shape,
    // per[k] its time, *sum their total
    auto tune_groups = [&](int rj, std::vector<Group>& gs, std::vector<double>& per, double* sum) -> int {
      per.assign(gs.size(), -1.0);
      *sum = 0;
      for (size_t k = 0; k < gs.size(); k++) {
        Group& g = gs[k];
        if (!g.fast) continue;
        double best = -1;
        std::vector<int> cand = {g.n};
        for (int w : {8, 4, 2, 1})
          if (w * cu < g.n) cand.push_back(w * cu);
        std::vector<int> kerns = {kKernRow};
        if (g.half_ok) kerns.push_back(kKernHalf);
        if (g.half2_ok) kerns.push_back(kKernHalf2);
        int best_kern = g.kern, best_grid = g.grid, best_ncr = g.ncr;
        const int ncr0 = std::min(std::max(g.ncr, 1), kMaxCR);
        for (int kern : kerns) {
          g.kern = kern;
          double kbest = -1;
          int kgrid = cand[0];
          auto trial = [&](int grid, int ncr) -> int {
            double t = 0;
            const int rc = timed([&]() { return launch_group(x, g, grid, 0, ncr, rj); }, &t);
            if (rc) return rc;
            if (kbest < 0 || t < kbest) {
              kbest = t;
              kgrid = grid;
            }
            if (best < 0 || t < best) {
              best = t;
              best_grid = grid;
              best_ncr = ncr;
              best_kern = kern;
            }
            return HAMSTR_OK;
          };
          int rc;
          for (int grid : cand)
            if ((rc = trial(grid, ncr0))) return rc;
          // the half-LDS kernels always run the first reduction level
          const int grid = kgrid;
          for (int ncr = kern == kKernRow ? 0 : 1; ncr <= kMaxCR; ncr++)
            if (ncr != ncr0 && (rc = trial(grid, ncr))) return rc;
        }
        g.kern = best_kern;
        g.grid = best_grid;
        g.ncr = best_ncr;
        if (g.final == kFinalOpen && !g.two) {
          // open chains: the PCR levels before the Thomas finish, at the best
          // kernel, grid and depth, then the depths again at the best level
          const int t0 = g.tpcr;
          int best_t = t0;
          auto ttrial = [&](int tp, int ncr) -> int {
            double t = 0;
            g.tpcr = tp;
            const int rc = timed([&]() { return launch_group(x, g, best_grid, 0, ncr, rj); }, &t);
            if (rc) return rc;
            if (t < best) {
              best = t;
              best_t = tp;
              best_ncr = ncr;
            }
            return HAMSTR_OK;
          };
          int rc;
          for (int tp : {-1, 2, 3, 4})
            if (tp != t0 && tp <= kMaxThomasPcr && (rc = ttrial(tp, best_ncr))) return rc;
          const int n1 = best_ncr;
          for (int ncr = g.kern == kKernRow ? 0 : 1; ncr <= kMaxCR; ncr++)
            if (ncr != n1 && (rc = ttrial(best_t, ncr))) return rc;
          g.tpcr = best_t;
          g.ncr = best_ncr;
        }
        per[k] = best;
        *sum += best;
      }
      return HAMSTR_OK;
    };
    int rc;
    std::vector<Group> gs = x->groups;
    std::vector<double> per;
    double sum = 0;
    int rj = HAMSTR_JAC_STORED;
    if ((rc = tune_groups(HAMSTR_JAC_STORED, gs, per, &sum))) return rc;
    if (jac_recompute_possible(x)) {
      // the fused iteration stores every face's Jacobians and the ADI kernels
      // read them, or it stores the orphan faces' only and the ADI kernels
      // compute the chain faces' again, from stored face states or from q
      // (or phase 0 reads stored ones and the other phases use states).
      // Each source's groups are tuned for it; then the iteration runs as a
      // captured block runs it (kernel times measured one by one leave out
      // how each kernel finds the caches), and the fastest wins: with fused
      // iterations their steady state, computeLinearRHS and the update fused
      // with the next computeRHS, then the first sweep; else computeRHS and
      // the first sweep.  r, r0, dq, q and sigma are restored below.
      if ((rc = ensure_face_states(x))) return rc;
      const std::vector<Group> orig = x->groups;
      const bool fused = can_fuse(x);
      auto seq = [&](int src, const std::vector<Group>& g, double* ms) -> int {
        x->groups = g;
        int rc2 = 0;
        if (fused) {
          // the source's stored arrays and a valid dq first, then the steady state
          rc2 = enqueue_rhs_sweep0(x, nullptr, nullptr, x->partial, src);
          // the least of five timings: sources within a few percent of each
          // other are told apart by their best, not by one noisy sample
          for (int k = 0; k < 5 && !rc2; k++) {
            double t = 0;
            rc2 = timed(
                [&]() {
                  const int rc3 = launch_lin_rhs(x, src, x->partial, x->partial, nullptr);
                  return rc3 ? rc3 : enqueue_rhs_sweep0(x, nullptr, nullptr, x->partial, src, true);
                },
                &t);
            if (!rc2) *ms = k == 0 ? t : std::min(*ms, t);
          }
        } else {
          for (int k = 0; k < 5 && !rc2; k++) {
            double t = 0;
            rc2 = timed([&]() { return enqueue_rhs_sweep0(x, nullptr, nullptr, x->partial, src); }, &t);
            if (!rc2) *ms = k == 0 ? t : std::min(*ms, t);
          }
        }
        x->groups = orig;
        return rc2;
      };
      double best = 0;
      if ((rc = seq(HAMSTR_JAC_STORED, gs, &best))) return rc;
      for (int src : {HAMSTR_JAC_FROM_STATES, HAMSTR_JAC_FROM_Q, HAMSTR_JAC_STATES_LINES}) {
        if (src == HAMSTR_JAC_STATES_LINES && x->jac_to <= 0) continue;
        std::vector<Group> gq = orig;
        std::vector<double> perq;
        double sumq = 0, t = 0;
        if ((rc = tune_groups(src, gq, perq, &sumq))) return rc;
        if ((rc = seq(src, gq, &t))) return rc;
        if (t < best) {
          best = t;
          rj = src;
          gs.swap(gq);
          per.swap(perq);
        }
      }
      // the state the trials stepped restored, and every face's Jacobians
      // stored again for it, as hamstr_compute_rhs leaves them
      HIPCHK(hipMemcpyAsync(x->q, q_save, c4, hipMemcpyDeviceToDevice, x->stream));
      if ((rc = launch_rhs_faces(x, true, 0))) return rc;
    }
    x->groups = gs;
    x->jac_src = rj;
    for (size_t k = 0; k < per.size() && best_ms && (int)k < cap; k++) best_ms[k] = per[k];
    return HAMSTR_OK;
  };
  const int rc = search();
  if (saved) {
    (void)hipMemcpyAsync(x->r, r_save, c4, hipMemcpyDeviceToDevice, x->stream);
    (void)hipMemcpyAsync(x->dq, dq_save, c4, hipMemcpyDeviceToDevice, x->stream);
    (void)hipMemcpyAsync(x->r0, r0_save, c4, hipMemcpyDeviceToDevice, x->stream);
    (void)hipMemcpyAsync(x->q, q_save, c4, hipMemcpyDeviceToDevice, x->stream);
    (void)hipMemcpyAsync(x->sigma, sg_save, c4 / 4, hipMemcpyDeviceToDevice, x->stream);
    (void)hipMemcpyAsync(x->nanflag, nan_save, sizeof(int), hipMemcpyDeviceToDevice, x->stream);
  }
  const hipError_t se = hipStreamSynchronize(x->stream);
  if (r_save) (void)hipFree(r_save);
  if (dq_save) (void)hipFree(dq_save);
  if (r0_save) (void)hipFree(r0_save);
  if (q_save) (void)hipFree(q_save);
  if (sg_save) (void)hipFree(sg_save);
  if (nan_save) (void)hipFree(nan_save);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc) return rc;
  HIPCHK(se);
  if (ngroups) *ngroups = (int)x->groups.size();
  return HAMSTR_OK;
}

// launch groups: per group (phase, fast, final, chains, maxN, grid, block)
int hamstr_launch_groups(const hamstr_ctx* x, int* out8, int cap, int* n) {
  if (!x || !n) return fail(HAMSTR_ERR_ARG, "null argument");
  int k = 0;
  for (const Group& g : x->groups) {
    if (out8 && k < cap) {
      int* o = out8 + 8 * k;
      o[0] = g.phase;
      o[1] = !g.fast ? kKernExact : (g.two ? kKernTwo : (g.kern ? g.kern : kKernRow));
      o[2] = g.final; o[3] = g.n; o[4] = g.maxN; o[5] = g.grid; o[6] = g.block;
      o[7] = g.ncr;
    }
    k++;
  }
  *n = k;
  return HAMSTR_OK;
}

int hamstr_sync(hamstr_ctx* x) {
  if (!x) return fail(HAMSTR_ERR_ARG, "null context");
  HIPCHK(hipStreamSynchronize(x->stream));
  return HAMSTR_OK;
}

int hamstr_set_timing(hamstr_ctx* x, int on) {
  if (!x) return fail(HAMSTR_ERR_ARG, "null context");
  x->timing = on != 0;
  return HAMSTR_OK;
}

int hamstr_last_timing_names(hamstr_ctx* x, char* buf, int cap) {
  if (!x || !buf || cap <= 0) return fail(HAMSTR_ERR_ARG, "null argument");
  std::string all;
  for (size_t k = 0; k < x->stage_names.size(); k++) {
    if (k) all += ",";
    all += x->stage_names[k];
  }
  std::snprintf(buf, cap, "%s", all.c_str());
  return HAMSTR_OK;
}

int hamstr_last_timings(hamstr_ctx* x, double* ms, int cap, int* n) {
  if (!x || !ms || !n) return fail(HAMSTR_ERR_ARG, "null argument");
  int k = 0;
  for (; k < cap && k < (int)x->last_ms.size(); k++) ms[k] = x->last_ms[k];
  *n = k;
  return HAMSTR_OK;
}

}  // extern "C"

// ---- unit entry points ------------------------------------------------------
namespace {
struct Scoped {
  std::vector<void*> ptrs;
  ~Scoped() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename V>
  int up(V** d, const V* h, size_t n) {
    HIPCHK(hipMalloc(d, sizeof(V) * std::max<size_t>(n, 1)));
    ptrs.push_back(*d);
    if (h) HIPCHK(hipMemcpy(*d, h, sizeof(V) * n, hipMemcpyHostToDevice));
    return HAMSTR_OK;
  }
};
}  // namespace

extern "C" {

int hamstr_flux_roe2d(int n, const double* nvec, const double* ql, const double* qr, double gamma, double* flux,
                      double* sr) {
  if (n <= 0 || !nvec || !ql || !qr || !flux || !sr) return fail(HAMSTR_ERR_ARG, "bad arguments");
  Scoped s;
  double *dn, *dl, *dr, *df, *ds;
  int* dflag;
  int rc;
  const int zero = 0;
  if ((rc = s.up(&dn, nvec, 2 * (size_t)n)) || (rc = s.up(&dl, ql, 4 * (size_t)n)) ||
      (rc = s.up(&dr, qr, 4 * (size_t)n)) || (rc = s.up(&df, (const double*)nullptr, 4 * (size_t)n)) ||
      (rc = s.up(&ds, (const double*)nullptr, (size_t)n)) || (rc = s.up(&dflag, &zero, 1)))
    return rc;
  hipLaunchKernelGGL(k_unit_flux, dim3(nblk(n)), dim3(kBlock), 0, 0, n, dn, dl, dr, gamma, df, ds, dflag);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(flux, df, sizeof(double) * 4 * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(sr, ds, sizeof(double) * n, hipMemcpyDeviceToHost));
  int flag = 0;
  HIPCHK(hipMemcpy(&flag, dflag, sizeof(int), hipMemcpyDeviceToHost));
  return flag ? fail(HAMSTR_ERR_NAN, "NaN spectral radius") : HAMSTR_OK;
}

int hamstr_jacobian_roe2d(int n, const double* nvec, const double* ql, const double* qr, double gamma, int lhs,
                          double* lmat, double* rmat) {
  if (n <= 0 || !nvec || !ql || !qr || !lmat || !rmat) return fail(HAMSTR_ERR_ARG, "bad arguments");
  Scoped s;
  double *dn, *dl, *dr, *dlm, *drm;
  int rc;
  if ((rc = s.up(&dn, nvec, 2 * (size_t)n)) || (rc = s.up(&dl, ql, 4 * (size_t)n)) ||
      (rc = s.up(&dr, qr, 4 * (size_t)n)) || (rc = s.up(&dlm, (const double*)nullptr, 16 * (size_t)n)) ||
      (rc = s.up(&drm, (const double*)nullptr, 16 * (size_t)n)))
    return rc;
  if (lhs == HAMSTR_LHS_FP32)
    hipLaunchKernelGGL(k_unit_jac<float>, dim3(nblk(n)), dim3(kBlock), 0, 0, n, dn, dl, dr, gamma, dlm, drm);
  else
    hipLaunchKernelGGL(k_unit_jac<double>, dim3(nblk(n)), dim3(kBlock), 0, 0, n, dn, dl, dr, gamma, dlm, drm);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(lmat, dlm, sizeof(double) * 16 * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(rmat, drm, sizeof(double) * 16 * n, hipMemcpyDeviceToHost));
  return HAMSTR_OK;
}

}  // extern "C"
static int unit_recon(int kind, int nchain, int imax, int is, int iend, double qt, double eps, const double* f,
                      const double* dq, double* ql, double* qr) {
  if (nchain <= 0 || imax <= 0 || !f || !ql || !qr) return fail(HAMSTR_ERR_ARG, "bad arguments");
  const bool weno = (kind == 1 || kind == 3);
  if (is < (weno ? 2 : 1) || iend > (weno ? imax - 2 : imax - 1) || iend <= is)
    return fail(HAMSTR_ERR_ARG, "row range outside the chain");
  if ((kind >= 2) && !dq) return fail(HAMSTR_ERR_ARG, "dq required");
  const size_t tot = (size_t)nchain * imax * 4;
  Scoped s;
  double *df, *dd = nullptr, *dl, *dr;
  int rc;
  if ((rc = s.up(&df, f, tot)) || (rc = s.up(&dl, ql, tot)) || (rc = s.up(&dr, qr, tot))) return rc;
  if (kind >= 2 && (rc = s.up(&dd, dq, tot))) return rc;
  const int nth = nchain * (iend - is) * 4;
  hipLaunchKernelGGL(k_unit_recon, dim3(nblk(nth)), dim3(kBlock), 0, 0, kind, nchain, imax, is, iend, qt, eps, df,
                     dd, dl, dr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(ql, dl, sizeof(double) * tot, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(qr, dr, sizeof(double) * tot, hipMemcpyDeviceToHost));
  return HAMSTR_OK;
}

extern "C" {
// MUSCL rows is..ie-1 (row ie uses a stale slope in the reference and is
// never read by its callers); first order (qt == 0) copies rows is..ie.
int hamstr_muscl(int nchain, int imax, int is, int ie, double th, double qt, double eps, const double* f,
                 double* ql, double* qr) {
  if (th != kThird) return fail(HAMSTR_ERR_ARG, "th must be 1/3 (computeRHS.cpp:268)");
  if (qt == 0.0) {
    if (is < 0 || ie >= imax || ie < is) return fail(HAMSTR_ERR_ARG, "row range outside the chain");
    for (int c = 0; c < nchain; c++)
      for (int i = is; i <= ie; i++)
        for (int k = 0; k < 4; k++) {
          const size_t o = ((size_t)c * imax + i) * 4 + k;
          ql[o] = qr[o] = f[o];
        }
    return HAMSTR_OK;
  }
  return unit_recon(0, nchain, imax, is, ie, qt, eps, f, nullptr, ql, qr);
}

int hamstr_weno(int nchain, int imax, int is, int ie, double eps, const double* f, double* ql, double* qr) {
  (void)ie;
  (void)eps;
  return unit_recon(1, nchain, imax, is, imax - 2, 0.0, 1e-15, f, nullptr, ql, qr);
}

int hamstr_muscl_deriv(int nchain, int imax, int is, int ie, double th, double qt, double eps, const double* f,
                       const double* dq, double* ql, double* qr) {
  if (th != kThird) return fail(HAMSTR_ERR_ARG, "th must be 1/3 (computeLinearRHS.cpp:283)");
  if (qt == 0.0) {
    if (is < 0 || ie >= imax || ie < is) return fail(HAMSTR_ERR_ARG, "row range outside the chain");
    for (int c = 0; c < nchain; c++)
      for (int i = is; i <= ie; i++)
        for (int k = 0; k < 4; k++) {
          const size_t o = ((size_t)c * imax + i) * 4 + k;
          ql[o] = qr[o] = dq[o];
        }
    return HAMSTR_OK;
  }
  return unit_recon(2, nchain, imax, is, ie, qt, eps, f, dq, ql, qr);
}

int hamstr_weno_deriv(int nchain, int imax, int is, int ie, double eps, const double* f, const double* dq,
                      double* ql, double* qr) {
  (void)ie;
  return unit_recon(3, nchain, imax, is, imax - 2, 0.0, eps, f, dq, ql, qr);
}

}  // extern "C"
static int unit_tridag(int periodic, int nsys, int N, const double* a, const double* b, const double* c,
                       double* f) {
  if (nsys <= 0 || N < (periodic ? 3 : 1) || !a || !b || !c || !f) return fail(HAMSTR_ERR_ARG, "bad arguments");
  Scoped s;
  double *da, *db, *dc, *df, *ds;
  int rc;
  const size_t m = (size_t)nsys * N;
  if ((rc = s.up(&da, a, 16 * m)) || (rc = s.up(&db, b, 16 * m)) || (rc = s.up(&dc, c, 16 * m)) ||
      (rc = s.up(&df, f, 4 * m)) || (rc = s.up(&ds, (const double*)nullptr, 36 * m)))
    return rc;
  hipLaunchKernelGGL(k_unit_tridag, dim3(nblk(nsys, 64)), dim3(64), 0, 0, nsys, N, periodic, da, db, dc, df, ds);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(f, df, sizeof(double) * 4 * m, hipMemcpyDeviceToHost));
  return HAMSTR_OK;
}

// ---- the 3D / strand building blocks ------------------------------------------
extern "C" {
int hamstr_flux_roe3d(int n, const double* nxyz, const double* ql, const double* qr, double gamma, double* flux,
                      double* sr) {
  if (n <= 0 || !nxyz || !ql || !qr || !flux || !sr) return fail(HAMSTR_ERR_ARG, "bad arguments");
  Scoped s;
  double *dn, *dl, *dr, *df, *ds;
  int rc;
  if ((rc = s.up(&dn, nxyz, 3 * (size_t)n)) || (rc = s.up(&dl, ql, 5 * (size_t)n)) ||
      (rc = s.up(&dr, qr, 5 * (size_t)n)) || (rc = s.up(&df, (const double*)nullptr, 5 * (size_t)n)) ||
      (rc = s.up(&ds, (const double*)nullptr, (size_t)n)))
    return rc;
  hipLaunchKernelGGL(k_unit_flux3d, dim3(nblk(n)), dim3(kBlock), 0, 0, n, dn, dl, dr, gamma, df, ds);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(flux, df, sizeof(double) * 5 * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(sr, ds, sizeof(double) * n, hipMemcpyDeviceToHost));
  return HAMSTR_OK;
}

int hamstr_jacobian_roe3d(int n, const double* nxyz, const double* ql, const double* qr, double gamma,
                          double* lmat, double* rmat) {
  if (n <= 0 || !nxyz || !ql || !qr || !lmat || !rmat) return fail(HAMSTR_ERR_ARG, "bad arguments");
  Scoped s;
  double *dn, *dl, *dr, *dlm, *drm;
  int rc;
  if ((rc = s.up(&dn, nxyz, 3 * (size_t)n)) || (rc = s.up(&dl, ql, 5 * (size_t)n)) ||
      (rc = s.up(&dr, qr, 5 * (size_t)n)) || (rc = s.up(&dlm, (const double*)nullptr, 25 * (size_t)n)) ||
      (rc = s.up(&drm, (const double*)nullptr, 25 * (size_t)n)))
    return rc;
  hipLaunchKernelGGL(k_unit_jac3d, dim3(nblk(n)), dim3(kBlock), 0, 0, n, dn, dl, dr, gamma, dlm, drm);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(lmat, dlm, sizeof(double) * 25 * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(rmat, drm, sizeof(double) * 25 * n, hipMemcpyDeviceToHost));
  return HAMSTR_OK;
}
}  // extern "C"

static int unit_tridag5(int periodic, int nsys, int N, const double* a, const double* b, const double* c,
                        double* f) {
  if (nsys <= 0 || N < (periodic ? 3 : 1) || !a || !b || !c || !f) return fail(HAMSTR_ERR_ARG, "bad arguments");
  Scoped s;
  double *da, *db, *dc, *df;
  int rc;
  const size_t m = (size_t)nsys * N;
  if ((rc = s.up(&da, a, 25 * m)) || (rc = s.up(&db, b, 25 * m)) || (rc = s.up(&dc, c, 25 * m)) ||
      (rc = s.up(&df, f, 5 * m)))
    return rc;
  hipLaunchKernelGGL(k_unit_tridag5, dim3(nblk(nsys, 64)), dim3(64), 0, 0, nsys, N, periodic, da, db, dc, df);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(f, df, sizeof(double) * 5 * m, hipMemcpyDeviceToHost));
  return HAMSTR_OK;
}

extern "C" {
int hamstr_block_tridag5(int nsys, int N, const double* a, const double* b, const double* c, double* f) {
  return unit_tridag5(0, nsys, N, a, b, c, f);
}

int hamstr_block_tridag_periodic5(int nsys, int N, const double* a, const double* b, const double* c, double* f) {
  return unit_tridag5(1, nsys, N, a, b, c, f);
}
}  // extern "C"

// ---- block solves that allocate nothing -------------------------------------
// The scratch rows of k_unit_tridag (C', F' and the spike column, 36 doubles
// per block row) live in this library-owned array instead of a hipMalloc per
// call.  Each GPU holds its own copy for the life of the process; every call
// on one GPU shares it, so calls that overlap on one GPU race.  The kernel
// writes every scratch row before it reads it, so nothing needs clearing.
static __device__ __align__(256) double scratch[36 * (size_t)HAMSTR_NOALLOC_MAX_ROWS];

namespace {
std::mutex g_scratch_mu;
std::vector<double*> g_scratch_dev;  // device address of `scratch`, per device

// The device address of `scratch` on the current device, looked up once per
// device.  Host code uses only this address: the array's name is a host-side
// stand-in.
int scratch_address(double** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_dev.size() <= dev) g_scratch_dev.resize(dev + 1, nullptr);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(HAMSTR_ERR_HIP, "hipGetSymbolAddress(scratch) gave a null address");
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[dev];
  return HAMSTR_OK;
}
}  // namespace

extern "C" {
int hamstr_block_tridag4(int nsys, int N, const double* a, const double* b, const double* c, double* f) {
  return unit_tridag(0, nsys, N, a, b, c, f);
}

int hamstr_block_tridag4_noalloc(int periodic, int nsys, int N, const double* a, const double* b, const double* c,
                                 double* f, void* stream) {
  if ((periodic != 0 && periodic != 1) || nsys <= 0 || N < (periodic ? 3 : 1) || !a || !b || !c || !f)
    return fail(HAMSTR_ERR_ARG, "bad arguments");
  if ((long long)nsys * N > HAMSTR_NOALLOC_MAX_ROWS)
    return fail(HAMSTR_ERR_ARG, "nsys*N = " + std::to_string((long long)nsys * N) + " block rows exceed the " +
                                    std::to_string(HAMSTR_NOALLOC_MAX_ROWS) + " the allocation-free solve holds");
  double* S = nullptr;
  const int rc = scratch_address(&S);
  if (rc) return rc;
  hipLaunchKernelGGL(k_unit_tridag, dim3(nblk(nsys, 64)), dim3(64), 0, static_cast<hipStream_t>(stream), nsys, N,
                     periodic, a, b, c, f, S);
  HIPCHK(hipGetLastError());
  return HAMSTR_OK;
}

int hamstr_block_tridag_periodic4(int nsys, int N, const double* a, const double* b, const double* c, double* f) {
  return unit_tridag(1, nsys, N, a, b, c, f);
}

}  // extern "C"
#endif  // !HAMSTR_LONG_TU

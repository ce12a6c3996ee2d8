// adi_long.hip -- the FAST ADI kernel for chains longer than a workgroup
// (adi_fast_chain_long, k_adi_fast_long in hamstr_hip.hip), compiled as a
// translation unit of its own so that its code does not change how the
// common k_adi_fast is compiled, and its launcher.
#define HAMSTR_LONG_TU
#include "hamstr_hip.hip"

hipError_t hamstr_adi_long_launch(bool fp32, int final_type, unsigned grid, unsigned block, size_t lds,
                                  hipStream_t stream, const void* ct_, const void* lv_, const int* chains, int nch,
                                  const void* ff, int nProc, double cfl, const double* sigma, double* r, double* dq,
                                  int dqmode, int ncr, void* hs, int hstride, const int* hbase) {
  ChainTab ct;
  Live lv;
  std::memcpy(&ct, ct_, sizeof ct);
  std::memcpy(&lv, lv_, sizeof lv);
  const dim3 g(grid), b(block);
#define LONGL(T, F, M)                                                                                     \
  hipLaunchKernelGGL((k_adi_fast_long<T, F, M>), g, b, lds, stream, ct, lv, chains, nch, (const T*)ff, nProc, \
                     cfl, sigma, r, dq, dqmode, ncr, (T*)hs, hstride, hbase)
  if (fp32) {
    if (final_type == kFinalOpen) LONGL(float, kFinalOpen, 1024);
    else if (final_type == kFinalPair) LONGL(float, kFinalPair, 1024);
    else LONGL(float, kFinalOdd, 1024);
  } else {
    if (final_type == kFinalOpen) LONGL(double, kFinalOpen, 512);
    else if (final_type == kFinalPair) LONGL(double, kFinalPair, 512);
    else LONGL(double, kFinalOdd, 512);
  }
#undef LONGL
  return hipGetLastError();
}

// launcher of k_adi_fast_half (FP64) / k_adi_fast_half2 (FP64 or FP32, pairs
// != 0: two open chains per 768-thread workgroup); chains of at most 384 rows,
// one row per thread, at most 168 VGPRs
hipError_t hamstr_adi_half_launch(bool fp32, int final_type, int pairs, unsigned grid, unsigned block, size_t lds,
                                  hipStream_t stream, const void* ct_, const void* lv_, const void* js_,
                                  const int* chains, int nch, const void* ff, int nProc, double cfl,
                                  const double* sigma, double* r, double* dq, int dqmode, int ncr, void* hs,
                                  int hstride, const int* hbase) {
  ChainTab ct;
  Live lv;
  JacSrc js;
  std::memcpy(&ct, ct_, sizeof ct);
  std::memcpy(&lv, lv_, sizeof lv);
  std::memcpy(&js, js_, sizeof js);
  const bool jq = js.order != 0;  // the Jacobians computed from q
  const dim3 g(grid), b(block);
  if (pairs) {
    if (final_type != kFinalOpen || block != 2 * 384) return hipErrorInvalidValue;
#define PAIRL(T, J)                                                                                              \
  hipLaunchKernelGGL((k_adi_fast_half2<T, 384, J>), g, b, lds, stream, ct, lv, js, chains, nch, (const T*)ff, nProc, \
                     cfl, sigma, r, dq, dqmode, ncr, (T*)hs, hstride, hbase)
    if (fp32) {
      if (jq) PAIRL(float, true); else PAIRL(float, false);
    } else {
      if (jq) PAIRL(double, true); else PAIRL(double, false);
    }
#undef PAIRL
    return hipGetLastError();
  }
  if (fp32 || block > 384) return hipErrorInvalidValue;
#define HALFJ(F, M, J)                                                                                          \
  hipLaunchKernelGGL((k_adi_fast_half<double, F, M, 3, J>), g, b, lds, stream, ct, lv, js, chains, nch,             \
                     (const double*)ff, nProc, cfl, sigma, r, dq, dqmode, ncr, (double*)hs, hstride, hbase)
#define HALFL(F, M)          \
  do {                       \
    if (jq)                  \
      HALFJ(F, M, true);     \
    else                     \
      HALFJ(F, M, false);    \
  } while (0)
  // 192-thread instance (a 28.5 KB LDS slice) for chains of up to 192 rows
  if (block <= 192) {
    if (final_type == kFinalOpen) HALFL(kFinalOpen, 192);
    else if (final_type == kFinalPair) HALFL(kFinalPair, 192);
    else HALFL(kFinalOdd, 192);
  } else {
    if (final_type == kFinalOpen) HALFL(kFinalOpen, 384);
    else if (final_type == kFinalPair) HALFL(kFinalPair, 384);
    else HALFL(kFinalOdd, 384);
  }
#undef HALFL
#undef HALFJ
  return hipGetLastError();
}

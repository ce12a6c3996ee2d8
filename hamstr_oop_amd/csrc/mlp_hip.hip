// mlp_hip.hip -- the learned warm start: an MLP (6 -> 256 x 4 SiLU -> 4) that
// maps (Mach, alpha, thickness, cell centroid x, y, wall distance) to the
// primitive state (rho, u, v, p).  The three 256x256 layers run on gfx950's
// BF16 matrix cores (v_mfma_f32_32x32x16_bf16, FP32 accumulation) for the
// forward pass and both backward products, with 16-byte operand loads (the
// activations row-major, the weights stored both ways) and, for the weight
// gradients, batch tiles transposed through LDS; the 6-wide input and 4-wide
// output layers run on the vector ALUs in FP32.  Weights are FP32 masters with
// BF16 copies; training is Adam on mean-squared error.  Every reduction over
// the batch is split into fixed chunks whose partials one kernel sums in a
// fixed order, so training is deterministic.
//
// The network only supplies the initial state of a solve (hamstr_warm_start);
// the ADI iteration decides convergence, and cells whose prediction is not a
// valid state fall back to the freestream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/hamstr_hip.h"

namespace {

thread_local std::string m_err;
int mfail(int code, const std::string& msg) {
  m_err = msg;
  return code;
}
#define MCHK(x)                                                                                \
  do {                                                                                         \
    hipError_t e_ = (x);                                                                       \
    if (e_ != hipSuccess) return mfail(HAMSTR_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int kIn = 6, kH = 256, kOut = 4, kHidden = 4;
// flat parameter layout (floats)
constexpr int oW1 = 0, oB1 = oW1 + kIn * kH;
constexpr int oW2 = oB1 + kH, oB2 = oW2 + kH * kH;
constexpr int oW3 = oB2 + kH, oB3 = oW3 + kH * kH;
constexpr int oW4 = oB3 + kH, oB4 = oW4 + kH * kH;
constexpr int oW5 = oB4 + kH, oB5 = oW5 + kH * kOut;
constexpr int kParams = oB5 + kOut;  // 200196
constexpr int kSplit = 16;           // at most this many batch chunks per reduction

typedef unsigned short bf16;
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef unsigned short v8u __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

// batch chunks of the reductions: `rows` per chunk (a multiple of the
// matrix-core K step, 16), `n` chunks, n <= kSplit
struct Chunks {
  int rows, n;
};
inline Chunks chunks_of(int B) {
  const int rows = ((B + kSplit - 1) / kSplit + 15) / 16 * 16;
  return Chunks{rows, (B + rows - 1) / rows};
}

__device__ __forceinline__ bf16 f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (bf16)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float(((unsigned)h) << 16); }
__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }

// --------------------------------------------------------------------------
// C[M x 256] = A[M x 256] * B on matrix cores, B given as Bt[n][k] (both with
// k contiguous, so every operand fragment is one 16-byte load).  One 32x32
// tile per wave; a workgroup of four waves covers 32 rows and 128 columns
// (grid M/32 x 2); K in steps of 16.  Lane l supplies
// A[m0 + l%32][k0 + 8*(l/32) + j] and Bt[n0 + l%32][k0 + 8*(l/32) + j],
// j = 0..7, and holds D[m0 + 8*(v/4) + 4*(l/32) + v%4][n0 + l%32] in v.
// EPI 0 (forward): Z = acc + bias -> zout, H = silu(Z) -> hout (bf16)
// EPI 1 (backward): acc is dH of the layer below; dZ = dH silu'(Z) with Z
//   from zin -> dzout (f32) and dzbout (bf16)
// --------------------------------------------------------------------------
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_rows(const bf16* __restrict__ A, const bf16* __restrict__ Bt,
                                                   float* __restrict__ zout, const float* __restrict__ bias,
                                                   bf16* __restrict__ hout, const float* __restrict__ zin,
                                                   float* __restrict__ dzout, bf16* __restrict__ dzbout) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m0 = blockIdx.x * 32, n0 = blockIdx.y * 128 + wave * 32;
  const int r = lane & 31, kh = 8 * (lane >> 5);
  const bf16* arow = A + (long)(m0 + r) * kH + kh;
  const bf16* brow = Bt + (long)(n0 + r) * kH + kh;
  v16f acc = {};
#pragma unroll
  for (int k0 = 0; k0 < kH; k0 += 16) {
    const v8bf a = *reinterpret_cast<const v8bf*>(arow + k0);
    const v8bf b = *reinterpret_cast<const v8bf*>(brow + k0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 16; v++) {
    const int row = m0 + 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
    const int col = n0 + r;
    const long o = (long)row * kH + col;
    if (EPI == 0) {
      const float z = acc[v] + bias[col];
      zout[o] = z;
      hout[o] = f2bf(z * sigm(z));
    } else {
      const float zz = zin[o];
      const float sg = sigm(zz);
      const float g = acc[v] * (sg + zz * sg * (1.0f - sg));
      dzout[o] = g;
      dzbout[o] = f2bf(g);
    }
  }
}

// --------------------------------------------------------------------------
// Weight gradient of a hidden layer over one batch chunk, dW = H^T dZ
// (m = input unit i, n = output unit j, k = batch row), into the chunk's
// partials part[z][oW + i*256 + j].  A workgroup covers a 64x64 block of dW
// (four waves of 32x32).  The chunk's rows go through LDS up to 256 at a
// time, transposed so that every fragment is one 16-byte LDS read: thread t
// takes the 16-byte segment t/32 of the block's units in rows 2p, 2p+1 (+64m),
// p = t%32, issues all its loads first and writes each unit's pair of rows as
// one 4-byte word.  The workgroups of the first row of blocks also sum the
// chunk's FP32 dZ into the bias partials part[z][oB + j], in a fixed order.
// --------------------------------------------------------------------------
constexpr int kDwRows = 256;           // batch rows per LDS stage
constexpr int kDwPitch = kDwRows + 8;  // LDS row pitch (elements); 16-byte aligned rows
__global__ __launch_bounds__(256) void k_gemm_dw(const bf16* __restrict__ H, const bf16* __restrict__ dZb,
                                                 const float* __restrict__ dz, int B, int rows,
                                                 float* __restrict__ part, int oW, int oB) {
  __shared__ __attribute__((aligned(16))) bf16 Ht[64][kDwPitch];
  __shared__ __attribute__((aligned(16))) bf16 Zt[64][kDwPitch];
  __shared__ float red[4][64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 64, z = blockIdx.z;
  const int r0 = z * rows, r1 = min(B, r0 + rows);
  const int wi = wave & 1, wj = wave >> 1;
  const int r = lane & 31, kh = 8 * (lane >> 5);
  const int seg = t >> 5, pr = t & 31;
  float* out = part + (long)z * kParams;
  v16f acc = {};
  for (int s0 = r0; s0 < r1; s0 += kDwRows) {
    const int len = min(kDwRows, r1 - s0);  // a multiple of 16
    v8u hv[8], zv[8];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int w = 0; w < 2; w++) {
        const int q = 64 * m + 2 * pr + w;
        if (q < len) {
          const long base = (long)(s0 + q) * kH + seg * 8;
          hv[2 * m + w] = *reinterpret_cast<const v8u*>(H + base + i0);
          zv[2 * m + w] = *reinterpret_cast<const v8u*>(dZb + base + j0);
        }
      }
    __syncthreads();  // the previous stage's reads are done
#pragma unroll
    for (int m = 0; m < 4; m++) {
      const int q = 64 * m + 2 * pr;
      if (q < len) {
#pragma unroll
        for (int e = 0; e < 8; e++) {
          *reinterpret_cast<unsigned*>(&Ht[seg * 8 + e][q]) = (unsigned)hv[2 * m][e] | ((unsigned)hv[2 * m + 1][e] << 16);
          *reinterpret_cast<unsigned*>(&Zt[seg * 8 + e][q]) = (unsigned)zv[2 * m][e] | ((unsigned)zv[2 * m + 1][e] << 16);
        }
      }
    }
    __syncthreads();
    for (int k0 = 0; k0 < len; k0 += 16) {
      const v8bf a = *reinterpret_cast<const v8bf*>(&Ht[wi * 32 + r][k0 + kh]);
      const v8bf b = *reinterpret_cast<const v8bf*>(&Zt[wj * 32 + r][k0 + kh]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
  }
#pragma unroll
  for (int v = 0; v < 16; v++) {
    const int i = i0 + wi * 32 + 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
    const int j = j0 + wj * 32 + r;
    out[oW + (long)i * kH + j] = acc[v];
  }
  if (blockIdx.x == 0) {
    // bias: four interleaved partial sums per column, added in a fixed order
    const int j = j0 + (t & 63), q = t >> 6;
    float sum = 0.f;
#pragma unroll 8
    for (int row = r0 + q; row < r1; row += 4) sum += dz[(long)row * kH + j];
    red[q][t & 63] = sum;
    __syncthreads();
    if (q == 0) out[oB + j] = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
  }
}

// layer 1 (K = 6) on the vector ALUs: Z1 = X W1 + b1, H1 = silu(Z1)
__global__ void k_layer1(const float* __restrict__ x, const float* __restrict__ p, int B, float* __restrict__ z,
                         bf16* __restrict__ h) {
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (long)B * kH) return;
  const int row = id / kH, col = id % kH;
  float s = p[oB1 + col];
#pragma unroll
  for (int k = 0; k < kIn; k++) s += x[row * kIn + k] * p[oW1 + k * kH + col];
  z[id] = s;
  h[id] = f2bf(s * sigm(s));
}

// output layer (N = 4), one thread per output: Y = H4 W5 + b5, summed over
// the 256 hidden units in order.  With t: dY = 2 (Y - T) / (B * 4) and the
// per-block loss partials, summed in a fixed tree.
__global__ __launch_bounds__(256) void k_out_loss(const bf16* __restrict__ h, const float* __restrict__ p, int B,
                                                  float* __restrict__ y, const float* __restrict__ t,
                                                  float* __restrict__ dy, float* __restrict__ lpart) {
  __shared__ float sh[256];
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (id < B * kOut) {
    const int row = id / kOut, j = id % kOut;
    float s = p[oB5 + j];
    const bf16* hr = h + (long)row * kH;
    for (int k0 = 0; k0 < kH; k0 += 8) {
      const v8u hv = *reinterpret_cast<const v8u*>(hr + k0);
#pragma unroll
      for (int e = 0; e < 8; e++) s += bf2f(hv[e]) * p[oW5 + (k0 + e) * kOut + j];
    }
    y[id] = s;
    if (t) {
      const float d = s - t[id];
      dy[id] = 2.0f * d / (float)(B * kOut);
      l = d * d / (float)(B * kOut);
    }
  }
  if (!t) return;  // the same for the whole launch
  sh[threadIdx.x] = l;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) lpart[blockIdx.x] = sh[0];
}

// dZ4 = (dY W5^T) silu'(Z4), as f32 and bf16
__global__ void k_out_bwd_dz(const float* __restrict__ dy, const float* __restrict__ p,
                             const float* __restrict__ z, int B, float* __restrict__ dz,
                             bf16* __restrict__ dzb) {
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (long)B * kH) return;
  const int row = id / kH, col = id % kH;
  float dh = 0.f;
#pragma unroll
  for (int j = 0; j < kOut; j++) dh += dy[row * kOut + j] * p[oW5 + col * kOut + j];
  const float zz = z[id];
  const float sg = sigm(zz);
  const float g = dh * (sg + zz * sg * (1.0f - sg));
  dz[id] = g;
  dzb[id] = f2bf(g);
}

// W5 and b5 gradient partials of one batch chunk (block z; thread = hidden
// unit i and row group g = t/256, which takes the chunk's rows g, g+4, ...):
// part[z][oW5 + 4i + j] = sum_rows H4[row][i] dY[row][j] and
// part[z][oB5 + j] = sum_rows dY[row][j], the four groups added in order
__global__ __launch_bounds__(1024) void k_outer_h4(const bf16* __restrict__ h4, const float* __restrict__ dy, int B,
                                                   int rows, float* __restrict__ part) {
  __shared__ float red[4][kH][kOut + 1];
  const int i = threadIdx.x & (kH - 1), g = threadIdx.x >> 8, z = blockIdx.x;
  const int r0 = z * rows, r1 = min(B, r0 + rows);
  float s[kOut] = {0.f, 0.f, 0.f, 0.f}, sb = 0.f;
#pragma unroll 4
  for (int row = r0 + g; row < r1; row += 4) {
    const float hv = bf2f(h4[(long)row * kH + i]);
    const float4 d = *reinterpret_cast<const float4*>(dy + row * kOut);
    s[0] += hv * d.x;
    s[1] += hv * d.y;
    s[2] += hv * d.z;
    s[3] += hv * d.w;
    if (i < kOut) sb += (i == 0 ? d.x : i == 1 ? d.y : i == 2 ? d.z : d.w);
  }
#pragma unroll
  for (int j = 0; j < kOut; j++) red[g][i][j] = s[j];
  red[g][i][kOut] = sb;
  __syncthreads();
  if (g != 0) return;
  float* out = part + (long)z * kParams;
#pragma unroll
  for (int j = 0; j < kOut; j++) out[oW5 + i * kOut + j] = ((red[0][i][j] + red[1][i][j]) + red[2][i][j]) + red[3][i][j];
  if (i < kOut) out[oB5 + i] = ((red[0][i][kOut] + red[1][i][kOut]) + red[2][i][kOut]) + red[3][i][kOut];
}

// W1 and b1 gradient partials of one batch chunk (block z; thread = hidden
// unit j and row group g): part[z][oW1 + 256k + j] = sum_rows X[row][k] dZ1[row][j]
// and part[z][oB1 + j] = sum_rows dZ1[row][j], the four groups added in order
__global__ __launch_bounds__(1024) void k_outer_x(const float* __restrict__ x, const float* __restrict__ dz, int B,
                                                  int rows, float* __restrict__ part) {
  __shared__ float red[4][kIn + 1][kH];
  const int j = threadIdx.x & (kH - 1), g = threadIdx.x >> 8, z = blockIdx.x;
  const int r0 = z * rows, r1 = min(B, r0 + rows);
  float s[kIn] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, sb = 0.f;
#pragma unroll 4
  for (int row = r0 + g; row < r1; row += 4) {
    const float d = dz[(long)row * kH + j];
#pragma unroll
    for (int k = 0; k < kIn; k++) s[k] += x[row * kIn + k] * d;
    sb += d;
  }
#pragma unroll
  for (int k = 0; k < kIn; k++) red[g][k][j] = s[k];
  red[g][kIn][j] = sb;
  __syncthreads();
  if (g != 0) return;
  float* out = part + (long)z * kParams;
#pragma unroll
  for (int k = 0; k < kIn; k++) out[oW1 + k * kH + j] = ((red[0][k][j] + red[1][k][j]) + red[2][k][j]) + red[3][k][j];
  out[oB1 + j] = ((red[0][kIn][j] + red[1][kIn][j]) + red[2][kIn][j]) + red[3][kIn][j];
}

// the gradient: g[i] = sum over the nz chunks of part[z][i], in chunk order
__global__ void k_reduce_all(const float* __restrict__ part, int nz, float* __restrict__ g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kParams) return;
  float s = 0.f;
  for (int z = 0; z < nz; z++) s += part[(long)z * kParams + i];
  g[i] = s;
}

// Adam with the step's learning rate (cosine from lr to lr_final over `steps`)
// and bias corrections computed from the device step counter, so that a
// captured training step can be replayed unchanged.  The new W2..W4 also go
// to their BF16 copies (row-major and transposed) for the next step.
struct Bf16Copies {
  bf16 *wb[3], *wtb[3];
};
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, const int* __restrict__ step, long t0, int steps, float lr0,
                       float lr_final, float b1, float b2, float eps, Bf16Copies cp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = *step;
  const float t = (float)(t0 + s + 1);
  const float c1 = 1.0f - powf(b1, t), c2 = 1.0f - powf(b2, t);
  const float lr = lr_final + 0.5f * (lr0 - lr_final) * (1.0f + cosf(3.14159265358979f * (float)s / (float)max(1, steps - 1)));
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pn = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pn;
#pragma unroll
  for (int l = 0; l < 3; l++) {
    const int w = i - (oW2 + l * (kH * kH + kH));  // W2, W3, W4 start kH*kH + kH apart
    if (w >= 0 && w < kH * kH) {
      const bf16 h = f2bf(pn);
      cp.wb[l][w] = h;
      cp.wtb[l][(w % kH) * kH + w / kH] = h;
    }
  }
}

// BF16 copies of a 256x256 weight: w (row-major K x N) and wt (N x K)
__global__ void k_w_bf16(const float* __restrict__ w, bf16* __restrict__ wb, bf16* __restrict__ wtb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kH * kH) return;
  const int k = i / kH, n = i % kH;
  const bf16 h = f2bf(w[i]);
  wb[i] = h;
  wtb[n * kH + k] = h;
}

// gather a training batch by a per-step hash of (seed, step, i); the step
// comes from the device counter
__global__ void k_gather(const float* __restrict__ X, const float* __restrict__ T, long n, int B, unsigned seed,
                         const int* __restrict__ stepp, float* __restrict__ xb, float* __restrict__ tb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B) return;
  const unsigned step = (unsigned)*stepp;
  unsigned long long h = ((unsigned long long)seed << 32) ^ ((unsigned long long)step * 0x9E3779B97F4A7C15ull) ^
                         (unsigned long long)i * 0xBF58476D1CE4E5B9ull;
  h ^= h >> 31;
  h *= 0x94D049BB133111EBull;
  h ^= h >> 29;
  const long j = (long)(h % (unsigned long long)n);
#pragma unroll
  for (int k = 0; k < kIn; k++) xb[i * kIn + k] = X[j * kIn + k];
#pragma unroll
  for (int k = 0; k < kOut; k++) tb[i * kOut + k] = T[j * kOut + k];
}

// end of a training step: the batch loss into losses[step] (losses may be
// null), then step += 1
__global__ void k_step_end(const float* __restrict__ lpart, int nlb, float* __restrict__ losses,
                           int* __restrict__ step) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (losses) {
    float s = 0.f;
    for (int z = 0; z < nlb; z++) s += lpart[z];
    losses[*step] = s;
  }
  *step += 1;
}

// features -> normalised -> (forward done by the caller) ; y normalised -> q
// primitive (rho, u, v, p) -> conserved with gamma = 1.4; invalid -> freestream
__global__ void k_norm_features(const float* __restrict__ f, const float* __restrict__ nrm, int n,
                                float* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * kIn) return;
  const int k = i % kIn;
  x[i] = (f[i] - nrm[k]) / nrm[kIn + k];
}

__global__ void k_to_state(const float* __restrict__ y, const float* __restrict__ nrm, int n,
                           const double4 qinf, double* __restrict__ q, int* __restrict__ nfallback) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // the network predicts the deviation from the case's freestream primitive state
  const double pinf[kOut] = {qinf.x, qinf.y / qinf.x, qinf.z / qinf.x,
                             0.4 * (qinf.w - 0.5 * (qinf.y * qinf.y + qinf.z * qinf.z) / qinf.x)};
  double pr[kOut];
#pragma unroll
  for (int k = 0; k < kOut; k++)
    pr[k] = pinf[k] + (double)y[i * kOut + k] * nrm[2 * kIn + kOut + k] + nrm[2 * kIn + k];
  const double rho = pr[0], u = pr[1], v = pr[2], p = pr[3];
  double4 s;
  const bool ok = rho > 1e-3 && p > 1e-3 && rho == rho && u == u && v == v && p == p && fabs(u) < 10 && fabs(v) < 10;
  if (ok) {
    s = make_double4(rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v));
  } else {
    s = qinf;
    atomicAdd(nfallback, 1);
  }
  *reinterpret_cast<double4*>(q + 4 * (size_t)i) = s;
}

}  // namespace

struct hamstr_mlp {
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  bf16 *wb[3] = {nullptr, nullptr, nullptr}, *wtb[3] = {nullptr, nullptr, nullptr};
  // activations for a batch of up to cap rows
  int cap = 0;
  float *x = nullptr, *t = nullptr, *y = nullptr, *dy = nullptr, *part = nullptr, *lpart = nullptr;
  float *z[kHidden] = {nullptr, nullptr, nullptr, nullptr};
  bf16* h[kHidden] = {nullptr, nullptr, nullptr, nullptr};
  float* dz[2] = {nullptr, nullptr};  // dZ of the current and the next layer down (f32 and bf16)
  bf16* dzb[2] = {nullptr, nullptr};
  // training set
  float *X = nullptr, *T = nullptr;
  long n = 0;
  long adam_t = 0;
  int* dstep = nullptr;  // device step counter of the running hamstr_mlp_train
  bool use_graph = true;
  hipStream_t stream = nullptr;
};

namespace {

const int kHW[3] = {oW2, oW3, oW4};
const int kHB[3] = {oB2, oB3, oB4};

int refresh_bf16(hamstr_mlp* M) {
  for (int l = 0; l < 3; l++)
    hipLaunchKernelGGL(k_w_bf16, dim3((kH * kH + 255) / 256), dim3(256), 0, M->stream, M->p + kHW[l], M->wb[l],
                       M->wtb[l]);
  MCHK(hipGetLastError());
  return HAMSTR_OK;
}

int ensure_cap(hamstr_mlp* M, int B) {
  if (M->cap >= B) return HAMSTR_OK;
  void* old[] = {M->x, M->t, M->y, M->dy, M->part, M->lpart, M->dz[0], M->dz[1], M->dzb[0], M->dzb[1],
                 M->z[0], M->z[1], M->z[2], M->z[3], M->h[0], M->h[1], M->h[2], M->h[3]};
  for (void* q : old)
    if (q) (void)hipFree(q);
  const long bh = (long)B * kH;
  MCHK(hipMalloc(&M->x, sizeof(float) * B * kIn));
  MCHK(hipMalloc(&M->t, sizeof(float) * B * kOut));
  MCHK(hipMalloc(&M->y, sizeof(float) * B * kOut));
  MCHK(hipMalloc(&M->dy, sizeof(float) * B * kOut));
  MCHK(hipMalloc(&M->part, sizeof(float) * (long)kSplit * kParams));
  MCHK(hipMalloc(&M->lpart, sizeof(float) * ((B * kOut + 255) / 256)));
  for (int k = 0; k < 2; k++) {
    MCHK(hipMalloc(&M->dz[k], sizeof(float) * bh));
    MCHK(hipMalloc(&M->dzb[k], sizeof(bf16) * bh));
  }
  for (int l = 0; l < kHidden; l++) {
    MCHK(hipMalloc(&M->z[l], sizeof(float) * bh));
    MCHK(hipMalloc(&M->h[l], sizeof(bf16) * bh));
  }
  M->cap = B;
  return HAMSTR_OK;
}

// the hidden layers on M->x (B rows, B a multiple of 32): Z1..Z4, H1..H4
int forward_hidden(hamstr_mlp* M, int B) {
  const long bh = (long)B * kH;
  hipLaunchKernelGGL(k_layer1, dim3((bh + 255) / 256), dim3(256), 0, M->stream, M->x, M->p, B, M->z[0], M->h[0]);
  for (int l = 0; l < 3; l++)
    hipLaunchKernelGGL((k_gemm_rows<0>), dim3(B / 32, 2), dim3(256), 0, M->stream, M->h[l], M->wtb[l], M->z[l + 1],
                       M->p + kHB[l], M->h[l + 1], (const float*)nullptr, (float*)nullptr, (bf16*)nullptr);
  MCHK(hipGetLastError());
  return HAMSTR_OK;
}

// forward on M->x -> M->y
int forward(hamstr_mlp* M, int B) {
  int rc = forward_hidden(M, B);
  if (rc) return rc;
  hipLaunchKernelGGL(k_out_loss, dim3((B * kOut + 255) / 256), dim3(256), 0, M->stream, M->h[3], M->p, B, M->y,
                     (const float*)nullptr, (float*)nullptr, (float*)nullptr);
  MCHK(hipGetLastError());
  return HAMSTR_OK;
}

// forward, loss and gradient (into M->g) for the batch in M->x, M->t; *loss
// on the host when asked.  Every reduction over the batch writes per-chunk
// partials into M->part, which k_reduce_all sums in chunk order.
int forward_backward(hamstr_mlp* M, int B, float* loss) {
  int rc = forward_hidden(M, B);
  if (rc) return rc;
  const long bh = (long)B * kH;
  const int nlb = (B * kOut + 255) / 256;
  const Chunks c = chunks_of(B);
  hipLaunchKernelGGL(k_out_loss, dim3(nlb), dim3(256), 0, M->stream, M->h[3], M->p, B, M->y, M->t, M->dy, M->lpart);
  // output layer: dW5, db5; dZ4 = (dY W5^T) silu'(Z4)
  hipLaunchKernelGGL(k_outer_h4, dim3(c.n), dim3(1024), 0, M->stream, M->h[3], M->dy, B, c.rows, M->part);
  hipLaunchKernelGGL(k_out_bwd_dz, dim3((bh + 255) / 256), dim3(256), 0, M->stream, M->dy, M->p, M->z[3], B,
                     M->dz[0], M->dzb[0]);
  // hidden layers 4, 3, 2 on matrix cores: dW = H^T dZ (+ db), then the next
  // layer's dZ = (dZ W^T) silu'(Z) into the other buffer
  int cur = 0;
  for (int l = 2; l >= 0; l--) {
    hipLaunchKernelGGL(k_gemm_dw, dim3(kH / 64, kH / 64, c.n), dim3(256), 0, M->stream, M->h[l], M->dzb[cur],
                       M->dz[cur], B, c.rows, M->part, kHW[l], kHB[l]);
    hipLaunchKernelGGL((k_gemm_rows<1>), dim3(B / 32, 2), dim3(256), 0, M->stream, M->dzb[cur], M->wb[l],
                       (float*)nullptr, (const float*)nullptr, (bf16*)nullptr, M->z[l], M->dz[1 - cur],
                       M->dzb[1 - cur]);
    cur = 1 - cur;
  }
  // layer 1: dW1 = X^T dZ1, db1 = sum dZ1; then the whole gradient
  hipLaunchKernelGGL(k_outer_x, dim3(c.n), dim3(1024), 0, M->stream, M->x, M->dz[cur], B, c.rows, M->part);
  hipLaunchKernelGGL(k_reduce_all, dim3((kParams + 255) / 256), dim3(256), 0, M->stream, M->part, c.n, M->g);
  MCHK(hipGetLastError());
  if (loss) {
    std::vector<float> lp(nlb);
    MCHK(hipMemcpyAsync(lp.data(), M->lpart, sizeof(float) * nlb, hipMemcpyDeviceToHost, M->stream));
    MCHK(hipStreamSynchronize(M->stream));
    double s = 0;
    for (float v : lp) s += v;
    *loss = (float)s;
  }
  return HAMSTR_OK;
}

}  // namespace

extern "C" {

const char* hamstr_mlp_last_error(void) { return m_err.c_str(); }

int hamstr_mlp_create(unsigned seed, hamstr_mlp** out) {
  if (!out) return mfail(HAMSTR_ERR_ARG, "null argument");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return mfail(HAMSTR_ERR_NODEV, "no HIP device");
  hamstr_mlp* M = new hamstr_mlp();
  // Kaiming-uniform-like initialisation from a host RNG, zero biases
  std::mt19937 rng(seed);
  std::vector<float> p(kParams, 0.0f);
  auto fill = [&](int off, int fan_in, int n) {
    std::uniform_real_distribution<float> d(-std::sqrt(6.0f / fan_in), std::sqrt(6.0f / fan_in));
    for (int i = 0; i < n; i++) p[off + i] = d(rng);
  };
  fill(oW1, kIn, kIn * kH);
  fill(oW2, kH, kH * kH);
  fill(oW3, kH, kH * kH);
  fill(oW4, kH, kH * kH);
  fill(oW5, kH, kH * kOut);
  MCHK(hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking));
  MCHK(hipMalloc(&M->p, sizeof(float) * kParams));
  MCHK(hipMalloc(&M->g, sizeof(float) * kParams));
  MCHK(hipMalloc(&M->m, sizeof(float) * kParams));
  MCHK(hipMalloc(&M->v, sizeof(float) * kParams));
  for (int l = 0; l < 3; l++) {
    MCHK(hipMalloc(&M->wb[l], sizeof(bf16) * kH * kH));
    MCHK(hipMalloc(&M->wtb[l], sizeof(bf16) * kH * kH));
  }
  MCHK(hipMemcpy(M->p, p.data(), sizeof(float) * kParams, hipMemcpyHostToDevice));
  MCHK(hipMemset(M->g, 0, sizeof(float) * kParams));
  MCHK(hipMemset(M->m, 0, sizeof(float) * kParams));
  MCHK(hipMemset(M->v, 0, sizeof(float) * kParams));
  int rc = refresh_bf16(M);
  if (rc) return rc;
  MCHK(hipStreamSynchronize(M->stream));
  *out = M;
  return HAMSTR_OK;
}

int hamstr_mlp_destroy(hamstr_mlp* M) {
  if (!M) return HAMSTR_OK;
  (void)hipStreamSynchronize(M->stream);
  void* ptrs[] = {M->p, M->g, M->m, M->v, M->wb[0], M->wb[1], M->wb[2], M->wtb[0], M->wtb[1], M->wtb[2],
                  M->x, M->t, M->y, M->dy, M->part, M->lpart, M->dz[0], M->dz[1], M->dzb[0], M->dzb[1],
                  M->z[0], M->z[1], M->z[2], M->z[3], M->h[0], M->h[1], M->h[2], M->h[3], M->X, M->T};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (M->dstep) (void)hipFree(M->dstep);
  if (M->stream) (void)hipStreamDestroy(M->stream);
  delete M;
  return HAMSTR_OK;
}

int hamstr_mlp_num_params(void) { return kParams; }

int hamstr_mlp_set_graph(hamstr_mlp* M, int on) {
  if (!M) return mfail(HAMSTR_ERR_ARG, "null argument");
  M->use_graph = on != 0;
  return HAMSTR_OK;
}

int hamstr_mlp_get_params(hamstr_mlp* M, float* host) {
  if (!M || !host) return mfail(HAMSTR_ERR_ARG, "null argument");
  MCHK(hipMemcpy(host, M->p, sizeof(float) * kParams, hipMemcpyDeviceToHost));
  return HAMSTR_OK;
}

int hamstr_mlp_set_params(hamstr_mlp* M, const float* host) {
  if (!M || !host) return mfail(HAMSTR_ERR_ARG, "null argument");
  MCHK(hipMemcpy(M->p, host, sizeof(float) * kParams, hipMemcpyHostToDevice));
  int rc = refresh_bf16(M);
  if (rc) return rc;
  MCHK(hipStreamSynchronize(M->stream));
  M->adam_t = 0;
  MCHK(hipMemset(M->m, 0, sizeof(float) * kParams));
  MCHK(hipMemset(M->v, 0, sizeof(float) * kParams));
  return HAMSTR_OK;
}

// forward pass on n rows of (normalised) features; n padded internally to 32
int hamstr_mlp_forward(hamstr_mlp* M, int n, const float* x, float* y) {
  if (!M || n <= 0 || !x || !y) return mfail(HAMSTR_ERR_ARG, "bad arguments");
  const int B = (n + 31) / 32 * 32;
  int rc = ensure_cap(M, B);
  if (rc) return rc;
  MCHK(hipMemsetAsync(M->x, 0, sizeof(float) * B * kIn, M->stream));
  MCHK(hipMemcpyAsync(M->x, x, sizeof(float) * n * kIn, hipMemcpyHostToDevice, M->stream));
  if ((rc = forward(M, B))) return rc;
  MCHK(hipMemcpyAsync(y, M->y, sizeof(float) * n * kOut, hipMemcpyDeviceToHost, M->stream));
  MCHK(hipStreamSynchronize(M->stream));
  return HAMSTR_OK;
}

// mean-squared-error loss and its gradient on n rows (n a multiple of 32)
int hamstr_mlp_loss_grad(hamstr_mlp* M, int n, const float* x, const float* t, float* loss, float* grad) {
  if (!M || n <= 0 || n % 32 || !x || !t || !loss || !grad) return mfail(HAMSTR_ERR_ARG, "bad arguments");
  int rc = ensure_cap(M, n);
  if (rc) return rc;
  MCHK(hipMemcpyAsync(M->x, x, sizeof(float) * n * kIn, hipMemcpyHostToDevice, M->stream));
  MCHK(hipMemcpyAsync(M->t, t, sizeof(float) * n * kOut, hipMemcpyHostToDevice, M->stream));
  if ((rc = forward_backward(M, n, loss))) return rc;
  MCHK(hipMemcpyAsync(grad, M->g, sizeof(float) * kParams, hipMemcpyDeviceToHost, M->stream));
  MCHK(hipStreamSynchronize(M->stream));
  return HAMSTR_OK;
}

// device-resident training set: n rows of normalised features and targets
int hamstr_mlp_set_data(hamstr_mlp* M, long n, const float* x, const float* t) {
  if (!M || n <= 0 || !x || !t) return mfail(HAMSTR_ERR_ARG, "bad arguments");
  if (M->X) (void)hipFree(M->X);
  if (M->T) (void)hipFree(M->T);
  M->X = M->T = nullptr;
  MCHK(hipMalloc(&M->X, sizeof(float) * n * kIn));
  MCHK(hipMalloc(&M->T, sizeof(float) * n * kOut));
  MCHK(hipMemcpy(M->X, x, sizeof(float) * n * kIn, hipMemcpyHostToDevice));
  MCHK(hipMemcpy(M->T, t, sizeof(float) * n * kOut, hipMemcpyHostToDevice));
  M->n = n;
  return HAMSTR_OK;
}

// Adam on random batches of the training set, learning rate decaying from lr
// to lr_final on a cosine; the batch losses stay on the device until the end.
// One training step (gather, forward, backward, Adam with the BF16 weight
// copies, step counter) is captured as a HIP graph and replayed `steps` times.
int hamstr_mlp_train(hamstr_mlp* M, int steps, int batch, float lr, float lr_final, unsigned seed, float* losses) {
  if (!M || steps < 0 || batch <= 0 || batch % 32 || M->n <= 0 || !(lr > 0) || !(lr_final >= 0))
    return mfail(HAMSTR_ERR_ARG, "bad arguments");
  if (steps == 0) return HAMSTR_OK;
  int rc = ensure_cap(M, batch);
  if (rc) return rc;
  if (!M->dstep) MCHK(hipMalloc(&M->dstep, sizeof(int)));
  float* dloss = nullptr;
  if (losses) MCHK(hipMalloc(&dloss, sizeof(float) * steps));
  MCHK(hipMemsetAsync(M->dstep, 0, sizeof(int), M->stream));
  const int nlb = (batch * kOut + 255) / 256;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const long t0 = M->adam_t;
  auto one_step = [&]() -> int {
    hipLaunchKernelGGL(k_gather, dim3((batch + 255) / 256), dim3(256), 0, M->stream, M->X, M->T, M->n, batch, seed,
                       M->dstep, M->x, M->t);
    int r;
    if ((r = forward_backward(M, batch, nullptr))) return r;
    const Bf16Copies cp{{M->wb[0], M->wb[1], M->wb[2]}, {M->wtb[0], M->wtb[1], M->wtb[2]}};
    hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, M->stream, M->p, M->g, M->m, M->v, kParams,
                       M->dstep, t0, steps, lr, lr_final, b1, b2, eps, cp);
    hipLaunchKernelGGL(k_step_end, dim3(1), dim3(64), 0, M->stream, M->lpart, nlb, dloss, M->dstep);
    MCHK(hipGetLastError());
    return HAMSTR_OK;
  };
  if (!M->use_graph) {
    for (int st = 0; st < steps; st++)
      if ((rc = one_step())) return rc;
    M->adam_t += steps;
    if (dloss) MCHK(hipMemcpyAsync(losses, dloss, sizeof(float) * steps, hipMemcpyDeviceToHost, M->stream));
    MCHK(hipStreamSynchronize(M->stream));
    if (dloss) (void)hipFree(dloss);
    return HAMSTR_OK;
  }
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  MCHK(hipStreamBeginCapture(M->stream, hipStreamCaptureModeThreadLocal));
  rc = one_step();
  const hipError_t e = hipStreamEndCapture(M->stream, &g);
  if (rc) {
    if (e == hipSuccess && g) (void)hipGraphDestroy(g);
    return rc;
  }
  MCHK(e);
  const hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  MCHK(ei);
  for (int st = 0; st < steps; st++) {
    const hipError_t el = hipGraphLaunch(ge, M->stream);
    if (el != hipSuccess) {
      (void)hipGraphExecDestroy(ge);
      return mfail(HAMSTR_ERR_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(el));
    }
  }
  M->adam_t += steps;
  if (dloss) MCHK(hipMemcpyAsync(losses, dloss, sizeof(float) * steps, hipMemcpyDeviceToHost, M->stream));
  MCHK(hipStreamSynchronize(M->stream));
  (void)hipGraphExecDestroy(ge);
  if (dloss) (void)hipFree(dloss);
  return HAMSTR_OK;
}

}  // extern "C"

// ---- warm start of a solver context ------------------------------------------
// hamstr_ctx is defined in hamstr_hip.hip; these helpers reach the state
// through the C ABI only.
extern "C" int hamstr_warm_start(hamstr_ctx* ctx, hamstr_mlp* M, const float* feat, const float* nrm,
                                 int* fallback_cells) {
  if (!ctx || !M || !feat || !nrm) return mfail(HAMSTR_ERR_ARG, "null argument");
  int ncase = 0;
  int rc = hamstr_batch_info(ctx, &ncase, nullptr);
  if (rc) return rc;
  std::vector<int> off(ncase + 1);
  hamstr_batch_info(ctx, &ncase, off.data());
  const int n = off[ncase];
  const int B = (n + 31) / 32 * 32;
  if ((rc = ensure_cap(M, B))) return rc;
  float *f = nullptr, *nr = nullptr;
  double* q = nullptr;
  int* nfb = nullptr;
  MCHK(hipMalloc(&f, sizeof(float) * (size_t)n * kIn));
  MCHK(hipMalloc(&nr, sizeof(float) * (2 * kIn + 2 * kOut)));
  MCHK(hipMalloc(&q, sizeof(double) * 4 * (size_t)n));
  MCHK(hipMalloc(&nfb, sizeof(int)));
  MCHK(hipMemcpyAsync(f, feat, sizeof(float) * (size_t)n * kIn, hipMemcpyHostToDevice, M->stream));
  MCHK(hipMemcpyAsync(nr, nrm, sizeof(float) * (2 * kIn + 2 * kOut), hipMemcpyHostToDevice, M->stream));
  MCHK(hipMemsetAsync(nfb, 0, sizeof(int), M->stream));
  MCHK(hipMemsetAsync(M->x, 0, sizeof(float) * (size_t)B * kIn, M->stream));
  hipLaunchKernelGGL(k_norm_features, dim3((n * kIn + 255) / 256), dim3(256), 0, M->stream, f, nr, n, M->x);
  if ((rc = forward(M, B))) return rc;
  // each case's freestream is its fallback
  std::vector<double> qinf(4 * (size_t)ncase);
  for (int k = 0; k < ncase; k++) {
    double s[4];
    rc = hamstr_case_freestream(ctx, k, s);
    if (rc) return rc;
    const double4 qi = make_double4(s[0], s[1], s[2], s[3]);
    const int c0 = off[k], c1 = off[k + 1];
    hipLaunchKernelGGL(k_to_state, dim3((c1 - c0 + 255) / 256), dim3(256), 0, M->stream, M->y + (size_t)c0 * kOut,
                       nr, c1 - c0, qi, q + 4 * (size_t)c0, nfb);
  }
  MCHK(hipGetLastError());
  std::vector<double> hq(4 * (size_t)n);
  int hfb = 0;
  MCHK(hipMemcpyAsync(hq.data(), q, sizeof(double) * hq.size(), hipMemcpyDeviceToHost, M->stream));
  MCHK(hipMemcpyAsync(&hfb, nfb, sizeof(int), hipMemcpyDeviceToHost, M->stream));
  MCHK(hipStreamSynchronize(M->stream));
  (void)hipFree(f);
  (void)hipFree(nr);
  (void)hipFree(q);
  (void)hipFree(nfb);
  if (fallback_cells) *fallback_cells = hfb;
  return hamstr_set_field(ctx, HAMSTR_FIELD_Q, hq.data());
}

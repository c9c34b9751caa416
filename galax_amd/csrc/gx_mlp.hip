// gx_mlp.hip -- BF16 MFMA kernels for the learned acceleration surrogate.
//
// The surrogate is an MLP (5 inputs: rotating-frame x, y, z, R, r; four hidden
// layers of 256 with SiLU; 3 outputs) standing in for the expensive part of a
// composite field (a rotating bar plus dozens of co-rotating subhaloes).
// Weights and activations are BF16; every product accumulates in FP32 on the
// matrix cores (v_mfma_f32_16x16x32_bf16).
//
// One GEMM kernel serves forward and backward passes:
//   C[m, n] = sum_k A(m, k) B(k, n),  A(m,k) = A[m*sam + k*sak],
//                                     B(k,n) = B[k*sbk + n*sbn]   (BF16)
// with a fused epilogue (bias + SiLU storing the pre-activation, bias only,
// SiLU backward, or a plain FP32 store).  A 256-thread block computes a 64x64
// tile; each wave a 32x32 quarter as 2x2 MFMA tiles; K advances in slabs of
// 32 staged through LDS (16-byte loads when the reduction axis is
// contiguous).
//
// MFMA operand layout (16x16x32 BF16, wave64): lane l supplies
// A[l%16][8(l/16) .. 8(l/16)+7] and B[8(l/16) .. +7][l%16]; it receives
// D[4(l/16) + j][l%16], j = 0..3.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "gx_internal.hpp"
#include "gx_potential.hpp"
#include "gx_tableau.hpp"

namespace gx {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
typedef __attribute__((__vector_size__(4 * sizeof(unsigned)))) unsigned u32x4;

// x, or zero where keep is false, by a bit mask: loads feeding it stay
// unconditional (a select between a loaded value and zero compiled to a
// branch around the load and a wait after it).
__device__ __forceinline__ float keep_or_zero(float x, bool keep) {
  return __builtin_bit_cast(float, __builtin_bit_cast(unsigned, x) & (keep ? ~0u : 0u));
}

constexpr int BM = 64, BN = 64, BK = 32;
constexpr int LDSK = BK + 8;  // padded row (bf16 elements) against bank conflicts

enum : int {
  EPI_STORE = GX_EPI_STORE,
  EPI_BIAS = GX_EPI_BIAS,
  EPI_BIAS_SILU = GX_EPI_BIAS_SILU,
  EPI_DSILU = GX_EPI_DSILU
};

struct GemmArgs {
  const __bf16* A;
  long long sam, sak;
  const __bf16* B;
  long long sbk, sbn;
  long long M, N, K;
  int epi;
  int accumulate;       // EPI_STORE: add into C (atomically when split over K)
  int ksplit;           // blocks along K (grid z); > 1 only with EPI_STORE
  const float* bias;    // [N]
  float* C;             // fp32 output (EPI_STORE / EPI_BIAS); Z for BIAS_SILU
  long long ldc;
  __bf16* H;            // bf16 output (BIAS_SILU: silu(Z); DSILU: dZ)
  long long ldh;
  const float* Zin;     // DSILU: pre-activation
  long long ldzin;
};

// SiLU and its derivative in FP32 with the hardware reciprocal (1 ulp): the
// results are rounded to BF16 (8 bits), and a correctly rounded division
// took ten instructions per activation.  exp(-z) = +inf for very negative
// z gives 1/inf = 0, so silu -> -0 and dsilu -> 0 there, as in the limit.
__device__ __forceinline__ float sigmoid(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = sigmoid(z);
  return s * (1.0f + z * (1.0f - s));
}

// Stages a 64 x 32 slab of a strided BF16 operand X(r, k) = X[r*sr + k*sk]
// (rows r0.., reduction index k0..) into LDS as T[r][k], zero-padding out of
// range entries.  Three load shapes: 16-byte loads along k (sk == 1), 16-byte
// loads along r (sr == 1, e.g. the transposed operands of the weight
// gradient), or scalar loads.
template <bool VEC>
__device__ __forceinline__ void stage_tile(__bf16 (*T)[LDSK], const __bf16* X,
                                           long long sr, long long sk,
                                           long long r0, long long R,
                                           long long k0, long long K, int tid) {
  if (VEC && sk == 1) {
    const int r = tid >> 2, kc = (tid & 3) * 8;
    const long long rr = r0 + r, k = k0 + kc;
    const __bf16* src = X + rr * sr + k;
    if (rr < R && k + 8 <= K && ((reinterpret_cast<uintptr_t>(src) & 15) == 0)) {
      *reinterpret_cast<bf16x8*>(&T[r][kc]) = *reinterpret_cast<const bf16x8*>(src);
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j)
        T[r][kc + j] = (rr < R && k + j < K) ? X[rr * sr + k + j] : (__bf16)0.0f;
    }
  } else if (VEC && sr == 1) {
    const int kq = tid >> 3, rc = (tid & 7) * 8;
    const long long k = k0 + kq, rr = r0 + rc;
    const __bf16* src = X + k * sk + rr;
    if (k < K && rr + 8 <= R && ((reinterpret_cast<uintptr_t>(src) & 15) == 0)) {
      const bf16x8 v = *reinterpret_cast<const bf16x8*>(src);
#pragma unroll
      for (int j = 0; j < 8; ++j) T[rc + j][kq] = v[j];
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j)
        T[rc + j][kq] = (k < K && rr + j < R) ? X[k * sk + rr + j] : (__bf16)0.0f;
    }
  } else {
    const int r = tid >> 2, kc = (tid & 3) * 8;
    const long long rr = r0 + r;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const long long k = k0 + kc + j;
      T[r][kc + j] = (rr < R && k < K) ? X[rr * sr + k * sk] : (__bf16)0.0f;
    }
  }
}

template <bool VEC>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
  __shared__ __bf16 As[BM][LDSK];
  __shared__ __bf16 Bs[BN][LDSK];  // B transposed: [n][k]
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  // m tiles on x (up to 2^31 - 1 blocks), n tiles on y
  const long long m0 = (long long)blockIdx.x * BM, n0 = (long long)blockIdx.y * BN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

  // this block's share of the reduction (split-K for the weight gradients,
  // whose reduction runs over the whole batch)
  const long long kper = ((g.K + g.ksplit - 1) / g.ksplit + BK - 1) / BK * BK;
  const long long kbeg = (long long)blockIdx.z * kper;
  const long long kend = kbeg + kper < g.K ? kbeg + kper : g.K;
  for (long long k0 = kbeg; k0 < kend; k0 += BK) {
    // ---- stage A (64 x 32) and B^T (64 x 32) into LDS, [row][k] order
    stage_tile<VEC>(As, g.A, g.sam, g.sak, m0, g.M, k0, kend, tid);
    stage_tile<VEC>(Bs, g.B, g.sbn, g.sbk, n0, g.N, k0, kend, tid);
    __syncthreads();
    // ---- 2x2 MFMA tiles per wave
    const int kk = (lane >> 4) * 8;
    bf16x8 af[2], bfr[2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
      af[a] = *reinterpret_cast<const bf16x8*>(&As[wm + a * 16 + (lane & 15)][kk]);
#pragma unroll
    for (int b = 0; b < 2; ++b)
      bfr[b] = *reinterpret_cast<const bf16x8*>(&Bs[wn + b * 16 + (lane & 15)][kk]);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
    __syncthreads();
  }
  // ---- epilogue
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const long long m = m0 + wm + a * 16 + (lane >> 4) * 4 + j;
        const long long n = n0 + wn + b * 16 + (lane & 15);
        if (m >= g.M || n >= g.N) continue;
        float v = acc[a][b][j];
        switch (g.epi) {
          case EPI_STORE:
            if (g.ksplit > 1) {
              atomicAdd(g.C + m * g.ldc + n, v);
            } else {
              if (g.accumulate) v += g.C[m * g.ldc + n];
              g.C[m * g.ldc + n] = v;
            }
            break;
          case EPI_BIAS:
            g.C[m * g.ldc + n] = v + g.bias[n];
            break;
          case EPI_BIAS_SILU: {
            const float z = v + g.bias[n];
            if (g.C) g.C[m * g.ldc + n] = z;
            g.H[m * g.ldh + n] = (__bf16)silu(z);
            break;
          }
          default: {  // EPI_DSILU
            const float z = g.Zin[m * g.ldzin + n];
            g.H[m * g.ldh + n] = (__bf16)(v * dsilu(z));
            break;
          }
        }
      }
}

// The whole inference pass of an MLP of hidden width 256 in one kernel: a
// block takes 64 points and keeps their activations in LDS (two 64 x 256
// BF16 buffers, used in turn), so that activations never go to HBM; every
// hidden-to-hidden layer runs on the matrix cores with the same operand
// layout as k_gemm, the layer's BF16 weights read from L2 (each layer's
// 128 KB is shared by every block).  Each wave owns 64 output columns of a
// layer: 4 x 4 MFMA tiles.  The input layer (8 padded inputs) is a short
// dot product in FP32 over the block's inputs staged in LDS; the output
// layer (at most 16 columns, zero-padded) is one MFMA tile column per wave.
// Rounding as the layer-by-layer path: FP32 accumulation, bias + SiLU in
// FP32, BF16 activations.
constexpr int FH = 256;      // hidden width
constexpr int FM = 64;       // points per block
constexpr int FA = FM / 16;  // MFMA row tiles per block
constexpr int FLD = FH + 8;  // padded LDS row (bf16 elements)
constexpr int kMaxFusedLayers = 8;

struct FusedArgs {
  const __bf16* X;  // [n][ldx] inputs (BF16)
  int ldx, nin;     // row stride, inputs used
  long long n;
  int depth;        // hidden layers
  int nout;
  const __bf16* W[kMaxFusedLayers + 1];  // layer l: [out][in] row-major
  const float* b[kMaxFusedLayers + 1];
  float* Y;         // [n][ldy] FP32 outputs
  int ldy;
};

__global__ void __launch_bounds__(256) k_mlp_fused(FusedArgs f) {
  __shared__ __bf16 act[2][FM][FLD];
  __shared__ float xin[FM][8];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const long long m0 = (long long)blockIdx.x * FM;
  // ---- the block's inputs, once (zero past n and past nin)
  static_assert(FM * 8 % 256 == 0, "input staging: whole rounds of the block");
#pragma unroll
  for (int j = 0; j < FM * 8 / 256; ++j) {
    const int t = tid + 256 * j;
    const int r = t >> 3, k = t & 7;
    const long long m = m0 + r;
    const bool in = m < f.n && k < f.nin;  // else an in-range stand-in, masked
    xin[r][k] = keep_or_zero((float)f.X[(m < f.n ? m : f.n - 1) * f.ldx + (k < f.nin ? k : 0)], in);
  }
  __syncthreads();
  // ---- input layer: thread tid owns hidden unit tid for all 64 points
  {
    const int c = tid;
    float w[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
      w[k] = keep_or_zero((float)f.W[0][c * f.nin + (k < f.nin ? k : 0)], k < f.nin);
    const float bias = f.b[0][c];
    for (int r = 0; r < FM; ++r) {
      float z = 0.0f;
#pragma unroll
      for (int k = 0; k < 8; ++k) z = fmaf(xin[r][k], w[k], z);
      act[0][r][c] = (__bf16)silu(z + bias);
    }
  }
  int cur = 0;
  // ---- hidden-to-hidden layers on the matrix cores
  for (int l = 1; l < f.depth; ++l) {
    __syncthreads();
    const __bf16* W = f.W[l];
    f32x4 acc[FA][4];
#pragma unroll
    for (int a = 0; a < FA; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int k0 = 0; k0 < FH; k0 += 32) {
      const int kk = k0 + (lane >> 4) * 8;
      bf16x8 af[FA], bw[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int n = wave * 64 + b * 16 + (lane & 15);
        bw[b] = *reinterpret_cast<const bf16x8*>(W + n * FH + kk);
      }
#pragma unroll
      for (int a = 0; a < FA; ++a)
        af[a] = *reinterpret_cast<const bf16x8*>(&act[cur][a * 16 + (lane & 15)][kk]);
#pragma unroll
      for (int a = 0; a < FA; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bw[b], acc[a][b], 0, 0, 0);
    }
    const float* bias = f.b[l];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int n = wave * 64 + b * 16 + (lane & 15);
      const float bn = bias[n];
#pragma unroll
      for (int a = 0; a < FA; ++a)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          act[cur ^ 1][a * 16 + (lane >> 4) * 4 + j][n] = (__bf16)silu(acc[a][b][j] + bn);
    }
    cur ^= 1;
  }
  __syncthreads();
  // ---- output layer on the matrix cores: wave w takes the 16-point row
  // tiles w, w + 4, ... against the outputs padded to 16 columns (zero
  // weights past nout)
  {
    const __bf16* Wo = f.W[f.depth];
    const int n = lane & 15;
    // all eight weight fragments in flight at once (row 0 stands in for the
    // padding columns, whose fragments are then masked to zero)
    const int nr = n < f.nout ? n : 0;
    const unsigned mask = n < f.nout ? ~0u : 0u;
    bf16x8 bw[FH / 32];
#pragma unroll
    for (int s = 0; s < FH / 32; ++s) {
      const u32x4 v = *reinterpret_cast<const u32x4*>(Wo + nr * FH + s * 32 + (lane >> 4) * 8);
      bw[s] = __builtin_bit_cast(bf16x8, v & mask);
    }
    // wave w takes row tiles w, w + 4, ...
#pragma unroll
    for (int a = wave; a < FA; a += 4) {
      f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < FH / 32; ++s) {
        const int kk = s * 32 + (lane >> 4) * 8;
        const bf16x8 af = *reinterpret_cast<const bf16x8*>(&act[cur][a * 16 + n][kk]);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bw[s], acc, 0, 0, 0);
      }
      if (n < f.nout) {
        const float bo = f.b[f.depth][n];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const long long m = m0 + a * 16 + (lane >> 4) * 4 + j;
          if (m < f.n) f.Y[m * f.ldy + n] = acc[j] + bo;
        }
      }
    }
  }
}

// db[n] += sum_m dZ[m, n] (bf16 in, fp32 out) -- partial column sums per
// block of kColsumRows rows, combined with atomics.  A lane keeps four
// independent sums so that four rows' loads are in flight at once; the
// grid has one block row per chunk of rows, enough to fill the device for
// a training batch (65,536 rows: 512 chunks).
constexpr int kColsumRows = 128;
__global__ void k_colsum(const __bf16* __restrict__ dz, long long ld, long long M,
                         long long N, float* __restrict__ out) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const long long rows = (M + gridDim.y - 1) / gridDim.y;
  const long long r0 = (long long)blockIdx.y * rows;
  const long long r1 = r0 + rows < M ? r0 + rows : M;
  float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
  long long m = r0;
  for (; m + 3 < r1; m += 4) {
    s0 += (float)dz[m * ld + n];
    s1 += (float)dz[(m + 1) * ld + n];
    s2 += (float)dz[(m + 2) * ld + n];
    s3 += (float)dz[(m + 3) * ld + n];
  }
  for (; m < r1; ++m) s0 += (float)dz[m * ld + n];
  atomicAdd(out + n, (s0 + s1) + (s2 + s3));
}

// Adam on fp32 master parameters (bias-corrected, as Kingma & Ba 2015);
// writes the bf16 copy the GEMMs read.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v,
                       __bf16* __restrict__ pb, long long n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  p[i] = pi;
  if (pb) pb[i] = (__bf16)pi;
}

// Pseudo-Huber loss on the network's outputs: r = Y - T (first 3 columns),
// loss = delta^2 (sqrt(1 + (r/delta)^2) - 1) summed, gradient
// r / sqrt(1 + (r/delta)^2) times `scale` into dY (bf16, other columns 0).
// The bounded gradient keeps rare large targets (points close to a
// subhalo's core) from dominating training.
__global__ void k_huber(const float* __restrict__ Y, long long ldy,
                        const float* __restrict__ T, long long B, float delta,
                        float scale, __bf16* __restrict__ dY, long long ldd,
                        double* __restrict__ loss) {
  __shared__ double part[256];
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (i < B) {
    const float id = 1.0f / delta;
    for (int c = 0; c < 3; ++c) {
      const float r = Y[i * ldy + c] - T[i * 3 + c];
      const float q = sqrtf(1.0f + (r * id) * (r * id));
      l += (double)(delta * delta * (q - 1.0f));
      dY[i * ldd + c] = (__bf16)(scale * r / q);
    }
    for (long long c = 3; c < ldd; ++c) dY[i * ldd + c] = (__bf16)0.0f;
  }
  part[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, part[0]);
}

// ------------------------------------------------ surrogate field plumbing
// The perturbing part of the field (a bar and subhaloes rotating rigidly at
// omega) is static in the frame rotating with angle th = omega t, so the
// network sees rotating-frame positions and predicts the rotating-frame
// acceleration:  x' = Rz(-th) x,  a = Rz(th) a'.

__device__ __forceinline__ double stage_time(const double* t, int t_stride,
                                             long long i, double dt) {
  return t[t_stride ? i : 0] + dt;
}

// X[i] = (x', y', z, R, r) / L as bf16, zero-padded to width ldx.
__global__ void k_sur_inputs(const double* __restrict__ xyz,
                             const double* __restrict__ t, int t_stride,
                             double dt, long long n, double inv_len, double omega,
                             __bf16* __restrict__ X, long long ldx) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s, c;
  sincos(omega * stage_time(t, t_stride, i, dt), &s, &c);
  const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
  const double R2 = x * x + y * y;
  __bf16* row = X + i * ldx;
  row[0] = (__bf16)(float)((c * x + s * y) * inv_len);
  row[1] = (__bf16)(float)((c * y - s * x) * inv_len);
  row[2] = (__bf16)(float)(z * inv_len);
  row[3] = (__bf16)(float)(sqrt(R2) * inv_len);
  row[4] = (__bf16)(float)(sqrt(R2 + z * z) * inv_len);
  for (long long k = 5; k < ldx; ++k) row[k] = (__bf16)0.0f;
}

// Training targets: T[i] = Rz(-th) (a_full - a_base)(x_i, t_i) / scale.
__global__ void k_sur_targets(const double* __restrict__ t, int t_stride,
                              long long n, const double* __restrict__ a_full,
                              const double* __restrict__ a_base, double omega,
                              double inv_scale, float* __restrict__ T) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s, c;
  sincos(omega * t[t_stride ? i : 0], &s, &c);
  const double dx = a_full[3 * i] - a_base[3 * i];
  const double dy = a_full[3 * i + 1] - a_base[3 * i + 1];
  const double dz = a_full[3 * i + 2] - a_base[3 * i + 2];
  T[3 * i] = (float)((c * dx + s * dy) * inv_scale);
  T[3 * i + 1] = (float)((c * dy - s * dx) * inv_scale);
  T[3 * i + 2] = (float)(dz * inv_scale);
}

// The surrogate field: a = a_base(x) (exact FP64) + scale Rz(th) Y[i], or the
// exact full field where the gate says so (outside the trained ball
// |x| > r_max, or an orbit flagged by the periodic check).  EB and EF
// evaluate the base and full fields: the compile-time evaluators when the
// tables have their signatures (the Milky Way model; it plus a rotating
// group), else the table loop.
template <class EB, class EF>
__global__ void k_sur_combine(DPot base, DPot full, const double* __restrict__ xyz,
                              const double* __restrict__ t, int t_stride, double dt,
                              long long n, const float* __restrict__ Y,
                              long long ldy, double scale, double omega,
                              double r_max, const int* __restrict__ flag,
                              double* __restrict__ acc) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
  const double ti = stage_time(t, t_stride, i, dt);
  const bool exact = (x * x + y * y + z * z > r_max * r_max) || (flag && flag[i]);
  double ax, ay, az;
  if (exact) {
    EF::acc(full, x, y, z, ti, ax, ay, az);
  } else {
    EB::acc(base, x, y, z, ti, ax, ay, az);
    double s, c;
    sincos(omega * ti, &s, &c);
    const double px = scale * (double)Y[i * ldy], py = scale * (double)Y[i * ldy + 1];
    ax += c * px - s * py;
    ay += s * px + c * py;
    az += scale * (double)Y[i * ldy + 2];
  }
  acc[3 * i] = ax;
  acc[3 * i + 1] = ay;
  acc[3 * i + 2] = az;
}

// The periodic gate check: the surrogate field at the current positions
// against the exact field; an orbit whose relative error exceeds `tol` is
// flagged (sticky) and integrated with the exact field from then on.
template <class EF>
__global__ void k_sur_check(DPot full, const double* __restrict__ xyz,
                            const double* __restrict__ t, int t_stride, long long n,
                            const double* __restrict__ a_sur, double tol,
                            int* __restrict__ flag, float* __restrict__ err) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double ax, ay, az;
  EF::acc(full, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], t[t_stride ? i : 0], ax, ay, az);
  const double ex = a_sur[3 * i] - ax, ey = a_sur[3 * i + 1] - ay, ez = a_sur[3 * i + 2] - az;
  const double e = sqrt((ex * ex + ey * ey + ez * ez) / (ax * ax + ay * ay + az * az));
  if (err) err[i] = (float)e;
  if (!(e <= tol)) flag[i] = 1;
}

// Lock-step fixed-step Dopri8 (Runge-Kutta-Nystrom form, as the persistent
// integrator) over n orbits sharing one time t[0]; the stage accelerations
// live in HBM, G = [13][n][3], so any batched field (the surrogate) can
// produce them.  Stage positions Q_i = q + h (rs_i p + h sum_l (A^2)_il g_l):
__global__ void k_stage_point(int stage, const double* __restrict__ q,
                              const double* __restrict__ p,
                              const double* __restrict__ G, long long n, double h,
                              double* __restrict__ Q) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr auto& K = Dopri8::k;
  double s[3] = {0.0, 0.0, 0.0};
  for (int l = 0; l < stage; ++l) {
    const double a = K.aa[stage][l];
    if (a == 0.0) continue;
    for (int m = 0; m < 3; ++m) s[m] = fma(a, G[((long long)l * n + i) * 3 + m], s[m]);
  }
  const double rs = K.rs[stage];
  for (int m = 0; m < 3; ++m)
    Q[3 * i + m] = fma(h, fma(rs, p[3 * i + m], h * s[m]), q[3 * i + m]);
}

// Step update q += h (sb p + h sum bA_l g_l), p += h sum b_l g_l; thread 0
// also advances the shared time t[0] += h.
__global__ void k_step_update(double* __restrict__ q, double* __restrict__ p,
                              const double* __restrict__ G, long long n, double h,
                              double* __restrict__ t) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) t[0] += h;
  if (i >= n) return;
  constexpr auto& K = Dopri8::k;
  double b[3] = {0.0, 0.0, 0.0}, c[3] = {0.0, 0.0, 0.0};
  for (int l = 0; l < 13; ++l) {
    for (int m = 0; m < 3; ++m) {
      const double gl = G[((long long)l * n + i) * 3 + m];
      if (K.b[l] != 0.0) b[m] = fma(K.b[l], gl, b[m]);
      if (K.ba[l] != 0.0) c[m] = fma(K.ba[l], gl, c[m]);
    }
  }
  for (int m = 0; m < 3; ++m) {
    const double pm = p[3 * i + m];
    q[3 * i + m] = fma(h, fma(K.sb, pm, h * c[m]), q[3 * i + m]);
    p[3 * i + m] = fma(h, b[m], pm);
  }
}

long long blocks_for(long long n, int bs) { return (n + bs - 1) / bs; }

bool grid_ok(long long n) { return blocks_for(n, 256) <= 0x7fffffffLL; }

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" {

int gx_mlp_gemm(const void* A, int64_t sam, int64_t sak, const void* B,
                int64_t sbk, int64_t sbn, int64_t M, int64_t N, int64_t K,
                int32_t epi, int32_t accumulate, int32_t ksplit, const float* bias,
                float* C, int64_t ldc, void* H, int64_t ldh, const float* Zin,
                int64_t ldzin, void* stream) {
  if (M < 0 || N < 0 || K < 0 || !A || !B) return GX_EINVAL;
  if (M == 0 || N == 0) return GX_OK;
  if (epi < EPI_STORE || epi > EPI_DSILU) return GX_EINVAL;
  if ((epi == EPI_STORE || epi == EPI_BIAS) && !C) return GX_EINVAL;
  if ((epi == EPI_BIAS || epi == EPI_BIAS_SILU) && !bias) return GX_EINVAL;
  if ((epi == EPI_BIAS_SILU || epi == EPI_DSILU) && !H) return GX_EINVAL;
  if (epi == EPI_DSILU && !Zin) return GX_EINVAL;
  if (ksplit < 1 || ksplit > 65535 || (ksplit > 1 && epi != EPI_STORE)) return GX_EINVAL;
  GemmArgs g{};
  g.A = static_cast<const __bf16*>(A);
  g.sam = sam;
  g.sak = sak;
  g.B = static_cast<const __bf16*>(B);
  g.sbk = sbk;
  g.sbn = sbn;
  g.M = M;
  g.N = N;
  g.K = K;
  g.epi = epi;
  g.accumulate = accumulate;
  g.ksplit = ksplit;
  g.bias = bias;
  g.C = C;
  g.ldc = ldc;
  g.H = static_cast<__bf16*>(H);
  g.ldh = ldh;
  g.Zin = Zin;
  g.ldzin = ldzin;
  const long long mt = (M + BM - 1) / BM, nt = (N + BN - 1) / BN;
  if (mt > 0x7fffffffLL || nt > 65535) return GX_EINVAL;
  dim3 grid((unsigned)mt, (unsigned)nt, (unsigned)ksplit);
  hipLaunchKernelGGL(k_gemm<true>, grid, dim3(256), 0, (hipStream_t)stream, g);
  return check(hipGetLastError());
}

int gx_mlp_forward_fused(const void* X, int32_t ldx, int32_t nin, int64_t n,
                         int32_t depth, const void* const* W,
                         const float* const* b, int32_t nout, float* Y,
                         int32_t ldy, void* stream) {
  if (n < 0 || depth < 1 || depth > kMaxFusedLayers || nin < 1 || nin > 8 ||
      ldx < nin || nout < 1 || nout > 16 || ldy < nout)
    return GX_EINVAL;
  if (n == 0) return GX_OK;
  if (!X || !W || !b || !Y) return GX_EINVAL;
  FusedArgs f{};
  f.X = static_cast<const __bf16*>(X);
  f.ldx = ldx;
  f.nin = nin;
  f.n = n;
  f.depth = depth;
  f.nout = nout;
  f.Y = Y;
  f.ldy = ldy;
  for (int l = 0; l <= depth; ++l) {
    if (!W[l] || !b[l]) return GX_EINVAL;
    // 16-byte fragment loads of the hidden and output layers' rows
    if (l > 0 && (reinterpret_cast<uintptr_t>(W[l]) & 15)) return GX_EINVAL;
    f.W[l] = static_cast<const __bf16*>(W[l]);
    f.b[l] = b[l];
  }
  const long long blocks = (n + FM - 1) / FM;
  if (blocks > 0x7fffffffLL) return GX_EINVAL;
  hipLaunchKernelGGL(k_mlp_fused, dim3((unsigned)blocks), dim3(256), 0,
                     (hipStream_t)stream, f);
  return check(hipGetLastError());
}

int gx_mlp_colsum(const void* dz, int64_t ld, int64_t M, int64_t N, float* out,
                  void* stream) {
  if (!dz || !out || M < 0 || N < 0 || ld < N) return GX_EINVAL;
  if (M == 0 || N == 0) return GX_OK;
  long long ychunks = (M + kColsumRows - 1) / kColsumRows;
  if (ychunks > 65535) ychunks = 65535;  // grid y limit; chunks then grow
  dim3 grid((unsigned)blocks_for(N, 256), (unsigned)ychunks);
  hipLaunchKernelGGL(k_colsum, grid, dim3(256), 0, (hipStream_t)stream,
                     static_cast<const __bf16*>(dz), (long long)ld, (long long)M,
                     (long long)N, out);
  return check(hipGetLastError());
}

int gx_mlp_adam(float* p, const float* g, float* m, float* v, void* pb,
                int64_t n, float lr, float b1, float b2, float eps, int64_t step,
                void* stream) {
  if (!p || !g || !m || !v || n < 0 || step < 1) return GX_EINVAL;
  if (n == 0) return GX_OK;
  if (!grid_ok(n)) return GX_EINVAL;
  const float c1 = 1.0f / (1.0f - powf(b1, (float)step));
  const float c2 = 1.0f / (1.0f - powf(b2, (float)step));
  hipLaunchKernelGGL(k_adam, dim3((unsigned)blocks_for(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, p, g, m, v, static_cast<__bf16*>(pb),
                     (long long)n, lr, b1, b2, eps, c1, c2);
  return check(hipGetLastError());
}

int gx_mlp_huber(const float* Y, int64_t ldy, const float* T, int64_t B, float delta,
                 float scale, void* dY, int64_t ldd, double* loss, void* stream) {
  if (!Y || !T || !dY || !loss || B < 0 || ldy < 3 || ldd < 3 || !(delta > 0.f))
    return GX_EINVAL;
  if (B == 0) return GX_OK;
  if (!grid_ok(B)) return GX_EINVAL;
  hipLaunchKernelGGL(k_huber, dim3((unsigned)blocks_for(B, 256)), dim3(256), 0,
                     (hipStream_t)stream, Y, (long long)ldy, T, (long long)B, delta,
                     scale, static_cast<__bf16*>(dY), (long long)ldd, loss);
  return check(hipGetLastError());
}

int gx_sur_inputs(const double* xyz, const double* t, int32_t t_stride, double dt,
                  int64_t n, double inv_len, double omega, void* X, int64_t ldx,
                  void* stream) {
  if (!xyz || !t || !X || n < 0 || ldx < 5) return GX_EINVAL;
  if (n == 0) return GX_OK;
  if (!grid_ok(n)) return GX_EINVAL;
  hipLaunchKernelGGL(k_sur_inputs, dim3((unsigned)blocks_for(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, xyz, t, t_stride, dt, (long long)n, inv_len,
                     omega, static_cast<__bf16*>(X), (long long)ldx);
  return check(hipGetLastError());
}

int gx_sur_targets(const double* t, int32_t t_stride, int64_t n, const double* a_full,
                   const double* a_base, double omega, double inv_scale, float* T,
                   void* stream) {
  if (!t || !a_full || !a_base || !T || n < 0) return GX_EINVAL;
  if (n == 0) return GX_OK;
  if (!grid_ok(n)) return GX_EINVAL;
  hipLaunchKernelGGL(k_sur_targets, dim3((unsigned)blocks_for(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, t, t_stride, (long long)n, a_full, a_base, omega,
                     inv_scale, T);
  return check(hipGetLastError());
}

int gx_sur_combine(const gx_pot* base, const gx_pot* full, const double* xyz,
                   const double* t, int32_t t_stride, double dt, int64_t n,
                   const float* Y, int64_t ldy, double scale, double omega,
                   double r_max, const int32_t* flag, double* acc, void* stream) {
  if (!base || !full || !xyz || !t || !Y || !acc || n < 0 || ldy < 3) return GX_EINVAL;
  if (!base->basic || base->device != full->device) return GX_EINVAL;
  if (n == 0) return GX_OK;
  if (!grid_ok(n)) return GX_EINVAL;
  const dim3 grid((unsigned)blocks_for(n, 256)), blk(256);
  const bool mw = base->sig == kSigMilkyWay, rot = full->sig == kSigMilkyWayRotating;
  const auto kern = mw ? (rot ? k_sur_combine<EvalMW, EvalMilkyWayRotating>
                              : k_sur_combine<EvalMW, EvalTable<true>>)
                       : (rot ? k_sur_combine<EvalTable<false>, EvalMilkyWayRotating>
                              : k_sur_combine<EvalTable<false>, EvalTable<true>>);
  hipLaunchKernelGGL(kern, grid, blk, 0, (hipStream_t)stream, base->dev, full->dev, xyz, t,
                     t_stride, dt, (long long)n, Y, (long long)ldy, scale, omega, r_max, flag,
                     acc);
  return check(hipGetLastError());
}

int gx_sur_check(const gx_pot* full, const double* xyz, const double* t,
                 int32_t t_stride, int64_t n, const double* a_sur, double tol,
                 int32_t* flag, float* err, void* stream) {
  if (!full || !xyz || !t || !a_sur || !flag || n < 0) return GX_EINVAL;
  if (n == 0) return GX_OK;
  if (!grid_ok(n)) return GX_EINVAL;
  const auto kern = full->sig == kSigMilkyWayRotating ? k_sur_check<EvalMilkyWayRotating>
                                                      : k_sur_check<EvalTable<true>>;
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks_for(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, full->dev, xyz, t, t_stride, (long long)n, a_sur,
                     tol, flag, err);
  return check(hipGetLastError());
}

int gx_method_nodes(int32_t method, double* c, int32_t cap) {
  if (method == GX_DOPRI8) {
    if (c)
      for (int i = 0; i < 13 && i < cap; ++i) c[i] = Dopri8::k.c[i];
    return 13;
  }
  if (method == GX_DOPRI5) {
    if (c)
      for (int i = 0; i < 7 && i < cap; ++i) c[i] = Dopri5::k.c[i];
    return 7;
  }
  return GX_EINVAL;
}

int gx_lockstep_stage(int32_t stage, const double* q, const double* p,
                      const double* G, int64_t n, double h, double* Q, void* stream) {
  if (stage < 0 || stage >= 13 || !q || !p || !G || !Q || n < 0) return GX_EINVAL;
  if (n == 0) return GX_OK;
  if (!grid_ok(n)) return GX_EINVAL;
  hipLaunchKernelGGL(k_stage_point, dim3((unsigned)blocks_for(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, (int)stage, q, p, G, (long long)n, h, Q);
  return check(hipGetLastError());
}

int gx_lockstep_update(double* q, double* p, const double* G, int64_t n, double h,
                       double* t, void* stream) {
  if (!q || !p || !G || !t || n < 0) return GX_EINVAL;
  if (!grid_ok(n)) return GX_EINVAL;
  hipLaunchKernelGGL(k_step_update, dim3((unsigned)(n > 0 ? blocks_for(n, 256) : 1)),
                     dim3(256), 0, (hipStream_t)stream, q, p, G, (long long)n, h, t);
  return check(hipGetLastError());
}

}  // extern "C"

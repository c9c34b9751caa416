// gx_mlp.hip -- BF16 MFMA kernels for the learned acceleration surrogate.
//
// The surrogate is an MLP (5 inputs: x, y, z, cos wt, sin wt; four hidden
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

constexpr int BM = 64, BN = 64, BK = 32;
constexpr int LDSK = BK + 8;  // padded row (bf16 elements) against bank conflicts

enum : int { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_SILU = 2, EPI_DSILU = 3 };

struct GemmArgs {
  const __bf16* A;
  long long sam, sak;
  const __bf16* B;
  long long sbk, sbn;
  long long M, N, K;
  int epi;
  int accumulate;       // EPI_STORE: add into C
  const float* bias;    // [N]
  float* C;             // fp32 output (EPI_STORE / EPI_BIAS); Z for BIAS_SILU
  long long ldc;
  __bf16* H;            // bf16 output (BIAS_SILU: silu(Z); DSILU: dZ)
  long long ldh;
  const float* Zin;     // DSILU: pre-activation
  long long ldzin;
};

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

template <bool VEC>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
  __shared__ __bf16 As[BM][LDSK];
  __shared__ __bf16 Bs[BN][LDSK];  // B transposed: [n][k]
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const long long m0 = (long long)blockIdx.y * BM, n0 = (long long)blockIdx.x * BN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (long long k0 = 0; k0 < g.K; k0 += BK) {
    // ---- stage A (64 x 32) and B^T (64 x 32): 8 elements per thread each
    {
      const int r = tid >> 2, kc = (tid & 3) * 8;
      const long long m = m0 + r, n = n0 + r;
      if (VEC && g.sak == 1 && m < g.M && k0 + kc + 8 <= g.K &&
          ((reinterpret_cast<uintptr_t>(g.A + m * g.sam + k0 + kc) & 15) == 0)) {
        *reinterpret_cast<bf16x8*>(&As[r][kc]) =
            *reinterpret_cast<const bf16x8*>(g.A + m * g.sam + k0 + kc);
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const long long k = k0 + kc + j;
          As[r][kc + j] = (m < g.M && k < g.K) ? g.A[m * g.sam + k * g.sak] : (__bf16)0.0f;
        }
      }
      if (VEC && g.sbk == 1 && n < g.N && k0 + kc + 8 <= g.K &&
          ((reinterpret_cast<uintptr_t>(g.B + n * g.sbn + k0 + kc) & 15) == 0)) {
        *reinterpret_cast<bf16x8*>(&Bs[r][kc]) =
            *reinterpret_cast<const bf16x8*>(g.B + n * g.sbn + k0 + kc);
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const long long k = k0 + kc + j;
          Bs[r][kc + j] = (n < g.N && k < g.K) ? g.B[k * g.sbk + n * g.sbn] : (__bf16)0.0f;
        }
      }
    }
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
            if (g.accumulate) v += g.C[m * g.ldc + n];
            g.C[m * g.ldc + n] = v;
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

// db[n] += sum_m dZ[m, n] (bf16 in, fp32 out) -- partial column sums per
// block combined with atomics.
__global__ void k_colsum(const __bf16* __restrict__ dz, long long ld, long long M,
                         long long N, float* __restrict__ out) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const long long rows = (M + gridDim.y - 1) / gridDim.y;
  const long long r0 = (long long)blockIdx.y * rows;
  const long long r1 = r0 + rows < M ? r0 + rows : M;
  float s = 0.f;
  for (long long m = r0; m < r1; ++m) s += (float)dz[m * ld + n];
  atomicAdd(out + n, s);
}

// Adam on fp32 master parameters; writes the bf16 copy used by the GEMMs.
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

// ------------------------------------------------ surrogate field plumbing
// Network inputs for points xyz[n*3] at times t (t_stride 0: one time):
// X[i] = (x/L, y/L, z/L, cos wt, sin wt, 0 ...) as bf16 rows of width ldx.
__global__ void k_sur_inputs(const double* __restrict__ xyz,
                             const double* __restrict__ t, int t_stride,
                             long long n, double inv_len, double omega,
                             __bf16* __restrict__ X, long long ldx) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double ti = t[t_stride ? i : 0];
  double s, c;
  sincos(omega * ti, &s, &c);
  __bf16* row = X + i * ldx;
  row[0] = (__bf16)(float)(xyz[3 * i] * inv_len);
  row[1] = (__bf16)(float)(xyz[3 * i + 1] * inv_len);
  row[2] = (__bf16)(float)(xyz[3 * i + 2] * inv_len);
  row[3] = (__bf16)(float)c;
  row[4] = (__bf16)(float)s;
  for (long long k = 5; k < ldx; ++k) row[k] = (__bf16)0.0f;
}

// a = a_base(x) (exact FP64) + scale * Y[i] (network), or the exact full
// field where the gate says so: outside the trained ball |x| > r_max, or for
// orbits flagged by the periodic exact check.
__global__ void k_sur_combine(DPot base, DPot full, const double* __restrict__ xyz,
                              const double* __restrict__ t, int t_stride,
                              long long n, const float* __restrict__ Y,
                              long long ldy, double scale, double r_max,
                              const int* __restrict__ flag,
                              double* __restrict__ acc, int* __restrict__ used_exact) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
  const double ti = t[t_stride ? i : 0];
  const bool exact = (x * x + y * y + z * z > r_max * r_max) || (flag && flag[i]);
  double ax, ay, az;
  if (exact) {
    accel<true>(full, x, y, z, ti, ax, ay, az);
    if (used_exact) used_exact[i] = 1;
  } else {
    accel<false>(base, x, y, z, ti, ax, ay, az);
    ax += scale * (double)Y[i * ldy];
    ay += scale * (double)Y[i * ldy + 1];
    az += scale * (double)Y[i * ldy + 2];
  }
  acc[3 * i] = ax;
  acc[3 * i + 1] = ay;
  acc[3 * i + 2] = az;
}

// Lock-step fixed-step Dopri8 over n orbits with stage accelerations held in
// HBM: stage positions Q_i, and the step update.  G is [S][n][3].
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
  for (int m = 0; m < 3; ++m) Q[3 * i + m] = fma(h, fma(rs, p[3 * i + m], h * s[m]), q[3 * i + m]);
}

__global__ void k_step_update(double* __restrict__ q, double* __restrict__ p,
                              const double* __restrict__ G, long long n, double h) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
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

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" {

int gx_mlp_gemm(const void* A, int64_t sam, int64_t sak, const void* B,
                int64_t sbk, int64_t sbn, int64_t M, int64_t N, int64_t K,
                int32_t epi, int32_t accumulate, const float* bias, float* C,
                int64_t ldc, void* H, int64_t ldh, const float* Zin,
                int64_t ldzin, void* stream) {
  if (M < 0 || N < 0 || K < 0 || !A || !B) return GX_EINVAL;
  if (M == 0 || N == 0) return GX_OK;
  if (epi < EPI_STORE || epi > EPI_DSILU) return GX_EINVAL;
  if ((epi == EPI_STORE || epi == EPI_BIAS) && !C) return GX_EINVAL;
  if ((epi == EPI_BIAS || epi == EPI_BIAS_SILU) && !bias) return GX_EINVAL;
  if ((epi == EPI_BIAS_SILU || epi == EPI_DSILU) && !H) return GX_EINVAL;
  if (epi == EPI_DSILU && !Zin) return GX_EINVAL;
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
  g.bias = bias;
  g.C = C;
  g.ldc = ldc;
  g.H = static_cast<__bf16*>(H);
  g.ldh = ldh;
  g.Zin = Zin;
  g.ldzin = ldzin;
  dim3 grid((unsigned)((N + BN - 1) / BN), (unsigned)((M + BM - 1) / BM));
  if (grid.y > 65535u * 1024u) return GX_EINVAL;
  hipLaunchKernelGGL(k_gemm<true>, grid, dim3(256), 0, (hipStream_t)stream, g);
  return check(hipGetLastError());
}

int gx_mlp_colsum(const void* dz, int64_t ld, int64_t M, int64_t N, float* out,
                  void* stream) {
  if (!dz || !out || M < 0 || N < 0) return GX_EINVAL;
  if (M == 0 || N == 0) return GX_OK;
  long long ychunks = M / 4096 + 1;
  if (ychunks > 1024) ychunks = 1024;
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
  const float c1 = 1.0f / (1.0f - powf(b1, (float)step));
  const float c2 = 1.0f / (1.0f - powf(b2, (float)step));
  hipLaunchKernelGGL(k_adam, dim3((unsigned)blocks_for(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, p, g, m, v, static_cast<__bf16*>(pb),
                     (long long)n, lr, b1, b2, eps, c1, c2);
  return check(hipGetLastError());
}

int gx_sur_inputs(const double* xyz, const double* t, int32_t t_stride, int64_t n,
                  double inv_len, double omega, void* X, int64_t ldx, void* stream) {
  if (!xyz || !t || !X || n < 0 || ldx < 5) return GX_EINVAL;
  if (n == 0) return GX_OK;
  hipLaunchKernelGGL(k_sur_inputs, dim3((unsigned)blocks_for(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, xyz, t, t_stride, (long long)n, inv_len,
                     omega, static_cast<__bf16*>(X), (long long)ldx);
  return check(hipGetLastError());
}

int gx_sur_combine(const gx_pot* base, const gx_pot* full, const double* xyz,
                   const double* t, int32_t t_stride, int64_t n, const float* Y,
                   int64_t ldy, double scale, double r_max, const int32_t* flag,
                   double* acc, int32_t* used_exact, void* stream) {
  if (!base || !full || !xyz || !t || !Y || !acc || n < 0) return GX_EINVAL;
  if (!base->basic) return GX_EINVAL;
  if (n == 0) return GX_OK;
  hipLaunchKernelGGL(k_sur_combine, dim3((unsigned)blocks_for(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, base->dev, full->dev, xyz, t, t_stride,
                     (long long)n, Y, (long long)ldy, scale, r_max, flag, acc,
                     used_exact);
  return check(hipGetLastError());
}

int gx_lockstep_stage(int32_t stage, const double* q, const double* p,
                      const double* G, int64_t n, double h, double* Q, void* stream) {
  if (stage < 0 || stage >= 13 || !q || !p || !G || !Q || n < 0) return GX_EINVAL;
  if (n == 0) return GX_OK;
  hipLaunchKernelGGL(k_stage_point, dim3((unsigned)blocks_for(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, (int)stage, q, p, G, (long long)n, h, Q);
  return check(hipGetLastError());
}

int gx_lockstep_update(double* q, double* p, const double* G, int64_t n, double h,
                       void* stream) {
  if (!q || !p || !G || n < 0) return GX_EINVAL;
  if (n == 0) return GX_OK;
  hipLaunchKernelGGL(k_step_update, dim3((unsigned)blocks_for(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, q, p, G, (long long)n, h);
  return check(hipGetLastError());
}

}  // extern "C"

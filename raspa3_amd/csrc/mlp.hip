// Learned insertion proposal: an MLP on periodic Fourier features of fractional coordinates, trained and evaluated
// with hand-written BF16 MFMA kernels (gfx950 v_mfma_f32_16x16x32_bf16, FP32 accumulation).
//
// Network: x(s) = [sin 2 pi m s_a, cos 2 pi m s_a]_{a = 0..2, m = 1..F} (6F inputs; F = 32 -> 192),
// three hidden layers of width 256 with SiLU, one linear output f(s) = an unnormalised log-density.
// Training minimises, per minibatch of points s_i with energies U_i, the cross-entropy between the Boltzmann weights
// p_i = softmax(-beta U)_i and the model's q_i = softmax(f)_i:  L = -sum_i p_i log q_i,  dL/df_i = q_i - p_i.
// Weights are kept as FP32 masters (updated by Adam) and rounded to BF16 for the products; activations are stored
// in BF16 for the next product and pre-activations in FP32 for the backward pass.
//
// A training step is four kernels: the fused forward pass (features, three layers and the output for 16 rows per
// workgroup, activations in LDS), the fused loss and backward pass down to the first layer's pre-activation
// gradients, the parameter gradients (three weight-gradient products in 64 x 64 tiles, split K, operands stored
// feature-major so every MFMA operand is one 16-byte LDS read, and in the same launch the bias and output-weight
// gradients as row sums), and Adam.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "autotune.hpp"
#include "rg_internal.hpp"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace rg
{
namespace
{

constexpr int WIDTH = 256;
constexpr int NFREQ = 32;
constexpr int NIN = 6 * NFREQ;  // 192

__device__ __forceinline__ float sigmoidf(float z) { return 1.0f / (1.0f + __expf(-z)); }

// Weight gradient C (M x N) += A B^T over k, with A (M x K) and B (N x K) row-major (k contiguous: the layer inputs
// and pre-activation gradients stored feature-major).  The workgroup computes the 64 x 64 tile (m0, n0) over k in
// [kBegin, kEnd) and adds it atomically (split K).  K is taken in stages of up to KS = 256: every global load of a
// stage is issued before any of them is written to LDS, so a stage costs one memory latency.
constexpr int KS = 256;
__device__ __forceinline__ void weightGradTile(int N, int kBegin, int kEnd, const __bf16* __restrict__ A, int lda,
                                               const __bf16* __restrict__ B, int ldb, float* __restrict__ C, int m0,
                                               int n0, __bf16 (*As)[KS + 8], __bf16 (*Bs)[KS + 8])
{
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kBegin; k0 < kEnd; k0 += KS)
  {
    const int kk = min(KS, kEnd - k0);  // a multiple of 32
    const int nv = 8 * kk;               // 8-element vectors per operand tile (64 rows x kk)
    const int rowv = kk / 8;             // vectors along k per row
    bf16x8 va[KS / 32], vb[KS / 32];
#pragma unroll
    for (int q = 0; q < KS / 32; ++q)
    {
      const int v = t + 256 * q;
      if (v < nv)
      {
        va[q] = *reinterpret_cast<const bf16x8*>(A + size_t(m0 + v / rowv) * lda + k0 + (v % rowv) * 8);
        vb[q] = *reinterpret_cast<const bf16x8*>(B + size_t(n0 + v / rowv) * ldb + k0 + (v % rowv) * 8);
      }
    }
#pragma unroll
    for (int q = 0; q < KS / 32; ++q)
    {
      const int v = t + 256 * q;
      if (v < nv)
      {
        *reinterpret_cast<bf16x8*>(&As[v / rowv][(v % rowv) * 8]) = va[q];
        *reinterpret_cast<bf16x8*>(&Bs[v / rowv][(v % rowv) * 8]) = vb[q];
      }
    }
    __syncthreads();
    for (int ks = 0; ks < kk; ks += 32)
    {
#pragma unroll
      for (int i = 0; i < 2; ++i)
      {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[wm + 16 * i + (lane & 15)][ks + 8 * (lane >> 4)]);
#pragma unroll
        for (int j = 0; j < 2; ++j)
        {
          const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[wn + 16 * j + (lane & 15)][ks + 8 * (lane >> 4)]);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i][j], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        atomicAdd(&C[size_t(m0 + wm + 16 * i + 4 * (lane >> 4) + r) * N + n0 + wn + 16 * j + (lane & 15)],
                  acc[i][j][r]);
}

// Parameter layout in one FP32 array: W1 (192 x 256), b1, W2 (256 x 256), b2, W3 (256 x 256), b3, w4 (256), b4;
// each W stored in x out, row-major.  The fused forward pass reads transposed BF16 copies (out x in, row-major) of
// W1..W3, laid out one after another.
constexpr size_t OFF_W1 = 0;
constexpr size_t OFF_B1 = OFF_W1 + size_t(NIN) * WIDTH;
constexpr size_t OFF_W2 = OFF_B1 + WIDTH;
constexpr size_t OFF_B2 = OFF_W2 + size_t(WIDTH) * WIDTH;
constexpr size_t OFF_W3 = OFF_B2 + WIDTH;
constexpr size_t OFF_B3 = OFF_W3 + size_t(WIDTH) * WIDTH;
constexpr size_t OFF_W4 = OFF_B3 + WIDTH;
constexpr size_t OFF_B4 = OFF_W4 + WIDTH;
constexpr size_t NPARAM = OFF_B4 + 1;
constexpr size_t OFF_T1 = 0;
constexpr size_t OFF_T2 = OFF_T1 + size_t(WIDTH) * NIN;
constexpr size_t OFF_T3 = OFF_T2 + size_t(WIDTH) * WIDTH;
constexpr size_t NTRANS = OFF_T3 + size_t(WIDTH) * WIDTH;

// index of parameter i in the transposed copies, or -1 for biases and w4
__device__ __forceinline__ long long transposedIndex(size_t i)
{
  if (i < OFF_B1) return (long long)(OFF_T1 + (i % WIDTH) * NIN + i / WIDTH);
  if (i >= OFF_W2 && i < OFF_B2) return (long long)(OFF_T2 + ((i - OFF_W2) % WIDTH) * WIDTH + (i - OFF_W2) / WIDTH);
  if (i >= OFF_W3 && i < OFF_B3) return (long long)(OFF_T3 + ((i - OFF_W3) % WIDTH) * WIDTH + (i - OFF_W3) / WIDTH);
  return -1;
}

// Fused kernels: one workgroup of four wavefronts takes FROWS batch rows through the network.  Activations stay in
// LDS (row-major, k contiguous, row stride LDSW); every layer is computed in transposed form, Z^T = W^T H^T
// (forward) and dH^T = W dZ^T (backward), so that the A operand is rows of a weight matrix read straight from global
// memory (shared by all workgroups through the L2) and the B operand is rows of the LDS tile.  Wavefront w owns
// output features 64 w .. 64 w + 63 (four 16-row A blocks) for all FROWS rows (RB 16-column B blocks); a lane's
// four results per block are four consecutive features of one row, written to LDS as one 8-byte store.  16 rows per
// workgroup spread a minibatch of 4096 over 256 workgroups, one per compute unit: the per-row work of the epilogues
// (SiLU, stores) then runs on every SIMD of the chip rather than on a quarter of them.
constexpr int FROWS = 16;
constexpr int RB = FROWS / 16;
constexpr int LDSW = WIDTH + 8;
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// acc[i][j] += Wrows(64 w + 16 i ..) x tile rows (16 j ..) over k in [0, K): Wrows row-major with row stride K.
// All of the wavefront's A fragments (K / 32 steps x 4 blocks, 16 B each per lane) are loaded before the first
// product, so a layer waits for one memory latency rather than one per K step.
template <int K>
__device__ __forceinline__ void tileProduct(const __bf16* __restrict__ Wrows, const __bf16 (*tile)[LDSW], int w,
                                            int lane, f32x4 (&acc)[4][RB])
{
  constexpr int NS = K / 32;
  const int kq = 8 * (lane >> 4);
  const __bf16* arow = Wrows + size_t(64 * w + (lane & 15)) * K + kq;
  bf16x8 a[NS][4];
#pragma unroll
  for (int k = 0; k < NS; ++k)
#pragma unroll
    for (int i = 0; i < 4; ++i) a[k][i] = *reinterpret_cast<const bf16x8*>(arow + size_t(16 * i) * K + 32 * k);
#pragma unroll
  for (int k = 0; k < NS; ++k)
  {
    bf16x8 b[RB];
#pragma unroll
    for (int j = 0; j < RB; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&tile[16 * j + (lane & 15)][32 * k + kq]);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < RB; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[k][i], b[j], acc[i][j], 0, 0, 0);
  }
}

// Writes nf features [f0, f0 + nf) of an LDS tile feature-major to global memory (out[f * ld + row0 + row]); a lane
// takes row lane % FROWS of feature lane / FROWS + (64 / FROWS) k, so each store covers FROWS consecutive values of
// 64 / FROWS features.
__device__ __forceinline__ void storeFeatureMajor(const __bf16 (*tile)[LDSW], int f0, int nf, __bf16* __restrict__ out,
                                                  int ld, int row0, int lane)
{
  constexpr int FPI = 64 / FROWS;
  const int row = lane % FROWS;
#pragma unroll 4
  for (int f = f0 + lane / FROWS; f < f0 + nf; f += FPI) out[size_t(f) * ld + row0 + row] = tile[row][f];
}

// Forward pass of FROWS rows per workgroup: Fourier features (from frac, or gathered through perm with the energies
// copied to bU), three SiLU layers and the output f = H3 . w4 + b4.  TRAIN also stores the pre-activations Z1..Z3
// (row-major FP32) and the layer outputs X, H1, H2, H3 feature-major (ld = ldt) for the parameter gradients.
template <bool TRAIN>
__global__ __launch_bounds__(256) void k_mlp_forward(const double* __restrict__ frac, const int* __restrict__ perm,
                                                     const double* __restrict__ Uall, double* __restrict__ bU,
                                                     const __bf16* __restrict__ wt, const __bf16* __restrict__ wb,
                                                     const float* __restrict__ w, float* __restrict__ Z1,
                                                     float* __restrict__ Z2, float* __restrict__ Z3, __bf16* __restrict__ Xt,
                                                     __bf16* __restrict__ H1t, __bf16* __restrict__ H2t,
                                                     __bf16* __restrict__ H3t, int ldt, float* __restrict__ f,
                                                     double beta, double* __restrict__ part)
{
  __shared__ __attribute__((aligned(16))) __bf16 bufA[FROWS][LDSW];
  __shared__ __attribute__((aligned(16))) __bf16 bufB[FROWS][LDSW];
  __shared__ double us[FROWS];  // -beta U of the rows (training)
  __shared__ float fs[FROWS];
  const int t = threadIdx.x, lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const int row0 = blockIdx.x * FROWS;
  // features: thread (row, axis) evaluates sin and cos of 2 pi m s for m = 1..32 by the angle-addition recurrence in
  // FP64 from one FP64 sincos (exact far below the BF16 resolution of the stored feature)
  if (t < 3 * FROWS)
  {
    const int row = t & (FROWS - 1), ax = t / FROWS;
    const int gi = row0 + row;
    const int src = perm ? perm[gi] : gi;
    if (TRAIN && ax == 0)
    {
      const double u = perm ? Uall[src] : bU[gi];
      if (perm) bU[gi] = u;
      us[row] = -beta * u;
    }
    double s1, c1;
    sincos(2.0 * M_PI * frac[3 * size_t(src) + ax], &s1, &c1);
    double sp = 0.0, cp = 1.0, sm = s1, cm = c1;
#pragma unroll
    for (int m8 = 0; m8 < NFREQ / 4; ++m8)
    {
      bf16x8 v;
#pragma unroll
      for (int q = 0; q < 4; ++q)
      {
        v[2 * q] = (__bf16)sm;
        v[2 * q + 1] = (__bf16)cm;
        const double sn = 2.0 * c1 * sm - sp, cn = 2.0 * c1 * cm - cp;
        sp = sm;
        cp = cm;
        sm = sn;
        cm = cn;
      }
      *reinterpret_cast<bf16x8*>(&bufA[row][2 * NFREQ * ax + 8 * m8]) = v;
    }
  }
  __syncthreads();
  if (TRAIN) storeFeatureMajor(bufA, 48 * wv, 48, Xt, ldt, row0, lane);
  // three layers: A operand = transposed weights, B operand = the LDS tile of the layer input
  __bf16 (*in)[LDSW] = bufA;
  __bf16 (*out)[LDSW] = bufB;
#pragma unroll 1
  for (int l = 0; l < 3; ++l)
  {
    f32x4 acc[4][RB];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < RB; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    if (l == 0)
      tileProduct<NIN>(wt + OFF_T1, in, wv, lane, acc);
    else
      tileProduct<WIDTH>(wt + (l == 1 ? OFF_T2 : OFF_T3), in, wv, lane, acc);
    const float* bias = w + (l == 0 ? OFF_B1 : l == 1 ? OFF_B2 : OFF_B3);
    float* Z = l == 0 ? Z1 : l == 1 ? Z2 : Z3;
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
      const int o0 = 64 * wv + 16 * i + 4 * (lane >> 4);
      const float4 bv = *reinterpret_cast<const float4*>(bias + o0);
#pragma unroll
      for (int j = 0; j < RB; ++j)
      {
        const int row = 16 * j + (lane & 15);
        const float z0 = acc[i][j][0] + bv.x, z1 = acc[i][j][1] + bv.y, z2 = acc[i][j][2] + bv.z,
                    z3 = acc[i][j][3] + bv.w;
        if (TRAIN) *reinterpret_cast<float4*>(Z + size_t(row0 + row) * WIDTH + o0) = float4{z0, z1, z2, z3};
        *reinterpret_cast<bf16x4*>(&out[row][o0]) =
            bf16x4{(__bf16)(z0 * sigmoidf(z0)), (__bf16)(z1 * sigmoidf(z1)), (__bf16)(z2 * sigmoidf(z2)),
                   (__bf16)(z3 * sigmoidf(z3))};
      }
    }
    __syncthreads();
    if (TRAIN) storeFeatureMajor(out, 64 * wv, 64, l == 0 ? H1t : l == 1 ? H2t : H3t, ldt, row0, lane);
    __bf16 (*nin)[LDSW] = out;
    out = in;
    in = nin;
  }
  // output: f = H3 . w4 + b4, TPR consecutive threads per row over WIDTH / TPR features each
  {
    constexpr int TPR = 256 / FROWS, FPT = WIDTH / TPR;
    const int row = t / TPR, part = t % TPR;
    float s = 0.f;
#pragma unroll 8
    for (int k = FPT * part; k < FPT * part + FPT; ++k) s += float(in[row][k]) * float(wb[OFF_W4 + k]);
#pragma unroll
    for (int o = 1; o < TPR; o <<= 1) s += __shfl_xor(s, o, 64);
    if (part == 0)
    {
      f[row0 + row] = s + w[OFF_B4];
      fs[row] = s + w[OFF_B4];
    }
  }
  if (TRAIN)
  {
    // the workgroup's partial softmax statistics (the backward pass combines them): max f, sum exp(f - max f),
    // max u, sum exp(u - max u) and sum exp(u - max u) f over its rows, u = -beta U
    __syncthreads();
    if (t < 64)
    {
      const bool live = t < FROWS;
      const double fv = live ? double(fs[t]) : -INFINITY, uv = live ? us[t] : -INFINITY;
      double mf = fv, mu = uv;
      for (int o = 1; o < 64; o <<= 1)
      {
        mf = fmax(mf, __shfl_xor(mf, o, 64));
        mu = fmax(mu, __shfl_xor(mu, o, 64));
      }
      // rows of infinite energy (u = -inf) weigh nothing, also when they are all of the workgroup's rows
      const double eu = (live && uv > -INFINITY) ? double(expf(float(uv - mu))) : 0.0;
      double sf = live ? double(expf(float(fv - mf))) : 0.0, su = eu, tu = eu > 0.0 ? eu * fv : 0.0;
      for (int o = 1; o < 64; o <<= 1)
      {
        sf += __shfl_xor(sf, o, 64);
        su += __shfl_xor(su, o, 64);
        tu += __shfl_xor(tu, o, 64);
      }
      if (t == 0)
      {
        double* q = part + 5 * size_t(blockIdx.x);
        q[0] = mf;
        q[1] = sf;
        q[2] = mu;
        q[3] = su;
        q[4] = tu;
      }
    }
  }
}

// Loss of a batch: p = softmax(u), u = -beta U, q = softmax(f), L = -sum p log q, dL/df = q - p; the maxima and
// the exponent arguments in FP64, the exponentials of rows in FP32 (df is FP32), the sums in FP64.  Every workgroup
// of the forward pass leaves partial statistics of its rows (max f, sum exp(f - max f), max u, sum exp(u - max u),
// sum exp(u - max u) f); every workgroup of the backward pass combines all of them in the same order (so every
// workgroup obtains the same values): M = max of the maxima, S = sum of the partial sums times exp(max - M) in FP64.
// Then lse = Mf + log Sf and L = lse - T / Su, T the combined sum exp(u - Mu) f.
__device__ __forceinline__ void reduceBlock3(double& a, double& b, double& c, bool isMax, double (*red)[4])
{
  const int t = threadIdx.x;
  for (int o = 32; o > 0; o >>= 1)
  {
    const double x = __shfl_xor(a, o, 64), y = __shfl_xor(b, o, 64), z = __shfl_xor(c, o, 64);
    a = isMax ? fmax(a, x) : a + x;
    b = isMax ? fmax(b, y) : b + y;
    c = isMax ? fmax(c, z) : c + z;
  }
  if ((t & 63) == 0)
  {
    red[0][t >> 6] = a;
    red[1][t >> 6] = b;
    red[2][t >> 6] = c;
  }
  __syncthreads();
  a = red[0][0];
  b = red[1][0];
  c = red[2][0];
  for (int i = 1; i < 4; ++i)
  {
    a = isMax ? fmax(a, red[0][i]) : a + red[0][i];
    b = isMax ? fmax(b, red[1][i]) : b + red[1][i];
    c = isMax ? fmax(c, red[2][i]) : c + red[2][i];
  }
  __syncthreads();
}

struct LossStats
{
  double mf, mu, su, lse, loss;
};

__device__ __forceinline__ LossStats combineStats(int nparts, const double* __restrict__ part, double (*red)[4])
{
  const int t = threadIdx.x;
  double mf = -INFINITY, mu = -INFINITY, unused = 0.0;
  for (int j = t; j < nparts; j += 256)
  {
    mf = fmax(mf, part[5 * j]);
    mu = fmax(mu, part[5 * j + 2]);
  }
  reduceBlock3(mf, mu, unused, true, red);
  double sf = 0.0, su = 0.0, tu = 0.0;
  for (int j = t; j < nparts; j += 256)
  {
    const double* q = part + 5 * j;
    const double a = exp(q[0] - mf), b = exp(q[2] - mu);
    sf += q[1] * a;
    su += q[3] * b;
    tu += q[4] * b;
  }
  reduceBlock3(sf, su, tu, false, red);
  LossStats L;
  L.mf = mf;
  L.mu = mu;
  L.su = su;
  L.lse = mf + log(sf);
  L.loss = L.lse - tu / su;
  return L;
}

// Backward through one hidden layer for the workgroup's rows: dH^T = W dZ^T (W row-major in x out, dZ the LDS tile),
// dZprev = dH * SiLU'(Zprev) rounded to BF16 into the output tile.
__device__ __forceinline__ void backLayer(const __bf16* __restrict__ W, const float* __restrict__ Zprev,
                                          const __bf16 (*din)[LDSW], __bf16 (*dout)[LDSW], int row0, int wv, int lane)
{
  f32x4 acc[4][RB];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < RB; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  tileProduct<WIDTH>(W, din, wv, lane, acc);
#pragma unroll
  for (int i = 0; i < 4; ++i)
  {
    const int k0 = 64 * wv + 16 * i + 4 * (lane >> 4);
#pragma unroll
    for (int j = 0; j < RB; ++j)
    {
      const int row = 16 * j + (lane & 15);
      const float4 zv = *reinterpret_cast<const float4*>(Zprev + size_t(row0 + row) * WIDTH + k0);
      const float z[4] = {zv.x, zv.y, zv.z, zv.w};
      bf16x4 h;
#pragma unroll
      for (int r = 0; r < 4; ++r)
      {
        const float s = sigmoidf(z[r]);
        h[r] = (__bf16)(acc[i][j][r] * s * (1.0f + z[r] * (1.0f - s)));
      }
      *reinterpret_cast<bf16x4*>(&dout[row][k0]) = h;
    }
  }
}

// Backward pass of FROWS rows per workgroup, down to the gradient of the first layer's pre-activations: output layer
// (dZ3 = df w4 SiLU'(Z3)), then layers 3 and 2 (dZ2, dZ1), all stored feature-major (ld = ldt) for the parameter
// gradients.
// The loss and the backward pass of n rows: every workgroup first combines the forward pass's partial statistics,
// then forms the loss gradient dL/df = q - p of its own FROWS rows (stored to df for the parameter gradients) and
// propagates it down to the first layer's pre-activation gradients.  Workgroup 0 also stores the batch loss
// (loss[0] = L, loss[1] += L).
__global__ __launch_bounds__(256) void k_mlp_backward(int n, const float* __restrict__ f, const double* __restrict__ U,
                                                      double beta, const double* __restrict__ part,
                                                      float* __restrict__ df, double* __restrict__ loss,
                                                      const float* __restrict__ Z1, const float* __restrict__ Z2,
                                                      const float* __restrict__ Z3, const __bf16* __restrict__ wb,
                                                      __bf16* __restrict__ dZ1t, __bf16* __restrict__ dZ2t,
                                                      __bf16* __restrict__ dZ3t, int ldt)
{
  __shared__ __attribute__((aligned(16))) __bf16 bufA[FROWS][LDSW];
  __shared__ __attribute__((aligned(16))) __bf16 bufB[FROWS][LDSW];
  __shared__ float dfs[FROWS];
  __shared__ double red[3][4];
  const int t = threadIdx.x, lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const int row0 = blockIdx.x * FROWS;
  {
    const LossStats L = combineStats(n / FROWS, part, red);
    if (t < FROWS)
    {
      const int i = row0 + t;
      const double p = double(expf(float(-beta * U[i] - L.mu))) / L.su;
      const double lq = double(f[i]) - L.lse;
      const float d = float(double(expf(float(lq))) - p);
      dfs[t] = d;
      df[i] = d;
    }
    if (blockIdx.x == 0 && t == 0)
    {
      loss[0] = L.loss;
      loss[1] += L.loss;
    }
  }
  __syncthreads();
  {
    // output layer, one thread per hidden unit j
    const int j = t;
    const float wj = float(wb[OFF_W4 + j]);
#pragma unroll 16
    for (int row = 0; row < FROWS; ++row)
    {
      const float z = Z3[size_t(row0 + row) * WIDTH + j], s = sigmoidf(z), d = dfs[row];
      bufA[row][j] = (__bf16)(d * wj * s * (1.0f + z * (1.0f - s)));
    }
  }
  __syncthreads();
  storeFeatureMajor(bufA, 64 * wv, 64, dZ3t, ldt, row0, lane);
  backLayer(wb + OFF_W3, Z2, bufA, bufB, row0, wv, lane);
  __syncthreads();
  storeFeatureMajor(bufB, 64 * wv, 64, dZ2t, ldt, row0, lane);
  backLayer(wb + OFF_W2, Z1, bufB, bufA, row0, wv, lane);
  __syncthreads();
  storeFeatureMajor(bufA, 64 * wv, 64, dZ1t, ldt, row0, lane);
}

// Bias gradients and the output weights' gradient as row sums over the batch, one wavefront per sum (no atomics, a
// fixed order): task 0..4 WIDTH - 1 takes row j = task % WIDTH of dZ1t, dZ2t, dZ3t (db1, db2, db3) or of H3t
// weighted by df (dw4); task 4 WIDTH sums df (db4).  Every sum is stored, not added: the caller cleared g.
__device__ __forceinline__ void biasSum(int task, int lane, int n, int ld, const __bf16* __restrict__ dZ1t,
                                        const __bf16* __restrict__ dZ2t, const __bf16* __restrict__ dZ3t,
                                        const __bf16* __restrict__ H3t, const float* __restrict__ df,
                                        float* __restrict__ g)
{
  const int which = task / WIDTH, j = task % WIDTH;
  // n is a multiple of 16: each lane reads 8 consecutive values per step (16-byte loads, all in flight together)
  float s = 0.f;
  if (which == 4)
  {
    for (int i = 8 * lane; i < n; i += 512)
    {
      const float4 a = *reinterpret_cast<const float4*>(df + i), b = *reinterpret_cast<const float4*>(df + i + 4);
      s += (a.x + a.y) + (a.z + a.w) + (b.x + b.y) + (b.z + b.w);
    }
  }
  else
  {
    const __bf16* row = (which == 0 ? dZ1t : which == 1 ? dZ2t : which == 2 ? dZ3t : H3t) + size_t(j) * ld;
#pragma unroll 4
    for (int i = 8 * lane; i < n; i += 512)
    {
      const bf16x8 v = *reinterpret_cast<const bf16x8*>(row + i);
      if (which == 3)
      {
        const float4 a = *reinterpret_cast<const float4*>(df + i), b = *reinterpret_cast<const float4*>(df + i + 4);
        s += (float(v[0]) * a.x + float(v[1]) * a.y) + (float(v[2]) * a.z + float(v[3]) * a.w) +
             (float(v[4]) * b.x + float(v[5]) * b.y) + (float(v[6]) * b.z + float(v[7]) * b.w);
      }
      else
        s += (float(v[0]) + float(v[1])) + (float(v[2]) + float(v[3])) + (float(v[4]) + float(v[5])) +
             (float(v[6]) + float(v[7]));
    }
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0)
  {
    const size_t off = which == 0 ? OFF_B1 : which == 1 ? OFF_B2 : which == 2 ? OFF_B3 : which == 3 ? OFF_W4 : OFF_B4;
    g[off + (which == 4 ? 0 : j)] = s;
  }
}

// The parameter gradients of a minibatch in one launch: z-slices 0 .. 3 nsplit - 1 are the three weight-gradient
// products (layer l = z / nsplit, K chunk z % nsplit of Kc rows; 64 x 64 tiles, workgroups beyond a layer's M
// return at once), the last BIAS_Z slices the row sums of biasSum spread over their wavefronts.
constexpr int BIAS_Z = 4;
struct ParamGradArgs
{
  const __bf16* A[3];  // layer inputs, feature-major: H2t, H1t, Xt
  const __bf16* B[3];  // pre-activation gradients, feature-major: dZ3t, dZ2t, dZ1t
  float* C[3];         // gradient blocks: W3, W2, W1
  int M[3];            // rows of each block: WIDTH, WIDTH, NIN
  const __bf16 *dZ1t, *dZ2t, *dZ3t, *H3t;
  const float* df;
  float* g;
};
__global__ __launch_bounds__(256) void k_param_grads(ParamGradArgs a, int K, int Kc, int nsplit, int ld)
{
  __shared__ __attribute__((aligned(16))) __bf16 As[64][KS + 8];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[64][KS + 8];
  const int z = blockIdx.z;
  if (z >= 3 * nsplit)
  {
    const int lane = threadIdx.x & 63;
    const int nw = 4 * gridDim.x * gridDim.y * BIAS_Z;
    const int wid = 4 * ((z - 3 * nsplit) * gridDim.x * gridDim.y + blockIdx.y * gridDim.x + blockIdx.x) +
                    int(threadIdx.x >> 6);
    for (int task = wid; task <= 4 * WIDTH; task += nw) biasSum(task, lane, K, ld, a.dZ1t, a.dZ2t, a.dZ3t, a.H3t, a.df, a.g);
    return;
  }
  const int l = z / nsplit, sp = z % nsplit;
  if (int(blockIdx.y) * 64 >= a.M[l]) return;  // the whole workgroup leaves before any barrier
  weightGradTile(WIDTH, sp * Kc, min(K, (sp + 1) * Kc), a.A[l], ld, a.B[l], ld, a.C[l], blockIdx.y * 64,
                 blockIdx.x * 64, As, Bs);
}

__global__ void k_zero(float* p, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

// Adam (Kingma & Ba, 2015) on the FP32 masters, then the BF16 copies used by the products (row-major and, for
// W1..W3, transposed); the gradient is cleared for the next minibatch's atomic sums.
// One Adam update of parameter i (the gradient is cleared for the next step); returns the new weight.
__device__ __forceinline__ float adamOne(size_t i, float* __restrict__ w, float* __restrict__ g, float* __restrict__ m,
                                        float* __restrict__ v, __bf16* __restrict__ wb, float lr, float b1, float b2,
                                        float eps, float c1, float c2)
{
  const float gi = g[i];
  g[i] = 0.f;
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  wb[i] = (__bf16)wi;
  return wi;
}

// Adam over all parameters, with the BF16 copies refreshed: workgroups 0 .. ADAM_TILES - 1 take 32 x 32 tiles of
// W1, W2 and W3 (rows k, columns j; coalesced along j) and write the transposed copy through an LDS tile (coalesced
// along k); the remaining workgroups take the biases, w4 and b4 one parameter per thread.
constexpr int ADAM_T1 = (NIN / 32) * (WIDTH / 32), ADAM_T2 = (WIDTH / 32) * (WIDTH / 32);
constexpr int ADAM_TILES = ADAM_T1 + 2 * ADAM_T2;
constexpr int ADAM_REST = int(OFF_W2 - OFF_B1) + int(OFF_W3 - OFF_B2) + int(NPARAM - OFF_B3);
constexpr int ADAM_BLOCKS = ADAM_TILES + (ADAM_REST + 255) / 256;
static_assert(NIN % 32 == 0 && WIDTH % 32 == 0, "Adam tiles are 32 x 32");
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ w, float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, __bf16* __restrict__ wb, __bf16* __restrict__ wt,
                                              float lr, float b1, float b2, float eps, float c1, float c2)
{
  __shared__ __bf16 tile[32][34];
  const int t = threadIdx.x;
  const int b = blockIdx.x;
  if (b < ADAM_TILES)
  {
    size_t off, toff;
    int K, tb;
    if (b < ADAM_T1)
    {
      off = OFF_W1, toff = OFF_T1, K = NIN, tb = b;
    }
    else if (b < ADAM_T1 + ADAM_T2)
    {
      off = OFF_W2, toff = OFF_T2, K = WIDTH, tb = b - ADAM_T1;
    }
    else
    {
      off = OFF_W3, toff = OFF_T3, K = WIDTH, tb = b - ADAM_T1 - ADAM_T2;
    }
    const int k0 = (tb / (WIDTH / 32)) * 32, j0 = (tb % (WIDTH / 32)) * 32;
    const int tx = t & 31, ty = t >> 5;
#pragma unroll
    for (int q = 0; q < 4; ++q)
    {
      const int k = k0 + ty + 8 * q;
      tile[tx][ty + 8 * q] = (__bf16)adamOne(off + size_t(k) * WIDTH + j0 + tx, w, g, m, v, wb, lr, b1, b2, eps, c1, c2);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q)
    {
      const int j = j0 + ty + 8 * q;
      wt[toff + size_t(j) * K + k0 + tx] = tile[ty + 8 * q][tx];
    }
    return;
  }
  const int e = (b - ADAM_TILES) * 256 + t;
  if (e >= ADAM_REST) return;
  const int nb1 = int(OFF_W2 - OFF_B1), nb2 = int(OFF_W3 - OFF_B2);
  const size_t i = e < nb1 ? OFF_B1 + e : (e < nb1 + nb2 ? OFF_B2 + (e - nb1) : OFF_B3 + (e - nb1 - nb2));
  adamOne(i, w, g, m, v, wb, lr, b1, b2, eps, c1, c2);
}

__global__ void k_zero_d(double* p, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

__global__ void k_to_bf16(int n, const float* __restrict__ a, __bf16* __restrict__ b, __bf16* __restrict__ bt)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  b[i] = (__bf16)a[i];
  const long long ti = transposedIndex(size_t(i));
  if (ti >= 0) bt[ti] = (__bf16)a[i];
}

}  // namespace
}  // namespace rg

using namespace rg;

struct rg_mlp
{
  rg_ctx* ctx = nullptr;
  int device = 0;
  float* w = nullptr;   // FP32 masters
  __bf16* wb = nullptr; // BF16 copies
  __bf16* wt = nullptr; // BF16 copies of W1..W3 transposed (out x in)
  float* g = nullptr;   // gradients
  float* m = nullptr;   // Adam moments
  float* v = nullptr;
  long long adamStep = 0;
  // activations for the current batch capacity
  int cap = 0;
  // layer inputs and pre-activation gradients feature-major (features x cap), for the weight gradients
  __bf16 *Xt = nullptr, *H1t = nullptr, *H2t = nullptr, *H3t = nullptr, *dZ1t = nullptr, *dZ2t = nullptr,
         *dZ3t = nullptr;
  float *Z1 = nullptr, *Z2 = nullptr, *Z3 = nullptr, *f = nullptr, *df = nullptr;
  double *frac = nullptr, *U = nullptr, *loss = nullptr;
  double* part = nullptr;  // the forward pass's partial softmax statistics, 5 per workgroup
};

namespace
{

void freeBatch(rg_mlp* p)
{
  void* bufs[] = {p->Xt, p->H1t, p->H2t, p->H3t, p->dZ1t, p->dZ2t, p->dZ3t, p->Z1, p->Z2, p->Z3, p->f, p->df,
                  p->frac, p->U, p->loss, p->part};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  p->Xt = p->H1t = p->H2t = p->H3t = p->dZ1t = p->dZ2t = p->dZ3t = nullptr;
  p->Z1 = p->Z2 = p->Z3 = p->f = p->df = nullptr;
  p->frac = p->U = p->loss = p->part = nullptr;
  p->cap = 0;
}

rg_status ensureBatch(rg_mlp* p, int n)
{
  rg_ctx* ctx = p->ctx;
  const int need = (n + FROWS - 1) / FROWS * FROWS;
  if (p->cap >= need) return RG_OK;
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  freeBatch(p);
  const size_t nb = need;
  RG_HIP(ctx, hipMalloc(&p->Xt, nb * NIN * sizeof(__bf16)));
  RG_HIP(ctx, hipMalloc(&p->H1t, nb * WIDTH * sizeof(__bf16)));
  RG_HIP(ctx, hipMalloc(&p->H2t, nb * WIDTH * sizeof(__bf16)));
  RG_HIP(ctx, hipMalloc(&p->H3t, nb * WIDTH * sizeof(__bf16)));
  RG_HIP(ctx, hipMalloc(&p->dZ1t, nb * WIDTH * sizeof(__bf16)));
  RG_HIP(ctx, hipMalloc(&p->dZ2t, nb * WIDTH * sizeof(__bf16)));
  RG_HIP(ctx, hipMalloc(&p->dZ3t, nb * WIDTH * sizeof(__bf16)));
  RG_HIP(ctx, hipMalloc(&p->Z1, nb * WIDTH * sizeof(float)));
  RG_HIP(ctx, hipMalloc(&p->Z2, nb * WIDTH * sizeof(float)));
  RG_HIP(ctx, hipMalloc(&p->Z3, nb * WIDTH * sizeof(float)));
  RG_HIP(ctx, hipMalloc(&p->f, nb * sizeof(float)));
  RG_HIP(ctx, hipMalloc(&p->df, nb * sizeof(float)));
  RG_HIP(ctx, hipMalloc(&p->frac, nb * 3 * sizeof(double)));
  RG_HIP(ctx, hipMalloc(&p->U, nb * sizeof(double)));
  RG_HIP(ctx, hipMalloc(&p->loss, 8 * sizeof(double)));
  RG_HIP(ctx, hipMalloc(&p->part, nb / FROWS * 5 * sizeof(double)));
  p->cap = need;
  return RG_OK;
}

// Forward pass of nrow rows (a multiple of FROWS, at most the batch capacity): fractional coordinates from p->frac,
// or (perm != nullptr) gathered from fracAll through perm with the energies copied from Uall to p->U.  train keeps
// what the backward pass needs, including the partial loss statistics at inverse temperature beta.
rg_status forward(rg_mlp* p, int nrow, bool train, double beta = 0.0, const int* perm = nullptr,
                  const double* fracAll = nullptr, const double* Uall = nullptr)
{
  if (nrow % FROWS != 0 || nrow > p->cap) return fail(p->ctx, RG_ERR_INVALID, "mlp forward: %d rows", nrow);
  const double* fr = perm ? fracAll : p->frac;
  if (train)
    hipLaunchKernelGGL(k_mlp_forward<true>, dim3(nrow / FROWS), dim3(256), 0, p->ctx->stream, fr, perm, Uall, p->U,
                       p->wt, p->wb, p->w, p->Z1, p->Z2, p->Z3, p->Xt, p->H1t, p->H2t, p->H3t, p->cap, p->f, beta,
                       p->part);
  else
    hipLaunchKernelGGL(k_mlp_forward<false>, dim3(nrow / FROWS), dim3(256), 0, p->ctx->stream, fr, perm, Uall, p->U,
                       p->wt, p->wb, p->w, p->Z1, p->Z2, p->Z3, p->Xt, p->H1t, p->H2t, p->H3t, p->cap, p->f, beta,
                       nullptr);
  RG_HIP(p->ctx, hipGetLastError());
  return RG_OK;
}

// The parameter gradients (k_param_grads) of n rows: the K chunk of the weight-gradient products is tuned online
// among 256 .. 4096 rows (results differ only in the order of the FP32 atomic additions, which is unordered anyway).
rg_status paramGrads(rg_mlp* p, int n)
{
  rg_ctx* ctx = p->ctx;
  ParamGradArgs a{};
  a.A[0] = p->H2t, a.B[0] = p->dZ3t, a.C[0] = p->g + OFF_W3, a.M[0] = WIDTH;
  a.A[1] = p->H1t, a.B[1] = p->dZ2t, a.C[1] = p->g + OFF_W2, a.M[1] = WIDTH;
  a.A[2] = p->Xt, a.B[2] = p->dZ1t, a.C[2] = p->g + OFF_W1, a.M[2] = NIN;
  a.dZ1t = p->dZ1t, a.dZ2t = p->dZ2t, a.dZ3t = p->dZ3t, a.H3t = p->H3t, a.df = p->df, a.g = p->g;
  static const int kChunks[5] = {256, 512, 1024, 2048, 4096};
  auto launch = [&](int kc)
  {
    kc = std::min(kc, n);
    const int nsplit = (n + kc - 1) / kc;
    const dim3 grid(WIDTH / 64, WIDTH / 64, 3 * nsplit + BIAS_Z);
    hipLaunchKernelGGL(k_param_grads, grid, dim3(256), 0, ctx->stream, a, n, kc, nsplit, p->cap);
    return hipGetLastError();
  };
  Autotuner& tuner = Autotuner::instance();
  const std::string key = "mlp_param_grads/n=" + std::to_string(n);
  const int cand = tuner.next(key, 5, 2);
  if (cand >= 0)
  {
    float ms = 0.0f;
    RG_HIP(ctx, timeLaunch(ctx->stream, [&] { return launch(kChunks[cand]); }, ms));
    tuner.report(key, cand, ms);
    return RG_OK;
  }
  RG_HIP(ctx, launch(kChunks[tuner.best(key)]));
  return RG_OK;
}

// Loss and gradients (added into p->g, which the caller has cleared) of a batch of n rows (a multiple of FROWS)
// whose forward pass ran with train = true: the backward pass with the loss fused in, then the parameter gradients.
rg_status backward(rg_mlp* p, int n, double beta)
{
  hipStream_t st = p->ctx->stream;
  hipLaunchKernelGGL(k_mlp_backward, dim3(n / FROWS), dim3(256), 0, st, n, p->f, p->U, beta, p->part, p->df, p->loss,
                     p->Z1, p->Z2, p->Z3, p->wb, p->dZ1t, p->dZ2t, p->dZ3t, p->cap);
  rg_status ws;
  if ((ws = paramGrads(p, n)) != RG_OK) return ws;
  RG_HIP(p->ctx, hipGetLastError());
  return RG_OK;
}

}  // namespace

extern "C"
{

size_t rg_mlp_num_params(void) { return NPARAM; }

rg_status rg_mlp_create(rg_ctx* ctx, unsigned long long seed, rg_mlp** out)
{
  if (!ctx || !out) return fail(ctx, RG_ERR_INVALID, "rg_mlp_create: invalid arguments");
  RG_ENTER(ctx);
  rg_mlp* p = new rg_mlp();
  p->ctx = ctx;
  p->device = ctx->device;
  RG_HIP(ctx, hipMalloc(&p->w, NPARAM * sizeof(float)));
  RG_HIP(ctx, hipMalloc(&p->wb, NPARAM * sizeof(__bf16)));
  RG_HIP(ctx, hipMalloc(&p->wt, NTRANS * sizeof(__bf16)));
  RG_HIP(ctx, hipMalloc(&p->g, NPARAM * sizeof(float)));
  RG_HIP(ctx, hipMalloc(&p->m, NPARAM * sizeof(float)));
  RG_HIP(ctx, hipMalloc(&p->v, NPARAM * sizeof(float)));
  // He-uniform initialisation from a seeded generator (splitmix64), zero biases
  std::vector<float> w(NPARAM, 0.f);
  unsigned long long s = seed;
  auto next = [&]()
  {
    s += 0x9E3779B97F4A7C15ull;
    unsigned long long z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return double(z >> 11) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](size_t off, size_t count, int fanIn)
  {
    const double a = std::sqrt(6.0 / fanIn);
    for (size_t i = 0; i < count; ++i) w[off + i] = float(a * (2.0 * next() - 1.0));
  };
  fill(OFF_W1, size_t(NIN) * WIDTH, NIN);
  fill(OFF_W2, size_t(WIDTH) * WIDTH, WIDTH);
  fill(OFF_W3, size_t(WIDTH) * WIDTH, WIDTH);
  fill(OFF_W4, WIDTH, WIDTH);
  RG_HIP(ctx, hipMemcpyAsync(p->w, w.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_to_bf16, dim3((NPARAM + 255) / 256), dim3(256), 0, ctx->stream, int(NPARAM), p->w, p->wb,
                     p->wt);
  hipLaunchKernelGGL(k_zero, dim3((NPARAM + 255) / 256), dim3(256), 0, ctx->stream, p->g, int(NPARAM));
  hipLaunchKernelGGL(k_zero, dim3((NPARAM + 255) / 256), dim3(256), 0, ctx->stream, p->m, int(NPARAM));
  hipLaunchKernelGGL(k_zero, dim3((NPARAM + 255) / 256), dim3(256), 0, ctx->stream, p->v, int(NPARAM));
  RG_HIP(ctx, hipGetLastError());
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *out = p;
  return RG_OK;
}

rg_status rg_mlp_destroy(rg_mlp* p)
{
  if (!p) return RG_ERR_INVALID;
  rg_status st = rgReleaseStatus(hipSetDevice(p->device), RG_OK);
  // the context (and its stream) may already be gone: wait for the whole device instead
  st = rgReleaseStatus(hipDeviceSynchronize(), st);
  freeBatch(p);
  void* bufs[] = {p->w, p->wb, p->wt, p->g, p->m, p->v};
  for (void* b : bufs)
    if (b) st = rgReleaseStatus(hipFree(b), st);
  delete p;
  return st;
}

rg_status rg_mlp_get_params(rg_ctx* ctx, const rg_mlp* p, float* out)
{
  if (!ctx || !p || !out) return fail(ctx, RG_ERR_INVALID, "rg_mlp_get_params: invalid arguments");
  RG_HIP(ctx, hipMemcpyAsync(out, p->w, NPARAM * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RG_OK;
}

rg_status rg_mlp_set_params(rg_ctx* ctx, rg_mlp* p, const float* in)
{
  if (!ctx || !p || !in) return fail(ctx, RG_ERR_INVALID, "rg_mlp_set_params: invalid arguments");
  RG_HIP(ctx, hipMemcpyAsync(p->w, in, NPARAM * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_to_bf16, dim3((NPARAM + 255) / 256), dim3(256), 0, ctx->stream, int(NPARAM), p->w, p->wb,
                     p->wt);
  RG_HIP(ctx, hipGetLastError());
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RG_OK;
}

rg_status rg_mlp_forward(rg_ctx* ctx, rg_mlp* p, int n, const double* frac, double* logdensity)
{
  if (!ctx || !p || n < 0 || (n > 0 && (!frac || !logdensity)))
    return fail(ctx, RG_ERR_INVALID, "rg_mlp_forward: invalid arguments");
  if (n == 0) return RG_OK;
  RG_ENTER(ctx);
  const int chunk = 1 << 16;
  rg_status st = ensureBatch(p, std::min(n, chunk));
  if (st != RG_OK) return st;
  std::vector<float> f(chunk);
  for (int i0 = 0; i0 < n; i0 += chunk)
  {
    const int c = std::min(chunk, n - i0);
    const int rows = (c + FROWS - 1) / FROWS * FROWS;
    RG_HIP(ctx, hipMemsetAsync(p->frac, 0, size_t(rows) * 3 * sizeof(double), ctx->stream));
    RG_HIP(ctx, hipMemcpyAsync(p->frac, frac + 3 * size_t(i0), size_t(c) * 3 * sizeof(double), hipMemcpyHostToDevice,
                               ctx->stream));
    if ((st = forward(p, rows, false)) != RG_OK) return st;
    RG_HIP(ctx, hipMemcpyAsync(f.data(), p->f, c * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < c; ++i) logdensity[i0 + i] = f[i];
  }
  return RG_OK;
}

rg_status rg_mlp_gradients(rg_ctx* ctx, rg_mlp* p, int n, const double* frac, const double* energy, double beta,
                           float* grads, double* loss)
{
  if (!ctx || !p || n <= 0 || n % 64 != 0 || n > (1 << 16) || !frac || !energy || !grads || !loss)
    return fail(ctx, RG_ERR_INVALID, "rg_mlp_gradients: n must be a positive multiple of 64 up to 65536");
  RG_ENTER(ctx);
  rg_status st = ensureBatch(p, n);
  if (st != RG_OK) return st;
  RG_HIP(ctx, hipMemcpyAsync(p->frac, frac, size_t(n) * 3 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  RG_HIP(ctx, hipMemcpyAsync(p->U, energy, size_t(n) * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_zero, dim3((NPARAM + 255) / 256), dim3(256), 0, ctx->stream, p->g, int(NPARAM));
  if ((st = forward(p, n, true, beta)) != RG_OK) return st;
  if ((st = backward(p, n, beta)) != RG_OK) return st;
  RG_HIP(ctx, hipMemcpyAsync(grads, p->g, NPARAM * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  RG_HIP(ctx, hipMemcpyAsync(loss, p->loss, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RG_OK;
}

rg_status rg_mlp_train(rg_ctx* ctx, rg_mlp* p, int n, const double* frac, const double* energy, double beta,
                       int batch, int epochs, double lr, unsigned long long seed, double* epoch_loss)
{
  if (!ctx || !p || n <= 0 || !frac || !energy || batch <= 0 || batch % 64 != 0 || batch > (1 << 16) ||
      batch > n || epochs < 0)
    return fail(ctx, RG_ERR_INVALID, "rg_mlp_train: invalid arguments (batch: multiple of 64, at most 65536 and n)");
  RG_ENTER(ctx);
  rg_status st = ensureBatch(p, batch);
  if (st != RG_OK) return st;
  // the whole training set stays on the device; minibatches are gathered through a per-epoch permutation
  double *dfrac = nullptr, *dU = nullptr;
  int* dperm = nullptr;
  hipError_t e = hipMalloc(&dfrac, size_t(n) * 3 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&dU, size_t(n) * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&dperm, size_t(n) * sizeof(int));
  if (e == hipSuccess)
    e = hipMemcpyAsync(dfrac, frac, size_t(n) * 3 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dU, energy, size_t(n) * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
  // epoch ep trains on the ep + 1 times shuffled identity; the next epoch's permutation is shuffled on the host (into
  // the other of two host arrays) while the current epoch's kernels run
  std::vector<int> perm[2] = {std::vector<int>(n), std::vector<int>()};
  for (int i = 0; i < n; ++i) perm[0][i] = i;
  unsigned long long s = seed ? seed : 1;
  auto rnd = [&]()
  {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
  };
  auto shuffle = [&](std::vector<int>& v)
  {
    for (int i = n - 1; i > 0; --i) std::swap(v[i], v[rnd() % (unsigned long long)(i + 1)]);
  };
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const int nb = n / batch;
  // the gradient sums start from zero; k_adam clears them again after every step
  hipLaunchKernelGGL(k_zero, dim3((NPARAM + 255) / 256), dim3(256), 0, ctx->stream, p->g, int(NPARAM));
  if (epochs > 0)
  {
    shuffle(perm[0]);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dperm, perm[0].data(), size_t(n) * sizeof(int), hipMemcpyHostToDevice, ctx->stream);
  }
  for (int ep = 0; ep < epochs && e == hipSuccess && st == RG_OK; ++ep)
  {
    hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(64), 0, ctx->stream, p->loss, 2);
    for (int b = 0; b < nb; ++b)
    {
      if ((st = forward(p, batch, true, beta, dperm + size_t(b) * batch, dfrac, dU)) != RG_OK ||
          (st = backward(p, batch, beta)) != RG_OK)
        break;
      p->adamStep += 1;
      const float c1 = 1.f - std::pow(b1, float(p->adamStep)), c2 = 1.f - std::pow(b2, float(p->adamStep));
      hipLaunchKernelGGL(k_adam, dim3(ADAM_BLOCKS), dim3(256), 0, ctx->stream, p->w, p->g, p->m, p->v, p->wb, p->wt,
                         float(lr), b1, b2, eps, c1, c2);
    }
    e = hipGetLastError();
    if (e == hipSuccess && st == RG_OK && ep + 1 < epochs)
    {
      // the host array of the next epoch was last copied before the previous epoch's synchronisation; the device
      // copy is ordered after this epoch's kernels, which read dperm
      std::vector<int>& next = perm[(ep + 1) & 1];
      next = perm[ep & 1];
      shuffle(next);
      e = hipMemcpyAsync(dperm, next.data(), size_t(n) * sizeof(int), hipMemcpyHostToDevice, ctx->stream);
    }
    double l[2] = {0.0, 0.0};
    if (e == hipSuccess) e = hipMemcpyAsync(l, p->loss, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (epoch_loss) epoch_loss[ep] = l[1] / std::max(1, nb);
  }
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(dfrac);
  (void)hipFree(dU);
  (void)hipFree(dperm);
  if (st != RG_OK) return st;
  if (e != hipSuccess) return hipfail(ctx, e, "rg_mlp_train");
  return RG_OK;
}

// Voxel probabilities of the proposal: p_v proportional to (1 - floor) exp(f(centre_v)) / sum + floor / nvox.
rg_status rg_mlp_proposal(rg_ctx* ctx, rg_mlp* p, const int* n3, double floor_fraction, double* probability)
{
  if (!ctx || !p || !n3 || !probability || n3[0] <= 0 || n3[1] <= 0 || n3[2] <= 0 || floor_fraction <= 0.0 ||
      floor_fraction > 1.0)
    return fail(ctx, RG_ERR_INVALID, "rg_mlp_proposal: invalid arguments");
  const int nvox = n3[0] * n3[1] * n3[2];
  std::vector<double> frac(3 * size_t(nvox)), f(nvox);
  for (int v = 0; v < nvox; ++v)
  {
    const int x = v % n3[0], y = (v / n3[0]) % n3[1], z = v / (n3[0] * n3[1]);
    frac[3 * v] = (x + 0.5) / n3[0];
    frac[3 * v + 1] = (y + 0.5) / n3[1];
    frac[3 * v + 2] = (z + 0.5) / n3[2];
  }
  rg_status st = rg_mlp_forward(ctx, p, nvox, frac.data(), f.data());
  if (st != RG_OK) return st;
  double mx = -INFINITY;
  for (double x : f) mx = std::max(mx, x);
  double sum = 0.0;
  for (int v = 0; v < nvox; ++v) sum += std::exp(f[v] - mx);
  for (int v = 0; v < nvox; ++v)
    probability[v] = (1.0 - floor_fraction) * std::exp(f[v] - mx) / sum + floor_fraction / nvox;
  return RG_OK;
}

}  // extern "C"

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
// GEMM tile: 64 x 64 outputs per workgroup, four wavefronts of 32 x 32 (2 x 2 MFMA 16 x 16 x 32 blocks); the whole K
// extent of a stage (up to 256) is staged through LDS with the K index contiguous, so every MFMA operand is one
// 16-byte LDS read, and all of a stage's global loads are in flight together.
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

enum Epilogue
{
  EPI_BIAS_SILU = 0,  // Z = acc + bias (FP32 out), H = SiLU(Z) (BF16 out)
  EPI_STORE_F32 = 1,  // C = acc (FP32 out)
  EPI_DSILU = 2,      // dZ = acc * SiLU'(Zprev) (BF16 out), column sums of dZ added into outF (bias gradient)
  EPI_ATOMIC_F32 = 3  // C += acc (FP32, split-K partial products of the weight gradients)
};

__device__ __forceinline__ float sigmoidf(float z) { return 1.0f / (1.0f + __expf(-z)); }

// C (M x N) = op(A) (M x K) * op(B) (K x N).  TA: A is stored K x M; TB: B is stored N x K.  All row-major.
// Workgroup (x, y, z) computes the 64 x 64 tile (y, x) over k in [z Kc, min(K, (z + 1) Kc)).  K is taken in stages
// of up to KS = 256: every global load of a stage is issued before any of them is written to LDS, so a stage costs
// one memory latency (the network's K extents, 192 and 256, and the weight-gradient chunks are one stage each).
constexpr int KS = 256;
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, int Kc, const __bf16* __restrict__ A, int lda,
                                              const __bf16* __restrict__ B, int ldb, const float* __restrict__ bias,
                                              const float* __restrict__ Zprev, float* __restrict__ outF,
                                              __bf16* __restrict__ outH)
{
  __shared__ __attribute__((aligned(16))) __bf16 As[64][KS + 8];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[64][KS + 8];
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int kEnd = min(K, int(blockIdx.z + 1) * Kc);
  for (int k0 = blockIdx.z * Kc; k0 < kEnd; k0 += KS)
  {
    const int kk = min(KS, kEnd - k0);  // a multiple of 32
    const int nv = 8 * kk;               // 8-element vectors per operand tile (64 rows x kk)
    const int rowv = kk / 8;             // vectors along k per row of a k-contiguous operand
    bf16x8 va[KS / 32], vb[KS / 32];
#pragma unroll
    for (int q = 0; q < KS / 32; ++q)
    {
      const int v = t + 256 * q;
      if (v < nv)
      {
        va[q] = !TA ? *reinterpret_cast<const bf16x8*>(A + size_t(m0 + v / rowv) * lda + k0 + (v % rowv) * 8)
                    : *reinterpret_cast<const bf16x8*>(A + size_t(k0 + (v >> 3)) * lda + m0 + (v & 7) * 8);
        vb[q] = !TB ? *reinterpret_cast<const bf16x8*>(B + size_t(k0 + (v >> 3)) * ldb + n0 + (v & 7) * 8)
                    : *reinterpret_cast<const bf16x8*>(B + size_t(n0 + v / rowv) * ldb + k0 + (v % rowv) * 8);
      }
    }
#pragma unroll
    for (int q = 0; q < KS / 32; ++q)
    {
      const int v = t + 256 * q;
      if (v < nv)
      {
        if (!TA)
          *reinterpret_cast<bf16x8*>(&As[v / rowv][(v % rowv) * 8]) = va[q];
        else
        {
#pragma unroll
          for (int i = 0; i < 8; ++i) As[(v & 7) * 8 + i][v >> 3] = va[q][i];
        }
        if (!TB)
        {
#pragma unroll
          for (int i = 0; i < 8; ++i) Bs[(v & 7) * 8 + i][v >> 3] = vb[q][i];
        }
        else
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
  float csum[2] = {0.f, 0.f};  // EPI_DSILU: this lane's share of the column sums of dZ
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
      {
        const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
        const int n = n0 + wn + 16 * j + (lane & 15);
        const size_t o = size_t(m) * N + n;
        const float v = acc[i][j][r];
        if (EPI == EPI_BIAS_SILU)
        {
          const float z = v + bias[n];
          outF[o] = z;
          outH[o] = (__bf16)(z * sigmoidf(z));
        }
        else if (EPI == EPI_STORE_F32)
          outF[o] = v;
        else if (EPI == EPI_ATOMIC_F32)
          atomicAdd(&outF[o], v);
        else
        {
          const float z = Zprev[o];
          const float s = sigmoidf(z);
          const __bf16 h = (__bf16)(v * s * (1.0f + z * (1.0f - s)));
          outH[o] = h;
          csum[j] += float(h);
        }
      }
  if (EPI == EPI_DSILU)
  {
    // bias gradient: the tile's column sums (lanes l, l + 16, l + 32, l + 48 hold column l), added into outF[n]
#pragma unroll
    for (int j = 0; j < 2; ++j)
    {
      float cs = csum[j];
      cs += __shfl_xor(cs, 16, 64);
      cs += __shfl_xor(cs, 32, 64);
      if (lane < 16) atomicAdd(&outF[n0 + wn + 16 * j + lane], cs);
    }
  }
}

// Fourier features of fractional coordinates
__global__ void k_features(int n, const double* __restrict__ frac, __bf16* __restrict__ X)
{
  const int i = blockIdx.x * blockDim.y + threadIdx.y;
  if (i >= n) return;
  for (int q = threadIdx.x; q < NIN; q += blockDim.x)
  {
    // the phase is reduced to one period in FP64 (m s - floor(m s)), then evaluated in FP32, whose 1e-7 accuracy
    // is far below the BF16 resolution of the stored feature
    const int a = q / (2 * NFREQ), m = (q % (2 * NFREQ)) / 2 + 1, c = q & 1;
    const double ms = m * frac[3 * i + a];
    const float ph = float(2.0 * M_PI * (ms - floor(ms)));
    float sn, cs;
    sincosf(ph, &sn, &cs);
    X[size_t(i) * NIN + q] = (__bf16)(c ? cs : sn);
  }
}

// f_i = H3_i . w4 + b4 (one wavefront per row)
__global__ void k_out_forward(int n, const __bf16* __restrict__ H, const __bf16* __restrict__ w4,
                              const float* __restrict__ b4, float* __restrict__ f)
{
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  float s = 0.f;
  for (int j = lane; j < WIDTH; j += 64) s += float(H[size_t(i) * WIDTH + j]) * float(w4[j]);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) f[i] = s + b4[0];
}

// One workgroup: p = softmax(-beta U), q = softmax(f), df = q - p, loss = -sum p log q
__global__ __launch_bounds__(1024) void k_loss(int n, const float* __restrict__ f, const double* __restrict__ U,
                                               double beta, float* __restrict__ df, double* __restrict__ loss)
{
  __shared__ double red[16];
  const int t = threadIdx.x;
  auto blockMax = [&](double v)
  {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    double r = -INFINITY;
    for (int i = 0; i < int(blockDim.x >> 6); ++i) r = fmax(r, red[i]);
    __syncthreads();
    return r;
  };
  auto blockSum = [&](double v)
  {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    double r = 0.0;
    for (int i = 0; i < int(blockDim.x >> 6); ++i) r += red[i];
    __syncthreads();
    return r;
  };
  double mf = -INFINITY, mp = -INFINITY;
  for (int i = t; i < n; i += blockDim.x)
  {
    mf = fmax(mf, double(f[i]));
    mp = fmax(mp, -beta * U[i]);
  }
  mf = blockMax(mf);
  mp = blockMax(mp);
  double sf = 0.0, sp = 0.0;
  for (int i = t; i < n; i += blockDim.x)
  {
    sf += exp(double(f[i]) - mf);
    sp += exp(-beta * U[i] - mp);
  }
  sf = blockSum(sf);
  sp = blockSum(sp);
  const double lse = mf + log(sf);
  double l = 0.0;
  for (int i = t; i < n; i += blockDim.x)
  {
    const double p = exp(-beta * U[i] - mp) / sp;
    const double q = exp(double(f[i]) - lse);
    df[i] = float(q - p);
    l -= p * (double(f[i]) - lse);
  }
  l = blockSum(l);
  if (t == 0)
  {
    loss[0] = l;
    loss[1] += l;
  }
}

// Output layer backward over a chunk of OB_ROWS rows per workgroup, one thread per hidden unit j:
// dw4_j += sum_i H3_ij df_i, db4 += sum_i df_i, dZ3_ij = df_i w4_j SiLU'(Z3_ij) (BF16), db3_j += sum_i dZ3_ij.
constexpr int OB_ROWS = 32;
__global__ void k_out_backward(int n, const __bf16* __restrict__ H, const float* __restrict__ Z,
                               const __bf16* __restrict__ w4, const float* __restrict__ df, float* __restrict__ gw4,
                               float* __restrict__ gb4, __bf16* __restrict__ dZ, float* __restrict__ gb3)
{
  const int j = threadIdx.x;  // blockDim = WIDTH
  const int i0 = blockIdx.x * OB_ROWS;
  const int i1 = min(n, i0 + OB_ROWS);
  float g = 0.f, gb = 0.f, cs = 0.f;
  const float wj = float(w4[j]);
#pragma unroll 4
  for (int i = i0; i < i1; ++i)
  {
    const size_t o = size_t(i) * WIDTH + j;
    const float d = df[i];
    g += float(H[o]) * d;
    gb += d;
    const float z = Z[o], s = sigmoidf(z);
    const __bf16 h = (__bf16)(d * wj * s * (1.0f + z * (1.0f - s)));
    dZ[o] = h;
    cs += float(h);
  }
  atomicAdd(&gw4[j], g);
  atomicAdd(&gb3[j], cs);
  if (j == 0) atomicAdd(gb4, gb);
}

__global__ void k_zero(float* p, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

// Adam (Kingma & Ba, 2015) on the FP32 masters, then the BF16 copies used by the products
__global__ void k_adam(int n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, __bf16* __restrict__ wb, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = (__bf16)wi;
}

// minibatch rows gathered from the device-resident training set through a permutation
__global__ void k_gather(int n, const int* __restrict__ perm, const double* __restrict__ frac,
                         const double* __restrict__ U, double* __restrict__ bfrac, double* __restrict__ bU)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int q = perm[i];
  bfrac[3 * i] = frac[3 * size_t(q)];
  bfrac[3 * i + 1] = frac[3 * size_t(q) + 1];
  bfrac[3 * i + 2] = frac[3 * size_t(q) + 2];
  bU[i] = U[q];
}

__global__ void k_zero_d(double* p, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

__global__ void k_to_bf16(int n, const float* __restrict__ a, __bf16* __restrict__ b)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (__bf16)a[i];
}

}  // namespace
}  // namespace rg

using namespace rg;

// parameter layout in one FP32 array: W1 (192x256), b1, W2 (256x256), b2, W3 (256x256), b3, w4 (256), b4
namespace
{
constexpr size_t OFF_W1 = 0;
constexpr size_t OFF_B1 = OFF_W1 + size_t(NIN) * WIDTH;
constexpr size_t OFF_W2 = OFF_B1 + WIDTH;
constexpr size_t OFF_B2 = OFF_W2 + size_t(WIDTH) * WIDTH;
constexpr size_t OFF_W3 = OFF_B2 + WIDTH;
constexpr size_t OFF_B3 = OFF_W3 + size_t(WIDTH) * WIDTH;
constexpr size_t OFF_W4 = OFF_B3 + WIDTH;
constexpr size_t OFF_B4 = OFF_W4 + WIDTH;
constexpr size_t NPARAM = OFF_B4 + 1;
}  // namespace

struct rg_mlp
{
  rg_ctx* ctx = nullptr;
  int device = 0;
  float* w = nullptr;   // FP32 masters
  __bf16* wb = nullptr; // BF16 copies
  float* g = nullptr;   // gradients
  float* m = nullptr;   // Adam moments
  float* v = nullptr;
  long long adamStep = 0;
  // activations for the current batch capacity
  int cap = 0;
  __bf16 *X = nullptr, *H1 = nullptr, *H2 = nullptr, *H3 = nullptr, *dZa = nullptr, *dZb = nullptr;
  float *Z1 = nullptr, *Z2 = nullptr, *Z3 = nullptr, *f = nullptr, *df = nullptr;
  double *frac = nullptr, *U = nullptr, *loss = nullptr;
};

namespace
{

void freeBatch(rg_mlp* p)
{
  void* bufs[] = {p->X, p->H1, p->H2, p->H3, p->dZa, p->dZb, p->Z1, p->Z2, p->Z3, p->f, p->df, p->frac, p->U, p->loss};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  p->X = p->H1 = p->H2 = p->H3 = p->dZa = p->dZb = nullptr;
  p->Z1 = p->Z2 = p->Z3 = p->f = p->df = nullptr;
  p->frac = p->U = p->loss = nullptr;
  p->cap = 0;
}

rg_status ensureBatch(rg_mlp* p, int n)
{
  rg_ctx* ctx = p->ctx;
  const int need = (n + 63) / 64 * 64;
  if (p->cap >= need) return RG_OK;
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  freeBatch(p);
  const size_t nb = need;
  RG_HIP(ctx, hipMalloc(&p->X, nb * NIN * sizeof(__bf16)));
  RG_HIP(ctx, hipMalloc(&p->H1, nb * WIDTH * sizeof(__bf16)));
  RG_HIP(ctx, hipMalloc(&p->H2, nb * WIDTH * sizeof(__bf16)));
  RG_HIP(ctx, hipMalloc(&p->H3, nb * WIDTH * sizeof(__bf16)));
  RG_HIP(ctx, hipMalloc(&p->dZa, nb * WIDTH * sizeof(__bf16)));
  RG_HIP(ctx, hipMalloc(&p->dZb, nb * WIDTH * sizeof(__bf16)));
  RG_HIP(ctx, hipMalloc(&p->Z1, nb * WIDTH * sizeof(float)));
  RG_HIP(ctx, hipMalloc(&p->Z2, nb * WIDTH * sizeof(float)));
  RG_HIP(ctx, hipMalloc(&p->Z3, nb * WIDTH * sizeof(float)));
  RG_HIP(ctx, hipMalloc(&p->f, nb * sizeof(float)));
  RG_HIP(ctx, hipMalloc(&p->df, nb * sizeof(float)));
  RG_HIP(ctx, hipMalloc(&p->frac, nb * 3 * sizeof(double)));
  RG_HIP(ctx, hipMalloc(&p->U, nb * sizeof(double)));
  RG_HIP(ctx, hipMalloc(&p->loss, 8 * sizeof(double)));
  p->cap = need;
  return RG_OK;
}

// forward pass of nrow (multiple of 64) rows whose fractional coordinates are in p->frac
rg_status forward(rg_mlp* p, int nrow)
{
  hipStream_t st = p->ctx->stream;
  const dim3 blk(256);
  hipLaunchKernelGGL(k_features, dim3((nrow + 3) / 4), dim3(64, 4), 0, st, nrow, p->frac, p->X);
  const dim3 grid(WIDTH / 64, nrow / 64);
  hipLaunchKernelGGL((k_gemm<false, false, EPI_BIAS_SILU>), grid, blk, 0, st, nrow, WIDTH, NIN, NIN, p->X, NIN,
                     p->wb + OFF_W1, WIDTH, p->w + OFF_B1, nullptr, p->Z1, p->H1);
  hipLaunchKernelGGL((k_gemm<false, false, EPI_BIAS_SILU>), grid, blk, 0, st, nrow, WIDTH, WIDTH, WIDTH, p->H1, WIDTH,
                     p->wb + OFF_W2, WIDTH, p->w + OFF_B2, nullptr, p->Z2, p->H2);
  hipLaunchKernelGGL((k_gemm<false, false, EPI_BIAS_SILU>), grid, blk, 0, st, nrow, WIDTH, WIDTH, WIDTH, p->H2, WIDTH,
                     p->wb + OFF_W3, WIDTH, p->w + OFF_B3, nullptr, p->Z3, p->H3);
  hipLaunchKernelGGL(k_out_forward, dim3((nrow + 3) / 4), blk, 0, st, nrow, p->H3, p->wb + OFF_W4, p->w + OFF_B4, p->f);
  RG_HIP(p->ctx, hipGetLastError());
  return RG_OK;
}

// Weight gradient G (M x WIDTH) += A^T dZ over n rows, A stored n x M, split over K chunks of kc rows whose partial
// products are added atomically.  The chunk is tuned online among 256 .. 4096 rows (results differ only in the
// order of the FP32 atomic additions, which is unordered anyway).
rg_status weightGrad(rg_ctx* ctx, int M, int n, const __bf16* A, const __bf16* dZ, float* G)
{
  static const int kChunks[5] = {256, 512, 1024, 2048, 4096};
  auto launch = [&](int kc)
  {
    kc = std::min(kc, n);
    const dim3 grid(WIDTH / 64, M / 64, (n + kc - 1) / kc);
    hipLaunchKernelGGL((k_gemm<true, false, EPI_ATOMIC_F32>), grid, dim3(256), 0, ctx->stream, M, WIDTH, n, kc, A, M,
                       dZ, WIDTH, nullptr, nullptr, G, nullptr);
    return hipGetLastError();
  };
  Autotuner& tuner = Autotuner::instance();
  const std::string key = "mlp_weight_grad/M=" + std::to_string(M) + "/n=" + std::to_string(n);
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

// loss and gradients of a batch of n rows (n multiple of 64) already in p->frac / p->U
rg_status backward(rg_mlp* p, int n, double beta)
{
  hipStream_t st = p->ctx->stream;
  const dim3 blk(256);
  hipLaunchKernelGGL(k_zero, dim3((NPARAM + 255) / 256), blk, 0, st, p->g, int(NPARAM));
  hipLaunchKernelGGL(k_loss, dim3(1), dim3(1024), 0, st, n, p->f, p->U, beta, p->df, p->loss);
  hipLaunchKernelGGL(k_out_backward, dim3((n + OB_ROWS - 1) / OB_ROWS), dim3(WIDTH), 0, st, n, p->H3, p->Z3,
                     p->wb + OFF_W4, p->df, p->g + OFF_W4, p->g + OFF_B4, p->dZa, p->g + OFF_B3);
  // layer 3: dW3 = H2^T dZ3; dZ2 = (dZ3 W3^T) * SiLU'(Z2) with db2 = column sums of dZ2
  rg_status ws;
  if ((ws = weightGrad(p->ctx, WIDTH, n, p->H2, p->dZa, p->g + OFF_W3)) != RG_OK) return ws;
  hipLaunchKernelGGL((k_gemm<false, true, EPI_DSILU>), dim3(WIDTH / 64, n / 64), blk, 0, st, n, WIDTH, WIDTH, WIDTH,
                     p->dZa, WIDTH, p->wb + OFF_W3, WIDTH, nullptr, p->Z2, p->g + OFF_B2, p->dZb);
  // layer 2: dW2 = H1^T dZ2; dZ1 = (dZ2 W2^T) * SiLU'(Z1) with db1
  if ((ws = weightGrad(p->ctx, WIDTH, n, p->H1, p->dZb, p->g + OFF_W2)) != RG_OK) return ws;
  hipLaunchKernelGGL((k_gemm<false, true, EPI_DSILU>), dim3(WIDTH / 64, n / 64), blk, 0, st, n, WIDTH, WIDTH, WIDTH,
                     p->dZb, WIDTH, p->wb + OFF_W2, WIDTH, nullptr, p->Z1, p->g + OFF_B1, p->dZa);
  // layer 1: dW1 = X^T dZ1
  if ((ws = weightGrad(p->ctx, NIN, n, p->X, p->dZa, p->g + OFF_W1)) != RG_OK) return ws;
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
  hipLaunchKernelGGL(k_to_bf16, dim3((NPARAM + 255) / 256), dim3(256), 0, ctx->stream, int(NPARAM), p->w, p->wb);
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
  void* bufs[] = {p->w, p->wb, p->g, p->m, p->v};
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
  hipLaunchKernelGGL(k_to_bf16, dim3((NPARAM + 255) / 256), dim3(256), 0, ctx->stream, int(NPARAM), p->w, p->wb);
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
    const int rows = (c + 63) / 64 * 64;
    RG_HIP(ctx, hipMemsetAsync(p->frac, 0, size_t(rows) * 3 * sizeof(double), ctx->stream));
    RG_HIP(ctx, hipMemcpyAsync(p->frac, frac + 3 * size_t(i0), size_t(c) * 3 * sizeof(double), hipMemcpyHostToDevice,
                               ctx->stream));
    if ((st = forward(p, rows)) != RG_OK) return st;
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
  if ((st = forward(p, n)) != RG_OK) return st;
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
  std::vector<int> perm(n);
  for (int i = 0; i < n; ++i) perm[i] = i;
  unsigned long long s = seed ? seed : 1;
  auto rnd = [&]()
  {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
  };
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const int nb = n / batch;
  for (int ep = 0; ep < epochs && e == hipSuccess && st == RG_OK; ++ep)
  {
    for (int i = n - 1; i > 0; --i) std::swap(perm[i], perm[rnd() % (unsigned long long)(i + 1)]);
    // the previous epoch's kernels read dperm: the copy is ordered after them on the same stream
    e = hipMemcpyAsync(dperm, perm.data(), size_t(n) * sizeof(int), hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) break;
    hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(64), 0, ctx->stream, p->loss, 2);
    for (int b = 0; b < nb; ++b)
    {
      hipLaunchKernelGGL(k_gather, dim3((batch + 255) / 256), dim3(256), 0, ctx->stream, batch,
                         dperm + size_t(b) * batch, dfrac, dU, p->frac, p->U);
      if ((st = forward(p, batch)) != RG_OK || (st = backward(p, batch, beta)) != RG_OK) break;
      p->adamStep += 1;
      const float c1 = 1.f - std::pow(b1, float(p->adamStep)), c2 = 1.f - std::pow(b2, float(p->adamStep));
      hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, ctx->stream, int(NPARAM), p->w, p->g, p->m,
                         p->v, p->wb, float(lr), b1, b2, eps, c1, c2);
    }
    e = hipGetLastError();
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

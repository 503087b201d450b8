// DeepSets warm-start network for the Sinkhorn solver: forward, backward and
// Adam kernels (gfx950).
//
// OTT's learned initializer (initializers/neural/meta_initializer.py:35-213)
// is an MLP on the weight vectors of a fixed geometry, trained by flax/optax.
// This network instead reads the point clouds themselves:
//   encoder  phi: R^d -> R^256, 3 layers (ReLU, ReLU, linear), shared by x and y
//   pooling  e_x = mean_i phi(x~_i),  e_y = mean_j phi(y~_j)
//   context  ctx = [e_x, e_y, kappa, s_x - 1, s_y - 1, mean y~, 0 ...] (CTX wide)
//   head     g~_j = psi(y~_j, ctx), 3 layers (ReLU, ReLU, linear),
//            plus a context-conditioned linear + quadratic term z . in_j with
//            z = Wq ctx and in_j = [y~_j, |y~_j|^2 - 1, 0 ...] the head input
// with x~ = (x - mu) / sigma, y~ = (y - mu) / sigma per problem
// (mu = (mean x + mean y) / 2, sigma^2 = mean(C) / 2), kappa = eps / sigma^2,
// s_x = mean |x~|^2, s_y = mean |y~|^2 (exact moments over every point; the
// mean of x~ is minus that of y~).  The quadratic term carries the part of the
// potential that is quadratic in the point (exact for Gaussian clouds); the MLP
// the rest.  The prediction is g_j = sigma^2 * g~_j (a potential, up to a
// constant).
//
// Every matrix product runs through one bf16 MFMA GEMM, C = A . Bt^T with both
// operands K-contiguous ("NT"), f32 accumulation, and an epilogue that adds a
// column bias and/or a per-row-group bias, applies ReLU or a ReLU mask, and
// writes f32, bf16 and transposed-bf16 copies of the result as requested.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/ottx.h"

namespace ottx_ds {

typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4;
typedef uint16_t bf16_t;

__device__ __forceinline__ bf16_t to_bf16(float v) {
  uint32_t u = __float_as_uint(v);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<bf16_t>(u >> 16);
}
__device__ __forceinline__ float from_bf16(bf16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// ---------------------------------------------------------------------------
// GEMM: C[M, N] (+)= A[M, K] . Bt[N, K]^T.  M, N multiples of 64, K multiple
// of 32 (per split).  Tile 64 x 64, 4 waves each owning 32 x 32.
// ---------------------------------------------------------------------------
struct GemmArgs {
  const bf16_t* A;
  const bf16_t* Bt;
  int64_t lda, ldb;      // row strides (elements)
  int M, N, K;           // K = reduction length of one split
  int splits;            // gridDim.z; split z covers k in [z K, (z + 1) K)
  const float* bias;     // [N] column bias, or null
  const float* gbias;    // [groups, N] per-row-group bias, or null
  const int* row_group;  // [M] group of each row (for gbias), or null
  int relu;              // apply max(0, .) after the biases
  const bf16_t* mask;    // [M, ldm] multiply by (mask > 0), or null
  int64_t ldm;
  float* Cf;             // f32 output [M, ldc] (or [splits][M, ldc] partials)
  int64_t ldc;
  bf16_t* Cb;            // bf16 output [M, ldcb], or null
  int64_t ldcb;
  bf16_t* CbT;           // bf16 transposed output [N, ldct], or null
  int64_t ldct;
  int valid_rows;        // rows >= valid_rows are written as zero
};

// K is staged through LDS in chunks of KC (NBUF buffers): the loads of a
// whole chunk are issued before any is used, and with two buffers the next
// chunk's loads are in flight while the current one is on MFMA.
template <int KC, int NBUF>
__global__ __launch_bounds__(256) void gemm_nt_kernel(const GemmArgs g) {
  constexpr int LDS_STRIDE = KC + 8;  // bf16 per LDS row (padding staggers the banks)
  constexpr int PER = KC / 32;        // 16-byte groups per thread per operand and chunk
  __shared__ bf16_t As[NBUF][64 * LDS_STRIDE];
  __shared__ bf16_t Bs[NBUF][64 * LDS_STRIDE];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int64_t kbase = static_cast<int64_t>(blockIdx.z) * g.K;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  u32x4 ra[PER], rb[PER];
  // chunk kc holds kn = min(KC, K - kc) columns (a multiple of 32): kn / 8
  // groups per row, kn / 32 per thread
  auto load = [&](int kc) {
    const int kn = min(KC, g.K - kc), per_row = kn / 8;
#pragma unroll
    for (int i = 0; i < PER; ++i)
      if (i < kn / 32) {
        const int gi = tid + 256 * i, r = gi / per_row, c8 = gi % per_row;
        ra[i] = *reinterpret_cast<const u32x4*>(g.A + (m0 + r) * g.lda + kbase + kc + c8 * 8);
        rb[i] = *reinterpret_cast<const u32x4*>(g.Bt + (n0 + r) * g.ldb + kbase + kc + c8 * 8);
      }
  };
  load(0);
  for (int kc = 0, ci = 0; kc < g.K; kc += KC, ++ci) {
    const int buf = NBUF == 1 ? 0 : (ci & 1);
    const int kn = min(KC, g.K - kc), per_row = kn / 8;
    if (NBUF == 1) __syncthreads();  // the previous chunk's reads of the buffer are done
#pragma unroll
    for (int i = 0; i < PER; ++i)
      if (i < kn / 32) {
        const int gi = tid + 256 * i, r = gi / per_row, c8 = gi % per_row;
        *reinterpret_cast<u32x4*>(&As[buf][r * LDS_STRIDE + c8 * 8]) = ra[i];
        *reinterpret_cast<u32x4*>(&Bs[buf][r * LDS_STRIDE + c8 * 8]) = rb[i];
      }
    __syncthreads();
    if (kc + KC < g.K) load(kc + KC);
    for (int k0 = 0; k0 < kn; k0 += 32) {
      bf16x8 af[2], bf[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        af[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(
            &As[buf][(wr + i * 16 + (lane & 15)) * LDS_STRIDE + k0 + 8 * (lane >> 4)]));
        bf[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(
            &Bs[buf][(wc + i * 16 + (lane & 15)) * LDS_STRIDE + k0 + 8 * (lane >> 4)]));
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
  }
  // Epilogue: element (row 4 (lane / 16) + r, column lane % 16) of each block.
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wr + i * 16 + 4 * (lane >> 4) + r;
        const int col = n0 + wc + j * 16 + (lane & 15);
        float v = acc[i][j][r];
        if (g.splits > 1) {
          g.Cf[static_cast<int64_t>(blockIdx.z) * g.M * g.ldc + static_cast<int64_t>(row) * g.ldc + col] = v;
          continue;
        }
        if (g.bias) v += g.bias[col];
        if (g.gbias) v += g.gbias[static_cast<int64_t>(g.row_group[row]) * g.N + col];
        if (g.relu) v = fmaxf(v, 0.f);
        if (g.mask && !(from_bf16(g.mask[static_cast<int64_t>(row) * g.ldm + col]) > 0.f)) v = 0.f;
        if (row >= g.valid_rows) v = 0.f;
        if (g.Cf) g.Cf[static_cast<int64_t>(row) * g.ldc + col] = v;
        if (g.Cb) g.Cb[static_cast<int64_t>(row) * g.ldcb + col] = to_bf16(v);
        if (g.CbT) g.CbT[static_cast<int64_t>(col) * g.ldct + row] = to_bf16(v);
      }
}

// Sum of split-K partials: out[i] = sum_z part[z][i] (deterministic order).
__global__ void sum_splits_kernel(const float* __restrict__ part, int splits, int64_t count, float* __restrict__ out,
                                  int accumulate) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float s = 0.f;
  for (int z = 0; z < splits; ++z) s += part[z * count + i];
  out[i] = accumulate ? out[i] + s : s;
}

// ---------------------------------------------------------------------------
// Normalisation statistics per problem: mu = (mean x + mean y) / 2 and
// sigma^2 = mean(C) / 2 with mean(C) = mean|x|^2 + mean|y|^2 - 2 <mean x, mean y>.
// One workgroup per problem (thread per feature, rows strided).
// ---------------------------------------------------------------------------
constexpr int kStatChunks = 128;

// Per-(problem, chunk) float64 partial sums of the coordinates and their
// squares, for x and y: part[((b * kStatChunks + c) * 4 + q) * d + k] with
// q = sum x, sum x^2, sum y, sum y^2.  Grid (kStatChunks, B).
__global__ void stats_partial_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                     const int* __restrict__ xoff, const int* __restrict__ yoff, int d,
                                     double* __restrict__ part) {
  __shared__ double fs[4][256];
  const int c = blockIdx.x, b = blockIdx.y, t = threadIdx.x;
  const int G = max(1, 256 / d);
  const int k = t % d, gi = t / d;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (t < G * d) {
    // the chunk's rows of both clouds, U rows of each per trip: the sums are
    // short, so they are bound by latency, and 2 U loads are in flight
    constexpr int U = 8;
    int lo[2], hi[2];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int* off = side ? yoff : xoff;
      const int r0 = off[b], r1 = off[b + 1];
      const int per = (r1 - r0 + kStatChunks - 1) / kStatChunks;
      lo[side] = r0 + c * per + gi;
      hi[side] = min(r1, r0 + c * per + per);
    }
    for (int i = 0; lo[0] + i * G < hi[0] || lo[1] + i * G < hi[1]; i += U) {
      float v[2][U];
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        const float* p = side ? y : x;
#pragma unroll
        for (int u = 0; u < U; ++u) {
          // rows past the chunk load its last row and are zeroed after the load,
          // so every load is unconditional (hi > lo when the chunk has rows)
          const int r = lo[side] + (i + u) * G;
          const float w = p[static_cast<int64_t>(min(r, max(hi[side] - 1, 0))) * d + k];
          v[side][u] = r < hi[side] ? w : 0.f;
        }
      }
#pragma unroll
      for (int side = 0; side < 2; ++side)
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const double w = v[side][u];
          acc[2 * side] += w;
          acc[2 * side + 1] += w * w;
        }
    }
  }
  for (int q = 0; q < 4; ++q) fs[q][t] = acc[q];
  __syncthreads();
  if (t < d) {
    for (int q = 0; q < 4; ++q) {
      double s = 0.0;
      for (int g = 0; g < G; ++g) s += fs[q][g * d + t];
      part[((static_cast<int64_t>(b) * kStatChunks + c) * 4 + q) * d + t] = s;
    }
  }
}

// Context width: [e_x (256), e_y (256), kappa, s_x - 1, s_y - 1, mean y~ (d)]
// padded to a multiple of 32.
__host__ __device__ inline int ctx_width(int d) { return 512 + (3 + d + 31) / 32 * 32; }

// mu_b = (mean x + mean y) / 2, sigma_b^2 = mean(C) / 2 with
// mean(C) = var x + var y + |mean x - mean y|^2 (var = mean |p - mean p|^2), and
// the moments mom_b = [s_x - 1, s_y - 1, mean y~ (d)] with
// s_x = (var x + |mean x - mean y|^2 / 4) / sigma^2 (s_y alike) and
// mean y~ = (mean y - mean x) / (2 sigma).  Grid B of kFinalThreads threads:
// thread (value v = (q, k), group g) sums the chunks c = g mod G of one of the
// 4 d values, all its loads in flight at once, then the groups are reduced.
constexpr int kFinalThreads = 1024;

__global__ __launch_bounds__(kFinalThreads) void stats_final_kernel(
    const double* __restrict__ part, const int* __restrict__ xoff, const int* __restrict__ yoff, int d,
    float* __restrict__ mu, float* __restrict__ sig2, float* __restrict__ mom) {
  __shared__ double red[kFinalThreads];
  __shared__ double sums[4][256];
  __shared__ double tot[3][256];
  const int b = blockIdx.x, t = threadIdx.x;
  const int nx = xoff[b + 1] - xoff[b], ny = yoff[b + 1] - yoff[b];
  const int V = 4 * d;  // values per chunk, in part's order (q, k)
  const int G = max(1, kFinalThreads / V);
  const int v = t % V, g = t / V;
  double a = 0.0;
  if (t < G * V) {
    const double* pb = part + static_cast<int64_t>(b) * kStatChunks * V + v;
    constexpr int U = 16;
    for (int c = g; c < kStatChunks; c += U * G) {
      double w[U];
#pragma unroll
      for (int u = 0; u < U; ++u) w[u] = c + u * G < kStatChunks ? pb[static_cast<int64_t>(c + u * G) * V] : 0.0;
#pragma unroll
      for (int u = 0; u < U; ++u) a += w[u];
    }
  }
  red[t] = a;
  __syncthreads();
  for (int e = t; e < V; e += kFinalThreads) {
    double s2 = 0.0;
    for (int gg = 0; gg < G; ++gg) s2 += red[gg * V + e];
    sums[e / d][e % d] = s2;
  }
  __syncthreads();
  double va = 0.0, vb = 0.0, dm = 0.0, mxk = 0.0, myk = 0.0;
  if (t < d) {
    mxk = sums[0][t] / nx;
    myk = sums[2][t] / ny;
    va = sums[1][t] / nx - mxk * mxk;
    vb = sums[3][t] / ny - myk * myk;
    dm = (mxk - myk) * (mxk - myk);
  }
  if (t < 256) {
    tot[0][t] = va;
    tot[1][t] = vb;
    tot[2][t] = dm;
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w)
#pragma unroll
      for (int j = 0; j < 3; ++j) tot[j][t] += tot[j][t + w];
    __syncthreads();
  }
  const double A = tot[0][0], Bv = tot[1][0], Cc = tot[2][0];
  const double s2 = fmax(0.5 * (A + Bv + Cc), 1e-30);
  const double inv_sig = 1.0 / sqrt(s2);
  if (t < d) {
    mu[b * d + t] = static_cast<float>(0.5 * (mxk + myk));
    mom[static_cast<int64_t>(b) * (d + 2) + 2 + t] = static_cast<float>(0.5 * (myk - mxk) * inv_sig);
  }
  if (t == 0) {
    sig2[b] = static_cast<float>(s2);
    mom[static_cast<int64_t>(b) * (d + 2)] = static_cast<float>((A + 0.25 * Cc) / s2 - 1.0);
    mom[static_cast<int64_t>(b) * (d + 2) + 1] = static_cast<float>((Bv + 0.25 * Cc) / s2 - 1.0);
  }
}

// Rows to normalise into bf16 network inputs: encoder rows (src / seg_of_row
// give the point and its problem and side) or head rows (prob gives the
// problem of target point r).
struct NormJob {
  const int* src;
  const int* seg_of_row;
  const int* prob;
  int64_t rows, rows_pad;
  bf16_t* X;
  bf16_t* XT;  // transposed copy for the weight gradients, or null
};

// Both jobs in one launch: blocks [0, blocks0) take job 0, the rest job 1.
__global__ void normalize_kernel(const float* __restrict__ x, const float* __restrict__ y, NormJob j0, NormJob j1,
                                 int blocks0, int d, int kpad, const float* __restrict__ mu,
                                 const float* __restrict__ sig2) {
  // one 64-lane wave per row: coalesced loads and stores, wave-reduced |p~|^2
  const bool second = static_cast<int>(blockIdx.x) >= blocks0;
  const NormJob J = second ? j1 : j0;
  const int lane = threadIdx.x & 63;
  const int64_t blk = second ? blockIdx.x - blocks0 : blockIdx.x;
  const int64_t r = blk * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (r >= J.rows_pad) return;
  const float* p = nullptr;
  int b = 0;
  if (r < J.rows) {
    if (J.src) {
      const int seg = J.seg_of_row[r];
      b = seg >> 1;
      p = ((seg & 1) ? y : x) + static_cast<int64_t>(J.src[r]) * d;
    } else {
      b = J.prob[r];
      p = y + r * d;
    }
  }
  const float inv = p ? rsqrtf(sig2[b]) : 0.f;
  float part = 0.f;
  for (int k = lane; k < d; k += 64) {
    if (p) {
      const float v = (p[k] - mu[b * d + k]) * inv;
      part = fmaf(v, v, part);
    }
  }
  for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
  for (int k = lane; k < kpad; k += 64) {
    float v = 0.f;
    if (p) {
      if (k < d) v = (p[k] - mu[b * d + k]) * inv;
      else if (k == d) v = part - 1.f;
    }
    const bf16_t h = to_bf16(v);
    J.X[r * kpad + k] = h;
    if (J.XT) J.XT[static_cast<int64_t>(k) * J.rows_pad + r] = h;
  }
}

// Mean pooling of the encoder output H [P, 256] per (problem, side), with
// float partial sums over row chunks: part[(b * 2 + side) * chunks + c][256].
__global__ void pool_partial_kernel(const float* __restrict__ H, const int* __restrict__ seg_off, int chunks,
                                    float* __restrict__ part) {
  const int seg = blockIdx.y, c = blockIdx.x, f = threadIdx.x;
  const int r0 = seg_off[seg], r1 = seg_off[seg + 1];
  const int per = (r1 - r0 + chunks - 1) / chunks;
  const int a = r0 + c * per, e = min(r1, a + per);
  // U rows per trip, every load unconditional (rows past the chunk reload its
  // last row and count zero): the sum is short and bound by latency
  constexpr int U = 16;
  float sum = 0.f;
  for (int r = a; r < e; r += U) {
    float v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const float w = H[static_cast<int64_t>(min(r + u, e - 1)) * 256 + f];
      v[u] = r + u < e ? w : 0.f;
    }
#pragma unroll
    for (int u = 0; u < U; u += 2) sum += v[u] + v[u + 1];
  }
  part[(static_cast<int64_t>(seg) * chunks + c) * 256 + f] = sum;
}

// The context vector of problem b and its heads, in one launch:
//   ctx[b] = [e_x (256), e_y (256), kappa = eps / sigma^2, mom_b (d + 2), 0 pad ...]
//   (ld ctx_width(d)), e the means of the pooled encoder outputs;
//   rows o < 256:        hctx[b, o] = bh[o] + sum_c Wc[o, c] ctx[b, c]
//   rows o = 256 + k:    zq[b, k]   = sum_c Wq[k, c] ctx[b, c]   (k < kin)
// Every workgroup forms ctx[b] in LDS from the pooling partials (thread f
// sums feature f's chunks of both sides, kCtxTrip of each side per trip, all
// in flight), workgroup 0 also stores it for the backward pass, and each wave
// then takes one output row.  Grid ((256 + kin) / 4, B), block 256.
constexpr int kCtxTrip = 32;
constexpr int kCtxMax = 512 + (3 + 256 + 31) / 32 * 32;  // ctx_width(256)

__global__ __launch_bounds__(256) void ctx_forward_kernel(
    const float* __restrict__ part, const int* __restrict__ seg_off, int chunks, const float* __restrict__ eps,
    const float* __restrict__ eps_rel, const float* __restrict__ sig2, const float* __restrict__ mom, int d,
    const float* __restrict__ Wc, const float* __restrict__ bh, const float* __restrict__ Wq, int kin,
    float* __restrict__ ctx, float* __restrict__ hctx, float* __restrict__ zq) {
  __shared__ float cb[kCtxMax];
  const int b = blockIdx.y, f = threadIdx.x;
  const int W = ctx_width(d);
  const int seg0 = 2 * b;
  float sum[2] = {0.f, 0.f};
  for (int c = 0; c < chunks; c += kCtxTrip) {
    float v[2][kCtxTrip];
#pragma unroll
    for (int side = 0; side < 2; ++side)
#pragma unroll
      for (int u = 0; u < kCtxTrip; ++u)
        v[side][u] = c + u < chunks ? part[(static_cast<int64_t>(seg0 + side) * chunks + c + u) * 256 + f] : 0.f;
#pragma unroll
    for (int side = 0; side < 2; ++side)
#pragma unroll
      for (int u = 0; u < kCtxTrip; ++u) sum[side] += v[side][u];
  }
#pragma unroll
  for (int side = 0; side < 2; ++side)
    cb[side * 256 + f] = sum[side] / static_cast<float>(seg_off[seg0 + side + 1] - seg_off[seg0 + side]);
  // kappa = eps / sigma^2; with eps = c mean(C) and sigma^2 = mean(C) / 2 it is 2 c.
  const float kappa = (eps_rel && eps_rel[b] > 0.f) ? 2.f * eps_rel[b] : eps[b] / sig2[b];
  for (int i = f; i < W - 512; i += 256) {
    float v = 0.f;
    if (i == 0) v = kappa;
    else if (i <= d + 2) v = mom[static_cast<int64_t>(b) * (d + 2) + i - 1];
    cb[512 + i] = v;
  }
  __syncthreads();
  if (blockIdx.x == 0)
    for (int i = f; i < W; i += 256) ctx[static_cast<int64_t>(b) * W + i] = cb[i];
  const int lane = threadIdx.x & 63;
  const int o = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (o >= 256 + kin) return;
  const float* w = o < 256 ? Wc + static_cast<int64_t>(o) * W : Wq + static_cast<int64_t>(o - 256) * W;
  float s = 0.f;
  for (int c = lane; c < W; c += 64) s = fmaf(w[c], cb[c], s);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) {
    if (o < 256) hctx[b * 256 + o] = s + bh[o];
    else zq[b * kin + (o - 256)] = s;
  }
}

// Head output, one wave per row q of problem b = prob[q]:
//   out[q] = b5 + sum_k w5[k] A[q, k] + sum_{k < kin} zq[b, k] Y0[q, k]
// (A f32 [P, 256], Y0 the bf16 head input [P, kin]); g[q] = sigma_b^2 out[q]
// when g is given.  Grid ceil(rows / 4), block 256.
__global__ void head_out_kernel(const float* __restrict__ A, const float* __restrict__ w5, const float* __restrict__ b5,
                                const bf16_t* __restrict__ Y0, int kin, const float* __restrict__ zq,
                                const int* __restrict__ prob, int64_t rows, float* __restrict__ out,
                                const float* __restrict__ sig2, float* __restrict__ g) {
  const int lane = threadIdx.x & 63;
  const int64_t q = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (q >= rows) return;
  const int b = prob[q];
  const float* a = A + q * 256;
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < 4; ++i) s = fmaf(w5[lane + 64 * i], a[lane + 64 * i], s);
  for (int k = lane; k < kin; k += 64) s = fmaf(zq[b * kin + k], from_bf16(Y0[q * kin + k]), s);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) {
    const float v = s + b5[0];
    out[q] = v;
    if (g) g[q] = sig2[b] * v;
  }
}

// One layer over a tile of 16 RB rows for the fused inference kernels:
// acc = In . W^T over K, In a [16 RB x K] bf16 tile in LDS (row stride LI),
// W [256 x K] bf16 rows in global memory (read by every workgroup, so from
// L2).  Wave cw / 64 owns the output columns [cw, cw + 64): acc[i][j] holds
// rows 16 i + 4 (lane / 16) + r and column cw + 16 j + lane % 16.  The wave's
// whole slice of W (K / 32 steps x 4 fragments, at most 128 registers) is
// loaded before the first MFMA: one L2 round trip per layer instead of one per
// step (the fused kernels run one workgroup per CU, so registers are free).
template <int RB, int K, int LI>
__device__ __forceinline__ void layer_tile(const bf16_t* In, const bf16_t* __restrict__ W, int cw, int lane,
                                           f32x4 (&acc)[RB][4]) {
  constexpr int KS = K / 32;
  const int lr = lane & 15, lk = 8 * (lane >> 4);
  bf16x8 bw[KS][4];
#pragma unroll
  for (int ks = 0; ks < KS; ++ks)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      bw[ks][j] =
          __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(W + (cw + 16 * j + lr) * K + 32 * ks + lk));
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    bf16x8 ai[RB];
#pragma unroll
    for (int i = 0; i < RB; ++i)
      ai[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(&In[(16 * i + lr) * LI + 32 * ks + lk]));
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ai[i], bw[ks][j], acc[i][j], 0, 0, 0);
  }
}

// A [ROWS x K] bf16 row tile of a row-major matrix into LDS (row stride LI).
template <int ROWS, int K, int LI>
__device__ __forceinline__ void load_tile(const bf16_t* __restrict__ src, int64_t r0, bf16_t* dst, int tid) {
  for (int i = tid; i < ROWS * K / 8; i += 256) {
    const int r = i / (K / 8), c8 = i % (K / 8);
    *reinterpret_cast<u32x4*>(&dst[r * LI + c8 * 8]) = *reinterpret_cast<const u32x4*>(src + (r0 + r) * K + c8 * 8);
  }
}

// relu(acc + bias) rounded to bf16 into an LDS tile; rows >= valid are zero
// (as the unfused GEMM epilogue writes them).  The lane's four bias values are
// loaded before the first store: the stores go through a generic pointer, so a
// load left after one of them would be issued (and waited for) after it.
template <int RB, int LO>
__device__ __forceinline__ void relu_bf16_tile(const f32x4 (&acc)[RB][4], const float* __restrict__ bias, int cw,
                                               int lane, int64_t r0, int64_t valid, bf16_t* dst) {
  const int lr = lane & 15;
  float bv[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) bv[j] = bias[cw + 16 * j + lr];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * i + 4 * (lane >> 4) + r;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int col = cw + 16 * j + lr;
        const float v = r0 + row < valid ? fmaxf(acc[i][j][r] + bv[j], 0.f) : 0.f;
        dst[row * LO + col] = to_bf16(v);
      }
    }
}

// The encoder for inference, one workgroup per 32 encoder rows (the encoder
// sees a few thousand rows: small tiles spread them over more CUs):
//   A0 = relu(X0 W0^T + b0), A1 = relu(A0 W1^T + b1) (bf16, as the unfused
//   GEMMs round them), H2 = A1 W2^T + b2 (float32, for the pooling).
// Grid rows_pad / 32.
template <int KIN>
__global__ __launch_bounds__(256) void encoder_fused_kernel(const bf16_t* __restrict__ X0,
                                                            const bf16_t* __restrict__ W0, const float* __restrict__ b0,
                                                            const bf16_t* __restrict__ W1, const float* __restrict__ b1,
                                                            const bf16_t* __restrict__ W2, const float* __restrict__ b2,
                                                            int64_t rows, float* __restrict__ H2) {
  constexpr int RB = 2, LX = KIN + 8, LZ = 256 + 8;
  __shared__ bf16_t Xs[16 * RB * LX];
  __shared__ bf16_t Za[16 * RB * LZ];
  __shared__ bf16_t Zb[16 * RB * LZ];
  const int tid = threadIdx.x, lane = tid & 63, cw = (tid >> 6) * 64;
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * 16 * RB;
  load_tile<16 * RB, KIN, LX>(X0, r0, Xs, tid);
  __syncthreads();
  f32x4 acc[RB][4];
  layer_tile<RB, KIN, LX>(Xs, W0, cw, lane, acc);
  relu_bf16_tile<RB, LZ>(acc, b0, cw, lane, r0, rows, Za);
  __syncthreads();
  layer_tile<RB, 256, LZ>(Za, W1, cw, lane, acc);
  relu_bf16_tile<RB, LZ>(acc, b1, cw, lane, r0, rows, Zb);
  __syncthreads();
  layer_tile<RB, 256, LZ>(Zb, W2, cw, lane, acc);
  const int lr = lane & 15;
  float bv[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) bv[j] = b2[cw + 16 * j + lr];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * i + 4 * (lane >> 4) + r;
      const int64_t q = r0 + row;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int col = cw + 16 * j + lr;
        H2[q * 256 + col] = q < rows ? acc[i][j][r] + bv[j] : 0.f;
      }
    }
}

// The whole head for inference, one workgroup per 64 head rows:
//   Z1 = relu(Y0 Wy^T + hctx[b])  (bf16, as the unfused GEMM rounds it)
//   A2 = relu(Z1 W4^T + b4)       (float32)
//   out = A2 w5 + b5 + zq[b] . Y0,  g = sigma_b^2 out
// The row tile and Z1 stay in LDS; wave w owns output columns [64 w, 64 w + 64)
// and reads its weight rows straight from L2 (every workgroup shares them).
// KIN = the padded input width (64, 128, 192 or 256).  Grid rows_pad / 64.
template <int KIN>
__global__ __launch_bounds__(256) void head_fused_kernel(const bf16_t* __restrict__ Y0, const bf16_t* __restrict__ Wy,
                                                         const float* __restrict__ hctx,
                                                         const bf16_t* __restrict__ W4, const float* __restrict__ b4,
                                                         const float* __restrict__ w5, const float* __restrict__ b5,
                                                         const float* __restrict__ zq, const int* __restrict__ prob,
                                                         int64_t rows, const float* __restrict__ sig2,
                                                         float* __restrict__ out, float* __restrict__ g) {
  constexpr int LY = KIN + 8, LZ = 256 + 8;  // LDS row strides (bf16)
  __shared__ bf16_t Ys[64 * LY];
  __shared__ bf16_t Zs[64 * LZ];
  __shared__ float rsum[4][64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * 64;
  load_tile<64, KIN, LY>(Y0, r0, Ys, tid);  // padded rows are zero
  __syncthreads();
  const int cw = wave * 64;  // first output column of this wave
  const int lr = lane & 15;
  f32x4 acc[4][4];
  // layer 1, with the per-problem bias of each row.  The lane's 16 rows'
  // problems and their 64 bias values are all loaded before the first store
  // into LDS (a generic-pointer store orders every later load after it).
  layer_tile<4, KIN, LY>(Ys, Wy, cw, lane, acc);
  int pb[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      // clamped, not skipped: a load under a branch would be waited for alone
      const int64_t q = min(r0 + 16 * i + 4 * (lane >> 4) + r, rows - 1);
      pb[i][r] = prob[q];
    }
  float hv[4][4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int j = 0; j < 4; ++j) hv[i][r][j] = hctx[static_cast<int64_t>(pb[i][r]) * 256 + cw + 16 * j + lr];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * i + 4 * (lane >> 4) + r;
      const int64_t q = r0 + row;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int col = cw + 16 * j + lr;
        const float v = q < rows ? fmaxf(acc[i][j][r] + hv[i][r][j], 0.f) : 0.f;
        Zs[row * LZ + col] = to_bf16(v);
      }
    }
  __syncthreads();
  // layer 2
  layer_tile<4, 256, LZ>(Zs, W4, cw, lane, acc);
  // w5 . relu(. + b4) over this wave's 64 columns, per row
  float wv[4], bv[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    wv[j] = w5[cw + 16 * j + lr];
    bv[j] = b4[cw + 16 * j + lr];
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float sacc = 0.f;
#pragma unroll
      for (int j = 0; j < 4; ++j) sacc = fmaf(wv[j], fmaxf(acc[i][j][r] + bv[j], 0.f), sacc);
      // sum over the 16 lanes that share the row (lane bits 0..3)
      sacc += __shfl_xor(sacc, 1);
      sacc += __shfl_xor(sacc, 2);
      sacc += __shfl_xor(sacc, 4);
      sacc += __shfl_xor(sacc, 8);
      if (lr == 0) rsum[wave][16 * i + 4 * (lane >> 4) + r] = sacc;
    }
  __syncthreads();
  if (tid < 64) {
    const int64_t q = r0 + tid;
    if (q < rows) {
      const int b = prob[q];
      float v = ((rsum[0][tid] + rsum[1][tid]) + (rsum[2][tid] + rsum[3][tid])) + b5[0];
      const float* z = zq + static_cast<int64_t>(b) * KIN;
      for (int k = 0; k < KIN; ++k) v = fmaf(z[k], from_bf16(Ys[tid * LY + k]), v);
      out[q] = v;
      if (g) g[q] = sig2[b] * v;
    }
  }
}

// Centred MSE per problem on the head points:
// L = (1/B) sum_b (1/m_b) sum_j (r_j - mean_b r)^2,  r = out - target;
// dout_j = 2 (r_j - mean_b r) / (B m_b).  loss[b] receives L_b.
__global__ void loss_kernel(const float* __restrict__ out, const float* __restrict__ target,
                            const int* __restrict__ yoff, int nprob, float* __restrict__ dout,
                            float* __restrict__ loss) {
  __shared__ double red[256];
  const int b = blockIdx.x, t = threadIdx.x;
  const int q0 = yoff[b], q1 = yoff[b + 1];
  double s = 0;
  for (int q = q0 + t; q < q1; q += 256) s += static_cast<double>(out[q]) - target[q];
  red[t] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (t < k) red[t] += red[t + k];
    __syncthreads();
  }
  const double mean = red[0] / (q1 - q0);
  __syncthreads();
  double l = 0;
  for (int q = q0 + t; q < q1; q += 256) {
    const double r = static_cast<double>(out[q]) - target[q] - mean;
    l += r * r;
    if (dout) dout[q] = static_cast<float>(2.0 * r / (static_cast<double>(nprob) * (q1 - q0)));
  }
  red[t] = l;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (t < k) red[t] += red[t + k];
    __syncthreads();
  }
  if (t == 0) loss[b] = static_cast<float>(red[0] / (q1 - q0));
}

// Backward of the last layer: dZ[q, k] = dout[q] w5[k] (A[q, k] > 0), bf16 and
// transposed bf16; rows >= rows are zero.
__global__ void head_out_backward_kernel(const float* __restrict__ dout, const float* __restrict__ w5,
                                         const float* __restrict__ A, int64_t rows, int64_t rows_pad,
                                         bf16_t* __restrict__ dZ, bf16_t* __restrict__ dZT) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= rows_pad * 256) return;
  const int64_t q = i / 256;
  const int k = static_cast<int>(i % 256);
  float v = 0.f;
  if (q < rows && A[q * 256 + k] > 0.f) v = dout[q] * w5[k];
  const bf16_t h = to_bf16(v);
  dZ[i] = h;
  dZT[static_cast<int64_t>(k) * rows_pad + q] = h;
}

// Column sums over row chunks (bias gradients and per-group sums):
// part[c][N] = sum_{rows of chunk c} M[r, :], with M bf16 [rows, N] or f32.
__global__ void colsum_partial_kernel(const bf16_t* __restrict__ Mb, const float* __restrict__ Mf, int64_t rows,
                                      int N, int chunks, float* __restrict__ part) {
  const int c = blockIdx.x;
  const int64_t per = (rows + chunks - 1) / chunks;
  const int64_t r0 = c * per, r1 = min(rows, r0 + per);
  for (int f = threadIdx.x; f < N; f += blockDim.x) {
    float s = 0.f;
    for (int64_t r = r0; r < r1; ++r) s += Mb ? from_bf16(Mb[r * N + f]) : Mf[r * N + f];
    part[static_cast<int64_t>(c) * N + f] = s;
  }
}

// Gradient of the head's second bias in float32 (the centred loss makes it a
// sum with heavy cancellation, so it is not taken from the bf16 copy of dZ):
// part[c][k] = w5[k] sum_{q in chunk c} dout[q] (A[q, k] > 0).
__global__ void head_bias_grad_partial_kernel(const float* __restrict__ dout, const float* __restrict__ w5,
                                              const float* __restrict__ A, int64_t rows, int chunks,
                                              float* __restrict__ part) {
  const int c = blockIdx.x, k = threadIdx.x;
  const int64_t per = (rows + chunks - 1) / chunks;
  const int64_t r0 = c * per, r1 = min(rows, r0 + per);
  float s = 0.f;
  for (int64_t q = r0; q < r1; ++q)
    if (A[q * 256 + k] > 0.f) s += dout[q];
  part[static_cast<int64_t>(c) * 256 + k] = w5[k] * s;
}

// Per-group row sums: out[b, :] = sum_{r in [off[b], off[b+1])} M[r, :] (bf16 M).
__global__ void group_sum_kernel(const bf16_t* __restrict__ M, const int* __restrict__ off, int N,
                                 float* __restrict__ out) {
  const int b = blockIdx.x;
  for (int f = threadIdx.x; f < N; f += blockDim.x) {
    float s = 0.f;
    for (int r = off[b]; r < off[b + 1]; ++r) s += from_bf16(M[static_cast<int64_t>(r) * N + f]);
    out[static_cast<int64_t>(b) * N + f] = s;
  }
}

// dw5[k] = sum_q dout[q] A[q, k] and db5 = sum_q dout[q], over row chunks.
__global__ void head_out_wgrad_kernel(const float* __restrict__ dout, const float* __restrict__ A, int64_t rows,
                                      int chunks, float* __restrict__ part) {
  const int c = blockIdx.x, k = threadIdx.x;
  const int64_t per = (rows + chunks - 1) / chunks;
  const int64_t r0 = c * per, r1 = min(rows, r0 + per);
  float s = 0.f, sb = 0.f;
  for (int64_t q = r0; q < r1; ++q) {
    s = fmaf(dout[q], A[q * 256 + k], s);
    if (k == 0) sb += dout[q];
  }
  part[static_cast<int64_t>(c) * 257 + k] = s;
  if (k == 0) part[static_cast<int64_t>(c) * 257 + 256] = sb;
}

// Gradient of the head's quadratic term with respect to z:
// part[(b * chunks + c) * kin + k] = sum_{q in chunk c of problem b} dout[q] Y0[q, k].
// Grid (chunks, B), block 256 (k = thread < kin).
__global__ void quad_grad_partial_kernel(const float* __restrict__ dout, const bf16_t* __restrict__ Y0, int kin,
                                         const int* __restrict__ yoff, int chunks, float* __restrict__ part) {
  const int c = blockIdx.x, b = blockIdx.y, k = threadIdx.x;
  if (k >= kin) return;
  const int q0 = yoff[b], q1 = yoff[b + 1];
  const int per = (q1 - q0 + chunks - 1) / chunks;
  const int lo = q0 + c * per, hi = min(q1, lo + per);
  float s = 0.f;
  for (int q = lo; q < hi; ++q) s = fmaf(dout[q], from_bf16(Y0[static_cast<int64_t>(q) * kin + k]), s);
  part[(static_cast<int64_t>(b) * chunks + c) * kin + k] = s;
}

// dzq[b, k] = sum_c part[(b * chunks + c) * kin + k].  Grid B, block 256.
__global__ void quad_grad_final_kernel(const float* __restrict__ part, int kin, int chunks, float* __restrict__ dzq) {
  const int b = blockIdx.x, k = threadIdx.x;
  if (k >= kin) return;
  float s = 0.f;
  for (int c = 0; c < chunks; ++c) s += part[(static_cast<int64_t>(b) * chunks + c) * kin + k];
  dzq[b * kin + k] = s;
}

// Context backward, thread per context column c (W = ctx_width(d)):
//   dWc[o, c] = sum_b dh[b, o] ctx[b, c],  dWq[k, c] = sum_b dzq[b, k] ctx[b, c],
//   dctx[b, c] = sum_o Wc[o, c] dh[b, o] + sum_k Wq[k, c] dzq[b, k].
__global__ void ctx_backward_kernel(const float* __restrict__ Wc, const float* __restrict__ Wq,
                                    const float* __restrict__ ctx, const float* __restrict__ dh,
                                    const float* __restrict__ dzq, int nprob, int d, int kin,
                                    float* __restrict__ dWc, float* __restrict__ dWq, float* __restrict__ dctx) {
  const int W = ctx_width(d);
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= W) return;
  for (int o = 0; o < 256; ++o) {
    float s = 0.f;
    for (int b = 0; b < nprob; ++b) s = fmaf(dh[b * 256 + o], ctx[static_cast<int64_t>(b) * W + c], s);
    dWc[static_cast<int64_t>(o) * W + c] = s;
  }
  for (int k = 0; k < kin; ++k) {
    float s = 0.f;
    for (int b = 0; b < nprob; ++b) s = fmaf(dzq[b * kin + k], ctx[static_cast<int64_t>(b) * W + c], s);
    dWq[static_cast<int64_t>(k) * W + c] = s;
  }
  for (int b = 0; b < nprob; ++b) {
    float s = 0.f;
    for (int o = 0; o < 256; ++o) s = fmaf(Wc[static_cast<int64_t>(o) * W + c], dh[b * 256 + o], s);
    for (int k = 0; k < kin; ++k) s = fmaf(Wq[static_cast<int64_t>(k) * W + c], dzq[b * kin + k], s);
    dctx[static_cast<int64_t>(b) * W + c] = s;
  }
}

// Pool backward: dH[r, f] = dctx[b(r), side(r) * 256 + f] / count(segment),
// bf16 and transposed bf16, zero on padded rows.
__global__ void pool_backward_kernel(const float* __restrict__ dctx, int ctx_w, const int* __restrict__ seg_of_row,
                                     const int* __restrict__ seg_off, int64_t rows, int64_t rows_pad,
                                     bf16_t* __restrict__ dH, bf16_t* __restrict__ dHT) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= rows_pad * 256) return;
  const int64_t r = i / 256;
  const int f = static_cast<int>(i % 256);
  float v = 0.f;
  if (r < rows) {
    const int seg = seg_of_row[r];
    v = dctx[static_cast<int64_t>(seg >> 1) * ctx_w + (seg & 1) * 256 + f] /
        static_cast<float>(seg_off[seg + 1] - seg_off[seg]);
  }
  const bf16_t h = to_bf16(v);
  dH[i] = h;
  dHT[static_cast<int64_t>(f) * rows_pad + r] = h;
}

// Adam on a flat f32 parameter buffer (optax.adam semantics with bias
// correction), then the bf16 copies: W (row-major) and WT (transposed) for the
// tensors listed in the copy table.
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int64_t count, float lr, float b1, float b2, float eps,
                            float bc1, float bc2) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

__global__ void to_bf16_kernel(const float* __restrict__ src, int rows, int cols, int ld_src, bf16_t* __restrict__ dst,
                               int ld_dst, bf16_t* __restrict__ dstT, int ld_dstT) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<int64_t>(rows) * cols) return;
  const int r = static_cast<int>(i / cols), c = static_cast<int>(i % cols);
  const bf16_t h = to_bf16(src[static_cast<int64_t>(r) * ld_src + c]);
  if (dst) dst[static_cast<int64_t>(r) * ld_dst + c] = h;
  if (dstT) dstT[static_cast<int64_t>(c) * ld_dstT + r] = h;
}

}  // namespace ottx_ds

using namespace ottx_ds;

namespace {

#define DS_LAUNCHED()                                     \
  do {                                                    \
    if (hipGetLastError() != hipSuccess) return OTTX_EHIP; \
  } while (0)

inline int blocks_for(int64_t n, int t = 256) { return static_cast<int>((n + t - 1) / t); }
inline int64_t pad64(int64_t v) { return (v + 63) / 64 * 64; }
inline size_t al(size_t v) { return (v + 255) & ~static_cast<size_t>(255); }

// Parameter layout inside the flat f32 buffer (offsets in floats).
struct ParamLayout {
  int kin;  // padded input width (multiple of 64: a GEMM N dimension in the gradients)
  int ctx;  // context width
  int64_t W0, b0, W1, b1, W2, b2, Wy, Wc, bh, W4, b4, w5, b5, Wq, total;
};

ParamLayout param_layout(int d) {
  ParamLayout L;
  L.kin = (d + 1 + 63) / 64 * 64;  // d coordinates + the squared-norm feature
  L.ctx = ctx_width(d);
  int64_t o = 0;
  L.W0 = o; o += 256LL * L.kin;
  L.b0 = o; o += 256;
  L.W1 = o; o += 256LL * 256;
  L.b1 = o; o += 256;
  L.W2 = o; o += 256LL * 256;
  L.b2 = o; o += 256;
  L.Wy = o; o += 256LL * L.kin;
  L.Wc = o; o += 256LL * L.ctx;
  L.bh = o; o += 256;
  L.W4 = o; o += 256LL * 256;
  L.b4 = o; o += 256;
  L.w5 = o; o += 256;
  L.b5 = o; o += 64;
  L.Wq = o; o += static_cast<int64_t>(L.kin) * L.ctx;
  L.total = o;
  return L;
}

// bf16 weight copies (row-major W and transposed WT) for the GEMM operands.
struct WeightCopies {
  int64_t W0, W0T, W1, W1T, W2, W2T, Wy, WyT, W4, W4T, total;  // offsets in bf16 elements
};

WeightCopies weight_copies(const ParamLayout& L) {
  WeightCopies C;
  int64_t o = 0;
  C.W0 = o; o += 256LL * L.kin;
  C.W0T = o; o += 256LL * L.kin;
  C.W1 = o; o += 65536;
  C.W1T = o; o += 65536;
  C.W2 = o; o += 65536;
  C.W2T = o; o += 65536;
  C.Wy = o; o += 256LL * L.kin;
  C.WyT = o; o += 256LL * L.kin;
  C.W4 = o; o += 65536;
  C.W4T = o; o += 65536;
  C.total = o;
  return C;
}

// Workspace layout for one batch.
struct DsLayout {
  int64_t Pe, Pe_pad, Ph, Ph_pad;  // encoder / head rows (valid, padded)
  int nprob, chunks, kchunks;
  size_t mu, sig2, mom, spart, kappa, X0, Y0, A0, A0T, A1, A1T, H2f, part, ctx, hctx, zq, Z1h, Z1hT, A2h, A2hT, A2f,
      out, dout, loss, dZ, dZT, dA, dAT, dctx, dh, qpart, dzq, gpart, total;
};

DsLayout ds_layout(const ottx_ds_batch* bt, const ParamLayout& P, bool training) {
  DsLayout L{};
  L.nprob = bt->count;
  L.Pe = bt->enc_rows;
  L.Pe_pad = pad64(L.Pe);
  L.Ph = bt->m_total;
  L.Ph_pad = pad64(L.Ph);
  L.chunks = 64;
  L.kchunks = training ? 64 : 0;
  size_t o = 0;
  auto take = [&](size_t b) {
    const size_t at = o;
    o = al(o + b);
    return at;
  };
  auto take_t = [&](size_t b) { return take(training ? b : 0); };  // training-only buffers
  const int64_t Pe = L.Pe_pad, Ph = L.Ph_pad;
  const int B = L.nprob;
  L.mu = take(static_cast<size_t>(B) * bt->d * 4);
  L.sig2 = take(static_cast<size_t>(B) * 4);
  L.mom = take(static_cast<size_t>(B) * (bt->d + 2) * 4);
  L.spart = take(static_cast<size_t>(B) * kStatChunks * 4 * bt->d * 8);
  L.kappa = take(static_cast<size_t>(B) * 4);
  L.X0 = take(static_cast<size_t>(Pe) * P.kin * 2 * (training ? 2 : 1));  // X0 (and X0T)
  L.Y0 = take(static_cast<size_t>(Ph) * P.kin * 2 * (training ? 2 : 1));  // Y0 (and Y0T)
  L.A0 = take(static_cast<size_t>(Pe) * 256 * 2);
  L.A0T = take_t(static_cast<size_t>(Pe) * 256 * 2);
  L.A1 = take(static_cast<size_t>(Pe) * 256 * 2);
  L.A1T = take_t(static_cast<size_t>(Pe) * 256 * 2);
  L.H2f = take(static_cast<size_t>(Pe) * 256 * 4);
  L.part = take(static_cast<size_t>(2 * B) * L.chunks * 256 * 4);
  L.ctx = take(static_cast<size_t>(B) * P.ctx * 4);
  L.hctx = take(static_cast<size_t>(B) * 256 * 4);
  L.zq = take(static_cast<size_t>(B) * P.kin * 4);
  L.Z1h = take(static_cast<size_t>(Ph) * 256 * 2);
  L.Z1hT = take_t(static_cast<size_t>(Ph) * 256 * 2);
  L.A2h = take(static_cast<size_t>(Ph) * 256 * 2);
  L.A2hT = take_t(static_cast<size_t>(Ph) * 256 * 2);
  L.A2f = take(static_cast<size_t>(Ph) * 256 * 4);
  L.out = take(static_cast<size_t>(Ph) * 4);
  L.dout = take_t(static_cast<size_t>(Ph) * 4);
  L.loss = take(static_cast<size_t>(B) * 4);
  const int64_t Pm = Pe > Ph ? Pe : Ph;
  L.dZ = take_t(static_cast<size_t>(Pm) * 256 * 2);
  L.dZT = take_t(static_cast<size_t>(Pm) * 256 * 2);
  L.dA = take_t(static_cast<size_t>(Pm) * 256 * 2);
  L.dAT = take_t(static_cast<size_t>(Pm) * 256 * 2);
  L.dctx = take_t(static_cast<size_t>(B) * P.ctx * 4);
  L.dh = take_t(static_cast<size_t>(B) * 256 * 4);
  L.qpart = take_t(static_cast<size_t>(B) * L.chunks * P.kin * 4);
  L.dzq = take_t(static_cast<size_t>(B) * P.kin * 4);
  // split-K partials of the weight gradients (at most 256 x 256 per chunk)
  L.gpart = take_t(static_cast<size_t>(L.kchunks) * 256 * 256 * 4);
  L.total = o;
  return L;
}

int gemm(hipStream_t st, const GemmArgs& a) {
  if (a.M % 64 || a.N % 64 || a.K % 32 || a.splits < 1) return OTTX_EINVAL;
  // one 256-wide chunk for short K and few row tiles; otherwise 64-wide
  // double-buffered chunks (more workgroups resident, loads overlapped)
  if (a.K <= 256 && a.M < 8192)
    hipLaunchKernelGGL((gemm_nt_kernel<256, 1>), dim3(a.N / 64, a.M / 64, a.splits), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL((gemm_nt_kernel<64, 2>), dim3(a.N / 64, a.M / 64, a.splits), dim3(256), 0, st, a);
  DS_LAUNCHED();
  return OTTX_OK;
}

GemmArgs gemm_args(const bf16_t* A, int64_t lda, const bf16_t* Bt, int64_t ldb, int M, int N, int K) {
  GemmArgs g{};
  g.A = A;
  g.Bt = Bt;
  g.lda = lda;
  g.ldb = ldb;
  g.M = M;
  g.N = N;
  g.K = K;
  g.splits = 1;
  g.valid_rows = M;
  return g;
}

// dW[N_out, K_in] = dZ^T . X  as split-K GEMM over the rows, summed into grad.
int weight_grad(hipStream_t st, const bf16_t* dZT, int64_t rows_pad, const bf16_t* XT, int n_out, int k_in,
                float* gpart, int kchunks, float* grad) {
  int splits = kchunks;
  while (splits > 1 && (rows_pad % splits != 0 || (rows_pad / splits) % 32 != 0)) --splits;
  GemmArgs g = gemm_args(dZT, rows_pad, XT, rows_pad, n_out, k_in, static_cast<int>(rows_pad / splits));
  g.splits = splits;
  g.Cf = gpart;
  g.ldc = k_in;
  int rc = gemm(st, g);
  if (rc != OTTX_OK) return rc;
  const int64_t count = static_cast<int64_t>(n_out) * k_in;
  hipLaunchKernelGGL(sum_splits_kernel, dim3(blocks_for(count)), dim3(256), 0, st, gpart, splits, count, grad, 0);
  DS_LAUNCHED();
  return OTTX_OK;
}

// grad[0:N] = column sums of M (bf16 [rows, N]).
int bias_grad(hipStream_t st, const bf16_t* M, int64_t rows, int N, float* part, int chunks, float* grad) {
  hipLaunchKernelGGL(colsum_partial_kernel, dim3(chunks), dim3(256), 0, st, M, static_cast<const float*>(nullptr), rows,
                     N, chunks, part);
  hipLaunchKernelGGL(sum_splits_kernel, dim3(blocks_for(N)), dim3(256), 0, st, part, chunks, static_cast<int64_t>(N),
                     grad, 0);
  DS_LAUNCHED();
  return OTTX_OK;
}

struct Forward {
  DsLayout L;
  ParamLayout P;
  WeightCopies C;
  char* ws;
  const float* params;
  const bf16_t* wb;
};

template <typename T>
T* at(char* base, size_t off) {
  return reinterpret_cast<T*>(base + off);
}

// The forward pass up to the head output (out[Ph]).
// g_out (optional) receives the prediction as a potential, sigma^2 * out.
int forward(const ottx_ds_batch* bt, const ottx_ds_model* md, void* workspace, hipStream_t st, bool training,
            float* g_out, Forward* F) {
  F->P = param_layout(bt->d);
  F->C = weight_copies(F->P);
  F->L = ds_layout(bt, F->P, training);
  const ParamLayout& P = F->P;
  const WeightCopies& C = F->C;
  const DsLayout& L = F->L;
  char* ws = static_cast<char*>(workspace);
  F->ws = ws;
  F->params = md->params;
  F->wb = reinterpret_cast<const bf16_t*>(md->weights_bf16);
  const float* prm = md->params;
  const bf16_t* wb = F->wb;
  const int B = L.nprob;
  const int64_t Pe = L.Pe_pad, Ph = L.Ph_pad;
  float* mu = at<float>(ws, L.mu);
  float* sig2 = at<float>(ws, L.sig2);
  double* spart = at<double>(ws, L.spart);
  hipLaunchKernelGGL(stats_partial_kernel, dim3(kStatChunks, B), dim3(256), 0, st, bt->x, bt->y, bt->x_off,
                     bt->y_off, bt->d, spart);
  float* mom = at<float>(ws, L.mom);
  hipLaunchKernelGGL(stats_final_kernel, dim3(B), dim3(kFinalThreads), 0, st, spart, bt->x_off, bt->y_off, bt->d,
                     mu, sig2, mom);
  DS_LAUNCHED();
  bf16_t* X0 = at<bf16_t>(ws, L.X0);
  bf16_t* X0T = X0 + Pe * P.kin;
  bf16_t* Y0 = at<bf16_t>(ws, L.Y0);
  bf16_t* Y0T = Y0 + Ph * P.kin;
  // encoder and head inputs in one launch
  const NormJob enc{bt->enc_src, bt->seg_of_row, nullptr, L.Pe, Pe, X0, training ? X0T : nullptr};
  const NormJob head{nullptr, nullptr, bt->y_prob, L.Ph, Ph, Y0, training ? Y0T : nullptr};
  const int enc_blocks = static_cast<int>((Pe + 3) / 4);
  hipLaunchKernelGGL(normalize_kernel, dim3(static_cast<unsigned>(enc_blocks + (Ph + 3) / 4)), dim3(256), 0, st,
                     bt->x, bt->y, enc, head, enc_blocks, bt->d, P.kin, mu, sig2);
  DS_LAUNCHED();
  int rc;
  GemmArgs g;
  if (!training) {
    // inference: the encoder in one launch
    auto launch = [&](auto kernel) {
      hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(Pe / 32)), dim3(256), 0, st, X0, wb + C.W0, prm + P.b0,
                         wb + C.W1, prm + P.b1, wb + C.W2, prm + P.b2, L.Pe, at<float>(ws, L.H2f));
    };
    switch (P.kin) {
      case 64: launch(encoder_fused_kernel<64>); break;
      case 128: launch(encoder_fused_kernel<128>); break;
      case 192: launch(encoder_fused_kernel<192>); break;
      default: launch(encoder_fused_kernel<256>); break;
    }
    DS_LAUNCHED();
  } else {
    // encoder layer 0: A0 = relu(X0 W0^T + b0)
    g = gemm_args(X0, P.kin, wb + C.W0, P.kin, static_cast<int>(Pe), 256, P.kin);
    g.bias = prm + P.b0;
    g.relu = 1;
    g.Cb = at<bf16_t>(ws, L.A0);
    g.ldcb = 256;
    g.CbT = training ? at<bf16_t>(ws, L.A0T) : nullptr;
    g.ldct = Pe;
    g.valid_rows = static_cast<int>(L.Pe);
    if ((rc = gemm(st, g)) != OTTX_OK) return rc;
    // encoder layer 1: A1 = relu(A0 W1^T + b1)
    g = gemm_args(at<bf16_t>(ws, L.A0), 256, wb + C.W1, 256, static_cast<int>(Pe), 256, 256);
    g.bias = prm + P.b1;
    g.relu = 1;
    g.Cb = at<bf16_t>(ws, L.A1);
    g.ldcb = 256;
    g.CbT = training ? at<bf16_t>(ws, L.A1T) : nullptr;
    g.ldct = Pe;
    g.valid_rows = static_cast<int>(L.Pe);
    if ((rc = gemm(st, g)) != OTTX_OK) return rc;
    // encoder layer 2 (linear): H2 = A1 W2^T + b2 (f32 for the pooling)
    g = gemm_args(at<bf16_t>(ws, L.A1), 256, wb + C.W2, 256, static_cast<int>(Pe), 256, 256);
    g.bias = prm + P.b2;
    g.Cf = at<float>(ws, L.H2f);
    g.ldc = 256;
    g.valid_rows = static_cast<int>(L.Pe);
    if ((rc = gemm(st, g)) != OTTX_OK) return rc;
  }
  // pooling per (problem, side) and the context vector
  hipLaunchKernelGGL(pool_partial_kernel, dim3(L.chunks, 2 * B), dim3(256), 0, st, at<float>(ws, L.H2f), bt->seg_off,
                     L.chunks, at<float>(ws, L.part));
  hipLaunchKernelGGL(ctx_forward_kernel, dim3((256 + P.kin + 3) / 4, B), dim3(256), 0, st, at<float>(ws, L.part),
                     bt->seg_off, L.chunks, bt->eps, bt->eps_rel, sig2, mom, bt->d, prm + P.Wc, prm + P.bh,
                     prm + P.Wq, P.kin, at<float>(ws, L.ctx), at<float>(ws, L.hctx), at<float>(ws, L.zq));
  DS_LAUNCHED();
  if (!training) {
    // inference: the head in one launch
    const unsigned grid = static_cast<unsigned>(Ph / 64);
    auto launch = [&](auto kernel) {
      hipLaunchKernelGGL(kernel, dim3(grid), dim3(256), 0, st, Y0, wb + C.Wy, at<float>(ws, L.hctx), wb + C.W4,
                         prm + P.b4, prm + P.w5, prm + P.b5, at<float>(ws, L.zq), bt->y_prob, L.Ph, sig2,
                         at<float>(ws, L.out), g_out);
    };
    switch (P.kin) {
      case 64: launch(head_fused_kernel<64>); break;
      case 128: launch(head_fused_kernel<128>); break;
      case 192: launch(head_fused_kernel<192>); break;
      default: launch(head_fused_kernel<256>); break;
    }
    DS_LAUNCHED();
    return OTTX_OK;
  }
  // head layer 1: Z1 = relu(Y0 Wy^T + hctx[b])
  g = gemm_args(Y0, P.kin, wb + C.Wy, P.kin, static_cast<int>(Ph), 256, P.kin);
  g.gbias = at<float>(ws, L.hctx);
  g.row_group = bt->y_prob;
  g.relu = 1;
  g.Cb = at<bf16_t>(ws, L.Z1h);
  g.ldcb = 256;
  g.CbT = training ? at<bf16_t>(ws, L.Z1hT) : nullptr;
  g.ldct = Ph;
  g.valid_rows = static_cast<int>(L.Ph);
  if ((rc = gemm(st, g)) != OTTX_OK) return rc;
  // head layer 2: A2 = relu(Z1 W4^T + b4)
  g = gemm_args(at<bf16_t>(ws, L.Z1h), 256, wb + C.W4, 256, static_cast<int>(Ph), 256, 256);
  g.bias = prm + P.b4;
  g.relu = 1;
  g.Cf = at<float>(ws, L.A2f);
  g.ldc = 256;
  g.Cb = at<bf16_t>(ws, L.A2h);
  g.ldcb = 256;
  g.CbT = training ? at<bf16_t>(ws, L.A2hT) : nullptr;
  g.ldct = Ph;
  g.valid_rows = static_cast<int>(L.Ph);
  if ((rc = gemm(st, g)) != OTTX_OK) return rc;
  hipLaunchKernelGGL(head_out_kernel, dim3(static_cast<unsigned>((L.Ph + 3) / 4)), dim3(256), 0, st,
                     at<float>(ws, L.A2f), prm + P.w5, prm + P.b5, Y0, P.kin, at<float>(ws, L.zq), bt->y_prob, L.Ph,
                     at<float>(ws, L.out), sig2, g_out);
  DS_LAUNCHED();
  return OTTX_OK;
}

}  // namespace

extern "C" {

int64_t ottx_ds_param_count(int32_t d) {
  if (d < 1 || d > 255) return 0;
  return param_layout(d).total;
}

int64_t ottx_ds_weight_bf16_count(int32_t d) {
  if (d < 1 || d > 255) return 0;
  return weight_copies(param_layout(d)).total;
}

size_t ottx_ds_workspace_bytes(const ottx_ds_batch* bt, int32_t training) {
  if (!bt || bt->count < 1 || bt->d < 1 || bt->d > 255 || bt->n_total < 1 || bt->m_total < 1) return 0;
  // each of the 2 B segments holds at least one encoder row, and no more rows than points
  if (bt->enc_rows < 2 * bt->count ||
      static_cast<int64_t>(bt->enc_rows) > static_cast<int64_t>(bt->n_total) + bt->m_total)
    return 0;
  return ds_layout(bt, param_layout(bt->d), training != 0).total;
}

int ottx_ds_refresh_bf16(const ottx_ds_model* md, void* stream) {
  if (!md || !md->params || !md->weights_bf16 || md->d < 1 || md->d > 255) return OTTX_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const ParamLayout P = param_layout(md->d);
  const WeightCopies C = weight_copies(P);
  bf16_t* wb = reinterpret_cast<bf16_t*>(md->weights_bf16);
  const float* p = md->params;
  struct Item {
    int64_t src, dst, dstT;
    int rows, cols;
  } items[] = {{P.W0, C.W0, C.W0T, 256, P.kin}, {P.W1, C.W1, C.W1T, 256, 256}, {P.W2, C.W2, C.W2T, 256, 256},
               {P.Wy, C.Wy, C.WyT, 256, P.kin}, {P.W4, C.W4, C.W4T, 256, 256}};
  for (const Item& it : items) {
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(static_cast<int64_t>(it.rows) * it.cols)), dim3(256), 0, st,
                       p + it.src, it.rows, it.cols, it.cols, wb + it.dst, it.cols, wb + it.dstT, it.rows);
  }
  DS_LAUNCHED();
  return OTTX_OK;
}

int ottx_ds_forward(const ottx_ds_model* md, const ottx_ds_batch* bt, float* g_out, void* workspace,
                    size_t workspace_bytes, void* stream) {
  if (!md || !bt || !g_out || !workspace || md->d != bt->d) return OTTX_EINVAL;
  const size_t need = ottx_ds_workspace_bytes(bt, 0);
  if (need == 0) return OTTX_EINVAL;
  if (workspace_bytes < need || (reinterpret_cast<uintptr_t>(workspace) & 255)) return OTTX_EWORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  Forward F;
  return forward(bt, md, workspace, st, false, g_out, &F);
}

int ottx_ds_loss_backward(const ottx_ds_model* md, const ottx_ds_batch* bt, const float* target, float* grad,
                          float* loss_out, void* workspace, size_t workspace_bytes, void* stream) {
  if (!md || !bt || !target || !grad || !workspace || md->d != bt->d) return OTTX_EINVAL;
  const size_t need = ottx_ds_workspace_bytes(bt, 1);
  if (need == 0) return OTTX_EINVAL;
  if (workspace_bytes < need || (reinterpret_cast<uintptr_t>(workspace) & 255)) return OTTX_EWORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  Forward F;
  int rc = forward(bt, md, workspace, st, true, nullptr, &F);
  if (rc != OTTX_OK) return rc;
  const ParamLayout& P = F.P;
  const WeightCopies& C = F.C;
  const DsLayout& L = F.L;
  char* ws = F.ws;
  const float* prm = md->params;
  const bf16_t* wb = F.wb;
  const int B = L.nprob;
  const int64_t Pe = L.Pe_pad, Ph = L.Ph_pad;
  float* dout = at<float>(ws, L.dout);
  float* lossb = loss_out ? loss_out : at<float>(ws, L.loss);
  hipLaunchKernelGGL(loss_kernel, dim3(B), dim3(256), 0, st, at<float>(ws, L.out), target, bt->y_off, B, dout, lossb);
  DS_LAUNCHED();
  float* gpart = at<float>(ws, L.gpart);
  // last layer
  hipLaunchKernelGGL(head_out_wgrad_kernel, dim3(L.kchunks), dim3(256), 0, st, dout, at<float>(ws, L.A2f), L.Ph,
                     L.kchunks, gpart);
  hipLaunchKernelGGL(sum_splits_kernel, dim3(blocks_for(257)), dim3(256), 0, st, gpart, L.kchunks,
                     static_cast<int64_t>(257), grad + P.w5, 0);
  DS_LAUNCHED();
  bf16_t* dZ = at<bf16_t>(ws, L.dZ);
  bf16_t* dZT = at<bf16_t>(ws, L.dZT);
  bf16_t* dA = at<bf16_t>(ws, L.dA);
  bf16_t* dAT = at<bf16_t>(ws, L.dAT);
  hipLaunchKernelGGL(head_out_backward_kernel, dim3(blocks_for(Ph * 256)), dim3(256), 0, st, dout, prm + P.w5,
                     at<float>(ws, L.A2f), L.Ph, Ph, dZ, dZT);
  DS_LAUNCHED();
  // head layer 2: dW4 = dZ^T Z1, db4 = colsum dZ, dZ1 = (dZ W4) * (Z1 > 0)
  if ((rc = weight_grad(st, dZT, Ph, at<bf16_t>(ws, L.Z1hT), 256, 256, gpart, L.kchunks, grad + P.W4)) != OTTX_OK)
    return rc;
  hipLaunchKernelGGL(head_bias_grad_partial_kernel, dim3(L.kchunks), dim3(256), 0, st, dout, prm + P.w5,
                     at<float>(ws, L.A2f), L.Ph, L.kchunks, gpart);
  hipLaunchKernelGGL(sum_splits_kernel, dim3(1), dim3(256), 0, st, gpart, L.kchunks, static_cast<int64_t>(256),
                     grad + P.b4, 0);
  DS_LAUNCHED();
  GemmArgs g = gemm_args(dZ, 256, wb + C.W4T, 256, static_cast<int>(Ph), 256, 256);
  g.mask = at<bf16_t>(ws, L.Z1h);
  g.ldm = 256;
  g.Cb = dA;
  g.ldcb = 256;
  g.CbT = dAT;
  g.ldct = Ph;
  g.valid_rows = static_cast<int>(L.Ph);
  if ((rc = gemm(st, g)) != OTTX_OK) return rc;
  // head layer 1: dWy = dZ1^T Y0, dh[b] = per-problem sums of dZ1, bh grad = sum_b dh
  if ((rc = weight_grad(st, dAT, Ph, at<bf16_t>(ws, L.Y0) + Ph * P.kin, 256, P.kin, gpart, L.kchunks,
                        grad + P.Wy)) != OTTX_OK)
    return rc;
  float* dh = at<float>(ws, L.dh);
  hipLaunchKernelGGL(group_sum_kernel, dim3(B), dim3(256), 0, st, dA, bt->y_off, 256, dh);
  hipLaunchKernelGGL(sum_splits_kernel, dim3(1), dim3(256), 0, st, dh, B, static_cast<int64_t>(256), grad + P.bh, 0);
  // quadratic term: dz[b] = sum_{rows of b} dout Y0
  float* dzq = at<float>(ws, L.dzq);
  hipLaunchKernelGGL(quad_grad_partial_kernel, dim3(L.chunks, B), dim3(256), 0, st, dout, at<bf16_t>(ws, L.Y0),
                     P.kin, bt->y_off, L.chunks, at<float>(ws, L.qpart));
  hipLaunchKernelGGL(quad_grad_final_kernel, dim3(B), dim3(256), 0, st, at<float>(ws, L.qpart), P.kin, L.chunks,
                     dzq);
  float* dctx = at<float>(ws, L.dctx);
  hipLaunchKernelGGL(ctx_backward_kernel, dim3(blocks_for(P.ctx)), dim3(256), 0, st, prm + P.Wc, prm + P.Wq,
                     at<float>(ws, L.ctx), dh, dzq, B, bt->d, P.kin, grad + P.Wc, grad + P.Wq, dctx);
  DS_LAUNCHED();
  // pool backward -> dH2 (encoder output gradient)
  hipLaunchKernelGGL(pool_backward_kernel, dim3(blocks_for(Pe * 256)), dim3(256), 0, st, dctx, P.ctx,
                     bt->seg_of_row, bt->seg_off, L.Pe, Pe, dZ, dZT);
  DS_LAUNCHED();
  // encoder layer 2 (linear): dW2 = dH2^T A1, db2, dA1 = (dH2 W2) * (A1 > 0)
  if ((rc = weight_grad(st, dZT, Pe, at<bf16_t>(ws, L.A1T), 256, 256, gpart, L.kchunks, grad + P.W2)) != OTTX_OK)
    return rc;
  if ((rc = bias_grad(st, dZ, Pe, 256, gpart, L.kchunks, grad + P.b2)) != OTTX_OK) return rc;
  g = gemm_args(dZ, 256, wb + C.W2T, 256, static_cast<int>(Pe), 256, 256);
  g.mask = at<bf16_t>(ws, L.A1);
  g.ldm = 256;
  g.Cb = dA;
  g.ldcb = 256;
  g.CbT = dAT;
  g.ldct = Pe;
  g.valid_rows = static_cast<int>(L.Pe);
  if ((rc = gemm(st, g)) != OTTX_OK) return rc;
  // encoder layer 1: dW1 = dZ1^T A0, db1, dA0 = (dZ1 W1) * (A0 > 0)
  if ((rc = weight_grad(st, dAT, Pe, at<bf16_t>(ws, L.A0T), 256, 256, gpart, L.kchunks, grad + P.W1)) != OTTX_OK)
    return rc;
  if ((rc = bias_grad(st, dA, Pe, 256, gpart, L.kchunks, grad + P.b1)) != OTTX_OK) return rc;
  g = gemm_args(dA, 256, wb + C.W1T, 256, static_cast<int>(Pe), 256, 256);
  g.mask = at<bf16_t>(ws, L.A0);
  g.ldm = 256;
  g.Cb = dZ;
  g.ldcb = 256;
  g.CbT = dZT;
  g.ldct = Pe;
  g.valid_rows = static_cast<int>(L.Pe);
  if ((rc = gemm(st, g)) != OTTX_OK) return rc;
  // encoder layer 0: dW0 = dZ0^T X0, db0
  if ((rc = weight_grad(st, dZT, Pe, at<bf16_t>(ws, L.X0) + Pe * P.kin, 256, P.kin, gpart, L.kchunks,
                        grad + P.W0)) != OTTX_OK)
    return rc;
  if ((rc = bias_grad(st, dZ, Pe, 256, gpart, L.kchunks, grad + P.b0)) != OTTX_OK) return rc;
  return OTTX_OK;
}

int ottx_ds_adam_step(const ottx_ds_model* md, const float* grad, float* m, float* v, int32_t step, float lr,
                      float beta1, float beta2, float eps, void* stream) {
  if (!md || !grad || !m || !v || step < 1) return OTTX_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t count = param_layout(md->d).total;
  const float bc1 = 1.f - powf(beta1, static_cast<float>(step));
  const float bc2 = 1.f - powf(beta2, static_cast<float>(step));
  hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(count)), dim3(256), 0, st, md->params, grad, m, v, count, lr, beta1,
                     beta2, eps, bc1, bc2);
  DS_LAUNCHED();
  return ottx_ds_refresh_bf16(md, stream);
}

}  // extern "C"

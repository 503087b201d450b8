// ottx: log-domain Sinkhorn on squared-Euclidean point clouds for gfx950.
// Host side of the C ABI declared in include/ottx.h plus the small kernels
// around the fused half-step (statistics, packing, loop state, dual terms).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <type_traits>
#include <vector>

#include "ottx_common.h"
#include "ottx_kernels.h"

namespace ottx {

// ---------------------------------------------------------------------------
// Small kernels
// ---------------------------------------------------------------------------

// Thread t's float64 sums of feature t % d and its square over rows
// r0 + t / d, r0 + t / d + G, ... < r1.  U rows per trip with every load
// unconditional (rows past the chunk reload its last row and count zero):
// the sums are short, so they are bound by latency.
__device__ __forceinline__ void chunk_sums(const float* __restrict__ pts, int64_t r0, int64_t r1, int d, int G,
                                           int t, double& a1, double& a2) {
  if (r1 <= r0) return;
  constexpr int U = 16;
  const int k = t % d, g = t / d;
  for (int64_t r = r0 + g; r < r1; r += static_cast<int64_t>(U) * G) {
    float v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t ru = r + static_cast<int64_t>(u) * G;
      const float w = pts[min(ru, r1 - 1) * d + k];
      v[u] = ru < r1 ? w : 0.f;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double w = v[u];
      a1 += w;
      a2 += w * w;
    }
  }
}

// Per-chunk, per-feature float64 sums of x and x^2 (deterministic two-stage
// reduction for mean(C), pointcloud.py:170-186, and for the centring shift).
// Thread t < G*d (G = 256 / d row groups) owns feature t % d and rows
// r0 + t / d, r0 + t / d + G, ...; groups are summed in a fixed order.
__global__ void stats_partial_kernel(const float* __restrict__ pts, int64_t np, int d,
                                     int64_t rows_per_chunk, double* __restrict__ part) {
  __shared__ double s1[256], s2[256];
  const int t = threadIdx.x;
  const int G = max(1, 256 / d);
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * rows_per_chunk;
  const int64_t r1 = min(np, r0 + rows_per_chunk);
  double a1 = 0.0, a2 = 0.0;
  if (t < G * d) chunk_sums(pts, r0, r1, d, G, t, a1, a2);
  s1[t] = a1;
  s2[t] = a2;
  __syncthreads();
  if (t < d) {
    double b1 = 0.0, b2 = 0.0;
    for (int g = 0; g < G; ++g) {
      b1 += s1[g * d + t];
      b2 += s2[g * d + t];
    }
    part[(static_cast<int64_t>(blockIdx.x) * d + t) * 2 + 0] = b1;
    part[(static_cast<int64_t>(blockIdx.x) * d + t) * 2 + 1] = b2;
  }
}

// One workgroup of kFinalThreads: combine the chunks, write mean(C) (out[0]), the common
// centring shift mu = (mean x + mean y) / 2 (float, [d]) and, when ``sc`` is
// given, the solve's scalars: eps = rel * mean(C) if rel > 0 else eps_abs.
constexpr int kFinalThreads = 1024;

__global__ __launch_bounds__(kFinalThreads) void stats_final_kernel(
    const double* __restrict__ px, int cx, const double* __restrict__ py, int cy, int d, int64_t n, int64_t m,
    double s, double* __restrict__ out, float* __restrict__ mu, SolveScalars* sc, double rel, double eps_abs,
    double tau_a, double tau_b) {
  __shared__ double red[kFinalThreads];
  __shared__ double sums[4][256];
  __shared__ double acc3[3][256];
  const int t = threadIdx.x;
  // thread (value v, group g): values v < 2 d are x's (feature v / 2, sum or
  // sum of squares), the rest y's, in the partials' order; group g sums the
  // chunks g, g + G, ... with all the loads of a trip in flight
  const int V = 4 * d;
  const int G = max(1, kFinalThreads / V);
  const int v = t % V, g = t / V;
  double a = 0.0;
  if (t < G * V) {
    const bool side_y = v >= 2 * d;
    const double* p = (side_y ? py : px) + (side_y ? v - 2 * d : v);
    const int C = side_y ? cy : cx;
    constexpr int U = 16;
    for (int c = g; c < C; c += U * G) {
      double w[U];
#pragma unroll
      for (int u = 0; u < U; ++u) w[u] = c + u * G < C ? p[static_cast<int64_t>(c + u * G) * 2 * d] : 0.0;
#pragma unroll
      for (int u = 0; u < U; ++u) a += w[u];
    }
  }
  red[t] = a;
  __syncthreads();
  for (int e = t; e < V; e += kFinalThreads) {
    double s2 = 0.0;
    for (int gg = 0; gg < G; ++gg) s2 += red[gg * V + e];
    // e = side * 2 d + 2 k + j -> sums[2 side + j][k]
    sums[2 * (e / (2 * d)) + (e & 1)][(e % (2 * d)) >> 1] = s2;
  }
  __syncthreads();
  double nx2 = 0.0, ny2 = 0.0, dot = 0.0;
  if (t < d) {
    const double mxk = sums[0][t] / static_cast<double>(n);
    const double myk = sums[2][t] / static_cast<double>(m);
    // any common shift leaves the cost unchanged: a feature whose mean is not
    // finite (a NaN or inf input) is not shifted, so a bad row stays confined
    // to its own costs, as in the reference, instead of reaching every point
    const double shift = 0.5 * (mxk + myk);
    if (mu) mu[t] = isfinite(shift) ? static_cast<float>(shift) : 0.f;
    nx2 = sums[1][t];
    ny2 = sums[3][t];
    dot = mxk * myk;
  }
  if (t < 256) {
    acc3[0][t] = nx2;
    acc3[1][t] = ny2;
    acc3[2][t] = dot;
  }
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st)
      for (int q = 0; q < 3; ++q) acc3[q][t] += acc3[q][t + st];
    __syncthreads();
  }
  if (t == 0) {
    const double mean_c =
        s * (acc3[0][0] / static_cast<double>(n) + acc3[1][0] / static_cast<double>(m) - 2.0 * acc3[2][0]);
    out[0] = mean_c;
    if (sc) {
      const double eps = rel > 0.0 ? rel * mean_c : eps_abs;
      sc->eps = eps;
      sc->eps_f = static_cast<float>(eps);
      sc->c = static_cast<float>(2.0 * s * 1.4426950408889634 / eps);
      sc->inv_rho_a = tau_a == 1.0 ? 0.f : static_cast<float>((1.0 - tau_a) / (eps * tau_a));
      sc->inv_rho_b = tau_b == 1.0 ? 0.f : static_cast<float>((1.0 - tau_b) / (eps * tau_b));
    }
  }
}

// ---------------------------------------------------------------------------
// Standard deviation of the cost matrix without forming it
// (geometry/geometry.py:119-136: sqrt(relu(mean(C^2) - mean(C)^2)), the means
// uniform over the n m entries).  With u = x - mu, v = y - mu (the cost is
// invariant under a common shift) and D_ij = |u_i - v_j|^2 = C_ij / s:
//   E[D]   = E|u|^2 + E|v|^2 - 2 <E u, E v>
//   E[D^2] = E|u|^4 + E|v|^4 + 2 E|u|^2 E|v|^2 + 4 tr(S_u S_v)
//            - 4 <E[|u|^2 u], E v> - 4 <E u, E[|v|^2 v]>
// with S_u = E[u u^T].  Per cloud the terms are row sums (vector segment) and
// the Gram matrix U^T U (Gram segment, 32 x 32 tiles of the upper triangle);
// everything is float64, so the cancellation in E[D^2] - E[D]^2 costs no
// accuracy that float32 inputs have.
// ---------------------------------------------------------------------------
constexpr int kMomGramTile = 32;

struct MomentArgs {
  const float* pts[2];
  int64_t np[2];
  int d;
  int nkb;                // feature blocks of 32
  int tiles;              // Gram tiles (kb <= lb pairs)
  const float* mu;
  int vchunks[2];
  int64_t vrows[2];
  double* vpart[2];       // [chunk][2 d + 2]: sum u_k, sum |u|^2 u_k, sum |u|^2, sum |u|^4
  int gchunks[2];
  int64_t grows[2];
  double* gpart[2];       // [chunk][tile][32 * 32]
  int seg[5];             // first block of: vec x, vec y, Gram x, Gram y; seg[4] = grid
};

// Rows of each moment chunk: a vector chunk holds at most 32 rows per wave,
// a Gram chunk a share of the rows that keeps about 8 chunks per cloud.
inline void moment_plan(int64_t np, int d, int* vchunks, int64_t* vrows, int* gchunks, int64_t* grows) {
  *vrows = std::max<int64_t>(128, (np + 127) / 128);
  *vchunks = static_cast<int>((np + *vrows - 1) / *vrows);
  *grows = std::max<int64_t>(256, ((np + 7) / 8 + 63) / 64 * 64);
  *gchunks = static_cast<int>((np + *grows - 1) / *grows);
  (void)d;
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Vector moments of rows [r0, r1): wave w takes rows r0 + w, r0 + w + 4, ...,
// eight at a time with every load issued first; lane l holds features
// l, l + 64, l + 128, l + 192.
__device__ void moment_vec_block(const float* __restrict__ pts, int64_t r0, int64_t r1, int d,
                                 const float* __restrict__ mu, double* __restrict__ out) {
  constexpr int Q = 4, U = 8;
  __shared__ double s_red[4][2 * 256 + 2];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  double mk[Q], au[Q], awu[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int k = lane + 64 * q;
    mk[q] = k < d ? static_cast<double>(mu[k]) : 0.0;
    au[q] = awu[q] = 0.0;
  }
  double aw = 0.0, aw2 = 0.0;
  for (int64_t rb = r0 + w; rb < r1; rb += 4 * U) {
    float v[U][Q];
#pragma unroll
    for (int uu = 0; uu < U; ++uu) {
      const int64_t r = min(rb + 4 * uu, r1 - 1);
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int k = lane + 64 * q;
        v[uu][q] = k < d ? pts[r * d + k] : 0.f;
      }
    }
#pragma unroll
    for (int uu = 0; uu < U; ++uu) {
      if (rb + 4 * uu >= r1) break;  // uniform over the wave
      double u[Q], sq = 0.0;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        u[q] = (lane + 64 * q < d) ? static_cast<double>(v[uu][q]) - mk[q] : 0.0;
        sq += u[q] * u[q];
      }
      sq = wave_sum_f64(sq);
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        au[q] += u[q];
        awu[q] += sq * u[q];
      }
      aw += sq;
      aw2 += sq * sq;
    }
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int k = lane + 64 * q;
    if (k < d) {
      s_red[w][k] = au[q];
      s_red[w][d + k] = awu[q];
    }
  }
  if (lane == 0) {
    s_red[w][2 * d] = aw;
    s_red[w][2 * d + 1] = aw2;
  }
  __syncthreads();
  for (int e = tid; e < 2 * d + 2; e += 256) out[e] = (s_red[0][e] + s_red[1][e]) + (s_red[2][e] + s_red[3][e]);
}

// Gram tile (kb, lb) of rows [r0, r1): sum_i u_ik u_il for k in block kb, l in
// block lb; 64 rows at a time through LDS, each thread four entries.
__device__ void moment_gram_block(const float* __restrict__ pts, int64_t r0, int64_t r1, int d,
                                  const float* __restrict__ mu, int kb, int lb, double* __restrict__ out) {
  constexpr int TT = kMomGramTile;
  __shared__ double s_k[64][TT + 1], s_l[64][TT + 1];
  const int tid = threadIdx.x, kk = tid & (TT - 1), l4 = (tid >> 5) * 4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const int c = tid & (TT - 1);
  const int kq = kb * TT + c, lq = lb * TT + c;
  const double mkq = kq < d ? static_cast<double>(mu[kq]) : 0.0;
  const double mlq = lq < d ? static_cast<double>(mu[lq]) : 0.0;
  for (int64_t rs = r0; rs < r1; rs += 64) {
    float vk[8], vl[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int r = (tid >> 5) + 8 * i;
      const int64_t row = min(rs + r, r1 - 1);
      vk[i] = kq < d ? pts[row * d + kq] : 0.f;
      vl[i] = lq < d ? pts[row * d + lq] : 0.f;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int r = (tid >> 5) + 8 * i;
      const bool in = rs + r < r1;
      s_k[r][c] = (in && kq < d) ? static_cast<double>(vk[i]) - mkq : 0.0;
      s_l[r][c] = (in && lq < d) ? static_cast<double>(vl[i]) - mlq : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int r = 0; r < 64; ++r) {
      const double uk = s_k[r][kk];
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[e] = fma(uk, s_l[r][l4 + e], acc[e]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) out[kk * TT + l4 + e] = acc[e];
}

// Gram tile index t -> (kb, lb), kb <= lb, row-major over the upper triangle.
__device__ __forceinline__ void gram_tile_blocks(int t, int nkb, int* kb, int* lb) {
  int k = 0;
  while (t >= nkb - k) {
    t -= nkb - k;
    ++k;
  }
  *kb = k;
  *lb = k + t;
}

__global__ __launch_bounds__(256) void moments_kernel(const MomentArgs a) {
  const int bid = blockIdx.x;
  if (bid < a.seg[2]) {
    const int side = bid >= a.seg[1] ? 1 : 0;
    const int chunk = bid - a.seg[side];
    const int64_t r0 = static_cast<int64_t>(chunk) * a.vrows[side];
    const int64_t r1 = min(a.np[side], r0 + a.vrows[side]);
    moment_vec_block(a.pts[side], r0, r1, a.d, a.mu, a.vpart[side] + static_cast<int64_t>(chunk) * (2 * a.d + 2));
    return;
  }
  const int side = bid >= a.seg[3] ? 1 : 0;
  const int local = bid - a.seg[2 + side];
  const int chunk = local / a.tiles, tile = local % a.tiles;
  int kb, lb;
  gram_tile_blocks(tile, a.nkb, &kb, &lb);
  const int64_t r0 = static_cast<int64_t>(chunk) * a.grows[side];
  const int64_t r1 = min(a.np[side], r0 + a.grows[side]);
  moment_gram_block(a.pts[side], r0, r1, a.d, a.mu, kb, lb,
                    a.gpart[side] + (static_cast<int64_t>(chunk) * a.tiles + tile) * kMomGramTile * kMomGramTile);
}

// One workgroup: sums the moment partials, writes std(C) to out[1] and, when
// ``sc`` is given and rel > 0, the solve's scalars with eps = rel * std(C).
__global__ __launch_bounds__(1024) void std_final_kernel(const MomentArgs a, double s, double* __restrict__ out,
                                                        SolveScalars* sc, double rel, double tau_a, double tau_b) {
  constexpr int NT = 1024, TT2 = kMomGramTile * kMomGramTile;
  __shared__ double s_vec[2][2 * 256 + 2];
  __shared__ double s_red[4][NT];
  const int t = threadIdx.x;
  const int d = a.d, V = 2 * d + 2;
  for (int e = t; e < 2 * V; e += NT) {
    const int side = e / V, v = e % V;
    const double* p = a.vpart[side] + v;
    double acc = 0.0;
    for (int c = 0; c < a.vchunks[side]; ++c) acc += p[static_cast<int64_t>(c) * V];
    s_vec[side][v] = acc;
  }
  // tr(G_u G_v): both Gram matrices summed over their chunks, entry by entry;
  // an off-diagonal tile stands for its mirror image too
  double tr = 0.0;
  for (int e = t; e < a.tiles * TT2; e += NT) {
    const int tile = e / TT2;
    int kb, lb;
    gram_tile_blocks(tile, a.nkb, &kb, &lb);
    double gx = 0.0, gy = 0.0;
    for (int c = 0; c < a.gchunks[0]; ++c) gx += a.gpart[0][static_cast<int64_t>(c) * a.tiles * TT2 + e];
    for (int c = 0; c < a.gchunks[1]; ++c) gy += a.gpart[1][static_cast<int64_t>(c) * a.tiles * TT2 + e];
    tr += (kb == lb ? 1.0 : 2.0) * gx * gy;
  }
  __syncthreads();
  const double n = static_cast<double>(a.np[0]), m = static_cast<double>(a.np[1]);
  double d_uv = 0.0, d_wu_v = 0.0, d_u_wv = 0.0;
  if (t < d) {
    const double mu_u = s_vec[0][t] / n, mu_v = s_vec[1][t] / m;
    d_uv = mu_u * mu_v;
    d_wu_v = (s_vec[0][d + t] / n) * mu_v;
    d_u_wv = mu_u * (s_vec[1][d + t] / m);
  }
  s_red[0][t] = tr;
  s_red[1][t] = d_uv;
  s_red[2][t] = d_wu_v;
  s_red[3][t] = d_u_wv;
  __syncthreads();
  for (int st = NT / 2; st > 0; st >>= 1) {
    if (t < st)
#pragma unroll
      for (int q = 0; q < 4; ++q) s_red[q][t] += s_red[q][t + st];
    __syncthreads();
  }
  if (t == 0) {
    const double e2u = s_vec[0][2 * d] / n, e4u = s_vec[0][2 * d + 1] / n;
    const double e2v = s_vec[1][2 * d] / m, e4v = s_vec[1][2 * d + 1] / m;
    const double trs = s_red[0][0] / (n * m);
    const double ed = e2u + e2v - 2.0 * s_red[1][0];
    const double ed2 = e4u + e4v + 2.0 * e2u * e2v + 4.0 * trs - 4.0 * s_red[2][0] - 4.0 * s_red[3][0];
    const double var = ed2 - ed * ed;
    const double sd = s * sqrt(var > 0.0 ? var : 0.0);
    out[1] = sd;
    if (sc && rel > 0.0) {
      const double eps = rel * sd;
      sc->eps = eps;
      sc->eps_f = static_cast<float>(eps);
      sc->c = static_cast<float>(2.0 * s * 1.4426950408889634 / eps);
      sc->inv_rho_a = tau_a == 1.0 ? 0.f : static_cast<float>((1.0 - tau_a) / (eps * tau_a));
      sc->inv_rho_b = tau_b == 1.0 ? 0.f : static_cast<float>((1.0 - tau_b) / (eps * tau_b));
    }
  }
}

// bf16 hi/lo fragments of scale (p - mu): [rb][kb][hi|lo][lane][8], lane ->
// row rb*16 + lane%16, features kb*32 + 8*(lane/16) + e.  The y cloud is
// packed with scale = c (2 s log2(e) / eps), the x cloud with 1, so the
// half-step's MFMA products come out as base-2 exponents.
__device__ __forceinline__ void pack_bf16_one(const float* __restrict__ pts, int64_t np, int d,
                                              const float* __restrict__ mu, float scale, int KB, int64_t gid,
                                              uint4* __restrict__ frag) {
  const int lane = static_cast<int>(gid & 63);
  const int64_t chunk = gid >> 6;
  const int kb = static_cast<int>(chunk % KB);
  const int64_t rb = chunk / KB;
  const int64_t row = rb * 16 + (lane & 15);
  const int k0 = kb * 32 + 8 * (lane >> 4);
  uint32_t hw[4], lw[4];
  // the 8 features as two 16-byte loads when the row holds all of them
  float in[8];
  if (row < np && k0 + 8 <= d && (d & 3) == 0 && (reinterpret_cast<uintptr_t>(pts) & 15) == 0) {
    const float4 p0 = *reinterpret_cast<const float4*>(pts + row * d + k0);
    const float4 p1 = *reinterpret_cast<const float4*>(pts + row * d + k0 + 4);
    in[0] = p0.x; in[1] = p0.y; in[2] = p0.z; in[3] = p0.w;
    in[4] = p1.x; in[5] = p1.y; in[6] = p1.z; in[7] = p1.w;
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e) in[e] = (row < np && k0 + e < d) ? pts[row * d + k0 + e] : 0.f;
  }
#pragma unroll
  for (int e = 0; e < 8; e += 2) {
    uint32_t h2[2], l2[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int k = k0 + e + q;
      const float v = (row < np && k < d) ? (in[e + q] - mu[k]) * scale : 0.f;
      const uint32_t h = f32_to_bf16_rne(v);
      const uint32_t l = f32_to_bf16_rne(v - bf16_to_f32(h));
      h2[q] = h;
      l2[q] = l;
    }
    hw[e / 2] = h2[0] | (h2[1] << 16);
    lw[e / 2] = l2[0] | (l2[1] << 16);
  }
  frag[(chunk * 2 + 0) * 64 + lane] = make_uint4(hw[0], hw[1], hw[2], hw[3]);
  frag[(chunk * 2 + 1) * 64 + lane] = make_uint4(lw[0], lw[1], lw[2], lw[3]);
}

// fp32 fragments of scale (p - mu): [rb][kq][lane][4], element r -> feature
// (4 kq + r) * 4 + lane/16.
__device__ __forceinline__ void pack_f32_one(const float* __restrict__ pts, int64_t np, int d,
                                             const float* __restrict__ mu, float scale, int KQ, int64_t gid,
                                             float4* __restrict__ frag) {
  const int lane = static_cast<int>(gid & 63);
  const int64_t chunk = gid >> 6;
  const int kq = static_cast<int>(chunk % KQ);
  const int64_t rb = chunk / KQ;
  const int64_t row = rb * 16 + (lane & 15);
  float v[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int k = (4 * kq + r) * 4 + (lane >> 4);
    v[r] = (row < np && k < d) ? (pts[row * d + k] - mu[k]) * scale : 0.f;
  }
  frag[chunk * 64 + lane] = make_float4(v[0], v[1], v[2], v[3]);
}

// s * |p - mu|^2 per point (float64 accumulation), zero on padded rows.
__global__ void norms_kernel(const float* __restrict__ pts, int64_t np, int d, const float* __restrict__ mu,
                             double s, int64_t np_pad, float* __restrict__ norm) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= np_pad) return;
  if (i >= np) {
    norm[i] = 0.f;
    return;
  }
  double acc = 0.0;
  for (int k = 0; k < d; ++k) {
    const double v = static_cast<double>(pts[i * d + k] - mu[k]);
    acc += v * v;
  }
  norm[i] = static_cast<float>(s * acc);
}

// Optional loop-state reset folded into the preparation launch.
struct PrepInit {
  SolveState* st = nullptr;
  float* errors = nullptr;
  int outer = 0;
  uint32_t* counters = nullptr;
  int n_counters = 0;
  int zero_init = 0;  // bit 0: f, bit 1: g start as DefaultInitializer's (0, -inf at zero weight)
  float* f = nullptr;
  float* g = nullptr;
  // fixed shift of the first f-update estimated from the given f (or null)
  float* lse_f = nullptr;
  double tau_a = 1.0;
  // fixed shift of the first g-update (or null): kShiftFromPot estimates it
  // from the given g, kShiftProbe from a row sample with the given f
  float* lse_g = nullptr;
  int lse_g_mode = 0;
  double tau_b = 1.0;
  const SolveScalars* scal = nullptr;
};

constexpr int kShiftFromPot = 1;
// Threads sharing one row in the preparation's norm segments (d <= 256).
constexpr int kNormLanes = 16;
constexpr int kShiftProbe = 2;
// Rows of x sampled by the first g-update's shift probe.
constexpr int kProbeRows = 64;

// Potentials the solve initialises itself for an init mode (PrepInit::zero_init bits).
inline int zero_init_bits(int init_mode) {
  return init_mode == OTTX_INIT_ZEROS ? 3 : init_mode == OTTX_INIT_F_GIVEN ? 2 : 0;
}

// Row r < K of the shift probe's sample: evenly strided over the n rows.
__device__ __forceinline__ int64_t probe_row(int r, int64_t n, int K) {
  const int64_t p = static_cast<int64_t>(r) * n;
  return K == kProbeRows ? p / kProbeRows : p / K;  // K = 64 unless n < 64: a shift, not a 64-bit division
}

constexpr int kProbeCols = 64;
constexpr int kProbeChunk = 64;

// First g-update's fixed shift from a row sample, for a solve that starts
// from a given f and no g.  For column j this stores the largest of
//   <x_r - x_0, y_j - x_0> + f_r / (2 s) - |x_r - x_0|^2 / 2
// over the K = min(64, n) evenly strided rows r: the kernel's acc_rj with the
// points centred on x_0 instead of mu, which differs from it by a term of j
// alone; prep_pack_kernel adds that term and half of log2(n) / c, since the
// log-sum-exp over all rows lies between the max and the max + log2(n) / c.
// Needing neither mu nor eps, the probe runs beside the statistics.  The value
// only sets the streaming shift of the first g-update: a tile whose sums
// leave range falls back to the running max.
//
// A workgroup of four waves takes 64 columns and wave w the sampled rows
// 16w .. 16w + 15: a 64 x 64 product over d, on the fp32 MFMA
// (v_mfma_f32_16x16x4f32, four 16 x 16 tiles a wave).  Lane (q, l) of a wave
// (q = lane / 16, l = lane % 16) feeds features 16q .. 16q + 15 of each 64-wide
// chunk, of sampled row 16w + l and of column 16nb + l; both operands come
// from LDS, where the chunk of the sampled rows, of the block's columns
// (read coalesced) and of x_0 is staged with padded rows.  Each chunk takes
// one round trip of loads, all issued before the first wait.  The waves
// merge their maxima through LDS.
__device__ void probe_max_block(int blk, const float* __restrict__ x, int64_t n, const float* __restrict__ y,
                                int64_t m, int64_t m_pad, int d, const float* __restrict__ f, double s,
                                float* __restrict__ out) {
  constexpr int KC = kProbeChunk;
  constexpr int PS = KC / 4 + 1;               // padded row of s_x and s_y in float4: no bank conflicts
  constexpr int NX = kProbeRows * KC / 256;    // floats of the row chunk per thread
  constexpr int NY = kProbeCols * KC / 1024;   // float4 of the column chunk per thread
  static_assert(kProbeRows == 64 && kProbeCols == 64 && KC == 64, "four waves of four 16 x 16 tiles");
  __shared__ float4 s_x[kProbeRows][PS];
  __shared__ float4 s_y[kProbeCols][PS];
  __shared__ float4 s_c[KC / 4];
  __shared__ float s_f[kProbeRows], s_nrm[kProbeRows];
  __shared__ float s_mx[4][kProbeCols];
  const int tid = threadIdx.x, lane = tid & 63, q = lane >> 4, l16 = lane & 15;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int K = static_cast<int>(min<int64_t>(kProbeRows, n));
  const int64_t col0 = static_cast<int64_t>(blk) * kProbeCols;
  const bool vec16 = (d & 3) == 0 && (reinterpret_cast<uintptr_t>(y) & 15) == 0;
  if (tid < kProbeRows) {
    s_f[tid] = tid < K ? f[probe_row(tid, n, K)] : 0.f;
    s_nrm[tid] = 0.f;
  }
  f32x4 acc[4];
  float ycp[4];
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
    ycp[nb] = 0.f;
  }
  for (int k0 = 0; k0 < d; k0 += KC) {
    const int kc = min(KC, d - k0);
    // the chunk of the block's columns, of the sampled rows and of x_0, zero past d, m and K
    float4 ys[NY];
#pragma unroll
    for (int i = 0; i < NY; ++i) {
      const int e = tid + 256 * i, k = 4 * (e % (KC / 4));
      const int64_t j = col0 + e / (KC / 4);
      const float* src = y + j * d + k0 + k;
      if (j >= m || k >= kc) {
        ys[i] = make_float4(0.f, 0.f, 0.f, 0.f);
      } else if (vec16) {
        ys[i] = *reinterpret_cast<const float4*>(src);
      } else {
        ys[i].x = src[0];
        ys[i].y = k + 1 < kc ? src[1] : 0.f;
        ys[i].z = k + 2 < kc ? src[2] : 0.f;
        ys[i].w = k + 3 < kc ? src[3] : 0.f;
      }
    }
    float xs[NX];
#pragma unroll
    for (int i = 0; i < NX; ++i) {
      const int e = tid + 256 * i, r = e / KC, k = e % KC;
      xs[i] = (r < K && k < kc) ? x[probe_row(r, n, K) * d + k0 + k] : 0.f;
    }
    const float ck = (tid < KC && tid < kc) ? x[k0 + tid] : 0.f;
    __syncthreads();  // the previous chunk's readers are done
    float* sxf = reinterpret_cast<float*>(s_x);
#pragma unroll
    for (int i = 0; i < NX; ++i) {
      const int e = tid + 256 * i;
      sxf[(e / KC) * (4 * PS) + e % KC] = xs[i];
    }
#pragma unroll
    for (int i = 0; i < NY; ++i) {
      const int e = tid + 256 * i;
      s_y[e / (KC / 4)][e % (KC / 4)] = ys[i];
    }
    if (tid < KC) reinterpret_cast<float*>(s_c)[tid] = ck;
    __syncthreads();
    // |x_r - x_0|^2 over the chunk: four threads a row
    {
      const int r = tid >> 2, qr = tid & 3;
      float nr = 0.f;
#pragma unroll
      for (int i = 0; i < KC / 16; ++i) {
        const float4 v = s_x[r][qr * (KC / 16) + i], u = s_c[qr * (KC / 16) + i];
        const float a = v.x - u.x, b = v.y - u.y, c = v.z - u.z, e = v.w - u.w;
        nr += a * a + b * b + c * c + e * e;
      }
      nr += __shfl_xor(nr, 1);
      nr += __shfl_xor(nr, 2);
      if (qr == 0) s_nrm[r] += nr;
    }
    // <x_r, y_j - x_0> on the MFMA, and the lane's part of <x_0, y_j - x_0>
    float4 cv[4], av[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      cv[i] = s_c[4 * q + i];
      av[i] = s_x[16 * w + l16][4 * q + i];
    }
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float4 v = s_y[16 * nb + l16][4 * q + i];
        v.x -= cv[i].x;
        v.y -= cv[i].y;
        v.z -= cv[i].z;
        v.w -= cv[i].w;
        ycp[nb] += cv[i].x * v.x + cv[i].y * v.y + cv[i].z * v.z + cv[i].w * v.w;
        acc[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i].x, v.x, acc[nb], 0, 0, 0);
        acc[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i].y, v.y, acc[nb], 0, 0, 0);
        acc[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i].z, v.z, acc[nb], 0, 0, 0);
        acc[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i].w, v.w, acc[nb], 0, 0, 0);
      }
    }
  }
  __syncthreads();
  // element i of tile nb in lane (q, l): sampled row 16w + 4q + i, column 16nb + l
  const float inv2s = static_cast<float>(0.5 / s);
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    float yc = ycp[nb];
    yc += __shfl_xor(yc, 16);
    yc += __shfl_xor(yc, 32);
    float M = -INFINITY;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = 16 * w + 4 * q + i;
      if (r >= K) continue;
      const float a = acc[nb][i] - yc + s_f[r] * inv2s - 0.5f * s_nrm[r];
      if (isfinite(a)) M = fmaxf(M, a);
    }
    M = fmaxf(M, __shfl_xor(M, 16));
    M = fmaxf(M, __shfl_xor(M, 32));
    if (q == 0) s_mx[w][16 * nb + l16] = M;
  }
  __syncthreads();
  if (tid < kProbeCols) {
    const int64_t j = col0 + tid;
    if (j >= m_pad) return;
    const float M = fmaxf(fmaxf(s_mx[0][tid], s_mx[1][tid]), fmaxf(s_mx[2][tid], s_mx[3][tid]));
    out[j] = j < m ? M : -INFINITY;
  }
}

// Statistics partials of both clouds in one launch: blocks [0, cx) take x,
// blocks [cx, cx + cy) take y (same per-chunk sums as stats_partial_kernel).
// With ``probe`` set, the blocks past them run probe_max_block with the
// given f.
__global__ __launch_bounds__(256) void prep_stats_kernel(const float* __restrict__ x, int64_t n, int64_t rows_x, int cx,
                                  const float* __restrict__ y, int64_t m, int64_t rows_y, int cy, int d,
                                  double* __restrict__ part_x, double* __restrict__ part_y,
                                  const float* __restrict__ f, double s, int64_t m_pad, float* __restrict__ probe) {
  __shared__ double s1[256], s2[256];
  const int t = threadIdx.x;
  if (static_cast<int>(blockIdx.x) >= cx + cy) {
    probe_max_block(static_cast<int>(blockIdx.x) - cx - cy, x, n, y, m, m_pad, d, f, s, probe);
    return;
  }
  const bool side_y = static_cast<int>(blockIdx.x) >= cx;
  const int chunk = side_y ? static_cast<int>(blockIdx.x) - cx : static_cast<int>(blockIdx.x);
  const float* pts = side_y ? y : x;
  const int64_t np = side_y ? m : n;
  const int64_t rows = side_y ? rows_y : rows_x;
  double* part = side_y ? part_y : part_x;
  const int G = max(1, 256 / d);
  const int64_t r0 = static_cast<int64_t>(chunk) * rows;
  const int64_t r1 = min(np, r0 + rows);
  double a1 = 0.0, a2 = 0.0;
  if (t < G * d) chunk_sums(pts, r0, r1, d, G, t, a1, a2);
  s1[t] = a1;
  s2[t] = a2;
  __syncthreads();
  if (t < d) {
    double b1 = 0.0, b2 = 0.0;
    for (int g = 0; g < G; ++g) {
      b1 += s1[g * d + t];
      b2 += s2[g * d + t];
    }
    part[(static_cast<int64_t>(chunk) * d + t) * 2 + 0] = b1;
    part[(static_cast<int64_t>(chunk) * d + t) * 2 + 1] = b2;
  }
}

// Everything that follows the statistics, in one launch: fragments and norms
// of both clouds, uniform weights where the caller passed none, and (for a
// solve) the loop-state reset.  Blocks are split into consecutive segments.
struct PrepArgs {
  const float* pts[2];
  int64_t np[2], np_pad[2];
  void* frag[2];
  float* norm[2];
  float* wfill[2];       // uniform weights to write, or null
  const float* wuser[2]; // the caller's weights (null: uniform)
  int64_t pack_threads[2];
  int seg[7];            // first block of: pack x, pack y, norm x, norm y, w x, w y, init; seg[6] + blocks_init = grid
  int blocks_init;
  const float* mu;
  const SolveScalars* scal;  // c of the y fragments
  int d, chunks, prec;
  double s;
  PrepInit init;
};

__global__ void prep_pack_kernel(const PrepArgs p) {
  const int bid = blockIdx.x;
  const int t = threadIdx.x;
  int segi = 0;
  while (segi < 6 && bid >= p.seg[segi + 1]) ++segi;
  const int64_t local = static_cast<int64_t>(bid - p.seg[segi]) * blockDim.x + t;
  if (segi < 2) {  // fragments
    const int side = segi;
    if (local >= p.pack_threads[side]) return;
    const float scale = side == 1 ? p.scal->c : 1.f;
    if (p.prec == 1)
      pack_bf16_one(p.pts[side], p.np[side], p.d, p.mu, scale, p.chunks, local, static_cast<uint4*>(p.frag[side]));
    else
      pack_f32_one(p.pts[side], p.np[side], p.d, p.mu, scale, p.chunks, local, static_cast<float4*>(p.frag[side]));
  } else if (segi < 4) {  // norms s |p - mu|^2, kNormLanes threads a row
    const int side = segi - 2;
    const int64_t row = local / kNormLanes;
    const int ln = static_cast<int>(local % kNormLanes);
    if (row >= p.np_pad[side]) return;  // whole rows leave together
    float* norm = p.norm[side];
    const PrepInit& in = p.init;
    const bool est_f = side == 0 && in.lse_f != nullptr;
    const bool est_g = side == 1 && in.lse_g != nullptr && in.lse_g_mode == kShiftFromPot;
    const bool probe = side == 1 && in.lse_g != nullptr && in.lse_g_mode == kShiftProbe;
    if (row >= p.np[side]) {
      if (ln == 0) {
        norm[row] = 0.f;
        if (est_f) in.lse_f[row] = -INFINITY;
        if (est_g) in.lse_g[row] = -INFINITY;
      }
      return;
    }
    // each lane's features, all loads of the row in flight at once; with the
    // probe also the term that moves its centring from x_0 to mu,
    // <mu - x_0, (mu + x_0) / 2 - y_j>
    const float* pt = p.pts[side] + row * p.d;
    const float* x0 = p.pts[0];
    const bool vec16 = (p.d & 3) == 0 && (reinterpret_cast<uintptr_t>(p.pts[side]) & 15) == 0 &&
                       (reinterpret_cast<uintptr_t>(x0) & 15) == 0;
    double a = 0.0, cr = 0.0;
    if (vec16) {
#pragma unroll
      for (int i = 0; i < 256 / (4 * kNormLanes); ++i) {
        const int k = 4 * (ln + kNormLanes * i);
        if (k < p.d) {
          const float4 v = *reinterpret_cast<const float4*>(pt + k);
          const float4 u = *reinterpret_cast<const float4*>(p.mu + k);
          const double e0 = static_cast<double>(v.x - u.x), e1 = static_cast<double>(v.y - u.y);
          const double e2 = static_cast<double>(v.z - u.z), e3 = static_cast<double>(v.w - u.w);
          a += (e0 * e0 + e1 * e1) + (e2 * e2 + e3 * e3);
          if (probe) {
            const float4 z = *reinterpret_cast<const float4*>(x0 + k);
            cr += (static_cast<double>(u.x) - z.x) * (0.5 * (static_cast<double>(u.x) + z.x) - v.x) +
                  (static_cast<double>(u.y) - z.y) * (0.5 * (static_cast<double>(u.y) + z.y) - v.y) +
                  (static_cast<double>(u.z) - z.z) * (0.5 * (static_cast<double>(u.z) + z.z) - v.z) +
                  (static_cast<double>(u.w) - z.w) * (0.5 * (static_cast<double>(u.w) + z.w) - v.w);
          }
        }
      }
    } else {
      for (int k = ln; k < p.d; k += kNormLanes) {
        const double e = static_cast<double>(pt[k] - p.mu[k]);
        a += e * e;
        if (probe) {
          const double u = p.mu[k], z = x0[k];
          cr += (u - z) * (0.5 * (u + z) - static_cast<double>(pt[k]));
        }
      }
    }
#pragma unroll
    for (int off = 1; off < kNormLanes; off <<= 1) {
      a += __shfl_xor(a, off, kNormLanes);
      cr += __shfl_xor(cr, off, kNormLanes);
    }
    if (ln != 0) return;
    const float nrm = static_cast<float>(p.s * a);
    norm[row] = nrm;
    if (est_f) {
      // The first f-update's fixed shift from the given f: the column
      // log-sum-exp (base 2, in exponent units) that the f-update's epilogue
      // would map back to f itself, L = c (eps log a - f / tau_a + s|x - mu|^2) / (2 s).
      // A warm start puts the first update's result close to it; non-finite
      // values get -inf, so their tiles use the running max, and a tile whose
      // sums leave range falls back to it as well.
      const double w = p.wuser[0] ? static_cast<double>(p.wuser[0][row]) : 1.0 / static_cast<double>(p.np[0]);
      const double v = static_cast<double>(in.scal->c) *
                       (in.scal->eps * log(w) - static_cast<double>(in.f[row]) / in.tau_a + nrm) / (2.0 * p.s);
      in.lse_f[row] = isfinite(v) ? static_cast<float>(v) : -INFINITY;
    }
    if (est_g) {
      // the same estimate for the first g-update from the given g
      const double w = p.wuser[1] ? static_cast<double>(p.wuser[1][row]) : 1.0 / static_cast<double>(p.np[1]);
      const double v = static_cast<double>(in.scal->c) *
                       (in.scal->eps * log(w) - static_cast<double>(in.g[row]) / in.tau_b + nrm) / (2.0 * p.s);
      in.lse_g[row] = isfinite(v) ? static_cast<float>(v) : -INFINITY;
    }
    if (probe) {
      // the probe's max, centred on x_0, moved to the centring on mu, in
      // exponent units, plus half of log2(n)
      const float M = in.lse_g[row];
      if (isfinite(M)) {
        const double L = (static_cast<double>(M) + cr) * static_cast<double>(in.scal->c) +
                         0.5 * log2(static_cast<double>(p.np[0]));
        in.lse_g[row] = isfinite(L) ? static_cast<float>(L) : -INFINITY;
      }
    }
  } else if (segi < 6) {  // uniform weights
    const int side = segi - 4;
    if (local < p.np[side]) p.wfill[side][local] = static_cast<float>(1.0 / static_cast<double>(p.np[side]));
  } else {  // loop state
    const PrepInit& in = p.init;
    const int64_t stride = static_cast<int64_t>(p.blocks_init) * blockDim.x;
    if (local == 0) {
      SolveState* st = in.st;
      st->done = 0;
      st->block = 0;
      st->n_blocks = 0;
      st->finalized = 0;
      st->tiles_done = 0u;
      st->shift_ok[0] = in.lse_f != nullptr ? 1 : 0;
      st->shift_ok[1] = in.lse_g != nullptr ? 1 : 0;
      st->fin_count = 0u;
      st->total_mass = 0.0;
      st->last_err = 0.0;
    }
    for (int64_t k = local; k < in.outer; k += stride) in.errors[k] = -1.f;
    for (int64_t k = local; k < in.n_counters; k += stride) in.counters[k] = 0u;
    // DefaultInitializer: 0, -inf where the weight is 0
    if (in.zero_init & 1)
      for (int64_t k = local; k < p.np[0]; k += stride)
        in.f[k] = (p.wuser[0] == nullptr || p.wuser[0][k] > 0.f) ? 0.f : -INFINITY;
    if (in.zero_init & 2)
      for (int64_t k = local; k < p.np[1]; k += stride)
        in.g[k] = (p.wuser[1] == nullptr || p.wuser[1][k] > 0.f) ? 0.f : -INFINITY;
  }
}

// Loop state reset: errors = -1 (sinkhorn.py:964-972), counters = 0.  With
// ``zero_init`` bit 0 (f) / bit 1 (g) the potentials start as
// DefaultInitializer's: 0, or -inf where the weight is 0 (initializers.py:70-99).
__global__ void init_state_kernel(SolveState* st, float* errors, int outer, uint32_t* counters, int n_counters,
                                  int zero_init, float* f, const float* a, int64_t n, float* g, const float* b,
                                  int64_t m) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  if (i == 0) {
    st->done = 0;
    st->block = 0;
    st->n_blocks = 0;
    st->finalized = 0;
    st->tiles_done = 0u;
    st->shift_ok[0] = 0;
    st->shift_ok[1] = 0;
    st->fin_count = 0u;
    st->total_mass = 0.0;
    st->last_err = 0.0;
  }
  for (int64_t k = i; k < outer; k += stride) errors[k] = -1.f;
  for (int64_t k = i; k < n_counters; k += stride) counters[k] = 0u;
  if (zero_init & 1)
    for (int64_t k = i; k < n; k += stride) f[k] = a[k] > 0.f ? 0.f : -INFINITY;
  if (zero_init & 2)
    for (int64_t k = i; k < m; k += stride) g[k] = b[k] > 0.f ? 0.f : -INFINITY;
}

// out[0] = sum of v[0, n) (one workgroup, fixed order).
__global__ void sum_f64_kernel(const double* __restrict__ v, int n, double* __restrict__ out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += v[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

__global__ void copy_f32_kernel(const float* __restrict__ src, float* __restrict__ dst, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// Float64 dual terms of compute_kl_reg_cost (sinkhorn.py:224-277) and the
// errors summary of output_from_state (sinkhorn.py:1020-1023, 431-434), as a
// deterministic two-stage reduction.  Stage 1, grid (kDualBlocks, 3):
//   y = 0: partial div_a and sum a over a slice;  y = 1: same for b / g
//   y = 2 (block 0 only): #NaN and #(errors != -1)
constexpr int kDualBlocks = 32;

__global__ void dual_partial_kernel(const float* __restrict__ f, const float* __restrict__ a, int64_t n,
                                    const float* __restrict__ g, const float* __restrict__ b, int64_t m,
                                    const SolveScalars* __restrict__ sc, double tau_a, double tau_b,
                                    const float* __restrict__ errors, int outer, double* __restrict__ part) {
  __shared__ double red[2][256];
  const int tid = threadIdx.x;
  const int side = blockIdx.y;
  double s0 = 0.0, s1 = 0.0;
  if (side < 2) {
    const float* pot = side == 0 ? f : g;
    const float* w = side == 0 ? a : b;
    const int64_t len = side == 0 ? n : m;
    const double tau = side == 0 ? tau_a : tau_b;
    const double eps = sc->eps;
    const double rho = eps * tau / (1.0 - tau);
    const int64_t per = (len + kDualBlocks - 1) / kDualBlocks;
    const int64_t i0 = blockIdx.x * per, i1 = min(len, i0 + per);
    for (int64_t i = i0 + tid; i < i1; i += blockDim.x) {
      const double wi = w[i];
      s1 += wi;
      if (wi > 0.0) {
        const double h = static_cast<double>(pot[i]) - eps * log(wi);
        if (tau == 1.0) {
          s0 += wi * h;
        } else {
          s0 -= wi * rho * (exp(-h / rho) - 1.0);
        }
      }
    }
  } else if (blockIdx.x == 0) {
    for (int i = tid; i < outer; i += blockDim.x) {
      const float e = errors[i];
      if (e != e) s0 += 1.0;
      if (e != -1.f) s1 += 1.0;
    }
  }
  red[0][tid] = s0;
  red[1][tid] = s1;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) {
      red[0][tid] += red[0][tid + st];
      red[1][tid] += red[1][tid + st];
    }
    __syncthreads();
  }
  if (tid == 0) {
    part[(side * kDualBlocks + blockIdx.x) * 2 + 0] = red[0][0];
    part[(side * kDualBlocks + blockIdx.x) * 2 + 1] = red[1][0];
  }
}

// Stage 2 (one thread): out[0] = div_a, [1] = sum a, [2] = div_b, [3] = sum b,
// [4] = #NaN errors, [5] = #(errors != -1), [6] = errors[-1], [7] = eps.
__global__ void dual_final_kernel(const double* __restrict__ part, const float* __restrict__ errors, int outer,
                                  const SolveScalars* __restrict__ sc, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  for (int side = 0; side < 3; ++side) {
    double s0 = 0.0, s1 = 0.0;
    for (int b = 0; b < kDualBlocks; ++b) {
      s0 += part[(side * kDualBlocks + b) * 2 + 0];
      s1 += part[(side * kDualBlocks + b) * 2 + 1];
    }
    out[side * 2 + 0] = s0;
    out[side * 2 + 1] = s1;
  }
  out[6] = outer > 0 ? static_cast<double>(errors[outer - 1]) : 0.0;
  out[7] = sc->eps;
}

// Cost maxima for the scale_cost modes: the largest squared norm of a cloud
// and the largest squared distance of a pair, as float bits (non-negative
// floats order as their bits) combined with one atomicMax per workgroup.
__global__ void zero_words_kernel(uint32_t* p, int count) {
  if (static_cast<int>(threadIdx.x) < count) p[threadIdx.x] = 0u;
}

__device__ __forceinline__ void block_max_store(float v, uint32_t* out) {
  __shared__ float red[4];
  // NaN-propagating (jnp.max returns NaN); a NaN's bits exceed every
  // non-negative float's, so the unsigned atomicMax keeps it too
  for (int off = 32; off > 0; off >>= 1) v = nan_max(v, __shfl_xor(v, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(out, __float_as_uint(nan_max(nan_max(red[0], red[1]), nan_max(red[2], red[3]))));
}

__global__ __launch_bounds__(256) void max_sq_norm_kernel(const float* __restrict__ p, int64_t np, int d,
                                                          uint32_t* out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  float v = 0.f;
  if (i < np) {
    const float* r = p + i * d;
    float a0 = 0.f, a1 = 0.f;
    int k = 0;
    for (; k + 1 < d; k += 2) {
      a0 = fmaf(r[k], r[k], a0);
      a1 = fmaf(r[k + 1], r[k + 1], a1);
    }
    if (k < d) a0 = fmaf(r[k], r[k], a0);
    v = a0 + a1;
  }
  block_max_store(v, out);
}

__global__ __launch_bounds__(256) void max_cost_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                       int64_t n, int64_t m, int d, uint32_t* out) {
  const int64_t i = static_cast<int64_t>(blockIdx.y) * 4 + (threadIdx.x >> 6);
  const int64_t j = static_cast<int64_t>(blockIdx.x) * 64 + (threadIdx.x & 63);
  float v = 0.f;
  if (i < n && j < m) {
    const float* xi = x + i * d;
    const float* yj = y + j * d;
    float c0 = 0.f, c1 = 0.f;
    int k = 0;
    for (; k + 1 < d; k += 2) {
      const float u = xi[k] - yj[k], w = xi[k + 1] - yj[k + 1];
      c0 = fmaf(u, u, c0);
      c1 = fmaf(w, w, c1);
    }
    if (k < d) {
      const float u = xi[k] - yj[k];
      c0 = fmaf(u, u, c0);
    }
    v = c0 + c1;
  }
  block_max_store(v, out);
}

// Host view of a problem's cost: c = alpha (sum_k phi(z_k))^beta.
struct CostForm {
  int phi;
  float p, alpha, beta;
};

inline bool cost_form(int cost, double p, CostForm* cf) {
  switch (cost) {
    case OTTX_COST_SQEUCLIDEAN: *cf = CostForm{2, 2.f, 1.f, 1.f}; return true;
    case OTTX_COST_EUCLIDEAN: *cf = CostForm{2, 2.f, 1.f, 0.5f}; return true;
    default: break;
  }
  if (!(p >= 1.0 && p <= 1e4)) return false;
  const float pf = static_cast<float>(p);
  switch (cost) {
    case OTTX_COST_PNORMP:  // sum |z|^p / p
      *cf = p == 1.0 ? CostForm{1, 1.f, 1.f, 1.f}
                     : (p == 2.0 ? CostForm{2, 2.f, 0.5f, 1.f} : CostForm{3, pf, static_cast<float>(1.0 / p), 1.f});
      return true;
    case OTTX_COST_SQPNORM:  // (sum |z|^p)^(2 / p) / 2
      *cf = p == 1.0 ? CostForm{1, 1.f, 0.5f, 2.f}
                     : (p == 2.0 ? CostForm{2, 2.f, 0.5f, 1.f} : CostForm{3, pf, 0.5f, static_cast<float>(2.0 / p)});
      return true;
    case OTTX_COST_EUCLIDEANP:  // (sum z^2)^(p / 2)
      *cf = CostForm{2, 2.f, 1.f, static_cast<float>(0.5 * p)};
      return true;
    default: return false;
  }
}

// Key of the half-step kernel family: the precision for the squared-Euclidean
// cost, 10 + phi for the vector-ALU cost stream (PREC 2, one variant).
inline int kernel_key(const ottx_problem* pb) {
  if (pb->cost == OTTX_COST_SQEUCLIDEAN) return pb->precision;
  CostForm cf;
  return cost_form(pb->cost, pb->cost_p, &cf) ? 10 + cf.phi : -1;
}

// Dense transport matrix exp((f_i + g_j - s c(x_i, y_j)) / eps), row-major
// [n, m]: four rows by 64 columns per workgroup, one entry per thread, the
// cost summed from the differences (cost_phi_acc / cost_psi).
template <int PHI>
__global__ void transport_matrix_kernel(const float* __restrict__ x, const float* __restrict__ y, int64_t n,
                                        int64_t m, int d, float s, const float* __restrict__ f,
                                        const float* __restrict__ g, float inv_eps, float* __restrict__ out,
                                        CostForm cf) {
  const int64_t i = static_cast<int64_t>(blockIdx.y) * 4 + (threadIdx.x >> 6);
  const int64_t j = static_cast<int64_t>(blockIdx.x) * 64 + (threadIdx.x & 63);
  if (i >= n || j >= m) return;
  const float* xi = x + i * d;
  const float* yj = y + j * d;
  float c0 = 0.f, c1 = 0.f;
  int k = 0;
  for (; k + 1 < d; k += 2) {
    c0 = cost_phi_acc<PHI>(c0, xi[k] - yj[k], cf.p);
    c1 = cost_phi_acc<PHI>(c1, xi[k + 1] - yj[k + 1], cf.p);
  }
  if (k < d) c0 = cost_phi_acc<PHI>(c0, xi[k] - yj[k], cf.p);
  out[i * m + j] = expf((f[i] + g[j] - s * cost_psi(c0 + c1, cf.alpha, cf.beta)) * inv_eps);
}

// <P, C'> in one pass over the pairs (SinkhornOutput.primal_cost and
// transport_cost_at_geom, sinkhorn.py:345-411): P_ij = exp((f_i + g_j - s
// c(x_i, y_j)) / eps) and C'_ij = s c(x_i, y_j) itself or s' |x'_i - y'_j|^2
// of another pair of clouds, costs summed from the differences in float32.  A workgroup takes 64 rows by 64
// columns; each thread 4 x 4 pairs, with the features staged through LDS,
// feature-major, 32 at a time.  One float64 partial per workgroup.
constexpr int kTcTile = 64;
constexpr int kTcChunk = 32;
constexpr int kTcPad = kTcTile + 4;  // keeps float4 rows aligned, spreads the staging stores

// Per-feature cost terms of a 64 x 64 block of pairs, sum_k phi(x_ik - y_jk)
// (phi as cost_phi_acc), 4 x 4 pairs per thread, features staged through LDS.
template <int PHI>
__device__ __forceinline__ void tc_cost_acc(const float* __restrict__ x, const float* __restrict__ y, int64_t n,
                                            int64_t m, int d, int64_t i0, int64_t j0, float (*xs)[kTcPad],
                                            float (*ys)[kTcPad], int tid, float (&acc)[4][4], float p) {
  const int r0 = (tid >> 4) * 4, c0 = (tid & 15) * 4;
  for (int k0 = 0; k0 < d; k0 += kTcChunk) {
    const int kc = min(kTcChunk, d - k0);
    __syncthreads();
    for (int e = tid; e < kTcTile * kTcChunk; e += 256) {
      const int r = e / kTcChunk, k = e % kTcChunk;
      const int64_t i = i0 + r, j = j0 + r;
      xs[k][r] = (k < kc && i < n) ? x[i * d + k0 + k] : 0.f;
      ys[k][r] = (k < kc && j < m) ? y[j * d + k0 + k] : 0.f;
    }
    __syncthreads();
    for (int k = 0; k < kc; ++k) {
      const float4 xv = *reinterpret_cast<const float4*>(&xs[k][r0]);
      const float4 yv = *reinterpret_cast<const float4*>(&ys[k][c0]);
      const float xa[4] = {xv.x, xv.y, xv.z, xv.w};
      const float ya[4] = {yv.x, yv.y, yv.z, yv.w};
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = cost_phi_acc<PHI>(acc[a][b], xa[a] - ya[b], p);
    }
  }
}

__device__ __forceinline__ void tc_sq_dist(const float* __restrict__ x, const float* __restrict__ y, int64_t n,
                                           int64_t m, int d, int64_t i0, int64_t j0, float (*xs)[kTcPad],
                                           float (*ys)[kTcPad], int tid, float (&acc)[4][4]) {
  tc_cost_acc<2>(x, y, n, m, d, i0, j0, xs, ys, tid, acc, 2.f);
}

template <bool SAME, int PHI>
__global__ __launch_bounds__(256) void transport_cost_kernel(
    const float* __restrict__ x, const float* __restrict__ y, int64_t n, int64_t m, int d, float s,
    const float* __restrict__ f, const float* __restrict__ g, float inv_eps, const float* __restrict__ xo,
    const float* __restrict__ yo, int d_o, float s_o, double* __restrict__ partial, CostForm cf) {
  __shared__ __align__(16) float xs[kTcChunk][kTcPad];
  __shared__ __align__(16) float ys[kTcChunk][kTcPad];
  __shared__ double red[256];
  const int tid = threadIdx.x;
  const int64_t i0 = static_cast<int64_t>(blockIdx.y) * kTcTile;
  const int64_t j0 = static_cast<int64_t>(blockIdx.x) * kTcTile;
  float acc[4][4] = {}, acc_o[4][4] = {};
  tc_cost_acc<PHI>(x, y, n, m, d, i0, j0, xs, ys, tid, acc, cf.p);
  if constexpr (!SAME) tc_sq_dist(xo, yo, n, m, d_o, i0, j0, xs, ys, tid, acc_o);
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = cost_psi(acc[a][b], cf.alpha, cf.beta);
  const int r0 = (tid >> 4) * 4, c0 = (tid & 15) * 4;
  float part = 0.f;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int64_t i = i0 + r0 + a;
    if (i >= n) continue;
    const float fi = f[i];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t j = j0 + c0 + b;
      if (j >= m) continue;
      const float p = expf((fi + g[j] - s * acc[a][b]) * inv_eps);
      const float c = SAME ? s * acc[a][b] : s_o * acc_o[a][b];
      if (p > 0.f) part = fmaf(p, c, part);
    }
  }
  red[tid] = static_cast<double>(part);
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) partial[static_cast<int64_t>(blockIdx.y) * gridDim.x + blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// Median of the unscaled cost matrix (geometry.py:106-109 median_cost_matrix,
// used by scale_cost="median", pointcloud.py:351-353: jnp.median over the
// n m entries, the mean of the two middle ones when n m is even) without
// forming the matrix: an exact radix select over the float bits of the costs.
// The costs are summed from the differences, so they are >= 0 and order like
// their bit patterns.  Three passes bin bits [31:21], [20:10] and [9:0] of the
// entries whose higher bits equal the prefix selected so far; after each, one
// workgroup picks the bin that holds the wanted rank.  For an even count a
// fourth pass takes the smallest entry above the lower middle one, unless
// the lower middle value also holds the next rank.  Any NaN entry makes the
// median NaN, as in jnp.median.
// ---------------------------------------------------------------------------
struct MedianState {
  unsigned long long rank;      // rank wanted among the entries matching the prefix
  unsigned long long total;     // n m
  unsigned long long eq_count;  // entries equal to the lower middle value
  uint32_t prefix;              // selected high bits
  uint32_t mask;                // which bits of prefix are selected
  uint32_t lo_bits;             // the lower middle entry
  uint32_t hi_bits;             // the smallest entry above it (fourth pass)
  uint32_t need_hi;             // the fourth pass is needed
  uint32_t nan_seen;
};
constexpr int kMedBins = 2048;

__global__ void median_init_kernel(MedianState* st, unsigned long long* hist, unsigned long long total) {
  for (int e = threadIdx.x; e < kMedBins; e += blockDim.x) hist[e] = 0ull;
  if (threadIdx.x == 0) {
    st->rank = (total - 1) / 2;
    st->total = total;
    st->eq_count = 0ull;
    st->prefix = 0u;
    st->mask = 0u;
    st->lo_bits = 0u;
    st->hi_bits = 0xffffffffu;
    st->need_hi = 0u;
    st->nan_seen = 0u;
  }
}

// One 64 x 64 block of costs per workgroup (tc_cost_acc); PASS 0-2 add the
// block's matching entries to the global histogram, PASS 3 takes the minimum
// above the lower middle entry.
template <int PASS, int PHI>
__global__ __launch_bounds__(256) void median_pass_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                          int64_t n, int64_t m, int d, MedianState* st,
                                                          unsigned long long* __restrict__ hist, CostForm cf) {
  __shared__ __align__(16) float xs[kTcChunk][kTcPad];
  __shared__ __align__(16) float ys[kTcChunk][kTcPad];
  __shared__ uint32_t h[PASS < 3 ? kMedBins : 1];
  __shared__ uint32_t s_flag;
  if (PASS == 3 && st->need_hi == 0u) return;  // the whole grid
  const int tid = threadIdx.x;
  constexpr int kShift = PASS == 0 ? 21 : (PASS == 1 ? 10 : 0);
  constexpr uint32_t kBinMask = PASS == 2 ? 1023u : 2047u;
  if (PASS < 3)
    for (int e = tid; e < kMedBins; e += 256) h[e] = 0u;
  if (tid == 0) s_flag = PASS < 3 ? 0u : 0xffffffffu;
  const int64_t i0 = static_cast<int64_t>(blockIdx.y) * kTcTile;
  const int64_t j0 = static_cast<int64_t>(blockIdx.x) * kTcTile;
  float acc[4][4] = {};
  tc_cost_acc<PHI>(x, y, n, m, d, i0, j0, xs, ys, tid, acc, cf.p);  // begins with a barrier
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = cost_psi(acc[a][b], cf.alpha, cf.beta);
  const uint32_t prefix = st->prefix, mask = st->mask, lo = st->lo_bits;
  const int r0 = (tid >> 4) * 4, c0 = (tid & 15) * 4;
  bool nan = false;
  uint32_t mn = 0xffffffffu;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      if (i0 + r0 + a >= n || j0 + c0 + b >= m) continue;
      const uint32_t u = __float_as_uint(acc[a][b]);
      if (PASS == 3) {
        if (u > lo && u < mn) mn = u;
      } else {
        if (PASS == 0 && acc[a][b] != acc[a][b]) nan = true;
        if ((u & mask) == prefix) atomicAdd(&h[(u >> kShift) & kBinMask], 1u);
      }
    }
  }
  if (PASS == 3) {
    if (mn != 0xffffffffu) atomicMin(&s_flag, mn);
    __syncthreads();
    if (tid == 0 && s_flag != 0xffffffffu) atomicMin(&st->hi_bits, s_flag);
    return;
  }
  if (nan) s_flag = 1u;
  __syncthreads();
  for (int e = tid; e <= static_cast<int>(kBinMask); e += 256)
    if (h[e]) atomicAdd(&hist[e], static_cast<unsigned long long>(h[e]));
  if (PASS == 0 && tid == 0 && s_flag) atomicOr(&st->nan_seen, 1u);
}

// One workgroup: the bin of the histogram that holds the wanted rank extends
// the prefix; the histogram is cleared for the next pass.
template <int PASS>
__global__ __launch_bounds__(1024) void median_select_kernel(MedianState* st, unsigned long long* hist) {
  constexpr int kShift = PASS == 0 ? 21 : (PASS == 1 ? 10 : 0);
  constexpr int kBins = PASS == 2 ? 1024 : 2048;
  constexpr int kPer = kBins / 1024;
  __shared__ unsigned long long s_sum[1024];
  const int t = threadIdx.x;
  unsigned long long c[kPer];
  unsigned long long own = 0ull;
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    c[q] = hist[t * kPer + q];
    hist[t * kPer + q] = 0ull;
    own += c[q];
  }
  s_sum[t] = own;
  __syncthreads();
  // inclusive scan of the per-thread sums
  for (int off = 1; off < 1024; off <<= 1) {
    const unsigned long long v = t >= off ? s_sum[t - off] : 0ull;
    __syncthreads();
    s_sum[t] += v;
    __syncthreads();
  }
  const unsigned long long rank = st->rank;
  unsigned long long before = s_sum[t] - own;
  __syncthreads();
  if (rank >= before && rank < before + own) {
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      if (rank < before + c[q]) {
        const uint32_t bin = static_cast<uint32_t>(t * kPer + q);
        st->prefix |= bin << kShift;
        st->mask |= static_cast<uint32_t>(kBins - 1) << kShift;
        st->rank = rank - before;
        if (PASS == 2) {
          st->lo_bits = st->prefix;
          st->eq_count = c[q];
          // an even count also wants the next rank: the same value if it
          // repeats beyond this rank, else the smallest larger entry
          st->need_hi = (st->total % 2ull == 0ull && rank - before + 1ull >= c[q]) ? 1u : 0u;
        }
        break;
      }
      before += c[q];
    }
  }
}

// Sum, sum of squares and maximum of the unscaled costs of 64 x 64 pair
// blocks (the mean and std of the cost matrix, geometry.py:112-136, and
// scale_cost="mean" / "max_cost", pointcloud.py:342-383, for costs without a
// closed form): float64 partials per block, the maximum through its float
// bits (costs are >= 0).
template <int PHI>
__global__ __launch_bounds__(256) void cost_summary_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                           int64_t n, int64_t m, int d, CostForm cf,
                                                           double* __restrict__ partial, int64_t blocks,
                                                           uint32_t* __restrict__ max_bits) {
  __shared__ __align__(16) float xs[kTcChunk][kTcPad];
  __shared__ __align__(16) float ys[kTcChunk][kTcPad];
  __shared__ double red[2][256];
  __shared__ uint32_t s_max;
  const int tid = threadIdx.x;
  const int64_t i0 = static_cast<int64_t>(blockIdx.y) * kTcTile;
  const int64_t j0 = static_cast<int64_t>(blockIdx.x) * kTcTile;
  if (tid == 0) s_max = 0u;
  float acc[4][4] = {};
  tc_cost_acc<PHI>(x, y, n, m, d, i0, j0, xs, ys, tid, acc, cf.p);  // begins with a barrier
  const int r0 = (tid >> 4) * 4, c0 = (tid & 15) * 4;
  double s1 = 0.0, s2 = 0.0;
  uint32_t mx = 0u;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      if (i0 + r0 + a >= n || j0 + c0 + b >= m) continue;
      const float c = cost_psi(acc[a][b], cf.alpha, cf.beta);
      s1 += c;
      s2 += static_cast<double>(c) * c;
      const uint32_t u = __float_as_uint(c);
      mx = u > mx ? u : mx;
    }
  red[0][tid] = s1;
  red[1][tid] = s2;
  atomicMax(&s_max, mx);
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) {
      red[0][tid] += red[0][tid + w];
      red[1][tid] += red[1][tid + w];
    }
    __syncthreads();
  }
  if (tid == 0) {
    const int64_t bid = static_cast<int64_t>(blockIdx.y) * gridDim.x + blockIdx.x;
    partial[bid] = red[0][0];
    partial[blocks + bid] = red[1][0];
    atomicMax(max_bits, s_max);
  }
}

// Sum of the partials in a fixed order (one workgroup of 1024 threads).
__global__ __launch_bounds__(1024) void sum_partials_kernel(const double* __restrict__ partial, int64_t count,
                                                            double* __restrict__ out) {
  __shared__ double red[1024];
  const int tid = threadIdx.x;
  double v = 0.0;
  for (int64_t k = tid; k < count; k += 1024) v += partial[k];
  red[tid] = v;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) out[0] = red[0];
}

// Danskin gradient of reg_ot_cost with respect to one cloud (sinkhorn.py:
// 329-335 with compute_kl_reg_cost, :224-277): with f and g held fixed only
// the coupling's total mass depends on the points, and
//   d reg / d p_i = 2 s sum_q P_iq (p_i - q_q)
// for the cloud p against the other cloud q.  A workgroup owns 64 points p
// and streams a slice of the q points 64 at a time: the 64 x 64 block of P
// goes through LDS (pair terms from the differences, as above), then each
// thread accumulates sum_q P_iq (p_ik - q_qk) for 4 points by 4 features.
// Partials per slice, summed in float64 by grad_reduce_kernel.  d <= 64.
constexpr int kGdMaxD = 64;
constexpr int kGdRow = kGdMaxD + 4;

__global__ __launch_bounds__(256) void grad_partial_kernel(
    const float* __restrict__ p, const float* __restrict__ pot_p, int64_t np, const float* __restrict__ q,
    const float* __restrict__ pot_q, int64_t nq, int d, float s, float inv_eps, int64_t slice,
    float* __restrict__ partial) {
  __shared__ __align__(16) float ps[kGdMaxD][kTcPad];  // own points, feature-major
  __shared__ __align__(16) float qs[kGdMaxD][kTcPad];  // streamed points, feature-major
  __shared__ __align__(16) float qr[kTcTile][kGdRow];  // streamed points, point-major
  __shared__ __align__(16) float pt[kTcTile][kTcPad];  // P block, [q][p]
  const int tid = threadIdx.x;
  const int64_t i0 = static_cast<int64_t>(blockIdx.x) * kTcTile;
  const int64_t q_begin = static_cast<int64_t>(blockIdx.y) * slice;
  const int64_t q_end = min(nq, q_begin + slice);
  for (int e = tid; e < kTcTile * kGdMaxD; e += 256) {
    const int r = e / kGdMaxD, k = e % kGdMaxD;
    const int64_t i = i0 + r;
    ps[k][r] = (k < d && i < np) ? p[i * d + k] : 0.f;
  }
  const int r0 = (tid >> 4) * 4, c0 = (tid & 15) * 4;
  float fi[4];
  float own[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int64_t i = i0 + r0 + a;
    fi[a] = i < np ? pot_p[i] : -INFINITY;
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) own[a][b] = (c0 + b < d) ? ps[c0 + b][r0 + a] : 0.f;
  float acc[4][4] = {};
  for (int64_t j0 = q_begin; j0 < q_end; j0 += kTcTile) {
    __syncthreads();
    for (int e = tid; e < kTcTile * kGdMaxD; e += 256) {
      const int r = e / kGdMaxD, k = e % kGdMaxD;
      const int64_t j = j0 + r;
      const float v = (k < d && j < q_end) ? q[j * d + k] : 0.f;
      qs[k][r] = v;
      qr[r][k] = v;
    }
    __syncthreads();
    // P block: rows r0..r0+3 (own points) x columns c0..c0+3 (streamed points)
    float dist[4][4] = {};
    for (int k = 0; k < d; ++k) {
      const float4 xv = *reinterpret_cast<const float4*>(&ps[k][r0]);
      const float4 yv = *reinterpret_cast<const float4*>(&qs[k][c0]);
      const float xa[4] = {xv.x, xv.y, xv.z, xv.w};
      const float ya[4] = {yv.x, yv.y, yv.z, yv.w};
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const float u = xa[a] - ya[b];
          dist[a][b] = fmaf(u, u, dist[a][b]);
        }
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t j = j0 + c0 + b;
      const float gj = j < q_end ? pot_q[j] : -INFINITY;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const float e = (fi[a] + gj - s * dist[a][b]) * inv_eps;
        pt[c0 + b][r0 + a] = (e == e && e > -INFINITY) ? expf(e) : 0.f;
      }
    }
    __syncthreads();
    // acc[a][b] += sum_q P(r0 + a, q) (own[a][b] - q_(c0 + b)): rows r0.., features c0..
    const int jn = static_cast<int>(min<int64_t>(kTcTile, q_end - j0));
    for (int jj = 0; jj < jn; ++jj) {
      const float4 pv = *reinterpret_cast<const float4*>(&pt[jj][r0]);
      const float4 qv = *reinterpret_cast<const float4*>(&qr[jj][c0]);
      const float pa[4] = {pv.x, pv.y, pv.z, pv.w};
      const float qa[4] = {qv.x, qv.y, qv.z, qv.w};
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = fmaf(pa[a], own[a][b] - qa[b], acc[a][b]);
    }
  }
  // partial[slice][i][k] for the thread's 4 points x 4 features
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int64_t i = i0 + r0 + a;
    if (i >= np) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b)
      if (c0 + b < d) partial[(static_cast<int64_t>(blockIdx.y) * np + i) * d + c0 + b] = acc[a][b];
  }
}

// grad[i][k] = 2 s sum over slices of partial[slice][i][k], summed in float64.
__global__ void grad_reduce_kernel(const float* __restrict__ partial, int64_t count, int slices, double two_s,
                                   float* __restrict__ grad) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= count) return;
  double v = 0.0;
  for (int t = 0; t < slices; ++t) v += static_cast<double>(partial[static_cast<int64_t>(t) * count + e]);
  grad[e] = static_cast<float>(two_s * v);
}

// f <- g for the parallel updates of a symmetric problem (f = g throughout);
// launches after the loop has stopped return at once.
__global__ void copy_potential_kernel(const float* __restrict__ src, float* __restrict__ dst, int64_t n,
                                      const SolveState* __restrict__ st) {
  if (st->done) return;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// Adaptive momentum, the iteration ``start`` that a block-ending g-update
// computes when inner divides start: that launch measures the error its
// weight depends on, so it writes the plain update (and g_save the previous
// g); once the block has been finalized, this kernel blends them with the
// weight (acceleration.py:161-172).  Every other launch exits at once.
__global__ void momentum_fixup_kernel(float* __restrict__ g, const float* __restrict__ g_save, int64_t m,
                                      const SolveState* __restrict__ st, const float* __restrict__ errors,
                                      int start, int inner, float value, float threshold) {
  if (st->done || st->block * inner != start) return;
  const float w = adaptive_momentum(errors, start, inner, value, threshold);
  if (w == 1.f) return;
  const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < m) {
    const float h_old = g_save[j];
    g[j] = (1.f - w) * (isfinite(h_old) ? h_old : 0.f) + w * g[j];
  }
}

// Outputs of a stopped loop, computed on the device right after the block
// that stopped it so the host only has to read pinned memory: the float64
// dual terms (as dual_partial_kernel + dual_final_kernel), g <- g_save, and
// the SolveMail written last.  Grid (kDualBlocks, 3); every launch before the
// loop stops, or after its outputs were written, exits at once.  The last
// workgroup to arrive combines the partials.
__global__ void solve_finalize_kernel(const float* f, const float* __restrict__ a, int64_t n,
                                      const float* __restrict__ g_save, float* __restrict__ g_out,
                                      const float* __restrict__ b, int64_t m, const SolveScalars* __restrict__ sc,
                                      double tau_a, double tau_b, const float* __restrict__ errors, int outer,
                                      double* __restrict__ part, SolveState* st, SolveMail* mail, uint32_t seq,
                                      int recenter, float* __restrict__ f_out) {
  __shared__ double red[2][256];
  __shared__ double cen[2][kDualBlocks];
  __shared__ int s_last;
  if (!st->done || st->finalized) return;
  const int tid = threadIdx.x;
  const int side = blockIdx.y;
  double s0 = 0.0, s1 = 0.0;
  double c0 = 0.0, c1 = 0.0;  // recentering: sum and count of the finite f
  if (side < 2) {
    const float* pot = side == 0 ? f : g_save;
    const float* w = side == 0 ? a : b;
    const int64_t len = side == 0 ? n : m;
    const double tau = side == 0 ? tau_a : tau_b;
    const double eps = sc->eps;
    const double rho = eps * tau / (1.0 - tau);
    const int64_t per = (len + kDualBlocks - 1) / kDualBlocks;
    const int64_t i0 = blockIdx.x * per, i1 = min(len, i0 + per);
    for (int64_t i = i0 + tid; i < i1; i += blockDim.x) {
      const double wi = w[i];
      if (side == 0 && recenter && isfinite(pot[i])) {
        c0 += pot[i];
        c1 += 1.0;
      }
      s1 += wi;
      if (wi > 0.0) {
        const double h = static_cast<double>(pot[i]) - eps * log(wi);
        if (tau == 1.0) {
          s0 += wi * h;
        } else {
          s0 -= wi * rho * (exp(-h / rho) - 1.0);
        }
      }
      if (side == 1) g_out[i] = g_save[i];
      if (side == 0 && f_out != f) f_out[i] = pot[i];
    }
  } else if (blockIdx.x == 0) {
    for (int i = tid; i < outer; i += blockDim.x) {
      const float e = errors[i];
      if (e != e) s0 += 1.0;
      if (e != -1.f) s1 += 1.0;
    }
  }
  red[0][tid] = s0;
  red[1][tid] = s1;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (tid < k) {
      red[0][tid] += red[0][tid + k];
      red[1][tid] += red[1][tid + k];
    }
    __syncthreads();
  }
  if (tid == 0) {
    store_f64_through(part + (side * kDualBlocks + blockIdx.x) * 2 + 0, red[0][0]);
    store_f64_through(part + (side * kDualBlocks + blockIdx.x) * 2 + 1, red[1][0]);
  }
  if (recenter && side == 0) {
    __syncthreads();
    red[0][tid] = c0;
    red[1][tid] = c1;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
      if (tid < k) {
        red[0][tid] += red[0][tid + k];
        red[1][tid] += red[1][tid + k];
      }
      __syncthreads();
    }
    if (tid == 0) {
      store_f64_through(part + (3 * kDualBlocks + blockIdx.x) * 2 + 0, red[0][0]);
      store_f64_through(part + (3 * kDualBlocks + blockIdx.x) * 2 + 1, red[1][0]);
    }
  }
  if (tid == 0) {
    wait_own_stores();  // this thread stored every partial of the workgroup
    const uint32_t prev = atomicAdd(&st->fin_count, 1u);
    s_last = (prev == 3u * kDualBlocks - 1u) ? 1 : 0;
  }
  __syncthreads();
  if (!s_last) return;
  // the last workgroup: all partials loaded at once (one per thread), then a
  // fixed-order tree per (side, term)
  static_assert(3 * kDualBlocks * 2 <= 256, "one partial per thread");
  const int q = tid % kDualBlocks, sd = (tid / kDualBlocks) % 3, term = tid / (3 * kDualBlocks);
  red[0][tid] = (tid < 3 * kDualBlocks * 2) ? load_f64_coherent(part + (sd * kDualBlocks + q) * 2 + term) : 0.0;
  __syncthreads();
  if (recenter && tid < kDualBlocks) {
    cen[0][tid] = load_f64_coherent(part + (3 * kDualBlocks + tid) * 2 + 0);
    cen[1][tid] = load_f64_coherent(part + (3 * kDualBlocks + tid) * 2 + 1);
  }
  __syncthreads();
  for (int k = kDualBlocks / 2; k > 0; k >>= 1) {
    if (tid < 3 * kDualBlocks * 2 && q < k) red[0][tid] += red[0][tid + k];
    if (recenter && tid < k) {
      cen[0][tid] += cen[0][tid + k];
      cen[1][tid] += cen[1][tid + k];
    }
    __syncthreads();
  }
  // SinkhornState.recenter, balanced branch (sinkhorn.py:98-102): f - c and
  // g + c with c the mean of the finite f; the dual terms follow the shift.
  const double shift = (recenter && cen[1][0] > 0.0) ? cen[0][0] / cen[1][0] : 0.0;
  if (recenter) {
    const float cf = static_cast<float>(shift);
    for (int64_t i = tid; i < n; i += blockDim.x) f_out[i] -= cf;
    for (int64_t j = tid; j < m; j += blockDim.x) g_out[j] = g_save[j] + cf;
  }
  if (tid != 0) return;
  double du[8];
  for (int sd2 = 0; sd2 < 3; ++sd2) {
    du[sd2 * 2 + 0] = red[0][(0 * 3 + sd2) * kDualBlocks];
    du[sd2 * 2 + 1] = red[0][(1 * 3 + sd2) * kDualBlocks];
  }
  // sum_a a (f - c - eps log a) = div_a - c sum a; likewise + c sum b for g
  du[0] -= shift * du[1];
  du[2] += shift * du[3];
  du[6] = outer > 0 ? static_cast<double>(errors[outer - 1]) : 0.0;
  du[7] = sc->eps;
  for (int q = 0; q < 8; ++q) mail->du[q] = du[q];
  mail->total_mass = st->total_mass;
  mail->n_blocks = st->n_blocks;
  st->finalized = 1;
  st->fin_count = 0u;
  __threadfence_system();
  mail->ready = seq;
}

// ---------------------------------------------------------------------------
// Recentred unbalanced iterations (both taus < 1 with recenter_potentials):
// Sinkhorn.lse_step's translation-invariant terms (sinkhorn.py:803-857) and
// the recentring of SinkhornState (sinkhorn.py:76-114).  The half-steps write
// their log-sum-exp terms (kModeWriteApp); one workgroup per side then forms
//   t = tau (eps log w - app) - k smin(other),  new = t + xi smin(t),
// blends it with momentum and records log sum w exp(-pot / rho) of the new
// potential, from which the next update's smin and the error's shift follow.
// smin(p, w) = -rho log sum_i w_i exp(-p_i / rho) (sinkhorn.py:818-822);
// entries with w_i = 0 are left out, as jax.scipy.special.logsumexp with b
// does.  These passes are O(n) against the O(n m d) half-steps.
// ---------------------------------------------------------------------------
constexpr int kRcThreads = 1024;

// Running log-sum-exp (max, sum) in float64.
__device__ __forceinline__ void lse_push(double& mx, double& sm, double z) {
  if (!(z > -INFINITY)) {
    if (z != z) mx = sm = NAN;
    return;
  }
  if (z > mx) {
    sm = sm * exp(mx - z) + 1.0;
    mx = z;
  } else {
    sm += exp(z - mx);
  }
}

// Workgroup combine of the threads' (max, sum); the result in every thread.
__device__ double block_lse(double mx, double sm, double (*s_red)[kRcThreads]) {
  const int tid = threadIdx.x;
  __syncthreads();
  s_red[0][tid] = mx;
  s_red[1][tid] = sm;
  __syncthreads();
  for (int k = kRcThreads / 2; k > 0; k >>= 1) {
    if (tid < k) {
      const double m1 = s_red[0][tid], s1 = s_red[1][tid];
      const double m2 = s_red[0][tid + k], s2 = s_red[1][tid + k];
      if (m1 != m1 || m2 != m2 || s1 != s1 || s2 != s2) {
        s_red[0][tid] = s_red[1][tid] = NAN;
      } else {
        const double m = fmax(m1, m2);
        s_red[0][tid] = m;
        s_red[1][tid] = (m == -INFINITY) ? 0.0 : s1 * exp(m1 - m) + s2 * exp(m2 - m);
      }
    }
    __syncthreads();
  }
  const double m = s_red[0][0], s = s_red[1][0];
  return (s > 0.0) ? m + log(s) : (s != s ? NAN : -INFINITY);
}

// log sum_i w_i exp(-pot_i / rho) of the starting f.
__device__ __forceinline__ double rc_rho(const SolveScalars* sc, double tau) { return sc->eps * tau / (1.0 - tau); }

__global__ __launch_bounds__(kRcThreads) void rc_lse_kernel(const float* __restrict__ pot,
                                                            const float* __restrict__ w, int64_t n, double tau,
                                                            const SolveScalars* __restrict__ sc, double* out) {
  __shared__ double s_red[2][kRcThreads];
  const double rho = rc_rho(sc, tau);
  double mx = -INFINITY, sm = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kRcThreads) {
    const double wi = w[i];
    if (wi > 0.0) lse_push(mx, sm, log(wi) - static_cast<double>(pot[i]) / rho);
  }
  const double r = block_lse(mx, sm, s_red);
  if (threadIdx.x == 0) *out = r;
}

struct RcFinishArgs {
  const float* app;   // eps LSE terms of the side's half-step
  float* pot;         // old potential in, new potential out
  const float* w;     // the side's weights
  int64_t n;
  double tau, tau_other;
  double kc, xc;      // k(tau_i, tau_i) and xi(tau_i, tau_j) of the side
  const SolveScalars* sc;
  const double* lse_other;  // log sum w exp(-pot / rho) of the other side, for its smin
  double* lse_self;         // written: the same sum for this side's new potential
  float momentum;
  int32_t mom_start, inner, it_offset;
  float mom_threshold;
  const float* errors;
  const SolveState* st;
};

__global__ __launch_bounds__(kRcThreads) void rc_finish_kernel(const RcFinishArgs p) {
  __shared__ double s_red[2][kRcThreads];
  if (p.st->done) return;
  float wm = p.momentum;
  if (p.mom_start > 0 && p.st->block * p.inner + p.it_offset >= p.mom_start)
    wm = adaptive_momentum(p.errors, p.mom_start, p.inner, p.momentum, p.mom_threshold);
  const double eps = p.sc->eps, rho_self = rc_rho(p.sc, p.tau);
  const double smin_other = -rc_rho(p.sc, p.tau_other) * *p.lse_other;
  // t_i as the plain update forms it (the kernel epilogue's order), minus k smin
  auto t_of = [&](int64_t i) -> double {
    const double wi = p.w[i];
    if (!(wi > 0.0)) return -INFINITY;
    const float a = p.app[i];
    const float upd = static_cast<float>(eps) * logf(static_cast<float>(wi)) - (isfinite(a) ? a : 0.f);
    return static_cast<double>(upd) * p.tau - p.kc * smin_other;
  };
  double mx = -INFINITY, sm = 0.0;
  for (int64_t i = threadIdx.x; i < p.n; i += kRcThreads) {
    const double wi = p.w[i];
    if (wi > 0.0) lse_push(mx, sm, log(wi) - t_of(i) / rho_self);
  }
  const double smin_t = -rho_self * block_lse(mx, sm, s_red);
  mx = -INFINITY;
  sm = 0.0;
  for (int64_t i = threadIdx.x; i < p.n; i += kRcThreads) {
    const double nw = t_of(i) + p.xc * smin_t;
    const float old = p.pot[i];
    float v = static_cast<float>(nw);
    if (wm != 1.f) v = (1.f - wm) * (isfinite(old) ? old : 0.f) + wm * v;
    p.pot[i] = v;
    const double wi = p.w[i];
    if (wi > 0.0) lse_push(mx, sm, log(wi) - static_cast<double>(v) / rho_self);
  }
  const double r = block_lse(mx, sm, s_red);
  if (threadIdx.x == 0) *p.lse_self = r;
}

// Errors of the state (f, g) recentred (sinkhorn.py:57-68, 160-221): the
// coupling, and so both marginals, do not change under the translation; the
// targets a exp(-(f + s) / rho_a) and b exp(-(g - s) / rho_b) do.  Column
// marginals come from a marginal half-step, row marginals from the f-update's
// log-sum-exp terms.  Thread 0 then runs the block finalizer.
__global__ __launch_bounds__(kRcThreads) void rc_error_kernel(const HalfStepArgs fin, const float* __restrict__ colmarg,
                                                              const float* __restrict__ app_f,
                                                              const float* __restrict__ f, const float* __restrict__ g,
                                                              const float* __restrict__ a, const float* __restrict__ b,
                                                              int64_t n, int64_t m, double tau_a, double tau_b,
                                                              const SolveScalars* __restrict__ sc,
                                                              const RcScalars* __restrict__ rc) {
  __shared__ double s_red[2][kRcThreads];
  __shared__ double s_m[kRcThreads];
  const SolveState* st = fin.state;
  if (st->done) return;
  const double eps = sc->eps, rho_a = rc_rho(sc, tau_a), rho_b = rc_rho(sc, tau_b);
  const int tid = threadIdx.x;
  const double shift = rho_a * rho_b / (rho_a + rho_b) * (rc->lse[0] - rc->lse[1]);
  const double pn = static_cast<double>(fin.norm_p);
  double col = 0.0, row = 0.0, mass = 0.0;
  for (int64_t j = tid; j < m; j += kRcThreads) {
    const float marg = colmarg[j];
    const float w = b[j];
    const double target = (w > 0.f) ? static_cast<double>(w) * exp(-(static_cast<double>(g[j]) - shift) / rho_b) : 0.0;
    const double diff = fabs(static_cast<double>(marg) - target);
    col += (fin.norm_p == 1) ? diff : pow(diff, pn);
    mass += static_cast<double>(marg);
  }
  for (int64_t i = tid; i < n; i += kRcThreads) {
    const float marg = expf((f[i] + app_f[i]) / static_cast<float>(eps));
    const float w = a[i];
    const double target = (w > 0.f) ? static_cast<double>(w) * exp(-(static_cast<double>(f[i]) + shift) / rho_a) : 0.0;
    const double diff = fabs(static_cast<double>(marg) - target);
    row += (fin.norm_p == 1) ? diff : pow(diff, pn);
  }
  s_red[0][tid] = col;
  s_red[1][tid] = row;
  s_m[tid] = mass;
  __syncthreads();
  for (int k = kRcThreads / 2; k > 0; k >>= 1) {
    if (tid < k) {
      s_red[0][tid] += s_red[0][tid + k];
      s_red[1][tid] += s_red[1][tid + k];
      s_m[tid] += s_m[tid + k];
    }
    __syncthreads();
  }
  if (tid == 0) finalize_block(fin, s_red[0][0], s_red[1][0], s_m[0]);
}

// Outputs of the stopped loop, recentred (sinkhorn.py:1008-1009): f + s in
// place and g - s into g_out, before solve_finalize_kernel reads them.
__global__ __launch_bounds__(kRcThreads) void rc_output_kernel(float* f, const float* __restrict__ g,
                                                               float* __restrict__ g_out, int64_t n, int64_t m,
                                                               double tau_a, double tau_b,
                                                               const SolveScalars* __restrict__ sc,
                                                               const RcScalars* __restrict__ rc, const SolveState* st) {
  if (!st->done || st->finalized) return;
  const double rho_a = rc_rho(sc, tau_a), rho_b = rc_rho(sc, tau_b);
  const float s = static_cast<float>(rho_a * rho_b / (rho_a + rho_b) * (rc->lse[0] - rc->lse[1]));
  for (int64_t i = threadIdx.x; i < n; i += kRcThreads) f[i] += s;
  for (int64_t j = threadIdx.x; j < m; j += kRcThreads) g_out[j] = g[j] - s;
}

// ---------------------------------------------------------------------------
// Kernel selection
// ---------------------------------------------------------------------------
typedef void (*HalfStepFn)(const HalfStepArgs);
typedef void (*BatchedFn)(const HalfStepArgs*, const int4*, const uint32_t*, uint32_t);

struct KernelCfg {
  int prec;
  int kk;       // bf16: K blocks of 32; fp32: k-steps of 4
  int nb, mt, w;
  HalfStepFn fn;
  BatchedFn bfn;
  int cols() const { return 16 * nb * w; }
  // 1 KiB chunks per 16 rows (none for the cost stream, which reads the points)
  int frag_chunks() const { return prec == 1 ? kk : (prec == 0 ? (kk + 3) / 4 : 0); }
};

template <int PREC, int KK, int NB, int MT, int W>
KernelCfg make_cfg() {
  return KernelCfg{PREC, KK, NB, MT, W, &halfstep_kernel<PREC, KK, NB, MT, W>,
                   &halfstep_batched_kernel<PREC, KK, NB, MT, W>};
}

// Tile variants of the fused half-step per (precision, K): the search space
// of the autotuner.  Index 0 is the default.
int kernel_variants(int prec, int d, KernelCfg* out, int max_out) {
  int n = 0;
  auto add = [&](const KernelCfg& k) {
    if (n < max_out) out[n] = k;
    ++n;
  };
  if (d < 1 || d > 256) return 0;
  // costs other than the squared Euclidean: the vector-ALU stream, 64 columns
  // by 64 rows per step (kernel_key 10 + phi)
  switch (prec) {
    case 11: add(make_cfg<2, 1, 1, 4, 4>()); return n;
    case 12: add(make_cfg<2, 2, 1, 4, 4>()); return n;
    case 13: add(make_cfg<2, 3, 1, 4, 4>()); return n;
    default: break;
  }
  if (prec == OTTX_PREC_BF16_SPLIT) {
    const int kb = (d + 31) / 32;
    switch (kb) {
      case 1:
        add(make_cfg<1, 1, 4, 4, 4>()); add(make_cfg<1, 1, 4, 2, 4>());
        add(make_cfg<1, 1, 2, 4, 4>()); add(make_cfg<1, 1, 8, 2, 4>());
        break;
      case 2:
        add(make_cfg<1, 2, 4, 4, 4>()); add(make_cfg<1, 2, 4, 2, 4>());
        add(make_cfg<1, 2, 2, 4, 4>()); add(make_cfg<1, 2, 8, 2, 4>());
        add(make_cfg<1, 2, 4, 2, 8>()); add(make_cfg<1, 2, 2, 4, 8>()); add(make_cfg<1, 2, 4, 4, 8>());
        break;
      case 3: add(make_cfg<1, 3, 2, 4, 4>()); add(make_cfg<1, 3, 2, 2, 4>()); add(make_cfg<1, 3, 1, 4, 4>()); break;
      case 4: add(make_cfg<1, 4, 2, 4, 4>()); add(make_cfg<1, 4, 2, 2, 4>()); add(make_cfg<1, 4, 1, 4, 4>()); break;
      case 5: add(make_cfg<1, 5, 1, 2, 4>()); break;
      case 6: add(make_cfg<1, 6, 1, 2, 4>()); break;
      case 7: add(make_cfg<1, 7, 1, 2, 4>()); break;
      case 8: add(make_cfg<1, 8, 1, 2, 4>()); break;
      default: return 0;
    }
    return n;
  }
  if (prec == OTTX_PREC_FP32) {
    const int ks = (d + 3) / 4;
    if (ks <= 1) {
      add(make_cfg<0, 1, 4, 4, 4>()); add(make_cfg<0, 1, 4, 2, 4>()); add(make_cfg<0, 1, 8, 2, 4>());
    } else if (ks <= 2) {
      add(make_cfg<0, 2, 4, 4, 4>()); add(make_cfg<0, 2, 4, 2, 4>()); add(make_cfg<0, 2, 8, 2, 4>());
    } else if (ks <= 4) {
      add(make_cfg<0, 4, 4, 4, 4>()); add(make_cfg<0, 4, 4, 2, 4>()); add(make_cfg<0, 4, 2, 4, 4>());
    } else if (ks <= 8) {
      add(make_cfg<0, 8, 4, 4, 4>()); add(make_cfg<0, 8, 4, 2, 4>()); add(make_cfg<0, 8, 2, 4, 4>());
    } else if (ks <= 12) {
      add(make_cfg<0, 12, 4, 4, 4>()); add(make_cfg<0, 12, 4, 2, 4>()); add(make_cfg<0, 12, 2, 4, 4>());
    } else if (ks <= 16) {
      add(make_cfg<0, 16, 4, 4, 4>()); add(make_cfg<0, 16, 4, 2, 4>()); add(make_cfg<0, 16, 2, 4, 4>());
      add(make_cfg<0, 16, 2, 4, 8>()); add(make_cfg<0, 16, 4, 2, 8>());
    } else if (ks <= 24) {
      add(make_cfg<0, 24, 2, 4, 4>()); add(make_cfg<0, 24, 2, 2, 4>()); add(make_cfg<0, 24, 1, 4, 4>());
    } else if (ks <= 32) {
      add(make_cfg<0, 32, 2, 4, 4>()); add(make_cfg<0, 32, 2, 2, 4>()); add(make_cfg<0, 32, 1, 4, 4>());
    } else if (ks <= 48) {
      add(make_cfg<0, 48, 1, 2, 4>());
    } else if (ks <= 64) {
      add(make_cfg<0, 64, 1, 2, 4>());
    } else {
      return 0;
    }
    return n;
  }
  return 0;
}

// Occupancy targets (workgroups per CU the row splits aim for).
constexpr int kTargets[] = {4, 2, 8, 1, 16, 3, 6};
constexpr int kNumTargets = 7;

// tile_config: -1 = default, else variant | (target index << 8).
bool select_kernel(int prec, int d, int tile_config, KernelCfg* out, int* target_mult) {
  KernelCfg v[8];
  const int n = kernel_variants(prec, d, v, 8);
  if (n <= 0) return false;
  int vi = 0, ti = 0;
  if (tile_config >= 0) {
    vi = tile_config & 0xff;
    ti = (tile_config >> 8) & 0xff;
    if (vi >= n || ti >= kNumTargets) return false;
  }
  *out = v[vi];
  *target_mult = kTargets[ti];
  return true;
}

int device_cus() {
  static int cached[64] = {0};  // per device; a racing first query writes the same value
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (dev >= 0 && dev < 64 && cached[dev] > 0) return cached[dev];
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) return 256;
  if (dev >= 0 && dev < 64) cached[dev] = v;
  return v;
}

int64_t pad_rows(int64_t n, int64_t pad = kRowPad) { return ((n + pad - 1) / pad) * pad; }

struct PassPlan {
  int tiles = 0;
  int splits = 1;
  int rb_per_split = 0;
  int n_in_rb = 0;
};

PassPlan plan_pass(const KernelCfg& k, int64_t n_in, int64_t n_out, int cus, int target_mult,
                   int64_t pad = kRowPad) {
  PassPlan p;
  const int64_t in_pad = pad_rows(n_in, pad), out_pad = pad_rows(n_out, pad);
  p.tiles = static_cast<int>(out_pad / k.cols());
  p.n_in_rb = static_cast<int>(in_pad / 16);
  const int in_tiles = p.n_in_rb / k.mt;
  const int target = target_mult * cus;
  int splits = (target + p.tiles - 1) / p.tiles;
  splits = std::max(1, std::min(splits, in_tiles / 2));
  // an even number of streamed tiles per split (the stream loop computes two
  // per trip); in_tiles is even since the row padding is a multiple of 32 MT
  int tiles_per_split = (in_tiles + splits - 1) / splits;
  tiles_per_split += tiles_per_split & 1;
  p.rb_per_split = tiles_per_split * k.mt;
  p.splits = (p.n_in_rb + p.rb_per_split - 1) / p.rb_per_split;
  return p;
}

size_t align_up(size_t v) { return (v + 255) & ~static_cast<size_t>(255); }

// The stream loop computes two tiles of MT row blocks per trip: every split
// (and so the padded row count) must hold an even number of tiles.
bool even_tiles(const KernelCfg& k, const PassPlan& p) {
  return p.rb_per_split > 0 && p.rb_per_split % (2 * k.mt) == 0 && p.n_in_rb % (2 * k.mt) == 0;
}

// Partial buffers of the std(C) moments (offsets relative to their region).
struct MomentLayout {
  int vchunks[2], gchunks[2];
  int64_t vrows[2], grows[2];
  int nkb, tiles;
  size_t vpart[2], gpart[2], total;
};

MomentLayout moment_layout(int64_t n, int64_t m, int d) {
  MomentLayout M{};
  moment_plan(n, d, &M.vchunks[0], &M.vrows[0], &M.gchunks[0], &M.grows[0]);
  moment_plan(m, d, &M.vchunks[1], &M.vrows[1], &M.gchunks[1], &M.grows[1]);
  M.nkb = (d + kMomGramTile - 1) / kMomGramTile;
  M.tiles = M.nkb * (M.nkb + 1) / 2;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off = (off + bytes + 255) & ~static_cast<size_t>(255);
    return at;
  };
  for (int side = 0; side < 2; ++side)
    M.vpart[side] = take(static_cast<size_t>(M.vchunks[side]) * (2 * d + 2) * sizeof(double));
  for (int side = 0; side < 2; ++side)
    M.gpart[side] = take(static_cast<size_t>(M.gchunks[side]) * M.tiles * kMomGramTile * kMomGramTile *
                         sizeof(double));
  M.total = off;
  return M;
}

// Moments of both clouds and std(C) (two launches after the statistics, which
// wrote mu): out[1] = std(C); with sc and rel > 0 the solve's scalars follow it.
int launch_std(const float* x, int64_t n, const float* y, int64_t m, int d, const float* mu, double s,
               char* region, const MomentLayout& M, double* out, SolveScalars* sc, double rel, double tau_a,
               double tau_b, hipStream_t st) {
  MomentArgs a{};
  a.pts[0] = x;
  a.pts[1] = y;
  a.np[0] = n;
  a.np[1] = m;
  a.d = d;
  a.nkb = M.nkb;
  a.tiles = M.tiles;
  a.mu = mu;
  int cur = 0;
  for (int side = 0; side < 2; ++side) {
    a.vchunks[side] = M.vchunks[side];
    a.vrows[side] = M.vrows[side];
    a.vpart[side] = reinterpret_cast<double*>(region + M.vpart[side]);
    a.gchunks[side] = M.gchunks[side];
    a.grows[side] = M.grows[side];
    a.gpart[side] = reinterpret_cast<double*>(region + M.gpart[side]);
    a.seg[side] = cur;
    cur += M.vchunks[side];
  }
  for (int side = 0; side < 2; ++side) {
    a.seg[2 + side] = cur;
    cur += M.gchunks[side] * M.tiles;
  }
  a.seg[4] = cur;
  hipLaunchKernelGGL(moments_kernel, dim3(cur), dim3(256), 0, st, a);
  hipLaunchKernelGGL(std_final_kernel, dim3(1), dim3(1024), 0, st, a, s, out, sc, rel, tau_a, tau_b);
  return hipGetLastError() == hipSuccess ? OTTX_OK : OTTX_EHIP;
}

struct Layout {
  size_t frag_x, frag_y, norm_x, norm_y, g_save, f_save, lse_x, lse_y, partial, counters, err_col, mass, err_row, state,
      dual, dual_part, scal, stats_x, stats_y, stats_out, mu, w_x, w_y, moments, rc_app, rc_col, rc, total;
  int chunks_x, chunks_y;
  int64_t rows_chunk_x, rows_chunk_y;
  MomentLayout mom;
};

// A relative epsilon on std(C) needs the moment partials.
inline bool wants_std(const ottx_problem* pb) {
  return pb->relative_epsilon > 0.0 && pb->epsilon_statistic == OTTX_EPS_STAT_STD;
}

// At most 128 chunks: the single-workgroup final pass reads every chunk, and
// its latency grows with their number.
void stats_chunks(int64_t np, int* chunks, int64_t* rows) {
  int64_t r = std::max<int64_t>(64, (np + 127) / 128);
  *rows = r;
  *chunks = static_cast<int>((np + r - 1) / r);
}

Layout make_layout(const ottx_problem* pb, const KernelCfg& k, const PassPlan& gp, const PassPlan& fp, int outer,
                   int64_t pad = kRowPad) {
  Layout L{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off = align_up(off + bytes);
    return at;
  };
  const int64_t n_pad = pad_rows(pb->n, pad), m_pad = pad_rows(pb->m, pad);
  const size_t chunk_bytes = static_cast<size_t>(k.frag_chunks()) * (k.prec == 1 ? 2048 : 1024);
  L.frag_x = take(static_cast<size_t>(n_pad / 16) * chunk_bytes);
  L.frag_y = take(static_cast<size_t>(m_pad / 16) * chunk_bytes);
  L.norm_x = take(n_pad * sizeof(float));
  L.norm_y = take(m_pad * sizeof(float));
  L.g_save = take(m_pad * sizeof(float));
  L.f_save = take(n_pad * sizeof(float));  // parallel updates: the f the g-update reads
  L.lse_x = take(n_pad * sizeof(float));
  L.lse_y = take(m_pad * sizeof(float));
  const size_t part_g = static_cast<size_t>(gp.splits) * gp.tiles * k.cols();
  const size_t part_f = static_cast<size_t>(fp.splits) * fp.tiles * k.cols();
  L.partial = take(std::max(part_g, part_f) * sizeof(float2));
  const int max_tiles = std::max(gp.tiles, fp.tiles);
  L.counters = take(static_cast<size_t>(max_tiles) * sizeof(uint32_t));
  L.err_col = take(static_cast<size_t>(max_tiles) * sizeof(double));
  L.mass = take(static_cast<size_t>(max_tiles) * sizeof(double));
  L.err_row = take(static_cast<size_t>(max_tiles) * sizeof(double));
  L.state = take(sizeof(SolveState));
  L.dual = take(16 * sizeof(double));
  L.dual_part = take(4 * kDualBlocks * 2 * sizeof(double));  // 3 dual sides + the recentering sums
  L.scal = take(sizeof(SolveScalars));
  if (pb->tau_a < 1.0 && pb->tau_b < 1.0) {
    // recentred unbalanced iterations: the f-update's log-sum-exp terms, the
    // column marginals (then the shifted output g) and the soft minima
    L.rc_app = take(n_pad * sizeof(float));
    L.rc_col = take(m_pad * sizeof(float));
    L.rc = take(sizeof(RcScalars));
  }
  stats_chunks(pb->n, &L.chunks_x, &L.rows_chunk_x);
  stats_chunks(pb->m, &L.chunks_y, &L.rows_chunk_y);
  L.stats_x = take(static_cast<size_t>(L.chunks_x) * pb->d * 2 * sizeof(double));
  L.stats_y = take(static_cast<size_t>(L.chunks_y) * pb->d * 2 * sizeof(double));
  L.stats_out = take(8 * sizeof(double));
  L.mu = take(static_cast<size_t>(pb->d) * sizeof(float));
  // uniform weights when the caller passes none (linear_problem.py:65-79)
  L.w_x = pb->a ? 0 : take(static_cast<size_t>(pb->n) * sizeof(float));
  L.w_y = pb->b ? 0 : take(static_cast<size_t>(pb->m) * sizeof(float));
  if (wants_std(pb)) {
    L.mom = moment_layout(pb->n, pb->m, pb->d);
    L.moments = take(L.mom.total);
  }
  (void)outer;
  L.total = off;
  return L;
}

struct Plan {
  KernelCfg k;
  PassPlan gp, fp;
  Layout L;
};

// The tile configuration of a problem that names none (-1), which is also what
// the operator entry points (no tile_config argument) run with.  From 2^24
// pairs on: the choices the tuner makes on MI355X at n = m = 16384, d = 64 and
// at n = 32768, m = 8192, d = 128 (bf16 split: variant 1, NB 4 or 2 by d, MT
// 2, two workgroups per CU; fp32: variant 2, NB 2, MT 4, four per CU, or
// beyond 64 features variant 1 with three per CU).  Smaller problems: variant
// 0 at four workgroups per CU.
int default_tile_config(int key, int d, int64_t n, int64_t m) {
  if (static_cast<double>(n) * static_cast<double>(m) < 16777216.0) return 0;
  if (key == OTTX_PREC_BF16_SPLIT && d <= 128) return 1 | (1 << 8);
  if (key == OTTX_PREC_FP32) {
    const int ks = (d + 3) / 4;
    if (ks >= 3 && ks <= 16) return 2;
    if (ks > 16 && ks <= 32) return 1 | (5 << 8);
  }
  return 0;
}

bool make_plan(const ottx_problem* pb, int tile_config, Plan* P) {
  int tm = 4;
  if (tile_config < 0) tile_config = default_tile_config(kernel_key(pb), pb->d, pb->n, pb->m);
  if (!select_kernel(kernel_key(pb), pb->d, tile_config, &P->k, &tm)) return false;
  const int cus = device_cus();
  // Every column tile (16 NB W points) and streamed tile (16 MT rows) must
  // divide the row padding, so tiles never straddle the end of an array.
  if (kRowPad % P->k.cols() != 0 || kRowPad % (32 * P->k.mt) != 0) return false;
  P->gp = plan_pass(P->k, pb->n, pb->m, cus, tm);
  P->fp = plan_pass(P->k, pb->m, pb->n, cus, tm);
  if (!even_tiles(P->k, P->gp) || !even_tiles(P->k, P->fp)) return false;
  P->L = make_layout(pb, P->k, P->gp, P->fp, 1);
  return true;
}

// Workspace large enough for every tile configuration of the problem.
size_t max_workspace(const ottx_problem* pb) {
  KernelCfg v[8];
  const int n = kernel_variants(kernel_key(pb), pb->d, v, 8);
  size_t best = 0;
  for (int vi = 0; vi < n; ++vi)
    for (int ti = 0; ti < kNumTargets; ++ti) {
      Plan P;
      if (make_plan(pb, vi | (ti << 8), &P)) best = std::max(best, P.L.total);
    }
  return best;
}

int check_problem(const ottx_problem* pb) {
  if (!pb || pb->n <= 0 || pb->m <= 0 || pb->d <= 0 || pb->d > 256) return OTTX_EINVAL;
  if (pb->n > (int64_t{1} << 30) || pb->m > (int64_t{1} << 30)) return OTTX_EINVAL;
  if (!pb->x || !pb->y) return OTTX_EINVAL;  // a or b NULL: uniform weights
  if (!(pb->epsilon > 0.0) && !(pb->relative_epsilon > 0.0)) return OTTX_EINVAL;
  if (!(pb->inv_scale_cost > 0.0) || pb->relative_epsilon < 0.0) return OTTX_EINVAL;
  if (!(pb->tau_a > 0.0 && pb->tau_a <= 1.0) || !(pb->tau_b > 0.0 && pb->tau_b <= 1.0)) return OTTX_EINVAL;
  if (pb->precision != OTTX_PREC_FP32 && pb->precision != OTTX_PREC_BF16_SPLIT) return OTTX_EINVAL;
  if (pb->epsilon_statistic != OTTX_EPS_STAT_MEAN && pb->epsilon_statistic != OTTX_EPS_STAT_STD) return OTTX_EINVAL;
  CostForm cf;
  if (!cost_form(pb->cost, pb->cost_p, &cf)) return OTTX_EINVAL;
  // without a closed form for mean(C) and std(C) the caller resolves a
  // relative epsilon (ottx_cost_summary)
  if (pb->cost != OTTX_COST_SQEUCLIDEAN && pb->relative_epsilon > 0.0) return OTTX_EUNSUPPORTED;
  return OTTX_OK;
}


#define OTTX_HIP(call)                              \
  do {                                              \
    if ((call) != hipSuccess) return OTTX_EHIP;     \
  } while (0)

#define OTTX_LAUNCHED()                                  \
  do {                                                   \
    if (hipGetLastError() != hipSuccess) return OTTX_EHIP; \
  } while (0)

struct Prepared {
  char* base;
  Layout L;
  void* frag_x;
  void* frag_y;
  float* norm_x;
  float* norm_y;
  float* mu;
  double* stats_out;
  SolveScalars* scal;
  const float* a;  // the caller's weights, or the uniform ones in the workspace
  const float* b;
};


// The prepared buffers of a workspace (no launch).
void prepared_view(const ottx_problem* pb, const Layout& L, void* ws, Prepared* P) {
  char* base = static_cast<char*>(ws);
  P->base = base;
  P->L = L;
  P->frag_x = base + L.frag_x;
  P->frag_y = base + L.frag_y;
  P->norm_x = reinterpret_cast<float*>(base + L.norm_x);
  P->norm_y = reinterpret_cast<float*>(base + L.norm_y);
  P->mu = reinterpret_cast<float*>(base + L.mu);
  P->stats_out = reinterpret_cast<double*>(base + L.stats_out);
  P->scal = reinterpret_cast<SolveScalars*>(base + L.scal);
  P->a = pb->a ? pb->a : reinterpret_cast<const float*>(base + L.w_x);
  P->b = pb->b ? pb->b : reinterpret_cast<const float*>(base + L.w_y);
}

// Statistics, centring shift, fragments and norms of both clouds (three
// launches), plus the loop-state reset of a solve when ``init`` names one.
int prepare(const ottx_problem* pb, const KernelCfg& k, const Layout& L, void* ws, hipStream_t st,
            Prepared* P, const PrepInit* init = nullptr) {
  prepared_view(pb, L, ws, P);
  char* base = P->base;
  double* sx = reinterpret_cast<double*>(base + L.stats_x);
  double* sy = reinterpret_cast<double*>(base + L.stats_y);
  // the first g-update's shift probe runs beside the statistics
  const bool probe = init != nullptr && init->st != nullptr && init->lse_g != nullptr &&
                     init->lse_g_mode == kShiftProbe;
  const int64_t m_pad = pad_rows(pb->m);
  const int probe_blocks = probe ? static_cast<int>((m_pad + kProbeCols - 1) / kProbeCols) : 0;
  hipLaunchKernelGGL(prep_stats_kernel, dim3(L.chunks_x + L.chunks_y + probe_blocks), dim3(256), 0, st, pb->x,
                     pb->n, L.rows_chunk_x, L.chunks_x, pb->y, pb->m, L.rows_chunk_y, L.chunks_y, pb->d, sx, sy,
                     probe ? init->f : nullptr, pb->inv_scale_cost, m_pad, probe ? init->lse_g : nullptr);
  hipLaunchKernelGGL(stats_final_kernel, dim3(1), dim3(kFinalThreads), 0, st, sx, L.chunks_x, sy, L.chunks_y, pb->d,
                     pb->n, pb->m, pb->inv_scale_cost, P->stats_out, P->mu, P->scal,
                     wants_std(pb) ? 0.0 : pb->relative_epsilon, pb->epsilon, pb->tau_a, pb->tau_b);
  OTTX_LAUNCHED();
  if (wants_std(pb)) {
    // epsilon = relative_epsilon * std(C), resolved on the device like the mean
    const int rc = launch_std(pb->x, pb->n, pb->y, pb->m, pb->d, P->mu, pb->inv_scale_cost, base + L.moments, L.mom,
                              P->stats_out, P->scal, pb->relative_epsilon, pb->tau_a, pb->tau_b, st);
    if (rc != OTTX_OK) return rc;
  }
  PrepArgs pa{};
  const int chunks = k.frag_chunks();
  pa.pts[0] = pb->x;
  pa.pts[1] = pb->y;
  pa.np[0] = pb->n;
  pa.np[1] = pb->m;
  pa.np_pad[0] = pad_rows(pb->n);
  pa.np_pad[1] = pad_rows(pb->m);
  pa.frag[0] = P->frag_x;
  pa.frag[1] = P->frag_y;
  pa.norm[0] = P->norm_x;
  pa.norm[1] = P->norm_y;
  pa.wuser[0] = pb->a;
  pa.wuser[1] = pb->b;
  pa.wfill[0] = pb->a ? nullptr : reinterpret_cast<float*>(base + L.w_x);
  pa.wfill[1] = pb->b ? nullptr : reinterpret_cast<float*>(base + L.w_y);
  pa.mu = P->mu;
  pa.scal = P->scal;
  pa.d = pb->d;
  pa.chunks = chunks;
  pa.prec = k.prec;
  pa.s = k.prec == 2 ? 0.0 : pb->inv_scale_cost;  // no norm term in the cost stream's exponent
  auto blocks = [](int64_t threads) { return static_cast<int>((threads + 255) / 256); };
  int cur = 0;
  for (int side = 0; side < 2; ++side) {
    pa.pack_threads[side] = (pa.np_pad[side] / 16) * chunks * 64;
    pa.seg[side] = cur;
    cur += blocks(pa.pack_threads[side]);
  }
  for (int side = 0; side < 2; ++side) {
    pa.seg[2 + side] = cur;
    cur += blocks(pa.np_pad[side] * kNormLanes);
  }
  for (int side = 0; side < 2; ++side) {
    pa.seg[4 + side] = cur;
    if (pa.wfill[side]) cur += blocks(pa.np[side]);
  }
  pa.seg[6] = cur;
  if (init != nullptr && init->st != nullptr) {
    pa.init = *init;
    pa.init.scal = P->scal;
    const int64_t span = std::max<int64_t>(std::max<int64_t>(pb->n, pb->m),
                                           std::max<int64_t>(init->outer, init->n_counters));
    pa.blocks_init = static_cast<int>(std::min<int64_t>(256, blocks(span)));
    cur += pa.blocks_init;
  }
  hipLaunchKernelGGL(prep_pack_kernel, dim3(cur), dim3(256), 0, st, pa);
  OTTX_LAUNCHED();
  return OTTX_OK;
}

// Arguments common to every half-step of one problem.  side_out = 1 for the
// g-update (out = y), 0 for the f-update (out = x).
HalfStepArgs base_args(const ottx_problem* pb, const KernelCfg& k, const PassPlan& plan, const Prepared& P,
                       int side_out, float* in_pot, float* out_pot, double eps) {
  HalfStepArgs a{};
  const bool g_upd = side_out == 1;
  a.in_frag = g_upd ? P.frag_x : P.frag_y;
  a.in_pot = in_pot;
  a.in_norm = g_upd ? P.norm_x : P.norm_y;
  a.out_frag = g_upd ? P.frag_y : P.frag_x;
  a.out_pot = out_pot;
  a.out_norm = g_upd ? P.norm_y : P.norm_x;
  a.out_w = g_upd ? P.b : P.a;
  a.n_in = g_upd ? pb->n : pb->m;
  a.n_out = g_upd ? pb->m : pb->n;
  a.n_in_rb = plan.n_in_rb;
  a.rb_per_split = plan.rb_per_split;
  a.splits = plan.splits;
  a.tiles = plan.tiles;
  a.norm_p = 1;
  const double s = pb->inv_scale_cost;
  a.inv2s = static_cast<float>(1.0 / (2.0 * s));
  a.two_s = 2.0 * s;
  a.c = static_cast<float>(2.0 * s * 1.4426950408889634 / (eps > 0.0 ? eps : 1.0));
  a.eps = static_cast<float>(eps);
  a.eps_marg = static_cast<float>(pb->epsilon);
  const double tau = g_upd ? pb->tau_b : pb->tau_a;
  a.tau = static_cast<float>(tau);
  a.inv_rho = tau == 1.0 ? 0.f : static_cast<float>((1.0 - tau) / (pb->epsilon * tau));
  a.momentum = 1.f;
  a.balanced = (pb->tau_a == 1.0 && pb->tau_b == 1.0) ? 1 : 0;
  char* base = P.base;
  a.partial = reinterpret_cast<float2*>(base + P.L.partial);
  a.tile_counter = reinterpret_cast<uint32_t*>(base + P.L.counters);
  a.err_partial = reinterpret_cast<double*>(base + (g_upd ? P.L.err_col : P.L.err_row));
  a.mass_partial = g_upd ? reinterpret_cast<double*>(base + P.L.mass) : nullptr;
  a.state = reinterpret_cast<SolveState*>(base + P.L.state);
  a.scal = P.scal;
  a.pack_scal = P.scal;
  a.side = side_out;
  if (k.prec == 2) {
    // the cost stream reads the raw points and has no norm term
    CostForm cf;
    cost_form(pb->cost, pb->cost_p, &cf);
    a.in_frag = g_upd ? pb->x : pb->y;
    a.out_frag = g_upd ? pb->y : pb->x;
    a.d = pb->d;
    a.cost_p = cf.p;
    a.cost_alpha = cf.alpha;
    a.cost_beta = cf.beta;
  }
  return a;
}

int launch_halfstep(const KernelCfg& k, const HalfStepArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(k.fn, dim3(a.tiles, a.splits), dim3(64 * k.w), 0, st, a);
  OTTX_LAUNCHED();
  return OTTX_OK;
}

// Per-thread host helpers: a non-blocking stream for status reads and a
// pinned mailbox they land in.
struct HostHelpers {
  hipStream_t side = nullptr;
  void* mailbox = nullptr;  // 4 KiB pinned: [0, 1024) copies, [1024, ...) the SolveMail
  int device = -1;
  uint32_t seq = 0;         // sequence number of the last solve on this thread
};

int host_helpers(HostHelpers** out) {
  static thread_local HostHelpers h;
  int dev = 0;
  OTTX_HIP(hipGetDevice(&dev));
  if (h.device != dev) {
    if (h.side) (void)hipStreamDestroy(h.side);
    if (h.mailbox) (void)hipHostFree(h.mailbox);
    h.side = nullptr;
    h.mailbox = nullptr;
    OTTX_HIP(hipStreamCreateWithFlags(&h.side, hipStreamNonBlocking));
    OTTX_HIP(hipHostMalloc(&h.mailbox, 4096, hipHostMallocDefault));
    h.device = dev;
  }
  *out = &h;
  return OTTX_OK;
}

}  // namespace ottx

using namespace ottx;

extern "C" {

int ottx_abi_version(void) { return OTTX_ABI_VERSION; }

const char* ottx_strerror(int status) {
  switch (status) {
    case OTTX_OK: return "ok";
    case OTTX_EINVAL: return "invalid argument";
    case OTTX_EUNSUPPORTED: return "unsupported option";
    case OTTX_EHIP: return "HIP runtime error";
    case OTTX_EWORKSPACE: return "workspace too small or misaligned";
    case OTTX_ESCRATCH: return "problem larger than the no-allocation scratch (OTTX_NOALLOC_SCRATCH_BYTES)";
    default: return "unknown status";
  }
}

size_t ottx_mean_cost_workspace_bytes(int64_t n, int64_t m, int32_t d) {
  if (n <= 0 || m <= 0 || d <= 0) return 0;
  int cx, cy;
  int64_t rx, ry;
  stats_chunks(n, &cx, &rx);
  stats_chunks(m, &cy, &ry);
  return align_up(static_cast<size_t>(cx) * d * 16) + align_up(static_cast<size_t>(cy) * d * 16) +
         align_up(8 * sizeof(double));
}

int ottx_mean_cost(int64_t n, int64_t m, int32_t d, const float* x, const float* y, double inv_scale_cost,
                   double* out_host, void* workspace, size_t workspace_bytes, void* stream) {
  if (n <= 0 || m <= 0 || d <= 0 || !x || !y || !out_host || !workspace) return OTTX_EINVAL;
  if (workspace_bytes < ottx_mean_cost_workspace_bytes(n, m, d)) return OTTX_EWORKSPACE;
  if (reinterpret_cast<uintptr_t>(workspace) & 255) return OTTX_EWORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  int cx, cy;
  int64_t rx, ry;
  stats_chunks(n, &cx, &rx);
  stats_chunks(m, &cy, &ry);
  char* base = static_cast<char*>(workspace);
  double* sx = reinterpret_cast<double*>(base);
  double* sy = reinterpret_cast<double*>(base + align_up(static_cast<size_t>(cx) * d * 16));
  double* o = reinterpret_cast<double*>(base + align_up(static_cast<size_t>(cx) * d * 16) +
                                        align_up(static_cast<size_t>(cy) * d * 16));
  hipLaunchKernelGGL(stats_partial_kernel, dim3(cx), dim3(256), 0, st, x, n, d, rx, sx);
  hipLaunchKernelGGL(stats_partial_kernel, dim3(cy), dim3(256), 0, st, y, m, d, ry, sy);
  hipLaunchKernelGGL(stats_final_kernel, dim3(1), dim3(kFinalThreads), 0, st, sx, cx, sy, cy, d, n, m,
                     inv_scale_cost, o, static_cast<float*>(nullptr), static_cast<SolveScalars*>(nullptr), 0.0, 0.0,
                     1.0, 1.0);
  OTTX_LAUNCHED();
  OTTX_HIP(hipMemcpyAsync(out_host, o, sizeof(double), hipMemcpyDeviceToHost, st));
  OTTX_HIP(hipStreamSynchronize(st));
  return OTTX_OK;
}

size_t ottx_cost_moments_workspace_bytes(int64_t n, int64_t m, int32_t d) {
  if (n <= 0 || m <= 0 || d <= 0 || d > 256) return 0;
  return ottx_mean_cost_workspace_bytes(n, m, d) + align_up(static_cast<size_t>(d) * sizeof(float)) +
         moment_layout(n, m, d).total;
}

int ottx_cost_moments(int64_t n, int64_t m, int32_t d, const float* x, const float* y, double inv_scale_cost,
                      double* out_host, void* workspace, size_t workspace_bytes, void* stream) {
  if (n <= 0 || m <= 0 || d <= 0 || d > 256 || !x || !y || !out_host || !workspace) return OTTX_EINVAL;
  if (!(inv_scale_cost > 0.0)) return OTTX_EINVAL;
  if (workspace_bytes < ottx_cost_moments_workspace_bytes(n, m, d)) return OTTX_EWORKSPACE;
  if (reinterpret_cast<uintptr_t>(workspace) & 255) return OTTX_EWORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  int cx, cy;
  int64_t rx, ry;
  stats_chunks(n, &cx, &rx);
  stats_chunks(m, &cy, &ry);
  char* base = static_cast<char*>(workspace);
  double* sx = reinterpret_cast<double*>(base);
  double* sy = reinterpret_cast<double*>(base + align_up(static_cast<size_t>(cx) * d * 16));
  const size_t stats_bytes = ottx_mean_cost_workspace_bytes(n, m, d);
  double* o = reinterpret_cast<double*>(base + stats_bytes - align_up(8 * sizeof(double)));
  float* mu = reinterpret_cast<float*>(base + stats_bytes);
  char* region = base + stats_bytes + align_up(static_cast<size_t>(d) * sizeof(float));
  hipLaunchKernelGGL(stats_partial_kernel, dim3(cx), dim3(256), 0, st, x, n, d, rx, sx);
  hipLaunchKernelGGL(stats_partial_kernel, dim3(cy), dim3(256), 0, st, y, m, d, ry, sy);
  hipLaunchKernelGGL(stats_final_kernel, dim3(1), dim3(kFinalThreads), 0, st, sx, cx, sy, cy, d, n, m,
                     inv_scale_cost, o, mu, static_cast<SolveScalars*>(nullptr), 0.0, 0.0, 1.0, 1.0);
  OTTX_LAUNCHED();
  const int rc = launch_std(x, n, y, m, d, mu, inv_scale_cost, region, moment_layout(n, m, d), o, nullptr, 0.0, 1.0,
                            1.0, st);
  if (rc != OTTX_OK) return rc;
  double host[2];
  OTTX_HIP(hipMemcpyAsync(host, o, sizeof(host), hipMemcpyDeviceToHost, st));
  OTTX_HIP(hipStreamSynchronize(st));
  out_host[0] = host[0];
  out_host[1] = host[1];
  return OTTX_OK;
}

int ottx_cost_maxima(int64_t n, int64_t m, int32_t d, const float* x, const float* y, int32_t with_max_cost,
                     double* out_host, void* workspace, size_t workspace_bytes, void* stream) {
  if (n <= 0 || m <= 0 || d <= 0 || !x || !y || !out_host || !workspace) return OTTX_EINVAL;
  if (workspace_bytes < 256) return OTTX_EWORKSPACE;
  if (reinterpret_cast<uintptr_t>(workspace) & 255) return OTTX_EWORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  uint32_t* bits = static_cast<uint32_t*>(workspace);
  hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, st, bits, 4);
  hipLaunchKernelGGL(max_sq_norm_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, x, n, d,
                     bits + 0);
  hipLaunchKernelGGL(max_sq_norm_kernel, dim3(static_cast<unsigned>((m + 255) / 256)), dim3(256), 0, st, y, m, d,
                     bits + 1);
  if (with_max_cost) {
    const dim3 grid(static_cast<unsigned>((m + 63) / 64), static_cast<unsigned>((n + 3) / 4));
    hipLaunchKernelGGL(max_cost_kernel, grid, dim3(256), 0, st, x, y, n, m, d, bits + 2);
  }
  OTTX_LAUNCHED();
  uint32_t host[4];
  OTTX_HIP(hipMemcpyAsync(host, bits, sizeof(host), hipMemcpyDeviceToHost, st));
  OTTX_HIP(hipStreamSynchronize(st));
  float v[3];
  std::memcpy(v, host, sizeof(v));
  out_host[0] = v[0];
  out_host[1] = v[1];
  out_host[2] = with_max_cost ? static_cast<double>(v[2]) : 0.0;
  return OTTX_OK;
}

size_t ottx_cost_median_workspace_bytes(void) {
  return align_up(sizeof(MedianState)) + align_up(sizeof(unsigned long long) * kMedBins);
}

int ottx_cost_median(const ottx_problem* pb, double* out_host, void* workspace, size_t workspace_bytes,
                     void* stream) {
  int rc = check_problem(pb);
  if (rc != OTTX_OK) return rc;
  if (!out_host || !workspace) return OTTX_EINVAL;
  const int64_t n = pb->n, m = pb->m;
  const int d = pb->d;
  CostForm cf;
  if (!cost_form(pb->cost, pb->cost_p, &cf)) return OTTX_EINVAL;
  if (n > (int64_t{1} << 40) / m) return OTTX_EINVAL;  // n m <= 2^40 entries
  if (workspace_bytes < ottx_cost_median_workspace_bytes()) return OTTX_EWORKSPACE;
  if (reinterpret_cast<uintptr_t>(workspace) & 255) return OTTX_EWORKSPACE;
  if ((n + kTcTile - 1) / kTcTile > 65535) return OTTX_EINVAL;  // row blocks index the grid's y dimension
  hipStream_t st = static_cast<hipStream_t>(stream);
  MedianState* ms = static_cast<MedianState*>(workspace);
  unsigned long long* hist =
      reinterpret_cast<unsigned long long*>(static_cast<char*>(workspace) + align_up(sizeof(MedianState)));
  const dim3 grid(static_cast<unsigned>((m + kTcTile - 1) / kTcTile), static_cast<unsigned>((n + kTcTile - 1) / kTcTile));
  const float* x = pb->x;
  const float* y = pb->y;
  hipLaunchKernelGGL(median_init_kernel, dim3(1), dim3(256), 0, st, ms, hist,
                     static_cast<unsigned long long>(n) * static_cast<unsigned long long>(m));
  auto passes = [&](auto phi) {
    constexpr int PHI = decltype(phi)::value;
    hipLaunchKernelGGL((median_pass_kernel<0, PHI>), grid, dim3(256), 0, st, x, y, n, m, d, ms, hist, cf);
    hipLaunchKernelGGL(median_select_kernel<0>, dim3(1), dim3(1024), 0, st, ms, hist);
    hipLaunchKernelGGL((median_pass_kernel<1, PHI>), grid, dim3(256), 0, st, x, y, n, m, d, ms, hist, cf);
    hipLaunchKernelGGL(median_select_kernel<1>, dim3(1), dim3(1024), 0, st, ms, hist);
    hipLaunchKernelGGL((median_pass_kernel<2, PHI>), grid, dim3(256), 0, st, x, y, n, m, d, ms, hist, cf);
    hipLaunchKernelGGL(median_select_kernel<2>, dim3(1), dim3(1024), 0, st, ms, hist);
    hipLaunchKernelGGL((median_pass_kernel<3, PHI>), grid, dim3(256), 0, st, x, y, n, m, d, ms, hist, cf);
  };
  if (cf.phi == 1) passes(std::integral_constant<int, 1>{});
  else if (cf.phi == 2) passes(std::integral_constant<int, 2>{});
  else passes(std::integral_constant<int, 3>{});
  OTTX_LAUNCHED();
  MedianState host;
  OTTX_HIP(hipMemcpyAsync(&host, ms, sizeof(host), hipMemcpyDeviceToHost, st));
  OTTX_HIP(hipStreamSynchronize(st));
  float lo, hi;
  std::memcpy(&lo, &host.lo_bits, sizeof(lo));
  std::memcpy(&hi, &host.hi_bits, sizeof(hi));
  if (host.nan_seen) {
    out_host[0] = std::numeric_limits<double>::quiet_NaN();
  } else if (host.total % 2ull == 1ull) {
    out_host[0] = lo;
  } else {
    out_host[0] = 0.5 * (static_cast<double>(lo) + static_cast<double>(host.need_hi ? hi : lo));
  }
  return OTTX_OK;
}

size_t ottx_cost_summary_workspace_bytes(int64_t n, int64_t m) {
  if (n <= 0 || m <= 0) return 0;
  const int64_t blocks = ((n + kTcTile - 1) / kTcTile) * ((m + kTcTile - 1) / kTcTile);
  return static_cast<size_t>(256 + align_up(static_cast<size_t>(blocks) * 2 * sizeof(double)));
}

int ottx_cost_summary(const ottx_problem* pb, double* out_host, void* workspace, size_t workspace_bytes,
                      void* stream) {
  int rc = check_problem(pb);
  if (rc != OTTX_OK) return rc;
  if (!out_host || !workspace) return OTTX_EINVAL;
  CostForm cf;
  if (!cost_form(pb->cost, pb->cost_p, &cf)) return OTTX_EINVAL;
  if (workspace_bytes < ottx_cost_summary_workspace_bytes(pb->n, pb->m)) return OTTX_EWORKSPACE;
  if (reinterpret_cast<uintptr_t>(workspace) & 255) return OTTX_EWORKSPACE;
  const int64_t gx = (pb->m + kTcTile - 1) / kTcTile, gy = (pb->n + kTcTile - 1) / kTcTile;
  if (gy > 65535) return OTTX_EUNSUPPORTED;
  hipStream_t st = static_cast<hipStream_t>(stream);
  char* base = static_cast<char*>(workspace);
  double* result = reinterpret_cast<double*>(base);            // [0] sum, [1] sum of squares
  uint32_t* max_bits = reinterpret_cast<uint32_t*>(base + 64);
  double* partial = reinterpret_cast<double*>(base + 256);
  const dim3 grid(static_cast<unsigned>(gx), static_cast<unsigned>(gy));
  hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, st, max_bits, 1);
  if (cf.phi == 1)
    hipLaunchKernelGGL(cost_summary_kernel<1>, grid, dim3(256), 0, st, pb->x, pb->y, pb->n, pb->m, pb->d, cf,
                       partial, gx * gy, max_bits);
  else if (cf.phi == 2)
    hipLaunchKernelGGL(cost_summary_kernel<2>, grid, dim3(256), 0, st, pb->x, pb->y, pb->n, pb->m, pb->d, cf,
                       partial, gx * gy, max_bits);
  else
    hipLaunchKernelGGL(cost_summary_kernel<3>, grid, dim3(256), 0, st, pb->x, pb->y, pb->n, pb->m, pb->d, cf,
                       partial, gx * gy, max_bits);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(1024), 0, st, static_cast<const double*>(partial), gx * gy,
                     result);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(1024), 0, st,
                     static_cast<const double*>(partial + gx * gy), gx * gy, result + 1);
  OTTX_LAUNCHED();
  double host[9];
  OTTX_HIP(hipMemcpyAsync(host, result, sizeof(host), hipMemcpyDeviceToHost, st));
  OTTX_HIP(hipStreamSynchronize(st));
  uint32_t mb;
  std::memcpy(&mb, &host[8], sizeof(mb));
  float mx;
  std::memcpy(&mx, &mb, sizeof(mx));
  out_host[0] = host[0];
  out_host[1] = host[1];
  out_host[2] = mx;
  return OTTX_OK;
}

size_t ottx_op_workspace_bytes(const ottx_problem* pb) {
  if (check_problem(pb) != OTTX_OK) return 0;
  return max_workspace(pb);
}

int ottx_transport_matrix(const ottx_problem* pb, const float* f, const float* g, double eps, float* out,
                          void* stream) {
  int rc = check_problem(pb);
  if (rc != OTTX_OK) return rc;
  if (!f || !g || !out || !(eps > 0.0)) return OTTX_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 grid(static_cast<unsigned>((pb->m + 63) / 64), static_cast<unsigned>((pb->n + 3) / 4));
  CostForm cf;
  if (!cost_form(pb->cost, pb->cost_p, &cf)) return OTTX_EINVAL;
  const float s = static_cast<float>(pb->inv_scale_cost), inv_eps = static_cast<float>(1.0 / eps);
  if (cf.phi == 1)
    hipLaunchKernelGGL(transport_matrix_kernel<1>, grid, dim3(256), 0, st, pb->x, pb->y, pb->n, pb->m, pb->d, s, f,
                       g, inv_eps, out, cf);
  else if (cf.phi == 2)
    hipLaunchKernelGGL(transport_matrix_kernel<2>, grid, dim3(256), 0, st, pb->x, pb->y, pb->n, pb->m, pb->d, s, f,
                       g, inv_eps, out, cf);
  else
    hipLaunchKernelGGL(transport_matrix_kernel<3>, grid, dim3(256), 0, st, pb->x, pb->y, pb->n, pb->m, pb->d, s, f,
                       g, inv_eps, out, cf);
  OTTX_LAUNCHED();
  return OTTX_OK;
}

size_t ottx_transport_cost_workspace_bytes(int64_t n, int64_t m) {
  if (n <= 0 || m <= 0) return 0;
  const int64_t blocks = ((n + kTcTile - 1) / kTcTile) * ((m + kTcTile - 1) / kTcTile);
  return static_cast<size_t>(256 + ((blocks * 8 + 255) / 256) * 256);
}

int ottx_transport_cost(const ottx_problem* pb, const float* f, const float* g, double eps, const float* x_other,
                        const float* y_other, int32_t d_other, double other_inv_scale_cost, double* out_host,
                        void* workspace, size_t workspace_bytes, void* stream) {
  int rc = check_problem(pb);
  if (rc != OTTX_OK) return rc;
  if (!f || !g || !out_host || !workspace || !(eps > 0.0)) return OTTX_EINVAL;
  const bool same = x_other == nullptr && y_other == nullptr;
  if (!same && (!x_other || !y_other || d_other <= 0)) return OTTX_EINVAL;
  if (workspace_bytes < ottx_transport_cost_workspace_bytes(pb->n, pb->m)) return OTTX_EWORKSPACE;
  if (reinterpret_cast<uintptr_t>(workspace) & 255) return OTTX_EWORKSPACE;
  const int64_t gx = (pb->m + kTcTile - 1) / kTcTile, gy = (pb->n + kTcTile - 1) / kTcTile;
  if (gy > 65535) return OTTX_EUNSUPPORTED;
  hipStream_t st = static_cast<hipStream_t>(stream);
  double* result = static_cast<double*>(workspace);
  double* partial = reinterpret_cast<double*>(static_cast<char*>(workspace) + 256);
  const dim3 grid(static_cast<unsigned>(gx), static_cast<unsigned>(gy));
  const float s = static_cast<float>(pb->inv_scale_cost);
  const float inv_eps = static_cast<float>(1.0 / eps);
  CostForm cf;
  if (!cost_form(pb->cost, pb->cost_p, &cf)) return OTTX_EINVAL;
  const float so = static_cast<float>(other_inv_scale_cost);
  auto launch = [&](auto same_c, auto phi_c) {
    constexpr bool SAME = decltype(same_c)::value;
    constexpr int PHI = decltype(phi_c)::value;
    hipLaunchKernelGGL((transport_cost_kernel<SAME, PHI>), grid, dim3(256), 0, st, pb->x, pb->y, pb->n, pb->m, pb->d,
                       s, f, g, inv_eps, SAME ? pb->x : x_other, SAME ? pb->y : y_other, SAME ? pb->d : d_other,
                       SAME ? s : so, partial, cf);
  };
  using T = std::true_type;
  using F = std::false_type;
  using P1 = std::integral_constant<int, 1>;
  using P2 = std::integral_constant<int, 2>;
  using P3 = std::integral_constant<int, 3>;
  if (same) {
    if (cf.phi == 1) launch(T{}, P1{});
    else if (cf.phi == 2) launch(T{}, P2{});
    else launch(T{}, P3{});
  } else {
    if (cf.phi == 1) launch(F{}, P1{});
    else if (cf.phi == 2) launch(F{}, P2{});
    else launch(F{}, P3{});
  }
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(1024), 0, st, static_cast<const double*>(partial), gx * gy,
                     result);
  OTTX_LAUNCHED();
  double host = 0.0;
  OTTX_HIP(hipMemcpyAsync(&host, result, sizeof(double), hipMemcpyDeviceToHost, st));
  OTTX_HIP(hipStreamSynchronize(st));
  *out_host = host;
  return OTTX_OK;
}

// Slices of the streamed cloud for the gradient: enough workgroups for the
// chip (about 2048), each slice a multiple of 64 points.
static void grad_plan(int64_t np, int64_t nq, int64_t* slice, int* slices) {
  const int64_t tiles = (np + kTcTile - 1) / kTcTile;
  const int64_t q_tiles = (nq + kTcTile - 1) / kTcTile;
  int64_t want = std::max<int64_t>(1, std::min<int64_t>(q_tiles, (2048 + tiles - 1) / tiles));
  const int64_t per = (q_tiles + want - 1) / want;
  *slice = per * kTcTile;
  *slices = static_cast<int>((q_tiles + per - 1) / per);
}

size_t ottx_cost_gradient_workspace_bytes(const ottx_problem* pb, int32_t side) {
  if (check_problem(pb) != OTTX_OK || (side != 0 && side != 1) || pb->d > kGdMaxD) return 0;
  const int64_t np = side == 0 ? pb->n : pb->m, nq = side == 0 ? pb->m : pb->n;
  int64_t slice;
  int slices;
  grad_plan(np, nq, &slice, &slices);
  return static_cast<size_t>(((static_cast<int64_t>(slices) * np * pb->d * 4 + 255) / 256) * 256);
}

int ottx_cost_gradient(const ottx_problem* pb, const float* f, const float* g, double eps, int32_t side, float* grad,
                       void* workspace, size_t workspace_bytes, void* stream) {
  int rc = check_problem(pb);
  if (rc != OTTX_OK) return rc;
  if (!f || !g || !grad || !workspace || !(eps > 0.0) || (side != 0 && side != 1)) return OTTX_EINVAL;
  if (pb->d > kGdMaxD || pb->cost != OTTX_COST_SQEUCLIDEAN) return OTTX_EUNSUPPORTED;
  if (workspace_bytes < ottx_cost_gradient_workspace_bytes(pb, side)) return OTTX_EWORKSPACE;
  const bool rows = side == 0;
  const int64_t np = rows ? pb->n : pb->m, nq = rows ? pb->m : pb->n;
  int64_t slice;
  int slices;
  grad_plan(np, nq, &slice, &slices);
  const int64_t tiles = (np + kTcTile - 1) / kTcTile;
  if (tiles > 0x7fffffff || slices > 65535) return OTTX_EUNSUPPORTED;
  hipStream_t st = static_cast<hipStream_t>(stream);
  float* partial = static_cast<float*>(workspace);
  hipLaunchKernelGGL(grad_partial_kernel, dim3(static_cast<unsigned>(tiles), static_cast<unsigned>(slices)), dim3(256),
                     0, st, rows ? pb->x : pb->y, rows ? f : g, np, rows ? pb->y : pb->x, rows ? g : f, nq, pb->d,
                     static_cast<float>(pb->inv_scale_cost), static_cast<float>(1.0 / eps), slice, partial);
  const int64_t count = np * pb->d;
  hipLaunchKernelGGL(grad_reduce_kernel, dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256), 0, st,
                     static_cast<const float*>(partial), count, slices, 2.0 * pb->inv_scale_cost, grad);
  OTTX_LAUNCHED();
  return OTTX_OK;
}

}  // extern "C"

namespace ottx {

// Arguments of ottx_pointcloud_op checked, the default plan made: OTTX_OK
// and *PL, or the status to return.
int pointcloud_op_plan(const ottx_problem* pb, int32_t op, int32_t axis, const float* f, const float* g,
                       const float* log_marginal, double eps, float* out, Plan* PL) {
  int rc = check_problem(pb);
  if (rc != OTTX_OK) return rc;
  if (!f || !g || !out || !(eps > 0.0)) return OTTX_EINVAL;
  if (axis != 0 && axis != 1) return OTTX_EINVAL;
  if (op != OTTX_OP_APPLY_LSE && op != OTTX_OP_UPDATE && op != OTTX_OP_MARGINAL) return OTTX_EINVAL;
  if (op == OTTX_OP_UPDATE && !log_marginal) return OTTX_EINVAL;
  if (!make_plan(pb, -1, PL)) return OTTX_EUNSUPPORTED;
  return OTTX_OK;
}

// The operator call on a checked plan with its workspace (caller memory or
// the no-allocation scratch).
int pointcloud_op_run(const ottx_problem* pb, const Plan& PL, int32_t op, int32_t axis, const float* f,
                      const float* g, const float* log_marginal, double eps, float* out, void* workspace,
                      hipStream_t st) {
  int rc;
  const KernelCfg& k = PL.k;
  const PassPlan& gp = PL.gp;
  const PassPlan& fp = PL.fp;
  const Layout& L = PL.L;
  Prepared P;
  rc = prepare(pb, k, L, workspace, st, &P);
  if (rc != OTTX_OK) return rc;
  uint32_t* counters = reinterpret_cast<uint32_t*>(P.base + L.counters);
  SolveState* state = reinterpret_cast<SolveState*>(P.base + L.state);
  float* dummy_err = reinterpret_cast<float*>(P.base + L.dual);
  hipLaunchKernelGGL(init_state_kernel, dim3(1), dim3(256), 0, st, state, dummy_err, 0, counters,
                     std::max(gp.tiles, fp.tiles), 0, static_cast<float*>(nullptr), static_cast<const float*>(nullptr),
                     int64_t{0}, static_cast<float*>(nullptr), static_cast<const float*>(nullptr), int64_t{0});
  OTTX_LAUNCHED();
  const bool g_upd = axis == 0;
  const PassPlan& plan = g_upd ? gp : fp;
  const float* h = g_upd ? g : f;
  const float* in_pot = g_upd ? f : g;
  HalfStepArgs a = base_args(pb, k, plan, P, g_upd ? 1 : 0, const_cast<float*>(in_pot), const_cast<float*>(h), eps);
  a.state = nullptr;
  a.scal = nullptr;
  const int64_t n_out = g_upd ? pb->m : pb->n;
  if (op == OTTX_OP_UPDATE) {
    hipLaunchKernelGGL(copy_f32_kernel, dim3(static_cast<int>((n_out + 255) / 256)), dim3(256), 0, st, h, out,
                       n_out);
    OTTX_LAUNCHED();
    a.out_pot = out;
    a.out_logw = log_marginal;
    a.tau = 1.f;
    a.mode = kModeWritePot;
  } else if (op == OTTX_OP_APPLY_LSE) {
    a.out_vec = out;
    a.mode = kModeApply;
  } else {
    a.out_vec = out;
    a.eps_marg = static_cast<float>(eps);
    a.mode = kModeMarginal;
  }
  return launch_halfstep(k, a, st);
}

// The no-allocation scratch of ottx_pointcloud_op_noalloc: one array per
// device, held until the process ends.  In host code ``scratch`` names a
// host-side stand-in, never device memory: only the address that
// hipGetSymbolAddress returns for the current device (kept per device in
// scratch_ptr) may reach a kernel or a copy.
static __device__ __align__(256) unsigned char scratch[OTTX_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch (resolved once per device,
// with that device current).
int scratch_ptr(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return OTTX_EHIP;
  if (dev < 0 || dev >= kMaxDevices) return OTTX_EUNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return OTTX_EHIP;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return OTTX_OK;
}

}  // namespace ottx

extern "C" {

int ottx_pointcloud_op(const ottx_problem* pb, int32_t op, int32_t axis, const float* f, const float* g,
                       const float* log_marginal, double eps, float* out, void* workspace, size_t workspace_bytes,
                       void* stream) {
  Plan PL;
  const int rc = pointcloud_op_plan(pb, op, axis, f, g, log_marginal, eps, out, &PL);
  if (rc != OTTX_OK) return rc;
  if (!workspace) return OTTX_EINVAL;
  if (workspace_bytes < PL.L.total) return OTTX_EWORKSPACE;
  if (reinterpret_cast<uintptr_t>(workspace) & 255) return OTTX_EWORKSPACE;
  return pointcloud_op_run(pb, PL, op, axis, f, g, log_marginal, eps, out, workspace,
                           static_cast<hipStream_t>(stream));
}

int ottx_pointcloud_op_noalloc(const ottx_problem* pb, int32_t op, int32_t axis, const float* f, const float* g,
                               const float* log_marginal, double eps, float* out, void* stream) {
  Plan PL;
  int rc = pointcloud_op_plan(pb, op, axis, f, g, log_marginal, eps, out, &PL);
  if (rc != OTTX_OK) return rc;
  if (PL.L.total > static_cast<size_t>(OTTX_NOALLOC_SCRATCH_BYTES)) return OTTX_ESCRATCH;
  void* ws = nullptr;
  rc = scratch_ptr(&ws);
  if (rc != OTTX_OK) return rc;
  return pointcloud_op_run(pb, PL, op, axis, f, g, log_marginal, eps, out, ws, static_cast<hipStream_t>(stream));
}

size_t ottx_pointcloud_op_noalloc_bytes(const ottx_problem* pb) {
  Plan PL;
  if (check_problem(pb) != OTTX_OK || !make_plan(pb, -1, &PL)) return 0;
  return PL.L.total;
}

size_t ottx_sinkhorn_workspace_bytes(const ottx_problem* pb, const ottx_sinkhorn_opts* o) {
  if (check_problem(pb) != OTTX_OK || !o) return 0;
  Plan PL;
  if (!make_plan(pb, o->tile_config, &PL)) return 0;
  return PL.L.total;
}

int ottx_sinkhorn_solve(const ottx_problem* pb, const ottx_sinkhorn_opts* o, ottx_sinkhorn_out* out, void* workspace,
                        size_t workspace_bytes, void* stream) {
  int rc = check_problem(pb);
  if (rc != OTTX_OK) return rc;
  if (!o || !out || !out->f || !out->g || !out->errors || !workspace) return OTTX_EINVAL;
  if (o->inner_iterations < 1 || o->max_iterations < 0 || o->min_iterations < 0 || o->norm_error < 1)
    return OTTX_EINVAL;
  if (!(o->momentum > 0.0 && o->momentum < 2.0)) return OTTX_EINVAL;
  Plan PL;
  if (!make_plan(pb, o->tile_config, &PL)) return o->tile_config >= 0 ? OTTX_EINVAL : OTTX_EUNSUPPORTED;
  const KernelCfg& k = PL.k;
  const PassPlan& gp = PL.gp;
  const PassPlan& fp = PL.fp;
  const Layout& L = PL.L;
  if (workspace_bytes < L.total) return OTTX_EWORKSPACE;
  if (reinterpret_cast<uintptr_t>(workspace) & 255) return OTTX_EWORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  HostHelpers* hh = nullptr;
  rc = host_helpers(&hh);
  if (rc != OTTX_OK) return rc;

  const int inner = o->inner_iterations;
  const int outer = (o->max_iterations + inner - 1) / inner;
  const bool force_scan = o->min_iterations == o->max_iterations;
  const int total_blocks = force_scan ? o->max_iterations / inner : outer;

  if (o->init_mode < 0 || o->init_mode > 3) return OTTX_EINVAL;
  // recenter_potentials: balanced problems are recentred by the finalizer;
  // with one tau = 1 the reference leaves the potentials as they are
  // (sinkhorn.py:104-106); both taus < 1 run the translation-invariant
  // iteration of lse_step (sinkhorn.py:809-854), recentred errors and outputs.
  const bool both_unbalanced = pb->tau_a < 1.0 && pb->tau_b < 1.0;
  if (o->recenter_potentials && total_blocks == 0) return OTTX_EUNSUPPORTED;
  const int recenter = (o->recenter_potentials && pb->tau_a == 1.0 && pb->tau_b == 1.0) ? 1 : 0;
  const bool rc_mode = o->recenter_potentials && both_unbalanced;
  // parallel dual updates (sinkhorn.py:840-846): on the symmetric problems of
  // a Sinkhorn divergence (x against x, a = b, balanced, zero start) f = g
  // throughout and one pass per iteration updates both; any other problem
  // runs both updates from the previous state, the f-update first (saving
  // the old f, which the g-update then reads).
  const bool parallel = o->parallel_dual_updates != 0;
  if (parallel && o->recenter_potentials) return OTTX_EUNSUPPORTED;
  const bool symmetric = parallel && pb->x == pb->y && pb->n == pb->m && pb->a == pb->b && pb->tau_a == 1.0 &&
                         pb->tau_b == 1.0 && o->init_mode == OTTX_INIT_ZEROS;
  const bool par_general = parallel && !symmetric;
  // epsilon schedule (epsilon_scheduler.py:47-60) with sequential updates, not
  // with the recentred iteration (whose soft minima take rho at the target)
  const double eps_init = o->epsilon_init > 1.0 ? o->epsilon_init : 1.0;
  const double eps_decay = o->epsilon_decay > 0.0 ? o->epsilon_decay : 1.0;
  if (!(eps_decay <= 1.0) || !(eps_init < 1e30)) return OTTX_EINVAL;
  const bool sched = eps_init > 1.0;
  if (sched && (parallel || rc_mode)) return OTTX_EUNSUPPORTED;
  // adaptive momentum (acceleration.py:138-160) on L1 errors, from an
  // iteration that leaves two recorded blocks before it
  const int mom_start = o->momentum_start;
  if (mom_start < 0) return OTTX_EINVAL;
  const bool adaptive = mom_start > 0;
  // With norm_error != 1 the reference adds an L1 column to its errors
  // (sinkhorn.py:1032-1038) but records the first norm's error in every
  // column (one_iteration sets errors[it // inner, :] to solution_error(...)[0]),
  // so the weight reads the errors this solve records.
  if (adaptive && (parallel || mom_start < 2 * inner)) return OTTX_EUNSUPPORTED;
  const bool per_launch = sched || adaptive;
  char* base = static_cast<char*>(workspace);
  SolveState* state = reinterpret_cast<SolveState*>(base + L.state);
  uint32_t* counters = reinterpret_cast<uint32_t*>(base + L.counters);
  float* g_save = reinterpret_cast<float*>(base + L.g_save);
  double* dual = reinterpret_cast<double*>(base + L.dual);
  PrepInit pinit;
  pinit.st = state;
  pinit.errors = out->errors;
  pinit.outer = outer;
  pinit.counters = counters;
  pinit.n_counters = std::max(gp.tiles, fp.tiles);
  pinit.zero_init = zero_init_bits(o->init_mode);
  pinit.f = out->f;
  pinit.g = out->g;
  // a given f (warm start or restart) estimates the first f-update's shift;
  // the first g-update's comes from the given g, or from a row sample of the
  // log-sum-exp with the given f when g is not given
  if (o->init_mode == OTTX_INIT_GIVEN || o->init_mode == OTTX_INIT_F_GIVEN) {
    pinit.lse_f = reinterpret_cast<float*>(base + L.lse_x);
    pinit.tau_a = pb->tau_a;
    pinit.lse_g = reinterpret_cast<float*>(base + L.lse_y);
    pinit.lse_g_mode = o->init_mode == OTTX_INIT_GIVEN ? kShiftFromPot : kShiftProbe;
    pinit.tau_b = pb->tau_b;
  }
  Prepared P;
  rc = prepare(pb, k, L, workspace, st, &P, &pinit);
  if (rc != OTTX_OK) return rc;

  // Results reach the host through pinned memory: the block finalizer
  // publishes each block's progress, the output finalizer the results.
  SolveMail* mail = reinterpret_cast<SolveMail*>(static_cast<char*>(hh->mailbox) + 1024);
  const uint32_t seq = ++hh->seq;
  mail->ready = 0u;
  __atomic_store_n(&mail->progress, static_cast<uint64_t>(0), __ATOMIC_RELEASE);

  const double eps = pb->epsilon;
  HalfStepArgs ga = base_args(pb, k, gp, P, 1, out->f, out->g, eps);
  HalfStepArgs fa = base_args(pb, k, fp, P, 0, out->g, out->f, eps);
  for (HalfStepArgs* a : {&ga, &fa}) {
    a->mail = mail;
    a->seq = seq;
    a->norm_p = o->norm_error;
    a->momentum = static_cast<float>(o->momentum);
    a->errors = out->errors;
    a->row_err_partial = reinterpret_cast<const double*>(base + L.err_row);
    a->row_tiles = fp.tiles;
    a->inner = inner;
    a->min_it = o->min_iterations;
    a->max_it = o->max_iterations;
    a->outer = outer;
    a->force_scan = force_scan ? 1 : 0;
    a->threshold = static_cast<float>(o->threshold);
  }
  ga.out_save = g_save;
  ga.lse_shift = ga.lse_store = reinterpret_cast<float*>(base + L.lse_y);
  fa.lse_shift = fa.lse_store = reinterpret_cast<float*>(base + L.lse_x);
  float* f_save = reinterpret_cast<float*>(base + L.f_save);
  if (par_general) {
    // the g-update reads the f of the previous state; the error is the row
    // plus the column marginal error (sinkhorn.py:160-189) of that state,
    // measured by the next iteration's two passes
    ga.in_pot = f_save;
    fa.out_save = f_save;
    ga.balanced = fa.balanced = 0;
  }
  const bool balanced = (pb->tau_a == 1.0 && pb->tau_b == 1.0);
  HalfStepArgs g_plain = ga, g_err = ga, f_plain = fa, f_err = fa;
  g_plain.mode = kModeWritePot | kModeCheckDone;
  g_err.mode = kModeWritePot | kModeCheckDone | kModeSaveOld | kModeErrCol;
  f_plain.mode = kModeWritePot | kModeCheckDone;
  f_err.mode = kModeWritePot | kModeCheckDone | (balanced ? 0u : kModeErrRow);
  if (par_general) {
    f_plain.mode |= kModeSaveOld;
    f_err.mode = kModeWritePot | kModeCheckDone | kModeSaveOld | kModeErrRow | kModeErrOld;
  }

  if (symmetric) g_err.symmetric = 1;
  // With a schedule every launch works out its iteration t = block * inner +
  // it_offset on the device, and the g-update that ends a block measures the
  // error only once eps_t is the target; before that a pass at the target eps
  // (no potential written: h_old is g_t, saved by that g-update) measures it.
  // The g-update after the f-update of iteration it belongs to iteration it + 1.
  HalfStepArgs g_meas = ga;
  g_meas.mode = kModeCheckDone | kModeErrCol;
  g_meas.out_pot = g_save;
  g_meas.lse_shift = g_meas.lse_store = nullptr;
  g_meas.err_pass = 1;
  g_meas.it_offset = inner;
  // Unbalanced problems also need the row error at the target eps: a pass over
  // the rows of the state after the f-update of iteration it (f not written),
  // before the g-update that finalizes the block reads its partials.
  HalfStepArgs f_meas = fa;
  f_meas.mode = kModeCheckDone | kModeErrRow;
  f_meas.lse_shift = f_meas.lse_store = nullptr;
  f_meas.err_pass = 1;
  f_meas.it_offset = inner - 1;
  for (HalfStepArgs* a : {&g_plain, &g_err, &f_plain, &f_err, &g_meas, &f_meas}) {
    a->eps_sched = sched ? 1 : 0;
    a->eps_init = eps_init;
    a->eps_decay = eps_decay;
    a->mom_start = mom_start;
    a->mom_threshold = static_cast<float>(o->momentum_error_threshold);
  }
  // Recentred unbalanced iterations, in plain order: g-update terms, g
  // finished; f-update terms, f finished; at the block's last iteration the
  // column marginals of the new state and the recentred errors.
  RcScalars* rcs = rc_mode ? reinterpret_cast<RcScalars*>(base + L.rc) : nullptr;
  float* rc_app = rc_mode ? reinterpret_cast<float*>(base + L.rc_app) : nullptr;
  float* rc_col = rc_mode ? reinterpret_cast<float*>(base + L.rc_col) : nullptr;
  HalfStepArgs g_rc = ga, f_rc = fa, m_rc = ga;
  g_rc.mode = kModeWriteApp | kModeCheckDone;
  g_rc.out_vec = g_save;
  f_rc.mode = kModeWriteApp | kModeCheckDone;
  f_rc.out_vec = rc_app;
  m_rc.mode = kModeMarginal | kModeCheckDone;
  m_rc.out_vec = rc_col;
  RcFinishArgs rg{}, rf{};
  if (rc_mode) {
    const double ta = pb->tau_a, tb = pb->tau_b;
    auto kk = [&](double ti, double tj) {
      return -tj * (ta - 1.0) * (tb - 1.0) * (ti - 1.0) / ((tj - 1.0) * (ta * (tb - 1.0) + tb * (ta - 1.0)));
    };
    auto xi = [&](double ti, double tj) { return kk(ti, tj) / (1.0 - kk(ti, tj)); };
    for (RcFinishArgs* r : {&rg, &rf}) {
      r->sc = P.scal;
      r->momentum = static_cast<float>(o->momentum);
      r->mom_start = mom_start;
      r->inner = inner;
      r->mom_threshold = static_cast<float>(o->momentum_error_threshold);
      r->errors = out->errors;
      r->st = state;
    }
    rg.app = g_save;
    rg.pot = out->g;
    rg.w = P.b;
    rg.n = pb->m;
    rg.tau = tb;
    rg.tau_other = ta;
    rg.kc = kk(tb, tb);
    rg.xc = xi(tb, ta);
    rg.lse_other = &rcs->lse[0];
    rg.lse_self = &rcs->lse[1];
    rf.app = rc_app;
    rf.pot = out->f;
    rf.w = P.a;
    rf.n = pb->n;
    rf.tau = ta;
    rf.tau_other = tb;
    rf.kc = kk(ta, ta);
    rf.xc = xi(ta, tb);
    rf.lse_other = &rcs->lse[1];
    rf.lse_self = &rcs->lse[0];
  }
  auto launch_rc_block = [&](hipStream_t s) -> int {
    for (int it = 0; it < inner; ++it) {
      int r;
      RcFinishArgs rg_it = rg, rf_it = rf;
      rg_it.it_offset = rf_it.it_offset = it;
      if ((r = launch_halfstep(k, g_rc, s)) != OTTX_OK) return r;
      hipLaunchKernelGGL(rc_finish_kernel, dim3(1), dim3(kRcThreads), 0, s, rg_it);
      OTTX_LAUNCHED();
      if ((r = launch_halfstep(k, f_rc, s)) != OTTX_OK) return r;
      hipLaunchKernelGGL(rc_finish_kernel, dim3(1), dim3(kRcThreads), 0, s, rf_it);
      OTTX_LAUNCHED();
      if (it == inner - 1) {
        if ((r = launch_halfstep(k, m_rc, s)) != OTTX_OK) return r;
        hipLaunchKernelGGL(rc_error_kernel, dim3(1), dim3(kRcThreads), 0, s, ga, static_cast<const float*>(rc_col),
                           static_cast<const float*>(rc_app), static_cast<const float*>(out->f),
                           static_cast<const float*>(out->g), P.a, P.b, pb->n, pb->m, pb->tau_a, pb->tau_b,
                           static_cast<const SolveScalars*>(P.scal), static_cast<const RcScalars*>(rcs));
        OTTX_LAUNCHED();
      }
    }
    return OTTX_OK;
  };
  auto launch_block = [&](hipStream_t s) -> int {
    if (rc_mode) return launch_rc_block(s);
    for (int it = 0; it < inner; ++it) {
      int r;
      if (per_launch) {
        HalfStepArgs fa_it = it == inner - 1 ? f_err : f_plain;
        HalfStepArgs ga_it = it == inner - 1 ? g_err : g_plain;
        fa_it.it_offset = it;
        ga_it.it_offset = it + 1;
        if ((r = launch_halfstep(k, fa_it, s)) != OTTX_OK) return r;
        if (sched && !balanced && it == inner - 1 && (r = launch_halfstep(k, f_meas, s)) != OTTX_OK) return r;
        if ((r = launch_halfstep(k, ga_it, s)) != OTTX_OK) return r;
        continue;
      }
      if (symmetric) {
        // f <- g: the f-update of the parallel step is the g-update by symmetry
        hipLaunchKernelGGL(copy_potential_kernel, dim3(static_cast<unsigned>((pb->n + 255) / 256)), dim3(256), 0, s,
                           static_cast<const float*>(out->g), out->f, pb->n, static_cast<const SolveState*>(state));
        OTTX_LAUNCHED();
      } else {
        r = launch_halfstep(k, it == inner - 1 ? f_err : f_plain, s);
        if (r != OTTX_OK) return r;
      }
      r = launch_halfstep(k, it == inner - 1 ? g_err : g_plain, s);
      if (r != OTTX_OK) return r;
    }
    if (sched) {
      const int r = launch_halfstep(k, g_meas, s);
      if (r != OTTX_OK) return r;
    }
    if (adaptive && mom_start % inner == 0) {
      hipLaunchKernelGGL(momentum_fixup_kernel, dim3(static_cast<unsigned>((pb->m + 255) / 256)), dim3(256), 0, s,
                         out->g, static_cast<const float*>(g_save), pb->m, static_cast<const SolveState*>(state),
                         static_cast<const float*>(out->errors), mom_start, inner, static_cast<float>(o->momentum),
                         static_cast<float>(o->momentum_error_threshold));
      OTTX_LAUNCHED();
    }
    return OTTX_OK;
  };

  if (o->init_mode == OTTX_INIT_F_FROM_G) {
    // f <- tau_a (eps log a - eps LSE_j((g_j - C_ij) / eps)), a plain update
    HalfStepArgs f_init = fa;
    f_init.mode = kModeWritePot;
    f_init.momentum = 1.f;
    f_init.state = nullptr;
    rc = launch_halfstep(k, f_init, st);
    if (rc != OTTX_OK) return rc;
  }

  // The output finalizer writes the results (once) after the block that
  // stops the loop.  It is queued behind each of the first kEagerFinalize
  // blocks (and behind every block of a graph); a longer loop saves its
  // launches and the host queues it once it has seen the loop stop.
  constexpr int kEagerFinalize = 3;
  double* dual_part = reinterpret_cast<double*>(base + L.dual_part);
  auto launch_finalize = [&](hipStream_t s) -> int {
    if (rc_mode) {
      // the recentred outputs: f in place, g into the column-marginal buffer,
      // which the finalizer then copies to g
      hipLaunchKernelGGL(rc_output_kernel, dim3(1), dim3(kRcThreads), 0, s, out->f,
                         static_cast<const float*>(out->g), rc_col, pb->n, pb->m, pb->tau_a, pb->tau_b,
                         static_cast<const SolveScalars*>(P.scal), static_cast<const RcScalars*>(rcs),
                         static_cast<const SolveState*>(state));
      OTTX_LAUNCHED();
    }
    hipLaunchKernelGGL(solve_finalize_kernel, dim3(kDualBlocks, 3), dim3(256), 0, s,
                       par_general ? static_cast<const float*>(f_save) : out->f, P.a, pb->n,
                       rc_mode ? rc_col : g_save, out->g, P.b, pb->m, P.scal, pb->tau_a, pb->tau_b, out->errors,
                       outer, dual_part, state, mail, seq, recenter, out->f);
    OTTX_LAUNCHED();
    return OTTX_OK;
  };
  auto mail_ready = [&]() { return __atomic_load_n(&mail->ready, __ATOMIC_ACQUIRE) == seq; };
  // blocks completed by this solve, as the finalizers report them, and
  // whether the loop has stopped (the results then follow in the mailbox)
  auto progress = [&](bool* stopped) -> int64_t {
    const uint64_t v = __atomic_load_n(&mail->progress, __ATOMIC_ACQUIRE);
    if ((v >> 32) != seq) return -1;
    *stopped = (v & 0x80000000u) != 0;
    return static_cast<int64_t>(v & 0x7fffffffu);
  };
  if (total_blocks > 0) {
    if (rc_mode) {
      // log sum a exp(-f / rho_a) of the starting f, for the first g-update
      hipLaunchKernelGGL(rc_lse_kernel, dim3(1), dim3(kRcThreads), 0, st, static_cast<const float*>(out->f), P.a,
                         pb->n, pb->tau_a, static_cast<const SolveScalars*>(P.scal), &rcs->lse[0]);
      OTTX_LAUNCHED();
    } else {
      if (par_general) {
        rc = launch_halfstep(k, f_plain, st);  // f-update of iteration 0 (parallel)
        if (rc != OTTX_OK) return rc;
      }
      rc = launch_halfstep(k, g_plain, st);  // g-update of iteration 0
      if (rc != OTTX_OK) return rc;
    }
    bool fin_queued = false;  // an output finalizer follows the last block queued
    auto launch_all = [&](hipStream_t s, bool with_finalize) -> int {
      const int r = launch_block(s);
      fin_queued = with_finalize;
      return (r != OTTX_OK || !with_finalize) ? r : launch_finalize(s);
    };
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    hipStream_t cap = nullptr;
    // A graph pays for its capture and instantiation (about a millisecond)
    // only when a block has several launches that are short against the
    // host's launch cost: with inner_iterations = 1 (3 launches), or with
    // half-steps of more than 2^24 pairs (the device then runs behind the
    // host anyway), direct launches return sooner.
    if (o->use_graph && inner >= 4 && static_cast<double>(pb->n) * static_cast<double>(pb->m) <= 16777216.0) {
      OTTX_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
      OTTX_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
      rc = launch_all(cap, true);
      hipGraph_t g = nullptr;
      const hipError_t e1 = hipStreamEndCapture(cap, &g);
      if (rc != OTTX_OK || e1 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipStreamDestroy(cap);
        return rc != OTTX_OK ? rc : OTTX_EHIP;
      }
      graph = g;
      if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(graph);
        (void)hipStreamDestroy(cap);
        return OTTX_EHIP;
      }
    }
    int status = OTTX_OK;
    // Waits until `blocks` blocks have finished (or the loop has stopped),
    // spinning on the pinned mailbox that every block finalizer updates (no
    // event records in the stream, whose system-scope release would put a gap
    // between blocks, and no blocking wait with its wake-up latency).  The
    // stream itself is queried now and then, so an error or a drained stream
    // ends the wait.  Returns true when the loop has stopped.
    auto wait_blocks = [&](int blocks) -> bool {
      for (uint32_t spin = 1;; ++spin) {
        if (mail_ready()) return true;
        bool stopped = false;
        if (progress(&stopped) >= blocks) return stopped;
        if ((spin & 1023u) == 0) {
          const hipError_t q = hipStreamQuery(st);
          if (q == hipSuccess) return mail_ready();
          if (q != hipErrorNotReady) {
            status = OTTX_EHIP;
            return true;
          }
        }
      }
    };
    double err_prev = -1.0, err_prev2 = -1.0;
    for (int blk = 0; blk < total_blocks; ++blk) {
      if (exec) {
        if (hipGraphLaunch(exec, st) != hipSuccess) { status = OTTX_EHIP; break; }
        fin_queued = true;
      } else {
        status = launch_all(st, blk < kEagerFinalize);
        if (status != OTTX_OK) break;
      }
      if (o->progress_fn != nullptr) {
        // progress reporting: every block is waited for, so the mailbox's
        // error is this block's when the callback reads it
        const bool stopped = wait_blocks(blk + 1);
        if (status != OTTX_OK) break;
        bool st_bit = false;
        if (progress(&st_bit) >= blk + 1) {
          double e = 0.0;
          __atomic_load(&mail->last_err, &e, __ATOMIC_ACQUIRE);
          o->progress_fn(o->progress_user, blk, e);
        }
        if (stopped) break;
        continue;
      }
      if (blk == 0) continue;
      // One block runs ahead: wait for the previous block.
      if (wait_blocks(blk)) break;
      // Its error predicts the running block's (a geometric decrease, or a
      // sixteenfold one from a single error).  When that is below the threshold
      // the running block is likely the last: wait for it instead of queuing
      // a speculative block that would only exit early.
      err_prev2 = err_prev;
      __atomic_load(&mail->last_err, &err_prev, __ATOMIC_ACQUIRE);
      if (!force_scan && blk + 1 < total_blocks && blk + 1 >= (o->min_iterations + inner - 1) / inner &&
          std::isfinite(err_prev) && err_prev >= 0.0) {
        const double pred = (err_prev2 > 0.0) ? err_prev * (err_prev / err_prev2) : err_prev / 16.0;
        if (pred < o->threshold && wait_blocks(blk + 1)) break;
      }
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
    if (status != OTTX_OK) return status;
    // The loop stops at the latest in the last block queued; the output
    // finalizer behind it (queued now when none is) writes the mailbox.
    if (!mail_ready()) {
      if (!fin_queued && (rc = launch_finalize(st)) != OTTX_OK) return rc;
      for (uint32_t spin = 1; !mail_ready(); ++spin) {
        if ((spin & 1023u) == 0) {
          const hipError_t q = hipStreamQuery(st);
          if (q == hipSuccess) break;
          if (q != hipErrorNotReady) return OTTX_EHIP;
        }
      }
      if (!mail_ready()) return OTTX_EHIP;
    }
    const SolveMail* fm = mail;
    const double* du = fm->du;
    out->total_mass = fm->total_mass;
    out->epsilon = du[7];
    out->reg_ot_cost = du[0] + du[2] + du[7] * (du[1] * du[3] - fm->total_mass);
    const bool any_nan = du[4] > 0.0;
    const float last = static_cast<float>(du[6]);
    out->converged = (!any_nan && outer > 0 && last < static_cast<float>(o->threshold)) ? 1 : 0;
    out->n_iters = static_cast<int32_t>(du[5]) * inner;
    out->n_blocks = fm->n_blocks;
    return OTTX_OK;
  }

  // No iteration runs: the outputs are the initial potentials.
  HalfStepArgs mass = ga;
  mass.mode = kModeSaveOld | kModeErrCol;
  mass.state = nullptr;
  rc = launch_halfstep(k, mass, st);
  if (rc != OTTX_OK) return rc;
  double* dpart = reinterpret_cast<double*>(base + L.dual_part);
  hipLaunchKernelGGL(dual_partial_kernel, dim3(kDualBlocks, 3), dim3(256), 0, st, out->f, P.a, pb->n, g_save,
                     P.b, pb->m, P.scal, pb->tau_a, pb->tau_b, out->errors, outer, dpart);
  hipLaunchKernelGGL(dual_final_kernel, dim3(1), dim3(64), 0, st, dpart, out->errors, outer, P.scal, dual);
  OTTX_LAUNCHED();
  hipLaunchKernelGGL(sum_f64_kernel, dim3(1), dim3(256), 0, st, reinterpret_cast<const double*>(base + L.mass),
                     gp.tiles, dual + 8);
  OTTX_LAUNCHED();
  char* mb = static_cast<char*>(hh->mailbox);
  OTTX_HIP(hipMemcpyAsync(mb + 256, dual, 9 * sizeof(double), hipMemcpyDeviceToHost, st));
  OTTX_HIP(hipStreamSynchronize(st));
  const double* du = reinterpret_cast<const double*>(mb + 256);
  const double total = du[8];
  out->total_mass = total;
  out->epsilon = du[7];
  out->reg_ot_cost = du[0] + du[2] + du[7] * (du[1] * du[3] - total);
  out->converged = 0;
  out->n_iters = 0;
  out->n_blocks = 0;
  return OTTX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Transport applied to vectors (geometry.py:445-486): a column log-sum-exp
// pass of the half-step, then apply_pv_kernel for each block of 16 vectors.
// ---------------------------------------------------------------------------
typedef void (*ApplyFn)(const ApplyArgs);
struct ApplyCfg {
  ApplyFn fn;
  int nb, mt, w;
};

template <int PREC, int KK, int NB, int MT>
ApplyCfg apply_cfg() {
  return ApplyCfg{&apply_pv_kernel<PREC, KK, NB, MT, 4>, NB, MT, 4};
}

// One tile shape per fragment layout; wide features keep fewer column blocks
// and rows in registers.
bool select_apply(int prec, int kk, ApplyCfg* out) {
  if (prec != OTTX_PREC_FP32 && prec != OTTX_PREC_BF16_SPLIT) return false;  // the MFMA cross term only
  if (prec == OTTX_PREC_BF16_SPLIT) {
    switch (kk) {
      case 1: *out = apply_cfg<1, 1, 2, 2>(); return true;
      case 2: *out = apply_cfg<1, 2, 2, 2>(); return true;
      case 3: *out = apply_cfg<1, 3, 2, 2>(); return true;
      case 4: *out = apply_cfg<1, 4, 2, 2>(); return true;
      case 5: *out = apply_cfg<1, 5, 1, 1>(); return true;
      case 6: *out = apply_cfg<1, 6, 1, 1>(); return true;
      case 7: *out = apply_cfg<1, 7, 1, 1>(); return true;
      case 8: *out = apply_cfg<1, 8, 1, 1>(); return true;
      default: return false;
    }
  }
  switch (kk) {
    case 1: *out = apply_cfg<0, 1, 2, 2>(); return true;
    case 2: *out = apply_cfg<0, 2, 2, 2>(); return true;
    case 4: *out = apply_cfg<0, 4, 2, 2>(); return true;
    case 8: *out = apply_cfg<0, 8, 2, 2>(); return true;
    case 12: *out = apply_cfg<0, 12, 2, 2>(); return true;
    case 16: *out = apply_cfg<0, 16, 2, 2>(); return true;
    case 24: *out = apply_cfg<0, 24, 1, 2>(); return true;
    case 32: *out = apply_cfg<0, 32, 1, 2>(); return true;
    case 48: *out = apply_cfg<0, 48, 1, 1>(); return true;
    case 64: *out = apply_cfg<0, 64, 1, 1>(); return true;
    default: return false;
  }
}

// Fragments of up to 16 vectors over n rows (zero past n and past kvec):
// frag[(rb * 64 + 16 q + v) * 4 + r] = vecs[v][16 rb + 4 q + r].
__global__ void pack_vectors_kernel(const float* __restrict__ vecs, int64_t n, int kvec, int64_t n_pad,
                                    float* __restrict__ frag) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n_pad * 16) return;
  const int r = static_cast<int>(e & 3);
  const int l = static_cast<int>((e >> 2) & 63);
  const int64_t rb = e >> 8;
  const int v = l & 15;
  const int64_t i = rb * 16 + 4 * (l >> 4) + r;
  frag[e] = (v < kvec && i < n) ? vecs[static_cast<int64_t>(v) * n + i] : 0.f;
}

__global__ void zero_u32_kernel(uint32_t* __restrict__ p, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

struct ApplyLayout {
  ApplyCfg ac;
  PassPlan pv;
  size_t vfrag, opart, counters, total;
};

bool apply_layout(const ottx_problem* pb, int axis, const Plan& PL, ApplyLayout* AL) {
  if (!select_apply(PL.k.prec, PL.k.kk, &AL->ac)) return false;
  KernelCfg kpv = PL.k;
  kpv.nb = AL->ac.nb;
  kpv.mt = AL->ac.mt;
  kpv.w = AL->ac.w;
  const int64_t n_in = axis == 0 ? pb->n : pb->m, n_out = axis == 0 ? pb->m : pb->n;
  AL->pv = plan_pass(kpv, n_in, n_out, device_cus(), 2);
  size_t off = PL.L.total;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off = align_up(off + bytes);
    return at;
  };
  const int64_t n_in_pad = pad_rows(n_in), n_out_pad = pad_rows(n_out);
  AL->vfrag = take(static_cast<size_t>(n_in_pad) * 16 * sizeof(float));
  AL->opart = take(static_cast<size_t>(AL->pv.splits) * n_out_pad * 16 * sizeof(float));
  AL->counters = take(static_cast<size_t>(AL->pv.tiles) * sizeof(uint32_t));
  AL->total = off;
  return static_cast<int64_t>(AL->pv.tiles) * kpv.cols() == n_out_pad && AL->pv.n_in_rb * 16 == n_in_pad;
}

extern "C" size_t ottx_apply_workspace_bytes(const ottx_problem* pb, int32_t axis) {
  if (check_problem(pb) != OTTX_OK || (axis != 0 && axis != 1)) return 0;
  Plan PL;
  if (!make_plan(pb, -1, &PL)) return 0;
  ApplyLayout AL;
  if (!apply_layout(pb, axis, PL, &AL)) return 0;
  return AL.total;
}

extern "C" int ottx_pointcloud_apply(const ottx_problem* pb, int32_t axis, const float* f, const float* g,
                                     const float* vecs, int32_t kvec, double eps, int32_t flags, float* out,
                                     void* workspace, size_t workspace_bytes, void* stream) {
  int rc = check_problem(pb);
  if (rc != OTTX_OK) return rc;
  if (!f || !g || !vecs || !out || kvec < 1 || !(eps > 0.0) || (axis != 0 && axis != 1)) return OTTX_EINVAL;
  if (!workspace || (reinterpret_cast<uintptr_t>(workspace) & 255)) return OTTX_EWORKSPACE;
  Plan PL;
  if (!make_plan(pb, -1, &PL)) return OTTX_EUNSUPPORTED;
  ApplyLayout AL;
  if (!apply_layout(pb, axis, PL, &AL)) return OTTX_EUNSUPPORTED;
  if (workspace_bytes < AL.total) return OTTX_EWORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const KernelCfg& k = PL.k;
  const Layout& L = PL.L;
  Prepared P;
  if (flags & OTTX_APPLY_PREPARED) {
    prepared_view(pb, L, workspace, &P);
  } else {
    rc = prepare(pb, k, L, workspace, st, &P);
    if (rc != OTTX_OK) return rc;
  }
  char* base = P.base;
  uint32_t* counters = reinterpret_cast<uint32_t*>(base + L.counters);
  uint32_t* pv_counters = reinterpret_cast<uint32_t*>(base + AL.counters);
  const int max_tiles = std::max(PL.gp.tiles, PL.fp.tiles);
  hipLaunchKernelGGL(zero_u32_kernel, dim3(static_cast<unsigned>((max_tiles + 255) / 256)), dim3(256), 0, st,
                     counters, static_cast<int64_t>(max_tiles));
  hipLaunchKernelGGL(zero_u32_kernel, dim3(static_cast<unsigned>((AL.pv.tiles + 255) / 256)), dim3(256), 0, st,
                     pv_counters, static_cast<int64_t>(AL.pv.tiles));
  OTTX_LAUNCHED();
  // the column log-sum-exp in exponent units (no epilogue output)
  const bool g_upd = axis == 0;
  const float* h = g_upd ? g : f;
  const float* in_pot = g_upd ? f : g;
  float* lse = reinterpret_cast<float*>(base + (g_upd ? L.lse_y : L.lse_x));
  HalfStepArgs a = base_args(pb, k, g_upd ? PL.gp : PL.fp, P, g_upd ? 1 : 0, const_cast<float*>(in_pot),
                             const_cast<float*>(h), eps);
  a.state = nullptr;
  a.scal = nullptr;
  a.mode = 0;
  a.lse_shift = nullptr;
  a.lse_store = lse;
  if (!(flags & OTTX_APPLY_LSE_READY)) {
    rc = launch_halfstep(k, a, st);
    if (rc != OTTX_OK) return rc;
  }
  // P V, 16 vectors per launch
  const int64_t n_in = g_upd ? pb->n : pb->m, n_out = g_upd ? pb->m : pb->n;
  const int64_t n_in_pad = pad_rows(n_in);
  ApplyArgs A{};
  A.h = a;
  A.h.tiles = AL.pv.tiles;
  A.h.splits = AL.pv.splits;
  A.h.rb_per_split = AL.pv.rb_per_split;
  A.h.n_in_rb = AL.pv.n_in_rb;
  A.h.lse_shift = lse;
  A.h.lse_store = nullptr;
  A.h.tile_counter = pv_counters;
  float* vfrag = reinterpret_cast<float*>(base + AL.vfrag);
  A.v_frag = vfrag;
  A.o_part = reinterpret_cast<float*>(base + AL.opart);
  A.out_stride = n_out;
  for (int v0 = 0; v0 < kvec; v0 += 16) {
    const int kb = std::min(16, kvec - v0);
    hipLaunchKernelGGL(pack_vectors_kernel, dim3(static_cast<unsigned>((n_in_pad * 16 + 255) / 256)), dim3(256), 0,
                       st, vecs + static_cast<int64_t>(v0) * n_in, n_in, kb, n_in_pad, vfrag);
    A.out = out + static_cast<int64_t>(v0) * n_out;
    A.kvec = kb;
    hipLaunchKernelGGL(AL.ac.fn, dim3(AL.pv.tiles, AL.pv.splits), dim3(64 * AL.ac.w), 0, st, A);
    OTTX_LAUNCHED();
  }
  return OTTX_OK;
}

extern "C" int ottx_tile_configs(int32_t precision, int32_t d, int32_t* desc, int32_t max_configs) {
  KernelCfg v[8];
  const int n = kernel_variants(precision, d, v, 8);
  if (n <= 0) return 0;
  int out = 0;
  for (int vi = 0; vi < n; ++vi)
    for (int ti = 0; ti < kNumTargets; ++ti) {
      if (desc && out < max_configs) {
        int32_t* e = desc + 5 * out;
        e[0] = vi | (ti << 8);
        e[1] = v[vi].nb;
        e[2] = v[vi].mt;
        e[3] = v[vi].w;
        e[4] = kTargets[ti];
      }
      ++out;
    }
  return out;
}

extern "C" int ottx_benchmark_halfstep(const ottx_problem* pb, int32_t axis, int32_t tile_config, int32_t reps,
                                       float* f, float* g, double* avg_ms_host, void* workspace,
                                       size_t workspace_bytes, void* stream) {
  int rc = check_problem(pb);
  if (rc != OTTX_OK) return rc;
  if ((axis != 0 && axis != 1) || reps < 1 || !f || !g || !avg_ms_host || !workspace) return OTTX_EINVAL;
  Plan PL;
  if (!make_plan(pb, tile_config, &PL)) return tile_config >= 0 ? OTTX_EINVAL : OTTX_EUNSUPPORTED;
  const KernelCfg& k = PL.k;
  const PassPlan& gp = PL.gp;
  const PassPlan& fp = PL.fp;
  const Layout& L = PL.L;
  if (workspace_bytes < L.total) return OTTX_EWORKSPACE;
  if (reinterpret_cast<uintptr_t>(workspace) & 255) return OTTX_EWORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  Prepared P;
  rc = prepare(pb, k, L, workspace, st, &P);
  if (rc != OTTX_OK) return rc;
  uint32_t* counters = reinterpret_cast<uint32_t*>(P.base + L.counters);
  SolveState* state = reinterpret_cast<SolveState*>(P.base + L.state);
  float* dummy_err = reinterpret_cast<float*>(P.base + L.dual);
  hipLaunchKernelGGL(init_state_kernel, dim3(1), dim3(256), 0, st, state, dummy_err, 0, counters,
                     std::max(gp.tiles, fp.tiles), 0, static_cast<float*>(nullptr), static_cast<const float*>(nullptr),
                     int64_t{0}, static_cast<float*>(nullptr), static_cast<const float*>(nullptr), int64_t{0});
  OTTX_LAUNCHED();
  HalfStepArgs a = axis == 0 ? base_args(pb, k, gp, P, 1, f, g, pb->epsilon)
                             : base_args(pb, k, fp, P, 0, g, f, pb->epsilon);
  a.lse_shift = a.lse_store = reinterpret_cast<float*>(P.base + (axis == 0 ? L.lse_y : L.lse_x));
  a.mode = kModeWritePot;
  for (int w = 0; w < 2; ++w) {
    rc = launch_halfstep(k, a, st);
    if (rc != OTTX_OK) return rc;
  }
  hipEvent_t e0, e1;
  OTTX_HIP(hipEventCreate(&e0));
  OTTX_HIP(hipEventCreate(&e1));
  OTTX_HIP(hipEventRecord(e0, st));
  for (int r = 0; r < reps; ++r) {
    rc = launch_halfstep(k, a, st);
    if (rc != OTTX_OK) break;
  }
  OTTX_HIP(hipEventRecord(e1, st));
  OTTX_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  const hipError_t e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != OTTX_OK) return rc;
  if (e != hipSuccess) return OTTX_EHIP;
  *avg_ms_host = static_cast<double>(ms) / reps;
  return OTTX_OK;
}

// ===========================================================================
// Batched solves: many independent problems in shared launches.
// Reference: tools/segment_sinkhorn.py:24-138 / geometry/segment.py:23-190,
// which pad every problem to one size and vmap the solve; here every problem
// keeps its own size, epsilon, loop state and errors, and each half-step is a
// single launch over a work table of (problem, column tile, row split).
// ===========================================================================
namespace ottx {

// Per-problem device view of one batched problem (pointers into the
// problem's inputs and its own region of the workspace).
struct BatchItem {
  const float* x;
  const float* y;
  const float* a;
  const float* b;
  float* a_fill;  // OTTX_BATCH_UNIFORM_A: a, written with 1 / n by the init kernel (else null)
  float* b_fill;
  float* f;
  float* g;
  float* g_save;
  float* errors;
  int64_t n, m, n_pad, m_pad;
  double n_div, m_div;   // row counts the mean-cost statistics divide by
  double eps_abs;        // absolute epsilon (relative epsilon when 0)
  void* frag_x;
  void* frag_y;
  float* norm_x;
  float* norm_y;
  float* mu;
  SolveScalars* scal;
  SolveState* state;
  uint32_t* counters;
  double* dual_part;
  int32_t n_counters;
  int32_t pad_;
};

// Statistics of every problem in two stages.  Stage 1, grid (chunks, K, 2):
// float64 per-feature sums of x and x^2 over 256-row chunks of one cloud
// (chunk_sums' fixed order), so that the largest problem is spread over many
// workgroups.  part[((k * 2 + side) * chunks + c) * d + feature][2].
constexpr int kBatchStatRows = 256;

__global__ __launch_bounds__(256) void batch_stats_partial_kernel(const BatchItem* __restrict__ items, int d,
                                                                  int chunks, double* __restrict__ part) {
  __shared__ double s1[256], s2[256];
  const BatchItem& it = items[blockIdx.y];
  const int side = blockIdx.z, c = blockIdx.x, t = threadIdx.x;
  const int64_t np = side ? it.m : it.n;
  const int64_t r0 = static_cast<int64_t>(c) * kBatchStatRows;
  if (r0 >= np) return;  // the whole workgroup
  const int64_t r1 = min(np, r0 + kBatchStatRows);
  const int G = max(1, 256 / d);
  double a1 = 0.0, a2 = 0.0;
  if (t < G * d) chunk_sums(side ? it.y : it.x, r0, r1, d, G, t, a1, a2);
  s1[t] = a1;
  s2[t] = a2;
  __syncthreads();
  if (t < d) {
    double b1 = 0.0, b2 = 0.0;
    for (int g = 0; g < G; ++g) {
      b1 += s1[g * d + t];
      b2 += s2[g * d + t];
    }
    double* q = part + ((static_cast<int64_t>(blockIdx.y) * 2 + side) * chunks + c) * d * 2 + t * 2;
    q[0] = b1;
    q[1] = b2;
  }
}

// Stage 2, one workgroup per problem: the chunk partials summed in order, the
// centring shift, mean(C) over (n_div, m_div) rows and the solve scalars
// (pointcloud.py:170-186, geometry.py:147-171).
__global__ __launch_bounds__(256) void batch_stats_final_kernel(const BatchItem* __restrict__ items, int d,
                                                                int chunks, const double* __restrict__ part,
                                                                double s, double rel, double tau_a, double tau_b) {
  __shared__ double acc3[3][256];
  const BatchItem& it = items[blockIdx.x];
  const int t = threadIdx.x;
  double nx2 = 0.0, ny2 = 0.0, dot = 0.0;
  if (t < d) {
    double a1 = 0.0, a2 = 0.0, b1 = 0.0, b2 = 0.0;
    const int cx = static_cast<int>((it.n + kBatchStatRows - 1) / kBatchStatRows);
    const int cy = static_cast<int>((it.m + kBatchStatRows - 1) / kBatchStatRows);
    const double* px = part + (static_cast<int64_t>(blockIdx.x) * 2 + 0) * chunks * d * 2;
    const double* py = part + (static_cast<int64_t>(blockIdx.x) * 2 + 1) * chunks * d * 2;
    for (int c = 0; c < cx; ++c) {
      a1 += px[(c * d + t) * 2 + 0];
      a2 += px[(c * d + t) * 2 + 1];
    }
    for (int c = 0; c < cy; ++c) {
      b1 += py[(c * d + t) * 2 + 0];
      b2 += py[(c * d + t) * 2 + 1];
    }
    const double shift = 0.5 * (a1 / static_cast<double>(it.n) + b1 / static_cast<double>(it.m));
    it.mu[t] = isfinite(shift) ? static_cast<float>(shift) : 0.f;  // as stats_final_kernel
    nx2 = a2;
    ny2 = b2;
    dot = (a1 / it.n_div) * (b1 / it.m_div);
  }
  acc3[0][t] = nx2;
  acc3[1][t] = ny2;
  acc3[2][t] = dot;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st)
      for (int q = 0; q < 3; ++q) acc3[q][t] += acc3[q][t + st];
    __syncthreads();
  }
  if (t == 0) {
    const double mean_c = s * (acc3[0][0] / it.n_div + acc3[1][0] / it.m_div - 2.0 * acc3[2][0]);
    const double eps = rel > 0.0 ? rel * mean_c : it.eps_abs;
    SolveScalars* sc = it.scal;
    sc->eps = eps;
    sc->eps_f = static_cast<float>(eps);
    sc->c = static_cast<float>(2.0 * s * 1.4426950408889634 / eps);
    sc->inv_rho_a = tau_a == 1.0 ? 0.f : static_cast<float>((1.0 - tau_a) / (eps * tau_a));
    sc->inv_rho_b = tau_b == 1.0 ? 0.f : static_cast<float>((1.0 - tau_b) / (eps * tau_b));
  }
}

// Fragments and norms of every problem in one launch over a work table of
// (problem, side, block, kind): kind 0 packs 256 fragment threads, kind 1
// computes 256 norms (only the blocks a problem needs are launched).
__global__ __launch_bounds__(256) void batch_prep_kernel(const BatchItem* __restrict__ items,
                                                         const int4* __restrict__ table, int d, int chunks,
                                                         int prec, double s) {
  __shared__ float s_mu[256];
  const int4 w = table[blockIdx.x];
  const BatchItem& it = items[w.x];
  const int side = w.y;
  const int64_t np = side ? it.m : it.n;
  const int64_t np_pad = side ? it.m_pad : it.n_pad;
  const float* pts = side ? it.y : it.x;
  if (w.w == 0) {
    const int64_t gid = static_cast<int64_t>(w.z) * 256 + threadIdx.x;
    if (gid >= (np_pad / 16) * chunks * 64) return;
    void* frag = side ? it.frag_y : it.frag_x;
    const float scale = side ? it.scal->c : 1.f;
    if (prec == 1)
      pack_bf16_one(pts, np, d, it.mu, scale, chunks, gid, static_cast<uint4*>(frag));
    else
      pack_f32_one(pts, np, d, it.mu, scale, chunks, gid, static_cast<float4*>(frag));
    return;
  }
  for (int k = threadIdx.x; k < d; k += 256) s_mu[k] = it.mu[k];
  __syncthreads();
  const int64_t i = static_cast<int64_t>(w.z) * 256 + threadIdx.x;
  if (i >= np_pad) return;
  float* norm = side ? it.norm_y : it.norm_x;
  if (i >= np) {
    norm[i] = 0.f;
    return;
  }
  const float* row = pts + i * d;
  double acc = 0.0;
  if ((d & 3) == 0 && (reinterpret_cast<uintptr_t>(row) & 15) == 0) {
    for (int k = 0; k < d; k += 16) {
      float4 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        v[u] = (k + 4 * u < d) ? *reinterpret_cast<const float4*>(row + k + 4 * u) : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (k + 4 * u >= d) continue;
        const double e0 = v[u].x - s_mu[k + 4 * u], e1 = v[u].y - s_mu[k + 4 * u + 1];
        const double e2 = v[u].z - s_mu[k + 4 * u + 2], e3 = v[u].w - s_mu[k + 4 * u + 3];
        acc += (e0 * e0 + e1 * e1) + (e2 * e2 + e3 * e3);
      }
    }
  } else {
    for (int k = 0; k < d; ++k) {
      const double v = static_cast<double>(row[k] - s_mu[k]);
      acc += v * v;
    }
  }
  norm[i] = static_cast<float>(s * acc);
}

// Progress of a batched solve for the host, written into pinned memory after
// every block: the number of finished problems, then (seq << 32) | blocks.
struct BatchMail {
  uint64_t progress;
  uint32_t done;
  uint32_t live[2];  // live work items of the g- and f-updates after the block
};

// End of a block.  When problems have finished since the last rebuild, the
// live work lists are rebuilt from the full ones without the items of
// finished problems, keeping their longest-first order (a stable compaction:
// wave ballots and a scan of the per-wave counts, 8192 items per round, eight
// per thread).  Then the
// progress (finished problems, live items, completed blocks) goes to the
// mailbox.  counters: batch_done (kBatchDoneBits).
constexpr int kCompactThreads = 1024;
constexpr int kCompactPer = 8;  // items per thread per round of loads

__global__ __launch_bounds__(kCompactThreads) void batch_progress_kernel(
    uint32_t* counters, BatchMail* mail, uint32_t seq, int K,
    const int4* __restrict__ work_g, const int4* __restrict__ work_f, int4* __restrict__ live_g,
    int4* __restrict__ live_f, uint32_t* __restrict__ live_n, int n_g, int n_f) {
  constexpr int kWaves = kCompactThreads / 64;
  static_assert(kCompactPer * kWaves == 128, "one wave scans the round's counts, two per lane");
  __shared__ uint32_t s_cnt[kCompactPer * kWaves];  // kept items per (sub-round, wave), list order
  __shared__ uint32_t s_off[kCompactPer * kWaves];  // where each (sub-round, wave) writes
  __shared__ uint32_t s_total;
  __shared__ uint32_t s_done[65536 / 32];  // finished problems, one bit each (K <= 65535)
  const int tid = threadIdx.x;
  const uint32_t dn = __hip_atomic_load(&counters[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (dn != counters[2]) {
    for (int i = tid; i < (K + 31) / 32; i += kCompactThreads) s_done[i] = counters[kBatchDoneBits + i];
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6;
    for (int dir = 0; dir < 2; ++dir) {
      const int4* __restrict__ src = dir ? work_f : work_g;
      int4* __restrict__ dst = dir ? live_f : live_g;
      const int n = dir ? n_f : n_g;
      uint32_t base = 0;
      for (int c0 = 0; c0 < n; c0 += kCompactThreads * kCompactPer) {
        int4 w[kCompactPer];
        bool keep[kCompactPer];
        uint32_t rank[kCompactPer];  // kept items before this one in its wave
#pragma unroll
        for (int u = 0; u < kCompactPer; ++u) {
          const int i = c0 + u * kCompactThreads + tid;
          w[u] = i < n ? src[i] : make_int4(0, 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < kCompactPer; ++u) {
          const int i = c0 + u * kCompactThreads + tid;
          keep[u] = i < n && ((s_done[w[u].x >> 5] >> (w[u].x & 31)) & 1u) == 0u;
          const uint64_t mask = __ballot(keep[u]);
          if (lane == 0) s_cnt[u * kWaves + wave] = static_cast<uint32_t>(__popcll(mask));
          rank[u] = static_cast<uint32_t>(__popcll(mask & ((uint64_t{1} << lane) - 1u)));
        }
        __syncthreads();
        if (wave == 0) {
          // exclusive scan of the 128 counts in list order, two per lane
          const uint32_t a = s_cnt[2 * lane], b = s_cnt[2 * lane + 1];
          uint32_t incl = a + b;
#pragma unroll
          for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
          }
          const uint32_t excl = base + incl - (a + b);
          s_off[2 * lane] = excl;
          s_off[2 * lane + 1] = excl + a;
          if (lane == 63) s_total = incl;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kCompactPer; ++u)
          if (keep[u]) dst[s_off[u * kWaves + wave] + rank[u]] = w[u];
        base += s_total;
        __syncthreads();  // s_cnt, s_off and s_total are reused by the next round
      }
      if (tid == 0) live_n[dir] = base;
    }
    if (tid == 0) counters[2] = dn;
  }
  if (tid != 0) return;
  const uint32_t blocks = counters[1] + 1u;
  counters[1] = blocks;
  __hip_atomic_store(&mail->live[0], live_n[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&mail->live[1], live_n[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&mail->done, dn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&mail->progress, (static_cast<uint64_t>(seq) << 32) | blocks, __ATOMIC_RELEASE,
                     __HIP_MEMORY_SCOPE_SYSTEM);
}

// Loop state of every problem (sinkhorn.py:964-972) and the batch counters.
__global__ void batch_init_kernel(const BatchItem* __restrict__ items, int outer, uint32_t* batch_done,
                                  int zero_init) {
  const BatchItem& it = items[blockIdx.x];
  const int t = threadIdx.x;
  if (t == 0) {
    SolveState* st = it.state;
    st->done = 0;
    st->block = 0;
    st->n_blocks = 0;
    st->finalized = 0;
    st->tiles_done = 0u;
    st->shift_ok[0] = 0;
    st->shift_ok[1] = 0;
    st->fin_count = 0u;
    st->total_mass = 0.0;
    st->last_err = 0.0;
    if (blockIdx.x == 0) {
      batch_done[0] = 0u;  // finished problems
      batch_done[1] = 0u;  // completed blocks
      batch_done[2] = 0u;  // finished problems at the last rebuild of the live work lists
    }
    if ((blockIdx.x & 31) == 0) batch_done[kBatchDoneBits + (blockIdx.x >> 5)] = 0u;
  }
  for (int k = t; k < outer; k += blockDim.x) it.errors[k] = -1.f;
  for (int k = t; k < it.n_counters; k += blockDim.x) it.counters[k] = 0u;
  // default weights (segment.py:100-104), float64 1 / n rounded to float32
  if (it.a_fill != nullptr) {
    const float w = static_cast<float>(1.0 / static_cast<double>(it.n));
    for (int64_t k = t; k < it.n; k += blockDim.x) it.a_fill[k] = w;
  }
  if (it.b_fill != nullptr) {
    const float w = static_cast<float>(1.0 / static_cast<double>(it.m));
    for (int64_t k = t; k < it.m; k += blockDim.x) it.b_fill[k] = w;
  }
  // DefaultInitializer: 0, or -inf where the weight is 0 (bit 0: f, bit 1: g);
  // each weight written above is read back by the thread that wrote it
  if (zero_init & 1)
    for (int64_t k = t; k < it.n; k += blockDim.x) it.f[k] = it.a[k] > 0.f ? 0.f : -INFINITY;
  if (zero_init & 2)
    for (int64_t k = t; k < it.m; k += blockDim.x) it.g[k] = it.b[k] > 0.f ? 0.f : -INFINITY;
}

// Stage 1 of the float64 dual terms of every problem: grid (nblk, 3, K),
// nblk <= kDualBlocks blocks per side, as dual_partial_kernel for one problem.
__global__ void batch_dual_partial_kernel(const BatchItem* __restrict__ items, double tau_a, double tau_b,
                                          int outer, int nblk) {
  __shared__ double red[2][256];
  const BatchItem& it = items[blockIdx.z];
  const int tid = threadIdx.x;
  const int side = blockIdx.y;
  double s0 = 0.0, s1 = 0.0;
  if (side < 2) {
    const float* pot = side == 0 ? it.f : it.g_save;
    const float* w = side == 0 ? it.a : it.b;
    const int64_t len = side == 0 ? it.n : it.m;
    const double tau = side == 0 ? tau_a : tau_b;
    const double eps = it.scal->eps;
    const double rho = eps * tau / (1.0 - tau);
    const int64_t per = (len + nblk - 1) / nblk;
    const int64_t i0 = blockIdx.x * per, i1 = min(len, i0 + per);
    for (int64_t i = i0 + tid; i < i1; i += blockDim.x) {
      const double wi = w[i];
      s1 += wi;
      if (wi > 0.0) {
        const double h = static_cast<double>(pot[i]) - eps * log(wi);
        if (tau == 1.0) {
          s0 += wi * h;
        } else {
          s0 -= wi * rho * (exp(-h / rho) - 1.0);
        }
      }
    }
  } else if (blockIdx.x == 0) {
    for (int i = tid; i < outer; i += blockDim.x) {
      const float e = it.errors[i];
      if (e != e) s0 += 1.0;
      if (e != -1.f) s1 += 1.0;
    }
  }
  red[0][tid] = s0;
  red[1][tid] = s1;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) {
      red[0][tid] += red[0][tid + st];
      red[1][tid] += red[1][tid + st];
    }
    __syncthreads();
  }
  if (tid == 0) {
    it.dual_part[(side * kDualBlocks + blockIdx.x) * 2 + 0] = red[0][0];
    it.dual_part[(side * kDualBlocks + blockIdx.x) * 2 + 1] = red[1][0];
  }
}

// Stage 2: out[16 k + ...] = div_a, sum a, div_b, sum b, #NaN errors,
// #(errors != -1), errors[-1], eps, total mass, blocks run.  Grid K.
__global__ void batch_dual_final_kernel(const BatchItem* __restrict__ items, int outer, int nblk,
                                        double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  const BatchItem& it = items[blockIdx.x];
  double* o = out + 16 * static_cast<int64_t>(blockIdx.x);
  for (int side = 0; side < 3; ++side) {
    double s0 = 0.0, s1 = 0.0;
    for (int b = 0; b < nblk; ++b) {
      s0 += it.dual_part[(side * kDualBlocks + b) * 2 + 0];
      s1 += it.dual_part[(side * kDualBlocks + b) * 2 + 1];
    }
    o[side * 2 + 0] = s0;
    o[side * 2 + 1] = s1;
  }
  o[6] = outer > 0 ? static_cast<double>(it.errors[outer - 1]) : 0.0;
  o[7] = it.scal->eps;
  o[8] = it.state->total_mass;
  o[9] = static_cast<double>(it.state->n_blocks);
}

// g <- g_save for every problem: grid (x blocks, K).
__global__ void batch_copy_g_kernel(const BatchItem* __restrict__ items) {
  const BatchItem& it = items[blockIdx.y];
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < it.m) it.g[i] = it.g_save[i];
}

// Host plan of a batched solve: one kernel configuration for all problems,
// each problem padded to a multiple of ``pad`` rows and given its own layout.
struct BatchPlan {
  KernelCfg k;
  int64_t pad = 0;
  int K = 0;
  int outer = 0;
  std::vector<ottx_problem> pbs;
  std::vector<PassPlan> gp, fp;
  std::vector<Layout> L;
  std::vector<size_t> base;  // offset of each problem's layout
  size_t off_done = 0, off_items = 0, off_args_g = 0, off_args_f = 0, off_work_g = 0, off_work_f = 0,
         off_live_g = 0, off_live_f = 0, off_live_n = 0, off_out = 0, off_stats = 0, off_prep = 0, total = 0;
  int stat_chunks = 0;
  int64_t n_prep = 0;  // (problem, side, block, kind) items of the fragment / norm preparation
  int64_t n_work_g = 0, n_work_f = 0;
  int64_t max_np = 0;
};

int64_t round_up(int64_t v, int64_t q) { return (v + q - 1) / q * q; }

// Rows of the input side one workgroup streams: sized so the whole batch
// makes about ``target`` workgroups per half-step direction.
int batch_rows_per_split(const KernelCfg& k, const BatchPlan& B, bool g_dir, int target) {
  double pairs = 0.0;
  int64_t max_in = 0;
  for (int i = 0; i < B.K; ++i) {
    // streamed rows are padded to the split unit, columns to the tile
    const int64_t n_in = pad_rows(g_dir ? B.pbs[i].n : B.pbs[i].m, 32 * k.mt);
    const int64_t n_out = pad_rows(g_dir ? B.pbs[i].m : B.pbs[i].n, k.cols());
    pairs += static_cast<double>(n_in) * static_cast<double>(n_out);
    max_in = std::max(max_in, n_in);
  }
  const int64_t unit = 32 * k.mt;  // an even number of streamed tiles per split
  int64_t rows = static_cast<int64_t>(pairs / (static_cast<double>(k.cols()) * target));
  rows = std::max<int64_t>(unit, round_up(std::max<int64_t>(rows, 1), unit));
  return static_cast<int>(std::min<int64_t>(rows, round_up(max_in, unit)));
}

int check_batch(const ottx_batch* bt) {
  // count <= 65535: problems index the y and z grid dimensions of the batched kernels
  if (!bt || bt->count < 1 || bt->count > 65535 || bt->d < 1 || bt->d > 256) return OTTX_EINVAL;
  if (!bt->x || !bt->y || !bt->a || !bt->b || !bt->x_off || !bt->y_off) return OTTX_EINVAL;
  if (bt->flags & ~(OTTX_BATCH_UNIFORM_A | OTTX_BATCH_UNIFORM_B)) return OTTX_EINVAL;
  if (bt->precision != OTTX_PREC_FP32 && bt->precision != OTTX_PREC_BF16_SPLIT) return OTTX_EINVAL;
  if (!(bt->inv_scale_cost > 0.0) || bt->relative_epsilon < 0.0 || bt->mean_rows_x < 0.0 || bt->mean_rows_y < 0.0)
    return OTTX_EINVAL;
  if (!(bt->relative_epsilon > 0.0) && !bt->epsilon) return OTTX_EINVAL;
  if (!(bt->tau_a > 0.0 && bt->tau_a <= 1.0) || !(bt->tau_b > 0.0 && bt->tau_b <= 1.0)) return OTTX_EINVAL;
  if (bt->x_off[0] != 0 || bt->y_off[0] != 0) return OTTX_EINVAL;
  for (int i = 0; i < bt->count; ++i) {
    const int64_t n = bt->x_off[i + 1] - bt->x_off[i], m = bt->y_off[i + 1] - bt->y_off[i];
    if (n < 1 || m < 1 || n > (int64_t{1} << 30) || m > (int64_t{1} << 30)) return OTTX_EINVAL;
    if ((bt->mean_rows_x > 0.0 && bt->mean_rows_x < static_cast<double>(n)) ||
        (bt->mean_rows_y > 0.0 && bt->mean_rows_y < static_cast<double>(m)))
      return OTTX_EINVAL;
    if (!(bt->relative_epsilon > 0.0) && !(bt->epsilon[i] > 0.0)) return OTTX_EINVAL;
  }
  return OTTX_OK;
}

int check_opts(const ottx_sinkhorn_opts* o) {
  if (!o) return OTTX_EINVAL;
  if (o->inner_iterations < 1 || o->max_iterations < 0 || o->min_iterations < 0 || o->norm_error < 1)
    return OTTX_EINVAL;
  if (!(o->momentum > 0.0 && o->momentum < 2.0)) return OTTX_EINVAL;
  return OTTX_OK;
}

// Kernel choice for a batch: the variant with the smallest row padding (the
// least waste on small problems) unless ``tile_config`` names one.
bool batch_kernel(int prec, int d, int tile_config, KernelCfg* k, int* target_mult) {
  KernelCfg v[8];
  const int n = kernel_variants(prec, d, v, 8);
  if (n <= 0) return false;
  if (tile_config >= 0) return select_kernel(prec, d, tile_config, k, target_mult);
  int best = 0;
  for (int i = 1; i < n; ++i) {
    const int pi = std::max(v[i].cols(), 32 * v[i].mt), pb = std::max(v[best].cols(), 32 * v[best].mt);
    if (pi < pb) best = i;
  }
  *k = v[best];
  *target_mult = 8;
  return true;
}

int make_batch_plan(const ottx_batch* bt, const ottx_sinkhorn_opts* o, BatchPlan* B) {
  int tm = 8;
  if (!batch_kernel(bt->precision, bt->d, o->tile_config, &B->k, &tm))
    return o->tile_config >= 0 ? OTTX_EINVAL : OTTX_EUNSUPPORTED;
  const KernelCfg& k = B->k;
  B->pad = std::max<int64_t>(k.cols(), 32 * k.mt);
  B->K = bt->count;
  B->outer = (o->max_iterations + o->inner_iterations - 1) / o->inner_iterations;
  B->pbs.resize(B->K);
  for (int i = 0; i < B->K; ++i) {
    ottx_problem& p = B->pbs[i];
    p = ottx_problem{};
    p.n = bt->x_off[i + 1] - bt->x_off[i];
    p.m = bt->y_off[i + 1] - bt->y_off[i];
    p.d = bt->d;
    p.precision = bt->precision;
    p.x = bt->x + bt->x_off[i] * bt->d;
    p.y = bt->y + bt->y_off[i] * bt->d;
    p.a = bt->a + bt->x_off[i];
    p.b = bt->b + bt->y_off[i];
    p.epsilon = bt->epsilon ? bt->epsilon[i] : 0.0;
    p.relative_epsilon = bt->relative_epsilon;
    p.inv_scale_cost = bt->inv_scale_cost;
    p.tau_a = bt->tau_a;
    p.tau_b = bt->tau_b;
    B->max_np = std::max(B->max_np, std::max(p.n, p.m));
  }
  const int target = tm * device_cus();
  const int rows_g = batch_rows_per_split(k, *B, true, target);
  const int rows_f = batch_rows_per_split(k, *B, false, target);
  B->gp.resize(B->K);
  B->fp.resize(B->K);
  B->L.resize(B->K);
  B->base.resize(B->K);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off = align_up(off + bytes);
    return at;
  };
  B->off_done = take(sizeof(uint32_t) * (kBatchDoneBits + (B->K + 31) / 32));
  B->off_items = take(sizeof(BatchItem) * B->K);
  B->off_args_g = take(sizeof(HalfStepArgs) * B->K);
  B->off_args_f = take(sizeof(HalfStepArgs) * B->K);
  B->n_work_g = B->n_work_f = 0;
  for (int i = 0; i < B->K; ++i) {
    for (int dir = 0; dir < 2; ++dir) {
      const int64_t n_in = dir == 0 ? B->pbs[i].n : B->pbs[i].m;
      const int64_t n_out = dir == 0 ? B->pbs[i].m : B->pbs[i].n;
      PassPlan p;
      p.tiles = static_cast<int>(pad_rows(n_out, k.cols()) / k.cols());
      // rows past n_in carry a -inf bias, so a split only streams the rows up
      // to the next multiple of its unit (two tiles), not the storage padding
      p.n_in_rb = static_cast<int>(pad_rows(n_in, 32 * k.mt) / 16);
      p.rb_per_split = (dir == 0 ? rows_g : rows_f) / 16;
      p.splits = (p.n_in_rb + p.rb_per_split - 1) / p.rb_per_split;
      if (!even_tiles(k, p)) return OTTX_EINVAL;
      (dir == 0 ? B->gp[i] : B->fp[i]) = p;
      (dir == 0 ? B->n_work_g : B->n_work_f) += static_cast<int64_t>(p.tiles) * p.splits;
    }
  }
  if (B->n_work_g > (int64_t{1} << 31) - 1 || B->n_work_f > (int64_t{1} << 31) - 1) return OTTX_EINVAL;
  B->off_work_g = take(sizeof(int4) * B->n_work_g);
  B->off_work_f = take(sizeof(int4) * B->n_work_f);
  B->off_live_g = take(sizeof(int4) * B->n_work_g);
  B->off_live_f = take(sizeof(int4) * B->n_work_f);
  B->off_live_n = take(sizeof(uint32_t) * 2);
  B->n_prep = 0;
  for (int i = 0; i < B->K; ++i)
    for (int side = 0; side < 2; ++side) {
      const int64_t np_pad = pad_rows(side ? B->pbs[i].m : B->pbs[i].n, B->pad);
      B->n_prep += (np_pad / 16 * k.frag_chunks() * 64 + 255) / 256 + (np_pad + 255) / 256;
    }
  if (B->n_prep > (int64_t{1} << 31) - 1) return OTTX_EINVAL;
  B->off_prep = take(sizeof(int4) * B->n_prep);
  B->off_out = take(sizeof(double) * 16 * B->K);
  B->stat_chunks = static_cast<int>((B->max_np + kBatchStatRows - 1) / kBatchStatRows);
  B->off_stats = take(sizeof(double) * 2 * static_cast<size_t>(B->K) * 2 * B->stat_chunks * bt->d);
  for (int i = 0; i < B->K; ++i) {
    B->L[i] = make_layout(&B->pbs[i], k, B->gp[i], B->fp[i], 1, B->pad);
    B->base[i] = take(B->L[i].total);
  }
  B->total = off;
  return OTTX_OK;
}

// Pointers of one problem's layout inside the workspace (no launches).
Prepared prepared_view(char* base, const Layout& L, const ottx_problem& pb) {
  Prepared P;
  P.a = pb.a;
  P.b = pb.b;
  P.base = base;
  P.L = L;
  P.frag_x = base + L.frag_x;
  P.frag_y = base + L.frag_y;
  P.norm_x = reinterpret_cast<float*>(base + L.norm_x);
  P.norm_y = reinterpret_cast<float*>(base + L.norm_y);
  P.mu = reinterpret_cast<float*>(base + L.mu);
  P.stats_out = reinterpret_cast<double*>(base + L.stats_out);
  P.scal = reinterpret_cast<SolveScalars*>(base + L.scal);
  return P;
}

// Pinned host staging for the batch tables, kept per thread and grown on demand.
struct HostStaging {
  void* ptr = nullptr;
  size_t bytes = 0;
};

int host_staging(size_t bytes, void** out) {
  static thread_local HostStaging h;
  if (h.bytes < bytes) {
    if (h.ptr) (void)hipHostFree(h.ptr);
    h.ptr = nullptr;
    h.bytes = 0;
    OTTX_HIP(hipHostMalloc(&h.ptr, bytes, hipHostMallocDefault));
    h.bytes = bytes;
  }
  *out = h.ptr;
  return OTTX_OK;
}

}  // namespace ottx

extern "C" size_t ottx_sinkhorn_batched_workspace_bytes(const ottx_batch* bt, const ottx_sinkhorn_opts* o) {
  if (check_batch(bt) != OTTX_OK || check_opts(o) != OTTX_OK) return 0;
  BatchPlan B;
  if (make_batch_plan(bt, o, &B) != OTTX_OK) return 0;
  return B.total;
}

extern "C" int ottx_sinkhorn_batched_solve(const ottx_batch* bt, const ottx_sinkhorn_opts* o, ottx_batch_out* out,
                                           void* workspace, size_t workspace_bytes, void* stream) {
  int rc = check_batch(bt);
  if (rc != OTTX_OK) return rc;
  if ((rc = check_opts(o)) != OTTX_OK) return rc;
  if (!out || !out->f || !out->g || !out->errors || !out->reg_ot_cost || !out->converged || !out->n_iters ||
      !workspace)
    return OTTX_EINVAL;
  if (o->init_mode == OTTX_INIT_F_FROM_G || o->recenter_potentials || o->parallel_dual_updates ||
      o->epsilon_init > 1.0 || o->momentum_start != 0)
    return OTTX_EUNSUPPORTED;
  if (o->init_mode < 0 || o->init_mode > 3) return OTTX_EINVAL;
  const int inner = o->inner_iterations;
  const int outer = (o->max_iterations + inner - 1) / inner;
  if (outer < 1) return OTTX_EUNSUPPORTED;
  BatchPlan B;
  if ((rc = make_batch_plan(bt, o, &B)) != OTTX_OK) return rc;
  if (workspace_bytes < B.total) return OTTX_EWORKSPACE;
  if (reinterpret_cast<uintptr_t>(workspace) & 255) return OTTX_EWORKSPACE;
  const KernelCfg& k = B.k;
  const int K = B.K;
  hipStream_t st = static_cast<hipStream_t>(stream);
  HostHelpers* hh = nullptr;
  if ((rc = host_helpers(&hh)) != OTTX_OK) return rc;
  char* ws = static_cast<char*>(workspace);
  const bool force_scan = o->min_iterations == o->max_iterations;
  const int total_blocks = force_scan ? o->max_iterations / inner : outer;
  const bool balanced = (bt->tau_a == 1.0 && bt->tau_b == 1.0);
  uint32_t* batch_done = reinterpret_cast<uint32_t*>(ws + B.off_done);

  // Host tables (items, per-problem arguments of both directions, work lists)
  // are staged in pinned memory with the same offsets they have in the
  // workspace, so one copy moves them.  The staging buffer is free whenever no
  // call is running on this thread: every call ends with a stream sync after
  // its last copy.
  const size_t tbl_bytes = B.off_out - B.off_items;
  void* stage = nullptr;
  if ((rc = host_staging(std::max(tbl_bytes, sizeof(double) * 16 * K), &stage)) != OTTX_OK) return rc;
  char* sb = static_cast<char*>(stage);
  BatchItem* items = reinterpret_cast<BatchItem*>(sb);
  HalfStepArgs* args_g = reinterpret_cast<HalfStepArgs*>(sb + (B.off_args_g - B.off_items));
  HalfStepArgs* args_f = reinterpret_cast<HalfStepArgs*>(sb + (B.off_args_f - B.off_items));
  int4* work_g = reinterpret_cast<int4*>(sb + (B.off_work_g - B.off_items));
  int4* work_f = reinterpret_cast<int4*>(sb + (B.off_work_f - B.off_items));
  int4* prep = reinterpret_cast<int4*>(sb + (B.off_prep - B.off_items));
  int64_t wg = 0, wf = 0, wp = 0;
  for (int i = 0; i < K; ++i) {
    const ottx_problem& pb = B.pbs[i];
    const Layout& L = B.L[i];
    char* base = ws + B.base[i];
    const Prepared P = prepared_view(base, L, pb);
    BatchItem& it = items[i];
    it = BatchItem{};
    it.x = pb.x;
    it.y = pb.y;
    it.a = pb.a;
    it.b = pb.b;
    it.a_fill = (bt->flags & OTTX_BATCH_UNIFORM_A) ? bt->a + bt->x_off[i] : nullptr;
    it.b_fill = (bt->flags & OTTX_BATCH_UNIFORM_B) ? bt->b + bt->y_off[i] : nullptr;
    it.f = out->f + bt->x_off[i];
    it.g = out->g + bt->y_off[i];
    it.g_save = reinterpret_cast<float*>(base + L.g_save);
    it.errors = out->errors + static_cast<int64_t>(i) * outer;
    it.n = pb.n;
    it.m = pb.m;
    it.n_pad = pad_rows(pb.n, B.pad);
    it.m_pad = pad_rows(pb.m, B.pad);
    it.n_div = bt->mean_rows_x > 0.0 ? bt->mean_rows_x : static_cast<double>(pb.n);
    it.m_div = bt->mean_rows_y > 0.0 ? bt->mean_rows_y : static_cast<double>(pb.m);
    it.eps_abs = pb.epsilon;
    it.frag_x = P.frag_x;
    it.frag_y = P.frag_y;
    it.norm_x = P.norm_x;
    it.norm_y = P.norm_y;
    it.mu = P.mu;
    it.scal = P.scal;
    it.state = reinterpret_cast<SolveState*>(base + L.state);
    it.counters = reinterpret_cast<uint32_t*>(base + L.counters);
    it.dual_part = reinterpret_cast<double*>(base + L.dual_part);
    it.n_counters = std::max(B.gp[i].tiles, B.fp[i].tiles);
    const double eps_host = pb.epsilon > 0.0 ? pb.epsilon : 1.0;  // replaced on the device by scal
    HalfStepArgs ga = base_args(&pb, k, B.gp[i], P, 1, it.f, it.g, eps_host);
    HalfStepArgs fa = base_args(&pb, k, B.fp[i], P, 0, it.g, it.f, eps_host);
    for (HalfStepArgs* a : {&ga, &fa}) {
      a->norm_p = o->norm_error;
      a->momentum = static_cast<float>(o->momentum);
      a->errors = it.errors;
      a->row_err_partial = reinterpret_cast<const double*>(base + L.err_row);
      a->row_tiles = B.fp[i].tiles;
      a->inner = inner;
      a->min_it = o->min_iterations;
      a->max_it = o->max_iterations;
      a->outer = outer;
      a->force_scan = force_scan ? 1 : 0;
      a->threshold = static_cast<float>(o->threshold);
      a->batch_done = batch_done;
      a->problem = i;
    }
    ga.out_save = it.g_save;
    ga.lse_shift = ga.lse_store = reinterpret_cast<float*>(base + L.lse_y);
    fa.lse_shift = fa.lse_store = reinterpret_cast<float*>(base + L.lse_x);
    args_g[i] = ga;
    args_f[i] = fa;
    for (int t = 0; t < B.gp[i].tiles; ++t)
      for (int sp = 0; sp < B.gp[i].splits; ++sp) work_g[wg++] = make_int4(i, t, sp, 0);
    for (int t = 0; t < B.fp[i].tiles; ++t)
      for (int sp = 0; sp < B.fp[i].splits; ++sp) work_f[wf++] = make_int4(i, t, sp, 0);
    for (int side = 0; side < 2; ++side) {
      const int64_t np_pad = side ? it.m_pad : it.n_pad;
      const int pack_blocks = static_cast<int>((np_pad / 16 * k.frag_chunks() * 64 + 255) / 256);
      const int norm_blocks = static_cast<int>((np_pad + 255) / 256);
      for (int q = 0; q < pack_blocks; ++q) prep[wp++] = make_int4(i, side, q, 0);
      for (int q = 0; q < norm_blocks; ++q) prep[wp++] = make_int4(i, side, q, 1);
    }
  }
  // Longest items first (the classic longest-processing-time order): an
  // item's cost is the rows its split streams times the tile's columns, the
  // same for every column tile of a problem; dispatching the large problems'
  // items before the small ones keeps the launch's tail short.
  auto by_cost = [&](int4* work, int64_t count, bool g_dir) {
    std::vector<std::pair<int64_t, int4>> v(static_cast<size_t>(count));
    for (int64_t q = 0; q < count; ++q) {
      const int4 w = work[q];
      const PassPlan& pp = g_dir ? B.gp[w.x] : B.fp[w.x];
      const int64_t rb0 = static_cast<int64_t>(w.z) * pp.rb_per_split;
      const int64_t rows = std::min<int64_t>(pp.rb_per_split, pp.n_in_rb - rb0);
      v[static_cast<size_t>(q)] = {rows, w};
    }
    std::stable_sort(v.begin(), v.end(), [](const std::pair<int64_t, int4>& l, const std::pair<int64_t, int4>& r) {
      return l.first > r.first;
    });
    for (int64_t q = 0; q < count; ++q) work[q] = v[static_cast<size_t>(q)].second;
  };
  by_cost(work_g, wg, true);
  by_cost(work_f, wf, false);
  // the live lists start as the full ones; the queue heads at 0
  std::memcpy(sb + (B.off_live_g - B.off_items), work_g, sizeof(int4) * wg);
  std::memcpy(sb + (B.off_live_f - B.off_items), work_f, sizeof(int4) * wf);
  uint32_t* live_n = reinterpret_cast<uint32_t*>(sb + (B.off_live_n - B.off_items));
  live_n[0] = static_cast<uint32_t>(wg);
  live_n[1] = static_cast<uint32_t>(wf);
  BatchItem* d_items = reinterpret_cast<BatchItem*>(ws + B.off_items);
  HalfStepArgs* d_args_g = reinterpret_cast<HalfStepArgs*>(ws + B.off_args_g);
  HalfStepArgs* d_args_f = reinterpret_cast<HalfStepArgs*>(ws + B.off_args_f);
  int4* d_work_g = reinterpret_cast<int4*>(ws + B.off_work_g);
  int4* d_work_f = reinterpret_cast<int4*>(ws + B.off_work_f);
  int4* d_live_g = reinterpret_cast<int4*>(ws + B.off_live_g);
  int4* d_live_f = reinterpret_cast<int4*>(ws + B.off_live_f);
  uint32_t* d_live_n = reinterpret_cast<uint32_t*>(ws + B.off_live_n);
  OTTX_HIP(hipMemcpyAsync(d_items, stage, tbl_bytes, hipMemcpyHostToDevice, st));

  // Preparation of every problem.
  const double s = bt->inv_scale_cost;
  double* d_stats = reinterpret_cast<double*>(ws + B.off_stats);
  hipLaunchKernelGGL(batch_stats_partial_kernel, dim3(B.stat_chunks, K, 2), dim3(256), 0, st, d_items, bt->d,
                     B.stat_chunks, d_stats);
  hipLaunchKernelGGL(batch_stats_final_kernel, dim3(K), dim3(256), 0, st, d_items, bt->d, B.stat_chunks,
                     static_cast<const double*>(d_stats), s, bt->relative_epsilon, bt->tau_a, bt->tau_b);
  hipLaunchKernelGGL(batch_prep_kernel, dim3(static_cast<unsigned>(B.n_prep)), dim3(256), 0, st, d_items,
                     reinterpret_cast<const int4*>(ws + B.off_prep), bt->d, k.frag_chunks(), k.prec, s);
  hipLaunchKernelGGL(batch_init_kernel, dim3(K), dim3(256), 0, st, d_items, outer, batch_done,
                     zero_init_bits(o->init_mode));
  OTTX_LAUNCHED();

  const uint32_t g_plain = kModeWritePot | kModeCheckDone;
  const uint32_t g_err = kModeWritePot | kModeCheckDone | kModeSaveOld | kModeErrCol;
  const uint32_t f_plain = kModeWritePot | kModeCheckDone;
  const uint32_t f_err = kModeWritePot | kModeCheckDone | (balanced ? 0u : kModeErrRow);
  const dim3 block(64 * k.w);
  // One workgroup per live item.  Launches made by the host loop take the
  // live counts of the last block the host has seen finish; the lists only
  // shrink, so those counts bound the ones the launches will read.  A
  // captured block keeps the full counts.
  uint32_t grid_g = static_cast<uint32_t>(B.n_work_g), grid_f = static_cast<uint32_t>(B.n_work_f);
  auto launch = [&](bool g_dir, uint32_t mode, hipStream_t s2) -> int {
    hipLaunchKernelGGL(k.bfn, dim3(std::max(1u, g_dir ? grid_g : grid_f)), block, 0, s2,
                       g_dir ? d_args_g : d_args_f, g_dir ? d_live_g : d_live_f,
                       static_cast<const uint32_t*>(d_live_n + (g_dir ? 0 : 1)), mode);
    OTTX_LAUNCHED();
    return OTTX_OK;
  };
  BatchMail* bmail = reinterpret_cast<BatchMail*>(static_cast<char*>(hh->mailbox) + 2048);
  const uint32_t seq = ++hh->seq;
  __atomic_store_n(&bmail->progress, static_cast<uint64_t>(0), __ATOMIC_RELEASE);
  auto launch_block = [&](hipStream_t s2) -> int {
    for (int itn = 0; itn < inner; ++itn) {
      int r = launch(false, itn == inner - 1 ? f_err : f_plain, s2);
      if (r != OTTX_OK) return r;
      r = launch(true, itn == inner - 1 ? g_err : g_plain, s2);
      if (r != OTTX_OK) return r;
    }
    hipLaunchKernelGGL(batch_progress_kernel, dim3(1), dim3(kCompactThreads), 0, s2, batch_done, bmail, seq,
                       K, static_cast<const int4*>(d_work_g),
                       static_cast<const int4*>(d_work_f), d_live_g, d_live_f, d_live_n,
                       static_cast<int>(B.n_work_g), static_cast<int>(B.n_work_f));
    OTTX_LAUNCHED();
    return OTTX_OK;
  };

  if ((rc = launch(true, g_plain, st)) != OTTX_OK) return rc;  // g-update of iteration 0
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  hipStream_t cap = nullptr;
  // As for one problem, a graph pays for its capture and instantiation only
  // when a block has several launches.
  if (o->use_graph && inner >= 4) {
    OTTX_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    OTTX_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    rc = launch_block(cap);
    hipGraph_t g = nullptr;
    const hipError_t e1 = hipStreamEndCapture(cap, &g);
    if (rc != OTTX_OK || e1 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      (void)hipStreamDestroy(cap);
      return rc != OTTX_OK ? rc : OTTX_EHIP;
    }
    graph = g;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(graph);
      (void)hipStreamDestroy(cap);
      return OTTX_EHIP;
    }
  }
  int status = OTTX_OK;
  int launched = 0;
  // One block runs ahead of the host, which spins on the pinned mailbox that
  // every block's progress kernel writes (no event in the stream, no blocking
  // wait); the stream is queried now and then so that an error or a drained
  // stream ends the wait.  Returns false on an error.
  auto wait_block = [&](uint32_t blocks) -> bool {
    for (uint32_t spin = 1;; ++spin) {
      const uint64_t v = __atomic_load_n(&bmail->progress, __ATOMIC_ACQUIRE);
      if ((v >> 32) == seq && static_cast<uint32_t>(v) >= blocks) return true;
      if ((spin & 1023u) == 0) {
        const hipError_t q = hipStreamQuery(st);
        if (q == hipSuccess) {
          const uint64_t w = __atomic_load_n(&bmail->progress, __ATOMIC_ACQUIRE);
          return (w >> 32) == seq && static_cast<uint32_t>(w) >= blocks;
        }
        if (q != hipErrorNotReady) return false;
      }
    }
  };
  for (int blk = 0; blk < total_blocks; ++blk) {
    if (exec) {
      if (hipGraphLaunch(exec, st) != hipSuccess) { status = OTTX_EHIP; break; }
    } else {
      status = launch_block(st);
      if (status != OTTX_OK) break;
    }
    ++launched;
    if (blk >= 1) {
      if (!wait_block(static_cast<uint32_t>(blk))) { status = OTTX_EHIP; break; }
      if (__atomic_load_n(&bmail->done, __ATOMIC_ACQUIRE) >= static_cast<uint32_t>(K)) break;
      grid_g = __atomic_load_n(&bmail->live[0], __ATOMIC_ACQUIRE);
      grid_f = __atomic_load_n(&bmail->live[1], __ATOMIC_ACQUIRE);
    }
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (cap) (void)hipStreamDestroy(cap);
  if (status != OTTX_OK) return status;

  double* d_out = reinterpret_cast<double*>(ws + B.off_out);
  // one workgroup per 4096 points of the longest cloud (per side and problem)
  const int dual_blocks = static_cast<int>(std::min<int64_t>(kDualBlocks, (B.max_np + 4095) / 4096));
  hipLaunchKernelGGL(batch_dual_partial_kernel, dim3(dual_blocks, 3, K), dim3(256), 0, st, d_items, bt->tau_a,
                     bt->tau_b, outer, dual_blocks);
  hipLaunchKernelGGL(batch_dual_final_kernel, dim3(K), dim3(64), 0, st, d_items, outer, dual_blocks, d_out);
  hipLaunchKernelGGL(batch_copy_g_kernel, dim3(static_cast<unsigned>((B.max_np + 255) / 256), K), dim3(256), 0, st,
                     d_items);
  OTTX_LAUNCHED();
  // The tables' copy precedes these launches on the stream, so the staging
  // buffer can take the results.
  void* res = stage;
  OTTX_HIP(hipMemcpyAsync(res, d_out, sizeof(double) * 16 * K, hipMemcpyDeviceToHost, st));
  OTTX_HIP(hipStreamSynchronize(st));
  const double* du = static_cast<const double*>(res);
  for (int i = 0; i < K; ++i) {
    const double* q = du + 16 * i;
    out->reg_ot_cost[i] = q[0] + q[2] + q[7] * (q[1] * q[3] - q[8]);
    if (out->epsilon) out->epsilon[i] = q[7];
    const bool any_nan = q[4] > 0.0;
    const float last = static_cast<float>(q[6]);
    out->converged[i] = (!any_nan && last < static_cast<float>(o->threshold)) ? 1 : 0;
    out->n_iters[i] = static_cast<int32_t>(q[5]) * inner;
  }
  out->n_blocks = launched;
  return OTTX_OK;
}

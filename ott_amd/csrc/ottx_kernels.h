// Device kernels of the point-cloud Sinkhorn path (gfx950 / CDNA4).
//
// The central kernel is the fused log-domain half-step.  For the g-update
// (reference: geometry/geometry.py:376-398 -> pointcloud.py:74-83 ->
// geometry.py:260-299, 423-442 with C from costs.py:554-561) it computes, for
// every column j of the out side,
//
//   LSE_i((f_i - C_ij) / eps),   C_ij = s (|x_i|^2 + |y_j|^2 - 2 <x_i, y_j>)
//
// without materialising C: the cross term runs on MFMA.  The y fragments are
// packed already multiplied by c = 2 s log2(e) / eps (the target epsilon),
// and the row bias c beta_i, beta_i = (f_i - s|x_i|^2) / (2 s), is loaded as
// the accumulator input, so the MFMA result acc_ij = c (<x_i, y_j> + beta_i)
// is the base-2 exponent itself ("exponent units"; a launch at another
// epsilon, an epsilon schedule, multiplies it by k = eps / eps_t).  With a
// fixed per-column shift m_j one VOP2 subtraction per pair remains, and none
// when a workgroup's shifts are close enough to share one, folded into the
// bias.  A streaming max / sum-of-exp2 per
// column reduces the tile, wave shuffles combine lane groups, and a split-K
// combine (last-arriving workgroup per column tile) merges row splits.  The
// epilogue turns (max, sum) into the new potential, the marginal of the old
// state and the float64 error terms, and the last column tile of an error
// pass finalizes the block: errors[], convergence, the stop flag.
#pragma once

#include "ottx_common.h"

namespace ottx {

__device__ __forceinline__ float exp2_fast(float x) { return __builtin_amdgcn_exp2f(x); }

__device__ __forceinline__ uint32_t f32_to_bf16_rne(float v) {
  uint32_t u = __float_as_uint(v);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (u >> 16) | 0x40u;  // NaN stays a (quiet) NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

__device__ __forceinline__ float bf16_to_f32(uint32_t h) { return __uint_as_float(h << 16); }

// NaN-propagating maximum (IEEE 754-2019 maximum; v_maximum3_f32 on gfx950).
// fmaxf drops a NaN operand, which would let a NaN row vanish from a column's
// log-sum-exp; jax.scipy.special.logsumexp (math/utils.py:132-138) returns NaN
// there, and so must the kernels.
__device__ __forceinline__ float nan_max(float a, float b) { return __builtin_elementwise_maximum(a, b); }

// Combine two (max, sum) pairs of base-2 log-sum-exp accumulators (max in
// exponent units).
__device__ __forceinline__ void lse_merge(float& mx, float& sm, float om, float os) {
  const float mn = nan_max(mx, om);
  float s = 0.f;
  if (mn != -INFINITY) s = sm * exp2_fast(mx - mn) + os * exp2_fast(om - mn);
  mx = mn;
  sm = s;
}

// Cross-workgroup hand-off without touching the L2 as a whole.  On gfx950 an
// agent-scope release fence is buffer_wbl2 (write back every dirty line of
// the XCD's L2) and __threadfence() adds buffer_inv; issued by every
// workgroup of a launch, the write-backs alone cost about 16 us per half-step
// at n = m = 16384.  Instead, the data handed to another workgroup is stored
// write-through at agent scope (global_store ... sc1), each thread waits for
// its own stores to complete (s_waitcnt 0) before the workgroup barrier that
// precedes the arrival counter's atomic, and readers take the data with
// agent-coherent (sc1) loads, which bypass a stale L2, after the counter's
// atomic returned.
__device__ __forceinline__ void store_f2_through(float2* p, float2 v) {
  const uint64_t u = static_cast<uint64_t>(__float_as_uint(v.x)) |
                     (static_cast<uint64_t>(__float_as_uint(v.y)) << 32);
  __hip_atomic_store(reinterpret_cast<uint64_t*>(p), u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void store_f64_through(double* p, double v) {
  __hip_atomic_store(reinterpret_cast<uint64_t*>(p), static_cast<uint64_t>(__double_as_longlong(v)),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// This thread's stores (and loads) have completed; the compiler keeps memory
// operations on their side of it.
__device__ __forceinline__ void wait_own_stores() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ double load_f64_coherent(const double* p) {
  const uint64_t v = __hip_atomic_load(reinterpret_cast<const uint64_t*>(p), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double(static_cast<long long>(v));
}

__device__ __forceinline__ float2 load_f2_coherent(const float2* p) {
  const uint64_t v = __hip_atomic_load(reinterpret_cast<const uint64_t*>(p), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  float2 r;
  r.x = __uint_as_float(static_cast<uint32_t>(v));
  r.y = __uint_as_float(static_cast<uint32_t>(v >> 32));
  return r;
}

// Momentum.weight / Momentum.lehmann (acceleration.py:138-160) once iteration
// start is reached: the constant value unless the error of block
// idx - 1 (idx = start / inner) is below the threshold, then
// 2 / (1 + sqrt(1 - r^(1 / inner))), r = min(err[idx - 1] / err[idx - 2], 0.99),
// in float32 like the reference's errors.  start >= 2 inner (checked by the host).
__device__ inline float adaptive_momentum(const float* errors, int start, int inner, float value,
                                          float threshold) {
  const int idx = start / inner;
  const float e1 = errors[idx - 1];
  if (!(e1 < threshold)) return value;
  float r = e1 / errors[idx - 2];
  if (r > 0.99f) r = 0.99f;  // NaN stays NaN, as with jnp.minimum
  return 2.f / (1.f + sqrtf(1.f - powf(r, 1.f / static_cast<float>(inner))));
}

// ---------------------------------------------------------------------------
// Block finalizer: sinkhorn.py:914-931 (errors layout), :940-953 and
// fixed_point_loop.py:63-70 (continue test), run by one thread of the last
// column tile of an error pass with the pass's sums: col = sum |colmarg -
// target|^p, row (unbalanced: the row term) and the total mass.
// ---------------------------------------------------------------------------
__device__ inline void finalize_block(const HalfStepArgs& p, double col, double row, double mass) {
  const double inv_p = 1.0 / static_cast<double>(p.norm_p);
  double err = (p.norm_p == 1) ? col : pow(col, inv_p);
  if (!p.balanced) err += (p.norm_p == 1) ? row : pow(row, inv_p);
  // parallel updates of a symmetric problem: the row marginal's error equals
  // the column one (sinkhorn.py:160-189 adds both)
  if (p.symmetric) err *= 2.0;
  SolveState* st = p.state;
  const int k = st->block;
  const int t_last = k * p.inner + p.inner - 1;
  const bool record = (t_last >= p.min_it) || (t_last == p.max_it - 1);
  const float e = record ? static_cast<float>(err) : INFINITY;
  if (k < p.outer) p.errors[k] = e;
  st->block = k + 1;
  st->last_err = err;
  st->total_mass = mass;
  const int t_next = (k + 1) * p.inner;
  bool cont;
  if (p.force_scan) {
    cont = (k + 1) < (p.max_it / p.inner);
  } else {
    cont = (t_next < p.max_it) &&
           ((t_next < p.min_it) || (isfinite(e) && !(e < p.threshold)));
  }
  if (!cont) {
    st->n_blocks = k + 1;
    st->done = 1;
    if (p.batch_done != nullptr) {
      atomicOr(p.batch_done + kBatchDoneBits + (p.problem >> 5), 1u << (p.problem & 31));
      atomicAdd(p.batch_done, 1u);
    }
  }
  if (p.mail != nullptr) {
    // the loop's progress (and this block's error) for the host, which keeps
    // at most one block queued ahead; the stop bit tells it the results are
    // due (the output finalizer writes them)
    __hip_atomic_store(&p.mail->last_err, err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    const uint64_t v = (static_cast<uint64_t>(p.seq) << 32) | (cont ? 0u : 0x80000000u) | static_cast<uint32_t>(k + 1);
    __hip_atomic_store(&p.mail->progress, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// ---------------------------------------------------------------------------
// Streaming core of the fused half-step, shared by the single-problem and the
// batched kernels.
//   PREC 1 (bf16 split): KK = K blocks of 32 features, fragments [rb][kb][hi|lo][lane][8 bf16]
//   PREC 0 (fp32)      : KK = k-steps of 4 features,  fragments [rb][kq][lane][4 f32]
//   NB  = 16-column blocks per wave, MT = 16-row blocks per streamed tile.
// MFMA operand layouts (16x16 output): lane l holds A row (l % 16), B column
// (l % 16), and output rows 4 (l / 16) + r, r < 4, of column l % 16.
// ---------------------------------------------------------------------------
template <int PREC, int KK, int NB>
struct OutOperand {
  static constexpr int KQ = (KK + 3) / 4;
  bf16x8 bh[PREC == 1 ? NB : 1][PREC == 1 ? KK : 1];
  bf16x8 bl[PREC == 1 ? NB : 1][PREC == 1 ? KK : 1];
  float4 bq[PREC == 0 ? NB : 1][PREC == 0 ? KQ : 1];

  // Fragments of column blocks cb0 .. cb0 + NB - 1 (16 points each).
  __device__ __forceinline__ void load(const void* frag, int64_t cb0, int lane) {
    if constexpr (PREC == 1) {
      const uint4* __restrict__ of = static_cast<const uint4*>(frag);
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int kb = 0; kb < KK; ++kb) {
          const size_t idx = ((static_cast<size_t>(cb0 + nb) * KK + kb) * 2) * 64 + lane;
          bh[nb][kb] = __builtin_bit_cast(bf16x8, of[idx]);
          bl[nb][kb] = __builtin_bit_cast(bf16x8, of[idx + 64]);
        }
    } else {
      const float4* __restrict__ of = static_cast<const float4*>(frag);
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int q = 0; q < KQ; ++q) bq[nb][q] = of[(static_cast<size_t>(cb0 + nb) * KQ + q) * 64 + lane];
    }
  }
};

__device__ __forceinline__ float f4_get(const float4& v, int e) {
  return (e == 0) ? v.x : (e == 1) ? v.y : (e == 2) ? v.z : v.w;
}

// Number of 1 KiB fragment chunks per 16-row block.
template <int PREC, int KK>
struct FragChunks {
  static constexpr int value = (PREC == 1) ? KK * 2 : (PREC == 0 ? (KK + 3) / 4 : 1);
};

// Loads one streamed tile (contiguous fragment chunks of MT row blocks) and
// the raw potential and norm of its rows into registers.  Every thread issues
// the same loads unconditionally (row indices are clamped, tiles past the end
// of the split are clamped by the caller): the compiler counts outstanding
// loads per wave, and a load skipped by a branch would make it fall back to
// waiting for all of them (vmcnt(0)), which drains the prefetch.  For the same
// reason the row bias (h_i - norm_i) / (2 s) is only formed when the tile is
// copied to LDS.
template <int TILE_U4, int NT, int PER_T, int ROWS, int CH>
__device__ __forceinline__ void tile_prefetch(const u32x4* __restrict__ src, const float* __restrict__ in_pot,
                                              const float* __restrict__ in_norm, int64_t n_in, int rb, int tid,
                                              u32x4 (&pre)[PER_T], float2& pre_row) {
  const u32x4* base = src + static_cast<size_t>(rb) * CH * 64;
#pragma unroll
  for (int q = 0; q < PER_T; ++q) pre[q] = base[min(tid + q * NT, TILE_U4 - 1)];
  const int64_t i = min(static_cast<int64_t>(rb) * 16 + (tid & (ROWS - 1)), n_in - 1);
  pre_row.x = in_pot[i];
  pre_row.y = in_norm[i];
}

// Bias of row rb * 16 + tid (tid < ROWS) from its prefetched potential and
// norm, in exponent units (kb = c / (2 s)), less a workgroup-uniform shift;
// rows past n_in get -inf so they add nothing.
__device__ __forceinline__ float row_bias(float2 r, int64_t row, int64_t n_in, float kb, float bshift) {
  return row < n_in ? fmaf(r.x - r.y, kb, -bshift) : -INFINITY;
}

// How the exponent of a pair comes from its accumulator in fixed-shift mode:
// fmaf(acc, k, -m_j) at a launch epsilon other than the packing one,
// acc - m_j otherwise, or acc alone when the shift is workgroup-uniform (in
// the bias).
enum ExpMode : int { kExpFma = 0, kExpSub = 1, kExpNone = 2 };

// One streamed tile (in LDS at L, row biases at LB) against the NB column
// blocks in B: the exponents of MT x NB 16 x 16 pairs into acc, the bias
// entering as the accumulator input.
// The A operands (streamed rows) and row biases of one tile, read from LDS
// into registers.
template <int PREC, int KK, int MT>
struct InTile {
  static constexpr int KQ = (KK + 3) / 4;
  bf16x8 ah[PREC == 1 ? MT : 1][PREC == 1 ? KK : 1];
  bf16x8 al[PREC == 1 ? MT : 1][PREC == 1 ? KK : 1];
  float4 aq[PREC == 0 ? MT : 1][PREC == 0 ? KQ : 1];
  f32x4 bias[MT];

  __device__ __forceinline__ void load(const u32x4* L, const float* LB, int lane) {
    const int rsub = (lane >> 4) * 4;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const float4 bv = *reinterpret_cast<const float4*>(LB + mt * 16 + rsub);
      bias[mt][0] = bv.x;
      bias[mt][1] = bv.y;
      bias[mt][2] = bv.z;
      bias[mt][3] = bv.w;
      if constexpr (PREC == 1) {
#pragma unroll
        for (int kb = 0; kb < KK; ++kb) {
          ah[mt][kb] = __builtin_bit_cast(bf16x8, L[((mt * KK + kb) * 2) * 64 + lane]);
          al[mt][kb] = __builtin_bit_cast(bf16x8, L[((mt * KK + kb) * 2 + 1) * 64 + lane]);
        }
      } else {
#pragma unroll
        for (int q = 0; q < KQ; ++q) aq[mt][q] = __builtin_bit_cast(float4, L[(mt * KQ + q) * 64 + lane]);
      }
    }
  }
};

template <int PREC, int KK, int NB, int MT>
__device__ __forceinline__ void tile_mfma(const OutOperand<PREC, KK, NB>& B, const InTile<PREC, KK, MT>& A,
                                          f32x4 (&acc)[MT][NB]) {
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    if constexpr (PREC == 1) {
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
        acc[mt][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.ah[mt][0], B.bh[nb][0], A.bias[mt], 0, 0, 0);
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
        acc[mt][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.ah[mt][0], B.bl[nb][0], acc[mt][nb], 0, 0, 0);
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
        acc[mt][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.al[mt][0], B.bh[nb][0], acc[mt][nb], 0, 0, 0);
#pragma unroll
      for (int kb = 1; kb < KK; ++kb) {
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
          acc[mt][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.ah[mt][kb], B.bh[nb][kb], acc[mt][nb], 0, 0, 0);
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
          acc[mt][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.ah[mt][kb], B.bl[nb][kb], acc[mt][nb], 0, 0, 0);
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
          acc[mt][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.al[mt][kb], B.bh[nb][kb], acc[mt][nb], 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
        acc[mt][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(A.aq[mt][0].x, B.bq[nb][0].x, A.bias[mt], 0, 0, 0);
#pragma unroll
      for (int ks = 1; ks < KK; ++ks) {
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
          acc[mt][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(f4_get(A.aq[mt][ks >> 2], ks & 3),
                                                             f4_get(B.bq[nb][ks >> 2], ks & 3), acc[mt][nb], 0,
                                                             0, 0);
      }
    }
  }
}

// Reads a tile from LDS and runs its MFMAs.
template <int PREC, int KK, int NB, int MT>
__device__ __forceinline__ void tile_mfma(const OutOperand<PREC, KK, NB>& B, const u32x4* L, const float* LB,
                                          int lane, f32x4 (&acc)[MT][NB]) {
  InTile<PREC, KK, MT> A;
  A.load(L, LB, lane);
  tile_mfma<PREC, KK, NB, MT>(B, A, acc);
}

// exp2 of one tile's exponents against the fixed per-column shift mx[],
// summed into sm[] (fixed-shift mode); EM: how the exponent is formed
// (ExpMode).  k multiplies the accumulator (kExpFma only).
template <int NB, int MT, int EM>
__device__ __forceinline__ void tile_exp_shift(const f32x4 (&acc)[MT][NB], float k, const float (&mx)[NB],
                                               float (&sm)[NB]) {
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const float m = mx[nb];
    auto ex = [&](float v) {
      if constexpr (EM == kExpFma) return exp2_fast(fmaf(v, k, -m));
      else if constexpr (EM == kExpSub) return exp2_fast(v - m);
      else return exp2_fast(v);
    };
    float part[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) part[r] = ex(acc[0][nb][r]);
#pragma unroll
    for (int mt = 1; mt < MT; ++mt)
#pragma unroll
      for (int r = 0; r < 4; ++r) part[r] += ex(acc[mt][nb][r]);
    sm[nb] += (part[0] + part[1]) + (part[2] + part[3]);
  }
}

// Streaming log-sum-exp of one tile's accumulators, base 2, with a running
// per-column max (the mode without a fixed shift): one rescale per column per
// tile.  The exponent is k acc (k > 0), mx[] in exponent units.  The tile max
// is a tree and the sum uses four partial accumulators, so no VALU chain is
// longer than MT + 2 dependent operations.
template <int NB, int MT>
__device__ __forceinline__ void tile_exp_max(const f32x4 (&acc)[MT][NB], float k, float (&mx)[NB],
                                           float (&sm)[NB]) {
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    // NaN-propagating maxima: a NaN accumulator makes the column's sum NaN
    float m4[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
      m4[mt] = nan_max(nan_max(acc[mt][nb][0], acc[mt][nb][1]), nan_max(acc[mt][nb][2], acc[mt][nb][3]));
#pragma unroll
    for (int w = 1; w < MT; w <<= 1)
#pragma unroll
      for (int mt = 0; mt + w < MT; mt += 2 * w) m4[mt] = nan_max(m4[mt], m4[mt + w]);
    const float mnew = nan_max(mx[nb], k * m4[0]);
    if (mnew != -INFINITY) {
      float part[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) part[r] = exp2_fast(fmaf(acc[0][nb][r], k, -mnew));
#pragma unroll
      for (int mt = 1; mt < MT; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) part[r] += exp2_fast(fmaf(acc[mt][nb][r], k, -mnew));
      sm[nb] = fmaf(sm[nb], exp2_fast(mx[nb] - mnew), (part[0] + part[1]) + (part[2] + part[3]));
      mx[nb] = mnew;
    }
  }
}

// exp / sum of one tile's accumulators in either mode.
template <int NB, int MT, bool SHIFT, int EM>
__device__ __forceinline__ void tile_exp(const f32x4 (&acc)[MT][NB], float k, float (&mx)[NB], float (&sm)[NB]) {
  if constexpr (SHIFT)
    tile_exp_shift<NB, MT, EM>(acc, k, mx, sm);
  else
    tile_exp_max<NB, MT>(acc, k, mx, sm);
}

// Streams in-side row blocks [rb_begin, rb_end) (multiples of MT) against the
// NB column blocks held in B, updating each lane's running (max, sum) of
// exp2(k acc) per column (EM).  in_frag / in_pot / in_norm are relative to the
// first row of the in side; rows i >= n_in get a -inf bias.
//
// SHIFT: mx[] holds a fixed per-column shift (the column's log-sum-exp from
// the previous half-step of the same side) and only the sum is accumulated;
// otherwise mx[] is the running max, rescaled once per tile.
//
// The W waves of the workgroup stream the same rows: each tile of MT row
// blocks (contiguous in the fragment layout) and its row biases are copied
// once into LDS and every wave reads its A operands from there.  Two tiles are
// in flight in registers ahead of the one on MFMA (prefetch distance 2) and
// LDS is double-buffered, with one barrier per tile.  [rb_begin, rb_end)
// holds an even number of tiles (the planners guarantee it), so each loop
// trip computes two tiles without a branch in between.
template <int PREC, int KK, int NB, int MT, int W, bool SHIFT, int EM>
__device__ __forceinline__ void stream_rows(const OutOperand<PREC, KK, NB>& B, const void* in_frag,
                                            const float* __restrict__ in_pot, const float* __restrict__ in_norm,
                                            int64_t n_in, int rb_begin, int rb_end, float kb, float bshift,
                                            float k, int tid, float (&mx)[NB], float (&sm)[NB], u32x4* s_tile,
                                            float* s_bias) {
  constexpr int CH = FragChunks<PREC, KK>::value;
  constexpr int NT = W * 64;
  constexpr int TILE_U4 = MT * CH * 64;
  constexpr int PER_T = (TILE_U4 + NT - 1) / NT;
  constexpr int ROWS = MT * 16;
  static_assert((ROWS & (ROWS - 1)) == 0 && ROWS <= NT, "MT * 16 rows must be a power of two");
  const int lane = tid & 63;
  const u32x4* __restrict__ src = static_cast<const u32x4*>(in_frag);
  u32x4* L0 = s_tile;
  u32x4* L1 = s_tile + TILE_U4;
  float* B0 = s_bias;
  float* B1 = s_bias + ROWS;
  const int rb_last = rb_end - MT;  // start of the last tile; prefetches past it reload it

  u32x4 pre0[PER_T], pre1[PER_T];
  float2 pb0, pb1;
  tile_prefetch<TILE_U4, NT, PER_T, ROWS, CH>(src, in_pot, in_norm, n_in, rb_begin, tid, pre0, pb0);
  tile_prefetch<TILE_U4, NT, PER_T, ROWS, CH>(src, in_pot, in_norm, n_in, min(rb_begin + MT, rb_last), tid, pre1,
                                              pb1);
  // Software pipeline across tiles: the exp2 / sum of tile t runs after the
  // MFMAs of tile t + 1 are issued.  The MFMAs run at raised wave priority, so
  // of the two waves on a SIMD the one with matrix work keeps the matrix core
  // fed while the other's exp2 / sum fills the VALU.
  f32x4 acc0[MT][NB], acc1[MT][NB];
#pragma unroll
  for (int q = 0; q < PER_T; ++q)
    if (TILE_U4 % NT == 0 || tid + q * NT < TILE_U4) L0[tid + q * NT] = pre0[q];
  if (tid < ROWS) B0[tid] = row_bias(pb0, static_cast<int64_t>(rb_begin) * 16 + tid, n_in, kb, bshift);
  __syncthreads();
  tile_prefetch<TILE_U4, NT, PER_T, ROWS, CH>(src, in_pot, in_norm, n_in, min(rb_begin + 2 * MT, rb_last), tid,
                                              pre0, pb0);
  __builtin_amdgcn_s_setprio(1);
  tile_mfma<PREC, KK, NB, MT>(B, L0, B0, lane, acc0);
  __builtin_amdgcn_s_setprio(0);
  // One half trip: stage tile rb_stage (prefetched into pre) into buffer L,
  // prefetch tile rb_pref into pre, MFMA the staged tile into acc_new and
  // finish the previous tile in acc_old.  When the tile's A operands fit in
  // about 40 registers (PRE_A), they are read from LDS right after the
  // barrier and the previous tile's exp2 / sum runs while the reads are in
  // flight, so the MFMAs issue with every operand in registers; the
  // scheduling barriers keep the compiler from sinking the reads to their
  // uses, which would expose the LDS latency inside the MFMA phase.  Larger
  // tiles read their operands per row block (holding all of them would cost
  // occupancy) and finish the previous tile after the MFMAs.
  constexpr int A_REGS = MT * 4 + (PREC == 1 ? MT * KK * 8 : MT * ((KK + 3) / 4) * 4);
  constexpr bool PRE_A = A_REGS <= 40;
  for (int rb = rb_begin; rb < rb_end; rb += 2 * MT) {
    // tile rb + MT: stage, MFMA, and finish tile rb
#pragma unroll
    for (int q = 0; q < PER_T; ++q)
      if (TILE_U4 % NT == 0 || tid + q * NT < TILE_U4) L1[tid + q * NT] = pre1[q];
    if (tid < ROWS) B1[tid] = row_bias(pb1, static_cast<int64_t>(rb + MT) * 16 + tid, n_in, kb, bshift);
    __syncthreads();
    if constexpr (PRE_A) {
      InTile<PREC, KK, MT> A;
      A.load(L1, B1, lane);
      tile_prefetch<TILE_U4, NT, PER_T, ROWS, CH>(src, in_pot, in_norm, n_in, min(rb + 3 * MT, rb_last), tid,
                                                  pre1, pb1);
      __builtin_amdgcn_sched_barrier(0);
      tile_exp<NB, MT, SHIFT, EM>(acc0, k, mx, sm);
      __builtin_amdgcn_sched_barrier(0);
      __builtin_amdgcn_s_setprio(1);
      tile_mfma<PREC, KK, NB, MT>(B, A, acc1);
      __builtin_amdgcn_s_setprio(0);
    } else {
      tile_prefetch<TILE_U4, NT, PER_T, ROWS, CH>(src, in_pot, in_norm, n_in, min(rb + 3 * MT, rb_last), tid,
                                                  pre1, pb1);
      __builtin_amdgcn_s_setprio(1);
      tile_mfma<PREC, KK, NB, MT>(B, L1, B1, lane, acc1);
      __builtin_amdgcn_s_setprio(0);
      tile_exp<NB, MT, SHIFT, EM>(acc0, k, mx, sm);
    }
    if (rb + 2 * MT >= rb_end) {
      tile_exp<NB, MT, SHIFT, EM>(acc1, k, mx, sm);
      break;
    }
    // tile rb + 2 MT: stage, MFMA, and finish tile rb + MT
#pragma unroll
    for (int q = 0; q < PER_T; ++q)
      if (TILE_U4 % NT == 0 || tid + q * NT < TILE_U4) L0[tid + q * NT] = pre0[q];
    if (tid < ROWS) B0[tid] = row_bias(pb0, static_cast<int64_t>(rb + 2 * MT) * 16 + tid, n_in, kb, bshift);
    __syncthreads();
    if constexpr (PRE_A) {
      InTile<PREC, KK, MT> A;
      A.load(L0, B0, lane);
      tile_prefetch<TILE_U4, NT, PER_T, ROWS, CH>(src, in_pot, in_norm, n_in, min(rb + 4 * MT, rb_last), tid,
                                                  pre0, pb0);
      __builtin_amdgcn_sched_barrier(0);
      tile_exp<NB, MT, SHIFT, EM>(acc1, k, mx, sm);
      __builtin_amdgcn_sched_barrier(0);
      __builtin_amdgcn_s_setprio(1);
      tile_mfma<PREC, KK, NB, MT>(B, A, acc0);
      __builtin_amdgcn_s_setprio(0);
    } else {
      tile_prefetch<TILE_U4, NT, PER_T, ROWS, CH>(src, in_pot, in_norm, n_in, min(rb + 4 * MT, rb_last), tid,
                                                  pre0, pb0);
      __builtin_amdgcn_s_setprio(1);
      tile_mfma<PREC, KK, NB, MT>(B, L0, B0, lane, acc0);
      __builtin_amdgcn_s_setprio(0);
      tile_exp<NB, MT, SHIFT, EM>(acc1, k, mx, sm);
    }
  }
}

// Ping-pong variant of stream_rows for 8-wave workgroups.  Waves 0-3 (group 0)
// and 4-7 (group 1) sit two per SIMD and work half a tile apart, locked by
// one barrier per phase: in phase k, group g runs the MFMAs of tile
// (k - g) / 2 when k - g is even and the exp2 / sum of tile (k - g - 1) / 2
// when it is odd.  So in every phase each SIMD has one wave on the matrix
// core and its partner on the VALU, and each wave needs one accumulator set.
// Tile t + 1 is staged into LDS buffer (t + 1) & 1 at the start of phase
// 2t + 1, when the buffer's previous tile (t - 1) has been read by both
// groups; its global prefetch was issued two phases earlier.
template <int PREC, int KK, int NB, int MT, bool SHIFT, int EM>
__device__ __forceinline__ void stream_rows_pp(const OutOperand<PREC, KK, NB>& B, const void* in_frag,
                                               const float* __restrict__ in_pot, const float* __restrict__ in_norm,
                                               int64_t n_in, int rb_begin, int rb_end, float kb, float bshift,
                                               float k, int tid, float (&mx)[NB], float (&sm)[NB], u32x4* s_tile,
                                               float* s_bias) {
  constexpr int W = 8;
  constexpr int CH = FragChunks<PREC, KK>::value;
  constexpr int NT = W * 64;
  constexpr int TILE_U4 = MT * CH * 64;
  constexpr int PER_T = (TILE_U4 + NT - 1) / NT;
  constexpr int ROWS = MT * 16;
  static_assert((ROWS & (ROWS - 1)) == 0 && ROWS <= NT, "MT * 16 rows must be a power of two");
  const int lane = tid & 63;
  const int group = tid >> 8;  // waves 0-3 / 4-7
  const u32x4* __restrict__ src = static_cast<const u32x4*>(in_frag);
  const int T = (rb_end - rb_begin) / MT;  // tiles of this split
  const int rb_last = rb_end - MT;
  u32x4 pre[PER_T];
  float2 pb;
  auto stage = [&](int t) {
    u32x4* L = s_tile + (t & 1) * TILE_U4;
#pragma unroll
    for (int q = 0; q < PER_T; ++q)
      if (TILE_U4 % NT == 0 || tid + q * NT < TILE_U4) L[tid + q * NT] = pre[q];
    if (tid < ROWS) s_bias[(t & 1) * ROWS + tid] = row_bias(pb, static_cast<int64_t>(rb_begin + t * MT) * 16 + tid, n_in, kb, bshift);
  };
  tile_prefetch<TILE_U4, NT, PER_T, ROWS, CH>(src, in_pot, in_norm, n_in, rb_begin, tid, pre, pb);
  stage(0);
  tile_prefetch<TILE_U4, NT, PER_T, ROWS, CH>(src, in_pot, in_norm, n_in, min(rb_begin + MT, rb_last), tid, pre,
                                              pb);
  __syncthreads();
  f32x4 acc[MT][NB];
  for (int ph = 0; ph <= 2 * T; ++ph) {
    if (ph & 1) {
      // stage tile (ph + 1) / 2, prefetch the one after it
      const int tn = (ph >> 1) + 1;
      if (tn < T) {
        stage(tn);
        tile_prefetch<TILE_U4, NT, PER_T, ROWS, CH>(src, in_pot, in_norm, n_in, min(rb_begin + (tn + 1) * MT, rb_last),
                                                    tid, pre, pb);
      }
    }
    const int kg = ph - group;
    if (kg >= 0) {
      const int t = kg >> 1;
      if (t < T) {
        if ((kg & 1) == 0) {
          __builtin_amdgcn_s_setprio(1);
          tile_mfma<PREC, KK, NB, MT>(B, s_tile + (t & 1) * TILE_U4, s_bias + (t & 1) * ROWS, lane, acc);
          __builtin_amdgcn_s_setprio(0);
        } else {
          tile_exp<NB, MT, SHIFT, EM>(acc, k, mx, sm);
        }
      }
    }
    __syncthreads();
  }
}

// The streaming loop of a W-wave workgroup: the ping-pong schedule for eight
// waves, the double-buffered one otherwise.
template <int PREC, int KK, int NB, int MT, int W, bool SHIFT, ExpMode EM>
__device__ __forceinline__ void stream_any(const OutOperand<PREC, KK, NB>& B, const void* in_frag,
                                           const float* __restrict__ in_pot, const float* __restrict__ in_norm,
                                           int64_t n_in, int rb_begin, int rb_end, float kb, float bshift, float k,
                                           int tid, float (&mx)[NB], float (&sm)[NB], u32x4* s_tile,
                                           float* s_bias) {
  if constexpr (W == 8)
    stream_rows_pp<PREC, KK, NB, MT, SHIFT, EM>(B, in_frag, in_pot, in_norm, n_in, rb_begin, rb_end, kb, bshift, k,
                                                tid, mx, sm, s_tile, s_bias);
  else
    stream_rows<PREC, KK, NB, MT, W, SHIFT, EM>(B, in_frag, in_pot, in_norm, n_in, rb_begin, rb_end, kb, bshift, k,
                                                tid, mx, sm, s_tile, s_bias);
}

// Combine the four 16-lane row groups of each column; lanes < 16 store the
// column's (max, sum) to s_ms[(wave * NB + nb) * 16 + lane].
template <int NB>
__device__ __forceinline__ void combine_lanes(float (&mx)[NB], float (&sm)[NB], int lane, int wave,
                                              float2* s_ms) {
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
#pragma unroll
    for (int off = 16; off <= 32; off <<= 1) {
      const float om = __shfl_xor(mx[nb], off, 64);
      const float os = __shfl_xor(sm[nb], off, 64);
      lse_merge(mx[nb], sm[nb], om, os);
    }
    if (lane < 16) s_ms[(wave * NB + nb) * 16 + lane] = make_float2(mx[nb], sm[nb]);
  }
}

// Scalars of the column epilogue.
struct EpiScalars {
  uint32_t mode;
  int32_t norm_p;
  float eps, eps_marg, tau, inv_rho, momentum;
  double conv;  // exponent units of the launch to cost units: 2 s / (k c_pack)
};

// Epilogue of one out-side point j (pointers already offset to the problem):
// turns the column's (max, sum) into apply_lse_kernel's value ``app`` =
// eps * LSE_i((h_i - C_ij) / eps), the update of geometry.py:376-398 scaled by
// tau (sinkhorn.py:836-855) and blended by constant momentum
// (acceleration.py:161-172), and the marginal-error / mass terms of
// sinkhorn.py:191-221 with the unbalanced target of unbalanced_functions.py:30-50.
__device__ __forceinline__ void column_epilogue(const EpiScalars& e, float2 ms, int64_t j,
                                                const float* __restrict__ out_norm, float* out_pot,
                                                const float* __restrict__ out_w, const float* __restrict__ out_logw,
                                                float* out_save, float* out_vec, double& e_acc, double& m_acc) {
  float app;
  if (ms.x == -INFINITY || !(ms.y > 0.f)) {
    app = (ms.y != ms.y) ? ms.y : -INFINITY;
  } else {
    // in float64: the conversion factor is the exact inverse of the packing
    // scale, so the large term carries no systematic relative error
    app = static_cast<float>(e.conv * (static_cast<double>(ms.x) + static_cast<double>(log2f(ms.y)))) - out_norm[j];
  }
  const float h_old = out_pot[j];
  float h_marg = h_old;
  if (e.mode & kModeWritePot) {
    const float lw = out_logw ? out_logw[j] : logf(out_w[j]);
    float upd = e.eps * lw - (isfinite(app) ? app : 0.f);
    upd *= e.tau;
    float nv = upd;
    if (e.momentum != 1.f) nv = (1.f - e.momentum) * (isfinite(h_old) ? h_old : 0.f) + e.momentum * upd;
    out_pot[j] = nv;
    if ((e.mode & kModeErrRow) && !(e.mode & kModeErrOld)) h_marg = nv;
  }
  if (e.mode & kModeSaveOld) out_save[j] = h_old;
  if (e.mode & (kModeErrCol | kModeErrRow)) {
    const float marg = expf((h_marg + app) / e.eps_marg);
    const float w = out_w[j];
    const float target = (e.tau == 1.f) ? w : ((w > 0.f) ? w * expf(-h_marg * e.inv_rho) : 0.f);
    const double diff = fabs(static_cast<double>(marg) - static_cast<double>(target));
    e_acc += (e.norm_p == 1) ? diff : pow(diff, static_cast<double>(e.norm_p));
    m_acc += static_cast<double>(marg);
  }
  if (e.mode & kModeApply) out_vec[j] = (h_old == -INFINITY || h_old != h_old) ? h_old : app;
  if (e.mode & kModeMarginal) out_vec[j] = expf((h_old + app) / e.eps_marg);
  if (e.mode & kModeWriteApp) out_vec[j] = app;
}

// Deterministic tree sum of two float64 values over the workgroup; the
// result is valid in thread 0.
template <int NT>
__device__ __forceinline__ void block_sum2(double (*s_red)[NT], int tid, double a, double b) {
  s_red[0][tid] = a;
  s_red[1][tid] = b;
  __syncthreads();
#pragma unroll
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (tid < s) {
      s_red[0][tid] += s_red[0][tid + s];
      s_red[1][tid] += s_red[1][tid + s];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Costs other than the squared Euclidean (geometry/costs.py TICost family and
// Euclidean): c(x, y) = alpha (sum_k phi(x_k - y_k))^beta with phi(z) = |z|
// (PHI 1), z^2 (PHI 2) or |z|^p (PHI 3).  No cross term exists for MFMA, so a
// workgroup of 256 threads sums the per-feature terms on the vector ALUs for
// 64 rows by 64 columns at a time, features staged through LDS 32 at a time
// (4 x 4 pairs per thread), then folds the rows into each column's running
// (max, sum) in base-2 exponent units, z_ij = (h_i - s c_ij) log2(e) / eps.
// ---------------------------------------------------------------------------
constexpr int kGcTile = 64;
constexpr int kGcChunk = 32;
constexpr int kGcPad = kGcTile + 4;

template <int PHI>
__device__ __forceinline__ float cost_phi_acc(float acc, float u, float p) {
  if constexpr (PHI == 1) {
    return acc + fabsf(u);
  } else if constexpr (PHI == 2) {
    return fmaf(u, u, acc);
  } else {
    const float a = fabsf(u);
    return acc + (a > 0.f ? exp2f(p * log2f(a)) : (a == 0.f ? 0.f : a));  // NaN stays NaN
  }
}

// alpha a^beta for a >= 0 (NaN propagates)
__device__ __forceinline__ float cost_psi(float a, float alpha, float beta) {
  if (beta == 1.f) return alpha * a;
  if (beta == 0.5f) return alpha * sqrtf(a);
  if (beta == 2.f) return alpha * a * a;
  return alpha * (a > 0.f ? exp2f(beta * log2f(a)) : (a == 0.f ? 0.f : a));
}

// Column (max, sum) of the rows [row_begin, row_end) of the in side against
// the 64 out-side points from col0, into s_ms[64].
template <int PHI>
__device__ __forceinline__ void stream_cost(const HalfStepArgs& p, int64_t col0, int64_t row_begin, int64_t row_end,
                                            float cz, float s, int tid, float2* s_ms,
                                            float (*xs)[kGcPad], float (*ys)[kGcPad], float2 (*part)[kGcTile]) {
  const float* __restrict__ X = static_cast<const float*>(p.in_frag);
  const float* __restrict__ Y = static_cast<const float*>(p.out_frag);
  const int d = p.d;
  const float cp = p.cost_p, alpha = p.cost_alpha, beta = p.cost_beta;
  const int r0 = (tid >> 4) * 4, c0 = (tid & 15) * 4;
  const int64_t n_in = p.n_in, n_out = p.n_out;
  const int64_t rend = row_end < n_in ? row_end : n_in;
  float mx[4], sm[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    mx[b] = -INFINITY;
    sm[b] = 0.f;
  }
  for (int64_t i0 = row_begin; i0 < rend; i0 += kGcTile) {
    float acc[4][4] = {};
    for (int k0 = 0; k0 < d; k0 += kGcChunk) {
      const int kc = min(kGcChunk, d - k0);
      __syncthreads();
      for (int e = tid; e < kGcTile * kGcChunk; e += 256) {
        const int r = e / kGcChunk, k = e % kGcChunk;
        const int64_t i = i0 + r, j = col0 + r;
        xs[k][r] = (k < kc && i < rend) ? X[i * d + k0 + k] : 0.f;
        ys[k][r] = (k < kc && j < n_out) ? Y[j * d + k0 + k] : 0.f;
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
          for (int b = 0; b < 4; ++b) acc[a][b] = cost_phi_acc<PHI>(acc[a][b], xa[a] - ya[b], cp);
      }
    }
    float z[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int64_t i = i0 + r0 + a;
      const float hi = i < rend ? p.in_pot[i] : -INFINITY;
#pragma unroll
      for (int b = 0; b < 4; ++b) z[a][b] = (hi - s * cost_psi(acc[a][b], alpha, beta)) * cz;
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      float cm = nan_max(nan_max(z[0][b], z[1][b]), nan_max(z[2][b], z[3][b]));
      const float nm = nan_max(mx[b], cm);
      if (nm == -INFINITY) continue;  // every term so far is exp(-inf) = 0
      float acc_s = sm[b] * exp2_fast(mx[b] - nm);
#pragma unroll
      for (int a = 0; a < 4; ++a) acc_s += exp2_fast(z[a][b] - nm);
      sm[b] = acc_s;
      mx[b] = nm;
    }
  }
  // the 16 row groups of each column, merged in a fixed order
#pragma unroll
  for (int b = 0; b < 4; ++b) part[tid >> 4][c0 + b] = make_float2(mx[b], sm[b]);
  __syncthreads();
  if (tid < kGcTile) {
    float M = -INFINITY, S = 0.f;
    for (int g = 0; g < 16; ++g) lse_merge(M, S, part[g][tid].x, part[g][tid].y);
    s_ms[tid] = make_float2(M, S);
  }
}

// ---------------------------------------------------------------------------
// One workgroup of a half-step: column tile ``tile`` of the output side and
// row split ``split`` of the input side.  W waves each own NB column blocks;
// the last split to arrive on a column tile merges the partial (max, sum)
// pairs and runs the epilogue.
// ---------------------------------------------------------------------------
template <int PREC, int KK, int NB, int MT, int W>
__device__ __forceinline__ void halfstep_block(const HalfStepArgs& p, uint32_t mode, const int tile,
                                               const int split, const int tid) {
  if ((mode & kModeCheckDone) && p.state->done) return;
  constexpr int NT = W * 64;
  constexpr int COLS = W * NB * 16;
  constexpr int TILE_U4 = MT * FragChunks<PREC, KK>::value * 64;
  __shared__ u32x4 s_tile[2 * TILE_U4];
  __shared__ float s_bias[2 * MT * 16];
  __shared__ float2 s_ms[COLS];
  __shared__ double s_red[2][NT];
  __shared__ int s_last;

  __shared__ float s_mm[W][2];

  const int lane = tid & 63;
  const int wave = tid >> 6;
  float c = p.c, eps = p.eps, eps_marg = p.eps_marg, inv_rho = p.inv_rho;
  if (p.scal != nullptr) {
    c = p.scal->c;
    eps = eps_marg = p.scal->eps_f;
    inv_rho = p.side ? p.scal->inv_rho_b : p.scal->inv_rho_a;
  }
  // the exponent constant the y fragments were packed with
  const float c_pack = p.pack_scal != nullptr ? p.pack_scal->c : p.c;
  if (p.eps_sched) {
    // epsilon schedule (epsilon_scheduler.py:47-60): this launch's iteration
    // t from the block index, eps_t = max(init decay^t, 1) eps.  While eps_t >
    // eps the marginal error, taken at the target eps, cannot come from this
    // pass: the error pass that follows the g-update measures it instead.
    const int t = p.state->block * p.inner + p.it_offset;
    const double scale = fmax(p.eps_init * pow(p.eps_decay, t), 1.0);
    if (p.err_pass) {
      if (!(scale > 1.0)) return;  // the g-update measured it
    } else {
      if (scale > 1.0) mode &= ~(kModeErrCol | kModeErrRow);
      eps = static_cast<float>(static_cast<double>(eps) * scale);
      c = static_cast<float>(static_cast<double>(c) / scale);
    }
  }
  float momentum = p.momentum;
  if (p.mom_start > 0) {
    // adaptive momentum: iterations from mom_start on use Lehmann's weight,
    // taken from the errors of blocks idx - 2 and idx - 1 (idx = mom_start /
    // inner).  The g-update ending block idx - 1 is itself iteration
    // mom_start when inner divides it; its block's error is not final yet, so
    // it writes the plain update and momentum_fixup_kernel blends it after
    // the block has been finalized.
    const int t = p.state->block * p.inner + p.it_offset;
    if (t >= p.mom_start)
      momentum = (p.mom_start / p.inner - 1 == p.state->block)
                     ? 1.f
                     : adaptive_momentum(p.errors, p.mom_start, p.inner, p.momentum, p.mom_threshold);
  }

  // Exponent units: acc = c_pack (<x, y> + beta); this launch's exponent is
  // k acc, k = 1 exactly at the packing epsilon; the row bias is
  // c_pack beta = (f - norm) kb.
  const float k = c / c_pack;
  const float kb = static_cast<float>(static_cast<double>(c_pack) / p.two_s);

  const int rb_begin = split * p.rb_per_split;
  const int rb_end = min(rb_begin + p.rb_per_split, p.n_in_rb);
  if constexpr (PREC == 2) {
    static_assert(COLS == kGcTile && NT == 256, "the cost stream covers 64 columns with 256 threads");
    __shared__ __align__(16) float s_xs[kGcChunk][kGcPad];
    __shared__ __align__(16) float s_ys[kGcChunk][kGcPad];
    __shared__ float2 s_part[16][kGcTile];
    (void)kb;
    stream_cost<KK>(p, static_cast<int64_t>(tile) * COLS, static_cast<int64_t>(rb_begin) * 16,
                    static_cast<int64_t>(rb_end) * 16, static_cast<float>(static_cast<double>(c) / p.two_s),
                    static_cast<float>(0.5 * p.two_s), tid, s_ms, s_xs, s_ys, s_part);
  } else {
  OutOperand<PREC, KK, NB> B;
  B.load(p.out_frag, static_cast<int64_t>(tile * W + wave) * NB, lane);
  float mx[NB], sm[NB];
  const int64_t cbase = static_cast<int64_t>(tile * W + wave) * NB * 16 + (lane & 15);
  // Fixed-shift streaming when the previous half-step of this side left a
  // finite log-sum-exp for every column of the tile; the workgroup falls back
  // to the running max if the resulting sums leave [2^-60, 2^60].  When the
  // workgroup's shifts all lie within kUniformRange of their largest S (and
  // k = 1), every column is summed relative to S itself: S goes into the row
  // bias and a pair costs no operation before its exp2; the sums then only
  // need to stay above 2^-100.  The exponents are then rounded at magnitudes
  // up to the range, and that rounding changes from pass to pass with S: it
  // sets a floor under the marginal error of about ln 2 ulp(range) / 2, 3.3e-7
  // for a range of 8 (at 64 the floor rises to 2.6e-6, visible against the
  // oracle's errors).  The fp32 path keeps per-column shifts.
  constexpr float kUniformRange = PREC == 1 ? 8.f : -1.f;
  bool shifted = p.lse_shift != nullptr && p.state != nullptr && p.state->shift_ok[p.side] != 0;
  bool uniform = false;
  float wg_shift = 0.f;
  if (shifted) {
    if (tid == 0) s_last = 1;
    __syncthreads();
    bool ok = true;
    float lmax = -INFINITY, lmin = INFINITY;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      mx[nb] = k * p.lse_shift[cbase + nb * 16];
      sm[nb] = 0.f;
      ok = ok && isfinite(mx[nb]);
      lmax = fmaxf(lmax, mx[nb]);
      lmin = fminf(lmin, mx[nb]);
    }
    if (!ok) s_last = 0;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      lmax = fmaxf(lmax, __shfl_xor(lmax, off, 64));
      lmin = fminf(lmin, __shfl_xor(lmin, off, 64));
    }
    if (lane == 0) {
      s_mm[wave][0] = lmax;
      s_mm[wave][1] = lmin;
    }
    __syncthreads();
    shifted = s_last != 0;
    if (shifted && k == 1.f) {
      float M = s_mm[0][0], m = s_mm[0][1];
#pragma unroll
      for (int w = 1; w < W; ++w) {
        M = fmaxf(M, s_mm[w][0]);
        m = fminf(m, s_mm[w][1]);
      }
      uniform = M - m <= kUniformRange;
      wg_shift = M;
    }
    __syncthreads();
  }
  if (shifted) {
    if (uniform) {
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) mx[nb] = wg_shift;
      stream_any<PREC, KK, NB, MT, W, true, kExpNone>(B, p.in_frag, p.in_pot, p.in_norm, p.n_in, rb_begin, rb_end,
                                                      kb, wg_shift, k, tid, mx, sm, s_tile, s_bias);
    } else if (k == 1.f) {
      stream_any<PREC, KK, NB, MT, W, true, kExpSub>(B, p.in_frag, p.in_pot, p.in_norm, p.n_in, rb_begin, rb_end,
                                                     kb, 0.f, k, tid, mx, sm, s_tile, s_bias);
    } else {
      stream_any<PREC, KK, NB, MT, W, true, kExpFma>(B, p.in_frag, p.in_pot, p.in_norm, p.n_in, rb_begin, rb_end,
                                                     kb, 0.f, k, tid, mx, sm, s_tile, s_bias);
    }
    combine_lanes<NB>(mx, sm, lane, wave, s_ms);
    if (tid == 0) s_last = 1;
    __syncthreads();
    const bool has_rows = static_cast<int64_t>(rb_begin) * 16 < p.n_in;
    const float lo = uniform ? 7.888609e-31f : 8.673617e-19f;  // 2^-100 : 2^-60
    for (int t = tid; t < COLS; t += NT) {
      const float S = s_ms[t].y;
      if (has_rows && !(S >= lo && S <= 1.1529215e18f)) s_last = 0;
    }
    __syncthreads();
    shifted = s_last != 0;
    __syncthreads();
  }
  if (!shifted) {
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      mx[nb] = -INFINITY;
      sm[nb] = 0.f;
    }
    stream_any<PREC, KK, NB, MT, W, false, kExpFma>(B, p.in_frag, p.in_pot, p.in_norm, p.n_in, rb_begin, rb_end,
                                                    kb, 0.f, k, tid, mx, sm, s_tile, s_bias);
    combine_lanes<NB>(mx, sm, lane, wave, s_ms);
  }
  }  // PREC 0 / 1
  __syncthreads();

  const int64_t col0 = static_cast<int64_t>(tile) * COLS;
  if (p.splits > 1) {
    const int64_t stride = static_cast<int64_t>(p.tiles) * COLS;
    float2* part = p.partial + static_cast<int64_t>(split) * stride;
    for (int t = tid; t < COLS; t += NT) store_f2_through(part + col0 + t, s_ms[t]);
    wait_own_stores();
    __syncthreads();
    if (tid == 0) {
      const uint32_t prev = atomicAdd(p.tile_counter + tile, 1u);
      s_last = (prev == static_cast<uint32_t>(p.splits - 1)) ? 1 : 0;
    }
    __syncthreads();
    if (!s_last) return;
    for (int t = tid; t < COLS; t += NT) {
      // the splits' pairs loaded eight at a time, one round trip per eight
      float M = -INFINITY, S = 0.f;
      for (int sp0 = 0; sp0 < p.splits; sp0 += 8) {
        float2 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
          v[u] = (sp0 + u < p.splits) ? load_f2_coherent(p.partial + (sp0 + u) * stride + col0 + t)
                                      : make_float2(-INFINITY, 0.f);
#pragma unroll
        for (int u = 0; u < 8; ++u) lse_merge(M, S, v[u].x, v[u].y);
      }
      s_ms[t] = make_float2(M, S);
    }
    if (tid == 0) p.tile_counter[tile] = 0u;
    __syncthreads();
  }

  if (p.lse_store != nullptr) {
    // stored in packing units (independent of this launch's k)
    const float inv_k = 1.f / k;
    for (int t = tid; t < COLS; t += NT) {
      const float2 ms = s_ms[t];
      p.lse_store[col0 + t] = (ms.y > 0.f) ? (ms.x + log2f(ms.y)) * inv_k : -INFINITY;
    }
    if (tid == 0 && p.state != nullptr) p.state->shift_ok[p.side] = 1;
  }
  const EpiScalars es{mode,     p.norm_p, eps, eps_marg, p.tau, inv_rho, momentum,
                      p.two_s / (static_cast<double>(k) * static_cast<double>(c_pack))};
  const bool want_err = (mode & (kModeErrCol | kModeErrRow)) != 0;
  double e_acc = 0.0, m_acc = 0.0;
  for (int t = tid; t < COLS; t += NT) {
    const int64_t j = col0 + t;
    if (j >= p.n_out) continue;
    column_epilogue(es, s_ms[t], j, p.out_norm, p.out_pot, p.out_w, p.out_logw, p.out_save, p.out_vec, e_acc,
                    m_acc);
  }

  if (want_err) {
    block_sum2<NT>(s_red, tid, e_acc, m_acc);
    if (tid == 0) {
      store_f64_through(p.err_partial + tile, s_red[0][0]);
      if (p.mass_partial) store_f64_through(p.mass_partial + tile, s_red[1][0]);
    }
    if ((mode & kModeErrCol) && p.state != nullptr) {
      if (tid == 0) {
        wait_own_stores();  // the partials above are stored by this thread
        const uint32_t prev = atomicAdd(&p.state->tiles_done, 1u);
        s_last = (prev == static_cast<uint32_t>(p.tiles - 1)) ? 1 : 0;
      }
      __syncthreads();
      if (s_last) {
        // the last tile: every partial of the pass loaded at once (a few per
        // thread) and summed in a fixed-order tree, then the block's finalizer
        double col = 0.0, mass = 0.0, row = 0.0;
        for (int t = tid; t < p.tiles; t += NT) {
          col += load_f64_coherent(p.err_partial + t);
          if (p.mass_partial) mass += load_f64_coherent(p.mass_partial + t);
        }
        if (!p.balanced)
          for (int t = tid; t < p.row_tiles; t += NT) row += p.row_err_partial[t];
        block_sum2<NT>(s_red, tid, col, mass);
        col = s_red[0][0];
        mass = s_red[1][0];
        __syncthreads();
        if (!p.balanced) {
          block_sum2<NT>(s_red, tid, row, 0.0);
          row = s_red[0][0];
        }
        if (tid == 0) {
          finalize_block(p, col, row, mass);
          p.state->tiles_done = 0u;
        }
      }
    }
  }
}

// Single-problem half-step: grid (column tiles, row splits).
template <int PREC, int KK, int NB, int MT, int W>
__global__ __launch_bounds__(W * 64) void halfstep_kernel(const HalfStepArgs p) {
  // Workgroups reach the 8 XCDs round robin in dispatch order (x fastest), and
  // each XCD has its own L2.  In grid order an XCD would take every split of
  // its tiles: all of the streamed cloud and 1/8 of the other.  Here XCD k
  // takes half of x and a quarter of y, the same parts in both directions
  // (g-update: split group k % 2 of x's rows and tile group k / 2 of y's
  // columns; f-update: split group k / 2 of y's rows and tile group k % 2 of
  // x's), 3/8 of the fragments instead of 9/16.  Any bijection is correct: the
  // grouping only moves which L2 serves a workgroup.  Grids below 512
  // workgroups keep grid order: their problems fit every L2, and at n = m =
  // 1024, d = 2 (256 workgroups) the remapped order measured 2 % slower.
  int tile = blockIdx.x, split = blockIdx.y;
  const int T = gridDim.x, S = gridDim.y;
  const bool g_dir = p.side == 1;
  const int sg_n = g_dir ? 2 : 4, tg_n = 8 / sg_n;
  if (T * S >= 512 && S % sg_n == 0 && T % tg_n == 0) {
    const int b = blockIdx.x + blockIdx.y * T, k = b & 7, local = b >> 3;
    const int Tg = T / tg_n;
    split = (g_dir ? (k & 1) : (k >> 1)) * (S / sg_n) + local / Tg;
    tile = (g_dir ? (k >> 1) : (k & 1)) * Tg + local % Tg;
  }
  halfstep_block<PREC, KK, NB, MT, W>(p, p.mode, tile, split, threadIdx.x);
}

// ---------------------------------------------------------------------------
// Transport applied to vectors (geometry.py:445-486): out[v][j] = sum_i P_ij
// V[i][v] with P_ij = exp((h_i + h_j - C_ij) / eps), h_i the in side's and h_j
// the out side's potential, for 16 vectors per launch, signs and all.  The
// exponents come from the half-step's MFMA cross term; with the column's
// log-sum-exp L_j (exponent units, from a half-step pass) the weights
// exp2(acc_ij - L_j) lie in [0, 1] and sum to one over i, the products with
// V run on fp32 MFMA (16x16x4: for r < 4, lane l gives P[4 (l / 16) + r][l %
// 16] as A and V[4 (l / 16) + r][l % 16] as B), and the column factor
// exp((L_j 2 s / c + h_j - s|y_j|^2) / eps) is applied once at the end.
// V arrives in fragments of 16 rows: [rb][lane = 16 q + v][r] = V[16 rb + 4 q + r][v].
// ---------------------------------------------------------------------------
struct ApplyArgs {
  HalfStepArgs h;      // in / out sides, potentials, norms, plan; lse_shift = the column log-sum-exp
  const void* v_frag;  // the 16 vectors' fragments over the in side's rows
  float* o_part;       // [splits][n_out_pad][16] partial sums
  float* out;          // out[v * out_stride + j]
  int64_t out_stride;
  int32_t kvec;        // vectors of this launch (<= 16)
};

template <int PREC, int KK, int NB, int MT, int W>
__global__ __launch_bounds__(W * 64) void apply_pv_kernel(const ApplyArgs A) {
  const HalfStepArgs& p = A.h;
  constexpr int NT = W * 64;
  constexpr int COLS = W * NB * 16;
  constexpr int CH = FragChunks<PREC, KK>::value;
  constexpr int TILE_U4 = MT * CH * 64;
  constexpr int VT_U4 = MT * 64;
  constexpr int ROWS = MT * 16;
  __shared__ u32x4 s_tile[TILE_U4];
  __shared__ u32x4 s_v[VT_U4];
  __shared__ float s_bias[ROWS];
  __shared__ int s_last;
  const int tile = blockIdx.x, split = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const float c_pack = p.pack_scal != nullptr ? p.pack_scal->c : p.c;
  const float c = p.scal != nullptr ? p.scal->c : p.c;
  const float eps = p.scal != nullptr ? p.scal->eps_f : p.eps;
  const float k = c / c_pack;
  const float kb = static_cast<float>(static_cast<double>(c_pack) / p.two_s);

  OutOperand<PREC, KK, NB> B;
  B.load(p.out_frag, static_cast<int64_t>(tile * W + wave) * NB, lane);
  const int64_t cbase = static_cast<int64_t>(tile * W + wave) * NB * 16 + (lane & 15);
  float m[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const float L = p.lse_shift[cbase + nb * 16] * k;
    m[nb] = isfinite(L) ? L : 0.f;  // -inf: every weight of the column is 0 already
  }
  f32x4 O[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) O[nb] = f32x4{0.f, 0.f, 0.f, 0.f};

  const u32x4* __restrict__ src = static_cast<const u32x4*>(p.in_frag);
  const u32x4* __restrict__ vsrc = static_cast<const u32x4*>(A.v_frag);
  const int rb_begin = split * p.rb_per_split;
  const int rb_end = min(rb_begin + p.rb_per_split, p.n_in_rb);
  for (int rb = rb_begin; rb < rb_end; rb += MT) {
    for (int q = tid; q < TILE_U4; q += NT) s_tile[q] = src[static_cast<size_t>(rb) * CH * 64 + q];
    for (int q = tid; q < VT_U4; q += NT) s_v[q] = vsrc[static_cast<size_t>(rb) * 64 + q];
    if (tid < ROWS) {
      const int64_t row = static_cast<int64_t>(rb) * 16 + tid;
      const int64_t i = min(row, p.n_in - 1);
      s_bias[tid] = row_bias(make_float2(p.in_pot[i], p.in_norm[i]), row, p.n_in, kb, 0.f);
    }
    __syncthreads();
    InTile<PREC, KK, MT> Ain;
    Ain.load(s_tile, s_bias, lane);
    f32x4 acc[MT][NB];
    tile_mfma<PREC, KK, NB, MT>(B, Ain, acc);
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const float4 vv = __builtin_bit_cast(float4, s_v[mt * 64 + lane]);
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        O[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(exp2_fast(fmaf(acc[mt][nb][0], k, -m[nb])), vv.x, O[nb], 0, 0, 0);
        O[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(exp2_fast(fmaf(acc[mt][nb][1], k, -m[nb])), vv.y, O[nb], 0, 0, 0);
        O[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(exp2_fast(fmaf(acc[mt][nb][2], k, -m[nb])), vv.z, O[nb], 0, 0, 0);
        O[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(exp2_fast(fmaf(acc[mt][nb][3], k, -m[nb])), vv.w, O[nb], 0, 0, 0);
      }
    }
    __syncthreads();
  }

  // Split-K: partials stored write-through, the last split to arrive sums them.
  const int64_t n_out_pad = static_cast<int64_t>(p.tiles) * COLS;
  const int64_t col0 = static_cast<int64_t>(tile) * COLS;
#pragma unroll
  for (int nb = 0; nb < NB; ++nb)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int64_t j = col0 + (wave * NB + nb) * 16 + 4 * (lane >> 4) + t;
      __hip_atomic_store(reinterpret_cast<uint32_t*>(A.o_part + (split * n_out_pad + j) * 16 + (lane & 15)),
                         __float_as_uint(O[nb][t]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  wait_own_stores();
  __syncthreads();
  if (tid == 0) {
    const uint32_t prev = atomicAdd(p.tile_counter + tile, 1u);
    s_last = (prev == static_cast<uint32_t>(p.splits - 1)) ? 1 : 0;
  }
  __syncthreads();
  if (!s_last) return;
  for (int e = tid; e < COLS * 16; e += NT) {
    const int64_t j = col0 + e / 16;
    const int v = e % 16;
    float s = 0.f;
    for (int sp = 0; sp < p.splits; ++sp)
      s += __uint_as_float(__hip_atomic_load(reinterpret_cast<const uint32_t*>(A.o_part + (sp * n_out_pad + j) * 16 + v),
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    if (j >= p.n_out || v >= A.kvec) continue;
    const float L = p.lse_shift[j];
    const float h = p.out_pot[j];
    float r;
    if (L == -INFINITY || h == -INFINITY) {
      r = (s != s) ? s : 0.f;
    } else {
      const double ex = (static_cast<double>(L) * p.two_s / static_cast<double>(c_pack) + static_cast<double>(h) -
                         static_cast<double>(p.out_norm[j])) / static_cast<double>(eps);
      r = static_cast<float>(static_cast<double>(s) * exp(ex));
    }
    A.out[static_cast<int64_t>(v) * A.out_stride + j] = r;
  }
  if (tid == 0) p.tile_counter[tile] = 0u;
}

// Batched half-step over many independent problems.  Workgroup b runs item
// (problem, column tile, row split) = work[b].x, .y, .z of the live work
// list, which holds the items of unfinished problems only, longest first (it
// is rebuilt between blocks); workgroups past its length (*count) exit at
// once.  ``mode`` replaces the per-problem mode.
template <int PREC, int KK, int NB, int MT, int W>
__global__ __launch_bounds__(W * 64) void halfstep_batched_kernel(const HalfStepArgs* __restrict__ args,
                                                                  const int4* __restrict__ work,
                                                                  const uint32_t* __restrict__ count, uint32_t mode) {
  if (blockIdx.x >= *count) return;
  const int4 w = work[blockIdx.x];
  halfstep_block<PREC, KK, NB, MT, W>(args[w.x], mode, w.y, w.z, threadIdx.x);
}

}  // namespace ottx

// Shared device/host definitions of the ottx library (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ottx.h"

namespace ottx {

typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4;

constexpr float kLog2e = 1.4426950408889634f;
constexpr float kLn2 = 0.6931471805599453f;
constexpr int kWave = 64;

// Rows of every packed point array are padded to this multiple, so that any
// streamed row tile and any output column tile stays inside the allocation.
constexpr int kRowPad = 512;

// Half-step mode bits (uniform per launch).
enum : uint32_t {
  kModeWritePot = 1u << 0,   // write the (momentum-blended) new potential
  kModeSaveOld = 1u << 1,    // copy the old out-side potential to out_save
  kModeErrCol = 1u << 2,     // accumulate the column error + mass, finalize the block
  kModeErrRow = 1u << 3,     // accumulate the row error of the new potential
  kModeApply = 1u << 4,      // out_vec = apply_lse_kernel
  kModeMarginal = 1u << 5,   // out_vec = marginal_from_potentials
  kModeCheckDone = 1u << 6,  // return at once when state->done is set
  kModeWriteApp = 1u << 7,   // out_vec = eps LSE_i((h_i - C_ij) / eps) as it is (recentred updates)
  kModeErrOld = 1u << 8,     // kModeErrRow measures the old potential (parallel updates)
};

// Batched solves: batch_done[0] counts finished problems, [1] completed
// blocks, [2] finished problems at the last rebuild of the live work lists;
// bit k of the words from kBatchDoneBits on is set when problem k finished.
constexpr int kBatchDoneBits = 4;

// Loop state shared by the kernels of one solve (device memory).
struct SolveState {
  int32_t done;        // set by the block finalizer when the loop must stop
  int32_t block;       // index of the next block whose error is finalized
  int32_t n_blocks;    // blocks run when the loop stopped
  int32_t finalized;   // the outputs of the stopped loop have been written
  uint32_t tiles_done; // arrival counter of the column tiles of an error pass
  int32_t shift_ok[2]; // a log-sum-exp per column exists for side 0 (f) / 1 (g)
  uint32_t fin_count;  // arrival counter of the output finalizer's workgroups
  double total_mass;   // sum of the coupling at the last error pass
  double last_err;     // last raw error (before the inf/record rule)
};

// Results of a solve, written by the device into pinned host memory once the
// loop has stopped (``ready`` last, equal to the solve's sequence number).
struct SolveMail {
  double du[8];        // div_a, sum a, div_b, sum b, #NaN errors, #(errors != -1), errors[-1], eps
  double total_mass;
  int32_t n_blocks;
  uint32_t ready;
  uint64_t progress;   // (seq << 32) | blocks completed, written by every block's finalizer
  double last_err;     // error of the last completed block, written just before progress
};

// Scalars resolved on the device at solve start (relative epsilon).
struct SolveScalars {
  double eps;        // target epsilon
  float eps_f;       // float copy used by the kernels
  float c;           // 2 s log2(e) / eps
  float inv_rho_a;   // (1 - tau_a) / (eps tau_a), 0 when tau_a == 1
  float inv_rho_b;
};

// Recentred unbalanced iterations: log sum_i w_i exp(-pot_i / rho) of f (0,
// weights a, rho_a) and of g (1, weights b, rho_b) at their latest update.
struct RcScalars {
  double lse[2];
};

// Arguments of one half-step launch: the out side gets the new potential, the
// in side is streamed.  For the g-update out = y (m columns), in = x (n rows).
struct HalfStepArgs {
  const void* in_frag;
  const float* in_pot;
  const float* in_norm;
  const void* out_frag;
  float* out_pot;       // read (old) and written (new) when kModeWritePot
  float* out_save;      // kModeSaveOld target
  float* out_vec;       // kModeApply / kModeMarginal target
  const float* out_norm;
  const float* out_w;   // weights of the out side
  const float* out_logw; // optional log-marginal replacing log(out_w) in the update
  int64_t n_in;
  int64_t n_out;
  int32_t n_in_rb;      // padded in rows / 16
  int32_t rb_per_split; // row blocks streamed by one split (multiple of MT)
  int32_t splits;
  int32_t tiles;        // column tiles of this launch
  uint32_t mode;
  int32_t norm_p;
  float inv2s;          // 1 / (2 s)
  double two_s;         // 2 s
  float c;              // 2 s log2(e) / eps (the launch's; the y side of the fragments carries the packing c)
  float eps;            // epsilon of this launch's exponent
  float eps_marg;       // epsilon of the marginal (target epsilon)
  float tau;            // tau of the out side
  float inv_rho;        // 1 / rho_out (0 when tau == 1)
  float momentum;       // constant momentum weight
  int32_t balanced;
  int32_t symmetric;    // symmetric problem with parallel updates: error = 2 x column error
  int32_t eps_sched;    // epsilon schedule: eps_t = max(eps_init eps_decay^t, 1) eps
  int32_t err_pass;     // the schedule's error pass (target eps, after the g-update)
  int32_t it_offset;    // iteration of this launch = state->block * inner + it_offset
  double eps_init;
  double eps_decay;
  int32_t mom_start;    // adaptive momentum from this iteration on (0: constant momentum)
  float mom_threshold;  // ... if the error of block mom_start / inner - 1 is below this
  int32_t side;         // 1: out = y (tau_b), 0: out = x (tau_a)
  const SolveScalars* scal;  // when set, overrides c, eps, eps_marg and inv_rho
  const SolveScalars* pack_scal;  // c the out fragments were packed with (null: c above)
  const float* lse_shift;    // per-column log-sum-exp (exponent units at the packing c) of the previous same-side pass
  float* lse_store;          // where this pass stores its per-column log-sum-exp
  float2* partial;      // [splits][n_out_pad] running (max, sum) per column
  uint32_t* tile_counter;   // [tiles]
  double* err_partial;  // [tiles] column (or row) error partials
  double* mass_partial; // [tiles]
  // block finalizer (kModeErrCol)
  SolveState* state;
  float* errors;
  const double* row_err_partial;
  int32_t row_tiles;
  int32_t inner;
  int32_t min_it;
  int32_t max_it;
  int32_t outer;
  int32_t force_scan;
  float threshold;
  SolveMail* mail;       // single solves: the block finalizer publishes the loop's progress here
  uint32_t seq;          // ... tagged with the solve's sequence number
  uint32_t* batch_done;  // batched solves: count of problems whose loop finished, then one bit per problem
  int32_t problem;       // batched solves: index of the problem (its bit in batch_done + kBatchDoneBits)
  // costs other than the squared Euclidean (PREC 2): in_frag / out_frag are
  // the raw points [rows, d]; a pair's cost is s alpha (sum_k phi(z_k))^beta
  int32_t d;
  float cost_p;         // exponent of phi(z) = |z|^p
  float cost_alpha;
  float cost_beta;
};

}  // namespace ottx

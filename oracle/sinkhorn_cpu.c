/*
 * FP64 CPU oracle kernel: the log-sum-exp operator of OTT's PointCloud
 * geometry, with the squared-Euclidean cost recomputed on the fly.
 *
 * TEST INFRASTRUCTURE ONLY.  Loaded through ctypes by oracle/sinkhorn64.py
 * (PointCloudLSE), which is imported only by tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py.  The product library never links it.
 *
 * ox_lse restates geometry/geometry.py:423-442 (``_center`` then ``_softmax``
 * with vec=None) for C_ij = s * ||x_i - y_j||^2 (geometry/costs.py:554-561,
 * pointcloud.py:334-340), in the online form of pointcloud.py:85-103: the cost
 * block is recomputed for each block of points and reduced with a max-shifted
 * log-sum-exp (math/utils.py:132-138), so the n x m matrix is never stored.
 *
 *   axis = 0:  out[j] = eps * LSE_i((f_i + g_j - C_ij) / eps),  j < m
 *   axis = 1:  out[i] = eps * LSE_j((f_i + g_j - C_ij) / eps),  i < n
 *
 * An all -inf slice gives -inf, as jax.scipy.special.logsumexp does.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <omp.h>

#define OX_OUT_BLOCK 8
#define OX_IN_BLOCK 256

static void lse_rows(const double* out_pts, const double* out_pot, int64_t n_out,
                     const double* in_pts, const double* in_pot, int64_t n_in,
                     int32_t d, double scale, double eps, double* out) {
  const double inv_eps = 1.0 / eps;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t j0 = 0; j0 < n_out; j0 += OX_OUT_BLOCK) {
    const int64_t jn = (n_out - j0) < OX_OUT_BLOCK ? (n_out - j0) : OX_OUT_BLOCK;
    double run_max[OX_OUT_BLOCK], run_sum[OX_OUT_BLOCK];
    double z[OX_OUT_BLOCK][OX_IN_BLOCK];
    for (int64_t jj = 0; jj < jn; ++jj) {
      run_max[jj] = -INFINITY;
      run_sum[jj] = 0.0;
    }
    for (int64_t i0 = 0; i0 < n_in; i0 += OX_IN_BLOCK) {
      const int64_t in = (n_in - i0) < OX_IN_BLOCK ? (n_in - i0) : OX_IN_BLOCK;
      for (int64_t jj = 0; jj < jn; ++jj) {
        const double* q = out_pts + (j0 + jj) * (int64_t)d;
        const double hq = out_pot[j0 + jj];
        double bmax = -INFINITY;
        for (int64_t ii = 0; ii < in; ++ii) {
          const double* p = in_pts + (i0 + ii) * (int64_t)d;
          double c = 0.0;
          for (int32_t k = 0; k < d; ++k) {
            const double t = p[k] - q[k];
            c += t * t;
          }
          const double v = (in_pot[i0 + ii] + hq - scale * c) * inv_eps;
          z[jj][ii] = v;
          if (v > bmax) bmax = v;
        }
        if (bmax == -INFINITY) continue;
        double mx = run_max[jj];
        if (bmax > mx) {
          if (mx != -INFINITY) run_sum[jj] *= exp(mx - bmax);
          mx = bmax;
          run_max[jj] = mx;
        }
        double s = 0.0;
        for (int64_t ii = 0; ii < in; ++ii) s += exp(z[jj][ii] - mx);
        run_sum[jj] += s;
      }
    }
    for (int64_t jj = 0; jj < jn; ++jj) {
      out[j0 + jj] = (run_max[jj] == -INFINITY)
                         ? -INFINITY
                         : eps * (run_max[jj] + log(run_sum[jj]));
    }
  }
}

int ox_lse(const double* x, const double* y, int64_t n, int64_t m, int32_t d,
           double scale, const double* f, const double* g, double eps, int axis,
           double* out, int nthreads) {
  if (n <= 0 || m <= 0 || d <= 0 || !(eps > 0.0)) return -1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  if (axis == 0) {
    lse_rows(y, g, m, x, f, n, d, scale, eps, out);
  } else if (axis == 1) {
    lse_rows(x, f, n, y, g, m, d, scale, eps, out);
  } else {
    return -1;
  }
  return 0;
}

int ox_max_threads(void) { return omp_get_max_threads(); }

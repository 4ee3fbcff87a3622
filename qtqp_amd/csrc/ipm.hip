// Device-resident interior-point driver for one dense QP (MI355X, gfx950).
//
// qtqp's iteration loop (src/qtqp/__init__.py:473-578) with the KKT engine of
// kkt_engine.h underneath: the iterate (x, y, tau, s), K^{-1}q, the Newton
// directions and every scalar of an iteration (mu, tau+, alpha, sigma, the
// termination quantities) live in HBM.
//
// An iteration is one fixed sequence of kernel launches, captured once in a
// HIP graph and replayed: one graph launch and one record read back per IPM
// iteration.  What the host loop decided between launches is decided on the
// device instead, in the control block `Ctl`:
//  - the factor precision (an FP32 chain and an FP64 chain are both in the
//    sequence; a gate word lets one of them run, and an FP32 breakdown opens
//    the FP64 one);
//  - each refinement loop of DirectKktSolver.solve (direct.py:243-291): the
//    sequence holds a fixed number of unrolled steps, each switched on by the
//    loop's gate while the loop is still running, and a one-thread kernel
//    after every step applies the reference's stopping rules to the FP64
//    residual norm the step left in device memory;
//  - the tau method (__init__.py:741-752) from the refinement result.
// When an iteration needs what the sequence does not hold (more refinement
// steps than unrolled, an FP64 refactor after a slowly contracting FP32
// step, or a retry of an FP64 factor that broke down), the control block
// says so; the host restores the iteration's saved start (x, y, s, tau,
// K^{-1}q) and runs that iteration again with the host-driven launches of
// the same kernels, which loop as long as the reference does.
//
// Each vector phase is a grid-stride kernel whose workgroups write partial
// sums / maxima / minima to `vpart` (kSlots per workgroup), followed by a
// one-workgroup "fin" kernel that combines them in a fixed order and applies
// the scalar logic of the reference, storing the result in the device
// scalars `Scal`.  Reductions are deterministic for a given problem size.
//
// Termination uses the equilibrated data: with A_eq = D A E, P_eq = E P E and
// the equilibrated iterate, A x_u = (A_eq x)/d, A' y_u = (A_eq' y)/e and
// P x_u = (P_eq x)/e, while b and c are the unscaled vectors.
#include <cstddef>

#include "kkt_engine.h"
#include "qtqp_ipm.h"

using namespace qk;

namespace {

constexpr double kEps = 1e-15;  // _EPS (__init__.py:40)
constexpr int kSlots = 24;      // partial results per workgroup
constexpr int kMaxVecBlocks = 240;

// Scalars of the iteration in progress (device memory).
struct Scal {
  double tau, scale, mu, sigma, alpha_p, alpha, step, tau_p, tau_c;
  double mean_sy, mean_soy;
  double norm_b, norm_c;  // ||b||_inf, ||c||_inf of the unscaled data
};

__device__ inline void put_sum(double v, double* part, int k, double* red) {
  v = block_sum(v, red);
  if (threadIdx.x == 0) part[blockIdx.x * kSlots + k] = v;
}
__device__ inline void put_max(double v, double* part, int k, double* red) {
  v = block_max(v, red);
  if (threadIdx.x == 0) part[blockIdx.x * kSlots + k] = v;
}
__device__ inline void put_min(double v, double* part, int k, double* red) {
  v = -block_max(-v, red);
  if (threadIdx.x == 0) part[blockIdx.x * kSlots + k] = v;
}
__device__ inline double fin_sum(const double* part, int nb, int k, double* red) {
  double v = 0.0;
  for (int b = threadIdx.x; b < nb; b += WG) v += part[b * kSlots + k];
  return block_sum(v, red);
}
__device__ inline double fin_max(const double* part, int nb, int k, double* red) {
  double v = -INFINITY;
  for (int b = threadIdx.x; b < nb; b += WG) v = fmax(v, part[b * kSlots + k]);
  return block_max(v, red);
}
__device__ inline double fin_min(const double* part, int nb, int k, double* red) {
  double v = INFINITY;
  for (int b = threadIdx.x; b < nb; b += WG) v = fmin(v, part[b * kSlots + k]);
  return -block_max(-v, red);
}
// K partial results of this workgroup into slots k0.. of part, combined across
// the workgroup exactly as put_sum / put_max would, with one pair of barriers;
// bit k of MAXMASK marks a maximum.  red: >= 4 K doubles of LDS.
template <int K, unsigned MAXMASK = 0u>
__device__ inline void put_n(double (&v)[K], double* part, int k0, double* red) {
  block_reduce_n<K, MAXMASK>(v, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) part[blockIdx.x * kSlots + k0 + k] = v[k];
  }
}
// The K slots k0.. over nb workgroups, as fin_sum / fin_max would combine them.
template <int K, unsigned MAXMASK = 0u>
__device__ inline void fin_n(const double* part, int nb, int k0, double (&v)[K], double* red) {
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = ((MAXMASK >> k) & 1u) ? -INFINITY : 0.0;
  for (int b = threadIdx.x; b < nb; b += WG) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const double x = part[b * kSlots + k0 + k];
      v[k] = ((MAXMASK >> k) & 1u) ? fmax(v[k], x) : v[k] + x;
    }
  }
  block_reduce_n<K, MAXMASK>(v, red);
}

// |v| for an inf-norm; NaN is reported as inf (fmax drops NaN).
__device__ inline double absn(double v) { return v != v ? INFINITY : fabs(v); }

#define GRID_STRIDE(k, count) \
  for (int64_t k = blockIdx.x * (int64_t)WG + threadIdx.x; k < (count); k += (int64_t)gridDim.x * WG)

// True in every thread of the last workgroup of the grid to get here, which
// then sees the partial results every other workgroup wrote before it; that
// workgroup resets *ctr for the next launch.  Lets one launch both write
// per-workgroup partials and combine them (the "part" and "fin" halves of a
// reduction below).
__device__ inline bool last_block(unsigned* ctr) {
  __shared__ unsigned sh_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    sh_last = atomicAdd(ctr, 1u) == gridDim.x - 1 ? 1u : 0u;
    if (sh_last) {
      *ctr = 0u;
      __threadfence();
    }
  }
  __syncthreads();
  return sh_last != 0u;
}

// ---------------------------------------------------------------------------
// Equilibration (__init__.py:616-663) and start (__init__.py:299-307)
// ---------------------------------------------------------------------------

// Workgroup b takes rows b*RB .. +RB of the row-major M (rows x ncol): the
// row inf-norms into rowmax (if given) and the column maxima over its rows
// into cpart[b * ncol + j].
__global__ __launch_bounds__(WG) void k_ruiz_rows(const double* __restrict__ M, int rows, int ncol,
                                                  double* __restrict__ rowmax,
                                                  double* __restrict__ cpart) {
  __shared__ double red[RB][4];
  const int r0 = blockIdx.x * RB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double rm[RB];
#pragma unroll
  for (int q = 0; q < RB; ++q) rm[q] = 0.0;
  for (int j = tid; j < ncol; j += WG) {
    double cm = 0.0;
#pragma unroll
    for (int q = 0; q < RB; ++q) {
      if (r0 + q < rows) {
        const double a = fabs(M[(int64_t)(r0 + q) * ncol + j]);
        rm[q] = fmax(rm[q], a);
        cm = fmax(cm, a);
      }
    }
    cpart[(int64_t)blockIdx.x * ncol + j] = cm;
  }
  if (rowmax) {
#pragma unroll
    for (int q = 0; q < RB; ++q) {
      const double v = wave_max(rm[q]);
      if (lane == 0) red[q][wave] = v;
    }
    __syncthreads();
    if (tid < RB && r0 + tid < rows)
      rowmax[r0 + tid] = fmax(fmax(red[tid][0], red[tid][1]), fmax(red[tid][2], red[tid][3]));
  }
}

__device__ inline double ruiz_factor(double nrm, double lo, double hi) {
  // np.clip(1 / sqrt(where(nrm == 0, 1, nrm)), lo, hi)
  const double v = 1.0 / sqrt(nrm == 0.0 ? 1.0 : nrm);
  return fmin(fmax(v, lo), hi);
}

// One pass's scaling factors: d_i from the row norms of A, e_i from the
// column norms of A and P (P = P_off + diag(pdiag)); d *= d_i, e *= e_i.
__global__ __launch_bounds__(WG) void k_ruiz_factors(int m, int n, const double* __restrict__ rowmax,
                                                     const double* __restrict__ cpa, int nba,
                                                     const double* __restrict__ cpp, int nbp,
                                                     const double* __restrict__ pdiag,
                                                     double* __restrict__ di, double* __restrict__ ei,
                                                     double* __restrict__ d, double* __restrict__ e) {
  GRID_STRIDE(k, (int64_t)n + m) {
    if (k < n) {
      double col = fabs(pdiag[k]);
      for (int b = 0; b < nba; ++b) col = fmax(col, cpa[(int64_t)b * n + k]);
      for (int b = 0; b < nbp; ++b) col = fmax(col, cpp[(int64_t)b * n + k]);
      const double f = ruiz_factor(col, 1e-3, 1e3);
      ei[k] = f;
      e[k] *= f;
    } else {
      const int i = (int)(k - n);
      const double f = ruiz_factor(rowmax[i], 1e-3, 1e3);
      di[i] = f;
      d[i] *= f;
    }
  }
}

// M[i][j] *= (fr[i] * fc[j])  (the reference's a.data *= d_i[rows] * e_i[cols]);
// workgroups take rows in turn, threads stride along the row.
__global__ void k_scale_rows_cols(double* __restrict__ M, int rows, int ncol,
                                  const double* __restrict__ fr, const double* __restrict__ fc) {
  for (int i = blockIdx.x; i < rows; i += gridDim.x) {
    double* row = M + (int64_t)i * ncol;
    const double f = fr[i];
    for (int j = threadIdx.x; j < ncol; j += WG) row[j] *= f * fc[j];
  }
}

__global__ void k_scale_pdiag(double* __restrict__ pdiag, int n, const double* __restrict__ ei) {
  GRID_STRIDE(k, n) pdiag[k] *= ei[k] * ei[k];
}

// q = [c*e; b*d]; default start x = 0, y[z:] = 1/d, s[z:] = d (the
// equilibrated image of y = s = 1), tau = 1; K^{-1}q = 0.
__global__ void k_start(int n, int m, int z, const double* __restrict__ b,
                        const double* __restrict__ c, const double* __restrict__ d,
                        const double* __restrict__ e, double* __restrict__ q,
                        double* __restrict__ xy, double* __restrict__ s, double* __restrict__ kq,
                        Scal* S, double norm_b, double norm_c) {
  GRID_STRIDE(k, (int64_t)n + m) {
    if (k < n) {
      q[k] = c[k] * e[k];
      xy[k] = 0.0 / e[k];
    } else {
      const int i = (int)(k - n);
      q[k] = b[i] * d[i];
      const double one = i < z ? 0.0 : 1.0;
      xy[k] = one / d[i];
      s[i] = one * d[i];
    }
    kq[k] = 0.0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S->tau = 1.0;
    S->norm_b = norm_b;
    S->norm_c = norm_c;
  }
}

// ---------------------------------------------------------------------------
// One iteration
// ---------------------------------------------------------------------------

// ||x||^2, ||y||^2, y's  (__init__.py:859-880, 477).
__device__ inline void norm_part(int n, int m, const double* __restrict__ xy,
                                                  const double* __restrict__ s, double* part) {
  __shared__ double red[4];
  double xx = 0.0, yy = 0.0, ys = 0.0;
  GRID_STRIDE(k, (int64_t)n + m) {
    const double v = xy[k];
    if (k < n) {
      xx += v * v;
    } else {
      yy += v * v;
      ys += v * s[k - n];
    }
  }
  put_sum(xx, part, 0, red);
  put_sum(yy, part, 1, red);
  put_sum(ys, part, 2, red);
}

// scale = sqrt(m - z + 1) / max(eps, ||(x, y, tau)||), tau *= scale,
// mu = (scale y)'(scale s) / (m - z).
__device__ inline void norm_fin(int m, int z, const double* part, int nb, Scal* S) {
  __shared__ double red[4];
  const double xx = fin_sum(part, nb, 0, red);
  const double yy = fin_sum(part, nb, 1, red);
  const double ys = fin_sum(part, nb, 2, red);
  if (threadIdx.x == 0) {
    const double tau = S->tau;
    const double scale = sqrt((double)(m - z + 1)) / fmax(kEps, sqrt(xx + yy + tau * tau));
    S->scale = scale;
    S->tau = tau * scale;
    S->mu = ys * (scale * scale) / (m - z);
  }
}

// Applies the normalisation and forms the regularised KKT diagonal and its
// correction (DirectKktSolver.update, direct.py:196-216):
//   t = [diag(P) + mu; mu (z rows); s/y + mu],  reg = max(t, min_reg),
//   diag = [reg_x; -reg_y],  corr = diag - [t_x; -t_y].
__global__ void k_scale_diag(int n, int m, int z, double* __restrict__ xy, double* __restrict__ s,
                             const double* __restrict__ pdiag, double min_reg, const Scal* S,
                             double* __restrict__ diag, double* __restrict__ corr) {
  const double sc = S->scale, mu = S->mu;
  GRID_STRIDE(k, (int64_t)n + m) {
    double t;
    if (k < n) {
      xy[k] *= sc;
      t = pdiag[k] + mu;
    } else {
      const int i = (int)(k - n);
      const double yv = xy[k] * sc, sv = s[i] * sc;
      xy[k] = yv;
      s[i] = sv;
      t = i < z ? mu : sv / yv + mu;
    }
    const double rg = fmax(t, min_reg);
    if (k < n) {
      diag[k] = rg;
      corr[k] = rg - t;
    } else {
      diag[k] = -rg;
      corr[k] = -rg - (-t);
    }
  }
}

// out0 = P v0 (and out1 = P v1), P = P_off + diag(pdiag), one wave per row.
template <int NV>
__global__ __launch_bounds__(WG) void k_pmul(const double* __restrict__ Poff,
                                             const double* __restrict__ pdiag, int n,
                                             const double* __restrict__ v0, double* __restrict__ o0,
                                             const double* __restrict__ v1, double* __restrict__ o1) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (WG / 64) + (threadIdx.x >> 6);
  if (row >= n) return;
  const double* pr = Poff + (int64_t)row * n;
  double a0 = 0.0, a1 = 0.0;
  for (int j = lane; j < n; j += 64) {
    const double p = pr[j];
    a0 += p * v0[j];
    if (NV == 2) a1 += p * v1[j];
  }
  a0 = wave_sum(a0);
  if (NV == 2) a1 = wave_sum(a1);
  if (lane == 0) {
    o0[row] = a0 + pdiag[row] * v0[row];
    if (NV == 2) o1[row] = a1 + pdiag[row] * v1[row];
  }
}

// Right-hand side of a Newton step (__init__.py:722-731):
//   r = (mu - mu_t) anchor;  r[n+z:] += mu_t / y (if mu_t != 0) + s + corr.
// Predictor (corrector = 0): mu_t = 0, anchor = [x; y], copied into xp.
// Corrector: mu_t = sigma mu, anchor = xp = [x_p; y_p], and the Mehrotra term
// corr = -d_s (y_p - y) / y (__init__.py:528) is formed into cm.
// kr (optional) gets a copy of the anchor: the refined solve's warm start.
__global__ void k_newton_rhs(int n, int m, int z, int corrector, const double* __restrict__ xy,
                             const double* __restrict__ s, const double* __restrict__ ds,
                             double* __restrict__ xp, double* __restrict__ cm,
                             double* __restrict__ r, const Scal* S, double* __restrict__ kr) {
  const double mu = S->mu;
  const double mu_t = corrector ? S->sigma * mu : 0.0;
  GRID_STRIDE(k, (int64_t)n + m) {
    double a;
    if (corrector) {
      a = xp[k];
    } else {
      a = xy[k];
      xp[k] = a;
    }
    if (kr) kr[k] = a;
    double v = (mu - mu_t) * a;
    if (k >= n) {
      const int i = (int)(k - n);
      if (i >= z) {
        const double yv = xy[k];
        if (mu_t != 0.0) v += mu_t / yv;
        v += s[i];
        if (corrector) {
          const double c = -ds[i] * (a - yv) / yv;
          cm[i] = c;
          v += c;
        }
      } else if (corrector) {
        cm[i] = 0.0;
      }
    }
    r[k] = v;
  }
}

// Dot products of the tau equations (__init__.py:759-857):
//   0 kq.q  1 kr.q  2 kq_x.pq  3 kr_x.pq  4 kq_x.pr  5 kr_x.pr
//   6 q_x.x  7 q_y.y  8 x.px  9 px.kq_x  10 px.kr_x
// with pq = P kq_x, pr = P kr_x, px = P x (slots 2-5, 8-10 only with P).
__device__ inline void tau_part(int n, int m, int has_p, const double* __restrict__ q,
                                                 const double* __restrict__ kq,
                                                 const double* __restrict__ kr,
                                                 const double* __restrict__ xy,
                                                 const double* __restrict__ pq,
                                                 const double* __restrict__ pr,
                                                 const double* __restrict__ px, double* part) {
  __shared__ double red[4 * 11];
  double v[11];
#pragma unroll
  for (int t = 0; t < 11; ++t) v[t] = 0.0;
  GRID_STRIDE(k, (int64_t)n + m) {
    const double qk = q[k], kqk = kq[k], krk = kr[k], xk = xy[k];
    v[0] += kqk * qk;
    v[1] += krk * qk;
    if (k < n) {
      v[6] += qk * xk;
      if (has_p) {
        v[2] += kqk * pq[k];
        v[3] += krk * pq[k];
        v[4] += kqk * pr[k];
        v[5] += krk * pr[k];
        v[8] += xk * px[k];
        v[9] += px[k] * kqk;
        v[10] += px[k] * krk;
      }
    } else {
      v[7] += qk * xk;
    }
  }
  put_n<11>(v, part, 0, red);
}

// tau+ of a Newton step: the stable root of t_a tau^2 + t_b tau + t_c = 0
// (_solve_for_tau, __init__.py:759-817) when `quadratic` is allowed and the
// root is valid, else the linearised step with its [0.1, 10] trust region
// (_solve_for_tau_linearized_fallback, __init__.py:819-857).
// quad_res: when given, the refined solve's (status, residual) on the device,
// which decide `quadratic` as the host would (__init__.py:741-752).
__device__ inline void tau_fin(int corrector, int quadratic, int has_p,
                                                const double* part, int nb, Scal* S,
                                                qtqp_ipm_record* rec, const int* quad_status,
                                                const double* quad_res) {
  __shared__ double red[4 * 11];
  double v[11];
  fin_n<11>(part, nb, 0, v, red);
  if (threadIdx.x != 0) return;
  if (quad_status) quadratic = *quad_status == 0 || *quad_res < 1e-7;
  const double mu = S->mu, tau = S->tau;
  const double mu_t = corrector ? S->sigma * mu : 0.0;
  const double tau_anchor = corrector ? S->tau_p : tau;
  const double kq_q = v[0], kr_q = v[1];
  bool have = false;
  double tp = 0.0;
  if (quadratic) {
    const double r_tau = (mu - mu_t) * tau_anchor;
    double t_a = mu + kq_q;
    double t_b = -r_tau - kr_q;
    double t_c = -mu_t;
    if (has_p) {
      t_a -= v[2];
      t_b += v[3] + v[4];
      t_c -= v[5];
    }
    if (fabs(t_a) < kEps) {
      if (!(fabs(t_b) < kEps)) {
        const double root = -t_c / t_b;
        if (isfinite(root) && !(root < -1e-10)) {
          tp = fmax(0.0, root);
          have = true;
        }
      }
    } else {
      const double disc = t_b * t_b - 4 * t_a * t_c;
      if (!(disc < -1e-9)) {
        const double sq = sqrt(fmax(0.0, disc));
        const double root = t_b > 0 ? t_c / (-0.5 * (t_b + sq)) : (-0.5 * (t_b - sq)) / t_a;
        if (isfinite(root) && !(root < -1e-10)) {
          tp = fmax(0.0, root);
          have = true;
        }
      }
    }
  }
  if (!have) {
    const double x_px = has_p ? v[8] : 0.0, px_kq = has_p ? v[9] : 0.0, px_kr = has_p ? v[10] : 0.0;
    const double q_z = v[6] + v[7];
    const double q_kq = kq_q;
    const double q_rz = kr_q - tau * q_kq - q_z;
    const double px_rz = px_kr - tau * px_kq - x_px;
    const double g = mu * tau * tau + (mu_t - mu) * tau_anchor * tau - tau * q_z - mu_t - x_px;
    const double num = g - tau * q_rz - 2.0 * px_rz;
    const double den = 2.0 * mu * tau + (mu_t - mu) * tau_anchor - q_z + tau * q_kq + 2.0 * px_kq;
    const double tn = tau + (fabs(den) < 1e-16 ? 0.0 : -num / den);
    tp = !isfinite(tn) ? tau : fmin(fmax(tn, 0.1 * tau), 10.0 * tau);
  }
  if (corrector) S->tau_c = tp;
  else S->tau_p = tp;
  rec->tau_method[corrector] = have ? 1 : 0;
}

// [x+; y+] = kr - kq tau+  (__init__.py:755-757).
__global__ void k_newton_fin(int64_t count, int corrector, const double* __restrict__ kr,
                             const double* __restrict__ kq, double* __restrict__ xp, const Scal* S) {
  const double tp = corrector ? S->tau_c : S->tau_p;
  GRID_STRIDE(k, count) xp[k] = kr[k] - kq[k] * tp;
}

// d_s over the inequality rows and the fraction-to-boundary minima
// (__init__.py:507-510, 546-547, 671-680):
//   predictor: d_s = -y_p s / y;  corrector: d_s = sigma mu / y + corr - y_c s / y;
//   slot 0: min -s/d_s over d_s < -eps;  slot 1: min -y/d_y over d_y < -eps.
__device__ inline void step_part(int n, int m, int z, int corrector,
                                                  const double* __restrict__ xy,
                                                  const double* __restrict__ s,
                                                  const double* __restrict__ xp,
                                                  const double* __restrict__ cm,
                                                  double* __restrict__ ds, const Scal* S,
                                                  double* part) {
  __shared__ double red[4];
  const double smu = corrector ? S->sigma * S->mu : 0.0;
  double mn_s = INFINITY, mn_y = INFINITY;
  GRID_STRIDE(i, m) {
    double d = 0.0;
    if (i >= z) {
      const double yv = xy[n + i], sv = s[i], yn = xp[n + i];
      d = corrector ? smu / yv + cm[i] - yn * sv / yv : -yn * sv / yv;
      if (d < -kEps) mn_s = fmin(mn_s, -sv / d);
      const double dy = yn - yv;
      if (dy < -kEps) mn_y = fmin(mn_y, -yv / dy);
    }
    ds[i] = d;
  }
  put_min(mn_s, part, 0, red);
  put_min(mn_y, part, 1, red);
}

// alpha = min(max step on s, max step on y) (__init__.py:882-886); for the
// corrector also step = scale * alpha and tau = max(tau + step (tau_c - tau), 1e-30).
__device__ inline void alpha_fin(int corrector, double step_scale,
                                                  const double* part, int nb, Scal* S) {
  __shared__ double red[4];
  const double ms = fin_min(part, nb, 0, red), my = fin_min(part, nb, 1, red);
  if (threadIdx.x != 0) return;
  const double as = ms == INFINITY ? 1.0 : fmin(1.0, ms);
  const double ay = my == INFINITY ? 1.0 : fmin(1.0, my);
  const double alpha = fmin(as, ay);
  if (!corrector) {
    S->alpha_p = alpha;
    return;
  }
  S->alpha = alpha;
  const double step = step_scale * alpha;
  S->step = step;
  S->tau = fmax(S->tau + step * (S->tau_c - S->tau), 1e-30);
}

// Affine trial point norms for sigma (__init__.py:682-706).
__device__ inline void sigma_part(int n, int m, const double* __restrict__ xy,
                                                   const double* __restrict__ s,
                                                   const double* __restrict__ xp,
                                                   const double* __restrict__ ds, const Scal* S,
                                                   double* part) {
  __shared__ double red[4];
  const double al = S->alpha_p;
  double sx = 0.0, sy = 0.0, sys = 0.0;
  GRID_STRIDE(k, (int64_t)n + m) {
    const double v = xy[k];
    const double va = v + al * (xp[k] - v);
    if (k < n) {
      sx += va * va;
    } else {
      const int i = (int)(k - n);
      const double sa = s[i] + al * ds[i];
      sy += va * va;
      sys += va * sa;
    }
  }
  put_sum(sx, part, 0, red);
  put_sum(sy, part, 1, red);
  put_sum(sys, part, 2, red);
}

__device__ inline void sigma_fin(int m, int z, const double* part, int nb, Scal* S) {
  __shared__ double red[4];
  const double sx = fin_sum(part, nb, 0, red), sy = fin_sum(part, nb, 1, red);
  const double sys = fin_sum(part, nb, 2, red);
  if (threadIdx.x != 0) return;
  const double tau = S->tau, al = S->alpha_p;
  const double ta = tau + al * (S->tau_p - tau);
  const double nrm2 = sx + sy + ta * ta;
  const int k = m - z;
  const double mu_aff = (k + 1) / fmax(kEps * kEps, nrm2) * sys / k;
  const double base = mu_aff / fmax(kEps, S->mu);
  S->sigma = fmin(fmax(base * base * base, 0.0), 1.0);
}

// x += step (x_c - x), y += step (y_c - y), s += step d_s, clamps
// (__init__.py:556-565).
__global__ void k_update(int n, int m, int z, double* __restrict__ xy, double* __restrict__ s,
                         const double* __restrict__ xp, const double* __restrict__ ds,
                         const Scal* S) {
  const double step = S->step;
  GRID_STRIDE(k, (int64_t)n + m) {
    const double v = xy[k];
    double nv = v + step * (xp[k] - v);
    if (k >= n) {
      const int i = (int)(k - n);
      double sv = s[i] + step * ds[i];
      if (i >= z) {
        nv = fmax(nv, 1e-30);
        sv = fmax(sv, 1e-30);
      }
      s[i] = sv;
    }
    xy[k] = nv;
  }
}

// Slots of the termination reduction that are maxima (the rest are sums).
constexpr unsigned kTermMax = (((1u << 15) - 1u) & ~0xFu) | (1u << 16) | (1u << 17);

// Termination quantities on the unscaled problem (__init__.py:888-965).
// ax = A_eq x (m), aty = A_eq' y (n), pxe = P_eq x (n).
//   sums: 0 x'Px  1 c'x  2 b'y  3 y's     maxima: 4 |Px| 5 |A'y| 6 dres
//   7 |x| 8 |Ax| 9 |s| 10 pres 11 |Ax+s| 12 |y|
//   collect_stats, rows z..m of sy = s*y, soy = s/max(eps, y):
//   13 max sy 14 min sy 15 sum sy 16 max soy 17 min soy 18 sum soy
__device__ inline void term_part(int n, int m, int z, int has_p, int collect,
                                                  const double* __restrict__ xy,
                                                  const double* __restrict__ s,
                                                  const double* __restrict__ d,
                                                  const double* __restrict__ e,
                                                  const double* __restrict__ ax,
                                                  const double* __restrict__ aty,
                                                  const double* __restrict__ pxe,
                                                  const double* __restrict__ b,
                                                  const double* __restrict__ c, const Scal* S,
                                                  double* part) {
  __shared__ double red[4 * 19];
  const double inv_tau = 1.0 / fmax(S->tau, kEps);
  double xpx = 0.0, ctx = 0.0, bty = 0.0, ys = 0.0;
  double n_px = 0.0, n_aty = 0.0, dres = 0.0, n_x = 0.0;
  double n_ax = 0.0, n_s = 0.0, pres = 0.0, n_axs = 0.0, n_y = 0.0;
  double sy_max = -INFINITY, sy_min = INFINITY, sy_sum = 0.0;
  double so_max = -INFINITY, so_min = INFINITY, so_sum = 0.0;
  GRID_STRIDE(k, (int64_t)n + m) {
    if (k < n) {
      const double xu = e[k] * xy[k];
      const double pxu = has_p ? pxe[k] / e[k] : 0.0;
      const double atyu = aty[k] / e[k];
      xpx += xu * pxu;
      ctx += c[k] * xu;
      n_px = fmax(n_px, absn(pxu));
      n_aty = fmax(n_aty, absn(atyu));
      dres = fmax(dres, absn((pxu + atyu) * inv_tau + c[k]));
      n_x = fmax(n_x, absn(xu));
    } else {
      const int i = (int)(k - n);
      const double yu = d[i] * xy[k], su = s[i] / d[i], axu = ax[i] / d[i];
      bty += b[i] * yu;
      ys += yu * su;
      n_ax = fmax(n_ax, absn(axu));
      n_s = fmax(n_s, absn(su));
      pres = fmax(pres, absn((axu + su) * inv_tau - b[i]));
      n_axs = fmax(n_axs, absn(axu + su));
      n_y = fmax(n_y, absn(yu));
      if (collect && i >= z) {
        const double sy = su * yu, so = su / fmax(kEps, yu);
        sy_max = fmax(sy_max, sy);
        sy_min = fmin(sy_min, sy);
        sy_sum += sy;
        so_max = fmax(so_max, so);
        so_min = fmin(so_min, so);
        so_sum += so;
      }
    }
  }
  // minima (slots 14, 17) travel negated as maxima
  double v[19] = {xpx,   ctx,   bty,  ys,      n_px,    n_aty,   dres,   n_x,     n_ax, n_s,
                  pres,  n_axs, n_y,  sy_max,  -sy_min, sy_sum,  so_max, -so_min, so_sum};
  put_n<19, kTermMax>(v, part, 0, red);
}

struct Tol {
  double atol, rtol, atol_inf, rtol_inf;
};

// Costs, residuals, certificates and the status (__init__.py:905-1008).
__device__ inline void term_fin(int m, int z, int has_p, int collect, Tol tl,
                                                 const double* part, int nb, Scal* S,
                                                 qtqp_ipm_record* rec) {
  __shared__ double red[4 * 19];
  double v[19];
  fin_n<19, kTermMax>(part, nb, 0, v, red);
  v[14] = -v[14];
  v[17] = -v[17];
  if (threadIdx.x != 0) return;
  const double tau = S->tau;
  const double inv_tau = 1.0 / fmax(tau, kEps);
  const double xpx = has_p ? v[0] : 0.0, ctx = v[1], bty = v[2];
  const double norm_px = v[4], norm_aty = v[5], dres = v[6], norm_x = v[7];
  const double norm_ax = v[8], norm_s = v[9], pres = v[10], norm_axs = v[11], norm_y = v[12];
  const double pcost = (ctx + 0.5 * xpx * inv_tau) * inv_tau;
  const double dcost = (-bty - 0.5 * xpx * inv_tau) * inv_tau;
  const double gap = fabs((ctx + bty + xpx * inv_tau) * inv_tau);
  const double dinfeas_a = norm_axs / (fabs(ctx) + kEps);
  const double dinfeas_p = norm_px / (fabs(ctx) + kEps);
  const double dinfeas = fmax(dinfeas_a, dinfeas_p);
  const double pinfeas = norm_aty / (fabs(bty) + kEps);
  const double prelrhs = fmax(fmax(norm_ax * inv_tau, norm_s * inv_tau), S->norm_b);
  const double drelrhs = fmax(fmax(norm_px * inv_tau, norm_aty * inv_tau), S->norm_c);
  int status = QTQP_IPM_UNFINISHED;
  if (gap < tl.atol + tl.rtol * fmin(fabs(pcost), fabs(dcost)) &&
      pres < tl.atol + tl.rtol * prelrhs && dres < tl.atol + tl.rtol * drelrhs)
    status = QTQP_IPM_SOLVED;
  else if (ctx < -kEps && dinfeas < tl.atol_inf + tl.rtol_inf * norm_x / fabs(ctx))
    status = QTQP_IPM_UNBOUNDED;
  else if (bty < -kEps && pinfeas < tl.atol_inf + tl.rtol_inf * norm_y / fabs(bty))
    status = QTQP_IPM_INFEASIBLE;
  rec->status = status;
  rec->ctx = ctx;
  rec->bty = bty;
  rec->pcost = pcost;
  rec->dcost = dcost;
  rec->pres = pres;
  rec->dres = dres;
  rec->gap = gap;
  rec->pinfeas = pinfeas;
  rec->dinfeas = dinfeas;
  rec->dinfeas_a = dinfeas_a;
  rec->dinfeas_p = dinfeas_p;
  rec->mu = S->mu;
  rec->sigma = S->sigma;
  rec->alpha = S->alpha;
  rec->tau = tau;
  rec->norm_x = norm_x;
  rec->norm_y = norm_y;
  rec->prelrhs = prelrhs;
  rec->drelrhs = drelrhs;
  if (collect) {
    const int k = m - z;
    rec->complementarity = fabs(v[3] * inv_tau * inv_tau);
    rec->norm_s = norm_s;
    rec->max_sy = v[13];
    rec->min_sy = v[14];
    rec->max_s_over_y = v[16];
    rec->min_s_over_y = v[17];
    S->mean_sy = v[15] / k;
    S->mean_soy = v[18] / k;
    rec->mean_s_over_y = S->mean_soy;
  }
}

// Second pass of the standard deviations (np.std: sqrt(mean(|v - mean|^2))).
__device__ inline void std_part(int n, int m, int z, const double* __restrict__ xy,
                                                 const double* __restrict__ s,
                                                 const double* __restrict__ d, const Scal* S,
                                                 double* part) {
  __shared__ double red[4];
  const double msy = S->mean_sy, mso = S->mean_soy;
  double a = 0.0, b = 0.0;
  GRID_STRIDE(i, m) {
    if (i >= z) {
      const double yu = d[i] * xy[n + i], su = s[i] / d[i];
      const double dsy = su * yu - msy, dso = su / fmax(kEps, yu) - mso;
      a += dsy * dsy;
      b += dso * dso;
    }
  }
  put_sum(a, part, 0, red);
  put_sum(b, part, 1, red);
}

__device__ inline void std_fin(int m, int z, const double* part, int nb,
                                                qtqp_ipm_record* rec) {
  __shared__ double red[4];
  const double a = fin_sum(part, nb, 0, red), b = fin_sum(part, nb, 1, red);
  if (threadIdx.x != 0) return;
  rec->std_sy = sqrt(a / (m - z));
  rec->std_s_over_y = sqrt(b / (m - z));
}


// Each reduction in one launch: the per-workgroup partials, then (in the last
// workgroup) the combination and the scalar logic, in the same order as two
// launches would.
__global__ __launch_bounds__(WG) void k_norm(int n, int m, int z, const double* __restrict__ xy,
                                             const double* __restrict__ s, double* part, Scal* S,
                                             unsigned* ctr) {
  norm_part(n, m, xy, s, part);
  if (last_block(ctr)) norm_fin(m, z, part, gridDim.x, S);
}

__global__ __launch_bounds__(WG) void k_tau(int n, int m, int has_p, const double* __restrict__ q,
                                            const double* __restrict__ kq,
                                            const double* __restrict__ kr,
                                            const double* __restrict__ xy,
                                            const double* __restrict__ pq,
                                            const double* __restrict__ pr,
                                            const double* __restrict__ px, double* part,
                                            int corrector, int quadratic, Scal* S,
                                            qtqp_ipm_record* rec, const int* quad_status,
                                            const double* quad_res, unsigned* ctr) {
  tau_part(n, m, has_p, q, kq, kr, xy, pq, pr, px, part);
  if (last_block(ctr))
    tau_fin(corrector, quadratic, has_p, part, gridDim.x, S, rec, quad_status, quad_res);
}

__global__ __launch_bounds__(WG) void k_step_alpha(int n, int m, int z, int corrector,
                                                   const double* __restrict__ xy,
                                                   const double* __restrict__ s,
                                                   const double* __restrict__ xp,
                                                   const double* __restrict__ cm,
                                                   double* __restrict__ ds, Scal* S, double* part,
                                                   double step_scale, unsigned* ctr) {
  step_part(n, m, z, corrector, xy, s, xp, cm, ds, S, part);
  if (last_block(ctr)) alpha_fin(corrector, step_scale, part, gridDim.x, S);
}

__global__ __launch_bounds__(WG) void k_sigma(int n, int m, int z, const double* __restrict__ xy,
                                              const double* __restrict__ s,
                                              const double* __restrict__ xp,
                                              const double* __restrict__ ds, Scal* S, double* part,
                                              unsigned* ctr) {
  sigma_part(n, m, xy, s, xp, ds, S, part);
  if (last_block(ctr)) sigma_fin(m, z, part, gridDim.x, S);
}

__global__ __launch_bounds__(WG) void k_term(int n, int m, int z, int has_p, int collect,
                                             const double* __restrict__ xy,
                                             const double* __restrict__ s,
                                             const double* __restrict__ d,
                                             const double* __restrict__ e,
                                             const double* __restrict__ ax,
                                             const double* __restrict__ aty,
                                             const double* __restrict__ pxe,
                                             const double* __restrict__ b,
                                             const double* __restrict__ c, Scal* S, double* part,
                                             Tol tl, qtqp_ipm_record* rec, unsigned* ctr) {
  term_part(n, m, z, has_p, collect, xy, s, d, e, ax, aty, pxe, b, c, S, part);
  if (last_block(ctr)) term_fin(m, z, has_p, collect, tl, part, gridDim.x, S, rec);
}

__global__ __launch_bounds__(WG) void k_std(int n, int m, int z, const double* __restrict__ xy,
                                            const double* __restrict__ s,
                                            const double* __restrict__ d, Scal* S, double* part,
                                            qtqp_ipm_record* rec, unsigned* ctr) {
  std_part(n, m, z, xy, s, d, S, part);
  if (last_block(ctr)) std_fin(m, z, part, gridDim.x, rec);
}

// ---------------------------------------------------------------------------
// Device control of an iteration replayed from a graph
// ---------------------------------------------------------------------------
// An FP32 step contracting the residual by less than this asks for FP64 (the
// same bound as the host-driven refinement's); kF32StepsPerSolve
// (kkt_engine.h) bounds the mean FP32 steps per refined solve.
constexpr double kPoorContraction = 0.25;

// One refinement loop (DirectKktSolver.solve, direct.py:243-307).
struct IrCtl {
  unsigned long long norm[2];  // ||res||_inf bits and NaN flag of the last residual
  unsigned long long rhs_bits; // ||krhs||_inf bits
  double res, old, tol, rhs_norm;
  int status, solves, nan, pad;
  Gate gate;  // active while the loop runs; prec = precision of the factor
};

struct Ctl {
  IrCtl ir[3];  // kinv_q, predictor, corrector
  Gate fgate;   // prec selects the factor chain that runs
  unsigned fin_ctr, pad2;  // workgroups done in a one-launch reduction (last_block)
  int info32, info64;
  // state across iterations (the host sets it after an iteration it ran itself)
  int force64, fprec, prev_steps, prev_fprec;
  // precision probe of a mixed handle: 0 before its first FP32 iteration, 1
  // after it (the next iteration runs FP64); then 2 if the FP64 iteration was
  // faster (FP64 from then on), 3 if the FP32 one was (FP32 kept)
  int probe, pad1;
  unsigned long long t_iter0, t32, t64;
  // this iteration
  int redo;      // 1: refinement needs more steps than unrolled, 2: FP32 contracted
                 // too slowly (FP64 from now on), 3: FP64 factor broke down
  int fact_col;  // breakdown column of the FP64 factor
  int steps, f32_steps;
  // cumulative counters
  long long factorizations, fp64_fallbacks, backend_solves, inner_steps;
  unsigned long long t0, factor_ticks;
};

__device__ inline double bits_to_double(unsigned long long b) { return __longlong_as_double((long long)b); }

// Saves the iteration's start (xy, s, K^{-1}q, scalars) and resets the
// control block; decides the factor precision of this iteration.
__global__ void k_iter_begin(Ctl* c, int mixed, int64_t N, int64_t m, const double* __restrict__ xy,
                             const double* __restrict__ s, const double* __restrict__ kq,
                             const Scal* S, double* __restrict__ xy_sv, double* __restrict__ s_sv,
                             double* __restrict__ kq_sv, Scal* S_sv) {
  GRID_STRIDE(k, N) {
    xy_sv[k] = xy[k];
    kq_sv[k] = kq[k];
    if (k < m) s_sv[k] = s[k];
  }
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  *S_sv = *S;
  for (int r = 0; r < 3; ++r) {
    IrCtl& ir = c->ir[r];
    ir.norm[0] = ir.norm[1] = 0ull;
    ir.rhs_bits = 0ull;
    ir.status = 1;
    ir.solves = 0;
    ir.nan = 0;
    ir.gate.active = 0;
  }
  c->info32 = c->info64 = 0;
  c->redo = 0;
  c->fact_col = 0;
  // A mixed handle times one iteration with the FP32 factor (its first) and
  // the next with the FP64 factor, and keeps the faster: the FP32 factor
  // saves MFMA time but its solves need more refinement steps, and which
  // wins depends on the size and the conditioning.
  int use32 = mixed && !c->force64;
  if (use32 && c->probe == 1) use32 = 0;
  if (use32 && c->probe == 2) {
    use32 = 0;  // FP64 was faster (else the probe would not have ended here)
    c->force64 = 1;
    c->fp64_fallbacks++;
  }
  c->steps = c->f32_steps = 0;
  c->fgate.active = 1;
  c->fgate.prec = use32 ? 32 : 64;
  c->factorizations++;
  c->t0 = wall_clock64();
  c->t_iter0 = c->t0;
}

// After the FP32 chain: a breakdown opens the FP64 chain (and FP64 stays).
__global__ void k_fact_mid(Ctl* c) {
  if (c->fgate.prec == 32 && c->info32 != 0) {
    c->fgate.prec = 64;
    c->force64 = 1;
    c->fp64_fallbacks++;
  }
}

// After the FP64 chain: a breakdown asks the host for the retried factor.
__global__ void k_fact_end(Ctl* c) {
  if (c->fgate.prec == 64 && c->info64 != 0) {
    c->redo = 3;
    c->fact_col = c->info64;
  }
  c->fprec = c->fgate.prec;
  c->factor_ticks += wall_clock64() - c->t0;
}

// Refinement control after the residual of step `step` (0: the initial
// residual of the warm start).  Stops as DirectKktSolver.solve does
// (direct.py:259-283): converged below tol, stalled when the residual stops
// decreasing, or max_steps solves; an FP32 step contracting by less than
// kPoorContraction, or a loop that needs more than the `unrolled` steps in
// the sequence, hands the iteration back to the host.
__device__ inline unsigned long long load_bits(unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void ir_control(Ctl* c, int r, int step, int max_steps, int unrolled,
                                  double atol, double rtol) {
  IrCtl& ir = c->ir[r];
  const double nv = bits_to_double(load_bits(&ir.norm[0]));
  const int nan = load_bits(&ir.norm[1]) != 0ull;
  if (step == 0) {
    ir.rhs_norm = bits_to_double(load_bits(&ir.rhs_bits));
    ir.tol = fma(rtol, ir.rhs_norm, atol);  // one rounding, as the host path computes it
    ir.res = nv;
    ir.nan = nan;
    ir.norm[0] = ir.norm[1] = 0ull;
    ir.gate.prec = c->fprec;
    ir.gate.active = 1;
    return;
  }
  if (!ir.gate.active) return;
  ir.norm[0] = ir.norm[1] = 0ull;
  ir.old = ir.res;
  ir.res = nv;
  ir.nan = nan;
  ir.solves = step;
  c->steps++;
  c->backend_solves++;
  if (ir.gate.prec == 32) {
    c->f32_steps++;
    c->inner_steps++;
  }
  if (ir.res < ir.tol) {
    ir.status = 0;
    ir.gate.active = 0;
  } else if (ir.res >= ir.old) {
    ir.status = 2;
    ir.gate.active = 0;
  } else if (ir.gate.prec == 32 && ir.res > kPoorContraction * ir.old) {
    c->redo = 2;
    c->force64 = 1;
    c->fp64_fallbacks++;
    ir.gate.active = 0;
  } else if (step >= max_steps) {
    ir.gate.active = 0;
  } else if (step >= unrolled) {
    // more steps than the sequence holds; with an FP32 factor that is more
    // than an FP64 factor would need, so FP64 from now on
    c->redo = ir.gate.prec == 32 ? 2 : 1;
    if (ir.gate.prec == 32) {
      c->force64 = 1;
      c->fp64_fallbacks++;
    }
    ir.gate.active = 0;
  }
}

// The residual of a refinement step (k_kmv_finish) and, in the last
// workgroup to finish, the loop's control for that step.
__global__ __launch_bounds__(WG) void k_kmv_finish_ctl(
    int n, int m, const double* __restrict__ x, const double* __restrict__ diag,
    const double* __restrict__ rpp, const double* __restrict__ rpa, int ncb,
    const double* __restrict__ cpart, int nbr, const double* __restrict__ rhs,
    double* __restrict__ krhs, const double* __restrict__ corr, double* __restrict__ out, Ctl* c,
    int r, int step, int max_steps, int unrolled, double atol, double rtol) {
  IrCtl& ir = c->ir[r];
  if (step > 0 && gated_off(&ir.gate, 0)) return;
  kmv_finish(n, m, x, diag, rpp, rpa, ncb, cpart, nbr, rhs, krhs, krhs ? &ir.rhs_bits : nullptr,
             corr, out, ir.norm);
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&c->fin_ctr, 1u) == gridDim.x - 1) {
      c->fin_ctr = 0u;
      __threadfence();
      ir_control(c, r, step, max_steps, unrolled, atol, rtol);
    }
  }
}

// The refined solves' statistics into the record; the state the next
// iteration's precision rule reads.
__global__ void k_iter_end(Ctl* c, qtqp_ipm_record* rec) {
  for (int r = 0; r < 3; ++r) {
    const IrCtl& ir = c->ir[r];
    rec->lin[r].solves = ir.solves < 1 ? 1 : ir.solves;
    rec->lin[r].status = ir.status;
    rec->lin[r].final_residual_norm = ir.res;
    rec->lin[r].rhs_norm = ir.rhs_norm;
    rec->lin[r].tolerance = ir.tol;
  }
  c->prev_steps = c->steps;
  c->prev_fprec = c->fprec;
  const unsigned long long dt = wall_clock64() - c->t_iter0;
  if (c->probe == 0 && c->fprec == 32) {
    c->t32 = dt;
    c->probe = 1;
  } else if (c->probe == 1 && c->fprec == 64) {
    c->t64 = dt;
    c->probe = 2;
    // FP32 faster: keep it (probe 3 = decided, FP32)
    if (c->t32 < c->t64) c->probe = 3;
  }
}

__global__ void k_copy(int64_t count, const double* __restrict__ src, double* __restrict__ dst) {
  GRID_STRIDE(k, count) dst[k] = src[k];
}

// Puts back the start saved by k_iter_begin.
__global__ void k_iter_restore(int64_t N, int64_t m, double* __restrict__ xy, double* __restrict__ s,
                               double* __restrict__ kq, Scal* S, const double* __restrict__ xy_sv,
                               const double* __restrict__ s_sv, const double* __restrict__ kq_sv,
                               const Scal* S_sv) {
  GRID_STRIDE(k, N) {
    xy[k] = xy_sv[k];
    kq[k] = kq_sv[k];
    if (k < m) s[k] = s_sv[k];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *S = *S_sv;
}

}  // namespace

struct qtqp_ipm {
  qtqp_kkt* k = nullptr;  // KKT engine: equilibrated A and P_off, factor, refinement buffers
  int64_t n = 0, m = 0, z = 0;
  int has_p = 0;
  int nv = 1;  // workgroups of the vector kernels
  double *pdiag = nullptr, *b = nullptr, *c = nullptr, *d = nullptr, *e = nullptr;
  double *q = nullptr, *xy = nullptr, *s = nullptr, *kq = nullptr, *xp = nullptr, *kr = nullptr;
  double *r = nullptr, *ds = nullptr, *cm = nullptr, *px = nullptr, *pq = nullptr, *pr = nullptr;
  double *ax = nullptr, *aty = nullptr, *pxe = nullptr;
  double *vpart = nullptr, *cpartp = nullptr, *rowmax = nullptr, *di = nullptr, *ei = nullptr;
  Scal* S = nullptr;
  qtqp_ipm_record* drec = nullptr;
  // graph path: control block, the saved start of the iteration, graphs
  Ctl* ctl = nullptr;
  double *xy_sv = nullptr, *s_sv = nullptr, *kq_sv = nullptr;
  Scal* S_sv = nullptr;
  struct Graph {
    hipGraphExec_t exec = nullptr;
    qtqp_ipm_opts opts{};
    int steps = 0;
  } graphs[2];  // [0]: FP64 factor only, [1]: FP32 and FP64 chains
  int steps_cur = 0;  // refinement steps per solve in the graphs (grows after a redo)
  int wall_khz = 100000;  // device wall clock rate
  int64_t redos = 0;      // iterations handed back to the host-driven launches
};

namespace {

void ipm_release(qtqp_ipm* h) {
  void* bufs[] = {h->pdiag, h->b,  h->c,  h->d,  h->e,   h->q,      h->xy,     h->s,
                  h->kq,    h->xp, h->kr, h->r,  h->ds,  h->cm,     h->px,     h->pq,
                  h->pr,    h->ax, h->aty, h->pxe, h->vpart, h->cpartp, h->rowmax, h->di,
                  h->ei,    h->S,  h->drec, h->ctl, h->xy_sv, h->s_sv, h->kq_sv, h->S_sv};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  for (auto& g : h->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (h->k) {
    release(h->k);
    delete h->k;
  }
}

unsigned pmul_blocks(int64_t n) { return (unsigned)((n + WG / 64 - 1) / (WG / 64)); }

// Device Ruiz equilibration (_equilibrate, __init__.py:616-663), in place on
// the engine's A and P_off and on pdiag; d and e accumulate the scalings.
int ruiz(qtqp_ipm* h) {
  qtqp_kkt* k = h->k;
  const int n = (int)h->n, m = (int)h->m;
  const int nba = nblk_rows(m), nbp = nblk_rows(n);
  for (int pass = 0; pass < 10; ++pass) {
    hipLaunchKernelGGL(k_ruiz_rows, dim3(nba), dim3(WG), 0, k->stream, k->A64, m, n, h->rowmax,
                       k->part);
    hipLaunchKernelGGL(k_ruiz_rows, dim3(nbp), dim3(WG), 0, k->stream, k->Poff64, n, n, nullptr,
                       h->cpartp);
    hipLaunchKernelGGL(k_ruiz_factors, dim3(blocks_for(n + m)), dim3(WG), 0, k->stream, m, n,
                       h->rowmax, k->part, nba, h->cpartp, nbp, h->pdiag, h->di, h->ei, h->d, h->e);
    hipLaunchKernelGGL(k_scale_rows_cols, dim3(2048), dim3(WG), 0, k->stream, k->A64, m, n, h->di,
                       h->ei);
    hipLaunchKernelGGL(k_scale_rows_cols, dim3(1024), dim3(WG), 0, k->stream, k->Poff64, n, n,
                       h->ei, h->ei);
    hipLaunchKernelGGL(k_scale_pdiag, dim3(blocks_for(n)), dim3(WG), 0, k->stream, h->pdiag, n,
                       h->ei);
    HIPC(hipGetLastError());
  }
  return QTQP_OK;
}

__global__ void k_fill(double* v, int64_t count, double val) {
  GRID_STRIDE(k, count) v[k] = val;
}

}  // namespace

extern "C" {

int qtqp_ipm_create(int64_t n, int64_t m, int64_t z, const double* a, const double* p,
                    int32_t col_major, int32_t has_p, const double* b, const double* c,
                    int32_t precision, int32_t equilibrate, qtqp_ipm** out) {
  if (!out) return fail(QTQP_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (!b || !c) return fail(QTQP_ERR_ARG, "b and c are required");
  if (z < 0 || z >= m) return fail(QTQP_ERR_ARG, "need 0 <= z < m");
  if (m > (int64_t)1 << 30) return fail(QTQP_ERR_ARG, "problem too large");
  qtqp_kkt* k = nullptr;
  int rc = engine_create(n, m, a, p, precision, false, &k, col_major != 0);
  if (rc) return rc;
  auto* h = new qtqp_ipm();
  h->k = k;
  h->n = n;
  h->m = m;
  h->z = z;
  h->has_p = has_p ? 1 : 0;
  const int64_t N = n + m;
  h->nv = (int)std::min<int64_t>(kMaxVecBlocks, std::max<int64_t>(1, (N + 4 * WG - 1) / (4 * WG)));
  hipError_t e = hipSuccess;
  auto bad = [&](const char* what) {
    std::string msg = std::string("allocation/copy failed: ") + what + ": " + hipGetErrorString(e);
    ipm_release(h);
    delete h;
    return fail(QTQP_ERR_HIP, msg);
  };
#define ALLOC(field, T, count) \
  if ((h->field = dalloc<T>((count), &e)), e != hipSuccess) return bad(#field)
  ALLOC(pdiag, double, n);
  ALLOC(b, double, m);
  ALLOC(c, double, n);
  ALLOC(d, double, m);
  ALLOC(e, double, n);
  ALLOC(q, double, N);
  ALLOC(xy, double, N);
  ALLOC(s, double, m);
  ALLOC(kq, double, N);
  ALLOC(xp, double, N);
  ALLOC(kr, double, N);
  ALLOC(r, double, N);
  ALLOC(ds, double, m);
  ALLOC(cm, double, m);
  ALLOC(px, double, n);
  ALLOC(pq, double, n);
  ALLOC(pr, double, n);
  ALLOC(ax, double, m);
  ALLOC(aty, double, n);
  ALLOC(pxe, double, n);
  ALLOC(vpart, double, (int64_t)kMaxVecBlocks * kSlots);
  ALLOC(cpartp, double, (int64_t)nblk_rows(n) * n);
  ALLOC(rowmax, double, m);
  ALLOC(di, double, m);
  ALLOC(ei, double, n);
  ALLOC(S, Scal, 1);
  ALLOC(drec, qtqp_ipm_record, 1);
  ALLOC(ctl, Ctl, 1);
  ALLOC(xy_sv, double, N);
  ALLOC(s_sv, double, m);
  ALLOC(kq_sv, double, N);
  ALLOC(S_sv, Scal, 1);
#undef ALLOC
  {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0)
      h->wall_khz = khz;
  }
  const hipStream_t st = k->stream;
  std::vector<double> pd((size_t)n, 0.0);
  if (p)
    for (int64_t i = 0; i < n; ++i) pd[i] = p[i * n + i];
  double norm_b = 0.0, norm_c = 0.0;
  for (int64_t i = 0; i < m; ++i) norm_b = std::max(norm_b, std::fabs(b[i]));
  for (int64_t j = 0; j < n; ++j) norm_c = std::max(norm_c, std::fabs(c[j]));
  if ((e = hipMemcpyAsync(h->pdiag, pd.data(), sizeof(double) * n, hipMemcpyHostToDevice, st)))
    return bad("pdiag");
  if ((e = hipMemcpyAsync(h->b, b, sizeof(double) * m, hipMemcpyHostToDevice, st))) return bad("b");
  if ((e = hipMemcpyAsync(h->c, c, sizeof(double) * n, hipMemcpyHostToDevice, st))) return bad("c");
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(m)), dim3(WG), 0, st, h->d, m, 1.0);
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(n)), dim3(WG), 0, st, h->e, n, 1.0);
  if ((e = hipMemsetAsync(h->S, 0, sizeof(Scal), st))) return bad("scalars");
  if ((e = hipMemsetAsync(h->drec, 0, sizeof(qtqp_ipm_record), st))) return bad("record");
  if ((e = hipMemsetAsync(h->ctl, 0, sizeof(Ctl), st))) return bad("control");
  if ((e = hipGetLastError())) return bad("fill");
  if (equilibrate && (rc = ruiz(h))) {
    ipm_release(h);
    delete h;
    return rc;
  }
  if ((rc = cast_fp32(k))) {
    ipm_release(h);
    delete h;
    return rc;
  }
  hipLaunchKernelGGL(k_start, dim3(blocks_for(N)), dim3(WG), 0, st, (int)n, (int)m, (int)z, h->b,
                     h->c, h->d, h->e, h->q, h->xy, h->s, h->kq, h->S, norm_b, norm_c);
  if ((e = hipGetLastError())) return bad("start");
  if ((e = hipStreamSynchronize(st))) return bad("sync");
  *out = h;
  return QTQP_OK;
}

}  // extern "C"

namespace {

// One iteration with host-driven launches: the host reads the factor's
// breakdown flag and every refinement residual back and loops as the
// reference does.
int iterate_eager(qtqp_ipm* h, const qtqp_ipm_opts* o, qtqp_ipm_record* rec) {
  qtqp_kkt* k = h->k;
  const hipStream_t st = k->stream;
  const int n = (int)h->n, m = (int)h->m, z = (int)h->z;
  const int64_t N = h->n + h->m;
  const int nv = h->nv;
  const dim3 gv(nv), gb(WG), g1(1);
  const unsigned gN = blocks_for(N);
  double* part = h->vpart;
  unsigned* fin = &h->ctl->fin_ctr;
  int rc;
  qtqp_ipm_record out{};

  // normalise, mu, regularised diagonal (__init__.py:474-478, direct.py:183-216)
  hipLaunchKernelGGL(k_norm, gv, gb, 0, st, n, m, z, h->xy, h->s, part, h->S, fin);
  hipLaunchKernelGGL(k_scale_diag, dim3(gN), gb, 0, st, n, m, z, h->xy, h->s, h->pdiag,
                     o->min_static_regularization, h->S, k->diag, k->corr);
  HIPC(hipGetLastError());
  k->have_diag = true;
  k->have_factor = false;
  if ((rc = engine_factorize(k))) return rc;

  const int steps = o->max_iterative_refinement_steps;
  const double latol = o->linear_solver_atol, lrtol = o->linear_solver_rtol;
  // kinv_q = K^{-1} q, warm-started from the previous one (__init__.py:479-481)
  if ((rc = ir_refine(k, h->q, h->kq, steps, latol, lrtol, &out.lin[0]))) return rc;
  if (h->has_p)
    hipLaunchKernelGGL(k_pmul<2>, dim3(pmul_blocks(n)), gb, 0, st, k->Poff64, h->pdiag, n, h->xy,
                       h->px, h->kq, h->pq);

  for (int corrector = 0; corrector < 2; ++corrector) {
    if (corrector) {
      // sigma from the affine trial point (__init__.py:503-520)
      hipLaunchKernelGGL(k_step_alpha, gv, gb, 0, st, n, m, z, 0, h->xy, h->s, h->xp, h->cm, h->ds,
                         h->S, part, o->step_size_scale, fin);
      hipLaunchKernelGGL(k_sigma, gv, gb, 0, st, n, m, z, h->xy, h->s, h->xp, h->ds, h->S, part, fin);
    }
    // Newton step (__init__.py:708-757)
    hipLaunchKernelGGL(k_newton_rhs, dim3(gN), gb, 0, st, n, m, z, corrector, h->xy, h->s, h->ds,
                       h->xp, h->cm, h->r, h->S, nullptr);
    HIPC(hipMemcpyAsync(h->kr, h->xp, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    qtqp_ir_stats& ls = out.lin[1 + corrector];
    if ((rc = ir_refine(k, h->r, h->kr, steps, latol, lrtol, &ls))) return rc;
    const int quadratic = ls.status == 0 || ls.final_residual_norm < 1e-7;
    if (h->has_p)
      hipLaunchKernelGGL(k_pmul<1>, dim3(pmul_blocks(n)), gb, 0, st, k->Poff64, h->pdiag, n, h->kr,
                         h->pr, nullptr, nullptr);
    hipLaunchKernelGGL(k_tau, gv, gb, 0, st, n, m, h->has_p, h->q, h->kq, h->kr, h->xy, h->pq,
                       h->pr, h->px, part, corrector, quadratic, h->S, h->drec, nullptr, nullptr, fin);
    hipLaunchKernelGGL(k_newton_fin, dim3(gN), gb, 0, st, N, corrector, h->kr, h->kq, h->xp, h->S);
    HIPC(hipGetLastError());
  }
  // step and update (__init__.py:546-565)
  hipLaunchKernelGGL(k_step_alpha, gv, gb, 0, st, n, m, z, 1, h->xy, h->s, h->xp, h->cm, h->ds, h->S,
                     part, o->step_size_scale, fin);
  hipLaunchKernelGGL(k_update, dim3(gN), gb, 0, st, n, m, z, h->xy, h->s, h->xp, h->ds, h->S);
  HIPC(hipGetLastError());

  // termination (__init__.py:888-1008)
  if ((rc = tile_pass<double>(k, k->A64, m, n, h->xy, k->rpa, h->xy + n, nullptr))) return rc;
  hipLaunchKernelGGL(k_rowfin, dim3(blocks_for(m)), gb, 0, st, k->rpa, tiles_c(n), m, h->ax);
  hipLaunchKernelGGL(k_colfin, dim3(colfin_blocks(n)), gb, 0, st, k->part, tiles_rx(m), n, nullptr,
                     h->aty);
  if (h->has_p)
    hipLaunchKernelGGL(k_pmul<1>, dim3(pmul_blocks(n)), gb, 0, st, k->Poff64, h->pdiag, n, h->xy,
                       h->pxe, nullptr, nullptr);
  const int collect = o->collect_stats ? 1 : 0;
  const Tol tl{o->atol, o->rtol, o->atol_infeas, o->rtol_infeas};
  hipLaunchKernelGGL(k_term, gv, gb, 0, st, n, m, z, h->has_p, collect, h->xy, h->s, h->d, h->e,
                     h->ax, h->aty, h->pxe, h->b, h->c, h->S, part, tl, h->drec, fin);
  if (collect)
    hipLaunchKernelGGL(k_std, gv, gb, 0, st, n, m, z, h->xy, h->s, h->d, h->S, part, h->drec, fin);
  HIPC(hipGetLastError());
  qtqp_ipm_record dev{};
  HIPC(hipMemcpyAsync(&dev, h->drec, sizeof(dev), hipMemcpyDeviceToHost, st));
  HIPC(hipStreamSynchronize(st));
  for (int t = 0; t < 3; ++t) dev.lin[t] = out.lin[t];
  *rec = dev;
  return QTQP_OK;
}


// One refinement loop of the fixed sequence: krhs = [rhs_x; -rhs_y], the
// residual of the warm start in sol, then `steps` gated steps sol +=
// K_reg^{-1} res with the factor in its current precision (both precisions'
// kernels are in the sequence when `both`), each followed by its control.
int refine_seq(qtqp_ipm* h, int r, const double* rhs, double* sol, bool both, int steps,
               const qtqp_ipm_opts* o) {
  qtqp_kkt* k = h->k;
  const hipStream_t st = k->stream;
  IrCtl* ic = &h->ctl->ir[r];
  const Gate* g = &ic->gate;
  const int max_steps = o->max_iterative_refinement_steps;
  const double atol = o->linear_solver_atol, rtol = o->linear_solver_rtol;
  int rc;
  const int64_t n = h->n, m = h->m, N = n + m;
  // res = krhs - K_reg sol + corr * sol and the loop's control after it, with
  // the passes over A and P_off in one launch
  auto residual = [&](int step, const double* rhs_in, double* krhs_out) {
    const dim3 grid(tiles_rx(m) + tiles_r(n), tiles_c(n));
    if (n % 2 == 0)
      hipLaunchKernelGGL(k_tile_pass_kx<true>, grid, dim3(WG), 0, st, k->A64, k->Poff64, (int)m,
                         (int)n, sol, k->rpa, k->rpp, k->part, step > 0 ? g : nullptr);
    else
      hipLaunchKernelGGL(k_tile_pass_kx<false>, grid, dim3(WG), 0, st, k->A64, k->Poff64, (int)m,
                         (int)n, sol, k->rpa, k->rpp, k->part, step > 0 ? g : nullptr);
    hipLaunchKernelGGL(k_kmv_finish_ctl, dim3(blocks_for(N)), dim3(WG), 0, st, (int)n, (int)m, sol,
                       k->diag, k->rpp, k->rpa, tiles_c(n), k->part, tiles_rx(m), rhs_in, krhs_out,
                       k->corr, k->vres, h->ctl, r, step, max_steps, steps, atol, rtol);
  };
  // the warm start's residual against krhs = [rhs_x; -rhs_y] (written to vin)
  residual(0, rhs, k->vin);
  for (int step = 1; step <= std::min(steps, max_steps); ++step) {
    // sol += K_reg^{-1} res
    if (both && (rc = solve_with<float>(k, k->A32, k->G32, k->vres, nullptr, 0, g, 32, sol)))
      return rc;
    if ((rc = solve_with<double>(k, k->A64, k->G64, k->vres, nullptr, 0, g, 64, sol))) return rc;
    residual(step, k->vin, nullptr);
  }
  HIPC(hipGetLastError());
  return QTQP_OK;
}

// The launches of one whole iteration with no host synchronisation: the
// sequence a graph captures.  `both`: the FP32 factor chain and solves are
// in it (a mixed handle that has not fallen back to FP64).
int launch_iteration(qtqp_ipm* h, const qtqp_ipm_opts* o, bool both, int steps) {
  qtqp_kkt* k = h->k;
  const hipStream_t st = k->stream;
  const int n = (int)h->n, m = (int)h->m, z = (int)h->z;
  const int64_t N = h->n + h->m;
  const int nv = h->nv;
  const dim3 gv(nv), gb(WG), g1(1);
  const unsigned gN = blocks_for(N);
  double* part = h->vpart;
  Ctl* c = h->ctl;
  unsigned* fin = &c->fin_ctr;
  int rc;
  hipLaunchKernelGGL(k_iter_begin, dim3(gN), gb, 0, st, c, both ? 1 : 0, N, (int64_t)m, h->xy, h->s,
                     h->kq, h->S, h->xy_sv, h->s_sv, h->kq_sv, h->S_sv);
  // normalise, mu, regularised diagonal (__init__.py:474-478, direct.py:183-216)
  hipLaunchKernelGGL(k_norm, gv, gb, 0, st, n, m, z, h->xy, h->s, part, h->S, fin);
  hipLaunchKernelGGL(k_scale_diag, dim3(gN), gb, 0, st, n, m, z, h->xy, h->s, h->pdiag,
                     o->min_static_regularization, h->S, k->diag, k->corr);
  HIPC(hipGetLastError());
  // factor: the FP32 chain, then the FP64 chain if FP64 or after an FP32 breakdown
  if (both && (rc = factor_launch<float>(k, k->A32, k->Poff32, k->G32, &c->info32, kJitterRel[0],
                                         &c->fgate, 32)))
    return rc;
  hipLaunchKernelGGL(k_fact_mid, g1, g1, 0, st, c);
  if ((rc = factor_launch<double>(k, k->A64, k->Poff64, k->G64, &c->info64, kJitterRel[0],
                                  &c->fgate, 64)))
    return rc;
  hipLaunchKernelGGL(k_fact_end, g1, g1, 0, st, c);
  // kinv_q = K^{-1} q, warm-started from the previous one (__init__.py:479-481)
  if ((rc = refine_seq(h, 0, h->q, h->kq, both, steps, o))) return rc;
  if (h->has_p)
    hipLaunchKernelGGL(k_pmul<2>, dim3(pmul_blocks(n)), gb, 0, st, k->Poff64, h->pdiag, n, h->xy,
                       h->px, h->kq, h->pq);
  for (int corrector = 0; corrector < 2; ++corrector) {
    if (corrector) {
      hipLaunchKernelGGL(k_step_alpha, gv, gb, 0, st, n, m, z, 0, h->xy, h->s, h->xp, h->cm, h->ds,
                         h->S, part, o->step_size_scale, fin);
      hipLaunchKernelGGL(k_sigma, gv, gb, 0, st, n, m, z, h->xy, h->s, h->xp, h->ds, h->S, part, fin);
    }
    hipLaunchKernelGGL(k_newton_rhs, dim3(gN), gb, 0, st, n, m, z, corrector, h->xy, h->s, h->ds,
                       h->xp, h->cm, h->r, h->S, h->kr);
    if ((rc = refine_seq(h, 1 + corrector, h->r, h->kr, both, steps, o))) return rc;
    if (h->has_p)
      hipLaunchKernelGGL(k_pmul<1>, dim3(pmul_blocks(n)), gb, 0, st, k->Poff64, h->pdiag, n, h->kr,
                         h->pr, nullptr, nullptr);
    const IrCtl* ic = &c->ir[1 + corrector];
    hipLaunchKernelGGL(k_tau, gv, gb, 0, st, n, m, h->has_p, h->q, h->kq, h->kr, h->xy, h->pq,
                       h->pr, h->px, part, corrector, 0, h->S, h->drec, &ic->status, &ic->res, fin);
    hipLaunchKernelGGL(k_newton_fin, dim3(gN), gb, 0, st, N, corrector, h->kr, h->kq, h->xp, h->S);
    HIPC(hipGetLastError());
  }
  hipLaunchKernelGGL(k_step_alpha, gv, gb, 0, st, n, m, z, 1, h->xy, h->s, h->xp, h->cm, h->ds, h->S,
                     part, o->step_size_scale, fin);
  hipLaunchKernelGGL(k_update, dim3(gN), gb, 0, st, n, m, z, h->xy, h->s, h->xp, h->ds, h->S);
  HIPC(hipGetLastError());
  // termination (__init__.py:888-1008)
  if ((rc = tile_pass<double>(k, k->A64, m, n, h->xy, k->rpa, h->xy + n, nullptr))) return rc;
  hipLaunchKernelGGL(k_rowfin, dim3(blocks_for(m)), gb, 0, st, k->rpa, tiles_c(n), m, h->ax);
  hipLaunchKernelGGL(k_colfin, dim3(colfin_blocks(n)), gb, 0, st, k->part, tiles_rx(m), n, nullptr,
                     h->aty, nullptr, 0);
  if (h->has_p)
    hipLaunchKernelGGL(k_pmul<1>, dim3(pmul_blocks(n)), gb, 0, st, k->Poff64, h->pdiag, n, h->xy,
                       h->pxe, nullptr, nullptr);
  const int collect = o->collect_stats ? 1 : 0;
  const Tol tl{o->atol, o->rtol, o->atol_infeas, o->rtol_infeas};
  hipLaunchKernelGGL(k_term, gv, gb, 0, st, n, m, z, h->has_p, collect, h->xy, h->s, h->d, h->e,
                     h->ax, h->aty, h->pxe, h->b, h->c, h->S, part, tl, h->drec, fin);
  if (collect)
    hipLaunchKernelGGL(k_std, gv, gb, 0, st, n, m, z, h->xy, h->s, h->d, h->S, part, h->drec, fin);
  hipLaunchKernelGGL(k_iter_end, g1, g1, 0, st, c, h->drec);
  HIPC(hipGetLastError());
  return QTQP_OK;
}

// Options that change the captured sequence.
bool same_graph_opts(const qtqp_ipm_opts& a, const qtqp_ipm_opts& b) {
  return a.atol == b.atol && a.rtol == b.rtol && a.atol_infeas == b.atol_infeas &&
         a.rtol_infeas == b.rtol_infeas && a.step_size_scale == b.step_size_scale &&
         a.min_static_regularization == b.min_static_regularization &&
         a.linear_solver_atol == b.linear_solver_atol &&
         a.linear_solver_rtol == b.linear_solver_rtol &&
         a.max_iterative_refinement_steps == b.max_iterative_refinement_steps &&
         a.collect_stats == b.collect_stats && a.graph_steps == b.graph_steps;
}

// The graph of one iteration for these options (captured on first use).
int iteration_graph(qtqp_ipm* h, const qtqp_ipm_opts* o, bool both, hipGraphExec_t* out) {
  auto& slot = h->graphs[both ? 1 : 0];
  // An FP32 factor needs two steps where an FP64 one needs one.  A step held
  // in the graph but not needed costs a few empty launches; a refinement that
  // needs more steps than the graph holds costs the whole iteration again, so
  // an expensive iteration (factor work above 2e9 flops) starts with two.
  const double nn = (double)h->n, work = (double)h->m * nn * nn + nn * nn * nn / 3.0;
  const int base_steps = (both || work > 2e9) ? 2 : 1;
  const int steps = o->graph_steps > 0 ? o->graph_steps : std::max(base_steps, h->steps_cur);
  if (slot.exec && same_graph_opts(slot.opts, *o) && slot.steps == steps) {
    *out = slot.exec;
    return QTQP_OK;
  }
  if (slot.exec) {
    HIPC(hipGraphExecDestroy(slot.exec));
    slot.exec = nullptr;
  }
  const hipStream_t st = h->k->stream;
  HIPC(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const int rc = launch_iteration(h, o, both, steps);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &graph);
  if (rc != QTQP_OK || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    if (rc != QTQP_OK) return rc;
    return fail(QTQP_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess)
    return fail(QTQP_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
  slot.exec = exec;
  slot.opts = *o;
  slot.steps = steps;
  *out = exec;
  return QTQP_OK;
}

// The state of the control block the next iteration's precision rule reads,
// after an iteration the host ran itself.
struct CtlCarry {
  int force64, fprec, prev_steps, prev_fprec;
};

int iterate_graph(qtqp_ipm* h, const qtqp_ipm_opts* o, qtqp_ipm_record* rec) {
  qtqp_kkt* k = h->k;
  const hipStream_t st = k->stream;
  const bool both = k->precision == QTQP_PREC_MIXED && !k->force64;
  hipGraphExec_t exec = nullptr;
  int rc = iteration_graph(h, o, both, &exec);
  if (rc) return rc;
  HIPC(hipGraphLaunch(exec, st));
  qtqp_ipm_record dev{};
  Ctl c{};
  HIPC(hipMemcpyAsync(&dev, h->drec, sizeof(dev), hipMemcpyDeviceToHost, st));
  HIPC(hipMemcpyAsync(&c, h->ctl, sizeof(c), hipMemcpyDeviceToHost, st));
  HIPC(hipStreamSynchronize(st));
  k->force64 = c.force64 != 0;
  if (c.redo) {
    // Back to the iteration's start, then the host-driven launches, which
    // loop as long as the reference does (and retry a broken-down factor).
    // A refinement that needed more steps than the graph holds makes the next
    // graphs hold one more (up to 8).
    h->redos++;
    if (c.redo == 1 && o->graph_steps == 0) {
      const int used = h->graphs[both ? 1 : 0].steps;
      h->steps_cur = std::min(std::min(used + 1, 8), o->max_iterative_refinement_steps);
    }
    const int64_t N = h->n + h->m;
    hipLaunchKernelGGL(k_iter_restore, dim3(blocks_for(N)), dim3(WG), 0, st, N, h->m, h->xy, h->s,
                       h->kq, h->S, h->xy_sv, h->s_sv, h->kq_sv, h->S_sv);
    HIPC(hipGetLastError());
    k->f32_solves = k->f32_steps = 0;
    k->fprec = c.fprec;
    rc = iterate_eager(h, o, rec);
    if (rc != QTQP_OK) return rc;
    int steps = 0;
    for (int t = 0; t < 3; ++t) steps += rec->lin[t].solves;
    const CtlCarry carry{k->force64 ? 1 : 0, k->fprec, steps, k->fprec};
    HIPC(hipMemcpyAsync(reinterpret_cast<char*>(h->ctl) + offsetof(Ctl, force64), &carry,
                        sizeof(carry), hipMemcpyHostToDevice, st));
    HIPC(hipStreamSynchronize(st));
    return QTQP_OK;
  }
  for (int t = 0; t < 3; ++t)
    if (c.ir[t].nan) return fail(QTQP_ERR_NAN, "Linear solver returned NaNs.");
  k->fprec = c.fprec;
  k->have_diag = true;
  k->have_factor = true;
  *rec = dev;
  return QTQP_OK;
}

}  // namespace

extern "C" {

int qtqp_ipm_iterate(qtqp_ipm* h, const qtqp_ipm_opts* o, qtqp_ipm_record* rec) {
  if (!h || !o || !rec) return fail(QTQP_ERR_ARG, "NULL argument");
  if (o->max_iterative_refinement_steps < 1)
    return fail(QTQP_ERR_ARG, "max_iterative_refinement_steps must be >= 1");
  if (o->graph_steps < 0) return fail(QTQP_ERR_ARG, "graph_steps must be >= 0");
  return o->use_graph ? iterate_graph(h, o, rec) : iterate_eager(h, o, rec);
}

int qtqp_ipm_get_iterates(qtqp_ipm* h, double* x, double* y, double* s, double* tau) {
  if (!h || !x || !y || !s || !tau) return fail(QTQP_ERR_ARG, "NULL argument");
  const int64_t n = h->n, m = h->m;
  const hipStream_t st = h->k->stream;
  std::vector<double> d(m), e(n);
  Scal sc{};
  HIPC(hipMemcpyAsync(x, h->xy, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  HIPC(hipMemcpyAsync(y, h->xy + n, sizeof(double) * m, hipMemcpyDeviceToHost, st));
  HIPC(hipMemcpyAsync(s, h->s, sizeof(double) * m, hipMemcpyDeviceToHost, st));
  HIPC(hipMemcpyAsync(d.data(), h->d, sizeof(double) * m, hipMemcpyDeviceToHost, st));
  HIPC(hipMemcpyAsync(e.data(), h->e, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  HIPC(hipMemcpyAsync(&sc, h->S, sizeof(sc), hipMemcpyDeviceToHost, st));
  HIPC(hipStreamSynchronize(st));
  for (int64_t j = 0; j < n; ++j) x[j] = e[j] * x[j];
  for (int64_t i = 0; i < m; ++i) {
    y[i] = d[i] * y[i];
    s[i] = s[i] / d[i];
  }
  *tau = sc.tau;
  return QTQP_OK;
}

int qtqp_ipm_get_info(qtqp_ipm* h, qtqp_kkt_info* info) {
  if (!h || !info) return fail(QTQP_ERR_ARG, "NULL argument");
  Ctl c{};
  HIPC(hipMemcpy(&c, h->ctl, sizeof(c), hipMemcpyDeviceToHost));
  *info = h->k->info_acc;
  info->factor_precision = h->k->fprec;
  // the graph path's counters (the host-driven path counts in info_acc)
  info->factorizations += c.factorizations;
  info->fp64_fallbacks += c.fp64_fallbacks;
  info->backend_solves += c.backend_solves;
  info->inner_steps += c.inner_steps;
  info->factor_ms += (double)c.factor_ticks / h->wall_khz;
  info->graph_redos = h->redos;
  return QTQP_OK;
}

void qtqp_ipm_free(qtqp_ipm* h) {
  if (!h) return;
  ipm_release(h);
  delete h;
}

}  // extern "C"

// k_potential.hip — the nonlocal Poisson solve for phi (replaces the 11th Eigen GMRES solve
// of main.cpp:764-779 on K11, utils.h:1794-1848, with residual F11 of utils.h:798-841).
//
// Unknowns are the free nodes (liquid, >= 1 liquid-liquid bond, not Dirichlet for phi). Rows
// are scaled by V_i/V0 so the operator M is symmetric positive definite:
//   (M p)_i = dd_i p_i + sum_{free j} w_ij (p_i - p_j),  w_ij = sc_i Beta V_j / xi^(3+2 s111),
// dd_i = sc_i * (weights of the Dirichlet liquid neighbours); right-hand side
//   b_i = sc_i ((1/(K11prop1 sigma1)) F CSolid sum_k z_k U_k,i + sum_{Dirichlet j} w_ij phi_j).
// The bond sum is taken over differences p_i - p_j, which keeps the FP32 inner solve accurate
// for smooth vectors.
//
// FP64: Jacobi-preconditioned CG, three kernels per iteration (SpMV + dot, update + dots,
// direction update); all scalars stay on the device and a converged flag makes the remaining
// launches of a batch exit at once, so the host synchronises once per batch.
// MIXED: iterative refinement; the residual b - M x is formed in FP64 (FP64 weights and
// vectors), the correction solve runs in FP32 (FP32 weights and vectors, Kahan-compensated
// bond sums, FP64 dot products), the correction is added in FP64.
#include <cmath>

#include "nnpp_internal.hpp"

namespace {

constexpr int CG_BATCH = 32;

struct Slices {
  const int64_t* off;
  const int32_t* len;
  const int32_t* col;
};

__device__ __forceinline__ bool cg_done(const double* scal) {
  return nnpp_ld_coherent(scal + SC_DONE) != 0.0;
}

__device__ __forceinline__ double spmv64_row(const Slices& sl, const double* __restrict__ w,
                                             const double* __restrict__ dd,
                                             const double* __restrict__ v, int64_t i) {
  const int64_t base = sl.off[i / NNPP_WAVE] + (i % NNPP_WAVE);
  const int L = sl.len[i / NNPP_WAVE];
  const double vi = v[i];
  double acc = 0.0;
  for (int t = 0; t < L; t++) {
    const int64_t slot = base + (int64_t)t * NNPP_WAVE;
    const int32_t j = sl.col[slot];
    if (j < 0) break;
    acc += w[slot] * (vi - v[j]);
  }
  return dd[i] * vi + acc;
}

__device__ __forceinline__ float spmv32_row(const Slices& sl, const float* __restrict__ w,
                                            const float* __restrict__ dd,
                                            const float* __restrict__ v, int64_t i) {
  const int64_t base = sl.off[i / NNPP_WAVE] + (i % NNPP_WAVE);
  const int L = sl.len[i / NNPP_WAVE];
  const float vi = v[i];
  float acc = 0.0f, cmp = 0.0f;
  for (int t = 0; t < L; t++) {
    const int64_t slot = base + (int64_t)t * NNPP_WAVE;
    const int32_t j = sl.col[slot];
    if (j < 0) break;
    const float y = w[slot] * (vi - v[j]) - cmp;
    const float s = acc + y;
    cmp = (s - acc) - y;
    acc = s;
  }
  return dd[i] * vi + acc;
}

// b (free rows), warm start x = phi, ||b||^2.
__global__ __launch_bounds__(NNPP_BLOCK) void k_rhs(int32_t n, int S, const double* __restrict__ U,
                                                    const double* __restrict__ zs, double cq,
                                                    const uint8_t* __restrict__ pfree,
                                                    const double* __restrict__ psc,
                                                    const double* __restrict__ pbdir,
                                                    const double* __restrict__ phi,
                                                    double* __restrict__ b, double* __restrict__ x,
                                                    double* part, unsigned* ticket, double* scal) {
  const int64_t i = (int64_t)blockIdx.x * NNPP_BLOCK + threadIdx.x;
  double v[1] = {0.0};
  if (i < n) {
    double bi = 0.0, xi = 0.0;
    if (pfree[i]) {
      double q = 0.0;
      for (int k = 0; k < S; k++) q += zs[k] * U[(int64_t)k * n + i];
      bi = psc[i] * (cq * q + pbdir[i]);
      xi = phi[i];
    }
    b[i] = bi;
    x[i] = xi;
    v[0] = bi * bi;
  }
  if (nnpp_grid_sum<1>(v, part, ticket, gridDim.x, blockIdx.x)) scal[SC_BNORM2] = v[0];
}

// r = b - M x in FP64. START: also z = D^-1 r, p = z, rho, done flag (target from ||b||).
template <bool START>
__global__ __launch_bounds__(NNPP_BLOCK) void k_resid(int32_t n, Slices sl,
                                                      const double* __restrict__ w,
                                                      const double* __restrict__ dd,
                                                      const double* __restrict__ pinv,
                                                      const double* __restrict__ b,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ r, double* __restrict__ zv,
                                                      double* __restrict__ pv, double rtol2,
                                                      double* part, unsigned* ticket,
                                                      double* scal) {
  const int64_t i = nnpp_xcd_block(blockIdx.x, gridDim.x) * NNPP_BLOCK + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < n) {
    const double ri = (pinv[i] != 0.0) ? b[i] - spmv64_row(sl, w, dd, x, i) : 0.0;
    r[i] = ri;
    v[1] = ri * ri;
    if (START) {
      const double zi = ri * pinv[i];
      zv[i] = zi;
      pv[i] = zi;
      v[0] = ri * zi;
    }
  }
  if (nnpp_grid_sum<2>(v, part, ticket, gridDim.x, blockIdx.x)) {
    scal[SC_RR] = v[1];
    if (START) {
      scal[SC_RHO] = v[0];
      const double t2 = rtol2 * scal[SC_BNORM2];
      scal[SC_TARGET2] = t2;
      scal[SC_ITERS] = 0.0;
      scal[SC_DONE] = (v[1] <= t2) ? 1.0 : 0.0;
    }
  }
}

__global__ __launch_bounds__(NNPP_BLOCK) void k_cg_spmv(int32_t n, Slices sl,
                                                        const double* __restrict__ w,
                                                        const double* __restrict__ dd,
                                                        const double* __restrict__ p,
                                                        double* __restrict__ q, double* part,
                                                        unsigned* ticket, double* scal) {
  if (cg_done(scal)) return;
  const int64_t i = nnpp_xcd_block(blockIdx.x, gridDim.x) * NNPP_BLOCK + threadIdx.x;
  double v[1] = {0.0};
  if (i < n) {
    const double qi = spmv64_row(sl, w, dd, p, i);
    q[i] = qi;
    v[0] = p[i] * qi;
  }
  if (nnpp_grid_sum<1>(v, part, ticket, gridDim.x, blockIdx.x)) {
    scal[SC_PQ] = v[0];
    scal[SC_ALPHA] = scal[SC_RHO] / v[0];
  }
}

__global__ __launch_bounds__(NNPP_BLOCK) void k_cg_update(int32_t n, const double* __restrict__ pinv,
                                                          const double* __restrict__ p,
                                                          const double* __restrict__ q,
                                                          double* __restrict__ x,
                                                          double* __restrict__ r,
                                                          double* __restrict__ zv, double* part,
                                                          unsigned* ticket, double* scal) {
  if (cg_done(scal)) return;
  const double alpha = nnpp_ld_coherent(scal + SC_ALPHA);
  const int64_t i = (int64_t)blockIdx.x * NNPP_BLOCK + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < n) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    const double zi = ri * pinv[i];
    zv[i] = zi;
    v[0] = ri * zi;
    v[1] = ri * ri;
  }
  if (nnpp_grid_sum<2>(v, part, ticket, gridDim.x, blockIdx.x)) {
    scal[SC_BETA] = v[0] / scal[SC_RHO];
    scal[SC_RHO] = v[0];
    scal[SC_RR] = v[1];
    scal[SC_ITERS] += 1.0;
    if (v[1] <= scal[SC_TARGET2]) scal[SC_DONE] = 1.0;
  }
}

__global__ void k_cg_dir(int32_t n, const double* __restrict__ zv, double* __restrict__ p,
                         const double* scal) {
  if (cg_done(scal)) return;
  const double beta = nnpp_ld_coherent(scal + SC_BETA);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = zv[i] + beta * p[i];
}

// ---- mixed precision inner solve ----

__global__ __launch_bounds__(NNPP_BLOCK) void k_m_start(int32_t n, const double* __restrict__ r64,
                                                        double inv_s, const float* __restrict__ pinv,
                                                        float* __restrict__ e, float* __restrict__ r,
                                                        float* __restrict__ zv, float* __restrict__ p,
                                                        double inner_rtol2, double* part,
                                                        unsigned* ticket, double* scal) {
  const int64_t i = (int64_t)blockIdx.x * NNPP_BLOCK + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < n) {
    const float ri = (float)(r64[i] * inv_s);
    const float zi = ri * pinv[i];
    e[i] = 0.0f;
    r[i] = ri;
    zv[i] = zi;
    p[i] = zi;
    v[0] = (double)ri * (double)zi;
    v[1] = (double)ri * (double)ri;
  }
  if (nnpp_grid_sum<2>(v, part, ticket, gridDim.x, blockIdx.x)) {
    scal[SC_RHO] = v[0];
    scal[SC_RR] = v[1];
    scal[SC_TARGET2] = inner_rtol2 * v[1];
    scal[SC_ITERS] = 0.0;
    scal[SC_DONE] = (v[1] == 0.0) ? 1.0 : 0.0;
  }
}

__global__ __launch_bounds__(NNPP_BLOCK) void k_m_spmv(int32_t n, Slices sl,
                                                       const float* __restrict__ w,
                                                       const float* __restrict__ dd,
                                                       const float* __restrict__ p,
                                                       float* __restrict__ q, double* part,
                                                       unsigned* ticket, double* scal) {
  if (cg_done(scal)) return;
  const int64_t i = nnpp_xcd_block(blockIdx.x, gridDim.x) * NNPP_BLOCK + threadIdx.x;
  double v[1] = {0.0};
  if (i < n) {
    const float qi = spmv32_row(sl, w, dd, p, i);
    q[i] = qi;
    v[0] = (double)p[i] * (double)qi;
  }
  if (nnpp_grid_sum<1>(v, part, ticket, gridDim.x, blockIdx.x)) {
    scal[SC_PQ] = v[0];
    scal[SC_ALPHA] = scal[SC_RHO] / v[0];
  }
}

__global__ __launch_bounds__(NNPP_BLOCK) void k_m_update(int32_t n, const float* __restrict__ pinv,
                                                         const float* __restrict__ p,
                                                         const float* __restrict__ q,
                                                         float* __restrict__ e,
                                                         float* __restrict__ r,
                                                         float* __restrict__ zv, double* part,
                                                         unsigned* ticket, double* scal) {
  if (cg_done(scal)) return;
  const float alpha = (float)nnpp_ld_coherent(scal + SC_ALPHA);
  const int64_t i = (int64_t)blockIdx.x * NNPP_BLOCK + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < n) {
    e[i] += alpha * p[i];
    const float ri = r[i] - alpha * q[i];
    r[i] = ri;
    const float zi = ri * pinv[i];
    zv[i] = zi;
    v[0] = (double)ri * (double)zi;
    v[1] = (double)ri * (double)ri;
  }
  if (nnpp_grid_sum<2>(v, part, ticket, gridDim.x, blockIdx.x)) {
    scal[SC_BETA] = v[0] / scal[SC_RHO];
    scal[SC_RHO] = v[0];
    scal[SC_RR] = v[1];
    scal[SC_ITERS] += 1.0;
    if (v[1] <= scal[SC_TARGET2]) scal[SC_DONE] = 1.0;
  }
}

__global__ void k_m_dir(int32_t n, const float* __restrict__ zv, float* __restrict__ p,
                        const double* scal) {
  if (cg_done(scal)) return;
  const float beta = (float)nnpp_ld_coherent(scal + SC_BETA);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = zv[i] + beta * p[i];
}

__global__ void k_m_apply(int32_t n, double s, const float* __restrict__ e,
                          double* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += s * (double)e[i];
}

// phi = x on free nodes, Dirichlet value on Dirichlet nodes, unchanged elsewhere.
__global__ void k_scatter_phi(int32_t n, int S, const uint8_t* __restrict__ pfree,
                              const uint16_t* __restrict__ bcmask, const double* __restrict__ bcphi,
                              const double* __restrict__ x, const double* __restrict__ phi_old,
                              double* __restrict__ phi) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = phi_old[i];
  if (pfree[i]) v = x[i];
  else if ((bcmask[i] >> S) & 1u) v = bcphi[i];
  phi[i] = v;
}

}  // namespace

static int read_scalars(nnpp_ctx* c) {
  NNPP_HIP(c, hipMemcpyAsync(c->h_scal, c->scal, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost,
                             c->stream));
  NNPP_HIP(c, hipStreamSynchronize(c->stream));
  return NNPP_OK;
}

// Uq: state whose species give the charge; phi_in: warm start / kept values; phi_out: result.
int nnpp_potential_solve_dev(nnpp_ctx* c, int precision, const double* Uq, double* phi,
                             int32_t* iters, double* relres, int64_t* spmv64, int64_t* spmv32) {
  const int32_t n = c->n;
  const int S = c->p.n_species;
  hipStream_t s = c->stream;
  if (iters) *iters = 0;
  if (relres) *relres = 0.0;
  if (!n) return NNPP_OK;
  const int64_t nb = nnpp_nblocks(n);
  Slices sl{c->pot.off, c->pot.len, c->pot.col};
  double zs[NNPP_MAX_SPECIES] = {0};
  for (int k = 0; k < S; k++) zs[k] = c->p.z[k];
  double* dz = c->cg_q;  // z_k staging: copy valences to device scratch
  NNPP_HIP(c, hipMemcpyAsync(c->scal + SC_COUNT, zs, sizeof(double) * NNPP_MAX_SPECIES,
                             hipMemcpyHostToDevice, s));
  (void)dz;
  const double cq = (1. / (c->p.K_phi * c->p.sigma1)) * c->p.F * c->p.CSolid;
  k_rhs<<<nb, NNPP_BLOCK, 0, s>>>(n, S, Uq, c->scal + SC_COUNT, cq, c->pfree, c->psc, c->pbdir,
                                  phi, c->cg_b, c->cg_x, c->red_part, c->red_ticket, c->scal);
  NNPP_HIP(c, hipGetLastError());
  int st = read_scalars(c);
  if (st) return st;
  const double bnorm2 = c->h_scal[SC_BNORM2];
  if (!std::isfinite(bnorm2))
    return nnpp_set_error(c, NNPP_ERR_NONFINITE, "non-finite potential right-hand side");
  const double* bcphi = c->bcval + (int64_t)S * n;
  int64_t n64 = 0, n32 = 0;
  int total = 0;
  int status = NNPP_OK;
  double rr = 0.0;
  if (bnorm2 == 0.0) {
    // zero right-hand side: x = 0 on the free nodes (already set to 0 by k_rhs? no: x = phi)
    NNPP_HIP(c, hipMemsetAsync(c->cg_x, 0, sizeof(double) * n, s));
  } else if (precision != NNPP_MIXED) {
    const double rtol2 = c->rtol * c->rtol;
    k_resid<true><<<nb, NNPP_BLOCK, 0, s>>>(n, sl, c->pot.w64, c->pdd, c->pinv, c->cg_b, c->cg_x,
                                            c->cg_r, c->cg_z, c->cg_p, rtol2, c->red_part,
                                            c->red_ticket, c->scal);
    NNPP_HIP(c, hipGetLastError());
    n64++;
    st = read_scalars(c);
    if (st) return st;
    bool done = c->h_scal[SC_DONE] != 0.0;
    while (!done) {
      const int left = c->maxit - total;
      if (left <= 0) {
        status = NNPP_ERR_NOCONV;
        break;
      }
      const int batch = left < CG_BATCH ? left : CG_BATCH;
      for (int it = 0; it < batch; it++) {
        k_cg_spmv<<<nb, NNPP_BLOCK, 0, s>>>(n, sl, c->pot.w64, c->pdd, c->cg_p, c->cg_q,
                                            c->red_part, c->red_ticket, c->scal);
        k_cg_update<<<nb, NNPP_BLOCK, 0, s>>>(n, c->pinv, c->cg_p, c->cg_q, c->cg_x, c->cg_r,
                                              c->cg_z, c->red_part, c->red_ticket, c->scal);
        k_cg_dir<<<nb, NNPP_BLOCK, 0, s>>>(n, c->cg_z, c->cg_p, c->scal);
      }
      NNPP_HIP(c, hipGetLastError());
      st = read_scalars(c);
      if (st) return st;
      total = (int)c->h_scal[SC_ITERS];
      done = c->h_scal[SC_DONE] != 0.0;
      if (!std::isfinite(c->h_scal[SC_RR]))
        return nnpp_set_error(c, NNPP_ERR_NONFINITE, "potential solve diverged (non-finite)");
    }
    n64 += total;
    // FP64 residual check of the final iterate
    k_resid<false><<<nb, NNPP_BLOCK, 0, s>>>(n, sl, c->pot.w64, c->pdd, c->pinv, c->cg_b, c->cg_x,
                                             c->cg_r, nullptr, nullptr, 0.0, c->red_part,
                                             c->red_ticket, c->scal);
    NNPP_HIP(c, hipGetLastError());
    n64++;
    st = read_scalars(c);
    if (st) return st;
    rr = c->h_scal[SC_RR];
  } else {
    const double target2 = c->rtol * c->rtol * bnorm2;
    const double irt2 = c->inner_rtol * c->inner_rtol;
    while (true) {
      k_resid<false><<<nb, NNPP_BLOCK, 0, s>>>(n, sl, c->pot.w64, c->pdd, c->pinv, c->cg_b,
                                               c->cg_x, c->cg_r, nullptr, nullptr, 0.0,
                                               c->red_part, c->red_ticket, c->scal);
      NNPP_HIP(c, hipGetLastError());
      n64++;
      st = read_scalars(c);
      if (st) return st;
      rr = c->h_scal[SC_RR];
      if (!std::isfinite(rr))
        return nnpp_set_error(c, NNPP_ERR_NONFINITE, "potential solve diverged (non-finite)");
      if (rr <= target2) break;
      if (total >= c->maxit) {
        status = NNPP_ERR_NOCONV;
        break;
      }
      const double sc = std::sqrt(rr);
      k_m_start<<<nb, NNPP_BLOCK, 0, s>>>(n, c->cg_r, 1.0 / sc, c->pinv32, c->f_e, c->f_r, c->f_z,
                                          c->f_p, irt2, c->red_part, c->red_ticket, c->scal);
      NNPP_HIP(c, hipGetLastError());
      int inner = 0;
      bool done = false;
      while (!done) {
        const int left = c->maxit - total - inner;
        if (left <= 0) break;
        const int batch = left < CG_BATCH ? left : CG_BATCH;
        for (int it = 0; it < batch; it++) {
          k_m_spmv<<<nb, NNPP_BLOCK, 0, s>>>(n, sl, c->pot.w32, c->pdd32, c->f_p, c->f_q,
                                             c->red_part, c->red_ticket, c->scal);
          k_m_update<<<nb, NNPP_BLOCK, 0, s>>>(n, c->pinv32, c->f_p, c->f_q, c->f_e, c->f_r,
                                               c->f_z, c->red_part, c->red_ticket, c->scal);
          k_m_dir<<<nb, NNPP_BLOCK, 0, s>>>(n, c->f_z, c->f_p, c->scal);
        }
        NNPP_HIP(c, hipGetLastError());
        st = read_scalars(c);
        if (st) return st;
        inner = (int)c->h_scal[SC_ITERS];
        done = c->h_scal[SC_DONE] != 0.0;
      }
      total += inner;
      n32 += inner;
      k_m_apply<<<nb, NNPP_BLOCK, 0, s>>>(n, sc, c->f_e, c->cg_x);
      NNPP_HIP(c, hipGetLastError());
      if (inner == 0) {  // no progress possible (budget spent)
        status = NNPP_ERR_NOCONV;
        break;
      }
    }
  }
  k_scatter_phi<<<nb, NNPP_BLOCK, 0, s>>>(n, S, c->pfree, c->bcmask, bcphi, c->cg_x, phi, phi);
  NNPP_HIP(c, hipGetLastError());
  NNPP_HIP(c, hipStreamSynchronize(s));
  if (iters) *iters = total;
  if (relres) *relres = bnorm2 > 0 ? std::sqrt(rr / bnorm2) : 0.0;
  if (spmv64) *spmv64 += n64;
  if (spmv32) *spmv32 += n32;
  return status;
}

// One potential-operator application (with its dot product) on the solver's direction vector;
// used by the kernel timing hook. The converged flag is cleared first so the launch runs.
int nnpp_pot_spmv_launch(nnpp_ctx* c, int precision) {
  const int32_t n = c->n;
  if (!n) return NNPP_OK;
  const int64_t nb = nnpp_nblocks(n);
  Slices sl{c->pot.off, c->pot.len, c->pot.col};
  if (precision == NNPP_MIXED)
    k_m_spmv<<<nb, NNPP_BLOCK, 0, c->stream>>>(n, sl, c->pot.w32, c->pdd32, c->f_p, c->f_q,
                                               c->red_part, c->red_ticket, c->scal);
  else
    k_cg_spmv<<<nb, NNPP_BLOCK, 0, c->stream>>>(n, sl, c->pot.w64, c->pdd, c->cg_p, c->cg_q,
                                                c->red_part, c->red_ticket, c->scal);
  NNPP_HIP(c, hipGetLastError());
  return NNPP_OK;
}

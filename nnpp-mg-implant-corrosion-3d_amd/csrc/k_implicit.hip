// k_implicit.hip — the reference's implicit time step (main.cpp:547-1025) on the device.
//
// Modified Newton with the Jacobian of utils::NewtonMatrices (utils.h:844-1849) frozen at U^n.
// The reference assembles eleven Eigen sparse matrices and runs eleven serial GMRES solves; here
// the Jacobian is never stored: its species blocks are applied by the flux kernel's JAC mode
// (one pass over the bonds applies all S blocks at once, the vectors staged in the AoS gather
// records), so the S species systems are solved together by one batched BiCGSTAB (per-species
// scalars, a species leaves the batch when it converges). The potential block K11 is the
// potential operator: the Newton update of phi is one potential CG solve with the Dirichlet
// coupling moved to the right-hand side. Dot products use the deterministic two-launch
// reduction (block partials, then one block sums them in block order).
#include <cmath>
#include <cstring>

#include "nnpp_internal.hpp"

namespace {

constexpr int NV = 2 * NNPP_MAX_SPECIES;  // values per reduction: two per species

struct Coef {
  double v[NNPP_MAX_SPECIES];
};

// stride of the Jacobian's gather records: [v_0 .. v_{S-1}, phi^n], padded to even
static inline int jac_rec_len(int S) { return (S + 2) & ~1; }

// The species kernels run on a (gx, S) grid: blockIdx.y is the species, so a thread keeps two
// accumulators instead of 2 S (the all-species form held one wave per SIMD). gx is a multiple
// of the XCD count, so the XCD of a block is blockIdx.x % 8 as in nnpp_blocks_of.
static inline int64_t bs_grid(int64_t nb, int S) {
  int64_t g = (NNPP_RED_GRID / S) & ~(int64_t)(NNPP_NXCD - 1);
  if (g < NNPP_NXCD) g = NNPP_NXCD;
  return nb < g ? nb : g;
}

// Block partials of species blockIdx.y: a in column k, b in column NNPP_MAX_SPECIES + k of row
// blockIdx.x (k_bs_final sums each column over the rows in row order).
__device__ __forceinline__ void bs_partial(double a, double b, double* __restrict__ part) {
  __shared__ double lds[2 * (NNPP_BLOCK / NNPP_WAVE)];
  double v[2] = {a, b};
  nnpp_block_sum<2>(v, lds);
  if (threadIdx.x == 0) {
    part[(int64_t)blockIdx.x * NV + blockIdx.y] = v[0];
    part[(int64_t)blockIdx.x * NV + NNPP_MAX_SPECIES + blockIdx.y] = v[1];
  }
}

// Gather record of node i for the column form: species k's value, and (species 0's block) phi^n
// and the padding.
__device__ __forceinline__ void bs_record(double* __restrict__ rec, int R, int S, int64_t i,
                                          int k, double val, const double* __restrict__ phin) {
  double* o = rec + i * R;
  o[k] = val;
  if (k == 0) {
    o[S] = phin[i];
    for (int q = S + 1; q < R; q++) o[q] = 0.0;
  }
}

// b = -F (species fields of the residual), r = rh = b, x = p = v = 0; per species b.b
__global__ __launch_bounds__(NNPP_BLOCK) void k_bs_start(int32_t n, const double* __restrict__ F,
                                                         double* __restrict__ x,
                                                         double* __restrict__ r,
                                                         double* __restrict__ rh,
                                                         double* __restrict__ p,
                                                         double* __restrict__ v,
                                                         double* __restrict__ part) {
  const int64_t o = (int64_t)blockIdx.y * n;
  double acc = 0.0;
  const int64_t nb = nnpp_nblocks(n);
  for (nnpp_block_iter it = nnpp_blocks_of(nb); it.cur < it.end; it.cur += it.stride) {
    const int64_t i = it.cur * NNPP_BLOCK + threadIdx.x;
    if (i >= n) continue;
    const int64_t q = o + i;
    const double b = -F[q];
    r[q] = b;
    rh[q] = b;
    x[q] = 0.0;
    p[q] = 0.0;
    v[q] = 0.0;
    acc += b * b;
  }
  bs_partial(acc, 0.0, part);
}

// p = r + beta (p - omega v) for the active species; hat (and the records, if given) get
// p / diag, 0 for inactive species. With mask 0 this only stages zeros (and phi^n).
// Grid (node blocks, S).
__global__ __launch_bounds__(NNPP_BLOCK) void k_bs_dir(int32_t n, int S, uint32_t mask,
                                                       Coef beta, Coef omega,
                                                       const double* __restrict__ r,
                                                       double* __restrict__ p,
                                                       const double* __restrict__ v,
                                                       const double* __restrict__ diag,
                                                       const double* __restrict__ phin,
                                                       double* __restrict__ rec, int R,
                                                       double* __restrict__ hat) {
  const int k = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * NNPP_BLOCK + threadIdx.x;
  if (i >= n) return;
  const int64_t q = (int64_t)k * n + i;
  double ph = 0.0;
  if ((mask >> k) & 1u) {
    const double pn = r[q] + beta.v[k] * (p[q] - omega.v[k] * v[q]);
    p[q] = pn;
    ph = pn / diag[q];
  }
  hat[q] = ph;
  if (rec) bs_record(rec, R, S, i, k, ph, phin);
}

// s = r - alpha v for the active species; hat (and the records, if given) get s / diag;
// per species s.s
__global__ __launch_bounds__(NNPP_BLOCK) void k_bs_s(int32_t n, int S, uint32_t mask, Coef alpha,
                                                     const double* __restrict__ r,
                                                     const double* __restrict__ v,
                                                     double* __restrict__ s,
                                                     const double* __restrict__ diag,
                                                     const double* __restrict__ phin,
                                                     double* __restrict__ rec, int R,
                                                     double* __restrict__ hat,
                                                     double* __restrict__ part) {
  const int k = blockIdx.y;
  const bool on = (mask >> k) & 1u;
  const double al = alpha.v[k];
  const int64_t o = (int64_t)k * n;
  double acc = 0.0;
  const int64_t nb = nnpp_nblocks(n);
  for (nnpp_block_iter it = nnpp_blocks_of(nb); it.cur < it.end; it.cur += it.stride) {
    const int64_t i = it.cur * NNPP_BLOCK + threadIdx.x;
    if (i >= n) continue;
    const int64_t q = o + i;
    double sh = 0.0;
    if (on) {
      const double sv = r[q] - al * v[q];
      s[q] = sv;
      sh = sv / diag[q];
      acc += sv * sv;
    }
    hat[q] = sh;
    if (rec) bs_record(rec, R, S, i, k, sh, phin);
  }
  bs_partial(acc, 0.0, part);
}

// per species a0.b0 (first half) and a1.b1 (second half, if a1)
__global__ __launch_bounds__(NNPP_BLOCK) void k_bs_dots(int32_t n,
                                                        const double* __restrict__ a0,
                                                        const double* __restrict__ b0,
                                                        const double* __restrict__ a1,
                                                        const double* __restrict__ b1,
                                                        double* __restrict__ part) {
  const int64_t o = (int64_t)blockIdx.y * n;
  double acc0 = 0.0, acc1 = 0.0;
  const int64_t nb = nnpp_nblocks(n);
  for (nnpp_block_iter it = nnpp_blocks_of(nb); it.cur < it.end; it.cur += it.stride) {
    const int64_t i = it.cur * NNPP_BLOCK + threadIdx.x;
    if (i >= n) continue;
    const int64_t q = o + i;
    acc0 += a0[q] * b0[q];
    if (a1) acc1 += a1[q] * b1[q];
  }
  bs_partial(acc0, acc1, part);
}

// x += alpha p/diag + omega s/diag, r = s - omega t for the active species; per species rh.r
// (first half) and r.r (second half) of the (possibly unchanged) residual
__global__ __launch_bounds__(NNPP_BLOCK) void k_bs_update(
    int32_t n, uint32_t mask, Coef alpha, Coef omega, double* __restrict__ x,
    double* __restrict__ r, const double* __restrict__ p, const double* __restrict__ s,
    const double* __restrict__ t, const double* __restrict__ diag,
    const double* __restrict__ rh, double* __restrict__ part) {
  const int k = blockIdx.y;
  const bool on = (mask >> k) & 1u;
  const double al = alpha.v[k], om = omega.v[k];
  const int64_t o = (int64_t)k * n;
  double acc0 = 0.0, acc1 = 0.0;
  const int64_t nb = nnpp_nblocks(n);
  for (nnpp_block_iter it = nnpp_blocks_of(nb); it.cur < it.end; it.cur += it.stride) {
    const int64_t i = it.cur * NNPP_BLOCK + threadIdx.x;
    if (i >= n) continue;
    const int64_t q = o + i;
    double rv = r[q];
    if (on) {
      const double dg = diag[q];
      x[q] += al * (p[q] / dg) + om * (s[q] / dg);
      rv = s[q] - om * t[q];
      r[q] = rv;
    }
    acc0 += rh[q] * rv;
    acc1 += rv * rv;
  }
  bs_partial(acc0, acc1, part);
}

// Sum of the NV-wide partials of nblk blocks in a fixed order: thread t adds column t % NV of
// blocks t / NV, t / NV + G, ... (G = NNPP_BLOCK / NV groups), then the G group sums are added
// in group order. One block of NNPP_BLOCK threads.
__global__ void k_bs_final(const double* __restrict__ part, int64_t nblk,
                           double* __restrict__ out) {
  constexpr int G = NNPP_BLOCK / NV;
  __shared__ double grp[G][NV];
  const int k = threadIdx.x % NV, q = threadIdx.x / NV;
  double a = 0.0;
  for (int64_t b = q; b < nblk; b += G) a += part[b * NV + k];
  grp[q][k] = a;
  __syncthreads();
  if (threadIdx.x < NV) {
    double t = 0.0;
    for (int j = 0; j < G; j++) t += grp[j][threadIdx.x];
    out[threadIdx.x] = t;
  }
}

// main.cpp:556-640: the first iterate is U^n with the Dirichlet values
__global__ void k_newton_init(int32_t n, int F, const uint16_t* __restrict__ bcmask,
                              const double* __restrict__ bcval, const double* __restrict__ Un,
                              double* __restrict__ U1) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned m = bcmask[i];
  for (int f = 0; f < F; f++) {
    const int64_t q = (int64_t)f * n + i;
    U1[q] = ((m >> f) & 1u) ? bcval[q] : Un[q];
  }
}

// Newton update U1 += dU (species from the BiCGSTAB solution, phi from dphi), per field
// ||dU_f||^2, then the clamp of main.cpp:805-837
__global__ __launch_bounds__(NNPP_BLOCK) void k_newton_update(
    int32_t n, int S, double* __restrict__ U1, const double* __restrict__ dUs,
    const double* __restrict__ dphi, const uint8_t* __restrict__ phase, uint32_t clamp_always,
    uint32_t clamp_liquid, double floor_v, double* __restrict__ part) {
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) acc[k] = 0.0;
  const int64_t nb = nnpp_nblocks(n);
  for (nnpp_block_iter it = nnpp_blocks_of(nb); it.cur < it.end; it.cur += it.stride) {
    const int64_t i = it.cur * NNPP_BLOCK + threadIdx.x;
    if (i >= n) continue;
    const int liquid = (phase[i] == 0);
#pragma unroll
    for (int f = 0; f <= NNPP_MAX_SPECIES; f++) {
      if (f <= S) {
        const int64_t q = (int64_t)f * n + i;
        const double d = (f < S) ? dUs[q] : dphi[i];
        double u = U1[q] + d;
        acc[f] += d * d;
        if (f < S && (((clamp_always >> f) & 1u) || (((clamp_liquid >> f) & 1u) && liquid)) &&
            u < floor_v)
          u = floor_v;
        U1[q] = u;
      }
    }
  }
  nnpp_block_partial<NV>(acc, part);
}

// Potential block: the Dirichlet part of the update (identity rows: dU = -F) as a vector
// vector; *any is set to 1 if an entry is nonzero (it must be 0 before the launch; lanes that
// set it all store the same value)
__global__ void k_newton_phi_dir(int32_t n, int S, const uint16_t* __restrict__ bcmask,
                                 const double* __restrict__ Fphi, double* __restrict__ dd,
                                 int32_t* __restrict__ any) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = ((bcmask[i] >> S) & 1u) ? -Fphi[i] : 0.0;
  dd[i] = d;
  if (d != 0.0) *any = 1;
}

// Right-hand side of the free rows, psc (-F - K11_fD dU_D) (coup = the operator applied to the
// Dirichlet part), and dU = -F on every other row (identity rows of K11)
__global__ void k_newton_phi_rhs(int32_t n, const uint8_t* __restrict__ pfree,
                                 const double* __restrict__ psc, const double* __restrict__ Fphi,
                                 const double* __restrict__ coup, double* __restrict__ rhs,
                                 double* __restrict__ dphi) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (pfree[i]) {
    rhs[i] = psc[i] * (-Fphi[i] - coup[i]);
    dphi[i] = 0.0;
  } else {
    rhs[i] = 0.0;
    dphi[i] = -Fphi[i];
  }
}

// in (species fields) and phi^n into the Jacobian's gather records
__global__ void k_jac_stage(int32_t n, int S, const double* __restrict__ in,
                            const double* __restrict__ phin, double* __restrict__ rec, int R) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double* o = rec + i * R;
  for (int k = 0; k < S; k++) o[k] = in[(int64_t)k * n + i];
  o[S] = phin[i];
  for (int k = S + 1; k < R; k++) o[k] = 0.0;
}

int imp_alloc(nnpp_ctx* c) {
  const int S = c->p.n_species;
  const int64_t n = c->n;
  if (c->imp && c->imp_n == n && c->imp_S == S) return NNPP_OK;
  nnpp_implicit_free(c);
  const int64_t m = n > 0 ? n : 1;
  NNPP_HIP(c, hipMalloc((void**)&c->imp, sizeof(double) * 10 * S * m));
  NNPP_HIP(c, hipMalloc((void**)&c->imp_prev, sizeof(double) * (S + 1) * m));
  NNPP_HIP(c, hipMalloc((void**)&c->imp_dphi, sizeof(double) * 3 * m));
  NNPP_HIP(c, hipMalloc((void**)&c->imp_part, sizeof(double) * NNPP_RED_GRID * NV));
  NNPP_HIP(c, hipMalloc((void**)&c->imp_red, sizeof(double) * NV));
  NNPP_HIP(c, hipMalloc((void**)&c->imp_flag, sizeof(int32_t)));
  NNPP_HIP(c, hipHostMalloc((void**)&c->h_imp, sizeof(double) * NV, hipHostMallocDefault));
  c->imp_n = n;
  c->imp_S = S;
  c->imp_first = true;
  return NNPP_OK;
}

// Sum the partials of the last reducing launch (g blocks) into out[NV] on the host.
int imp_reduce(nnpp_ctx* c, int64_t g, double* out) {
  k_bs_final<<<1, NNPP_BLOCK, 0, c->stream>>>(c->imp_part, g, c->imp_red);
  NNPP_HIP(c, hipGetLastError());
  NNPP_HIP(c, hipMemcpyAsync(c->h_imp, c->imp_red, sizeof(double) * NV, hipMemcpyDeviceToHost,
                             c->stream));
  NNPP_HIP(c, hipStreamSynchronize(c->stream));
  std::memcpy(out, c->h_imp, sizeof(double) * NV);
  return NNPP_OK;
}

// Species vectors inside c->imp
// hat: the vector the Jacobian is applied to (p or s over the diagonal), species-major; the
// records (c->aos) hold the same values for the column form.
struct ImpVecs {
  double *x, *r, *rh, *p, *v, *s, *t, *d0, *diag, *hat;
};
ImpVecs imp_vecs(nnpp_ctx* c) {
  const int64_t sn = (int64_t)c->p.n_species * c->n;
  double* b = c->imp;
  return ImpVecs{b, b + sn, b + 2 * sn, b + 3 * sn, b + 4 * sn, b + 5 * sn, b + 6 * sn,
                 b + 7 * sn, b + 8 * sn, b + 9 * sn};
}

// K_k x_k = -F_k (species fields of c->Fres) for every species together by Jacobi-
// preconditioned BiCGSTAB from x = 0; x in imp_vecs(c).x. The frozen Jacobian (d0, diag) must
// be set up.
int imp_bicgstab(nnpp_ctx* c, double BV, double rtol, int32_t maxit, int32_t* iters,
                 double* relmax) {
  const int S = c->p.n_species;
  const int32_t n = c->n;
  hipStream_t st = c->stream;
  const int R = jac_rec_len(S);
  const int64_t nb = nnpp_nblocks(n);
  const int64_t gx = bs_grid(nb, S);
  const dim3 g((unsigned)gx, (unsigned)S), gn((unsigned)nb, (unsigned)S);
  const ImpVecs V = imp_vecs(c);
  const double* Un = c->U[NNPP_SLOT_N];
  const double* phin = Un + (int64_t)S * n;
  // the column form of the Jacobian reads the gather records, the stencil form reads hat
  double* rec = nnpp_flux_stencil_active(c) ? nullptr : c->aos;
  double red[NV];
  k_bs_start<<<g, NNPP_BLOCK, 0, st>>>(n, c->Fres, V.x, V.r, V.rh, V.p, V.v, c->imp_part);
  NNPP_HIP(c, hipGetLastError());
  int e = imp_reduce(c, gx, red);
  if (e) return e;
  double bn2[NNPP_MAX_SPECIES], rr[NNPP_MAX_SPECIES], rho[NNPP_MAX_SPECIES];
  double rho_prev[NNPP_MAX_SPECIES], al[NNPP_MAX_SPECIES], om[NNPP_MAX_SPECIES];
  uint32_t mask = 0;
  for (int k = 0; k < S; k++) {
    bn2[k] = rr[k] = rho[k] = red[k];
    rho_prev[k] = al[k] = om[k] = 1.0;
    if (!std::isfinite(bn2[k]))
      return nnpp_set_error(c, NNPP_ERR_NONFINITE, "non-finite Newton residual");
    if (bn2[k] > 0.0) mask |= 1u << k;
  }
  int it = 0;
  while (mask && it < maxit) {
    Coef beta{}, omega{}, alpha{};
    for (int k = 0; k < S; k++) {
      if (!((mask >> k) & 1u)) continue;
      beta.v[k] = (rho[k] / rho_prev[k]) * (al[k] / om[k]);
      omega.v[k] = om[k];
      rho_prev[k] = rho[k];
    }
    k_bs_dir<<<gn, NNPP_BLOCK, 0, st>>>(n, S, mask, beta, omega, V.r, V.p, V.v, V.diag, phin,
                                        rec, R, V.hat);
    NNPP_HIP(c, hipGetLastError());
    if ((e = nnpp_launch_jac(c, BV, Un, V.hat, V.d0, V.v))) return e;
    k_bs_dots<<<g, NNPP_BLOCK, 0, st>>>(n, V.rh, V.v, nullptr, nullptr, c->imp_part);
    NNPP_HIP(c, hipGetLastError());
    if ((e = imp_reduce(c, gx, red))) return e;
    for (int k = 0; k < S; k++) {
      if (!((mask >> k) & 1u)) continue;
      if (red[k] == 0.0 || !std::isfinite(red[k])) {  // breakdown: keep the current x_k
        mask &= ~(1u << k);
        continue;
      }
      al[k] = rho[k] / red[k];
      alpha.v[k] = al[k];
    }
    k_bs_s<<<g, NNPP_BLOCK, 0, st>>>(n, S, mask, alpha, V.r, V.v, V.s, V.diag, phin, rec, R,
                                     V.hat, c->imp_part);
    NNPP_HIP(c, hipGetLastError());
    if ((e = imp_reduce(c, gx, red))) return e;
    uint32_t fin = 0;  // ||s|| already small: x += alpha p^, r = s (omega = 0)
    for (int k = 0; k < S; k++)
      if (((mask >> k) & 1u) && red[k] <= rtol * rtol * bn2[k]) fin |= 1u << k;
    if ((e = nnpp_launch_jac(c, BV, Un, V.hat, V.d0, V.t))) return e;
    k_bs_dots<<<g, NNPP_BLOCK, 0, st>>>(n, V.t, V.s, V.t, V.t, c->imp_part);
    NNPP_HIP(c, hipGetLastError());
    if ((e = imp_reduce(c, gx, red))) return e;
    for (int k = 0; k < S; k++) {
      if (!((mask >> k) & 1u)) continue;
      const double tt = red[NNPP_MAX_SPECIES + k];
      om[k] = ((fin >> k) & 1u) || !(tt > 0.0) ? 0.0 : red[k] / tt;
      omega.v[k] = om[k];
    }
    k_bs_update<<<g, NNPP_BLOCK, 0, st>>>(n, mask, alpha, omega, V.x, V.r, V.p, V.s, V.t,
                                          V.diag, V.rh, c->imp_part);
    NNPP_HIP(c, hipGetLastError());
    if ((e = imp_reduce(c, gx, red))) return e;
    it++;
    for (int k = 0; k < S; k++) {
      if (!((mask >> k) & 1u)) continue;
      rho[k] = red[k];
      rr[k] = red[NNPP_MAX_SPECIES + k];
      if (!std::isfinite(rr[k]))
        return nnpp_set_error(c, NNPP_ERR_NONFINITE, "non-finite species solve");
      if (((fin >> k) & 1u) || rr[k] <= rtol * rtol * bn2[k] || om[k] == 0.0 || rho[k] == 0.0)
        mask &= ~(1u << k);
    }
  }
  double worst = 0.0;
  for (int k = 0; k < S; k++)
    if (bn2[k] > 0.0) worst = std::fmax(worst, std::sqrt(rr[k] / bn2[k]));
  *iters = it;
  *relmax = worst;
  return mask ? NNPP_ERR_NOCONV : NNPP_OK;
}

// Frozen Jacobian at U^n: d0 = 1/dt - R' and the species block diagonals (hat and the column
// form's records: zeros and phi^n; the output goes to the t vector, unused).
int imp_setup(nnpp_ctx* c, double BV) {
  const int S = c->p.n_species;
  const int32_t n = c->n;
  const ImpVecs V = imp_vecs(c);
  const double* Un = c->U[NNPP_SLOT_N];
  double* rec = nnpp_flux_stencil_active(c) ? nullptr : c->aos;
  const dim3 gn((unsigned)nnpp_nblocks(n), (unsigned)S);
  k_bs_dir<<<gn, NNPP_BLOCK, 0, c->stream>>>(n, S, 0u, Coef{}, Coef{}, V.r, V.p, V.v, V.diag,
                                             Un + (int64_t)S * n, rec, jac_rec_len(S), V.hat);
  NNPP_HIP(c, hipGetLastError());
  return nnpp_launch_jac(c, BV, Un, V.hat, nullptr, V.t, V.d0, V.diag);
}

}  // namespace

void nnpp_implicit_free(nnpp_ctx* c) {
  void* ps[] = {c->imp, c->imp_prev, c->imp_dphi, c->imp_part, c->imp_red, c->imp_flag};
  for (void* q : ps)
    if (q) hipFree(q);
  c->imp_flag = nullptr;
  if (c->h_imp) hipHostFree(c->h_imp);
  c->imp = c->imp_prev = c->imp_dphi = c->imp_part = c->imp_red = c->h_imp = nullptr;
  c->imp_n = -1;
  c->imp_S = 0;
  c->imp_first = true;
}

int nnpp_implicit_once(nnpp_ctx* c, int precision, const nnpp_newton_opts& o,
                       nnpp_newton_stats& st) {
  int e = imp_alloc(c);
  if (e) return e;
  const int S = c->p.n_species;
  const int F = S + 1;
  const int32_t n = c->n;
  hipStream_t s = c->stream;
  double* Un = c->U[NNPP_SLOT_N];
  double* U1 = c->U[NNPP_SLOT_N1];
  const int64_t nb = nnpp_nblocks(n);
  const int64_t g = nnpp_red_grid(nb);
  int status = NNPP_OK;
  if (c->imp_first) {
    if (n) NNPP_HIP(c, hipMemsetAsync(c->imp_prev, 0, sizeof(double) * F * n, s));
  }
  if ((c->imp_first || o.start == NNPP_NEWTON_START_STATE) && n) {
    k_newton_init<<<nb, NNPP_BLOCK, 0, s>>>(n, F, c->bcmask, c->bcval, Un, U1);
    NNPP_HIP(c, hipGetLastError());
  }
  const double BV = std::pow(10., -c->p.vol_red * (1. - c->voln1rel));  // utils.h:100-116
  if (n && (e = imp_setup(c, BV))) return e;
  const ImpVecs V = imp_vecs(c);
  double* dphi = c->imp_dphi;
  double* dtmp = c->imp_dphi + n;
  double* coup = c->imp_dphi + 2 * (int64_t)n;
  uint32_t clamp_always = 0, clamp_liquid = 0;
  for (int k = 0; k < S; k++) {
    if (c->p.clamp[k] == NNPP_CLAMP_ALWAYS) clamp_always |= 1u << k;
    if (c->p.clamp[k] == NNPP_CLAMP_LIQUID) clamp_liquid |= 1u << k;
  }
  double res_old[NNPP_MAX_SPECIES + 1] = {0.0};
  double red[NV];
  bool converged = false;
  double resmax = 0.0;
  for (int it = 0; it < o.newton_maxit && n; it++) {
    // residual with the Dirichlet rows (main.cpp:667-711)
    if ((e = nnpp_launch_flux(c, 1, NNPP_FP64, BV, U1, Un, c->Fres))) return e;
    if ((e = nnpp_bc_override_dev(c, U1, c->Fres))) return e;
    // species blocks
    int32_t lit = 0;
    double lrel = 0.0;
    e = imp_bicgstab(c, BV, o.lin_rtol, o.lin_maxit, &lit, &lrel);
    if (e == NNPP_ERR_NOCONV) status = NNPP_ERR_NOCONV;
    else if (e) return e;
    st.lin_iterations += lit;
    st.lin_relres_max = std::fmax(st.lin_relres_max, lrel);
    // potential block: K11 dU = -F with the Dirichlet rows' update moved to the right-hand side
    const double* Fphi = c->Fres + (int64_t)S * n;
    // the coupling is 0 (exactly: each bond adds w (0 - 0)) unless a Dirichlet row still
    // moves; the iterate carries the Dirichlet values after the first Newton update
    NNPP_HIP(c, hipMemsetAsync(c->imp_flag, 0, sizeof(int32_t), s));
    k_newton_phi_dir<<<nb, NNPP_BLOCK, 0, s>>>(n, S, c->bcmask, Fphi, dtmp, c->imp_flag);
    NNPP_HIP(c, hipGetLastError());
    if ((e = nnpp_potential_apply_dev(c, dtmp, coup, c->imp_flag))) return e;
    k_newton_phi_rhs<<<nb, NNPP_BLOCK, 0, s>>>(n, c->pfree, c->psc, Fphi, coup, dtmp, dphi);
    NNPP_HIP(c, hipGetLastError());
    const double rtol_keep = c->rtol;
    c->rtol = o.lin_rtol;
    int32_t cit = 0;
    double crr = 0.0;
    e = nnpp_potential_solve_dev(c, precision, nullptr, dphi, &cit, &crr, nullptr, nullptr, dtmp);
    c->rtol = rtol_keep;
    if (e == NNPP_ERR_NOCONV) status = NNPP_ERR_NOCONV;
    else if (e) return e;
    st.cg_iterations += cit;
    // update, norms, clamp
    k_newton_update<<<g, NNPP_BLOCK, 0, s>>>(n, S, U1, V.x, dphi, c->phase, clamp_always,
                                             clamp_liquid, c->p.clamp_floor, c->imp_part);
    NNPP_HIP(c, hipGetLastError());
    if ((e = imp_reduce(c, g, red))) return e;
    st.newton_iterations++;
    resmax = 0.0;
    double res[NNPP_MAX_SPECIES + 1];
    for (int f = 0; f < F; f++) {
      res[f] = std::sqrt(red[f]);
      if (!std::isfinite(res[f]))
        return nnpp_set_error(c, NNPP_ERR_NONFINITE, "non-finite Newton update");
      resmax = std::fmax(resmax, std::fabs(res[f] - res_old[f]));
    }
    if (resmax < o.newton_tol) {  // main.cpp:863-911
      converged = true;
      break;
    }
    for (int f = 0; f < F; f++) res_old[f] = res[f];
  }
  if (!converged && n) st.unconverged_steps++;
  st.last_resmax = resmax;
  // front update on the iterate and the volume cadence (main.cpp:915-949)
  int32_t flips = 0;
  double volsum = 0.0;
  if ((e = nnpp_front_dev(c, U1, c->p.front_threshold, 1, &flips, &volsum))) return e;
  const bool vol_update = c->p.sub_steps > 0 && (c->step % c->p.sub_steps) == 0 &&
                          c->vol0rel != 0.0;
  if (vol_update) c->voln1rel = volsum / c->vol0rel;
  c->step++;
  // rotation (main.cpp:1002-1024): (U^{n-1}, U^n, iterate) <- (U^n, iterate, U^{n-1})
  double* old_prev = c->imp_prev;
  c->imp_prev = c->U[NNPP_SLOT_N];
  c->U[NNPP_SLOT_N] = c->U[NNPP_SLOT_N1];
  c->U[NNPP_SLOT_N1] = old_prev;
  c->imp_first = false;
  c->phi_hist = 0;  // the explicit step's potential history no longer applies
  st.steps++;
  st.flipped += flips;
  return status;
}

int nnpp_jacobian_apply_dev(nnpp_ctx* c, double voln1rel, const double* in, double* out,
                            double* diag) {
  int e = imp_alloc(c);
  if (e) return e;
  const int S = c->p.n_species;
  const int32_t n = c->n;
  if (!n) return NNPP_OK;
  const double BV = std::pow(10., -c->p.vol_red * (1. - voln1rel));
  const ImpVecs V = imp_vecs(c);
  const double* Un = c->U[NNPP_SLOT_N];
  k_jac_stage<<<nnpp_nblocks(n), NNPP_BLOCK, 0, c->stream>>>(n, S, in, Un + (int64_t)S * n,
                                                             c->aos, jac_rec_len(S));
  NNPP_HIP(c, hipGetLastError());
  if ((e = nnpp_launch_jac(c, BV, Un, in, nullptr, out, V.d0, V.diag))) return e;
  if (diag)
    NNPP_HIP(c, hipMemcpyAsync(diag, V.diag, sizeof(double) * S * n, hipMemcpyDeviceToDevice,
                               c->stream));
  return nnpp_jac_phi_dev(c, in + (int64_t)S * n, out + (int64_t)S * n,
                          diag ? diag + (int64_t)S * n : nullptr);
}

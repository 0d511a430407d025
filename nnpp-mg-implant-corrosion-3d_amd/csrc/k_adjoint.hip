// k_adjoint.hip — reverse mode of the explicit FP64 step (nnpp_tape_adjoint).
//
// Forward step (k_flux, explicit update) for species k at node i, bond classes c = LL / SS
// (same phase) and SL (mixed), w_ij = Beta(xi) V_j / xi^(3+2s) the stored bond weight:
//   G_k,i = -sum_c [ A_k,c sum_{j in c} w_ij (U_k,j - U_k,i) + M_k,c sum_{j in c} w_ij U_k,j
//                    (phi_j - phi_i) ] + r_diss U_0,i [generic sink, interface solid nodes]
//   U_k,i^{n+1} = U_k,i - dt G_k,i, then Dirichlet rows and the clamp.
// With g_k,i = dL/dG_k,i = -dt dL/dU_k,i^{n+1} on entries that pass (not Dirichlet, not
// clamped), the transposed bond terms at node i gather over the same bond list (bonds are
// symmetric and w_ji = w_ij V_i / V_j); with h_k,j = g_k,j / V_j staged per node:
//   dL/dU_k,i += A (g_k,i W - V_i TH_k) - M V_i TM_k
//   dL/dphi_i += sum_k M (g_k,i PC_k - V_i U_k,i TH_k)
// where, per class, W = sum w_ij, TH_k = sum w_ij h_k,j, TM_k = sum w_ij h_k,j (phi_i - phi_j),
// PC_k = sum w_ij U_k,j. The class sums of the forward (sd_k = PC_k - U_k,i W,
// SM_k = sum w_ij U_k,j (phi_j - phi_i)) give the parameter derivatives:
//   dL/dD_liquid,k += -g_k,i (K1 sd + K2 z F/(RT) SM)          (LL class)
//   dL/dD_solid,k  += the same over SS bonds (species with solid bonds), and BV times it over
//                     SL bonds; dL/dBV += D_solid times it over SL bonds (solid-bond species).
// The potential enters as phi^{n+1} = M^{-1} b(U^n) on free nodes: psi = M^{-1} dL/dphi^{n+1}
// (M symmetric), dL/dU_k,i += psi_i s_i (F CSolid / (K11 sigma1)) z_k.
#include <cmath>
#include <cstring>
#include <vector>

#include "nnpp_internal.hpp"

namespace {

struct AdjArgs {
  int32_t n;
  const int64_t* off;
  const int32_t* len;
  const int32_t* col;
  const uint16_t* d16;
  const int32_t* first;
  const double* w64;
  const double* rec;    // per node [U_0 .. U_{S-1}, phi, h_0 .. h_{S-1}] (k_adj_pack)
  const uint8_t* phase;
  const uint16_t* pass;
  const double* lam;    // dL/dU^{n+1}, field-major
  const double* psi;    // adjoint potential (free nodes)
  const uint8_t* pfree;
  const double* psc;
  const uint16_t* bcmask;
  const double* vol;    // nullptr: every node has volume V0
  double V0, dt, r_diss;
  int32_t generic;
  uint32_t sb;          // species with solid-solid bonds
  double* lam_out;      // dL/dU^n
  double* part;         // per-block parameter partials
  double A_LL[NNPP_MAX_SPECIES], A_SL[NNPP_MAX_SPECIES], A_SS[NNPP_MAX_SPECIES];
  double M_LL[NNPP_MAX_SPECIES], M_SL[NNPP_MAX_SPECIES], M_SS[NNPP_MAX_SPECIES];
  double P1[NNPP_MAX_SPECIES], P2[NNPP_MAX_SPECIES];  // dA/dD = K1, dM/dD = K2 z F/(RT)
  double bvSL[NNPP_MAX_SPECIES], Ds[NNPP_MAX_SPECIES];
  double cqz[NNPP_MAX_SPECIES];                        // F CSolid z_k / (K11 sigma1)
};

template <int S>
struct AdjRec {
  static constexpr int used = 2 * S + 1;
  static constexpr int len = (used + 1) & ~1;
};

// Gather records: U^n and phi^n of every node with h_k = g_k / V.
template <int S>
__global__ void k_adj_pack(int32_t n, const double* __restrict__ U, const double* __restrict__ lam,
                           const uint16_t* __restrict__ pass, const double* __restrict__ vol,
                           double V0, double dt, double* __restrict__ rec) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr int R = AdjRec<S>::len;
  double v[R];
#pragma unroll
  for (int k = 0; k <= S; k++) v[k] = U[(int64_t)k * n + i];
  const double Vi = vol ? vol[i] : V0;
  const unsigned pb = pass[i];
#pragma unroll
  for (int k = 0; k < S; k++)
    v[S + 1 + k] = ((pb >> k) & 1u) ? (-dt * lam[(int64_t)k * n + i]) / Vi : 0.0;
#pragma unroll
  for (int k = AdjRec<S>::used; k < R; k++) v[k] = 0.0;
  double2* out = reinterpret_cast<double2*>(rec + i * R);
#pragma unroll
  for (int k = 0; k < R / 2; k++) out[k] = make_double2(v[2 * k], v[2 * k + 1]);
}

template <int R>
__device__ __forceinline__ void load_rec(const double* __restrict__ rec, int64_t j, double (&v)[R]) {
  const double2* r = reinterpret_cast<const double2*>(rec + j * R);
#pragma unroll
  for (int k = 0; k < R / 2; k++) {
    const double2 t = r[k];
    v[2 * k] = t.x;
    v[2 * k + 1] = t.y;
  }
}

template <int S, bool WIDE>
__device__ __forceinline__ void adj_node(const AdjArgs& a, int64_t i, double (&gp)[2 * S + 2]) {
  constexpr int R = AdjRec<S>::len;
  const int64_t n = a.n;
  const int64_t base = a.off[i / NNPP_WAVE] + 4 * (i % NNPP_WAVE);
  const int G = a.len[i / NNPP_WAVE];
  const int pi = a.phase[i];
  double ci[S], phii;
  {
    double v[R];
    load_rec<R>(a.rec, i, v);
#pragma unroll
    for (int k = 0; k < S; k++) ci[k] = v[k];
    phii = v[S];
  }
  const double Vi = a.vol ? a.vol[i] : a.V0;
  const unsigned pb = a.pass[i];
  double gi[S];
#pragma unroll
  for (int k = 0; k < S; k++) gi[k] = ((pb >> k) & 1u) ? -a.dt * a.lam[(int64_t)k * n + i] : 0.0;

  double lc[S];
#pragma unroll
  for (int k = 0; k < S; k++) lc[k] = 0.0;
  double lphi = 0.0;
  int mixed = 0;
  // pass 0: same-phase bonds (others weigh 0); pass 1 (interface nodes): solid-liquid bonds
  for (int ps = 0; ps < 2; ps++) {
    if (ps == 1 && !mixed) break;
    double W = 0.0, TH[S], TM[S], PC[S], SM[S];
#pragma unroll
    for (int k = 0; k < S; k++) TH[k] = TM[k] = PC[k] = SM[k] = 0.0;
    int32_t jr = WIDE ? 0 : a.first[i];
    for (int gq = 0; gq < G; gq++) {
      const int64_t s0 = base + 256 * (int64_t)gq;
      int32_t cq[4];
      if (WIDE) {
        const int4 q = *reinterpret_cast<const int4*>(a.col + s0);
        cq[0] = q.x;
        cq[1] = q.y;
        cq[2] = q.z;
        cq[3] = q.w;
      } else {
        const uint2 d = *reinterpret_cast<const uint2*>(a.d16 + s0);
        cq[0] = jr + (int32_t)(d.x & 0xffffu);
        cq[1] = cq[0] + (int32_t)(d.x >> 16);
        cq[2] = cq[1] + (int32_t)(d.y & 0xffffu);
        cq[3] = cq[2] + (int32_t)(d.y >> 16);
        jr = cq[3];
      }
      const double2 t0 = *reinterpret_cast<const double2*>(a.w64 + s0);
      const double2 t1 = *reinterpret_cast<const double2*>(a.w64 + s0 + 2);
      const double wq[4] = {t0.x, t0.y, t1.x, t1.y};
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int32_t j = cq[e];
        const bool same = (a.phase[j] == pi);
        double w = wq[e];
        if (ps == 0) {
          mixed |= (!same && w != 0.0);
          if (!same) w = 0.0;
        } else if (same || w == 0.0) {
          continue;
        }
        double v[R];
        load_rec<R>(a.rec, j, v);
        const double dphi = v[S] - phii;  // phi_j - phi_i
        W += w;
#pragma unroll
        for (int k = 0; k < S; k++) {
          const double wh = w * v[S + 1 + k];
          const double wc = w * v[k];
          TH[k] += wh;
          TM[k] -= wh * dphi;
          PC[k] += wc;
          SM[k] += wc * dphi;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < S; k++) {
      double A, M;
      if (ps == 0) {
        A = pi ? a.A_SS[k] : a.A_LL[k];
        M = pi ? a.M_SS[k] : a.M_LL[k];
      } else {
        A = a.A_SL[k];
        M = a.M_SL[k];
      }
      lc[k] += A * (gi[k] * W - Vi * TH[k]) - M * Vi * TM[k];
      lphi += M * (gi[k] * PC[k] - Vi * ci[k] * TH[k]);
      const double dG = -(a.P1[k] * (PC[k] - ci[k] * W) + a.P2[k] * SM[k]);
      const bool sbk = (a.sb >> k) & 1u;
      if (ps == 0) {
        if (!pi) gp[k] += gi[k] * dG;
        else if (sbk) gp[S + k] += gi[k] * dG;
      } else {
        gp[S + k] += gi[k] * a.bvSL[k] * dG;
        if (sbk) gp[2 * S + 1] += gi[k] * a.Ds[k] * dG;
      }
    }
  }
  if (a.generic && pi == 1 && mixed) {  // interface solid node: G_0 += r_diss U_0
    lc[0] += gi[0] * a.r_diss;
    gp[2 * S] += gi[0] * ci[0];
  }
  const bool fr = a.pfree[i] != 0;
  const double pq = fr ? a.psi[i] * a.psc[i] : 0.0;
#pragma unroll
  for (int k = 0; k < S; k++)
    a.lam_out[(int64_t)k * n + i] =
        lc[k] + (((pb >> k) & 1u) ? a.lam[(int64_t)k * n + i] : 0.0) + pq * a.cqz[k];
  // phi^{n+1} = phi^n on nodes that are neither free nor Dirichlet for phi
  const bool keep = !fr && !((a.bcmask[i] >> S) & 1u);
  a.lam_out[(int64_t)S * n + i] = lphi + (keep ? a.lam[(int64_t)S * n + i] : 0.0);
}

// Grid-stride over node blocks (XCD-contiguous ranges), one partial per block of the 2S+2
// parameter sums [dD_liquid(S), dD_solid(S), dr_diss, dBV].
template <int S, bool WIDE>
__global__ __launch_bounds__(NNPP_BLOCK) void k_adj_flux(const AdjArgs a) {
  double gp[2 * S + 2];
#pragma unroll
  for (int q = 0; q < 2 * S + 2; q++) gp[q] = 0.0;
  const int64_t nb = nnpp_nblocks(a.n);
  for (nnpp_block_iter it = nnpp_blocks_of(nb); it.cur < it.end; it.cur += it.stride) {
    const int64_t i = it.cur * NNPP_BLOCK + threadIdx.x;
    if (i >= a.n) continue;
    adj_node<S, WIDE>(a, i, gp);
  }
  nnpp_block_partial<2 * S + 2>(gp, a.part);
}

// Sum of the block partials in a fixed order: acc[0..2S] += parameter sums; acc[2S+2] holds
// mu = dL/d(voln1rel in force), which takes this step's BV derivative (reset first when the
// step recomputed voln1rel at its end, whose derivative was injected into dL/dU_0 already).
template <int S>
__global__ void k_adj_reduce(const double* __restrict__ part, int64_t nblk, double* acc,
                             int reset_mu, double dbv_dv) {
  double v[2 * S + 2];
  nnpp_final_sum<2 * S + 2>(v, part, nblk);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q <= 2 * S; q++) acc[q] += v[q];
    const double mu = reset_mu ? 0.0 : acc[2 * S + 2];
    acc[2 * S + 2] = mu + dbv_dv * v[2 * S + 1];
  }
}

// voln1rel = sum over solid nodes of (U_0 - CSat/CSolid) / vol0rel, recomputed at the end of
// the step: its derivative mu reaches dL/dU_0 on the solid nodes of the phase after the step.
__global__ void k_mu_inject(int32_t n, double* __restrict__ lam0, const uint8_t* __restrict__ phase1,
                            const double* __restrict__ mu, double inv_vol0) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (phase1[i] == 1) lam0[i] += nnpp_ld_scalar(mu) * inv_vol0;
}

template <int S>
hipError_t launch_adj(const AdjArgs& a, int wide, const double* U, const double* lam,
                      const uint16_t* pass, double* rec, double dt, double* acc, int reset_mu,
                      double dbv_dv, hipStream_t s) {
  const int64_t nb = nnpp_nblocks(a.n);
  const int64_t g = nnpp_red_grid(nb);
  k_adj_pack<S><<<nb, NNPP_BLOCK, 0, s>>>(a.n, U, lam, pass, a.vol, a.V0, dt, rec);
  if (wide) k_adj_flux<S, true><<<g, NNPP_BLOCK, 0, s>>>(a);
  else k_adj_flux<S, false><<<g, NNPP_BLOCK, 0, s>>>(a);
  k_adj_reduce<S><<<1, NNPP_BLOCK, 0, s>>>(a.part, g, acc, reset_mu, dbv_dv);
  return hipGetLastError();
}

// Device buffers of one adjoint call, freed on every return path.
struct DevBufs {
  std::vector<void*> ptrs;
  ~DevBufs() {
    for (void* p : ptrs) hipFree(p);
  }
  template <typename T>
  bool alloc(T** p, size_t bytes) {
    *p = nullptr;
    if (hipMalloc((void**)p, bytes ? bytes : 8) != hipSuccess) return false;
    ptrs.push_back((void*)*p);
    return true;
  }
};

}  // namespace

int nnpp_adjoint_run(nnpp_ctx* c, const double* lam_h, double* grad_h, double* lam0_h,
                     nnpp_adjoint_stats* stats) {
  const nnpp_params& p = c->p;
  const int32_t n = c->n;
  const int S = p.n_species;
  const int NG = 2 * S + 2;
  const int64_t F = S + 1;
  nnpp_tape& T = c->tape;
  const int64_t nrec = (int64_t)T.steps.size();
  const int64_t E = T.every;
  hipStream_t s = c->stream;
  const size_t m = n ? n : 1;
  const size_t fb = sizeof(double) * m * F;
  nnpp_adjoint_stats st_{};
  if (p.chemistry != NNPP_CHEM_GENERIC)
    return nnpp_set_error(c, NNPP_ERR_UNSUPPORTED, "the adjoint covers the generic chemistry");

  DevBufs B;
  double *lamA, *lamB, *psi, *rec, *part, *acc, *Usave, *seg, *PhiSave;
  uint8_t *Psave, *segP;
  uint16_t* pass;
  const int RA = (2 * S + 2) & ~1;
  const int64_t segn = std::max<int64_t>(std::min<int64_t>(E, nrec), 1);
  if (!B.alloc(&lamA, fb) || !B.alloc(&lamB, fb) || !B.alloc(&psi, sizeof(double) * m) ||
      !B.alloc(&rec, sizeof(double) * m * RA) || !B.alloc(&pass, sizeof(uint16_t) * m) ||
      !B.alloc(&part, sizeof(double) * NG * NNPP_RED_GRID) ||
      !B.alloc(&acc, sizeof(double) * (NG + 1)) || !B.alloc(&Usave, fb) || !B.alloc(&Psave, m) ||
      !B.alloc(&PhiSave, sizeof(double) * 2 * m) ||
      !B.alloc(&seg, fb * segn) || !B.alloc(&segP, m * (segn + 1)))
    return nnpp_set_error(c, NNPP_ERR_HIP, "adjoint buffer allocation");
  st_.tape_bytes = T.bytes + (int64_t)(fb * segn + m * (segn + 1));

  NNPP_HIP(c, hipMemcpyAsync(lamA, lam_h, sizeof(double) * n * F, hipMemcpyHostToDevice, s));
  NNPP_HIP(c, hipMemsetAsync(acc, 0, sizeof(double) * (NG + 1), s));
  NNPP_HIP(c, hipMemsetAsync(psi, 0, sizeof(double) * m, s));
  // the current state, restored at the end
  NNPP_HIP(c, hipMemcpyAsync(Usave, c->U[NNPP_SLOT_N], fb, hipMemcpyDeviceToDevice, s));
  NNPP_HIP(c, hipMemcpyAsync(Psave, c->phase, m, hipMemcpyDeviceToDevice, s));
  const int save_hist = c->phi_hist;
  if (save_hist > 0)
    NNPP_HIP(c, hipMemcpyAsync(PhiSave, c->phi_prev, sizeof(double) * 2 * m,
                               hipMemcpyDeviceToDevice, s));
  const int64_t save_step = c->step, save_solid = c->solid_count;
  const double save_vr = c->voln1rel;

  // Every exit after this point restores the state saved above: the reversal overwrites the
  // context's state, phase and step scalars with checkpoint and recomputed states.
  auto restore = [&]() -> hipError_t {
    hipError_t e = hipMemcpyAsync(c->U[NNPP_SLOT_N], Usave, fb, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(c->phase, Psave, m, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess && save_hist > 0)
      e = hipMemcpyAsync(c->phi_prev, PhiSave, sizeof(double) * 2 * m, hipMemcpyDeviceToDevice,
                         s);
    const hipError_t e2 = hipStreamSynchronize(s);
    c->phi_hist = save_hist;
    c->step = save_step;
    c->solid_count = save_solid;
    c->voln1rel = save_vr;
    c->pot_dirty = true;
    return e != hipSuccess ? e : e2;
  };
  std::vector<double> hacc(NG + 1);
  const int run_status = [&]() -> int {

  AdjArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n = n;
  a.off = c->bonds.off;
  a.len = c->bonds.len;
  a.col = c->bonds.col;
  a.d16 = c->bonds.d16;
  a.first = c->bonds.first;
  a.w64 = c->bonds.w64;
  a.rec = rec;
  a.phase = c->phase;
  a.pass = pass;
  a.psi = psi;
  a.pfree = c->pfree;
  a.psc = c->psc;
  a.bcmask = c->bcmask;
  a.vol = c->vol;
  a.V0 = p.dx * p.dy * p.dz;
  a.dt = p.dt;
  a.r_diss = p.r_diss;
  a.generic = 1;
  a.part = part;
  const double FRT = p.F / (p.R * p.T);
  const double cq = (1. / (p.K_phi * p.sigma1)) * p.F * p.CSolid;
  for (int k = 0; k < S; k++) {
    const bool sb = p.solid_bonds[k] != 0;
    a.sb |= sb ? (1u << k) : 0u;
    a.A_LL[k] = p.K1[k] * p.D_liquid[k];
    a.A_SS[k] = sb ? p.K1[k] * p.D_solid[k] : 0.0;
    a.M_LL[k] = p.K2[k] * p.D_liquid[k] * p.z[k] * FRT;
    a.M_SS[k] = sb ? p.K2[k] * p.D_solid[k] * p.z[k] * FRT : 0.0;
    a.P1[k] = p.K1[k];
    a.P2[k] = p.K2[k] * p.z[k] * FRT;
    a.Ds[k] = p.D_solid[k];
    a.cqz[k] = cq * p.z[k];
  }

  const double ln10 = std::log(10.0);
  const int64_t nck = (int64_t)T.ckpts.size();
  int status = NNPP_OK;
  for (int64_t ck = nck - 1; ck >= 0; ck--) {
    const int64_t r0 = ck * E, r1 = std::min<int64_t>(r0 + E, nrec);
    if (r1 <= r0) continue;
    // phase after the segment's last step
    const uint8_t* ph_end = (ck + 1 < nck) ? T.ckpts[ck + 1].phase : Psave;
    NNPP_HIP(c, hipMemcpyAsync(segP + (size_t)(r1 - r0) * m, ph_end, m, hipMemcpyDeviceToDevice, s));
    // restore the checkpoint and recompute the segment's states
    const nnpp_tape_ckpt& K = T.ckpts[ck];
    NNPP_HIP(c, hipMemcpyAsync(c->U[NNPP_SLOT_N], K.U, fb, hipMemcpyDeviceToDevice, s));
    NNPP_HIP(c, hipMemcpyAsync(c->phase, K.phase, m, hipMemcpyDeviceToDevice, s));
    c->phi_hist = K.phi_hist;
    if (K.phi_hist > 0)
      NNPP_HIP(c, hipMemcpyAsync(c->phi_prev, K.phi_prev, sizeof(double) * 2 * m,
                                 hipMemcpyDeviceToDevice, s));
    c->step = K.step;
    c->voln1rel = K.voln1rel;
    c->solid_count = K.solid_count;
    c->pot_dirty = true;
    for (int64_t r = r0; r < r1; r++) {
      NNPP_HIP(c, hipMemcpyAsync(seg + (size_t)(r - r0) * m * F, c->U[NNPP_SLOT_N], fb,
                                 hipMemcpyDeviceToDevice, s));
      NNPP_HIP(c, hipMemcpyAsync(segP + (size_t)(r - r0) * m, c->phase, m,
                                 hipMemcpyDeviceToDevice, s));
      if (r + 1 < r1) {
        nnpp_step_stats tmp;
        std::memset(&tmp, 0, sizeof(tmp));
        const int e = nnpp_step_once(c, NNPP_FP64, tmp, nullptr);
        if (e != NNPP_OK && e != NNPP_ERR_NOCONV) return e;
        st_.recomputed_steps++;
      }
    }
    // reverse the segment
    c->pot_dirty = true;
    for (int64_t r = r1 - 1; r >= r0; r--) {
      const nnpp_tape_step& R = T.steps[r];
      const double* Ur = seg + (size_t)(r - r0) * m * F;
      NNPP_HIP(c, hipMemcpyAsync(c->phase, segP + (size_t)(r - r0) * m, m,
                                 hipMemcpyDeviceToDevice, s));
      if (r + 1 < r1 && R.flips > 0) c->pot_dirty = true;  // the phase differs from step r+1's
      if (R.vol_update && n)
        k_mu_inject<<<nnpp_nblocks(n), NNPP_BLOCK, 0, s>>>(
            n, lamA, segP + (size_t)(r + 1 - r0) * m, acc + NG, 1.0 / c->vol0rel);
      const double BV = std::pow(10., -p.vol_red * (1. - R.voln1rel));
      // forward update of step r again, for the mask of entries the derivative passes
      int e = nnpp_launch_flux(c, 0, NNPP_FP64, BV, Ur, Ur, c->Unew, pass);
      if (e) return e;
      // adjoint potential: M psi = dL/dphi^{r+1} on the free rows
      if (c->pot_dirty && (e = nnpp_build_potential(c))) return e;
      int32_t its = 0;
      double rr = 0.0;
      e = nnpp_potential_solve_dev(c, NNPP_FP64, nullptr, psi, &its, &rr, nullptr, nullptr,
                                   lamA + (size_t)S * n);
      if (e == NNPP_ERR_NOCONV) status = NNPP_ERR_NOCONV;
      else if (e) return e;
      st_.cg_iterations += its;
      st_.last_relres = rr;
      if (n) {
        for (int k = 0; k < S; k++) {
          const bool sb = p.solid_bonds[k] != 0;
          const double bvSL = sb ? BV : 1.0;
          a.A_SL[k] = p.K1[k] * p.D_solid[k] * bvSL;
          a.M_SL[k] = p.K2[k] * p.D_solid[k] * bvSL * p.z[k] * FRT;
          a.bvSL[k] = bvSL;
        }
        a.lam = lamA;
        a.lam_out = lamB;
        hipError_t he = hipSuccess;
        const double dbv = BV * ln10 * p.vol_red;
        switch (S) {
#define NNPP_CASE(SS)                                                                      \
  case SS:                                                                                 \
    he = launch_adj<SS>(a, c->bonds.wide, Ur, lamA, pass, rec, p.dt, acc, R.vol_update, dbv, s); \
    break;
          NNPP_CASE(1) NNPP_CASE(2) NNPP_CASE(3) NNPP_CASE(4) NNPP_CASE(5) NNPP_CASE(6)
          NNPP_CASE(7) NNPP_CASE(8) NNPP_CASE(9) NNPP_CASE(10) NNPP_CASE(11) NNPP_CASE(12)
          NNPP_CASE(13) NNPP_CASE(14) NNPP_CASE(15) NNPP_CASE(16)
#undef NNPP_CASE
          default:
            return nnpp_set_error(c, NNPP_ERR_ARG, "n_species out of range");
        }
        if (he != hipSuccess) return nnpp_set_error(c, NNPP_ERR_HIP, hipGetErrorString(he));
      }
      std::swap(lamA, lamB);
      st_.steps++;
    }
  }

  NNPP_HIP(c, hipMemcpyAsync(hacc.data(), acc, sizeof(double) * (NG + 1), hipMemcpyDeviceToHost, s));
  if (lam0_h)
    NNPP_HIP(c, hipMemcpyAsync(lam0_h, lamA, sizeof(double) * n * F, hipMemcpyDeviceToHost, s));
  return status;
  }();
  // leave the state as it was before the call, on success and on every error
  const hipError_t re = restore();
  if (run_status != NNPP_OK && run_status != NNPP_ERR_NOCONV) return run_status;
  if (re != hipSuccess) return nnpp_set_error(c, NNPP_ERR_HIP, hipGetErrorString(re));
  for (int q = 0; q <= 2 * S; q++) grad_h[q] = hacc[q];
  if (stats) *stats = st_;
  return run_status;
}

// k_flux.hip — fused multi-species nonlocal Nernst-Planck flux (utils.h:27-842, all fields in
// one pass over the bonds).
//
// One thread per node, 64 consecutive nodes per wavefront reading their bonds from the
// wave-height slices in the quad layout (one 8-byte load of four 16-bit column steps and one
// 16/32-byte weight load per four bonds). The fields a bond gathers are staged once per call in
// an AoS record per node (k_pack), so a neighbour costs R/2 16-byte loads instead of S+1
// separate 8-byte gathers; on a lattice the 64 lanes gather 64 consecutive records. Per node
// and species the kernel keeps
//   sd_k = sum_b w_b (c_k,j - c_k,i)            (diffusion, coefficient applied afterwards)
//   sm_k = sum_b w_b c_k,j (phi_j - phi_i)      (electromigration)
// with w_b = Beta(xi) V_j / xi^(3+2s). The diffusivity and BV factor depend only on the bond's
// phase pair (utils.h:100-159, 211-238), so bonds to nodes of the same phase are summed in a
// first, branch-free pass (other bonds and padding weigh 0) and the few solid-liquid bonds of
// interface nodes in a second pass, each pass applying one coefficient per species at the end
// instead of one select per bond.
//
// RESID: F_k = (U1_k - Un_k)/dt + (sum_KR1 + sum_KR2) - R_k, phi row of utils.h:798-841.
// !RESID: explicit update U_k^{n+1} = U_k^n - dt (sum_KR1 + sum_KR2 - R_k), Dirichlet rows,
//         clamp of main.cpp:805-837.
// MIXED: FP32 weights, FP32 differences taken from the FP64 state, FP32 sums over groups of 16
// bond slots added with Kahan compensation per node and species; reactions, the update and the
// output stay FP64.
// JAC: the species blocks of the frozen Jacobian of utils.h:844-1849 applied to v (the gathered
//      record holds v instead of the iterate): out_k = d0_k v_k + sum_KR_k1(v) + sum_KR_k2(v)
//      with d0_k = 1/dt - R'_k (R' at U^n), identity rows where field k is Dirichlet. The
//      bond sums are linear in v with the lagged phi^n, so they are the Jacobian's off-diagonal
//      and diffusion-diagonal terms exactly. With d0_out set the kernel also computes d0 from
//      U^n and the diagonal d0 + sum of the diffusion weights (the Jacobi preconditioner).
#include "nnpp_internal.hpp"

namespace {

struct RxConst {  // reference chemistry constants, pow(10,-pK)/CSolid computed on the host
  double KMgOH2, KMgCO3, KMgHCO3, KMgPO4, KCaCO3, KCa5, KH2O, KHCO3, KHPO4;
  double k1b, k2b, k3b, k4b, k5b, k6b, k7b, k8b, k9b;
  double JHp_c;     // JHp / (z2 F CSolid)
  double RT_Fz1;    // R T / (F z1)
  double sat_rel;   // CSat / CSolid
  double phim_se;   // phim + phimse
  double alphaC, F, R, T;
};

struct FluxArgs {
  int32_t n;
  int32_t chemistry;
  const int64_t* off;
  const int32_t* len;  // groups of 4 entries per lane
  const int32_t* col;
  const uint16_t* d16;
  const int32_t* first;
  const double* w64;
  const float* w32;
  const double* rec;   // AoS gather records (k_pack)
  double* rec_out;
  const uint8_t* phase;
  const uint16_t* bcmask;
  const double* bcval;
  const double* U1;  // iterate (bond sums, reaction n1 terms)
  const double* Un;  // previous state (lagged phi, reaction n terms, time term)
  double* out;
  uint16_t* passmask;  // explicit update: bit k set where U_k^{n+1} is not a Dirichlet or clamped
                       // value (the derivative passes); nullptr: not written
  const double* d0;    // JAC: 1/dt - R'_k per species and node (S*n)
  double* d0_out;      // JAC setup: d0 written here
  double* diag_out;    // JAC setup: diagonal of each species block (1 on Dirichlet rows)
  double A_LL[NNPP_MAX_SPECIES], A_SL[NNPP_MAX_SPECIES], A_SS[NNPP_MAX_SPECIES];
  double M_LL[NNPP_MAX_SPECIES], M_SL[NNPP_MAX_SPECIES], M_SS[NNPP_MAX_SPECIES];
  double z[NNPP_MAX_SPECIES];
  double dt;
  double phi_q;         // (1/(K11prop1 sigma1)) F CSolid
  double clamp_floor;
  uint32_t clamp_always, clamp_liquid;
  double r_diss;
  RxConst rx;
};

__device__ __forceinline__ void kahan_add(float& s, float& cmp, float v) {
  const float y = v - cmp;
  const float t = s + y;
  cmp = (t - s) - y;
  s = t;
}

// Coefficient-weighted sum of one pass, A sd + M sm, with the rounding written out so the
// column and stencil forms of the flux round alike.
__device__ __forceinline__ double pass_sum(double A, double sd, double M, double sm) {
  return __fma_rn(A, sd, M * sm);
}

// Reference reactions (utils.h:162-174, 241-253, 321-330, 396-401, 466-476, 541-547, 612-618,
// 683-690); u1 = iterate at node i, un = previous state at node i.
__device__ double ref_reaction(int k, const RxConst& c, int phase, int intf, const double* u1,
                               const double* un, double phin) {
  const int liquid = (phase == 0);
  double R = 0.0;
  switch (k) {
    case 0:
      if (intf == 1 || liquid)
        R = -c.k1b * fmax((u1[0] * un[2] * un[2] - c.KMgOH2), 0.0)
            - c.k2b * fmax((u1[0] * un[4] - c.KMgCO3), 0.0)
            - c.k3b * fmax((u1[0] * un[4] - c.KMgHCO3), 0.0)
            - c.k4b * fmax((un[0] * un[0] * un[0] * un[5] * un[5] - c.KMgPO4), 0.0);
      break;
    case 1:
      if (intf == 1) {
        const double phic = c.RT_Fz1 * log(un[0] / c.sat_rel);
        const double eta = c.phim_se - phin - phic;
        R += c.JHp_c * exp(-c.alphaC * c.F * eta / (c.R * c.T));
      } else if (liquid) {
        R = c.k7b * (c.KH2O - u1[1] * un[2]) + c.k8b * (c.KHCO3 - u1[1] * un[4]) +
            c.k9b * (c.KHPO4 - u1[1] * un[6]);
      }
      break;
    case 2: {
      if (intf == 1 || liquid) {
        R = -c.k1b * fmax((un[2] * un[2] * un[0] - c.KMgOH2), 0.0)
            - c.k6b * fmax((u1[2] * un[7] * un[7] * un[7] * un[7] * un[7] * un[6] * un[6] * un[6] - c.KCa5), 0.0);
        if (liquid) R = R + c.k7b * (c.KH2O - u1[2] * un[1]);
      }
      break;
    }
    case 3:
      if (liquid) R = c.k8b * (c.KHCO3 - u1[3] * un[1]);
      break;
    case 4:
      if (intf == 1 || liquid) {
        R = -c.k2b * fmax((u1[4] * un[0] - c.KMgCO3), 0.0)
            - c.k3b * fmax((u1[4] * un[0] - c.KMgHCO3), 0.0)
            - c.k5b * fmax((u1[4] * un[7] - c.KCaCO3), 0.0);
        if (liquid) R = R + c.k8b * (c.KHCO3 - un[1] * u1[4]);
      }
      break;
    case 5:
      if (intf == 1 || liquid) {
        R = -c.k4b * fmax((un[0] * un[0] * un[0] * un[5] * un[5] - c.KMgPO4), 0.0);
        if (liquid) R = R + c.k9b * (-c.KHPO4 + un[6] * un[1]);
      }
      break;
    case 6:
      if (intf == 1 || liquid) {
        R = -c.k6b * fmax((un[2] * un[7] * un[7] * un[7] * un[7] * un[7] * un[6] * un[6] * un[6] - c.KCa5), 0.0);
        if (liquid) R = R + c.k9b * (c.KHPO4 - u1[6] * un[1]);
      }
      break;
    case 7:
      if (intf == 1 || liquid)
        R = -c.k5b * fmax((u1[7] * un[4] - c.KCaCO3), 0.0)
            - c.k6b * fmax((un[2] * un[7] * un[7] * un[7] * un[7] * un[7] * un[6] * un[6] * un[6] - c.KCa5), 0.0);
      break;
    default:
      break;
  }
  return R;
}

// R'_k of the frozen Jacobian's diagonal at U^n (reference chemistry: utils.h:1001-1006,
// 1089-1100, 1184-1190, 1274-1279, 1364-1369, 1454-1459, 1545-1550, 1635-1640; none for Na+
// and Cl-). Generic chemistry: the derivative of its sink, -r_diss at interface solid nodes.
__device__ double jac_reaction(int k, int chemistry, const RxConst& c, double r_diss, int phase,
                               int intf, const double* un) {
  const int liquid = (phase == 0);
  if (chemistry != NNPP_CHEM_REFERENCE) return (k == 0 && intf == 1) ? -r_diss : 0.0;
  switch (k) {
    case 1:
      return (intf == 1 || liquid) ? c.k7b * (-un[2]) + c.k8b * (-un[4]) + c.k9b * (-un[6])
                                   : 0.0;
    case 2:
      return (intf != 1 && liquid) ? c.k7b * (-un[1]) : 0.0;
    case 3:
    case 4:
      return liquid ? c.k8b * (-un[1]) : 0.0;
    case 6:
      return liquid ? c.k9b * (-un[1]) : 0.0;
    default:
      return 0.0;
  }
}

// Gather record of node j: [c_0 .. c_{S-1} (iterate), phi^n (lagged), phi^{n+1} (RESID)],
// padded to an even number of doubles so every record is 16-byte aligned.
template <int S, bool RESID>
struct Rec {
  static constexpr int used = RESID ? S + 2 : S + 1;
  static constexpr int len = (used + 1) & ~1;
};

// AoS staging copy of the fields the bonds gather (one pass over the SoA state).
template <int S, bool RESID>
__global__ void k_pack(int32_t n, const double* __restrict__ U1, const double* __restrict__ Un,
                       double* __restrict__ rec) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr int R = Rec<S, RESID>::len;
  double v[R];
#pragma unroll
  for (int k = 0; k < S; k++) v[k] = U1[(int64_t)k * n + i];
  v[S] = Un[(int64_t)S * n + i];
  if (RESID) v[S + 1] = U1[(int64_t)S * n + i];
#pragma unroll
  for (int k = Rec<S, RESID>::used; k < R; k++) v[k] = 0.0;
  double2* out = reinterpret_cast<double2*>(rec + (int64_t)i * R);
#pragma unroll
  for (int k = 0; k < R / 2; k++) out[k] = make_double2(v[2 * k], v[2 * k + 1]);
}

// Explicit update of node i from its bond sums (the previous state is the iterate, U^n):
// U_k^{n+1} = U_k^n - dt (sum_KR1 + sum_KR2 - R_k), Dirichlet rows, clamp of main.cpp:805-837,
// and the rows whose derivative passes (passmask).
template <int S>
__device__ __forceinline__ void explicit_update(const FluxArgs& a, int64_t i, int pi, int intf,
                                                const double (&ci)[S], const double (&acc)[S],
                                                double phii) {
  const int64_t n = a.n;
  const unsigned bm = a.bcmask[i];
  unsigned pass_bits = 0;
#pragma unroll
  for (int k = 0; k < S; k++) {
    double Rk = 0.0;
    if constexpr (S == 10) {
      if (a.chemistry == NNPP_CHEM_REFERENCE) Rk = ref_reaction(k, a.rx, pi, intf, ci, ci, phii);
    }
    if (a.chemistry == NNPP_CHEM_GENERIC && k == 0 && intf) Rk = -a.r_diss * ci[0];
    const double flux = -acc[k];  // sum_KR_k1 + sum_KR_k2
    double v = ci[k] - a.dt * (flux - Rk);
    bool keep = !((bm >> k) & 1u);
    if (!keep) v = a.bcval[(int64_t)k * n + i];
    const bool cl = ((a.clamp_always >> k) & 1u) || (((a.clamp_liquid >> k) & 1u) && pi == 0);
    if (cl && v < a.clamp_floor) {
      v = a.clamp_floor;
      keep = false;
    }
    a.out[(int64_t)k * n + i] = v;
    pass_bits |= keep ? (1u << k) : 0u;
  }
  if (a.passmask) a.passmask[i] = (uint16_t)pass_bits;
}

// Jacobian rows of node i from its bond sums (acc) and weight sums (sw: same-phase and
// solid-liquid passes): out_k = d0_k v_k - acc_k, identity rows where field k is Dirichlet; with
// d0_out set, d0 = 1/dt - R'_k at U^n (un) and the block diagonal are written first.
template <int S>
__device__ __forceinline__ void jac_update(const FluxArgs& a, int64_t i, int pi, int intf,
                                           const double (&v)[S], const double (&acc)[S],
                                           const double (&sw)[2], const double* un) {
  const int64_t n = a.n;
  const unsigned bm = a.bcmask[i];
#pragma unroll
  for (int k = 0; k < S; k++) {
    const bool dir = (bm >> k) & 1u;
    double d0;
    if (a.d0_out) {
      d0 = (1. / a.dt) - jac_reaction(k, a.chemistry, a.rx, a.r_diss, pi, intf, un);
      a.d0_out[(int64_t)k * n + i] = d0;
      const double A0 = pi ? a.A_SS[k] : a.A_LL[k];
      const double dg = d0 + __fma_rn(A0, sw[0], a.A_SL[k] * sw[1]);
      a.diag_out[(int64_t)k * n + i] = dir ? 1.0 : dg;
    } else {
      d0 = a.d0[(int64_t)k * n + i];
    }
    a.out[(int64_t)k * n + i] = dir ? v[k] : __fma_rn(d0, v[k], -acc[k]);
  }
}

template <int S, bool RESID, bool MIXED, bool WIDE, bool JAC = false>
__global__ __launch_bounds__(NNPP_BLOCK) void k_flux(const FluxArgs a) {
  static_assert(!JAC || (!RESID && !MIXED), "the Jacobian apply is FP64 and not a residual");
  const int64_t i = nnpp_xcd_block(blockIdx.x, gridDim.x) * NNPP_BLOCK + threadIdx.x;
  if (i >= a.n) return;
  constexpr int R = Rec<S, RESID>::len;
  const int64_t n = a.n;
  const int64_t base = a.off[i / NNPP_WAVE] + 4 * (i % NNPP_WAVE);
  const int G = a.len[i / NNPP_WAVE];
  const int pi = a.phase[i];
  const double* __restrict__ rec = a.rec;

  double ci[S];
  double phii, phi1i = 0.0;
  {
    const double2* ri = reinterpret_cast<const double2*>(rec + i * R);
    double v[R];
#pragma unroll
    for (int k = 0; k < R / 2; k++) {
      const double2 t = ri[k];
      v[2 * k] = t.x;
      v[2 * k + 1] = t.y;
    }
#pragma unroll
    for (int k = 0; k < S; k++) ci[k] = v[k];
    phii = v[S];
    if (RESID) phi1i = v[S + 1];
  }

  double acc[S];
  int mixed = 0;
  double sphi = 0.0;
  int flagLL = 0;
  double sw[2] = {0.0, 0.0};  // JAC: weight sums of the two passes (diffusion diagonal)

  // pass 0: same-phase bonds (branch-free: other bonds get weight 0); pass 1 (interface nodes
  // only): solid-liquid bonds
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1 && !mixed) break;
    double sd[S], sm[S];
    float fd[S], fm[S], cd[S], cm[S], pd[S], pm[S];
#pragma unroll
    for (int k = 0; k < S; k++) {
      sd[k] = sm[k] = 0.0;
      fd[k] = fm[k] = cd[k] = cm[k] = pd[k] = pm[k] = 0.0f;
    }
    int32_t jr = WIDE ? 0 : a.first[i];
    for (int gq = 0; gq < G; gq++) {
      const int64_t s0 = base + 256 * (int64_t)gq;
      int32_t cq[4];
      if (WIDE) {
        const int4 q = nnpp_ld_stream(reinterpret_cast<const int4*>(a.col + s0));
        cq[0] = q.x;
        cq[1] = q.y;
        cq[2] = q.z;
        cq[3] = q.w;
      } else {
        const uint2 d = nnpp_ld_stream(reinterpret_cast<const uint2*>(a.d16 + s0));
        cq[0] = jr + (int32_t)(d.x & 0xffffu);
        cq[1] = cq[0] + (int32_t)(d.x >> 16);
        cq[2] = cq[1] + (int32_t)(d.y & 0xffffu);
        cq[3] = cq[2] + (int32_t)(d.y >> 16);
        jr = cq[3];
      }
      double wq[4];
      if (MIXED) {
        const float4 t = nnpp_ld_stream(reinterpret_cast<const float4*>(a.w32 + s0));
        wq[0] = t.x;
        wq[1] = t.y;
        wq[2] = t.z;
        wq[3] = t.w;
      } else {
        const double2 t0 = nnpp_ld_stream(reinterpret_cast<const double2*>(a.w64 + s0));
        const double2 t1 = nnpp_ld_stream(reinterpret_cast<const double2*>(a.w64 + s0 + 2));
        wq[0] = t0.x;
        wq[1] = t0.y;
        wq[2] = t1.x;
        wq[3] = t1.y;
      }
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int32_t j = cq[e];
        const int pj = a.phase[j];
        const bool same = (pj == pi);
        double w;
        if (pass == 0) {
          mixed |= (!same && wq[e] != 0.0);
          w = same ? wq[e] : 0.0;
        } else {
          if (same || wq[e] == 0.0) continue;  // rare: interface nodes only
          w = wq[e];
        }
        const double2* rj = reinterpret_cast<const double2*>(rec + (int64_t)j * R);
        double v[R];
#pragma unroll
        for (int k = 0; k < R / 2; k++) {
          const double2 t = rj[k];
          v[2 * k] = t.x;
          v[2 * k + 1] = t.y;
        }
        const double dphi = v[S] - phii;
        if (JAC) sw[pass] += w;
        if (MIXED) {
          const float wf = (float)w, wdp = wf * (float)dphi;
#pragma unroll
          for (int k = 0; k < S; k++) {
            pd[k] = __fmaf_rn(wf, (float)(v[k] - ci[k]), pd[k]);
            pm[k] = __fmaf_rn(wdp, (float)v[k], pm[k]);
          }
        } else {
          const double wdphi = w * dphi;
#pragma unroll
          for (int k = 0; k < S; k++) {
            sd[k] = __fma_rn(w, v[k] - ci[k], sd[k]);
            sm[k] = __fma_rn(wdphi, v[k], sm[k]);
          }
        }
        if (RESID && pass == 0 && pi == 0) {  // liquid-liquid bond of the phi row
          sphi += w * (v[S + 1] - phi1i);
          flagLL |= (w != 0.0);
        }
      }
      if (MIXED && ((gq & 3) == 3 || gq == G - 1)) {  // fold the group of 16 bond slots
#pragma unroll
        for (int k = 0; k < S; k++) {
          kahan_add(fd[k], cd[k], pd[k]);
          kahan_add(fm[k], cm[k], pm[k]);
          pd[k] = pm[k] = 0.0f;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < S; k++) {
      if (MIXED) {
        sd[k] = (double)fd[k];
        sm[k] = (double)fm[k];
      }
      if (pass == 0) {
        const double A = pi ? a.A_SS[k] : a.A_LL[k];
        const double M = pi ? a.M_SS[k] : a.M_LL[k];
        acc[k] = pass_sum(A, sd[k], M, sm[k]);
      } else {
        acc[k] += pass_sum(a.A_SL[k], sd[k], a.M_SL[k], sm[k]);
      }
    }
  }

  const int intf = (pi == 1 && mixed) ? 1 : 0;  // utils.h:82-89
  const double* __restrict__ U1 = a.U1;
  const double* __restrict__ Un = a.Un;
  double un[S];
#pragma unroll
  for (int k = 0; k < S; k++)
    un[k] = (RESID || (JAC && a.d0_out)) ? Un[(int64_t)k * n + i] : ci[k];
  const unsigned bm = a.bcmask[i];
  (void)U1;
  if constexpr (JAC) {
    (void)bm;
    jac_update<S>(a, i, pi, intf, ci, acc, sw, un);
    return;
  }

  if constexpr (!RESID) {
    explicit_update<S>(a, i, pi, intf, ci, acc, phii);
    return;
  }
#pragma unroll
  for (int k = 0; k < S; k++) {
    double Rk = 0.0;
    if constexpr (S == 10) {
      if (a.chemistry == NNPP_CHEM_REFERENCE) Rk = ref_reaction(k, a.rx, pi, intf, ci, un, phii);
    }
    if (a.chemistry == NNPP_CHEM_GENERIC && k == 0 && intf) Rk = -a.r_diss * ci[0];
    a.out[(int64_t)k * n + i] = ((ci[k] - un[k]) / a.dt) + (-acc[k]) - Rk;
  }
  double o = 0.0;
  if (flagLL) {
    double q = 0.0;
#pragma unroll
    for (int k = 0; k < S; k++) q += a.z[k] * un[k];
    o = -sphi - a.phi_q * q;
  }
  a.out[(int64_t)S * n + i] = o;
}

template <int S>
hipError_t launch_s(const FluxArgs& a, int resid, int mixed, int wide, int64_t nb, hipStream_t s) {
  // stage the gathered fields, then the bond pass
  const int64_t npk = (a.n + NNPP_BLOCK - 1) / NNPP_BLOCK;
  if (resid) k_pack<S, true><<<npk, NNPP_BLOCK, 0, s>>>(a.n, a.U1, a.Un, a.rec_out);
  else k_pack<S, false><<<npk, NNPP_BLOCK, 0, s>>>(a.n, a.U1, a.Un, a.rec_out);
#define NNPP_FLUX(R_, M_, W_) k_flux<S, R_, M_, W_><<<nb, NNPP_BLOCK, 0, s>>>(a)
  if (resid) {
    if (mixed) { if (wide) NNPP_FLUX(true, true, true); else NNPP_FLUX(true, true, false); }
    else { if (wide) NNPP_FLUX(true, false, true); else NNPP_FLUX(true, false, false); }
  } else {
    if (mixed) { if (wide) NNPP_FLUX(false, true, true); else NNPP_FLUX(false, true, false); }
    else { if (wide) NNPP_FLUX(false, false, true); else NNPP_FLUX(false, false, false); }
  }
#undef NNPP_FLUX
  return hipGetLastError();
}

template <int S>
hipError_t launch_jac(const FluxArgs& a, int wide, int64_t nb, hipStream_t s) {
  if (wide) k_flux<S, false, false, true, true><<<nb, NNPP_BLOCK, 0, s>>>(a);
  else k_flux<S, false, false, false, true><<<nb, NNPP_BLOCK, 0, s>>>(a);
  return hipGetLastError();
}


// ---------------------------------------------------------------- stencil form -------------
// Lattice form of the explicit update (every bond j - i in the stencil table, one flux weight per
// offset; the tables of the potential operator's stencil form). A block takes one 256-row tile.
// The tile's neighbour window (the segments of the stencil, as in the potential SpMV) is staged
// into LDS as pairs (c_j, phi^n_j), phi once, c species by species, and the phases once; each row
// derives its same-phase bonds (m0) and solid-liquid bonds (m1) from the staged phases. The row
// sums run over the bonds in ascending offset order (the neighbour-list order) with the
// arithmetic of the column form, so the result is the column form's bit for bit; a bond absent
// from a row (or of the other pass) takes weight +0, which adds +-0. The window is read from L2
// once per tile and field instead of once per bond.
constexpr int ST_SLOTS = 24;  // staged values per thread and field: windows up to 24 * 256

struct FluxSt {
  const uint4* bmask;        // bonds of each row
  const uint8_t* perm;       // row of each lane within its tile (rows with every bond first)
  const int32_t* lc;         // LDS position of entry t relative to the lane (padded with 0)
  const double* tf;          // flux weight of entry t (padded with 0)
  const int32_t* seg_start;  // segments: first offset, span, LDS base
  const int32_t* seg_len;
  const int32_t* seg_base;
  int K, R, E;               // entries, segments, staged values per field (E <= ST_SLOTS * 256)
  int off_min;               // smallest offset
  uint32_t nz[4];            // entries with a nonzero weight
};

// Thread t stages window positions e = t + 256 k (slot k), e < E. off[k] is the vector position
// of e relative to the window start r0 + off_min (fixed for the kernel; positions of the last
// slot past the window read the window start and are not stored). A buffer resource based at the window
// start covers [window start, n): its hardware range check returns 0 past n, so the loads need
// no bounds tests; a window that starts before row 0 keeps the resource at row 0 and shifts the
// offsets by the (negative) start, which wraps them past 2^31, beyond the resource's size.
struct StSlots {
  uint32_t off[ST_SLOTS];

  // Segments outermost: a segment's three table words are read together (one scalar round
  // trip per segment; per slot and segment, with the reads behind the range tests, they were
  // several hundred dependent round trips per block), then every slot is tested against it.
  __device__ __forceinline__ void init(const FluxSt& st) {
    int32_t pos[ST_SLOTS];
#pragma unroll
    for (int k = 0; k < ST_SLOTS; k++) pos[k] = st.off_min;
    for (int r = 0; r < st.R; r++) {
      const int b = nnpp_ld_table(st.seg_base, r);
      const int end = b + NNPP_BLOCK + nnpp_ld_table(st.seg_len, r) - 1;
      const int s0 = nnpp_ld_table(st.seg_start, r);
#pragma unroll
      for (int k = 0; k < ST_SLOTS; k++) {
        const int e = (int)threadIdx.x + k * NNPP_BLOCK;
        if (e >= b && e < end) pos[k] = s0 + (e - b);
      }
    }
#pragma unroll
    for (int k = 0; k < ST_SLOTS; k++) off[k] = (uint32_t)(pos[k] - st.off_min);
  }

  template <typename T>
  __device__ __forceinline__ static __amdgpu_buffer_rsrc_t rsrc(const T* src, int64_t n,
                                                               int64_t w0, uint32_t& add) {
    const int64_t b0 = w0 >= 0 ? w0 : 0;
    add = (uint32_t)((w0 - b0) * (int64_t)sizeof(T));
    const int64_t bytes = (n - b0) * (int64_t)sizeof(T);
    return __builtin_amdgcn_make_buffer_rsrc((void*)(src + b0), (short)0,
                                             (int)(bytes < 0x7fffffff ? bytes : 0x7fffffff),
                                             0x00020000);
  }

  // the window of field src (FP64) into component x (c) or y (phi) of the pair array
  template <bool PHI>
  __device__ __forceinline__ void stage(const FluxSt& st, const double* src, int64_t n,
                                        int64_t w0, double2* lds) const {
    uint32_t add;
    const __amdgpu_buffer_rsrc_t rs = rsrc(src, n, w0, add);
    double v[ST_SLOTS];
#pragma unroll
    for (int k = 0; k < ST_SLOTS; k++)
      if (k * NNPP_BLOCK < st.E)
        v[k] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(
                                              rs, off[k] * 8u + add, 0, 0));
#pragma unroll
    for (int k = 0; k < ST_SLOTS; k++) {
      const int e = (int)threadIdx.x + k * NNPP_BLOCK;
      if (k * NNPP_BLOCK < st.E && e < st.E) {  // the last slot is partly outside the window
        if (PHI) lds[e].y = v[k];
        else lds[e].x = v[k];
      }
    }
  }

  __device__ __forceinline__ void stage_u8(const FluxSt& st, const uint8_t* src, int64_t n,
                                           int64_t w0, uint8_t* lds) const {
    uint32_t add;
    const __amdgpu_buffer_rsrc_t rs = rsrc(src, n, w0, add);
    uint8_t v[ST_SLOTS];
#pragma unroll
    for (int k = 0; k < ST_SLOTS; k++)
      if (k * NNPP_BLOCK < st.E) v[k] = __builtin_amdgcn_raw_buffer_load_b8(rs, off[k] + add, 0, 0);
#pragma unroll
    for (int k = 0; k < ST_SLOTS; k++) {
      const int e = (int)threadIdx.x + k * NNPP_BLOCK;
      if (k * NNPP_BLOCK < st.E && e < st.E) lds[e] = v[k];
    }
  }
};

// sd += w (c_j - c_i), sm += (w (phi_j - phi_i)) c_j over the bonds of mask m, in chunks of 16
// entries (LDS positions and weights wave-uniform). Bit c of anyc: some lane has a bond in chunk
// c; of fullc: every lane has every bond of nonzero weight in chunk c (no masks).
__device__ __forceinline__ void st_sums(const FluxSt& st, const double2* rcp,
                                        const uint32_t (&m)[4], uint32_t anyc, uint32_t fullc,
                                        double ci, double phii, double& sd, double& sm) {
  constexpr int BT = 16;
  const int K = st.K;
  for (int c = 0; c * BT < K; c++) {
    if (!((anyc >> c) & 1u)) continue;
    const int t0 = c * BT;
    int lv[BT];
    double wv[BT];
    double2 v[BT];
#pragma unroll
    for (int k = 0; k < BT; k++) {
      lv[k] = nnpp_ld_table(st.lc, t0 + k);
      wv[k] = nnpp_ld_table(st.tf, t0 + k);
    }
#pragma unroll
    for (int k = 0; k < BT; k++) v[k] = rcp[lv[k]];
    if ((fullc >> c) & 1u) {
#pragma unroll
      for (int k = 0; k < BT; k++) {
        sd = __fma_rn(wv[k], v[k].x - ci, sd);
        const double wdphi = wv[k] * (v[k].y - phii);
        sm = __fma_rn(wdphi, v[k].x, sm);
      }
    } else {
      const int wd = t0 >> 5;  // a chunk never straddles a 32-bit mask word
      const uint32_t bits =
          (wd == 0 ? m[0] : (wd == 1 ? m[1] : (wd == 2 ? m[2] : m[3]))) >> (t0 & 31);
#pragma unroll
      for (int k = 0; k < BT; k++) {
        const double w = nnpp_mask_weight<double>(wv[k], (int)(bits << (31 - k)) >> 31);
        sd = __fma_rn(w, v[k].x - ci, sd);
        const double wdphi = w * (v[k].y - phii);
        sm = __fma_rn(wdphi, v[k].x, sm);
      }
    }
  }
}

// sw = sum of the weights of the bonds of mask m in ascending order (the column form's weight
// sums of the Jacobian diagonal; absent bonds add +0).
__device__ __forceinline__ double st_wsum(const FluxSt& st, const uint32_t (&m)[4]) {
  const int K = st.K;
  double sw = 0.0;
  for (int t0 = 0; t0 < K; t0 += 32) {
    const int wd = t0 >> 5;
    const uint32_t bits = wd == 0 ? m[0] : (wd == 1 ? m[1] : (wd == 2 ? m[2] : m[3]));
    const int cnt = K - t0 < 32 ? K - t0 : 32;
    for (int k = 0; k < cnt; k++)
      sw += nnpp_mask_weight<double>(nnpp_ld_table(st.tf, t0 + k), (int)(bits << (31 - k)) >> 31);
  }
  return sw;
}

// Per chunk of 16 entries: some lane has a bond of mask m there (anyc), every lane has every
// bond of nonzero weight there (fullc). Wave-uniform.
__device__ __forceinline__ void st_chunks(const FluxSt& st, const uint32_t (&m)[4],
                                          uint32_t& anyc, uint32_t& fullc) {
  constexpr int BT = 16;
  const int K = st.K;
  anyc = fullc = 0u;
  for (int c = 0; c * BT < K; c++) {
    const int t0 = c * BT;
    const int wd = t0 >> 5;
    const uint32_t bits =
        (wd == 0 ? m[0] : (wd == 1 ? m[1] : (wd == 2 ? m[2] : m[3]))) >> (t0 & 31);
    const uint32_t nzw =
        (wd == 0 ? st.nz[0] : (wd == 1 ? st.nz[1] : (wd == 2 ? st.nz[2] : st.nz[3]))) >> (t0 & 31);
    const uint32_t valid = (K - t0 >= BT ? 0xffffu : (1u << (K - t0)) - 1u) & nzw;
    if (__any(bits & valid)) anyc |= 1u << c;
    if (__all((bits & valid) == valid)) fullc |= 1u << c;
  }
}

// JAC: the species blocks of the frozen Jacobian applied to v (a.U1, species-major) with the
// lagged phi^n, as the column kernel's JAC mode (identical additions and epilogue).
template <int S, bool JAC>
__global__ __launch_bounds__(NNPP_BLOCK) void k_flux_st(const FluxArgs a, const FluxSt st) {
  extern __shared__ double2 fst_lds[];
  double2* lcp = fst_lds;                                         // [E] (c_j, phi_j)
  uint8_t* lph = reinterpret_cast<uint8_t*>(fst_lds + st.E);      // [E] phases
  const int64_t n = a.n;
  const int64_t r0 = nnpp_xcd_block(blockIdx.x, gridDim.x) * NNPP_BLOCK;
  const int64_t w0 = r0 + st.off_min;
  // lanes take the tile's rows with every bond first, so most waves sum without masks
  const int l = st.perm[r0 + threadIdx.x];
  const int64_t i = r0 + l;
  const bool in = i < n;
  // explicit update: the fields of U^n; JAC: the vector v
  const double* __restrict__ U = JAC ? a.U1 : a.Un;
  StSlots sl;
  sl.init(st);
  sl.stage<true>(st, a.Un + (int64_t)S * n, n, w0, lcp);
  sl.stage_u8(st, a.phase, n, w0, lph);
  double ci[S];
#pragma unroll
  for (int k = 0; k < S; k++) ci[k] = in ? U[(int64_t)k * n + i] : 0.0;
  const double phii = in ? a.Un[(int64_t)S * n + i] : 0.0;
  const int pi = in ? a.phase[i] : 0;
  const uint4 bq = in ? st.bmask[i] : make_uint4(0u, 0u, 0u, 0u);
  __syncthreads();
  // same-phase (m0) and solid-liquid (m1) bonds of nonzero weight
  const uint32_t bw[4] = {bq.x, bq.y, bq.z, bq.w};
  uint32_t m0[4], m1[4];
  const uint8_t* rph = lph + l;
  const int K = st.K;
#pragma unroll
  for (int wd = 0; wd < 4; wd++) {
    uint32_t same = 0u;
    if (wd * 32 < K) {
#pragma unroll 8
      for (int t = 0; t < 32; t++) {
        const int tt = wd * 32 + t;
        if (tt < K) same |= (uint32_t)((int)rph[nnpp_ld_table(st.lc, tt)] == pi) << t;
      }
    }
    m0[wd] = bw[wd] & st.nz[wd] & same;
    m1[wd] = bw[wd] & st.nz[wd] & ~same;
  }
  const bool mixed = (m1[0] | m1[1] | m1[2] | m1[3]) != 0u;
  const int intf = (pi == 1 && mixed) ? 1 : 0;  // utils.h:82-89
  uint32_t any0, full0, any1, full1;
  st_chunks(st, m0, any0, full0);
  st_chunks(st, m1, any1, full1);
  const double2* rcp = lcp + l;
  double acc[S];
#pragma unroll
  for (int k = 0; k < S; k++) {
    __syncthreads();  // the previous species' reads of the window are complete
    sl.stage<false>(st, U + (int64_t)k * n, n, w0, lcp);
    __syncthreads();
    double sd = 0.0, sm = 0.0;
    st_sums(st, rcp, m0, any0, full0, ci[k], phii, sd, sm);
    acc[k] = pass_sum(pi ? a.A_SS[k] : a.A_LL[k], sd, pi ? a.M_SS[k] : a.M_LL[k], sm);
    if (any1) {  // interface nodes: the solid-liquid bonds
      double sd1 = 0.0, sm1 = 0.0;
      st_sums(st, rcp, m1, any1, full1, ci[k], phii, sd1, sm1);
      if (mixed) acc[k] += pass_sum(a.A_SL[k], sd1, a.M_SL[k], sm1);
    }
  }
  if constexpr (JAC) {
    double sw[2] = {0.0, 0.0};
    double un[S];
    if (a.d0_out) {  // Jacobian set-up: weight sums and U^n for the diagonal
      sw[0] = st_wsum(st, m0);
      sw[1] = mixed ? st_wsum(st, m1) : 0.0;
#pragma unroll
      for (int k = 0; k < S; k++) un[k] = in ? a.Un[(int64_t)k * n + i] : 0.0;
    }
    if (in) jac_update<S>(a, i, pi, intf, ci, acc, sw, un);
  } else {
    if (in) explicit_update<S>(a, i, pi, intf, ci, acc, phii);
  }
}

template <int S, bool JAC>
hipError_t launch_st(const FluxArgs& a, const FluxSt& st, size_t lds, hipStream_t s) {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_flux_st<S, JAC>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  k_flux_st<S, JAC><<<nnpp_nblocks(a.n), NNPP_BLOCK, lds, s>>>(a, st);
  return hipGetLastError();
}

// ------------------------------------------------------- stencil form, mixed precision -------
// The explicit update of the stencil form with FP32 bond sums. The window of each species is
// staged as FP32 pairs (c^_j, phi^_j c^_j), where phi^ = phi - phi_r is the difference from the
// tile's first row and c^_j = c_j - c_r(p_j) the difference from the window's first node of
// the same phase p_j (both taken in FP64, then rounded). Per pass (same-phase bonds m0,
// solid-liquid bonds m1; c_r = c_r(p_i), c_r' = c_r(p_j) of the pass):
//   sd = sum w (c_j - c_i)          = D + (c_r' - c_r) W
//   sm = sum w (phi_j - phi_i) c_j  = (S2 - phi^_i (D + c^_i W)) + c_r' P
// with D = sum w (c^_j - c^_i) and S2 = sum w phi^_j c^_j in FP32 (one 8-byte LDS read, one
// subtraction and two FMAs per bond; FP32 partial sums per plane of the lattice ball or per 16
// table entries, accumulated in FP64), and W = sum w and P = sum w (phi_j - phi_i) in FP64 from
// the FP64 phi window, once per row. w is the flux weight rounded to FP32 in every sum. The
// diffusion sum adds differences, as the column form's does, so its rounding follows the
// field's local variation (a sum of staged values would round with the variation over the
// whole window, which over 1000 steps at 128^3 moved a front flip by a step); the jumps between
// the phases go into the FP64 terms. Reactions, the update, the Dirichlet rows and the clamp
// stay FP64.

// W += sum w, P += sum w (phi_j - phi_i) over the bonds of mask m (FP64, phi window in LDS).
__device__ __forceinline__ void st_wp(const FluxSt& st, const float* tf32, const double* rphi,
                                      const uint32_t (&m)[4], uint32_t anyc, double phii,
                                      double& W, double& P) {
  constexpr int BT = 16;
  const int K = st.K;
  for (int c = 0; c * BT < K; c++) {
    if (!((anyc >> c) & 1u)) continue;
    const int t0 = c * BT;
    const int wd = t0 >> 5;
    const uint32_t bits =
        (wd == 0 ? m[0] : (wd == 1 ? m[1] : (wd == 2 ? m[2] : m[3]))) >> (t0 & 31);
    int lv[BT];
    float wv[BT];
    double v[BT];
#pragma unroll
    for (int k = 0; k < BT; k++) {
      lv[k] = nnpp_ld_table(st.lc, t0 + k);
      wv[k] = nnpp_ld_table(tf32, t0 + k);
    }
#pragma unroll
    for (int k = 0; k < BT; k++) v[k] = rphi[lv[k]];
#pragma unroll
    for (int k = 0; k < BT; k++) {
      const double w = (double)nnpp_mask_weight<float>(wv[k], (int)(bits << (31 - k)) >> 31);
      W += w;
      P = __fma_rn(w, v[k] - phii, P);
    }
  }
}

// S1 += sum w (c^_j - chi), S2 += sum w q_j over the bonds of mask m from the staged pairs
// (c^_j, q_j): FP32 sums over each chunk of 16 entries, added in FP64.
__device__ __forceinline__ void st_sums32(const FluxSt& st, const float* tf32, const float2* rcq,
                                          const uint32_t (&m)[4], uint32_t anyc, uint32_t fullc,
                                          float chi, double& S1, double& S2) {
  constexpr int BT = 16;
  const int K = st.K;
  for (int c = 0; c * BT < K; c++) {
    if (!((anyc >> c) & 1u)) continue;
    const int t0 = c * BT;
    int lv[BT];
    float wv[BT];
    float2 v[BT];
#pragma unroll
    for (int k = 0; k < BT; k++) {
      lv[k] = nnpp_ld_table(st.lc, t0 + k);
      wv[k] = nnpp_ld_table(tf32, t0 + k);
    }
#pragma unroll
    for (int k = 0; k < BT; k++) v[k] = rcq[lv[k]];
    float a1 = 0.0f, a2 = 0.0f;
    if ((fullc >> c) & 1u) {
#pragma unroll
      for (int k = 0; k < BT; k++) {
        a1 = __builtin_fmaf(wv[k], v[k].x - chi, a1);
        a2 = __builtin_fmaf(wv[k], v[k].y, a2);
      }
    } else {
      const int wd = t0 >> 5;
      const uint32_t bits =
          (wd == 0 ? m[0] : (wd == 1 ? m[1] : (wd == 2 ? m[2] : m[3]))) >> (t0 & 31);
#pragma unroll
      for (int k = 0; k < BT; k++) {
        const float w = nnpp_mask_weight<float>(wv[k], (int)(bits << (31 - k)) >> 31);
        a1 = __builtin_fmaf(w, v[k].x - chi, a1);
        a2 = __builtin_fmaf(w, v[k].y, a2);
      }
    }
    S1 += (double)a1;
    S2 += (double)a2;
  }
}

// Radius-3 lattice ball with flux weights by class (the box lattice case): the entries walked
// line by line in ascending order, one LDS base per line and the entries at constant offsets
// from it, the weights by class dx^2 + dy^2 + dz^2 in scalar registers.
struct FluxBall {
  int32_t lc[NNPP_BALL_LINES];   // LDS position of each line's first entry, relative to the lane
  float w32[NNPP_BALL_CLASSES];  // FP32 flux weight by class
  double w64[NNPP_BALL_CLASSES];  // the same values in FP64
  double wsum;                   // sum of w32 over the ball (FP64)
};

// S1 += sum w (c^_j - chi), S2 += sum w q_j over the ball's entries (FULL) or those of mask m:
// FP32 partial sums per plane of the ball (at most 28 entries), added in FP64.
template <bool FULL>
__device__ __forceinline__ void ball_sums32(const FluxBall& fb, const float2* lds, int lane,
                                            const uint32_t (&mi)[4], float chi, double& S1,
                                            double& S2) {
  constexpr nnpp_ball<NNPP_BALL_M> B{};
  // opaque copies of the mask: the per-entry weights are formed here, not speculated into the
  // caller ahead of its branches and held in registers
  uint32_t m[4];
#pragma unroll
  for (int wd = 0; wd < 4; wd++) {
    m[wd] = mi[wd];
    if (!FULL) asm volatile("" : "+v"(m[wd]));
  }
  float a1 = 0.0f, a2 = 0.0f;
#pragma unroll
  for (int t = 0; t < B.K; t++) {
    const int l = B.line[t];
    const float2 v = lds[lane + fb.lc[l] + B.pos[t]];
    float w = fb.w32[B.d2[t]];
    if (!FULL) w = nnpp_mask_weight<float>(w, (int)(m[t >> 5] << (31 - (t & 31))) >> 31);
    a1 = __builtin_fmaf(w, v.x - chi, a1);
    a2 = __builtin_fmaf(w, v.y, a2);
    if (t == B.K - 1 || B.ldz[B.line[t < B.K - 1 ? t + 1 : t]] != B.ldz[l]) {
      S1 += (double)a1;
      S2 += (double)a2;
      a1 = a2 = 0.0f;
      // one plane at a time: the next plane's addresses and weights depend on values made
      // opaque after this plane's sums, so they are not formed ahead and held in registers
      if (FULL) asm volatile("" : "+v"(lane) : "v"(S1), "v"(S2));
      else
        asm volatile("" : "+v"(lane), "+v"(m[0]), "+v"(m[1]), "+v"(m[2]), "+v"(m[3])
                     : "v"(S1), "v"(S2));
    }
  }
}

// W += sum w, P += sum w (phi_j - phi_i) over the ball's entries (FULL: W += wsum) or those of
// mask m, in FP64 from the phi window.
template <bool FULL>
__device__ __forceinline__ void ball_wp(const FluxBall& fb, const double* lds, int lane,
                                        const uint32_t (&mi)[4], double phii, double& W, double& P) {
  constexpr nnpp_ball<NNPP_BALL_M> B{};
  uint32_t m[4];  // opaque copies, as in ball_sums32
#pragma unroll
  for (int wd = 0; wd < 4; wd++) {
    m[wd] = mi[wd];
    if (!FULL) asm volatile("" : "+v"(m[wd]));
  }
#pragma unroll
  for (int t = 0; t < B.K; t++) {
    const double v = lds[lane + fb.lc[B.line[t]] + B.pos[t]];
    double w = fb.w64[B.d2[t]];
    if (!FULL) {
      w = nnpp_mask_weight<double>(w, (int)(m[t >> 5] << (31 - (t & 31))) >> 31);
      W += w;
    }
    P = __fma_rn(w, v - phii, P);
    if (t < B.K - 1 && B.ldz[B.line[t + 1]] != B.ldz[B.line[t]]) {  // plane by plane
      if (FULL) asm volatile("" : "+v"(lane) : "v"(P));
      else
        asm volatile("" : "+v"(lane), "+v"(m[0]), "+v"(m[1]), "+v"(m[2]), "+v"(m[3])
                     : "v"(P), "v"(W));
    }
  }
  if (FULL) W += fb.wsum;
}

// Bits of the ball's entries whose node has phase pi (staged phases in LDS).
__device__ __forceinline__ void ball_same(const FluxBall& fb, const uint8_t* lds, int lane, int pi,
                                          uint32_t (&same)[4]) {
  constexpr nnpp_ball<NNPP_BALL_M> B{};
#pragma unroll
  for (int w = 0; w < 4; w++) same[w] = 0u;
#pragma unroll
  for (int t = 0; t < B.K; t++) {
    same[t >> 5] |= (uint32_t)((int)lds[lane + fb.lc[B.line[t]] + B.pos[t]] == pi) << (t & 31);
    if (t < B.K - 1 && B.ldz[B.line[t + 1]] != B.ldz[B.line[t]])
      asm volatile("" : "+v"(lane) : "v"(same[0]), "v"(same[1]), "v"(same[2]), "v"(same[3]));
  }
}

// BALL: the ball walks above (the stencil is the radius-3 lattice ball with flux weights by
// class); otherwise the table-driven chunks of 16 entries. The species loop is not unrolled (one
// copy of the bond walks); acc[k] is written by selects.
template <int S, bool BALL>
__global__ __launch_bounds__(NNPP_BLOCK, 3) void k_flux_st32(const FluxArgs a, const FluxSt st,
                                                             const float* __restrict__ tf32,
                                                             const FluxBall fb) {
  extern __shared__ float2 fst32_lds[];
  float2* lcq = fst32_lds;                                      // [E] (c^_j, phi^_j c^_j)
  double* lphi = reinterpret_cast<double*>(fst32_lds);          // [E] phi_j (before the species)
  uint8_t* lph = reinterpret_cast<uint8_t*>(fst32_lds + st.E);  // [E] phases
  const int64_t n = a.n;
  const int64_t r0 = nnpp_xcd_block(blockIdx.x, gridDim.x) * NNPP_BLOCK;
  const int64_t w0 = r0 + st.off_min;
  const int l = st.perm[r0 + threadIdx.x];
  const int64_t i = r0 + l;
  const bool in = i < n;
  const double* __restrict__ U = a.Un;
  const double* __restrict__ phi = a.Un + (int64_t)S * n;
  StSlots sl;
  sl.init(st);
  const double phir = phi[r0];  // the tile's phi reference: its first row (r0 < n)
  // the species references: per phase, the window's first node of that phase (row r0 where the
  // window has none), so the FP32 values of the same-phase sums stay away from the jumps
  // between the phases
  __shared__ uint32_t refpos[2];
  if (threadIdx.x < 2) refpos[threadIdx.x] = 0xffffffffu;
  __syncthreads();
  // the phi window (FP64) and, per staged position, phi^ (FP32) for the species pairs; the
  // phases, and whether the window holds one phase only (the phase of row r0)
  float php[ST_SLOTS];
  uint32_t sbits = 0u;  // staged positions (slots) of a solid node
  bool uni;
  {
    uint32_t add;
    const __amdgpu_buffer_rsrc_t rs = StSlots::rsrc(phi, n, w0, add);
    double v[ST_SLOTS];
#pragma unroll
    for (int k = 0; k < ST_SLOTS; k++)
      if (k * NNPP_BLOCK < st.E)
        v[k] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(
                                              rs, sl.off[k] * 8u + add, 0, 0));
    const __amdgpu_buffer_rsrc_t rp = StSlots::rsrc(a.phase, n, w0, add);
    uint8_t q[ST_SLOTS];
#pragma unroll
    for (int k = 0; k < ST_SLOTS; k++)
      if (k * NNPP_BLOCK < st.E) q[k] = __builtin_amdgcn_raw_buffer_load_b8(rp, sl.off[k] + add, 0, 0);
    const int p0 = a.phase[r0];
    bool u = true;
    uint32_t lo[2] = {0xffffffffu, 0xffffffffu};
#pragma unroll
    for (int k = 0; k < ST_SLOTS; k++) {
      const int e = (int)threadIdx.x + k * NNPP_BLOCK;
      if (k * NNPP_BLOCK < st.E) {
        php[k] = (float)(v[k] - phir);
        sbits |= (q[k] != 0 ? 1u : 0u) << k;
        if (e < st.E) {
          lphi[e] = v[k];
          lph[e] = q[k];
          u = u && (int)q[k] == p0;
          const int64_t j = w0 + (int64_t)sl.off[k];
          if (j >= 0 && j < n) {
            const int ph = q[k] != 0 ? 1 : 0;
            lo[ph] = sl.off[k] < lo[ph] ? sl.off[k] : lo[ph];
          }
        }
      }
    }
    if (lo[0] != 0xffffffffu) atomicMin(&refpos[0], lo[0]);
    if (lo[1] != 0xffffffffu) atomicMin(&refpos[1], lo[1]);
    uni = __syncthreads_and(u) != 0;
  }
  const int64_t jr0 = refpos[0] != 0xffffffffu ? w0 + (int64_t)refpos[0] : r0;
  const int64_t jr1 = refpos[1] != 0xffffffffu ? w0 + (int64_t)refpos[1] : r0;
  const double phii = in ? phi[i] : 0.0;
  const float phih = (float)(phii - phir);
  const int pi = in ? a.phase[i] : 0;
  const uint4 bq = in ? st.bmask[i] : make_uint4(0u, 0u, 0u, 0u);
  const uint32_t bw[4] = {bq.x, bq.y, bq.z, bq.w};
  uint32_t m0[4], m1[4];
  {
    const uint8_t* rph = lph + l;
    const int K = st.K;
    uint32_t same[4];
    if (uni) {  // every staged node has the phase of row r0
#pragma unroll
      for (int wd = 0; wd < 4; wd++) same[wd] = pi == a.phase[r0] ? ~0u : 0u;
    } else if (BALL) {
      ball_same(fb, lph, l, pi, same);
    } else {
#pragma unroll
      for (int wd = 0; wd < 4; wd++) {
        same[wd] = 0u;
        if (wd * 32 < K) {
#pragma unroll 8
          for (int t = 0; t < 32; t++) {
            const int tt = wd * 32 + t;
            if (tt < K) same[wd] |= (uint32_t)((int)rph[nnpp_ld_table(st.lc, tt)] == pi) << t;
          }
        }
      }
    }
#pragma unroll
    for (int wd = 0; wd < 4; wd++) {
      m0[wd] = bw[wd] & st.nz[wd] & same[wd];
      m1[wd] = bw[wd] & st.nz[wd] & ~same[wd];
    }
  }
  const bool mixed = (m1[0] | m1[1] | m1[2] | m1[3]) != 0u;
  const int intf = (pi == 1 && mixed) ? 1 : 0;  // utils.h:82-89
  uint32_t any0, full0, any1, full1;
  bool wfull0 = false;  // BALL: every lane of the wave has every same-phase entry
  if (BALL) {
    wfull0 = __all(m0[0] == st.nz[0] && m0[1] == st.nz[1] && m0[2] == st.nz[2] &&
                   m0[3] == st.nz[3]);
    any0 = __any((m0[0] | m0[1] | m0[2] | m0[3]) != 0u) ? ~0u : 0u;
    any1 = __any(mixed) ? ~0u : 0u;
    full0 = full1 = 0u;
  } else {
    st_chunks(st, m0, any0, full0);
    st_chunks(st, m1, any1, full1);
  }
  double W0 = 0.0, P0 = 0.0, W1 = 0.0, P1 = 0.0;
  if (BALL) {
    if (wfull0) ball_wp<true>(fb, lphi, l, m0, phii, W0, P0);
    else if (any0) ball_wp<false>(fb, lphi, l, m0, phii, W0, P0);
    if (any1) ball_wp<false>(fb, lphi, l, m1, phii, W1, P1);
  } else {
    st_wp(st, tf32, lphi + l, m0, any0, phii, W0, P0);
    if (any1) st_wp(st, tf32, lphi + l, m1, any1, phii, W1, P1);
  }
  double acc[S];
#pragma unroll
  for (int k = 0; k < S; k++) acc[k] = 0.0;
#pragma unroll 1
  for (int k = 0; k < S; k++) {
    __syncthreads();  // the previous reads of the buffer (phi window, species k - 1) are complete
    const double* __restrict__ Uk = U + (int64_t)k * n;
    const double cr0 = Uk[jr0], cr1 = Uk[jr1];
    {
      uint32_t add;
      const __amdgpu_buffer_rsrc_t rs = StSlots::rsrc(Uk, n, w0, add);
      constexpr int H = ST_SLOTS / 2;  // two rounds of loads (registers)
#pragma unroll
      for (int h = 0; h < ST_SLOTS; h += H) {
        double v[H];
#pragma unroll
        for (int s = 0; s < H; s++)
          if ((h + s) * NNPP_BLOCK < st.E)
            v[s] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(
                                                  rs, sl.off[h + s] * 8u + add, 0, 0));
#pragma unroll
        for (int s = 0; s < H; s++) {
          const int e = (int)threadIdx.x + (h + s) * NNPP_BLOCK;
          if ((h + s) * NNPP_BLOCK < st.E && e < st.E) {
            const float ch = (float)(v[s] - (((sbits >> (h + s)) & 1u) ? cr1 : cr0));
            lcq[e] = make_float2(ch, php[h + s] * ch);
          }
        }
      }
    }
    __syncthreads();
    const double cri = pi ? cr1 : cr0, cro = pi ? cr0 : cr1;  // own phase, other phase
    const float chf = (float)((in ? Uk[i] : 0.0) - cri);
    const double ch_i = (double)chf;
    double S1 = 0.0, S2 = 0.0;
    // the lane's row as an opaque value per species: the per-line LDS addresses are formed
    // inside the loop instead of being hoisted out of it and held in registers
    int lk = l;
    asm volatile("" : "+v"(lk));
    const float2* rcq = lcq + lk;
    if (BALL) {
      if (wfull0) ball_sums32<true>(fb, lcq, lk, m0, chf, S1, S2);
      else if (any0) ball_sums32<false>(fb, lcq, lk, m0, chf, S1, S2);
    } else {
      st_sums32(st, tf32, rcq, m0, any0, full0, chf, S1, S2);
    }
    double ak;
    {
      const double sd = S1;
      const double sm = (S2 - (double)phih * (S1 + ch_i * W0)) + cri * P0;
      ak = pass_sum(pi ? a.A_SS[k] : a.A_LL[k], sd, pi ? a.M_SS[k] : a.M_LL[k], sm);
    }
    if (any1) {  // interface nodes: the solid-liquid bonds
      double T1 = 0.0, T2 = 0.0;
      if (BALL) ball_sums32<false>(fb, lcq, lk, m1, chf, T1, T2);
      else st_sums32(st, tf32, rcq, m1, any1, full1, chf, T1, T2);
      const double sd = T1 + (cro - cri) * W1;
      const double sm = (T2 - (double)phih * (T1 + ch_i * W1)) + cro * P1;
      if (mixed) ak += pass_sum(a.A_SL[k], sd, a.M_SL[k], sm);
    }
#pragma unroll
    for (int j = 0; j < S; j++)
      if (j == k) acc[j] = ak;
  }
  if (in) {
    double ci[S];
#pragma unroll
    for (int k = 0; k < S; k++) ci[k] = U[(int64_t)k * n + i];
    explicit_update<S>(a, i, pi, intf, ci, acc, phii);
  }
}

template <int S>
hipError_t launch_st32(const FluxArgs& a, const FluxSt& st, const float* tf32, const FluxBall* fb,
                       size_t lds, hipStream_t s) {
  const void* f = fb ? reinterpret_cast<const void*>(&k_flux_st32<S, true>)
                     : reinterpret_cast<const void*>(&k_flux_st32<S, false>);
  hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  if (fb) k_flux_st32<S, true><<<nnpp_nblocks(a.n), NNPP_BLOCK, lds, s>>>(a, st, tf32, *fb);
  else k_flux_st32<S, false><<<nnpp_nblocks(a.n), NNPP_BLOCK, lds, s>>>(a, st, tf32, FluxBall{});
  return hipGetLastError();
}

// ------------------------------------------------- z-marching form (full box lattices) -------
// The explicit update on a full box lattice with the radius-3 ball of offsets and flux weights by
// class dx^2 + dy^2 + dz^2 (FP64 sums). Blocks take z-marching work items (a 32 x 8 column tile
// and a chunk of lz output planes, ZmGeom; lz chosen by the launch) and the species one at a
// time. The phases of every input plane of the chunk (2 outside the box) are staged into LDS
// once per block; per species the block walks its input planes, staging each plane's tile with a
// 3-node halo as pairs (c_j, q_j = phi^n_j c_j) into LDS, and each row of the thread's column in
// flight (up to 7) takes the plane's bonds,
//   acc += A(p_o, p_j) sum w (c_j - c_o) + M(p_o, p_j) sum w (q_j - phi_o c_j),
// with the phase pair's coefficients (utils.h:100-159: D by phase pair, BV on solid-liquid
// bonds); A and M are applied per plane. A row is updated for the species when it completes
// (U_k^{n+1}, Dirichlet row, clamp: explicit_update for one field).
//
// Per plane, every active slot takes all its bonds with the coefficients of the row's own phase
// from plain ring sums of c and q (positions outside the box are staged as c = q = 0 and add
// nothing; the row's weight sum counts the bonds inside the box only). Where the wave's disc
// region holds both phases, or rows of the slot differ in phase, the slot then takes the
// difference of the solid-liquid coefficients over its bonds to nodes of the other phase, bond by
// bond (this also finds the interface rows).

// output planes per block at most (the phases of every input plane stay in LDS)
constexpr int ZM_FLZ = 16;

struct ZmFluxW {
  double w[NNPP_BALL_CLASSES];  // FP64 flux weight by class dx^2 + dy^2 + dz^2
  double wz[NNPP_BALL_M + 1];   // weight sum of the disc at plane distance |dz| (the center
                                // included for dz != 0)
};

template <int S>
__device__ __forceinline__ void zm_update_one(const FluxArgs& a, int k, uint32_t i, int pi,
                                              int intf, double co, double acc) {
  const int64_t n = a.n;
  double Rk = 0.0;
  if constexpr (S == 10) {
    if (a.chemistry == NNPP_CHEM_REFERENCE) {
      double ci[S];
#pragma unroll
      for (int f = 0; f < S; f++) ci[f] = a.Un[(int64_t)f * n + i];
      Rk = ref_reaction(k, a.rx, pi, intf, ci, ci, a.Un[(int64_t)S * n + i]);
    }
  }
  if (a.chemistry == NNPP_CHEM_GENERIC && k == 0 && intf) Rk = -a.r_diss * co;
  // explicit_update's operations for field k
  const double flux = -acc;
  double v = co - a.dt * (flux - Rk);
  const unsigned bm = a.bcmask[i];
  bool keep = !((bm >> k) & 1u);
  if (!keep) v = a.bcval[(int64_t)k * n + i];
  const bool cl = ((a.clamp_always >> k) & 1u) || (((a.clamp_liquid >> k) & 1u) && pi == 0);
  if (cl && v < a.clamp_floor) {
    v = a.clamp_floor;
    keep = false;
  }
  a.out[(int64_t)k * n + i] = v;
  if (a.passmask) {
    const uint16_t bit = keep ? (uint16_t)(1u << k) : (uint16_t)0;
    a.passmask[i] = k == 0 ? bit : (uint16_t)(a.passmask[i] | bit);
  }
}

// The Jacobian apply of the implicit step (frozen Jacobian, lagged phi^n; as the stencil and
// column kernels' JAC mode, with d0 given): out_k = v_k on Dirichlet rows, else d0_k v_k - acc_k,
// the bond sums of v in place of U^n.
__device__ __forceinline__ void zm_jac_one(const FluxArgs& a, int k, uint32_t i, double v,
                                           double acc) {
  const int64_t j = (int64_t)k * a.n + i;
  const bool dir = (a.bcmask[i] >> k) & 1u;
  a.out[j] = dir ? v : __fma_rn(a.d0[j], v, -acc);
}

// JAC: the Jacobian apply to v = a.U1 (species-major) instead of the explicit update.
// (held at 3 waves per SIMD: the register demand of the rows in flight, FP64)
template <int S, bool JAC>
__global__ __launch_bounds__(NNPP_BLOCK) __attribute__((amdgpu_waves_per_eu(3, 3))) void
k_zm_flux(const FluxArgs a, const ZmGeom g, const ZmFluxW fw) {
  __shared__ double2 lcp[2][ZM_PTS];                  // (c_j, phi_j c_j); 0 outside the box
  __shared__ uint8_t lcl[ZM_FLZ + 2 * ZM_M][ZM_PTS];  // phases of the input planes, 2 outside
  const int64_t n = a.n;
  const int64_t item = nnpp_xcd_block(blockIdx.x, g.items);
  const int zc = (int)(item % g.nzc);
  const int64_t tile = item / g.nzc;
  const int x0 = (int)(tile % g.tiles_x) * ZM_TX, y0 = (int)(tile / g.tiles_x) * ZM_TY;
  const int z0 = zc * g.lz;
  const int zend = min(z0 + g.lz, g.nz);
  const int tx = threadIdx.x % ZM_TX, ty = threadIdx.x / ZM_TX;
  const int x = x0 + tx, y = y0 + ty;
  const bool col_ok = x < g.nx && y < g.ny;
  const int wv = threadIdx.x >> 6;  // the wave's rows: ty = 2 wv, 2 wv + 1
  int32_t soff[ZM_STAGE];
#pragma unroll
  for (int k = 0; k < ZM_STAGE; k++) {
    const int e = threadIdx.x + k * NNPP_BLOCK;
    const int gx = x0 + e % ZM_W - ZM_M, gy = y0 + e / ZM_W - ZM_M;
    soff[k] = (e < ZM_PTS && gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny) ? gy * g.nx + gx : -1;
  }
  const int zs = max(z0 - ZM_M, 0), ze = min(zend + ZM_M, g.nz);  // input planes [zs, ze)
  const uint32_t col = col_ok ? (uint32_t)(y * g.nx + x) : 0u;
  const uint32_t sz = (uint32_t)g.sz;
  const double* __restrict__ phi = a.Un + (int64_t)S * n;
  const auto in_chunk = [&](int zo) { return zo >= z0 && zo < zend; };
  // weight of the row's bonds to in-plane positions inside the box, by plane distance |dz|
  // (in LDS: the registers go to the rows in flight)
  __shared__ double lwin[ZM_M + 1][NNPP_BLOCK];
#pragma unroll
  for (int az = 0; az <= ZM_M; az++) {
    double s = 0.0;
#pragma unroll
    for (int dy = -ZM_M; dy <= ZM_M; dy++)
#pragma unroll
      for (int dx = -ZM_M; dx <= ZM_M; dx++) {
        const int d2 = dx * dx + dy * dy + az * az;
        if (d2 > ZM_M * ZM_M || d2 == 0) continue;
        const bool out = x + dx < 0 || x + dx >= g.nx || y + dy < 0 || y + dy >= g.ny;
        if (out) s += fw.w[d2];
      }
    lwin[az][threadIdx.x] = fw.wz[az] - s;
  }
  // the phases of every input plane, staged once for all species, and per plane whether the
  // wave's disc region (staged lines 2 wv .. 2 wv + 7, 304 points) holds liquid / solid nodes
  // (bit zi - zs of vl / vs; five class bytes per lane, then a vote)
  for (int zi = zs; zi < ze; zi++) {
    const uint32_t base = (uint32_t)zi * sz;
#pragma unroll
    for (int e = 0; e < ZM_STAGE; e++) {
      const int pt = threadIdx.x + e * NNPP_BLOCK;
      if (pt < ZM_PTS) lcl[zi - zs][pt] = soff[e] >= 0 ? a.phase[base + (uint32_t)soff[e]] : 2;
    }
  }
  __syncthreads();
  uint32_t vl = 0u, vs = 0u;
  for (int zi = zs; zi < ze; zi++) {
    bool hl = false, hs = false;
    const uint8_t* __restrict__ rg = &lcl[zi - zs][2 * wv * ZM_W];
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int r = 0; r < (8 * ZM_W + 63) / 64; r++) {
      const int pt = lane + 64 * r;
      if (pt < 8 * ZM_W) {
        const uint32_t cq = rg[pt];
        hl = hl || cq == 0u;
        hs = hs || cq == 1u;
      }
    }
    vl |= (__any(hl) ? 1u : 0u) << (zi - zs);
    vs |= (__any(hs) ? 1u : 0u) << (zi - zs);
  }
  vl = __builtin_amdgcn_readfirstlane(vl);  // (uniform over the wave)
  vs = __builtin_amdgcn_readfirstlane(vs);
  const int ctr = (ty + ZM_M) * ZM_W + tx + ZM_M;  // the thread's column in a staged plane

  int buf = 0;
#pragma unroll 1
  for (int k = 0; k < S; k++) {
    const double* __restrict__ Uk = (JAC ? a.U1 : a.Un) + (int64_t)k * n;
    const double A_same[2] = {a.A_LL[k], a.A_SS[k]}, M_same[2] = {a.M_LL[k], a.M_SS[k]};
    const double A_x = a.A_SL[k], M_x = a.M_SL[k];
    // rows in flight: slot q holds output plane zi - M + q while input plane zi is processed
    double acc[ZM_SL], co[ZM_SL], pho[ZM_SL];
    uint32_t pos = 0u, cross = 0u;  // bit q: the row of slot q is solid / has a solid-liquid bond
    const auto row_load = [&](int zo, double& c, double& p, uint32_t& solid) {
      if (col_ok && zo >= max(z0, 0) && zo < zend) {
        const uint32_t i = (uint32_t)zo * sz + col;
        c = Uk[i];
        p = phi[i];
        solid = lcl[zo - zs][ctr] != 0 ? 1u : 0u;
      } else {
        c = p = 0.0;
        solid = 0u;
      }
    };
#pragma unroll
    for (int q = 0; q < ZM_SL; q++) {
      acc[q] = 0.0;
      uint32_t sb;
      row_load(zs - ZM_M + q, co[q], pho[q], sb);
      pos |= sb << q;
    }
    // the row that enters the last slot at the next plane, loaded a plane ahead
    double nc, np;
    uint32_t nsb;
    row_load(zs + 1 + ZM_M, nc, np, nsb);
    // staged loads of the next plane
    double sc[ZM_STAGE], sp[ZM_STAGE];
    const auto fetch = [&](int zi) {
      const uint32_t base = (uint32_t)zi * sz;
#pragma unroll
      for (int e = 0; e < ZM_STAGE; e++) {
        const bool ok = soff[e] >= 0;
        const uint32_t j = base + (uint32_t)(ok ? soff[e] : 0);
        sc[e] = ok ? Uk[j] : 0.0;
        sp[e] = ok ? phi[j] : 0.0;
      }
    };
    fetch(zs);
    const auto finish = [&](int zi) {
      const int zo = zi - ZM_M;
      if (col_ok && in_chunk(zo)) {
        if constexpr (JAC) {
          zm_jac_one(a, k, (uint32_t)zo * sz + col, co[0], acc[0]);
        } else {
          const int pi = (int)(pos & 1u);
          const int intf = (pi == 1 && (cross & 1u)) ? 1 : 0;  // utils.h:82-89
          zm_update_one<S>(a, k, (uint32_t)zo * sz + col, pi, intf, co[0], acc[0]);
        }
      }
#pragma unroll
      for (int q = 0; q + 1 < ZM_SL; q++) {
        acc[q] = acc[q + 1];
        co[q] = co[q + 1];
        pho[q] = pho[q + 1];
      }
      pos >>= 1;
      cross >>= 1;
      acc[ZM_SL - 1] = 0.0;
      co[ZM_SL - 1] = nc;
      pho[ZM_SL - 1] = np;
      pos |= nsb << (ZM_SL - 1);
      row_load(zi + 2 + ZM_M, nc, np, nsb);
    };
#pragma unroll 1
    for (int zi = zs; zi < ze; zi++) {
#pragma unroll
      for (int e = 0; e < ZM_STAGE; e++) {
        const int pt = threadIdx.x + e * NNPP_BLOCK;
        if (pt < ZM_PTS) lcp[buf][pt] = make_double2(sc[e], sp[e] * sc[e]);
      }
      __syncthreads();
      if (zi + 1 < ze) fetch(zi + 1);
      const bool hasL = (vl >> (zi - zs)) & 1u, hasS = (vs >> (zi - zs)) & 1u;
      const uint32_t cstar = hasS ? 1u : 0u;
      const bool single = !(hasL && hasS);
      // the thread's first disc line (dy = -M) in the staged plane
      const int lb0 = ty * ZM_W + tx + ZM_M;
      // per slot: active (row in the chunk), fast (one phase: every bond, no tests)
      // (bit q of actm / fastm; scalar masks, uniform over the wave)
      uint32_t actm = 0u, fastm = 0u;
#pragma unroll
      for (int q = 0; q < ZM_SL; q++) {
        const bool act = in_chunk(zi - ZM_M + q);
        const uint32_t sb = (pos >> q) & 1u;
        actm |= (act ? 1u : 0u) << q;
        fastm |= (act && single && __all(!col_ok || sb == cstar) ? 1u : 0u) << q;
      }
      actm = __builtin_amdgcn_readfirstlane(actm);
      fastm = __builtin_amdgcn_readfirstlane(fastm);
      const auto act = [&](int q) { return ((actm >> q) & 1u) != 0u; };
      const auto fast = [&](int q) { return ((fastm >> q) & 1u) != 0u; };
      {
        // every active slot, all bonds with the coefficients of the row's own phase: the disc
        // by rings of equal k = dx^2 + dy^2 around the column, SC_k = sum c_j, SQ_k = sum q_j
        // (q = phi c), and a slot at plane distance dz, with Tc = sum_k w SC_k + w(dz^2) c_c and
        // Tq alike (the column's own point for dz != 0), takes
        //   sum w (c_j - c_o)          = Tc - c_o (W_dz - w_out)
        //   sum w (phi_j - phi_o) c_j  = Tq - phi_o Tc
        // (positions outside the box are staged as 0 and add nothing to Tc and Tq; W_dz - w_out
        // is the weight of the row's bonds inside the box in that plane)
        if (actm != 0u) {
          const double2 cen = lcp[buf][lb0 + ZM_M * ZM_W];
          double SC[ZM_M * ZM_M + 1], SQ[ZM_M * ZM_M + 1];
#pragma unroll
          for (int dy = -ZM_M; dy <= ZM_M; dy++) {
            // one disc line at a time: the scheduler does not batch the whole disc's reads
            // ahead of the sums and hold them in registers
            __builtin_amdgcn_sched_barrier(0);
            const double2* __restrict__ line = &lcp[buf][lb0 + (dy + ZM_M) * ZM_W];
#pragma unroll
            for (int dx = -ZM_M; dx <= ZM_M; dx++) {
              const int kk = dx * dx + dy * dy;
              if (kk == 0 || kk > ZM_M * ZM_M) continue;
              const double2 v = line[dx];
              const bool first = zm_ring_first<ZM_M>(dx, dy);
              SC[kk] = first ? v.x : SC[kk] + v.x;
              SQ[kk] = first ? v.y : SQ[kk] + v.y;
            }
          }
          __builtin_amdgcn_sched_barrier(0);
          // the sums of one plane distance |dz|, shared by the slots dz = +-|dz|
#pragma unroll
          for (int az = 0; az <= ZM_M; az++) {
            const int qa = ZM_M - az, qb = ZM_M + az;  // the slots at dz = az, -az
            if (!act(qa) && !act(qb)) continue;        // uniform over the wave
            double G = 0.0, H = 0.0;
            bool first = true;
#pragma unroll
            for (int kk = 1; kk <= ZM_M * ZM_M; kk++) {
              if (!zm_ring_exists<ZM_M>(kk) || kk + az * az > ZM_M * ZM_M) continue;
              const double w = fw.w[kk + az * az];
              G = first ? w * SC[kk] : __fma_rn(w, SC[kk], G);
              H = first ? w * SQ[kk] : __fma_rn(w, SQ[kk], H);
              first = false;
            }
            if (az != 0) {  // the column's own point
              G = __fma_rn(fw.w[az * az], cen.x, G);
              H = __fma_rn(fw.w[az * az], cen.y, H);
            }
            const double win = lwin[az][threadIdx.x];
#pragma unroll
            for (int h = 0; h < (az ? 2 : 1); h++) {
              const int q = h ? qb : qa;
              if (!act(q)) continue;  // uniform over the wave
              const bool po = (pos >> q) & 1u;
              const double sd = __fma_rn(-co[q], win, G);
              const double sm = __fma_rn(-pho[q], G, H);
              acc[q] = __fma_rn(po ? A_same[1] : A_same[0], sd, acc[q]);
              acc[q] = __fma_rn(po ? M_same[1] : M_same[0], sm, acc[q]);
            }
          }
        }
      }
      // the slots whose wave region holds both phases: the bonds to nodes of the other phase
      // take the solid-liquid coefficients instead, a correction summed bond by bond
#pragma unroll
      for (int q = 0; q < ZM_SL; q++) {
        if (!act(q) || fast(q)) continue;  // uniform over the wave
        const int dz = ZM_M - q;
        const uint32_t po = (pos >> q) & 1u;
        double dxs = 0.0, mxs = 0.0;  // the bonds to the other phase
        bool anyx = false;
        int lb = lb0;
        __builtin_amdgcn_sched_barrier(0);  // one slot at a time
#pragma unroll
        for (int dy = -ZM_M; dy <= ZM_M; dy++) {
          if (dz * dz + dy * dy > ZM_M * ZM_M) {
            lb += ZM_W;
            continue;
          }
          __builtin_amdgcn_sched_barrier(0);  // one disc line at a time
          const double2* __restrict__ line = &lcp[buf][lb];
          const uint8_t* __restrict__ linec = &lcl[zi - zs][lb];
          lb += ZM_W;
#pragma unroll
          for (int dx = -ZM_M; dx <= ZM_M; dx++) {
            const int d2 = dx * dx + dy * dy + dz * dz;
            if (d2 > ZM_M * ZM_M || d2 == 0) continue;
            const double2 v = line[dx];
            const uint32_t cj = linec[dx];
            const bool x = cj == (po ^ 1u);
            const double wx = x ? fw.w[d2] : 0.0;
            anyx = anyx || x;
            const double dc = v.x - co[q], dq = __fma_rn(-pho[q], v.x, v.y);  // (phi_j - phi_o) c_j
            dxs = __fma_rn(wx, dc, dxs);
            mxs = __fma_rn(wx, dq, mxs);
          }
        }
        acc[q] = __fma_rn(A_x - A_same[po], dxs, acc[q]);
        acc[q] = __fma_rn(M_x - M_same[po], mxs, acc[q]);
        cross |= (anyx ? 1u : 0u) << q;
      }
      finish(zi);
      buf ^= 1;
    }
    // planes past the top of the box add nothing: the rows still in flight are complete
#pragma unroll 1
    for (int zi = ze; zi < zend + ZM_M; zi++) finish(zi);
  }
}

// Work items: lz chosen so that every block is resident at once (one round: a second, partly
// filled round would run at a fraction of the occupancy), 4 <= lz <= ZM_FLZ.
template <int S, bool JAC>
hipError_t launch_zm(const FluxArgs& a, ZmGeom g, const ZmFluxW& fw, hipStream_t s,
                     int device) {
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_zm_flux<S, JAC>, NNPP_BLOCK, 0) ==
          hipSuccess &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) ==
          hipSuccess &&
      per_cu > 0 && cus > 0) {
    const int64_t tiles = (int64_t)g.tiles_x * g.tiles_y, cap = (int64_t)per_cu * cus;
    int64_t lz = (g.nz * tiles + cap - 1) / cap;
    lz = std::min<int64_t>(std::max<int64_t>(lz, 4), std::min(ZM_FLZ, g.nz));
    g.lz = (int)lz;
    g.nzc = (int)((g.nz + lz - 1) / lz);
    g.items = tiles * g.nzc;
  }
  k_zm_flux<S, JAC><<<g.items, NNPP_BLOCK, 0, s>>>(a, g, fw);
  return hipGetLastError();
}

}  // namespace

static FluxArgs flux_args(nnpp_ctx* c, double BV, const double* U1, const double* Un,
                          double* out, uint16_t* passmask) {
  const nnpp_params& p = c->p;
  const int S = p.n_species;
  FluxArgs a;
  a.d0 = nullptr;
  a.d0_out = nullptr;
  a.diag_out = nullptr;
  a.n = c->n;
  a.chemistry = p.chemistry;
  a.off = c->bonds.off;
  a.len = c->bonds.len;
  a.col = c->bonds.col;
  a.d16 = c->bonds.d16;
  a.first = c->bonds.first;
  a.w64 = c->bonds.w64;
  a.w32 = c->bonds.w32;
  a.rec = c->aos;
  a.rec_out = c->aos;
  a.phase = c->phase;
  a.bcmask = c->bcmask;
  a.bcval = c->bcval;
  a.U1 = U1;
  a.Un = Un;
  a.out = out;
  a.passmask = passmask;
  const double FRT = p.F / (p.R * p.T);
  a.clamp_always = a.clamp_liquid = 0;
  for (int k = 0; k < NNPP_MAX_SPECIES; k++) {
    a.A_LL[k] = a.A_SL[k] = a.A_SS[k] = a.M_LL[k] = a.M_SL[k] = a.M_SS[k] = a.z[k] = 0.0;
    if (k >= S) continue;
    const double bvSL = p.solid_bonds[k] ? BV : 1.0;
    a.A_LL[k] = p.K1[k] * p.D_liquid[k];
    a.A_SL[k] = p.K1[k] * p.D_solid[k] * bvSL;
    a.A_SS[k] = p.solid_bonds[k] ? p.K1[k] * p.D_solid[k] : 0.0;  // BV = 1 (utils.h:113-116)
    a.M_LL[k] = p.K2[k] * p.D_liquid[k] * p.z[k] * FRT;
    a.M_SL[k] = p.K2[k] * p.D_solid[k] * bvSL * p.z[k] * FRT;
    a.M_SS[k] = p.solid_bonds[k] ? p.K2[k] * p.D_solid[k] * p.z[k] * FRT : 0.0;
    a.z[k] = p.z[k];
    if (p.clamp[k] == NNPP_CLAMP_ALWAYS) a.clamp_always |= 1u << k;
    if (p.clamp[k] == NNPP_CLAMP_LIQUID) a.clamp_liquid |= 1u << k;
  }
  a.dt = p.dt;
  a.phi_q = (1. / (p.K_phi * p.sigma1)) * p.F * p.CSolid;
  a.clamp_floor = p.clamp_floor;
  a.r_diss = p.r_diss;
  RxConst& r = a.rx;
  const double CS = p.CSolid;
  r.KMgOH2 = pow(10., -p.pKMgOH2) / CS;
  r.KMgCO3 = pow(10., -p.pKMgCO3) / CS;
  r.KMgHCO3 = pow(10., -p.pKMgHCO3) / CS;
  r.KMgPO4 = pow(10., -p.pKMgPO4) / CS;
  r.KCaCO3 = pow(10., -p.pKCaCO3) / CS;
  r.KCa5 = pow(10., -p.pKCa5OHPO4) / CS;
  r.KH2O = pow(10., -p.pKH2O) / CS;
  r.KHCO3 = pow(10., -p.pKHCO3) / CS;
  r.KHPO4 = pow(10., -p.pKHPO4) / CS;
  r.k1b = p.k1b;
  r.k2b = p.k2b;
  r.k3b = p.k3b;
  r.k4b = p.k4b;
  r.k5b = p.k5b;
  r.k6b = p.k6b;
  r.k7b = p.k7b;
  r.k8b = p.k8b;
  r.k9b = p.k9b;
  r.JHp_c = S > 1 ? p.JHp / (p.z[1] * p.F * p.CSolid) : 0.0;
  r.RT_Fz1 = p.R * p.T / (p.F * p.z[0]);
  r.sat_rel = p.CSat / p.CSolid;
  r.phim_se = p.phim + p.phimse;
  r.alphaC = p.alphaC;
  r.F = p.F;
  r.R = p.R;
  r.T = p.T;
  return a;
}

// LDS bytes of the stencil flux (FP64 or, with fp32, the mixed form) for the current
// geometry, or 0 where it does not apply.
static size_t flux_stencil_lds(nnpp_ctx* c, bool fp32 = false) {
  if (c->format & NNPP_FORMAT_NO_STENCIL) return 0;
  if (nnpp_stencil_ensure(c) != NNPP_OK) return 0;
  const nnpp_stencil& T = c->pst;
  if (T.state != 1 || !T.flux_ok || T.lds_elems > ST_SLOTS * NNPP_BLOCK) return 0;
  const size_t bytes = ((fp32 ? sizeof(float2) : 2 * sizeof(double)) + 1) * (size_t)T.lds_elems;
  int max_lds = 0;
  if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) !=
          hipSuccess ||
      bytes + 1024 > (size_t)max_lds)
    return 0;
  return (bytes + 7) & ~(size_t)7;
}

// The stencil form of the explicit update (JAC = false; FP32 bond sums with fp32) or of the
// Jacobian apply (JAC = true).
template <bool JAC>
static int launch_stencil(nnpp_ctx* c, const FluxArgs& a, size_t st_lds, bool fp32 = false) {
  const int S = c->p.n_species;
  const nnpp_stencil& T = c->pst;
  FluxSt st;
  st.bmask = T.bmask;
  st.perm = T.fperm;
  st.lc = T.lc;
  st.tf = T.tf64;
  st.seg_start = T.seg_start;
  st.seg_len = T.seg_len;
  st.seg_base = T.seg_base;
  st.K = T.K;
  st.R = T.R;
  st.E = T.lds_elems;
  st.off_min = T.off_min;
  for (int w = 0; w < 4; w++) st.nz[w] = T.fnz[w];
  FluxBall fball;
  const FluxBall* fbp = nullptr;
  if (T.fball && !(c->format & NNPP_FORMAT_NO_BALL)) {  // the ball walks of the mixed form
    for (int l = 0; l < NNPP_BALL_LINES; l++) fball.lc[l] = T.ball_args.lc[l];
    for (int d = 0; d < NNPP_BALL_CLASSES; d++) {
      fball.w32[d] = T.fball_w32[d];
      fball.w64[d] = (double)T.fball_w32[d];
    }
    fball.wsum = T.fball_wsum;
    fbp = &fball;
  }
  hipError_t e = hipSuccess;
  switch (S) {
#define NNPP_CASE(SS)                                                                  \
  case SS:                                                                             \
    e = (!JAC && fp32) ? launch_st32<SS>(a, st, T.tf32, fbp, st_lds, c->stream)        \
                       : launch_st<SS, JAC>(a, st, st_lds, c->stream);                 \
    break;
    NNPP_CASE(1) NNPP_CASE(2) NNPP_CASE(3) NNPP_CASE(4) NNPP_CASE(5) NNPP_CASE(6)
    NNPP_CASE(7) NNPP_CASE(8) NNPP_CASE(9) NNPP_CASE(10) NNPP_CASE(11) NNPP_CASE(12)
    NNPP_CASE(13) NNPP_CASE(14) NNPP_CASE(15) NNPP_CASE(16)
#undef NNPP_CASE
    default:
      return nnpp_set_error(c, NNPP_ERR_ARG, "n_species out of range");
  }
  if (e != hipSuccess) return nnpp_set_error(c, NNPP_ERR_HIP, hipGetErrorString(e));
  return NNPP_OK;
}

int nnpp_flux_stencil_active(nnpp_ctx* c) { return flux_stencil_lds(c) != 0; }

// The z-marching form of the explicit update (FP64 sums): a full box lattice with the radius-3
// ball of offsets and flux weights by class, unless a format flag selects the tile forms.
int nnpp_flux_zmarch_active(nnpp_ctx* c) {
  if (c->format & (NNPP_FORMAT_NO_STENCIL | NNPP_FORMAT_NO_BALL | NNPP_FORMAT_NO_ZMARCH)) return 0;
  if (flux_stencil_lds(c) == 0) return 0;
  const nnpp_stencil& T = c->pst;
  return (T.box && T.fball && T.ball == NNPP_BALL_M) ? 1 : 0;
}

static int launch_zmarch(nnpp_ctx* c, const FluxArgs& a, bool jac = false) {
  const int S = c->p.n_species;
  ZmGeom g = nnpp_zm_geom(c->pst, 1024);
  if (g.lz > ZM_FLZ) {  // (no reduction partials here: the grid may grow)
    g.lz = ZM_FLZ;
    g.nzc = (g.nz + ZM_FLZ - 1) / ZM_FLZ;
    g.items = (int64_t)g.tiles_x * g.tiles_y * g.nzc;
  }
  ZmFluxW fw;
  for (int d = 0; d < NNPP_BALL_CLASSES; d++) fw.w[d] = c->pst.fball_w64[d];
  constexpr int M = NNPP_BALL_M;
  for (int az = 0; az <= M; az++) {
    double sw = 0.0;
    for (int dy = -M; dy <= M; dy++)
      for (int dx = -M; dx <= M; dx++) {
        const int d2 = dx * dx + dy * dy + az * az;
        if (d2 <= M * M && d2 != 0) sw += fw.w[d2];
      }
    fw.wz[az] = sw;
  }
  hipError_t e = hipSuccess;
  switch (S) {
#define NNPP_CASE(SS) \
  case SS:                                                                   \
    e = jac ? launch_zm<SS, true>(a, g, fw, c->stream, c->device)            \
            : launch_zm<SS, false>(a, g, fw, c->stream, c->device);          \
    break;
    NNPP_CASE(1) NNPP_CASE(2) NNPP_CASE(3) NNPP_CASE(4) NNPP_CASE(5) NNPP_CASE(6)
    NNPP_CASE(7) NNPP_CASE(8) NNPP_CASE(9) NNPP_CASE(10) NNPP_CASE(11) NNPP_CASE(12)
    NNPP_CASE(13) NNPP_CASE(14) NNPP_CASE(15) NNPP_CASE(16)
#undef NNPP_CASE
    default:
      return nnpp_set_error(c, NNPP_ERR_ARG, "n_species out of range");
  }
  if (e != hipSuccess) return nnpp_set_error(c, NNPP_ERR_HIP, hipGetErrorString(e));
  return NNPP_OK;
}

int nnpp_launch_flux(nnpp_ctx* c, int resid, int precision, double BV, const double* U1,
                     const double* Un, double* out, uint16_t* passmask) {
  const int S = c->p.n_species;
  const FluxArgs a = flux_args(c, BV, U1, Un, out, passmask);
  if (!c->n) return NNPP_OK;
  const int64_t nb = nnpp_nblocks(c->n);
  const int mixed = (precision == NNPP_MIXED);
  hipError_t e = hipSuccess;
  // explicit update on a full box lattice in FP64: the z-marching form
  if (!resid && !mixed && nnpp_flux_zmarch_active(c)) return launch_zmarch(c, a);
  // explicit update on a lattice: the stencil form, with FP32 bond sums in mixed precision
  const size_t st_lds = resid ? 0 : flux_stencil_lds(c, mixed);
  if (st_lds) return launch_stencil<false>(c, a, st_lds, mixed);
  switch (S) {
#define NNPP_CASE(SS) \
  case SS: e = launch_s<SS>(a, resid, mixed, c->bonds.wide, nb, c->stream); break;
    NNPP_CASE(1) NNPP_CASE(2) NNPP_CASE(3) NNPP_CASE(4) NNPP_CASE(5) NNPP_CASE(6)
    NNPP_CASE(7) NNPP_CASE(8) NNPP_CASE(9) NNPP_CASE(10) NNPP_CASE(11) NNPP_CASE(12)
    NNPP_CASE(13) NNPP_CASE(14) NNPP_CASE(15) NNPP_CASE(16)
#undef NNPP_CASE
    default:
      return nnpp_set_error(c, NNPP_ERR_ARG, "n_species out of range");
  }
  if (e != hipSuccess) return nnpp_set_error(c, NNPP_ERR_HIP, hipGetErrorString(e));
  return NNPP_OK;
}

int nnpp_launch_jac(nnpp_ctx* c, double BV, const double* Un, const double* vin,
                    const double* d0, double* out, double* d0_out, double* diag_out) {
  const int S = c->p.n_species;
  FluxArgs a = flux_args(c, BV, vin, Un, out, nullptr);
  a.d0 = d0;
  a.d0_out = d0_out;
  a.diag_out = diag_out;
  if (!c->n) return NNPP_OK;
  // full box lattice, d0 given: the z-marching form; otherwise on lattices the stencil form
  // (both read v species-major)
  if (vin && d0 && !d0_out && nnpp_flux_zmarch_active(c)) return launch_zmarch(c, a, true);
  const size_t st_lds = vin ? flux_stencil_lds(c) : 0;
  if (st_lds) return launch_stencil<true>(c, a, st_lds);
  const int64_t nb = nnpp_nblocks(c->n);
  hipError_t e = hipSuccess;
  switch (S) {
#define NNPP_CASE(SS) \
  case SS: e = launch_jac<SS>(a, c->bonds.wide, nb, c->stream); break;
    NNPP_CASE(1) NNPP_CASE(2) NNPP_CASE(3) NNPP_CASE(4) NNPP_CASE(5) NNPP_CASE(6)
    NNPP_CASE(7) NNPP_CASE(8) NNPP_CASE(9) NNPP_CASE(10) NNPP_CASE(11) NNPP_CASE(12)
    NNPP_CASE(13) NNPP_CASE(14) NNPP_CASE(15) NNPP_CASE(16)
#undef NNPP_CASE
    default:
      return nnpp_set_error(c, NNPP_ERR_ARG, "n_species out of range");
  }
  if (e != hipSuccess) return nnpp_set_error(c, NNPP_ERR_HIP, hipGetErrorString(e));
  return NNPP_OK;
}

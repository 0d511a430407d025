// k_flux.hip — fused multi-species nonlocal Nernst-Planck flux (utils.h:27-842, all fields in
// one pass over the bonds).
//
// One thread per node, 64 consecutive nodes per wavefront reading their bonds from the
// wave-height slices (coalesced 256 B column reads per bond step; on a lattice the neighbour
// gathers of the 64 lanes are 64 consecutive nodes too). Per node and species the kernel keeps
//   sd_k = sum_b w_b (c_k,j - c_k,i)            (diffusion, coefficient applied afterwards)
//   sm_k = sum_b w_b c_k,j (phi_j - phi_i)      (electromigration)
// with w_b = Beta(xi) V_j / xi^(3+2s). The diffusivity and BV factor depend only on the bond's
// phase pair (utils.h:100-159, 211-238), so bonds to nodes of the same phase are summed in a
// first pass and the few solid-liquid bonds of interface nodes in a second pass, each pass
// applying one coefficient per species at the end instead of one select per bond.
//
// RESID: F_k = (U1_k - Un_k)/dt + (sum_KR1 + sum_KR2) - R_k, phi row of utils.h:798-841.
// !RESID: explicit update U_k^{n+1} = U_k^n - dt (sum_KR1 + sum_KR2 - R_k), Dirichlet rows,
//         clamp of main.cpp:805-837.
// MIXED: FP32 weights, FP32 differences taken from the FP64 state, Kahan-compensated FP32 sums
// per node and species; reactions, the update and the output stay FP64.
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
  const int32_t* len;
  const int32_t* col;
  const double* w64;
  const float* w32;
  const uint8_t* phase;
  const uint16_t* bcmask;
  const double* bcval;
  const double* U1;  // iterate (bond sums, reaction n1 terms)
  const double* Un;  // previous state (lagged phi, reaction n terms, time term)
  double* out;
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

template <int S, bool RESID, bool MIXED>
__global__ __launch_bounds__(NNPP_BLOCK) void k_flux(const FluxArgs a) {
  const int64_t i = nnpp_xcd_block(blockIdx.x, gridDim.x) * NNPP_BLOCK + threadIdx.x;
  if (i >= a.n) return;
  const int64_t n = a.n;
  const int64_t base = a.off[i / NNPP_WAVE] + (i % NNPP_WAVE);
  const int L = a.len[i / NNPP_WAVE];
  const int pi = a.phase[i];
  const double* __restrict__ U1 = a.U1;
  const double* __restrict__ Un = a.Un;
  const double* __restrict__ phin = Un + (int64_t)S * n;

  double ci[S];
#pragma unroll
  for (int k = 0; k < S; k++) ci[k] = U1[(int64_t)k * n + i];
  const double phii = phin[i];
  const double phi1i = RESID ? U1[(int64_t)S * n + i] : 0.0;

  double acc[S];
  int mixed = 0;
  double sphi = 0.0;
  int flagLL = 0;

  // pass 1: same-phase bonds; pass 2 (interface nodes only): solid-liquid bonds
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1 && !mixed) break;
    double sd[S], sm[S];
    float fd[S], fm[S], cd[S], cm[S];
#pragma unroll
    for (int k = 0; k < S; k++) {
      sd[k] = 0.0;
      sm[k] = 0.0;
      fd[k] = 0.0f;
      fm[k] = 0.0f;
      cd[k] = 0.0f;
      cm[k] = 0.0f;
    }
    for (int t = 0; t < L; t++) {
      const int64_t slot = base + (int64_t)t * NNPP_WAVE;
      const int32_t j = a.col[slot];
      if (j < 0) break;
      const int pj = a.phase[j];
      if (pass == 0) {
        if (pj != pi) {
          mixed = 1;
          continue;
        }
      } else if (pj == pi) {
        continue;
      }
      const double phij = phin[j];
      if (MIXED) {
        const float w = a.w32[slot];
        const float dphi = (float)(phij - phii);
#pragma unroll
        for (int k = 0; k < S; k++) {
          const double cj = U1[(int64_t)k * n + j];
          kahan_add(fd[k], cd[k], w * (float)(cj - ci[k]));
          kahan_add(fm[k], cm[k], (w * (float)cj) * dphi);
        }
      } else {
        const double w = a.w64[slot];
        const double wdphi = w * (phij - phii);
#pragma unroll
        for (int k = 0; k < S; k++) {
          const double cj = U1[(int64_t)k * n + j];
          sd[k] += w * (cj - ci[k]);
          sm[k] += wdphi * cj;
        }
      }
      if (RESID && pass == 0 && pi == 0) {
        if (MIXED)
          sphi += (double)(a.w32[slot] * (float)(U1[(int64_t)S * n + j] - phi1i));
        else
          sphi += a.w64[slot] * (U1[(int64_t)S * n + j] - phi1i);
        flagLL = 1;
      }
    }
#pragma unroll
    for (int k = 0; k < S; k++) {
      if (MIXED) {
        sd[k] = (double)fd[k];
        sm[k] = (double)fm[k];
      }
      double A, M;
      if (pass == 0) {
        A = pi ? a.A_SS[k] : a.A_LL[k];
        M = pi ? a.M_SS[k] : a.M_LL[k];
        acc[k] = A * sd[k] + M * sm[k];
      } else {
        acc[k] += a.A_SL[k] * sd[k] + a.M_SL[k] * sm[k];
      }
    }
  }

  const int intf = (pi == 1 && mixed) ? 1 : 0;  // utils.h:82-89
  double un[S];
#pragma unroll
  for (int k = 0; k < S; k++) un[k] = RESID ? Un[(int64_t)k * n + i] : ci[k];
  const unsigned bm = a.bcmask[i];

#pragma unroll
  for (int k = 0; k < S; k++) {
    double R = 0.0;
    if constexpr (S == 10) {
      if (a.chemistry == NNPP_CHEM_REFERENCE) R = ref_reaction(k, a.rx, pi, intf, ci, un, phii);
    }
    if (a.chemistry == NNPP_CHEM_GENERIC && k == 0 && intf) R = -a.r_diss * ci[0];
    const double flux = -acc[k];  // sum_KR_k1 + sum_KR_k2
    if (RESID) {
      a.out[(int64_t)k * n + i] = ((ci[k] - un[k]) / a.dt) + flux - R;
    } else {
      double v = un[k] - a.dt * (flux - R);
      if ((bm >> k) & 1u) v = a.bcval[(int64_t)k * n + i];
      const bool cl = ((a.clamp_always >> k) & 1u) || (((a.clamp_liquid >> k) & 1u) && pi == 0);
      if (cl && v < a.clamp_floor) v = a.clamp_floor;
      a.out[(int64_t)k * n + i] = v;
    }
  }
  if (RESID) {
    double o = 0.0;
    if (flagLL) {
      double q = 0.0;
#pragma unroll
      for (int k = 0; k < S; k++) q += a.z[k] * un[k];
      o = -sphi - a.phi_q * q;
    }
    a.out[(int64_t)S * n + i] = o;
  }
}

template <int S>
hipError_t launch_s(const FluxArgs& a, int resid, int mixed, int64_t nb, hipStream_t s) {
  if (resid) {
    if (mixed) k_flux<S, true, true><<<nb, NNPP_BLOCK, 0, s>>>(a);
    else k_flux<S, true, false><<<nb, NNPP_BLOCK, 0, s>>>(a);
  } else {
    if (mixed) k_flux<S, false, true><<<nb, NNPP_BLOCK, 0, s>>>(a);
    else k_flux<S, false, false><<<nb, NNPP_BLOCK, 0, s>>>(a);
  }
  return hipGetLastError();
}

}  // namespace

int nnpp_launch_flux(nnpp_ctx* c, int resid, int precision, double BV, const double* U1,
                     const double* Un, double* out) {
  const nnpp_params& p = c->p;
  const int S = p.n_species;
  FluxArgs a;
  a.n = c->n;
  a.chemistry = p.chemistry;
  a.off = c->bonds.off;
  a.len = c->bonds.len;
  a.col = c->bonds.col;
  a.w64 = c->bonds.w64;
  a.w32 = c->bonds.w32;
  a.phase = c->phase;
  a.bcmask = c->bcmask;
  a.bcval = c->bcval;
  a.U1 = U1;
  a.Un = Un;
  a.out = out;
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

  if (!c->n) return NNPP_OK;
  const int64_t nb = nnpp_nblocks(c->n);
  const int mixed = (precision == NNPP_MIXED);
  hipError_t e = hipSuccess;
  switch (S) {
#define NNPP_CASE(SS) \
  case SS: e = launch_s<SS>(a, resid, mixed, nb, c->stream); break;
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

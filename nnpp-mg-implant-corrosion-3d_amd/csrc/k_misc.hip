// k_misc.hip — device exclusive scan, bond-slice construction, front update (main.cpp:915-922
// plus the volume sum of main.cpp:944-949), Dirichlet residual rows (main.cpp:676-711) and
// the non-finite check.
#include <algorithm>
#include <cstring>
#include <vector>

#include "nnpp_internal.hpp"

namespace {

constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = NNPP_BLOCK * SCAN_ITEMS;

// Exclusive scan of one tile of 2048 values; tile totals go to bsum.
template <typename TI>
__global__ __launch_bounds__(NNPP_BLOCK) void k_scan_tile(const TI* __restrict__ in, int64_t n,
                                                          int64_t* __restrict__ out,
                                                          int64_t* __restrict__ bsum) {
  __shared__ int64_t s[NNPP_BLOCK];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  int64_t v[SCAN_ITEMS];
  int64_t tsum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    v[k] = (base + k < n) ? (int64_t)in[base + k] : 0;
    tsum += v[k];
  }
  s[threadIdx.x] = tsum;
  __syncthreads();
  for (int o = 1; o < NNPP_BLOCK; o <<= 1) {  // Hillis-Steele inclusive scan of thread sums
    const int64_t add = (threadIdx.x >= (unsigned)o) ? s[threadIdx.x - o] : 0;
    __syncthreads();
    s[threadIdx.x] += add;
    __syncthreads();
  }
  int64_t run = s[threadIdx.x] - tsum;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == NNPP_BLOCK - 1) bsum[blockIdx.x] = s[threadIdx.x];
}

__global__ void k_scan_add(int64_t n, int64_t* __restrict__ out,
                           const int64_t* __restrict__ boff) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += boff[i / SCAN_TILE];
}

__global__ void k_set_total(int64_t n, int64_t* __restrict__ out,
                            const int64_t* __restrict__ bsum_scan, int64_t nt,
                            const int64_t* __restrict__ bsum) {
  if (threadIdx.x == 0 && blockIdx.x == 0) out[n] = bsum_scan[nt - 1] + bsum[nt - 1];
}

// Longest row per 64-row slice.
__global__ void k_slice_len(int32_t n, const int64_t* __restrict__ row_ptr,
                            int32_t* __restrict__ len) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  atomicMax(&len[i / NNPP_WAVE], (int32_t)((row_ptr[i + 1] - row_ptr[i] + 3) / 4));
}

__global__ void k_times(int64_t ns, const int32_t* __restrict__ len, int32_t mult,
                        int32_t* __restrict__ out) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ns) out[c] = len[c] * mult;
}

struct WeightArgs {
  double delta, dx, V0, expo;
  const double* vol;
};

__device__ __forceinline__ double bond_weight(const WeightArgs& wa, const double* x,
                                              const double* y, const double* z, int64_t i,
                                              int32_t j) {
  const double xx = x[j] - x[i], yy = y[j] - y[i], zz = z[j] - z[i];
  const double xi = sqrt(xx * xx + yy * yy + zz * zz);
  const double V = wa.vol ? wa.vol[j] : wa.V0;
  return ((1.0 / pow(xi, wa.expo)) * nnpp_beta(xi, wa.delta, wa.dx)) * V;
}

// Fill the bond slices from CSR in the quad layout (entry t of row i at
// off[i/64] + 256 (t/4) + 4 (i%64) + t%4), with 16-bit column steps from first[i]; padding has
// weight 0 and repeats the last column. Steps above 65535 raise *maxstep.
__global__ void k_slice_fill(int32_t n, int64_t nrows_pad, const int64_t* __restrict__ row_ptr,
                             const int32_t* __restrict__ col, const double* __restrict__ x,
                             const double* __restrict__ y, const double* __restrict__ z,
                             WeightArgs wa, const int64_t* __restrict__ off,
                             const int32_t* __restrict__ len, int32_t* __restrict__ first,
                             int32_t* __restrict__ scol, uint16_t* __restrict__ d16,
                             double* __restrict__ w64, float* __restrict__ w32,
                             int* __restrict__ maxstep) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows_pad) return;
  const int64_t c = i / NNPP_WAVE;
  const int64_t base = off[c] + 4 * (i % NNPP_WAVE);
  const int L = 4 * len[c];
  const int64_t b0 = (i < n) ? row_ptr[i] : 0;
  const int rl = (i < n) ? (int)(row_ptr[i + 1] - b0) : 0;
  int32_t prev = (int32_t)(i < n ? i : 0);
  if (rl > 0) prev = col[b0];
  if (i < n) first[i] = prev;
  int big = 0;
  for (int t = 0; t < L; t++) {
    const int64_t slot = base + 256 * (int64_t)(t >> 2) + (t & 3);
    if (t < rl) {
      const int32_t j = col[b0 + t];
      const double w = bond_weight(wa, x, y, z, i, j);
      const int step = j - prev;
      big = step > big ? step : big;
      scol[slot] = j;
      d16[slot] = (uint16_t)(step & 0xffff);
      w64[slot] = w;
      w32[slot] = (float)w;
      prev = j;
    } else {
      scol[slot] = prev;
      d16[slot] = 0;
      w64[slot] = 0.0;
      w32[slot] = 0.0f;
    }
  }
  if (big > 65535) atomicMax(maxstep, big);
}

// Potential operator, step 1: free flags (liquid, not Dirichlet for phi, >= 1 liquid bond).
__global__ void k_pot_flags(int32_t n, int S, const int64_t* __restrict__ row_ptr,
                            const int32_t* __restrict__ col, const uint8_t* __restrict__ phase,
                            const uint16_t* __restrict__ bcmask, uint8_t* __restrict__ pfree) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint8_t f = 0;
  if (phase[i] == 0 && !((bcmask[i] >> S) & 1u)) {
    for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++)
      if (phase[col[b]] == 0) {
        f = 1;
        break;
      }
  }
  pfree[i] = f;
}

// Step 2: entries per free row (free-free liquid bonds) -> slice lengths.
__global__ void k_pot_len(int32_t n, const int64_t* __restrict__ row_ptr,
                          const int32_t* __restrict__ col, const uint8_t* __restrict__ phase,
                          const uint8_t* __restrict__ pfree, int32_t* __restrict__ len,
                          int32_t* __restrict__ rcount) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int cnt = 0;
  if (pfree[i])
    for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
      const int32_t j = col[b];
      if (phase[j] == 0 && pfree[j]) cnt++;
    }
  rcount[i] = cnt;
  atomicMax(&len[i / NNPP_WAVE], (cnt + 3) / 4);  // groups of 4 entries per lane
}

// Step 3: entries (quad layout), scaled diagonal, Dirichlet-neighbour diagonal part and the
// Dirichlet part of the right-hand side. Row i is scaled by V_i / V0 (symmetrises the operator
// for per-node volumes; 1 for the reference's constant volume). Column steps that do not fit
// 16 bits raise *maxstep above 65535 (the wide format is used then).
__global__ void k_pot_fill(int32_t n, int64_t nrows_pad, int S, const int64_t* __restrict__ row_ptr,
                           const int32_t* __restrict__ col, const uint8_t* __restrict__ phase,
                           const uint8_t* __restrict__ pfree, const uint16_t* __restrict__ bcmask,
                           const double* __restrict__ bcphi, const double* __restrict__ x,
                           const double* __restrict__ y, const double* __restrict__ z,
                           WeightArgs wa, const int64_t* __restrict__ off,
                           const int32_t* __restrict__ len, int32_t* __restrict__ first,
                           int32_t* __restrict__ scol, uint16_t* __restrict__ d16,
                           double* __restrict__ w64, float* __restrict__ w32,
                           double* __restrict__ pdd, float* __restrict__ pdd32,
                           double* __restrict__ pinv, float* __restrict__ pinv32,
                           double* __restrict__ pbdir, double* __restrict__ psc,
                           int* __restrict__ maxstep) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows_pad) return;
  const int64_t c = i / NNPP_WAVE;
  const int64_t base = off[c] + 4 * (i % NNPP_WAVE);
  const int L = 4 * len[c];
  int t = 0;
  int32_t prev = (int32_t)(i < n ? i : 0), f0 = prev;
  int big = 0;
  if (i < n) {
    const double sc = wa.vol ? wa.vol[i] / wa.V0 : 1.0;
    double diag = 0.0, ddir = 0.0, bdir = 0.0;
    if (pfree[i]) {
      for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
        const int32_t j = col[b];
        if (phase[j] != 0) continue;
        const double w = bond_weight(wa, x, y, z, i, j);
        diag += w;
        if (pfree[j]) {
          const int64_t slot = base + 256 * (int64_t)(t >> 2) + (t & 3);
          if (t == 0) f0 = prev = j;
          const int step = j - prev;  // rows are sorted: step >= 0
          big = step > big ? step : big;
          scol[slot] = j;
          d16[slot] = (uint16_t)(step & 0xffff);
          w64[slot] = sc * w;
          w32[slot] = (float)(sc * w);
          prev = j;
          t++;
        } else {  // a liquid neighbour of a free node is free or Dirichlet
          ddir += w;
          if ((bcmask[j] >> S) & 1u) bdir += w * bcphi[j];
        }
      }
    }
    first[i] = f0;
    pdd[i] = sc * ddir;
    pdd32[i] = (float)(sc * ddir);
    pinv[i] = pfree[i] ? 1.0 / (sc * diag) : 0.0;
    pinv32[i] = pfree[i] ? (float)(1.0 / (sc * diag)) : 0.0f;
    pbdir[i] = bdir;
    psc[i] = sc;
  }
  for (; t < L; t++) {  // padding: weight 0, column unchanged
    const int64_t slot = base + 256 * (int64_t)(t >> 2) + (t & 3);
    scol[slot] = prev;
    d16[slot] = 0;
    w64[slot] = 0.0;
    w32[slot] = 0.0f;
  }
  if (big > 65535) atomicMax(maxstep, big);
}

// main.cpp:915-922 front update, then the volume sum of main.cpp:944-949 over the new phase.
__global__ __launch_bounds__(NNPP_BLOCK) void k_front(int32_t n, const double* __restrict__ U0,
                                                      uint8_t* __restrict__ phase, double thr,
                                                      double sat, double* __restrict__ part) {
  double v[3] = {0.0, 0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * NNPP_BLOCK + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * NNPP_BLOCK) {
    uint8_t ph = phase[i];
    const double u = U0[i];
    if (ph == 1 && u <= thr) {
      ph = 0;
      phase[i] = 0;
      v[0] += 1.0;
    }
    if (ph == 1) {
      v[1] += (u - sat);
      v[2] += 1.0;
    }
  }
  nnpp_block_partial<3>(v, part);
}

// main.cpp:676-711: F = U1 - bc on Dirichlet rows/fields (bc already divided by CSolid for
// species).
__global__ void k_bc_rows(int32_t n, int F, const uint16_t* __restrict__ bcmask,
                          const double* __restrict__ bcval, const double* __restrict__ U1,
                          double* __restrict__ Fo) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned m = bcmask[i];
  if (!m) return;
  for (int f = 0; f < F; f++)
    if ((m >> f) & 1u) Fo[(int64_t)f * n + i] = U1[(int64_t)f * n + i] - bcval[(int64_t)f * n + i];
}

__global__ __launch_bounds__(NNPP_BLOCK) void k_nonfinite(int64_t m, const double* __restrict__ U,
                                                          double* __restrict__ part) {
  double v[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * NNPP_BLOCK + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * NNPP_BLOCK)
    if (!isfinite(U[i])) v[0] += 1.0;
  nnpp_block_partial<1>(v, part);
}

// One block: sum the partials of the previous kernel and update the device scalars.
__global__ __launch_bounds__(NNPP_BLOCK) void k_finalize(int op, const double* __restrict__ part,
                                                         int64_t nblk, double* __restrict__ scal,
                                                         double a0) {
  if (op == FIN_BNORM2 || op == FIN_NONFINITE) {
    double v[1];
    nnpp_final_sum<1>(v, part, nblk);
    if (threadIdx.x != 0) return;
    if (op == FIN_BNORM2) scal[SC_BNORM2] = v[0];
    else scal[SC_R3] = v[0];
  } else if (op == FIN_FRONT) {
    double v[3];
    nnpp_final_sum<3>(v, part, nblk);
    if (threadIdx.x != 0) return;
    scal[SC_R0] = v[0];
    scal[SC_R1] = v[1];
    scal[SC_R2] = v[2];
  } else {
    double v[2];
    nnpp_final_sum<2>(v, part, nblk);
    if (threadIdx.x != 0) return;
    switch (op) {
      case FIN_RESID:
        scal[SC_RR] = v[1];
        break;
      case FIN_RESID_START: {
        scal[SC_RHO] = v[0];
        scal[SC_RR] = v[1];
        const double t2 = a0 * scal[SC_BNORM2];
        scal[SC_TARGET2] = t2;
        scal[SC_ITERS] = 0.0;
        scal[SC_DONE] = (v[1] <= t2) ? 1.0 : 0.0;
        break;
      }
      case FIN_M_REPLACE:
      case FIN_M_REPLACE_START:
        scal[SC_RHO] = v[0];
        scal[SC_RR] = v[1];
        scal[SC_RRMAX] = v[1];
        scal[SC_TARGET2] = a0;
        if (op == FIN_M_REPLACE_START) scal[SC_ITERS] = 0.0;
        // after a replacement the FP64 test has just failed: run at least one more iteration
        scal[SC_DONE] = (op == FIN_M_REPLACE_START && v[1] <= a0) ? 1.0 : 0.0;
        break;
      default:
        break;
    }
  }
}

// Poisson operator action for nnpp_poisson_apply: K11 rows of utils.h:1794-1848 without the
// identity rows, weights computed from the coordinates. A flag `any` that reads 0 says `in` is
// all zeros: every sum is then +0 (each bond adds w (0 - 0)) and is stored without the bonds.
__global__ void k_pot_apply_csr(int32_t n, const int64_t* __restrict__ row_ptr,
                                const int32_t* __restrict__ col,
                                const uint8_t* __restrict__ phase, const double* __restrict__ x,
                                const double* __restrict__ y, const double* __restrict__ z,
                                WeightArgs wa, const double* __restrict__ in,
                                double* __restrict__ out, const int32_t* __restrict__ any) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  if (any && *any == 0) {
    out[i] = acc;
    return;
  }
  if (phase[i] == 0)
    for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
      const int32_t j = col[b];
      if (phase[j] != 0) continue;
      acc += bond_weight(wa, x, y, z, i, j) * (in[i] - in[j]);
    }
  out[i] = acc;
}

// K11 of utils.h:1794-1848 with its identity rows (Dirichlet rows, rows without a liquid-liquid
// bond): out = K11 in, diag = its diagonal. w = Beta(xi) V_j / xi^(3+2 s_phi).
__global__ void k_jac_phi(int32_t n, int S, const int64_t* __restrict__ row_ptr,
                          const int32_t* __restrict__ col, const uint8_t* __restrict__ phase,
                          const uint16_t* __restrict__ bcmask, const double* __restrict__ x,
                          const double* __restrict__ y, const double* __restrict__ z,
                          WeightArgs wa, const double* __restrict__ in, double* __restrict__ out,
                          double* __restrict__ diag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0, d = 0.0;
  int flag = 0;
  if (phase[i] == 0 && !((bcmask[i] >> S) & 1u))
    for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
      const int32_t j = col[b];
      if (phase[j] != 0) continue;
      const double w = bond_weight(wa, x, y, z, i, j);
      acc += w * (in[i] - in[j]);
      d += w;
      flag = 1;
    }
  out[i] = flag ? acc : in[i];
  if (diag) diag[i] = flag ? d : 1.0;
}

constexpr int DICT_HASH = 4096;

// Collect the distinct nonzero FP64 weights in a small open-addressing hash set; gives up
// (flag) as soon as a probe chain gets long or too many keys arrive.
__global__ void k_dict_insert(int64_t slots, const double* __restrict__ w,
                              unsigned long long* __restrict__ table, int* __restrict__ count,
                              int* __restrict__ fail) {
  unsigned long long last = 0;
  int64_t seen = 0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < slots;
       k += (int64_t)gridDim.x * blockDim.x) {
    if ((++seen & 63) == 0 && __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      return;
    const double v = w[k];
    if (v == 0.0) continue;
    const unsigned long long key = (unsigned long long)__double_as_longlong(v);
    if (key == last) continue;
    last = key;
    unsigned h = (unsigned)((key * 0x9E3779B97F4A7C15ull) >> 52);
    int probe = 0;
    for (; probe < 32; probe++, h = (h + 1) & (DICT_HASH - 1)) {
      const unsigned long long cur = table[h];
      if (cur == key) break;
      if (cur == 0ull) {
        const unsigned long long prev = atomicCAS(&table[h], 0ull, key);
        if (prev == 0ull) {
          if (atomicAdd(count, 1) >= NNPP_DICT_MAX - 1) atomicExch(fail, 1);
          break;
        }
        if (prev == key) break;
      }
    }
    if (probe == 32) {
      atomicExch(fail, 1);
      return;
    }
  }
}

// Class of every entry: index of its weight in the sorted table (class 0 = weight 0).
__global__ void k_dict_map(int64_t slots, const double* __restrict__ w,
                           const double* __restrict__ tab, int ntab, uint8_t* __restrict__ cls) {
  __shared__ double t[NNPP_DICT_MAX];
  for (int k = threadIdx.x; k < NNPP_DICT_MAX; k += blockDim.x) t[k] = k < ntab ? tab[k] : 0.0;
  __syncthreads();
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < slots;
       k += (int64_t)gridDim.x * blockDim.x) {
    const double v = w[k];
    int lo = 0, hi = ntab - 1;  // t is sorted ascending and contains v
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (t[mid] < v) lo = mid + 1;
      else hi = mid;
    }
    cls[k] = (uint8_t)lo;
  }
}

// ---- stencil form of the potential operator ----

// Offsets j - i of every bond into a hash set (keys stored as the 32-bit offset, never 0 since
// rows exclude themselves), with one row that has each offset; gives up (fail) beyond
// NNPP_STENCIL_MAX distinct offsets or on a long probe chain.
__global__ void k_st_collect(int32_t n, const int64_t* __restrict__ row_ptr,
                             const int32_t* __restrict__ col, unsigned long long* __restrict__ table,
                             int32_t* __restrict__ rep, int* __restrict__ count,
                             int* __restrict__ fail) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (__hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  unsigned long long last = 0;
  for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
    const unsigned long long key = (unsigned long long)(uint32_t)(col[b] - (int32_t)i);
    if (key == last) continue;
    last = key;
    unsigned h = (unsigned)((key * 0x9E3779B97F4A7C15ull) >> 52);
    int probe = 0;
    for (; probe < 32; probe++, h = (h + 1) & (DICT_HASH - 1)) {
      const unsigned long long cur = table[h];
      if (cur == key) break;
      if (cur == 0ull) {
        const unsigned long long prev = atomicCAS(&table[h], 0ull, key);
        if (prev == 0ull) {
          rep[h] = (int32_t)i;
          if (atomicAdd(count, 1) >= NNPP_STENCIL_MAX) atomicExch(fail, 1);
          break;
        }
        if (prev == key) break;
      }
    }
    if (probe == 32) {
      atomicExch(fail, 1);
      return;
    }
  }
}

// Weight of each offset from one bond that has it (the same arithmetic as k_pot_fill and
// k_slice_fill): potential exponent (tw64/tw32) and flux exponent (tf64).
__global__ void k_st_weights(int K, const int32_t* __restrict__ toff,
                             const int32_t* __restrict__ rep, const double* __restrict__ x,
                             const double* __restrict__ y, const double* __restrict__ z,
                             WeightArgs wa, WeightArgs wf, double* __restrict__ tw64,
                             float* __restrict__ tw32, double* __restrict__ tf64) {
  const int t = threadIdx.x;
  if (t >= K) return;
  const int64_t i = rep[t];
  const double w = bond_weight(wa, x, y, z, i, (int32_t)(i + toff[t]));
  tw64[t] = w;
  tw32[t] = (float)w;
  tf64[t] = bond_weight(wf, x, y, z, i, (int32_t)(i + toff[t]));
}

// Row masks of all bonds; fails if an entry's offset is not in the table, the entries of a
// row are not in ascending order, or a bond's weight differs (in any bit) from its offset's.
// fail_flux: some bond's flux weight differs from its offset's.
__global__ void k_st_mask(int32_t n, int K, const int64_t* __restrict__ row_ptr,
                          const int32_t* __restrict__ col, const int32_t* __restrict__ toff,
                          const double* __restrict__ tw64, const double* __restrict__ tf64,
                          const double* __restrict__ x, const double* __restrict__ y,
                          const double* __restrict__ z, WeightArgs wa, WeightArgs wf,
                          uint4* __restrict__ bmask, int* __restrict__ fail,
                          int* __restrict__ fail_flux) {
  __shared__ int32_t t_off[NNPP_STENCIL_MAX];
  __shared__ double t_w[NNPP_STENCIL_MAX];
  __shared__ double t_f[NNPP_STENCIL_MAX];
  for (int t = threadIdx.x; t < K; t += blockDim.x) {
    t_off[t] = toff[t];
    t_w[t] = tw64[t];
    t_f[t] = tf64[t];
  }
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t m[4] = {0u, 0u, 0u, 0u};
  int prev = -1, bad = 0;
  for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
    const int32_t j = col[b];
    const int32_t o = j - (int32_t)i;
    int lo = 0, hi = K - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (t_off[mid] < o) lo = mid + 1;
      else hi = mid;
    }
    if (t_off[lo] != o || lo <= prev) {
      bad = 1;
      break;
    }
    const double w = bond_weight(wa, x, y, z, i, j);
    if (__double_as_longlong(w) != __double_as_longlong(t_w[lo])) {
      bad = 1;
      break;
    }
    if (__double_as_longlong(bond_weight(wf, x, y, z, i, j)) != __double_as_longlong(t_f[lo]))
      atomicExch(fail_flux, 1);
    m[lo >> 5] |= 1u << (lo & 31);
    prev = lo;
  }
  if (bad) atomicExch(fail, 1);
  bmask[i] = make_uint4(m[0], m[1], m[2], m[3]);
}

// Stencil form of k_pot_flags/k_pot_len/k_pot_fill: entries are the bonds of the row mask
// whose far end is free, in ascending offset order (the neighbour-list order), so the diagonal
// sums match the column form bit for bit. Counts entries and free rows per block.
__global__ __launch_bounds__(NNPP_BLOCK) void k_pot_st_fill(
    int32_t n, int S, int K, const int32_t* __restrict__ toff, const double* __restrict__ tw64,
    const uint4* __restrict__ bmask, const uint8_t* __restrict__ phase,
    const uint8_t* __restrict__ pfree, const uint16_t* __restrict__ bcmask,
    const double* __restrict__ bcphi, uint4* __restrict__ pmask, double* __restrict__ pdd,
    float* __restrict__ pdd32, double* __restrict__ pinv, float* __restrict__ pinv32,
    double* __restrict__ pbdir, double* __restrict__ psc, double* __restrict__ part) {
  double v[3] = {0.0, 0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * NNPP_BLOCK + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * NNPP_BLOCK) {
    uint32_t pm[4] = {0u, 0u, 0u, 0u};
    double diag = 0.0, ddir = 0.0, bdir = 0.0;
    int cnt = 0;
    const bool fr = pfree[i] != 0;
    if (fr) {
      const uint4 bm = bmask[i];
      const uint32_t words[4] = {bm.x, bm.y, bm.z, bm.w};
      for (int t = 0; t < K; t++) {
        if (!((words[t >> 5] >> (t & 31)) & 1u)) continue;
        const int32_t j = (int32_t)i + toff[t];
        if (phase[j] != 0) continue;
        const double w = tw64[t];
        diag += w;
        if (pfree[j]) {
          pm[t >> 5] |= 1u << (t & 31);
          cnt++;
        } else {  // a liquid neighbour of a free node is free or Dirichlet
          ddir += w;
          if ((bcmask[j] >> S) & 1u) bdir += w * bcphi[j];
        }
      }
    }
    pmask[i] = make_uint4(pm[0], pm[1], pm[2], pm[3]);
    const double sc = 1.0;  // constant node volume (the stencil form requires it)
    pdd[i] = sc * ddir;
    pdd32[i] = (float)(sc * ddir);
    pinv[i] = fr ? 1.0 / (sc * diag) : 0.0;
    pinv32[i] = fr ? (float)(1.0 / (sc * diag)) : 0.0f;
    pbdir[i] = bdir;
    psc[i] = sc;
    v[0] += (double)cnt;
    v[1] += fr ? 1.0 : 0.0;
  }
  nnpp_block_partial<3>(v, part);
}

// Box-lattice check: row i = x + nx (y + ny z) must hold exactly the NNPP_BALL_M ball offsets
// whose far end lies inside the nx x ny x nz box (bit t of bmask = entry t of nnpp_ball's
// enumeration, the order of the offsets).
__global__ void k_st_box(int32_t n, int nx, int ny, int nz, const uint4* __restrict__ bmask,
                         int* __restrict__ fail) {
  constexpr nnpp_ball<NNPP_BALL_M> B{};
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((int64_t)nx * ny));
  uint32_t m[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int t = 0; t < B.K; t++) {
    const int xx = x + B.dx[t], yy = y + B.ldy[B.line[t]], zz = z + B.ldz[B.line[t]];
    if (xx >= 0 && xx < nx && yy >= 0 && yy < ny && zz >= 0 && zz < nz)
      m[t >> 5] |= 1u << (t & 31);
  }
  const uint4 b = bmask[i];
  if (b.x != m[0] || b.y != m[1] || b.z != m[2] || b.w != m[3]) atomicExch(fail, 1);
}

// Lane order of each 256-row tile of the stencil SpMV: rows whose mask holds every entry first,
// then the other rows with entries, then rows without any (and rows past n), each group in
// row order. Waves made of complete rows then run without per-entry mask tests.
__global__ __launch_bounds__(NNPP_BLOCK) void k_st_perm(int32_t n, int K,
                                                        const uint4* __restrict__ pmask,
                                                        uint8_t* __restrict__ perm) {
  __shared__ int wcount[3][NNPP_BLOCK / NNPP_WAVE];
  const int64_t i = (int64_t)blockIdx.x * NNPP_BLOCK + threadIdx.x;
  int cls = 2;
  if (i < n) {
    const uint4 m = pmask[i];
    const int cnt = __popc(m.x) + __popc(m.y) + __popc(m.z) + __popc(m.w);
    cls = cnt == K ? 0 : (cnt ? 1 : 2);
  }
  const int lane = threadIdx.x % NNPP_WAVE, wid = threadIdx.x / NNPP_WAVE;
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (NNPP_WAVE - lane));
  int rank = 0;
  for (int c = 0; c < 3; c++) {
    const unsigned long long b = __ballot(cls == c);
    if (cls == c) rank = __popcll(b & below);
    if (lane == 0) wcount[c][wid] = __popcll(b);
  }
  __syncthreads();
  int pos = 0;
  for (int c = 0; c < 3; c++)
    for (int w = 0; w < NNPP_BLOCK / NNPP_WAVE; w++)
      if (c < cls || (c == cls && w < wid)) pos += wcount[c][w];
  perm[(int64_t)blockIdx.x * NNPP_BLOCK + pos + rank] = (uint8_t)threadIdx.x;
}

}  // namespace

// ---------------------------------------------------------------- host launchers ---------

template <typename TI>
static int scan_impl(nnpp_ctx* c, const TI* in, int64_t n, int64_t* out) {
  hipStream_t s = c->stream;
  if (n == 0) {
    NNPP_HIP(c, hipMemsetAsync(out, 0, sizeof(int64_t), s));
    return NNPP_OK;
  }
  const int64_t nt = (n + SCAN_TILE - 1) / SCAN_TILE;
  int64_t *bsum = nullptr, *boff = nullptr;
  NNPP_HIP(c, hipMalloc((void**)&bsum, sizeof(int64_t) * nt));
  NNPP_HIP(c, hipMalloc((void**)&boff, sizeof(int64_t) * (nt + 1)));
  k_scan_tile<TI><<<nt, NNPP_BLOCK, 0, s>>>(in, n, out, bsum);
  int st = NNPP_OK;
  if (nt > 1) {
    st = scan_impl<int64_t>(c, bsum, nt, boff);
    if (st == NNPP_OK) {
      k_scan_add<<<nnpp_nblocks(n), NNPP_BLOCK, 0, s>>>(n, out, boff);
      k_set_total<<<1, 1, 0, s>>>(n, out, boff, nt, bsum);
    }
  } else {
    NNPP_HIP(c, hipMemcpyAsync(out + n, bsum, sizeof(int64_t), hipMemcpyDeviceToDevice, s));
  }
  hipError_t e = hipStreamSynchronize(s);
  hipFree(bsum);
  hipFree(boff);
  if (e != hipSuccess) return nnpp_set_error(c, NNPP_ERR_HIP, hipGetErrorString(e));
  return st;
}

int nnpp_device_exclusive_scan_i32(nnpp_ctx* c, const int32_t* in, int64_t n, int64_t* out) {
  return scan_impl<int32_t>(c, in, n, out);
}

static int slices_alloc(nnpp_ctx* c, nnpp_slices& sl, int64_t nslice) {
  if (sl.nslice != nslice || !sl.off) {
    if (sl.off) hipFree(sl.off);
    if (sl.len) hipFree(sl.len);
    NNPP_HIP(c, hipMalloc((void**)&sl.off, sizeof(int64_t) * (nslice + 1)));
    NNPP_HIP(c, hipMalloc((void**)&sl.len, sizeof(int32_t) * (nslice ? nslice : 1)));
    sl.nslice = nslice;
  }
  return NNPP_OK;
}

static int slices_reserve(nnpp_ctx* c, nnpp_slices& sl, int64_t slots) {
  if (slots > sl.cap_slots) {
    if (sl.col) hipFree(sl.col);
    if (sl.w64) hipFree(sl.w64);
    if (sl.w32) hipFree(sl.w32);
    if (sl.d16) hipFree(sl.d16);
    if (sl.cls) hipFree(sl.cls);
    sl.cls = nullptr;
    sl.col = nullptr;
    sl.w64 = nullptr;
    sl.w32 = nullptr;
    sl.d16 = nullptr;
    const int64_t cap = slots + slots / 16 + 256;
    NNPP_HIP(c, hipMalloc((void**)&sl.col, sizeof(int32_t) * cap));
    NNPP_HIP(c, hipMalloc((void**)&sl.w64, sizeof(double) * cap));
    NNPP_HIP(c, hipMalloc((void**)&sl.w32, sizeof(float) * cap));
    NNPP_HIP(c, hipMalloc((void**)&sl.d16, sizeof(uint16_t) * cap));
    sl.cap_slots = cap;
  }
  sl.slots = slots;
  return NNPP_OK;
}

// Slice offsets from per-slice lengths: off = exclusive scan of 64 * len.
static int slices_offsets(nnpp_ctx* c, nnpp_slices& sl, int32_t entries_per_len) {
  int32_t* tmp = nullptr;
  NNPP_HIP(c, hipMalloc((void**)&tmp, sizeof(int32_t) * (sl.nslice ? sl.nslice : 1)));
  k_times<<<nnpp_nblocks(sl.nslice ? sl.nslice : 1), NNPP_BLOCK, 0, c->stream>>>(
      sl.nslice, sl.len, entries_per_len, tmp);
  int st = scan_impl<int32_t>(c, tmp, sl.nslice, sl.off);
  hipFree(tmp);
  if (st) return st;
  int64_t slots = 0;
  NNPP_HIP(c, hipMemcpy(&slots, sl.off + sl.nslice, sizeof(int64_t), hipMemcpyDeviceToHost));
  return slices_reserve(c, sl, slots);
}

static WeightArgs weight_args(const nnpp_ctx* c, double expo) {
  WeightArgs wa;
  wa.delta = c->p.delta;
  wa.dx = c->p.dx;
  wa.V0 = c->p.dx * c->p.dy * c->p.dz;
  wa.expo = expo;
  wa.vol = c->vol;
  return wa;
}

int nnpp_build_bond_slices(nnpp_ctx* c) {
  const int32_t n = c->n;
  const int64_t ns = (n + NNPP_WAVE - 1) / NNPP_WAVE;
  int st = slices_alloc(c, c->bonds, ns);
  if (st) return st;
  hipStream_t s = c->stream;
  NNPP_HIP(c, hipMemsetAsync(c->bonds.len, 0, sizeof(int32_t) * (ns ? ns : 1), s));
  if (n) k_slice_len<<<nnpp_nblocks(n), NNPP_BLOCK, 0, s>>>(n, c->row_ptr, c->bonds.len);
  NNPP_HIP(c, hipGetLastError());
  st = slices_offsets(c, c->bonds, 4 * NNPP_WAVE);
  if (st) return st;
  if (!c->bonds.first)
    NNPP_HIP(c, hipMalloc((void**)&c->bonds.first, sizeof(int32_t) * (n ? n : 1)));
  int* dmax = nullptr;
  NNPP_HIP(c, hipMalloc((void**)&dmax, sizeof(int)));
  NNPP_HIP(c, hipMemsetAsync(dmax, 0, sizeof(int), s));
  const int64_t npad = ns * NNPP_WAVE;
  if (npad)
    k_slice_fill<<<nnpp_nblocks(npad), NNPP_BLOCK, 0, s>>>(
        n, npad, c->row_ptr, c->col, c->x, c->y, c->z, weight_args(c, c->w_exponent),
        c->bonds.off, c->bonds.len, c->bonds.first, c->bonds.col, c->bonds.d16, c->bonds.w64,
        c->bonds.w32, dmax);
  NNPP_HIP(c, hipGetLastError());
  int hmax = 0;
  NNPP_HIP(c, hipMemcpyAsync(&hmax, dmax, sizeof(int), hipMemcpyDeviceToHost, s));
  NNPP_HIP(c, hipStreamSynchronize(s));
  hipFree(dmax);
  c->bonds.wide = (hmax > 65535 || (c->format & NNPP_FORMAT_WIDE_COLUMNS)) ? 1 : 0;
  c->bonds.nnz = c->nnz;
  c->geom_dirty = false;
  return NNPP_OK;
}

void nnpp_free_stencil(nnpp_stencil& t) {
  void* ps[] = {t.toff, t.lc, t.seg_start, t.seg_len, t.seg_base, t.piece, t.tw64, t.tw32, t.tf64,
                t.tf32, t.bmask, t.fperm, t.pmask, t.perm, t.zmfull};
  for (void* q : ps)
    if (q) hipFree(q);
  t = nnpp_stencil();
}

// Whether the stencil is the NNPP_BALL_M lattice ball, for the potential SpMV's specialised
// form: offset t is entry t of nnpp_ball's enumeration at dz*SZ + dy*SY + dx (SY, SZ read off
// the entries (0, 1, 0) and (1, 0, 0)), a line's entries sit at consecutive LDS positions, and
// the weights depend on dx^2 + dy^2 + dz^2 only, bit for bit. Fills T.ball and T.ball_args.
static void detect_ball(nnpp_stencil& T, const std::vector<int32_t>& toff,
                        const std::vector<int32_t>& lc, const std::vector<float>& w32,
                        const std::vector<double>& w64) {
  constexpr nnpp_ball<NNPP_BALL_M> B{};
  T.ball = 0;
  if (T.K != B.K) return;
  int ty = -1, tz = -1;
  for (int t = 0; t < B.K; t++) {
    const int l = B.line[t];
    if (B.ldz[l] == 0 && B.ldy[l] == 1 && B.dx[t] == 0) ty = t;
    if (B.ldz[l] == 1 && B.ldy[l] == 0 && B.dx[t] == 0) tz = t;
  }
  if (ty < 0 || tz < 0) return;
  const int64_t SY = toff[ty], SZ = toff[tz];
  nnpp_ball_args a = {};
  bool seen[NNPP_BALL_CLASSES] = {};
  for (int t = 0; t < B.K; t++) {
    const int l = B.line[t];
    if ((int64_t)toff[t] != B.ldz[l] * SZ + B.ldy[l] * SY + B.dx[t]) return;
    if (B.pos[t] == 0) a.lc[l] = lc[t];
    else if (lc[t] != a.lc[l] + B.pos[t]) return;
    const int d2 = B.d2[t];
    if (!seen[d2]) {
      seen[d2] = true;
      a.w32[d2] = w32[t];
      a.w64[d2] = w64[t];
    } else if (std::memcmp(&a.w32[d2], &w32[t], sizeof(float)) ||
               std::memcmp(&a.w64[d2], &w64[t], sizeof(double))) {
      return;
    }
  }
  T.ball_args = a;
  T.ball = NNPP_BALL_M;
}

// Detect the stencil form of the potential operator for the current geometry (sets
// c->pst.state to 1 or -1).
static int build_stencil(nnpp_ctx* c) {
  nnpp_stencil& T = c->pst;
  const int32_t n = c->n;
  hipStream_t s = c->stream;
  nnpp_free_stencil(T);
  T.state = -1;  // until shown otherwise
  if (n == 0 || c->nnz <= 0 || c->vol) return NNPP_OK;
  unsigned long long* table = nullptr;
  int32_t* rep = nullptr;
  int* flags = nullptr;  // [0] count, [1] fail, [2] flux weights fail
  NNPP_HIP(c, hipMalloc((void**)&table, sizeof(unsigned long long) * DICT_HASH));
  NNPP_HIP(c, hipMalloc((void**)&rep, sizeof(int32_t) * DICT_HASH));
  NNPP_HIP(c, hipMalloc((void**)&flags, sizeof(int) * 3));
  NNPP_HIP(c, hipMemsetAsync(table, 0, sizeof(unsigned long long) * DICT_HASH, s));
  NNPP_HIP(c, hipMemsetAsync(flags, 0, sizeof(int) * 2, s));
  k_st_collect<<<nnpp_nblocks(n), NNPP_BLOCK, 0, s>>>(n, c->row_ptr, c->col, table, rep, flags,
                                                       flags + 1);
  NNPP_HIP(c, hipGetLastError());
  std::vector<unsigned long long> ht(DICT_HASH);
  std::vector<int32_t> hr(DICT_HASH);
  int hf[3] = {0, 0, 0};
  NNPP_HIP(c, hipMemcpyAsync(ht.data(), table, sizeof(unsigned long long) * DICT_HASH,
                             hipMemcpyDeviceToHost, s));
  NNPP_HIP(c, hipMemcpyAsync(hr.data(), rep, sizeof(int32_t) * DICT_HASH, hipMemcpyDeviceToHost, s));
  NNPP_HIP(c, hipMemcpyAsync(hf, flags, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
  NNPP_HIP(c, hipStreamSynchronize(s));
  hipFree(table);
  hipFree(rep);
  if (hf[1] || hf[0] > NNPP_STENCIL_MAX || hf[0] == 0) {
    hipFree(flags);
    return NNPP_OK;
  }
  std::vector<std::pair<int32_t, int32_t>> offs;  // (offset, representative row)
  for (int h = 0; h < DICT_HASH; h++)
    if (ht[h]) offs.push_back({(int32_t)(uint32_t)ht[h], hr[h]});
  std::sort(offs.begin(), offs.end());
  const int K = (int)offs.size();
  std::vector<int32_t> toff(K), reps(K), lc(K), rs, rl, rb, pc;
  for (int t = 0; t < K; t++) {
    toff[t] = offs[t].first;
    reps[t] = offs[t].second;
  }
  int base = 0;
  for (int t = 0; t < K;) {  // segments: ranges [toff + r0, toff + r0 + 255] that overlap or touch
    int u = t + 1;
    while (u < K && (int64_t)toff[u] <= (int64_t)toff[u - 1] + NNPP_BLOCK) u++;
    const int span = toff[u - 1] - toff[t] + 1;
    const int64_t staged = (int64_t)NNPP_BLOCK + span - 1;
    if ((int64_t)base + staged > (int64_t)1 << 20) {  // far beyond any LDS: not a stencil
      hipFree(flags);
      return NNPP_OK;
    }
    rs.push_back(toff[t]);
    rl.push_back(span);
    rb.push_back(base);
    for (int q = t; q < u; q++) lc[q] = base + (toff[q] - toff[t]);
    for (int64_t k = 0; k < staged; k += NNPP_BLOCK) {
      pc.push_back(toff[t] + (int32_t)k);
      pc.push_back(base + (int32_t)k);
      pc.push_back((int32_t)(staged - k < NNPP_BLOCK ? staged - k : NNPP_BLOCK));
    }
    base += (int)staged;
    t = u;
  }
  const int R = (int)rs.size();
  const int P = (int)pc.size() / 3;
  // the FP64 tile (plus the kernels' few static LDS words) must fit one workgroup's LDS; it
  // also holds the tile's 256 results at the end
  int max_lds = 65536;
  if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) !=
      hipSuccess)
    max_lds = 65536;
  if (base < NNPP_BLOCK || (size_t)base * sizeof(double) + 1024 > (size_t)max_lds) {
    hipFree(flags);
    return NNPP_OK;
  }
  T.K = K;
  T.R = R;
  T.P = P;
  T.off_min = toff[0];
  T.off_max = toff[K - 1];
  T.lds_elems = base;
  // the SpMV reads lc and the weights in chunks of NNPP_STENCIL_CHUNK consecutive entries: both
  // tables are padded to a multiple of it (LDS position 0, weight 0)
  const int Kp = (K + NNPP_STENCIL_CHUNK - 1) / NNPP_STENCIL_CHUNK * NNPP_STENCIL_CHUNK;
  lc.resize(Kp, 0);
  int32_t* reps_d = nullptr;
  NNPP_HIP(c, hipMalloc((void**)&T.toff, sizeof(int32_t) * K));
  NNPP_HIP(c, hipMalloc((void**)&T.lc, sizeof(int32_t) * Kp));
  NNPP_HIP(c, hipMalloc((void**)&reps_d, sizeof(int32_t) * K));
  NNPP_HIP(c, hipMalloc((void**)&T.seg_start, sizeof(int32_t) * R));
  NNPP_HIP(c, hipMalloc((void**)&T.seg_len, sizeof(int32_t) * R));
  NNPP_HIP(c, hipMalloc((void**)&T.seg_base, sizeof(int32_t) * R));
  const int Ppad = (P + NNPP_PIECE_BATCH - 1) / NNPP_PIECE_BATCH * NNPP_PIECE_BATCH;
  pc.resize(3 * (size_t)Ppad, 0);  // empty pieces: the SpMV reads whole batches of the table
  NNPP_HIP(c, hipMalloc((void**)&T.piece, sizeof(int32_t) * 3 * Ppad));
  NNPP_HIP(c, hipMalloc((void**)&T.tw64, sizeof(double) * Kp));
  NNPP_HIP(c, hipMalloc((void**)&T.tw32, sizeof(float) * Kp));
  NNPP_HIP(c, hipMemsetAsync(T.tw64, 0, sizeof(double) * Kp, s));
  NNPP_HIP(c, hipMemsetAsync(T.tw32, 0, sizeof(float) * Kp, s));
  NNPP_HIP(c, hipMalloc((void**)&T.tf64, sizeof(double) * Kp));
  NNPP_HIP(c, hipMemsetAsync(T.tf64, 0, sizeof(double) * Kp, s));
  NNPP_HIP(c, hipMalloc((void**)&T.tf32, sizeof(float) * Kp));
  NNPP_HIP(c, hipMalloc((void**)&T.bmask, sizeof(uint4) * n));
  NNPP_HIP(c, hipMalloc((void**)&T.pmask, sizeof(uint4) * n));
  NNPP_HIP(c, hipMalloc((void**)&T.perm, (size_t)nnpp_nblocks(n) * NNPP_BLOCK));
  NNPP_HIP(c, hipMalloc((void**)&T.fperm, (size_t)nnpp_nblocks(n) * NNPP_BLOCK));
  NNPP_HIP(c, hipMemcpyAsync(T.toff, toff.data(), sizeof(int32_t) * K, hipMemcpyHostToDevice, s));
  NNPP_HIP(c, hipMemcpyAsync(T.lc, lc.data(), sizeof(int32_t) * Kp, hipMemcpyHostToDevice, s));
  NNPP_HIP(c, hipMemcpyAsync(reps_d, reps.data(), sizeof(int32_t) * K, hipMemcpyHostToDevice, s));
  NNPP_HIP(c, hipMemcpyAsync(T.seg_start, rs.data(), sizeof(int32_t) * R, hipMemcpyHostToDevice, s));
  NNPP_HIP(c, hipMemcpyAsync(T.seg_len, rl.data(), sizeof(int32_t) * R, hipMemcpyHostToDevice, s));
  NNPP_HIP(c, hipMemcpyAsync(T.seg_base, rb.data(), sizeof(int32_t) * R, hipMemcpyHostToDevice, s));
  NNPP_HIP(c, hipMemcpyAsync(T.piece, pc.data(), sizeof(int32_t) * 3 * Ppad, hipMemcpyHostToDevice,
                             s));
  const WeightArgs wa = weight_args(c, 3.0 + 2.0 * c->p.s_phi);
  const WeightArgs wf = weight_args(c, c->w_exponent);
  k_st_weights<<<1, NNPP_STENCIL_MAX, 0, s>>>(K, T.toff, reps_d, c->x, c->y, c->z, wa, wf,
                                              T.tw64, T.tw32, T.tf64);
  NNPP_HIP(c, hipMemsetAsync(flags, 0, sizeof(int) * 3, s));
  k_st_mask<<<nnpp_nblocks(n), NNPP_BLOCK, 0, s>>>(n, K, c->row_ptr, c->col, T.toff, T.tw64,
                                                    T.tf64, c->x, c->y, c->z, wa, wf, T.bmask,
                                                    flags + 1, flags + 2);
  NNPP_HIP(c, hipGetLastError());
  std::vector<double> tf(K), pw64(K);
  std::vector<float> pw32(K);
  NNPP_HIP(c, hipMemcpyAsync(hf, flags, sizeof(int) * 3, hipMemcpyDeviceToHost, s));
  NNPP_HIP(c, hipMemcpyAsync(tf.data(), T.tf64, sizeof(double) * K, hipMemcpyDeviceToHost, s));
  NNPP_HIP(c, hipMemcpyAsync(pw64.data(), T.tw64, sizeof(double) * K, hipMemcpyDeviceToHost, s));
  NNPP_HIP(c, hipMemcpyAsync(pw32.data(), T.tw32, sizeof(float) * K, hipMemcpyDeviceToHost, s));
  NNPP_HIP(c, hipStreamSynchronize(s));
  hipFree(reps_d);
  hipFree(flags);
  if (hf[1]) {
    nnpp_free_stencil(T);
    T.state = -1;
    return NNPP_OK;
  }
  T.flux_ok = hf[2] ? 0 : 1;
  {
    std::vector<float> tf32(Kp, 0.0f);
    for (int t = 0; t < K; t++) tf32[t] = (float)tf[t];
    NNPP_HIP(c, hipMemcpyAsync(T.tf32, tf32.data(), sizeof(float) * Kp, hipMemcpyHostToDevice, s));
    NNPP_HIP(c, hipStreamSynchronize(s));
  }
  k_st_perm<<<nnpp_nblocks(n), NNPP_BLOCK, 0, s>>>(n, K, T.bmask, T.fperm);
  NNPP_HIP(c, hipGetLastError());
  for (int t = 0; t < K; t++)
    if (tf[t] != 0.0) T.fnz[t >> 5] |= 1u << (t & 31);
  detect_ball(T, toff, lc, pw32, pw64);
  if (T.ball == NNPP_BALL_M && T.flux_ok) {  // the flux weights by class as well?
    constexpr nnpp_ball<NNPP_BALL_M> B{};
    double wc[NNPP_BALL_CLASSES] = {};
    bool seen[NNPP_BALL_CLASSES] = {};
    bool ok = true;
    T.fball_wsum = 0.0;
    for (int t = 0; t < B.K && ok; t++) {
      const int d2 = B.d2[t];
      if (!seen[d2]) {
        seen[d2] = true;
        wc[d2] = tf[t];
      } else if (std::memcmp(&wc[d2], &tf[t], sizeof(double))) {
        ok = false;
      }
      T.fball_wsum += (double)(float)tf[t];
    }
    for (int d = 0; d < NNPP_BALL_CLASSES; d++) {
      T.fball_w32[d] = (float)wc[d];
      T.fball_w64[d] = wc[d];
    }
    T.fball = ok ? 1 : 0;
  }
  if (T.ball == NNPP_BALL_M) {  // the full box, numbered x fastest?
    constexpr nnpp_ball<NNPP_BALL_M> B{};
    int64_t sy = -1, sz = -1;
    for (int t = 0; t < B.K; t++) {
      const int l = B.line[t];
      if (B.ldz[l] == 0 && B.ldy[l] == 1 && B.dx[t] == 0) sy = toff[t];
      if (B.ldz[l] == 1 && B.ldy[l] == 0 && B.dx[t] == 0) sz = toff[t];
    }
    if (sy > 2 * NNPP_BALL_M && sz > 0 && sz % sy == 0 && n % sz == 0) {
      const int nx = (int)sy, ny = (int)(sz / sy), nz = (int)(n / sz);
      int* bf = nullptr;
      NNPP_HIP(c, hipMalloc((void**)&bf, sizeof(int)));
      NNPP_HIP(c, hipMemsetAsync(bf, 0, sizeof(int), s));
      k_st_box<<<nnpp_nblocks(n), NNPP_BLOCK, 0, s>>>(n, nx, ny, nz, T.bmask, bf);
      int hb = 1;
      NNPP_HIP(c, hipGetLastError());
      NNPP_HIP(c, hipMemcpyAsync(&hb, bf, sizeof(int), hipMemcpyDeviceToHost, s));
      NNPP_HIP(c, hipStreamSynchronize(s));
      hipFree(bf);
      // (the z-marching kernels address nodes by 32-bit byte offsets: n * 8 < 2^32)
      if (!hb && ny > 2 * NNPP_BALL_M && nz >= 1 && (int64_t)n * 8 < ((int64_t)1 << 32)) {
        T.box = 1;
        T.nx = nx;
        T.ny = ny;
        T.nz = nz;
      }
    }
  }
  T.state = 1;
  return NNPP_OK;
}

// Stencil form of the potential operator for the current phase.
static int build_potential_stencil(nnpp_ctx* c) {
  const int32_t n = c->n;
  const int S = c->p.n_species;
  hipStream_t s = c->stream;
  nnpp_stencil& T = c->pst;
  const int64_t g = nnpp_red_grid(nnpp_nblocks(n));
  k_pot_flags<<<nnpp_nblocks(n), NNPP_BLOCK, 0, s>>>(n, S, c->row_ptr, c->col, c->phase,
                                                     c->bcmask, c->pfree);
  k_pot_st_fill<<<g, NNPP_BLOCK, 0, s>>>(n, S, T.K, T.toff, T.tw64, T.bmask, c->phase, c->pfree,
                                         c->bcmask, c->bcval + (int64_t)S * n, T.pmask, c->pdd,
                                         c->pdd32, c->pinv, c->pinv32, c->pbdir, c->psc,
                                         c->red_part);
  NNPP_HIP(c, hipGetLastError());
  k_st_perm<<<nnpp_nblocks(n), NNPP_BLOCK, 0, s>>>(n, T.K, T.pmask, T.perm);
  NNPP_HIP(c, hipGetLastError());
  if (T.box) {
    const int zst = nnpp_zm_prepare(c);
    if (zst) return zst;
  }
  int st = nnpp_finalize(c, FIN_FRONT, g, 0.0);
  if (st) return st;
  NNPP_HIP(c, hipMemcpyAsync(c->h_scal, c->scal, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost, s));
  NNPP_HIP(c, hipStreamSynchronize(s));
  c->pot.nnz = (int64_t)c->h_scal[SC_R0];
  c->pot_nfree = (int64_t)c->h_scal[SC_R1];
  T.active = 1;
  c->pot_dirty = false;
  return NNPP_OK;
}

int nnpp_stencil_ensure(nnpp_ctx* c) {
  if (c->pst.state == 0 && c->n && !c->vol) return build_stencil(c);
  return NNPP_OK;
}

int nnpp_build_potential(nnpp_ctx* c) {
  const int32_t n = c->n;
  const int S = c->p.n_species;
  const int64_t ns = (n + NNPP_WAVE - 1) / NNPP_WAVE;
  hipStream_t s = c->stream;
  c->pst.active = 0;
  if (n && !(c->format & NNPP_FORMAT_NO_STENCIL) && !c->vol) {
    if (c->pst.state == 0) {
      const int st = build_stencil(c);
      if (st) return st;
    }
    if (c->pst.state == 1) return build_potential_stencil(c);
  }
  int st = slices_alloc(c, c->pot, ns);
  if (st) return st;
  int32_t* rcount = nullptr;
  NNPP_HIP(c, hipMalloc((void**)&rcount, sizeof(int32_t) * (n ? n : 1)));
  NNPP_HIP(c, hipMemsetAsync(c->pot.len, 0, sizeof(int32_t) * (ns ? ns : 1), s));
  if (n) {
    k_pot_flags<<<nnpp_nblocks(n), NNPP_BLOCK, 0, s>>>(n, S, c->row_ptr, c->col, c->phase,
                                                       c->bcmask, c->pfree);
    k_pot_len<<<nnpp_nblocks(n), NNPP_BLOCK, 0, s>>>(n, c->row_ptr, c->col, c->phase, c->pfree,
                                                     c->pot.len, rcount);
  }
  NNPP_HIP(c, hipGetLastError());
  st = slices_offsets(c, c->pot, 4 * NNPP_WAVE);
  if (st) {
    hipFree(rcount);
    return st;
  }
  if (!c->pot.first) NNPP_HIP(c, hipMalloc((void**)&c->pot.first, sizeof(int32_t) * (n ? n : 1)));
  int* dmax = nullptr;
  NNPP_HIP(c, hipMalloc((void**)&dmax, sizeof(int)));
  NNPP_HIP(c, hipMemsetAsync(dmax, 0, sizeof(int), s));
  const int64_t npad = ns * NNPP_WAVE;
  const double expo = 3.0 + 2.0 * c->p.s_phi;
  if (npad)
    k_pot_fill<<<nnpp_nblocks(npad), NNPP_BLOCK, 0, s>>>(
        n, npad, S, c->row_ptr, c->col, c->phase, c->pfree, c->bcmask, c->bcval + (int64_t)S * n,
        c->x, c->y, c->z, weight_args(c, expo), c->pot.off, c->pot.len, c->pot.first,
        c->pot.col, c->pot.d16, c->pot.w64, c->pot.w32, c->pdd, c->pdd32, c->pinv, c->pinv32,
        c->pbdir, c->psc, dmax);
  NNPP_HIP(c, hipGetLastError());
  int hmax = 0;
  NNPP_HIP(c, hipMemcpyAsync(&hmax, dmax, sizeof(int), hipMemcpyDeviceToHost, s));
  // count real entries and free rows
  std::vector<int32_t> rc(n);
  std::vector<uint8_t> fr(n);
  if (n) {
    NNPP_HIP(c, hipMemcpyAsync(rc.data(), rcount, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
    NNPP_HIP(c, hipMemcpyAsync(fr.data(), c->pfree, n, hipMemcpyDeviceToHost, s));
  }
  NNPP_HIP(c, hipStreamSynchronize(s));
  hipFree(rcount);
  hipFree(dmax);
  c->pot.wide = (hmax > 65535 || (c->format & NNPP_FORMAT_WIDE_COLUMNS)) ? 1 : 0;
  c->pot.ncls = 0;
  if (!(c->format & NNPP_FORMAT_NO_DICTIONARY)) {
    st = nnpp_build_dictionary(c, c->pot);
    if (st) return st;
  }
  int64_t nnz = 0, nf = 0;
  for (int32_t i = 0; i < n; i++) {
    nnz += rc[i];
    nf += fr[i];
  }
  c->pot.nnz = nnz;
  c->pot_nfree = nf;
  c->pot_dirty = false;
  return NNPP_OK;
}

int nnpp_front_dev(nnpp_ctx* c, double* U, double thr, int do_vol, int32_t* flips,
                   double* volsum) {
  (void)do_vol;
  const int32_t n = c->n;
  hipStream_t s = c->stream;
  if (n) {
    const int64_t g = nnpp_red_grid(nnpp_nblocks(n));
    k_front<<<g, NNPP_BLOCK, 0, s>>>(n, U, c->phase, thr, c->p.CSat / c->p.CSolid, c->red_part);
    NNPP_HIP(c, hipGetLastError());
    const int st = nnpp_finalize(c, FIN_FRONT, g, 0.0);
    if (st) return st;
    NNPP_HIP(c, hipMemcpyAsync(c->h_scal, c->scal, sizeof(double) * SC_COUNT,
                               hipMemcpyDeviceToHost, s));
  }
  NNPP_HIP(c, hipStreamSynchronize(s));
  const int32_t fl = n ? (int32_t)c->h_scal[SC_R0] : 0;
  if (flips) *flips = fl;
  if (volsum) *volsum = n ? c->h_scal[SC_R1] : 0.0;
  if (n) c->solid_count = (int64_t)c->h_scal[SC_R2];
  if (fl > 0) c->pot_dirty = true;
  return NNPP_OK;
}

int nnpp_bc_override_dev(nnpp_ctx* c, const double* U1, double* Fo) {
  if (!c->n || !c->have_bc) return NNPP_OK;
  k_bc_rows<<<nnpp_nblocks(c->n), NNPP_BLOCK, 0, c->stream>>>(c->n, c->p.n_species + 1, c->bcmask,
                                                              c->bcval, U1, Fo);
  NNPP_HIP(c, hipGetLastError());
  return NNPP_OK;
}

int nnpp_check_finite_dev(nnpp_ctx* c, const double* U, int32_t* bad) {
  const int64_t m = (int64_t)c->n * (c->p.n_species + 1);
  *bad = 0;
  if (!m) return NNPP_OK;
  const int64_t nb = nnpp_red_grid(nnpp_nblocks(m));
  k_nonfinite<<<nb, NNPP_BLOCK, 0, c->stream>>>(m, U, c->red_part);
  NNPP_HIP(c, hipGetLastError());
  const int st = nnpp_finalize(c, FIN_NONFINITE, nb, 0.0);
  if (st) return st;
  NNPP_HIP(c, hipMemcpyAsync(c->h_scal, c->scal, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost,
                             c->stream));
  NNPP_HIP(c, hipStreamSynchronize(c->stream));
  *bad = (int32_t)c->h_scal[SC_R3];
  return NNPP_OK;
}

int nnpp_jac_phi_dev(nnpp_ctx* c, const double* in, double* out, double* diag) {
  if (!c->n) return NNPP_OK;
  k_jac_phi<<<nnpp_nblocks(c->n), NNPP_BLOCK, 0, c->stream>>>(
      c->n, c->p.n_species, c->row_ptr, c->col, c->phase, c->bcmask, c->x, c->y, c->z,
      weight_args(c, 3.0 + 2.0 * c->p.s_phi), in, out, diag);
  NNPP_HIP(c, hipGetLastError());
  return NNPP_OK;
}

int nnpp_potential_apply_dev(nnpp_ctx* c, const double* in, double* out, const int32_t* any) {
  if (!c->n) return NNPP_OK;
  k_pot_apply_csr<<<nnpp_nblocks(c->n), NNPP_BLOCK, 0, c->stream>>>(
      c->n, c->row_ptr, c->col, c->phase, c->x, c->y, c->z,
      weight_args(c, 3.0 + 2.0 * c->p.s_phi), in, out, any);
  NNPP_HIP(c, hipGetLastError());
  return NNPP_OK;
}

int nnpp_finalize(nnpp_ctx* c, int op, int64_t nblk, double a0) {
  k_finalize<<<1, NNPP_BLOCK, 0, c->stream>>>(op, c->red_part, nblk, c->scal, a0);
  NNPP_HIP(c, hipGetLastError());
  return NNPP_OK;
}

int nnpp_build_dictionary(nnpp_ctx* c, nnpp_slices& sl) {
  sl.ncls = 0;
  if (!sl.slots) return NNPP_OK;
  hipStream_t s = c->stream;
  unsigned long long* table = nullptr;
  int* flags = nullptr;  // [0] = count, [1] = fail
  NNPP_HIP(c, hipMalloc((void**)&table, sizeof(unsigned long long) * DICT_HASH));
  NNPP_HIP(c, hipMalloc((void**)&flags, sizeof(int) * 2));
  NNPP_HIP(c, hipMemsetAsync(table, 0, sizeof(unsigned long long) * DICT_HASH, s));
  NNPP_HIP(c, hipMemsetAsync(flags, 0, sizeof(int) * 2, s));
  k_dict_insert<<<1024, NNPP_BLOCK, 0, s>>>(sl.slots, sl.w64, table, flags, flags + 1);
  NNPP_HIP(c, hipGetLastError());
  std::vector<unsigned long long> ht(DICT_HASH);
  int hf[2] = {0, 0};
  NNPP_HIP(c, hipMemcpyAsync(ht.data(), table, sizeof(unsigned long long) * DICT_HASH,
                             hipMemcpyDeviceToHost, s));
  NNPP_HIP(c, hipMemcpyAsync(hf, flags, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
  NNPP_HIP(c, hipStreamSynchronize(s));
  hipFree(table);
  hipFree(flags);
  if (hf[1] || hf[0] > NNPP_DICT_MAX - 1) return NNPP_OK;  // not a small set: keep FP weights
  std::vector<double> tab(1, 0.0);
  for (unsigned long long k : ht) {
    if (!k) continue;
    double v;
    std::memcpy(&v, &k, sizeof(v));
    tab.push_back(v);
  }
  std::sort(tab.begin(), tab.end());
  const int ntab = (int)tab.size();
  if (ntab > NNPP_DICT_MAX || tab[1] <= 0.0) return NNPP_OK;  // weights are positive
  std::vector<float> tab32(ntab);
  for (int k = 0; k < ntab; k++) tab32[k] = (float)tab[k];
  if (!sl.tab64) NNPP_HIP(c, hipMalloc((void**)&sl.tab64, sizeof(double) * NNPP_DICT_MAX));
  if (!sl.tab32) NNPP_HIP(c, hipMalloc((void**)&sl.tab32, sizeof(float) * NNPP_DICT_MAX));
  if (!sl.cls) NNPP_HIP(c, hipMalloc((void**)&sl.cls, sl.cap_slots));
  NNPP_HIP(c, hipMemsetAsync(sl.tab64, 0, sizeof(double) * NNPP_DICT_MAX, s));
  NNPP_HIP(c, hipMemsetAsync(sl.tab32, 0, sizeof(float) * NNPP_DICT_MAX, s));
  NNPP_HIP(c, hipMemcpyAsync(sl.tab64, tab.data(), sizeof(double) * ntab, hipMemcpyHostToDevice, s));
  NNPP_HIP(c, hipMemcpyAsync(sl.tab32, tab32.data(), sizeof(float) * ntab, hipMemcpyHostToDevice, s));
  k_dict_map<<<2048, NNPP_BLOCK, 0, s>>>(sl.slots, sl.w64, sl.tab64, ntab, sl.cls);
  NNPP_HIP(c, hipGetLastError());
  NNPP_HIP(c, hipStreamSynchronize(s));
  sl.ncls = ntab;
  return NNPP_OK;
}

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
// direction update). The scalars stay on the device: each kernel that needs alpha or beta sums
// the previous kernel's block partials itself (every block in the same order, so all agree; no
// one-block reduction launch in between), and a converged flag makes the remaining launches of
// a batch exit at once, so the host synchronises once per batch.
// MIXED: CG with reliable updates. The iteration runs in FP32 (FP32 weights and vectors,
// Kahan-compensated bond sums, or on box lattices the ring sums of the z-marching walk; FP64
// dot products) on the residual normalised by ||b||; on box lattices in the single-reduction
// form, one fused kernel per iteration. The
// solution is kept in FP64. Whenever the FP32 recursive residual has dropped by inner_rtol (1e-3)
// (or reached the target), the FP32 correction is added to the FP64 solution, the residual is
// recomputed in FP64 (FP64 weights and vectors) and replaces the FP32 one, and the iteration
// continues with the same search direction. The FP64 residual is the stopping test.
//
// Every kernel with a reduction runs on at most NNPP_RED_GRID blocks that stride over the node
// blocks (nnpp_blocks_of), so the last-block sum handles at most 2048 partials.
#include <algorithm>
#include <cmath>

#include <cstring>
#include <type_traits>

#include "nnpp_internal.hpp"

namespace {

constexpr int CG_BATCH = 32;

struct Slices {
  const int64_t* off;
  const int32_t* len;  // groups of 4 entries per lane
  const int32_t* col;
  const uint16_t* d16;
  const int32_t* first;
  const uint8_t* cls;   // weight classes (dictionary variant)
  const double* tab64;  // dictionary, FP64 / FP32
  const float* tab32;
  // stencil variant
  const uint4* pmask;  // bit t: entry t of the row present
  const int32_t* lc;   // LDS position of entry t relative to the lane (padded to 8 entries)
  const int32_t* seg_start;  // staging segments: first offset, span, LDS base
  const int32_t* seg_len;
  const int32_t* seg_base;
  const int32_t* piece;  // staging pieces: vector position - r0, LDS position, count
  const double* tw64;  // weight of entry t (padded with 0 to 8 entries)
  const float* tw32;
  const uint8_t* perm;  // row of each lane within its tile
  int K, R, P;
  int off_min, off_max;  // smallest and largest offset
  nnpp_ball_args ball;   // the lattice-ball form's line positions and class weights
};

template <typename T>
__device__ __forceinline__ T ball_weight(const Slices& sl, int d2);
template <>
__device__ __forceinline__ float ball_weight<float>(const Slices& sl, int d2) {
  return sl.ball.w32[d2];
}
template <>
__device__ __forceinline__ double ball_weight<double>(const Slices& sl, int d2) {
  return sl.ball.w64[d2];
}

template <typename T>
__device__ __forceinline__ const T* stencil_weights(const Slices& sl);
template <>
__device__ __forceinline__ const double* stencil_weights<double>(const Slices& sl) {
  return sl.tw64;
}
template <>
__device__ __forceinline__ const float* stencil_weights<float>(const Slices& sl) {
  return sl.tw32;
}

__device__ __forceinline__ bool cg_done(const double* scal) {
  return nnpp_ld_scalar(scal + SC_DONE) != 0.0;
}

// Stream words of one quad (4 consecutive entries of a row): the columns (four int32, or four
// 16-bit steps from the previous column) and the weights (FP64/FP32 values, or four 1-byte
// dictionary classes). The FP32 variants load them non-temporally so the gathered vector keeps
// its place in L2; for the FP64 variants plain loads measure faster.
template <typename T, bool WIDE, bool CLS>
struct Quad {
  static constexpr bool NT = sizeof(T) == 4;
  int4 col;
  uint2 d;
  uint32_t k;
  T w[4];

  __device__ __forceinline__ void load(const Slices& sl, const T* __restrict__ wp, int64_t s0) {
    if constexpr (WIDE) {
      const int4* p = reinterpret_cast<const int4*>(sl.col + s0);
      col = NT ? nnpp_ld_stream(p) : *p;
    } else {
      const uint2* p = reinterpret_cast<const uint2*>(sl.d16 + s0);
      d = NT ? nnpp_ld_stream(p) : *p;
    }
    if constexpr (CLS) {
      const uint32_t* p = reinterpret_cast<const uint32_t*>(sl.cls + s0);
      k = NT ? __builtin_nontemporal_load(p) : *p;
    } else if constexpr (NT) {
      const float4 t = nnpp_ld_stream(reinterpret_cast<const float4*>(wp + s0));
      w[0] = t.x;
      w[1] = t.y;
      w[2] = t.z;
      w[3] = t.w;
    } else {
      const double2 a = *reinterpret_cast<const double2*>(wp + s0);
      const double2 b = *reinterpret_cast<const double2*>(wp + s0 + 2);
      w[0] = a.x;
      w[1] = a.y;
      w[2] = b.x;
      w[3] = b.y;
    }
  }

  // columns, advancing the running column j in the step format
  __device__ __forceinline__ void cols(int32_t& j, int32_t (&c)[4]) const {
    if constexpr (WIDE) {
      c[0] = col.x;
      c[1] = col.y;
      c[2] = col.z;
      c[3] = col.w;
    } else {
      c[0] = j + (int32_t)(d.x & 0xffffu);
      c[1] = c[0] + (int32_t)(d.x >> 16);
      c[2] = c[1] + (int32_t)(d.y & 0xffffu);
      c[3] = c[2] + (int32_t)(d.y >> 16);
      j = c[3];
    }
  }

  __device__ __forceinline__ void weights(const T* tab, T (&wq)[4]) const {
    if constexpr (CLS) {
      wq[0] = tab[k & 0xffu];
      wq[1] = tab[(k >> 8) & 0xffu];
      wq[2] = tab[(k >> 16) & 0xffu];
      wq[3] = tab[k >> 24];
    } else {
      wq[0] = w[0];
      wq[1] = w[1];
      wq[2] = w[2];
      wq[3] = w[3];
    }
  }
};

// Dictionary table copied to LDS at kernel start (all threads of the block take part).
template <typename T>
__device__ __forceinline__ void load_tab(T* lds, const T* __restrict__ tab) {
  for (int k = threadIdx.x; k < NNPP_DICT_MAX; k += NNPP_BLOCK) lds[k] = tab[k];
  __syncthreads();
}

// Per-row accumulator of the terms w (v_i - v_j): plain FP64 sum, or Kahan-compensated FP32
// sum. The fused multiply-adds are written out, so every storage form of the operator (and its
// row loop) rounds the same way whatever the compiler would contract on its own.
template <typename T>
struct RowAcc;
template <>
struct RowAcc<double> {
  double s = 0.0;
  __device__ __forceinline__ void add(double w, double d) { s = __fma_rn(w, d, s); }
  __device__ __forceinline__ double result(double dd, double vi) const {
    return __fma_rn(dd, vi, s);
  }
};
template <>
struct RowAcc<float> {
  float s = 0.0f, c = 0.0f;
  __device__ __forceinline__ void add(float w, float d) {
    const float y = __fmaf_rn(w, d, -c);
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ float result(float dd, float vi) const {
    return __fmaf_rn(dd, vi, s);
  }
};

// Quads per batch: with stored weights (6-12 bytes per entry) deep batches keep enough bytes in
// flight; with dictionary classes (3 bytes per entry) small batches keep occupancy and measure
// faster (MI355X, 128^3 lattice and jittered screw cases).
template <typename T, bool CLS>
constexpr int spmv_batch() {
  return !CLS ? 8 : (sizeof(T) == 4 ? 2 : 1);
}

// y_i = dd_i v_i + sum_t w_t (v_i - v_{c_t}) over the row's slice entries (padding weighs 0).
// Quads are taken in batches: all stream words of a batch are loaded first, then all its
// gathers are issued, then summed, so each wave keeps a batch's worth of loads in flight per
// memory round trip. G is uniform over the wave, so the batch tail is a uniform branch; the
// tail's loads are clamped to the row's last quad and not summed.
template <typename T, bool WIDE, bool CLS>
__device__ __forceinline__ T spmv_row(const Slices& sl, const T* __restrict__ w, const T* tab,
                                      const T* __restrict__ dd, const T* __restrict__ v,
                                      int64_t i) {
  constexpr int B = spmv_batch<T, CLS>();
  const int64_t base = sl.off[i / NNPP_WAVE] + 4 * (i % NNPP_WAVE);
  const int G = sl.len[i / NNPP_WAVE];
  const T vi = v[i];
  int32_t j = WIDE ? 0 : sl.first[i];
  RowAcc<T> acc;
  for (int g = 0; g < G; g += B) {
    Quad<T, WIDE, CLS> q[B];
#pragma unroll
    for (int b = 0; b < B; b++)
      q[b].load(sl, w, base + 256 * (int64_t)(g + b < G ? g + b : G - 1));
    int32_t c[B][4];
#pragma unroll
    for (int b = 0; b < B; b++) {
      if (g + b < G) q[b].cols(j, c[b]);
      else c[b][0] = c[b][1] = c[b][2] = c[b][3] = j;
    }
    T pv[B][4];
#pragma unroll
    for (int b = 0; b < B; b++)
#pragma unroll
      for (int e = 0; e < 4; e++) pv[b][e] = v[c[b][e]];
#pragma unroll
    for (int b = 0; b < B; b++) {
      if (g + b >= G) break;
      T wq[4];
      q[b].weights(tab, wq);
#pragma unroll
      for (int e = 0; e < 4; e++) acc.add(wq[e], vi - pv[b][e]);
    }
  }
  return acc.result(dd[i], vi);
}

// Stencil variant: one 256-row tile. Every segment of offsets a..b (offsets whose ranges overlap)
// needs the contiguous range v[r0 + a .. r0 + b + 255], staged into LDS with coalesced loads in
// pieces of up to 256 values; lane l then reads entry t at lds[l + lc[t]]. Entries go in chunks
// (32 in FP32, 16 in FP64): the chunk's LDS positions and weights are wave-uniform (scalar
// loads), its LDS reads are issued together, and an absent entry
// takes weight +0 (its bits cleared), which leaves the sum unchanged, so no lane branches.
// FP64: one fused multiply-add chain in ascending offset order, the additions of the column
// form (bit for bit). FP32: two fused multiply-add chains per chunk (even and odd entries),
// the chunk sums Kahan-compensated. The leading barrier orders the staging after the previous tile's reads;
// tiles without any entry skip the staging. All threads of the block call this.
// ORDERED: the result returned to lane j is row r0 + j's (the column form's lane order, which
// the FP64 dot products need to match it bit for bit); otherwise lane j gets the result of the
// row it computed, *row_out.
// The chunk loop of stencil_tile for the radius-M lattice ball (M > 0): the same chunks, masks
// and order of additions as the table-driven loop, so the same sums bit for bit (that loop's
// padding entries add w = +0 terms to partial sums that are never -0, which changes nothing),
// with the entries' layout known at compile time: one LDS address per line, the line's entries
// read at immediate offsets from it, and the weights from the per-class kernel arguments, so no
// table loads and no per-entry address arithmetic.
template <typename T, int M>
__device__ __forceinline__ void ball_rows(const Slices& sl, const T* row, const uint4 m,
                                          const T vi, RowAcc<T>& acc) {
  constexpr nnpp_ball<M> B{};
  constexpr int K = nnpp_ball<M>::K;
  constexpr int BT = sizeof(T) == 4 ? 32 : 16;
#pragma unroll
  for (int t0 = 0; t0 < K; t0 += BT) {
    const int wd = t0 >> 5;
    const uint32_t bits =
        (wd == 0 ? m.x : (wd == 1 ? m.y : (wd == 2 ? m.z : m.w))) >> (t0 & 31);
    const int nk = K - t0 < BT ? K - t0 : BT;
    const uint32_t valid = nk >= 32 ? 0xffffffffu : (1u << nk) - 1u;
    if (!__any(bits & valid)) continue;
    const bool full = __all((bits & valid) == valid);
    T pv[BT];
#pragma unroll
    for (int k = 0; k < BT; k++)
      if (k < nk) pv[k] = row[sl.ball.lc[B.line[t0 + k]] + B.pos[t0 + k]];
    T part[2] = {T(0), T(0)};
    if (full) {
#pragma unroll
      for (int k = 0; k < BT; k++) {
        if (k >= nk) break;
        const T w = ball_weight<T>(sl, B.d2[t0 + k]);
        if constexpr (sizeof(T) == 8) acc.add(w, vi - pv[k]);
        else part[k & 1] = __fmaf_rn(w, vi - pv[k], part[k & 1]);
      }
    } else {
#pragma unroll
      for (int k = 0; k < BT; k++) {
        if (k >= nk) break;
        const int keep = (int)(bits << (31 - k)) >> 31;
        const T wk = nnpp_mask_weight<T>(ball_weight<T>(sl, B.d2[t0 + k]), keep);
        if constexpr (sizeof(T) == 8) acc.add(wk, vi - pv[k]);
        else part[k & 1] = __fmaf_rn(wk, vi - pv[k], part[k & 1]);
      }
    }
    if constexpr (sizeof(T) == 4) acc.add(T(1), part[0] + part[1]);
  }
}

template <typename T, bool ORDERED, int M = 0>
__device__ __forceinline__ T stencil_tile(const Slices& sl, const T* __restrict__ dd,
                                          const T* __restrict__ v, int32_t n, int64_t tile,
                                          T* lds, int64_t* row_out) {
  const int64_t r0 = tile * NNPP_BLOCK;
  const int l = sl.perm[r0 + threadIdx.x];  // this lane's row within the tile
  const int64_t i = r0 + l;
  const bool in = i < n;
  const uint4 m = in ? sl.pmask[i] : make_uint4(0u, 0u, 0u, 0u);
  const uint32_t any = m.x | m.y | m.z | m.w;
  const T vi = in ? v[i] : T(0);
  if (__syncthreads_or(any != 0u)) {
    // tiles whose every staged index lies in [0, n) need no bounds checks (all but the first
    // and last few); a piece's source and destination are wave-uniform
    const bool inner = r0 + sl.off_min >= 0 && r0 + NNPP_BLOCK - 1 + sl.off_max < n;
    if (inner) {
      // the loads of 8 pieces are issued together, then stored: one memory latency per 8
      // pieces. The batch's table words are read first, unconditionally (the table is padded
      // to whole batches with empty pieces), so they arrive in one scalar round trip instead
      // of two dependent ones per piece.
      constexpr int PB = NNPP_PIECE_BATCH;
      const int P = sl.P;
      for (int pb = 0; pb < P; pb += PB) {
        int32_t pt[3 * PB];
#pragma unroll
        for (int k = 0; k < 3 * PB; k++) pt[k] = nnpp_ld_table(sl.piece, 3 * pb + k);
        T val[PB];
#pragma unroll
        for (int k = 0; k < PB; k++)
          if ((int)threadIdx.x < pt[3 * k + 2]) val[k] = v[r0 + pt[3 * k] + threadIdx.x];
#pragma unroll
        for (int k = 0; k < PB; k++)
          if ((int)threadIdx.x < pt[3 * k + 2]) lds[pt[3 * k + 1] + threadIdx.x] = val[k];
      }
    } else {
      const int R = sl.R;
      for (int r = 0; r < R; r++) {
        const int64_t g0 = r0 + nnpp_ld_table(sl.seg_start, r);
        const int len = NNPP_BLOCK + nnpp_ld_table(sl.seg_len, r) - 1;
        T* dst = lds + nnpp_ld_table(sl.seg_base, r);
        for (int u = threadIdx.x; u < len; u += NNPP_BLOCK) {
          const int64_t g = g0 + u;
          dst[u] = (g >= 0 && g < n) ? v[g] : T(0);
        }
      }
    }
    __syncthreads();
  }
  // entries per chunk: 32 in FP32 (half the wave-level round trips of 16, measured 16% faster
  // on a 128^3 lattice), 16 in FP64 (32 costs occupancy there)
  constexpr int BT = sizeof(T) == 4 ? 32 : 16;
  static_assert(NNPP_STENCIL_CHUNK % BT == 0, "tables are padded to whole chunks");
  const T* __restrict__ tw = stencil_weights<T>(sl);
  const T* row = lds + l;
  RowAcc<T> acc;
  const int K = sl.K;
  if constexpr (M > 0) {
    if (__any(any != 0u)) ball_rows<T, M>(sl, row, m, vi, acc);
  } else if (__any(any != 0u)) {
    for (int t0 = 0; t0 < K; t0 += BT) {  // a chunk never straddles a 32-bit mask word
      const int wd = t0 >> 5;
      const uint32_t bits =
          (wd == 0 ? m.x : (wd == 1 ? m.y : (wd == 2 ? m.z : m.w))) >> (t0 & 31);
      constexpr uint32_t all = BT >= 32 ? 0xffffffffu : ((1u << (BT & 31)) - 1u);
      const uint32_t valid = K - t0 >= BT ? all : (1u << (K - t0)) - 1u;
      if (!__any(bits & valid)) continue;  // no lane of the wave has an entry in this chunk
      const bool full = __all((bits & valid) == valid);
      int lv[BT];
      T wv[BT], pv[BT];
#pragma unroll
      for (int k = 0; k < BT; k++) {  // tables padded to a multiple of BT entries (weight 0)
        lv[k] = nnpp_ld_table(sl.lc, t0 + k);
        wv[k] = nnpp_ld_table(tw, t0 + k);
      }
#pragma unroll
      for (int k = 0; k < BT; k++) pv[k] = row[lv[k]];
      T part[2] = {T(0), T(0)};  // FP32: even and odd entries, two short dependency chains
      if (full) {  // every lane has every entry of the chunk
#pragma unroll
        for (int k = 0; k < BT; k++) {
          if constexpr (sizeof(T) == 8) acc.add(wv[k], vi - pv[k]);
          else part[k & 1] = __fmaf_rn(wv[k], vi - pv[k], part[k & 1]);
        }
      } else {
#pragma unroll
        for (int k = 0; k < BT; k++) {
          const int keep = (int)(bits << (31 - k)) >> 31;  // bit k -> all ones, else 0
          const T wk = nnpp_mask_weight<T>(wv[k], keep);
          if constexpr (sizeof(T) == 8) acc.add(wk, vi - pv[k]);
          else part[k & 1] = __fmaf_rn(wk, vi - pv[k], part[k & 1]);
        }
      }
      if constexpr (sizeof(T) == 4) acc.add(T(1), part[0] + part[1]);
    }
  }
  const T y = in ? acc.result(dd[i], vi) : T(0);
  if constexpr (!ORDERED) {
    *row_out = i;
    return y;
  } else {
    // hand the results back in row order (lane j of the caller owns row r0 + j), through the
    // start of the staging area once every lane is done reading it
    __syncthreads();
    lds[l] = y;
    __syncthreads();
    *row_out = r0 + threadIdx.x;
    return lds[threadIdx.x];
  }
}

// ---------------------------------------------------------------- z-marching form -------
// Box lattices (the radius-3 ball on the full nx x ny x nz box, numbered x fastest): a block
// owns a ZM_TX x ZM_TY column tile of (x, y) and a chunk of lz output planes along z. It walks
// the input planes z0-3 .. z0+lz+2; each plane's tile with a 3-node halo is staged into LDS
// once (FP64: as (p_j, m_j) pairs with m_j = 1 on free nodes and 0 elsewhere, outside the box
// too). A thread reads the 29-point disc dx^2 + dy^2 <= 9 around its column from LDS once per
// plane and feeds the (up to 7) output rows of its column that are in flight, held in
// registers: output z_o takes the bonds with dz = z_i - z_o from input plane z_i. Per row this
// is 29 LDS reads per plane instead of one per bond.
//
// The entries of row i are the bonds to free nodes inside the box (the potential operator's
// free-free entries; its rows are the free rows): an absent entry takes weight w * m_j = +0,
// exactly the masked weight of the stencil form. FP64 (zm_walk): every row is one fused
// multiply-add chain over its bonds in ascending offset order (planes in z order, then dy,
// then dx), the additions of the stencil and column forms, bit for bit. FP32 (zm_walk_ring):
// the bonds by rings of equal in-plane distance, one FP32 sum per row.
// 16-byte staging (VEC, nx a multiple of 4): each staged row of the plane starts at x0 - 4 and
// is ZM_W4 wide, whole quads of 4 nodes, one quad per thread
constexpr int ZM_X4 = 4;  // the staged row's first column relative to x0 (a whole quad)
constexpr int ZM_W4 = ZM_TX + 2 * ZM_X4;
constexpr int ZM_Q4 = ZM_W4 / 4 * ZM_H;  // quads per plane
static_assert(ZM_X4 >= ZM_M && ZM_Q4 <= NNPP_BLOCK, "one staged quad per thread");

// Loads and stores at a 32-bit byte offset from the array's start (node indices times the
// element size stay below 2^32 on the z-marching geometries), so the address is a scalar base
// plus a vector offset.
template <typename T>
__device__ __forceinline__ T zm_ld(const T* __restrict__ p, uint32_t j) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + j * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void zm_st(T* __restrict__ p, uint32_t j, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + j * (uint32_t)sizeof(T)) = v;
}


template <typename T>
struct ZmPair {
  T v, m;
};

template <typename T>
struct ZmRow;  // per in-flight row of the FP64 walk: its running sum
template <>
struct ZmRow<double> {
  double s;
  __device__ __forceinline__ void reset() { s = 0.0; }
  __device__ __forceinline__ void fold(double part) { s = part; }  // FP64: one chain
  __device__ __forceinline__ double start() const { return s; }
  __device__ __forceinline__ double result(double dd, double vi) const {
    return __fma_rn(dd, vi, s);
  }
};

template <typename T>
__device__ __forceinline__ T zm_weight(const nnpp_ball_args& b, int d2);
template <>
__device__ __forceinline__ float zm_weight<float>(const nnpp_ball_args& b, int d2) {
  return b.w32[d2];
}
template <>
__device__ __forceinline__ double zm_weight<double>(const nnpp_ball_args& b, int d2) {
  return b.w64[d2];
}

// Input vector of the z-marching walk: load(j) issues the loads of node j's value into a Raw,
// value(raw) forms the value (zero from Raw{}); the loads are issued a plane ahead of use.
template <typename T>
struct ZmLoad {  // a stored vector
  using Raw = T;
  const T* __restrict__ v;
  __device__ __forceinline__ Raw load(uint32_t j) const { return v[j]; }
  __device__ __forceinline__ T value(const Raw& r) const { return r; }
};
// Row handler without prefetched operands: f(i, y, p_i).
template <typename F>
struct ZmEmit {
  struct Pre {};
  F f;
  __device__ __forceinline__ Pre pre(uint32_t) const { return Pre{}; }
  template <typename T>
  __device__ __forceinline__ void operator()(int64_t i, T y, T pi, const Pre&) { f(i, y, pi); }
};
template <typename F>
__device__ __forceinline__ ZmEmit<F> zm_emit(F f) {
  return ZmEmit<F>{f};
}

// Walks the block's work item and hands every finished row to emit(i, y, p_i, pre) (rows inside
// the box only), pre = emit.pre(i) loaded a plane ahead. All threads of the block must call it
// (barriers inside). full[tile * nz + z]: every staged point of plane z of the tile is a free
// node (no weight needs its m_j). The FP64 operator: each row one fused multiply-add chain over
// its bonds in ascending offset order, the additions of the other forms.
template <typename T, typename Src, typename Emit>
__device__ __forceinline__ void zm_walk(const ZmGeom& g, const nnpp_ball_args& ball,
                                        const uint8_t* __restrict__ pfree,
                                        const uint8_t* __restrict__ full_tab,
                                        const T* __restrict__ dd, const Src& src, Emit& emit) {
  using Raw = typename Src::Raw;
  using Pre = typename Emit::Pre;
  __shared__ ZmPair<T> lds[2][ZM_PTS];
  int64_t tile;
  int zc;
  zm_item(g, blockIdx.x, tile, zc);
  const int x0 = (int)(tile % g.tiles_x) * ZM_TX, y0 = (int)(tile / g.tiles_x) * ZM_TY;
  const int z0 = zc * g.lz;
  const int zend = min(z0 + g.lz, g.nz);
  const int tx = threadIdx.x % ZM_TX, ty = threadIdx.x / ZM_TX;
  const int x = x0 + tx, y = y0 + ty;
  const bool col_ok = x < g.nx && y < g.ny;
  // this thread's staged points: in-plane offset (-1: outside the box or past the tile)
  int32_t soff[ZM_STAGE];
#pragma unroll
  for (int k = 0; k < ZM_STAGE; k++) {
    const int e = threadIdx.x + k * NNPP_BLOCK;
    const int gx = x0 + e % ZM_W - ZM_M, gy = y0 + e / ZM_W - ZM_M;
    soff[k] = (e < ZM_PTS && gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny) ? gy * g.nx + gx : -1;
  }
  const uint8_t* __restrict__ ftab = full_tab + tile * g.nz;
  const int zs = max(z0 - ZM_M, 0), ze = min(zend + ZM_M, g.nz);  // input planes [zs, ze)
  // row (x, y, zo) is node zo * sz + col (node indices fit 32 bits: n is an int32_t)
  const uint32_t col = col_ok ? (uint32_t)(y * g.nx + x) : 0u;
  const uint32_t sz = (uint32_t)g.sz;
  // slot q holds output plane zi - M + q while input plane zi is processed; rows outside the
  // chunk are never summed (uniform tests) nor emitted
  ZmRow<T> row[ZM_SL];
  T po[ZM_SL];
  const auto in_chunk = [&](int zo) { return zo >= z0 && zo < zend; };
  const auto row_raw = [&](int zo) {
    return (col_ok && in_chunk(zo)) ? src.load((uint32_t)zo * sz + col) : Raw{};
  };
#pragma unroll
  for (int q = 0; q < ZM_SL; q++) {
    row[q].reset();
    po[q] = src.value(row_raw(zs - ZM_M + q));
  }
  // the input value of the row that enters the last slot at the next plane, loaded a plane ahead
  Raw pend = row_raw(zs + 1 + ZM_M);
  // diagonal term, free flag and handler operands of the row in slot 0 (the next to finish),
  // loaded a plane ahead
  T d0 = T(0);
  bool f0 = false;
  Pre pre0{};
  const auto load_first = [&](int zo) {
    if (col_ok && in_chunk(zo)) {
      const uint32_t i = (uint32_t)zo * sz + col;
      d0 = zm_ld(dd, i);
      f0 = zm_ld(pfree, i) != 0;
      pre0 = emit.pre(i);
    }
  };
  load_first(zs - ZM_M);
  // loads of the next plane's staged points
  Raw sv[ZM_STAGE];
  uint8_t sm[ZM_STAGE];
  const auto fetch = [&](int zi) {
    const uint32_t base = (uint32_t)zi * sz;
#pragma unroll
    for (int k = 0; k < ZM_STAGE; k++) {
      sv[k] = soff[k] >= 0 ? src.load(base + (uint32_t)soff[k]) : Raw{};
      sm[k] = soff[k] >= 0 ? pfree[base + (uint32_t)soff[k]] : (uint8_t)0;
    }
  };
  fetch(zs);
  // after input plane zi: output plane zi - M is complete; the slots move down by one and the
  // last one takes output plane zi + 1 + M
  const auto finish = [&](int zi) {
    const int zo = zi - ZM_M;
    if (col_ok && in_chunk(zo)) {
      const T y = f0 ? row[0].result(d0, po[0]) : ZmRow<T>{}.result(d0, po[0]);
      emit((int64_t)((uint32_t)zo * sz + col), y, po[0], pre0);
    }
#pragma unroll
    for (int q = 0; q + 1 < ZM_SL; q++) {
      row[q] = row[q + 1];
      po[q] = po[q + 1];
    }
    row[ZM_SL - 1].reset();
    po[ZM_SL - 1] = src.value(pend);
    pend = row_raw(zi + 2 + ZM_M);
    load_first(zo + 1);
  };
  int buf = 0;
  for (int zi = zs; zi < ze; zi++) {
#pragma unroll
    for (int k = 0; k < ZM_STAGE; k++) {
      const int e = threadIdx.x + k * NNPP_BLOCK;
      if (e < ZM_PTS) {
        ZmPair<T> pr;
        pr.v = src.value(sv[k]);
        pr.m = sm[k] ? T(1) : T(0);
        lds[buf][e] = pr;
      }
    }
    const bool full = nnpp_ld_table(ftab, zi) != 0;
    // this plane staged (and the buffer written two planes ago read by every thread)
    __syncthreads();
    if (zi + 1 < ze) fetch(zi + 1);
    const ZmPair<T>* __restrict__ me = &lds[buf][(ty + ZM_M) * ZM_W + tx + ZM_M];
    // slot q (output plane zi - M + q) takes the bonds with dz = M - q
    bool act[ZM_SL];
#pragma unroll
    for (int q = 0; q < ZM_SL; q++) act[q] = in_chunk(zi - ZM_M + q);
    // MASK: some staged point of the plane is not a free node (weights times m_j). Separate
    // instances, so the common case (free interior planes) is one subtraction and one fused
    // multiply-add per bond.
    const auto body = [&](auto mask_c) {
      constexpr bool MASK = decltype(mask_c)::value;
      T part[ZM_SL];
#pragma unroll
      for (int q = 0; q < ZM_SL; q++) part[q] = row[q].start();
#pragma unroll
      for (int dy = -ZM_M; dy <= ZM_M; dy++) {
        ZmPair<T> lv[2 * ZM_M + 1];
#pragma unroll
        for (int dx = -ZM_M; dx <= ZM_M; dx++)
          if (dx * dx + dy * dy <= ZM_M * ZM_M) lv[dx + ZM_M] = me[dy * ZM_W + dx];
#pragma unroll
        for (int q = 0; q < ZM_SL; q++) {
          const int dz = ZM_M - q;
          if (dz * dz + dy * dy > ZM_M * ZM_M) continue;
          if (!act[q]) continue;  // uniform over the block
#pragma unroll
          for (int dx = -ZM_M; dx <= ZM_M; dx++) {
            const int d2 = dx * dx + dy * dy + dz * dz;
            if (d2 > ZM_M * ZM_M || d2 == 0) continue;
            const T w = zm_weight<T>(ball, d2);
            const ZmPair<T> pj = lv[dx + ZM_M];
            if constexpr (sizeof(T) == 8)
              part[q] = __fma_rn(MASK ? w * pj.m : w, po[q] - pj.v, part[q]);
            else
              part[q] = __fmaf_rn(MASK ? w * pj.m : w, po[q] - pj.v, part[q]);
          }
        }
      }
#pragma unroll
      for (int q = 0; q < ZM_SL; q++)
        if (act[q]) row[q].fold(part[q]);
    };
    // (an instance without the per-slot tests for the planes where every slot is active
    // measured slower: its straight-line code spilled at four waves per SIMD)
    if (full) body(std::false_type{});
    else body(std::true_type{});
    finish(zi);
    buf ^= 1;
  }
  // planes past the top of the box add nothing: the rows still in flight are complete
  for (int zi = ze; zi < zend + ZM_M; zi++) finish(zi);
}


// Emit::RawTag true: the handler takes the row's own operands (Src::Raw) from the walk's
// staging instead of loading them (emit(i, y, p_i, pre, raw)).
template <typename E, typename = void>
struct ZmEmitRaw : std::false_type {};
template <typename E>
struct ZmEmitRaw<E, std::void_t<typename E::RawTag>> : E::RawTag {};
// planes whose tile operands the walk keeps for the rows (a row completes 3 planes after its
// own; a slot is written again 5 planes later, two barriers after its last read)
constexpr int ZM_RR = 5;

template <typename Src, bool VEC>
struct ZmRaw4 {
  using type = typename Src::Raw;
};
template <typename Src>
struct ZmRaw4<Src, true> {
  using type = typename Src::Raw4;
};

// The FP32 walk of the CG vectors (the unmasked operator, rings). The input vector is 0 at
// every node that is not free (the CG vectors D^-1 r and their combinations, or a vector masked
// as it is loaded), and dd is the diagonal less the weights of the row's bonds to non-free nodes
// and to in-plane positions outside the box (pdd32nf): every bond then takes its weight unmasked
// (the bonds to those nodes add w (p_i - 0), which the diagonal takes back) and only the values
// are staged. Rows that are not free give dd p_i = 0.
//
// Bonds by rings: around the thread's column (value p_c in input plane z_i) the disc
// dx^2 + dy^2 <= M^2 splits into rings of equal k = dx^2 + dy^2, and the weight of a bond
// depends on k + dz^2 only, so with R_k = sum over ring k of (p_j - p_c) and W_dz the weight sum
// of the disc at plane distance dz, the row of output plane z_o takes from input plane z_i
//   sum over the plane's bonds of w (p_o - p_j) = W_dz (p_o - p_c) - sum_k w(k + dz^2) R_k,
// dz = z_i - z_o. The ring sums are formed once per plane and shared by the rows in flight. The
// ring terms are added to the row's sum as the planes pass; the first terms, over the planes,
// are the bonds along the row's own z-line, sum over a = 1..M of
//   W_a ((p_o - p_{o+a}) + (p_o - p_{o-a}))   (planes inside the box),
// taken when the row completes from the column values of the walked planes. The ring sums are
// symmetric sums of local differences and each z-line pair a symmetric difference, so every
// term is of the size of the row's result (second differences of p), and the row is one FP32
// sum without compensation. Per plane and thread: 28 subtractions and 22 additions for the
// rings, at most 6 multiply-adds per row in flight; against a subtraction and a multiply-add
// per bond (122).
//
// VEC (nx a multiple of 4): the plane is staged in 16-byte quads (Src::load4 / value4), the
// quads of the next two planes in flight; otherwise single values (Src::load / value), the
// next plane's in flight. With Emit::RawTag (VEC), the staged operand quads of the tile's own
// columns also go to an LDS ring of ZM_RR planes, from which each row takes its own operands
// when it completes: they are not loaded again (by then L2 has dropped most of them).
template <bool VEC, typename Src, typename Emit>
__device__ __forceinline__ void zm_walk_ring(const ZmGeom& g, const nnpp_ball_args& ball,
                                             const uint8_t* __restrict__ pfree,
                                             const float* __restrict__ dd, const Src& src,
                                             Emit& emit) {
  using Raw = typename Src::Raw;
  using Pre = typename Emit::Pre;
  constexpr int LW = VEC ? ZM_W4 : ZM_W;  // staged row width in LDS
  constexpr int LX = VEC ? ZM_X4 : ZM_M;  // the thread's column within the staged row
  __shared__ __align__(16) float ldsv[2][VEC ? LW * ZM_H : ZM_PTS];
  constexpr bool RAWROW = VEC && ZmEmitRaw<Emit>::value;
  // [plane slot][operand][tile column ty * ZM_TX + tx]: the rows' own operands
  __shared__ __align__(16) float lraw[RAWROW ? ZM_RR : 1][4][RAWROW ? NNPP_BLOCK : 4];
  int64_t tile;
  int zc;
  zm_item(g, blockIdx.x, tile, zc);
  const int x0 = (int)(tile % g.tiles_x) * ZM_TX, y0 = (int)(tile / g.tiles_x) * ZM_TY;
  const int z0 = zc * g.lz;
  const int zend = min(z0 + g.lz, g.nz);
  const int tx = threadIdx.x % ZM_TX, ty = threadIdx.x / ZM_TX;
  const int x = x0 + tx, y = y0 + ty;
  const bool col_ok = x < g.nx && y < g.ny;
  // this thread's staged points: in-plane offset (-1: outside the box or past the tile); VEC:
  // the first node of its quad (a quad lies wholly inside or outside the box: nx % 4 == 0)
  constexpr int NST = VEC ? 1 : ZM_STAGE;
  int32_t soff[NST];
#pragma unroll
  for (int k = 0; k < NST; k++) {
    if (VEC) {
      const int e = threadIdx.x;
      const int gx = x0 + (e % (ZM_W4 / 4)) * 4 - ZM_X4, gy = y0 + e / (ZM_W4 / 4) - ZM_M;
      soff[k] = (e < ZM_Q4 && gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny) ? gy * g.nx + gx
                                                                             : -1;
    } else {
      const int e = threadIdx.x + k * NNPP_BLOCK;
      const int gx = x0 + e % ZM_W - ZM_M, gy = y0 + e / ZM_W - ZM_M;
      soff[k] = (e < ZM_PTS && gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny) ? gy * g.nx + gx
                                                                             : -1;
    }
  }
  const int zs = max(z0 - ZM_M, 0), ze = min(zend + ZM_M, g.nz);  // input planes [zs, ze)
  const uint32_t col = col_ok ? (uint32_t)(y * g.nx + x) : 0u;
  const uint32_t sz = (uint32_t)g.sz;
  const auto in_chunk = [&](int zo) { return zo >= z0 && zo < zend; };
  // rs[q]: the ring terms of output plane zi - M + q so far; cv[k]: the column value of plane
  // zi - k (walked planes; 0 past the top of the box, never used there)
  float rs[ZM_SL], cv[ZM_SL];
#pragma unroll
  for (int q = 0; q < ZM_SL; q++) rs[q] = cv[q] = 0.0f;
  // weight sums of the disc by plane distance |dz| (FP32, fixed order)
  float wz[ZM_M + 1];
#pragma unroll
  for (int a = 0; a <= ZM_M; a++) {
    float s = 0.0f;
#pragma unroll
    for (int dy = -ZM_M; dy <= ZM_M; dy++)
#pragma unroll
      for (int dx = -ZM_M; dx <= ZM_M; dx++) {
        const int d2 = dx * dx + dy * dy + a * a;
        if (d2 > ZM_M * ZM_M || d2 == 0) continue;
        s += zm_weight<float>(ball, d2);
      }
    wz[a] = s;
  }
  // diagonal term, free flag and handler operands of the next row to complete, a plane ahead
  float d0 = 0.0f;
  bool f0 = false;
  Pre pre0{};
  const auto load_first = [&](int zo) {
    if (col_ok && in_chunk(zo)) {
      const uint32_t i = (uint32_t)zo * sz + col;
      d0 = zm_ld(dd, i);
      f0 = zm_ld(pfree, i) != 0;
      pre0 = emit.pre(i);
    }
  };
  load_first(zs - ZM_M);
  // staged loads: VEC the quads of the next two planes (svq[0] for the planes zs, zs + 2, ...,
  // svq[1] for the others), else the next plane's values
  using SvRaw = typename std::conditional<VEC, typename ZmRaw4<Src, VEC>::type, Raw>::type;
  SvRaw svq[VEC ? 2 : 1][NST];
  const auto fetch = [&](int zi, SvRaw (&dst)[NST]) {
    const uint32_t base = (uint32_t)zi * sz;
#pragma unroll
    for (int k = 0; k < NST; k++) {
      if constexpr (VEC)
        dst[k] = soff[k] >= 0 ? src.load4(base + (uint32_t)soff[k]) : SvRaw{};
      else
        dst[k] = soff[k] >= 0 ? src.load(base + (uint32_t)soff[k]) : SvRaw{};
    }
  };
  fetch(zs, svq[0]);
  if (VEC && zs + 1 < ze) fetch(zs + 1, svq[VEC ? 1 : 0]);
  // after input plane zi (cv[0] its column value): output plane zo = zi - M is complete
  const auto finish = [&](int zi) {
    const int zo = zi - ZM_M;
    if (col_ok && in_chunk(zo)) {
      const float uo = cv[ZM_M];
      float line = 0.0f;
#pragma unroll
      for (int a = 1; a <= ZM_M; a++) {
        const bool up = zo + a < g.nz, dn = zo - a >= 0;  // uniform over the block
        float t;
        if (up && dn) t = (uo - cv[ZM_M - a]) + (uo - cv[ZM_M + a]);
        else if (up) t = uo - cv[ZM_M - a];
        else if (dn) t = uo - cv[ZM_M + a];
        else continue;
        line = __fmaf_rn(wz[a], t, line);
      }
      const float yv = f0 ? __fmaf_rn(d0, uo, rs[0] + line) : d0 * uo;
      if constexpr (RAWROW) {
        const int sl = (zo - zs) % ZM_RR;
        const Raw raw{lraw[sl][0][threadIdx.x], lraw[sl][1][threadIdx.x],
                      lraw[sl][2][threadIdx.x], lraw[sl][3][threadIdx.x]};
        emit((int64_t)((uint32_t)zo * sz + col), yv, uo, pre0, raw);
      } else {
        emit((int64_t)((uint32_t)zo * sz + col), yv, uo, pre0);
      }
    }
#pragma unroll
    for (int q = 0; q + 1 < ZM_SL; q++) rs[q] = rs[q + 1];
    rs[ZM_SL - 1] = 0.0f;
    load_first(zo + 1);
  };
  const auto shift_cv = [&](float v) {
#pragma unroll
    for (int k = ZM_SL - 1; k > 0; k--) cv[k] = cv[k - 1];
    cv[0] = v;
  };
  int buf = 0;
  // input plane zi; its staged loads are in cur, which then takes plane zi + 2 (VEC) or zi + 1
  const auto plane = [&](int zi, SvRaw (&cur)[NST]) {
    if constexpr (VEC) {
      if (threadIdx.x < ZM_Q4) {
        reinterpret_cast<float4*>(&ldsv[buf][0])[threadIdx.x] = src.value4(cur[0]);
        if constexpr (RAWROW) {
          // quads 1 .. ZM_TX / 4 of staged lines M .. M + ZM_TY - 1: the tile's own columns
          const int qx = threadIdx.x % (ZM_W4 / 4), qy = threadIdx.x / (ZM_W4 / 4);
          if (qx >= 1 && qx <= ZM_TX / 4 && qy >= ZM_M && qy < ZM_M + ZM_TY) {
            const int o = (qy - ZM_M) * ZM_TX + (qx - 1) * 4, sl = (zi - zs) % ZM_RR;
            *reinterpret_cast<float4*>(&lraw[sl][0][o]) = cur[0].r;
            *reinterpret_cast<float4*>(&lraw[sl][1][o]) = cur[0].w;
            *reinterpret_cast<float4*>(&lraw[sl][2][o]) = cur[0].s;
            *reinterpret_cast<float4*>(&lraw[sl][3][o]) = cur[0].d;
          }
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < NST; k++) {
        const int e = threadIdx.x + k * NNPP_BLOCK;
        if (e < ZM_PTS) ldsv[buf][e] = src.value(cur[k]);
      }
    }
    // this plane staged (and the buffer written two planes ago read by every thread)
    __syncthreads();
    if (zi + (VEC ? 2 : 1) < ze) fetch(zi + (VEC ? 2 : 1), cur);
    const float* __restrict__ mev = &ldsv[buf][(ty + ZM_M) * LW + tx + LX];
    const float uc = mev[0];
    float R[ZM_M * ZM_M + 1];
#pragma unroll
    for (int dy = -ZM_M; dy <= ZM_M; dy++)
#pragma unroll
      for (int dx = -ZM_M; dx <= ZM_M; dx++) {
        const int k = dx * dx + dy * dy;
        if (k == 0 || k > ZM_M * ZM_M) continue;
        const float d = mev[dy * LW + dx] - uc;
        R[k] = zm_ring_first<ZM_M>(dx, dy) ? d : R[k] + d;
      }
    // slot q (output plane zi - M + q) takes the rings at dz = M - q
#pragma unroll
    for (int q = 0; q < ZM_SL; q++) {
      if (!in_chunk(zi - ZM_M + q)) continue;  // uniform over the block
      const int dz = ZM_M - q, a = dz < 0 ? -dz : dz;
#pragma unroll
      for (int k = 1; k <= ZM_M * ZM_M; k++) {
        if (!zm_ring_exists<ZM_M>(k) || k + a * a > ZM_M * ZM_M) continue;
        rs[q] = __fmaf_rn(-zm_weight<float>(ball, k + a * a), R[k], rs[q]);
      }
    }
    shift_cv(uc);
    finish(zi);
    buf ^= 1;
  };
  if constexpr (VEC) {
    for (int zi = zs; zi < ze; zi += 2) {
      plane(zi, svq[0]);
      if (zi + 1 < ze) plane(zi + 1, svq[VEC ? 1 : 0]);
    }
  } else {
    for (int zi = zs; zi < ze; zi++) plane(zi, svq[0]);
  }
  // planes past the top of the box add nothing: the rows still in flight are complete
  for (int zi = ze; zi < zend + ZM_M; zi++) {
    shift_cv(0.0f);
    finish(zi);
  }
}

// Per column tile and plane: 1 when every point the tile stages from that plane (tile plus
// halo, all inside the box) is a free node; the z-marching kernels then skip the m_j factors.
__global__ void k_zm_full(ZmGeom g, const uint8_t* __restrict__ pfree, uint8_t* __restrict__ full) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t tiles = (int64_t)g.tiles_x * g.tiles_y;
  if (k >= tiles * g.nz) return;
  const int64_t tile = k / g.nz;
  const int z = (int)(k % g.nz);
  const int x0 = (int)(tile % g.tiles_x) * ZM_TX - ZM_M, y0 = (int)(tile / g.tiles_x) * ZM_TY - ZM_M;
  uint8_t all = 1;
  if (x0 < 0 || y0 < 0 || x0 + ZM_W > g.nx || y0 + ZM_H > g.ny) all = 0;
  const uint8_t* __restrict__ f = pfree + (int64_t)z * g.sz;
  for (int yy = 0; yy < ZM_H && all; yy++)
    for (int xx = 0; xx < ZM_W; xx++)
      if (!f[(int64_t)(y0 + yy) * g.nx + x0 + xx]) {
        all = 0;
        break;
      }
  full[k] = all;
}

// The diagonal of the unmasked z-marching walk (NOMASK): per free row, pdd32 less the FP32
// weights (summed in FP64) of its bonds, within the box's z range, to nodes that are not free
// or to in-plane positions outside the box; rows that are not free keep pdd32 (their input
// value is 0).
__global__ void k_zm_ddnf(ZmGeom g, nnpp_ball_args ball, const uint8_t* __restrict__ pfree,
                          const float* __restrict__ dd, float* __restrict__ ddnf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.sz * g.nz) return;
  if (!pfree[i]) {
    ddnf[i] = dd[i];
    return;
  }
  const int x = (int)(i % g.nx), y = (int)((i / g.nx) % g.ny), z = (int)(i / g.sz);
  constexpr nnpp_ball<NNPP_BALL_M> B{};
  double W = 0.0;
#pragma unroll
  for (int t = 0; t < B.K; t++) {
    const int l = B.line[t];
    const int zj = z + B.ldz[l], yj = y + B.ldy[l], xj = x + B.dx[t];
    if (zj < 0 || zj >= g.nz) continue;  // planes outside the box are never walked
    const bool out = xj < 0 || xj >= g.nx || yj < 0 || yj >= g.ny;
    if (out || !pfree[(int64_t)zj * g.sz + (int64_t)yj * g.nx + xj])
      W += (double)ball.w32[B.d2[t]];
  }
  ddnf[i] = (float)((double)dd[i] - W);
}

// A stored vector taken as 0 at the nodes that are not free (the input of the unmasked walk).
struct ZmLoadFree {
  struct Raw {
    float v;
    uint32_t f;
  };
  struct Raw4 {
    float4 v;
    uint32_t f;  // the 4 free flags, one per byte
  };
  const float* __restrict__ v;
  const uint8_t* __restrict__ pfree;
  __device__ __forceinline__ Raw load(uint32_t j) const { return Raw{zm_ld(v, j), zm_ld(pfree, j)}; }
  __device__ __forceinline__ float value(const Raw& r) const { return r.f ? r.v : 0.0f; }
  __device__ __forceinline__ Raw4 load4(uint32_t j) const {
    return Raw4{zm_ld(reinterpret_cast<const float4*>(v), j >> 2),
                zm_ld(reinterpret_cast<const uint32_t*>(pfree), j >> 2)};
  }
  __device__ __forceinline__ float4 value4(const Raw4& r) const {
    return make_float4((r.f & 0xffu) ? r.v.x : 0.0f, (r.f & 0xff00u) ? r.v.y : 0.0f,
                       (r.f & 0xff0000u) ? r.v.z : 0.0f, (r.f & 0xff000000u) ? r.v.w : 0.0f);
  }
};

// FP32 SpMV: q = M p on the free rows (0 elsewhere), partial p.q (FP64 over FP32 products).
// The unmasked walk (rings, dd = pdd32nf) on p with its values at non-free nodes taken as 0.
template <bool VEC>
__global__ __launch_bounds__(NNPP_BLOCK) void k_zm_m_spmv(ZmGeom g, nnpp_ball_args ball,
                                                          const uint8_t* __restrict__ pfree,
                                                          const uint8_t* __restrict__ full,
                                                          const float* __restrict__ ddnf,
                                                          const float* __restrict__ p,
                                                          float* __restrict__ q,
                                                          double* __restrict__ part,
                                                          const double* __restrict__ scal) {
  if (cg_done(scal)) return;
  double v[1] = {0.0};
  auto em = zm_emit([&](int64_t i, float y, float pi) {
    zm_st(q, (uint32_t)i, y);
    v[0] += (double)pi * (double)y;
  });
  zm_walk_ring<VEC>(g, ball, pfree, ddnf, ZmLoadFree{p, pfree}, em);
  nnpp_block_partial<1>(v, part);
}

// FP64 CG SpMV: q = M p, partial p.q.
__global__ __launch_bounds__(NNPP_BLOCK) void k_zm_cg_spmv(ZmGeom g, nnpp_ball_args ball,
                                                           const uint8_t* __restrict__ pfree,
                                                           const uint8_t* __restrict__ full,
                                                           const double* __restrict__ dd,
                                                           const double* __restrict__ p,
                                                           double* __restrict__ q,
                                                           double* __restrict__ part,
                                                           const double* __restrict__ scal) {
  if (cg_done(scal)) return;
  double v[1] = {0.0};
  auto em = zm_emit([&](int64_t i, double y, double pi) {
    q[i] = y;
    v[0] += pi * y;
  });
  zm_walk<double>(g, ball, pfree, full, dd, ZmLoad<double>{p}, em);
  nnpp_block_partial<1>(v, part);
}

// r = b - M x in FP64 (partials r.z, r.r); START: also z = D^-1 r, p = z.
template <bool START>
__global__ __launch_bounds__(NNPP_BLOCK) void k_zm_resid(ZmGeom g, nnpp_ball_args ball,
                                                         const uint8_t* __restrict__ pfree,
                                                         const uint8_t* __restrict__ full,
                                                         const double* __restrict__ dd,
                                                         const double* __restrict__ pinv,
                                                         const double* __restrict__ b,
                                                         const double* __restrict__ x,
                                                         double* __restrict__ r,
                                                         double* __restrict__ zv,
                                                         double* __restrict__ pv,
                                                         double* __restrict__ part) {
  double v[2] = {0.0, 0.0};
  auto em = zm_emit([&](int64_t i, double mx, double) {
    const double pi_ = pinv[i];
    const double ri = (pi_ != 0.0) ? b[i] - mx : 0.0;
    r[i] = ri;
    v[1] += ri * ri;
    if (START) {
      const double zi = ri * pi_;
      zv[i] = zi;
      pv[i] = zi;
      v[0] += ri * zi;
    }
  });
  zm_walk<double>(g, ball, pfree, full, dd, ZmLoad<double>{x}, em);
  nnpp_block_partial<2>(v, part);
}

// ------------------------------------------- single-reduction FP32 CG (z-marching) ---------
// The mixed solve's inner iterations on a box lattice in the single-reduction form of CG
// (Chronopoulos and Gear, J. Comput. Appl. Math. 25 (1989) 153-168). With u = D^-1 r and
// w = M u:
//   gamma = r.u, delta = u.w, beta = gamma / gamma_old, alpha = gamma / (delta - beta gamma / alpha_old),
//   p = u + beta p, s = w + beta s (= M p), x += alpha p, r -= alpha s,
// so one SpMV and one grid reduction (gamma, delta and r.r from the same pass) make an
// iteration. A launch applies the previous iteration's vector updates while it stages its
// input: at every point it stages, s_{k-1} = w_{k-1} + beta s_{k-2}, r_k = r_{k-1} - alpha s_{k-1}
// and u_k = D^-1 r_k are formed from the stored vectors (at the halo points of other blocks
// too), the SpMV gives w_k = M u_k, and the rows of the block store s_{k-1}, r_k, w_k,
// p_{k-1} = u_{k-1} + beta p_{k-2} and x_k = x_{k-1} + alpha p_{k-1}. r, w and s alternate
// between two buffers by launch parity (other blocks read the old values at their halos); p and
// x are read and written by their own rows only. alpha, beta and the stop test come from the
// previous launch's partials, which every block sums itself: an iteration is one kernel and one
// reduction instead of an SpMV, an update and a direction kernel with a reduction each.
//
// RESTART (at the start, and after a residual replacement): u = D^-1 r from the stored r,
// w = M u, r and s copied to the other parity, p and x untouched, no stop test. The next launch
// then continues the recurrence with the replaced residual (beta from the gamma before the
// replacement, the direction kept), or at the start (gamma_old = inf, p = s = 0) begins it.
// The rows outside the operator (D^-1 = 0) carry r = u = w = s = p = 0.
struct ZmCgVec {
  const float* r_in;  // r_{k-1}, w_{k-1}, s_{k-2} (parity of the launch)
  const float* w_in;
  const float* s_in;
  float* r_out;       // r_k, w_k, s_{k-1} (the other parity)
  float* w_out;
  float* s_out;
  float* p;           // p_{k-2} -> p_{k-1}
  float* e;           // x_{k-1} -> x_k (FP32 correction, units of ||b||)
  const float* pinv;  // D^-1
};

// timing: no stop test and no scalar writes (the kernel timing hook launches it repeatedly)
template <bool RESTART, bool VEC>
__global__ __launch_bounds__(NNPP_BLOCK) void k_zm_m_cgs(ZmGeom g, nnpp_ball_args ball,
                                                         const uint8_t* __restrict__ pfree,
                                                         const uint8_t* __restrict__ full,
                                                         const float* __restrict__ dd, ZmCgVec vec,
                                                         const double* __restrict__ part_in,
                                                         int64_t nblk,
                                                         double* __restrict__ part_out,
                                                         double* scal, int par, double reliable,
                                                         int timing) {
  if (cg_done(scal)) return;
  float alpha = 0.0f, beta = 0.0f;
  if (!RESTART) {
    double v[3];
    nnpp_final_sum_all<3>(v, part_in, nblk);  // gamma, delta, r.r of the previous launch
    // RRMAX is a running maximum: a block that reads it after block 0 updated it decides alike
    const double mx = fmax(nnpp_ld_scalar(scal + SC_RRMAX), v[2]);
    const bool done = !timing && (v[2] <= nnpp_ld_scalar(scal + SC_TARGET2) ||
                                  (reliable > 0.0 && v[2] <= reliable * mx));
    const double g_old = nnpp_ld_scalar(scal + SC_CG_GAMMA0 + 2 * par);
    const double a_old = nnpp_ld_scalar(scal + SC_CG_ALPHA0 + 2 * par);
    const double b = v[0] / g_old;
    const double a = v[0] / (v[1] - b * v[0] / a_old);
    if (blockIdx.x == 0 && threadIdx.x == 0 && !timing) {
      scal[SC_RR] = v[2];
      scal[SC_RRMAX] = mx;
      if (done) {
        scal[SC_DONE] = 1.0;
        scal[SC_CG_PARDONE] = (double)par;
      } else {
        scal[SC_CG_GAMMA0 + 2 * (par ^ 1)] = v[0];
        scal[SC_CG_ALPHA0 + 2 * (par ^ 1)] = a;
        scal[SC_ITERS] += 1.0;
      }
    }
    if (done) return;
    alpha = (float)a;
    beta = (float)b;
  } else if (blockIdx.x == 0 && threadIdx.x == 0 && !timing) {
    scal[SC_CG_GAMMA0 + 2 * (par ^ 1)] = nnpp_ld_scalar(scal + SC_CG_GAMMA0 + 2 * par);
    scal[SC_CG_ALPHA0 + 2 * (par ^ 1)] = nnpp_ld_scalar(scal + SC_CG_ALPHA0 + 2 * par);
  }
  // r_k at node j from the stored r_{k-1}, w_{k-1}, s_{k-2} (the staged points and the rows
  // form it alike, bit for bit), u_k = D^-1 r_k
  struct Src {
    struct Raw {
      float r, w, s, d;
    };
    struct Raw4 {
      float4 r, w, s, d;
    };
    const float* __restrict__ r_in;
    const float* __restrict__ w_in;
    const float* __restrict__ s_in;
    const float* __restrict__ pinv;
    float alpha, beta;
    __device__ __forceinline__ Raw load(uint32_t j) const {
      return Raw{zm_ld(r_in, j), RESTART ? 0.0f : zm_ld(w_in, j),
                 RESTART ? 0.0f : zm_ld(s_in, j), zm_ld(pinv, j)};
    }
    __device__ __forceinline__ float r_new(float r, float w, float s) const {
      return RESTART ? r : __fmaf_rn(-alpha, __fmaf_rn(beta, s, w), r);
    }
    __device__ __forceinline__ float r_new(const Raw& a) const { return r_new(a.r, a.w, a.s); }
    __device__ __forceinline__ float value(const Raw& a) const { return a.d * r_new(a); }
    __device__ __forceinline__ Raw4 load4(uint32_t j) const {
      const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      const uint32_t q = j >> 2;
      return Raw4{zm_ld(reinterpret_cast<const float4*>(r_in), q),
                  RESTART ? z : zm_ld(reinterpret_cast<const float4*>(w_in), q),
                  RESTART ? z : zm_ld(reinterpret_cast<const float4*>(s_in), q),
                  zm_ld(reinterpret_cast<const float4*>(pinv), q)};
    }
    __device__ __forceinline__ float4 value4(const Raw4& a) const {
      return make_float4(a.d.x * r_new(a.r.x, a.w.x, a.s.x), a.d.y * r_new(a.r.y, a.w.y, a.s.y),
                         a.d.z * r_new(a.r.z, a.w.z, a.s.z), a.d.w * r_new(a.r.w, a.w.w, a.s.w));
    }
  } src{vec.r_in, vec.w_in, vec.s_in, vec.pinv, alpha, beta};
  // rows: the stored vectors' old values, loaded a plane ahead
  double v[3] = {0.0, 0.0, 0.0};
  // (RawTag, 16-byte staging: r, w, s and D^-1 of the row from the walk's staging; RESTART
  // loads them, its staged w and s being zero)
  struct Em {
    using RawTag = std::bool_constant<!RESTART>;
    struct Pre {
      typename Src::Raw a;
      float p, e;
    };
    const Src& src;
    const ZmCgVec& vec;
    double (&v)[3];
    __device__ __forceinline__ Pre pre(uint32_t i) const {
      Pre q;
      if (!(RawTag::value && VEC)) {
        q.a = src.load(i);
        q.a.s = zm_ld(vec.s_in, i);  // restart: s is copied
      }
      q.p = RESTART ? 0.0f : zm_ld((const float*)vec.p, i);
      q.e = RESTART ? 0.0f : zm_ld((const float*)vec.e, i);
      return q;
    }
    __device__ __forceinline__ void operator()(int64_t i64, float y, float ui, const Pre& q,
                                               const typename Src::Raw& a) {
      const uint32_t i = (uint32_t)i64;
      const float ri = src.r_new(a);
      if (RESTART) {
        zm_st(vec.s_out, i, a.s);
      } else {
        const float si = __fmaf_rn(src.beta, a.s, a.w);
        const float pi_ = __fmaf_rn(src.beta, q.p, a.d * a.r);
        zm_st(vec.p, i, pi_);
        zm_st(vec.e, i, __fmaf_rn(src.alpha, pi_, q.e));
        zm_st(vec.s_out, i, si);
      }
      zm_st(vec.r_out, i, ri);
      zm_st(vec.w_out, i, y);
      v[0] += (double)ri * (double)ui;
      v[1] += (double)ui * (double)y;
      v[2] += (double)ri * (double)ri;
    }
    __device__ __forceinline__ void operator()(int64_t i64, float y, float ui, const Pre& q) {
      (*this)(i64, y, ui, q, q.a);
    }
  } em{src, vec, v};
  zm_walk_ring<VEC>(g, ball, pfree, dd, src, em);
  nnpp_block_partial<3>(v, part_out);
}

// Start (START) or residual replacement of the single-reduction FP32 CG: r = r64 / ||b|| into
// the parity-0 buffer, x = 0 (the correction is in the FP64 solution); START: p = s = 0 and
// gamma_old = inf, so the first iteration takes beta = 0; otherwise s and (gamma, alpha) move
// from the parity of the launch that stopped to parity 0. Partials (0, r.r).
template <bool START>
__global__ __launch_bounds__(NNPP_BLOCK) void k_zm_m_replace(int32_t n,
                                                             const double* __restrict__ r64,
                                                             double sb, float* __restrict__ r0,
                                                             float* __restrict__ e,
                                                             float* __restrict__ p,
                                                             float* __restrict__ s0,
                                                             const float* __restrict__ s1,
                                                             double* __restrict__ part,
                                                             double* scal) {
  const int q = START ? 0 : (int)nnpp_ld_scalar(scal + SC_CG_PARDONE);
  double v[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * NNPP_BLOCK + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * NNPP_BLOCK) {
    const float ri = (float)(r64[i] * sb);
    r0[i] = ri;
    e[i] = 0.0f;
    if (START) {
      p[i] = 0.0f;
      s0[i] = 0.0f;
    } else if (q) {
      s0[i] = s1[i];
    }
    v[1] += (double)ri * (double)ri;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (START) {
      scal[SC_CG_GAMMA0] = __builtin_inf();
      scal[SC_CG_ALPHA0] = 1.0;
    } else if (q) {
      scal[SC_CG_GAMMA0] = nnpp_ld_scalar(scal + SC_CG_GAMMA1);
      scal[SC_CG_ALPHA0] = nnpp_ld_scalar(scal + SC_CG_ALPHA1);
    }
  }
  nnpp_block_partial<2>(v, part);
}

// b (free rows), warm start x = phi, ||b||^2.
__global__ __launch_bounds__(NNPP_BLOCK) void k_rhs(int32_t n, int S, const double* __restrict__ U,
                                                    const double* __restrict__ zs, double cq,
                                                    const uint8_t* __restrict__ pfree,
                                                    const double* __restrict__ psc,
                                                    const double* __restrict__ pbdir,
                                                    const double* __restrict__ phi,
                                                    const double* __restrict__ rhs,
                                                    double* __restrict__ b, double* __restrict__ x,
                                                    double* __restrict__ part) {
  double v[1] = {0.0};
  const int64_t nb = nnpp_nblocks(n);
  for (nnpp_block_iter it = nnpp_blocks_of(nb); it.cur < it.end; it.cur += it.stride) {
    const int64_t i = it.cur * NNPP_BLOCK + threadIdx.x;
    if (i >= n) continue;
    double bi = 0.0, xi = 0.0;
    if (pfree[i]) {
      if (rhs) {
        bi = rhs[i];
      } else {
        double q = 0.0;
        for (int k = 0; k < S; k++) q += zs[k] * U[(int64_t)k * n + i];
        bi = psc[i] * (cq * q + pbdir[i]);
      }
      xi = phi[i];
    }
    b[i] = bi;
    x[i] = xi;
    v[0] += bi * bi;
  }
  nnpp_block_partial<1>(v, part);
}

// r = b - M x in FP64 (RR = ||r||^2). START: also z = D^-1 r, p = z, rho, the target from
// ||b|| and the converged flag.
template <bool START, bool WIDE, bool CLS, bool STEN>
__global__ __launch_bounds__(NNPP_BLOCK) void k_resid(int32_t n, Slices sl,
                                                      const double* __restrict__ w,
                                                      const double* __restrict__ dd,
                                                      const double* __restrict__ pinv,
                                                      const double* __restrict__ b,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ r, double* __restrict__ zv,
                                                      double* __restrict__ pv,
                                                      double* __restrict__ part) {
  __shared__ double tab[CLS ? NNPP_DICT_MAX : 1];
  extern __shared__ double st_lds[];
  if (CLS) load_tab(tab, sl.tab64);
  double v[2] = {0.0, 0.0};
  const int64_t nb = nnpp_nblocks(n);
  for (nnpp_block_iter it = nnpp_blocks_of(nb); it.cur < it.end; it.cur += it.stride) {
    const int64_t i = it.cur * NNPP_BLOCK + threadIdx.x;
    double mx = 0.0;
    int64_t ir = i;
    if constexpr (STEN) mx = stencil_tile<double, true>(sl, dd, x, n, it.cur, st_lds, &ir);
    if (i >= n) continue;
    if constexpr (!STEN)
      if (pinv[i] != 0.0) mx = spmv_row<double, WIDE, CLS>(sl, w, tab, dd, x, i);
    const double ri = (pinv[i] != 0.0) ? b[i] - mx : 0.0;
    r[i] = ri;
    v[1] += ri * ri;
    if (START) {
      const double zi = ri * pinv[i];
      zv[i] = zi;
      pv[i] = zi;
      v[0] += ri * zi;
    }
  }
  nnpp_block_partial<2>(v, part);
}

template <bool WIDE, bool CLS, bool STEN, int M = 0>
__global__ __launch_bounds__(NNPP_BLOCK) void k_cg_spmv(int32_t n, Slices sl,
                                                        const double* __restrict__ w,
                                                        const double* __restrict__ dd,
                                                        const double* __restrict__ p,
                                                        double* __restrict__ q, double* __restrict__ part, const double* __restrict__ scal) {
  if (cg_done(scal)) return;
  __shared__ double tab[CLS ? NNPP_DICT_MAX : 1];
  extern __shared__ double st_lds[];
  if (CLS) load_tab(tab, sl.tab64);
  double v[1] = {0.0};
  const int64_t nb = nnpp_nblocks(n);
  for (nnpp_block_iter it = nnpp_blocks_of(nb); it.cur < it.end; it.cur += it.stride) {
    const int64_t i = it.cur * NNPP_BLOCK + threadIdx.x;
    double qi = 0.0;
    int64_t ir = i;
    if constexpr (STEN) qi = stencil_tile<double, true, M>(sl, dd, p, n, it.cur, st_lds, &ir);
    if (i >= n) continue;
    if constexpr (!STEN) qi = spmv_row<double, WIDE, CLS>(sl, w, tab, dd, p, i);
    q[i] = qi;
    v[0] += p[i] * qi;
  }
  nnpp_block_partial<1>(v, part);
}

// Slot of rho in iteration parity par.
__host__ __device__ constexpr int rho_slot(int par) { return par ? SC_RHO1 : SC_RHO; }

// alpha = rho / p.Ap from the SpMV's block partials (every block sums them in the same order;
// block 0 records PQ and ALPHA).
__device__ __forceinline__ double cg_alpha(const double* __restrict__ part_s, int64_t nblk_s,
                                           double* scal, int par) {
  double pq[1];
  nnpp_final_sum_all<1>(pq, part_s, nblk_s);
  const double alpha = nnpp_ld_scalar(scal + rho_slot(par)) / pq[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[SC_PQ] = pq[0];
    scal[SC_ALPHA] = alpha;
  }
  return alpha;
}

// beta, ||r||^2 and the stop test from the update's block partials (r.z, r.r). Block 0 records
// BETA, RR, RRMAX, ITERS, DONE and the new rho in the other slot (no block reads that slot in
// this kernel). reliable > 0 (mixed): also stop once ||r||^2 <= reliable * RRMAX; RRMAX is a
// running maximum, so a block that reads it after block 0 updated it takes the same decision.
__device__ __forceinline__ bool cg_beta(const double* __restrict__ part_u, int64_t nblk_u,
                                        double* scal, int par, double reliable, double& beta) {
  double v[2];
  nnpp_final_sum_all<2>(v, part_u, nblk_u);
  beta = v[0] / nnpp_ld_scalar(scal + rho_slot(par));
  const double mx = fmax(nnpp_ld_scalar(scal + SC_RRMAX), v[1]);
  const bool done = v[1] <= nnpp_ld_scalar(scal + SC_TARGET2) || (reliable > 0.0 && v[1] <= reliable * mx);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[SC_BETA] = beta;
    scal[rho_slot(par ^ 1)] = v[0];
    scal[SC_RR] = v[1];
    scal[SC_ITERS] += 1.0;
    if (reliable > 0.0) scal[SC_RRMAX] = mx;
    if (done) scal[SC_DONE] = 1.0;
  }
  return done;
}

// x += alpha p, r -= alpha q, z = D^-1 r; partials r.z, r.r into part_u.
__global__ __launch_bounds__(NNPP_BLOCK) void k_cg_update(int32_t n, const double* __restrict__ pinv,
                                                          const double* __restrict__ p,
                                                          const double* __restrict__ q,
                                                          double* __restrict__ x,
                                                          double* __restrict__ r,
                                                          double* __restrict__ zv,
                                                          double* __restrict__ part_u,
                                                          const double* __restrict__ part_s,
                                                          int64_t nblk_s, double* scal, int par) {
  if (cg_done(scal)) return;
  const double alpha = cg_alpha(part_s, nblk_s, scal, par);
  double v[2] = {0.0, 0.0};
  // two consecutive nodes per thread and 16-byte accesses (every vector is its own 256-byte
  // aligned allocation); an odd last node goes to lane 0 of block 0
  const int64_t n2 = n >> 1;
  for (int64_t j = (int64_t)blockIdx.x * NNPP_BLOCK + threadIdx.x; j < n2;
       j += (int64_t)gridDim.x * NNPP_BLOCK) {
    const double2 pj = reinterpret_cast<const double2*>(p)[j];
    const double2 qj = reinterpret_cast<const double2*>(q)[j];
    const double2 dj = reinterpret_cast<const double2*>(pinv)[j];
    double2 xj = reinterpret_cast<const double2*>(x)[j];
    double2 rj = reinterpret_cast<const double2*>(r)[j];
    double2 zj;
    xj.x += alpha * pj.x;
    xj.y += alpha * pj.y;
    rj.x = rj.x - alpha * qj.x;
    rj.y = rj.y - alpha * qj.y;
    zj.x = rj.x * dj.x;
    zj.y = rj.y * dj.y;
    reinterpret_cast<double2*>(x)[j] = xj;
    reinterpret_cast<double2*>(r)[j] = rj;
    reinterpret_cast<double2*>(zv)[j] = zj;
    v[0] += rj.x * zj.x;
    v[0] += rj.y * zj.y;
    v[1] += rj.x * rj.x;
    v[1] += rj.y * rj.y;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) {
    const int64_t i = n - 1;
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    const double zi = ri * pinv[i];
    zv[i] = zi;
    v[0] += ri * zi;
    v[1] += ri * ri;
  }
  nnpp_block_partial<2>(v, part_u);
}

// p = z + beta p, after the stop test (the iteration that stops skips it).
__global__ __launch_bounds__(NNPP_BLOCK) void k_cg_dir(int32_t n, const double* __restrict__ zv,
                                                       double* __restrict__ p,
                                                       const double* __restrict__ part_u,
                                                       int64_t nblk_u, double* scal, int par) {
  if (cg_done(scal)) return;
  double beta;
  if (cg_beta(part_u, nblk_u, scal, par, 0.0, beta)) return;
  // two nodes per thread, 16-byte accesses; an odd last node in lane 0 of block 0
  const int64_t n2 = n >> 1;
  for (int64_t j = (int64_t)blockIdx.x * NNPP_BLOCK + threadIdx.x; j < n2;
       j += (int64_t)gridDim.x * NNPP_BLOCK) {
    const double2 zj = reinterpret_cast<const double2*>(zv)[j];
    double2 pj = reinterpret_cast<const double2*>(p)[j];
    pj.x = zj.x + beta * pj.x;
    pj.y = zj.y + beta * pj.y;
    reinterpret_cast<double2*>(p)[j] = pj;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) p[n - 1] = zv[n - 1] + beta * p[n - 1];
}

// ---- mixed precision: CG with reliable updates ----

// FP32 residual from the FP64 one (normalised by sb = 1/||b||), z = D^-1 r, rho, ||r||^2.
// START also clears the correction e and sets p = z and the iteration count.
template <bool START>
__global__ __launch_bounds__(NNPP_BLOCK) void k_m_replace(int32_t n, const double* __restrict__ r64,
                                                          double sb, const float* __restrict__ pinv,
                                                          float* __restrict__ e,
                                                          float* __restrict__ r,
                                                          float* __restrict__ zv,
                                                          float* __restrict__ p,
                                                          double* __restrict__ part) {
  double v[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * NNPP_BLOCK + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * NNPP_BLOCK) {
    const float ri = (float)(r64[i] * sb);
    const float zi = ri * pinv[i];
    r[i] = ri;
    zv[i] = zi;
    if (START) {
      e[i] = 0.0f;
      p[i] = zi;
    }
    v[0] += (double)ri * (double)zi;
    v[1] += (double)ri * (double)ri;
  }
  nnpp_block_partial<2>(v, part);
}

template <bool WIDE, bool CLS, bool STEN, int M = 0>
__device__ __forceinline__ void m_spmv(int32_t n, const Slices& sl, const float* __restrict__ w,
                                       const float* __restrict__ dd, const float* __restrict__ p,
                                       float* __restrict__ q, double* __restrict__ part,
                                       const double* __restrict__ scal) {
  if (cg_done(scal)) return;
  __shared__ float tab[CLS ? NNPP_DICT_MAX : 1];
  extern __shared__ float st_lds32[];
  if (CLS) load_tab(tab, sl.tab32);
  double v[1] = {0.0};
  const int64_t nb = nnpp_nblocks(n);
  for (nnpp_block_iter it = nnpp_blocks_of(nb); it.cur < it.end; it.cur += it.stride) {
    int64_t i = it.cur * NNPP_BLOCK + threadIdx.x;
    float qi = 0.0f;
    // FP32: the stencil rows stay in their lane order (the dot is FP64 over FP32 products)
    if constexpr (STEN) qi = stencil_tile<float, false, M>(sl, dd, p, n, it.cur, st_lds32, &i);
    if (i >= n) continue;
    if constexpr (!STEN) qi = spmv_row<float, WIDE, CLS>(sl, w, tab, dd, p, i);
    q[i] = qi;
    v[0] += (double)p[i] * (double)qi;
  }
  nnpp_block_partial<1>(v, part);
}

template <bool WIDE, bool CLS>
__global__ __launch_bounds__(NNPP_BLOCK) void k_m_spmv(int32_t n, Slices sl,
                                                       const float* __restrict__ w,
                                                       const float* __restrict__ dd,
                                                       const float* __restrict__ p,
                                                       float* __restrict__ q,
                                                       double* __restrict__ part,
                                                       const double* __restrict__ scal) {
  m_spmv<WIDE, CLS, false>(n, sl, w, dd, p, q, part, scal);
}

// Stencil form, at most 72 VGPRs (7 waves per SIMD): measured 4% faster than the compiler's 76
// (6 waves), with two registers spilled outside the row sums. M > 0: the lattice-ball form.
template <int M>
__global__ __launch_bounds__(NNPP_BLOCK, 7) void k_m_spmv_st(int32_t n, Slices sl,
                                                             const float* __restrict__ w,
                                                             const float* __restrict__ dd,
                                                             const float* __restrict__ p,
                                                             float* __restrict__ q,
                                                             double* __restrict__ part,
                                                             const double* __restrict__ scal) {
  m_spmv<false, false, true, M>(n, sl, w, dd, p, q, part, scal);
}

// e += alpha p, r -= alpha q, z = D^-1 r; stops the batch (DONE) at the target or when the
// residual has dropped by trig2 (squared) below its maximum since the last replacement.
__global__ __launch_bounds__(NNPP_BLOCK) void k_m_update(int32_t n, const float* __restrict__ pinv,
                                                         const float* __restrict__ p,
                                                         const float* __restrict__ q,
                                                         float* __restrict__ e,
                                                         float* __restrict__ r,
                                                         float* __restrict__ zv,
                                                         double* __restrict__ part_u,
                                                         const double* __restrict__ part_s,
                                                         int64_t nblk_s, double* scal, int par) {
  if (cg_done(scal)) return;
  const float alpha = (float)cg_alpha(part_s, nblk_s, scal, par);
  double v[2] = {0.0, 0.0};
  // four consecutive nodes per thread and 16-byte accesses (every vector is its own 256-byte
  // aligned allocation); the last n % 4 nodes go to the first lanes of block 0
  const auto node = [&](int64_t i) {
    e[i] += alpha * p[i];
    const float ri = r[i] - alpha * q[i];
    r[i] = ri;
    const float zi = ri * pinv[i];
    zv[i] = zi;
    v[0] += (double)ri * (double)zi;
    v[1] += (double)ri * (double)ri;
  };
  const int64_t n4 = n >> 2;
  for (int64_t j = (int64_t)blockIdx.x * NNPP_BLOCK + threadIdx.x; j < n4;
       j += (int64_t)gridDim.x * NNPP_BLOCK) {
    const float4 pj = reinterpret_cast<const float4*>(p)[j];
    const float4 qj = reinterpret_cast<const float4*>(q)[j];
    const float4 dj = reinterpret_cast<const float4*>(pinv)[j];
    float4 ej = reinterpret_cast<const float4*>(e)[j];
    float4 rj = reinterpret_cast<const float4*>(r)[j];
    float4 zj;
    ej.x += alpha * pj.x;
    ej.y += alpha * pj.y;
    ej.z += alpha * pj.z;
    ej.w += alpha * pj.w;
    rj.x = rj.x - alpha * qj.x;
    rj.y = rj.y - alpha * qj.y;
    rj.z = rj.z - alpha * qj.z;
    rj.w = rj.w - alpha * qj.w;
    zj.x = rj.x * dj.x;
    zj.y = rj.y * dj.y;
    zj.z = rj.z * dj.z;
    zj.w = rj.w * dj.w;
    reinterpret_cast<float4*>(e)[j] = ej;
    reinterpret_cast<float4*>(r)[j] = rj;
    reinterpret_cast<float4*>(zv)[j] = zj;
    v[0] += (double)rj.x * (double)zj.x;
    v[0] += (double)rj.y * (double)zj.y;
    v[0] += (double)rj.z * (double)zj.z;
    v[0] += (double)rj.w * (double)zj.w;
    v[1] += (double)rj.x * (double)rj.x;
    v[1] += (double)rj.y * (double)rj.y;
    v[1] += (double)rj.z * (double)rj.z;
    v[1] += (double)rj.w * (double)rj.w;
  }
  if (blockIdx.x == 0 && threadIdx.x < n - 4 * n4) node(4 * n4 + threadIdx.x);
  nnpp_block_partial<2>(v, part_u);
}

// p = z + beta p after the stop test (beta, the test and the scalars from the update's
// partials). The iteration that stops a batch skips it, so the host runs k_m_dir_force once
// before a residual replacement: continuing along a direction that missed its update would take
// the step length r.z / p.Ap, which only holds for r orthogonal to the previous direction, and
// diverges.
__global__ __launch_bounds__(NNPP_BLOCK) void k_m_dir(int32_t n, const float* __restrict__ zv,
                                                      float* __restrict__ p,
                                                      const double* __restrict__ part_u,
                                                      int64_t nblk_u, double* scal, int par,
                                                      double reliable) {
  if (cg_done(scal)) return;
  double b;
  if (cg_beta(part_u, nblk_u, scal, par, reliable, b)) return;
  const float beta = (float)b;
  // four nodes per thread, 16-byte accesses; the last n % 4 nodes in block 0
  const int64_t n4 = n >> 2;
  for (int64_t j = (int64_t)blockIdx.x * NNPP_BLOCK + threadIdx.x; j < n4;
       j += (int64_t)gridDim.x * NNPP_BLOCK) {
    const float4 zj = reinterpret_cast<const float4*>(zv)[j];
    float4 pj = reinterpret_cast<const float4*>(p)[j];
    pj.x = zj.x + beta * pj.x;
    pj.y = zj.y + beta * pj.y;
    pj.z = zj.z + beta * pj.z;
    pj.w = zj.w + beta * pj.w;
    reinterpret_cast<float4*>(p)[j] = pj;
  }
  if (blockIdx.x == 0 && threadIdx.x < n - 4 * n4) {
    const int64_t i = 4 * n4 + threadIdx.x;
    p[i] = zv[i] + beta * p[i];
  }
}

// p = z + BETA p with the BETA of the iteration that stopped (before a residual replacement).
__global__ void k_m_dir_force(int32_t n, const float* __restrict__ zv, float* __restrict__ p,
                              const double* scal) {
  const float beta = (float)nnpp_ld_scalar(scal + SC_BETA);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = zv[i] + beta * p[i];
}

// x += e / sb (FP64), e = 0.
__global__ void k_m_apply(int32_t n, double inv_sb, float* __restrict__ e,
                          double* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    x[i] += inv_sb * (double)e[i];
    e[i] = 0.0f;
  }
}

// phi = x on free nodes, Dirichlet value on Dirichlet nodes, unchanged elsewhere.
__global__ void k_phi_guess(int32_t n, int hist, const uint8_t* __restrict__ pfree,
                            const double* __restrict__ phi, const double* __restrict__ prev,
                            const double* __restrict__ prev2, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = phi[i];
  double g = a;
  if (pfree[i] && hist >= 2) g = 3.0 * (a - prev[i]) + prev2[i];
  else if (pfree[i] && hist == 1) g = 2.0 * a - prev[i];
  out[i] = g;
}

__global__ void k_scatter_phi(int32_t n, int S, const uint8_t* __restrict__ pfree,
                              const uint16_t* __restrict__ bcmask, const double* __restrict__ bcphi,
                              const double* __restrict__ x, double* __restrict__ phi) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (pfree[i]) phi[i] = x[i];
  else if (bcphi && ((bcmask[i] >> S) & 1u)) phi[i] = bcphi[i];
}

// variant: 2 * wide + dictionary (column form), 4 = stencil form (lds bytes of dynamic LDS)
template <bool START, typename... A>
void launch_resid(int var, size_t lds, int64_t g, hipStream_t s, A... a) {
  switch (var) {
    case 0: k_resid<START, false, false, false><<<g, NNPP_BLOCK, 0, s>>>(a...); break;
    case 1: k_resid<START, false, true, false><<<g, NNPP_BLOCK, 0, s>>>(a...); break;
    case 2: k_resid<START, true, false, false><<<g, NNPP_BLOCK, 0, s>>>(a...); break;
    case 3: k_resid<START, true, true, false><<<g, NNPP_BLOCK, 0, s>>>(a...); break;
    default: k_resid<START, false, false, true><<<g, NNPP_BLOCK, lds, s>>>(a...); break;
  }
}

template <typename... A>
void launch_cg_spmv(int var, size_t lds, int64_t g, hipStream_t s, A... a) {
  switch (var) {
    case 0: k_cg_spmv<false, false, false><<<g, NNPP_BLOCK, 0, s>>>(a...); break;
    case 1: k_cg_spmv<false, true, false><<<g, NNPP_BLOCK, 0, s>>>(a...); break;
    case 2: k_cg_spmv<true, false, false><<<g, NNPP_BLOCK, 0, s>>>(a...); break;
    case 3: k_cg_spmv<true, true, false><<<g, NNPP_BLOCK, 0, s>>>(a...); break;
    case 5: k_cg_spmv<false, false, true, NNPP_BALL_M><<<g, NNPP_BLOCK, lds, s>>>(a...); break;
    default: k_cg_spmv<false, false, true><<<g, NNPP_BLOCK, lds, s>>>(a...); break;
  }
}

template <typename... A>
void launch_m_spmv(int var, size_t lds, int64_t g, hipStream_t s, A... a) {
  switch (var) {
    case 0: k_m_spmv<false, false><<<g, NNPP_BLOCK, 0, s>>>(a...); break;
    case 1: k_m_spmv<false, true><<<g, NNPP_BLOCK, 0, s>>>(a...); break;
    case 2: k_m_spmv<true, false><<<g, NNPP_BLOCK, 0, s>>>(a...); break;
    case 3: k_m_spmv<true, true><<<g, NNPP_BLOCK, 0, s>>>(a...); break;
    case 5: k_m_spmv_st<NNPP_BALL_M><<<g, NNPP_BLOCK, lds, s>>>(a...); break;
    default: k_m_spmv_st<0><<<g, NNPP_BLOCK, lds, s>>>(a...); break;
  }
}

// Operator view for the SpMV kernels: variant, slices/stencil tables, and the weight arrays.
struct PotOp {
  Slices sl;
  int var;
  size_t lds32, lds64;
  const double* w64;
  const float* w32;
  int64_t g32, g64;  // grids of the FP32 / FP64 SpMV-type kernels
  bool zm;           // z-marching kernels (box lattice)
  ZmGeom zg;
};

// Work items of the z-marching kernels: column tiles times chunks of lz output planes, lz
// chosen for about 1024 items (at most NNPP_RED_GRID: the reduction partials), at least 4.
ZmGeom zm_geom(const nnpp_stencil& T) { return nnpp_zm_geom(T, 1024); }

// 16-byte staging of the FP32 ring kernels: every staged quad lies wholly inside or outside the
// box and starts 16-byte aligned.
bool zm_vec4(const ZmGeom& g) { return g.nx % 4 == 0; }

// Blocks of the stencil SpMV kernels that fit on the device at once (LDS bounds them): a grid of
// exactly that many blocks walks the tiles in one round, where a larger grid would leave a
// partly filled second round. A multiple of 8 (one share per XCD), at most nnpp_red_grid.
template <typename K>
int64_t resident_grid(K kernel, size_t lds, int device, int64_t cap) {
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, NNPP_BLOCK, lds) !=
          hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      per_cu <= 0 || cus <= 0)
    return cap;
  int64_t g = (int64_t)per_cu * cus;
  g -= g % NNPP_NXCD;
  return g < NNPP_NXCD ? cap : (g < cap ? g : cap);
}

PotOp pot_op(const nnpp_ctx* c) {
  PotOp o;
  std::memset(&o, 0, sizeof(o));
  const nnpp_slices& P = c->pot;
  const nnpp_stencil& T = c->pst;
  o.sl = Slices{P.off, P.len, P.col, P.d16, P.first, P.cls, P.tab64, P.tab32,
                T.pmask, T.lc, T.seg_start, T.seg_len, T.seg_base, T.piece, T.tw64, T.tw32,
                T.perm, T.K, T.R, T.P, T.off_min, T.off_max, T.ball_args};
  const int form = nnpp_pot_form(c);
  const bool ball = form == NNPP_BALL_M;
  o.zm = form == NNPP_POT_ZMARCH;
  if (o.zm) o.zg = zm_geom(T);
  if (T.active) {
    o.var = ball ? 5 : 4;
    o.lds32 = sizeof(float) * T.lds_elems;
    o.lds64 = sizeof(double) * T.lds_elems;
    o.w64 = T.tw64;
    o.w32 = T.tw32;
  } else {
    o.var = 2 * (P.wide != 0) + (P.ncls > 0);
    o.w64 = P.w64;
    o.w32 = P.w32;
  }
  const int64_t cap = nnpp_red_grid(nnpp_nblocks(c->n));
  o.g32 = o.g64 = cap;
  if (o.zm) {
    o.g32 = o.g64 = o.zg.items;
  } else if (T.active) {
    o.g32 = ball ? resident_grid(k_m_spmv_st<NNPP_BALL_M>, o.lds32, c->device, cap)
                 : resident_grid(k_m_spmv_st<0>, o.lds32, c->device, cap);
    o.g64 = ball ? resident_grid(k_cg_spmv<false, false, true, NNPP_BALL_M>, o.lds64, c->device,
                                 cap)
                 : resident_grid(k_cg_spmv<false, false, true>, o.lds64, c->device, cap);
  }
  return o;
}

}  // namespace

static int read_scalars(nnpp_ctx* c) {
  NNPP_HIP(c, hipMemcpyAsync(c->h_scal, c->scal, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost,
                             c->stream));
  NNPP_HIP(c, hipStreamSynchronize(c->stream));
  return NNPP_OK;
}


// Iterations of the next launch batch: the host sees the residual only between batches, and
// iterations launched after the stop exit at once but still cost their launches (about 16 us
// for the three of an iteration at 128^3, against 30-50 us for a host round trip). A phase
// (solve start or residual replacement up to the next stop) is first sized by the same phase of
// the previous solve, plus one: successive steps repeat the phase lengths closely. Beyond that,
// from the squared-residual reduction per iteration of the batches so far, predict the
// iterations to the stop threshold (plus one); without either, CG_BATCH.
struct BatchPlan {
  double rate = 0.0;
  const std::vector<int32_t>* prev = nullptr;  // phase lengths of the previous solve
  std::vector<int32_t> lens;                   // phase lengths of this solve
  int phase = 0, phase_start = 0;
  void observe(double rr0, double rr1, int its) {
    if (its > 0 && rr0 > 0.0 && rr1 > 0.0 && rr1 < rr0) rate = std::pow(rr1 / rr0, 1.0 / its);
  }
  // a phase ended after `total` iterations of the solve
  void end_phase(int total) {
    lens.push_back(total - phase_start);
    phase++;
    phase_start = total;
  }
  int next(double rr, double stop, int left, int total) const {
    int k = 0;
    if (prev && phase < (int)prev->size()) k = (*prev)[phase] - (total - phase_start) + 1;
    if (k < 2) {
      k = CG_BATCH;
      if (rate > 0.0 && rate < 1.0 && stop > 0.0 && rr > stop)
        k = (int)std::ceil(std::log(stop / rr) / std::log(rate)) + 1;
    }
    k = k < 2 ? 2 : (k > 4 * CG_BATCH ? 4 * CG_BATCH : k);
    return k < left ? k : left;
  }
};

// Uq: state whose species give the charge; phi: warm start in, result out (non-free,
// non-Dirichlet entries are left as they are).
int nnpp_potential_solve_dev(nnpp_ctx* c, int precision, const double* Uq, double* phi,
                             int32_t* iters, double* relres, int64_t* spmv64, int64_t* spmv32,
                             const double* rhs) {
  const int32_t n = c->n;
  const int S = c->p.n_species;
  hipStream_t s = c->stream;
  if (iters) *iters = 0;
  if (relres) *relres = 0.0;
  if (!n) return NNPP_OK;
  const int64_t nb = nnpp_nblocks(n);
  const int64_t g = nnpp_red_grid(nb);  // grid of every kernel with a reduction
  const PotOp op = pot_op(c);
  const Slices& sl = op.sl;
  const int var = op.var;
  double* part = c->red_part;
  const double* scal = c->scal;
  double zs[NNPP_MAX_SPECIES] = {0};
  for (int k = 0; k < S; k++) zs[k] = c->p.z[k];
  NNPP_HIP(c, hipMemcpyAsync(c->scal + SC_COUNT, zs, sizeof(double) * NNPP_MAX_SPECIES,
                             hipMemcpyHostToDevice, s));
  const double cq = (1. / (c->p.K_phi * c->p.sigma1)) * c->p.F * c->p.CSolid;
  k_rhs<<<g, NNPP_BLOCK, 0, s>>>(n, S, Uq, c->scal + SC_COUNT, cq, c->pfree, c->psc, c->pbdir,
                                 phi, rhs, c->cg_b, c->cg_x, part);
  NNPP_HIP(c, hipGetLastError());
  int st = nnpp_finalize(c, FIN_BNORM2, g, 0.0);
  if (st) return st;
  st = read_scalars(c);
  if (st) return st;
  const double bnorm2 = c->h_scal[SC_BNORM2];
  if (!std::isfinite(bnorm2))
    return nnpp_set_error(c, NNPP_ERR_NONFINITE, "non-finite potential right-hand side");
  const double* bcphi = rhs ? nullptr : c->bcval + (int64_t)S * n;
  const double rtol2 = c->rtol * c->rtol;
  int64_t n64 = 0, n32 = 0;
  int total = 0;
  int status = NNPP_OK;
  double rr = 0.0;
  // r = b - M x in FP64 (START: also z, p, rho, target, converged flag)
  const nnpp_ball_args& ball = c->pst.ball_args;
  auto resid = [&](bool start) -> int {
    if (op.zm) {
      if (start)
        k_zm_resid<true><<<op.g64, NNPP_BLOCK, 0, s>>>(op.zg, ball, c->pfree, c->pst.zmfull, c->pdd, c->pinv,
                                                       c->cg_b, c->cg_x, c->cg_r, c->cg_z,
                                                       c->cg_p, part);
      else
        k_zm_resid<false><<<op.g64, NNPP_BLOCK, 0, s>>>(op.zg, ball, c->pfree, c->pst.zmfull, c->pdd, c->pinv,
                                                        c->cg_b, c->cg_x, c->cg_r, nullptr,
                                                        nullptr, part);
    } else if (start)
      launch_resid<true>(var, op.lds64, op.g64, s, n, sl, op.w64, c->pdd, c->pinv, c->cg_b, c->cg_x,
                         c->cg_r, c->cg_z, c->cg_p, part);
    else
      launch_resid<false>(var, op.lds64, op.g64, s, n, sl, op.w64, c->pdd, c->pinv, c->cg_b, c->cg_x,
                          c->cg_r, nullptr, nullptr, part);
    NNPP_HIP(c, hipGetLastError());
    n64++;
    int e = nnpp_finalize(c, start ? FIN_RESID_START : FIN_RESID, op.g64, rtol2);
    if (e) return e;
    return read_scalars(c);
  };
  // CG iterations: SpMV (partials p.Ap into part_s), update (alpha from them, partials r.z and
  // r.r into part), direction (beta and the stop test from those); rho alternates between two
  // device slots by iteration parity, reset to the first slot by every (re)start
  double* part_s = part + 2 * NNPP_RED_GRID;
  int par = 0;
  if (bnorm2 == 0.0) {  // zero right-hand side: the solution is 0 on the free nodes
    NNPP_HIP(c, hipMemsetAsync(c->cg_x, 0, sizeof(double) * n, s));
  } else if (precision != NNPP_MIXED) {
    if ((st = resid(true))) return st;
    bool done = c->h_scal[SC_DONE] != 0.0;
    BatchPlan plan;
    if (!rhs) plan.prev = &c->cg_phase_len[0];
    double rr_b = c->h_scal[SC_RR];
    int it_b = 0;
    while (!done) {
      const int left = c->maxit - total;
      if (left <= 0) {
        status = NNPP_ERR_NOCONV;
        break;
      }
      const int batch = plan.next(rr_b, c->h_scal[SC_TARGET2], left, total);
      for (int it = 0; it < batch; it++, par ^= 1) {
        if (op.zm)
          k_zm_cg_spmv<<<op.g64, NNPP_BLOCK, 0, s>>>(op.zg, ball, c->pfree, c->pst.zmfull, c->pdd, c->cg_p,
                                                     c->cg_q, part_s, scal);
        else
          launch_cg_spmv(var, op.lds64, op.g64, s, n, sl, op.w64, c->pdd, c->cg_p, c->cg_q,
                         part_s, scal);
        k_cg_update<<<g, NNPP_BLOCK, 0, s>>>(n, c->pinv, c->cg_p, c->cg_q, c->cg_x, c->cg_r,
                                             c->cg_z, part, part_s, op.g64, c->scal, par);
        k_cg_dir<<<g, NNPP_BLOCK, 0, s>>>(n, c->cg_z, c->cg_p, part, g, c->scal, par);
      }
      NNPP_HIP(c, hipGetLastError());
      if ((st = read_scalars(c))) return st;
      total = (int)c->h_scal[SC_ITERS];
      done = c->h_scal[SC_DONE] != 0.0;
      if (!std::isfinite(c->h_scal[SC_RR]))
        return nnpp_set_error(c, NNPP_ERR_NONFINITE, "potential solve diverged (non-finite)");
      plan.observe(rr_b, c->h_scal[SC_RR], total - it_b);
      rr_b = c->h_scal[SC_RR];
      it_b = total;
    }
    n64 += total;
    if (!rhs && status == NNPP_OK) {
      plan.end_phase(total);
      c->cg_phase_len[0] = plan.lens;
    }
    if ((st = resid(false))) return st;  // FP64 residual check of the final iterate
    rr = c->h_scal[SC_RR];
  } else {
    const double sb = 1.0 / std::sqrt(bnorm2);  // FP32 vectors live in units of ||b||
    const double target2 = rtol2 * bnorm2;
    // a refinement pass ends (reliable update) once the FP32 residual norm has dropped by
    // inner_rtol: its square by inner_rtol^2
    const double drop2 = c->inner_rtol * c->inner_rtol;
    if ((st = resid(false))) return st;
    rr = c->h_scal[SC_RR];
    if (rr > target2 && op.zm) {
      // box lattice: the single-reduction iterations (one fused kernel each); r, w and s
      // alternate between two buffers by launch parity, and so do the partials
      float* const Rv[2] = {c->f_r, c->f_z};
      float* const Wv[2] = {c->f_q, c->f_w1};
      float* const Sv[2] = {c->f_s0, c->f_s1};
      double* const Pv[2] = {part, part + 3 * NNPP_RED_GRID};
      const auto launch = [&](bool restart, int q) {
        const ZmCgVec v{Rv[q], Wv[q], Sv[q], Rv[q ^ 1], Wv[q ^ 1], Sv[q ^ 1],
                        c->f_p, c->f_e, c->pinv32};
        const bool v4 = zm_vec4(op.zg);
        auto kern = restart ? (v4 ? k_zm_m_cgs<true, true> : k_zm_m_cgs<true, false>)
                            : (v4 ? k_zm_m_cgs<false, true> : k_zm_m_cgs<false, false>);
        kern<<<op.g32, NNPP_BLOCK, 0, s>>>(op.zg, ball, c->pfree, c->pst.zmfull, c->pdd32nf, v,
                                          Pv[q], op.g32, Pv[q ^ 1], c->scal, q, drop2, 0);
      };
      k_zm_m_replace<true><<<g, NNPP_BLOCK, 0, s>>>(n, c->cg_r, sb, Rv[0], c->f_e, c->f_p,
                                                    Sv[0], Sv[1], part, c->scal);
      NNPP_HIP(c, hipGetLastError());
      if ((st = nnpp_finalize(c, FIN_M_REPLACE_START, g, rtol2))) return st;
      launch(true, 0);
      n32++;
      par = 1;
      BatchPlan plan;
      if (!rhs) plan.prev = &c->cg_phase_len[1];
      double rr_b = rr / bnorm2, stop_b = std::max(rtol2, drop2 * rr_b);
      int it_b = 0;
      while (true) {
        const int left = c->maxit - total;
        if (left <= 0) {
          status = NNPP_ERR_NOCONV;
          break;
        }
        const int batch = plan.next(rr_b, stop_b, left, total);
        for (int it = 0; it < batch; it++, par ^= 1) launch(false, par);
        NNPP_HIP(c, hipGetLastError());
        if ((st = read_scalars(c))) return st;
        total = (int)c->h_scal[SC_ITERS];
        if (!std::isfinite(c->h_scal[SC_RR]))
          return nnpp_set_error(c, NNPP_ERR_NONFINITE, "potential solve diverged (non-finite)");
        plan.observe(rr_b, c->h_scal[SC_RR], total - it_b);
        rr_b = c->h_scal[SC_RR];
        stop_b = std::max(c->h_scal[SC_TARGET2], drop2 * c->h_scal[SC_RRMAX]);
        it_b = total;
        if (c->h_scal[SC_DONE] == 0.0) continue;
        plan.end_phase(total);
        // reliable update: fold the correction into x, recompute the residual in FP64, and
        // continue the recurrence from it (the direction kept)
        k_m_apply<<<nb, NNPP_BLOCK, 0, s>>>(n, 1.0 / sb, c->f_e, c->cg_x);
        NNPP_HIP(c, hipGetLastError());
        if ((st = resid(false))) return st;
        rr = c->h_scal[SC_RR];
        if (!std::isfinite(rr))
          return nnpp_set_error(c, NNPP_ERR_NONFINITE, "potential solve diverged (non-finite)");
        if (rr <= target2) break;
        k_zm_m_replace<false><<<g, NNPP_BLOCK, 0, s>>>(n, c->cg_r, sb, Rv[0], c->f_e, c->f_p,
                                                       Sv[0], Sv[1], part, c->scal);
        NNPP_HIP(c, hipGetLastError());
        if ((st = nnpp_finalize(c, FIN_M_REPLACE, g, rtol2))) return st;
        launch(true, 0);
        n32++;
        par = 1;
        rr_b = rr / bnorm2;  // the replaced residual starts the next phase
        stop_b = std::max(rtol2, drop2 * rr_b);
      }
      n32 += total;
      if (!rhs && status == NNPP_OK) c->cg_phase_len[1] = plan.lens;
      if (status != NNPP_OK) {  // keep the progress made so far
        k_m_apply<<<nb, NNPP_BLOCK, 0, s>>>(n, 1.0 / sb, c->f_e, c->cg_x);
        NNPP_HIP(c, hipGetLastError());
      }
    } else if (rr > target2) {
      k_m_replace<true><<<g, NNPP_BLOCK, 0, s>>>(n, c->cg_r, sb, c->pinv32, c->f_e, c->f_r,
                                                 c->f_z, c->f_p, part);
      NNPP_HIP(c, hipGetLastError());
      if ((st = nnpp_finalize(c, FIN_M_REPLACE_START, g, rtol2))) return st;
      par = 0;
      BatchPlan plan;
      if (!rhs) plan.prev = &c->cg_phase_len[1];
      double rr_b = rr / bnorm2, stop_b = std::max(rtol2, drop2 * rr_b);
      int it_b = 0;
      while (true) {
        const int left = c->maxit - total;
        if (left <= 0) {
          status = NNPP_ERR_NOCONV;
          break;
        }
        const int batch = plan.next(rr_b, stop_b, left, total);
        for (int it = 0; it < batch; it++, par ^= 1) {
          if (op.zm)
            (zm_vec4(op.zg) ? k_zm_m_spmv<true> : k_zm_m_spmv<false>)<<<op.g32, NNPP_BLOCK, 0, s>>>(op.zg, ball, c->pfree, c->pst.zmfull, c->pdd32nf, c->f_p,
                                                      c->f_q, part_s, scal);
          else
            launch_m_spmv(var, op.lds32, op.g32, s, n, sl, op.w32, c->pdd32, c->f_p, c->f_q,
                          part_s, scal);
          k_m_update<<<g, NNPP_BLOCK, 0, s>>>(n, c->pinv32, c->f_p, c->f_q, c->f_e, c->f_r,
                                              c->f_z, part, part_s, op.g32, c->scal, par);
          k_m_dir<<<g, NNPP_BLOCK, 0, s>>>(n, c->f_z, c->f_p, part, g, c->scal, par,
                                           drop2);
        }
        NNPP_HIP(c, hipGetLastError());
        if ((st = read_scalars(c))) return st;
        total = (int)c->h_scal[SC_ITERS];
        if (!std::isfinite(c->h_scal[SC_RR]))
          return nnpp_set_error(c, NNPP_ERR_NONFINITE, "potential solve diverged (non-finite)");
        plan.observe(rr_b, c->h_scal[SC_RR], total - it_b);
        rr_b = c->h_scal[SC_RR];
        stop_b = std::max(c->h_scal[SC_TARGET2], drop2 * c->h_scal[SC_RRMAX]);
        it_b = total;
        if (c->h_scal[SC_DONE] == 0.0) continue;
        plan.end_phase(total);
        // reliable update: fold the correction into x, recompute the residual in FP64
        k_m_apply<<<nb, NNPP_BLOCK, 0, s>>>(n, 1.0 / sb, c->f_e, c->cg_x);
        NNPP_HIP(c, hipGetLastError());
        if ((st = resid(false))) return st;
        rr = c->h_scal[SC_RR];
        if (!std::isfinite(rr))
          return nnpp_set_error(c, NNPP_ERR_NONFINITE, "potential solve diverged (non-finite)");
        if (rr <= target2) break;
        k_m_dir_force<<<nb, NNPP_BLOCK, 0, s>>>(n, c->f_z, c->f_p, scal);
        k_m_replace<false><<<g, NNPP_BLOCK, 0, s>>>(n, c->cg_r, sb, c->pinv32, c->f_e, c->f_r,
                                                    c->f_z, c->f_p, part);
        NNPP_HIP(c, hipGetLastError());
        if ((st = nnpp_finalize(c, FIN_M_REPLACE, g, rtol2))) return st;
        par = 0;
        rr_b = rr / bnorm2;  // the replaced residual starts the next phase
        stop_b = std::max(rtol2, drop2 * rr_b);
      }
      n32 += total;
      if (!rhs && status == NNPP_OK) c->cg_phase_len[1] = plan.lens;
      if (status != NNPP_OK) {  // keep the progress made so far
        k_m_apply<<<nb, NNPP_BLOCK, 0, s>>>(n, 1.0 / sb, c->f_e, c->cg_x);
        NNPP_HIP(c, hipGetLastError());
      }
    }
  }
  k_scatter_phi<<<nb, NNPP_BLOCK, 0, s>>>(n, S, c->pfree, c->bcmask, bcphi, c->cg_x, phi);
  NNPP_HIP(c, hipGetLastError());
  NNPP_HIP(c, hipStreamSynchronize(s));
  if (iters) *iters = total;
  if (relres) *relres = bnorm2 > 0 ? std::sqrt(rr / bnorm2) : 0.0;
  if (spmv64) *spmv64 += n64;
  if (spmv32) *spmv32 += n32;
  return status;
}

int nnpp_zm_prepare(nnpp_ctx* c) {
  nnpp_stencil& T = c->pst;
  const ZmGeom g = zm_geom(T);
  const int64_t m = (int64_t)g.tiles_x * g.tiles_y * g.nz;
  if (!T.zmfull) NNPP_HIP(c, hipMalloc((void**)&T.zmfull, (size_t)(m ? m : 1)));
  if (m)
    k_zm_full<<<(m + NNPP_BLOCK - 1) / NNPP_BLOCK, NNPP_BLOCK, 0, c->stream>>>(g, c->pfree,
                                                                              T.zmfull);
  if (c->n)
    k_zm_ddnf<<<nnpp_nblocks(c->n), NNPP_BLOCK, 0, c->stream>>>(g, T.ball_args, c->pfree,
                                                                c->pdd32, c->pdd32nf);
  NNPP_HIP(c, hipGetLastError());
  return NNPP_OK;
}

int nnpp_phi_guess_dev(nnpp_ctx* c, int hist, const double* phi, const double* prev,
                       const double* prev2, double* out) {
  if (!c->n) return NNPP_OK;
  k_phi_guess<<<nnpp_nblocks(c->n), NNPP_BLOCK, 0, c->stream>>>(c->n, hist, c->pfree, phi, prev,
                                                                 prev2, out);
  NNPP_HIP(c, hipGetLastError());
  return NNPP_OK;
}

// One fused inner iteration of the mixed solve's single-reduction CG on a box lattice, without
// its stop test and scalar updates (the kernel timing hook; the vectors' values are arbitrary).
// NNPP_ERR_ARG where the solve has no such kernel (FP64, or not a box lattice).
int nnpp_cg_iter_launch(nnpp_ctx* c, int precision) {
  if (!c->n) return NNPP_OK;
  const PotOp op = pot_op(c);
  if (!op.zm || precision != NNPP_MIXED)
    return nnpp_set_error(c, NNPP_ERR_ARG,
                          "the fused CG iteration runs in mixed precision on box lattices only");
  const ZmCgVec v{c->f_z, c->f_w1, c->f_s1, c->f_r, c->f_q, c->f_s0, c->f_p, c->f_e, c->pinv32};
  (zm_vec4(op.zg) ? k_zm_m_cgs<false, true> : k_zm_m_cgs<false, false>)<<<op.g32, NNPP_BLOCK, 0, c->stream>>>(
      op.zg, c->pst.ball_args, c->pfree, c->pst.zmfull, c->pdd32nf, v,
      c->red_part + 3 * NNPP_RED_GRID, op.g32, c->red_part, c->scal, 1, 0.0, 1);
  NNPP_HIP(c, hipGetLastError());
  return NNPP_OK;
}

// One potential-operator application (with its dot product) on the solver's direction vector;
// used by the kernel timing hook. The converged flag is cleared first by the caller.
int nnpp_pot_spmv_launch(nnpp_ctx* c, int precision) {
  const int32_t n = c->n;
  if (!n) return NNPP_OK;
  const PotOp op = pot_op(c);
  const Slices& sl = op.sl;
  const int var = op.var;
  const nnpp_ball_args& ball = c->pst.ball_args;
  if (op.zm && precision == NNPP_MIXED)
    (zm_vec4(op.zg) ? k_zm_m_spmv<true> : k_zm_m_spmv<false>)<<<op.g32, NNPP_BLOCK, 0, c->stream>>>(op.zg, ball, c->pfree, c->pst.zmfull, c->pdd32nf, c->f_p,
                                                      c->f_q, c->red_part, (const double*)c->scal);
  else if (op.zm)
    k_zm_cg_spmv<<<op.g64, NNPP_BLOCK, 0, c->stream>>>(op.zg, ball, c->pfree, c->pst.zmfull, c->pdd, c->cg_p,
                                                       c->cg_q, c->red_part, (const double*)c->scal);
  else if (precision == NNPP_MIXED)
    launch_m_spmv(var, op.lds32, op.g32, c->stream, n, sl, op.w32, c->pdd32, c->f_p, c->f_q, c->red_part,
                  (const double*)c->scal);
  else
    launch_cg_spmv(var, op.lds64, op.g64, c->stream, n, sl, op.w64, c->pdd, c->cg_p, c->cg_q, c->red_part,
                   (const double*)c->scal);
  NNPP_HIP(c, hipGetLastError());
  return NNPP_OK;
}

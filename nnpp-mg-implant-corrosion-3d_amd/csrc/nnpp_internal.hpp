// nnpp_internal.hpp — context layout and shared device helpers of libnnpp (gfx950).
//
// HBM layout (all arrays node-indexed in the caller's node order):
//   x, y, z                 FP64 SoA coordinates
//   vol                     FP64 per-node volume (optional)
//   row_ptr/col             CSR neighbour list (API-facing, sorted by id, self excluded)
//   bond slices             "sliced ELL" with a slice height of one wavefront (64 rows):
//                           slice c holds rows 64c..64c+63; entry t of lane l sits at
//                           off[c] + 64 t + l, so the 64 lanes of a wave read 256 contiguous
//                           bytes of column ids per bond step. Padding entries hold -1.
//   bond weights            Beta(xi) V_j / xi^(3+2s) per slice entry, FP64 and FP32
//   U[slot]                 F = S+1 fields, field-major FP64 (U[f*n + i])
//   phase, bcmask, bcval    uint8 phase, uint16 per-node Dirichlet field mask, FP64 values
//   potential slices        same slice format restricted to free-free liquid bonds, weights
//                           pre-scaled by the row's volume ratio; diagonal separately
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/nnpp.h"

#define NNPP_WAVE 64
#define NNPP_BLOCK 256
#define NNPP_NXCD 8

struct nnpp_slices {
  int64_t nslice = 0;      // number of 64-row slices
  int64_t slots = 0;       // total entries (incl. padding)
  int64_t nnz = 0;         // real entries
  int64_t* off = nullptr;  // [nslice + 1]
  int32_t* len = nullptr;  // [nslice] entries per lane (quad layout: groups of 4 per lane)
  int32_t* col = nullptr;  // [slots]
  double* w64 = nullptr;   // [slots]
  float* w32 = nullptr;    // [slots]
  int64_t cap_slots = 0;
  // quad layout (potential operator): entry t of row i sits at
  //   off[i/64] + 256 (t/4) + 4 (i%64) + t%4,
  // so one lane reads its next 4 entries with one 8/16/32-byte load; padding has weight 0
  // and repeats the last column. Columns are also kept as 16-bit steps from the previous
  // entry of the row (first entry: from first[i]) unless a step exceeds 65535 (wide = 1).
  int32_t* first = nullptr;  // [n] column before entry 0 (the row's first column, or i)
  uint16_t* d16 = nullptr;   // [slots]
  int wide = 0;
  // Lossless weight dictionary: when the operator has at most NNPP_DICT_MAX - 1 distinct FP64
  // weights (a regular lattice has a handful), entry weights are stored as 1-byte classes into
  // tab64/tab32 (class 0 = padding, weight 0). ncls = 0: no dictionary, w64/w32 are read.
  uint8_t* cls = nullptr;    // [slots]
  double* tab64 = nullptr;   // [NNPP_DICT_MAX]
  float* tab32 = nullptr;    // [NNPP_DICT_MAX]
  int ncls = 0;
};
#define NNPP_DICT_MAX 256

// Stencil form of the potential operator, used when every bond (i, j) of the neighbour lists
// has j - i in a table of at most NNPP_STENCIL_MAX offsets and a weight that depends only on the
// offset (a lattice numbered in any fixed order, constant node volume). Row i then needs one
// 128-bit mask (bit t: entry i + toff[t] present) instead of its column and weight streams.
// The offsets are sorted ascending (the order of the neighbour lists). A 256-row tile of the
// SpMV reads, for offsets a..b, the vector range [r0 + a, r0 + b + 255]; offsets whose ranges
// overlap or touch form one segment, whose range the tile stages into LDS once (on a lattice
// numbered x-fastest, one segment per z-plane of the stencil), and lane l reads entry t at
// lds[l + lc[t]]. The staging walks the segments in pieces of at most 256 values, one value per
// thread and piece.
#define NNPP_STENCIL_MAX 128
#define NNPP_STENCIL_CHUNK 32  // tables padded to a multiple of the SpMV chunks (32 / 16)
#define NNPP_PIECE_BATCH 8     // staging pieces per batch; the piece table is padded to whole
                               // batches with empty pieces (count 0)

// Lattice ball of radius M: the offsets (dz, dy, dx) with dx^2 + dy^2 + dz^2 <= M^2 other than
// (0, 0, 0), in lexicographic order, which is ascending order of dz*SZ + dy*SY + dx on a
// lattice with row strides SY > 2M and SZ > 2M (SY + 1). Entry t lies on line `line[t]` (the
// offsets of one (dz, dy), consecutive), `pos[t]` places after the line's first entry.
template <int M>
struct nnpp_ball {
  static constexpr int count(bool lines) {
    int k = 0;
    for (int z = -M; z <= M; z++)
      for (int y = -M; y <= M; y++) {
        if (z * z + y * y > M * M) continue;
        if (lines) k++;
        else
          for (int x = -M; x <= M; x++)
            if (z * z + y * y + x * x <= M * M && (x | y | z) != 0) k++;
      }
    return k;
  }
  static constexpr int K = count(false), L = count(true);
  int line[K], pos[K], dx[K], d2[K];
  int ldz[L], ldy[L];
  constexpr nnpp_ball() : line(), pos(), dx(), d2(), ldz(), ldy() {
    int t = 0, l = 0;
    for (int z = -M; z <= M; z++)
      for (int y = -M; y <= M; y++) {
        if (z * z + y * y > M * M) continue;
        ldz[l] = z;
        ldy[l] = y;
        int x0 = M + 1;
        for (int x = -M; x <= M; x++) {
          if (z * z + y * y + x * x > M * M || (x | y | z) == 0) continue;
          if (x0 > M) x0 = x;
          line[t] = l;
          pos[t] = x - x0;
          dx[t] = x;
          d2[t] = z * z + y * y + x * x;
          t++;
        }
        l++;
      }
  }
};
// the radius the potential SpMV is specialised for (the lattice case with horizon 3 spacings)
#define NNPP_BALL_M 3
#define NNPP_BALL_LINES (nnpp_ball<NNPP_BALL_M>::L)
#define NNPP_BALL_CLASSES (NNPP_BALL_M * NNPP_BALL_M + 1)
// LDS position of each line's first entry and the weights by dx^2 + dy^2 + dz^2
struct nnpp_ball_args {
  int32_t lc[NNPP_BALL_LINES];
  float w32[NNPP_BALL_CLASSES];
  double w64[NNPP_BALL_CLASSES];
};
struct nnpp_stencil {
  int state = 0;               // 0: not built for this geometry, 1: valid, -1: not a stencil
  int K = 0, R = 0, P = 0;     // offsets, segments, staging pieces
  int off_min = 0, off_max = 0;  // toff[0], toff[K-1]
  int lds_elems = 0;           // LDS entries per 256-row tile
  int32_t* toff = nullptr;     // [K] j - i, ascending
  int32_t* lc = nullptr;       // [K rounded up to NNPP_STENCIL_CHUNK] LDS position of entry t
                               // relative to the lane (padding: 0)
  int32_t* seg_start = nullptr;  // [R] first offset of segment r
  int32_t* seg_len = nullptr;    // [R] last offset - first offset + 1
  int32_t* seg_base = nullptr;   // [R] LDS base of segment r
  int32_t* piece = nullptr;      // [3P, padded to whole NNPP_PIECE_BATCH batches] per piece:
                                 // vector position relative to the tile's first row, LDS
                                 // position, value count (1..256; 0 in the padding)
  double* tw64 = nullptr;      // [K rounded up to NNPP_STENCIL_CHUNK] potential weight of
  float* tw32 = nullptr;       // offset t (FP64 / FP32; padding: 0)
  double* tf64 = nullptr;      // [K rounded up to NNPP_STENCIL_CHUNK] flux weight of offset t
  float* tf32 = nullptr;       // (flux exponent; FP64 / FP32; padding: 0)
  int flux_ok = 0;             // every bond's flux weight equals its offset's (bit for bit)
  uint32_t fnz[4] = {0u, 0u, 0u, 0u};  // offsets with a nonzero flux weight
  uint4* bmask = nullptr;      // [n] bonds of each row
  uint8_t* fperm = nullptr;    // [256 per tile] lane order of the stencil flux: rows with every
                               // bond first (as perm, by bmask)
  uint4* pmask = nullptr;      // [n] potential-operator entries of each row
  uint8_t* perm = nullptr;     // [256 per tile] row of each lane: rows with every entry first,
                               // then the other rows with entries, then the rest (stable)
  int active = 0;              // the current potential operator uses the stencil form
  int ball = 0;  // the offsets, LDS positions and weights fit the NNPP_BALL_M ball (bit for bit)
  nnpp_ball_args ball_args = {};
  int fball = 0;  // ball == NNPP_BALL_M and the flux weights depend on dx^2 + dy^2 + dz^2 only
  float fball_w32[NNPP_BALL_CLASSES] = {};  // (bit for bit): the FP32 flux weight by class
  double fball_w64[NNPP_BALL_CLASSES] = {};  // the FP64 flux weight by class
  double fball_wsum = 0.0;  // sum of the FP32 flux weights of the ball's offsets (in FP64)
  // Box lattice (ball == NNPP_BALL_M only): the nodes are the full nx x ny x nz box numbered
  // x fastest (row strides sy = nx, sz = nx ny, n = nx ny nz) and every row's bonds are exactly
  // the ball offsets whose far end lies inside the box. The z-marching kernels need this: they
  // address neighbours by (x, y, z) and treat positions outside the box as absent.
  int box = 0;
  int nx = 0, ny = 0, nz = 0;
  uint8_t* zmfull = nullptr;  // [column tiles x nz] per tile and plane: every point the
                              // z-marching SpMV stages there is a free node (rebuilt with the
                              // potential operator)
};

// Differentiation tape (nnpp_tape_begin .. nnpp_tape_end).
struct nnpp_tape_step {  // host record of one recorded step
  int64_t step;          // step counter before the step
  double voln1rel;       // voln1rel in force during the step (its BV)
  int32_t flips;         // solid -> liquid flips at its end
  int32_t vol_update;    // voln1rel recomputed at its end
};
struct nnpp_tape_ckpt {  // state before recorded step `every * index`
  double* U = nullptr;   // F*n
  uint8_t* phase = nullptr;
  double* phi_prev = nullptr;  // 2n: the two previous potentials (warm start)
  int phi_hist = 0;             // how many of them are valid
  int64_t step = 0, solid_count = 0;
  double voln1rel = 0.0;
};
struct nnpp_tape {
  bool active = false;
  int32_t every = 1;
  std::vector<nnpp_tape_step> steps;
  std::vector<nnpp_tape_ckpt> ckpts;
  int64_t bytes = 0;
};

struct nnpp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  nnpp_params p{};
  bool have_params = false;
  int32_t n = 0;
  double* x = nullptr;
  double* y = nullptr;
  double* z = nullptr;
  double* vol = nullptr;

  // CSR
  int64_t nnz = -1;
  int64_t* row_ptr = nullptr;
  int32_t* col = nullptr;
  int64_t col_cap = 0;
  int32_t max_len = 0;

  nnpp_slices bonds;       // every bond, weights with the flux exponent
  nnpp_slices pot;         // free-free liquid bonds of the potential operator
  nnpp_stencil pst;        // stencil form of the potential operator (lattices)
  bool geom_dirty = true;  // bond slices need a rebuild
  bool pot_dirty = true;   // potential operator needs a rebuild (phase/bc changed)
  int32_t format = 0;      // NNPP_FORMAT_* flags
  double w_exponent = 3.0; // 3 + 2 s shared by every field

  uint8_t* phase = nullptr;
  uint16_t* bcmask = nullptr;
  double* bcval = nullptr;  // F*n, species already divided by CSolid
  bool have_bc = false;

  double* U[2] = {nullptr, nullptr};  // slots N, N1
  double* Unew = nullptr;             // F*n scratch of the explicit step
  double* Fres = nullptr;             // F*n residual output
  double* aos = nullptr;              // n * (S+3 rounded to even) AoS gather records

  // potential solve
  uint8_t* pfree = nullptr;
  double* pdd = nullptr;     // scaled weight sum of Dirichlet liquid neighbours (FP64)
  float* pdd32 = nullptr;
  float* pdd32nf = nullptr;  // box lattice: pdd32 less the weights of the row's bonds to
                             // non-free or out-of-box nodes in the box's z range (FP32 CG)
  double* pinv = nullptr;    // inverse of the scaled diagonal, 0 on non-free rows
  float* pinv32 = nullptr;
  double* pbdir = nullptr;   // Dirichlet part of the right-hand side (unscaled)
  double* psc = nullptr;     // row scale V_i / (dx dy dz)
  double* cg_b = nullptr;
  double* cg_x = nullptr;
  double* cg_r = nullptr;
  double* cg_z = nullptr;
  double* cg_p = nullptr;
  double* cg_q = nullptr;
  float* f_e = nullptr;      // FP32 inner-solve vectors (mixed mode)
  float* f_r = nullptr;
  float* f_z = nullptr;
  float* f_p = nullptr;
  float* f_q = nullptr;
  float* f_w1 = nullptr;     // single-reduction FP32 CG: second buffers of w and s (A p)
  float* f_s0 = nullptr;
  float* f_s1 = nullptr;
  int64_t pot_nfree = 0;

  // reductions: per-block partials + scalars; pinned host mirror of the scalars
  double* red_part = nullptr;   // [NNPP_RED_GRID * 4]
  double* scal = nullptr;       // device scalars (see enum CgScalar)
  double* h_scal = nullptr;     // pinned host copy

  double rtol = 1e-10;
  int32_t maxit = 20000;
  // iterations of each phase (between residual replacements) of the last potential solve of
  // the step kind, per precision (0 FP64, 1 mixed): the next solve sizes its launch batches by
  // them (successive steps repeat the pattern closely)
  std::vector<int32_t> cg_phase_len[2];
  double inner_rtol = 1e-3;  // residual-norm reduction per FP32 refinement pass

  // implicit (reference Newton) steps, allocated on first use
  double* imp = nullptr;       // species vectors, S*n each: x r rh p v s t d0 diag
  double* imp_prev = nullptr;  // F*n: U^{n-1} of the iterate rotation
  double* imp_dphi = nullptr;  // 3*n: potential update, Dirichlet part, coupling
  double* imp_part = nullptr;  // [NNPP_RED_GRID * 32] block partials
  double* imp_red = nullptr;   // [32] totals
  double* h_imp = nullptr;     // pinned host copy of the totals
  int32_t* imp_flag = nullptr;  // set when the Dirichlet part of the potential update is nonzero
  int64_t imp_n = -1;          // n and S the buffers were allocated for
  int imp_S = 0;
  bool imp_first = true;       // next implicit step starts from U^n with the Dirichlet values

  // potentials of the two previous explicit steps ([phi^{n-1}, phi^{n-2}]), for the
  // extrapolation that warm-starts the next potential solve; phi_hist of them are valid (0
  // after anything but an explicit step changed the state)
  double* phi_prev = nullptr;
  int phi_hist = 0;

  int64_t step = 1;
  double voln1rel = 0.0;
  double vol0rel = 0.0;
  int64_t solid_count = 0;

  nnpp_tape tape;

  // device address of the static scratch array of nnpp_implant_classify_noalloc on this
  // context's device (hipGetSymbolAddress at nnpp_create)
  void* noalloc_scratch = nullptr;
};

// Device scalars used by the CG kernels and reductions.
enum CgScalar {
  SC_RHO = 0, SC_ALPHA, SC_BETA, SC_PQ, SC_RR, SC_TARGET2, SC_DONE, SC_ITERS, SC_BNORM2,
  SC_R0, SC_R1, SC_R2, SC_R3, SC_RRMAX,
  SC_RHO1,  // second slot of rho: CG iterations alternate between SC_RHO and SC_RHO1
  // single-reduction FP32 CG (z-marching kernels): gamma = r.u and alpha of the last
  // iteration, one pair per launch parity, and the parity of the launch that stopped
  SC_CG_GAMMA0, SC_CG_ALPHA0, SC_CG_GAMMA1, SC_CG_ALPHA1, SC_CG_PARDONE,
  SC_COUNT
};

int nnpp_set_error(nnpp_ctx* c, int code, const std::string& msg);

#define NNPP_HIP(ctx, call)                                                              \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return nnpp_set_error((ctx), NNPP_ERR_HIP,                                         \
                            std::string(#call) + ": " + hipGetErrorString(e_));          \
  } while (0)

// ---------------------------------------------------------------- device helpers -------

// Blocks are handed to the 8 XCDs round-robin; remap so each XCD walks a contiguous range of
// node blocks (neighbour data then stays in that XCD's L2). Bijective for any block count.
__device__ __forceinline__ int64_t nnpp_xcd_block(int64_t b, int64_t nb) {
  const int64_t q = nb / NNPP_NXCD, r = nb % NNPP_NXCD;
  const int64_t xcd = b % NNPP_NXCD, k = b / NNPP_NXCD;
  return xcd * q + (xcd < r ? xcd : r) + k;
}

// Non-temporal loads of streams read once per kernel (slice columns and weights), so they do
// not push gathered vectors out of L2.
typedef unsigned int nnpp_u32x2 __attribute__((ext_vector_type(2)));
typedef double nnpp_f64x2 __attribute__((ext_vector_type(2)));
typedef float nnpp_f32x4 __attribute__((ext_vector_type(4)));
typedef int nnpp_i32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint2 nnpp_ld_stream(const uint2* p) {
  const nnpp_u32x2 v = __builtin_nontemporal_load(reinterpret_cast<const nnpp_u32x2*>(p));
  return make_uint2(v.x, v.y);
}
__device__ __forceinline__ double2 nnpp_ld_stream(const double2* p) {
  const nnpp_f64x2 v = __builtin_nontemporal_load(reinterpret_cast<const nnpp_f64x2*>(p));
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ float4 nnpp_ld_stream(const float4* p) {
  const nnpp_f32x4 v = __builtin_nontemporal_load(reinterpret_cast<const nnpp_f32x4*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ int4 nnpp_ld_stream(const int4* p) {
  const nnpp_i32x4 v = __builtin_nontemporal_load(reinterpret_cast<const nnpp_i32x4*>(p));
  return make_int4(v.x, v.y, v.z, v.w);
}

// utils.h:17-25
__device__ __forceinline__ double nnpp_beta(double xi, double delta, double dx) {
  if (xi <= delta - dx / 2.0) return 1.0;
  if (xi <= delta + dx / 2.0) return (delta + dx / 2.0 - xi) / dx;
  return 0.0;
}

__device__ __forceinline__ double nnpp_wave_sum(double v) {
  for (int o = NNPP_WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, NNPP_WAVE);
  return v;
}

// Block-wide sum of K values (blockDim.x == NNPP_BLOCK); result valid in thread 0.
template <int K>
__device__ __forceinline__ void nnpp_block_sum(double (&v)[K], double* lds /*[K*4]*/) {
  const int lane = threadIdx.x % NNPP_WAVE, wid = threadIdx.x / NNPP_WAVE;
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = nnpp_wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; k++) lds[k * (NNPP_BLOCK / NNPP_WAVE) + wid] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; k++) {
      double s = 0.0;
      for (int w = 0; w < NNPP_BLOCK / NNPP_WAVE; w++) s += lds[k * (NNPP_BLOCK / NNPP_WAVE) + w];
      v[k] = s;
    }
}

// Load of a value another block of the SAME kernel may have written (grid reductions).
__device__ __forceinline__ double nnpp_ld_coherent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Load of a device scalar written by an EARLIER kernel on the same stream: caches are
// invalidated at every kernel start, so a plain load is coherent; with a uniform address it
// becomes one scalar-cache load per wavefront instead of a per-lane L2 access to one line.
__device__ __forceinline__ double nnpp_ld_scalar(const double* __restrict__ p) { return *p; }

// Reductions run on a fixed grid of at most NNPP_RED_GRID blocks (a multiple of 8); every
// block walks several 256-node blocks. XCD x (= physical block % 8) owns the contiguous range
// [x*per, (x+1)*per) of node blocks, so neighbour data stays in its L2.
#define NNPP_RED_GRID 2048

// Load from a table that no kernel writes while it reads it, through the constant address
// space: with a wave-uniform index it becomes a scalar load (batched by the compiler) instead of
// a vector load of one address.
template <typename T>
__device__ __forceinline__ T nnpp_ld_table(const T* p, int i) {
  return ((const __attribute__((address_space(4))) T*)p)[i];
}

// w where keep is all ones, +0 where keep is 0. The weight must be wave-uniform (a table value
// read with nnpp_ld_table), the keep mask is per lane: one v_and_b32 per 32 bits with the weight
// as the scalar operand. Written out because the compiler otherwise turns the AND with a
// sign-extended bit into a compare and a select (twice the instructions).
template <typename T>
__device__ __forceinline__ T nnpp_mask_weight(T w, int keep);
template <>
__device__ __forceinline__ float nnpp_mask_weight<float>(float w, int keep) {
  float r;
  asm("v_and_b32 %0, %1, %2" : "=v"(r) : "s"(w), "v"(keep));
  return r;
}
template <>
__device__ __forceinline__ double nnpp_mask_weight<double>(double w, int keep) {
  const uint64_t b = (uint64_t)__double_as_longlong(w);
  uint32_t lo, hi;
  asm("v_and_b32 %0, %1, %2" : "=v"(lo) : "s"((uint32_t)b), "v"(keep));
  asm("v_and_b32 %0, %1, %2" : "=v"(hi) : "s"((uint32_t)(b >> 32)), "v"(keep));
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

struct nnpp_block_iter {
  int64_t cur, end, stride;
};

__device__ __forceinline__ nnpp_block_iter nnpp_blocks_of(int64_t nb) {
  const int64_t g = blockIdx.x, G = gridDim.x;
  nnpp_block_iter it;
  if (G >= nb) {  // one node block per physical block
    it.cur = nnpp_xcd_block(g, nb);
    it.end = it.cur + 1;
    it.stride = 1;
  } else {
    const int64_t per = (nb + NNPP_NXCD - 1) / NNPP_NXCD, x = g % NNPP_NXCD;
    it.cur = x * per + g / NNPP_NXCD;
    it.end = (x + 1) * per < nb ? (x + 1) * per : nb;
    it.stride = G / NNPP_NXCD;
  }
  return it;
}

static inline int64_t nnpp_red_grid(int64_t nb) {
  return nb <= NNPP_RED_GRID ? nb : NNPP_RED_GRID;
}

// Grid reductions in two launches. The producing kernel calls nnpp_block_partial (plain
// stores of one partial per block) and a one-block finalize kernel (nnpp_finalize) sums them in
// block order and updates the device scalars. No block waits for or reads another block's data
// inside a kernel: on gfx950 the 8 XCDs have separate L2s, and an in-kernel agent-scope fence
// per block would write back and invalidate its XCD's L2. The kernel boundary orders the partial
// stores before the finalize kernel at the cost of one launch.
template <int K>
__device__ __forceinline__ void nnpp_block_partial(double (&v)[K], double* __restrict__ part) {
  __shared__ double lds[K * (NNPP_BLOCK / NNPP_WAVE)];
  nnpp_block_sum<K>(v, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; k++) part[(int64_t)blockIdx.x * K + k] = v[k];
}

// Sum of the partials of nblk (<= NNPP_RED_GRID) blocks in a fixed order; one block of
// NNPP_BLOCK threads; the totals are valid in thread 0.
template <int K>
__device__ __forceinline__ void nnpp_final_sum(double (&v)[K], const double* __restrict__ part,
                                               int64_t nblk) {
  __shared__ double lds[K * (NNPP_BLOCK / NNPP_WAVE)];
  constexpr int PER = NNPP_RED_GRID / NNPP_BLOCK;
  // every load is issued (from block 0 past nblk, then replaced by 0) before any sum, so the
  // K * PER loads take one memory round trip; loads guarded by b < nblk became one branch each
  // and the first sum waited inside it
  double t[K][PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int64_t b = threadIdx.x + (int64_t)q * NNPP_BLOCK;
    const int64_t bb = b < nblk ? b : 0;
#pragma unroll
    for (int k = 0; k < K; k++) t[k][q] = part[bb * K + k];
  }
#pragma unroll
  for (int k = 0; k < K; k++) {
    double a = 0.0;
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int64_t b = threadIdx.x + (int64_t)q * NNPP_BLOCK;
      a += (b < nblk) ? t[k][q] : 0.0;
    }
    v[k] = a;
  }
  nnpp_block_sum<K>(v, lds);
}

// As nnpp_final_sum, with the totals in every thread of the block. Every block of a kernel that
// calls this on the same partials obtains the same totals (same order of additions).
template <int K>
__device__ __forceinline__ void nnpp_final_sum_all(double (&v)[K],
                                                   const double* __restrict__ part, int64_t nblk) {
  __shared__ double tot[K];
  nnpp_final_sum<K>(v, part, nblk);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; k++) tot[k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = tot[k];
}

// Finalize operations (what the one-block kernel does with the summed partials).
enum NnppFin {
  FIN_BNORM2 = 0,     // K=1: BNORM2
  FIN_RESID,          // K=2: RR
  FIN_RESID_START,    // K=2: RHO, RR, TARGET2 = a0 * BNORM2, ITERS = 0, DONE
  FIN_M_REPLACE,      // K=2: RHO, RR, RRMAX, TARGET2 = a0, DONE = 0
  FIN_M_REPLACE_START,// K=2: as above, ITERS = 0, DONE = (RR <= a0)
  FIN_FRONT,          // K=3: R0, R1, R2
  FIN_NONFINITE       // K=1: R3
};

__host__ __device__ static inline int64_t nnpp_nblocks(int64_t n) { return (n + NNPP_BLOCK - 1) / NNPP_BLOCK; }

// ---------------------------------------------------------------- launchers (host) -----
int nnpp_build_neighbors_dev(nnpp_ctx* c, double r, double cdx, double cdy, double cdz,
                             int32_t max_len, int64_t* nnz_out, int32_t* max_len_out);
int nnpp_build_bond_slices(nnpp_ctx* c);
int nnpp_implant_count_dev(nnpp_ctx* c, int64_t nb, const double* bx, const double* by,
                           const double* bz, const double bbox[6], double r, int32_t max_count,
                           int32_t* counts);
int nnpp_voronoi_dev(nnpp_ctx* c, double hp, double margin, double radius,
                     unsigned long long* count, int64_t dims[3]);
int nnpp_build_potential(nnpp_ctx* c);
int nnpp_launch_flux(nnpp_ctx* c, int resid, int precision, double BV, const double* U1,
                     const double* Un, double* out, uint16_t* passmask = nullptr);
// rhs == nullptr: b from the charge of Uq, Dirichlet nodes of phi set to their values;
// rhs given: b = rhs on the free rows (the adjoint solve), only free nodes of phi written.
int nnpp_potential_solve_dev(nnpp_ctx* c, int precision, const double* Uq, double* phi,
                             int32_t* iters, double* relres, int64_t* spmv64, int64_t* spmv32,
                             const double* rhs = nullptr);
// any (device flag, optional): when it reads 0, `in` is all zeros and out is set to 0.
int nnpp_potential_apply_dev(nnpp_ctx* c, const double* in, double* out,
                             const int32_t* any = nullptr);
// Warm start of the potential solve on the free rows (phi elsewhere): hist 2 -> quadratic
// extrapolation 3 phi - 3 prev + prev2, hist 1 -> linear 2 phi - prev, hist 0 -> phi.
int nnpp_phi_guess_dev(nnpp_ctx* c, int hist, const double* phi, const double* prev,
                       const double* prev2, double* out);
// Species blocks of the frozen Jacobian (utils.h:844-1849) applied to v: out = d0 v + bond
// sums. v is given species-major (vin, S*n; the lattice stencil form reads it) and staged in the
// gather records c->aos ([v_0 .. v_{S-1}, phi^n] per node; the column form reads those). With
// d0_out set, d0 = 1/dt - R' at Un and the block diagonals are written first (d0 is then not
// read).
int nnpp_launch_jac(nnpp_ctx* c, double BV, const double* Un, const double* vin,
                    const double* d0, double* out, double* d0_out = nullptr,
                    double* diag_out = nullptr);
// phi block K11 (utils.h:1794-1848) with its identity rows: out = K11 in, diag = its diagonal.
int nnpp_jac_phi_dev(nnpp_ctx* c, const double* in, double* out, double* diag);
// Implicit (reference Newton) time step; the potential operator must be current.
int nnpp_implicit_once(nnpp_ctx* c, int precision, const nnpp_newton_opts& o,
                       nnpp_newton_stats& st);
// Frozen-Jacobian apply of every field block on device vectors (F*n).
int nnpp_jacobian_apply_dev(nnpp_ctx* c, double voln1rel, const double* in, double* out,
                            double* diag);
void nnpp_implicit_free(nnpp_ctx* c);
int nnpp_front_dev(nnpp_ctx* c, double* U, double thr, int do_vol, int32_t* flips,
                   double* volsum);
int nnpp_bc_override_dev(nnpp_ctx* c, const double* U1, double* Fo);
int nnpp_check_finite_dev(nnpp_ctx* c, const double* U, int32_t* bad);
int nnpp_pot_spmv_launch(nnpp_ctx* c, int precision);
int nnpp_cg_iter_launch(nnpp_ctx* c, int precision);
// Tables of the z-marching potential SpMV for the current free flags (box lattices).
int nnpp_zm_prepare(nnpp_ctx* c);
// Address of the no-allocation scratch array on the context's device (current device).
int nnpp_noalloc_setup(nnpp_ctx* c);
// Initial values of main.cpp:422-457 from implant counts (host, 10 reference species + phi).
void nnpp_implant_icvalues_host(const nnpp_ctx* c, const int32_t* counts, double* ic_out);
int nnpp_finalize(nnpp_ctx* c, int op, int64_t nblk, double a0);
int nnpp_build_dictionary(nnpp_ctx* c, nnpp_slices& sl);
// One explicit step of nnpp_step (flux, potential, front, voln1rel); fills *rec if non-null.
int nnpp_step_once(nnpp_ctx* c, int precision, nnpp_step_stats& loc, nnpp_tape_step* rec);
int nnpp_adjoint_run(nnpp_ctx* c, const double* lam, double* grad, double* lam0,
                     nnpp_adjoint_stats* stats);
void nnpp_tape_free(nnpp_ctx* c);
void nnpp_free_stencil(nnpp_stencil& t);
// Builds the stencil tables of the current geometry if not yet known (c->pst.state 0 -> 1 or
// -1); NNPP_OK either way.
int nnpp_stencil_ensure(nnpp_ctx* c);
// 1 where the explicit flux takes the lattice stencil form (FP64 in both precisions) for the
// current geometry.
int nnpp_flux_stencil_active(nnpp_ctx* c);
int nnpp_flux_zmarch_active(nnpp_ctx* c);
// The potential SpMV's form: 0 column slices, 1 stencil masks with the table-driven loop,
// ------------------------------------------------------------ z-marching geometry -------
// Box lattices (nnpp_stencil::box): the z-marching kernels of the potential operator and of the
// flux give a block of 256 threads a ZM_TX x ZM_TY column tile of (x, y) and a chunk of lz
// output planes; it walks the input planes of the chunk with a halo of ZM_M planes, each
// staged with a ZM_M-node halo in x and y.
constexpr int ZM_TX = 32, ZM_TY = 8;  // a half-wave reads one x-row of 32 (no bank conflicts)
constexpr int ZM_M = NNPP_BALL_M;
constexpr int ZM_W = ZM_TX + 2 * ZM_M, ZM_H = ZM_TY + 2 * ZM_M;
constexpr int ZM_PTS = ZM_W * ZM_H;
constexpr int ZM_STAGE = (ZM_PTS + NNPP_BLOCK - 1) / NNPP_BLOCK;  // staged points per thread
constexpr int ZM_SL = 2 * ZM_M + 1;                               // rows in flight

struct ZmGeom {
  int nx, ny, nz;
  int tiles_x, tiles_y;
  int lz, nzc;       // output planes per block, chunks along z
  int64_t sz;        // nx * ny
  int64_t items;     // tiles_x * tiles_y * nzc (the grid)
};

// Work item of a potential-operator block. The blocks of one XCD take a contiguous range of
// items (nnpp_xcd_block) and the items run chunk-major, so each XCD walks whole z-slabs (every
// column tile of its chunks): the tiles that stage one another's halo rows read the same planes
// at the same time, through that XCD's L2.
__device__ __forceinline__ void zm_item(const ZmGeom& g, int64_t b, int64_t& tile, int& zc) {
  const int64_t item = nnpp_xcd_block(b, g.items);
  const int64_t tiles = (int64_t)g.tiles_x * g.tiles_y;
  zc = (int)(item / tiles);
  tile = item % tiles;
}

// Rings of the disc dx^2 + dy^2 <= M^2: the points of equal k = dx^2 + dy^2.
template <int M>
__device__ __forceinline__ constexpr bool zm_ring_first(int dx, int dy) {
  // (dx, dy) is the first point of its ring in the walk order (dy outer, dx inner, ascending)
  for (int y = -M; y <= M; y++)
    for (int x = -M; x <= M; x++) {
      if (x * x + y * y != dx * dx + dy * dy) continue;
      return x == dx && y == dy;
    }
  return false;
}
template <int M>
__device__ __forceinline__ constexpr bool zm_ring_exists(int k) {
  for (int y = -M; y <= M; y++)
    for (int x = -M; x <= M; x++)
      if (x * x + y * y == k) return true;
  return false;
}

// Work items: column tiles times chunks of lz output planes, lz chosen for about `target` items
// (at most NNPP_RED_GRID: the reduction partials), at least 4 planes.
inline ZmGeom nnpp_zm_geom(const nnpp_stencil& T, int64_t target) {
  ZmGeom g;
  g.nx = T.nx;
  g.ny = T.ny;
  g.nz = T.nz;
  g.tiles_x = (T.nx + ZM_TX - 1) / ZM_TX;
  g.tiles_y = (T.ny + ZM_TY - 1) / ZM_TY;
  const int64_t tiles = (int64_t)g.tiles_x * g.tiles_y;
  int64_t lz = (T.nz * tiles + target - 1) / target;
  lz = std::max<int64_t>(lz, (T.nz * tiles + NNPP_RED_GRID - 1) / NNPP_RED_GRID);
  lz = std::max<int64_t>(lz, 4);
  lz = std::min<int64_t>(lz, T.nz);
  g.lz = (int)lz;
  g.nzc = (int)((T.nz + lz - 1) / lz);
  g.sz = (int64_t)T.nx * T.ny;
  g.items = tiles * g.nzc;
  return g;
}

// NNPP_BALL_M stencil masks with the lattice-ball loop, NNPP_POT_ZMARCH the z-marching kernels
// of a full box lattice.
#define NNPP_POT_ZMARCH 4
inline int nnpp_pot_form(const nnpp_ctx* c) {
  if (!c->pst.active) return 0;
  if (c->pst.ball != NNPP_BALL_M || (c->format & NNPP_FORMAT_NO_BALL)) return 1;
  return (c->pst.box && !(c->format & NNPP_FORMAT_NO_ZMARCH)) ? NNPP_POT_ZMARCH : NNPP_BALL_M;
}

// k_neighbors.hip — spatial-hash neighbour construction (replaces build_clouds_hash,
// grid.h:43-204) and the per-row sort that turns it into id-sorted CSR.
//
// Bit-exactness: the hash (grid.h:84), the cell corners (grid.h:134-139), the sphere-cube test
// (grid.h:27-41) and the candidate predicate sqr(dx)+sqr(dy)+sqr(dz) <= sqr(r) (grid.h:186) are
// evaluated with the same FP64 operations in the same order and without FMA contraction (this
// file is compiled with -ffp-contract=off and the pragma below). The reference tests every
// cell against the sphere (grid.h:163); here a node tests only the cells of the index box one
// cell wider than its sphere's bounding box, which contains every cell the test can accept.
#pragma clang fp contract(off)

#include <climits>
#include <cmath>
#include <vector>

#include "nnpp_internal.hpp"

namespace {

__device__ __forceinline__ double sqr_(double v) { return v * v; }

// grid.h:27-41
__device__ __forceinline__ bool sphere_cube(double cx, double cy, double cz, double cr,
                                            double bx, double by, double bz, double w, double h,
                                            double d) {
  const double bmin[3] = {bx, by, bz};
  const double bmax[3] = {bx + w, by + h, bz + d};
  const double c[3] = {cx, cy, cz};
  const double r2 = sqr_(cr);
  double dmin = 0.0;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (c[a] < bmin[a]) dmin += sqr_(c[a] - bmin[a]);
    else if (c[a] > bmax[a]) dmin += sqr_(c[a] - bmax[a]);
  }
  return dmin <= r2;
}

struct Grid {
  double x0, y0, z0, cdx, cdy, cdz, r;
  int mx, my, mz_last, max_hash;
};

// grid.h:82-85
__global__ void k_hash(int32_t n, const double* __restrict__ x, const double* __restrict__ y,
                       const double* __restrict__ z, Grid g, int* __restrict__ hash,
                       int* __restrict__ max_hash) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int h = (int)((x[i] - g.x0) / g.cdx) + g.mx * (int)((y[i] - g.y0) / g.cdy) +
                g.mx * g.my * (int)((z[i] - g.z0) / g.cdz);
  hash[i] = h;
  atomicMax(max_hash, h);
}

__global__ void k_cell_count(int32_t n, const int* __restrict__ hash, int* __restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  atomicAdd(&cnt[hash[i]], 1);
}

__global__ void k_cell_scatter(int32_t n, const double* __restrict__ x,
                               const double* __restrict__ y, const double* __restrict__ z,
                               const int* __restrict__ hash, int* __restrict__ cursor,
                               double* __restrict__ cx, double* __restrict__ cy,
                               double* __restrict__ cz, int32_t* __restrict__ cid) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int pos = atomicAdd(&cursor[hash[i]], 1);
  cx[pos] = x[i];
  cy[pos] = y[i];
  cz[pos] = z[i];
  cid[pos] = (int32_t)i;
}

// ---- implant classification (kd_tree_interface.h:156-252 as called at main.cpp:316-320) ----
// A uniform grid of cells at least 1.001 r wide over the implant points' bounding box; the
// points are counting-sorted by cell, and a node tests the points of the 3x3x3 cells around
// its own (every point within r of it lies there). The predicate is the reference's
// Node::distance_euclidean: sqrt(x*x + y*y + z*z) <= r with (x, y, z) = point - node.
struct IGrid {
  double x0, y0, z0, h;
  int64_t mx, my, mz;
};

__device__ __forceinline__ int64_t icell_axis(double v, double v0, double h, int64_t m) {
  const double f = floor((v - v0) / h);
  return f < 0.0 ? 0 : (f > (double)(m - 1) ? m - 1 : (int64_t)f);
}

__global__ void k_imp_cell(int64_t nb, const double* __restrict__ bx,
                           const double* __restrict__ by, const double* __restrict__ bz,
                           IGrid g, int32_t* __restrict__ cell, int32_t* __restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const int64_t c = icell_axis(bx[i], g.x0, g.h, g.mx) +
                    g.mx * (icell_axis(by[i], g.y0, g.h, g.my) +
                            g.my * icell_axis(bz[i], g.z0, g.h, g.mz));
  cell[i] = (int32_t)c;
  atomicAdd(&cnt[c], 1);
}

__global__ void k_imp_scatter(int64_t nb, const double* __restrict__ bx,
                              const double* __restrict__ by, const double* __restrict__ bz,
                              const int32_t* __restrict__ cell, const int64_t* __restrict__ start,
                              int32_t* __restrict__ cursor, double* __restrict__ sx,
                              double* __restrict__ sy, double* __restrict__ sz,
                              int32_t* __restrict__ sid) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const int32_t c = cell[i];
  const int64_t pos = start[c] + atomicAdd(&cursor[c], 1);
  sx[pos] = bx[i];
  sy[pos] = by[i];
  sz[pos] = bz[i];
  if (sid) sid[pos] = (int32_t)i;
}

// Voronoi-approximate volumes: probe p of the lattice (centres lo + (k + 1/2) h) goes to the
// nearest node within radius (squared distances; ties to the lowest node index).
__global__ void k_vor_probe(int64_t np, int64_t px, int64_t py, double lx, double ly, double lz,
                            double hp, IGrid g, const int64_t* __restrict__ start,
                            const double* __restrict__ sx, const double* __restrict__ sy,
                            const double* __restrict__ sz, const int32_t* __restrict__ sid,
                            double r2, unsigned long long* __restrict__ count) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int64_t kx = p % px, ky = (p / px) % py, kz = p / (px * py);
  const double qx = lx + ((double)kx + 0.5) * hp, qy = ly + ((double)ky + 0.5) * hp,
               qz = lz + ((double)kz + 0.5) * hp;
  const double fx = floor((qx - g.x0) / g.h), fy = floor((qy - g.y0) / g.h),
               fz = floor((qz - g.z0) / g.h);
  if (!(fx >= -1.0 && fx <= (double)g.mx && fy >= -1.0 && fy <= (double)g.my && fz >= -1.0 &&
        fz <= (double)g.mz))
    return;
  const int64_t ax = (int64_t)fx, ay = (int64_t)fy, az = (int64_t)fz;
  const int64_t x0 = ax - 1 < 0 ? 0 : ax - 1, x1 = ax + 1 > g.mx - 1 ? g.mx - 1 : ax + 1;
  const int64_t y0 = ay - 1 < 0 ? 0 : ay - 1, y1 = ay + 1 > g.my - 1 ? g.my - 1 : ay + 1;
  const int64_t z0 = az - 1 < 0 ? 0 : az - 1, z1 = az + 1 > g.mz - 1 ? g.mz - 1 : az + 1;
  double best = r2;
  int32_t who = -1;
  for (int64_t cz = z0; cz <= z1; cz++)
    for (int64_t cy = y0; cy <= y1; cy++)
      for (int64_t cx = x0; cx <= x1; cx++) {
        const int64_t c = cx + g.mx * (cy + g.my * cz);
        for (int64_t k = start[c]; k < start[c + 1]; k++) {
          const double dx = sx[k] - qx, dy = sy[k] - qy, dz = sz[k] - qz;
          const double d2 = dx * dx + dy * dy + dz * dz;
          const int32_t j = sid[k];
          if (d2 < best || (d2 == best && who >= 0 && j < who) || (d2 == best && who < 0)) {
            best = d2;
            who = j;
          }
        }
      }
  if (who >= 0) atomicAdd(&count[who], 1ull);
}

__global__ void k_imp_count(int32_t n, const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ z, IGrid g,
                            const int64_t* __restrict__ start, const double* __restrict__ sx,
                            const double* __restrict__ sy, const double* __restrict__ sz,
                            double r, int32_t max_count, int32_t* __restrict__ counts,
                            int* __restrict__ over) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xi = x[i], yi = y[i], zi = z[i];
  // index range of the neighbouring cells, clipped to the grid (empty when the node is far)
  const double fx = floor((xi - g.x0) / g.h), fy = floor((yi - g.y0) / g.h),
               fz = floor((zi - g.z0) / g.h);
  int32_t cnt = 0;
  if (fx >= -1.0 && fx <= (double)g.mx && fy >= -1.0 && fy <= (double)g.my && fz >= -1.0 &&
      fz <= (double)g.mz) {
    const int64_t ax = (int64_t)fx, ay = (int64_t)fy, az = (int64_t)fz;
    const int64_t x0 = ax - 1 < 0 ? 0 : ax - 1, x1 = ax + 1 > g.mx - 1 ? g.mx - 1 : ax + 1;
    const int64_t y0 = ay - 1 < 0 ? 0 : ay - 1, y1 = ay + 1 > g.my - 1 ? g.my - 1 : ay + 1;
    const int64_t z0 = az - 1 < 0 ? 0 : az - 1, z1 = az + 1 > g.mz - 1 ? g.mz - 1 : az + 1;
    for (int64_t cz = z0; cz <= z1; cz++)
      for (int64_t cy = y0; cy <= y1; cy++)
        for (int64_t cx = x0; cx <= x1; cx++) {
          const int64_t c = cx + g.mx * (cy + g.my * cz);
          for (int64_t k = start[c]; k < start[c + 1]; k++) {
            const double dx = sx[k] - xi, dy = sy[k] - yi, dz = sz[k] - zi;
            if (__dsqrt_rn(dx * dx + dy * dy + dz * dz) <= r) cnt++;
          }
        }
  }
  counts[i] = cnt;
  if (cnt > max_count) atomicOr(over, 1);
}

// Candidate scan of one node; FILL == false counts, FILL == true writes the hits.
template <bool FILL>
__global__ __launch_bounds__(NNPP_BLOCK) void k_scan(
    int32_t n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, Grid g, const int* __restrict__ cstart,
    const double* __restrict__ cx, const double* __restrict__ cy, const double* __restrict__ cz,
    const int32_t* __restrict__ cid, int32_t* __restrict__ counts,
    const int64_t* __restrict__ row_ptr, int32_t* __restrict__ col, int32_t max_len) {
  const int64_t i = nnpp_xcd_block(blockIdx.x, gridDim.x) * NNPP_BLOCK + threadIdx.x;
  if (i >= n) return;
  const double xi = x[i], yi = y[i], zi = z[i];
  const double r2 = sqr_(g.r);
  int ixl = (int)floor((xi - g.r - g.x0) / g.cdx) - 1, ixh = (int)floor((xi + g.r - g.x0) / g.cdx) + 1;
  int iyl = (int)floor((yi - g.r - g.y0) / g.cdy) - 1, iyh = (int)floor((yi + g.r - g.y0) / g.cdy) + 1;
  int izl = (int)floor((zi - g.r - g.z0) / g.cdz) - 1, izh = (int)floor((zi + g.r - g.z0) / g.cdz) + 1;
  ixl = max(ixl, 0);
  iyl = max(iyl, 0);
  izl = max(izl, 0);
  ixh = min(ixh, g.mx - 1);
  iyh = min(iyh, g.my - 1);
  izh = min(izh, g.mz_last);
  const int mxy = g.mx * g.my;
  int32_t cnt = 0;
  const int64_t out = FILL ? row_ptr[i] : 0;
  const int64_t row_end = FILL ? row_ptr[i + 1] : 0;
  for (int iz = izl; iz <= izh; iz++)
    for (int iy = iyl; iy <= iyh; iy++)
      for (int ix = ixl; ix <= ixh; ix++) {
        const int m = ix + g.mx * iy + mxy * iz;
        if (m >= g.max_hash) continue;
        const int rr = (m % mxy) % g.mx, cc = (m % mxy) / g.mx, pp = m / mxy;
        const double bx = g.x0 + rr * g.cdx, by = g.y0 + cc * g.cdy, bz = g.z0 + pp * g.cdz;
        if (!sphere_cube(xi, yi, zi, g.r, bx, by, bz, g.cdx, g.cdy, g.cdz)) continue;
        const int k1 = cstart[m + 1];
        for (int k = cstart[m]; k < k1; k++) {
          if (sqr_(xi - cx[k]) + sqr_(yi - cy[k]) + sqr_(zi - cz[k]) <= r2) {
            const int32_t cand = cid[k];
            if (cand != i) {
              if (FILL) {
                if (out + cnt < row_end) col[out + cnt] = cand;
              }
              cnt++;
            }
          }
        }
      }
  if (!FILL) counts[i] = cnt;
}

// One wavefront per row: bitonic sort of the row in LDS (P = power of two >= longest row).
__global__ __launch_bounds__(NNPP_BLOCK) void k_sort_rows(int32_t n, int P,
                                                          const int64_t* __restrict__ row_ptr,
                                                          int32_t* __restrict__ col) {
  extern __shared__ int32_t s_rows[];
  const int wid = threadIdx.x / NNPP_WAVE, lane = threadIdx.x % NNPP_WAVE;
  const int64_t row = (int64_t)blockIdx.x * (NNPP_BLOCK / NNPP_WAVE) + wid;
  int32_t* s = s_rows + wid * P;
  int64_t b0 = 0;
  int len = 0;
  if (row < n) {
    b0 = row_ptr[row];
    len = (int)(row_ptr[row + 1] - b0);
  }
  for (int t = lane; t < P; t += NNPP_WAVE) s[t] = (t < len) ? col[b0 + t] : INT_MAX;
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = lane; t < P; t += NNPP_WAVE) {
        const int u = t ^ j;
        if (u > t) {
          const int32_t a = s[t], b = s[u];
          const bool up = ((t & k) == 0);
          if ((a > b) == up) {
            s[t] = b;
            s[u] = a;
          }
        }
      }
      __syncthreads();
    }
  for (int t = lane; t < len; t += NNPP_WAVE) col[b0 + t] = s[t];
}

__global__ void k_rowptr_from_counts(int32_t n, const int64_t* __restrict__ excl,
                                     int64_t* __restrict__ row_ptr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= n) row_ptr[i] = excl[i];
}

}  // namespace

int nnpp_device_exclusive_scan_i32(nnpp_ctx* c, const int32_t* in, int64_t n, int64_t* out);

template <typename T>
static int dev_alloc(nnpp_ctx* c, T** p, size_t count) {
  NNPP_HIP(c, hipMalloc((void**)p, sizeof(T) * (count ? count : 1)));
  return NNPP_OK;
}

int nnpp_build_neighbors_dev(nnpp_ctx* c, double r, double cdx, double cdy, double cdz,
                             int32_t max_len, int64_t* nnz_out, int32_t* max_len_out) {
  const int32_t n = c->n;
  if (!(cdx > 0) || !(cdy > 0) || !(cdz > 0) || max_len < 1 || !(r >= 0))
    return nnpp_set_error(c, NNPP_ERR_ARG, "build_neighbors: bad radius/cell size/max_len");
  hipStream_t s = c->stream;
  if (!c->row_ptr) {
    int st = dev_alloc(c, &c->row_ptr, (size_t)n + 1);
    if (st) return st;
  }
  if (n == 0) {
    NNPP_HIP(c, hipMemsetAsync(c->row_ptr, 0, sizeof(int64_t), s));
    NNPP_HIP(c, hipStreamSynchronize(s));
    c->nnz = 0;
    c->max_len = 0;
    if (nnz_out) *nnz_out = 0;
    if (max_len_out) *max_len_out = 0;
    c->geom_dirty = c->pot_dirty = true;
    c->pst.state = 0;
    return NNPP_OK;
  }
  // Step 0 (grid.h:59-71): bounding box of the coordinates; min/max are order independent,
  // so the host copies kept at upload time give the same values.
  std::vector<double> hx(n), hy(n), hz(n);
  NNPP_HIP(c, hipMemcpyAsync(hx.data(), c->x, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  NNPP_HIP(c, hipMemcpyAsync(hy.data(), c->y, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  NNPP_HIP(c, hipMemcpyAsync(hz.data(), c->z, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  NNPP_HIP(c, hipStreamSynchronize(s));
  double x0 = hx[0], x1 = hx[0], y0 = hy[0], y1 = hy[0], z0 = hz[0], z1 = hz[0];
  for (int32_t i = 0; i < n; i++) {
    x0 = hx[i] < x0 ? hx[i] : x0;
    x1 = x1 < hx[i] ? hx[i] : x1;
    y0 = hy[i] < y0 ? hy[i] : y0;
    y1 = y1 < hy[i] ? hy[i] : y1;
    z0 = hz[i] < z0 ? hz[i] : z0;
    z1 = z1 < hz[i] ? hz[i] : z1;
  }
  Grid g;
  g.x0 = x0;
  g.y0 = y0;
  g.z0 = z0;
  g.cdx = cdx;
  g.cdy = cdy;
  g.cdz = cdz;
  g.r = r;
  g.mx = (int)((x1 - x0) / cdx) + 1;  // grid.h:78-79
  g.my = (int)((y1 - y0) / cdy) + 1;
  const int64_t nb = nnpp_nblocks(n);

  int *hash = nullptr, *dmax = nullptr, *cell = nullptr, *cursor = nullptr;
  double *cx = nullptr, *cy = nullptr, *cz = nullptr;
  int32_t *cid = nullptr, *counts = nullptr;
  int64_t* excl = nullptr;
  int st = NNPP_OK;
#define NB_TRY(expr)           \
  do {                         \
    st = (expr);               \
    if (st) goto cleanup;      \
  } while (0)
#define NB_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      st = nnpp_set_error(c, NNPP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
      goto cleanup;                                                                    \
    }                                                                                  \
  } while (0)
  {
    NB_TRY(dev_alloc(c, &hash, n));
    NB_TRY(dev_alloc(c, &dmax, 1));
    NB_HIP(hipMemsetAsync(dmax, 0, sizeof(int), s));
    k_hash<<<nb, NNPP_BLOCK, 0, s>>>(n, c->x, c->y, c->z, g, hash, dmax);
    NB_HIP(hipGetLastError());
    int hmax = 0;
    NB_HIP(hipMemcpyAsync(&hmax, dmax, sizeof(int), hipMemcpyDeviceToHost, s));
    NB_HIP(hipStreamSynchronize(s));
    g.max_hash = hmax + 1;  // grid.h:95
    g.mz_last = (g.max_hash - 1) / (g.mx * g.my);
    NB_TRY(dev_alloc(c, &cell, (size_t)g.max_hash + 1));
    NB_TRY(dev_alloc(c, &cursor, (size_t)g.max_hash + 1));
    NB_HIP(hipMemsetAsync(cell, 0, sizeof(int) * ((size_t)g.max_hash + 1), s));
    k_cell_count<<<nb, NNPP_BLOCK, 0, s>>>(n, hash, cell);
    NB_HIP(hipGetLastError());
    {
      std::vector<int> hc((size_t)g.max_hash + 1);
      NB_HIP(hipMemcpyAsync(hc.data(), cell, sizeof(int) * g.max_hash, hipMemcpyDeviceToHost, s));
      NB_HIP(hipStreamSynchronize(s));
      int acc = 0;
      for (int m = 0; m < g.max_hash; m++) {
        const int v = hc[m];
        hc[m] = acc;
        acc += v;
      }
      hc[g.max_hash] = acc;
      NB_HIP(hipMemcpyAsync(cell, hc.data(), sizeof(int) * ((size_t)g.max_hash + 1),
                            hipMemcpyHostToDevice, s));
      NB_HIP(hipMemcpyAsync(cursor, hc.data(), sizeof(int) * ((size_t)g.max_hash + 1),
                            hipMemcpyHostToDevice, s));
      NB_HIP(hipStreamSynchronize(s));  // hc leaves scope
    }
    NB_TRY(dev_alloc(c, &cx, n));
    NB_TRY(dev_alloc(c, &cy, n));
    NB_TRY(dev_alloc(c, &cz, n));
    NB_TRY(dev_alloc(c, &cid, n));
    k_cell_scatter<<<nb, NNPP_BLOCK, 0, s>>>(n, c->x, c->y, c->z, hash, cursor, cx, cy, cz, cid);
    NB_HIP(hipGetLastError());
    NB_TRY(dev_alloc(c, &counts, n));
    k_scan<false><<<nb, NNPP_BLOCK, 0, s>>>(n, c->x, c->y, c->z, g, cell, cx, cy, cz, cid, counts,
                                            nullptr, nullptr, max_len);
    NB_HIP(hipGetLastError());
    NB_TRY(dev_alloc(c, &excl, (size_t)n + 1));
    NB_TRY(nnpp_device_exclusive_scan_i32(c, counts, n, excl));
    int64_t nnz = 0;
    NB_HIP(hipMemcpyAsync(&nnz, excl + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    // longest row (self excluded) for the overflow test and the sort width
    std::vector<int32_t> hcnt(n);
    NB_HIP(hipMemcpyAsync(hcnt.data(), counts, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
    NB_HIP(hipStreamSynchronize(s));
    int32_t longest = 0;
    for (int32_t i = 0; i < n; i++) longest = hcnt[i] > longest ? hcnt[i] : longest;
    if (longest >= max_len) {  // grid.h:187-189
      st = nnpp_set_error(c, NNPP_ERR_OVERFLOW,
                          "Not enough room to store clouds. Increase max_node_no.");
      goto cleanup;
    }
    k_rowptr_from_counts<<<nnpp_nblocks((int64_t)n + 1), NNPP_BLOCK, 0, s>>>(n, excl, c->row_ptr);
    NB_HIP(hipGetLastError());
    if (nnz > c->col_cap) {
      if (c->col) NB_HIP(hipFree(c->col));
      c->col = nullptr;
      NB_TRY(dev_alloc(c, &c->col, (size_t)nnz));
      c->col_cap = nnz;
    }
    k_scan<true><<<nb, NNPP_BLOCK, 0, s>>>(n, c->x, c->y, c->z, g, cell, cx, cy, cz, cid, nullptr,
                                           c->row_ptr, c->col, max_len);
    NB_HIP(hipGetLastError());
    int P = 1;
    while (P < longest) P <<= 1;
    if (P > 4096) {
      st = nnpp_set_error(c, NNPP_ERR_UNSUPPORTED, "build_neighbors: rows longer than 4096");
      goto cleanup;
    }
    const int rows_per_block = NNPP_BLOCK / NNPP_WAVE;
    k_sort_rows<<<(n + rows_per_block - 1) / rows_per_block, NNPP_BLOCK,
                  sizeof(int32_t) * P * rows_per_block, s>>>(n, P, c->row_ptr, c->col);
    NB_HIP(hipGetLastError());
    NB_HIP(hipStreamSynchronize(s));
    c->nnz = nnz;
    c->max_len = longest + 1;
    if (nnz_out) *nnz_out = nnz;
    if (max_len_out) *max_len_out = longest + 1;
    c->geom_dirty = c->pot_dirty = true;
    c->pst.state = 0;
  }
cleanup:
  hipStreamSynchronize(s);
  if (hash) hipFree(hash);
  if (dmax) hipFree(dmax);
  if (cell) hipFree(cell);
  if (cursor) hipFree(cursor);
  if (cx) hipFree(cx);
  if (cy) hipFree(cy);
  if (cz) hipFree(cz);
  if (cid) hipFree(cid);
  if (counts) hipFree(counts);
  if (excl) hipFree(excl);
#undef NB_TRY
#undef NB_HIP
  return st;
}

// A uniform grid (cells at least cell_min wide, at most 2^25 cells) over np points with the
// given bounding box: the points counting-sorted by cell (sx, sy, sz, source index sid).
struct PointGrid {
  IGrid g;
  int64_t* start = nullptr;
  double *sx = nullptr, *sy = nullptr, *sz = nullptr;
  int32_t* sid = nullptr;
  void release() {
    void* ps[] = {start, sx, sy, sz, sid};
    for (void* q : ps)
      if (q) hipFree(q);
    start = nullptr;
    sx = sy = sz = nullptr;
    sid = nullptr;
  }
};

static int build_point_grid(nnpp_ctx* c, int64_t np, const double* px, const double* py,
                            const double* pz, const double bbox[6], double cell_min,
                            bool want_id, PointGrid& G) {
  hipStream_t s = c->stream;
  IGrid& g = G.g;
  g.x0 = bbox[0];
  g.y0 = bbox[2];
  g.z0 = bbox[4];
  g.h = cell_min > 0 ? cell_min : 1.0;
  if (g.h < 1e-300) g.h = 1e-300;
  const double ext[3] = {bbox[1] - bbox[0], bbox[3] - bbox[2], bbox[5] - bbox[4]};
  for (;;) {  // widen the cells (never below cell_min) until the grid has at most 2^25 cells
    g.mx = (int64_t)std::floor(ext[0] / g.h) + 1;
    g.my = (int64_t)std::floor(ext[1] / g.h) + 1;
    g.mz = (int64_t)std::floor(ext[2] / g.h) + 1;
    if ((double)g.mx * (double)g.my * (double)g.mz <= (double)(1 << 25)) break;
    g.h *= 2.0;
  }
  const int64_t ncell = g.mx * g.my * g.mz;
  int32_t *cell = nullptr, *cnt = nullptr, *cursor = nullptr;
  int st = NNPP_OK;
  auto fail = [&](int e) {
    if (cell) hipFree(cell);
    if (cnt) hipFree(cnt);
    if (cursor) hipFree(cursor);
    G.release();
    return e;
  };
#define PG_TRY(expr)          \
  do {                        \
    st = (expr);              \
    if (st) return fail(st);  \
  } while (0)
#define PG_HIP(call)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(nnpp_set_error(c, NNPP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_))); \
  } while (0)
  PG_TRY(dev_alloc(c, &cell, np));
  PG_TRY(dev_alloc(c, &cnt, ncell));
  PG_TRY(dev_alloc(c, &cursor, ncell));
  PG_TRY(dev_alloc(c, &G.start, ncell + 1));
  PG_TRY(dev_alloc(c, &G.sx, np));
  PG_TRY(dev_alloc(c, &G.sy, np));
  PG_TRY(dev_alloc(c, &G.sz, np));
  if (want_id) PG_TRY(dev_alloc(c, &G.sid, np));
  PG_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t) * ncell, s));
  PG_HIP(hipMemsetAsync(cursor, 0, sizeof(int32_t) * ncell, s));
  const int64_t gb = (np + NNPP_BLOCK - 1) / NNPP_BLOCK;
  k_imp_cell<<<gb, NNPP_BLOCK, 0, s>>>(np, px, py, pz, g, cell, cnt);
  PG_HIP(hipGetLastError());
  PG_TRY(nnpp_device_exclusive_scan_i32(c, cnt, ncell, G.start));
  k_imp_scatter<<<gb, NNPP_BLOCK, 0, s>>>(np, px, py, pz, cell, G.start, cursor, G.sx, G.sy,
                                          G.sz, G.sid);
  PG_HIP(hipGetLastError());
  PG_HIP(hipStreamSynchronize(s));
  hipFree(cell);
  hipFree(cnt);
  hipFree(cursor);
  return NNPP_OK;
#undef PG_TRY
#undef PG_HIP
}

// Implant points within r of every node (device arrays bx/by/bz of nb points, bounding box
// given); counts on the device. NNPP_ERR_OVERFLOW when a count exceeds max_count.
int nnpp_implant_count_dev(nnpp_ctx* c, int64_t nb, const double* bx, const double* by,
                           const double* bz, const double bbox[6], double r, int32_t max_count,
                           int32_t* counts) {
  const int32_t n = c->n;
  hipStream_t s = c->stream;
  if (n == 0) return NNPP_OK;
  if (nb == 0) {
    NNPP_HIP(c, hipMemsetAsync(counts, 0, sizeof(int32_t) * n, s));
    NNPP_HIP(c, hipStreamSynchronize(s));
    return NNPP_OK;
  }
  PointGrid G;
  int st = build_point_grid(c, nb, bx, by, bz, bbox, r * 1.001, false, G);
  if (st) return st;
  int* over = nullptr;
  if (hipMalloc((void**)&over, sizeof(int)) != hipSuccess ||
      hipMemsetAsync(over, 0, sizeof(int), s) != hipSuccess) {
    if (over) hipFree(over);
    G.release();
    return nnpp_set_error(c, NNPP_ERR_HIP, "implant flag allocation");
  }
  k_imp_count<<<nnpp_nblocks(n), NNPP_BLOCK, 0, s>>>(n, c->x, c->y, c->z, G.g, G.start, G.sx,
                                                     G.sy, G.sz, r, max_count, counts, over);
  hipError_t e = hipGetLastError();
  int h_over = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&h_over, over, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  hipFree(over);
  G.release();
  if (e != hipSuccess) return nnpp_set_error(c, NNPP_ERR_HIP, hipGetErrorString(e));
  if (h_over)
    return nnpp_set_error(c, NNPP_ERR_OVERFLOW,
                          "an implant count exceeds max_count (kd_tree_interface.h:214-217)");
  return NNPP_OK;
}

// Voronoi-approximate node volumes (probe lattice of spacing hp over the nodes' bounding box
// widened by margin, nearest node within radius); counts on the device (n values).
int nnpp_voronoi_dev(nnpp_ctx* c, double hp, double margin, double radius,
                     unsigned long long* count, int64_t dims[3]) {
  const int32_t n = c->n;
  hipStream_t s = c->stream;
  dims[0] = dims[1] = dims[2] = 0;
  if (n == 0) return NNPP_OK;
  std::vector<double> hx(n), hy(n), hz(n);
  NNPP_HIP(c, hipMemcpyAsync(hx.data(), c->x, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  NNPP_HIP(c, hipMemcpyAsync(hy.data(), c->y, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  NNPP_HIP(c, hipMemcpyAsync(hz.data(), c->z, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  NNPP_HIP(c, hipStreamSynchronize(s));
  double bbox[6] = {hx[0], hx[0], hy[0], hy[0], hz[0], hz[0]};
  for (int32_t i = 0; i < n; i++) {
    bbox[0] = hx[i] < bbox[0] ? hx[i] : bbox[0];
    bbox[1] = hx[i] > bbox[1] ? hx[i] : bbox[1];
    bbox[2] = hy[i] < bbox[2] ? hy[i] : bbox[2];
    bbox[3] = hy[i] > bbox[3] ? hy[i] : bbox[3];
    bbox[4] = hz[i] < bbox[4] ? hz[i] : bbox[4];
    bbox[5] = hz[i] > bbox[5] ? hz[i] : bbox[5];
  }
  const double lo[3] = {bbox[0] - margin, bbox[2] - margin, bbox[4] - margin};
  for (int a = 0; a < 3; a++) {
    const double e = (bbox[2 * a + 1] + margin) - lo[a];
    dims[a] = (int64_t)std::llround(e / hp);
    if (dims[a] < 1) dims[a] = 1;
  }
  const double np_d = (double)dims[0] * (double)dims[1] * (double)dims[2];
  if (np_d > 4e9)
    return nnpp_set_error(c, NNPP_ERR_ARG, "voronoi volumes: more than 4e9 probes (raise h)");
  const int64_t np = dims[0] * dims[1] * dims[2];
  PointGrid G;
  int st = build_point_grid(c, n, c->x, c->y, c->z, bbox, radius * 1.001, true, G);
  if (st) return st;
  NNPP_HIP(c, hipMemsetAsync(count, 0, sizeof(unsigned long long) * n, s));
  k_vor_probe<<<(np + NNPP_BLOCK - 1) / NNPP_BLOCK, NNPP_BLOCK, 0, s>>>(
      np, dims[0], dims[1], lo[0], lo[1], lo[2], hp, G.g, G.start, G.sx, G.sy, G.sz, G.sid,
      radius * radius, count);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  G.release();
  if (e != hipSuccess) return nnpp_set_error(c, NNPP_ERR_HIP, hipGetErrorString(e));
  return NNPP_OK;
}

// ---- implant classification without device allocation ----
// Scratch of nnpp_implant_classify_noalloc: a file-scope device array, one copy per GPU. Host
// code reaches it only through the address hipGetSymbolAddress returns on each device (kept
// in the context); the array's name in host code is not a device address.
namespace {
constexpr int64_t NA_P = NNPP_NOALLOC_MAX_POINTS, NA_C = NNPP_NOALLOC_MAX_CELLS,
                  NA_N = NNPP_NOALLOC_MAX_NODES;
constexpr size_t na_round(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t NA_OFF_PTS = 0;                                        // 3 NA_P doubles
constexpr size_t NA_OFF_SORT = NA_OFF_PTS + na_round(3 * NA_P * 8);     // 3 NA_P doubles
constexpr size_t NA_OFF_CELL = NA_OFF_SORT + na_round(3 * NA_P * 8);    // NA_P int32
constexpr size_t NA_OFF_CNT = NA_OFF_CELL + na_round(NA_P * 4);         // NA_C int32
constexpr size_t NA_OFF_CUR = NA_OFF_CNT + na_round(NA_C * 4);          // NA_C int32
constexpr size_t NA_OFF_START = NA_OFF_CUR + na_round(NA_C * 4);        // NA_C + 1 int64
constexpr size_t NA_OFF_COUNTS = NA_OFF_START + na_round((NA_C + 1) * 8);  // NA_N int32
constexpr size_t NA_OFF_FLAG = NA_OFF_COUNTS + na_round(NA_N * 4);      // 1 int32
constexpr size_t NA_BYTES = NA_OFF_FLAG + 256;
}  // namespace

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {

__global__ void k_na_fill(int32_t* __restrict__ p, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Exclusive scan of n <= NA_C counts into start[0..n] by one block of 1024 threads: each
// thread sums a contiguous chunk, the chunk sums are scanned in LDS, then each thread writes
// its chunk's prefixes.
__global__ __launch_bounds__(1024) void k_na_scan(const int32_t* __restrict__ in, int64_t n,
                                                  int64_t* __restrict__ start) {
  __shared__ int64_t part[1024];
  const int t = threadIdx.x;
  const int64_t per = (n + 1023) / 1024;
  const int64_t a = t * per, b = a + per < n ? a + per : n;
  int64_t sum = 0;
  for (int64_t k = a; k < b; k++) sum += in[k];
  part[t] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {  // inclusive Hillis-Steele scan of the chunk sums
    const int64_t v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int64_t run = t ? part[t - 1] : 0;
  for (int64_t k = a; k < b; k++) {
    start[k] = run;
    run += in[k];
  }
  if (t == 1023) start[n] = part[1023];
}

}  // namespace

void nnpp_implant_icvalues_host(const nnpp_ctx* c, const int32_t* counts, double* ic_out) {
  const int32_t n = c->n;
  const double CS = c->p.CSolid;
  for (int32_t i = 0; i < n; i++) {  // main.cpp:422-457
    double v[11] = {0.0};
    if (counts[i] == 0) {
      v[1] = 1e-4 / CS;
      v[2] = 1e-4 / CS;
      v[3] = c->p.C0HCO3 / CS;
      v[5] = c->p.C0HPO4 / CS;
      v[7] = c->p.C0Ca / CS;
      v[8] = c->p.C0Na / CS;
      v[9] = c->p.C0Cl / CS;
    } else {
      v[0] = CS / CS;
    }
    for (int f = 0; f < 11; f++) ic_out[(int64_t)f * n + i] = v[f];
  }
}

int nnpp_noalloc_setup(nnpp_ctx* c) {
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
    return nnpp_set_error(c, NNPP_ERR_HIP, "scratch symbol address");
  c->noalloc_scratch = p;
  return NNPP_OK;
}

int nnpp_implant_classify_noalloc(nnpp_ctx* c, int64_t nb, const double* bx, const double* by,
                                  const double* bz, double r, int32_t max_count,
                                  int32_t* counts, double* ic_out) {
  if (!c || nb < 0 || (nb && (!bx || !by || !bz)) || (!counts && c->n) || !(r >= 0) ||
      max_count < 0)
    return NNPP_ERR_ARG;
  if (!c->have_params) return nnpp_set_error(c, NNPP_ERR_STATE, "nnpp_set_params not called");
  if (!c->x) return nnpp_set_error(c, NNPP_ERR_STATE, "nnpp_set_nodes not called");
  if (ic_out && c->p.n_species != 10)
    return nnpp_set_error(c, NNPP_ERR_UNSUPPORTED,
                          "implant initial values need the 10 reference species");
  if (nb > NA_P)
    return nnpp_set_error(c, NNPP_ERR_ARG,
                          "nnpp_implant_classify_noalloc: more implant points than "
                          "NNPP_NOALLOC_MAX_POINTS");
  if ((int64_t)c->n > NA_N)
    return nnpp_set_error(c, NNPP_ERR_ARG,
                          "nnpp_implant_classify_noalloc: more nodes than NNPP_NOALLOC_MAX_NODES");
  if (!c->noalloc_scratch) return nnpp_set_error(c, NNPP_ERR_STATE, "no scratch address");
  NNPP_HIP(c, hipSetDevice(c->device));
  const int32_t n = c->n;
  if (n == 0) return NNPP_OK;
  hipStream_t s = c->stream;
  int over_h = 0;
  if (nb == 0) {
    for (int32_t i = 0; i < n; i++) counts[i] = 0;
  } else {
    double bbox[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t k = 0; k < nb; k++) {
      const double v[3] = {bx[k], by[k], bz[k]};
      for (int a = 0; a < 3; a++) {
        if (k == 0 || v[a] < bbox[2 * a]) bbox[2 * a] = v[a];
        if (k == 0 || v[a] > bbox[2 * a + 1]) bbox[2 * a + 1] = v[a];
      }
    }
    for (int a = 0; a < 6; a++)
      if (!std::isfinite(bbox[a]))
        return nnpp_set_error(c, NNPP_ERR_ARG, "non-finite implant coordinates");
    // cells at least 1.001 r wide (as nnpp_implant_classify), widened to fit the scratch grid
    IGrid g;
    g.x0 = bbox[0];
    g.y0 = bbox[2];
    g.z0 = bbox[4];
    g.h = r * 1.001 > 0 ? r * 1.001 : 1.0;
    if (g.h < 1e-300) g.h = 1e-300;
    for (;;) {
      g.mx = (int64_t)std::floor((bbox[1] - bbox[0]) / g.h) + 1;
      g.my = (int64_t)std::floor((bbox[3] - bbox[2]) / g.h) + 1;
      g.mz = (int64_t)std::floor((bbox[5] - bbox[4]) / g.h) + 1;
      if ((double)g.mx * (double)g.my * (double)g.mz <= (double)NA_C) break;
      g.h *= 2.0;
    }
    const int64_t ncell = g.mx * g.my * g.mz;
    unsigned char* base = static_cast<unsigned char*>(c->noalloc_scratch);
    double* pts = reinterpret_cast<double*>(base + NA_OFF_PTS);
    double* srt = reinterpret_cast<double*>(base + NA_OFF_SORT);
    int32_t* cell = reinterpret_cast<int32_t*>(base + NA_OFF_CELL);
    int32_t* cnt = reinterpret_cast<int32_t*>(base + NA_OFF_CNT);
    int32_t* cur = reinterpret_cast<int32_t*>(base + NA_OFF_CUR);
    int64_t* start = reinterpret_cast<int64_t*>(base + NA_OFF_START);
    int32_t* dcounts = reinterpret_cast<int32_t*>(base + NA_OFF_COUNTS);
    int32_t* over = reinterpret_cast<int32_t*>(base + NA_OFF_FLAG);
    // the array keeps the last call's contents: clear what must start at zero (fill kernels,
    // which graph capture records as kernel nodes)
    k_na_fill<<<(ncell + NNPP_BLOCK - 1) / NNPP_BLOCK, NNPP_BLOCK, 0, s>>>(cnt, ncell, 0);
    k_na_fill<<<(ncell + NNPP_BLOCK - 1) / NNPP_BLOCK, NNPP_BLOCK, 0, s>>>(cur, ncell, 0);
    k_na_fill<<<1, 64, 0, s>>>(over, 1, 0);
    NNPP_HIP(c, hipGetLastError());
    NNPP_HIP(c, hipMemcpyAsync(pts, bx, sizeof(double) * nb, hipMemcpyHostToDevice, s));
    NNPP_HIP(c, hipMemcpyAsync(pts + NA_P, by, sizeof(double) * nb, hipMemcpyHostToDevice, s));
    NNPP_HIP(c, hipMemcpyAsync(pts + 2 * NA_P, bz, sizeof(double) * nb, hipMemcpyHostToDevice, s));
    const int64_t gb = (nb + NNPP_BLOCK - 1) / NNPP_BLOCK;
    k_imp_cell<<<gb, NNPP_BLOCK, 0, s>>>(nb, pts, pts + NA_P, pts + 2 * NA_P, g, cell, cnt);
    k_na_scan<<<1, 1024, 0, s>>>(cnt, ncell, start);
    k_imp_scatter<<<gb, NNPP_BLOCK, 0, s>>>(nb, pts, pts + NA_P, pts + 2 * NA_P, cell, start, cur,
                                            srt, srt + NA_P, srt + 2 * NA_P, nullptr);
    k_imp_count<<<nnpp_nblocks(n), NNPP_BLOCK, 0, s>>>(n, c->x, c->y, c->z, g, start, srt,
                                                       srt + NA_P, srt + 2 * NA_P, r, max_count,
                                                       dcounts, over);
    NNPP_HIP(c, hipGetLastError());
    NNPP_HIP(c, hipMemcpyAsync(counts, dcounts, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
    NNPP_HIP(c, hipMemcpyAsync(&over_h, over, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    NNPP_HIP(c, hipStreamSynchronize(s));
  }
  if (ic_out) nnpp_implant_icvalues_host(c, counts, ic_out);
  if (over_h)
    return nnpp_set_error(c, NNPP_ERR_OVERFLOW,
                          "an implant count exceeds max_count (kd_tree_interface.h:214-217)");
  return NNPP_OK;
}

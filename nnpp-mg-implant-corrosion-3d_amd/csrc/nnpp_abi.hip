// nnpp_abi.hip — extern "C" entry points of libnnpp (see include/nnpp.h for the reference
// interface each one replaces). Host code only; kernels live in k_*.hip.
#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

#include "nnpp_internal.hpp"

static_assert(offsetof(nnpp_params, vol_red) - offsetof(nnpp_params, t0) ==
                  (NNPP_DATA2_LEN - 1) * sizeof(double),
              ".data2 block must be contiguous");

int nnpp_set_error(nnpp_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

namespace {

template <typename T>
void dfree(T*& p) {
  if (p) hipFree((void*)p);
  p = nullptr;
}

void free_slices(nnpp_slices& s) {
  dfree(s.off);
  dfree(s.len);
  dfree(s.col);
  dfree(s.w64);
  dfree(s.w32);
  dfree(s.first);
  dfree(s.d16);
  dfree(s.cls);
  dfree(s.tab64);
  dfree(s.tab32);
  s = nnpp_slices();
}

void free_nodes(nnpp_ctx* c) {
  dfree(c->x);
  dfree(c->y);
  dfree(c->z);
  dfree(c->vol);
  dfree(c->row_ptr);
  dfree(c->col);
  c->col_cap = 0;
  c->nnz = -1;
  free_slices(c->bonds);
  free_slices(c->pot);
  nnpp_free_stencil(c->pst);
  dfree(c->phase);
  dfree(c->bcmask);
  dfree(c->bcval);
  dfree(c->U[0]);
  dfree(c->U[1]);
  dfree(c->Unew);
  dfree(c->phi_prev);
  c->phi_hist = 0;
  nnpp_implicit_free(c);
  dfree(c->Fres);
  dfree(c->aos);
  dfree(c->pfree);
  dfree(c->pdd);
  dfree(c->pdd32);
  dfree(c->pdd32nf);
  dfree(c->pinv);
  dfree(c->pinv32);
  dfree(c->pbdir);
  dfree(c->psc);
  dfree(c->cg_b);
  dfree(c->cg_x);
  dfree(c->cg_r);
  dfree(c->cg_z);
  dfree(c->cg_p);
  dfree(c->cg_q);
  dfree(c->f_e);
  dfree(c->f_r);
  dfree(c->f_z);
  dfree(c->f_p);
  dfree(c->f_q);
  dfree(c->f_w1);
  dfree(c->f_s0);
  dfree(c->f_s1);
  dfree(c->red_part);
}

int need_params(nnpp_ctx* c) {
  if (!c->have_params) return nnpp_set_error(c, NNPP_ERR_STATE, "nnpp_set_params not called");
  return NNPP_OK;
}

int need_graph(nnpp_ctx* c) {
  int st = need_params(c);
  if (st) return st;
  if (c->nnz < 0) return nnpp_set_error(c, NNPP_ERR_STATE, "no neighbour list (build or set it)");
  return NNPP_OK;
}

int ensure_bonds(nnpp_ctx* c) {
  int st = need_graph(c);
  if (st) return st;
  if (c->geom_dirty) return nnpp_build_bond_slices(c);
  return NNPP_OK;
}

int ensure_potential(nnpp_ctx* c) {
  int st = need_graph(c);
  if (st) return st;
  if (c->pot_dirty) return nnpp_build_potential(c);
  return NNPP_OK;
}

int64_t fields(const nnpp_ctx* c) { return (int64_t)c->p.n_species + 1; }

// The residual kernel evaluates the phi row (utils.h:798-841) with the species bond weights,
// so it needs the potential exponent s111 to equal the species exponent (every shipped
// parameter set uses 0 for both).
int phi_row_guard(nnpp_ctx* c) {
  if (c->p.s_phi != c->p.s1[0])
    return nnpp_set_error(c, NNPP_ERR_UNSUPPORTED,
                          "residuals need the potential kernel exponent equal to the species'");
  return NNPP_OK;
}

// Calls that would change what a recorded step depends on are refused while a tape exists.
int tape_guard(nnpp_ctx* c) {
  if (c->tape.active)
    return nnpp_set_error(c, NNPP_ERR_STATE, "a differentiation tape is recording: nnpp_tape_end first");
  return NNPP_OK;
}

// Device copy of the state before a recorded step.
int tape_checkpoint(nnpp_ctx* c) {
  nnpp_tape_ckpt k;
  const size_t m = c->n ? c->n : 1;
  const size_t bytes = sizeof(double) * m * fields(c);
  NNPP_HIP(c, hipMalloc((void**)&k.U, bytes));
  if (hipMalloc((void**)&k.phase, m) != hipSuccess) {
    hipFree(k.U);
    return nnpp_set_error(c, NNPP_ERR_HIP, "tape checkpoint allocation");
  }
  NNPP_HIP(c, hipMemcpyAsync(k.U, c->U[NNPP_SLOT_N], bytes, hipMemcpyDeviceToDevice, c->stream));
  NNPP_HIP(c, hipMemcpyAsync(k.phase, c->phase, m, hipMemcpyDeviceToDevice, c->stream));
  k.phi_hist = c->phi_hist;
  if (k.phi_hist > 0) {
    if (hipMalloc((void**)&k.phi_prev, sizeof(double) * 2 * m) != hipSuccess) {
      hipFree(k.U);
      hipFree(k.phase);
      return nnpp_set_error(c, NNPP_ERR_HIP, "tape checkpoint allocation");
    }
    NNPP_HIP(c, hipMemcpyAsync(k.phi_prev, c->phi_prev, sizeof(double) * 2 * m,
                               hipMemcpyDeviceToDevice, c->stream));
  }
  k.step = c->step;
  k.solid_count = c->solid_count;
  k.voln1rel = c->voln1rel;
  c->tape.ckpts.push_back(k);
  c->tape.bytes += (int64_t)(bytes + m + (k.phi_prev ? sizeof(double) * 2 * m : 0));
  return NNPP_OK;
}

}  // namespace

void nnpp_tape_free(nnpp_ctx* c) {
  for (nnpp_tape_ckpt& k : c->tape.ckpts) {
    dfree(k.U);
    dfree(k.phase);
    dfree(k.phi_prev);
  }
  c->tape = nnpp_tape();
}

extern "C" {

int nnpp_create(nnpp_ctx** out, int32_t device) {
  if (!out) return NNPP_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return NNPP_ERR_HIP;
  if (device < 0 || device >= ndev) return NNPP_ERR_ARG;
  nnpp_ctx* c = new nnpp_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc((void**)&c->scal, sizeof(double) * (SC_COUNT + NNPP_MAX_SPECIES)) != hipSuccess ||
      hipHostMalloc((void**)&c->h_scal, sizeof(double) * (SC_COUNT + NNPP_MAX_SPECIES)) !=
          hipSuccess ||
      hipMemset(c->scal, 0, sizeof(double) * (SC_COUNT + NNPP_MAX_SPECIES)) != hipSuccess ||
      nnpp_noalloc_setup(c) != NNPP_OK) {
    nnpp_destroy(c);
    return NNPP_ERR_HIP;
  }
  *out = c;
  return NNPP_OK;
}

int nnpp_destroy(nnpp_ctx* c) {
  if (!c) return NNPP_OK;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  nnpp_tape_free(c);
  free_nodes(c);
  dfree(c->scal);
  if (c->h_scal) hipHostFree(c->h_scal);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return NNPP_OK;
}

const char* nnpp_last_error(const nnpp_ctx* c) { return c ? c->err.c_str() : "null context"; }

int nnpp_set_params(nnpp_ctx* c, const nnpp_params* p) {
  if (!c || !p) return NNPP_ERR_ARG;
  if (int g_ = tape_guard(c)) return g_;
  const int S = p->n_species;
  if (S < 1 || S > NNPP_MAX_SPECIES)
    return nnpp_set_error(c, NNPP_ERR_ARG, "n_species must be 1..16");
  if (p->chemistry == NNPP_CHEM_REFERENCE && S != 10)
    return nnpp_set_error(c, NNPP_ERR_UNSUPPORTED, "reference chemistry needs 10 species");
  if (p->chemistry != NNPP_CHEM_REFERENCE && p->chemistry != NNPP_CHEM_GENERIC)
    return nnpp_set_error(c, NNPP_ERR_ARG, "unknown chemistry");
  if (!(p->dt > 0) || !(p->dx > 0) || !(p->dy > 0) || !(p->dz > 0) || !(p->delta > 0) ||
      !(p->CSolid != 0) || !(p->R * p->T != 0) || !(p->K_phi * p->sigma1 != 0))
    return nnpp_set_error(c, NNPP_ERR_ARG, "dt, dx, dy, dz, delta, CSolid, R*T, K11*sigma1");
  // one bond kernel exponent for every species term (all shipped parameter sets use s = 0)
  const double s0 = p->s1[0];
  for (int k = 0; k < S; k++)
    if (p->s1[k] != s0 || p->s2[k] != s0)
      return nnpp_set_error(c, NNPP_ERR_UNSUPPORTED,
                            "species kernel exponents s1/s2 must all be equal");
  const bool changed_geom = !c->have_params || c->p.delta != p->delta || c->p.dx != p->dx ||
                            c->p.dy != p->dy || c->p.dz != p->dz || c->p.s1[0] != s0 ||
                            c->p.s_phi != p->s_phi || c->p.n_species != S;
  const bool changed_bc = c->have_params && (c->p.CSolid != p->CSolid || c->p.n_species != S);
  if (c->have_params && c->p.n_species != S && c->n > 0)
    return nnpp_set_error(c, NNPP_ERR_STATE, "n_species changed: call nnpp_set_nodes again");
  c->p = *p;
  c->have_params = true;
  c->w_exponent = 3.0 + 2.0 * s0;
  if (changed_geom) {
    c->geom_dirty = c->pot_dirty = true;
    c->pst.state = 0;
  }
  if (changed_bc) c->have_bc = false;
  return NNPP_OK;
}

int nnpp_set_nodes(nnpp_ctx* c, int32_t n, const double* x, const double* y, const double* z,
                   const double* volumes) {
  if (!c || n < 0 || (n && (!x || !y || !z))) return NNPP_ERR_ARG;
  if (int g_ = tape_guard(c)) return g_;
  int st = need_params(c);
  if (st) return st;
  NNPP_HIP(c, hipSetDevice(c->device));
  NNPP_HIP(c, hipStreamSynchronize(c->stream));
  free_nodes(c);
  c->n = n;
  const size_t m = n ? n : 1;
  const size_t F = fields(c);
  NNPP_HIP(c, hipMalloc((void**)&c->x, sizeof(double) * m));
  NNPP_HIP(c, hipMalloc((void**)&c->y, sizeof(double) * m));
  NNPP_HIP(c, hipMalloc((void**)&c->z, sizeof(double) * m));
  if (volumes) NNPP_HIP(c, hipMalloc((void**)&c->vol, sizeof(double) * m));
  NNPP_HIP(c, hipMalloc((void**)&c->phase, m));
  NNPP_HIP(c, hipMalloc((void**)&c->bcmask, sizeof(uint16_t) * m));
  NNPP_HIP(c, hipMalloc((void**)&c->bcval, sizeof(double) * m * F));
  NNPP_HIP(c, hipMalloc((void**)&c->U[0], sizeof(double) * m * F));
  NNPP_HIP(c, hipMalloc((void**)&c->U[1], sizeof(double) * m * F));
  NNPP_HIP(c, hipMalloc((void**)&c->Unew, sizeof(double) * m * F));
  NNPP_HIP(c, hipMalloc((void**)&c->Fres, sizeof(double) * m * F));
  NNPP_HIP(c, hipMalloc((void**)&c->aos, sizeof(double) * m * ((F + 2) & ~(size_t)1)));
  NNPP_HIP(c, hipMalloc((void**)&c->pfree, m));
  NNPP_HIP(c, hipMalloc((void**)&c->pdd, sizeof(double) * m));
  NNPP_HIP(c, hipMalloc((void**)&c->pdd32, sizeof(float) * m));
  NNPP_HIP(c, hipMalloc((void**)&c->pdd32nf, sizeof(float) * m));
  NNPP_HIP(c, hipMalloc((void**)&c->pinv, sizeof(double) * m));
  NNPP_HIP(c, hipMalloc((void**)&c->pinv32, sizeof(float) * m));
  NNPP_HIP(c, hipMalloc((void**)&c->pbdir, sizeof(double) * m));
  NNPP_HIP(c, hipMalloc((void**)&c->psc, sizeof(double) * m));
  double** dv[] = {&c->cg_b, &c->cg_x, &c->cg_r, &c->cg_z, &c->cg_p, &c->cg_q};
  for (double** d : dv) NNPP_HIP(c, hipMalloc((void**)d, sizeof(double) * m));
  float** fv[] = {&c->f_e, &c->f_r, &c->f_z, &c->f_p, &c->f_q, &c->f_w1, &c->f_s0, &c->f_s1};
  for (float** f : fv) NNPP_HIP(c, hipMalloc((void**)f, sizeof(float) * m));
  const int64_t nblk = std::max<int64_t>(nnpp_nblocks(n), 4096);
  NNPP_HIP(c, hipMalloc((void**)&c->red_part, sizeof(double) * 4 * nblk));
  hipStream_t s = c->stream;
  if (n) {
    NNPP_HIP(c, hipMemcpyAsync(c->x, x, sizeof(double) * n, hipMemcpyHostToDevice, s));
    NNPP_HIP(c, hipMemcpyAsync(c->y, y, sizeof(double) * n, hipMemcpyHostToDevice, s));
    NNPP_HIP(c, hipMemcpyAsync(c->z, z, sizeof(double) * n, hipMemcpyHostToDevice, s));
    if (volumes)
      NNPP_HIP(c, hipMemcpyAsync(c->vol, volumes, sizeof(double) * n, hipMemcpyHostToDevice, s));
  }
  NNPP_HIP(c, hipMemsetAsync(c->phase, 0, m, s));
  NNPP_HIP(c, hipMemsetAsync(c->bcmask, 0, sizeof(uint16_t) * m, s));
  NNPP_HIP(c, hipMemsetAsync(c->bcval, 0, sizeof(double) * m * F, s));
  NNPP_HIP(c, hipMemsetAsync(c->U[0], 0, sizeof(double) * m * F, s));
  NNPP_HIP(c, hipMemsetAsync(c->U[1], 0, sizeof(double) * m * F, s));
  NNPP_HIP(c, hipMemsetAsync(c->Unew, 0, sizeof(double) * m * F, s));
  NNPP_HIP(c, hipStreamSynchronize(s));
  c->have_bc = false;
  c->geom_dirty = c->pot_dirty = true;
  c->pst.state = 0;
  c->step = 1;
  c->voln1rel = 0.0;
  c->vol0rel = 0.0;
  c->solid_count = 0;
  return NNPP_OK;
}

int nnpp_build_neighbors(nnpp_ctx* c, double r, double cdx, double cdy, double cdz,
                         int32_t max_len, int64_t* nnz_out, int32_t* max_len_out) {
  if (!c) return NNPP_ERR_ARG;
  if (int g_ = tape_guard(c)) return g_;
  int st = need_params(c);
  if (st) return st;
  if (!c->x) return nnpp_set_error(c, NNPP_ERR_STATE, "nnpp_set_nodes not called");
  NNPP_HIP(c, hipSetDevice(c->device));
  return nnpp_build_neighbors_dev(c, r, cdx, cdy, cdz, max_len, nnz_out, max_len_out);
}

int nnpp_get_neighbors(nnpp_ctx* c, int64_t* row_ptr, int32_t* col) {
  if (!c || !row_ptr) return NNPP_ERR_ARG;
  if (c->nnz < 0) return nnpp_set_error(c, NNPP_ERR_STATE, "no neighbour list");
  NNPP_HIP(c, hipSetDevice(c->device));
  NNPP_HIP(c, hipMemcpy(row_ptr, c->row_ptr, sizeof(int64_t) * ((size_t)c->n + 1),
                        hipMemcpyDeviceToHost));
  if (c->nnz && col)
    NNPP_HIP(c, hipMemcpy(col, c->col, sizeof(int32_t) * c->nnz, hipMemcpyDeviceToHost));
  return NNPP_OK;
}

int nnpp_set_neighbors(nnpp_ctx* c, const int64_t* row_ptr, const int32_t* col) {
  if (!c || !row_ptr) return NNPP_ERR_ARG;
  if (int g_ = tape_guard(c)) return g_;
  if (!c->x) return nnpp_set_error(c, NNPP_ERR_STATE, "nnpp_set_nodes not called");
  const int32_t n = c->n;
  if (row_ptr[0] != 0) return nnpp_set_error(c, NNPP_ERR_ARG, "row_ptr[0] != 0");
  int32_t longest = 0;
  for (int32_t i = 0; i < n; i++) {
    const int64_t l = row_ptr[i + 1] - row_ptr[i];
    if (l < 0) return nnpp_set_error(c, NNPP_ERR_ARG, "row_ptr not monotone");
    if (l > longest) longest = (int32_t)l;
  }
  const int64_t nnz = row_ptr[n];
  if (nnz && !col) return NNPP_ERR_ARG;
  for (int64_t k = 0; k < nnz; k++)
    if (col[k] < 0 || col[k] >= n) return nnpp_set_error(c, NNPP_ERR_ARG, "column out of range");
  NNPP_HIP(c, hipSetDevice(c->device));
  if (!c->row_ptr) NNPP_HIP(c, hipMalloc((void**)&c->row_ptr, sizeof(int64_t) * ((size_t)n + 1)));
  if (nnz > c->col_cap) {
    dfree(c->col);
    NNPP_HIP(c, hipMalloc((void**)&c->col, sizeof(int32_t) * nnz));
    c->col_cap = nnz;
  }
  NNPP_HIP(c, hipMemcpy(c->row_ptr, row_ptr, sizeof(int64_t) * ((size_t)n + 1),
                        hipMemcpyHostToDevice));
  if (nnz) NNPP_HIP(c, hipMemcpy(c->col, col, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
  c->nnz = nnz;
  c->max_len = longest + 1;
  c->geom_dirty = c->pot_dirty = true;
  c->pst.state = 0;
  return NNPP_OK;
}

int nnpp_set_bc(nnpp_ctx* c, const int32_t* bctypes, const double* bcvalues) {
  if (!c) return NNPP_ERR_ARG;
  if (int g_ = tape_guard(c)) return g_;
  int st = need_params(c);
  if (st) return st;
  if (!c->x) return nnpp_set_error(c, NNPP_ERR_STATE, "nnpp_set_nodes not called");
  const int32_t n = c->n;
  const int S = c->p.n_species;
  const int64_t F = fields(c);
  std::vector<uint16_t> mask(n, 0);
  std::vector<double> val((size_t)n * F, 0.0);
  if (bctypes && bcvalues)
    for (int32_t i = 0; i < n; i++) {
      if (bctypes[i] != 0) continue;  // main.cpp:570
      for (int f = 0; f < F; f++) {
        const double v = bcvalues[(int64_t)f * n + i];
        if (v == -1) continue;
        mask[i] |= (uint16_t)(1u << f);
        val[(int64_t)f * n + i] = (f < S) ? v / c->p.CSolid : v;  // main.cpp:571-625
      }
    }
  NNPP_HIP(c, hipSetDevice(c->device));
  if (n) {
    NNPP_HIP(c, hipMemcpy(c->bcmask, mask.data(), sizeof(uint16_t) * n, hipMemcpyHostToDevice));
    NNPP_HIP(c, hipMemcpy(c->bcval, val.data(), sizeof(double) * n * F, hipMemcpyHostToDevice));
  }
  c->have_bc = true;
  c->pot_dirty = true;
  return NNPP_OK;
}

int nnpp_set_state(nnpp_ctx* c, int32_t slot, const double* U) {
  if (!c || (!U && c->n) || slot < 0 || slot > 1) return NNPP_ERR_ARG;
  if (int g_ = tape_guard(c)) return g_;
  int st = need_params(c);
  if (st) return st;
  if (!c->x) return nnpp_set_error(c, NNPP_ERR_STATE, "nnpp_set_nodes not called");
  NNPP_HIP(c, hipSetDevice(c->device));
  NNPP_HIP(c, hipStreamSynchronize(c->stream));
  if (c->n)
    NNPP_HIP(c, hipMemcpy(c->U[slot], U, sizeof(double) * c->n * fields(c), hipMemcpyHostToDevice));
  c->imp_first = true;  // an implicit step starts again from U^n with the Dirichlet values
  c->phi_hist = 0;
  return NNPP_OK;
}

int nnpp_get_state(nnpp_ctx* c, int32_t slot, double* U) {
  if (!c || (!U && c->n) || slot < 0 || slot > 1) return NNPP_ERR_ARG;
  if (!c->x) return nnpp_set_error(c, NNPP_ERR_STATE, "nnpp_set_nodes not called");
  NNPP_HIP(c, hipSetDevice(c->device));
  NNPP_HIP(c, hipStreamSynchronize(c->stream));
  if (c->n)
    NNPP_HIP(c, hipMemcpy(U, c->U[slot], sizeof(double) * c->n * fields(c), hipMemcpyDeviceToHost));
  return NNPP_OK;
}

int nnpp_set_phase(nnpp_ctx* c, const int32_t* phase) {
  if (!c || (!phase && c->n)) return NNPP_ERR_ARG;
  if (int g_ = tape_guard(c)) return g_;
  if (!c->x) return nnpp_set_error(c, NNPP_ERR_STATE, "nnpp_set_nodes not called");
  std::vector<uint8_t> h(c->n);
  int64_t sc = 0;
  for (int32_t i = 0; i < c->n; i++) {
    h[i] = phase[i] ? 1 : 0;
    sc += h[i];
  }
  NNPP_HIP(c, hipSetDevice(c->device));
  if (c->n) NNPP_HIP(c, hipMemcpy(c->phase, h.data(), c->n, hipMemcpyHostToDevice));
  c->solid_count = sc;
  c->pot_dirty = true;
  return NNPP_OK;
}

int nnpp_get_phase(nnpp_ctx* c, int32_t* phase) {
  if (!c || (!phase && c->n)) return NNPP_ERR_ARG;
  if (!c->x) return nnpp_set_error(c, NNPP_ERR_STATE, "nnpp_set_nodes not called");
  std::vector<uint8_t> h(c->n);
  NNPP_HIP(c, hipSetDevice(c->device));
  NNPP_HIP(c, hipStreamSynchronize(c->stream));
  if (c->n) NNPP_HIP(c, hipMemcpy(h.data(), c->phase, c->n, hipMemcpyDeviceToHost));
  for (int32_t i = 0; i < c->n; i++) phase[i] = h[i];
  return NNPP_OK;
}

int nnpp_init_state(nnpp_ctx* c, const double* ic) {
  if (!c || (!ic && c->n)) return NNPP_ERR_ARG;
  if (int g_ = tape_guard(c)) return g_;
  int st = need_params(c);
  if (st) return st;
  if (!c->x) return nnpp_set_error(c, NNPP_ERR_STATE, "nnpp_set_nodes not called");
  const int32_t n = c->n;
  const int S = c->p.n_species;
  const int64_t F = fields(c);
  std::vector<double> U(ic, ic + (size_t)n * F);
  std::vector<int32_t> ph(n);
  int64_t sc = 0;
  double v0 = 0.0;  // main.cpp:476-479, from the initial values before any Dirichlet value
  for (int32_t i = 0; i < n; i++) {  // main.cpp:413-421
    ph[i] = (U[i] > 0.0 + c->p.phase_eps) ? 1 : 0;
    sc += ph[i];
    if (ph[i] == 1) v0 += (U[i] - c->p.CSat / c->p.CSolid);
  }
  if (c->have_bc && n) {  // Dirichlet values (main.cpp:569-625)
    std::vector<uint16_t> mask(n);
    std::vector<double> val((size_t)n * F);
    NNPP_HIP(c, hipMemcpy(mask.data(), c->bcmask, sizeof(uint16_t) * n, hipMemcpyDeviceToHost));
    NNPP_HIP(c, hipMemcpy(val.data(), c->bcval, sizeof(double) * n * F, hipMemcpyDeviceToHost));
    for (int32_t i = 0; i < n; i++)
      for (int f = 0; f < F; f++)
        if ((mask[i] >> f) & 1u) U[(int64_t)f * n + i] = val[(int64_t)f * n + i];
  }
  (void)S;
  st = nnpp_set_state(c, NNPP_SLOT_N, U.data());
  if (st) return st;
  st = nnpp_set_state(c, NNPP_SLOT_N1, U.data());
  if (st) return st;
  st = nnpp_set_phase(c, ph.data());
  if (st) return st;
  c->solid_count = sc;
  c->vol0rel = v0;
  c->voln1rel = 0.0;
  c->step = 1;
  return NNPP_OK;
}

int nnpp_set_time(nnpp_ctx* c, int64_t step, double voln1rel, double vol0rel) {
  if (!c) return NNPP_ERR_ARG;
  if (int g_ = tape_guard(c)) return g_;
  c->step = step;
  c->voln1rel = voln1rel;
  c->vol0rel = vol0rel;
  return NNPP_OK;
}

int nnpp_get_time(nnpp_ctx* c, int64_t* step, double* voln1rel, double* vol0rel) {
  if (!c) return NNPP_ERR_ARG;
  if (step) *step = c->step;
  if (voln1rel) *voln1rel = c->voln1rel;
  if (vol0rel) *vol0rel = c->vol0rel;
  return NNPP_OK;
}

int nnpp_residuals(nnpp_ctx* c, double voln1rel, int32_t apply_bc, int32_t precision,
                   double* F_out) {
  if (!c || (!F_out && c->n)) return NNPP_ERR_ARG;
  NNPP_HIP(c, hipSetDevice(c->device));
  int st = ensure_bonds(c);
  if (st) return st;
  if ((st = phi_row_guard(c))) return st;
  const double BV = std::pow(10., -c->p.vol_red * (1. - voln1rel));  // utils.h:100-116
  st = nnpp_launch_flux(c, 1, precision, BV, c->U[NNPP_SLOT_N1], c->U[NNPP_SLOT_N], c->Fres);
  if (st) return st;
  if (apply_bc) {
    st = nnpp_bc_override_dev(c, c->U[NNPP_SLOT_N1], c->Fres);
    if (st) return st;
  }
  NNPP_HIP(c, hipStreamSynchronize(c->stream));
  if (c->n)
    NNPP_HIP(c, hipMemcpy(F_out, c->Fres, sizeof(double) * c->n * fields(c), hipMemcpyDeviceToHost));
  return NNPP_OK;
}

int nnpp_poisson_apply(nnpp_ctx* c, const double* x, double* y) {
  if (!c || ((!x || !y) && c->n)) return NNPP_ERR_ARG;
  NNPP_HIP(c, hipSetDevice(c->device));
  int st = need_graph(c);
  if (st) return st;
  if (!c->n) return NNPP_OK;
  NNPP_HIP(c, hipMemcpy(c->cg_q, x, sizeof(double) * c->n, hipMemcpyHostToDevice));
  st = nnpp_potential_apply_dev(c, c->cg_q, c->cg_z);
  if (st) return st;
  NNPP_HIP(c, hipStreamSynchronize(c->stream));
  NNPP_HIP(c, hipMemcpy(y, c->cg_z, sizeof(double) * c->n, hipMemcpyDeviceToHost));
  return NNPP_OK;
}

int nnpp_set_solver(nnpp_ctx* c, double rtol, int32_t maxit, double inner_rtol) {
  if (!c || !(rtol >= 0) || maxit < 0 || !(inner_rtol > 0 && inner_rtol < 1)) return NNPP_ERR_ARG;
  if (int g_ = tape_guard(c)) return g_;
  c->rtol = rtol;
  c->maxit = maxit;
  c->inner_rtol = inner_rtol;
  return NNPP_OK;
}

int nnpp_set_format(nnpp_ctx* c, int32_t flags) {
  if (!c || (flags & ~(NNPP_FORMAT_WIDE_COLUMNS | NNPP_FORMAT_NO_DICTIONARY |
                       NNPP_FORMAT_NO_STENCIL | NNPP_FORMAT_NO_BALL | NNPP_FORMAT_NO_ZMARCH)))
    return NNPP_ERR_ARG;
  if (flags != c->format) c->geom_dirty = c->pot_dirty = true;
  c->format = flags;
  return NNPP_OK;
}

int nnpp_poisson_solve(nnpp_ctx* c, int32_t precision, int32_t* iters, double* relres) {
  if (!c) return NNPP_ERR_ARG;
  if (int g_ = tape_guard(c)) return g_;
  NNPP_HIP(c, hipSetDevice(c->device));
  int st = ensure_potential(c);
  if (st) return st;
  double* U = c->U[NNPP_SLOT_N];
  c->phi_hist = 0;  // phi^n no longer follows the explicit steps' history
  return nnpp_potential_solve_dev(c, precision, U, U + (int64_t)c->p.n_species * c->n, iters,
                                  relres, nullptr, nullptr);
}

int nnpp_front_update(nnpp_ctx* c, double threshold, int32_t* n_flipped) {
  if (!c) return NNPP_ERR_ARG;
  if (int g_ = tape_guard(c)) return g_;
  NNPP_HIP(c, hipSetDevice(c->device));
  int st = need_params(c);
  if (st) return st;
  if (!c->x) return nnpp_set_error(c, NNPP_ERR_STATE, "nnpp_set_nodes not called");
  return nnpp_front_dev(c, c->U[NNPP_SLOT_N], threshold, 0, n_flipped, nullptr);
}

}  // extern "C"

int nnpp_step_once(nnpp_ctx* c, int precision, nnpp_step_stats& loc, nnpp_tape_step* rec) {
  int st = ensure_potential(c);
  if (st) return st;
  const int32_t n = c->n;
  const int S = c->p.n_species;
  double* U = c->U[NNPP_SLOT_N];
  if (rec) {
    rec->step = c->step;
    rec->voln1rel = c->voln1rel;
  }
  // 1. species update from U^n (lagged phi^n)
  const double BV = std::pow(10., -c->p.vol_red * (1. - c->voln1rel));
  st = nnpp_launch_flux(c, 0, precision, BV, U, U, c->Unew);
  if (st) return st;
  // 2. phi^{n+1} from the charge of U^n, written into Unew; warm start on the free rows by
  //    extrapolation of phi^n and the previous potentials (quadratic once two are known)
  double* phi_new = c->Unew + (int64_t)S * n;
  const double* phi_n = U + (int64_t)S * n;
  if (n && !c->phi_prev)
    NNPP_HIP(c, hipMalloc((void**)&c->phi_prev, sizeof(double) * 2 * n));
  if (n && (st = nnpp_phi_guess_dev(c, c->phi_hist, phi_n, c->phi_prev, c->phi_prev + n,
                                    phi_new)))
    return st;
  int32_t its = 0;
  double rr = 0.0;
  int status = NNPP_OK;
  st = nnpp_potential_solve_dev(c, precision, U, phi_new, &its, &rr, &loc.cg_spmv_fp64,
                                &loc.cg_spmv_fp32);
  if (st == NNPP_ERR_NOCONV) status = NNPP_ERR_NOCONV;
  else if (st) return st;
  if (n) {  // history: [phi^n, phi^{n-1}]
    NNPP_HIP(c, hipMemcpyAsync(c->phi_prev + n, c->phi_prev, sizeof(double) * n,
                               hipMemcpyDeviceToDevice, c->stream));
    NNPP_HIP(c, hipMemcpyAsync(c->phi_prev, phi_n, sizeof(double) * n, hipMemcpyDeviceToDevice,
                               c->stream));
  }
  c->phi_hist = n > 0 ? std::min(c->phi_hist + 1, 2) : 0;
  std::swap(c->U[NNPP_SLOT_N], c->Unew);
  // 3. front update and 4. volume cadence
  int32_t flips = 0;
  double volsum = 0.0;
  st = nnpp_front_dev(c, c->U[NNPP_SLOT_N], c->p.front_threshold, 1, &flips, &volsum);
  if (st) return st;
  const bool vol_update = c->p.sub_steps > 0 && (c->step % c->p.sub_steps) == 0 &&
                          c->vol0rel != 0.0;
  if (vol_update) c->voln1rel = volsum / c->vol0rel;
  if (rec) {
    rec->flips = flips;
    rec->vol_update = vol_update ? 1 : 0;
  }
  c->step++;
  loc.steps++;
  loc.cg_iterations += its;
  loc.last_relres = rr;
  loc.flipped += flips;
  loc.nnz_flux = c->bonds.nnz;
  loc.nnz_ll = c->pot.nnz;
  loc.pot_stencil = nnpp_pot_form(c);
  loc.flux_stencil = !nnpp_flux_stencil_active(c) ? 0
                     : precision == NNPP_MIXED      ? 2
                     : nnpp_flux_zmarch_active(c)   ? 3
                                                    : 1;
  loc.pot_col_bytes = c->pst.active ? 0 : (c->pot.wide ? 4 : 2);
  loc.pot_weight_bytes = c->pst.active ? 0 : (c->pot.ncls > 0 ? 1 : 4);
  return status;
}

extern "C" {

int nnpp_step(nnpp_ctx* c, int32_t nsteps, int32_t precision, int32_t check_finite,
              nnpp_step_stats* stats) {
  if (!c || nsteps < 0) return NNPP_ERR_ARG;
  if (precision != NNPP_FP64 && precision != NNPP_MIXED) return NNPP_ERR_ARG;
  if (c->tape.active && precision != NNPP_FP64)
    return nnpp_set_error(c, NNPP_ERR_UNSUPPORTED, "a differentiation tape records FP64 steps only");
  NNPP_HIP(c, hipSetDevice(c->device));
  int st = ensure_bonds(c);
  if (st) return st;
  c->imp_first = true;  // the implicit rotation buffers no longer follow the state
  nnpp_step_stats loc;
  std::memset(&loc, 0, sizeof(loc));
  int status = NNPP_OK;
  for (int32_t k = 0; k < nsteps; k++) {
    nnpp_tape_step rec{};
    if (c->tape.active && c->tape.steps.size() % (size_t)c->tape.every == 0) {
      st = tape_checkpoint(c);
      if (st) return st;
    }
    st = nnpp_step_once(c, precision, loc, c->tape.active ? &rec : nullptr);
    if (st == NNPP_ERR_NOCONV) status = NNPP_ERR_NOCONV;
    else if (st) {
      if (stats) *stats = loc;
      return st;
    }
    if (c->tape.active) c->tape.steps.push_back(rec);
    if (c->solid_count <= 0) break;  // main.cpp:924-932
  }
  loc.solid_count = c->solid_count;
  loc.voln1rel = c->voln1rel;
  if (check_finite) {
    int32_t bad = 0;
    st = nnpp_check_finite_dev(c, c->U[NNPP_SLOT_N], &bad);
    if (st) return st;
    if (bad) status = nnpp_set_error(c, NNPP_ERR_NONFINITE, "non-finite values in the state");
  }
  NNPP_HIP(c, hipStreamSynchronize(c->stream));
  if (stats) *stats = loc;
  return status;
}

int nnpp_implant_classify(nnpp_ctx* c, int64_t nb, const double* bx, const double* by,
                          const double* bz, double r, int32_t max_count, int32_t* counts,
                          double* ic_out) {
  if (!c || nb < 0 || (nb && (!bx || !by || !bz)) || (!counts && c->n) || !(r >= 0) ||
      max_count < 0)
    return NNPP_ERR_ARG;
  int st = need_params(c);
  if (st) return st;
  if (!c->x) return nnpp_set_error(c, NNPP_ERR_STATE, "nnpp_set_nodes not called");
  if (ic_out && c->p.n_species != 10)
    return nnpp_set_error(c, NNPP_ERR_UNSUPPORTED, "implant initial values need the 10 reference species");
  NNPP_HIP(c, hipSetDevice(c->device));
  const int32_t n = c->n;
  if (n == 0) return NNPP_OK;
  double bbox[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t k = 0; k < nb; k++) {
    const double v[3] = {bx[k], by[k], bz[k]};
    for (int a = 0; a < 3; a++) {
      if (k == 0 || v[a] < bbox[2 * a]) bbox[2 * a] = v[a];
      if (k == 0 || v[a] > bbox[2 * a + 1]) bbox[2 * a + 1] = v[a];
    }
  }
  if (!(std::isfinite(bbox[0]) && std::isfinite(bbox[1]) && std::isfinite(bbox[2]) &&
        std::isfinite(bbox[3]) && std::isfinite(bbox[4]) && std::isfinite(bbox[5])))
    return nnpp_set_error(c, NNPP_ERR_ARG, "non-finite implant coordinates");
  double* d = nullptr;
  int32_t* dc = nullptr;
  const size_t m = nb ? (size_t)nb : 1;
  NNPP_HIP(c, hipMalloc((void**)&d, sizeof(double) * 3 * m));
  if (hipMalloc((void**)&dc, sizeof(int32_t) * n) != hipSuccess) {
    hipFree(d);
    return nnpp_set_error(c, NNPP_ERR_HIP, "implant classification allocation");
  }
  auto done = [&](int e) {
    hipFree(d);
    hipFree(dc);
    return e;
  };
  if (nb && (hipMemcpy(d, bx, sizeof(double) * nb, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(d + m, by, sizeof(double) * nb, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(d + 2 * m, bz, sizeof(double) * nb, hipMemcpyHostToDevice) != hipSuccess))
    return done(nnpp_set_error(c, NNPP_ERR_HIP, "implant point upload"));
  st = nnpp_implant_count_dev(c, nb, d, d + m, d + 2 * m, bbox, r, max_count, dc);
  if (st && st != NNPP_ERR_OVERFLOW) return done(st);
  if (hipMemcpy(counts, dc, sizeof(int32_t) * n, hipMemcpyDeviceToHost) != hipSuccess)
    return done(nnpp_set_error(c, NNPP_ERR_HIP, "count download"));
  if (ic_out) nnpp_implant_icvalues_host(c, counts, ic_out);
  return done(st);
}


int nnpp_voronoi_volumes(nnpp_ctx* c, double h, double margin, double radius,
                         double* vol_out) {
  if (!c || !(h > 0) || !(margin >= 0) || !(radius > 0) || (!vol_out && c->n))
    return NNPP_ERR_ARG;
  if (!c->x) return nnpp_set_error(c, NNPP_ERR_STATE, "nnpp_set_nodes not called");
  NNPP_HIP(c, hipSetDevice(c->device));
  const int32_t n = c->n;
  if (!n) return NNPP_OK;
  unsigned long long* cnt = nullptr;
  NNPP_HIP(c, hipMalloc((void**)&cnt, sizeof(unsigned long long) * n));
  int64_t dims[3];
  int st = nnpp_voronoi_dev(c, h, margin, radius, cnt, dims);
  std::vector<unsigned long long> hc(n);
  if (!st && hipMemcpy(hc.data(), cnt, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost) !=
                 hipSuccess)
    st = nnpp_set_error(c, NNPP_ERR_HIP, "count download");
  hipFree(cnt);
  if (st) return st;
  const double pv = h * h * h;
  for (int32_t i = 0; i < n; i++) vol_out[i] = (double)hc[i] * pv;
  return NNPP_OK;
}

int nnpp_newton_defaults(nnpp_newton_opts* o) {
  if (!o) return NNPP_ERR_ARG;
  o->newton_tol = 10E-6;  // EPS, main.cpp:35
  o->lin_rtol = 10E-6;
  o->newton_maxit = 200;  // MAX_ITER, main.cpp:36
  o->lin_maxit = 200;
  o->start = NNPP_NEWTON_START_ROTATE;
  o->reserved = 0;
  return NNPP_OK;
}

int nnpp_step_implicit(nnpp_ctx* c, int32_t nsteps, int32_t precision, int32_t check_finite,
                       const nnpp_newton_opts* opts, nnpp_newton_stats* stats) {
  if (!c || nsteps < 0) return NNPP_ERR_ARG;
  if (precision != NNPP_FP64 && precision != NNPP_MIXED) return NNPP_ERR_ARG;
  nnpp_newton_opts o;
  nnpp_newton_defaults(&o);
  if (opts) o = *opts;
  if (!(o.newton_tol >= 0) || !(o.lin_rtol >= 0) || o.newton_maxit < 0 || o.lin_maxit < 0 ||
      (o.start != NNPP_NEWTON_START_ROTATE && o.start != NNPP_NEWTON_START_STATE))
    return nnpp_set_error(c, NNPP_ERR_ARG, "invalid Newton options");
  if (c->tape.active)
    return nnpp_set_error(c, NNPP_ERR_UNSUPPORTED, "a differentiation tape records explicit steps only");
  NNPP_HIP(c, hipSetDevice(c->device));
  int st = ensure_bonds(c);
  if (st) return st;
  if ((st = phi_row_guard(c))) return st;
  nnpp_newton_stats loc;
  std::memset(&loc, 0, sizeof(loc));
  int status = NNPP_OK;
  for (int32_t k = 0; k < nsteps; k++) {
    if ((st = ensure_potential(c))) return st;
    st = nnpp_implicit_once(c, precision, o, loc);
    if (st == NNPP_ERR_NOCONV) status = NNPP_ERR_NOCONV;
    else if (st) {
      if (stats) *stats = loc;
      return st;
    }
    if (c->solid_count <= 0) break;  // main.cpp:924-932
  }
  loc.solid_count = c->solid_count;
  loc.voln1rel = c->voln1rel;
  if (check_finite) {
    int32_t bad = 0;
    st = nnpp_check_finite_dev(c, c->U[NNPP_SLOT_N], &bad);
    if (st) return st;
    if (bad) status = nnpp_set_error(c, NNPP_ERR_NONFINITE, "non-finite values in the state");
  }
  NNPP_HIP(c, hipStreamSynchronize(c->stream));
  if (stats) *stats = loc;
  return status;
}

int nnpp_jacobian_apply(nnpp_ctx* c, double voln1rel, const double* in, double* out,
                        double* diag) {
  if (!c || ((!in || !out) && c->n)) return NNPP_ERR_ARG;
  NNPP_HIP(c, hipSetDevice(c->device));
  int st = ensure_bonds(c);
  if (st) return st;
  if (!c->n) return NNPP_OK;
  const size_t bytes = sizeof(double) * c->n * fields(c);
  double* d = nullptr;
  NNPP_HIP(c, hipMalloc((void**)&d, 3 * bytes));
  double* din = d;
  double* dout = (double*)((char*)d + bytes);
  double* ddg = (double*)((char*)d + 2 * bytes);
  auto done = [&](int e) {
    hipFree(d);
    return e;
  };
  if (hipMemcpy(din, in, bytes, hipMemcpyHostToDevice) != hipSuccess)
    return done(nnpp_set_error(c, NNPP_ERR_HIP, "copy"));
  st = nnpp_jacobian_apply_dev(c, voln1rel, din, dout, diag ? ddg : nullptr);
  if (st) return done(st);
  if (hipStreamSynchronize(c->stream) != hipSuccess ||
      hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
      (diag && hipMemcpy(diag, ddg, bytes, hipMemcpyDeviceToHost) != hipSuccess))
    return done(nnpp_set_error(c, NNPP_ERR_HIP, "copy"));
  return done(NNPP_OK);
}

int64_t nnpp_abi_newton_stats_size(void) { return (int64_t)sizeof(nnpp_newton_stats); }

int nnpp_tape_begin(nnpp_ctx* c, int32_t every) {
  if (!c || every < 1) return NNPP_ERR_ARG;
  if (int g_ = tape_guard(c)) return g_;
  int st = need_graph(c);
  if (st) return st;
  if (!c->x) return nnpp_set_error(c, NNPP_ERR_STATE, "nnpp_set_nodes not called");
  if (c->p.chemistry != NNPP_CHEM_GENERIC)
    return nnpp_set_error(c, NNPP_ERR_UNSUPPORTED, "the adjoint covers the generic chemistry");
  nnpp_tape_free(c);
  c->tape.active = true;
  c->tape.every = every;
  return NNPP_OK;
}

int nnpp_tape_adjoint(nnpp_ctx* c, const double* lam, double* grad, double* lam0,
                      nnpp_adjoint_stats* stats) {
  if (!c || (!lam && c->n) || !grad) return NNPP_ERR_ARG;
  if (!c->tape.active) return nnpp_set_error(c, NNPP_ERR_STATE, "no tape (nnpp_tape_begin)");
  NNPP_HIP(c, hipSetDevice(c->device));
  int st = ensure_bonds(c);
  if (st) return st;
  return nnpp_adjoint_run(c, lam, grad, lam0, stats);
}

int nnpp_tape_end(nnpp_ctx* c) {
  if (!c) return NNPP_ERR_ARG;
  NNPP_HIP(c, hipSetDevice(c->device));
  NNPP_HIP(c, hipStreamSynchronize(c->stream));
  nnpp_tape_free(c);
  return NNPP_OK;
}

int64_t nnpp_abi_adjoint_stats_size(void) { return (int64_t)sizeof(nnpp_adjoint_stats); }

int nnpp_state_device_ptr(nnpp_ctx* c, int32_t slot, int32_t field, void** dptr) {
  if (!c || !dptr || slot < 0 || slot > 1 || field < 0 || field > c->p.n_species)
    return NNPP_ERR_ARG;
  if (!c->x) return nnpp_set_error(c, NNPP_ERR_STATE, "nnpp_set_nodes not called");
  *dptr = (void*)(c->U[slot] + (int64_t)field * c->n);
  return NNPP_OK;
}

int nnpp_stream(nnpp_ctx* c, void** stream) {
  if (!c || !stream) return NNPP_ERR_ARG;
  *stream = (void*)c->stream;
  return NNPP_OK;
}

int nnpp_potential_operator_apply(nnpp_ctx* c, int32_t precision, const double* x, double* y) {
  if (!c || ((!x || !y) && c->n)) return NNPP_ERR_ARG;
  if (precision != NNPP_FP64 && precision != NNPP_MIXED) return NNPP_ERR_ARG;
  NNPP_HIP(c, hipSetDevice(c->device));
  int st = ensure_potential(c);
  if (st) return st;
  const int32_t n = c->n;
  if (!n) return NNPP_OK;
  hipStream_t s = c->stream;
  const double zero = 0.0;
  NNPP_HIP(c, hipMemcpyAsync(c->scal + SC_DONE, &zero, sizeof(double), hipMemcpyHostToDevice, s));
  std::vector<float> xf;
  if (precision == NNPP_MIXED) {
    xf.assign(x, x + n);
    NNPP_HIP(c, hipMemcpyAsync(c->f_p, xf.data(), sizeof(float) * n, hipMemcpyHostToDevice, s));
  } else {
    NNPP_HIP(c, hipMemcpyAsync(c->cg_p, x, sizeof(double) * n, hipMemcpyHostToDevice, s));
  }
  st = nnpp_pot_spmv_launch(c, precision);
  if (st) return st;
  if (precision == NNPP_MIXED) {
    NNPP_HIP(c, hipMemcpyAsync(xf.data(), c->f_q, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    NNPP_HIP(c, hipStreamSynchronize(s));
    for (int32_t i = 0; i < n; i++) y[i] = xf[i];
  } else {
    NNPP_HIP(c, hipMemcpyAsync(y, c->cg_q, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    NNPP_HIP(c, hipStreamSynchronize(s));
  }
  return NNPP_OK;
}

int nnpp_time_kernel(nnpp_ctx* c, int32_t kernel, int32_t precision, int32_t reps,
                     double* avg_ms) {
  if (!c || !avg_ms || reps < 1) return NNPP_ERR_ARG;
  if (precision != NNPP_FP64 && precision != NNPP_MIXED) return NNPP_ERR_ARG;
  NNPP_HIP(c, hipSetDevice(c->device));
  int st = ensure_bonds(c);
  if (st) return st;
  st = ensure_potential(c);
  if (st) return st;
  hipStream_t s = c->stream;
  const double zero = 0.0;
  NNPP_HIP(c, hipMemcpyAsync(c->scal + SC_DONE, &zero, sizeof(double), hipMemcpyHostToDevice, s));
  NNPP_HIP(c, hipStreamSynchronize(s));
  hipEvent_t e0, e1;
  NNPP_HIP(c, hipEventCreate(&e0));
  NNPP_HIP(c, hipEventCreate(&e1));
  const double BV = std::pow(10., -c->p.vol_red * (1. - c->voln1rel));
  double* U = c->U[NNPP_SLOT_N];
  // one untimed launch, then the timed ones
  for (int r = 0; r <= reps; r++) {
    if (r == 1) NNPP_HIP(c, hipEventRecord(e0, s));
    if (kernel == NNPP_KERNEL_FLUX) st = nnpp_launch_flux(c, 0, precision, BV, U, U, c->Unew);
    else if (kernel == NNPP_KERNEL_POT_SPMV) st = nnpp_pot_spmv_launch(c, precision);
    else if (kernel == NNPP_KERNEL_CG_ITER) st = nnpp_cg_iter_launch(c, precision);
    else st = NNPP_ERR_ARG;
    if (st) break;
  }
  NNPP_HIP(c, hipEventRecord(e1, s));
  NNPP_HIP(c, hipEventSynchronize(e1));
  float ms = 0.f;
  NNPP_HIP(c, hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (st) return st;
  *avg_ms = (double)ms / reps;
  return NNPP_OK;
}

}  // extern "C"

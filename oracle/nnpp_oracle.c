/*
 * nnpp_oracle.c — CPU FP64 restatement of the reference's NNPP step path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline as the checker; the product library never uses it.
 *
 * Each function names the reference lines it restates (paths relative to the reference
 * repository). Expressions keep the reference's operand order so that FP64 results agree to
 * the last few ulps; summation over a node's bonds runs in ascending neighbour id (the
 * reference runs in its cell order, grid.h:163-194), which is the only intended difference.
 * Build with -ffp-contract=off (the reference is built without FMA, CMakeLists.txt:52).
 */
#include "nnpp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline double sqr_(double v) { return v * v; } /* grid.h:22-24 */

/* utils.h:17-25 */
double or_beta(double xi, double delta, double dx) {
  if (xi <= delta - dx / 2.0) {
    return 1.0;
  } else if ((xi > delta - dx / 2.0) && (xi <= delta + dx / 2.0)) {
    return (delta + dx / 2.0 - xi) / (dx);
  } else {
    return 0.0;
  }
}

/* grid.h:27-41 */
int or_sphere_cube_intersects(double cx, double cy, double cz, double cr, double rx, double ry,
                              double rz, double rw, double rh, double rd) {
  const double bmin[3] = {rx, ry, rz};
  const double bmax[3] = {rx + rw, ry + rh, rz + rd};
  const double c[3] = {cx, cy, cz};
  const double r2 = sqr_(cr);
  double dmin = 0.0;
  for (int a = 0; a < 3; a++) {
    if (c[a] < bmin[a]) dmin += sqr_(c[a] - bmin[a]);
    else if (c[a] > bmax[a]) dmin += sqr_(c[a] - bmax[a]);
  }
  return dmin <= r2;
}

static int cmp_i32(const void* a, const void* b) {
  int32_t u = *(const int32_t*)a, v = *(const int32_t*)b;
  return (u > v) - (u < v);
}

/*
 * grid.h:43-204 (build_clouds_hash). Steps 0-4 are restated exactly: bounding box, hash
 * ix + mx*iy + mx*my*iz with truncating casts, cell index, cell corners x0 + r*dx. Step 5
 * applies the same sphere-cube test and the same candidate predicate; instead of testing every
 * cell (grid.h:163) it tests the cells of the index box one cell wider than the sphere's
 * bounding box, which contains every cell the test can accept (a cell outside it is at least
 * one cell width beyond the sphere). Within a row the accepted ids are sorted ascending.
 * col == NULL: fill row_ptr (and max_len_out) only.
 */
int or_build_neighbors(int32_t n, const double* x, const double* y, const double* z, double r,
                       double cdx, double cdy, double cdz, int32_t max_len, int64_t* row_ptr,
                       int32_t* col, int32_t* max_len_out) {
  if (n < 0 || max_len < 1 || !(cdx > 0) || !(cdy > 0) || !(cdz > 0)) return NNPP_ERR_ARG;
  row_ptr[0] = 0;
  if (max_len_out) *max_len_out = 0;
  if (n == 0) return NNPP_OK;

  double x0 = x[0], x1 = x[0], y0 = y[0], y1 = y[0], z0 = z[0], z1 = z[0];
  for (int32_t i = 0; i < n; i++) { /* grid.h:59-71 */
    x0 = x[i] < x0 ? x[i] : x0;
    x1 = x1 < x[i] ? x[i] : x1;
    y0 = y[i] < y0 ? y[i] : y0;
    y1 = y1 < y[i] ? y[i] : y1;
    z0 = z[i] < z0 ? z[i] : z0;
    z1 = z1 < z[i] ? z[i] : z1;
  }
  const int mx = (int)((x1 - x0) / cdx) + 1; /* grid.h:78-79 */
  const int my = (int)((y1 - y0) / cdy) + 1;
  int* hash = (int*)malloc(sizeof(int) * (size_t)n);
  int max_hash = 0;
  for (int32_t i = 0; i < n; i++) { /* grid.h:82-85 */
    hash[i] = (int)((x[i] - x0) / cdx) + mx * (int)((y[i] - y0) / cdy) +
              mx * my * (int)((z[i] - z0) / cdz);
    if (hash[i] + 1 > max_hash) max_hash = hash[i] + 1;
  }
  /* grid.h:90-124: nodes grouped by cell (counting sort; order inside a cell is irrelevant
   * once rows are sorted by id). */
  int64_t* start = (int64_t*)calloc((size_t)max_hash + 1, sizeof(int64_t));
  int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  for (int32_t i = 0; i < n; i++) start[hash[i] + 1]++;
  for (int m = 0; m < max_hash; m++) start[m + 1] += start[m];
  {
    int64_t* fillp = (int64_t*)malloc(sizeof(int64_t) * (size_t)max_hash);
    memcpy(fillp, start, sizeof(int64_t) * (size_t)max_hash);
    for (int32_t i = 0; i < n; i++) order[fillp[hash[i]]++] = i;
    free(fillp);
  }
  const int mxy = mx * my;
  const int mz_last = (max_hash - 1) / mxy;
  const double r2 = sqr_(r);
  int32_t* counts = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int status = NNPP_OK;

#pragma omp parallel
  {
    int32_t cap = 4096;
    int32_t* buf = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap);
#pragma omp for schedule(dynamic, 256)
    for (int32_t i = 0; i < n; i++) {
      const double xi = x[i], yi = y[i], zi = z[i];
      int ixl = (int)floor((xi - r - x0) / cdx) - 1, ixh = (int)floor((xi + r - x0) / cdx) + 1;
      int iyl = (int)floor((yi - r - y0) / cdy) - 1, iyh = (int)floor((yi + r - y0) / cdy) + 1;
      int izl = (int)floor((zi - r - z0) / cdz) - 1, izh = (int)floor((zi + r - z0) / cdz) + 1;
      if (ixl < 0) ixl = 0;
      if (iyl < 0) iyl = 0;
      if (izl < 0) izl = 0;
      if (ixh > mx - 1) ixh = mx - 1;
      if (iyh > my - 1) iyh = my - 1;
      if (izh > mz_last) izh = mz_last;
      int32_t cnt = 0;
      for (int iz = izl; iz <= izh; iz++)
        for (int iy = iyl; iy <= iyh; iy++)
          for (int ix = ixl; ix <= ixh; ix++) {
            const int m = ix + mx * iy + mxy * iz;
            if (m >= max_hash) continue;
            const int rr = (m % mxy) % mx, cc = (m % mxy) / mx, pp = m / mxy; /* grid.h:134-139 */
            const double bx = x0 + rr * cdx, by = y0 + cc * cdy, bz = z0 + pp * cdz;
            if (!or_sphere_cube_intersects(xi, yi, zi, r, bx, by, bz, cdx, cdy, cdz)) continue;
            for (int64_t k = start[m]; k < start[m + 1]; k++) { /* grid.h:184-193 */
              const int32_t cand = order[k];
              if (sqr_(xi - x[cand]) + sqr_(yi - y[cand]) + sqr_(zi - z[cand]) <= r2 &&
                  cand != i) {
                if (cnt == cap) {
                  cap *= 2;
                  buf = (int32_t*)realloc(buf, sizeof(int32_t) * (size_t)cap);
                }
                buf[cnt++] = cand;
              }
            }
          }
      counts[i] = cnt;
      if (cnt >= max_len) { /* grid.h:187-189: fill (self included) may not reach max_node_no */
#pragma omp atomic write
        status = NNPP_ERR_OVERFLOW;
      } else if (col) {
        qsort(buf, (size_t)cnt, sizeof(int32_t), cmp_i32);
        memcpy(col + row_ptr[i], buf, sizeof(int32_t) * (size_t)cnt);
      }
    }
    free(buf);
  }
  if (!col) {
    int32_t mlen = 0;
    for (int32_t i = 0; i < n; i++) {
      row_ptr[i + 1] = row_ptr[i] + counts[i];
      if (counts[i] + 1 > mlen) mlen = counts[i] + 1; /* grid.h:198-203, lengths include self */
    }
    if (max_len_out) *max_len_out = mlen;
  } else if (max_len_out) {
    int32_t mlen = 0;
    for (int32_t i = 0; i < n; i++)
      if (counts[i] + 1 > mlen) mlen = counts[i] + 1;
    *max_len_out = mlen;
  }
  free(counts);
  free(order);
  free(start);
  free(hash);
  return status;
}

/* Volume of neighbour nn in a bond term: the reference multiplies by dx*dy*dz
 * (utils.h:122-124); with per-node volumes it is vol[nn]. */
static inline double bond_vol(const double* vol, const nnpp_params* p, double bval, int32_t nn) {
  if (vol) return bval * vol[nn];
  return bval * p->dx * p->dy * p->dz;
}

/* Reference reactions R1..R10 (utils.h:162-174, 241-253, 321-330, 396-401, 466-476, 541-547,
 * 612-618, 683-690; none for Na+ and Cl-). u1 = iterate U^{n+1} at node i, un = U^n at node i
 * (indexed by 0-based field), phin = lagged potential at node i. */
static double reference_reaction(int k, const nnpp_params* p, int phase, int intf,
                                 const double* u1, const double* un, double phin) {
  const double CS = p->CSolid;
  const int liquid = (phase == 0);
  double R = 0.0;
  switch (k) {
    case 0: /* Mg2+ */
      if (intf == 1 || liquid)
        R = -p->k1b * fmax((u1[0] * un[2] * un[2] - pow(10., -p->pKMgOH2) / CS), 0.0)
            - p->k2b * fmax((u1[0] * un[4] - pow(10., -p->pKMgCO3) / CS), 0.0)
            - p->k3b * fmax((u1[0] * un[4] - pow(10., -p->pKMgHCO3) / CS), 0.0)
            - p->k4b * fmax((un[0] * un[0] * un[0] * un[5] * un[5] - pow(10., -p->pKMgPO4) / CS), 0.0);
      break;
    case 1: /* H+ */
      if (intf == 1) {
        double phic = (p->R * p->T / (p->F * p->z[0])) * log(un[0] / (p->CSat / p->CSolid));
        double eta = p->phim + p->phimse - phin - phic;
        R += (p->JHp / (p->z[1] * p->F * p->CSolid)) * exp(-p->alphaC * p->F * eta / (p->R * p->T));
      } else if (liquid) {
        R = p->k7b * (pow(10., -p->pKH2O) / CS - u1[1] * un[2])
            + p->k8b * (pow(10., -p->pKHCO3) / CS - u1[1] * un[4])
            + p->k9b * (pow(10., -p->pKHPO4) / CS - u1[1] * un[6]);
      }
      break;
    case 2: /* OH- */
      if (intf == 1) {
        R = -p->k1b * fmax((un[2] * un[2] * un[0] - pow(10., -p->pKMgOH2) / CS), 0.0)
            - p->k6b * fmax((u1[2] * un[7] * un[7] * un[7] * un[7] * un[7] * un[6] * un[6] * un[6] - pow(10., -p->pKCa5OHPO4) / CS), 0.0);
      } else if (liquid) {
        R = -p->k1b * fmax((un[2] * un[2] * un[0] - pow(10., -p->pKMgOH2) / CS), 0.0)
            - p->k6b * fmax((u1[2] * un[7] * un[7] * un[7] * un[7] * un[7] * un[6] * un[6] * un[6] - pow(10., -p->pKCa5OHPO4) / CS), 0.0)
            + p->k7b * (pow(10., -p->pKH2O) / CS - u1[2] * un[1]);
      }
      break;
    case 3: /* HCO3- */
      if (phase == 1 && intf == 1) R = 0.0;
      else if (liquid) R = p->k8b * (pow(10., -p->pKHCO3) / CS - u1[3] * un[1]);
      break;
    case 4: /* CO3 2- */
      if (phase == 1 && intf == 1)
        R = -p->k2b * fmax((u1[4] * un[0] - pow(10., -p->pKMgCO3) / CS), 0.0)
            - p->k3b * fmax((u1[4] * un[0] - pow(10., -p->pKMgHCO3) / CS), 0.0)
            - p->k5b * fmax((u1[4] * un[7] - pow(10., -p->pKCaCO3) / CS), 0.0);
      else if (liquid)
        R = -p->k2b * fmax((u1[4] * un[0] - pow(10., -p->pKMgCO3) / CS), 0.0)
            - p->k3b * fmax((u1[4] * un[0] - pow(10., -p->pKMgHCO3) / CS), 0.0)
            - p->k5b * fmax((u1[4] * un[7] - pow(10., -p->pKCaCO3) / CS), 0.0)
            + p->k8b * (pow(10., -p->pKHCO3) / CS - un[1] * u1[4]);
      break;
    case 5: /* HPO4 2- */
      if (phase == 1 && intf == 1)
        R = -p->k4b * fmax((un[0] * un[0] * un[0] * un[5] * un[5] - pow(10., -p->pKMgPO4) / CS), 0.0);
      else if (liquid)
        R = -p->k4b * fmax((un[0] * un[0] * un[0] * un[5] * un[5] - pow(10., -p->pKMgPO4) / CS), 0.0)
            + p->k9b * (-pow(10., -p->pKHPO4) / CS + un[6] * un[1]);
      break;
    case 6: /* PO4 3- */
      if (phase == 1 && intf == 1)
        R = -p->k6b * fmax((un[2] * un[7] * un[7] * un[7] * un[7] * un[7] * un[6] * un[6] * un[6] - pow(10., -p->pKCa5OHPO4) / CS), 0.0);
      else if (liquid)
        R = -p->k6b * fmax((un[2] * un[7] * un[7] * un[7] * un[7] * un[7] * un[6] * un[6] * un[6] - pow(10., -p->pKCa5OHPO4) / CS), 0.0)
            + p->k9b * (pow(10., -p->pKHPO4) / CS - u1[6] * un[1]);
      break;
    case 7: /* Ca2+ */
      if ((phase == 1 && intf == 1) || liquid)
        R = -p->k5b * fmax((u1[7] * un[4] - pow(10., -p->pKCaCO3) / CS), 0.0)
            - p->k6b * fmax((un[2] * un[7] * un[7] * un[7] * un[7] * un[7] * un[6] * un[6] * un[6] - pow(10., -p->pKCa5OHPO4) / CS), 0.0);
      break;
    default: /* Na+, Cl- */
      break;
  }
  return R;
}

/* Generic chemistry (no reference counterpart): first-order sink of species 0 at interface
 * solid nodes. */
static double generic_reaction(int k, const nnpp_params* p, int phase, int intf,
                               const double* u1) {
  if (k == 0 && phase == 1 && intf == 1) return -p->r_diss * u1[0];
  return 0.0;
}

/*
 * utils.h:27-842 (Residuals), one pass per field as in the reference, generalised to S species:
 *   species k: F = (U1_k,i - Un_k,i)/dt + sum_KR_k1 + sum_KR_k2 - R_k,
 *     sum_KR_k1 += -(K1 D BV (U1_k,n - U1_k,i) / xi^(3+2 s1)) Beta dx dy dz,
 *     sum_KR_k2 += -(K2 D BV z F U1_k,n (phin_n - phin_i) / (R T xi^(3+2 s2))) Beta dx dy dz,
 *     D/BV by phase pair (utils.h:100-159 for species with solid_bonds, 211-238 otherwise);
 *   phi: F = (1/(K sigma1)) sum_LL -(K sigma1 (phi1_n - phi1_i)/xi^(3+2 s111)) Beta dx dy dz
 *            - (1/(K sigma1)) F CSolid sum_k z_k Un_k   on rows with a liquid-liquid bond,
 *        0 elsewhere (utils.h:798-841).
 */
int or_residuals(int32_t n, const nnpp_params* p, const double* x, const double* y,
                 const double* z, const double* vol, const int64_t* row_ptr, const int32_t* col,
                 const int32_t* phase, const double* U1, const double* Un, double voln1rel,
                 double* Fout) {
  const int S = p->n_species;
  if (S < 1 || S > NNPP_MAX_SPECIES) return NNPP_ERR_ARG;
  if (p->chemistry == NNPP_CHEM_REFERENCE && S != 10) return NNPP_ERR_UNSUPPORTED;
  const double BV = pow(10., -p->vol_red * (1. - voln1rel));
  const double* phin = Un + (size_t)S * n;  /* lagged potential U11n */
  const double* phi1 = U1 + (size_t)S * n;  /* iterate potential U11n1 */

  for (int k = 0; k < S; k++) {
    const double* c1 = U1 + (size_t)k * n;
    const double* cn = Un + (size_t)k * n;
    const double K1 = p->K1[k], K2 = p->K2[k], Ds = p->D_solid[k], Dl = p->D_liquid[k];
    const double zk = p->z[k], ex1 = 3.0 + 2.0 * p->s1[k], ex2 = 3.0 + 2.0 * p->s2[k];
    const int sb = p->solid_bonds[k];
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; i++) {
      int intf = 0;
      if (phase[i] == 1) /* utils.h:82-89 */
        for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++)
          if (phase[col[b]] == 0) {
            intf = 1;
            break;
          }
      double s1 = 0.0, s2 = 0.0;
      for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
        const int32_t nn = col[b];
        const double xx = x[nn] - x[i], yy = y[nn] - y[i], zz = z[nn] - z[i];
        const double xi = sqrt(xx * xx + yy * yy + zz * zz);
        const int pi = phase[i], pn = phase[nn];
        double D, bv;
        if (pi == 1 && pn == 1) { /* utils.h:100-116: BV = 1 on solid-solid bonds */
          if (!sb) continue;
          D = Ds;
          bv = 1.0;
        } else if ((pi == 1 && pn == 0) || (pi == 0 && pn == 1)) {
          D = Ds;
          bv = sb ? BV : 1.0;
        } else {
          D = Dl;
          bv = 1.0;
        }
        const double be = or_beta(xi, p->delta, p->dx);
        s1 += bond_vol(vol, p, -(K1 * D * bv * (c1[nn] - c1[i]) / pow(xi, ex1)) * be, nn);
        s2 += bond_vol(vol, p, -(K2 * D * bv * zk * p->F * c1[nn] * (phin[nn] - phin[i]) /
                                 (p->R * p->T * pow(xi, ex2))) * be, nn);
      }
      double u1[NNPP_MAX_SPECIES], un[NNPP_MAX_SPECIES];
      for (int q = 0; q < S; q++) {
        u1[q] = U1[(size_t)q * n + i];
        un[q] = Un[(size_t)q * n + i];
      }
      const double R = (p->chemistry == NNPP_CHEM_REFERENCE)
                           ? reference_reaction(k, p, phase[i], intf, u1, un, phin[i])
                           : generic_reaction(k, p, phase[i], intf, u1);
      Fout[(size_t)k * n + i] = ((c1[i] - cn[i]) / p->dt) + s1 + s2 - R;
    }
  }

  const double exphi = 3.0 + 2.0 * p->s_phi;
  const double ks = p->K_phi * p->sigma1;
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < n; i++) {
    double sum = 0.0;
    int flag = 0;
    for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
      const int32_t nn = col[b];
      if (!(phase[i] == 0 && phase[nn] == 0)) continue;
      const double xx = x[nn] - x[i], yy = y[nn] - y[i], zz = z[nn] - z[i];
      const double xi = sqrt(xx * xx + yy * yy + zz * zz);
      sum += bond_vol(vol, p, -(ks * (phi1[nn] - phi1[i]) / pow(xi, exphi)) *
                                  or_beta(xi, p->delta, p->dx), nn);
      flag = 1;
    }
    double out = 0.0;
    if (flag) {
      double q = 0.0;
      for (int k = 0; k < S; k++) q += p->z[k] * Un[(size_t)k * n + i];
      out = (1. / ks) * sum - (1. / ks) * p->F * p->CSolid * q;
    }
    Fout[(size_t)S * n + i] = out;
  }
  return NNPP_OK;
}

/* main.cpp:676-711: Dirichlet rows of the residual. */
int or_bc_override(int32_t n, const nnpp_params* p, const int32_t* bctypes,
                   const double* bcvalues, const double* U1, double* Fout) {
  const int S = p->n_species;
  for (int32_t i = 0; i < n; i++) {
    if (bctypes[i] != 0) continue;
    for (int f = 0; f <= S; f++) {
      const double v = bcvalues[(size_t)f * n + i];
      if (v == -1) continue;
      Fout[(size_t)f * n + i] = (f < S) ? U1[(size_t)f * n + i] - v / p->CSolid
                                        : U1[(size_t)f * n + i] - v;
    }
  }
  return NNPP_OK;
}

/* Potential bond weight of the K11 block (utils.h:1818-1827 without the K11prop1*sigma1
 * factor that utils.h:1824-1826 divides out again): Beta(xi) V_n / xi^(3+2 s111). */
static inline double phi_weight(const nnpp_params* p, const double* vol, double xi, int32_t nn) {
  return bond_vol(vol, p, (1.0 / pow(xi, 3.0 + 2.0 * p->s_phi)) * or_beta(xi, p->delta, p->dx),
                  nn);
}

static inline double bond_xi(const double* x, const double* y, const double* z, int32_t i,
                             int32_t nn) {
  const double xx = x[nn] - x[i], yy = y[nn] - y[i], zz = z[nn] - z[i];
  return sqrt(xx * xx + yy * yy + zz * zz);
}

/* K11 action without identity rows: out_i = sum_LL w_in (in_i - in_n) on rows with a
 * liquid-liquid bond, 0 elsewhere (utils.h:1794-1848). */
int or_poisson_apply(int32_t n, const nnpp_params* p, const double* x, const double* y,
                     const double* z, const double* vol, const int64_t* row_ptr,
                     const int32_t* col, const int32_t* phase, const double* in, double* out) {
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < n; i++) {
    double acc = 0.0;
    if (phase[i] == 0)
      for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
        const int32_t nn = col[b];
        if (phase[nn] != 0) continue;
        acc += phi_weight(p, vol, bond_xi(x, y, z, i, nn), nn) * (in[i] - in[nn]);
      }
    out[i] = acc;
  }
  return NNPP_OK;
}

static double dot_(int32_t n, const double* a, const double* b) {
  double s = 0.0;
  for (int32_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

/*
 * Potential solve replacing the 11th GMRES solve (main.cpp:764-779) on K11 (utils.h:1794-1848)
 * with residual F11 (utils.h:798-841). Unknowns: free nodes (liquid, at least one
 * liquid-liquid bond, not Dirichlet for phi). Dirichlet columns move to the right-hand side and
 * row i is scaled by V_i / (dx dy dz), which makes the operator symmetric positive
 * (semi-)definite; Jacobi-preconditioned CG from the current phi, stopping at
 * ||r||_2 <= rtol ||b||_2. Dirichlet nodes get phi = bc, other non-free nodes keep phi.
 * *relres is recomputed from the final iterate.
 */
static int poisson_solve_core(int32_t n, const nnpp_params* p, const double* x, const double* y,
                              const double* z, const double* vol, const int64_t* row_ptr,
                              const int32_t* col, const int32_t* phase, const int32_t* bctypes,
                              const double* bcvalues, const double* U, const double* rhs,
                              double* phi, double rtol, int32_t maxit, int32_t* iters,
                              double* relres);

int or_poisson_solve(int32_t n, const nnpp_params* p, const double* x, const double* y,
                     const double* z, const double* vol, const int64_t* row_ptr,
                     const int32_t* col, const int32_t* phase, const int32_t* bctypes,
                     const double* bcvalues, const double* U, double* phi, double rtol,
                     int32_t maxit, int32_t* iters, double* relres) {
  return poisson_solve_core(n, p, x, y, z, vol, row_ptr, col, phase, bctypes, bcvalues, U, NULL,
                            phi, rtol, maxit, iters, relres);
}

/* The same scaled operator M with a given right-hand side on the free rows (the adjoint
 * solve: M is symmetric, so M^T psi = rhs is M psi = rhs). psi holds the warm start on entry
 * and the solution on the free nodes on exit; other entries are left as they are. */
int or_poisson_solve_rhs(int32_t n, const nnpp_params* p, const double* x, const double* y,
                         const double* z, const double* vol, const int64_t* row_ptr,
                         const int32_t* col, const int32_t* phase, const int32_t* bctypes,
                         const double* bcvalues, const double* rhs, double* psi, double rtol,
                         int32_t maxit, int32_t* iters, double* relres) {
  return poisson_solve_core(n, p, x, y, z, vol, row_ptr, col, phase, bctypes, bcvalues, NULL, rhs,
                            psi, rtol, maxit, iters, relres);
}

static int poisson_solve_core(int32_t n, const nnpp_params* p, const double* x, const double* y,
                              const double* z, const double* vol, const int64_t* row_ptr,
                              const int32_t* col, const int32_t* phase, const int32_t* bctypes,
                              const double* bcvalues, const double* U, const double* rhs,
                              double* phi, double rtol, int32_t maxit, int32_t* iters,
                              double* relres) {
  const int S = p->n_species;
  const double V0 = p->dx * p->dy * p->dz;
  const double cb = p->F * p->CSolid / (p->K_phi * p->sigma1);
  const double* bphi = bcvalues ? bcvalues + (size_t)S * n : NULL;
  char* dir = (char*)calloc((size_t)n, 1);
  char* fre = (char*)calloc((size_t)n, 1);
  double* diag = (double*)calloc((size_t)n, sizeof(double));
  double* b = (double*)calloc((size_t)n, sizeof(double));
  double* r = (double*)calloc((size_t)n, sizeof(double));
  double* zz = (double*)calloc((size_t)n, sizeof(double));
  double* pp = (double*)calloc((size_t)n, sizeof(double));
  double* q = (double*)calloc((size_t)n, sizeof(double));
  for (int32_t i = 0; i < n; i++)
    if (bctypes && bphi && bctypes[i] == 0 && bphi[i] != -1) {
      dir[i] = 1;
      if (!rhs) phi[i] = bphi[i];
    }
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < n; i++) {
    if (dir[i] || phase[i] != 0) continue;
    double d = 0.0, bd = 0.0;
    int flag = 0;
    for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
      const int32_t nn = col[k];
      if (phase[nn] != 0) continue;
      const double w = phi_weight(p, vol, bond_xi(x, y, z, i, nn), nn);
      d += w;
      if (dir[nn]) bd += w * phi[nn];
      flag = 1;
    }
    if (!flag) continue;
    const double sc = vol ? vol[i] / V0 : 1.0;
    fre[i] = 1;
    diag[i] = sc * d;
    if (rhs) {
      b[i] = rhs[i];
    } else {
      double qc = 0.0;
      for (int k = 0; k < S; k++) qc += p->z[k] * U[(size_t)k * n + i];
      b[i] = sc * (cb * qc + bd);
    }
  }
  /* r = b - M x */
#define APPLY_M(in, out)                                                         \
  _Pragma("omp parallel for schedule(static)") for (int32_t i = 0; i < n; i++) { \
    double acc = 0.0;                                                            \
    if (fre[i]) {                                                                \
      for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {                    \
        const int32_t nn = col[k];                                               \
        if (phase[nn] != 0) continue;                                            \
        const double w = phi_weight(p, vol, bond_xi(x, y, z, i, nn), nn);        \
        acc += w * (in[i] - (fre[nn] ? in[nn] : 0.0));                           \
      }                                                                          \
      acc *= vol ? vol[i] / V0 : 1.0;                                            \
    }                                                                            \
    out[i] = acc;                                                                \
  }
  double* xv = (double*)calloc((size_t)n, sizeof(double));
  for (int32_t i = 0; i < n; i++) xv[i] = fre[i] ? phi[i] : 0.0;
  const double bnorm = sqrt(dot_(n, b, b));
  int it = 0, status = NNPP_OK;
  if (bnorm == 0.0) {
    for (int32_t i = 0; i < n; i++)
      if (fre[i]) xv[i] = 0.0;
  } else {
    APPLY_M(xv, q);
    for (int32_t i = 0; i < n; i++) {
      r[i] = fre[i] ? b[i] - q[i] : 0.0;
      zz[i] = fre[i] ? r[i] / diag[i] : 0.0;
      pp[i] = zz[i];
    }
    double rho = dot_(n, r, zz);
    const double target = rtol * bnorm;
    while (1) {
      if (sqrt(dot_(n, r, r)) <= target) break;
      if (it >= maxit) {
        status = NNPP_ERR_NOCONV;
        break;
      }
      APPLY_M(pp, q);
      const double alpha = rho / dot_(n, pp, q);
      for (int32_t i = 0; i < n; i++) {
        xv[i] += alpha * pp[i];
        r[i] -= alpha * q[i];
        zz[i] = fre[i] ? r[i] / diag[i] : 0.0;
      }
      const double rho1 = dot_(n, r, zz);
      const double beta = rho1 / rho;
      rho = rho1;
      for (int32_t i = 0; i < n; i++) pp[i] = zz[i] + beta * pp[i];
      it++;
    }
  }
  APPLY_M(xv, q);
  for (int32_t i = 0; i < n; i++) r[i] = fre[i] ? b[i] - q[i] : 0.0;
#undef APPLY_M
  if (relres) *relres = bnorm > 0 ? sqrt(dot_(n, r, r)) / bnorm : 0.0;
  if (iters) *iters = it;
  for (int32_t i = 0; i < n; i++)
    if (fre[i]) phi[i] = xv[i];
  free(xv);
  free(dir);
  free(fre);
  free(diag);
  free(b);
  free(r);
  free(zz);
  free(pp);
  free(q);
  return status;
}

/* main.cpp:400-421 (icvalues IC, phase = U_Mg > EPS), main.cpp:556-640 (Dirichlet values on
 * the first iterate, written here into the state), main.cpp:474-479 (vol0rel). */
int or_init_state(int32_t n, const nnpp_params* p, const int32_t* bctypes,
                  const double* bcvalues, const double* icvalues, double* U, int32_t* phase,
                  int64_t* solid_count, double* vol0rel) {
  const int S = p->n_species;
  memcpy(U, icvalues, sizeof(double) * (size_t)(S + 1) * n);
  int64_t sc = 0;
  double v0 = 0.0; /* main.cpp:476-479: from the initial values, before the Dirichlet values */
  for (int32_t i = 0; i < n; i++) {
    phase[i] = (U[i] > 0.0 + p->phase_eps) ? 1 : 0;
    sc += phase[i];
    if (phase[i] == 1) v0 += (U[i] - p->CSat / p->CSolid);
  }
  if (bctypes && bcvalues)
    for (int32_t i = 0; i < n; i++) {
      if (bctypes[i] != 0) continue;
      for (int f = 0; f <= S; f++) {
        const double v = bcvalues[(size_t)f * n + i];
        if (v != -1) U[(size_t)f * n + i] = (f < S) ? v / p->CSolid : v;
      }
    }
  if (solid_count) *solid_count = sc;
  if (vol0rel) *vol0rel = v0;
  return NNPP_OK;
}

/*
 * One explicit step (the reference's step main.cpp:547-1025 with the Newton/GMRES solve
 * replaced by one explicit update; see include/nnpp.h nnpp_step):
 *   species: U^{n+1} = U^n - dt * (Residuals(U1 := U^n, Un := U^n) - 0), Dirichlet, clamp;
 *   phi^{n+1}: potential solve with the charge of U^n;
 *   front update (main.cpp:915-922), voln1rel every sub_steps steps (main.cpp:944-949).
 */
int or_step(int32_t n, const nnpp_params* p, const double* x, const double* y, const double* z,
            const double* vol, const int64_t* row_ptr, const int32_t* col,
            const int32_t* bctypes, const double* bcvalues, double* U, int32_t* phase,
            int64_t* step, double* voln1rel, double vol0rel, int64_t* solid_count, double rtol,
            int32_t maxit, int32_t* cg_iters, double* relres, int32_t* flips) {
  const int S = p->n_species;
  const size_t nF = (size_t)(S + 1) * n;
  double* G = (double*)malloc(sizeof(double) * nF);
  int st = or_residuals(n, p, x, y, z, vol, row_ptr, col, phase, U, U, *voln1rel, G);
  if (st != NNPP_OK) {
    free(G);
    return st;
  }
  double* phi = (double*)malloc(sizeof(double) * (size_t)n);
  memcpy(phi, U + (size_t)S * n, sizeof(double) * (size_t)n);
  int32_t its = 0;
  double rr = 0.0;
  const int cst = or_poisson_solve(n, p, x, y, z, vol, row_ptr, col, phase, bctypes, bcvalues, U,
                                   phi, rtol, maxit, &its, &rr);
  for (int k = 0; k < S; k++) {
    double* u = U + (size_t)k * n;
    const double* g = G + (size_t)k * n;
    const double* bv = bcvalues ? bcvalues + (size_t)k * n : NULL;
    for (int32_t i = 0; i < n; i++) {
      double v = u[i] - p->dt * g[i];
      if (bctypes && bv && bctypes[i] == 0 && bv[i] != -1) v = bv[i] / p->CSolid;
      if ((p->clamp[k] == NNPP_CLAMP_ALWAYS || (p->clamp[k] == NNPP_CLAMP_LIQUID && phase[i] == 0)) &&
          v < p->clamp_floor)
        v = p->clamp_floor;
      u[i] = v;
    }
  }
  memcpy(U + (size_t)S * n, phi, sizeof(double) * (size_t)n);
  int32_t fl = 0;
  for (int32_t i = 0; i < n; i++)
    if (phase[i] == 1 && U[i] <= p->front_threshold) {
      phase[i] = 0;
      fl++;
    }
  if (solid_count) *solid_count -= fl;
  if (p->sub_steps > 0 && (*step % p->sub_steps) == 0) {
    double v = 0.0;
    for (int32_t i = 0; i < n; i++)
      if (phase[i] == 1) v += (U[i] - p->CSat / p->CSolid);
    *voln1rel = v / vol0rel;
  }
  (*step)++;
  if (cg_iters) *cg_iters = its;
  if (relres) *relres = rr;
  if (flips) *flips = fl;
  free(phi);
  free(G);
  return cst;
}

/*
 * Reverse-mode derivative of nsteps explicit steps (or_step) with respect to the per-species
 * diffusivities D_liquid[k], D_solid[k], the generic dissolution rate r_diss, and the starting
 * state. The reference has no adjoint; this differentiates the oracle's own or_step as the
 * plain reverse of its loops (a scatter over the bonds of every row), so it shares no code
 * path with the device adjoint (a gather over the transposed bonds). It is pinned by central
 * finite differences of or_step.
 *   - the phase trajectory is held fixed (front flips are piecewise constant);
 *   - Dirichlet and clamped entries pass no derivative;
 *   - the potential solve is differentiated at convergence (M symmetric: M psi = dL/dphi);
 *   - BV = 10^(-vol_red (1 - voln1rel)) is differentiated through the voln1rel updates
 *     (every sub_steps steps, voln1rel = sum over solid nodes of (U_0 - CSat/CSolid) / vol0rel).
 * Steps stop early, as in or_step's callers, once no solid node is left.
 * grad[0..S-1] = dL/dD_liquid, grad[S..2S-1] = dL/dD_solid, grad[2S] = dL/dr_diss;
 * lam0 (optional) = dL/dU^0. Generic chemistry only (NNPP_ERR_UNSUPPORTED otherwise).
 */
int or_adjoint(int32_t n, const nnpp_params* p, const double* x, const double* y, const double* z,
               const double* vol, const int64_t* row_ptr, const int32_t* col,
               const int32_t* bctypes, const double* bcvalues, const double* U0,
               const int32_t* phase0, int64_t step0, double voln1rel0, double vol0rel,
               int32_t nsteps, double rtol, int32_t maxit, const double* lamN, double* grad,
               double* lam0, int32_t* steps_done) {
  const int S = p->n_species;
  if (S < 1 || S > NNPP_MAX_SPECIES || nsteps < 0) return NNPP_ERR_ARG;
  if (p->chemistry != NNPP_CHEM_GENERIC) return NNPP_ERR_UNSUPPORTED;
  const size_t nF = (size_t)(S + 1) * n;
  const double V0 = p->dx * p->dy * p->dz;
  const double cb = p->F * p->CSolid / (p->K_phi * p->sigma1);
  double* Ut = (double*)malloc(sizeof(double) * nF * (size_t)(nsteps + 1));
  int32_t* Pt = (int32_t*)malloc(sizeof(int32_t) * (size_t)n * (size_t)(nsteps + 1));
  double* Vt = (double*)malloc(sizeof(double) * (size_t)(nsteps + 1));
  int64_t* St = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nsteps + 1));
  memcpy(Ut, U0, sizeof(double) * nF);
  memcpy(Pt, phase0, sizeof(int32_t) * (size_t)n);
  int64_t step = step0, sc = 0;
  double vr = voln1rel0;
  for (int32_t i = 0; i < n; i++) sc += phase0[i] == 1;
  int32_t done = 0;
  int status = NNPP_OK;
  for (int32_t m = 0; m < nsteps; m++) {
    Vt[m] = vr;
    St[m] = step;
    double* Un = Ut + (size_t)(m + 1) * nF;
    int32_t* Pn = Pt + (size_t)(m + 1) * n;
    memcpy(Un, Ut + (size_t)m * nF, sizeof(double) * nF);
    memcpy(Pn, Pt + (size_t)m * n, sizeof(int32_t) * (size_t)n);
    int32_t its = 0, fl = 0;
    double rr = 0.0;
    const int st = or_step(n, p, x, y, z, vol, row_ptr, col, bctypes, bcvalues, Un, Pn, &step,
                           &vr, vol0rel, &sc, rtol, maxit, &its, &rr, &fl);
    if (st != NNPP_OK && st != NNPP_ERR_NOCONV) {
      status = st;
      break;
    }
    if (st == NNPP_ERR_NOCONV) status = st;
    done = m + 1;
    if (sc <= 0) break;
  }

  double* lam = (double*)malloc(sizeof(double) * nF);
  double* ln = (double*)malloc(sizeof(double) * nF);
  double* G = (double*)malloc(sizeof(double) * nF);
  double* g = (double*)malloc(sizeof(double) * nF);
  double* rhs = (double*)calloc((size_t)n, sizeof(double));
  double* psi = (double*)calloc((size_t)n, sizeof(double));
  char* fre = (char*)calloc((size_t)n, 1);
  char* dir = (char*)calloc((size_t)n, 1);
  memcpy(lam, lamN, sizeof(double) * nF);
  for (int q = 0; q <= 2 * S; q++) grad[q] = 0.0;
  const double* bphi = bcvalues ? bcvalues + (size_t)S * n : NULL;
  for (int32_t i = 0; i < n; i++) dir[i] = bctypes && bphi && bctypes[i] == 0 && bphi[i] != -1;
  double mu = 0.0;  /* dL/d(voln1rel in force after the step being reversed) */
  for (int32_t m = done - 1; m >= 0 && status != NNPP_ERR_ARG; m--) {
    const double* Um = Ut + (size_t)m * nF;
    const int32_t* ph = Pt + (size_t)m * n;
    const int32_t* ph1 = Pt + (size_t)(m + 1) * n;
    /* voln1rel update at the end of step m */
    if (p->sub_steps > 0 && (St[m] % p->sub_steps) == 0) {
      for (int32_t i = 0; i < n; i++)
        if (ph1[i] == 1) lam[i] += mu / vol0rel;
      mu = 0.0;
    }
    const double BV = pow(10., -p->vol_red * (1. - Vt[m]));
    or_residuals(n, p, x, y, z, vol, row_ptr, col, ph, Um, Um, Vt[m], G);
    memset(ln, 0, sizeof(double) * nF);
    /* explicit update, Dirichlet rows and clamp */
    for (int k = 0; k < S; k++)
      for (int32_t i = 0; i < n; i++) {
        const size_t e = (size_t)k * n + i;
        double v = Um[e] - p->dt * G[e];
        int pass = 1;
        if (bctypes && bcvalues && bctypes[i] == 0 && bcvalues[e] != -1) {
          v = bcvalues[e] / p->CSolid;
          pass = 0;
        }
        if ((p->clamp[k] == NNPP_CLAMP_ALWAYS || (p->clamp[k] == NNPP_CLAMP_LIQUID && ph[i] == 0)) &&
            v < p->clamp_floor)
          pass = 0;
        g[e] = pass ? -p->dt * lam[e] : 0.0;
        ln[e] += pass ? lam[e] : 0.0;
      }
    /* potential: phi^{m+1} = M^-1 b(U^m) on free nodes, bc on Dirichlet nodes, phi^m elsewhere */
    for (int32_t i = 0; i < n; i++) {
      fre[i] = 0;
      if (dir[i] || ph[i] != 0) continue;
      for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++)
        if (ph[col[b]] == 0) {
          fre[i] = 1;
          break;
        }
    }
    for (int32_t i = 0; i < n; i++) rhs[i] = fre[i] ? lam[(size_t)S * n + i] : 0.0;
    int32_t its = 0;
    double rr = 0.0;
    const int pst = or_poisson_solve_rhs(n, p, x, y, z, vol, row_ptr, col, ph, bctypes, bcvalues,
                                         rhs, psi, rtol, maxit, &its, &rr);
    if (pst == NNPP_ERR_NOCONV) status = pst;
    for (int32_t i = 0; i < n; i++) {
      if (fre[i]) {
        const double s_ = vol ? vol[i] / V0 : 1.0;
        for (int k = 0; k < S; k++) ln[(size_t)k * n + i] += psi[i] * s_ * cb * p->z[k];
      } else if (!dir[i]) {
        ln[(size_t)S * n + i] += lam[(size_t)S * n + i];
      }
    }
    /* bond terms of G (or_residuals, species rows), reversed bond by bond */
    double gBV = 0.0;
    const double* phm = Um + (size_t)S * n;
    for (int k = 0; k < S; k++) {
      const double* c = Um + (size_t)k * n;
      const double K1 = p->K1[k], K2 = p->K2[k], Ds = p->D_solid[k], Dl = p->D_liquid[k];
      const double ex1 = 3.0 + 2.0 * p->s1[k], ex2 = 3.0 + 2.0 * p->s2[k];
      const double mz = K2 * p->z[k] * p->F / (p->R * p->T);
      const int sb = p->solid_bonds[k];
      for (int32_t i = 0; i < n; i++) {
        const double gi = g[(size_t)k * n + i];
        int intf = 0;
        for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
          const int32_t j = col[b];
          const int pi = ph[i], pj = ph[j];
          if (pi == 1 && pj == 0) intf = 1;
          if (gi == 0.0) continue;
          double D, bv;
          int liquid = 0, bvdep = 0;
          if (pi == 1 && pj == 1) {
            if (!sb) continue;
            D = Ds;
            bv = 1.0;
          } else if (pi != pj) {
            D = Ds;
            bv = sb ? BV : 1.0;
            bvdep = sb;
          } else {
            D = Dl;
            bv = 1.0;
            liquid = 1;
          }
          const double xx = x[j] - x[i], yy = y[j] - y[i], zz = z[j] - z[i];
          const double xi = sqrt(xx * xx + yy * yy + zz * zz);
          const double be = or_beta(xi, p->delta, p->dx);
          const double Vj = vol ? vol[j] : V0;
          const double a0 = K1 * be * Vj / pow(xi, ex1);          /* per unit D bv */
          const double m0 = mz * be * Vj / pow(xi, ex2);
          const double dph = phm[j] - phm[i];
          const double T0 = -a0 * (c[j] - c[i]) - m0 * c[j] * dph;  /* term = D bv T0 */
          const double f = gi * D * bv;
          ln[(size_t)k * n + j] += f * (-a0 - m0 * dph);
          ln[(size_t)k * n + i] += f * a0;
          ln[(size_t)S * n + j] += f * (-m0 * c[j]);
          ln[(size_t)S * n + i] += f * (m0 * c[j]);
          if (liquid) grad[k] += gi * bv * T0;
          else grad[S + k] += gi * bv * T0;
          if (bvdep) gBV += gi * D * T0;
        }
        /* generic reaction: G_0 += r_diss U_0 at interface solid nodes */
        if (k == 0 && intf) {
          ln[i] += gi * p->r_diss;
          grad[2 * S] += gi * c[i];
        }
      }
    }
    mu += gBV * BV * log(10.) * p->vol_red;
    double* t = lam;
    lam = ln;
    ln = t;
  }
  if (lam0) memcpy(lam0, lam, sizeof(double) * nF);
  if (steps_done) *steps_done = done;
  free(Ut);
  free(Pt);
  free(Vt);
  free(St);
  free(lam);
  free(ln);
  free(G);
  free(g);
  free(rhs);
  free(psi);
  free(fre);
  free(dir);
  return status;
}

/* ------------------------------------------------------------------------------------------
 * Implicit (reference) step: utils.h:844-1849 (NewtonMatrices) and main.cpp:547-1025.
 * ------------------------------------------------------------------------------------------ */

/* R'_k of the Jacobian diagonal (reference chemistry: utils.h:1001-1006, 1089-1100, 1184-1190,
 * 1274-1279, 1364-1369, 1454-1459, 1545-1550, 1635-1640; none for Na+ and Cl-). un = U^n at the
 * node. Generic chemistry: the exact derivative of its sink, -r_diss at interface solid nodes. */
static double jac_reaction(int k, const nnpp_params* p, int phase, int intf, const double* un) {
  const int liquid = (phase == 0);
  if (p->chemistry != NNPP_CHEM_REFERENCE)
    return (k == 0 && phase == 1 && intf == 1) ? -p->r_diss : 0.0;
  switch (k) {
    case 1: /* H+: interface or liquid */
      if (intf == 1 || liquid) return p->k7b * (-un[2]) + p->k8b * (-un[4]) + p->k9b * (-un[6]);
      return 0.0;
    case 2: /* OH-: liquid only (interface 0) */
      return (intf != 1 && liquid) ? p->k7b * (-un[1]) : 0.0;
    case 3: /* HCO3- */
    case 4: /* CO3 2- */
      return (!(phase == 1 && intf == 1) && liquid) ? p->k8b * (-un[1]) : 0.0;
    case 6: /* PO4 3- */
      return (!(phase == 1 && intf == 1) && liquid) ? p->k9b * (-un[1]) : 0.0;
    default: /* Mg2+ (0 although R1 depends on U1n1), HPO4 2-, Ca2+, Na+, Cl- */
      return 0.0;
  }
}

static int is_dirichlet(const int32_t* bctypes, const double* bcvalues, int32_t n, int f,
                        int32_t i) {
  return bctypes && bcvalues && bctypes[i] == 0 && bcvalues[(size_t)f * n + i] != -1;
}

/* Block f of the frozen Jacobian applied to in (one n-vector); diag (optional) gets K_ii. */
static void jac_field(int32_t n, const nnpp_params* p, const double* x, const double* y,
                      const double* z, const double* vol, const int64_t* row_ptr,
                      const int32_t* col, const int32_t* phase, const int32_t* bctypes,
                      const double* bcvalues, const double* Un, double voln1rel, int f,
                      const double* in, double* out, double* diag) {
  const int S = p->n_species;
  const double BV = pow(10., -p->vol_red * (1. - voln1rel));
  const double* phin = Un + (size_t)S * n;
  if (f == S) { /* utils.h:1794-1848 */
    const double ks = p->K_phi * p->sigma1, ex = 3.0 + 2.0 * p->s_phi;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; i++) {
      double sum = 0.0, off = 0.0;
      int flag = 0;
      if (!is_dirichlet(bctypes, bcvalues, n, S, i))
        for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
          const int32_t nn = col[b];
          if (!(phase[i] == 0 && phase[nn] == 0)) continue;
          const double xi = bond_xi(x, y, z, i, nn), be = or_beta(xi, p->delta, p->dx);
          sum += bond_vol(vol, p, -(ks / pow(xi, ex)) * be, nn);
          off += bond_vol(vol, p, -(1. / ks) * (ks / pow(xi, ex)) * be, nn) * in[nn];
          flag = 1;
        }
      const double d = flag ? -(1. / ks) * sum : 1.0;
      out[i] = d * in[i] + off;
      if (diag) diag[i] = d;
    }
    return;
  }
  const double K1 = p->K1[f], K2 = p->K2[f], Ds = p->D_solid[f], Dl = p->D_liquid[f];
  const double zk = p->z[f], ex1 = 3.0 + 2.0 * p->s1[f], ex2 = 3.0 + 2.0 * p->s2[f];
  const int sb = p->solid_bonds[f];
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < n; i++) {
    if (is_dirichlet(bctypes, bcvalues, n, f, i)) { /* identity row */
      out[i] = in[i];
      if (diag) diag[i] = 1.0;
      continue;
    }
    int intf = 0;
    if (phase[i] == 1)
      for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++)
        if (phase[col[b]] == 0) {
          intf = 1;
          break;
        }
    double sum = 0.0, off = 0.0;
    for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
      const int32_t nn = col[b];
      const int pi = phase[i], pn = phase[nn];
      double D, bv;
      if (pi == 1 && pn == 1) { /* only species with solid bonds, BV = 1 */
        if (!sb) continue;
        D = Ds;
        bv = 1.0;
      } else if (pi != pn) {
        D = Ds;
        bv = sb ? BV : 1.0;
      } else {
        D = Dl;
        bv = 1.0;
      }
      const double xi = bond_xi(x, y, z, i, nn), be = or_beta(xi, p->delta, p->dx);
      sum += bond_vol(vol, p, -(K1 * D * bv / pow(xi, ex1)) * be, nn);
      off += bond_vol(vol, p,
                      -((K1 * D * bv / pow(xi, ex1)) +
                        (K2 * D * bv * zk * p->F / (p->R * p->T) * (phin[nn] - phin[i]) /
                         pow(xi, ex2))) * be, nn) * in[nn];
    }
    double un[NNPP_MAX_SPECIES];
    for (int q = 0; q < S; q++) un[q] = Un[(size_t)q * n + i];
    const double d = (1. / p->dt) - sum - jac_reaction(f, p, phase[i], intf, un);
    out[i] = d * in[i] + off;
    if (diag) diag[i] = d;
  }
}

/*
 * utils.h:844-1849 as an operator: out_f = K_f in_f for every field f (the reference's 11
 * matrices are block-diagonal in the fields), diag (optional) = the diagonals. The Jacobian is
 * frozen at U^n: its off-diagonal entries use the lagged potential U^n_S, its diagonal
 * 1/dt - sum_K - R'_k the reaction derivatives at U^n. Dirichlet rows are identity rows.
 */
int or_jacobian_apply(int32_t n, const nnpp_params* p, const double* x, const double* y,
                      const double* z, const double* vol, const int64_t* row_ptr,
                      const int32_t* col, const int32_t* phase, const int32_t* bctypes,
                      const double* bcvalues, const double* Un, double voln1rel,
                      const double* in, double* out, double* diag) {
  const int S = p->n_species;
  if (S < 1 || S > NNPP_MAX_SPECIES) return NNPP_ERR_ARG;
  if (p->chemistry == NNPP_CHEM_REFERENCE && S != 10) return NNPP_ERR_UNSUPPORTED;
  for (int f = 0; f <= S; f++)
    jac_field(n, p, x, y, z, vol, row_ptr, col, phase, bctypes, bcvalues, Un, voln1rel, f,
              in + (size_t)f * n, out + (size_t)f * n, diag ? diag + (size_t)f * n : NULL);
  return NNPP_OK;
}

/* Jacobi-preconditioned BiCGSTAB on block f from x = 0 (the reference's GMRES also starts from
 * zero), stopping at ||b - K x||_2 <= rtol ||b||_2 or after maxit iterations. */
static int jac_solve(int32_t n, const nnpp_params* p, const double* x, const double* y,
                     const double* z, const double* vol, const int64_t* row_ptr,
                     const int32_t* col, const int32_t* phase, const int32_t* bctypes,
                     const double* bcvalues, const double* Un, double voln1rel, int f,
                     const double* b, double* xs, double rtol, int32_t maxit, int32_t* iters) {
  double* dg = (double*)malloc(sizeof(double) * (size_t)n);
  double* w = (double*)malloc(sizeof(double) * (size_t)n * 8);
  double *r = w, *rh = w + n, *pp = w + 2 * (size_t)n, *v = w + 3 * (size_t)n,
         *s = w + 4 * (size_t)n, *t = w + 5 * (size_t)n, *ph = w + 6 * (size_t)n,
         *sh = w + 7 * (size_t)n;
  jac_field(n, p, x, y, z, vol, row_ptr, col, phase, bctypes, bcvalues, Un, voln1rel, f, b, v,
            dg);
  for (int32_t i = 0; i < n; i++) {
    xs[i] = 0.0;
    r[i] = rh[i] = b[i];
    pp[i] = v[i] = 0.0;
  }
  const double bn = sqrt(dot_(n, b, b));
  double rho = 1.0, alpha = 1.0, omega = 1.0;
  int it = 0, status = NNPP_OK;
  if (bn > 0.0)
    while (1) {
      if (sqrt(dot_(n, r, r)) <= rtol * bn) break;
      if (it >= maxit) {
        status = NNPP_ERR_NOCONV;
        break;
      }
      const double rho1 = dot_(n, rh, r);
      if (rho1 == 0.0) break;
      const double beta = (rho1 / rho) * (alpha / omega);
      rho = rho1;
      for (int32_t i = 0; i < n; i++) {
        pp[i] = r[i] + beta * (pp[i] - omega * v[i]);
        ph[i] = pp[i] / dg[i];
      }
      jac_field(n, p, x, y, z, vol, row_ptr, col, phase, bctypes, bcvalues, Un, voln1rel, f, ph,
                v, NULL);
      const double rv = dot_(n, rh, v);
      if (rv == 0.0) break;
      alpha = rho / rv;
      for (int32_t i = 0; i < n; i++) {
        s[i] = r[i] - alpha * v[i];
        sh[i] = s[i] / dg[i];
      }
      it++;
      if (sqrt(dot_(n, s, s)) <= rtol * bn) {
        for (int32_t i = 0; i < n; i++) {
          xs[i] += alpha * ph[i];
          r[i] = s[i];
        }
        break;
      }
      jac_field(n, p, x, y, z, vol, row_ptr, col, phase, bctypes, bcvalues, Un, voln1rel, f, sh,
                t, NULL);
      const double tt = dot_(n, t, t);
      omega = tt > 0.0 ? dot_(n, t, s) / tt : 0.0;
      for (int32_t i = 0; i < n; i++) {
        xs[i] += alpha * ph[i] + omega * sh[i];
        r[i] = s[i] - omega * t[i];
      }
      if (omega == 0.0) break;
    }
  if (iters) *iters += it;
  free(w);
  free(dg);
  return status;
}

/*
 * One implicit step of main.cpp:547-1025 (modified Newton with the Jacobian of utils.h:844-1849
 * frozen at U^n):
 *   iterate: first != 0 -> U^n with the Dirichlet values (main.cpp:556-640), else Uit as left
 *            by the previous step's rotation;
 *   per Newton iteration (at most newton_maxit, main.cpp:665): F = Residuals(Uit, U^n)
 *            (utils.h:27-842) with the Dirichlet rows F = U - bc (main.cpp:676-711); per field
 *            K_f dU_f = -F_f (jac_solve); Uit += dU; res_f = ||dU_f||_2; clamp
 *            (main.cpp:805-837); stop when max_f |res_f - res_f,prev| < newton_tol
 *            (main.cpp:863-911, res_prev = 0 at the first iteration);
 *   front update on the iterate (main.cpp:915-922), voln1rel every sub_steps steps
 *            (main.cpp:944-949);
 *   rotation (main.cpp:1002-1024): Uprev <- U^n, U <- U^{n+1}, Uit <- old Uprev.
 * newton_iters: Newton iterations of the step; lin_iters: BiCGSTAB iterations summed over the
 * fields; converged: 1 when the Newton test passed.
 */
int or_step_implicit(int32_t n, const nnpp_params* p, const double* x, const double* y,
                     const double* z, const double* vol, const int64_t* row_ptr,
                     const int32_t* col, const int32_t* bctypes, const double* bcvalues,
                     double* U, double* Uprev, double* Uit, int32_t first, int32_t* phase,
                     int64_t* step, double* voln1rel, double vol0rel, int64_t* solid_count,
                     double newton_tol, int32_t newton_maxit, double lin_rtol, int32_t lin_maxit,
                     int32_t* newton_iters, int32_t* lin_iters, int32_t* converged,
                     double* resmax_out, int32_t* flips) {
  const int S = p->n_species;
  if (S < 1 || S > NNPP_MAX_SPECIES) return NNPP_ERR_ARG;
  if (p->chemistry == NNPP_CHEM_REFERENCE && S != 10) return NNPP_ERR_UNSUPPORTED;
  const size_t nF = (size_t)(S + 1) * n;
  if (first) {
    memcpy(Uit, U, sizeof(double) * nF);
    if (bctypes && bcvalues)
      for (int32_t i = 0; i < n; i++) {
        if (bctypes[i] != 0) continue;
        for (int f = 0; f <= S; f++) {
          const double v = bcvalues[(size_t)f * n + i];
          if (v != -1) Uit[(size_t)f * n + i] = (f < S) ? v / p->CSolid : v;
        }
      }
  }
  double* Fv = (double*)malloc(sizeof(double) * nF);
  double* bvec = (double*)malloc(sizeof(double) * (size_t)n);
  double* dU = (double*)malloc(sizeof(double) * (size_t)n);
  double res_old[NNPP_MAX_SPECIES + 1], res[NNPP_MAX_SPECIES + 1];
  for (int f = 0; f <= S; f++) res_old[f] = 0.0;
  int32_t nit = 0, lit = 0, conv = 0;
  double resmax = 0.0;
  int status = NNPP_OK;
  for (int it = 0; it < newton_maxit; it++) {
    int st = or_residuals(n, p, x, y, z, vol, row_ptr, col, phase, Uit, U, *voln1rel, Fv);
    if (st) {
      status = st;
      break;
    }
    if (bctypes && bcvalues) or_bc_override(n, p, bctypes, bcvalues, Uit, Fv);
    for (int f = 0; f <= S; f++) {
      for (int32_t i = 0; i < n; i++) bvec[i] = -Fv[(size_t)f * n + i];
      const int ls = jac_solve(n, p, x, y, z, vol, row_ptr, col, phase, bctypes, bcvalues, U,
                               *voln1rel, f, bvec, dU, lin_rtol, lin_maxit, &lit);
      if (ls == NNPP_ERR_NOCONV) status = NNPP_ERR_NOCONV;
      double* u = Uit + (size_t)f * n;
      for (int32_t i = 0; i < n; i++) u[i] += dU[i];
      res[f] = sqrt(dot_(n, dU, dU));
    }
    for (int k = 0; k < S; k++) {
      double* u = Uit + (size_t)k * n;
      for (int32_t i = 0; i < n; i++)
        if ((p->clamp[k] == NNPP_CLAMP_ALWAYS ||
             (p->clamp[k] == NNPP_CLAMP_LIQUID && phase[i] == 0)) &&
            u[i] < p->clamp_floor)
          u[i] = p->clamp_floor;
    }
    nit++;
    resmax = 0.0;
    for (int f = 0; f <= S; f++) {
      const double d = fabs(res[f] - res_old[f]);
      if (d > resmax) resmax = d;
    }
    if (resmax < newton_tol) {
      conv = 1;
      break;
    }
    for (int f = 0; f <= S; f++) res_old[f] = res[f];
  }
  int32_t fl = 0;
  for (int32_t i = 0; i < n; i++)
    if (phase[i] == 1 && Uit[i] <= p->front_threshold) {
      phase[i] = 0;
      fl++;
    }
  if (solid_count) *solid_count -= fl;
  if (p->sub_steps > 0 && (*step % p->sub_steps) == 0) {
    double v = 0.0;
    for (int32_t i = 0; i < n; i++)
      if (phase[i] == 1) v += (Uit[i] - p->CSat / p->CSolid);
    *voln1rel = v / vol0rel;
  }
  (*step)++;
  /* rotation: (Uprev, U, Uit) <- (U, Uit, Uprev) */
  for (size_t q = 0; q < nF; q++) {
    const double a = Uprev[q];
    Uprev[q] = U[q];
    U[q] = Uit[q];
    Uit[q] = a;
  }
  if (newton_iters) *newton_iters = nit;
  if (lin_iters) *lin_iters = lit;
  if (converged) *converged = conv;
  if (resmax_out) *resmax_out = resmax;
  if (flips) *flips = fl;
  free(Fv);
  free(bvec);
  free(dU);
  return status;
}

/* ------------------------------------------------------------------------------------------
 * Implant classification: build_clouds (kd_tree_interface.h:156-252) as main.cpp:316-320 calls
 * it (min_node_no = 0, dr = 0) and the initial values it selects (main.cpp:422-457).
 * ------------------------------------------------------------------------------------------ */

/* counts[i] = number of implant points p with sqrt(x*x + y*y + z*z) <= r, (x, y, z) = p - node
 * (Node::distance_euclidean, kd_tree_interface.h:44-50), by brute force over all pairs.
 * NNPP_ERR_OVERFLOW if a count exceeds max_count (the reference throws,
 * kd_tree_interface.h:214-217). */
int or_implant_classify(int32_t n, const double* x, const double* y, const double* z,
                        int64_t nb, const double* bx, const double* by, const double* bz,
                        double r, int32_t max_count, int32_t* counts) {
  int status = NNPP_OK;
#pragma omp parallel for schedule(dynamic, 64)
  for (int32_t i = 0; i < n; i++) {
    int32_t c = 0;
    for (int64_t k = 0; k < nb; k++) {
      const double dx = bx[k] - x[i], dy = by[k] - y[i], dz = bz[k] - z[i];
      if (sqrt(dx * dx + dy * dy + dz * dz) <= r) c++;
    }
    counts[i] = c;
    if (c > max_count) {
#pragma omp atomic write
      status = NNPP_ERR_OVERFLOW;
    }
  }
  return status;
}

/* main.cpp:422-457: no implant point within r -> liquid (Mg2+ 0, H+ and OH- 1e-4/CSolid,
 * HCO3- C0HCO3/CSolid, CO3 2- 0, HPO4 2- C0HPO4/CSolid, PO4 3- 0, Ca2+ C0Ca/CSolid,
 * Na+ C0Na/CSolid, Cl- C0Cl/CSolid, phi 0); otherwise solid (Mg2+ CSolid/CSolid, the rest 0).
 * The ten reference species only. ic: 11*n, field-major. */
int or_implant_icvalues(int32_t n, const nnpp_params* p, const int32_t* counts, double* ic) {
  if (p->n_species != 10) return NNPP_ERR_UNSUPPORTED;
  const double CS = p->CSolid;
  for (int32_t i = 0; i < n; i++) {
    double v[11] = {0.0};
    if (counts[i] == 0) {
      v[1] = 1e-4 / CS;
      v[2] = 1e-4 / CS;
      v[3] = p->C0HCO3 / CS;
      v[5] = p->C0HPO4 / CS;
      v[7] = p->C0Ca / CS;
      v[8] = p->C0Na / CS;
      v[9] = p->C0Cl / CS;
    } else {
      v[0] = CS / CS;
    }
    for (int f = 0; f < 11; f++) ic[(size_t)f * n + i] = v[f];
  }
  return NNPP_OK;
}

/* Voronoi-approximate node volumes (no reference counterpart): probes on a lattice of spacing h
 * over the nodes' bounding box widened by margin (round(extent / h) per axis, centres at
 * lo + (k + 1/2) h), each assigned to its nearest node within radius by brute force (squared
 * distances, ties to the lowest index); vol[i] = probes of node i * h^3. */
int or_voronoi_volumes(int32_t n, const double* x, const double* y, const double* z, double h,
                       double margin, double radius, double* vol) {
  if (n <= 0) return NNPP_OK;
  double lo[3] = {x[0], y[0], z[0]}, hi[3] = {x[0], y[0], z[0]};
  for (int32_t i = 0; i < n; i++) {
    const double v[3] = {x[i], y[i], z[i]};
    for (int a = 0; a < 3; a++) {
      lo[a] = v[a] < lo[a] ? v[a] : lo[a];
      hi[a] = v[a] > hi[a] ? v[a] : hi[a];
    }
  }
  int64_t dims[3];
  for (int a = 0; a < 3; a++) {
    const double l = lo[a] - margin;
    dims[a] = llround(((hi[a] + margin) - l) / h);
    if (dims[a] < 1) dims[a] = 1;
    lo[a] = l;
  }
  const int64_t np = dims[0] * dims[1] * dims[2];
  int64_t* cnt = (int64_t*)calloc((size_t)n, sizeof(int64_t));
  const double r2 = radius * radius;
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < np; p++) {
    const int64_t kx = p % dims[0], ky = (p / dims[0]) % dims[1], kz = p / (dims[0] * dims[1]);
    const double qx = lo[0] + ((double)kx + 0.5) * h, qy = lo[1] + ((double)ky + 0.5) * h,
                 qz = lo[2] + ((double)kz + 0.5) * h;
    double best = r2;
    int32_t who = -1;
    for (int32_t j = 0; j < n; j++) { /* ascending j: the first of equal distances wins */
      const double dx = x[j] - qx, dy = y[j] - qy, dz = z[j] - qz;
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 < best || (d2 == best && who < 0)) {
        best = d2;
        who = j;
      }
    }
    if (who >= 0) {
#pragma omp atomic
      cnt[who]++;
    }
  }
  for (int32_t i = 0; i < n; i++) vol[i] = (double)cnt[i] * (h * h * h);
  free(cnt);
  return NNPP_OK;
}

/*
 * nnpp_oracle.h — CPU FP64 restatement of the reference's NNPP step path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load this library, and only as the checker (or the timed CPU baseline). The product
 * path (libnnpp.so) never links, loads or calls it.
 *
 * Arrays follow include/nnpp.h: per-field arrays packed field-major (f * n + i), F = S + 1
 * fields with field S the potential phi. Neighbour lists are CSR (int64 row_ptr[n+1], int32
 * col[nnz]), rows sorted by id, self excluded.
 */
#ifndef NNPP_ORACLE_H
#define NNPP_ORACLE_H

#include <stdint.h>
#include "../include/nnpp.h"

#ifdef __cplusplus
extern "C" {
#endif

double or_beta(double xi, double delta, double dx);
int or_sphere_cube_intersects(double cx, double cy, double cz, double cr, double rx, double ry,
                              double rz, double rw, double rh, double rd);

int or_build_neighbors(int32_t n, const double* x, const double* y, const double* z, double r,
                       double cdx, double cdy, double cdz, int32_t max_len, int64_t* row_ptr,
                       int32_t* col, int32_t* max_len_out);

int or_residuals(int32_t n, const nnpp_params* p, const double* x, const double* y,
                 const double* z, const double* vol, const int64_t* row_ptr, const int32_t* col,
                 const int32_t* phase, const double* U1, const double* Un, double voln1rel,
                 double* Fout);

int or_bc_override(int32_t n, const nnpp_params* p, const int32_t* bctypes,
                   const double* bcvalues, const double* U1, double* Fout);

int or_poisson_apply(int32_t n, const nnpp_params* p, const double* x, const double* y,
                     const double* z, const double* vol, const int64_t* row_ptr,
                     const int32_t* col, const int32_t* phase, const double* in, double* out);

int or_poisson_solve(int32_t n, const nnpp_params* p, const double* x, const double* y,
                     const double* z, const double* vol, const int64_t* row_ptr,
                     const int32_t* col, const int32_t* phase, const int32_t* bctypes,
                     const double* bcvalues, const double* U, double* phi, double rtol,
                     int32_t maxit, int32_t* iters, double* relres);

int or_poisson_solve_rhs(int32_t n, const nnpp_params* p, const double* x, const double* y,
                         const double* z, const double* vol, const int64_t* row_ptr,
                         const int32_t* col, const int32_t* phase, const int32_t* bctypes,
                         const double* bcvalues, const double* rhs, double* psi, double rtol,
                         int32_t maxit, int32_t* iters, double* relres);
int or_init_state(int32_t n, const nnpp_params* p, const int32_t* bctypes,
                  const double* bcvalues, const double* icvalues, double* U, int32_t* phase,
                  int64_t* solid_count, double* vol0rel);

int or_step(int32_t n, const nnpp_params* p, const double* x, const double* y, const double* z,
            const double* vol, const int64_t* row_ptr, const int32_t* col,
            const int32_t* bctypes, const double* bcvalues, double* U, int32_t* phase,
            int64_t* step, double* voln1rel, double vol0rel, int64_t* solid_count, double rtol,
            int32_t maxit, int32_t* cg_iters, double* relres, int32_t* flips);

int or_adjoint(int32_t n, const nnpp_params* p, const double* x, const double* y, const double* z,
               const double* vol, const int64_t* row_ptr, const int32_t* col,
               const int32_t* bctypes, const double* bcvalues, const double* U0,
               const int32_t* phase0, int64_t step0, double voln1rel0, double vol0rel,
               int32_t nsteps, double rtol, int32_t maxit, const double* lamN, double* grad,
               double* lam0, int32_t* steps_done);

int or_jacobian_apply(int32_t n, const nnpp_params* p, const double* x, const double* y,
                      const double* z, const double* vol, const int64_t* row_ptr,
                      const int32_t* col, const int32_t* phase, const int32_t* bctypes,
                      const double* bcvalues, const double* Un, double voln1rel,
                      const double* in, double* out, double* diag);

int or_step_implicit(int32_t n, const nnpp_params* p, const double* x, const double* y,
                     const double* z, const double* vol, const int64_t* row_ptr,
                     const int32_t* col, const int32_t* bctypes, const double* bcvalues,
                     double* U, double* Uprev, double* Uit, int32_t first, int32_t* phase,
                     int64_t* step, double* voln1rel, double vol0rel, int64_t* solid_count,
                     double newton_tol, int32_t newton_maxit, double lin_rtol, int32_t lin_maxit,
                     int32_t* newton_iters, int32_t* lin_iters, int32_t* converged,
                     double* resmax_out, int32_t* flips);

int or_implant_classify(int32_t n, const double* x, const double* y, const double* z,
                        int64_t nb, const double* bx, const double* by, const double* bz,
                        double r, int32_t max_count, int32_t* counts);
int or_implant_icvalues(int32_t n, const nnpp_params* p, const int32_t* counts, double* ic);
int or_voronoi_volumes(int32_t n, const double* x, const double* y, const double* z, double h,
                       double margin, double radius, double* vol);

#ifdef __cplusplus
}
#endif

#endif

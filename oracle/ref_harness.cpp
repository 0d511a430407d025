// ref_harness.cpp — extern "C" entry points around the reference's own headers, so the tests
// can run the reference's build_clouds_hash (src/helpers/grid.h:43-204), utils::Residuals
// (src/helpers/utils.h:27-842), utils::NewtonMatrices (utils.h:844-1849) and the implant
// classification build_clouds (src/helpers/kd_tree_interface.h:156-252) on small inputs.
//
// TEST INFRASTRUCTURE ONLY. Compiled by oracle/Makefile (target `ref`) from the reference
// sources where they lie (-I <reference>/src/helpers -I <reference>/eigen); the output goes to
// oracle/_ref/ and is never part of the product. This file holds only adapter code: flat C
// arrays in, Eigen containers for the reference templates, flat C arrays out.
#include <Eigen/Dense>
#include <Eigen/Sparse>

#include <omp.h>

#include <cstdio>
#include <stdexcept>
#include <vector>

#include "grid.h"
#include "kd_tree_interface.h"
#include "utils.h"

typedef Eigen::SparseMatrix<double> SpMat;

namespace {

// Silence the progress lines grid.h prints on std::cout.
struct CoutSilencer {
  std::streambuf* old;
  CoutSilencer() : old(std::cout.rdbuf(nullptr)) {}
  ~CoutSilencer() { std::cout.rdbuf(old); }
};

void make_params(const double* data1, const double* data2, const double* datac,
                 const double* datac11, utils::parameter& p) {
  Eigen::VectorXd c1(4), c2(51), cc[10], c11(4);
  for (int i = 0; i < 4; i++) c1[i] = data1[i];
  for (int i = 0; i < 51; i++) c2[i] = data2[i];
  for (int k = 0; k < 10; k++) {
    cc[k].resize(7);
    for (int i = 0; i < 7; i++) cc[k][i] = datac[k * 7 + i];
  }
  for (int i = 0; i < 4; i++) c11[i] = datac11[i];
  p.set_int(c1);
  p.set_double(c2, cc[0], cc[1], cc[2], cc[3], cc[4], cc[5], cc[6], cc[7], cc[8], cc[9], c11);
}

void make_nodes(int n, const double* x, const double* y, const double* z, Eigen::MatrixXd& nodes) {
  nodes.resize(n, 3);
  for (int i = 0; i < n; i++) {
    nodes(i, 0) = x[i];
    nodes(i, 1) = y[i];
    nodes(i, 2) = z[i];
  }
}

// ELL (row-major n x max_len as handed over by the caller) -> Eigen col-major matrix.
void make_horizons(int n, int max_len, const int* ell, const int* lengths, Eigen::MatrixXi& h,
                   Eigen::VectorXi& hl) {
  h.resize(n, max_len);
  hl.resize(n);
  for (int i = 0; i < n; i++) {
    hl[i] = lengths[i];
    for (int j = 0; j < max_len; j++) h(i, j) = ell[(size_t)i * max_len + j];
  }
}

void vecs(int n, const double* a, Eigen::VectorXd* v) {
  for (int f = 0; f < 11; f++) {
    v[f].resize(n);
    for (int i = 0; i < n; i++) v[f][i] = a[(size_t)f * n + i];
  }
}

}  // namespace

extern "C" {

// build_clouds_hash(max_len, r, cdx, cdy, cdz, nodes, clouds, lengths, max) -> ell is written
// row-major (n x max_len) in the reference's emission order. Returns 0, or 2 on the
// reference's overflow exception.
int ref_build_clouds_hash(int n, const double* x, const double* y, const double* z, double r,
                          double cdx, double cdy, double cdz, int max_len, int* ell,
                          int* lengths, int* max_len_out) {
  Eigen::MatrixXd nodes;
  make_nodes(n, x, y, z, nodes);
  Eigen::MatrixXi clouds;
  Eigen::VectorXi cl;
  int mx = 0;
  try {
    CoutSilencer quiet;
    build_clouds_hash(max_len, r, cdx, cdy, cdz, nodes, clouds, cl, mx);
  } catch (const std::runtime_error&) {
    return 2;
  }
  for (int i = 0; i < n; i++) {
    lengths[i] = cl[i];
    for (int j = 0; j < max_len; j++) ell[(size_t)i * max_len + j] = clouds(i, j);
  }
  *max_len_out = mx;
  return 0;
}

// utils::Residuals for the reference's 10 species + phi. Field arrays are 11*n, field-major.
int ref_residuals(int n, const double* data1, const double* data2, const double* datac,
                  const double* datac11, const double* x, const double* y, const double* z,
                  int max_len, const int* ell, const int* lengths, const double* bcv,
                  const double* U1, const double* Un, const double* Un0, const int* phase,
                  const int* bctypes, double voln1rel, double* Fout) {
  utils::parameter p;
  make_params(data1, data2, datac, datac11, p);
  p.nodesNo = n;
  Eigen::MatrixXd nodes;
  make_nodes(n, x, y, z, nodes);
  Eigen::MatrixXi h;
  Eigen::VectorXi hl;
  make_horizons(n, max_len, ell, lengths, h, hl);
  Eigen::VectorXd B[11], A[11], N[11], O[11], F[11];
  vecs(n, bcv, B);
  vecs(n, U1, A);
  vecs(n, Un, N);
  vecs(n, Un0, O);
  for (int f = 0; f < 11; f++) F[f].resize(n);
  Eigen::VectorXi ph(n), bt(n);
  for (int i = 0; i < n; i++) {
    ph[i] = phase[i];
    bt[i] = bctypes[i];
  }
  utils::Residuals(nodes, h, B[0], B[1], B[2], B[3], B[4], B[5], B[6], B[7], B[8], B[9], B[10],
                   A[0], A[1], A[2], A[3], A[4], A[5], A[6], A[7], A[8], A[9], A[10], N[0], N[1],
                   N[2], N[3], N[4], N[5], N[6], N[7], N[8], N[9], N[10], O[0], O[1], O[2], O[3],
                   O[4], O[5], O[6], O[7], O[8], O[9], O[10], F[0], F[1], F[2], F[3], F[4], F[5],
                   F[6], F[7], F[8], F[9], F[10], ph, hl, bt, voln1rel, p);
  for (int f = 0; f < 11; f++)
    for (int i = 0; i < n; i++) Fout[(size_t)f * n + i] = F[f][i];
  return 0;
}

// utils::Residuals as above, called `reps` times on the same inputs; returns the wall seconds
// of those calls alone (the conversions to Eigen containers are not timed). For the CPU
// baseline of the benchmark: the reference's own bond loops on the host cores.
double ref_residuals_timed(int n, const double* data1, const double* data2, const double* datac,
                           const double* datac11, const double* x, const double* y,
                           const double* z, int max_len, const int* ell, const int* lengths,
                           const double* bcv, const double* U1, const double* Un,
                           const int* phase, const int* bctypes, double voln1rel, int reps,
                           double* Fout) {
  utils::parameter p;
  make_params(data1, data2, datac, datac11, p);
  p.nodesNo = n;
  Eigen::MatrixXd nodes;
  make_nodes(n, x, y, z, nodes);
  Eigen::MatrixXi h;
  Eigen::VectorXi hl;
  make_horizons(n, max_len, ell, lengths, h, hl);
  Eigen::VectorXd B[11], A[11], N[11], O[11], F[11];
  vecs(n, bcv, B);
  vecs(n, U1, A);
  vecs(n, Un, N);
  vecs(n, Un, O);
  for (int f = 0; f < 11; f++) F[f].resize(n);
  Eigen::VectorXi ph(n), bt(n);
  for (int i = 0; i < n; i++) {
    ph[i] = phase[i];
    bt[i] = bctypes[i];
  }
  const double t0 = omp_get_wtime();
  for (int r = 0; r < reps; r++)
    utils::Residuals(nodes, h, B[0], B[1], B[2], B[3], B[4], B[5], B[6], B[7], B[8], B[9],
                     B[10], A[0], A[1], A[2], A[3], A[4], A[5], A[6], A[7], A[8], A[9], A[10],
                     N[0], N[1], N[2], N[3], N[4], N[5], N[6], N[7], N[8], N[9], N[10], O[0],
                     O[1], O[2], O[3], O[4], O[5], O[6], O[7], O[8], O[9], O[10], F[0], F[1],
                     F[2], F[3], F[4], F[5], F[6], F[7], F[8], F[9], F[10], ph, hl, bt, voln1rel,
                     p);
  const double t = omp_get_wtime() - t0;
  for (int f = 0; f < 11; f++)
    for (int i = 0; i < n; i++) Fout[(size_t)f * n + i] = F[f][i];
  return t;
}

// utils::NewtonMatrices; returns matrix K<which> (1..11) as coordinate triplets in Eigen's
// column-major storage order. Returns 1 if cap is too small (*nnz_out holds the need).
int ref_newton_matrix(int n, const double* data1, const double* data2, const double* datac,
                      const double* datac11, const double* x, const double* y, const double* z,
                      int max_len, const int* ell, const int* lengths, const double* bcv,
                      const double* U1, const double* Un, const int* phase, const int* bctypes,
                      double voln1rel, int which, int cap, int* rows, int* cols, double* vals,
                      int* nnz_out) {
  utils::parameter p;
  make_params(data1, data2, datac, datac11, p);
  p.nodesNo = n;
  Eigen::MatrixXd nodes;
  make_nodes(n, x, y, z, nodes);
  Eigen::MatrixXi h;
  Eigen::VectorXi hl;
  make_horizons(n, max_len, ell, lengths, h, hl);
  Eigen::VectorXd B[11], A[11], N[11];
  vecs(n, bcv, B);
  vecs(n, U1, A);
  vecs(n, Un, N);
  Eigen::VectorXi ph(n), bt(n);
  for (int i = 0; i < n; i++) {
    ph[i] = phase[i];
    bt[i] = bctypes[i];
  }
  std::vector<SpMat> K(11, SpMat(n, n));
  utils::NewtonMatrices(K[0], K[1], K[2], K[3], K[4], K[5], K[6], K[7], K[8], K[9], K[10], nodes,
                        h, B[0], B[1], B[2], B[3], B[4], B[5], B[6], B[7], B[8], B[9], B[10], A[0],
                        A[1], A[2], A[3], A[4], A[5], A[6], A[7], A[8], A[9], A[10], N[0], N[1],
                        N[2], N[3], N[4], N[5], N[6], N[7], N[8], N[9], N[10], ph, hl, bt,
                        voln1rel, p);
  const SpMat& M = K[which - 1];
  *nnz_out = (int)M.nonZeros();
  if (M.nonZeros() > cap) return 1;
  int c = 0;
  for (int k = 0; k < M.outerSize(); ++k)
    for (SpMat::InnerIterator it(M, k); it; ++it) {
      rows[c] = (int)it.row();
      cols[c] = (int)it.col();
      vals[c] = it.value();
      c++;
    }
  return 0;
}

// build_clouds(min_node_no, max_node_no, r, dr, nodes, nodes2, ...) (kd_tree_interface.h:156-252)
// with nodes2 = the implant points, as main.cpp:316-320 calls it with min_node_no = 0, dr = 0:
// lengths[i] = number of implant points within r of node i. Returns 0, or 2 on the reference's
// overflow exception.
int ref_build_clouds_implant(int n, const double* x, const double* y, const double* z, int nb,
                             const double* bx, const double* by, const double* bz, double r,
                             int max_len, int* lengths) {
  Eigen::MatrixXd nodes, bunny;
  make_nodes(n, x, y, z, nodes);
  make_nodes(nb, bx, by, bz, bunny);
  Eigen::MatrixXi clouds;
  Eigen::VectorXi cl;
  int mx = 0;
  double mean = 0.0;
  try {
    build_clouds(0, max_len, r, 0.0, nodes, bunny, clouds, cl, mx, mean);
  } catch (const std::runtime_error&) {
    return 2;
  }
  for (int i = 0; i < n; i++) lengths[i] = cl[i];
  return 0;
}

// utils::Beta
double ref_beta(double xi, double delta, double dx) { return utils::Beta(xi, delta, dx); }

}  // extern "C"

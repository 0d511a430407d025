/*
 * nnpp.h — C ABI of the MI355X-native NNPP (nonlocal Nernst–Planck–Poisson) time step.
 *
 * Drop-in boundary for the step path of alhermann/nnpp-mg-implant-corrosion-3d. Every entry
 * point below names the reference interface it replaces (paths relative to the reference
 * repository). The reference's interface is a set of header-only C++ templates over Eigen
 * containers called from main(); here it becomes plain C: host pointers + sizes, integer
 * status codes (0 = NNPP_OK), no exceptions and no C++/torch types across the boundary.
 *
 * Ownership: host arrays belong to the caller and are only accessed during the call.
 * Device buffers belong to the context. Threading: one context per host thread; all work is
 * issued on the context's own HIP stream and every call blocks until it is complete.
 *
 * Field numbering: a problem has S species (the reference hard-codes S = 10: Mg2+, H+, OH-,
 * HCO3-, CO3 2-, HPO4 2-, PO4 3-, Ca2+, Na+, Cl-; main.cpp:526-529) plus the potential phi,
 * i.e. F = S + 1 fields; field f = 0..S-1 is species f+1 of the reference's 1-based naming,
 * and field S is phi (the reference's U11). Per-field host arrays are packed field-major:
 * array[f * n_nodes + i].
 */
#ifndef NNPP_H
#define NNPP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NNPP_MAX_SPECIES 16
#define NNPP_DATA2_LEN 51   /* utils.h:2348-2398 reads 51 values from <job>.data2 */
#define NNPP_DATAC_LEN 7    /* utils.h:2400-2478: z, DSolid, DLiquid, s1, s2, Kprop1, Kprop2 */
#define NNPP_DATAC_PHI_LEN 4 /* utils.h:2480-2483: sigma0, sigma1, s111, K11prop1 */

/* Status codes. */
enum {
  NNPP_OK = 0,
  NNPP_ERR_ARG = 1,          /* bad argument / size */
  NNPP_ERR_OVERFLOW = 2,     /* a neighbourhood does not fit max_len (grid.h:187-189 throws) */
  NNPP_ERR_NONFINITE = 3,    /* NaN/Inf detected in the state after a step */
  NNPP_ERR_NOCONV = 4,       /* potential solve did not reach its tolerance within maxit */
  NNPP_ERR_HIP = 5,          /* HIP runtime error (message in nnpp_last_error) */
  NNPP_ERR_STATE = 6,        /* call order violated (e.g. step before neighbours exist) */
  NNPP_ERR_UNSUPPORTED = 7,  /* parameter combination this build does not implement */
  NNPP_ERR_IO = 8            /* input file missing or short (readers only) */
};

/* Reaction model. */
enum {
  NNPP_CHEM_GENERIC = 0,     /* no bulk reactions; optional first-order Mg sink r_diss at
                                interface solid nodes (no reference counterpart) */
  NNPP_CHEM_REFERENCE = 1    /* the reference's R1..R8 terms, utils.h:162-174,241-253,...;
                                requires n_species == 10 */
};

/* Arithmetic of the step. */
enum {
  NNPP_FP64 = 0,             /* every operation in FP64 */
  NNPP_MIXED = 1             /* FP32 CG iterations (FP32 operator and vectors) with FP64
                                dots, residual replacement and solution; FP32 flux bond sums
                                (column form: FP32 differences from the FP64 state, Kahan sums
                                over groups of 16 bonds; lattice stencil form: FP32 sums of
                                differences from each tile's first row, per 16 bonds, added in
                                FP64); FP64 reactions and state update */
};

/* Lower clamp applied to a species after each update (main.cpp:805-837). */
enum { NNPP_CLAMP_NONE = 0, NNPP_CLAMP_ALWAYS = 1, NNPP_CLAMP_LIQUID = 2 };

/*
 * Problem parameters; mirrors utils::parameter (utils.h:2186-2485). Fields in the .data2 block
 * keep the reference's order and names (pKH2O is the reference's pKH20). Species arrays hold
 * .datac1 .. .datac<S>; sigma0 .. K_phi hold .datac<S+1>. Fields after "extensions" have no
 * counterpart in the reference files; nnpp_params_defaults() sets them to the reference's
 * hard-coded behaviour.
 */
typedef struct nnpp_params {
  int32_t n_nodes;     /* .data1[0] nodesNo */
  int32_t n_steps;     /* .data1[1] stepsNo */
  int32_t n_species;   /* S, 1..NNPP_MAX_SPECIES (reference: 10) */
  int32_t chemistry;   /* NNPP_CHEM_* */

  /* .data2 (51 values, utils.h:2348-2398) */
  double t0, t1, dt, dx, dy, dz, delta, CSolid, CSat, C0Na, C0Ca, C0HCO3, C0HPO4, C0Cl;
  double pKMgOH2, pKMgCO3, pKMgHCO3, pKMgPO4, pKCaCO3, pKCa5OHPO4, pKH2O, pKHCO3, pKHPO4;
  double phim, phimse, beta, alphaA, alphaC;
  double k1f, k1b, k2f, k2b, k3f, k3b, k4f, k4b, k5f, k5b, k6f, k6b, k7f, k7b, k8f, k8b, k9f, k9b;
  double JHp, T, R, F, vol_red;

  /* .datac1 .. .datac<S> (7 values each, utils.h:2400-2478) */
  double z[NNPP_MAX_SPECIES];
  double D_solid[NNPP_MAX_SPECIES];
  double D_liquid[NNPP_MAX_SPECIES];
  double s1[NNPP_MAX_SPECIES];   /* diffusion kernel exponent: 1/xi^(3+2 s1) */
  double s2[NNPP_MAX_SPECIES];   /* migration kernel exponent: 1/xi^(3+2 s2) */
  double K1[NNPP_MAX_SPECIES];   /* Kprop1 */
  double K2[NNPP_MAX_SPECIES];   /* Kprop2 */

  /* .datac<S+1> (4 values, utils.h:2480-2483) */
  double sigma0, sigma1, s_phi, K_phi;

  /* extensions */
  int32_t solid_bonds[NNPP_MAX_SPECIES]; /* 1: species diffuses over solid-solid bonds and its
                                            solid-liquid bonds are scaled by BV (reference:
                                            Mg only, utils.h:100-159); 0: solid-solid bonds
                                            are skipped (utils.h:211-213) */
  int32_t clamp[NNPP_MAX_SPECIES];       /* NNPP_CLAMP_* per species (main.cpp:805-837) */
  double clamp_floor;       /* reference: 1e-8 / CSolid */
  double front_threshold;   /* solid -> liquid when U_Mg <= this; reference CSat / CSolid */
  double phase_eps;         /* icvalues IC: solid iff U_Mg > phase_eps; reference EPS = 1e-5 */
  double r_diss;            /* NNPP_CHEM_GENERIC interface sink rate (1/time) */
  int32_t sub_steps;        /* voln1rel refresh cadence; reference sub_step_no = 8 */
  int32_t reserved;
} nnpp_params;

/* Step statistics (accumulated over the steps of one nnpp_step call). */
typedef struct nnpp_step_stats {
  int64_t steps;            /* steps taken */
  int64_t cg_iterations;    /* total CG iterations (inner iterations in mixed mode) */
  int64_t cg_spmv_fp64;     /* FP64 Poisson operator applications */
  int64_t cg_spmv_fp32;     /* FP32 Poisson operator applications (mixed mode) */
  double last_relres;       /* final relative residual of the last solve (FP64 check) */
  int64_t flipped;          /* solid -> liquid flips */
  int64_t solid_count;      /* solid nodes after the last step */
  double voln1rel;          /* current relative Mg volume feeding BV */
  int64_t nnz_flux;         /* directed bonds per flux evaluation */
  int64_t nnz_ll;           /* liquid-liquid directed bonds in the last Poisson operator */
  int32_t pot_col_bytes;    /* bytes per stored column of that operator: 2 (steps) or 4 */
  int32_t pot_weight_bytes; /* bytes per stored weight read in mixed mode: 1 (dictionary
                               class) or 4 (FP32 value); 0 for the stencil form */
  int32_t pot_stencil;      /* the operator used the stencil form (16 bytes per row, no
                               per-entry stream; pot_col_bytes = 0): 1 with the table-driven
                               loop, 3 with the loop specialised for the radius-3 lattice
                               ball, 4 with the z-marching kernels of a full box lattice (no
                               row masks: a free flag per node); 0 for the column form */
  int32_t flux_stencil;     /* the explicit flux used the lattice stencil form: 1 with FP64
                               bond sums, 2 with the FP32 bond sums of NNPP_MIXED, 3 the
                               z-marching form of a full box lattice (FP64 sums); 0 for the
                               column form */
} nnpp_step_stats;

typedef struct nnpp_ctx nnpp_ctx;

/* ---------------------------------------------------------------- host-only helpers ----- */

/* sizeof(nnpp_params) and sizeof(nnpp_step_stats) as compiled into the library, so bindings
 * in other languages can check their struct mirrors. No reference counterpart. */
int64_t nnpp_abi_params_size(void);
int64_t nnpp_abi_stats_size(void);

/* Fill *p with the reference's hard-coded behaviour for S species: solid_bonds = {1,0,...},
 * clamp = ALWAYS for species 1..9 and LIQUID for species 10 (main.cpp:805-837),
 * clamp_floor = 1e-8/CSolid, front_threshold = CSat/CSolid, phase_eps = 1e-5,
 * sub_steps = 8. Call after the numeric fields are set (it reads CSolid and CSat). */
int nnpp_params_defaults(nnpp_params* p, int32_t n_species, int32_t chemistry);

/* Read <job>.data1, <job>.data2, <job>.datac1 .. <job>.datac<S+1> into *p and apply
 * nnpp_params_defaults. Replaces main.cpp:210-231 + utils::parameter::set_int/set_double
 * (utils.h:2334-2484). Unlike the reference, a missing file is an error (NNPP_ERR_IO).
 * Then, if <job>.nnpp exists, its "key = value" lines set the extension fields that the
 * reference files cannot hold: r_diss, front_threshold, clamp_floor, phase_eps, sub_steps,
 * solid_bonds (S integers), clamp (S integers); '#' starts a comment. */
int nnpp_read_params(const char* job, int32_t n_species, int32_t chemistry, nnpp_params* p);

/* Whitespace-separated readers; replace utils::read_vector_from_file and
 * utils::read_matrix_from_file (utils.h:1851-1887). nodes: n lines "x y z" (main.cpp:234). */
int nnpp_read_vector(const char* path, int64_t n, double* out);
int nnpp_read_ivector(const char* path, int64_t n, int32_t* out);
int nnpp_read_nodes(const char* path, int64_t n, double* x, double* y, double* z);

/* Output / restart CSV (utils.h:2053-2100 writer, utils.h:1928-2027 reader): header
 * "X,Y,Z,<species names>,Phi,Phase", 16 significant digits, one row per node. Species names
 * are the reference's (Mg++,H+,OH-,HCO3-,CO32-,HPO42-,PO43-,Ca++,Na+,Cl-) for S = 10 and
 * U1..U<S> otherwise. The reader looks columns up by name; a missing column leaves zeros
 * (the reference zeroes the state before reading, main.cpp:513-524). Unlike the reference,
 * a missing file is NNPP_ERR_IO instead of a silent zero state. */
int nnpp_write_output_csv(const char* path, int32_t n, int32_t n_species, const double* x,
                          const double* y, const double* z, const double* U,
                          const int32_t* phase);
int nnpp_read_start_csv(const char* path, int32_t n, int32_t n_species, double* U,
                        int32_t* phase);

/* ELL adapter: turns a CSR neighbour list (rows sorted by id, no self) into the reference's
 * layout, an n x max_len column-major int matrix with self at column 0, neighbours in
 * ascending id after it and -1 padding, plus lengths that count self (grid.h:145-149,170-195,
 * main.cpp:34,138). Returns NNPP_ERR_OVERFLOW if a row does not fit. */
int nnpp_neighbors_to_ell(int32_t n, const int64_t* row_ptr, const int32_t* col,
                          int32_t max_len, int32_t* ell, int32_t* lengths);

/* The inverse adapter: the reference's horizons and horizonsLengths (n x max_len column-major,
 * self at column 0, lengths counting self; grid.h:145-149,170-195) to CSR with rows sorted by
 * id and self dropped, e.g. for nnpp_set_neighbors. row_ptr (n+1) is always written; col
 * (sum of lengths - n entries) when not NULL, so a first call with col = NULL sizes it.
 * NNPP_ERR_ARG on a malformed row (self not first, an id out of range, a repeated id). */
int nnpp_neighbors_from_ell(int32_t n, int32_t max_len, const int32_t* ell,
                            const int32_t* lengths, int64_t* row_ptr, int32_t* col);

/* ------------------------------------------------------------------- context ------------- */

/* Create a context on HIP device `device`. Replaces nothing (the reference has no device). */
int nnpp_create(nnpp_ctx** ctx, int32_t device);
int nnpp_destroy(nnpp_ctx* ctx);
const char* nnpp_last_error(const nnpp_ctx* ctx);

/* Set the parameters; replaces utils::parameter (utils.h:2186-2485). */
int nnpp_set_params(nnpp_ctx* ctx, const nnpp_params* p);

/* Upload node coordinates (SoA) and optional per-node volumes (NULL: dx*dy*dz for every node,
 * as utils.h:122-124). Replaces the nodes matrix read at main.cpp:234. */
int nnpp_set_nodes(nnpp_ctx* ctx, int32_t n, const double* x, const double* y, const double* z,
                   const double* volumes);

/* Spatial-hash neighbour construction on the GPU for the nodes set by nnpp_set_nodes.
 * Replaces build_clouds_hash(max_node_no, r, dx, dy, dz, nodes, clouds, cloud_lengths,
 * max_cloud_length) (grid.h:43-204; called at main.cpp:304-305 with r = delta*(1+1e-5) and
 * cells 10*dx x 10*dy x 10*dz). Same predicate, same cell logic; the result is kept on the
 * device as CSR (rows sorted by neighbour id, self excluded). *nnz_out = directed bonds,
 * *max_len_out = longest row INCLUDING self (the reference's max_cloud_length).
 * NNPP_ERR_OVERFLOW when a row (with self) would exceed max_len, as grid.h:187-189. */
int nnpp_build_neighbors(nnpp_ctx* ctx, double r, double cell_dx, double cell_dy,
                         double cell_dz, int32_t max_len, int64_t* nnz_out,
                         int32_t* max_len_out);

/* Implant classification on the GPU. Replaces build_clouds(0, max_node_no, r, 0, nodes,
 * bunnyNodes, horizonsBunny, horizonsBunnyLengths, ...) (kd_tree_interface.h:156-252) as
 * main.cpp:316-320 calls it, r = 3 * 0.4 * dx * (1 + 1e-5): counts[i] (n ints) = the number of
 * the nb implant points p with sqrt(x*x + y*y + z*z) <= r, (x, y, z) = p - node i (the
 * reference's Node::distance_euclidean), over a uniform grid of the points instead of a
 * kd-tree. NNPP_ERR_OVERFLOW when a count exceeds max_count (the reference throws,
 * kd_tree_interface.h:214-217); counts are still written. ic_out (F*n doubles, may be NULL;
 * the ten reference species only) receives the initial values main.cpp:422-457 selects: count
 * 0 -> liquid (H+ and OH- 1e-4, HCO3- C0HCO3, HPO4 2- C0HPO4, Ca2+ C0Ca, Na+ C0Na, Cl- C0Cl,
 * all / CSolid, the rest 0), else solid (Mg2+ CSolid/CSolid = 1, the rest 0); pass them to
 * nnpp_init_state. Uses the nodes of nnpp_set_nodes. */
int nnpp_implant_classify(nnpp_ctx* ctx, int64_t nb, const double* bx, const double* by,
                          const double* bz, double r, int32_t max_count, int32_t* counts,
                          double* ic_out);

/* The implant classification of nnpp_implant_classify without any device allocation. No
 * counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc). Same arguments, results
 * and errors; its scratch (the points, their cell grid, the counts) is a static device array
 * of NNPP_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until the process
 * ends, so nb <= NNPP_NOALLOC_MAX_POINTS, the node count <= NNPP_NOALLOC_MAX_NODES
 * (NNPP_ERR_ARG otherwise), and the cell grid is coarsened to at most NNPP_NOALLOC_MAX_CELLS
 * cells (the counts do not change). Every call on one GPU uses the same array: calls that
 * overlap on one GPU (other threads, other contexts of the same device) race. */
#define NNPP_NOALLOC_MAX_POINTS (1 << 19)
#define NNPP_NOALLOC_MAX_NODES (1 << 22)
#define NNPP_NOALLOC_MAX_CELLS (1 << 18)
int nnpp_implant_classify_noalloc(nnpp_ctx* ctx, int64_t nb, const double* bx, const double* by,
                                  const double* bz, double r, int32_t max_count,
                                  int32_t* counts, double* ic_out);

/* Voronoi-approximate node volumes for irregular clouds (the reference multiplies every bond
 * by dx*dy*dz, utils.h:122-124; pass the result to nnpp_set_nodes as per-node volumes).
 * Probes on a lattice of spacing h fill the nodes' bounding box widened by margin on every
 * side (round((extent + 2 margin) / h) probes per axis, centres at lo + (k + 1/2) h); every
 * probe goes to its nearest node within radius (ties to the lowest index), and vol_out[i]
 * (n doubles) = probes of node i * h^3. Probes with no node within radius count for none.
 * On a lattice of spacing dx with margin dx/2, radius >= dx and h = dx/k every node gets dx^3
 * exactly. No reference counterpart. */
int nnpp_voronoi_volumes(nnpp_ctx* ctx, double h, double margin, double radius,
                         double* vol_out);

/* Copy the device CSR out (row_ptr: n+1, col: nnz). */
int nnpp_get_neighbors(nnpp_ctx* ctx, int64_t* row_ptr, int32_t* col);

/* Upload a CSR built elsewhere (rows sorted by id, no self) instead of building it. */
int nnpp_set_neighbors(nnpp_ctx* ctx, const int64_t* row_ptr, const int32_t* col);

/* Boundary conditions: bctypes (n ints; 0 = Dirichlet node) and the F per-field value arrays
 * packed field-major (F*n doubles), raw as read from <job>.bcvalues1..F. A node/field is
 * Dirichlet iff bctypes==0 and value != -1; species values are divided by CSolid, phi is
 * used raw (main.cpp:571-625,676-711). */
int nnpp_set_bc(nnpp_ctx* ctx, const int32_t* bctypes, const double* bcvalues);

/* State slots. */
enum { NNPP_SLOT_N = 0,  /* U^n: the state the explicit step advances */
       NNPP_SLOT_N1 = 1  /* U^{n+1}: the Newton iterate used by nnpp_residuals */ };

/* Upload / download all F fields of a slot (F*n doubles, field-major). */
int nnpp_set_state(nnpp_ctx* ctx, int32_t slot, const double* U);
int nnpp_get_state(nnpp_ctx* ctx, int32_t slot, double* U);

/* Node phase, 1 = solid, 0 = liquid (main.cpp:413-457). */
int nnpp_set_phase(nnpp_ctx* ctx, const int32_t* phase);
int nnpp_get_phase(nnpp_ctx* ctx, int32_t* phase);

/* Initial state from icvalues (F*n, field-major), main.cpp:400-421: U^n = icvalues, phase =
 * (U_Mg > phase_eps), then Dirichlet values are written into every field (the reference
 * writes them into the first Newton iterate, main.cpp:556-640). Sets solid_count, vol0rel =
 * sum over solid nodes of (icvalue_Mg - CSat/CSolid), before any Dirichlet value
 * (main.cpp:474-479), and voln1rel = 0, and
 * resets the step counter to s0 = 1. */
int nnpp_init_state(nnpp_ctx* ctx, const double* icvalues);

/* Set the relative volume and step counter explicitly (restart, main.cpp:506-542). */
int nnpp_set_time(nnpp_ctx* ctx, int64_t step, double voln1rel, double vol0rel);

/* Read them back (the step counter, voln1rel and vol0rel variables of main.cpp:474-479,
 * 944-949); any pointer may be NULL. */
int nnpp_get_time(nnpp_ctx* ctx, int64_t* step, double* voln1rel, double* vol0rel);

/* Residuals F_k of every field for the iterate U^{n+1} (slot N1) and the previous state
 * (slot N), exactly as utils::Residuals (utils.h:27-842): bond sums of the iterate, lagged
 * potential, reactions, (U1-Un)/dt, phi residual on rows with a liquid-liquid bond.
 * apply_bc != 0 also applies main.cpp:676-711 (Dirichlet rows: F = U1 - bc).
 * F_out: F*n doubles, field-major. precision: NNPP_FP64 or NNPP_MIXED. */
int nnpp_residuals(nnpp_ctx* ctx, double voln1rel, int32_t apply_bc, int32_t precision,
                   double* F_out);

/* Poisson operator (K11 of utils.h:1794-1848 without its identity rows): y_i =
 * sum over liquid-liquid bonds of w_ij (x_i - x_j) for rows with such a bond, y_i = 0
 * elsewhere; w_ij = Beta(xi) V_j / xi^(3+2 s_phi). Uses the phase on the device. */
int nnpp_poisson_apply(nnpp_ctx* ctx, const double* x, double* y);

/* The operator the potential solve iterates with, applied once in the solver's precision
 * (NNPP_FP64, or NNPP_MIXED: FP32 values and Kahan-compensated FP32 row sums) and storage
 * format: y_i = s_i (sum over liquid bonds of w_ij x_i - sum over free-free bonds of w_ij x_j)
 * on free rows (s_i = V_i/(dx dy dz)), 0 elsewhere; x and y are n doubles (converted to and
 * from FP32 in mixed mode). Checks the hot SpMV kernel in isolation. No reference
 * counterpart (K11 of utils.h:1794-1848 restricted to the free rows and scaled). */
int nnpp_potential_operator_apply(nnpp_ctx* ctx, int32_t precision, const double* x, double* y);

/* Solver controls for the potential (replaces Eigen GMRES at main.cpp:764-779):
 * relative residual rtol (default 1e-10), maxit (default 20000); for NNPP_MIXED the inner
 * FP32 solve reduces the residual norm by inner_rtol per refinement pass before the FP64
 * residual replaces it (default 1e-3). */
int nnpp_set_solver(nnpp_ctx* ctx, double rtol, int32_t maxit, double inner_rtol);

/* Storage format of the bond slices and the potential operator. NNPP_FORMAT_AUTO (default):
 * the stencil form of the potential operator when the neighbour lists are a lattice stencil
 * (every j - i in a table of at most 128 offsets, weights depending on the offset only, no
 * per-node volumes): one 128-bit mask per row; otherwise 16-bit column steps unless a row
 * needs a larger step, and a 1-byte weight-class dictionary when the operator has at most 255
 * distinct nonzero weights. The flags force 32-bit columns, plain FP64/FP32 weights and/or the
 * column form. Every format gives bit-identical operator rows (the z-marching explicit flux of a
 * full box lattice sums in another order, within 1e-12); they differ in bytes read per bond. No reference counterpart (the reference stores an
 * Eigen::SparseMatrix, utils.h:1794-1848). Takes effect at the next operator build. */
#define NNPP_FORMAT_AUTO 0
#define NNPP_FORMAT_WIDE_COLUMNS 1
#define NNPP_FORMAT_NO_DICTIONARY 2
#define NNPP_FORMAT_NO_STENCIL 4
#define NNPP_FORMAT_NO_BALL 8 /* stencil form: the table-driven SpMV loop even where the
                                 offsets are the radius-3 lattice ball (same results) */
#define NNPP_FORMAT_NO_ZMARCH 16 /* lattice-ball stencil on a full box lattice: the tile SpMV
                                    (one 256-row tile per LDS window) instead of the z-marching
                                    one (column tiles walking z); FP64 rows identical, dot
                                    products summed in another order. Also keeps the explicit
                                    flux in its tile stencil form (flux_stencil 1, the column
                                    form's sums) instead of the z-marching one (3), and the
                                    implicit step's Jacobian apply alike */
int nnpp_set_format(nnpp_ctx* ctx, int32_t flags);

/* Solve the potential for the charge of slot N: phi^{n+1} on free nodes (liquid, with a
 * liquid-liquid bond, not Dirichlet), Dirichlet phi = bc, other nodes keep their value.
 * Warm-starts from the phi in slot N and writes the result there. */
int nnpp_poisson_solve(nnpp_ctx* ctx, int32_t precision, int32_t* iters, double* relres);

/* Advance nsteps explicit steps (the reference's time loop main.cpp:547-1025 with the
 * implicit Newton/GMRES solve replaced by one explicit update):
 *   1. species: U^{n+1} = U^n - dt * G(U^n), G = utils::Residuals without (U1-Un)/dt and
 *      with U^{n+1} := U^n, then Dirichlet rows, then the clamp (main.cpp:805-837);
 *   2. phi^{n+1} from the charge of U^n (utils.h:831-840) by CG, warm-started;
 *   3. front update: solid -> liquid where U_Mg^{n+1} <= front_threshold (main.cpp:915-922);
 *   4. every sub_steps steps: voln1rel = sum_solid (U_Mg - CSat/CSolid) / vol0rel
 *      (main.cpp:944-949), which feeds BV = 10^(-vol_red (1 - voln1rel)).
 * stats may be NULL. Returns NNPP_ERR_NOCONV if a solve fails (state is still advanced),
 * NNPP_ERR_NONFINITE if check_finite != 0 and the state holds NaN/Inf. */
int nnpp_step(nnpp_ctx* ctx, int32_t nsteps, int32_t precision, int32_t check_finite,
              nnpp_step_stats* stats);

/* -------------------------------------------------------- implicit (reference) step ----
 * The reference's own time step, main.cpp:547-1025: modified Newton with the Jacobian of
 * utils::NewtonMatrices (utils.h:844-1849) frozen at U^n for the whole step, every field solved
 * on its own block K_f dU_f = -F_f. Per step:
 *   - iterate: on the first implicit step after the state was set, U^n with the Dirichlet
 *     values (main.cpp:556-640); afterwards, with start = NNPP_NEWTON_START_ROTATE, the buffer
 *     the reference's rotation (main.cpp:1002-1024) leaves there: U^{n-1}, which is all zeros on
 *     the second step (U1n0.setZero(), main.cpp:388); with NNPP_NEWTON_START_STATE, U^n with
 *     the Dirichlet values every step;
 *   - at most newton_maxit iterations (MAX_ITER = 200, main.cpp:36) of: F = utils::Residuals
 *     (iterate, U^n) with the Dirichlet rows F = U - bc (main.cpp:667-711); species blocks
 *     solved together by Jacobi-preconditioned BiCGSTAB, the potential block by the
 *     potential CG (both from zero, to ||b - K x|| <= lin_rtol ||b||; the reference runs
 *     GMRES(30) with a diagonal preconditioner to EPS = 1e-5, main.cpp:714-779); U += dU;
 *     res_f = ||dU_f||_2; the clamp (main.cpp:805-837); stop once
 *     max_f |res_f - res_f of the previous iteration| < newton_tol (main.cpp:863-911; the
 *     previous value is 0 at the first iteration);
 *   - the front update on the iterate (main.cpp:915-922), voln1rel every sub_steps steps
 *     (main.cpp:944-949), the step counter, and the rotation (U^{n-1}, U^n, iterate) <-
 *     (U^n, iterate, U^{n-1}); slot N then holds U^{n+1} and slot N1 the next start.
 * Species solves are FP64; precision selects the potential solve (NNPP_FP64 or NNPP_MIXED). */
#define NNPP_NEWTON_START_ROTATE 0
#define NNPP_NEWTON_START_STATE 1
typedef struct nnpp_newton_opts {
  double newton_tol;     /* EPS = 1e-5 (main.cpp:35) */
  double lin_rtol;       /* relative residual of every linear solve (default 1e-5 = EPS) */
  int32_t newton_maxit;  /* MAX_ITER = 200 */
  int32_t lin_maxit;     /* MAX_ITER = 200 */
  int32_t start;         /* NNPP_NEWTON_START_ROTATE (default) or NNPP_NEWTON_START_STATE */
  int32_t reserved;
} nnpp_newton_opts;

typedef struct nnpp_newton_stats {
  int32_t steps;
  int32_t newton_iterations;  /* summed over the steps */
  int32_t unconverged_steps;  /* steps that ran out of Newton iterations */
  int32_t lin_iterations;     /* BiCGSTAB iterations (the species advance together) */
  int32_t cg_iterations;      /* potential CG iterations */
  int32_t flipped;            /* solid -> liquid flips */
  int64_t solid_count;
  double voln1rel;
  double last_resmax;         /* Newton test value of the last iteration */
  double lin_relres_max;      /* largest final relative residual of a species solve */
} nnpp_newton_stats;

/* Fill opts with the reference's values. */
int nnpp_newton_defaults(nnpp_newton_opts* opts);

/* Advance nsteps implicit steps. opts NULL: the reference's values; stats may be NULL.
 * Returns NNPP_ERR_NOCONV when a linear solve stops at its iteration limit (the step is still
 * taken, as the reference takes it), NNPP_ERR_NONFINITE if check_finite != 0 and the state
 * holds NaN/Inf, NNPP_ERR_UNSUPPORTED while a differentiation tape is recording. Replaces the
 * body of the time loop main.cpp:547-1025 (Jacobian, Newton loop, front, volume). */
int nnpp_step_implicit(nnpp_ctx* ctx, int32_t nsteps, int32_t precision, int32_t check_finite,
                       const nnpp_newton_opts* opts, nnpp_newton_stats* stats);

/* The frozen Jacobian at slot N (utils::NewtonMatrices, utils.h:844-1849) applied block by
 * block: out_f = K_f in_f for every field (F*n doubles each, field-major); diag (F*n, may be
 * NULL) receives the diagonals. Species blocks: diagonal 1/dt + sum of the diffusion weights
 * - R'_k (R' at U^n), off-diagonals -(diffusion + migration with the lagged phi^n) by phase
 * pair; the potential block is K11 including its identity rows; Dirichlet rows are identity
 * rows. The reference assembles eleven Eigen sparse matrices; here they are never stored. */
int nnpp_jacobian_apply(nnpp_ctx* ctx, double voln1rel, const double* in, double* out,
                        double* diag);

int64_t nnpp_abi_newton_stats_size(void);

/* Front update alone (main.cpp:915-922) on slot N. */
int nnpp_front_update(nnpp_ctx* ctx, double threshold, int32_t* n_flipped);

/* Raw device pointer of a field in a slot (for zero-copy interop). nnpp_step swaps the
 * buffers of slot N with its scratch buffer, so the pointer is valid until the next
 * nnpp_step, nnpp_set_nodes or nnpp_destroy call. */
int nnpp_state_device_ptr(nnpp_ctx* ctx, int32_t slot, int32_t field, void** dptr);

/* HIP stream the context issues its work on (hipStream_t as void*). */
int nnpp_stream(nnpp_ctx* ctx, void** stream);

/* Measurement hook, no reference counterpart: launches one hot kernel `reps` times on the
 * context's stream between two HIP events and returns the mean duration per launch in ms.
 * kernel: NNPP_KERNEL_FLUX (explicit flux/update kernel on slot N), NNPP_KERNEL_POT_SPMV
 * (potential operator application with its dot product) or NNPP_KERNEL_CG_ITER (one fused
 * inner iteration of the potential solve: NNPP_MIXED on a full box lattice only, where the
 * solve runs single-reduction CG iterations as one kernel each; NNPP_ERR_ARG otherwise);
 * precision selects the FP64 or the mixed variant. Leaves the state of slot N unchanged. */
enum { NNPP_KERNEL_FLUX = 0, NNPP_KERNEL_POT_SPMV = 1, NNPP_KERNEL_CG_ITER = 2 };
int nnpp_time_kernel(nnpp_ctx* ctx, int32_t kernel, int32_t precision, int32_t reps,
                     double* avg_ms);

/* ------------------------------------------------------------- differentiation ----------
 * Adjoint of nnpp_step's explicit FP64 step (no reference counterpart: the reference has no
 * derivative code). A tape records the steps taken between nnpp_tape_begin and nnpp_tape_end;
 * nnpp_tape_adjoint then propagates dL/dU backwards through them:
 *   - species update: the bond sums are linear in the state of the neighbours and in the
 *     diffusivities; their transpose is evaluated as a gather over the same (symmetric) bond
 *     lists; Dirichlet and clamped entries pass no derivative;
 *   - potential: phi^{n+1} = M^{-1} b(U^n) on free nodes, so dL/db = M^{-1} dL/dphi^{n+1}
 *     (M is symmetric) by the same CG;
 *   - BV = 10^(-vol_red (1 - voln1rel)) is differentiated through the voln1rel updates;
 *   - the phase trajectory is held fixed (front flips are piecewise constant).
 * Generic chemistry only (NNPP_ERR_UNSUPPORTED for the reference chemistry). */
typedef struct nnpp_adjoint_stats {
  int64_t steps;             /* steps reversed */
  int64_t recomputed_steps;  /* forward steps recomputed from checkpoints */
  int64_t cg_iterations;     /* iterations of the adjoint potential solves */
  int64_t tape_bytes;        /* device bytes held by checkpoints and the segment buffer */
  double last_relres;        /* relative residual of the last adjoint potential solve */
} nnpp_adjoint_stats;

/* Start recording at the current state. Every checkpoint_every steps (1 = every step, the full
 * tape of states) nnpp_step stores a device copy of the state, phase and step scalars; the
 * parameters and solver settings are frozen into the tape. While a tape exists, calls that
 * change the parameters, nodes, neighbours, boundary conditions, state, phase or time return
 * NNPP_ERR_STATE, and nnpp_step with NNPP_MIXED returns NNPP_ERR_UNSUPPORTED. */
int nnpp_tape_begin(nnpp_ctx* ctx, int32_t checkpoint_every);

/* Reverse every recorded step. lam: dL/dU at the current state (F*n doubles, field-major).
 * grad (2S+1 doubles): dL/dD_liquid[0..S-1], dL/dD_solid[0..S-1], dL/dr_diss. lam0 (F*n, may
 * be NULL): dL/dU at the state where the tape began (phase held fixed). Segments between
 * checkpoints are recomputed forward (bit-identical to the recorded run). The current state is
 * left as it was; the tape stays valid for further calls. stats may be NULL. */
int nnpp_tape_adjoint(nnpp_ctx* ctx, const double* lam, double* grad, double* lam0,
                      nnpp_adjoint_stats* stats);

/* Stop recording and free the tape. */
int nnpp_tape_end(nnpp_ctx* ctx);

int64_t nnpp_abi_adjoint_stats_size(void);

#ifdef __cplusplus
}
#endif

#endif /* NNPP_H */

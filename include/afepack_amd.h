/*
 * afepack_amd.h -- C ABI of the MI355X (gfx950) assemble-and-solve path.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.
 * Host arrays (meshes, patterns, fixtures) are borrowed for the duration of a
 * call.  Vectors passed as `d_*` are device pointers (HBM) on the context's
 * device; the library never keeps them after the call returns.  All device
 * work of a context is issued on the context's stream; every call returns
 * with that stream synchronised unless its comment says otherwise.
 *
 * Each entry point names the AFEPack interface it replaces (paths relative to
 * the AFEPack source tree).  The C++ drop-in classes in
 * afepack_amd/include/AFEPack/ (StiffMatrix, MassMatrix, AMGSolver, ...)
 * are thin facades over these calls.
 */
#ifndef AFEPACK_AMD_H
#define AFEPACK_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define AFE_OK 0
#define AFE_E_ARG (-1)          /* invalid argument / size mismatch          */
#define AFE_E_NOMEM (-2)        /* host or device allocation failed          */
#define AFE_E_HIP (-3)          /* a HIP runtime call failed                 */
#define AFE_E_SINGULAR (-4)     /* coarse matrix singular (dgetrf failure)    */
#define AFE_E_NOTCONVERGED (-5) /* Krylov solve hit max_it above rtol        */
#define AFE_E_PATTERN (-6)      /* lazyReinit with a different pattern        */
#define AFE_E_BUDGET (-7)       /* request exceeds the HBM admission budget   */
#define AFE_E_NODEVICE (-8)     /* no HIP device / device index out of range  */

/* element kinds (template elements triangle.1, triangle.2, tetrahedron.1) */
#define AFE_P1_TRI 1
#define AFE_P2_TRI 2
#define AFE_P1_TET 3

/* bilinear operators (BilinearOperator.templates.h:342-406 and the
 * variable-coefficient override of examples/elliptic_equation) */
#define AFE_OP_STIFF 0
#define AFE_OP_MASS 1
#define AFE_OP_VARSTIFF 2

/* right-hand sides for afe_rhs (Operator::L2Discretize) */
#define AFE_F_POISSON 0 /* f = dim*pi^2*prod sin(pi x_d)                   */
#define AFE_F_VARCOEF 1 /* f = -div(a grad u), u = prod sin(pi x_d)         */
#define AFE_F_NODAL 2   /* f = interpolant of DOF values                    */
#define AFE_F_ONE 3     /* f = 1                                            */

/* solve modes for afe_amg_pcg */
#define AFE_PREC_FP64 0 /* FP64 PCG, FP64 symmetric V-cycle preconditioner */
#define AFE_PREC_FP32 1 /* FP64 PCG, FP32 symmetric V-cycle preconditioner */

typedef struct afe_ctx afe_ctx;
typedef struct afe_space afe_space;
typedef struct afe_matrix afe_matrix;
typedef struct afe_amg afe_amg;

/* Seeded Fourier field g(x) = scale * sum_m w_m sin(2 pi k_m.x + phi_m);
 * a(x) = exp(g(x)).  k is n_modes x dim, row-major. */
typedef struct afe_field {
  int32_t n_modes;
  int32_t dim;
  const double *k;
  const double *phi;
  const double *w;
  double scale;
} afe_field;

typedef struct afe_amg_params {
  double alpha;      /* strength threshold, AMGSolver.h:87-92 (0.25)        */
  uint32_t nmpo;     /* max coarsest order (50)                            */
  double nmpsd;      /* min coarsest density (0.382)                       */
  int32_t lazy;      /* 1: lazyReinit/lazyProject hierarchy               */
  int32_t smooth;    /* GS sweeps per side (AMGSolver::smoothStep, 3)      */
  int32_t build_fp32;/* also keep FP32 copies for the mixed V-cycle        */
} afe_amg_params;

typedef struct afe_solve_stats {
  int32_t iterations;     /* V-cycles (stationary) or PCG iterations       */
  int32_t converged;      /* 1 if the tolerance was reached                */
  double rel_res_l2;      /* ||b - A x||_2 / ||b||_2 at exit (true residual) */
  double rel_res_l1;      /* ||A x - b||_1 / ||A x0 - b||_1 (AMGSolver.cpp:564,617) */
  double res_l1;          /* ||A x - b||_1 at exit                         */
  double seconds;         /* device wall time of the solve                 */
} afe_solve_stats;

/* ---- context ------------------------------------------------------------ */
const char *afe_last_error(void);
int afe_version(void);
int afe_ctx_create(int device, afe_ctx **out);
int afe_ctx_destroy(afe_ctx *ctx);
int afe_ctx_synchronize(afe_ctx *ctx);
/* bytes of device memory held by live objects of this context */
int afe_ctx_device_bytes(afe_ctx *ctx, int64_t *bytes);

/* ---- device memory (plain HBM buffers for the d_* arguments) ------------- */
int afe_dev_alloc(afe_ctx *ctx, int64_t bytes, void **d_ptr);
int afe_dev_free(afe_ctx *ctx, void *d_ptr);
int afe_dev_upload(afe_ctx *ctx, void *d_dst, const void *h_src, int64_t bytes);
int afe_dev_download(afe_ctx *ctx, void *h_dst, const void *d_src,
                     int64_t bytes);
int afe_dev_copy(afe_ctx *ctx, void *d_dst, const void *d_src, int64_t bytes);
int afe_dev_zero(afe_ctx *ctx, void *d_ptr, int64_t bytes);

/* ---- synthetic meshes (the benchmark inputs) ---------------------------- */
/* 2D: n x n squares on the unit square, split (n00,n10,n11),(n00,n11,n01) as
 * examples/FAST/src/poisson_lshape_stretched_afepack.cpp:262-286; interior
 * nodes jittered by U(-jitter h, jitter h) per coordinate from seed.
 * Sizes: (n+1)^2 vertices, 2 n^2 triangles. */
int afe_mesh_square(int32_t n, double jitter, uint64_t seed, double *coords,
                    int32_t *elem_vert, int32_t *vert_mark);
/* 3D: n^3 cubes on the unit cube, 6 tetrahedra each as tools/tricube.cpp:67-90.
 * Sizes: (n+1)^3 vertices, 6 n^3 tetrahedra. */
int afe_mesh_cube(int32_t n, double jitter, uint64_t seed, double *coords,
                  int32_t *elem_vert, int32_t *vert_mark);
/* displace interior vertices by amp_h*sin(2 pi k x_d) per coordinate
 * (moving-mesh update; amp_h is absolute, e.g. 0.1 h) */
int afe_mesh_displace(int32_t dim, int64_t n_vert, double *coords,
                      const int32_t *vert_mark, double amp_h, int32_t k);
/* seeded n_modes-mode field with |g| <= amplitude; k, phi, w are outputs of
 * size n_modes*dim, n_modes, n_modes; *scale receives the scale factor. */
int afe_field_random(int32_t n_modes, int32_t dim, double amplitude,
                     uint64_t seed, double *k, double *phi, double *w,
                     double *scale);

/* ---- FE space: FEMSpace::buildDof + buildDofBoundaryMark --------------- */
/* Replaces FEMSpace::buildElement/buildDof/buildDofBoundaryMark
 * (include/AFEPack/FEMSpace.templates.h:183-258,689-880) and
 * BilinearOperator::buildSparsityPattern (BilinearOperator.templates.h:64-172):
 * first-touch DOF numbering, the deal.II-style CSR pattern (diagonal first,
 * rest ascending) and the element-to-CSR gather map, built on the host and
 * uploaded to HBM once. */
/* ctx may be NULL: a host-only space (numbering and pattern only) */
int afe_space_create(afe_ctx *ctx, int32_t kind, int64_t n_vert,
                     const double *coords, int64_t n_elem,
                     const int32_t *elem_vert, const int32_t *vert_mark,
                     afe_space **out);
/* the same with the mesh's side list (Mesh geometry(1): n_side vertex pairs
 * side_vert and their boundary marks side_mark, as EasyMesh .s files and
 * the Mesh text format give them): a P2 edge DOF takes the mark of its side
 * (FEMSpace::buildDofBoundaryMark, FEMSpace.templates.h:689-701); an element
 * edge missing from the list is an error.  afe_space_create without sides
 * marks an edge 1 when it lies in a single element. */
int afe_space_create_sides(afe_ctx *ctx, int32_t kind, int64_t n_vert,
                           const double *coords, int64_t n_elem,
                           const int32_t *elem_vert, const int32_t *vert_mark,
                           int64_t n_side, const int32_t *side_vert,
                           const int32_t *side_mark, afe_space **out);
int afe_space_destroy(afe_space *s);
int afe_space_info(const afe_space *s, int64_t *n_dof, int64_t *nnz,
                   int32_t *n_loc);
/* copies out DOF data; any pointer may be NULL */
int afe_space_dofs(const afe_space *s, int32_t *elem_dof, int32_t *dof_mark,
                   double *dof_point);
/* the pattern with the reference's size_t index width */
int afe_space_pattern(const afe_space *s, int64_t *row_ptr, int64_t *col);
/* moving mesh: new vertex coordinates, same connectivity (values-only
 * re-assembly keeps pattern, gather map and lazy hierarchy) */
int afe_space_update_coords(afe_space *s, const double *coords);

/* ---- assembly: BilinearOperator::build ---------------------------------- */
/* Replaces BilinearOperator::build / buildSparseMatrix / getElementMatrix
 * (BilinearOperator.templates.h:55-60,176-255,342-406): element matrices in
 * FP64 on the device, gathered into the CSR values in ascending element order
 * (no atomics).  coef is used by AFE_OP_VARSTIFF only. */
int afe_matrix_create(afe_space *s, afe_matrix **out);
/* a square matrix given as the reference's CSR (SparseMatrix<double> over a
 * compressed SparsityPattern: diagonal first in each row), for solvers on
 * matrices assembled elsewhere; val may be NULL (zeros) */
int afe_matrix_create_csr(afe_ctx *ctx, int64_t m, const int64_t *row_ptr,
                          const int64_t *col, const double *val,
                          afe_matrix **out);
/* BilinearOperator(sp0, sp1) with two FE spaces on one mesh (the same-mesh
 * branch of buildSparsityPattern / buildDofInfo, BilinearOperator.templates.h:
 * 64-172): rows are the DOFs of s0, columns those of s1; element e couples
 * each s0 DOF of e with each s1 DOF of e; rows sorted ascending (diagonal
 * first when the pattern is square, as SparsityPattern.h:39-70 stores it).
 * The spaces must share context, elements and coordinates.  afe_assemble
 * takes AFE_OP_MASS only (L2InnerProduct, BilinearOperator.templates.h:
 * 300-326: Jxw * phi0_j * phi1_k on the quadrature of element 0);
 * afe_assemble_elements takes n_elem x n_loc0 x n_loc1 element matrices;
 * vmult / tvmult / values work as on a square matrix.  Dirichlet
 * elimination and AMG reject it. */
int afe_matrix_create_pair(afe_space *s0, afe_space *s1, afe_matrix **out);
/* rows, columns and stored entries; max_coupling = nMaxCouplingDof
 * (BilinearOperator.templates.h:169-170) of a two-space matrix, -1 else */
int afe_matrix_info(const afe_matrix *a, int64_t *m, int64_t *n, int64_t *nnz,
                    int32_t *max_coupling);
/* the compressed pattern (row_ptr: m + 1 entries, col: nnz) */
int afe_matrix_pattern(const afe_matrix *a, int64_t *row_ptr, int64_t *col);
int afe_matrix_destroy(afe_matrix *a);
int afe_assemble(afe_matrix *a, int32_t op, int32_t acc, const afe_field *coef);
/* user-supplied element matrices (a getElementMatrix override evaluated by
 * the caller): h_elem_mat is n_elem x n_loc x n_loc row-major on the host
 * (n_loc0 x n_loc1 for a two-space matrix); the CSR gather runs on the
 * device. */
int afe_assemble_elements(afe_matrix *a, const double *h_elem_mat);
int afe_matrix_values(const afe_matrix *a, double *h_val);     /* D2H copy */
int afe_matrix_set_values(afe_matrix *a, const double *h_val); /* H2D copy */
double *afe_matrix_device_values(afe_matrix *a);
/* SparseMatrix::vmult (SparseMatrix.h:100-114): d_y = A d_x (m entries),
 * summed in stored order as the reference (bitwise its result) */
int afe_matrix_vmult(afe_matrix *a, const double *d_x, double *d_y);
/* SparseMatrix::Tvmult (SparseMatrix.h:116-132): d_y = A^T d_x (n entries),
 * each entry summed in ascending row order as the reference's scatter */
int afe_matrix_tvmult(afe_matrix *a, const double *d_x, double *d_y);

/* ---- right-hand side and Dirichlet elimination -------------------------- */
/* Operator::L2Discretize (Operator.discretize.templates.h:80-152) into d_f. */
int afe_rhs(afe_space *s, int32_t acc, int32_t fkind, const afe_field *coef,
            const double *d_nodal, double *d_f);
/* BoundaryConditionAdmin::apply, preserve_symmetry = true, zero Dirichlet
 * data on every DOF with a nonzero boundary mark
 * (FEMSpace.templates.h:1338-1373).  d_u is not touched (as the reference). */
int afe_apply_dirichlet(afe_matrix *a, double *d_u, double *d_f);

/* Callbacks of the reference (f, boundary data, exact solutions) are host C
 * functions; these entry points take their values and keep the assembly on
 * the device.  afe_space_qpoints: physical quadrature points (n_elem x n_qp x
 * dim) and Jxw (n_elem x n_qp) of rule acc; either output may be NULL, *n_qp
 * receives the rule size. */
int afe_space_qpoints(afe_space *s, int32_t acc, int64_t *n_qp, double *h_pts,
                      double *h_jxw);
/* Operator::L2Discretize(f, space, rhs, acc) (Operator.discretize.templates.h:
 * 121-152) with f given at the quadrature points (n_elem x n_qp) */
int afe_rhs_qp(afe_space *s, int32_t acc, const double *h_fq, double *d_f);
/* BoundaryConditionAdmin::apply(A, u, f) (FEMSpace.templates.h:1338-1373) for
 * the DOFs whose boundary mark is in marks[0..n_marks), Dirichlet data h_g
 * (one value per DOF, used on those DOFs); same elimination arithmetic */
int afe_apply_dirichlet_values(afe_matrix *a, const int32_t *marks,
                               int32_t n_marks, const double *h_g, double *d_u,
                               double *d_f);
/* BoundaryConditionAdmin::apply(A, u, f, preserve_symmetry)
 * (FEMSpace.templates.h:1338-1373) with the data of afe_apply_dirichlet_values:
 * preserve_symmetry = 1 is afe_apply_dirichlet_values; 0 sets the flagged
 * rows (f_i = A_ii g_i, off-diagonals zero) and leaves every other row and
 * its right-hand side as they are. */
int afe_apply_dirichlet_bc(afe_matrix *a, const int32_t *marks, int32_t n_marks,
                           const double *h_g, double *d_u, double *d_f,
                           int32_t preserve_symmetry);
/* Functional::L2Error (h1 = 0, exact values n_elem x n_qp) and H1SemiError
 * (h1 = 1, exact gradients n_elem x n_qp x dim) */
int afe_error_qp(afe_space *s, const double *d_uh, int32_t acc, int32_t h1,
                 const double *h_exact, double *err);

/* The same functional with everything on the device, allocating nothing:
 * no counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  d_exact holds the exact values (gradients for h1 = 1) at the
 * quadrature points as afe_error_qp's h_exact, d_err receives the error (one
 * double, device memory).  The reduction's partial sums live in a
 * file-scope device array of 1024 doubles, one per GPU, held until the
 * process ends; the array's size does not depend on the mesh, so no mesh is
 * too large for it.  Every partial sum is overwritten by each call (nothing
 * to clear); calls that overlap on one GPU race on the array.  Returns with
 * the context stream synchronised; the result equals afe_error_qp's bit for
 * bit (same kernels). */
int afe_error_qp_noalloc(afe_space *s, const double *d_uh, int32_t acc,
                         int32_t h1, const double *d_exact, double *d_err);

/* ---- error functionals -------------------------------------------------- */
/* Functional::L2Error / H1SemiError (Functional.templates.h:296-317,399-426)
 * of the FE function d_uh against u = prod sin(pi x_d). */
int afe_error_l2(afe_space *s, const double *d_uh, int32_t acc, double *err);
int afe_error_h1(afe_space *s, const double *d_uh, int32_t acc, double *err);

/* ---- AMG ---------------------------------------------------------------- */
/* AMGSolver(M, tol, s, nmpo, nmpsd, alpha) / reinit (src/AMGSolver.cpp:384-438,
 * 683-699) or lazyReinit on a fresh solver (:442-496): the hierarchy is built
 * on the host (Project / lazyProject + getPMPt, identical patterns), coloured,
 * converted to sliced-ELL and uploaded once.  As in the reference, the fine
 * matrix is borrowed (projected_matrix[0] = &M): it must outlive the
 * hierarchy or its next afe_amg_lazy_reinit. */
int afe_amg_create(afe_matrix *a, const afe_amg_params *p, afe_amg **out);
int afe_amg_destroy(afe_amg *g);
/* host-only hierarchy of a CSR matrix (the coarsening half of
 * AMGSolver::reinit / lazyInit, src/AMGSolver.cpp:384-496) without a device;
 * inspect it with afe_amg_info / afe_amg_level_csr / afe_amg_coarse_inverse */
int afe_amg_create_host(int64_t m, const int64_t *row_ptr, const int64_t *col,
                        const double *val, const afe_amg_params *p,
                        afe_amg **out);
/* lazyReinit of a host-only lazy hierarchy with new values on the same
 * pattern (src/AMGSolver.cpp:505-545) */
int afe_amg_lazy_reinit_host(afe_amg *g, const double *val);
/* the coarsest dense inverse, row-major n_c x n_c (DenseMatrix::gauss_jordan
 * of AMGSolver::reinit, src/AMGSolver.cpp:423-435) */
int afe_amg_coarse_inverse(const afe_amg *g, double *minv);
/* AMGSolver::lazyReinit (src/AMGSolver.cpp:505-545): same pattern object
 * required (AFE_E_PATTERN where the reference aborts); P is kept and the
 * Galerkin products are recomputed.  The pattern object is the afe_space of
 * a space matrix (any matrix of that space qualifies, so a re-assembled
 * matrix may replace the destroyed one the hierarchy was built on) and the
 * afe_matrix itself for a matrix made by afe_matrix_create_csr. */
int afe_amg_lazy_reinit(afe_amg *g, afe_matrix *a);
/* Structure sharing (no counterpart in the reference): the numbering,
 * pattern and gather map of a space, and the interpolation, colouring and
 * device layouts of a lazy hierarchy, depend only on the mesh connectivity.
 * afe_space_clone copies a space for a mesh of the same connectivity with
 * new vertex coordinates (NULL: the same ones); afe_amg_clone copies a lazy
 * hierarchy for matrix a, which must have its pattern (AFE_E_PATTERN
 * otherwise), and computes a's Galerkin products on the device -- the result
 * equals afe_amg_create(a, lazy) bit for bit, without the host coarsening,
 * colouring and layout.  The copies are independent of their sources. */
int afe_space_clone(const afe_space *src, const double *coords,
                    afe_space **out);
int afe_amg_clone(const afe_amg *src, afe_matrix *a, afe_amg **out);
int afe_amg_info(const afe_amg *g, int32_t *n_levels, int64_t *level_n,
                 int64_t *level_nnz, int32_t *coarse_full);
/* host copies of a level of the hierarchy (reference patterns, including the
 * explicit zeros): which = 0 -> A_l, 1 -> P_l; NULL pointers skip arrays */
int afe_amg_level_csr(const afe_amg *g, int32_t level, int32_t which,
                      int64_t *m, int64_t *n, int64_t *nnz, int64_t *row_ptr,
                      int64_t *col, double *val);
/* the multicolour Gauss-Seidel colouring of a device level: colour[i] of
 * every row i of A_l (n_l entries).  The sweeps visit the colours in
 * increasing order (the symmetric cycle's post-smoother in decreasing
 * order); no two rows coupled by a stored entry share a colour.  The
 * colouring replaces the row order of AMGSolver::GaussSidel
 * (src/AMGSolver.cpp:362-381); no counterpart in the reference's interface. */
int afe_amg_level_colours(const afe_amg *g, int32_t level, int32_t *colour);
/* AMGSolver::solve(x, r, tol, step, mode) (src/AMGSolver.cpp:549-639):
 * exactly step+1 V(s,s) cycles from the initial guess d_x, FP64, multicolour
 * Gauss-Seidel in forward colour order on both sides.  If rtol > 0 the
 * cycles stop early once ||b-Ax||_2/||b||_2 <= rtol (benchmark mode). */
int afe_amg_solve(afe_amg *g, double *d_x, const double *d_r, double tol,
                  int64_t step, int32_t mode, double rtol,
                  afe_solve_stats *st);
/* AMGPreconditioner::vmult (AMGSolver.h:193-195) = solve(x, b, 0, is, 1) */
int afe_amg_precondition(afe_amg *g, double *d_x, const double *d_b,
                         int32_t iterate_step);
/* FP64 preconditioned CG with one symmetric V-cycle (forward colours before,
 * reverse colours after the coarse correction) per iteration, in FP64 or in
 * FP32 (prec); stops at ||b-Ax||_2/||b||_2 <= rtol (true residual). */
int afe_amg_pcg(afe_amg *g, double *d_x, const double *d_b, double rtol,
                int32_t max_it, int32_t prec, afe_solve_stats *st);
/* one symmetric V-cycle z = M^{-1} r from zero (prec as above) */
int afe_amg_vcycle(afe_amg *g, double *d_z, const double *d_r, int32_t prec);
/* AMGSolver::smoothStep() = s (AMGSolver.h:98-99) after construction */
int afe_amg_set_smooth(afe_amg *g, int32_t smooth);
/* AMGSolver::isSolveMostProjectExactly() = exact (AMGSolver.h:100-101):
 * exact = 1 (the default) solves the coarsest level with the dense inverse
 * (or by diagonal scaling when the coarsening stalled); exact = 0 runs
 * smoothStep() Gauss-Seidel sweeps from zero there instead
 * (src/AMGSolver.cpp:590-605), multicolour in forward colour order. */
int afe_amg_set_coarse_exact(afe_amg *g, int32_t exact);

/* ---- batched solves (no counterpart in the reference) --------------------
 * Continuous batching of afe_amg_pcg solves on hierarchies of one context:
 * every afe_batch_step runs one CG iteration of every active member with one
 * launch per step kind and level/colour covering all members.  A member's
 * arithmetic does not depend on the other members: its result is bitwise
 * the same alone and in any mix.  Members may be added between steps; a
 * member finishes when its delivered vector reaches rtol (as afe_amg_pcg) or
 * max_it; its solution is then in its d_x.  The V-cycle uses `smooth`
 * sweeps per side for every member (the member hierarchy's smoothing count
 * is set to it).  One active solve per hierarchy at a time. */
/* A member's hierarchy and vectors must stay alive until the member is
 * removed or the batch destroyed.  The member's CG state lives in its
 * hierarchy's work vectors: no afe_amg_pcg / _solve / _vcycle /
 * _precondition on that hierarchy while it is an active member. */
typedef struct afe_batch afe_batch;
int afe_batch_create(afe_ctx *ctx, int32_t prec, int32_t smooth,
                     afe_batch **out);
int afe_batch_destroy(afe_batch *b);
/* add a member: its b and x are copied and its residual issued; its first
 * preconditioning runs inside the next afe_batch_step (not counted as an
 * iteration), together with the other members */
int afe_batch_add(afe_batch *b, afe_amg *g, double *d_x, const double *d_b,
                  double rtol, int32_t max_it, int32_t *slot);
/* one batched CG iteration; *n_active = members still iterating */
int afe_batch_step(afe_batch *b, int32_t *n_active);
/* steps until no member is active */
int afe_batch_run(afe_batch *b);
/* *done, stats (seconds = time from afe_batch_add to completion); returns
 * the member's status once done (AFE_E_NOTCONVERGED if it missed rtol) */
int afe_batch_result(afe_batch *b, int32_t slot, int32_t *done,
                     afe_solve_stats *st);
int afe_batch_remove(afe_batch *b, int32_t slot);
/* launch shape of the batched V-cycle level kernels, by level bucket: slices
 * per wavefront on level 0 (tile_level0) and on the coarser levels run as
 * launches (tile_coarse), each 1, 2, 4 or 8; chosen by timing (the C4
 * autotuner).  Results do not depend on it. */
int afe_batch_set_tiles(afe_batch *b, int32_t tile_level0, int32_t tile_coarse);

/* ---- solve server (no counterpart in the reference) ----------------------
 * One device, a pool of host set-up threads and one continuously running
 * batch.  A request names a structured mesh (n x n squares split in two
 * triangles, or n^3 cubes split in six tetrahedra, interior jitter and
 * seed) or brings a mesh of its own, under a cache key; a later request with
 * the same key re-solves that mesh after a moving-mesh update (structured:
 * interior vertices displaced by move_amp * h * sin(2 pi move_k x_d) from the
 * base mesh; own mesh: the request's coordinates): the pattern, gather map
 * and lazy hierarchy are kept, only values are reassembled and
 * AMGSolver::lazyReinit refreshes the Galerkin products.  A new mesh whose
 * connectivity equals a cached one's copies that entry's structure
 * (afe_space_clone, afe_amg_clone) instead of numbering and coarsening on
 * the host.  Requests are set up concurrently, admitted into the batch as
 * slots free up, and solved with FP64 CG + FP32 V-cycle to rtol; cached
 * meshes stay in HBM within hbm_budget bytes (idle ones evicted, least
 * recently used first); a request that cannot fit even with everything else
 * evicted fails with AFE_E_BUDGET.  A request's result does not depend on
 * the requests solved with it, nor on the entry its structure came from. */
typedef struct afe_request {
  int64_t mesh_key;  /* cache key of the mesh                              */
  int32_t kind;      /* AFE_P1_TRI, AFE_P2_TRI, AFE_P1_TET                  */
  int32_t n;         /* squares (cubes) per side                            */
  double jitter;     /* interior jitter of the base mesh, fraction of h    */
  uint64_t seed;     /* base-mesh seed                                     */
  double move_amp;   /* moving-mesh displacement amplitude / h (0: none)   */
  int32_t move_k;    /* its wave number                                    */
  int32_t acc;       /* quadrature accuracy                                */
  int32_t op;        /* AFE_OP_STIFF or AFE_OP_VARSTIFF                     */
  int32_t fkind;     /* AFE_F_POISSON, AFE_F_VARCOEF or AFE_F_ONE           */
  afe_field coef;    /* coefficient field (copied at submit)               */
  double rtol;       /* relative residual target                           */
  int32_t max_it;    /* CG iteration limit                                 */
  int32_t keep_x;    /* 1: keep the solution on the host for afe_server_wait */
  /* a mesh of the caller's instead of a structured one (coords != NULL):
   * n_vert vertices (coords, vert_mark) and n_elem elements of `kind`
   * (elem_vert), copied at submit; n, jitter, seed and move_* are unused.
   * A later request with the same mesh_key and the same elements and marks
   * re-solves that mesh on its new coordinates (values only). */
  int64_t n_vert;
  const double *coords;
  int64_t n_elem;
  const int32_t *elem_vert;
  const int32_t *vert_mark;
} afe_request;
typedef struct afe_server afe_server;
int afe_server_create(int32_t device, int64_t hbm_budget, int32_t max_slots,
                      int32_t setup_threads, int32_t smooth,
                      afe_server **out);
int afe_server_destroy(afe_server *s);
int afe_server_submit(afe_server *s, const afe_request *r, int64_t *id);
int afe_server_poll(afe_server *s, int64_t id, int32_t *done);
/* blocks until the request is done; latency = completion - submission;
 * h_x (n_dof doubles, may be NULL) receives the solution if keep_x */
int afe_server_wait(afe_server *s, int64_t id, double *h_x, int64_t *n_dof,
                    afe_solve_stats *st, double *latency_s);
int afe_server_release(afe_server *s, int64_t id);
int afe_server_info(afe_server *s, int64_t *device_bytes, int64_t *cached,
                    int64_t *completed, int64_t *evictions);
/* sums over finished requests (seconds): out[0] set-up of new meshes,
 * out[1] set-up of moving-mesh re-solves, out[2] waiting for a set-up
 * thread, out[3] waiting for a batch slot, out[4] solving; out[5], out[6]
 * the counts of new and moving-mesh requests */
int afe_server_timing(afe_server *s, double *out);
/* the same sums, the first n_out of: the seven above, then out[7] the set-up
 * of new meshes that copied a cached entry's structure (part of out[0]),
 * out[8] their count, out[9] waiting for HBM room before set-ups (a set-up
 * thread holds the request meanwhile), out[10] and out[11] the set-up
 * threads' user and system CPU seconds during set-ups, out[12] their waits
 * for the context lock during set-ups */
int afe_server_timing_n(afe_server *s, double *out, int32_t n_out);

/* ---- measurement hooks (no counterpart in the reference) ------------------ */
/* average duration of `reps` back-to-back launches of a level-0 kernel as the
 * solvers launch it, timed with HIP events on the context stream; which:
 * 0 FP64 SpMV (CG q = A p), 1 FP32 Gauss-Seidel sweep (all colours),
 * 2 FP64 sweep, 3 FP32 residual.  *bytes = algorithmic bytes per launch. */
int afe_amg_kernel_timing(afe_amg *g, int32_t which, int32_t reps,
                          double *avg_ms, double *bytes);
/* afe_assemble(a, op, acc, coef) `reps` times: the average durations of its
 * two kernels (element matrices, CSR gather) from HIP events on the context
 * stream, and its algorithmic bytes per assembly (vertex coordinates,
 * element vertex and DOF indices read, CSR values written) */
int afe_assemble_kernel_timing(afe_matrix *a, int32_t op, int32_t acc,
                               const afe_field *coef, int32_t reps,
                               double *elem_ms, double *gather_ms,
                               double *bytes);
/* afe_assemble(a, op, acc, coef) `reps` times as afe_assemble runs it (the
 * element chunks and the gathers of the rows they complete overlapping on
 * two streams): *ms = the average device time from a HIP event before the
 * first kernel to one after the last gather, on the context stream */
int afe_assemble_timing(afe_matrix *a, int32_t op, int32_t acc,
                        const afe_field *coef, int32_t reps, double *ms);
/* live timing of the level-0 FP64 SpMV inside afe_amg_pcg (a start/stop
 * HIP event pair attached to that kernel's dispatch on the solver stream in
 * every CG iteration; the iterate is unchanged): enable = 1 resets and
 * starts, 0 stops; *avg_ms / *launches = the average so far */
int afe_amg_live_timing(afe_amg *g, int32_t enable, double *avg_ms,
                        int64_t *launches);
/* graphs-on timing of the two captured parts of every CG iteration of
 * afe_amg_pcg: part 0 (SpMV with its dot, x/r update, scalars) and part 1
 * (the preconditioner V-cycle, z, beta, p), from HIP events recorded between
 * the graph launches; enable = 1 resets and starts, 0 stops; the averages
 * so far and the iterations timed are returned */
int afe_amg_part_timing(afe_amg *g, int32_t enable, double *part0_ms,
                        double *part1_ms, int64_t *iterations);
/* average duration of `reps` launches of a plain streaming copy of `bytes`
 * (read + write, 16-byte vectors): calibrates the HBM byte counters and
 * gives the achievable copy bandwidth next to the solver kernels */
int afe_copy_kernel_timing(afe_ctx *ctx, int64_t bytes, int32_t reps,
                           double *avg_ms);
/* launch tuning of the sliced-ELL kernels (process-wide, set by the
 * autotuner): xcd = 0 plain workgroup order; 1 orders workgroups so that
 * each XCD works on a contiguous slice range; 2 interleaves the slices of the
 * grid-stride level-0 kernels (neighbouring slices read at the same time);
 * spmv_blocks = grid of the grid-stride level-0 kernels (multiple of 8, at
 * most 16384); tail_rows = levels of at most this many rows run in the
 * single-workgroup coarse tail.  Applies to hierarchies built afterwards. */
int afe_set_tuning(int32_t xcd, int32_t spmv_blocks, int64_t tail_rows);
/* move the boundary of the single-workgroup coarse tail of one hierarchy:
 * levels of at most `rows` rows run in the tail (batches of small problems
 * run one tail workgroup per member in parallel) */
int afe_amg_set_tail_rows(afe_amg *g, int64_t rows);
/* device layout of a level: rows, colours, stored sliced-ELL entries, the
 * first level handled by the single-workgroup coarse tail, and the lanes per
 * row of the level matrix's kernels (1 on level 0; up to 16 on small levels
 * of long rows, whose row sums then add per-lane partials) */
int afe_amg_device_info(const afe_amg *g, int32_t level, int64_t *rows,
                        int32_t *colours, int64_t *stored, int32_t *tail_level,
                        int32_t *row_lanes);

#ifdef __cplusplus
}
#endif
#endif

/*
 * afe_oracle.h -- CPU oracle for AFEPack's assemble-and-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it restates, in plain
 * C, the algorithms of the reference files cited at each function
 * (BilinearOperator::build/getElementMatrix, FEMSpace::buildDof,
 * SparsityPattern/SparseMatrix, Operator::L2Discretize,
 * BoundaryConditionAdmin::apply, AMGSolver::{Project, lazyProject, getPMPt,
 * reinit, lazyReinit, GaussSidel, solve}, Functional::{L2Error, H1SemiError}).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library (afepack_amd) never links or calls it.
 *
 * Index width follows the reference: size_t row pointers / column numbers,
 * stored here as int64_t.
 *
 * Parity pinning (how each part is checked before it is trusted) is listed in
 * tests/test_oracle_pinning.py; every entry point below has at least one check
 * there that does not restate its own formula.
 */
#ifndef AFE_ORACLE_H
#define AFE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_P1_TRI = 1, OR_P2_TRI = 2, OR_P1_TET = 3 };
enum { OR_OP_STIFF = 0, OR_OP_MASS = 1, OR_OP_VARSTIFF = 2 };
/* right-hand sides for L2Discretize */
enum {
  OR_F_POISSON = 0, /* f = d*pi^2 * prod sin(pi x_i)  (u = prod sin(pi x_i))   */
  OR_F_VARCOEF = 1, /* f = -div(a grad u), a = exp(g), u = prod sin(pi x_i)    */
  OR_F_NODAL = 2,   /* f = finite-element interpolant of given DOF values      */
  OR_F_ONE = 3      /* f = 1                                                   */
};
/* exact solutions for the error functionals */
enum { OR_U_SINPROD = 0, OR_U_ZERO = 1 };

/* Seeded Fourier field g(x) = scale * sum_m w_m sin(2 pi k_m . x + phi_m);
 * the coefficient is a(x) = exp(g(x)). k is n_modes x dim (row-major). */
typedef struct or_field {
  int n_modes;
  int dim;
  const double *k;
  const double *phi;
  const double *w;
  double scale;
} or_field;

typedef struct or_space or_space;
typedef struct or_mat or_mat;
typedef struct or_amg or_amg;

/* ---- FE space: DOF numbering (FEMSpace::buildDof) ---------------------- */
or_space *or_space_new(int kind, int64_t n_vert, const double *coords,
                       int64_t n_elem, const int32_t *elem_vert,
                       const int32_t *vert_mark);
or_space *or_space_new_sides(int kind, int64_t n_vert, const double *coords,
                             int64_t n_elem, const int32_t *elem_vert,
                             const int32_t *vert_mark, int64_t n_side,
                             const int32_t *side_vert, const int32_t *side_mark);
void or_space_free(or_space *s);
int64_t or_space_n_dof(const or_space *s);
int or_space_n_loc(const or_space *s);
void or_space_dofs(const or_space *s, int32_t *elem_dof, int32_t *dof_mark,
                   double *dof_point);

/* ---- quadrature (QuadratureInfoAdmin::find) ----------------------------- */
int or_quad_n(int tdim, int acc);
void or_quad_get(int tdim, int acc, double *pts, double *w);

/* ---- template-element plugins (for pinning against the reference's .so) -- */
void or_basis_eval(int kind, const double *p, const double *verts, double *val,
                   double *grad);
void or_local_to_global(int tdim, const double *lp, const double *verts,
                        double *gp);
double or_l2g_jacobian(int tdim, const double *verts);
double or_template_volume(int tdim);

/* ---- element matrices and assembly -------------------------------------- */
void or_element_matrix(int kind, int op, int acc, const double *vert_coords,
                       const or_field *coef, double *K);
or_mat *or_build(const or_space *s, int op, int acc, const or_field *coef);
/* two spaces on one mesh: L2InnerProduct (op OR_OP_MASS) or pattern only (op < 0) */
void or_element_matrix_pair(int kind0, int kind1, int acc, const double *vert_coords, double *K);
or_mat *or_build_pair(const or_space *s0, const or_space *s1, int op, int acc, int64_t *max_coupling);

/* ---- CSR access ---------------------------------------------------------- */
or_mat *or_mat_from_csr(int64_t m, int64_t n, const int64_t *rp,
                        const int64_t *col, const double *val);
or_mat *or_mat_share_pattern(const or_mat *a, const double *val);
void or_mat_free(or_mat *a);
int64_t or_mat_m(const or_mat *a);
int64_t or_mat_n(const or_mat *a);
int64_t or_mat_nnz(const or_mat *a);
void or_mat_get(const or_mat *a, int64_t *rp, int64_t *col, double *val);
void or_mat_vmult(const or_mat *a, const double *x, double *y);
void or_mat_tvmult(const or_mat *a, const double *x, double *y);

/* ---- right-hand side, boundary conditions, error functionals ------------ */
void or_rhs(const or_space *s, int acc, int fkind, const or_field *coef,
            const double *nodal, double *f);
void or_apply_dirichlet(const or_space *s, or_mat *a, double *u, double *f);
void or_apply_dirichlet_values(const or_space *s, or_mat *a, const int32_t *marks, int n_marks,
                               const double *g, double *f);
void or_apply_dirichlet_bc(const or_space *s, or_mat *a, const int32_t *marks, int n_marks,
                           const double *g, double *f, int preserve_symmetry);
double or_l2_error(const or_space *s, const double *uh, int acc, int ukind);
double or_h1_error(const or_space *s, const double *uh, int acc, int ukind);

/* ---- AMG ------------------------------------------------------------------ */
or_amg *or_amg_new(const or_mat *a, int lazy, double alpha, unsigned nmpo,
                   double nmpsd);
int or_amg_lazy_reinit(or_amg *g, const or_mat *a);
void or_amg_free(or_amg *g);
int or_amg_n_project(const or_amg *g);
int or_amg_coarse_full(const or_amg *g);
const or_mat *or_amg_A(const or_amg *g, int level);
const or_mat *or_amg_P(const or_amg *g, int level);
const or_mat *or_amg_Pt(const or_amg *g, int level);
void or_amg_coarse_inverse(const or_amg *g, double *minv);
/* exactly step+1 V-cycles; res_hist gets step+2 L1 residuals (initial first) */
void or_amg_solve(const or_amg *g, double *x, const double *r, unsigned smooth,
                  int64_t step, int solve_coarse_exactly, double *res_hist);
void or_gauss_seidel(const or_mat *a, double *x, const double *r, int s);
int or_dense_inverse(int n, double *a);

#ifdef __cplusplus
}
#endif
#endif

/*
 * cdx_oracle.h -- CPU restatement of CardilloCxx's Moreau contact solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in cardillocxx_amd/.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product path never does.
 *
 * Every function restates the reference algorithm from its source text
 * (citations are relative to the CardilloCxx source tree).  Input layout is
 * the same flat SoA layout the C ABI in include/cdx/cdx.h takes:
 *
 *   bodies   : nb, dof[nb] in {3,6}; stacked arrays of length nv = sum(dof):
 *              minv, v, fext, fgyro (velocity layout [v_world; omega_body]).
 *   contacts : nc, body_a/body_b (-1 = static side), mu (>0 => frictional,
 *              3 rows; else 1 row), frame[9] = n,t1,t2 (world),
 *              angA/angB[9] = r_body x d_body for d in {n,t1,t2} (unsigned),
 *              rest[3] per-row restitution, vbias[3] per-row static-side
 *              velocity term, warm[3] last impulse (normal >= 0).
 */
#ifndef CDX_ORACLE_H
#define CDX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_OK = 0,
    ORC_NOT_FACTORIZED = 1, /* singular S (zero mass on a dynamic dof) */
    ORC_NONFINITE = 2,
    ORC_DIVERGED = 3, /* PGS / colored GS: NaN or Inf residual */
    ORC_BAD_ARG = 4,
};

typedef struct {
    int nb;
    const int* dof;
    const double* minv;
    const double* v;
    const double* fext;
    const double* fgyro;
} orc_bodies;

typedef struct {
    int nc;
    const int* body_a;
    const int* body_b;
    const double* mu;
    const double* frame;
    const double* angA;
    const double* angB;
    const double* rest;
    const double* vbias;
    const double* warm;
} orc_contacts;

typedef struct {
    double alpha;      /* pj.alpha */
    double relaxation; /* pj.relaxation (PGS / colored GS only) */
    int max_iter;      /* pj.max_iterations */
    double tol_abs;    /* pj.tol_abs */
    double tol_rel;    /* pj.tol_rel */
    int warmstart;     /* pj.warmstart */
    int nesterov;      /* pj.nesterov */
    double nesterov_beta_threshold;
    int nesterov_restart_limit;
    int chebyshev;     /* pj.chebyshev (ignored when nesterov is set, projected_jacobi.cpp:469) */
    int block_precond; /* pj.rdiag_true_delassus: exact per-contact Delassus blocks instead of the row scaling */
    int anderson;      /* pj.anderson (ignored when nesterov or chebyshev is set, projected_jacobi.cpp:470-472) */
    int anderson_m;    /* pj.anderson_m: history window (>= 1) */
    int implicit_gyro; /* moreau.implicit_gyroscopy: PJ solves with S_rot = diag(I) - dt G_rot per rigid
                          body and leaves dt f_gyro out of the right-hand side; PGS only leaves it out
                          (refused by orc_pj_solve_joints) */
} orc_params;

typedef struct {
    int iters;
    double residual; /* last residual evaluated (NaN if none) */
    int nrows;
    int ncolors; /* colored GS only */
    double rho;  /* Chebyshev loop: the spectral-radius estimate it used (NaN otherwise) */
} orc_result;

/* Sparse contact Jacobian W in CSR, rows ordered frictionless-then-frictional. */
typedef struct {
    int nrows, nv, nnfc;
    int* rowptr;
    int* col;
    double* val;
    double* mu;    /* per row */
    double* rest;  /* per row */
    double* cvec;  /* per row static-side velocity term */
    int* row_base; /* per contact, -1 if static-static */
    int* row_size; /* per contact: 1 or 3 (0 if skipped) */
    int* voff;     /* per body velocity offset, nb+1 entries */
} orc_W;

int orc_build_W(const orc_bodies* b, const orc_contacts* c, orc_W* out);
void orc_free_W(orc_W* w);

void orc_tangent_frame(const double n_in[3], double t1[3], double t2[3]);
void orc_cross(const double a[3], const double b[3], double out[3]);

/* Projected Jacobi: v_out[nv], imp_out[3*nc] (canonical normal >= 0). */
/* min ||A g - b||_2 for A (rows x cols, column-major, overwritten by its factorisation): Householder QR
 * with column pivoting and Eigen's pivot count (ColPivHouseholderQR::compute / solve); returns the
 * number of pivots used (columns beyond it get g = 0). */
int orc_lstsq_colpiv(int rows, int cols, double* A, const double* b, double* g);

int orc_pj_solve(const orc_bodies* b, const orc_contacts* c, double dt, double theta, const orc_params* prm,
                 int nthreads, double* v_out, double* imp_out, orc_result* res);

/* Bilateral rows of the extended system, as DynamicsAssembler::rebuildInteractionW_ emits them
 * (dynamics_assembler.cpp:341-439): ns spring rows (Wg) then nd damper rows (Wgamma).  Row r
 * couples body_a[r] and body_b[r] (-1: static, its velocity already folded into vsrc) with the
 * coefficients wa[6r..6r+5] / wb[6r..6r+5] on the body's velocity dofs (first dof entries used);
 * comp[r] is C (spring) or A (damper), vsrc[r] C_v / A_v, gerr[r] the position error of a spring
 * row, lam_prev[r] Lambda_g / Lambda_gamma of the previous step. */
typedef struct {
    int ns, nd;
    const int* body_a;
    const int* body_b;
    const double* wa;
    const double* wb;
    const double* comp;
    const double* vsrc;
    const double* gerr;
    const double* lam_prev;
    double bias_factor; /* constraint_bias_factor (Baumgarte, pj_assembler.cpp:163-164) */
} orc_joints;

/* Projected Jacobi on the extended system S [v; Lg; Lgamma] (PjAssembler::buildAndFactorS / rhs /
 * solveS, pj_assembler.cpp:9-182, with implicit gyroscopy and lambda-theta off), standard loop.
 * S is factorised once by dense LU with partial pivoting (the reference uses SparseLU).
 * lam_out[ns+nd] receives the new Lambda_g / Lambda_gamma. */
int orc_pj_solve_joints(const orc_bodies* b, const orc_contacts* c, const orc_joints* jt, double dt, double theta,
                        const orc_params* prm, double* v_out, double* imp_out, double* lam_out, orc_result* res);

/* Projected Gauss-Seidel (sequential, lambda <= 0 internally). */
int orc_pgs_solve(const orc_bodies* b, const orc_contacts* c, double dt, double theta, const orc_params* prm,
                  double* v_out, double* imp_out, orc_result* res);

/* Greedy Welsh-Powell colouring of the shared-body block graph. color_of[nblk]. */
int orc_color_blocks(int nb, int nc, const int* body_a, const int* body_b, int* color_of);

/* Condensed-style colored Gauss-Seidel. */
int orc_colored_gs_solve(const orc_bodies* b, const orc_contacts* c, double dt, double theta, const orc_params* prm,
                         double* v_out, double* imp_out, orc_result* res);

/* Explicit position update q += h B(q) v for every body (stacked q: 3 or 7 per body). */
void orc_position_update(int nb, const int* dof, double* q, const double* v, double h);

/* Gyroscopic force -omega x (I omega) per rigid body; zero for point masses. */
void orc_gyro_force(int nb, const int* dof, const double* minv, const double* v, double* fgyro);

/* Complementarity / cone residual of a contact impulse against post-solve velocities. */
double orc_complementarity(const orc_bodies* b, const orc_contacts* c, const double* v, const double* imp);

#ifdef __cplusplus
}
#endif
#endif

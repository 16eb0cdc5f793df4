/*
 * cdx_oracle.c -- CPU restatement of CardilloCxx's Moreau contact solve.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP path.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Written from the reference source as text (no code copied); each function
 * cites the file:line it follows.  Arithmetic order mirrors the reference
 * where it is cheap to do so (row-major SpMV summed in column order, W^T
 * products accumulated row by row) so that the restatement is as close as
 * possible to Eigen's evaluation; bit identity is not claimed (Eigen's
 * SparseLU and vectorised reductions are not reproduced).
 *
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off -fopenmp).
 */
#include "cdx_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

void orc_cross(const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

/* Frisvad-style tangent frame with south-pole branch.
 * Follows src/collision/tangent_frame.hpp:23-50. */
void orc_tangent_frame(const double n_in[3], double t1[3], double t2[3]) {
    double n[3] = {n_in[0], n_in[1], n_in[2]};
    const double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (len <= 2.220446049250313e-16) { /* tangent_frame.hpp:26-31 */
        t1[0] = 1; t1[1] = 0; t1[2] = 0;
        t2[0] = 0; t2[1] = 1; t2[2] = 0;
        return;
    }
    n[0] /= len; n[1] /= len; n[2] /= len;
    if (n[2] < -0.9999999) { /* tangent_frame.hpp:34-41 */
        double t[3] = {0, -1, 0};
        const double d = n[0] * t[0] + n[1] * t[1] + n[2] * t[2];
        t[0] -= n[0] * d; t[1] -= n[1] * d; t[2] -= n[2] * d;
        const double tl = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        if (tl > 0) { t[0] /= tl; t[1] /= tl; t[2] /= tl; }
        t1[0] = t[0]; t1[1] = t[1]; t1[2] = t[2];
        orc_cross(n, t1, t2);
        return;
    }
    const double a = 1.0 / (1.0 + n[2]); /* tangent_frame.hpp:43-49 */
    const double b = -n[0] * n[1] * a;
    double t[3] = {1.0 - n[0] * n[0] * a, b, -n[0]};
    const double tl = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    if (tl > 0) { t[0] /= tl; t[1] /= tl; t[2] /= tl; }
    t1[0] = t[0]; t1[1] = t[1]; t1[2] = t[2];
    orc_cross(n, t1, t2);
}

/* ---------------------------------------------------------------------------------------- */
/* W assembly: src/physics/assembly/dynamics_assembler.cpp:211-332 and                     */
/* src/physics/assembly/contact_jacobian.hpp:17-23.                                         */
/* ---------------------------------------------------------------------------------------- */

typedef struct {
    int row, col;
    double val;
} trip;

static int trip_cmp(const void* pa, const void* pb) {
    const trip* a = (const trip*)pa;
    const trip* b = (const trip*)pb;
    if (a->row != b->row) return a->row < b->row ? -1 : 1;
    if (a->col != b->col) return a->col < b->col ? -1 : 1;
    return 0;
}

int orc_build_W(const orc_bodies* B, const orc_contacts* C, orc_W* W) {
    memset(W, 0, sizeof(*W));
    const int nb = B->nb, nc = C->nc;
    W->voff = (int*)malloc(sizeof(int) * (size_t)(nb + 1));
    int nv = 0;
    for (int b = 0; b < nb; ++b) {
        if (B->dof[b] != 3 && B->dof[b] != 6) return ORC_BAD_ARG;
        W->voff[b] = nv;
        nv += B->dof[b];
    }
    W->voff[nb] = nv;
    W->nv = nv;
    W->row_base = (int*)malloc(sizeof(int) * (size_t)(nc > 0 ? nc : 1));
    W->row_size = (int*)malloc(sizeof(int) * (size_t)(nc > 0 ? nc : 1));
    const size_t cap = (size_t)nc * 3 + 1;
    W->mu = (double*)calloc(cap, sizeof(double));
    W->rest = (double*)calloc(cap, sizeof(double));
    W->cvec = (double*)calloc(cap, sizeof(double));
    trip* trips = (trip*)malloc(sizeof(trip) * ((size_t)nc * 36 + 1));
    size_t nt = 0;
    int row = 0, nnfc = 0;
    for (int ci = 0; ci < nc; ++ci) { W->row_base[ci] = -1; W->row_size[ci] = 0; }

    /* Two passes: frictionless rows first, then frictional (dynamic_assembler.cpp:236-244). */
    for (int pass = 0; pass < 2; ++pass) {
        for (int ci = 0; ci < nc; ++ci) {
            const int frictional = C->mu[ci] > 0;
            if (pass == 0 && frictional) continue;
            if (pass == 1 && !frictional) continue;
            const int a = C->body_a[ci], bb = C->body_b[ci];
            if (a < -1 || a >= nb || bb < -1 || bb >= nb || (a >= 0 && a == bb)) { free(trips); return ORC_BAD_ARG; }
            if (a < 0 && bb < 0) continue; /* static-static skipped (:259-260) */
            const int nr = frictional ? 3 : 1;
            W->row_base[ci] = row;
            W->row_size[ci] = nr;
            for (int k = 0; k < nr; ++k) {
                const int r = row + k;
                const double* d = C->frame + 9 * (size_t)ci + 3 * k;
                for (int side = 0; side < 2; ++side) {
                    const int body = side == 0 ? a : bb;
                    const double s = side == 0 ? -1.0 : 1.0; /* A: -1, B: +1 (:295-296) */
                    if (body < 0) continue; /* static side goes to cvec via vbias */
                    const int col0 = W->voff[body];
                    const int dof = B->dof[body];
                    const double* ang = (side == 0 ? C->angA : C->angB) + 9 * (size_t)ci + 3 * k;
                    double rowv[6];
                    /* buildContactRowByDof: [s*d_world ; r_body x (s*d_body)] (contact_jacobian.hpp:17-23) */
                    for (int j = 0; j < 3; ++j) rowv[j] = s * d[j];
                    for (int j = 0; j < 3; ++j) rowv[3 + j] = s * ang[j];
                    for (int j = 0; j < dof; ++j) {
                        if (rowv[j] != 0.0) { /* zero entries dropped (:274-277) */
                            trips[nt].row = r; trips[nt].col = col0 + j; trips[nt].val = rowv[j];
                            ++nt;
                        }
                    }
                }
                W->cvec[r] = C->vbias[3 * (size_t)ci + k];
                W->rest[r] = C->rest[3 * (size_t)ci + k];
                W->mu[r] = frictional ? C->mu[ci] : 0.0; /* :301-312 */
            }
            row += nr;
            if (!frictional) ++nnfc;
        }
    }
    W->nrows = row;
    W->nnfc = nnfc;
    /* Compress like Eigen's setFromTriplets: sort by (row, col), sum duplicates. */
    qsort(trips, nt, sizeof(trip), trip_cmp);
    W->rowptr = (int*)calloc((size_t)row + 1, sizeof(int));
    W->col = (int*)malloc(sizeof(int) * (nt + 1));
    W->val = (double*)malloc(sizeof(double) * (nt + 1));
    size_t nnz = 0;
    for (size_t t = 0; t < nt; ++t) {
        if (nnz > 0 && t > 0 && trips[t].row == trips[t - 1].row && trips[t].col == trips[t - 1].col) {
            W->val[nnz - 1] += trips[t].val;
            continue;
        }
        W->col[nnz] = trips[t].col;
        W->val[nnz] = trips[t].val;
        W->rowptr[trips[t].row + 1]++;
        ++nnz;
    }
    for (int r = 0; r < row; ++r) W->rowptr[r + 1] += W->rowptr[r];
    free(trips);
    return ORC_OK;
}

void orc_free_W(orc_W* w) {
    free(w->rowptr); free(w->col); free(w->val); free(w->mu); free(w->rest); free(w->cvec);
    free(w->row_base); free(w->row_size); free(w->voff);
    memset(w, 0, sizeof(*w));
}

/* y = W x (row-wise, column order) */
static void spmv(const orc_W* W, const double* x, double* y) {
    for (int r = 0; r < W->nrows; ++r) {
        double s = 0;
        for (int k = W->rowptr[r]; k < W->rowptr[r + 1]; ++k) s += W->val[k] * x[W->col[k]];
        y[r] = s;
    }
}

/* y += W^T p, accumulated row by row (Eigen row-major transpose product). */
static void spmv_t_acc(const orc_W* W, const double* p, double* y) {
    for (int r = 0; r < W->nrows; ++r) {
        const double pr = p[r];
        for (int k = W->rowptr[r]; k < W->rowptr[r + 1]; ++k) y[W->col[k]] += W->val[k] * pr;
    }
}

/* Threaded variants for the fully threaded CPU baseline (OpenMP over rows of W and of W^T);
 * the reference itself runs its SpMV on one thread (examples/example_main.hpp:33). */
static void spmv_mt(const orc_W* W, const double* x, double* y, int nthreads) {
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int r = 0; r < W->nrows; ++r) {
        double s = 0;
        for (int k = W->rowptr[r]; k < W->rowptr[r + 1]; ++k) s += W->val[k] * x[W->col[k]];
        y[r] = s;
    }
}

/* W^T as CSR over the nv columns of W: Wt_ptr[j]..Wt_ptr[j+1] lists (row, entry index). */
typedef struct { int* ptr; int* row; int* ent; } orc_Wt;

static void build_Wt(const orc_W* W, orc_Wt* T) {
    const int nv = W->nv, nnz = W->rowptr[W->nrows];
    T->ptr = (int*)calloc((size_t)nv + 1, sizeof(int));
    T->row = (int*)malloc(sizeof(int) * (size_t)(nnz + 1));
    T->ent = (int*)malloc(sizeof(int) * (size_t)(nnz + 1));
    for (int k = 0; k < nnz; ++k) T->ptr[W->col[k] + 1]++;
    for (int j = 0; j < nv; ++j) T->ptr[j + 1] += T->ptr[j];
    int* fill = (int*)malloc(sizeof(int) * (size_t)(nv + 1));
    memcpy(fill, T->ptr, sizeof(int) * (size_t)nv);
    for (int r = 0; r < W->nrows; ++r)
        for (int k = W->rowptr[r]; k < W->rowptr[r + 1]; ++k) {
            const int at = fill[W->col[k]]++;
            T->row[at] = r;
            T->ent[at] = k;
        }
    free(fill);
}

/* y = W^T p, one column of W per iteration. */
static void spmv_t_mt(const orc_W* W, const orc_Wt* T, const double* p, double* y, int nthreads) {
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int j = 0; j < W->nv; ++j) {
        double s = 0;
        for (int k = T->ptr[j]; k < T->ptr[j + 1]; ++k) s += W->val[T->ent[k]] * p[T->row[k]];
        y[j] = s;
    }
}

/* D_i = sum_j W_ij^2 Minv_j (projected_jacobi.cpp:57-69; pgs_assembler.cpp:226-238). */
static void delassus_diag(const orc_W* W, const double* minv, double* D) {
    for (int r = 0; r < W->nrows; ++r) {
        double s = 0;
        for (int k = W->rowptr[r]; k < W->rowptr[r + 1]; ++k) {
            const double w = W->val[k];
            s += w * w * minv[W->col[k]];
        }
        D[r] = s;
    }
}

/* sum_j W_rj W_sj M^-1_j over the common columns of rows r and s, in column order */
static double row_pair_dot(const orc_W* W, int r, int s, const double* minv) {
    double acc = 0;
    int a = W->rowptr[r], b = W->rowptr[s];
    const int ae = W->rowptr[r + 1], be = W->rowptr[s + 1];
    while (a < ae && b < be) {
        if (W->col[a] < W->col[b]) ++a;
        else if (W->col[b] < W->col[a]) ++b;
        else { acc += W->val[a] * W->val[b] * minv[W->col[a]]; ++a; ++b; }
    }
    return acc;
}

/* Inverse of a small SPD block by Cholesky (Eigen::LLT::solve of the identity, invertSmallSpd,
 * block_diagonal.hpp:15-27): A = L L^T, then L y = e_i and L^T x = y per column.  Returns 0, or 1
 * when a pivot is not positive (the LLT failure the reference then falls back from). */
static int llt_inverse(int n, const double* A, double* Ainv) {
    double L[9] = {0};
    for (int k = 0; k < n; ++k) {
        double d = A[k * n + k];
        for (int j = 0; j < k; ++j) d -= L[k * 3 + j] * L[k * 3 + j];
        if (!(d > 0)) return 1;
        L[k * 3 + k] = sqrt(d);
        for (int i = k + 1; i < n; ++i) {
            double t = A[i * n + k];
            for (int j = 0; j < k; ++j) t -= L[i * 3 + j] * L[k * 3 + j];
            L[i * 3 + k] = t / L[k * 3 + k];
        }
    }
    for (int col = 0; col < n; ++col) {
        double y[3], x[3];
        for (int i = 0; i < n; ++i) {
            double t = (i == col) ? 1.0 : 0.0;
            for (int j = 0; j < i; ++j) t -= L[i * 3 + j] * y[j];
            y[i] = t / L[i * 3 + i];
        }
        for (int i = n - 1; i >= 0; --i) {
            double t = y[i];
            for (int j = i + 1; j < n; ++j) t -= L[j * 3 + i] * x[j];
            x[i] = t / L[i * 3 + i];
        }
        for (int i = 0; i < n; ++i) Ainv[i * n + col] = x[i];
    }
    return 0;
}

/* Exact per-contact Delassus block preconditioner (pj.rdiag_true_delassus:
 * buildBlockPreconditioner, projected_jacobi.cpp:71-98, with BlockDiagonal::calculateInverse,
 * block_diagonal.hpp:81-93): the block (W_c M^-1 W_c^T) / alpha of each contact (1x1 for a
 * frictionless row, 3x3 for a frictional contact's rows) is inverted -- elementwise when every
 * block is diagonal, otherwise each by Cholesky, falling back to the inverse of its diagonal when
 * the factorisation fails.  (The reference tries LDLT in between; a block with a dynamic side is
 * >= m^-1 I on its orthonormal frame rows, so the Cholesky step does not fail for contacts.)
 * Rb holds, per row, that row of its block inverse: 3 entries (unused ones 0). */
static void build_rblock(const orc_W* W, const double* minv, double alpha, double* Rb) {
    const int nr = W->nrows, nnfc = W->nnfc;
    const double inv_alpha = alpha > 0 ? 1.0 / alpha : 0.0;
    double* blk = (double*)malloc(sizeof(double) * 3 * (size_t)(nr + 1));
    int is_diag = 1;
    for (int r = 0; r < nnfc; ++r) blk[3 * r] = row_pair_dot(W, r, r, minv) * inv_alpha;
    for (int r = nnfc; r < nr; r += 3)
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < 3; ++j) {
                const double v = row_pair_dot(W, r + k, r + j, minv) * inv_alpha;
                blk[3 * (r + k) + j] = v;
                if (k != j && v != 0.0) is_diag = 0;
            }
    memset(Rb, 0, sizeof(double) * 3 * (size_t)nr);
    for (int r = 0; r < nnfc; ++r) {
        double inv = 1.0 / blk[3 * r];
        if (!is_diag && llt_inverse(1, &blk[3 * r], &inv)) inv = 1.0 / blk[3 * r];
        Rb[3 * r] = inv;
    }
    for (int r = nnfc; r < nr; r += 3) {
        double A[9], Ai[9];
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < 3; ++j) A[k * 3 + j] = blk[3 * (r + k) + j];
        if (is_diag || llt_inverse(3, A, Ai)) {
            memset(Ai, 0, sizeof(Ai));
            for (int k = 0; k < 3; ++k) Ai[k * 3 + k] = 1.0 / A[k * 3 + k];
        }
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < 3; ++j) Rb[3 * (r + k) + j] = Ai[k * 3 + j];
    }
    free(blk);
}

/* out = R_block in (BlockDiagonal::applyTo, block_diagonal.hpp:103-124); out may alias in */
static void apply_rblock(const orc_W* W, const double* Rb, double* v) {
    for (int r = 0; r < W->nnfc; ++r) v[r] = Rb[3 * r] * v[r];
    for (int r = W->nnfc; r < W->nrows; r += 3) {
        double t[3];
        for (int k = 0; k < 3; ++k) {
            const double* b = Rb + 3 * (r + k);
            t[k] = b[0] * v[r] + b[1] * v[r + 1] + b[2] * v[r + 2];
        }
        v[r] = t[0]; v[r + 1] = t[1]; v[r + 2] = t[2];
    }
}

/* Residual (projected_jacobi.cpp:140-144). */
static double pj_residual(int nv, const double* v, const double* vp, double tol_rel, double tol_abs) {
    double s = 0;
    for (int j = 0; j < nv; ++j) {
        const double q = dmax(fabs(v[j]), fabs(vp[j])) * tol_rel + tol_abs;
        const double e = (v[j] - vp[j]) / q;
        s += e * e;
    }
    return 1.0 / sqrt((double)nv) * sqrt(s);
}

/* Projection (projected_jacobi.cpp:100-119): normals >= 0, tangents into the Coulomb disk. */
static void pj_project(const orc_W* W, double* p, int nthreads) {
    for (int i = 0; i < W->nnfc; ++i) p[i] = dmax(p[i], 0.0);
    const int nc = W->nrows;
    const int nnfc = W->nnfc;
    const double* mu = W->mu;
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int i = nnfc; i < nc; i += 3) {
        p[i] = dmax(p[i], 0.0);
        const double t1 = p[i + 1], t2 = p[i + 2];
        const double tn2 = t1 * t1 + t2 * t2;
        if (tn2 <= 0) continue;
        const double s = dmin(1.0, (mu[i] * p[i]) / sqrt(tn2));
        p[i + 1] = s * t1;
        p[i + 2] = s * t2;
    }
}

static void fill_mdiag(int nv, const double* minv, double* M) {
    /* rebuildMass_: M = 1/Minv where Minv > 0 else 0 (dynamics_assembler.cpp:56-60) */
    for (int j = 0; j < nv; ++j) M[j] = minv[j] > 0 ? 1.0 / minv[j] : 0.0;
}

/* S of the contact-only system: diag(M), or with implicit gyroscopy (blk != NULL) the inverse
 * rotational block of every rigid body in blk (9 doubles per body, row-major). */
typedef struct {
    int nb;
    const int* dof;
    const int* voff;
    const double* M;
    const double* blk;
} orc_sinv;

/* The inverse rotational blocks of S with implicit gyroscopy (PjAssembler::buildAndFactorS,
 * pj_assembler.cpp:32-61): a rigid body (6 DOFs) with inertia I = 1 / minv[3..5] and angular
 * velocity w = v[3..5] has S_rot = diag(I) + corr, corr = -dt G_rot, G_rot = 1/2 ([I w]_x -
 * [w]_x diag(I)), [a]_x b = a x b; inverted here by its adjugate (the reference factors the whole
 * S by SparseLU).  Returns ORC_NOT_FACTORIZED for a singular or non-finite block.  The
 * ORC_MUT_GYRO_* macros each plant one deliberate mistake for the mutation tests. */
static int gyro_blocks(const orc_bodies* B, const int* voff, double dt, double* blk) {
    for (int b = 0; b < B->nb; ++b) {
        double* K = blk + 9 * (size_t)b;
        for (int k = 0; k < 9; ++k) K[k] = 0.0;
        if (B->dof[b] != 6) continue;
        const int o = voff[b];
        double I[3], w[3], Iw[3];
        for (int k = 0; k < 3; ++k) {
            I[k] = B->minv[o + 3 + k] > 0 ? 1.0 / B->minv[o + 3 + k] : 0.0;
            w[k] = B->v[o + 3 + k];
            Iw[k] = I[k] * w[k];
        }
        const double sw[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
        const double sIw[9] = {0, -Iw[2], Iw[1], Iw[2], 0, -Iw[0], -Iw[1], Iw[0], 0};
        double S[9];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
#if defined(ORC_MUT_GYRO_T)
                const double g = 0.5 * (sIw[3 * c + r] - sw[3 * c + r] * I[r]);
#elif defined(ORC_MUT_GYRO_HALF)
                const double g = sIw[3 * r + c] - sw[3 * r + c] * I[c];
#else
                const double g = 0.5 * (sIw[3 * r + c] - sw[3 * r + c] * I[c]);
#endif
#if defined(ORC_MUT_GYRO_SIGN)
                const double corr = dt * g;
#else
                const double corr = -dt * g;
#endif
                S[3 * r + c] = r == c ? I[r] + corr : corr;
            }
        const double c0 = S[4] * S[8] - S[5] * S[7], c1 = S[5] * S[6] - S[3] * S[8], c2 = S[3] * S[7] - S[4] * S[6];
        const double det = S[0] * c0 + S[1] * c1 + S[2] * c2;
        if (!(det != 0.0) || !isfinite(det)) return ORC_NOT_FACTORIZED;
        K[0] = c0 / det;
        K[1] = (S[2] * S[7] - S[1] * S[8]) / det;
        K[2] = (S[1] * S[5] - S[2] * S[4]) / det;
        K[3] = c1 / det;
        K[4] = (S[0] * S[8] - S[2] * S[6]) / det;
        K[5] = (S[2] * S[3] - S[0] * S[5]) / det;
        K[6] = c2 / det;
        K[7] = (S[1] * S[6] - S[0] * S[7]) / det;
        K[8] = (S[0] * S[4] - S[1] * S[3]) / det;
    }
    return ORC_OK;
}

/* y <- S^-1 y (solveS, pj_assembler.cpp:173-182) */
static void sinv_inplace(const orc_sinv* s, double* y) {
    for (int b = 0; b < s->nb; ++b) {
        const int o = s->voff[b], nd = s->dof[b];
        if (s->blk && nd == 6) {
            const double* K = s->blk + 9 * (size_t)b;
            for (int j = o; j < o + 3; ++j) y[j] = y[j] / s->M[j];
            const double y3 = y[o + 3], y4 = y[o + 4], y5 = y[o + 5];
            for (int r = 0; r < 3; ++r) y[o + 3 + r] = K[3 * r] * y3 + K[3 * r + 1] * y4 + K[3 * r + 2] * y5;
        } else {
            for (int j = o; j < o + nd; ++j) y[j] = y[j] / s->M[j];
        }
    }
}

/* Spectral radius of the linear sweep map v -> v - S^-1 W^T R W v, R the row scaling or R_block
 * (estimateSpectralRadius, projected_jacobi.cpp:155-178): 8 power iterations from the normalised ones vector, clamped to
 * 0.995. */
static double estimate_rho(const orc_W* W, const orc_W* RW, const double* Rb, const orc_sinv* SI, int nv) {
    if (nv <= 0) return 0.0;
    double* v = (double*)malloc(sizeof(double) * (size_t)nv);
    double* r = (double*)malloc(sizeof(double) * (size_t)nv);
    double* d = (double*)malloc(sizeof(double) * (size_t)(W->nrows + 1));
    const double inv = 1.0 / sqrt((double)nv);
    for (int j = 0; j < nv; ++j) v[j] = inv;
    double lambda = 0.0;
    for (int it = 0; it < 8; ++it) {
        spmv(RW, v, d);
        if (Rb) apply_rblock(W, Rb, d);
        memset(r, 0, sizeof(double) * (size_t)nv);
        spmv_t_acc(W, d, r);
        sinv_inplace(SI, r);
        double n2 = 0.0;
        for (int j = 0; j < nv; ++j) {
            v[j] = v[j] - r[j];
            n2 += v[j] * v[j];
        }
        const double n = sqrt(n2);
        if (!(n > 1e-300)) { lambda = 0.0; break; }
        lambda = n;
        for (int j = 0; j < nv; ++j) v[j] = v[j] / n;
    }
    free(v); free(r); free(d);
    return dmin(lambda, 0.995);
}

/* Householder vector of v = [c0; tail] (length n) in place (Eigen's makeHouseholderInPlace): v becomes
 * [beta; essential], returns tau; tau = 0 when the tail is zero. */
static double householder_in_place(int n, double* v) {
    double tail2 = 0;
    for (int i = 1; i < n; ++i) tail2 += v[i] * v[i];
    const double c0 = v[0];
    if (tail2 <= DBL_MIN) {
        for (int i = 1; i < n; ++i) v[i] = 0;
        return 0.0;
    }
    double beta = sqrt(c0 * c0 + tail2);
    if (c0 >= 0) beta = -beta;
    for (int i = 1; i < n; ++i) v[i] = v[i] / (c0 - beta);
    v[0] = beta;
    return (beta - c0) / beta;
}

/* y <- (I - tau [1; e][1; e]^T) y for y of length n (Eigen's applyHouseholderOnTheLeft) */
static void apply_householder(int n, const double* e, double tau, double* y) {
    if (n == 1) { y[0] *= 1.0 - tau; return; }
    if (tau == 0.0) return;
    double t = 0;
    for (int i = 1; i < n; ++i) t += e[i] * y[i];
    t += y[0];
    y[0] -= tau * t;
    for (int i = 1; i < n; ++i) y[i] -= (tau * e[i]) * t;
}

/* ColPivHouseholderQR::compute and solve (Eigen, the reference's pinned 5.0.1; Businger-Golub column
 * pivoting on the downdated column norms, a pivot counted while the largest remaining column norm^2
 * is at least (maxColNorm eps)^2 (rows - k) / rows) for the least-squares step of anderson_loop. */
int orc_lstsq_colpiv(int rows, int cols, double* A, const double* b, double* g) {
    const int size = rows < cols ? rows : cols;
    double* nd = (double*)malloc(sizeof(double) * (size_t)(cols + 1)); /* direct norms */
    double* nu = (double*)malloc(sizeof(double) * (size_t)(cols + 1)); /* updated norms */
    double* tau = (double*)malloc(sizeof(double) * (size_t)(size + 1));
    int* perm = (int*)malloc(sizeof(int) * (size_t)(cols + 1));
    double* c = (double*)malloc(sizeof(double) * (size_t)(rows + 1));
    double maxnorm = 0;
    for (int j = 0; j < cols; ++j) {
        double s = 0;
        for (int i = 0; i < rows; ++i) s += A[(size_t)j * rows + i] * A[(size_t)j * rows + i];
        nd[j] = nu[j] = sqrt(s);
        perm[j] = j;
        if (nd[j] > maxnorm) maxnorm = nd[j];
    }
    const double eps = DBL_EPSILON;
    const double thr = (maxnorm * eps) * (maxnorm * eps) / (double)rows;
    const double downdate_thr = sqrt(eps);
    int nz = size;
    for (int k = 0; k < size; ++k) {
        int big = k;
        for (int j = k + 1; j < cols; ++j)
            if (nu[j] > nu[big]) big = j;
        const double big2 = nu[big] * nu[big];
        if (nz == size && big2 < thr * (double)(rows - k)) nz = k;
        if (big != k) {
            for (int i = 0; i < rows; ++i) {
                const double t = A[(size_t)k * rows + i];
                A[(size_t)k * rows + i] = A[(size_t)big * rows + i];
                A[(size_t)big * rows + i] = t;
            }
            double t = nu[k]; nu[k] = nu[big]; nu[big] = t;
            t = nd[k]; nd[k] = nd[big]; nd[big] = t;
            const int ti = perm[k]; perm[k] = perm[big]; perm[big] = ti;
        }
        double* ck = A + (size_t)k * rows + k;
        tau[k] = householder_in_place(rows - k, ck);
        for (int j = k + 1; j < cols; ++j) apply_householder(rows - k, ck, tau[k], A + (size_t)j * rows + k);
        for (int j = k + 1; j < cols; ++j) {
            if (nu[j] == 0.0) continue;
            double t = fabs(A[(size_t)j * rows + k]) / nu[j];
            t = (1.0 + t) * (1.0 - t);
            if (t < 0) t = 0;
            const double t2 = t * (nu[j] / nd[j]) * (nu[j] / nd[j]);
            if (t2 <= downdate_thr) {
                double s = 0;
                for (int i = k + 1; i < rows; ++i) s += A[(size_t)j * rows + i] * A[(size_t)j * rows + i];
                nd[j] = nu[j] = sqrt(s);
            } else {
                nu[j] *= sqrt(t);
            }
        }
    }
    for (int j = 0; j < cols; ++j) g[j] = 0;
    if (nz > 0) {
        memcpy(c, b, sizeof(double) * (size_t)rows);
        for (int k = 0; k < nz; ++k) apply_householder(rows - k, A + (size_t)k * rows + k, tau[k], c + k);
        for (int i = nz - 1; i >= 0; --i) { /* R(0:nz, 0:nz) upper triangular */
            double s = c[i];
            for (int j = i + 1; j < nz; ++j) s -= A[(size_t)j * rows + i] * c[j];
            c[i] = s / A[(size_t)i * rows + i];
        }
        for (int i = 0; i < nz; ++i) g[perm[i]] = c[i];
    }
    free(nd); free(nu); free(tau); free(perm); free(c);
    return nz;
}

/* PJ solve: ProjectedJacobiSolver::solve (projected_jacobi.cpp:451-500) with build_workspace
 * (:316-369), standard_loop (:371-377) / nesterov_loop (:379-449), pj_sweep (:121-138),
 * rhs (pj_assembler.cpp:127-170, contact-only), solveS on S = diag(M) (pj_assembler.cpp:26-30,173-182),
 * warm start / store (warmstart.hpp:26-56). */
int orc_pj_solve(const orc_bodies* B, const orc_contacts* C, double dt, double theta, const orc_params* prm,
                 int nthreads, double* v_out, double* imp_out, orc_result* res) {
    (void)theta; /* theta only enters spring/damper rows and S for bilateral scenes */
    /* nthreads < 0: fully threaded sweep (both products and the vector updates on |nthreads|
     * threads) -- a timing variant only; the sums of W^T p then run column-wise. */
    const int full = nthreads < 0;
    if (full) nthreads = -nthreads;
    if (nthreads < 1) nthreads = 1;
    orc_W W;
    int st = orc_build_W(B, C, &W);
    if (st) { orc_free_W(&W); return st; }
    const int nv = W.nv, nr = W.nrows;
    res->iters = 0;
    res->residual = NAN;
    res->nrows = nr;
    res->rho = NAN;
    res->ncolors = 0;
    double* M = (double*)malloc(sizeof(double) * (size_t)(nv + 1));
    fill_mdiag(nv, B->minv, M);
    for (int j = 0; j < nv; ++j)
        if (M[j] == 0.0) { free(M); orc_free_W(&W); return ORC_NOT_FACTORIZED; } /* singular diagonal S */
    double* blk = NULL;
    if (prm->implicit_gyro) {
        blk = (double*)malloc(sizeof(double) * 9 * (size_t)(B->nb + 1));
        if (gyro_blocks(B, W.voff, dt, blk)) { free(blk); free(M); orc_free_W(&W); return ORC_NOT_FACTORIZED; }
    }
    const orc_sinv SI = {B->nb, B->dof, W.voff, M, blk};

    /* rhs_free = M v + dt f_ext (+ dt f_gyro unless implicit) (pj_assembler.cpp:152-153) */
    double* x = (double*)malloc(sizeof(double) * (size_t)(nv + 1));
    for (int j = 0; j < nv; ++j) x[j] = M[j] * B->v[j] + dt * B->fext[j];
#if defined(ORC_MUT_GYRO_RHS)
    for (int j = 0; j < nv; ++j) x[j] += dt * B->fgyro[j];
#else
    if (!prm->implicit_gyro)
        for (int j = 0; j < nv; ++j) x[j] += dt * B->fgyro[j];
#endif

    if (nr == 0) { /* no contacts: free solution (:460-467) */
        memcpy(v_out, x, sizeof(double) * (size_t)nv);
        sinv_inplace(&SI, v_out);
        for (int c = 0; c < C->nc; ++c) { imp_out[3 * c] = imp_out[3 * c + 1] = imp_out[3 * c + 2] = 0; }
        free(x); free(blk); free(M); orc_free_W(&W);
        return ORC_OK;
    }

    double* R = (double*)malloc(sizeof(double) * (size_t)nr);
    double* bias = (double*)malloc(sizeof(double) * (size_t)nr);
    double* p = (double*)calloc((size_t)nr, sizeof(double));
    double* pp = (double*)malloc(sizeof(double) * (size_t)nr);
    double* dl = (double*)malloc(sizeof(double) * (size_t)nr);
    double* xp = (double*)malloc(sizeof(double) * (size_t)nv);
    double* rhs = (double*)malloc(sizeof(double) * (size_t)nv);
    double* RWv = (double*)malloc(sizeof(double) * (W.rowptr[nr] + 1));

    /* v_prev = W v + c; biasImpulse = c + e .* v_prev; R = alpha / D; bias = R .* bias (:337-347) */
    spmv(&W, B->v, bias);
    for (int r = 0; r < nr; ++r) bias[r] += W.cvec[r];
    for (int r = 0; r < nr; ++r) bias[r] = W.cvec[r] + W.rest[r] * bias[r];
    double* Rb = NULL;
    if (prm->block_precond) { /* RW = W, bias unscaled, R_block applied in the sweep (:341-343) */
        Rb = (double*)malloc(sizeof(double) * 3 * (size_t)nr);
        build_rblock(&W, B->minv, prm->alpha, Rb);
        for (int k = 0; k < W.rowptr[nr]; ++k) RWv[k] = W.val[k];
    } else {
        delassus_diag(&W, B->minv, R);
        for (int r = 0; r < nr; ++r) R[r] = (R[r] > 0) ? (prm->alpha / R[r]) : 0.0;
        for (int r = 0; r < nr; ++r) bias[r] = R[r] * bias[r];
        for (int r = 0; r < nr; ++r)
            for (int k = W.rowptr[r]; k < W.rowptr[r + 1]; ++k) RWv[k] = R[r] * W.val[k];
    }

    /* warm start (:358-359, warmstart.hpp:26-37) */
    if (prm->warmstart)
        for (int ci = 0; ci < C->nc; ++ci) {
            const int base = W.row_base[ci];
            if (base < 0) continue;
            p[base] = C->warm[3 * ci];
            if (W.row_size[ci] > 1) { p[base + 1] = C->warm[3 * ci + 1]; p[base + 2] = C->warm[3 * ci + 2]; }
        }
    /* x0 = S^-1 (rhs_free + W^T p) (:361-363) */
    spmv_t_acc(&W, p, x);
    sinv_inplace(&SI, x);

    orc_W RW = W;
    RW.val = RWv;
    orc_Wt WT = {0, 0, 0};
    if (full) build_Wt(&W, &WT);

#define PJ_SWEEP()                                                          \
    do {                                                                    \
        if (full) {                                                         \
            _Pragma("omp parallel for schedule(static) num_threads(nthreads)") \
            for (int j = 0; j < nv; ++j) xp[j] = x[j];                      \
            spmv_mt(&RW, x, dl, nthreads);                                  \
            _Pragma("omp parallel for schedule(static) num_threads(nthreads)") \
            for (int r = 0; r < nr; ++r) dl[r] += bias[r];                  \
            if (Rb) apply_rblock(&W, Rb, dl);                               \
            _Pragma("omp parallel for schedule(static) num_threads(nthreads)") \
            for (int r = 0; r < nr; ++r) {                                  \
                pp[r] = p[r];                                               \
                p[r] -= dl[r];                                              \
            }                                                               \
            pj_project(&W, p, nthreads);                                    \
            _Pragma("omp parallel for schedule(static) num_threads(nthreads)") \
            for (int r = 0; r < nr; ++r) dl[r] = p[r] - pp[r];              \
            spmv_t_mt(&W, &WT, dl, rhs, nthreads);                          \
            if (blk) {                                                      \
                sinv_inplace(&SI, rhs);                                     \
                for (int j = 0; j < nv; ++j) x[j] += rhs[j];                \
                break;                                                      \
            }                                                               \
            _Pragma("omp parallel for schedule(static) num_threads(nthreads)") \
            for (int j = 0; j < nv; ++j) x[j] += rhs[j] / M[j];             \
            break;                                                          \
        }                                                                   \
        memcpy(xp, x, sizeof(double) * (size_t)nv);                         \
        memcpy(pp, p, sizeof(double) * (size_t)nr);                         \
        spmv(&RW, x, dl);                                                   \
        for (int r = 0; r < nr; ++r) dl[r] += bias[r];                      \
        if (Rb) apply_rblock(&W, Rb, dl);                                   \
        for (int r = 0; r < nr; ++r) p[r] -= dl[r];                         \
        pj_project(&W, p, nthreads);                                        \
        for (int r = 0; r < nr; ++r) dl[r] = p[r] - pp[r];                  \
        memset(rhs, 0, sizeof(double) * (size_t)nv);                        \
        spmv_t_acc(&W, dl, rhs);                                            \
        sinv_inplace(&SI, rhs);                                             \
        for (int j = 0; j < nv; ++j) x[j] += rhs[j];                        \
    } while (0)

    if (!prm->nesterov && !prm->chebyshev && prm->anderson) {
        /* anderson_loop (:242-310): z = [x; p]; one plain sweep w = g(z) per iteration, residual every
         * iteration on the velocity part; the history keeps the last m + 1 pairs (z_i, f_i = g(z_i) - z_i)
         * and z_{k+1} = w - (dZ + dF) gamma with gamma = argmin ||f_k - dF gamma|| (column-pivoted
         * Householder QR); a non-finite gamma or candidate takes the plain step and drops the history. */
        const int m = prm->anderson_m < 1 ? 1 : prm->anderson_m;
        const int N = nv + nr;
        double* Zh = (double*)malloc(sizeof(double) * (size_t)(m + 1) * N);
        double* Fh = (double*)malloc(sizeof(double) * (size_t)(m + 1) * N);
        double* z = (double*)malloc(sizeof(double) * (size_t)N);
        double* w = (double*)malloc(sizeof(double) * (size_t)N);
        double* dF = (double*)malloc(sizeof(double) * (size_t)m * N);
        double* dZ = (double*)malloc(sizeof(double) * (size_t)m * N);
        double* gam = (double*)malloc(sizeof(double) * (size_t)m);
        double* cand = (double*)malloc(sizeof(double) * (size_t)N);
        double* tmp = (double*)malloc(sizeof(double) * (size_t)N);
        memcpy(z, x, sizeof(double) * nv);
        memcpy(z + nv, p, sizeof(double) * nr);
        int cnt = 0; /* pairs in the history, oldest at index 0 */
        for (int iter = 0; iter < prm->max_iter; ++iter) {
            memcpy(x, z, sizeof(double) * nv);
            memcpy(p, z + nv, sizeof(double) * nr);
            PJ_SWEEP();
            memcpy(w, x, sizeof(double) * nv);
            memcpy(w + nv, p, sizeof(double) * nr);
            res->iters = iter + 1;
            const double err = pj_residual(nv, x, xp, prm->tol_rel, prm->tol_abs);
            res->residual = err;
            if (err <= 1) {
                memcpy(z, w, sizeof(double) * N);
                break;
            }
            if (cnt == m + 1) { /* pop_front after the push below */
                memmove(Zh, Zh + N, sizeof(double) * (size_t)m * N);
                memmove(Fh, Fh + N, sizeof(double) * (size_t)m * N);
                --cnt;
            }
            memcpy(Zh + (size_t)cnt * N, z, sizeof(double) * N);
            for (int i = 0; i < N; ++i) Fh[(size_t)cnt * N + i] = w[i] - z[i];
            ++cnt;
            const int mk = cnt - 1;
            if (mk < 1) {
                memcpy(z, w, sizeof(double) * N);
                continue;
            }
            for (int j = 0; j < mk; ++j)
                for (int i = 0; i < N; ++i) {
                    dF[(size_t)j * N + i] = Fh[(size_t)(j + 1) * N + i] - Fh[(size_t)j * N + i];
                    dZ[(size_t)j * N + i] = Zh[(size_t)(j + 1) * N + i] - Zh[(size_t)j * N + i];
                }
            const double* f = Fh + (size_t)mk * N;
            /* (dZ + dF) before dF is overwritten by its factorisation */
            for (int j = 0; j < mk; ++j)
                for (int i = 0; i < N; ++i) dZ[(size_t)j * N + i] = dZ[(size_t)j * N + i] + dF[(size_t)j * N + i];
            orc_lstsq_colpiv(N, mk, dF, f, gam);
            int finite = 1;
            for (int j = 0; j < mk; ++j) finite = finite && isfinite(gam[j]);
            for (int i = 0; i < N; ++i) tmp[i] = 0;
            for (int j = 0; j < mk; ++j)
                for (int i = 0; i < N; ++i) tmp[i] += dZ[(size_t)j * N + i] * gam[j];
            for (int i = 0; i < N; ++i) {
                cand[i] = w[i] - tmp[i];
                finite = finite && isfinite(cand[i]);
            }
            if (finite) {
                memcpy(z, cand, sizeof(double) * N);
            } else {
                memcpy(z, w, sizeof(double) * N);
                cnt = 0;
            }
        }
        memcpy(x, z, sizeof(double) * nv);
        memcpy(p, z + nv, sizeof(double) * nr);
        free(Zh); free(Fh); free(z); free(w); free(dF); free(dZ); free(gam); free(cand); free(tmp);
    } else if (!prm->nesterov && prm->chebyshev) {
        /* chebyshev_loop (:180-228): residual every sweep against the sweep's input; on
         * convergence the loop returns that input (x = xk1), not the swept iterate */
        const double rho = estimate_rho(&W, &RW, Rb, &SI, nv);
        const double rho2 = rho * rho;
        res->rho = rho;
        double* xk = (double*)malloc(sizeof(double) * (size_t)nv);
        double* pk = (double*)malloc(sizeof(double) * (size_t)nr);
        double* xk1 = (double*)malloc(sizeof(double) * (size_t)nv);
        double* pk1 = (double*)malloc(sizeof(double) * (size_t)nr);
        memcpy(xk, x, sizeof(double) * nv); memcpy(xk1, x, sizeof(double) * nv);
        memcpy(pk, p, sizeof(double) * nr); memcpy(pk1, p, sizeof(double) * nr);
        double omega = 1.0;
        for (int iter = 0; iter < prm->max_iter; ++iter) {
            memcpy(x, xk1, sizeof(double) * nv);
            memcpy(p, pk1, sizeof(double) * nr);
            PJ_SWEEP();
            res->iters = iter + 1;
            const double err = pj_residual(nv, x, xk1, prm->tol_rel, prm->tol_abs);
            res->residual = err;
            if (err <= 1) break;
            double omega_next = (iter == 0) ? (1.0 / (1.0 - rho2 * 0.5)) : (1.0 / (1.0 - (rho2 * 0.25) * omega));
            int finite = 1;
            for (int j = 0; j < nv && finite; ++j) finite = isfinite(x[j]);
            if (!isfinite(omega_next) || omega_next <= 0.0 || !finite) {
                memcpy(xk, xk1, sizeof(double) * nv); memcpy(pk, pk1, sizeof(double) * nr);
                memcpy(xk1, x, sizeof(double) * nv); memcpy(pk1, p, sizeof(double) * nr);
                omega = 1.0;
            } else {
                for (int j = 0; j < nv; ++j) { const double t = omega_next * (x[j] - xk[j]) + xk[j]; xk[j] = xk1[j]; xk1[j] = t; }
                for (int r = 0; r < nr; ++r) { const double t = omega_next * (p[r] - pk[r]) + pk[r]; pk[r] = pk1[r]; pk1[r] = t; }
                omega = omega_next;
            }
        }
        memcpy(x, xk1, sizeof(double) * nv);
        memcpy(p, pk1, sizeof(double) * nr);
        free(xk); free(pk); free(xk1); free(pk1);
    } else if (!prm->nesterov) {
        /* standard_loop (:371-377): residual only when iter % 10 == 0 */
        for (int iter = 0; iter < prm->max_iter; ++iter) {
            PJ_SWEEP();
            res->iters = iter + 1;
            if (iter % 10 == 0) {
                const double e = pj_residual(nv, x, xp, prm->tol_rel, prm->tol_abs);
                res->residual = e;
                if (e <= 1) break;
            }
        }
    } else {
        /* nesterov_loop (:379-449).  The ORC_MUT_NES_* macros each plant one deliberate mistake
         * (theta or beta recurrence, growth factor, direction test, restart limit, extrapolation
         * anchor) for the mutation tests; the oracle itself is built with none of them. */
        double* xk = (double*)malloc(sizeof(double) * (size_t)nv);
        double* pk = (double*)malloc(sizeof(double) * (size_t)nr);
        double* xy = (double*)malloc(sizeof(double) * (size_t)nv);
        double* py = (double*)malloc(sizeof(double) * (size_t)nr);
        double* xk1 = (double*)malloc(sizeof(double) * (size_t)nv);
        double* pk1 = (double*)malloc(sizeof(double) * (size_t)nr);
        memcpy(xk, x, sizeof(double) * nv); memcpy(xy, x, sizeof(double) * nv); memcpy(xk1, x, sizeof(double) * nv);
        memcpy(pk, p, sizeof(double) * nr); memcpy(py, p, sizeof(double) * nr); memcpy(pk1, p, sizeof(double) * nr);
        double thk = 1.0, err_prev = INFINITY;
        int restarts = 0, momentum_disabled = 0;
        for (int iter = 0; iter < prm->max_iter; ++iter) {
            memcpy(x, xy, sizeof(double) * nv);
            memcpy(p, py, sizeof(double) * nr);
            PJ_SWEEP();
            memcpy(xk1, x, sizeof(double) * nv);
            memcpy(pk1, p, sizeof(double) * nr);
            res->iters = iter + 1;
            const double err = pj_residual(nv, xk1, xy, prm->tol_rel, prm->tol_abs);
            res->residual = err;
            if (err <= 1) break;
#if defined(ORC_MUT_NES_THETA)
            const double thk1 = 0.5 * (1.0 + sqrt(thk * thk + 1.0));
#else
            const double thk1 = 0.5 * (1.0 + sqrt(4.0 * thk * thk + 1.0));
#endif
#if defined(ORC_MUT_NES_BETA)
            const double betak1 = thk / thk1;
#else
            const double betak1 = (thk - 1.0) / thk1;
#endif
            int restart = 0;
            if (!isfinite(err)) restart = 1;
#if defined(ORC_MUT_NES_GROWTH)
            else if (err_prev < INFINITY && err > 1.5 * err_prev) restart = 1;
#else
            else if (err_prev < INFINITY && err > 1.05 * err_prev) restart = 1;
#endif
            else if (betak1 > prm->nesterov_beta_threshold) restart = 1;
            else if (!isfinite(betak1) || betak1 < 0.0 || betak1 > 1.0) restart = 1;
            else {
                double dot = 0;
                for (int j = 0; j < nv; ++j) dot += (xy[j] - xk1[j]) * (xk1[j] - xk[j]);
#if defined(ORC_MUT_NES_DIRECTION)
                if (dot < 0.0) restart = 1;
#else
                if (dot > 0.0) restart = 1;
#endif
            }
            if (restart) {
                memcpy(xy, xk1, sizeof(double) * nv); memcpy(py, pk1, sizeof(double) * nr);
                memcpy(xk, xk1, sizeof(double) * nv); memcpy(pk, pk1, sizeof(double) * nr);
                thk = 1.0;
#if defined(ORC_MUT_NES_LIMIT)
                if (++restarts > prm->nesterov_restart_limit) momentum_disabled = 1;
#else
                if (++restarts >= prm->nesterov_restart_limit) momentum_disabled = 1;
#endif
            } else {
                if (momentum_disabled) {
                    memcpy(xy, xk1, sizeof(double) * nv); memcpy(py, pk1, sizeof(double) * nr);
                    thk = 1.0;
                } else {
#if defined(ORC_MUT_NES_ANCHOR)
                    for (int j = 0; j < nv; ++j) xy[j] = xk1[j] + betak1 * (xk1[j] - xy[j]);
#else
                    for (int j = 0; j < nv; ++j) xy[j] = xk1[j] + betak1 * (xk1[j] - xk[j]);
#endif
                    for (int r = 0; r < nr; ++r) py[r] = pk1[r] + betak1 * (pk1[r] - pk[r]);
                    thk = thk1;
                }
                memcpy(xk, xk1, sizeof(double) * nv); memcpy(pk, pk1, sizeof(double) * nr);
            }
            err_prev = err;
        }
        memcpy(x, xk1, sizeof(double) * nv);
        memcpy(p, pk1, sizeof(double) * nr);
        free(xk); free(pk); free(xy); free(py); free(xk1); free(pk1);
    }
#undef PJ_SWEEP

    /* storeImpulse (warmstart.hpp:41-56) */
    for (int ci = 0; ci < C->nc; ++ci) {
        double* o = imp_out + 3 * (size_t)ci;
        o[0] = o[1] = o[2] = 0;
        const int base = W.row_base[ci];
        if (base < 0) continue;
        o[0] = p[base];
        if (W.row_size[ci] > 1) { o[1] = p[base + 1]; o[2] = p[base + 2]; }
    }
    memcpy(v_out, x, sizeof(double) * (size_t)nv);
    free(R); free(Rb); free(bias); free(p); free(pp); free(dl); free(xp); free(rhs); free(RWv); free(x); free(blk); free(M);
    free(WT.ptr); free(WT.row); free(WT.ent);
    orc_free_W(&W);
    return ORC_OK;
}

/* ---- extended system with bilateral rows ---------------------------------------------- */

typedef struct {
    int n;
    double* a;  /* n x n row-major, LU in place */
    int* piv;
} orc_lu;

/* Doolittle LU with partial pivoting; returns 0 on success, -1 on a zero pivot. */
static int lu_factor(orc_lu* f) {
    const int n = f->n;
    double* a = f->a;
    for (int k = 0; k < n; ++k) {
        int pr = k;
        double best = fabs(a[(size_t)k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(a[(size_t)i * n + k]) > best) best = fabs(a[(size_t)i * n + k]), pr = i;
        if (best == 0.0) return -1;
        f->piv[k] = pr;
        if (pr != k)
            for (int j = 0; j < n; ++j) {
                const double t = a[(size_t)k * n + j];
                a[(size_t)k * n + j] = a[(size_t)pr * n + j];
                a[(size_t)pr * n + j] = t;
            }
        const double d = a[(size_t)k * n + k];
        for (int i = k + 1; i < n; ++i) {
            const double l = a[(size_t)i * n + k] / d;
            a[(size_t)i * n + k] = l;
            if (l != 0.0)
                for (int j = k + 1; j < n; ++j) a[(size_t)i * n + j] -= l * a[(size_t)k * n + j];
        }
    }
    return 0;
}

static void lu_solve(const orc_lu* f, const double* b, double* x) {
    const int n = f->n;
    const double* a = f->a;
    memcpy(x, b, sizeof(double) * (size_t)n);
    for (int k = 0; k < n; ++k) {
        const int pr = f->piv[k];
        if (pr != k) { const double t = x[k]; x[k] = x[pr]; x[pr] = t; }
    }
    for (int i = 1; i < n; ++i) {
        double s = x[i];
        for (int j = 0; j < i; ++j) s -= a[(size_t)i * n + j] * x[j];
        x[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = x[i];
        for (int j = i + 1; j < n; ++j) s -= a[(size_t)i * n + j] * x[j];
        x[i] = s / a[(size_t)i * n + i];
    }
}

/* y[r] = sum_j G_rj v_j over the body dofs of row r */
static double joint_row_dot(const orc_joints* jt, const int* voff, const int* dof, int r, const double* v) {
    double s = 0;
    for (int side = 0; side < 2; ++side) {
        const int b = side ? jt->body_b[r] : jt->body_a[r];
        if (b < 0) continue;
        const double* w = (side ? jt->wb : jt->wa) + 6 * (size_t)r;
        for (int j = 0; j < dof[b]; ++j) s += w[j] * v[voff[b] + j];
    }
    return s;
}

int orc_pj_solve_joints(const orc_bodies* B, const orc_contacts* C, const orc_joints* jt, double dt, double theta,
                        const orc_params* prm, double* v_out, double* imp_out, double* lam_out, orc_result* res) {
    /* the loops of ProjectedJacobiSolver::solve (:469-476) over x = [v; Lambda]: Nesterov before
     * Chebyshev; Anderson, the block preconditioner and implicit gyroscopy are not restated with
     * joint rows */
    if (!jt || jt->ns < 0 || jt->nd < 0 || prm->block_precond || (prm->anderson && !prm->nesterov && !prm->chebyshev) ||
        prm->implicit_gyro)
        return ORC_BAD_ARG;
    orc_W W;
    int st = orc_build_W(B, C, &W);
    if (st) { orc_free_W(&W); return st; }
    const int nv = W.nv, nr = W.nrows, ns = jt->ns, nd = jt->nd, nj = ns + nd, N = nv + nj;
    res->iters = 0;
    res->residual = NAN;
    res->nrows = nr;
    res->rho = NAN;
    res->ncolors = 0;
    for (int r = 0; r < nj; ++r) {
        const int a = jt->body_a[r], b = jt->body_b[r];
        if (a < -1 || a >= B->nb || b < -1 || b >= B->nb || (a >= 0 && a == b)) { orc_free_W(&W); return ORC_BAD_ARG; }
    }
    double* M = (double*)malloc(sizeof(double) * (size_t)(nv + 1));
    fill_mdiag(nv, B->minv, M);
    /* S = [M, Wg^T, Wgamma^T; Wg, -C/(theta dt^2), 0; Wgamma, 0, -A/(theta dt)] (:66-73) */
    orc_lu lu = {N, (double*)calloc((size_t)N * N + 1, sizeof(double)), (int*)malloc(sizeof(int) * (size_t)(N + 1))};
    for (int i = 0; i < nv; ++i) lu.a[(size_t)i * N + i] = M[i];
    const double cScale = -1.0 / (theta * dt * dt), aScale = -1.0 / (theta * dt);
    for (int r = 0; r < nj; ++r) {
        const int R = nv + r;
        for (int side = 0; side < 2; ++side) {
            const int b = side ? jt->body_b[r] : jt->body_a[r];
            if (b < 0) continue;
            const double* w = (side ? jt->wb : jt->wa) + 6 * (size_t)r;
            for (int j = 0; j < B->dof[b]; ++j) {
                if (w[j] == 0.0) continue; /* zero entries dropped (emitColRef) */
                lu.a[(size_t)R * N + W.voff[b] + j] += w[j];
                lu.a[(size_t)(W.voff[b] + j) * N + R] += w[j];
            }
        }
        lu.a[(size_t)R * N + R] = jt->comp[r] * (r < ns ? cScale : aScale);
    }
    if (lu_factor(&lu)) { free(lu.a); free(lu.piv); free(M); orc_free_W(&W); return ORC_NOT_FACTORIZED; }

    /* rhs (PjAssembler::rhs, :127-170) */
    double* rhs = (double*)calloc((size_t)N + 1, sizeof(double));
    for (int j = 0; j < nv; ++j) rhs[j] = M[j] * B->v[j] + dt * B->fext[j];
    for (int j = 0; j < nv; ++j) rhs[j] += dt * B->fgyro[j];
    for (int r = 0; r < nj; ++r) {
        const double wv = joint_row_dot(jt, W.voff, B->dof, r, B->v);
        if (r < ns)
            rhs[nv + r] = -(1.0 / (theta * dt * dt)) * (jt->comp[r] * jt->lam_prev[r]) - ((1.0 - theta) / theta) * wv -
                          (1.0 / theta) * jt->vsrc[r];
        else
            rhs[nv + r] = -((1.0 - theta) / theta) * wv - (1.0 / theta) * jt->vsrc[r];
    }
    if (jt->bias_factor > 0)
        for (int r = 0; r < ns; ++r)
            rhs[nv + r] -= (-(jt->comp[r] * jt->lam_prev[r]) / dt + jt->gerr[r]) * (jt->bias_factor / (dt * theta));

    double* x = (double*)malloc(sizeof(double) * ((size_t)N + 1));
    if (nr == 0) { /* no contact rows: free solution (:460-467) */
        lu_solve(&lu, rhs, x);
        memcpy(v_out, x, sizeof(double) * (size_t)nv);
        memcpy(lam_out, x + nv, sizeof(double) * (size_t)nj);
        for (int c = 0; c < C->nc; ++c) { imp_out[3 * c] = imp_out[3 * c + 1] = imp_out[3 * c + 2] = 0; }
        free(x); free(rhs); free(lu.a); free(lu.piv); free(M); orc_free_W(&W);
        return ORC_OK;
    }
    double* Rr = (double*)malloc(sizeof(double) * (size_t)nr);
    double* bias = (double*)malloc(sizeof(double) * (size_t)nr);
    double* p = (double*)calloc((size_t)nr, sizeof(double));
    double* pp = (double*)malloc(sizeof(double) * (size_t)nr);
    double* dl = (double*)malloc(sizeof(double) * (size_t)nr);
    double* xp = (double*)malloc(sizeof(double) * (size_t)nv);
    double* d = (double*)calloc((size_t)N + 1, sizeof(double));
    double* dx = (double*)malloc(sizeof(double) * ((size_t)N + 1));
    double* RWv = (double*)malloc(sizeof(double) * (W.rowptr[nr] + 1));
    spmv(&W, B->v, bias);
    for (int r = 0; r < nr; ++r) bias[r] += W.cvec[r];
    for (int r = 0; r < nr; ++r) bias[r] = W.cvec[r] + W.rest[r] * bias[r];
    delassus_diag(&W, B->minv, Rr);
    for (int r = 0; r < nr; ++r) Rr[r] = (Rr[r] > 0) ? (prm->alpha / Rr[r]) : 0.0;
    for (int r = 0; r < nr; ++r) bias[r] = Rr[r] * bias[r];
    for (int r = 0; r < nr; ++r)
        for (int k = W.rowptr[r]; k < W.rowptr[r + 1]; ++k) RWv[k] = Rr[r] * W.val[k];
    if (prm->warmstart)
        for (int ci = 0; ci < C->nc; ++ci) {
            const int base = W.row_base[ci];
            if (base < 0) continue;
            p[base] = C->warm[3 * ci];
            if (W.row_size[ci] > 1) { p[base + 1] = C->warm[3 * ci + 1]; p[base + 2] = C->warm[3 * ci + 2]; }
        }
    /* x0 = S^-1 (rhs_free + [W^T p0; 0]) (:358-363) */
    spmv_t_acc(&W, p, rhs);
    lu_solve(&lu, rhs, x);
    orc_W RW = W;
    RW.val = RWv;
    /* pj_sweep (:121-138) on the extended state: p from the velocity part, x += S^-1 [W^T dp; 0] */
#define JSWEEP()                                                        \
    do {                                                                \
        memcpy(pp, p, sizeof(double) * (size_t)nr);                     \
        spmv(&RW, x, dl);                                               \
        for (int r = 0; r < nr; ++r) dl[r] += bias[r];                  \
        for (int r = 0; r < nr; ++r) p[r] -= dl[r];                     \
        pj_project(&W, p, 1);                                           \
        for (int r = 0; r < nr; ++r) dl[r] = p[r] - pp[r];              \
        memset(d, 0, sizeof(double) * (size_t)N);                       \
        spmv_t_acc(&W, dl, d);                                          \
        lu_solve(&lu, d, dx);                                           \
        for (int k = 0; k < N; ++k) x[k] += dx[k];                      \
    } while (0)
    if (prm->nesterov) {
        /* nesterov_loop (:379-449) with x = [v; Lambda]: extrapolation over the whole of x, the
         * residual and the restart direction test over the velocities */
        double* xk = (double*)malloc(sizeof(double) * (size_t)N);
        double* pk = (double*)malloc(sizeof(double) * (size_t)nr);
        double* xy = (double*)malloc(sizeof(double) * (size_t)N);
        double* py = (double*)malloc(sizeof(double) * (size_t)nr);
        double* xk1 = (double*)malloc(sizeof(double) * (size_t)N);
        double* pk1 = (double*)malloc(sizeof(double) * (size_t)nr);
        memcpy(xk, x, sizeof(double) * N); memcpy(xy, x, sizeof(double) * N); memcpy(xk1, x, sizeof(double) * N);
        memcpy(pk, p, sizeof(double) * nr); memcpy(py, p, sizeof(double) * nr); memcpy(pk1, p, sizeof(double) * nr);
        double thk = 1.0, err_prev = INFINITY;
        int restarts = 0, momentum_disabled = 0;
        for (int iter = 0; iter < prm->max_iter; ++iter) {
            memcpy(x, xy, sizeof(double) * N);
            memcpy(p, py, sizeof(double) * nr);
            JSWEEP();
            memcpy(xk1, x, sizeof(double) * N);
            memcpy(pk1, p, sizeof(double) * nr);
            res->iters = iter + 1;
            const double err = pj_residual(nv, xk1, xy, prm->tol_rel, prm->tol_abs);
            res->residual = err;
            if (err <= 1) break;
            const double thk1 = 0.5 * (1.0 + sqrt(4.0 * thk * thk + 1.0));
            const double betak1 = (thk - 1.0) / thk1;
            int restart = 0;
            if (!isfinite(err)) restart = 1;
            else if (err_prev < INFINITY && err > 1.05 * err_prev) restart = 1;
            else if (betak1 > prm->nesterov_beta_threshold) restart = 1;
            else if (!isfinite(betak1) || betak1 < 0.0 || betak1 > 1.0) restart = 1;
            else {
                double dot = 0;
                for (int j = 0; j < nv; ++j) dot += (xy[j] - xk1[j]) * (xk1[j] - xk[j]);
                if (dot > 0.0) restart = 1;
            }
            if (restart || momentum_disabled) {
                memcpy(xy, xk1, sizeof(double) * N); memcpy(py, pk1, sizeof(double) * nr);
                thk = 1.0;
                if (restart && ++restarts >= prm->nesterov_restart_limit) momentum_disabled = 1;
            } else {
                for (int k = 0; k < N; ++k) xy[k] = xk1[k] + betak1 * (xk1[k] - xk[k]);
                for (int r = 0; r < nr; ++r) py[r] = pk1[r] + betak1 * (pk1[r] - pk[r]);
                thk = thk1;
            }
            memcpy(xk, xk1, sizeof(double) * N); memcpy(pk, pk1, sizeof(double) * nr);
            err_prev = err;
        }
        memcpy(x, xk1, sizeof(double) * N);
        memcpy(p, pk1, sizeof(double) * nr);
        free(xk); free(pk); free(xy); free(py); free(xk1); free(pk1);
    } else if (prm->chebyshev) {
        /* estimateSpectralRadius (:155-178) of v -> v - S^-1 [W^T R W v_head; 0] over the full
         * length of x, then chebyshev_loop (:180-228) with x = [v; Lambda] */
        double* v = (double*)malloc(sizeof(double) * ((size_t)N + 1));
        const double inv = 1.0 / sqrt((double)N);
        for (int k = 0; k < N; ++k) v[k] = inv;
        double lambda = 0.0;
        for (int it = 0; it < 8; ++it) {
            spmv(&RW, v, dl);
            memset(d, 0, sizeof(double) * (size_t)N);
            spmv_t_acc(&W, dl, d);
            lu_solve(&lu, d, dx);
            double n2 = 0.0;
            for (int k = 0; k < N; ++k) {
                v[k] = v[k] - dx[k];
                n2 += v[k] * v[k];
            }
            const double n = sqrt(n2);
            if (!(n > 1e-300)) { lambda = 0.0; break; }
            lambda = n;
            for (int k = 0; k < N; ++k) v[k] = v[k] / n;
        }
        free(v);
        const double rho = dmin(lambda, 0.995), rho2 = rho * rho;
        res->rho = rho;
        double* xk = (double*)malloc(sizeof(double) * (size_t)N);
        double* pk = (double*)malloc(sizeof(double) * (size_t)nr);
        double* xk1 = (double*)malloc(sizeof(double) * (size_t)N);
        double* pk1 = (double*)malloc(sizeof(double) * (size_t)nr);
        memcpy(xk, x, sizeof(double) * N); memcpy(xk1, x, sizeof(double) * N);
        memcpy(pk, p, sizeof(double) * nr); memcpy(pk1, p, sizeof(double) * nr);
        double omega = 1.0;
        for (int iter = 0; iter < prm->max_iter; ++iter) {
            memcpy(x, xk1, sizeof(double) * N);
            memcpy(p, pk1, sizeof(double) * nr);
            JSWEEP();
            res->iters = iter + 1;
            const double err = pj_residual(nv, x, xk1, prm->tol_rel, prm->tol_abs);
            res->residual = err;
            if (err <= 1) break;
            double omega_next = (iter == 0) ? (1.0 / (1.0 - rho2 * 0.5)) : (1.0 / (1.0 - (rho2 * 0.25) * omega));
            int finite = 1;
            for (int k = 0; k < N && finite; ++k) finite = isfinite(x[k]);
            if (!isfinite(omega_next) || omega_next <= 0.0 || !finite) {
                memcpy(xk, xk1, sizeof(double) * N); memcpy(pk, pk1, sizeof(double) * nr);
                memcpy(xk1, x, sizeof(double) * N); memcpy(pk1, p, sizeof(double) * nr);
                omega = 1.0;
            } else {
                for (int k = 0; k < N; ++k) { const double t = omega_next * (x[k] - xk[k]) + xk[k]; xk[k] = xk1[k]; xk1[k] = t; }
                for (int r = 0; r < nr; ++r) { const double t = omega_next * (p[r] - pk[r]) + pk[r]; pk[r] = pk1[r]; pk1[r] = t; }
                omega = omega_next;
            }
        }
        memcpy(x, xk1, sizeof(double) * N);
        memcpy(p, pk1, sizeof(double) * nr);
        free(xk); free(pk); free(xk1); free(pk1);
    } else {
        /* standard_loop (:371-377) */
        for (int iter = 0; iter < prm->max_iter; ++iter) {
            memcpy(xp, x, sizeof(double) * (size_t)nv);
            JSWEEP();
            res->iters = iter + 1;
            if (iter % 10 == 0) {
                const double e = pj_residual(nv, x, xp, prm->tol_rel, prm->tol_abs);
                res->residual = e;
                if (e <= 1) break;
            }
        }
    }
#undef JSWEEP
    for (int ci = 0; ci < C->nc; ++ci) {
        double* o = imp_out + 3 * (size_t)ci;
        o[0] = o[1] = o[2] = 0;
        const int base = W.row_base[ci];
        if (base < 0) continue;
        o[0] = p[base];
        if (W.row_size[ci] > 1) { o[1] = p[base + 1]; o[2] = p[base + 2]; }
    }
    memcpy(v_out, x, sizeof(double) * (size_t)nv);
    memcpy(lam_out, x + nv, sizeof(double) * (size_t)nj);
    free(Rr); free(bias); free(p); free(pp); free(dl); free(xp); free(d); free(dx); free(RWv);
    free(x); free(rhs); free(lu.a); free(lu.piv); free(M);
    orc_free_W(&W);
    return ORC_OK;
}

/* u_free = v + Minv .* (dt f_ext) (+ Minv .* (dt f_gyro) unless moreau.implicit_gyroscopy, which
 * the PGS assembler does not otherwise model: pgs_assembler.cpp:33-36,75-82) */
static void u_free_of(int nv, const orc_bodies* B, double dt, int implicit_gyro, double* uf) {
    for (int j = 0; j < nv; ++j) uf[j] = B->v[j] + B->minv[j] * (dt * B->fext[j]);
    if (!implicit_gyro)
        for (int j = 0; j < nv; ++j) uf[j] += B->minv[j] * (dt * B->fgyro[j]);
}

/* Residual on velocities u_free - u_corr (projected_gauss_seidel.cpp:124-131) */
static double gs_residual(int nv, const double* uf, const double* uk, const double* uk1, double tol_rel, double tol_abs) {
    double s = 0;
    for (int j = 0; j < nv; ++j) {
        const double d = uk1[j] - uk[j];
        const double v0 = uf[j] - uk[j], v1 = uf[j] - uk1[j];
        const double q = dmax(fabs(v0), fabs(v1)) * tol_rel + tol_abs;
        const double e = d / q;
        s += e * e;
    }
    return 1.0 / sqrt((double)nv) * sqrt(s);
}

/* PGS projection on one block, lambda <= 0 convention, correcting dlambda
 * (projected_gauss_seidel.cpp:11-26). */
static void pgs_project(double* lam, double* dl, double mu, int off, int n) {
    const double ln = lam[off];
    lam[off] = dmin(ln, 0.0);
    dl[off] += lam[off] - ln;
    if (n <= 1) return;
    const double t1 = lam[off + 1], t2 = lam[off + 2];
    const double tn = sqrt(t1 * t1 + t2 * t2);
    if (tn <= 0) return;
    const double s = dmin(1.0, -(mu * lam[off]) / tn);
    const double p1 = s * t1, p2 = s * t2;
    dl[off + 1] += p1 - t1;
    dl[off + 2] += p2 - t2;
    lam[off + 1] = p1;
    lam[off + 2] = p2;
}

/* PGS: ProjectedGaussSeidelSolver::solve (projected_gauss_seidel.cpp:56-268), standard loop,
 * with PgsAssembler::ufree/rhs/DinvDiag (pgs_assembler.cpp:12-82,220-288) for contact rows.
 * `colored` visits the blocks colour class by colour class (condensed_solver.cpp:220-251)
 * instead of in row order; the block update is unchanged. */
/* pgs_sweep (projected_gauss_seidel.cpp:144-169): the blocks in `visit` order (row order, or
 * colour by colour), each with r = rhs - W_j u_corr (C = 0 for contact rows), dl = (D^-1 r) *
 * relaxation * alpha, lambda += dl, the projection with dl corrected, u_corr += M^-1 W_j^T dl. */
static void pgs_sweep(const orc_W* W, const orc_bodies* B, const orc_params* prm, int nblk, const int* boff,
                      const int* visit, const double* Dinv, const double* rhs, double* lam, double* dl, double* uc) {
    for (int v = 0; v < nblk; ++v) {
        const int off = boff[visit[v]];
        const int bs = off < W->nnfc ? 1 : 3;
        for (int k = 0; k < bs; ++k) {
            const int r = off + k;
            double s = 0;
            for (int q = W->rowptr[r]; q < W->rowptr[r + 1]; ++q) s += W->val[q] * uc[W->col[q]];
            const double rr = rhs[r] - s;
            dl[r] = (Dinv[r] * rr) * prm->relaxation;
        }
        for (int k = 0; k < bs; ++k) dl[off + k] *= prm->alpha;
        for (int k = 0; k < bs; ++k) lam[off + k] += dl[off + k];
        pgs_project(lam, dl, W->mu[off], off, bs);
        for (int k = 0; k < bs; ++k) {
            const int r = off + k;
            for (int q = W->rowptr[r]; q < W->rowptr[r + 1]; ++q) uc[W->col[q]] += (B->minv[W->col[q]] * W->val[q]) * dl[r];
        }
    }
}

static int pgs_core(const orc_bodies* B, const orc_contacts* C, double dt, const orc_params* prm, int colored,
                    double* v_out, double* imp_out, orc_result* res) {
    orc_W W;
    int st = orc_build_W(B, C, &W);
    if (st) { orc_free_W(&W); return st; }
    const int nv = W.nv, nr = W.nrows;
    res->iters = 0; res->residual = NAN; res->nrows = nr; res->ncolors = 0; res->rho = NAN;
    double* uf = (double*)malloc(sizeof(double) * (size_t)(nv + 1));
    u_free_of(nv, B, dt, prm->implicit_gyro, uf);
    if (nr == 0) {
        memcpy(v_out, uf, sizeof(double) * (size_t)nv);
        for (int c = 0; c < C->nc; ++c) imp_out[3 * c] = imp_out[3 * c + 1] = imp_out[3 * c + 2] = 0;
        free(uf); orc_free_W(&W);
        return ORC_OK;
    }
    double* D = (double*)malloc(sizeof(double) * (size_t)nr);
    double* Dinv = (double*)malloc(sizeof(double) * (size_t)nr);
    double* rhs = (double*)malloc(sizeof(double) * (size_t)nr);
    double* lam = (double*)calloc((size_t)nr, sizeof(double));
    double* dl = (double*)calloc((size_t)nr, sizeof(double));
    double* uc = (double*)calloc((size_t)nv, sizeof(double));
    double* uk = (double*)malloc(sizeof(double) * (size_t)nv);
    delassus_diag(&W, B->minv, D);
    for (int r = 0; r < W.nnfc; ++r) Dinv[r] = D[r] > 0 ? 1.0 / D[r] : 0.0; /* :277-280 */
    for (int r = W.nnfc; r < nr; ++r) Dinv[r] = 1.0 / D[r];                   /* :282-285 (no guard) */
    /* rhs = W u_free + c (+ e .* (W v + c) when any e != 0) (pgs_assembler.cpp:61-70) */
    int any_rest = 0;
    for (int r = 0; r < nr; ++r) if (W.rest[r] != 0.0) any_rest = 1;
    double* bi = (double*)malloc(sizeof(double) * (size_t)nr);
    for (int r = 0; r < nr; ++r) bi[r] = W.cvec[r];
    if (any_rest) {
        spmv(&W, B->v, rhs);
        for (int r = 0; r < nr; ++r) rhs[r] += W.cvec[r];
        for (int r = 0; r < nr; ++r) bi[r] += W.rest[r] * rhs[r];
    }
    spmv(&W, uf, rhs);
    for (int r = 0; r < nr; ++r) rhs[r] += bi[r];
    free(bi);
    /* warm start with inverted normal (:95-101) */
    if (prm->warmstart)
        for (int ci = 0; ci < C->nc; ++ci) {
            const int base = W.row_base[ci];
            if (base < 0) continue;
            lam[base] = -C->warm[3 * ci];
            if (W.row_size[ci] > 1) { lam[base + 1] = C->warm[3 * ci + 1]; lam[base + 2] = C->warm[3 * ci + 2]; }
        }
    /* u_corr = (Minv W^T) lambda (:81,108) */
    for (int r = 0; r < nr; ++r)
        for (int k = W.rowptr[r]; k < W.rowptr[r + 1]; ++k) uc[W.col[k]] += (B->minv[W.col[k]] * W.val[k]) * lam[r];
    memcpy(uk, uc, sizeof(double) * (size_t)nv);

    /* block list: row offsets in row order (1-row frictionless, 3-row frictional) */
    const int nblk = W.nnfc + (nr - W.nnfc) / 3;
    int* boff = (int*)malloc(sizeof(int) * ((size_t)nblk + 1));
    for (int k = 0, off = 0; k < nblk; ++k) { boff[k] = off; off += off < W.nnfc ? 1 : 3; }
    int* visit = (int*)malloc(sizeof(int) * ((size_t)nblk + 1));
    for (int k = 0; k < nblk; ++k) visit[k] = k;
    if (colored) {
        /* blocks in row order are the contacts with rows, frictionless first */
        int* ba = (int*)malloc(sizeof(int) * ((size_t)nblk + 1));
        int* bb = (int*)malloc(sizeof(int) * ((size_t)nblk + 1));
        for (int ci = 0; ci < C->nc; ++ci) {
            const int base = W.row_base[ci];
            if (base < 0) continue;
            const int k = base < W.nnfc ? base : W.nnfc + (base - W.nnfc) / 3;
            ba[k] = C->body_a[ci]; bb[k] = C->body_b[ci];
        }
        int* col = (int*)malloc(sizeof(int) * ((size_t)nblk + 1));
        const int ncol = orc_color_blocks(B->nb, nblk, ba, bb, col);
        res->ncolors = ncol;
        int m = 0;
        for (int c = 0; c < ncol; ++c)
            for (int k = 0; k < nblk; ++k)
                if (col[k] == c) visit[m++] = k;
        free(ba); free(bb); free(col);
    }

    if (!prm->nesterov) {
        for (int iter = 0; iter < prm->max_iter; ++iter) {
            pgs_sweep(&W, B, prm, nblk, boff, visit, Dinv, rhs, lam, dl, uc);
            res->iters = iter + 1;
            const double e = gs_residual(nv, uf, uc, uk, prm->tol_rel, prm->tol_abs);
            res->residual = e;
            if (isnan(e) || isinf(e)) { st = ORC_DIVERGED; break; } /* :133-138 */
            if (e < 1.0) break;
            memcpy(uk, uc, sizeof(double) * (size_t)nv);
        }
    } else {
        /* Nesterov outer loop (:196-262) on (lambda, u_corr); the divergence test of
         * check_residual (:133-138) precedes the restart rules, so a non-finite residual ends the
         * solve.  The ORC_MUT_NES_* macros plant the same mistakes as in the PJ loop. */
        double* lk = (double*)malloc(sizeof(double) * (size_t)nr);
        double* ly = (double*)malloc(sizeof(double) * (size_t)nr);
        double* lk1 = (double*)malloc(sizeof(double) * (size_t)nr);
        double* uy = (double*)malloc(sizeof(double) * (size_t)nv);
        double* uk1 = (double*)malloc(sizeof(double) * (size_t)nv);
        memcpy(lk, lam, sizeof(double) * nr); memcpy(ly, lam, sizeof(double) * nr);
        memcpy(uy, uc, sizeof(double) * nv); /* uk holds u_corr already */
        double thk = 1.0, err_prev = INFINITY;
        int restarts = 0, momentum_disabled = 0;
        for (int iter = 0; iter < prm->max_iter; ++iter) {
            memcpy(lam, ly, sizeof(double) * nr);
            memcpy(uc, uy, sizeof(double) * nv);
            pgs_sweep(&W, B, prm, nblk, boff, visit, Dinv, rhs, lam, dl, uc);
            memcpy(lk1, lam, sizeof(double) * nr);
            memcpy(uk1, uc, sizeof(double) * nv);
            res->iters = iter + 1;
            const double err = gs_residual(nv, uf, uy, uk1, prm->tol_rel, prm->tol_abs);
            res->residual = err;
            if (isnan(err) || isinf(err)) { st = ORC_DIVERGED; break; }
            if (err <= 1) break;
#if defined(ORC_MUT_NES_THETA)
            const double thk1 = 0.5 * (1.0 + sqrt(thk * thk + 1.0));
#else
            const double thk1 = 0.5 * (1.0 + sqrt(4.0 * thk * thk + 1.0));
#endif
#if defined(ORC_MUT_NES_BETA)
            const double betak1 = thk / thk1;
#else
            const double betak1 = (thk - 1.0) / thk1;
#endif
            int restart = 0;
#if defined(ORC_MUT_NES_GROWTH)
            if (err_prev < INFINITY && err > 1.5 * err_prev) restart = 1;
#else
            if (err_prev < INFINITY && err > 1.05 * err_prev) restart = 1;
#endif
            else if (betak1 > prm->nesterov_beta_threshold) restart = 1;
            else if (!isfinite(betak1) || betak1 < 0.0 || betak1 > 1.0) restart = 1;
            else {
                double dot = 0;
                for (int j = 0; j < nv; ++j) dot += (uy[j] - uk1[j]) * (uk1[j] - uk[j]);
#if defined(ORC_MUT_NES_DIRECTION)
                if (dot < 0.0) restart = 1;
#else
                if (dot > 0.0) restart = 1;
#endif
            }
            if (restart) {
                memcpy(ly, lk1, sizeof(double) * nr); memcpy(uy, uk1, sizeof(double) * nv);
                thk = 1.0;
#if defined(ORC_MUT_NES_LIMIT)
                if (++restarts > prm->nesterov_restart_limit) momentum_disabled = 1;
#else
                if (++restarts >= prm->nesterov_restart_limit) momentum_disabled = 1;
#endif
            } else if (momentum_disabled) {
                memcpy(ly, lk1, sizeof(double) * nr); memcpy(uy, uk1, sizeof(double) * nv);
                thk = 1.0;
            } else {
                for (int r = 0; r < nr; ++r) ly[r] = lk1[r] + betak1 * (lk1[r] - lk[r]);
#if defined(ORC_MUT_NES_ANCHOR)
                for (int j = 0; j < nv; ++j) uy[j] = uk1[j] + betak1 * (uk1[j] - uy[j]);
#else
                for (int j = 0; j < nv; ++j) uy[j] = uk1[j] + betak1 * (uk1[j] - uk[j]);
#endif
                thk = thk1;
            }
            memcpy(lk, lk1, sizeof(double) * nr); memcpy(uk, uk1, sizeof(double) * nv);
            err_prev = err;
        }
        free(lk); free(ly); free(lk1); free(uy); free(uk1);
    }
    for (int ci = 0; ci < C->nc; ++ci) {
        double* o = imp_out + 3 * (size_t)ci;
        o[0] = o[1] = o[2] = 0;
        const int base = W.row_base[ci];
        if (base < 0) continue;
        o[0] = -lam[base];
        if (W.row_size[ci] > 1) { o[1] = lam[base + 1]; o[2] = lam[base + 2]; }
    }
    for (int j = 0; j < nv; ++j) v_out[j] = uf[j] - uc[j]; /* :267 */
    free(uf); free(D); free(Dinv); free(rhs); free(lam); free(dl); free(uc); free(uk); free(boff); free(visit);
    orc_free_W(&W);
    return st;
}

int orc_pgs_solve(const orc_bodies* B, const orc_contacts* C, double dt, double theta, const orc_params* prm,
                  double* v_out, double* imp_out, orc_result* res) {
    (void)theta;
    return pgs_core(B, C, dt, prm, 0, v_out, imp_out, res);
}

int orc_colored_gs_solve(const orc_bodies* B, const orc_contacts* C, double dt, double theta, const orc_params* prm,
                         double* v_out, double* imp_out, orc_result* res) {
    (void)theta;
    return pgs_core(B, C, dt, prm, 1, v_out, imp_out, res);
}

/* ---------------------------------------------------------------------------------------- */
/* Colouring for the coloured GS: condensed_solver.cpp:195-215, graph_coloring.cpp:8-41.  */
/* ---------------------------------------------------------------------------------------- */

typedef struct {
    int deg, idx;
} degidx;

static int degidx_cmp(const void* pa, const void* pb) {
    const degidx* a = (const degidx*)pa;
    const degidx* b = (const degidx*)pb;
    if (a->deg != b->deg) return a->deg > b->deg ? -1 : 1; /* descending degree */
    return a->idx < b->idx ? -1 : (a->idx > b->idx); /* ties: ascending index (stable) */
}

static int int_cmp(const void* a, const void* b) {
    const int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

/* Block adjacency = blocks sharing a body (buildBlockAdjacency, condensed_solver.cpp:195-215),
 * greedy colouring in descending-degree order (graph_coloring.cpp:8-41).  The reference sorts
 * with std::sort (unstable); ties here are broken by block index. */
int orc_color_blocks(int nb, int nc, const int* body_a, const int* body_b, int* color_of) {
    /* incidence lists */
    int* cnt = (int*)calloc((size_t)nb + 1, sizeof(int));
    for (int c = 0; c < nc; ++c) {
        if (body_a[c] >= 0) cnt[body_a[c] + 1]++;
        if (body_b[c] >= 0) cnt[body_b[c] + 1]++;
    }
    for (int b = 0; b < nb; ++b) cnt[b + 1] += cnt[b];
    int* inc = (int*)malloc(sizeof(int) * ((size_t)cnt[nb] + 1));
    int* fill = (int*)calloc((size_t)nb + 1, sizeof(int));
    for (int c = 0; c < nc; ++c) {
        if (body_a[c] >= 0) inc[cnt[body_a[c]] + fill[body_a[c]]++] = c;
        if (body_b[c] >= 0) inc[cnt[body_b[c]] + fill[body_b[c]]++] = c;
    }
    /* adjacency as sorted unique lists */
    int* adjcnt = (int*)calloc((size_t)nc + 1, sizeof(int));
    for (int b = 0; b < nb; ++b) {
        const int d = cnt[b + 1] - cnt[b];
        for (int i = 0; i < d; ++i) adjcnt[inc[cnt[b] + i] + 1] += d - 1;
    }
    for (int c = 0; c < nc; ++c) adjcnt[c + 1] += adjcnt[c];
    int* adj = (int*)malloc(sizeof(int) * ((size_t)adjcnt[nc] + 1));
    int* afill = (int*)calloc((size_t)nc + 1, sizeof(int));
    for (int b = 0; b < nb; ++b) {
        const int s = cnt[b], d = cnt[b + 1] - cnt[b];
        for (int i = 0; i < d; ++i)
            for (int j = 0; j < d; ++j)
                if (i != j) { const int u = inc[s + i]; adj[adjcnt[u] + afill[u]++] = inc[s + j]; }
    }
    int* deg = (int*)calloc((size_t)nc + 1, sizeof(int));
    for (int c = 0; c < nc; ++c) {
        int* a = adj + adjcnt[c];
        const int n = afill[c];
        qsort(a, (size_t)n, sizeof(int), int_cmp);
        int u = 0;
        for (int i = 0; i < n; ++i)
            if (u == 0 || a[u - 1] != a[i]) a[u++] = a[i];
        deg[c] = u;
    }
    degidx* order = (degidx*)malloc(sizeof(degidx) * ((size_t)nc + 1));
    for (int c = 0; c < nc; ++c) { order[c].deg = deg[c]; order[c].idx = c; color_of[c] = -1; }
    qsort(order, (size_t)nc, sizeof(degidx), degidx_cmp);
    int* stamp = (int*)malloc(sizeof(int) * ((size_t)nc + 2));
    for (int i = 0; i < nc + 2; ++i) stamp[i] = -1;
    int ncolors = 0;
    for (int k = 0; k < nc; ++k) {
        const int node = order[k].idx;
        for (int i = 0; i < deg[node]; ++i) {
            const int cc = color_of[adj[adjcnt[node] + i]];
            if (cc >= 0) stamp[cc] = k;
        }
        int chosen = 0;
        while (stamp[chosen] == k) ++chosen;
        color_of[node] = chosen;
        if (chosen + 1 > ncolors) ncolors = chosen + 1;
    }
    free(cnt); free(inc); free(fill); free(adjcnt); free(adj); free(afill); free(deg); free(order); free(stamp);
    return ncolors;
}

/* ---------------------------------------------------------------------------------------- */
/* Moreau drift: integration_base.cpp:6-30, math_helper.hpp:11-17.                          */
/* ---------------------------------------------------------------------------------------- */
void orc_position_update(int nb, const int* dof, double* q, const double* v, double h) {
    int qo = 0, vo = 0;
    for (int b = 0; b < nb; ++b) {
        double* x = q + qo;
        const double* u = v + vo;
        x[0] += h * u[0]; x[1] += h * u[1]; x[2] += h * u[2];
        if (dof[b] == 6) {
            double* P = x + 3; /* (x, y, z, w) coefficient order */
            const double qp[4] = {P[0], P[1], P[2], P[3]};
            const double* w = u + 3;
            const double w0 = P[3];
            P[3] -= h * 0.5 * (P[0] * w[0] + P[1] * w[1] + P[2] * w[2]);
            double cr[3];
            orc_cross(P, w, cr);
            const double n0 = P[0] + h * 0.5 * (w0 * w[0] + cr[0]);
            const double n1 = P[1] + h * 0.5 * (w0 * w[1] + cr[1]);
            const double n2 = P[2] + h * 0.5 * (w0 * w[2] + cr[2]);
            P[0] = n0; P[1] = n1; P[2] = n2;
            /* alignQuaternionTo: non-finite -> previous; normalize; flip to q_prev hemisphere */
            if (!(isfinite(P[0]) && isfinite(P[1]) && isfinite(P[2]) && isfinite(P[3]))) {
                P[0] = qp[0]; P[1] = qp[1]; P[2] = qp[2]; P[3] = qp[3];
            } else {
                const double n2s = P[0] * P[0] + P[1] * P[1] + P[2] * P[2] + P[3] * P[3];
                if (n2s > 0) {
                    const double nn = sqrt(n2s);
                    P[0] /= nn; P[1] /= nn; P[2] /= nn; P[3] /= nn;
                }
                const double dt = qp[0] * P[0] + qp[1] * P[1] + qp[2] * P[2] + qp[3] * P[3];
                if (dt < 0) { P[0] = -P[0]; P[1] = -P[1]; P[2] = -P[2]; P[3] = -P[3]; }
            }
            qo += 7;
        } else {
            qo += 3;
        }
        vo += dof[b];
    }
}

/* f_gyro = -omega x (I omega) with body-frame omega, I = diag (world.cpp:152-173). */
void orc_gyro_force(int nb, const int* dof, const double* minv, const double* v, double* f) {
    int vo = 0;
    for (int b = 0; b < nb; ++b) {
        for (int j = 0; j < dof[b]; ++j) f[vo + j] = 0.0;
        if (dof[b] == 6) {
            const double* w = v + vo + 3;
            double I[3], Iw[3], c[3];
            for (int j = 0; j < 3; ++j) I[j] = minv[vo + 3 + j] > 0 ? 1.0 / minv[vo + 3 + j] : 0.0;
            for (int j = 0; j < 3; ++j) Iw[j] = I[j] * w[j];
            orc_cross(w, Iw, c);
            for (int j = 0; j < 3; ++j) f[vo + 3 + j] = -c[j];
        }
        vo += dof[b];
    }
}

/* Natural-map complementarity residual with diagonal Delassus scaling:
 * ||p - Pi_K(p - D^-1 (W v + c + e .* (W v_n + c)))|| / max(||p||, tiny).  Fixed points of PJ,
 * PGS and colored GS (all of which use a positive multiple of D^-1) are its zeros. */
double orc_complementarity(const orc_bodies* B, const orc_contacts* C, const double* v, const double* imp) {
    orc_W W;
    if (orc_build_W(B, C, &W)) { orc_free_W(&W); return NAN; }
    const int nr = W.nrows;
    double* g = (double*)malloc(sizeof(double) * ((size_t)nr + 1));
    double* g0 = (double*)malloc(sizeof(double) * ((size_t)nr + 1));
    double* D = (double*)malloc(sizeof(double) * ((size_t)nr + 1));
    spmv(&W, v, g);
    spmv(&W, B->v, g0);
    delassus_diag(&W, B->minv, D);
    double num = 0, den = 0;
    for (int ci = 0; ci < C->nc; ++ci) {
        const int base = W.row_base[ci];
        if (base < 0) continue;
        const int n = W.row_size[ci];
        double p[3] = {imp[3 * ci], n > 1 ? imp[3 * ci + 1] : 0, n > 1 ? imp[3 * ci + 2] : 0};
        double z[3];
        for (int k = 0; k < n; ++k) {
            const int r = base + k;
            const double gam = g[r] + W.cvec[r] + W.rest[r] * (g0[r] + W.cvec[r]);
            z[k] = p[k] - (D[r] > 0 ? gam / D[r] : 0.0);
        }
        z[0] = dmax(z[0], 0.0);
        if (n > 1) {
            const double tn = sqrt(z[1] * z[1] + z[2] * z[2]);
            if (tn > 0) {
                const double s = dmin(1.0, W.mu[base] * z[0] / tn);
                z[1] *= s; z[2] *= s;
            }
        }
        for (int k = 0; k < n; ++k) { num += (p[k] - z[k]) * (p[k] - z[k]); den += p[k] * p[k]; }
    }
    free(g); free(g0); free(D);
    orc_free_W(&W);
    return sqrt(num) / dmax(sqrt(den), 1e-300);
}

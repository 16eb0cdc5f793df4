// cdx_pj_device.hpp -- per-contact and per-body work of one projected-Jacobi solve, shared by the
// grid-wide kernels (cdx_pj.hip) and the batched persistent kernel (cdx_batch.hip) so both run
// the same arithmetic.
#pragma once

#include "cdx_device.hpp"

namespace cdx {

// Entry (k, l) of the contact's Delassus block W_c M^-1 W_c^T: sums over each side's columns in
// column order, the side with the lower body index first (as row_D).
__device__ inline double block_entry(const ContactJ& j, int k, int l, double4 mA, double4 mB, int a, int b) {
    double DA = 0, DB = 0;
#pragma unroll
    for (int m = 0; m < 3; ++m) DA += j.d[k][m] * j.d[l][m] * mA.x;
    DA += j.aA[k][0] * j.aA[l][0] * mA.y;
    DA += j.aA[k][1] * j.aA[l][1] * mA.z;
    DA += j.aA[k][2] * j.aA[l][2] * mA.w;
#pragma unroll
    for (int m = 0; m < 3; ++m) DB += j.d[k][m] * j.d[l][m] * mB.x;
    DB += j.aB[k][0] * j.aB[l][0] * mB.y;
    DB += j.aB[k][1] * j.aB[l][1] * mB.z;
    DB += j.aB[k][2] * j.aB[l][2] * mB.w;
    return (a >= 0 && b >= 0 && b < a) ? DB + DA : DA + DB;
}

// Inverse of an SPD 3x3 block by Cholesky and two triangular solves per column (invertSmallSpd,
// block_diagonal.hpp:15-27); the inverse of the diagonal when a pivot is not positive.
__device__ inline void spd_inverse3(const double A[3][3], double Ai[3][3]) {
    double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double d = A[k][k];
        for (int j = 0; j < k; ++j) d -= L[k][j] * L[k][j];
        ok = ok && d > 0;
        L[k][k] = sqrt(fmax(d, 0.0));
        for (int i = k + 1; i < 3; ++i) {
            double t = A[i][k];
            for (int j = 0; j < k; ++j) t -= L[i][j] * L[k][j];
            L[i][k] = t / L[k][k];
        }
    }
    if (!ok) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) Ai[i][j] = i == j ? 1.0 / A[i][i] : 0.0;
        return;
    }
#pragma unroll
    for (int col = 0; col < 3; ++col) {
        double y[3], x[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double t = i == col ? 1.0 : 0.0;
            for (int j = 0; j < i; ++j) t -= L[i][j] * y[j];
            y[i] = t / L[i][i];
        }
#pragma unroll
        for (int i = 2; i >= 0; --i) {
            double t = y[i];
            for (int j = i + 1; j < 3; ++j) t -= L[j][i] * x[j];
            x[i] = t / L[i][i];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) Ai[i][col] = x[i];
    }
}

// Per-contact setup of one PJ solve (see k_pj_setup_contacts).  With Rb (block mode,
// pj.rdiag_true_delassus, projected_jacobi.cpp:330-348) the bias stays unscaled and Rb receives
// alpha (W_c M^-1 W_c^T)^-1 (buildBlockPreconditioner, :81-98): the 1x1 block of a frictionless
// row or the 3x3 block of a frictional contact, entry (k, l) at (3k + l) nc + c.
__device__ inline void pj_setup_contact(int c, int nc, const int* __restrict__ ida, const int* __restrict__ idb,
                                        const double* __restrict__ J, const double* __restrict__ mu,
                                        const double* __restrict__ rest, const double* __restrict__ vb,
                                        const double* __restrict__ minv, const double* __restrict__ v, double alpha,
                                        int warmstart, double* __restrict__ R, double* __restrict__ bias,
                                        double* __restrict__ p, double* __restrict__ imp,
                                        double* __restrict__ Rb = nullptr) {
    const int a = ida[c], b = idb[c];
    if (Rb)
        for (int k = 0; k < 9; ++k) Rb[(size_t)k * nc + c] = 0;
    if (a < 0 && b < 0) {
        for (int k = 0; k < 3; ++k) {
            R[(size_t)k * nc + c] = 0;
            bias[(size_t)k * nc + c] = 0;
            p[(size_t)k * nc + c] = 0;
        }
        return;
    }
    const int nr = mu[c] > 0 ? 3 : 1;
    ContactJ j;
    load_J(J, nc, c, nr, j);
    double ua[6], ub[6];
    load_body(v, a, ua);
    load_body(v, b, ub);
    const double4 mA = load_minv4(minv, a), mB = load_minv4(minv, b);
    double pw[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) {
        if (k >= nr) {
            R[(size_t)k * nc + c] = 0;
            bias[(size_t)k * nc + c] = 0;
            p[(size_t)k * nc + c] = 0;
            continue;
        }
        const double D = row_D(j.d[k], j.aA[k], j.aB[k], mA, mB, a, b);
        const double Rk = D > 0 ? alpha / D : 0.0;
        const double ck = vb[(size_t)k * nc + c];
        const double vprev = row_dot(j.d[k], j.aA[k], j.aB[k], ua, ub) + ck;
        const double bk = ck + rest[(size_t)k * nc + c] * vprev;
        R[(size_t)k * nc + c] = Rk;
        bias[(size_t)k * nc + c] = Rb ? bk : Rk * bk;
        pw[k] = warmstart ? p[(size_t)k * nc + c] : 0.0;
        p[(size_t)k * nc + c] = pw[k];
    }
    if (Rb) {
        const double inv_alpha = alpha > 0 ? 1.0 / alpha : 0.0;
        if (nr == 1) {
            Rb[c] = 1.0 / (block_entry(j, 0, 0, mA, mB, a, b) * inv_alpha);
        } else {
            double A[3][3], Ai[3][3];
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int l = 0; l < 3; ++l) A[k][l] = block_entry(j, k, l, mA, mB, a, b) * inv_alpha;
            spd_inverse3(A, Ai);
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int l = 0; l < 3; ++l) Rb[(size_t)(3 * k + l) * nc + c] = Ai[k][l];
        }
    }
    write_side_impulses9(imp, c, a, b, nr, j, pw);
}

// Per-body x0 (see k_pj_setup_bodies).
// Implicit gyroscopy (moreau.implicit_gyroscopy, pj_assembler.cpp:32-61): the inverse of the
// rotational block diag(I) - dt G_rot(w) of S for a rigid body, G_rot = 1/2 ([I w]_x - [w]_x diag(I)),
// I = 1 / minv[1..3] of the body, w = v[3..5]; inverted by its adjugate into K (row-major).
// Returns false for a singular or non-finite block (S not factorized).
__device__ inline bool gyro_block_inverse(const double* __restrict__ mi, const double* __restrict__ w, double dt,
                                          double K[9]) {
    double I[3], Iw[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        I[k] = mi[1 + k] > 0 ? 1.0 / mi[1 + k] : 0.0;
        Iw[k] = I[k] * w[k];
    }
    const double sw[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    const double sIw[9] = {0, -Iw[2], Iw[1], Iw[2], 0, -Iw[0], -Iw[1], Iw[0], 0};
    double S[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double corr = -dt * (0.5 * (sIw[3 * r + c] - sw[3 * r + c] * I[c]));
            S[3 * r + c] = r == c ? I[r] + corr : corr;
        }
    const double c0 = S[4] * S[8] - S[5] * S[7], c1 = S[5] * S[6] - S[3] * S[8], c2 = S[3] * S[7] - S[4] * S[6];
    const double det = S[0] * c0 + S[1] * c1 + S[2] * c2;
    if (!(det != 0.0) || !isfinite(det)) return false;
    K[0] = c0 / det;
    K[1] = (S[2] * S[7] - S[1] * S[8]) / det;
    K[2] = (S[1] * S[5] - S[2] * S[4]) / det;
    K[3] = c1 / det;
    K[4] = (S[0] * S[8] - S[2] * S[6]) / det;
    K[5] = (S[2] * S[3] - S[0] * S[5]) / det;
    K[6] = c2 / det;
    K[7] = (S[1] * S[6] - S[0] * S[7]) / det;
    K[8] = (S[0] * S[4] - S[1] * S[3]) / det;
    return true;
}

// Row r of a rigid body's inverse rotational block times (a3, a4, a5): the rotational part of
// S^-1 a with implicit gyroscopy (the other DOFs take a_k / M_k).
__device__ __forceinline__ double gyro_row(const double* __restrict__ K, double a3, double a4, double a5) {
    return K[0] * a3 + K[1] * a4 + K[2] * a5;
}

// sinv: with implicit gyroscopy the bodies' inverse rotational blocks (9 doubles per body, zero for
// bodies without 6 DOFs), and dt f_gyro leaves the right-hand side; nullptr otherwise.
__device__ inline void pj_setup_body(int bi, const int* __restrict__ dof, const double* __restrict__ minv,
                                     const double* __restrict__ v, const double* __restrict__ fext,
                                     const double* __restrict__ fgyro, const int* __restrict__ incptr,
                                     const int* __restrict__ incidx, const double* __restrict__ imp, double dt,
                                     int have_contacts, double* __restrict__ x, const int* __restrict__ jslot = nullptr,
                                     double* __restrict__ tr = nullptr, const double* __restrict__ sinv = nullptr) {
    const int nd = dof[bi];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (have_contacts)
        for (int i = incptr[bi]; i < incptr[bi + 1]; ++i) gather_side9(imp, incidx[i], acc);
    const double* vv = v + (size_t)bi * kBodyStride;
    const double* fe = fext + (size_t)bi * kBodyStride;
    const double* fg = fgyro + (size_t)bi * kBodyStride;
    const double* mi = minv + (size_t)bi * kMinvStride;
    double out[6] = {0, 0, 0, 0, 0, 0}, rhs[6] = {0, 0, 0, 0, 0, 0};
    const bool jointed = jslot && jslot[bi] >= 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if (k >= nd) break;
        const double mk = k < 3 ? mi[0] : mi[k - 2];
        const double M = mk > 0 ? 1.0 / mk : 0.0;
        double r = M * vv[k] + dt * fe[k];
        if (!sinv) r += dt * fg[k];
        r += acc[k];
        rhs[k] = r;
        out[k] = jointed ? r : r / M;
    }
    if (sinv && nd == 6 && !jointed) {  // the rotational block of S^-1
        const double* K = sinv + (size_t)bi * 9;
#pragma unroll
        for (int k = 0; k < 3; ++k) out[3 + k] = gyro_row(K + 3 * k, rhs[3], rhs[4], rhs[5]);
    }
    // a jointed body's right-hand side goes to the chain solve, which writes its x0
    if (jointed) store6(tr + (size_t)bi * 6, out);
    else store6(x + (size_t)bi * kBodyStride, out);
}

// w (x - xk) + xk with separate roundings, as the reference's vector expression (the Chebyshev
// extrapolation, projected_jacobi.cpp:208-209)
__device__ inline double cheb_extrap(double w, double x, double xk) {
#pragma clang fp contract(off)
    return w * (x - xk) + xk;
}

// Body b's twist as the input of a Chebyshev sweep: X, or w (X - XK) + XK (zeros for a static side).
__device__ inline void cheb_input6(const double* __restrict__ X, const double* __restrict__ XK, int b, double w, bool take,
                                   double u[6]) {
    load_body(X, b, u);
    if (take || b < 0) return;
    double k6[6];
    load_body(XK, b, k6);
#pragma unroll
    for (int m = 0; m < 6; ++m) u[m] = cheb_extrap(w, u[m], k6[m]);
}

// The next Chebyshev weight after sweep `it` (projected_jacobi.cpp:202): 1 / (1 - rho^2 / 2) after
// the first sweep, 1 / (1 - rho^2 / 4 w) after later ones.
__device__ inline double cheb_next_w(int it, double rho2, double w) {
#pragma clang fp contract(off)
    return it == 0 ? 1.0 / (1.0 - rho2 * 0.5) : 1.0 / (1.0 - (rho2 * 0.25) * w);
}

// p_k - sum_l Rb(k, l) t_l over the nr rows of contact c (block mode: the inverse Delassus block,
// entry (k, l) at (3k + l) nc + c; BlockDiagonal::applyTo, block_diagonal.hpp:103-124), in place
// of the scaled rows p_k - t_k.
__device__ inline void block_apply_rows(int c, int nc, int nr, const double* __restrict__ Rb, const double pold[3],
                                        double t[3]) {
    const double t0 = t[0], t1 = t[1], t2 = t[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= nr) break;
        const double* r = Rb + (size_t)(3 * k) * nc + c;
        t[k] = pold[k] - (nr == 1 ? r[0] * t0 : r[0] * t0 + r[nc] * t1 + r[2 * (size_t)nc] * t2);
    }
}

// One contact of a PJ sweep (pj_sweep, projected_jacobi.cpp:121-138, with project(), :100-119).
// The bias is only read when the problem has one (`use_bias`: some restitution or velocity bias).
// With Rb (block mode) R is unused and the step is p - Rb (W x + b).
__device__ inline void pj_contact_sweep(int c, int nc, const int* __restrict__ ida, const int* __restrict__ idb,
                                        const double* __restrict__ J, const double* __restrict__ R,
                                        const double* __restrict__ bias, int use_bias, const double* __restrict__ mu,
                                        double* __restrict__ p, const double* __restrict__ x,
                                        double* __restrict__ imp, const double* __restrict__ Rb = nullptr) {
    const int a = ida[c], b = idb[c];
    if (a < 0 && b < 0) return;
    const double muc = mu[c];
    const int nr = muc > 0 ? 3 : 1;
    ContactJ j;
    load_J(J, nc, c, nr, j);
    double ua[6], ub[6];
    load_body(x, a, ua);
    load_body(x, b, ub);
    double pold[3] = {0, 0, 0}, pn[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= nr) break;
        const double g = row_dot(j.d[k], j.aA[k], j.aB[k], ua, ub);
        pold[k] = p[(size_t)k * nc + c];
        if (Rb)
            pn[k] = g + (use_bias ? bias[(size_t)k * nc + c] : 0.0);
        else
            pn[k] = pold[k] - (R[(size_t)k * nc + c] * g + (use_bias ? bias[(size_t)k * nc + c] : 0.0));
    }
    if (Rb) block_apply_rows(c, nc, nr, Rb, pold, pn);
    // project() (projected_jacobi.cpp:100-119)
    pn[0] = fmax(pn[0], 0.0);
    if (nr == 3) {
        const double t1 = pn[1], t2 = pn[2];
        const double tn2 = t1 * t1 + t2 * t2;
        if (tn2 > 0) {
            const double s = fmin(1.0, (muc * pn[0]) / sqrt(tn2));
            pn[1] = s * t1;
            pn[2] = s * t2;
        }
    }
    double dp[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= nr) break;
        dp[k] = pn[k] - pold[k];
        p[(size_t)k * nc + c] = pn[k];
    }
    write_side_impulses9(imp, c, a, b, nr, j, dp);
}

// One body of a PJ sweep: x += (W^T dp) / M over its incidences in a fixed order; returns the
// body's part of the residual sum (projected_jacobi.cpp:140-144) when `check`.  A jointed body
// (jslot >= 0) instead hands W^T dp to the chain solve through tr, which updates its x.
__device__ inline double pj_body_sweep(int bi, const int* __restrict__ dof, const double* __restrict__ minv,
                                       const int* __restrict__ incptr, const int* __restrict__ incidx,
                                       const double* __restrict__ imp, double* __restrict__ x, int check,
                                       double tol_rel, double tol_abs, const int* __restrict__ jslot = nullptr,
                                       double* __restrict__ tr = nullptr) {
    double e2 = 0;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const int i0 = incptr[bi], i1 = incptr[bi + 1];
    for (int i = i0; i < i1; ++i) gather_side9(imp, incidx[i], acc);
    if (jslot && jslot[bi] >= 0) {
        store6(tr + (size_t)bi * 6, acc);
        return 0.0;
    }
    double u[6];
    load_body(x, bi, u);
    const double4 m4 = *reinterpret_cast<const double4*>(minv + (size_t)bi * kMinvStride);
    const int nd = dof[bi];
    double un[6];
    const double Ml = m4.x > 0 ? 1.0 / m4.x : 0.0;
    const double Ma[3] = {m4.y > 0 ? 1.0 / m4.y : 0.0, m4.z > 0 ? 1.0 / m4.z : 0.0, m4.w > 0 ? 1.0 / m4.w : 0.0};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double M = k < 3 ? Ml : Ma[k - 3];
        un[k] = (k < nd) ? u[k] + acc[k] / M : u[k];
    }
    if (i1 > i0) store6(x + (size_t)bi * kBodyStride, un);
    if (check) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            if (k >= nd) break;
            const double q = fmax(fabs(un[k]), fabs(u[k])) * tol_rel + tol_abs;
            const double e = (un[k] - u[k]) / q;
            e2 += e * e;
        }
    }
    return e2;
}

}  // namespace cdx

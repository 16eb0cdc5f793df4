// cdx_joints_device.hpp -- per-chain work of the joint rows (block LDL^T of the Schur complement
// K = G M^-1 G^T + E and its substitutions), shared by the grid-wide kernels (cdx_joints.hip) and
// the batched kernel (cdx_batch.hip).
#pragma once

#include "cdx_device.hpp"

namespace cdx {

constexpr int kMaxJRows = 12;  // rows per joint block (a full 6-dof joint with dampers)
constexpr int kJRowStride = 16;  // doubles per row: wa[6], wb[6], comp, vsrc, gerr, lam_prev

// inverse mass of dof k of a rigid body
__device__ inline double minv_k(const double* __restrict__ minv, int b, int k) {
    return minv[(size_t)b * kMinvStride + (k < 3 ? 0 : k - 2)];
}

// coefficient row i of block j on body `side` (0: a, 1: b)
__device__ inline const double* jrow(const double* __restrict__ jr, int row, int side) {
    return jr + (size_t)row * kJRowStride + 6 * side;
}

// Once per solve, for chain ch: the joint-row right-hand side of PjAssembler::rhs (:158-164;
// implicit gyroscopy and lambda-theta off) and the block LDL^T factors of K.  Returns
// CDX_NOT_FACTORIZED when a diagonal block is singular (S cannot be factorised, pj_assembler.cpp:175).
__device__ inline int joint_factor_chain(int ch, const int* __restrict__ jc_off, const int* __restrict__ jb_off,
                                         const int* __restrict__ jb_body, const int* __restrict__ jr_kind,
                                         const double* __restrict__ jr, const double* __restrict__ minv,
                                         const double* __restrict__ v, double dt, double theta, double beta,
                                         double* __restrict__ rlam, double* __restrict__ Lm, double* __restrict__ Dinv) {
    double K[kMaxJRows * kMaxJRows], Dp[kMaxJRows * kMaxJRows], A[kMaxJRows * 2 * kMaxJRows];
    int mp = 0;
    for (int j = jc_off[ch]; j < jc_off[ch + 1]; ++j) {
        const int r0 = jb_off[j], m = jb_off[j + 1] - r0;
        const int ba = jb_body[2 * j], bb = jb_body[2 * j + 1];
        // right-hand side of the rows
        for (int i = 0; i < m; ++i) {
            const int r = r0 + i;
            double gv = 0;
            for (int side = 0; side < 2; ++side) {
                const int b = side ? bb : ba;
                if (b < 0) continue;
                const double* w = jrow(jr, r, side);
                for (int k = 0; k < 6; ++k) gv += w[k] * v[(size_t)b * kBodyStride + k];
            }
            const double* rr = jr + (size_t)r * kJRowStride;
            const double comp = rr[12], vsrc = rr[13], gerr = rr[14], lprev = rr[15];
            double o;
            if (jr_kind[r] == 0) {
                o = -(1.0 / (theta * dt * dt)) * (comp * lprev) - ((1.0 - theta) / theta) * gv - (1.0 / theta) * vsrc;
                if (beta > 0) o -= (-(comp * lprev) / dt + gerr) * (beta / (dt * theta));
            } else {
                o = -((1.0 - theta) / theta) * gv - (1.0 / theta) * vsrc;
            }
            rlam[r] = o;
        }
        // K_jj = sum over the block's bodies of G M^-1 G^T, plus E
        for (int i = 0; i < m; ++i)
            for (int l = 0; l < m; ++l) {
                double s = 0;
                for (int side = 0; side < 2; ++side) {
                    const int b = side ? bb : ba;
                    if (b < 0) continue;
                    const double* wi = jrow(jr, r0 + i, side);
                    const double* wl = jrow(jr, r0 + l, side);
                    for (int k = 0; k < 6; ++k) s += wi[k] * minv_k(minv, b, k) * wl[k];
                }
                K[i * m + l] = s;
            }
        for (int i = 0; i < m; ++i) {
            const double* rr = jr + (size_t)(r0 + i) * kJRowStride;
            K[i * m + i] += rr[12] * (jr_kind[r0 + i] == 0 ? 1.0 / (theta * dt * dt) : 1.0 / (theta * dt));
        }
        if (j > jc_off[ch]) {
            // K_{j,j-1} through the body shared with the previous block, L_j = K_{j,j-1} D_{j-1}^-1,
            // D_j = K_jj - L_j K_{j,j-1}^T
            const int p0 = jb_off[j - 1];
            const int pa = jb_body[2 * (j - 1)], pb = jb_body[2 * (j - 1) + 1];
            const int s = (ba >= 0 && (ba == pa || ba == pb)) ? ba : bb;
            const int side_j = s == ba ? 0 : 1, side_p = s == pa ? 0 : 1;
            double Kjp[kMaxJRows * kMaxJRows];
            for (int i = 0; i < m; ++i)
                for (int l = 0; l < mp; ++l) {
                    const double* wi = jrow(jr, r0 + i, side_j);
                    const double* wl = jrow(jr, p0 + l, side_p);
                    double t = 0;
                    for (int k = 0; k < 6; ++k) t += wi[k] * minv_k(minv, s, k) * wl[k];
                    Kjp[i * mp + l] = t;
                }
            double* L = Lm + (size_t)j * kMaxJRows * kMaxJRows;
            for (int i = 0; i < m; ++i)
                for (int l = 0; l < mp; ++l) {
                    double t = 0;
                    for (int q = 0; q < mp; ++q) t += Kjp[i * mp + q] * Dp[q * mp + l];
                    L[i * mp + l] = t;
                }
            for (int i = 0; i < m; ++i)
                for (int l = 0; l < m; ++l) {
                    double t = 0;
                    for (int q = 0; q < mp; ++q) t += L[i * mp + q] * Kjp[l * mp + q];
                    K[i * m + l] -= t;
                }
        }
        // D_j^-1 by Gauss-Jordan with partial pivoting on [D_j | I]
        const int w2 = 2 * m;
        for (int i = 0; i < m; ++i)
            for (int l = 0; l < w2; ++l) A[i * w2 + l] = l < m ? K[i * m + l] : (l - m == i ? 1.0 : 0.0);
        for (int c = 0; c < m; ++c) {
            int pr = c;
            for (int i = c + 1; i < m; ++i)
                if (fabs(A[i * w2 + c]) > fabs(A[pr * w2 + c])) pr = i;
            if (pr != c)
                for (int l = 0; l < w2; ++l) {
                    const double t = A[c * w2 + l];
                    A[c * w2 + l] = A[pr * w2 + l];
                    A[pr * w2 + l] = t;
                }
            const double d = A[c * w2 + c];
            if (!(fabs(d) > 0.0)) return CDX_NOT_FACTORIZED;  // singular joint block
            for (int l = 0; l < w2; ++l) A[c * w2 + l] /= d;
            for (int i = 0; i < m; ++i) {
                if (i == c) continue;
                const double f = A[i * w2 + c];
                if (f != 0.0)
                    for (int l = 0; l < w2; ++l) A[i * w2 + l] -= f * A[c * w2 + l];
            }
        }
        double* Di = Dinv + (size_t)j * kMaxJRows * kMaxJRows;
        for (int i = 0; i < m; ++i)
            for (int l = 0; l < m; ++l) {
                Di[i * m + l] = A[i * w2 + m + l];
                Dp[i * m + l] = A[i * w2 + m + l];
            }
        mp = m;
    }
    return CDX_OK;
}

// x += S^-1 [r; -r_lambda] restricted to the jointed bodies of chain ch.
//   setup = 1: x0 and Lambda_0 are set (absolute; tr holds the full body right-hand side);
//   setup = 0: a sweep increment (tr holds W^T dp of each jointed body); x and Lambda accumulate.
// Returns the chain bodies' part of the residual sum (projected_jacobi.cpp:140-144) when `check`.
__device__ inline double joint_solve_chain(int ch, int setup, int check, const int* __restrict__ jc_off,
                                           const int* __restrict__ jcb_off, const int* __restrict__ jb_off,
                                           const int* __restrict__ jb_body, const double* __restrict__ jr,
                                           const double* __restrict__ rlam, const double* __restrict__ Lm,
                                           const double* __restrict__ Dinv, const int* __restrict__ jbody,
                                           const int* __restrict__ jbody_blk, const double* __restrict__ minv,
                                           const double* __restrict__ tr, double* __restrict__ z,
                                           double* __restrict__ lam, double* __restrict__ x, double tol_rel,
                                           double tol_abs) {
    double e2 = 0;
    {
        // forward: z_j = s_j - L_j z_{j-1}, then y_j = D_j^-1 z_j (stored in z)
        double zp[kMaxJRows], zc[kMaxJRows];
        int mp = 0;
        for (int j = jc_off[ch]; j < jc_off[ch + 1]; ++j) {
            const int r0 = jb_off[j], m = jb_off[j + 1] - r0;
            const int ba = jb_body[2 * j], bb = jb_body[2 * j + 1];
            for (int i = 0; i < m; ++i) {
                double s = 0;
                for (int side = 0; side < 2; ++side) {
                    const int b = side ? bb : ba;
                    if (b < 0) continue;
                    const double* w = jrow(jr, r0 + i, side);
                    const double* t = tr + (size_t)b * 6;
                    for (int k = 0; k < 6; ++k) s += w[k] * (minv_k(minv, b, k) * t[k]);
                }
                zc[i] = setup ? s - rlam[r0 + i] : s;
            }
            if (j > jc_off[ch]) {
                const double* L = Lm + (size_t)j * kMaxJRows * kMaxJRows;
                for (int i = 0; i < m; ++i) {
                    double t = 0;
                    for (int l = 0; l < mp; ++l) t += L[i * mp + l] * zp[l];
                    zc[i] -= t;
                }
            }
            for (int i = 0; i < m; ++i) zp[i] = zc[i];
            const double* Di = Dinv + (size_t)j * kMaxJRows * kMaxJRows;
            for (int i = 0; i < m; ++i) {
                double t = 0;
                for (int l = 0; l < m; ++l) t += Di[i * m + l] * zc[l];
                z[r0 + i] = t;
            }
            mp = m;
        }
        // backward: Lambda_j = y_j - L_{j+1}^T Lambda_{j+1} (kept in z)
        for (int j = jc_off[ch + 1] - 2; j >= jc_off[ch]; --j) {
            const int r0 = jb_off[j], m = jb_off[j + 1] - r0;
            const int n0 = jb_off[j + 1], mn = jb_off[j + 2] - n0;
            const double* L = Lm + (size_t)(j + 1) * kMaxJRows * kMaxJRows;  // mn x m
            for (int i = 0; i < m; ++i) {
                double t = 0;
                for (int l = 0; l < mn; ++l) t += L[l * m + i] * z[n0 + l];
                z[r0 + i] -= t;
            }
        }
        for (int j = jc_off[ch]; j < jc_off[ch + 1]; ++j)
            for (int r = jb_off[j]; r < jb_off[j + 1]; ++r) lam[r] = setup ? z[r] : lam[r] + z[r];
        // bodies: dv = M^-1 (r - sum over its blocks of G_b^T Lambda)
        for (int q = jcb_off[ch]; q < jcb_off[ch + 1]; ++q) {
            const int b = jbody[q];
            double acc[6];
            for (int k = 0; k < 6; ++k) acc[k] = tr[(size_t)b * 6 + k];
            for (int e = 0; e < 2; ++e) {
                const int j = jbody_blk[4 * (size_t)q + 2 * e], side = jbody_blk[4 * (size_t)q + 2 * e + 1];
                if (j < 0) continue;
                for (int r = jb_off[j]; r < jb_off[j + 1]; ++r) {
                    const double* w = jrow(jr, r, side);
                    for (int k = 0; k < 6; ++k) acc[k] -= w[k] * z[r];
                }
            }
            double* xb = x + (size_t)b * kBodyStride;
            for (int k = 0; k < 6; ++k) {
                const double dv = minv_k(minv, b, k) * acc[k];
                const double u = xb[k];
                const double un = setup ? dv : u + dv;
                xb[k] = un;
                if (check) {
                    const double qd = fmax(fabs(un), fabs(u)) * tol_rel + tol_abs;
                    const double ee = (un - u) / qd;
                    e2 += ee * ee;
                }
            }
        }
    }
    return e2;
}

}  // namespace cdx

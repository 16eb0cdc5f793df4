// cdx_adjoint.hip -- vector-Jacobian product of the FP64 projected-Jacobi solve.
//
// Forward (cdx_pj.hip, projected_jacobi.cpp:121-138,316-377): u_free = v + h M^-1 (f_ext + f_gyro),
// p_0 = warm start, x_0 = u_free + M^-1 W^T p_0, and per sweep
//     z_k = p_k - R (W x_k + c''),  p_{k+1} = Pi_mu(z_k),  x_{k+1} = x_k + M^-1 W^T (p_{k+1} - p_k)
// with R = alpha / diag(W M^-1 W^T) and c'' = c + e .* (W v + c) (so b = R c'').  The geometry
// (W, c, e) is held fixed.  Two backward modes share one final stage:
//
//  implicit  At the fixed point p* = Pi(p* - R(W x* + c'')), x* = u_free + M^-1 W^T p*,
//            the adjoint lambda solves (I - P (I - R A))^T lambda = g_p with g_p = W M^-1 L_x + L_p,
//            A = W M^-1 W^T and P = dPi/dz per contact.  It is iterated as
//            lambda <- g_p + xi - W M^-1 W^T (R xi),  xi = P^T lambda,
//            which is the transposed linearised sweep: one contact pass and one body pass per
//            iteration, the same memory pattern as the forward.  Then with s = R xi:
//              x_bar = L_x - W^T s,  c''_bar = -s,  R_bar = -xi .* (W x* + c'').
//  unrolled  Reverse mode through the K recorded sweeps (x_k, p_k).
//
// Final stage: u_free_bar = x_bar_0, M^-1_bar from x = ... + M^-1 W^T p and R(M^-1), v_bar from
// u_free (with the gyroscopic Jacobian) and the restitution term of c'', f_ext_bar = h M^-1 u_bar.
#include <algorithm>
#include <cmath>
#include <vector>

#include "cdx_pj_device.hpp"

namespace cdx {

// contact scratch layout, component-major [k * nc + c]
enum {
    A_P = 0,       // 9: dPi/dz, row-major (row = output n,t1,t2; column = input)
    A_Y = 9,       // 3: W x* + c''
    A_DMU = 12,    // 3: dp/dmu
    A_GP = 15,     // 3: g_p
    A_LAM = 18,    // 3: lambda
    A_XI = 21,     // 3: xi = P^T lambda
    A_RBAR = 24,   // 3: accumulated R_bar (unrolled)
    A_CBAR = 27,   // 3: accumulated c''_bar (unrolled)
    A_MUBAR = 30,  // 1: accumulated mu_bar (unrolled)
    A_COUNT = 31
};

// Generalised Jacobian of project() (projected_jacobi.cpp:100-119) at z, following the same
// branches: normal clamp, then the disk scaling s = min(1, mu p_n / |t|) when |t| > 0.
__device__ inline void projection_jacobian(const double z[3], int nr, double mu, double P[9], double dmu[3]) {
    for (int k = 0; k < 9; ++k) P[k] = 0.0;
    dmu[0] = dmu[1] = dmu[2] = 0.0;
    const double pn = fmax(z[0], 0.0);
    const double dn = z[0] > 0.0 ? 1.0 : 0.0;
    P[0] = dn;
    if (nr == 1) return;
    const double t1 = z[1], t2 = z[2];
    const double tn2 = t1 * t1 + t2 * t2;
    if (!(tn2 > 0)) {  // no scaling applied: p_t = t
        P[4] = 1.0;
        P[8] = 1.0;
        return;
    }
    const double tn = sqrt(tn2);
    const double sr = (mu * pn) / tn;
    if (sr >= 1.0) {  // inside the disk (stick)
        P[4] = 1.0;
        P[8] = 1.0;
        return;
    }
    // slip: p_t = mu p_n t / |t|
    const double u1 = t1 / tn, u2 = t2 / tn;
    P[4] = sr * (1.0 - u1 * u1);
    P[5] = sr * (-u1 * u2);
    P[7] = sr * (-u2 * u1);
    P[8] = sr * (1.0 - u2 * u2);
    P[3] = mu * u1 * dn;
    P[6] = mu * u2 * dn;
    dmu[1] = pn * u1;
    dmu[2] = pn * u2;
}

__device__ inline void mat_t_vec(const double P[9], const double v[3], double o[3]) {
    for (int c = 0; c < 3; ++c) o[c] = P[c] * v[0] + P[3 + c] * v[1] + P[6 + c] * v[2];
}

__device__ inline void minv_body(const double* __restrict__ minv, int b, double m[6]) {
    if (b < 0) {
        for (int k = 0; k < 6; ++k) m[k] = 0;
        return;
    }
    const double* mi = minv + (size_t)b * kMinvStride;
    m[0] = m[1] = m[2] = mi[0];
    m[3] = mi[1];
    m[4] = mi[2];
    m[5] = mi[3];
}

// unscaled bias c'' = c + e .* (W v + c) of row k
__device__ inline double cbias(const ContactJ& j, int k, int nc, int c, const double* __restrict__ rest,
                               const double* __restrict__ vb, const double va[6], const double vbb[6]) {
    const double ck = vb[(size_t)k * nc + c];
    return ck + rest[(size_t)k * nc + c] * (row_dot(j.d[k], j.aA[k], j.aB[k], va, vbb) + ck);
}

// Implicit setup: P, y, dmu at the fixed point, g_p, lambda_0 = g_p, xi_0, and the side impulses
// of W^T (R xi_0).  g_p = W S^-1_vv L_x (+ L_p): S^-1_vv = M^-1 applied here, or (lx_applied, with
// joint rows) already applied by the caller.
__global__ void k_adj_setup(int nc, const int* __restrict__ ida, const int* __restrict__ idb, const double* __restrict__ J,
                            const double* __restrict__ R, const double* __restrict__ bias, const double* __restrict__ mu,
                            const double* __restrict__ rest, const double* __restrict__ vb, const double* __restrict__ minv,
                            const double* __restrict__ v, const double* __restrict__ x, const double* __restrict__ p,
                            const double* __restrict__ lx, const double* __restrict__ lp, double* __restrict__ S,
                            double* __restrict__ imp, int lx_applied) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int a = ida[c], b = idb[c];
    if (a < 0 && b < 0) {
        for (int k = 0; k < A_COUNT; ++k) S[(size_t)k * nc + c] = 0;
        return;
    }
    const double muc = mu[c];
    const int nr = muc > 0 ? 3 : 1;
    ContactJ j;
    load_J(J, nc, c, nr, j);
    double xa[6], xb[6], va[6], vbb[6], la[6], lb[6], ma[6], mb[6];
    load_body(x, a, xa);
    load_body(x, b, xb);
    load_body(v, a, va);
    load_body(v, b, vbb);
    load_body(lx, a, la);
    load_body(lx, b, lb);
    minv_body(minv, a, ma);
    minv_body(minv, b, mb);
    if (!lx_applied)  // lx holds L_x; with joint rows it already holds S^-1_vv L_x
        for (int k = 0; k < 6; ++k) {
            la[k] *= ma[k];
            lb[k] *= mb[k];
        }
    double z[3] = {0, 0, 0}, y[3] = {0, 0, 0}, gp[3] = {0, 0, 0};
    for (int k = 0; k < nr; ++k) {
        const double g = row_dot(j.d[k], j.aA[k], j.aB[k], xa, xb);
        const double pk = p[(size_t)k * nc + c];
        z[k] = pk - (R[(size_t)k * nc + c] * g + bias[(size_t)k * nc + c]);
        y[k] = g + cbias(j, k, nc, c, rest, vb, va, vbb);
        gp[k] = row_dot(j.d[k], j.aA[k], j.aB[k], la, lb) + (lp ? lp[3 * (size_t)c + k] : 0.0);
    }
    double P[9], dmu[3], xi[3];
    projection_jacobian(z, nr, muc, P, dmu);
    mat_t_vec(P, gp, xi);
    for (int k = 0; k < 9; ++k) S[(size_t)(A_P + k) * nc + c] = P[k];
    double s[3];
    for (int k = 0; k < 3; ++k) {
        S[(size_t)(A_Y + k) * nc + c] = y[k];
        S[(size_t)(A_DMU + k) * nc + c] = dmu[k];
        S[(size_t)(A_GP + k) * nc + c] = gp[k];
        S[(size_t)(A_LAM + k) * nc + c] = gp[k];
        S[(size_t)(A_XI + k) * nc + c] = xi[k];
        S[(size_t)(A_RBAR + k) * nc + c] = 0.0;
        S[(size_t)(A_CBAR + k) * nc + c] = 0.0;
        s[k] = k < nr ? R[(size_t)k * nc + c] * xi[k] : 0.0;
    }
    S[(size_t)A_MUBAR * nc + c] = 0.0;
    write_side_impulses(imp, c, a, b, nr, j, s);
}

// t = M^-1 (sum of side impulses); a jointed body (jslot >= 0) gets t = 0 and hands the sum to
// the chain solve through tr, which adds its share of S^-1.
__global__ void k_adj_body_t(int nb, const int* __restrict__ dof, const double* __restrict__ minv,
                             const int* __restrict__ incptr, const int* __restrict__ incidx, const double* __restrict__ imp,
                             double* __restrict__ t, const int* __restrict__ ctrl, const int* __restrict__ jslot,
                             double* __restrict__ tr) {
    if (ctrl && ctrl[0]) return;
    const int bi = blockIdx.x * blockDim.x + threadIdx.x;
    if (bi >= nb) return;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int i = incptr[bi]; i < incptr[bi + 1]; ++i) {
        const double2* q = reinterpret_cast<const double2*>(imp + (size_t)incidx[i] * 6);
        const double2 q0 = q[0], q1 = q[1], q2 = q[2];
        acc[0] += q0.x;
        acc[1] += q0.y;
        acc[2] += q1.x;
        acc[3] += q1.y;
        acc[4] += q2.x;
        acc[5] += q2.y;
    }
    double m[6];
    minv_body(minv, bi, m);
    double o[6];
    if (jslot && jslot[bi] >= 0) {
        store6(tr + (size_t)bi * 6, acc);
        for (int k = 0; k < 6; ++k) o[k] = 0.0;
    } else {
        for (int k = 0; k < 6; ++k) o[k] = k < dof[bi] ? m[k] * acc[k] : 0.0;
    }
    store6(t + (size_t)bi * kBodyStride, o);
}

// out = M^-1 in for the bodies outside joint chains; a jointed body gets out = 0 and tr = in
// (the chain solve then adds S^-1 [tr; 0]).
__global__ void k_adj_minv_apply(int nb, const int* __restrict__ dof, const double* __restrict__ minv,
                                 const double* __restrict__ in, double* __restrict__ out, const int* __restrict__ jslot,
                                 double* __restrict__ tr) {
    const int bi = blockIdx.x * blockDim.x + threadIdx.x;
    if (bi >= nb) return;
    double u[6], m[6], o[6];
    load_body(in, bi, u);
    minv_body(minv, bi, m);
    if (jslot[bi] >= 0) {
        store6(tr + (size_t)bi * 6, u);
        for (int k = 0; k < 6; ++k) o[k] = 0.0;
    } else {
        for (int k = 0; k < 6; ++k) o[k] = k < dof[bi] ? m[k] * u[k] : 0.0;
    }
    store6(out + (size_t)bi * kBodyStride, o);
}

// lambda <- g_p + xi - W t; xi <- P^T lambda; side impulses of W^T (R xi); convergence partials.
__global__ void __launch_bounds__(kBlock) k_adj_iter(int nc, const int* __restrict__ ida, const int* __restrict__ idb,
                                                     const double* __restrict__ J, const double* __restrict__ R,
                                                     const double* __restrict__ mu, const double* __restrict__ t,
                                                     double* __restrict__ S, double* __restrict__ imp, int check,
                                                     double* __restrict__ partial, const int* __restrict__ ctrl) {
    if (ctrl[0]) return;
    __shared__ double red[kBlock / 64];
    const int c = blockIdx.x * kBlock + threadIdx.x;
    double d2 = 0, n2 = 0;
    if (c < nc) {
        const int a = ida[c], b = idb[c];
        if (a >= 0 || b >= 0) {
            const int nr = mu[c] > 0 ? 3 : 1;
            ContactJ j;
            load_J(J, nc, c, nr, j);
            double ta[6], tb[6];
            load_body(t, a, ta);
            load_body(t, b, tb);
            double lam[3] = {0, 0, 0}, P[9], xi[3], s[3] = {0, 0, 0};
            for (int k = 0; k < nr; ++k) {
                const double old = S[(size_t)(A_LAM + k) * nc + c];
                lam[k] = S[(size_t)(A_GP + k) * nc + c] + S[(size_t)(A_XI + k) * nc + c] -
                         row_dot(j.d[k], j.aA[k], j.aB[k], ta, tb);
                d2 += (lam[k] - old) * (lam[k] - old);
                n2 += lam[k] * lam[k];
            }
            for (int k = 0; k < 9; ++k) P[k] = S[(size_t)(A_P + k) * nc + c];
            mat_t_vec(P, lam, xi);
            for (int k = 0; k < nr; ++k) {
                S[(size_t)(A_LAM + k) * nc + c] = lam[k];
                S[(size_t)(A_XI + k) * nc + c] = xi[k];
                s[k] = R[(size_t)k * nc + c] * xi[k];
            }
            write_side_impulses(imp, c, a, b, nr, j, s);
        }
    }
    if (check) {
        const double sd = block_sum(d2, red);
        const double sn = block_sum(n2, red);
        if (threadIdx.x == 0) {
            partial[2 * blockIdx.x] = sd;
            partial[2 * blockIdx.x + 1] = sn;
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_adj_finalize(int nparts, const double* __restrict__ partial, int it, double tol,
                                                         int* __restrict__ ctrl, double* __restrict__ res,
                                                         int* __restrict__ hflag) {
    if (ctrl[0]) return;
    __shared__ double red[kBlock / 64];
    double d = 0, n = 0;
    for (int i = threadIdx.x; i < nparts; i += kBlock) {
        d += partial[2 * i];
        n += partial[2 * i + 1];
    }
    d = block_sum(d, red);
    n = block_sum(n, red);
    if (threadIdx.x == 0) {
        const double r = n > 0 ? sqrt(d / n) : sqrt(d);
        res[0] = r;
        if (r <= tol) {
            ctrl[1] = it + 1;
            ctrl[0] = 1;
            hflag[0] = 1;
        }
    }
}

// ---- unrolled reverse sweep k: contact pass
// Inputs: recorded x_k (xk), p_k (pk), p_{k+1} (pk1), x_bar_{k+1} (xbar, body), p_bar_{k+1} (pbar,
// contact, 3 per contact in S[A_LAM]).  Outputs: p_bar_k in S[A_LAM], side impulses of R z_bar (imp)
// and of dp_k = p_{k+1} - p_k (imp2), accumulated R_bar, c''_bar, mu_bar.
__global__ void __launch_bounds__(kBlock) k_unr_contacts(int nc, const int* __restrict__ ida, const int* __restrict__ idb,
                                                         const double* __restrict__ J, const double* __restrict__ R,
                                                         const double* __restrict__ bias, const double* __restrict__ mu,
                                                         const double* __restrict__ rest, const double* __restrict__ vb,
                                                         const double* __restrict__ minv, const double* __restrict__ v,
                                                         const double* __restrict__ xk, const double* __restrict__ pk,
                                                         const double* __restrict__ pk1, const double* __restrict__ xbar,
                                                         double* __restrict__ S, double* __restrict__ imp,
                                                         double* __restrict__ imp2) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= nc) return;
    const int a = ida[c], b = idb[c];
    if (a < 0 && b < 0) return;
    const double muc = mu[c];
    const int nr = muc > 0 ? 3 : 1;
    ContactJ j;
    load_J(J, nc, c, nr, j);
    double ua[6], ub[6], ga[6], gb[6], ma[6], mb[6], va[6], vbb[6];
    load_body(xk, a, ua);
    load_body(xk, b, ub);
    load_body(xbar, a, ga);
    load_body(xbar, b, gb);
    load_body(v, a, va);
    load_body(v, b, vbb);
    minv_body(minv, a, ma);
    minv_body(minv, b, mb);
    for (int k = 0; k < 6; ++k) {
        ga[k] *= ma[k];
        gb[k] *= mb[k];
    }
    double z[3] = {0, 0, 0}, y[3] = {0, 0, 0}, ptot[3] = {0, 0, 0}, w[3] = {0, 0, 0}, dp[3] = {0, 0, 0};
    for (int k = 0; k < nr; ++k) {
        const double g = row_dot(j.d[k], j.aA[k], j.aB[k], ua, ub);
        const double p0 = pk[(size_t)k * nc + c];
        z[k] = p0 - (R[(size_t)k * nc + c] * g + bias[(size_t)k * nc + c]);
        y[k] = g + cbias(j, k, nc, c, rest, vb, va, vbb);
        w[k] = row_dot(j.d[k], j.aA[k], j.aB[k], ga, gb);  // W M^-1 x_bar_{k+1}
        ptot[k] = S[(size_t)(A_LAM + k) * nc + c] + w[k];
        dp[k] = pk1[(size_t)k * nc + c] - p0;
    }
    double P[9], dmu[3], zb[3], s[3] = {0, 0, 0};
    projection_jacobian(z, nr, muc, P, dmu);
    mat_t_vec(P, ptot, zb);
    double mub = 0;
    for (int k = 0; k < nr; ++k) mub += dmu[k] * ptot[k];
    S[(size_t)A_MUBAR * nc + c] += mub;
    for (int k = 0; k < nr; ++k) {
        const double Rk = R[(size_t)k * nc + c];
        S[(size_t)(A_LAM + k) * nc + c] = zb[k] - w[k];  // p_bar_k
        S[(size_t)(A_RBAR + k) * nc + c] += -zb[k] * y[k];
        S[(size_t)(A_CBAR + k) * nc + c] += -Rk * zb[k];
        s[k] = Rk * zb[k];
    }
    write_side_impulses(imp, c, a, b, nr, j, s);
    write_side_impulses(imp2, c, a, b, nr, j, dp);
}

// unrolled reverse sweep k: body pass.  M^-1_bar += x_bar_{k+1} .* W^T dp_k ; x_bar_k = x_bar_{k+1} - W^T (R z_bar)
__global__ void k_unr_bodies(int nb, const int* __restrict__ dof, const int* __restrict__ incptr,
                             const int* __restrict__ incidx, const double* __restrict__ imp, const double* __restrict__ imp2,
                             double* __restrict__ xbar, double* __restrict__ gminv) {
    const int bi = blockIdx.x * blockDim.x + threadIdx.x;
    if (bi >= nb) return;
    double a1[6] = {0, 0, 0, 0, 0, 0}, a2[6] = {0, 0, 0, 0, 0, 0};
    for (int i = incptr[bi]; i < incptr[bi + 1]; ++i) {
        const double* q = imp + (size_t)incidx[i] * 6;
        const double* r = imp2 + (size_t)incidx[i] * 6;
        for (int k = 0; k < 6; ++k) {
            a1[k] += q[k];
            a2[k] += r[k];
        }
    }
    double g[6], gm[6];
    load_body(xbar, bi, g);
    load_body(gminv, bi, gm);
    for (int k = 0; k < dof[bi]; ++k) {
        gm[k] += g[k] * a2[k];
        g[k] -= a1[k];
    }
    store6(xbar + (size_t)bi * kBodyStride, g);
    store6(gminv + (size_t)bi * kBodyStride, gm);
}

// Final contact pass (both modes): side vectors of (a) W^T s, (b) W^T pv, (c) sum_i rho_i W_ij^2,
// (d) W^T (e .* c''_bar); and mu_bar per contact.
__global__ void k_adj_final_contacts(int nc, int implicit, double alpha, const int* __restrict__ ida,
                                     const int* __restrict__ idb, const double* __restrict__ J, const double* __restrict__ R,
                                     const double* __restrict__ mu, const double* __restrict__ rest,
                                     const double* __restrict__ pv, const double* __restrict__ S, double* __restrict__ ia,
                                     double* __restrict__ ib, double* __restrict__ ic, double* __restrict__ id,
                                     double* __restrict__ gmu) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int a = ida[c], b = idb[c];
    if (a < 0 && b < 0) {
        if (gmu) gmu[c] = 0.0;
        return;
    }
    const int nr = mu[c] > 0 ? 3 : 1;
    ContactJ j;
    load_J(J, nc, c, nr, j);
    double s[3] = {0, 0, 0}, pp[3] = {0, 0, 0}, rho[3] = {0, 0, 0}, ec[3] = {0, 0, 0};
    double mub = 0;
    for (int k = 0; k < nr; ++k) {
        const double Rk = R[(size_t)k * nc + c];
        pp[k] = pv[(size_t)k * nc + c];
        double cbar;
        if (implicit) {
            const double xi = S[(size_t)(A_XI + k) * nc + c];
            s[k] = Rk * xi;
            rho[k] = xi * S[(size_t)(A_Y + k) * nc + c] * Rk * Rk / alpha;
            cbar = -s[k];
            mub += S[(size_t)(A_DMU + k) * nc + c] * S[(size_t)(A_LAM + k) * nc + c];
        } else {
            rho[k] = -S[(size_t)(A_RBAR + k) * nc + c] * Rk * Rk / alpha;
            cbar = S[(size_t)(A_CBAR + k) * nc + c];
        }
        ec[k] = rest[(size_t)k * nc + c] * cbar;
    }
    if (!implicit) mub = S[(size_t)A_MUBAR * nc + c];
    if (gmu) gmu[c] = mub;
    write_side_impulses(ia, c, a, b, nr, j, s);
    write_side_impulses(ib, c, a, b, nr, j, pp);
    write_side_impulses(id, c, a, b, nr, j, ec);
    // (c): per side, sum_k rho_k W_kj^2 (signs square away)
    double la[6] = {0, 0, 0, 0, 0, 0}, lbv[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < nr; ++k)
        for (int m = 0; m < 3; ++m) {
            const double d2 = j.d[k][m] * j.d[k][m] * rho[k];
            la[m] += d2;
            lbv[m] += d2;
            la[3 + m] += j.aA[k][m] * j.aA[k][m] * rho[k];
            lbv[3 + m] += j.aB[k][m] * j.aB[k][m] * rho[k];
        }
    if (a >= 0) store6(ic + (size_t)(2 * c) * 6, la);
    if (b >= 0) store6(ic + (size_t)(2 * c + 1) * 6, lbv);
}

// Final body pass: x_bar = base - (a); M^-1_bar = acc + x_bar .* (b) + (c); u_free backward.
__global__ void k_adj_final_bodies(int nb, const int* __restrict__ dof, const double* __restrict__ minv,
                                   const double* __restrict__ v, const double* __restrict__ fext,
                                   const double* __restrict__ fgyro, const int* __restrict__ incptr,
                                   const int* __restrict__ incidx, const double* __restrict__ ia,
                                   const double* __restrict__ ib, const double* __restrict__ ic,
                                   const double* __restrict__ id, const double* __restrict__ base,
                                   const double* __restrict__ gacc, double h, double* __restrict__ gv,
                                   double* __restrict__ gm, double* __restrict__ gf, const int* __restrict__ jslot,
                                   double* __restrict__ tr) {
    const int bi = blockIdx.x * blockDim.x + threadIdx.x;
    if (bi >= nb) return;
    double sa[6] = {0, 0, 0, 0, 0, 0}, sb[6] = {0, 0, 0, 0, 0, 0}, sc[6] = {0, 0, 0, 0, 0, 0}, sd[6] = {0, 0, 0, 0, 0, 0};
    for (int i = incptr[bi]; i < incptr[bi + 1]; ++i) {
        const size_t o = (size_t)incidx[i] * 6;
        for (int k = 0; k < 6; ++k) {
            sa[k] += ia[o + k];
            sb[k] += ib[o + k];
            sc[k] += ic[o + k];
            sd[k] += id[o + k];
        }
    }
    const int nd = dof[bi];
    double g[6], acc[6], m[6], vv[6], fe[6], fg[6];
    load_body(base, bi, g);
    load_body(gacc, bi, acc);
    load_body(v, bi, vv);
    load_body(fext, bi, fe);
    load_body(fgyro, bi, fg);
    minv_body(minv, bi, m);
    double xb[6], gmo[6] = {0, 0, 0, 0, 0, 0}, gvo[6] = {0, 0, 0, 0, 0, 0}, gfo[6] = {0, 0, 0, 0, 0, 0};
    if (jslot && jslot[bi] >= 0) {
        // jointed body: r_bar = S^-1 [x_bar; 0] comes from the chain solve; here x_bar goes to tr,
        // dL/dv_n starts from the restitution term, and dL/dM^-1 is not offered (NaN)
        for (int k = 0; k < 6; ++k) {
            xb[k] = g[k] - sa[k];
            gvo[k] = sd[k];
            gmo[k] = NAN;
        }
        store6(tr + (size_t)bi * 6, xb);
        store6(gv + (size_t)bi * kBodyStride, gvo);
        store6(gm + (size_t)bi * kBodyStride, gmo);
        store6(gf + (size_t)bi * kBodyStride, gfo);
        return;
    }
    for (int k = 0; k < nd; ++k) {
        xb[k] = g[k] - sa[k];  // u_free_bar
        gmo[k] = acc[k] + xb[k] * sb[k] + sc[k] + h * xb[k] * (fe[k] + fg[k]);
        gfo[k] = h * m[k] * xb[k];
        gvo[k] = xb[k] + sd[k];
    }
    if (nd == 6) {
        // u_free_ang = w + h I^-1 (tau + f_gyro(w)), f_gyro = -w x (I w):
        // v_bar_w += -h ( -I (w x g') + (I w) x g' ),  g' = I^-1 u_bar;  I^-1_bar += h I_j^2 w_j [g' x w]_j
        const double w[3] = {vv[3], vv[4], vv[5]};
        double I[3], Iw[3], gp[3];
        for (int k = 0; k < 3; ++k) {
            I[k] = m[3 + k] > 0 ? 1.0 / m[3 + k] : 0.0;
            Iw[k] = I[k] * w[k];
            gp[k] = m[3 + k] * xb[3 + k];
        }
        const double wxg[3] = {w[1] * gp[2] - w[2] * gp[1], w[2] * gp[0] - w[0] * gp[2], w[0] * gp[1] - w[1] * gp[0]};
        const double Iwxg[3] = {Iw[1] * gp[2] - Iw[2] * gp[1], Iw[2] * gp[0] - Iw[0] * gp[2], Iw[0] * gp[1] - Iw[1] * gp[0]};
        const double gxw[3] = {-wxg[0], -wxg[1], -wxg[2]};
        for (int k = 0; k < 3; ++k) {
            gvo[3 + k] += -h * (-I[k] * wxg[k] + Iwxg[k]);
            gmo[3 + k] += h * I[k] * I[k] * w[k] * gxw[k];
        }
    }
    store6(gv + (size_t)bi * kBodyStride, gvo);
    store6(gm + (size_t)bi * kBodyStride, gmo);
    store6(gf + (size_t)bi * kBodyStride, gfo);
}

// Convergence of the body-space adjoint t = S^-1_vv W^T R xi between two iterations: per-block sums
// of |t - t_prev|^2 and |t|^2.  Where contacts are statically indeterminate (A = W S^-1 W^T
// singular on the stick rows) lambda may drift along directions with W^T R v = 0, which leave t,
// and with it every velocity-space gradient, unchanged.
__global__ void __launch_bounds__(kBlock) k_adj_tcheck(int nb, const double* __restrict__ t, const double* __restrict__ tp,
                                                       double* __restrict__ partial, const int* __restrict__ ctrl) {
    if (ctrl[0]) return;
    __shared__ double red[kBlock / 64];
    const int bi = blockIdx.x * kBlock + threadIdx.x;
    double d2 = 0, n2 = 0;
    if (bi < nb) {
        double a[6], b[6];
        load_body(t, bi, a);
        load_body(tp, bi, b);
        for (int k = 0; k < 6; ++k) {
            d2 += (a[k] - b[k]) * (a[k] - b[k]);
            n2 += a[k] * a[k];
        }
    }
    const double sd = block_sum(d2, red);
    const double sn = block_sum(n2, red);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = sd;
        partial[2 * blockIdx.x + 1] = sn;
    }
}

// ---- GMRES on the implicit adjoint system B lambda = g_p, B = I - (I - A R) P^T, A = W S^-1_vv W^T
// (used with joint rows, where the fixed-point iteration of the contact adjoint can be too slow).
// Vectors live on the contact rows, component-major [k * nc + c] (rows a contact does not have
// stay zero).  One application of B: xi = P^T v and the side impulses of W^T (R xi) (k_gm_pt),
// t = S^-1_vv W^T R xi (body pass + chain solve), B v = v - xi + W t (k_gm_out).
constexpr int kGmRestart = 30;

__global__ void k_gm_pt(int nc, const int* __restrict__ ida, const int* __restrict__ idb, const double* __restrict__ J,
                        const double* __restrict__ R, const double* __restrict__ mu, const double* __restrict__ S,
                        const double* __restrict__ v, double* __restrict__ xi_out, double* __restrict__ imp) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int a = ida[c], b = idb[c];
    if (a < 0 && b < 0) {
        for (int k = 0; k < 3; ++k) xi_out[(size_t)k * nc + c] = 0.0;
        return;
    }
    const int nr = mu[c] > 0 ? 3 : 1;
    ContactJ j;
    load_J(J, nc, c, nr, j);
    double P[9], vv[3] = {0, 0, 0}, xi[3], s[3] = {0, 0, 0};
    for (int k = 0; k < 9; ++k) P[k] = S[(size_t)(A_P + k) * nc + c];
    for (int k = 0; k < nr; ++k) vv[k] = v[(size_t)k * nc + c];
    mat_t_vec(P, vv, xi);
    for (int k = 0; k < 3; ++k) {
        xi_out[(size_t)k * nc + c] = k < nr ? xi[k] : 0.0;
        s[k] = k < nr ? R[(size_t)k * nc + c] * xi[k] : 0.0;
    }
    write_side_impulses(imp, c, a, b, nr, j, s);
}

__global__ void k_gm_out(int nc, const int* __restrict__ ida, const int* __restrict__ idb, const double* __restrict__ J,
                         const double* __restrict__ mu, const double* __restrict__ t, const double* __restrict__ v,
                         const double* __restrict__ xi, double* __restrict__ w) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int a = ida[c], b = idb[c];
    const int nr = (a < 0 && b < 0) ? 0 : (mu[c] > 0 ? 3 : 1);
    ContactJ j;
    if (nr) load_J(J, nc, c, nr, j);
    double ta[6], tb[6];
    load_body(t, a, ta);
    load_body(t, b, tb);
    for (int k = 0; k < 3; ++k) {
        const size_t r = (size_t)k * nc + c;
        w[r] = k < nr ? v[r] - xi[r] + row_dot(j.d[k], j.aA[k], j.aB[k], ta, tb) : 0.0;
    }
}

// partial[i * gridDim.x + block] = this block's part of <w, V_i>, i < nvec
__global__ void __launch_bounds__(kBlock) k_gm_dots(int n, int nvec, const double* __restrict__ V, size_t ld,
                                                    const double* __restrict__ w, double* __restrict__ partial) {
    // all nvec products reduced within each wave by shuffles, then one barrier for the block's waves
    __shared__ double red[kBlock / 64][kGmRestart + 1];
    const int r = blockIdx.x * kBlock + threadIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const double wr = r < n ? w[r] : 0.0;
    for (int i = 0; i < nvec; ++i) {
        double d = r < n ? wr * V[i * ld + r] : 0.0;
        for (int o = 32; o > 0; o >>= 1) d += __shfl_xor(d, o, 64);
        if (lane == 0) red[wv][i] = d;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nvec; i += kBlock) {
        double s = 0;
        for (int q = 0; q < kBlock / 64; ++q) s += red[q][i];
        partial[(size_t)i * gridDim.x + blockIdx.x] = s;
    }
}

// h[i] = (h[i] +) sum of the partials of dot i; one workgroup, a wave per dot (no barriers)
__global__ void __launch_bounds__(kBlock) k_gm_reduce(int nvec, int nblk, const double* __restrict__ partial,
                                                      double* __restrict__ h, int accumulate) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int i = wv; i < nvec; i += kBlock / 64) {
        double s = 0;
        for (int b = lane; b < nblk; b += 64) s += partial[(size_t)i * nblk + b];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (lane == 0) h[i] = accumulate ? h[i] + s : s;
    }
}

// w += sign * sum_i h_i V_i
__global__ void k_gm_axpy(int n, int nvec, const double* __restrict__ V, size_t ld, const double* __restrict__ h,
                          double sign, double* __restrict__ w) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double acc = 0;
    for (int i = 0; i < nvec; ++i) acc += h[i] * V[i * ld + r];
    w[r] += sign * acc;
}

// out = w / sqrt(*nrm2) (zero after a breakdown); *beta = sqrt(*nrm2)
__global__ void k_gm_normalize(int n, const double* __restrict__ w, const double* __restrict__ nrm2, double* __restrict__ beta,
                               double* __restrict__ out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const double b = sqrt(fmax(*nrm2, 0.0));
    if (r == 0) *beta = b;
    if (r < n) out[r] = b > 1e-300 ? w[r] / b : 0.0;
}

// out[i] += a[i], i < n (one workgroup)
__global__ void k_gm_add(int n, const double* __restrict__ a, double* __restrict__ out) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) out[i] += a[i];
}

__global__ void k_gm_sub(int n, const double* __restrict__ g, const double* __restrict__ w, double* __restrict__ out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) out[r] = g[r] - w[r];
}

// ---- scene-segmented GMRES: the scenes of a batch are independent (B is block diagonal), and each
// gets its own Krylov space, coefficients and convergence test, so every scene's adjoint equals the
// one GMRES finds for that scene alone.  Scene s owns the contacts [coff[s], coff[s+1]); its
// vector entries are rows k * nc + c of those contacts.

// out[s * ostride + i] = (out +) <w, V_i> over scene s's entries, i < nvec; one workgroup per scene
__global__ void __launch_bounds__(kBlock) k_gms_dots(int ns, const int* __restrict__ coff, int nc, int nvec,
                                                     const double* __restrict__ V, size_t ld, const double* __restrict__ w,
                                                     double* __restrict__ out, size_t ostride) {
    __shared__ double red[kBlock / 64];
    const int s = blockIdx.x;
    if (s >= ns) return;
    const int c0 = coff[s], c1 = coff[s + 1];
    for (int i = 0; i < nvec; ++i) {
        double d = 0;
        for (int c = c0 + (int)threadIdx.x; c < c1; c += kBlock)
            for (int k = 0; k < 3; ++k) {
                const size_t r = (size_t)k * nc + c;
                d += w[r] * V[i * ld + r];
            }
        d = block_sum(d, red);
        if (threadIdx.x == 0) out[(size_t)s * ostride + i] = d;
    }
}

// out = w / sqrt(nrm2[scene]) per scene (zero after a breakdown); beta[s * bstride] = sqrt(nrm2[s])
__global__ void k_gms_normalize(int nc, const int* __restrict__ cscene, const double* __restrict__ w,
                                const double* __restrict__ nrm2, size_t nstride, double* __restrict__ beta, size_t bstride,
                                double* __restrict__ out, int ns) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < ns) beta[(size_t)r * bstride] = sqrt(fmax(nrm2[(size_t)r * nstride], 0.0));
    if (r >= 3 * nc) return;
    const int c = r % nc;
    const double b = sqrt(fmax(nrm2[(size_t)cscene[c] * nstride], 0.0));
    out[r] = b > 1e-300 ? w[r] / b : 0.0;
}

// w += sign * sum_i h[scene * hstride + i] V_i
__global__ void k_gms_axpy(int nc, const int* __restrict__ cscene, int nvec, const double* __restrict__ V, size_t ld,
                           const double* __restrict__ h, size_t hstride, double sign, double* __restrict__ w) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= 3 * nc) return;
    const double* hs = h + (size_t)cscene[r % nc] * hstride;
    double acc = 0;
    for (int i = 0; i < nvec; ++i) acc += hs[i] * V[i * ld + r];
    w[r] += sign * acc;
}

// out[s * stride + i] += a[s * stride + i], i < n
__global__ void k_gms_add(int ns, int n, const double* __restrict__ a, size_t astride, double* __restrict__ out,
                          size_t ostride) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ns * n) return;
    const int s = t / n, i = t % n;
    out[(size_t)s * ostride + i] += a[(size_t)s * astride + i];
}

// Jointed bodies after the chain solve r = S^-1 [x_bar; 0] (velocity part rv): dL/dv_n += M r_v
// (r_v = M v + h f_ext + h f_gyro(v) + ...), the gyroscopic Jacobian, and the (1 - theta) G v term
// of the joint rows' right-hand side; dL/df_ext = h r_v.
__global__ void k_adj_final_jointed(int njb, const int* __restrict__ jbody, const int* __restrict__ jbody_blk,
                                    const int* __restrict__ jb_off, const double* __restrict__ jr,
                                    const double* __restrict__ rlb, const double* __restrict__ minv,
                                    const double* __restrict__ v, const double* __restrict__ rv, double h, double theta,
                                    double* __restrict__ gv, double* __restrict__ gf) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= njb) return;
    const int bi = jbody[q];
    double r[6], m[6], vv[6], gvo[6], gfo[6];
    load_body(rv, bi, r);
    minv_body(minv, bi, m);
    load_body(v, bi, vv);
    load_body(gv, bi, gvo);
    for (int k = 0; k < 6; ++k) {
        gvo[k] += (m[k] > 0 ? 1.0 / m[k] : 0.0) * r[k];
        gfo[k] = h * r[k];
    }
    const double w[3] = {vv[3], vv[4], vv[5]};
    double I[3], Iw[3], gp[3];
    for (int k = 0; k < 3; ++k) {
        I[k] = m[3 + k] > 0 ? 1.0 / m[3 + k] : 0.0;
        Iw[k] = I[k] * w[k];
        gp[k] = r[3 + k];
    }
    const double wxg[3] = {w[1] * gp[2] - w[2] * gp[1], w[2] * gp[0] - w[0] * gp[2], w[0] * gp[1] - w[1] * gp[0]};
    const double Iwxg[3] = {Iw[1] * gp[2] - Iw[2] * gp[1], Iw[2] * gp[0] - Iw[0] * gp[2], Iw[0] * gp[1] - Iw[1] * gp[0]};
    for (int k = 0; k < 3; ++k) gvo[3 + k] += -h * (-I[k] * wxg[k] + Iwxg[k]);
    if (theta != 1.0) {
        const double f = -(1.0 - theta) / theta;
        for (int e = 0; e < 2; ++e) {
            const int j = jbody_blk[4 * (size_t)q + 2 * e], side = jbody_blk[4 * (size_t)q + 2 * e + 1];
            if (j < 0) continue;
            for (int row = jb_off[j]; row < jb_off[j + 1]; ++row) {
                const double* wr = jr + (size_t)row * 16 + 6 * side;
                for (int k = 0; k < 6; ++k) gvo[k] += f * wr[k] * rlb[row];
            }
        }
    }
    store6(gv + (size_t)bi * kBodyStride, gvo);
    store6(gf + (size_t)bi * kBodyStride, gfo);
}

// dL/dcomp per joint row (block order): E = C / (theta h^2) for a spring row, A / (theta h) for a
// damper row enters S as -E, so dL/dE = r_bar_Lambda Lambda; a spring row's right-hand side
// -C Lambda_prev / (theta h^2) - beta (-C Lambda_prev / h + g) / (h theta) adds r_bar dr/dC.
__global__ void k_adj_comp(int n, const int* __restrict__ kind, const double* __restrict__ jr,
                           const double* __restrict__ lam, const double* __restrict__ rlb, double dt, double theta,
                           double beta, double* __restrict__ out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const double lprev = jr[(size_t)r * 16 + 15];
    double g;
    if (kind[r] == 0) {
        g = rlb[r] * lam[r] / (theta * dt * dt);
        g += rlb[r] * (-lprev / (theta * dt * dt) + (beta > 0 ? lprev * beta / (dt * dt * theta) : 0.0));
    } else {
        g = rlb[r] * lam[r] / (theta * dt);
    }
    out[r] = g;
}

}  // namespace cdx

using namespace cdx;

static int adj_alloc(cdx_ctx* ctx) {
    const size_t nc = (size_t)std::max(ctx->nc, 1), nb = (size_t)ctx->nb;  // contact scratch is cleared even without contacts
    CDX_TRY(ensure_alloc(ctx, &ctx->d_adj, (size_t)A_COUNT * nc, &ctx->adj_cap, "adjoint contact scratch"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_adj_imp, 48 * nc, &ctx->adj_imp_cap, "adjoint side vectors"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_adj_body, 8 * kBodyStride * nb, &ctx->adj_body_cap, "adjoint body scratch"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_adj_partial, 2 * (size_t)std::max(grid_for((int)nc), grid_for((int)nb)) + 2,
                         &ctx->adj_partial_cap, "adjoint partials"));
    return CDX_OK;
}

// Restarted GMRES(kGmRestart) on B lambda = g_p, B = I - (I - A R) P^T the transposed linearised
// sweep (lambda_0 = g_p from the setup in S): classical Gram-Schmidt with one reorthogonalisation
// on the device, the small least-squares problem on the host at each restart.  With joint rows the
// body pass hands the jointed bodies' share to the chain solve (jl: its row scratch).  Stops at a
// relative residual <= aprm->tol, after aprm->max_iter applications, or after three restarts in a
// row without a 1 % gain (a (nearly) inconsistent system: lambda is then the least-squares
// solution of the last cycle).  t, I0: body and side-impulse scratch of one application.
static int adjoint_gmres(cdx_ctx* ctx, const cdx_adjoint_params* aprm, bool joints, double* jl, double* S, double* I0,
                         double* t, int* applies_out, double* rel_out, bool* converged_out) {
    const int nb = ctx->nb, nc = ctx->nc;
    const int gc = grid_for(nc), gb = grid_for(nb);
    // restarted GMRES on B lambda = g_p (lambda_0 = g_p from the setup), classical Gram-Schmidt
    // with one reorthogonalisation; the small least-squares problem is solved on the host
    const int n = 3 * nc, m = kGmRestart;
    const int gn = grid_for(n);
    CDX_TRY(ensure_alloc(ctx, &ctx->d_gm, (size_t)(m + 3) * n + (size_t)(m + 1) * (m + 1) + 8, &ctx->gm_cap,
                         "adjoint GMRES vectors"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_gm_part, (size_t)(m + 1) * gn, &ctx->gm_part_cap, "adjoint GMRES partials"));
    double* V = ctx->d_gm;                        // m + 1 basis vectors
    double* Wk = V + (size_t)(m + 1) * n;         // work vector
    double* Xi = Wk + n;                          // P^T v of the last application
    double* H = Xi + n;                           // (m + 1) x m, column-major, then scalars
    double* sc = H + (size_t)(m + 1) * m;         // [0] |.|^2, [1] beta, [2..] second-pass dots
    double* lamv = S + (size_t)A_LAM * nc;
    const double* gvec = S + (size_t)A_GP * nc;
    double* part = ctx->d_gm_part;
    auto apply = [&](const double* v, double* w) -> int {
        CDX_LAUNCH(ctx, "adj_iter", k_gm_pt, gc, kBlock, nc, ctx->d_ida, ctx->d_idb, ctx->d_J, ctx->d_R, ctx->d_mu, S, v, Xi,
                   I0);
        CDX_LAUNCH(ctx, "adj_body", k_adj_body_t, gb, kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_incptr, ctx->d_incidx,
                   I0, t, nullptr, joints ? ctx->d_jslot : nullptr, joints ? ctx->d_jtr : nullptr);
        if (joints) CDX_TRY(joints_apply(ctx, t, jl, nullptr));
        CDX_LAUNCH(ctx, "adj_iter", k_gm_out, gc, kBlock, nc, ctx->d_ida, ctx->d_idb, ctx->d_J, ctx->d_mu, t, v, Xi, w);
        return hip_check(ctx, hipGetLastError(), "adjoint GMRES apply");
    };
    auto dots = [&](int nvec, const double* Vb, const double* w, double* out, int accumulate) -> int {
        CDX_LAUNCH(ctx, "adj_residual", k_gm_dots, gn, kBlock, n, nvec, Vb, (size_t)n, w, part);
        CDX_LAUNCH(ctx, "adj_residual", k_gm_reduce, 1, kBlock, nvec, gn, part, out, accumulate);
        return hip_check(ctx, hipGetLastError(), "adjoint GMRES dots");
    };
    auto host_scalar = [&](const double* d, double* h) -> int {
        CDX_HIP(ctx, hipMemcpyAsync(h, d, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return CDX_OK;
    };
    // the restart cycle's graph (one per call: it holds this call's pointers)
    const bool use_graph = ctx->graphs && !ctx->profiling;
    hipGraphExec_t cycle = nullptr;
    struct CycleGuard {
        hipGraphExec_t* e;
        ~CycleGuard() {
            if (*e) hipGraphExecDestroy(*e);
        }
    } cycle_guard{&cycle};
    double gnorm2 = 0;
    CDX_TRY(dots(1, gvec, gvec, sc, 0));
    CDX_TRY(host_scalar(sc, &gnorm2));
    const double gnorm = std::sqrt(gnorm2);
    int applies = 0, stalls = 0;
    bool converged = gnorm == 0.0, stalled = false;
    double rel = 0.0, rel_prev = INFINITY;
    std::vector<double> Hh((size_t)(m + 1) * m), y(m);
    while (!converged && !stalled && applies < aprm->max_iter) {
        CDX_TRY(apply(lamv, Wk));
        ++applies;
        CDX_LAUNCH(ctx, "adj_residual", k_gm_sub, gn, kBlock, n, gvec, Wk, Wk);
        CDX_TRY(dots(1, Wk, Wk, sc, 0));
        CDX_LAUNCH(ctx, "adj_residual", k_gm_normalize, gn, kBlock, n, Wk, sc, sc + 1, V);
        double beta = 0;
        CDX_TRY(host_scalar(sc + 1, &beta));
        rel = beta / gnorm;
        if (rel <= aprm->tol) {
            converged = true;
            break;
        }
        // three restarts in a row without a 1% gain: the system is (nearly) inconsistent -- contacts
        // on a stick/slip or contact/separation boundary, where the projection is not differentiable
        stalls = rel > 0.99 * rel_prev ? stalls + 1 : 0;
        rel_prev = rel;
        if (stalls >= 3) {
            stalled = true;
            break;
        }
        CDX_HIP(ctx, hipMemsetAsync(H, 0, sizeof(double) * (size_t)(m + 1) * m, ctx->stream));
        // Arnoldi step k: w = B v_k, two Gram-Schmidt passes against v_0..v_k, v_{k+1} = w / |w|
        auto arnoldi = [&](int k) -> int {
            double* vk = V + (size_t)k * n;
            CDX_TRY(apply(vk, Wk));
            double* hk = H + (size_t)k * (m + 1);
            CDX_TRY(dots(k + 1, V, Wk, hk, 0));
            CDX_LAUNCH(ctx, "adj_residual", k_gm_axpy, gn, kBlock, n, k + 1, V, (size_t)n, hk, -1.0, Wk);
            CDX_TRY(dots(k + 1, V, Wk, sc + 2, 0));
            CDX_LAUNCH(ctx, "adj_residual", k_gm_axpy, gn, kBlock, n, k + 1, V, (size_t)n, sc + 2, -1.0, Wk);
            CDX_LAUNCH(ctx, "adj_residual", k_gm_add, 1, kBlock, k + 1, sc + 2, hk);
            CDX_TRY(dots(1, Wk, Wk, sc, 0));
            CDX_LAUNCH(ctx, "adj_residual", k_gm_normalize, gn, kBlock, n, Wk, sc, hk + k + 1, V + (size_t)(k + 1) * n);
            return hip_check(ctx, hipGetLastError(), "adjoint GMRES Arnoldi step");
        };
        int k = 0;
        if (use_graph && aprm->max_iter - applies >= m) {
            // a full cycle is the same launches every restart: captured once, replayed
            if (!cycle) {
                CDX_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                int cst = CDX_OK;
                for (int kk = 0; kk < m && cst == CDX_OK; ++kk) cst = arnoldi(kk);
                hipGraph_t g = nullptr;
                const hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
                if (cst) {
                    if (g) hipGraphDestroy(g);
                    return cst;
                }
                CDX_HIP(ctx, ce);
                const hipError_t ie = hipGraphInstantiate(&cycle, g, nullptr, nullptr, 0);
                hipGraphDestroy(g);
                CDX_HIP(ctx, ie);
            }
            CDX_HIP(ctx, hipGraphLaunch(cycle, ctx->stream));
            applies += m;
            k = m;
        } else {
            for (; k < m && applies < aprm->max_iter; ++k) {
                CDX_TRY(arnoldi(k));
                ++applies;
            }
        }
        // least squares min |beta e1 - H y| by Givens rotations
        CDX_HIP(ctx, hipMemcpyAsync(Hh.data(), H, sizeof(double) * Hh.size(), hipMemcpyDeviceToHost, ctx->stream));
        CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        std::vector<double> gq(k + 1, 0.0), cs(k), sn(k);
        gq[0] = beta;
        for (int c = 0; c < k; ++c) {
            double* col = &Hh[(size_t)c * (m + 1)];
            for (int i = 0; i < c; ++i) {
                const double a0 = col[i], a1 = col[i + 1];
                col[i] = cs[i] * a0 + sn[i] * a1;
                col[i + 1] = -sn[i] * a0 + cs[i] * a1;
            }
            const double r = std::hypot(col[c], col[c + 1]);
            cs[c] = r > 0 ? col[c] / r : 1.0;
            sn[c] = r > 0 ? col[c + 1] / r : 0.0;
            col[c] = r;
            col[c + 1] = 0.0;
            const double g0 = gq[c];
            gq[c] = cs[c] * g0;
            gq[c + 1] = -sn[c] * g0;
        }
        for (int i = k - 1; i >= 0; --i) {
            double s = gq[i];
            for (int l = i + 1; l < k; ++l) s -= Hh[(size_t)l * (m + 1) + i] * y[l];
            const double d = Hh[(size_t)i * (m + 1) + i];
            y[i] = d != 0.0 ? s / d : 0.0;
        }
        CDX_HIP(ctx, hipMemcpyAsync(sc + 2, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, ctx->stream));
        CDX_LAUNCH(ctx, "adj_residual", k_gm_axpy, gn, kBlock, n, k, V, (size_t)n, sc + 2, 1.0, lamv);
        rel = std::fabs(gq[k]) / gnorm;
        CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    *applies_out = applies;
    *rel_out = rel;
    *converged_out = converged;
    return CDX_OK;
}

// adjoint_gmres with one Krylov space per scene of the context's batch (cdx_set_scenes): the
// applications of B are shared, the dot products, coefficients, least-squares problems and
// convergence / stall tests are per scene.  A scene that has converged or stalled keeps its lambda.
// *rel_out is the worst scene's relative residual.
static int adjoint_gmres_scenes(cdx_ctx* ctx, const cdx_adjoint_params* aprm, double* S, double* I0, double* t,
                                int* applies_out, double* rel_out, bool* converged_out) {
    const int nb = ctx->nb, nc = ctx->nc, ns = ctx->nscenes;
    const int gc = grid_for(nc), gb = grid_for(nb);
    const int n = 3 * nc, m = kGmRestart;
    const int gn = grid_for(n);
    const size_t hs = (size_t)(m + 1) * m;  // one scene's Hessenberg, column-major
    const size_t ss = (size_t)(m + 4);      // one scene's scalars: [0] |.|^2, [1] beta, [2..] second-pass dots
    CDX_TRY(ensure_alloc(ctx, &ctx->d_gm, (size_t)(m + 3) * n + hs * ns + ss * ns + (size_t)(m + 1) * ns + (size_t)nc + 8,
                         &ctx->gm_cap, "adjoint GMRES vectors"));
    double* V = ctx->d_gm;
    double* Wk = V + (size_t)(m + 1) * n;
    double* Xi = Wk + n;
    double* H = Xi + n;               // ns Hessenbergs
    double* sc = H + hs * ns;         // ns scalar blocks
    double* Y = sc + ss * ns;         // ns coefficient vectors
    int* cscene = reinterpret_cast<int*>(Y + (size_t)(m + 1) * ns);  // scene of each contact
    {
        std::vector<int> h(nc);
        for (int q = 0; q < ns; ++q)
            for (int c = ctx->scene_coff[q]; c < ctx->scene_coff[q + 1]; ++c) h[c] = q;
        if (nc) CDX_HIP(ctx, hipMemcpyAsync(cscene, h.data(), sizeof(int) * nc, hipMemcpyHostToDevice, ctx->stream));
    }
    const int* coff = ctx->d_scene_i + ns + ns + 1;  // scene table: [order ns | boff ns+1 | coff ns+1 | ...]
    double* lamv = S + (size_t)A_LAM * nc;
    const double* gvec = S + (size_t)A_GP * nc;
    auto apply = [&](const double* v, double* w) -> int {
        CDX_LAUNCH(ctx, "adj_iter", k_gm_pt, gc, kBlock, nc, ctx->d_ida, ctx->d_idb, ctx->d_J, ctx->d_R, ctx->d_mu, S, v, Xi,
                   I0);
        CDX_LAUNCH(ctx, "adj_body", k_adj_body_t, gb, kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_incptr, ctx->d_incidx,
                   I0, t, nullptr, nullptr, nullptr);
        CDX_LAUNCH(ctx, "adj_iter", k_gm_out, gc, kBlock, nc, ctx->d_ida, ctx->d_idb, ctx->d_J, ctx->d_mu, t, v, Xi, w);
        return hip_check(ctx, hipGetLastError(), "adjoint GMRES apply");
    };
    auto sdots = [&](int nvec, const double* Vb, const double* w, double* out, size_t ostride) -> int {
        CDX_LAUNCH(ctx, "adj_residual", k_gms_dots, ns, kBlock, ns, coff, nc, nvec, Vb, (size_t)n, w, out, ostride);
        return hip_check(ctx, hipGetLastError(), "adjoint GMRES scene dots");
    };
    const int gnorm_grid = grid_for(std::max(n, ns));
    std::vector<double> gn2(ns), hsc((size_t)ns * ss);
    CDX_TRY(sdots(1, gvec, gvec, sc, ss));
    CDX_HIP(ctx, hipMemcpyAsync(hsc.data(), sc, sizeof(double) * hsc.size(), hipMemcpyDeviceToHost, ctx->stream));
    CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<double> gnorm(ns), rel(ns, 0.0), rel_prev(ns, INFINITY);
    std::vector<int> stalls(ns, 0);
    std::vector<char> done(ns, 0), conv(ns, 0);
    for (int q = 0; q < ns; ++q) {
        gnorm[q] = std::sqrt(hsc[(size_t)q * ss]);
        if (!(gnorm[q] > 0)) done[q] = conv[q] = 1;
    }
    const bool use_graph = ctx->graphs && !ctx->profiling;
    hipGraphExec_t cycle = nullptr;
    struct CycleGuard {
        hipGraphExec_t* e;
        ~CycleGuard() {
            if (*e) hipGraphExecDestroy(*e);
        }
    } cycle_guard{&cycle};
    int applies = 0;
    std::vector<double> Hh(hs * ns), ys((size_t)(m + 1) * ns);
    auto all_done = [&]() {
        for (char d : done)
            if (!d) return false;
        return true;
    };
    while (!all_done() && applies < aprm->max_iter) {
        CDX_TRY(apply(lamv, Wk));
        ++applies;
        CDX_LAUNCH(ctx, "adj_residual", k_gm_sub, gn, kBlock, n, gvec, Wk, Wk);
        CDX_TRY(sdots(1, Wk, Wk, sc, ss));
        CDX_LAUNCH(ctx, "adj_residual", k_gms_normalize, gnorm_grid, kBlock, nc, cscene, Wk, sc, ss, sc + 1, ss, V, ns);
        CDX_HIP(ctx, hipMemcpyAsync(hsc.data(), sc, sizeof(double) * hsc.size(), hipMemcpyDeviceToHost, ctx->stream));
        CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (int q = 0; q < ns; ++q) {
            if (done[q]) continue;
            rel[q] = hsc[(size_t)q * ss + 1] / gnorm[q];
            if (rel[q] <= aprm->tol) {
                done[q] = conv[q] = 1;
                continue;
            }
            // three restarts in a row without a 1 % gain: (nearly) inconsistent -- keep the
            // least-squares lambda of the last cycle
            stalls[q] = rel[q] > 0.99 * rel_prev[q] ? stalls[q] + 1 : 0;
            rel_prev[q] = rel[q];
            if (stalls[q] >= 3) done[q] = 1;
        }
        if (all_done()) break;
        CDX_HIP(ctx, hipMemsetAsync(H, 0, sizeof(double) * hs * ns, ctx->stream));
        auto arnoldi = [&](int k) -> int {
            double* vk = V + (size_t)k * n;
            CDX_TRY(apply(vk, Wk));
            double* hk = H + (size_t)k * (m + 1);  // scene q's column k at hk + q * hs
            CDX_TRY(sdots(k + 1, V, Wk, hk, hs));
            CDX_LAUNCH(ctx, "adj_residual", k_gms_axpy, gn, kBlock, nc, cscene, k + 1, V, (size_t)n, hk, hs, -1.0, Wk);
            CDX_TRY(sdots(k + 1, V, Wk, sc + 2, ss));
            CDX_LAUNCH(ctx, "adj_residual", k_gms_axpy, gn, kBlock, nc, cscene, k + 1, V, (size_t)n, sc + 2, ss, -1.0, Wk);
            CDX_LAUNCH(ctx, "adj_residual", k_gms_add, grid_for(ns * (k + 1)), kBlock, ns, k + 1, sc + 2, ss, hk, hs);
            CDX_TRY(sdots(1, Wk, Wk, sc, ss));
            CDX_LAUNCH(ctx, "adj_residual", k_gms_normalize, gnorm_grid, kBlock, nc, cscene, Wk, sc, ss, hk + k + 1, hs,
                       V + (size_t)(k + 1) * n, ns);
            return hip_check(ctx, hipGetLastError(), "adjoint GMRES Arnoldi step");
        };
        int k = 0;
        if (use_graph && aprm->max_iter - applies >= m) {
            if (!cycle) {
                CDX_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                int cst = CDX_OK;
                for (int kk = 0; kk < m && cst == CDX_OK; ++kk) cst = arnoldi(kk);
                hipGraph_t g = nullptr;
                const hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
                if (cst) {
                    if (g) hipGraphDestroy(g);
                    return cst;
                }
                CDX_HIP(ctx, ce);
                const hipError_t ie = hipGraphInstantiate(&cycle, g, nullptr, nullptr, 0);
                hipGraphDestroy(g);
                CDX_HIP(ctx, ie);
            }
            CDX_HIP(ctx, hipGraphLaunch(cycle, ctx->stream));
            applies += m;
            k = m;
        } else {
            for (; k < m && applies < aprm->max_iter; ++k) {
                CDX_TRY(arnoldi(k));
                ++applies;
            }
        }
        // per scene: least squares min |beta e1 - H y| by Givens rotations (zero update once done)
        CDX_HIP(ctx, hipMemcpyAsync(Hh.data(), H, sizeof(double) * Hh.size(), hipMemcpyDeviceToHost, ctx->stream));
        CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        std::fill(ys.begin(), ys.end(), 0.0);
        for (int q = 0; q < ns; ++q) {
            if (done[q]) continue;
            double* Hq = &Hh[(size_t)q * hs];
            double* y = &ys[(size_t)q * (m + 1)];
            std::vector<double> gq(k + 1, 0.0), cs(k), sn(k);
            gq[0] = hsc[(size_t)q * ss + 1];
            for (int c = 0; c < k; ++c) {
                double* col = Hq + (size_t)c * (m + 1);
                for (int i = 0; i < c; ++i) {
                    const double a0 = col[i], a1 = col[i + 1];
                    col[i] = cs[i] * a0 + sn[i] * a1;
                    col[i + 1] = -sn[i] * a0 + cs[i] * a1;
                }
                const double r = std::hypot(col[c], col[c + 1]);
                cs[c] = r > 0 ? col[c] / r : 1.0;
                sn[c] = r > 0 ? col[c + 1] / r : 0.0;
                col[c] = r;
                col[c + 1] = 0.0;
                const double g0 = gq[c];
                gq[c] = cs[c] * g0;
                gq[c + 1] = -sn[c] * g0;
            }
            for (int i = k - 1; i >= 0; --i) {
                double sum = gq[i];
                for (int l = i + 1; l < k; ++l) sum -= Hq[(size_t)l * (m + 1) + i] * y[l];
                const double d = Hq[(size_t)i * (m + 1) + i];
                y[i] = d != 0.0 ? sum / d : 0.0;
            }
            rel[q] = std::fabs(gq[k]) / gnorm[q];
        }
        CDX_HIP(ctx, hipMemcpyAsync(Y, ys.data(), sizeof(double) * ys.size(), hipMemcpyHostToDevice, ctx->stream));
        CDX_LAUNCH(ctx, "adj_residual", k_gms_axpy, gn, kBlock, nc, cscene, k, V, (size_t)n, Y, (size_t)(m + 1), 1.0, lamv);
        CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    double worst = 0;
    bool all_conv = true;
    for (int q = 0; q < ns; ++q) {
        worst = std::max(worst, rel[q]);
        all_conv = all_conv && conv[q];
    }
    *applies_out = applies;
    *rel_out = worst;
    *converged_out = all_conv;
    return CDX_OK;
}

static int adjoint_run(cdx_ctx* ctx, double dt, const cdx_solver_params* prm, const cdx_adjoint_params* aprm,
                       const double* g_v, const double* g_p, double* gv_out, double* gminv_out, double* gfext_out,
                       double* gmu_out, double** dlp_out, double** dgmu_out, int* iters_out, double* res_out, int* mode_io) {
    const int nb = ctx->nb, nc = ctx->nc;
    double* B = ctx->d_adj_body;  // [lx | t (later gf) | xbar | gminv_acc | gv | gm]
    double* lx = B;
    double* t = B + (size_t)kBodyStride * nb;
    double* xbar = B + 2 * (size_t)kBodyStride * nb;
    double* gacc = B + 3 * (size_t)kBodyStride * nb;
    double* gvd = B + 4 * (size_t)kBodyStride * nb;
    double* gmd = B + 5 * (size_t)kBodyStride * nb;
    double* gfd = t;  // t is free again by the final stage
    {
        std::vector<double> h((size_t)kBodyStride * nb, 0.0);
        for (int b = 0; b < nb; ++b)
            for (int k = 0; k < ctx->dof[b]; ++k) h[(size_t)b * kBodyStride + k] = g_v[ctx->voff[b] + k];
        if (nb) CDX_HIP(ctx, hipMemcpyAsync(lx, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        if (nb) CDX_HIP(ctx, hipMemsetAsync(gacc, 0, sizeof(double) * kBodyStride * (size_t)nb, ctx->stream));
        CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (g_p && nc) {
        CDX_HIP(ctx, hipMalloc((void**)dlp_out, sizeof(double) * 3 * (size_t)nc));
        CDX_HIP(ctx, hipMemcpy(*dlp_out, g_p, sizeof(double) * 3 * (size_t)nc, hipMemcpyHostToDevice));
    }
    double* dlp = *dlp_out;
    CDX_HIP(ctx, hipMalloc((void**)dgmu_out, sizeof(double) * (size_t)(nc ? nc : 1)));
    double* dgmu = *dgmu_out;
    double* S = ctx->d_adj;
    double* I0 = ctx->d_adj_imp;  // 4 side-vector buffers of 12 doubles per contact
    double* I1 = I0 + 12 * (size_t)nc;
    double* I2 = I0 + 24 * (size_t)nc;
    double* I3 = I0 + 36 * (size_t)nc;
    const int gc = grid_for(nc), gb = grid_for(nb);
    const bool rows = nc > 0 && ctx->nrows > 0;
    int& mode = *mode_io;
    if (rows && mode == 2) {
        // the same linear system by restarted GMRES (the least-squares adjoint where it is
        // inconsistent: impulses that are not unique, contacts on a stick/slip boundary)
        CDX_LAUNCH(ctx, "adj_setup", k_adj_setup, gc, kBlock, nc, ctx->d_ida, ctx->d_idb, ctx->d_J, ctx->d_R, ctx->d_bias,
                   ctx->d_mu, ctx->d_rest, ctx->d_vb, ctx->d_minv, ctx->d_v, ctx->d_x, ctx->d_p, lx, dlp, S, I0, 0);
        bool converged = false;
        if (ctx->nscenes >= 1 && !ctx->scene_coff.empty() && ctx->scene_coff.back() == nc)
            CDX_TRY(adjoint_gmres_scenes(ctx, aprm, S, I0, t, iters_out, res_out, &converged));
        else
            CDX_TRY(adjoint_gmres(ctx, aprm, false, nullptr, S, I0, t, iters_out, res_out, &converged));
        CDX_LAUNCH(ctx, "adj_iter", k_gm_pt, gc, kBlock, nc, ctx->d_ida, ctx->d_idb, ctx->d_J, ctx->d_R, ctx->d_mu, S,
                   S + (size_t)A_LAM * nc, S + (size_t)A_XI * nc, I0);
        CDX_HIP(ctx, hipGetLastError());
    }
    if (rows && mode == 0) {
        CDX_TRY(reset_control(ctx));
        CDX_LAUNCH(ctx, "adj_setup", k_adj_setup, gc, kBlock, nc, ctx->d_ida, ctx->d_idb, ctx->d_J, ctx->d_R, ctx->d_bias,
                   ctx->d_mu, ctx->d_rest, ctx->d_vb, ctx->d_minv, ctx->d_v, ctx->d_x, ctx->d_p, lx, dlp, S, I0, 0);
        CDX_LAUNCH(ctx, "adj_body", k_adj_body_t, gb, kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_incptr, ctx->d_incidx, I0, t,
                   nullptr, nullptr, nullptr);
        auto iter = [&](int it) -> int {
            const int check = (it % 10) == 9;
            CDX_LAUNCH(ctx, "adj_iter", k_adj_iter, gc, kBlock, nc, ctx->d_ida, ctx->d_idb, ctx->d_J, ctx->d_R, ctx->d_mu, t, S,
                       I0, check, ctx->d_adj_partial, ctx->d_ctrl);
            CDX_LAUNCH(ctx, "adj_body", k_adj_body_t, gb, kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_incptr, ctx->d_incidx,
                       I0, t, ctx->d_ctrl, nullptr, nullptr);
            if (check)
                CDX_LAUNCH(ctx, "adj_residual", k_adj_finalize, 1, kBlock, gc, ctx->d_adj_partial, it, aprm->tol, ctx->d_ctrl,
                           ctx->d_res, ctx->h_flag_dev);
            return hip_check(ctx, hipGetLastError(), "adjoint iteration launch");
        };
        int it_done = 0;
        CDX_TRY(run_sweeps(ctx, aprm->max_iter, 10, 4, iter, &it_done));
        int ctrl[3];
        CDX_HIP(ctx, hipMemcpy(ctrl, ctx->d_ctrl, sizeof(ctrl), hipMemcpyDeviceToHost));
        CDX_TRY(read_residual(ctx, res_out));
        *iters_out = it_done;
        if (!ctrl[0])  // not converged within max_iter
            return set_err(ctx, CDX_DIVERGED, "cdx_pj_adjoint: implicit adjoint iteration did not converge");
    }
    if (rows && mode == 1) {
        if (!ctx->record_valid)
            return set_err(ctx, CDX_BAD_ARG, "cdx_pj_adjoint: unrolled mode needs the forward solve recorded (cdx_set_recording)");
        const int K = ctx->recorded;
        CDX_HIP(ctx, hipMemsetAsync(S, 0, sizeof(double) * A_COUNT * (size_t)nc, ctx->stream));
        if (dlp) {  // p_bar_K = L_p (component-major)
            std::vector<double> hl(3 * (size_t)nc), cm(3 * (size_t)nc);
            for (int c = 0; c < nc; ++c)
                for (int k = 0; k < 3; ++k) cm[(size_t)k * nc + c] = g_p[3 * (size_t)c + k];
            CDX_HIP(ctx, hipMemcpyAsync(S + (size_t)A_LAM * nc, cm.data(), sizeof(double) * cm.size(), hipMemcpyHostToDevice,
                                        ctx->stream));
            CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        }
        CDX_HIP(ctx, hipMemcpyAsync(xbar, lx, sizeof(double) * kBodyStride * (size_t)nb, hipMemcpyDeviceToDevice, ctx->stream));
        for (int k = K - 1; k >= 0; --k) {
            const double* xk = ctx->d_rec_x + (size_t)k * kBodyStride * nb;
            const double* pk = ctx->d_rec_p + (size_t)k * 3 * nc;
            const double* pk1 = (k + 1 < K) ? ctx->d_rec_p + (size_t)(k + 1) * 3 * nc : ctx->d_p;
            CDX_LAUNCH(ctx, "unr_contacts", k_unr_contacts, gc, kBlock, nc, ctx->d_ida, ctx->d_idb, ctx->d_J, ctx->d_R,
                       ctx->d_bias, ctx->d_mu, ctx->d_rest, ctx->d_vb, ctx->d_minv, ctx->d_v, xk, pk, pk1, xbar, S, I0, I1);
            CDX_LAUNCH(ctx, "unr_bodies", k_unr_bodies, gb, kBlock, nb, ctx->d_dof, ctx->d_incptr, ctx->d_incidx, I0, I1, xbar,
                       gacc);
        }
        CDX_HIP(ctx, hipGetLastError());
        *iters_out = K;
        *res_out = 0.0;
    }
    // final stage (shared by the modes; GMRES leaves the implicit state)
    const bool implicit = mode != 1;
    const double* pv = implicit ? ctx->d_p : ctx->d_rec_p;  // p* (implicit) or p_0 (unrolled)
    const double* base = implicit ? lx : xbar;
    if (rows) {
        CDX_LAUNCH(ctx, "adj_final", k_adj_final_contacts, gc, kBlock, nc, implicit ? 1 : 0, prm->alpha, ctx->d_ida, ctx->d_idb,
                   ctx->d_J, ctx->d_R, ctx->d_mu, ctx->d_rest, pv, S, I0, I1, I2, I3, dgmu);
    } else {
        CDX_HIP(ctx, hipMemsetAsync(I0, 0, sizeof(double) * 48 * (size_t)(nc ? nc : 1), ctx->stream));
        CDX_HIP(ctx, hipMemsetAsync(dgmu, 0, sizeof(double) * (size_t)(nc ? nc : 1), ctx->stream));
    }
    if (nb)
        CDX_LAUNCH(ctx, "adj_final_bodies", k_adj_final_bodies, gb, kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_v, ctx->d_fext,
                   ctx->d_fgyro, ctx->d_incptr, ctx->d_incidx, I0, I1, I2, I3, base, gacc, dt, gvd, gmd, gfd, nullptr, nullptr);
    CDX_HIP(ctx, hipGetLastError());
    CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<double> hv((size_t)kBodyStride * nb), hm((size_t)kBodyStride * nb), hf((size_t)kBodyStride * nb);
    if (nb) {
        CDX_HIP(ctx, hipMemcpy(hv.data(), gvd, hv.size() * sizeof(double), hipMemcpyDeviceToHost));
        CDX_HIP(ctx, hipMemcpy(hm.data(), gmd, hm.size() * sizeof(double), hipMemcpyDeviceToHost));
        CDX_HIP(ctx, hipMemcpy(hf.data(), gfd, hf.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    for (int b = 0; b < nb; ++b)
        for (int k = 0; k < ctx->dof[b]; ++k) {
            const int o = ctx->voff[b] + k;
            if (gv_out) gv_out[o] = hv[(size_t)b * kBodyStride + k];
            if (gminv_out) gminv_out[o] = hm[(size_t)b * kBodyStride + k];
            if (gfext_out) gfext_out[o] = hf[(size_t)b * kBodyStride + k];
        }
    if (gmu_out && nc) CDX_HIP(ctx, hipMemcpy(gmu_out, dgmu, sizeof(double) * (size_t)nc, hipMemcpyDeviceToHost));
    return CDX_OK;
}

// Implicit adjoint of a solve with joint rows (see cdx_pj_adjoint_joints): the iteration and the
// final stage of adjoint_run with S^-1_vv in place of M^-1 (the chain solve applies the joint
// rows' share), then r_bar = S^-1 [x_bar; 0] gives dL/dv_n, dL/df_ext and dL/dcomp.
static int adjoint_joints_run(cdx_ctx* ctx, double dt, const cdx_solver_params* prm, const cdx_adjoint_params* aprm,
                              const double* g_v, const double* g_p, double* gv_out, double* gfext_out, double* gmu_out,
                              double* gcomp_out, double** dlp_out, double** dgmu_out, int* iters_out, double* res_out) {
    const int nb = ctx->nb, nc = ctx->nc, nj = ctx->njrows;
    const double theta = ctx->last_theta;
    const size_t n8 = (size_t)kBodyStride * nb;
    double* B = ctx->d_adj_body;  // [lx | t | tlx | gacc | gv | gm | gf | rv]
    double *lx = B, *t = B + n8, *tlx = B + 2 * n8, *gacc = B + 3 * n8, *gvd = B + 4 * n8, *gmd = B + 5 * n8,
           *gfd = B + 6 * n8, *rv = B + 7 * n8;
    CDX_TRY(ensure_alloc(ctx, &ctx->d_adj_jl, 3 * (size_t)nj, &ctx->adj_jl_cap, "joint adjoint rows"));
    double *jl = ctx->d_adj_jl, *rlb = jl + nj, *gcomp = jl + 2 * (size_t)nj;
    {
        std::vector<double> h(n8, 0.0);
        for (int b = 0; b < nb; ++b)
            for (int k = 0; k < ctx->dof[b]; ++k) h[(size_t)b * kBodyStride + k] = g_v[ctx->voff[b] + k];
        CDX_HIP(ctx, hipMemcpyAsync(lx, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        CDX_HIP(ctx, hipMemsetAsync(gacc, 0, sizeof(double) * n8, ctx->stream));
        CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (g_p && nc) {
        CDX_HIP(ctx, hipMalloc((void**)dlp_out, sizeof(double) * 3 * (size_t)nc));
        CDX_HIP(ctx, hipMemcpy(*dlp_out, g_p, sizeof(double) * 3 * (size_t)nc, hipMemcpyHostToDevice));
    }
    double* dlp = *dlp_out;
    CDX_HIP(ctx, hipMalloc((void**)dgmu_out, sizeof(double) * (size_t)(nc ? nc : 1)));
    double* dgmu = *dgmu_out;
    double* S = ctx->d_adj;
    double* I0 = ctx->d_adj_imp;
    double* I1 = I0 + 12 * (size_t)nc;
    double* I2 = I0 + 24 * (size_t)nc;
    double* I3 = I0 + 36 * (size_t)nc;
    const int gc = grid_for(nc), gb = grid_for(nb);
    const bool rows = nc > 0 && ctx->nrows > 0;
    *iters_out = 0;
    *res_out = 0.0;
    if (rows) {
        // tlx = S^-1_vv L_x
        CDX_LAUNCH(ctx, "adj_body", k_adj_minv_apply, gb, kBlock, nb, ctx->d_dof, ctx->d_minv, lx, tlx, ctx->d_jslot, ctx->d_jtr);
        CDX_HIP(ctx, hipMemsetAsync(jl, 0, sizeof(double) * nj, ctx->stream));
        CDX_TRY(joints_apply(ctx, tlx, jl, nullptr));
        CDX_TRY(reset_control(ctx));
        CDX_LAUNCH(ctx, "adj_setup", k_adj_setup, gc, kBlock, nc, ctx->d_ida, ctx->d_idb, ctx->d_J, ctx->d_R, ctx->d_bias,
                   ctx->d_mu, ctx->d_rest, ctx->d_vb, ctx->d_minv, ctx->d_v, ctx->d_x, ctx->d_p, tlx, dlp, S, I0, 1);
        CDX_LAUNCH(ctx, "adj_body", k_adj_body_t, gb, kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_incptr, ctx->d_incidx, I0, t,
                   nullptr, ctx->d_jslot, ctx->d_jtr);
        CDX_TRY(joints_apply(ctx, t, jl, nullptr));
        int applies = 0;
        double rel = 0.0;
        bool converged = false;
        CDX_TRY(adjoint_gmres(ctx, aprm, true, jl, S, I0, t, &applies, &rel, &converged));
        *iters_out = applies;
        *res_out = rel;
        if (!converged && aprm->mode != CDX_ADJ_AUTO)
            return set_err(ctx, CDX_DIVERGED,
                           "cdx_pj_adjoint_joints: adjoint GMRES did not converge (relative residual " + std::to_string(rel) +
                               " after " + std::to_string(applies) + " applications)");
        // xi = P^T lambda for the final stage
        CDX_LAUNCH(ctx, "adj_iter", k_gm_pt, gc, kBlock, nc, ctx->d_ida, ctx->d_idb, ctx->d_J, ctx->d_R, ctx->d_mu, S,
                   S + (size_t)A_LAM * nc, S + (size_t)A_XI * nc, I0);
        CDX_LAUNCH(ctx, "adj_final", k_adj_final_contacts, gc, kBlock, nc, 1, prm->alpha, ctx->d_ida, ctx->d_idb, ctx->d_J,
                   ctx->d_R, ctx->d_mu, ctx->d_rest, ctx->d_p, S, I0, I1, I2, I3, dgmu);
    } else {
        CDX_HIP(ctx, hipMemsetAsync(I0, 0, sizeof(double) * 48 * (size_t)(nc ? nc : 1), ctx->stream));
        CDX_HIP(ctx, hipMemsetAsync(dgmu, 0, sizeof(double) * (size_t)(nc ? nc : 1), ctx->stream));
    }
    // x_bar = L_x - W^T s; non-jointed bodies finish here, jointed ones hand x_bar to the chains
    CDX_LAUNCH(ctx, "adj_final_bodies", k_adj_final_bodies, gb, kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_v, ctx->d_fext,
               ctx->d_fgyro, ctx->d_incptr, ctx->d_incidx, I0, I1, I2, I3, lx, gacc, dt, gvd, gmd, gfd, ctx->d_jslot,
               ctx->d_jtr);
    CDX_HIP(ctx, hipMemsetAsync(rv, 0, sizeof(double) * n8, ctx->stream));
    CDX_HIP(ctx, hipMemsetAsync(rlb, 0, sizeof(double) * nj, ctx->stream));
    CDX_TRY(joints_apply(ctx, rv, rlb, nullptr));
    CDX_LAUNCH(ctx, "adj_final_bodies", k_adj_final_jointed, grid_for(ctx->njbodies), kBlock, ctx->njbodies, ctx->d_jbody,
               ctx->d_jbody_blk, ctx->d_jb_off, ctx->d_jr, rlb, ctx->d_minv, ctx->d_v, rv, dt, theta, gvd, gfd);
    CDX_LAUNCH(ctx, "adj_final_bodies", k_adj_comp, grid_for(nj), kBlock, nj, ctx->d_jr_kind, ctx->d_jr, ctx->d_jlam, rlb,
               dt, theta, ctx->joint_bias, gcomp);
    CDX_HIP(ctx, hipGetLastError());
    CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<double> hv(n8), hf(n8), hc(nj);
    CDX_HIP(ctx, hipMemcpy(hv.data(), gvd, n8 * sizeof(double), hipMemcpyDeviceToHost));
    CDX_HIP(ctx, hipMemcpy(hf.data(), gfd, n8 * sizeof(double), hipMemcpyDeviceToHost));
    CDX_HIP(ctx, hipMemcpy(hc.data(), gcomp, (size_t)nj * sizeof(double), hipMemcpyDeviceToHost));
    for (int b = 0; b < nb; ++b)
        for (int k = 0; k < ctx->dof[b]; ++k) {
            const int o = ctx->voff[b] + k;
            if (gv_out) gv_out[o] = hv[(size_t)b * kBodyStride + k];
            if (gfext_out) gfext_out[o] = hf[(size_t)b * kBodyStride + k];
        }
    if (gcomp_out)
        for (int i = 0; i < nj; ++i) gcomp_out[ctx->h_jorig[i]] = hc[i];
    if (gmu_out && nc) CDX_HIP(ctx, hipMemcpy(gmu_out, dgmu, sizeof(double) * (size_t)nc, hipMemcpyDeviceToHost));
    return CDX_OK;
}

extern "C" {

int cdx_set_recording(cdx_ctx* ctx, int enable, int max_sweeps) {
    if (!ctx) return CDX_BAD_ARG;
    if (enable && max_sweeps < 1) return set_err(ctx, CDX_BAD_ARG, "cdx_set_recording: max_sweeps must be >= 1");
    ctx->record = enable != 0;
    ctx->record_max = enable ? max_sweeps : 0;
    ctx->recorded = 0;
    return CDX_OK;
}

int cdx_pj_adjoint(cdx_ctx* ctx, double dt, const cdx_solver_params* prm, const cdx_adjoint_params* aprm, const double* g_v,
                   const double* g_p, double* gv_out, double* gminv_out, double* gfext_out, double* gmu_out,
                   cdx_adjoint_info* info) {
    if (!ctx || !prm || !aprm || !g_v) return CDX_BAD_ARG;
    hipSetDevice(ctx->device);
    if (!ctx->solved || !ctx->pj_last) return set_err(ctx, CDX_NO_STATE, "cdx_pj_adjoint: needs a preceding cdx_pj_solve");
    if (ctx->gyro_last || prm->implicit_gyro)
        return set_err(ctx, CDX_BAD_ARG, "cdx_pj_adjoint: differentiates the explicit-gyroscopy solve only");
    if (aprm->mode < 0 || aprm->mode > 2 || aprm->max_iter < 1) return set_err(ctx, CDX_BAD_ARG, "cdx_pj_adjoint: bad params");
    if (info) {
        info->iters = 0;
        info->residual = std::nan("");
        info->mode_used = aprm->mode;
    }
    CDX_TRY(adj_alloc(ctx));
    double* dlp = nullptr;
    double* dgmu = nullptr;
    int iters = 0;
    double res = std::nan("");
    int mode = aprm->mode == CDX_ADJ_UNROLLED ? 1 : (aprm->mode == CDX_ADJ_AUTO ? 2 : 0);
    const int st = adjoint_run(ctx, dt, prm, aprm, g_v, g_p, gv_out, gminv_out, gfext_out, gmu_out, &dlp, &dgmu, &iters,
                               &res, &mode);
    if (dlp) hipFree(dlp);
    if (dgmu) hipFree(dgmu);
    if (info) {
        info->iters = iters;
        info->residual = res;
        info->mode_used = aprm->mode;
    }
    return st;
}

int cdx_pj_adjoint_joints(cdx_ctx* ctx, double dt, const cdx_solver_params* prm, const cdx_adjoint_params* aprm,
                          const double* g_v, const double* g_p, double* gv_out, double* gfext_out, double* gmu_out,
                          double* gcomp_out, cdx_adjoint_info* info) {
    if (!ctx || !prm || !aprm || !g_v) return CDX_BAD_ARG;
    hipSetDevice(ctx->device);
    if (!ctx->solved || !ctx->pjj_last || ctx->njrows == 0)
        return set_err(ctx, CDX_NO_STATE, "cdx_pj_adjoint_joints: needs a preceding joint solve (cdx_pj_solve or cdx_batch_pj_solve)");
    if (aprm->mode == CDX_ADJ_UNROLLED || aprm->mode < 0 || aprm->mode > 2 || aprm->max_iter < 1)
        return set_err(ctx, CDX_BAD_ARG, "cdx_pj_adjoint_joints: implicit mode only");
    if (info) {
        info->iters = 0;
        info->residual = std::nan("");
        info->mode_used = CDX_ADJ_IMPLICIT;
    }
    CDX_TRY(adj_alloc(ctx));
    double* dlp = nullptr;
    double* dgmu = nullptr;
    int iters = 0;
    double res = std::nan("");
    const int st = adjoint_joints_run(ctx, dt, prm, aprm, g_v, g_p, gv_out, gfext_out, gmu_out, gcomp_out, &dlp, &dgmu,
                                      &iters, &res);
    if (dlp) hipFree(dlp);
    if (dgmu) hipFree(dgmu);
    if (info) {
        info->iters = iters;
        info->residual = res;
    }
    return st;
}

}  // extern "C"

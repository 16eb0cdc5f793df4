// cdx_pj.hip -- FP64 projected Jacobi contact solve on gfx950.
//
// Replaces ProjectedJacobiSolver::solve (src/physics/solver/projected_jacobi.cpp:451-500) for
// contact scenes, where S = diag(M) (pj_assembler.cpp:26-30) so every S^-1 is a division by the
// mass diagonal.  One sweep (pj_sweep, projected_jacobi.cpp:121-138) is two kernels:
//
//   pj_contacts : per contact (one lane each, contacts in the caller's order): gather the two
//                 body twists, gamma = W_c x, p <- Pi(p - (R gamma + b)) with the Coulomb
//                 projection of project() (:100-119), and write the side impulses
//                 J_side^T (p_new - p_old) for the scatter.
//   pj_bodies   : per body: gather its incident side impulses through the CSR incidence list
//                 in a fixed order (no atomics), x += (W^T dp) / M, and every `check_every`
//                 sweeps the per-body part of the residual (:140-144), reduced per workgroup.
//   pj_residual : one workgroup reduces the partial sums, and sets the device "done" flag when
//                 res <= 1 (standard_loop, :371-377); every later launch returns immediately.
//
// Both sweep kernels are HBM-bound (about 0.5 flop per byte); there is no MFMA work.
#include <cmath>
#include <deque>

#include "cdx_device.hpp"

namespace cdx {

// ---------------------------------------------------------------------------------------------
// Setup (build_workspace, projected_jacobi.cpp:316-369)
// ---------------------------------------------------------------------------------------------

// R = alpha / D (rdiag_sparse, :57-69); b = R .* (c + e .* (W v + c)) (:337-347);
// p0 = warm start (:358-359); side impulses of W^T p0 for x0 (:361-363).
__global__ void k_pj_setup_contacts(int nc, const int* __restrict__ ida, const int* __restrict__ idb,
                                    const double* __restrict__ J, const double* __restrict__ mu,
                                    const double* __restrict__ rest, const double* __restrict__ vb,
                                    const double* __restrict__ minv, const double* __restrict__ v, double alpha,
                                    int warmstart, double* __restrict__ R, double* __restrict__ bias,
                                    double* __restrict__ p, double* __restrict__ imp) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int a = ida[c], b = idb[c];
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
    double mA[4] = {0, 0, 0, 0}, mB[4] = {0, 0, 0, 0};
    if (a >= 0)
        for (int m = 0; m < 4; ++m) mA[m] = minv[(size_t)a * kMinvStride + m];
    if (b >= 0)
        for (int m = 0; m < 4; ++m) mB[m] = minv[(size_t)b * kMinvStride + m];
    double pw[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) {
        if (k >= nr) {
            R[(size_t)k * nc + c] = 0;
            bias[(size_t)k * nc + c] = 0;
            p[(size_t)k * nc + c] = 0;
            continue;
        }
        // D summed side by side in column order (lower body index first), as the row-major
        // compressed W stores the row (rdiag_sparse iterates the row's entries in column order).
        double DA = 0, DB = 0;
        for (int m = 0; m < 3; ++m) DA += j.d[k][m] * j.d[k][m] * mA[0];
        for (int m = 0; m < 3; ++m) DA += j.aA[k][m] * j.aA[k][m] * mA[1 + m];
        for (int m = 0; m < 3; ++m) DB += j.d[k][m] * j.d[k][m] * mB[0];
        for (int m = 0; m < 3; ++m) DB += j.aB[k][m] * j.aB[k][m] * mB[1 + m];
        const double D = (a >= 0 && b >= 0 && b < a) ? DB + DA : DA + DB;
        const double Rk = D > 0 ? alpha / D : 0.0;
        const double ck = vb[(size_t)k * nc + c];
        const double vprev = row_dot(j.d[k], j.aA[k], j.aB[k], ua, ub) + ck;
        R[(size_t)k * nc + c] = Rk;
        bias[(size_t)k * nc + c] = Rk * (ck + rest[(size_t)k * nc + c] * vprev);
        pw[k] = warmstart ? p[(size_t)k * nc + c] : 0.0;
        p[(size_t)k * nc + c] = pw[k];
    }
    write_side_impulses(imp, c, a, b, nr, j, pw);
}

// x0 = S^-1 (M v + dt f_ext + dt f_gyro + W^T p0) (pj_assembler.cpp:150-153,173-182).
// Also used for the contact-free free solve (:460-467) with an empty incidence list.
__global__ void k_pj_setup_bodies(int nb, const int* __restrict__ dof, const double* __restrict__ minv,
                                  const double* __restrict__ v, const double* __restrict__ fext,
                                  const double* __restrict__ fgyro, const int* __restrict__ incptr,
                                  const int* __restrict__ incidx, const double* __restrict__ imp, double dt,
                                  int have_contacts, double* __restrict__ x) {
    const int bi = blockIdx.x * blockDim.x + threadIdx.x;
    if (bi >= nb) return;
    const int nd = dof[bi];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (have_contacts) {
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
    }
    const double* vv = v + (size_t)bi * kBodyStride;
    const double* fe = fext + (size_t)bi * kBodyStride;
    const double* fg = fgyro + (size_t)bi * kBodyStride;
    const double* mi = minv + (size_t)bi * kMinvStride;
    double out[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < nd; ++k) {
        const double mk = k < 3 ? mi[0] : mi[k - 2];
        const double M = mk > 0 ? 1.0 / mk : 0.0;
        double r = M * vv[k] + dt * fe[k];
        r += dt * fg[k];
        r += acc[k];
        out[k] = r / M;
    }
    store6(x + (size_t)bi * kBodyStride, out);
}

// ---------------------------------------------------------------------------------------------
// Sweep
// ---------------------------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) k_pj_contacts(int nc, int nblocks, const int* __restrict__ ida,
                                                        const int* __restrict__ idb, const double* __restrict__ J,
                                                        const double* __restrict__ R, const double* __restrict__ bias,
                                                        const double* __restrict__ mu, double* __restrict__ p,
                                                        const double* __restrict__ x, double* __restrict__ imp,
                                                        const int* __restrict__ ctrl) {
    if (ctrl[0]) return;  // converged: the rest of the enqueued sweeps are no-ops
    const int c = xcd_tile(nblocks) * kBlock + threadIdx.x;
    if (c >= nc) return;
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
        pn[k] = pold[k] - (R[(size_t)k * nc + c] * g + bias[(size_t)k * nc + c]);
    }
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
    write_side_impulses(imp, c, a, b, nr, j, dp);
}

__global__ void __launch_bounds__(kBlock) k_pj_bodies(int nb, const int* __restrict__ dof,
                                                      const double* __restrict__ minv, const int* __restrict__ incptr,
                                                      const int* __restrict__ incidx, const double* __restrict__ imp,
                                                      double* __restrict__ x, int check, double tol_rel, double tol_abs,
                                                      double* __restrict__ partial, const int* __restrict__ ctrl) {
    if (ctrl[0]) return;
    __shared__ double red[kBlock / 64];
    const int bi = blockIdx.x * kBlock + threadIdx.x;
    double e2 = 0;
    if (bi < nb) {
        double acc[6] = {0, 0, 0, 0, 0, 0};
        const int i0 = incptr[bi], i1 = incptr[bi + 1];
        for (int i = i0; i < i1; ++i) {
            const double2* q = reinterpret_cast<const double2*>(imp + (size_t)incidx[i] * 6);
            const double2 q0 = q[0], q1 = q[1], q2 = q[2];
            acc[0] += q0.x;
            acc[1] += q0.y;
            acc[2] += q1.x;
            acc[3] += q1.y;
            acc[4] += q2.x;
            acc[5] += q2.y;
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
    }
    if (check) {
        const double s = block_sum(e2, red);
        if (threadIdx.x == 0) partial[blockIdx.x] = s;
    }
}

// mode 0: PJ (break when res <= 1); mode 1: GS (break when res < 1, NaN/Inf -> diverged).
__global__ void __launch_bounds__(kBlock) k_residual_finalize(int nparts, const double* __restrict__ partial, int nv,
                                                              int it, int mode, int* __restrict__ ctrl,
                                                              double* __restrict__ res, int* __restrict__ hflag) {
    if (ctrl[0]) return;
    __shared__ double red[kBlock / 64];
    double s = 0;
    for (int i = threadIdx.x; i < nparts; i += kBlock) s += partial[i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) {
        const double r = 1.0 / sqrt((double)nv) * sqrt(s);
        res[0] = r;
        int done = 0, status = 0;
        if (mode == 0) {
            done = r <= 1.0;
        } else {
            if (isnan(r) || isinf(r)) {
                done = 1;
                status = CDX_DIVERGED;
            } else {
                done = r < 1.0;
            }
        }
        if (done) {
            ctrl[1] = it + 1;
            ctrl[2] = status;
            ctrl[0] = 1;
            hflag[0] = 1;
        }
    }
}

int reset_control(cdx_ctx* ctx) {
    CDX_HIP(ctx, hipMemsetAsync(ctx->d_ctrl, 0, 4 * sizeof(int), ctx->stream));
    const double nan = std::nan("");
    CDX_HIP(ctx, hipMemcpyAsync(ctx->d_res, &nan, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    __atomic_store_n(&ctx->h_flag[0], 0, __ATOMIC_RELEASE);
    return CDX_OK;
}

int launch_residual_finalize(cdx_ctx* ctx, int nparts, int it, int mode) {
    CDX_LAUNCH(ctx, mode == 0 ? "pj_residual" : "gs_residual", k_residual_finalize, 1, kBlock, nparts, ctx->d_partial, ctx->nv,
               it, mode, ctx->d_ctrl, ctx->d_res, ctx->h_flag_dev);
    return hip_check(ctx, hipGetLastError(), "residual finalize launch");
}

int read_residual(cdx_ctx* ctx, double* r) {
    CDX_HIP(ctx, hipMemcpy(r, ctx->d_res, sizeof(double), hipMemcpyDeviceToHost));
    return CDX_OK;
}

int check_solve_args(cdx_ctx* ctx, const cdx_solver_params* prm) {
    if (!prm) return set_err(ctx, CDX_BAD_ARG, "null params");
    if (ctx->nb <= 0 && ctx->nc > 0) return set_err(ctx, CDX_NO_STATE, "contacts without bodies");
    if (prm->max_iter < 1) return set_err(ctx, CDX_BAD_ARG, "max_iter must be >= 1");
    if (prm->check_every < 1) return set_err(ctx, CDX_BAD_ARG, "check_every must be >= 1");
    if (!std::isfinite(prm->alpha) || !std::isfinite(prm->tol_abs) || !std::isfinite(prm->tol_rel))
        return set_err(ctx, CDX_BAD_ARG, "non-finite solver parameter");
    if (ctx->singular_mass)
        return set_err(ctx, CDX_NOT_FACTORIZED, "S matrix is not factorized: a dynamic dof has zero mass");
    return CDX_OK;
}

}  // namespace cdx

using namespace cdx;

extern "C" int cdx_pj_solve(cdx_ctx* ctx, double dt, double theta, const cdx_solver_params* prm, cdx_solve_info* info) {
    (void)theta;  // theta only enters bilateral rows and S of spring/damper scenes
    if (!ctx) return CDX_BAD_ARG;
    hipSetDevice(ctx->device);
    CDX_TRY(check_solve_args(ctx, prm));
    if (prm->accel != CDX_ACCEL_NONE) return set_err(ctx, CDX_BAD_ARG, "cdx_pj_solve: acceleration mode not supported on the device");
    if (info) {
        info->iters = 0;
        info->residual = std::nan("");
        info->nrows = ctx->nrows;
        info->ncolors = 0;
    }
    ctx->solved = false;
    const int nb = ctx->nb, nc = ctx->nc;
    if (nb == 0) {
        ctx->solved = true;
        return CDX_OK;
    }
    CDX_TRY(build_incidence(ctx));
    CDX_TRY(reset_control(ctx));
    const bool have_contacts = ctx->nrows > 0;
    if (have_contacts)
        CDX_LAUNCH(ctx, "pj_setup_contacts", k_pj_setup_contacts, grid_for(nc), kBlock, nc, ctx->d_ida, ctx->d_idb, ctx->d_J,
                   ctx->d_mu, ctx->d_rest, ctx->d_vb, ctx->d_minv, ctx->d_v, prm->alpha, prm->warmstart, ctx->d_R,
                   ctx->d_bias, ctx->d_p, ctx->d_imp);
    CDX_LAUNCH(ctx, "pj_setup_bodies", k_pj_setup_bodies, grid_for(nb), kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_v,
               ctx->d_fext, ctx->d_fgyro, ctx->d_incptr, ctx->d_incidx, ctx->d_imp, dt, have_contacts ? 1 : 0, ctx->d_x);
    CDX_HIP(ctx, hipGetLastError());
    if (!have_contacts) {  // no contacts: free solution only (projected_jacobi.cpp:460-467)
        CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ctx->solved = true;
        return CDX_OK;
    }
    const int gb = grid_for(nb), gc = grid_for(nc);
    CDX_TRY(ensure_alloc(ctx, &ctx->d_partial, (size_t)gb, &ctx->partial_cap, "residual partials"));
    const int check_every = prm->check_every;
    auto sweep = [&](int it) -> int {
        const int check = (it % check_every) == 0;
        CDX_LAUNCH(ctx, "pj_contacts", k_pj_contacts, gc, kBlock, nc, gc, ctx->d_ida, ctx->d_idb, ctx->d_J, ctx->d_R,
                   ctx->d_bias, ctx->d_mu, ctx->d_p, ctx->d_x, ctx->d_imp, ctx->d_ctrl);
        CDX_LAUNCH(ctx, "pj_bodies", k_pj_bodies, gb, kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_incptr, ctx->d_incidx,
                   ctx->d_imp, ctx->d_x, check, prm->tol_rel, prm->tol_abs, ctx->d_partial, ctx->d_ctrl);
        if (check) CDX_TRY(launch_residual_finalize(ctx, gb, it, 0));
        return hip_check(ctx, hipGetLastError(), "pj sweep launch");
    };
    int iters = 0;
    const int st = run_sweeps(ctx, prm->max_iter, check_every, 4, sweep, &iters);
    if (st) return set_err(ctx, st, "cdx_pj_solve failed");
    double r = 0;
    CDX_TRY(read_residual(ctx, &r));
    if (info) {
        info->iters = iters;
        info->residual = r;
    }
    ctx->solved = true;
    return CDX_OK;
}

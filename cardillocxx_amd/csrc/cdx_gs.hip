// cdx_gs.hip -- FP64 graph-coloured projected Gauss-Seidel on gfx950.
//
// The block update is the one of ProjectedGaussSeidelSolver's pgs_sweep
// (src/physics/solver/projected_gauss_seidel.cpp:144-169; lambda <= 0 internally, warm start and
// stored impulse with the normal sign inverted, :95-101,265): residual r = rhs - W_j u_corr,
// dl = (D_j^-1 r) * relaxation * alpha, lambda += dl, projection (:11-26) with dl corrected,
// u_corr += M^-1 W_j^T dl.  The blocks are visited colour by colour, as CondensedSolver's
// coloredSweep does (condensed_solver.cpp:220-251): blocks sharing a body never share a colour
// (shared-body adjacency, :195-215; greedy Welsh-Powell colouring, graph_coloring.cpp:8-41), so a
// colour is one kernel launch with one lane per block and no write conflicts.  Within a colour
// the visiting order does not change the result.  The residual is evaluated after every sweep
// (:124-141); NaN/Inf ends the solve with CDX_DIVERGED.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "cdx_device.hpp"

namespace cdx {

// Greedy Welsh-Powell colouring of the shared-body block graph.  Blocks are the contacts in the
// solver's row order (frictionless first, then frictional; dynamics_assembler.cpp:236-244).
// Ties in degree are broken by block index (the reference's std::sort leaves them unspecified).
static void color_blocks(int nb, const std::vector<int>& blk_a, const std::vector<int>& blk_b, std::vector<int>& color_of,
                         int& ncolors) {
    const int n = (int)blk_a.size();
    std::vector<int> cnt(nb + 1, 0);
    for (int k = 0; k < n; ++k) {
        if (blk_a[k] >= 0) cnt[blk_a[k] + 1]++;
        if (blk_b[k] >= 0) cnt[blk_b[k] + 1]++;
    }
    for (int b = 0; b < nb; ++b) cnt[b + 1] += cnt[b];
    std::vector<int> inc(cnt[nb] + 1), fill(cnt.begin(), cnt.end() - 1);
    for (int k = 0; k < n; ++k) {
        if (blk_a[k] >= 0) inc[fill[blk_a[k]]++] = k;
        if (blk_b[k] >= 0) inc[fill[blk_b[k]]++] = k;
    }
    std::vector<std::vector<int>> adj(n);
    for (int b = 0; b < nb; ++b)
        for (int i = cnt[b]; i < cnt[b + 1]; ++i)
            for (int j = cnt[b]; j < cnt[b + 1]; ++j)
                if (i != j) adj[inc[i]].push_back(inc[j]);
    for (auto& v : adj) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return adj[x].size() > adj[y].size(); });
    color_of.assign(n, -1);
    std::vector<int> stamp;
    ncolors = 0;
    for (int k = 0; k < n; ++k) {
        const int node = order[k];
        for (int nbr : adj[node]) {
            const int c = color_of[nbr];
            if (c < 0) continue;
            if (c >= (int)stamp.size()) stamp.resize(c + 1, -1);
            stamp[c] = k;
        }
        int chosen = 0;
        while (chosen < (int)stamp.size() && stamp[chosen] == k) ++chosen;
        color_of[node] = chosen;
        ncolors = std::max(ncolors, chosen + 1);
    }
}

static int build_coloring(cdx_ctx* ctx) {
    if (ctx->coloring_valid) return CDX_OK;
    std::vector<int> cidx, ba, bb;
    for (int pass = 0; pass < 2; ++pass)
        for (int c = 0; c < ctx->nc; ++c) {
            const bool fr = ctx->h_mu[c] > 0;
            if ((pass == 1) != fr) continue;
            if (ctx->h_a[c] < 0 && ctx->h_b[c] < 0) continue;
            cidx.push_back(c);
            ba.push_back(ctx->h_a[c]);
            bb.push_back(ctx->h_b[c]);
        }
    std::vector<int> color_of;
    int ncol = 0;
    color_blocks(ctx->nb, ba, bb, color_of, ncol);
    const int n = (int)cidx.size();
    std::vector<int> ptr(ncol + 1, 0), sorted(n + 1);
    for (int k = 0; k < n; ++k) ptr[color_of[k] + 1]++;
    for (int c = 0; c < ncol; ++c) ptr[c + 1] += ptr[c];
    std::vector<int> f(ptr.begin(), ptr.end() - 1);
    for (int k = 0; k < n; ++k) sorted[f[color_of[k]]++] = cidx[k];
    CDX_TRY(ensure_alloc(ctx, &ctx->d_color_blocks, (size_t)n, &ctx->color_cap, "colour classes"));
    if (n) CDX_HIP(ctx, hipMemcpy(ctx->d_color_blocks, sorted.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    ctx->color_ptr = ptr;
    ctx->ncolors = ncol;
    ctx->coloring_valid = true;
    return CDX_OK;
}

// u_free = v + M^-1 (dt f_ext) + M^-1 (dt f_gyro) (pgs_assembler.cpp:75-82).
__global__ void k_gs_ufree(int nb, const int* __restrict__ dof, const double* __restrict__ minv,
                           const double* __restrict__ v, const double* __restrict__ fext,
                           const double* __restrict__ fgyro, double dt, double* __restrict__ uf) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const double* vv = v + (size_t)b * kBodyStride;
    const double* fe = fext + (size_t)b * kBodyStride;
    const double* fg = fgyro + (size_t)b * kBodyStride;
    const double* mi = minv + (size_t)b * kMinvStride;
    double o[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < dof[b]; ++k) {
        const double m = k < 3 ? mi[0] : mi[k - 2];
        double u = vv[k] + m * (dt * fe[k]);
        u += m * (dt * fg[k]);
        o[k] = u;
    }
    store6(uf + (size_t)b * kBodyStride, o);
}

// D^-1 diagonal (PgsAssembler::DinvDiag, pgs_assembler.cpp:220-288: frictionless guarded,
// frictional not), rhs = W u_free + c (+ e .* (W v + c) when any e != 0, :61-70), warm start with
// inverted normal, and the side impulses of W^T lambda0 for u_corr = M^-1 W^T lambda0 (:108).
__global__ void k_gs_setup_contacts(int nc, const int* __restrict__ ida, const int* __restrict__ idb,
                                    const double* __restrict__ J, const double* __restrict__ mu,
                                    const double* __restrict__ rest, const double* __restrict__ vb,
                                    const double* __restrict__ minv, const double* __restrict__ v,
                                    const double* __restrict__ uf, int any_rest, int warmstart,
                                    double* __restrict__ Dinv, double* __restrict__ rhs, double* __restrict__ lam,
                                    double* __restrict__ imp) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int a = ida[c], b = idb[c];
    if (a < 0 && b < 0) {
        for (int k = 0; k < 3; ++k) {
            Dinv[(size_t)k * nc + c] = 0;
            rhs[(size_t)k * nc + c] = 0;
            lam[(size_t)k * nc + c] = 0;
        }
        return;
    }
    const int nr = mu[c] > 0 ? 3 : 1;
    ContactJ j;
    load_J(J, nc, c, nr, j);
    double va[6], vbb[6], fa[6], fb[6];
    load_body(v, a, va);
    load_body(v, b, vbb);
    load_body(uf, a, fa);
    load_body(uf, b, fb);
    double mA[4] = {0, 0, 0, 0}, mB[4] = {0, 0, 0, 0};
    if (a >= 0)
        for (int m = 0; m < 4; ++m) mA[m] = minv[(size_t)a * kMinvStride + m];
    if (b >= 0)
        for (int m = 0; m < 4; ++m) mB[m] = minv[(size_t)b * kMinvStride + m];
    double l0[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) {
        if (k >= nr) {
            Dinv[(size_t)k * nc + c] = 0;
            rhs[(size_t)k * nc + c] = 0;
            lam[(size_t)k * nc + c] = 0;
            continue;
        }
        double DA = 0, DB = 0;
        for (int m = 0; m < 3; ++m) DA += j.d[k][m] * j.d[k][m] * mA[0];
        for (int m = 0; m < 3; ++m) DA += j.aA[k][m] * j.aA[k][m] * mA[1 + m];
        for (int m = 0; m < 3; ++m) DB += j.d[k][m] * j.d[k][m] * mB[0];
        for (int m = 0; m < 3; ++m) DB += j.aB[k][m] * j.aB[k][m] * mB[1 + m];
        const double D = (a >= 0 && b >= 0 && b < a) ? DB + DA : DA + DB;
        Dinv[(size_t)k * nc + c] = (nr == 1) ? (D > 0 ? 1.0 / D : 0.0) : 1.0 / D;
        const double ck = vb[(size_t)k * nc + c];
        double bi = ck;
        if (any_rest) bi += rest[(size_t)k * nc + c] * (row_dot(j.d[k], j.aA[k], j.aB[k], va, vbb) + ck);
        rhs[(size_t)k * nc + c] = row_dot(j.d[k], j.aA[k], j.aB[k], fa, fb) + bi;
        if (warmstart) {
            const double w = lam[(size_t)k * nc + c];  // canonical warm start
            l0[k] = k == 0 ? -w : w;
        }
        lam[(size_t)k * nc + c] = l0[k];
    }
    write_side_impulses(imp, c, a, b, nr, j, l0);
}

// u_corr = M^-1 (W^T lambda0); also the residual baseline copy.
__global__ void k_gs_init_ucorr(int nb, const int* __restrict__ dof, const double* __restrict__ minv,
                                const int* __restrict__ incptr, const int* __restrict__ incidx,
                                const double* __restrict__ imp, double* __restrict__ uc, double* __restrict__ up) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int i = incptr[b]; i < incptr[b + 1]; ++i) {
        const double2* q = reinterpret_cast<const double2*>(imp + (size_t)incidx[i] * 6);
        const double2 q0 = q[0], q1 = q[1], q2 = q[2];
        acc[0] += q0.x;
        acc[1] += q0.y;
        acc[2] += q1.x;
        acc[3] += q1.y;
        acc[4] += q2.x;
        acc[5] += q2.y;
    }
    const double* mi = minv + (size_t)b * kMinvStride;
    double o[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < dof[b]; ++k) o[k] = (k < 3 ? mi[0] : mi[k - 2]) * acc[k];
    store6(uc + (size_t)b * kBodyStride, o);
    store6(up + (size_t)b * kBodyStride, o);
}

// PGS block update of contact c (projected_gauss_seidel.cpp:144-169 with project(), :11-26),
// writing lambda and the u_corr of both bodies directly.
__device__ inline void gs_block_update(int c, const int* __restrict__ ida, const int* __restrict__ idb,
                                       const double* __restrict__ J, const double* __restrict__ Dinv,
                                       const double* __restrict__ rhs, const double* __restrict__ mu,
                                       const double* __restrict__ minv, int nc, double relax, double alpha,
                                       double* __restrict__ lam, double* __restrict__ uc) {
    const int a = ida[c], b = idb[c];
    const double muc = mu[c];
    const int nr = muc > 0 ? 3 : 1;
    ContactJ j;
    load_J(J, nc, c, nr, j);
    double ua[6], ub[6];
    load_body(uc, a, ua);
    load_body(uc, b, ub);
    double l[3] = {0, 0, 0}, dl[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= nr) break;
        const double r = rhs[(size_t)k * nc + c] - row_dot(j.d[k], j.aA[k], j.aB[k], ua, ub);
        dl[k] = (Dinv[(size_t)k * nc + c] * r) * relax;
        dl[k] *= alpha;
        l[k] = lam[(size_t)k * nc + c] + dl[k];
    }
    // project (projected_gauss_seidel.cpp:11-26)
    const double ln = l[0];
    l[0] = fmin(ln, 0.0);
    dl[0] += l[0] - ln;
    if (nr == 3) {
        const double t1 = l[1], t2 = l[2];
        const double tn = sqrt(t1 * t1 + t2 * t2);
        if (tn > 0) {
            const double s = fmin(1.0, -(muc * l[0]) / tn);
            const double p1 = s * t1, p2 = s * t2;
            dl[1] += p1 - t1;
            dl[2] += p2 - t2;
            l[1] = p1;
            l[2] = p2;
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= nr) break;
        lam[(size_t)k * nc + c] = l[k];
    }
    double lin[3] = {0, 0, 0}, angA[3] = {0, 0, 0}, angB[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= nr) break;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            lin[m] += j.d[k][m] * dl[k];
            angA[m] += j.aA[k][m] * dl[k];
            angB[m] += j.aB[k][m] * dl[k];
        }
    }
    if (a >= 0) {
        const double4 m4 = *reinterpret_cast<const double4*>(minv + (size_t)a * kMinvStride);
        const double o[6] = {ua[0] + m4.x * -lin[0], ua[1] + m4.x * -lin[1], ua[2] + m4.x * -lin[2],
                             ua[3] + m4.y * -angA[0], ua[4] + m4.z * -angA[1], ua[5] + m4.w * -angA[2]};
        store6(uc + (size_t)a * kBodyStride, o);
    }
    if (b >= 0) {
        const double4 m4 = *reinterpret_cast<const double4*>(minv + (size_t)b * kMinvStride);
        const double o[6] = {ub[0] + m4.x * lin[0], ub[1] + m4.x * lin[1], ub[2] + m4.x * lin[2],
                             ub[3] + m4.y * angB[0], ub[4] + m4.z * angB[1], ub[5] + m4.w * angB[2]};
        store6(uc + (size_t)b * kBodyStride, o);
    }
}

// One colour class: each lane runs the PGS block update of one contact.  Blocks of a colour
// share no body, so the read-modify-write of u_corr is race-free.
__global__ void __launch_bounds__(kBlock) k_gs_color(int n, const int* __restrict__ blocks, const int* __restrict__ ida,
                                                     const int* __restrict__ idb, const double* __restrict__ J,
                                                     const double* __restrict__ Dinv, const double* __restrict__ rhs,
                                                     const double* __restrict__ mu, const double* __restrict__ minv,
                                                     int nc, double relax, double alpha, double* __restrict__ lam,
                                                     double* __restrict__ uc, const int* __restrict__ ctrl) {
    if (ctrl[0]) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    gs_block_update(blocks[i], ida, idb, J, Dinv, rhs, mu, minv, nc, relax, alpha, lam, uc);
}

// Per-body part of the residual on v = u_free - u_corr (projected_gauss_seidel.cpp:124-131), and
// the copy of the iterate for the next sweep.
__device__ inline double gs_body_residual(int b, const int* __restrict__ dof, const double* __restrict__ uf,
                                          const double* __restrict__ uc, double* __restrict__ up, double tol_rel,
                                          double tol_abs) {
    double f[6], u1[6], u0[6], e2 = 0;
    load_body(uf, b, f);
    load_body(uc, b, u1);
    load_body(up, b, u0);
    for (int k = 0; k < dof[b]; ++k) {
        const double d = u0[k] - u1[k];
        const double v0 = f[k] - u1[k], v1 = f[k] - u0[k];
        const double q = fmax(fabs(v0), fabs(v1)) * tol_rel + tol_abs;
        const double e = d / q;
        e2 += e * e;
    }
    store6(up + (size_t)b * kBodyStride, u1);
    return e2;
}

__global__ void __launch_bounds__(kBlock) k_gs_residual(int nb, const int* __restrict__ dof, const double* __restrict__ uf,
                                                        const double* __restrict__ uc, double* __restrict__ up,
                                                        double tol_rel, double tol_abs, double* __restrict__ partial,
                                                        const int* __restrict__ ctrl) {
    if (ctrl[0]) return;
    __shared__ double red[kBlock / 64];
    const int b = blockIdx.x * kBlock + threadIdx.x;
    const double e2 = b < nb ? gs_body_residual(b, dof, uf, uc, up, tol_rel, tol_abs) : 0.0;
    const double s = block_sum(e2, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// v = u_free - u_corr (:267) and canonical impulses (normal sign flipped back, :265).
__global__ void k_gs_finish_bodies(int nb, const double* __restrict__ uf, double* __restrict__ x) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double f[6], u[6], o[6];
    load_body(uf, b, f);
    load_body(x, b, u);
    for (int k = 0; k < 6; ++k) o[k] = f[k] - u[k];
    store6(x + (size_t)b * kBodyStride, o);
}

__global__ void k_gs_finish_contacts(int nc, double* __restrict__ lam) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    lam[c] = -lam[c];
}

}  // namespace cdx

using namespace cdx;

constexpr int kGsChunk = 8;

extern "C" int cdx_gs_colored_solve(cdx_ctx* ctx, double dt, double theta, const cdx_solver_params* prm,
                                    cdx_solve_info* info) {
    (void)theta;
    if (!ctx) return CDX_BAD_ARG;
    hipSetDevice(ctx->device);
    CDX_TRY(check_solve_args(ctx, prm));
    if (ctx->njrows > 0) return set_err(ctx, CDX_BAD_ARG, "cdx_gs_colored_solve: joint rows are solved by cdx_pj_solve only");
    if (prm->accel != CDX_ACCEL_NONE) return set_err(ctx, CDX_BAD_ARG, "cdx_gs_colored_solve: acceleration mode not supported");
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
    ctx->pj_last = false;
    ctx->pjj_last = false;
        return CDX_OK;
    }
    CDX_TRY(build_incidence(ctx));
    CDX_TRY(reset_control(ctx));
    const int gb = grid_for(nb);
    CDX_LAUNCH(ctx, "gs_ufree", k_gs_ufree, gb, kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_v, ctx->d_fext, ctx->d_fgyro, dt,
               ctx->d_ufree);
    if (ctx->nrows == 0) {  // numLambda == 0: return u_free (projected_gauss_seidel.cpp:73)
        CDX_HIP(ctx, hipMemcpyAsync(ctx->d_x, ctx->d_ufree, sizeof(double) * (size_t)nb * kBodyStride, hipMemcpyDeviceToDevice,
                                    ctx->stream));
        CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ctx->solved = true;
    ctx->pj_last = false;
    ctx->pjj_last = false;
        return CDX_OK;
    }
    CDX_TRY(build_coloring(ctx));
    CDX_LAUNCH(ctx, "gs_setup_contacts", k_gs_setup_contacts, grid_for(nc), kBlock, nc, ctx->d_ida, ctx->d_idb, ctx->d_J,
               ctx->d_mu, ctx->d_rest, ctx->d_vb, ctx->d_minv, ctx->d_v, ctx->d_ufree, ctx->any_rest ? 1 : 0, prm->warmstart,
               ctx->d_R, ctx->d_bias, ctx->d_p, ctx->d_imp);
    CDX_LAUNCH(ctx, "gs_init", k_gs_init_ucorr, gb, kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_incptr, ctx->d_incidx,
               ctx->d_imp, ctx->d_x, ctx->d_uprev);
    CDX_HIP(ctx, hipGetLastError());
    CDX_TRY(ensure_alloc(ctx, &ctx->d_partial, (size_t)gb, &ctx->partial_cap, "residual partials"));
    auto sweep = [&](int it, int) -> int {
        for (int c = 0; c < ctx->ncolors; ++c) {
            const int n = ctx->color_ptr[c + 1] - ctx->color_ptr[c];
            if (n == 0) continue;
            CDX_LAUNCH(ctx, "gs_color", k_gs_color, grid_for(n), kBlock, n, ctx->d_color_blocks + ctx->color_ptr[c], ctx->d_ida,
                       ctx->d_idb, ctx->d_J, ctx->d_R, ctx->d_bias, ctx->d_mu, ctx->d_minv, nc, prm->relaxation, prm->alpha,
                       ctx->d_p, ctx->d_x, ctx->d_ctrl);
        }
        CDX_LAUNCH(ctx, "gs_body_residual", k_gs_residual, gb, kBlock, nb, ctx->d_dof, ctx->d_ufree, ctx->d_x, ctx->d_uprev,
                   prm->tol_rel, prm->tol_abs, ctx->d_partial, ctx->d_ctrl);
        CDX_TRY(launch_residual_finalize(ctx, gb, it, 1));
        return hip_check(ctx, hipGetLastError(), "gs sweep launch");
    };
    int iters = 0;
    // residual every sweep (projected_gauss_seidel.cpp:130-140); one launch per colour, chunks of
    // 8 sweeps replayed as graphs.  (A single cooperative launch with grid barriers between the
    // colours measured twice as slow per sweep: each device-scope barrier writes back and
    // invalidates the L2 of every XCD, so every colour re-fetched its data.)
    const int st = run_sweeps_graphed(ctx, prm->max_iter, kGsChunk, true, sweep, &iters);
    CDX_LAUNCH(ctx, "gs_finish", k_gs_finish_bodies, gb, kBlock, nb, ctx->d_ufree, ctx->d_x);
    CDX_LAUNCH(ctx, "gs_finish", k_gs_finish_contacts, grid_for(nc), kBlock, nc, ctx->d_p);
    CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double r = 0;
    CDX_TRY(read_residual(ctx, &r));
    if (info) {
        info->iters = iters;
        info->residual = r;
        info->ncolors = ctx->ncolors;
    }
    ctx->solved = true;
    ctx->pj_last = false;
    ctx->pjj_last = false;
    if (st) return set_err(ctx, st, "Projected Gauss-Seidel diverged");
    return CDX_OK;
}

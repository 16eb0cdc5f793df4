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

// u_free = v + M^-1 (dt f_ext) + M^-1 (dt f_gyro) (pgs_assembler.cpp:75-82); fgyro = nullptr with
// implicit gyroscopy, whose only effect on the Gauss-Seidel solve is to drop that term (:33-36).
__global__ void k_gs_ufree(int nb, const int* __restrict__ dof, const double* __restrict__ minv,
                           const double* __restrict__ v, const double* __restrict__ fext,
                           const double* __restrict__ fgyro, double dt, double* __restrict__ uf) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const double* vv = v + (size_t)b * kBodyStride;
    const double* fe = fext + (size_t)b * kBodyStride;
    const double* fg = fgyro ? fgyro + (size_t)b * kBodyStride : nullptr;
    const double* mi = minv + (size_t)b * kMinvStride;
    double o[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < dof[b]; ++k) {
        const double m = k < 3 ? mi[0] : mi[k - 2];
        double u = vv[k] + m * (dt * fe[k]);
        if (fg) u += m * (dt * fg[k]);
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

// One colour class of the PGS block update (projected_gauss_seidel.cpp:144-169 with project(),
// :11-26), 4 lanes per contact as the projected-Jacobi contact pass: lane k < 3 loads row k's
// 9 Jacobian entries and computes that row's correction; the group's values meet through
// DPP exchanges for the projection, and the rows' side impulses are summed by quad_sum,
// (r0 + r1) + (r2 + 0), which is the sequential sum over the rows (lane 3 holds exact zeros).  Lane 0
// writes u_corr of body A, lane 1 that of body B.  Blocks of a colour share no body, so the
// read-modify-write of u_corr is race-free.  The contact data is in colour order (k_gs_pack):
// the class is the contiguous range [off, off + n) of arrays with row stride ns, so every load
// is coalesced; four lanes per contact keep four times the waves in flight on the small classes,
// whose time is load latency.
__global__ void __launch_bounds__(kBlock) k_gs_color4(int n, int off, int ns, const int* __restrict__ ids,
                                                      const double* __restrict__ pack, const double* __restrict__ minv,
                                                      double relax, double alpha, double* __restrict__ lam,
                                                      double* __restrict__ uc, const int* __restrict__ ctrl) {
    if (ctrl[0]) return;
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int i = t / kLanesPerContact, k = t % kLanesPerContact;
    const bool live = i < n;
    const int c = off + i;
    const double* J = pack;
    const double* Dinv = pack + 27 * (size_t)ns;
    const double* rhs = pack + 30 * (size_t)ns;
    const double* mu = pack + 36 * (size_t)ns;
    int a = -1, b = -1;
    double muc = 0;
    if (live) {
        a = ids[c];
        b = ids[ns + c];
        muc = mu[c];
    }
    const int nr = muc > 0 ? 3 : 1;
    const bool row = live && k < nr;
    double d[3] = {0, 0, 0}, aA[3] = {0, 0, 0}, aB[3] = {0, 0, 0};
    double ua[6] = {0, 0, 0, 0, 0, 0}, ub[6] = {0, 0, 0, 0, 0, 0};
    double l = 0, dl = 0;
    if (live) {
        load_body(uc, a, ua);
        load_body(uc, b, ub);
    }
    if (row) {
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            d[m] = J[(size_t)(3 * k + m) * ns + c];
            aA[m] = J[(size_t)(9 + 3 * k + m) * ns + c];
            aB[m] = J[(size_t)(18 + 3 * k + m) * ns + c];
        }
        const double r = rhs[(size_t)k * ns + c] - row_dot(d, aA, aB, ua, ub);
        dl = (Dinv[(size_t)k * ns + c] * r) * relax;
        dl *= alpha;
        l = lam[(size_t)k * ns + c] + dl;
    }
    // project (projected_gauss_seidel.cpp:11-26): lambda_n <= 0, tangential part scaled into the cone
    const double l0 = quad_bcast<0>(l);
    const double l1 = quad_bcast<1>(l);
    const double l2 = quad_bcast<2>(l);
    const double ln = fmin(l0, 0.0);
    if (k == 0) {
        dl += ln - l;
        l = ln;
    } else if (nr == 3 && k < 3) {
        const double tn = sqrt(l1 * l1 + l2 * l2);
        if (tn > 0) {
            const double sc = fmin(1.0, -(muc * ln) / tn);
            const double pv = sc * l;
            dl += pv - l;
            l = pv;
        }
    }
    if (row) lam[(size_t)k * ns + c] = l;
    double v9[9];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        v9[m] = d[m] * dl;
        v9[3 + m] = aA[m] * dl;
        v9[6 + m] = aB[m] * dl;
    }
#pragma unroll
    for (int m = 0; m < 9; ++m) v9[m] = quad_sum(v9[m]);
    if (live && k == 0 && a >= 0) {
        const double4 m4 = *reinterpret_cast<const double4*>(minv + (size_t)a * kMinvStride);
        const double o[6] = {ua[0] + m4.x * -v9[0], ua[1] + m4.x * -v9[1], ua[2] + m4.x * -v9[2],
                             ua[3] + m4.y * -v9[3], ua[4] + m4.z * -v9[4], ua[5] + m4.w * -v9[5]};
        store6(uc + (size_t)a * kBodyStride, o);
    }
    if (live && k == 1 && b >= 0) {
        const double4 m4 = *reinterpret_cast<const double4*>(minv + (size_t)b * kMinvStride);
        const double o[6] = {ub[0] + m4.x * v9[0], ub[1] + m4.x * v9[1], ub[2] + m4.x * v9[2],
                             ub[3] + m4.y * v9[6], ub[4] + m4.z * v9[7], ub[5] + m4.w * v9[8]};
        store6(uc + (size_t)b * kBodyStride, o);
    }
}

// Colour-order copy of the sweep's contact data: position i of the colour-sorted list holds
// contact blocks[i] (layout of cdx_ctx::d_gs_pack).
__global__ void k_gs_pack(int ns, int nc, const int* __restrict__ blocks, const int* __restrict__ ida,
                          const int* __restrict__ idb, const double* __restrict__ J, const double* __restrict__ Dinv,
                          const double* __restrict__ rhs, const double* __restrict__ lam, const double* __restrict__ mu,
                          double* __restrict__ pack, int* __restrict__ ids) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    const int c = blocks[i];
    for (int r = 0; r < 27; ++r) pack[(size_t)r * ns + i] = J[(size_t)r * nc + c];
    for (int k = 0; k < 3; ++k) {
        pack[(size_t)(27 + k) * ns + i] = Dinv[(size_t)k * nc + c];
        pack[(size_t)(30 + k) * ns + i] = rhs[(size_t)k * nc + c];
        pack[(size_t)(33 + k) * ns + i] = lam[(size_t)k * nc + c];
    }
    pack[(size_t)36 * ns + i] = mu[c];
    ids[i] = ida[c];
    ids[ns + i] = idb[c];
}

// lambda back from colour order to contact order.
__global__ void k_gs_unpack(int ns, int nc, const int* __restrict__ blocks, const double* __restrict__ pack,
                            double* __restrict__ lam) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    const int c = blocks[i];
    for (int k = 0; k < 3; ++k) lam[(size_t)k * nc + c] = pack[(size_t)(33 + k) * ns + i];
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

// ---------------------------------------------------------------------------------------------
// Nesterov-accelerated coloured Gauss-Seidel (the pj_nesterov branch of
// ProjectedGaussSeidelSolver::solve, projected_gauss_seidel.cpp:196-262).
//
// Buffers: X = u_corr (the sweep works in place), XK = u_k, UY = the sweep input u_y (kept for the
// residual and the restart test), lambda in colour order (the sweep works in place), LK = lambda_k.
// Each iteration: extrapolate (lambda_y, u_y) = k1 + beta (k1 - k) with beta from the previous
// iteration (0 on a restart, with momentum off, or before the first sweep), rotate k <- k1, run
// the colour launches, then reduce the residual between u_y and u_k1 and the restart dot
// (u_y - u_k1) . (u_k1 - u_k) and decide beta (the same rules and nest[] state as k_pjn_finalize).
// ---------------------------------------------------------------------------------------------

// Threads [0, nb): bodies; [nb, nb + ns): contacts in colour order.
__global__ void k_gsn_extrapolate(int nb, int ns, double* __restrict__ x, double* __restrict__ xk,
                                  double* __restrict__ uy, double* __restrict__ lam, double* __restrict__ lamk,
                                  const double* __restrict__ nest, const int* __restrict__ ctrl) {
    if (ctrl[0]) return;
    const double beta = nest[0];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nb) {
        double u[6], uk[6], y[6];
        load_body(x, t, u);
        load_body(xk, t, uk);
#pragma unroll
        for (int k = 0; k < 6; ++k) y[k] = beta == 0.0 ? u[k] : u[k] + beta * (u[k] - uk[k]);
        store6(xk + (size_t)t * kBodyStride, u);
        store6(x + (size_t)t * kBodyStride, y);
        store6(uy + (size_t)t * kBodyStride, y);
    } else if (t < nb + ns) {
        const int i = t - nb;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double l = lam[(size_t)k * ns + i], lk = lamk[(size_t)k * ns + i];
            lamk[(size_t)k * ns + i] = l;
            lam[(size_t)k * ns + i] = beta == 0.0 ? l : l + beta * (l - lk);
        }
    }
}

// Residual partials between u_y and u_k1 (projected_gauss_seidel.cpp:124-131) and the restart dot
// (u_y - u_k1) . (u_k1 - u_k) over the body's dofs.
__global__ void __launch_bounds__(kBlock) k_gsn_residual(int nb, const int* __restrict__ dof, const double* __restrict__ uf,
                                                         const double* __restrict__ uk1, const double* __restrict__ uy,
                                                         const double* __restrict__ uk, double tol_rel, double tol_abs,
                                                         double* __restrict__ partial, const int* __restrict__ ctrl) {
    if (ctrl[0]) return;
    __shared__ double red[kBlock / 64];
    const int b = blockIdx.x * kBlock + threadIdx.x;
    double e2 = 0, dt = 0;
    if (b < nb) {
        double f[6], u1[6], u0[6], k0[6];
        load_body(uf, b, f);
        load_body(uk1, b, u1);
        load_body(uy, b, u0);
        load_body(uk, b, k0);
        for (int k = 0; k < dof[b]; ++k) {
            const double d = u0[k] - u1[k];
            const double v0 = f[k] - u1[k], v1 = f[k] - u0[k];
            const double q = fmax(fabs(v0), fabs(v1)) * tol_rel + tol_abs;
            const double e = d / q;
            e2 += e * e;
            dt += (u0[k] - u1[k]) * (u1[k] - k0[k]);
        }
    }
    const double s = block_sum(e2, red);
    const double d = block_sum(dt, red);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s;
        partial[gridDim.x + blockIdx.x] = d;
    }
}

// nest[]: 0 beta, 1 thk, 2 err_prev, 3 restarts, 4 momentum_disabled.  A non-finite residual ends
// the solve with CDX_DIVERGED (check_residual throws before the restart rules, :133-138); the loop
// stops when res <= 1 (:224).  it < 0: the sweep index comes from the chunk counter ctrl[3].
__global__ void __launch_bounds__(kBlock) k_gsn_finalize(int nparts, const double* __restrict__ partial, int nv, int it,
                                                         double beta_threshold, int restart_limit, int* __restrict__ ctrl,
                                                         double* __restrict__ res, int* __restrict__ hflag,
                                                         double* __restrict__ nest) {
    if (ctrl[0]) return;
    __shared__ double red[kBlock / 64];
    double s = strided_sum(partial, nparts), d = strided_sum(partial + nparts, nparts);
    s = block_sum(s, red);
    d = block_sum(d, red);
    if (threadIdx.x != 0) return;
    if (it < 0) it = ctrl[3];
    ctrl[3] = ctrl[3] + 1;
    const double err = 1.0 / sqrt((double)nv) * sqrt(s);
    res[0] = err;
    if (isnan(err) || isinf(err) || err <= 1.0) {
        ctrl[1] = it + 1;
        ctrl[2] = (isnan(err) || isinf(err)) ? CDX_DIVERGED : 0;
        ctrl[0] = 1;
        hflag[0] = 1;
        return;
    }
    double thk = nest[1];
    const double err_prev = nest[2];
    double restarts = nest[3];
    int disabled = nest[4] != 0.0;
    const double thk1 = 0.5 * (1.0 + sqrt(4.0 * thk * thk + 1.0));
    const double betak1 = (thk - 1.0) / thk1;
    bool restart = false;
    if (err_prev < INFINITY && err > 1.05 * err_prev) restart = true;
    else if (betak1 > beta_threshold) restart = true;
    else if (!isfinite(betak1) || betak1 < 0.0 || betak1 > 1.0) restart = true;
    else if (d > 0.0) restart = true;
    double beta = 0.0;
    if (restart) {
        thk = 1.0;
        restarts += 1.0;
        if (restarts >= (double)restart_limit) disabled = 1;
    } else if (disabled) {
        thk = 1.0;
    } else {
        beta = betak1;
        thk = thk1;
    }
    nest[0] = beta;
    nest[1] = thk;
    nest[2] = err;
    nest[3] = restarts;
    nest[4] = disabled ? 1.0 : 0.0;
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
    if (prm->accel != CDX_ACCEL_NONE && prm->accel != CDX_ACCEL_NESTEROV)
        return set_err(ctx, CDX_BAD_ARG, "cdx_gs_colored_solve: only the Nesterov acceleration is offered (pj.nesterov)");
    const bool nesterov = prm->accel == CDX_ACCEL_NESTEROV;
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
    CDX_LAUNCH(ctx, "gs_ufree", k_gs_ufree, gb, kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_v, ctx->d_fext,
               prm->implicit_gyro ? nullptr : ctx->d_fgyro, dt, ctx->d_ufree);
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
    const int ns = ctx->color_ptr.empty() ? 0 : ctx->color_ptr.back();
    if (ns > 0) {
        CDX_TRY(ensure_alloc(ctx, &ctx->d_gs_pack, (size_t)37 * ns, &ctx->gs_pack_cap, "colour-ordered contacts"));
        CDX_TRY(ensure_alloc(ctx, &ctx->d_gs_ids, (size_t)2 * ns, &ctx->gs_ids_cap, "colour-ordered contact ids"));
        CDX_LAUNCH(ctx, "gs_pack", k_gs_pack, grid_for(ns), kBlock, ns, nc, ctx->d_color_blocks, ctx->d_ida, ctx->d_idb,
                   ctx->d_J, ctx->d_R, ctx->d_bias, ctx->d_p, ctx->d_mu, ctx->d_gs_pack, ctx->d_gs_ids);
    }
    CDX_HIP(ctx, hipGetLastError());
    CDX_TRY(ensure_alloc(ctx, &ctx->d_partial, (size_t)gb, &ctx->partial_cap, "residual partials"));
    double* lam_ns = ctx->d_gs_pack + (size_t)33 * ns;
    auto sweep = [&](int it, int) -> int {
        for (int c = 0; c < ctx->ncolors; ++c) {
            const int n = ctx->color_ptr[c + 1] - ctx->color_ptr[c];
            if (n == 0) continue;
            CDX_LAUNCH(ctx, "gs_color", k_gs_color4, grid_for(n * kLanesPerContact), kBlock, n, ctx->color_ptr[c], ns,
                       ctx->d_gs_ids, ctx->d_gs_pack, ctx->d_minv, prm->relaxation, prm->alpha, lam_ns, ctx->d_x,
                       ctx->d_ctrl);
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
    int st = CDX_OK;
    if (!nesterov) {
        st = run_sweeps_graphed(ctx, prm->max_iter, kGsChunk, true, sweep, &iters);
    } else {
        // (lambda_k, u_k) = (lambda_y, u_y) = the warm-started start; beta 0, thk 1, err_prev inf
        CDX_TRY(ensure_alloc(ctx, &ctx->d_xk1, (size_t)nb * kBodyStride, &ctx->xk1_cap, "gs nesterov u_k"));
        CDX_TRY(ensure_alloc(ctx, &ctx->d_pk, 3 * (size_t)std::max(ns, 1), &ctx->pk_cap, "gs nesterov lambda_k"));
        CDX_TRY(ensure_alloc(ctx, &ctx->d_nest, 8, nullptr, "nesterov scalars"));
        CDX_TRY(ensure_alloc(ctx, &ctx->d_partial2, 2 * (size_t)gb, &ctx->partial2_cap, "gs nesterov partials"));
        CDX_HIP(ctx, hipMemcpyAsync(ctx->d_xk1, ctx->d_x, sizeof(double) * (size_t)nb * kBodyStride, hipMemcpyDeviceToDevice,
                                    ctx->stream));
        if (ns > 0)
            CDX_HIP(ctx, hipMemcpyAsync(ctx->d_pk, lam_ns, sizeof(double) * 3 * (size_t)ns, hipMemcpyDeviceToDevice,
                                        ctx->stream));
        const double n0[8] = {0.0, 1.0, INFINITY, 0.0, 0.0, 0.0, 0.0, 0.0};
        CDX_HIP(ctx, hipMemcpyAsync(ctx->d_nest, n0, sizeof(n0), hipMemcpyHostToDevice, ctx->stream));
        auto nsweep = [&](int it, int) -> int {
            CDX_LAUNCH(ctx, "gsn_extrapolate", k_gsn_extrapolate, grid_for(nb + ns), kBlock, nb, ns, ctx->d_x, ctx->d_xk1,
                       ctx->d_uprev, lam_ns, ctx->d_pk, ctx->d_nest, ctx->d_ctrl);
            for (int c = 0; c < ctx->ncolors; ++c) {
                const int n = ctx->color_ptr[c + 1] - ctx->color_ptr[c];
                if (n == 0) continue;
                CDX_LAUNCH(ctx, "gs_color", k_gs_color4, grid_for(n * kLanesPerContact), kBlock, n, ctx->color_ptr[c], ns,
                           ctx->d_gs_ids, ctx->d_gs_pack, ctx->d_minv, prm->relaxation, prm->alpha, lam_ns, ctx->d_x,
                           ctx->d_ctrl);
            }
            CDX_LAUNCH(ctx, "gsn_residual", k_gsn_residual, gb, kBlock, nb, ctx->d_dof, ctx->d_ufree, ctx->d_x, ctx->d_uprev,
                       ctx->d_xk1, prm->tol_rel, prm->tol_abs, ctx->d_partial2, ctx->d_ctrl);
            CDX_LAUNCH(ctx, "gs_residual", k_gsn_finalize, 1, kBlock, gb, ctx->d_partial2, ctx->nv, it,
                       prm->nesterov_beta_threshold, prm->nesterov_restart_limit, ctx->d_ctrl, ctx->d_res,
                       ctx->h_flag_dev, ctx->d_nest);
            return hip_check(ctx, hipGetLastError(), "gs nesterov sweep launch");
        };
        st = run_sweeps_graphed(ctx, prm->max_iter, kGsChunk, true, nsweep, &iters);
    }
    if (ns > 0)
        CDX_LAUNCH(ctx, "gs_unpack", k_gs_unpack, grid_for(ns), kBlock, ns, nc, ctx->d_color_blocks, ctx->d_gs_pack,
                   ctx->d_p);
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

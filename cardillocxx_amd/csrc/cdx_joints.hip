// cdx_joints.hip -- bilateral joint rows in the projected-Jacobi solve (the extended system S).
//
// The reference puts joint rows into the system matrix and factorises it with SparseLU
// (PjAssembler::buildAndFactorS / rhs / solveS, pj_assembler.cpp:9-182; rows from
// DynamicsAssembler::rebuildInteractionW_, dynamics_assembler.cpp:341-439):
//
//     S = [ M   G^T ]      G = [Wg; Wgamma],  E = diag(C / (theta h^2), A / (theta h))
//         [ G   -E  ]
//
// and every PJ sweep adds x += S^-1 [W^T dp; 0] (projected_jacobi.cpp:121-138).  Here S^-1 is
// applied through the Schur complement on the joint rows,
//     K = G M^-1 G^T + E,   Lambda = K^-1 (G M^-1 r_v - r_lambda),   v = M^-1 (r_v - G^T Lambda),
// which for joints forming chains (each body in at most two joint blocks, no loops: capsule chains,
// beams) is block tridiagonal.  It is factorised once per solve as K = L D L^T (block rows of
// up to kMaxJRows) by one thread per chain, and each application is a forward and a backward
// substitution along the chain -- the same arithmetic as an LU of S up to rounding.
//
// A joint block is the set of rows that couple the same pair of bodies (a body and the world
// when one side is static).  Bodies in no block keep the diagonal solve of the plain path.
#include <algorithm>
#include <map>
#include <vector>

#include "cdx_joints_device.hpp"

namespace cdx {

// One 64-lane workgroup per chain (see joint_factor_chain); a singular block stops the solve.
__global__ void __launch_bounds__(64) k_joint_factor(int nchains, const int* __restrict__ jc_off, const int* __restrict__ jb_off,
                                                     const int* __restrict__ jb_body, const int* __restrict__ jr_kind,
                                                     const int* __restrict__ jr_blk, const double* __restrict__ jr,
                                                     const double* __restrict__ minv, const double* __restrict__ v, double dt,
                                                     double theta, double beta, double* __restrict__ rlam,
                                                     const int* __restrict__ jl_off, double* __restrict__ Lm,
                                                     double* __restrict__ Dinv, int nrow,
                                                     int* __restrict__ ctrl) {
    __shared__ double lds[kFactorLds];
    const int ch = blockIdx.x;
    if (ch >= nchains) return;
    if (joint_factor_chain(ch, threadIdx.x, jc_off, jb_off, jb_body, jr_kind, jr_blk, jr, minv, v, dt, theta, beta, rlam, jl_off,
                           Lm, Dinv, nrow, lds) &&
        threadIdx.x == 0) {
        ctrl[2] = CDX_NOT_FACTORIZED;
        ctrl[0] = 1;
    }
}

// One 64-lane workgroup per chain (see joint_solve_chain); with `check` one residual partial per chain.
__global__ void __launch_bounds__(64) k_joint_solve(
    int nchains, int setup, int check, const int* __restrict__ jc_off, const int* __restrict__ jcb_off,
    const int* __restrict__ jb_off, const int* __restrict__ jl_off, const int* __restrict__ jr_info,
    const double* __restrict__ jr, const double* __restrict__ jw, int nrow,
    const double* __restrict__ rlam, const double* __restrict__ Lm, const double* __restrict__ Dinv,
    const int* __restrict__ jbody_info, const double* __restrict__ minv, const double* __restrict__ tr,
    double* __restrict__ z, double* __restrict__ lam,
    double* __restrict__ yb, double* __restrict__ x, double tol_rel, double tol_abs, double* __restrict__ partial,
    const int* __restrict__ ctrl) {
    __shared__ double lds[kSolveLdsL + kSolveLdsR];
    if (ctrl && ctrl[0]) return;
    const int ch = blockIdx.x;
    if (ch >= nchains) return;
    double e2 = joint_solve_chain<true>(ch, threadIdx.x, setup, check, jc_off, jcb_off, jb_off, jl_off, jr_info, jr, jw, nrow, rlam, Lm, Dinv,
                                  jbody_info, minv, tr, z, yb, lam, x, tol_rel, tol_abs, lds, kSolveLdsL, kSolveLdsR);
    if (check) {
        for (int off = 32; off > 0; off >>= 1) e2 += __shfl_xor(e2, off, 64);
        if (threadIdx.x == 0) partial[ch] = e2;
    }
}

// jw[k R + r] = jr[r kJRowStride + k], k < 12: the coefficients by coefficient
__global__ void k_joint_by_coef(int R, const double* __restrict__ jr, double* __restrict__ jw) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 12 * R) return;
    const int k = t / R, r = t - (t / R) * R;
    jw[t] = jr[(size_t)r * kJRowStride + k];
}

int joints_factor(cdx_ctx* ctx, double dt, double theta) {
    if (ctx->nchains == 0) return CDX_OK;
    CDX_LAUNCH(ctx, "joint_factor", k_joint_factor, ctx->nchains, 64, ctx->nchains, ctx->d_jc_off, ctx->d_jb_off,
               ctx->d_jb_body, ctx->d_jr_kind, ctx->d_jr_blk, ctx->d_jr, ctx->d_minv, ctx->d_v, dt, theta, ctx->joint_bias,
               ctx->d_jrlam, ctx->d_jl_off, ctx->d_jL, ctx->d_jDinv, ctx->njrows, ctx->d_ctrl);
    return hip_check(ctx, hipGetLastError(), "joint factor launch");
}

int joints_solve(cdx_ctx* ctx, int setup, int check, double tol_rel, double tol_abs, double* partial) {
    if (ctx->nchains == 0) return CDX_OK;
    CDX_LAUNCH(ctx, "joint_solve", k_joint_solve, ctx->nchains, 64, ctx->nchains, setup, check, ctx->d_jc_off, ctx->d_jcb_off,
               ctx->d_jb_off, ctx->d_jl_off, ctx->d_jr_info, ctx->d_jr, ctx->d_jw, ctx->njrows, ctx->d_jrlam, ctx->d_jL, ctx->d_jDinv, ctx->d_jbody_info, ctx->d_minv,
               ctx->d_jtr, ctx->d_jz, ctx->d_jlam, ctx->d_jy, ctx->d_x, tol_rel, tol_abs, partial,
               setup ? nullptr : ctx->d_ctrl);
    return hip_check(ctx, hipGetLastError(), "joint solve launch");
}

// out += S^-1 [tr; 0] for the jointed bodies (velocity part into x_out, joint part into lam_out;
// both accumulate): the joint rows' share of S^-1 applied to a body vector, no right-hand side.
int joints_apply(cdx_ctx* ctx, double* x_out, double* lam_out, const int* ctrl) {
    if (ctx->nchains == 0) return CDX_OK;
    CDX_LAUNCH(ctx, "joint_solve", k_joint_solve, ctx->nchains, 64, ctx->nchains, 0, 0, ctx->d_jc_off, ctx->d_jcb_off,
               ctx->d_jb_off, ctx->d_jl_off, ctx->d_jr_info, ctx->d_jr, ctx->d_jw, ctx->njrows, ctx->d_jrlam, ctx->d_jL, ctx->d_jDinv, ctx->d_jbody_info, ctx->d_minv,
               ctx->d_jtr, ctx->d_jz, lam_out, ctx->d_jy, x_out, 0.0, 0.0, nullptr, ctrl);
    return hip_check(ctx, hipGetLastError(), "joint apply launch");
}

}  // namespace cdx

using namespace cdx;

// Row values in block order (jorig: block-ordered row -> caller's row): jr, 16 doubles per row,
// uploaded, and on the device the same coefficients by coefficient (jw[k R + r]), so that a wave
// reading one coefficient of 64 rows touches 4 cache lines instead of 64.
static int upload_joint_values(cdx_ctx* ctx, const std::vector<int>& jorig, const double* wa, const double* wb,
                               const double* comp, const double* vsrc, const double* gerr, const double* lam_prev) {
    const size_t R = jorig.size(), n = R * kJRowStride;
    // gathered straight into the pinned staging buffer: one DMA copy at the link's rate
    CDX_TRY(ensure_stage(ctx, n > 0 ? n : 1));
    double* jr = ctx->h_stage;
    parallel_ranges((int)R, host_threads((int)R), [&](int, int i0, int i1) {
        for (size_t i = (size_t)i0; i < (size_t)i1; ++i) {
            const size_t r = (size_t)jorig[i];
            double* o = jr + i * kJRowStride;
            for (int k = 0; k < 6; ++k) o[k] = wa[6 * r + k];
            for (int k = 0; k < 6; ++k) o[6 + k] = wb[6 * r + k];
            o[12] = comp[r];
            o[13] = vsrc[r];
            o[14] = gerr[r];
            o[15] = lam_prev[r];
        }
    });
    CDX_TRY(ensure_alloc(ctx, &ctx->d_jr, n, &ctx->jr_cap, "joint rows"));
    if (n > 0) CDX_HIP(ctx, hipMemcpyAsync(ctx->d_jr, jr, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_jw, 12 * R, &ctx->jw_cap, "joint rows"));
    CDX_LAUNCH(ctx, "joint_by_coef", k_joint_by_coef, grid_for((int)(12 * R)), kBlock, (int)R, ctx->d_jr, ctx->d_jw);
    return hip_check(ctx, hipGetLastError(), "joint coefficient launch");
}

extern "C" {

int cdx_set_joints(cdx_ctx* ctx, int ns, int nd, const int* body_a, const int* body_b, const double* wa, const double* wb,
                   const double* comp, const double* vsrc, const double* gerr, const double* lam_prev, double bias_factor) {
    if (!ctx) return CDX_BAD_ARG;
    hipSetDevice(ctx->device);
    ctx->solved = false;
    ctx->nchains = 0;
    ctx->njrows = 0;
    const bool had_tables = ctx->jsig_ns >= 0;
    const int sig_ns = ctx->jsig_ns, sig_nd = ctx->jsig_nd, sig_nb = ctx->jsig_nb;
    ctx->jsig_ns = -1;  // the device tables count as stale until this call completes
    if (ns < 0 || nd < 0) return set_err(ctx, CDX_BAD_ARG, "cdx_set_joints: negative row count");
    const int n = ns + nd;
    if (n == 0) return CDX_OK;
    if (!body_a || !body_b || !wa || !wb || !comp || !vsrc || !gerr || !lam_prev)
        return set_err(ctx, CDX_BAD_ARG, "cdx_set_joints: null array");
    if (!std::isfinite(bias_factor) || bias_factor < 0) return set_err(ctx, CDX_BAD_ARG, "cdx_set_joints: bad bias factor");
    const int nb = ctx->nb;
    for (int r = 0; r < n; ++r) {
        const int a = body_a[r], b = body_b[r];
        if (a < -1 || a >= nb || b < -1 || b >= nb || (a < 0 && b < 0) || a == b)
            return set_err(ctx, CDX_BAD_ARG, "cdx_set_joints: bad body index");
        for (int side = 0; side < 2; ++side) {
            const int body = side ? b : a;
            if (body >= 0 && ctx->dof[body] != 6)
                return set_err(ctx, CDX_BAD_ARG, "cdx_set_joints: jointed bodies must be 6-dof rigid bodies");
        }
        for (int k = 0; k < 6; ++k)
            if (!std::isfinite(wa[6 * (size_t)r + k]) || !std::isfinite(wb[6 * (size_t)r + k]))
                return set_err(ctx, CDX_NONFINITE, "cdx_set_joints: non-finite row");
        if (!std::isfinite(comp[r]) || comp[r] < 0 || !std::isfinite(vsrc[r]) || !std::isfinite(gerr[r]) ||
            !std::isfinite(lam_prev[r]))
            return set_err(ctx, CDX_NONFINITE, "cdx_set_joints: bad compliance or source term");
    }
    // the same row bodies as the tables on the device: only the row values change
    if (had_tables && ns == sig_ns && nd == sig_nd && nb == sig_nb && std::equal(body_a, body_a + n, ctx->h_jsig_a.begin()) &&
        std::equal(body_b, body_b + n, ctx->h_jsig_b.begin())) {
        CDX_TRY(upload_joint_values(ctx, ctx->h_jorig, wa, wb, comp, vsrc, gerr, lam_prev));
        ctx->jsig_ns = ns;
        ctx->njrows = n;
        ctx->nchains = ctx->jsig_nchains;
        ctx->njbodies = ctx->jsig_njb;
        ctx->joint_bias = bias_factor;
        return CDX_OK;
    }
    // blocks: rows grouped by body pair, in order of first appearance
    std::map<std::pair<int, int>, int> key;
    std::vector<std::vector<int>> blk_rows;
    std::vector<std::pair<int, int>> blk_body;
    for (int r = 0; r < n; ++r) {
        const auto k = std::make_pair(body_a[r], body_b[r]);
        auto it = key.find(k);
        if (it == key.end()) {
            it = key.emplace(k, (int)blk_rows.size()).first;
            blk_rows.emplace_back();
            blk_body.push_back(k);
        }
        blk_rows[it->second].push_back(r);
    }
    const int nblk = (int)blk_rows.size();
    for (const auto& rows : blk_rows)
        if ((int)rows.size() > kMaxJRows) return set_err(ctx, CDX_BAD_ARG, "cdx_set_joints: more than 12 rows on one body pair");
    // chains: every dynamic body in at most two blocks, no loops
    std::vector<std::vector<int>> body_blk(nb);
    for (int j = 0; j < nblk; ++j)
        for (int b : {blk_body[j].first, blk_body[j].second})
            if (b >= 0) body_blk[b].push_back(j);
    for (int b = 0; b < nb; ++b)
        if (body_blk[b].size() > 2)
            return set_err(ctx, CDX_BAD_ARG, "cdx_set_joints: a body in more than two joint blocks (joints must form chains)");
    auto neighbours = [&](int j, int other[2]) {
        int c = 0;
        for (int b : {blk_body[j].first, blk_body[j].second}) {
            if (b < 0) continue;
            for (int k : body_blk[b])
                if (k != j) other[c++] = k;
        }
        return c;
    };
    std::vector<int> order, chain_off{0}, visited(nblk, 0);
    for (int pass = 0; pass < 2; ++pass)
        for (int j0 = 0; j0 < nblk; ++j0) {
            if (visited[j0]) continue;
            int nbrs[4];
            if (pass == 0 && neighbours(j0, nbrs) > 1) continue;  // start chains at their ends
            if (pass == 1) return set_err(ctx, CDX_BAD_ARG, "cdx_set_joints: joint blocks form a loop");
            int prev = -1, cur = j0;
            while (cur >= 0) {
                visited[cur] = 1;
                order.push_back(cur);
                const int cnt = neighbours(cur, nbrs);
                int next = -1;
                for (int k = 0; k < cnt; ++k)
                    if (nbrs[k] != prev && !visited[nbrs[k]]) next = nbrs[k];
                prev = cur;
                cur = next;
            }
            chain_off.push_back((int)order.size());
        }
    const int nchains = (int)chain_off.size() - 1;
    // block-ordered rows and per-block data
    std::vector<int> jb_off{0}, jb_body, jr_kind, jorig, jr_blk;
    for (int j : order) {
        for (int r : blk_rows[j]) {
            jr_blk.push_back((int)jb_off.size() - 1);
            jorig.push_back(r);
            jr_kind.push_back(r < ns ? 0 : 1);
        }
        jb_off.push_back((int)jorig.size());
        jb_body.push_back(blk_body[j].first);
        jb_body.push_back(blk_body[j].second);
    }
    // jointed bodies in chain order, each with its (block, side) pairs
    std::vector<int> pos_of(nblk);
    for (int i = 0; i < nblk; ++i) pos_of[order[i]] = i;
    std::vector<int> jbody, jbody_blk, jcb_off{0}, jslot(nb, -1);
    for (int c = 0; c < nchains; ++c) {
        for (int i = chain_off[c]; i < chain_off[c + 1]; ++i)
            for (int b : {jb_body[2 * i], jb_body[2 * i + 1]}) {
                if (b < 0 || jslot[b] >= 0) continue;
                jslot[b] = (int)jbody.size();
                jbody.push_back(b);
                int e = 0, ent[4] = {-1, -1, -1, -1};
                for (int k : body_blk[b]) {
                    const int pk = pos_of[k];
                    ent[2 * e] = pk;
                    ent[2 * e + 1] = jb_body[2 * pk] == b ? 0 : 1;
                    ++e;
                }
                jbody_blk.insert(jbody_blk.end(), ent, ent + 4);
            }
        jcb_off.push_back((int)jbody.size());
    }
    const size_t R = jorig.size(), NJB = jbody.size();
    auto upi = [&](int** d, const std::vector<int>& h, size_t* cap, const char* what) -> int {
        CDX_TRY(ensure_alloc(ctx, d, h.size(), cap, what));
        return hip_check(ctx, hipMemcpy(*d, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice), what);
    };
    CDX_TRY(upi(&ctx->d_jc_off, chain_off, &ctx->jc_off_cap, "joint chains"));
    CDX_TRY(upi(&ctx->d_jcb_off, jcb_off, &ctx->jcb_off_cap, "joint chain bodies"));
    CDX_TRY(upi(&ctx->d_jb_off, jb_off, &ctx->jb_off_cap, "joint blocks"));
    // L_j (m_j x m_{j-1}) packed back to back along each chain, none for a chain's first block
    std::vector<int> jl_off(jb_off.size(), 0);
    for (int c = 0; c < nchains; ++c)
        for (int i = chain_off[c]; i < chain_off[c + 1]; ++i) {
            const size_t ml = i > chain_off[c] ? (size_t)(jb_off[i + 1] - jb_off[i]) * (jb_off[i] - jb_off[i - 1]) : 0;
            if ((size_t)jl_off[i] + ml >= (1u << 31)) return set_err(ctx, CDX_BAD_ARG, "cdx_set_joints: too many joint rows");
            jl_off[i + 1] = jl_off[i] + (int)ml;
        }
    CDX_TRY(upi(&ctx->d_jl_off, jl_off, &ctx->jl_off_cap, "joint factor offsets"));
    CDX_TRY(upi(&ctx->d_jb_body, jb_body, &ctx->jb_body_cap, "joint block bodies"));
    CDX_TRY(upi(&ctx->d_jr_kind, jr_kind, &ctx->jr_kind_cap, "joint row kinds"));
    CDX_TRY(upi(&ctx->d_jr_blk, jr_blk, &ctx->jr_blk_cap, "joint row blocks"));
    CDX_TRY(upi(&ctx->d_jbody, jbody, &ctx->jbody_cap, "jointed bodies"));
    CDX_TRY(upi(&ctx->d_jbody_blk, jbody_blk, &ctx->jbody_blk_cap, "jointed body blocks"));
    // per-row and per-jointed-body records of the chain solve (one load instead of an index chain):
    // row r: {body a, body b, block, r0 << 4 | m}; jointed body q: {body, rows of its first block,
    // rows of its second block (-1: none), 0} with rows encoded r0 << 5 | m << 1 | side
    std::vector<int> jr_info(4 * jorig.size()), jbody_info(4 * jbody.size());
    for (size_t j = 0; j + 1 < jb_off.size(); ++j)
        for (int r = jb_off[j]; r < jb_off[j + 1]; ++r) {
            jr_info[4 * (size_t)r] = jb_body[2 * j];
            jr_info[4 * (size_t)r + 1] = jb_body[2 * j + 1];
            jr_info[4 * (size_t)r + 2] = (int)j;
            jr_info[4 * (size_t)r + 3] = (jb_off[j] << 4) | (jb_off[j + 1] - jb_off[j]);
        }
    for (size_t q = 0; q < jbody.size(); ++q) {
        jbody_info[4 * q] = jbody[q];
        for (int e = 0; e < 2; ++e) {
            const int j = jbody_blk[4 * q + 2 * e], side = jbody_blk[4 * q + 2 * e + 1];
            jbody_info[4 * q + 1 + e] = j < 0 ? -1 : (jb_off[j] << 5) | ((jb_off[j + 1] - jb_off[j]) << 1) | side;
        }
        jbody_info[4 * q + 3] = 0;
    }
    if (jorig.size() >= (1u << 26)) return set_err(ctx, CDX_BAD_ARG, "cdx_set_joints: too many joint rows");
    CDX_TRY(upi(&ctx->d_jr_info, jr_info, &ctx->jr_info_cap, "joint row records"));
    CDX_TRY(upi(&ctx->d_jbody_info, jbody_info, &ctx->jbody_info_cap, "jointed body records"));
    CDX_TRY(upi(&ctx->d_jslot, jslot, &ctx->jslot_cap, "joint slots"));
    CDX_TRY(upload_joint_values(ctx, jorig, wa, wb, comp, vsrc, gerr, lam_prev));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_jrlam, R, &ctx->jrlam_cap, "joint rhs"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_jz, R, &ctx->jz_cap, "joint scratch"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_jy, R, &ctx->jy_cap, "joint scratch"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_jlam, R, &ctx->jlam_cap, "joint impulses"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_jL, (size_t)std::max(jl_off.back(), 1), &ctx->jL_cap, "joint factors"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_jDinv, (size_t)kMaxJRows * R, &ctx->jDinv_cap, "joint factors"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_jtr, (size_t)nb * 6, &ctx->jtr_cap, "joint body rhs"));
    ctx->h_jorig = jorig;
    ctx->h_jc_off = chain_off;
    ctx->h_jb_body = jb_body;
    ctx->njrows = (int)R;
    ctx->nchains = nchains;
    ctx->njbodies = (int)NJB;
    ctx->joint_bias = bias_factor;
    ctx->h_jsig_a.assign(body_a, body_a + n);
    ctx->h_jsig_b.assign(body_b, body_b + n);
    ctx->jsig_ns = ns;
    ctx->jsig_nd = nd;
    ctx->jsig_nb = nb;
    ctx->jsig_nchains = nchains;
    ctx->jsig_njb = (int)NJB;
    return CDX_OK;
}

int cdx_get_joint_impulses(cdx_ctx* ctx, double* lam_out) {
    if (!ctx) return CDX_BAD_ARG;
    if (ctx->njrows == 0) return CDX_OK;
    if (!lam_out) return CDX_BAD_ARG;
    if (!ctx->solved || !ctx->joints_solved) return set_err(ctx, CDX_NO_STATE, "cdx_get_joint_impulses: no joint solve yet");
    std::vector<double> h(ctx->njrows);
    CDX_HIP(ctx, hipMemcpy(h.data(), ctx->d_jlam, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < ctx->njrows; ++i) lam_out[ctx->h_jorig[i]] = h[i];
    return CDX_OK;
}

}  // extern "C"

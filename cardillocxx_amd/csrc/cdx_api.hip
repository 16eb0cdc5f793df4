// cdx_api.hip -- context, state upload/download, Moreau position update and profiling.
//
// The state set here is what DynamicsAssembler stacks for a solver
// (src/physics/assembly/dynamics_assembler.cpp:41-120,211-332): bodies in body order and a
// contact buffer in Contact[] order.  Everything is copied once into HBM and stays there.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iterator>

#include "cdx_device.hpp"

namespace cdx {

int set_err(cdx_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int hip_check(cdx_ctx* ctx, hipError_t e, const char* what) {
    if (e == hipSuccess) return CDX_OK;
    return set_err(ctx, CDX_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

int grid_for(int n) { return (n + kBlock - 1) / kBlock; }

static hipEvent_t pool_event(cdx_ctx* ctx) {
    hipEvent_t e;
    if (!ctx->event_pool.empty()) {
        e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
    } else {
        hipEventCreate(&e);
    }
    return e;
}

void prof_events(cdx_ctx* ctx, hipEvent_t* start, hipEvent_t* stop) {
    *start = pool_event(ctx);
    *stop = pool_event(ctx);
}

void prof_push(cdx_ctx* ctx, const char* name, hipEvent_t start, hipEvent_t stop) {
    ctx->pending.push_back(PendingEvent{name, start, stop});
    if (ctx->pending.size() >= 4096) harvest_events(ctx);
}

int harvest_events(cdx_ctx* ctx) {
    if (ctx->pending.empty()) return CDX_OK;
    hipEventSynchronize(ctx->pending.back().stop);
    for (auto& p : ctx->pending) {
        float ms = 0;
        hipEventElapsedTime(&ms, p.start, p.stop);
        auto& s = ctx->stats[p.name];
        s.calls += 1;
        s.ms += ms;
        ctx->event_pool.push_back(p.start);
        ctx->event_pool.push_back(p.stop);
    }
    ctx->pending.clear();
    return CDX_OK;
}

// Body -> incident contact-side CSR, sides listed in contact order so the per-body sum has a
// fixed order.  Entry value = 2*c + side.
// Grows the device and pinned host staging buffers of the uploads and downloads to n doubles.
// 0 when contact c passes cdx_set_contacts's checks, else 1 (body index), 2 (mu), 3 (frame or
// angular rows), 4 (restitution, bias or warm start), in the order they are tested.
static int contact_fault(int nb, int c, const int* body_a, const int* body_b, const double* mu, const double* frame,
                         const double* angA, const double* angB, const double* rest, const double* vbias,
                         const double* warm) {
    const int a = body_a[c], b = body_b[c];
    if (a < -1 || a >= nb || b < -1 || b >= nb || (a >= 0 && a == b)) return 1;
    if (!std::isfinite(mu[c])) return 2;
    for (int k = 0; k < 9; ++k)
        if (!std::isfinite(frame[9 * (size_t)c + k]) || !std::isfinite(angA[9 * (size_t)c + k]) ||
            !std::isfinite(angB[9 * (size_t)c + k]))
            return 3;
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(rest[3 * (size_t)c + k]) || !std::isfinite(vbias[3 * (size_t)c + k]) ||
            !std::isfinite(warm[3 * (size_t)c + k]))
            return 4;
    return 0;
}

int ensure_stage(cdx_ctx* ctx, size_t n) {
    // grown by half again at least: a contact count that creeps up step by step would otherwise
    // pin and map a new host buffer (tens of ms for 100 MB) on every step
    if (ctx->stage_cap < n) CDX_TRY(ensure_alloc(ctx, &ctx->d_stage, std::max(n, ctx->stage_cap + ctx->stage_cap / 2),
                                                 &ctx->stage_cap, "staging"));
    if (ctx->h_stage_cap < n) {
        const size_t want = std::max(n, ctx->h_stage_cap + ctx->h_stage_cap / 2);
        CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->h_stage) hipHostFree(ctx->h_stage);
        ctx->h_stage = nullptr;
        ctx->h_stage_cap = 0;
        if (hipHostMalloc((void**)&ctx->h_stage, want * sizeof(double), hipHostMallocDefault) != hipSuccess) {
            ctx->h_stage = nullptr;
            return set_err(ctx, CDX_HIP_ERROR, "staging: pinned host allocation failed");
        }
        ctx->h_stage_cap = want;
    }
    return CDX_OK;
}

// The stacked array packed on the device goes to the caller through the pinned staging buffer:
// one DMA copy at the link's rate and one host copy, where a pageable copy of the padded array
// and a repack on the host took several times as long.
static int download_stage(cdx_ctx* ctx, size_t n, double* out) {
    CDX_HIP(ctx, hipMemcpyAsync(ctx->h_stage, ctx->d_stage, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (out)
        parallel_ranges((int)n, host_threads((int)(n / 8)), [&](int, int i0, int i1) {
            std::memcpy(out + i0, ctx->h_stage + i0, (size_t)(i1 - i0) * sizeof(double));
        });
    return CDX_OK;
}

int build_incidence(cdx_ctx* ctx) {
    if (ctx->incidence_valid) return CDX_OK;
    const int nb = ctx->nb, nc = ctx->nc;
    std::vector<int> ptr(nb + 1, 0), idx(2 * (size_t)nc + 1);
    for (int c = 0; c < nc; ++c) {
        const int a = ctx->h_a[c], b = ctx->h_b[c];
        if (a >= 0) ptr[a + 1]++;
        if (b >= 0) ptr[b + 1]++;
    }
    for (int i = 0; i < nb; ++i) ptr[i + 1] += ptr[i];
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int c = 0; c < nc; ++c) {
        const int a = ctx->h_a[c], b = ctx->h_b[c];
        if (a >= 0) idx[fill[a]++] = 2 * c;
        if (b >= 0) idx[fill[b]++] = 2 * c + 1;
    }
    // d_incptr is sized with the body capacity, d_incidx with the contact capacity; both go
    // through the pinned staging buffer
    const size_t nptr = (size_t)nb + 1, nidx = 2 * (size_t)nc;
    CDX_TRY(ensure_stage(ctx, (nptr + nidx + 1) / 2));
    int* hi = reinterpret_cast<int*>(ctx->h_stage);
    std::memcpy(hi, ptr.data(), nptr * sizeof(int));
    std::memcpy(hi + nptr, idx.data(), nidx * sizeof(int));
    CDX_HIP(ctx, hipMemcpyAsync(ctx->d_incptr, hi, sizeof(int) * nptr, hipMemcpyHostToDevice, ctx->stream));
    if (nc > 0)
        CDX_HIP(ctx, hipMemcpyAsync(ctx->d_incidx, hi + nptr, sizeof(int) * nidx, hipMemcpyHostToDevice, ctx->stream));
    CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    idx.resize(2 * (size_t)nc);
    ctx->h_incptr = std::move(ptr);
    ctx->h_incidx = std::move(idx);
    ctx->incidence_valid = true;
    return CDX_OK;
}

// ------------------------------------------------------------------------------------------
// Kernels
// ------------------------------------------------------------------------------------------

// f_gyro = -omega x (I omega), I = diag from the inverse inertia (world.cpp:152-173,
// M = 1/Minv as rebuildMass_ does, dynamics_assembler.cpp:56-60).
__global__ void k_gyro(int nb, const int* __restrict__ dof, const double* __restrict__ minv,
                       const double* __restrict__ v, double* __restrict__ fg) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double* o = fg + (size_t)b * kBodyStride;
    for (int k = 0; k < kBodyStride; ++k) o[k] = 0.0;
    if (dof[b] != 6) return;
    gyro_force3(v + (size_t)b * kBodyStride + 3, minv + (size_t)b * kMinvStride + 1, o + 3);
}

// The caller's contact-major contact arrays (staged as frame | angA | angB [nc][9], rest | vbias |
// warm [nc][3], mu [nc], then body_a | body_b) to the component-major device arrays, with
// rebuildW_'s masks (dynamics_assembler.cpp:211-332): no angular rows on a static side or a point
// mass (buildContactRowByDof, dof 3), no tangent rows or friction without mu > 0, nothing on a
// static-static contact (:259-260).
__global__ void k_contacts_soa(int nc, const int* __restrict__ dof, const double* __restrict__ raw,
                               const int* __restrict__ ids, int* __restrict__ ida, int* __restrict__ idb,
                               double* __restrict__ J, double* __restrict__ mu_o, double* __restrict__ rest_o,
                               double* __restrict__ vb_o, double* __restrict__ p_o) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const size_t n = (size_t)nc;
    const double *frame = raw, *angA = raw + 9 * n, *angB = raw + 18 * n, *rest = raw + 27 * n,
                 *vbias = raw + 30 * n, *warm = raw + 33 * n, *mu = raw + 36 * n;
    int a = ids[c], b = ids[n + c];
    const bool has_rows = a >= 0 || b >= 0;
    if (!has_rows) a = b = -1;
    ida[c] = a;
    idb[c] = b;
    const bool rotA = a >= 0 && dof[a] != 3, rotB = b >= 0 && dof[b] != 3;
    for (int k = 0; k < 9; ++k) {
        J[k * n + c] = frame[9 * (size_t)c + k];
        J[(9 + k) * n + c] = rotA ? angA[9 * (size_t)c + k] : 0.0;
        J[(18 + k) * n + c] = rotB ? angB[9 * (size_t)c + k] : 0.0;
    }
    const bool fr = mu[c] > 0;
    mu_o[c] = fr ? mu[c] : 0.0;
    for (int k = 0; k < 3; ++k) {
        const bool row = has_rows && (k == 0 || fr);
        rest_o[k * n + c] = row ? rest[3 * (size_t)c + k] : 0.0;
        vb_o[k * n + c] = row ? vbias[3 * (size_t)c + k] : 0.0;
        p_o[k * n + c] = row ? warm[3 * (size_t)c + k] : 0.0;
    }
}

// Packs a per-body padded array (kBodyStride doubles per body) into the stacked layout: body b's
// first n entries (n = its DOFs, or 7 / 3 position entries with `positions`) at off[b].
__global__ void k_pack_stacked(int nb, const int* __restrict__ dof, const int* __restrict__ off, int positions,
                               const double* __restrict__ src, double* __restrict__ dst) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb * kBodyStride) return;
    const int b = t / kBodyStride, k = t % kBodyStride;
    const int n = positions ? (dof[b] == 6 ? 7 : 3) : dof[b];
    if (k < n) dst[off[b] + k] = src[t];
}

// Component-major impulses [3][nc] to contact-major [nc][3].
__global__ void k_pack_impulses(int nc, const double* __restrict__ p, double* __restrict__ dst) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * nc) return;
    const int c = t / 3, k = t % 3;
    dst[t] = p[(size_t)k * nc + c];
}

// Bodies with a non-finite twist or pose entry, one workgroup sum each.
__global__ void __launch_bounds__(kBlock) k_nonfinite_bodies(int nb, const int* __restrict__ dof,
                                                            const double* __restrict__ x, const double* __restrict__ q,
                                                            int have_q, double* __restrict__ partial) {
    __shared__ double red[kBlock / 64];
    const int b = blockIdx.x * kBlock + threadIdx.x;
    double bad = 0;
    if (b < nb) {
        const int nd = dof[b], nq = nd == 6 ? 7 : 3;
        bool ok = true;
        for (int k = 0; k < nd; ++k) ok = ok && isfinite(x[(size_t)b * kBodyStride + k]);
        if (have_q)
            for (int k = 0; k < nq; ++k) ok = ok && isfinite(q[(size_t)b * kBodyStride + k]);
        bad = ok ? 0.0 : 1.0;
    }
    const double s = block_sum(bad, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Explicit position update (integration_base.cpp:6-30) with alignQuaternionTo
// (math_helper.hpp:11-17).  When `vsrc` differs from v, the solved velocities are first
// written back as the body velocities (writeVelocityToSystem, dynamics_assembler.cpp:145-188).
__global__ void k_position_update(int nb, const int* __restrict__ dof, double* __restrict__ q,
                                  double* __restrict__ v, const double* __restrict__ vsrc, double h, int write_v) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double* u = v + (size_t)b * kBodyStride;
    if (write_v) {
        const double* s = vsrc + (size_t)b * kBodyStride;
        for (int k = 0; k < 6; ++k) u[k] = s[k];
    }
    double* x = q + (size_t)b * kBodyStride;
    x[0] += h * u[0];
    x[1] += h * u[1];
    x[2] += h * u[2];
    if (dof[b] != 6) return;
    const double w0 = x[6];
    const double P0 = x[3], P1 = x[4], P2 = x[5];
    const double wx = u[3], wy = u[4], wz = u[5];
    const double hh = h * 0.5;
    double nw = w0 - hh * (P0 * wx + P1 * wy + P2 * wz);
    const double c0 = P1 * wz - P2 * wy;
    const double c1 = P2 * wx - P0 * wz;
    const double c2 = P0 * wy - P1 * wx;
    double n0 = P0 + hh * (w0 * wx + c0);
    double n1 = P1 + hh * (w0 * wy + c1);
    double n2 = P2 + hh * (w0 * wz + c2);
    if (!(isfinite(n0) && isfinite(n1) && isfinite(n2) && isfinite(nw))) return;  // keep q_prev
    const double s2 = n0 * n0 + n1 * n1 + n2 * n2 + nw * nw;
    if (s2 > 0) {
        const double nn = sqrt(s2);
        n0 /= nn;
        n1 /= nn;
        n2 /= nn;
        nw /= nn;
    }
    if (P0 * n0 + P1 * n1 + P2 * n2 + w0 * nw < 0) {
        n0 = -n0;
        n1 = -n1;
        n2 = -n2;
        nw = -nw;
    }
    x[3] = n0;
    x[4] = n1;
    x[5] = n2;
    x[6] = nw;
}

}  // namespace cdx

using namespace cdx;

extern "C" {

void cdx_default_params(cdx_solver_params* p) {
    if (!p) return;
    p->alpha = 0.3;
    p->relaxation = 1.0;
    p->max_iter = 200000;
    p->tol_abs = 1e-5;
    p->tol_rel = 1e-5;
    p->check_every = 10;
    p->warmstart = 1;
    p->accel = CDX_ACCEL_NONE;
    p->nesterov_beta_threshold = 1.0;
    p->nesterov_restart_limit = 10000;
    p->block_precond = 0;
    p->anderson_m = 2;
    p->implicit_gyro = 0;
}

int cdx_ctx_create(int device, cdx_ctx** out) {
    if (!out) return CDX_BAD_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return CDX_HIP_ERROR;
    if (device < 0 || device >= n) return CDX_BAD_ARG;
    cdx_ctx* ctx = new cdx_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return CDX_HIP_ERROR;
    }
    if (hipMalloc((void**)&ctx->d_ctrl, 4 * sizeof(int)) != hipSuccess || hipMalloc((void**)&ctx->d_res, 4 * sizeof(double)) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_flag, 4 * sizeof(int), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer((void**)&ctx->h_flag_dev, ctx->h_flag, 0) != hipSuccess) {
        cdx_ctx_destroy(ctx);
        return CDX_HIP_ERROR;
    }
    *out = ctx;
    return CDX_OK;
}

void cdx_ctx_destroy(cdx_ctx* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    void* bufs[] = {ctx->d_x, ctx->d_v, ctx->d_minv, ctx->d_fext, ctx->d_fgyro, ctx->d_q, ctx->d_uprev, ctx->d_ufree, ctx->d_dof,
                    ctx->d_ida, ctx->d_idb, ctx->d_J, ctx->d_mu, ctx->d_rest, ctx->d_vb, ctx->d_p, ctx->d_R, ctx->d_bias,
                    ctx->d_imp, ctx->d_incptr, ctx->d_incidx, ctx->d_color_blocks, ctx->d_ctrl, ctx->d_res, ctx->d_partial,
                    ctx->d_partial2, ctx->d_pk, ctx->d_nest, ctx->d_xk1, ctx->d_pk1, ctx->d_cheb, ctx->d_Rb, ctx->d_scene_i, ctx->d_scene_res,
                    ctx->d_rec_x, ctx->d_rec_p, ctx->d_adj, ctx->d_adj_imp, ctx->d_adj_body, ctx->d_adj_partial,
                    ctx->d_jc_off, ctx->d_jcb_off, ctx->d_jb_off, ctx->d_jb_body, ctx->d_jr_kind, ctx->d_jbody,
                    ctx->d_jbody_blk, ctx->d_jslot, ctx->d_jr, ctx->d_jrlam, ctx->d_jz, ctx->d_jlam, ctx->d_jL,
                    ctx->d_jDinv, ctx->d_jtr, ctx->d_sc_off, ctx->d_sc_list, ctx->d_adj_jl, ctx->d_jr_blk, ctx->d_jy, ctx->d_gm, ctx->d_gm_part, ctx->d_jr_info, ctx->d_jbody_info,
                    ctx->d_jl_off, ctx->d_jw, ctx->d_and, ctx->d_and_gam, ctx->d_and_part, ctx->d_and_st,
                    ctx->d_gs_pack, ctx->d_gs_ids, ctx->d_rg_p, ctx->d_rg_l, ctx->d_pr_ptr, ctx->d_pr_out, ctx->d_pr_inc,
                    ctx->d_bp_ptr, ctx->d_bp_ovf, ctx->d_part, ctx->d_accj, ctx->d_sinv, ctx->d_geo, ctx->d_geo_pt,
                    ctx->d_geo_bad, ctx->d_voff, ctx->d_qoff, ctx->d_stage};
    for (void* p : bufs)
        if (p) hipFree(p);
    if (ctx->h_flag) hipHostFree(ctx->h_flag);
    if (ctx->h_stage) hipHostFree(ctx->h_stage);
    for (auto& p : ctx->pending) {
        hipEventDestroy(p.start);
        hipEventDestroy(p.stop);
    }
    for (auto e : ctx->event_pool) hipEventDestroy(e);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* cdx_last_error(const cdx_ctx* ctx) { return ctx ? ctx->err.c_str() : noalloc_last_error(); }

int cdx_get_stream(cdx_ctx* ctx, void** stream) {
    if (!ctx || !stream) return CDX_BAD_ARG;
    *stream = static_cast<void*>(ctx->stream);
    return CDX_OK;
}

int cdx_synchronize(cdx_ctx* ctx) {
    if (!ctx) return CDX_BAD_ARG;
    CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CDX_OK;
}

int cdx_set_bodies(cdx_ctx* ctx, int nb, const int* dof, const double* minv, const double* v, const double* fext,
                   const double* fgyro, const double* q) {
    if (!ctx) return CDX_BAD_ARG;
    if (nb < 0 || (nb > 0 && (!dof || !minv || !v || !fext))) return set_err(ctx, CDX_BAD_ARG, "cdx_set_bodies: null array");
    hipSetDevice(ctx->device);
    std::vector<int> voff(nb + 1), qoff(nb + 1);
    int nv = 0, nq = 0;
    for (int b = 0; b < nb; ++b) {
        if (dof[b] != 3 && dof[b] != 6) return set_err(ctx, CDX_BAD_ARG, "cdx_set_bodies: dof must be 3 or 6");
        voff[b] = nv;
        qoff[b] = nq;
        nv += dof[b];
        nq += dof[b] == 6 ? 7 : 3;
    }
    voff[nb] = nv;
    qoff[nb] = nq;
    std::vector<double> hx((size_t)nb * kBodyStride, 0.0), hf((size_t)nb * kBodyStride, 0.0),
        hg((size_t)nb * kBodyStride, 0.0), hq((size_t)nb * kBodyStride, 0.0), hm((size_t)nb * kMinvStride, 0.0);
    bool singular = false;
    for (int b = 0; b < nb; ++b) {
        const int o = voff[b], d = dof[b];
        // The reference's mass diagonal has one mass on all translational dofs (world.cpp:49-64).
        if (minv[o] != minv[o + 1] || minv[o] != minv[o + 2])
            return set_err(ctx, CDX_BAD_ARG, "cdx_set_bodies: translational inverse mass must be isotropic");
        hm[(size_t)b * kMinvStride] = minv[o];
        for (int k = 0; k < d; ++k) {
            if (!std::isfinite(minv[o + k]) || !std::isfinite(v[o + k]) || !std::isfinite(fext[o + k]) ||
                (fgyro && !std::isfinite(fgyro[o + k])))
                return set_err(ctx, CDX_NONFINITE, "cdx_set_bodies: non-finite body input");
            if (!(minv[o + k] > 0)) singular = true;
            hx[(size_t)b * kBodyStride + k] = v[o + k];
            hf[(size_t)b * kBodyStride + k] = fext[o + k];
            if (fgyro) hg[(size_t)b * kBodyStride + k] = fgyro[o + k];
        }
        if (d == 6)
            for (int k = 0; k < 3; ++k) hm[(size_t)b * kMinvStride + 1 + k] = minv[o + 3 + k];
        if (q) {
            const int qo = qoff[b], nqb = d == 6 ? 7 : 3;
            for (int k = 0; k < nqb; ++k) hq[(size_t)b * kBodyStride + k] = q[qo + k];
        }
    }
    const bool resized = nb != ctx->nb;
    ctx->nb = nb;
    ctx->nv = nv;
    ctx->dof.assign(dof, dof + nb);
    ctx->voff = voff;
    ctx->qoff = qoff;
    ctx->have_q = q != nullptr;
    ctx->singular_mass = singular;
    const size_t n8 = (size_t)nb * kBodyStride;
    size_t cap = ctx->body_cap;
    // all body buffers share one capacity
    if ((size_t)nb > ctx->body_cap || ctx->d_x == nullptr) {
        double** bufs8[] = {&ctx->d_x, &ctx->d_v, &ctx->d_fext, &ctx->d_fgyro, &ctx->d_q, &ctx->d_uprev, &ctx->d_ufree};
        for (double** p : bufs8) {
            if (*p) hipFree(*p);
            *p = nullptr;
            CDX_TRY(ensure_alloc(ctx, p, n8, nullptr, "body state"));
        }
        if (ctx->d_minv) hipFree(ctx->d_minv);
        ctx->d_minv = nullptr;
        CDX_TRY(ensure_alloc(ctx, &ctx->d_minv, (size_t)nb * kMinvStride, nullptr, "minv"));
        if (ctx->d_dof) hipFree(ctx->d_dof);
        ctx->d_dof = nullptr;
        CDX_TRY(ensure_alloc(ctx, &ctx->d_dof, (size_t)nb, nullptr, "dof"));
        if (ctx->d_incptr) hipFree(ctx->d_incptr);
        ctx->d_incptr = nullptr;
        CDX_TRY(ensure_alloc(ctx, &ctx->d_incptr, (size_t)nb + 1, nullptr, "incidence ptr"));
        for (int** p : {&ctx->d_voff, &ctx->d_qoff}) {
            if (*p) hipFree(*p);
            *p = nullptr;
            CDX_TRY(ensure_alloc(ctx, p, (size_t)nb + 1, nullptr, "stacked offsets"));
        }
        cap = nb;
    }
    ctx->body_cap = cap;
    if (nb > 0) {
        CDX_HIP(ctx, hipMemcpyAsync(ctx->d_v, hx.data(), n8 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        CDX_HIP(ctx, hipMemcpyAsync(ctx->d_x, hx.data(), n8 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        CDX_HIP(ctx, hipMemcpyAsync(ctx->d_fext, hf.data(), n8 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        CDX_HIP(ctx, hipMemcpyAsync(ctx->d_q, hq.data(), n8 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        CDX_HIP(ctx, hipMemcpyAsync(ctx->d_minv, hm.data(), (size_t)nb * kMinvStride * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        CDX_HIP(ctx, hipMemcpyAsync(ctx->d_dof, dof, (size_t)nb * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        CDX_HIP(ctx, hipMemcpyAsync(ctx->d_voff, ctx->voff.data(), (size_t)(nb + 1) * sizeof(int), hipMemcpyHostToDevice,
                                    ctx->stream));
        CDX_HIP(ctx, hipMemcpyAsync(ctx->d_qoff, ctx->qoff.data(), (size_t)(nb + 1) * sizeof(int), hipMemcpyHostToDevice,
                                    ctx->stream));
        if (fgyro) {
            CDX_HIP(ctx, hipMemcpyAsync(ctx->d_fgyro, hg.data(), n8 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        } else {
            CDX_LAUNCH(ctx, "gyro", k_gyro, grid_for(nb), kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_v, ctx->d_fgyro);
        }
        CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (resized) {
        ctx->incidence_valid = false;
        ctx->ranges_valid = false;
        ctx->coloring_valid = false;
    }
    ctx->nscenes = 0;
    ctx->njrows = ctx->nchains = ctx->njbodies = 0;  // joint rows depend on the body state: set them again
    ctx->joints_solved = false;
    ctx->solved = false;
    ctx->geo_ok = false;  // the lever arms depend on the positions
    ctx->geo_bad = -1;
    return CDX_OK;
}

int cdx_set_contacts(cdx_ctx* ctx, int nc, const int* body_a, const int* body_b, const double* mu, const double* frame,
                     const double* angA, const double* angB, const double* rest, const double* vbias, const double* warm) {
    if (!ctx) return CDX_BAD_ARG;
    if (nc < 0 || (nc > 0 && (!body_a || !body_b || !mu || !frame || !angA || !angB || !rest || !vbias || !warm)))
        return set_err(ctx, CDX_BAD_ARG, "cdx_set_contacts: null array");
    hipSetDevice(ctx->device);
    // the checks, the row count and the host copies of the ids and friction, over contiguous
    // chunks in host threads; the first failing contact (the lowest chunk's first) names the error
    const int T = host_threads(nc);
    struct Scan {
        int bad = -1, fault = 0, nrows = 0;
        bool rest = false, vbias = false;
    };
    std::vector<Scan> sc(T);
    std::vector<int> ia(nc), ib(nc);
    std::vector<double> m(nc);
    parallel_ranges(nc, T, [&](int t, int c0, int c1) {
        Scan& r = sc[t];
        for (int c = c0; c < c1; ++c) {
            const int f = contact_fault(ctx->nb, c, body_a, body_b, mu, frame, angA, angB, rest, vbias, warm);
            if (f) {
                r.bad = c;
                r.fault = f;
                return;
            }
            const bool fr = mu[c] > 0, rows = body_a[c] >= 0 || body_b[c] >= 0;
            m[c] = fr ? mu[c] : 0.0;
            ia[c] = rows ? body_a[c] : -1;
            ib[c] = rows ? body_b[c] : -1;
            if (!rows) continue;
            r.nrows += fr ? 3 : 1;
            for (int k = 0; k < (fr ? 3 : 1); ++k) {
                r.rest = r.rest || rest[3 * (size_t)c + k] != 0.0;
                r.vbias = r.vbias || vbias[3 * (size_t)c + k] != 0.0;
            }
        }
    });
    int nrows = 0;
    bool any_rest = false, any_vbias = false;
    for (const Scan& r : sc) {
        if (r.bad >= 0) {
            static const char* const msg[] = {"", "cdx_set_contacts: body index out of range (bodies must be set first)",
                                              "cdx_set_contacts: non-finite mu", "cdx_set_contacts: non-finite contact frame",
                                              "cdx_set_contacts: non-finite contact data"};
            return set_err(ctx, r.fault == 1 ? CDX_BAD_ARG : CDX_NONFINITE, msg[r.fault]);
        }
        nrows += r.nrows;
        any_rest = any_rest || r.rest;
        any_vbias = any_vbias || r.vbias;
    }
    if ((size_t)nc > ctx->contact_cap || !ctx->d_ida) {
        int** ibufs[] = {&ctx->d_ida, &ctx->d_idb};
        for (int** p : ibufs) {
            if (*p) hipFree(*p);
            *p = nullptr;
            CDX_TRY(ensure_alloc(ctx, p, (size_t)nc, nullptr, "contact ids"));
        }
        struct {
            double** p;
            size_t n;
        } dbufs[] = {{&ctx->d_J, (size_t)kJ * nc}, {&ctx->d_mu, (size_t)nc},        {&ctx->d_rest, 3 * (size_t)nc},
                     {&ctx->d_vb, 3 * (size_t)nc}, {&ctx->d_p, 3 * (size_t)nc},   {&ctx->d_R, 3 * (size_t)nc},
                     {&ctx->d_bias, 3 * (size_t)nc}, {&ctx->d_imp, 12 * (size_t)nc}};
        for (auto& d : dbufs) {
            if (*d.p) hipFree(*d.p);
            *d.p = nullptr;
            CDX_TRY(ensure_alloc(ctx, d.p, d.n, nullptr, "contact buffer"));
        }
        if (ctx->d_incidx) hipFree(ctx->d_incidx);
        ctx->d_incidx = nullptr;
        CDX_TRY(ensure_alloc(ctx, &ctx->d_incidx, 2 * (size_t)nc, nullptr, "incidence"));
        ctx->contact_cap = nc;
    }
    ctx->nc = nc;
    ctx->nrows = nrows;
    ctx->any_rest = any_rest;
    ctx->any_bias = any_rest || any_vbias;
    ctx->h_a = std::move(ia);
    ctx->h_b = std::move(ib);
    ctx->h_mu = std::move(m);
    if (nc > 0) {
        // the caller's arrays as they are, through pinned staging in one copy; the device
        // transposes and masks them (k_contacts_soa)
        const size_t n = (size_t)nc, nd = 37 * n;
        CDX_TRY(ensure_stage(ctx, nd + n));
        const struct {
            const double* src;
            size_t per;  // doubles per contact
        } parts[] = {{frame, 9}, {angA, 9}, {angB, 9}, {rest, 3}, {vbias, 3}, {warm, 3}, {mu, 1}};
        int* hi = reinterpret_cast<int*>(ctx->h_stage + nd);
        parallel_ranges(nc, T, [&](int, int c0, int c1) {
            double* h = ctx->h_stage;
            for (const auto& part : parts) {
                std::memcpy(h + part.per * c0, part.src + part.per * c0, part.per * (size_t)(c1 - c0) * sizeof(double));
                h += part.per * n;
            }
            std::memcpy(hi + c0, body_a + c0, (size_t)(c1 - c0) * sizeof(int));
            std::memcpy(hi + n + c0, body_b + c0, (size_t)(c1 - c0) * sizeof(int));
        });
        CDX_HIP(ctx, hipMemcpyAsync(ctx->d_stage, ctx->h_stage, (nd + n) * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        CDX_LAUNCH(ctx, "contacts_soa", k_contacts_soa, grid_for(nc), kBlock, nc, ctx->d_dof, ctx->d_stage,
                   reinterpret_cast<const int*>(ctx->d_stage + nd), ctx->d_ida, ctx->d_idb, ctx->d_J, ctx->d_mu,
                   ctx->d_rest, ctx->d_vb, ctx->d_p);
    }
    CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->incidence_valid = false;
    ctx->ranges_valid = false;
    ctx->coloring_valid = false;
    ctx->nscenes = 0;
    ctx->solved = false;
    ctx->geo_ok = false;
    ctx->geo_bad = -1;
    return build_incidence(ctx);
}

int cdx_get_velocities(cdx_ctx* ctx, double* v_out) {
    if (!ctx || (!v_out && ctx->nv > 0)) return CDX_BAD_ARG;
    if (!ctx->solved) return set_err(ctx, CDX_NO_STATE, "cdx_get_velocities: no solve yet");
    if (ctx->nv == 0) return CDX_OK;
    CDX_TRY(ensure_stage(ctx, (size_t)ctx->nv));
    CDX_LAUNCH(ctx, "pack_velocities", k_pack_stacked, grid_for(ctx->nb * kBodyStride), kBlock, ctx->nb, ctx->d_dof,
               ctx->d_voff, 0, ctx->d_x, ctx->d_stage);
    return download_stage(ctx, (size_t)ctx->nv, v_out);
}

int cdx_get_impulses(cdx_ctx* ctx, double* imp_out) {
    if (!ctx || (!imp_out && ctx->nc > 0)) return CDX_BAD_ARG;
    if (!ctx->solved) return set_err(ctx, CDX_NO_STATE, "cdx_get_impulses: no solve yet");
    const int nc = ctx->nc;
    if (nc == 0) return CDX_OK;
    CDX_TRY(ensure_stage(ctx, 3 * (size_t)nc));
    CDX_LAUNCH(ctx, "pack_impulses", k_pack_impulses, grid_for(3 * nc), kBlock, nc, ctx->d_p, ctx->d_stage);
    return download_stage(ctx, 3 * (size_t)nc, imp_out);
}

int cdx_get_positions(cdx_ctx* ctx, double* q_out) {
    if (!ctx || (!q_out && ctx->nb > 0)) return CDX_BAD_ARG;
    if (!ctx->have_q) return set_err(ctx, CDX_NO_STATE, "cdx_get_positions: no positions were set");
    const size_t nq = (size_t)ctx->qoff[ctx->nb];
    if (nq == 0) return CDX_OK;
    CDX_TRY(ensure_stage(ctx, nq));
    CDX_LAUNCH(ctx, "pack_positions", k_pack_stacked, grid_for(ctx->nb * kBodyStride), kBlock, ctx->nb, ctx->d_dof,
               ctx->d_qoff, 1, ctx->d_q, ctx->d_stage);
    return download_stage(ctx, nq, q_out);
}

int cdx_nonfinite_bodies(cdx_ctx* ctx, int* count) {
    if (!ctx || !count) return CDX_BAD_ARG;
    if (!ctx->solved) return set_err(ctx, CDX_NO_STATE, "cdx_nonfinite_bodies: no solve yet");
    *count = 0;
    if (ctx->nb == 0) return CDX_OK;
    const int grid = (ctx->nb + kBlock - 1) / kBlock;
    CDX_TRY(ensure_stage(ctx, (size_t)grid));
    CDX_LAUNCH(ctx, "nonfinite_bodies", k_nonfinite_bodies, grid, kBlock, ctx->nb, ctx->d_dof, ctx->d_x, ctx->d_q,
               ctx->have_q ? 1 : 0, ctx->d_stage);
    CDX_TRY(download_stage(ctx, (size_t)grid, nullptr));
    double n = 0;
    for (int k = 0; k < grid; ++k) n += ctx->h_stage[k];
    *count = (int)n;
    return CDX_OK;
}

int cdx_moreau_predrift(cdx_ctx* ctx, double dt, double theta) {
    if (!ctx) return CDX_BAD_ARG;
    if (!ctx->have_q) return set_err(ctx, CDX_NO_STATE, "cdx_moreau_predrift: no positions were set");
    const double h = (1.0 - theta) * dt;
    if (h != 0.0 && ctx->nb > 0)
        CDX_LAUNCH(ctx, "moreau_drift", k_position_update, grid_for(ctx->nb), kBlock, ctx->nb, ctx->d_dof, ctx->d_q, ctx->d_v,
                   ctx->d_v, h, 0);
    if (h != 0.0) {
        ctx->geo_ok = false;
        ctx->geo_bad = -1;
    }
    CDX_HIP(ctx, hipGetLastError());
    return CDX_OK;
}

int cdx_moreau_update(cdx_ctx* ctx, double dt, double theta) {
    if (!ctx) return CDX_BAD_ARG;
    if (!ctx->have_q) return set_err(ctx, CDX_NO_STATE, "cdx_moreau_update: no positions were set");
    if (!ctx->solved) return set_err(ctx, CDX_NO_STATE, "cdx_moreau_update: no solve yet");
    if (ctx->nb > 0) {
        CDX_LAUNCH(ctx, "moreau_update", k_position_update, grid_for(ctx->nb), kBlock, ctx->nb, ctx->d_dof, ctx->d_q, ctx->d_v,
                   ctx->d_x, theta * dt, 1);
        CDX_LAUNCH(ctx, "gyro", k_gyro, grid_for(ctx->nb), kBlock, ctx->nb, ctx->d_dof, ctx->d_minv, ctx->d_v, ctx->d_fgyro);
    }
    ctx->geo_ok = false;
    ctx->geo_bad = -1;
    CDX_HIP(ctx, hipGetLastError());
    return CDX_OK;
}

int cdx_set_profiling(cdx_ctx* ctx, int enable) {
    if (!ctx) return CDX_BAD_ARG;
    if (!enable) harvest_events(ctx);
    ctx->profiling = enable != 0;
    return CDX_OK;
}

int cdx_kernel_stat(cdx_ctx* ctx, int index, char* name_buf, int buf_len, int* calls, double* total_ms) {
    if (!ctx || index < 0) return CDX_BAD_ARG;
    harvest_events(ctx);
    if (index >= (int)ctx->stats.size()) return CDX_BAD_ARG;
    auto it = ctx->stats.begin();
    std::advance(it, index);
    if (name_buf && buf_len > 0) {
        std::strncpy(name_buf, it->first.c_str(), (size_t)buf_len - 1);
        name_buf[buf_len - 1] = '\0';
    }
    if (calls) *calls = it->second.calls;
    if (total_ms) *total_ms = it->second.ms;
    return CDX_OK;
}

int cdx_set_graphs(cdx_ctx* ctx, int enable) {
    if (!ctx) return CDX_BAD_ARG;
    ctx->graphs = enable != 0;
    return CDX_OK;
}

int cdx_kernel_time(cdx_ctx* ctx, const char* kernel, int* calls, double* total_ms) {
    if (!ctx || !kernel) return CDX_BAD_ARG;
    harvest_events(ctx);
    auto it = ctx->stats.find(kernel);
    if (calls) *calls = it == ctx->stats.end() ? 0 : it->second.calls;
    if (total_ms) *total_ms = it == ctx->stats.end() ? 0.0 : it->second.ms;
    if (it != ctx->stats.end()) ctx->stats.erase(it);
    return CDX_OK;
}

}  // extern "C"

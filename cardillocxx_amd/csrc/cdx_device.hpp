// cdx_device.hpp -- device helpers shared by the solve kernels, and the host sweep driver.
#pragma once

#include <algorithm>
#include <deque>

#include "cdx_common.hpp"

namespace cdx {

constexpr int kLanesPerContact = 4;  // lanes per contact in the contact passes: one per row, one idle

// Exchanges inside aligned groups of lanes through DPP (a VALU modifier; ds_bpermute would go
// through the LDS crossbar).  quad_perm<CTRL>: lane j of each group of 4 reads lane
// (CTRL >> 2j) & 3 of its group.  half_mirror: lane j of each group of 8 reads lane 7 - j.
// Every lane of the group must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_double(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int I>
__device__ __forceinline__ double quad_bcast(double v) {  // lane I of the group of 4
    return dpp_double<I * 0x55>(v);
}
__device__ __forceinline__ double quad_xor1(double v) { return dpp_double<0xB1>(v); }  // quad_perm [1,0,3,2]
__device__ __forceinline__ double quad_xor2(double v) { return dpp_double<0x4E>(v); }  // quad_perm [2,3,0,1]
__device__ __forceinline__ double half_mirror(double v) { return dpp_double<0x141>(v); }

// Sum over a group of 4 lanes, the same value in every lane: (v0 + v1) + (v2 + v3).
__device__ __forceinline__ double quad_sum(double v) {
    v += quad_xor1(v);
    return v + quad_xor2(v);
}

// Sum over a group of 8 lanes, the same value in every lane: the two quad sums added.
__device__ __forceinline__ double oct_sum(double v) {
    v = quad_sum(v);
    return v + half_mirror(v);
}

// Sum of v[i] over i = threadIdx.x + j kBlock (j = 0, 1, ...) with eight loads in flight per step
// (a one-workgroup reduction over thousands of partials otherwise waits on one load at a time); the
// accumulators meet in a fixed order, so the sum repeats bit for bit.
__device__ inline double strided_sum(const double* __restrict__ v, int n) {
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int i = threadIdx.x;
    for (; i + 7 * kBlock < n; i += 8 * kBlock)
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] += v[i + u * kBlock];
#pragma unroll
    for (int u = 0; u < 7; ++u)
        if (i + u * kBlock < n) a[u] += v[i + u * kBlock];
    return ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
}

__device__ inline void load_body(const double* __restrict__ x, int b, double u[6]) {
    if (b < 0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) u[k] = 0.0;
        return;
    }
    const double2* p = reinterpret_cast<const double2*>(x + (size_t)b * kBodyStride);
    const double2 a0 = p[0], a1 = p[1], a2 = p[2];
    u[0] = a0.x;
    u[1] = a0.y;
    u[2] = a1.x;
    u[3] = a1.y;
    u[4] = a2.x;
    u[5] = a2.y;
}

__device__ inline void store6(double* __restrict__ dst, const double v[6]) {
    double2* p = reinterpret_cast<double2*>(dst);
    p[0] = make_double2(v[0], v[1]);
    p[1] = make_double2(v[2], v[3]);
    p[2] = make_double2(v[4], v[5]);
}

// Row k of W_c times the stacked twist: side A carries sign -1, side B +1
// (dynamics_assembler.cpp:295-296; contact_jacobian.hpp:17-23).
__device__ inline double row_dot(const double d[3], const double aA[3], const double aB[3], const double ua[6],
                                 const double ub[6]) {
    const double sA = d[0] * ua[0] + d[1] * ua[1] + d[2] * ua[2] + aA[0] * ua[3] + aA[1] * ua[4] + aA[2] * ua[5];
    const double sB = d[0] * ub[0] + d[1] * ub[1] + d[2] * ub[2] + aB[0] * ub[3] + aB[1] * ub[4] + aB[2] * ub[5];
    return sB - sA;
}

struct ContactJ {
    double d[3][3], aA[3][3], aB[3][3];
};

__device__ inline void load_J(const double* __restrict__ J, int nc, int c, int nr, ContactJ& j) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= nr) break;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            j.d[k][m] = J[(size_t)(3 * k + m) * nc + c];
            j.aA[k][m] = J[(size_t)(9 + 3 * k + m) * nc + c];
            j.aB[k][m] = J[(size_t)(18 + 3 * k + m) * nc + c];
        }
    }
}

// Side impulses J_A^T dp = -[sum d_k dp_k ; sum angA_k dp_k], J_B^T dp = +[... ; sum angB_k dp_k].
__device__ inline void write_side_impulses(double* __restrict__ imp, int c, int a, int b, int nr, const ContactJ& j,
                                           const double dp[3]) {
    double lin[3] = {0, 0, 0}, angA[3] = {0, 0, 0}, angB[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= nr) break;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            lin[m] += j.d[k][m] * dp[k];
            angA[m] += j.aA[k][m] * dp[k];
            angB[m] += j.aB[k][m] * dp[k];
        }
    }
    if (a >= 0) {
        const double va[6] = {-lin[0], -lin[1], -lin[2], -angA[0], -angA[1], -angA[2]};
        store6(imp + (size_t)(2 * c) * 6, va);
    }
    if (b >= 0) {
        const double vb[6] = {lin[0], lin[1], lin[2], angB[0], angB[1], angB[2]};
        store6(imp + (size_t)(2 * c + 1) * 6, vb);
    }
}

// Compact side impulses of the projected-Jacobi sweep: one 9-double record per contact,
// [lin | angA | angB] at imp9 + 9c with lin = sum_k d_k dp_k and angX = sum_k aX_k dp_k.  Side A
// receives -[lin ; angA] and side B +[lin ; angB] (W row signs A -1, B +1), so the linear part is
// stored once; a static side's angular part is not written.
__device__ inline void write_side_impulses9(double* __restrict__ imp9, int c, int a, int b, int nr, const ContactJ& j,
                                            const double dp[3]) {
    double lin[3] = {0, 0, 0}, angA[3] = {0, 0, 0}, angB[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= nr) break;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            lin[m] += j.d[k][m] * dp[k];
            angA[m] += j.aA[k][m] * dp[k];
            angB[m] += j.aB[k][m] * dp[k];
        }
    }
    double* r = imp9 + 9 * (size_t)c;
    if (a < 0 && b < 0) return;
#pragma unroll
    for (int m = 0; m < 3; ++m) r[m] = lin[m];
    if (a >= 0)
#pragma unroll
        for (int m = 0; m < 3; ++m) r[3 + m] = angA[m];
    if (b >= 0)
#pragma unroll
        for (int m = 0; m < 3; ++m) r[6 + m] = angB[m];
}

// acc += J_side^T dp of incidence e = 2c + side from the compact records.  Adding -v and
// subtracting v round identically, so the sums equal those over explicitly negated records.
__device__ inline void gather_side9(const double* __restrict__ imp9, int e, double acc[6]) {
    const double* r = imp9 + 9 * (size_t)(e >> 1);
    if (e & 1) {
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            acc[m] += r[m];
            acc[3 + m] += r[6 + m];
        }
    } else {
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            acc[m] -= r[m];
            acc[3 + m] -= r[3 + m];
        }
    }
}

// f_gyro = -omega x (I omega) of a rigid body from its body-frame omega and inverse principal
// inertia (world.cpp:152-173; I = 1 / I^-1 as rebuildMass_ inverts it, dynamics_assembler.cpp:56-60).
__device__ inline void gyro_force3(const double* __restrict__ w, const double* __restrict__ iinv, double* __restrict__ o) {
    double I[3], Iw[3];
    for (int k = 0; k < 3; ++k) I[k] = iinv[k] > 0 ? 1.0 / iinv[k] : 0.0;
    for (int k = 0; k < 3; ++k) Iw[k] = I[k] * w[k];
    o[0] = -(w[1] * Iw[2] - w[2] * Iw[1]);
    o[1] = -(w[2] * Iw[0] - w[0] * Iw[2]);
    o[2] = -(w[0] * Iw[1] - w[1] * Iw[0]);
}

// Inverse mass of a body as [m^-1, Ixx^-1, Iyy^-1, Izz^-1]; zeros for a static side.
__device__ inline double4 load_minv4(const double* __restrict__ minv, int b) {
    if (b < 0) return make_double4(0.0, 0.0, 0.0, 0.0);
    return *reinterpret_cast<const double4*>(minv + (size_t)b * kMinvStride);
}

// D = sum_j W_kj^2 M^-1_j of one row (rdiag_sparse, projected_jacobi.cpp:57-69): summed side by
// side in column order (lower body index first), as the row-major compressed W stores the row.
// Without contraction: the same operations as the oracle's delassus_diag.
__device__ inline double row_D(const double d[3], const double aA[3], const double aB[3], double4 mA, double4 mB, int a,
                               int b) {
#pragma clang fp contract(off)
    double DA = 0, DB = 0;
    DA += d[0] * d[0] * mA.x;
    DA += d[1] * d[1] * mA.x;
    DA += d[2] * d[2] * mA.x;
    DA += aA[0] * aA[0] * mA.y;
    DA += aA[1] * aA[1] * mA.z;
    DA += aA[2] * aA[2] * mA.w;
    DB += d[0] * d[0] * mB.x;
    DB += d[1] * d[1] * mB.x;
    DB += d[2] * d[2] * mB.x;
    DB += aB[0] * aB[0] * mB.y;
    DB += aB[1] * aB[1] * mB.z;
    DB += aB[2] * aB[2] * mB.w;
    return (a >= 0 && b >= 0 && b < a) ? DB + DA : DA + DB;
}

// Runs the sweep loop of a solve: enqueues sweeps in chunks (aligned to the residual cadence) and
// keeps at most `window` chunks in flight, polling the pinned done flag.  Enqueued sweeps after
// convergence return on their first instruction.  On every exit path, errors included, the
// stream is drained and the chunk events are destroyed before returning, so a caller may free
// what the queued work uses (a graph exec, buffers) as soon as this returns.
template <typename SweepFn>
int run_sweeps(cdx_ctx* ctx, int max_iter, int chunk, int window, SweepFn&& sweep, int* iters_out) {
    struct Drain {
        cdx_ctx* ctx;
        std::deque<hipEvent_t> inflight;
        ~Drain() {
            hipStreamSynchronize(ctx->stream);
            for (hipEvent_t e : inflight) hipEventDestroy(e);
        }
    } drain{ctx, {}};
    auto& inflight = drain.inflight;
    int it = 0;
    bool done = false;
    while (it < max_iter && !done) {
        const int end = std::min(max_iter, it + chunk);
        for (; it < end; ++it) CDX_TRY(sweep(it));
        hipEvent_t ev;
        CDX_HIP(ctx, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        inflight.push_back(ev);
        CDX_HIP(ctx, hipEventRecord(ev, ctx->stream));
        if ((int)inflight.size() >= window) {
            hipEvent_t e = inflight.front();
            inflight.pop_front();
            const hipError_t se = hipEventSynchronize(e);
            hipEventDestroy(e);
            CDX_HIP(ctx, se);
            if (__atomic_load_n(&ctx->h_flag[0], __ATOMIC_ACQUIRE)) done = true;
        }
    }
    CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int ctrl[3] = {0, 0, 0};
    CDX_HIP(ctx, hipMemcpy(ctrl, ctx->d_ctrl, sizeof(ctrl), hipMemcpyDeviceToHost));
    *iters_out = ctrl[0] ? ctrl[1] : max_iter;
    return ctrl[0] ? ctrl[2] : CDX_OK;
}


// Like run_sweeps, with every full chunk of `chunk` sweeps captured once as a HIP graph and
// replayed: the host submits one graph per chunk instead of several kernels per sweep.
// launch(it, pos) enqueues sweep `it` at position `pos` of its chunk; inside the capture it is
// called with it = -1, so kernels that need the sweep index read the device sweep counter
// (k_residual_finalize).  The tail shorter than a chunk is launched directly.  Without graphs
// (profiling events, recording, or a single chunk) every sweep is launched directly.
template <typename LaunchFn>
int run_sweeps_graphed(cdx_ctx* ctx, int max_iter, int chunk, bool allow_graph, LaunchFn&& launch, int* iters_out) {
    const int nchunks = max_iter / chunk;
    if (!allow_graph || !ctx->graphs || ctx->profiling || nchunks < 2)
        return run_sweeps(ctx, max_iter, chunk, 4, [&](int it) { return launch(it, it % chunk); }, iters_out);
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    CDX_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int cst = CDX_OK;
    for (int i = 0; i < chunk && cst == CDX_OK; ++i) cst = launch(-1, i);
    hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
    if (cst) {
        if (graph) hipGraphDestroy(graph);
        return cst;
    }
    CDX_HIP(ctx, ce);
    ce = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    CDX_HIP(ctx, ce);
    auto step = [&](int it) -> int {
        if (it / chunk < nchunks) {
            if (it % chunk != 0) return CDX_OK;  // the whole chunk goes with its first sweep
            return hip_check(ctx, hipGraphLaunch(exec, ctx->stream), "hipGraphLaunch");
        }
        return launch(it, it % chunk);
    };
    const int st = run_sweeps(ctx, max_iter, chunk, 4, step, iters_out);
    hipGraphExecDestroy(exec);
    return st;
}

}  // namespace cdx

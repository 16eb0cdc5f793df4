// cdx_batch.hip -- thousands of independent scenes solved by projected Jacobi in one launch.
//
// The reference runs one PhysicsEngine per scene (physics_engine.hpp:215-221), so a batch is a
// sequence of independent ProjectedJacobiSolver::solve calls.  Here every scene is a contiguous
// range of the context's body and contact arrays (cdx_set_scenes) and one launch solves them all:
// a workgroup of B threads hosts G scenes (B/G threads each, a multiple of the wave size) and
// runs each scene's whole solve -- setup, sweeps, residual every `check_every` sweeps, own
// convergence test -- with workgroup barriers between the contact and body phases.  The
// per-contact and per-body arithmetic is the same device code as the grid-wide solve
// (cdx_pj_device.hpp), so every scene's result equals a standalone solve of that scene.
// Scenes are launched largest first (longest-processing-time order) so the hardware's dynamic
// workgroup dispatch balances uneven scene sizes.  (B, G) comes from the autotuner below.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>

#include "cdx_joints_device.hpp"
#include "cdx_pj_device.hpp"

namespace cdx {

constexpr int kMaxBatchBlock = 512;

// Joint rows of the batch (null sc_off: none).  The chains of scene s are sc_list[sc_off[s] ..
// sc_off[s+1]); each is factorised and solved by one wave of the scene's group, as the grid-wide
// kernels do with one workgroup per chain.
struct BatchJoints {
    const int* sc_off;
    const int* sc_list;
    const int* jc_off;
    const int* jcb_off;
    const int* jb_off;
    const int* jl_off;
    const int* jb_body;
    const int* jr_kind;
    const int* jr_blk;
    const double* jr;
    const double* jw;
    int nrow;
    double* rlam;
    double* Lm;
    double* Dinv;
    const int* jbody;
    const int* jbody_blk;
    const int* jr_info;
    const int* jbody_info;
    const int* jslot;
    double* tr;
    double* z;
    double* y;
    double* lam;
    double theta;
    double beta;
    int wave_lds;  // doubles of dynamic LDS per wave (>= kMinWaveLds): factorisation, then the chain window
};

// Joints = false compiles none of the chain code (no scratch for its per-lane block matrices).
template <bool Joints>
__global__ void __launch_bounds__(kMaxBatchBlock) k_batch_pj(
    int ns, int groups, const int* __restrict__ order, const int* __restrict__ boff, const int* __restrict__ coff,
    const int* __restrict__ snv, const int* __restrict__ srows, int nc, const int* __restrict__ ida,
    const int* __restrict__ idb, const double* __restrict__ J, const double* __restrict__ mu,
    const double* __restrict__ rest, const double* __restrict__ vb, double* __restrict__ R, double* __restrict__ bias,
    double* __restrict__ p, double* __restrict__ imp, const int* __restrict__ dof, const double* __restrict__ minv,
    const double* __restrict__ v, const double* __restrict__ fext, const double* __restrict__ fgyro,
    const int* __restrict__ incptr, const int* __restrict__ incidx, double* __restrict__ x, double dt, double alpha,
    int use_bias, int warmstart, int max_iter, int check_every, double tol_rel, double tol_abs, int* __restrict__ iters_out,
    double* __restrict__ res_out, double* __restrict__ Rb, BatchJoints ja) {
    __shared__ double wsum[kMaxBatchBlock / 64];
    __shared__ int gdone[kMaxBatchBlock / 64];
    __shared__ int gfail[kMaxBatchBlock / 64];
    __shared__ int all_done;
    // per wave (dynamic, ja.wave_lds doubles each): the chain factorisation, then the staging
    // window of the chain solve
    extern __shared__ double jdyn[];
    double* jlds = jdyn + (size_t)(threadIdx.x >> 6) * (Joints ? ja.wave_lds : 0);
    const int capR = kSolveLdsR, capL = Joints ? ja.wave_lds - kSolveLdsR : 0;
    const int gsize = blockDim.x / groups;
    const int g = threadIdx.x / gsize, lt = threadIdx.x - g * gsize;
    const int slot = blockIdx.x * groups + g;
    const int s = slot < ns ? order[slot] : -1;
    const int b0 = s >= 0 ? boff[s] : 0, b1 = s >= 0 ? boff[s + 1] : 0;
    const int c0 = s >= 0 ? coff[s] : 0, c1 = s >= 0 ? coff[s + 1] : 0;
    const bool rows = s >= 0 && srows[s] > 0;
    // setup: R, b, p0 and the side impulses of W^T p0, then x0 (also for a scene without rows,
    // whose static-static contacts get R = 0 and p = 0, as in the grid-wide setup)
    if (s >= 0)
        for (int c = c0 + lt; c < c1; c += gsize)
            pj_setup_contact(c, nc, ida, idb, J, mu, rest, vb, minv, v, alpha, warmstart, R, bias, p, imp, Rb);
    __syncthreads();
    for (int b = b0 + lt; b < b1; b += gsize)
        pj_setup_body(b, dof, minv, v, fext, fgyro, incptr, incidx, imp, dt, rows ? 1 : 0, x, ja.jslot, ja.tr);
    if (lt == 0) gfail[g] = 0;
    const int wv = lt >> 6, nwv = gsize >> 6, lane = lt & 63;  // the group's waves take its chains
    if constexpr (Joints) {  // factor the scene's chains, then x0 and Lambda_0 of its jointed bodies
        __syncthreads();
        if (s >= 0)
            for (int q = ja.sc_off[s] + wv; q < ja.sc_off[s + 1]; q += nwv) {
                const int ch = ja.sc_list[q];
                if (joint_factor_chain(ch, lane, ja.jc_off, ja.jb_off, ja.jb_body, ja.jr_kind, ja.jr_blk, ja.jr, minv, v, dt,
                                       ja.theta, ja.beta, ja.rlam, ja.jl_off, ja.Lm, ja.Dinv, ja.nrow, jlds))
                    gfail[g] = 1;
                else
                    joint_solve_chain<true>(ch, lane, 1, 0, ja.jc_off, ja.jcb_off, ja.jb_off, ja.jl_off, ja.jr_info, ja.jr, ja.jw, ja.nrow, ja.rlam, ja.Lm,
                                      ja.Dinv, ja.jbody_info, minv, ja.tr, ja.z, ja.y, ja.lam, x, tol_rel, tol_abs,
                                      jlds, capL, capR);
            }
    }
    __syncthreads();
    const bool failed = gfail[g] != 0;
    if (lt == 0) gdone[g] = (rows && !failed) ? 0 : 1;
    bool done = !rows || failed;
    int iters = 0;
    double res = rows ? NAN : NAN;
    __syncthreads();
    const int waves_per_group = gsize >> 6;
    for (int it = 0; it < max_iter; ++it) {
        if (!done)
            for (int c = c0 + lt; c < c1; c += gsize)
                pj_contact_sweep(c, nc, ida, idb, J, R, bias, use_bias, mu, p, x, imp, Rb);
        __syncthreads();
        const int check = (it % check_every) == 0;
        double e2 = 0;
        if (!done)
            for (int b = b0 + lt; b < b1; b += gsize)
                e2 += pj_body_sweep(b, dof, minv, incptr, incidx, imp, x, check, tol_rel, tol_abs, ja.jslot, ja.tr);
        if constexpr (Joints) {  // the chains apply the jointed bodies' W^T dp through S^-1
            __syncthreads();
            if (!done) {
                for (int q = ja.sc_off[s] + wv; q < ja.sc_off[s + 1]; q += nwv)
                    e2 += joint_solve_chain<true>(ja.sc_list[q], lane, 0, check, ja.jc_off, ja.jcb_off, ja.jb_off, ja.jl_off, ja.jr_info,
                                            ja.jr, ja.jw, ja.nrow, ja.rlam, ja.Lm, ja.Dinv, ja.jbody_info, minv, ja.tr, ja.z, ja.y, ja.lam, x,
                                            tol_rel, tol_abs, jlds, capL, capR);
            }
        }
        if (check) {
            for (int off = 32; off > 0; off >>= 1) e2 += __shfl_xor(e2, off, 64);
            if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = e2;
        }
        __syncthreads();
        if (check && !done) {
            double t = 0;
            for (int w = 0; w < waves_per_group; ++w) t += wsum[g * waves_per_group + w];
            res = 1.0 / sqrt((double)snv[s]) * sqrt(t);
            if (res <= 1.0) {
                done = true;
                iters = it + 1;
            }
        }
        if (!done) iters = it + 1;
        if (lt == 0) gdone[g] = done ? 1 : 0;
        if (threadIdx.x == 0) all_done = 0;
        __syncthreads();
        if (threadIdx.x == 0) {
            int a = 1;
            for (int k = 0; k < groups; ++k) a &= gdone[k];
            all_done = a;
        }
        __syncthreads();
        if (all_done) break;
    }
    if (s >= 0 && lt == 0) {
        iters_out[s] = failed ? -1 : (rows ? iters : 0);  // -1: a singular joint block
        res_out[s] = res;
    }
}

// ---- Chebyshev-accelerated standard loop per scene (chebyshev_loop with estimateSpectralRadius,
// projected_jacobi.cpp:145-228): each scene's group estimates its own spectral radius (8
// normalised power iterations of the linear sweep map from ones/sqrt(nv), rho = min(lambda,
// 0.995), 0 when the norm vanishes) and runs its own w recurrence, residual every sweep.  As in
// the grid-wide loop the extrapolation is fused into the next sweep: X holds the swept iterate,
// XK = xk, XK1 the previous input; the passes form the input u = w (X - XK) + XK (u = X on the
// first sweep and after a fallback), rotate XK <- XK1, XK1 <- u and sweep from u.  A converged
// scene returns its sweep input (XK1, PK1); after max_iter sweeps the pending update.

// linear contact pass of the power iteration: dp = -R (W v), no projection, no bias
__device__ inline void cheb_lin_contact(int c, int nc, const int* __restrict__ ida, const int* __restrict__ idb,
                                        const double* __restrict__ J, const double* __restrict__ R,
                                        const double* __restrict__ mu, const double* __restrict__ v,
                                        double* __restrict__ imp, const double* __restrict__ Rb) {
    const int a = ida[c], b = idb[c];
    if (a < 0 && b < 0) return;
    const int nr = mu[c] > 0 ? 3 : 1;
    ContactJ j;
    load_J(J, nc, c, nr, j);
    double ua[6], ub[6];
    load_body(v, a, ua);
    load_body(v, b, ub);
    double dp[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= nr) break;
        const double g = row_dot(j.d[k], j.aA[k], j.aB[k], ua, ub);
        dp[k] = Rb ? g : -(R[(size_t)k * nc + c] * g);
    }
    if (Rb) {  // -Rb (W v)
        const double zero[3] = {0, 0, 0};
        block_apply_rows(c, nc, nr, Rb, zero, dp);
    }
    write_side_impulses9(imp, c, a, b, nr, j, dp);
}

__device__ inline void cheb_input(const double* __restrict__ X, const double* __restrict__ XK, int b, double w,
                                  bool take, double u[6]) {
    load_body(X, b, u);
    if (take || b < 0) return;
    double k6[6];
    load_body(XK, b, k6);
#pragma unroll
    for (int m = 0; m < 6; ++m) u[m] = cheb_extrap(w, u[m], k6[m]);
}

// contact pass of a Chebyshev sweep (the grid-wide k_pjc_contacts4, one thread per contact)
__device__ inline void pjc_contact_sweep(int c, int nc, const int* __restrict__ ida, const int* __restrict__ idb,
                                         const double* __restrict__ J, const double* __restrict__ R,
                                         const double* __restrict__ bias, int use_bias, const double* __restrict__ mu,
                                         double* __restrict__ P, double* __restrict__ PK, double* __restrict__ PK1,
                                         const double* __restrict__ X, const double* __restrict__ XK, double w,
                                         bool take, double* __restrict__ imp, const double* __restrict__ Rb) {
    const int a = ida[c], b = idb[c];
    if (a < 0 && b < 0) return;
    const double muc = mu[c];
    const int nr = muc > 0 ? 3 : 1;
    ContactJ j;
    load_J(J, nc, c, nr, j);
    double ua[6], ub[6];
    cheb_input(X, XK, a, w, take, ua);
    cheb_input(X, XK, b, w, take, ub);
    double pin[3] = {0, 0, 0}, pn[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= nr) break;
        const size_t r = (size_t)k * nc + c;
        const double ps = P[r];
        pin[k] = take ? ps : cheb_extrap(w, ps, PK[r]);
        PK[r] = PK1[r];
        PK1[r] = pin[k];
        const double g = row_dot(j.d[k], j.aA[k], j.aB[k], ua, ub);
        pn[k] = Rb ? g + (use_bias ? bias[r] : 0.0) : pin[k] - (R[r] * g + (use_bias ? bias[r] : 0.0));
    }
    if (Rb) block_apply_rows(c, nc, nr, Rb, pin, pn);
    pn[0] = fmax(pn[0], 0.0);
    if (nr == 3) {
        const double t1 = pn[1], t2 = pn[2];
        const double tn2 = t1 * t1 + t2 * t2;
        if (tn2 > 0) {
            const double sc = fmin(1.0, (muc * pn[0]) / sqrt(tn2));
            pn[1] = sc * t1;
            pn[2] = sc * t2;
        }
    }
    double dp[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= nr) break;
        dp[k] = pn[k] - pin[k];
        P[(size_t)k * nc + c] = pn[k];
    }
    write_side_impulses9(imp, c, a, b, nr, j, dp);
}

__global__ void __launch_bounds__(kMaxBatchBlock) k_batch_pjc(
    int ns, int groups, const int* __restrict__ order, const int* __restrict__ boff, const int* __restrict__ coff,
    const int* __restrict__ snv, const int* __restrict__ srows, int nc, const int* __restrict__ ida,
    const int* __restrict__ idb, const double* __restrict__ J, const double* __restrict__ mu,
    const double* __restrict__ rest, const double* __restrict__ vb, double* __restrict__ R, double* __restrict__ bias,
    double* __restrict__ p, double* __restrict__ imp, const int* __restrict__ dof, const double* __restrict__ minv,
    const double* __restrict__ v, const double* __restrict__ fext, const double* __restrict__ fgyro,
    const int* __restrict__ incptr, const int* __restrict__ incidx, double* __restrict__ x, double* __restrict__ XK,
    double* __restrict__ XK1, double* __restrict__ PK, double* __restrict__ PK1, double dt, double alpha, int use_bias,
    int warmstart, int max_iter, double tol_rel, double tol_abs, int* __restrict__ iters_out,
    double* __restrict__ res_out, double* __restrict__ Rb) {
    __shared__ double wsum[kMaxBatchBlock / 64];
    __shared__ double wbad[kMaxBatchBlock / 64];
    __shared__ int gdone[kMaxBatchBlock / 64];
    __shared__ int all_done;
    const int gsize = blockDim.x / groups;
    const int g = threadIdx.x / gsize, lt = threadIdx.x - g * gsize;
    const int slot = blockIdx.x * groups + g;
    const int s = slot < ns ? order[slot] : -1;
    const int b0 = s >= 0 ? boff[s] : 0, b1 = s >= 0 ? boff[s + 1] : 0;
    const int c0 = s >= 0 ? coff[s] : 0, c1 = s >= 0 ? coff[s + 1] : 0;
    const bool rows = s >= 0 && srows[s] > 0;
    const int wpg = gsize >> 6;
    // the group's sum of one value per thread, in every thread of the group (two barriers)
    auto group_sum2 = [&](double a, double b, double* sa, double* sb) {
        for (int off = 32; off > 0; off >>= 1) {
            a += __shfl_xor(a, off, 64);
            b += __shfl_xor(b, off, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            wsum[threadIdx.x >> 6] = a;
            wbad[threadIdx.x >> 6] = b;
        }
        __syncthreads();
        double ta = 0, tb = 0;
        for (int w = 0; w < wpg; ++w) {
            ta += wsum[g * wpg + w];
            tb += wbad[g * wpg + w];
        }
        *sa = ta;
        *sb = tb;
        __syncthreads();
    };
    if (s >= 0)
        for (int c = c0 + lt; c < c1; c += gsize)
            pj_setup_contact(c, nc, ida, idb, J, mu, rest, vb, minv, v, alpha, warmstart, R, bias, p, imp, Rb);
    __syncthreads();
    for (int b = b0 + lt; b < b1; b += gsize)
        pj_setup_body(b, dof, minv, v, fext, fgyro, incptr, incidx, imp, dt, rows ? 1 : 0, x);
    // spectral radius, the power vector in XK1
    const double inv = 1.0 / sqrt((double)(s >= 0 ? snv[s] : 1));
    for (int i = lt; i < 6 * (b1 - b0); i += gsize) {
        const int b = b0 + i / 6, l = i % 6;
        XK1[(size_t)b * kBodyStride + l] = l < dof[b] ? inv : 0.0;
    }
    __syncthreads();
    bool stopped = !rows;
    double lambda = 0;
    for (int pit = 0; pit < 8; ++pit) {
        if (!stopped)
            for (int c = c0 + lt; c < c1; c += gsize) cheb_lin_contact(c, nc, ida, idb, J, R, mu, XK1, imp, Rb);
        __syncthreads();
        double s2 = 0;
        if (!stopped)
            for (int b = b0 + lt; b < b1; b += gsize) {
                double acc[6] = {0, 0, 0, 0, 0, 0};
                for (int i = incptr[b]; i < incptr[b + 1]; ++i) gather_side9(imp, incidx[i], acc);
                const double* mi = minv + (size_t)b * kMinvStride;
                for (int l = 0; l < dof[b]; ++l) {
                    const double mk = l < 3 ? mi[0] : mi[l - 2];
                    const double M = mk > 0 ? 1.0 / mk : 0.0;
                    double* vp = XK1 + (size_t)b * kBodyStride + l;
                    const double un = *vp + acc[l] / M;
                    *vp = un;
                    s2 += un * un;
                }
            }
        double n2 = 0, unused = 0;
        group_sum2(s2, 0.0, &n2, &unused);
        if (!stopped) {
            const double n = sqrt(n2);
            if (!(n > 1e-300)) {
                stopped = true;
                lambda = 0.0;
            } else {
                lambda = n;
                for (int b = b0 + lt; b < b1; b += gsize)
                    for (int l = 0; l < dof[b]; ++l) XK1[(size_t)b * kBodyStride + l] /= n;
            }
        }
        __syncthreads();
    }
    const double rho = lambda > 0 ? fmin(lambda, 0.995) : 0.0;
    const double rho2 = rho * rho;
    // xk = xk1 = x0, pk = pk1 = p0
    for (int i = lt; i < 6 * (b1 - b0); i += gsize) {
        const size_t o = (size_t)(b0 + i / 6) * kBodyStride + i % 6;
        XK[o] = XK1[o] = x[o];
    }
    for (int i = lt; i < 3 * (c1 - c0); i += gsize) {
        const size_t r = (size_t)(i / (c1 - c0)) * nc + c0 + i % (c1 - c0);
        PK[r] = PK1[r] = p[r];
    }
    if (lt == 0) gdone[g] = rows ? 0 : 1;
    bool done = !rows, converged = false;
    int iters = 0;
    double res = NAN, w = 1.0;
    bool take = true;
    __syncthreads();
    for (int it = 0; it < max_iter; ++it) {
        if (!done)
            for (int c = c0 + lt; c < c1; c += gsize)
                pjc_contact_sweep(c, nc, ida, idb, J, R, bias, use_bias, mu, p, PK, PK1, x, XK, w, take, imp, Rb);
        __syncthreads();
        double e2 = 0, bad = 0;
        if (!done)
            for (int b = b0 + lt; b < b1; b += gsize) {
                double acc[6] = {0, 0, 0, 0, 0, 0};
                const int i0 = incptr[b], i1 = incptr[b + 1];
                for (int i = i0; i < i1; ++i) gather_side9(imp, incidx[i], acc);
                const double* mi = minv + (size_t)b * kMinvStride;
                for (int l = 0; l < dof[b]; ++l) {
                    const size_t o = (size_t)b * kBodyStride + l;
                    const double xs = x[o];
                    const double u = take ? xs : cheb_extrap(w, xs, XK[o]);
                    XK[o] = XK1[o];
                    XK1[o] = u;
                    const double mk = l < 3 ? mi[0] : mi[l - 2];
                    const double M = mk > 0 ? 1.0 / mk : 0.0;
                    const double un = i1 > i0 ? u + acc[l] / M : u;
                    x[o] = un;
                    const double q = fmax(fabs(un), fabs(u)) * tol_rel + tol_abs;
                    const double e = (un - u) / q;
                    e2 += e * e;
                    bad += isfinite(un) ? 0.0 : 1.0;
                }
            }
        double t2 = 0, nbad = 0;
        group_sum2(e2, bad, &t2, &nbad);
        if (!done) {
            iters = it + 1;
            res = 1.0 / sqrt((double)snv[s]) * sqrt(t2);
            if (res <= 1.0) {
                done = converged = true;
            } else {
                const double wn = cheb_next_w(it, rho2, w);
                if (!isfinite(wn) || wn <= 0.0 || nbad != 0.0) {
                    w = 1.0;
                    take = true;
                } else {
                    w = wn;
                    take = false;
                }
            }
        }
        if (lt == 0) gdone[g] = done ? 1 : 0;
        if (threadIdx.x == 0) all_done = 0;
        __syncthreads();
        if (threadIdx.x == 0) {
            int a = 1;
            for (int k = 0; k < groups; ++k) a &= gdone[k];
            all_done = a;
        }
        __syncthreads();
        if (all_done) break;
    }
    // the converged sweep's input, or the pending update after max_iter sweeps
    if (rows) {
        for (int i = lt; i < 6 * (b1 - b0); i += gsize) {
            const int b = b0 + i / 6, l = i % 6;
            if (l >= dof[b]) continue;
            const size_t o = (size_t)b * kBodyStride + l;
            if (converged) x[o] = XK1[o];
            else if (!take) x[o] = cheb_extrap(w, x[o], XK[o]);
        }
        for (int i = lt; i < 3 * (c1 - c0); i += gsize) {
            const int k = i / (c1 - c0), c = c0 + i % (c1 - c0);
            if ((ida[c] < 0 && idb[c] < 0) || k >= (mu[c] > 0 ? 3 : 1)) continue;
            const size_t r = (size_t)k * nc + c;
            if (converged) p[r] = PK1[r];
            else if (!take) p[r] = cheb_extrap(w, p[r], PK[r]);
        }
    }
    if (s >= 0 && lt == 0) {
        iters_out[s] = rows ? iters : 0;
        res_out[s] = res;
    }
}

// ---- one scene per workgroup with its velocities and side-impulse records in LDS
//
// When every scene's bodies (6 doubles each), contact records (9 doubles each) and incidence
// lists fit in kBatchLdsBytes, a workgroup hosting one scene keeps them in LDS through the sweeps:
// a sweep then reads from global memory only the contact rows (ids, J, R, bias, mu, p) and the
// bodies' inverse masses, and the body pass and the contact pass's body gathers run at LDS
// latency.  A stagnating
// scene that runs to the iteration cap sets a batch step's time, and its sweep is a chain of
// such latencies.  The arithmetic is that of pj_contact_sweep / pj_body_sweep, on local indices.
constexpr int kBatchLdsBytes = 60 * 1024;  // dynamic LDS, under the 64 KiB a workgroup gets by default

__device__ inline void lds_body6(const double* __restrict__ xs, int lb, double u[6]) {
    if (lb < 0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) u[k] = 0.0;
        return;
    }
    const double2* q = reinterpret_cast<const double2*>(xs + 6 * (size_t)lb);
    const double2 a0 = q[0], a1 = q[1], a2 = q[2];
    u[0] = a0.x;
    u[1] = a0.y;
    u[2] = a1.x;
    u[3] = a1.y;
    u[4] = a2.x;
    u[5] = a2.y;
}

// pj_contact_sweep with x read from the scene's LDS copy (body b at xs + 6 (b - b0)) and the
// record written to the LDS records (contact c at rs + 9 (c - c0)).
__device__ inline void contact_sweep_lds(int c, int c0, int b0, int nc, const int* __restrict__ ida,
                                         const int* __restrict__ idb, const double* __restrict__ J,
                                         const double* __restrict__ R, const double* __restrict__ bias, int use_bias,
                                         const double* __restrict__ mu, double* __restrict__ p,
                                         const double* __restrict__ xs, double* __restrict__ rs,
                                         const double* __restrict__ Rb) {
    const int a = ida[c], b = idb[c];
    if (a < 0 && b < 0) return;
    const double muc = mu[c];
    const int nr = muc > 0 ? 3 : 1;
    ContactJ j;
    load_J(J, nc, c, nr, j);
    double ua[6], ub[6];
    lds_body6(xs, a < 0 ? -1 : a - b0, ua);
    lds_body6(xs, b < 0 ? -1 : b - b0, ub);
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
    pn[0] = fmax(pn[0], 0.0);
    if (nr == 3) {
        const double t1 = pn[1], t2 = pn[2];
        const double tn2 = t1 * t1 + t2 * t2;
        if (tn2 > 0) {
            const double sc = fmin(1.0, (muc * pn[0]) / sqrt(tn2));
            pn[1] = sc * t1;
            pn[2] = sc * t2;
        }
    }
    double dp[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= nr) break;
        dp[k] = pn[k] - pold[k];
        p[(size_t)k * nc + c] = pn[k];
    }
    write_side_impulses9(rs, c - c0, a, b, nr, j, dp);
}

// pj_body_sweep (no joints) on the LDS copies: the scene's incidence lists with local offsets
// (ip, body bi at ip[bi - b0]) and local entries e - 2 c0 (ix), which read rs at 9 (c - c0).
__device__ inline double body_sweep_lds(int bi, int b0, const double4 m4, const int nd, const int* __restrict__ ip,
                                        const int* __restrict__ ix, const double* __restrict__ rs,
                                        double* __restrict__ xs, int check, double tol_rel, double tol_abs) {
    double e2 = 0;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const int i0 = ip[bi - b0], i1 = ip[bi - b0 + 1];
    for (int i = i0; i < i1; ++i) gather_side9(rs, ix[i], acc);
    double u[6];
    lds_body6(xs, bi - b0, u);
    double un[6];
    const double Ml = m4.x > 0 ? 1.0 / m4.x : 0.0;
    const double Ma[3] = {m4.y > 0 ? 1.0 / m4.y : 0.0, m4.z > 0 ? 1.0 / m4.z : 0.0, m4.w > 0 ? 1.0 / m4.w : 0.0};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double M = k < 3 ? Ml : Ma[k - 3];
        un[k] = (k < nd) ? u[k] + acc[k] / M : u[k];
    }
    if (i1 > i0) {
        double2* q = reinterpret_cast<double2*>(xs + 6 * (size_t)(bi - b0));
        q[0] = make_double2(un[0], un[1]);
        q[1] = make_double2(un[2], un[3]);
        q[2] = make_double2(un[4], un[5]);
    }
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

// A contact's sweep inputs held in registers through a scene's sweeps (one contact per lane): its
// rows, scaling or inverse Delassus block, bias and impulses, loaded once after the setup.
struct ContactReg {
    int a, b, nr;
    double muc;
    ContactJ j;
    double R[3], bias[3], p[3], rb[9];
};

__device__ inline void load_contact_reg(int c, int nc, const int* __restrict__ ida, const int* __restrict__ idb,
                                        const double* __restrict__ J, const double* __restrict__ R,
                                        const double* __restrict__ bias, int use_bias, const double* __restrict__ mu,
                                        const double* __restrict__ p, const double* __restrict__ Rb, ContactReg& r) {
    r.a = ida[c];
    r.b = idb[c];
    r.muc = mu[c];
    r.nr = r.muc > 0 ? 3 : 1;
    if (r.a < 0 && r.b < 0) return;
    load_J(J, nc, c, r.nr, r.j);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= r.nr) break;
        r.p[k] = p[(size_t)k * nc + c];
        r.R[k] = Rb ? 0.0 : R[(size_t)k * nc + c];
        r.bias[k] = use_bias ? bias[(size_t)k * nc + c] : 0.0;
#pragma unroll
        for (int l = 0; l < 3; ++l) r.rb[3 * k + l] = Rb ? Rb[(size_t)(3 * k + l) * nc + c] : 0.0;
    }
}

// contact_sweep_lds on the registers of load_contact_reg: the same arithmetic, the impulses kept
// in r.p (written back after the last sweep).
__device__ inline void contact_sweep_reg(ContactReg& r, int cl, int b0, bool block, const double* __restrict__ xs,
                                         double* __restrict__ rs) {
    if (r.a < 0 && r.b < 0) return;
    const int nr = r.nr;
    double ua[6], ub[6];
    lds_body6(xs, r.a < 0 ? -1 : r.a - b0, ua);
    lds_body6(xs, r.b < 0 ? -1 : r.b - b0, ub);
    double pold[3] = {0, 0, 0}, pn[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= nr) break;
        const double g = row_dot(r.j.d[k], r.j.aA[k], r.j.aB[k], ua, ub);
        pold[k] = r.p[k];
        if (block)
            pn[k] = g + r.bias[k];
        else
            pn[k] = pold[k] - (r.R[k] * g + r.bias[k]);
    }
    if (block) {
        const double t0 = pn[0], t1 = pn[1], t2 = pn[2];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (k >= nr) break;
            const double* q = r.rb + 3 * k;
            pn[k] = pold[k] - (nr == 1 ? q[0] * t0 : q[0] * t0 + q[1] * t1 + q[2] * t2);
        }
    }
    pn[0] = fmax(pn[0], 0.0);
    if (nr == 3) {
        const double t1 = pn[1], t2 = pn[2];
        const double tn2 = t1 * t1 + t2 * t2;
        if (tn2 > 0) {
            const double sc = fmin(1.0, (r.muc * pn[0]) / sqrt(tn2));
            pn[1] = sc * t1;
            pn[2] = sc * t2;
        }
    }
    double dp[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= nr) break;
        dp[k] = pn[k] - pold[k];
        r.p[k] = pn[k];
    }
    write_side_impulses9(rs, cl, r.a, r.b, nr, r.j, dp);
}

// Resident: every scene has at most blockDim contacts and bodies, so each lane owns at most one
// of each and keeps its contact's sweep inputs and its body's masses in registers through the
// sweeps (a stagnating scene's sweep is then LDS traffic and barriers only).
template <bool Resident>
__global__ void __launch_bounds__(kMaxBatchBlock) k_batch_pj_lds(
    int ns, int lds_bodies, int lds_contacts, const int* __restrict__ order, const int* __restrict__ boff, const int* __restrict__ coff,
    const int* __restrict__ snv, const int* __restrict__ srows, int nc, const int* __restrict__ ida,
    const int* __restrict__ idb, const double* __restrict__ J, const double* __restrict__ mu,
    const double* __restrict__ rest, const double* __restrict__ vb, double* __restrict__ R, double* __restrict__ bias,
    double* __restrict__ p, double* __restrict__ imp, const int* __restrict__ dof, const double* __restrict__ minv,
    const double* __restrict__ v, const double* __restrict__ fext, const double* __restrict__ fgyro,
    const int* __restrict__ incptr, const int* __restrict__ incidx, double* __restrict__ x, double dt, double alpha,
    int use_bias, int warmstart, int max_iter, int check_every, double tol_rel, double tol_abs, int* __restrict__ iters_out,
    double* __restrict__ res_out, double* __restrict__ Rb) {
    extern __shared__ double sdyn[];
    __shared__ double wsum[kMaxBatchBlock / 64];
    if ((int)blockIdx.x >= ns) return;  // uniform over the workgroup
    const int s = order[blockIdx.x];
    const int b0 = boff[s], b1 = boff[s + 1], c0 = coff[s], c1 = coff[s + 1];
    const bool rows = srows[s] > 0;
    const int lt = threadIdx.x, gsize = blockDim.x;
    double* xs = sdyn;                                // [lds_bodies][6]
    double* rs = sdyn + 6 * (size_t)lds_bodies;      // [lds_contacts][9]
    int* ip = reinterpret_cast<int*>(rs + 9 * (size_t)lds_contacts);  // [lds_bodies + 1]
    int* ix = ip + lds_bodies + 1;                                     // [2 lds_contacts]
    // setup in global memory, as the grid-wide solve, then x0 into LDS
    for (int c = c0 + lt; c < c1; c += gsize)
        pj_setup_contact(c, nc, ida, idb, J, mu, rest, vb, minv, v, alpha, warmstart, R, bias, p, imp, Rb);
    __syncthreads();
    for (int b = b0 + lt; b < b1; b += gsize)
        pj_setup_body(b, dof, minv, v, fext, fgyro, incptr, incidx, imp, dt, rows ? 1 : 0, x);
    __syncthreads();
    for (int i = lt; i < 6 * (b1 - b0); i += gsize) xs[i] = x[(size_t)(b0 + i / 6) * kBodyStride + i % 6];
    const int e0 = incptr[b0], ne = incptr[b1] - e0;  // the scene's incidences: contiguous, at most 2 per contact
    for (int i = lt; i <= b1 - b0; i += gsize) ip[i] = incptr[b0 + i] - e0;
    for (int i = lt; i < ne; i += gsize) ix[i] = incidx[e0 + i] - 2 * c0;
    __syncthreads();
    bool done = !rows;  // the same in every thread: one scene per workgroup
    int iters = 0;
    double res = NAN;
    ContactReg cr;
    double4 bm4 = make_double4(0, 0, 0, 0);
    int bnd = 0;
    const bool own_c = Resident && c0 + lt < c1, own_b = Resident && b0 + lt < b1;
    if (own_c) load_contact_reg(c0 + lt, nc, ida, idb, J, R, bias, use_bias, mu, p, Rb, cr);
    if (own_b) {
        bm4 = *reinterpret_cast<const double4*>(minv + (size_t)(b0 + lt) * kMinvStride);
        bnd = dof[b0 + lt];
    }
    for (int it = 0; it < max_iter && !done; ++it) {
        if constexpr (Resident) {
            if (own_c) contact_sweep_reg(cr, lt, b0, Rb != nullptr, xs, rs);
        } else {
            for (int c = c0 + lt; c < c1; c += gsize)
                contact_sweep_lds(c, c0, b0, nc, ida, idb, J, R, bias, use_bias, mu, p, xs, rs, Rb);
        }
        __syncthreads();
        const int check = (it % check_every) == 0;
        double e2 = 0;
        if constexpr (Resident) {
            if (own_b) e2 = body_sweep_lds(b0 + lt, b0, bm4, bnd, ip, ix, rs, xs, check, tol_rel, tol_abs);
        } else {
            for (int b = b0 + lt; b < b1; b += gsize)
                e2 += body_sweep_lds(b, b0, *reinterpret_cast<const double4*>(minv + (size_t)b * kMinvStride), dof[b],
                                     ip, ix, rs, xs, check, tol_rel, tol_abs);
        }
        if (check) {
            for (int off = 32; off > 0; off >>= 1) e2 += __shfl_xor(e2, off, 64);
            if ((lt & 63) == 0) wsum[lt >> 6] = e2;
        }
        __syncthreads();
        iters = it + 1;
        if (check) {
            double t = 0;
            for (int w = 0; w < (gsize >> 6); ++w) t += wsum[w];
            res = 1.0 / sqrt((double)snv[s]) * sqrt(t);
            if (res <= 1.0) done = true;
        }
    }
    for (int i = lt; i < 6 * (b1 - b0); i += gsize) x[(size_t)(b0 + i / 6) * kBodyStride + i % 6] = xs[i];
    if (own_c && (cr.a >= 0 || cr.b >= 0))
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (k < cr.nr) p[(size_t)k * nc + c0 + lt] = cr.p[k];
    if (lt == 0) {
        iters_out[s] = rows ? iters : 0;
        res_out[s] = res;
    }
}

int upload_scenes(cdx_ctx* ctx) {
    const int ns = ctx->nscenes;
    CDX_TRY(ensure_alloc(ctx, &ctx->d_scene_i, 6 * (size_t)ns + 2, &ctx->scene_cap, "scene tables"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_scene_res, (size_t)ns, &ctx->scene_res_cap, "scene residuals"));
    std::vector<int> buf(6 * (size_t)ns + 2, 0);
    // [order ns | boff ns+1 | coff ns+1 | nv ns | rows ns | iters ns]
    std::copy(ctx->scene_order.begin(), ctx->scene_order.end(), buf.begin());
    std::copy(ctx->scene_boff.begin(), ctx->scene_boff.end(), buf.begin() + ns);
    std::copy(ctx->scene_coff.begin(), ctx->scene_coff.end(), buf.begin() + 2 * ns + 1);
    std::copy(ctx->scene_nv.begin(), ctx->scene_nv.end(), buf.begin() + 3 * ns + 2);
    std::copy(ctx->scene_rows.begin(), ctx->scene_rows.end(), buf.begin() + 4 * ns + 2);
    CDX_HIP(ctx, hipMemcpy(ctx->d_scene_i, buf.data(), sizeof(int) * buf.size(), hipMemcpyHostToDevice));
    return CDX_OK;
}

// Chains of each scene for the batched kernel: a chain belongs to the scene of its bodies (all of
// its bodies must lie in one scene).
static int upload_scene_chains(cdx_ctx* ctx) {
    const int ns = ctx->nscenes, nch = ctx->nchains;
    std::vector<int> scene_of(nch, -1), sc_off(ns + 1, 0), sc_list(nch);
    for (int c = 0; c < nch; ++c) {
        for (int j = ctx->h_jc_off[c]; j < ctx->h_jc_off[c + 1]; ++j)
            for (int side = 0; side < 2; ++side) {
                const int b = ctx->h_jb_body[2 * j + side];
                if (b < 0) continue;
                const int sc = int(std::upper_bound(ctx->scene_boff.begin(), ctx->scene_boff.end(), b) - ctx->scene_boff.begin()) - 1;
                if (scene_of[c] >= 0 && scene_of[c] != sc)
                    return set_err(ctx, CDX_BAD_ARG, "cdx_batch_pj_solve: a joint chain spans two scenes");
                scene_of[c] = sc;
            }
        ++sc_off[scene_of[c] + 1];
    }
    for (int s = 0; s < ns; ++s) sc_off[s + 1] += sc_off[s];
    std::vector<int> fill(sc_off.begin(), sc_off.end() - 1);
    for (int c = 0; c < nch; ++c) sc_list[fill[scene_of[c]]++] = c;
    CDX_TRY(ensure_alloc(ctx, &ctx->d_sc_off, (size_t)ns + 1, &ctx->sc_off_cap, "scene chains"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_sc_list, (size_t)std::max(nch, 1), &ctx->sc_list_cap, "scene chains"));
    CDX_HIP(ctx, hipMemcpy(ctx->d_sc_off, sc_off.data(), sizeof(int) * sc_off.size(), hipMemcpyHostToDevice));
    CDX_HIP(ctx, hipMemcpy(ctx->d_sc_list, sc_list.data(), sizeof(int) * nch, hipMemcpyHostToDevice));
    return CDX_OK;
}

static BatchJoints batch_joints(cdx_ctx* ctx, double theta, int wave_lds) {
    BatchJoints ja{};
    if (ctx->njrows == 0) return ja;
    ja = BatchJoints{ctx->d_sc_off,  ctx->d_sc_list,   ctx->d_jc_off,  ctx->d_jcb_off,     ctx->d_jb_off, ctx->d_jl_off, ctx->d_jb_body,
                     ctx->d_jr_kind, ctx->d_jr_blk,    ctx->d_jr,      ctx->d_jw,          ctx->njrows,   ctx->d_jrlam,       ctx->d_jL,     ctx->d_jDinv,
                     ctx->d_jbody,   ctx->d_jbody_blk, ctx->d_jr_info, ctx->d_jbody_info,  ctx->d_jslot,  ctx->d_jtr,
                     ctx->d_jz,      ctx->d_jy,        ctx->d_jlam,    theta,              ctx->joint_bias,
                     wave_lds};
    return ja;
}

// Doubles of dynamic LDS per wave for a batched launch of `block` threads with joint rows: a
// single-window chain solve's worth where the device's per-workgroup LDS allows it (32-link chains
// of damped spherical joints then stay in LDS for the whole solve), less for large blocks.
static int batch_wave_lds(cdx_ctx* ctx, int block, size_t* shm) {
    if (ctx->max_lds_block == 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) != hipSuccess || v <= 0)
            v = 64 * 1024;
        ctx->max_lds_block = v;
    }
    const int nwv = block / 64;
    const int avail = ctx->max_lds_block - 512;  // the kernel's own sums and flags
    const int per = std::min(kSolveLdsL + kSolveLdsR, avail / (int)(sizeof(double) * nwv));
    *shm = (size_t)per * sizeof(double) * nwv;
    return per;
}

static int launch_batch(cdx_ctx* ctx, double dt, double theta, const cdx_solver_params* prm, int block, int groups) {
    const int ns = ctx->nscenes;
    const int* d = ctx->d_scene_i;
    const int grid = (ns + groups - 1) / groups;
    double* Rb = nullptr;  // block mode: the contacts' inverse Delassus blocks
    if (prm->block_precond && ctx->nc > 0) {
        CDX_TRY(ensure_alloc(ctx, &ctx->d_Rb, 9 * (size_t)ctx->nc, &ctx->Rb_cap, "inverse Delassus blocks"));
        Rb = ctx->d_Rb;
    }
    if (prm->accel == CDX_ACCEL_CHEBYSHEV) {
        const size_t n8 = (size_t)ctx->nb * kBodyStride, n3 = 3 * (size_t)ctx->nc;
        CDX_TRY(ensure_alloc(ctx, &ctx->d_xk1, n8, &ctx->xk1_cap, "chebyshev xk1"));
        CDX_TRY(ensure_alloc(ctx, &ctx->d_pk, n3, &ctx->pk_cap, "chebyshev pk"));
        CDX_TRY(ensure_alloc(ctx, &ctx->d_pk1, n3, &ctx->pk1_cap, "chebyshev pk1"));
        CDX_LAUNCH(ctx, "batch_pjc", k_batch_pjc, grid, block, ns, groups, d, d + ns, d + 2 * ns + 1, d + 3 * ns + 2,
                   d + 4 * ns + 2, ctx->nc, ctx->d_ida, ctx->d_idb, ctx->d_J, ctx->d_mu, ctx->d_rest, ctx->d_vb, ctx->d_R,
                   ctx->d_bias, ctx->d_p, ctx->d_imp, ctx->d_dof, ctx->d_minv, ctx->d_v, ctx->d_fext, ctx->d_fgyro,
                   ctx->d_incptr, ctx->d_incidx, ctx->d_x, ctx->d_uprev, ctx->d_xk1, ctx->d_pk, ctx->d_pk1, dt,
                   prm->alpha, ctx->any_bias ? 1 : 0, prm->warmstart, prm->max_iter, prm->tol_rel, prm->tol_abs,
                   ctx->d_scene_i + 5 * ns + 2, ctx->d_scene_res, Rb);
        return hip_check(ctx, hipGetLastError(), "batch launch");
    }
    size_t shm = 0;
    const int wave_lds = ctx->njrows > 0 ? batch_wave_lds(ctx, block, &shm) : 0;
    if (ctx->njrows > 0 && wave_lds < kMinWaveLds)
        return set_err(ctx, CDX_BAD_ARG, "batched solve: too little LDS per wave for the joint chains at this block size");
#define CDX_BATCH_ARGS                                                                                               \
    ns, groups, d, d + ns, d + 2 * ns + 1, d + 3 * ns + 2, d + 4 * ns + 2, ctx->nc, ctx->d_ida, ctx->d_idb, ctx->d_J,     \
        ctx->d_mu, ctx->d_rest, ctx->d_vb, ctx->d_R, ctx->d_bias, ctx->d_p, ctx->d_imp, ctx->d_dof, ctx->d_minv,         \
        ctx->d_v, ctx->d_fext, ctx->d_fgyro, ctx->d_incptr, ctx->d_incidx, ctx->d_x, dt, prm->alpha,                     \
        ctx->any_bias ? 1 : 0, prm->warmstart, prm->max_iter, prm->check_every, prm->tol_rel, prm->tol_abs,              \
        ctx->d_scene_i + 5 * ns + 2, ctx->d_scene_res, Rb, batch_joints(ctx, theta, wave_lds)
    const size_t lds_bytes =
        sizeof(double) * (6 * (size_t)ctx->scene_max_bodies + 9 * (size_t)ctx->scene_max_contacts) +
        sizeof(int) * ((size_t)ctx->scene_max_bodies + 1 + 2 * (size_t)ctx->scene_max_contacts);
    if (ctx->njrows == 0 && groups == 1 && lds_bytes <= (size_t)kBatchLdsBytes) {
        // at least one lane per contact and per body of the largest scene (up to 512), so that a
        // stagnating scene's passes take one round each: its sweep latency sets the step time
        const int need = (std::max(ctx->scene_max_bodies, ctx->scene_max_contacts) + 63) / 64 * 64;
        const int lds_block = std::min(kMaxBatchBlock, std::max(block, need));
#define CDX_LDS_LAUNCH(RES)                                                                                          \
    CDX_LAUNCH_SHM(ctx, "batch_pj", k_batch_pj_lds<RES>, ns, lds_block, lds_bytes, ns, ctx->scene_max_bodies,          \
                   ctx->scene_max_contacts, d, d + ns, d + 2 * ns + 1, d + 3 * ns + 2, d + 4 * ns + 2, ctx->nc,       \
                   ctx->d_ida, ctx->d_idb, ctx->d_J, ctx->d_mu, ctx->d_rest, ctx->d_vb, ctx->d_R, ctx->d_bias,       \
                   ctx->d_p, ctx->d_imp, ctx->d_dof, ctx->d_minv, ctx->d_v, ctx->d_fext, ctx->d_fgyro,                \
                   ctx->d_incptr, ctx->d_incidx, ctx->d_x, dt, prm->alpha, ctx->any_bias ? 1 : 0, prm->warmstart,     \
                   prm->max_iter, prm->check_every, prm->tol_rel, prm->tol_abs, ctx->d_scene_i + 5 * ns + 2,          \
                   ctx->d_scene_res, Rb)
        if (lds_block >= need)
            CDX_LDS_LAUNCH(true);
        else
            CDX_LDS_LAUNCH(false);
#undef CDX_LDS_LAUNCH
        return hip_check(ctx, hipGetLastError(), "batch launch");
    }
    if (ctx->njrows > 0) {
        if (shm > 64 * 1024)
            CDX_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void*>(&k_batch_pj<true>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        CDX_LAUNCH_SHM(ctx, "batch_pj", k_batch_pj<true>, grid, block, shm, CDX_BATCH_ARGS);
    }
    else
        CDX_LAUNCH(ctx, "batch_pj", k_batch_pj<false>, grid, block, CDX_BATCH_ARGS);
#undef CDX_BATCH_ARGS
    return hip_check(ctx, hipGetLastError(), "batch launch");
}

static bool valid_launch(int block, int groups) {
    return (block == 64 || block == 128 || block == 256 || block == 512) && groups >= 1 && block % groups == 0 &&
           (block / groups) % 64 == 0;
}

}  // namespace cdx

using namespace cdx;

extern "C" {

int cdx_set_scenes(cdx_ctx* ctx, int nscenes, const int* body_off, const int* contact_off) {
    if (!ctx) return CDX_BAD_ARG;
    if (nscenes < 1 || !body_off || !contact_off) return set_err(ctx, CDX_BAD_ARG, "cdx_set_scenes: bad arguments");
    if (body_off[0] != 0 || contact_off[0] != 0 || body_off[nscenes] != ctx->nb || contact_off[nscenes] != ctx->nc)
        return set_err(ctx, CDX_BAD_ARG, "cdx_set_scenes: offsets must start at 0 and cover all bodies and contacts");
    std::vector<int> nv(nscenes, 0), rows(nscenes, 0), cost(nscenes, 0);
    for (int s = 0; s < nscenes; ++s) {
        if (body_off[s + 1] < body_off[s] || contact_off[s + 1] < contact_off[s])
            return set_err(ctx, CDX_BAD_ARG, "cdx_set_scenes: offsets must be non-decreasing");
        for (int b = body_off[s]; b < body_off[s + 1]; ++b) nv[s] += ctx->dof[b];
        for (int c = contact_off[s]; c < contact_off[s + 1]; ++c) {
            const int a = ctx->h_a[c], b = ctx->h_b[c];
            if ((a >= 0 && (a < body_off[s] || a >= body_off[s + 1])) || (b >= 0 && (b < body_off[s] || b >= body_off[s + 1])))
                return set_err(ctx, CDX_BAD_ARG, "cdx_set_scenes: a contact references a body of another scene");
            if (a >= 0 || b >= 0) rows[s] += ctx->h_mu[c] > 0 ? 3 : 1;
        }
        cost[s] = 4 * (contact_off[s + 1] - contact_off[s]) + (body_off[s + 1] - body_off[s]);
    }
    ctx->nscenes = nscenes;
    ctx->scene_boff.assign(body_off, body_off + nscenes + 1);
    ctx->scene_coff.assign(contact_off, contact_off + nscenes + 1);
    ctx->scene_nv = nv;
    ctx->scene_rows = rows;
    ctx->scene_max_bodies = ctx->scene_max_contacts = 0;
    for (int s = 0; s < nscenes; ++s) {
        ctx->scene_max_bodies = std::max(ctx->scene_max_bodies, body_off[s + 1] - body_off[s]);
        ctx->scene_max_contacts = std::max(ctx->scene_max_contacts, contact_off[s + 1] - contact_off[s]);
    }
    ctx->scene_order.resize(nscenes);
    std::iota(ctx->scene_order.begin(), ctx->scene_order.end(), 0);
    std::stable_sort(ctx->scene_order.begin(), ctx->scene_order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    return upload_scenes(ctx);
}

int cdx_batch_set_launch(cdx_ctx* ctx, int block, int groups) {
    if (!ctx) return CDX_BAD_ARG;
    if (!valid_launch(block, groups)) return set_err(ctx, CDX_BAD_ARG, "cdx_batch_set_launch: invalid (block, groups)");
    ctx->batch_block = block;
    ctx->batch_groups = groups;
    return CDX_OK;
}

int cdx_batch_pj_solve(cdx_ctx* ctx, double dt, double theta, const cdx_solver_params* prm, int* iters, double* residuals) {
    if (!ctx) return CDX_BAD_ARG;
    hipSetDevice(ctx->device);
    CDX_TRY(check_solve_args(ctx, prm));
    const bool joints = ctx->njrows > 0;
    if (joints && !(theta > 0)) return set_err(ctx, CDX_BAD_ARG, "cdx_batch_pj_solve: theta must be positive with joint rows");
    if (prm->accel != CDX_ACCEL_NONE && prm->accel != CDX_ACCEL_CHEBYSHEV)
        return set_err(ctx, CDX_BAD_ARG, "cdx_batch_pj_solve: the batched solve runs the standard or the Chebyshev loop");
    if (prm->implicit_gyro)
        return set_err(ctx, CDX_BAD_ARG, "cdx_batch_pj_solve: the batched solve takes explicit gyroscopy only");
    if (prm->accel == CDX_ACCEL_CHEBYSHEV && ctx->njrows > 0)
        return set_err(ctx, CDX_BAD_ARG, "cdx_batch_pj_solve: the batched Chebyshev loop does not take joint rows");
    if (prm->block_precond && ctx->njrows > 0)
        return set_err(ctx, CDX_BAD_ARG, "cdx_batch_pj_solve: the block preconditioner does not take joint rows");
    if (ctx->nscenes < 1 || ctx->scene_boff.back() != ctx->nb || ctx->scene_coff.back() != ctx->nc)
        return set_err(ctx, CDX_NO_STATE, "cdx_batch_pj_solve: scenes not set for the current bodies/contacts");
    ctx->solved = false;
    ctx->pj_last = false;
    ctx->pjj_last = false;
    ctx->joints_solved = false;
    ctx->record_valid = false;
    CDX_TRY(build_incidence(ctx));
    if (joints) CDX_TRY(upload_scene_chains(ctx));
    CDX_TRY(launch_batch(ctx, dt, theta, prm, ctx->batch_block, ctx->batch_groups));
    CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const int ns = ctx->nscenes;
    std::vector<int> it(ns);
    CDX_HIP(ctx, hipMemcpy(it.data(), ctx->d_scene_i + 5 * ns + 2, sizeof(int) * ns, hipMemcpyDeviceToHost));
    if (iters) std::copy(it.begin(), it.end(), iters);
    if (residuals) CDX_HIP(ctx, hipMemcpy(residuals, ctx->d_scene_res, sizeof(double) * ns, hipMemcpyDeviceToHost));
    for (int s = 0; s < ns; ++s)
        if (it[s] < 0) return set_err(ctx, CDX_NOT_FACTORIZED, "cdx_batch_pj_solve: S is not factorized (singular joint block)");
    ctx->solved = true;
    ctx->joints_solved = joints;
    // a converged state of the same fixed point: the implicit adjoint applies (to the row-scaled
    // solve, as after cdx_pj_solve)
    ctx->pj_last = !joints && !prm->block_precond;
    ctx->pjj_last = joints;
    ctx->last_theta = theta;
    ctx->record_valid = false;
    ctx->recorded = 0;
    return CDX_OK;
}

// Times every valid (block, groups) on the current batch and keeps the fastest.  The warm-start
// impulses are restored before each trial so every candidate solves the same problem.
int cdx_batch_autotune(cdx_ctx* ctx, double dt, double theta, const cdx_solver_params* prm, int* block_out, int* groups_out,
                       double* best_ms) {
    if (!ctx) return CDX_BAD_ARG;
    hipSetDevice(ctx->device);
    CDX_TRY(check_solve_args(ctx, prm));
    if (prm->block_precond && ctx->njrows > 0)
        return set_err(ctx, CDX_BAD_ARG, "cdx_batch_autotune: the block preconditioner does not take joint rows");
    if (prm->accel != CDX_ACCEL_NONE && prm->accel != CDX_ACCEL_CHEBYSHEV)
        return set_err(ctx, CDX_BAD_ARG, "cdx_batch_autotune: the batched solve runs the standard or the Chebyshev loop");
    if (prm->implicit_gyro)
        return set_err(ctx, CDX_BAD_ARG, "cdx_batch_autotune: the batched solve takes explicit gyroscopy only");
    if (ctx->nscenes < 1) return set_err(ctx, CDX_NO_STATE, "cdx_batch_autotune: scenes not set");
    if (ctx->njrows > 0 && !(theta > 0)) return set_err(ctx, CDX_BAD_ARG, "cdx_batch_autotune: theta must be positive with joint rows");
    CDX_TRY(build_incidence(ctx));
    if (ctx->njrows > 0) CDX_TRY(upload_scene_chains(ctx));
    const size_t np = 3 * (size_t)ctx->nc;
    double* saved = nullptr;
    CDX_HIP(ctx, hipMalloc((void**)&saved, sizeof(double) * (np ? np : 1)));
    if (np) CDX_HIP(ctx, hipMemcpyAsync(saved, ctx->d_p, sizeof(double) * np, hipMemcpyDeviceToDevice, ctx->stream));
    const int cands[][2] = {{64, 1}, {128, 1}, {128, 2}, {256, 1}, {256, 2}, {256, 4}, {512, 1}, {512, 2}, {512, 4}, {512, 8}};
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    double best = 1e300;
    int bb = ctx->batch_block, bg = ctx->batch_groups;
    int st = CDX_OK;
    for (auto& cd : cands) {
        float ms = 0;
        for (int rep = 0; rep < 2 && st == CDX_OK; ++rep) {  // first run warms caches and code
            if (np) hipMemcpyAsync(ctx->d_p, saved, sizeof(double) * np, hipMemcpyDeviceToDevice, ctx->stream);
            hipEventRecord(e0, ctx->stream);
            st = launch_batch(ctx, dt, theta, prm, cd[0], cd[1]);
            hipEventRecord(e1, ctx->stream);
            hipEventSynchronize(e1);
            hipEventElapsedTime(&ms, e0, e1);
        }
        if (st) break;
        if (ms < best) {
            best = ms;
            bb = cd[0];
            bg = cd[1];
        }
    }
    if (np) hipMemcpyAsync(ctx->d_p, saved, sizeof(double) * np, hipMemcpyDeviceToDevice, ctx->stream);
    hipStreamSynchronize(ctx->stream);
    hipFree(saved);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (st) return st;
    ctx->batch_block = bb;
    ctx->batch_groups = bg;
    if (block_out) *block_out = bb;
    if (groups_out) *groups_out = bg;
    if (best_ms) *best_ms = best;
    return CDX_OK;
}

}  // extern "C"

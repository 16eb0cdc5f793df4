// cdx_joints_device.hpp -- per-chain work of the joint rows (block LDL^T of the Schur complement
// K = G M^-1 G^T + E and its substitutions), shared by the grid-wide kernels (cdx_joints.hip) and
// the batched kernel (cdx_batch.hip).
//
// One wave (64 lanes) works on one chain.  The parts that are independent per row or per body
// (right-hand sides, G M^-1 r, the body updates) use every lane; the block recurrences along the
// chain (factorisation, forward and backward substitution) put block component i on lane i and
// pass the previous block's components between lanes with readlane, so a step along the chain
// costs a few register moves and FMAs instead of a chain of dependent memory loads.  Every dot
// product is summed in the same order as a one-thread-per-chain loop would sum it.
#pragma once

#include <type_traits>

#include "cdx_device.hpp"

namespace cdx {

constexpr int kMaxJRows = 12;    // rows per joint block (a full 6-dof joint with dampers)
constexpr int kJRowStride = 16;  // doubles per row: wa[6], wb[6], comp, vsrc, gerr, lam_prev
constexpr int kJBlk = kMaxJRows * kMaxJRows;
// LDS of one wave's factorisation: K_jj, K_{j,j-1}, the previous D^-1, [D | I] and the
// elimination factors of one Gauss-Jordan column
constexpr int kFactorLds = 3 * kJBlk + 2 * kJBlk + kMaxJRows;
// staging window of the grid-wide chain solve (L entries, rows): a 32-link chain of spherical
// joints with dampers (31 blocks of 6 rows) fits in one window
constexpr int kSolveLdsL = 1536, kSolveLdsR = 192;
// the batched kernel gives each wave at least this many doubles of LDS (the factorisation's
// and a staging window of at least one block)
constexpr int kMinWaveLds = kFactorLds > kSolveLdsR + kJBlk ? kFactorLds : kSolveLdsR + kJBlk;

// inverse mass of dof k of a rigid body
__device__ inline double minv_k(const double* __restrict__ minv, int b, int k) {
    return minv[(size_t)b * kMinvStride + (k < 3 ? 0 : k - 2)];
}

// coefficient row i of block j on body `side` (0: a, 1: b)
__device__ inline const double* jrow(const double* __restrict__ jr, int row, int side) {
    return jr + (size_t)row * kJRowStride + 6 * side;
}

// value of v held by lane l (l uniform across the wave)
__device__ inline double readlane_d(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// the wave's earlier global and LDS writes are visible to all of its lanes
__device__ inline void wave_sync() {
    __threadfence_block();
    __builtin_amdgcn_wave_barrier();
}

// Right-hand side of joint row r in block j (PjAssembler::rhs, pj_assembler.cpp:158-164;
// implicit gyroscopy and lambda-theta off).
__device__ inline void joint_rhs_row(int r, int j, const int* __restrict__ jb_body, const int* __restrict__ jr_kind,
                                     const double* __restrict__ jr, const double* __restrict__ v, double dt, double theta,
                                     double beta, double* __restrict__ rlam) {
    const int ba = jb_body[2 * j], bb = jb_body[2 * j + 1];
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

// Once per solve, chain ch on one wave: the joint-row right-hand sides and the block LDL^T
// factors of K (L_j = K_{j,j-1} D_{j-1}^-1, D_j = K_jj - L_j K_{j,j-1}^T, D_j^-1 by Gauss-Jordan
// with partial pivoting).  `lds` holds kFactorLds doubles private to the wave.  Returns
// CDX_NOT_FACTORIZED (on every lane) when a diagonal block is singular, i.e. S cannot be
// factorised (pj_assembler.cpp:175).
__device__ inline int joint_factor_chain(int ch, int lane, const int* __restrict__ jc_off, const int* __restrict__ jb_off,
                                         const int* __restrict__ jb_body, const int* __restrict__ jr_kind,
                                         const int* __restrict__ jr_blk, const double* __restrict__ jr,
                                         const double* __restrict__ minv, const double* __restrict__ v, double dt,
                                         double theta, double beta, double* __restrict__ rlam,
                                         const int* __restrict__ jl_off, double* __restrict__ Lm,
                                         double* __restrict__ Dinv, int nrow, double* __restrict__ lds) {
    const int j0 = jc_off[ch], j1 = jc_off[ch + 1];
    for (int r = jb_off[j0] + lane; r < jb_off[j1]; r += 64)
        joint_rhs_row(r, jr_blk[r], jb_body, jr_kind, jr, v, dt, theta, beta, rlam);
    double* K = lds;
    double* Kjp = lds + kJBlk;
    double* Dp = lds + 2 * kJBlk;
    double* A = lds + 3 * kJBlk;  // 2 * kJBlk
    double* F = lds + 5 * kJBlk;  // kMaxJRows
    int mp = 0;
    for (int j = j0; j < j1; ++j) {
        const int r0 = jb_off[j], m = jb_off[j + 1] - r0;
        const int ba = jb_body[2 * j], bb = jb_body[2 * j + 1];
        // K_jj = sum over the block's bodies of G M^-1 G^T, plus E
        for (int e = lane; e < m * m; e += 64) {
            const int i = e / m, l = e - (e / m) * m;
            double s = 0;
            for (int side = 0; side < 2; ++side) {
                const int b = side ? bb : ba;
                if (b < 0) continue;
                const double* wi = jrow(jr, r0 + i, side);
                const double* wl = jrow(jr, r0 + l, side);
                for (int k = 0; k < 6; ++k) s += wi[k] * minv_k(minv, b, k) * wl[k];
            }
            if (i == l) {
                const double* rr = jr + (size_t)(r0 + i) * kJRowStride;
                s += rr[12] * (jr_kind[r0 + i] == 0 ? 1.0 / (theta * dt * dt) : 1.0 / (theta * dt));
            }
            K[e] = s;
        }
        if (j > j0) {
            // K_{j,j-1} through the body shared with the previous block
            const int p0 = jb_off[j - 1];
            const int pa = jb_body[2 * (j - 1)], pb = jb_body[2 * (j - 1) + 1];
            const int sb = (ba >= 0 && (ba == pa || ba == pb)) ? ba : bb;
            const int side_j = sb == ba ? 0 : 1, side_p = sb == pa ? 0 : 1;
            for (int e = lane; e < m * mp; e += 64) {
                const int i = e / mp, l = e - (e / mp) * mp;
                const double* wi = jrow(jr, r0 + i, side_j);
                const double* wl = jrow(jr, p0 + l, side_p);
                double t = 0;
                for (int k = 0; k < 6; ++k) t += wi[k] * minv_k(minv, sb, k) * wl[k];
                Kjp[e] = t;
            }
            wave_sync();
            double* L = Lm + jl_off[j];  // packed: m x mp entries
            for (int e = lane; e < m * mp; e += 64) {
                const int i = e / mp, l = e - (e / mp) * mp;
                double t = 0;
                for (int q = 0; q < mp; ++q) t += Kjp[i * mp + q] * Dp[q * mp + l];
                L[e] = t;
                A[e] = t;
            }
            wave_sync();
            for (int e = lane; e < m * m; e += 64) {
                const int i = e / m, l = e - (e / m) * m;
                double t = 0;
                for (int q = 0; q < mp; ++q) t += A[i * mp + q] * Kjp[l * mp + q];
                K[e] -= t;
            }
        }
        wave_sync();
        // D_j^-1 by Gauss-Jordan with partial pivoting on [D_j | I]
        const int w2 = 2 * m;
        for (int e = lane; e < m * w2; e += 64) {
            const int i = e / w2, l = e - (e / w2) * w2;
            A[e] = l < m ? K[i * m + l] : (l - m == i ? 1.0 : 0.0);
        }
        wave_sync();
        for (int c = 0; c < m; ++c) {
            int pr = c;
            for (int i = c + 1; i < m; ++i)
                if (fabs(A[i * w2 + c]) > fabs(A[pr * w2 + c])) pr = i;
            if (pr != c)
                for (int l = lane; l < w2; l += 64) {
                    const double t = A[c * w2 + l];
                    A[c * w2 + l] = A[pr * w2 + l];
                    A[pr * w2 + l] = t;
                }
            wave_sync();
            const double d = A[c * w2 + c];
            if (!(fabs(d) > 0.0)) return CDX_NOT_FACTORIZED;
            wave_sync();
            for (int l = lane; l < w2; l += 64) A[c * w2 + l] /= d;
            for (int i = lane; i < m; i += 64) F[i] = A[i * w2 + c];
            wave_sync();
            for (int e = lane; e < m * w2; e += 64) {
                const int i = e / w2, l = e - (e / w2) * w2;
                const double f = F[i];
                if (i != c && f != 0.0) A[e] -= f * A[c * w2 + l];
            }
            wave_sync();
        }
        // entry (i, l) of D_j^-1 at Dinv[l nrow + r0 + i] (by column, so that the rows of a wave
        // read each column in one coalesced load), zero beyond column m: a substitution may sum
        // all kMaxJRows columns of a row
        for (int e = lane; e < m * kMaxJRows; e += 64) {
            const int i = e / kMaxJRows, l = e - (e / kMaxJRows) * kMaxJRows;
            Dinv[(size_t)l * nrow + r0 + i] = l < m ? A[i * w2 + m + l] : 0.0;
        }
        for (int e = lane; e < m * m; e += 64) {
            const int i = e / m, l = e - (e / m) * m;
            Dp[e] = A[i * w2 + m + l];
        }
        wave_sync();
        mp = m;
    }
    return CDX_OK;
}

// Staging windows along a chain (uniform across the wave; a window always holds at least one
// block, whose L has at most kJBlk entries and whose rows at most kMaxJRows).  off(j) is the
// first row of block j.
// Forward: blocks [j, je) with their L_jj (m_jj x m_jj-1; none for the chain's first block j0).
template <typename Off>
__device__ inline int stage_fwd_end(int j, int j1, int j0, Off off, int capL, int capR) {
    int nl = 0, nr = 0, je = j;
    while (je < j1) {
        const int m = off(je + 1) - off(je);
        const int ml = je > j0 ? m * (off(je) - off(je - 1)) : 0;
        if (je != j && (nl + ml > capL || nr + m > capR)) break;
        nl += ml;
        nr += m;
        ++je;
    }
    return je;
}

// Backward: blocks jj = j-1 down to je (>= j0) with L_{jj+1} (m_{jj+1} x m_jj); returns je.
template <typename Off>
__device__ inline int stage_bwd_end(int j, int j0, Off off, int capL, int capR) {
    int nl = 0, nr = 0, je = j;
    while (je > j0) {
        const int jj = je - 1;
        const int m = off(jj + 1) - off(jj);
        const int ml = (off(jj + 2) - off(jj + 1)) * m;
        if (je != j && (nl + ml > capL || nr + m > capR)) break;
        nl += ml;
        nr += m;
        --je;
    }
    return je;
}

// sum over l < N of a[l * stride] * (v of lane l), in order of l; all N loads are issued first
template <int N>
__device__ inline double lane_dot(const double* __restrict__ a, int stride, double v) {
    double ar[N];
#pragma unroll
    for (int l = 0; l < N; ++l) ar[l] = a[l * stride];
    double t = 0;
#pragma unroll
    for (int l = 0; l < N; ++l) t = fma(ar[l], readlane_d(v, l), t);
    return t;
}

// Block widths with a straight-line body of their own (a spherical joint: 3 rows, with its
// dampers: 6); any other width takes the kMaxJRows-wide form, masked to its width.
__device__ inline bool special_width(int n) { return n == 3 || n == 6; }

// f(std::integral_constant<int, N>) for a wave-uniform n in [1, kMaxJRows], with N = n for the
// special widths and N = kMaxJRows >= n otherwise: f must mask columns beyond n itself.
template <typename F>
__device__ inline auto with_width(int n, F&& f) {
    using std::integral_constant;
    switch (n) {
        case 3: return f(integral_constant<int, 3>{});
        case 6: return f(integral_constant<int, 6>{});
        default: return f(integral_constant<int, kMaxJRows>{});
    }
}

// lane_dot for a wave-uniform n in [1, kMaxJRows]; reads a[l * stride] for l < n only
__device__ inline double lane_dot_n(int n, const double* __restrict__ a, int stride, double v) {
    if (n == 3) return lane_dot<3>(a, stride, v);
    if (n == 6) return lane_dot<6>(a, stride, v);
    double ar[kMaxJRows];
#pragma unroll
    for (int l = 0; l < kMaxJRows; ++l) ar[l] = a[(l < n ? l : n - 1) * stride];
    double t = 0;
#pragma unroll
    for (int l = 0; l < kMaxJRows; ++l) {
        const double vl = readlane_d(v, l);
        t = l < n ? fma(ar[l], vl, t) : t;
    }
    return t;
}

// The two recurrences along a chain of nblk blocks that all have W rows, held in LDS: block b's
// rows at sR[b W, b W + W), L_b (b >= 1, W x W row-major) at sL[(b - 1) W W].  The next block's L
// row (column, going back) and row value are read before the current step's FMAs, so LDS
// latency stays off the chain of dependent steps.  Same arithmetic, in the same order, as the
// general steps.
//   forward:  z_b = s_b - L_b z_{b-1}          (z_0 = s_0)
template <int W>
__device__ inline void fwd_uniform(int nblk, int lane, const double* __restrict__ sL, double* __restrict__ sR) {
    const bool act = lane < W;
    const int li = act ? lane : 0;
    double zp = sR[li];
    double lr[W], sn = 0;
#pragma unroll
    for (int l = 0; l < W; ++l) lr[l] = nblk > 1 ? sL[li * W + l] : 0.0;
    if (nblk > 1) sn = sR[W + li];
    for (int b = 1; b < nblk; ++b) {
        double cl[W];
#pragma unroll
        for (int l = 0; l < W; ++l) cl[l] = lr[l];
        const double cs = sn;
        if (b + 1 < nblk) {
            const double* ln = sL + b * W * W + li * W;
#pragma unroll
            for (int l = 0; l < W; ++l) lr[l] = ln[l];
            sn = sR[(b + 1) * W + li];
        }
        double t = 0;
#pragma unroll
        for (int l = 0; l < W; ++l) t = fma(cl[l], readlane_d(zp, l), t);
        const double zc = cs - t;
        if (act) sR[b * W + lane] = zc;
        zp = act ? zc : 0.0;
    }
}

//   backward: Lambda_b = y_b - L_{b+1}^T Lambda_{b+1}   (Lambda_{nblk-1} = y_{nblk-1})
template <int W>
__device__ inline void bwd_uniform(int nblk, int lane, const double* __restrict__ sL, double* __restrict__ sR) {
    const bool act = lane < W;
    const int li = act ? lane : 0;
    double ln = sR[(nblk - 1) * W + li];
    double lc[W], yn = 0;
    const int b0 = nblk - 2;
#pragma unroll
    for (int i = 0; i < W; ++i) lc[i] = b0 >= 0 ? sL[b0 * W * W + i * W + li] : 0.0;
    if (b0 >= 0) yn = sR[b0 * W + li];
    for (int b = b0; b >= 0; --b) {
        double cl[W];
#pragma unroll
        for (int i = 0; i < W; ++i) cl[i] = lc[i];
        const double cy = yn;
        if (b >= 1) {
            const double* lcol = sL + (b - 1) * W * W + li;
#pragma unroll
            for (int i = 0; i < W; ++i) lc[i] = lcol[i * W];
            yn = sR[(b - 1) * W + li];
        }
        double t = 0;
#pragma unroll
        for (int i = 0; i < W; ++i) t = fma(cl[i], readlane_d(ln, i), t);
        const double lj = cy - t;
        if (act) sR[b * W + lane] = lj;
        ln = act ? lj : 0.0;
    }
}

// dst[i] = src[i] for i < n by one wave; every load of a lane is issued before its first store,
// so a copy costs about one memory latency per 512 elements instead of one per element.
__device__ inline void wave_copy(double* __restrict__ dst, const double* __restrict__ src, int n, int lane) {
    constexpr int U = 8;
    for (int i0 = 0; i0 < n; i0 += 64 * U) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + 64 * u + lane;
            v[u] = src[i < n ? i : n - 1];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + 64 * u + lane;
            if (i < n) dst[i] = v[u];
        }
    }
}

// rows and (body, dof) items of one lane in the single-window solve
constexpr int kWinRowsPerLane = kSolveLdsR / 64;
constexpr int kWinItemsPerLane = 6;

// x += S^-1 [r; -r_lambda] restricted to the jointed bodies of chain ch, on one wave.
//   setup = 1: x0 and Lambda_0 are set (absolute; tr holds the full body right-hand side);
//   setup = 0: a sweep increment (tr holds W^T dp of each jointed body); x and Lambda accumulate.
// Returns this lane's share of the chain bodies' residual sum (projected_jacobi.cpp:140-144)
// when `check`.  z and y are scratch of one double per joint row; `lds` holds capL + capR
// doubles private to the wave (capR <= kSolveLdsR).
//
//   s = G M^-1 t (rows in parallel);  z_j = s_j - L_j z_{j-1} (along the chain);
//   y_j = D_j^-1 z_j (rows in parallel);  Lambda_j = y_j - L_{j+1}^T Lambda_{j+1} (back along it);
//   Lambda, then the bodies' dv = M^-1 (t - G^T Lambda) (in parallel).
// A chain whose rows fit in capR and whose L fits in capL keeps s, z, y and Lambda in LDS for
// the whole solve; a longer one goes through z and y in memory, a staged window at a time.
// L_j (m_j x m_{j-1}, row-major) of the chain's blocks are packed back to back from jl_off[j].
// jr_info per row: {body a, body b, block, r0 << 4 | m}; jbody_info per jointed body: {body,
// rows of its first block, rows of its second block (-1: none), 0}, rows as r0 << 5 | m << 1 | side.
// The first 64 block and L offsets of the chain sit in registers, one per lane, read by readlane.
// Widths = true specialises the single-window path's block steps and row sums on the block
// width, a scalar value there (register reads of the chain's offsets); the windowed path always
// uses the fixed kMaxJRows-wide form.
template <bool Widths>
__device__ inline double joint_solve_chain(int ch, int lane, int setup, int check, const int* __restrict__ jc_off,
                                           const int* __restrict__ jcb_off, const int* __restrict__ jb_off,
                                           const int* __restrict__ jl_off, const int* __restrict__ jr_info,
                                           const double* __restrict__ jr, const double* __restrict__ jw, int nrow,
                                           const double* __restrict__ rlam,
                                           const double* __restrict__ Lm, const double* __restrict__ Dinv,
                                           const int* __restrict__ jbody_info, const double* __restrict__ minv,
                                           const double* __restrict__ tr, double* __restrict__ z, double* __restrict__ yb,
                                           double* __restrict__ lam, double* __restrict__ x, double tol_rel, double tol_abs,
                                           double* __restrict__ lds, int capL, int capR) {
    const int j0 = jc_off[ch], j1 = jc_off[ch + 1];
    const int boff = lane <= j1 - j0 ? jb_off[j0 + lane] : 0;
    const int loffr = lane <= j1 - j0 ? jl_off[j0 + lane] : 0;
    auto off = [&](int jj) -> int {
        const int i = jj - j0;
        return i < 64 ? __builtin_amdgcn_readlane(boff, i) : jb_off[jj];
    };
    auto loff = [&](int jj) -> int {
        const int i = jj - j0;
        return i < 64 ? __builtin_amdgcn_readlane(loffr, i) : jl_off[jj];
    };
    const int rb = off(j0), re = off(j1);
    const int lb = loff(j0), le = loff(j1);
    const int4* ri4 = reinterpret_cast<const int4*>(jr_info);
    const int4* bi4 = reinterpret_cast<const int4*>(jbody_info);
    const int qb = jcb_off[ch], nq = jcb_off[ch + 1] - qb;
    // uniform: the chain's L, its rows (also 6 products per row) and its jointed bodies fit, and
    // all of its block offsets sit in the registers
    const bool one = j1 - j0 < 64 && re - rb <= capR && le - lb <= capL && 6 * (re - rb) <= capL &&
                     6 * nq <= 64 * kWinItemsPerLane;
    auto off_r = [&](int jj) -> int { return __builtin_amdgcn_readlane(boff, jj - j0); };
    auto loff_r = [&](int jj) -> int { return __builtin_amdgcn_readlane(loffr, jj - j0); };
    double* sL = lds;         // staged L entries (packed as in Lm)
    double* sR = lds + capL;  // staged row values
    // z_j = s_j - L_j z_{j-1} over blocks [ja, jb) held in LDS from L offset lbase and row rs
    double zp = 0.0;
    int mp = 0;
    // (block offsets through offf / lofff; Spec: width-specialised steps, only where offf and
    // lofff are register reads, i.e. the widths are scalar values)
    auto fwd_blocks = [&](int ja, int jb, int lbase, int rs, auto offf, auto lofff, auto Spec) {
        for (int jj = ja; jj < jb; ++jj) {
            const int r0 = offf(jj), m = offf(jj + 1) - r0;
            const bool act = lane < m;
            const int li = act ? lane : 0;
            double t = 0;
            if constexpr (decltype(Spec)::value) {
                if (jj > j0) t = lane_dot_n(mp, sL + lofff(jj) - lbase + li * mp, 1, zp);
            } else if (jj > j0) {
                const int o = lofff(jj) - lbase + li * mp;
                double lr[kMaxJRows];
#pragma unroll
                for (int l = 0; l < kMaxJRows; ++l) lr[l] = sL[o + (l < mp ? l : mp - 1)];
#pragma unroll
                for (int l = 0; l < kMaxJRows; ++l) {
                    const double zl = readlane_d(zp, l);
                    t = l < mp ? fma(lr[l], zl, t) : t;
                }
            }
            const double zc = sR[r0 - rs + li] - t;
            if (act) sR[r0 - rs + lane] = zc;
            zp = act ? zc : 0.0;
            mp = m;
        }
    };
    // Lambda_jj = y_jj - L_{jj+1}^T Lambda_{jj+1} for jj = jhi-1 down to jlo
    double lnext = 0.0;
    int mnx = 0;
    auto bwd_blocks = [&](int jhi, int jlo, int lbase, int rs, auto offf, auto lofff, auto Spec) {
        for (int jj = jhi - 1; jj >= jlo; --jj) {
            const int r0 = offf(jj), m = offf(jj + 1) - r0;
            const bool act = lane < m;
            const int li = act ? lane : 0;
            double t = 0;
            if constexpr (decltype(Spec)::value) {
                t = lane_dot_n(mnx, sL + lofff(jj + 1) - lbase + li, m, lnext);
            } else {
                const int o = lofff(jj + 1) - lbase + li;
                double lr[kMaxJRows];
#pragma unroll
                for (int l = 0; l < kMaxJRows; ++l) lr[l] = sL[o + (l < mnx ? l : mnx - 1) * m];
#pragma unroll
                for (int l = 0; l < kMaxJRows; ++l) {
                    const double ll = readlane_d(lnext, l);
                    t = l < mnx ? fma(lr[l], ll, t) : t;
                }
            }
            const double lj = sR[r0 - rs + li] - t;
            if (act) sR[r0 - rs + lane] = lj;
            lnext = act ? lj : 0.0;
            mnx = m;
        }
    };
    // y_r = (D_j^-1 z_j)_i for row r = r0 + i of block j, z read through zat
    auto y_row = [&](int r, auto zat) -> double {
        const int4 ri = ri4[r];
        const int r0 = ri.w >> 4, m = ri.w & 15;
        double t = 0;  // D^-1 is zero beyond column m
#pragma unroll
        for (int l = 0; l < kMaxJRows; ++l) t = fma(Dinv[(size_t)l * nrow + r], zat(r0 + (l < m ? l : m - 1)), t);
        return t;
    };
    if (one) {
        // widest block of the chain (wave-uniform)
        int kw = kMaxJRows;
        bool same = false;  // every block of the chain has kw rows, a special width
        if constexpr (Widths) {
            const int bnext = __shfl_down(boff, 1, 64);  // every lane takes part in the shuffles
            int mw = lane < j1 - j0 ? bnext - boff : 0;
            int mn = lane < j1 - j0 ? bnext - boff : kMaxJRows + 1;
            for (int o = 32; o > 0; o >>= 1) {
                mw = max(mw, __shfl_xor(mw, o, 64));
                mn = min(mn, __shfl_xor(mn, o, 64));
            }
            kw = __builtin_amdgcn_readfirstlane(mw);
            same = __builtin_amdgcn_readfirstlane(mn) == kw && special_width(kw);
        }
        // this lane's rows rb + lane + 64 u: records and coefficients stay in registers
        int4 ri[kWinRowsPerLane];
        double w[kWinRowsPerLane][12];
#pragma unroll
        for (int u = 0; u < kWinRowsPerLane; ++u) {
            const int r = rb + lane + 64 * u, rc = r < re ? r : re - 1;
            ri[u] = ri4[rc];
#pragma unroll
            for (int k = 0; k < 12; ++k) w[u][k] = jw[(size_t)k * nrow + rc];
        }
        // s = G M^-1 t (minus r_lambda at setup)
#pragma unroll
        for (int u = 0; u < kWinRowsPerLane; ++u) {
            const int r = rb + lane + 64 * u, rc = r < re ? r : re - 1;
            double s = 0;
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const int b = side ? ri[u].y : ri[u].x;
                const int bc = b < 0 ? 0 : b;
                const double* t = tr + (size_t)bc * 6;
#pragma unroll
                for (int k = 0; k < 6; ++k) {  // a static side adds exact zeros
                    const double wk = b >= 0 ? w[u][6 * side + k] : 0.0;
                    s = fma(wk, minv_k(minv, bc, k) * t[k], s);
                }
            }
            const double val = setup ? s - rlam[rc] : s;
            if (r < re) sR[r - rb] = val;
        }
        wave_copy(sL, Lm + lb, le - lb, lane);
        wave_sync();
        if (same)
            with_width(kw, [&](auto N) {
                fwd_uniform<decltype(N)::value>(j1 - j0, lane, sL, sR);
                return 0;
            });
        else
            fwd_blocks(j0, j1, lb, rb, off_r, loff_r, std::integral_constant<bool, Widths>{});
        wave_sync();
        // y = D^-1 z of this lane's rows, n >= the chain's widest block columns (D^-1 is zero
        // beyond a block's width); one straight-line copy per width
        double yv[kWinRowsPerLane];
        auto ycalc = [&](auto N) {
            constexpr int n = decltype(N)::value;
#pragma unroll
            for (int u = 0; u < kWinRowsPerLane; ++u) {
                const int r = rb + lane + 64 * u, rc = r < re ? r : re - 1;
                const int r0 = ri[u].w >> 4, m = ri[u].w & 15;
                double d[n], t = 0;
#pragma unroll
                for (int l = 0; l < n; ++l) d[l] = Dinv[(size_t)l * nrow + rc];
#pragma unroll
                for (int l = 0; l < n; ++l) t = fma(d[l], sR[r0 - rb + (l < m ? l : m - 1)], t);
                yv[u] = t;
            }
            return 0;
        };
        if constexpr (Widths)
            with_width(kw, ycalc);
        else
            ycalc(std::integral_constant<int, kMaxJRows>{});
        wave_sync();
#pragma unroll
        for (int u = 0; u < kWinRowsPerLane; ++u) {
            const int r = rb + lane + 64 * u;
            if (r < re) sR[r - rb] = yv[u];
        }
        wave_sync();
        if (same) {
            with_width(kw, [&](auto N) {
                bwd_uniform<decltype(N)::value>(j1 - j0, lane, sL, sR);
                return 0;
            });
        } else {
            const int rl = off_r(j1 - 1);
            lnext = lane < re - rl ? sR[rl - rb + lane] : 0.0;
            mnx = re - rl;
            bwd_blocks(j1 - 1, j0, lb, rb, off_r, loff_r, std::integral_constant<bool, Widths>{});
        }
        wave_sync();
        double lv[kWinRowsPerLane];
#pragma unroll
        for (int u = 0; u < kWinRowsPerLane; ++u) {
            const int r = rb + lane + 64 * u;
            lv[u] = r < re ? sR[r - rb] : 0.0;
            if (r < re) lam[r] = setup ? lv[u] : lam[r] + lv[u];
        }
        // bodies: t - G_b^T Lambda from the row products w Lambda, staged in the (now free) L
        // window one side at a time; one (body, dof) item per lane and slot
        const int ni = 6 * nq;
        int4 bi[kWinItemsPerLane];
        double acc[kWinItemsPerLane];
#pragma unroll
        for (int i = 0; i < kWinItemsPerLane; ++i) {
            const int e = lane + 64 * i, ec = e < ni ? e : ni - 1;
            const int q = qb + ec / 6, k = ec - (ec / 6) * 6;
            bi[i] = bi4[q];
            acc[i] = tr[(size_t)bi[i].x * 6 + k];
        }
        double* P = sL;
#pragma unroll
        for (int side = 0; side < 2; ++side) {
#pragma unroll
            for (int u = 0; u < kWinRowsPerLane; ++u) {
                const int r = rb + lane + 64 * u;
                if (r < re)
#pragma unroll
                    for (int k = 0; k < 6; ++k) P[(r - rb) * 6 + k] = w[u][6 * side + k] * lv[u];
            }
            wave_sync();
            // each item subtracts the products of its block rows on this side (n >= the widest block)
            auto isum = [&](auto N) {
                constexpr int n = decltype(N)::value;
#pragma unroll
                for (int i = 0; i < kWinItemsPerLane; ++i) {
                    const int e = lane + 64 * i, k = e - (e / 6) * 6;
                    if (e >= ni) continue;
#pragma unroll
                    for (int ee = 0; ee < 2; ++ee) {
                        const int enc = ee ? bi[i].z : bi[i].y;
                        const bool use = enc >= 0 && (enc & 1) == side;
                        const int r0 = use ? enc >> 5 : rb, m = use ? (enc >> 1) & 15 : 0;
                        double pv[n];
#pragma unroll
                        for (int l = 0; l < n; ++l) pv[l] = P[(r0 - rb + (l < m ? l : 0)) * 6 + k];
#pragma unroll
                        for (int l = 0; l < n; ++l) acc[i] = l < m ? acc[i] - pv[l] : acc[i];
                    }
                }
                return 0;
            };
            if constexpr (Widths)
                with_width(kw, isum);
            else
                isum(std::integral_constant<int, kMaxJRows>{});
            wave_sync();
        }
        double e2 = 0;
#pragma unroll
        for (int i = 0; i < kWinItemsPerLane; ++i) {
            const int e = lane + 64 * i, k = e - (e / 6) * 6;
            if (e >= ni) continue;
            const int b = bi[i].x;
            double* xb = x + (size_t)b * kBodyStride;
            const double dv = minv_k(minv, b, k) * acc[i];
            const double u = xb[k];
            const double un = setup ? dv : u + dv;
            xb[k] = un;
            if (check) {
                const double qd = fmax(fabs(un), fabs(u)) * tol_rel + tol_abs;
                const double ee2 = (un - u) / qd;
                e2 += ee2 * ee2;
            }
        }
        return e2;
    } else {
        // s = G M^-1 t (minus r_lambda at setup) into z
        for (int r = rb + lane; r < re; r += 64) {
            const int4 ri = ri4[r];
            double s = 0;
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const int b = side ? ri.y : ri.x;
                const int bc = b < 0 ? 0 : b;
                const double* w = jrow(jr, r, side);
                const double* t = tr + (size_t)bc * 6;
#pragma unroll
                for (int k = 0; k < 6; ++k) s = fma(b >= 0 ? w[k] : 0.0, minv_k(minv, bc, k) * t[k], s);
            }
            z[r] = setup ? s - rlam[r] : s;
        }
        wave_sync();
        // forward along the chain, a staged window at a time
        for (int j = j0; j < j1;) {
            const int je = stage_fwd_end(j, j1, j0, off, capL, capR);
            const int rs = off(j), nr = off(je) - rs;
            wave_copy(sL, Lm + loff(j), loff(je) - loff(j), lane);
            wave_copy(sR, z + rs, nr, lane);
            wave_sync();
            fwd_blocks(j, je, loff(j), rs, off, loff, std::false_type{});
            wave_sync();
            wave_copy(z + rs, sR, nr, lane);
            wave_sync();
            j = je;
        }
        for (int r = rb + lane; r < re; r += 64) yb[r] = y_row(r, [&](int rr) { return z[rr]; });
        wave_sync();
        // backward: the last block keeps y; windows of blocks [je, j) going down
        const int rl = off(j1 - 1), ml = re - rl;
        lnext = lane < ml ? yb[rl + lane] : 0.0;
        if (lane < ml) z[rl + lane] = lnext;
        mnx = ml;
        for (int j = j1 - 1; j > j0;) {
            const int je = stage_bwd_end(j, j0, off, capL, capR);
            const int rs = off(je), nr = off(j) - rs;
            wave_copy(sL, Lm + loff(je + 1), loff(j + 1) - loff(je + 1), lane);
            wave_copy(sR, yb + rs, nr, lane);
            wave_sync();
            bwd_blocks(j, je, loff(je + 1), rs, off, loff, std::false_type{});
            wave_sync();
            wave_copy(z + rs, sR, nr, lane);
            wave_sync();
            j = je;
        }
        for (int r = rb + lane; r < re; r += 64) lam[r] = setup ? z[r] : lam[r] + z[r];
    }
    // bodies: dv = M^-1 (r - sum over its blocks of G_b^T Lambda), one (body, dof) per lane
    auto bodies = [&](auto lamat) -> double {
        double e2 = 0;
        for (int e = lane; e < 6 * nq; e += 64) {
            const int q = qb + e / 6, k = e - (e / 6) * 6;
            const int4 bi = bi4[q];
            const int b = bi.x;
            double acc = tr[(size_t)b * 6 + k];
#pragma unroll
            for (int ee = 0; ee < 2; ++ee) {
                const int enc = ee ? bi.z : bi.y;
                const bool has = enc >= 0;
                const int r0 = has ? enc >> 5 : rb, m = has ? (enc >> 1) & 15 : 1, side = enc & 1;
#pragma unroll
                for (int l = 0; l < kMaxJRows; ++l) {
                    const int r = r0 + (l < m ? l : m - 1);
                    const double w = jrow(jr, r, side)[k];
                    const double lv = lamat(r);
                    acc = has && l < m ? fma(-w, lv, acc) : acc;
                }
            }
            double* xb = x + (size_t)b * kBodyStride;
            const double dv = minv_k(minv, b, k) * acc;
            const double u = xb[k];
            const double un = setup ? dv : u + dv;
            xb[k] = un;
            if (check) {
                const double qd = fmax(fabs(un), fabs(u)) * tol_rel + tol_abs;
                const double ee2 = (un - u) / qd;
                e2 += ee2 * ee2;
            }
        }
        return e2;
    };
    return bodies([&](int r) { return z[r]; });
}

}  // namespace cdx

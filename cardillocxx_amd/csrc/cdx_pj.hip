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
#include <algorithm>
#include <cmath>
#include <deque>
#include <mutex>
#include <string>

#include "cdx_pj_device.hpp"

namespace cdx {

// ---------------------------------------------------------------------------------------------
// Setup (build_workspace, projected_jacobi.cpp:316-369)
// ---------------------------------------------------------------------------------------------

// R = alpha / D (rdiag_sparse, :57-69); b = R .* (c + e .* (W v + c)) (:337-347), or with Rb the
// inverse Delassus blocks and b unscaled (block mode, :341-343);
// p0 = warm start (:358-359); side impulses of W^T p0 for x0 (:361-363).
__global__ void k_pj_setup_contacts(int nc, const int* __restrict__ ida, const int* __restrict__ idb,
                                    const double* __restrict__ J, const double* __restrict__ mu,
                                    const double* __restrict__ rest, const double* __restrict__ vb,
                                    const double* __restrict__ minv, const double* __restrict__ v, double alpha,
                                    int warmstart, double* __restrict__ R, double* __restrict__ bias,
                                    double* __restrict__ p, double* __restrict__ imp, double* __restrict__ Rb) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    pj_setup_contact(c, nc, ida, idb, J, mu, rest, vb, minv, v, alpha, warmstart, R, bias, p, imp, Rb);
}

// x0 = S^-1 (M v + dt f_ext + dt f_gyro + W^T p0) (pj_assembler.cpp:150-153,173-182).
// Also used for the contact-free free solve (:460-467) with an empty incidence list.
__global__ void k_pj_setup_bodies(int nb, const int* __restrict__ dof, const double* __restrict__ minv,
                                  const double* __restrict__ v, const double* __restrict__ fext,
                                  const double* __restrict__ fgyro, const int* __restrict__ incptr,
                                  const int* __restrict__ incidx, const double* __restrict__ imp, double dt,
                                  int have_contacts, double* __restrict__ x, const int* __restrict__ jslot,
                                  double* __restrict__ tr, const double* __restrict__ sinv) {
    const int bi = blockIdx.x * blockDim.x + threadIdx.x;
    if (bi >= nb) return;
    pj_setup_body(bi, dof, minv, v, fext, fgyro, incptr, incidx, imp, dt, have_contacts, x, jslot, tr, sinv);
}

// Implicit gyroscopy: the inverse rotational block of S of every body (zero for bodies without 6
// DOFs), one body per thread; a singular or non-finite block stops the solve as not factorized
// (the reference's SparseLU of S fails).
__global__ void k_pj_gyro_blocks(int nb, const int* __restrict__ dof, const double* __restrict__ minv,
                                 const double* __restrict__ v, double dt, double* __restrict__ sinv,
                                 int* __restrict__ ctrl) {
    const int bi = blockIdx.x * blockDim.x + threadIdx.x;
    if (bi >= nb) return;
    double K[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (dof[bi] == 6 &&
        !gyro_block_inverse(minv + (size_t)bi * kMinvStride, v + (size_t)bi * kBodyStride + 3, dt, K)) {
        ctrl[2] = CDX_NOT_FACTORIZED;
        ctrl[0] = 1;
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) sinv[(size_t)bi * 9 + k] = K[k];
}

// ---------------------------------------------------------------------------------------------
// Sweep
// ---------------------------------------------------------------------------------------------

// Contact pass with 4 lanes per contact (lane k = row k of n, t1, t2; lane 3 idle): each lane
// loads only its row's 9 Jacobian entries, so the kernel needs about half the registers of a
// one-lane-per-contact form and runs at twice the occupancy (64 VGPRs: recomputing R = alpha / D
// in the sweep instead of streaming it costs 16 more and two waves per SIMD, a net loss).  The
// bias is read only when the problem has one.  The normal and tangential values meet through
// DPP exchanges inside the group of 4 for the Coulomb projection (every lane evaluates the same
// scale factor), the side impulses J^T dp are summed over the rows by quad_sum, and lanes 0-2
// write the contact's compact record [lin | angA | angB] (72 contiguous bytes).

// project() (projected_jacobi.cpp:100-119) of one contact held by a group of 4 lanes (lane k has
// the trial value of row k): every lane evaluates the same scale factor from the group's values.
__device__ __forceinline__ double project4(double pt, int k, double muc) {
    const double zn = quad_bcast<0>(pt);
    const double z1 = quad_bcast<1>(pt);
    const double z2 = quad_bcast<2>(pt);
    const double pn = fmax(zn, 0.0);
    double pnew = pn;
    if (k > 0) {
        pnew = k == 1 ? z1 : z2;
        const double tn2 = z1 * z1 + z2 * z2;
        if (tn2 > 0) {
            const double sc = fmin(1.0, (muc * pn) / sqrt(tn2));
            pnew = sc * pnew;
        }
    }
    return pnew;
}

// Row k of R_block t for the contact held by a group of 4 lanes (lane k has t_k; block mode):
// sum_l Rb(k, l) t_l over the contact's rows, in row order (BlockDiagonal::applyTo,
// block_diagonal.hpp:103-124).  Every lane takes part in the exchanges; rows only read Rb.
__device__ __forceinline__ double block_apply4(double t, int k, int c, int nc, int nr, bool row,
                                               const double* __restrict__ Rb) {
    const double t0 = quad_bcast<0>(t);
    const double t1 = quad_bcast<1>(t);
    const double t2 = quad_bcast<2>(t);
    if (!row) return 0.0;
    const double* r = Rb + (size_t)(3 * k) * nc + c;
    if (nr == 1) return r[0] * t0;
    return r[0] * t0 + r[nc] * t1 + r[2 * (size_t)nc] * t2;
}

// The contact's compact record [lin | angA | angB] = sum over rows of (d_k, aA_k, aB_k) dp_k,
// summed over the group of 4 by quad_sum; lane 0 writes lin, lane 1 angA (dynamic A only),
// lane 2 angB (dynamic B only).
__device__ __forceinline__ void write_record4(bool contact, int k, int c, int a, int b, const double d[3],
                                              const double aA[3], const double aB[3], double dp,
                                              double* __restrict__ imp9) {
    double v9[9];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        v9[m] = d[m] * dp;
        v9[3 + m] = aA[m] * dp;
        v9[6 + m] = aB[m] * dp;
    }
#pragma unroll
    for (int m = 0; m < 9; ++m) v9[m] = quad_sum(v9[m]);
    if (contact && k < 3 && (k == 0 || (k == 1 ? a >= 0 : b >= 0))) {
        const double o0 = k == 0 ? v9[0] : (k == 1 ? v9[3] : v9[6]);
        const double o1 = k == 0 ? v9[1] : (k == 1 ? v9[4] : v9[7]);
        const double o2 = k == 0 ? v9[2] : (k == 1 ? v9[5] : v9[8]);
        double* dst = imp9 + 9 * (size_t)c + 3 * k;
        dst[0] = o0;
        dst[1] = o1;
        dst[2] = o2;
    }
}

// Block = true: R points to the inverse Delassus blocks and the bias is unscaled
// (pj.rdiag_true_delassus): dp = Pi(p - Rb (W x + b)) - p.
template <bool Block>
__global__ void __launch_bounds__(kBlock) k_pj_contacts4(int nc, int nblocks, const int* __restrict__ ida,
                                                         const int* __restrict__ idb, const double* __restrict__ J,
                                                         const double* __restrict__ R,
                                                         const double* __restrict__ bias, int use_bias,
                                                         const double* __restrict__ mu, double* __restrict__ p,
                                                         const double* __restrict__ x, double* __restrict__ imp9,
                                                         const int* __restrict__ ctrl) {
    if (ctrl[0]) return;
    const int tid = xcd_tile(nblocks) * kBlock + threadIdx.x;
    const int c = tid / kLanesPerContact, k = tid % kLanesPerContact;
    const bool live = c < nc;
    int a = -1, b = -1;
    double muc = 0;
    if (live) {
        a = ida[c];
        b = idb[c];
        muc = mu[c];
    }
    const int nr = muc > 0 ? 3 : 1;
    const bool row = live && (a >= 0 || b >= 0) && k < nr;
    double d[3] = {0, 0, 0}, aA[3] = {0, 0, 0}, aB[3] = {0, 0, 0};
    double pold = 0, pt = 0;
    if (row) {
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            d[m] = J[(size_t)(3 * k + m) * nc + c];
            aA[m] = J[(size_t)(9 + 3 * k + m) * nc + c];
            aB[m] = J[(size_t)(18 + 3 * k + m) * nc + c];
        }
        double ua[6], ub[6];
        load_body(x, a, ua);
        load_body(x, b, ub);
        const double g = row_dot(d, aA, aB, ua, ub);
        pold = p[(size_t)k * nc + c];
        if constexpr (Block)
            pt = g + (use_bias ? bias[(size_t)k * nc + c] : 0.0);
        else
            pt = pold - (R[(size_t)k * nc + c] * g + (use_bias ? bias[(size_t)k * nc + c] : 0.0));
    }
    if constexpr (Block) pt = pold - block_apply4(pt, k, c, nc, nr, row, R);
    const double pnew = project4(pt, k, muc);
    double dp = 0;
    if (row) {
        dp = pnew - pold;
        p[(size_t)k * nc + c] = pnew;
    }
    write_record4(live && (a >= 0 || b >= 0), k, c, a, b, d, aA, aB, dp, imp9);
}

// Body pass with 8 lanes per body: lane l sums incidences l, l+8, ... of its body, the 8 partial
// twists are combined by a fixed DPP tree (oct_sum, deterministic), and lanes 0..5 each update one
// velocity component.  Eight times the lanes of one-lane-per-body keeps enough waves in flight to
// hide the latency of the index -> side-impulse gather chain.
constexpr int kLanesPerBody = 8;

__global__ void __launch_bounds__(kBlock) k_pj_bodies8(int nb, const int* __restrict__ dof, const double* __restrict__ minv,
                                                       const int* __restrict__ incptr, const int* __restrict__ incidx,
                                                       const double* __restrict__ imp9, double* __restrict__ x, int check,
                                                       double tol_rel, double tol_abs, double* __restrict__ partial,
                                                       const int* __restrict__ ctrl, const int* __restrict__ jslot,
                                                       double* __restrict__ tr, const double* __restrict__ sinv) {
    if (ctrl[0]) return;
    __shared__ double red[kBlock / 64];
    // XCD-aware tiles, as the contact pass: a contiguous range of bodies per XCD gathers the
    // records its own XCD's contact tiles wrote (contacts are ordered by body), from its L2
    const int tid = xcd_tile(gridDim.x) * kBlock + threadIdx.x;
    const int bi = tid / kLanesPerBody, l = tid % kLanesPerBody;
    const bool live = bi < nb;
    // the body's own values are loaded before the gather chain, so their latency overlaps it
    int nd = 0, js = -1;
    double mk = 0, u = 0;
    if (live && l < 6) {
        nd = dof[bi];
        if (jslot) js = jslot[bi];
        mk = l < 3 ? minv[(size_t)bi * kMinvStride] : minv[(size_t)bi * kMinvStride + l - 2];
        u = x[(size_t)bi * kBodyStride + l];
    }
    double acc[6] = {0, 0, 0, 0, 0, 0};
    int i0 = 0, i1 = 0;
    if (live) {
        i0 = incptr[bi];
        i1 = incptr[bi + 1];
        for (int i = i0 + l; i < i1; i += kLanesPerBody) gather_side9(imp9, incidx[i], acc);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) acc[k] = oct_sum(acc[k]);
    double e2 = 0;
    if (live && l < 6) {
        double a = acc[0];
#pragma unroll
        for (int k = 1; k < 6; ++k)
            if (l == k) a = acc[k];
        if (js >= 0) {
            if (l < nd) tr[(size_t)bi * 6 + l] = a;  // W^T dp of a jointed body: the chain solve applies it
        } else if (l < nd) {
            const double M = mk > 0 ? 1.0 / mk : 0.0;
            // S^-1: a / M, or a row of the inverse rotational block (implicit gyroscopy)
            const double d = sinv && nd == 6 && l >= 3
                                 ? gyro_row(sinv + (size_t)bi * 9 + 3 * (l - 3), acc[3], acc[4], acc[5])
                                 : a / M;
            const double un = u + d;
            if (i1 > i0) x[(size_t)bi * kBodyStride + l] = un;
            if (check) {
                const double q = fmax(fabs(un), fabs(u)) * tol_rel + tol_abs;
                const double e = (un - u) / q;
                e2 = e * e;
            }
        }
    }
    if (check) {
        const double s = block_sum(e2, red);
        if (threadIdx.x == 0) partial[blockIdx.x] = s;
    }
}

// ---------------------------------------------------------------------------------------------
// Range sweep: per-body partial sums of each workgroup's contact range, reduced in LDS
// ---------------------------------------------------------------------------------------------
//
// One workgroup takes a contiguous range of contacts.  It runs their contact pass (as
// k_pj_contacts4) with the side-impulse records in LDS instead of device memory, and then, after
// a workgroup barrier, sums for every body its range touches the records of its incidences in the
// range, and writes that sum as one 6-double partial.  The body pass (k_pj_bodies_parts) adds a
// body's partials, one per range that touches it, and updates x: together the Jacobi sweep of
// projected_jacobi.cpp:121-138.  With contacts in the order of their lower body (the contact
// generator's order) a range touches few bodies outside its own span, so a body has about two
// partials where the separate passes move one record per incidence (eight per body in a pile).
// Partials are stored body by body, ranges in ascending order, so every sum has a fixed order
// and results repeat bit for bit.
constexpr int kRangeMaxContacts = 768;     // LDS records: 9 doubles per contact, 55,296 bytes
constexpr int kRangeAutoContacts = 32768;  // range mode 1 from this many contacts
constexpr int kRangeLdsBytes = 65536;      // LDS of one range workgroup: records and tables

__host__ __device__ constexpr int pad64(int n) { return (n + 63) & ~63; }

// Copies n ints from src to the LDS array dst by LDS DMA (global_load_lds_dword: lane i of a wave
// writes dst + 4 i bytes), one 64-int chunk per wave and step; a lane past the end loads the last
// element again into padding (dst holds pad64(n) ints).  Completed by s_waitcnt + barrier.
__device__ __forceinline__ void dma_ints(const int* __restrict__ src, int n, int* dst) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, W = blockDim.x >> 6;
    for (int ch = w; ch * 64 < n; ch += W) {
        const int i = min(ch * 64 + lane, n - 1);
        __builtin_amdgcn_global_load_lds(src + i, dst + ch * 64, 4, 0, 0);
    }
}

// ---------------------------------------------------------------------------------------------
// Geometry form of the contact rows
// ---------------------------------------------------------------------------------------------
//
// A rigid side's angular row k is r_body x (R^T d_k) (buildContactRowByDof, contact_jacobian.hpp:
// 17-23, with makeContact's body-frame lever arm r_body = R^T (p - x_com)), R the body's rotation.
// Given the world contact points, k_geo_pack derives each side's r_body from the body's position and
// quaternion and rebuilds every angular row; where all of them equal the rows passed to
// cdx_set_contacts bit for bit, the range sweep rebuilds them from (r_body, quaternion, frame) in
// every sweep instead of reading them: 48 bytes of lever arms per contact in place of 144 bytes of
// angular rows, the quaternions being gathered with the twists; the second tangent is rebuilt as
// n x t1 where that holds bit for bit too.  The arithmetic is the host contact generator's,
// operation by operation and without contraction, so that the rebuilt rows are the same numbers.

// R^T v for the unit quaternion (qx, qy, qz, qw) (Eigen's coefficient order), matrix entries first.
__device__ inline void geo_rot_t(const double* __restrict__ q, const double v[3], double out[3]) {
#pragma clang fp contract(off)
    const double x = q[0], y = q[1], z = q[2], w = q[3];
    const double r00 = 1 - 2 * (y * y + z * z), r01 = 2 * (x * y - z * w), r02 = 2 * (x * z + y * w);
    const double r10 = 2 * (x * y + z * w), r11 = 1 - 2 * (x * x + z * z), r12 = 2 * (y * z - x * w);
    const double r20 = 2 * (x * z - y * w), r21 = 2 * (y * z + x * w), r22 = 1 - 2 * (x * x + y * y);
    out[0] = r00 * v[0] + r10 * v[1] + r20 * v[2];
    out[1] = r01 * v[0] + r11 * v[1] + r21 * v[2];
    out[2] = r02 * v[0] + r12 * v[1] + r22 * v[2];
}

__device__ inline void geo_cross(const double a[3], const double b[3], double out[3]) {
#pragma clang fp contract(off)
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
}

// The angular row r x (R^T d) of one side (q: the body's quaternion).
__device__ inline void geo_row(const double* __restrict__ q, const double r[3], const double d[3], double out[3]) {
    double t[3];
    geo_rot_t(q, d, t);
    geo_cross(r, t, out);
}

// One thread per contact: the lever arms r_body = R^T (p - x_com) of its rigid sides into G
// (component-major, rA then rB; NaN for a static or point-mass side, which has no angular row) and
// a count of the contacts whose rebuilt angular rows or second tangent n x t1 differ from J's in any
// bit, per workgroup.
__global__ void __launch_bounds__(kBlock) k_geo_pack(int nc, const int* __restrict__ ida, const int* __restrict__ idb,
                                                     const int* __restrict__ dof, const double* __restrict__ q,
                                                     const double* __restrict__ J, const double* __restrict__ point,
                                                     double* __restrict__ G, int* __restrict__ bad_part) {
    __shared__ double red[kBlock / 64];
    const int c = blockIdx.x * kBlock + threadIdx.x;
    double bad = 0;
    if (c < nc) {
        for (int side = 0; side < 2; ++side) {
            const int b = side == 0 ? ida[c] : idb[c];
            double r[3] = {NAN, NAN, NAN};
            if (b >= 0 && dof[b] == 6) {
                const double* qb = q + (size_t)b * kBodyStride;
                double rel[3];
                {
#pragma clang fp contract(off)
                    rel[0] = point[3 * (size_t)c] - qb[0];
                    rel[1] = point[3 * (size_t)c + 1] - qb[1];
                    rel[2] = point[3 * (size_t)c + 2] - qb[2];
                }
                geo_rot_t(qb + 3, rel, r);
#pragma unroll 1
                for (int k = 0; k < 3; ++k) {
                    double d[3], a[3];
                    for (int m = 0; m < 3; ++m) d[m] = J[(size_t)(3 * k + m) * nc + c];
                    geo_row(qb + 3, r, d, a);
                    for (int m = 0; m < 3; ++m)
                        if (__double_as_longlong(a[m]) != __double_as_longlong(J[(size_t)(9 + 9 * side + 3 * k + m) * nc + c]))
                            bad = 1;
                }
            }
            for (int m = 0; m < 3; ++m) G[(size_t)(3 * side + m) * nc + c] = r[m];
        }
        // the second tangent is rebuilt as n x t1
        double n[3], t1[3], t2[3];
        for (int m = 0; m < 3; ++m) {
            n[m] = J[(size_t)m * nc + c];
            t1[m] = J[(size_t)(3 + m) * nc + c];
        }
        geo_cross(n, t1, t2);
        for (int m = 0; m < 3; ++m)
            if (__double_as_longlong(t2[m]) != __double_as_longlong(J[(size_t)(6 + m) * nc + c])) bad = 1;
    }
    const double s = block_sum(bad, red);
    if (threadIdx.x == 0) bad_part[blockIdx.x] = (int)s;
}

// The Nesterov extrapolation y = u + beta (u - uk) (u itself while beta is 0), as nesterov_loop
// forms it (projected_jacobi.cpp:379-449).
__device__ inline double extrapolate1(double u, double uk, double beta) {
    return beta == 0.0 ? u : u + beta * (u - uk);
}
__device__ inline void extrapolate6(const double u[6], const double uk[6], double beta, double out[6]) {
#pragma unroll
    for (int k = 0; k < 6; ++k) out[k] = extrapolate1(u[k], uk[k], beta);
}

// Loops the range pass serves: the standard loop, a Chebyshev sweep, a Nesterov sweep.
constexpr int kLoopStd = 0, kLoopCheb = 1, kLoopNest = 2;

// project() (projected_jacobi.cpp:100-119) of one contact's trial values z (nr = 1 or 3 rows).
__device__ __forceinline__ void project3(double z[3], int nr, double muc) {
    z[0] = fmax(z[0], 0.0);
    if (nr == 3) {
        const double tn2 = z[1] * z[1] + z[2] * z[2];
        if (tn2 > 0) {
            const double sc = fmin(1.0, (muc * z[0]) / sqrt(tn2));
            z[1] = sc * z[1];
            z[2] = sc * z[2];
        }
    }
}

// One lane per contact and one lane per (body, range) pair: the pass issues few instructions per
// contact (no cross-lane sums), which is what bounds it -- with four lanes per contact the DPP
// exchanges of the records and of the projection took most of its instruction issue.
//
// Block = true: R holds the inverse Delassus blocks and the bias is unscaled
// (pj.rdiag_true_delassus): dp = Pi(p - Rb (W x + b)) - p.  Geo = true: the angular rows are
// rebuilt from the lever arms G and the quaternions q (the geometry form above) and the second
// tangent as n x t1 instead of being read from J.  Loop = kLoopCheb (row scaling only): a sweep of
// the Chebyshev loop (chebyshev_loop, projected_jacobi.cpp:180-228) -- its input is the
// extrapolation w (x - xk) + xk of the twists and w (p - pk) + pk of the impulses (or x, p after a
// reset: scal[1]), formed here as k_pjc_contacts4 forms it, with the impulse history rotated
// (pk <- pk1, pk1 <- input); the twists' history is rotated by k_pjc_bodies_parts.  Loop =
// kLoopNest (row scaling only): a sweep of the Nesterov loop from y = x + beta (x - xk),
// py = p + beta (p - pk) (beta = scal[0]), with pk <- p, as k_pjn_contacts; k_pjn_bodies_parts
// rotates the twists.
template <bool Block, bool Geo, int Loop>
__global__ void __launch_bounds__(256, 4) k_pj_range(
    int nranges, int nc, int per, int maxs, int maxp, const int* __restrict__ rg_p, const int* __restrict__ rg_l,
    const int* __restrict__ pr_ptr, const int* __restrict__ pr_out, const int* __restrict__ pr_inc,
    const int* __restrict__ ida, const int* __restrict__ idb, const double* __restrict__ J,
    const double* __restrict__ R, const double* __restrict__ bias, int use_bias, const double* __restrict__ mu,
    double* __restrict__ p, const double* __restrict__ x, double* __restrict__ part, const int* __restrict__ ctrl,
    const double* __restrict__ G, const double* __restrict__ q, double* __restrict__ PK, double* __restrict__ PK1,
    const double* __restrict__ XK, const double* __restrict__ scal) {
    static_assert(!Block || Loop == kLoopStd, "the range form of the accelerated loops takes the row scaling");
    if (ctrl[0]) return;
    double w = 1.0, beta = 0.0;
    bool take = true;
    if constexpr (Loop == kLoopCheb) {
        w = scal[0];
        take = scal[1] != 0.0;
    }
    if constexpr (Loop == kLoopNest) beta = scal[0];
    // LDS: the records (9 doubles per contact of the range: lin | angA | angB, summed over its
    // rows), then the range's tables, copied by DMA while the contacts are swept: the incidence
    // offsets of its pairs (body, range), their partial slots, and the incidences (contact in the
    // range << 1 | side)
    extern __shared__ double lrec[];
    const int g = xcd_tile(nranges);
    const int s0 = g * per, s1 = min(nc, s0 + per);
    const int q0 = rg_p[g], q1 = rg_p[g + 1], l0 = rg_l[g], l1 = rg_l[g + 1];
    int* Lptr = reinterpret_cast<int*>(lrec + 9 * (size_t)maxs);
    int* Lout = Lptr + pad64(maxp + 1);
    int* Linc = Lout + pad64(maxp);
    dma_ints(pr_ptr + q0, q1 - q0 + 1, Lptr);
    dma_ints(pr_out + q0, q1 - q0, Lout);
    dma_ints(pr_inc + l0, l1 - l0, Linc);

#pragma unroll 1
    for (int c = s0 + threadIdx.x; c < s1; c += blockDim.x) {
        const int a = ida[c], b = idb[c];
        const double muc = mu[c];
        const int nr = muc > 0 ? 3 : 1;
        double rec[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (a >= 0 || b >= 0) {
            double d[3][3], aA[3][3], aB[3][3];
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    if (!Geo || k < 2) d[k][m] = J[(size_t)(3 * k + m) * nc + c];
                    if constexpr (!Geo) {
                        aA[k][m] = J[(size_t)(9 + 3 * k + m) * nc + c];
                        aB[k][m] = J[(size_t)(18 + 3 * k + m) * nc + c];
                    } else {
                        aA[k][m] = 0.0;
                        aB[k][m] = 0.0;
                    }
                }
            if constexpr (Geo) {
                geo_cross(d[0], d[1], d[2]);  // t2 = n x t1 (checked bit for bit by k_geo_pack)
                double rA[3], rB[3];
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    rA[m] = G[(size_t)m * nc + c];
                    rB[m] = G[(size_t)(3 + m) * nc + c];
                }
                if (rA[0] == rA[0])
#pragma unroll
                    for (int k = 0; k < 3; ++k) geo_row(q + (size_t)a * kBodyStride + 3, rA, d[k], aA[k]);
                if (rB[0] == rB[0])
#pragma unroll
                    for (int k = 0; k < 3; ++k) geo_row(q + (size_t)b * kBodyStride + 3, rB, d[k], aB[k]);
            }
            double ua[6], ub[6];
            if constexpr (Loop == kLoopCheb) {
                cheb_input6(x, XK, a, w, take, ua);
                cheb_input6(x, XK, b, w, take, ub);
            } else if constexpr (Loop == kLoopNest) {
                double t0[6], t1[6];
                load_body(x, a, t0);
                load_body(XK, a, t1);
                extrapolate6(t0, t1, beta, ua);
                load_body(x, b, t0);
                load_body(XK, b, t1);
                extrapolate6(t0, t1, beta, ub);
            } else {
                load_body(x, a, ua);
                load_body(x, b, ub);
            }
            double pold[3] = {0, 0, 0}, z[3] = {0, 0, 0};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if (k >= nr) break;
                pold[k] = p[(size_t)k * nc + c];
                if constexpr (Loop == kLoopCheb) {  // the sweep's input impulse, and the history rotated
                    const size_t r = (size_t)k * nc + c;
                    pold[k] = take ? pold[k] : cheb_extrap(w, pold[k], PK[r]);
                    PK[r] = PK1[r];
                    PK1[r] = pold[k];
                }
                if constexpr (Loop == kLoopNest) {
                    const size_t r = (size_t)k * nc + c;
                    const double pc = pold[k];
                    pold[k] = extrapolate1(pc, PK[r], beta);
                    PK[r] = pc;
                }
                const double gk = row_dot(d[k], aA[k], aB[k], ua, ub);
                const double bk = use_bias ? bias[(size_t)k * nc + c] : 0.0;
                if constexpr (Block)
                    z[k] = gk + bk;
                else
                    z[k] = pold[k] - (R[(size_t)k * nc + c] * gk + bk);
            }
            if constexpr (Block) {  // p - R_block t, in row order (BlockDiagonal::applyTo)
                const double t0 = z[0], t1 = z[1], t2 = z[2];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    if (k >= nr) break;
                    const double* r = R + (size_t)(3 * k) * nc + c;
                    z[k] = pold[k] - (nr == 1 ? r[0] * t0 : r[0] * t0 + r[nc] * t1 + r[2 * (size_t)nc] * t2);
                }
            }
            project3(z, nr, muc);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if (k >= nr) break;
                p[(size_t)k * nc + c] = z[k];
                const double dp = z[k] - pold[k];
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    rec[m] += d[k][m] * dp;
                    rec[3 + m] += aA[k][m] * dp;  // zero for a static side
                    rec[6 + m] += aB[k][m] * dp;
                }
            }
        }
        double* l = lrec + 9 * (size_t)(c - s0);
#pragma unroll
        for (int m = 0; m < 9; ++m) l[m] = rec[m];
    }
    __builtin_amdgcn_s_waitcnt(0);  // this wave's table DMA has landed
    __syncthreads();                // ... and every wave's records and DMA
    // One partial per (body, range) pair: the pair's incidences in order, one lane per pair.
    const int npairs = q1 - q0;
#pragma unroll 1
    for (int j = threadIdx.x; j < npairs; j += blockDim.x) {
        const int i0 = Lptr[j] - l0, i1 = Lptr[j + 1] - l0, out = Lout[j];
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (int i = i0; i < i1; ++i) {
            const int e = Linc[i];
            const double* r = lrec + 9 * (e >> 1);
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
        double2* o = reinterpret_cast<double2*>(part + (size_t)out * 6);
        o[0] = make_double2(acc[0], acc[1]);
        o[1] = make_double2(acc[2], acc[3]);
        o[2] = make_double2(acc[4], acc[5]);
    }
}

// Body pass over the partials.  Body b's first kPartSlots partials sit in slots b kPartSlots + j
// and any further ones in an overflow area (ovf[b] + j - kPartSlots), so a body's count, twist,
// inverse mass and first partials are all loaded at once.  Lane l < 6 of a body's 8 adds
// component l of its partials in order and updates x (x += (W^T dp) / M), with the residual
// partials every check sweep.
constexpr int kPartSlots = 3;

__global__ void __launch_bounds__(kBlock) k_pj_bodies_parts(int nb, const int* __restrict__ dof,
                                                            const double* __restrict__ minv,
                                                            const int* __restrict__ pcnt, const int* __restrict__ povf,
                                                            const double* __restrict__ part, double* __restrict__ x,
                                                            int check, double tol_rel, double tol_abs,
                                                            double* __restrict__ partial, const int* __restrict__ ctrl,
                                                            const double* __restrict__ sinv) {
    if (ctrl[0]) return;
    __shared__ double red[kBlock / 64];
    const int tid = xcd_tile(gridDim.x) * kBlock + threadIdx.x;
    const int bi = tid / kLanesPerBody, l = tid % kLanesPerBody;
    double e2 = 0;
    if (bi < nb && l < 6) {
        const int nd = dof[bi], n = pcnt[bi];
        const double mk = l < 3 ? minv[(size_t)bi * kMinvStride] : minv[(size_t)bi * kMinvStride + l - 2];
        const double u = x[(size_t)bi * kBodyStride + l];
        double ps[kPartSlots];
#pragma unroll
        for (int j = 0; j < kPartSlots; ++j) ps[j] = part[((size_t)bi * kPartSlots + j) * 6 + l];
        if (l < nd) {
            double a = 0;
#pragma unroll
            for (int j = 0; j < kPartSlots; ++j)
                if (j < n) a += ps[j];
            if (n > kPartSlots) {
                const size_t o = (size_t)nb * kPartSlots + povf[bi];
                for (int j = kPartSlots; j < n; ++j) a += part[(o + j - kPartSlots) * 6 + l];
            }
            const double M = mk > 0 ? 1.0 / mk : 0.0;
            double d = a / M;
            if (sinv && nd == 6 && l >= 3) {
                // implicit gyroscopy: a row of the inverse rotational block; the lane sums the three
                // rotational components itself, in the order every lane sums its own
                double a3[3] = {0, 0, 0};
                for (int j = 0; j < n; ++j) {
                    const size_t o = j < kPartSlots ? ((size_t)bi * kPartSlots + j) * 6
                                                    : ((size_t)nb * kPartSlots + povf[bi] + j - kPartSlots) * 6;
#pragma unroll
                    for (int k = 0; k < 3; ++k) a3[k] += part[o + 3 + k];
                }
                d = gyro_row(sinv + (size_t)bi * 9 + 3 * (l - 3), a3[0], a3[1], a3[2]);
            }
            const double un = u + d;
            if (n > 0) x[(size_t)bi * kBodyStride + l] = un;
            if (check) {
                const double q = fmax(fabs(un), fabs(u)) * tol_rel + tol_abs;
                const double ee = (un - u) / q;
                e2 = ee * ee;
            }
        }
    }
    if (check) {
        const double s = block_sum(e2, red);
        if (threadIdx.x == 0) partial[blockIdx.x] = s;
    }
}

// Body pass of a Chebyshev range sweep (as k_pjc_bodies8 over the partials): input
// u = w (X - XK) + XK (or X), XK <- XK1, XK1 <- u, X <- u + S^-1 (sum of the body's partials); per-block
// sums of the residual terms (every sweep) and of non-finite entries of X.
__global__ void __launch_bounds__(kBlock) k_pjc_bodies_parts(int nb, const int* __restrict__ dof,
                                                             const double* __restrict__ minv, const int* __restrict__ pcnt,
                                                             const int* __restrict__ povf, const double* __restrict__ part,
                                                             double* __restrict__ X, double* __restrict__ XK,
                                                             double* __restrict__ XK1, double tol_rel, double tol_abs,
                                                             double* __restrict__ partial, const int* __restrict__ ctrl,
                                                             const double* __restrict__ cheb,
                                                             const double* __restrict__ sinv) {
    if (ctrl[0]) return;
    __shared__ double red[kBlock / 64];
    const double w = cheb[0];
    const bool take = cheb[1] != 0.0;
    const int tid = xcd_tile(gridDim.x) * kBlock + threadIdx.x;
    const int bi = tid / kLanesPerBody, l = tid % kLanesPerBody;
    double e2 = 0, bad = 0;
    if (bi < nb && l < 6) {
        const int nd = dof[bi], n = pcnt[bi];
        const double mk = l < 3 ? minv[(size_t)bi * kMinvStride] : minv[(size_t)bi * kMinvStride + l - 2];
        const size_t o = (size_t)bi * kBodyStride + l;
        const double xs = X[o];
        double ps[kPartSlots];
#pragma unroll
        for (int j = 0; j < kPartSlots; ++j) ps[j] = part[((size_t)bi * kPartSlots + j) * 6 + l];
        if (l < nd) {
            const double u = take ? xs : cheb_extrap(w, xs, XK[o]);
            XK[o] = XK1[o];
            XK1[o] = u;
            double a = 0;
#pragma unroll
            for (int j = 0; j < kPartSlots; ++j)
                if (j < n) a += ps[j];
            if (n > kPartSlots) {
                const size_t ov = (size_t)nb * kPartSlots + povf[bi];
                for (int j = kPartSlots; j < n; ++j) a += part[(ov + j - kPartSlots) * 6 + l];
            }
            const double M = mk > 0 ? 1.0 / mk : 0.0;
            double d = a / M;
            if (sinv && nd == 6 && l >= 3) {  // implicit gyroscopy, as k_pj_bodies_parts
                double a3[3] = {0, 0, 0};
                for (int j = 0; j < n; ++j) {
                    const size_t oo = j < kPartSlots ? ((size_t)bi * kPartSlots + j) * 6
                                                     : ((size_t)nb * kPartSlots + povf[bi] + j - kPartSlots) * 6;
#pragma unroll
                    for (int k = 0; k < 3; ++k) a3[k] += part[oo + 3 + k];
                }
                d = gyro_row(sinv + (size_t)bi * 9 + 3 * (l - 3), a3[0], a3[1], a3[2]);
            }
            const double un = n > 0 ? u + d : u;
            X[o] = un;
            const double qq = fmax(fabs(un), fabs(u)) * tol_rel + tol_abs;
            const double e = (un - u) / qq;
            e2 = e * e;
            bad = isfinite(un) ? 0.0 : 1.0;
        }
    }
    const double s = block_sum(e2, red);
    const double nf = block_sum(bad, red);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s;
        partial[gridDim.x + blockIdx.x] = nf;
    }
}

// Body pass of a Nesterov range sweep (as k_pjn_bodies over the partials), one lane per DOF:
// y = X + beta (X - XK), X <- y + S^-1 (sum of the body's partials), XK <- the old X; per-block sums
// of the residual terms between the new X and y and of the restart dot (y - X_new) . (X_new - X_old).
__global__ void __launch_bounds__(kBlock) k_pjn_bodies_parts(int nb, const int* __restrict__ dof,
                                                             const double* __restrict__ minv, const int* __restrict__ pcnt,
                                                             const int* __restrict__ povf, const double* __restrict__ part,
                                                             double* __restrict__ X, double* __restrict__ XK,
                                                             double tol_rel, double tol_abs, double* __restrict__ partial,
                                                             const int* __restrict__ ctrl, const double* __restrict__ nest,
                                                             const double* __restrict__ sinv) {
    if (ctrl[0]) return;
    __shared__ double red[kBlock / 64];
    const double beta = nest[0];
    const int tid = xcd_tile(gridDim.x) * kBlock + threadIdx.x;
    const int bi = tid / kLanesPerBody, l = tid % kLanesPerBody;
    double e2 = 0, dt = 0;
    if (bi < nb && l < 6) {
        const int nd = dof[bi], n = pcnt[bi];
        const double mk = l < 3 ? minv[(size_t)bi * kMinvStride] : minv[(size_t)bi * kMinvStride + l - 2];
        const size_t o = (size_t)bi * kBodyStride + l;
        const double u = X[o], uk = XK[o];
        double ps[kPartSlots];
#pragma unroll
        for (int j = 0; j < kPartSlots; ++j) ps[j] = part[((size_t)bi * kPartSlots + j) * 6 + l];
        if (l < nd) {
            const double y = extrapolate1(u, uk, beta);
            double a = 0;
#pragma unroll
            for (int j = 0; j < kPartSlots; ++j)
                if (j < n) a += ps[j];
            if (n > kPartSlots) {
                const size_t ov = (size_t)nb * kPartSlots + povf[bi];
                for (int j = kPartSlots; j < n; ++j) a += part[(ov + j - kPartSlots) * 6 + l];
            }
            const double M = mk > 0 ? 1.0 / mk : 0.0;
            double d = a / M;
            if (sinv && nd == 6 && l >= 3) {  // implicit gyroscopy, as k_pj_bodies_parts
                double a3[3] = {0, 0, 0};
                for (int j = 0; j < n; ++j) {
                    const size_t oo = j < kPartSlots ? ((size_t)bi * kPartSlots + j) * 6
                                                     : ((size_t)nb * kPartSlots + povf[bi] + j - kPartSlots) * 6;
#pragma unroll
                    for (int k = 0; k < 3; ++k) a3[k] += part[oo + 3 + k];
                }
                d = gyro_row(sinv + (size_t)bi * 9 + 3 * (l - 3), a3[0], a3[1], a3[2]);
            }
            const double un = y + d;
            const double qq = fmax(fabs(un), fabs(y)) * tol_rel + tol_abs;
            const double e = (un - y) / qq;
            e2 = e * e;
            dt = (y - un) * (un - u);
            XK[o] = u;
            X[o] = un;
        }
    }
    const double s = block_sum(e2, red);
    const double d = block_sum(dt, red);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s;
        partial[gridDim.x + blockIdx.x] = d;
    }
}

// mode 0: PJ (break when res <= 1); mode 1: GS (break when res < 1, NaN/Inf -> diverged).
// it < 0: the sweep index comes from the device chunk counter ctrl[3] (graph-replayed chunks of
// `cadence` sweeps whose residual check is the chunk's first sweep); the counter advances on
// every call.
__global__ void __launch_bounds__(kBlock) k_residual_finalize(int nparts, const double* __restrict__ partial, int nv,
                                                              int it, int cadence, int mode, int* __restrict__ ctrl,
                                                              double* __restrict__ res, int* __restrict__ hflag) {
    if (ctrl[0]) return;
    __shared__ double red[kBlock / 64];
    double s = 0;
    s = strided_sum(partial, nparts);
    s = block_sum(s, red);
    if (threadIdx.x == 0) {
        if (it < 0) it = ctrl[3] * cadence;
        ctrl[3] = ctrl[3] + 1;
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

int launch_residual_finalize(cdx_ctx* ctx, int nparts, int it, int mode, int cadence) {
    CDX_LAUNCH(ctx, mode == 0 ? "pj_residual" : "gs_residual", k_residual_finalize, 1, kBlock, nparts, ctx->d_partial, ctx->nv,
               it, cadence, mode, ctx->d_ctrl, ctx->d_res, ctx->h_flag_dev);
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


// ---------------------------------------------------------------------------------------------
// Nesterov-accelerated loop (nesterov_loop, projected_jacobi.cpp:379-449), fused into the sweep.
//
// Buffers after iteration i: X = xk1_i, XK = xk_i, P = pk1_i, PK = pk_i, and beta_i from the
// finalize kernel (0 on a restart or once momentum is disabled).  Iteration i+1 forms
// xy = X + beta (X - XK) and py = P + beta (P - PK) on the fly, sweeps from (xy, py), and
// rotates XK <- X, PK <- P.  The residual compares xk1 with xy and the restart test uses
// (xy - xk1) . (xk1 - xk) over the velocity entries, exactly as the reference loop does.
// ---------------------------------------------------------------------------------------------

// nest[]: 0 beta, 1 thk, 2 err_prev, 3 restarts, 4 momentum_disabled

__global__ void __launch_bounds__(kBlock) k_pjn_contacts(int nc, int nblocks, const int* __restrict__ ida,
                                                         const int* __restrict__ idb, const double* __restrict__ J,
                                                         const double* __restrict__ R, const double* __restrict__ bias,
                                                         const double* __restrict__ mu, double* __restrict__ p,
                                                         double* __restrict__ pk, const double* __restrict__ x,
                                                         const double* __restrict__ xk, double* __restrict__ imp,
                                                         const int* __restrict__ ctrl, const double* __restrict__ nest,
                                                         const double* __restrict__ Rb) {
    if (ctrl[0]) return;
    const double beta = nest[0];
    const int c = xcd_tile(nblocks) * kBlock + threadIdx.x;
    if (c >= nc) return;
    const int a = ida[c], b = idb[c];
    if (a < 0 && b < 0) return;
    const double muc = mu[c];
    const int nr = muc > 0 ? 3 : 1;
    ContactJ j;
    load_J(J, nc, c, nr, j);
    double ua[6], ub[6], t0[6], t1[6];
    load_body(x, a, t0);
    load_body(xk, a, t1);
    extrapolate6(t0, t1, beta, ua);
    load_body(x, b, t0);
    load_body(xk, b, t1);
    extrapolate6(t0, t1, beta, ub);
    double py[3] = {0, 0, 0}, pn[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= nr) break;
        const double pc = p[(size_t)k * nc + c], pp = pk[(size_t)k * nc + c];
        py[k] = beta == 0.0 ? pc : pc + beta * (pc - pp);
        pk[(size_t)k * nc + c] = pc;
        const double g = row_dot(j.d[k], j.aA[k], j.aB[k], ua, ub);
        pn[k] = Rb ? g + bias[(size_t)k * nc + c] : py[k] - (R[(size_t)k * nc + c] * g + bias[(size_t)k * nc + c]);
    }
    if (Rb) {  // p_y - R_block (W y + b)
        double t[3] = {pn[0], pn[1], pn[2]};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (k >= nr) break;
            const double* r = Rb + (size_t)(3 * k) * nc + c;
            pn[k] = py[k] - (nr == 1 ? r[0] * t[0] : r[0] * t[0] + r[nc] * t[1] + r[2 * (size_t)nc] * t[2]);
        }
    }
    pn[0] = fmax(pn[0], 0.0);
    if (nr == 3) {
        const double a1 = pn[1], a2 = pn[2];
        const double tn2 = a1 * a1 + a2 * a2;
        if (tn2 > 0) {
            const double s = fmin(1.0, (muc * pn[0]) / sqrt(tn2));
            pn[1] = s * a1;
            pn[2] = s * a2;
        }
    }
    double dp[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= nr) break;
        dp[k] = pn[k] - py[k];
        p[(size_t)k * nc + c] = pn[k];
    }
    write_side_impulses(imp, c, a, b, nr, j, dp);
}

__global__ void __launch_bounds__(kBlock) k_pjn_bodies(int nb, const int* __restrict__ dof,
                                                       const double* __restrict__ minv, const int* __restrict__ incptr,
                                                       const int* __restrict__ incidx, const double* __restrict__ imp,
                                                       double* __restrict__ x, double* __restrict__ xk, double tol_rel,
                                                       double tol_abs, double* __restrict__ partial,
                                                       const int* __restrict__ ctrl, const double* __restrict__ nest,
                                                       const double* __restrict__ sinv) {
    if (ctrl[0]) return;
    __shared__ double red[kBlock / 64];
    const double beta = nest[0];
    const int bi = blockIdx.x * kBlock + threadIdx.x;
    double e2 = 0, dt = 0;
    if (bi < nb) {
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
        double u[6], uk[6], y[6], un[6];
        load_body(x, bi, u);
        load_body(xk, bi, uk);
        extrapolate6(u, uk, beta, y);
        const double4 m4 = *reinterpret_cast<const double4*>(minv + (size_t)bi * kMinvStride);
        const int nd = dof[bi];
        const double Ml = m4.x > 0 ? 1.0 / m4.x : 0.0;
        const double Ma[3] = {m4.y > 0 ? 1.0 / m4.y : 0.0, m4.z > 0 ? 1.0 / m4.z : 0.0, m4.w > 0 ? 1.0 / m4.w : 0.0};
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const double M = k < 3 ? Ml : Ma[k - 3];
            un[k] = (k < nd) ? y[k] + acc[k] / M : y[k];
        }
        if (sinv && nd == 6) {  // implicit gyroscopy: the inverse rotational block
            const double* K = sinv + (size_t)bi * 9;
#pragma unroll
            for (int k = 0; k < 3; ++k) un[3 + k] = y[3 + k] + gyro_row(K + 3 * k, acc[3], acc[4], acc[5]);
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            if (k >= nd) break;
            const double q = fmax(fabs(un[k]), fabs(y[k])) * tol_rel + tol_abs;
            const double e = (un[k] - y[k]) / q;
            e2 += e * e;
            dt += (y[k] - un[k]) * (un[k] - u[k]);
        }
        store6(xk + (size_t)bi * kBodyStride, u);
        store6(x + (size_t)bi * kBodyStride, un);
    }
    const double s = block_sum(e2, red);
    const double d = block_sum(dt, red);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s;
        partial[gridDim.x + blockIdx.x] = d;
    }
}

__global__ void __launch_bounds__(kBlock) k_pjn_finalize(int nparts, const double* __restrict__ partial, int nv, int it,
                                                         double beta_threshold, int restart_limit, int* __restrict__ ctrl,
                                                         double* __restrict__ res, int* __restrict__ hflag,
                                                         double* __restrict__ nest) {
    if (ctrl[0]) return;
    __shared__ double red[kBlock / 64];
    double s = strided_sum(partial, nparts), d = strided_sum(partial + nparts, nparts);
    s = block_sum(s, red);
    d = block_sum(d, red);
    if (threadIdx.x != 0) return;
    if (it < 0) it = ctrl[3];  // graph-replayed chunk: the sweep index from the device counter
    ctrl[3] = ctrl[3] + 1;
    const double err = 1.0 / sqrt((double)nv) * sqrt(s);
    res[0] = err;
    if (err <= 1.0) {
        ctrl[1] = it + 1;
        ctrl[2] = 0;
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
    if (!isfinite(err)) restart = true;
    else if (err_prev < INFINITY && err > 1.05 * err_prev) restart = true;
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

// ---------------------------------------------------------------------------------------------
// Chebyshev-accelerated loop (chebyshev_loop, projected_jacobi.cpp:180-228) and its spectral
// radius estimate (estimateSpectralRadius, :155-178).
//
// Spectral radius: 8 normalised power iterations of the linear sweep map v -> v - M^-1 W^T R W v
// from ones/sqrt(nv), each one linear contact pass (dp = -R W v, no projection, no bias), one body
// pass (v' = v + W^T dp / M and per-block sums of v'^2), a one-workgroup norm and a scaling pass.
// A norm <= 1e-300 (or NaN) ends the estimate at rho = 0; otherwise rho = min(lambda_8, 0.995).
//
// Loop: the reference keeps xk, xk1 (the sweep input) and x = sweep(xk1), and after a sweep that
// did not converge sets xk2 = w (x - xk) + xk (or xk2 = x on a fallback), xk <- xk1, xk1 <- xk2.
// Here that update is fused into the next sweep: with X = the swept iterate, XK = xk and XK1 =
// the previous input, the contact and body passes form the new input u = w (X - XK) + XK (u = X
// on a fallback, and on the first sweep) on the fly, the body pass rotates XK <- XK1, XK1 <- u
// and stores X <- sweep(u), and the contact pass does the same for P / PK / PK1.  The residual
// (every sweep) compares X with u.  The finalize kernel ends the loop (res <= 1) or computes the
// next w; the finish kernels return XK1 / PK1 after convergence (the reference returns the sweep
// input) and apply the pending update after max_iter sweeps.
//
// cheb[]: 0 w, 1 take the swept iterate as the next input (1 on the first sweep and after a
// fallback), 2 rho^2, 3 rho, 4 power iteration stopped, 5 lambda, 6 current norm.
// ---------------------------------------------------------------------------------------------


__global__ void k_cheb_fill(int nb, const int* __restrict__ dof, double inv, double* __restrict__ v) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb * 6) return;
    const int bi = t / 6, l = t % 6;
    v[(size_t)bi * kBodyStride + l] = l < dof[bi] ? inv : 0.0;
}

// linear contact pass of the power iteration: dp_k = -R_k (W_k v)
__global__ void __launch_bounds__(kBlock) k_cheb_lin_contacts4(int nc, int nblocks, const int* __restrict__ ida,
                                                               const int* __restrict__ idb, const double* __restrict__ J,
                                                               const double* __restrict__ R, const double* __restrict__ mu,
                                                               const double* __restrict__ v, double* __restrict__ imp9,
                                                               const double* __restrict__ cheb, const double* __restrict__ Rb) {
    if (cheb[4] != 0.0) return;
    const int tid = xcd_tile(nblocks) * kBlock + threadIdx.x;
    const int c = tid / kLanesPerContact, k = tid % kLanesPerContact;
    const bool live = c < nc;
    int a = -1, b = -1;
    double muc = 0;
    if (live) {
        a = ida[c];
        b = idb[c];
        muc = mu[c];
    }
    const int nr = muc > 0 ? 3 : 1;
    const bool row = live && (a >= 0 || b >= 0) && k < nr;
    double d[3] = {0, 0, 0}, aA[3] = {0, 0, 0}, aB[3] = {0, 0, 0};
    double dp = 0;
    if (row) {
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            d[m] = J[(size_t)(3 * k + m) * nc + c];
            aA[m] = J[(size_t)(9 + 3 * k + m) * nc + c];
            aB[m] = J[(size_t)(18 + 3 * k + m) * nc + c];
        }
        double ua[6], ub[6];
        load_body(v, a, ua);
        load_body(v, b, ub);
        dp = row_dot(d, aA, aB, ua, ub);
        if (!Rb) dp = -(R[(size_t)k * nc + c] * dp);
    }
    if (Rb) dp = -block_apply4(dp, k, c, nc, nr, row, Rb);
    write_record4(live && (a >= 0 || b >= 0), k, c, a, b, d, aA, aB, dp, imp9);
}

// Gathered W^T dp of body bi into acc (lanes combined), as in the sweep's body pass.
__device__ __forceinline__ void gather_body8(int bi, int l, bool live, const int* __restrict__ incptr,
                                             const int* __restrict__ incidx, const double* __restrict__ imp9,
                                             double acc[6], bool* touched) {
    int i0 = 0, i1 = 0;
    if (live) {
        i0 = incptr[bi];
        i1 = incptr[bi + 1];
        for (int i = i0 + l; i < i1; i += kLanesPerBody) gather_side9(imp9, incidx[i], acc);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) acc[k] = oct_sum(acc[k]);
    *touched = i1 > i0;
}

__device__ __forceinline__ double lane_value(const double acc[6], int l) {
    double a = acc[0];
#pragma unroll
    for (int k = 1; k < 6; ++k)
        if (l == k) a = acc[k];
    return a;
}

__device__ __forceinline__ double lane_mass(const double* __restrict__ minv, int bi, int l) {
    const double mk = l < 3 ? minv[(size_t)bi * kMinvStride] : minv[(size_t)bi * kMinvStride + l - 2];
    return mk > 0 ? 1.0 / mk : 0.0;
}

// Component l of S^-1 acc for a body of nd DOFs: acc_l / M_l, or with implicit gyroscopy (sinv) a
// row of a rigid body's inverse rotational block.
__device__ __forceinline__ double lane_sinv(const double acc[6], int l, int nd, int bi, const double* __restrict__ minv,
                                            const double* __restrict__ sinv) {
    if (sinv && nd == 6 && l >= 3) return gyro_row(sinv + (size_t)bi * 9 + 3 * (l - 3), acc[3], acc[4], acc[5]);
    return lane_value(acc, l) / lane_mass(minv, bi, l);
}

// linear body pass: v <- v + W^T dp / M, per-block sums of the new v^2
__global__ void __launch_bounds__(kBlock) k_cheb_lin_bodies8(int nb, const int* __restrict__ dof, const double* __restrict__ minv,
                                                             const int* __restrict__ incptr, const int* __restrict__ incidx,
                                                             const double* __restrict__ imp9, double* __restrict__ v,
                                                             double* __restrict__ partial, const double* __restrict__ cheb,
                                                             const double* __restrict__ sinv) {
    if (cheb[4] != 0.0) return;
    __shared__ double red[kBlock / 64];
    const int tid = xcd_tile(gridDim.x) * kBlock + threadIdx.x;
    const int bi = tid / kLanesPerBody, l = tid % kLanesPerBody;
    const bool live = bi < nb;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    bool touched = false;
    gather_body8(bi, l, live, incptr, incidx, imp9, acc, &touched);
    double s2 = 0;
    const int nd = live ? dof[bi] : 0;
    if (live && l < nd) {
        double* vp = v + (size_t)bi * kBodyStride + l;
        const double un = *vp + lane_sinv(acc, l, nd, bi, minv, sinv);
        *vp = un;
        s2 = un * un;
    }
    const double s = block_sum(s2, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kBlock) k_cheb_norm(int nparts, const double* __restrict__ partial, double* __restrict__ cheb) {
    if (cheb[4] != 0.0) return;
    __shared__ double red[kBlock / 64];
    double s = 0;
    s = strided_sum(partial, nparts);
    s = block_sum(s, red);
    if (threadIdx.x == 0) {
        const double n = sqrt(s);
        if (!(n > 1e-300)) {
            cheb[4] = 1.0;  // nothing to accelerate: rho = 0
            cheb[5] = 0.0;
        } else {
            cheb[5] = n;
            cheb[6] = n;
        }
    }
}

__global__ void k_cheb_scale(int nb, const int* __restrict__ dof, double* __restrict__ v, const double* __restrict__ cheb) {
    if (cheb[4] != 0.0) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb * 6) return;
    const int bi = t / 6, l = t % 6;
    if (l < dof[bi]) v[(size_t)bi * kBodyStride + l] = v[(size_t)bi * kBodyStride + l] / cheb[6];
}

__global__ void k_cheb_start(double* __restrict__ cheb) {
    const double rho = cheb[4] != 0.0 ? 0.0 : fmin(cheb[5], 0.995);
    cheb[0] = 1.0;
    cheb[1] = 1.0;
    cheb[2] = rho * rho;
    cheb[3] = rho;
}

// Contact pass of a Chebyshev sweep: input p = w (P - PK) + PK (or P), rotation PK <- PK1,
// PK1 <- input, then the projected update of the plain sweep from the input velocities.
__global__ void __launch_bounds__(kBlock) k_pjc_contacts4(int nc, int nblocks, const int* __restrict__ ida,
                                                          const int* __restrict__ idb, const double* __restrict__ J,
                                                          const double* __restrict__ R, const double* __restrict__ bias,
                                                          int use_bias, const double* __restrict__ mu, double* __restrict__ P,
                                                          double* __restrict__ PK, double* __restrict__ PK1,
                                                          const double* __restrict__ X, const double* __restrict__ XK,
                                                          double* __restrict__ imp9, const int* __restrict__ ctrl,
                                                          const double* __restrict__ cheb, const double* __restrict__ Rb) {
    if (ctrl[0]) return;
    const double w = cheb[0];
    const bool take = cheb[1] != 0.0;
    const int tid = xcd_tile(nblocks) * kBlock + threadIdx.x;
    const int c = tid / kLanesPerContact, k = tid % kLanesPerContact;
    const bool live = c < nc;
    int a = -1, b = -1;
    double muc = 0;
    if (live) {
        a = ida[c];
        b = idb[c];
        muc = mu[c];
    }
    const int nr = muc > 0 ? 3 : 1;
    const bool row = live && (a >= 0 || b >= 0) && k < nr;
    double d[3] = {0, 0, 0}, aA[3] = {0, 0, 0}, aB[3] = {0, 0, 0};
    double pin = 0, pt = 0;
    if (row) {
        const size_t r = (size_t)k * nc + c;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            d[m] = J[(size_t)(3 * k + m) * nc + c];
            aA[m] = J[(size_t)(9 + 3 * k + m) * nc + c];
            aB[m] = J[(size_t)(18 + 3 * k + m) * nc + c];
        }
        const double ps = P[r];
        pin = take ? ps : cheb_extrap(w, ps, PK[r]);
        PK[r] = PK1[r];
        PK1[r] = pin;
        double ua[6], ub[6];
        cheb_input6(X, XK, a, w, take, ua);
        cheb_input6(X, XK, b, w, take, ub);
        const double g = row_dot(d, aA, aB, ua, ub);
        pt = Rb ? g + (use_bias ? bias[r] : 0.0) : pin - (R[r] * g + (use_bias ? bias[r] : 0.0));
    }
    if (Rb) pt = pin - block_apply4(pt, k, c, nc, nr, row, Rb);
    const double pnew = project4(pt, k, muc);
    double dp = 0;
    if (row) {
        dp = pnew - pin;
        P[(size_t)k * nc + c] = pnew;
    }
    write_record4(live && (a >= 0 || b >= 0), k, c, a, b, d, aA, aB, dp, imp9);
}

// Body pass of a Chebyshev sweep: input u = w (X - XK) + XK (or X), XK <- XK1, XK1 <- u,
// X <- u + W^T dp / M; per-block sums of the residual terms and of non-finite entries of X.
__global__ void __launch_bounds__(kBlock) k_pjc_bodies8(int nb, const int* __restrict__ dof, const double* __restrict__ minv,
                                                        const int* __restrict__ incptr, const int* __restrict__ incidx,
                                                        const double* __restrict__ imp9, double* __restrict__ X,
                                                        double* __restrict__ XK, double* __restrict__ XK1, double tol_rel,
                                                        double tol_abs, double* __restrict__ partial,
                                                        const int* __restrict__ ctrl, const double* __restrict__ cheb,
                                                        const double* __restrict__ sinv) {
    if (ctrl[0]) return;
    __shared__ double red[kBlock / 64];
    const double w = cheb[0];
    const bool take = cheb[1] != 0.0;
    const int tid = xcd_tile(gridDim.x) * kBlock + threadIdx.x;
    const int bi = tid / kLanesPerBody, l = tid % kLanesPerBody;
    const bool live = bi < nb;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    bool touched = false;
    gather_body8(bi, l, live, incptr, incidx, imp9, acc, &touched);
    double e2 = 0, bad = 0;
    const int nd = live ? dof[bi] : 0;
    if (live && l < nd) {
        const size_t o = (size_t)bi * kBodyStride + l;
        const double xs = X[o];
        const double u = take ? xs : cheb_extrap(w, xs, XK[o]);
        XK[o] = XK1[o];
        XK1[o] = u;
        const double un = touched ? u + lane_sinv(acc, l, nd, bi, minv, sinv) : u;
        X[o] = un;
        const double q = fmax(fabs(un), fabs(u)) * tol_rel + tol_abs;
        const double e = (un - u) / q;
        e2 = e * e;
        bad = isfinite(un) ? 0.0 : 1.0;
    }
    const double s = block_sum(e2, red);
    const double nf = block_sum(bad, red);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s;
        partial[gridDim.x + blockIdx.x] = nf;
    }
}

// res <= 1 ends the loop; otherwise the next w (or the fallback to the swept iterate when w is
// not a positive finite number or the swept iterate holds a non-finite entry).  it < 0: the sweep
// index is the device sweep counter ctrl[3] (graph-replayed chunks).
__global__ void __launch_bounds__(kBlock) k_pjc_finalize(int nparts, const double* __restrict__ partial, int nv, int it,
                                                         int* __restrict__ ctrl, double* __restrict__ res,
                                                         int* __restrict__ hflag, double* __restrict__ cheb) {
#pragma clang fp contract(off)
    if (ctrl[0]) return;
    __shared__ double red[kBlock / 64];
    double s = strided_sum(partial, nparts), nf = strided_sum(partial + nparts, nparts);
    s = block_sum(s, red);
    nf = block_sum(nf, red);
    if (threadIdx.x != 0) return;
    if (it < 0) it = ctrl[3];
    ctrl[3] = ctrl[3] + 1;
    const double err = 1.0 / sqrt((double)nv) * sqrt(s);
    res[0] = err;
    if (err <= 1.0) {
        ctrl[1] = it + 1;
        ctrl[2] = 0;
        ctrl[0] = 1;
        hflag[0] = 1;
        return;
    }
    const double rho2 = cheb[2];
    const double wn = it == 0 ? 1.0 / (1.0 - rho2 * 0.5) : 1.0 / (1.0 - (rho2 * 0.25) * cheb[0]);
    if (!isfinite(wn) || wn <= 0.0 || nf != 0.0) {
        cheb[0] = 1.0;
        cheb[1] = 1.0;
    } else {
        cheb[0] = wn;
        cheb[1] = 0.0;
    }
}

// After the loop: the sweep input of the converged sweep (XK1, PK1), or the pending update of
// the last sweep after max_iter; written to X / P (the solve's velocity and impulse buffers).
__global__ void k_pjc_finish_bodies(int nb, const int* __restrict__ dof, double* __restrict__ X,
                                    const double* __restrict__ XK, const double* __restrict__ XK1,
                                    const int* __restrict__ ctrl, const double* __restrict__ cheb) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb * 6) return;
    const int bi = t / 6, l = t % 6;
    if (l >= dof[bi]) return;
    const size_t o = (size_t)bi * kBodyStride + l;
    if (ctrl[0]) X[o] = XK1[o];
    else if (cheb[1] == 0.0) X[o] = cheb_extrap(cheb[0], X[o], XK[o]);
}

__global__ void k_pjc_finish_contacts(int nc, const int* __restrict__ ida, const int* __restrict__ idb,
                                      const double* __restrict__ mu, double* __restrict__ P, const double* __restrict__ PK,
                                      const double* __restrict__ PK1, const int* __restrict__ ctrl,
                                      const double* __restrict__ cheb) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * nc) return;
    const int k = t / nc, c = t % nc;
    if ((ida[c] < 0 && idb[c] < 0) || k >= (mu[c] > 0 ? 3 : 1)) return;
    if (ctrl[0]) P[t] = PK1[t];
    else if (cheb[1] == 0.0) P[t] = cheb_extrap(cheb[0], P[t], PK[t]);
}

// ---------------------------------------------------------------------------------------------
// Nesterov and Chebyshev loops with joint rows (nesterov_loop, chebyshev_loop and
// estimateSpectralRadius, projected_jacobi.cpp:155-228,379-449, over x = [v; Lambda]).
//
// The sweep is the standard one (contact pass, body pass, chain solve of the jointed bodies and
// Lambda); these kernels keep the loops' state around it.  The state has three segments: the
// padded twists (nb x 8), Lambda (one per joint row) and the impulses (3 nc).  C is the working
// copy the sweep updates in place (d_x, d_jlam, d_p), K and K1 the loops' older iterates.
//   Nesterov: after sweep i, C = x_{k+1}; the update forms y = C + beta (C - K) (y = C when beta
//   is 0: a restart, momentum off, or the first sweep), K <- C, C <- y, and keeps y's twists
//   for the next residual.
//   Chebyshev: K = x_k, K1 = x_{k+1} (the sweep input), C = sweep(K1); the update forms
//   x_{k+2} = w (C - K) + K (C after a fallback), K <- K1, K1 <- x_{k+2}, C <- x_{k+2}.
// The residual and the restart direction test run over the velocities, the extrapolations over
// everything, as in the reference.
// ---------------------------------------------------------------------------------------------

// Residual partials between C and the sweep input IN over the velocities, and per mode the restart
// dot (IN - C) . (C - K) (0) or the count of non-finite entries of C and of the joint impulses
// (1: Chebyshev's fallback test on the whole of x).
__global__ void __launch_bounds__(kBlock) k_accj_residual(int mode, int nb, const int* __restrict__ dof,
                                                          const double* __restrict__ C, const double* __restrict__ IN,
                                                          const double* __restrict__ K, int nj,
                                                          const double* __restrict__ LC, double tol_rel, double tol_abs,
                                                          double* __restrict__ partial, const int* __restrict__ ctrl) {
    if (ctrl[0]) return;
    __shared__ double red[kBlock / 64];
    const int b = blockIdx.x * kBlock + threadIdx.x;
    double e2 = 0, aux = 0;
    if (b < nb) {
        double c[6], y[6], k6[6];
        load_body(C, b, c);
        load_body(IN, b, y);
        load_body(K, b, k6);
        for (int m = 0; m < dof[b]; ++m) {
            const double q = fmax(fabs(c[m]), fabs(y[m])) * tol_rel + tol_abs;
            const double e = (c[m] - y[m]) / q;
            e2 += e * e;
            if (mode == 0) aux += (y[m] - c[m]) * (c[m] - k6[m]);
            else if (!isfinite(c[m])) aux += 1.0;
        }
    }
    if (mode == 1)
        for (int r = b; r < nj; r += gridDim.x * kBlock)
            if (!isfinite(LC[r])) aux += 1.0;
    const double s = block_sum(e2, red);
    const double d = block_sum(aux, red);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s;
        partial[gridDim.x + blockIdx.x] = d;
    }
}

// The loop update over the three segments (lengths n0, n1, n2).  mode 0 (Nesterov): s = nest
// (s[0] = beta), A = the kept sweep input (twists only, A0).  mode 1 (Chebyshev): s = cheb
// (s[0] = w, s[1] = take the swept iterate), A = K1.
__global__ void k_accj_update(int mode, int n0, int n1, int n2, double* __restrict__ C0, double* __restrict__ C1,
                              double* __restrict__ C2, double* __restrict__ K0, double* __restrict__ K1,
                              double* __restrict__ K2, double* __restrict__ A0, double* __restrict__ A1,
                              double* __restrict__ A2, const double* __restrict__ s, const int* __restrict__ ctrl) {
#pragma clang fp contract(off)
    if (ctrl[0]) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    double *c, *k, *a;
    if (t < n0) {
        c = C0; k = K0; a = A0;
    } else if ((t -= n0) < n1) {
        c = C1; k = K1; a = A1;
    } else if ((t -= n1) < n2) {
        c = C2; k = K2; a = A2;
    } else {
        return;
    }
    const double cv = c[t], kv = k[t];
    if (mode == 0) {
        const double beta = s[0];
        const double y = beta == 0.0 ? cv : cv + beta * (cv - kv);
        k[t] = cv;
        c[t] = y;
        if (a) a[t] = y;
    } else {
        const double n = s[1] != 0.0 ? cv : cheb_extrap(s[0], cv, kv);
        k[t] = a[t];
        a[t] = n;
        c[t] = n;
    }
}

// After the loop: Nesterov returns x_{k+1} (K after an update, so when the loop ran out of
// sweeps), Chebyshev the sweep input K1.
__global__ void k_accj_finish(int mode, int n0, int n1, int n2, double* __restrict__ C0, double* __restrict__ C1,
                              double* __restrict__ C2, const double* __restrict__ S0, const double* __restrict__ S1,
                              const double* __restrict__ S2, const int* __restrict__ ctrl) {
    if (mode == 0 && ctrl[0]) return;  // converged: the swept iterate is in place
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n0) C0[t] = S0[t];
    else if ((t -= n0) < n1) C1[t] = S1[t];
    else if ((t -= n1) < n2) C2[t] = S2[t];
}

// Power iteration with joint rows: the vector over the twists and Lambda.
__global__ void k_accj_fill(int nb, const int* __restrict__ dof, int nj, double inv, double* __restrict__ vx,
                            double* __restrict__ vl) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nb * kBodyStride) {
        const int bi = t / kBodyStride, l = t % kBodyStride;
        vx[t] = l < dof[bi] ? inv : 0.0;
    } else if (t - nb * kBodyStride < nj) {
        vl[t - nb * kBodyStride] = inv;
    }
}

__global__ void __launch_bounds__(kBlock) k_accj_sq(int nb, const int* __restrict__ dof, const double* __restrict__ vx,
                                                    int nj, const double* __restrict__ vl, double* __restrict__ partial,
                                                    const double* __restrict__ cheb) {
    if (cheb[4] != 0.0) return;
    __shared__ double red[kBlock / 64];
    const int b = blockIdx.x * kBlock + threadIdx.x;
    double s2 = 0;
    if (b < nb)
        for (int m = 0; m < dof[b]; ++m) s2 += vx[(size_t)b * kBodyStride + m] * vx[(size_t)b * kBodyStride + m];
    for (int r = b; r < nj; r += gridDim.x * kBlock) s2 += vl[r] * vl[r];
    const double s = block_sum(s2, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void k_accj_scale(int nb, int nj, double* __restrict__ vx, double* __restrict__ vl,
                             const double* __restrict__ cheb) {
    if (cheb[4] != 0.0) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nb * kBodyStride) vx[t] = vx[t] / cheb[6];
    else if (t - nb * kBodyStride < nj) vl[t - nb * kBodyStride] = vl[t - nb * kBodyStride] / cheb[6];
}

// ---------------------------------------------------------------------------------------------
// Anderson mixing (anderson_loop, projected_jacobi.cpp:224-310).
//
// The mixed vector z = [x; p] over the entries the reference has: x (b, k) for k < dof[b] (entry
// i = 6 b + k), then the contact rows (entry 6 nb + k nc + c, row k < 1 or 3 of a contact with a
// dynamic side).  History rings Z and F of S = m + 1 slots each hold (z_i, f_i = g(z_i) - z_i); the
// state st = {cnt, head, mk, gamma ok, non-finite candidate} lives on the device, so a chunk of
// iterations replays as one graph.  Per iteration, after the sweep w = g(z) (contact and body
// passes, residual every iteration):
//   push:      f = w - z into F[zs] (zs = the slot after head, which holds z), the dot products of the
//              difference columns dF_j = F[s_{j+1}] - F[s_j] with each other and with f;
//   solve:     gamma = argmin ||f - dF gamma|| by a pivoted Cholesky of the Gram matrix with the
//              pivot rule of Eigen's ColPivHouseholderQR; advance the ring;
//   candidate: w - (dZ + dF) gamma, flagging a non-finite entry;
//   commit:    z_next = the candidate, or w (and the history dropped) when gamma or the candidate is
//              not finite; z_next also goes into the ring slot after the new head.
// The dot products and the least squares run in a different order from Eigen's Householder QR,
// so the device matches the oracle to rounding, not bitwise.
// ---------------------------------------------------------------------------------------------
constexpr int kAndMaxM = 8;
constexpr int kAndDots = kAndMaxM * (kAndMaxM + 1) / 2 + kAndMaxM;
constexpr int kAndBlocks = 1024;  // grid of the dot-product pass (fixed: deterministic partials)

// address of mixed entry i, or null where the reference has no such row
__device__ inline double* and_entry(int i, int nb, int nc, const int* __restrict__ dof, const int* __restrict__ ida,
                                    const int* __restrict__ idb, const double* __restrict__ mu, double* x, double* p) {
    if (i < 6 * nb) {
        const int b = i / 6, k = i - 6 * (i / 6);
        return k < dof[b] ? x + (size_t)b * kBodyStride + k : nullptr;
    }
    const int t = i - 6 * nb, k = t / nc, c = t - (t / nc) * nc;
    if ((ida[c] < 0 && idb[c] < 0) || k >= (mu[c] > 0 ? 3 : 1)) return nullptr;
    return p + t;
}

// oldest-first slot j (0 .. cnt-1) of a ring whose newest entry is `head`
__device__ inline int and_slot(int j, int cnt, int head, int S) { return (head - (cnt - 1) + j + 2 * S) % S; }

__global__ void __launch_bounds__(kBlock) k_and_init(int n, int nb, int nc, const int* __restrict__ dof,
                                                     const int* __restrict__ ida, const int* __restrict__ idb,
                                                     const double* __restrict__ mu, double* x, double* p,
                                                     double* __restrict__ Z0) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double* e = and_entry(i, nb, nc, dof, ida, idb, mu, x, p);
        Z0[i] = e ? *e : 0.0;
    }
}

template <int MK>
__device__ inline void and_push_body(int n, int nb, int nc, const int* __restrict__ dof, const int* __restrict__ ida,
                                     const int* __restrict__ idb, const double* __restrict__ mu, double* x, double* p,
                                     double* __restrict__ Z, double* __restrict__ F, int S, int zs,
                                     double* __restrict__ part) {
    constexpr int ND = MK * (MK + 1) / 2 + MK;
    double acc[ND > 0 ? ND : 1];
#pragma unroll
    for (int d = 0; d < ND; ++d) acc[d] = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double* e = and_entry(i, nb, nc, dof, ida, idb, mu, x, p);
        if (!e) continue;
        const double f = *e - Z[(size_t)zs * n + i];
        F[(size_t)zs * n + i] = f;
        if constexpr (MK > 0) {
            double Fo[MK + 1];  // F of the history, oldest first; the newest is f
#pragma unroll
            for (int j = 0; j < MK; ++j) Fo[j] = F[(size_t)((zs - MK + j + S) % S) * n + i];
            Fo[MK] = f;
            double dF[MK];
#pragma unroll
            for (int j = 0; j < MK; ++j) dF[j] = Fo[j + 1] - Fo[j];
            int d = 0;
#pragma unroll
            for (int j = 0; j < MK; ++j) {
#pragma unroll
                for (int l = j; l < MK; ++l) acc[d++] += dF[j] * dF[l];
            }
#pragma unroll
            for (int j = 0; j < MK; ++j) acc[d++] += dF[j] * f;
        }
    }
    if constexpr (ND > 0) {
        __shared__ double red[kBlock / 64][ND];
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            double v = acc[d];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            if (lane == 0) red[wv][d] = v;
        }
        __syncthreads();
        if (threadIdx.x < ND) {
            double v = 0;
            for (int w = 0; w < kBlock / 64; ++w) v += red[w][threadIdx.x];
            part[(size_t)blockIdx.x * kAndDots + threadIdx.x] = v;
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_and_push(int n, int nb, int nc, const int* __restrict__ dof,
                                                     const int* __restrict__ ida, const int* __restrict__ idb,
                                                     const double* __restrict__ mu, double* x, double* p,
                                                     double* __restrict__ ring, int m, const int* __restrict__ st,
                                                     double* __restrict__ part, const int* __restrict__ ctrl) {
    if (ctrl[0]) return;
    const int S = m + 1;
    const int zs = (st[1] + 1) % S;
    const int mk = min(st[0] + 1, S) - 1;
    double* Z = ring;
    double* F = ring + (size_t)S * n;
    switch (mk) {
        case 0: and_push_body<0>(n, nb, nc, dof, ida, idb, mu, x, p, Z, F, S, zs, part); break;
        case 1: and_push_body<1>(n, nb, nc, dof, ida, idb, mu, x, p, Z, F, S, zs, part); break;
        case 2: and_push_body<2>(n, nb, nc, dof, ida, idb, mu, x, p, Z, F, S, zs, part); break;
        case 3: and_push_body<3>(n, nb, nc, dof, ida, idb, mu, x, p, Z, F, S, zs, part); break;
        case 4: and_push_body<4>(n, nb, nc, dof, ida, idb, mu, x, p, Z, F, S, zs, part); break;
        case 5: and_push_body<5>(n, nb, nc, dof, ida, idb, mu, x, p, Z, F, S, zs, part); break;
        case 6: and_push_body<6>(n, nb, nc, dof, ida, idb, mu, x, p, Z, F, S, zs, part); break;
        case 7: and_push_body<7>(n, nb, nc, dof, ida, idb, mu, x, p, Z, F, S, zs, part); break;
        default: and_push_body<8>(n, nb, nc, dof, ida, idb, mu, x, p, Z, F, S, zs, part); break;
    }
}

// One workgroup: the Gram matrix G_jl = dF_j . dF_l and r_j = dF_j . f from the partials, then
// gamma.  Pivot k takes the remaining column of largest norm (Schur complement diagonal of G);
// as in ColPivHouseholderQR the solve uses the pivots before the first whose remaining norm^2 is
// below (max column norm * eps)^2 (rows - k) / rows, the other coefficients are 0.  A pivot that
// passes that test with a non-positive remaining norm^2 (possible only through the rounding of
// the Gram matrix; the Householder solve would divide by a zero or rounding-level diagonal)
// makes gamma non-finite, i.e. the plain step.  Advances the ring: head = zs, cnt += 1 (<= S).
__global__ void __launch_bounds__(kBlock) k_and_solve(int nparts, const double* __restrict__ part, int m, double rows,
                                                      int* __restrict__ st, double* __restrict__ gam,
                                                      const int* __restrict__ ctrl) {
    if (ctrl[0]) return;
    __shared__ double red[kBlock / 64];
    __shared__ double dots[kAndDots];
    const int S = m + 1;
    const int cnt = min(st[0] + 1, S), mk = cnt - 1, zs = (st[1] + 1) % S;
    const int nd = mk * (mk + 1) / 2 + mk;
    for (int d = 0; d < nd; ++d) {
        double v = 0;
        for (int b = threadIdx.x; b < nparts; b += kBlock) v += part[(size_t)b * kAndDots + d];
        v = block_sum(v, red);
        if (threadIdx.x == 0) dots[d] = v;
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    double A[kAndMaxM][kAndMaxM], c[kAndMaxM], g[kAndMaxM];
    int perm[kAndMaxM];
    int d = 0;
    for (int j = 0; j < mk; ++j)
        for (int l = j; l < mk; ++l) A[j][l] = A[l][j] = dots[d++];
    for (int j = 0; j < mk; ++j) c[j] = dots[d++];
    for (int j = 0; j < mk; ++j) perm[j] = j;
    double maxn2 = 0;
    for (int j = 0; j < mk; ++j) maxn2 = fmax(maxn2, A[j][j]);
    const double eps = 2.220446049250313e-16;
    const double thr = (sqrt(maxn2) * eps) * (sqrt(maxn2) * eps) / rows;
    int nz = mk, ok = 1;
    for (int k = 0; k < mk; ++k) {
        int big = k;
        for (int j = k + 1; j < mk; ++j)
            if (A[j][j] > A[big][big]) big = j;
        if (A[big][big] < thr * (rows - k)) {
            nz = k;
            break;
        }
        if (big != k) {
            for (int j = 0; j < mk; ++j) {
                const double t = A[k][j]; A[k][j] = A[big][j]; A[big][j] = t;
            }
            for (int j = 0; j < mk; ++j) {
                const double t = A[j][k]; A[j][k] = A[j][big]; A[j][big] = t;
            }
            const double t = c[k]; c[k] = c[big]; c[big] = t;
            const int ti = perm[k]; perm[k] = perm[big]; perm[big] = ti;
        }
        if (!(A[k][k] > 0.0)) {
            ok = 0;
            break;
        }
        const double rkk = sqrt(A[k][k]);
        A[k][k] = rkk;
        for (int j = k + 1; j < mk; ++j) A[k][j] /= rkk;  // row k of R
        for (int j = k + 1; j < mk; ++j)
            for (int l = k + 1; l < mk; ++l) A[j][l] -= A[k][j] * A[k][l];
    }
    for (int j = 0; j < mk; ++j) g[j] = 0.0;
    if (ok && nz > 0) {
        double y[kAndMaxM];
        for (int k = 0; k < nz; ++k) {  // R^T y = c
            double t = c[k];
            for (int i = 0; i < k; ++i) t -= A[i][k] * y[i];
            y[k] = t / A[k][k];
        }
        for (int k = nz - 1; k >= 0; --k) {  // R gamma' = y
            double t = y[k];
            for (int j = k + 1; j < nz; ++j) t -= A[k][j] * g[j];
            g[k] = t / A[k][k];
        }
        for (int k = 0; k < nz; ++k) ok = ok && isfinite(g[k]);
    }
    double out[kAndMaxM];
    for (int j = 0; j < mk; ++j) out[j] = 0.0;
    for (int k = 0; k < nz; ++k) out[perm[k]] = g[k];
    for (int j = 0; j < mk; ++j) gam[j] = out[j];
    st[0] = cnt;
    st[1] = zs;
    st[2] = mk;
    st[3] = ok;
    st[4] = 0;
}

__global__ void __launch_bounds__(kBlock) k_and_candidate(int n, int nb, int nc, const int* __restrict__ dof,
                                                          const int* __restrict__ ida, const int* __restrict__ idb,
                                                          const double* __restrict__ mu, double* x, double* p,
                                                          const double* __restrict__ ring, int m, int* __restrict__ st,
                                                          const double* __restrict__ gam, double* __restrict__ C,
                                                          const int* __restrict__ ctrl) {
    if (ctrl[0]) return;
    const int mk = st[2];
    if (mk < 1 || !st[3]) return;
    const int S = m + 1, cnt = st[0], head = st[1];
    const double* Z = ring;
    const double* F = ring + (size_t)S * n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double* e = and_entry(i, nb, nc, dof, ida, idb, mu, x, p);
        if (!e) continue;
        double t = 0.0;
        for (int j = 0; j < mk; ++j) {
            const size_t s0 = (size_t)and_slot(j, cnt, head, S) * n + i, s1 = (size_t)and_slot(j + 1, cnt, head, S) * n + i;
            const double a = (Z[s1] - Z[s0]) + (F[s1] - F[s0]);
            t += a * gam[j];
        }
        const double v = *e - t;
        C[i] = v;
        if (!isfinite(v)) atomicOr(st + 4, 1);
    }
}

__global__ void __launch_bounds__(kBlock) k_and_commit(int n, int nb, int nc, const int* __restrict__ dof,
                                                       const int* __restrict__ ida, const int* __restrict__ idb,
                                                       const double* __restrict__ mu, double* x, double* p,
                                                       double* __restrict__ ring, int m, int* __restrict__ st,
                                                       const double* __restrict__ C, const int* __restrict__ ctrl) {
    if (ctrl[0]) return;
    const int S = m + 1, mk = st[2];
    const bool ok = mk >= 1 && st[3] && !st[4];
    const int zn = (st[1] + 1) % S;
    double* Z = ring;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double* e = and_entry(i, nb, nc, dof, ida, idb, mu, x, p);
        if (!e) {
            Z[(size_t)zn * n + i] = 0.0;
            continue;
        }
        double v = *e;
        if (ok) {
            v = C[i];
            *e = v;
        }
        Z[(size_t)zn * n + i] = v;
    }
    if (mk >= 1 && !ok && blockIdx.x == 0 && threadIdx.x == 0) st[0] = 0;  // drop the history
}

}  // namespace cdx

using namespace cdx;

static int run_plain_sweeps(cdx_ctx* ctx, const cdx_solver_params* prm, int gb, int gc, int* iters_out);
static size_t range_lds_bytes(int maxs, int maxp, int maxl);
static int build_ranges(cdx_ctx* ctx);

// The inverse rotational blocks of S of the current solve with implicit gyroscopy, else nullptr.
static const double* gyro_sinv(const cdx_ctx* ctx, const cdx_solver_params* prm) {
    return prm->implicit_gyro ? ctx->d_sinv : nullptr;
}
static int run_chebyshev(cdx_ctx* ctx, const cdx_solver_params* prm, int* iters_out);
static int run_anderson(cdx_ctx* ctx, const cdx_solver_params* prm, int* iters_out);
static int run_accel_joints(cdx_ctx* ctx, const cdx_solver_params* prm, int mode, int* iters_out);

extern "C" int cdx_pj_solve(cdx_ctx* ctx, double dt, double theta, const cdx_solver_params* prm, cdx_solve_info* info) {
    // theta only enters the joint rows of S (PjAssembler::buildAndFactorS / rhs)
    if (!ctx) return CDX_BAD_ARG;
    hipSetDevice(ctx->device);
    CDX_TRY(check_solve_args(ctx, prm));
    if (prm->accel != CDX_ACCEL_NONE && prm->accel != CDX_ACCEL_NESTEROV && prm->accel != CDX_ACCEL_CHEBYSHEV &&
        prm->accel != CDX_ACCEL_ANDERSON)
        return set_err(ctx, CDX_BAD_ARG, "cdx_pj_solve: unknown acceleration mode");
    if (prm->accel == CDX_ACCEL_ANDERSON && (prm->anderson_m < 1 || prm->anderson_m > kAndMaxM))
        return set_err(ctx, CDX_BAD_ARG, "cdx_pj_solve: anderson_m must be in 1..8");
    const bool joints = ctx->njrows > 0;
    const bool block = prm->block_precond != 0;
    if (joints && prm->accel == CDX_ACCEL_ANDERSON)
        return set_err(ctx, CDX_BAD_ARG, "cdx_pj_solve: the Anderson loop does not take joint rows");
    if (joints && block)
        return set_err(ctx, CDX_BAD_ARG, "cdx_pj_solve: the block preconditioner does not take joint rows");
    if (joints && !(theta > 0)) return set_err(ctx, CDX_BAD_ARG, "cdx_pj_solve: theta must be positive with joint rows");
    if (joints && prm->implicit_gyro)
        return set_err(ctx, CDX_BAD_ARG, "cdx_pj_solve: implicit gyroscopy does not take joint rows");
    const bool gyro = prm->implicit_gyro != 0;
    ctx->joints_solved = false;
    ctx->gyro_last = false;
    if (info) {
        info->iters = 0;
        info->residual = std::nan("");
        info->nrows = ctx->nrows;
        info->ncolors = 0;
    }
    ctx->solved = false;
    ctx->ranges_used = false;
    ctx->geo_used = false;
    ctx->pj_last = false;
    ctx->pjj_last = false;
    ctx->last_theta = theta;
    ctx->record_valid = false;
    ctx->recorded = 0;
    const int nb = ctx->nb, nc = ctx->nc;
    if (nb == 0) {
        ctx->solved = true;
        return CDX_OK;
    }
    CDX_TRY(build_incidence(ctx));
    CDX_TRY(reset_control(ctx));
    const bool have_contacts = ctx->nrows > 0;
    if (have_contacts && block) CDX_TRY(ensure_alloc(ctx, &ctx->d_Rb, 9 * (size_t)nc, &ctx->Rb_cap, "inverse Delassus blocks"));
    if (have_contacts)
        CDX_LAUNCH(ctx, "pj_setup_contacts", k_pj_setup_contacts, grid_for(nc), kBlock, nc, ctx->d_ida, ctx->d_idb, ctx->d_J,
                   ctx->d_mu, ctx->d_rest, ctx->d_vb, ctx->d_minv, ctx->d_v, prm->alpha, prm->warmstart, ctx->d_R,
                   ctx->d_bias, ctx->d_p, ctx->d_imp, block ? ctx->d_Rb : nullptr);
    if (gyro) {  // S's rotational blocks; a singular one ends the solve before the loop
        CDX_TRY(ensure_alloc(ctx, &ctx->d_sinv, 9 * (size_t)nb, &ctx->sinv_cap, "gyroscopic blocks"));
        CDX_LAUNCH(ctx, "pj_gyro_blocks", k_pj_gyro_blocks, grid_for(nb), kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_v, dt,
                   ctx->d_sinv, ctx->d_ctrl);
        int ctrl[3];
        CDX_HIP(ctx, hipMemcpyAsync(ctrl, ctx->d_ctrl, sizeof(ctrl), hipMemcpyDeviceToHost, ctx->stream));
        CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (ctrl[0]) return set_err(ctx, CDX_NOT_FACTORIZED, "cdx_pj_solve: S is not factorized (singular gyroscopic block)");
    }
    CDX_LAUNCH(ctx, "pj_setup_bodies", k_pj_setup_bodies, grid_for(nb), kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_v,
               ctx->d_fext, ctx->d_fgyro, ctx->d_incptr, ctx->d_incidx, ctx->d_imp, dt, have_contacts ? 1 : 0, ctx->d_x,
               joints ? ctx->d_jslot : nullptr, joints ? ctx->d_jtr : nullptr, gyro_sinv(ctx, prm));
    CDX_HIP(ctx, hipGetLastError());
    if (joints) {  // x0 = S^-1 (rhs + [W^T p0; 0]) for the jointed bodies, and Lambda_0
        CDX_TRY(joints_factor(ctx, dt, theta));
        CDX_TRY(joints_solve(ctx, 1, 0, 0.0, 0.0, nullptr));
    }
    if (!have_contacts) {  // no contacts: free solution only (projected_jacobi.cpp:460-467)
        CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (joints) {
            int ctrl[3];
            CDX_HIP(ctx, hipMemcpy(ctrl, ctx->d_ctrl, sizeof(ctrl), hipMemcpyDeviceToHost));
            if (ctrl[0] && ctrl[2]) return set_err(ctx, ctrl[2], "cdx_pj_solve: S is not factorized (singular joint block)");
        }
        ctx->solved = true;
        ctx->pj_last = !joints && !block;
        ctx->pjj_last = joints;
        ctx->joints_solved = joints;
        ctx->gyro_last = gyro;
        ctx->record_valid = ctx->record && !joints && !block && !gyro;
        return CDX_OK;
    }
    const int gb = grid_for(nb), gc = grid_for(nc);
    CDX_TRY(ensure_alloc(ctx, &ctx->d_partial, (size_t)gb, &ctx->partial_cap, "residual partials"));
    int iters = 0;
    int st = CDX_OK;
    if (joints && (prm->accel == CDX_ACCEL_NESTEROV || prm->accel == CDX_ACCEL_CHEBYSHEV)) {
        st = run_accel_joints(ctx, prm, prm->accel == CDX_ACCEL_NESTEROV ? 0 : 1, &iters);
    } else if (prm->accel == CDX_ACCEL_NESTEROV) {
        // xk = x0, pk = p0, beta = 0, thk = 1, err_prev = inf, restarts = 0, momentum on
        // the range sweep (one lane per contact, per-body partials) on large problems, row scaling only
        bool ranges = false;
        if (!block && (ctx->range_mode == 2 || (ctx->range_mode == 1 && nc >= kRangeAutoContacts))) {
            if (!ctx->ranges_valid) CDX_TRY(build_ranges(ctx));
            ranges = ctx->ranges_ok;
        }
        ctx->ranges_used = ranges;
        const bool geo = ranges && ctx->geo_ok && ctx->geo_mode;
        ctx->geo_used = geo;
        const int gb8 = grid_for(nb * kLanesPerBody), ng = ctx->nranges, per = ctx->range_per_used;
        const int nthr = ranges ? std::max(64, std::min({ctx->range_block, 256, pad64(std::min(per, nc))})) : 0;
        const size_t shm = ranges ? range_lds_bytes(std::min(per, nc), ctx->range_maxp, ctx->range_maxl) : 0;
        const int nparts = ranges ? gb8 : gb;
        CDX_TRY(ensure_alloc(ctx, &ctx->d_pk, 3 * (size_t)nc, &ctx->pk_cap, "nesterov pk"));
        CDX_TRY(ensure_alloc(ctx, &ctx->d_nest, 8, nullptr, "nesterov scalars"));
        CDX_TRY(ensure_alloc(ctx, &ctx->d_partial2, 2 * (size_t)nparts, &ctx->partial2_cap, "nesterov partials"));
        CDX_HIP(ctx, hipMemcpyAsync(ctx->d_uprev, ctx->d_x, sizeof(double) * (size_t)nb * kBodyStride, hipMemcpyDeviceToDevice,
                                    ctx->stream));
        CDX_HIP(ctx, hipMemcpyAsync(ctx->d_pk, ctx->d_p, sizeof(double) * 3 * (size_t)nc, hipMemcpyDeviceToDevice, ctx->stream));
        const double n0[8] = {0.0, 1.0, INFINITY, 0.0, 0.0, 0.0, 0.0, 0.0};
        CDX_HIP(ctx, hipMemcpyAsync(ctx->d_nest, n0, sizeof(n0), hipMemcpyHostToDevice, ctx->stream));
        // one sweep; it < 0 inside a captured chunk (the finalize kernel reads the sweep counter)
        auto nsweep = [&](int it, int) -> int {
            if (ranges) {
#define CDX_NEST_RANGE(GEO)                                                                                             \
    CDX_LAUNCH_SHM(ctx, "pjn_range", (k_pj_range<false, GEO, kLoopNest>), ng, nthr, shm, ng, nc, per, std::min(per, nc), \
                   ctx->range_maxp, ctx->d_rg_p, ctx->d_rg_l, ctx->d_pr_ptr, ctx->d_pr_out, ctx->d_pr_inc, ctx->d_ida,  \
                   ctx->d_idb, ctx->d_J, ctx->d_R, ctx->d_bias, ctx->any_bias ? 1 : 0, ctx->d_mu, ctx->d_p, ctx->d_x,   \
                   ctx->d_part, ctx->d_ctrl, ctx->d_geo, ctx->d_q, ctx->d_pk, nullptr, ctx->d_uprev, ctx->d_nest)
                if (geo)
                    CDX_NEST_RANGE(true);
                else
                    CDX_NEST_RANGE(false);
#undef CDX_NEST_RANGE
                CDX_LAUNCH(ctx, "pjn_bodies_parts", k_pjn_bodies_parts, gb8, kBlock, nb, ctx->d_dof, ctx->d_minv,
                           ctx->d_bp_ptr, ctx->d_bp_ovf, ctx->d_part, ctx->d_x, ctx->d_uprev, prm->tol_rel, prm->tol_abs,
                           ctx->d_partial2, ctx->d_ctrl, ctx->d_nest, gyro_sinv(ctx, prm));
            } else {
                CDX_LAUNCH(ctx, "pj_contacts", k_pjn_contacts, gc, kBlock, nc, gc, ctx->d_ida, ctx->d_idb, ctx->d_J, ctx->d_R,
                           ctx->d_bias, ctx->d_mu, ctx->d_p, ctx->d_pk, ctx->d_x, ctx->d_uprev, ctx->d_imp, ctx->d_ctrl,
                           ctx->d_nest, block ? ctx->d_Rb : nullptr);
                CDX_LAUNCH(ctx, "pj_bodies", k_pjn_bodies, gb, kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_incptr,
                           ctx->d_incidx, ctx->d_imp, ctx->d_x, ctx->d_uprev, prm->tol_rel, prm->tol_abs, ctx->d_partial2,
                           ctx->d_ctrl, ctx->d_nest, gyro_sinv(ctx, prm));
            }
            CDX_LAUNCH(ctx, "pj_residual", k_pjn_finalize, 1, kBlock, nparts, ctx->d_partial2, ctx->nv, it,
                       prm->nesterov_beta_threshold, prm->nesterov_restart_limit, ctx->d_ctrl, ctx->d_res, ctx->h_flag_dev,
                       ctx->d_nest);
            return hip_check(ctx, hipGetLastError(), "pj nesterov sweep launch");
        };
        st = run_sweeps_graphed(ctx, prm->max_iter, 8, true, nsweep, &iters);
    } else if (prm->accel == CDX_ACCEL_CHEBYSHEV) {
        st = run_chebyshev(ctx, prm, &iters);
    } else if (prm->accel == CDX_ACCEL_ANDERSON) {
        st = run_anderson(ctx, prm, &iters);
    } else {
        st = run_plain_sweeps(ctx, prm, gb, gc, &iters);
    }
    if (st) return set_err(ctx, st, st == CDX_NOT_FACTORIZED ? "cdx_pj_solve: S is not factorized (singular joint block)"
                                                             : "cdx_pj_solve failed");
    double r = 0;
    CDX_TRY(read_residual(ctx, &r));
    if (info) {
        info->iters = iters;
        info->residual = r;
    }
    ctx->solved = true;
    ctx->joints_solved = joints;
    ctx->pj_last = !joints && !block;  // the adjoint differentiates the row-scaled, diagonal-S solve only
    ctx->pjj_last = joints && prm->accel == CDX_ACCEL_NONE;  // joint rows refuse block mode and acceleration
    ctx->recorded = iters;
    ctx->gyro_last = gyro;
    ctx->record_valid = ctx->record && prm->accel == CDX_ACCEL_NONE && !block && !gyro && iters <= ctx->record_max;
    return CDX_OK;
}

// LDS of one range workgroup: 9 doubles per contact of the range, then the int tables
// (k_pj_range).
static size_t range_lds_bytes(int maxs, int maxp, int maxl) {
    return sizeof(double) * 9 * (size_t)maxs + sizeof(int) * (size_t)(pad64(maxp + 1) + pad64(maxp) + pad64(maxl));
}

template <typename T>
static int upload_exact(cdx_ctx* ctx, T** dst, const std::vector<T>& src, const char* what) {
    if (*dst) hipFree(*dst);
    *dst = nullptr;
    CDX_TRY(ensure_alloc(ctx, dst, src.size(), nullptr, what));
    if (!src.empty())
        CDX_HIP(ctx, hipMemcpyAsync(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice, ctx->stream));
    return CDX_OK;
}

// Tables of the range sweep (k_pj_range, k_pj_bodies_parts) for the current contacts: range g
// holds contacts [g per, (g + 1) per).  A (body, range) pair exists for every range holding an
// incidence of the body; a body's partials are numbered in ascending range order, the first
// kPartSlots in the body's fixed slots and the rest in the overflow area.
// Per range, its pairs in body order (rg_p), each with its partial slot (pr_out) and its
// incidences in contact order (pr_ptr, pr_inc: contact in the range << 1 | side).
static int build_ranges(cdx_ctx* ctx) {
    ctx->ranges_valid = true;
    ctx->ranges_ok = false;
    const int nb = ctx->nb, nc = ctx->nc;
    if (nb <= 0 || nc <= 0) return CDX_OK;
    const int per = std::max(16, std::min(ctx->range_per > 0 ? ctx->range_per : 128, kRangeMaxContacts));
    const int ng = (nc + per - 1) / per;
    const std::vector<int>& ip = ctx->h_incptr;
    const std::vector<int>& ix = ctx->h_incidx;
    // pairs and incidences per range; partial slots per body
    std::vector<int> npr(ng + 1, 0), nin(ng + 1, 0), bp(nb + 1, 0);
    for (int b = 0; b < nb; ++b) {
        int last = -1;
        for (int i = ip[b]; i < ip[b + 1]; ++i) {
            const int g = (ix[i] >> 1) / per;  // incidences run in contact order: ranges ascend
            if (g != last) {
                npr[g + 1]++;
                bp[b + 1]++;
                last = g;
            }
            nin[g + 1]++;
        }
    }
    for (int g = 0; g < ng; ++g) {
        npr[g + 1] += npr[g];
        nin[g + 1] += nin[g];
    }
    // partial slots: the first kPartSlots of body b at b kPartSlots + j, the rest in the overflow
    // area after nb kPartSlots slots
    std::vector<int> pcnt(nb), povf(nb, 0);
    int novf = 0;
    for (int b = 0; b < nb; ++b) {
        pcnt[b] = bp[b + 1];
        povf[b] = novf;
        novf += std::max(0, pcnt[b] - kPartSlots);
    }
    const int npairs = npr[ng], ninc = nin[ng];
    int maxp = 0, maxl = 0;
    for (int g = 0; g < ng; ++g) {
        maxp = std::max(maxp, npr[g + 1] - npr[g]);
        maxl = std::max(maxl, nin[g + 1] - nin[g]);
    }
    const size_t shm = range_lds_bytes(std::min(per, nc), maxp, maxl);
    if (shm > (size_t)kRangeLdsBytes) return CDX_OK;  // tables beyond the LDS budget: separate passes
    std::vector<int> pr_ptr(npairs + 1), pr_out(npairs), pr_inc(std::max(ninc, 1));
    std::vector<int> fp(npr.begin(), npr.end() - 1), fl(nin.begin(), nin.end() - 1);
    for (int b = 0; b < nb; ++b) {
        int last = -1, k = 0;
        for (int i = ip[b]; i < ip[b + 1]; ++i) {
            const int e = ix[i], c = e >> 1, g = c / per;
            if (g != last) {
                const int q = fp[g]++;
                pr_out[q] = k < kPartSlots ? b * kPartSlots + k : nb * kPartSlots + povf[b] + k - kPartSlots;
                ++k;
                pr_ptr[q] = fl[g];
                last = g;
            }
            pr_inc[fl[g]++] = ((c - g * per) << 1) | (e & 1);
        }
    }
    pr_ptr[npairs] = ninc;
    CDX_TRY(upload_exact(ctx, &ctx->d_rg_p, npr, "range pairs"));
    CDX_TRY(upload_exact(ctx, &ctx->d_rg_l, nin, "range incidences"));
    CDX_TRY(upload_exact(ctx, &ctx->d_pr_ptr, pr_ptr, "pair incidence offsets"));
    CDX_TRY(upload_exact(ctx, &ctx->d_pr_out, npairs ? pr_out : std::vector<int>{0}, "pair partial slots"));
    CDX_TRY(upload_exact(ctx, &ctx->d_pr_inc, pr_inc, "pair incidences"));
    CDX_TRY(upload_exact(ctx, &ctx->d_bp_ptr, pcnt, "body partial counts"));
    CDX_TRY(upload_exact(ctx, &ctx->d_bp_ovf, povf, "body partial overflow offsets"));
    if (ctx->d_part) hipFree(ctx->d_part);
    ctx->d_part = nullptr;
    CDX_TRY(ensure_alloc(ctx, &ctx->d_part, 6 * ((size_t)nb * kPartSlots + novf), nullptr, "partial sums"));
    CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->nranges = ng;
    ctx->npart = npairs;
    ctx->range_per_used = per;
    ctx->range_maxp = maxp;
    ctx->range_maxl = maxl;
    ctx->ranges_ok = true;
    return CDX_OK;
}

// Standard loop with the range sweep: per sweep one k_pj_range launch and one k_pj_bodies_parts
// launch, the residual every check_every sweeps; chunks between two checks replayed as graphs.
static int run_range_sweeps(cdx_ctx* ctx, const cdx_solver_params* prm, int* iters_out) {
    const int nb = ctx->nb, nc = ctx->nc, ng = ctx->nranges, per = ctx->range_per_used;
    const int check_every = prm->check_every;
    const bool block = prm->block_precond != 0;
    const int gb8 = grid_for(nb * kLanesPerBody);
    // one lane per contact of a range (and per pair afterwards): up to 256 threads, looping over
    // longer ranges
    const int nthr = std::max(64, std::min({ctx->range_block, 256, pad64(std::min(per, nc))}));
    const int maxs = std::min(per, nc);
    const size_t shm = range_lds_bytes(maxs, ctx->range_maxp, ctx->range_maxl);
    CDX_TRY(ensure_alloc(ctx, &ctx->d_partial, (size_t)gb8, &ctx->partial_cap, "residual partials"));
    ctx->ranges_used = true;
    const bool geo = ctx->geo_ok && ctx->geo_mode;
    ctx->geo_used = geo;
    auto launch_sweep = [&](int it, int check) -> int {
#define CDX_RANGE_LAUNCH(BLK, GEO)                                                                                    \
    CDX_LAUNCH_SHM(ctx, "pj_range", (k_pj_range<BLK, GEO, kLoopStd>), ng, nthr, shm, ng, nc, per, maxs, ctx->range_maxp, \
                   ctx->d_rg_p, ctx->d_rg_l, ctx->d_pr_ptr, ctx->d_pr_out, ctx->d_pr_inc, ctx->d_ida, ctx->d_idb,     \
                   ctx->d_J, BLK ? ctx->d_Rb : ctx->d_R, ctx->d_bias, ctx->any_bias ? 1 : 0, ctx->d_mu, ctx->d_p,     \
                   ctx->d_x, ctx->d_part, ctx->d_ctrl, ctx->d_geo, ctx->d_q, nullptr, nullptr, nullptr, nullptr)
        if (block && geo)
            CDX_RANGE_LAUNCH(true, true);
        else if (block)
            CDX_RANGE_LAUNCH(true, false);
        else if (geo)
            CDX_RANGE_LAUNCH(false, true);
        else
            CDX_RANGE_LAUNCH(false, false);
#undef CDX_RANGE_LAUNCH
        CDX_LAUNCH(ctx, "pj_bodies_parts", k_pj_bodies_parts, gb8, kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_bp_ptr,
                   ctx->d_bp_ovf, ctx->d_part, ctx->d_x, check, prm->tol_rel, prm->tol_abs, ctx->d_partial,
                   ctx->d_ctrl, gyro_sinv(ctx, prm));
        if (check) CDX_TRY(launch_residual_finalize(ctx, gb8, it, 0, check_every));
        return hip_check(ctx, hipGetLastError(), "pj range sweep launch");
    };
    return run_sweeps_graphed(ctx, prm->max_iter, check_every, true,
                              [&](int it, int pos) { return launch_sweep(it, pos == 0); }, iters_out);
}

static int run_plain_sweeps(cdx_ctx* ctx, const cdx_solver_params* prm, int gb, int gc, int* iters_out) {
    const int nb = ctx->nb, nc = ctx->nc;
    // range sweeps: standard loop without joint rows or recording, on large enough problems
    const bool want_ranges = ctx->range_mode == 2 || (ctx->range_mode == 1 && nc >= kRangeAutoContacts);
    if (want_ranges && ctx->njrows == 0 && !ctx->record) {
        if (!ctx->ranges_valid) CDX_TRY(build_ranges(ctx));
        if (ctx->ranges_ok) return run_range_sweeps(ctx, prm, iters_out);
    }
    ctx->ranges_used = false;
    const int check_every = prm->check_every;
    const bool block = prm->block_precond != 0;
    const int gb8 = grid_for(nb * kLanesPerBody);
    const int gc4 = grid_for(nc * kLanesPerContact);
    (void)gc;
    const int gj = ctx->njrows > 0 ? ctx->nchains : 0;  // one partial per chain follows the bodies'
    CDX_TRY(ensure_alloc(ctx, &ctx->d_partial, (size_t)(gb8 + gj), &ctx->partial_cap, "residual partials"));
    (void)gb;
    if (ctx->record) {
        CDX_TRY(ensure_alloc(ctx, &ctx->d_rec_x, (size_t)ctx->record_max * kBodyStride * nb, &ctx->rec_x_cap, "recorded x"));
        CDX_TRY(ensure_alloc(ctx, &ctx->d_rec_p, (size_t)ctx->record_max * 3 * nc, &ctx->rec_p_cap, "recorded p"));
    }
    // one sweep; it < 0 inside a captured chunk (the residual kernel then reads the chunk counter)
    auto launch_sweep = [&](int it, int check) -> int {
        if (block)
            CDX_LAUNCH(ctx, "pj_contacts", k_pj_contacts4<true>, gc4, kBlock, nc, gc4, ctx->d_ida, ctx->d_idb, ctx->d_J,
                       ctx->d_Rb, ctx->d_bias, ctx->any_bias ? 1 : 0, ctx->d_mu, ctx->d_p, ctx->d_x, ctx->d_imp, ctx->d_ctrl);
        else
            CDX_LAUNCH(ctx, "pj_contacts", k_pj_contacts4<false>, gc4, kBlock, nc, gc4, ctx->d_ida, ctx->d_idb, ctx->d_J,
                       ctx->d_R, ctx->d_bias, ctx->any_bias ? 1 : 0, ctx->d_mu, ctx->d_p, ctx->d_x, ctx->d_imp, ctx->d_ctrl);
        CDX_LAUNCH(ctx, "pj_bodies", k_pj_bodies8, gb8, kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_incptr, ctx->d_incidx,
                   ctx->d_imp, ctx->d_x, check, prm->tol_rel, prm->tol_abs, ctx->d_partial, ctx->d_ctrl,
                   gj ? ctx->d_jslot : nullptr, gj ? ctx->d_jtr : nullptr, gyro_sinv(ctx, prm));
        if (gj) CDX_TRY(joints_solve(ctx, 0, check, prm->tol_rel, prm->tol_abs, ctx->d_partial + gb8));
        if (check) CDX_TRY(launch_residual_finalize(ctx, gb8 + gj, it, 0, check_every));
        return hip_check(ctx, hipGetLastError(), "pj sweep launch");
    };
    auto sweep = [&](int it) -> int {
        if (ctx->record && it < ctx->record_max) {  // (x_k, p_k) before sweep k, for the unrolled backward
            CDX_HIP(ctx, hipMemcpyAsync(ctx->d_rec_x + (size_t)it * kBodyStride * nb, ctx->d_x,
                                        sizeof(double) * kBodyStride * (size_t)nb, hipMemcpyDeviceToDevice, ctx->stream));
            CDX_HIP(ctx, hipMemcpyAsync(ctx->d_rec_p + (size_t)it * 3 * nc, ctx->d_p, sizeof(double) * 3 * (size_t)nc,
                                        hipMemcpyDeviceToDevice, ctx->stream));
        }
        return launch_sweep(it, (it % check_every) == 0);
    };
    if (ctx->record) return run_sweeps(ctx, prm->max_iter, check_every, 4, sweep, iters_out);
    // one chunk = the sweeps between two residual checks
    return run_sweeps_graphed(ctx, prm->max_iter, check_every, true,
                              [&](int it, int pos) { return launch_sweep(it, pos == 0); }, iters_out);
}

// Nesterov (mode 0) or Chebyshev (mode 1) loop with joint rows over x = [v; Lambda]: the
// standard sweep (contact pass, body pass, chain solve) around the k_accj_* state kernels.
// Chebyshev's spectral radius: 8 power iterations of v -> v + S^-1 [W^T dp; 0], dp = -R W v_head,
// over the twists and Lambda, from ones / sqrt(nv + joint rows).
static int run_accel_joints(cdx_ctx* ctx, const cdx_solver_params* prm, int mode, int* iters_out) {
    const int nb = ctx->nb, nc = ctx->nc, nj = ctx->njrows;
    const int n0 = nb * kBodyStride, n1 = nj, n2 = 3 * nc, m1 = std::max(n1, 1);
    const int gb = grid_for(nb), gb8 = grid_for(nb * kLanesPerBody), gc4 = grid_for(nc * kLanesPerContact);
    const int gn = grid_for(n0 + n1 + n2), gv = grid_for(n0 + n1);
    CDX_TRY(ensure_alloc(ctx, &ctx->d_accj, 3 * (size_t)n0 + 3 * (size_t)m1 + 2 * (size_t)n2, &ctx->accj_cap,
                         "accelerated loop state"));
    double* Kx = ctx->d_accj;
    double* K1x = Kx + n0;
    double* Yx = K1x + n0;  // Nesterov: the sweep input's twists; Chebyshev: the power-iteration vector
    double* Kl = Yx + n0;
    double* K1l = Kl + m1;
    double* Vl = K1l + m1;
    double* Kp = Vl + m1;
    double* K1p = Kp + n2;
    CDX_TRY(ensure_alloc(ctx, &ctx->d_partial2, 2 * (size_t)gb, &ctx->partial2_cap, "accelerated loop partials"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_partial, (size_t)(gb8 + ctx->nchains), &ctx->partial_cap, "residual partials"));
    const size_t bx = sizeof(double) * (size_t)n0, bl = sizeof(double) * (size_t)n1, bp = sizeof(double) * (size_t)n2;
    for (double* d : {Kx, K1x, Yx}) CDX_HIP(ctx, hipMemcpyAsync(d, ctx->d_x, bx, hipMemcpyDeviceToDevice, ctx->stream));
    if (n1)
        for (double* d : {Kl, K1l}) CDX_HIP(ctx, hipMemcpyAsync(d, ctx->d_jlam, bl, hipMemcpyDeviceToDevice, ctx->stream));
    for (double* d : {Kp, K1p}) CDX_HIP(ctx, hipMemcpyAsync(d, ctx->d_p, bp, hipMemcpyDeviceToDevice, ctx->stream));
    double* st_scalars = nullptr;
    if (mode == 0) {
        CDX_TRY(ensure_alloc(ctx, &ctx->d_nest, 8, nullptr, "nesterov scalars"));
        const double n0s[8] = {0.0, 1.0, INFINITY, 0.0, 0.0, 0.0, 0.0, 0.0};
        CDX_HIP(ctx, hipMemcpyAsync(ctx->d_nest, n0s, sizeof(n0s), hipMemcpyHostToDevice, ctx->stream));
        st_scalars = ctx->d_nest;
    } else {
        CDX_TRY(ensure_alloc(ctx, &ctx->d_cheb, 8, nullptr, "chebyshev scalars"));
        const double z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        CDX_HIP(ctx, hipMemcpyAsync(ctx->d_cheb, z, sizeof(z), hipMemcpyHostToDevice, ctx->stream));
        st_scalars = ctx->d_cheb;
        CDX_LAUNCH(ctx, "cheb_power", k_accj_fill, gv, kBlock, nb, ctx->d_dof, nj, 1.0 / std::sqrt((double)(ctx->nv + nj)), Yx,
                   Vl);
        for (int k = 0; k < 8; ++k) {
            CDX_LAUNCH(ctx, "cheb_power", k_cheb_lin_contacts4, gc4, kBlock, nc, gc4, ctx->d_ida, ctx->d_idb, ctx->d_J,
                       ctx->d_R, ctx->d_mu, Yx, ctx->d_imp, ctx->d_cheb, nullptr);
            CDX_LAUNCH(ctx, "cheb_power", k_pj_bodies8, gb8, kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_incptr,
                       ctx->d_incidx, ctx->d_imp, Yx, 0, 0.0, 0.0, ctx->d_partial, ctx->d_ctrl, ctx->d_jslot, ctx->d_jtr,
                       nullptr);
            CDX_TRY(joints_apply(ctx, Yx, Vl, ctx->d_ctrl));
            CDX_LAUNCH(ctx, "cheb_power", k_accj_sq, gb, kBlock, nb, ctx->d_dof, Yx, nj, Vl, ctx->d_partial2, ctx->d_cheb);
            CDX_LAUNCH(ctx, "cheb_power", k_cheb_norm, 1, kBlock, gb, ctx->d_partial2, ctx->d_cheb);
            CDX_LAUNCH(ctx, "cheb_power", k_accj_scale, gv, kBlock, nb, nj, Yx, Vl, ctx->d_cheb);
        }
        CDX_LAUNCH(ctx, "cheb_power", k_cheb_start, 1, 1, ctx->d_cheb);
    }
    auto sweep = [&](int it) -> int {
        CDX_LAUNCH(ctx, "pj_contacts", k_pj_contacts4<false>, gc4, kBlock, nc, gc4, ctx->d_ida, ctx->d_idb, ctx->d_J,
                   ctx->d_R, ctx->d_bias, ctx->any_bias ? 1 : 0, ctx->d_mu, ctx->d_p, ctx->d_x, ctx->d_imp, ctx->d_ctrl);
        CDX_LAUNCH(ctx, "pj_bodies", k_pj_bodies8, gb8, kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_incptr, ctx->d_incidx,
                   ctx->d_imp, ctx->d_x, 0, prm->tol_rel, prm->tol_abs, ctx->d_partial, ctx->d_ctrl, ctx->d_jslot, ctx->d_jtr,
                   nullptr);
        CDX_TRY(joints_solve(ctx, 0, 0, prm->tol_rel, prm->tol_abs, nullptr));
        CDX_LAUNCH(ctx, "accj_residual", k_accj_residual, gb, kBlock, mode, nb, ctx->d_dof, ctx->d_x, mode == 0 ? Yx : K1x,
                   Kx, nj, ctx->d_jlam, prm->tol_rel, prm->tol_abs, ctx->d_partial2, ctx->d_ctrl);
        if (mode == 0)
            CDX_LAUNCH(ctx, "pj_residual", k_pjn_finalize, 1, kBlock, gb, ctx->d_partial2, ctx->nv, it,
                       prm->nesterov_beta_threshold, prm->nesterov_restart_limit, ctx->d_ctrl, ctx->d_res, ctx->h_flag_dev,
                       ctx->d_nest);
        else
            CDX_LAUNCH(ctx, "pjc_residual", k_pjc_finalize, 1, kBlock, gb, ctx->d_partial2, ctx->nv, it, ctx->d_ctrl,
                       ctx->d_res, ctx->h_flag_dev, ctx->d_cheb);
        CDX_LAUNCH(ctx, "accj_update", k_accj_update, gn, kBlock, mode, n0, n1, n2, ctx->d_x, ctx->d_jlam, ctx->d_p, Kx, Kl,
                   Kp, mode == 0 ? Yx : K1x, mode == 0 ? nullptr : K1l, mode == 0 ? nullptr : K1p, st_scalars, ctx->d_ctrl);
        return hip_check(ctx, hipGetLastError(), "pj accelerated joint sweep launch");
    };
    const int st = run_sweeps(ctx, prm->max_iter, 10, 4, sweep, iters_out);
    if (st) return st;
    CDX_LAUNCH(ctx, "accj_finish", k_accj_finish, gn, kBlock, mode, n0, n1, n2, ctx->d_x, ctx->d_jlam, ctx->d_p,
               mode == 0 ? Kx : K1x, mode == 0 ? Kl : K1l, mode == 0 ? Kp : K1p, ctx->d_ctrl);
    CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return hip_check(ctx, hipGetLastError(), "pj accelerated joint finish");
}

// chebyshev_loop with estimateSpectralRadius (projected_jacobi.cpp:155-228, 470-476): buffers
// X = d_x, XK = d_uprev, XK1 = d_xk1, P = d_p, PK = d_pk, PK1 = d_pk1.
static int run_chebyshev(cdx_ctx* ctx, const cdx_solver_params* prm, int* iters_out) {
    const int nb = ctx->nb, nc = ctx->nc;
    const int gb8 = grid_for(nb * kLanesPerBody);
    const int gc4 = grid_for(nc * kLanesPerContact);
    const int gb6 = grid_for(nb * 6);
    const size_t n8 = (size_t)nb * kBodyStride;
    const double* Rb = prm->block_precond ? ctx->d_Rb : nullptr;
    CDX_TRY(ensure_alloc(ctx, &ctx->d_xk1, n8, &ctx->xk1_cap, "chebyshev xk1"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_pk, 3 * (size_t)nc, &ctx->pk_cap, "chebyshev pk"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_pk1, 3 * (size_t)nc, &ctx->pk1_cap, "chebyshev pk1"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_cheb, 8, nullptr, "chebyshev scalars"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_partial2, 2 * (size_t)gb8, &ctx->partial2_cap, "chebyshev partials"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_partial, (size_t)gb8, &ctx->partial_cap, "power-iteration partials"));
    // spectral radius (the vector lives in d_xk1 until the loop starts)
    const double c0[8] = {1.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    CDX_HIP(ctx, hipMemcpyAsync(ctx->d_cheb, c0, sizeof(c0), hipMemcpyHostToDevice, ctx->stream));
    CDX_LAUNCH(ctx, "cheb_power", k_cheb_fill, gb6, kBlock, nb, ctx->d_dof, 1.0 / std::sqrt((double)ctx->nv), ctx->d_xk1);
    for (int it = 0; it < 8; ++it) {
        CDX_LAUNCH(ctx, "cheb_power", k_cheb_lin_contacts4, gc4, kBlock, nc, gc4, ctx->d_ida, ctx->d_idb, ctx->d_J, ctx->d_R,
                   ctx->d_mu, ctx->d_xk1, ctx->d_imp, ctx->d_cheb, Rb);
        CDX_LAUNCH(ctx, "cheb_power", k_cheb_lin_bodies8, gb8, kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_incptr,
                   ctx->d_incidx, ctx->d_imp, ctx->d_xk1, ctx->d_partial, ctx->d_cheb, gyro_sinv(ctx, prm));
        CDX_LAUNCH(ctx, "cheb_power", k_cheb_norm, 1, kBlock, gb8, ctx->d_partial, ctx->d_cheb);
        CDX_LAUNCH(ctx, "cheb_power", k_cheb_scale, gb6, kBlock, nb, ctx->d_dof, ctx->d_xk1, ctx->d_cheb);
    }
    CDX_LAUNCH(ctx, "cheb_power", k_cheb_start, 1, 1, ctx->d_cheb);
    CDX_HIP(ctx, hipGetLastError());
    // xk = xk1 = x0, pk = pk1 = p0
    CDX_HIP(ctx, hipMemcpyAsync(ctx->d_uprev, ctx->d_x, sizeof(double) * n8, hipMemcpyDeviceToDevice, ctx->stream));
    CDX_HIP(ctx, hipMemcpyAsync(ctx->d_xk1, ctx->d_x, sizeof(double) * n8, hipMemcpyDeviceToDevice, ctx->stream));
    CDX_HIP(ctx, hipMemcpyAsync(ctx->d_pk, ctx->d_p, sizeof(double) * 3 * (size_t)nc, hipMemcpyDeviceToDevice, ctx->stream));
    CDX_HIP(ctx, hipMemcpyAsync(ctx->d_pk1, ctx->d_p, sizeof(double) * 3 * (size_t)nc, hipMemcpyDeviceToDevice, ctx->stream));
    // the range sweep (one lane per contact, per-body partials) on large problems, row scaling only
    bool ranges = false;
    if (!Rb && (ctx->range_mode == 2 || (ctx->range_mode == 1 && nc >= kRangeAutoContacts))) {
        if (!ctx->ranges_valid) CDX_TRY(build_ranges(ctx));
        ranges = ctx->ranges_ok;
    }
    ctx->ranges_used = ranges;
    const bool geo = ranges && ctx->geo_ok && ctx->geo_mode;
    ctx->geo_used = geo;
    const int ng = ctx->nranges, per = ctx->range_per_used;
    const int nthr = ranges ? std::max(64, std::min({ctx->range_block, 256, pad64(std::min(per, nc))})) : 0;
    const size_t shm = ranges ? range_lds_bytes(std::min(per, nc), ctx->range_maxp, ctx->range_maxl) : 0;
    auto launch_sweep = [&](int it, int) -> int {
        if (ranges) {
#define CDX_CHEB_RANGE(GEO)                                                                                             \
    CDX_LAUNCH_SHM(ctx, "pjc_range", (k_pj_range<false, GEO, kLoopCheb>), ng, nthr, shm, ng, nc, per, std::min(per, nc), \
                   ctx->range_maxp, ctx->d_rg_p, ctx->d_rg_l, ctx->d_pr_ptr, ctx->d_pr_out, ctx->d_pr_inc, ctx->d_ida,  \
                   ctx->d_idb, ctx->d_J, ctx->d_R, ctx->d_bias, ctx->any_bias ? 1 : 0, ctx->d_mu, ctx->d_p, ctx->d_x,   \
                   ctx->d_part, ctx->d_ctrl, ctx->d_geo, ctx->d_q, ctx->d_pk, ctx->d_pk1, ctx->d_uprev, ctx->d_cheb)
            if (geo)
                CDX_CHEB_RANGE(true);
            else
                CDX_CHEB_RANGE(false);
#undef CDX_CHEB_RANGE
            CDX_LAUNCH(ctx, "pjc_bodies_parts", k_pjc_bodies_parts, gb8, kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_bp_ptr,
                       ctx->d_bp_ovf, ctx->d_part, ctx->d_x, ctx->d_uprev, ctx->d_xk1, prm->tol_rel, prm->tol_abs,
                       ctx->d_partial2, ctx->d_ctrl, ctx->d_cheb, gyro_sinv(ctx, prm));
        } else {
            CDX_LAUNCH(ctx, "pjc_contacts", k_pjc_contacts4, gc4, kBlock, nc, gc4, ctx->d_ida, ctx->d_idb, ctx->d_J, ctx->d_R,
                       ctx->d_bias, ctx->any_bias ? 1 : 0, ctx->d_mu, ctx->d_p, ctx->d_pk, ctx->d_pk1, ctx->d_x,
                       ctx->d_uprev, ctx->d_imp, ctx->d_ctrl, ctx->d_cheb, Rb);
            CDX_LAUNCH(ctx, "pjc_bodies", k_pjc_bodies8, gb8, kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_incptr,
                       ctx->d_incidx, ctx->d_imp, ctx->d_x, ctx->d_uprev, ctx->d_xk1, prm->tol_rel, prm->tol_abs,
                       ctx->d_partial2, ctx->d_ctrl, ctx->d_cheb, gyro_sinv(ctx, prm));
        }
        CDX_LAUNCH(ctx, "pjc_residual", k_pjc_finalize, 1, kBlock, gb8, ctx->d_partial2, ctx->nv, it, ctx->d_ctrl, ctx->d_res,
                   ctx->h_flag_dev, ctx->d_cheb);
        return hip_check(ctx, hipGetLastError(), "pj chebyshev sweep launch");
    };
    const int st = run_sweeps_graphed(ctx, prm->max_iter, 8, true, launch_sweep, iters_out);
    if (st) return st;
    CDX_LAUNCH(ctx, "pjc_finish", k_pjc_finish_bodies, gb6, kBlock, nb, ctx->d_dof, ctx->d_x, ctx->d_uprev, ctx->d_xk1,
               ctx->d_ctrl, ctx->d_cheb);
    CDX_LAUNCH(ctx, "pjc_finish", k_pjc_finish_contacts, grid_for(3 * nc), kBlock, nc, ctx->d_ida, ctx->d_idb, ctx->d_mu,
               ctx->d_p, ctx->d_pk, ctx->d_pk1, ctx->d_ctrl, ctx->d_cheb);
    CDX_HIP(ctx, hipGetLastError());
    CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CDX_OK;
}

// anderson_loop (projected_jacobi.cpp:224-310): plain sweeps in place on (x, p) with the residual
// every iteration, then push / solve / candidate / commit; 8 iterations per graph chunk.
static int run_anderson(cdx_ctx* ctx, const cdx_solver_params* prm, int* iters_out) {
    const int nb = ctx->nb, nc = ctx->nc, m = prm->anderson_m, S = m + 1;
    const int n = 6 * nb + 3 * nc;
    const bool block = prm->block_precond != 0;
    const int gb8 = grid_for(nb * kLanesPerBody);
    const int gc4 = grid_for(nc * kLanesPerContact);
    const int gn = std::min(grid_for(n), kAndBlocks);
    CDX_TRY(ensure_alloc(ctx, &ctx->d_and, (size_t)(2 * S + 1) * n, &ctx->and_cap, "anderson history"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_and_gam, kAndMaxM, nullptr, "anderson coefficients"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_and_part, (size_t)kAndBlocks * kAndDots, nullptr, "anderson partials"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_and_st, 8, nullptr, "anderson state"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_partial, (size_t)gb8, &ctx->partial_cap, "residual partials"));
    double* ring = ctx->d_and;  // Z: S slots, F: S slots, then the candidate
    double* cand = ctx->d_and + (size_t)2 * S * n;
    const int st0[8] = {0, m, 0, 0, 0, 0, 0, 0};  // empty history, head = m so that z_0 goes to slot 0
    CDX_HIP(ctx, hipMemcpyAsync(ctx->d_and_st, st0, sizeof(st0), hipMemcpyHostToDevice, ctx->stream));
    CDX_LAUNCH(ctx, "and_init", k_and_init, gn, kBlock, n, nb, nc, ctx->d_dof, ctx->d_ida, ctx->d_idb, ctx->d_mu, ctx->d_x,
               ctx->d_p, ring);
    CDX_HIP(ctx, hipGetLastError());
    const double rows = (double)ctx->nv + (double)ctx->nrows;  // the rows of the reference's dF
    auto launch_iter = [&](int it, int) -> int {
        if (block)
            CDX_LAUNCH(ctx, "pj_contacts", k_pj_contacts4<true>, gc4, kBlock, nc, gc4, ctx->d_ida, ctx->d_idb, ctx->d_J,
                       ctx->d_Rb, ctx->d_bias, ctx->any_bias ? 1 : 0, ctx->d_mu, ctx->d_p, ctx->d_x, ctx->d_imp, ctx->d_ctrl);
        else
            CDX_LAUNCH(ctx, "pj_contacts", k_pj_contacts4<false>, gc4, kBlock, nc, gc4, ctx->d_ida, ctx->d_idb, ctx->d_J,
                       ctx->d_R, ctx->d_bias, ctx->any_bias ? 1 : 0, ctx->d_mu, ctx->d_p, ctx->d_x, ctx->d_imp, ctx->d_ctrl);
        CDX_LAUNCH(ctx, "pj_bodies", k_pj_bodies8, gb8, kBlock, nb, ctx->d_dof, ctx->d_minv, ctx->d_incptr, ctx->d_incidx,
                   ctx->d_imp, ctx->d_x, 1, prm->tol_rel, prm->tol_abs, ctx->d_partial, ctx->d_ctrl, nullptr, nullptr,
                   gyro_sinv(ctx, prm));
        CDX_TRY(launch_residual_finalize(ctx, gb8, it, 0, 1));
        CDX_LAUNCH(ctx, "and_push", k_and_push, gn, kBlock, n, nb, nc, ctx->d_dof, ctx->d_ida, ctx->d_idb, ctx->d_mu,
                   ctx->d_x, ctx->d_p, ring, m, ctx->d_and_st, ctx->d_and_part, ctx->d_ctrl);
        CDX_LAUNCH(ctx, "and_solve", k_and_solve, 1, kBlock, gn, ctx->d_and_part, m, rows, ctx->d_and_st, ctx->d_and_gam,
                   ctx->d_ctrl);
        CDX_LAUNCH(ctx, "and_candidate", k_and_candidate, gn, kBlock, n, nb, nc, ctx->d_dof, ctx->d_ida, ctx->d_idb,
                   ctx->d_mu, ctx->d_x, ctx->d_p, ring, m, ctx->d_and_st, ctx->d_and_gam, cand, ctx->d_ctrl);
        CDX_LAUNCH(ctx, "and_commit", k_and_commit, gn, kBlock, n, nb, nc, ctx->d_dof, ctx->d_ida, ctx->d_idb, ctx->d_mu,
                   ctx->d_x, ctx->d_p, ring, m, ctx->d_and_st, cand, ctx->d_ctrl);
        return hip_check(ctx, hipGetLastError(), "pj anderson iteration launch");
    };
    const int st = run_sweeps_graphed(ctx, prm->max_iter, 8, true, launch_iter, iters_out);
    if (st) return st;
    CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CDX_OK;
}

// ---------------------------------------------------------------------------------------------
// Allocation-free projected Jacobi (cdx_pj_solve_noalloc): the standard loop of cdx_pj_solve on
// caller-owned device arrays.  Every buffer the solve needs beyond its inputs and outputs (padded
// body state, the contact rows in component-major order, R, b, p, the side-impulse records, the
// incidence lists, residual partials, control words) lives in one file-scope device array,
// `scratch`, laid out for the maximum sizes below; the calls allocate nothing.  Each GPU holds its
// own copy for the life of the process; calls on one GPU share it, so calls that overlap on one
// GPU (two host threads, or two streams) race.
// ---------------------------------------------------------------------------------------------
namespace cdx {

constexpr int kNaMaxBodies = CDX_NOALLOC_MAX_BODIES;
constexpr int kNaMaxContacts = CDX_NOALLOC_MAX_CONTACTS;
constexpr int kNaMaxDevices = 64;
constexpr size_t kNaNB = kNaMaxBodies, kNaNC = kNaMaxContacts;
constexpr size_t kNaPartials = (kNaNB * kLanesPerBody + kBlock - 1) / kBlock;
// double region (offsets in doubles)
constexpr size_t kNaX = 0;                          // solver state x, padded [nb][8]
constexpr size_t kNaV = kNaX + kNaNB * kBodyStride;  // v_n
constexpr size_t kNaF = kNaV + kNaNB * kBodyStride;  // f_ext
constexpr size_t kNaG = kNaF + kNaNB * kBodyStride;  // f_gyro
constexpr size_t kNaM = kNaG + kNaNB * kBodyStride;  // [m^-1, I^-1] per body
constexpr size_t kNaJ = kNaM + kNaNB * kMinvStride;  // J[27][nc]
constexpr size_t kNaMu = kNaJ + kJ * kNaNC;
constexpr size_t kNaRest = kNaMu + kNaNC;
constexpr size_t kNaVb = kNaRest + 3 * kNaNC;
constexpr size_t kNaP = kNaVb + 3 * kNaNC;
constexpr size_t kNaR = kNaP + 3 * kNaNC;
constexpr size_t kNaBias = kNaR + 3 * kNaNC;
constexpr size_t kNaImp = kNaBias + 3 * kNaNC;       // side-impulse records [nc][9]
constexpr size_t kNaPart = kNaImp + 9 * kNaNC;       // residual partials
constexpr size_t kNaRes = kNaPart + kNaPartials;     // [0] residual
constexpr size_t kNaDoubles = kNaRes + 4;
// int region (offsets in ints, after the doubles)
constexpr size_t kNaDof = 0;
constexpr size_t kNaVoff = kNaDof + kNaNB;           // [nb+1] stacked offsets
constexpr size_t kNaIncPtr = kNaVoff + kNaNB + 1;    // [nb+1]
constexpr size_t kNaCnt = kNaIncPtr + kNaNB + 1;     // [nb] incidence counts, then fill cursors
constexpr size_t kNaIncIdx = kNaCnt + kNaNB;         // [2nc]
constexpr size_t kNaIda = kNaIncIdx + 2 * kNaNC;
constexpr size_t kNaIdb = kNaIda + kNaNC;
constexpr size_t kNaCtrl = kNaIdb + kNaNC;           // [4] done, iterations, status, chunk counter
constexpr size_t kNaFlags = kNaCtrl + 4;             // [0] input error, [1] rows, [2] nv, [4] done mirror
constexpr size_t kNaInts = kNaFlags + 8;
constexpr size_t kNaBytes = kNaDoubles * sizeof(double) + kNaInts * sizeof(int);

static __device__ __align__(256) double scratch[(kNaBytes + sizeof(double) - 1) / sizeof(double)];

struct NaBuf {
    double *x, *v, *f, *g, *m, *J, *mu, *rest, *vb, *p, *R, *bias, *imp, *part, *res;
    int *dof, *voff, *incptr, *cnt, *incidx, *ida, *idb, *ctrl, *flags;
};

static NaBuf na_buf(double* base) {
    int* ib = reinterpret_cast<int*>(base + kNaDoubles);
    return NaBuf{base + kNaX,    base + kNaV,    base + kNaF,    base + kNaG,   base + kNaM,    base + kNaJ,
                 base + kNaMu,   base + kNaRest, base + kNaVb,   base + kNaP,   base + kNaR,    base + kNaBias,
                 base + kNaImp,  base + kNaPart, base + kNaRes,  ib + kNaDof,   ib + kNaVoff,   ib + kNaIncPtr,
                 ib + kNaCnt,    ib + kNaIncIdx, ib + kNaIda,    ib + kNaIdb,   ib + kNaCtrl,   ib + kNaFlags};
}

// Input errors are merged by atomicMax, so a call reports the same status as cdx_set_bodies /
// cdx_set_contacts / cdx_pj_solve would on the same input: CDX_BAD_ARG before CDX_NONFINITE
// before CDX_NOT_FACTORIZED.
__device__ inline void na_flag(int* flags, int code) { atomicMax(&flags[0], code); }

// Clears what must start at zero: control words, flags, incidence counts; residual = NaN.
__global__ void k_na_clear(int nb, int* __restrict__ cnt, int* __restrict__ ctrl, int* __restrict__ flags,
                           double* __restrict__ res) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb) cnt[i] = 0;
    if (i < 4) ctrl[i] = 0;
    if (i < 8) flags[i] = 0;
    if (i == 0) res[0] = __longlong_as_double(0x7ff8000000000000ll);
}

// Exclusive prefix sum of n values of `in` into out[0..n] by one workgroup of kNaScan threads
// (each thread sums a contiguous segment, then a scan over the segment sums in LDS).
constexpr int kNaScan = 1024;
__device__ inline void na_block_scan(int n, const int* __restrict__ in, int* __restrict__ out, int* __restrict__ lds) {
    const int t = threadIdx.x;
    const int per = (n + kNaScan - 1) / kNaScan;
    const int i0 = min(n, t * per), i1 = min(n, i0 + per);
    int s = 0;
    for (int i = i0; i < i1; ++i) s += in[i];
    lds[t] = s;
    __syncthreads();
    for (int off = 1; off < kNaScan; off <<= 1) {
        const int v = t >= off ? lds[t - off] : 0;
        __syncthreads();
        lds[t] += v;
        __syncthreads();
    }
    int run = lds[t] - s;  // exclusive prefix of this segment
    for (int i = i0; i < i1; ++i) {
        out[i] = run;
        run += in[i];
    }
    if (t == kNaScan - 1) out[n] = lds[t];
}

// dof per body (3 or 6, else CDX_BAD_ARG) and the stacked offsets voff; flags[2] = nv.
__global__ void __launch_bounds__(kNaScan) k_na_offsets(int nb, const int* __restrict__ dof_in, int* __restrict__ dof,
                                                        int* __restrict__ voff, int* __restrict__ flags) {
    __shared__ int lds[kNaScan];
    for (int b = threadIdx.x; b < nb; b += kNaScan) {
        int d = dof_in[b];
        if (d != 3 && d != 6) {
            na_flag(flags, CDX_BAD_ARG);
            d = 6;
        }
        dof[b] = d;
    }
    __syncthreads();
    na_block_scan(nb, dof, voff, lds);
    if (threadIdx.x == kNaScan - 1) flags[2] = voff[nb];
}

// Bodies into the padded layout (cdx_set_bodies): [v; omega], f_ext, f_gyro (computed as
// -omega x (I omega) when not given), [m^-1, I^-1].
__global__ void k_na_pack_bodies(int nb, const int* __restrict__ dof, const int* __restrict__ voff,
                                 const double* __restrict__ minv, const double* __restrict__ v,
                                 const double* __restrict__ fext, const double* __restrict__ fgyro,
                                 double* __restrict__ xv, double* __restrict__ vv, double* __restrict__ fv,
                                 double* __restrict__ gv, double* __restrict__ mv, int* __restrict__ flags) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int o = voff[b], d = dof[b];
    double* X = xv + (size_t)b * kBodyStride;
    double* V = vv + (size_t)b * kBodyStride;
    double* F = fv + (size_t)b * kBodyStride;
    double* G = gv + (size_t)b * kBodyStride;
    double* M = mv + (size_t)b * kMinvStride;
    if (minv[o] != minv[o + 1] || minv[o] != minv[o + 2]) na_flag(flags, CDX_BAD_ARG);
    M[0] = minv[o];
    M[1] = M[2] = M[3] = 0.0;
    for (int k = 0; k < kBodyStride; ++k) X[k] = V[k] = F[k] = G[k] = 0.0;
    for (int k = 0; k < d; ++k) {
        const double mk = minv[o + k], vk = v[o + k], fk = fext[o + k], gk = fgyro ? fgyro[o + k] : 0.0;
        if (!isfinite(mk) || !isfinite(vk) || !isfinite(fk) || !isfinite(gk)) na_flag(flags, CDX_NONFINITE);
        if (!(mk > 0)) na_flag(flags, CDX_NOT_FACTORIZED);
        X[k] = V[k] = vk;
        F[k] = fk;
        G[k] = gk;
    }
    if (d == 6) {
        for (int k = 0; k < 3; ++k) M[1 + k] = minv[o + 3 + k];
        if (!fgyro) gyro_force3(V + 3, M + 1, G + 3);
    }
}

// Contacts into the component-major layout of cdx_set_contacts: static-static contacts carry no
// rows (both ids -1), a static side and a point mass no angular Jacobian; mu, restitution, bias
// and warm start only on the rows the contact has.  Counts rows and each body's incidences.
__global__ void k_na_pack_contacts(int nc, int nb, const int* __restrict__ dof, const int* __restrict__ body_a,
                                   const int* __restrict__ body_b, const double* __restrict__ mu,
                                   const double* __restrict__ frame, const double* __restrict__ angA,
                                   const double* __restrict__ angB, const double* __restrict__ rest,
                                   const double* __restrict__ vbias, const double* __restrict__ warm,
                                   int* __restrict__ ida, int* __restrict__ idb, double* __restrict__ J,
                                   double* __restrict__ mus, double* __restrict__ rs, double* __restrict__ vbs,
                                   double* __restrict__ ps, int* __restrict__ cnt, int* __restrict__ flags) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    int a = body_a[c], b = body_b[c];
    if (a < -1 || a >= nb || b < -1 || b >= nb || (a >= 0 && a == b)) {
        na_flag(flags, CDX_BAD_ARG);
        a = b = -1;
    }
    const double m = mu[c];
    bool finite = isfinite(m);
    const bool fr = m > 0;
    const bool angAok = a >= 0 && dof[a] == 6, angBok = b >= 0 && dof[b] == 6;
    for (int k = 0; k < 9; ++k) {
        const double fk = frame[9 * (size_t)c + k], ak = angA[9 * (size_t)c + k], bk = angB[9 * (size_t)c + k];
        finite = finite && isfinite(fk) && isfinite(ak) && isfinite(bk);
        J[(size_t)k * nc + c] = fk;
        J[(size_t)(9 + k) * nc + c] = angAok ? ak : 0.0;
        J[(size_t)(18 + k) * nc + c] = angBok ? bk : 0.0;
    }
    const bool has_rows = a >= 0 || b >= 0;
    for (int k = 0; k < 3; ++k) {
        const double rk = rest[3 * (size_t)c + k], vk = vbias[3 * (size_t)c + k];
        const double wk = warm ? warm[3 * (size_t)c + k] : 0.0;
        finite = finite && isfinite(rk) && isfinite(vk) && isfinite(wk);
        const bool row = has_rows && (k == 0 || fr);
        rs[(size_t)k * nc + c] = row ? rk : 0.0;
        vbs[(size_t)k * nc + c] = row ? vk : 0.0;
        ps[(size_t)k * nc + c] = row ? wk : 0.0;
    }
    if (!finite) na_flag(flags, CDX_NONFINITE);
    mus[c] = fr ? m : 0.0;
    ida[c] = has_rows ? a : -1;
    idb[c] = has_rows ? b : -1;
    if (has_rows) atomicAdd(&flags[1], fr ? 3 : 1);
    if (a >= 0) atomicAdd(&cnt[a], 1);
    if (b >= 0) atomicAdd(&cnt[b], 1);
}

// incptr = exclusive prefix of the counts; the counts become fill cursors.
__global__ void __launch_bounds__(kNaScan) k_na_incidence_ptr(int nb, int* __restrict__ cnt, int* __restrict__ incptr) {
    __shared__ int lds[kNaScan];
    na_block_scan(nb, cnt, incptr, lds);
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += kNaScan) cnt[b] = incptr[b];
}

__global__ void k_na_incidence_fill(int nc, const int* __restrict__ ida, const int* __restrict__ idb,
                                    int* __restrict__ cursor, int* __restrict__ incidx) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int a = ida[c], b = idb[c];
    if (a >= 0) incidx[atomicAdd(&cursor[a], 1)] = 2 * c;
    if (b >= 0) incidx[atomicAdd(&cursor[b], 1)] = 2 * c + 1;
}

// Each body's incidences into ascending order (= contact order, as build_incidence lists them),
// so the body pass sums in the same fixed order as cdx_pj_solve: insertion sort for short lists,
// heap sort beyond.
__global__ void k_na_incidence_sort(int nb, const int* __restrict__ incptr, int* __restrict__ incidx) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    int* e = incidx + incptr[b];
    const int n = incptr[b + 1] - incptr[b];
    if (n <= 32) {
        for (int i = 1; i < n; ++i) {
            const int v = e[i];
            int j = i - 1;
            while (j >= 0 && e[j] > v) {
                e[j + 1] = e[j];
                --j;
            }
            e[j + 1] = v;
        }
        return;
    }
    auto sift = [&](int root, int end) {
        while (2 * root + 1 < end) {
            int child = 2 * root + 1;
            if (child + 1 < end && e[child] < e[child + 1]) ++child;
            if (e[root] >= e[child]) return;
            const int t = e[root];
            e[root] = e[child];
            e[child] = t;
            root = child;
        }
    };
    for (int i = n / 2 - 1; i >= 0; --i) sift(i, n);
    for (int end = n - 1; end > 0; --end) {
        const int t = e[0];
        e[0] = e[end];
        e[end] = t;
        sift(0, end);
    }
}

// Solved velocities back to the stacked layout, impulses to [nc][3] (cdx_get_velocities /
// cdx_get_impulses).
__global__ void k_na_unpack(int nb, int nc, const int* __restrict__ dof, const int* __restrict__ voff,
                            const double* __restrict__ x, const double* __restrict__ p, double* __restrict__ v_out,
                            double* __restrict__ imp_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb)
        for (int k = 0; k < dof[i]; ++k) v_out[voff[i] + k] = x[(size_t)i * kBodyStride + k];
    if (i < nc)
        for (int k = 0; k < 3; ++k) imp_out[3 * (size_t)i + k] = p[(size_t)k * nc + i];
}

static thread_local std::string g_na_err;
static std::mutex g_na_mutex;
static double* g_na_scratch[kNaMaxDevices] = {};

const char* noalloc_last_error() { return g_na_err.c_str(); }

static int na_fail(int code, const std::string& msg) {
    g_na_err = msg;
    return code;
}

static int na_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return CDX_OK;
    return na_fail(CDX_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}
#define NA_HIP(expr) CDX_TRY(na_hip((expr), #expr))

// Device address of this GPU's scratch array, looked up once per device with that device current.
// The array's name is only a host-side handle: kernels and copies get this address.
static int na_scratch_ptr(int dev, double** out) {
    if (dev < 0 || dev >= kNaMaxDevices) return na_fail(CDX_BAD_ARG, "cdx_pj_solve_noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (!g_na_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return na_fail(CDX_HIP_ERROR, std::string("cdx_pj_solve_noalloc: hipGetSymbolAddress(scratch): ") +
                                              hipGetErrorString(e));
        g_na_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_na_scratch[dev];
    return CDX_OK;
}

static int na_read_flags(hipStream_t s, const int* d_flags, int* h_flags) {
    NA_HIP(hipMemcpyAsync(h_flags, d_flags, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
    NA_HIP(hipStreamSynchronize(s));
    return CDX_OK;
}

}  // namespace cdx

extern "C" int cdx_pj_solve_noalloc(const cdx_device_problem* pb, double dt, double theta, const cdx_solver_params* prm,
                                    double* v_out, double* imp_out, cdx_solve_info* info, void* stream) {
    (void)theta;  // theta only enters joint rows, which this entry point does not take
    g_na_err.clear();
    if (info) {
        info->iters = 0;
        info->residual = std::nan("");
        info->nrows = 0;
        info->ncolors = 0;
    }
    if (!pb || !prm) return na_fail(CDX_BAD_ARG, "cdx_pj_solve_noalloc: null problem or params");
    const int nb = pb->nb, nc = pb->nc;
    if (nb < 0 || nc < 0) return na_fail(CDX_BAD_ARG, "cdx_pj_solve_noalloc: negative size");
    if (nb > kNaMaxBodies || nc > kNaMaxContacts)
        return na_fail(CDX_BAD_ARG, "cdx_pj_solve_noalloc: " + std::to_string(nb) + " bodies / " + std::to_string(nc) +
                                        " contacts exceed the allocation-free maximum of " +
                                        std::to_string(kNaMaxBodies) + " bodies / " + std::to_string(kNaMaxContacts) +
                                        " contacts");
    if (nb > 0 && (!pb->dof || !pb->minv || !pb->v || !pb->fext || !v_out))
        return na_fail(CDX_BAD_ARG, "cdx_pj_solve_noalloc: null body array");
    if (nc > 0 && (!pb->body_a || !pb->body_b || !pb->mu || !pb->frame || !pb->angA || !pb->angB || !pb->rest ||
                   !pb->vbias || !imp_out))
        return na_fail(CDX_BAD_ARG, "cdx_pj_solve_noalloc: null contact array");
    if (nb == 0 && nc > 0) return na_fail(CDX_NO_STATE, "cdx_pj_solve_noalloc: contacts without bodies");
    if (prm->max_iter < 1) return na_fail(CDX_BAD_ARG, "max_iter must be >= 1");
    if (prm->check_every < 1) return na_fail(CDX_BAD_ARG, "check_every must be >= 1");
    if (!std::isfinite(prm->alpha) || !std::isfinite(prm->tol_abs) || !std::isfinite(prm->tol_rel))
        return na_fail(CDX_BAD_ARG, "non-finite solver parameter");
    if (prm->accel != CDX_ACCEL_NONE || prm->block_precond || prm->implicit_gyro)
        return na_fail(CDX_BAD_ARG, "cdx_pj_solve_noalloc: standard loop with the row scaling and explicit gyroscopy only");
    if (nb == 0) return CDX_OK;
    int dev = 0;
    NA_HIP(hipGetDevice(&dev));
    double* base = nullptr;
    CDX_TRY(na_scratch_ptr(dev, &base));
    const NaBuf s = na_buf(base);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int n_clear = std::max(nb, 8);
    hipLaunchKernelGGL(k_na_clear, dim3(grid_for(n_clear)), dim3(kBlock), 0, st, nb, s.cnt, s.ctrl, s.flags, s.res);
    hipLaunchKernelGGL(k_na_offsets, dim3(1), dim3(kNaScan), 0, st, nb, pb->dof, s.dof, s.voff, s.flags);
    NA_HIP(hipGetLastError());
    int fl[4];
    CDX_TRY(na_read_flags(st, s.flags, fl));
    if (fl[0]) return na_fail(fl[0], "cdx_pj_solve_noalloc: dof must be 3 or 6");
    hipLaunchKernelGGL(k_na_pack_bodies, dim3(grid_for(nb)), dim3(kBlock), 0, st, nb, s.dof, s.voff, pb->minv, pb->v,
                       pb->fext, pb->fgyro, s.x, s.v, s.f, s.g, s.m, s.flags);
    if (nc > 0) {
        hipLaunchKernelGGL(k_na_pack_contacts, dim3(grid_for(nc)), dim3(kBlock), 0, st, nc, nb, s.dof, pb->body_a,
                           pb->body_b, pb->mu, pb->frame, pb->angA, pb->angB, pb->rest, pb->vbias, pb->warm, s.ida, s.idb,
                           s.J, s.mu, s.rest, s.vb, s.p, s.cnt, s.flags);
    }
    hipLaunchKernelGGL(k_na_incidence_ptr, dim3(1), dim3(kNaScan), 0, st, nb, s.cnt, s.incptr);
    if (nc > 0) {
        hipLaunchKernelGGL(k_na_incidence_fill, dim3(grid_for(nc)), dim3(kBlock), 0, st, nc, s.ida, s.idb, s.cnt, s.incidx);
        hipLaunchKernelGGL(k_na_incidence_sort, dim3(grid_for(nb)), dim3(kBlock), 0, st, nb, s.incptr, s.incidx);
    }
    NA_HIP(hipGetLastError());
    CDX_TRY(na_read_flags(st, s.flags, fl));
    if (fl[0] == CDX_BAD_ARG)
        return na_fail(CDX_BAD_ARG, "cdx_pj_solve_noalloc: bad input (anisotropic translational inverse mass, or a "
                                    "contact body index out of range)");
    if (fl[0] == CDX_NONFINITE) return na_fail(CDX_NONFINITE, "cdx_pj_solve_noalloc: non-finite input");
    if (fl[0] == CDX_NOT_FACTORIZED)
        return na_fail(CDX_NOT_FACTORIZED, "S matrix is not factorized: a dynamic dof has zero mass");
    const int nrows = fl[1], nv = fl[2];
    if (info) info->nrows = nrows;
    const bool have_contacts = nrows > 0;
    if (have_contacts)
        hipLaunchKernelGGL(k_pj_setup_contacts, dim3(grid_for(nc)), dim3(kBlock), 0, st, nc, s.ida, s.idb, s.J, s.mu,
                           s.rest, s.vb, s.m, s.v, prm->alpha, prm->warmstart, s.R, s.bias, s.p, s.imp, nullptr);
    hipLaunchKernelGGL(k_pj_setup_bodies, dim3(grid_for(nb)), dim3(kBlock), 0, st, nb, s.dof, s.m, s.v, s.f, s.g,
                       s.incptr, s.incidx, s.imp, dt, have_contacts ? 1 : 0, s.x, nullptr, nullptr, nullptr);
    NA_HIP(hipGetLastError());
    int iters = 0;
    if (have_contacts) {
        const int gb8 = grid_for(nb * kLanesPerBody), gc4 = grid_for(nc * kLanesPerContact);
        const int ce = prm->check_every;
        const int chunk = 4 * ce;  // sweeps between two polls of the done flag
        int ctrl[4] = {0, 0, 0, 0};
        for (int it = 0; it < prm->max_iter && !ctrl[0];) {
            const int end = std::min(prm->max_iter, it + chunk);
            for (; it < end; ++it) {
                const int check = (it % ce) == 0;
                // the bias array is always read: it holds exact zeros where a row has no bias,
                // and p - (R g + 0) rounds as p - R g
                hipLaunchKernelGGL(k_pj_contacts4<false>, dim3(gc4), dim3(kBlock), 0, st, nc, gc4, s.ida, s.idb, s.J,
                                   s.R, s.bias, 1, s.mu, s.p, s.x, s.imp, s.ctrl);
                hipLaunchKernelGGL(k_pj_bodies8, dim3(gb8), dim3(kBlock), 0, st, nb, s.dof, s.m, s.incptr, s.incidx,
                                   s.imp, s.x, check, prm->tol_rel, prm->tol_abs, s.part, s.ctrl, nullptr, nullptr, nullptr);
                if (check)
                    hipLaunchKernelGGL(k_residual_finalize, dim3(1), dim3(kBlock), 0, st, gb8, s.part, nv, it, ce, 0,
                                       s.ctrl, s.res, s.flags + 4);
            }
            NA_HIP(hipGetLastError());
            NA_HIP(hipMemcpyAsync(ctrl, s.ctrl, 3 * sizeof(int), hipMemcpyDeviceToHost, st));
            NA_HIP(hipStreamSynchronize(st));
        }
        iters = ctrl[0] ? ctrl[1] : prm->max_iter;
    }
    hipLaunchKernelGGL(k_na_unpack, dim3(grid_for(std::max(nb, nc))), dim3(kBlock), 0, st, nb, nc, s.dof, s.voff, s.x,
                       s.p, v_out, nc > 0 ? imp_out : nullptr);
    NA_HIP(hipGetLastError());
    double r = std::nan("");
    NA_HIP(hipMemcpyAsync(&r, s.res, sizeof(double), hipMemcpyDeviceToHost, st));
    NA_HIP(hipStreamSynchronize(st));
    if (info) {
        info->iters = iters;
        info->residual = r;
    }
    return CDX_OK;
}

extern "C" int cdx_set_range_sweep(cdx_ctx* ctx, int mode, int contacts_per_range, int block) {
    if (!ctx) return CDX_BAD_ARG;
    if (mode < 0 || mode > 2) return set_err(ctx, CDX_BAD_ARG, "cdx_set_range_sweep: mode must be 0, 1 or 2");
    if (contacts_per_range != 0 && (contacts_per_range < 16 || contacts_per_range > kRangeMaxContacts))
        return set_err(ctx, CDX_BAD_ARG, "cdx_set_range_sweep: contacts_per_range must be 0 or in 16..768");
    if (block != 0 && block != 256 && block != 512)
        return set_err(ctx, CDX_BAD_ARG, "cdx_set_range_sweep: block must be 0, 256 or 512");
    const int nblock = block ? block : 512;
    if (contacts_per_range != ctx->range_per) ctx->ranges_valid = false;
    ctx->range_mode = mode;
    ctx->range_per = contacts_per_range;
    ctx->range_block = nblock;
    return CDX_OK;
}

extern "C" int cdx_range_sweep_info(cdx_ctx* ctx, int* used, int* nranges, int* npartials, int* contacts_per_range) {
    if (!ctx) return CDX_BAD_ARG;
    if (used) *used = ctx->ranges_used ? 1 : 0;
    if (nranges) *nranges = ctx->ranges_ok ? ctx->nranges : 0;
    if (npartials) *npartials = ctx->ranges_ok ? ctx->npart : 0;
    if (contacts_per_range) *contacts_per_range = ctx->ranges_ok ? ctx->range_per_used : 0;
    return CDX_OK;
}

// Geometry form of the contact rows (k_geo_pack, k_pj_range<*, true>).
extern "C" int cdx_set_contact_points(cdx_ctx* ctx, const double* point) {
    if (!ctx) return CDX_BAD_ARG;
    hipSetDevice(ctx->device);
    const int nc = ctx->nc;
    ctx->geo_ok = false;
    ctx->geo_bad = -1;
    if (nc > 0 && !point) return set_err(ctx, CDX_BAD_ARG, "cdx_set_contact_points: null array");
    if (nc > 0 && !ctx->have_q) return set_err(ctx, CDX_NO_STATE, "cdx_set_contact_points: no positions were set");
    for (size_t i = 0; i < 3 * (size_t)nc; ++i)
        if (!std::isfinite(point[i])) return set_err(ctx, CDX_NONFINITE, "cdx_set_contact_points: non-finite point");
    if (nc == 0) {
        ctx->geo_bad = 0;
        return CDX_OK;
    }
    const int gc = grid_for(nc);
    CDX_TRY(ensure_alloc(ctx, &ctx->d_geo, 6 * (size_t)nc, &ctx->geo_cap, "contact lever arms"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_geo_pt, 3 * (size_t)nc, &ctx->geo_pt_cap, "contact points"));
    CDX_TRY(ensure_alloc(ctx, &ctx->d_geo_bad, (size_t)gc, &ctx->geo_bad_cap, "contact point check"));
    CDX_HIP(ctx, hipMemcpyAsync(ctx->d_geo_pt, point, sizeof(double) * 3 * (size_t)nc, hipMemcpyHostToDevice, ctx->stream));
    CDX_LAUNCH(ctx, "geo_pack", k_geo_pack, gc, kBlock, nc, ctx->d_ida, ctx->d_idb, ctx->d_dof, ctx->d_q, ctx->d_J,
               ctx->d_geo_pt, ctx->d_geo, ctx->d_geo_bad);
    std::vector<int> bad(gc);
    CDX_HIP(ctx, hipMemcpyAsync(bad.data(), ctx->d_geo_bad, sizeof(int) * (size_t)gc, hipMemcpyDeviceToHost, ctx->stream));
    CDX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    long long nbad = 0;
    for (int v : bad) nbad += v;
    ctx->geo_bad = (int)nbad;
    ctx->geo_ok = nbad == 0;
    return CDX_OK;
}

extern "C" int cdx_set_geometry_rows(cdx_ctx* ctx, int mode) {
    if (!ctx) return CDX_BAD_ARG;
    if (mode != 0 && mode != 1) return set_err(ctx, CDX_BAD_ARG, "cdx_set_geometry_rows: mode must be 0 or 1");
    ctx->geo_mode = mode;
    return CDX_OK;
}

extern "C" int cdx_geometry_rows_info(cdx_ctx* ctx, int* usable, int* mismatches, int* used) {
    if (!ctx) return CDX_BAD_ARG;
    if (usable) *usable = ctx->geo_ok ? 1 : 0;
    if (mismatches) *mismatches = ctx->geo_bad;
    if (used) *used = ctx->geo_used ? 1 : 0;
    return CDX_OK;
}

// cdx_common.hpp -- internal context and helpers of the gfx950 contact solve library.
//
// Device data layout (all FP64, HBM-resident for the life of a context):
//   bodies   : x/v/fext/fgyro/q padded to 8 doubles (64 B) per body so that a contact gathers
//              one body twist with three 16-byte loads from one 64-byte segment;
//              minv packed as 4 doubles (m^-1, Ixx^-1, Iyy^-1, Izz^-1).
//   contacts : structure of arrays, component-major ([k * nc + c]) so that every per-contact
//              scalar is one fully coalesced 512-byte wave load:
//              J[27][nc] (frame 9 | angA 9 | angB 9), R[3][nc], bias[3][nc], mu[nc], p[3][nc].
//   scatter  : imp[(2c + side) * 6 + k], the generalised impulse J_side^T dp of one contact side,
//              written by the contact pass and gathered per body through a CSR incidence list
//              (inc_ptr/inc_idx) so the body pass needs no atomics and sums in a fixed order.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cdx/cdx.h"

namespace cdx {

constexpr int kBodyStride = 8;  // doubles per padded body vector
constexpr int kMinvStride = 4;  // doubles per body inverse mass
constexpr int kJ = 27;          // Jacobian doubles per contact
constexpr int kBlock = 256;     // threads per workgroup (4 wave64)

struct KernelStat {
    int calls = 0;
    double ms = 0;
};

struct PendingEvent {
    std::string name;
    hipEvent_t start, stop;
};

}  // namespace cdx

struct cdx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // ---- bodies
    int nb = 0, nv = 0;
    std::vector<int> dof, voff, qoff;
    bool have_q = false;
    bool singular_mass = false;  // a dynamic dof with zero mass: S not factorizable
    double* d_x = nullptr;       // solver velocity state [nb*8]
    double* d_v = nullptr;       // body velocity v_n [nb*8]
    double* d_minv = nullptr;    // [nb*4]
    double* d_fext = nullptr;    // [nb*8]
    double* d_fgyro = nullptr;   // [nb*8]
    double* d_q = nullptr;       // [nb*8]
    int* d_voff = nullptr;       // [nb+1] stacked velocity offsets (downloads)
    int* d_qoff = nullptr;       // [nb+1] stacked position offsets
    double* d_stage = nullptr;   // downloads: the stacked output packed on the device
    size_t stage_cap = 0;
    double* h_stage = nullptr;   // ... and its pinned host copy
    size_t h_stage_cap = 0;
    double* d_uprev = nullptr;   // [nb*8] previous iterate (GS residual)
    double* d_ufree = nullptr;   // [nb*8]
    int* d_dof = nullptr;        // [nb]
    size_t body_cap = 0;

    // ---- contacts
    int nc = 0;
    std::vector<int> h_a, h_b;
    std::vector<double> h_mu;
    int nrows = 0;
    // bilateral joint rows (cdx_set_joints): blocks of rows per body pair, ordered along chains
    int njrows = 0, nchains = 0, njbodies = 0;
    double joint_bias = 0.0;
    bool joints_solved = false;
    std::vector<int> h_jorig;  // block-ordered row -> caller's row index
    std::vector<int> h_jc_off, h_jb_body;  // host copies of the chain and block tables
    // the row bodies of the last cdx_set_joints call whose tables are on the device: a call with the
    // same bodies (the same joints at a new step) only refreshes the row values
    std::vector<int> h_jsig_a, h_jsig_b;
    int jsig_ns = -1, jsig_nd = -1, jsig_nb = -1, jsig_nchains = 0, jsig_njb = 0;
    int *d_sc_off = nullptr, *d_sc_list = nullptr;  // batched solve: chains of each scene
    size_t sc_off_cap = 0, sc_list_cap = 0;
    int *d_jc_off = nullptr, *d_jcb_off = nullptr, *d_jb_off = nullptr, *d_jb_body = nullptr, *d_jr_kind = nullptr;
    int *d_jbody = nullptr, *d_jbody_blk = nullptr, *d_jslot = nullptr, *d_jr_blk = nullptr;
    int *d_jr_info = nullptr, *d_jbody_info = nullptr;  // chain-solve records per row / per jointed body
    size_t jr_info_cap = 0, jbody_info_cap = 0;
    int* d_jl_off = nullptr;  // [nblocks+1] packed offset of each block's L factor in d_jL
    size_t jl_off_cap = 0;
    double* d_jw = nullptr;  // [12*njrows] row coefficients (wa, wb) by coefficient: d_jw[k * njrows + r]
    size_t jw_cap = 0;
    int max_lds_block = 0;  // LDS bytes a workgroup may use (queried at the first batched joint solve)
    // Anderson loop: history rings Z and F ((m+1) slots each) and the candidate, over the mixed
    // vector [x; p]; least-squares coefficients; dot-product partials; ring state
    double *d_and = nullptr, *d_and_gam = nullptr, *d_and_part = nullptr;
    int* d_and_st = nullptr;
    size_t and_cap = 0;
    double *d_jr = nullptr, *d_jrlam = nullptr, *d_jz = nullptr, *d_jlam = nullptr, *d_jL = nullptr, *d_jDinv = nullptr;
    double* d_jtr = nullptr;  // [nb*6] right-hand side of each jointed body for the chain solve
    double* d_jy = nullptr;   // [njrows] chain-solve scratch (D^-1 z)
    size_t jy_cap = 0;
    size_t jc_off_cap = 0, jcb_off_cap = 0, jb_off_cap = 0, jb_body_cap = 0, jr_kind_cap = 0, jbody_cap = 0,
           jbody_blk_cap = 0, jslot_cap = 0, jr_cap = 0, jrlam_cap = 0, jz_cap = 0, jlam_cap = 0, jL_cap = 0, jDinv_cap = 0,
           jtr_cap = 0, jr_blk_cap = 0;
    bool any_rest = false;  // any non-zero restitution row (PgsAssembler::rhs, pgs_assembler.cpp:64-67)
    bool any_bias = false;  // any row with restitution or a velocity bias: b = R (c + e (W v + c)) != 0
    int* d_ida = nullptr;
    int* d_idb = nullptr;
    double* d_J = nullptr;     // [27*nc]
    double* d_mu = nullptr;    // [nc]
    double* d_rest = nullptr;  // [3*nc]
    double* d_vb = nullptr;    // [3*nc]
    double* d_p = nullptr;     // [3*nc] impulses (PJ: normal >= 0, GS: lambda <= 0 during the solve)
    double* d_R = nullptr;     // [3*nc]
    double* d_bias = nullptr;  // [3*nc]
    double* d_Rb = nullptr;    // [9*nc] inverse Delassus blocks, entry (k, j) at (3k + j) * nc + c (block mode)
    size_t Rb_cap = 0;
    double* d_imp = nullptr;   // [12*nc]
    int* d_incptr = nullptr;   // [nb+1]
    int* d_incidx = nullptr;   // [2*nc]
    std::vector<int> h_incptr, h_incidx;  // host copies of the incidence lists
    size_t contact_cap = 0;
    bool incidence_valid = false;

    // ---- colored GS topology (rebuilt when contacts change)
    bool coloring_valid = false;
    int ncolors = 0;
    std::vector<int> color_ptr;  // [ncolors+1] into the colour-sorted block list
    int* d_color_blocks = nullptr;  // [nblocks] contact index per colour-sorted block
    size_t color_cap = 0;
    // the sweep's per-contact data copied into colour order, so that a colour class is one
    // contiguous range: [27][n] J, [3][n] D^-1, [3][n] rhs, [3][n] lambda, [n] mu; ids [2][n]
    double* d_gs_pack = nullptr;
    int* d_gs_ids = nullptr;
    size_t gs_pack_cap = 0, gs_ids_cap = 0;

    // ---- batched scenes (cdx_set_scenes); cleared whenever bodies or contacts are set
    int nscenes = 0;
    std::vector<int> scene_boff, scene_coff, scene_nv, scene_rows, scene_order;
    int scene_max_bodies = 0, scene_max_contacts = 0;
    int* d_scene_i = nullptr;  // [order | boff | coff | nv | rows | iters]
    size_t scene_cap = 0;
    double* d_scene_res = nullptr;
    size_t scene_res_cap = 0;
    int batch_block = 256, batch_groups = 1;

    // ---- differentiation
    bool pj_last = false;  // the last solve was cdx_pj_solve (adjoint available)
    bool pjj_last = false;  // the last solve was a standard-loop joint solve (cdx_pj_adjoint_joints available)
    bool gyro_last = false;  // the last PJ solve ran with implicit gyroscopy (no adjoint)
    double last_theta = 1.0;
    bool record = false, record_valid = false;
    int record_max = 0, recorded = 0;
    double* d_rec_x = nullptr;  // [record_max][nb*8]
    double* d_rec_p = nullptr;  // [record_max][3*nc]
    size_t rec_x_cap = 0, rec_p_cap = 0;
    double* d_adj = nullptr;
    size_t adj_cap = 0;
    double* d_adj_imp = nullptr;
    size_t adj_imp_cap = 0;
    double* d_adj_body = nullptr;
    size_t adj_body_cap = 0;
    double* d_adj_jl = nullptr;  // joint adjoint: [lambda scratch | r_bar_Lambda | dL/dcomp] per joint row
    double* d_gm = nullptr;       // joint adjoint GMRES: basis, work vectors, Hessenberg, scalars
    size_t gm_cap = 0;
    double* d_gm_part = nullptr;  // GMRES dot-product partials
    size_t gm_part_cap = 0;
    size_t adj_jl_cap = 0;
    double* d_adj_partial = nullptr;
    size_t adj_partial_cap = 0;

    // ---- solve control
    int* d_ctrl = nullptr;         // [0] done flag, [1] iterations, [2] status
    double* d_res = nullptr;       // [0] last residual
    double* d_partial = nullptr;   // per-block residual partial sums
    size_t partial_cap = 0;
    double* d_partial2 = nullptr;  // Nesterov: residual and restart-dot partials
    size_t partial2_cap = 0;
    double* d_pk = nullptr;        // Nesterov: previous impulse iterate [3*nc]
    size_t pk_cap = 0;
    double* d_nest = nullptr;      // Nesterov scalars: beta, thk, err_prev, restarts, disabled
    double* d_xk1 = nullptr;       // Chebyshev: sweep input [nb*8] (power-iteration vector first)
    size_t xk1_cap = 0;
    double* d_pk1 = nullptr;       // Chebyshev: impulse sweep input [3*nc]
    size_t pk1_cap = 0;
    double* d_cheb = nullptr;      // Chebyshev scalars: w, take-swept, rho^2, rho, stopped, lambda, norm
    double* d_accj = nullptr;      // accelerated loops with joint rows: K, K1, Y (twists, Lambda, impulses)
    size_t accj_cap = 0;
    double* d_sinv = nullptr;      // implicit gyroscopy: inverse rotational blocks of S [nb*9]
    size_t sinv_cap = 0;
    // geometry form of the contact rows (cdx_set_contact_points): lever arms [6*nc], points [3*nc]
    double* d_geo = nullptr;
    double* d_geo_pt = nullptr;
    int* d_geo_bad = nullptr;
    size_t geo_cap = 0, geo_pt_cap = 0, geo_bad_cap = 0;
    bool geo_ok = false;    // every rebuilt angular row equals the current one bit for bit
    int geo_mode = 1;       // 0: never use the geometry form
    bool geo_used = false;  // the last standard-loop solve rebuilt the rows
    int geo_bad = -1;       // contacts whose rebuilt rows differ (-1: no points for these contacts)
    int* h_flag = nullptr;         // pinned mirror of the done flag
    int* h_flag_dev = nullptr;
    bool solved = false;

    // ---- replay whole chunks of sweeps as HIP graphs (off while profiling or recording)
    bool graphs = true;

    // ---- range sweep of projected Jacobi: one workgroup per contiguous contact range, per-body
    // partial sums reduced in LDS; tables rebuilt at the first standard-loop solve after the
    // contacts change
    int range_mode = 1;         // 0 off, 1 from kRangeAutoContacts contacts, 2 for every problem
    int range_per = 0;          // contacts per range (0: 128)
    int range_block = 512;      // threads per range workgroup
    bool ranges_valid = false;  // the tables match the current contacts and settings
    bool ranges_ok = false;     // the tables fit the LDS budget
    bool ranges_used = false;   // the last standard-loop solve swept ranges
    int nranges = 0, npart = 0, range_per_used = 0, range_maxp = 0, range_maxl = 0;
    int *d_rg_p = nullptr, *d_rg_l = nullptr, *d_pr_ptr = nullptr, *d_pr_out = nullptr, *d_pr_inc = nullptr;
    int* d_bp_ptr = nullptr;   // [nb] partial sums per body
    int* d_bp_ovf = nullptr;   // [nb] offset of a body's partials beyond its fixed slots
    double* d_part = nullptr;  // partial sums: fixed slots per body, then the overflow area

    // ---- profiling
    bool profiling = false;
    std::vector<cdx::PendingEvent> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, cdx::KernelStat> stats;
};

namespace cdx {

int set_err(cdx_ctx* ctx, int code, const std::string& msg);
int hip_check(cdx_ctx* ctx, hipError_t e, const char* what);
template <typename T>
int ensure_alloc(cdx_ctx* ctx, T** ptr, size_t count, size_t* cap_or_null, const char* what);
void prof_events(cdx_ctx* ctx, hipEvent_t* start, hipEvent_t* stop);
void prof_push(cdx_ctx* ctx, const char* name, hipEvent_t start, hipEvent_t stop);
int harvest_events(cdx_ctx* ctx);
int build_incidence(cdx_ctx* ctx);
// Host threads for a pass over n items (contacts, joint rows): one below 32k, at most 8.
inline int host_threads(int n) {
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    return std::max(1, std::min({8, hw, n / 32768}));
}

// fn(t, c0, c1) over T contiguous chunks of [0, n), chunk t in thread t (chunk 0 on the caller's).
template <class F>
void parallel_ranges(int n, int T, F fn) {
    auto lo = [&](int t) { return (int)((long long)n * t / T); };
    std::vector<std::thread> th;
    for (int t = 1; t < T; ++t) th.emplace_back(fn, t, lo(t), lo(t + 1));
    fn(0, 0, lo(1));
    for (auto& x : th) x.join();
}

// Grows the device (d_stage) and pinned host (h_stage) staging buffers to n doubles each; the
// host buffer is idle again once the stream has been synchronised after its last copy.
int ensure_stage(cdx_ctx* ctx, size_t n);
int grid_for(int n);
int reset_control(cdx_ctx* ctx);
int read_residual(cdx_ctx* ctx, double* r);
int check_solve_args(cdx_ctx* ctx, const cdx_solver_params* prm);
// Reduces ctx->d_partial[0:nparts) into the residual; mode 0 = PJ test (res <= 1),
// mode 1 = Gauss-Seidel test (res < 1, NaN/Inf -> CDX_DIVERGED).
int launch_residual_finalize(cdx_ctx* ctx, int nparts, int it, int mode, int cadence = 1);
// Message of the calling thread's last failed cdx_pj_solve_noalloc ("" after a success).
const char* noalloc_last_error();
int joints_factor(cdx_ctx* ctx, double dt, double theta);
int joints_solve(cdx_ctx* ctx, int setup, int check, double tol_rel, double tol_abs, double* partial);
int joints_apply(cdx_ctx* ctx, double* x_out, double* lam_out, const int* ctrl);

// Launch helper: brackets the launch with profiling events when enabled.
// With profiling on, the start/stop events travel in the kernel's own dispatch
// (hipExtLaunchKernelGGL), so the measured interval is the kernel's execution, not the gap
// between separate marker packets.
#define CDX_LAUNCH_SHM(ctx, name, kernel, grid, block, shm, ...)                                         \
    do {                                                                                                  \
        if ((ctx)->profiling) {                                                                           \
            hipEvent_t cdx_ev_start__, cdx_ev_stop__;                                                     \
            ::cdx::prof_events((ctx), &cdx_ev_start__, &cdx_ev_stop__);                                   \
            hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), (shm), (ctx)->stream, cdx_ev_start__, \
                                  cdx_ev_stop__, 0, __VA_ARGS__);                                         \
            ::cdx::prof_push((ctx), (name), cdx_ev_start__, cdx_ev_stop__);                               \
        } else {                                                                                          \
            hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), (shm), (ctx)->stream, __VA_ARGS__);       \
        }                                                                                                 \
    } while (0)
#define CDX_LAUNCH(ctx, name, kernel, grid, block, ...) CDX_LAUNCH_SHM(ctx, name, kernel, grid, block, 0, __VA_ARGS__)

#define CDX_TRY(expr)               \
    do {                            \
        int cdx_st__ = (expr);      \
        if (cdx_st__) return cdx_st__; \
    } while (0)

#define CDX_HIP(ctx, expr) CDX_TRY(::cdx::hip_check((ctx), (expr), #expr))

template <typename T>
int ensure_alloc(cdx_ctx* ctx, T** ptr, size_t count, size_t* cap, const char* what) {
    if (count == 0) count = 1;
    if (*ptr && cap && *cap >= count) return CDX_OK;
    if (*ptr && !cap) return CDX_OK;
    if (*ptr) {
        hipFree(*ptr);
        *ptr = nullptr;
    }
    hipError_t e = hipMalloc((void**)ptr, count * sizeof(T));
    if (e != hipSuccess) return set_err(ctx, CDX_HIP_ERROR, std::string("hipMalloc failed for ") + what + ": " + hipGetErrorString(e));
    if (cap) *cap = count;
    return CDX_OK;
}

// ------------------------------------------------------------------------------------------
// Device helpers
// ------------------------------------------------------------------------------------------

// Wave64 + workgroup sum of one double per thread.  Result valid in thread 0.  Fixed tree
// order, so the value is deterministic for a given launch shape.
__device__ inline double block_sum(double v, double* lds /* >= blockDim.x/64 */) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int w = 0; w < nw; ++w) s += lds[w];
    }
    __syncthreads();
    return s;
}

// Maps the hardware workgroup id to a logical tile so that each of the 8 XCDs (which receive
// workgroups round-robin) walks one contiguous range of the spatially sorted contact list,
// keeping neighbouring contacts, and the bodies they share, in one XCD's L2.
__device__ inline int xcd_tile(int nblocks) {
    const int hw = blockIdx.x;
    const int per = (nblocks + 7) >> 3;
    const int xcd = hw & 7, slot = hw >> 3;
    const int full = nblocks - (per - 1) * 8;  // number of XCDs that get `per` tiles
    int logical;
    if (xcd < full || per * 8 == nblocks) {
        logical = xcd * per + slot;
    } else {
        logical = full * per + (xcd - full) * (per - 1) + slot;
    }
    return logical;
}

}  // namespace cdx

/*
 * cdx.h -- C ABI of the MI355X (gfx950) Moreau contact solve.
 *
 * This is the device boundary that replaces CardilloCxx's solver backends
 * behind cardillo::solver::SolverBase (src/physics/solver/solver_base.hpp:14-36):
 *
 *   scene config -> contact list -> impulses -> velocity/position update
 *
 * All arithmetic is FP64 (the reference's real_t is double: src/misc/types.hpp:16).
 * Every entry point returns an int status (CDX_OK == 0); cdx_last_error() gives
 * the message of the most recent failure on a context.  The reference signals
 * failures with exceptions (pj_assembler.cpp:175,178-180; projected_gauss_seidel.cpp:137);
 * the C++ adapter in cardillocxx_amd/csrc/host/cdx_solvers.hpp turns non-zero
 * statuses back into std::runtime_error.
 *
 * Threading: a context owns one HIP stream (default priority) and is not
 * thread-safe; use one context per host thread.  Host pointers passed in are
 * only read for the duration of the call (device copies are owned by the context).
 *
 * Layouts (the same flat SoA layout the reference's DynamicsAssembler getters
 * expose, src/physics/assembly/dynamics_assembler.hpp:24-56):
 *   bodies  : nb bodies, dof[b] in {3 (point mass), 6 (rigid)}; stacked arrays of
 *             length nv = sum(dof) in body order (bodyVelOffsets):
 *             velocity [v_world(3); omega_body(3)], minv diag [m^-1 x3; I^-1 diag],
 *             fext [f_world; tau_body], fgyro [0; -omega x I omega].
 *             Positions q are stacked 3 (point mass) or 7 (x,y,z, qx,qy,qz,qw) per body.
 *   contacts: nc contacts in the reference's Contact[] order (src/collision/types.hpp:12-46):
 *             body_a/body_b body index or -1 for a static side; mu > 0 marks a
 *             frictional contact (3 rows n,t1,t2), otherwise 1 normal row
 *             (dynamics_assembler.cpp:236-324);
 *             frame[9*c]  = n, t1, t2 (world);
 *             angA[9*c]   = r_A,body x d_A,body for d in {n,t1,t2} (unsigned; the
 *                           side sign -1 for A, +1 for B is applied inside, as
 *                           buildContactRowByDof does, contact_jacobian.hpp:17-23);
 *             angB[9*c]   = same for side B;
 *             rest[3*c]   = restitution per row; vbias[3*c] = static-side velocity
 *                           term per row (DynamicsAssembler::contactVVec);
 *             warm[3*c]   = previous impulse (canonical: normal >= 0, warmstart.hpp:17-24).
 */
#ifndef CDX_CDX_H
#define CDX_CDX_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cdx_ctx cdx_ctx;

enum cdx_status {
    CDX_OK = 0,
    CDX_NOT_FACTORIZED = 1, /* S = diag(M) singular: zero mass on a dynamic dof (pj_assembler.cpp:98-113,175) */
    CDX_NONFINITE = 2,      /* non-finite input */
    CDX_DIVERGED = 3,       /* Gauss-Seidel NaN/Inf residual (projected_gauss_seidel.cpp:133-138) */
    CDX_BAD_ARG = 4,
    CDX_HIP_ERROR = 5,
    CDX_NO_STATE = 6 /* called before bodies / contacts / a solve were set */
};

enum cdx_accel {
    CDX_ACCEL_NONE = 0,     /* standard_loop (projected_jacobi.cpp:371-377) */
    CDX_ACCEL_NESTEROV = 1, /* nesterov_loop (projected_jacobi.cpp:379-449) */
    CDX_ACCEL_CHEBYSHEV = 2, /* chebyshev_loop (projected_jacobi.cpp:180-228), pj.chebyshev without pj.nesterov */
    CDX_ACCEL_ANDERSON = 3   /* anderson_loop (projected_jacobi.cpp:224-310), pj.anderson without the two above */
};

/* Solver parameters; the Config keys they stand for (src/config/config.hpp:61-113). */
typedef struct {
    double alpha;      /* pj.alpha (default 0.3) */
    double relaxation; /* pj.relaxation (default 1.0; Gauss-Seidel only) */
    int max_iter;      /* pj.max_iterations (default 200000) */
    double tol_abs;    /* pj.tol_abs (default 1e-5) */
    double tol_rel;    /* pj.tol_rel (default 1e-5) */
    int check_every;   /* PJ residual cadence (reference: every 10 sweeps, projected_jacobi.cpp:375) */
    int warmstart;     /* pj.warmstart (default 1) */
    int accel;         /* enum cdx_accel (pj.nesterov, pj.chebyshev) */
    double nesterov_beta_threshold; /* pj.nesterov_beta_threshold (default 1.0) */
    int nesterov_restart_limit;     /* pj.nesterov_restart_limit (default 10000) */
    int block_precond; /* pj.rdiag_true_delassus (default 0): per-contact inverse Delassus blocks alpha (W_c M^-1 W_c^T)^-1
                          instead of the row scaling alpha / D_ii (projected_jacobi.cpp:71-98, 330-348); projected
                          Jacobi only (ignored by Gauss-Seidel, refused with joint rows and by cdx_pj_adjoint) */
    int anderson_m;    /* pj.anderson_m (default 2): Anderson history window, 1..8 */
    int implicit_gyro; /* moreau.implicit_gyroscopy (default 0, config.hpp:233): projected Jacobi solves with
                          the rotational block diag(I) - dt G_rot(w_n) of S per rigid body, G_rot =
                          1/2 ([I w]_x - [w]_x diag(I)), and without dt f_gyro on the right-hand side
                          (pj_assembler.cpp:32-61,150-153); Gauss-Seidel only drops dt f_gyro from u_free
                          (pgs_assembler.cpp:33-36,75-82).  Refused with joint rows, by the batched solve
                          and by cdx_pj_solve_noalloc; cdx_pj_adjoint refuses to follow such a solve. */
} cdx_solver_params;

typedef struct {
    int iters;       /* sweeps performed (SolverBase::lastIterations, solver_base.hpp:29) */
    double residual; /* last residual evaluated (NaN if none) */
    int nrows;       /* contact rows (frictionless + 3 x frictional) */
    int ncolors;     /* colored Gauss-Seidel: colour classes used */
} cdx_solve_info;

/* Fills *p with the reference Config defaults (config.hpp:61-113). */
void cdx_default_params(cdx_solver_params* p);

/* Context lifetime.  device: HIP ordinal. */
int cdx_ctx_create(int device, cdx_ctx** out);
void cdx_ctx_destroy(cdx_ctx* ctx);
/* Message of the most recent failure on ctx; with ctx = NULL, of the calling thread's most recent
 * context-free call (cdx_pj_solve_noalloc). */
const char* cdx_last_error(const cdx_ctx* ctx);
int cdx_synchronize(cdx_ctx* ctx);
/* The context's HIP stream (a hipStream_t) in *stream: every kernel of the context runs on it, so
 * a caller can order its own work, or record timing events, against the solve. */
int cdx_get_stream(cdx_ctx* ctx, void** stream);

/* Body state: stacks what DynamicsAssembler::refreshState builds
 * (dynamics_assembler.cpp:41-120, 469-513).  fgyro may be NULL, in which case it is
 * computed on the device as -omega x (I omega) (world.cpp:152-173).  q may be NULL
 * when no position update is needed. */
int cdx_set_bodies(cdx_ctx* ctx, int nb, const int* dof, const double* minv, const double* v, const double* fext,
                   const double* fgyro, const double* q);

/* Contact buffer: the input of DynamicsAssembler::rebuildW_ (dynamics_assembler.cpp:211-332),
 * filled by CPU contact generation and kept device-resident. */
int cdx_set_contacts(cdx_ctx* ctx, int nc, const int* body_a, const int* body_b, const double* mu, const double* frame,
                     const double* angA, const double* angB, const double* rest, const double* vbias,
                     const double* warm);

/* Bilateral rows of the extended system: the output of DynamicsAssembler::rebuildInteractionW_
 * (dynamics_assembler.cpp:341-439) that PjAssembler puts into S = [M, G^T; G, -E]
 * (pj_assembler.cpp:9-182).  ns spring rows (Wg, compliance C) then nd damper rows (Wgamma,
 * compliance A).  Row r couples body_a[r] and body_b[r] (-1: static; its velocity is folded into
 * vsrc) with wa[6r..6r+5] / wb[6r..6r+5] on the body's [v; omega]; comp[r] = C or A (finite,
 * >= 0); vsrc[r] = C_v / A_v; gerr[r] = position error of a spring row (Baumgarte bias, with
 * bias_factor = constraint_bias_factor, pj_assembler.cpp:163-164); lam_prev[r] = Lambda_g /
 * Lambda_gamma of the previous step.  Rows of one body pair form a joint block (<= 12 rows);
 * blocks must form chains (each dynamic body in at most two blocks, no loops) -- capsule chains,
 * beams -- and jointed bodies must be 6-dof.  Taken by cdx_pj_solve (standard loop); cleared by
 * cdx_set_bodies. */
int cdx_set_joints(cdx_ctx* ctx, int ns, int nd, const int* body_a, const int* body_b, const double* wa, const double* wb,
                   const double* comp, const double* vsrc, const double* gerr, const double* lam_prev, double bias_factor);

/* Lambda_g / Lambda_gamma of the last cdx_pj_solve with joint rows, in the caller's row order
 * (setLambda_g / setLambda_gamma, projected_jacobi.cpp:495-497). */
int cdx_get_joint_impulses(cdx_ctx* ctx, double* lam_out);

/* Projected Jacobi: replaces ProjectedJacobiSolver::solve (projected_jacobi.cpp:451-500).
 * Results stay on the device: fetch with cdx_get_velocities / cdx_get_impulses.
 * The new impulses also become the next warm start (storeImpulse, warmstart.hpp:41-56).
 * With joint rows (cdx_set_joints) x = [v; Lambda] solves the extended system, in the standard,
 * Nesterov and Chebyshev loops (extrapolating the whole of x, testing the residual on v, as the
 * reference does); Anderson and the block preconditioner are refused with joint rows. */
int cdx_pj_solve(cdx_ctx* ctx, double dt, double theta, const cdx_solver_params* prm, cdx_solve_info* info);

/* Graph-coloured projected Gauss-Seidel: the PGS block update
 * (projected_gauss_seidel.cpp:144-169) run colour by colour as in
 * CondensedSolver's coloredSweep (condensed_solver.cpp:220-251), Welsh-Powell colouring
 * (graph_coloring.cpp:8-41).  Residual every sweep, CDX_DIVERGED on NaN/Inf.  prm->accel:
 * CDX_ACCEL_NONE, or CDX_ACCEL_NESTEROV for the reference's Nesterov outer loop over (lambda,
 * u_corr) with its restart rules (pj.nesterov, projected_gauss_seidel.cpp:196-262); the other
 * modes are refused (CDX_BAD_ARG): the reference's PGS reads pj.nesterov only. */
int cdx_gs_colored_solve(cdx_ctx* ctx, double dt, double theta, const cdx_solver_params* prm, cdx_solve_info* info);

/* Outputs of the last solve.  v_out: nv stacked velocities (the SolverBase::solve return
 * value); imp_out: 3*nc impulses per contact, canonical normal >= 0, zero tangents for
 * frictionless contacts (Contact::last_impulse). */
int cdx_get_velocities(cdx_ctx* ctx, double* v_out);
int cdx_get_impulses(cdx_ctx* ctx, double* imp_out);
int cdx_get_positions(cdx_ctx* ctx, double* q_out);

/* Number of bodies whose solved twist (cdx_get_velocities) or pose (cdx_get_positions, when
 * positions were set) holds a NaN or Inf entry, counted on the device so that a caller can test a
 * step without downloading the state.  No counterpart in the reference, whose PGS reports a
 * non-finite residual (projected_gauss_seidel.cpp:133-138) and whose projected Jacobi returns the
 * state as it is; for batched callers that put a scene that blew up back at rest. */
int cdx_nonfinite_bodies(cdx_ctx* ctx, int* count);

/* Moreau theta-method around the solve (moreau.cpp:6-13, integration_base.cpp:6-30):
 *   cdx_moreau_predrift : q += (1-theta) h B(q) v_n
 *   cdx_moreau_update   : v_n <- solved velocities (writeVelocityToSystem,
 *                         dynamics_assembler.cpp:145-188), q += theta h B(q) v_{n+1},
 *                         quaternion normalised and sign-aligned (math_helper.hpp:11-17),
 *                         f_gyro recomputed from the new velocities (world.cpp:152-173). */
int cdx_moreau_predrift(cdx_ctx* ctx, double dt, double theta);
int cdx_moreau_update(cdx_ctx* ctx, double dt, double theta);

/* ---- Batched scenes (no counterpart in the reference, which runs one PhysicsEngine per scene,
 * physics_engine.hpp:215-221).  The context's bodies and contacts are partitioned into nscenes
 * independent scenes: scene s owns bodies [body_off[s], body_off[s+1]) and contacts
 * [contact_off[s], contact_off[s+1]); contacts may only reference bodies of their own scene.
 * Must be called after cdx_set_bodies / cdx_set_contacts (which clear the partition). */
int cdx_set_scenes(cdx_ctx* ctx, int nscenes, const int* body_off, const int* contact_off);

/* Solves every scene with the projected-Jacobi standard or Chebyshev loop (prm->accel NONE or
 * CHEBYSHEV), row scaling or block preconditioner (prm->block_precond, not with joint rows), in
 * one persistent launch; each scene has its own residual test, spectral radius and iteration count
 * (iters[s], residuals[s]; either may be NULL).  Results are read with cdx_get_velocities /
 * cdx_get_impulses like a single solve. */
int cdx_batch_pj_solve(cdx_ctx* ctx, double dt, double theta, const cdx_solver_params* prm, int* iters,
                       double* residuals);

/* Launch shape of the batched solve: `block` threads per workgroup (64..512) hosting `groups`
 * scenes (block/groups must be a multiple of 64).  With groups = 1, no joint rows and every
 * scene's velocities, side-impulse records and incidence lists within 60 KiB, each scene is kept
 * in LDS through its sweeps, and the workgroup gets at least one thread per contact and per body
 * of the largest scene (up to 512). */
int cdx_batch_set_launch(cdx_ctx* ctx, int block, int groups);

/* Autotuner: times every launch shape on the current batch (warm start restored before each
 * trial), keeps the fastest for later cdx_batch_pj_solve calls and reports it. */
int cdx_batch_autotune(cdx_ctx* ctx, double dt, double theta, const cdx_solver_params* prm, int* block,
                       int* groups, double* best_ms);

/* ---- Differentiation of the projected-Jacobi solve (no counterpart in the reference).
 * The solve maps (v_n, M^-1, f_ext, mu) to (v_{n+1} = x*, p*) through the fixed point
 * p* = Pi_mu(p* - R (W x* + b)), x* = u_free + M^-1 W^T p* (projected_jacobi.cpp:121-138).
 * The backward of the last cdx_pj_solve on a context is computed with the contact geometry
 * (W, c, e) held fixed:
 *   CDX_ADJ_IMPLICIT : implicit differentiation of the fixed point -- one linear adjoint solve
 *                      lambda = g_p + (I - A R) P^T lambda (A = W M^-1 W^T, P the generalized
 *                      Jacobian of the Coulomb projection per active-set case), iterated with the
 *                      same contact/body passes as the forward sweep;
 *   CDX_ADJ_UNROLLED : reverse mode through the recorded forward sweeps (cdx_set_recording must
 *                      have been on during the solve);
 *   CDX_ADJ_AUTO     : the implicit system solved by restarted GMRES(30) from lambda_0 = g_p instead
 *                      of the fixed-point iteration: it converges where that iteration drifts (a
 *                      singular or near-singular system: impulses that are not unique, as on a box
 *                      face resting on four coplanar contacts), and where the system is
 *                      inconsistent (contacts on a stick/slip or contact/separation boundary) it
 *                      returns the least-squares adjoint, with its relative residual in
 *                      info->residual (GMRES stops after three restarts without a 1 % gain). */
enum cdx_adjoint_mode { CDX_ADJ_IMPLICIT = 0, CDX_ADJ_UNROLLED = 1, CDX_ADJ_AUTO = 2 };

typedef struct {
    int mode;     /* enum cdx_adjoint_mode */
    int max_iter; /* implicit: iteration cap */
    double tol;   /* implicit: stop when ||lambda_k+1 - lambda_k|| <= tol ||lambda_k+1|| (checked every 10) */
} cdx_adjoint_params;

typedef struct {
    int iters;       /* implicit iterations, GMRES applications (CDX_ADJ_AUTO), or unrolled sweeps */
    double residual; /* last relative change of the implicit iterate, or GMRES's relative residual */
    int mode_used;
} cdx_adjoint_info;

/* Record every sweep's (x, p) of subsequent cdx_pj_solve calls (standard loop) for the unrolled
 * backward; `max_sweeps` bounds the memory (sweeps beyond it are not recorded and make the
 * unrolled backward fail with CDX_BAD_ARG). */
int cdx_set_recording(cdx_ctx* ctx, int enable, int max_sweeps);

/* Vector-Jacobian product of the last cdx_pj_solve or cdx_batch_pj_solve (the batched solve
 * leaves the same converged state; it records nothing, so it supports CDX_ADJ_IMPLICIT only --
 * scenes are independent, so one implicit iteration over the whole batch is every scene's
 * adjoint at once).  Inputs: g_v = dL/dv_{n+1} (nv, required),
 * g_p = dL/dp (3*nc canonical impulses, or NULL).  Outputs (each may be NULL):
 * gv_out = dL/dv_n (nv), gminv_out = dL/dM^-1 (nv, f_ext held fixed), gfext_out = dL/df_ext (nv),
 * gmu_out = dL/dmu per contact (nc). */
int cdx_pj_adjoint(cdx_ctx* ctx, double dt, const cdx_solver_params* prm, const cdx_adjoint_params* aprm,
                   const double* g_v, const double* g_p, double* gv_out, double* gminv_out, double* gfext_out,
                   double* gmu_out, cdx_adjoint_info* info);

/* Vector-Jacobian product of the last cdx_pj_solve or cdx_batch_pj_solve with joint rows
 * (implicit mode; CDX_ADJ_AUTO has no unrolled fallback here).  The same adjoint iteration as
 * cdx_pj_adjoint with S^-1 (the extended system of PjAssembler, pj_assembler.cpp:9-124) in place
 * of M^-1, contact geometry and joint Jacobians held fixed; dt must be the solve's step (theta is
 * taken from the solve).  Inputs: g_v = dL/dv_{n+1} (nv, required), g_p = dL/dp (3*nc or NULL).
 * Outputs (each may be NULL): gv_out = dL/dv_n (nv), gfext_out = dL/df_ext (nv), gmu_out = dL/dmu
 * per contact (nc), gcomp_out = dL/dcomp per joint row in the order given to cdx_set_joints (the
 * compliance C of a spring row, A of a damper row; a damper of damping D has A = 1/D, so
 * dL/dD = -dL/dA / D^2).  Mass gradients are not offered with joint rows.  The adjoint system
 * (I - (I - A R) P^T) lambda = g_p is solved by restarted GMRES to a relative residual of
 * aprm->tol within aprm->max_iter applications (each one transposed sweep with the chain solve).
 * Where contacts sit on a stick/slip or contact/separation boundary the system can be
 * inconsistent and GMRES stalls: CDX_ADJ_IMPLICIT then fails with CDX_DIVERGED, CDX_ADJ_AUTO
 * returns the least-squares adjoint (info->residual gives its relative residual).  No
 * counterpart in the reference, which has no differentiable solve. */
int cdx_pj_adjoint_joints(cdx_ctx* ctx, double dt, const cdx_solver_params* prm, const cdx_adjoint_params* aprm,
                          const double* g_v, const double* g_p, double* gv_out, double* gfext_out, double* gmu_out,
                          double* gcomp_out, cdx_adjoint_info* info);

/* Kernel timing with HIP events on the context's stream (the analogue of the reference's
 * TimingManager scopes, src/misc/timings/TimingManager.hpp:16-58).  When enabled every
 * launch is bracketed by events; cdx_kernel_time reads back the call count and total ms for
 * one kernel name ("pj_contacts", "pj_bodies", "gs_color", ...) and resets it. */
int cdx_set_profiling(cdx_ctx* ctx, int enable);
int cdx_kernel_time(cdx_ctx* ctx, const char* kernel, int* calls, double* total_ms);
/* The index-th kernel with recorded time (in name order), without resetting it: its name (copied
 * into name_buf, at most buf_len bytes with the terminating NUL), call count and total ms.
 * CDX_BAD_ARG past the last one.  For a TimingManager-style breakdown over all kernels. */
int cdx_kernel_stat(cdx_ctx* ctx, int index, char* name_buf, int buf_len, int* calls, double* total_ms);

/* Replay of sweep chunks as HIP graphs (default on).  With enable = 0 every sweep's kernels are
 * launched one by one (as with profiling on): the same kernels and results, more host launches.
 * For profilers and tools that intercept kernel dispatches.  No counterpart in the reference. */
int cdx_set_graphs(cdx_ctx* ctx, int enable);

/* Range sweep of the standard loop (cdx_pj_solve without acceleration, joint rows or sweep
 * recording; row scaling or block preconditioner).  One workgroup takes a contiguous range of
 * contacts: it runs their contact pass with the side impulses in LDS and writes, for each body its
 * range touches, the sum of those side impulses as one partial; a body pass adds each body's
 * partials and updates the velocities.  The same sweep as the two-pass form (pj_sweep,
 * projected_jacobi.cpp:121-138), with per-body sums in another fixed order: results equal to
 * rounding, bit-reproducible run to run.  It moves the least data when contacts come in the order
 * of their lower body and body indices follow space; any order is correct.
 * mode: 0 = off (two passes), 1 = on from 32,768 contacts (default), 2 = on for every problem.
 * contacts_per_range: 0 (= 128) or 16..768.  block: threads per range workgroup, 0 (= 512), 256
 * or 512.  When a range's tables exceed the workgroup's LDS the solve uses the two-pass form.
 * No counterpart in the reference. */
int cdx_set_range_sweep(cdx_ctx* ctx, int mode, int contacts_per_range, int block);
/* Whether the last cdx_pj_solve swept ranges, and the range tables of the current contacts:
 * ranges, partial sums per sweep (one per body and range touching it), contacts per range (0
 * until a range solve has built them). */
int cdx_range_sweep_info(cdx_ctx* ctx, int* used, int* nranges, int* npartials, int* contacts_per_range);

/* World contact points of the current contacts (3 doubles per contact, the reference Contact's
 * point), for the geometry form of the rows: a rigid side's angular row k is r_body x (R^T d_k),
 * r_body = R^T (p - x_com) (makeContact + buildContactRowByDof, contact_jacobian.hpp:17-23;
 * rebuildW_, dynamics_assembler.cpp:211-332).  The library derives every rigid side's r_body from
 * the point and the body's current position and quaternion, rebuilds the angular rows from them
 * with the contact generator's arithmetic (R from the quaternion entry by entry, then R^T d, then
 * the cross product, no contraction) and compares them with the rows passed to cdx_set_contacts.
 * It also checks that each frame's second tangent is n x t1.  When every one matches bit for bit,
 * the range sweep rebuilds the rows in each sweep from the lever arms, the quaternions and the
 * normal and first tangent instead of reading them (96 instead of 216 bytes per contact); the
 * results are identical.  Otherwise the sweep reads the rows as always.  Call after
 * cdx_set_contacts; cdx_set_bodies, cdx_set_contacts and the Moreau position updates discard the
 * points.  No counterpart in the reference (which builds W from these points each step). */
int cdx_set_contact_points(cdx_ctx* ctx, const double* point);
/* mode 1 (default): the range sweep uses the geometry form when the points match; 0: never.
 * No counterpart in the reference. */
int cdx_set_geometry_rows(cdx_ctx* ctx, int mode);
/* usable: the current points reproduce every angular row; mismatches: contacts whose rebuilt rows
 * differ in any bit (-1 when no points were given for the current state); used: the last
 * cdx_pj_solve swept with the geometry form. */
int cdx_geometry_rows_info(cdx_ctx* ctx, int* usable, int* mismatches, int* used);

/* ---- Allocation-free projected Jacobi.  No counterpart in the reference: this entry point is for
 * callers that cannot allocate device memory (tight serving loops, harnesses that forbid hipMalloc,
 * code that must not grow the device heap).  It runs the standard loop of cdx_pj_solve on arrays
 * the caller already holds on the device, with every buffer the solve needs beyond its inputs and
 * outputs in one statically sized device array of the library (about 32 MiB per GPU, held from
 * the library's first call on that GPU until the process ends): no hipMalloc, hipFree or other
 * allocating call.  Limits: at most CDX_NOALLOC_MAX_BODIES bodies and CDX_NOALLOC_MAX_CONTACTS
 * contacts (larger problems fail with CDX_BAD_ARG), standard loop with the row scaling only (no
 * acceleration, block preconditioner or joint rows: CDX_BAD_ARG).  All calls on one GPU share the
 * array, so calls that overlap on one GPU (two host threads, or two streams) race; it keeps the
 * last call's contents, and each call clears what must start at zero.  The call synchronises
 * `stream` to read the convergence flag every 4 x check_every sweeps, so it cannot itself be
 * captured into a graph. */
#define CDX_NOALLOC_MAX_BODIES 16384
#define CDX_NOALLOC_MAX_CONTACTS 65536

/* A contact problem on the device, in the layouts of cdx_set_bodies / cdx_set_contacts (device
 * pointers, read only).  fgyro may be NULL (computed as -omega x (I omega)); warm may be NULL
 * (no warm start). */
typedef struct {
    int nb, nc;
    const int* dof;
    const double* minv;
    const double* v;
    const double* fext;
    const double* fgyro;
    const int* body_a;
    const int* body_b;
    const double* mu;
    const double* frame;
    const double* angA;
    const double* angB;
    const double* rest;
    const double* vbias;
    const double* warm;
} cdx_device_problem;

/* cdx_pj_solve + cdx_get_velocities + cdx_get_impulses of `pb` in one call, on `stream` (a
 * hipStream_t; NULL = the default stream) of the current device: v_out (nv = sum(dof) doubles) and
 * imp_out (3 * nc doubles, canonical normal >= 0) are device pointers the caller owns.  The same
 * statuses as the context path; the message of a failure is returned by cdx_last_error(NULL) on
 * the calling thread. */
int cdx_pj_solve_noalloc(const cdx_device_problem* pb, double dt, double theta, const cdx_solver_params* prm,
                         double* v_out, double* imp_out, cdx_solve_info* info, void* stream);

#ifdef __cplusplus
}
#endif
#endif

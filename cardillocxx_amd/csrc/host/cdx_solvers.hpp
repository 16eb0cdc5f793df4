// cdx_solvers.hpp -- C++ host adapter over the C ABI, shaped like the reference's solver
// interface so that it drops in behind MoreauIntegrator::step (src/physics/integration/moreau.cpp:11).
//
//   cdx::SolverBase                 <- cardillo::solver::SolverBase (solver_base.hpp:14-36)
//   cdx::GpuProjectedJacobiSolver   <- ProjectedJacobiSolver (projected_jacobi.hpp:120-163)
//   cdx::GpuColoredGaussSeidelSolver<- ProjectedGaussSeidelSolver (projected_gauss_seidel.hpp)
//   cdx::ContactProblem             <- the DynamicsAssembler getters a solver reads
//                                      (dynamics_assembler.hpp:24-103): stacked bodies + contacts
//
// Non-zero statuses become std::runtime_error, as the reference throws (pj_assembler.cpp:175,178-180;
// projected_gauss_seidel.cpp:137).  Impulses are written back into ContactProblem::warm, the
// equivalent of storeImpulse into Contact::last_impulse (warmstart.hpp:41-56).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "cdx/cdx.h"
#include "cdx_timing.hpp"

namespace cdx {

struct Config {  // hot-path keys of cardillo::config::Config (config.hpp:46-113,232)
    bool friction_enable = false;
    int pj_max_iterations = 200000;
    double pj_tol_abs = 1e-5;
    double pj_tol_rel = 1e-5;
    double pj_relaxation = 1.0;
    double pj_alpha = 0.3;
    bool pj_nesterov = false;
    double pj_nesterov_beta_threshold = 1.0;
    int pj_nesterov_restart_limit = 10000;
    bool pj_chebyshev = false;
    bool pj_anderson = false;
    int pj_anderson_m = 2;
    bool pj_rdiag_true_delassus = false;
    bool pj_warmstart = true;
    double moreau_theta = 1.0;
    bool moreau_implicit_gyroscopy = false;  // config.hpp:233

    cdx_solver_params params() const {
        cdx_solver_params p;
        cdx_default_params(&p);
        p.alpha = pj_alpha;
        p.relaxation = pj_relaxation;
        p.max_iter = pj_max_iterations < 1 ? 1 : pj_max_iterations;
        p.tol_abs = pj_tol_abs;
        p.tol_rel = pj_tol_rel;
        p.warmstart = pj_warmstart ? 1 : 0;
        // precedence pj_nesterov > pj_chebyshev > pj_anderson (projected_jacobi.cpp:469-472)
        p.accel = pj_nesterov ? CDX_ACCEL_NESTEROV
                              : (pj_chebyshev ? CDX_ACCEL_CHEBYSHEV : (pj_anderson ? CDX_ACCEL_ANDERSON : CDX_ACCEL_NONE));
        p.anderson_m = pj_anderson_m < 1 ? 1 : pj_anderson_m;
        p.block_precond = pj_rdiag_true_delassus ? 1 : 0;
        p.nesterov_beta_threshold = pj_nesterov_beta_threshold;
        p.nesterov_restart_limit = pj_nesterov_restart_limit;
        p.implicit_gyro = moreau_implicit_gyroscopy ? 1 : 0;
        return p;
    }
};

struct ContactProblem {
    // bodies (stacked, nv = sum(dof))
    std::vector<int> dof;
    std::vector<double> minv, v, fext, fgyro, q;
    // contacts (Contact[] order)
    std::vector<int> body_a, body_b;
    std::vector<double> mu, frame, angA, angB, rest, vbias, warm;

    int numBodies() const { return (int)dof.size(); }
    int numContacts() const { return (int)body_a.size(); }
    int numV() const {
        int n = 0;
        for (int d : dof) n += d;
        return n;
    }
};

class SolverBase {
   public:
    virtual ~SolverBase() = default;
    virtual std::vector<double> solve(double dt, double theta) = 0;
    virtual int lastIterations() const { return m_last_iters; }
    virtual double lastError() const { return 0.0; }
    virtual const char* name() const { return "SolverBase"; }

   protected:
    int m_last_iters{-1};
};

// One device context per solver (one HIP stream); not thread-safe, like the reference solvers.
class GpuSolver : public SolverBase {
   public:
    GpuSolver(ContactProblem& problem, const Config& cfg, int device = 0) : m_problem(problem), m_cfg(cfg) {
        check(cdx_ctx_create(device, &m_ctx), "cdx_ctx_create");
    }
    ~GpuSolver() override { cdx_ctx_destroy(m_ctx); }
    GpuSolver(const GpuSolver&) = delete;
    GpuSolver& operator=(const GpuSolver&) = delete;

    double lastError() const override { return m_info.residual; }
    const cdx_solve_info& lastInfo() const { return m_info; }
    cdx_ctx* context() { return m_ctx; }

    // Section timings of every solve (the reference's DynamicsAssembler::timings()): the solve's
    // wall time under name(), the upload under RebuildContactJacobians, the device kernels under
    // their sections (kernel events on: sweeps are then launched one by one, not graph-replayed).
    void setTimings(TimingManager* t) {
        m_timings = t;
        check(cdx_set_profiling(m_ctx, t ? 1 : 0), "cdx_set_profiling");
    }

   protected:
    void check(int st, const char* what) const {
        if (st != CDX_OK)
            throw std::runtime_error(std::string(what) + " failed (" + std::to_string(st) + "): " +
                                     (m_ctx ? cdx_last_error(m_ctx) : "no context"));
    }

    void upload() {
        TimingManager::Scope sc = timed("RebuildContactJacobians");
        ContactProblem& p = m_problem;
        const int nb = p.numBodies(), nc = p.numContacts();
        check(cdx_set_bodies(m_ctx, nb, p.dof.data(), p.minv.data(), p.v.data(), p.fext.data(),
                             p.fgyro.empty() ? nullptr : p.fgyro.data(), p.q.empty() ? nullptr : p.q.data()),
              "cdx_set_bodies");
        std::vector<double> mu = p.mu;
        if (!m_cfg.friction_enable) mu.assign(mu.size(), 0.0);  // friction.enable=false: 1 row per contact
        check(cdx_set_contacts(m_ctx, nc, p.body_a.data(), p.body_b.data(), mu.data(), p.frame.data(), p.angA.data(),
                               p.angB.data(), p.rest.data(), p.vbias.data(), p.warm.data()),
              "cdx_set_contacts");
    }

    // A scope on the timings when they are enabled (a throwaway manager otherwise).
    TimingManager::Scope timed(const char* section) { return (m_timings ? *m_timings : m_idle).scope(section); }

    std::vector<double> finish() {
        if (m_timings) m_timings->harvestDevice(m_ctx);
        m_last_iters = m_info.iters;
        std::vector<double> v(m_problem.numV());
        check(cdx_get_velocities(m_ctx, v.data()), "cdx_get_velocities");
        m_problem.warm.resize(3 * (size_t)m_problem.numContacts());
        check(cdx_get_impulses(m_ctx, m_problem.warm.data()), "cdx_get_impulses");
        return v;
    }

    ContactProblem& m_problem;
    Config m_cfg;
    cdx_ctx* m_ctx = nullptr;
    cdx_solve_info m_info{};
    TimingManager* m_timings = nullptr;
    TimingManager m_idle;
};

class GpuProjectedJacobiSolver : public GpuSolver {
   public:
    using GpuSolver::GpuSolver;
    std::vector<double> solve(double dt, double theta) override {
        TimingManager::Scope sc = timed(name());
        upload();
        const cdx_solver_params prm = m_cfg.params();
        check(cdx_pj_solve(m_ctx, dt, theta, &prm, &m_info), "cdx_pj_solve");
        return finish();
    }
    const char* name() const override { return "ProjectedJacobi"; }
};

class GpuColoredGaussSeidelSolver : public GpuSolver {
   public:
    using GpuSolver::GpuSolver;
    std::vector<double> solve(double dt, double theta) override {
        TimingManager::Scope sc = timed(name());
        upload();
        cdx_solver_params prm = m_cfg.params();
        // ProjectedGaussSeidelSolver reads pj.nesterov only (projected_gauss_seidel.cpp:196)
        prm.accel = m_cfg.pj_nesterov ? CDX_ACCEL_NESTEROV : CDX_ACCEL_NONE;
        const int st = cdx_gs_colored_solve(m_ctx, dt, theta, &prm, &m_info);
        if (st == CDX_DIVERGED) throw std::runtime_error("Projected Gauss-Seidel diverged");
        check(st, "cdx_gs_colored_solve");
        return finish();
    }
    const char* name() const override { return "ColoredGaussSeidel"; }
};

}  // namespace cdx

// cdx_adapter_demo.cpp -- C++ end-to-end driver of the adapter: the 1000-sphere drop
// (8x8 jittered grid, pitch 0.12 m, r=0.05 m, m=1 kg, mu=0.5, 1x1x2 m box), Moreau steps with
// CPU contact generation, warm-start matching, the device PJ solve behind SolverBase, and the
// device position update.  Prints a per-step line, a final fingerprint line, the section breakdown
// of TimingManager::printBreakdown (TimingManager.hpp:107-163; host sections plus the device
// kernels booked under theirs), and, when CARDILLO_DUMP_STATE is set, the reference's
// "[STATE-DUMP] t= numBodies= totalKE= posNormSum=" line on stderr (examples/example_main.hpp:56-69).
// With a VTK directory it writes the body and contact frames every `vtk_every` steps (VtkWriter,
// src/io/vtk_writer.hpp:39-112) under the section OutputWrite.
//
// usage: cdx_adapter_demo [steps=200] [sweeps=100] [solver=pj|gs] [vtk_dir] [vtk_every=10]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <vector>

#include "cdx_scene.h"
#include "cdx_solvers.hpp"
#include "cdx_vtk.hpp"

int main(int argc, char** argv) {
    const int steps = argc > 1 ? std::atoi(argv[1]) : 200;
    const int sweeps = argc > 2 ? std::atoi(argv[2]) : 100;
    const bool gs = argc > 3 && std::strcmp(argv[3], "gs") == 0;
    const std::string vtk_dir = argc > 4 ? argv[4] : "";
    const int vtk_every = argc > 5 ? std::atoi(argv[5]) : 10;
    if (!vtk_dir.empty()) std::filesystem::create_directories(vtk_dir);
    cdx::VtkWriter vtk(vtk_dir, "spheres", vtk_every);
    const int n = 1000;
    const double r = 0.05, m = 1.0, mu = 0.5, h = 1e-3;
    std::vector<double> q(7 * (size_t)n);
    cdx_scene_sphere_grid(n, 8, 8, 0.12, 0.06, 0.01, 0, q.data());
    std::vector<double> radius(n, r), mub(n, mu);
    const double wn[15] = {0, 0, 1, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1, 0};
    const double wd[5] = {0.0, -0.5, -0.5, -0.5, -0.5};
    const double wmu[5] = {mu, mu, mu, mu, mu};

    cdx::ContactProblem P;
    P.dof.assign(n, 6);
    const double I = 0.4 * m * r * r;
    for (int i = 0; i < n; ++i) {
        const double mi[6] = {1 / m, 1 / m, 1 / m, 1 / I, 1 / I, 1 / I};
        const double fe[6] = {0, 0, -9.81 * m, 0, 0, 0};
        P.minv.insert(P.minv.end(), mi, mi + 6);
        P.fext.insert(P.fext.end(), fe, fe + 6);
    }
    P.v.assign(6 * (size_t)n, 0.0);
    P.fgyro.assign(6 * (size_t)n, 0.0);
    P.q = q;

    cdx::Config cfg;
    cfg.friction_enable = true;
    cfg.pj_max_iterations = sweeps;
    cfg.pj_tol_abs = cfg.pj_tol_rel = 0.0;  // exactly `sweeps` sweeps per step
    cdx::GpuProjectedJacobiSolver pj(P, cfg);
    cdx::GpuColoredGaussSeidelSolver gsv(P, cfg);
    cdx::GpuSolver& solver = gs ? (cdx::GpuSolver&)gsv : (cdx::GpuSolver&)pj;
    cdx::TimingManager timings;
    solver.setTimings(&timings);
    double t = 0.0;
    cdx::TimingManager::Scope total = timings.scope("Total");

    std::vector<int> pa, pb;
    std::vector<double> ppt, pimp;
    const int cap = 16 * n;
    std::vector<int> ia(cap), ib(cap);
    std::vector<double> cmu(cap), fr(9 * (size_t)cap), aA(9 * (size_t)cap), aB(9 * (size_t)cap), pt(3 * (size_t)cap), dep(cap);
    for (int s = 0; s < steps; ++s) {
        int nc = 0;
        {
            auto sc = timings.scope("CollisionNarrowphase");
            nc = cdx_sphere_contacts(n, P.q.data(), radius.data(), mub.data(), 5, wn, wd, wmu, 0.0, cap, ia.data(), ib.data(),
                                     cmu.data(), fr.data(), aA.data(), aB.data(), pt.data(), dep.data());
        }
        if (nc < 0 || nc > cap) {
            std::fprintf(stderr, "contact generation failed\n");
            return 1;
        }
        P.body_a.assign(ia.begin(), ia.begin() + nc);
        P.body_b.assign(ib.begin(), ib.begin() + nc);
        P.mu.assign(cmu.begin(), cmu.begin() + nc);
        P.frame.assign(fr.begin(), fr.begin() + 9 * (size_t)nc);
        P.angA.assign(aA.begin(), aA.begin() + 9 * (size_t)nc);
        P.angB.assign(aB.begin(), aB.begin() + 9 * (size_t)nc);
        P.rest.assign(3 * (size_t)nc, 0.0);
        P.vbias.assign(3 * (size_t)nc, 0.0);
        P.warm.assign(3 * (size_t)nc, 0.0);
        {
            auto sc = timings.scope("CollisionMatching");
            cdx_match_warmstart((int)pa.size(), pa.data(), pb.data(), ppt.data(), pimp.data(), nc, P.body_a.data(),
                                P.body_b.data(), pt.data(), 0.02, P.warm.data());
        }
        std::vector<double> v = solver.solve(h, 1.0);
        pa = P.body_a;
        pb = P.body_b;
        ppt.assign(pt.begin(), pt.begin() + 3 * (size_t)nc);
        pimp = P.warm;
        // v_{n+1} and q_{n+1} = q_n + h B(q) v_{n+1} on the device, then back to the host state
        {
            auto sc = timings.scope("Integration");
            if (cdx_moreau_update(solver.context(), h, 1.0) != CDX_OK ||
                cdx_get_positions(solver.context(), P.q.data()) != CDX_OK) {
                std::fprintf(stderr, "moreau update failed: %s\n", cdx_last_error(solver.context()));
                return 1;
            }
            timings.harvestDevice(solver.context());
        }
        t += h;
        P.v = v;
        if (!vtk_dir.empty()) {
            auto sc = timings.scope("OutputWrite");
            if (!vtk.maybeWrite(s, t, P, radius, ppt)) {
                std::fprintf(stderr, "vtk write failed in %s\n", vtk_dir.c_str());
                return 1;
            }
        }
        double ke = 0;
        for (int i = 0; i < 6 * n; ++i) ke += 0.5 * v[i] * v[i] / P.minv[i];
        if (s % 20 == 0 || s == steps - 1)
            std::printf("step %d contacts %d iters %d KE %.12e\n", s, nc, solver.lastIterations(), ke);
    }
    { cdx::TimingManager::Scope done = std::move(total); }
    double ke = 0, pos = 0;
    for (int i = 0; i < 6 * n; ++i) ke += 0.5 * P.v[i] * P.v[i] / P.minv[i];
    for (int i = 0; i < n; ++i) pos += std::sqrt(P.q[7 * i] * P.q[7 * i] + P.q[7 * i + 1] * P.q[7 * i + 1] + P.q[7 * i + 2] * P.q[7 * i + 2]);
    std::fflush(stdout);
    timings.printBreakdown(std::cout);
    std::cout << std::flush;
    std::printf("totalKE=%.15e posNormSum=%.15e numBodies=%d solver=%s\n", ke, pos, n, solver.name());
    std::fflush(stdout);
    if (std::getenv("CARDILLO_DUMP_STATE"))
        std::cerr << std::setprecision(15) << "[STATE-DUMP] t=" << t << " numBodies=" << n << " totalKE=" << ke
                  << " posNormSum=" << pos << std::endl;
    return 0;
}

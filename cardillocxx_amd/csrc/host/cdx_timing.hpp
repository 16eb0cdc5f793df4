// cdx_timing.hpp -- section timing with an exclusive/inclusive breakdown, in the shape of the
// reference's TimingManager (src/misc/timings/TimingManager.hpp:16-163): RAII scopes on the host
// (steady clock), device kernel time booked from the library's per-kernel events
// (cdx_set_profiling / cdx_kernel_stat) under the section of its kernel family, and
// printBreakdown with the same columns (Section, Calls, Excl(s), %, Avg(us)) and the same
// "Other" residual and "Total (inclusive)" line.
#pragma once

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <iomanip>
#include <map>
#include <ostream>
#include <string>
#include <vector>

#include "cdx/cdx.h"

namespace cdx {

class TimingManager {
   public:
    struct Stats {
        double totalSec = 0.0;      // exclusive
        double inclusiveSec = 0.0;  // inclusive
        std::uint64_t calls = 0;
    };

    class Scope {
       public:
        Scope(TimingManager& mgr, const std::string& name) : m_mgr(&mgr) { m_mgr->begin(name); }
        Scope(const Scope&) = delete;
        Scope& operator=(const Scope&) = delete;
        Scope(Scope&& o) noexcept : m_mgr(o.m_mgr) { o.m_mgr = nullptr; }
        ~Scope() {
            if (m_mgr) m_mgr->end();
        }

       private:
        TimingManager* m_mgr;
    };

    Scope scope(const std::string& name) { return Scope(*this, name); }
    void reset() {
        m_stats.clear();
        m_stack.clear();
    }
    const Stats& stats(const std::string& name) const {
        static const Stats empty;
        auto it = m_stats.find(name);
        return it == m_stats.end() ? empty : it->second;
    }

    // Time measured elsewhere (device kernels): booked as exclusive time of `name` and as child
    // time of the innermost open scope, so that scope's exclusive time is what the host spent
    // around the device work (launches, polling, copies).
    void addExclusive(const std::string& name, double sec, std::uint64_t calls) {
        Stats& s = m_stats[name];
        s.totalSec += sec;
        s.inclusiveSec += sec;
        s.calls += calls;
        if (!m_stack.empty()) m_stack.back().childSec += sec;
    }

    // Section of a device kernel (the names the library records with cdx_set_profiling on).
    static std::string sectionOf(const std::string& k) {
        auto starts = [&](const char* p) { return k.rfind(p, 0) == 0; };
        if (starts("pj_setup")) return "ProjectedJacobiSetup";
        if (starts("pj_") || starts("pjc_") || starts("cheb_") || starts("and_") || k == "joint_solve" ||
            k == "joint_by_coef")
            return "ProjectedJacobiSweep";
        if (k == "joint_factor") return "BuildAndFactorS";
        if (k == "gs_color" || k == "gs_body_residual" || k == "gs_residual") return "ProjectedGaussSeidelSweep";
        if (starts("gs_")) return "ProjectedGaussSeidelSetup";
        if (starts("moreau_") || k == "gyro") return "Integration";
        if (starts("batch_")) return "ProjectedJacobiBatched";
        if (starts("adj_") || starts("unr_") || starts("gm_")) return "Adjoint";
        return "Device:" + k;
    }

    // Books every kernel timed on ctx since the last harvest, then clears the library's counters.
    void harvestDevice(cdx_ctx* ctx) {
        std::vector<std::string> names;
        char buf[64];
        int calls = 0;
        double ms = 0;
        for (int i = 0; cdx_kernel_stat(ctx, i, buf, (int)sizeof(buf), &calls, &ms) == CDX_OK; ++i) names.push_back(buf);
        for (const std::string& n : names) {
            if (cdx_kernel_time(ctx, n.c_str(), &calls, &ms) != CDX_OK) continue;
            addExclusive(sectionOf(n), ms * 1e-3, (std::uint64_t)calls);
        }
    }

    void printBreakdown(std::ostream& os) const {
        double grandInclusive = -1.0;
        if (auto it = m_stats.find("Total"); it != m_stats.end()) grandInclusive = it->second.inclusiveSec;
        double knownExclusive = 0.0;
        for (const auto& kv : m_stats)
            if (kv.first != "Total" && kv.first != "Other") knownExclusive += kv.second.totalSec;
        const double grand = grandInclusive >= 0.0 ? grandInclusive : knownExclusive;
        if (grand <= 0.0) {
            os << "\n[Timings] No recorded timings.\n";
            return;
        }
        const double other = std::max(0.0, grand - knownExclusive);
        struct Row {
            std::string name;
            double sec;
            std::uint64_t calls;
            double pct, avg;
        };
        std::vector<Row> rows;
        for (const auto& kv : m_stats) {
            if (kv.first == "Total" || kv.first == "Other") continue;
            const Stats& st = kv.second;
            rows.push_back({kv.first, st.totalSec, st.calls, st.totalSec / grand * 100.0,
                            st.calls ? st.totalSec / (double)st.calls : 0.0});
        }
        if (other > 0.0) rows.push_back({"Other", other, 0, other / grand * 100.0, 0.0});
        std::sort(rows.begin(), rows.end(), [](const Row& a, const Row& b) { return a.sec > b.sec; });
        const int w = 24 + 10 + 12 + 10 + 12;
        os << "\n[Timing Breakdown - Exclusive]\n";
        os << std::left << std::setw(24) << "Section" << std::right << std::setw(10) << "Calls" << std::setw(12)
           << "Excl(s)" << std::setw(10) << "%" << std::setw(12) << "Avg(us)" << "\n";
        os << std::string(w, '-') << "\n";
        for (const Row& r : rows)
            os << std::left << std::setw(24) << r.name << std::right << std::setw(10) << r.calls << std::setw(12)
               << std::fixed << std::setprecision(6) << r.sec << std::setw(10) << std::setprecision(2) << r.pct
               << std::setw(12) << std::setprecision(2) << r.avg * 1e6 << "\n";
        os << std::string(w, '-') << "\n";
        os << std::left << std::setw(24) << "Total (inclusive)" << std::right << std::setw(10) << "" << std::setw(12)
           << std::fixed << std::setprecision(4) << grand << std::setw(10) << 100.0 << std::setw(12) << "" << "\n";
        os.unsetf(std::ios::floatfield);
    }

   private:
    using clock = std::chrono::steady_clock;
    struct Frame {
        std::string name;
        clock::time_point start;
        double childSec = 0.0;
    };

    void begin(const std::string& name) { m_stack.push_back(Frame{name, clock::now(), 0.0}); }
    void end() {
        if (m_stack.empty()) return;
        const Frame f = m_stack.back();
        m_stack.pop_back();
        const double incl = std::chrono::duration<double>(clock::now() - f.start).count();
        Stats& s = m_stats[f.name];
        s.inclusiveSec += incl;
        s.totalSec += std::max(0.0, incl - f.childSec);
        s.calls += 1;
        if (!m_stack.empty()) m_stack.back().childSec += incl;
    }

    std::map<std::string, Stats> m_stats;
    std::vector<Frame> m_stack;
};

}  // namespace cdx

// cdx_vtk.hpp -- VTK XML PolyData (.vtp) frames and .pvd time collections of the body states
// and contacts of a contact problem, for ParaView.  Counterpart of cardillo::io::VtkWriter
// (src/io/vtk_writer.hpp:39-112, vtk_writer.cpp:92-152,240-503) without the VTK library: the same stream
// names ("<base>_geo", "<base>_contact_manifolds"), frequency rule and point-data array names,
// written as ASCII XML by this header.  Bodies are points at their centres carrying a "radius"
// array (for glyphs) instead of the reference's triangulated entity meshes.
//
//   <dir>/<base>_geo_<step>.vtp                  one point per body: velocity, angular_velocity
//                                                (body frame), angular_velocity_world, entity_id,
//                                                radius
//   <dir>/<base>_contact_manifolds_<step>.vtp    one point per contact at its position: normal,
//                                                tangent1, tangent2, friction_mu, id_a, id_b, pn,
//                                                pt_mag, percussion (world impulse pn n + pt1 t1 + pt2 t2)
//   <dir>/<base>_geo.pvd, <dir>/<base>_contact_manifolds.pvd   the frames written so far
#pragma once

#include <cmath>
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

#include "cdx_solvers.hpp"

namespace cdx {

class VtkWriter {
   public:
    VtkWriter(std::string dir, std::string base, int frequency = 1)
        : m_dir(std::move(dir)), m_base(std::move(base)), m_frequency(frequency < 1 ? 1 : frequency) {}

    // VtkWriter::maybeWrite (vtk_writer.cpp:92-95): a frame every `frequency` steps.
    bool maybeWrite(int step, double time, const ContactProblem& P, const std::vector<double>& radius,
                    const std::vector<double>& contact_points) {
        if (step % m_frequency != 0) return true;
        return write(step, time, P, radius, contact_points);
    }

    // One frame of both streams.  `contact_points` holds 3 doubles per contact of P (world
    // positions); impulses are P.warm (canonical normal >= 0, the solver's write-back).
    bool write(int step, double time, const ContactProblem& P, const std::vector<double>& radius,
               const std::vector<double>& contact_points) {
        const std::string g = frame_name("_geo", step), c = frame_name("_contact_manifolds", step);
        if (!write_bodies(m_dir + "/" + g, P, radius)) return false;
        if (!write_contacts(m_dir + "/" + c, P, contact_points)) return false;
        m_geo.emplace_back(time, g);
        m_con.emplace_back(time, c);
        return write_pvd(m_dir + "/" + m_base + "_geo.pvd", m_geo) &&
               write_pvd(m_dir + "/" + m_base + "_contact_manifolds.pvd", m_con);
    }

   private:
    std::string m_dir, m_base;
    int m_frequency;
    std::vector<std::pair<double, std::string>> m_geo, m_con;

    std::string frame_name(const char* stream, int step) const {
        char buf[32];
        std::snprintf(buf, sizeof buf, "_%06d.vtp", step);
        return m_base + stream + buf;
    }

    static void array(FILE* f, const char* type, const char* name, int comps, const std::vector<double>& v) {
        std::fprintf(f, "        <DataArray type=\"%s\" Name=\"%s\" NumberOfComponents=\"%d\" format=\"ascii\">\n", type, name,
                     comps);
        for (size_t i = 0; i < v.size(); ++i) std::fprintf(f, i + 1 < v.size() ? "%.17g " : "%.17g", v[i]);
        std::fprintf(f, "\n        </DataArray>\n");
    }

    static void int_array(FILE* f, const char* name, const std::vector<int>& v) {
        std::fprintf(f, "        <DataArray type=\"Int32\" Name=\"%s\" NumberOfComponents=\"1\" format=\"ascii\">\n", name);
        for (size_t i = 0; i < v.size(); ++i) std::fprintf(f, i + 1 < v.size() ? "%d " : "%d", v[i]);
        std::fprintf(f, "\n        </DataArray>\n");
    }

    // Points with one vertex cell each, then the point-data arrays.
    static FILE* open_polydata(const std::string& path, const std::vector<double>& xyz) {
        FILE* f = std::fopen(path.c_str(), "w");
        if (!f) return nullptr;
        const size_t n = xyz.size() / 3;
        std::fprintf(f, "<?xml version=\"1.0\"?>\n");
        std::fprintf(f, "<VTKFile type=\"PolyData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n");
        std::fprintf(f, "  <PolyData>\n");
        std::fprintf(f, "    <Piece NumberOfPoints=\"%zu\" NumberOfVerts=\"%zu\" NumberOfLines=\"0\" NumberOfStrips=\"0\" "
                        "NumberOfPolys=\"0\">\n", n, n);
        std::fprintf(f, "      <Points>\n");
        array(f, "Float64", "Points", 3, xyz);
        std::fprintf(f, "      </Points>\n      <Verts>\n");
        std::vector<int> conn(n), off(n);
        for (size_t i = 0; i < n; ++i) {
            conn[i] = (int)i;
            off[i] = (int)i + 1;
        }
        int_array(f, "connectivity", conn);
        int_array(f, "offsets", off);
        std::fprintf(f, "      </Verts>\n      <PointData>\n");
        return f;
    }

    static bool close_polydata(FILE* f) {
        std::fprintf(f, "      </PointData>\n    </Piece>\n  </PolyData>\n</VTKFile>\n");
        return std::fclose(f) == 0;
    }

    // World angular velocity R(q) w_body from the body-frame twist (quaternion x, y, z, w).
    static void rotate(const double* q, const double* w, double* o) {
        const double x = q[0], y = q[1], z = q[2], s = q[3];
        const double R[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - z * s),     2 * (x * z + y * s),
                             2 * (x * y + z * s),     1 - 2 * (x * x + z * z), 2 * (y * z - x * s),
                             2 * (x * z - y * s),     2 * (y * z + x * s),     1 - 2 * (x * x + y * y)};
        for (int i = 0; i < 3; ++i) o[i] = R[3 * i] * w[0] + R[3 * i + 1] * w[1] + R[3 * i + 2] * w[2];
    }

    static bool write_bodies(const std::string& path, const ContactProblem& P, const std::vector<double>& radius) {
        const int nb = (int)P.dof.size();
        std::vector<double> xyz, vel, ang, angw, rad;
        std::vector<int> id;
        size_t vo = 0, qo = 0;
        for (int b = 0; b < nb; ++b) {
            const int d = P.dof[b];
            const double* v = P.v.data() + vo;
            const double* q = P.q.data() + qo;
            xyz.insert(xyz.end(), q, q + 3);
            vel.insert(vel.end(), v, v + 3);
            double w[3] = {0, 0, 0}, ww[3] = {0, 0, 0};
            if (d == 6) {
                for (int k = 0; k < 3; ++k) w[k] = v[3 + k];
                rotate(q + 3, w, ww);
            }
            ang.insert(ang.end(), w, w + 3);
            angw.insert(angw.end(), ww, ww + 3);
            id.push_back(b);
            rad.push_back(b < (int)radius.size() ? radius[b] : 0.0);
            vo += d;
            qo += d == 6 ? 7 : 3;
        }
        FILE* f = open_polydata(path, xyz);
        if (!f) return false;
        array(f, "Float64", "velocity", 3, vel);
        array(f, "Float64", "angular_velocity", 3, ang);
        array(f, "Float64", "angular_velocity_world", 3, angw);
        int_array(f, "entity_id", id);
        array(f, "Float64", "radius", 1, rad);
        return close_polydata(f);
    }

    static bool write_contacts(const std::string& path, const ContactProblem& P, const std::vector<double>& pts) {
        const size_t nc = P.body_a.size();
        std::vector<double> xyz(pts.begin(), pts.begin() + std::min(pts.size(), 3 * nc)), nrm, t1, t2, mu, pn, ptm, perc;
        xyz.resize(3 * nc, 0.0);
        for (size_t c = 0; c < nc; ++c) {
            const double* fr = P.frame.data() + 9 * c;
            const double* w = P.warm.data() + 3 * c;
            nrm.insert(nrm.end(), fr, fr + 3);
            t1.insert(t1.end(), fr + 3, fr + 6);
            t2.insert(t2.end(), fr + 6, fr + 9);
            mu.push_back(P.mu[c]);
            pn.push_back(w[0]);
            ptm.push_back(std::sqrt(w[1] * w[1] + w[2] * w[2]));
            for (int k = 0; k < 3; ++k) perc.push_back(w[0] * fr[k] + w[1] * fr[3 + k] + w[2] * fr[6 + k]);
        }
        FILE* f = open_polydata(path, xyz);
        if (!f) return false;
        array(f, "Float64", "normal", 3, nrm);
        array(f, "Float64", "tangent1", 3, t1);
        array(f, "Float64", "tangent2", 3, t2);
        array(f, "Float64", "friction_mu", 1, mu);
        int_array(f, "id_a", P.body_a);
        int_array(f, "id_b", P.body_b);
        array(f, "Float64", "pn", 1, pn);
        array(f, "Float64", "pt_mag", 1, ptm);
        array(f, "Float64", "percussion", 3, perc);
        return close_polydata(f);
    }

    // PvdWriter (src/io/pvd_writer.cpp): the collection of frames with their times.
    static bool write_pvd(const std::string& path, const std::vector<std::pair<double, std::string>>& frames) {
        FILE* f = std::fopen(path.c_str(), "w");
        if (!f) return false;
        std::fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
        std::fprintf(f, "  <Collection>\n");
        for (const auto& fr : frames)
            std::fprintf(f, "    <DataSet timestep=\"%.17g\" group=\"\" part=\"0\" file=\"%s\"/>\n", fr.first, fr.second.c_str());
        std::fprintf(f, "  </Collection>\n</VTKFile>\n");
        return std::fclose(f) == 0;
    }
};

}  // namespace cdx

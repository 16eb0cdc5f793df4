"""The VTK writer of the C++ adapter (cardillocxx_amd/csrc/host/cdx_vtk.hpp, the counterpart of
cardillo::io::VtkWriter, src/io/vtk_writer.hpp:39-112): a small C++ driver built with g++ writes
frames of a three-body, two-contact problem; the XML is parsed back and its arrays are checked
against values fixed by hand: the frame cadence of maybeWrite (step % frequency == 0,
vtk_writer.cpp:92-95), the world angular velocity R(q) w of a quarter turn about z, the contact
percussion pn n + pt1 t1 + pt2 t2 and its tangential magnitude, and the .pvd time collection."""
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#include <cmath>
#include "cdx_vtk.hpp"
int main(int, char** argv) {
    cdx::ContactProblem P;
    P.dof = {6, 3, 6};
    const double s = std::sqrt(0.5);
    P.q = {0, 0, 1, 0, 0, s, s,   1, 2, 3,   -1, 0, 0.5, 0, 0, 0, 1};
    P.v = {0.1, 0.2, 0.3, 1, 0, 0,   -1, -2, -3,   0, 0, 0, 0, 0, 2};
    P.body_a = {0, 1};
    P.body_b = {-1, 2};
    P.mu = {0.5, 0.0};
    P.frame = {0, 0, 1, 1, 0, 0, 0, 1, 0,   1, 0, 0, 0, 1, 0, 0, 0, 1};
    P.warm = {2, 0.3, -0.4,   1.5, 0, 0};
    std::vector<double> radius = {0.05, 0.02, 0.1};
    std::vector<double> pts = {0, 0, 0.95,   0.5, 1, 1.75};
    cdx::VtkWriter w(argv[1], "t", 2);
    for (int step = 0; step < 5; ++step)
        if (!w.maybeWrite(step, 0.001 * (step + 1), P, radius, pts)) return 1;
    return 0;
}
"""


@pytest.fixture(scope="module")
def frames(tmp_path_factory):
    d = tmp_path_factory.mktemp("vtk")
    src = d / "drv.cpp"
    src.write_text(DRIVER)
    exe = d / "drv"
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "cardillocxx_amd", "csrc", "host")]
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Wextra", "-Werror", *inc, "-o", str(exe), str(src)], check=True)
    out = d / "out"
    out.mkdir()
    subprocess.run([str(exe), str(out)], check=True)
    return out


def arrays(path):
    root = ET.parse(path).getroot()
    piece = root.find("PolyData/Piece")
    out = {"Points": np.array(piece.find("Points/DataArray").text.split(), float).reshape(-1, 3)}
    for a in piece.find("PointData"):
        vals = np.array(a.text.split(), float if a.get("type") == "Float64" else int)
        nc = int(a.get("NumberOfComponents"))
        out[a.get("Name")] = vals.reshape(-1, nc) if nc > 1 else vals
    out["npoints"] = int(piece.get("NumberOfPoints"))
    out["verts"] = np.array(piece.find("Verts/DataArray[@Name='connectivity']").text.split(), int)
    return out


def test_frames_follow_the_frequency(frames):
    names = sorted(os.listdir(frames))
    assert names == ["t_contact_manifolds.pvd", "t_contact_manifolds_000000.vtp", "t_contact_manifolds_000002.vtp",
                     "t_contact_manifolds_000004.vtp", "t_geo.pvd", "t_geo_000000.vtp", "t_geo_000002.vtp",
                     "t_geo_000004.vtp"]
    for stream in ("t_geo", "t_contact_manifolds"):
        ds = ET.parse(frames / f"{stream}.pvd").getroot().findall("Collection/DataSet")
        assert [d.get("file") for d in ds] == [f"{stream}_{s:06d}.vtp" for s in (0, 2, 4)]
        assert [float(d.get("timestep")) for d in ds] == pytest.approx([0.001, 0.003, 0.005], rel=0, abs=1e-18)


def test_body_frame(frames):
    a = arrays(frames / "t_geo_000002.vtp")
    assert a["npoints"] == 3 and list(a["verts"]) == [0, 1, 2]
    np.testing.assert_array_equal(a["Points"], [[0, 0, 1], [1, 2, 3], [-1, 0, 0.5]])
    np.testing.assert_array_equal(a["velocity"], [[0.1, 0.2, 0.3], [-1, -2, -3], [0, 0, 0]])
    np.testing.assert_array_equal(a["angular_velocity"], [[1, 0, 0], [0, 0, 0], [0, 0, 2]])
    # a quarter turn about z takes the body x axis to world y; the identity leaves z alone
    np.testing.assert_allclose(a["angular_velocity_world"], [[0, 1, 0], [0, 0, 0], [0, 0, 2]], atol=1e-15)
    assert list(a["entity_id"]) == [0, 1, 2]
    np.testing.assert_array_equal(a["radius"], [0.05, 0.02, 0.1])


def test_contact_frame(frames):
    a = arrays(frames / "t_contact_manifolds_000004.vtp")
    assert a["npoints"] == 2
    np.testing.assert_array_equal(a["Points"], [[0, 0, 0.95], [0.5, 1, 1.75]])
    np.testing.assert_array_equal(a["normal"], [[0, 0, 1], [1, 0, 0]])
    np.testing.assert_array_equal(a["tangent1"], [[1, 0, 0], [0, 1, 0]])
    np.testing.assert_array_equal(a["tangent2"], [[0, 1, 0], [0, 0, 1]])
    assert list(a["id_a"]) == [0, 1] and list(a["id_b"]) == [-1, 2]
    np.testing.assert_array_equal(a["friction_mu"], [0.5, 0.0])
    np.testing.assert_array_equal(a["pn"], [2, 1.5])
    np.testing.assert_allclose(a["pt_mag"], [0.5, 0.0], rtol=1e-15)
    # percussion = pn n + pt1 t1 + pt2 t2 in world coordinates
    np.testing.assert_allclose(a["percussion"], [[0.3, -0.4, 2.0], [1.5, 0, 0]], rtol=1e-15)

"""VTK PolyData snapshots (the reference's io::VTKWriter and its VTKTest,
tests/integration/VTKTest.cpp: four cells written as vertices with a Level
array)."""
import numpy as np

from fluidloom_amd import io_vtk


def test_vtk_round_trip(tmp_path):
    # the reference test's cells: (0,0,0), (1,0,0), (0,1,0), (0,0,1) at level 0, plus a level-2 cell
    x = np.array([0, 1, 0, 0, 5]); y = np.array([0, 0, 1, 0, 6]); z = np.array([0, 0, 0, 1, 7])
    lv = np.array([0, 0, 0, 0, 2])
    rho = np.linspace(0.9, 1.1, 5)
    u = np.arange(15, dtype=np.float64).reshape(5, 3) * 1e-3
    p = io_vtk.write_vtp(tmp_path / "snapshot_000000.vtp", x, y, z, lv, {"rho": rho, "velocity": u})
    txt = p.read_text()
    assert '<VTKFile type="PolyData"' in txt and 'NumberOfPoints="5"' in txt and 'NumberOfVerts="5"' in txt
    d = io_vtk.read_vtp(p)
    assert np.allclose(d["Points"][4], [5.5 / 4, 6.5 / 4, 7.5 / 4])
    assert np.allclose(d["Points"][1], [1.5, 0.5, 0.5])
    assert d["Level"].tolist() == lv.tolist()
    assert np.array_equal(d["connectivity"], np.arange(5)) and np.array_equal(d["offsets"], np.arange(1, 6))
    assert np.allclose(d["rho"], rho.astype(np.float32)) and np.allclose(d["velocity"], u.astype(np.float32))

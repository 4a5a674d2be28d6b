"""Run directories of the measured configurations (tlbm_amd/cases.py)."""
import os

import numpy as np

from tlbm_amd import cases, geometry, io_formats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_channel_case_is_the_reference_fixture_lattice(tmp_path):
    """The 24x24x96 case writes the node types the reference ran on (the
    fixture made with the reference executable)."""
    d = str(tmp_path / "c1")
    p, nd = cases.write("channel", d, parts=4, binary=True)
    g = np.load(os.path.join(ROOT, "tests", "golden", "channel_d3q15_24x24x96_f64.npz"))
    assert np.array_equal(nd, g["ndType"])
    rp = io_formats.read_params(os.path.join(d, "params.lbm"))
    assert (rp.latticeType, rp.nx, rp.ny, rp.nz, rp.numTs) == ("D3Q15", 24, 24, 96, 2000)
    assert rp.omega == 1.0 / 0.6 and rp.uLBM == 0.05
    assert np.array_equal(io_formats.read_int_lines(os.path.join(d, "ndType.lbm")), nd)
    assert np.array_equal(io_formats.read_ndtype_bin(os.path.join(d, "ndType.bin")), nd)
    parts = io_formats.read_int_lines(os.path.join(d, "parts.lbm"))
    assert np.array_equal(np.bincount(parts), [24 * 24 * 24] * 4)
    assert np.all(np.diff(parts) >= 0)  # contiguous z slabs


def test_cylinder_case_geometry():
    p, nd = cases.cylinders()
    g = nd.reshape(512, 128, 128)
    z, y = np.meshgrid(np.arange(512), np.arange(128), indexing="ij")
    inside = (y - 64) ** 2 + (z - 128) ** 2 < 144
    assert np.array_equal(g[:, :, 5] == geometry.SOLID, inside)
    assert np.all(g == g[:, :, :1])  # uniform along the cylinder axis x
    assert np.all(g[0][~inside[0]] == geometry.INLET) and np.all(g[-1] == geometry.OUTLET)
    assert (p.nx, p.ny, p.nz, p.numTs, p.uLBM) == (128, 128, 512, 20000, 0.04)

"""Host-side inputs: params.lbm / ndType.lbm / parts.lbm and the node
classification rules, checked by brute force on small lattices."""
import os

import numpy as np

from tlbm_amd import geometry, io_formats


def test_params_round_trip(tmp_path):
    p = io_formats.Params(latticeType="D3Q27", dynamics=1, numTs=2000, tsRepFreq=100,
                          warmupTs=3, plotFreq=500, cs=0.1, rhoLBM=1.0, uLBM=0.05,
                          omega=1.0 / 0.6, nx=24, ny=25, nz=96, restartFlag=0, timeAvgFlag=1,
                          pRefIdx=17)
    io_formats.write_params(tmp_path / "params.lbm", p)
    text = (tmp_path / "params.lbm").read_text().split()
    assert len(text) == 23 and text[0] == "D3Q27" and text[10:13] == ["24", "25", "96"]
    q = io_formats.read_params(tmp_path / "params.lbm")
    assert q == p and q.omega == 1.0 / 0.6 and q.lattice_q == 27


def test_int_lines_and_slab_parts(tmp_path):
    v = np.array([0, 1, 2, 3, 5, 1])
    io_formats.write_int_lines(tmp_path / "x.lbm", v)
    assert np.array_equal(io_formats.read_int_lines(tmp_path / "x.lbm"), v)
    parts = io_formats.slab_parts(2, 3, 8, 4)
    assert np.array_equal(np.bincount(parts), [12, 12, 12, 12])
    assert np.all(np.diff(parts) >= 0)


def _brute_channel(nx, ny, nz, sphere=None):
    out = np.zeros(nx * ny * nz, int)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                g = x + nx * y + nx * ny * z
                wall = x in (0, nx - 1) or y in (0, ny - 1)
                obst = False
                if sphere is not None:
                    cx, cy, cz, r = sphere
                    obst = (x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 < r * r
                if wall or obst:
                    out[g] = 1
                elif z == 0:
                    out[g] = 2
                elif z == nz - 1:
                    out[g] = 3
    return out


def test_channel_classification_brute_force():
    nx, ny, nz = 7, 6, 9
    got = geometry.channel(nx, ny, nz, obstacles=[geometry.sphere(3, 2, 1, 2)])
    want = _brute_channel(nx, ny, nz, (3, 2, 1, 2))
    assert np.array_equal(got, want)
    # the sphere reaches z=0: those inlet nodes become solid
    assert got[3 + nx * 2] == 1


def test_sphere_is_strict_and_spheroid_inclusive():
    x, y, z = geometry.grid(9, 9, 9)
    s = geometry.sphere(4, 4, 4, 2)(x, y, z)
    assert not s[4, 4, 6] and s[4, 4, 5]          # d^2 = r^2 excluded
    e = geometry.prolate_spheroid(4, 4, 4, 2, 3)(x, y, z)
    assert e[7, 4, 4] and not e[8, 4, 4]          # z semi-axis 3 included
    assert e[4, 4, 6] and not e[4, 4, 7]          # x semi-axis 2 included
    c = geometry.cylinder_x(4, 4, 2)(x, y, z)
    assert c[4, 4, :].all() and not c[4, 6, 0]


def test_lid_cavity_and_periodic_box():
    nd = geometry.lid_cavity(5, 6, 7).reshape(7, 6, 5)
    assert np.all(nd[:, 5, :] == 5)
    assert np.all(nd[0, :5, :] == 1) and np.all(nd[:, :5, 0] == 1)
    assert np.all(nd[1:6, 1:5, 1:4] == 0)
    box = geometry.periodic_box(4, 4, 4)
    assert np.all(box == 0)


def test_rsa_pack_is_seeded_and_non_overlapping():
    c, r = geometry.rsa_spheres(40, 40, 40, 3, 5, porosity=0.85, seed=7)
    c2, r2 = geometry.rsa_spheres(40, 40, 40, 3, 5, porosity=0.85, seed=7)
    assert np.array_equal(c, c2) and np.array_equal(r, r2)
    d = np.sqrt(((c[:, None, :] - c[None, :, :]) ** 2).sum(-1))
    np.fill_diagonal(d, np.inf)
    assert np.all(d >= r[:, None] + r[None, :])
    assert 4.0 / 3.0 * np.pi * (r ** 3).sum() >= 0.15 * 40 ** 3


def test_rsa_pack_reaches_lattice_porosity():
    n = 48
    solid = np.zeros((n, n, n), bool)
    solid[:, :, 0] = solid[:, :, -1] = solid[:, 0, :] = solid[:, -1, :] = True
    walls = solid.copy()
    c, r = geometry.rsa_spheres(n, n, n, 3, 5, porosity=0.8, seed=3, solid_mask=solid)
    mask = geometry.sphere_pack_mask(n, n, n, c, r)
    assert np.array_equal(solid, mask | walls)
    porosity = 1.0 - solid.mean()
    # stops with the first sphere that crosses the target
    last = geometry.sphere_pack_mask(n, n, n, c[-1:], r[-1:]).sum()
    assert 0.8 - last / n ** 3 <= porosity <= 0.8


def test_int_lines_match_printf_and_read_back(tmp_path):
    rng = np.random.default_rng(4)
    for v in (rng.integers(0, 6, 500), rng.integers(-5000, 200000, 2000), np.array([0]),
              np.array([-1, 10, 9, 100000000])):
        want = "".join("%d \n" % x for x in v.tolist()).encode()
        assert io_formats.int_lines_bytes(v) == want
        path = os.path.join(str(tmp_path), "x.lbm")
        io_formats.write_int_lines(path, v)
        assert np.array_equal(io_formats.read_int_lines(path), v)


def test_ndtype_sidecar_round_trip(tmp_path):
    """Header: magic, node count, then the size and CRC-32 of the text file
    it was written from (zlib's CRC-32, the IEEE polynomial the driver
    computes), zeros without one."""
    import zlib
    nd = geometry.channel(7, 5, 9, obstacles=[geometry.sphere(3, 2, 4, 2)])
    path = os.path.join(str(tmp_path), "ndType.bin")
    io_formats.write_ndtype_bin(path, nd)
    assert np.array_equal(io_formats.read_ndtype_bin(path), nd)
    raw = open(path, "rb").read()
    assert raw[:8] == b"TLBMNDT2" and int.from_bytes(raw[8:16], "little") == nd.size
    assert raw[16:32] == bytes(16)
    assert len(raw) == 32 + nd.size
    text = os.path.join(str(tmp_path), "ndType.lbm")
    io_formats.write_int_lines(text, nd)
    io_formats.write_ndtype_bin(path, nd, text_path=text)
    raw = open(path, "rb").read()
    body = open(text, "rb").read()
    assert int.from_bytes(raw[16:24], "little") == len(body)
    assert int.from_bytes(raw[24:32], "little") == zlib.crc32(body)
    # a known CRC-32 value: the standard check string
    assert zlib.crc32(b"123456789") == 0xCBF43926

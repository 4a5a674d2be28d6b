"""Drop-in driver and restart tool on the CPU: input validation happens
before any GPU call, so every error path is exercised here."""
import os
import subprocess

import numpy as np
import pytest

from exec_util import RESTART, case_dir, require, run_exec
from tlbm_amd import geometry, io_formats


def good_params(**kw):
    p = io_formats.Params(latticeType="D3Q19", numTs=4, tsRepFreq=2, plotFreq=2, omega=1.5,
                          uLBM=0.05, nx=6, ny=5, nz=8)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def nd_for(p):
    return geometry.channel(p.nx, p.ny, p.nz)


def test_missing_params_file(tmp_path):
    r = run_exec(str(tmp_path))
    assert r.returncode == 1 and "Could not open params.lbm!" in r.stderr


@pytest.mark.parametrize("kw,msg", [
    (dict(latticeType="D3Q13"), "Invalid Lattice Structure!"),
    (dict(plotFreq=0), "plotFreq and tsRepFreq must be positive"),
    (dict(tsRepFreq=0), "plotFreq and tsRepFreq must be positive"),
    (dict(nx=0), "nx, ny, nz must be positive"),
])
def test_bad_parameters_rejected(tmp_path, kw, msg):
    p = good_params(**kw)
    d = case_dir(tmp_path, good_params(), nd_for(good_params()))
    io_formats.write_params(os.path.join(d, "params.lbm"), p)
    r = run_exec(d)
    assert r.returncode == 1 and msg in r.stderr


def test_truncated_params_rejected(tmp_path):
    d = case_dir(tmp_path, good_params(), nd_for(good_params()))
    text = open(os.path.join(d, "params.lbm")).read().split("\n")
    open(os.path.join(d, "params.lbm"), "w").write("\n".join(text[:10]))
    r = run_exec(d)
    assert r.returncode == 1 and "missing or malformed value for nx" in r.stderr


def test_unknown_node_type_rejected(tmp_path):
    p = good_params()
    nd = nd_for(p).astype(int)
    nd[7] = 4
    d = case_dir(tmp_path, p, nd)
    r = run_exec(d)
    assert r.returncode == 1 and "node 7 has type 4" in r.stderr


@pytest.mark.parametrize("delta", [-1, 1])
def test_ndtype_length_must_match(tmp_path, delta):
    p = good_params()
    nd = nd_for(p)
    d = case_dir(tmp_path, p, nd)
    io_formats.write_int_lines(os.path.join(d, "ndType.lbm"),
                               nd[:delta] if delta < 0 else np.append(nd, 0))
    r = run_exec(d)
    assert r.returncode == 1 and "ndType.lbm has" in r.stderr


def test_missing_ndtype_rejected(tmp_path):
    p = good_params()
    d = case_dir(tmp_path, p, nd_for(p))
    os.remove(os.path.join(d, "ndType.lbm"))
    r = run_exec(d)
    assert r.returncode == 1 and "Could not open ndType.lbm" in r.stderr


def test_parts_length_must_match(tmp_path):
    p = good_params()
    d = case_dir(tmp_path, p, nd_for(p))
    io_formats.write_int_lines(os.path.join(d, "parts.lbm"), np.zeros(p.num_nodes - 3, int))
    r = run_exec(d)
    assert r.returncode == 1 and "parts.lbm has" in r.stderr


def test_bad_storage_rejected(tmp_path):
    p = good_params()
    d = case_dir(tmp_path, p, nd_for(p))
    r = run_exec(d, storage="bf16")
    assert r.returncode == 1 and "TLBM_STORAGE must be" in r.stderr


def _h5_values(path, dataset, tmp):
    out = os.path.join(tmp, "v.bin")
    subprocess.run(["/opt/conda/bin/h5dump", "-d", dataset, "-b", "LE", "-o", out, path],
                   check=True, capture_output=True)
    return np.fromfile(out, dtype=np.float32)


@pytest.mark.parametrize("dtype,flag", [(np.float32, "f32"), (np.float64, "f64")])
def test_restart_tool_reorders_and_writes_float32(tmp_path, dtype, flag):
    d = str(tmp_path)
    n = 12
    order = np.random.default_rng(1).permutation(n).astype(np.int32)
    order.tofile(os.path.join(d, "ordering.b_dat"))
    vals = {}
    for i, f in enumerate(("ux", "uy", "uz", "density")):
        v = (np.arange(n) * (i + 1) * 0.01 + 0.1).astype(dtype)
        v.tofile(os.path.join(d, f"{f}3.b_dat"))
        g = np.empty(n, dtype)
        g[order] = v
        vals[f] = g.astype(np.float32)
    r = subprocess.run([require(RESTART), "3", flag], cwd=d, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    h5 = os.path.join(d, "restart.h5")
    for f, ds in (("ux", "/velocity/x"), ("uy", "/velocity/y"), ("uz", "/velocity/z"),
                  ("density", "/density/rho")):
        assert np.array_equal(_h5_values(h5, ds, d), vals[f])
    # the reference's script opens with 'w-': an existing file is not overwritten
    r2 = subprocess.run([RESTART, "3", flag], cwd=d, capture_output=True, text=True)
    assert r2.returncode == 1


def test_ndtype_sidecar_is_read_and_validated(tmp_path):
    p = good_params()
    nd = nd_for(p).astype(int)
    d = case_dir(tmp_path, p, nd)
    text = os.path.join(d, "ndType.lbm")
    bad = nd.copy()
    bad[7] = 4
    io_formats.write_ndtype_bin(os.path.join(d, "ndType.bin"), bad, text_path=text)
    r = run_exec(d)
    assert r.returncode == 1 and "ndType.bin: node 7 has type 4" in r.stderr
    io_formats.write_ndtype_bin(os.path.join(d, "ndType.bin"), nd[:-1], text_path=text)
    r = run_exec(d)
    assert r.returncode == 1 and "ndType.bin has" in r.stderr
    with open(os.path.join(d, "ndType.bin"), "wb") as fh:
        fh.write(b"NOTASIDE" + bytes(24))
    r = run_exec(d)
    assert r.returncode == 1 and "not an ndType sidecar" in r.stderr
    # without a text file the sidecar is read as it is
    os.remove(text)
    io_formats.write_ndtype_bin(os.path.join(d, "ndType.bin"), bad)
    r = run_exec(d)
    assert r.returncode == 1 and "ndType.bin: node 7 has type 4" in r.stderr


@pytest.mark.parametrize("touch", ["newer_sidecar", "older_sidecar"])
def test_stale_ndtype_sidecar_is_ignored(tmp_path, touch):
    """A sidecar written from another ndType.lbm is not used, whatever the
    files' times: here the text file was rewritten (same node count and
    length, one node changed) after the sidecar was made from the old one,
    and the driver reads (and rejects) the new text."""
    p = good_params()
    nd = nd_for(p).astype(int)
    d = case_dir(tmp_path, p, nd)
    text = os.path.join(d, "ndType.lbm")
    io_formats.write_ndtype_bin(os.path.join(d, "ndType.bin"), nd, text_path=text)
    bad = nd.copy()
    bad[7] = 4 if nd[7] != 4 else 5
    io_formats.write_int_lines(text, bad)
    assert os.path.getsize(text) == int.from_bytes(
        open(os.path.join(d, "ndType.bin"), "rb").read()[16:24], "little")
    t = os.stat(text).st_mtime
    side = t + 10 if touch == "newer_sidecar" else t - 10
    os.utime(os.path.join(d, "ndType.bin"), (side, side))
    r = run_exec(d)
    assert r.returncode == 1 and "ndType.lbm: node 7 has type 4" in r.stderr
    assert "ndType.bin was not written from this ndType.lbm" in r.stderr


@pytest.fixture(scope="module")
def stub_exec(tmp_path_factory):
    """The driver's host sources linked against the C-ABI test double
    (tests/cabi_stub.cpp) instead of the GPU library: the driver's control
    flow on a machine without a GPU."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    host = os.path.join(root, "tlbm_amd", "host")
    exe = str(tmp_path_factory.mktemp("stub") / "tlbm_exec_stub")
    srcs = [os.path.join(host, f) for f in ("Problem.cpp", "TLBM_Partition.cpp", "h5_io.cpp",
                                             "tlbm_main.cpp")]
    srcs.append(os.path.join(root, "tests", "cabi_stub.cpp"))
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(root, "include"), "-I", host,
                    "-I", os.path.join(root, "tlbm_amd", "csrc"), *srcs, "-o", exe], check=True)
    return exe


def _run_stub(exe, d, env):
    e = dict(os.environ)
    e.update(env)
    return subprocess.run([exe], cwd=d, capture_output=True, text=True, env=e, timeout=60)


def test_driver_stops_when_the_state_is_not_finite(tmp_path, stub_exec):
    """Divergence guard: the driver checks the state every tsRepFreq steps
    and, when the library reports NaN or infinite populations, stops with a
    non-zero status naming the time step (here the library reports them
    from step 12 on; the first check after that is at step 15)."""
    p = good_params(numTs=40, tsRepFreq=5, plotFreq=100)
    d = case_dir(tmp_path, p, nd_for(p))
    r = _run_stub(stub_exec, d, {"TLBM_STUB_NONFINITE_FROM": "12"})
    assert r.returncode == 1
    assert "time step 15:" in r.stderr and "NaN or infinite" in r.stderr
    assert "Executing time step: 15" in r.stdout and "Executing time step: 20" not in r.stdout
    assert "Test complete." not in r.stdout


def test_driver_guard_can_be_switched_off(tmp_path, stub_exec):
    p = good_params(numTs=40, tsRepFreq=5, plotFreq=100)
    d = case_dir(tmp_path, p, nd_for(p))
    r = _run_stub(stub_exec, d, {"TLBM_STUB_NONFINITE_FROM": "12", "TLBM_CHECK_FINITE": "0"})
    assert r.returncode == 0, r.stderr
    assert "Test complete." in r.stdout
    r = _run_stub(stub_exec, d, {})
    assert r.returncode == 0 and "Executing time step: 40" in r.stdout

"""The drop-in driver end to end on the GPU: tlbm_exec reads the reference's
input files and writes the reference's dump files.  In FP64 storage the
dumps match the reference's own dumps (golden fixtures) and the oracle."""
import os
import subprocess

import numpy as np
import pytest

import goldens
from exec_util import RESTART, case_dir, read_dump, read_raw, require, run_exec
from pyoracle import Oracle
from tlbm_amd import geometry, io_formats

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def rel_err(got, ref):
    scale = np.max(np.abs(ref))
    return float(np.max(np.abs(got - ref)) / scale) if scale else float(np.max(np.abs(got)))


def params_of(case):
    p = case.params
    return io_formats.Params(**{k: p[k] for k in io_formats.PARAM_ORDER})


EXEC_CASES = [c for c in goldens.load_cases() if c.precision == "f64" and
              (c.params["numTs"] >= 20)]


@pytest.mark.parametrize("case", EXEC_CASES, ids=[c.name for c in EXEC_CASES])
def test_exec_dumps_match_reference_dumps(gpu_lib, tmp_path, case):
    print(f"exec {case.name}", flush=True)
    d = case_dir(tmp_path, params_of(case), case.nd)
    r = run_exec(d, storage="f64")
    assert r.returncode == 0, r.stderr
    assert "Estimated LPU/s" in r.stdout
    n = case.nd.size
    assert np.array_equal(read_raw(d, "ordering.b_dat", np.int32), np.arange(n, dtype=np.int32))
    got = read_dump(d, 0, np.float64)
    for f, ref in case.out.items():
        assert rel_err(got[f], ref) <= TOL64, f


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f32x", 1e-5), ("f16", 1e-3)])
def test_exec_dump_schedule_and_float_dumps(gpu_lib, tmp_path, storage, tol):
    """Dumps at (ts+1) % plotFreq == 0 after warm-up (tLBM.cpp:55-59); FP32,
    FP32X and FP16 storage write float .b_dat files."""
    nx, ny, nz = 10, 8, 16
    nd = geometry.channel(nx, ny, nz, obstacles=[geometry.sphere(5, 4, 6, 2)])
    p = io_formats.Params(latticeType="D3Q19", numTs=50, tsRepFreq=25, warmupTs=20, plotFreq=10,
                          uLBM=0.05, omega=1.6, nx=nx, ny=ny, nz=nz)
    d = case_dir(tmp_path, p, nd)
    r = run_exec(d, storage=storage)
    assert r.returncode == 0, r.stderr
    assert "Executing time step: 25" in r.stdout and "Executing time step: 50" in r.stdout
    files = sorted(f for f in os.listdir(d) if f.startswith("ux"))
    assert files == ["ux0.b_dat", "ux1.b_dat", "ux2.b_dat"]
    o = Oracle(19, nx, ny, nz, nd, omega=1.6, u_lbm=0.05)
    for k, step in enumerate((30, 40, 50)):
        o.step(30 if k == 0 else 10)
        ref = o.macros()
        got = read_dump(d, k, np.float32)
        fluid = nd == 0
        u_ref = np.stack([ref[c][fluid] for c in ("ux", "uy", "uz")])
        u_got = np.stack([got[c][fluid] for c in ("ux", "uy", "uz")]).astype(np.float64)
        assert np.linalg.norm(u_got - u_ref) / np.linalg.norm(u_ref) < tol, k


def test_exec_time_average_sums(gpu_lib, tmp_path):
    nx, ny, nz = 8, 7, 12
    nd = geometry.channel(nx, ny, nz, obstacles=[geometry.sphere(3, 3, 4, 2)])
    p = io_formats.Params(latticeType="D3Q15", numTs=20, tsRepFreq=20, plotFreq=20, uLBM=0.05,
                          omega=1.6, nx=nx, ny=ny, nz=nz, timeAvgFlag=1)
    d = case_dir(tmp_path, p, nd)
    r = run_exec(d, storage="f64")
    assert r.returncode == 0, r.stderr
    assert "Writing time average data." in r.stdout
    o = Oracle(15, nx, ny, nz, nd, omega=1.6, u_lbm=0.05)
    sums = {k: np.zeros(nd.size) for k in ("ux", "uy", "uz", "rho")}
    for _ in range(20):
        o.step(1)
        m = o.macros()
        for k in sums:
            sums[k] += m[k]
    for name, k in (("uAvg", "ux"), ("vAvg", "uy"), ("wAvg", "uz"), ("rhoAvg", "rho")):
        got = read_raw(d, f"{name}.b_dat", np.float64)
        assert rel_err(got, sums[k]) <= TOL64, name


def _h5(path, dataset, tmp):
    out = os.path.join(tmp, "h5.bin")
    subprocess.run(["/opt/conda/bin/h5dump", "-d", dataset, "-b", "LE", "-o", out, path],
                   check=True, capture_output=True)
    return np.fromfile(out, dtype=np.float32)


def test_exec_hdf5_output_layout_and_units(gpu_lib, tmp_path):
    nx, ny, nz = 8, 7, 12
    nd = geometry.channel(nx, ny, nz, obstacles=[geometry.sphere(3, 3, 4, 2)])
    p = io_formats.Params(latticeType="D3Q19", numTs=10, tsRepFreq=10, plotFreq=10, uLBM=0.05,
                          omega=1.6, nx=nx, ny=ny, nz=nz, tConvFact=0.002, lConvFact=0.01,
                          pConvFact=3.5, pRefIdx=17, lx_p=0.7)
    d = case_dir(tmp_path, p, nd)
    r = run_exec(d, storage="f64", env={"TLBM_HDF5": "1"})
    assert r.returncode == 0, r.stderr
    h5 = os.path.join(d, "out0.h5")
    listing = subprocess.run(["/opt/conda/bin/h5ls", "-r", h5], capture_output=True, text=True).stdout
    for ds in ("/velo_group/x_velo", "/velo_group/velmag", "/force_group/z_force",
               "/pres_group/presmag"):
        assert ds in listing
    assert "{12, 7, 8}" in listing  # velmag/presmag shaped (Nz, Ny, Nx)
    dump = read_dump(d, 0, np.float64)
    u_conv = 0.002 / 0.01
    assert np.allclose(_h5(h5, "/velo_group/x_velo", d), dump["ux"] / u_conv, rtol=1e-6, atol=1e-9)
    pres = dump["density"] * 3.5
    assert np.allclose(_h5(h5, "/pres_group/presmag", d), pres - pres[17], rtol=1e-5, atol=1e-6)
    xmf = open(os.path.join(d, "data0.xmf")).read()
    assert "out0.h5:/velo_group/x_velo" in xmf and 'Dimensions="12 7 8"' in xmf


def _feq(w, e, rho, ux, uy, uz):
    out = np.empty((rho.size, w.size))
    for s in range(w.size):
        cu = 3.0 * (e[s, 0] * ux + e[s, 1] * uy + e[s, 2] * uz)
        out[:, s] = w[s] * rho * (1.0 + cu + 0.5 * (cu * cu) - 3.0 / 2.0 * (ux * ux + uy * uy + uz * uz))
    return out


def test_exec_restart_from_dump(gpu_lib, tmp_path):
    """Run, make restart.h5 from the dump, restart: the restarted run starts
    from f_eq of the float32 dump values (TLBM_Partition.cpp:481-540)."""
    from test_gpu_parity import _velocities
    nx, ny, nz = 8, 7, 12
    nd = geometry.channel(nx, ny, nz, obstacles=[geometry.sphere(3, 3, 4, 2)])
    kw = dict(latticeType="D3Q19", tsRepFreq=30, uLBM=0.05, omega=1.6, nx=nx, ny=ny, nz=nz)
    p = io_formats.Params(numTs=30, plotFreq=30, **kw)
    d = case_dir(tmp_path, p, nd)
    assert run_exec(d, storage="f64").returncode == 0
    rr = subprocess.run([require(RESTART), "0", "f64"], cwd=d, capture_output=True, text=True)
    assert rr.returncode == 0, rr.stderr
    dump = read_dump(d, 0, np.float64)
    io_formats.write_params(os.path.join(d, "params.lbm"),
                            io_formats.Params(numTs=5, plotFreq=5, restartFlag=1, **kw))
    r = run_exec(d, storage="f64")
    assert r.returncode == 0, r.stderr
    assert "Loading Restart Data" in r.stdout
    o = Oracle(19, nx, ny, nz, nd, omega=1.6, u_lbm=0.05)
    f32 = {k: dump[k].astype(np.float32).astype(np.float64) for k in ("ux", "uy", "uz", "density")}
    o.set_populations(_feq(o.weights(), _velocities(19), f32["density"], f32["ux"], f32["uy"], f32["uz"]))
    o.step(5)
    ref = o.macros()
    # the restarted run numbers its dumps from 0 again, as the reference does
    got = read_dump(d, 0, np.float64)
    for f, k in (("ux", "ux"), ("uy", "uy"), ("uz", "uz"), ("density", "rho")):
        assert rel_err(got[f], ref[k]) <= TOL64, f


def test_exec_reads_the_binary_node_type_sidecar(gpu_lib, tmp_path):
    """The configs' 24x24x96 D3Q15 case written with ndType.bin (tlbm_amd.cases):
    with the text ndType.lbm that it records, and with ndType.lbm removed,
    the run uses the sidecar and matches the reference's 4-rank dumps."""
    from tlbm_amd import cases
    case = [c for c in goldens.load_cases() if c.name.startswith("channel_d3q15_24x24x96_f64")][0]
    d = str(tmp_path)
    p, nd = cases.write("channel", d, parts=4, binary=True, steps=case.params["numTs"],
                        plot_freq=case.params["numTs"])
    for text in ("kept", "removed"):
        if text == "removed":
            os.remove(os.path.join(d, "ndType.lbm"))
        r = run_exec(d, storage="f64")
        assert r.returncode == 0, r.stderr
        assert "not written from this ndType.lbm" not in r.stderr
        got = read_dump(d, 0, np.float64)
        for f, ref in case.out.items():
            assert rel_err(got[f], ref) <= TOL64, (text, f)


def test_exec_stops_when_the_state_is_not_finite(gpu_lib, tmp_path):
    """Divergence guard end to end: a restart.h5 with one NaN velocity
    (written by the restart tool from a dump) starts the run from a NaN
    equilibrium at that node; the driver's check after step 10 (tsRepFreq)
    finds it, and the driver stops with status 1 naming the step."""
    nx, ny, nz = 8, 7, 12
    nd = geometry.channel(nx, ny, nz, obstacles=[geometry.sphere(3, 3, 4, 2)])
    kw = dict(latticeType="D3Q19", tsRepFreq=10, uLBM=0.05, omega=1.6, nx=nx, ny=ny, nz=nz)
    d = case_dir(tmp_path, io_formats.Params(numTs=10, plotFreq=10, **kw), nd)
    r = run_exec(d, storage="f64")
    assert r.returncode == 0, r.stderr
    ux = read_raw(d, "ux0.b_dat", np.float64)
    ux[nx * ny * 6 + nx * 3 + 4] = np.nan
    ux.tofile(os.path.join(d, "ux0.b_dat"))
    rr = subprocess.run([require(RESTART), "0", "f64"], cwd=d, capture_output=True, text=True)
    assert rr.returncode == 0, rr.stderr
    io_formats.write_params(os.path.join(d, "params.lbm"),
                            io_formats.Params(numTs=40, plotFreq=40, restartFlag=1, **kw))
    for storage in ("f64", "f32", "f16"):
        r = run_exec(d, storage=storage)
        assert r.returncode == 1, (storage, r.stdout, r.stderr)
        assert "time step 10:" in r.stderr and "NaN or infinite" in r.stderr, storage
        assert "Test complete." not in r.stdout


def test_exec_time_average_hdf5(gpu_lib, tmp_path):
    """With TLBM_HDF5=1 the driver also writes timeAvg.h5 / timeAvg.xmf, the
    files python/processTLBM.py:134-186 makes from the time-average sums:
    each sum divided by Num_ts, velocity in physical units, pressure
    relative to node pRefIdx, the speed, and no force group."""
    nx, ny, nz = 8, 7, 12
    nd = geometry.channel(nx, ny, nz, obstacles=[geometry.sphere(3, 3, 4, 2)])
    p = io_formats.Params(latticeType="D3Q19", numTs=20, tsRepFreq=20, plotFreq=20, uLBM=0.05,
                          omega=1.6, nx=nx, ny=ny, nz=nz, timeAvgFlag=1, tConvFact=0.002,
                          lConvFact=0.01, pConvFact=3.5, pRefIdx=17, lx_p=0.7)
    d = case_dir(tmp_path, p, nd)
    r = run_exec(d, storage="f64", env={"TLBM_HDF5": "1"})
    assert r.returncode == 0, r.stderr
    h5 = os.path.join(d, "timeAvg.h5")
    listing = subprocess.run(["/opt/conda/bin/h5ls", "-r", h5], capture_output=True, text=True).stdout
    for ds in ("/velo_group/x_velo", "/velo_group/y_velo", "/velo_group/z_velo",
               "/velo_group/velmag", "/pres_group/presmag"):
        assert ds in listing, ds
    assert "force_group" not in listing
    assert "{12, 7, 8}" in listing
    u_conv = 0.002 / 0.01
    sums = {k: read_raw(d, f"{k}Avg.b_dat", np.float64) for k in ("u", "v", "w", "rho")}
    u = sums["u"] / 20 / u_conv
    v = sums["v"] / 20 / u_conv
    w = sums["w"] / 20 / u_conv
    for ds, want in (("/velo_group/x_velo", u), ("/velo_group/y_velo", v), ("/velo_group/z_velo", w)):
        assert np.allclose(_h5(h5, ds, d), want, rtol=1e-6, atol=1e-9), ds
    pres = sums["rho"] / 20 * 3.5
    assert np.allclose(_h5(h5, "/pres_group/presmag", d), pres - pres[17], rtol=1e-5, atol=1e-6)
    assert np.allclose(_h5(h5, "/velo_group/velmag", d), np.sqrt(u * u + v * v + w * w),
                       rtol=1e-5, atol=1e-9)
    xmf = open(os.path.join(d, "timeAvg.xmf")).read()
    assert "timeAvg.h5:/velo_group/x_velo" in xmf and 'Dimensions="12 7 8"' in xmf
    assert "force" not in xmf

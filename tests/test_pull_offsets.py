"""The stream-collide kernels' relative gathers (PullRel and PullRow in
tlbm_amd/csrc/lbm_kernels.h) restated in numpy with the kernels' 32-bit
unsigned arithmetic, against the chunked layout's own index of each
population's periodic source, for every interior workgroup of the test,
benchmark and configuration lattices (sampled planes of the largest).

Each lane's byte offset must land exactly on the source population, and
every offset plus the direction's scalar offset must lie inside the buffer
descriptor's range (base at chunk c0 - kb, the rest of the state buffer),
so no gather reads out of bounds."""
import numpy as np
import pytest

from lattice_tables import tables

# storage -> (bytes per value, log2 of the chunk: 32 KiB of values)
STORAGES = {"f64": (8, 12), "f32": (4, 13), "f16": (2, 14)}
M32 = np.uint64(0xFFFFFFFF)


def _setup(dims, q, storage):
    nx, ny, nz = dims
    ks, cb = STORAGES[storage]
    ch = 1 << cb
    nxny, n = nx * ny, nx * ny * nz
    kb = (2 * nxny + ch - 1) // ch + 1
    rel = (2 * kb + 2) * ch * q * ks < 2 ** 32
    nchunks = (n + ch - 1) // ch
    fbytes = nchunks * q * ch * ks
    return nx, ny, nz, nxny, n, ks, cb, ch, kb, rel, fbytes


def _nodes(dims, bs, ch, nxny, n, planes):
    """Global indices of the nodes of interior workgroups (pull_rel_ok) in
    the given z planes."""
    nx, ny, nz = dims
    g = np.concatenate([np.arange(z * nxny, (z + 1) * nxny, dtype=np.int64) for z in planes])
    g0 = (g // bs) * bs
    ok = (g0 >= nxny) & (g0 + bs + nxny <= n)
    return g[ok], g0[ok]


def _expected(g, e, dims, ch, cb, q, s):
    """Element offset of population s of node g's source in the chunked
    layout: ((src >> cb) q + s) ch + (src & (ch - 1))."""
    nx, ny, nz = dims
    nxny = nx * ny
    z, r = np.divmod(g, nxny)
    y, x = np.divmod(r, nx)
    src = (x - e[0]) % nx + nx * ((y - e[1]) % ny) + nxny * ((z - e[2]) % nz)
    return ((src >> cb) * q + s) * ch + (src & (ch - 1))


def _u32(v):
    return np.asarray(v, dtype=np.int64).astype(np.uint64) & M32


def _lane_parts(g, g0, dims, ks, cb, ch, kb):
    nx, ny, nz = dims
    nxny = nx * ny
    klog = {8: 3, 4: 2, 2: 1}[ks]
    z, r = np.divmod(g, nxny)
    y, x = np.divmod(r, nx)
    c0 = g0 >> cb
    xs = [_u32(np.where(x + 1 == nx, 0, x + 1) << klog), _u32(x << klog),
          _u32(np.where(x == 0, nx - 1, x - 1) << klog)]
    ys = [_u32((np.where(y + 1 == ny, 0, y + 1) * nx) << klog), _u32((y * nx) << klog),
          _u32((np.where(y == 0, ny - 1, y - 1) * nx) << klog)]
    zr = _u32(g - r - (c0 - kb) * ch)  # zplane - (c0 - kb) CH, modulo 2^32
    return x, c0, xs, ys, zr, klog


@pytest.mark.parametrize("dims,planes", [
    ((128, 128, 512), (1, 2, 255, 510)),
    ((64, 64, 96), tuple(range(96))),
    ((16, 12, 24), tuple(range(24))),
    ((128, 64, 64), tuple(range(64))),
    ((256, 128, 128), (1, 64, 126)),
    ((1024, 512, 512), (1, 256, 510)),
    ((1024, 512, 64), (1, 32, 62)),
    ((384, 384, 384), (1, 200, 382)),
    ((24, 24, 96), tuple(range(96))),
])
@pytest.mark.parametrize("storage,bs", [("f64", 1024), ("f64", 512), ("f32", 256), ("f16", 256),
                                        ("f16", 512), ("f32", 512)])
@pytest.mark.parametrize("q", [15, 19, 27])
def test_relative_gathers_land_on_the_source(dims, planes, storage, bs, q):
    nx, ny, nz, nxny, n, ks, cb, ch, kb, rel, fbytes = _setup(dims, q, storage)
    if not rel or ch % bs:
        pytest.skip("the lattice takes the generic gather")
    e, _, _, _ = tables(q)
    g, g0 = _nodes(dims, bs, ch, nxny, n, planes)
    x, c0, xs, ys, zr, klog = _lane_parts(g, g0, dims, ks, cb, ch, kb)
    back = (c0 - kb) * q * ch * ks
    num_records = np.minimum(fbytes - back, 0xFFFFFFFF)  # make_rsrc's range
    row_path = ch % nx == 0
    mult = (q - 1) * ch * ks
    zs = [_u32(((zr.astype(np.int64) + nxny) << klog)), _u32(zr.astype(np.int64) << klog),
          _u32(((zr.astype(np.int64) - nxny) << klog))]
    for s in range(q):
        ex, ey, ez = (int(v) for v in e[s])
        want = _expected(g, e[s], dims, ch, cb, q, s) * ks - back  # bytes from the base
        soff = s * ch * ks
        # PullRel: per population rb = xs + ys + zs, off = rb + (rb >> (CB + log)) mult
        rb = (xs[ex + 1] + ys[ey + 1] + zs[ez + 1]) & M32
        off = (rb + (rb >> np.uint64(cb + klog)) * np.uint64(mult)) & M32
        assert np.array_equal(off.astype(np.int64) + soff, want), ("PullRel", s)
        assert np.all(off.astype(np.int64) + soff + ks <= num_records), ("PullRel range", s)
        if row_path:
            # PullRow: per source row rowb = ys + zs, row = rowb + (rowb >> (CB + log)) mult,
            # then + xs per population
            rowb = (ys[ey + 1] + zs[ez + 1]) & M32
            row = (rowb + (rowb >> np.uint64(cb + klog)) * np.uint64(mult)) & M32
            off2 = (row + xs[ex + 1]) & M32
            assert np.array_equal(off2.astype(np.int64) + soff, want), ("PullRow", s)
            assert np.all(off2.astype(np.int64) + soff + ks <= num_records), ("PullRow range", s)
    # the 24-bit multiply-add's operands fit
    assert mult < 2 ** 24

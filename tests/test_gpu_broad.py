"""Device broad phase and NL2set (icf_broad_phase, icf_nl2set) against the host versions
(icfh_kdop_pairs, icfh_nl2set): identical pair sets in the same (a, b) order, identical
(node, tet) tuples in first-occurrence order, and the closest faces that follow."""
import ctypes

import numpy as np
import pytest

from icflow_amd import _lib
from icflow_amd.gpu_functionality import broad_phase, stride_of
from workloads import Pipeline, make_state, small_specimen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("exclude_shared", [1, 0])
def test_device_broad_phase_and_nl2set_match_host(specimen_consts, exclude_shared):
    col = small_specimen(cells=(6, 5, 4), size=(0.06, 0.05, 0.04), grains=6, seed=14)
    st = make_state(col, specimen_consts, seed=15, platen_dz_frac=0.3)
    p = Pipeline(col, specimen_consts, st)
    gf, lib, ctx = p.gf, p.gf.lib, p.gf.ctx
    nS = len(col.surf)
    host = broad_phase(gf, gf.c_dn, gf.nd_stride, gf.c_ie, gf.el_all_stride, nS, exclude_shared=bool(exclude_shared))
    Q = len(host)
    assert Q > 0
    d_pairs = gf.buf("bp_pairs", 2 * Q, np.int32)
    n = ctypes.c_longlong()
    r = lib.icf_broad_phase(ctx, gf.g_dn.p, gf.nd_stride, gf.g_ie.p, gf.el_all_stride, nS, exclude_shared,
                            d_pairs.p, Q - 1, ctypes.byref(n))
    assert r == _lib.ICF_E_CAPACITY and n.value == Q
    _lib.check(lib.icf_broad_phase(ctx, gf.g_dn.p, gf.nd_stride, gf.g_ie.p, gf.el_all_stride, nS, exclude_shared,
                                   d_pairs.p, Q, ctypes.byref(n)), ctx, "broad_phase")
    assert n.value == Q
    got = d_pairs.download(2 * Q).reshape(-1, 2)
    assert np.array_equal(got, host)
    # narrow phase on a copy, then NL2set on the device and on the host
    d_nl = gf.buf("bp_nl", 2 * Q, np.int32)
    d_nl.upload(host.ravel())
    _lib.check(lib.icf_narrow_phase(ctx, Q, gf.g_dn.p, gf.g_ie.p, d_nl.p, gf.el_all_stride, gf.nd_stride), ctx,
               "narrow")
    masks = d_nl.download(2 * Q)
    cap = int(np.count_nonzero(masks[0::2])) * 8 + 8
    ts = stride_of(cap)
    ref = np.zeros(2 * ts, np.int32)
    nh = ctypes.c_int()
    _lib.check(lib.icfh_nl2set(_lib.ptr(masks), _lib.ptr(host.ravel().copy()), Q, _lib.ptr(gf.c_ie),
                               gf.el_all_stride, _lib.ptr(ref), ts, ts, ctypes.byref(nh)))
    assert nh.value > 0
    d_nl2 = gf.buf("bp_nl2", 4 * ts, np.int32)
    nd = ctypes.c_int()
    _lib.check(lib.icf_nl2set(ctx, d_nl.p, d_pairs.p, Q, gf.g_ie.p, gf.el_all_stride, d_nl2.p, ts, ctypes.byref(nd)),
               ctx, "nl2set")
    assert nd.value == nh.value
    out = d_nl2.download(4 * ts)
    k = nh.value
    assert np.array_equal(out[:k], ref[:k])
    assert np.array_equal(out[ts:ts + k], ref[ts:ts + k])
    small = gf.buf("bp_small", 4 * 64, np.int32)
    r = lib.icf_nl2set(ctx, d_nl.p, d_pairs.p, Q, gf.g_ie.p, gf.el_all_stride, small.p, 1, ctypes.byref(nd))
    assert r == _lib.ICF_E_CAPACITY and nd.value == nh.value
    gf.close()


def test_device_broad_phase_crowded_cells():
    """100 tets on top of one another (distinct node ids) and 20 apart: every one of the 100
    overlaps every other, so one cell holds more own tets than a wave stages at once and more
    pairs than a wave's buffer holds before it flushes.  Same pairs, same order as the host
    grid."""
    from icflow_amd.gpu_functionality import GPUFunctionality
    rng = np.random.default_rng(3)
    unit = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * 0.01
    tets = [unit + 1e-6 * k for k in range(100)]
    tets += [unit + np.array([0.05 * (k + 1), 0.0, 0.0]) + 1e-4 * rng.random((4, 3)) for k in range(20)]
    nT = len(tets)
    nN = 4 * nT
    ns, es = stride_of(nN), stride_of(nT)
    c_dn = np.zeros(18 * ns)
    for t, x in enumerate(tets):
        for i in range(4):
            for k in range(3):
                c_dn[4 * t + i + ns * (15 + k)] = x[i, k]
    c_ie = np.zeros(5 * es, np.int32)
    for t in range(nT):
        for i in range(4):
            c_ie[t + es * i] = 4 * t + i
    gf = GPUFunctionality(0)
    lib, ctx = gf.lib, gf.ctx
    d_dn, d_ie = gf.buf("cr_dn", c_dn.size, np.float64), gf.buf("cr_ie", c_ie.size, np.int32)
    d_dn.upload(c_dn)
    d_ie.upload(c_ie)
    for excl in (0, 1):
        host = broad_phase(gf, c_dn, ns, c_ie, es, nT, exclude_shared=bool(excl))
        Q = len(host)
        assert Q >= 100 * 99 // 2
        d_pairs = gf.buf(f"cr_pairs{excl}", 2 * Q, np.int32)
        n = ctypes.c_longlong()
        _lib.check(lib.icf_broad_phase(ctx, d_dn.p, ns, d_ie.p, es, nT, excl, d_pairs.p, Q, ctypes.byref(n)), ctx,
                   "broad_phase")
        assert n.value == Q
        assert np.array_equal(d_pairs.download(2 * Q).reshape(-1, 2), host)
    gf.close()


def test_device_broad_phase_sparse_scene_coarsens_the_grid():
    """Five clusters of small tets at the corners of a 3 m cube: cells of the largest extent
    (1 cm) would need about 300^3 grid cells, more than the 2^24 the device grid takes, so the
    cells are coarsened; the pairs (within the clusters) are the host grid's, in order."""
    from icflow_amd.gpu_functionality import GPUFunctionality
    rng = np.random.default_rng(5)
    unit = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * 0.01
    corners = [np.array(c, float) for c in ((0, 0, 0), (3, 0, 0), (0, 3, 0), (0, 0, 3), (3, 3, 3))]
    tets = [unit + c + 0.004 * rng.random(3) for c in corners for _ in range(12)]
    nT = len(tets)
    nN = 4 * nT
    ns, es = stride_of(nN), stride_of(nT)
    c_dn = np.zeros(18 * ns)
    for t, x in enumerate(tets):
        for i in range(4):
            for k in range(3):
                c_dn[4 * t + i + ns * (15 + k)] = x[i, k]
    c_ie = np.zeros(5 * es, np.int32)
    for t in range(nT):
        for i in range(4):
            c_ie[t + es * i] = 4 * t + i
    gf = GPUFunctionality(0)
    lib, ctx = gf.lib, gf.ctx
    d_dn, d_ie = gf.buf("sp_dn", c_dn.size, np.float64), gf.buf("sp_ie", c_ie.size, np.int32)
    d_dn.upload(c_dn)
    d_ie.upload(c_ie)
    host = broad_phase(gf, c_dn, ns, c_ie, es, nT, exclude_shared=True)
    Q = len(host)
    assert 0 < Q <= 5 * 12 * 11 // 2
    d_pairs = gf.buf("sp_pairs", 2 * Q, np.int32)
    n = ctypes.c_longlong()
    _lib.check(lib.icf_broad_phase(ctx, d_dn.p, ns, d_ie.p, es, nT, 1, d_pairs.p, Q, ctypes.byref(n)), ctx,
               "broad_phase")
    assert n.value == Q
    assert np.array_equal(d_pairs.download(2 * Q).reshape(-1, 2), host)
    gf.close()


@pytest.mark.parametrize("nT", [1, 2, 3])
def test_device_broad_phase_few_tets(nT):
    """A chain of one to three overlapping tets (each overlaps the next only) on a fresh
    context, the smallest candidate buffers: the pairs are the host grid's, none for one tet."""
    from icflow_amd.gpu_functionality import GPUFunctionality
    unit = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * 0.01
    tets = [unit + np.array([0.004 * t, 0.002 * t, 0.0]) for t in range(nT)]
    nN = 4 * nT
    ns, es = stride_of(nN), stride_of(nT)
    c_dn = np.zeros(18 * ns)
    for t, x in enumerate(tets):
        for i in range(4):
            for k in range(3):
                c_dn[4 * t + i + ns * (15 + k)] = x[i, k]
    c_ie = np.zeros(5 * es, np.int32)
    for t in range(nT):
        for i in range(4):
            c_ie[t + es * i] = 4 * t + i
    gf = GPUFunctionality(0)
    lib, ctx = gf.lib, gf.ctx
    d_dn, d_ie = gf.buf("ft_dn", c_dn.size, np.float64), gf.buf("ft_ie", c_ie.size, np.int32)
    d_dn.upload(c_dn)
    d_ie.upload(c_ie)
    host = broad_phase(gf, c_dn, ns, c_ie, es, nT, exclude_shared=True)
    Q = len(host)
    assert Q == nT - 1
    d_pairs = gf.buf("ft_pairs", 2 * max(Q, 1), np.int32)
    n = ctypes.c_longlong()
    _lib.check(lib.icf_broad_phase(ctx, d_dn.p, ns, d_ie.p, es, nT, 1, d_pairs.p, max(Q, 1), ctypes.byref(n)), ctx,
               "broad_phase")
    assert n.value == Q
    if Q:
        assert np.array_equal(d_pairs.download(2 * Q).reshape(-1, 2), host)
    gf.close()

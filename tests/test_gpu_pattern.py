"""Device pattern build and pcsr packing (icf_build_pattern, icf_pack_pcsr) against the host
restatement of CSRDictionary and TransferPCSR / collisionResponse packing (icfh_csr_build,
icfh_pack_*), bit-exact, on a specimen with CZs, anchored nodes and contact; then a fused
assembly from the device-packed buffers against the oracle."""
import ctypes

import numpy as np
import pytest

from conftest import rel_err
from icflow_amd import _lib
from workloads import H, Pipeline, make_state, oracle_assemble, small_specimen

pytestmark = pytest.mark.gpu


def test_device_pattern_and_packing_match_host(orc, specimen_consts):
    col = small_specimen(cells=(5, 4, 4), size=(0.05, 0.04, 0.04), grains=5, seed=9)
    st = make_state(col, specimen_consts, seed=10, platen_dz_frac=0.3)
    p = Pipeline(col, specimen_consts, st, contact=True)
    gf, lib, ctx = p.gf, p.gf.lib, p.gf.ctx
    assert gf.nImpacts > 0 and col.anchored.any()
    icr_host, cr = gf.pack_icr(p.impacts)
    T, C, P = len(col.elastic), len(col.czs), gf.nImpacts
    es, cs = gf.el_elastic_stride, gf.cz_stride
    alt = np.ascontiguousarray(col.alt, np.int32)

    def dev(a):
        b = gf.buf(f"t{id(a)}", a.size, a.dtype)
        b.upload(a)
        return b

    # unit buffers with only their node ids (slot and row fields scrambled)
    ie = gf.c_ie_pcsr.copy()
    ie[4 * es:24 * es] = -7
    icz = gf.c_icz.copy()
    icz[10 * cs:52 * cs] = -7
    icz[0:cs] = 5                                # flags TransferPCSR resets
    icz[2 * cs:4 * cs] = 5
    icr = icr_host.copy()
    icr[4 * cr:24 * cr] = -7
    d_alt, d_ie, d_icz, d_icr = dev(alt), dev(ie), dev(icz), dev(icr)
    N = gf.csr.N
    d_rows = gf.buf("p_rows", N + 1, np.int32)
    d_cols = gf.buf("p_cols", gf.csr.nnz, np.int32)
    nnz = ctypes.c_int()
    # too small a capacity reports the size
    r = lib.icf_build_pattern(ctx, d_alt.p, len(alt), d_ie.p, T, es, d_icz.p, C, cs, d_icr.p, P, cr, N, d_rows.p,
                              d_cols.p, 1, ctypes.byref(nnz))
    assert r == _lib.ICF_E_CAPACITY and nnz.value == gf.csr.nnz
    _lib.check(lib.icf_build_pattern(ctx, d_alt.p, len(alt), d_ie.p, T, es, d_icz.p, C, cs, d_icr.p, P, cr, N,
                                     d_rows.p, d_cols.p, gf.csr.nnz, ctypes.byref(nnz)), ctx, "build_pattern")
    assert nnz.value == gf.csr.nnz
    assert np.array_equal(d_rows.download(N + 1), gf.csr.rows)
    assert np.array_equal(d_cols.download(gf.csr.nnz), gf.csr.cols)
    _lib.check(lib.icf_pack_pcsr(ctx, d_alt.p, len(alt), d_rows.p, d_cols.p, N, d_ie.p, T, es, d_icz.p, C, cs,
                                 d_icr.p, P, cr), ctx, "pack_pcsr")

    def fields(a, stride, n, f0, f1):      # fields f0..f1-1 of the first n units (padding excluded)
        return a.reshape(-1, stride)[f0:f1, :n]

    assert np.array_equal(fields(d_ie.download(ie.size), es, T, 0, 24), fields(gf.c_ie_pcsr, es, T, 0, 24))
    got_icz = d_icz.download(icz.size)
    assert np.array_equal(fields(got_icz, cs, C, 0, 4), fields(gf.c_icz, cs, C, 0, 4))
    assert np.array_equal(fields(got_icz, cs, C, 10, 52), fields(gf.c_icz, cs, C, 10, 52))
    assert np.array_equal(fields(d_icr.download(icr.size), cr, P, 0, 24), fields(icr_host, cr, P, 0, 24))
    # altIds beyond the declared row count are an index error
    d_ie.upload(ie)
    r = lib.icf_pack_pcsr(ctx, d_alt.p, len(alt), d_rows.p, d_cols.p, 1, d_ie.p, T, es, None, 0, 0, None, 0, 0)
    assert r == _lib.ICF_E_INDEX
    gf.close()


def test_pattern_static_part_kept_and_rechecked(orc, specimen_consts):
    """icf_build_pattern keeps the static part (elements and CZs) between calls and merges the
    impacts in: with impacts, without, with them again, and after one element's node ids
    change in place (same sizes, so only the kept copy of the node ids can tell), each result
    equals the host pattern of the same inputs."""
    from icflow_amd.gpu_functionality import CSR
    col = small_specimen(cells=(5, 4, 4), size=(0.05, 0.04, 0.04), grains=5, seed=9)
    st = make_state(col, specimen_consts, seed=10, platen_dz_frac=0.3)
    p = Pipeline(col, specimen_consts, st, contact=True)
    gf, lib, ctx = p.gf, p.gf.lib, p.gf.ctx
    icr_host, cr = gf.pack_icr(p.impacts)
    T, C, P = len(col.elastic), len(col.czs), gf.nImpacts
    es, cs = gf.el_elastic_stride, gf.cz_stride
    alt = np.ascontiguousarray(col.alt, np.int32)
    assert P > 0

    def dev(name, a):
        b = gf.buf(name, a.size, a.dtype)
        b.upload(a)
        return b

    ie = gf.c_ie_pcsr.copy()
    d_alt, d_ie = dev("c_alt", alt), dev("c_ie", ie)
    d_icz, d_icr = dev("c_icz", gf.c_icz.copy()), dev("c_icr", icr_host.copy())
    N = gf.csr.N
    cap = 4 * gf.csr.nnz
    d_rows, d_cols = gf.buf("c_rows", N + 1, np.int32), gf.buf("c_cols", cap, np.int32)

    def device(nimp):
        nnz = ctypes.c_int()
        _lib.check(lib.icf_build_pattern(ctx, d_alt.p, len(alt), d_ie.p, T, es, d_icz.p, C, cs, d_icr.p, nimp, cr, N,
                                         d_rows.p, d_cols.p, cap, ctypes.byref(nnz)), ctx, "build_pattern")
        return d_rows.download(N + 1), d_cols.download(nnz.value)

    def host(elastic, imp):
        h = CSR(lib).build(alt[elastic], alt[col.czs], alt[imp] if imp is not None else None)
        assert h.N == N
        return h.rows, h.cols

    elastic = np.asarray(col.elastic)
    for nimp, imp in ((P, p.impacts), (0, None), (P, p.impacts)):
        rows, cols = device(nimp)
        hr, hc = host(elastic, imp)
        assert np.array_equal(rows, hr) and np.array_equal(cols, hc), nimp
    # element 0 takes the node ids of the last element: same sizes, different static part
    el2 = elastic.copy()
    el2[0] = el2[T - 1]
    ie2 = ie.copy()
    for k in range(4):
        ie2[0 + es * k] = ie2[T - 1 + es * k]
    assert np.array_equal(ie2.reshape(-1, es)[:4, :T].T, el2)
    d_ie.upload(ie2)
    rows, cols = device(P)
    hr, hc = host(el2, p.impacts)
    assert np.array_equal(rows, hr) and np.array_equal(cols, hc)
    gf.close()


def test_assembly_from_device_packed_buffers(orc, specimen_consts):
    """The fused assembly after icf_pack_pcsr rewrote the slot fields in place: the library's
    cached maps are re-checked against the new fields and the result matches the oracle."""
    col = small_specimen(seed=12)
    st = make_state(col, specimen_consts, seed=13)
    p = Pipeline(col, specimen_consts, st)
    gf, lib, ctx = p.gf, p.gf.lib, p.gf.ctx
    first = p.assemble()
    T, C = len(col.elastic), len(col.czs)
    es, cs = gf.el_elastic_stride, gf.cz_stride
    alt = np.ascontiguousarray(col.alt, np.int32)
    d_alt = gf.buf("alt", alt.size, np.int32)
    d_alt.upload(alt)
    N = gf.csr.N
    d_rows = gf.buf("p_rows", N + 1, np.int32)
    d_cols = gf.buf("p_cols", gf.csr.nnz, np.int32)
    d_rows.upload(gf.csr.rows)
    d_cols.upload(gf.csr.cols)
    _lib.check(lib.icf_pack_pcsr(ctx, d_alt.p, len(alt), d_rows.p, d_cols.p, N, gf.g_ie_pcsr.p, T, es,
                                 gf.g_icz.p, C, cs, None, 0, 0), ctx, "pack_pcsr")
    gf.TransferNodesToGPU(st)
    again = p.assemble()
    for k in ("vals", "rhs", "F"):
        assert np.array_equal(first[k], again[k]), k
    ref = oracle_assemble(orc, p)
    assert rel_err(again["vals"], ref["vals"]) <= 1e-9
    gf.close()

"""Contexts, captured graphs and in-place writes.

 - Two contexts on one device with different constants, their fused calls interleaved
   without waiting: each context's results equal the oracle's under its own constants (the
   constants live in one per-device block that the library reloads, after the device's queued
   work, when the calling context changes).
 - A captured iteration graph is refused (ICF_E_CAPTURE) once it may be stale: after
   icf_pattern_changed, after a larger uncaptured call grew an internal buffer, after a write
   to slot fields, and on another context.  A fresh capture replays again.
 - A buffer written in place by the contact kernels (icf_nl2set rows 0-1, the closest face
   rows 1-3) and used as the impact array of icf_collision_response: the node-force slot map
   built from it follows a second contact set written the same way.
"""
import ctypes

import numpy as np
import pytest

from conftest import rel_err
from icflow_amd import _lib
from workloads import DIST_EPS, H, PENALTY_K, Pipeline, make_state, oracle_assemble, small_specimen

pytestmark = pytest.mark.gpu


def _asm(p):
    gf, lib, ctx = p.gf, p.gf.lib, p.gf.ctx
    _lib.check(lib.icf_assemble_elems_and_czs(
        ctx, gf.g_ie_pcsr.p, len(p.col.elastic), gf.el_elastic_stride, gf.g_dcz.p, gf.g_icz.p, len(p.col.czs),
        gf.cz_stride, gf.g_dn.p, gf.nd_stride, H, gf.g_dvals.p, gf.csr.nnz, gf.g_drhs.p, gf.csr.N, None, None), ctx,
        "icf_assemble_elems_and_czs")


def _results(p):
    gf = p.gf
    vals, rhs = gf.TransferLinearSystemToHost()
    ns = gf.nd_stride
    return dict(vals=vals, rhs=rhs, F=gf.g_dn.download(3 * ns, offset_elems=12 * ns))


def test_two_contexts_with_different_constants(orc, specimen_consts):
    col = small_specimen(cells=(5, 4, 4), size=(0.05, 0.04, 0.04), grains=5, seed=31)
    other = _lib.consts_from_prms(**{**_lib.SPECIMEN_PRMS, "Y": 2.0e9, "sigma_max": 0.6e6, "rho": 1500.0})
    st_a = make_state(col, specimen_consts, seed=32)
    st_b = make_state(col, other, seed=32)
    pa = Pipeline(col, specimen_consts, st_a)
    pb = Pipeline(col, other, st_b)
    for p in (pa, pb):
        p.assemble()                      # buffers, slot maps
    ref = {id(p): oracle_assemble(orc, p, contact=False) for p in (pa, pb)}
    assert rel_err(ref[id(pa)]["vals"], ref[id(pb)]["vals"]) > 0.1      # the constants matter
    ns = pa.gf.nd_stride
    for p in (pa, pb):
        _lib.check(p.gf.lib.icf_set_blocking(p.gf.ctx, 0))
    for rep in range(3):
        for p in (pa, pb):
            p.gf.g_dn.upload(np.zeros(3 * ns), offset_elems=12 * ns)
        for _ in range(2):                # interleaved, nothing waits in between
            _asm(pa)
            _asm(pb)
        for p in (pa, pb):
            _lib.check(p.gf.lib.icf_synchronize(p.gf.ctx), p.gf.ctx, "sync")
        for p in (pa, pb):
            got, want = _results(p), ref[id(p)]
            assert rel_err(got["vals"], want["vals"]) <= 1e-9, rep
            assert rel_err(got["rhs"], want["rhs"]) <= 1e-9, rep
            # node forces add up over the two calls of the repetition
            assert rel_err(got["F"], 2 * want["F"]) <= 1e-9, rep
    for p in (pa, pb):
        p.gf.close()


def test_stale_graphs_are_refused(specimen_consts):
    col = small_specimen(cells=(5, 4, 4), size=(0.05, 0.04, 0.04), grains=5, seed=33)
    st = make_state(col, specimen_consts, seed=34)
    p = Pipeline(col, specimen_consts, st)
    gf, lib, ctx = p.gf, p.gf.lib, p.gf.ctx
    p.assemble()
    _lib.check(lib.icf_set_blocking(ctx, 0))

    def capture():
        _lib.check(lib.icf_graph_begin(ctx), ctx, "begin")
        _asm(p)
        g = ctypes.c_void_p()
        _lib.check(lib.icf_graph_end(ctx, ctypes.byref(g)), ctx, "end")
        return g

    def launch(g):
        return lib.icf_graph_launch(ctx, g)

    _asm(p)
    _lib.check(lib.icf_synchronize(ctx))
    ref = _results(p)

    # fresh capture: replays equal the eager result (node forces accumulate: compare A and b)
    g = capture()
    for _ in range(2):
        _lib.check(launch(g), ctx, "launch")
    _lib.check(lib.icf_synchronize(ctx))
    got = _results(p)
    assert np.array_equal(got["vals"], ref["vals"]) and np.array_equal(got["rhs"], ref["rhs"])

    # 1. icf_pattern_changed
    _lib.check(lib.icf_pattern_changed(ctx), ctx, "pattern_changed")
    assert launch(g) == _lib.ICF_E_CAPTURE
    assert b"stale" in lib.icf_last_error_string(ctx)
    lib.icf_graph_destroy(g)
    _asm(p)                                                  # checks the maps again, uncaptured
    g = capture()
    _lib.check(launch(g), ctx, "launch after recapture")

    # 2. an internal buffer grows: the element-only call with more tets than staged so far
    #    (a second copy of the tets' pcsr array, twice as long)
    T, es = len(col.elastic), gf.el_elastic_stride
    ie2 = np.zeros(28 * 2 * es, np.int32)
    src = gf.c_ie_pcsr.reshape(28, es)
    dst = ie2.reshape(28, 2 * es)
    dst[:, :T] = src[:, :T]
    dst[:, T:2 * T] = src[:, :T]
    d_ie2 = gf.buf("ie2", len(ie2), np.int32)
    d_ie2.upload(ie2)
    A2 = gf.buf("A2", 9 * gf.csr.nnz, np.float64)
    b2 = gf.buf("b2", 3 * gf.csr.N, np.float64)
    _lib.check(lib.icf_elem_force(ctx, d_ie2.p, gf.g_dn.p, H, A2.p, b2.p, 2 * T, 2 * es, gf.nd_stride), ctx,
               "elem_force")
    _lib.check(lib.icf_synchronize(ctx))
    assert launch(g) == _lib.ICF_E_CAPTURE
    lib.icf_graph_destroy(g)

    # 3. a write to the slot fields the maps were built from
    _asm(p)
    g = capture()
    _lib.check(launch(g), ctx, "launch")
    gf.g_ie_pcsr.upload(gf.c_ie_pcsr)                         # icf_h2d over the slots
    assert launch(g) == _lib.ICF_E_CAPTURE
    lib.icf_graph_destroy(g)

    # 4. another context
    _asm(p)
    g = capture()
    q = Pipeline(col, specimen_consts, st)
    r = q.gf.lib.icf_graph_launch(q.gf.ctx, g)
    assert r == _lib.ICF_E_CAPTURE and b"another context" in lib.icf_last_error_string(q.gf.ctx)
    _lib.check(launch(g), ctx, "launch on its own context")
    _lib.check(lib.icf_synchronize(ctx))
    got = _results(p)
    assert np.array_equal(got["vals"], ref["vals"]) and np.array_equal(got["rhs"], ref["rhs"])
    lib.icf_graph_destroy(g)
    q.gf.close()
    gf.close()


def test_captured_copy_into_slot_fields_is_refused(specimen_consts):
    col = small_specimen(cells=(4, 4, 3), size=(0.04, 0.04, 0.03), grains=3, seed=35)
    st = make_state(col, specimen_consts, seed=36, cells_z=3, size_z=0.03)
    p = Pipeline(col, specimen_consts, st)
    gf, lib, ctx = p.gf, p.gf.lib, p.gf.ctx
    p.assemble()
    es = gf.el_elastic_stride
    copy_src = gf.buf("slot_copy", 16 * es, np.int32)
    copy_src.upload(gf.c_ie_pcsr[4 * es:20 * es])
    _lib.check(lib.icf_set_blocking(ctx, 0))
    _lib.check(lib.icf_graph_begin(ctx), ctx, "begin")
    _asm(p)
    r = lib.icf_d2d(ctx, gf.g_ie_pcsr.p, 16 * es, copy_src.p, 0, 64 * es)
    assert r == _lib.ICF_E_CAPTURE
    assert b"cannot be captured" in lib.icf_last_error_string(ctx)
    g = ctypes.c_void_p()
    _lib.check(lib.icf_graph_end(ctx, ctypes.byref(g)), ctx, "end")
    lib.icf_graph_destroy(g)
    _lib.check(lib.icf_set_blocking(ctx, 1))
    gf.close()


def test_contact_rows_written_in_place_reach_the_force_map(orc, specimen_consts):
    """nl2set + closest face write an impact array's node rows in place (the slots and rhs rows
    are then uploaded by the host, rows 4-27 only); collision response over two different
    contact sets in a row must use each set's nodes for the node forces."""
    col = small_specimen(cells=(5, 4, 4), size=(0.05, 0.04, 0.04), grains=5, seed=37)
    outs = []
    gf = None
    for k, dz in enumerate((0.30, 0.45)):
        st = make_state(col, specimen_consts, seed=38, platen_dz_frac=dz)
        p = Pipeline(col, specimen_consts, st, contact=True)
        imp = p.impacts
        assert len(imp) > 0
        if gf is None:
            gf, lib, ctx = p.gf, p.gf.lib, p.gf.ctx
            cr = 4 * 1024
            X = gf.buf("X_icr", 28 * cr, np.int32)
            pairs = np.ascontiguousarray(p.pairs, np.int32)
        else:
            # same context, new node positions and pattern (the second contact set)
            p.gf.close()
            gf.TransferNodesToGPU(st)
            gf.CreateStructure(imp)
            gf.TransferPCSR(st)
            p.gf = gf
            pairs = np.ascontiguousarray(p.pairs, np.int32)
        nT = len(imp)
        assert nT <= cr
        Q = len(pairs)
        d_pairs = gf.buf("X_pairs", 2 * Q, np.int32)
        d_pairs.upload(pairs.ravel())
        d_list = gf.buf("X_list", 2 * Q, np.int32)
        d_list.upload(pairs.ravel())
        _lib.check(lib.icf_narrow_phase(ctx, Q, gf.g_dn.p, gf.g_ie.p, d_list.p, gf.el_all_stride, gf.nd_stride),
                   ctx, "narrow")
        n = ctypes.c_int()
        _lib.check(lib.icf_nl2set(ctx, d_list.p, d_pairs.p, Q, gf.g_ie.p, gf.el_all_stride, X.p, cr,
                                  ctypes.byref(n)), ctx, "nl2set")
        assert n.value == nT
        _lib.check(lib.icf_find_closest_face(ctx, nT, cr, X.p, gf.g_dn.p, gf.g_ie.p, gf.g_ifc.p, gf.nd_stride,
                                             gf.el_all_stride, gf.fc_stride), ctx, "closest")
        rows03 = X.download(4 * cr)
        got_imp = np.stack([rows03[q * cr:q * cr + nT] for q in range(4)], axis=1)
        assert np.array_equal(got_imp, imp)
        icr, cr2 = gf.pack_icr(imp)
        full = np.zeros(28 * cr, np.int32)
        for q in range(28):
            full[q * cr:q * cr + nT] = icr[q * cr2:q * cr2 + nT]
        X.upload(full[4 * cr:], offset_elems=4 * cr)       # slots and rows only
        vals = gf.buf("X_vals", 9 * gf.csr.nnz, np.float64)
        rhs = gf.buf("X_rhs", 3 * gf.csr.N, np.float64)
        vals.upload(np.zeros(9 * gf.csr.nnz))
        rhs.upload(np.zeros(3 * gf.csr.N))
        ns = gf.nd_stride
        gf.g_dn.upload(np.zeros(3 * ns), offset_elems=12 * ns)
        _lib.check(lib.icf_collision_response(ctx, X.p, gf.g_dn.p, H, vals.p, rhs.p, nT, cr, ns, PENALTY_K, DIST_EPS),
                   ctx, "contact")
        F = gf.g_dn.download(3 * ns, offset_elems=12 * ns)
        dn = gf.c_dn.copy()
        dn[12 * ns:15 * ns] = 0
        A = np.zeros(9 * gf.csr.nnz)
        b = np.zeros(3 * gf.csr.N)
        orc.set_constants(specimen_consts)
        orc.collision_response(full, dn, H, A, b, nT, cr, ns, PENALTY_K, DIST_EPS)
        assert rel_err(F, dn[12 * ns:15 * ns]) <= 1e-9, k
        assert rel_err(vals.download(9 * gf.csr.nnz), A) <= 1e-9, k
        outs.append(nT)
    gf.close()

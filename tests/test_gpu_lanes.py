"""Calls on two lanes.  With the context non-blocking the contact-detection calls run beside the
assembly; a whole contact iteration -- restore the in-place candidate lists, narrow phase,
closest face, fused assembly, collision response -- issued without waiting (with and without
overlap, twice in a row), and as one captured graph replayed three times, must equal the
blocking sequence bit for bit (narrow-phase masks, closest faces, BCSR values, rhs rows, node
forces, CZ counts)."""
import ctypes

import numpy as np
import pytest

from icflow_amd import _lib
from workloads import DIST_EPS, H, PENALTY_K, Pipeline, make_state, small_specimen

pytestmark = pytest.mark.gpu


def test_iteration_on_two_lanes_matches_blocking(specimen_consts):
    col = small_specimen(cells=(5, 4, 4), size=(0.05, 0.04, 0.04), grains=5, seed=21)
    st = make_state(col, specimen_consts, seed=22, platen_dz_frac=0.3)
    p = Pipeline(col, specimen_consts, st, contact=True)
    gf, lib, ctx = p.gf, p.gf.lib, p.gf.ctx
    assert gf.nImpacts > 0
    p.assemble()                                      # creates the BCSR, rhs and impact buffers
    Q, ts, ns = len(p.pairs), gf.tet_stride, gf.nd_stride
    itet_src = gf.buf("lane_itet_src", 2 * Q, np.int32)
    itet_src.upload(np.ascontiguousarray(p.pairs, np.int32).ravel())
    itet = gf.buf("lane_itet", 2 * Q, np.int32)
    nl2_host = np.zeros(4 * ts, np.int32)
    nl2_host[:2 * ts] = gf.c_itet[:2 * ts]            # (node, tet) rows of NL2set
    nl2_src = gf.buf("lane_nl2_src", 4 * ts, np.int32)
    nl2_src.upload(nl2_host)
    nl2 = gf.buf("lane_nl2", 4 * ts, np.int32)
    zeroF = np.zeros(3 * ns)

    def step():
        _lib.check(lib.icf_d2d(ctx, itet.p, 0, itet_src.p, 0, 8 * Q), ctx, "d2d")
        _lib.check(lib.icf_d2d(ctx, nl2.p, 0, nl2_src.p, 0, 16 * ts), ctx, "d2d")
        _lib.check(lib.icf_narrow_phase(ctx, Q, gf.g_dn.p, gf.g_ie.p, itet.p, gf.el_all_stride, ns), ctx, "narrow")
        _lib.check(lib.icf_find_closest_face(ctx, gf.nImpacts, ts, nl2.p, gf.g_dn.p, gf.g_ie.p, gf.g_ifc.p, ns,
                                             gf.el_all_stride, gf.fc_stride), ctx, "closest")
        _lib.check(lib.icf_assemble_elems_and_czs(
            ctx, gf.g_ie_pcsr.p, len(col.elastic), gf.el_elastic_stride, gf.g_dcz.p, gf.g_icz.p, len(col.czs),
            gf.cz_stride, gf.g_dn.p, ns, H, gf.g_dvals.p, gf.csr.nnz, gf.g_drhs.p, gf.csr.N, None, None), ctx, "asm")
        _lib.check(lib.icf_collision_response(ctx, gf.g_icr.p, gf.g_dn.p, H, gf.g_dvals.p, gf.g_drhs.p, gf.nImpacts,
                                              gf.cr_stride, ns, PENALTY_K, DIST_EPS), ctx, "contact")

    def results():
        vals, rhs = gf.TransferLinearSystemToHost()
        d, f = ctypes.c_int(), ctypes.c_int()
        _lib.check(lib.icf_cz_counts(ctx, ctypes.byref(d), ctypes.byref(f)), ctx, "counts")
        return dict(itet=itet.download(2 * Q), nl2=nl2.download(4 * ts), vals=vals, rhs=rhs,
                    F=gf.g_dn.download(3 * ns, offset_elems=12 * ns), counts=(d.value, f.value))

    def same(a, b, what):
        for k in a:
            assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), (what, k)

    gf.g_dn.upload(zeroF, offset_elems=12 * ns)
    step()
    ref = results()
    assert np.count_nonzero(ref["itet"][0::2]) > 0
    # the closest faces of the blocking run are those NarrowPhaseCollisionDetection found
    assert np.array_equal(ref["nl2"][:gf.nImpacts], gf.c_itet[:gf.nImpacts])

    for overlap in (1, 0):
        _lib.check(lib.icf_set_overlap(ctx, overlap), ctx, "overlap")
        _lib.check(lib.icf_set_blocking(ctx, 0))
        for sync in (True, False):                    # host reads order themselves after the lanes
            gf.g_dn.upload(zeroF, offset_elems=12 * ns)
            step()
            if sync:
                _lib.check(lib.icf_synchronize(ctx), ctx, "sync")
            same(results(), ref, ("eager", overlap, sync))
        # two iterations back to back: the second one restores the lists while the first one's
        # contact work may still use them (node forces add up over the two; the rest repeats)
        step()
        step()
        got = results()
        got.pop("F")
        same(got, ref, ("back to back", overlap))
        # captured
        _lib.check(lib.icf_graph_begin(ctx), ctx, "begin")
        step()
        g = ctypes.c_void_p()
        _lib.check(lib.icf_graph_end(ctx, ctypes.byref(g)), ctx, "end")
        _lib.check(lib.icf_set_blocking(ctx, 1))
        for rep in range(3):
            gf.g_dn.upload(zeroF, offset_elems=12 * ns)
            gf.g_dvals.zero()
            itet.upload(np.full(2 * Q, -3, np.int32))
            _lib.check(lib.icf_graph_launch(ctx, g), ctx, "launch")
            same(results(), ref, ("graph", overlap, rep))
        lib.icf_graph_destroy(g)
    _lib.check(lib.icf_set_overlap(ctx, 1), ctx, "overlap")
    gf.close()

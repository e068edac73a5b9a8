"""Edge cases of the C ABI on the GPU: empty unit lists (no pairs, no impacts, no CZs, no
elements), an elements-only and a CZs-only fused call against the oracle, and invalid
arguments refused with ICF_E_INVALID before any launch."""
import ctypes

import numpy as np
import pytest

from conftest import rel_err
from icflow_amd import _lib
from workloads import DIST_EPS, H, PENALTY_K, Pipeline, make_state, oracle_assemble, small_specimen

pytestmark = pytest.mark.gpu


def test_empty_lists_are_no_ops(specimen_consts):
    col = small_specimen(cells=(4, 4, 3), size=(0.04, 0.04, 0.03), grains=3, seed=41)
    st = make_state(col, specimen_consts, seed=42, cells_z=3, size_z=0.03)
    p = Pipeline(col, specimen_consts, st)
    gf, lib, ctx = p.gf, p.gf.lib, p.gf.ctx
    ref = p.assemble()
    vals0, rhs0 = gf.TransferLinearSystemToHost()
    ns = gf.nd_stride
    F0 = gf.g_dn.download(3 * ns, offset_elems=12 * ns)
    dummy = gf.buf("edge_dummy", 64, np.int32)
    assert lib.icf_narrow_phase(ctx, 0, gf.g_dn.p, gf.g_ie.p, dummy.p, gf.el_all_stride, ns) == 0
    assert lib.icf_find_closest_face(ctx, 0, 16, dummy.p, gf.g_dn.p, gf.g_ie.p, gf.g_ifc.p, ns, gf.el_all_stride,
                                     gf.fc_stride) == 0
    assert lib.icf_collision_response(ctx, dummy.p, gf.g_dn.p, H, gf.g_dvals.p, gf.g_drhs.p, 0, 16, ns, PENALTY_K,
                                      DIST_EPS) == 0
    n = ctypes.c_longlong(-1)
    assert lib.icf_broad_phase(ctx, gf.g_dn.p, ns, gf.g_ie.p, gf.el_all_stride, 0, 1, dummy.p, 32,
                               ctypes.byref(n)) == 0 and n.value == 0
    nt = ctypes.c_int(-1)
    assert lib.icf_nl2set(ctx, dummy.p, dummy.p, 0, gf.g_ie.p, gf.el_all_stride, dummy.p, 16, ctypes.byref(nt)) == 0
    assert nt.value == 0
    _lib.check(lib.icf_synchronize(ctx), ctx, "sync")
    vals, rhs = gf.TransferLinearSystemToHost()
    assert np.array_equal(vals, vals0) and np.array_equal(rhs, rhs0)
    assert np.array_equal(gf.g_dn.download(3 * ns, offset_elems=12 * ns), F0)
    assert np.array_equal(vals, ref["vals"])
    gf.close()


def test_elements_only_and_czs_only(orc, specimen_consts):
    # one grain: no CZs at all
    col = small_specimen(cells=(5, 4, 3), size=(0.05, 0.04, 0.03), grains=1, seed=43)
    assert len(col.czs) == 0
    st = make_state(col, specimen_consts, seed=44, cells_z=3, size_z=0.03)
    p = Pipeline(col, specimen_consts, st)
    got = p.assemble()
    ref = oracle_assemble(orc, p, contact=False)
    for k in ("vals", "rhs", "F"):
        assert rel_err(got[k], ref[k]) <= 1e-9, k
    p.gf.close()
    # CZs only: the fused call with no elements, against the oracle's CZ pass alone
    col = small_specimen(cells=(5, 4, 3), size=(0.05, 0.04, 0.03), grains=4, seed=45)
    st = make_state(col, specimen_consts, seed=46, cells_z=3, size_z=0.03)
    p = Pipeline(col, specimen_consts, st)
    gf, lib, ctx = p.gf, p.gf.lib, p.gf.ctx
    p.assemble()                                     # buffers
    gf.TransferNodesToGPU(st)
    gf.TransferPCSR(st)
    C, ns = len(col.czs), gf.nd_stride
    _lib.check(lib.icf_assemble_elems_and_czs(
        ctx, gf.g_ie_pcsr.p, 0, gf.el_elastic_stride, gf.g_dcz.p, gf.g_icz.p, C, gf.cz_stride, gf.g_dn.p, ns, H,
        gf.g_dvals.p, gf.csr.nnz, gf.g_drhs.p, gf.csr.N, None, None), ctx, "cz only")
    vals, rhs = gf.TransferLinearSystemToHost()
    orc.set_constants(specimen_consts)
    A, b = np.zeros(gf.csr.nnz * 9), np.zeros(gf.csr.N * 3)
    dn, icz, dcz = gf.c_dn.copy(), gf.c_icz.copy(), gf.c_dcz.copy()
    orc.cz_force(dcz, icz, dn, A, b, H, C, gf.cz_stride, ns)
    assert np.abs(A).max() > 0
    assert rel_err(vals, A) <= 1e-9 and rel_err(rhs, b) <= 1e-9
    assert np.array_equal(gf.g_icz.download(4 * gf.cz_stride)[gf.cz_stride:], icz[gf.cz_stride:4 * gf.cz_stride])
    gf.close()


def test_invalid_arguments_are_refused(specimen_consts):
    col = small_specimen(cells=(4, 4, 3), size=(0.04, 0.04, 0.03), grains=3, seed=47)
    st = make_state(col, specimen_consts, seed=48, cells_z=3, size_z=0.03)
    p = Pipeline(col, specimen_consts, st)
    gf, lib, ctx = p.gf, p.gf.lib, p.gf.ctx
    p.assemble()
    ns, es = gf.nd_stride, gf.el_elastic_stride
    E = _lib.ICF_E_INVALID
    assert lib.icf_narrow_phase(ctx, -1, gf.g_dn.p, gf.g_ie.p, gf.g_ie.p, gf.el_all_stride, ns) == E
    assert lib.icf_narrow_phase(ctx, 4, None, gf.g_ie.p, gf.g_ie.p, gf.el_all_stride, ns) == E
    assert lib.icf_elem_force(ctx, gf.g_ie_pcsr.p, gf.g_dn.p, H, gf.g_dvals.p, gf.g_drhs.p, -3, es, ns) == E
    assert lib.icf_assemble_elems_and_czs(ctx, gf.g_ie_pcsr.p, len(col.elastic), es, None, None, 5, gf.cz_stride,
                                          gf.g_dn.p, ns, H, gf.g_dvals.p, gf.csr.nnz, gf.g_drhs.p, gf.csr.N,
                                          None, None) == E          # CZs without their buffers
    assert lib.icf_assemble_elems_and_czs(ctx, gf.g_ie_pcsr.p, len(col.elastic), es, None, None, 0, 0, gf.g_dn.p,
                                          ns, 0.0, gf.g_dvals.p, gf.csr.nnz, gf.g_drhs.p, gf.csr.N,
                                          None, None) == E          # h = 0
    assert lib.icf_collision_response(ctx, None, gf.g_dn.p, H, gf.g_dvals.p, gf.g_drhs.p, 3, 16, ns, PENALTY_K,
                                      DIST_EPS) == E
    assert b"invalid" in lib.icf_last_error_string(ctx)
    # the context still works after refusals
    again = p.assemble()
    assert np.isfinite(again["vals"]).all()
    gf.close()

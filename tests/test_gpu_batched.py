"""Batched specimens through icf_assemble_batched: several independent cubes (lognormal sizes,
their own grains) back to back in one set of buffers.  The batched call must equal the fused
call over the whole lists bit for bit (BCSR values, rhs rows, node forces, CZ state), count
damaged and failed CZs per specimen as the host does from the tentative flags, refuse a
worklist whose block ranges do not match the units' slots, and replay from a captured graph."""
import ctypes

import numpy as np
import pytest

from icflow_amd import _lib, specimens
from icflow_amd.gpu_functionality import GPUFunctionality
from workloads import H

pytestmark = pytest.mark.gpu


def _setup(consts, precision):
    col = specimens.lognormal_cubes(5, 3, median_tets=1500, min_tets=300, max_tets=6000, sigma=0.9, grains=(3, 8),
                                    size=0.01)
    st = specimens.synthetic_state(col, consts, 8, equal_gp=True)
    gf = GPUFunctionality(0)
    gf.SetConstants(consts)
    gf.set_precision(precision)
    gf.TransferStaticDataToDevice(col)
    gf.TransferNodesToGPU(st)
    gf.CreateStructure(None)
    gf.TransferPCSR(st)
    tet, cz, row = specimens.specimen_offsets(col)
    offs = [tet, cz, row, gf.csr.rows[row].astype(np.int32)]
    return col, st, gf, offs


def _outputs(gf):
    vals, rhs = gf.TransferLinearSystemToHost()
    out = gf.TransferUpdatedStateToHost()
    return dict(vals=vals, rhs=rhs, F=out["F"], dcz=out["dcz"], icz=out["icz"])


def _same(a, b, what):
    for k in a:
        assert np.array_equal(a[k], b[k]), (what, k)


@pytest.mark.parametrize("precision", [_lib.ICF_FP64, _lib.ICF_MIXED], ids=["fp64", "mixed"])
def test_batched_equals_fused_call(specimen_consts, precision):
    col, st, gf, offs = _setup(specimen_consts, precision)
    n = len(offs[0]) - 1
    assert n == 5 and np.all(np.diff(offs[0]) > 0) and np.all(np.diff(offs[1]) > 0)
    dmg_all, fld_all = gf.AssembleElemsAndCZs(H)
    ref = _outputs(gf)
    gf.TransferNodesToGPU(st)                       # F back to zero, the same inputs
    gf.TransferPCSR(st)
    dmg, fld = gf.AssembleBatched(H, *offs)
    got = _outputs(gf)
    _same(got, ref, "batched")
    # per-specimen sums of the tentative flags, as the host would take them
    cs, C = gf.cz_stride, len(col.czs)
    icz = got["icz"]
    tdmg, tfld = icz[2 * cs:2 * cs + C], icz[3 * cs:3 * cs + C]
    for s in range(n):
        a, b = offs[1][s], offs[1][s + 1]
        assert dmg[s] == tdmg[a:b].sum() and fld[s] == tfld[a:b].sum(), s
    assert dmg.sum() == dmg_all and fld.sum() == fld_all
    gf.close()


def test_batched_layout_check_counts_and_capture(specimen_consts):
    col, st, gf, offs = _setup(specimen_consts, _lib.ICF_MIXED)
    lib, ctx = gf.lib, gf.ctx
    n = len(offs[0]) - 1
    dmg, fld = gf.AssembleBatched(H, *offs)
    ref = _outputs(gf)
    # a block range shifted by one block: specimen 1's units write a block of specimen 0
    bad = [a.copy() for a in offs]
    bad[3][1] += 1
    wl = _lib.IcfWorklist(n, *[_lib.ptr(a) for a in bad])
    r = lib.icf_assemble_batched(ctx, ctypes.byref(wl), gf.g_ie_pcsr.p, gf.el_elastic_stride, gf.g_dcz.p,
                                 gf.g_icz.p, gf.cz_stride, gf.g_dn.p, gf.nd_stride, H, gf.g_dvals.p, gf.g_drhs.p,
                                 None, None)
    assert r == _lib.ICF_E_INDEX
    # decreasing offsets
    bad = [a.copy() for a in offs]
    bad[0][2] = bad[0][1] - 1
    wl = _lib.IcfWorklist(n, *[_lib.ptr(a) for a in bad])
    r = lib.icf_assemble_batched(ctx, ctypes.byref(wl), gf.g_ie_pcsr.p, gf.el_elastic_stride, gf.g_dcz.p,
                                 gf.g_icz.p, gf.cz_stride, gf.g_dn.p, gf.nd_stride, H, gf.g_dvals.p, gf.g_drhs.p,
                                 None, None)
    assert r == _lib.ICF_E_INVALID
    # non-blocking, then the counts of the last batched call
    zeroF = np.zeros(3 * gf.nd_stride)
    gf.g_dn.upload(zeroF, offset_elems=12 * gf.nd_stride)
    _lib.check(lib.icf_set_blocking(ctx, 0))
    wl = _lib.IcfWorklist(n, *[_lib.ptr(a) for a in offs])

    def call():
        _lib.check(lib.icf_assemble_batched(ctx, ctypes.byref(wl), gf.g_ie_pcsr.p, gf.el_elastic_stride,
                                            gf.g_dcz.p, gf.g_icz.p, gf.cz_stride, gf.g_dn.p, gf.nd_stride, H,
                                            gf.g_dvals.p, gf.g_drhs.p, None, None), ctx, "icf_assemble_batched")

    call()
    d2, f2 = np.zeros(n, np.int32), np.zeros(n, np.int32)
    _lib.check(lib.icf_batch_cz_counts(ctx, _lib.ptr(d2), _lib.ptr(f2)), ctx, "icf_batch_cz_counts")
    assert np.array_equal(d2, dmg) and np.array_equal(f2, fld)
    _same(_outputs(gf), ref, "non-blocking")
    # captured once, replayed with the node forces cleared in between
    _lib.check(lib.icf_graph_begin(ctx), ctx, "begin")
    call()
    g = ctypes.c_void_p()
    _lib.check(lib.icf_graph_end(ctx, ctypes.byref(g)), ctx, "end")
    _lib.check(lib.icf_set_blocking(ctx, 1))
    for rep in range(3):
        gf.g_dn.upload(zeroF, offset_elems=12 * gf.nd_stride)
        gf.g_dvals.zero()
        _lib.check(lib.icf_graph_launch(ctx, g), ctx, "launch")
        _same(_outputs(gf), ref, ("graph", rep))
        _lib.check(lib.icf_batch_cz_counts(ctx, _lib.ptr(d2), _lib.ptr(f2)), ctx, "icf_batch_cz_counts")
        assert np.array_equal(d2, dmg) and np.array_equal(f2, fld)
    lib.icf_graph_destroy(g)
    gf.close()

"""Parity at the full size of the compression-with-contact workload that bench.py times: 40^3
hex x 6 tets (384,000), 500 grains, 93,101 CZs, two rigid platens pressed 0.3 cell in,
1.6M candidate pairs and ~61k impacts.

- Narrow-phase masks and closest faces bit-exact against the oracle on every pair.
- The fused FP64 iteration (elements, CZs, contact) against the oracle: <= 1e-9 normwise per
  assembled array; CZ flags bit-exact.
- The fused call equals the reference's separate launches (icf_elem_force, icf_cz_force into
  zeroed arrays) bit for bit, in FP64 and in mixed mode: both add every block's, row's and
  node's contributions in the same list order, through different kernels (staged chunks and
  per-node tet sums in the fused call, accumulate-mode reductions in the separate ones).
- Mixed against FP64 on the same inputs: <= 1e-5 normwise for the BCSR values, residuals
  identical."""
import ctypes

import numpy as np
import pytest

from conftest import rel_err
from icflow_amd import _lib, specimens
from workloads import DIST_EPS, H, PENALTY_K, Pipeline, oracle_assemble

pytestmark = pytest.mark.gpu

CELLS, GRAINS, SEED, PLATEN_FRAC, SIZE = 40, 500, 1, 0.3, 0.1


@pytest.fixture(scope="module")
def workload():
    consts = _lib.consts_from_prms(**_lib.SPECIMEN_PRMS)
    col = specimens.make_specimen((CELLS,) * 3, (SIZE,) * 3, GRAINS, SEED, platens=(32, 32)).finalize()
    st = specimens.synthetic_state(col, consts, SEED, platen_dz=PLATEN_FRAC * SIZE / CELLS, equal_gp=True)
    assert len(col.elastic) == 384000
    return consts, col, st


def _separate(p):
    """The reference's AssembleElemsAndCZs launches (GPU_Functionality.cs:439-479) on zeroed
    arrays, then the collision response."""
    gf, lib, col = p.gf, p.gf.lib, p.col
    nnz, N, ns = gf.csr.nnz, gf.csr.N, gf.nd_stride
    gf.g_dvals.zero()
    gf.g_drhs.zero()
    gf.g_dn.upload(np.zeros(3 * ns), offset_elems=12 * ns)
    _lib.check(lib.icf_elem_force(gf.ctx, gf.g_ie_pcsr.p, gf.g_dn.p, H, gf.g_dvals.p, gf.g_drhs.p, len(col.elastic),
                                  gf.el_elastic_stride, ns), gf.ctx, "icf_elem_force")
    _lib.check(lib.icf_cz_force(gf.ctx, gf.g_dcz.p, gf.g_icz.p, gf.g_dn.p, gf.g_dvals.p, gf.g_drhs.p, H,
                                len(col.czs), gf.cz_stride, ns), gf.ctx, "icf_cz_force")
    _lib.check(lib.icf_collision_response(gf.ctx, gf.g_icr.p, gf.g_dn.p, H, gf.g_dvals.p, gf.g_drhs.p, gf.nImpacts,
                                          gf.cr_stride, ns, PENALTY_K, DIST_EPS), gf.ctx, "icf_collision_response")
    vals, rhs = gf.TransferLinearSystemToHost()
    return dict(vals=vals, rhs=rhs, F=gf.g_dn.download(3 * ns, offset_elems=12 * ns))


def test_fullsize_fp64_matches_oracle(workload, orc):
    consts, col, st = workload
    p = Pipeline(col, consts, st, contact=True)
    gf = p.gf
    Q, nT = len(p.pairs), gf.nImpacts
    print(f"full size: {len(col.elastic)} tets, {len(col.czs)} CZs, {Q} pairs, {nT} impacts", flush=True)
    assert Q > 1_000_000 and nT > 10_000
    # narrow phase: every mask bit-exact
    nl = p.pairs.astype(np.int32).ravel().copy()
    orc.narrow_phase(Q, gf.c_dn, gf.c_ie, nl, gf.el_all_stride, gf.nd_stride)
    assert np.array_equal(nl[0::2], gf.masks)
    # closest faces of every (node, tet) tuple bit-exact
    ts = gf.tet_stride
    masks2 = np.zeros(2 * Q, np.int32)
    masks2[0::2] = gf.masks
    out = np.zeros(4 * ts, np.int32)
    n = ctypes.c_int()
    _lib.check(gf.lib.icfh_nl2set(_lib.ptr(masks2), _lib.ptr(p.pairs.astype(np.int32).ravel().copy()), Q,
                                  _lib.ptr(gf.c_ie), gf.el_all_stride, _lib.ptr(out), ts, ts, ctypes.byref(n)))
    assert n.value == nT
    orc.find_closest_face(nT, ts, out, gf.c_dn, gf.c_ie, gf.c_ifc, gf.nd_stride, gf.el_all_stride, gf.fc_stride)
    for k in range(4):
        assert np.array_equal(out[k * ts:k * ts + nT], gf.c_itet[k * ts:k * ts + nT]), k
    # the fused iteration with contact against the oracle
    got = p.assemble()
    ref = oracle_assemble(orc, p)
    errs = {k: rel_err(got[k], ref[k]) for k in ("vals", "rhs", "F")}
    print("normwise rel errors:", errs, flush=True)
    assert max(errs.values()) <= 1e-9
    C, cs = len(col.czs), gf.cz_stride
    for f in (1, 2, 3):   # tentative contact, damaged, failed
        assert np.array_equal(got["icz"][f * cs:f * cs + C], ref["icz"][f * cs:f * cs + C]), f
    gf.close()


@pytest.mark.parametrize("precision", [_lib.ICF_FP64, _lib.ICF_MIXED], ids=["fp64", "mixed"])
def test_fullsize_fused_equals_separate_launches(workload, precision):
    consts, col, st = workload
    p = Pipeline(col, consts, st, precision=precision, contact=True)
    fused = p.assemble()
    sep = _separate(p)
    for k in ("vals", "rhs", "F"):
        assert np.array_equal(fused[k], sep[k]), k
    p.gf.close()


def test_fullsize_mixed_against_fp64(workload):
    consts, col, st = workload
    p = Pipeline(col, consts, st, precision=_lib.ICF_FP64, contact=True)
    a = p.assemble()
    p.gf.set_precision(_lib.ICF_MIXED)
    p.gf.TransferNodesToGPU(st)
    p.gf.TransferPCSR(st)
    b = p.assemble()
    err = rel_err(b["vals"], a["vals"])
    print("mixed vs fp64 values:", err, flush=True)
    assert 0 < err <= 1e-5
    assert np.array_equal(a["rhs"], b["rhs"])
    assert np.array_equal(a["F"], b["F"])
    p.gf.close()

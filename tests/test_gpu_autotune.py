"""Autotuner of the BCSR-value reductions' chunk shape (icf_set_autotune, icf_get_autotune,
icf_set_chunk_keys): key limit and staging size.  The shape changes how the keys are chunked
over workgroups, never the order in which a key's contributions are added: the fused iteration
with contact gives the same bits (BCSR values, rhs rows, node forces, CZ flags) under every
candidate of every role, and after tuning; so do the reference's separate launches
(accumulate maps)."""
import numpy as np
import pytest

from icflow_amd import _lib
from workloads import DIST_EPS, H, PENALTY_K, Pipeline, make_state, small_specimen

pytestmark = pytest.mark.gpu

CANDS = tuple((k, s) for k in (64, 128, 256, 512) for s in (2, 3))


def _iteration(p):
    gf = p.gf
    gf.g_dn.upload(np.zeros(3 * gf.nd_stride), offset_elems=12 * gf.nd_stride)
    return p.assemble()


def _separate(p):
    """The reference's separate launches (icf_elem_force, icf_cz_force into zeroed arrays, then
    the collision response): accumulate-map reductions."""
    gf, lib, col = p.gf, p.gf.lib, p.col
    ns = gf.nd_stride
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
    out = gf.TransferUpdatedStateToHost()
    return dict(vals=vals, rhs=rhs, F=out["F"], icz=out["icz"])


def _same(a, b, what):
    for k in ("vals", "rhs", "F", "icz"):
        assert np.array_equal(a[k], b[k]), (what, k)


@pytest.mark.parametrize("precision", [_lib.ICF_FP64, _lib.ICF_MIXED], ids=["fp64", "mixed"])
def test_every_chunk_shape_gives_the_same_bits(specimen_consts, precision):
    col = small_specimen(cells=(10, 8, 6), size=(0.1, 0.08, 0.06), grains=12, seed=31)
    st = make_state(col, specimen_consts, seed=32, platen_dz_frac=0.3, cells_z=6, size_z=0.06)
    p = Pipeline(col, specimen_consts, st, precision=precision, contact=True)
    gf = p.gf
    assert gf.nImpacts > 0
    fp64 = precision == _lib.ICF_FP64
    stage = "fp64" if fp64 else "fp32"
    state = gf.autotune_state()
    for r, (shape, tuning, _) in state.items():          # defaults before tuning
        assert shape == ((256, 3) if r.endswith("fp64") else (256, 2)) and not tuning, (r, shape)
    ref = _iteration(p)
    for keys, stg in CANDS:
        for acc in (0, 1):
            gf.set_chunk_keys(acc, fp64, keys, stg)
        assert gf.autotune_state()[f"overwrite,{stage}"][0] == (keys, stg)
        _same(_iteration(p), ref, (keys, stg))
    # the separate single-kernel launches (accumulate maps) under every shape
    for keys, stg in CANDS:
        gf.set_chunk_keys(1, fp64, keys, stg)
        _same(_separate(p), ref, ("separate", keys, stg))
    # mixed choices (overwrite and accumulate maps tuned apart); stage 0 keeps the staging
    gf.set_chunk_keys(0, fp64, 64, 2)
    gf.set_chunk_keys(1, fp64, 512, 0)
    assert gf.autotune_state()[f"accumulate,{stage}"][0] == (512, 3)
    _same(_iteration(p), ref, "64/512")
    # tuning: every candidate of the two roles this iteration uses is timed, then one is kept
    gf.set_autotune(2)
    used = [f"overwrite,{stage}", f"accumulate,{stage}"]
    for it in range(2 * len(CANDS) + 1):
        _same(_iteration(p), ref, ("tuning", it))
    state = gf.autotune_state()
    for r in used:
        shape, tuning, best = state[r]
        assert not tuning, (r, state[r])
        assert shape in CANDS
        assert all(ms > 0 for ms in best.values()), (r, best)
        assert best[shape] == min(best.values()), (r, best)
    _same(_iteration(p), ref, "tuned")
    # bad shapes are refused; rounds = 0 stops tuning and keeps a timed candidate
    assert gf.lib.icf_set_chunk_keys(gf.ctx, 0, 0, 100, 2) == _lib.ICF_E_INVALID
    assert gf.lib.icf_set_chunk_keys(gf.ctx, 0, 0, 256, 4) == _lib.ICF_E_INVALID
    assert gf.lib.icf_set_autotune(gf.ctx, -1) == _lib.ICF_E_INVALID
    gf.set_autotune(1)
    _iteration(p)
    gf.set_autotune(0)
    state = gf.autotune_state()
    for r in used:
        shape, tuning, best = state[r]
        assert not tuning and best[shape] > 0, (r, state[r])
    _same(_iteration(p), ref, "stopped")
    # the tuned shapes carry over to a new context through JSON: a different shape per role,
    # so that a role taken for another shows; the new context assembles the same bits with them
    import json
    for i, (keys, stg) in enumerate(((64, 2), (512, 3), (128, 3), (256, 2))):
        gf.set_chunk_keys(i >> 1, i & 1, keys, stg)
    saved = json.loads(json.dumps(gf.tuned_shapes()))
    assert len({tuple(saved[r]) for r in gf.ROLES}) == 4
    gf.close()
    p2 = Pipeline(col, specimen_consts, st, precision=precision, contact=True)
    p2.gf.use_shapes(saved)
    for r, (shape, tuning, _) in p2.gf.autotune_state().items():
        assert list(shape) == saved[r] and not tuning, (r, shape)
    _same(_iteration(p2), ref, "restored shapes")
    p2.gf.close()


SHAPES = {"phase1": (64, 128, 256), "narrow": (64, 128, 256), "closest": (4, 8, 16), "impact": (64, 128, 256),
          "order": (0, 1)}


@pytest.mark.parametrize("precision", [_lib.ICF_FP64, _lib.ICF_MIXED], ids=["fp64", "mixed"])
def test_every_launch_shape_gives_the_same_bits(specimen_consts, precision):
    """The other tuned roles (icf_set_shape / icf_get_shape): phase 1's, the narrow phase's and
    the impacts' workgroup sizes, the closest face's lanes per tuple and the order of the
    fused value launch's phases.  The whole iteration
    with contact (narrow phase and closest face included, their outputs compared) gives the
    same bits under every value of every role and after tuning them; tuning times every
    candidate of each role this iteration launches and keeps the fastest."""
    col = small_specimen(cells=(10, 8, 6), size=(0.1, 0.08, 0.06), grains=12, seed=31)
    st = make_state(col, specimen_consts, seed=32, platen_dz_frac=0.3, cells_z=6, size_z=0.06)
    p = Pipeline(col, specimen_consts, st, precision=precision, contact=True)
    gf = p.gf
    assert gf.nImpacts > 0
    defaults = {"phase1": 256, "narrow": 256, "closest": 8, "impact": 256, "order": 1}
    for r, (v, tuning, _) in gf.shape_state().items():
        assert v == defaults[r] and not tuning, (r, v)

    def contact():
        imp = gf.NarrowPhaseCollisionDetection(p.pairs)
        return gf.masks.copy(), imp.copy()

    ref = _iteration(p)
    ref_masks, ref_imp = contact()
    assert np.array_equal(ref_imp, p.impacts)
    for role, cands in SHAPES.items():
        for v in cands:
            gf.set_shape(role, v)
            assert gf.shape_state()[role][0] == v
            masks, imp = contact()
            assert np.array_equal(masks, ref_masks) and np.array_equal(imp, ref_imp), (role, v)
            _same(_iteration(p), ref, (role, v))
        gf.set_shape(role, defaults[role])
    for role, bad in (("phase1", 32), ("narrow", 512), ("closest", 2), ("impact", 100), ("order", 2)):
        assert gf.lib.icf_set_shape(gf.ctx, gf.SHAPE_ROLES.index(role), bad) == _lib.ICF_E_INVALID
    assert gf.lib.icf_set_shape(gf.ctx, 7, 64) == _lib.ICF_E_INVALID
    # tuning every role at once: each iteration launches all four shape roles
    gf.set_autotune(2)
    for it in range(2 * len(CANDS) + 1):
        masks, imp = contact()
        assert np.array_equal(masks, ref_masks) and np.array_equal(imp, ref_imp), ("tuning", it)
        _same(_iteration(p), ref, ("tuning", it))
    for role, (v, tuning, best) in gf.shape_state().items():
        assert v in SHAPES[role] and not tuning, (role, v, tuning)
        assert all(ms > 0 for ms in best.values()), (role, best)
        assert best[v] == min(best.values()), (role, best)
    _same(_iteration(p), ref, "tuned")
    gf.set_autotune(0)
    gf.close()

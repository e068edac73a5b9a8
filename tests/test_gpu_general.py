"""The general-material path: an elasticity matrix that is not of the isotropic form (two
diagonal entries scaled) makes the library stage whole rotation + gradient records and
evaluate R K0 R^T blocks (tet_block_gen) instead of the isotropic closed form.  The fused
iteration with contact against the oracle, in FP64 (1e-9) and mixed (1e-5 values, FP64
residuals)."""
import ctypes

import pytest

from conftest import rel_err
from icflow_amd import _lib
from workloads import Pipeline, make_state, oracle_assemble, small_specimen

pytestmark = pytest.mark.gpu


def _anisotropic(consts):
    c = _lib.IcfConsts()
    ctypes.memmove(ctypes.byref(c), ctypes.byref(consts), ctypes.sizeof(c))
    c.E[0] *= 1.10          # E[0][0]
    c.E[7] *= 1.05          # E[1][1]; still symmetric
    return c


@pytest.mark.parametrize("precision", [_lib.ICF_FP64, _lib.ICF_MIXED], ids=["fp64", "mixed"])
def test_general_material_path(orc, specimen_consts, precision):
    consts = _anisotropic(specimen_consts)
    col = small_specimen(cells=(5, 4, 4), size=(0.05, 0.04, 0.04), grains=5, seed=51)
    st = make_state(col, consts, seed=52, platen_dz_frac=0.3)
    p = Pipeline(col, consts, st, precision=precision, contact=True)
    assert p.gf.nImpacts > 0
    got = p.assemble()
    ref = oracle_assemble(orc, p)
    errs = {k: rel_err(got[k], ref[k]) for k in ("vals", "rhs", "F")}
    print("general path errors", errs, flush=True)
    tol_vals = 1e-9 if precision == _lib.ICF_FP64 else 1e-5
    assert errs["vals"] <= tol_vals and errs["rhs"] <= 1e-9 and errs["F"] <= 1e-9
    p.gf.close()

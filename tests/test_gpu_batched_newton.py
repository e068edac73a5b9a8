"""The batched Newton driver (BatchedImplicitModel: one icf_assemble_batched per iteration
over every specimen, per-specimen SolveDouble3 on host threads, per-specimen convergence)
against each specimen run alone through ImplicitModel (icf_assemble_elems_and_czs and the
same solve): displacements and failed CZs bit for bit, FP64 and mixed, through steps that
damage and fail CZs."""
import numpy as np
import pytest

from icflow_amd import _lib, newton, specimens

pytestmark = pytest.mark.gpu

CUBES = dict(median_tets=1500, min_tets=600, max_tets=3000, grains=(4, 12))
STEPS, STRAIN, RAMP = 6, 5e-3, 4


@pytest.mark.parametrize("mode", [_lib.ICF_FP64, _lib.ICF_MIXED], ids=["fp64", "mixed"])
def test_batched_equals_specimens_alone(specimen_consts, mode):
    col, presc, ids = newton.tension(specimens.lognormal_cubes(3, 11, **CUBES), STRAIN, RAMP)
    m = newton.BatchedImplicitModel(col, specimen_consts, precision=mode, step=newton.StepParams(h=1e-4),
                                    prescribed=presc, prescribed_nodes=ids, threads=3)
    try:
        for _ in range(STEPS):
            m.step()
        u = m.displacements()
        failed = m.failed_set()
        spec_of_cz = col.node_mesh[np.asarray(col.czs)[:, 0]]
        its = np.array([i for i, _ in m.spec_records])
        damaged = [r.damaged for r in m.records]
    finally:
        m.close()
    print(f"batched: {len(col.elastic)} tets, {len(col.czs)} CZs, {len(failed)} failed, damaged per step {damaged}, "
          f"iterations per step and specimen {its.tolist()}", flush=True)
    assert len(failed) > 0 or max(damaged) > 0, "the loading should damage CZs"
    for k, ck in enumerate(specimens.lognormal_specimens(3, 11, **CUBES)):
        ck, pk, idk = newton.tension(ck, STRAIN, RAMP)
        mk = newton.ImplicitModel(ck, specimen_consts, precision=mode, step=newton.StepParams(h=1e-4),
                                  prescribed=pk, prescribed_nodes=idk)
        try:
            itk = [mk.step().iterations for _ in range(STEPS)]
            uk = mk.displacements()
            fk = mk.failed_set()
        finally:
            mk.close()
        assert itk == its[:, k].tolist()
        assert np.array_equal(uk, u[col.node_mesh == k])
        first = np.flatnonzero(spec_of_cz == k)[0]
        assert np.array_equal(fk, failed[spec_of_cz[failed] == k] - first)

"""Host parts of the Newton driver (icflow_amd/newton.py): the symmetric expansion of the
upper-triangular BCSR (PARDISO mtype -2 input) and the substitute solve, against a dense
matrix built independently; and the three-point-bend specimen's supports and load line."""
import numpy as np

from icflow_amd import newton


def test_upper_bcsr_expansion_and_solve():
    rng = np.random.default_rng(0)
    N = 12
    M = rng.normal(size=(3 * N, 3 * N))
    A = M @ M.T + 3 * N * np.eye(3 * N)
    # sparsify at block level, keeping symmetry and every diagonal block
    keep = rng.random((N, N)) < 0.35
    keep = keep | keep.T | np.eye(N, dtype=bool)
    mask = np.kron(keep, np.ones((3, 3), bool))
    A = np.where(mask, A, 0.0)
    rows, cols, vals = [0], [], []
    for i in range(N):
        for j in range(i, N):
            if keep[i, j]:
                cols.append(j)
                vals.append(A[3 * i:3 * i + 3, 3 * j:3 * j + 3])
        rows.append(len(cols))
    rows, cols, vals = np.array(rows), np.array(cols), np.array(vals).ravel()
    S = newton.upper_bcsr_to_csc(rows, cols, vals).toarray()
    assert np.array_equal(S, A)
    b = rng.normal(size=3 * N)
    x = newton.solve_upper_bcsr(rows, cols, vals, b)
    assert np.allclose(A @ x, b, rtol=1e-10, atol=1e-10)


def test_three_point_bend_supports_and_load_line(lib):
    from icflow_amd import _lib
    consts = _lib.consts_from_prms(**_lib.SPECIMEN_PRMS)
    col, prescribed = newton.three_point_bend(consts, cells=(40, 8, 8), grains=16, seed=0, ramp_steps=200, h=1e-4)
    assert len(col.elastic) == 15360
    ids = np.flatnonzero(col.anchored)
    x = col.x0[ids]
    on_support = np.isclose(x[:, 2], 0) & (np.isclose(x[:, 0], 0.02) | np.isclose(x[:, 0], 0.38))
    on_top = np.isclose(x[:, 2], 0.05) & np.isclose(x[:, 0], 0.20)
    assert np.all(on_support | on_top)
    assert on_support.sum() >= 18 and on_top.sum() >= 9       # 9 node columns across y, more where split
    u_half = prescribed(100 * 1e-4)
    assert np.allclose(u_half[on_top, 2], -1e-3) and not u_half[on_support].any()
    assert np.allclose(prescribed(1.0)[on_top, 2], -2e-3)


def test_banded_solver_matches_dense_and_falls_back():
    rng = np.random.default_rng(1)
    N = 30
    keep = rng.random((N, N)) < 0.15
    keep = keep | keep.T | np.eye(N, dtype=bool)
    M = rng.normal(size=(3 * N, 3 * N))
    spd = np.where(np.kron(keep, np.ones((3, 3), bool)), M @ M.T, 0.0) + 40 * N * np.eye(3 * N)
    indef = spd - 80 * N * np.eye(3 * N)
    for A in (spd, indef):
        rows, cols, vals = [0], [], []
        for i in range(N):
            for j in range(i, N):
                if keep[i, j]:
                    cols.append(j)
                    vals.append(A[3 * i:3 * i + 3, 3 * j:3 * j + 3])
            rows.append(len(cols))
        rows, cols, vals = np.array(rows), np.array(cols), np.array(vals).ravel()
        s = newton.UpperBcsrSolver(rows, cols)
        b = rng.normal(size=3 * N)
        x = s.solve(vals, b)
        assert np.allclose(A @ x, b, rtol=1e-9, atol=1e-9 * np.abs(b).max())


def test_time_step_control_rules():
    """adjust_time_step against the cases of ImplicitModel3._adjustTimeStep (:258-288) and the
    acceptance rule of Step (:415-420), written out by hand."""
    from icflow_amd.newton import HOLD_FACTOR_DELAY as HOLD, PARTS, adjust_time_step as adj
    # converged at factor 1: accepted, nothing changes
    assert adj(1, 0, False, False, True, 10) == (True, 1, 0)
    # exploding: discarded, factor doubled, hold set
    assert adj(4, 1, True, False, True, 10) == (False, 8, HOLD)
    # diverging or not converged below the finest step: discarded, doubled
    assert adj(4, 0, False, True, False, 10) == (False, 8, HOLD)
    assert adj(4, 0, False, False, False, 10) == (False, 8, HOLD)
    # converged with a raised factor: the hold counts down first, then the factor halves
    assert adj(8, 3, False, False, True, 10) == (True, 8, 2)
    assert adj(8, 0, False, False, True, 10) == (True, 4, 2)
    # at the finest step a diverging attempt is accepted anyway, factor unchanged
    assert adj(PARTS, 0, False, True, False, 10) == (True, PARTS, 0)
    # with a single iteration per step every attempt is accepted
    assert adj(2, 0, False, False, False, 1) == (True, 2, 0)

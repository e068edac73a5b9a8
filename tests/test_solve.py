"""SolveDouble3 (the host solve behind the reference's PardisoLoader2 signature; a sparse
LDL^T, not MKL PARDISO) against dense solves of the same symmetric matrices: an assembled
Newmark system of a small beam, indefinite shifts of it, a block-diagonal matrix of two
disconnected parts, a matrix whose first pivot block is singular (handled by perturbation and
refinement), alternating patterns (the kept analysis is redone), and the refused arguments."""
import numpy as np
import pytest

from icflow_amd import _lib, newton
from icflow_amd._lib import ptr


def _solve(lib, rows, cols, vals, b, mtype=-2, dim=3, check=1):
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    a = np.ascontiguousarray(vals, np.float64)
    bb = np.ascontiguousarray(b, np.float64)
    b0 = bb.copy()
    x = np.zeros_like(bb)
    r = lib.SolveDouble3(ptr(cols), ptr(rows), ptr(a), len(rows) - 1, ptr(bb), ptr(x), mtype, 0, dim, 0, check)
    assert np.array_equal(bb, b0)          # b is not changed
    return r, x


def _dense(rows, cols, vals):
    return newton.upper_bcsr_to_csc(rows, cols, vals).toarray()


@pytest.fixture(scope="module")
def beam_system(orc, lib, specimen_consts):
    """A 10x4x4 three-point-bend beam's element + CZ Newmark system (the oracle's assembly)."""
    from icflow_amd.gpu_functionality import CSR, stride_of
    col, _ = newton.three_point_bend(specimen_consts, cells=(10, 4, 4), size=(0.1, 0.025, 0.025), grains=4, seed=0)
    alt = np.ascontiguousarray(col.alt, np.int32)
    csr = CSR(lib).build(alt[col.elastic], alt[col.czs])
    T, C, Nn = len(col.elastic), len(col.czs), len(col.x0)
    es, cs, ns = stride_of(T), stride_of(C), stride_of(Nn)
    ie = np.zeros(28 * es, np.int32)
    lib.icfh_pack_elems(csr.h, ptr(np.ascontiguousarray(col.elastic, np.int32)), ptr(alt), T, es, ptr(ie))
    icz = np.zeros(58 * cs, np.int32)
    lib.icfh_pack_czs(csr.h, ptr(np.ascontiguousarray(col.czs, np.int32)), ptr(alt), C, cs, ptr(icz))
    dn = np.zeros(18 * ns)
    rng = np.random.default_rng(1)
    for k in range(3):
        dn[k * ns:k * ns + Nn] = col.x0[:, k]
        dn[(3 + k) * ns:(3 + k) * ns + Nn] = rng.normal(size=Nn) * 1e-7
    A = np.zeros(9 * csr.nnz)
    b = np.zeros(3 * csr.N)
    orc.set_constants(specimen_consts)
    orc.elem_force(ie, dn, 1e-4, A, b, T, es, ns)
    orc.cz_force(np.zeros(16 * cs), icz, dn, A, b, 1e-4, C, cs, ns)
    return csr.rows.copy(), csr.cols.copy(), A, b


def test_matches_dense_solve(lib, beam_system):
    rows, cols, A, b = beam_system
    r, x = _solve(lib, rows, cols, A, b)
    assert r == 0
    M = _dense(rows, cols, A)
    xd = np.linalg.solve(M, b)
    assert np.linalg.norm(M @ x - b) <= 1e-13 * np.linalg.norm(b)
    assert np.abs(x - xd).max() <= 1e-10 * np.abs(xd).max()


@pytest.mark.parametrize("frac", [0.3, 0.7])
def test_indefinite_shift(lib, beam_system, frac):
    """A - s I with s between eigenvalues: symmetric indefinite (PARDISO's mtype -2)."""
    rows, cols, A, b = beam_system
    M = _dense(rows, cols, A)
    w = np.linalg.eigvalsh(M)
    k = int(frac * len(w))
    s = 0.5 * (w[k] + w[k + 1])
    As = A.copy().reshape(-1, 3, 3)
    br = np.repeat(np.arange(len(rows) - 1), np.diff(rows))
    As[br == cols] -= s * np.eye(3)
    r, x = _solve(lib, rows, cols, As.ravel(), b)
    Ms = M - s * np.eye(len(M))
    assert r == 0
    xd = np.linalg.solve(Ms, b)
    assert np.linalg.norm(Ms @ x - b) <= 1e-9 * np.linalg.norm(b)
    assert np.abs(x - xd).max() <= 1e-7 * np.abs(xd).max()


def test_disconnected_parts_and_alternating_patterns(lib, beam_system):
    rows, cols, A, b = beam_system
    N = len(rows) - 1
    rows2 = np.r_[rows, rows[1:] + rows[-1]]
    cols2 = np.r_[cols, cols + N]
    A2 = np.r_[A, 2 * A]
    b2 = np.r_[b, -b]
    for _ in range(2):                      # the kept analysis follows the pattern in use
        r, x2 = _solve(lib, rows2, cols2, A2, b2)
        assert r == 0
        M2 = _dense(rows2, cols2, A2)
        assert np.linalg.norm(M2 @ x2 - b2) <= 1e-13 * np.linalg.norm(b2)
        r, x = _solve(lib, rows, cols, A, b)
        assert r == 0 and np.linalg.norm(_dense(rows, cols, A) @ x - b) <= 1e-13 * np.linalg.norm(b)


def test_singular_pivot_block_is_perturbed_and_refined(lib):
    """[[0, B], [B^T, C]] with B invertible: the first 3x3 pivot block is zero, so the
    factorisation perturbs it; iterative refinement then recovers the solution."""
    rng = np.random.default_rng(7)
    B = rng.normal(size=(3, 3)) + 3 * np.eye(3)
    Cm = rng.normal(size=(3, 3))
    Cm = Cm @ Cm.T + np.eye(3)
    rows = np.array([0, 2, 3])
    cols = np.array([0, 1, 1])
    vals = np.r_[np.zeros(9), B.ravel(), Cm.ravel()]
    b = rng.normal(size=6)
    r, x = _solve(lib, rows, cols, vals, b)
    M = _dense(rows, cols, vals)
    assert r == 0
    assert np.linalg.norm(M @ x - b) <= 1e-6 * np.linalg.norm(b)


def test_refused_arguments(lib, beam_system):
    rows, cols, A, b = beam_system
    assert _solve(lib, rows, cols, A, b, mtype=11)[0] == -1      # unsymmetric: not supported
    assert _solve(lib, rows, cols, A, b, dim=1)[0] == -1         # scalar CSR: not supported
    bad = cols.copy()
    bad[rows[1]] = 0                                             # a block below the diagonal
    assert _solve(lib, rows, bad, A, b)[0] == -1
    # block row 0 without its diagonal block (found by the analysis even without the check)
    eye = np.eye(3).ravel()
    assert _solve(lib, [0, 1, 2], [1, 1], np.r_[eye, eye], np.ones(6), check=0)[0] == -1


def test_newton_solver_class(lib, beam_system):
    rows, cols, A, b = beam_system
    x = newton.SolveDouble3Solver(rows, cols).solve(A, b)
    xb = newton.UpperBcsrSolver(rows, cols).solve(A, b)
    assert np.abs(x - xb).max() <= 1e-10 * np.abs(xb).max()


def _cube_system(orc, lib, consts, cells, seed):
    """An 8-grain cube's element + CZ Newmark system, bottom face fixed (the oracle's assembly)."""
    from icflow_amd import specimens
    from icflow_amd.gpu_functionality import CSR, stride_of
    col = specimens.make_specimen((cells,) * 3, (0.02,) * 3, 8, seed)
    x = col.meshes[0].x0
    col.anchored = np.isclose(x[:, 2], 0.0)
    col.finalize()
    alt = np.ascontiguousarray(col.alt, np.int32)
    csr = CSR(lib).build(alt[col.elastic], alt[col.czs])
    T, C, Nn = len(col.elastic), len(col.czs), len(col.x0)
    es, cs, ns = stride_of(T), stride_of(C), stride_of(Nn)
    ie = np.zeros(28 * es, np.int32)
    lib.icfh_pack_elems(csr.h, ptr(np.ascontiguousarray(col.elastic, np.int32)), ptr(alt), T, es, ptr(ie))
    icz = np.zeros(58 * cs, np.int32)
    lib.icfh_pack_czs(csr.h, ptr(np.ascontiguousarray(col.czs, np.int32)), ptr(alt), C, cs, ptr(icz))
    dn = np.zeros(18 * ns)
    rng = np.random.default_rng(seed)
    for k in range(3):
        dn[k * ns:k * ns + Nn] = col.x0[:, k]
        dn[(3 + k) * ns:(3 + k) * ns + Nn] = rng.normal(size=Nn) * 1e-7
    A = np.zeros(9 * csr.nnz)
    b = np.zeros(3 * csr.N)
    orc.set_constants(consts)
    orc.elem_force(ie, dn, 1e-4, A, b, T, es, ns)
    orc.cz_force(np.zeros(16 * cs), icz, dn, A, b, 1e-4, C, cs, ns)
    return csr.rows.copy(), csr.cols.copy(), A, b


def test_cube_against_sparse_lu(orc, lib, specimen_consts):
    """A compact 3D system (nested dissection, large supernodes) against scipy's sparse LU."""
    import scipy.sparse.linalg as sla
    rows, cols, A, b = _cube_system(orc, lib, specimen_consts, 7, 3)
    r, x = _solve(lib, rows, cols, A, b)
    assert r == 0
    M = newton.upper_bcsr_to_csc(rows, cols, A).tocsc()
    xr = sla.spsolve(M, b)
    assert np.linalg.norm(M @ x - b) <= 1e-13 * np.linalg.norm(b)
    assert np.abs(x - xr).max() <= 1e-10 * np.abs(xr).max()


def test_concurrent_calls_and_thread_counts_agree(orc, lib, specimen_consts, beam_system):
    """Several threads solving different systems at once (each pattern's analysis shared from
    the cache) give the sequential results bit for bit, with and without the solve's own
    threads; a negative thread count is refused."""
    from concurrent.futures import ThreadPoolExecutor
    systems = [beam_system] + [_cube_system(orc, lib, specimen_consts, c, s) for c, s in ((4, 1), (5, 2), (6, 4))]
    assert lib.icfh_solve_threads(-1) == -1
    try:
        lib.icfh_solve_threads(0)
        seq = [_solve(lib, *s)[1] for s in systems]
        lib.icfh_solve_threads(1)
        with ThreadPoolExecutor(4) as pool:
            par = list(pool.map(lambda s: _solve(lib, *s), systems * 3))
    finally:
        lib.icfh_solve_threads(0)
    for k, (r, x) in enumerate(par):
        assert r == 0
        assert np.array_equal(x, seq[k % len(systems)])


def test_batched_solver_per_specimen(orc, lib, specimen_consts, beam_system):
    """BatchedSolver over a block-diagonal matrix of three systems gives each system's own
    SolveDouble3 result bit for bit, leaves inactive specimens at zero, and refuses a matrix
    that is not block-diagonal over the given rows."""
    parts = [beam_system] + [_cube_system(orc, lib, specimen_consts, c, s) for c, s in ((4, 1), (5, 2))]
    rows, cols, vals, rhs, row_off = [0], [], [], [], [0]
    for r, c, a, b in parts:
        rows.extend((np.asarray(r[1:]) + rows[-1]).tolist())
        cols.extend((np.asarray(c) + row_off[-1]).tolist())
        vals.append(a)
        rhs.append(b)
        row_off.append(row_off[-1] + len(r) - 1)
    rows, cols = np.array(rows), np.array(cols)
    vals, rhs = np.concatenate(vals), np.concatenate(rhs)
    bs = newton.BatchedSolver(rows, cols, row_off, threads=2)
    try:
        dx = bs.solve(vals, rhs)
        for k, (r, c, a, b) in enumerate(parts):
            assert np.array_equal(dx[3 * row_off[k]:3 * row_off[k + 1]], _solve(lib, r, c, a, b)[1])
        dx2 = bs.solve(vals, rhs, active=np.array([True, False, True]))
        assert np.array_equal(dx2[3 * row_off[1]:3 * row_off[2]], np.zeros(3 * (row_off[2] - row_off[1])))
        assert np.array_equal(dx2[3 * row_off[2]:], dx[3 * row_off[2]:])
    finally:
        bs.close()
    with pytest.raises(ValueError):
        newton.BatchedSolver(rows, cols, [0, row_off[1] - 1, row_off[-1]])


def _analyse(lib, rows, cols):
    import ctypes
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    fd, pk, fl = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_double()
    r = lib.icfh_solve_analyse(ptr(cols), ptr(rows), len(rows) - 1, ctypes.byref(fd), ctypes.byref(pk),
                               ctypes.byref(fl))
    return r, fd.value, pk.value, fl.value


def test_analysis_sizes(lib, beam_system):
    """icfh_solve_analyse: a fully coupled pattern of n block rows is one dense front -- (3n)^2
    doubles of panel and (3n)^3 multiply-adds (M K^2 with M = K = 3n); a block-diagonal pattern
    has no fill (9 doubles and 27 multiply-adds per block row); the beam's figures are
    consistent (peak >= factor), and a solve after the analysis (the cached analysis) still
    matches the dense solve."""
    n = 10
    rows = np.arange(n + 1, dtype=np.int32) * 0
    cols = []
    for r in range(n):
        cols += list(range(r, n))
        rows[r + 1] = len(cols)
    r, fd, pk, fl = _analyse(lib, rows, np.array(cols, np.int32))
    assert r == 0 and fd == (3 * n) ** 2 and fl == (3 * n) ** 3 and pk >= fd
    diag_rows, diag_cols = np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32)
    r, fd, pk, fl = _analyse(lib, diag_rows, diag_cols)
    assert r == 0 and fd == 9 * n and fl == 27 * n
    brows, bcols, A, b = beam_system
    r, fd, pk, fl = _analyse(lib, brows, bcols)
    assert r == 0 and fd > 9 * (len(brows) - 1) and pk >= fd and fl > 0
    r, x = _solve(lib, brows, bcols, A, b)
    assert r == 0
    xd = np.linalg.solve(_dense(brows, bcols, A), b)
    assert np.abs(x - xd).max() <= 1e-9 * np.abs(xd).max()
    import ctypes
    z = ctypes.c_longlong()
    assert lib.icfh_solve_analyse(None, None, n, ctypes.byref(z), ctypes.byref(z), None) == -1

"""Pins the parts of the oracle that a restatement could get wrong without any other test
noticing: the Newmark coefficients of the element kernel (element_kernels2.h:228-251), the
consistent mass (ModelPrms.cs:239-248), the CZ Gauss points and shape values
(ModelPrms.cs:288-307, 324-349) and the CZ force scale w * a_J (cz_kernels.h:343-364).

Every expected value here comes from mathematics that does not use the oracle's formulas:

 - Newmark consistency: the effective tangent is minus the derivative of the residual with
   respect to the displacement increment, when vn and an follow from un by the Newmark
   relations (orc_nodes_infer, itself pinned in test_oracle_nodes.py). Checked by central
   differences with a step size h small enough, and damping coefficients large and unequal
   enough, that the mass, damping and stiffness terms are all of the same order;
 - resultants on a free tet: a uniform velocity and acceleration are rigid translations, so
   the summed residual is m (g - alpha_M v - a) with m = rho V, V from the coordinates;
   a rate-of-strain velocity gives the Rayleigh damping force -(alpha_M M + alpha_K K) v with
   K = V B^T E B and M from quadrature;
 - the consistent mass equals rho * integral(N_i N_j) from a 4-point rule exact for quadratics,
   and gives the kinetic energy of a rigid rotation;
 - the 3-point CZ rule integrates products of linear shape functions exactly:
   sum_gp w sf_i sf_j = (1 + delta_ij) / 12, sum_gp w sf_i = 1/3;
 - a flat CZ opened uniformly along its normal carries the traction Tn(delta, 0) over its area:
   the three top nodes receive -area * Tn * n / 3 each, the bottom nodes the opposite, and the
   summed normal stiffness is area * Dnn. Tn and Dnn are pinned by the PPR closed forms.

Each test fails under the mutation named in its docstring.
"""
import numpy as np
import pytest

from conftest import rel_err
from test_oracle_element import rand_rotation, rand_tet, voigt_B

# 4-point rule on a tet, exact for polynomials of degree 2 (Keast / Hammer-Stroud):
# barycentric points (a, b, b, b) and permutations, weights 1/4 of the volume.
_QA, _QB = 0.5854101966249685, 0.1381966011250105
_QPTS = np.array([[_QA, _QB, _QB, _QB], [_QB, _QA, _QB, _QB], [_QB, _QB, _QA, _QB], [_QB, _QB, _QB, _QA]])


def _quad_mass(rho):
    """rho * integral(N_i N_j dV) / V over any tet (4x4)."""
    return rho * sum(0.25 * np.outer(p, p) for p in _QPTS)


@pytest.fixture(scope="module")
def Cd(lib):
    """Specimen material with unequal, large damping coefficients (the reference's defaults
    are equal, 0.0005 each, which would hide a swap of alpha_M and alpha_K)."""
    from icflow_amd import _lib
    return _lib.consts_from_prms(**{**_lib.SPECIMEN_PRMS, "dampingMass": 4.0e5, "dampingStiffness": 7.0e-7})


def _one_tet(orc, consts, x0, un, vn, an, h):
    """orc_elem_force on a single tet with every block written (16 slots, 4 rows):
    returns rhs (12,), LHS (12, 12) and the node force F (4, 3)."""
    orc.set_constants(consts)
    ns = 4
    ie = np.zeros(28, np.int32)
    ie[0:4] = np.arange(4)
    ie[4:20] = np.arange(16)
    ie[20:24] = np.arange(4)
    dn = np.zeros(18 * ns)
    for k in range(3):
        dn[k * ns:k * ns + 4] = x0[:, k]
        dn[(3 + k) * ns:(3 + k) * ns + 4] = un[:, k]
        dn[(6 + k) * ns:(6 + k) * ns + 4] = vn[:, k]
        dn[(9 + k) * ns:(9 + k) * ns + 4] = an[:, k]
    A = np.zeros(16 * 9)
    b = np.zeros(12)
    orc.elem_force(ie, dn, h, A, b, 1, 1, ns)
    L = np.zeros((12, 12))
    for i in range(4):
        for j in range(4):
            L[3 * i:3 * i + 3, 3 * j:3 * j + 3] = A[9 * (4 * i + j):9 * (4 * i + j) + 9].reshape(3, 3)
    F = np.stack([dn[(12 + k) * ns:(12 + k) * ns + 4] for k in range(3)], axis=1)
    return b.copy(), L, F


def _infer(orc, u, v, a, du, h, beta, gamma):
    """un, vn, an of 4 nodes from the start-of-step state and the increment (Node.cs:79-97)."""
    ss = ns = 4
    st = np.zeros(12 * ss)
    for k in range(3):
        st[k * ss:k * ss + 4] = u[:, k]
        st[(3 + k) * ss:(3 + k) * ss + 4] = v[:, k]
        st[(6 + k) * ss:(6 + k) * ss + 4] = a[:, k]
        st[(9 + k) * ss:(9 + k) * ss + 4] = du[:, k]
    dn = np.zeros(18 * ns)
    orc.nodes_infer(st, ss, dn, ns, 4, h, beta, gamma)
    get = lambda f: np.stack([dn[(f + k) * ns:(f + k) * ns + 4] for k in range(3)], axis=1)
    return get(3), get(6), get(9)


def test_newmark_tangent_is_minus_residual_derivative(orc, Cd):
    """LHS = -d rhs / d un along the Newmark path. Fails if beta or gamma sits in the wrong
    place in either mass or stiffness coefficient, if 1/h^2 becomes 1/h, or if the tangent's
    damping terms do not match the residual's."""
    rng = np.random.default_rng(11)
    h = 1.0e-6
    bt, gm = Cd.NewmarkBeta, Cd.NewmarkGamma
    for _ in range(6):
        x0 = rand_tet(rng)
        # small strain around a rigid rotation: there Df = d f / d un up to O(strain)
        Q = rand_rotation(rng)
        u = (x0 @ Q.T - x0) + rng.normal(size=(4, 3)) * 1e-11
        v = rng.normal(size=(4, 3)) * 1e-3
        a = rng.normal(size=(4, 3))
        du0 = rng.normal(size=(4, 3)) * 1e-11
        un, vn, an = _infer(orc, u, v, a, du0, h, bt, gm)
        _, L, _ = _one_tet(orc, Cd, x0, un, vn, an, h)
        J = np.zeros((12, 12))
        step = 1e-10
        for c in range(12):
            d = np.zeros(12)
            d[c] = step
            r = []
            for sgn in (1, -1):
                un_, vn_, an_ = _infer(orc, u, v, a, du0 + sgn * d.reshape(4, 3), h, bt, gm)
                r.append(_one_tet(orc, Cd, x0, un_, vn_, an_, h)[0])
            J[:, c] = (r[0] - r[1]) / (2 * step)
        # the three parts of the tangent are of the same order at this h
        B, V = voigt_B(x0)
        K = V * B.T @ np.array(Cd.E[:]).reshape(6, 6) @ B
        m = Cd.rho * V / 20 / (bt * h * h)
        assert 0.05 < m / np.max(np.abs(K)) < 20
        assert 0.05 < Cd.dampingStiffness * gm / (bt * h) < 20
        assert 0.05 < Cd.dampingMass * gm * h < 20
        assert rel_err(-J, L) <= 1e-6


def test_free_tet_resultants(orc, Cd):
    """Uniform v and a (rigid translation) on an unstrained tet: the summed residual is
    rho V (g e_z - alpha_M v - a). Fails if gravity changes sign, if the inertia or the mass
    damping term changes sign, if alpha_M and alpha_K are swapped, or if the mass loses rho."""
    rng = np.random.default_rng(12)
    h = 1e-4
    for _ in range(10):
        x0 = rand_tet(rng)
        _, V = voigt_B(x0)
        v = rng.normal(size=3) * 0.1
        a = rng.normal(size=3) * 3.0
        rhs, _, F = _one_tet(orc, Cd, x0, np.zeros((4, 3)), np.tile(v, (4, 1)), np.tile(a, (4, 1)), h)
        m = Cd.rho * V
        want = m * (np.array([0.0, 0.0, Cd.gravity]) - Cd.dampingMass * v - a)
        got = rhs.reshape(4, 3).sum(axis=0)
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
        assert rel_err(F, rhs.reshape(4, 3)) == 0.0      # node force carries the whole residual
    # at rest with no motion: only gravity, a quarter of the weight on each node, downward
    rhs, _, _ = _one_tet(orc, Cd, x0, np.zeros((4, 3)), np.zeros((4, 3)), np.zeros((4, 3)), h)
    assert Cd.gravity < 0
    # (the unstrained tet's elastic force is round-off of R K (R^T x - x0), ~1e-6 of the weight)
    wq = Cd.rho * V * Cd.gravity / 4
    assert np.max(np.abs(rhs.reshape(4, 3) - [[0, 0, wq]] * 4)) <= 1e-5 * abs(wq)


def test_rayleigh_damping_of_a_strain_rate(orc, Cd):
    """An affine velocity v(x) = D x on an unstrained tet: rhs = -(alpha_M M_q + alpha_K K) v
    with K = V B^T E B and M_q from the 4-point rule (no oracle piece is used for the
    expectation). Fails if the damping coefficients are swapped or the stiffness damping uses
    anything but the tangent."""
    rng = np.random.default_rng(13)
    h = 1e-4
    Cz = Cd
    E = np.array(Cz.E[:]).reshape(6, 6)
    for _ in range(10):
        x0 = rand_tet(rng)
        B, V = voigt_B(x0)
        K = V * B.T @ E @ B
        Mq = np.kron(_quad_mass(Cz.rho) * V, np.eye(3))
        D = rng.normal(size=(3, 3)) * 10.0
        vel = x0 @ D.T
        rhs, _, _ = _one_tet(orc, Cz, x0, np.zeros((4, 3)), vel, np.zeros((4, 3)), h)
        rhs[2::3] -= Cz.rho * V * Cz.gravity / 4          # gravity pinned above
        want = -(Cz.dampingMass * Mq + Cz.dampingStiffness * K) @ vel.ravel()
        assert rel_err(rhs, want) <= 1e-8        # round-off of the unstrained elastic force


def test_consistent_mass_by_quadrature(orc, specimen_consts):
    """M / V = rho * integral(N_i N_j) per direction and zero across directions, checked with a
    quadrature rule and through the kinetic energy of a rigid rotation. Fails for a lumped
    mass (rho V / 4 on the diagonal), for a missing rho or for a wrong off-diagonal weight."""
    c = specimen_consts
    M = np.array(c.M[:]).reshape(12, 12)
    Mq = _quad_mass(c.rho)
    for d in range(3):
        for e in range(3):
            blk = M[d::3, e::3]
            assert np.max(np.abs(blk - (Mq if d == e else 0))) <= 1e-12 * c.rho
    rng = np.random.default_rng(14)
    for _ in range(5):
        x0 = rand_tet(rng)
        _, V = voigt_B(x0)
        w = rng.normal(size=3)
        vel = np.cross(w, x0)                              # rigid rotation about the origin
        ke = 0.5 * vel.ravel() @ (V * M) @ vel.ravel()
        # exact: 1/2 rho integral |w x r|^2 dV, the integrand is quadratic: the rule is exact
        qp = _QPTS @ x0
        ke_q = 0.5 * c.rho * V * 0.25 * np.sum(np.sum(np.cross(w, qp) ** 2, axis=1))
        assert abs(ke - ke_q) <= 1e-12 * ke_q
        # and the oracle kernel applies it: a unit acceleration of one dof gives -V M[:, k]
        # (differences with the motionless tet remove gravity and the elastic round-off)
        z = np.zeros((4, 3))
        rest, _, _ = _one_tet(orc, c, x0, z, z, z, 1e-4)
        for k in (0, 4, 11):
            an = np.zeros(12)
            an[k] = 1.0
            rhs, _, _ = _one_tet(orc, c, x0, z, z, an.reshape(4, 3), 1e-4)
            rhs -= rest
            want = -V * np.kron(Mq, np.eye(3))[:, k]
            assert rel_err(rhs, want) <= 1e-12


def test_cz_gauss_rule_is_exact_for_quadratics(specimen_consts):
    """sum_gp (1/3) sf_i sf_j = (1 + delta_ij)/12 and sum_gp (1/3) sf_i = 1/3: the Gauss
    points integrate linear-times-linear shape functions exactly over the triangle. B is
    [s (x) I3, -s (x) I3] per Gauss point. Fails for wrong Gauss points (e.g. the vertices or
    the centroid) or a permuted shape-value table."""
    sf = np.array(specimen_consts.sf[:]).reshape(3, 3)     # sf[node][gp]
    w = np.full(3, 1.0 / 3.0)
    G = (sf * w) @ sf.T
    assert np.allclose(G, (np.ones((3, 3)) + np.eye(3)) / 12, rtol=0, atol=1e-15)
    assert np.allclose(sf @ w, 1.0 / 3.0, rtol=0, atol=1e-15)
    # each Gauss point's shape values are the barycentric coordinates of an interior point
    assert np.allclose(sf.sum(axis=0), 1.0) and np.all(sf > 0)
    Bm = np.array(specimen_consts.B[:]).reshape(3, 3, 18)
    for g in range(3):
        want = np.hstack([np.kron(sf[:, g][None, :], np.eye(3)), -np.kron(sf[:, g][None, :], np.eye(3))])
        assert np.array_equal(Bm[g], want)


def _cz_case(orc, C, x, u, pmax=0.0, tmax=0.0):
    from test_oracle_cz import _one_cz
    return _one_cz(orc, C, x, u, [pmax] * 3, [tmax] * 3)


@pytest.fixture(scope="module")
def Cs(orc, specimen_consts):
    orc.set_constants(specimen_consts)
    return specimen_consts


def test_cz_traction_resultant(orc, Cs):
    """Flat CZ, top face opened uniformly by delta along the normal (or slid by delta along
    e1), undamaged history: each top node gets -area T / 3, each bottom node +area T / 3, and
    the summed normal stiffness of the top face is area * Dnn. Fails if the Gauss weight 1/3
    is dropped, the area doubled, the traction's sign flipped or the faces exchanged."""
    rng = np.random.default_rng(15)
    for trial in range(8):
        x = rng.normal(size=(3, 3)) * 0.01 + rng.uniform(-0.1, 0.1, size=3)
        area = 0.5 * np.linalg.norm(np.cross(x[1] - x[0], x[2] - x[0]))
        n = np.cross(x[1] - x[0], x[0] - x[2])
        n /= np.linalg.norm(n)
        e1 = (x[1] - x[0]) / np.linalg.norm(x[1] - x[0])
        # normal opening, pre-peak and post-peak (loading: history below the opening)
        for frac in (0.3, 2.0):
            d = frac * Cs.rn * Cs.deln
            u = np.zeros((6, 3))
            u[:3] = d * n
            K, b, icz, _ = _cz_case(orc, Cs, x, u)
            assert icz[3] == 0 and icz[1] == 0
            Tn = orc.ppr("Tn", d, 0.0)
            assert Tn > 0
            top, bot = b[:9].reshape(3, 3), b[9:].reshape(3, 3)
            scale = area * Tn
            assert np.max(np.abs(top - (-area * Tn / 3) * n)) <= 1e-9 * scale
            assert np.max(np.abs(bot - (area * Tn / 3) * n)) <= 1e-9 * scale
            Ktt = sum(K[3 * i:3 * i + 3, 3 * j:3 * j + 3] for i in range(3) for j in range(3))
            Dnn = orc.ppr("Dnn", d, 0.0)
            assert abs(n @ Ktt @ n - area * Dnn) <= 1e-9 * area * abs(Dnn)
        # tangential slide along e1, no normal opening
        d = 0.5 * Cs.rt * Cs.delt
        u = np.zeros((6, 3))
        u[:3] = d * e1
        K, b, icz, _ = _cz_case(orc, Cs, x, u)
        assert icz[3] == 0
        Tt = orc.ppr("Tt", 0.0, d)
        top = b[:9].reshape(3, 3)
        assert np.max(np.abs(top - (-area * Tt / 3) * e1)) <= 1e-9 * area * Tt
        assert np.max(np.abs(b[9:].reshape(3, 3) + top)) <= 1e-9 * area * Tt

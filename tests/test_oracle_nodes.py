"""Pins the oracle's node-state functions (orc_nodes_infer, orc_nodes_add_dx,
orc_nodes_accept: v3Library/Geometry/Node.cs:62-97, ImplicitModel3.cs:247-256) against the
Newmark relations they implement, written in a different algebraic form:
  un = u + du,
  un = u + h v + h^2 ((1/2 - beta) a + beta an),
  vn = v + h ((1 - gamma) a + gamma an),
and, for beta = 1/4, gamma = 1/2, exact reproduction of constant-acceleration motion."""
import numpy as np

from icflow_amd.gpu_functionality import stride_of


def _state(rng, n, ss):
    st = np.zeros(12 * ss)
    for f in range(12):
        st[f * ss:f * ss + n] = rng.normal(0, 1e-3 * (10 ** (f // 3)), n)
    return st


def _fields(buf, stride, f0, n):
    return np.stack([buf[(f0 + k) * stride:(f0 + k) * stride + n] for k in range(3)], 1)


def test_infer_satisfies_newmark_relations(orc):
    rng = np.random.default_rng(1)
    n = 300
    ss = ns = stride_of(n)
    for beta, gamma in ((0.25, 0.5), (0.3025, 0.6), (0.4, 0.7)):
        st = _state(rng, n, ss)
        dn = np.zeros(18 * ns)
        dn[:3 * ns] = rng.normal(0, 0.05, 3 * ns)
        dn[12 * ns:15 * ns] = 7.0                      # F must come back zeroed
        h = 1e-4
        orc.nodes_infer(st, ss, dn, ns, n, h, beta, gamma)
        u, v, a, du = (_fields(st, ss, 3 * q, n) for q in range(4))
        x0, un, vn, an = (_fields(dn, ns, f, n) for f in (0, 3, 6, 9))
        assert np.allclose(un, u + du, rtol=1e-15, atol=0)
        lhs = u + h * v + h * h * ((0.5 - beta) * a + beta * an)
        assert np.allclose(un, lhs, rtol=1e-9, atol=1e-9 * np.abs(un).max())
        assert np.allclose(vn, v + h * ((1 - gamma) * a + gamma * an), rtol=1e-12, atol=1e-12 * np.abs(vn).max())
        assert np.array_equal(_fields(dn, ns, 15, n), x0 + un)
        assert not _fields(dn, ns, 12, n).any()


def test_constant_acceleration_is_exact_for_average_acceleration(orc):
    rng = np.random.default_rng(2)
    n = 50
    ss = ns = stride_of(n)
    h = 1e-3
    st = np.zeros(12 * ss)
    v = rng.normal(0, 1, (n, 3))
    a = rng.normal(0, 10, (n, 3))
    for k in range(3):
        st[(3 + k) * ss:(3 + k) * ss + n] = v[:, k]
        st[(6 + k) * ss:(6 + k) * ss + n] = a[:, k]
        st[(9 + k) * ss:(9 + k) * ss + n] = v[:, k] * h + a[:, k] * h * h / 2
    dn = np.zeros(18 * ns)
    orc.nodes_infer(st, ss, dn, ns, n, h, 0.25, 0.5)
    assert np.allclose(_fields(dn, ns, 9, n), a, rtol=1e-6, atol=1e-6)
    assert np.allclose(_fields(dn, ns, 6, n), v + a * h, rtol=1e-9, atol=1e-9)


def test_add_dx_skips_anchored_nodes(orc):
    rng = np.random.default_rng(3)
    n = 40
    ss = stride_of(n)
    st = _state(rng, n, ss)
    alt = np.where(rng.random(n) < 0.3, -1, 0).astype(np.int32)
    alt[alt >= 0] = rng.permutation(int((alt >= 0).sum()))
    dx = rng.normal(0, 1, 3 * int((alt >= 0).sum()))
    before = st.copy()
    orc.nodes_add_dx(st, ss, alt, n, dx)
    du0, du1 = _fields(before, ss, 9, n), _fields(st, ss, 9, n)
    for i in range(n):
        want = du0[i] + (dx[3 * alt[i]:3 * alt[i] + 3] if alt[i] >= 0 else 0)
        assert np.array_equal(du1[i], want)
    assert np.array_equal(st[:9 * ss], before[:9 * ss])


def test_accept_backward_differences(orc):
    rng = np.random.default_rng(4)
    n = 60
    ss = ns = stride_of(n)
    h = 2e-4
    st = _state(rng, n, ss)
    dn = np.zeros(18 * ns)
    orc.nodes_infer(st, ss, dn, ns, n, h, 0.25, 0.5)
    u0, v0 = _fields(st, ss, 0, n), _fields(st, ss, 3, n)
    un = _fields(dn, ns, 3, n)
    orc.nodes_accept(st, ss, dn, ns, n, h)
    u1, v1, a1, du1 = (_fields(st, ss, 3 * q, n) for q in range(4))
    assert np.array_equal(u1, un)
    assert np.allclose(u0 + h * v1, un, rtol=1e-12, atol=1e-15)
    assert np.allclose(v0 + h * a1, v1, rtol=1e-9, atol=1e-12)
    assert not du1.any()

"""Pins the oracle's cohesive-zone math (orc_Tn/Tt/Dnn/Dtt/Dnt, orc_cohesive_law,
orc_cz_rotation, orc_cz_force; restating cz_kernels9/PPR_math.h and cz_kernels.h):

 - PPR closed forms (Park-Paulino-Roesler 2009): peak tractions sigma_max / tau_max at the
   openings lambda delta, zero slope there, fracture energy = integral of the traction, and
   zero traction at the final openings delta_n / delta_t;
 - the tangent functions are the derivatives of the tractions (central differences), and
   Dnt is both dTn/dDt and dTt/dDn;
 - law branches: unloading / reloading and contact are secants through the origin,
   softening follows the envelope and updates the history, openings past delta fail;
 - CZ element: the frame is orthonormal with the area of the triangle; a rigid motion gives
   no force; in the linear (secant) branch the residual equals -Keff u exactly, which pins
   the sign, B structure and rotation of both rhs and tangent.
"""
import numpy as np
import pytest
from scipy.integrate import quad

from conftest import rel_err


@pytest.fixture(scope="module")
def C(orc, specimen_consts):
    orc.set_constants(specimen_consts)
    return specimen_consts


@pytest.fixture(scope="module")
def Cdef(orc):
    """Reference default parameters (alpha = beta = 3, phi = 30, sigma 0.4 MPa, tau 1.5 MPa)."""
    from icflow_amd import _lib
    return _lib.consts_from_prms()


@pytest.mark.parametrize("which", ["specimen", "default"])
def test_ppr_closed_forms(orc, C, Cdef, which):
    c = C if which == "specimen" else Cdef
    orc.set_constants(c)
    ln, lt = c.rn * c.deln, c.rt * c.delt
    assert abs(orc.ppr("Tn", ln, 0.0) - c.f_tn) <= 1e-12 * c.f_tn
    assert abs(orc.ppr("Tt", 0.0, lt) - c.f_tt) <= 1e-12 * c.f_tt
    assert abs(orc.ppr("Dnn", ln, 0.0)) <= 1e-9 * c.f_tn / c.deln        # peak: zero slope
    assert abs(orc.ppr("Dtt", 0.0, lt)) <= 1e-9 * c.f_tt / c.delt
    assert abs(orc.ppr("Tn", c.deln, 0.0)) <= 1e-12 * c.f_tn              # complete separation
    phi_n = quad(lambda d: orc.ppr("Tn", d, 0.0), 0, c.deln, points=[ln], limit=200, epsabs=0, epsrel=1e-12)[0]
    phi_t = quad(lambda d: orc.ppr("Tt", 0.0, d), 0, c.delt, points=[lt], limit=200, epsabs=0, epsrel=1e-12)[0]
    assert abs(phi_n - c.G_fn) <= 1e-8 * c.G_fn
    assert abs(phi_t - c.G_ft) <= 1e-8 * c.G_ft
    orc.set_constants(C)


def test_ppr_tangents_are_derivatives(orc, C):
    rng = np.random.default_rng(0)
    for _ in range(200):
        dn_ = rng.uniform(0.001, 0.95) * C.deln
        dt_ = rng.uniform(0.001, 0.95) * C.delt
        hn, ht = 1e-6 * C.deln, 1e-6 * C.delt
        fd_nn = (orc.ppr("Tn", dn_ + hn, dt_) - orc.ppr("Tn", dn_ - hn, dt_)) / (2 * hn)
        fd_tt = (orc.ppr("Tt", dn_, dt_ + ht) - orc.ppr("Tt", dn_, dt_ - ht)) / (2 * ht)
        fd_nt = (orc.ppr("Tn", dn_, dt_ + ht) - orc.ppr("Tn", dn_, dt_ - ht)) / (2 * ht)
        fd_tn = (orc.ppr("Tt", dn_ + hn, dt_) - orc.ppr("Tt", dn_ - hn, dt_)) / (2 * hn)
        sc_n, sc_t = C.f_tn / C.deln, C.f_tt / C.delt
        Dnn, Dtt, Dnt = (orc.ppr(k, dn_, dt_) for k in ("Dnn", "Dtt", "Dnt"))
        # central differences: truncation error is relative to the (steep near zero) slope
        assert abs(Dnn - fd_nn) <= 1e-5 * max(abs(Dnn), sc_n)
        assert abs(Dtt - fd_tt) <= 1e-5 * max(abs(Dtt), sc_t)
        assert abs(Dnt - fd_nt) <= 1e-5 * max(abs(Dnt), sc_n, sc_t)
        assert abs(Dnt - fd_tn) <= 1e-5 * max(abs(Dnt), sc_n, sc_t)


def test_law_branches(orc, C):
    dn_, dt_ = C.deln, C.delt
    # failure past the final opening
    r = orc.cohesive_law(0, 0, 1.01 * dn_, 0.0)
    assert r["failed"] and r["Tn"] == 0 and r["Dnn"] == 0
    r = orc.cohesive_law(0, 0, 0.0, 1.01 * dt_)
    assert r["failed"]
    # softening: envelope value, history updated
    o_n, o_t = 0.3 * dn_, 0.2 * dt_
    r = orc.cohesive_law(0.1 * dn_, 0.1 * dt_, o_n, o_t)
    assert not r["failed"] and r["pmax"] == o_n and r["tmax"] == o_t
    assert r["Tn"] == orc.ppr("Tn", o_n, o_t) and r["Dnn"] == orc.ppr("Dnn", o_n, o_t)
    assert r["Dnt"] == orc.ppr("Dnt", o_n, o_t)
    # unloading in both directions: secants through the origin, linear in the opening
    pm, tm = 0.4 * dn_, 0.3 * dt_
    r1 = orc.cohesive_law(pm, tm, 0.05 * dn_, 0.04 * dt_)
    r2 = orc.cohesive_law(pm, tm, 0.10 * dn_, 0.08 * dt_)
    assert not r1["failed"] and r1["pmax"] == pm and r1["tmax"] == tm
    assert abs(r2["Tn"] - 2 * r1["Tn"]) <= 1e-12 * abs(r2["Tn"])
    assert abs(r2["Tt"] - 2 * r1["Tt"]) <= 1e-12 * abs(r2["Tt"])
    assert abs(r1["Dnn"] - orc.ppr("Tn", pm, tm) / pm) <= 1e-12 * r1["Dnn"]
    assert abs(r1["Dtt"] - orc.ppr("Tt", pm, tm) / tm) <= 1e-12 * r1["Dtt"]
    assert r1["Dnt"] == 0
    # contact: negative normal opening, compressive secant traction
    r = orc.cohesive_law(pm, tm, -0.01 * dn_, 0.0)
    assert r["contact"] and r["Tn"] < 0 and r["Dnn"] > 0 and not r["failed"]
    assert abs(r["Tn"] - orc.ppr("Tn", pm, tm) * (-0.01 * dn_) / pm) <= 1e-12 * abs(r["Tn"])
    # undamaged contact uses the initial stiffness Dnn(0, tmax)
    r = orc.cohesive_law(0.0, 0.0, -0.01 * dn_, 0.0)
    assert abs(r["Dnn"] - orc.ppr("Dnn", 0.0, 0.0)) <= 1e-12 * r["Dnn"]
    # weak-traction failure beyond lambda delta / 5 (cz_kernels.h:93-95)
    r = orc.cohesive_law(0.0, 0.0, 0.97 * dn_, 0.0)
    assert orc.ppr("Tn", 0.97 * dn_, 0.0) < 0.05 * C.f_tn and r["failed"]


def test_cz_frame(orc):
    rng = np.random.default_rng(1)
    for _ in range(20):
        p = rng.normal(size=(3, 3))
        R, area = orc.cz_rotation(p.ravel())
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(R) - 1) < 1e-14
        assert abs(area - 0.5 * np.linalg.norm(np.cross(p[1] - p[0], p[2] - p[0]))) <= 1e-14 * area
        assert np.allclose(R[0], (p[1] - p[0]) / np.linalg.norm(p[1] - p[0]))
        n = np.cross(p[1] - p[0], p[0] - p[2])
        assert np.allclose(R[2], n / np.linalg.norm(n))


def _one_cz(orc, C, x_top, u, pmax, tmax):
    """Run orc_cz_force on one CZ whose bottom nodes coincide with its top nodes at rest."""
    cs, ns = 1, 6
    x0 = np.r_[x_top, x_top]
    dn = np.zeros(18 * ns)
    for k in range(3):
        dn[k * ns:k * ns + 6] = x0[:, k]
        dn[(3 + k) * ns:(3 + k) * ns + 6] = u[:, k]
    icz = np.zeros(58 * cs, np.int32)
    icz[4:10] = np.arange(6)
    slots = np.arange(36).reshape(6, 6)
    icz[10:46] = slots.ravel()
    icz[46:52] = np.arange(6)
    dcz = np.zeros(16 * cs)
    dcz[0:3] = pmax
    dcz[3:6] = tmax
    A = np.zeros(36 * 9)
    b = np.zeros(18)
    orc.cz_force(dcz, icz, dn, A, b, 1e-4, 1, cs, ns)
    K = np.zeros((18, 18))
    for i in range(6):
        for j in range(6):
            K[3 * i:3 * i + 3, 3 * j:3 * j + 3] = A[9 * slots[i, j]:9 * slots[i, j] + 9].reshape(3, 3)
    return K, b, icz, dcz


def test_cz_rigid_motion_no_force(orc, C):
    rng = np.random.default_rng(2)
    x = rng.normal(size=(3, 3)) * 0.01
    Q = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    xx = np.r_[x, x]
    u = (xx @ Q.T + 0.003) - xx
    K, b, icz, _ = _one_cz(orc, C, x, u, [0.3 * C.deln] * 3, [0.2 * C.delt] * 3)
    assert icz[3] == 0
    assert np.max(np.abs(b)) <= 1e-9 * C.f_tn * 1e-4


def test_cz_secant_branch_rhs_equals_minus_K_u(orc, C):
    """Reloading branch with a damaged history: T = D Delta exactly, so rhs = -K u when the
    midplane frame is unchanged by the motion (top face lifted along its normal, slid in plane
    by less than tmax)."""
    rng = np.random.default_rng(3)
    x = np.array([[0.0, 0.0, 0.0], [0.01, 0.0, 0.0], [0.002, 0.009, 0.0]])
    R, _ = orc.cz_rotation(x.ravel())
    n = R[2]
    t1 = R[0]
    pm, tm = 0.4 * C.deln, 0.3 * C.delt
    for d_n, d_t in ((0.1, 0.0), (0.05, 0.1), (-0.02, 0.05)):
        u = np.zeros((6, 3))
        u[:3] = d_n * C.deln * n + d_t * C.delt * t1        # top face rigidly displaced
        K, b, icz, _ = _one_cz(orc, C, x, u, [pm] * 3, [tm] * 3)
        assert icz[3] == 0
        assert rel_err(b, -K @ u.ravel()) <= 1e-9
        assert rel_err(K, K.T) <= 1e-12
        # block structure of B = [s x I, -s x I]: K_top,bot = -K_top,top for coincident faces
        assert rel_err(K[:9, 9:], -K[:9, :9]) <= 1e-12
        assert rel_err(K[9:, 9:], K[:9, :9]) <= 1e-12


def test_law_leaves_logged_over_crafted_states(orc, lib, specimen_consts):
    """The oracle's leaf log (orc_set_branch_log) over the inputs the device branch test uses
    (a small specimen's displacements scaled 0.3x to 10x, load histories of 0 or two
    magnitudes per Gauss point): every reachable leaf appears, and the five that need pmax == 0
    without normal softening or tmax == 0 without tangential softening never do, although
    zero histories are among the inputs."""
    from icflow_amd._lib import ptr
    from icflow_amd.gpu_functionality import CSR, stride_of
    from workloads import make_state, small_specimen
    orc.set_constants(specimen_consts)
    col = small_specimen(cells=(6, 5, 4), grains=6, seed=0)
    st = make_state(col, specimen_consts, seed=1)
    alt = np.ascontiguousarray(col.alt, np.int32)
    csr = CSR(lib).build(alt[col.elastic], alt[col.czs])
    C, Nn = len(col.czs), len(col.x0)
    cs, ns = stride_of(C), stride_of(Nn)
    icz = np.zeros(58 * cs, np.int32)
    lib.icfh_pack_czs(csr.h, ptr(np.ascontiguousarray(col.czs, np.int32)), ptr(alt), C, cs, ptr(icz))
    deln, delt = specimen_consts.deln, specimen_consts.delt
    rng = np.random.default_rng(11)
    seen = set()
    for scale in (0.3, 1.0, 3.0, 10.0):
        for hist in (0, 1, 2):
            dn = np.zeros(18 * ns)
            for k in range(3):
                dn[k * ns:k * ns + Nn] = col.x0[:, k]
                dn[(3 + k) * ns:(3 + k) * ns + Nn] = st.un[:, k] * scale
            dcz = np.zeros(16 * cs)
            if hist:
                for j in range(3):
                    dcz[j * cs:j * cs + C] = rng.choice([0.0, 0.02 * hist * deln, 0.2 * hist * deln], C)
                    dcz[(3 + j) * cs:(3 + j) * cs + C] = rng.choice([0.0, 0.02 * hist * delt, 0.2 * hist * delt], C)
            log = orc.cz_branches(dcz, icz, dn, np.zeros(9 * csr.nnz), np.zeros(3 * csr.N), 1e-4, C, cs, ns)
            seen |= set(log[log >= 0].tolist())
    assert seen == {0, 1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 16}

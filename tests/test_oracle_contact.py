"""Pins the oracle's contact math (orc_pt, orc_dtn, orc_bvalues/ctest, orc_narrow_phase,
orc_find_closest_face, orc_collision_response; restating collision_kernels9):

 - squared point-triangle distance equals the minimum over the closed triangle computed a
   different way (interior projection, three segment projections, three vertices);
 - the 12-gradient and 12x12 Hessian equal central differences of the distance function in
   every one of the eight region branches (a dropped chain-rule term or a transposed index in
   the point-plane Hessian fails this);
 - the returned (s, t) are the barycentric coordinates of the closest point;
 - narrow-phase containment equals a strict barycentric test by numpy.linalg.solve, and pairs
   sharing a node give 0; the closest face is the brute-force first minimum;
 - the contact force obeys momentum balance and its tangent is the (k/2) Hessian.
"""
import numpy as np
import pytest

from conftest import rel_err


def closest_sq(p, a, b, c):
    """Closed-triangle squared distance: interior projection if inside, else min over edges."""
    e0, e1 = b - a, c - a
    n = np.cross(e0, e1)
    best = np.inf
    # interior
    M = np.array([e0, e1]).T
    st, *_ = np.linalg.lstsq(M, p - a, rcond=None)
    if st[0] >= 0 and st[1] >= 0 and st.sum() <= 1:
        best = np.dot(p - a, n) ** 2 / np.dot(n, n)
    for q0, q1 in ((a, b), (b, c), (c, a)):
        d = q1 - q0
        t = np.clip(np.dot(p - q0, d) / np.dot(d, d), 0, 1)
        best = min(best, np.sum((q0 + t * d - p) ** 2))
    return best


def rand_cfg(rng, kind):
    tri = rng.normal(size=(3, 3))
    a, b, c = tri
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n)
    if kind == "interior":
        w = rng.dirichlet([2, 2, 2])
        p = w @ tri + n * rng.normal() * 0.5
    else:
        p = (a + b + c) / 3 + rng.normal(size=3) * 2.5
    return np.r_[p, tri.ravel()]


def test_distance_and_barycentrics(orc):
    rng = np.random.default_rng(0)
    seen = set()
    for i in range(600):
        x = rand_cfg(rng, "interior" if i % 3 == 0 else "outside")
        d, fd, sd, s, t, br = orc.pt(x)
        seen.add(br)
        p, a, b, c = x[:3], x[3:6], x[6:9], x[9:12]
        ref = closest_sq(p, a, b, c)
        assert abs(d - ref) <= 1e-10 * max(ref, 1e-30) + 1e-14, (br, d, ref)
        q = (1 - s - t) * a + s * b + t * c
        assert abs(np.sum((q - p) ** 2) - d) <= 1e-9 * max(d, 1e-30) + 1e-14
    assert seen == set(range(8)), seen


def _d2(orc, x):
    return orc.pt(x)[0]


def test_gradient_and_hessian_every_branch(orc):
    rng = np.random.default_rng(1)
    hit = {}
    tries = 0
    while len(hit) < 8 or min(hit.values()) < 5:
        tries += 1
        assert tries < 20000
        x = rand_cfg(rng, "interior" if tries % 3 == 0 else "outside")
        d, fd, sd, s, t, br = orc.pt(x)
        h = 1e-5
        # stay inside one branch for all perturbations
        if any(orc.pt(x + h * np.eye(12)[k] * sg)[5] != br for k in range(12) for sg in (-1, 1)):
            continue
        g = np.array([(_d2(orc, x + h * e) - _d2(orc, x - h * e)) / (2 * h) for e in np.eye(12)])
        assert rel_err(fd, g) <= 1e-7, br
        H = np.zeros((12, 12))
        for k, e in enumerate(np.eye(12)):
            H[:, k] = (orc.pt(x + h * e)[1] - orc.pt(x - h * e)[1]) / (2 * h)
        assert rel_err(sd, H) <= 1e-6, br
        assert rel_err(sd, sd.T) <= 1e-9
        hit[br] = hit.get(br, 0) + 1


def dtn_expected(p, a, b, c):
    """What dist_node2tri.h:14-129 returns: the true squared distance, except in its region-6
    branch with a + d <= b + e (dist_node2tri.h:103-106), where the reference evaluates the
    edge parameter as clamp(-e/c) instead of Eberly's clamp(-d/a) on the edge t = 0."""
    e0, e1, v0 = b - a, c - a, a - p
    A, B, Cc, D, Ee = e0 @ e0, e0 @ e1, e1 @ e1, e0 @ v0, e1 @ v0
    det, s, t = A * Cc - B * B, B * Ee - Cc * D, B * D - A * Ee
    if not (s + t < det) and s >= 0 and t < 0 and not (A + D > B + Ee):
        s = np.clip(-Ee / Cc, 0, 1)
        return float(np.sum((a + s * e0 - p) ** 2)), True
    return closest_sq(p, a, b, c), False


def test_dtn_matches_distance_and_reference_quirk(orc):
    rng = np.random.default_rng(2)
    quirk = 0
    for _ in range(600):
        x = rand_cfg(rng, "outside")
        p, f = x[:3], x[3:]
        ref, q = dtn_expected(p, f[:3], f[3:6], f[6:])
        quirk += q
        assert abs(orc.dtn(f, p) - ref) <= 1e-10 * max(ref, 1e-30) + 1e-14
    assert quirk > 5


def test_containment_test(orc):
    rng = np.random.default_rng(3)
    n_in = 0
    for _ in range(2000):
        tet = rng.normal(size=(4, 3))
        q = rng.normal(size=3) * 0.7
        bm = orc.bvalues(tet.ravel())
        M = (tet[1:] - tet[0]).T
        lam = np.linalg.solve(M, q - tet[0])
        strict = (lam > 1e-10).all() and lam.sum() < 1 - 1e-10
        margin = min(lam.min(), 1 - lam.sum())
        if abs(margin) < 1e-8:
            continue                      # too close to a face to decide in floating point
        got = orc.ctest(bm, *(q - tet[0]))
        assert got == strict
        n_in += got
    assert n_in > 40


def test_narrow_phase_masks(orc):
    rng = np.random.default_rng(4)
    ns, es = 128, 128
    nodes = rng.normal(size=(16, 3)) * 0.5
    nodes[8:] = nodes[:8] * 0.3 + rng.normal(size=(8, 3)) * 0.2
    dn = np.zeros(18 * ns)
    for k in range(3):
        dn[(15 + k) * ns:(15 + k) * ns + 16] = nodes[:, k]
    tets = np.array([[0, 1, 2, 3], [4, 5, 6, 7], [8, 9, 10, 11], [12, 13, 14, 15], [0, 9, 10, 11]])
    ie = np.zeros(6 * es, np.int32)
    for i in range(4):
        ie[i * es:i * es + len(tets)] = tets[:, i]
    pairs = [(a, b) for a in range(len(tets)) for b in range(len(tets)) if a != b]
    nl = np.array(pairs, np.int32).ravel().copy()
    orc.narrow_phase(len(pairs), dn, ie, nl, es, ns)
    for k, (a, b) in enumerate(pairs):
        if set(tets[a]) & set(tets[b]):
            assert nl[2 * k] == 0
            continue
        want = 0
        for side, (A, Bt) in enumerate(((a, b), (b, a))):
            PA = nodes[tets[A]]
            M = (PA[1:] - PA[0]).T
            for i, v in enumerate(tets[Bt]):
                lam = np.linalg.solve(M, nodes[v] - PA[0])
                if (lam > 1e-10).all() and lam.sum() < 1 - 1e-10:
                    want |= (1 << i) << (4 * side)
        assert nl[2 * k] == want, (a, b)


def test_closest_face_first_minimum(orc):
    rng = np.random.default_rng(5)
    ns, es, fs, ts = 128, 128, 128, 128
    nodes = rng.normal(size=(40, 3))
    dn = np.zeros(18 * ns)
    for k in range(3):
        dn[(15 + k) * ns:(15 + k) * ns + 40] = nodes[:, k]
    faces = rng.integers(0, 40, size=(30, 3))
    faces = faces[[len(set(f)) == 3 for f in faces]]
    faces = np.r_[faces, faces[:2]]                      # duplicates: ties, first must win
    F = len(faces)
    ifc = np.zeros(3 * fs, np.int32)
    for v in range(3):
        ifc[v * fs:v * fs + F] = faces[:, v]
    nadj = 6
    ie = np.zeros((5 + nadj) * es, np.int32)
    adj = [rng.choice(F, size=nadj, replace=False) for _ in range(10)]
    for t in range(10):
        ie[4 * es + t] = nadj
        for k in range(nadj):
            ie[(5 + k) * es + t] = adj[t][k]
    nT = 10
    nl2 = np.zeros(4 * ts, np.int32)
    nl2[:nT] = rng.integers(0, 40, size=nT)
    nl2[ts:ts + nT] = np.arange(nT)
    src = nl2.copy()
    orc.find_closest_face(nT, ts, nl2, dn, ie, ifc, ns, es, fs)
    for u in range(nT):
        p = nodes[src[u]]
        # selection rule on the (separately pinned) dtn values: first strict minimum wins
        d = [orc.dtn(nodes[faces[f]].ravel(), p) for f in adj[src[ts + u]]]
        best = adj[src[ts + u]][int(np.argmin(d))]
        assert list(nl2[[u + ts, u + 2 * ts, u + 3 * ts]]) == list(faces[best])


def test_collision_response_force_and_tangent(orc):
    rng = np.random.default_rng(6)
    k = 50.0
    for _ in range(50):
        x = rand_cfg(rng, "interior")
        d, fd, sd, s, t, br = orc.pt(x)
        cs, ns = 1, 4
        dn = np.zeros(18 * ns)
        for kk in range(3):
            dn[(15 + kk) * ns:(15 + kk) * ns + 4] = x.reshape(4, 3)[:, kk]
        icr = np.zeros(28, np.int32)
        icr[0:4] = np.arange(4)
        slots = np.arange(16).reshape(4, 4)
        icr[4:20] = slots.ravel()
        icr[20:24] = np.arange(4)
        A = np.zeros(16 * 9)
        b = np.zeros(12)
        orc.collision_response(icr, dn, 1e-4, A, b, 1, cs, ns, k, 1e-15)
        Fnode = dn[12 * ns:15 * ns].reshape(3, ns).T
        assert np.allclose(Fnode.ravel(), b)
        assert np.allclose(b.reshape(4, 3).sum(axis=0), 0, atol=1e-12 * np.abs(b).max())   # momentum
        assert np.allclose(b[:3], -0.5 * k * fd[:3])
        w = np.array([1 - s - t, s, t])
        for i in range(3):
            assert np.allclose(b[3 + 3 * i:6 + 3 * i], w[i] * 0.5 * k * fd[:3])
        Hs = np.triu(sd) + np.triu(sd, 1).T
        for i in range(4):
            for j in range(4):
                blk = A[9 * slots[i, j]:9 * slots[i, j] + 9].reshape(3, 3)
                assert np.allclose(blk, 0.5 * k * Hs[3 * i:3 * i + 3, 3 * j:3 * j + 3], rtol=1e-14, atol=0)

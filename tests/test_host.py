"""Host-side helpers of libicflow_hip.so (include/icflow_host.h) and the specimen generator,
on the CPU only: model constants, BCSR pattern and slot lookup, pcsr packing, NL2set
de-duplication, the k-DOP broad phase, and the CZ insertion conventions."""
import ctypes

import numpy as np
import pytest

from icflow_amd import _lib, specimens
from icflow_amd._lib import ptr
from icflow_amd.gpu_functionality import CSR, broad_phase, stride_of


def test_constants_match_model_parameters(lib):
    c = _lib.consts_from_prms(**_lib.SPECIMEN_PRMS)
    Y, nu = 9e9, 0.3
    lam = Y * nu / ((1 + nu) * (1 - 2 * nu))
    mu = Y / (2 * (1 + nu))
    E = np.array(c.E[:]).reshape(6, 6)
    assert np.allclose(E[:3, :3], lam + 2 * mu * np.eye(3), rtol=1e-14)
    assert np.allclose(np.diag(E)[3:], mu, rtol=1e-14)
    # uniaxial stress: sigma = E eps with eps = (1, -nu, -nu) / Y gives (1, 0, 0)
    s = E @ np.array([1, -nu, -nu, 0, 0, 0]) / Y
    assert np.allclose(s, [1, 0, 0, 0, 0, 0], atol=1e-15)
    # PPR exponents (Park-Paulino-Roesler): m = a (a-1) l^2 / (1 - a l^2)
    assert abs(c.p_m - 4 * 3 * 0.02 ** 2 / (1 - 4 * 0.02 ** 2)) < 1e-16
    sf = np.array(c.sf[:]).reshape(3, 3)
    assert np.allclose(sf.sum(axis=0), 1)
    B = np.array(c.B[:]).reshape(3, 3, 18)
    for g in range(3):
        assert np.allclose(B[g][:, :9], np.kron(sf[:, g], np.eye(3)))
        assert np.allclose(B[g][:, 9:], -np.kron(sf[:, g], np.eye(3)))


def test_set_constants_rejects_unsupported_B(lib):
    """icf_set_constants validates the B table before touching the device; no GPU needed to
    reach that check, but icf_create needs one -- so only the host copy is checked here."""
    c = _lib.consts_from_prms()
    B = np.array(c.B[:])
    assert np.count_nonzero(B) == 3 * 3 * 6


def _units(rng, n_nodes, n_units, k):
    return np.array([rng.choice(n_nodes, size=k, replace=False) for _ in range(n_units)])


def test_csr_pattern_and_lookup(lib):
    rng = np.random.default_rng(0)
    nn = 60
    alt = np.arange(nn)
    alt[rng.choice(nn, 8, replace=False)] = -1
    act = np.flatnonzero(alt >= 0)
    alt[act] = np.arange(len(act))
    el, cz, im = _units(rng, nn, 40, 4), _units(rng, nn, 10, 6), _units(rng, nn, 7, 4)
    el = np.r_[el, np.arange(nn).reshape(-1, 4)]        # every node in some element, as in a mesh
    csr = CSR(lib).build(alt[el], alt[cz], alt[im])
    want = set()
    for units, strict in ((el, False), (cz, False), (im, True)):
        for u in alt[units]:
            for a in u:
                for b in u:
                    if a >= 0 and b >= 0 and (b > a or (b == a and not strict)):
                        want.add((a, b))
    got = {(i, int(csr.cols[s])) for i in range(csr.N) for s in range(csr.rows[i], csr.rows[i + 1])}
    assert got == want
    # CSRDictionary.Assert: rows increasing, diagonal first, columns strictly increasing
    assert csr.rows[0] == 0 and csr.rows[-1] == csr.nnz
    assert (np.diff(csr.rows) > 0).all()
    for i in range(csr.N):
        c = csr.cols[csr.rows[i]:csr.rows[i + 1]]
        assert c[0] == i and (np.diff(c) > 0).all()
    for (i, j) in list(want)[:50]:
        s = lib.icfh_csr_lookup(csr.h, int(i), int(j))
        assert csr.cols[s] == j and csr.rows[i] <= s < csr.rows[i + 1]
    assert lib.icfh_csr_lookup(csr.h, 0, -3) == -1


def test_pack_elems_and_czs(lib):
    rng = np.random.default_rng(1)
    nn = 50
    alt = np.arange(nn)
    alt[[3, 7, 11]] = -1
    act = np.flatnonzero(alt >= 0)
    alt[act] = np.arange(len(act))
    el, cz = _units(rng, nn, 30, 4), _units(rng, nn, 9, 6)
    csr = CSR(lib).build(alt[el], alt[cz])
    es, cs = stride_of(len(el)), stride_of(len(cz))
    ie = np.full(28 * es, -7, np.int32)
    icz = np.full(58 * cs, -7, np.int32)
    a32 = alt.astype(np.int32)
    assert lib.icfh_pack_elems(csr.h, ptr(el.astype(np.int32)), ptr(a32), len(el), es, ptr(ie)) == 0
    assert lib.icfh_pack_czs(csr.h, ptr(cz.astype(np.int32)), ptr(a32), len(cz), cs, ptr(icz)) == 0
    for units, buf, st, k, slot0, row0, node0 in ((el, ie, es, 4, 4, 20, 0), (cz, icz, cs, 6, 10, 46, 4)):
        for u, nodes in enumerate(units):
            for i in range(k):
                assert buf[u + st * (node0 + i)] == nodes[i]
                assert buf[u + st * (row0 + i)] == alt[nodes[i]]
                for j in range(k):
                    ai, aj = alt[nodes[i]], alt[nodes[j]]
                    s = buf[u + st * (slot0 + i * k + j)]
                    if ai >= 0 and aj >= 0 and aj >= ai:
                        assert csr.cols[s] == aj and csr.rows[ai] <= s < csr.rows[ai + 1]
                    else:
                        assert s == -1
    # flags the reference resets on upload (GPU_Functionality.cs:384-386)
    assert (icz[:len(cz)] == 0).all() and (icz[2 * cs:2 * cs + len(cz)] == 0).all()
    # a pair the pattern does not hold is an error (the reference's Dictionary lookup throws)
    missing = None
    for a in range(len(act)):
        for b in range(a + 1, len(act)):
            if lib.icfh_csr_lookup(csr.h, a, b) < 0:
                missing = (a, b)
                break
        if missing:
            break
    na, nb = np.flatnonzero(alt == missing[0])[0], np.flatnonzero(alt == missing[1])[0]
    others = [n for n in act if n not in (na, nb)][:2]
    bad = np.array([[na, nb, *others]], np.int32)
    assert lib.icfh_pack_elems(csr.h, ptr(bad), ptr(a32), 1, 128, ptr(np.zeros(28 * 128, np.int32))) == -5


def test_nl2set_first_insertion_order(lib):
    rng = np.random.default_rng(2)
    S, es = 12, 128
    tet_nodes = rng.integers(0, 30, size=(S, 4))
    ie = np.zeros(4 * es, np.int32)
    for i in range(4):
        ie[i * es:i * es + S] = tet_nodes[:, i]
    pairs = np.array([(a, b) for a in range(S) for b in range(a + 1, S)][:40], np.int32)
    masks = np.zeros(2 * len(pairs), np.int32)
    masks[0::2] = rng.integers(0, 256, size=len(pairs)) * (rng.uniform(size=len(pairs)) < 0.4)
    want = {}
    for k, (a, b) in enumerate(pairs):        # GPU_Functionality.cs:567-580
        bits = masks[2 * k]
        for i in range(4):
            if bits & (1 << i):
                want.setdefault((int(tet_nodes[b][i]), int(a)), None)
        for i in range(4):
            if bits & (16 << i):
                want.setdefault((int(tet_nodes[a][i]), int(b)), None)
    cap = 512
    out = np.zeros(2 * cap, np.int32)
    n = ctypes.c_int()
    assert lib.icfh_nl2set(ptr(masks), ptr(pairs.ravel().copy()), len(pairs), ptr(ie), es, ptr(out), cap, cap,
                           ctypes.byref(n)) == 0
    got = list(zip(out[:n.value].tolist(), out[cap:cap + n.value].tolist()))
    assert got == list(want)


def test_kdop_broad_phase_is_exact(lib):
    col = specimens.make_specimen((3, 3, 2), (0.03, 0.03, 0.02), 3, 0, platens=(3, 3)).finalize()
    c = _lib.consts_from_prms(**_lib.SPECIMEN_PRMS)
    st = specimens.synthetic_state(col, c, 0, platen_dz=0.003)
    S = len(col.surf)
    es, ns = stride_of(S), stride_of(len(col.x0))
    ie = np.zeros(4 * es, np.int32)
    for i in range(4):
        ie[i * es:i * es + S] = col.surf[:, i]
    dn = np.zeros(18 * ns)
    for k in range(3):
        dn[(15 + k) * ns:(15 + k) * ns + len(col.x0)] = st.xc[:, k]
    pts = st.xc[col.surf]                         # (S, 4, 3)
    x, y, z = pts[..., 0], pts[..., 1], pts[..., 2]
    pr = np.stack([x, y, z, x + y, x + z, y + z, x - y, x - z, y - z, x + y - z, x + z - y, y + z - x], -1)
    lo, hi = pr.min(axis=1), pr.max(axis=1)
    for excl in (False, True):
        want = []
        for a in range(S):
            ov = (~(lo[a] > hi[a + 1:]) & ~(hi[a] < lo[a + 1:])).all(axis=1)
            for b in np.flatnonzero(ov) + a + 1:
                if excl and set(col.surf[a]) & set(col.surf[b]):
                    continue
                want.append((a, b))
        got = broad_phase(lib, dn, ns, ie, es, S, exclude_shared=excl)
        assert [tuple(p) for p in got.tolist()] == want


def test_specimen_cz_conventions():
    col = specimens.make_specimen((4, 3, 3), (0.04, 0.03, 0.03), 4, 3).finalize()
    x, T, czs = col.x0, col.elastic, col.czs
    assert (specimens.jdet(x[T]) > 0).all()
    assert np.array_equal(x[czs[:, :3]], x[czs[:, 3:]])           # zero thickness, paired vertices
    m = col.meshes[0]
    # every CZ joins two tets of different grains, and its two faces are tet faces
    tet_faces = {}
    for t, tet in enumerate(T):
        for f in specimens.CONV:
            tet_faces[tuple(sorted(tet[f]))] = t
    for c, (f0, f1) in enumerate(m.cz_faces):
        t0 = tet_faces[tuple(sorted(m.faces[f0]))]
        t1 = tet_faces[tuple(sorted(m.faces[f1]))]
        assert m.grain[t0] != m.grain[t1] and t0 < t1
        assert set(czs[c, :3]) == set(m.faces[f0]) and set(czs[c, 3:]) == set(m.faces[f1])
    # the CZ normal points into e0 (top side): separating the grains opens the CZ positively
    p = x[czs[:, :3]]
    n = np.cross(p[:, 1] - p[:, 0], p[:, 0] - p[:, 2])
    cent0 = np.array([x[T[tet_faces[tuple(sorted(m.faces[f0]))]]].mean(axis=0) for f0, _ in m.cz_faces])
    assert (np.einsum("ij,ij->i", n, cent0 - p.mean(axis=1)) > 0).all()
    # each node in the grid appears once per grain of its tets when it lies on a CZ face
    assert len(col.x0) > 4 * 5 * 4 * 4 / 4


def test_batched_specimen_offsets(lib):
    """specimen_offsets of a batch: each specimen's tets, CZs and active rows form one run, in
    specimen order, and the rows' BCSR blocks stay inside the specimen's block range (the
    block-diagonal layout icf_assemble_batched checks on the device)."""
    from icflow_amd import specimens
    from icflow_amd.gpu_functionality import CSR
    col = specimens.lognormal_cubes(4, 3, median_tets=800, min_tets=300, max_tets=3000, sigma=0.9, grains=(3, 6),
                                    size=0.01)
    tet, cz, row = specimens.specimen_offsets(col)
    n = len(col.meshes)
    assert len(tet) == len(cz) == len(row) == n + 1
    assert tet[-1] == len(col.elastic) and cz[-1] == len(col.czs) and row[-1] == int((~col.anchored).sum())
    for s in range(n):
        assert np.all(col.node_mesh[col.elastic[tet[s]:tet[s + 1]]] == s)
        assert np.all(col.node_mesh[col.czs[cz[s]:cz[s + 1]]] == s)
        act = np.flatnonzero((col.node_mesh == s) & ~col.anchored)
        assert np.array_equal(np.sort(col.alt[act]), np.arange(row[s], row[s + 1]))
    csr = CSR(lib).build(col.alt[col.elastic], col.alt[col.czs])
    br = np.repeat(np.arange(csr.N), np.diff(csr.rows))
    blk = csr.rows[row]
    for s in range(n):
        cols = csr.cols[blk[s]:blk[s + 1]]
        assert np.all((cols >= row[s]) & (cols < row[s + 1])), s
        assert np.all((br[blk[s]:blk[s + 1]] >= row[s]) & (br[blk[s]:blk[s + 1]] < row[s + 1]))

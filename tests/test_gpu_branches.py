"""Branch coverage of the device kernels against the oracle.

The point-triangle distance of the contact response (pt, point_triange_v2.h:589-842) has eight
leaves, the oracle's `branch` codes 0-7: the face interior, the three vertices, and the three
edges (edge B-C reached by two paths).  The PPR law (cohesive_law_thermodynamically_consistent,
cz_kernels.h:37-174) has the leaves the oracle logs (`orc_set_branch_log`); twelve are reachable:
opening beyond deln or delt; in contact, pmax non-zero or zero times tangential loading,
unloading or failure; away from contact, both softening, tangential softening only with
pmax != 0, normal softening only with tmax != 0, neither softening with both non-zero, and
failure.  The other five need pmax == 0 without normal softening or tmax == 0 without
tangential softening, which the law's own tests rule out (nsoft = opn >= pmax holds for opn >= 0
when pmax == 0, and tsoft = opt >= tmax for the opening magnitude opt >= 0 when tmax == 0).

Each test builds inputs that reach every leaf, shows that with the oracle's classification,
and compares the device result with the oracle leaf group by leaf group (contact) or over
inputs covering all leaves together (CZs)."""
import ctypes

import numpy as np
import pytest

from conftest import rel_err
from icflow_amd import _lib
from workloads import DIST_EPS, H, PENALTY_K, Pipeline, make_state, small_specimen

pytestmark = pytest.mark.gpu

XC, FF = 15, 12


def test_every_point_triangle_branch_matches_oracle(orc, specimen_consts):
    col = small_specimen(seed=3)
    st = make_state(col, specimen_consts, seed=4)
    p = Pipeline(col, specimen_consts, st)
    gf, lib, ctx = p.gf, p.gf.lib, p.gf.ctx
    ns = gf.nd_stride
    alt = np.asarray(col.alt)
    free = np.flatnonzero(alt >= 0)
    tri = next(e[:3] for e in np.asarray(col.elastic) if (alt[e] >= 0).all())   # a face of a free tet
    pts = free[~np.isin(free, tri)][:400]
    rng = np.random.default_rng(7)
    dn = gf.c_dn.copy()
    pos = lambda n: np.array([dn[n + ns * (XC + k)] for k in range(3)])   # noqa: E731
    # the triangle reshaped to an obtuse angle at B (about 140 degrees, slightly tilted): the
    # edge B-C leaf reached from region 6 needs it
    A0 = pos(tri[0])
    Lt = np.linalg.norm(pos(tri[1]) - A0)
    B0, C0 = A0 + Lt * np.array([1.0, 0.0, 0.0]), A0 + Lt * np.array([1.6, 0.5, 0.05])
    for n, x in ((tri[1], B0), (tri[2], C0)):
        for k in range(3):
            dn[n + ns * (XC + k)] = x[k]
    e1, e2 = B0 - A0, C0 - A0
    nrm = np.cross(e1, e2)
    nrm /= np.linalg.norm(nrm)
    L = np.linalg.norm(e1)
    for n in pts:     # points around the triangle: in-plane (-3..3, -3..3), off-plane +-1 edge
        u, v = rng.uniform(-3.0, 3.0, 2)
        x = A0 + u * e1 + v * e2 + rng.uniform(-1.0, 1.0) * L * nrm
        for k in range(3):
            dn[n + ns * (XC + k)] = x[k]
    impacts = np.array([[n, tri[0], tri[1], tri[2]] for n in pts], np.int32)
    br = np.array([orc.pt(np.concatenate([pos(n) for n in imp]))[5] for imp in impacts])
    assert set(br.tolist()) == set(range(8)), np.unique(br, return_counts=True)
    gf.CreateStructure(impacts)
    nnz, N = gf.csr.nnz, gf.csr.N
    d_A, d_b = gf.buf("br_A", 9 * nnz, np.float64), gf.buf("br_b", 3 * N, np.float64)
    for code in range(8):
        sel = impacts[br == code][:40]
        icr, cr = gf.pack_icr(sel)
        d_icr = gf.buf(f"br_icr{code}", icr.size, np.int32)
        d_icr.upload(icr)
        dn_dev = dn.copy()
        dn_dev[FF * ns:(FF + 3) * ns] = 0
        gf.g_dn.upload(dn_dev)
        d_A.zero()
        d_b.zero()
        _lib.check(lib.icf_collision_response(ctx, d_icr.p, gf.g_dn.p, H, d_A.p, d_b.p, len(sel), cr, ns,
                                              PENALTY_K, DIST_EPS), ctx, "icf_collision_response")
        A_o, b_o, dn_o = np.zeros(9 * nnz), np.zeros(3 * N), dn_dev.copy()
        orc.collision_response(icr, dn_o, H, A_o, b_o, len(sel), cr, ns, PENALTY_K, DIST_EPS)
        assert np.abs(A_o).max() > 0, code
        assert rel_err(d_A.download(9 * nnz), A_o) <= 1e-12, code
        assert rel_err(d_b.download(3 * N), b_o) <= 1e-12, code
        assert rel_err(gf.g_dn.download(3 * ns, offset_elems=FF * ns), dn_o[FF * ns:(FF + 3) * ns]) <= 1e-12, code
    gf.close()


REACHABLE_LAW_LEAVES = {0, 1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 16}


def test_every_reachable_ppr_leaf_matches_oracle(orc, specimen_consts):
    """Displacements scaled from small to past failure and random load histories (pmax, tmax
    each 0 or one of two magnitudes per Gauss point): icf_cz_force on the device against the
    oracle's cz_force on each input -- blocks, rhs, tentative flags bit for bit, tentative
    history -- and the union of the leaves the oracle logs is every reachable one."""
    col = small_specimen(cells=(6, 5, 4), grains=6, seed=0)
    st = make_state(col, specimen_consts, seed=1)
    p = Pipeline(col, specimen_consts, st)
    gf, lib, ctx = p.gf, p.gf.lib, p.gf.ctx
    orc.set_constants(specimen_consts)
    C, cs, ns = len(col.czs), gf.cz_stride, gf.nd_stride
    nnz, N = gf.csr.nnz, gf.csr.N
    deln, delt = specimen_consts.deln, specimen_consts.delt
    d_A, d_b = gf.buf("pl_A", 9 * nnz, np.float64), gf.buf("pl_b", 3 * N, np.float64)
    rng = np.random.default_rng(11)
    seen = set()
    UN = 3
    for scale in (0.3, 1.0, 3.0, 10.0):
        for hist in (0, 1, 2):
            dn = gf.c_dn.copy()
            dn[UN * ns:(UN + 3) * ns] *= scale
            dcz = gf.c_dcz.copy()
            if hist:
                for j in range(3):
                    dcz[j * cs:j * cs + C] = rng.choice([0.0, 0.02 * hist * deln, 0.2 * hist * deln], C)
                    dcz[(3 + j) * cs:(3 + j) * cs + C] = rng.choice([0.0, 0.02 * hist * delt, 0.2 * hist * delt], C)
            icz = gf.c_icz.copy()
            A_o, b_o = np.zeros(9 * nnz), np.zeros(3 * N)
            log = orc.cz_branches(dcz, icz, dn, A_o, b_o, H, C, cs, ns)
            seen |= set(log[log >= 0].tolist())
            dcz_o, icz_o, dn_o = dcz.copy(), icz.copy(), dn.copy()
            orc.cz_force(dcz_o, icz_o, dn_o, A_o, b_o, H, C, cs, ns)
            gf.g_dn.upload(dn)
            gf.g_dcz.upload(dcz)
            gf.g_icz.upload(icz)
            d_A.zero()
            d_b.zero()
            _lib.check(lib.icf_cz_force(ctx, gf.g_dcz.p, gf.g_icz.p, gf.g_dn.p, d_A.p, d_b.p, H, C, cs, ns), ctx,
                       "icf_cz_force")
            tag = (scale, hist)
            assert rel_err(d_A.download(9 * nnz), A_o) <= 1e-12, tag
            assert rel_err(d_b.download(3 * N), b_o) <= 1e-12, tag
            gi, gd = gf.g_icz.download(icz.size), gf.g_dcz.download(dcz.size)
            for f in (1, 2, 3):       # tentative contact, damaged, failed
                assert np.array_equal(gi[f * cs:f * cs + C], icz_o[f * cs:f * cs + C]), (tag, f)
            for f in range(6, 12):    # tentative pmax, tmax
                assert rel_err(gd[f * cs:f * cs + C], dcz_o[f * cs:f * cs + C]) <= 1e-12, (tag, f)
    assert seen == REACHABLE_LAW_LEAVES, sorted(seen)
    gf.close()

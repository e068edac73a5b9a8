"""Parity of the HIP path (through the C ABI) against the CPU oracle on seeded specimens.

Tolerances: FP64 mode, 1e-9 normwise relative per assembled array (the assembled BCSR
values, rhs and node forces); integer outputs (flags, masks, closest faces) bit-exact.
Mixed mode: tangent blocks within 1e-5 normwise relative (FP32 blocks, FP64 accumulation),
residual and node forces within 1e-9 (computed in FP64).
"""
import ctypes

import numpy as np
import pytest

from conftest import rel_err
from icflow_amd import _lib
from workloads import H, Pipeline, make_state, oracle_assemble, small_specimen

pytestmark = pytest.mark.gpu
TOL64 = 1e-9


@pytest.fixture(scope="module")
def col():
    return small_specimen()


def _compare(got, ref, tol_vals=TOL64, tol_rhs=TOL64):
    errs = dict(vals=rel_err(got["vals"], ref["vals"]), rhs=rel_err(got["rhs"], ref["rhs"]),
                F=rel_err(got["F"], ref["F"]))
    print("normwise rel errors:", errs, flush=True)
    assert errs["vals"] <= tol_vals
    assert errs["rhs"] <= tol_rhs
    assert errs["F"] <= tol_rhs
    return errs


def _cz_fields(gf, got, ref, C):
    cs = gf.cz_stride
    gi, ri = got["icz"], ref["icz"]
    for f in (1, 2, 3):   # tentative contact, damaged, failed: bit-exact
        assert np.array_equal(gi[f * cs:f * cs + C], ri[f * cs:f * cs + C]), f"icz field {f}"
    gd, rd = got["dcz"], ref["dcz"]
    for f in range(6, 12):   # tentative pmax / tmax
        assert rel_err(gd[f * cs:f * cs + C], rd[f * cs:f * cs + C]) <= 1e-12, f"dcz field {f}"
    for f in range(12, 16):  # averages (written for non-failed CZs)
        assert rel_err(gd[f * cs:f * cs + C], rd[f * cs:f * cs + C]) <= TOL64, f"dcz field {f}"


def test_fused_assembly_fp64_matches_oracle(col, orc, specimen_consts):
    st = make_state(col, specimen_consts, seed=1)
    p = Pipeline(col, specimen_consts, st)
    got = p.assemble()
    ref = oracle_assemble(orc, p)
    _compare(got, ref)
    _cz_fields(p.gf, got, ref, len(col.czs))
    C = len(col.czs)
    cs = p.gf.cz_stride
    assert got["counts"] == (int(ref["icz"][2 * cs:2 * cs + C].sum()), int(ref["icz"][3 * cs:3 * cs + C].sum()))
    # the state exercises the failure branch and keeps most CZs alive
    assert 0 < got["counts"][1] < C
    p.gf.close()


def test_separate_kernels_accumulate_like_reference(col, orc, specimen_consts):
    """icf_elem_force then icf_cz_force on zeroed buffers = the reference's two launches."""
    st = make_state(col, specimen_consts, seed=2)
    p = Pipeline(col, specimen_consts, st)
    gf, lib = p.gf, p.gf.lib
    nnz, N = gf.csr.nnz, gf.csr.N
    vals = gf.buf("vals", 2 * nnz * 9, np.float64)
    rhs = gf.buf("rhs", N * 3, np.float64)
    vals.zero()
    rhs.zero()
    _lib.check(lib.icf_elem_force(gf.ctx, gf.g_ie_pcsr.p, gf.g_dn.p, H, vals.p, rhs.p, len(col.elastic),
                                  gf.el_elastic_stride, gf.nd_stride), gf.ctx, "icf_elem_force")
    _lib.check(lib.icf_cz_force(gf.ctx, gf.g_dcz.p, gf.g_icz.p, gf.g_dn.p, vals.p, rhs.p, H, len(col.czs),
                                gf.cz_stride, gf.nd_stride), gf.ctx, "icf_cz_force")
    got = dict(vals=vals.download(nnz * 9), rhs=rhs.download(N * 3),
               F=gf.g_dn.download(3 * gf.nd_stride, offset_elems=12 * gf.nd_stride))
    ref = oracle_assemble(orc, p)
    _compare(got, ref)
    gf.close()


def test_assembly_is_bitwise_deterministic(col, specimen_consts):
    st = make_state(col, specimen_consts, seed=3)
    p = Pipeline(col, specimen_consts, st)
    a = p.assemble()
    p.gf.TransferNodesToGPU(st)          # F back to zero, same inputs
    p.gf.TransferPCSR(st)                # re-upload: rebuilds the slot maps
    b = p.assemble()
    for k in ("vals", "rhs", "F"):
        assert np.array_equal(a[k], b[k]), k
    # one stream instead of two (and profiling, which also runs on one stream): same bits
    gf = p.gf
    for setter in (gf.lib.icf_set_overlap, gf.lib.icf_set_profiling):
        setter(gf.ctx, 1 if setter is gf.lib.icf_set_profiling else 0)
        gf.TransferNodesToGPU(st)
        c = p.assemble()
        for k in ("vals", "rhs", "F"):
            assert np.array_equal(a[k], c[k]), k
    gf.lib.icf_set_profiling(gf.ctx, 0)
    gf.lib.icf_set_overlap(gf.ctx, 1)
    p.gf.close()


def test_contact_pipeline_matches_oracle(orc, specimen_consts):
    col = small_specimen(cells=(5, 4, 4), size=(0.05, 0.04, 0.04), grains=5, seed=4)
    st = make_state(col, specimen_consts, seed=5, platen_dz_frac=0.3)
    p = Pipeline(col, specimen_consts, st, contact=True)
    gf = p.gf
    # narrow phase: masks bit-exact against the oracle on the same candidate pairs
    Q = len(p.pairs)
    nl = p.pairs.astype(np.int32).ravel().copy()
    orc.narrow_phase(Q, gf.c_dn, gf.c_ie, nl, gf.el_all_stride, gf.nd_stride)
    assert np.array_equal(nl[0::2], gf.masks)
    assert np.count_nonzero(gf.masks) > 0
    # closest face: bit-exact
    ts = gf.tet_stride
    nT = gf.nImpacts
    # rebuild the (node, tet) input rows from NL2set
    masks2 = np.zeros(2 * Q, np.int32)
    masks2[0::2] = gf.masks
    out = np.zeros(4 * ts, np.int32)
    n = ctypes.c_int()
    _lib.check(gf.lib.icfh_nl2set(_lib.ptr(masks2), _lib.ptr(p.pairs.astype(np.int32).ravel().copy()), Q,
                                  _lib.ptr(gf.c_ie), gf.el_all_stride, _lib.ptr(out), ts, ts, ctypes.byref(n)))
    assert n.value == nT
    orc.find_closest_face(nT, ts, out, gf.c_dn, gf.c_ie, gf.c_ifc, gf.nd_stride, gf.el_all_stride, gf.fc_stride)
    for k in range(4):
        assert np.array_equal(out[k * ts:k * ts + nT], gf.c_itet[k * ts:k * ts + nT]), k
    # full iteration with contact response
    got = p.assemble()
    ref = oracle_assemble(orc, p)
    _compare(got, ref)
    gf.close()


def test_mixed_precision_tolerances(col, orc, specimen_consts):
    st = make_state(col, specimen_consts, seed=6)
    p = Pipeline(col, specimen_consts, st, precision=_lib.ICF_MIXED)
    got = p.assemble()
    ref = oracle_assemble(orc, p)
    errs = _compare(got, ref, tol_vals=1e-5, tol_rhs=TOL64)
    assert errs["vals"] > 0          # really ran in FP32
    _cz_fields(p.gf, got, ref, len(col.czs))
    p.gf.close()


def test_graph_replays_match_eager(orc, specimen_consts):
    """One captured iteration graph (fused assembly + contact response), replayed three times
    with eager calls in between; every replay must equal the eager result bit for bit."""
    col = small_specimen(cells=(5, 4, 4), size=(0.05, 0.04, 0.04), grains=5, seed=7)
    st = make_state(col, specimen_consts, seed=8, platen_dz_frac=0.3)
    p = Pipeline(col, specimen_consts, st, contact=True)
    gf, lib, ctx = p.gf, p.gf.lib, p.gf.ctx
    eager = p.assemble()
    ns = gf.nd_stride
    zeroF = np.zeros(3 * ns)
    from workloads import DIST_EPS, PENALTY_K

    def calls():
        _lib.check(lib.icf_assemble_elems_and_czs(
            ctx, gf.g_ie_pcsr.p, len(col.elastic), gf.el_elastic_stride, gf.g_dcz.p, gf.g_icz.p, len(col.czs),
            gf.cz_stride, gf.g_dn.p, ns, H, gf.g_dvals.p, gf.csr.nnz, gf.g_drhs.p, gf.csr.N, None, None), ctx, "asm")
        _lib.check(lib.icf_collision_response(ctx, gf.g_icr.p, gf.g_dn.p, H, gf.g_dvals.p, gf.g_drhs.p, gf.nImpacts,
                                              gf.cr_stride, ns, PENALTY_K, DIST_EPS), ctx, "contact")

    _lib.check(lib.icf_set_blocking(ctx, 0))
    _lib.check(lib.icf_graph_begin(ctx), ctx, "begin")
    calls()
    g = ctypes.c_void_p()
    _lib.check(lib.icf_graph_end(ctx, ctypes.byref(g)), ctx, "end")
    _lib.check(lib.icf_set_blocking(ctx, 1))
    for rep in range(3):
        gf.g_dn.upload(zeroF, offset_elems=12 * ns)                 # eager work between replays
        gf.g_dvals.zero()
        _lib.check(lib.icf_graph_launch(ctx, g), ctx, "launch")
        vals, rhs = gf.TransferLinearSystemToHost()
        F = gf.TransferUpdatedStateToHost()["F"]
        assert np.array_equal(vals, eager["vals"]), rep
        assert np.array_equal(rhs, eager["rhs"]), rep
        assert np.array_equal(F, eager["F"]), rep
        d, f = ctypes.c_int(), ctypes.c_int()
        _lib.check(lib.icf_cz_counts(ctx, ctypes.byref(d), ctypes.byref(f)), ctx, "counts")
        assert (d.value, f.value) == eager["counts"]
    # building a slot map is refused during capture
    _lib.check(lib.icf_set_blocking(ctx, 0))
    _lib.check(lib.icf_graph_begin(ctx), ctx, "begin")
    r = lib.icf_elem_force(ctx, gf.g_ie_pcsr.p, gf.g_dn.p, H, gf.g_dvals.p, gf.g_drhs.p, len(col.elastic),
                           gf.el_elastic_stride, ns)
    g2 = ctypes.c_void_p()
    lib.icf_graph_end(ctx, ctypes.byref(g2))
    assert r == -6
    if g2.value:
        lib.icf_graph_destroy(g2)
    lib.icf_graph_destroy(g)
    gf.close()


def test_slot_with_more_contributions_than_one_chunk(orc, specimen_consts):
    """1000 copies of one tet on the same pcsr slots: every slot then has more contributions
    than the staged reduction holds in LDS (768), which takes its unstaged path.  Checked in
    both the overwrite (fused) and the accumulate (single-kernel) modes."""
    from icflow_amd.gpu_functionality import stride_of
    lib = _lib.load()
    ctx = ctypes.c_void_p()
    _lib.check(lib.icf_create(0, ctypes.byref(ctx)))
    try:
        _lib.check(lib.icf_set_constants(ctx, ctypes.byref(specimen_consts)), ctx, "consts")
        rng = np.random.default_rng(11)
        nn, ns = 4, stride_of(4)
        dn = np.zeros(18 * ns)
        x0 = np.array([[0, 0, 0], [0.01, 0, 0], [0, 0.01, 0], [0, 0, 0.01]]) + rng.normal(0, 1e-4, (4, 3))
        un = rng.normal(0, 1e-5, (4, 3))
        for k in range(3):
            dn[k * ns:k * ns + nn] = x0[:, k]
            dn[(3 + k) * ns:(3 + k) * ns + nn] = un[:, k]
            dn[(6 + k) * ns:(6 + k) * ns + nn] = rng.normal(0, 1e-3, 4)
            dn[(9 + k) * ns:(9 + k) * ns + nn] = rng.normal(0, 1e-2, 4)
            dn[(15 + k) * ns:(15 + k) * ns + nn] = x0[:, k] + un[:, k]
        T = 1000
        es = stride_of(T)
        ie = np.zeros(28 * es, np.int32)
        up = {}
        for i in range(4):
            for j in range(i, 4):
                up[(i, j)] = len(up)
        for t in range(T):
            for i in range(4):
                ie[t + es * i] = i
                ie[t + es * (20 + i)] = i
                for j in range(4):
                    ie[t + es * (4 + 4 * i + j)] = up[(i, j)] if j >= i else -1
        nnz, N = len(up), 4
        ref_A, ref_b = np.zeros(nnz * 9), np.zeros(N * 3)
        orc.set_constants(specimen_consts)
        dref = dn.copy()
        orc.elem_force(ie, dref, H, ref_A, ref_b, T, es, ns)

        def dev(a):
            p = ctypes.c_void_p()
            _lib.check(lib.icf_alloc(ctx, a.nbytes, ctypes.byref(p)), ctx, "alloc")
            _lib.check(lib.icf_h2d(ctx, p, 0, _lib.ptr(a), a.nbytes), ctx, "h2d")
            return p

        d_dn, d_ie = dev(dn), dev(ie)
        d_A, d_b = dev(np.zeros(nnz * 9)), dev(np.zeros(N * 3))
        for fused in (True, False):
            _lib.check(lib.icf_h2d(ctx, d_dn, 0, _lib.ptr(dn), dn.nbytes), ctx, "h2d")
            _lib.check(lib.icf_memset(ctx, d_A, 0, nnz * 72), ctx, "memset")
            _lib.check(lib.icf_memset(ctx, d_b, 0, N * 24), ctx, "memset")
            if fused:
                _lib.check(lib.icf_assemble_elems_and_czs(ctx, d_ie, T, es, None, None, 0, 0, d_dn, ns, H, d_A,
                                                          nnz, d_b, N, None, None), ctx, "fused")
            else:
                _lib.check(lib.icf_elem_force(ctx, d_ie, d_dn, H, d_A, d_b, T, es, ns), ctx, "elem")
            A, b = np.zeros(nnz * 9), np.zeros(N * 3)
            _lib.check(lib.icf_d2h(ctx, _lib.ptr(A), d_A, 0, A.nbytes), ctx, "d2h")
            _lib.check(lib.icf_d2h(ctx, _lib.ptr(b), d_b, 0, b.nbytes), ctx, "d2h")
            print("fused" if fused else "single", rel_err(A, ref_A), rel_err(b, ref_b), flush=True)
            assert rel_err(A, ref_A) <= TOL64
            assert rel_err(b, ref_b) <= TOL64
    finally:
        lib.icf_destroy(ctx)


def test_mixed_precision_with_contact(orc, specimen_consts):
    """Mixed mode on the contact pipeline: tet, CZ and impact blocks are FP32 values added in
    FP64; residuals and node forces stay at the FP64 bar."""
    col = small_specimen(cells=(5, 4, 4), size=(0.05, 0.04, 0.04), grains=5, seed=16)
    st = make_state(col, specimen_consts, seed=17, platen_dz_frac=0.3)
    p = Pipeline(col, specimen_consts, st, precision=_lib.ICF_MIXED, contact=True)
    assert p.gf.nImpacts > 0
    got = p.assemble()
    ref = oracle_assemble(orc, p)
    errs = _compare(got, ref, tol_vals=1e-5, tol_rhs=TOL64)
    assert errs["vals"] > 0
    p.gf.close()


def test_closest_face_follows_adjacency_changes(orc, specimen_consts):
    """The closest-face search reads a compact copy of the surface tets' adjacent-face lists,
    built on first use.  Re-uploading different lists through icf_h2d must drop that copy:
    with every list reversed (the first strict minimum then picks the other face of a tie)
    and one face removed per list, the device still matches the oracle on the new lists."""
    col = small_specimen(cells=(5, 4, 4), size=(0.05, 0.04, 0.04), grains=5, seed=4)
    st = make_state(col, specimen_consts, seed=5, platen_dz_frac=0.3)
    p = Pipeline(col, specimen_consts, st, contact=True)
    gf, lib, ctx = p.gf, p.gf.lib, p.gf.ctx
    nT, ts, es = gf.nImpacts, gf.tet_stride, gf.el_all_stride
    assert nT > 0
    tuples = np.zeros(4 * ts, np.int32)
    tuples[:2 * ts] = gf.c_itet[:2 * ts]
    d_nl2 = gf.buf("cf_nl2", 4 * ts, np.int32)

    def device(ie_host):
        d_nl2.upload(tuples)
        _lib.check(lib.icf_find_closest_face(ctx, nT, ts, d_nl2.p, gf.g_dn.p, gf.g_ie.p, gf.g_ifc.p, gf.nd_stride, es,
                                             gf.fc_stride), ctx, "closest")
        return d_nl2.download(4 * ts)

    def oracle(ie_host):
        out = tuples.copy()
        orc.find_closest_face(nT, ts, out, gf.c_dn, ie_host, gf.c_ifc, gf.nd_stride, es, gf.fc_stride)
        return out

    first = device(gf.c_ie)
    assert np.array_equal(first[:4 * ts], oracle(gf.c_ie)[:4 * ts])
    ie2 = gf.c_ie.copy()
    nS = len(col.surf)
    for t in range(nS):
        n = ie2[t + es * 4]
        lst = [ie2[t + es * (5 + k)] for k in range(n)][::-1]
        if n > 1:
            lst = lst[1:]
        ie2[t + es * 4] = len(lst)
        for k, f in enumerate(lst):
            ie2[t + es * (5 + k)] = f
    gf.g_ie.upload(ie2)
    second = device(ie2)
    ref2 = oracle(ie2)
    for k in range(4):
        assert np.array_equal(second[k * ts:k * ts + nT], ref2[k * ts:k * ts + nT]), k
    # the changed lists change the answer, so a stale copy would have failed the check above
    assert not np.array_equal(ref2[ts:4 * ts], oracle(gf.c_ie)[ts:4 * ts])
    gf.close()

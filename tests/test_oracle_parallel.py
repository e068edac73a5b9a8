"""The parallel oracle pass timed as the CPU baseline (orc_assemble_par: OpenMP threads,
compare-and-swap scatter like the reference's atomicAdd2) assembles what the sequential
passes assemble: the same BCSR values, rhs rows and node forces up to the order of the
additions, and the same CZ flags and tentative history bit for bit."""
import numpy as np
import pytest

from conftest import rel_err


@pytest.mark.parametrize("threads", [1, 4])
def test_parallel_pass_equals_sequential(orc, lib, specimen_consts, threads):
    from icflow_amd import specimens
    from icflow_amd._lib import ptr
    from icflow_amd.gpu_functionality import CSR, stride_of
    col = specimens.make_specimen((6, 5, 4), (0.06, 0.05, 0.04), 6, 0).finalize()
    st = specimens.synthetic_state(col, specimen_consts, 1)
    rng = np.random.default_rng(5)
    Nn = len(col.x0)
    # impacts: a random node against a random face (any geometry exercises the pt branches)
    faces = col.faces[rng.choice(len(col.faces), 300)]
    pts = rng.integers(0, Nn, 300)
    keep = np.array([p not in f for p, f in zip(pts, faces)])
    imp = np.c_[pts[keep], faces[keep]].astype(np.int32)
    alt = np.ascontiguousarray(col.alt, np.int32)
    csr = CSR(lib).build(alt[col.elastic], alt[col.czs], alt[imp])
    T, C, P = len(col.elastic), len(col.czs), len(imp)
    es, cs, crs, ns = stride_of(T), stride_of(C), stride_of(P), stride_of(Nn)
    ie = np.zeros(28 * es, np.int32)
    icz = np.zeros(58 * cs, np.int32)
    icr = np.zeros(28 * crs, np.int32)
    assert lib.icfh_pack_elems(csr.h, ptr(np.ascontiguousarray(col.elastic, np.int32)), ptr(alt), T, es, ptr(ie)) == 0
    assert lib.icfh_pack_czs(csr.h, ptr(np.ascontiguousarray(col.czs, np.int32)), ptr(alt), C, cs, ptr(icz)) == 0
    nl2 = np.zeros(4 * crs, np.int32)
    for k in range(4):
        nl2[k * crs:k * crs + P] = imp[:, k]
    assert lib.icfh_pack_impacts(csr.h, ptr(nl2), crs, ptr(alt), P, crs, ptr(icr)) == 0
    dcz = np.zeros(16 * cs)
    for k in range(3):
        dcz[k * cs:k * cs + C] = st.pmax[:, k] if st.pmax.ndim == 2 else st.pmax
        dcz[(3 + k) * cs:(3 + k) * cs + C] = st.tmax[:, k] if st.tmax.ndim == 2 else st.tmax
    dn = np.zeros(18 * ns)
    for k in range(3):
        for f, arr in ((0, col.x0), (3, st.un), (6, st.vn), (9, st.an), (15, st.xc)):
            dn[(f + k) * ns:(f + k) * ns + Nn] = arr[:, k]
    h, kp, eps = 1e-4, 1e11 * 0.01 ** 2, 1e-15
    orc.set_constants(specimen_consts)

    seq = dict(dn=dn.copy(), dcz=dcz.copy(), icz=icz.copy(), A=np.zeros(9 * csr.nnz), b=np.zeros(3 * csr.N))
    orc.elem_force(ie, seq["dn"], h, seq["A"], seq["b"], T, es, ns)
    orc.cz_force(seq["dcz"], seq["icz"], seq["dn"], seq["A"], seq["b"], h, C, cs, ns)
    orc.collision_response(icr, seq["dn"], h, seq["A"], seq["b"], P, crs, ns, kp, eps)

    par = dict(dn=dn.copy(), dcz=dcz.copy(), icz=icz.copy(), A=np.full(9 * csr.nnz, 7.0), b=np.full(3 * csr.N, 7.0))
    orc.assemble_par(ie, T, es, par["dcz"], par["icz"], C, cs, icr, P, crs, par["dn"], ns, h, kp, eps, par["A"],
                     par["b"], threads)
    assert np.count_nonzero(seq["icz"][3 * cs:3 * cs + C]) + np.count_nonzero(seq["icz"][2 * cs:2 * cs + C]) > 0
    assert rel_err(par["A"], seq["A"]) <= 1e-13
    assert rel_err(par["b"], seq["b"]) <= 1e-13
    assert rel_err(par["dn"][12 * ns:15 * ns], seq["dn"][12 * ns:15 * ns]) <= 1e-13
    assert np.array_equal(par["icz"], seq["icz"])
    assert np.array_equal(par["dcz"], seq["dcz"])

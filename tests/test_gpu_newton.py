"""The device Newton driver (icflow_amd/newton.py) against the same loop run on the CPU
oracle with the same host solve: displacements, CZ history and failures after a few steps
of a small three-point bend, FP64 (1e-9 normwise on displacements)."""
import numpy as np
import pytest

from icflow_amd import _lib, newton
from icflow_amd.gpu_functionality import stride_of

pytestmark = pytest.mark.gpu


def _oracle_run(orc, m, steps):
    """ImplicitModel.step restated on the oracle, reading the model's static arrays."""
    col, p = m.col, m.p
    n, ns = m.n, m.ns
    st = np.zeros(12 * ns)
    dn = np.zeros(18 * ns)
    for k in range(3):
        dn[k * ns:k * ns + n] = col.x0[:, k]
    T, es = m.T, m.es
    ie = m.g_ie.download(28 * es)
    alive = np.arange(len(m.cz_nodes))
    pmax = np.zeros((len(m.cz_nodes), 3))
    tmax = np.zeros((len(m.cz_nodes), 3))
    alt = m.alt
    t = 0.0
    orc.set_constants(m.consts)
    for _ in range(steps):
        h = p.h
        u = np.stack([st[k * ns:k * ns + n] for k in range(3)], 1)
        v = np.stack([st[(3 + k) * ns:(3 + k) * ns + n] for k in range(3)], 1)
        du = np.zeros_like(u)
        du[alt >= 0] = v[alt >= 0] * h
        du[m.anchored_ids] = m.prescribed(t + h) - u[m.anchored_ids]
        for k in range(3):
            st[(9 + k) * ns:(9 + k) * ns + n] = du[:, k]
        # pattern and packing of the live CZs (host helpers)
        from icflow_amd.gpu_functionality import CSR
        csr = CSR(m.lib).build(alt[col.elastic], alt[m.cz_nodes[alive]])
        import ctypes
        iep = np.zeros(28 * es, np.int32)
        _lib.check(m.lib.icfh_pack_elems(csr.h, _lib.ptr(np.ascontiguousarray(col.elastic, np.int32)),
                                         _lib.ptr(np.ascontiguousarray(alt, np.int32)), T, es, _lib.ptr(iep)))
        C = len(alive)
        cs = stride_of(C)
        icz = np.zeros(58 * cs, np.int32)
        _lib.check(m.lib.icfh_pack_czs(csr.h, _lib.ptr(np.ascontiguousarray(m.cz_nodes[alive], np.int32)),
                                       _lib.ptr(np.ascontiguousarray(alt, np.int32)), C, cs, _lib.ptr(icz)))
        it, err0, conv, div = 0, None, False, False
        while True:
            orc.nodes_infer(st, ns, dn, ns, n, h, p.beta, p.gamma)
            dcz = np.zeros(16 * cs)
            for j in range(3):
                dcz[j * cs:j * cs + C] = pmax[alive, j]
                dcz[(3 + j) * cs:(3 + j) * cs + C] = tmax[alive, j]
            iczi = icz.copy()
            A = np.zeros(9 * csr.nnz)
            b = np.zeros(3 * csr.N)
            orc.elem_force(iep, dn, h, A, b, T, es, ns)
            orc.cz_force(dcz, iczi, dn, A, b, h, C, cs, ns)
            dx = newton.UpperBcsrSolver(csr.rows, csr.cols).solve(A, b)
            norm = float(dx @ dx)
            if it == 0:
                err0 = norm if norm != 0 else p.conv_eps
            elif norm < p.conv_cutoff:
                conv = True
            else:
                rel = np.sqrt(norm / err0)
                conv, div = rel <= p.conv_eps, rel > 1.01
            orc.nodes_add_dx(st, ns, np.ascontiguousarray(alt, np.int32), n, dx)
            it += 1
            if not (it < p.min_iter or (not div and not conv and it < p.max_iter)):
                break
        orc.nodes_accept(st, ns, dn, ns, n, h)
        for j in range(3):
            pmax[alive, j] = dcz[(6 + j) * cs:(6 + j) * cs + C]
            tmax[alive, j] = dcz[(9 + j) * cs:(9 + j) * cs + C]
        failed = iczi[3 * cs:3 * cs + C] != 0
        alive = alive[~failed]
        t += h
    u = np.stack([st[k * ns:k * ns + n] for k in range(3)], 1)
    return u, pmax, tmax, alive


def test_newton_driver_matches_oracle_loop(orc, specimen_consts):
    steps = 4
    col, presc = newton.three_point_bend(specimen_consts, cells=(20, 4, 4), size=(0.2, 0.025, 0.025), grains=4,
                                         seed=0, deflection=2e-4, ramp_steps=4, h=1e-4)
    m = newton.ImplicitModel(col, specimen_consts, precision=_lib.ICF_FP64, step=newton.StepParams(h=1e-4),
                             prescribed=presc)
    for _ in range(steps):
        m.step()
    u_dev = m.displacements()
    u_ref, pmax, tmax, alive = _oracle_run(orc, m, steps)
    err = np.abs(u_dev - u_ref).max() / np.abs(u_ref).max()
    print("rel displacement difference", err, "iterations", [r.iterations for r in m.records], flush=True)
    assert np.abs(u_ref).max() > 0
    assert err <= 1e-9
    assert np.array_equal(np.sort(m.cz_alive), np.sort(alive))
    # history: normwise (openings are differences of nearly equal coordinates, so small ones
    # carry the coordinates' rounding relative to their own size)
    from conftest import rel_err
    print("history differences", rel_err(m.pmax, pmax), rel_err(m.tmax, tmax), flush=True)
    assert rel_err(m.pmax, pmax) <= 1e-9 and rel_err(m.tmax, tmax) <= 1e-9
    m.close()


def test_time_step_control_discards_and_refines(specimen_consts):
    """The reference's time-step control (StepParams.adapt): a deflection ramped fast enough to
    damage or fail many CZs at once makes attempts explode or diverge; a discarded attempt leaves the
    accepted state (displacements, failed set, time) as it was, the retry runs at half the
    step, and the simulated time only advances by accepted steps."""
    col, presc = newton.three_point_bend(specimen_consts, cells=(20, 4, 4), size=(0.2, 0.025, 0.025), grains=4,
                                         seed=0, deflection=2e-3, ramp_steps=4, h=1e-4)
    # thresholds of 5% (the default 1% is less than one CZ of this small beam)
    m = newton.ImplicitModel(col, specimen_consts, precision=_lib.ICF_FP64,
                             step=newton.StepParams(h=1e-4, adapt=True, max_damage=0.05, max_fail=0.05),
                             prescribed=presc)
    discarded = 0
    for _ in range(40):
        u0, f0, k0, tsf0 = m.displacements(), m.failed_set(), m.int_time, m.tsf
        rec = m.step()
        assert rec.h == 1e-4 * (newton.PARTS // tsf0) / newton.PARTS
        if rec.accepted:
            assert m.int_time == k0 + newton.PARTS // tsf0
        else:
            discarded += 1
            assert np.array_equal(m.displacements(), u0) and np.array_equal(m.failed_set(), f0)
            assert m.int_time == k0 and m.tsf == 2 * tsf0
    print("attempts 40, discarded", discarded, "failed", len(m.failed_set()), "t", m.t, "tsf", m.tsf, flush=True)
    assert 0 < discarded < 40 and m.t > 0
    m.close()


def test_newton_with_contact(specimen_consts):
    """The Newton driver with contact (platen_compression: a free cube between rigid platens,
    the top one moving): every iteration's candidate pairs, narrow phase, NL2set, closest face
    and impact records are device calls.  The impacts of the last iteration equal the host
    pipeline's (k-DOP grid, narrow phase, host NL2set, closest face) on the same positions, two
    FP64 runs give the same bits, and the mixed run stays within 1e-5 of FP64 with the same
    impact count."""
    from icflow_amd.gpu_functionality import broad_phase
    steps = 3

    def run(mode):
        col, presc, nodes, k = newton.platen_compression(specimen_consts, cells=(6, 6, 6), size=(0.015,) * 3,
                                                         grains=5, seed=2)
        m = newton.ImplicitModel(col, specimen_consts, precision=mode, step=newton.StepParams(h=1e-4),
                                 prescribed=presc, prescribed_nodes=nodes, contact=True, penalty=k)
        for _ in range(steps):
            m.step()
        return m

    m = run(_lib.ICF_FP64)
    gf = m.gf
    assert m.nImpacts > 0
    got = np.stack([m.g_icr.download(4 * m.cr_stride)[r * m.cr_stride:r * m.cr_stride + m.nImpacts]
                    for r in range(4)], 1)
    gf.c_dn = m.g_dn.download(18 * m.ns)
    pairs = broad_phase(gf, gf.c_dn, m.ns, gf.c_ie, gf.el_all_stride, len(m.col.surf))
    host = gf.NarrowPhaseCollisionDetection(pairs)
    print("impacts", m.nImpacts, "host", len(host), flush=True)
    assert np.array_equal(got, host)
    u64 = m.displacements()
    m.close()
    m2 = run(_lib.ICF_FP64)
    assert np.array_equal(m2.displacements(), u64)
    m2.close()
    mx = run(_lib.ICF_MIXED)
    spec = np.flatnonzero(mx.col.node_mesh == 0)
    err = np.linalg.norm(mx.displacements()[spec] - u64[spec]) / np.linalg.norm(u64[spec])
    print("mixed vs fp64", err, "impacts", mx.nImpacts, flush=True)
    assert err <= 1e-5 and mx.nImpacts == len(host)
    mx.close()

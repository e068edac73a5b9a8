"""The Newton loop of ImplicitModel3.Step (ImplicitModel3.cs:377-435, fixed step) run on the CPU
oracle, from the static arrays of a device model (icflow_amd.newton.ImplicitModel), with the
same host solve.  With contact, every iteration detects contacts like the reference: k-DOP
candidate pairs (host grid), the oracle's narrow phase, the host NL2set, the oracle's closest
face, then the pattern with the impacts' nodes and the oracle's collision response
(ImplicitModel3.cs:383-398, GPU_Functionality.cs:539-676).  Test infrastructure only."""
import ctypes

import numpy as np

from icflow_amd import _lib, newton
from icflow_amd.gpu_functionality import CSR, broad_phase, stride_of


def _pack(lib, csr, units, alt, nq, stride, fn):
    n = len(units)
    out = np.zeros(nq * max(stride, 1), np.int32)
    if n:
        _lib.check(fn(csr.h, _lib.ptr(np.ascontiguousarray(units, np.int32)), _lib.ptr(alt), n, stride,
                      _lib.ptr(out)))
    return out


def _contacts(orc, m, dn):
    """Impacts (node, face v0, v1, v2) of the current positions, as the reference finds them."""
    gf, lib = m.gf, m.lib
    ns, es = m.ns, gf.el_all_stride
    pairs = broad_phase(gf, dn, ns, gf.c_ie, es, len(m.col.surf))
    Q = len(pairs)
    if Q == 0:
        return np.zeros((0, 4), np.int32)
    nl = np.ascontiguousarray(pairs, np.int32).ravel().copy()
    orc.narrow_phase(Q, dn, gf.c_ie, nl, es, ns)
    masks = np.zeros(2 * Q, np.int32)
    masks[0::2] = nl[0::2]
    cap = int(np.count_nonzero(masks)) * 8 + 1
    ts = stride_of(cap)
    out = np.zeros(4 * ts, np.int32)
    n = ctypes.c_int()
    _lib.check(lib.icfh_nl2set(_lib.ptr(masks), _lib.ptr(np.ascontiguousarray(pairs, np.int32).ravel().copy()), Q,
                               _lib.ptr(gf.c_ie), es, _lib.ptr(out), ts, cap, ctypes.byref(n)), what="icfh_nl2set")
    nT = n.value
    if nT == 0:
        return np.zeros((0, 4), np.int32)
    orc.find_closest_face(nT, ts, out, dn, gf.c_ie, gf.c_ifc, ns, es, gf.fc_stride)
    return np.stack([out[k * ts:k * ts + nT] for k in range(4)], 1)


def oracle_newton(orc, m, steps, contact=False, log=None):
    """Returns displacements, pmax, tmax, the surviving CZ indices and per-step impact counts of
    `steps` fixed steps (the device model's StepParams, prescribed motion and constants)."""
    col, p, lib = m.col, m.p, m.lib
    n, ns = m.n, m.ns
    st = np.zeros(12 * ns)
    dn = np.zeros(18 * ns)
    for k in range(3):
        dn[k * ns:k * ns + n] = col.x0[:, k]
    T, es = m.T, m.es
    alive = np.arange(len(m.cz_nodes))
    pmax = np.zeros((len(m.cz_nodes), 3))
    tmax = np.zeros((len(m.cz_nodes), 3))
    alt = np.ascontiguousarray(m.alt, np.int32)
    elastic = np.ascontiguousarray(col.elastic, np.int32)
    t = 0.0
    impacts_per_step = []
    orc.set_constants(m.consts)
    for step in range(steps):
        h = p.h
        u = np.stack([st[k * ns:k * ns + n] for k in range(3)], 1)
        v = np.stack([st[(3 + k) * ns:(3 + k) * ns + n] for k in range(3)], 1)
        du = np.zeros_like(u)
        du[alt >= 0] = v[alt >= 0] * h
        if m.prescribed is not None and len(m.anchored_ids):
            du[m.anchored_ids] = m.prescribed(t + h) - u[m.anchored_ids]
        for k in range(3):
            st[(9 + k) * ns:(9 + k) * ns + n] = du[:, k]
        C = len(alive)
        cs = stride_of(C) if C else 0
        czn = m.cz_nodes[alive]
        it, err0, conv, div = 0, None, False, False
        nimp = 0
        while True:
            orc.nodes_infer(st, ns, dn, ns, n, h, p.beta, p.gamma)
            imp = _contacts(orc, m, dn) if contact else np.zeros((0, 4), np.int32)
            nimp = len(imp)
            csr = CSR(lib).build(alt[elastic], alt[czn] if C else np.zeros((0, 6)), alt[imp] if nimp else None)
            iep = _pack(lib, csr, elastic, alt, 28, es, lib.icfh_pack_elems)
            icz = _pack(lib, csr, czn, alt, 58, cs, lib.icfh_pack_czs) if C else None
            dcz = np.zeros(16 * max(cs, 1))
            for j in range(3):
                dcz[j * cs:j * cs + C] = pmax[alive, j]
                dcz[(3 + j) * cs:(3 + j) * cs + C] = tmax[alive, j]
            A = np.zeros(9 * csr.nnz)
            b = np.zeros(3 * csr.N)
            orc.elem_force(iep, dn, h, A, b, T, es, ns)
            if C:
                orc.cz_force(dcz, icz, dn, A, b, h, C, cs, ns)
            if nimp:
                cr = stride_of(nimp)
                nl2 = np.zeros(4 * cr, np.int32)
                for k in range(4):
                    nl2[k * cr:k * cr + nimp] = imp[:, k]
                icr = np.zeros(28 * cr, np.int32)
                _lib.check(lib.icfh_pack_impacts(csr.h, _lib.ptr(nl2), cr, _lib.ptr(alt), nimp, cr, _lib.ptr(icr)),
                           what="icfh_pack_impacts")
                orc.collision_response(icr, dn, h, A, b, nimp, cr, ns, m.penalty, m.dist_eps)
            dx = newton.UpperBcsrSolver(csr.rows, csr.cols).solve(A, b)
            norm = float(dx @ dx)
            if it == 0:
                err0 = norm if norm != 0 else p.conv_eps
            elif norm < p.conv_cutoff:
                conv = True
            else:
                rel = np.sqrt(norm / err0)
                conv, div = rel <= p.conv_eps, rel > 1.01
            orc.nodes_add_dx(st, ns, alt, n, dx)
            it += 1
            if not (it < p.min_iter or (not div and not conv and it < p.max_iter)):
                break
        orc.nodes_accept(st, ns, dn, ns, n, h)
        if C:
            for j in range(3):
                pmax[alive, j] = dcz[(6 + j) * cs:(6 + j) * cs + C]
                tmax[alive, j] = dcz[(9 + j) * cs:(9 + j) * cs + C]
            failed = icz[3 * cs:3 * cs + C] != 0
            alive = alive[~failed]
        impacts_per_step.append(nimp)
        if log:
            log(f"oracle step {step}: {it} iterations, {nimp} impacts, {len(m.cz_nodes) - len(alive)} failed")
        t += h
    u = np.stack([st[k * ns:k * ns + n] for k in range(3)], 1)
    return u, pmax, tmax, alive, impacts_per_step

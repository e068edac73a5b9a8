"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker.  The product library never loads it.

All kernel entry points take the reference SoA buffers as numpy arrays and modify them in
place, exactly like the device library's entry points do on device memory.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

c_int, c_double, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p


class OrcConsts(ctypes.Structure):
    """Same field layout as icf_consts / orc_consts."""
    _fields_ = [("E", c_double * 36), ("M", c_double * 144),
                ("NewmarkBeta", c_double), ("NewmarkGamma", c_double), ("dampingMass", c_double),
                ("dampingStiffness", c_double), ("rho", c_double), ("gravity", c_double),
                ("YoungsModulus", c_double),
                ("G_fn", c_double), ("G_ft", c_double), ("f_tn", c_double), ("f_tt", c_double),
                ("alpha", c_double), ("beta", c_double), ("rn", c_double), ("rt", c_double),
                ("p_m", c_double), ("p_n", c_double), ("deln", c_double), ("delt", c_double),
                ("pMtn", c_double), ("pMnt", c_double), ("gam_n", c_double), ("gam_t", c_double),
                ("B", c_double * 162), ("sf", c_double * 9)]


_SIGS = {
    "orc_set_constants": (None, [ctypes.POINTER(OrcConsts)]),
    "orc_F_and_Df": (None, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "orc_elem_force": (None, [c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_int, c_int, c_int]),
    "orc_Tn": (c_double, [c_double, c_double]),
    "orc_Tt": (c_double, [c_double, c_double]),
    "orc_Dnn": (c_double, [c_double, c_double]),
    "orc_Dtt": (c_double, [c_double, c_double]),
    "orc_Dnt": (c_double, [c_double, c_double]),
    "orc_cohesive_law": (None, [c_void_p, c_void_p, c_void_p, c_double, c_double, c_void_p]),
    "orc_set_branch_log": (None, [c_void_p]),
    "orc_cz_rotation": (None, [c_void_p, c_void_p, c_void_p]),
    "orc_cz_force": (None, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_int, c_int, c_int]),
    "orc_ctest": (c_int, [c_void_p, c_double, c_double, c_double]),
    "orc_bvalues": (None, [c_void_p, c_void_p]),
    "orc_narrow_phase": (None, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_int]),
    "orc_dtn": (c_double, [c_void_p, c_void_p]),
    "orc_find_closest_face": (None, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int]),
    "orc_pt": (c_double, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "orc_collision_response": (None, [c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_int, c_int, c_int,
                                      c_double, c_double]),
    "orc_assemble_par": (None, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int, c_int, c_void_p, c_int, c_int,
                                c_void_p, c_int, c_double, c_double, c_double, c_void_p, ctypes.c_longlong, c_void_p,
                                ctypes.c_longlong, c_int]),
    "orc_nodes_infer": (None, [c_void_p, c_int, c_void_p, c_int, c_int, c_double, c_double, c_double]),
    "orc_nodes_add_dx": (None, [c_void_p, c_int, c_void_p, c_int, c_void_p]),
    "orc_nodes_accept": (None, [c_void_p, c_int, c_void_p, c_int, c_int, c_double]),
}

_lib = None


def _open(path):
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} missing: run `make -C oracle`")
        _lib = _open(LIB_PATH)
    return _lib


class swapped:
    """Context manager: every function of this module calls the oracle built at `path`
    instead (the mutation tests build deliberately wrong copies of the oracle this way)."""

    def __init__(self, path):
        self.path = path

    def __enter__(self):
        global _lib
        self.saved = load()
        _lib = _open(self.path)
        return self

    def __exit__(self, *exc):
        global _lib
        _lib = self.saved
        return False


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(c_void_p)


def set_constants(consts):
    """consts: any ctypes structure with the icf_consts layout."""
    oc = OrcConsts.from_buffer_copy(bytes(consts))
    load().orc_set_constants(ctypes.byref(oc))


def F_and_Df(x0, un):
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    un = np.ascontiguousarray(un, dtype=np.float64)
    f = np.zeros(12)
    Df = np.zeros((12, 12))
    V = np.zeros(1)
    load().orc_F_and_Df(_p(x0), _p(un), _p(f), _p(Df), _p(V))
    return f, Df, V[0]


def elem_force(ie, dn, h, A, b, n, es, ns):
    load().orc_elem_force(_p(ie), _p(dn), h, _p(A), _p(b), n, es, ns)


def cz_force(dcz, icz, dn, A, b, h, n, cs, ns):
    load().orc_cz_force(_p(dcz), _p(icz), _p(dn), _p(A), _p(b), h, n, cs, ns)


def cz_branches(dcz, icz, dn, A, b, h, n, cs, ns):
    """The PPR-law leaf each Gauss point of the first n CZs takes (codes in
    orc_set_branch_log's comment), from cz_force on copies of the arrays; -1 where a CZ is
    skipped (already failed)."""
    log = np.full(3 * n, -1, np.int32)
    lib = load()
    lib.orc_set_branch_log(_p(log))
    try:
        cz_force(dcz.copy(), icz.copy(), dn.copy(), A.copy(), b.copy(), h, n, cs, ns)
    finally:
        lib.orc_set_branch_log(None)
    return log


def narrow_phase(nPairs, dn, ie, nl, es, ns):
    load().orc_narrow_phase(nPairs, _p(dn), _p(ie), _p(nl), es, ns)


def find_closest_face(nPairs, ts, nl2, dn, ie, faces, ns, es, fs):
    load().orc_find_closest_face(nPairs, ts, _p(nl2), _p(dn), _p(ie), _p(faces), ns, es, fs)


def collision_response(icr, dn, h, A, b, n, cs, ns, k, eps):
    load().orc_collision_response(_p(icr), _p(dn), h, _p(A), _p(b), n, cs, ns, k, eps)


def ppr(name, dn_, dt_):
    return getattr(load(), "orc_" + name)(float(dn_), float(dt_))


def cohesive_law(pmax, tmax, opn, opt, contact=0, failed=0):
    flags = np.array([contact, failed], dtype=np.int32)
    pm = np.array([pmax], dtype=np.float64)
    tm = np.array([tmax], dtype=np.float64)
    out = np.zeros(6)
    load().orc_cohesive_law(_p(flags), _p(pm), _p(tm), float(opn), float(opt), _p(out))
    return dict(contact=int(flags[0]), failed=int(flags[1]), pmax=pm[0], tmax=tm[0], Tn=out[0], Tt=out[1],
                Dnn=out[2], Dtt=out[3], Dnt=out[4], Dtn=out[5])


def cz_rotation(p9):
    p9 = np.ascontiguousarray(p9, dtype=np.float64)
    R = np.zeros(9)
    a = np.zeros(1)
    load().orc_cz_rotation(_p(p9), _p(R), _p(a))
    return R.reshape(3, 3), a[0]


def bvalues(p12):
    p12 = np.ascontiguousarray(p12, dtype=np.float64)
    b = np.zeros(9)
    load().orc_bvalues(_p(p12), _p(b))
    return b


def ctest(bm, x, y, z):
    bm = np.ascontiguousarray(bm, dtype=np.float64)
    return bool(load().orc_ctest(_p(bm), float(x), float(y), float(z)))


def dtn(f9, p3):
    f9 = np.ascontiguousarray(f9, dtype=np.float64)
    p3 = np.ascontiguousarray(p3, dtype=np.float64)
    return load().orc_dtn(_p(f9), _p(p3))


def pt(x12):
    x12 = np.ascontiguousarray(x12, dtype=np.float64)
    fd = np.zeros(12)
    sd = np.zeros((12, 12))
    s = np.zeros(1)
    t = np.zeros(1)
    br = np.zeros(1, dtype=np.int32)
    d = load().orc_pt(_p(x12), _p(fd), _p(sd), _p(s), _p(t), _p(br))
    return d, fd, sd, s[0], t[0], int(br[0])


def assemble_par(ie, nE, es, dcz, icz, nC, cs, icr, nP, crs, dn, ns, h, k, eps, A, b, threads):
    """The iteration's element, CZ and contact passes on `threads` OpenMP threads with
    compare-and-swap scatter (A and b are zeroed first); dcz/icz/icr may be None when their
    counts are 0."""
    load().orc_assemble_par(_p(ie), nE, es, _p(dcz), _p(icz), nC, cs, _p(icr), nP, crs, _p(dn), ns, h, k, eps,
                            _p(A), A.size, _p(b), b.size, threads)


def nodes_infer(st, ss, dn, ns, n, h, beta, gamma):
    load().orc_nodes_infer(_p(st), ss, _p(dn), ns, n, h, beta, gamma)


def nodes_add_dx(st, ss, alt, n, dx):
    load().orc_nodes_add_dx(_p(st), ss, _p(alt), n, _p(dx))


def nodes_accept(st, ss, dn, ns, n, h):
    load().orc_nodes_accept(_p(st), ss, _p(dn), ns, n, h)

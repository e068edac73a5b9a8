"""ctypes binding of libicflow_hip.so (include/icflow_hip.h and include/icflow_host.h).

The shared library is built in-tree by ``icflow_amd.build.build()``.  There is no fallback:
if the library is missing, loading raises, so nothing can silently run on the CPU instead.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libicflow_hip.so")

c_int, c_double, c_void_p, c_size_t = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t
P_int = ctypes.POINTER(ctypes.c_int)

ICF_FP64, ICF_MIXED = 0, 1
ICF_E_INVALID, ICF_E_NOCONST, ICF_E_CAPACITY, ICF_E_CONSTANTS, ICF_E_INDEX, ICF_E_CAPTURE = -1, -2, -3, -4, -5, -6
ERRORS = {-1: "ICF_E_INVALID", -2: "ICF_E_NOCONST", -3: "ICF_E_CAPACITY", -4: "ICF_E_CONSTANTS",
          -5: "ICF_E_INDEX", -6: "ICF_E_CAPTURE"}


class IcfConsts(ctypes.Structure):
    """Mirror of icf_consts (include/icflow_hip.h)."""
    _fields_ = [("E", c_double * 36), ("M", c_double * 144),
                ("NewmarkBeta", c_double), ("NewmarkGamma", c_double), ("dampingMass", c_double),
                ("dampingStiffness", c_double), ("rho", c_double), ("gravity", c_double),
                ("YoungsModulus", c_double),
                ("G_fn", c_double), ("G_ft", c_double), ("f_tn", c_double), ("f_tt", c_double),
                ("alpha", c_double), ("beta", c_double), ("rn", c_double), ("rt", c_double),
                ("p_m", c_double), ("p_n", c_double), ("deln", c_double), ("delt", c_double),
                ("pMtn", c_double), ("pMnt", c_double), ("gam_n", c_double), ("gam_t", c_double),
                ("B", c_double * 162), ("sf", c_double * 9)]


class IcfModelPrms(ctypes.Structure):
    """Mirror of icf_model_prms (include/icflow_host.h); fields of ModelPrms.cs."""
    _fields_ = [(n, c_double) for n in (
        "Y", "nu", "rho", "dampingMass", "dampingStiffness", "NewmarkBeta", "NewmarkGamma", "gravity",
        "alpha", "beta", "lambda_n", "lambda_t", "phi_n", "phi_t", "sigma_max", "tau_max")]


class IcfWorklist(ctypes.Structure):
    """Mirror of icf_worklist (include/icflow_hip.h): per-specimen offsets, host arrays."""
    _fields_ = [("nSpecimens", c_int), ("tet_off", c_void_p), ("cz_off", c_void_p), ("row_off", c_void_p),
                ("nnzb_off", c_void_p)]


_SIGS = {
    "icf_create": (c_int, [c_int, ctypes.POINTER(c_void_p)]),
    "icf_destroy": (c_int, [c_void_p]),
    "icf_last_error_string": (ctypes.c_char_p, [c_void_p]),
    "icf_set_constants": (c_int, [c_void_p, ctypes.POINTER(IcfConsts)]),
    "icf_set_precision": (c_int, [c_void_p, c_int]),
    "icf_get_precision": (c_int, [c_void_p]),
    "icf_alloc": (c_int, [c_void_p, c_size_t, ctypes.POINTER(c_void_p)]),
    "icf_free": (c_int, [c_void_p, c_void_p]),
    "icf_h2d": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p, c_size_t]),
    "icf_d2h": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t, c_size_t]),
    "icf_memset": (c_int, [c_void_p, c_void_p, c_int, c_size_t]),
    "icf_synchronize": (c_int, [c_void_p]),
    "icf_elem_force": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_int, c_int, c_int]),
    "icf_cz_force": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double,
                             c_int, c_int, c_int]),
    "icf_narrow_phase": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int, c_int]),
    "icf_find_closest_face": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                      c_int, c_int, c_int]),
    "icf_find_closest_face_noalloc": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                              c_int, c_int, c_int]),
    "icf_collision_response": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_int,
                                       c_int, c_int, c_double, c_double]),
    "icf_assemble_elems_and_czs": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_int, c_int,
                                           c_void_p, c_int, c_double, c_void_p, c_int, c_void_p, c_int,
                                           P_int, P_int]),
    "icf_assemble_batched": (c_int, [c_void_p, ctypes.POINTER(IcfWorklist), c_void_p, c_int, c_void_p, c_void_p,
                                     c_int, c_void_p, c_int, c_double, c_void_p, c_void_p, c_void_p, c_void_p]),
    "icf_batch_cz_counts": (c_int, [c_void_p, c_void_p, c_void_p]),
    "icf_sum_int": (c_int, [c_void_p, c_void_p, c_int, P_int]),
    "icf_pattern_changed": (c_int, [c_void_p]),
    "icf_last_device_ms": (c_double, [c_void_p]),
    "icf_d2d": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p, c_size_t, c_size_t]),
    "icf_probe_valu": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p]),
    "icf_set_shape": (c_int, [c_void_p, c_int, c_int]),
    "icf_get_shape": (c_int, [c_void_p, c_int, P_int, P_int, c_void_p]),
    "icf_set_blocking": (c_int, [c_void_p, c_int]),
    "icf_cz_counts": (c_int, [c_void_p, P_int, P_int]),
    "icf_graph_begin": (c_int, [c_void_p]),
    "icf_graph_end": (c_int, [c_void_p, ctypes.POINTER(c_void_p)]),
    "icf_graph_launch": (c_int, [c_void_p, c_void_p]),
    "icf_graph_destroy": (c_int, [c_void_p]),
    "icf_set_profiling": (c_int, [c_void_p, c_int]),
    "icf_set_overlap": (c_int, [c_void_p, c_int]),
    "icf_set_autotune": (c_int, [c_void_p, c_int]),
    "icf_get_autotune": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "icf_set_chunk_keys": (c_int, [c_void_p, c_int, c_int, c_int, c_int]),
    "icf_profile_reset": (c_int, [c_void_p]),
    "icf_profile_count": (c_int, [c_void_p]),
    "icf_profile_entry": (c_int, [c_void_p, c_int, ctypes.c_char_p, c_int, ctypes.POINTER(c_double),
                                  ctypes.POINTER(ctypes.c_longlong)]),
    # host helpers
    "icfh_compute_constants": (c_int, [ctypes.POINTER(IcfModelPrms), ctypes.POINTER(IcfConsts)]),
    "icfh_csr_create": (c_int, [ctypes.POINTER(c_void_p)]),
    "icfh_csr_destroy": (c_int, [c_void_p]),
    "icfh_csr_build": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p, c_int]),
    "icfh_csr_dims": (c_int, [c_void_p, P_int, P_int]),
    "icfh_csr_copy": (c_int, [c_void_p, c_void_p, c_void_p]),
    "icfh_csr_lookup": (c_int, [c_void_p, c_int, c_int]),
    "icfh_pack_elems": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "icfh_pack_czs": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "icfh_pack_impacts": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_int, c_void_p]),
    "icfh_nl2set": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, c_int, P_int]),
    "icfh_kdop_pairs": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_int, c_void_p, ctypes.c_longlong,
                                ctypes.POINTER(ctypes.c_longlong)]),
    "SolveDouble3": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_int, c_int, c_int,
                             c_int]),
    "icfh_solve_threads": (c_int, [c_int]),
    "icfh_solve_analyse": (c_int, [c_void_p, c_void_p, c_int, ctypes.POINTER(ctypes.c_longlong),
                                   ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(c_double)]),
    "icf_nodes_infer": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_int, c_double, c_double, c_double]),
    "icf_nodes_add_dx": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p]),
    "icf_nodes_accept": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_int, c_double]),
    "icf_build_pattern": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_int, c_void_p, c_int, c_int,
                                  c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_int, ctypes.POINTER(c_int)]),
    "icf_pack_pcsr": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_int, c_int,
                              c_void_p, c_int, c_int, c_void_p, c_int, c_int]),
    "icf_broad_phase": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_int, c_int, c_void_p, ctypes.c_longlong,
                                ctypes.POINTER(ctypes.c_longlong)]),
    "icf_nl2set": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p, c_int,
                           ctypes.POINTER(c_int)]),
}

EXPORTED = tuple(_SIGS)
_lib = None


def load():
    """Load libicflow_hip.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with icflow_amd.build.build()")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def ptr(a):
    """Raw data pointer of a contiguous numpy array (None for None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(c_void_p)


class IcfError(RuntimeError):
    pass


def check(code, ctx=None, what=""):
    if code != 0:
        msg = ERRORS.get(code, f"hipError {code}")
        detail = ""
        if ctx is not None:
            s = load().icf_last_error_string(ctx)
            detail = s.decode() if s else ""
        raise IcfError(f"{what}: {msg} {detail}".strip())


def consts_from_prms(**kw):
    """icf_consts from ModelPrms-style keyword parameters (defaults: the reference's)."""
    lib = load()
    p = IcfModelPrms(**{**MODEL_DEFAULTS, **kw})
    c = IcfConsts()
    check(lib.icfh_compute_constants(ctypes.byref(p), ctypes.byref(c)), what="icfh_compute_constants")
    return c


# ModelPrms defaults (ModelPrms.cs:38-40, 137-151, 180-209)
MODEL_DEFAULTS = dict(Y=1e10, nu=0.3, rho=916.2, dampingMass=0.0005, dampingStiffness=0.0005,
                      NewmarkBeta=0.25, NewmarkGamma=0.5, gravity=-9.8,
                      alpha=3.0, beta=3.0, lambda_n=0.02, lambda_t=0.02, phi_n=30.0, phi_t=30.0,
                      sigma_max=4e5, tau_max=15e5)

# The material and PPR parameters of the benchmark specimens: E=9 GPa, nu=0.3, rho=916,
# sigma_max=1 MPa, tau_max=2 MPa, phi=10 J/m^2, alpha=beta=4, lambda=0.02.
SPECIMEN_PRMS = dict(Y=9e9, nu=0.3, rho=916.0, alpha=4.0, beta=4.0, lambda_n=0.02, lambda_t=0.02,
                     phi_n=10.0, phi_t=10.0, sigma_max=1e6, tau_max=2e6)


def consts_to_numpy(c):
    """Flatten an IcfConsts into a dict of numpy arrays / floats."""
    out = {}
    for name, typ in IcfConsts._fields_:
        v = getattr(c, name)
        out[name] = np.ctypeslib.as_array(v).copy() if hasattr(typ, "_length_") else float(v)
    return out

"""Device-resident node state (icf_nodes_infer / icf_nodes_add_dx / icf_nodes_accept)
against the oracle's restatement of Node.cs:62-97 and ImplicitModel3._XtoDU: bit-exact
(same expressions, no contraction), over two Newton iterations and an accepted step."""
import ctypes

import numpy as np
import pytest

from icflow_amd import _lib
from icflow_amd.gpu_functionality import stride_of

pytestmark = pytest.mark.gpu


def test_node_state_iteration_matches_oracle_bitwise(orc):
    lib = _lib.load()
    ctx = ctypes.c_void_p()
    _lib.check(lib.icf_create(0, ctypes.byref(ctx)))
    try:
        rng = np.random.default_rng(5)
        n = 5000
        ss = ns = stride_of(n)
        st = np.zeros(12 * ss)
        for f in range(12):
            st[f * ss:f * ss + n] = rng.normal(0, 1e-3 * 10 ** (f // 3), n)
        dn = np.zeros(18 * ns)
        dn[:3 * ns] = rng.normal(0, 0.05, 3 * ns)
        alt = np.full(n, -1, np.int32)
        act = rng.random(n) < 0.8
        alt[act] = rng.permutation(int(act.sum()))
        N = int(act.sum())
        h, beta, gamma = 5e-5, 0.25, 0.5

        def dev(a):
            p = ctypes.c_void_p()
            _lib.check(lib.icf_alloc(ctx, a.nbytes, ctypes.byref(p)), ctx, "alloc")
            _lib.check(lib.icf_h2d(ctx, p, 0, _lib.ptr(a), a.nbytes), ctx, "h2d")
            return p

        def host(p, like):
            out = np.empty_like(like)
            _lib.check(lib.icf_d2h(ctx, _lib.ptr(out), p, 0, out.nbytes), ctx, "d2h")
            return out

        d_st, d_dn, d_alt = dev(st), dev(dn), dev(alt)
        d_dx = dev(np.zeros(3 * N))
        for it in range(2):
            _lib.check(lib.icf_nodes_infer(ctx, d_st, ss, d_dn, ns, n, h, beta, gamma), ctx, "infer")
            orc.nodes_infer(st, ss, dn, ns, n, h, beta, gamma)
            assert np.array_equal(host(d_dn, dn), dn), it
            dx = rng.normal(0, 1e-6, 3 * N)
            _lib.check(lib.icf_h2d(ctx, d_dx, 0, _lib.ptr(dx), dx.nbytes), ctx, "h2d")
            _lib.check(lib.icf_nodes_add_dx(ctx, d_st, ss, d_alt, n, d_dx), ctx, "add_dx")
            orc.nodes_add_dx(st, ss, alt, n, dx)
            assert np.array_equal(host(d_st, st), st), it
        _lib.check(lib.icf_nodes_accept(ctx, d_st, ss, d_dn, ns, n, h), ctx, "accept")
        orc.nodes_accept(st, ss, dn, ns, n, h)
        assert np.array_equal(host(d_st, st), st)
    finally:
        lib.icf_destroy(ctx)

"""icf_find_closest_face_noalloc: the allocation-free variant of icf_find_closest_face.

Through the C ABI, on the device: two calls in a row (the second with fewer tuples and a
smaller surface) give exactly the existing entry point's closest faces (integer outputs:
bit-exact), which match the oracle; a call whose lists would not fit the fixed scratch array
is refused with ICF_E_CAPACITY before any work; a steady-state call leaves the device's free
memory unchanged.
"""
import ctypes

import numpy as np
import pytest

from icflow_amd import _lib
from workloads import Pipeline, make_state, small_specimen

pytestmark = pytest.mark.gpu


def _closest(lib, ctx, fn, gf, d_nl2, tuples, n, ts):
    d_nl2.upload(tuples)
    _lib.check(fn(ctx, n, ts, d_nl2.p, gf.g_dn.p, gf.g_ie.p, gf.g_ifc.p, gf.nd_stride, gf.el_all_stride,
                  gf.fc_stride), ctx, fn.__name__)
    return d_nl2.download(4 * ts)


def _case(orc, consts, cells, size, grains, seed):
    col = small_specimen(cells=cells, size=size, grains=grains, seed=seed)
    st = make_state(col, consts, seed=seed + 1, platen_dz_frac=0.3, cells_z=cells[2], size_z=size[2])
    p = Pipeline(col, consts, st, contact=True)
    gf = p.gf
    nT, ts = gf.nImpacts, gf.tet_stride
    assert nT > 0
    tuples = np.zeros(4 * ts, np.int32)
    tuples[:2 * ts] = gf.c_itet[:2 * ts]
    return p, gf, nT, ts, tuples


def test_noalloc_closest_face_matches_allocating_entry_point(orc, specimen_consts):
    lib = _lib.load()
    cases = [((6, 5, 4), (0.06, 0.05, 0.04), 6, 4), ((4, 4, 3), (0.04, 0.04, 0.03), 3, 7)]
    sizes = []
    for call, (cells, size, grains, seed) in enumerate(cases):
        p, gf, nT, ts, tuples = _case(orc, specimen_consts, cells, size, grains, seed)
        if call == 1:
            nT = max(1, nT // 2)          # the second call: fewer tuples on a smaller surface
        sizes.append((nT, gf.el_all_stride))
        print(f"noalloc call {call}: tuples {nT}, surface stride {gf.el_all_stride}", flush=True)
        d_a = gf.buf("cf_a", 4 * ts, np.int32)
        d_b = gf.buf("cf_b", 4 * ts, np.int32)
        got = _closest(lib, gf.ctx, lib.icf_find_closest_face_noalloc, gf, d_a, tuples, nT, ts)
        ref = _closest(lib, gf.ctx, lib.icf_find_closest_face, gf, d_b, tuples, nT, ts)
        assert np.array_equal(got, ref)
        # untouched beyond the tuples, and equal to the oracle on them
        assert np.array_equal(got[nT:ts], tuples[nT:ts])
        want = tuples.copy()
        orc.find_closest_face(nT, ts, want, gf.c_dn, gf.c_ie, gf.c_ifc, gf.nd_stride, gf.el_all_stride, gf.fc_stride)
        for k in range(4):
            assert np.array_equal(got[k * ts:k * ts + nT], want[k * ts:k * ts + nT]), (call, k)
        gf.close()
    assert sizes[1][0] < sizes[0][0] and sizes[1][1] < sizes[0][1]


def _hip_free_bytes():
    # the HIP runtime the library already loaded (same file, so the same runtime instance)
    path = next(line.split()[-1] for line in open("/proc/self/maps") if "libamdhip64.so" in line)
    hip = ctypes.CDLL(path)
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_noalloc_refuses_oversized_lists_and_allocates_nothing(orc, specimen_consts):
    lib = _lib.load()
    p, gf, nT, ts, tuples = _case(orc, specimen_consts, (5, 4, 4), (0.05, 0.04, 0.04), 5, 4)
    ctx = gf.ctx
    d_a = gf.buf("cf_a", 4 * ts, np.int32)
    _closest(lib, ctx, lib.icf_find_closest_face_noalloc, gf, d_a, tuples, nT, ts)      # warm-up
    before = _hip_free_bytes()
    for _ in range(3):
        _closest(lib, ctx, lib.icf_find_closest_face_noalloc, gf, d_a, tuples, nT, ts)
    assert _hip_free_bytes() == before
    # a surface of 2^20 tets with 16 adjacent-face rows: 4 (2^20 + 1) + 4 * 2^20 * 16 bytes
    # of lists exceed the 64 MiB scratch; refused before any kernel is issued
    es = 1 << 20
    big_ie = gf.buf("big_ie", es * 21, np.int32)
    r = lib.icf_find_closest_face_noalloc(ctx, 1, ts, d_a.p, gf.g_dn.p, big_ie.p, gf.g_ifc.p, gf.nd_stride, es,
                                          gf.fc_stride)
    assert r == _lib.ICF_E_CAPACITY
    assert b"scratch" in lib.icf_last_error_string(ctx)
    # the context still works afterwards
    got = _closest(lib, ctx, lib.icf_find_closest_face_noalloc, gf, d_a, tuples, nT, ts)
    want = tuples.copy()
    orc.find_closest_face(nT, ts, want, gf.c_dn, gf.c_ie, gf.c_ifc, gf.nd_stride, gf.el_all_stride, gf.fc_stride)
    assert np.array_equal(got[:4 * ts], want[:4 * ts])
    gf.close()

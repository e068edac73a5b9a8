"""Python mirror of icFlow's GPU_Functionality (v3Library/Simulation/GPU/GPU_Functionality.cs)
driving libicflow_hip.so through its C ABI.

Method names, call order and buffer layouts follow the C# class so that a driver written
against the reference reads the same here.  The packing loops the C# class runs with
Parallel.For are the C++ host helpers of the library (include/icflow_host.h).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, ptr

BLOCK = 128                       # GPU_Functionality.cs:38
FP_DATA_SIZE_NODE, INT_DATA_SIZE_ELEM, FP_DATA_SIZE_CZ, INT_DATA_SIZE_CZ = 18, 28, 16, 58
X0, UN, VN, AN, FF, XC = 0, 3, 6, 9, 12, 15


def stride_of(n):
    """ceil(n / 128) * 128 (GPU_Functionality.cs:139-149)."""
    return ((n + BLOCK - 1) // BLOCK) * BLOCK


class DeviceArray:
    """A CudaDeviceVariable<T> look-alike: device allocation owned through the context."""

    def __init__(self, gf, n, dtype):
        self.gf, self.n, self.dtype = gf, int(n), np.dtype(dtype)
        p = ctypes.c_void_p()
        check(gf.lib.icf_alloc(gf.ctx, max(1, self.n) * self.dtype.itemsize, ctypes.byref(p)), gf.ctx, "icf_alloc")
        self.p = p

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    def upload(self, host, offset_elems=0):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        assert offset_elems + host.size <= self.n
        check(self.gf.lib.icf_h2d(self.gf.ctx, self.p, offset_elems * self.dtype.itemsize, ptr(host), host.nbytes),
              self.gf.ctx, "icf_h2d")

    def download(self, n=None, offset_elems=0):
        n = self.n - offset_elems if n is None else n
        out = np.empty(n, self.dtype)
        check(self.gf.lib.icf_d2h(self.gf.ctx, ptr(out), self.p, offset_elems * self.dtype.itemsize, out.nbytes),
              self.gf.ctx, "icf_d2h")
        return out

    def zero(self):
        check(self.gf.lib.icf_memset(self.gf.ctx, self.p, 0, self.nbytes), self.gf.ctx, "icf_memset")

    def at(self, offset_elems):
        return ctypes.c_void_p(self.p.value + offset_elems * self.dtype.itemsize)

    def free(self):
        if self.p is not None and self.p.value:
            self.gf.lib.icf_free(self.gf.ctx, self.p)
        self.p = None


class CSR:
    """CSRDictionary (CSRDictionary.cs) through the host helpers."""

    def __init__(self, lib):
        self.lib = lib
        h = ctypes.c_void_p()
        check(lib.icfh_csr_create(ctypes.byref(h)), what="icfh_csr_create")
        self.h = h
        self.N = self.nnz = 0

    def build(self, alt_elem, alt_cz, alt_imp=None):
        ae = np.ascontiguousarray(alt_elem, np.int32)
        ac = np.ascontiguousarray(alt_cz, np.int32)
        ai = np.ascontiguousarray(alt_imp if alt_imp is not None else np.zeros((0, 4)), np.int32)
        check(self.lib.icfh_csr_build(self.h, ptr(ae), len(ae), ptr(ac), len(ac), ptr(ai), len(ai)),
              what="icfh_csr_build")
        N, nnz = ctypes.c_int(), ctypes.c_int()
        self.lib.icfh_csr_dims(self.h, ctypes.byref(N), ctypes.byref(nnz))
        self.N, self.nnz = N.value, nnz.value
        self.rows = np.zeros(self.N + 1, np.int32)
        self.cols = np.zeros(max(1, self.nnz), np.int32)
        self.lib.icfh_csr_copy(self.h, ptr(self.rows), ptr(self.cols))
        self.cols = self.cols[:self.nnz]
        return self

    def __del__(self):
        try:
            self.lib.icfh_csr_destroy(self.h)
        except Exception:
            pass


class GPUFunctionality:
    def __init__(self, device=0):
        self.lib = _lib.load()
        ctx = ctypes.c_void_p()
        check(self.lib.icf_create(device, ctypes.byref(ctx)), what="icf_create")
        self.ctx = ctx
        self.csr = CSR(self.lib)
        self.nImpacts = 0
        self._bufs = {}

    def close(self):
        for b in self._bufs.values():
            b.free()
        self._bufs.clear()
        if self.ctx is not None:
            self.lib.icf_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- buffers (grow-only, AllocateMemoryForMesh GPU_Functionality.cs:196-249) --
    def buf(self, name, n, dtype):
        b = self._bufs.get(name)
        if b is None or b.n < n or b.dtype != np.dtype(dtype):
            if b is not None:
                b.free()
            b = DeviceArray(self, n, dtype)
            self._bufs[name] = b
        return b

    # -- SetConstants (:151-194) --
    def SetConstants(self, consts):
        self.consts = consts
        check(self.lib.icf_set_constants(self.ctx, ctypes.byref(consts)), self.ctx, "icf_set_constants")

    def set_precision(self, mode):
        check(self.lib.icf_set_precision(self.ctx, mode), self.ctx, "icf_set_precision")

    # -- autotuner of the BCSR-value reductions' chunk shape (no reference counterpart) --
    ROLES = ("overwrite,fp32", "overwrite,fp64", "accumulate,fp32", "accumulate,fp64")
    KEY_CANDIDATES = (64, 128, 256, 512)
    STAGE_CANDIDATES = (2, 3)
    CANDIDATES = ((64, 2), (64, 3), (128, 2), (128, 3), (256, 2), (256, 3), (512, 2), (512, 3))   # index c

    def set_autotune(self, rounds):
        check(self.lib.icf_set_autotune(self.ctx, rounds), self.ctx, "icf_set_autotune")

    def autotune_state(self):
        """{role: ((keys, stage) in use, tuning?, {(keys, stage): fastest ms})} for the four
        roles."""
        nc = len(self.CANDIDATES)
        keys, stage, tuning = np.zeros(4, np.int32), np.zeros(4, np.int32), np.zeros(4, np.int32)
        best = np.zeros(4 * nc, np.float32)
        check(self.lib.icf_get_autotune(self.ctx, _lib.ptr(keys), _lib.ptr(stage), _lib.ptr(tuning), _lib.ptr(best)),
              self.ctx, "icf_get_autotune")
        return {r: ((int(keys[i]), int(stage[i])), bool(tuning[i]),
                    {c: float(best[i * nc + j]) for j, c in enumerate(self.CANDIDATES)})
                for i, r in enumerate(self.ROLES)}

    def set_chunk_keys(self, accumulate, fp64, keys, stage=0):
        check(self.lib.icf_set_chunk_keys(self.ctx, int(accumulate), int(fp64), keys, stage), self.ctx,
              "icf_set_chunk_keys")

    # launch-shape roles besides the value reductions (icf_set_shape / icf_get_shape)
    SHAPE_ROLES = ("phase1", "narrow", "closest", "impact", "order")
    SHAPE_CANDIDATES = {"phase1": (64, 128, 256), "narrow": (64, 128, 256), "closest": (4, 8, 16),
                        "impact": (64, 128, 256), "order": (0, 1)}

    def set_shape(self, role, value):
        check(self.lib.icf_set_shape(self.ctx, self.SHAPE_ROLES.index(role), int(value)), self.ctx, "icf_set_shape")

    def shape_state(self):
        """{role: (value in use, tuning?, {candidate: fastest ms})} for the shape roles."""
        out = {}
        for i, r in enumerate(self.SHAPE_ROLES):
            v, t = ctypes.c_int(), ctypes.c_int()
            best = np.zeros(3, np.float32)
            check(self.lib.icf_get_shape(self.ctx, i, ctypes.byref(v), ctypes.byref(t), _lib.ptr(best)), self.ctx,
                  "icf_get_shape")
            out[r] = (int(v.value), bool(t.value), {c: float(best[j]) for j, c in enumerate(self.SHAPE_CANDIDATES[r])})
        return out

    def tuned_shapes(self):
        """{role: [keys, stage]} for the value reductions and {role: value} for the other launch
        shapes in use, to keep for a later run on the same specimen."""
        out = {r: list(v[0]) for r, v in self.autotune_state().items()}
        out.update({r: v[0] for r, v in self.shape_state().items()})
        return out

    def use_shapes(self, shapes):
        """Fix the shapes of tuned_shapes() (for example read back from JSON) without tuning."""
        for i, r in enumerate(self.ROLES):
            if r in shapes:
                keys, stage = shapes[r]
                self.set_chunk_keys(i >> 1, i & 1, int(keys), int(stage))
        for r in self.SHAPE_ROLES:
            if r in shapes:
                self.set_shape(r, shapes[r])

    # -- TransferStaticDataToDevice (:255-302) --
    def TransferStaticDataToDevice(self, col):
        self.col = col
        self.nd_stride = stride_of(len(col.x0))
        self.el_all_stride = stride_of(len(col.surf))
        self.el_elastic_stride = stride_of(len(col.elastic))
        self.fc_stride = stride_of(len(col.faces))
        self.cz_stride = stride_of(len(col.czs))
        self.maxAdjacentFaces = int(np.diff(col.adj_ptr).max()) if len(col.surf) else 0
        es, S = self.el_all_stride, len(col.surf)
        c_ie = np.zeros((5 + self.maxAdjacentFaces) * es, np.int32)
        for i in range(4):
            c_ie[i * es:i * es + S] = col.surf[:, i]
        cnt = np.diff(col.adj_ptr)
        c_ie[4 * es:4 * es + S] = cnt
        rows = np.repeat(np.arange(S), cnt)
        pos = np.arange(len(col.adj_faces)) - col.adj_ptr[rows]
        c_ie[(5 + pos) * es + rows] = col.adj_faces
        self.c_ie = c_ie
        self.g_ie = self.buf("ie", len(c_ie), np.int32)
        self.g_ie.upload(c_ie)
        fs, F = self.fc_stride, len(col.faces)
        c_ifc = np.zeros(3 * fs, np.int32)
        for v in range(3):
            c_ifc[v * fs:v * fs + F] = col.faces[:, v]
        self.c_ifc = c_ifc
        self.g_ifc = self.buf("ifc", len(c_ifc), np.int32)
        self.g_ifc.upload(c_ifc)
        self.g_dn = self.buf("dn", FP_DATA_SIZE_NODE * self.nd_stride, np.float64)
        self.g_ie_pcsr = self.buf("ie_pcsr", INT_DATA_SIZE_ELEM * max(self.el_elastic_stride, 1), np.int32)
        if self.cz_stride:
            self.g_dcz = self.buf("dcz", FP_DATA_SIZE_CZ * self.cz_stride, np.float64)
            self.g_icz = self.buf("icz", INT_DATA_SIZE_CZ * self.cz_stride, np.int32)

    # -- TransferNodesToGPU (:307-343): F zeroed, un/vn/an/x0/x_current --
    def pack_nodes(self, state):
        ns, Nn = self.nd_stride, len(self.col.x0)
        c_dn = np.zeros(FP_DATA_SIZE_NODE * ns)
        for k in range(3):
            c_dn[(X0 + k) * ns:(X0 + k) * ns + Nn] = self.col.x0[:, k]
            c_dn[(UN + k) * ns:(UN + k) * ns + Nn] = state.un[:, k]
            c_dn[(VN + k) * ns:(VN + k) * ns + Nn] = state.vn[:, k]
            c_dn[(AN + k) * ns:(AN + k) * ns + Nn] = state.an[:, k]
            c_dn[(XC + k) * ns:(XC + k) * ns + Nn] = state.xc[:, k]
        return c_dn

    def TransferNodesToGPU(self, state):
        self.c_dn = self.pack_nodes(state)
        self.g_dn.upload(self.c_dn)

    # -- LinearSystem.CreateStructure (LinearSystem.cs:29-43) --
    def CreateStructure(self, impacts_nodes=None, static_czs=None):
        col = self.col
        alt = col.alt
        cz_for_pattern = col.czs if static_czs is None else static_czs
        alt_imp = alt[impacts_nodes] if impacts_nodes is not None and len(impacts_nodes) else None
        self.csr.build(alt[col.elastic], alt[cz_for_pattern], alt_imp)

    # -- TransferPCSR (:345-416) --
    def pack_pcsr(self, state, cz_current_failed=None):
        col, es, cs = self.col, self.el_elastic_stride, self.cz_stride
        alt32 = np.ascontiguousarray(col.alt, np.int32)
        c_ie_pcsr = np.zeros(INT_DATA_SIZE_ELEM * max(es, 1), np.int32)
        nodes = np.ascontiguousarray(col.elastic, np.int32)
        check(self.lib.icfh_pack_elems(self.csr.h, ptr(nodes), ptr(alt32), len(nodes), es, ptr(c_ie_pcsr)),
              what="icfh_pack_elems")
        c_icz = c_dcz = None
        if cs:
            c_icz = np.zeros(INT_DATA_SIZE_CZ * cs, np.int32)
            cz = np.ascontiguousarray(col.czs, np.int32)
            check(self.lib.icfh_pack_czs(self.csr.h, ptr(cz), ptr(alt32), len(cz), cs, ptr(c_icz)),
                  what="icfh_pack_czs")
            if cz_current_failed is not None:
                c_icz[:len(cz)] = cz_current_failed
            c_dcz = np.zeros(FP_DATA_SIZE_CZ * cs)
            C = len(cz)
            for j in range(3):
                c_dcz[j * cs:j * cs + C] = state.pmax[:, j]
                c_dcz[(3 + j) * cs:(3 + j) * cs + C] = state.tmax[:, j]
        return c_ie_pcsr, c_icz, c_dcz

    def TransferPCSR(self, state, cz_current_failed=None):
        self.c_ie_pcsr, self.c_icz, self.c_dcz = self.pack_pcsr(state, cz_current_failed)
        self.g_ie_pcsr.upload(self.c_ie_pcsr)
        if self.cz_stride:
            self.g_icz.upload(self.c_icz)
            self.g_dcz.upload(self.c_dcz)

    # -- AllocateMemoryForLinearSystem (:419-437) + AssembleElemsAndCZs (:439-479) --
    def AssembleElemsAndCZs(self, h):
        nnz, N = self.csr.nnz, self.csr.N
        self.g_dvals = self.buf("vals", max(1, 2 * nnz * 9), np.float64)
        self.g_drhs = self.buf("rhs", max(1, N * 3), np.float64)
        dmg, fld = ctypes.c_int(), ctypes.c_int()
        nE, nC = len(self.col.elastic), len(self.col.czs)
        check(self.lib.icf_assemble_elems_and_czs(
            self.ctx, self.g_ie_pcsr.p, nE, self.el_elastic_stride,
            self.g_dcz.p if nC else None, self.g_icz.p if nC else None, nC, self.cz_stride,
            self.g_dn.p, self.nd_stride, h, self.g_dvals.p, nnz, self.g_drhs.p, N,
            ctypes.byref(dmg), ctypes.byref(fld)), self.ctx, "icf_assemble_elems_and_czs")
        self.nCZDamaged, self.nCZFailedThisStep = dmg.value, fld.value
        return self.nCZDamaged, self.nCZFailedThisStep

    def AssembleBatched(self, h, tet_off, cz_off, row_off, nnzb_off):
        """icf_assemble_batched over the collection's specimens (per-specimen offsets, each
        n + 1 ints from 0); returns the per-specimen damaged and failed CZ counts."""
        nnz, N = self.csr.nnz, self.csr.N
        self.g_dvals = self.buf("vals", max(1, 2 * nnz * 9), np.float64)
        self.g_drhs = self.buf("rhs", max(1, N * 3), np.float64)
        offs = [np.ascontiguousarray(a, np.int32) for a in (tet_off, cz_off, row_off, nnzb_off)]
        n = len(offs[0]) - 1
        wl = _lib.IcfWorklist(n, *[ptr(a) for a in offs])
        dmg, fld = np.zeros(n, np.int32), np.zeros(n, np.int32)
        nC = int(offs[1][-1])
        check(self.lib.icf_assemble_batched(
            self.ctx, ctypes.byref(wl), self.g_ie_pcsr.p, self.el_elastic_stride,
            self.g_dcz.p if nC else None, self.g_icz.p if nC else None, self.cz_stride,
            self.g_dn.p, self.nd_stride, h, self.g_dvals.p, self.g_drhs.p, ptr(dmg), ptr(fld)),
            self.ctx, "icf_assemble_batched")
        return dmg, fld

    def last_device_ms(self):
        return self.lib.icf_last_device_ms(self.ctx)

    # -- NarrowPhaseCollisionDetection (:539-616) --
    def NarrowPhaseCollisionDetection(self, pairs):
        """pairs: (Q, 2) surface-tet ids from the broad phase.  Returns the impacts as
        (nImpacts, 4) global node ids (point, face v0, v1, v2)."""
        pairs = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
        Q = len(pairs)
        self.nImpacts = 0
        if Q == 0:
            self.impacts = np.zeros((0, 4), np.int32)
            return self.impacts
        g_itet = self.buf("itet", 2 * Q, np.int32)
        g_itet.upload(pairs.ravel())
        check(self.lib.icf_narrow_phase(self.ctx, Q, self.g_dn.p, self.g_ie.p, g_itet.p, self.el_all_stride,
                                        self.nd_stride), self.ctx, "icf_narrow_phase")
        masks = g_itet.download(2 * Q)
        self.masks = masks[0::2].copy()
        # NL2set
        cap = int(np.count_nonzero(self.masks)) * 8
        ts = stride_of(max(cap, 1))
        tuples = np.zeros(4 * ts, np.int32)
        n = ctypes.c_int()
        check(self.lib.icfh_nl2set(ptr(masks), ptr(pairs.ravel().copy()), Q, ptr(self.c_ie), self.el_all_stride,
                                   ptr(tuples), ts, cap, ctypes.byref(n)), what="icfh_nl2set")
        nT = n.value
        if nT == 0:
            self.impacts = np.zeros((0, 4), np.int32)
            return self.impacts
        self.tet_stride = stride_of(nT)
        c_itet = np.zeros(4 * self.tet_stride, np.int32)
        c_itet[:nT] = tuples[:nT]
        c_itet[self.tet_stride:self.tet_stride + nT] = tuples[ts:ts + nT]
        g2 = self.buf("itet2", 4 * self.tet_stride, np.int32)
        g2.upload(c_itet)
        check(self.lib.icf_find_closest_face(self.ctx, nT, self.tet_stride, g2.p, self.g_dn.p, self.g_ie.p,
                                             self.g_ifc.p, self.nd_stride, self.el_all_stride, self.fc_stride),
              self.ctx, "icf_find_closest_face")
        self.c_itet = g2.download(4 * self.tet_stride)
        self.nImpacts = nT
        self.impacts = np.stack([self.c_itet[k * self.tet_stride:k * self.tet_stride + nT] for k in range(4)], axis=1)
        return self.impacts

    # -- collisionResponse (:627-676) --
    def pack_icr(self, impacts):
        n = len(impacts)
        cr = stride_of(n)
        ts = cr
        nl2 = np.zeros(4 * ts, np.int32)
        for k in range(4):
            nl2[k * ts:k * ts + n] = impacts[:, k]
        icr = np.zeros(28 * cr, np.int32)
        alt32 = np.ascontiguousarray(self.col.alt, np.int32)
        check(self.lib.icfh_pack_impacts(self.csr.h, ptr(nl2), ts, ptr(alt32), n, cr, ptr(icr)),
              what="icfh_pack_impacts")
        return icr, cr

    def collisionResponse(self, h, k, eps):
        if self.nImpacts == 0:
            return
        self.c_icr, self.cr_stride = self.pack_icr(self.impacts)
        g_icr = self.buf("icr", len(self.c_icr), np.int32)
        g_icr.upload(self.c_icr)
        self.g_icr = g_icr
        check(self.lib.icf_collision_response(self.ctx, g_icr.p, self.g_dn.p, h, self.g_dvals.p, self.g_drhs.p,
                                              self.nImpacts, self.cr_stride, self.nd_stride, k, eps),
              self.ctx, "icf_collision_response")

    # -- TransferLinearSystemToHost (:481-485) --
    def TransferLinearSystemToHost(self):
        vals = self.g_dvals.download(self.csr.nnz * 9)
        rhs = self.g_drhs.download(self.csr.N * 3)
        return vals, rhs

    # -- TransferUpdatedStateToHost (:487-528) --
    def TransferUpdatedStateToHost(self):
        out = {}
        if self.cz_stride:
            out["dcz"] = self.g_dcz.download(FP_DATA_SIZE_CZ * self.cz_stride)
            out["icz"] = self.g_icz.download(4 * self.cz_stride)
        ns = self.nd_stride
        out["F"] = self.g_dn.download(3 * ns, offset_elems=FF * ns)
        return out


def broad_phase(gf_or_lib, c_dn, nd_stride, c_ie, el_all_stride, n_surf, exclude_shared=True):
    """Candidate surface-tet pairs (Q, 2) by k-DOP overlap (icfh_kdop_pairs)."""
    lib = gf_or_lib.lib if hasattr(gf_or_lib, "lib") else gf_or_lib
    cap = max(1024, 64 * n_surf)
    while True:
        out = np.zeros(2 * cap, np.int32)
        n = ctypes.c_longlong()
        r = lib.icfh_kdop_pairs(ptr(c_dn), nd_stride, ptr(c_ie), el_all_stride, n_surf, int(exclude_shared),
                                ptr(out), cap, ctypes.byref(n))
        if r == 0:
            return out[:2 * n.value].reshape(-1, 2)
        if r != -5:
            check(r, what="icfh_kdop_pairs")
        cap = int(n.value) + 16

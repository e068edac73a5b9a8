"""The implicit Newmark step of ImplicitModel3.Step (v3Library/Simulation/ImplicitModel3.cs:
377-435) driven through libicflow_hip.so, with everything but the linear solve on the device.

Per step:
  _beginStep (:117-156)          prescribed displacements of anchored nodes, du = v h for the rest
  loop (:383-413)                icf_nodes_infer (Node.InferTentativeValues), fused assembly
                                 (AssembleElemsAndCZs), solve, _checkDivergence, icf_nodes_add_dx
                                 (_XtoDU), while it < minIterations or (not converged and
                                 it < maxIterations and not diverging)
  accept (:417-427)              icf_nodes_accept (Node.AcceptTentativeValues); the CZs' tentative
                                 history becomes current; failed CZs leave the CZ set and the
                                 pattern is rebuilt (_acceptFrame, :290-300)

The linear solve stands in for MKL PARDISO (LinearSystem.Solve, LinearSystem.cs:45-64,
mtype -2 on the upper-triangular BCSR): the symmetric matrix is expanded from the upper
blocks and factored with SuperLU (scipy.sparse.linalg.splu).  It is a host substitute and not
PARDISO.  The time step is fixed (the reference's step-size adaptation, :258-288, is not
restated), so every step is accepted.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .gpu_functionality import GPUFunctionality, stride_of


@dataclass
class StepParams:
    """ModelPrms fields the step loop reads (ModelPrms.cs:180-209)."""
    h: float = 1e-4
    beta: float = 0.25
    gamma: float = 0.5
    conv_eps: float = 0.005      # ConvergenceEpsilon
    conv_cutoff: float = 1e-8    # ConvergenceCutoff
    min_iter: int = 3
    max_iter: int = 10


@dataclass
class StepRecord:
    iterations: int
    converged: bool
    failed_czs: int
    damaged: int
    norms: list = field(default_factory=list)


def upper_bcsr_to_csc(rows, cols, vals):
    """The symmetric matrix whose upper-triangular 3x3 blocks (diagonal blocks whole) are
    given in BCSR form -- the PARDISO mtype -2 input of LinearSystem.Solve -- as scipy CSC."""
    import scipy.sparse as sp
    N = len(rows) - 1
    V = np.asarray(vals, dtype=np.float64).reshape(-1, 3, 3)
    br = np.repeat(np.arange(N), np.diff(rows))
    r = (3 * br)[:, None, None] + np.arange(3)[None, :, None] + np.zeros((1, 1, 3), np.int64)
    c = (3 * np.asarray(cols))[:, None, None] + np.arange(3)[None, None, :] + np.zeros((1, 3, 1), np.int64)
    off = br != np.asarray(cols)
    ri = np.concatenate([r.ravel(), c[off].ravel()])
    ci = np.concatenate([c.ravel(), r[off].ravel()])
    vi = np.concatenate([V.ravel(), V[off].ravel()])
    return sp.coo_matrix((vi, (ri, ci)), shape=(3 * N, 3 * N)).tocsc()


def solve_upper_bcsr(rows, cols, vals, rhs):
    """x with A x = rhs for the symmetric A of upper_bcsr_to_csc.  Host substitute for
    PARDISO (LinearSystem.cs:45-64): SuperLU with a column minimum-degree ordering."""
    from scipy.sparse.linalg import splu
    return splu(upper_bcsr_to_csc(rows, cols, vals), permc_spec="COLAMD").solve(np.asarray(rhs, np.float64))


class UpperBcsrSolver:
    """Host substitute for PARDISO on one BCSR pattern: reverse Cuthill-McKee ordering of the
    3x3 block graph (computed once per pattern), then a banded Cholesky factorisation
    (LAPACK pbsv through scipy.linalg.solveh_banded); if the matrix is not positive definite,
    or its band is too wide, SuperLU (solve_upper_bcsr).  Not PARDISO."""

    def __init__(self, rows, cols, max_band_entries=2e8):
        import scipy.sparse as sp
        from scipy.sparse.csgraph import reverse_cuthill_mckee
        self.rows, self.cols = np.asarray(rows), np.asarray(cols)
        N = len(rows) - 1
        self.n = 3 * N
        br = np.repeat(np.arange(N), np.diff(rows))
        G = sp.coo_matrix((np.ones(len(cols)), (br, cols)), shape=(N, N)).tocsr()
        perm = reverse_cuthill_mckee((G + G.T).tocsr(), symmetric_mode=True)
        inv = np.empty(N, np.int64)
        inv[perm] = np.arange(N)
        # scalar positions of every stored entry after the permutation, folded to the upper
        # triangle; of a diagonal block only its upper half is taken (the block is symmetric)
        k = np.arange(3)
        r = (3 * inv[br])[:, None, None] + k[None, :, None] + np.zeros((1, 1, 3), np.int64)
        c = (3 * inv[self.cols])[:, None, None] + k[None, None, :] + np.zeros((1, 3, 1), np.int64)
        keep = np.ones(r.shape, bool)
        diag = br == self.cols
        keep[diag] = np.triu(np.ones((3, 3), bool))[None]
        lo, hi = np.minimum(r, c)[keep], np.maximum(r, c)[keep]
        self.take = np.flatnonzero(keep.ravel())
        self.bw = int((hi - lo).max()) if len(lo) else 0
        self.banded = (self.bw + 1) * self.n <= max_band_entries
        self.ab_row, self.ab_col = self.bw + lo - hi, hi
        self.perm = (3 * perm[:, None] + k[None, :]).ravel()

    def solve(self, vals, rhs):
        import scipy.linalg as la
        rhs = np.asarray(rhs, np.float64)
        if self.banded:
            ab = np.zeros((self.bw + 1, self.n))
            ab[self.ab_row, self.ab_col] = np.asarray(vals, np.float64).ravel()[self.take]
            try:
                y = la.solveh_banded(ab, rhs[self.perm], check_finite=False)
                x = np.empty_like(y)
                x[self.perm] = y
                return x
            except la.LinAlgError:
                pass
        return solve_upper_bcsr(self.rows, self.cols, vals, rhs)


class ImplicitModel:
    """One specimen's implicit dynamics on the device.  `prescribed(t)` returns, for the
    anchored deformable nodes, the target displacement (n_anchored, 3) at time t (the role
    of the anchored SurfaceFragments' dx, dy, dz with their application time)."""

    def __init__(self, col, consts, precision=_lib.ICF_FP64, step=None, prescribed=None):
        self.col, self.consts = col, consts
        self.p = step or StepParams()
        self.prescribed = prescribed
        gf = self.gf = GPUFunctionality(0)
        gf.SetConstants(consts)
        gf.set_precision(precision)
        self.lib, self.ctx = gf.lib, gf.ctx
        n = len(col.x0)
        self.n = n
        self.ns = stride_of(n)
        # node buffer: x0 (x_current follows from the state)
        dn = np.zeros(18 * self.ns)
        for k in range(3):
            dn[k * self.ns:k * self.ns + n] = col.x0[:, k]
        self.g_dn = gf.buf("dn", dn.size, np.float64)
        self.g_dn.upload(dn)
        self.st_stride = self.ns
        self.g_st = gf.buf("st", 12 * self.st_stride, np.float64)
        self.g_st.zero()
        self.alt = np.ascontiguousarray(col.alt, np.int32)
        self.g_alt = gf.buf("alt", max(1, n), np.int32)
        self.g_alt.upload(self.alt)
        self.N = int(self.alt.max()) + 1
        self.anchored_ids = np.flatnonzero(col.anchored & (col.node_mesh == 0))
        # elements (static), CZs (shrinks as they fail)
        T = len(col.elastic)
        self.T, self.es = T, stride_of(T)
        ie = np.zeros(28 * self.es, np.int32)
        for i in range(4):
            ie[i * self.es:i * self.es + T] = col.elastic[:, i]
        self.g_ie = gf.buf("ie", ie.size, np.int32)
        self.g_ie.upload(ie)
        self.cz_nodes = np.asarray(col.czs, np.int64)
        self.pmax = np.zeros((len(self.cz_nodes), 3))
        self.tmax = np.zeros((len(self.cz_nodes), 3))
        self.cz_alive = np.arange(len(self.cz_nodes))   # indices into the original CZ list
        self.t = 0.0
        self.records = []
        self._setup_czs()

    # -- MeshCollection.UpdateStaticStructureData + CreateStructure + TransferPCSR --
    def _setup_czs(self):
        gf = self.gf
        C = len(self.cz_alive)
        self.C, self.cs = C, stride_of(C) if C else 0
        if C:
            icz = np.zeros(58 * self.cs, np.int32)
            nodes = self.cz_nodes[self.cz_alive]
            for i in range(6):
                icz[(4 + i) * self.cs:(4 + i) * self.cs + C] = nodes[:, i]
            self.g_icz = gf.buf("icz", icz.size, np.int32)
            self.g_icz.upload(icz)
            self.g_dcz = gf.buf("dcz", 16 * self.cs, np.float64)
        cap = 10 * self.T + 21 * C + 1
        self.g_rows = gf.buf("rows", self.N + 1, np.int32)
        self.g_cols = gf.buf("cols", cap, np.int32)
        nnz = ctypes.c_int()
        _lib.check(self.lib.icf_build_pattern(
            self.ctx, self.g_alt.p, self.n, self.g_ie.p, self.T, self.es, self.g_icz.p if C else None, C, self.cs,
            None, 0, 0, self.N, self.g_rows.p, self.g_cols.p, cap, ctypes.byref(nnz)), self.ctx, "icf_build_pattern")
        self.nnz = nnz.value
        _lib.check(self.lib.icf_pack_pcsr(
            self.ctx, self.g_alt.p, self.n, self.g_rows.p, self.g_cols.p, self.N, self.g_ie.p, self.T, self.es,
            self.g_icz.p if C else None, C, self.cs, None, 0, 0), self.ctx, "icf_pack_pcsr")
        self.rows = self.g_rows.download(self.N + 1)
        self.cols = self.g_cols.download(self.nnz)
        self.solver = UpperBcsrSolver(self.rows, self.cols)
        self.g_vals = gf.buf("vals", 9 * self.nnz, np.float64)
        self.g_rhs = gf.buf("rhs", 3 * self.N, np.float64)
        self.g_dx = gf.buf("dx", 3 * self.N, np.float64)

    def _upload_history(self):
        if not self.C:
            return
        d = np.zeros(16 * self.cs)
        pm, tm = self.pmax[self.cz_alive], self.tmax[self.cz_alive]
        for j in range(3):
            d[j * self.cs:j * self.cs + self.C] = pm[:, j]
            d[(3 + j) * self.cs:(3 + j) * self.cs + self.C] = tm[:, j]
        self.g_dcz.upload(d)

    def _begin_step(self, h):
        """_beginStep (ImplicitModel3.cs:117-156): du of anchored nodes = target - u; du = v h
        for active nodes."""
        ss = self.st_stride
        st = self.g_st.download(12 * ss)
        u = np.stack([st[k * ss:k * ss + self.n] for k in range(3)], 1)
        v = np.stack([st[(3 + k) * ss:(3 + k) * ss + self.n] for k in range(3)], 1)
        du = np.zeros_like(u)
        act = self.alt >= 0
        du[act] = v[act] * h
        if self.prescribed is not None and len(self.anchored_ids):
            tgt = self.prescribed(self.t + h)
            du[self.anchored_ids] = tgt - u[self.anchored_ids]
        for k in range(3):
            st[(9 + k) * ss:(9 + k) * ss + self.n] = du[:, k]
        self.g_st.upload(st)

    def step(self):
        p, lib, ctx = self.p, self.lib, self.ctx
        h = p.h
        self._begin_step(h)
        it, err0, converged, diverges = 0, None, False, False
        norms = []
        dmg = ctypes.c_int()
        fld = ctypes.c_int()
        while True:
            _lib.check(lib.icf_nodes_infer(ctx, self.g_st.p, self.st_stride, self.g_dn.p, self.ns, self.n, h, p.beta,
                                           p.gamma), ctx, "icf_nodes_infer")
            if self.C:
                # TransferPCSR, each iteration: current history in, tentative fields and the CZ
                # flags cleared (GPU_Functionality.cs:376-412)
                self._upload_history()
                _lib.check(lib.icf_pack_pcsr(ctx, self.g_alt.p, self.n, self.g_rows.p, self.g_cols.p, self.N,
                                             None, 0, 0, self.g_icz.p, self.C, self.cs, None, 0, 0), ctx, "pack")
            _lib.check(lib.icf_assemble_elems_and_czs(
                ctx, self.g_ie.p, self.T, self.es, self.g_dcz.p if self.C else None,
                self.g_icz.p if self.C else None, self.C, self.cs, self.g_dn.p, self.ns, h, self.g_vals.p, self.nnz,
                self.g_rhs.p, self.N, ctypes.byref(dmg), ctypes.byref(fld)), ctx, "assemble")
            vals = self.g_vals.download(9 * self.nnz)
            rhs = self.g_rhs.download(3 * self.N)
            dx = self.solver.solve(vals, rhs)
            norm = float(dx @ dx)                       # LinearSystem.NormOfDx (:68-74)
            norms.append(norm)
            # _checkDivergence (ImplicitModel3.cs:214-245)
            if it == 0:
                err0 = norm if norm != 0 else p.conv_eps
            elif norm < p.conv_cutoff:
                converged = True
            else:
                rel = np.sqrt(norm / err0)
                converged = rel <= p.conv_eps
                diverges = rel > 1.01
            self.g_dx.upload(np.ascontiguousarray(dx))
            _lib.check(lib.icf_nodes_add_dx(ctx, self.g_st.p, self.st_stride, self.g_alt.p, self.n, self.g_dx.p),
                       ctx, "icf_nodes_add_dx")
            it += 1
            if not (it < p.min_iter or (not diverges and not converged and it < p.max_iter)):
                break
        _lib.check(lib.icf_nodes_accept(ctx, self.g_st.p, self.st_stride, self.g_dn.p, self.ns, self.n, h), ctx,
                   "icf_nodes_accept")
        failed_now = 0
        if self.C:
            d = self.g_dcz.download(16 * self.cs)
            icz = self.g_icz.download(4 * self.cs)
            for j in range(3):
                self.pmax[self.cz_alive, j] = d[(6 + j) * self.cs:(6 + j) * self.cs + self.C]
                self.tmax[self.cz_alive, j] = d[(9 + j) * self.cs:(9 + j) * self.cs + self.C]
            failed = icz[3 * self.cs:3 * self.cs + self.C] != 0
            failed_now = int(failed.sum())
            if failed_now:
                self.cz_alive = self.cz_alive[~failed]
                self._setup_czs()
        self.t += h
        rec = StepRecord(it, converged, failed_now, dmg.value, norms)
        self.records.append(rec)
        return rec

    def displacements(self):
        ss = self.st_stride
        st = self.g_st.download(3 * ss)
        return np.stack([st[k * ss:k * ss + self.n] for k in range(3)], 1)

    def failed_set(self):
        alive = np.zeros(len(self.cz_nodes), bool)
        alive[self.cz_alive] = True
        return np.flatnonzero(~alive)

    def close(self):
        self.gf.close()


def three_point_bend(consts, cells=(40, 8, 8), size=(0.40, 0.05, 0.05), grains=16, seed=0, deflection=2e-3,
                     ramp_steps=200, h=1e-4):
    """Three-point bend beam: bottom edge node lines at x = 0.02 and x = 0.38 pinned, the top
    midspan node line pushed down by `deflection` over `ramp_steps` steps (linear ramp, then
    held).  Returns the collection and the prescribed-displacement function."""
    from . import specimens
    col = specimens.make_specimen(cells, size, grains, seed)
    x = col.meshes[0].x0
    hx = size[0] / cells[0]
    sup = (np.isclose(x[:, 2], 0.0)) & (np.isclose(x[:, 0], 0.02, atol=hx * 1e-6) |
                                        np.isclose(x[:, 0], size[0] - 0.02, atol=hx * 1e-6))
    top = np.isclose(x[:, 2], size[2]) & np.isclose(x[:, 0], size[0] / 2, atol=hx * 1e-6)
    col.anchored = np.zeros(len(x), bool)
    col.anchored[sup | top] = True
    col.finalize()
    ids = np.flatnonzero(col.anchored & (col.node_mesh == 0))
    is_top = np.isclose(col.x0[ids, 2], size[2])
    T_ramp = ramp_steps * h

    def prescribed(t):
        out = np.zeros((len(ids), 3))
        out[is_top, 2] = -deflection * min(1.0, t / T_ramp)
        return out

    return col, prescribed


def precision_study(consts, steps=20, cells=(40, 8, 8), grains=16, seed=0, h=1e-4, ramp_steps=200, log=None):
    """The same three-point-bend run in FP64 and in mixed precision: relative displacement
    error of the mixed run after `steps` steps (normwise over all nodes), the failed-CZ sets,
    and Newton iterations per step."""
    out = {}
    runs = {}
    for name, mode in (("fp64", _lib.ICF_FP64), ("mixed", _lib.ICF_MIXED)):
        col, presc = three_point_bend(consts, cells=cells, grains=grains, seed=seed, ramp_steps=ramp_steps, h=h)
        m = ImplicitModel(col, consts, precision=mode, step=StepParams(h=h), prescribed=presc)
        for k in range(steps):
            rec = m.step()
            if log and (k + 1) % max(1, steps // 5) == 0:
                log(f"newton {name}: step {k + 1}/{steps}, {rec.iterations} iterations, "
                    f"{len(m.failed_set())} CZs failed")
        runs[name] = (m.displacements(), m.failed_set(), [r.iterations for r in m.records],
                      [r.converged for r in m.records])
        out[f"nodes"] = int(m.n)
        out[f"czs"] = int(len(m.cz_nodes))
        m.close()
    u64, f64, it64, c64 = runs["fp64"]
    umx, fmx, itmx, cmx = runs["mixed"]
    scale = float(np.linalg.norm(u64))
    out.update(steps=steps, dt=h,
               rel_err_displacement=float(np.linalg.norm(umx - u64) / scale) if scale > 0 else 0.0,
               max_abs_displacement=float(np.abs(u64).max()),
               failed_czs_fp64=int(len(f64)), failed_czs_mixed=int(len(fmx)),
               failed_set_identical=bool(np.array_equal(f64, fmx)),
               newton_iterations_fp64=int(sum(it64)), newton_iterations_mixed=int(sum(itmx)),
               steps_converged_fp64=int(sum(c64)), steps_converged_mixed=int(sum(cmx)))
    return out

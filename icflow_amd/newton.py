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
mtype -2 on the upper-triangular BCSR): the library's SolveDouble3, a supernodal sparse LDL^T
behind PARDISO's signature (SolveDouble3Solver; per specimen on host threads for batched
specimens, BatchedSolver).  It is a host substitute and not PARDISO.

Time-step control (StepParams.adapt): the reference's _checkDamage (:198-211), _adjustTimeStep
(:258-288), the acceptance rule of Step (:415-431) and FrameInfo.IncrementTime
(FrameInfo.cs:159-175).  An attempt runs at h0 / tsf (tsf a power of two up to PARTS); an
attempt whose Newton loop diverges, does not converge, or fails or damages more than 1% of the
initial CZs is discarded and retried at twice the factor, and after converged steps the factor
is halved again.  Without it every attempt is accepted at h0.
"""
import ctypes
import os
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .gpu_functionality import GPUFunctionality, stride_of


PARTS = 1024 * 32      # FrameInfo.Parts (= Factor2): the finest step is h0 / PARTS
HOLD_FACTOR_DELAY = 4  # ImplicitModel3.holdFactorDelay


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
    max_damage: float = 0.01     # maxDamagePerStep
    max_fail: float = 0.01       # maxFailPerStep
    adapt: bool = False          # the reference's time-step control (module docstring)


@dataclass
class StepRecord:
    iterations: int
    converged: bool
    failed_czs: int
    damaged: int
    norms: list = field(default_factory=list)
    accepted: bool = True
    h: float = 0.0
    explodes: bool = False


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


class SolveDouble3Solver:
    """LinearSystem.Solve (LinearSystem.cs:45-64) through the library's SolveDouble3 with the
    reference's arguments (ja = block columns, ia = block row starts, mtype -2, iparam4 0,
    dim 3, msglvl 0, check 0).  SolveDouble3 is a sparse LDL^T, not MKL PARDISO; it raises, as
    the reference does on a nonzero PARDISO error."""

    def __init__(self, rows, cols):
        self.lib = _lib.load()
        self.rows = np.ascontiguousarray(rows, np.int32)
        self.cols = np.ascontiguousarray(cols, np.int32)
        self.N = len(self.rows) - 1

    def solve(self, vals, rhs):
        a = np.ascontiguousarray(vals, np.float64)
        b = np.ascontiguousarray(rhs, np.float64)
        x = np.zeros(3 * self.N)
        r = self.lib.SolveDouble3(_lib.ptr(self.cols), _lib.ptr(self.rows), _lib.ptr(a), self.N, _lib.ptr(b),
                                  _lib.ptr(x), -2, 0, 3, 0, 0)
        if r != 0:
            raise RuntimeError(f"SolveDouble3 error {r}")
        return x


SOLVERS = {"SolveDouble3": SolveDouble3Solver, "banded": UpperBcsrSolver}


class BatchedSolver:
    """SolveDouble3 per specimen over the diagonal blocks of a block-diagonal upper BCSR matrix
    (specimen s owns the rows [row_off[s], row_off[s+1]) and the blocks they start), the
    specimens spread over a pool of host threads (largest first), each call single-threaded
    inside (icfh_solve_threads(1)).  The reference solves each specimen in its own process
    (LinearSystem.Solve, LinearSystem.cs:45-64)."""

    def __init__(self, rows, cols, row_off, threads=None):
        self.lib = _lib.load()
        rows = np.ascontiguousarray(rows, np.int64)
        cols = np.ascontiguousarray(cols, np.int64)
        self.row_off = np.asarray(row_off, np.int64)
        self.nz_off = rows[self.row_off]
        self.parts = []
        for s in range(len(self.row_off) - 1):
            r0, r1 = int(self.row_off[s]), int(self.row_off[s + 1])
            z0, z1 = int(self.nz_off[s]), int(self.nz_off[s + 1])
            ja = cols[z0:z1] - r0
            if len(ja) and (ja.min() < 0 or ja.max() >= r1 - r0):
                raise ValueError(f"specimen {s}: a block outside its own rows (the matrix is not block-diagonal)")
            self.parts.append((r0, r1, z0, z1, np.ascontiguousarray(rows[r0:r1 + 1] - z0, np.int32),
                               np.ascontiguousarray(ja, np.int32)))
        self.order = sorted(range(len(self.parts)), key=lambda s: -(self.parts[s][3] - self.parts[s][2]))
        self.threads = threads or min(16, len(os.sched_getaffinity(0)))
        self.pool = ThreadPoolExecutor(self.threads)
        self.solve_s = 0.0

    def solve(self, vals, rhs, active=None):
        """dx for the specimens with active[s] (all by default); zero rows for the others."""
        vals = np.ascontiguousarray(vals, np.float64)
        rhs = np.ascontiguousarray(rhs, np.float64)
        dx = np.zeros_like(rhs)
        lib = self.lib

        def one(s):
            r0, r1, z0, z1, ia, ja = self.parts[s]
            if r1 == r0:
                return 0
            x = np.zeros(3 * (r1 - r0))
            r = lib.SolveDouble3(_lib.ptr(ja), _lib.ptr(ia), _lib.ptr(vals[9 * z0:9 * z1]), r1 - r0,
                                 _lib.ptr(rhs[3 * r0:3 * r1]), _lib.ptr(x), -2, 0, 3, 0, 0)
            dx[3 * r0:3 * r1] = x
            return r

        todo = [s for s in self.order if active is None or active[s]]
        t0 = time.perf_counter()
        lib.icfh_solve_threads(1)
        try:
            codes = list(self.pool.map(one, todo))
        finally:
            lib.icfh_solve_threads(0)
        self.solve_s += time.perf_counter() - t0
        bad = [(s, c) for s, c in zip(todo, codes) if c]
        if bad:
            raise RuntimeError(f"SolveDouble3 failed for specimens {bad[:5]}")
        return dx

    def close(self):
        self.pool.shutdown()


def adjust_time_step(tsf, hold, explodes, diverges, converged, max_iter):
    """_adjustTimeStep (ImplicitModel3.cs:258-288) and the acceptance rule of Step (:415-420)
    for an attempt run at time-scale factor tsf with hold count hold.  Returns (accepted,
    factor, hold) for what comes next: the discarded frame's (cf) retry values, or the accepted
    frame's (tcf0) values.  Only cf is changed by an exploding, diverging or unconverged
    attempt; only tcf0 by a converged one."""
    cf_tsf, cf_hold, t_tsf, t_hold = tsf, hold, tsf, hold
    if explodes:
        cf_tsf, cf_hold = min(cf_tsf * 2, PARTS), HOLD_FACTOR_DELAY
    elif (diverges or not converged) and t_tsf < PARTS and max_iter > 1:
        cf_tsf, cf_hold = cf_tsf * 2, HOLD_FACTOR_DELAY
    elif converged and cf_tsf > 1:
        if t_hold > 0:
            t_hold -= 1
        else:
            t_tsf, t_hold = max(1, t_tsf // 2), 2
    accepted = (max_iter == 1 or t_tsf == PARTS or (not converged and t_tsf >= PARTS)
                or (not explodes and not diverges and converged))
    return (True, t_tsf, t_hold) if accepted else (False, cf_tsf, cf_hold)


class ImplicitModel:
    """One specimen's implicit dynamics on the device.  `prescribed(t)` returns, for the
    anchored deformable nodes, the target displacement (n_anchored, 3) at time t (the role
    of the anchored SurfaceFragments' dx, dy, dz with their application time)."""

    def __init__(self, col, consts, precision=_lib.ICF_FP64, step=None, prescribed=None, prescribed_nodes=None,
                 contact=False, penalty=None, dist_eps=1e-15, solver="SolveDouble3", perturb=0.0, perturb_seed=0):
        self.col, self.consts = col, consts
        # sensitivity studies: every assembled BCSR value is multiplied by (1 + perturb * N(0,1))
        # before the solve
        self.perturb, self._rng = perturb, np.random.default_rng(perturb_seed)
        self.solver_cls = SOLVERS[solver]
        self.p = step or StepParams()
        self.prescribed = prescribed
        gf = self.gf = GPUFunctionality(0)
        gf.SetConstants(consts)
        gf.set_precision(precision)
        self.lib, self.ctx = gf.lib, gf.ctx
        n = len(col.x0)
        self.n = n
        self.ns = stride_of(n)
        # contact (ImplicitModel3.cs:386-398): the collection's surface tets, adjacent faces and
        # faces on the device; candidate pairs, narrow phase, NL2set, closest face and the
        # impact records are all device calls each iteration
        self.contact = contact
        self.penalty, self.dist_eps = penalty, dist_eps
        self.nImpacts = 0
        if contact:
            gf.TransferStaticDataToDevice(col)
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
        self.anchored_ids = (np.flatnonzero(col.anchored & (col.node_mesh == 0)) if prescribed_nodes is None
                             else np.asarray(prescribed_nodes, np.int64))
        # elements (static), CZs (shrinks as they fail)
        T = len(col.elastic)
        self.T, self.es = T, stride_of(T)
        ie = np.zeros(28 * self.es, np.int32)
        for i in range(4):
            ie[i * self.es:i * self.es + T] = col.elastic[:, i]
        self.g_ie = gf.buf("ie_el", ie.size, np.int32)
        self.g_ie.upload(ie)
        self.cz_nodes = np.asarray(col.czs, np.int64)
        self.pmax = np.zeros((len(self.cz_nodes), 3))
        self.tmax = np.zeros((len(self.cz_nodes), 3))
        self.cz_alive = np.arange(len(self.cz_nodes))   # indices into the original CZ list
        self.t = 0.0
        self.records = []
        # FrameInfo fields of the current frame (time-step control)
        self.tsf, self.hold, self.int_time, self.frames = 1, 0, 0, 0
        self.nCZ_initial = len(self.cz_nodes)
        self.prev_damaged = 0
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
        self.solver = self.solver_cls(self.rows, self.cols)
        self.g_vals = gf.buf("vals", 9 * self.nnz, np.float64)
        self.g_rhs = gf.buf("rhs", 3 * self.N, np.float64)
        self.g_dx = gf.buf("dx", 3 * self.N, np.float64)

    def _detect_contacts(self):
        """bvh.ConstructAndTraverse + NarrowPhaseCollisionDetection + the impact packing of
        collisionResponse (ImplicitModel3.cs:389-391, GPU_Functionality.cs:539-616,627-665), on
        the device: icf_broad_phase, icf_narrow_phase on a copy of the pairs, icf_nl2set,
        icf_find_closest_face, and the impacts' node rows copied into icr."""
        gf, lib, ctx = self.gf, self.lib, self.ctx
        nS, es = len(self.col.surf), gf.el_all_stride
        n = ctypes.c_longlong()
        cap = getattr(self, "_pair_cap", 1024)
        while True:
            d_pairs = gf.buf("nw_pairs", 2 * cap, np.int32)
            r = lib.icf_broad_phase(ctx, self.g_dn.p, self.ns, gf.g_ie.p, es, nS, 1, d_pairs.p, cap, ctypes.byref(n))
            if r == _lib.ICF_E_CAPACITY:
                cap = int(n.value * 1.25) + 1024
                continue
            _lib.check(r, ctx, "icf_broad_phase")
            break
        self._pair_cap = cap
        Q = int(n.value)
        self.nImpacts = 0
        if Q == 0:
            return
        d_list = gf.buf("nw_list", 2 * cap, np.int32)
        _lib.check(lib.icf_d2d(ctx, d_list.p, 0, d_pairs.p, 0, 8 * Q), ctx, "icf_d2d")
        _lib.check(lib.icf_narrow_phase(ctx, Q, self.g_dn.p, gf.g_ie.p, d_list.p, es, self.ns), ctx,
                   "icf_narrow_phase")
        nt = ctypes.c_int()
        ts = getattr(self, "_tuple_stride", stride_of(1024))
        while True:
            d_nl2 = gf.buf("nw_nl2", 4 * ts, np.int32)
            r = lib.icf_nl2set(ctx, d_list.p, d_pairs.p, Q, gf.g_ie.p, es, d_nl2.p, ts, ctypes.byref(nt))
            if r == _lib.ICF_E_CAPACITY:
                ts = stride_of(int(nt.value * 1.25) + 1024)
                continue
            _lib.check(r, ctx, "icf_nl2set")
            break
        self._tuple_stride = ts
        nT = int(nt.value)
        if nT == 0:
            return
        _lib.check(lib.icf_find_closest_face(ctx, nT, ts, d_nl2.p, self.g_dn.p, gf.g_ie.p, gf.g_ifc.p, self.ns, es,
                                             gf.fc_stride), ctx, "icf_find_closest_face")
        # icr: rows 0-3 the impact's nodes (point, face), slots and rows packed with the pattern
        self.cr_stride = ts
        self.g_icr = gf.buf("nw_icr", 28 * ts, np.int32)
        _lib.check(lib.icf_d2d(ctx, self.g_icr.p, 0, d_nl2.p, 0, 16 * ts), ctx, "icf_d2d")
        self.nImpacts = nT

    def _pattern_with_impacts(self):
        """CreateStructure with the colliding nodes and TransferPCSR of every unit (elements,
        CZs, impacts) against it, on the device; the host solver is set up for the pattern."""
        gf, lib, ctx = self.gf, self.lib, self.ctx
        C, nT = self.C, self.nImpacts
        icr = self.g_icr.p if nT else None
        cr = self.cr_stride if nT else 0
        cap = 10 * self.T + 21 * C + 6 * nT + 1
        self.g_cols = gf.buf("cols", cap, np.int32)
        nnz = ctypes.c_int()
        _lib.check(lib.icf_build_pattern(ctx, self.g_alt.p, self.n, self.g_ie.p, self.T, self.es,
                                         self.g_icz.p if C else None, C, self.cs, icr, nT, cr, self.N, self.g_rows.p,
                                         self.g_cols.p, cap, ctypes.byref(nnz)), ctx, "icf_build_pattern")
        _lib.check(lib.icf_pack_pcsr(ctx, self.g_alt.p, self.n, self.g_rows.p, self.g_cols.p, self.N, self.g_ie.p,
                                     self.T, self.es, self.g_icz.p if C else None, C, self.cs, icr, nT, cr), ctx,
                   "icf_pack_pcsr")
        rows = self.g_rows.download(self.N + 1)
        cols = self.g_cols.download(nnz.value)
        if nnz.value != self.nnz or not np.array_equal(rows, self.rows) or not np.array_equal(cols, self.cols):
            self.nnz, self.rows, self.cols = nnz.value, rows, cols
            self.solver = self.solver_cls(rows, cols)
            self.g_vals = gf.buf("vals", 9 * self.nnz, np.float64)

    def _upload_history(self):
        if not self.C:
            return
        d = np.zeros(16 * self.cs)
        pm, tm = self.pmax[self.cz_alive], self.tmax[self.cz_alive]
        for j in range(3):
            d[j * self.cs:j * self.cs + self.C] = pm[:, j]
            d[(3 + j) * self.cs:(3 + j) * self.cs + self.C] = tm[:, j]
        self.g_dcz.upload(d)

    def _begin_step(self, h, t_new):
        """_beginStep (ImplicitModel3.cs:117-156): du of anchored nodes = target(t_new) - u;
        du = v h for active nodes."""
        ss = self.st_stride
        st = self.g_st.download(12 * ss)
        u = np.stack([st[k * ss:k * ss + self.n] for k in range(3)], 1)
        v = np.stack([st[(3 + k) * ss:(3 + k) * ss + self.n] for k in range(3)], 1)
        du = np.zeros_like(u)
        act = self.alt >= 0
        du[act] = v[act] * h
        if self.prescribed is not None and len(self.anchored_ids):
            tgt = self.prescribed(t_new)
            du[self.anchored_ids] = tgt - u[self.anchored_ids]
        for k in range(3):
            st[(9 + k) * ss:(9 + k) * ss + self.n] = du[:, k]
        self.g_st.upload(st)

    def _check_damage(self, damaged, failed_now):
        """_checkDamage (ImplicitModel3.cs:198-211): this attempt's new damage or failures above
        maxDamagePerStep / maxFailPerStep of the initial CZ count."""
        if self.tsf == PARTS or self.nCZ_initial == 0:
            return False
        dn_damaged = (damaged - self.prev_damaged) / self.nCZ_initial
        dn_failed = failed_now / self.nCZ_initial
        return dn_damaged > self.p.max_damage or dn_failed > self.p.max_fail

    def step(self):
        """One attempt of ImplicitModel3.Step; its record says whether it was accepted."""
        p, lib, ctx = self.p, self.lib, self.ctx
        if p.adapt:
            ticks = PARTS // self.tsf
            h = p.h * ticks / PARTS
            t_new = p.h * (self.int_time + ticks) / PARTS
        else:
            h, t_new = p.h, self.t + p.h
        self._begin_step(h, t_new)
        it, err0, converged, diverges, explodes = 0, None, False, False, False
        norms = []
        dmg = ctypes.c_int()
        fld = ctypes.c_int()
        while True:
            _lib.check(lib.icf_nodes_infer(ctx, self.g_st.p, self.st_stride, self.g_dn.p, self.ns, self.n, h, p.beta,
                                           p.gamma), ctx, "icf_nodes_infer")
            if self.contact:
                self._detect_contacts()
                self._upload_history()
                self._pattern_with_impacts()
            elif self.C:
                # TransferPCSR, each iteration: current history in, tentative fields and the CZ
                # flags cleared (GPU_Functionality.cs:376-412)
                self._upload_history()
                _lib.check(lib.icf_pack_pcsr(ctx, self.g_alt.p, self.n, self.g_rows.p, self.g_cols.p, self.N,
                                             None, 0, 0, self.g_icz.p, self.C, self.cs, None, 0, 0), ctx, "pack")
            _lib.check(lib.icf_assemble_elems_and_czs(
                ctx, self.g_ie.p, self.T, self.es, self.g_dcz.p if self.C else None,
                self.g_icz.p if self.C else None, self.C, self.cs, self.g_dn.p, self.ns, h, self.g_vals.p, self.nnz,
                self.g_rhs.p, self.N, ctypes.byref(dmg), ctypes.byref(fld)), ctx, "assemble")
            if self.contact and self.nImpacts:
                _lib.check(lib.icf_collision_response(ctx, self.g_icr.p, self.g_dn.p, h, self.g_vals.p, self.g_rhs.p,
                                                      self.nImpacts, self.cr_stride, self.ns, self.penalty,
                                                      self.dist_eps), ctx, "icf_collision_response")
            if p.adapt:
                explodes = self._check_damage(dmg.value, fld.value)
            vals = self.g_vals.download(9 * self.nnz)
            rhs = self.g_rhs.download(3 * self.N)
            if self.perturb:
                vals = vals * (1.0 + self.perturb * self._rng.standard_normal(vals.size))
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
            if not (it < p.min_iter or (not explodes and not diverges and not converged and it < p.max_iter)):
                break
        accepted = True
        if p.adapt:
            accepted, self.tsf, self.hold = adjust_time_step(self.tsf, self.hold, explodes, diverges, converged,
                                                             p.max_iter)
            if not accepted:
                rec = StepRecord(it, converged, 0, dmg.value, norms, accepted=False, h=h, explodes=explodes)
                self.records.append(rec)
                return rec
            self.int_time += ticks
            self.frames += 1
            if self.frames == 1:
                self.tsf = 1                            # _acceptFrame (:322)
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
        self.t = t_new
        self.prev_damaged = dmg.value
        rec = StepRecord(it, converged, failed_now, dmg.value, norms, accepted=True, h=h, explodes=explodes)
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


class BatchedImplicitModel(ImplicitModel):
    """Independent specimens of a batched collection (specimens.batch) advanced together with
    one fixed time step: each Newton iteration is one icf_assemble_batched over every specimen,
    each specimen's diagonal block is solved on the host (BatchedSolver), and each specimen
    iterates to its own convergence with the reference's criterion and iteration limits
    (ImplicitModel3.Step's loop, ImplicitModel3.cs:369-435, run per specimen); a specimen that
    has finished keeps its state (dx = 0) while the others iterate."""

    def __init__(self, col, consts, precision=_lib.ICF_FP64, step=None, prescribed=None, prescribed_nodes=None,
                 threads=None):
        from . import specimens
        self.nspec = len(col.meshes)
        tet, cz, row = specimens.specimen_offsets(col)
        self.tet_off, self.row_off = tet, row
        self.cz_spec = col.node_mesh[np.asarray(col.czs)[:, 0]] if len(col.czs) else np.zeros(0, np.int64)
        nm = np.asarray(col.node_mesh)
        if np.any(np.diff(nm) < 0):
            raise ValueError("nodes are not ordered by specimen")
        self.node_off = np.r_[0, np.cumsum(np.bincount(nm, minlength=self.nspec))]
        self.threads = threads
        self.timing = dict(assemble_ms=0.0, transfer_s=0.0, solve_s=0.0, iterations=0)
        super().__init__(col, consts, precision=precision, step=step or StepParams(adapt=False), prescribed=prescribed,
                         prescribed_nodes=prescribed_nodes)
        if self.p.adapt:
            raise ValueError("batched specimens share one fixed time step (adapt=False)")

    def _setup_czs(self):
        super()._setup_czs()
        cz = np.bincount(self.cz_spec[self.cz_alive], minlength=self.nspec) if self.C else np.zeros(self.nspec, int)
        self.cz_off = np.r_[0, np.cumsum(cz)].astype(np.int32)
        self.nnzb_off = self.rows[self.row_off].astype(np.int32)
        self.wl_arrays = [np.ascontiguousarray(a, np.int32) for a in (self.tet_off, self.cz_off, self.row_off,
                                                                        self.nnzb_off)]
        self.worklist = _lib.IcfWorklist(self.nspec, *[_lib.ptr(a) for a in self.wl_arrays])
        if getattr(self, "bsolver", None) is not None:
            self.bsolver.close()
        self.bsolver = BatchedSolver(self.rows, self.cols, self.row_off, self.threads)

    def step(self):
        """One fixed step of every specimen; the record carries the largest iteration count, and
        per-specimen iterations and convergence are kept in self.spec_records."""
        p, lib, ctx = self.p, self.lib, self.ctx
        h, t_new = p.h, self.t + p.h
        self._begin_step(h, t_new)
        S = self.nspec
        it = np.zeros(S, np.int64)
        err0 = np.zeros(S)
        conv = np.zeros(S, bool)
        div = np.zeros(S, bool)
        done = np.zeros(S, bool)
        dmg = (ctypes.c_int * S)()
        fld = (ctypes.c_int * S)()
        held = np.zeros(S, bool)
        dmg_last = np.zeros(S, np.int64)
        spec_of_row = np.repeat(np.arange(S), np.diff(self.row_off))
        while True:
            _lib.check(lib.icf_nodes_infer(ctx, self.g_st.p, self.st_stride, self.g_dn.p, self.ns, self.n, h, p.beta,
                                           p.gamma), ctx, "icf_nodes_infer")
            if self.C:
                self._upload_history()
                _lib.check(lib.icf_pack_pcsr(ctx, self.g_alt.p, self.n, self.g_rows.p, self.g_cols.p, self.N,
                                             None, 0, 0, self.g_icz.p, self.C, self.cs, None, 0, 0), ctx, "pack")
            t0 = time.perf_counter()
            _lib.check(lib.icf_assemble_batched(
                ctx, ctypes.byref(self.worklist), self.g_ie.p, self.es, self.g_dcz.p if self.C else None,
                self.g_icz.p if self.C else None, self.cs, self.g_dn.p, self.ns, h, self.g_vals.p, self.g_rhs.p,
                dmg, fld), ctx, "icf_assemble_batched")
            t1 = time.perf_counter()
            vals = self.g_vals.download(9 * self.nnz)
            rhs = self.g_rhs.download(3 * self.N)
            t2 = time.perf_counter()
            active = ~done
            dx = self.bsolver.solve(vals, rhs, active)
            sq = np.bincount(spec_of_row, weights=(dx * dx).reshape(-1, 3).sum(1), minlength=S)
            for s in np.flatnonzero(active):        # _checkDivergence per specimen (ImplicitModel3.cs:214-245)
                norm = sq[s]
                if it[s] == 0:
                    err0[s] = norm if norm != 0 else p.conv_eps
                elif norm < p.conv_cutoff:
                    conv[s] = True
                else:
                    rel = np.sqrt(norm / err0[s])
                    conv[s] = rel <= p.conv_eps
                    div[s] = rel > 1.01
            self.g_dx.upload(np.ascontiguousarray(dx))
            _lib.check(lib.icf_nodes_add_dx(ctx, self.g_st.p, self.st_stride, self.g_alt.p, self.n, self.g_dx.p),
                       ctx, "icf_nodes_add_dx")
            self.timing["assemble_ms"] += 1e3 * (t1 - t0)
            self.timing["transfer_s"] += t2 - t1
            self.timing["iterations"] += 1
            it[active] += 1
            done |= active & ~((it < p.min_iter) | (~div & ~conv & (it < p.max_iter)))
            newly = np.flatnonzero(active & done)
            if done.all():
                break
            # a specimen that has finished keeps the CZ state of its last assembly (the one its
            # own loop would have accepted), not that of later assemblies at its final iterate
            self._held_fields(newly, save=True)
            held[newly] = True
            dmg_last[newly] = np.asarray(dmg)[newly]
        self._held_fields(np.flatnonzero(held), save=False)
        dmg_last[~held] = np.asarray(dmg)[~held]
        self.timing["solve_s"] = self.bsolver.solve_s
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
        self.t = t_new
        self.spec_records = getattr(self, "spec_records", [])
        self.spec_records.append((it.copy(), conv.copy()))
        rec = StepRecord(int(it.max()), bool(conv.all()), failed_now, int(dmg_last.sum()), [], accepted=True, h=h)
        self.records.append(rec)
        return rec

    # what the accept reads of a finished specimen: the tentative displacement un as its last
    # inference left it (dn 3-5; Node.AcceptTentativeValues takes un as last inferred, before the
    # loop's final dx), and the CZ fields of its last assembly: tentative pmax, tmax and averages
    # (dcz 6-15), tentative contact, damaged and failed flags (icz 1-3)
    _DN_FIELDS, _DCZ_FIELDS, _ICZ_FIELDS = range(3, 6), range(6, 16), range(1, 4)

    def _held_fields(self, specs, save):
        """Copy those fields of the specimens `specs` into (save) or back from (restore) a
        device copy, so that later iterations of the other specimens do not change them."""
        if len(specs) == 0:
            return
        gf, lib, ctx, cs, ns = self.gf, self.lib, self.ctx, self.cs, self.ns
        groups = [(self._DN_FIELDS, self.g_dn, gf.buf("held_dn", 6 * ns, np.float64), 8, ns, self.node_off)]
        if self.C:
            groups += [(self._DCZ_FIELDS, self.g_dcz, gf.buf("held_dcz", 16 * cs, np.float64), 8, cs, self.cz_off),
                       (self._ICZ_FIELDS, self.g_icz, gf.buf("held_icz", 4 * cs, np.int32), 4, cs, self.cz_off)]
        for fields, live, keep, w, stride, offs in groups:
            for s in specs:
                u0, u1 = int(offs[s]), int(offs[s + 1])
                if u1 == u0:
                    continue
                for f in fields:
                    off = w * (f * stride + u0)
                    src, dst = (live, keep) if save else (keep, live)
                    _lib.check(lib.icf_d2d(ctx, dst.p, off, src.p, off, w * (u1 - u0)), ctx, "icf_d2d")

    def close(self):
        if getattr(self, "bsolver", None) is not None:
            self.bsolver.close()
        super().close()


def tension(col, strain=3e-4, ramp_steps=10, h=1e-4, size=0.02):
    """Specimens of edge `size` standing on z = 0, each pulled in z: the nodes of its bottom face
    fixed, those of its top face moved up by strain * size, ramped linearly over ramp_steps
    steps and then held.  Finalizes `col`; returns it, the prescribed displacement of the
    anchored nodes at time t, and those nodes."""
    if getattr(col, "x0", None) is None:
        col.finalize()
    z = col.x0[:, 2]
    bottom, top = np.isclose(z, 0.0), np.isclose(z, size)
    col.anchored = bottom | top
    col.finalize()
    ids = np.flatnonzero(col.anchored)
    is_top = np.isclose(col.x0[ids, 2], size)
    T_ramp = ramp_steps * h

    def prescribed(t):
        out = np.zeros((len(ids), 3))
        out[is_top, 2] = strain * size * min(1.0, t / T_ramp)
        return out

    return col, prescribed, ids


def batched_tension(consts, n_specimens=256, seed=2, strain=3e-4, ramp_steps=10, h=1e-4, size=0.02, **cubes):
    """tension of the batched specimens (specimens.lognormal_cubes, seed 2 for the benchmark's
    256)."""
    from . import specimens
    return tension(specimens.lognormal_cubes(n_specimens, seed, size=size, **cubes), strain, ramp_steps, h, size)


def batched_precision_study(consts, n_specimens=256, steps=4, seed=2, strain=3e-4, ramp_steps=10, h=1e-4,
                            threads=None, log=None):
    """batched_tension in FP64 and in mixed precision for `steps` fixed steps: the relative
    displacement error of the mixed run (over all nodes, and the largest over specimens), the
    failed-CZ sets, Newton iterations, and where the time of an iteration goes (the batched
    assembly, the values' and residual's transfer to the host, the host solves)."""
    out = {}
    runs = {}
    for name, mode in (("fp64", _lib.ICF_FP64), ("mixed", _lib.ICF_MIXED)):
        col, presc, ids = batched_tension(consts, n_specimens, seed, strain, ramp_steps, h)
        t0 = time.time()
        m = BatchedImplicitModel(col, consts, precision=mode, step=StepParams(h=h), prescribed=presc,
                                 prescribed_nodes=ids, threads=threads)
        for k in range(steps):
            rec = m.step()
            if log:
                log(f"batched newton {name}: step {k + 1}, {rec.iterations} iterations at most, "
                    f"{len(m.failed_set())} CZs failed, {m.timing}")
        runs[name] = (m.displacements(), m.failed_set(), [r.iterations for r in m.records],
                      int(sum(int(i.sum()) for i, _ in m.spec_records)), col.node_mesh.copy())
        tm = dict(m.timing)
        it = max(1, tm["iterations"])
        out[f"timing_{name}"] = dict(iterations=tm["iterations"], wall_s=round(time.time() - t0, 1),
                                     assemble_ms_per_iter=round(tm["assemble_ms"] / it, 3),
                                     transfer_ms_per_iter=round(1e3 * tm["transfer_s"] / it, 1),
                                     solve_ms_per_iter=round(1e3 * tm["solve_s"] / it, 1),
                                     solve_threads=m.bsolver.threads)
        out.update(specimens=n_specimens, tets=int(m.T), czs=int(len(m.cz_nodes)), rows=int(m.N), nnzb=int(m.nnz))
        m.close()
    u64, f64, it64, sit64, nm = runs["fp64"]
    umx, fmx, itmx, sitmx, _ = runs["mixed"]
    scale = float(np.linalg.norm(u64))
    per = []
    for s in range(n_specimens):
        sel = nm == s
        d = float(np.linalg.norm(u64[sel]))
        per.append(float(np.linalg.norm(umx[sel] - u64[sel]) / d) if d > 0 else 0.0)
    out.update(steps=steps, dt=h, strain=strain, ramp_steps=ramp_steps,
               rel_err_displacement=float(np.linalg.norm(umx - u64) / scale) if scale > 0 else 0.0,
               max_specimen_rel_err=float(max(per)), max_abs_displacement=float(np.abs(u64).max()),
               failed_czs_fp64=int(len(f64)), failed_czs_mixed=int(len(fmx)),
               failed_set_identical=bool(np.array_equal(f64, fmx)),
               newton_iterations_fp64=sit64, newton_iterations_mixed=sitmx)
    return out


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


def platen_compression(consts, cells=(8, 8, 8), size=(0.02, 0.02, 0.02), grains=8, seed=0, rate=5.0):
    """A cube between two rigid platens of its own footprint, one cell thick, meshed with 4/5 of
    the cube's cells per side (as 32 for 40) so that their vertices do not fall on the cube's
    tet faces: the bottom platen fixed, the top one moving down at rate * size_z per second,
    the specimen free (it meets both through point-triangle contact, and its grains meet each
    other).  Returns the collection, the prescribed displacement of the platen nodes at time t,
    those nodes, and the penalty stiffness per impact (1e11 N/m^3 times the platen cell
    area)."""
    from . import specimens
    pc = (max(1, cells[0] * 4 // 5), max(1, cells[1] * 4 // 5))
    col = specimens.make_specimen(cells, size, grains, seed, platens=pc).finalize()
    nodes = np.flatnonzero(col.node_mesh > 0)
    top = col.node_mesh[nodes] == 2
    v = rate * size[2]

    def prescribed(t):
        out = np.zeros((len(nodes), 3))
        out[top, 2] = -v * t
        return out

    return col, prescribed, nodes, 1e11 * (size[0] / pc[0]) * (size[1] / pc[1])


def contact_precision_study(consts, steps=20, h=1e-4, adapt=True, max_attempts=400, log=None, **scenario):
    """platen_compression in FP64 and in mixed precision up to time steps * h: relative
    displacement error at the end, the failed-CZ sets, impacts and Newton iterations."""
    out = {}
    runs = {}
    for name, mode in (("fp64", _lib.ICF_FP64), ("mixed", _lib.ICF_MIXED)):
        col, presc, nodes, k = platen_compression(consts, **scenario)
        m = ImplicitModel(col, consts, precision=mode, step=StepParams(h=h, adapt=adapt), prescribed=presc,
                          prescribed_nodes=nodes, contact=True, penalty=k)
        attempts, frames, impacts = 0, 0, []
        while attempts < max_attempts and (m.int_time < steps * PARTS if adapt else frames < steps):
            rec = m.step()
            attempts += 1
            if rec.accepted:
                frames += 1
                impacts.append(m.nImpacts)
                if log:
                    log(f"contact newton {name}: frame {frames} t={m.t:.6g} impacts {m.nImpacts}, "
                        f"{rec.iterations} iterations, {len(m.failed_set())} CZs failed")
        runs[name] = (m.displacements(), m.failed_set(), m.int_time if adapt else frames, m.t,
                      sum(r.iterations for r in m.records if r.accepted))
        out[f"attempts_{name}"] = attempts
        out[f"impacts_last_{name}"] = impacts[-1] if impacts else 0
        out[f"impacts_max_{name}"] = max(impacts) if impacts else 0
        out["nodes"] = int(m.n)
        out["tets"] = int(m.T)
        out["czs"] = int(len(m.cz_nodes))
        m.close()
    u64, f64, k64, t64, it64 = runs["fp64"]
    umx, fmx, kmx, tmx, itmx = runs["mixed"]
    spec = np.flatnonzero(col.node_mesh == 0)
    scale = float(np.linalg.norm(u64[spec]))
    out.update(steps=steps, dt=h, adapt=adapt, t_end_fp64=t64, t_end_mixed=tmx, same_end_time=bool(k64 == kmx),
               rel_err_displacement=float(np.linalg.norm(umx[spec] - u64[spec]) / scale) if scale > 0 else 0.0,
               failed_czs_fp64=int(len(f64)), failed_czs_mixed=int(len(fmx)),
               failed_set_identical=bool(np.array_equal(f64, fmx)),
               newton_iterations_fp64=int(it64), newton_iterations_mixed=int(itmx))
    return out


def bend_run(consts, mode, steps=20, cells=(40, 8, 8), grains=16, seed=0, h=1e-4, ramp_steps=200, adapt=False,
             max_attempts=4000, conv_eps=0.005, max_iter=10, conv_cutoff=1e-8, log=None, name=None, history=False,
             perturb=0.0, contact=False):
    """One three-point-bend run up to time steps * h in precision `mode` with the Newton
    tolerance `conv_eps` (relative) and `conv_cutoff` (absolute, on the squared norm of dx) and
    at most `max_iter` iterations per step.  Returns a dict with the
    displacements `u`, the failed CZs, the per-frame iterations, convergence and newly failed
    CZs, the end time and the attempts; with `history` also each frame's time and
    displacements (`frame_t`, `frame_u`); `perturb` as ImplicitModel's.  With `contact`, every
    exterior and CZ face of the beam is in the contact surface (self-contact, penalty 1e11 N/m^3
    times the square of the cell edge along the beam)."""
    col, presc = three_point_bend(consts, cells=cells, grains=grains, seed=seed, ramp_steps=ramp_steps, h=h)
    kw = dict(contact=True, penalty=1e11 * (0.40 / cells[0]) ** 2) if contact else {}   # the 0.4 m beam's cell
    m = ImplicitModel(col, consts, precision=mode, prescribed=presc, perturb=perturb, **kw,
                      step=StepParams(h=h, adapt=adapt, conv_eps=conv_eps, max_iter=max_iter, conv_cutoff=conv_cutoff))
    attempts, frames, max_tsf = 0, 0, 1
    frame_t, frame_u, impacts = [], [], []
    name = name or ("fp64" if mode == _lib.ICF_FP64 else "mixed")
    while attempts < max_attempts and (m.int_time < steps * PARTS if adapt else frames < steps):
        max_tsf = max(max_tsf, m.tsf)
        rec = m.step()
        attempts += 1
        frames += rec.accepted
        if history and rec.accepted:
            frame_t.append(m.t)
            frame_u.append(m.displacements())
        if contact and rec.accepted:
            impacts.append(m.nImpacts)
        if log and rec.accepted and (frames % max(1, steps // 5) == 0 or rec.failed_czs):
            log(f"newton {name}: frame {frames} t={m.t:.6g} (attempt {attempts}, tsf {max_tsf}), "
                f"{rec.iterations} iterations, {len(m.failed_set())} CZs failed")
    acc = [r for r in m.records if r.accepted]
    out = dict(u=m.displacements(), failed=m.failed_set(), iterations=[r.iterations for r in acc],
               converged=[r.converged for r in acc], failed_now=[r.failed_czs for r in acc],
               clock=m.int_time if adapt else frames, t=m.t, attempts=attempts, frames=frames,
               finest_step=h / max_tsf, nodes=int(m.n), czs=int(len(m.cz_nodes)), impacts=impacts)
    if history:
        out.update(frame_t=frame_t, frame_u=frame_u)
    m.close()
    return out


def compare_runs(a, b):
    """Relative displacement difference of run b from run a (normwise over all nodes), and
    whether the failed-CZ sets and the end times agree."""
    scale = float(np.linalg.norm(a["u"]))
    return dict(rel_err_displacement=float(np.linalg.norm(b["u"] - a["u"]) / scale) if scale > 0 else 0.0,
                same_end_time=bool(a["clock"] == b["clock"]),
                failed_set_identical=bool(np.array_equal(a["failed"], b["failed"])))


def precision_study(consts, steps=20, cells=(40, 8, 8), grains=16, seed=0, h=1e-4, ramp_steps=200, adapt=False,
                    max_attempts=4000, conv_eps=0.005, max_iter=10, conv_cutoff=1e-8, contact=False, log=None):
    """The same three-point-bend run in FP64 and in mixed precision up to time steps * h:
    relative displacement error of the mixed run at the end (normwise over all nodes), the
    failed-CZ sets, Newton iterations, and with adapt (the reference's time-step control) the
    attempts, discarded attempts and the finest step used."""
    out = {}
    runs = {}
    for name, mode in (("fp64", _lib.ICF_FP64), ("mixed", _lib.ICF_MIXED)):
        r = runs[name] = bend_run(consts, mode, steps=steps, cells=cells, grains=grains, seed=seed, h=h,
                                  ramp_steps=ramp_steps, adapt=adapt, max_attempts=max_attempts, conv_eps=conv_eps,
                                  max_iter=max_iter, conv_cutoff=conv_cutoff, log=log, name=name, contact=contact)
        out[f"attempts_{name}"] = r["attempts"]
        out[f"discarded_{name}"] = r["attempts"] - r["frames"]
        out[f"finest_step_{name}"] = r["finest_step"]
        out["nodes"] = r["nodes"]
        out["czs"] = r["czs"]
        if contact:
            out[f"impacts_max_{name}"] = int(max(r["impacts"], default=0))
            out[f"impacts_last_{name}"] = int(r["impacts"][-1]) if r["impacts"] else 0
    r64, rmx = runs["fp64"], runs["mixed"]
    out.update(steps=steps, dt=h, adapt=adapt, contact=contact, ramp_steps=ramp_steps, conv_eps=conv_eps,
               conv_cutoff=conv_cutoff, max_iter=max_iter, t_end_fp64=r64["t"],
               t_end_mixed=rmx["t"], **compare_runs(r64, rmx),
               max_abs_displacement=float(np.abs(r64["u"]).max()),
               failed_czs_fp64=int(len(r64["failed"])), failed_czs_mixed=int(len(rmx["failed"])),
               newton_iterations_fp64=int(sum(r64["iterations"])), newton_iterations_mixed=int(sum(rmx["iterations"])),
               steps_converged_fp64=int(sum(r64["converged"])), steps_converged_mixed=int(sum(rmx["converged"])))
    return out


def main(argv=None):
    """python -m icflow_amd.newton [--steps N] [--ramp R] [--adapt] [--self-contact] [--batched S]:
    the FP64-vs-mixed precision study as one JSON line (GPU required).  Examples: the 40x8x8
    beam for 200 steps, 2 mm over 200 steps, with time-step control: --steps 200 --ramp 200
    --adapt; the same beam with self-contact for 300 steps: --steps 300 --ramp 200 --adapt
    --self-contact; the 256 batched specimens, 5 fixed steps to 2e-3 strain over 4:
    --batched 256 --steps 5 --ramp 4 --strain 2e-3."""
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=10, help="simulated time in units of the initial step h")
    ap.add_argument("--ramp", type=int, default=200, help="steps of the deflection ramp (2 mm)")
    ap.add_argument("--adapt", action="store_true", help="the reference's time-step control")
    ap.add_argument("--max-attempts", type=int, default=4000)
    ap.add_argument("--conv-eps", type=float, default=0.005, help="Newton tolerance (ConvergenceEpsilon)")
    ap.add_argument("--max-iter", type=int, default=10, help="Newton iterations per step at most")
    ap.add_argument("--conv-cutoff", type=float, default=1e-8, help="absolute |dx|^2 cutoff (ConvergenceCutoff)")
    ap.add_argument("--contact", action="store_true",
                    help="platen compression with contact (platen_compression) instead of the three-point bend")
    ap.add_argument("--self-contact", action="store_true",
                    help="the three-point bend with self-contact (exterior and CZ faces)")
    ap.add_argument("--batched", type=int, default=0, help="batched specimens in tension instead (count)")
    ap.add_argument("--strain", type=float, default=3e-4, help="end strain of the batched tension")
    a = ap.parse_args(argv)
    consts = _lib.consts_from_prms(**_lib.SPECIMEN_PRMS)
    if a.batched:
        r = batched_precision_study(consts, n_specimens=a.batched, steps=a.steps, strain=a.strain,
                                    ramp_steps=a.ramp, log=lambda m: print(m, flush=True))
    elif a.contact:
        r = contact_precision_study(consts, steps=a.steps, adapt=a.adapt, max_attempts=a.max_attempts,
                                    log=lambda m: print(m, flush=True))
    else:
        r = precision_study(consts, steps=a.steps, ramp_steps=a.ramp, adapt=a.adapt, max_attempts=a.max_attempts,
                            conv_eps=a.conv_eps, max_iter=a.max_iter, conv_cutoff=a.conv_cutoff,
                            contact=a.self_contact, log=lambda m: print(m, flush=True))
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()

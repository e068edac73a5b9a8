#!/usr/bin/env python3
"""Newton-iteration assembly throughput of libicflow_hip.so on one MI355X.

Headline workload (default, --workload batched): 256 independent cubes with lognormal tet
counts (median 12k, truncated to 2k-40k, log-sd 0.6, seed 2), 8-64 Voronoi grains each, PPR
cohesive zones on every inter-grain face, no contact: 3.5M tets and 0.8M CZs assembled per
Newton iteration by icf_assemble_batched, one launch per phase over all specimens.

Further workloads (--extra, reported under "extra_workloads"; --workload selects one of them
as the headline instead):
  compression: a 0.1 m cube as a 40x40x40 hex grid split into 6 tets per hex (384,000 tets),
    500 Voronoi grains (seed 1), two rigid one-cell-thick tet platens (32x32 cells) pressed 0.3
    cell into the cube; contact: narrow phase over the broad-phase candidate pairs, closest
    face, the fused assembly and the contact response.  The in-place inputs of the two
    narrow-phase kernels are restored by device-to-device copies before each step.
  large: the same with a 0.2 m cube, 60^3 hex x 6 tets (1,296,000 tets), 2,000 grains
    (seed 3), 48x48 platens.
Material E=9 GPa, nu=0.3, rho=916; PPR sigma_max=1 MPa, tau_max=2 MPa, phi=10 J/m^2,
alpha=beta=4, lambda=0.02.  Synthetic seeded states.

One step = one Newton iteration's device work with every input already in HBM.
Throughput = (tets + active CZs + impacts) / step time.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# VALU issue peak of MI355X in wave64 instructions per second: 256 CUs x 4 SIMDs, one wave64
# FMA (FP32 unpacked or FP64) per SIMD every 4 cycles at the 2.4 GHz peak engine clock
SIMDS, CLOCK_HZ = 1024, 2.4e9
VALU_ISSUE_PEAK = SIMDS * CLOCK_HZ / 4


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_batched(n_specimens, seed):
    """Batched specimens: n cubes with lognormal tet counts (median 12k, truncated to 2k-40k,
    log-sd 0.6), 8-64 Voronoi grains each, laid side by side; no contact."""
    from icflow_amd import _lib, specimens
    consts = _lib.consts_from_prms(**_lib.SPECIMEN_PRMS)
    t = time.time()
    col = specimens.lognormal_cubes(n_specimens, seed)
    st = specimens.synthetic_state(col, consts, seed, equal_gp=True)
    log(f"batched specimens: {n_specimens} cubes, {len(col.x0)} nodes, {len(col.elastic)} tets, {len(col.czs)} CZs "
        f"({time.time() - t:.1f} s)")
    return consts, col, st


def build_workload(cells, grains, seed, platen_frac, size=0.1, platen_cells=32):
    from icflow_amd import _lib, specimens
    consts = _lib.consts_from_prms(**_lib.SPECIMEN_PRMS)
    t = time.time()
    col = specimens.make_specimen((cells, cells, cells), (size, size, size), grains, seed,
                                  platens=(platen_cells, platen_cells)).finalize()
    hz = size / cells
    st = specimens.synthetic_state(col, consts, seed, platen_dz=platen_frac * hz, equal_gp=True)
    log(f"specimen: {len(col.x0)} nodes, {len(col.elastic)} tets, {len(col.czs)} CZs, "
        f"{len(col.surf)} surface tets ({time.time() - t:.1f} s)")
    return consts, col, st


class Runner:
    """Device state for the timed step, built once through the GPU_Functionality mirror."""

    def __init__(self, consts, col, st, mode, contact=True):
        from icflow_amd import _lib
        from icflow_amd.gpu_functionality import GPUFunctionality, broad_phase, stride_of
        self.lib = _lib.load()
        self.check = _lib.check
        self.contact = contact
        self.col = col
        self.h = 5e-5
        gf = self.gf = GPUFunctionality(0)
        gf.SetConstants(consts)
        gf.set_precision(mode)
        gf.TransferStaticDataToDevice(col)
        gf.TransferNodesToGPU(st)
        if not contact:
            self.pairs, self.nT, self.Q = np.zeros((0, 2), np.int32), 0, 0
            self.contact_touched = {"vals": 0, "rhs": 0, "force": 0}
            t = time.time()
            gf.CreateStructure(None)
            gf.TransferPCSR(st)
            log(f"pattern: N={gf.csr.N} nnzb={gf.csr.nnz} ({time.time() - t:.1f} s)")
            self.vals = gf.buf("vals", 2 * gf.csr.nnz * 9, np.float64)
            self.rhs = gf.buf("rhs", gf.csr.N * 3, np.float64)
            # specimens of a batched collection: per-specimen offsets for icf_assemble_batched
            self.worklist = None
            if len(col.meshes) > 1:
                from icflow_amd import specimens
                tet, cz, row = specimens.specimen_offsets(col)
                offs = [tet, cz, row, gf.csr.rows[row].astype(np.int32)]
                self.worklist_arrays = offs
                self.worklist = _lib.IcfWorklist(len(tet) - 1, *[_lib.ptr(a) for a in offs])
            return
        t = time.time()
        self.pairs = broad_phase(gf, gf.c_dn, gf.nd_stride, gf.c_ie, gf.el_all_stride, len(col.surf))
        log(f"broad phase: {len(self.pairs)} candidate pairs ({time.time() - t:.1f} s)")
        impacts = gf.NarrowPhaseCollisionDetection(self.pairs)
        self.nT = gf.nImpacts
        log(f"narrow phase: {int(np.count_nonzero(gf.masks))} pairs with containment, {self.nT} impacts")
        t = time.time()
        gf.CreateStructure(impacts)
        gf.TransferPCSR(st)
        log(f"pattern: N={gf.csr.N} nnzb={gf.csr.nnz} ({time.time() - t:.1f} s)")
        # pristine copies of the in-place inputs of the narrow-phase kernels
        Q = len(self.pairs)
        self.Q = Q
        self.itet = gf.buf("itet", 2 * Q, np.int32)
        self.itet_src = gf.buf("itet_src", 2 * Q, np.int32)
        self.itet_src.upload(self.pairs.ravel())
        ts = gf.tet_stride
        self.nl2 = gf.buf("itet2", 4 * ts, np.int32)
        self.nl2_src = gf.buf("itet2_src", 2 * ts, np.int32)
        # rows 0-1 of the closest-face buffer are the (node, tet) tuples of NL2set
        tup = np.zeros(2 * ts, np.int32)
        masks = np.zeros(2 * Q, np.int32)
        masks[0::2] = gf.masks
        n = ctypes.c_int()
        self.check(self.lib.icfh_nl2set(_ptr(masks), _ptr(self.pairs.astype(np.int32).ravel().copy()), Q,
                                        _ptr(gf.c_ie), gf.el_all_stride, _ptr(tup), ts, ts, ctypes.byref(n)))
        self.nl2_src.upload(tup)
        # impacts: icr packed once (same impacts every step)
        gf.c_icr, gf.cr_stride = gf.pack_icr(impacts)
        self.icr = gf.buf("icr", len(gf.c_icr), np.int32)
        self.icr.upload(gf.c_icr)
        cs, nT = gf.cr_stride, self.nT
        fld = gf.c_icr.reshape(-1, cs)[:, :nT]
        self.contact_touched = {
            "vals": int(np.unique(fld[4:20][fld[4:20] >= 0]).size),
            "rhs": int(np.unique(fld[20:24][fld[20:24] >= 0]).size),
            "force": int(np.unique(fld[0:4]).size)}
        nnz, N = gf.csr.nnz, gf.csr.N
        self.vals = gf.buf("vals", 2 * nnz * 9, np.float64)
        self.rhs = gf.buf("rhs", N * 3, np.float64)
        self.col = col
        self.h = 5e-5
        self.k = 1e11 * (0.1 / 32) ** 2
        self.eps = 1e-15

    def step(self):
        """One iteration's device calls.  Blocking by default (the drop-in sequence); with
        icf_set_blocking(0) they only enqueue, and between icf_graph_begin/end they are captured."""
        gf, lib, ctx, chk = self.gf, self.lib, self.gf.ctx, self.check
        if not self.contact and self.worklist is not None:
            chk(lib.icf_assemble_batched(
                ctx, ctypes.byref(self.worklist), gf.g_ie_pcsr.p, gf.el_elastic_stride, gf.g_dcz.p, gf.g_icz.p,
                gf.cz_stride, gf.g_dn.p, gf.nd_stride, self.h, self.vals.p, self.rhs.p, None, None), ctx,
                "icf_assemble_batched")
            return
        if not self.contact:
            chk(lib.icf_assemble_elems_and_czs(
                ctx, gf.g_ie_pcsr.p, len(self.col.elastic), gf.el_elastic_stride, gf.g_dcz.p, gf.g_icz.p,
                len(self.col.czs), gf.cz_stride, gf.g_dn.p, gf.nd_stride, self.h, self.vals.p, gf.csr.nnz,
                self.rhs.p, gf.csr.N, None, None), ctx, "icf_assemble_elems_and_czs")
            return
        chk(lib.icf_d2d(ctx, self.itet.p, 0, self.itet_src.p, 0, 8 * self.Q), ctx, "icf_d2d")
        chk(lib.icf_d2d(ctx, self.nl2.p, 0, self.nl2_src.p, 0, 8 * gf.tet_stride), ctx, "icf_d2d")
        chk(lib.icf_narrow_phase(ctx, self.Q, gf.g_dn.p, gf.g_ie.p, self.itet.p, gf.el_all_stride, gf.nd_stride),
            ctx, "icf_narrow_phase")
        chk(lib.icf_find_closest_face(ctx, self.nT, gf.tet_stride, self.nl2.p, gf.g_dn.p, gf.g_ie.p, gf.g_ifc.p,
                                      gf.nd_stride, gf.el_all_stride, gf.fc_stride), ctx, "icf_find_closest_face")
        chk(lib.icf_assemble_elems_and_czs(
            ctx, gf.g_ie_pcsr.p, len(self.col.elastic), gf.el_elastic_stride, gf.g_dcz.p, gf.g_icz.p,
            len(self.col.czs), gf.cz_stride, gf.g_dn.p, gf.nd_stride, self.h, self.vals.p, gf.csr.nnz,
            self.rhs.p, gf.csr.N, None, None), ctx, "icf_assemble_elems_and_czs")
        chk(lib.icf_collision_response(ctx, self.icr.p, gf.g_dn.p, self.h, self.vals.p, self.rhs.p, self.nT,
                                       gf.cr_stride, gf.nd_stride, self.k, self.eps), ctx, "icf_collision_response")

    def transfers(self):
        """Host buffers for the per-iteration copies of the reference's sequence
        (ImplicitModel3.cs:381-397): up, the node buffer (TransferNodesToGPU), the elements' and
        CZs' pcsr arrays and the CZ history (TransferPCSR, GPU_Functionality.cs:345-416) and the
        impacts' rows (collisionResponse); down, the values and rhs
        (TransferLinearSystemToHost)."""
        gf = self.gf
        up = [(gf.g_dn, gf.c_dn), (gf.g_ie_pcsr, gf.c_ie_pcsr)]
        if gf.cz_stride:
            up += [(gf.g_icz, gf.c_icz), (gf.g_dcz, gf.c_dcz)]
        if self.nT:
            up.append((self.icr, gf.c_icr))
        self.up = [(d, np.ascontiguousarray(h)) for d, h in up]
        self.down = [(self.vals, np.empty(9 * gf.csr.nnz)), (self.rhs, np.empty(3 * gf.csr.N))]
        return sum(h.nbytes for _, h in self.up), sum(h.nbytes for _, h in self.down)

    def step_with_transfers(self):
        """One drop-in iteration with its host copies: uploads, the device calls, downloads."""
        lib, ctx, chk = self.lib, self.gf.ctx, self.check
        for d, h in self.up:
            chk(lib.icf_h2d(ctx, d.p, 0, _ptr(h), h.nbytes), ctx, "icf_h2d")
        self.step()
        for d, h in self.down:
            chk(lib.icf_d2h(ctx, _ptr(h), d.p, 0, h.nbytes), ctx, "icf_d2h")

    def failed_czs(self):
        d, f = ctypes.c_int(), ctypes.c_int()
        self.check(self.lib.icf_cz_counts(self.gf.ctx, ctypes.byref(d), ctypes.byref(f)), self.gf.ctx, "icf_cz_counts")
        return f.value

    def capture(self):
        """The whole iteration as one HIP graph (icf_graph_begin / icf_graph_end)."""
        lib, ctx = self.lib, self.gf.ctx
        self.check(lib.icf_graph_begin(ctx), ctx, "icf_graph_begin")
        self.step()
        g = ctypes.c_void_p()
        self.check(lib.icf_graph_end(ctx, ctypes.byref(g)), ctx, "icf_graph_end")
        return g

    def set_mode(self, mode):
        self.gf.set_precision(mode)

    def profile(self):
        lib, ctx = self.lib, self.gf.ctx
        out = {}
        buf = ctypes.create_string_buffer(128)
        for i in range(lib.icf_profile_count(ctx)):
            ms, n = ctypes.c_double(), ctypes.c_longlong()
            lib.icf_profile_entry(ctx, i, buf, 128, ctypes.byref(ms), ctypes.byref(n))
            out[buf.value.decode()] = (ms.value, n.value)
        return out


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def timed(runner, steps, warmup, how):
    """Seconds per step.  how = "dropin": blocking calls (each returns after its work, as the
    reference's Run()); "transfers": the same with the reference's per-iteration host copies
    around them (Runner.transfers); "graph": replays of one captured iteration graph; "async": the calls
    enqueued without waiting (icf_set_blocking(0)); "profile": the calls
    enqueued without waiting, each kernel bracketed by HIP events on the library's stream (the
    library runs the fused call on one stream while profiling, so each kernel is timed alone)."""
    lib, ctx = runner.lib, runner.gf.ctx
    g = None
    if how == "graph":
        lib.icf_set_blocking(ctx, 0)
        g = runner.capture()
        fn = lambda: lib.icf_graph_launch(ctx, g)      # noqa: E731
    elif how in ("profile", "async"):
        lib.icf_set_blocking(ctx, 0)
        fn = runner.step
    elif how == "transfers":
        lib.icf_set_blocking(ctx, 1)
        runner.transfers()
        fn = runner.step_with_transfers
    else:
        lib.icf_set_blocking(ctx, 1)
        fn = runner.step
    for _ in range(warmup):
        fn()
    lib.icf_synchronize(ctx)
    lib.icf_profile_reset(ctx)
    lib.icf_set_profiling(ctx, 1 if how == "profile" else 0)
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    lib.icf_synchronize(ctx)
    dt = time.perf_counter() - t0
    lib.icf_set_profiling(ctx, 0)
    lib.icf_set_blocking(ctx, 1)
    if g is not None:
        lib.icf_graph_destroy(g)
    log(f"  {how:8s}: {1e3 * dt / steps:.3f} ms/step over {steps} steps")
    return dt / steps, runner.profile()


def autotune(r, rounds, name):
    """Blocking iterations under icf_set_autotune(rounds) until the key limits of this mode's
    value reductions are chosen (set-up, before any timed step)."""
    gf = r.gf
    stage = "fp32" if name == "mixed" else "fp64"
    roles = [f"overwrite,{stage}"] + ([f"accumulate,{stage}"] if r.contact else [])
    gf.set_autotune(rounds)
    r.lib.icf_set_blocking(gf.ctx, 1)
    n = 0
    while n < 4 * rounds * len(gf.CANDIDATES) + 2:
        r.step()
        n += 1
        state = gf.autotune_state()
        if not any(state[k][1] for k in roles):
            break
    gf.set_autotune(0)
    state = gf.autotune_state()
    got = {k: dict(keys=state[k][0][0], stage=state[k][0][1],
                   best_us={f"{c[0]}/{c[1]}": round(1e3 * ms, 1) for c, ms in state[k][2].items()})
           for k in roles}
    shapes = gf.shape_state()
    got.update({f"shape,{k}": dict(value=v, best_us={str(c): round(1e3 * ms, 1) for c, ms in best.items()})
                for k, (v, _, best) in shapes.items() if any(ms > 0 for ms in best.values())})
    log(f"autotune {name} ({n} iterations): {got}")
    return got


def algorithmic_bytes(r):
    """Compulsory HBM bytes per launch of each kernel (every byte counted once; staging records
    and slot maps are implementation traffic and are not counted), keyed by the profile name
    without its template arguments; accumulate-mode launches (the contact reductions) key as
    "<name>+acc"."""
    col, gf = r.col, r.gf
    T, C, P, Q = len(col.elastic), len(col.czs), r.nT, r.Q
    nnzb, N = gf.csr.nnz, gf.csr.N
    n_def = int(np.unique(col.elastic).size)
    n_cz = int(np.unique(col.czs).size) if C else 0
    n_surf = int(np.unique(col.surf).size)
    adj = getattr(gf, "maxAdjacentFaces", 0)
    touched = r.contact_touched
    alg = {
        # ids 16 B/tet + x0,un,vn,an 96 B per node
        "k_tet_phase1": 16 * T + 96 * n_def,
        # failed flag + 6 ids + pmax/tmax in (76 B), tentative pmax/tmax/averages + 3 flags out (92 B);
        # x0, un per CZ node (48 B)
        "k_cz_phase1": 168 * C + 48 * n_cz,
        # BCSR values written once (72 B/block) + the slot fields the reference scatter reads
        "k_reduce_vals": 72 * nnzb + 64 * T + 144 * C,
        "k_reduce_vec3<rhs>": 24 * N + 16 * T + 24 * C,
        "k_reduce_vec3<force>": 48 * n_def + 16 * T,
        # 2 tet ids in, mask out (12 B/pair) + surface-tet node ids 16 B + x_current 24 B/node
        "k_narrow_phase": 12 * Q + 16 * len(col.surf) + 24 * n_surf,
        # tuple in/out 24 B + adjacency count and ids + face ids
        "k_closest_face": 24 * P + 4 * P * (1 + adj),
        # impact int record (96 B) + 4 nodes' x_current (96 B)
        "k_contact_impact": 96 * P + 96 * P,
        # read-modify-write of the touched blocks + the 16 slot fields per impact
        "k_reduce_vals+acc": 144 * touched["vals"] + 64 * P,
        "k_reduce_vec3<rhs+acc>": 48 * touched["rhs"] + 16 * P,
        "k_reduce_vec3<force+acc>": 48 * touched["force"] + 16 * P,
        # the two restores of a step (8Q and 8 tet_stride bytes), read + write, averaged
        "k_copy": 8 * Q + 8 * getattr(gf, "tet_stride", 0),
        "k_count_flags": 8 * C,
        "k_fill_int": 8,
    }
    # launches that run several of the phases above side by side
    alg["k_phase1"] = alg["k_tet_phase1"] + alg["k_cz_phase1"]
    alg["k_vals_force"] = alg["k_reduce_vals"] + alg["k_reduce_vec3<force>"]
    alg["k_rhs_counts"] = alg["k_reduce_vec3<rhs>"] + alg["k_count_flags"]
    alg["k_vals_force_rhs"] = alg["k_vals_force"] + alg["k_rhs_counts"]
    alg["k_contact_sums"] = alg["k_reduce_vals+acc"] + alg["k_reduce_vec3<rhs+acc>"] + alg["k_reduce_vec3<force+acc>"]
    return alg


def iteration_bytes(r):
    """Compulsory HBM bytes of one whole Newton iteration's assembly (FP64 reference buffers,
    every byte once): unit records, node fields in, node force out, BCSR values and rhs out,
    narrow-phase lists.  Divided by the step time this is the iteration's HBM roofline."""
    col, gf = r.col, r.gf
    T, C, P, Q = len(col.elastic), len(col.czs), r.nT, r.Q
    Nn = len(col.x0)
    return (96 * T + (196 + 48 + 80 + 12) * C + 96 * P + 144 * Nn + 72 * gf.csr.nnz + 24 * gf.csr.N
            + 12 * Q + 24 * P)


def kernel_bytes(alg, name):
    base = name.split("<")[0]
    if name.endswith(",acc>"):
        if base == "k_reduce_vec3":
            return alg.get(name.replace(",acc>", "+acc>"))
        return alg.get(base + "+acc")
    return alg.get(name, alg.get(base))


def _rocprof_symbol(name):
    """The kernel symbol rocprofv3 reports for a per-kernel profile name of the library
    (template arguments spelled out as the compiler prints them; for kernels templated on their
    workgroup size, without that trailing argument and the closing bracket)."""
    import re
    if name == "k_rhs_counts":
        return "k_rhs_counts_fused"
    m = re.match(r"(\w+)<(.*)>$", name)
    if not m:
        return name
    base, args = m.group(1), m.group(2).split(",")
    b = lambda x: "true" if x else "false"   # noqa: E731
    if base == "k_reduce_vals":
        if args[0] == "contact":
            mixed = args[1] == "mixed"
            return f"k_reduce_vals<{b(mixed)}, false, true, {'float' if mixed else 'double'}"
        mixed, iso, acc = args[0] == "mixed", args[1] == "iso", "acc" in args
        return f"k_reduce_vals<{b(mixed)}, {b(iso)}, {b(acc)}, {'float' if mixed else 'double'}"
    if base in ("k_tet_phase1",):
        return f"{base}<{b(args[0] == 'mixed')}, {b(args[1] == 'iso')}>"
    if base in ("k_cz_phase1",):
        return f"{base}<{b(args[0] == 'mixed')}>"
    if base == "k_contact_impact":
        return f"{base}<{'float' if args[0] == 'mixed' else 'double'}>"
    if base == "k_phase1":
        # the third argument (nontemporal staging stores) follows from the workload's size
        return f"k_phase1_fused<{b(args[0] == 'mixed')}, {b(args[1] == 'iso')}"
    if base in ("k_vals_force", "k_vals_force_rhs"):
        mixed = args[0] == "mixed"
        return f"{base}_fused<{b(mixed)}, {b(args[1] == 'iso')}, {'float' if mixed else 'double'}"
    if base == "k_contact_sums":
        mixed = args[0] == "mixed"
        return f"k_contact_sums<{b(mixed)}, {'float' if mixed else 'double'}"
    return name


def key_limit(gf, name):
    """The chunk shape (workgroup size, staged contributions per key) the autotuner chose for a
    value-reduction kernel, the lanes per tuple of the closest-face search, or None for kernels
    without one."""
    base = name.split("<")[0]
    if base == "k_closest_face":      # templated on its lanes per tuple, a tuned launch shape
        return gf.shape_state()["closest"][0]
    if base not in ("k_vals_force", "k_vals_force_rhs", "k_contact_sums", "k_reduce_vals"):
        return None
    stage = "fp64" if "fp64" in name else "fp32"
    acc = base == "k_contact_sums" or "acc" in name or "contact" in name
    return gf.autotune_state()[f"{'accumulate' if acc else 'overwrite'},{stage}"][0]


def _shaped_symbol(name, shape):
    """The rocprofv3 symbol of a kernel with its tuned shape: (workgroup size, staging) as the
    last two template arguments, or one integer template argument."""
    if isinstance(shape, int):
        return f"{_rocprof_symbol(name)}<{shape}>"
    return _rocprof_symbol(name) + (f", {shape[0]}, {shape[1]}>" if shape else "")


def profiled_traffic(name, shape=None, path=os.path.join(ROOT, "profiles", "assembly_compression40_kernels.md")):
    """HBM bytes per dispatch of a kernel from the committed rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE runs scaled by the copy-probe factors), or None.  shape: the
    (workgroup size, staging) of a kernel templated on them (its last two template arguments)."""
    sym = _shaped_symbol(name, shape)
    try:
        for line in open(path):
            head = f"| `{sym}"
            if line.startswith(head) and line[len(head)] in ",>`" and line.count("|") == 5:
                return float(line.split("|")[4]) * 1e6
    except OSError:
        return None
    return None


def profiled_valu(name, shape=None, path=None):
    """VALU wave-instructions per dispatch of a kernel from the committed rocprofv3 summary
    (SQ_INSTS_VALU of the wave-counter pass), or None."""
    sym = _shaped_symbol(name, shape)
    try:
        for line in open(path):
            head = f"| `{sym}"
            if line.startswith(head) and line[len(head)] in ",>`" and line.count("|") == 8:
                return float(line.split("|")[6]) * 1e6
    except (OSError, TypeError, ValueError):
        return None
    return None


def valu_probe(r, workgroups=4096, iters=4096, reps=5):
    """icf_probe_valu in FP64 and FP32: the achievable FMA rate (TFLOP/s) and the wave64
    VALU instructions per second it issues, best of `reps` blocking launches (HIP events)."""
    gf, lib, ctx = r.gf, r.lib, r.gf.ctx
    buf = gf.buf("valu_probe", workgroups * 256, np.float64)
    lib.icf_set_blocking(ctx, 1)
    out = {}
    for name, fp64 in (("fp64", 1), ("fp32", 0)):
        best = None
        for _ in range(reps):
            r.check(lib.icf_probe_valu(ctx, fp64, workgroups, iters, buf.p), ctx, "icf_probe_valu")
            ms = lib.icf_last_device_ms(ctx)
            best = ms if best is None else min(best, ms)
        instr = workgroups * 4 * iters * 16             # wave64 FMA instructions
        out[name] = dict(tflops=round(2 * 64 * instr / (best * 1e-3) / 1e12, 2),
                         wave_instr_per_s=round(instr / (best * 1e-3), 0),
                         frac_of_issue_peak=round(instr / (best * 1e-3) / VALU_ISSUE_PEAK, 4), ms=round(best, 4))
    buf.free()
    gf._bufs.pop("valu_probe", None)
    out["note"] = ("16 independent FMA chains per thread, 4096 x 256 threads; issue peak = 1024 SIMDs x 2.4 GHz / 4 "
                   "cycles per wave64 instruction (78.6 TFLOP/s FP64 or unpacked FP32)")
    return out


def copy_bandwidth(r, mib=1024, reps=10):
    """A STREAM-like copy through icf_d2d (the library's 16-byte-per-thread copy kernel): GB/s
    of bytes read plus bytes written, best of `reps` blocking copies of `mib` MiB."""
    gf, lib, ctx = r.gf, r.lib, r.gf.ctx
    nb = mib << 20
    a, b = gf.buf("bw_a", nb // 8, np.float64), gf.buf("bw_b", nb // 8, np.float64)
    a.zero()
    lib.icf_set_blocking(ctx, 1)
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        r.check(lib.icf_d2d(ctx, b.p, 0, a.p, 0, nb), ctx, "icf_d2d")
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    a.free()
    b.free()
    gf._bufs.pop("bw_a", None)
    gf._bufs.pop("bw_b", None)
    return dict(bytes=2 * nb, GBs=round(2 * nb / best / 1e9, 1), reps=reps,
                note="best blocking icf_d2d copy, host-timed (includes the call's launch and wait)")


def broad_phase_timing(r, reps=5):
    """icf_broad_phase over the bench's surface tets: pair count (equal to the host k-DOP
    grid's) and milliseconds per call (blocking, includes its host count read-back)."""
    gf, lib, ctx = r.gf, r.lib, r.gf.ctx
    nS = len(r.col.surf)
    d = gf.buf("bp_bench", 2 * r.Q, np.int32)
    n = ctypes.c_longlong()
    r.check(lib.icf_broad_phase(ctx, gf.g_dn.p, gf.nd_stride, gf.g_ie.p, gf.el_all_stride, nS, 1, d.p, r.Q,
                                ctypes.byref(n)), ctx, "icf_broad_phase")
    same = bool(n.value == r.Q and np.array_equal(d.download(2 * r.Q), r.pairs.ravel()))
    t0 = time.perf_counter()
    for _ in range(reps):
        lib.icf_broad_phase(ctx, gf.g_dn.p, gf.nd_stride, gf.g_ie.p, gf.el_all_stride, nS, 1, d.p, r.Q,
                            ctypes.byref(n))
    ms = 1e3 * (time.perf_counter() - t0) / reps
    return {"pairs": int(n.value), "same_as_host": same, "ms_per_call": round(ms, 3)}


def device_iteration_timing(r, reps=5):
    """One Newton iteration's device work in the reference's order (ImplicitModel3.cs:376-397),
    contact detection included: broad phase over the surface tets (icf_broad_phase, for the
    host BVH), the narrow phase on a copy of the pairs, NL2set, closest face, pattern and pcsr
    packing (icf_build_pattern, icf_pack_pcsr, for CreateStructure and TransferPCSR), the
    fused assembly, the damage counts and the collision response.  Blocking calls: the host
    reads the pair and tuple counts.  The impact rows keep the first iteration's node ids
    (the state does not move here, so the impacts are the same).  Milliseconds per part and
    per iteration (host clock)."""
    gf, lib, ctx, chk, col = r.gf, r.lib, r.gf.ctx, r.check, r.col
    nS, Q, ts, P = len(col.surf), r.Q, gf.tet_stride, r.nT
    T, C, N, nnzb = len(col.elastic), len(col.czs), gf.csr.N, gf.csr.nnz
    es, cs, cr = gf.el_elastic_stride, gf.cz_stride, gf.cr_stride
    pairs = gf.buf("it_pairs", 2 * Q, np.int32)
    alt = gf.buf("it_alt", len(col.alt), np.int32)
    alt.upload(np.ascontiguousarray(col.alt, np.int32))
    rows, cols = gf.buf("it_rows", N + 1, np.int32), gf.buf("it_cols", nnzb, np.int32)
    nq, nt, nz, nd, nf = ctypes.c_longlong(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib.icf_set_blocking(ctx, 1)

    def narrow():
        chk(lib.icf_d2d(ctx, r.itet.p, 0, pairs.p, 0, 8 * nq.value), ctx, "icf_d2d")
        return lib.icf_narrow_phase(ctx, nq.value, gf.g_dn.p, gf.g_ie.p, r.itet.p, gf.el_all_stride, gf.nd_stride)

    parts = [
        ("broad_phase", lambda: lib.icf_broad_phase(ctx, gf.g_dn.p, gf.nd_stride, gf.g_ie.p, gf.el_all_stride, nS,
                                                    1, pairs.p, Q, ctypes.byref(nq))),
        ("narrow_phase", narrow),
        ("nl2set", lambda: lib.icf_nl2set(ctx, r.itet.p, pairs.p, nq.value, gf.g_ie.p, gf.el_all_stride, r.nl2.p,
                                          ts, ctypes.byref(nt))),
        ("closest_face", lambda: lib.icf_find_closest_face(ctx, nt.value, ts, r.nl2.p, gf.g_dn.p, gf.g_ie.p,
                                                           gf.g_ifc.p, gf.nd_stride, gf.el_all_stride,
                                                           gf.fc_stride)),
        ("pattern", lambda: lib.icf_build_pattern(ctx, alt.p, len(col.alt), gf.g_ie_pcsr.p, T, es, gf.g_icz.p, C,
                                                  cs, r.icr.p, P, cr, N, rows.p, cols.p, nnzb, ctypes.byref(nz))),
        ("pack_pcsr", lambda: lib.icf_pack_pcsr(ctx, alt.p, len(col.alt), rows.p, cols.p, N, gf.g_ie_pcsr.p, T, es,
                                                gf.g_icz.p, C, cs, r.icr.p, P, cr)),
        ("assembly", lambda: lib.icf_assemble_elems_and_czs(
            ctx, gf.g_ie_pcsr.p, T, es, gf.g_dcz.p, gf.g_icz.p, C, cs, gf.g_dn.p, gf.nd_stride, r.h, r.vals.p, nnzb,
            r.rhs.p, N, None, None)),
        ("cz_counts", lambda: lib.icf_cz_counts(ctx, ctypes.byref(nd), ctypes.byref(nf))),
        ("collision_response", lambda: lib.icf_collision_response(ctx, r.icr.p, gf.g_dn.p, r.h, r.vals.p, r.rhs.p, P,
                                                                  cr, gf.nd_stride, r.k, r.eps)),
    ]
    ms = {name: 0.0 for name, _ in parts}
    for rep in range(reps + 1):                  # the first pass is a warm-up
        for name, fn in parts:
            t0 = time.perf_counter()
            chk(fn(), ctx, name)
            if rep:
                ms[name] += 1e3 * (time.perf_counter() - t0)
        if nq.value != Q or nt.value != P or nz.value != nnzb:
            raise RuntimeError(f"device iteration: {nq.value} pairs, {nt.value} tuples, {nz.value} blocks; "
                               f"expected {Q}, {P}, {nnzb}")
    out = {k: round(v / reps, 3) for k, v in ms.items()}
    return {"ms_per_iteration": round(sum(out.values()), 3), "parts_ms": out}


def host_solve_once(r):
    """One FP64 iteration assembled on the device, then solved once on the host by SolveDouble3
    (the reference's PardisoLoader2 signature, not PARDISO), timed apart from the assembly:
    the analysis alone (icfh_solve_analyse: factor size, peak memory, multiply-adds), then the
    call (factorisation and substitution, the analysis kept), and the relative residual."""
    from icflow_amd import _lib, newton
    lib, gf = r.lib, r.gf
    r.set_mode(_lib.ICF_FP64)
    r.step()
    N, nnz = gf.csr.N, gf.csr.nnz
    A = r.vals.download(9 * nnz).copy()
    b = r.rhs.download(3 * N).copy()
    rows = np.ascontiguousarray(gf.csr.rows, np.int32)
    cols = np.ascontiguousarray(gf.csr.cols, np.int32)
    fd, pk, fl = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_double()
    t = time.time()
    rc = lib.icfh_solve_analyse(_ptr(cols), _ptr(rows), N, ctypes.byref(fd), ctypes.byref(pk), ctypes.byref(fl))
    ta = time.time() - t
    log(f"host solve: analysis {ta:.1f} s, factor {fd.value * 8e-9:.1f} GB, {fl.value * 1e-9:.0f} G multiply-adds")
    x = np.zeros_like(b)
    t = time.time()
    rc2 = lib.SolveDouble3(_ptr(cols), _ptr(rows), _ptr(A), N, _ptr(b), _ptr(x), -2, 0, 3, 2, 1)
    ts = time.time() - t
    M = newton.upper_bcsr_to_csc(rows, cols, A)
    res = float(np.linalg.norm(M @ x - b) / np.linalg.norm(b))
    return {"rc": [int(rc), int(rc2)], "block_rows": int(N), "blocks": int(nnz), "analysis_s": round(ta, 2),
            "factorise_and_solve_s": round(ts, 1), "factor_GB": round(fd.value * 8e-9, 2),
            "peak_GB": round(pk.value * 8e-9, 2), "gmadds": round(fl.value * 1e-9, 1), "rel_residual": res,
            "threads": host_threads()}


def host_threads():
    """Threads for the all-core CPU baseline: OMP_NUM_THREADS when set (the GPU box sets it to
    the CPU share of its one GPU), else the CPUs this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else len(os.sched_getaffinity(0))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(runner, seconds, units):
    """The oracle's element, CZ and contact passes of the same iteration on the host
    (orc_assemble_par: OpenMP over units, compare-and-swap scatter additions as the reference's
    atomicAdd2; A and b zeroed inside, as the fused device call does), timed on all threads of
    this process's CPU share and on one thread.  Buffers are allocated once, outside the timed
    passes.  Units are the device value's: tets + active CZs + impacts."""
    from oracle import oracle as orc
    gf, col = runner.gf, runner.col
    orc.set_constants(gf.consts)
    T, C, P = len(col.elastic), len(col.czs), runner.nT
    dn = gf.c_dn.copy()
    A = np.zeros(gf.csr.nnz * 9)
    b = np.zeros(gf.csr.N * 3)
    ie = gf.c_ie_pcsr
    icz0 = gf.c_icz.copy() if C else None
    dcz0 = gf.c_dcz.copy() if C else None
    icz = icz0.copy() if C else None
    dcz = dcz0.copy() if C else None
    icr = getattr(gf, "c_icr", None) if P else None
    cr = getattr(gf, "cr_stride", 0) if P else 0
    k = getattr(runner, "k", 0.0)
    eps = getattr(runner, "eps", 0.0)

    def one_pass(threads):
        orc.assemble_par(ie, T, gf.el_elastic_stride, dcz, icz, C, gf.cz_stride, icr, P, cr, dn, gf.nd_stride,
                         runner.h, k, eps, A, b, threads)

    def rate(threads, budget, warm):
        if warm:
            one_pass(threads)                   # caches and the thread pool
        passes, t0 = 0, time.perf_counter()
        while True:
            one_pass(threads)
            passes += 1
            el = time.perf_counter() - t0
            if el >= budget or passes >= 50:
                return units * passes / el / 1e6, passes, el

    n = host_threads()
    v_all, p_all, t_all = rate(n, seconds / 3, True)
    v_one, p_one, t_one = rate(1, 2 * seconds / 3, False)
    return dict(value=round(v_all, 4), unit="Munits/s", cores=n, kind="port",
                value_1core=round(v_one, 4), nproc=os.cpu_count(), affinity=len(os.sched_getaffinity(0)),
                cpu_model=cpu_model(),
                sample=(f"the whole iteration (elements, CZs, contacts: {units} units) through the C oracle's "
                        f"OpenMP pass with CAS scatter: {p_all} passes on {n} threads in {t_all:.1f} s, "
                        f"{p_one} on 1 thread in {t_one:.1f} s; the reference's own C# + MKL host cannot run here"))


def measure_extra(name, consts, col, st, contact, args, traffic_md=None):
    """A further workload beside the headline: autotuned, both modes timed (graph replays and
    eager non-blocking calls, the faster kept), the dominant kernel's algorithmic HBM rate from
    a HIP-event pass."""
    from icflow_amd import _lib
    r = Runner(consts, col, st, _lib.ICF_MIXED, contact=contact)
    res = {}
    for mname, mode in (("fp64", _lib.ICF_FP64), ("mixed", _lib.ICF_MIXED)):
        r.set_mode(mode)
        autotune(r, max(1, args.tune_rounds - 1), mname)
        t = {how: timed(r, args.steps, args.warmup, how) for how in ("async", "graph", "profile")}
        issue = min(("async", "graph"), key=lambda h: t[h][0])
        res[mname] = (issue, t[issue][0], t["profile"][1])
    xfer = timed(r, min(args.steps, 5), 1, "transfers")[0] if "transfers" in args.timings.split(",") else None
    T, C, P = len(col.elastic), len(col.czs), r.nT
    units = T + C - r.failed_czs() + P
    alg = algorithmic_bytes(r)
    issue, sec, prof = res["mixed"]
    dname, (dms, dn) = max(prof.items(), key=lambda kv: kv[1][0])
    avg_s = dms / dn / 1e3
    dbytes = kernel_bytes(alg, dname) or 0
    dkeys = key_limit(r.gf, dname)
    traffic = profiled_traffic(dname, dkeys, traffic_md) if traffic_md else None
    util = {}
    for kname, (ms, n) in prof.items():
        us = 1e3 * ms / max(n, 1)
        vi = profiled_valu(kname, key_limit(r.gf, kname), traffic_md) if traffic_md else None
        tb = profiled_traffic(kname, key_limit(r.gf, kname), traffic_md) if traffic_md else None
        util[kname] = dict(avg_us=round(us, 2),
                           traffic_frac=round(tb / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4) if tb else None,
                           valu_frac=round(vi / (us * 1e-6) / VALU_ISSUE_PEAK, 4) if vi else None)
    out = {"workload": name, "tets": T, "czs": C, "active_czs": C - r.failed_czs(), "impacts": P,
           "candidate_pairs": r.Q, "nnzb": r.gf.csr.nnz, "rows": r.gf.csr.N,
           "value": round(units / sec / 1e6, 3), "ms_per_step": round(sec * 1e3, 4), "issue": issue,
           "fp64": {"value": round(units / res["fp64"][1] / 1e6, 3), "ms_per_step": round(res["fp64"][1] * 1e3, 4),
                    "issue": res["fp64"][0]},
           "mixed_over_fp64": round(res["fp64"][1] / sec, 3),
           "dropin_with_transfers_ms_per_step": round(xfer * 1e3, 3) if xfer else None,
           "transfer_bytes_per_step": dict(zip(("host_to_device", "device_to_host"), r.transfers())) if xfer else None,
           "iteration_roofline_frac": round(iteration_bytes(r) / sec / 1e9 / HBM_PEAK_GBS, 4),
           "kernels_us": {k: round(1e3 * ms / max(n, 1), 2) for k, (ms, n) in prof.items()},
           "kernels_util": util,
           "roofline": {"kernel": dname, "alg_bytes": int(dbytes), "avg_us": round(avg_s * 1e6, 2),
                        "achieved": round(dbytes / avg_s / 1e9, 1),
                        "frac": round(dbytes / avg_s / 1e9 / HBM_PEAK_GBS, 4),
                        "traffic": int(traffic) if traffic else None}}
    if getattr(args, "solve_once", False):
        out["host_solve"] = host_solve_once(r)
        log(f"{name}: host solve {out['host_solve']}")
    if contact:
        out["device_broad_phase"] = broad_phase_timing(r)
        log(f"{name}: device broad phase {out['device_broad_phase']}")
        # last: icf_pack_pcsr resets the CZ flags the unit count above reads
        out["device_iteration"] = device_iteration_timing(r)
        log(f"{name}: device iteration {out['device_iteration']}")
    r.gf.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=40)
    ap.add_argument("--size", type=float, default=0.1, help="cube edge (m)")
    ap.add_argument("--platen-cells", type=int, default=32)
    ap.add_argument("--grains", type=int, default=500)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--platen-frac", type=float, default=0.3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--modes", default="fp64,mixed", help="precision modes to time (fp64, mixed)")
    ap.add_argument("--workload", default="batched", choices=["compression", "batched"],
                    help="batched: --specimens cubes, lognormal sizes (seed 2), no contact; compression: the "
                         "--cells^3 cube with platens and contact")
    ap.add_argument("--extra", default="compression",
                    help="comma list of further workloads timed after the headline: compression (40^3 cube with "
                         "contact), large (60^3 cube, 2000 grains, 48x48 platens); empty: none")
    ap.add_argument("--specimens", type=int, default=256)
    ap.add_argument("--newton-damage-steps", type=int, default=26,
                    help="simulated time (in 1e-4 s steps) of the FP64-vs-mixed three-point-bend run into fracture "
                         "(2 mm deflection ramped over 30 steps, the reference's time-step control; 0: skip)")
    ap.add_argument("--newton-converged", type=int, default=1,
                    help="also run the fracture study with every Newton loop converged (relative tolerance 1e-9, "
                         "no absolute cutoff, at most 40 iterations): mixed against FP64 without the reference's "
                         "early stop (0: skip)")
    ap.add_argument("--newton-contact-attempts", type=int, default=0,
                    help="attempts of an FP64-vs-mixed Newton run with contact of a cube between rigid platens "
                         "(the reference's time-step control; 0: skip)")
    ap.add_argument("--newton-self-contact-steps", type=int, default=30,
                    help="simulated time (in 1e-4 s steps) of the FP64-vs-mixed three-point-bend run with "
                         "self-contact (every exterior and CZ face in the contact surface; 2 mm ramped over 20 "
                         "steps, the reference's time-step control; 0: skip)")
    ap.add_argument("--newton-batched-specimens", type=int, default=32,
                    help="specimens (the first of the headline's 256) of the FP64-vs-mixed batched Newton run: "
                         "each cube pulled to 3e-4 strain over 10 steps, per-specimen solves on host threads "
                         "(0: skip)")
    ap.add_argument("--newton-batched-steps", type=int, default=4, help="fixed steps of the batched Newton run")
    ap.add_argument("--newton-steps", type=int, default=10,
                    help="steps of the FP64-vs-mixed three-point-bend Newton run (0: skip)")
    ap.add_argument("--no-overlap", action="store_true",
                    help="icf_set_overlap(0): one kernel at a time, on the library's stream (no fused "
                         "launches, no second stream)")
    ap.add_argument("--timings", default="dropin,async,graph,profile,transfers",
                    help="ways of issuing the timed steps (profile is always run: it gives the kernel times)")
    ap.add_argument("--tune-rounds", type=int, default=3,
                    help="icf_set_autotune rounds before timing each mode (0: keep the default key limit)")
    ap.add_argument("--solve-once", action="store_true",
                    help="for each extra workload, one FP64 iteration solved on the host by SolveDouble3, timed")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run icf_d2d over this many MiB (a plain copy of known size, for "
                         "calibrating hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path (requested --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    from icflow_amd import _lib
    contact = args.workload == "compression"
    if contact:
        consts, col, st = build_workload(args.cells, args.grains, args.seed, args.platen_frac, args.size,
                                         args.platen_cells)
    else:
        consts, col, st = build_batched(args.specimens, 2)
    r = Runner(consts, col, st, _lib.ICF_MIXED, contact=contact)
    if args.no_overlap:
        r.check(r.lib.icf_set_overlap(r.gf.ctx, 0), r.gf.ctx, "icf_set_overlap")
    hows = [h for h in ("dropin", "async", "graph") if h in args.timings.split(",")] + ["profile"]
    if "transfers" in args.timings.split(","):
        hows.append("transfers")
    if not ({"async", "graph"} & set(hows)):
        hows.insert(0, "graph")

    # correctness guard of the two modes on this workload (assembled arrays, normwise)
    res = {}
    for mode in (_lib.ICF_FP64, _lib.ICF_MIXED):
        r.set_mode(mode)
        r.gf.TransferNodesToGPU(st)        # F back to zero
        r.step()
        res[mode] = (r.vals.download(r.gf.csr.nnz * 9), r.rhs.download(r.gf.csr.N * 3))
    errs = {}
    for i, k in enumerate(("vals", "rhs")):
        ref = res[_lib.ICF_FP64][i]
        errs[k] = float(np.max(np.abs(res[_lib.ICF_MIXED][i] - ref)) / max(np.max(np.abs(ref)), 1e-300))
    log(f"mixed vs fp64 normwise rel err: {errs}")

    out = {}
    tuned = {}
    modes = [m for m in args.modes.split(",") if m]
    for name, mode in (("fp64", _lib.ICF_FP64), ("mixed", _lib.ICF_MIXED)):
        if name not in modes:
            continue
        r.set_mode(mode)
        if args.tune_rounds > 0:
            tuned[name] = autotune(r, args.tune_rounds, name)
        log(f"timing {name}: {args.warmup} warmup + {args.steps} steps each way")
        res = {how: timed(r, min(args.steps, 5) if how == "transfers" else args.steps,
                          min(args.warmup, 1) if how == "transfers" else args.warmup, how) for how in hows}
        out[name] = res
        prof = res["profile"][1]
        for kname, (ms, n) in sorted(prof.items(), key=lambda kv: -kv[1][0]):
            log(f"    {kname:28s} {1e3 * ms / max(n, 1):9.1f} us/launch  x{n}")

    if args.copy_probe_mib > 0:
        nb = args.copy_probe_mib << 20
        a, b = r.gf.buf("probe_a", nb // 8, np.float64), r.gf.buf("probe_b", nb // 8, np.float64)
        a.zero()
        for _ in range(3):
            r.check(r.lib.icf_d2d(r.gf.ctx, b.p, 0, a.p, 0, nb), r.gf.ctx, "icf_d2d")
        log(f"copy probe: 3 x icf_d2d of {nb} bytes")

    T, C, P = len(col.elastic), len(col.czs), r.nT
    active_C = C - r.failed_czs()
    units = T + active_C + P
    head = "mixed" if "mixed" in out else next(iter(out))

    def best(name):
        """The faster of the two pipelined ways of issuing iterations (graph replays, or eager
        calls enqueued without waiting)."""
        t = {h: out[name][h][0] for h in ("graph", "async") if h in out[name]}
        h = min(t, key=t.get)
        return h, t[h]

    issue, sec = best(head)
    prof = out[head]["profile"][1]
    alg = algorithmic_bytes(r)
    # dominant kernel of the step (largest total time) and its roofline
    dname, (dms, dn) = max(prof.items(), key=lambda kv: kv[1][0])
    dbytes = kernel_bytes(alg, dname) or 0
    avg_s = dms / dn / 1e3
    achieved = dbytes / avg_s / 1e9
    dkeys = key_limit(r.gf, dname)
    # the committed counter summaries cover the default workloads only
    if contact:
        profiled = (args.cells, args.grains, args.seed, args.size, args.platen_cells) == (40, 500, 1, 0.1, 32)
        md = os.path.join(ROOT, "profiles", "assembly_compression40_kernels.md")
    else:
        profiled = args.specimens == 256
        md = os.path.join(ROOT, "profiles", "assembly_batched256_kernels.md")
    traffic = profiled_traffic(dname, dkeys, md) if profiled else None
    roof = dict(kernel=dname + (f" ({dkeys[0]}-lane workgroups, {dkeys[1]} staged per key)"
                                if isinstance(dkeys, tuple) else ""),
                bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=int(traffic) if traffic else None,
                alg_bytes=int(dbytes), avg_us=round(avg_s * 1e6, 2),
                traffic_source=(os.path.relpath(md, ROOT) + " (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)"
                                if traffic else None))
    it_bytes = iteration_bytes(r)
    # every kernel of the step against HBM peak: measured traffic (committed counter summary,
    # default workload only) and algorithmic bytes over its HIP-event duration
    vp = valu_probe(r)
    log(f"VALU probe: {vp}")
    util = {}
    for kname, (ms, n) in prof.items():
        us = 1e3 * ms / max(n, 1)
        tb = profiled_traffic(kname, key_limit(r.gf, kname), md) if profiled else None
        vi = profiled_valu(kname, key_limit(r.gf, kname), md) if profiled else None
        ab = kernel_bytes(alg, kname)
        util[kname] = dict(avg_us=round(us, 2),
                           traffic_frac=round(tb / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4) if tb else None,
                           alg_frac=round(ab / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4) if ab else None,
                           valu_frac=round(vi / (us * 1e-6) / VALU_ISSUE_PEAK, 4) if vi else None,
                           valu_frac_of_probe=(round(vi / (us * 1e-6) / vp["fp64"]["wave_instr_per_s"], 4)
                                               if vi else None))

    # device broad phase over the same surface tets (reported beside the step; the reference's
    # broad phase is a host BVH outside the assembly path)
    bw = copy_bandwidth(r)
    log(f"copy probe: {bw}")
    roof["frac_of_copy_probe"] = round(achieved / bw["GBs"], 4)
    roof["valu_frac"] = util.get(dname, {}).get("valu_frac")
    bp = broad_phase_timing(r) if contact else None
    if bp:
        log(f"device broad phase: {bp}")
    study = None
    if args.newton_steps > 0:
        from icflow_amd import newton
        t0 = time.time()
        study = newton.precision_study(consts, steps=args.newton_steps, log=log)
        study["wall_s"] = round(time.time() - t0, 1)
        log(f"precision study: {study}")
    last = [time.time()]

    def every_15s(msg):                     # progress of a long run, a line at a time
        if time.time() - last[0] > 15:
            last[0] = time.time()
            log(msg)

    study_damage = None
    if args.newton_damage_steps > 0:
        from icflow_amd import newton
        t0 = time.time()
        study_damage = newton.precision_study(consts, steps=args.newton_damage_steps, ramp_steps=30, adapt=True,
                                              log=every_15s)
        study_damage["wall_s"] = round(time.time() - t0, 1)
        log(f"precision study into fracture: {study_damage}")
    study_conv = None
    if args.newton_damage_steps > 0 and args.newton_converged:
        from icflow_amd import newton
        t0 = time.time()
        study_conv = newton.precision_study(consts, steps=args.newton_damage_steps, ramp_steps=30, adapt=True,
                                            conv_eps=1e-9, conv_cutoff=1e-30, max_iter=40, log=every_15s)
        study_conv["wall_s"] = round(time.time() - t0, 1)
        log(f"precision study into fracture, Newton converged: {study_conv}")
    study_contact = None
    if args.newton_contact_attempts > 0:
        from icflow_amd import newton
        t0 = time.time()
        study_contact = newton.contact_precision_study(consts, steps=20, adapt=True,
                                                       max_attempts=args.newton_contact_attempts)
        study_contact["wall_s"] = round(time.time() - t0, 1)
        log(f"precision study with contact: {study_contact}")
    study_self = None
    if args.newton_self_contact_steps > 0:
        from icflow_amd import newton
        t0 = time.time()
        study_self = newton.precision_study(consts, steps=args.newton_self_contact_steps, ramp_steps=20, adapt=True,
                                            contact=True, log=every_15s)
        study_self["wall_s"] = round(time.time() - t0, 1)
        log(f"precision study with self-contact: {study_self}")
    study_batched = None
    if args.newton_batched_specimens > 0:
        from icflow_amd import newton
        t0 = time.time()
        study_batched = newton.batched_precision_study(consts, n_specimens=args.newton_batched_specimens,
                                                       steps=args.newton_batched_steps, log=log)
        study_batched["wall_s"] = round(time.time() - t0, 1)
        log(f"batched precision study: {study_batched}")
    cpu = None if args.no_cpu else cpu_baseline(r, args.cpu_seconds, units)
    if cpu:
        log(f"cpu baseline: {cpu}")
    # last on this runner: icf_pack_pcsr resets the CZ flags the unit count reads
    dev_it = device_iteration_timing(r) if contact else None
    if dev_it:
        log(f"device iteration: {dev_it}")
    line = {
        "metric": "Newton-iter assembly Melem+CZ/s, mixed vs FP64 at <=1e-5 rel err; HBM %peak",
        "value": round(units / sec / 1e6, 3),
        "unit": "Melem+CZ/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if head == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": (f"compression with contact: {args.size} m cube, {args.cells}^3 hex x 6 tets, "
                                f"{args.grains} grains (seed {args.seed}), PPR CZs on inter-grain faces, 2 rigid "
                                f"{args.platen_cells}x{args.platen_cells} platens"
                                if contact else
                                f"batched: {args.specimens} cubes, lognormal tet counts (median 12k, 2k-40k, "
                                f"log-sd 0.6, seed 2), 8-64 grains each, PPR CZs, no contact, one launch per phase"),
                   "tets": T, "czs": C, "active_czs": active_C, "impacts": P, "candidate_pairs": r.Q,
                   "nodes": len(col.x0), "rows": r.gf.csr.N, "nnzb": r.gf.csr.nnz, "mode": head},
        "issue": issue,
        "timing": ("value from the faster of: graph = one captured HIP graph per Newton iteration, replayed"
                   + ("" if args.no_overlap else " (tets beside CZs, BCSR values beside node forces, rhs rows "
                                                  "beside CZ counts, the impact sums together: one launch each)")
                   + "; async = the iteration's calls enqueued without waiting"
                   + ("" if args.no_overlap else " (contact detection on a second stream beside the assembly)")
                   + ". Per-kernel times from a separate pass with HIP events around each kernel, all on "
                     "the library's stream"),
        "graph_ms_per_step": {k: round(out[k]["graph"][0] * 1e3, 4) for k in out if "graph" in out[k]},
        "async_ms_per_step": {k: round(out[k]["async"][0] * 1e3, 4) for k in out if "async" in out[k]},
        "dropin_ms_per_step": {k: round(out[k]["dropin"][0] * 1e3, 4) for k in out if "dropin" in out[k]},
        "dropin_with_transfers_ms_per_step": (
            {k: round(out[k]["transfers"][0] * 1e3, 3) for k in out if "transfers" in out[k]} or None),
        "transfer_bytes_per_step": (dict(zip(("host_to_device", "device_to_host"), r.transfers()))
                                    if "transfers" in hows else None),
        "fp64": ({"value": round(units / best("fp64")[1] / 1e6, 3), "ms_per_step": round(best("fp64")[1] * 1e3, 4),
                  "issue": best("fp64")[0]} if "fp64" in out else None),
        "mixed_over_fp64": (round(best("fp64")[1] / best("mixed")[1], 3)
                            if "fp64" in out and "mixed" in out else None),
        "iteration_roofline": {"alg_bytes": it_bytes, "achieved_GBs": round(it_bytes / sec / 1e9, 1),
                               "frac": round(it_bytes / sec / 1e9 / HBM_PEAK_GBS, 4)},
        "mixed_rel_err_vs_fp64": errs,
        "kernels_us": {k: round(1e3 * ms / max(n, 1), 2) for k, (ms, n) in prof.items()},
        "kernels_hbm": util,
        "kernels_util_note": ("per kernel: traffic_frac = copy-calibrated FETCH_SIZE + WRITE_SIZE bytes per dispatch "
                              "(committed rocprofv3 summary) / live HIP-event duration / 8 TB/s; alg_frac = "
                              "algorithmic bytes / duration / 8 TB/s; valu_frac = SQ_INSTS_VALU per dispatch (same "
                              "summary) / duration / VALU issue peak; valu_frac_of_probe = the same against the "
                              "FP64 FMA probe's measured issue rate"),
        "valu_probe": vp,
        "autotune": tuned or None,
        "roofline": roof,
        "hbm_copy_probe": bw,
        "device_broad_phase": bp,
        "device_iteration": dev_it,
        "precision_study": study,
        "precision_study_fracture": study_damage,
        "precision_study_fracture_converged": study_conv,
        "precision_study_contact": study_contact,
        "precision_study_self_contact": study_self,
        "precision_study_batched": study_batched,
        "cpu_baseline": cpu,
    }
    r.gf.close()
    extra = {}
    for ex in [e for e in args.extra.split(",") if e]:
        if ex == ("compression" if contact else "batched"):
            continue
        t0 = time.time()
        if ex == "compression":
            c2, col2, st2 = build_workload(40, 500, 1, 0.3, 0.1, 32)
            extra[ex] = measure_extra("compression with contact: 0.1 m cube, 40^3 hex x 6 tets, 500 grains (seed 1), "
                                      "2 rigid 32x32 platens", c2, col2, st2, True, args,
                                      os.path.join(ROOT, "profiles", "assembly_compression40_kernels.md"))
        elif ex == "large":
            c2, col2, st2 = build_workload(60, 2000, 3, 0.3, 0.2, 48)
            extra[ex] = measure_extra("compression with contact: 0.2 m cube, 60^3 hex x 6 tets, 2000 grains (seed 3), "
                                      "2 rigid 48x48 platens", c2, col2, st2, True, args)
        elif ex == "batched":
            c2, col2, st2 = build_batched(args.specimens, 2)
            extra[ex] = measure_extra(f"batched: {args.specimens} cubes", c2, col2, st2, False, args)
        else:
            raise SystemExit(f"unknown workload {ex}")
        extra[ex]["wall_s"] = round(time.time() - t0, 1)
        log(f"extra workload {ex}: {extra[ex]}")
    line["extra_workloads"] = extra or None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

/*
 * icflow_hip.h -- C ABI of libicflow_hip.so, the MI355X (gfx950) drop-in for the device side
 * of icFlow's per-Newton-iteration assembly path.
 *
 * It replaces the ManagedCUDA calls made by v3Library/Simulation/GPU/GPU_Functionality.cs:
 * one context object instead of CudaContext + three PTX modules, one call per kernel launch
 * with the reference kernel's own argument list, and the same SoA buffers and strides
 * (value(item, field) = buf[item + stride * field]).  Plain pointers and sizes only; the
 * caller may be C#, via [DllImport("icflow_hip")].
 *
 * Semantics kept from the reference:
 *  - vals (A), rhs (b) and the node force column F of dn are ACCUMULATED into, in the fixed
 *    order unit 0..n-1 (the reference used CAS-loop atomics, whose order is arbitrary; here
 *    the order is the sequential one, so results are bitwise reproducible run to run);
 *  - a slot or row index of -1 means "skip"; only the upper triangle (altId_j >= altId_i) is
 *    written; blocks are row-major 3x3, A[slot*9 + 3k + l];
 *  - every call is blocking (synchronous on the context's stream), like the reference's Run().
 *
 * All functions return 0 on success, otherwise a positive hipError_t value or a negative
 * ICF_E_* code; icf_last_error_string() describes the last failure on that context.
 * A context is not re-entrant: one host thread at a time.  Contexts on one device may be used
 * from different threads: the constants live in one per-device block, and a call that needs
 * them holds a per-device lock from loading its context's constants (after all queued device
 * work, when another context's were loaded) to its last launch.
 */
#ifndef ICFLOW_HIP_H
#define ICFLOW_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ICF_OK 0
#define ICF_E_INVALID (-1)      /* bad argument (null pointer, negative count, bad stride)  */
#define ICF_E_NOCONST (-2)      /* icf_set_constants has not been called                    */
#define ICF_E_CAPACITY (-3)     /* more units than the 24-bit unit field of the slot maps   */
#define ICF_E_CONSTANTS (-4)    /* constants outside the supported form (see icf_consts)    */
#define ICF_E_INDEX (-5)        /* a slot / row / node index beyond the declared size       */
#define ICF_E_CAPTURE (-6)      /* not allowed while a graph is being captured (see below)  */

#define ICF_FP64 0              /* plain FP64 throughout                                    */
#define ICF_MIXED 1             /* FP32 tangent blocks, FP64 residual, law and accumulation */

typedef struct icf_ctx icf_ctx;

/* Model constants, the __constant__ symbols that GPU_Functionality.SetConstants uploads by
 * name (GPU_Functionality.cs:151-194; elem_kernels9/kernel.cu:6-9; cz_kernels9/kernel.cu:5-20),
 * in the layout of the host arrays ModelPrms.SetComputedVariables builds
 * (v3Library/Simulation/DataSheets/ModelPrms.cs:237-322).
 * Supported form: E symmetric; B[gp][p][k] = +sf[k/3][gp] (k<9) or -sf[k/3-3][gp] (k>=9) when
 * p == k%3 and 0 otherwise, which is what ModelPrms.k_Bmatrix (ModelPrms.cs:324-349) produces. */
typedef struct icf_consts {
    double E[36];          /* 6x6 elasticity matrix, E[i*6+j]                    */
    double M[144];         /* 12x12 consistent mass per unit volume, M[i*12+j]   */
    double NewmarkBeta, NewmarkGamma, dampingMass, dampingStiffness, rho, gravity;
    double YoungsModulus;
    double G_fn, G_ft, f_tn, f_tt, alpha, beta, rn, rt;
    double p_m, p_n, deln, delt, pMtn, pMnt, gam_n, gam_t;
    double B[162];         /* CZ B[gp][3][18]                                     */
    double sf[9];          /* CZ shape functions at the Gauss points, sf[node*3+gp] */
} icf_consts;

/* --- context (replaces new CudaContext + LoadModulePTX, GPU_Functionality.cs:112-137) --- */
int icf_create(int device, icf_ctx **out);
int icf_destroy(icf_ctx *ctx);
const char *icf_last_error_string(const icf_ctx *ctx);
int icf_set_constants(icf_ctx *ctx, const icf_consts *c);     /* SetConstants, :151-194    */
int icf_set_precision(icf_ctx *ctx, int mode);                 /* ICF_FP64 | ICF_MIXED      */
int icf_get_precision(const icf_ctx *ctx);

/* --- device memory (replaces CudaDeviceVariable<T>, :211-248; CopyToDevice/CopyToHost;
 *     MemsetAsync(0, NullStream), :433-434) --- */
int icf_alloc(icf_ctx *ctx, size_t bytes, void **dptr);
int icf_free(icf_ctx *ctx, void *dptr);
int icf_h2d(icf_ctx *ctx, void *dst, size_t dst_off, const void *src, size_t bytes);
int icf_d2h(icf_ctx *ctx, void *dst, const void *src, size_t src_off, size_t bytes);
int icf_memset(icf_ctx *ctx, void *dst, int value, size_t bytes);
/* device-to-device copy (CudaDeviceVariable.CopyToDevice(CudaDeviceVariable)) */
int icf_d2d(icf_ctx *ctx, void *dst, size_t dst_off, const void *src, size_t src_off, size_t bytes);
int icf_synchronize(icf_ctx *ctx);

/* --- the five reference kernels, same argument lists (plus the context) --- */

/* kelElementElasticityForce (elem_kernels9/element_kernels2.h:195-263; launch
 * GPU_Functionality.cs:448-454): corotational tet, Newmark effective tangent + residual.
 * ie: int SoA 28 x el_stride (0-3 node ids, 4-19 slots, 20-23 rhs rows). */
int icf_elem_force(icf_ctx *ctx, const int *ie, double *dn, double h, double *A, double *b,
                   int nElems, int el_stride, int nd_stride);

/* kczCZForce (cz_kernels9/cz_kernels.h:210-434; launch GPU_Functionality.cs:463-468):
 * PPR cohesive zone, 3 Gauss points; writes tentative pmax/tmax/averages and flags.
 * icz: int SoA 58 x cz_stride; dcz: double SoA 16 x cz_stride. */
int icf_cz_force(icf_ctx *ctx, double *dcz, int *icz, double *dn, double *A, double *b, double h,
                 int nCZs, int cz_stride, int nd_stride);

/* kNarrowPhase_new (collision_kernels9/narrowphase.h:46-155; launch GPU_Functionality.cs:557-559):
 * narrowList[2i], [2i+1] = surface-tet ids in, narrowList[2i] = 8-bit containment mask out. */
int icf_narrow_phase(icf_ctx *ctx, int nPairs, const double *dn, const int *ie, int *narrowList,
                     int el_all_stride, int nd_stride);

/* kFindClosestFace (narrowphase.h:159-201; launch GPU_Functionality.cs:602-607):
 * nl2 SoA 4 x tet_stride, rows (node, tet) in -> rows 1..3 = closest face's node ids out. */
int icf_find_closest_face(icf_ctx *ctx, int nPairs, int tet_stride, int *nl2, const double *dn,
                          const int *ie, const int *faces, int nd_stride, int el_all_stride,
                          int fc_stride);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops): icf_find_closest_face with the same
 * arguments and results, bit for bit, that calls no hipMalloc/hipFree or allocating library
 * routine.  Its scratch (the compact adjacent-face lists of ie, rebuilt on every call) lives
 * in one fixed 64 MiB device array per GPU, shared by every context and call on that GPU
 * and held until the process ends.  It needs 4 (el_all_stride + 1) + 4 el_all_stride x
 * (adjacent-face rows of ie) bytes plus a small scan workspace; larger sizes return
 * ICF_E_CAPACITY before any work is issued.  Calls that overlap on one GPU (two contexts,
 * or non-blocking calls on different streams) race on the array: order them. */
int icf_find_closest_face_noalloc(icf_ctx *ctx, int nPairs, int tet_stride, int *nl2,
                                  const double *dn, const int *ie, const int *faces,
                                  int nd_stride, int el_all_stride, int fc_stride);

/* kCollisionResponseForce (collision_kernels9/collision_response.h:38-126; launch
 * GPU_Functionality.cs:668-672): point-triangle penalty force and tangent.
 * icr: int SoA 28 x cr_stride (0-3 node ids p,v0,v1,v2; 4-19 slots; 20-23 rows). */
int icf_collision_response(icf_ctx *ctx, const int *icr, double *dn, double h, double *A,
                           double *b, int nImpacts, int cr_stride, int nd_stride, double k,
                           double distanceEpsilon);

/* --- composite calls of the reference's host methods --- */

/* AssembleElemsAndCZs (GPU_Functionality.cs:439-479 with AllocateMemoryForLinearSystem
 * :419-437): zero A[0..nnzb*9) and b[0..N*3), element pass, CZ pass, then the two flag sums
 * (tentative damaged / failed counts, :470-475).  Each BCSR block and rhs entry is written
 * once.  nCZs may be 0 (dcz/icz may then be NULL).  Count outputs may be NULL. */
int icf_assemble_elems_and_czs(icf_ctx *ctx, const int *ie, int nElems, int el_stride,
                               double *dcz, int *icz, int nCZs, int cz_stride,
                               double *dn, int nd_stride, double h,
                               double *A, int nnzb, double *b, int N,
                               int *nCZDamaged, int *nCZFailed);

/* Batched specimens: one icf_assemble_elems_and_czs over many independent specimens whose
 * units lie back to back in its buffers.  Specimen s owns the elastic tets
 * [tet_off[s], tet_off[s+1]), the CZs [cz_off[s], cz_off[s+1]), the rhs rows
 * [row_off[s], row_off[s+1]) and the BCSR blocks [nnzb_off[s], nnzb_off[s+1]) (host arrays of
 * nSpecimens + 1 ints, from 0, non-decreasing); the totals tet_off[n], cz_off[n], row_off[n],
 * nnzb_off[n] are the call's nElems, nCZs, N and nnzb.  Every phase is one launch over all
 * specimens, and the results are those of icf_assemble_elems_and_czs over the whole lists, bit
 * for bit.  Beyond that call:
 *  - nCZDamaged[s], nCZFailed[s] (host arrays of nSpecimens, may be NULL; filled in blocking
 *    mode): the flag sums of GPU_Functionality.cs:470-475 per specimen; icf_batch_cz_counts
 *    reads those of the last batched call at any time (it waits for the device);
 *  - the layout is checked on the first call and whenever the slot or row fields change: a
 *    unit that writes a block or row outside its specimen's ranges fails with ICF_E_INDEX.
 * New offsets are copied to the device (not during capture).  No reference counterpart: the
 * reference runs one specimen per process (ImplicitModel3.Step, ImplicitModel3.cs:377-435). */
typedef struct icf_worklist {
    int nSpecimens;
    const int *tet_off;
    const int *cz_off;
    const int *row_off;
    const int *nnzb_off;
} icf_worklist;
int icf_assemble_batched(icf_ctx *ctx, const icf_worklist *wl, const int *ie, int el_stride,
                         double *dcz, int *icz, int cz_stride, double *dn, int nd_stride,
                         double h, double *A, double *b, int *nCZDamaged, int *nCZFailed);
int icf_batch_cz_counts(icf_ctx *ctx, int *nCZDamaged, int *nCZFailed);

/* Sum of n ints in device memory (GPU_Functionality.Sum, :93-107). */
int icf_sum_int(icf_ctx *ctx, const int *d, int n, int *out);

/* The library caches, per pcsr buffer, the slot -> contribution lists of its deterministic
 * reduction.  Writes through icf_h2d / icf_memset / icf_free invalidate the lists of the
 * buffers they touch; call this after writing pcsr arrays by any other means. */
int icf_pattern_changed(icf_ctx *ctx);

/* Time, in milliseconds, of the last call's device work (HIP events on the context's
 * stream); -1 if the last call launched nothing. */
double icf_last_device_ms(const icf_ctx *ctx);

/* VALU issue-rate probe (no counterpart in the reference: a measurement aid beside the
 * icf_d2d copy used as the HBM probe).  Launches `workgroups` x 256 threads that each run
 * `iters` rounds of 16 independent FMA chains in FP64 (fp64 != 0) or FP32 and store one value
 * each into `out` (device memory of workgroups * 256 doubles or floats).  Time it with
 * icf_last_device_ms in blocking mode: 2 * 16 * iters * 256 * workgroups flops. */
int icf_probe_valu(icf_ctx *ctx, int fp64, int workgroups, int iters, void *out);

/* Per-kernel timing (the role of FrameInfo's KerElemForce / KerCZForce / CollForce
 * stopwatches, FrameInfo.cs:75-105): when on, every kernel launch is bracketed by HIP events
 * on the context's stream and its duration is accumulated under the kernel's name.
 * icf_profile_entry fills name (NUL-terminated, truncated to cap), total milliseconds and
 * launch count of entry i < icf_profile_count(). */
/* --- asynchronous use and iteration graphs ---
 * icf_set_blocking(ctx, 0): kernel entry points return once their work is enqueued on the
 * context's stream (the reference's RunAsync, GPU_Functionality.cs:451); icf_synchronize()
 * waits.  The fused call then leaves its two CZ counts on the device; icf_cz_counts() reads
 * them (and waits).
 * icf_graph_begin / icf_graph_end capture every call made in between into one HIP graph;
 * icf_graph_launch replays it (blocking unless icf_set_blocking(ctx, 0)).  A replay repeats
 * the captured calls exactly -- same buffers, sizes, time step and penalty parameters -- so
 * capture again when any of them changes.  The slot maps of the captured calls must already
 * exist (run the iteration once uncaptured); calls that would build one, grow an internal
 * buffer, or copy to or from the host return ICF_E_CAPTURE during capture.  icf_d2d is
 * allowed and is recorded as a copy kernel, except into the slot / row / node-id fields that
 * a slot map was built from (ICF_E_CAPTURE: each replay would change them behind the maps).
 * icf_graph_launch refuses (ICF_E_CAPTURE) a graph whose memory may be stale: one captured
 * on another context, or one captured before the context freed or moved memory it owns
 * (an internal buffer grew, a slot map was rebuilt or dropped, the adjacency lists were
 * dropped, new batch offsets), or before a write to fields a slot map was built from
 * (icf_h2d, icf_memset, icf_d2d, icf_pack_pcsr, the contact kernels' in-place outputs,
 * icf_free, icf_pattern_changed).  Capture it again then. */
typedef struct icf_graph icf_graph;
int icf_set_blocking(icf_ctx *ctx, int blocking);
int icf_cz_counts(icf_ctx *ctx, int *nCZDamaged, int *nCZFailed);
int icf_graph_begin(icf_ctx *ctx);
int icf_graph_end(icf_ctx *ctx, icf_graph **out);
int icf_graph_launch(icf_ctx *ctx, icf_graph *g);
int icf_graph_destroy(icf_graph *g);

int icf_set_profiling(icf_ctx *ctx, int on);
/* icf_set_overlap(ctx, 1) (the default) lets independent kernels share the GPU:
 *  - non-blocking (icf_set_blocking(ctx, 0)), outside capture: the contact-detection work
 *    (icf_narrow_phase, icf_find_closest_face, the per-impact phase of icf_collision_response)
 *    runs on a second stream of the context beside the assembly.  Each call waits for the
 *    earlier calls that write bytes it reads, or read or write bytes it writes, so results are
 *    those of one stream; device memory touched outside the library needs icf_synchronize().
 *  - every mode: independent phases share one launch -- icf_assemble_elems_and_czs runs its
 *    tets with its CZs, its BCSR values with its node forces and its rhs rows with its CZ
 *    counts; icf_collision_response adds its node forces, BCSR blocks and rhs rows in one.
 * 0 keeps every kernel on the context's stream, one phase per launch.  While profiling is on everything runs on one
 * stream, so each kernel's events time it alone.  No reference counterpart (the reference's
 * Run() issues one stream, GPU_Functionality.cs:451-476). */
int icf_set_overlap(icf_ctx *ctx, int on);

/* Autotuner of the BCSR-value reductions.  Their workgroups reduce chunks of consecutive keys
 * (BCSR blocks), one key per lane, with the chunk's contributions staged in LDS.  A chunk
 * shape is a key limit (the workgroup size: 64, 128, 256 or 512) and a staging size (stage =
 * 2 or 3 FP32 contributions per key, half as many FP64 ones); the ICF_CHUNK_CANDIDATES = 8
 * shapes are numbered c = 2 * (index of the key limit) + (stage - 2).  Until tuned: 256 keys,
 * stage 2 for FP32 staging, 3 for FP64.  There are four roles, r = 2 * accumulate + fp64:
 * overwrite maps (icf_assemble_elems_and_czs) and accumulate maps (icf_collision_response and
 * the single-kernel calls), each with FP32 (mixed) or FP64 staging.
 * icf_set_autotune(ctx, rounds > 0) starts tuning every role: its next launches take the
 * candidates in turn, each timed with events (waiting for the previous timed launch of the
 * role), until every candidate has `rounds` timings; then the fastest (least time over its
 * timings) is kept.  rounds = 0 stops tuning and keeps the fastest timed so far.  Captured
 * calls are never timed.  Time it in blocking mode or with overlap off for clean timings.
 * icf_get_autotune: per role, keys[r] and stage[r] = the shape in use, tuning[r] = 1 while
 * tuning, best_ms[r * ICF_CHUNK_CANDIDATES + c] = fastest timing of candidate c (0 = none);
 * any of the four may be null.  icf_set_chunk_keys fixes one role's shape (stage 0: keep the
 * role's staging) and ends its tuning.  Results do not depend on the shape: each key's
 * contributions are added in list order whatever the chunking.  No reference counterpart (the
 * reference launches every kernel with a fixed block_size = 128, GPU_Functionality.cs:38,123-136). */
#define ICF_CHUNK_CANDIDATES 8
int icf_set_autotune(icf_ctx *ctx, int rounds);
int icf_get_autotune(icf_ctx *ctx, int *keys, int *stage, int *tuning, float *best_ms);
int icf_set_chunk_keys(icf_ctx *ctx, int accumulate, int fp64, int keys, int stage);

/* Launch shapes besides the value reductions (no counterpart in the reference, which launches
 * everything with 128-thread blocks, GPU_Functionality.cs:38), tuned with the value reductions
 * by icf_set_autotune (the candidates of each role timed in turn on the role's launches):
 *   ICF_SHAPE_PHASE1   workgroup size of phase 1 of the fused call (64, 128, 256)
 *   ICF_SHAPE_NARROW   workgroup size of the narrow phase (64, 128, 256)
 *   ICF_SHAPE_CLOSEST  lanes per (node, tet) tuple of the closest face (4, 8, 16)
 *   ICF_SHAPE_IMPACT   workgroup size of the impacts' phase 1 (64, 128, 256)
 *   ICF_SHAPE_ORDER    the fused value launch's phases: 0 node forces first, then the BCSR
 *                      values; 1 interleaved in rounds of eight workgroups (0, 1)
 * Every candidate gives the same results bit for bit.  icf_set_shape fixes a role's value on
 * this context (ICF_E_INVALID for other values); icf_get_shape reports the context's value
 * (the default until tuned or set), whether the role is still being tuned and the fastest
 * timing per candidate (best_ms[3], 0: not timed or no such candidate).  Finished tunings, of these roles and of the
 * value reductions, are kept by the library per device, role, precision and size class (units
 * of the launch to the next power of two); a context's launches of a role that it is neither
 * tuning nor has set use the kept choice for their size class, when there is one. */
#define ICF_SHAPE_PHASE1 0
#define ICF_SHAPE_NARROW 1
#define ICF_SHAPE_CLOSEST 2
#define ICF_SHAPE_IMPACT 3
#define ICF_SHAPE_ORDER 4
int icf_set_shape(icf_ctx *ctx, int role, int value);
int icf_get_shape(icf_ctx *ctx, int role, int *value, int *tuning, float *best_ms);
int icf_profile_reset(icf_ctx *ctx);
int icf_profile_count(const icf_ctx *ctx);
int icf_profile_entry(const icf_ctx *ctx, int i, char *name, int cap, double *total_ms, long long *launches);

/* --- node state kept on the device between iterations ---
 * The reference keeps u, v, a, du per node on the host (v3Library/Geometry/Node.cs) and packs
 * and uploads the whole node buffer every iteration (TransferNodesToGPU,
 * GPU_Functionality.cs:307-340).  With these calls the state stays on the device in
 * st: double SoA 12 x st_stride per node (u 0-2, v 3-5, a 6-8, du 9-11), and only the solver's
 * dx crosses the bus.  Results are the reference's doubles (same expressions, no contraction). */
/* Node.InferTentativeValues (Node.cs:79-97) for every node, written into dn as
 * TransferNodesToGPU packs it: un, vn, an, x_current = x0 + un (x0 read from dn), F = 0. */
int icf_nodes_infer(icf_ctx *ctx, const double *st, int st_stride, double *dn, int nd_stride,
                    int nNodes, double h, double beta, double gamma);
/* ImplicitModel3._XtoDU (ImplicitModel3.cs:247-256): du += dx[3 altId + k] for altId >= 0;
 * dx is the solver's increment (3 per active row) in device memory. */
int icf_nodes_add_dx(icf_ctx *ctx, double *st, int st_stride, const int *altId, int nNodes,
                     const double *dx);
/* Node.AcceptTentativeValues (Node.cs:62-77) with un as last inferred into dn:
 * v' = (un - u)/h, a = (v' - v)/h, v = v', u = un, du = 0. */
int icf_nodes_accept(icf_ctx *ctx, double *st, int st_stride, const double *dn, int nd_stride,
                     int nNodes, double h);

/* --- BCSR pattern and pcsr packing on the device ---
 * The reference builds the pattern on the host (CSRDictionary.CreateStructure,
 * CSRDictionary.cs:64-97, from MeshCollection.UpdateStaticStructureData,
 * MeshCollection.cs:91-109, and ImplicitModel3._addCollidingNodesToStructure,
 * ImplicitModel3.cs:177-196) and packs every unit's slots and rows on the host before each
 * upload (TransferPCSR, GPU_Functionality.cs:345-416; collisionResponse, :627-665).
 * altId: per node (all nodes), -1 for anchored; node ids are read from the unit buffers
 * (ie fields 0-3, icz 4-9, icr 0-3).  All arrays are device memory. */
/* Rows 0..N-1; row r's columns are the sorted union of c >= r over pairs (altId_i, altId_j)
 * of non-anchored vertices of each element and CZ, and c > r over each impact's nodes.
 * The context keeps the elements' and CZs' part with a copy of their node ids and the altIds,
 * and reuses it on a later call whose node ids and altIds compare equal (only the impacts are
 * then sorted and merged in); icf_pattern_changed drops it.
 * Writes csr_rows[N+1], csr_cols[*nnz].  Returns ICF_E_CAPACITY (with *nnz set) if the
 * pattern has more than max_nnz blocks, ICF_E_INDEX for a node id outside [0, nNodes) or an
 * altId >= N.  Any unit count may be 0 (its pointer may then be NULL). */
int icf_build_pattern(icf_ctx *ctx, const int *altId, int nNodes,
                      const int *ie, int nElems, int el_stride,
                      const int *icz, int nCZs, int cz_stride,
                      const int *icr, int nImpacts, int cr_stride,
                      int N, int *csr_rows, int *csr_cols, int max_nnz, int *nnz);
/* Slot fields (position of (altId_i, altId_j) if both >= 0 and altId_j >= altId_i, else -1)
 * and row fields (altId_i or -1) of every element, CZ and impact, written in place; for CZs
 * also the current-failed, tentative-failed and tentative-damaged flags are reset to 0, as
 * TransferPCSR does.  ICF_E_INDEX if a needed pair is not in the pattern. */
int icf_pack_pcsr(icf_ctx *ctx, const int *altId, int nNodes, const int *csr_rows,
                  const int *csr_cols, int N, int *ie, int nElems, int el_stride,
                  int *icz, int nCZs, int cz_stride, int *icr, int nImpacts, int cr_stride);

/* --- broad phase and NL2set on the device ---
 * The reference finds candidate pairs with a host k-DOP BVH (BVHT.ConstructAndTraverse,
 * v3Library/Collisions/BVH/BVHT.cs:29-77; BVHN.cs:166-196; kDOP24.cs:84-140) and de-duplicates
 * the narrow phase's (node, tet) tuples in a host HashSet (GPU_Functionality.cs:565-600). */
/* Every pair (a < b) of the first nTets surface tets of ie whose 12-direction k-DOPs on
 * x_current (dn fields 15-17) overlap, sorted by (a, b), into pairs[2k], pairs[2k+1] (device).
 * exclude_shared != 0 drops pairs that share a node (the narrow phase zeroes them anyway,
 * narrowphase.h:99-103).  ICF_E_CAPACITY with *nOut = the full count if it exceeds
 * max_pairs. */
int icf_broad_phase(icf_ctx *ctx, const double *dn, int nd_stride, const int *ie,
                    int el_all_stride, int nTets, int exclude_shared, int *pairs,
                    long long max_pairs, long long *nOut);
/* From the narrow phase's masks (narrowList[2i] after icf_narrow_phase) and the candidate
 * pairs it was given (pairs[2i], pairs[2i+1]): the (node, tet) tuples -- vertex k of tet B
 * inside A for mask bit k < 4, vertex k-4 of A inside B for bit k >= 4 -- each kept once, in
 * the order of first occurrence, written to rows 0 (node) and 1 (tet) of nl2 (SoA,
 * tet_stride), ready for icf_find_closest_face.  ICF_E_CAPACITY (with *nOut set) if there are
 * more than tet_stride tuples. */
int icf_nl2set(icf_ctx *ctx, const int *masks, const int *pairs, int nPairs, const int *ie,
               int el_all_stride, int *nl2, int tet_stride, int *nOut);

#ifdef __cplusplus
}
#endif
#endif

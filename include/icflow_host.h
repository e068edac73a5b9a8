/*
 * icflow_host.h -- C ABI of the host-side helpers in libicflow_hip.so.
 *
 * These restate, in C++, the host logic that sits between icFlow's C# model and its GPU
 * kernels on the assembly path: the derived model constants (ModelPrms), the upper-
 * triangular 3x3 BCSR pattern and its (row, column) -> slot map (CSRDictionary), the packing
 * of elements, CZs and impacts into the reference's SoA pcsr arrays (TransferPCSR,
 * collisionResponse), and the (node, tet) de-duplication between the two narrow-phase kernels
 * (the NL2set HashSet).  A C# caller keeps its own versions of these; they exist so that the
 * library can be driven end to end from C, C++ or Python with the reference's data layout.
 * No device memory is touched here.
 */
#ifndef ICFLOW_HOST_H
#define ICFLOW_HOST_H

#include "icflow_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* User-facing parameters of ModelPrms (v3Library/Simulation/DataSheets/ModelPrms.cs:17-211). */
typedef struct icf_model_prms {
    double Y, nu, rho, dampingMass, dampingStiffness;
    double NewmarkBeta, NewmarkGamma, gravity;
    double alpha, beta, lambda_n, lambda_t, phi_n, phi_t, sigma_max, tau_max;
} icf_model_prms;

/* ModelPrms.SetComputedVariables (ModelPrms.cs:237-322) + Recompute (:79-116):
 * E, M, PPR derived constants, Gauss-point shape functions and the CZ B table. */
int icfh_compute_constants(const icf_model_prms *p, icf_consts *out);

/* --- CSRDictionary (v3Library/Simulation/LinearSystem/CSRDictionary.cs) --- */
typedef struct icfh_csr icfh_csr;
int icfh_csr_create(icfh_csr **out);
int icfh_csr_destroy(icfh_csr *c);
/* Static structure from elements and CZs (MeshCollection.UpdateStaticStructureData,
 * MeshCollection.cs:91-109) plus dynamic pairs from impacts
 * (ImplicitModel3._addCollidingNodesToStructure, ImplicitModel3.cs:177-196), then
 * CreateStructure (CSRDictionary.cs:64-97): rows = altId, sorted columns >= row.
 * alt_* arrays hold per-unit altIds (-1 = anchored), AoS: unit*4+i or unit*6+i. */
int icfh_csr_build(icfh_csr *c, const int *alt_elem, int nElems, const int *alt_cz, int nCZs,
                   const int *alt_imp, int nImpacts);
int icfh_csr_dims(const icfh_csr *c, int *N, int *nnz);
int icfh_csr_copy(const icfh_csr *c, int *rows /* N+1 */, int *cols /* nnz */);
int icfh_csr_lookup(const icfh_csr *c, int row, int col);   /* slot, or -1 */

/* --- pcsr packing (GPU_Functionality.TransferPCSR, GPU_Functionality.cs:345-416, and
 *     collisionResponse, :627-665) ---
 * nodes: global node ids, AoS (unit*k + i); alt: per-node altId (-1 = anchored).
 * Output arrays are the reference SoA buffers; fields the reference leaves alone are left
 * alone (ROWSIZE, flags beyond those it sets). */
int icfh_pack_elems(const icfh_csr *c, const int *nodes, const int *alt, int nElems, int el_stride, int *ie);
int icfh_pack_czs(const icfh_csr *c, const int *nodes, const int *alt, int nCZs, int cz_stride, int *icz);
int icfh_pack_impacts(const icfh_csr *c, const int *nl2, int tet_stride, const int *alt, int nImpacts,
                      int cr_stride, int *icr);

/* NL2set (GPU_Functionality.cs:565-600): from the narrow-phase masks (narrowList[2i]) and
 * the candidate tet pairs (pairs[2i], pairs[2i+1], surface-tet ids), the (node, tet) tuples
 * in first-insertion order; tet_nodes is the surface-tet SoA (node i of tet t at
 * tet_nodes[t + el_all_stride*i]).  Writes node ids to out[k] and tet ids to
 * out[k + out_stride]; *nOut = count.  Returns ICF_E_INDEX if more than max_out tuples. */
int icfh_nl2set(const int *masks, const int *pairs, int nPairs, const int *tet_nodes, int el_all_stride,
                int *out, int out_stride, int max_out, int *nOut);

/* Broad phase: every pair (a < b) of tets whose 12-direction k-DOPs on current positions
 * overlap -- the set BVHT.ConstructAndTraverse / BVHN.SelfCollide reports
 * (v3Library/Collisions/BVH/BVHN.cs:166-196, kDOP24.cs:84-140), found here with a uniform
 * grid instead of a tree.  dn: node SoA (x_current at field 15); ie: surface-tet SoA.
 * exclude_shared != 0 drops pairs that share a node (the narrow phase zeroes them anyway,
 * narrowphase.h:99-103).  pairs[2k], pairs[2k+1], sorted by (a, b).  If more than max_pairs
 * pairs exist, *nOut receives the full count and ICF_E_INDEX is returned. */
int icfh_kdop_pairs(const double *dn, int nd_stride, const int *ie, int el_all_stride, int nTets,
                    int exclude_shared, int *pairs, long long max_pairs, long long *nOut);

/* --- the linear solve (LinearSystem.Solve, v3Library/Simulation/LinearSystem/LinearSystem.cs:45-64) ---
 * SolveDouble3 takes the arguments of the reference's PardisoLoader2 export
 * (PardisoLoader2/dllmain.cpp:29-79) and is NOT MKL PARDISO: it solves A x = b for the
 * symmetric matrix A given by its upper-triangular BCSR with 3x3 blocks (ja: block column
 * indices, ia: n + 1 block row starts, zero-based; a: 9 doubles per block, row-major) with a
 * supernodal multifrontal LDL^T (nested-dissection or reverse Cuthill-McKee ordering of the
 * block graph, whichever fills less; 3x3 pivot blocks, small pivots perturbed by 1e-8 of the
 * largest entry followed by iterative refinement).  mtype must be -2
 * (symmetric indefinite, the reference's value) or 2, dim must be 3; iparam4 is ignored (the
 * reference passes 0); msglvl > 0 prints sizes and timings to stderr (msglvl > 1 also a progress
 * line per tenth of the factorisation); check != 0 verifies the
 * structure first.  Returns 0, or PARDISO's codes: -1 inconsistent input, -2 out of memory,
 * -3 ordering failed, -4 zero pivot.  x is overwritten; b is not changed.  Reentrant: threads
 * may solve different systems at once; the analyses of the last 1024 patterns are kept. */
int SolveDouble3(int *ja, int *ia, double *a, const int n, double *b, double *x, int mtype, int iparam4,
                 int dim, int msglvl, int check);

/* Threads SolveDouble3 uses inside one call for the dense updates of large fronts (no
 * counterpart in the reference, whose MKL reads MKL_NUM_THREADS): 0 = OpenMP's default, 1 =
 * none (for callers that solve several systems on threads of their own).  Returns 0, or -1
 * for a negative count. */
int icfh_solve_threads(int threads);

/* The ordering and symbolic analysis SolveDouble3 would do for the pattern (ja, ia, n) alone,
 * kept in the same cache, with the sizes it implies -- the figures PARDISO reports after its
 * analysis phase (phase 11: factor size, peak memory, factorisation flops), which the
 * reference's wrapper does not expose: doubles of L's panels, doubles at the factorisation's
 * peak (panels, update stack, largest front and panel), floating-point operations of the
 * factorisation.  Returns 0, -1 for bad arguments, -2 out of memory, -3 ordering failed. */
int icfh_solve_analyse(const int *ja, const int *ia, int n, long long *factor_doubles,
                       long long *peak_doubles, double *flops);

#ifdef __cplusplus
}
#endif
#endif

// icf_reduce.inc.h -- deterministic scatter-free assembly (included by icf_lib.hip).
//
// The reference scatters every unit's blocks with a CAS-loop atomicAdd2
// (elem_kernels9/atomic_add.h:1-13, AssembleBCSR12/18), so the summation order in each BCSR
// slot is arbitrary.  Here assembly is a gather:
//
//   slot map  (built when a pcsr array changes): for every key (BCSR slot, rhs row or node)
//             the sorted list of 32-bit codes (kind, unit, local index) that contribute to it.
//             Sorting by code orders contributions exactly like a sequential loop over
//             kinds (elements, CZs, contacts), units and the reference's local loops.
//   phase 2   one thread per key walks its list, rebuilds each contribution from the unit's
//             staged record and adds it into an FP64 accumulator, then writes the key once.
//
// Results are therefore bitwise reproducible and independent of launch order.
#pragma once

namespace icf {

// One pcsr-style key source: keys are base[u + stride * (field0 + q)], q < nq, u < n.
struct MapSrc {
    const int *base;
    int n, stride, field0, nq;
    uint32_t kind;
};

// Largest key of a source: grid-stride over the q-major enumeration (adjacent threads read
// adjacent units of one field), block maximum, one atomic per block.
__global__ __launch_bounds__(256) void k_map_maxkey(MapSrc s, int *__restrict__ maxkey)
{
    __shared__ int part[4];
    const long long tot = (long long)s.n * s.nq;
    int key = -1;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (long long)gridDim.x * blockDim.x) {
        const int q = (int)(i / s.n), u = (int)(i % s.n);
        key = max(key, s.base[u + (size_t)s.stride * (s.field0 + q)]);
    }
    for (int off = 32; off > 0; off >>= 1) key = max(key, __shfl_xor(key, off));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = key;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int m = max(max(part[0], part[1]), max(part[2], part[3]));
        if (m >= 0) atomicMax(maxkey, m);
    }
}

__global__ __launch_bounds__(256) void k_map_count(MapSrc s, int nkeys, int *__restrict__ count, int *__restrict__ err)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)s.n * s.nq) return;
    const int q = (int)(i / s.n), u = (int)(i % s.n);
    const int key = s.base[u + (size_t)s.stride * (s.field0 + q)];
    if (key < 0) return;
    if (key >= nkeys) {
        atomicOr(err, 1);
        return;
    }
    atomicAdd(&count[key], 1);
}

__global__ __launch_bounds__(256) void k_map_fill(MapSrc s, int nkeys, const int *__restrict__ off,
                                                  int *__restrict__ cursor, uint32_t *__restrict__ codes)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)s.n * s.nq) return;
    const int q = (int)(i / s.n), u = (int)(i % s.n);
    const int key = s.base[u + (size_t)s.stride * (s.field0 + q)];
    if (key < 0 || key >= nkeys) return;
    const int pos = off[key] + atomicAdd(&cursor[key], 1);
    codes[pos] = make_code(s.kind, (uint32_t)u, (uint32_t)q);
}

// Sort each key's codes ascending (insertion sort; lists are short: a BCSR slot receives
// at most a few tens of contributions).
__global__ __launch_bounds__(256) void k_map_sort(const int *__restrict__ off, uint32_t *__restrict__ codes, int nkeys)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nkeys) return;
    const int b = off[k], e = off[k + 1];
    for (int i = b + 1; i < e; i++) {
        const uint32_t v = codes[i];
        int j = i - 1;
        while (j >= b && codes[j] > v) {
            codes[j + 1] = codes[j];
            j--;
        }
        codes[j + 1] = v;
    }
}

// Copy of a source's key fields (q-major), and its comparison with the current fields.
__global__ __launch_bounds__(256) void k_snap_copy(MapSrc s, int *__restrict__ snap)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)s.n * s.nq) return;
    const int q = (int)(i / s.n), u = (int)(i % s.n);
    snap[i] = s.base[u + (size_t)s.stride * (s.field0 + q)];
}

__global__ __launch_bounds__(256) void k_snap_cmp(MapSrc s, const int *__restrict__ snap, int *__restrict__ diff)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)s.n * s.nq) return;
    const int q = (int)(i / s.n), u = (int)(i % s.n);
    if (snap[i] != s.base[u + (size_t)s.stride * (s.field0 + q)]) atomicOr(diff, 1);
}

// flag[k] = 1 if key k has contributions
__global__ __launch_bounds__(256) void k_map_flag(const int *__restrict__ cnt, int nkeys, int *__restrict__ flag)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nkeys) flag[k] = cnt[k] > 0;
}

// keys[pos[k]] = k for every key with contributions (pos = exclusive scan of the flags)
__global__ __launch_bounds__(256) void k_map_compact(const int *__restrict__ flag, const int *__restrict__ pos,
                                                     int nkeys, int *__restrict__ keys)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nkeys && flag[k]) keys[pos[k]] = k;
}

// Chunk shape of the staged BCSR-value reduction: keys per workgroup (one thread each; the
// workgroup size) and contributions staged in LDS, STAGE FP32 or STAGE/2 FP64 values x 9 per
// key.  STAGE = 2 is the least that still holds the chunk's output blocks (9 doubles per key)
// when they leave through the same LDS; a smaller buffer lets more workgroups share a CU, a
// larger one ends fewer chunks on the contribution limit (keys average about 5
// contributions).  A value map has a chunk table for every candidate shape and staging type,
// and the library's autotuner picks the shape per role (overwrite or accumulate map, FP32 or
// FP64 staging) by timing the candidates.  Every shape gives the same values: each key's
// contributions are added in list order whatever the chunking.
constexpr int NKC = 4, NSC = 2, NCAND = NKC * NSC;       // candidate key limits x stagings
constexpr int KEY_CANDS[NKC] = {64, 128, 256, 512};
constexpr int STAGE_CANDS[NSC] = {2, 3};
__host__ __device__ constexpr int cand_keys(int c) { return KEY_CANDS[c / NSC]; }
__host__ __device__ constexpr int cand_stage(int c) { return STAGE_CANDS[c % NSC]; }
constexpr int CAND_DEFAULT_FP32 = 2 * NSC + 0;           // 256 keys, 2 per key, until tuned
constexpr int CAND_DEFAULT_FP64 = 2 * NSC + 1;           // 256 keys, 1.5 per key
template <typename L, int KEYS, int STAGE>
__host__ __device__ constexpr int flat_cap() { return sizeof(L) == 4 ? STAGE * KEYS : STAGE * KEYS / 2; }
__host__ __device__ constexpr int flat_cap_of(bool fp32, int keys, int stage)
{
    return fp32 ? stage * keys : stage * keys / 2;
}

// Staged per-unit records of the current call.
struct Stage {
    const double *tet_rhs;  // node part i of tet u at (i * tplane + u) * 3
    const void *tet_tan;    // isotropic: node i's (g~, V) of tet u at record i * tplane + u
    int tplane;             // tet plane stride (the staged tet count, rounded up to 64)
    const double *cz_rhs;
    const void *cz_tan;
    const void *ct_tan;     // impact tangent blocks (float in mixed mode)
    const double *ct_f;     // impact forces
    double stiff, mcoef;   // tet LHS coefficients (element_kernels2.h:239-240)
};

__device__ inline const double *rhs_of(const Stage &sv, uint32_t code)
{
    const uint32_t kind = code_kind(code), u = code_unit(code), q = code_q(code);
    if (kind == KIND_TET) return sv.tet_rhs + ((size_t)q * sv.tplane + u) * 3;
    if (kind == KIND_CZ) return sv.cz_rhs + (size_t)u * CZ_RHS_REC + 3 * q;
    return sv.ct_f + (size_t)u * CT_F_REC + 3 * q;
}

// One contribution's 3x3 block, as the values that are added into the slot's FP64
// accumulator (FP32 values in mixed mode for tets and CZs).
template <bool MIXED, bool ISO, typename L>
__device__ inline void contrib9(const Stage &sv, uint32_t code, L v[9])
{
    using T = typename std::conditional<MIXED, float, double>::type;
    const uint32_t kind = code_kind(code), u = code_unit(code), q = code_q(code);
    const T stiff = (T)sv.stiff, mcoef = (T)sv.mcoef;
    if (kind == KIND_TET && ISO) {
        TetIsoLoad<T> l;
        l.load((const T *)sv.tet_tan, sv.tplane, code);
        T w[9];
        l.values(w, stiff, mcoef);
#pragma unroll
        for (int k = 0; k < 9; k++) v[k] = (L)w[k];
        return;
    }
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (kind == KIND_TET)
        tet_block_gen<T>(acc, (const T *)sv.tet_tan + (size_t)u * TET_GEN_REC, q >> 2, q & 3, stiff, mcoef);
    else if (kind == KIND_CZ)
        cz_block<T>(acc, (const T *)sv.cz_tan + (size_t)u * CZ_TAN_REC, q / 6, q % 6);
    else
        contact_block<T>(acc, (const T *)sv.ct_tan + (size_t)u * CT_TAN_REC, q >> 2, q & 3);
#pragma unroll
    for (int k = 0; k < 9; k++) v[k] = (L)acc[k];
}

// Phase 2 for BCSR values, one workgroup per chunk of consecutive keys (at most KEYS keys and
// flat_cap<L, KEYS>() contributions; the chunk table, built with the map, holds each
// chunk's first list and code positions).  The chunk's contributions are contiguous in
// `codes`: first every thread evaluates contributions
// (uniform work across lanes) into LDS, then one thread per key adds its contributions from
// LDS in list order into an FP64 accumulator and writes the block.  The values and their
// order are those of a sequential loop over the key's list, so the result does not depend
// on the chunking.
//   keys   null: chunk positions are keys; else keys[position] (accumulate mode)
//   ACC    add to A (single-kernel calls); otherwise overwrite, zero where nothing contributes
// L is the staging type: float for tet and CZ blocks in mixed mode, double otherwise.
template <bool MIXED, bool ISO, bool ACC, typename L, int KEYS, int STAGE>
__device__ __forceinline__ void reduce_vals(int bid_, int nb_, const int *__restrict__ off,
        const uint32_t *__restrict__ codes, const int *__restrict__ keys, const int2 *__restrict__ chunks, int nchunks,
        double *__restrict__ A, Stage sv)
{
    constexpr int CAP = flat_cap<L, KEYS, STAGE>();
    static_assert(CAP * 9 * sizeof(L) >= KEYS * 9 * sizeof(double), "output staging fits the LDS");
    constexpr int PER = (CAP + KEYS - 1) / KEYS;
    __shared__ L lds[CAP * 9];
    const int ch = xcd_remap(bid_, nb_);
    if (ch >= nchunks) return;
    const int2 c0 = chunks[ch], c1 = chunks[ch + 1];
    const int p0 = c0.x, p1 = c1.x, cb = c0.y, ce = c1.y;
    const int nc = ce - cb;
    const bool staged = nc <= CAP;
    // this thread's key and its code range, loaded beside the chunk's codes
    const int p = p0 + threadIdx.x;
    const bool live = p < p1;
    int s = 0, b = 0, e = 0;
    double acc[9];
#pragma unroll
    for (int i = 0; i < 9; i++) acc[i] = 0.;
    if (live) {
        s = keys ? keys[p] : p;
        b = off[s];
        e = off[s + 1];
        if (ACC) {   // the block's current values, also loaded before the contributions
            const double *src = A + (size_t)s * 9;
#pragma unroll
            for (int i = 0; i < 9; i++) acc[i] = src[i];
        }
    }
    if (staged) {
        // codes first (coalesced), then the contributions, so their loads overlap
        uint32_t cd[PER];
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const int i = threadIdx.x + r * KEYS;
            cd[r] = i < nc ? codes[cb + i] : 0u;
        }
#pragma unroll
        for (int r = 0; r < PER; r++) {
            const int i = threadIdx.x + r * KEYS;
            if (i < nc) {
                L v[9];
                contrib9<MIXED, ISO, L>(sv, cd[r], v);
#pragma unroll
                for (int k = 0; k < 9; k++) lds[i * 9 + k] = v[k];
            }
        }
        __syncthreads();
    }
    if (live) {
        if (staged) {
            for (int c = b - cb; c < e - cb; c++) {
#pragma unroll
                for (int k = 0; k < 9; k++) acc[k] += (double)lds[c * 9 + k];
            }
        } else {
            for (int c = b; c < e; c++) {
                L v[9];
                contrib9<MIXED, ISO, L>(sv, codes[c], v);
#pragma unroll
                for (int k = 0; k < 9; k++) acc[k] += (double)v[k];
            }
        }
    }
    if (ACC) {
        if (live) {
            double *dst = A + (size_t)s * 9;
#pragma unroll
            for (int i = 0; i < 9; i++) dst[i] = acc[i];
        }
        return;
    }
    // overwrite mode: the chunk's blocks A[p0 .. p1) are contiguous; write them through LDS
    // as consecutive doubles (coalesced) instead of one 72-byte block per thread, as
    // nontemporal stores: A is written once and not read again here, so it need not take L2
    // lines from the staged records the other chunks still gather
    double *out = reinterpret_cast<double *>(lds);
    __syncthreads();
    if (live) {
#pragma unroll
        for (int k = 0; k < 9; k++) out[threadIdx.x * 9 + k] = acc[k];
    }
    __syncthreads();
    const int nout = (p1 - p0) * 9;
    double *dst = A + (size_t)p0 * 9;
    for (int i = threadIdx.x; i < nout; i += KEYS) __builtin_nontemporal_store(out[i], &dst[i]);
}

template <bool MIXED, bool ISO, bool ACC, typename L, int KEYS, int STAGE>
__global__ __launch_bounds__(KEYS) void k_reduce_vals(const int *__restrict__ off, const uint32_t *__restrict__ codes,
                                                           const int *__restrict__ keys, const int2 *__restrict__ chunks,
                                                           int nchunks, double *__restrict__ A, Stage sv)
{
    reduce_vals<MIXED, ISO, ACC, L, KEYS, STAGE>(blockIdx.x, gridDim.x, off, codes, keys, chunks, nchunks, A, sv);
}

// Phase 2 for rhs rows (b[row*3+k]) or node forces (dn[node + ns*(F+k)]): one thread per key.
// TSUM (node forces of the fused call): the sum of the key's contributions from zero is also
// written to tsum[3 key .. 3 key + 2], from which k_reduce_rhs_tsum starts the rhs rows.
template <bool ACC, bool TSUM>
__device__ __forceinline__ void reduce_vec3(int bid_, int nb_, const int *__restrict__ off,
        const uint32_t *__restrict__ codes, const int *__restrict__ keys, int nkeys, double *__restrict__ out,
        int comp_stride, int elem_stride, Stage sv, double *__restrict__ tsum)
{
    const int t = xcd_remap(bid_, nb_) * blockDim.x + threadIdx.x;
    if (t >= nkeys) return;
    const int s = keys ? keys[t] : t;
    const int b = off[s], e = off[s + 1];
    double *dst = out + (size_t)s * elem_stride;
    double a0, a1, a2, t0 = 0., t1 = 0., t2 = 0.;
    if (ACC) {
        if (b == e) return;
        a0 = dst[0]; a1 = dst[comp_stride]; a2 = dst[2 * comp_stride];
    } else {
        a0 = a1 = a2 = 0.;
    }
    auto part = [&](int c) { return rhs_of(sv, codes[c]); };
    int c = b;
    for (; c + 4 <= e; c += 4) {          // loads of four contributions in flight, added in order
        const double *r0 = part(c), *r1 = part(c + 1);
        const double *r2 = part(c + 2), *r3 = part(c + 3);
        const double x0 = r0[0], y0 = r0[1], z0 = r0[2], x1 = r1[0], y1 = r1[1], z1 = r1[2];
        const double x2 = r2[0], y2 = r2[1], z2 = r2[2], x3 = r3[0], y3 = r3[1], z3 = r3[2];
        a0 += x0; a1 += y0; a2 += z0;
        a0 += x1; a1 += y1; a2 += z1;
        a0 += x2; a1 += y2; a2 += z2;
        a0 += x3; a1 += y3; a2 += z3;
        if (TSUM) {
            t0 += x0; t1 += y0; t2 += z0;
            t0 += x1; t1 += y1; t2 += z1;
            t0 += x2; t1 += y2; t2 += z2;
            t0 += x3; t1 += y3; t2 += z3;
        }
    }
    for (; c < e; c++) {
        const double *r = part(c);
        a0 += r[0]; a1 += r[1]; a2 += r[2];
        if (TSUM) { t0 += r[0]; t1 += r[1]; t2 += r[2]; }
    }
    dst[0] = a0; dst[comp_stride] = a1; dst[2 * comp_stride] = a2;
    if (TSUM) {
        double *o = tsum + (size_t)s * 3;
        o[0] = t0; o[1] = t1; o[2] = t2;
    }
}

// Node forces and rhs rows of the fused call in one pass, one thread per node key: F[node] gets
// the node's tet contributions added in list order (as reduce_vec3<true, TSUM>), and the node's
// rhs row -- node_row[node], when not negative -- is the sum of the same contributions from zero
// followed by the row's CZ contributions in list order (as reduce_rhs_tsum from that tet sum).
// Used when every rhs row's tet contributions are its node's force list and every row has such
// a node (the host checks both once per pair of maps).
__device__ __forceinline__ void reduce_force_rhs(int bid_, int nb_, const int *__restrict__ foff,
        const uint32_t *__restrict__ fcodes, const int *__restrict__ fkeys, int nfkeys, double *__restrict__ F,
        int ns, const int *__restrict__ node_row, const int *__restrict__ roff, const uint32_t *__restrict__ rcodes,
        const int *__restrict__ cz_begin, double *__restrict__ rhs, Stage sv)
{
    const int t = xcd_remap(bid_, nb_) * blockDim.x + threadIdx.x;
    if (t >= nfkeys) return;
    const int s = fkeys ? fkeys[t] : t;
    const int b = foff[s], e = foff[s + 1];
    const int r = node_row[s];
    double *dst = F + s;
    double a0 = dst[0], a1 = dst[ns], a2 = dst[2 * (size_t)ns], t0 = 0., t1 = 0., t2 = 0.;
    auto part = [&](int c) { return rhs_of(sv, fcodes[c]); };
    int c = b;
    for (; c + 4 <= e; c += 4) {          // loads of four contributions in flight, added in order
        const double *r0 = part(c), *r1 = part(c + 1);
        const double *r2 = part(c + 2), *r3 = part(c + 3);
        const double x0 = r0[0], y0 = r0[1], z0 = r0[2], x1 = r1[0], y1 = r1[1], z1 = r1[2];
        const double x2 = r2[0], y2 = r2[1], z2 = r2[2], x3 = r3[0], y3 = r3[1], z3 = r3[2];
        a0 += x0; a1 += y0; a2 += z0;
        a0 += x1; a1 += y1; a2 += z1;
        a0 += x2; a1 += y2; a2 += z2;
        a0 += x3; a1 += y3; a2 += z3;
        t0 += x0; t1 += y0; t2 += z0;
        t0 += x1; t1 += y1; t2 += z1;
        t0 += x2; t1 += y2; t2 += z2;
        t0 += x3; t1 += y3; t2 += z3;
    }
    for (; c < e; c++) {
        const double *q = part(c);
        a0 += q[0]; a1 += q[1]; a2 += q[2];
        t0 += q[0]; t1 += q[1]; t2 += q[2];
    }
    dst[0] = a0; dst[ns] = a1; dst[2 * (size_t)ns] = a2;
    if (r < 0) return;
    const int re = roff[r + 1];
    for (int k = cz_begin[r]; k < re; k++) {
        const double *q = rhs_of(sv, rcodes[k]);
        t0 += q[0]; t1 += q[1]; t2 += q[2];
    }
    double *o = rhs + (size_t)r * 3;
    o[0] = t0; o[1] = t1; o[2] = t2;
}

template <bool ACC, bool TSUM = false>
__global__ __launch_bounds__(256) void k_reduce_vec3(const int *__restrict__ off, const uint32_t *__restrict__ codes,
                                                     const int *__restrict__ keys, int nkeys, double *__restrict__ out,
                                                     int comp_stride, int elem_stride, Stage sv,
                                                     double *__restrict__ tsum = nullptr)
{
    reduce_vec3<ACC, TSUM>(blockIdx.x, gridDim.x, off, codes, keys, nkeys, out, comp_stride, elem_stride, sv, tsum);
}

// For each rhs row r of the fused call: the length of its leading run of tet codes and the
// node those tets share.  The row's tet contributions must be exactly the force key's list
// of that node (same codes, same order); any other case sets *err.
__global__ __launch_bounds__(256) void k_rhs_check(const int *__restrict__ roff, const uint32_t *__restrict__ rcodes,
                                                   int N, const int *__restrict__ foff,
                                                   const uint32_t *__restrict__ fcodes, int nfkeys, MapSrc tet,
                                                   int *__restrict__ row_node, int *__restrict__ cz_begin,
                                                   int *__restrict__ err)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    const int b = roff[r], e = roff[r + 1];
    int L = 0;
    while (b + L < e && code_kind(rcodes[b + L]) == KIND_TET) L++;
    cz_begin[r] = b + L;
    if (L == 0) {
        row_node[r] = -1;
        return;
    }
    const uint32_t c0 = rcodes[b];
    const int u = (int)code_unit(c0), q = (int)code_q(c0);
    const int n = (u < tet.n && q < 4) ? tet.base[u + (size_t)tet.stride * (tet.field0 + q)] : -1;
    if (n < 0 || n >= nfkeys || foff[n + 1] - foff[n] != L) {
        atomicOr(err, 1);
        return;
    }
    const int fb = foff[n];
    for (int k = 0; k < L; k++)
        if (fcodes[fb + k] != rcodes[b + k]) {
            atomicOr(err, 1);
            return;
        }
    row_node[r] = n;
}

// node_row[row_node[r]] = r for every rhs row r (node_row filled with -1 before).  A row
// without a tet node, or a node with two rows, sets *err.
__global__ __launch_bounds__(256) void k_node_row(const int *__restrict__ row_node, int N, int *__restrict__ node_row,
                                                  int *__restrict__ err)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    const int n = row_node[r];
    if (n < 0 || atomicExch(node_row + n, r) != -1) atomicOr(err, 1);
}

// rhs rows of the fused call: the node's tet sum (k_reduce_vec3 TSUM), then the row's CZ
// contributions in order -- the sequential sum over the row's list.
__device__ __forceinline__ void reduce_rhs_tsum(int bid_, int nb_, const int *__restrict__ off,
        const uint32_t *__restrict__ codes, const int *__restrict__ cz_begin, const int *__restrict__ row_node,
        const double *__restrict__ tsum, int N, double *__restrict__ out, Stage sv)
{
    const int r = xcd_remap(bid_, nb_) * blockDim.x + threadIdx.x;
    if (r >= N) return;
    const int n = row_node[r];
    double a0 = 0., a1 = 0., a2 = 0.;
    if (n >= 0) {
        a0 = tsum[(size_t)n * 3];
        a1 = tsum[(size_t)n * 3 + 1];
        a2 = tsum[(size_t)n * 3 + 2];
    }
    const int e = off[r + 1];
    for (int c = cz_begin[r]; c < e; c++) {
        const double *p = rhs_of(sv, codes[c]);
        a0 += p[0]; a1 += p[1]; a2 += p[2];
    }
    double *o = out + (size_t)r * 3;
    o[0] = a0; o[1] = a1; o[2] = a2;
}

__global__ __launch_bounds__(256) void k_reduce_rhs_tsum(const int *__restrict__ off, const uint32_t *__restrict__ codes,
                                                         const int *__restrict__ cz_begin, const int *__restrict__ row_node,
                                                         const double *__restrict__ tsum, int N, double *__restrict__ out,
                                                         Stage sv)
{
    reduce_rhs_tsum(blockIdx.x, gridDim.x, off, codes, cz_begin, row_node, tsum, N, out, sv);
}

// Byte copy, 16 bytes per thread where both ends allow it (icf_d2d; a kernel so that it can
// be recorded in a captured graph).
__global__ __launch_bounds__(256) void k_copy(char *__restrict__ dst, const char *__restrict__ src, size_t bytes)
{
    const size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 16;
    if (i >= bytes) return;
    if (i + 16 <= bytes && (((uintptr_t)(dst + i) | (uintptr_t)(src + i)) & 15) == 0) {
        *(uint4 *)(dst + i) = *(const uint4 *)(src + i);
    } else {
        const size_t e = i + 16 < bytes ? i + 16 : bytes;
        for (size_t k = i; k < e; k++) dst[k] = src[k];
    }
}

// VALU issue-rate probe (icf_probe_valu): every thread runs `iters` rounds of 16 independent
// fused multiply-add chains x = x * m + c (one VALU instruction each, no memory traffic in the
// loop) and stores the sum of its chains, so that no chain can be dropped.  With enough waves
// per SIMD the loop is bound by VALU issue alone: its rate is the achievable FMA peak.
template <typename T>
__global__ __launch_bounds__(256) void k_valu_probe(T *__restrict__ out, int iters, T m, T c)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    T x[16];
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = (T)(tid & 7) * (T)0.125 + (T)k;
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int k = 0; k < 16; k++) x[k] = fma(x[k], m, c);
    }
    T s = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) s += x[k];
    out[tid] = s;
}

__global__ void k_fill_int(int *__restrict__ p, int n, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Block-wide integer sum (256 threads), result valid in thread 0.
__device__ inline int block_sum256(int x)
{
    __shared__ int part[4];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = x;
    __syncthreads();
    int r = 0;
    if (threadIdx.x == 0) r = part[0] + part[1] + part[2] + part[3];
    __syncthreads();
    return r;
}

// Integer sum (GPU_Functionality.Sum, :93-107): grid-stride, one atomic per block.
__global__ __launch_bounds__(256) void k_sum_int(const int *__restrict__ v, int n, int *__restrict__ out)
{
    int x = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x += v[i];
    x = block_sum256(x);
    if (threadIdx.x == 0 && x != 0) atomicAdd(out, x);
}

// The damaged and failed counts of one CZ pass in one sweep (out[0], out[1]).
__device__ __forceinline__ void count_flags(int bid_, int nb_, const int *__restrict__ damaged,
        const int *__restrict__ failed, int n, int *__restrict__ out)
{
    int d = 0, f = 0;
    for (int i = bid_ * blockDim.x + threadIdx.x; i < n; i += nb_ * blockDim.x) {
        d += damaged[i];
        f += failed[i];
    }
    d = block_sum256(d);
    f = block_sum256(f);
    if (threadIdx.x == 0) {
        if (d) atomicAdd(out, d);
        if (f) atomicAdd(out + 1, f);
    }
}

// count_flags for workgroups of NT threads (a multiple of 64, at most 1024).
template <int NT>
__device__ __forceinline__ void count_flags_n(int bid_, int nb_, const int *__restrict__ damaged,
        const int *__restrict__ failed, int n, int *__restrict__ out)
{
    __shared__ int part[2][NT / 64];
    int d = 0, f = 0;
    for (int i = bid_ * NT + threadIdx.x; i < n; i += nb_ * NT) {
        d += damaged[i];
        f += failed[i];
    }
    for (int o = 32; o > 0; o >>= 1) {
        d += __shfl_xor(d, o);
        f += __shfl_xor(f, o);
    }
    if ((threadIdx.x & 63) == 0) {
        part[0][threadIdx.x >> 6] = d;
        part[1][threadIdx.x >> 6] = f;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        d = f = 0;
        for (int w = 0; w < NT / 64; w++) {
            d += part[0][w];
            f += part[1][w];
        }
        if (d) atomicAdd(out, d);
        if (f) atomicAdd(out + 1, f);
    }
}

// Per-specimen counters: workgroup bid_ takes specimens bid_, bid_ + nb_, ...; each
// specimen's damaged and failed counts are written (no atomics) and added into the totals.
template <int NT>
__device__ __forceinline__ void count_flags_spec(int bid_, int nb_, const int *__restrict__ damaged,
        const int *__restrict__ failed, const int *__restrict__ cz_off, int nspec, int *__restrict__ spec,
        int *__restrict__ out)
{
    __shared__ int part[2][NT / 64];
    int td = 0, tf = 0;
    for (int s = bid_; s < nspec; s += nb_) {
        int d = 0, f = 0;
        for (int i = cz_off[s] + (int)threadIdx.x; i < cz_off[s + 1]; i += NT) {
            d += damaged[i];
            f += failed[i];
        }
        for (int o = 32; o > 0; o >>= 1) {
            d += __shfl_xor(d, o);
            f += __shfl_xor(f, o);
        }
        if ((threadIdx.x & 63) == 0) {
            part[0][threadIdx.x >> 6] = d;
            part[1][threadIdx.x >> 6] = f;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            d = f = 0;
            for (int w = 0; w < NT / 64; w++) {
                d += part[0][w];
                f += part[1][w];
            }
            spec[2 * s] = d;
            spec[2 * s + 1] = f;
            td += d;
            tf += f;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (td) atomicAdd(out, td);
        if (tf) atomicAdd(out + 1, tf);
    }
}

__global__ __launch_bounds__(256) void k_count_flags(const int *__restrict__ damaged, const int *__restrict__ failed,
                                                     int n, int *__restrict__ out)
{
    count_flags(blockIdx.x, gridDim.x, damaged, failed, n, out);
}

}  // namespace icf

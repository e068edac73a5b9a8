// icf_pattern.inc.h -- BCSR pattern and pcsr slot packing on the device (included by
// icf_lib.hip).
//
// Restates, for device-resident unit buffers:
//  - the static structure of MeshCollection.UpdateStaticStructureData
//    (v3Library/Geometry/Mesh/MeshCollection.cs:91-109): for every CZ and element, every pair
//    of non-anchored vertices with altId_2 >= altId_1;
//  - the dynamic structure of ImplicitModel3._addCollidingNodesToStructure
//    (ImplicitModel3.cs:177-196): every pair of an impact's non-anchored nodes with
//    altId_2 > altId_1;
//  - CSRDictionary.CreateStructure (CSRDictionary.cs:64-97): rows 0..N-1, each row's columns
//    the sorted union of its static and dynamic neighbours, slot = position;
//  - the slot / row packing of GPU_Functionality.TransferPCSR (GPU_Functionality.cs:345-416)
//    and collisionResponse (:627-665): slot(i,j) = position of (altId_i, altId_j) when both
//    nodes are non-anchored and altId_j >= altId_i, else -1; row_i = altId_i or -1.
// Pairs are emitted as 64-bit keys (row << 32 | column), sorted and de-duplicated with
// hipCUB, which gives the same rows and sorted columns as the dictionary.
#pragma once

namespace icf {

// Keys (row << cb | column) with cb bits per index, (1 << cb) > N; the largest 2cb-bit value
// marks a pair that adds no block and sorts after every block.
__host__ __device__ inline unsigned long long pat_none(int cb) { return (1ULL << (2 * cb)) - 1; }

// One unit source of the pattern: node ids base[u + stride * (field0 + i)], i < K.
struct PatSrc {
    const int *base;
    int n, stride, field0, K;
    int strict;            // 1: only altId_j > altId_i (impacts)
    long long out0;        // first output position of this source
};

// Number of candidate pairs per unit: K(K+1)/2, or K(K-1)/2 when strict.
__host__ __device__ inline int pat_pairs(int K, int strict) { return strict ? K * (K - 1) / 2 : K * (K + 1) / 2; }

__global__ __launch_bounds__(256) void k_pat_emit(PatSrc s, const int *__restrict__ alt, int nNodes, int N,
                                                  int cb, unsigned long long *__restrict__ keys,
                                                  int *__restrict__ err)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int P = pat_pairs(s.K, s.strict);
    if (t >= (long long)s.n * P) return;
    const int u = (int)(t / P);
    int p = (int)(t % P);
    // local pair (i, j), i <= j (i < j when strict), in row-major order
    int i = 0, row_len = s.strict ? s.K - 1 : s.K;
    while (p >= row_len) {
        p -= row_len;
        i++;
        row_len--;
    }
    const int j = s.strict ? i + 1 + p : i + p;
    const int ni = s.base[u + (size_t)s.stride * (s.field0 + i)];
    const int nj = s.base[u + (size_t)s.stride * (s.field0 + j)];
    unsigned long long key = pat_none(cb);
    if (ni < 0 || ni >= nNodes || nj < 0 || nj >= nNodes) {
        atomicOr(err, 1);
    } else {
        const int a = alt[ni], b = alt[nj];
        if (a >= 0 && b >= 0 && !(s.strict && a == b)) {
            const int r = a < b ? a : b, c = a < b ? b : a;
            if (c >= N) atomicOr(err, 1);
            else key = ((unsigned long long)r << cb) | (unsigned)c;
        }
    }
    keys[s.out0 + t] = key;
}

// The node-id fields a pattern's static part is built from, as one sequence: the altId of
// every node, then the 4 node ids of each element (field by field), then the 6 of each CZ.
struct PatFields {
    const int *alt;
    int nNodes;
    const int *ie;
    int nElems, el_stride;
    const int *icz;
    int nCZs, cz_stride;
};
__device__ inline int pat_field(const PatFields &f, long long i)
{
    if (i < f.nNodes) return f.alt[i];
    i -= f.nNodes;
    if (i < 4LL * f.nElems) return f.ie[i % f.nElems + (size_t)f.el_stride * (EL_NODES + i / f.nElems)];
    i -= 4LL * f.nElems;
    return f.icz[i % f.nCZs + (size_t)f.cz_stride * (CZI_VRTS + i / f.nCZs)];
}
// compare == 0: snap = the fields; else *diff = 1 if any field differs from snap.
__global__ __launch_bounds__(256) void k_pat_snap(PatFields f, long long n, int *__restrict__ snap, int compare,
                                                  int *__restrict__ diff)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = pat_field(f, i);
    if (!compare) snap[i] = v;
    else if (snap[i] != v) atomicOr(diff, 1);
}

// The sorted keys' distinct blocks: flag[i] = 1 where key i differs from key i-1 and is a
// block (flag[M] = 0 is set by the caller); then each flagged key goes to its scan position.
__global__ __launch_bounds__(256) void k_pat_flag(const unsigned long long *__restrict__ keys, int M, int cb,
                                                  int *__restrict__ flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const unsigned long long k = keys[i];
    flag[i] = (k != pat_none(cb) && (i == 0 || keys[i - 1] != k)) ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_pat_compact(const unsigned long long *__restrict__ keys, int M,
                                                     const int *__restrict__ flag, const int *__restrict__ pos,
                                                     unsigned long long *__restrict__ ukeys)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M && flag[i]) ukeys[pos[i]] = keys[i];
}

// csr_cols[k] = column of unique key k (k < nnz); csr_rows[r] = first k with row >= r.
__global__ __launch_bounds__(256) void k_pat_cols(const unsigned long long *__restrict__ ukeys, int nnz, int cb,
                                                  int *__restrict__ cols)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz) cols[k] = (int)(ukeys[k] & ((1ULL << cb) - 1));
}

__global__ __launch_bounds__(256) void k_pat_rows(const unsigned long long *__restrict__ ukeys, int nnz, int N, int cb,
                                                  int *__restrict__ rows)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > N) return;
    int lo = 0, hi = nnz;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((int)(ukeys[mid] >> cb) < r) lo = mid + 1;
        else hi = mid;
    }
    rows[r] = lo;
}

__device__ inline int pat_lookup(const int *__restrict__ rows, const int *__restrict__ cols, int r, int c)
{
    int lo = rows[r], hi = rows[r + 1];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int v = cols[mid];
        if (v == c) return mid;
        if (v < c) lo = mid + 1;
        else hi = mid;
    }
    return -1;
}

// Slot and row fields of one unit kind; one thread per (unit, local row i).
struct PackSrc {
    int *base;
    int n, stride, node0, slot0, row0, K;
};

__global__ __launch_bounds__(256) void k_pack_pcsr(PackSrc s, const int *__restrict__ alt, int nNodes,
                                                   const int *__restrict__ rows, const int *__restrict__ cols, int N,
                                                   int *__restrict__ err)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)s.n * s.K) return;
    const int i = (int)(t / s.n), u = (int)(t % s.n);     // i-major: adjacent threads, adjacent units
    const int ni = s.base[u + (size_t)s.stride * (s.node0 + i)];
    if (ni < 0 || ni >= nNodes) {
        atomicOr(err, 1);
        return;
    }
    const int ai = alt[ni];
    if (ai >= N) {
        atomicOr(err, 1);
        return;
    }
    s.base[u + (size_t)s.stride * (s.row0 + i)] = ai >= 0 ? ai : -1;
    for (int j = 0; j < s.K; j++) {
        const int nj = s.base[u + (size_t)s.stride * (s.node0 + j)];
        int slot = -1;
        if (nj < 0 || nj >= nNodes) {
            atomicOr(err, 1);
        } else {
            const int aj = alt[nj];
            if (ai >= 0 && aj >= 0 && aj >= ai) {
                slot = aj < N ? pat_lookup(rows, cols, ai, aj) : -1;
                if (slot < 0) atomicOr(err, 1);
            }
        }
        s.base[u + (size_t)s.stride * (s.slot0 + i * s.K + j)] = slot;
    }
}

// CZ flags that TransferPCSR resets (GPU_Functionality.cs:384-386): current failed,
// tentative failed, tentative damaged.
__global__ __launch_bounds__(256) void k_pack_cz_flags(int *__restrict__ icz, int n, int cs)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    icz[u + (size_t)cs * CZI_FAILED] = 0;
    icz[u + (size_t)cs * CZI_TFAILED] = 0;
    icz[u + (size_t)cs * CZI_TDAMAGED] = 0;
}

}  // namespace icf

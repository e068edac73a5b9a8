// icf_broad.inc.h -- broad phase and (node, tet) de-duplication on the device (included by
// icf_lib.hip).
//
// Broad phase: every pair (a < b) of surface tets whose 12-direction k-DOPs on the current
// positions overlap -- the set the reference's BVH reports (v3Library/Collisions/BVH/
// BVHN.cs:166-196 with kDOP24.cs:84-140), which is the set icfh_kdop_pairs finds on the host.
// The k-DOP extents are sums and differences of coordinates and the overlap test is a
// comparison, so both produce the same pairs; the order is (a, b) ascending.  The search is
// a uniform grid of cell size = the largest AABB extent: tets are binned by the cell of
// their AABB's minimum corner, so overlapping tets lie in neighbouring cells.
//
// NL2set: the (node, tet) tuples of GPU_Functionality.NarrowPhaseCollisionDetection
// (GPU_Functionality.cs:565-600), de-duplicated keeping each tuple's first occurrence, in
// the order of first occurrence (a HashSet without removals enumerates in insertion order).
#pragma once

namespace icf {

constexpr int BP_DIRS = 12;

__global__ __launch_bounds__(256) void k_bp_kdop(const double *__restrict__ dn, int ns, const int *__restrict__ ie,
                                                 int es, int nT, double *__restrict__ lo, double *__restrict__ hi,
                                                 double *__restrict__ ext, double *__restrict__ mn,
                                                 int *__restrict__ err, int nNodesBound)
{
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    double l[BP_DIRS], h[BP_DIRS];
#pragma unroll
    for (int d = 0; d < BP_DIRS; d++) {
        l[d] = 1e300;
        h[d] = -1e300;
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int n = ie[t + (size_t)es * v];
        if (n < 0 || (nNodesBound > 0 && n >= nNodesBound)) {
            atomicOr(err, 1);
            return;
        }
        const double x = dn[n + (size_t)ns * ND_XC], y = dn[n + (size_t)ns * (ND_XC + 1)],
                     z = dn[n + (size_t)ns * (ND_XC + 2)];
        const double p[BP_DIRS] = {x, y, z, x + y, x + z, y + z, x - y, x - z, y - z, x + y - z, x + z - y, y + z - x};
#pragma unroll
        for (int d = 0; d < BP_DIRS; d++) {
            l[d] = fmin(l[d], p[d]);
            h[d] = fmax(h[d], p[d]);
        }
    }
#pragma unroll
    for (int d = 0; d < BP_DIRS; d++) {
        lo[(size_t)d * nT + t] = l[d];
        hi[(size_t)d * nT + t] = h[d];
    }
    ext[t] = fmax(h[0] - l[0], fmax(h[1] - l[1], h[2] - l[2]));
#pragma unroll
    for (int k = 0; k < 3; k++) mn[(size_t)k * nT + t] = l[k];
}

// red = {largest extent, min x, min y, min z}
__device__ inline double bp_cell_size(const double *red) { return red[0] > 0 ? red[0] * 1.000001 : 1.0; }

constexpr int BP_CELL_BITS = 21;

__global__ __launch_bounds__(256) void k_bp_cell(const double *__restrict__ lo, int nT, const double *__restrict__ red,
                                                 unsigned long long *__restrict__ key, int *__restrict__ idx,
                                                 int *__restrict__ err)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    const double cs = bp_cell_size(red);
    unsigned long long k = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double c = floor((lo[(size_t)a * nT + t] - red[1 + a]) / cs);
        if (!(c >= 0 && c < (double)((1 << BP_CELL_BITS) - 2))) {
            atomicOr(err, 1);
            return;
        }
        k |= (unsigned long long)c << (BP_CELL_BITS * a);
    }
    key[t] = k;
    idx[t] = t;
}

// The k-DOPs and node ids in cell order: sorted position j holds tet sidx[j], so that the tets
// of one cell, and of neighbouring cells, are adjacent in memory for k_bp_pairs9.  Per position:
// the k-DOP as (lo, hi) pairs per direction (skd, 24 doubles), the node ids (snode), and the
// axis-aligned box in FP32 rounded outward, lo down and hi up (box): a prefilter that never
// rejects a pair whose FP64 boxes overlap.
__global__ __launch_bounds__(256) void k_bp_sorted(const double *__restrict__ lo, const double *__restrict__ hi,
                                                   int nT, const int *__restrict__ sidx, const int *__restrict__ ie,
                                                   int es, double2 *__restrict__ skd, int4 *__restrict__ snode,
                                                   float4 *__restrict__ box)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nT) return;
    const int t = sidx[j];
    double l[BP_DIRS], h[BP_DIRS];
#pragma unroll
    for (int d = 0; d < BP_DIRS; d++) {
        l[d] = lo[(size_t)d * nT + t];
        h[d] = hi[(size_t)d * nT + t];
        skd[(size_t)j * BP_DIRS + d] = make_double2(l[d], h[d]);
    }
    snode[j] = make_int4(ie[t], ie[t + (size_t)es], ie[t + 2 * (size_t)es], ie[t + 3 * (size_t)es]);
    box[2 * (size_t)j] = make_float4(__double2float_rd(l[0]), __double2float_rd(l[1]), __double2float_rd(l[2]), 0.f);
    box[2 * (size_t)j + 1] = make_float4(__double2float_ru(h[0]), __double2float_ru(h[1]), __double2float_ru(h[2]), 0.f);
}

// Visit each overlapping pair once, from the partner with the lower sorted position, over the
// cell-ordered copies, with nine lanes per sorted position, one per neighbouring row of cells
// (item i: position i / 9, row i % 9), so that a tet's rows are scanned side by side.  The 27
// neighbouring cells are 9 rows of three: a row (x - 1 .. x + 1; x is the key's low field) is
// one key range, so it takes one binary search, and only positions after the lane's own are
// visited.  A candidate passes the outward-rounded FP32 box prefilter, then (when asked) the
// shared-node exclusion, which removes most neighbours of a conforming mesh before the FP64
// 12-direction test (the predicate of the reference's k-DOP overlap).  Keys are
// (min << 32 | max) of the original indices.  Pairs are appended in any order (the caller sorts the keys): the
// lanes of a wave that find a pair in the same iteration take consecutive places from one
// atomicAdd on *total, and keys beyond cap are only counted (the caller then grows the
// buffer and searches again).
__global__ __launch_bounds__(256) void k_bp_pairs9(const double2 *__restrict__ skd, int nT,
                                                   const unsigned long long *__restrict__ skey,
                                                   const int *__restrict__ sidx, const int4 *__restrict__ snode,
                                                   const float4 *__restrict__ box, int exclude_shared,
                                                   const double *__restrict__ red,
                                                   unsigned long long *__restrict__ total,
                                                   unsigned long long *__restrict__ out, long long cap)
{
    const long long item = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= 9LL * nT) return;
    const int ja = (int)(item / 9), r = (int)(item % 9);
    const int lane = __lane_id();
    const double cs = bp_cell_size(red);
    int ca[3];
#pragma unroll
    for (int k = 0; k < 3; k++) ca[k] = (int)floor((skd[(size_t)ja * BP_DIRS + k].x - red[1 + k]) / cs);
    const int cy = ca[1] + (r % 3) - 1, cz = ca[2] + (r / 3) - 1;
    if (cy < 0 || cz < 0) return;
    const unsigned long long row = ((unsigned long long)cy << BP_CELL_BITS) | ((unsigned long long)cz << (2 * BP_CELL_BITS));
    const unsigned long long k0 = row | (unsigned long long)(ca[0] > 0 ? ca[0] - 1 : 0);
    const unsigned long long k1 = row | (unsigned long long)(ca[0] + 1);
    int lo_i = 0, hi_i = nT;
    while (lo_i < hi_i) {
        const int mid = (lo_i + hi_i) >> 1;
        if (skey[mid] < k0) lo_i = mid + 1;
        else hi_i = mid;
    }
    const float4 bl = box[2 * (size_t)ja], bh = box[2 * (size_t)ja + 1];
    const int4 na = snode[ja];
    const unsigned a = (unsigned)sidx[ja];
    for (int j = lo_i > ja ? lo_i : ja + 1; j < nT && skey[j] <= k1; j++) {
        bool hit = true;
        const float4 ol = box[2 * (size_t)j], oh = box[2 * (size_t)j + 1];
        if (bl.x > oh.x || bl.y > oh.y || bl.z > oh.z || bh.x < ol.x || bh.y < ol.y || bh.z < ol.z) hit = false;
        if (hit && exclude_shared) {
            const int4 nb = snode[j];
            const int m[4] = {nb.x, nb.y, nb.z, nb.w};
#pragma unroll
            for (int j2 = 0; j2 < 4; j2++) hit = hit && m[j2] != na.x && m[j2] != na.y && m[j2] != na.z && m[j2] != na.w;
        }
        if (hit) {
#pragma unroll
            for (int d = 0; d < BP_DIRS; d++) {
                const double2 u = skd[(size_t)ja * BP_DIRS + d], v = skd[(size_t)j * BP_DIRS + d];
                hit = hit && !(u.x > v.y) && !(u.y < v.x);
            }
        }
        const unsigned long long m = __ballot(hit);
        if (m) {
            const int leader = __ffsll((long long)m) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(total, (unsigned long long)__popcll(m));
            const unsigned lo32 = __shfl((unsigned)(base & 0xffffffffULL), leader, 64);
            const unsigned hi32 = __shfl((unsigned)(base >> 32), leader, 64);
            if (hit) {
                const unsigned long long at = (((unsigned long long)hi32 << 32) | lo32) +
                                              (unsigned long long)__popcll(m & ((1ULL << lane) - 1));
                const unsigned b = (unsigned)sidx[j];
                const unsigned p = a < b ? a : b, q = a < b ? b : a;
                if (at < (unsigned long long)cap) out[at] = ((unsigned long long)p << 32) | q;
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_bp_unpack(const unsigned long long *__restrict__ keys, long long n,
                                                   int *__restrict__ pairs)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const unsigned long long v = keys[k];
    pairs[2 * k] = (int)(v >> 32);
    pairs[2 * k + 1] = (int)(v & 0xffffffffULL);
}

// NL2set, pass 1: tuples per candidate pair (set bits of its mask).
__global__ __launch_bounds__(256) void k_nl2_count(const int *__restrict__ masks, int nPairs, int *__restrict__ cnt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPairs) return;
    cnt[i] = __popc((unsigned)masks[2LL * i] & 0xffu);
}

// Pass 2: (node << 32 | tet) keys with their insertion position 8 i + slot, in insertion order.
__global__ __launch_bounds__(256) void k_nl2_emit(const int *__restrict__ masks, const int *__restrict__ pairs,
                                                  int nPairs, const int *__restrict__ ie, int es,
                                                  const int *__restrict__ off, unsigned long long *__restrict__ key,
                                                  int *__restrict__ pos)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPairs) return;
    const int bits = masks[2LL * i] & 0xff;
    if (!bits) return;
    const int tA = pairs[2LL * i], tB = pairs[2LL * i + 1];
    int o = off[i];
    for (int s = 0; s < 8; s++) {
        if (!(bits & (1 << s))) continue;
        // bits 0-3: vertex s of B inside A; bits 4-7: vertex s-4 of A inside B
        const int node = s < 4 ? ie[tB + (size_t)es * s] : ie[tA + (size_t)es * (s - 4)];
        const int tet = s < 4 ? tA : tB;
        key[o] = ((unsigned long long)(unsigned)node << 32) | (unsigned)tet;
        pos[o] = 8 * i + s;
        o++;
    }
}

// nl2 rows 0-1 from the first-occurrence-ordered keys.
__global__ __launch_bounds__(256) void k_nl2_write(const unsigned long long *__restrict__ key, int n,
                                                   int *__restrict__ nl2, int ts)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    nl2[k] = (int)(key[k] >> 32);
    nl2[k + (size_t)ts] = (int)(key[k] & 0xffffffffULL);
}

}  // namespace icf

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
// of one cell, and of neighbouring cells, are adjacent in memory for k_bp_cellpairs.  Per position:
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

// Cells of the sorted positions: flag[j] = 1 where a new cell key starts.
__global__ __launch_bounds__(256) void k_bp_cellflag(const unsigned long long *__restrict__ skey, int nT,
                                                     int *__restrict__ flag)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nT) flag[j] = (j == 0 || skey[j] != skey[j - 1]) ? 1 : 0;
}

// From the exclusive scan of the flags: each cell's key and first position; cstart[ncells] = nT.
__global__ __launch_bounds__(256) void k_bp_cells(const unsigned long long *__restrict__ skey, int nT,
                                                  const int *__restrict__ flag, const int *__restrict__ cidx,
                                                  unsigned long long *__restrict__ ckey, int *__restrict__ cstart)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nT) return;
    if (flag[j]) {
        ckey[cidx[j]] = skey[j];
        cstart[cidx[j]] = j;
    }
    if (j == nT - 1) cstart[cidx[j] + flag[j]] = nT;
}

// One workgroup per cell: every overlapping pair with both tets in the cell, or one in the cell
// and the other in one of the 13 neighbouring cells whose key is larger (each pair of cells
// visited once), over the cell-ordered copies.  The candidates (the cell itself and the 13
// cells, concatenated) and the cell's own tets are brought into LDS, BP_C and BP_A at a time,
// with coalesced loads; the lanes then take (own, candidate) pairs in turn, BP_T at a time.  A
// pair passes the outward-rounded FP32 box prefilter, then (when asked) the shared-node
// exclusion, which removes most neighbours of a conforming mesh before the FP64 12-direction
// test (the predicate of the reference's k-DOP overlap).  Pairs collect in LDS as (min << 32
// | max) keys of the original indices and leave in batches: one atomicAdd on *total takes
// the batch's places in out (keys beyond cap are only counted; the caller then grows the
// buffer and searches again).  The order of the keys is the caller's to fix (it sorts them).
constexpr int BP_A = 128, BP_C = 256, BP_T = 256, BP_BUF = 1024;
__global__ __launch_bounds__(BP_T) void k_bp_cellpairs(const double2 *__restrict__ skd, int nT,
                                                       const int *__restrict__ sidx, const int4 *__restrict__ snode,
                                                       const float4 *__restrict__ box, int exclude_shared,
                                                       const unsigned long long *__restrict__ ckey,
                                                       const int *__restrict__ cstart, int ncells,
                                                       unsigned long long *__restrict__ total,
                                                       unsigned long long *__restrict__ out, long long cap)
{
    __shared__ float4 abox[2 * BP_A], cbox[2 * BP_C];
    __shared__ int4 anode[BP_A], cnode[BP_C];
    __shared__ int cj[BP_C];
    __shared__ int nb_lo[14], nb_hi[14], nb_pre[15];
    __shared__ unsigned long long pbuf[BP_BUF];
    __shared__ int wave_cnt[BP_T / 64];
    __shared__ int pcount;
    __shared__ unsigned long long flush_base;
    const int c = blockIdx.x, tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
    const int s0 = cstart[c], s1 = cstart[c + 1];
    if (tid < 14) {
        // neighbour t: t = 0 the cell itself, 1-13 the offsets (dx, dy, dz) whose key is larger
        int lo = 0, hi = 0;
        const unsigned long long k = ckey[c];
        const long long cx = (long long)(k & ((1ULL << BP_CELL_BITS) - 1));
        const long long cy = (long long)((k >> BP_CELL_BITS) & ((1ULL << BP_CELL_BITS) - 1));
        const long long cz = (long long)(k >> (2 * BP_CELL_BITS));
        if (tid == 0) {
            lo = s0;
            hi = s1;
        } else {
            const int q = tid + 13;            // offsets 14..26 of the 27 in (dz, dy, dx) order
            const long long nx = cx + q % 3 - 1, ny = cy + (q / 3) % 3 - 1, nz = cz + q / 9 - 1;
            if (nx >= 0 && ny >= 0 && nz >= 0) {
                const unsigned long long nk = (unsigned long long)nx | ((unsigned long long)ny << BP_CELL_BITS) |
                                              ((unsigned long long)nz << (2 * BP_CELL_BITS));
                int l = 0, h = ncells;
                while (l < h) {
                    const int m = (l + h) >> 1;
                    if (ckey[m] < nk) l = m + 1;
                    else h = m;
                }
                if (l < ncells && ckey[l] == nk) {
                    lo = cstart[l];
                    hi = cstart[l + 1];
                }
            }
        }
        nb_lo[tid] = lo;
        nb_hi[tid] = hi;
    }
    if (tid == 0) pcount = 0;
    __syncthreads();
    if (tid == 0) {
        nb_pre[0] = 0;
        for (int t = 0; t < 14; t++) nb_pre[t + 1] = nb_pre[t] + nb_hi[t] - nb_lo[t];
    }
    __syncthreads();
    const int K = nb_pre[14];
    // move the collected keys out: one atomicAdd for the batch, then a coalesced copy
    auto flush = [&]() {
        __syncthreads();
        const int n = pcount;
        if (tid == 0) flush_base = n ? atomicAdd(total, (unsigned long long)n) : 0ULL;
        __syncthreads();
        const unsigned long long base = flush_base;
        for (int i = tid; i < n; i += BP_T)
            if (base + i < (unsigned long long)cap) out[base + i] = pbuf[i];
        __syncthreads();
        if (tid == 0) pcount = 0;
        __syncthreads();
    };
    for (int a0 = s0; a0 < s1; a0 += BP_A) {
        const int na = min(BP_A, s1 - a0);
        __syncthreads();
        for (int i = tid; i < na; i += BP_T) {
            abox[2 * i] = box[2 * (size_t)(a0 + i)];
            abox[2 * i + 1] = box[2 * (size_t)(a0 + i) + 1];
            anode[i] = snode[a0 + i];
        }
        for (int k0 = 0; k0 < K; k0 += BP_C) {
            const int nk = min(BP_C, K - k0);
            __syncthreads();
            for (int i = tid; i < nk; i += BP_T) {
                const int k = k0 + i;
                int t = 0;
                while (nb_pre[t + 1] <= k) t++;
                const int jb = nb_lo[t] + (k - nb_pre[t]);
                cj[i] = t == 0 ? -1 - jb : jb;   // the cell's own tets: pairs with larger positions only
                cbox[2 * i] = box[2 * (size_t)jb];
                cbox[2 * i + 1] = box[2 * (size_t)jb + 1];
                cnode[i] = snode[jb];
            }
            __syncthreads();
            const int npairs = na * nk;
            for (int p0 = 0; p0 < npairs; p0 += BP_T) {
                const int p = p0 + tid;
                bool hit = false;
                int ja = 0, jb = 0;
                if (p < npairs) {
                    const int ia = p / nk, ik = p % nk;
                    ja = a0 + ia;
                    const int code = cj[ik];
                    jb = code < 0 ? -1 - code : code;
                    if (code >= 0 || jb > ja) {
                        const float4 bl = abox[2 * ia], bh = abox[2 * ia + 1];
                        const float4 ol = cbox[2 * ik], oh = cbox[2 * ik + 1];
                        hit = !(bl.x > oh.x || bl.y > oh.y || bl.z > oh.z || bh.x < ol.x || bh.y < ol.y ||
                                bh.z < ol.z);
                        if (hit && exclude_shared) {
                            const int4 na4 = anode[ia], nb4 = cnode[ik];
                            const int m[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
#pragma unroll
                            for (int j2 = 0; j2 < 4; j2++)
                                hit = hit && m[j2] != na4.x && m[j2] != na4.y && m[j2] != na4.z && m[j2] != na4.w;
                        }
                        if (hit) {
#pragma unroll
                            for (int d = 0; d < BP_DIRS; d++) {
                                const double2 u = skd[(size_t)ja * BP_DIRS + d], v = skd[(size_t)jb * BP_DIRS + d];
                                hit = hit && !(u.x > v.y) && !(u.y < v.x);
                            }
                        }
                    }
                }
                // this round's places in the batch: lane prefix within the wave, then the waves
                const unsigned long long m = __ballot(hit);
                if (lane == 0) wave_cnt[wave] = __popcll(m);
                __syncthreads();
                int at = pcount, round = 0;
                for (int w = 0; w < BP_T / 64; w++) {
                    if (w < wave) at += wave_cnt[w];
                    round += wave_cnt[w];
                }
                at += __popcll(m & ((1ULL << lane) - 1));
                if (hit) {
                    const unsigned a = (unsigned)sidx[ja], b = (unsigned)sidx[jb];
                    pbuf[at] = ((unsigned long long)(a < b ? a : b) << 32) | (a < b ? b : a);
                }
                __syncthreads();
                if (tid == 0) pcount += round;
                __syncthreads();
                if (pcount > BP_BUF - BP_T) flush();
            }
        }
    }
    flush();
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

// icf_broad.inc.h -- broad phase and (node, tet) de-duplication on the device (included by
// icf_lib.hip).
//
// Broad phase: every pair (a < b) of surface tets whose 12-direction k-DOPs on the current
// positions overlap -- the set the reference's BVH reports (v3Library/Collisions/BVH/
// BVHN.cs:166-196 with kDOP24.cs:84-140), which is the set icfh_kdop_pairs finds on the host.
// The k-DOP extents are sums and differences of coordinates and the overlap test is a
// comparison, so both produce the same pairs; the order is (a, b) ascending.  The search is
// a uniform grid of cells at least as large as the largest AABB extent: tets are binned by
// the cell of their AABB's minimum corner, so overlapping tets lie in neighbouring cells.
//
// NL2set: the (node, tet) tuples of GPU_Functionality.NarrowPhaseCollisionDetection
// (GPU_Functionality.cs:565-600), de-duplicated keeping each tuple's first occurrence, in
// the order of first occurrence (a HashSet without removals enumerates in insertion order):
// a device hash table keeps each key's smallest insertion position, and the positions that
// are their key's first are compacted in order.
#pragma once

namespace icf {

constexpr int BP_DIRS = 12;

// Bounds of the k-DOP pass, kept as order-preserving integer images of the doubles so that
// atomicMin / atomicMax on them are exact: {largest AABB extent, min of the AABB minimum
// corners (x, y, z), max of the same (x, y, z)}.
constexpr int BP_NBOUNDS = 7;
__host__ __device__ inline unsigned long long bp_ord(double v)
{
    unsigned long long u;
    memcpy(&u, &v, sizeof(u));
    return (u >> 63) ? ~u : (u | 0x8000000000000000ULL);
}
__host__ __device__ inline double bp_unord(unsigned long long u)
{
    u = (u >> 63) ? (u & 0x7fffffffffffffffULL) : ~u;
    double v;
    memcpy(&v, &u, sizeof(v));
    return v;
}

__global__ void k_bp_bounds_init(unsigned long long *__restrict__ b)
{
    const int k = threadIdx.x;
    // min slots start high, max slots low; slot BP_NBOUNDS is the bad-node-id flag
    if (k <= BP_NBOUNDS) b[k] = (k >= 1 && k <= 3) ? ~0ULL : 0ULL;
}

// Tet t's 12-direction k-DOP (lo l, hi h) from its 4 nodes' current positions; false if a
// node id is out of range.  Both kernels below compute it with this code, so their values agree
// bit for bit.
__device__ inline bool bp_tet_kdop(const double *__restrict__ dn, int ns, const int *__restrict__ ie, int es, int t,
                                   int nNodesBound, double l[BP_DIRS], double h[BP_DIRS])
{
#pragma clang fp contract(off)
#pragma unroll
    for (int d = 0; d < BP_DIRS; d++) {
        l[d] = 1e300;
        h[d] = -1e300;
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int n = ie[t + (size_t)es * v];
        if (n < 0 || (nNodesBound > 0 && n >= nNodesBound)) return false;
        const double x = dn[n + (size_t)ns * ND_XC], y = dn[n + (size_t)ns * (ND_XC + 1)],
                     z = dn[n + (size_t)ns * (ND_XC + 2)];
        const double p[BP_DIRS] = {x, y, z, x + y, x + z, y + z, x - y, x - z, y - z, x + y - z, x + z - y, y + z - x};
#pragma unroll
        for (int d = 0; d < BP_DIRS; d++) {
            l[d] = fmin(l[d], p[d]);
            h[d] = fmax(h[d], p[d]);
        }
    }
    return true;
}

// Per surface tet: the minimum corner of its AABB (the k-DOP's first three lo planes, which
// bin it into the grid), and the block's contribution to the bounds (one atomic per bound per
// block).  k_bp_sorted computes the whole k-DOP again rather than reading 24 planes back.
__global__ __launch_bounds__(256) void k_bp_kdop(const double *__restrict__ dn, int ns, const int *__restrict__ ie,
                                                 int es, int nT, double *__restrict__ lo,
                                                 unsigned long long *__restrict__ bounds, int nNodesBound)
{
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double b[BP_NBOUNDS] = {0, 1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
    if (t < nT) {
        double l[BP_DIRS], h[BP_DIRS];
        if (!bp_tet_kdop(dn, ns, ie, es, t, nNodesBound, l, h)) {
            atomicMax(&bounds[BP_NBOUNDS], 1ULL);
        } else {
#pragma unroll
            for (int d = 0; d < 3; d++) lo[(size_t)d * nT + t] = l[d];
            b[0] = fmax(h[0] - l[0], fmax(h[1] - l[1], h[2] - l[2]));
#pragma unroll
            for (int k = 0; k < 3; k++) b[1 + k] = b[4 + k] = l[k];
        }
    }
    // block reduction: across the wave, then the block's waves through LDS
    __shared__ double part[256 / 64][BP_NBOUNDS];
#pragma unroll
    for (int k = 0; k < BP_NBOUNDS; k++) {
        double v = b[k];
        for (int o = 32; o > 0; o >>= 1) {
            const double w = __shfl_xor(v, o);
            v = (k == 0 || k >= 4) ? fmax(v, w) : fmin(v, w);
        }
        b[k] = v;
    }
    const int wave = threadIdx.x >> 6;
    if (__lane_id() == 0)
        for (int k = 0; k < BP_NBOUNDS; k++) part[wave][k] = b[k];
    __syncthreads();
    if (threadIdx.x < BP_NBOUNDS) {
        const int k = threadIdx.x;
        double v = part[0][k];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) v = (k == 0 || k >= 4) ? fmax(v, part[w][k]) : fmin(v, part[w][k]);
        if (k == 0 || k >= 4) atomicMax(&bounds[k], bp_ord(v));
        else atomicMin(&bounds[k], bp_ord(v));
    }
}

// The search grid: cells of size cs (at least the largest AABB extent), n[k] cells per axis
// over the AABB minimum corners, stored with a margin -- one cell below and two above in x,
// one below and one above in y and z -- so that the neighbour rows a cell reads never wrap.
// Cell (x, y, z) of the margin grid has the linear key x + gx (y + gy z).  Tets are binned
// by the cell of their AABB's minimum corner; with cs >= every extent, overlapping tets lie
// in the same or neighbouring cells.
struct BpGrid {
    double cs, org[3];
    int n[3], gx, gy, gz;
};

// The rank of this lane's key among all ranks taken from count[key] (bumped atomically), for
// lanes with on set (-1 for the others): one atomicAdd per run of consecutive such lanes with
// the same key -- the run's first lane reserves the run's length and its lanes take
// consecutive ranks.  Where keys repeat along a wave this replaces up to 64 atomics on one
// address, which the L2 serialises, by one.  Lanes above an inactive lane must be inactive
// too (a kernel's early exit for its last indices).
__device__ inline int run_rank(bool on, unsigned key, int *__restrict__ count)
{
    const int lane = __lane_id();
    const unsigned long long om = __ballot(on);
    const unsigned prev = __shfl_up(key, 1);
    const bool head = on && (lane == 0 || !((om >> (lane - 1)) & 1) || prev != key);
    const unsigned long long heads = __ballot(head);
    const unsigned long long ends = heads | ~om;   // lanes where a run cannot continue
    int r = -1, h = 0;
    if (on) h = 63 - __clzll(heads & ((2ULL << lane) - 1));   // the run's first lane
    if (head) {
        const unsigned long long above = lane == 63 ? 0ULL : ends & ~((2ULL << lane) - 1);
        const int len = above ? __ffsll((long long)above) - 1 - lane : 64 - lane;
        r = atomicAdd(&count[key], len);
    }
    const int base = __shfl(r, h);
    return on ? base + lane - h : -1;
}

// Per tet: its cell key and its rank among the tets of its cell (count[key] is bumped
// atomically, so the order within a cell varies from call to call; the pairs do not).
__global__ __launch_bounds__(256) void k_bp_cell(const double *__restrict__ lo, int nT, BpGrid g,
                                                 unsigned *__restrict__ key, int *__restrict__ rank,
                                                 int *__restrict__ count)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    int c[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double f = floor((lo[(size_t)a * nT + t] - g.org[a]) / g.cs);
        c[a] = f < 0 ? 0 : (f > (double)(g.n[a] - 1) ? g.n[a] - 1 : (int)f);
    }
    const unsigned k = (unsigned)(c[0] + 1) + (unsigned)g.gx * ((unsigned)(c[1] + 1) + (unsigned)g.gy * (unsigned)(c[2] + 1));
    key[t] = k;
    rank[t] = run_rank(true, k, count);   // neighbouring tets of a mesh often share a cell
}

// The k-DOPs and node ids in cell order (a counting sort: first[key] is the exclusive scan
// of the cell counts): tet t goes to position j = first[key[t]] + rank[t], so that the tets
// of one cell, and of neighbouring cells, are adjacent in memory for the pair search.  Per
// position: the tet (sidx), the k-DOP as (lo, hi) pairs per direction (skd, 24 doubles), the
// node ids (snode), and the k-DOP in FP32 rounded outward, lo down and hi up (sk32: 12 lo then
// 12 hi): a prefilter that never rejects a pair whose FP64 k-DOPs overlap.  first[L] for
// L = 0 .. G is then the first position of the cells with keys >= L.
__global__ __launch_bounds__(256) void k_bp_sorted(const double *__restrict__ dn, int ns, int nNodesBound, int nT,
                                                   const unsigned *__restrict__ key,
                                                   const int *__restrict__ rank, const int *__restrict__ first,
                                                   const int *__restrict__ ie, int es, int *__restrict__ sidx,
                                                   double2 *__restrict__ skd, int4 *__restrict__ snode,
                                                   float4 *__restrict__ sk32)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    const int j = first[key[t]] + rank[t];
    sidx[j] = t;
    double l[BP_DIRS], h[BP_DIRS];
    (void)bp_tet_kdop(dn, ns, ie, es, t, nNodesBound, l, h);   // node ids checked by k_bp_kdop
    float f[2 * BP_DIRS];
#pragma unroll
    for (int d = 0; d < BP_DIRS; d++) {
        skd[(size_t)j * BP_DIRS + d] = make_double2(l[d], h[d]);
        f[d] = __double2float_rd(l[d]);
        f[BP_DIRS + d] = __double2float_ru(h[d]);
    }
    snode[j] = make_int4(ie[t], ie[t + (size_t)es], ie[t + 2 * (size_t)es], ie[t + 3 * (size_t)es]);
#pragma unroll
    for (int q = 0; q < 6; q++)
        sk32[(size_t)j * 6 + q] = make_float4(f[4 * q], f[4 * q + 1], f[4 * q + 2], f[4 * q + 3]);
}

// A wave's LDS writes made visible to the wave's other lanes (no workgroup barrier).
__device__ inline void bp_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The pair search, first pass: one wave per cell of the grid's interior; the waves of a
// workgroup meet only at its start.  The candidates are the cell's own tets and those of the 13 neighbouring cells
// with a larger key (each pair of cells visited once): 5 runs of consecutive keys -- (x, x+1)
// of the cell's own row, and the rows (x-1 .. x+1) at (y+1, z), (y-1, z+1), (y, z+1),
// (y+1, z+1) -- hence 5 ranges of sorted positions; in the first range only larger positions
// pair up.  The 64 lanes hold 64 candidates at a time (FP32 k-DOP and node ids in registers,
// loaded coalesced); the cell's own tets, BPW_OWN at a time, are staged in the wave's LDS and
// read by all lanes at once.  A pair passes if its outward-rounded FP32 k-DOPs overlap in all
// 12 directions and (when asked) the tets share no node.  The passing (own, candidate)
// positions collect in the wave's LDS buffer and leave in batches, one atomicAdd on the
// counter of the workgroup's buffer region per batch (positions beyond the region's capacity
// are only counted; the caller then grows the buffer and searches again).  What a wave holds at its end leaves with the rest of its workgroup's:
// the last wave of the workgroup to finish reserves space for all of them with one atomicAdd
// (the reservations on the one counter, one per wave, took a quarter of the kernel's time).
// Two waves per workgroup: a workgroup's LDS stays allocated until its slowest wave ends, and
// the work per cell varies several-fold, so with four waves the resident waves averaged about
// two per SIMD.  The second pass applies the exact FP64 test.
constexpr int BPW_WAVES = 2, BPW_OWN = 16, BPW_BUF = 256;
// The candidate buffer is split into BP_PARTS regions of cap / BP_PARTS entries, each with its
// own counter (BP_PART_STRIDE counters apart, so that no two share a cache line): workgroup w
// appends to region w mod BP_PARTS, which spreads the reservations over as many addresses.
constexpr int BP_PARTS = 8, BP_PART_STRIDE = 64;

// The later passes run a grid row per region (blockIdx.y): this thread's position in the
// candidate buffer, or -1 beyond the region's count (its counter, capped at the capacity).
__device__ inline long long bp_slot(long long cap, const unsigned long long *__restrict__ total)
{
    const long long capP = cap / BP_PARTS;
    const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)total[(size_t)blockIdx.y * BP_PART_STRIDE];
    return o < (n < capP ? n : capP) ? (long long)blockIdx.y * capP + o : -1;
}

// One wave's cell of the pair search (k_bp_cellwave): on return, cnt pairs are left in buf.
__device__ inline void bp_cell_pairs(const float4 *__restrict__ sk32, const int4 *__restrict__ snode,
                                     int exclude_shared, const BpGrid &g, const int *__restrict__ first,
                                     unsigned long long *__restrict__ total, unsigned long long *__restrict__ out,
                                     long long cap, int i, float4 (*ow)[7], unsigned long long *buf, int lane,
                                     unsigned long long below, int &cnt)
{
    const int x = i % g.n[0] + 1, y = (i / g.n[0]) % g.n[1] + 1, z = i / (g.n[0] * g.n[1]) + 1;
    const int L = x + g.gx * (y + g.gy * z);
    const int s0 = first[L], s1 = first[L + 1];
    if (s0 == s1) return;
    const int gxy = g.gx * g.gy;
    const int r_lo[5] = {s0, first[L - 1 + g.gx], first[L - 1 - g.gx + gxy], first[L - 1 + gxy],
                         first[L - 1 + g.gx + gxy]};
    const int r_hi[5] = {first[L + 2], first[L + 2 + g.gx], first[L + 2 - g.gx + gxy], first[L + 2 + gxy],
                         first[L + 2 + g.gx + gxy]};
    int r_pre[6];
    r_pre[0] = 0;
#pragma unroll
    for (int r = 0; r < 5; r++) r_pre[r + 1] = r_pre[r] + r_hi[r] - r_lo[r];
    const int K = r_pre[5];
    auto flush = [&]() {
        bp_wave_sync();
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(total, (unsigned long long)cnt);
        base = __shfl(base, 0);
        for (int q = lane; q < cnt; q += 64)
            if (base + q < (unsigned long long)cap) out[base + q] = buf[q];
        bp_wave_sync();
        cnt = 0;
    };
    for (int a0 = s0; a0 < s1; a0 += BPW_OWN) {
        const int na = min(BPW_OWN, s1 - a0);
        bp_wave_sync();   // the previous own tets are no longer read
        if (lane < na) {
#pragma unroll
            for (int q = 0; q < 6; q++) ow[lane][q] = sk32[(size_t)(a0 + lane) * 6 + q];
            const int4 n4 = snode[a0 + lane];
            ow[lane][6] = make_float4(__int_as_float(n4.x), __int_as_float(n4.y), __int_as_float(n4.z),
                                      __int_as_float(n4.w));
        }
        bp_wave_sync();
        for (int k0 = 0; k0 < K; k0 += 64) {
            const int k = k0 + lane;
            const bool valid = k < K;
            int jb = s0, own_row = 0;
#pragma unroll
            for (int r = 0; r < 5; r++)
                if (k >= r_pre[r] && k < r_pre[r + 1]) {
                    jb = r_lo[r] + (k - r_pre[r]);
                    own_row = r == 0;
                }
            float c[2 * BP_DIRS];
#pragma unroll
            for (int q = 0; q < 6; q++) {
                const float4 v = sk32[(size_t)jb * 6 + q];
                c[4 * q] = v.x;
                c[4 * q + 1] = v.y;
                c[4 * q + 2] = v.z;
                c[4 * q + 3] = v.w;
            }
            const int4 cn = snode[jb];
            for (int ia = 0; ia < na; ia++) {
                const int ja = a0 + ia;
                float a[2 * BP_DIRS];
#pragma unroll
                for (int q = 0; q < 6; q++) {
                    const float4 v = ow[ia][q];
                    a[4 * q] = v.x;
                    a[4 * q + 1] = v.y;
                    a[4 * q + 2] = v.z;
                    a[4 * q + 3] = v.w;
                }
                const float4 an4 = ow[ia][6];
                // the wave's mask of passing lanes, one ballot per comparison (each a compare
                // into a mask and a scalar AND)
                unsigned long long mk = __ballot(valid & (!own_row | (jb > ja)));
#pragma unroll
                for (int d = 0; d < BP_DIRS; d++) {
                    mk &= ~__ballot(a[d] > c[BP_DIRS + d]);
                    mk &= ~__ballot(a[BP_DIRS + d] < c[d]);
                }
                if (exclude_shared) {
                    const int an[4] = {__float_as_int(an4.x), __float_as_int(an4.y), __float_as_int(an4.z),
                                       __float_as_int(an4.w)};
                    const int m[4] = {cn.x, cn.y, cn.z, cn.w};
#pragma unroll
                    for (int j2 = 0; j2 < 4; j2++)
#pragma unroll
                        for (int j1 = 0; j1 < 4; j1++) mk &= ~__ballot(m[j2] == an[j1]);
                }
                if (mk == 0) continue;
                const bool hit = (mk >> lane) & 1;
                if (hit) buf[cnt + __popcll(mk & below)] = ((unsigned long long)ja << 32) | (unsigned)jb;
                cnt += __popcll(mk);
                if (cnt > BPW_BUF - 64) flush();
            }
        }
    }
}

// At most 80 registers, for 6 waves per SIMD (88 otherwise: 5 waves; a 12-byte spill; forcing
// 8 waves spills 100 bytes and doubles the time).
__global__ __launch_bounds__(64 * BPW_WAVES) __attribute__((amdgpu_waves_per_eu(6))) void k_bp_cellwave(const float4 *__restrict__ sk32,
                                                                const int4 *__restrict__ snode, int exclude_shared,
                                                                BpGrid g, const int *__restrict__ first,
                                                                unsigned long long *__restrict__ total,
                                                                unsigned long long *__restrict__ out, long long cap)
{
    __shared__ float4 own[BPW_WAVES][BPW_OWN][7];
    __shared__ unsigned long long pbuf[BPW_WAVES][BPW_BUF];
    __shared__ int wcnt[BPW_WAVES], done;
    const int wave = threadIdx.x >> 6, lane = __lane_id();
    if (threadIdx.x == 0) done = 0;
    __syncthreads();
    const unsigned long long below = (1ULL << lane) - 1;
    unsigned long long *buf = pbuf[wave];
    int cnt = 0;
    const int ncell = g.n[0] * g.n[1] * g.n[2];
    const int i = (int)blockIdx.x * BPW_WAVES + wave;
    const long long capP = cap / BP_PARTS;
    const int part = (int)(blockIdx.x % BP_PARTS);
    total += (size_t)part * BP_PART_STRIDE;
    out += (size_t)part * capP;
    if (i < ncell) bp_cell_pairs(sk32, snode, exclude_shared, g, first, total, out, capP, i, own[wave], buf, lane,
                                 below, cnt);
    // the workgroup's last wave out writes every wave's remaining pairs
    if (lane == 0) wcnt[wave] = cnt;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    int last = 0;
    if (lane == 0) last = atomicAdd(&done, 1) == BPW_WAVES - 1;
    if (!__shfl(last, 0)) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    int pre[BPW_WAVES + 1];
    pre[0] = 0;
#pragma unroll
    for (int w = 0; w < BPW_WAVES; w++) pre[w + 1] = pre[w] + wcnt[w];
    if (pre[BPW_WAVES] == 0) return;
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(total, (unsigned long long)pre[BPW_WAVES]);
    base = __shfl(base, 0);
#pragma unroll
    for (int w = 0; w < BPW_WAVES; w++)
        for (int q = lane; q < wcnt[w]; q += 64)
            if (base + pre[w] + q < (unsigned long long)capP) out[base + pre[w] + q] = pbuf[w][q];
}

// The pair search, second pass: per candidate (own, candidate) position pair of the first
// pass, the exact test -- the FP64 12-direction overlap, the predicate of the reference's
// k-DOP test.  The FP32 copies decide first where they can: the first pass's outward-rounded
// values, each moved one step inward (nextafter), still overlap only if the FP64 values do
// (lo <= nextafter(lo rounded down, +inf) and hi >= nextafter(hi rounded up, -inf)), so the
// FP64 k-DOPs are read only for the pairs that are within an FP32 step of touching.  A pair
// that passes becomes (a, b), a < b its original indices (io[i] = a << 32 | b), and takes its
// rank among the pairs of the same a (count[a] bumped atomically); rank[i] = -1 for a pair
// that fails.
__global__ __launch_bounds__(256) void k_bp_exact(const double2 *__restrict__ skd, const float4 *__restrict__ sk32,
                                                  const int *__restrict__ sidx,
                                                  const unsigned long long *__restrict__ ncand, long long cap,
                                                  unsigned long long *__restrict__ io, int *__restrict__ rank,
                                                  int *__restrict__ count)
{
    const long long i = bp_slot(cap, ncand);
    if (i < 0) return;   // the region's used positions come first, so only a wave's top lanes leave
    const unsigned long long v = io[i];
    const int ja = (int)(v >> 32), jb = (int)(v & 0xffffffffULL);
    float fa[2 * BP_DIRS], fb[2 * BP_DIRS];
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const float4 x = sk32[(size_t)ja * 6 + q], y = sk32[(size_t)jb * 6 + q];
        fa[4 * q] = x.x, fa[4 * q + 1] = x.y, fa[4 * q + 2] = x.z, fa[4 * q + 3] = x.w;
        fb[4 * q] = y.x, fb[4 * q + 1] = y.y, fb[4 * q + 2] = y.z, fb[4 * q + 3] = y.w;
    }
    bool sure = true;
#pragma unroll
    for (int d = 0; d < BP_DIRS; d++) {
        sure = sure & (nextafterf(fa[d], INFINITY) <= nextafterf(fb[BP_DIRS + d], -INFINITY)) &
               (nextafterf(fb[d], INFINITY) <= nextafterf(fa[BP_DIRS + d], -INFINITY));
    }
    bool hit = sure;
    if (!sure) {
        hit = true;
#pragma unroll
        for (int d = 0; d < BP_DIRS; d++) {
            const double2 u = skd[(size_t)ja * BP_DIRS + d], w = skd[(size_t)jb * BP_DIRS + d];
            hit = hit & !(u.x > w.y) & !(u.y < w.x);
        }
    }
    unsigned a = 0xffffffffu;
    if (hit) {
        const unsigned oa = (unsigned)sidx[ja], ob = (unsigned)sidx[jb];
        const unsigned b = oa < ob ? ob : oa;
        a = oa < ob ? oa : ob;
        io[i] = ((unsigned long long)a << 32) | b;
    }
    // the first pass emits the candidates of one own tet together, so a repeats along the wave
    rank[i] = run_rank(hit, a, count);
}

// The pairs grouped by a, in no particular order within a group: pair i goes to
// start[a] + rank[i] (start = exclusive scan of the counts).
__global__ __launch_bounds__(256) void k_bp_scatter(const unsigned long long *__restrict__ io,
                                                    const int *__restrict__ rank,
                                                    const unsigned long long *__restrict__ ncand, long long cap,
                                                    const int *__restrict__ start,
                                                    unsigned long long *__restrict__ grp)
{
    const long long i = bp_slot(cap, ncand);
    if (i < 0) return;
    const int r = rank[i];
    if (r < 0) return;
    const unsigned long long v = io[i];
    grp[start[v >> 32] + r] = v;
}

// Each grouped pair's final place: its group's start plus the number of b's in the group
// smaller than its own (the b's of a group are distinct), so the pairs come out in (a, b)
// order; those below max_pairs are written.
__global__ __launch_bounds__(256) void k_bp_place(const unsigned long long *__restrict__ grp,
                                                  const int *__restrict__ start, int nT, long long max_pairs,
                                                  int *__restrict__ pairs)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)start[nT]) return;
    const unsigned long long v = grp[p];
    const int a = (int)(v >> 32), b = (int)(v & 0xffffffffULL);
    const int s = start[a], e = start[a + 1];
    int r = 0;
    for (int q = s; q < e; q++) r += (int)(grp[q] & 0xffffffffULL) < b;
    const long long at = (long long)s + r;
    if (at < max_pairs) {
        pairs[2 * at] = a;
        pairs[2 * at + 1] = b;
    }
}

// NL2set, pass 1: tuples per candidate pair (set bits of its mask).
__global__ __launch_bounds__(256) void k_nl2_count(const int *__restrict__ masks, int nPairs, int *__restrict__ cnt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nPairs) return;
    cnt[i] = __popc((unsigned)masks[2LL * i] & 0xffu);
}

// Pass 2: the (node << 32 | tet) keys in insertion order (position o), each inserted into an
// open-addressing table (linear probing, at most half full) that keeps the smallest
// position of every key: its first occurrence.
constexpr unsigned long long NL2_EMPTY = ~0ULL;   // no key: node and tet ids are >= 0
__device__ inline unsigned long long nl2_hash(unsigned long long k)
{
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    return k;
}
__global__ __launch_bounds__(256) void k_nl2_emit(const int *__restrict__ masks, const int *__restrict__ pairs,
                                                  int nPairs, const int *__restrict__ ie, int es,
                                                  const int *__restrict__ off, unsigned long long *__restrict__ key,
                                                  unsigned long long *__restrict__ tkey, int *__restrict__ tpos,
                                                  unsigned long long tmask)
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
        const unsigned long long k = ((unsigned long long)(unsigned)node << 32) | (unsigned)tet;
        key[o] = k;
        for (unsigned long long h = nl2_hash(k) & tmask;; h = (h + 1) & tmask) {
            const unsigned long long prev = atomicCAS(&tkey[h], NL2_EMPTY, k);
            if (prev == NL2_EMPTY || prev == k) {
                atomicMin(&tpos[h], o);
                break;
            }
        }
        o++;
    }
}

// Pass 3: keep[o] = 1 where position o is its key's first occurrence (keep[n] = 0 is set by
// the caller, so the exclusive scan of keep ends with the count).
__global__ __launch_bounds__(256) void k_nl2_keep(const unsigned long long *__restrict__ key, int n,
                                                  const unsigned long long *__restrict__ tkey,
                                                  const int *__restrict__ tpos, unsigned long long tmask,
                                                  int *__restrict__ keep)
{
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    const unsigned long long k = key[o];
    unsigned long long h = nl2_hash(k) & tmask;
    while (tkey[h] != k) h = (h + 1) & tmask;
    keep[o] = tpos[h] == o ? 1 : 0;
}

// nl2 rows 0-1: the kept keys at their places in first-occurrence order.
__global__ __launch_bounds__(256) void k_nl2_write(const unsigned long long *__restrict__ key, int n,
                                                   const int *__restrict__ keep, const int *__restrict__ dst,
                                                   int *__restrict__ nl2, int ts)
{
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n || !keep[o]) return;
    const int k = dst[o];
    if (k >= ts) return;   // beyond the output's capacity: the caller reports the count
    nl2[k] = (int)(key[o] >> 32);
    nl2[k + (size_t)ts] = (int)(key[o] & 0xffffffffULL);
}

__global__ void k_fill_u64(unsigned long long *__restrict__ p, long long n, unsigned long long v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace icf

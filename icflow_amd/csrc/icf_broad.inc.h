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

// Visit the overlapping partners b > a of tet a; count them or write (a << 32 | b) keys.
template <bool WRITE>
__global__ __launch_bounds__(256) void k_bp_pairs(const double *__restrict__ lo, const double *__restrict__ hi, int nT,
                                                  const unsigned long long *__restrict__ skey,
                                                  const int *__restrict__ sidx, const int *__restrict__ ie, int es,
                                                  int exclude_shared, const double *__restrict__ red,
                                                  long long *__restrict__ count, const long long *__restrict__ off,
                                                  unsigned long long *__restrict__ out)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nT) return;
    const double cs = bp_cell_size(red);
    double la[BP_DIRS], ha[BP_DIRS];
#pragma unroll
    for (int d = 0; d < BP_DIRS; d++) {
        la[d] = lo[(size_t)d * nT + a];
        ha[d] = hi[(size_t)d * nT + a];
    }
    int ca[3];
#pragma unroll
    for (int k = 0; k < 3; k++) ca[k] = (int)floor((la[k] - red[1 + k]) / cs);
    int na[4];
#pragma unroll
    for (int i = 0; i < 4; i++) na[i] = ie[a + (size_t)es * i];
    long long pos = WRITE ? off[a] : 0;
    int cnt = 0;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const int cx = ca[0] + dx, cy = ca[1] + dy, cz = ca[2] + dz;
                if (cx < 0 || cy < 0 || cz < 0) continue;
                const unsigned long long key = (unsigned long long)cx | ((unsigned long long)cy << BP_CELL_BITS) |
                                               ((unsigned long long)cz << (2 * BP_CELL_BITS));
                int lo_i = 0, hi_i = nT;
                while (lo_i < hi_i) {
                    const int mid = (lo_i + hi_i) >> 1;
                    if (skey[mid] < key) lo_i = mid + 1;
                    else hi_i = mid;
                }
                for (int j = lo_i; j < nT && skey[j] == key; j++) {
                    const int b = sidx[j];
                    if (b <= a) continue;
                    bool ov = true;
#pragma unroll
                    for (int d = 0; d < BP_DIRS; d++)
                        ov = ov && !(la[d] > hi[(size_t)d * nT + b]) && !(ha[d] < lo[(size_t)d * nT + b]);
                    if (!ov) continue;
                    if (exclude_shared) {
                        bool sh = false;
#pragma unroll
                        for (int j2 = 0; j2 < 4; j2++) {
                            const int nb = ie[b + (size_t)es * j2];
                            sh = sh || nb == na[0] || nb == na[1] || nb == na[2] || nb == na[3];
                        }
                        if (sh) continue;
                    }
                    if (WRITE) out[pos + cnt] = ((unsigned long long)a << 32) | (unsigned)b;
                    cnt++;
                }
            }
    if (!WRITE) count[a] = cnt;
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

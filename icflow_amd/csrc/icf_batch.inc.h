// icf_batch.inc.h -- batched specimens (included by icf_lib.hip after icf_reduce.inc.h).
//
// Many independent specimens share the buffers of one fused call, back to back: specimen s owns
// the units [off[s], off[s + 1]) of each list (tets, CZs) and the rows and BCSR blocks in its
// own ranges.  These kernels check that layout (every unit writes only its specimen's blocks and
// rows) and count the CZ flags per specimen (the Sum of GPU_Functionality.cs:93-107,470-475,
// taken per specimen).
#pragma once

namespace icf {

// The specimen s with off[s] <= u < off[s + 1]; off holds n + 1 non-decreasing entries and
// off[0] <= u < off[n].  Empty specimens are never returned.
__device__ inline int specimen_of(const int *__restrict__ off, int n, int u)
{
    int lo = 0, hi = n;                         // off[lo] <= u < off[hi]
    while (hi - lo > 1) {
        const int m = (lo + hi) >> 1;
        if (off[m] <= u) lo = m;
        else hi = m;
    }
    return lo;
}

// Offsets of a batch on the device: four tables of n + 1 entries (tets, CZs, rows, blocks).
struct BatchOff {
    const int *tet, *cz, *row, *blk;
    int n;
};

// Threads [0, nT) check tets, [nT, nT + nC) CZs: each slot (row) a unit writes, when not
// negative (GPU_Functionality.cs:355-369: -1 = skip), must be a block (row) of the unit's
// specimen.  Any other value sets *err.
__global__ __launch_bounds__(256) void k_batch_check(const int *__restrict__ ie, int nT, int es,
                                                     const int *__restrict__ icz, int nC, int cs, BatchOff b,
                                                     int *__restrict__ err)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT + nC) return;
    const bool tet = t < nT;
    const int u = tet ? t : t - nT;
    const int s = specimen_of(tet ? b.tet : b.cz, b.n, u);
    const int *base = tet ? ie : icz;
    const size_t stride = (size_t)(tet ? es : cs);
    const int f_slot = tet ? EL_SLOTS : CZI_SLOTS, n_slot = tet ? 16 : 36;
    const int f_row = tet ? EL_ROWS : CZI_ROWS, n_row = tet ? 4 : 6;
    const int blo = b.blk[s], bhi = b.blk[s + 1], rlo = b.row[s], rhi = b.row[s + 1];
    bool bad = false;
    for (int i = 0; i < n_slot; i++) {
        const int v = base[u + stride * (f_slot + i)];
        bad |= v >= 0 && (v < blo || v >= bhi);
    }
    for (int i = 0; i < n_row; i++) {
        const int v = base[u + stride * (f_row + i)];
        bad |= v >= 0 && (v < rlo || v >= rhi);
    }
    if (bad) atomicOr(err, 1);
}

// counts[2 s], counts[2 s + 1] = sums of damaged[u] and failed[u] over the CZs u of specimen
// s = blockIdx.x (one 256-thread workgroup per specimen; no atomics, every count written).
__global__ __launch_bounds__(256) void k_batch_counts(const int *__restrict__ damaged, const int *__restrict__ failed,
                                                      const int *__restrict__ cz_off, int *__restrict__ counts)
{
    const int s = blockIdx.x;
    const int a = cz_off[s], b = cz_off[s + 1];
    int d = 0, f = 0;
    for (int u = a + (int)threadIdx.x; u < b; u += 256) {
        d += damaged[u];
        f += failed[u];
    }
    d = block_sum256(d);
    f = block_sum256(f);
    if (threadIdx.x == 0) {
        counts[2 * s] = d;
        counts[2 * s + 1] = f;
    }
}

}  // namespace icf

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

// q-major enumeration: adjacent threads read adjacent units of one field (coalesced).
__global__ __launch_bounds__(256) void k_map_maxkey(MapSrc s, int *__restrict__ maxkey)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long tot = (long long)s.n * s.nq;
    int key = -1;
    if (i < tot) {
        const int q = (int)(i / s.n), u = (int)(i % s.n);
        key = s.base[u + (size_t)s.stride * (s.field0 + q)];
    }
    // wavefront max, then one atomic per wavefront
    for (int off = 32; off > 0; off >>= 1) key = max(key, __shfl_xor(key, off));
    if ((threadIdx.x & 63) == 0 && key >= 0) atomicMax(maxkey, key);
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

// Staged per-unit records of the current call.
struct Stage {
    const double *tet_rhs;
    const void *tet_tan;
    const double *cz_rhs;
    const void *cz_tan;
    const double *ct;
    double stiff, mcoef;   // tet LHS coefficients (element_kernels2.h:239-240)
};

// Phase 2 for BCSR values: one thread per slot.  accumulate != 0 adds to A (reference
// semantics for a single kernel call); accumulate == 0 overwrites, zero for slots without
// contributions (the fused AssembleElemsAndCZs, whose memset is folded in here).
// keys != nullptr: thread t handles key keys[t] (t < nkeys), the keys that have contributions.
template <bool MIXED, bool ISO>
__global__ __launch_bounds__(256) void k_reduce_vals(const int *__restrict__ off, const uint32_t *__restrict__ codes,
                                                     const int *__restrict__ keys, int nkeys, double *__restrict__ A,
                                                     int accumulate, Stage sv)
{
    using T = typename std::conditional<MIXED, float, double>::type;
    const int t = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= nkeys) return;
    const int s = keys ? keys[t] : t;
    const int b = off[s], e = off[s + 1];
    double acc[9];
    double *dst = A + (size_t)s * 9;
    if (accumulate) {
        if (b == e) return;
#pragma unroll
        for (int i = 0; i < 9; i++) acc[i] = dst[i];
    } else {
#pragma unroll
        for (int i = 0; i < 9; i++) acc[i] = 0.;
    }
    const T stiff = (T)sv.stiff, mcoef = (T)sv.mcoef;
    int c = b;
    if (ISO) {
        // Tets come first in a slot's list (codes are sorted by kind).  Four at a time: all
        // loads issued before the in-order accumulation, for memory-level parallelism.
        const T *tt = (const T *)sv.tet_tan;
        for (; c + 4 <= e; c += 4) {
            const uint32_t k0 = codes[c], k1 = codes[c + 1], k2 = codes[c + 2], k3 = codes[c + 3];
            if (code_kind(k3) != KIND_TET) break;
            TetIsoLoad<T> l0, l1, l2, l3;
            l0.load(tt, k0);
            l1.load(tt, k1);
            l2.load(tt, k2);
            l3.load(tt, k3);
            l0.add(acc, stiff, mcoef);
            l1.add(acc, stiff, mcoef);
            l2.add(acc, stiff, mcoef);
            l3.add(acc, stiff, mcoef);
        }
    }
    for (; c < e; c++) {
        const uint32_t code = codes[c];
        const uint32_t kind = code_kind(code), u = code_unit(code), q = code_q(code);
        if (kind == KIND_TET) {
            if (ISO) tet_block_iso<T>(acc, (const T *)sv.tet_tan + (size_t)u * TET_ISO_REC, q >> 2, q & 3, stiff, mcoef);
            else tet_block_gen<T>(acc, (const T *)sv.tet_tan + (size_t)u * TET_GEN_REC, q >> 2, q & 3, stiff, mcoef);
        } else if (kind == KIND_CZ) {
            cz_block<T>(acc, (const T *)sv.cz_tan + (size_t)u * CZ_TAN_REC, q / 6, q % 6);
        } else {
            contact_block(acc, sv.ct + (size_t)u * CT_REC, q >> 2, q & 3);
        }
    }
#pragma unroll
    for (int i = 0; i < 9; i++) dst[i] = acc[i];
}

__device__ inline const double *rhs_of(const Stage &sv, uint32_t code)
{
    const uint32_t kind = code_kind(code), u = code_unit(code), q = code_q(code);
    if (kind == KIND_TET) return sv.tet_rhs + (size_t)u * TET_RHS_REC + 3 * q;
    if (kind == KIND_CZ) return sv.cz_rhs + (size_t)u * CZ_RHS_REC + 3 * q;
    return sv.ct + (size_t)u * CT_REC + CT_F_OFF + 3 * q;
}

// Phase 2 for rhs rows (b[row*3+k]) or node forces (dn[node + ns*(F+k)]): one thread per key.
__global__ __launch_bounds__(256) void k_reduce_vec3(const int *__restrict__ off, const uint32_t *__restrict__ codes,
                                                     const int *__restrict__ keys, int nkeys, double *__restrict__ out,
                                                     int comp_stride, int elem_stride, int accumulate, Stage sv)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nkeys) return;
    const int s = keys ? keys[t] : t;
    const int b = off[s], e = off[s + 1];
    double *dst = out + (size_t)s * elem_stride;
    double a0, a1, a2;
    if (accumulate) {
        if (b == e) return;
        a0 = dst[0]; a1 = dst[comp_stride]; a2 = dst[2 * comp_stride];
    } else {
        a0 = a1 = a2 = 0.;
    }
    int c = b;
    for (; c + 4 <= e; c += 4) {          // loads of four contributions in flight, added in order
        const double *r0 = rhs_of(sv, codes[c]), *r1 = rhs_of(sv, codes[c + 1]);
        const double *r2 = rhs_of(sv, codes[c + 2]), *r3 = rhs_of(sv, codes[c + 3]);
        const double x0 = r0[0], y0 = r0[1], z0 = r0[2], x1 = r1[0], y1 = r1[1], z1 = r1[2];
        const double x2 = r2[0], y2 = r2[1], z2 = r2[2], x3 = r3[0], y3 = r3[1], z3 = r3[2];
        a0 += x0; a1 += y0; a2 += z0;
        a0 += x1; a1 += y1; a2 += z1;
        a0 += x2; a1 += y2; a2 += z2;
        a0 += x3; a1 += y3; a2 += z3;
    }
    for (; c < e; c++) {
        const double *r = rhs_of(sv, codes[c]);
        a0 += r[0]; a1 += r[1]; a2 += r[2];
    }
    dst[0] = a0; dst[comp_stride] = a1; dst[2 * comp_stride] = a2;
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
__global__ __launch_bounds__(256) void k_count_flags(const int *__restrict__ damaged, const int *__restrict__ failed,
                                                     int n, int *__restrict__ out)
{
    int d = 0, f = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
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

}  // namespace icf

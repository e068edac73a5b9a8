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
template <bool MIXED, bool ISO>
__global__ __launch_bounds__(256) void k_reduce_vals(const int *__restrict__ off, const uint32_t *__restrict__ codes,
                                                     int nkeys, double *__restrict__ A, int accumulate, Stage sv)
{
    using T = typename std::conditional<MIXED, float, double>::type;
    const int s = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (s >= nkeys) return;
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
    for (int c = b; c < e; c++) {
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
                                                     int nkeys, double *__restrict__ out, int comp_stride, int elem_stride,
                                                     int accumulate, Stage sv)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nkeys) return;
    const int b = off[s], e = off[s + 1];
    double *dst = out + (size_t)s * elem_stride;
    double a0, a1, a2;
    if (accumulate) {
        if (b == e) return;
        a0 = dst[0]; a1 = dst[comp_stride]; a2 = dst[2 * comp_stride];
    } else {
        a0 = a1 = a2 = 0.;
    }
    for (int c = b; c < e; c++) {
        const double *r = rhs_of(sv, codes[c]);
        a0 += r[0]; a1 += r[1]; a2 += r[2];
    }
    dst[0] = a0; dst[comp_stride] = a1; dst[2 * comp_stride] = a2;
}

// Integer sum for the CZ flag counts (GPU_Functionality.Sum, :93-107).
__global__ __launch_bounds__(256) void k_sum_int(const int *__restrict__ v, int n, int *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int x = i < n ? v[i] : 0;
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if ((threadIdx.x & 63) == 0 && x != 0) atomicAdd(out, x);
}

}  // namespace icf

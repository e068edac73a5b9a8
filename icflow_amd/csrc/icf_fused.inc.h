// icf_fused.inc.h -- independent phases of one call in one launch (included by icf_lib.hip after
// icf_reduce.inc.h).
//
// The first workgroups of a launch run one phase and the rest the next one(s); each phase sees
// its own workgroup index and count, as if launched alone.  Phase boundaries are multiples of 8
// workgroups, so every phase keeps its XCD mapping (workgroup w runs on XCD w % 8).  Phases that
// write different arrays and read only what earlier launches wrote then share the GPU without a
// second stream, and without the cross-queue waits a fork and join cost.
#pragma once

namespace icf {

// Phase 1 of the fused call: tets (workgroups [0, nbt)), then CZs.  Workgroup 0 also clears the
// two CZ counters that the last launch of the call adds up.
template <bool MIXED, bool ISO>
__global__ __launch_bounds__(256) void k_phase1_fused(int nbt, const int *__restrict__ ie, const double *__restrict__ dn,
                                                      int nt, int tp, int es, int ns, double *__restrict__ tet_rhs,
                                                      void *__restrict__ tet_tan, double *__restrict__ dcz,
                                                      int *__restrict__ icz, int nc, int cs,
                                                      double *__restrict__ cz_rhs, void *__restrict__ cz_tan,
                                                      int *__restrict__ counts)
{
    if (blockIdx.x == 0 && threadIdx.x < 2) counts[threadIdx.x] = 0;
    if ((int)blockIdx.x < nbt)
        tet_phase1<MIXED, ISO>(blockIdx.x, nbt, ie, dn, nt, tp, es, ns, tet_rhs, tet_tan);
    else
        cz_phase1<MIXED>(blockIdx.x - nbt, gridDim.x - nbt, dcz, icz, dn, nc, cs, ns, cz_rhs, cz_tan);
}

// Node forces with each node's tet sum (k_reduce_vec3<true, true>; workgroups [0, nbf), the
// short phase first), then the BCSR values of the fused call (overwrite).  Workgroups of KEYS
// lanes: one key per lane in both phases.
template <bool MIXED, bool ISO, typename L, int KEYS, int STAGE>
__global__ __launch_bounds__(KEYS) void k_vals_force_fused(int nbf, const int *__restrict__ off,
                                                           const uint32_t *__restrict__ codes,
                                                           const int2 *__restrict__ chunks, int nchunks,
                                                           double *__restrict__ A, const int *__restrict__ foff,
                                                           const uint32_t *__restrict__ fcodes,
                                                           const int *__restrict__ fkeys, int nfkeys,
                                                           double *__restrict__ F, int ns, Stage sv,
                                                           double *__restrict__ tsum)
{
    if ((int)blockIdx.x < nbf)
        reduce_vec3<true, true>(blockIdx.x, nbf, foff, fcodes, fkeys, nfkeys, F, ns, 1, sv, tsum);
    else
        reduce_vals<MIXED, ISO, false, L, KEYS, STAGE>(blockIdx.x - nbf, gridDim.x - nbf, off, codes, nullptr, chunks,
                                                nchunks, A, sv);
}

// Node forces with the rhs rows (reduce_force_rhs; workgroups [0, nbf)), the CZ counters (the
// next nbc), then the BCSR values of the fused call (overwrite): the whole second half of the
// fused call as one launch.  Workgroups of KEYS lanes.
template <bool MIXED, bool ISO, typename L, int KEYS, int STAGE>
__global__ __launch_bounds__(KEYS) void k_vals_force_rhs_fused(int nbf, int nbc, const int *__restrict__ off,
        const uint32_t *__restrict__ codes, const int2 *__restrict__ chunks, int nchunks, double *__restrict__ A,
        const int *__restrict__ foff, const uint32_t *__restrict__ fcodes, const int *__restrict__ fkeys, int nfkeys,
        double *__restrict__ F, int ns, const int *__restrict__ node_row, const int *__restrict__ roff,
        const uint32_t *__restrict__ rcodes, const int *__restrict__ cz_begin, double *__restrict__ b,
        const int *__restrict__ damaged, const int *__restrict__ failed, int nc, int *__restrict__ counts, Stage sv)
{
    const int w = blockIdx.x;
    if (w < nbf)
        reduce_force_rhs(w, nbf, foff, fcodes, fkeys, nfkeys, F, ns, node_row, roff, rcodes, cz_begin, b, sv);
    else if (w < nbf + nbc)
        count_flags_n<KEYS>(w - nbf, nbc, damaged, failed, nc, counts);
    else
        reduce_vals<MIXED, ISO, false, L, KEYS, STAGE>(w - nbf - nbc, gridDim.x - nbf - nbc, off, codes, nullptr, chunks,
                                                nchunks, A, sv);
}

// rhs rows from the tet sums plus their CZ terms (workgroups [0, nbr)), then the CZ counters.
__global__ __launch_bounds__(256) void k_rhs_counts_fused(int nbr, const int *__restrict__ off,
                                                          const uint32_t *__restrict__ codes,
                                                          const int *__restrict__ cz_begin,
                                                          const int *__restrict__ row_node,
                                                          const double *__restrict__ tsum, int N,
                                                          double *__restrict__ b, Stage sv,
                                                          const int *__restrict__ damaged,
                                                          const int *__restrict__ failed, int nc,
                                                          int *__restrict__ counts)
{
    if ((int)blockIdx.x < nbr)
        reduce_rhs_tsum(blockIdx.x, nbr, off, codes, cz_begin, row_node, tsum, N, b, sv);
    else
        count_flags(blockIdx.x - nbr, gridDim.x - nbr, damaged, failed, nc, counts);
}

// The impact sums of the collision response, added into node forces (workgroups [0, nb0)),
// BCSR blocks (the next nb1) and rhs rows (the rest); KEYS lanes per workgroup, one key each.
template <bool MIXED, typename L, int KEYS, int STAGE>
__global__ __launch_bounds__(KEYS) void k_contact_sums(int nb0, int nb1, const int *__restrict__ foff,
                                                       const uint32_t *__restrict__ fcodes,
                                                       const int *__restrict__ fkeys, int nfkeys,
                                                       double *__restrict__ F, int ns, const int *__restrict__ voff,
                                                       const uint32_t *__restrict__ vcodes,
                                                       const int *__restrict__ vkeys, const int2 *__restrict__ chunks,
                                                       int nchunks, double *__restrict__ A,
                                                       const int *__restrict__ roff,
                                                       const uint32_t *__restrict__ rcodes,
                                                       const int *__restrict__ rkeys, int nrkeys,
                                                       double *__restrict__ b, Stage sv)
{
    const int w = blockIdx.x;
    if (w < nb0)
        reduce_vec3<true, false>(w, nb0, foff, fcodes, fkeys, nfkeys, F, ns, 1, sv, nullptr);
    else if (w < nb0 + nb1)
        reduce_vals<MIXED, false, true, L, KEYS, STAGE>(w - nb0, nb1, voff, vcodes, vkeys, chunks, nchunks, A, sv);
    else
        reduce_vec3<true, false>(w - nb0 - nb1, gridDim.x - nb0 - nb1, roff, rcodes, rkeys, nrkeys, b, 1, 3, sv,
                                 nullptr);
}

}  // namespace icf

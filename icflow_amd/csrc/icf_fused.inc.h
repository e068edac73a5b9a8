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

// Two phases interleaved in one launch: rounds of 8 consecutive workgroups (one per XCD) go to
// phase A (na workgroups, a multiple of 8) or phase B (nb workgroups), B's rounds spread evenly
// among A's, so that work that waits on memory and work that keeps the VALUs busy share the
// CUs for the whole launch instead of one after the other.  The grid has 8 (na/8 + ceil(nb/8))
// workgroups; `local` is the workgroup's index within its phase (round * 8 + XCD, so each phase
// keeps its own XCD mapping); B's padding workgroups get local >= nb.
__host__ __device__ inline int interleave2_grid(int na, int nb) { return na + (nb + 7) / 8 * 8; }
__device__ inline bool interleave2(int w, int na, int nb, int &local)
{
    const long long ra = na / 8, rb = (nb + 7) / 8, R = w / 8;
    const int x = w % 8;
    if (ra + rb == 0) {
        local = w;
        return false;
    }
    const long long b_before = R * rb / (ra + rb), b_upto = (R + 1) * rb / (ra + rb);
    if (b_upto > b_before) {
        local = (int)(b_before * 8 + x);
        return true;
    }
    local = (int)((R - b_before) * 8 + x);
    return false;
}

// interleave2 when `ilv`, else phase A's workgroups first, then phase B's (same grid; B's
// padding workgroups get local >= nb).
__device__ inline bool split2(int w, int na, int nb, bool ilv, int &local)
{
    if (ilv) return interleave2(w, na, nb, local);
    if (w < na) {
        local = w;
        return false;
    }
    local = w - na;
    return true;
}

// Phase 1 of the fused call: tets (nbt workgroups) interleaved with CZs (nbc).  Workgroup 0 also
// clears the two CZ counters that the last launch of the call adds up.
template <bool MIXED, bool ISO, bool NT>
__global__ __launch_bounds__(256) void k_phase1_fused(int nbt, int nbc, const int *__restrict__ ie,
                                                      const double *__restrict__ dn, int nt, int tp, int es, int ns,
                                                      double *__restrict__ tet_rhs, void *__restrict__ tet_tan,
                                                      double *__restrict__ dcz, int *__restrict__ icz, int nc, int cs,
                                                      double *__restrict__ cz_rhs, void *__restrict__ cz_tan,
                                                      int *__restrict__ counts)
{
    if (blockIdx.x == 0 && threadIdx.x < 2) counts[threadIdx.x] = 0;
    int local;
    if (!interleave2(blockIdx.x, nbt, nbc, local))
        tet_phase1<MIXED, ISO, NT>(local, nbt, ie, dn, nt, tp, es, ns, tet_rhs, tet_tan);
    else if (local < nbc)
        cz_phase1<MIXED>(local, nbc, dcz, icz, dn, nc, cs, ns, cz_rhs, cz_tan);
}

// Node forces with each node's tet sum (k_reduce_vec3<true, true>; nbf workgroups) interleaved
// with the BCSR values of the fused call (overwrite; nchunks workgroups).  Workgroups of KEYS
// lanes: one key per lane in both phases.
template <bool MIXED, bool ISO, typename L, int KEYS, int STAGE>
__global__ __launch_bounds__(KEYS) void k_vals_force_fused(int nbf, const int *__restrict__ off,
                                                           const uint32_t *__restrict__ codes,
                                                           const int2 *__restrict__ chunks, int nchunks,
                                                           double *__restrict__ A, const int *__restrict__ foff,
                                                           const uint32_t *__restrict__ fcodes,
                                                           const int *__restrict__ fkeys, int nfkeys,
                                                           double *__restrict__ F, int ns, Stage sv,
                                                           double *__restrict__ tsum, int ilv)
{
    int local;
    if (!split2(blockIdx.x, nbf, nchunks, ilv != 0, local))
        reduce_vec3<true, true>(local, nbf, foff, fcodes, fkeys, nfkeys, F, ns, 1, sv, tsum);
    else if (local < nchunks)
        reduce_vals<MIXED, ISO, false, L, KEYS, STAGE>(local, nchunks, off, codes, nullptr, chunks, nchunks, A, sv);
}

// The CZ counters (workgroups [0, nbc); per specimen when cz_off is given: the CZs of specimen s
// are cz_off[s] .. cz_off[s + 1], its counts go to spec_counts[2 s], [2 s + 1] and into the
// totals), then node forces with the rhs rows (reduce_force_rhs;
// nbf workgroups) interleaved with the BCSR values of the fused call (overwrite; nchunks
// workgroups): the whole second half of the fused call as one launch.  Workgroups of KEYS lanes.
template <bool MIXED, bool ISO, typename L, int KEYS, int STAGE>
__global__ __launch_bounds__(KEYS) void k_vals_force_rhs_fused(int nbf, int nbc, const int *__restrict__ off,
        const uint32_t *__restrict__ codes, const int2 *__restrict__ chunks, int nchunks, double *__restrict__ A,
        const int *__restrict__ foff, const uint32_t *__restrict__ fcodes, const int *__restrict__ fkeys, int nfkeys,
        double *__restrict__ F, int ns, const int *__restrict__ node_row, const int *__restrict__ roff,
        const uint32_t *__restrict__ rcodes, const int *__restrict__ cz_begin, double *__restrict__ b,
        const int *__restrict__ damaged, const int *__restrict__ failed, int nc, int *__restrict__ counts,
        const int *__restrict__ cz_off, int nspec, int *__restrict__ spec_counts, Stage sv, int ilv)
{
    const int w = blockIdx.x;
    int local;
    if (w < nbc) {
        if (cz_off) count_flags_spec<KEYS>(w, nbc, damaged, failed, cz_off, nspec, spec_counts, counts);
        else count_flags_n<KEYS>(w, nbc, damaged, failed, nc, counts);
    }
    else if (!split2(w - nbc, nbf, nchunks, ilv != 0, local))
        reduce_force_rhs(local, nbf, foff, fcodes, fkeys, nfkeys, F, ns, node_row, roff, rcodes, cz_begin, b, sv);
    else if (local < nchunks)
        reduce_vals<MIXED, ISO, false, L, KEYS, STAGE>(local, nchunks, off, codes, nullptr, chunks, nchunks, A, sv);
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

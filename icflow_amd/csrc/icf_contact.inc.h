// icf_contact.inc.h -- contact detection and penalty response (included by icf_lib.hip).
//
// Restated from collision_kernels9/narrowphase.h:3-201 (Bvalues, ctest, kNarrowPhase_new,
// kFindClosestFace), dist_node2tri.h:9-129 (dtn), point_triange_v2.h:13-842 (pt and its
// helpers) and collision_response.h:38-126 (kCollisionResponseForce).
//
// The narrow phase and the closest-face search produce integers from floating-point
// comparisons; they are evaluated without FMA contraction so that they round exactly like
// the reference expression order and match the CPU oracle bit for bit.
#pragma once

namespace icf {

constexpr double NP_EPS = 1e-10;  // narrowphase.h:1

__device__ inline void tet_inverse(const double *p, double bm[9])
{
#pragma clang fp contract(off)
    double a11 = p[3] - p[0], a12 = p[6] - p[0], a13 = p[9] - p[0];
    double a21 = p[4] - p[1], a22 = p[7] - p[1], a23 = p[10] - p[1];
    double a31 = p[5] - p[2], a32 = p[8] - p[2], a33 = p[11] - p[2];
    double det = a31 * (-a13 * a22 + a12 * a23) + a32 * (a13 * a21 - a11 * a23) + a33 * (-a12 * a21 + a11 * a22);
    bm[0] = (-a23 * a32 + a22 * a33) / det;
    bm[1] = (a13 * a32 - a12 * a33) / det;
    bm[2] = (-a13 * a22 + a12 * a23) / det;
    bm[3] = (a23 * a31 - a21 * a33) / det;
    bm[4] = (-a13 * a31 + a11 * a33) / det;
    bm[5] = (a13 * a21 - a11 * a23) / det;
    bm[6] = (-a22 * a31 + a21 * a32) / det;
    bm[7] = (a12 * a31 - a11 * a32) / det;
    bm[8] = (-a12 * a21 + a11 * a22) / det;
}

__device__ inline bool strictly_inside(const double bm[9], double x, double y, double z)
{
#pragma clang fp contract(off)
    double l1 = x * bm[0] + y * bm[1] + z * bm[2];
    double l2 = x * bm[3] + y * bm[4] + z * bm[5];
    double l3 = x * bm[6] + y * bm[7] + z * bm[8];
    return (l1 > NP_EPS && l2 > NP_EPS && l3 > NP_EPS && (l1 + l2 + l3) < (1 - NP_EPS));
}

// One thread per candidate pair of surface tets (narrowphase.h:46-155).
__global__ __launch_bounds__(256) void k_narrow_phase(int nPairs, const double *__restrict__ dn,
                                                      const int *__restrict__ ie, int *__restrict__ nl,
                                                      int es, int ns)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nPairs) return;
    const int t0 = nl[2 * u], t1 = nl[2 * u + 1];
    int ids[8];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        ids[i] = ie[t0 + es * i];
        ids[4 + i] = ie[t1 + es * i];
    }
    bool shared = false;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 4; j < 8; j++) shared |= (ids[i] == ids[j]);
    if (shared) {
        nl[2 * u] = 0;
        return;
    }
    double p[24];
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) p[3 * i + k] = dn[ids[i] + ns * (ND_XC + k)];
    double b0[9], b1[9];
    tet_inverse(&p[0], b0);
    tet_inverse(&p[12], b1);
    int mask = 0;
    {
#pragma clang fp contract(off)
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (strictly_inside(b0, p[12 + 3 * i] - p[0], p[13 + 3 * i] - p[1], p[14 + 3 * i] - p[2])) mask |= 1 << i;
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (strictly_inside(b1, p[3 * i] - p[12], p[3 * i + 1] - p[13], p[3 * i + 2] - p[14])) mask |= 16 << i;
    }
    nl[2 * u] = mask;
}

__device__ inline double clamp01(double v) { return v <= 0 ? 0 : v >= 1 ? 1 : v; }

// Squared point-triangle distance, Eberly regions (dist_node2tri.h:14-129).
__device__ inline double point_tri_dsq(const double f[9], const double p[3])
{
#pragma clang fp contract(off)
    double e0[3] = {f[3] - f[0], f[4] - f[1], f[5] - f[2]};
    double e1[3] = {f[6] - f[0], f[7] - f[1], f[8] - f[2]};
    double v0[3] = {f[0] - p[0], f[1] - p[1], f[2] - p[2]};
    double a = e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2];
    double b = e0[0] * e1[0] + e0[1] * e1[1] + e0[2] * e1[2];
    double c = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
    double d = e0[0] * v0[0] + e0[1] * v0[1] + e0[2] * v0[2];
    double e = e1[0] * v0[0] + e1[1] * v0[1] + e1[2] * v0[2];
    double det = a * c - b * b;
    double s = b * e - c * d;
    double t = b * d - a * e;
    if (s + t < det) {
        if (s < 0) {
            if (t < 0) {
                if (d < 0) { s = clamp01(-d / a); t = 0; }
                else { s = 0; t = clamp01(-e / c); }
            } else { s = 0; t = clamp01(-e / c); }
        } else if (t < 0) { s = clamp01(-d / a); t = 0; }
        else { double inv = 1.0 / det; s *= inv; t *= inv; }
    } else {
        if (s < 0) {
            double t0 = b + d, t1 = c + e;
            if (t1 > t0) { double num = t1 - t0, den = a - 2 * b + c; s = clamp01(num / den); t = 1 - s; }
            else { t = clamp01(-e / c); s = 0; }
        } else if (t < 0) {
            if (a + d > b + e) { double num = c + e - b - d, den = a - 2 * b + c; s = clamp01(num / den); t = 1 - s; }
            else { s = clamp01(-e / c); t = 0; }
        } else { double num = c + e - b - d, den = a - 2 * b + c; s = clamp01(num / den); t = 1 - s; }
    }
    double q0 = f[0] + s * e0[0] + t * e1[0] - p[0];
    double q1 = f[1] + s * e0[1] + t * e1[1] - p[1];
    double q2 = f[2] + s * e0[2] + t * e1[2] - p[2];
    return q0 * q0 + q1 * q1 + q2 * q2;
}

// Compact copy of the surface tets' adjacent-face lists (ie row 4 = count, rows 5.. = face
// ids; one SoA row per list position): tet t's list at ids[off[t] .. off[t+1]).  The closest
// face search then reads a tuple's list from one place instead of one row per face.
__global__ __launch_bounds__(256) void k_adj_count(const int *__restrict__ ie, int es, int max_rows,
                                                   int *__restrict__ cnt)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= es) return;
    const int n = ie[t + (size_t)es * 4];
    cnt[t] = n < 0 ? 0 : (n > max_rows ? max_rows : n);
}

__global__ __launch_bounds__(256) void k_adj_fill(const int *__restrict__ ie, int es, const int *__restrict__ off,
                                                  int *__restrict__ ids)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= es) return;
    const int o = off[t], n = off[t + 1] - o;
    for (int k = 0; k < n; k++) ids[o + k] = ie[t + (size_t)es * (5 + k)];
}

// (node, tet) tuples (narrowphase.h:159-201), CF_GROUP lanes per tuple (4, 8 or 16: the
// autotuned shape; at least 3, the lanes that write the face's ids): lane l evaluates the
// adjacent faces k = l, l + CF_GROUP, ...; the group then takes the lexicographic minimum of
// (distance, k).  That is the reference's sequential "first strict minimum" rule: a NaN
// distance at k > 0 never compares smaller (it is ranked +inf), and a NaN at k = 0 keeps
// face 0 (ranked -inf).  A tet without adjacent faces (not produced by the reference's
// surface detection) yields -1 ids instead of reading faces[-1].
template <int CF_GROUP>
__global__ __launch_bounds__(256) void k_closest_face(int nPairs, int ts, int *__restrict__ nl2,
                                                      const double *__restrict__ dn,
                                                      const int *__restrict__ adj_off,
                                                      const int *__restrict__ adj_ids,
                                                      const int *__restrict__ faces, int ns, int fs)
{
    const int u = (blockIdx.x * blockDim.x + threadIdx.x) / CF_GROUP;
    const int lane = threadIdx.x % CF_GROUP;
    const bool live = u < nPairs;
    double key = __builtin_inf();
    int bestk = 0x7fffffff, bestf = -1;
    if (live) {
        const int node = nl2[u], tet = nl2[u + ts];
        double p[3];
#pragma unroll
        for (int k = 0; k < 3; k++) p[k] = dn[node + ns * (ND_XC + k)];
        const int a0 = adj_off[tet], nf = adj_off[tet + 1] - a0;
        for (int k = lane; k < nf; k += CF_GROUP) {
            const int fid = adj_ids[a0 + k];
            double fc[9];
#pragma unroll
            for (int v = 0; v < 3; v++) {
                const int fn = faces[fid + fs * v];
#pragma unroll
                for (int j = 0; j < 3; j++) fc[3 * v + j] = dn[fn + ns * (ND_XC + j)];
            }
            double d = point_tri_dsq(fc, p);
            if (d != d) d = (k == 0) ? -__builtin_inf() : __builtin_inf();
            if (d < key || (d == key && k < bestk)) { key = d; bestk = k; bestf = fid; }
        }
    }
    // lexicographic (key, k) minimum over the group (all lanes take part in the shuffles)
#pragma unroll
    for (int off = CF_GROUP / 2; off > 0; off >>= 1) {
        const double ok = __shfl_xor(key, off, CF_GROUP);
        const int okk = __shfl_xor(bestk, off, CF_GROUP), of = __shfl_xor(bestf, off, CF_GROUP);
        if (ok < key || (ok == key && okk < bestk)) { key = ok; bestk = okk; bestf = of; }
    }
    if (!live || lane >= 3) return;
    nl2[u + ts * (1 + lane)] = bestf >= 0 ? faces[bestf + fs * lane] : -1;
}

// Phase 2: block (a,b) of an impact's block record (the transpose of block (b,a) when a > b).
template <typename T>
__device__ inline void contact_block(double acc[9], const T *__restrict__ tan, int a, int b)
{
    const T *blk = tan + ct_block(a < b ? a : b, a < b ? b : a) * 9;
    T v[9];
#pragma unroll
    for (int i = 0; i < 9; i++) v[i] = blk[i];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int l = 0; l < 3; l++) acc[3 * k + l] += (double)(a <= b ? v[3 * k + l] : v[3 * l + k]);
}

}  // namespace icf

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

// One thread per (node, tet) tuple (narrowphase.h:159-201).  A tet without adjacent faces
// (not produced by the reference's surface detection) yields -1 ids instead of reading
// faces[-1].
__global__ __launch_bounds__(256) void k_closest_face(int nPairs, int ts, int *__restrict__ nl2,
                                                      const double *__restrict__ dn, const int *__restrict__ ie,
                                                      const int *__restrict__ faces, int ns, int es, int fs)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nPairs) return;
    const int node = nl2[u], tet = nl2[u + ts];
    double p[3];
#pragma unroll
    for (int k = 0; k < 3; k++) p[k] = dn[node + ns * (ND_XC + k)];
    const int nf = ie[tet + es * 4];
    int best = -1;
    double bestd = 0;
    for (int k = 0; k < nf; k++) {
        const int fid = ie[tet + es * (5 + k)];
        double fc[9];
#pragma unroll
        for (int v = 0; v < 3; v++) {
            const int fn = faces[fid + fs * v];
#pragma unroll
            for (int j = 0; j < 3; j++) fc[3 * v + j] = dn[fn + ns * (ND_XC + j)];
        }
        const double d = point_tri_dsq(fc, p);
        if (k == 0 || d < bestd) { best = fid; bestd = d; }
    }
#pragma unroll
    for (int v = 0; v < 3; v++) nl2[u + ts * (1 + v)] = best >= 0 ? faces[best + fs * v] : -1;
}

// ---- point-triangle squared distance with derivatives (point_triange_v2.h) ---------------
// The Hessian is never held whole: entries are produced one at a time and the upper
// triangle, times k/2, is written into the impact's record (the reference keeps the upper
// triangle and mirrors it, collision_response.h:111-114).

struct HessOut {
    double *rec;   // 78 packed upper entries
    double half_k;
    __device__ inline void put(int r, int c, double v) const
    {
        if (r <= c) rec[upper12(r, c)] = half_k * v;
    }
};

// vertex-vertex branch: |x_i - x_j|^2 (point_triange_v2.h:115-150)
__device__ inline double pt_vv(const double x[12], int i, int j, double fd[12], const HessOut &H)
{
    double d[3];
#pragma unroll
    for (int k = 0; k < 3; k++) d[k] = x[3 * i + k] - x[3 * j + k];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        fd[3 * i + k] = 2 * d[k];
        fd[3 * j + k] = -fd[3 * i + k];
        H.put(3 * i + k, 3 * i + k, 2);
        H.put(3 * j + k, 3 * j + k, 2);
        H.put(3 * i + k, 3 * j + k, -2);
        H.put(3 * j + k, 3 * i + k, -2);
    }
    return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

// vertex-edge branch: distance^2 from x[0..2] to the line through points i and j, nearer
// endpoint first; exact derivatives of f1 - g/f2 (point_triange_v2.h:152-294).
__device__ inline double pt_ve(const double x[12], int i, int j, double fd[12], const HessOut &H)
{
    int a = 3 * i, b = 3 * j;
    double pa = 0, pb = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        pa += (x[k] - x[a + k]) * (x[k] - x[a + k]);
        pb += (x[k] - x[b + k]) * (x[k] - x[b + k]);
    }
    if (pa > pb) { int t = a; a = b; b = t; }
    const double y[9] = {x[0], x[1], x[2], x[a], x[a + 1], x[a + 2], x[b], x[b + 1], x[b + 2]};
    const int map[3] = {0, a, b};

    const double L2 = (y[6] - y[3]) * (y[6] - y[3]) + (y[7] - y[4]) * (y[7] - y[4]) + (y[8] - y[5]) * (y[8] - y[5]);
    const double tt = -((y[3] - y[0]) * (y[6] - y[3]) + (y[4] - y[1]) * (y[7] - y[4]) + (y[5] - y[2]) * (y[8] - y[5])) / L2;
    const double uu = 1 - tt;
    double value = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double r = y[k] - uu * y[3 + k] - tt * y[6 + k];
        value += r * r;
    }
    const double q = (y[3] - y[0]) * (y[6] - y[3]) + (y[4] - y[1]) * (y[7] - y[4]) + (y[5] - y[2]) * (y[8] - y[5]);
    double qd[9], f1d[9], f2d[9];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        qd[k] = y[3 + k] - y[6 + k];
        qd[3 + k] = y[k] - 2 * y[3 + k] + y[6 + k];
        qd[6 + k] = y[3 + k] - y[k];
        const double e1 = y[3 + k] - y[k], e2 = y[6 + k] - y[3 + k];
        f1d[k] = -2 * e1; f1d[3 + k] = 2 * e1; f1d[6 + k] = 0;
        f2d[k] = 0; f2d[3 + k] = -2 * e2; f2d[6 + k] = 2 * e2;
    }
    const double g = q * q, f2 = L2, f2sq = f2 * f2, f2cu = f2sq * f2;
    // constant second-derivative node patterns: q (sp_dot3), f1 = |y1-y0|^2, f2 = |y2-y1|^2
    const double QP[3][3] = {{0, 1, -1}, {1, -2, 1}, {-1, 1, 0}};
    const double F1P[3][3] = {{2, -2, 0}, {-2, 2, 0}, {0, 0, 0}};
    const double F2P[3][3] = {{0, 0, 0}, {0, 2, -2}, {0, -2, 2}};
#pragma unroll
    for (int I = 0; I < 9; I++) {
        const double gdI = 2 * q * qd[I];
        fd[map[I / 3] + I % 3] = f1d[I] + (g * f2d[I] - f2 * gdI) / f2sq;
#pragma unroll
        for (int J = 0; J < 9; J++) {
            const int r = map[I / 3] + I % 3, c = map[J / 3] + J % 3;
            if (r > c) continue;
            const bool same = (I % 3) == (J % 3);
            const double qdd = same ? QP[I / 3][J / 3] : 0, f1dd = same ? F1P[I / 3][J / 3] : 0;
            const double f2dd = same ? F2P[I / 3][J / 3] : 0;
            const double gdJ = 2 * q * qd[J];
            const double gdd = 2 * (qd[I] * qd[J] + q * qdd);
            const double v = -2 * g * f2d[I] * f2d[J] / f2cu + (gdI * f2d[J] + gdJ * f2d[I]) / f2sq + f1dd +
                             g * f2dd / f2sq - gdd / f2;
            H.put(r, c, v);
        }
    }
    return value;
}

// interior branch: point-plane distance^2, chain rule through s(x), t(x)
// (point_triange_v2.h:449-582; tables a2..e2 :298-371 as 4x4 node patterns x I3).
__device__ inline double pt_pp(const double x[12], double fd[12], const HessOut &H)
{
    const double *p = &x[0], *v0 = &x[3], *v1 = &x[6], *v2 = &x[9];
    double E1[3], E2[3], W[3];
#pragma unroll
    for (int k = 0; k < 3; k++) { E1[k] = v1[k] - v0[k]; E2[k] = v2[k] - v0[k]; W[k] = v0[k] - p[k]; }
    const double a = dot3(E1, E1), b = dot3(E2, E1), c = dot3(E2, E2), d = dot3(W, E1), e = dot3(W, E2);
    const double det = a * c - b * b, det2 = det * det, det3 = det2 * det;
    const double Ns = b * e - c * d, Nt = b * d - a * e;
    const double s = Ns * (1. / det), t = Nt * (1. / det), u = 1 - (s + t);
    // first derivatives of a..e over (p, v0, v1, v2)
    double a1[12], b1[12], c1[12], d1[12], e1[12];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        a1[k] = 0;      a1[3 + k] = -2 * E1[k];       a1[6 + k] = 2 * E1[k]; a1[9 + k] = 0;
        b1[k] = 0;      b1[3 + k] = -E1[k] - E2[k];   b1[6 + k] = E2[k];     b1[9 + k] = E1[k];
        c1[k] = 0;      c1[3 + k] = -2 * E2[k];       c1[6 + k] = 0;         c1[9 + k] = 2 * E2[k];
        d1[k] = -E1[k]; d1[3 + k] = E1[k] - W[k];     d1[6 + k] = W[k];      d1[9 + k] = 0;
        e1[k] = -E2[k]; e1[3 + k] = E2[k] - W[k];     e1[6 + k] = 0;         e1[9 + k] = W[k];
    }
    const double A2[4][4] = {{0, 0, 0, 0}, {0, 2, -2, 0}, {0, -2, 2, 0}, {0, 0, 0, 0}};
    const double B2[4][4] = {{0, 0, 0, 0}, {0, 2, -1, -1}, {0, -1, 0, 1}, {0, -1, 1, 0}};
    const double C2[4][4] = {{0, 0, 0, 0}, {0, 2, 0, -2}, {0, 0, 0, 0}, {0, -2, 0, 2}};
    const double D2[4][4] = {{0, 1, -1, 0}, {1, -2, 1, 0}, {-1, 1, 0, 0}, {0, 0, 0, 0}};
    const double E2p[4][4] = {{0, 1, 0, -1}, {1, -2, 0, 1}, {0, 0, 0, 0}, {-1, 1, 0, 0}};

    double s1[12], t1[12], u1[12], dt1[12];
    double r[3];
#pragma unroll
    for (int k = 0; k < 3; k++) r[k] = v1[k] * s + v2[k] * t + v0[k] * u - p[k];
    const double val = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
#pragma unroll
    for (int i = 0; i < 12; i++) {
        dt1[i] = c * a1[i] + a * c1[i] - 2 * b * b1[i];
        s1[i] = ((c * d - b * e) * dt1[i]) / det2 + ((e * b1[i] + b * e1[i]) - (d * c1[i] + c * d1[i])) / det;
        t1[i] = ((a * e - b * d) * dt1[i]) / det2 + ((d * b1[i] + b * d1[i]) - (a * e1[i] + e * a1[i])) / det;
        u1[i] = -(s1[i] + t1[i]);
    }
    double dr[3][12];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int i = 0; i < 12; i++) {
            double v = v1[k] * s1[i] + v2[k] * t1[i] + v0[k] * u1[i];
            const int node = i / 3;
            if (i % 3 == k) v += (node == 0 ? -1.0 : node == 1 ? u : node == 2 ? s : t);
            dr[k][i] = v;
        }
#pragma unroll
    for (int i = 0; i < 12; i++) fd[i] = 2 * (r[0] * dr[0][i] + r[1] * dr[1][i] + r[2] * dr[2][i]);

    for (int i = 0; i < 12; i++)
        for (int j = i; j < 12; j++) {
            const int ni = i / 3, nj = j / 3;
            const bool same = (i % 3) == (j % 3);
            const double a2 = same ? A2[ni][nj] : 0, b2 = same ? B2[ni][nj] : 0, c2 = same ? C2[ni][nj] : 0;
            const double d2 = same ? D2[ni][nj] : 0, e2 = same ? E2p[ni][nj] : 0;
            const double ddet = -2 * b1[i] * b1[j] + a1[j] * c1[i] + a1[i] * c1[j] + c * a2 - 2 * b * b2 + a * c2;
            const double s2 = (-(c1[j] * d1[i]) - c1[i] * d1[j] + b1[j] * e1[i] + b1[i] * e1[j] + e * b2 - d * c2 - c * d2 + b * e2) / det -
                              ((dt1[j] * (e * b1[i] - d * c1[i] - c * d1[i] + b * e1[i])) +
                               (dt1[i] * (e * b1[j] - d * c1[j] - c * d1[j] + b * e1[j])) + (Ns * ddet)) / det2 +
                              (2 * Ns * dt1[i] * dt1[j]) / det3;
            const double t2 = (b1[j] * d1[i] + b1[i] * d1[j] - a1[j] * e1[i] - a1[i] * e1[j] - e * a2 + d * b2 + b * d2 - a * e2) / det -
                              ((dt1[j] * (-(e * a1[i]) + d * b1[i] + b * d1[i] - a * e1[i])) +
                               (dt1[i] * (-(e * a1[j]) + d * b1[j] + b * d1[j] - a * e1[j])) + (Nt * ddet)) / det2 +
                              (2 * Nt * dt1[i] * dt1[j]) / det3;
            const double u2 = -(s2 + t2);
            double acc = 0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                double v = v1[k] * s2 + v2[k] * t2 + v0[k] * u2;
                if (i % 3 == k) v += (ni == 1 ? u1[j] : ni == 2 ? s1[j] : ni == 3 ? t1[j] : 0);
                if (j % 3 == k) v += (nj == 1 ? u1[i] : nj == 2 ? s1[i] : nj == 3 ? t1[i] : 0);
                acc += dr[k][i] * dr[k][j] + r[k] * v;
            }
            H.put(i, j, 2 * acc);
        }
    return val;
}

// Seven-region point-triangle classification (point_triange_v2.h:589-842).
__device__ inline double pt_all(const double x[12], double fd[12], const HessOut &H, double &so, double &to)
{
    double E1[3], E2[3], W[3];
#pragma unroll
    for (int k = 0; k < 3; k++) { E1[k] = x[6 + k] - x[3 + k]; E2[k] = x[9 + k] - x[3 + k]; W[k] = x[3 + k] - x[k]; }
    const double a = dot3(E1, E1), b = dot3(E2, E1), c = dot3(E2, E2), d = dot3(W, E1), e = dot3(W, E2);
    const double det = a * c - b * b;
    double s = b * e - c * d, t = b * d - a * e, dsq;
    if (s + t <= det) {
        if (s < 0) {
            if (t < 0 && d < 0) {
                t = 0;
                if (-d >= a) { s = 1; dsq = pt_vv(x, 0, 2, fd, H); }
                else { s = -d / a; dsq = pt_ve(x, 2, 1, fd, H); }
            } else {
                s = 0;
                if (e >= 0) { t = 0; dsq = pt_vv(x, 0, 1, fd, H); }
                else if (-e >= c) { t = 1; dsq = pt_vv(x, 0, 3, fd, H); }
                else { t = -e / c; dsq = pt_ve(x, 3, 1, fd, H); }
            }
        } else if (t < 0) {
            t = 0;
            if (d >= 0) { s = 0; dsq = pt_vv(x, 0, 1, fd, H); }
            else if (-d >= a) { s = 1; dsq = pt_vv(x, 0, 2, fd, H); }
            else { s = -d / a; dsq = pt_ve(x, 1, 2, fd, H); }
        } else {
            const double inv = 1. / det;
            s *= inv; t *= inv;
            dsq = pt_pp(x, fd, H);
        }
    } else {
        if (s < 0) {
            const double t0 = b + d, t1 = c + e;
            if (t1 > t0) {
                const double num = t1 - t0, den = a - 2 * b + c;
                if (num >= den) { s = 1; t = 0; dsq = pt_vv(x, 0, 2, fd, H); }
                else { s = num / den; t = 1 - s; dsq = pt_ve(x, 2, 3, fd, H); }
            } else {
                s = 0;
                if (t1 <= 0) { t = 1; dsq = pt_vv(x, 0, 3, fd, H); }
                else if (e >= 0) { t = 0; dsq = pt_vv(x, 0, 1, fd, H); }
                else { t = -e / c; dsq = pt_ve(x, 1, 3, fd, H); }
            }
        } else if (t < 0) {
            const double t0 = b + e, t1 = a + d;
            if (t1 > t0) {
                const double num = t1 - t0, den = a - 2 * b + c;
                if (num >= den) { t = 1; s = 0; dsq = pt_vv(x, 0, 3, fd, H); }
                else { t = num / den; s = 1 - t; dsq = pt_ve(x, 2, 3, fd, H); }
            } else {
                t = 0;
                if (t1 <= 0) { s = 1; dsq = pt_vv(x, 0, 2, fd, H); }
                else if (d >= 0) { s = 0; dsq = pt_vv(x, 0, 1, fd, H); }
                else { s = -d / a; dsq = pt_ve(x, 1, 2, fd, H); }
            }
        } else {
            const double num = c + e - b - d;
            if (num <= 0) { s = 0; t = 1; dsq = pt_vv(x, 0, 3, fd, H); }
            else {
                const double den = a - 2 * b + c;
                if (num >= den) { s = 1; t = 0; dsq = pt_vv(x, 0, 2, fd, H); }
                else { s = num / den; t = 1 - s; dsq = pt_ve(x, 2, 3, fd, H); }
            }
        }
    }
    so = s;
    to = t;
    return dsq;
}

// Phase 1 for impacts: record = 78 upper entries of (k/2) Hess d^2, then 12 forces
// (collision_response.h:38-126).  Rigid-rigid impacts and d^2 <= eps^2 stage zeros.
__global__ __launch_bounds__(128) void k_contact_phase1(const int *__restrict__ icr, const double *__restrict__ dn,
                                                        int n, int cs, int ns, double k, double eps,
                                                        double *__restrict__ st)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    double *rec = st + (size_t)u * CT_REC;
    for (int i = 0; i < CT_REC; i++) rec[i] = 0.;
    bool active = false;
#pragma unroll
    for (int i = 0; i < 4; i++) active |= icr[u + cs * (CR_ROWS + i)] >= 0;
    if (!active) return;
    double x[12];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int nd = icr[u + cs * (CR_NODES + i)];
#pragma unroll
        for (int j = 0; j < 3; j++) x[3 * i + j] = dn[nd + ns * (ND_XC + j)];
    }
    double fd[12];
#pragma unroll
    for (int i = 0; i < 12; i++) fd[i] = 0;
    HessOut H{rec, k / 2};
    double s, t;
    const double dsq = pt_all(x, fd, H, s, t);
    if (!(dsq > eps * eps)) {
        for (int i = 0; i < CT_REC; i++) rec[i] = 0.;
        return;
    }
    const double w0 = 1 - (s + t), w1 = s, w2 = t;
    const double f[3] = {k * 0.5 * fd[0], k * 0.5 * fd[1], k * 0.5 * fd[2]};
    double *F = rec + CT_F_OFF;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        F[c] = -f[c];
        F[3 + c] = w0 * f[c];
        F[6 + c] = w1 * f[c];
        F[9 + c] = w2 * f[c];
    }
}

// Phase 2: add block (a,b) of an impact record (symmetric from its upper triangle).
__device__ inline void contact_block(double acc[9], const double *__restrict__ rec, int a, int b)
{
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int l = 0; l < 3; l++) {
            const int r = 3 * a + k, c = 3 * b + l;
            acc[3 * k + l] += r <= c ? rec[upper12(r, c)] : rec[upper12(c, r)];
        }
}

}  // namespace icf

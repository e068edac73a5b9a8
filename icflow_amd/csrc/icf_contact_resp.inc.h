// icf_contact_resp.inc.h -- point-triangle penalty response, phase 1 (included by icf_lib.hip).
//
// Restated from collision_kernels9/collision_response.h:38-126 (kCollisionResponseForce) and
// point_triange_v2.h:13-842 (pt and its helpers).  The reference evaluates one impact per
// thread with 12x12 temporaries; here each thread also takes one impact, but computes every
// upper-triangle entry of the Hessian from the impact's few vectors and stores it at once,
// so no 12x12 array is held.
#pragma once

namespace icf {

__device__ inline double sel3(const double v[3], int c) { return c == 0 ? v[0] : (c == 1 ? v[1] : v[2]); }

// Quantities of one impact shared by its lanes.
struct Impact {
    double p[3], v0[3], v1[3], v2[3];
    int branch;          // 0 interior, 1..3 vertex, 4..7 edge (point_triange_v2.h:589-842)
    double dsq, s, t;
    // vertex branch: vertices (12-index node numbers 0..3) i0, i1
    int i0, i1;
    // edge branch: nodes of the line, nearer endpoint first
    int ea, eb;
    double y0[3], y1[3], y2[3];
    double q, g, f2, f2sq, f2cu, if2, if2sq, if2cu;   // if* = reciprocals
};

// coordinate k of impact node nd (0 = point, 1..3 = triangle vertices), by selection so that
// the Impact stays in registers
__device__ inline double xnode(const Impact &I, int nd, int k)
{
    return nd == 0 ? sel3(I.p, k) : (nd == 1 ? sel3(I.v0, k) : (nd == 2 ? sel3(I.v1, k) : sel3(I.v2, k)));
}

__device__ inline void setup_vv(Impact &I, int i0, int i1)
{
    I.i0 = i0;
    I.i1 = i1;
    double s = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double dd = xnode(I, i0, k) - xnode(I, i1, k);
        s += dd * dd;
    }
    I.dsq = s;
}

// vertex_edge_distance_and_derivs_12 (point_triange_v2.h:237-294): nearer endpoint first.
__device__ inline void setup_ve(Impact &I, int na, int nb)
{
    double pa = 0, pb = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double xa = xnode(I, na, k), xb = xnode(I, nb, k);
        pa += (I.p[k] - xa) * (I.p[k] - xa);
        pb += (I.p[k] - xb) * (I.p[k] - xb);
    }
    if (pa > pb) { int tmp = na; na = nb; nb = tmp; }
    I.ea = na;
    I.eb = nb;
#pragma unroll
    for (int k = 0; k < 3; k++) { I.y0[k] = I.p[k]; I.y1[k] = xnode(I, na, k); I.y2[k] = xnode(I, nb, k); }
    const double L2 = (I.y2[0] - I.y1[0]) * (I.y2[0] - I.y1[0]) + (I.y2[1] - I.y1[1]) * (I.y2[1] - I.y1[1]) +
                      (I.y2[2] - I.y1[2]) * (I.y2[2] - I.y1[2]);
    const double q = (I.y1[0] - I.y0[0]) * (I.y2[0] - I.y1[0]) + (I.y1[1] - I.y0[1]) * (I.y2[1] - I.y1[1]) +
                     (I.y1[2] - I.y0[2]) * (I.y2[2] - I.y1[2]);
    const double tt = -q / L2, uu = 1 - tt;
    double v = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double rr = I.y0[k] - uu * I.y1[k] - tt * I.y2[k];
        v += rr * rr;
    }
    I.dsq = v;
    I.q = q;
    I.g = q * q;
    I.f2 = L2;
    I.f2sq = L2 * L2;
    I.f2cu = I.f2sq * L2;
    I.if2 = 1.0 / I.f2;
    I.if2sq = 1.0 / I.f2sq;
    I.if2cu = 1.0 / I.f2cu;
}

// point_plane_distance (point_triange_v2.h:449-582): the squared distance through the
// projection's barycentric coordinates, as the reference evaluates it.
__device__ inline void setup_pp(Impact &I)
{
    double E1[3], E2[3], W[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        E1[k] = I.v1[k] - I.v0[k];
        E2[k] = I.v2[k] - I.v0[k];
        W[k] = I.v0[k] - I.p[k];
    }
    const double a = dot3(E1, E1), b = dot3(E2, E1), c = dot3(E2, E2), d = dot3(W, E1), e = dot3(W, E2);
    const double det = a * c - b * b;
    const double ps = (b * e - c * d) * (1. / det), pt = (b * d - a * e) * (1. / det), pu = 1 - (ps + pt);
    double v = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double r = I.v1[k] * ps + I.v2[k] * pt + I.v0[k] * pu - I.p[k];
        v += r * r;
    }
    I.dsq = v;
}

// Region classification of pt (point_triange_v2.h:589-842); fills branch data, s, t, dsq.
__device__ inline void classify(Impact &I)
{
    double E1[3], E2[3], W[3];
#pragma unroll
    for (int k = 0; k < 3; k++) { E1[k] = I.v1[k] - I.v0[k]; E2[k] = I.v2[k] - I.v0[k]; W[k] = I.v0[k] - I.p[k]; }
    const double a = dot3(E1, E1), b = dot3(E2, E1), c = dot3(E2, E2), d = dot3(W, E1), e = dot3(W, E2);
    const double det = a * c - b * b;
    double s = b * e - c * d, t = b * d - a * e;
    int br;
    if (s + t <= det) {
        if (s < 0) {
            if (t < 0 && d < 0) {
                t = 0;
                if (-d >= a) { br = 1; s = 1; setup_vv(I, 0, 2); }
                else { br = 5; s = -d / a; setup_ve(I, 2, 1); }
            } else {
                s = 0;
                if (e >= 0) { br = 3; t = 0; setup_vv(I, 0, 1); }
                else if (-e >= c) { br = 2; t = 1; setup_vv(I, 0, 3); }
                else { br = 4; t = -e / c; setup_ve(I, 3, 1); }
            }
        } else if (t < 0) {
            t = 0;
            if (d >= 0) { br = 3; s = 0; setup_vv(I, 0, 1); }
            else if (-d >= a) { br = 1; s = 1; setup_vv(I, 0, 2); }
            else { br = 5; s = -d / a; setup_ve(I, 1, 2); }
        } else {
            br = 0;
            const double inv = 1. / det;
            s *= inv;
            t *= inv;
            setup_pp(I);
        }
    } else {
        if (s < 0) {
            const double t0 = b + d, t1 = c + e;
            if (t1 > t0) {
                const double num = t1 - t0, den = a - 2 * b + c;
                if (num >= den) { br = 1; s = 1; t = 0; setup_vv(I, 0, 2); }
                else { br = 6; s = num / den; t = 1 - s; setup_ve(I, 2, 3); }
            } else {
                s = 0;
                if (t1 <= 0) { br = 2; t = 1; setup_vv(I, 0, 3); }
                else if (e >= 0) { br = 3; t = 0; setup_vv(I, 0, 1); }
                else { br = 4; t = -e / c; setup_ve(I, 1, 3); }
            }
        } else if (t < 0) {
            const double t0 = b + e, t1 = a + d;
            if (t1 > t0) {
                const double num = t1 - t0, den = a - 2 * b + c;
                if (num >= den) { br = 2; t = 1; s = 0; setup_vv(I, 0, 3); }
                else { br = 7; t = num / den; s = 1 - t; setup_ve(I, 2, 3); }
            } else {
                t = 0;
                if (t1 <= 0) { br = 1; s = 1; setup_vv(I, 0, 2); }
                else if (d >= 0) { br = 3; s = 0; setup_vv(I, 0, 1); }
                else { br = 5; s = -d / a; setup_ve(I, 1, 2); }
            }
        } else {
            const double num = c + e - b - d;
            if (num <= 0) { br = 2; s = 0; t = 1; setup_vv(I, 0, 3); }
            else {
                const double den = a - 2 * b + c;
                if (num >= den) { br = 1; s = 1; t = 0; setup_vv(I, 0, 2); }
                else { br = 6; s = num / den; t = 1 - s; setup_ve(I, 2, 3); }
            }
        }
    }
    I.branch = br;
    I.s = s;
    I.t = t;
}

// ---- vertex-vertex: |x_i0 - x_i1|^2 ----
__device__ inline double vv_grad(const Impact &I, int r)
{
    const int nd = r / 3, k = r % 3;
    const double dd = xnode(I, I.i0, k) - xnode(I, I.i1, k);
    return nd == I.i0 ? 2 * dd : (nd == I.i1 ? -(2 * dd) : 0.0);
}
__device__ inline double vv_hess(const Impact &I, int r, int c)
{
    if (r % 3 != c % 3) return 0;
    const int nr = r / 3, nc = c / 3;
    const bool ir = nr == I.i0 || nr == I.i1, ic = nc == I.i0 || nc == I.i1;
    if (!ir || !ic) return 0;
    return nr == nc ? 2.0 : -2.0;
}

// ---- vertex-edge: derivatives of f1 - g/f2 on (y0, y1, y2) ----
__device__ inline int ve_slot(const Impact &I, int nd) { return nd == 0 ? 0 : (nd == I.ea ? 1 : (nd == I.eb ? 2 : -1)); }
struct VeD { double qd, f1d, f2d; };
__device__ inline VeD ve_d1(const Impact &I, int P, int k)
{
    const double a = sel3(I.y0, k), b = sel3(I.y1, k), c = sel3(I.y2, k);
    VeD D;
    D.qd = P == 0 ? b - c : (P == 1 ? a - 2 * b + c : b - a);
    const double e1 = b - a, e2 = c - b;
    D.f1d = P == 0 ? -2 * e1 : (P == 1 ? 2 * e1 : 0.0);
    D.f2d = P == 0 ? 0.0 : (P == 1 ? -2 * e2 : 2 * e2);
    return D;
}
__device__ inline double ve_grad(const Impact &I, int r)
{
    const int P = ve_slot(I, r / 3);
    if (P < 0) return 0;
    const VeD D = ve_d1(I, P, r % 3);
    const double gd = 2 * I.q * D.qd;
    return D.f1d + (I.g * D.f2d - I.f2 * gd) * I.if2sq;
}
__device__ inline double ve_hess(const Impact &I, int r, int c)
{
    const int Pr = ve_slot(I, r / 3), Pc = ve_slot(I, c / 3);
    if (Pr < 0 || Pc < 0) return 0;
    const VeD A = ve_d1(I, Pr, r % 3), B = ve_d1(I, Pc, c % 3);
    const bool same = (r % 3) == (c % 3);
    // node patterns of q = (y1-y0).(y2-y1), f1 = |y1-y0|^2, f2 = |y2-y1|^2
    const int pc = Pr * 3 + Pc;
    const double qdd = same ? (pc == 1 || pc == 3 || pc == 5 || pc == 7 ? 1.0 : (pc == 4 ? -2.0 : (pc == 2 || pc == 6 ? -1.0 : 0.0))) : 0.0;
    const double f1dd = same ? (pc == 0 || pc == 4 ? 2.0 : (pc == 1 || pc == 3 ? -2.0 : 0.0)) : 0.0;
    const double f2dd = same ? (pc == 4 || pc == 8 ? 2.0 : (pc == 5 || pc == 7 ? -2.0 : 0.0)) : 0.0;
    const double gdI = 2 * I.q * A.qd, gdJ = 2 * I.q * B.qd;
    const double gdd = 2 * (A.qd * B.qd + I.q * qdd);
    return -2 * I.g * A.f2d * B.f2d * I.if2cu + (gdI * B.f2d + gdJ * A.f2d) * I.if2sq + f1dd + I.g * f2dd * I.if2sq -
           gdd * I.if2;
}

// Entry (r, c), r <= c, of the tangent into an impact's block record; inside a diagonal
// block the mirrored entry is written too, so that every block is whole.
template <typename T>
__device__ inline void ct_store(T *__restrict__ tan, int r, int c, double v)
{
    const int a = r / 3, b = c / 3, k = r % 3, l = c % 3;
    T *blk = tan + ct_block(a, b) * 9;
    blk[k * 3 + l] = (T)v;
    if (a == b && k != l) blk[l * 3 + k] = (T)v;
}

// ---- interior, closed form ----
// In the interior region the squared distance is that to the plane of the triangle,
// f = h^2 / m with N = E1 x E2, W = p - v0, h = N.W, m = N.N (E1 = v1 - v0, E2 = v2 - v0);
// the reference's chain rule through s(x), t(x) (point_triange_v2.h:449-582) evaluates the
// same function.  Here
//   grad f = (2h/m) gh - (h/m)^2 gm,
//   Hess f = (2/m) u u^T + (2h/m) Hh - (h/m)^2 Hm,     u = gh - (h/m) gm,
// with gh = grad h = [N, -(sum), E2 x W, W x E1], gm = grad m = [0, -(sum), 2 E2 x N, 2 N x E1],
//   Hh_ij = A T(W) + B T(E2) + C T(E1),   Hm_ij = 2 c_i.c_j + 2 A T(N),
// where i = (node na, axis ka), j = (nb, kb), T(V) = eps_{ka kb c} V_c, c_i = dN/dx_i, and with
// per-node coefficients a1 = dE1/dx, a2 = dE2/dx, w = dW/dx (p: 0,0,1; v0: -1,-1,-1; v1: 1,0,0;
// v2: 0,1,0): A = a1(na) a2(nb) - a1(nb) a2(na), B = a1(nb) w(na) - a1(na) w(nb),
// C = a2(na) w(nb) - a2(nb) w(na).  Each entry costs about 15 operations instead of the
// chain rule's ~200.
struct PpClosed {
    double N[3], W[3], E1[3], E2[3];
    double u[12];
    double cu, chh, chm, gp;   // 2/m, 2h/m, (h/m)^2, 2h/m (gradient w.r.t. p = gp N)
};

__device__ inline void cross3(const double a[3], const double b[3], double c[3])
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ __attribute__((always_inline)) inline void pp_closed_setup(const Impact &I, PpClosed &P)
{
#pragma unroll
    for (int k = 0; k < 3; k++) {
        P.W[k] = I.p[k] - I.v0[k];
        P.E1[k] = I.v1[k] - I.v0[k];
        P.E2[k] = I.v2[k] - I.v0[k];
    }
    cross3(P.E1, P.E2, P.N);
    const double h = dot3(P.N, P.W), m = dot3(P.N, P.N), im = 1.0 / m, r = h * im;
    double g2[3], g3[3], m2[3], m3[3];
    cross3(P.E2, P.W, g2);
    cross3(P.W, P.E1, g3);
    cross3(P.E2, P.N, m2);
    cross3(P.N, P.E1, m3);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        m2[k] *= 2;
        m3[k] *= 2;
        const double g1 = -(P.N[k] + g2[k] + g3[k]), m1 = -(m2[k] + m3[k]);
        P.u[k] = P.N[k];
        P.u[3 + k] = g1 - r * m1;
        P.u[6 + k] = g2[k] - r * m2[k];
        P.u[9 + k] = g3[k] - r * m3[k];
    }
    P.cu = 2 * im;
    P.chh = 2 * h * im;
    P.chm = r * r;
    P.gp = 2 * h * im;
}

__device__ inline double sel12(const double v[12], int i)
{
    double r = 0;
#pragma unroll
    for (int k = 0; k < 12; k++) r += (i == k ? 1.0 : 0.0) * v[k];
    return r;
}

// V[k] for a runtime k as a masked sum: a select between elements of a private array is
// turned into a load through a selected address, which puts the array in scratch memory.
__device__ inline double pick3(const double V[3], int k)
{
    return (k == 0 ? 1.0 : 0.0) * V[0] + (k == 1 ? 1.0 : 0.0) * V[1] + (k == 2 ? 1.0 : 0.0) * V[2];
}

// eps_{a b c} V_c
__device__ inline double tperm(int a, int b, const double V[3])
{
    const int d = (b - a + 3) % 3;   // 1: (a, a+1), 2: (a, a+2) = -(a+2, a)
    const double vc = pick3(V, 3 - a - b);
    return a == b ? 0.0 : (d == 1 ? vc : -vc);
}

// Row i of (k/2) Hess f, entries c >= i, into the impact's block record.
template <typename T>
__device__ __attribute__((always_inline)) inline void pp_closed_row(const PpClosed &P, int i, double hk, T *__restrict__ tan)
{
    const int na = i / 3, ka = i % 3;
    const double a1i = na == 1 ? -1.0 : (na == 2 ? 1.0 : 0.0);
    const double a2i = na == 1 ? -1.0 : (na == 3 ? 1.0 : 0.0);
    const double wi = na == 0 ? 1.0 : (na == 1 ? -1.0 : 0.0);
    const double ui = sel12(P.u, i);
    // c_i = a1 (e_ka x E2) - a2 (e_ka x E1)
    double ek2[3], ek1[3];
    ek2[0] = 0; ek2[1] = 0; ek2[2] = 0;
    ek1[0] = 0; ek1[1] = 0; ek1[2] = 0;
    {
        const int k1 = (ka + 1) % 3, k2 = (ka + 2) % 3;
        // e_k x V = (component k1) -V_k2 ... : (e_k x V)_k1 = -V_k2, (e_k x V)_k2 = V_k1
        const double e2k1 = pick3(P.E2, k1), e2k2 = pick3(P.E2, k2);
        const double e1k1 = pick3(P.E1, k1), e1k2 = pick3(P.E1, k2);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            ek2[c] = c == k1 ? -e2k2 : (c == k2 ? e2k1 : 0.0);
            ek1[c] = c == k1 ? -e1k2 : (c == k2 ? e1k1 : 0.0);
        }
    }
    double ci[3];
#pragma unroll
    for (int c = 0; c < 3; c++) ci[c] = a1i * ek2[c] - a2i * ek1[c];
    // T(ka, kb; V) for kb = 0, 1, 2
    double tw[3], t2[3], t1[3], tn[3];
#pragma unroll
    for (int kb = 0; kb < 3; kb++) {
        tw[kb] = tperm(ka, kb, P.W);
        t2[kb] = tperm(ka, kb, P.E2);
        t1[kb] = tperm(ka, kb, P.E1);
        tn[kb] = tperm(ka, kb, P.N);
    }
#pragma unroll
    for (int j = 0; j < 12; j++) {
        if (j < i) continue;
        const int nb = j / 3, kb = j % 3;
        const double a1j = nb == 1 ? -1.0 : (nb == 2 ? 1.0 : 0.0);
        const double a2j = nb == 1 ? -1.0 : (nb == 3 ? 1.0 : 0.0);
        const double wj = nb == 0 ? 1.0 : (nb == 1 ? -1.0 : 0.0);
        const double A = a1i * a2j - a1j * a2i, B = a1j * wi - a1i * wj, C = a2i * wj - a2j * wi;
        // c_j = a1j (e_kb x E2) - a2j (e_kb x E1), constant-coefficient in E1, E2
        double cj[3];
        {
            const int k1 = (kb + 1) % 3, k2 = (kb + 2) % 3;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double x2 = c == k1 ? -P.E2[k2] : (c == k2 ? P.E2[k1] : 0.0);
                const double x1 = c == k1 ? -P.E1[k2] : (c == k2 ? P.E1[k1] : 0.0);
                cj[c] = a1j * x2 - a2j * x1;
            }
        }
        const double hh = A * tw[kb] + B * t2[kb] + C * t1[kb];
        const double hm = 2 * (ci[0] * cj[0] + ci[1] * cj[1] + ci[2] * cj[2]) + 2 * A * tn[kb];
        ct_store(tan, i, j, hk * (P.cu * ui * P.u[j] + P.chh * hh - P.chm * hm));
    }
}

// Phase 1 for impacts, one thread per impact: the tangent blocks (k/2) Hess d^2 and the 12
// forces f = [-fp, w0 fp, w1 fp, w2 fp] with fp = (k/2) d(d^2)/dp (collision_response.h:89-114).
// Rigid-rigid impacts and d^2 <= eps^2 stage zeros.  Every (row, column) index is a
// compile-time constant after unrolling, so the entry formulas' index selections fold away.
template <typename T>
__global__ __launch_bounds__(256) void k_contact_impact(const int *__restrict__ icr, const double *__restrict__ dn,
                                                        int n, int cs, int ns, double kpen, double eps,
                                                        T *__restrict__ st_tan, double *__restrict__ st_f)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    T *tan = st_tan + (size_t)u * CT_TAN_REC;
    double *frc = st_f + (size_t)u * CT_F_REC;
    bool active = false;
#pragma unroll
    for (int i = 0; i < 4; i++) active |= icr[u + cs * (CR_ROWS + i)] >= 0;
    Impact I;
    if (active) {
        const int n0 = icr[u + cs * CR_NODES], n1 = icr[u + cs * (CR_NODES + 1)];
        const int n2 = icr[u + cs * (CR_NODES + 2)], n3 = icr[u + cs * (CR_NODES + 3)];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            I.p[j] = dn[n0 + ns * (ND_XC + j)];
            I.v0[j] = dn[n1 + ns * (ND_XC + j)];
            I.v1[j] = dn[n2 + ns * (ND_XC + j)];
            I.v2[j] = dn[n3 + ns * (ND_XC + j)];
        }
        classify(I);
        active = I.dsq > eps * eps;
    }
    const double hk = kpen / 2;
    if (!active) {
#pragma unroll
        for (int k = 0; k < CT_TAN_REC; k++) tan[k] = (T)0;
#pragma unroll
        for (int r = 0; r < 12; r++) frc[r] = 0.;
        return;
    }
    double gpk[3];    // (k/2) d(d^2)/dp
    if (I.branch == 0) {
        PpClosed P;
        pp_closed_setup(I, P);
#pragma unroll
        for (int r = 0; r < 12; r++) pp_closed_row(P, r, hk, tan);
#pragma unroll
        for (int k = 0; k < 3; k++) gpk[k] = hk * P.gp * P.N[k];
    } else if (I.branch <= 3) {
#pragma unroll
        for (int r = 0; r < 12; r++)
#pragma unroll
            for (int c = r; c < 12; c++) ct_store(tan, r, c, hk * vv_hess(I, r, c));
#pragma unroll
        for (int k = 0; k < 3; k++) gpk[k] = hk * vv_grad(I, k);
    } else {
#pragma unroll
        for (int r = 0; r < 12; r++)
#pragma unroll
            for (int c = r; c < 12; c++) ct_store(tan, r, c, hk * ve_hess(I, r, c));
#pragma unroll
        for (int k = 0; k < 3; k++) gpk[k] = hk * ve_grad(I, k);
    }
    const double w[4] = {-1.0, 1 - (I.s + I.t), I.s, I.t};
#pragma unroll
    for (int r = 0; r < 12; r++) frc[r] = w[r / 3] * gpk[r % 3];
}

}  // namespace icf

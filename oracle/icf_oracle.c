/*
 * icf_oracle.c -- CPU FP64 restatement of icFlow's assembly-path device math.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Nothing in the product library links or calls it.
 *
 * Built with -O2 -ffp-contract=off (see oracle/Makefile): no fused multiply-add contraction,
 * so every expression rounds in the order written.  Kernels become sequential loops over the
 * unit index; the reference's CAS-loop atomicAdd2 (elem_kernels9/atomic_add.h:1-13) becomes
 * a plain `+=`, which fixes the accumulation order to: unit 0..n-1, reference loop order
 * within a unit.
 *
 * Parity of each function is pinned by tests/test_oracle_*.py (closed forms, invariants,
 * finite differences, brute force) -- the reference ships no golden vectors and its CUDA
 * sources cannot be compiled in this environment.
 */
#include "icf_oracle.h"
#include <math.h>
#include <string.h>

static orc_consts C;

void orc_set_constants(const orc_consts *c) { C = *c; }

/* Scatter addition.  Sequential calls add in unit order.  Inside orc_assemble_par the units
 * run on several threads and every addition into A, b and the node forces is a
 * compare-and-swap loop on the 64-bit pattern of the double, as the reference's device
 * scatter does (atomicAdd2, elem_kernels9/atomic_add.h:1-13). */
static int g_par = 0;

static void acc(double *p, double v)
{
    if (!g_par) {
        *p += v;
        return;
    }
    unsigned long long *q = (unsigned long long *)p;
    unsigned long long old = __atomic_load_n(q, __ATOMIC_RELAXED), nw;
    do {
        double d;
        memcpy(&d, &old, sizeof d);
        d += v;
        memcpy(&nw, &d, sizeof nw);
    } while (!__atomic_compare_exchange_n(q, &old, nw, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED));
}

/* node-array field offsets: elem_kernels9/element_kernels2.h:8-12, cz_kernels.h:4-10 */
enum { X0 = 0, UN = 3, VN = 6, AN = 9, FF = 12, XC = 15 };

/* ------------------------------------------------------------------------------------------
 * Element: corotational linear tetrahedron
 * ------------------------------------------------------------------------------------------ */

/* Orthonormal frame from three points: column 0 along p1-p0, column 1 along
 * (p1-p0)x(p2-p0), column 2 = column1 x column0.  Output R[row*3+col].
 * Restates fastRotationMatrix, elem_kernels9/matrix_math1.h:32-71. */
static void frame3(const double *p0, const double *p1, const double *p2, double R[9])
{
    double d[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    R[0] = d[0] / len; R[3] = d[1] / len; R[6] = d[2] / len;
    double w[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    double c0 = -d[2] * w[1] + d[1] * w[2];
    double c1 = d[2] * w[0] - d[0] * w[2];
    double c2 = -d[1] * w[0] + d[0] * w[1];
    len = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    R[1] = c0 / len; R[4] = c1 / len; R[7] = c2 / len;
    double e0 = R[4] * R[6] - R[3] * R[7];
    double e1 = -R[1] * R[6] + R[0] * R[7];
    double e2 = R[1] * R[3] - R[0] * R[4];
    len = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
    R[2] = e0 / len; R[5] = e1 / len; R[8] = e2 / len;
}

/* f = R K (R^T x_c - x0), Df = R K R^T, V = rest volume, with K = B^T E B V.
 * Restates F_and_Df_Corotational, elem_kernels9/element_kernels2.h:19-159. */
void orc_F_and_Df(const double x0[12], const double un[12], double f[12], double Df[144], double *Vout)
{
    double xc[12];
    for (int i = 0; i < 12; i++) xc[i] = x0[i] + un[i];

    const double X1 = x0[0], Y1 = x0[1], Z1 = x0[2];
    const double X2 = x0[3], Y2 = x0[4], Z2 = x0[5];
    const double X3 = x0[6], Y3 = x0[7], Z3 = x0[8];
    const double X4 = x0[9], Y4 = x0[10], Z4 = x0[11];
    /* pairwise differences (element_kernels2.h:41-46) */
    double x12 = X1 - X2, x13 = X1 - X3, x14 = X1 - X4, x23 = X2 - X3, x24 = X2 - X4, x34 = X3 - X4;
    double y12 = Y1 - Y2, y13 = Y1 - Y3, y14 = Y1 - Y4, y23 = Y2 - Y3, y24 = Y2 - Y4, y34 = Y3 - Y4;
    double z12 = Z1 - Z2, z13 = Z1 - Z3, z14 = Z1 - Z4, z23 = Z2 - Z3, z24 = Z2 - Z4, z34 = Z3 - Z4;
    double x21 = -x12, x31 = -x13, x32 = -x23, x42 = -x24, x43 = -x34;
    double y21 = -y12, y31 = -y13, y32 = -y23, y42 = -y24, y43 = -y34;
    double z21 = -z12, z31 = -z13, z32 = -z23, z42 = -z24, z43 = -z34;
    (void)y21; (void)z21;
    double J = x21 * (y23 * z34 - y34 * z23) + x32 * (y34 * z12 - y12 * z34) + x43 * (y12 * z23 - y23 * z12);
    double V = J / 6.;
    /* shape-function gradients times J (element_kernels2.h:50-53) */
    double g[4][3];
    g[0][0] = y42 * z32 - y32 * z42; g[0][1] = x32 * z42 - x42 * z32; g[0][2] = x42 * y32 - x32 * y42;
    g[1][0] = y31 * z43 - y34 * z13; g[1][1] = x43 * z31 - x13 * z34; g[1][2] = x31 * y43 - x34 * y13;
    g[2][0] = y24 * z14 - y14 * z24; g[2][1] = x14 * z24 - x24 * z14; g[2][2] = x24 * y14 - x14 * y24;
    g[3][0] = y13 * z21 - y12 * z31; g[3][1] = x21 * z13 - x31 * z12; g[3][2] = x13 * y21 - x12 * y31;
    for (int n = 0; n < 4; n++) for (int k = 0; k < 3; k++) g[n][k] /= J;

    /* strain-displacement matrix, Voigt rows xx,yy,zz,xy,yz,zx (element_kernels2.h:59-65) */
    double Bm[6][12];
    memset(Bm, 0, sizeof(Bm));
    for (int n = 0; n < 4; n++) {
        double a = g[n][0], b = g[n][1], c = g[n][2];
        Bm[0][3 * n + 0] = a;
        Bm[1][3 * n + 1] = b;
        Bm[2][3 * n + 2] = c;
        Bm[3][3 * n + 0] = b; Bm[3][3 * n + 1] = a;
        Bm[4][3 * n + 1] = c; Bm[4][3 * n + 2] = b;
        Bm[5][3 * n + 0] = c; Bm[5][3 * n + 2] = a;
    }
    double BtE[12][6];
    for (int r = 0; r < 12; r++)
        for (int c = 0; c < 6; c++) {
            double s = 0;
            for (int i = 0; i < 6; i++) s += Bm[i][r] * C.E[i * 6 + c];
            BtE[r][c] = s;
        }
    double K[12][12];
    for (int r = 0; r < 12; r++)
        for (int c = 0; c < 12; c++) {
            double s = 0;
            for (int i = 0; i < 6; i++) s += BtE[r][i] * Bm[i][c] * V;
            K[r][c] = s;
        }

    /* R = R1 R0^T: rest frame R0 from x0, current frame R1 from x_c (element_kernels2.h:78-104) */
    double R0[9], R1[9], R[9];
    frame3(&x0[0], &x0[3], &x0[6], R0);
    frame3(&xc[0], &xc[3], &xc[6], R1);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            R[i * 3 + j] = R1[i * 3 + 0] * R0[j * 3 + 0] + R1[i * 3 + 1] * R0[j * 3 + 1] + R1[i * 3 + 2] * R0[j * 3 + 2];

    /* RK and RKR^T block by block (element_kernels2.h:106-125) */
    double RK[12][12], RKRt[12][12];
    memset(RK, 0, sizeof(RK));
    memset(RKRt, 0, sizeof(RKRt));
    for (int bi = 0; bi < 4; bi++)
        for (int bj = 0; bj < 4; bj++) {
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++)
                    for (int m = 0; m < 3; m++)
                        RK[3 * bi + k][3 * bj + l] += R[k * 3 + m] * K[3 * bi + m][3 * bj + l];
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++)
                    for (int m = 0; m < 3; m++)
                        RKRt[3 * bi + k][3 * bj + l] += RK[3 * bi + k][3 * bj + m] * R[l * 3 + m];
        }

    /* xr = R^T x_c - x0 (element_kernels2.h:127-150) */
    double xr[12];
    for (int n = 0; n < 4; n++) {
        const double *p = &xc[3 * n];
        xr[3 * n + 0] = p[0] * R[0] + p[1] * R[3] + p[2] * R[6];
        xr[3 * n + 1] = p[0] * R[1] + p[1] * R[4] + p[2] * R[7];
        xr[3 * n + 2] = p[0] * R[2] + p[1] * R[5] + p[2] * R[8];
    }
    for (int i = 0; i < 12; i++) xr[i] -= x0[i];

    /* f += RK xr; Df = RKR^T (element_kernels2.h:154-158) */
    for (int i = 0; i < 12; i++)
        for (int j = 0; j < 12; j++) {
            f[i] += RK[i][j] * xr[j];
            Df[i * 12 + j] = RKRt[i][j];
        }
    *Vout = V;
}

/* Scatter of a 4-node unit into BCSR + rhs.  pcsr points at the unit's first slot entry;
 * entries are strided.  Restates AssembleBCSR12, element_kernels2.h:162-189 (the contact
 * copy collision_response.h:6-33 is identical). */
static void scatter4(const int *pcsr, int stride, double *A, double *b, const double LHS[144], const double rhs[12])
{
    for (int i = 0; i < 4; i++) {
        for (int j = 0; j < 4; j++) {
            int slot = pcsr[stride * (i * 4 + j)];
            if (slot >= 0)
                for (int k = 0; k < 3; k++)
                    for (int l = 0; l < 3; l++) acc(&A[slot * 9 + 3 * k + l], LHS[(i * 3 + k) * 12 + j * 3 + l]);
        }
        int row = pcsr[stride * (16 + i)];
        if (row >= 0)
            for (int k = 0; k < 3; k++) acc(&b[row * 3 + k], rhs[i * 3 + k]);
    }
}

/* Newmark effective tangent and residual per tet, scattered.
 * Restates kelElementElasticityForce, elem_kernels9/element_kernels2.h:195-263. */
static void elem_unit(int u, const int *ie, double *dn, double h, double *A, double *b, int el_stride,
                      int nd_stride)
{
    {
        double x0[12], un[12], vn[12], an[12];
        int nid[4];
        for (int i = 0; i < 4; i++) {
            int n = ie[u + el_stride * i];
            nid[i] = n;
            for (int j = 0; j < 3; j++) {
                x0[3 * i + j] = dn[n + nd_stride * (X0 + j)];
                un[3 * i + j] = dn[n + nd_stride * (UN + j)];
                vn[3 * i + j] = dn[n + nd_stride * (VN + j)];
                an[3 * i + j] = dn[n + nd_stride * (AN + j)];
            }
        }
        double Df[144], f[12] = {0}, V;
        memset(Df, 0, sizeof(Df));
        orc_F_and_Df(x0, un, f, Df, &V);

        double rhs[12] = {0}, LHS[144];
        double gz = C.gravity * C.rho * V / 4;
        rhs[2] += gz; rhs[5] += gz; rhs[8] += gz; rhs[11] += gz;
        double massCoeff = V * (1.0 / (h * h) + C.dampingMass * C.NewmarkGamma / h) / C.NewmarkBeta;
        double stiffCoeff = 1.0 + C.dampingStiffness * C.NewmarkGamma / (h * C.NewmarkBeta);
        for (int i = 0; i < 12; i++) {
            rhs[i] -= f[i];
            for (int j = 0; j < 12; j++) {
                double Mij = C.M[i * 12 + j];
                rhs[i] -= (Mij * V * C.dampingMass + Df[i * 12 + j] * C.dampingStiffness) * vn[j] + (Mij * V * an[j]);
                LHS[i * 12 + j] = Df[i * 12 + j] * stiffCoeff + Mij * massCoeff;
            }
        }
        /* node force (element_kernels2.h:256-260), then BCSR (:262) */
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 3; j++) acc(&dn[nid[i] + nd_stride * (FF + j)], rhs[i * 3 + j]);
        scatter4(&ie[u + el_stride * 4], el_stride, A, b, LHS, rhs);
    }
}

void orc_elem_force(const int *ie, double *dn, double h, double *A, double *b,
                    int nElems, int el_stride, int nd_stride)
{
    for (int u = 0; u < nElems; u++) elem_unit(u, ie, dn, h, A, b, el_stride, nd_stride);
}

/* ------------------------------------------------------------------------------------------
 * Cohesive zones: Park-Paulino-Roesler potential
 * ------------------------------------------------------------------------------------------ */

/* Normal traction T_n(Dn, Dt).  Restates Tn_, cz_kernels9/PPR_math.h:2-12. */
double orc_Tn(double Dn, double Dt)
{
    double rn_ = Dn / C.deln, rt_ = Dt / C.delt;
    double q = C.p_m / C.alpha + rn_;
    double pre = C.gam_n / C.deln;
    double mid = (C.p_m * pow(1 - rn_, C.alpha) * pow(q, C.p_m - 1)) - (C.alpha * pow(1 - rn_, C.alpha - 1) * pow(q, C.p_m));
    double post = C.gam_t * pow(1 - rt_, C.beta) * pow(C.p_n / C.beta + rt_, C.p_n) + C.pMtn;
    return pre * mid * post;
}

/* Tangential traction T_t(Dn, Dt).  Restates Tt_, PPR_math.h:14-24. */
double orc_Tt(double Dn, double Dt)
{
    double rn_ = Dn / C.deln, rt_ = Dt / C.delt;
    double e1 = 1 - rt_, e2 = C.p_n / C.beta + rt_;
    double pre = C.gam_t / C.delt;
    double mid = C.p_n * pow(e1, C.beta) * pow(e2, C.p_n - 1) - C.beta * pow(e1, C.beta - 1) * pow(e2, C.p_n);
    double post = C.gam_n * pow(1 - rn_, C.alpha) * pow(C.p_m / C.alpha + rn_, C.p_m) + C.pMnt;
    return pre * mid * post;
}

/* dT_n/dDn.  Restates Dnn_, PPR_math.h:26-35. */
double orc_Dnn(double opn, double opt)
{
    double coeff = C.gam_n / (C.deln * C.deln);
    double a = 1.0 - (opn / C.deln), q = (C.p_m / C.alpha) + (opn / C.deln);
    double e1 = (C.p_m * C.p_m - C.p_m) * pow(a, C.alpha) * pow(q, C.p_m - 2.0);
    double e2 = (C.alpha * C.alpha - C.alpha) * pow(a, C.alpha - 2.0) * pow(q, C.p_m);
    double e3 = 2.0 * C.alpha * C.p_m * pow(a, C.alpha - 1.0) * pow(q, C.p_m - 1.0);
    double e4 = C.gam_t * pow((1.0 - (opt / C.delt)), C.beta) * pow(((C.p_n / C.beta) + (opt / C.delt)), C.p_n) + C.pMtn;
    return coeff * (e1 + e2 - e3) * e4;
}

/* dT_t/dDt.  Restates Dtt_, PPR_math.h:37-46. */
double orc_Dtt(double opn, double opt)
{
    double coeff = C.gam_t / (C.delt * C.delt);
    double a = 1.0 - (opt / C.delt), q = (C.p_n / C.beta) + (opt / C.delt);
    double e1 = (C.p_n * C.p_n - C.p_n) * pow(a, C.beta) * pow(q, C.p_n - 2.0);
    double e2 = (C.beta * C.beta - C.beta) * pow(a, C.beta - 2.0) * pow(q, C.p_n);
    double e3 = 2.0 * C.beta * C.p_n * pow(a, C.beta - 1.0) * pow(q, C.p_n - 1.0);
    double e4 = C.gam_n * pow(1.0 - (opn / C.deln), C.alpha) * pow((C.p_m / C.alpha) + (opn / C.deln), C.p_m) + C.pMnt;
    return coeff * (e1 + e2 - e3) * e4;
}

/* dT_n/dDt = dT_t/dDn.  Restates Dnt_, PPR_math.h:48-57. */
double orc_Dnt(double opn, double opt)
{
    double coeff = C.gam_n * C.gam_t / (C.deln * C.delt);
    double an = 1.0 - (opn / C.deln), qn = (C.p_m / C.alpha) + (opn / C.deln);
    double at = 1.0 - (opt / C.delt), qt = (C.p_n / C.beta) + (opt / C.delt);
    double e1 = C.p_m * pow(an, C.alpha) * pow(qn, C.p_m - 1.0);
    double e2 = C.alpha * pow(an, C.alpha - 1.0) * pow(qn, C.p_m);
    double e3 = C.p_n * pow(at, C.beta) * pow(qt, C.p_n - 1.0);
    double e4 = C.beta * pow(at, C.beta - 1.0) * pow(qt, C.p_n);
    return coeff * (e1 - e2) * (e3 - e4);
}

/* Test instrumentation of the law's leaves: while a log is set (orc_set_branch_log),
 * orc_cz_force writes the leaf each Gauss point took to log[3 * unit + gp]:
 *   0 opening beyond deln or delt (early failure)
 *   contact (opn < 0): 1-3 with pmax != 0, 4-6 with pmax == 0; within each, tangential
 *     loading, unloading, failure
 *   7 both softening; 8, 9 tangential softening only (pmax != 0, == 0); 10, 11 normal
 *     softening only (tmax != 0, == 0); 12-15 neither (pmax != 0 / == 0 x tmax != 0 / == 0);
 *   16 failure away from contact. */
static int *g_branch_log = 0;
static _Thread_local int g_branch_at = -1;   /* per thread: orc_assemble_par runs units in parallel */
void orc_set_branch_log(int *log) { g_branch_log = log; }
static void log_branch(int code)
{
    if (g_branch_log && g_branch_at >= 0) g_branch_log[g_branch_at] = code;
}

/* PPR law with loading history.  flags[0] = contact, flags[1] = failed (set, never cleared).
 * out = {Tn, Tt, Dnn, Dtt, Dnt, Dtn}.  On the early-failure return Dtn is left at 0 here;
 * the reference leaves it unset, which is harmless because a failed CZ is not assembled.
 * Restates cohesive_law_thermodynamically_consistent, cz_kernels9/cz_kernels.h:37-174. */
void orc_cohesive_law(int flags[2], double *pmax, double *tmax, double opn, double opt, double out[6])
{
    double Tn = 0, Tt = 0, Dnn = 0, Dtt = 0, Dnt = 0;
    out[5] = 0;
    if (opn > C.deln || opt > C.delt) {           /* :40-42 */
        flags[1] = 1;
        out[0] = out[1] = out[2] = out[3] = out[4] = 0;
        log_branch(0);
        return;
    }
    int contact = (opn < 0);
    flags[0] = contact;                            /* :43 assigns (not ORs) per Gauss point */
    const double eps = -1e-9, frac = 0.05;         /* :44-45 */
    double thr_t = C.f_tt * frac, thr_n = C.f_tn * frac;

    if (contact) {                                 /* :49-87 */
        Dnt = 0;
        const int base = *pmax != 0 ? 1 : 4;
        if (*pmax != 0) {
            double peak = orc_Tn(*pmax, *tmax);
            Tn = peak * opn / *pmax;
            Dnn = peak / *pmax;
        } else {
            Dnn = orc_Dnn(0, *tmax);
            Tn = Dnn * opn;
        }
        Tt = orc_Tt(0, opt);
        if (Tt >= eps && !(opt > C.delt * C.rt / 5 && Tt < thr_t)) {
            if (opt >= *tmax) {
                *tmax = opt;
                Dtt = orc_Dtt(0, opt);
                log_branch(base);
            } else {
                double peak = orc_Tt(0, *tmax);
                Tt = peak * opt / *tmax;
                Dtt = peak / *tmax;
                log_branch(base + 1);
            }
        } else {
            Tt = Dtt = Dnt = 0;
            Tn = Dnn = 0;
            flags[1] = 1;
            log_branch(base + 2);
        }
    } else {                                       /* :88-172 */
        Tt = orc_Tt(opn, opt);
        Tn = orc_Tn(opn, opt);
        if (Tt >= eps && Tn >= eps && !(opt > C.delt * C.rt / 5 && Tt < thr_t) &&
            !(opn > C.deln * C.rn / 5 && Tn < thr_n)) {
            int tsoft = (opt >= *tmax), nsoft = (opn >= *pmax);
            if (tsoft && nsoft) {                  /* :100-108 */
                log_branch(7);
                *tmax = opt;
                *pmax = opn;
                Dnn = orc_Dnn(opn, opt);
                Dnt = orc_Dnt(opn, opt);
                Dtt = orc_Dtt(opn, opt);
            } else if (tsoft && !nsoft) {          /* :109-124 */
                log_branch(*pmax != 0 ? 8 : 9);
                Dnt = 0;
                if (*pmax != 0) {
                    double peak = orc_Tn(*pmax, *tmax);
                    Tn = peak * opn / *pmax;
                    Dnn = peak / *pmax;
                } else {
                    Tn = 0;
                    Dnn = orc_Dnn(0, *tmax);
                }
                *tmax = opt;
                Tt = orc_Tt(*pmax, opt);
                Dtt = orc_Dtt(*pmax, opt);
            } else if (!tsoft && nsoft) {          /* :125-141 */
                log_branch(*tmax != 0 ? 10 : 11);
                Dnt = 0;
                if (*tmax != 0) {
                    double peak = orc_Tt(*pmax, *tmax);
                    Tt = peak * opt / *tmax;
                    Dtt = peak / *tmax;
                } else {
                    Tt = 0;
                    Dtt = orc_Dtt(*pmax, 0);
                }
                *pmax = opn;
                Tn = orc_Tn(*pmax, *tmax);
                Dnn = orc_Dnn(*pmax, *tmax);
            } else {                               /* :142-164 */
                log_branch(12 + (*pmax != 0 ? 0 : 2) + (*tmax != 0 ? 0 : 1));
                Dnt = 0;
                double peakN = orc_Tn(*pmax, *tmax);
                if (*pmax != 0) {
                    Tn = peakN * opn / *pmax;
                    Dnn = peakN / *pmax;
                } else {
                    Tn = 0;
                    Dnn = orc_Dnn(0, *tmax);
                }
                if (*tmax != 0) {
                    double peakT = orc_Tt(*pmax, *tmax);
                    Tt = peakT * opt / *tmax;
                    Dtt = peakT / *tmax;
                } else {
                    Tt = 0;
                    Dtt = orc_Dtt(*pmax, 0);
                }
            }
        } else {                                   /* :167-171 */
            flags[1] = 1;
            Tn = Tt = Dnn = Dtt = Dnt = 0;
            log_branch(16);
        }
    }
    out[0] = Tn; out[1] = Tt; out[2] = Dnn; out[3] = Dtt; out[4] = Dnt; out[5] = Dnt; /* :173 */
}

/* Midplane frame of a CZ: rows e1 = (x1-x0)/|.|, e2 = n x e1 (normalised), e3 = n with
 * n = (x1-x0) x (x0-x2) / |.|; area = |n|/2.  R[row*3+col].
 * Restates CZRotationMatrix, cz_kernels9/matrix_math.h:15-68. */
void orc_cz_rotation(const double p[9], double R[9], double *area)
{
    double a0 = p[3] - p[0], a1 = p[4] - p[1], a2 = p[5] - p[2];
    double c0 = p[0] - p[6], c1 = p[1] - p[7], c2 = p[2] - p[8];
    double la = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
    R[0] = a0 / la; R[1] = a1 / la; R[2] = a2 / la;
    double n0 = -a2 * c1 + a1 * c2;
    double n1 = a2 * c0 - a0 * c2;
    double n2 = -a1 * c0 + a0 * c1;
    double ln = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    *area = ln / 2;
    n0 /= ln; n1 /= ln; n2 /= ln;
    R[6] = n0; R[7] = n1; R[8] = n2;
    a0 /= la; a1 /= la; a2 /= la;
    double r0 = -n2 * a1 + n1 * a2;
    double r1 = n2 * a0 - n0 * a2;
    double r2 = -n1 * a0 + n0 * a1;
    double lr = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
    R[3] = r0 / lr; R[4] = r1 / lr; R[5] = r2 / lr;
}

/* CZ SoA field offsets: cz_kernels9/cz_kernels.h:13-32 */
enum { CZ_PMAX = 0, CZ_TMAX = 3, CZ_TPMAX = 6, CZ_TTMAX = 9, CZ_DN = 12, CZ_DT = 13, CZ_TN = 14, CZ_TT = 15 };
enum { CZ_FAILED = 0, CZ_TCONTACT = 1, CZ_TDAMAGED = 2, CZ_TFAILED = 3, CZ_VRTS = 4, CZ_PCSR = 10 };

/* 6-node zero-thickness CZ, 3 Gauss points.  Restates kczCZForce, cz_kernels.h:210-434
 * (with AssembleBCSR18, :177-203). */
static void cz_unit(int u, double *dcz, int *icz, double *dn, double *A, double *b, int cz_stride, int nd_stride)
{
    {
        if (icz[u + cz_stride * CZ_FAILED] != 0) return;      /* :219 */
        double xc[18], xr[18], pmax[3], tmax[3];
        for (int i = 0; i < 6; i++) {
            int v = icz[u + cz_stride * (CZ_VRTS + i)];
            for (int j = 0; j < 3; j++)
                xc[3 * i + j] = dn[v + nd_stride * (X0 + j)] + dn[v + nd_stride * (UN + j)];
        }
        for (int i = 0; i < 3; i++) {
            pmax[i] = dcz[u + cz_stride * (CZ_PMAX + i)];
            tmax[i] = dcz[u + cz_stride * (CZ_TMAX + i)];
        }
        double mid[9];
        for (int i = 0; i < 9; i++) mid[i] = (xc[i] + xc[i + 9]) * 0.5;
        double R[9], area;
        orc_cz_rotation(mid, R, &area);
        for (int i = 0; i < 6; i++)
            for (int r = 0; r < 3; r++)
                xr[3 * i + r] = xc[3 * i + 0] * R[r * 3 + 0] + xc[3 * i + 1] * R[r * 3 + 1] + xc[3 * i + 2] * R[r * 3 + 2];

        double Keff[18 * 18], rhs[18];
        memset(Keff, 0, sizeof(Keff));
        memset(rhs, 0, sizeof(rhs));
        int gp_contact[3] = {0, 0, 0}, gp_failed[3] = {0, 0, 0};
        double avgDn = 0, avgDt = 0, avgTn = 0, avgTt = 0;
        const double W = 1.0 / 3.0;

        for (int gp = 0; gp < 3; gp++) {
            double dt1 = 0, dt2 = 0, dnn = 0;           /* :281-290 */
            for (int i = 0; i < 3; i++) {
                double s = C.sf[i * 3 + gp];
                dt1 += (xr[i * 3 + 0] - xr[i * 3 + 9]) * s;
                dt2 += (xr[i * 3 + 1] - xr[i * 3 + 10]) * s;
                dnn += (xr[i * 3 + 2] - xr[i * 3 + 11]) * s;
            }
            double opn = dnn, opt = sqrt(dt1 * dt1 + dt2 * dt2);
            int fl[2] = {gp_contact[gp], gp_failed[gp]};
            double L[6];
            g_branch_at = 3 * u + gp;
            orc_cohesive_law(fl, &pmax[gp], &tmax[gp], opn, opt, L);
            g_branch_at = -1;
            gp_contact[gp] = fl[0];
            gp_failed[gp] = fl[1];
            double Tn = L[0], Tt = L[1], Dnn = L[2], Dtt = L[3], Dnt = L[4], Dtn = L[5];
            avgDn += opn / 3; avgDt += opt / 3; avgTn += Tn / 3; avgTt += Tt / 3;

            double T[3] = {0, 0, 0}, Td[9] = {0};      /* :301-340 */
            if (opt < 1e-20) {
                T[2] = Tn;
                Td[0] = Dtt; Td[4] = Dtt; Td[8] = Dnn;
                Td[3] = Td[1] = 0;
                Td[6] = Dtn; Td[2] = Dnt; Td[7] = Dtn; Td[5] = Dnt;
            } else {
                T[0] = Tt * dt1 / opt;
                T[1] = Tt * dt2 / opt;
                T[2] = Tn;
                double o2 = opt * opt, o3 = o2 * opt;
                double d00 = dt1 * dt1, d10 = dt2 * dt1, d11 = dt2 * dt2;
                Td[0] = Dtt * d00 / o2 + Tt * d11 / o3;
                Td[4] = Dtt * d11 / o2 + Tt * d00 / o3;
                Td[8] = Dnn;
                Td[3] = Td[1] = Dtt * d10 / o2 - Tt * d10 / o3;
                Td[6] = Dtn * dt1 / opt;
                Td[2] = Dnt * dt1 / opt;
                Td[7] = Dtn * dt2 / opt;
                Td[5] = Dnt * dt2 / opt;
            }
            const double *Bg = &C.B[gp * 54];           /* B[gp][3][18] */
            double BtT[18];                               /* :345-351 */
            for (int i = 0; i < 18; i++) {
                double s = 0;
                for (int j = 0; j < 3; j++) s += Bg[j * 18 + i] * T[j];
                BtT[i] = s * -(W * area);
            }
            for (int i = 0; i < 6; i++)                   /* :354-364: rhs += R^T BtT */
                for (int r = 0; r < 3; r++)
                    rhs[i * 3 + r] += BtT[i * 3 + 0] * R[0 * 3 + r] + BtT[i * 3 + 1] * R[1 * 3 + r] + BtT[i * 3 + 2] * R[2 * 3 + r];

            double BtTd[18][3];                           /* :368-373 */
            for (int row = 0; row < 18; row++)
                for (int col = 0; col < 3; col++) {
                    double s = 0;
                    for (int k = 0; k < 3; k++) s += Bg[k * 18 + row] * Td[k * 3 + col];
                    BtTd[row][col] = s * (W * area);
                }
            double BtTdB[18][18];                         /* :376-380 */
            for (int row = 0; row < 18; row++)
                for (int col = 0; col < 18; col++) {
                    double s = 0;
                    for (int k = 0; k < 3; k++) s += BtTd[row][k] * Bg[k * 18 + col];
                    BtTdB[row][col] = s;
                }
            for (int bi = 0; bi < 6; bi++)                /* :385-401: Keff += R^T BtTdB R */
                for (int bj = 0; bj < 6; bj++) {
                    double tmp[3][3] = {{0}};
                    for (int k = 0; k < 3; k++)
                        for (int l = 0; l < 3; l++)
                            for (int m = 0; m < 3; m++) tmp[k][l] += R[m * 3 + k] * BtTdB[3 * bi + m][3 * bj + l];
                    for (int k = 0; k < 3; k++)
                        for (int l = 0; l < 3; l++)
                            for (int m = 0; m < 3; m++) Keff[(3 * bi + k) * 18 + 3 * bj + l] += tmp[k][m] * R[m * 3 + l];
                }
        }

        /* tentative history: max over Gauss points into all three slots (:404-416) */
        int damaged = 0;
        double pm = fmax(fmax(pmax[0], pmax[1]), pmax[2]);
        double tm = fmax(fmax(tmax[0], tmax[1]), tmax[2]);
        for (int i = 0; i < 3; i++) {
            dcz[u + cz_stride * (CZ_TPMAX + i)] = pm;
            dcz[u + cz_stride * (CZ_TTMAX + i)] = tm;
            if (pmax[i] >= C.deln * C.rn || tmax[i] >= C.delt * C.rt) damaged = 1;
        }
        int failed = gp_failed[0] || gp_failed[1] || gp_failed[2];
        int contact = gp_contact[0] || gp_contact[1] || gp_contact[2];
        icz[u + cz_stride * CZ_TCONTACT] = contact ? 1 : 0;
        icz[u + cz_stride * CZ_TFAILED] = failed ? 1 : 0;
        icz[u + cz_stride * CZ_TDAMAGED] = failed ? 0 : damaged;
        if (failed) return;                                /* :424 */

        const int *pcsr = &icz[u + cz_stride * CZ_PCSR];   /* AssembleBCSR18, :177-203 */
        for (int i = 0; i < 6; i++) {
            for (int j = 0; j < 6; j++) {
                int slot = pcsr[cz_stride * (i * 6 + j)];
                if (slot >= 0)
                    for (int k = 0; k < 3; k++)
                        for (int l = 0; l < 3; l++) acc(&A[slot * 9 + 3 * k + l], Keff[(i * 3 + k) * 18 + j * 3 + l]);
            }
            int row = pcsr[cz_stride * (36 + i)];
            if (row >= 0)
                for (int k = 0; k < 3; k++) acc(&b[row * 3 + k], rhs[i * 3 + k]);
        }
        dcz[u + cz_stride * CZ_DN] = avgDn;                /* :430-433 */
        dcz[u + cz_stride * CZ_DT] = avgDt;
        dcz[u + cz_stride * CZ_TN] = avgTn;
        dcz[u + cz_stride * CZ_TT] = avgTt;
    }
}

void orc_cz_force(double *dcz, int *icz, double *dn, double *A, double *b, double h,
                  int nCZs, int cz_stride, int nd_stride)
{
    (void)h;
    for (int u = 0; u < nCZs; u++) cz_unit(u, dcz, icz, dn, A, b, cz_stride, nd_stride);
}

/* ------------------------------------------------------------------------------------------
 * Contact: narrow phase, closest face, point-triangle penalty
 * ------------------------------------------------------------------------------------------ */

#define ORC_EPS 1e-10   /* collision_kernels9/narrowphase.h:1 */

/* Inverse of the edge matrix [p1-p0 | p2-p0 | p3-p0] by cofactors.
 * Restates Bvalues, narrowphase.h:3-33.  bmat[row*3+col]. */
void orc_bvalues(const double p[12], double bm[9])
{
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

/* Strict interior test on barycentric coordinates.  Restates ctest, narrowphase.h:35-44. */
int orc_ctest(const double bm[9], double x, double y, double z)
{
    double l1 = x * bm[0] + y * bm[1] + z * bm[2];
    double l2 = x * bm[3] + y * bm[4] + z * bm[5];
    double l3 = x * bm[6] + y * bm[7] + z * bm[8];
    return (l1 > ORC_EPS && l2 > ORC_EPS && l3 > ORC_EPS && (l1 + l2 + l3) < (1 - ORC_EPS));
}

/* Per candidate tet pair: 8-bit mask of vertices strictly inside the other tet, 0 if the
 * tets share a node.  Restates kNarrowPhase_new, narrowphase.h:46-155. */
void orc_narrow_phase(int nPairs, const double *dn, const int *ie, int *narrowList,
                      int el_all_stride, int nd_stride)
{
    for (int u = 0; u < nPairs; u++) {
        double p[24];
        int ids[8];
        for (int e = 0; e < 2; e++) {
            int tet = narrowList[u * 2 + e];
            for (int i = 0; i < 4; i++) {
                int n = ie[tet + el_all_stride * i];
                ids[i + 4 * e] = n;
                for (int j = 0; j < 3; j++) p[j + 3 * i + 12 * e] = dn[n + nd_stride * (XC + j)];
            }
        }
        narrowList[u * 2] = 0;
        int shared = 0;
        for (int i = 0; i < 4; i++)
            for (int j = 4; j < 8; j++)
                if (ids[i] == ids[j]) shared = 1;
        if (shared) continue;
        double b0[9], b1[9];
        orc_bvalues(&p[0], b0);
        orc_bvalues(&p[12], b1);
        int mask = 0;
        for (int i = 0; i < 4; i++)   /* vertices of tet 1 inside tet 0: bits 0..3 */
            if (orc_ctest(b0, p[12 + 3 * i] - p[0], p[13 + 3 * i] - p[1], p[14 + 3 * i] - p[2])) mask |= 1 << i;
        for (int i = 0; i < 4; i++)   /* vertices of tet 0 inside tet 1: bits 4..7 */
            if (orc_ctest(b1, p[3 * i] - p[12], p[3 * i + 1] - p[13], p[3 * i + 2] - p[14])) mask |= 16 << i;
        narrowList[u * 2] = mask;
    }
}

static double clamp01(double v) { return v <= 0 ? 0 : v >= 1 ? 1 : v; }

/* Squared distance from point p to triangle f (Eberly region classification).
 * Restates dtn + clamp, collision_kernels9/dist_node2tri.h:9-129. */
double orc_dtn(const double f[9], const double p[3])
{
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
    double q[3];
    for (int k = 0; k < 3; k++) q[k] = f[k] + s * e0[k] + t * e1[k] - p[k];
    return q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
}

/* Per (node, tet) tuple: the adjacent face nearest the node (first strict minimum wins);
 * rows 1..3 of nl2 receive that face's node ids.  Restates kFindClosestFace,
 * narrowphase.h:159-201. */
void orc_find_closest_face(int nPairs, int tet_stride, int *nl2, const double *dn, const int *ie,
                           const int *faces, int nd_stride, int el_all_stride, int fc_stride)
{
    for (int u = 0; u < nPairs; u++) {
        int node = nl2[u], tet = nl2[u + tet_stride];
        double p[3];
        for (int i = 0; i < 3; i++) p[i] = dn[node + nd_stride * (XC + i)];
        int nf = ie[tet + el_all_stride * 4];
        int best = -1;
        double bestd = 0;
        for (int k = 0; k < nf; k++) {
            int fid = ie[tet + el_all_stride * (5 + k)];
            double fc[9];
            for (int v = 0; v < 3; v++) {
                int fn = faces[fid + fc_stride * v];
                for (int j = 0; j < 3; j++) fc[j + 3 * v] = dn[fn + nd_stride * (XC + j)];
            }
            double d = orc_dtn(fc, p);
            if (k == 0 || d < bestd) { best = fid; bestd = d; }
        }
        for (int v = 0; v < 3; v++) nl2[u + tet_stride * (1 + v)] = best >= 0 ? faces[best + fc_stride * v] : -1;
    }
}

/* --- point-triangle squared distance with gradient and Hessian (point_triange_v2.h) --- */

/* |x[i]-x[j]|^2 over 12-vectors with derivatives written into the (i,j) entries only.
 * Restates vertex_vertex_distance_and_derivs_12, point_triange_v2.h:115-150. */
static double vv12(int i, int j, const double x[12], double fd[12], double sd[144])
{
    double d[3];
    for (int k = 0; k < 3; k++) d[k] = x[3 * i + k] - x[3 * j + k];
    for (int k = 0; k < 3; k++) {
        fd[3 * i + k] = 2 * d[k];
        fd[3 * j + k] = -fd[3 * i + k];
        sd[(3 * i + k) * 12 + 3 * i + k] = 2;
        sd[(3 * j + k) * 12 + 3 * j + k] = 2;
        sd[(3 * i + k) * 12 + 3 * j + k] = -2;
        sd[(3 * j + k) * 12 + 3 * i + k] = -2;
    }
    return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

/* Squared distance from y0 to the line through y1, y2 (9 coordinates) with exact first and
 * second derivatives of f1 - g/f2, where f1 = |y1-y0|^2, f2 = |y2-y1|^2,
 * g = ((y1-y0).(y2-y1))^2.  Restates vertex_edge_distance_and_derivs and its helpers
 * sp_dot3 / function_squared / vertex_vertex_distance_and_derivs, point_triange_v2.h:13-232. */
static double ve9(const double y[9], double fd[9], double sd[81])
{
    /* value (:167-174) */
    double L2 = (y[6] - y[3]) * (y[6] - y[3]) + (y[7] - y[4]) * (y[7] - y[4]) + (y[8] - y[5]) * (y[8] - y[5]);
    double tt = -((y[3] - y[0]) * (y[6] - y[3]) + (y[4] - y[1]) * (y[7] - y[4]) + (y[5] - y[2]) * (y[8] - y[5])) / L2;
    double uu = 1 - tt;
    double r0 = y[0] - uu * y[3] - tt * y[6], r1 = y[1] - uu * y[4] - tt * y[7], r2 = y[2] - uu * y[5] - tt * y[8];
    double value = r0 * r0 + r1 * r1 + r2 * r2;

    /* q = (y1-y0).(y2-y1) with derivatives; pattern kron([[0,1,-1],[1,-2,1],[-1,1,0]], I3) */
    static const double qpat[3][3] = {{0, 1, -1}, {1, -2, 1}, {-1, 1, 0}};
    double q = (y[3] - y[0]) * (y[6] - y[3]) + (y[4] - y[1]) * (y[7] - y[4]) + (y[5] - y[2]) * (y[8] - y[5]);
    double qd[9];
    for (int k = 0; k < 3; k++) {
        qd[k] = y[3 + k] - y[6 + k];
        qd[3 + k] = y[k] - 2 * y[3 + k] + y[6 + k];
        qd[6 + k] = y[3 + k] - y[k];
    }
    /* g = q^2 */
    double g = q * q, gd[9], gdd[81];
    for (int i = 0; i < 9; i++) {
        gd[i] = 2 * q * qd[i];
        for (int j = 0; j < 9; j++) {
            double qdd = (i % 3 == j % 3) ? qpat[i / 3][j / 3] : 0;
            gdd[i * 9 + j] = 2 * (qd[i] * qd[j] + q * qdd);
        }
    }
    /* f1 = |y1-y0|^2 (points 1,0), f2 = |y2-y1|^2 (points 2,1) */
    double f1d[9] = {0}, f2d[9] = {0}, f1dd[81] = {0}, f2dd[81] = {0};
    for (int k = 0; k < 3; k++) {
        double a = y[3 + k] - y[k];
        f1d[3 + k] = 2 * a; f1d[k] = -2 * a;
        double c = y[6 + k] - y[3 + k];
        f2d[6 + k] = 2 * c; f2d[3 + k] = -2 * c;
        f1dd[(3 + k) * 9 + 3 + k] = 2; f1dd[k * 9 + k] = 2; f1dd[(3 + k) * 9 + k] = -2; f1dd[k * 9 + 3 + k] = -2;
        f2dd[(6 + k) * 9 + 6 + k] = 2; f2dd[(3 + k) * 9 + 3 + k] = 2; f2dd[(6 + k) * 9 + 3 + k] = -2; f2dd[(3 + k) * 9 + 6 + k] = -2;
    }
    double f2 = L2;   /* |y2-y1|^2, the value of vertex_vertex_distance_and_derivs(2,1) */
    double f2sq = f2 * f2, f2cu = f2sq * f2;
    for (int i = 0; i < 9; i++) {                /* combination, :216-229 */
        fd[i] = f1d[i] + (g * f2d[i] - f2 * gd[i]) / f2sq;
        for (int j = 0; j < 9; j++) {
            double t1 = -2 * g * f2d[i] * f2d[j] / f2cu;
            double t2 = (gd[i] * f2d[j] + gd[j] * f2d[i]) / f2sq;
            double t3 = f1dd[i * 9 + j];
            double t4 = g * f2dd[i * 9 + j] / f2sq;
            double t5 = -gdd[i * 9 + j] / f2;
            sd[i * 9 + j] = t1 + t2 + t3 + t4 + t5;
        }
    }
    return value;
}

/* Point (x[0..2]) to the edge between 12-vector points i and j, nearer endpoint first.
 * Restates vertex_edge_distance_and_derivs_12, point_triange_v2.h:237-294. */
static double ve12(const double x[12], int i, int j, double fd[12], double sd[144])
{
    int a = 3 * i, b = 3 * j;
    double pa = 0, pb = 0;
    for (int k = 0; k < 3; k++) {
        pa += (x[k] - x[a + k]) * (x[k] - x[a + k]);
        pb += (x[k] - x[b + k]) * (x[k] - x[b + k]);
    }
    if (pa > pb) { int tmp = a; a = b; b = tmp; }
    double y[9] = {x[0], x[1], x[2], x[a], x[a + 1], x[a + 2], x[b], x[b + 1], x[b + 2]};
    double yd[9] = {0}, ydd[81] = {0};
    double val = ve9(y, yd, ydd);
    const int map[3] = {0, a, b};
    for (int P = 0; P < 3; P++)
        for (int k = 0; k < 3; k++) {
            fd[map[P] + k] = yd[3 * P + k];
            for (int Q = 0; Q < 3; Q++)
                for (int l = 0; l < 3; l++) sd[(map[P] + k) * 12 + map[Q] + l] = ydd[(3 * P + k) * 9 + 3 * Q + l];
        }
    return val;
}

/* Interior (point-plane) branch: d^2 = |p - (u v0 + s v1 + t v2)|^2 with s, t the
 * unconstrained minimisers, differentiated through s(x), t(x) by the chain rule over the
 * quadratic-form coefficients a..e.  Restates point_plane_distance, point_triange_v2.h:449-582
 * (second-derivative tables a2..e2, :298-371, written here as 4x4 node patterns times I3). */
static double pp12(const double x[12], double fd[12], double sd[144])
{
    const double *p = &x[0], *v0 = &x[3], *v1 = &x[6], *v2 = &x[9];
    double E1[3], E2[3], W[3];
    for (int k = 0; k < 3; k++) { E1[k] = v1[k] - v0[k]; E2[k] = v2[k] - v0[k]; W[k] = v0[k] - p[k]; }
    double a = E1[0] * E1[0] + E1[1] * E1[1] + E1[2] * E1[2];
    double b = E2[0] * E1[0] + E2[1] * E1[1] + E2[2] * E1[2];
    double c = E2[0] * E2[0] + E2[1] * E2[1] + E2[2] * E2[2];
    double d = W[0] * E1[0] + W[1] * E1[1] + W[2] * E1[2];
    double e = W[0] * E2[0] + W[1] * E2[1] + W[2] * E2[2];
    double det = a * c - b * b, det2 = det * det, det3 = det2 * det;
    double s = (b * e - c * d) * (1. / det);
    double t = (b * d - a * e) * (1. / det);
    double u = 1 - (s + t);

    /* first derivatives of a..e; node order p, v0, v1, v2 */
    double a1[12], b1[12], c1[12], d1[12], e1[12];
    for (int k = 0; k < 3; k++) {
        a1[k] = 0;            a1[3 + k] = -2 * E1[k];          a1[6 + k] = 2 * E1[k];   a1[9 + k] = 0;
        b1[k] = 0;            b1[3 + k] = -E1[k] - E2[k];      b1[6 + k] = E2[k];       b1[9 + k] = E1[k];
        c1[k] = 0;            c1[3 + k] = -2 * E2[k];          c1[6 + k] = 0;           c1[9 + k] = 2 * E2[k];
        d1[k] = -E1[k];       d1[3 + k] = E1[k] - W[k];        d1[6 + k] = W[k];        d1[9 + k] = 0;
        e1[k] = -E2[k];       e1[3 + k] = E2[k] - W[k];        e1[6 + k] = 0;           e1[9 + k] = W[k];
    }
    /* constant second-derivative patterns (node x node), each times I3 */
    static const double A2[4][4] = {{0, 0, 0, 0}, {0, 2, -2, 0}, {0, -2, 2, 0}, {0, 0, 0, 0}};
    static const double B2[4][4] = {{0, 0, 0, 0}, {0, 2, -1, -1}, {0, -1, 0, 1}, {0, -1, 1, 0}};
    static const double C2[4][4] = {{0, 0, 0, 0}, {0, 2, 0, -2}, {0, 0, 0, 0}, {0, -2, 0, 2}};
    static const double D2[4][4] = {{0, 1, -1, 0}, {1, -2, 1, 0}, {-1, 1, 0, 0}, {0, 0, 0, 0}};
    static const double E2p[4][4] = {{0, 1, 0, -1}, {1, -2, 0, 1}, {0, 0, 0, 0}, {-1, 1, 0, 0}};

    double s1[12], t1[12], u1[12], dt1[12];
    double r[3];
    for (int k = 0; k < 3; k++) r[k] = v1[k] * s + v2[k] * t + v0[k] * u - p[k];
    double val = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    for (int i = 0; i < 12; i++) {
        dt1[i] = c * a1[i] + a * c1[i] - 2 * b * b1[i];
        s1[i] = ((c * d - b * e) * dt1[i]) / det2 + ((e * b1[i] + b * e1[i]) - (d * c1[i] + c * d1[i])) / det;
        t1[i] = ((a * e - b * d) * dt1[i]) / det2 + ((d * b1[i] + b * d1[i]) - (a * e1[i] + e * a1[i])) / det;
        u1[i] = -(s1[i] + t1[i]);
    }
    /* dr_k/dx_i */
    double dr[3][12];
    for (int k = 0; k < 3; k++)
        for (int i = 0; i < 12; i++) {
            double v = v1[k] * s1[i] + v2[k] * t1[i] + v0[k] * u1[i];
            int node = i / 3, comp = i % 3;
            if (comp == k) v += (node == 0 ? -1.0 : node == 1 ? u : node == 2 ? s : t);
            dr[k][i] = v;
        }
    for (int i = 0; i < 12; i++) fd[i] = 2 * (r[0] * dr[0][i] + r[1] * dr[1][i] + r[2] * dr[2][i]);

    for (int i = 0; i < 12; i++)
        for (int j = 0; j < 12; j++) {
            int ni = i / 3, nj = j / 3, same = (i % 3) == (j % 3);
            double a2 = same ? A2[ni][nj] : 0, b2 = same ? B2[ni][nj] : 0, c2 = same ? C2[ni][nj] : 0;
            double d2 = same ? D2[ni][nj] : 0, e2 = same ? E2p[ni][nj] : 0;
            double ddet = -2 * b1[i] * b1[j] + a1[j] * c1[i] + a1[i] * c1[j] + c * a2 - 2 * b * b2 + a * c2;
            double Ns = b * e - c * d, Nt = b * d - a * e;
            double s2 = (-(c1[j] * d1[i]) - c1[i] * d1[j] + b1[j] * e1[i] + b1[i] * e1[j] + e * b2 - d * c2 - c * d2 + b * e2) / det
                - ((dt1[j] * (e * b1[i] - d * c1[i] - c * d1[i] + b * e1[i])) + (dt1[i] * (e * b1[j] - d * c1[j] - c * d1[j] + b * e1[j])) + (Ns * ddet)) / det2
                + (2 * Ns * dt1[i] * dt1[j]) / det3;
            double t2 = (b1[j] * d1[i] + b1[i] * d1[j] - a1[j] * e1[i] - a1[i] * e1[j] - e * a2 + d * b2 + b * d2 - a * e2) / det
                - ((dt1[j] * (-(e * a1[i]) + d * b1[i] + b * d1[i] - a * e1[i])) + (dt1[i] * (-(e * a1[j]) + d * b1[j] + b * d1[j] - a * e1[j])) + (Nt * ddet)) / det2
                + (2 * Nt * dt1[i] * dt1[j]) / det3;
            double u2 = -(s2 + t2);
            double acc = 0;
            for (int k = 0; k < 3; k++) {
                /* d2r_k/dx_i dx_j */
                double v = v1[k] * s2 + v2[k] * t2 + v0[k] * u2;
                if (i % 3 == k) v += (ni == 1 ? u1[j] : ni == 2 ? s1[j] : ni == 3 ? t1[j] : 0);
                if (j % 3 == k) v += (nj == 1 ? u1[i] : nj == 2 ? s1[i] : nj == 3 ? t1[i] : 0);
                acc += dr[k][i] * dr[k][j] + r[k] * v;
            }
            sd[i * 12 + j] = 2 * acc;
        }
    return val;
}

/* Point-triangle squared distance with 12-gradient and 12x12 Hessian over the seven region
 * branches; returns d^2 and the barycentric (s, t) of the closest point.  fd/sd must be
 * zeroed by the caller.  Restates pt, point_triange_v2.h:589-842. */
double orc_pt(const double x[12], double fd[12], double sd[144], double *so, double *to, int *branch)
{
    const double *P = &x[0], *A = &x[3], *Bv = &x[6], *Cv = &x[9];
    double E1[3], E2[3], W[3];
    for (int k = 0; k < 3; k++) { E1[k] = Bv[k] - A[k]; E2[k] = Cv[k] - A[k]; W[k] = A[k] - P[k]; }
    double a = E1[0] * E1[0] + E1[1] * E1[1] + E1[2] * E1[2];
    double b = E2[0] * E1[0] + E2[1] * E1[1] + E2[2] * E1[2];
    double c = E2[0] * E2[0] + E2[1] * E2[1] + E2[2] * E2[2];
    double d = W[0] * E1[0] + W[1] * E1[1] + W[2] * E1[2];
    double e = W[0] * E2[0] + W[1] * E2[1] + W[2] * E2[2];
    double det = a * c - b * b;
    double s = b * e - c * d, t = b * d - a * e;
    double dsq;
    int br = -1;
    if (s + t <= det) {
        if (s < 0) {
            if (t < 0 && d < 0) {                    /* region 4, first half */
                t = 0;
                if (-d >= a) { br = 1; s = 1; dsq = vv12(0, 2, x, fd, sd); }
                else { br = 5; s = -d / a; dsq = ve12(x, 2, 1, fd, sd); }
            } else {                                 /* region 4 (d >= 0) and region 3 */
                s = 0;
                if (e >= 0) { br = 3; t = 0; dsq = vv12(0, 1, x, fd, sd); }
                else if (-e >= c) { br = 2; t = 1; dsq = vv12(0, 3, x, fd, sd); }
                else { br = 4; t = -e / c; dsq = ve12(x, 3, 1, fd, sd); }
            }
        } else if (t < 0) {                          /* region 5 */
            t = 0;
            if (d >= 0) { br = 3; s = 0; dsq = vv12(0, 1, x, fd, sd); }
            else if (-d >= a) { br = 1; s = 1; dsq = vv12(0, 2, x, fd, sd); }
            else { br = 5; s = -d / a; dsq = ve12(x, 1, 2, fd, sd); }
        } else {                                     /* region 0 */
            br = 0;
            double inv = (1) / det;
            s *= inv; t *= inv;
            dsq = pp12(x, fd, sd);
        }
    } else {
        if (s < 0) {                                 /* region 2 */
            double t0 = b + d, t1 = c + e;
            if (t1 > t0) {
                double num = t1 - t0, den = a - 2 * b + c;
                if (num >= den) { br = 1; s = 1; t = 0; dsq = vv12(0, 2, x, fd, sd); }
                else { br = 6; s = num / den; t = 1 - s; dsq = ve12(x, 2, 3, fd, sd); }
            } else {
                s = 0;
                if (t1 <= 0) { br = 2; t = 1; dsq = vv12(0, 3, x, fd, sd); }
                else if (e >= 0) { br = 3; t = 0; dsq = vv12(0, 1, x, fd, sd); }
                else { br = 4; t = -e / c; dsq = ve12(x, 1, 3, fd, sd); }
            }
        } else if (t < 0) {                          /* region 6 */
            double t0 = b + e, t1 = a + d;
            if (t1 > t0) {
                double num = t1 - t0, den = a - 2 * b + c;
                if (num >= den) { br = 2; t = 1; s = 0; dsq = vv12(0, 3, x, fd, sd); }
                else { br = 7; t = num / den; s = 1 - t; dsq = ve12(x, 2, 3, fd, sd); }
            } else {
                t = 0;
                if (t1 <= 0) { br = 1; s = 1; dsq = vv12(0, 2, x, fd, sd); }
                else if (d >= 0) { br = 3; s = 0; dsq = vv12(0, 1, x, fd, sd); }
                else { br = 5; s = -d / a; dsq = ve12(x, 1, 2, fd, sd); }
            }
        } else {                                     /* region 1 */
            double num = c + e - b - d;
            if (num <= 0) { br = 2; s = 0; t = 1; dsq = vv12(0, 3, x, fd, sd); }
            else {
                double den = a - 2 * b + c;
                if (num >= den) { br = 1; s = 1; t = 0; dsq = vv12(0, 2, x, fd, sd); }
                else { br = 6; s = num / den; t = 1 - s; dsq = ve12(x, 2, 3, fd, sd); }
            }
        }
    }
    *so = s; *to = t; *branch = br;
    return dsq;
}

/* Penalty response of one impact (point + triangle).  Restates kCollisionResponseForce,
 * collision_kernels9/collision_response.h:38-126. */
static void contact_unit(int u, const int *icr, double *dn, double *A, double *b, int cr_stride, int nd_stride,
                         double k, double distEps)
{
    {
        const int *pcsr = &icr[u + cr_stride * 4];
        int anyActive = 0;
        for (int i = 0; i < 4; i++) if (pcsr[cr_stride * (16 + i)] >= 0) anyActive = 1;
        if (!anyActive) return;                            /* :48-51 */
        double x[12];
        int nid[4];
        for (int i = 0; i < 4; i++) {
            int n = icr[u + cr_stride * i];
            nid[i] = n;
            for (int j = 0; j < 3; j++) x[3 * i + j] = dn[n + nd_stride * (XC + j)];
        }
        double fd[12] = {0}, sd[144] = {0}, s, t;
        int br;
        double dsq = orc_pt(x, fd, sd, &s, &t, &br);
        double w[3] = {1 - (s + t), s, t};
        if (!(dsq > distEps * distEps)) return;
        double fx = k * 0.5 * fd[0], fy = k * 0.5 * fd[1], fz = k * 0.5 * fd[2];
        double fi[12] = {-fx, -fy, -fz,
                         w[0] * fx, w[0] * fy, w[0] * fz,
                         w[1] * fx, w[1] * fy, w[1] * fz,
                         w[2] * fx, w[2] * fy, w[2] * fz};
        double H[144];
        for (int i = 0; i < 12; i++)
            for (int j = i; j < 12; j++) H[i * 12 + j] = H[j * 12 + i] = k * sd[i * 12 + j] / 2;
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 3; j++) acc(&dn[nid[i] + nd_stride * (FF + j)], fi[i * 3 + j]);
        scatter4(pcsr, cr_stride, A, b, H, fi);
    }
}

void orc_collision_response(const int *icr, double *dn, double h, double *A, double *b,
                            int nImpacts, int cr_stride, int nd_stride, double k, double distEps)
{
    (void)h;
    for (int u = 0; u < nImpacts; u++) contact_unit(u, icr, dn, A, b, cr_stride, nd_stride, k, distEps);
}

/* One Newton iteration's assembly on `threads` host threads (OpenMP): A[0..nA) and b[0..nb)
 * zeroed, then the element, CZ and contact passes with their units spread over the threads
 * and every scatter addition a compare-and-swap (see acc).  The order of the additions is
 * arbitrary, as on the reference's device; the CPU baseline times this. */
void orc_assemble_par(const int *ie, int nElems, int el_stride, double *dcz, int *icz, int nCZs, int cz_stride,
                      const int *icr, int nImpacts, int cr_stride, double *dn, int nd_stride, double h, double k,
                      double distEps, double *A, long long nA, double *b, long long nb, int threads)
{
    g_par = 1;
#pragma omp parallel num_threads(threads)
    {
#pragma omp for schedule(static)
        for (long long i = 0; i < nA; i++) A[i] = 0;
#pragma omp for schedule(static)
        for (long long i = 0; i < nb; i++) b[i] = 0;
#pragma omp for schedule(dynamic, 256)
        for (int u = 0; u < nElems; u++) elem_unit(u, ie, dn, h, A, b, el_stride, nd_stride);
#pragma omp for schedule(dynamic, 64) nowait
        for (int u = 0; u < nCZs; u++) cz_unit(u, dcz, icz, dn, A, b, cz_stride, nd_stride);
#pragma omp for schedule(dynamic, 64)
        for (int u = 0; u < nImpacts; u++) contact_unit(u, icr, dn, A, b, cr_stride, nd_stride, k, distEps);
    }
    g_par = 0;
}

/* ---------------------------------------------------------------------------------------
 * Node state between iterations.
 * ------------------------------------------------------------------------------------- */

/* Node.cs:79-86 (InferTentativeUVA), per component. */
static void infer_uva(double du, double u, double v, double a, double h, double beta, double gamma,
                      double *un, double *vn, double *an)
{
    *un = u + du;
    *an = a * (1.0 - 1.0 / (2 * beta)) + du / (h * h * beta) + v * (-1.0 / (h * beta));
    *vn = v + h * ((1.0 - gamma) * a + gamma * *an);
}

/* Node.cs:88-97 with the packing of GPU_Functionality.cs:313-337. */
void orc_nodes_infer(const double *st, int ss, double *dn, int ns, int n, double h, double beta, double gamma)
{
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            double un, vn, an;
            infer_uva(st[i + ss * (9 + k)], st[i + ss * k], st[i + ss * (3 + k)], st[i + ss * (6 + k)], h, beta,
                      gamma, &un, &vn, &an);
            dn[i + ns * (3 + k)] = un;
            dn[i + ns * (6 + k)] = vn;
            dn[i + ns * (9 + k)] = an;
            dn[i + ns * (12 + k)] = 0;
            dn[i + ns * (15 + k)] = dn[i + ns * k] + un;
        }
}

/* ImplicitModel3.cs:247-256 */
void orc_nodes_add_dx(double *st, int ss, const int *alt, int n, const double *dx)
{
    for (int i = 0; i < n; i++)
        if (alt[i] >= 0)
            for (int k = 0; k < 3; k++) st[i + ss * (9 + k)] += dx[alt[i] * 3 + k];
}

/* Node.cs:62-77 */
void orc_nodes_accept(double *st, int ss, const double *dn, int ns, int n, double h)
{
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            const double un = dn[i + ns * (3 + k)];
            const double vn = (un - st[i + ss * k]) / h;
            st[i + ss * (6 + k)] = (vn - st[i + ss * (3 + k)]) / h;
            st[i + ss * (3 + k)] = vn;
            st[i + ss * k] = un;
            st[i + ss * (9 + k)] = 0;
        }
}

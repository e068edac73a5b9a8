// icf_cz.inc.h -- 6-node zero-thickness cohesive zone with the Park-Paulino-Roesler (PPR)
// potential (included by icf_lib.hip).
//
// Restated from cz_kernels9/cz_kernels.h:37-434 (cohesive law, kczCZForce), PPR_math.h:2-57
// (Tn_, Tt_, Dnn_, Dtt_, Dnt_) and matrix_math.h:15-68 (CZRotationMatrix).
// Differences of formulation (same mathematics):
//  - the five PPR functions share their powers: at an opening (Dn, Dt) the law needs
//    A^{a, a-1, a-2} and Bq^{p, p-1, p-2} for A = 1 - D/delta, Bq = p/a + D/delta, in the
//    normal and tangential directions; these come from one log and one exp per base instead of
//    one pow per factor;
//  - B = [s (x) I3, -s (x) I3] (ModelPrms.cs:324-349), so the 18x18 tangent is
//    K_ab = sgn_a sgn_b sum_gp s_a,gp s_b,gp Q_gp with Q_gp = (w a_J) R^T T_d R; phase 1
//    stages the three symmetric Q_gp (18 values), phase 2 forms each block from them.
#pragma once

namespace icf {

// x^e, x^(e-1), x^(e-2)
__device__ inline void pow3(double x, double e, double &p0, double &p1, double &p2)
{
    if (x > 0.0) {
        double lx = log(x);
        p0 = exp(e * lx);
        double r = 1.0 / x;
        p1 = p0 * r;
        p2 = p1 * r;
    } else {
        p0 = pow(x, e);
        p1 = pow(x, e - 1.0);
        p2 = pow(x, e - 2.0);
    }
}

// Powers of one direction of the PPR potential at one opening value.
struct PprPow {
    double a0, a1, a2;  // (1 - D/delta)^{a, a-1, a-2}
    double b0, b1, b2;  // (p/a + D/delta)^{p, p-1, p-2}
};

__device__ __attribute__((always_inline)) inline PprPow ppr_pow_n(double Dn)
{
    PprPow r;
    if (Dn == 0.0) {
        r.a0 = r.a1 = r.a2 = 1.0;
        r.b0 = g_c.bn0[0]; r.b1 = g_c.bn0[1]; r.b2 = g_c.bn0[2];
        return r;
    }
    double d = Dn / g_c.deln;
    pow3(1.0 - d, g_c.alpha, r.a0, r.a1, r.a2);
    pow3(g_c.pm_alpha + d, g_c.p_m, r.b0, r.b1, r.b2);
    return r;
}

__device__ __attribute__((always_inline)) inline PprPow ppr_pow_t(double Dt)
{
    PprPow r;
    if (Dt == 0.0) {
        r.a0 = r.a1 = r.a2 = 1.0;
        r.b0 = g_c.bt0[0]; r.b1 = g_c.bt0[1]; r.b2 = g_c.bt0[2];
        return r;
    }
    double d = Dt / g_c.delt;
    pow3(1.0 - d, g_c.beta, r.a0, r.a1, r.a2);
    pow3(g_c.pn_beta + d, g_c.p_n, r.b0, r.b1, r.b2);
    return r;
}

// Factors of the PPR expressions (PPR_math.h):
//  X = p A^a Bq^(p-1) - a A^(a-1) Bq^p           (the bracket of Tn_ / Tt_)
//  Y = gam A^a Bq^p + Macaulay                    (the cross-direction factor)
//  Z = (p^2-p) A^a Bq^(p-2) + (a^2-a) A^(a-2) Bq^p - 2 a p A^(a-1) Bq^(p-1)   (Dnn_/Dtt_)
__device__ inline double pprX(const PprPow &w, double p, double a) { return p * w.a0 * w.b1 - a * w.a1 * w.b0; }
__device__ inline double pprZ(const PprPow &w, double p, double a)
{
    return (p * p - p) * w.a0 * w.b2 + (a * a - a) * w.a2 * w.b0 - 2.0 * a * p * w.a1 * w.b1;
}
__device__ inline double Tn_at(const PprPow &n, const PprPow &t)
{
    return g_c.k_tn * pprX(n, g_c.p_m, g_c.alpha) * (g_c.gam_t * t.a0 * t.b0 + g_c.pMtn);
}
__device__ inline double Tt_at(const PprPow &n, const PprPow &t)
{
    return g_c.k_tt * pprX(t, g_c.p_n, g_c.beta) * (g_c.gam_n * n.a0 * n.b0 + g_c.pMnt);
}
__device__ inline double Dnn_at(const PprPow &n, const PprPow &t)
{
    return g_c.k_dnn * pprZ(n, g_c.p_m, g_c.alpha) * (g_c.gam_t * t.a0 * t.b0 + g_c.pMtn);
}
__device__ inline double Dtt_at(const PprPow &n, const PprPow &t)
{
    return g_c.k_dtt * pprZ(t, g_c.p_n, g_c.beta) * (g_c.gam_n * n.a0 * n.b0 + g_c.pMnt);
}
__device__ inline double Dnt_at(const PprPow &n, const PprPow &t)
{
    return g_c.k_dnt * pprX(n, g_c.p_m, g_c.alpha) * pprX(t, g_c.p_n, g_c.beta);
}

// PPR law with history, one Gauss point.  Same branches and failure tests as
// cohesive_law_thermodynamically_consistent (cz_kernels.h:37-174).
// L = {Tn, Tt, Dnn, Dtt, Dnt}; Dtn == Dnt.
__device__ inline void ppr_law(bool &contact, bool &failed, double &pmax, double &tmax, double opn, double opt,
                               double L[5])
{
    double Tn = 0, Tt = 0, Dnn = 0, Dtt = 0, Dnt = 0;
    if (opn > g_c.deln || opt > g_c.delt) {
        failed = true;
        L[0] = L[1] = L[2] = L[3] = L[4] = 0;
        return;
    }
    contact = (opn < 0);
    const double eps = -1e-9;
    const double thr_t = g_c.thr_t, thr_n = g_c.thr_n;
    // Every branch below draws on the powers at these openings only: evaluated once, up
    // front, so that a wavefront whose lanes take different branches does not evaluate
    // them once per branch.  (At zero opening the powers are constants.)
    const PprPow n0 = ppr_pow_n(0.0), t0 = ppr_pow_t(0.0);
    const PprPow n_opn = ppr_pow_n(contact ? 0.0 : opn);
    const PprPow n_pm = ppr_pow_n(pmax);
    const PprPow t_opt = ppr_pow_t(opt);
    const PprPow t_tm = ppr_pow_t(tmax);
    if (contact) {
        if (pmax != 0) {
            double peak = Tn_at(n_pm, t_tm);
            Tn = peak * opn / pmax;
            Dnn = peak / pmax;
        } else {
            Dnn = Dnn_at(n0, t_tm);
            Tn = Dnn * opn;
        }
        Tt = Tt_at(n0, t_opt);
        if (Tt >= eps && !(opt > g_c.lim_t && Tt < thr_t)) {
            if (opt >= tmax) {
                tmax = opt;
                Dtt = Dtt_at(n0, t_opt);
            } else {
                double peak = Tt_at(n0, t_tm);
                Tt = peak * opt / tmax;
                Dtt = peak / tmax;
            }
        } else {
            Tt = Dtt = Dnt = 0;
            Tn = Dnn = 0;
            failed = true;
        }
    } else {
        Tt = Tt_at(n_opn, t_opt);
        Tn = Tn_at(n_opn, t_opt);
        if (Tt >= eps && Tn >= eps && !(opt > g_c.lim_t && Tt < thr_t) && !(opn > g_c.lim_n && Tn < thr_n)) {
            const bool tsoft = (opt >= tmax), nsoft = (opn >= pmax);
            if (tsoft && nsoft) {
                tmax = opt;
                pmax = opn;
                Dnn = Dnn_at(n_opn, t_opt);
                Dnt = Dnt_at(n_opn, t_opt);
                Dtt = Dtt_at(n_opn, t_opt);
            } else if (tsoft && !nsoft) {
                Dnt = 0;
                if (pmax != 0) {
                    double peak = Tn_at(n_pm, t_tm);
                    Tn = peak * opn / pmax;
                    Dnn = peak / pmax;
                } else {
                    Tn = 0;
                    Dnn = Dnn_at(n0, t_tm);
                }
                tmax = opt;
                Tt = Tt_at(n_pm, t_opt);
                Dtt = Dtt_at(n_pm, t_opt);
            } else if (!tsoft && nsoft) {
                Dnt = 0;
                if (tmax != 0) {
                    double peak = Tt_at(n_pm, t_tm);
                    Tt = peak * opt / tmax;
                    Dtt = peak / tmax;
                } else {
                    Tt = 0;
                    Dtt = Dtt_at(n_pm, t0);
                }
                pmax = opn;
                Tn = Tn_at(n_opn, t_tm);
                Dnn = Dnn_at(n_opn, t_tm);
            } else {
                Dnt = 0;
                double peakN = Tn_at(n_pm, t_tm);
                if (pmax != 0) {
                    Tn = peakN * opn / pmax;
                    Dnn = peakN / pmax;
                } else {
                    Tn = 0;
                    Dnn = Dnn_at(n0, t_tm);
                }
                if (tmax != 0) {
                    double peakT = Tt_at(n_pm, t_tm);
                    Tt = peakT * opt / tmax;
                    Dtt = peakT / tmax;
                } else {
                    Tt = 0;
                    Dtt = Dtt_at(n_pm, t0);
                }
            }
        } else {
            failed = true;
            Tn = Tt = Dnn = Dtt = Dnt = 0;
        }
    }
    L[0] = Tn; L[1] = Tt; L[2] = Dnn; L[3] = Dtt; L[4] = Dnt;
}

// CZ midplane frame (rows e1, e2 = n x e1, e3 = n) and area, rounded like
// CZRotationMatrix (matrix_math.h:94-147).
__device__ inline void cz_frame(const double p[9], double R[9], double &area)
{
#pragma clang fp contract(off)
    double a0 = p[3] - p[0], a1 = p[4] - p[1], a2 = p[5] - p[2];
    double c0 = p[0] - p[6], c1 = p[1] - p[7], c2 = p[2] - p[8];
    double la = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
    R[0] = a0 / la; R[1] = a1 / la; R[2] = a2 / la;
    double n0 = -a2 * c1 + a1 * c2, n1 = a2 * c0 - a0 * c2, n2 = -a1 * c0 + a0 * c1;
    double ln = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    area = ln / 2;
    n0 /= ln; n1 /= ln; n2 /= ln;
    R[6] = n0; R[7] = n1; R[8] = n2;
    a0 /= la; a1 /= la; a2 /= la;
    double r0 = -n2 * a1 + n1 * a2, r1 = n2 * a0 - n0 * a2, r2 = -n1 * a0 + n0 * a1;
    double lr = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
    R[3] = r0 / lr; R[4] = r1 / lr; R[5] = r2 / lr;
}

// Phase 2: add tangent block (a,b), a,b in 0..5, of CZ record `rec` to acc.
template <typename T>
__device__ inline void cz_block(double acc[9], const T *__restrict__ rec, int a, int b)
{
    const int ia = a % 3, ib = b % 3;
    const T sg = ((a < 3) == (b < 3)) ? (T)1 : (T)-1;
    T blk[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int gp = 0; gp < 3; gp++) {
        const T w = sg * (T)g_c.sf[ia * 3 + gp] * (T)g_c.sf[ib * 3 + gp];
#pragma unroll
        for (int q = 0; q < 6; q++) blk[q] += w * rec[6 * gp + q];
    }
    acc[0] += (double)blk[0]; acc[1] += (double)blk[1]; acc[2] += (double)blk[2];
    acc[3] += (double)blk[1]; acc[4] += (double)blk[3]; acc[5] += (double)blk[4];
    acc[6] += (double)blk[2]; acc[7] += (double)blk[4]; acc[8] += (double)blk[5];
}

}  // namespace icf

// icf_cz_phase1.inc.h -- cohesive-zone phase 1 (included by icf_lib.hip after icf_cz.inc.h).
//
// Restates kczCZForce (cz_kernels9/cz_kernels.h:210-434).  Each CZ is shared by three lanes of
// a wavefront, one per Gauss point (21 CZs per 64-lane wave, lane 63 idle).  All three lanes
// compute the midplane frame and the local top-bottom differences (cheap, identical); each
// evaluates its own Gauss point's opening, PPR law, traction and tangent, writes its own
// Q_gp into the tangent record, and the per-CZ results are combined with wavefront shuffles
// in Gauss-point order 0, 1, 2 -- the order in which the reference accumulates them.

#pragma once

namespace icf {

constexpr int CZ_PER_WAVE = 21;

__device__ inline double shfl_d(double v, int src) { return __shfl(v, src, 64); }

template <bool MIXED>
__device__ __forceinline__ void cz_phase1(int bid_, int nb_, double *__restrict__ dcz, int *__restrict__ icz,
        const double *__restrict__ dn, int n, int cs, int ns, double *__restrict__ st_rhs, void *__restrict__ st_tan)
{
    const int wl = threadIdx.x & 63;                              // lane in the wave
    const int wave = (bid_ * blockDim.x + threadIdx.x) >> 6;
    const int slot = wl / 3, gp = wl % 3;                         // CZ within the wave, Gauss point
    const int base = slot * 3;                                    // lane of this CZ's gp 0
    int u = wave * CZ_PER_WAVE + slot;
    const bool live = wl < 63 && u < n;
    if (!live) u = 0;                                             // dead lanes still shuffle
    const bool prev_failed = live && icz[u + cs * CZI_FAILED] != 0;     // cz_kernels.h:219
    const bool work = live && !prev_failed;

    double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, area = 0, dx[9];
#pragma unroll
    for (int i = 0; i < 9; i++) dx[i] = 0;
    if (work) {
        double xc[18];
#pragma unroll
        for (int i = 0; i < 6; i++) {
            const int v = icz[u + cs * (CZI_VRTS + i)];
#pragma unroll
            for (int k = 0; k < 3; k++) xc[3 * i + k] = dn[v + ns * (ND_X0 + k)] + dn[v + ns * (ND_UN + k)];
        }
        {
#pragma clang fp contract(off)
            double mid[9];
#pragma unroll
            for (int i = 0; i < 9; i++) mid[i] = (xc[i] + xc[i + 9]) * 0.5;
            cz_frame(mid, R, area);
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    double top = xc[3 * i] * R[r * 3] + xc[3 * i + 1] * R[r * 3 + 1] + xc[3 * i + 2] * R[r * 3 + 2];
                    double bot = xc[3 * i + 9] * R[r * 3] + xc[3 * i + 10] * R[r * 3 + 1] + xc[3 * i + 11] * R[r * 3 + 2];
                    dx[3 * i + r] = top - bot;
                }
        }
    }

    // this lane's Gauss point
    double pmax = 0, tmax = 0;
    bool gc = false, gf = false;
    double opn = 0, opt = 0, Tn = 0, Tt = 0;
    double rhs9[9];        // contribution to the top nodes; the bottom nodes get its negative
    double q[6];
#pragma unroll
    for (int i = 0; i < 9; i++) rhs9[i] = 0;
#pragma unroll
    for (int i = 0; i < 6; i++) q[i] = 0;
    if (work) {
        pmax = dcz[u + cs * (CZD_PMAX + gp)];
        tmax = dcz[u + cs * (CZD_TMAX + gp)];
        double dt1, dt2, dnn;
        {
#pragma clang fp contract(off)
            dt1 = 0; dt2 = 0; dnn = 0;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const double s = g_c.sf[i * 3 + gp];
                dt1 += dx[3 * i + 0] * s;
                dt2 += dx[3 * i + 1] * s;
                dnn += dx[3 * i + 2] * s;
            }
        }
        opn = dnn;
        opt = sqrt(dt1 * dt1 + dt2 * dt2);
        double L[5];
        ppr_law(gc, gf, pmax, tmax, opn, opt, L);
        Tn = L[0];
        Tt = L[1];
        const double Dnn = L[2], Dtt = L[3], Dnt = L[4];
        double T0, T1, T2 = Tn, Td[9];
        if (opt < 1e-20) {
            T0 = T1 = 0;
            Td[0] = Dtt; Td[1] = 0; Td[2] = Dnt;
            Td[3] = 0; Td[4] = Dtt; Td[5] = Dnt;
            Td[6] = Dnt; Td[7] = Dnt; Td[8] = Dnn;
        } else {
            T0 = Tt * dt1 / opt;
            T1 = Tt * dt2 / opt;
            const double o2 = opt * opt, o3 = o2 * opt;
            const double d00 = dt1 * dt1, d10 = dt2 * dt1, d11 = dt2 * dt2;
            Td[0] = Dtt * d00 / o2 + Tt * d11 / o3;
            Td[4] = Dtt * d11 / o2 + Tt * d00 / o3;
            Td[8] = Dnn;
            Td[1] = Td[3] = Dtt * d10 / o2 - Tt * d10 / o3;
            Td[2] = Td[6] = Dnt * dt1 / opt;
            Td[5] = Td[7] = Dnt * dt2 / opt;
        }
        const double wa = (1.0 / 3.0) * area;
        double Tg[3];
#pragma unroll
        for (int r = 0; r < 3; r++) Tg[r] = R[r] * T0 + R[3 + r] * T1 + R[6 + r] * T2;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const double c = g_c.sf[a * 3 + gp] * wa;
#pragma unroll
            for (int r = 0; r < 3; r++) rhs9[3 * a + r] = -(c * Tg[r]);
        }
        double TR[9];
#pragma unroll
        for (int m = 0; m < 3; m++)
#pragma unroll
            for (int l = 0; l < 3; l++) TR[m * 3 + l] = Td[m * 3] * R[l] + Td[m * 3 + 1] * R[3 + l] + Td[m * 3 + 2] * R[6 + l];
        q[0] = wa * (R[0] * TR[0] + R[3] * TR[3] + R[6] * TR[6]);
        q[1] = wa * (R[0] * TR[1] + R[3] * TR[4] + R[6] * TR[7]);
        q[2] = wa * (R[0] * TR[2] + R[3] * TR[5] + R[6] * TR[8]);
        q[3] = wa * (R[1] * TR[1] + R[4] * TR[4] + R[7] * TR[7]);
        q[4] = wa * (R[1] * TR[2] + R[4] * TR[5] + R[7] * TR[8]);
        q[5] = wa * (R[2] * TR[2] + R[5] * TR[5] + R[8] * TR[8]);
    }

    // ---- combine the three Gauss points (every lane ends with the CZ's totals) ----
    const int l0 = base, l1 = base + 1, l2 = base + 2;
    auto sum3 = [&](double v) { return (shfl_d(v, l0) + shfl_d(v, l1)) + shfl_d(v, l2); };
    auto max3 = [&](double v) { return fmax(fmax(shfl_d(v, l0), shfl_d(v, l1)), shfl_d(v, l2)); };
    auto any3 = [&](int v) { return __shfl(v, l0, 64) | __shfl(v, l1, 64) | __shfl(v, l2, 64); };
    const double pm = max3(pmax), tm = max3(tmax);
    const int dmg = any3((pmax >= g_c.deln * g_c.rn || tmax >= g_c.delt * g_c.rt) ? 1 : 0);
    const int failed = any3(gf ? 1 : 0), contact = any3(gc ? 1 : 0);
    const double avgDn = (((0.0 + shfl_d(opn / 3, l0)) + shfl_d(opn / 3, l1)) + shfl_d(opn / 3, l2));
    const double avgDt = (((0.0 + shfl_d(opt / 3, l0)) + shfl_d(opt / 3, l1)) + shfl_d(opt / 3, l2));
    const double avgTn = (((0.0 + shfl_d(Tn / 3, l0)) + shfl_d(Tn / 3, l1)) + shfl_d(Tn / 3, l2));
    const double avgTt = (((0.0 + shfl_d(Tt / 3, l0)) + shfl_d(Tt / 3, l1)) + shfl_d(Tt / 3, l2));
    // lane gp owns top node gp and bottom node gp + 3 of the residual
    double own[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        double a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
        for (int a = 0; a < 3; a++) {     // every lane shuffles every node; keeps node `gp`
            const double v0 = shfl_d(rhs9[3 * a + r], l0), v1 = shfl_d(rhs9[3 * a + r], l1), v2 = shfl_d(rhs9[3 * a + r], l2);
            if (a == gp) { a0 = v0; a1 = v1; a2 = v2; }
        }
        own[r] = ((0.0 + a0) + a1) + a2;
    }
    if (!live) return;
    double *orhs = st_rhs + (size_t)u * CZ_RHS_REC;
    if (!work || failed) {
#pragma unroll
        for (int r = 0; r < 3; r++) {
            orhs[3 * gp + r] = 0.;
            orhs[9 + 3 * gp + r] = 0.;
        }
        if (MIXED) {
            float *t = (float *)st_tan + (size_t)u * CZ_TAN_REC + 6 * gp;
#pragma unroll
            for (int i = 0; i < 6; i++) t[i] = 0.f;
        } else {
            double *t = (double *)st_tan + (size_t)u * CZ_TAN_REC + 6 * gp;
#pragma unroll
            for (int i = 0; i < 6; i++) t[i] = 0.;
        }
        if (!work) return;               // currently failed: no outputs at all (cz_kernels.h:219)
    }
    // tentative history and flags (cz_kernels.h:404-422): lane gp writes slot gp
    dcz[u + cs * (CZD_TPMAX + gp)] = pm;
    dcz[u + cs * (CZD_TTMAX + gp)] = tm;
    if (gp == 0) {
        icz[u + cs * CZI_TCONTACT] = contact;
        icz[u + cs * CZI_TFAILED] = failed;
        icz[u + cs * CZI_TDAMAGED] = failed ? 0 : dmg;
    }
    if (failed) return;
    if (gp == 0) {
        dcz[u + cs * CZD_DN] = avgDn;
        dcz[u + cs * CZD_DT] = avgDt;
        dcz[u + cs * CZD_TN] = avgTn;
        dcz[u + cs * CZD_TT] = avgTt;
    }
#pragma unroll
    for (int r = 0; r < 3; r++) {
        orhs[3 * gp + r] = own[r];
        orhs[9 + 3 * gp + r] = -own[r];
    }
    if (MIXED) {
        float *t = (float *)st_tan + (size_t)u * CZ_TAN_REC + 6 * gp;
#pragma unroll
        for (int i = 0; i < 6; i++) t[i] = (float)q[i];
    } else {
        double *t = (double *)st_tan + (size_t)u * CZ_TAN_REC + 6 * gp;
#pragma unroll
        for (int i = 0; i < 6; i++) t[i] = q[i];
    }
}

template <bool MIXED>
__global__ __launch_bounds__(256) void k_cz_phase1(double *__restrict__ dcz, int *__restrict__ icz,
                                                   const double *__restrict__ dn, int n, int cs, int ns,
                                                   double *__restrict__ st_rhs, void *__restrict__ st_tan)
{
    cz_phase1<MIXED>(blockIdx.x, gridDim.x, dcz, icz, dn, n, cs, ns, st_rhs, st_tan);
}

}  // namespace icf

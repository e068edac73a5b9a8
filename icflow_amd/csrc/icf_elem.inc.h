// icf_elem.inc.h -- corotational linear tetrahedron (included by icf_lib.hip).
//
// Phase 1 (k_tet_phase1, one thread per tet) computes the tet's residual contribution in FP64
// and a compact tangent record; phase 2 (the slot reduction in icf_reduce.inc.h) rebuilds
// each 3x3 tangent block from that record inside the BCSR slot that owns it.
//
// Restated from elem_kernels9/element_kernels2.h:19-263 (kelElementElasticityForce and
// F_and_Df_Corotational) and matrix_math1.h:32-71 (fastRotationMatrix).  Differences of
// formulation (same mathematics):
//  - K = B^T E B V is never formed: K y is applied as V B^T (E (B y)) (B is sparse);
//  - the tangent block (i,j) is R K_ij R^T; for isotropic E it equals
//    V (lam g~_i g~_j^T + mu g~_j g~_i^T + mu (g~_i . g~_j) I) with corotated gradients
//    g~ = R g, so only g~ and V are staged (16 values per tet).
#pragma once

namespace icf {

// XCD-aware block order: the hardware deals workgroups round-robin over the 8 XCDs, each of
// which has its own L2.  This maps the physical block id so that the blocks running on one
// XCD cover one contiguous range of units (neighbouring tets share nodes).  Bijective on
// [0, nb).
__device__ inline int xcd_remap(int b, int nb)
{
    const int X = 8;
    int q = nb / X, r = nb % X;
    int x = b % X, k = b / X;
    return x * q + (x < r ? x : r) + k;
}

// Orthonormal frame of three points; columns = (p1-p0), (p1-p0)x(p2-p0), col1 x col0, each
// normalised.  R[row*3+col].  Same construction as fastRotationMatrix (matrix_math1.h:57-96);
// each column is scaled by one reciprocal of its length.
__device__ inline void frame3(const double *p0, const double *p1, const double *p2, double R[9])
{
    double d0 = p1[0] - p0[0], d1 = p1[1] - p0[1], d2 = p1[2] - p0[2];
    double il = 1.0 / sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    R[0] = d0 * il; R[3] = d1 * il; R[6] = d2 * il;
    double w0 = p2[0] - p0[0], w1 = p2[1] - p0[1], w2 = p2[2] - p0[2];
    double c0 = -d2 * w1 + d1 * w2;
    double c1 = d2 * w0 - d0 * w2;
    double c2 = -d1 * w0 + d0 * w1;
    il = 1.0 / sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    R[1] = c0 * il; R[4] = c1 * il; R[7] = c2 * il;
    double e0 = R[4] * R[6] - R[3] * R[7];
    double e1 = -R[1] * R[6] + R[0] * R[7];
    double e2 = R[1] * R[3] - R[0] * R[4];
    il = 1.0 / sqrt(e0 * e0 + e1 * e1 + e2 * e2);
    R[2] = e0 * il; R[5] = e1 * il; R[8] = e2 * il;
}

// Rest gradients of the four linear shape functions, g[n*3+k], and the rest volume.
// Same cofactor formulas as element_kernels2.h:41-57.
__device__ inline void rest_gradients(const double x0[12], double g[12], double &V)
{
#pragma clang fp contract(off)
    const double X1 = x0[0], Y1 = x0[1], Z1 = x0[2], X2 = x0[3], Y2 = x0[4], Z2 = x0[5];
    const double X3 = x0[6], Y3 = x0[7], Z3 = x0[8], X4 = x0[9], Y4 = x0[10], Z4 = x0[11];
    double x12 = X1 - X2, x13 = X1 - X3, x14 = X1 - X4, x23 = X2 - X3, x24 = X2 - X4, x34 = X3 - X4;
    double y12 = Y1 - Y2, y13 = Y1 - Y3, y14 = Y1 - Y4, y23 = Y2 - Y3, y24 = Y2 - Y4, y34 = Y3 - Y4;
    double z12 = Z1 - Z2, z13 = Z1 - Z3, z14 = Z1 - Z4, z23 = Z2 - Z3, z24 = Z2 - Z4, z34 = Z3 - Z4;
    double x21 = -x12, x31 = -x13, x32 = -x23, x42 = -x24, x43 = -x34;
    double y31 = -y13, y32 = -y23, y42 = -y24, y43 = -y34, y21 = -y12;
    double z21 = -z12, z31 = -z13, z32 = -z23, z42 = -z24, z43 = -z34;
    double J = x21 * (y23 * z34 - y34 * z23) + x32 * (y34 * z12 - y12 * z34) + x43 * (y12 * z23 - y23 * z12);
    V = J / 6.;
    g[0] = y42 * z32 - y32 * z42;  g[1] = x32 * z42 - x42 * z32;  g[2] = x42 * y32 - x32 * y42;
    g[3] = y31 * z43 - y34 * z13;  g[4] = x43 * z31 - x13 * z34;  g[5] = x31 * y43 - x34 * y13;
    g[6] = y24 * z14 - y14 * z24;  g[7] = x14 * z24 - x24 * z14;  g[8] = x24 * y14 - x14 * y24;
    g[9] = y13 * z21 - y12 * z31;  g[10] = x21 * z13 - x31 * z12; g[11] = x13 * y21 - x12 * y31;
    const double iJ = 1.0 / J;
#pragma unroll
    for (int i = 0; i < 12; i++) g[i] *= iJ;
}

// out = V B^T E B y for the 12-vector y (the rest stiffness applied to y).  ISO: E is the
// isotropic lam/mu form (checked by icf_set_constants), so E e = lam tr(e) + 2 mu e (normal
// strains), mu e (engineering shear strains).
template <bool ISO>
__device__ inline void apply_K0(const double g[12], double V, const double y[12], double out[12])
{
    double e[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int n = 0; n < 4; n++) {
        const double a = g[3 * n], b = g[3 * n + 1], c = g[3 * n + 2];
        const double yx = y[3 * n], yy = y[3 * n + 1], yz = y[3 * n + 2];
        e[0] += a * yx;
        e[1] += b * yy;
        e[2] += c * yz;
        e[3] += b * yx + a * yy;
        e[4] += c * yy + b * yz;
        e[5] += c * yx + a * yz;
    }
    double s[6];
    if (ISO) {
        const double lv = g_c.lam * (e[0] + e[1] + e[2]), mu = g_c.mu;
        s[0] = (lv + 2 * mu * e[0]) * V;
        s[1] = (lv + 2 * mu * e[1]) * V;
        s[2] = (lv + 2 * mu * e[2]) * V;
        s[3] = mu * e[3] * V;
        s[4] = mu * e[4] * V;
        s[5] = mu * e[5] * V;
    } else {
#pragma unroll
        for (int r = 0; r < 6; r++) {
            double v = 0;
#pragma unroll
            for (int k = 0; k < 6; k++) v += g_c.E[r * 6 + k] * e[k];
            s[r] = v * V;
        }
    }
#pragma unroll
    for (int n = 0; n < 4; n++) {
        const double a = g[3 * n], b = g[3 * n + 1], c = g[3 * n + 2];
        out[3 * n + 0] = a * s[0] + b * s[3] + c * s[5];
        out[3 * n + 1] = b * s[1] + a * s[3] + c * s[4];
        out[3 * n + 2] = c * s[2] + b * s[4] + a * s[5];
    }
}

// Phase 1 for tets.  rhs (12 per tet) = gravity - f - dampK * Df vn - V M (dampM vn + an)
// (element_kernels2.h:228-251), written FP64 to st_rhs; tangent record to st_tan
// (double, or float when MIXED).
template <bool MIXED, bool ISO, bool NT>
__device__ __forceinline__ void tet_phase1(int bid_, int nb_, const int *__restrict__ ie,
        const double *__restrict__ dn, int n, int tp, int es, int ns, double *__restrict__ st_rhs,
        void *__restrict__ st_tan)
{
    const int nb = nb_;
    const int u = xcd_remap(bid_, nb) * blockDim.x + threadIdx.x;
    if (u >= n) return;

    int nid[4];
#pragma unroll
    for (int i = 0; i < 4; i++) nid[i] = ie[u + es * (EL_NODES + i)];
    double x0[12], un[12];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            x0[3 * i + k] = dn[nid[i] + ns * (ND_X0 + k)];
            un[3 * i + k] = dn[nid[i] + ns * (ND_UN + k)];
        }

    double g[12], V;
    rest_gradients(x0, g, V);

    double R[9];
    {
        double R0[9], R1[9], xc[9];
#pragma unroll
        for (int i = 0; i < 9; i++) xc[i] = x0[i] + un[i];
        frame3(&x0[0], &x0[3], &x0[6], R0);
        frame3(&xc[0], &xc[3], &xc[6], R1);
        {
#pragma clang fp contract(off)
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++)
                    R[i * 3 + j] = R1[i * 3 + 0] * R0[j * 3 + 0] + R1[i * 3 + 1] * R0[j * 3 + 1] + R1[i * 3 + 2] * R0[j * 3 + 2];
        }
    }

    // y = R^T x_c - x0, rounded like element_kernels2.h:127-150
    double y[12];
    {
#pragma clang fp contract(off)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            double p0 = x0[3 * i] + un[3 * i], p1 = x0[3 * i + 1] + un[3 * i + 1], p2 = x0[3 * i + 2] + un[3 * i + 2];
            y[3 * i + 0] = (p0 * R[0] + p1 * R[3] + p2 * R[6]) - x0[3 * i + 0];
            y[3 * i + 1] = (p0 * R[1] + p1 * R[4] + p2 * R[7]) - x0[3 * i + 1];
            y[3 * i + 2] = (p0 * R[2] + p1 * R[5] + p2 * R[8]) - x0[3 * i + 2];
        }
    }

    // elastic force and Rayleigh stiffness damping together: f + dampK Df vn
    // = R K0 (R^T x_c - x0) + dampK R K0 R^T vn = R K0 (y + dampK R^T vn)
    double rhs[12];
    {
        const double dK = g_c.dampK;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            double v0 = dn[nid[i] + ns * (ND_VN + 0)], v1 = dn[nid[i] + ns * (ND_VN + 1)], v2 = dn[nid[i] + ns * (ND_VN + 2)];
            y[3 * i + 0] += dK * (R[0] * v0 + R[3] * v1 + R[6] * v2);
            y[3 * i + 1] += dK * (R[1] * v0 + R[4] * v1 + R[7] * v2);
            y[3 * i + 2] += dK * (R[2] * v0 + R[5] * v1 + R[8] * v2);
        }
        double Ky[12];
        apply_K0<ISO>(g, V, y, Ky);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int k = 0; k < 3; k++)
                rhs[3 * i + k] = -(R[k * 3 + 0] * Ky[3 * i] + R[k * 3 + 1] * Ky[3 * i + 1] + R[k * 3 + 2] * Ky[3 * i + 2]);
    }
    {
        // inertia and mass damping: V M (dampM vn + an); gravity rho g V / 4 on z
        double q[12];
        const double dM = g_c.dampM;
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int k = 0; k < 3; k++)
                q[3 * i + k] = dM * dn[nid[i] + ns * (ND_VN + k)] + dn[nid[i] + ns * (ND_AN + k)];
        const double gz = g_c.grav * g_c.rho * V / 4;
#pragma unroll
        for (int r = 0; r < 12; r++) {
            double m = 0;
            if (ISO) {            // M = m_ij I3 per node pair
#pragma unroll
                for (int j = 0; j < 4; j++) m += g_c.m4[(r / 3) * 4 + j] * q[3 * j + r % 3];
            } else {
#pragma unroll
                for (int c = 0; c < 12; c++) m += g_c.M[r * 12 + c] * q[c];
            }
            rhs[r] -= V * m;
            if (r % 3 == 2) rhs[r] += gz;
        }
    }
    // node i's part at plane i (tet_rhs_at), so that each store is one coalesced run; NT:
    // nontemporal stores, when the staged records are too large to stay in the caches until
    // phase 2 reads them, so that they do not evict the node fields phase 1 gathers
#pragma unroll
    for (int i = 0; i < 4; i++) {
        double *o = st_rhs + ((size_t)i * tp + u) * 3;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if (NT) __builtin_nontemporal_store(rhs[3 * i + k], &o[k]);
            else o[k] = rhs[3 * i + k];
        }
    }

    if (ISO) {
        double gt[12];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int k = 0; k < 3; k++)
                gt[3 * i + k] = R[k * 3 + 0] * g[3 * i] + R[k * 3 + 1] * g[3 * i + 1] + R[k * 3 + 2] * g[3 * i + 2];
        // record: four (g~_i, V) quadruples, node i's at plane i, so that phase 2 fetches node
        // i's gradient and the volume with one 16-byte (FP32) or two 16-byte (FP64) loads and
        // phase 1 stores one coalesced run per node
        if (MIXED) {
            float4 *t = (float4 *)st_tan;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float4 v = make_float4((float)gt[3 * i], (float)gt[3 * i + 1], (float)gt[3 * i + 2], (float)V);
                float *tf = (float *)&t[(size_t)i * tp + u];
                if (NT) {
                    __builtin_nontemporal_store(v.x, tf);
                    __builtin_nontemporal_store(v.y, tf + 1);
                    __builtin_nontemporal_store(v.z, tf + 2);
                    __builtin_nontemporal_store(v.w, tf + 3);
                } else {
                    t[(size_t)i * tp + u] = v;
                }
            }
        } else {
            double2 *t = (double2 *)st_tan;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                t[((size_t)i * tp + u) * 2] = make_double2(gt[3 * i], gt[3 * i + 1]);
                t[((size_t)i * tp + u) * 2 + 1] = make_double2(gt[3 * i + 2], V);
            }
        }
    } else {
        if (MIXED) {
            float *t = (float *)st_tan + (size_t)u * TET_GEN_REC;
#pragma unroll
            for (int i = 0; i < 9; i++) t[i] = (float)R[i];
#pragma unroll
            for (int i = 0; i < 12; i++) t[9 + i] = (float)g[i];
            t[21] = (float)V;
            t[22] = t[23] = 0.f;
        } else {
            double *t = (double *)st_tan + (size_t)u * TET_GEN_REC;
#pragma unroll
            for (int i = 0; i < 9; i++) t[i] = R[i];
#pragma unroll
            for (int i = 0; i < 12; i++) t[9 + i] = g[i];
            t[21] = V;
            t[22] = t[23] = 0.;
        }
    }
}

template <bool MIXED, bool ISO, bool NT>
__global__ __launch_bounds__(256) void k_tet_phase1(const int *__restrict__ ie, const double *__restrict__ dn,
                                                    int n, int tp, int es, int ns, double *__restrict__ st_rhs,
                                                    void *__restrict__ st_tan)
{
    tet_phase1<MIXED, ISO, NT>(blockIdx.x, gridDim.x, ie, dn, n, tp, es, ns, st_rhs, st_tan);
}

// Phase 2: LHS block (i,j) of an isotropic tet record.  LHS = stiff * R K R^T + mcoef V M
// with stiff = 1 + dampK gamma / (h beta), mcoef = (1/h^2 + dampM gamma / h) / beta
// (element_kernels2.h:239-249); with isotropic E the stiffness block is
// V (lam g~_i g~_j^T + mu g~_j g~_i^T + mu (g~_i . g~_j) I), and the mass block is m_ij I
// (the supported form of M).
template <typename T>
struct TetIsoLoad;

template <>
struct TetIsoLoad<float> {
    float4 gi, gj;   // (g~, V)
    int ij;
    __device__ inline void load(const float *__restrict__ base, int tp, uint32_t code)
    {
        const uint32_t q = code_q(code);
        const float4 *rec = (const float4 *)base + code_unit(code);
        ij = (int)q;
        gi = rec[(size_t)(q >> 2) * tp];
        gj = rec[(size_t)(q & 3) * tp];
    }
    __device__ inline void values(float w[9], float stiff, float mcoef) const
    {
        const float V = gi.w, c1 = stiff * V, cl = c1 * g_c.lamf, cu = c1 * g_c.muf;
        const float dg = gi.x * gj.x + gi.y * gj.y + gi.z * gj.z;
        const float dd = cu * dg + mcoef * V * g_c.m4f[ij];
        const float a[3] = {gi.x, gi.y, gi.z}, b[3] = {gj.x, gj.y, gj.z};
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int l = 0; l < 3; l++) {
                float v = cl * a[k] * b[l] + cu * b[k] * a[l];
                if (k == l) v += dd;
                w[3 * k + l] = v;
            }
    }
};

template <>
struct TetIsoLoad<double> {
    double2 i0, i1, j0, j1;   // (g~x, g~y), (g~z, V)
    int ij;
    __device__ inline void load(const double *__restrict__ base, int tp, uint32_t code)
    {
        const uint32_t q = code_q(code);
        const double2 *rec = (const double2 *)base + (size_t)code_unit(code) * 2;
        ij = (int)q;
        i0 = rec[(size_t)(q >> 2) * tp * 2];
        i1 = rec[(size_t)(q >> 2) * tp * 2 + 1];
        j0 = rec[(size_t)(q & 3) * tp * 2];
        j1 = rec[(size_t)(q & 3) * tp * 2 + 1];
    }
    __device__ inline void values(double w[9], double stiff, double mcoef) const
    {
        const double V = i1.y, c1 = stiff * V, cl = c1 * g_c.lam, cu = c1 * g_c.mu;
        const double a[3] = {i0.x, i0.y, i1.x}, b[3] = {j0.x, j0.y, j1.x};
        const double dg = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
        const double dd = cu * dg + mcoef * V * g_c.m4[ij];
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int l = 0; l < 3; l++) {
                double v = cl * a[k] * b[l] + cu * b[k] * a[l];
                if (k == l) v += dd;
                w[3 * k + l] = v;
            }
    }
};

template <typename T>
__device__ inline void tet_block_gen(double acc[9], const T *__restrict__ rec, int i, int j, T stiff, T mcoef)
{
    T R[9];
#pragma unroll
    for (int q = 0; q < 9; q++) R[q] = rec[q];
    const T *g = rec + 9;
    const T V = rec[21];
    const T ai = g[3 * i], bi = g[3 * i + 1], ci = g[3 * i + 2];
    const T aj = g[3 * j], bj = g[3 * j + 1], cj = g[3 * j + 2];
    auto E = [](int r, int c) -> T { return sizeof(T) == 4 ? (T)g_c.Ef[r * 6 + c] : (T)g_c.E[r * 6 + c]; };
    // columns of B_j (Voigt xx,yy,zz,xy,yz,zx): x: (a,0,0,b,0,c)  y: (0,b,0,a,c,0)  z: (0,0,c,0,b,a)
    T EB[6][3];
#pragma unroll
    for (int r = 0; r < 6; r++) {
        EB[r][0] = E(r, 0) * aj + E(r, 3) * bj + E(r, 5) * cj;
        EB[r][1] = E(r, 1) * bj + E(r, 3) * aj + E(r, 4) * cj;
        EB[r][2] = E(r, 2) * cj + E(r, 4) * bj + E(r, 5) * aj;
    }
    T K[3][3];
#pragma unroll
    for (int l = 0; l < 3; l++) {
        K[0][l] = (ai * EB[0][l] + bi * EB[3][l] + ci * EB[5][l]) * V;
        K[1][l] = (bi * EB[1][l] + ai * EB[3][l] + ci * EB[4][l]) * V;
        K[2][l] = (ci * EB[2][l] + bi * EB[4][l] + ai * EB[5][l]) * V;
    }
    T RK[3][3];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int l = 0; l < 3; l++) RK[k][l] = R[k * 3] * K[0][l] + R[k * 3 + 1] * K[1][l] + R[k * 3 + 2] * K[2][l];
    const T cm = mcoef * V;
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int l = 0; l < 3; l++) {
            const int mi = (3 * i + k) * 12 + 3 * j + l;
            const T m = sizeof(T) == 4 ? (T)g_c.Mf[mi] : (T)g_c.M[mi];
            T v = stiff * (RK[k][0] * R[l * 3] + RK[k][1] * R[l * 3 + 1] + RK[k][2] * R[l * 3 + 2]) + cm * m;
            acc[3 * k + l] += (double)v;
        }
}

}  // namespace icf

// icf_device.h -- device-side constants, layouts and math helpers shared by every kernel of
// libicflow_hip.so.  Included once, by icf_lib.hip (single translation unit, so one
// __constant__ block serves all kernels).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace icf {

// ---- SoA field offsets of the reference buffers -------------------------------------------
// node array (double, 18 x nd_stride): GPU_Functionality.cs:55-61
constexpr int ND_X0 = 0, ND_UN = 3, ND_VN = 6, ND_AN = 9, ND_F = 12, ND_XC = 15;
// elastic tet (int, 28 x el_stride): elem_kernels9/element_kernels2.h:2-5
constexpr int EL_NODES = 0, EL_SLOTS = 4, EL_ROWS = 20;
// CZ int (58 x cz_stride) / double (16 x cz_stride): cz_kernels9/cz_kernels.h:13-32
constexpr int CZI_FAILED = 0, CZI_TCONTACT = 1, CZI_TDAMAGED = 2, CZI_TFAILED = 3, CZI_VRTS = 4,
              CZI_SLOTS = 10, CZI_ROWS = 46;
constexpr int CZD_PMAX = 0, CZD_TMAX = 3, CZD_TPMAX = 6, CZD_TTMAX = 9, CZD_DN = 12, CZD_DT = 13,
              CZD_TN = 14, CZD_TT = 15;
// impact (int, 28 x cr_stride): collision_response.h:48-50, GPU_Functionality.cs:643-663
constexpr int CR_NODES = 0, CR_SLOTS = 4, CR_ROWS = 20;

// ---- kinds of contributing units (top bits of a slot-map code) ----------------------------
constexpr uint32_t KIND_TET = 0, KIND_CZ = 1, KIND_CONTACT = 2;
constexpr int CODE_UNIT_BITS = 24;
constexpr uint32_t CODE_UNIT_MAX = (1u << CODE_UNIT_BITS) - 1;
__host__ __device__ inline uint32_t make_code(uint32_t kind, uint32_t unit, uint32_t q)
{
    return (kind << 30) | (unit << 6) | q;
}
__host__ __device__ inline uint32_t code_kind(uint32_t c) { return c >> 30; }
__host__ __device__ inline uint32_t code_unit(uint32_t c) { return (c >> 6) & CODE_UNIT_MAX; }
__host__ __device__ inline uint32_t code_q(uint32_t c) { return c & 63u; }

// ---- staged per-unit records (AoS, one record per unit) -----------------------------------
// tet tangent, isotropic E: corotated gradients g~ = R g (12), rest volume V, 3 pad.
constexpr int TET_ISO_REC = 16;
// tet tangent, general E: R (9, row-major), rest gradients g (12), V, 2 pad.
constexpr int TET_GEN_REC = 24;
// tet residual: 12 doubles.
constexpr int TET_RHS_REC = 12;
// CZ tangent: per Gauss point the symmetric 3x3 Q = (w a_J) R^T T_d R as xx,xy,xz,yy,yz,zz.
constexpr int CZ_TAN_REC = 18;
// CZ residual: 18 doubles.
constexpr int CZ_RHS_REC = 18;
// contact: 78 packed upper-triangle entries of (k/2) Hess d^2, then 12 forces, 6 pad.
// impact records: the 10 upper 3x3 blocks of the 12x12 tangent, block (a, b) a <= b at
// ct_block(a, b), each whole and row-major (90 values, FP32 in mixed mode), and the 12 node
// forces (FP64) in a separate array
constexpr int CT_TAN_REC = 90, CT_F_REC = 12;
__host__ __device__ inline int ct_block(int a, int b) { return a * 4 - (a * (a - 1)) / 2 + (b - a); }

// ---- model constants -----------------------------------------------------------------------
struct DevConsts {
    double E[36];
    double M[144];
    double nbeta, ngamma, dampM, dampK, rho, grav;
    double G_fn, G_ft, f_tn, f_tt, alpha, beta, rn, rt;
    double p_m, p_n, deln, delt, pMtn, pMnt, gam_n, gam_t;
    double sf[9];           // sf[node*3 + gp]
    double lam, mu;         // isotropic split of E (valid when iso != 0)
    double m4[16];          // M[(3i)*12 + 3j]: the mass block (i,j) is m4[4i+j] I3 when iso != 0
    int iso;
    // constant factors of the PPR expressions, formed on the host with the same operations
    // and order as in PPR_math.h / cz_kernels.h, so they are the same doubles
    double k_tn, k_tt;              // gam_n / deln, gam_t / delt
    double k_dnn, k_dtt, k_dnt;     // gam_n / (deln deln), gam_t / (delt delt), gam_n gam_t / (deln delt)
    double pm_alpha, pn_beta;       // p_m / alpha, p_n / beta
    double thr_n, thr_t;            // f_tn * 0.05, f_tt * 0.05
    double lim_n, lim_t;            // deln * rn / 5, delt * rt / 5
    // powers at zero opening: (1 - 0)^e = 1 and (p/a + 0)^e, used by Dnn_(0, .) etc.
    double bn0[3], bt0[3];  // (p_m/alpha)^{p_m, p_m-1, p_m-2}, (p_n/beta)^{p_n, p_n-1, p_n-2}
    float Ef[36];
    float Mf[144];
    float lamf, muf;
    float m4f[16];
};

// ---- small vector helpers --------------------------------------------------------------------
template <typename T>
__device__ inline T dot3(const T *a, const T *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Index of entry (r, c), r <= c, in a packed upper triangle of a 12x12 matrix.
__host__ __device__ inline int upper12(int r, int c) { return r * 12 - (r * (r - 1)) / 2 + (c - r); }

}  // namespace icf

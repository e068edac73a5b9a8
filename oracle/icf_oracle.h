/*
 * icf_oracle.h -- CPU (FP64, sequential) restatement of icFlow's per-Newton-iteration
 * assembly math.  TEST INFRASTRUCTURE ONLY: the checker that the HIP library is compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it;
 * the product path never does.
 *
 * Every function restates one reference routine and cites it (paths relative to the
 * reference repository root).  Kernels are restated as plain loops over the unit index in
 * increasing order, so every BCSR / rhs / node-force accumulation happens in one fixed
 * order: unit 0..n-1, and within a unit the reference's own loop order.
 *
 * Buffer layouts are the reference's SoA layouts: value(item, field) = buf[item + stride*field].
 */
#ifndef ICF_ORACLE_H
#define ICF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Model constants: the __constant__ symbols of elem_kernels9/kernel.cu:6-9 and
 * cz_kernels9/kernel.cu:5-20 (uploaded by GPU_Functionality.cs:151-194). */
typedef struct orc_consts {
    double E[36];          /* E[i][j] row-major 6x6 (elem_kernels9/kernel.cu:6)          */
    double M[144];         /* M[i][j] row-major 12x12, per unit volume (kernel.cu:7)       */
    double NewmarkBeta, NewmarkGamma, dampingMass, dampingStiffness, rho, gravity;
    double YoungsModulus;
    double G_fn, G_ft, f_tn, f_tt, alpha, beta, rn, rt;
    double p_m, p_n, deln, delt, pMtn, pMnt, gam_n, gam_t;
    double B[162];         /* B[gp][3][18] (cz_kernels9/kernel.cu:19)                      */
    double sf[9];          /* sf[node][gp] (cz_kernels9/kernel.cu:20)                      */
} orc_consts;

void orc_set_constants(const orc_consts *c);

/* --- element (elem_kernels9) --- */
void orc_F_and_Df(const double x0[12], const double un[12], double f[12], double Df[144], double *V);
void orc_elem_force(const int *ie, double *dn, double h, double *A, double *b,
                    int nElems, int el_stride, int nd_stride);

/* --- cohesive zones (cz_kernels9) --- */
double orc_Tn(double Dn, double Dt);
double orc_Tt(double Dn, double Dt);
double orc_Dnn(double Dn, double Dt);
double orc_Dtt(double Dn, double Dt);
double orc_Dnt(double Dn, double Dt);
/* out[6] = Tn, Tt, Dnn, Dtt, Dnt, Dtn; flags[2] = contact, failed (in/out, OR-ed) */
void orc_cohesive_law(int flags[2], double *pmax, double *tmax, double opn, double opt, double out[6]);
void orc_cz_rotation(const double p[9], double R[9], double *area);
void orc_cz_force(double *dcz, int *icz, double *dn, double *A, double *b, double h,
                  int nCZs, int cz_stride, int nd_stride);

/* --- contact (collision_kernels9) --- */
int    orc_ctest(const double bmat[9], double x, double y, double z);
void   orc_bvalues(const double p[12], double bmat[9]);
void   orc_narrow_phase(int nPairs, const double *dn, const int *ie, int *narrowList,
                        int el_all_stride, int nd_stride);
double orc_dtn(const double f[9], const double p[3]);
void   orc_find_closest_face(int nPairs, int tet_stride, int *nl2, const double *dn, const int *ie,
                             const int *faces, int nd_stride, int el_all_stride, int fc_stride);
double orc_pt(const double x[12], double fd[12], double sd[144], double *s, double *t, int *branch);
void   orc_collision_response(const int *icr, double *dn, double h, double *A, double *b,
                              int nImpacts, int cr_stride, int nd_stride, double k, double distEps);

/* --- node state between iterations (v3Library/Geometry/Node.cs, ImplicitModel3.cs) ---
 * st: double SoA 12 x st_stride per node: u 0-2, v 3-5, a 6-8, du 9-11 (Node.ux.., vx.., ax..,
 * dux..).  dn: the kernels' node buffer (18 x nd_stride). */
/* Node.InferTentativeValues / InferTentativeUVA (Node.cs:79-97) for every node, stored the way
 * GPU_Functionality.TransferNodesToGPU packs them (GPU_Functionality.cs:307-340): un, vn, an,
 * x_current = x0 + un, F = 0 (x0 is read from dn). */
/* The element, CZ and contact passes of one iteration on `threads` OpenMP threads, with
 * compare-and-swap scatter additions (the CPU baseline; results equal the sequential passes up
 * to the order of the additions). */
void orc_assemble_par(const int *ie, int nElems, int el_stride, double *dcz, int *icz, int nCZs, int cz_stride,
                      const int *icr, int nImpacts, int cr_stride, double *dn, int nd_stride, double h, double k,
                      double distEps, double *A, long long nA, double *b, long long nb, int threads);

void orc_nodes_infer(const double *st, int st_stride, double *dn, int nd_stride, int n, double h,
                     double beta, double gamma);
/* ImplicitModel3._XtoDU (ImplicitModel3.cs:247-256): du += dx[3 altId + k] for altId >= 0. */
void orc_nodes_add_dx(double *st, int st_stride, const int *alt, int n, const double *dx);
/* Node.AcceptTentativeValues (Node.cs:62-77), with un as last inferred (dn): v' = (un - u)/h,
 * a = (v' - v)/h, v = v', u = un, du = 0. */
void orc_nodes_accept(double *st, int st_stride, const double *dn, int nd_stride, int n, double h);

#ifdef __cplusplus
}
#endif
#endif

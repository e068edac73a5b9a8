// icf_nodes.inc.h -- node state kept on the device between Newton iterations (included by
// icf_lib.hip).
//
// The reference keeps u, v, a, du per node on the host (v3Library/Geometry/Node.cs) and
// re-packs and uploads the whole node buffer every iteration
// (GPU_Functionality.TransferNodesToGPU, GPU_Functionality.cs:307-340).  These kernels keep
// that state in a device SoA buffer st (12 x st_stride: u 0-2, v 3-5, a 6-8, du 9-11) and
// write the kernels' node buffer dn in place.  One thread per node and component; the
// expressions are the reference's, evaluated without FMA contraction, so the results are
// the reference's doubles.
#pragma once

namespace icf {

// Node.InferTentativeValues / InferTentativeUVA (Node.cs:79-97), packed as TransferNodesToGPU
// does: un, vn, an, x_current = x0 + un, F = 0.
__global__ __launch_bounds__(256) void k_nodes_infer(const double *__restrict__ st, int ss, double *__restrict__ dn,
                                                     int ns, int n, double h, double beta, double gamma)
{
#pragma clang fp contract(off)
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3LL * n) return;
    const int k = (int)(t / n), i = (int)(t % n);
    const double u = st[i + (size_t)ss * k], v = st[i + (size_t)ss * (3 + k)];
    const double a = st[i + (size_t)ss * (6 + k)], du = st[i + (size_t)ss * (9 + k)];
    const double un = u + du;
    const double an = a * (1.0 - 1.0 / (2 * beta)) + du / (h * h * beta) + v * (-1.0 / (h * beta));
    const double vn = v + h * ((1.0 - gamma) * a + gamma * an);
    dn[i + (size_t)ns * (ND_UN + k)] = un;
    dn[i + (size_t)ns * (ND_VN + k)] = vn;
    dn[i + (size_t)ns * (ND_AN + k)] = an;
    dn[i + (size_t)ns * (ND_F + k)] = 0.0;
    dn[i + (size_t)ns * (ND_XC + k)] = dn[i + (size_t)ns * (ND_X0 + k)] + un;
}

// ImplicitModel3._XtoDU (ImplicitModel3.cs:247-256): du += dx[3 altId + k] where altId >= 0.
__global__ __launch_bounds__(256) void k_nodes_add_dx(double *__restrict__ st, int ss, const int *__restrict__ alt,
                                                      int n, const double *__restrict__ dx)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3LL * n) return;
    const int k = (int)(t / n), i = (int)(t % n);
    const int r = alt[i];
    if (r >= 0) st[i + (size_t)ss * (9 + k)] += dx[(size_t)r * 3 + k];
}

// Node.AcceptTentativeValues (Node.cs:62-77), un as last inferred.
__global__ __launch_bounds__(256) void k_nodes_accept(double *__restrict__ st, int ss, const double *__restrict__ dn,
                                                      int ns, int n, double h)
{
#pragma clang fp contract(off)
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3LL * n) return;
    const int k = (int)(t / n), i = (int)(t % n);
    const double un = dn[i + (size_t)ns * (ND_UN + k)];
    const double u = st[i + (size_t)ss * k], v = st[i + (size_t)ss * (3 + k)];
    const double vn = (un - u) / h;
    st[i + (size_t)ss * (6 + k)] = (vn - v) / h;
    st[i + (size_t)ss * (3 + k)] = vn;
    st[i + (size_t)ss * k] = un;
    st[i + (size_t)ss * (9 + k)] = 0.0;
}

}  // namespace icf

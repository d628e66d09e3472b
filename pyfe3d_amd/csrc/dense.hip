// Small FP64 dense kernels for the eigen-solvers' Krylov bases: a basis is
// m columns of length n, column-major with leading dimension ld.  Every
// reduction has a fixed order (per-column workgroup tree), so reruns give
// the same bits.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pf3d.h"

int pf3d_fail(int code, const char *fmt, ...);
int pf3d_hip_check(hipError_t e, const char *where);

namespace {

// c[j] = alpha * V[:, j] . x + beta * c[j]   (one workgroup per column)
__global__ void __launch_bounds__(256) gemv_t(int64_t n, int m, const double *__restrict__ V,
                                              int64_t ld, const double *__restrict__ x,
                                              double *__restrict__ c, double alpha, double beta)
{
    __shared__ double sh[256];
    const int j = blockIdx.x;
    const double *v = V + (int64_t)j * ld;
    double s = 0.;
    for (int64_t i = threadIdx.x; i < n; i += 256) s += v[i] * x[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) c[j] = alpha * sh[0] + (beta != 0. ? beta * c[j] : 0.);
}

// y = alpha * V c + beta * y
__global__ void gemv_n(int64_t n, int m, const double *__restrict__ V, int64_t ld,
                       const double *__restrict__ c, double *__restrict__ y, double alpha,
                       double beta)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.;
    for (int j = 0; j < m; j++) s += V[(int64_t)j * ld + i] * c[j];
    y[i] = alpha * s + (beta != 0. ? beta * y[i] : 0.);
}

// y = a * x + b * y
__global__ void axpby(int64_t n, double a, const double *__restrict__ x, double b,
                      double *__restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + (b != 0. ? b * y[i] : 0.);
}

// FP32 vectors (the multigrid preconditioner in single precision)
__global__ void axpby_f32(int64_t n, float a, const float *__restrict__ x, float b,
                          float *__restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + (b != 0.f ? b * y[i] : 0.f);
}

__global__ void widen(int64_t n, const float *__restrict__ a, double *__restrict__ b)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}

// dot product in two fixed-order stages: DOT_BLOCKS partial sums over a
// grid stride, then one workgroup adds the partials in index order
constexpr int DOT_BLOCKS = 1024;

__global__ void __launch_bounds__(256) dot_partial(int64_t n, const double *__restrict__ a,
                                                   const double *__restrict__ b,
                                                   double *__restrict__ part)
{
    __shared__ double sh[256];
    double s = 0.;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256)
        s += a[i] * b[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(256) dot_finish(int nparts, const double *__restrict__ part,
                                                  double *__restrict__ out)
{
    __shared__ double sh[256];
    double s = 0.;
    for (int k = threadIdx.x; k < nparts; k += 256) s += part[k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

}  // namespace

extern "C" {

int pf3d_axpby_f32(int64_t n, float a, const float *x, float b, float *y, void *stream)
{
    if (n < 0 || (n && (!x || !y))) return pf3d_fail(PF3D_E_ARG, "pf3d_axpby_f32: bad arguments");
    if (!n) return PF3D_OK;
    hipLaunchKernelGGL(axpby_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, n, a, x, b, y);
    return pf3d_hip_check(hipGetLastError(), "pf3d_axpby_f32");
}

int pf3d_to_f64(int64_t n, const float *a, double *b, void *stream)
{
    if (n < 0 || (n && (!a || !b))) return pf3d_fail(PF3D_E_ARG, "pf3d_to_f64: bad arguments");
    if (!n) return PF3D_OK;
    hipLaunchKernelGGL(widen, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, n, a, b);
    return pf3d_hip_check(hipGetLastError(), "pf3d_to_f64");
}

int pf3d_dot(int64_t n, const double *a, const double *b, double *out, double *work,
             void *stream)
{
    if (n < 0 || (n && (!a || !b)) || !out || !work)
        return pf3d_fail(PF3D_E_ARG, "pf3d_dot: bad arguments");
    int64_t nb = (n + 255) / 256;
    const int blocks = (int)(nb < DOT_BLOCKS ? (nb < 1 ? 1 : nb) : DOT_BLOCKS);
    hipLaunchKernelGGL(dot_partial, dim3(blocks), dim3(256), 0, (hipStream_t)stream, n, a, b,
                       work);
    hipLaunchKernelGGL(dot_finish, dim3(1), dim3(256), 0, (hipStream_t)stream, blocks, work,
                       out);
    return pf3d_hip_check(hipGetLastError(), "pf3d_dot");
}


int pf3d_gemv_t(int64_t n, int m, const double *V, int64_t ld, const double *x, double *c,
                double alpha, double beta, void *stream)
{
    if (n < 0 || m < 0 || ld < n || (m && (!V || !x || !c)))
        return pf3d_fail(PF3D_E_ARG, "pf3d_gemv_t: bad arguments");
    if (!m) return PF3D_OK;
    hipLaunchKernelGGL(gemv_t, dim3(m), dim3(256), 0, (hipStream_t)stream, n, m, V, ld, x, c,
                       alpha, beta);
    return pf3d_hip_check(hipGetLastError(), "pf3d_gemv_t");
}

int pf3d_gemv_n(int64_t n, int m, const double *V, int64_t ld, const double *c, double *y,
                double alpha, double beta, void *stream)
{
    if (n < 0 || m < 0 || ld < n || (n && (!V || !c || !y)))
        return pf3d_fail(PF3D_E_ARG, "pf3d_gemv_n: bad arguments");
    if (!n) return PF3D_OK;
    hipLaunchKernelGGL(gemv_n, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, n, m, V, ld, c, y, alpha, beta);
    return pf3d_hip_check(hipGetLastError(), "pf3d_gemv_n");
}

int pf3d_axpby(int64_t n, double a, const double *x, double b, double *y, void *stream)
{
    if (n < 0 || (n && (!x || !y))) return pf3d_fail(PF3D_E_ARG, "pf3d_axpby: bad arguments");
    if (!n) return PF3D_OK;
    hipLaunchKernelGGL(axpby, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, n, a, x, b, y);
    return pf3d_hip_check(hipGetLastError(), "pf3d_axpby");
}

}  // extern "C"

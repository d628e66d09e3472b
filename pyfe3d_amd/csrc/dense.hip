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

// Conjugate-gradient updates with the scalars kept on the device (slots:
// s[0] = r.z, s[1] = p.q, s[3] = new r.z, s[4] = beta, s[6] = r.z_old), so
// a whole iteration runs without a host round trip and can be captured
__global__ void cg_update_xr(int64_t n, const double *__restrict__ s,
                             const double *__restrict__ p, const double *__restrict__ q,
                             double *__restrict__ x, double *__restrict__ r)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double alpha = s[1] != 0. ? s[0] / s[1] : 0.;
    if (i < n) {
        x[i] += alpha * p[i];
        r[i] -= alpha * q[i];
    }
}

__global__ void cg_beta(double *s, int flexible)
{
    if (threadIdx.x == 0) {
        const double num = s[3] - (flexible ? s[6] : 0.);
        s[4] = s[0] != 0. ? num / s[0] : 0.;
        s[0] = s[3];
    }
}

__global__ void cg_update_p(int64_t n, const double *__restrict__ s, const double *__restrict__ z,
                            double *__restrict__ p)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = z[i] + s[4] * p[i];
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

// C = A B^T for p vectors A[i][0:n] (row stride lda) and q vectors
// B[j][0:n] (row stride ldb), p, q <= 32: the Gram matrices of a block
// eigensolver.  Stage 1: block (g, ti, tj) accumulates a 4 x 4 tile over
// the g-th stride of t in registers and reduces it over its threads;
// stage 2 sums the BD_G partials of each entry in index order.
constexpr int BD_G = 256;

__global__ void __launch_bounds__(256) block_dot_partial(int64_t n, int p, const double *A,
                                                         int64_t lda, int q, const double *B,
                                                         int64_t ldb, double *part)
{
    __shared__ double sh[16][256];
    const int g = blockIdx.x, i0 = 4 * blockIdx.y, j0 = 4 * blockIdx.z;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = 0.;
    for (int64_t t = (int64_t)g * 256 + threadIdx.x; t < n; t += (int64_t)BD_G * 256) {
        double x[4], y[4];
#pragma unroll
        for (int a = 0; a < 4; a++) x[a] = i0 + a < p ? A[(int64_t)(i0 + a) * lda + t] : 0.;
#pragma unroll
        for (int b = 0; b < 4; b++) y[b] = j0 + b < q ? B[(int64_t)(j0 + b) * ldb + t] : 0.;
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) acc[a][b] += x[a] * y[b];
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) sh[4 * a + b][threadIdx.x] = acc[a][b];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
#pragma unroll
            for (int e = 0; e < 16; e++) sh[e][threadIdx.x] += sh[e][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 16) {
        const int a = threadIdx.x / 4, b = threadIdx.x % 4;
        if (i0 + a < p && j0 + b < q)
            part[((int64_t)g * p + i0 + a) * q + j0 + b] = sh[threadIdx.x][0];
    }
}

__global__ void __launch_bounds__(1024) block_dot_finish(int p, int q, const double *part,
                                                         double *C)
{
    const int e = threadIdx.x;
    if (e >= p * q) return;
    double s = 0.;
    for (int g = 0; g < BD_G; g++) s += part[(int64_t)g * p * q + e];
    C[e] = s;
}

// Y[j][t] = sum_i C[i][j] S[i][t]: k combinations of m vectors, C (m x k,
// row-major) staged in LDS, one thread per t (S read once)
__global__ void __launch_bounds__(256) block_comb(int64_t n, int m, const double *S, int64_t lds,
                                                  int k, const double *C, double *Y, int64_t ldy)
{
    __shared__ double cs[32 * 32];
    for (int e = threadIdx.x; e < m * k; e += 256) cs[e] = C[e];
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    double s[32];
#pragma unroll
    for (int i = 0; i < 32; i++) s[i] = i < m ? S[(int64_t)i * lds + t] : 0.;
    for (int j = 0; j < k; j++) {
        double y = 0.;
#pragma unroll
        for (int i = 0; i < 32; i++)
            if (i < m) y += cs[i * k + j] * s[i];
        Y[(int64_t)j * ldy + t] = y;
    }
}

}  // namespace

extern "C" {

int pf3d_block_comb(int64_t n, int m, const double *S, int64_t lds, int k, const double *C,
                    double *Y, int64_t ldy, void *stream)
{
    if (n < 0 || m < 1 || k < 1 || m > 32 || k > 32 || lds < n || ldy < n || !S || !C || !Y)
        return pf3d_fail(PF3D_E_ARG, "pf3d_block_comb: bad arguments");
    if (!n) return PF3D_OK;
    hipLaunchKernelGGL(block_comb, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, n, m, S, lds, k, C, Y, ldy);
    return pf3d_hip_check(hipGetLastError(), "pf3d_block_comb");
}


int pf3d_block_dot(int64_t n, int p, const double *A, int64_t lda, int q, const double *B,
                   int64_t ldb, double *C, double *work, void *stream)
{
    if (n < 0 || p < 1 || q < 1 || p > 32 || q > 32 || lda < n || ldb < n || !A || !B || !C ||
        !work)
        return pf3d_fail(PF3D_E_ARG, "pf3d_block_dot: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(block_dot_partial, dim3(BD_G, (p + 3) / 4, (q + 3) / 4), dim3(256), 0, s,
                       n, p, A, lda, q, B, ldb, work);
    hipLaunchKernelGGL(block_dot_finish, dim3(1), dim3(1024), 0, s, p, q, work, C);
    return pf3d_hip_check(hipGetLastError(), "pf3d_block_dot");
}


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

int pf3d_cg_update_xr(int64_t n, const double *s, const double *p, const double *q, double *x,
                      double *r, void *stream)
{
    if (n <= 0 || !s || !p || !q || !x || !r)
        return pf3d_fail(PF3D_E_ARG, "pf3d_cg_update_xr: bad arguments");
    hipLaunchKernelGGL(cg_update_xr, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, n, s, p, q, x, r);
    return pf3d_hip_check(hipGetLastError(), "pf3d_cg_update_xr");
}

int pf3d_cg_beta(double *s, int flexible, void *stream)
{
    if (!s) return pf3d_fail(PF3D_E_ARG, "pf3d_cg_beta: bad arguments");
    hipLaunchKernelGGL(cg_beta, dim3(1), dim3(64), 0, (hipStream_t)stream, s, flexible);
    return pf3d_hip_check(hipGetLastError(), "pf3d_cg_beta");
}

int pf3d_cg_update_p(int64_t n, const double *s, const double *z, double *p, void *stream)
{
    if (n <= 0 || !s || !z || !p) return pf3d_fail(PF3D_E_ARG, "pf3d_cg_update_p: bad arguments");
    hipLaunchKernelGGL(cg_update_p, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, n, s, z, p);
    return pf3d_hip_check(hipGetLastError(), "pf3d_cg_update_p");
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

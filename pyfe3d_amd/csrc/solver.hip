// 6x6 block-sparse (BSR) operator, block-Jacobi preconditioner and the
// preconditioned conjugate gradient solver (FP64, and FP32 inside an FP64
// iterative-refinement loop).  These replace the reference caller's
// KC0[bu, :][:, bu] slicing and scipy spsolve/cg
// (tests/test_quad4_static_point_load.py:84-102, tests/test_tria3r_static_point_load.py:131).
//
// Dirichlet conditions are applied to the assembled matrix in place
// (K <- P K P + (I - P), P = diag(free)), which keeps the 6x6 block
// structure; the free-DOF solution is that of the sliced system and the
// fixed DOFs come out as zero.
//
// All reductions are two-stage with a fixed tree (per-workgroup partials,
// then one workgroup), so repeated solves are bit-identical.
#include <math.h>
#include <stdint.h>

#include <mutex>

#include "../../include/pf3d.h"
#include "pf3d_device.hpp"

int pf3d_fail(int code, const char *fmt, ...);
int pf3d_hip_check(hipError_t e, const char *where);

namespace {

constexpr int VB = 256;        // threads per block of the vector kernels
constexpr int RED_BLOCKS = 1024;  // max partials of a reduction

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// ---------------------------------------------------------------- BSR SpMV
// One wavefront per RPW block rows: lane (j, i) multiplies row i of the
// j-th block of each of its rows (10 blocks per pass) with its 6-vector,
// then the wave sums the partial rows through LDS.  MT is the storage type
// of the matrix values (FP64, or FP32 to halve the dominant traffic), VT
// that of the vectors; products and sums are FP64.  A wave takes four rows
// in lockstep (one block of each per pass) to keep four loads per lane in
// flight: measured on the 6M-DOF plate's matrix, 556 -> 503 us (FP64) and
// 335 -> 301 us (FP32) against one row per wave, same bits.
#ifndef PF3D_SPMV_RPW64
#define PF3D_SPMV_RPW64 4
#endif
#ifndef PF3D_SPMV_RPW32
#define PF3D_SPMV_RPW32 4
#endif
template <typename MT>
constexpr int spmv_rpw() { return sizeof(MT) == 8 ? PF3D_SPMV_RPW64 : PF3D_SPMV_RPW32; }

template <typename MT, typename VT>
__global__ void __launch_bounds__(256) bsr_spmv(int64_t nb, const int32_t *__restrict__ rp,
                                                const int32_t *__restrict__ ci,
                                                const MT *__restrict__ val,
                                                const VT *__restrict__ x, VT *__restrict__ y)
{
    constexpr int RPW = spmv_rpw<MT>();
    __shared__ double part[4][RPW][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t row0 = ((int64_t)blockIdx.x * 4 + wave) * RPW;
    const int j = lane / 6, i = lane - 6 * (lane / 6);
    double acc[RPW];
    int b[RPW], b1[RPW];
#pragma unroll
    for (int q = 0; q < RPW; q++) {
        acc[q] = 0.;
        const int64_t row = row0 + q;
        const bool ok = row < nb && j < 10;
        b[q] = ok ? rp[row] + j : 0;
        b1[q] = ok ? rp[row + 1] : 0;
    }
    // the rows' blocks in lockstep: one load of each row per pass
    bool more = true;
    while (more) {
        more = false;
#pragma unroll
        for (int q = 0; q < RPW; q++) {
            if (b[q] < b1[q]) {
                const MT *v = val + (int64_t)b[q] * 36 + i * 6;
                const VT *xv = x + (int64_t)ci[b[q]] * 6;
#pragma unroll
                for (int k = 0; k < 6; k++) acc[q] += (double)v[k] * (double)xv[k];
                b[q] += 10;
                more |= b[q] < b1[q];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RPW; q++) part[wave][q][lane] = acc[q];
    __syncthreads();
    if (lane < 6 * RPW) {
        const int q = lane / 6, l = lane - 6 * q;
        const int64_t row = row0 + q;
        if (row < nb) {
            double s = 0.;
#pragma unroll
            for (int jj = 0; jj < 10; jj++) s += part[wave][q][l + 6 * jj];
            y[row * 6 + l] = (VT)s;
        }
    }
}

// ------------------------------------------ fused multigrid vector steps
// One pass over the block rows for the Chebyshev smoother, the residual
// and the prolongation of the V-cycle (same lane mapping as bsr_spmv):
//   q_i = (A xin)_i                      (xin may be NULL: q = 0)
//   r_i = sa * (b ? b_i : r_i) + sb * q_i
//   dnew_i = c1 * (Dinv r)_i + c2 * dold_i   (when dinv is given)
//   xacc_i += dold_i                     (when xacc is given)
// r, dnew and xacc are written by their own row only; xin must not alias
// any output and dold must not alias dnew.
template <typename MT, typename VT>
__global__ void __launch_bounds__(256)
    bsr_fused(int64_t nb, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
              const MT *__restrict__ val, const VT *__restrict__ xin, const VT *__restrict__ b,
              VT *r, const MT *__restrict__ dinv, const VT *__restrict__ dold,
              VT *__restrict__ dnew, VT *__restrict__ xacc, double sa, double sb, double c1,
              double c2)
{
    constexpr int RPW = spmv_rpw<MT>();
    __shared__ double part[4][RPW][64];
    __shared__ double rrow[4][RPW][6];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t row0 = ((int64_t)blockIdx.x * 4 + wave) * RPW;
    const int j = lane / 6, i = lane - 6 * (lane / 6);
    double acc[RPW];
    int bb[RPW], b1[RPW];
#pragma unroll
    for (int q = 0; q < RPW; q++) {
        acc[q] = 0.;
        const int64_t row = row0 + q;
        const bool ok = xin && row < nb && j < 10;
        bb[q] = ok ? rp[row] + j : 0;
        b1[q] = ok ? rp[row + 1] : 0;
    }
    bool more = true;
    while (more) {
        more = false;
#pragma unroll
        for (int q = 0; q < RPW; q++) {
            if (bb[q] < b1[q]) {
                const MT *v = val + (int64_t)bb[q] * 36 + i * 6;
                const VT *xv = xin + (int64_t)ci[bb[q]] * 6;
#pragma unroll
                for (int k = 0; k < 6; k++) acc[q] += (double)v[k] * (double)xv[k];
                bb[q] += 10;
                more |= bb[q] < b1[q];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RPW; q++) part[wave][q][lane] = acc[q];
    __syncthreads();
    const int q = lane / 6, l = lane - 6 * (lane / 6);
    const int64_t row = row0 + q;
    const bool own = lane < 6 * RPW && row < nb;
    if (own) {
        double s = 0.;
#pragma unroll
        for (int jj = 0; jj < 10; jj++) s += part[wave][q][l + 6 * jj];
        const int64_t k = row * 6 + l;
        const double rv = sa * (double)(b ? b[k] : r[k]) + sb * s;
        r[k] = (VT)rv;
        rrow[wave][q][l] = (double)(VT)rv;
    }
    if (!dinv && !xacc) return;
    __syncthreads();
    if (own) {
        const int64_t k = row * 6 + l;
        const double dv = dold ? (double)dold[k] : 0.;
        if (dinv) {
            const MT *D = dinv + row * 36 + l * 6;
            double t = 0.;
#pragma unroll
            for (int m = 0; m < 6; m++) t += (double)D[m] * rrow[wave][q][m];
            dnew[k] = (VT)(c1 * t + c2 * dv);
        }
        if (xacc) xacc[k] = (VT)((double)xacc[k] + dv);
    }
}

template <typename MT, typename VT>
int fused_launch(int64_t nb, const int32_t *rp, const int32_t *ci, const MT *val, const VT *xin,
                 const VT *b, VT *r, const MT *dinv, const VT *dold, VT *dnew, VT *xacc,
                 double sa, double sb, double c1, double c2, void *stream, const char *name)
{
    if (nb < 0 || (nb && (!r || (xin && (!rp || !ci || !val)) || (dinv && !dnew) ||
                          (xacc && !dold) || (dold && dold == dnew) ||
                          (xin && (xin == r || xin == dnew || xin == xacc)))))
        return pf3d_fail(PF3D_E_ARG, "%s: bad arguments", name);
    if (!nb) return PF3D_OK;
    hipLaunchKernelGGL((bsr_fused<MT, VT>), dim3(nblk(nb, 4 * spmv_rpw<MT>())), dim3(256), 0,
                       (hipStream_t)stream, nb, rp, ci, val, xin, b, r, dinv, dold, dnew, xacc,
                       sa, sb, c1, c2);
    return pf3d_hip_check(hipGetLastError(), name);
}

// Multi-vector form of bsr_fused for nv <= MAXV vectors stored one after
// another (vector v of each operand at v * ld): each block of A is read
// once for all of them.  One block row per wave; sa == 0 reads neither b
// nor r.
constexpr int MAXV = 8;

template <typename MT, typename VT>
__global__ void __launch_bounds__(256)
    bsr_fused_mv(int64_t nb, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                 const MT *__restrict__ val, int nv, int64_t ldx, const VT *__restrict__ xin,
                 int64_t ld, const VT *__restrict__ b, VT *r, const MT *__restrict__ dinv,
                 const VT *__restrict__ dold, VT *__restrict__ dnew, VT *__restrict__ xacc,
                 double sa, double sb, double c1, double c2)
{
    __shared__ double part[4][MAXV][64];
    __shared__ double rrow[4][MAXV][6];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + wave;
    const int j = lane / 6, i = lane - 6 * (lane / 6);
    double acc[MAXV];
#pragma unroll
    for (int v = 0; v < MAXV; v++) acc[v] = 0.;
    if (xin && row < nb && j < 10) {
        for (int bb = rp[row] + j; bb < rp[row + 1]; bb += 10) {
            const MT *m = val + (int64_t)bb * 36 + i * 6;
            double mv[6];
#pragma unroll
            for (int k = 0; k < 6; k++) mv[k] = (double)m[k];
            const int64_t c6 = (int64_t)ci[bb] * 6;
#pragma unroll
            for (int v = 0; v < MAXV; v++) {
                if (v < nv) {
                    const VT *xv = xin + v * ldx + c6;
#pragma unroll
                    for (int k = 0; k < 6; k++) acc[v] += mv[k] * (double)xv[k];
                }
            }
        }
    }
#pragma unroll
    for (int v = 0; v < MAXV; v++) part[wave][v][lane] = acc[v];
    __syncthreads();
    // lanes (v, l): vector v, component l of the row
    const int v = lane / 6, l = lane - 6 * (lane / 6);
    const bool own = v < nv && row < nb;
    if (own) {
        double s = 0.;
#pragma unroll
        for (int jj = 0; jj < 10; jj++) s += part[wave][v][l + 6 * jj];
        const int64_t k = v * ld + row * 6 + l;
        const double base = sa == 0. ? 0. : sa * (double)(b ? b[k] : r[k]);
        const double rv = base + sb * s;
        r[k] = (VT)rv;
        rrow[wave][v][l] = (double)(VT)rv;
    }
    if (!dinv && !xacc) return;
    __syncthreads();
    if (own) {
        const int64_t k = v * ld + row * 6 + l;
        const double dv = dold ? (double)dold[k] : 0.;
        if (dinv) {
            const MT *D = dinv + row * 36 + l * 6;
            double t = 0.;
#pragma unroll
            for (int m = 0; m < 6; m++) t += (double)D[m] * rrow[wave][v][m];
            dnew[k] = (VT)(c1 * t + c2 * dv);
        }
        if (xacc) xacc[k] = (VT)((double)xacc[k] + dv);
    }
}

template <typename MT, typename VT>
int fused_mv_launch(int64_t nb, const int32_t *rp, const int32_t *ci, const MT *val, int nv,
                    int64_t ldx, const VT *xin, int64_t ld, const VT *b, VT *r, const MT *dinv,
                    const VT *dold, VT *dnew, VT *xacc, double sa, double sb, double c1,
                    double c2, void *stream, const char *name)
{
    if (nb < 0 || nv < 1 || nv > MAXV || ld < 6 * nb || (xin && ldx < 6) ||
        (nb && (!r || (xin && (!rp || !ci || !val)) || (dinv && !dnew) || (xacc && !dold) ||
                (dold && dold == dnew) ||
                (xin && (xin == r || xin == dnew || xin == xacc)))))
        return pf3d_fail(PF3D_E_ARG, "%s: bad arguments", name);
    if (!nb) return PF3D_OK;
    hipLaunchKernelGGL((bsr_fused_mv<MT, VT>), dim3(nblk(nb, 4)), dim3(256), 0,
                       (hipStream_t)stream, nb, rp, ci, val, nv, ldx, xin, ld, b, r, dinv, dold,
                       dnew, xacc, sa, sb, c1, c2);
    return pf3d_hip_check(hipGetLastError(), name);
}

// ------------------------------------------------- compensated products
// Row sums of shell stiffness matrices cancel by many digits (stiff
// membrane and shear terms against soft bending and drilling terms), so
// b - A x in plain FP64 carries an error of ~eps |A| |x| that can exceed
// 1e-8 |b|.  These kernels accumulate with error-free transformations
// (product split by FMA, TwoSum for the running sum, compensation terms
// summed apart): the result is accurate to ~eps of itself.
__device__ __forceinline__ void two_sum(double &s, double &c, double p)
{
#pragma clang fp contract(off)
    const double t = s + p;
    const double bp = t - s;
    c += (s - (t - bp)) + (p - bp);
    s = t;
}

// r = b - A x (b may be null: r = -A x; sign = +1 gives r = A x when b is null)
__global__ void __launch_bounds__(256) bsr_residual_dd(int64_t nb, const int32_t *__restrict__ rp,
                                                       const int32_t *__restrict__ ci,
                                                       const double *__restrict__ val,
                                                       const double *__restrict__ x,
                                                       const double *__restrict__ b, double sign,
                                                       double *__restrict__ y)
{
#pragma clang fp contract(off)
    __shared__ double ps[4][64], pc[4][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + wave;
    const int j = lane / 6, i = lane - 6 * (lane / 6);
    double s = 0., c = 0.;
    if (row < nb && j < 10) {
        const int b0 = rp[row], b1 = rp[row + 1];
        for (int k = b0 + j; k < b1; k += 10) {
            const double *v = val + (int64_t)k * 36 + i * 6;
            const double *xv = x + (int64_t)ci[k] * 6;
#pragma unroll
            for (int q = 0; q < 6; q++) {
                const double p = sign * v[q] * xv[q];
                const double e = fma(sign * v[q], xv[q], -p);  // exact: sign is +-1
                c += e;
                two_sum(s, c, p);
            }
        }
    }
    ps[wave][lane] = s;
    pc[wave][lane] = c;
    __syncthreads();
    if (row < nb && lane < 6) {
        double S = b ? b[row * 6 + lane] : 0., C = 0.;
        for (int jj = 0; jj < 10; jj++) {
            two_sum(S, C, ps[wave][lane + 6 * jj]);
            C += pc[wave][lane + 6 * jj];
        }
        y[row * 6 + lane] = S + C;
    }
}

// y[v] = A x[v] for nv <= MAXV vectors (ld apart) with the compensated
// accumulation of bsr_residual_dd: each block of A read once for all of
// them; per vector the same operations in the same order (same bits).
__global__ void __launch_bounds__(256) bsr_product_dd_mv(int64_t nb,
                                                         const int32_t *__restrict__ rp,
                                                         const int32_t *__restrict__ ci,
                                                         const double *__restrict__ val, int nv,
                                                         int64_t ld, const double *__restrict__ x,
                                                         double *__restrict__ y)
{
#pragma clang fp contract(off)
    __shared__ double ps[4][MAXV][64], pc[4][MAXV][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + wave;
    const int j = lane / 6, i = lane - 6 * (lane / 6);
    double s[MAXV], c[MAXV];
#pragma unroll
    for (int v = 0; v < MAXV; v++) { s[v] = 0.; c[v] = 0.; }
    if (row < nb && j < 10) {
        const int b0 = rp[row], b1 = rp[row + 1];
        for (int k = b0 + j; k < b1; k += 10) {
            const double *m = val + (int64_t)k * 36 + i * 6;
            double mv[6];
#pragma unroll
            for (int q = 0; q < 6; q++) mv[q] = m[q];
            const int64_t c6 = (int64_t)ci[k] * 6;
#pragma unroll
            for (int v = 0; v < MAXV; v++) {
                if (v < nv) {
                    const double *xv = x + v * ld + c6;
#pragma unroll
                    for (int q = 0; q < 6; q++) {
                        const double p = mv[q] * xv[q];
                        const double e = fma(mv[q], xv[q], -p);
                        c[v] += e;
                        two_sum(s[v], c[v], p);
                    }
                }
            }
        }
    }
#pragma unroll
    for (int v = 0; v < MAXV; v++) {
        ps[wave][v][lane] = s[v];
        pc[wave][v][lane] = c[v];
    }
    __syncthreads();
    const int v = lane / 6, l = lane - 6 * (lane / 6);
    if (row < nb && v < nv) {
        double S = 0., C = 0.;
        for (int jj = 0; jj < 10; jj++) {
            two_sum(S, C, ps[wave][v][l + 6 * jj]);
            C += pc[wave][v][l + 6 * jj];
        }
        y[v * ld + row * 6 + l] = S + C;
    }
}

// ------------------------------------------------------ boundary conditions
// K <- P K P + (I - P) with P = diag(1 - fixed)
__global__ void bsr_bc(int64_t nb, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                       double *__restrict__ val, const uint8_t *__restrict__ fixed, double diag)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t row = t / 36;
    if (row >= nb) return;
    const int e = (int)(t - row * 36), i = e / 6, k = e % 6;
    const bool fr = fixed[row * 6 + i];
    for (int b = rp[row]; b < rp[row + 1]; b++) {
        const int64_t col = ci[b];
        const bool fc = fixed[col * 6 + k];
        double *v = val + (int64_t)b * 36 + e;
        if (fr || fc) *v = (col == row && i == k) ? diag : 0.;
    }
}

// ------------------------------------------------------------ block Jacobi
// inverse of each 6x6 diagonal block, Gauss-Jordan with partial pivoting
__global__ void bjacobi_setup(int64_t nb, const int32_t *__restrict__ rp,
                              const int32_t *__restrict__ ci, const double *__restrict__ val,
                              double *__restrict__ dinv, float *__restrict__ dinv32,
                              int *__restrict__ bad)
{
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nb) return;
    double a[6][12];
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 12; j++) a[i][j] = (j - 6 == i) ? 1. : 0.;
    for (int b = rp[row]; b < rp[row + 1]; b++)
        if (ci[b] == row)
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++) a[i][j] = val[(int64_t)b * 36 + i * 6 + j];
    for (int c = 0; c < 6; c++) {
        int p = c;
        double m = fabs(a[c][c]);
        for (int i = c + 1; i < 6; i++)
            if (fabs(a[i][c]) > m) { m = fabs(a[i][c]); p = i; }
        if (!(m > 0.)) { atomicAdd(bad, 1); m = 1.; a[p][c] = 1.; }
        if (p != c)
            for (int j = 0; j < 12; j++) { double t = a[c][j]; a[c][j] = a[p][j]; a[p][j] = t; }
        const double inv = 1. / a[c][c];
        for (int j = 0; j < 12; j++) a[c][j] *= inv;
        for (int i = 0; i < 6; i++)
            if (i != c) {
                const double f = a[i][c];
                for (int j = 0; j < 12; j++) a[i][j] -= f * a[c][j];
            }
    }
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            dinv[row * 36 + i * 6 + j] = a[i][j + 6];
            if (dinv32) dinv32[row * 36 + i * 6 + j] = (float)a[i][j + 6];
        }
}

// ---------------------------------------------------------- reductions
__device__ __forceinline__ double block_sum(double v, double *sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] += sh[t + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// state of one solve, in device memory
struct PcgState {
    double rz, rz_old, pq, alpha, beta, rr, bb, pad;
};

// partial[k][blk] for up to 2 dot products, grid-strided over n
template <typename T>
__global__ void __launch_bounds__(VB) dot2(int64_t n, const T *__restrict__ a,
                                           const T *__restrict__ b, const T *__restrict__ c,
                                           const T *__restrict__ d, double *__restrict__ partial)
{
    __shared__ double sh[VB];
    double s0 = 0., s1 = 0.;
    for (int64_t t = (int64_t)blockIdx.x * VB + threadIdx.x; t < n; t += (int64_t)gridDim.x * VB) {
        s0 += (double)a[t] * (double)b[t];
        if (c) s1 += (double)c[t] * (double)d[t];
    }
    s0 = block_sum(s0, sh);
    s1 = block_sum(s1, sh);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s0;
        partial[RED_BLOCKS + blockIdx.x] = s1;
    }
}

// single-workgroup finisher: mode 0: pq -> alpha; mode 1: rz, rr -> beta;
// mode 2: init (rz, rr) ; mode 3: bb ; mode 4: rr from the first sum
__global__ void __launch_bounds__(VB) finish(int nparts, const double *__restrict__ partial,
                                             PcgState *st, int mode)
{
    __shared__ double sh[VB];
    double s0 = 0., s1 = 0.;
    for (int k = threadIdx.x; k < nparts; k += VB) {
        s0 += partial[k];
        s1 += partial[RED_BLOCKS + k];
    }
    s0 = block_sum(s0, sh);
    s1 = block_sum(s1, sh);
    if (threadIdx.x) return;
    if (mode == 0) {
        st->pq = s0;
        st->alpha = st->rz / s0;
    } else if (mode == 1) {
        st->rz_old = st->rz;
        st->rz = s0;
        st->rr = s1;
        st->beta = s0 / st->rz_old;
    } else if (mode == 2) {
        st->rz = s0;
        st->rr = s1;
    } else if (mode == 3) {
        st->bb = s0;
    } else {
        st->rr = s0;
    }
}

// x += alpha p ; r -= alpha q ; z = Dinv r ; partial (r.z, r.r)
template <typename MT, typename T>
__global__ void __launch_bounds__(VB) update_xrz(int64_t nb, const PcgState *__restrict__ st,
                                                 T *__restrict__ x, T *__restrict__ r,
                                                 const T *__restrict__ p, const T *__restrict__ q,
                                                 const MT *__restrict__ dinv, T *__restrict__ z,
                                                 double *__restrict__ partial)
{
    __shared__ double sh[VB];
    const double alpha = st->alpha;
    double s0 = 0., s1 = 0.;
    for (int64_t row = (int64_t)blockIdx.x * VB + threadIdx.x; row < nb;
         row += (int64_t)gridDim.x * VB) {
        T rv[6];
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const int64_t t = row * 6 + k;
            x[t] = (T)((double)x[t] + alpha * (double)p[t]);
            rv[k] = (T)((double)r[t] - alpha * (double)q[t]);
            r[t] = rv[k];
        }
#pragma unroll
        for (int i = 0; i < 6; i++) {
            double s = 0;
#pragma unroll
            for (int k = 0; k < 6; k++) s += (double)dinv[row * 36 + i * 6 + k] * (double)rv[k];
            z[row * 6 + i] = (T)s;
            s0 += s * (double)rv[i];
            s1 += (double)rv[i] * (double)rv[i];
        }
    }
    s0 = block_sum(s0, sh);
    s1 = block_sum(s1, sh);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s0;
        partial[RED_BLOCKS + blockIdx.x] = s1;
    }
}

__global__ void set_alpha(PcgState *st, double alpha) { st->alpha = alpha; }

template <typename T>
__global__ void update_p(int64_t n, const PcgState *__restrict__ st, const T *__restrict__ z,
                         T *__restrict__ p)
{
    const int64_t t = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (t < n) p[t] = (T)((double)z[t] + st->beta * (double)p[t]);
}

// r = b - y (FP64), optionally also to a second precision
template <typename T>
__global__ void residual(int64_t n, const double *__restrict__ b, const double *__restrict__ y,
                         double *__restrict__ r, T *__restrict__ r2)
{
    const int64_t t = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (t >= n) return;
    const double v = b[t] - y[t];
    r[t] = v;
    if (r2) r2[t] = (T)v;
}

template <typename T>
__global__ void copy_cast(int64_t n, const double *__restrict__ a, T *__restrict__ b)
{
    const int64_t t = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (t < n) b[t] = (T)a[t];
}

__global__ void add_to(int64_t n, const double *__restrict__ d, double *__restrict__ x)
{
    const int64_t t = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (t < n) x[t] += d[t];
}

template <typename T>
__global__ void fill0(int64_t n, T *__restrict__ a)
{
    const int64_t t = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (t < n) a[t] = (T)0;
}

inline unsigned red_blocks(int64_t n)
{
    int64_t b = (n + VB - 1) / VB;
    return (unsigned)(b < RED_BLOCKS ? (b < 1 ? 1 : b) : RED_BLOCKS);
}

// ------------------------------------------------------------- PCG driver
struct Mat {
    int64_t nb;
    const int32_t *rp, *ci;
    const void *val, *dinv;
};

template <typename MT, typename VT>
hipError_t spmv(const Mat &A, const VT *x, VT *y, hipStream_t s)
{
    hipLaunchKernelGGL((bsr_spmv<MT, VT>), dim3(nblk(A.nb, 4 * spmv_rpw<MT>())), dim3(256), 0, s,
                       A.nb, A.rp, A.ci, (const MT *)A.val, x, y);
    return hipGetLastError();
}

// workspace: 4 vectors of T (r, z, p, q), partials, state
size_t pcg_bytes(int64_t n, size_t tsize)
{
    return ((4 * n * tsize + 255) / 256) * 256 + 2 * RED_BLOCKS * sizeof(double) + 256;
}

// Solve A x = b to ||r|| <= rtol ||b|| (recursive residual), starting from
// the true residual of x (x0 != 0) or from x = 0.  Matrix and block-Jacobi
// inverse are stored as MT, vectors as VT.
template <typename MT, typename VT>
int pcg_run(const Mat &A, const VT *b, VT *x, double rtol, int maxiter, int x0, void *work,
            hipStream_t s, int64_t *iters, double *relres, int check_every)
{
    const int64_t n = A.nb * 6;
    char *w = (char *)work;
    VT *r = (VT *)w, *z = r + n, *p = z + n, *q = p + n;
    double *partial = (double *)(w + ((4 * n * sizeof(VT) + 255) / 256) * 256);
    PcgState *st = (PcgState *)(partial + 2 * RED_BLOCKS);
    const unsigned nv = nblk(n, VB), nr = red_blocks(A.nb), nd = red_blocks(n);
    hipError_t e;
    hipLaunchKernelGGL(dot2<VT>, dim3(nd), dim3(VB), 0, s, n, b, b, (const VT *)nullptr,
                       (const VT *)nullptr, partial);
    hipLaunchKernelGGL(finish, dim3(1), dim3(VB), 0, s, (int)nd, partial, st, 3);
    if (!x0) hipLaunchKernelGGL(fill0<VT>, dim3(nv), dim3(VB), 0, s, n, x);
    if ((e = spmv<MT, VT>(A, x, q, s)) != hipSuccess) return pf3d_hip_check(e, "pcg spmv");
    // r = b - A x ; z = M r ; (r.z, r.r) through the update kernel with p = 0
    hipLaunchKernelGGL(fill0<VT>, dim3(nv), dim3(VB), 0, s, n, p);
    hipMemcpyAsync(r, b, sizeof(VT) * n, hipMemcpyDeviceToDevice, s);
    hipLaunchKernelGGL(set_alpha, dim3(1), dim3(1), 0, s, st, 1.);
    hipLaunchKernelGGL((update_xrz<MT, VT>), dim3(nr), dim3(VB), 0, s, A.nb, st, x, r, p, q,
                       (const MT *)A.dinv, z, partial);
    hipLaunchKernelGGL(finish, dim3(1), dim3(VB), 0, s, (int)nr, partial, st, 2);
    hipMemcpyAsync(p, z, sizeof(VT) * n, hipMemcpyDeviceToDevice, s);
    PcgState h;
    e = hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return pf3d_hip_check(e, "pcg init");
    const double bb = h.bb > 0 ? h.bb : 1.;
    double rel = sqrt(h.rr / bb);
    int it = 0;
    while (rel > rtol && it < maxiter) {
        for (int k = 0; k < check_every && it < maxiter; k++, it++) {
            spmv<MT, VT>(A, p, q, s);
            hipLaunchKernelGGL(dot2<VT>, dim3(nd), dim3(VB), 0, s, n, p, q, (const VT *)nullptr,
                               (const VT *)nullptr, partial);
            hipLaunchKernelGGL(finish, dim3(1), dim3(VB), 0, s, (int)nd, partial, st, 0);
            hipLaunchKernelGGL((update_xrz<MT, VT>), dim3(nr), dim3(VB), 0, s, A.nb, st, x, r,
                               p, q, (const MT *)A.dinv, z, partial);
            hipLaunchKernelGGL(finish, dim3(1), dim3(VB), 0, s, (int)nr, partial, st, 1);
            hipLaunchKernelGGL(update_p<VT>, dim3(nv), dim3(VB), 0, s, n, st, z, p);
        }
        e = hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return pf3d_hip_check(e, "pcg iterate");
        rel = sqrt(h.rr / bb);
        if (!isfinite(rel)) break;
    }
    *iters = it;
    *relres = rel;
    return PF3D_OK;
}

// FP64 PCG restarted from the true residual until the true residual meets
// rtol (the recursive residual of long runs drifts from b - A x)
int pcg64_restarted(const Mat &A, const double *b, double *x, double rtol, int64_t maxiter,
                    int x0, void *work, hipStream_t s, int64_t *iters, double *relres)
{
    int64_t total = 0;
    double rel = 0.;
    for (int restart = 0; restart < 8; restart++) {
        int64_t it = 0;
        int st = pcg_run<double, double>(A, b, x, rtol, (int)(maxiter - total), x0 || restart,
                                         work, s, &it, &rel, 25);
        if (st) return st;
        total += it;
        if (it == 0 || total >= maxiter) break;  // true residual already below rtol
    }
    *iters = total;
    *relres = rel;
    return PF3D_OK;
}

// Mixed-precision PCG with reliable updates: the Krylov recursion reads K
// and the block inverses in FP32 (vectors and arithmetic FP64); every
// `replace` iterations the recursive residual is replaced by the true FP64
// residual b - K64 x (search direction kept, r.z recomputed), so the
// iteration converges to the FP64 system's solution without the restarts
// of iterative refinement.  When `stall` consecutive replacements fail to
// halve the true residual the solve finishes with FP64 PCG from the current
// x.  stats: [0] iterations, [1] replacements, [2] FP64 iterations.
int pcg_mixed_run(const Mat &A64, const Mat &A32, const double *b, double *x, double rtol,
                  int64_t maxiter, int replace, int stall, int x0, void *work, hipStream_t s,
                  int64_t *stats, double *relres)
{
    const int64_t n = A64.nb * 6;
    char *w = (char *)work;
    double *r = (double *)w, *z = r + n, *p = z + n, *q = p + n;
    double *partial = (double *)(w + ((4 * n * 8 + 255) / 256) * 256);
    PcgState *st = (PcgState *)(partial + 2 * RED_BLOCKS);
    double *y = (double *)(w + pcg_bytes(n, sizeof(double)));
    const unsigned nv = nblk(n, VB), nr = red_blocks(A64.nb), nd = red_blocks(n);
    hipError_t e;
    hipLaunchKernelGGL(dot2<double>, dim3(nd), dim3(VB), 0, s, n, b, b, (const double *)nullptr,
                       (const double *)nullptr, partial);
    hipLaunchKernelGGL(finish, dim3(1), dim3(VB), 0, s, (int)nd, partial, st, 3);
    if (!x0) hipLaunchKernelGGL(fill0<double>, dim3(nv), dim3(VB), 0, s, n, x);
    hipLaunchKernelGGL(fill0<double>, dim3(nv), dim3(VB), 0, s, n, p);
    // r = b - K64 x ; z = M r ; (r.z, r.r) ; keep p (zero at the start)
    auto true_residual = [&]() -> hipError_t {
        hipError_t e2 = spmv<double, double>(A64, x, y, s);
        if (e2 != hipSuccess) return e2;
        hipLaunchKernelGGL(residual<double>, dim3(nv), dim3(VB), 0, s, n, b, y, r,
                           (double *)nullptr);
        hipLaunchKernelGGL(set_alpha, dim3(1), dim3(1), 0, s, st, 0.);
        hipLaunchKernelGGL((update_xrz<float, double>), dim3(nr), dim3(VB), 0, s, A64.nb, st, x,
                           r, p, q, (const float *)A32.dinv, z, partial);
        hipLaunchKernelGGL(finish, dim3(1), dim3(VB), 0, s, (int)nr, partial, st, 2);
        return hipGetLastError();
    };
    if ((e = true_residual()) != hipSuccess) return pf3d_hip_check(e, "mixed pcg init");
    hipMemcpyAsync(p, z, sizeof(double) * n, hipMemcpyDeviceToDevice, s);
    PcgState h;
    e = hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return pf3d_hip_check(e, "mixed pcg init");
    const double bb = h.bb > 0 ? h.bb : 1.;
    double rel = sqrt(h.rr / bb), best = rel;
    int64_t it = 0, nrep = 0;
    int since = 0;
    while (rel > rtol && it < maxiter && since < stall) {
        for (int k = 0; k < replace && it < maxiter; k++, it++) {
            spmv<float, double>(A32, p, q, s);
            hipLaunchKernelGGL(dot2<double>, dim3(nd), dim3(VB), 0, s, n, p, q,
                               (const double *)nullptr, (const double *)nullptr, partial);
            hipLaunchKernelGGL(finish, dim3(1), dim3(VB), 0, s, (int)nd, partial, st, 0);
            hipLaunchKernelGGL((update_xrz<float, double>), dim3(nr), dim3(VB), 0, s, A64.nb, st,
                               x, r, p, q, (const float *)A32.dinv, z, partial);
            hipLaunchKernelGGL(finish, dim3(1), dim3(VB), 0, s, (int)nr, partial, st, 1);
            hipLaunchKernelGGL(update_p<double>, dim3(nv), dim3(VB), 0, s, n, st, z, p);
        }
        if ((e = true_residual()) != hipSuccess) return pf3d_hip_check(e, "mixed pcg replace");
        nrep++;
        e = hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return pf3d_hip_check(e, "mixed pcg iterate");
        rel = sqrt(h.rr / bb);
        if (!isfinite(rel)) break;
        if (rel <= 0.5 * best) {
            best = rel;
            since = 0;
        } else {
            since++;
        }
    }
    stats[0] = it;
    stats[1] = nrep;
    stats[2] = 0;
    if (rel > rtol && isfinite(rel) && it < maxiter) {
        // stagnation of the FP32 operator: finish in FP64 from x
        int64_t it64 = 0;
        int st2 = pcg64_restarted(A64, b, x, rtol, maxiter - it, 1, work, s, &it64, &rel);
        if (st2) return st2;
        stats[2] = it64;
    }
    *relres = rel;
    return PF3D_OK;
}

}  // namespace

extern "C" {

int pf3d_bsr_spmv(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx, const double *val,
                  const double *x, double *y, void *stream)
{
    if (nb < 0 || (nb && (!row_ptr || !col_idx || !val || !x || !y)))
        return pf3d_fail(PF3D_E_ARG, "pf3d_bsr_spmv: bad arguments");
    if (!nb) return PF3D_OK;
    Mat A{nb, row_ptr, col_idx, val, nullptr};
    return pf3d_hip_check(spmv<double, double>(A, x, y, (hipStream_t)stream), "pf3d_bsr_spmv");
}

int pf3d_bsr_product_dd_mv(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx,
                           const double *val, int nv, int64_t ld, const double *x, double *y,
                           void *stream)
{
    if (nb < 0 || nv < 1 || nv > MAXV || ld < 6 * nb ||
        (nb && (!row_ptr || !col_idx || !val || !x || !y || x == y)))
        return pf3d_fail(PF3D_E_ARG, "pf3d_bsr_product_dd_mv: bad arguments");
    if (!nb) return PF3D_OK;
    hipLaunchKernelGGL(bsr_product_dd_mv, dim3(nblk(nb, 4)), dim3(256), 0, (hipStream_t)stream,
                       nb, row_ptr, col_idx, val, nv, ld, x, y);
    return pf3d_hip_check(hipGetLastError(), "pf3d_bsr_product_dd_mv");
}

int pf3d_bsr_residual_dd(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx,
                         const double *val, const double *x, const double *b, double *r,
                         void *stream)
{
    if (nb < 0 || (nb && (!row_ptr || !col_idx || !val || !x || !r)))
        return pf3d_fail(PF3D_E_ARG, "pf3d_bsr_residual_dd: bad arguments");
    if (!nb) return PF3D_OK;
    // b given: r = b - A x ; b null: r = A x
    hipLaunchKernelGGL(bsr_residual_dd, dim3(nblk(nb, 4)), dim3(256), 0, (hipStream_t)stream, nb,
                       row_ptr, col_idx, val, x, b, b ? -1.0 : 1.0, r);
    return pf3d_hip_check(hipGetLastError(), "pf3d_bsr_residual_dd");
}

int pf3d_bsr_spmv_f32(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx,
                      const float *val, const float *x, float *y, void *stream)
{
    if (nb < 0 || (nb && (!row_ptr || !col_idx || !val || !x || !y)))
        return pf3d_fail(PF3D_E_ARG, "pf3d_bsr_spmv_f32: bad arguments");
    if (!nb) return PF3D_OK;
    Mat A{nb, row_ptr, col_idx, val, nullptr};
    return pf3d_hip_check(spmv<float, float>(A, x, y, (hipStream_t)stream), "pf3d_bsr_spmv_f32");
}

int pf3d_bsr_fused(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx, const double *val,
                   const double *xin, const double *b, double *r, const double *dinv,
                   const double *dold, double *dnew, double *xacc, double sa, double sb,
                   double c1, double c2, void *stream)
{
    return fused_launch<double, double>(nb, row_ptr, col_idx, val, xin, b, r, dinv, dold, dnew,
                                        xacc, sa, sb, c1, c2, stream, "pf3d_bsr_fused");
}

int pf3d_bsr_fused_f32(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx,
                       const float *val, const float *xin, const float *b, float *r,
                       const float *dinv, const float *dold, float *dnew, float *xacc, double sa,
                       double sb, double c1, double c2, void *stream)
{
    return fused_launch<float, float>(nb, row_ptr, col_idx, val, xin, b, r, dinv, dold, dnew,
                                      xacc, sa, sb, c1, c2, stream, "pf3d_bsr_fused_f32");
}

int pf3d_bsr_fused_mv(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx,
                      const double *val, int nv, int64_t ldx, const double *xin, int64_t ld,
                      const double *b, double *r, const double *dinv, const double *dold,
                      double *dnew, double *xacc, double sa, double sb, double c1, double c2,
                      void *stream)
{
    return fused_mv_launch<double, double>(nb, row_ptr, col_idx, val, nv, ldx, xin, ld, b, r,
                                           dinv, dold, dnew, xacc, sa, sb, c1, c2, stream,
                                           "pf3d_bsr_fused_mv");
}

int pf3d_bsr_fused_mv_f32(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx,
                          const float *val, int nv, int64_t ldx, const float *xin, int64_t ld,
                          const float *b, float *r, const float *dinv, const float *dold,
                          float *dnew, float *xacc, double sa, double sb, double c1, double c2,
                          void *stream)
{
    return fused_mv_launch<float, float>(nb, row_ptr, col_idx, val, nv, ldx, xin, ld, b, r,
                                         dinv, dold, dnew, xacc, sa, sb, c1, c2, stream,
                                         "pf3d_bsr_fused_mv_f32");
}

int pf3d_bsr_dirichlet(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx, double *val,
                       const uint8_t *fixed, void *stream)
{
    if (nb <= 0 || !row_ptr || !col_idx || !val || !fixed)
        return pf3d_fail(PF3D_E_ARG, "pf3d_bsr_dirichlet: bad arguments");
    hipLaunchKernelGGL(bsr_bc, dim3(nblk(nb * 36, 256)), dim3(256), 0, (hipStream_t)stream, nb,
                       row_ptr, col_idx, val, fixed, 1.);
    return pf3d_hip_check(hipGetLastError(), "pf3d_bsr_dirichlet");
}

int pf3d_bsr_dirichlet_diag(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx,
                            double *val, const uint8_t *fixed, double diag, void *stream)
{
    if (nb <= 0 || !row_ptr || !col_idx || !val || !fixed)
        return pf3d_fail(PF3D_E_ARG, "pf3d_bsr_dirichlet_diag: bad arguments");
    hipLaunchKernelGGL(bsr_bc, dim3(nblk(nb * 36, 256)), dim3(256), 0, (hipStream_t)stream, nb,
                       row_ptr, col_idx, val, fixed, diag);
    return pf3d_hip_check(hipGetLastError(), "pf3d_bsr_dirichlet_diag");
}

int pf3d_to_f32(int64_t n, const double *a, float *b, void *stream)
{
    if (n < 0 || (n && (!a || !b))) return pf3d_fail(PF3D_E_ARG, "pf3d_to_f32: bad arguments");
    if (!n) return PF3D_OK;
    hipLaunchKernelGGL(copy_cast<float>, dim3(nblk(n, VB)), dim3(VB), 0, (hipStream_t)stream, n,
                       a, b);
    return pf3d_hip_check(hipGetLastError(), "pf3d_to_f32");
}

int pf3d_bjacobi(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx, const double *val,
                 double *dinv, float *dinv32, int *singular, void *stream)
{
    if (nb <= 0 || !row_ptr || !col_idx || !val || !dinv || !singular)
        return pf3d_fail(PF3D_E_ARG, "pf3d_bjacobi: bad arguments");
    hipLaunchKernelGGL(bjacobi_setup, dim3(nblk(nb, 64)), dim3(64), 0, (hipStream_t)stream, nb,
                       row_ptr, col_idx, val, dinv, dinv32, singular);
    return pf3d_hip_check(hipGetLastError(), "pf3d_bjacobi");
}

int pf3d_pcg_bytes(int64_t nb, int fp32, int64_t *bytes)
{
    if (nb <= 0 || !bytes) return pf3d_fail(PF3D_E_ARG, "pf3d_pcg_bytes: bad arguments");
    *bytes = (int64_t)pcg_bytes(6 * nb, fp32 ? sizeof(float) : sizeof(double));
    return PF3D_OK;
}

int pf3d_pcg(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx, const double *val,
             const double *dinv, const double *b, double *x, double rtol, int maxiter, int x0,
             void *work, int64_t work_bytes, int64_t *iters, double *relres, void *stream)
{
    if (nb <= 0 || !row_ptr || !col_idx || !val || !dinv || !b || !x || !work || !iters ||
        !relres)
        return pf3d_fail(PF3D_E_ARG, "pf3d_pcg: bad arguments");
    if ((int64_t)pcg_bytes(6 * nb, sizeof(double)) > work_bytes)
        return pf3d_fail(PF3D_E_ARG, "pf3d_pcg: workspace too small");
    Mat A{nb, row_ptr, col_idx, val, dinv};
    return pcg64_restarted(A, b, x, rtol, maxiter, x0, work, (hipStream_t)stream, iters, relres);
}

/* Allocation-free FP64 PCG (pf3d_pcg without the caller's workspace):
 * the four PCG vectors, the reduction partials and the solver state live
 * in `scratch`, a file-scope device array of fixed size (one copy per GPU,
 * held until the process ends).  Its address on each device comes from
 * hipGetSymbolAddress once, with that device current; host code uses that
 * address only, never the symbol itself.  Every call on one GPU shares the
 * array: calls that overlap on one GPU race. */
}  // extern "C"

namespace {
constexpr size_t PCG_STATIC_BYTES =
    ((4 * 6 * (size_t)PF3D_PCG_STATIC_MAX_ROWS * sizeof(double) + 255) / 256) * 256 +
    2 * RED_BLOCKS * sizeof(double) + 256;
}  // namespace

static __device__ __align__(256) char scratch[PCG_STATIC_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
void *scratch_dev[MAX_DEVICES];

// device address of `scratch` on the current device (looked up once per device)
int scratch_ptr(void **ptr)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return pf3d_hip_check(e, "pf3d_pcg_static: hipGetDevice");
    if (dev < 0 || dev >= MAX_DEVICES)
        return pf3d_fail(PF3D_E_ARG, "pf3d_pcg_static: device %d beyond %d", dev, MAX_DEVICES);
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return pf3d_hip_check(e, "pf3d_pcg_static: hipGetSymbolAddress");
        if (!p) return pf3d_fail(PF3D_E_HIP, "pf3d_pcg_static: no address for the scratch array");
        scratch_dev[dev] = p;
    }
    *ptr = scratch_dev[dev];
    return PF3D_OK;
}
}  // namespace

extern "C" {

int pf3d_pcg_static(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx,
                    const double *val, const double *dinv, const double *b, double *x,
                    double rtol, int maxiter, int x0, int64_t *iters, double *relres,
                    void *stream)
{
    if (nb <= 0 || !row_ptr || !col_idx || !val || !dinv || !b || !x || !iters || !relres)
        return pf3d_fail(PF3D_E_ARG, "pf3d_pcg_static: bad arguments");
    if (nb > PF3D_PCG_STATIC_MAX_ROWS)
        return pf3d_fail(PF3D_E_ARG,
                         "pf3d_pcg_static: %lld block rows exceed the static workspace "
                         "(at most %d; pf3d_pcg takes a caller workspace of any size)",
                         (long long)nb, PF3D_PCG_STATIC_MAX_ROWS);
    void *w = nullptr;
    int st = scratch_ptr(&w);
    if (st) return st;
    Mat A{nb, row_ptr, col_idx, val, dinv};
    return pcg64_restarted(A, b, x, rtol, maxiter, x0, w, (hipStream_t)stream, iters, relres);
}

/* FP64 iterative refinement around a PCG on the FP32-stored matrix:
 *   r = b - K64 x ; solve K32 d = r to inner_rtol ; x += d ; until
 *   ||b - K64 x|| <= rtol ||b||.  The inner PCG reads K and the block
 *   inverses in FP32 (half the bytes of the dominant stream) and keeps its
 *   vectors and arithmetic in FP64.  work: pf3d_ir_bytes(nb). */
int pf3d_ir_bytes(int64_t nb, int64_t *bytes)
{
    if (nb <= 0 || !bytes) return pf3d_fail(PF3D_E_ARG, "pf3d_ir_bytes: bad arguments");
    const int64_t n = 6 * nb;
    *bytes = (int64_t)pcg_bytes(n, sizeof(double)) + ((3 * n * 8 + 255) / 256) * 256 +
             2 * RED_BLOCKS * 8 + 256;
    return PF3D_OK;
}

int pf3d_ir_solve(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx,
                  const double *val64, const float *val32, const float *dinv32,
                  const double *b, double *x, double rtol, double inner_rtol, int max_outer,
                  int max_inner, int x0, void *work, int64_t work_bytes, int64_t *stats,
                  double *relres, void *stream)
{
    if (nb <= 0 || !row_ptr || !col_idx || !val64 || !val32 || !dinv32 || !b || !x || !work ||
        !stats || !relres)
        return pf3d_fail(PF3D_E_ARG, "pf3d_ir_solve: bad arguments");
    int64_t need;
    pf3d_ir_bytes(nb, &need);
    if (need > work_bytes) return pf3d_fail(PF3D_E_ARG, "pf3d_ir_solve: workspace too small");
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = 6 * nb;
    char *w = (char *)work;
    void *inner = w;
    char *o = w + pcg_bytes(n, sizeof(double));
    double *r64 = (double *)o, *y64 = r64 + n, *d64 = y64 + n;
    double *partial = (double *)(o + ((3 * n * 8 + 255) / 256) * 256);
    PcgState *st = (PcgState *)(partial + 2 * RED_BLOCKS);
    Mat A64{nb, row_ptr, col_idx, val64, nullptr};
    Mat A32{nb, row_ptr, col_idx, val32, dinv32};
    const unsigned nv = nblk(n, VB), nd = red_blocks(n);
    if (!x0) hipLaunchKernelGGL(fill0<double>, dim3(nv), dim3(VB), 0, s, n, x);
    hipLaunchKernelGGL(dot2<double>, dim3(nd), dim3(VB), 0, s, n, b, b, (const double *)nullptr,
                       (const double *)nullptr, partial);
    hipLaunchKernelGGL(finish, dim3(1), dim3(VB), 0, s, (int)nd, partial, st, 3);
    int64_t inner_total = 0;
    int outer = 0;
    double rel = 1.;
    for (;;) {
        hipError_t e = spmv<double, double>(A64, x, y64, s);
        if (e != hipSuccess) return pf3d_hip_check(e, "ir spmv");
        hipLaunchKernelGGL(residual<double>, dim3(nv), dim3(VB), 0, s, n, b, y64, r64,
                           (double *)nullptr);
        hipLaunchKernelGGL(dot2<double>, dim3(nd), dim3(VB), 0, s, n, r64, r64,
                           (const double *)nullptr, (const double *)nullptr, partial);
        hipLaunchKernelGGL(finish, dim3(1), dim3(VB), 0, s, (int)nd, partial, st, 4);
        PcgState h;
        e = hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return pf3d_hip_check(e, "ir residual");
        rel = sqrt(h.rr / (h.bb > 0 ? h.bb : 1.));
        if (rel <= rtol || outer >= max_outer || !isfinite(rel)) break;
        int64_t it = 0;
        double inner_rel = 0.;
        // inner solve on the FP32-stored matrix with FP64 vectors
        int st2 = pcg_run<float, double>(A32, r64, d64, inner_rtol, max_inner, 0, inner, s, &it,
                                         &inner_rel, 25);
        if (st2) return st2;
        inner_total += it;
        hipLaunchKernelGGL(add_to, dim3(nv), dim3(VB), 0, s, n, d64, x);
        outer++;
    }
    stats[0] = outer;
    stats[1] = inner_total;
    *relres = rel;
    if (!(rel <= rtol)) return pf3d_fail(PF3D_E_SOLVE, "IR stopped at relres %.3e", rel);
    return PF3D_OK;
}

int pf3d_pcg_mixed(int64_t nb, const int32_t *row_ptr, const int32_t *col_idx,
                   const double *val64, const double *dinv64, const float *val32,
                   const float *dinv32, const double *b, double *x, double rtol, int64_t maxiter,
                   int replace, int x0, void *work, int64_t work_bytes, int64_t *stats,
                   double *relres, void *stream)
{
    if (nb <= 0 || !row_ptr || !col_idx || !val64 || !dinv64 || !val32 || !dinv32 || !b || !x ||
        !work || !stats || !relres || replace < 1 || maxiter < 0)
        return pf3d_fail(PF3D_E_ARG, "pf3d_pcg_mixed: bad arguments");
    int64_t need;
    pf3d_ir_bytes(nb, &need);
    if (need > work_bytes) return pf3d_fail(PF3D_E_ARG, "pf3d_pcg_mixed: workspace too small");
    Mat A64{nb, row_ptr, col_idx, val64, dinv64};
    Mat A32{nb, row_ptr, col_idx, val32, dinv32};
    return pcg_mixed_run(A64, A32, b, x, rtol, maxiter, replace, 20, x0, work,
                         (hipStream_t)stream, stats, relres);
}

}  // extern "C"

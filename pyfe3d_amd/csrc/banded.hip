// Batched direct solves of many independent banded SPD systems: the
// design-sweep counterpart of the reference callers' scipy spsolve
// (tests/test_quad4_static_point_load.py:84-102), which factorises each
// free-DOF system with SuperLU.
//
// Each system v (a diagonal block of the block-diagonal BSR matrix, after
// the Dirichlet rows/columns have been replaced by identity ones) is
// renumbered by a node permutation that keeps its band narrow (for a
// structured nx x ny plate: nodes ordered along the longer side, so the
// band is ~6 (min(nx, ny) + 2) scalars) and copied into lower band storage
//     L[j*(b+1) + d] = A[j+d][j],  0 <= d <= b,
// then factorised in place by a blocked right-looking band Cholesky, one
// workgroup per system:
//   * panel of NB columns loaded into LDS and factorised by one wave (each
//     lane owns rows lane + 64q; pivots and multipliers travel by shuffle,
//     so the panel needs no workgroup barrier),
//   * symmetric rank-NB update of the trailing band window by the whole
//     workgroup.
// The triangular solves are blocked the same way (NB columns per barrier
// pair).  All sums have a fixed order: results do not depend on timing.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/pf3d.h"

int pf3d_fail(int code, const char *fmt, ...);
int pf3d_hip_check(hipError_t e, const char *where);

namespace {

constexpr int NB = 16;       // panel width
constexpr int MAXB = 640;    // largest supported bandwidth (LDS panel rows b + NB)
// launch configurations (threads per workgroup BT, largest bandwidth MB the
// LDS panel is sized for): a narrower panel fits more workgroups per CU
// (MB 640: 89 KB of LDS, one per CU; 448: 63 KB, two; 256: 37 KB, four)

struct BandSys {
    int64_t n_var;
    const int32_t *order;   // processing order (or null)
    const int64_t *dof0;    // [n_var+1] first DOF of each system (global numbering)
    const int32_t *bw;      // [n_var] bandwidth b
    const int64_t *boff;    // [n_var] offset of the band storage (doubles)
    double *band;
};

// ---------------------------------------------------------------- fill
__global__ void band_zero(int64_t n, double *__restrict__ p)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x)
        p[t] = 0.;
}

// one thread per (node row, entry of its 6x6 blocks): lower-triangle
// entries of the row's blocks go to the band of the node's system
__global__ void band_fill(int64_t n_nodes, const int32_t *__restrict__ rp,
                          const int32_t *__restrict__ ci, const double *__restrict__ val,
                          const int32_t *__restrict__ var, const int64_t *__restrict__ pinv,
                          const int64_t *__restrict__ dof0, const int32_t *__restrict__ bw,
                          const int64_t *__restrict__ boff, double *__restrict__ band)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_nodes * 36) return;
    const int64_t R = t / 36;
    const int e = (int)(t - R * 36), m = e / 6, nn = e % 6;
    const int v = var[R];
    const int64_t b = bw[v];
    const int64_t base = dof0[v];
    const int64_t i = 6 * pinv[R] + m - base;  // local row
    double *L = band + boff[v];
    for (int k = rp[R]; k < rp[R + 1]; k++) {
        const int64_t C = ci[k];
        const int64_t j = 6 * pinv[C] + nn - base;  // local column
        const int64_t d = i - j;
        if (d >= 0 && d <= b) L[j * (b + 1) + d] = val[(int64_t)k * 36 + e];
    }
}

// ------------------------------------------------------------ Cholesky
template <int BT, int MB>
__global__ void __launch_bounds__(BT) band_chol(BandSys s, int *__restrict__ bad)
{
    constexpr int PR = MB + NB;
    constexpr int QMAX = (PR + 63) / 64;  // panel rows per lane
    __shared__ double P[PR][NB + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t v = s.order ? (int64_t)s.order[blockIdx.x] : (int64_t)blockIdx.x;
    const int64_t n = s.dof0[v + 1] - s.dof0[v];
    const int b = s.bw[v];
    double *L = s.band + s.boff[v];
    const int64_t ld = b + 1;
    if (b > MB) {  // the panel would not fit: report, leave the band untouched
        if (tid == 0) atomicAdd(bad, 1);
        return;
    }
    bool failed = false;
    for (int64_t j0 = 0; j0 < n; j0 += NB) {
        const int jn = (int)(n - j0 < NB ? n - j0 : NB);
        const int rows = (int)(n - j0 < b + jn ? n - j0 : b + jn);
        // panel: P[r][c] = A[j0 + r][j0 + c], r >= c, r - c <= b
        for (int t = tid; t < rows * NB; t += BT) {
            const int c = t / rows, r = t - c * rows;
            double x = 0.;
            if (c < jn && r >= c && r - c <= b) x = L[(j0 + c) * ld + (r - c)];
            P[r][c] = x;
        }
        __syncthreads();
        if (wave == 0) {
            double l[QMAX];
            for (int c = 0; c < jn; c++) {
                // pivot (row c is owned by lane c, q = 0)
                double own = (lane == c) ? P[c][c] : 0.;
                const double d = __shfl(own, c, 64);
                if (!(d > 0.)) failed = true;
                const double sq = d > 0. ? sqrt(d) : 1.;
                const double inv = 1. / sq;
#pragma unroll
                for (int q = 0; q < QMAX; q++) {
                    const int r = lane + 64 * q;
                    double x = 0.;
                    if (r < rows && r >= c) {
                        x = (r == c) ? sq : P[r][c] * inv;
                        P[r][c] = x;
                    }
                    l[q] = x;
                }
                for (int c2 = c + 1; c2 < jn; c2++) {
                    const double mlt = __shfl(l[0], c2, 64);  // L[c2][c], row c2 < 64
#pragma unroll
                    for (int q = 0; q < QMAX; q++) {
                        const int r = lane + 64 * q;
                        if (r < rows && r >= c2) P[r][c2] -= l[q] * mlt;
                    }
                }
            }
        }
        __syncthreads();
        // write the panel's columns of L back
        for (int t = tid; t < rows * NB; t += BT) {
            const int c = t / rows, r = t - c * rows;
            if (c < jn && r >= c && r - c <= b) L[(j0 + c) * ld + (r - c)] = P[r][c];
        }
        // trailing update: A[j0+ri][j0+rk] -= sum_c P[ri][c] P[rk][c],
        // jn <= rk <= ri < rows, in 4 x 4 tiles per thread (8 LDS reads
        // per 16 products); consecutive threads take consecutive row tiles
        // of one column tile, so the global accesses coalesce
        const int T = rows - jn;
        const int nt = (T + 3) / 4;
        for (int64_t t = tid; t < (int64_t)nt * nt; t += BT) {
            const int tk = (int)(t / nt), ti = (int)(t % nt);
            if (ti < tk) continue;
            const int rk0 = jn + 4 * tk, ri0 = jn + 4 * ti;
            double acc[4][4];
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int q = 0; q < 4; q++) acc[a][q] = 0.;
#pragma unroll 4
            for (int c = 0; c < NB; c++) {
                double pk[4], pi[4];
#pragma unroll
                for (int a = 0; a < 4; a++) {
                    pk[a] = rk0 + a < rows ? P[rk0 + a][c] : 0.;
                    pi[a] = ri0 + a < rows ? P[ri0 + a][c] : 0.;
                }
#pragma unroll
                for (int a = 0; a < 4; a++)
#pragma unroll
                    for (int q = 0; q < 4; q++) acc[a][q] += pi[q] * pk[a];
            }
#pragma unroll
            for (int a = 0; a < 4; a++) {
                const int rk = rk0 + a;
                if (rk >= rows) continue;
                double *col = L + (j0 + rk) * ld;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int ri = ri0 + q;
                    if (ri >= rk && ri < rows) col[ri - rk] -= acc[a][q];
                }
            }
        }
        __syncthreads();
    }
    if (failed && lane == 0) atomicAdd(bad, 1);
}

// --------------------------------------------------------- triangular
// x <- A^-1 f for system v: permute f into y, L y = g, L^T x = y, scatter.
template <int BT>
__global__ void __launch_bounds__(BT) band_solve(BandSys s, const int64_t *__restrict__ pnode,
                                                  const double *__restrict__ f,
                                                  double *__restrict__ x,
                                                  double *__restrict__ work)
{
    __shared__ double yb[NB];
    __shared__ double part[NB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t v = s.order ? (int64_t)s.order[blockIdx.x] : (int64_t)blockIdx.x;
    const int64_t base = s.dof0[v];
    const int64_t n = s.dof0[v + 1] - base;
    const int b = s.bw[v];
    const double *L = s.band + s.boff[v];
    const int64_t ld = b + 1;
    double *y = work + base;  // local vector in permuted numbering
    // gather: y[6 p + d] = f[6 pnode[p] + d]
    for (int64_t t = tid; t < n; t += BT) {
        const int64_t p = t / 6;
        y[t] = f[6 * pnode[base / 6 + p] + (t - 6 * p)];
    }
    __syncthreads();
    // forward: L y = g
    for (int64_t j0 = 0; j0 < n; j0 += NB) {
        const int jn = (int)(n - j0 < NB ? n - j0 : NB);
        if (wave == 0) {
            // diagonal triangle: lane c holds y[j0 + c]
            double yc = lane < jn ? y[j0 + lane] : 0.;
            for (int c = 0; c < jn; c++) {
                double piv = lane == c ? yc / L[(j0 + c) * ld] : 0.;
                const double yv = __shfl(piv, c, 64);
                if (lane == c) yc = yv;
                if (lane > c && lane < jn && lane - c <= b) yc -= L[(j0 + c) * ld + (lane - c)] * yv;
            }
            if (lane < jn) {
                y[j0 + lane] = yc;
                yb[lane] = yc;
            }
        }
        __syncthreads();
        // rows below the block within the band
        const int64_t rend = (j0 + jn - 1 + b < n - 1) ? j0 + jn - 1 + b : n - 1;
        for (int64_t i = j0 + jn + tid; i <= rend; i += BT) {
            double acc = 0.;
            for (int c = 0; c < jn; c++) {
                const int64_t d = i - (j0 + c);
                if (d <= b) acc += L[(j0 + c) * ld + d] * yb[c];
            }
            y[i] -= acc;
        }
        __syncthreads();
    }
    // backward: L^T x = y, blocks from the end
    const int64_t nblk = (n + NB - 1) / NB;
    for (int64_t bk = nblk - 1; bk >= 0; bk--) {
        const int64_t j0 = bk * NB;
        const int jn = (int)(n - j0 < NB ? n - j0 : NB);
        // contributions of the solved rows below: s_c = sum_i L[i][j0+c] y[i]
        // for i in [j0+jn, j0+c+b]; 64 lanes per column, one wave per
        // column (columns c, c + BT/64, ... when there are fewer waves)
        for (int c = wave; c < NB; c += BT / 64) {
            double sacc = 0.;
            if (c < jn) {
                const int64_t j = j0 + c;
                const int64_t iend = (j + b < n - 1) ? j + b : n - 1;
                for (int64_t i = j0 + jn + lane; i <= iend; i += 64)
                    sacc += L[j * ld + (i - j)] * y[i];
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) sacc += __shfl_xor(sacc, o, 64);
            if (lane == 0) part[c] = sacc;
        }
        __syncthreads();
        if (wave == 0) {
            // back substitution in the diagonal triangle, lane c holds x[j0+c]
            double rhs = lane < jn ? y[j0 + lane] - part[lane] : 0.;
            double xc = 0.;
            for (int cc = jn - 1; cc >= 0; cc--) {
                double piv = lane == cc ? rhs / L[(j0 + cc) * ld] : 0.;
                const double xv = __shfl(piv, cc, 64);
                if (lane == cc) xc = xv;
                if (lane < cc && cc - lane <= b) rhs -= L[(j0 + lane) * ld + (cc - lane)] * xv;
            }
            if (lane < jn) y[j0 + lane] = xc;
        }
        __syncthreads();
    }
    // scatter back to the original numbering
    for (int64_t t = tid; t < n; t += BT) {
        const int64_t p = t / 6;
        x[6 * pnode[base / 6 + p] + (t - 6 * p)] = y[t];
    }
}

inline unsigned nb256(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

extern "C" {

int pf3d_band_max_width(void) { return MAXB; }

// band_chol<BT, MB> for a (threads, max_b) configuration; false if none
static bool launch_chol(int threads, int max_b, unsigned grid, hipStream_t s, BandSys sys,
                        int *bad)
{
#define PF3D_CHOL(BTV, MBV)                                                                 \
    if (threads == BTV && max_b == MBV) {                                                   \
        hipLaunchKernelGGL((band_chol<BTV, MBV>), dim3(grid), dim3(BTV), 0, s, sys, bad); \
        return true;                                                                        \
    }
    PF3D_CHOL(1024, 640) PF3D_CHOL(1024, 448) PF3D_CHOL(1024, 256)
    PF3D_CHOL(512, 640) PF3D_CHOL(512, 448) PF3D_CHOL(512, 256)
    PF3D_CHOL(256, 640) PF3D_CHOL(256, 448) PF3D_CHOL(256, 256)
#undef PF3D_CHOL
    return false;
}

int pf3d_band_factor(int64_t n_var, int64_t n_nodes, const int32_t *order, const int64_t *dof0,
                     const int32_t *bw, const int64_t *boff, int64_t band_size, double *band,
                     const int32_t *row_ptr, const int32_t *col_idx, const double *val,
                     const int32_t *var, const int64_t *pinv, int *bad, int threads, int max_b,
                     void *stream)
{
    if (threads == 0) threads = 1024;
    if (max_b == 0) max_b = MAXB;
    if ((threads != 256 && threads != 512 && threads != 1024) ||
        (max_b != 256 && max_b != 448 && max_b != 640))
        return pf3d_fail(PF3D_E_ARG, "pf3d_band_factor: threads must be 256/512/1024 and "
                                     "max_b 256/448/640 (got %d, %d)", threads, max_b);
    if (n_var < 0 || n_nodes < 0) return pf3d_fail(PF3D_E_ARG, "pf3d_band_factor: negative size");
    if (n_var == 0) return PF3D_OK;
    if (!dof0 || !bw || !boff || !band || !row_ptr || !col_idx || !val || !var || !pinv || !bad)
        return pf3d_fail(PF3D_E_ARG, "pf3d_band_factor: null argument");
    if (n_var > 2147483647) return pf3d_fail(PF3D_E_ARG, "too many systems");
    hipStream_t s = (hipStream_t)stream;
    const unsigned zb = nb256(band_size) < 1u << 20 ? nb256(band_size) : 1u << 20;
    hipLaunchKernelGGL(band_zero, dim3(zb), dim3(256), 0, s, band_size, band);
    hipLaunchKernelGGL(band_fill, dim3(nb256(n_nodes * 36)), dim3(256), 0, s, n_nodes, row_ptr,
                       col_idx, val, var, pinv, dof0, bw, boff, band);
    BandSys sys{n_var, order, dof0, bw, boff, band};
    launch_chol(threads, max_b, (unsigned)n_var, s, sys, bad);
    return pf3d_hip_check(hipGetLastError(), "pf3d_band_factor");
}

int pf3d_band_solve(int64_t n_var, const int32_t *order, const int64_t *dof0, const int32_t *bw,
                    const int64_t *boff, const double *band, const int64_t *pnode,
                    const double *f, double *x, double *work, int threads, void *stream)
{
    if (n_var < 0) return pf3d_fail(PF3D_E_ARG, "pf3d_band_solve: negative size");
    if (n_var == 0) return PF3D_OK;
    if (!dof0 || !bw || !boff || !band || !pnode || !f || !x || !work)
        return pf3d_fail(PF3D_E_ARG, "pf3d_band_solve: null argument");
    if (threads == 0) threads = 1024;
    BandSys sys{n_var, order, dof0, bw, boff, const_cast<double *>(band)};
    const dim3 grid((unsigned)n_var);
    hipStream_t s = (hipStream_t)stream;
    if (threads == 1024)
        hipLaunchKernelGGL(band_solve<1024>, grid, dim3(1024), 0, s, sys, pnode, f, x, work);
    else if (threads == 512)
        hipLaunchKernelGGL(band_solve<512>, grid, dim3(512), 0, s, sys, pnode, f, x, work);
    else if (threads == 256)
        hipLaunchKernelGGL(band_solve<256>, grid, dim3(256), 0, s, sys, pnode, f, x, work);
    else
        return pf3d_fail(PF3D_E_ARG, "pf3d_band_solve: threads must be 256, 512 or 1024");
    return pf3d_hip_check(hipGetLastError(), "pf3d_band_solve");
}

}  // extern "C"

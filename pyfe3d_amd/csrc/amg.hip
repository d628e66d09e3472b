// Multigrid setup on the GPU: the 6x6-block sparse products that build a
// smoothed-aggregation hierarchy (tentative prolongator smoothing, the
// Galerkin coarse operator R A P, R = P^T), so that only the per-aggregate
// work (geometric aggregation, orthonormalisation of the rigid-body modes)
// stays on the host.
//
// All matrices are 6x6 BSR with int32 row pointers and int32 block column
// indices ascending within each row; values are FP64, 36 per block,
// row-major.  Every result is computed in a fixed order (no floating-point
// atomics), so repeated setups give the same bits.
#include <hipcub/hipcub.hpp>

#include "../../include/pf3d.h"
#include "pf3d_device.hpp"

int pf3d_fail(int code, const char *fmt, ...);
int pf3d_hip_check(hipError_t e, const char *where);

namespace {

constexpr int MAXC = 128;   // distinct block columns per row of a product

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

// Sorted distinct block columns of row i of X Y (insertion into a sorted
// list; the rows here hold tens of columns).  Returns -1 past MAXC.
__device__ int product_columns(int i, const int32_t *__restrict__ xr, const int32_t *__restrict__ xc,
                               const int32_t *__restrict__ yr, const int32_t *__restrict__ yc,
                               int32_t *cols)
{
    int n = 0;
    for (int p = xr[i]; p < xr[i + 1]; p++) {
        const int k = xc[p];
        for (int q = yr[k]; q < yr[k + 1]; q++) {
            const int c = yc[q];
            int pos = n;
            while (pos > 0 && cols[pos - 1] > c) pos--;
            if (pos > 0 && cols[pos - 1] == c) continue;
            if (n == MAXC) return -1;
            for (int s = n; s > pos; s--) cols[s] = cols[s - 1];
            cols[pos] = c;
            n++;
        }
    }
    return n;
}

__global__ void spgemm_count(int64_t nr, const int32_t *__restrict__ xr,
                             const int32_t *__restrict__ xc, const int32_t *__restrict__ yr,
                             const int32_t *__restrict__ yc, int64_t *__restrict__ cnt,
                             int *__restrict__ over)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nr) return;
    if (i == nr) { cnt[nr] = 0; return; }
    int32_t cols[MAXC];
    const int n = product_columns((int)i, xr, xc, yr, yc, cols);
    if (n < 0) atomicAdd(over, 1);
    cnt[i] = n < 0 ? 0 : n;
}

__global__ void to_int32(int64_t n, const int64_t *__restrict__ a, int32_t *__restrict__ b)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (int32_t)a[i];
}

__global__ void spgemm_columns(int64_t nr, const int32_t *__restrict__ xr,
                               const int32_t *__restrict__ xc, const int32_t *__restrict__ yr,
                               const int32_t *__restrict__ yc, const int32_t *__restrict__ cr,
                               int32_t *__restrict__ cc)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nr) return;
    int32_t cols[MAXC];
    const int n = product_columns((int)i, xr, xc, yr, yc, cols);
    const int base = cr[i];
    for (int s = 0; s < n && base + s < cr[i + 1]; s++) cc[base + s] = cols[s];
}

// row of entry t: the last i with rp[i] <= t
__device__ __forceinline__ int row_of(const int32_t *__restrict__ rp, int nr, int64_t t)
{
    int lo = 0, hi = nr;   // answer in [0, nr)
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (rp[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

// position of column c in row k, or -1
__device__ __forceinline__ int find_col(const int32_t *__restrict__ rp,
                                        const int32_t *__restrict__ ci, int k, int c)
{
    int lo = rp[k], hi = rp[k + 1];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ci[mid] < c) lo = mid + 1; else hi = mid;
    }
    return (lo < rp[k + 1] && ci[lo] == c) ? lo : -1;
}

// C block t = sum over k of X(i, k) Y(k, c), k ascending
__global__ void spgemm_values(int64_t nr, int64_t nnz, const int32_t *__restrict__ xr,
                              const int32_t *__restrict__ xc, const double *__restrict__ xv,
                              const int32_t *__restrict__ yr, const int32_t *__restrict__ yc,
                              const double *__restrict__ yv, const int32_t *__restrict__ cr,
                              const int32_t *__restrict__ cc, double *__restrict__ cv)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nnz) return;
    const int i = row_of(cr, (int)nr, t);
    const int c = cc[t];
    double acc[36];
#pragma unroll
    for (int e = 0; e < 36; e++) acc[e] = 0.;
    for (int p = xr[i]; p < xr[i + 1]; p++) {
        const int q = find_col(yr, yc, xc[p], c);
        if (q < 0) continue;
        const double *X = xv + (int64_t)p * 36;
        const double *Y = yv + (int64_t)q * 36;
#pragma unroll
        for (int m = 0; m < 6; m++) {
            double xm[6], ym[6];
#pragma unroll
            for (int r = 0; r < 6; r++) { xm[r] = X[r * 6 + m]; ym[r] = Y[m * 6 + r]; }
#pragma unroll
            for (int r = 0; r < 6; r++)
#pragma unroll
                for (int s = 0; s < 6; s++) acc[r * 6 + s] = fma(xm[r], ym[s], acc[r * 6 + s]);
        }
    }
    double *C = cv + t * 36;
#pragma unroll
    for (int e = 0; e < 36; e++) C[e] = acc[e];
}

// P = P0 - omega D^-1 T with T = A P0; P0 row i holds q[i] at column agg[i]
__global__ void smooth_prolongator(int64_t nr, const int32_t *__restrict__ tr,
                                   const int32_t *__restrict__ tc, const double *__restrict__ tv,
                                   const int32_t *__restrict__ agg, const double *__restrict__ q,
                                   const double *__restrict__ dinv, double omega,
                                   double *__restrict__ pv, int *__restrict__ missing)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nr) return;
    double D[36];
#pragma unroll
    for (int e = 0; e < 36; e++) D[e] = dinv[i * 36 + e];
    bool found = false;
    for (int t = tr[i]; t < tr[i + 1]; t++) {
        const double *T = tv + (int64_t)t * 36;
        const bool own = tc[t] == agg[i];
        found |= own;
        for (int r = 0; r < 6; r++)
            for (int s = 0; s < 6; s++) {
                double a = 0.;
                for (int m = 0; m < 6; m++) a = fma(D[r * 6 + m], T[m * 6 + s], a);
                pv[(int64_t)t * 36 + r * 6 + s] = (own ? q[i * 36 + r * 6 + s] : 0.) - omega * a;
            }
    }
    if (!found) atomicAdd(missing, 1);
}

// out(i, j) = (V(i, j) + V(j, i)^T) / 2, plus 1 on the diagonal of every
// DOF flagged in dead (coarse DOFs without a near-null mode)
__global__ void symmetrize(int64_t nr, int64_t nnz, const int32_t *__restrict__ rp,
                           const int32_t *__restrict__ ci, const double *__restrict__ v,
                           const uint8_t *__restrict__ dead, double *__restrict__ out,
                           int *__restrict__ unpaired)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nnz) return;
    const int i = row_of(rp, (int)nr, t);
    const int j = ci[t];
    const int u = find_col(rp, ci, j, i);
    if (u < 0) atomicAdd(unpaired, 1);
    for (int r = 0; r < 6; r++)
        for (int s = 0; s < 6; s++) {
            double a = v[t * 36 + r * 6 + s];
            if (u >= 0) a = 0.5 * (a + v[(int64_t)u * 36 + s * 6 + r]);
            if (i == j && r == s && dead && dead[(int64_t)i * 6 + r]) a += 1.;
            out[t * 36 + r * 6 + s] = a;
        }
}

__global__ void transpose_keys(int64_t nr, const int32_t *__restrict__ rp,
                               const int32_t *__restrict__ ci, int64_t *__restrict__ keys,
                               int32_t *__restrict__ ids)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nr) return;
    for (int t = rp[i]; t < rp[i + 1]; t++) {
        keys[t] = (int64_t)ci[t] * nr + i;
        ids[t] = t;
    }
}

__global__ void transpose_emit(int64_t nr, int64_t nnz, const int64_t *__restrict__ keys,
                               const int32_t *__restrict__ ids, const double *__restrict__ v,
                               int32_t *__restrict__ tci, double *__restrict__ tv)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nnz) return;
    tci[s] = (int32_t)(keys[s] % nr);
    const double *V = v + (int64_t)ids[s] * 36;
    for (int r = 0; r < 6; r++)
        for (int c = 0; c < 6; c++) tv[s * 36 + r * 6 + c] = V[c * 6 + r];
}

// row pointer of the transpose: first sorted key of each column
__global__ void transpose_rows(int64_t nr, int64_t nc, int64_t nnz,
                               const int64_t *__restrict__ keys, int32_t *__restrict__ trp)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c > nc) return;
    const int64_t key = c * nr;
    int64_t lo = 0, hi = nnz;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (keys[mid] < key) lo = mid + 1; else hi = mid;
    }
    trp[c] = (int32_t)lo;
}

int bits_for(double maxkey)
{
    int b = 1;
    while (b < 63 && (double)(1ULL << b) < maxkey) b++;
    return b;
}

struct SpgemmWork {
    size_t cnt, scan, over, temp, total, temp_bytes;
};

hipError_t spgemm_layout(int64_t nr, SpgemmWork &w)
{
    size_t tb = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, tb, (int64_t *)nullptr,
                                                    (int64_t *)nullptr, (int)(nr + 1));
    if (e != hipSuccess) return e;
    w.temp_bytes = align256(tb);
    const size_t a = align256(sizeof(int64_t) * (size_t)(nr + 1));
    size_t o = 0;
    w.cnt = o; o += a;
    w.scan = o; o += a;
    w.over = o; o += 256;
    w.temp = o; o += w.temp_bytes;
    w.total = o;
    return hipSuccess;
}

struct TransposeWork {
    size_t keys_in, keys_out, ids_in, ids_out, temp, total, temp_bytes;
};

hipError_t transpose_layout(int64_t nnz, int64_t nr, int64_t nc, TransposeWork &w)
{
    size_t tb = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(
        nullptr, tb, (const int64_t *)nullptr, (int64_t *)nullptr, (const int32_t *)nullptr,
        (int32_t *)nullptr, (int)nnz, 0, bits_for((double)nr * (double)(nc + 1)));
    if (e != hipSuccess) return e;
    w.temp_bytes = align256(tb);
    const size_t a8 = align256(sizeof(int64_t) * (size_t)nnz);
    const size_t a4 = align256(sizeof(int32_t) * (size_t)nnz);
    size_t o = 0;
    w.keys_in = o; o += a8;
    w.keys_out = o; o += a8;
    w.ids_in = o; o += a4;
    w.ids_out = o; o += a4;
    w.temp = o; o += w.temp_bytes;
    w.total = o;
    return hipSuccess;
}

}  // namespace

extern "C" {

int pf3d_bsr_spgemm_bytes(int64_t nr, int64_t *bytes)
{
    if (nr <= 0 || nr >= INT32_MAX || !bytes)
        return pf3d_fail(PF3D_E_ARG, "pf3d_bsr_spgemm_bytes: bad arguments");
    SpgemmWork w;
    hipError_t e = spgemm_layout(nr, w);
    if (e != hipSuccess) return pf3d_hip_check(e, "pf3d_bsr_spgemm_bytes");
    *bytes = (int64_t)w.total;
    return PF3D_OK;
}

int pf3d_bsr_spgemm_symbolic(int64_t nr, const int32_t *xr, const int32_t *xc, const int32_t *yr,
                             const int32_t *yc, int32_t *cr, void *work, int64_t work_bytes,
                             int64_t *nnz, void *stream)
{
    if (nr <= 0 || nr >= INT32_MAX || !xr || !xc || !yr || !yc || !cr || !work || !nnz)
        return pf3d_fail(PF3D_E_ARG, "pf3d_bsr_spgemm_symbolic: bad arguments");
    SpgemmWork w;
    hipError_t e = spgemm_layout(nr, w);
    if (e != hipSuccess) return pf3d_hip_check(e, "pf3d_bsr_spgemm_symbolic");
    if ((int64_t)w.total > work_bytes)
        return pf3d_fail(PF3D_E_ARG, "pf3d_bsr_spgemm_symbolic: workspace too small");
    hipStream_t s = (hipStream_t)stream;
    char *base = (char *)work;
    int64_t *cnt = (int64_t *)(base + w.cnt), *scan = (int64_t *)(base + w.scan);
    int *over = (int *)(base + w.over);
    e = hipMemsetAsync(over, 0, sizeof(int), s);
    if (e != hipSuccess) return pf3d_hip_check(e, "pf3d_bsr_spgemm_symbolic");
    hipLaunchKernelGGL(spgemm_count, dim3(nblk(nr + 1, 128)), dim3(128), 0, s, nr, xr, xc, yr, yc,
                       cnt, over);
    size_t tb = w.temp_bytes;
    e = hipcub::DeviceScan::ExclusiveSum(base + w.temp, tb, cnt, scan, (int)(nr + 1), s);
    if (e != hipSuccess) return pf3d_hip_check(e, "pf3d_bsr_spgemm_symbolic");
    int64_t total = 0;
    int n_over = 0;
    e = hipMemcpyAsync(&total, scan + nr, sizeof(int64_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&n_over, over, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return pf3d_hip_check(e, "pf3d_bsr_spgemm_symbolic");
    if (n_over)
        return pf3d_fail(PF3D_E_ARG, "pf3d_bsr_spgemm_symbolic: %d rows with more than %d "
                         "block columns", n_over, MAXC);
    if (total >= INT32_MAX)
        return pf3d_fail(PF3D_E_ARG, "pf3d_bsr_spgemm_symbolic: %lld blocks exceed int32",
                         (long long)total);
    hipLaunchKernelGGL(to_int32, dim3(nblk(nr + 1, 256)), dim3(256), 0, s, nr + 1, scan, cr);
    *nnz = total;
    return pf3d_hip_check(hipGetLastError(), "pf3d_bsr_spgemm_symbolic");
}

int pf3d_bsr_spgemm_numeric(int64_t nr, const int32_t *xr, const int32_t *xc, const double *xv,
                            const int32_t *yr, const int32_t *yc, const double *yv,
                            const int32_t *cr, int32_t *cc, double *cv, int64_t nnz, void *stream)
{
    if (nr <= 0 || nr >= INT32_MAX || nnz < 0 || !xr || !xc || !xv || !yr || !yc || !yv || !cr ||
        (nnz > 0 && (!cc || !cv)))
        return pf3d_fail(PF3D_E_ARG, "pf3d_bsr_spgemm_numeric: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(spgemm_columns, dim3(nblk(nr, 128)), dim3(128), 0, s, nr, xr, xc, yr, yc,
                       cr, cc);
    if (nnz > 0)
        hipLaunchKernelGGL(spgemm_values, dim3(nblk(nnz, 128)), dim3(128), 0, s, nr, nnz, xr, xc,
                           xv, yr, yc, yv, cr, cc, cv);
    return pf3d_hip_check(hipGetLastError(), "pf3d_bsr_spgemm_numeric");
}

int pf3d_bsr_transpose_bytes(int64_t nr, int64_t nc, int64_t nnzb, int64_t *bytes)
{
    if (nr <= 0 || nc <= 0 || nnzb <= 0 || nnzb >= INT32_MAX || !bytes)
        return pf3d_fail(PF3D_E_ARG, "pf3d_bsr_transpose_bytes: bad arguments");
    TransposeWork w;
    hipError_t e = transpose_layout(nnzb, nr, nc, w);
    if (e != hipSuccess) return pf3d_hip_check(e, "pf3d_bsr_transpose_bytes");
    *bytes = (int64_t)w.total;
    return PF3D_OK;
}

int pf3d_bsr_transpose(int64_t nr, int64_t nc, int64_t nnzb, const int32_t *rp, const int32_t *ci,
                       const double *val, int32_t *trp, int32_t *tci, double *tval, void *work,
                       int64_t work_bytes, void *stream)
{
    if (nr <= 0 || nc <= 0 || nnzb <= 0 || nnzb >= INT32_MAX || !rp || !ci || !val || !trp ||
        !tci || !tval || !work)
        return pf3d_fail(PF3D_E_ARG, "pf3d_bsr_transpose: bad arguments");
    TransposeWork w;
    hipError_t e = transpose_layout(nnzb, nr, nc, w);
    if (e != hipSuccess) return pf3d_hip_check(e, "pf3d_bsr_transpose");
    if ((int64_t)w.total > work_bytes)
        return pf3d_fail(PF3D_E_ARG, "pf3d_bsr_transpose: workspace too small");
    hipStream_t s = (hipStream_t)stream;
    char *base = (char *)work;
    int64_t *kin = (int64_t *)(base + w.keys_in), *kout = (int64_t *)(base + w.keys_out);
    int32_t *iin = (int32_t *)(base + w.ids_in), *iout = (int32_t *)(base + w.ids_out);
    hipLaunchKernelGGL(transpose_keys, dim3(nblk(nr, 256)), dim3(256), 0, s, nr, rp, ci, kin, iin);
    size_t tb = w.temp_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(base + w.temp, tb, kin, kout, iin, iout, (int)nnzb, 0,
                                           bits_for((double)nr * (double)(nc + 1)), s);
    if (e != hipSuccess) return pf3d_hip_check(e, "pf3d_bsr_transpose");
    hipLaunchKernelGGL(transpose_emit, dim3(nblk(nnzb, 256)), dim3(256), 0, s, nr, nnzb, kout,
                       iout, val, tci, tval);
    hipLaunchKernelGGL(transpose_rows, dim3(nblk(nc + 1, 256)), dim3(256), 0, s, nr, nc, nnzb,
                       kout, trp);
    return pf3d_hip_check(hipGetLastError(), "pf3d_bsr_transpose");
}

int pf3d_amg_smooth_prolongator(int64_t nr, const int32_t *tr, const int32_t *tc, const double *tv,
                                const int32_t *agg, const double *q, const double *dinv,
                                double omega, double *pv, int *missing, void *stream)
{
    if (nr <= 0 || nr >= INT32_MAX || !tr || !tc || !tv || !agg || !q || !dinv || !pv || !missing)
        return pf3d_fail(PF3D_E_ARG, "pf3d_amg_smooth_prolongator: bad arguments");
    hipLaunchKernelGGL(smooth_prolongator, dim3(nblk(nr, 128)), dim3(128), 0, (hipStream_t)stream,
                       nr, tr, tc, tv, agg, q, dinv, omega, pv, missing);
    return pf3d_hip_check(hipGetLastError(), "pf3d_amg_smooth_prolongator");
}

int pf3d_bsr_symmetrize(int64_t nr, int64_t nnzb, const int32_t *rp, const int32_t *ci,
                        const double *val, const uint8_t *dead, double *out, int *unpaired,
                        void *stream)
{
    if (nr <= 0 || nr >= INT32_MAX || nnzb <= 0 || !rp || !ci || !val || !out || !unpaired ||
        val == out)
        return pf3d_fail(PF3D_E_ARG, "pf3d_bsr_symmetrize: bad arguments");
    hipLaunchKernelGGL(symmetrize, dim3(nblk(nnzb, 256)), dim3(256), 0, (hipStream_t)stream, nr,
                       nnzb, rp, ci, val, dead, out, unpaired);
    return pf3d_hip_check(hipGetLastError(), "pf3d_bsr_symmetrize");
}

}  // extern "C"

// Batched static solves of many independent systems (a design sweep) in one
// persistent launch.
//
// The systems are the diagonal blocks of one block-diagonal 6x6 BSR matrix:
// system v owns block rows [row0[v], row0[v+1]) and its blocks reference
// only those rows.  Each workgroup takes the next system from a queue
// (atomic counter; the caller may pass a processing order, e.g. largest
// first) and runs the whole solve on it with workgroup barriers only: no
// grid synchronisation and no host round trip, so the launch can be
// captured in a graph together with the assembly that precedes it.
//
// Per system, the reference caller's pattern (assemble, slice the free
// DOFs, spsolve; tests/test_quad4_static_point_load.py:80-102, and the
// re-assembly loop of the design-sweep tutorials) becomes:
//   plain   : PCG with 6x6 block-Jacobi on the FP64 matrix, restarted from
//             the true residual b - K x until ||b - K x|| <= rtol ||b||;
//   mixed   : FP64 iterative refinement, r = b - K64 x, inner PCG on the
//             FP32-stored matrix and block inverses (FP64 vectors and
//             arithmetic) to inner_rtol ||r||, x += d, until the FP64 true
//             residual meets rtol.
// x may hold an initial guess (x0 != 0).
//
// Work mapping inside a workgroup (256 threads, 4 waves): lanes 0..59 of
// each wave own 10 block rows, lane = 6*row_in_wave + component, so a
// block row's six threads sit in one wave and exchange values with
// shuffles.  Every phase uses this mapping; only the sparse product reads
// other threads' entries (after a barrier).  Reductions use a fixed tree
// (wave shuffles, then the 4 wave sums in order), so results do not depend
// on timing.
#include <math.h>
#include <stdint.h>

#include "../../include/pf3d.h"
#include "pf3d_device.hpp"

int pf3d_fail(int code, const char *fmt, ...);
int pf3d_hip_check(hipError_t e, const char *where);

namespace {

constexpr int BT = 256;        // threads per workgroup
constexpr int NW = BT / 64;    // waves per workgroup
constexpr int RPW = 10;        // block rows per wave and pass
constexpr int RPG = NW * RPW;  // block rows per workgroup pass

struct BatchArgs {
    int64_t n_var;
    const int64_t *row0;
    const int32_t *order;
    const int32_t *rp, *ci;
    const double *val64, *dinv64;
    const float *val32, *dinv32;
    const double *b;
    double *x;
    int x0;
    double rtol, inner_rtol;
    int max_iter, max_outer;
    double *r, *z, *p, *q, *d;
    int *counter;
    int32_t *iters, *outer;
    double *relres;
};

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// sums of two values over the workgroup; red has room for 2 buffers of
// 2*NW doubles, used alternately so one barrier per reduction suffices
__device__ __forceinline__ void wg_sum2(double &a, double &b, double *red, int &buf)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    a = wave_sum(a);
    b = wave_sum(b);
    double *s = red + buf * 2 * NW;
    if (lane == 0) {
        s[wave] = a;
        s[NW + wave] = b;
    }
    __syncthreads();
    a = 0.;
    b = 0.;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        a += s[w];
        b += s[NW + w];
    }
    buf ^= 1;
}

struct Map {
    int slot;  // block row within a pass, or -1 for idle lanes
    int comp;  // component 0..5
    int base;  // lane of component 0 of this block row
};

__device__ __forceinline__ Map lane_map()
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    Map m;
    if (lane < 6 * RPW) {
        m.slot = wave * RPW + lane / 6;
        m.comp = lane % 6;
        m.base = lane - m.comp;
    } else {
        m.slot = -1;
        m.comp = 0;
        m.base = lane;
    }
    return m;
}

// out = K in over rows [r0, r1); returns this thread's part of dot(out, w)
template <typename MT>
__device__ __forceinline__ double spmv_part(const Map &m, int64_t r0, int64_t r1,
                                            const int32_t *__restrict__ rp,
                                            const int32_t *__restrict__ ci,
                                            const MT *__restrict__ val,
                                            const double *__restrict__ in,
                                            double *__restrict__ out,
                                            const double *__restrict__ w)
{
    double dot = 0.;
    if (m.slot < 0) return 0.;
    for (int64_t row = r0 + m.slot; row < r1; row += RPG) {
        const int b0 = rp[row], b1 = rp[row + 1];
        double acc = 0.;
#pragma unroll 3
        for (int bk = b0; bk < b1; bk++) {
            const MT *v = val + (int64_t)bk * 36 + m.comp * 6;
            const double *xv = in + (int64_t)ci[bk] * 6;
#pragma unroll
            for (int k = 0; k < 6; k++) acc += (double)v[k] * xv[k];
        }
        const int64_t t = row * 6 + m.comp;
        out[t] = acc;
        if (w) dot += acc * w[t];
    }
    return dot;
}

// z = Dinv r for the thread's entries (r already final); returns parts of
// (r.z, r.r).  All lanes of the wave execute the shuffles.
template <typename MT>
__device__ __forceinline__ void precond_part(const Map &m, int64_t r0, int64_t r1,
                                             const MT *__restrict__ dinv,
                                             const double *__restrict__ r, double *__restrict__ z,
                                             double &rz, double &rr)
{
    for (int64_t pass = r0; pass < r1; pass += RPG) {
        const int64_t row = pass + (m.slot < 0 ? 0 : m.slot);
        const bool on = m.slot >= 0 && row < r1;
        const double rv = on ? r[row * 6 + m.comp] : 0.;
        double s = 0.;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const double rk = __shfl(rv, m.base + k, 64);
            if (on) s += (double)dinv[row * 36 + m.comp * 6 + k] * rk;
        }
        if (on) {
            z[row * 6 + m.comp] = s;
            rz += s * rv;
            rr += rv * rv;
        }
    }
}

// PCG on operator (val, dinv) for K sol = rhs given the initial residual
// in r (r = rhs - K sol).  Stops when r.r <= tol2 or after max_it
// iterations.  Returns iterations done; rr_out = final recursive r.r.
template <typename MT>
__device__ int pcg_wg(const Map &m, int64_t r0, int64_t r1, const int32_t *rp, const int32_t *ci,
                      const MT *val, const MT *dinv, double *sol, double *r, double *z,
                      double *p, double *q, double tol2, int max_it, double *red, int &buf,
                      double &rr_out)
{
    double rz = 0., rr = 0.;
    precond_part<MT>(m, r0, r1, dinv, r, z, rz, rr);
    wg_sum2(rz, rr, red, buf);
    // p = z
    if (m.slot >= 0)
        for (int64_t row = r0 + m.slot; row < r1; row += RPG) {
            const int64_t t = row * 6 + m.comp;
            p[t] = z[t];
        }
    int it = 0;
    while (rr > tol2 && it < max_it) {
        __syncthreads();  // p complete before the product reads it
        double pq = spmv_part<MT>(m, r0, r1, rp, ci, val, p, q, p), dummy = 0.;
        wg_sum2(pq, dummy, red, buf);
        if (!(pq > 0.) || !isfinite(pq)) break;
        const double alpha = rz / pq;
        if (m.slot >= 0)
            for (int64_t row = r0 + m.slot; row < r1; row += RPG) {
                const int64_t t = row * 6 + m.comp;
                sol[t] += alpha * p[t];
                r[t] -= alpha * q[t];
            }
        double rz_new = 0.;
        rr = 0.;
        precond_part<MT>(m, r0, r1, dinv, r, z, rz_new, rr);
        wg_sum2(rz_new, rr, red, buf);
        const double beta = rz_new / rz;
        rz = rz_new;
        if (m.slot >= 0)
            for (int64_t row = r0 + m.slot; row < r1; row += RPG) {
                const int64_t t = row * 6 + m.comp;
                p[t] = z[t] + beta * p[t];
            }
        it++;
    }
    rr_out = rr;
    __syncthreads();
    return it;
}

// r = b - K64 x (true residual); returns r.r and (optionally) b.b
__device__ __forceinline__ void true_residual(const Map &m, int64_t r0, int64_t r1,
                                              const BatchArgs &a, double &rr, double &bb,
                                              double *red, int &buf)
{
    __syncthreads();  // x complete
    spmv_part<double>(m, r0, r1, a.rp, a.ci, a.val64, a.x, a.q, nullptr);
    rr = 0.;
    bb = 0.;
    if (m.slot >= 0)
        for (int64_t row = r0 + m.slot; row < r1; row += RPG) {
            const int64_t t = row * 6 + m.comp;
            const double bv = a.b[t], rv = bv - a.q[t];
            a.r[t] = rv;
            rr += rv * rv;
            bb += bv * bv;
        }
    wg_sum2(rr, bb, red, buf);
}

__global__ void reset_counter(int *c) { *c = 0; }

template <bool MIXED>
__global__ void __launch_bounds__(BT) batch_solve(BatchArgs a)
{
    __shared__ double red[4 * NW];
    __shared__ int sk;
    const Map m = lane_map();
    int buf = 0;
    for (;;) {
        if (threadIdx.x == 0) sk = atomicAdd(a.counter, 1);
        __syncthreads();
        const int64_t k = sk;
        __syncthreads();
        if (k >= a.n_var) break;
        const int64_t v = a.order ? (int64_t)a.order[k] : k;
        const int64_t r0 = a.row0[v], r1 = a.row0[v + 1];
        if (!a.x0 && m.slot >= 0)
            for (int64_t row = r0 + m.slot; row < r1; row += RPG) a.x[row * 6 + m.comp] = 0.;
        double rr, bb;
        true_residual(m, r0, r1, a, rr, bb, red, buf);
        const double tol2 = a.rtol * a.rtol * (bb > 0. ? bb : 1.);
        int its = 0, outer = 0;
        while (rr > tol2 && its < a.max_iter && outer < a.max_outer) {
            double rr_in;
            if (MIXED) {
                // d = 0; inner solve K32 d = r to inner_rtol ||r||, x += d
                if (m.slot >= 0)
                    for (int64_t row = r0 + m.slot; row < r1; row += RPG) a.d[row * 6 + m.comp] = 0.;
                double itol2 = a.inner_rtol * a.inner_rtol * rr;
                if (itol2 < 0.25 * tol2) itol2 = 0.25 * tol2;
                its += pcg_wg<float>(m, r0, r1, a.rp, a.ci, a.val32, a.dinv32, a.d, a.r, a.z, a.p,
                                     a.q, itol2, a.max_iter - its, red, buf, rr_in);
                if (m.slot >= 0)
                    for (int64_t row = r0 + m.slot; row < r1; row += RPG) {
                        const int64_t t = row * 6 + m.comp;
                        a.x[t] += a.d[t];
                    }
            } else {
                // iterate on x directly from its true residual (a restart)
                its += pcg_wg<double>(m, r0, r1, a.rp, a.ci, a.val64, a.dinv64, a.x, a.r, a.z,
                                      a.p, a.q, 0.25 * tol2, a.max_iter - its, red, buf, rr_in);
            }
            outer++;
            true_residual(m, r0, r1, a, rr, bb, red, buf);
            if (!isfinite(rr)) break;
        }
        if (threadIdx.x == 0) {
            a.iters[v] = its;
            a.outer[v] = outer;
            a.relres[v] = sqrt(rr / (bb > 0. ? bb : 1.));
        }
    }
}

}  // namespace

extern "C" {

int pf3d_batch_bytes(int64_t nb, int64_t *bytes)
{
    if (nb <= 0 || !bytes) return pf3d_fail(PF3D_E_ARG, "pf3d_batch_bytes: bad arguments");
    *bytes = (int64_t)(5 * 6 * nb * sizeof(double)) + 256;
    return PF3D_OK;
}

int pf3d_batch_solve(int64_t n_var, int64_t nb, const int64_t *row0, const int32_t *order,
                     const int32_t *row_ptr, const int32_t *col_idx, const double *val64,
                     const double *dinv64, const float *val32, const float *dinv32,
                     const double *b, double *x, int x0, double rtol, double inner_rtol,
                     int max_iter, int max_outer, void *work, int64_t work_bytes,
                     int32_t *iters, int32_t *outer, double *relres, int n_groups, void *stream)
{
    if (n_var < 0 || nb < 0) return pf3d_fail(PF3D_E_ARG, "pf3d_batch_solve: negative size");
    if (n_var == 0) return PF3D_OK;
    if (!row0 || !row_ptr || !col_idx || !val64 || !b || !x || !work || !iters || !outer ||
        !relres)
        return pf3d_fail(PF3D_E_ARG, "pf3d_batch_solve: null argument");
    const bool mixed = val32 != nullptr;
    if (mixed && !dinv32) return pf3d_fail(PF3D_E_ARG, "mixed mode needs dinv32");
    if (!mixed && !dinv64) return pf3d_fail(PF3D_E_ARG, "plain mode needs dinv64");
    if (!(rtol > 0.) || !(inner_rtol > 0.) || max_iter < 0 || max_outer < 1)
        return pf3d_fail(PF3D_E_ARG, "pf3d_batch_solve: bad tolerance or limits");
    int64_t need;
    pf3d_batch_bytes(nb, &need);
    if (work_bytes < need)
        return pf3d_fail(PF3D_E_ARG, "pf3d_batch_solve: workspace %lld < %lld bytes",
                         (long long)work_bytes, (long long)need);
    BatchArgs a{};
    a.n_var = n_var; a.row0 = row0; a.order = order; a.rp = row_ptr; a.ci = col_idx;
    a.val64 = val64; a.dinv64 = dinv64; a.val32 = val32; a.dinv32 = dinv32; a.b = b; a.x = x;
    a.x0 = x0; a.rtol = rtol; a.inner_rtol = inner_rtol; a.max_iter = max_iter;
    a.max_outer = max_outer; a.iters = iters; a.outer = outer; a.relres = relres;
    const int64_t n = 6 * nb;
    double *w = (double *)work;
    a.r = w; a.z = w + n; a.p = w + 2 * n; a.q = w + 3 * n; a.d = w + 4 * n;
    a.counter = (int *)(w + 5 * n);
    hipStream_t s = (hipStream_t)stream;
    int groups = n_groups > 0 ? n_groups : 2048;
    if (groups > n_var) groups = (int)n_var;
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, s, a.counter);
    if (mixed)
        hipLaunchKernelGGL(batch_solve<true>, dim3(groups), dim3(BT), 0, s, a);
    else
        hipLaunchKernelGGL(batch_solve<false>, dim3(groups), dim3(BT), 0, s, a);
    return pf3d_hip_check(hipGetLastError(), "pf3d_batch_solve");
}

}  // extern "C"

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
// re-assembly loop of the design-sweep tutorials) becomes a 6x6
// block-Jacobi preconditioned CG on x with FP64 vectors and arithmetic:
//   plain : operator and block inverses in FP64;
//   mixed : operator and block inverses read in FP32 (about 0.6x the bytes
//           of an iteration), with reliable updates: every `replace`
//           iterations, and whenever the recursive residual claims
//           convergence, r is replaced by the FP64 true residual b - K64 x
//           and the iteration continues with the same search direction
//           (no restart, so the Krylov space is kept).
// Both stop when the FP64 true residual meets ||b - K64 x|| <= rtol ||b||.
// x may hold an initial guess (x0 != 0).
//
// Work mapping inside a workgroup (BT threads, BT/64 waves): lanes 0..59 of
// each wave own 10 block rows, lane = 6*row_in_wave + component, so a
// block row's six threads sit in one wave and exchange values with
// shuffles.  Every phase uses this mapping; only the sparse products read
// other threads' entries (after a barrier).  Reductions use a fixed tree
// (wave shuffles, then the wave sums in order), so results do not depend
// on timing.
#include <math.h>
#include <stdint.h>

#include <type_traits>

#include "../../include/pf3d.h"
#include "pf3d_device.hpp"

int pf3d_fail(int code, const char *fmt, ...);
int pf3d_hip_check(hipError_t e, const char *where);

namespace {

constexpr int RPW = 10;  // block rows per wave and pass

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
    double rtol;
    int replace;
    int max_iter;
    double *r, *z, *p, *q;
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
// (a buffer is rewritten only after the next reduction's barrier, which
// every thread reaches after reading it)
template <int NW>
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

// out = K in over rows [r0, r1) (RPG rows per pass); returns this thread's
// part of dot(out, w)
template <int RPG, typename MT>
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

// z = Dinv r for the thread's entries; adds this thread's parts of (r.z,
// r.r).  All lanes of the wave execute the shuffles.
template <int RPG, typename MT>
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

// r = b - K64 x (FP64 true residual) over the system's rows; returns the
// thread parts of r.r and b.b
template <int RPG>
__device__ __forceinline__ void true_residual_part(const Map &m, int64_t r0, int64_t r1,
                                                   const BatchArgs &a, double &rr, double &bb)
{
    __syncthreads();  // x complete
    spmv_part<RPG, double>(m, r0, r1, a.rp, a.ci, a.val64, a.x, a.q, nullptr);
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
}

__global__ void reset_counter(int *c) { *c = 0; }

template <int BT, bool MIXED>
__global__ void __launch_bounds__(BT) batch_solve(BatchArgs a)
{
    using MT = typename std::conditional<MIXED, float, double>::type;
    constexpr int NW = BT / 64, RPG = NW * RPW;
    __shared__ double red[4 * NW];
    __shared__ int sk;
    const Map m = lane_map();
    const MT *val = MIXED ? (const MT *)a.val32 : (const MT *)a.val64;
    const MT *dinv = MIXED ? (const MT *)a.dinv32 : (const MT *)a.dinv64;
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
        double rr, bb, bscratch;
        true_residual_part<RPG>(m, r0, r1, a, rr, bb);
        wg_sum2<NW>(rr, bb, red, buf);
        const double tol2 = a.rtol * a.rtol * (bb > 0. ? bb : 1.);
        int its = 0, replaced = 1;
        bool rr_true = true;
        double best = rr;  // smallest true r.r so far
        // operator of the iteration: MT (FP32 in mixed mode) until the true
        // residual stagnates, FP64 afterwards
        bool use64 = !MIXED;
        auto precond = [&](double &rz_out, double &rr_out) {
            if (use64)
                precond_part<RPG, double>(m, r0, r1, a.dinv64, a.r, a.z, rz_out, rr_out);
            else
                precond_part<RPG, MT>(m, r0, r1, dinv, a.r, a.z, rz_out, rr_out);
        };
        auto product = [&]() {
            return use64 ? spmv_part<RPG, double>(m, r0, r1, a.rp, a.ci, a.val64, a.p, a.q, a.p)
                         : spmv_part<RPG, MT>(m, r0, r1, a.rp, a.ci, val, a.p, a.q, a.p);
        };
        auto set_p = [&](double beta) {
            if (m.slot >= 0)
                for (int64_t row = r0 + m.slot; row < r1; row += RPG) {
                    const int64_t t = row * 6 + m.comp;
                    a.p[t] = a.z[t] + beta * a.p[t];
                }
        };
        double rz = 0., dummy = 0.;
        {
            double rr2 = 0.;
            precond(rz, rr2);
            wg_sum2<NW>(rz, dummy, red, buf);
        }
        set_p(0.);
        // stagnation watch (mixed mode): the true r.r must halve within
        // every 20 true-residual evaluations, else the FP32 operator has
        // reached its accuracy for this system
        double mark = rr;
        int since = 0;
        while (its < a.max_iter && isfinite(rr)) {
            bool restart = false;
            if (rr <= tol2) {
                if (rr_true) break;
                // the recursive residual claims convergence: check it
                true_residual_part<RPG>(m, r0, r1, a, rr, bscratch);
                wg_sum2<NW>(rr, bscratch, red, buf);
                replaced++;
                rr_true = true;
                if (rr <= tol2) break;
                restart = true;
            }
            if (rr_true) {
                if (rr < best) best = rr;
                bool switched = false;
                if (MIXED && !use64) {
                    if (rr < 0.5 * mark) {
                        mark = rr;
                        since = 0;
                    } else if (++since >= 20) {
                        use64 = true;  // finish this system with the FP64 operator
                        switched = restart = true;
                    }
                }
                if (restart) {
                    // new direction from the true residual; beta = 0 (a plain
                    // restart) when the operator changed or the residual rose
                    // far above the best one (lost conjugacy)
                    double rz_new = 0., rr2 = 0.;
                    precond(rz_new, rr2);
                    wg_sum2<NW>(rz_new, dummy, red, buf);
                    const double beta = (switched || rr > 1e4 * best) ? 0. : rz_new / rz;
                    rz = rz_new;
                    set_p(beta);
                }
            }
            __syncthreads();  // p complete before the product reads it
            double pq = product();
            wg_sum2<NW>(pq, dummy, red, buf);
            if (!(pq > 0.) || !isfinite(pq)) break;
            const double alpha = rz / pq;
            if (m.slot >= 0)
                for (int64_t row = r0 + m.slot; row < r1; row += RPG) {
                    const int64_t t = row * 6 + m.comp;
                    a.x[t] += alpha * a.p[t];
                    a.r[t] -= alpha * a.q[t];
                }
            its++;
            rr_true = false;
            if (MIXED && !use64 && a.replace > 0 && its % a.replace == 0) {
                // reliable update: r <- b - K64 x (x changed by other threads)
                true_residual_part<RPG>(m, r0, r1, a, rr, bscratch);
                replaced++;
                rr_true = true;
            }
            double rz_new = 0.;
            rr = 0.;
            precond(rz_new, rr);
            wg_sum2<NW>(rz_new, rr, red, buf);
            const double beta = (rr_true && rr > 1e4 * best) ? 0. : rz_new / rz;
            rz = rz_new;
            set_p(beta);
        }
        if (!rr_true) {
            true_residual_part<RPG>(m, r0, r1, a, rr, bscratch);
            wg_sum2<NW>(rr, bscratch, red, buf);
            replaced++;
        }
        if (threadIdx.x == 0) {
            a.iters[v] = its;
            a.outer[v] = replaced;
            a.relres[v] = sqrt(rr / (bb > 0. ? bb : 1.));
        }
        __syncthreads();
    }
}

template <int BT>
void launch_batch(bool mixed, int groups, hipStream_t s, const BatchArgs &a)
{
    if (mixed)
        hipLaunchKernelGGL((batch_solve<BT, true>), dim3(groups), dim3(BT), 0, s, a);
    else
        hipLaunchKernelGGL((batch_solve<BT, false>), dim3(groups), dim3(BT), 0, s, a);
}

}  // namespace

extern "C" {

int pf3d_batch_bytes(int64_t nb, int64_t *bytes)
{
    if (nb <= 0 || !bytes) return pf3d_fail(PF3D_E_ARG, "pf3d_batch_bytes: bad arguments");
    *bytes = (int64_t)(4 * 6 * nb * sizeof(double)) + 256;
    return PF3D_OK;
}

int pf3d_batch_solve(int64_t n_var, int64_t nb, const int64_t *row0, const int32_t *order,
                     const int32_t *row_ptr, const int32_t *col_idx, const double *val64,
                     const double *dinv64, const float *val32, const float *dinv32,
                     const double *b, double *x, int x0, double rtol, int replace, int max_iter,
                     void *work, int64_t work_bytes, int32_t *iters, int32_t *outer,
                     double *relres, int n_groups, int threads, void *stream)
{
    if (n_var < 0 || nb < 0) return pf3d_fail(PF3D_E_ARG, "pf3d_batch_solve: negative size");
    if (n_var == 0) return PF3D_OK;
    if (!row0 || !row_ptr || !col_idx || !val64 || !b || !x || !work || !iters || !outer ||
        !relres)
        return pf3d_fail(PF3D_E_ARG, "pf3d_batch_solve: null argument");
    const bool mixed = val32 != nullptr;
    if (mixed && !dinv32) return pf3d_fail(PF3D_E_ARG, "mixed mode needs dinv32");
    if (!mixed && !dinv64) return pf3d_fail(PF3D_E_ARG, "plain mode needs dinv64");
    if (!(rtol > 0.) || max_iter < 0 || replace < 0)
        return pf3d_fail(PF3D_E_ARG, "pf3d_batch_solve: bad tolerance or limits");
    if (threads == 0) threads = 1024;
    if (threads != 256 && threads != 512 && threads != 1024)
        return pf3d_fail(PF3D_E_ARG, "pf3d_batch_solve: threads must be 256, 512 or 1024");
    int64_t need;
    pf3d_batch_bytes(nb, &need);
    if (work_bytes < need)
        return pf3d_fail(PF3D_E_ARG, "pf3d_batch_solve: workspace %lld < %lld bytes",
                         (long long)work_bytes, (long long)need);
    BatchArgs a{};
    a.n_var = n_var; a.row0 = row0; a.order = order; a.rp = row_ptr; a.ci = col_idx;
    a.val64 = val64; a.dinv64 = dinv64; a.val32 = val32; a.dinv32 = dinv32; a.b = b; a.x = x;
    a.x0 = x0; a.rtol = rtol; a.replace = replace; a.max_iter = max_iter;
    a.iters = iters; a.outer = outer; a.relres = relres;
    const int64_t n = 6 * nb;
    double *w = (double *)work;
    a.r = w; a.z = w + n; a.p = w + 2 * n; a.q = w + 3 * n;
    a.counter = (int *)(w + 4 * n);
    hipStream_t s = (hipStream_t)stream;
    int groups = n_groups > 0 ? n_groups : 256 * (1024 / threads);
    if (groups > n_var) groups = (int)n_var;
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, s, a.counter);
    if (threads == 256) launch_batch<256>(mixed, groups, s, a);
    else if (threads == 512) launch_batch<512>(mixed, groups, s, a);
    else launch_batch<1024>(mixed, groups, s, a);
    return pf3d_hip_check(hipGetLastError(), "pf3d_batch_solve");
}

}  // extern "C"

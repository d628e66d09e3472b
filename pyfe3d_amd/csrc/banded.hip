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
//   * panel of 32 columns loaded into LDS: diagonal block factorised by one
//     wave, the rows below by all threads,
//   * symmetric rank-32 update of the trailing band window by the whole
//     workgroup;
// optionally with the forward substitution L y = f run panel by panel from
// the panel in LDS (the loads are known when the sweep factorises).  The
// triangular solves are blocked by 32 columns per barrier pair (both passes,
// or the backward one after a fused forward pass).  All sums have a fixed
// order: results do not depend on timing or launch shape.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/pf3d.h"

int pf3d_fail(int code, const char *fmt, ...);
int pf3d_hip_check(hipError_t e, const char *where);

namespace {

constexpr int MAXB = 576;    // largest supported bandwidth (LDS panel rows b + 32)
// launch configurations of the factorisation (threads per workgroup BT,
// largest bandwidth MB the LDS panel of 32 columns is sized for): a
// narrower panel fits more workgroups per CU (MB 576: 157 KB of LDS, one
// per CU; 448: 124 KB, one; 256: 74 KB, two)

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

// value of x in lane l (l uniform across the wave), as a scalar broadcast
__device__ __forceinline__ double readlane_d(double x, int l)
{
    const long long v = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)v, l);
    const int hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// ------------------------------------------------------------ Cholesky
// Per panel step of NF = 32 columns (j0 .. j0 + jn):
//   1. the panel (rows r = 0 .. b + jn of the columns) goes to LDS;
//   2. wave 0 factorises the jn x jn diagonal block in registers (two
//      lanes per row, 16 columns each; pivot by readlane, multipliers
//      broadcast through LDS);
//   3. every thread solves rows of the block below it, x L11^T = a, by
//      forward substitution against L11 read from LDS (broadcast);
//   4. the panel is written back and the trailing band window is updated,
//      A[ri][rk] -= sum_c L[ri][c] L[rk][c] for jn <= rk <= ri < rows, on
//      16 x 16 tiles by FP64 MFMA (one tile per wave at a time, operands
//      from LDS, the tile read from and written back to the band).
// A 32-column panel moves the trailing window through HBM once per 32
// columns.  Sums run in a fixed order, independent of the launch shape.
constexpr int NF = 32;   // factor panel width
constexpr int NH = NF / 2;  // diagonal block: columns per lane
typedef double v4d __attribute__((ext_vector_type(4)));

// A trailing-update tile (I, K) is computed transposed: the MFMA output
// D'[i][j] = A[ri][rk] - sum_c P[ri][c] P[rk][c] with rk = column block
// K's i-th window column and ri = row block I's j-th window row, so lane l
// holds window row ri = jn + 16 I + l % 16 of the columns rk0 + 4 v,
// rk0 = jn + 16 K + l / 16.  One load or store instruction then covers 16
// consecutive rows (128 contiguous bytes of the band) of 4 columns rather
// than 4 rows of 16 columns.  The A operand is P row jn + 16 K + l % 16
// (ra), the B operand P row jn + 16 I + l % 16 (rb); each output element
// receives the same products in the same order as untransposed.
struct TileAt {
    int ra, rb, ri, rk0;
};

__device__ __forceinline__ TileAt make_tile(int I, int K, int jn, int li, int lk)
{
    return TileAt{jn + 16 * K + li, jn + 16 * I + li, jn + 16 * I + li, jn + 16 * K + lk};
}

// tile t of the lower triangle in row-major order ((0,0), (1,0), (1,1),
// (2,0) ...; t uniform across the wave)
__device__ __forceinline__ TileAt tile_at(int t, int jn, int li, int lk)
{
    int I = (int)((sqrtf(8.f * t + 1.f) - 1.f) * 0.5f);
    I = __builtin_amdgcn_readfirstlane(I);
    while ((I + 1) * (I + 2) / 2 <= t) I++;
    while (I * (I + 1) / 2 > t) I--;
    const int K = t - I * (I + 1) / 2;
    return make_tile(I, K, jn, li, lk);
}

// tile u of the trailing update's column-major walk of the lower triangle
// of nt x nt tiles: down tile column K = 0 (rows I = K .. nt - 1), then
// K = 1, ... , so that the waves working at once stream down the same
// band columns (each column's entries are contiguous in the band) rather
// than across 16 columns a band column apart.  With look-ahead (la) the
// next diagonal block's tiles (0, 0), (1, 0), (1, 1) are left out (wave 0
// takes them): columns 0 and 1 start at row 2, then the walk continues on
// the triangle of columns and rows >= 2.  u uniform across the wave.
__device__ __forceinline__ TileAt tile_cm(int u, int nt, bool la, int jn, int li, int lk)
{
    int off = 0, m = nt, v = u;
    int I, K;
    if (la) {
        const int h = nt - 2;  // rows 2 .. nt - 1 of columns 0 and 1
        off = 2;
        m = h;
        if (u < 2 * h) {
            K = u < h ? 0 : 1;
            I = 2 + (u < h ? u : u - h);
            return make_tile(I, K, jn, li, lk);
        }
        v = u - 2 * h;
    }
    // column k of an m x m lower triangle starts at s_k = k m - k (k - 1) / 2
    const float q = (float)(2 * m + 1);
    int k = (int)((q - sqrtf(q * q - 8.f * v)) * 0.5f);
    k = __builtin_amdgcn_readfirstlane(k);
    k = k < 0 ? 0 : k;
    while (k > 0 && k * m - k * (k - 1) / 2 > v) k--;
    while ((k + 1) * m - (k + 1) * k / 2 <= v) k++;
    const int i = k + (v - (k * m - k * (k - 1) / 2));
    I = off + i;
    K = off + k;
    return make_tile(I, K, jn, li, lk);
}

// the tile's window entries A[ri][rk0 + 4 v] from the band (Lp: the
// panel's first column).  Addresses are clamped into the window, so every
// lane loads; entries outside the window's lower triangle are never stored
// and only reach their own output through the update.
__device__ __forceinline__ v4d tile_load(const TileAt &q, const double *Lp, int64_t ld,
                                         int rows)
{
    v4d a;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int rk = min(q.rk0 + 4 * v, rows - 1);
        const int ri = min(max(q.ri, rk), rows - 1);
        a[v] = Lp[rk * ld + (ri - rk)];
    }
    return a;
}

__device__ __forceinline__ void tile_store(const TileAt &q, const v4d &a, double *Lp,
                                           int64_t ld, int rows)
{
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int rk = q.rk0 + 4 * v;
        if (rk < rows && q.ri < rows && q.ri >= rk) Lp[rk * ld + (q.ri - rk)] = a[v];
    }
}

// a -= P[ra][0..31] P[rb][0..31]^T on the (transposed) tile; operand rows
// past the panel only feed outputs that are not stored
template <int PR>
__device__ __forceinline__ v4d tile_update(const TileAt &q, v4d a, const double (*P)[NF + 1],
                                           int lk)
{
    const int ra = min(q.ra, PR - 1), rb = min(q.rb, PR - 1);
    double x[NF / 4], y[NF / 4];
#pragma unroll
    for (int st = 0; st < NF / 4; st++) {
        x[st] = -P[ra][4 * st + lk];
        y[st] = P[rb][4 * st + lk];
    }
#pragma unroll
    for (int st = 0; st < NF / 4; st++)
        a = __builtin_amdgcn_mfma_f64_16x16x4f64(x[st], y[st], a, 0, 0, 0);
    return a;
}

// The jn x jn diagonal block held in P rows 0 .. 31, factorised in place
// by one wave (L11 and the reciprocal pivots dinv; *failed set on a
// non-positive pivot).  Lane (h, r) = (lane / 32, lane % 32) holds columns
// 16 h .. 16 h + 15 of row r.  Step k takes the pivot by readlane and
// broadcasts the multipliers L[r][k], k < r < jn, through LDS (zero
// elsewhere, two buffers so step k + 1 never overwrites what step k reads);
// the rank-1 update then runs on all 16 columns without masks: the lower
// triangle gets exactly its terms, the upper triangle and rows past jn hold
// values that are never broadcast.
__device__ __forceinline__ void diag_block(double (*P)[NF + 1], double *dinv, double (*Lk)[NF],
                                           int jn, int lane, bool &failed)
{
    int r = lane & 31, h = lane >> 5;
    // opaque per call, so the per-step lane masks and column offsets below
    // are formed where they are used instead of being kept live across
    // panels
    asm volatile("" : "+v"(r), "+v"(h));
    double l[NH];
#pragma unroll
    for (int i = 0; i < NH; i++) l[i] = P[r][NH * h + i];
#pragma unroll
    for (int k = 0; k < NF; k++) {
        if (k < jn) {
            const int kh = k / NH, ki = k % NH;
            const double d = readlane_d(l[ki], 32 * kh + k);
            if (!(d > 0.)) failed = true;
            const double sq = d > 0. ? sqrt(d) : 1.;
            const double inv = 1. / sq;
            if (h == kh) {
                l[ki] = r == k ? sq : l[ki] * inv;
                Lk[k & 1][r] = (r > k && r < jn) ? l[ki] : 0.;
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
            __builtin_amdgcn_wave_barrier();
            const double lr = Lk[k & 1][r];
#pragma unroll
            for (int i = 0; i < NH; i++) l[i] -= lr * Lk[k & 1][NH * h + i];
        }
    }
    // rows jn .. 31 exist in P only when jn = 32; the upper
    // triangle of the block is never read
#pragma unroll
    for (int i = 0; i < NH; i++) P[r][NH * h + i] = l[i];
    double dg = 1.;
#pragma unroll
    for (int i = 0; i < NH; i++)
        if (NH * h + i == r) dg = l[i];
    if (r < jn && (r >> 4) == h) dinv[r] = 1. / dg;
}

// forward substitution fused into the factorisation (y null: factor only):
// y[6 p + d] = f[6 pnode[p] + d] per system, then L y = y panel by panel
// from the panel in LDS, so the triangular solve afterwards only runs the
// backward half and reads the factor once
struct FwdArgs {
    const int64_t *pnode;
    const double *f;
    double *y;  // one double per DOF (permuted numbering)
};

template <int BT, int MB>
__global__ void __launch_bounds__(BT) band_chol(BandSys s, FwdArgs fw, int *__restrict__ bad)
{
    constexpr int PR = MB + NF;
    constexpr int RCH = (PR + 63) / 64;  // 64-row chunks of a panel column
    constexpr int NW = BT / 64;
    __shared__ double P[PR][NF + 1];
    __shared__ double dinv[NF];
    __shared__ double Lk[2][NF];
    __shared__ double yb[2][NF];  // solved forward unknowns of a block (by panel parity)
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
    const bool fwd = fw.y != nullptr;
    double *y = fwd ? fw.y + s.dof0[v] : nullptr;
    if (fwd) {
        // (visible to the first panel's substitution after its load barrier)
        const int64_t base = s.dof0[v];
        for (int64_t t = tid; t < n; t += BT) {
            const int64_t p = t / 6;
            y[t] = fw.f[6 * fw.pnode[base / 6 + p] + (t - 6 * p)];
        }
    }
    bool failed = false;
    bool ahead = false;  // this panel's diagonal block was factorised ahead
    for (int64_t j0 = 0; j0 < n; j0 += NF) {
        const int jn = (int)(n - j0 < NF ? n - j0 : NF);
        const int rows = (int)(n - j0 < b + jn ? n - j0 : b + jn);
        const int skip = ahead ? jn : 0;  // rows of P already holding L11
        // Forward substitution.  With a next panel, wave 0 finishes the next
        // block's rows (rows jn .. jn + 31 here) and solves through the next
        // diagonal block during its look-ahead (4b), and waves 1 .. NW - 1
        // update the rows below those; otherwise the block is solved in 3b
        // and all threads update the rows below.  The unknowns this step
        // reads (final since the last step) are loaded now, so the panel
        // load covers their latency.
        const bool la_step = n - (j0 + jn) > 0;
        const int fr0 = la_step ? jn + NF : jn;  // first row of 4a'
        const int ft0 = la_step ? tid - 64 : tid, fnt = la_step ? BT - 64 : BT;
        const int par = (int)((j0 / NF) & 1);
        double yblk = 0., yrow = 0., ynx = 0.;
        if (fwd) {
            if (wave == 0 && lane < jn && !ahead) yblk = y[j0 + lane];
            if (la_step && wave == 0 && lane < NF && jn + lane < n - j0) ynx = y[j0 + jn + lane];
            if (ft0 >= 0 && fr0 + ft0 < rows) yrow = y[j0 + fr0 + ft0];
        }
        // 1. panel: P[r][c] = A[j0 + r][j0 + c], r >= c, r - c <= b (zero
        //    elsewhere); wave w takes columns w, w + NW, ..., 64 rows per
        //    load, every load of a column in flight before its LDS stores
        for (int c = wave; c < NF; c += NW) {
            const double *col = L + (j0 + c) * ld;
            double x[RCH];
#pragma unroll
            for (int q = 0; q < RCH; q++) {
                const int r = 64 * q + lane;
                x[q] = (c < jn && r >= c && r >= skip && r < rows && r - c <= b) ? col[r - c] : 0.;
            }
#pragma unroll
            for (int q = 0; q < RCH; q++) {
                const int r = 64 * q + lane;
                if (r >= skip && r < rows) P[r][c] = x[q];
            }
        }
        __syncthreads();
        // 2. the diagonal block by wave 0 (unless the previous panel step
        //    factorised it ahead, see 4b)
        if (!ahead) {
            if (wave == 0) diag_block(P, dinv, Lk, jn, lane, failed);
            __syncthreads();
        }
        // 3. rows below the diagonal block: x L11^T = a, by columns: x[c] is
        //    final once scaled, then its term goes into every later x[c2].
        //    Each x[c2] receives its terms in the order c = 0, 1, ... as in
        //    forward substitution, while the terms of one column are
        //    independent.  (Rows exist below the block only when jn = NF.)
        for (int r = jn + tid; r < rows; r += BT) {
            // compiler barrier: L11 is re-read from LDS per row rather than
            // hoisted into ~500 registers for the few rows a thread takes
            asm volatile("" ::: "memory");
            double x[NF];
#pragma unroll
            for (int c = 0; c < NF; c++) x[c] = P[r][c];
#pragma unroll
            for (int c = 0; c < NF; c++) {
                x[c] *= dinv[c];
#pragma unroll
                for (int c2 = c + 1; c2 < NF; c2++) x[c2] -= x[c] * P[c2][c];
                // keep each column's L11 loads next to their use
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int c = 0; c < NF; c++) P[r][c] = x[c];
        }
        __syncthreads();
        // 3b. forward substitution through the panel's diagonal block by
        //     wave 0 (lane c finishes unknown c with its pivot's reciprocal,
        //     readlane hands it on), the same operations in the same order
        //     as band_solve's forward step
        if (fwd && wave == 0 && !ahead) {
            const int rr = lane & 31;
            double yv = yblk;
            for (int c = 0; c < jn; c++) {
                if (lane == c) yv *= dinv[c];
                const double yc = readlane_d(yv, c);
                if (lane > c && lane < jn) yv -= P[rr][c] * yc;
            }
            if (lane < jn) {
                y[j0 + lane] = yv;
                yb[par][lane] = yv;
            }
        }
        // 4a. the panel's columns of L back to the band (same mapping)
        for (int c = wave; c < jn; c += NW) {
            double *col = L + (j0 + c) * ld;
#pragma unroll
            for (int q = 0; q < RCH; q++) {
                const int r = 64 * q + lane;
                if (r >= c && r < rows && r - c <= b) col[r - c] = P[r][c];
            }
        }

        // 4b. trailing update on 16 x 16 tiles (I, K), K <= I, of the
        //     window on the FP64 matrix cores: rows ri = jn + 16 I + i,
        //     columns rk = jn + 16 K + j, D = A - P_I P_K^T as 8 steps of
        //     v_mfma_f64_16x16x4 over c = 0 .. 31 in order.  Lane l supplies
        //     operand row l % 16 at column 4 s + l / 16 and holds
        //     D[l / 16 + 4 v][l % 16], v = 0 .. 3 (band entries of column rk).
        // Look-ahead: the next panel's diagonal block is window rows and
        // columns jn .. jn + 31, i.e. tiles 0, 1, 2.  Wave 0 takes those
        // three tiles, keeps their values in P rows 0 .. 31 (free once the
        // write-back above is done) and factorises the block there while
        // the other waves update the rest of the window; the next step
        // then keeps those rows and skips its diagonal step.
        const int64_t j0n = j0 + jn;
        const int jnn = (int)(n - j0n < NF ? n - j0n : NF);  // next panel's width
        const bool la = jnn > 0;
        if (la || fwd) __syncthreads();
        // 4a'. the block's unknowns into the rows below it (one thread per
        //      row, terms in column order; entries outside the band are 0;
        //      P rows >= jn are not touched by 4b), by the waves that do not
        //      run the look-ahead
        if (fwd && ft0 >= 0) {
            for (int r = fr0 + ft0; r < rows; r += fnt) {
                asm volatile("" ::: "memory");  // yb read per row, not hoisted
                double acc = 0.;
#pragma unroll
                for (int c = 0; c < NF; c++) acc = fma(P[r][c], yb[par][c], acc);
                y[j0 + r] = (r == fr0 + ft0 ? yrow : y[j0 + r]) - acc;
            }
        }
        const int T = rows - jn;
        const int nt = (T + 15) >> 4;
        const int ntiles = nt * (nt + 1) / 2;
        const int li = lane & 15, lk = lane >> 4;
        double *Lp = L + j0 * ld;
        // each wave loads its next tile before it multiplies and stores the
        // current one (tiles are disjoint); two register sets, so the
        // prefetched tile is never copied, and a wave with no next tile
        // reloads its last one, so the load counter's order is the same on
        // every path and the multiply waits only for its own tile.  Without
        // look-ahead wave w takes walk positions w, w + NW, ...; with it
        // waves 1 .. NW - 1 take positions 0, 1, ... round robin.
        // the tiles of this loop, walked column by column (tile_cm)
        const int nwalk = la ? ntiles - 3 : ntiles;
        const int tstride = la ? NW - 1 : NW;
        int t = __builtin_amdgcn_readfirstlane(la ? wave - 1 : wave);
        if (!(la && wave == 0) && t < nwalk) {
            TileAt q0 = tile_cm(t, nt, la, jn, li, lk), q1;
            v4d a0 = tile_load(q0, Lp, ld, rows), a1;
            for (;;) {
                const int t1 = t + tstride;
                q1 = tile_cm(min(t1, nwalk - 1), nt, la, jn, li, lk);
                a1 = tile_load(q1, Lp, ld, rows);
                __builtin_amdgcn_sched_barrier(0);  // loads issue before the multiply
                a0 = tile_update<PR>(q0, a0, P, lk);
                tile_store(q0, a0, Lp, ld, rows);
                if (t1 >= nwalk) break;
                const int t2 = t1 + tstride;
                q0 = tile_cm(min(t2, nwalk - 1), nt, la, jn, li, lk);
                a0 = tile_load(q0, Lp, ld, rows);
                __builtin_amdgcn_sched_barrier(0);
                a1 = tile_update<PR>(q1, a1, P, lk);
                tile_store(q1, a1, Lp, ld, rows);
                if (t2 >= nwalk) break;
                t = t2;
            }
        }
        if (la && wave == 0) {
            // the next block from the band (entries outside this window are
            // untouched by this step; the window's are replaced below)
            const int rown = (int)(n - j0n < b + jnn ? n - j0n : b + jnn);
            const double *Ln = L + j0n * ld;
            int r = lane & 31, h = lane >> 5;
            asm volatile("" : "+v"(r), "+v"(h));
            double x[NF / 2];
#pragma unroll
            for (int q = 0; q < NF / 2; q++) {
                const int c = 2 * q + h;
                x[q] = (c < jnn && r >= c && r < rown && r - c <= b) ? Ln[c * ld + (r - c)] : 0.;
            }
#pragma unroll
            for (int q = 0; q < NF / 2; q++) P[r][2 * q + h] = x[q];
            // this step's update of the block: tiles 0 (0, 0), 1 (1, 0),
            // 2 (1, 1), also written back to the band
            const int nmine = ntiles < 3 ? ntiles : 3;
            for (int tt = 0; tt < nmine; tt++) {
                const TileAt q = tile_at(tt, jn, li, lk);
                v4d a = tile_load(q, Lp, ld, rows);
                a = tile_update<PR>(q, a, P, lk);
                tile_store(q, a, Lp, ld, rows);
#pragma unroll
                for (int v4 = 0; v4 < 4; v4++) {
                    const int rk = q.rk0 + 4 * v4;
                    if (rk < rows && q.ri < rows && q.ri >= rk) P[q.ri - jn][rk - jn] = a[v4];
                }
            }
            diag_block(P, dinv, Lk, jnn, lane, failed);
            if (fwd) {
                // the next block's rows: this block's terms (4a' for rows
                // jn .. jn + jnn - 1, from P rows jn .., untouched by the
                // look-ahead), then the substitution through the next
                // diagonal block now in P rows 0 .. 31 (3b of the next step)
                const int rr = lane & 31;
                double yv = 0.;
                if (lane < jnn) {
                    // (rows past this panel's band get no terms from it)
                    double acc = 0.;
                    if (jn + lane < rows) {
#pragma unroll
                        for (int c = 0; c < NF; c++) acc = fma(P[jn + rr][c], yb[par][c], acc);
                    }
                    yv = jn + lane < rows ? ynx - acc : ynx;
                }
                for (int c = 0; c < jnn; c++) {
                    if (lane == c) yv *= dinv[c];
                    const double yc = readlane_d(yv, c);
                    if (lane > c && lane < jnn) yv -= P[rr][c] * yc;
                }
                if (lane < jnn) {
                    y[j0n + lane] = yv;
                    yb[par ^ 1][lane] = yv;
                }
            }
        }
        __syncthreads();
        ahead = la;
    }
    if (failed && lane == 0) atomicAdd(bad, 1);
}

// --------------------------------------------------------- triangular
// x <- A^-1 f for system v: permute f into y, L y = g, L^T x = y, scatter
// (fwd_done: y = L^-1 g is already in work, from band_chol's fused pass).
// Blocks of NS = 32 columns: the block's diagonal triangle of L is staged
// in LDS by all threads, wave 0 substitutes through it (lane c holds
// unknown c and finishes it with its pivot's reciprocal; values travel by
// readlane), then all threads apply the block
// to the rows below it (forward) or gather the rows below into it
// (backward, one wave per column, fixed-order shuffle reduction).
constexpr int NS = 32;

template <int BT>
__global__ void __launch_bounds__(BT) band_solve(BandSys s, const int64_t *__restrict__ pnode,
                                                  const double *__restrict__ f,
                                                  double *__restrict__ x,
                                                  double *__restrict__ work, int fwd_done)
{
    constexpr int NW = BT / 64;
    __shared__ double T[NS][NS + 1];  // T[r][c] = L[j0 + r][j0 + c], r >= c (0 elsewhere)
    __shared__ double yb[NS];
    __shared__ double part[NS];
    // backward: the solved unknowns also in an LDS ring (x[i] at xr[i % RX]),
    // so the column sums read them from LDS; a slot is reused only after
    // every column that reads its row is done
    constexpr int RX = 1024;
    static_assert(RX >= MAXB + 2 * NS, "ring shorter than the rows a block reads");
    __shared__ double xr[RX];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t v = s.order ? (int64_t)s.order[blockIdx.x] : (int64_t)blockIdx.x;
    const int64_t base = s.dof0[v];
    const int64_t n = s.dof0[v + 1] - base;
    const int b = s.bw[v];
    const double *L = s.band + s.boff[v];
    const int64_t ld = b + 1;
    double *y = work + base;  // local vector in permuted numbering
    // (fwd_done: work already holds the forward result for every system,
    // from band_chol's fused substitution)
    if (!fwd_done) {
        // gather: y[6 p + d] = f[6 pnode[p] + d]
        for (int64_t t = tid; t < n; t += BT) {
            const int64_t p = t / 6;
            y[t] = f[6 * pnode[base / 6 + p] + (t - 6 * p)];
        }
    }
    // (loads unconditional, addresses clamped into the block's columns)
    auto stage = [&](int64_t j0, int jn) {
        static_assert((NS * NS) % BT == 0, "the stage covers the block in whole passes");
        constexpr int PER = NS * NS / BT;
        const double *Lb = L + j0 * ld;
        double sv[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int t = tid + q * BT, r = t / NS, c = t % NS;
            const bool in = c <= r && r < jn && r - c <= b;
            sv[q] = Lb[(c < jn ? c : 0) * (int)ld + (in ? r - c : 0)];
        }
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int t = tid + q * BT, r = t / NS, c = t % NS;
            T[r][c] = (c <= r && r < jn && r - c <= b) ? sv[q] : 0.;
        }
    };
    __syncthreads();
    if (!fwd_done) {
        // forward: L y = g
        for (int64_t j0 = 0; j0 < n; j0 += NS) {
            const int jn = (int)(n - j0 < NS ? n - j0 : NS);
            // the block's right-hand side, read with the stage's loads
            const double yt = y[j0 + (lane < jn ? lane : jn - 1)];
            stage(j0, jn);
            __syncthreads();
            if (wave == 0) {
                // lane c finishes unknown c with the reciprocal of its pivot
                // (formed by every lane at once), readlane hands it to all
                const double rd = lane < jn ? 1. / T[lane & 31][lane & 31] : 0.;
                double yc = lane < jn ? yt : 0.;
                for (int c = 0; c < jn; c++) {
                    if (lane == c) yc *= rd;
                    const double yv = readlane_d(yc, c);
                    if (lane > c && lane < jn) yc -= T[lane][c] * yv;
                }
                if (lane < jn) {
                    y[j0 + lane] = yc;
                    yb[lane] = yc;
                }
            }
            __syncthreads();
            // rows below the block within the band
            const int64_t rend = (j0 + jn - 1 + b < n - 1) ? j0 + jn - 1 + b : n - 1;
            // (branch-free: addresses clamped into the block's columns, masked
            // terms enter as fma(l, 0, acc) = acc; eight loads in flight per
            // round trip keep the registers of four workgroups per CU)
            const double *Lb = L + j0 * ld;
            for (int64_t i = j0 + jn + tid; i <= rend; i += BT) {
                const int di = (int)(i - j0);
                double acc = 0.;
    #pragma unroll 8
                for (int c = 0; c < NS; c++) {
                    const int cc = c < jn ? c : jn - 1;
                    const double l = *(const double *)((const char *)Lb +
                                                       (unsigned)(cc * (int)ld + min(di - cc, b)) * 8u);
                    acc = fma(l, (c < jn && di - c <= b) ? yb[c] : 0., acc);
                }
                y[i] -= acc;
            }
            __syncthreads();
        }
    }
    // backward: L^T x = y, blocks from the end
    const int64_t nblk = (n + NS - 1) / NS;
    for (int64_t bk = nblk - 1; bk >= 0; bk--) {
        const int64_t j0 = bk * NS;
        const int jn = (int)(n - j0 < NS ? n - j0 : NS);
        // the block's right-hand side (its forward result), read up front
        const double yt = y[j0 + (lane < jn ? lane : jn - 1)];
        // contributions of the solved rows below: s_c = sum_i L[i][j0+c] y[i]
        // for i in [j0+jn, j0+c+b]
        for (int c = wave; c < NS; c += NW) {
            double sacc = 0.;
            if (c < jn) {
                const int64_t j = j0 + c;
                const int64_t iend = (j + b < n - 1) ? j + b : n - 1;
                const double *Lj = L + j * ld;
                // two 64-row chunks in flight, addresses clamped, terms masked
                for (int64_t ib = j0 + jn; ib <= iend; ib += 128) {
                    const int64_t i1 = ib + lane, i2 = ib + 64 + lane;
                    const int64_t c1 = i1 <= iend ? i1 : iend, c2 = i2 <= iend ? i2 : iend;
                    const double l1 = Lj[(int)(c1 - j)], l2 = Lj[(int)(c2 - j)];
                    const double y1 = xr[c1 & (RX - 1)], y2 = xr[c2 & (RX - 1)];
                    sacc = fma(l1, i1 <= iend ? y1 : 0., sacc);
                    sacc = fma(l2, i2 <= iend ? y2 : 0., sacc);
                }
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) sacc += __shfl_xor(sacc, o, 64);
            if (lane == 0) part[c] = sacc;
        }
        stage(j0, jn);
        __syncthreads();
        if (wave == 0) {
            // back substitution in the diagonal triangle, lane c holds x[j0+c]
            const double rd = lane < jn ? 1. / T[lane & 31][lane & 31] : 0.;
            double rhs = lane < jn ? yt - part[lane & 31] : 0.;
            for (int cc = jn - 1; cc >= 0; cc--) {
                if (lane == cc) rhs *= rd;
                const double xv = readlane_d(rhs, cc);
                if (lane < cc) rhs -= T[cc][lane] * xv;
            }
            if (lane < jn) {
                y[j0 + lane] = rhs;
                xr[(j0 + lane) & (RX - 1)] = rhs;
            }
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
                        FwdArgs fw, int *bad)
{
#define PF3D_CHOL(BTV, MBV)                                                                     \
    if (threads == BTV && max_b == MBV) {                                                       \
        hipLaunchKernelGGL((band_chol<BTV, MBV>), dim3(grid), dim3(BTV), 0, s, sys, fw, bad); \
        return true;                                                                            \
    }
    PF3D_CHOL(1024, 576) PF3D_CHOL(1024, 448) PF3D_CHOL(1024, 256)
    PF3D_CHOL(512, 576) PF3D_CHOL(512, 448) PF3D_CHOL(512, 256)
    PF3D_CHOL(256, 576) PF3D_CHOL(256, 448) PF3D_CHOL(256, 256)
#undef PF3D_CHOL
    return false;
}

int pf3d_band_load(int64_t n_nodes, const int64_t *dof0, const int32_t *bw, const int64_t *boff,
                   int64_t band_size, double *band, const int32_t *row_ptr,
                   const int32_t *col_idx, const double *val, const int32_t *var,
                   const int64_t *pinv, void *stream)
{
    if (n_nodes < 0 || band_size < 0) return pf3d_fail(PF3D_E_ARG, "pf3d_band_load: negative size");
    if (n_nodes == 0) return PF3D_OK;
    if (!dof0 || !bw || !boff || !band || !row_ptr || !col_idx || !val || !var || !pinv)
        return pf3d_fail(PF3D_E_ARG, "pf3d_band_load: null argument");
    hipStream_t s = (hipStream_t)stream;
    const unsigned zb = nb256(band_size) < 1u << 20 ? nb256(band_size) : 1u << 20;
    hipLaunchKernelGGL(band_zero, dim3(zb), dim3(256), 0, s, band_size, band);
    hipLaunchKernelGGL(band_fill, dim3(nb256(n_nodes * 36)), dim3(256), 0, s, n_nodes, row_ptr,
                       col_idx, val, var, pinv, dof0, bw, boff, band);
    return pf3d_hip_check(hipGetLastError(), "pf3d_band_load");
}

int pf3d_band_chol(int64_t n_sys, const int32_t *order, const int64_t *dof0, const int32_t *bw,
                   const int64_t *boff, double *band, int *bad, int threads, int max_b,
                   const int64_t *pnode, const double *f, double *work, void *stream)
{
    if (threads == 0) threads = 1024;
    if (max_b == 0) max_b = MAXB;
    if ((threads != 256 && threads != 512 && threads != 1024) ||
        (max_b != 256 && max_b != 448 && max_b != 576))
        return pf3d_fail(PF3D_E_ARG, "pf3d_band_chol: threads must be 256/512/1024 and "
                                     "max_b 256/448/576 (got %d, %d)", threads, max_b);
    if (n_sys < 0) return pf3d_fail(PF3D_E_ARG, "pf3d_band_chol: negative size");
    if (n_sys == 0) return PF3D_OK;
    if (!dof0 || !bw || !boff || !band || !bad)
        return pf3d_fail(PF3D_E_ARG, "pf3d_band_chol: null argument");
    if (work && (!pnode || !f))
        return pf3d_fail(PF3D_E_ARG, "pf3d_band_chol: the forward substitution needs pnode and f");
    if (n_sys > 2147483647) return pf3d_fail(PF3D_E_ARG, "too many systems");
    BandSys sys{n_sys, order, dof0, bw, boff, band};
    FwdArgs fw{pnode, f, work};
    launch_chol(threads, max_b, (unsigned)n_sys, (hipStream_t)stream, sys, fw, bad);
    return pf3d_hip_check(hipGetLastError(), "pf3d_band_chol");
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
        (max_b != 256 && max_b != 448 && max_b != 576))
        return pf3d_fail(PF3D_E_ARG, "pf3d_band_factor: threads must be 256/512/1024 and "
                                     "max_b 256/448/576 (got %d, %d)", threads, max_b);
    if (n_var < 0 || n_nodes < 0) return pf3d_fail(PF3D_E_ARG, "pf3d_band_factor: negative size");
    if (n_var == 0) return PF3D_OK;
    if (!dof0 || !bw || !boff || !band || !row_ptr || !col_idx || !val || !var || !pinv || !bad)
        return pf3d_fail(PF3D_E_ARG, "pf3d_band_factor: null argument");
    if (n_var > 2147483647) return pf3d_fail(PF3D_E_ARG, "too many systems");
    int rc = pf3d_band_load(n_nodes, dof0, bw, boff, band_size, band, row_ptr, col_idx, val, var,
                            pinv, stream);
    if (rc != PF3D_OK) return rc;
    return pf3d_band_chol(n_var, order, dof0, bw, boff, band, bad, threads, max_b, nullptr,
                          nullptr, nullptr, stream);
}

// band_solve<BT> over n_var systems; fwd_done: work holds the forward result
static int launch_solve(const char *who, int64_t n_var, const int32_t *order, const int64_t *dof0,
                        const int32_t *bw, const int64_t *boff, const double *band,
                        const int64_t *pnode, const double *f, double *x, double *work,
                        int threads, int fwd_done, void *stream)
{
    if (threads == 0) threads = 1024;
    BandSys sys{n_var, order, dof0, bw, boff, const_cast<double *>(band)};
    const dim3 grid((unsigned)n_var);
    hipStream_t s = (hipStream_t)stream;
    if (threads == 1024)
        hipLaunchKernelGGL(band_solve<1024>, grid, dim3(1024), 0, s, sys, pnode, f, x, work,
                           fwd_done);
    else if (threads == 512)
        hipLaunchKernelGGL(band_solve<512>, grid, dim3(512), 0, s, sys, pnode, f, x, work,
                           fwd_done);
    else if (threads == 256)
        hipLaunchKernelGGL(band_solve<256>, grid, dim3(256), 0, s, sys, pnode, f, x, work,
                           fwd_done);
    else
        return pf3d_fail(PF3D_E_ARG, "%s: threads must be 256, 512 or 1024", who);
    return pf3d_hip_check(hipGetLastError(), who);
}

int pf3d_band_solve(int64_t n_var, const int32_t *order, const int64_t *dof0, const int32_t *bw,
                    const int64_t *boff, const double *band, const int64_t *pnode,
                    const double *f, double *x, double *work, int threads, void *stream)
{
    if (n_var < 0) return pf3d_fail(PF3D_E_ARG, "pf3d_band_solve: negative size");
    if (n_var == 0) return PF3D_OK;
    if (!dof0 || !bw || !boff || !band || !pnode || !f || !x || !work)
        return pf3d_fail(PF3D_E_ARG, "pf3d_band_solve: null argument");
    return launch_solve("pf3d_band_solve", n_var, order, dof0, bw, boff, band, pnode, f, x, work,
                        threads, 0, stream);
}

int pf3d_band_backsolve(int64_t n_var, const int32_t *order, const int64_t *dof0,
                        const int32_t *bw, const int64_t *boff, const double *band,
                        const int64_t *pnode, double *work, double *x, int threads, void *stream)
{
    if (n_var < 0) return pf3d_fail(PF3D_E_ARG, "pf3d_band_backsolve: negative size");
    if (n_var == 0) return PF3D_OK;
    if (!dof0 || !bw || !boff || !band || !pnode || !x || !work)
        return pf3d_fail(PF3D_E_ARG, "pf3d_band_backsolve: null argument");
    return launch_solve("pf3d_band_backsolve", n_var, order, dof0, bw, boff, band, pnode,
                        nullptr, x, work, threads, 1, stream);
}

}  // extern "C"

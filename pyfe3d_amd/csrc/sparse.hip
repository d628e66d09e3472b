// COO -> 6x6 block-sparse-row (BSR) conversion with a deterministic
// segmented sum, replacing scipy's coo_matrix((v, (r, c))).tocsc() on the
// reference's caller side (e.g. tests/test_quad4_static_point_load.py:80).
//
// Plan (once per mesh): sort the slots by (block row, block col, row%6,
// col%6), keep the slot permutation, the start of each run of equal
// (row, col) and the BSR position of each run.  Apply (every re-assembly):
// each output entry sums its run of slots in slot order, so results are
// bit-identical from run to run and independent of the launch shape.
#include <hipcub/hipcub.hpp>

#include "../../include/pf3d.h"
#include "pf3d_device.hpp"

int pf3d_fail(int code, const char *fmt, ...);
int pf3d_hip_check(hipError_t e, const char *where);

namespace {

struct PlanWork {
    // offsets (bytes) into the caller's workspace
    size_t keys_in, keys_out, ids_in, ids_out, uhead, bhead, temp, total;
    size_t temp_bytes;
};

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

int end_bit_for(int64_t n_nodes)
{
    // keys < n_nodes^2 * 36
    double maxkey = (double)n_nodes * (double)n_nodes * 36.0;
    int b = 1;
    while (b < 63 && (double)(1ULL << b) < maxkey) b++;
    return b;
}

hipError_t layout(int64_t n, int64_t n_nodes, PlanWork &w)
{
    size_t a8 = align256(sizeof(int64_t) * (size_t)n);
    size_t sort_bytes = 0, scan_bytes = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(
        nullptr, sort_bytes, (const int64_t *)nullptr, (int64_t *)nullptr,
        (const int64_t *)nullptr, (int64_t *)nullptr, (int)0, 0, end_bit_for(n_nodes));
    if (e != hipSuccess) return e;
    e = hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (int64_t *)nullptr,
                                         (int64_t *)nullptr, (int)0);
    if (e != hipSuccess) return e;
    // the size queries above do not depend on n for these primitives'
    // upper bounds at n = 0; query again with the real size
    sort_bytes = 0;
    scan_bytes = 0;
    e = hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const int64_t *)nullptr,
                                           (int64_t *)nullptr, (const int64_t *)nullptr,
                                           (int64_t *)nullptr, n, 0, end_bit_for(n_nodes));
    if (e != hipSuccess) return e;
    e = hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (int64_t *)nullptr,
                                         (int64_t *)nullptr, n);
    if (e != hipSuccess) return e;
    w.temp_bytes = align256(sort_bytes > scan_bytes ? sort_bytes : scan_bytes);
    size_t o = 0;
    w.keys_in = o; o += a8;
    w.keys_out = o; o += a8;
    w.ids_in = o; o += a8;
    w.ids_out = o; o += a8;
    w.uhead = o; o += a8;
    w.bhead = o; o += a8;
    w.temp = o; o += w.temp_bytes;
    w.total = o;
    return hipSuccess;
}

__global__ void make_keys(int64_t n, const int64_t *__restrict__ r, const int64_t *__restrict__ c,
                          int64_t n_nodes, int64_t *__restrict__ keys, int64_t *__restrict__ ids)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t ri = r[t], ci = c[t];
    keys[t] = ((ri / 6) * n_nodes + ci / 6) * 36 + (ri % 6) * 6 + ci % 6;
    ids[t] = t;
}

__global__ void mark_heads(int64_t n, const int64_t *__restrict__ keys, int64_t *__restrict__ uh,
                           int64_t *__restrict__ bh)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t k = keys[t];
    const int64_t kp = t ? keys[t - 1] : -1;
    uh[t] = (t == 0 || k != kp) ? 1 : 0;
    bh[t] = (t == 0 || k / 36 != kp / 36) ? 1 : 0;
}

// after the exclusive scans uh/bh hold run and block ids; heads are
// recognised from the keys again
__global__ void emit(int64_t n, int64_t n_nodes, const int64_t *__restrict__ keys,
                     const int64_t *__restrict__ uid, const int64_t *__restrict__ bid,
                     const int64_t *__restrict__ bflag, int64_t *__restrict__ seg, int64_t *__restrict__ dest,
                     int32_t *__restrict__ col_idx, int32_t *__restrict__ row_ptr,
                     int64_t n_unique, int64_t nnzb)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t k = keys[t];
    const int64_t kp = t ? keys[t - 1] : -1;
    // block of slot t: exclusive count of block heads before t, plus one if
    // t starts a block, minus one
    const int64_t b = bid[t] + bflag[t] - 1;
    if (t == 0 || k != kp) {
        seg[uid[t]] = t;
        dest[uid[t]] = b * 36 + k % 36;
    }
    if (t == 0 || k / 36 != kp / 36) {
        const int64_t bk = k / 36;
        const int64_t row = bk / n_nodes;
        col_idx[b] = (int32_t)(bk % n_nodes);
        const int64_t prow = t ? (kp / 36) / n_nodes : -1;
        for (int64_t rr = prow + 1; rr <= row; rr++) row_ptr[rr] = (int32_t)b;
        if (b == nnzb - 1)
            for (int64_t rr = row + 1; rr <= n_nodes; rr++) row_ptr[rr] = (int32_t)nnzb;
    }
    if (t == n - 1) seg[n_unique] = n;
}

__global__ void seg_sum(int64_t n_unique, const int64_t *__restrict__ perm,
                        const int64_t *__restrict__ seg, const int64_t *__restrict__ dest,
                        const double *__restrict__ v, double *__restrict__ out)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_unique) return;
    const int64_t a = seg[j], b = seg[j + 1];
    double s = 0.;
    for (int64_t t = a; t < b; t++) s += v[perm[t]];
    out[dest[j]] = s;
}

__global__ void fill_zero(int64_t n, double *__restrict__ p)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0.;
}

// ---- node incidence plan for the internal-force gather -----------------
// keys: node position of each (element, node) entry; entries whose DOF
// offset is not 6*node with node < n_nodes go to a bucket past the last node
__global__ void incidence_keys(int64_t n, const int64_t *__restrict__ conn, int64_t n_nodes,
                               int64_t *__restrict__ keys, int64_t *__restrict__ ids)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t c = conn[t];
    keys[t] = (c < 0 || c % 6 != 0 || c / 6 >= n_nodes) ? n_nodes : c / 6;
    ids[t] = t;
}

// offsets[p] = first sorted entry whose node >= p (binary search)
__global__ void incidence_offsets(int64_t n, const int64_t *__restrict__ keys, int64_t n_nodes,
                                  int64_t *__restrict__ offsets)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p > n_nodes) return;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (keys[mid] < p) lo = mid + 1;
        else hi = mid;
    }
    offsets[p] = lo;
}

// fint[6p + d] += sum of fe[id*6 + d] over the entries of node p, in
// ascending entry (= element) order
__global__ void gather_nodes(int64_t n_nodes, const int64_t *__restrict__ offsets,
                             const int64_t *__restrict__ ids, const double *__restrict__ fe,
                             double *__restrict__ fint)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 6 * n_nodes) return;
    const int64_t p = t / 6, d = t - 6 * p;
    double s = fint[t];
    for (int64_t k = offsets[p]; k < offsets[p + 1]; k++) s += fe[ids[k] * 6 + d];
    fint[t] = s;
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

extern "C" {

int pf3d_bsr_plan_bytes(int64_t n_slots, int64_t n_nodes, int64_t *bytes)
{
    if (n_slots <= 0 || n_nodes <= 0 || !bytes)
        return pf3d_fail(PF3D_E_ARG, "pf3d_bsr_plan_bytes: bad arguments");
    if (n_nodes > 2147483647) return pf3d_fail(PF3D_E_ARG, "too many nodes for int32 BSR");
    PlanWork w;
    hipError_t e = layout(n_slots, n_nodes, w);
    if (e != hipSuccess) return pf3d_hip_check(e, "pf3d_bsr_plan_bytes");
    *bytes = (int64_t)w.total;
    return PF3D_OK;
}

int pf3d_bsr_plan_sort(int64_t n_slots, const int64_t *r, const int64_t *c, int64_t n_nodes,
                       void *work, int64_t work_bytes, int64_t *counts, void *stream)
{
    if (n_slots <= 0 || !r || !c || !work || !counts)
        return pf3d_fail(PF3D_E_ARG, "pf3d_bsr_plan_sort: bad arguments");
    PlanWork w;
    hipError_t e = layout(n_slots, n_nodes, w);
    if (e != hipSuccess) return pf3d_hip_check(e, "pf3d_bsr_plan_sort");
    if ((int64_t)w.total > work_bytes)
        return pf3d_fail(PF3D_E_ARG, "workspace too small: %lld < %lld", (long long)work_bytes,
                         (long long)w.total);
    hipStream_t s = (hipStream_t)stream;
    char *base = (char *)work;
    int64_t *kin = (int64_t *)(base + w.keys_in), *kout = (int64_t *)(base + w.keys_out);
    int64_t *iin = (int64_t *)(base + w.ids_in), *iout = (int64_t *)(base + w.ids_out);
    int64_t *uh = (int64_t *)(base + w.uhead), *bh = (int64_t *)(base + w.bhead);
    void *tmp = base + w.temp;
    hipLaunchKernelGGL(make_keys, dim3(nblk(n_slots, 256)), dim3(256), 0, s, n_slots, r, c,
                       n_nodes, kin, iin);
    size_t tb = w.temp_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(tmp, tb, kin, kout, iin, iout, n_slots, 0,
                                           end_bit_for(n_nodes), s);
    if (e != hipSuccess) return pf3d_hip_check(e, "radix sort");
    hipLaunchKernelGGL(mark_heads, dim3(nblk(n_slots, 256)), dim3(256), 0, s, n_slots, kout, uh,
                       bh);
    // last flags are needed for the totals: keep copies before the scans
    int64_t last[2];
    e = hipMemcpyAsync(&last[0], uh + n_slots - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(&last[1], bh + n_slots - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return pf3d_hip_check(e, "copy flags");
    // run and block ids (exclusive scans, out of place into the input
    // buffers of the sort, which are free now)
    tb = w.temp_bytes;
    e = hipcub::DeviceScan::ExclusiveSum(tmp, tb, uh, kin, n_slots, s);
    if (e != hipSuccess) return pf3d_hip_check(e, "scan");
    tb = w.temp_bytes;
    e = hipcub::DeviceScan::ExclusiveSum(tmp, tb, bh, iin, n_slots, s);
    if (e != hipSuccess) return pf3d_hip_check(e, "scan");
    int64_t tot[2];
    e = hipMemcpyAsync(&tot[0], kin + n_slots - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(&tot[1], iin + n_slots - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return pf3d_hip_check(e, "plan totals");
    counts[0] = tot[0] + last[0];  // unique (row, col) entries
    counts[1] = tot[1] + last[1];  // 6x6 blocks
    if (counts[1] > 2147483647) return pf3d_fail(PF3D_E_ARG, "too many blocks for int32 BSR");
    return PF3D_OK;
}

int pf3d_bsr_plan_emit(int64_t n_slots, int64_t n_nodes, void *work, int64_t work_bytes,
                       int64_t n_unique, int64_t nnzb, int64_t *perm, int64_t *seg,
                       int64_t *dest, int32_t *row_ptr, int32_t *col_idx, void *stream)
{
    if (n_slots <= 0 || !work || !perm || !seg || !dest || !row_ptr || !col_idx)
        return pf3d_fail(PF3D_E_ARG, "pf3d_bsr_plan_emit: bad arguments");
    PlanWork w;
    hipError_t e = layout(n_slots, n_nodes, w);
    if (e != hipSuccess) return pf3d_hip_check(e, "pf3d_bsr_plan_emit");
    if ((int64_t)w.total > work_bytes) return pf3d_fail(PF3D_E_ARG, "workspace too small");
    hipStream_t s = (hipStream_t)stream;
    char *base = (char *)work;
    const int64_t *kout = (const int64_t *)(base + w.keys_out);
    const int64_t *iout = (const int64_t *)(base + w.ids_out);
    const int64_t *uid = (const int64_t *)(base + w.keys_in), *bid = (const int64_t *)(base + w.ids_in);
    const int64_t *bflag = (const int64_t *)(base + w.bhead);
    e = hipMemcpyAsync(perm, iout, sizeof(int64_t) * n_slots, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return pf3d_hip_check(e, "copy perm");
    hipLaunchKernelGGL(emit, dim3(nblk(n_slots, 256)), dim3(256), 0, s, n_slots, n_nodes, kout,
                       uid, bid, bflag, seg, dest, col_idx, row_ptr, n_unique, nnzb);
    return pf3d_hip_check(hipGetLastError(), "pf3d_bsr_plan_emit");
}

int pf3d_bsr_apply(int64_t n_unique, const int64_t *perm, const int64_t *seg, const int64_t *dest,
                   const double *v, int64_t nnzb, double *bsr_val, void *stream)
{
    if (n_unique <= 0 || !perm || !seg || !dest || !v || !bsr_val)
        return pf3d_fail(PF3D_E_ARG, "pf3d_bsr_apply: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(fill_zero, dim3(nblk(nnzb * 36, 256)), dim3(256), 0, s, nnzb * 36,
                       bsr_val);
    hipLaunchKernelGGL(seg_sum, dim3(nblk(n_unique, 256)), dim3(256), 0, s, n_unique, perm, seg,
                       dest, v, bsr_val);
    return pf3d_hip_check(hipGetLastError(), "pf3d_bsr_apply");
}

}  // extern "C"

// ---- internal-force gather ---------------------------------------------
namespace {
struct IncWork {
    size_t keys_in, keys_out, ids_in, temp, total, temp_bytes;
};

// radix bits for keys in [0, n_nodes]
int inc_bits(int64_t n_nodes)
{
    int b = 1;
    while (b < 63 && (int64_t(1) << b) <= n_nodes) b++;
    return b;
}

hipError_t inc_layout(int64_t n, int64_t n_nodes, IncWork &w)
{
    size_t a8 = align256(sizeof(int64_t) * (size_t)n);
    size_t sort_bytes = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(
        nullptr, sort_bytes, (const int64_t *)nullptr, (int64_t *)nullptr,
        (const int64_t *)nullptr, (int64_t *)nullptr, n, 0, inc_bits(n_nodes));
    if (e != hipSuccess) return e;
    w.temp_bytes = align256(sort_bytes);
    size_t o = 0;
    w.keys_in = o; o += a8;
    w.keys_out = o; o += a8;
    w.ids_in = o; o += a8;
    w.temp = o; o += w.temp_bytes;
    w.total = o;
    return hipSuccess;
}
}  // namespace

extern "C" {

int pf3d_incidence_bytes(int64_t n_entries, int64_t n_nodes, int64_t *bytes)
{
    if (n_entries <= 0 || n_nodes <= 0 || !bytes)
        return pf3d_fail(PF3D_E_ARG, "pf3d_incidence_bytes: bad arguments");
    IncWork w;
    hipError_t e = inc_layout(n_entries, n_nodes, w);
    if (e != hipSuccess) return pf3d_hip_check(e, "pf3d_incidence_bytes");
    *bytes = (int64_t)w.total;
    return PF3D_OK;
}

int pf3d_incidence_plan(int64_t n_entries, const int64_t *conn, int64_t n_nodes, void *work,
                        int64_t work_bytes, int64_t *offsets, int64_t *ids, void *stream)
{
    if (n_entries <= 0 || n_nodes <= 0 || !conn || !work || !offsets || !ids)
        return pf3d_fail(PF3D_E_ARG, "pf3d_incidence_plan: bad arguments");
    IncWork w;
    hipError_t e = inc_layout(n_entries, n_nodes, w);
    if (e != hipSuccess) return pf3d_hip_check(e, "pf3d_incidence_plan");
    if ((int64_t)w.total > work_bytes)
        return pf3d_fail(PF3D_E_ARG, "workspace too small: %lld < %lld", (long long)work_bytes,
                         (long long)w.total);
    hipStream_t s = (hipStream_t)stream;
    char *base = (char *)work;
    int64_t *kin = (int64_t *)(base + w.keys_in), *kout = (int64_t *)(base + w.keys_out);
    int64_t *iin = (int64_t *)(base + w.ids_in);
    hipLaunchKernelGGL(incidence_keys, dim3(nblk(n_entries, 256)), dim3(256), 0, s, n_entries,
                       conn, n_nodes, kin, iin);
    size_t tb = w.temp_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(base + w.temp, tb, kin, kout, iin, ids, n_entries, 0,
                                           inc_bits(n_nodes), s);
    if (e != hipSuccess) return pf3d_hip_check(e, "radix sort");
    hipLaunchKernelGGL(incidence_offsets, dim3(nblk(n_nodes + 1, 256)), dim3(256), 0, s,
                       n_entries, kout, n_nodes, offsets);
    return pf3d_hip_check(hipGetLastError(), "pf3d_incidence_plan");
}

int pf3d_gather_fint(int64_t n_nodes, const int64_t *offsets, const int64_t *ids,
                     const double *fe, double *fint, void *stream)
{
    if (n_nodes <= 0 || !offsets || !ids || !fe || !fint)
        return pf3d_fail(PF3D_E_ARG, "pf3d_gather_fint: bad arguments");
    hipLaunchKernelGGL(gather_nodes, dim3(nblk(6 * n_nodes, 256)), dim3(256), 0,
                       (hipStream_t)stream, n_nodes, offsets, ids, fe, fint);
    return pf3d_hip_check(hipGetLastError(), "pf3d_gather_fint");
}

}  // extern "C"

// ---- element-block plan --------------------------------------------------
// For COO arrays filled by the element kernels the slots of element e come
// in 6x6 node blocks: block (I, J) starts at init_k[e] + 36*nn*I + 6*J with
// row stride 6*nn (quad4.pyx:1287-1293; tria3r.pyx and beamc.pyx use the
// same row-major order).  A plan over these blocks needs one int64 per
// element block instead of one per slot, and its apply reads each block
// row as 6 contiguous values.  Block sums run over the sources in
// ascending (e, I, J), i.e. ascending slot order, like pf3d_bsr_apply.
namespace {

struct BlkWork {
    size_t keys_in, keys_out, ids_in, ids_out, head, temp, total, temp_bytes;
};

hipError_t blk_layout(int64_t n, int64_t n_nodes, BlkWork &w)
{
    const size_t a8 = align256(sizeof(int64_t) * (size_t)n);
    size_t sort_bytes = 0, scan_bytes = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(
        nullptr, sort_bytes, (const int64_t *)nullptr, (int64_t *)nullptr,
        (const int64_t *)nullptr, (int64_t *)nullptr, n, 0, 63);
    if (e != hipSuccess) return e;
    e = hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (int64_t *)nullptr,
                                         (int64_t *)nullptr, n);
    if (e != hipSuccess) return e;
    w.temp_bytes = align256(sort_bytes > scan_bytes ? sort_bytes : scan_bytes);
    size_t o = 0;
    w.keys_in = o; o += a8;
    w.keys_out = o; o += a8;
    w.ids_in = o; o += a8;
    w.ids_out = o; o += a8;
    w.head = o; o += a8;
    w.temp = o; o += w.temp_bytes;
    w.total = o;
    return hipSuccess;
}

// block key of entry id = e*nn*nn + I*nn + J (ids_in[t], or t)
__global__ void blk_keys(int64_t n_elem, int nn, const int64_t *__restrict__ conn,
                         int64_t n_nodes, const int64_t *__restrict__ ids_in,
                         int64_t *__restrict__ keys, int64_t *__restrict__ ids)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nb = (int64_t)nn * nn;
    if (t >= n_elem * nb) return;
    const int64_t id = ids_in ? ids_in[t] : t;
    const int64_t e = id / nb;
    const int ij = (int)(id - e * nb), I = ij / nn, J = ij - (ij / nn) * nn;
    const int64_t a = conn[e * nn + I] / 6, b = conn[e * nn + J] / 6;
    keys[t] = a * n_nodes + b;
    ids[t] = id;
}

// first slot of each element block (to order the sources by slot)
__global__ void blk_base(int64_t n_elem, int nn, const int64_t *__restrict__ init_k,
                         int64_t *__restrict__ keys, int64_t *__restrict__ ids)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nb = (int64_t)nn * nn;
    if (t >= n_elem * nb) return;
    const int64_t e = t / nb;
    const int ij = (int)(t - e * nb), I = ij / nn, J = ij - (ij / nn) * nn;
    keys[t] = init_k[e] + 36 * nn * I + 6 * J;
    ids[t] = t;
}

__global__ void blk_heads(int64_t n, const int64_t *__restrict__ keys, int64_t *__restrict__ h)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    h[t] = (t == 0 || keys[t] != keys[t - 1]) ? 1 : 0;
}

// src[t] = first slot of the t-th sorted element block; heads give the
// block ranges, the BSR column and the row pointers
__global__ void blk_emit(int64_t n, int nn, int64_t n_nodes, const int64_t *__restrict__ keys,
                         const int64_t *__restrict__ ids, const int64_t *__restrict__ bid,
                         const int64_t *__restrict__ init_k, int64_t nnzb,
                         int64_t *__restrict__ src, int64_t *__restrict__ src_ptr,
                         int32_t *__restrict__ row_ptr, int32_t *__restrict__ col_idx)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t nb = (int64_t)nn * nn;
    const int64_t id = ids[t], e = id / nb;
    const int ij = (int)(id - e * nb), I = ij / nn, J = ij - (ij / nn) * nn;
    const int64_t k0 = init_k ? init_k[e] : 36 * nb * e;
    src[t] = k0 + 36 * nn * I + 6 * J;
    const int64_t k = keys[t];
    const int64_t kp = t ? keys[t - 1] : -1;
    if (t == 0 || k != kp) {
        const int64_t b = bid[t];
        src_ptr[b] = t;
        const int64_t row = k / n_nodes;
        col_idx[b] = (int32_t)(k - row * n_nodes);
        const int64_t prow = t ? kp / n_nodes : -1;
        for (int64_t rr = prow + 1; rr <= row; rr++) row_ptr[rr] = (int32_t)b;
        if (b == nnzb - 1)
            for (int64_t rr = row + 1; rr <= n_nodes; rr++) row_ptr[rr] = (int32_t)nnzb;
    }
    if (t == n - 1) src_ptr[nnzb] = n;
}

// one thread per (BSR block, row m): 6 contiguous values per source
__global__ void blk_apply(int64_t nnzb, const int64_t *__restrict__ src_ptr,
                          const int64_t *__restrict__ src, int stride,
                          const double *__restrict__ v, double *__restrict__ out, int accumulate)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nnzb * 6) return;
    const int64_t b = t / 6;
    const int m = (int)(t - b * 6);
    double s[6];
    double *o = out + b * 36 + m * 6;
#pragma unroll
    for (int n = 0; n < 6; n++) s[n] = accumulate ? o[n] : 0.;
    for (int64_t k = src_ptr[b]; k < src_ptr[b + 1]; k++) {
        const double *p = v + src[k] + (int64_t)stride * m;
#pragma unroll
        for (int n = 0; n < 6; n++) s[n] += p[n];
    }
#pragma unroll
    for (int n = 0; n < 6; n++) o[n] = s[n];
}

}  // namespace

extern "C" {

int pf3d_block_plan_bytes(int64_t n_elem, int nn, int64_t n_nodes, int64_t *bytes)
{
    if (n_elem <= 0 || nn < 2 || nn > 4 || n_nodes <= 0 || !bytes)
        return pf3d_fail(PF3D_E_ARG, "pf3d_block_plan_bytes: bad arguments");
    if (n_nodes > 2147483647 || n_nodes > 3037000499LL)
        return pf3d_fail(PF3D_E_ARG, "too many nodes");
    BlkWork w;
    hipError_t e = blk_layout(n_elem * nn * nn, n_nodes, w);
    if (e != hipSuccess) return pf3d_hip_check(e, "pf3d_block_plan_bytes");
    *bytes = (int64_t)w.total;
    return PF3D_OK;
}

int pf3d_block_plan_sort(int64_t n_elem, int nn, const int64_t *conn, const int64_t *init_k,
                         int64_t n_nodes, void *work, int64_t work_bytes, int64_t *nnzb,
                         void *stream)
{
    if (n_elem <= 0 || nn < 2 || nn > 4 || !conn || !work || !nnzb)
        return pf3d_fail(PF3D_E_ARG, "pf3d_block_plan_sort: bad arguments");
    const int64_t n = n_elem * nn * nn;
    BlkWork w;
    hipError_t e = blk_layout(n, n_nodes, w);
    if (e != hipSuccess) return pf3d_hip_check(e, "pf3d_block_plan_sort");
    if ((int64_t)w.total > work_bytes)
        return pf3d_fail(PF3D_E_ARG, "workspace too small: %lld < %lld", (long long)work_bytes,
                         (long long)w.total);
    hipStream_t s = (hipStream_t)stream;
    char *base = (char *)work;
    int64_t *kin = (int64_t *)(base + w.keys_in), *kout = (int64_t *)(base + w.keys_out);
    int64_t *iin = (int64_t *)(base + w.ids_in), *iout = (int64_t *)(base + w.ids_out);
    int64_t *h = (int64_t *)(base + w.head);
    size_t tb = w.temp_bytes;
    const int64_t *first = nullptr;
    if (init_k) {
        // order the element blocks by slot first; the stable sort by block
        // key below keeps that order within a block
        hipLaunchKernelGGL(blk_base, dim3(nblk(n, 256)), dim3(256), 0, s, n_elem, nn, init_k,
                           kin, iin);
        e = hipcub::DeviceRadixSort::SortPairs(base + w.temp, tb, kin, kout, iin, iout, n, 0, 63,
                                               s);
        if (e != hipSuccess) return pf3d_hip_check(e, "radix sort");
        first = iout;
    }
    hipLaunchKernelGGL(blk_keys, dim3(nblk(n, 256)), dim3(256), 0, s, n_elem, nn, conn, n_nodes,
                       first, kin, iin);
    tb = w.temp_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(base + w.temp, tb, kin, kout, iin, iout, n, 0,
                                           inc_bits(n_nodes * n_nodes), s);
    if (e != hipSuccess) return pf3d_hip_check(e, "radix sort");
    hipLaunchKernelGGL(blk_heads, dim3(nblk(n, 256)), dim3(256), 0, s, n, kout, h);
    int64_t last = 0, tot = 0;
    e = hipMemcpyAsync(&last, h + n - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return pf3d_hip_check(e, "copy flag");
    // block ids: exclusive scan of the heads, into the sort's free input keys
    tb = w.temp_bytes;
    e = hipcub::DeviceScan::ExclusiveSum(base + w.temp, tb, h, kin, n, s);
    if (e != hipSuccess) return pf3d_hip_check(e, "scan");
    e = hipMemcpyAsync(&tot, kin + n - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return pf3d_hip_check(e, "plan totals");
    *nnzb = tot + last;
    if (*nnzb > 2147483647) return pf3d_fail(PF3D_E_ARG, "too many blocks for int32 BSR");
    return PF3D_OK;
}

int pf3d_block_plan_emit(int64_t n_elem, int nn, const int64_t *init_k, int64_t n_nodes,
                         void *work, int64_t work_bytes, int64_t nnzb, int64_t *src,
                         int64_t *src_ptr, int32_t *row_ptr, int32_t *col_idx, void *stream)
{
    if (n_elem <= 0 || nn < 2 || nn > 4 || nnzb <= 0 || !work || !src || !src_ptr || !row_ptr ||
        !col_idx)
        return pf3d_fail(PF3D_E_ARG, "pf3d_block_plan_emit: bad arguments");
    const int64_t n = n_elem * nn * nn;
    BlkWork w;
    hipError_t e = blk_layout(n, n_nodes, w);
    if (e != hipSuccess) return pf3d_hip_check(e, "pf3d_block_plan_emit");
    if ((int64_t)w.total > work_bytes) return pf3d_fail(PF3D_E_ARG, "workspace too small");
    char *base = (char *)work;
    hipLaunchKernelGGL(blk_emit, dim3(nblk(n, 256)), dim3(256), 0, (hipStream_t)stream, n, nn,
                       n_nodes, (const int64_t *)(base + w.keys_out),
                       (const int64_t *)(base + w.ids_out), (const int64_t *)(base + w.keys_in),
                       init_k, nnzb, src, src_ptr, row_ptr, col_idx);
    return pf3d_hip_check(hipGetLastError(), "pf3d_block_plan_emit");
}

int pf3d_block_apply(int64_t nnzb, const int64_t *src_ptr, const int64_t *src, int nn,
                     const double *v, double *bsr_val, int accumulate, void *stream)
{
    if (nnzb <= 0 || !src_ptr || !src || nn < 2 || nn > 4 || !v || !bsr_val)
        return pf3d_fail(PF3D_E_ARG, "pf3d_block_apply: bad arguments");
    hipLaunchKernelGGL(blk_apply, dim3(nblk(nnzb * 6, 256)), dim3(256), 0, (hipStream_t)stream,
                       nnzb, src_ptr, src, 6 * nn, v, bsr_val, accumulate);
    return pf3d_hip_check(hipGetLastError(), "pf3d_block_apply");
}

}  // extern "C"

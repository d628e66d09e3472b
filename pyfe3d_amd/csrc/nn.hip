// Kernels of the warm-start network: a small convolutional encoder-decoder
// that maps a plate's thickness field and pressure map to an initial guess
// for the PCG solve.  No reference counterpart (pyfe3d always starts its
// solves from zero); the network only changes the starting point of a
// solve whose result is still judged by the FP64 residual.
//
// Layout: activations NHWC in FP32, weights [C_out][9*C_in] (ky, kx, c_in
// fastest) in FP32 as the optimiser's master copy.  Every convolution is an
// im2col gather followed by one GEMM on the matrix cores: operands are
// rounded to bfloat16 when they are staged into LDS, products accumulate in
// FP32 (gfx950's v_mfma_f32_16x16x32_bf16: the whole 32-deep LDS tile in
// one instruction per 16x16 output tile).  The backward pass uses the same GEMM
// with other strides (dW = dY^T col, dcol = dY W) and a gather col2im, so
// every gradient is a fixed-order sum.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/pf3d.h"

int pf3d_fail(int code, const char *fmt, ...);
int pf3d_hip_check(hipError_t e, const char *where);

namespace {

using s16x8 = short __attribute__((ext_vector_type(8)));
using bf16x8 = __bf16 __attribute__((ext_vector_type(8)));
using f32x4 = float __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short to_bf16(float f)
{
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);  // inf/nan
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (unsigned short)(u >> 16);
}

// tile and LDS row (padded to 40 bf16 = 80 B: 16-byte aligned rows for the
// 8-value operand reads, rows 20 banks apart)
constexpr int TM = 64, TN = 64, TK = 32, KP = TK + 8;

struct Gemm {
    int M, N, K;
    const float *A;
    int64_t sam, sak;  // A(m, k) = A[m*sam + k*sak]
    const float *B;
    int64_t sbn, sbk;  // B(n, k) = B[n*sbn + k*sbk]
    float *C;
    int64_t ldc;       // C[m*ldc + n]
    int accumulate;    // C += result
    const float *bias; // [N] or null
    int relu;
    int kchunk;      // K range per grid.z slice (split-K), multiple of TK
    float *partial;  // split-K: raw sums [gridDim.z][M][N], else null
};

// C = A B^T (+ bias, ReLU).  256 threads: 4 waves in a 2x2 grid, each wave
// owns a 32x32 block of the 64x64 tile (2x2 MFMA tiles of 16x16).
__global__ void __launch_bounds__(256) gemm_bf16(Gemm g)
{
    __shared__ __align__(16) unsigned short As[TM][KP];
    __shared__ __align__(16) unsigned short Bs[TN][KP];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int kbeg = blockIdx.z * g.kchunk;
    const int kend = min(g.K, kbeg + g.kchunk);
    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        // stage 64x32 of A and of B (8 values per thread each); along k for
        // unit-stride k, along m/n otherwise, so global reads coalesce
        const bool ak = g.sak == 1, bk = g.sbk == 1;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int idx = tid + 256 * t;
            int r, kk;
            if (ak) { r = idx / TK; kk = idx % TK; }
            else { kk = idx / TM; r = idx % TM; }
            const int m = m0 + r, k = k0 + kk;
            float v = 0.f;
            if (m < g.M && k < kend) v = g.A[(int64_t)m * g.sam + (int64_t)k * g.sak];
            As[r][kk] = to_bf16(v);
            if (bk) { r = idx / TK; kk = idx % TK; }
            else { kk = idx / TN; r = idx % TN; }
            const int n = n0 + r, k2 = k0 + kk;
            float w = 0.f;
            if (n < g.N && k2 < kend) w = g.B[(int64_t)n * g.sbn + (int64_t)k2 * g.sbk];
            Bs[r][kk] = to_bf16(w);
        }
        __syncthreads();
        // operands: lane l supplies row l % 16 at k = 8 (l / 16) .. + 7 of
        // the tile, in A and in B alike, so each k meets its own partner
        {
            const int kq = 8 * (lane >> 4);
            s16x8 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                a[i] = *reinterpret_cast<const s16x8 *>(&As[wm + 16 * i + (lane & 15)][kq]);
                b[i] = *reinterpret_cast<const s16x8 *>(&Bs[wn + 16 * i + (lane & 15)][kq]);
            }
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(bf16x8, a[i]), __builtin_bit_cast(bf16x8, b[j]),
                        acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // D layout: lane holds rows 4*(lane/16) + q, column lane % 16
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + q;
                const int n = n0 + wn + 16 * j + (lane & 15);
                if (m >= g.M || n >= g.N) continue;
                float v = acc[i][j][q];
                if (g.partial) {
                    g.partial[((int64_t)blockIdx.z * g.M + m) * g.N + n] = v;
                    continue;
                }
                float *c = g.C + (int64_t)m * g.ldc + n;
                if (g.accumulate) v += *c;
                if (g.bias) v += g.bias[n];
                if (g.relu && v < 0.f) v = 0.f;
                *c = v;
            }
}

// C = sum over the split-K slices in slice order (+ C, bias, ReLU)
__global__ void splitk_reduce(Gemm g, int splits)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)g.M * g.N) return;
    const int m = (int)(t / g.N), n = (int)(t - (int64_t)m * g.N);
    float v = 0.f;
    for (int z = 0; z < splits; z++) v += g.partial[(int64_t)z * g.M * g.N + t];
    float *c = g.C + (int64_t)m * g.ldc + n;
    if (g.accumulate) v += *c;
    if (g.bias) v += g.bias[n];
    if (g.relu && v < 0.f) v = 0.f;
    *c = v;
}

// col[(b, oy, ox)][(ky, kx, c)] = x[b][oy*s + ky - 1][ox*s + kx - 1][c] (0 outside)
__global__ void im2col3(int B, int H, int W, int C, int s, const float *__restrict__ x,
                        float *__restrict__ col)
{
    const int Ho = (H - 1) / s + 1, Wo = (W - 1) / s + 1;
    const int64_t total = (int64_t)B * Ho * Wo * 9 * C;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int c = (int)(t % C);
    int64_t r = t / C;
    const int kx = (int)(r % 3); r /= 3;
    const int ky = (int)(r % 3); r /= 3;
    const int ox = (int)(r % Wo); r /= Wo;
    const int oy = (int)(r % Ho);
    const int b = (int)(r / Ho);
    const int y = oy * s + ky - 1, xx = ox * s + kx - 1;
    float v = 0.f;
    if (y >= 0 && y < H && xx >= 0 && xx < W) v = x[(((int64_t)b * H + y) * W + xx) * C + c];
    col[t] = v;
}

// dx[b][y][x][c] = sum over (ky, kx) of dcol at the output pixel that read it
__global__ void col2im3(int B, int H, int W, int C, int s, const float *__restrict__ dcol,
                        float *__restrict__ dx, int accumulate)
{
    const int Ho = (H - 1) / s + 1, Wo = (W - 1) / s + 1;
    const int64_t total = (int64_t)B * H * W * C;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int c = (int)(t % C);
    int64_t r = t / C;
    const int x = (int)(r % W); r /= W;
    const int y = (int)(r % H);
    const int b = (int)(r / H);
    float acc = 0.f;
    for (int ky = 0; ky < 3; ky++) {
        const int ny = y + 1 - ky;
        if (ny < 0 || ny % s) continue;
        const int oy = ny / s;
        if (oy >= Ho) continue;
        for (int kx = 0; kx < 3; kx++) {
            const int nx = x + 1 - kx;
            if (nx < 0 || nx % s) continue;
            const int ox = nx / s;
            if (ox >= Wo) continue;
            acc += dcol[((((int64_t)b * Ho + oy) * Wo + ox) * 9 + ky * 3 + kx) * C + c];
        }
    }
    dx[t] = accumulate ? dx[t] + acc : acc;
}

// dy <- dy * (y > 0)
__global__ void relu_bwd(int64_t n, const float *__restrict__ y, float *__restrict__ dy)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n && !(y[t] > 0.f)) dy[t] = 0.f;
}

// db[c] = sum over rows of dy[row][c] (one workgroup per 64 channels, fixed order)
// db[c] = sum_r dy[r][c] in a fixed order: block (k, channel tile) sums
// rows [k chunk, (k + 1) chunk) (four interleaved partial sums, added in
// order) into part[k][c]; bias_sum then adds the chunks in a fixed order.
__global__ void __launch_bounds__(256) bias_part(int64_t rows, int C, int64_t chunk,
                                                 const float *__restrict__ dy,
                                                 float *__restrict__ part)
{
    __shared__ float red[4][64];
    const int cl = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int c = blockIdx.y * 64 + cl;
    const int64_t r0 = (int64_t)blockIdx.x * chunk;
    const int64_t r1 = r0 + chunk < rows ? r0 + chunk : rows;
    float s = 0.f;
    if (c < C)
        for (int64_t r = r0 + q; r < r1; r += 4) s += dy[r * C + c];
    red[q][cl] = s;
    __syncthreads();
    if (q == 0 && c < C)
        part[(int64_t)blockIdx.x * C + c] = red[0][cl] + red[1][cl] + red[2][cl] + red[3][cl];
}

// one wave per channel: lane l adds chunks l, l + 64, ... in order, then a
// fixed butterfly over the lanes
__global__ void __launch_bounds__(256) bias_sum(int nchunk, int C, const float *__restrict__ part,
                                                float *__restrict__ db)
{
    const int c = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (c >= C) return;
    float s = 0.f;
    for (int k = lane; k < nchunk; k += 64) s += part[(int64_t)k * C + c];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) db[c] = s;
}

// y[b][y][x][c] = lo[b][y/2][x/2][c] + skip[b][y][x][c]   (H, W: fine grid)
__global__ void up2_add(int B, int H, int W, int C, const float *__restrict__ lo,
                        const float *__restrict__ skip, float *__restrict__ y)
{
    const int64_t total = (int64_t)B * H * W * C;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int c = (int)(t % C);
    int64_t r = t / C;
    const int x = (int)(r % W); r /= W;
    const int yy = (int)(r % H);
    const int b = (int)(r / H);
    const int Hl = (H + 1) / 2, Wl = (W + 1) / 2;
    y[t] = lo[(((int64_t)b * Hl + yy / 2) * Wl + x / 2) * C + c] + skip[t];
}

// dlo[b][yl][xl][c] = sum of the (up to) four fine-grid gradients it fed
__global__ void up2_bwd(int B, int H, int W, int C, const float *__restrict__ dy,
                        float *__restrict__ dlo)
{
    const int Hl = (H + 1) / 2, Wl = (W + 1) / 2;
    const int64_t total = (int64_t)B * Hl * Wl * C;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int c = (int)(t % C);
    int64_t r = t / C;
    const int xl = (int)(r % Wl); r /= Wl;
    const int yl = (int)(r % Hl);
    const int b = (int)(r / Hl);
    float s = 0.f;
    for (int dy_ = 0; dy_ < 2; dy_++)
        for (int dx_ = 0; dx_ < 2; dx_++) {
            const int yy = 2 * yl + dy_, xx = 2 * xl + dx_;
            if (yy < H && xx < W) s += dy[(((int64_t)b * H + yy) * W + xx) * C + c];
        }
    dlo[t] = s;
}

// element-centre outputs (B, n, n, C) -> interior node values of the
// (n+1)^2 node grid: u[b][i][j][c] = mean of the 4 elements around node
// (i, j), 1 <= i, j <= n-1; boundary nodes 0.  Loss 0.5*sum wc*(u - t)^2 /
// count and its gradient with respect to the element outputs.
__global__ void node_avg(int B, int n, int C, const float *__restrict__ e, float *__restrict__ u)
{
    const int nn = n + 1;
    const int64_t total = (int64_t)B * nn * nn * C;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int c = (int)(t % C);
    int64_t r = t / C;
    const int j = (int)(r % nn); r /= nn;
    const int i = (int)(r % nn);
    const int b = (int)(r / nn);
    float v = 0.f;
    if (i > 0 && j > 0 && i < n && j < n) {
        const float *p = e + ((int64_t)b * n * n) * C + c;
        v = 0.25f * (p[((i - 1) * n + j - 1) * C] + p[((i - 1) * n + j) * C] +
                     p[(i * n + j - 1) * C] + p[(i * n + j) * C]);
    }
    u[t] = v;
}

__global__ void node_loss_grad(int B, int n, int C, const float *__restrict__ u,
                               const float *__restrict__ target, const float *__restrict__ wc,
                               float scale, float *__restrict__ de)
{
    // de[b][ei][ej][c] = 0.25 * sum over the element's interior nodes of
    // scale * wc[c] * (u - target)
    const int nn = n + 1;
    const int64_t total = (int64_t)B * n * n * C;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int c = (int)(t % C);
    int64_t r = t / C;
    const int ej = (int)(r % n); r /= n;
    const int ei = (int)(r % n);
    const int b = (int)(r / n);
    float s = 0.f;
    for (int di = 0; di < 2; di++)
        for (int dj = 0; dj < 2; dj++) {
            const int i = ei + di, j = ej + dj;
            if (i > 0 && j > 0 && i < n && j < n) {
                const int64_t k = (((int64_t)b * nn + i) * nn + j) * C + c;
                s += u[k] - target[k];
            }
        }
    de[t] = 0.25f * scale * wc[c] * s;
}

// ---- energy-norm loss of the initial guess -----------------------------
// With u the network's node values scaled per channel, K the plate's
// stiffness (boundary conditions applied), f its loads and s = u*.f the
// energy of the solution u*, the best multiple alpha u of the guess
// (alpha = u.f / u.K u) leaves the relative energy-norm error
//   L = ||u* - alpha u||_K^2 / ||u*||_K^2 = 1 - a^2 / (c s),  a = u.f,
//   c = u.K u,
// the norm in which CG converges.  dL/du = (2a / (c s)) ((a / c) K u - f).

// x[t] = 0 on fixed DOFs, else u[t] * scale[t % C] (FP64 DOF vector)
__global__ void nn_to_dofs(int64_t n, int C, const float *__restrict__ u,
                           const float *__restrict__ scale, const uint8_t *__restrict__ fixed,
                           double *__restrict__ x)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    x[t] = fixed[t] ? 0. : (double)u[t] * (double)scale[t % C];
}

// one workgroup per plate: out[3b..3b+2] = (a, c, L) in a fixed order
__global__ void __launch_bounds__(256) energy_dots(int64_t per, const double *__restrict__ x,
                                                   const double *__restrict__ kx,
                                                   const double *__restrict__ f,
                                                   const double *__restrict__ s,
                                                   double *__restrict__ out)
{
    __shared__ double sa[256], sc[256];
    const int b = blockIdx.x;
    const double *xb = x + (int64_t)b * per, *kb = kx + (int64_t)b * per, *fb = f + (int64_t)b * per;
    double a = 0., c = 0.;
    for (int64_t t = threadIdx.x; t < per; t += 256) {
        a += xb[t] * fb[t];
        c += xb[t] * kb[t];
    }
    sa[threadIdx.x] = a;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sa[threadIdx.x] += sa[threadIdx.x + w];
            sc[threadIdx.x] += sc[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double A = sa[0], Cc = sc[0];
        out[3 * b] = A;
        out[3 * b + 1] = Cc;
        out[3 * b + 2] = (Cc > 0. && s[b] > 0.) ? 1. - A * A / (Cc * s[b]) : 1.;
    }
}

// de[b][ei][ej][c] = gscale * scale[c] * 0.25 * sum over the element's
// interior corner nodes of dL_b/du (the adjoint of node_avg)
__global__ void energy_grad(int B, int n, int C, const double *__restrict__ kx,
                            const double *__restrict__ f, const double *__restrict__ acl,
                            const double *__restrict__ s, const float *__restrict__ scale,
                            const uint8_t *__restrict__ fixed, float gscale,
                            float *__restrict__ de)
{
    const int nn = n + 1;
    const int64_t total = (int64_t)B * n * n * C;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int c = (int)(t % C);
    int64_t r = t / C;
    const int ej = (int)(r % n); r /= n;
    const int ei = (int)(r % n);
    const int b = (int)(r / n);
    const double A = acl[3 * b], Cc = acl[3 * b + 1], S = s[b];
    double g = 0.;
    if (Cc > 0. && S > 0. && isfinite(A) && isfinite(Cc)) {
        const double k1 = 2. * A / (Cc * S), k2 = A / Cc;
        for (int di = 0; di < 2; di++)
            for (int dj = 0; dj < 2; dj++) {
                const int i = ei + di, j = ej + dj;
                if (i > 0 && j > 0 && i < n && j < n) {
                    const int64_t k = (((int64_t)b * nn + i) * nn + j) * C + c;
                    if (!fixed[k]) g += k1 * (k2 * kx[k] - f[k]);
                }
            }
    }
    de[t] = (float)(0.25 * (double)gscale * (double)scale[c] * g);
}

// Adam (Kingma & Ba 2015) with bias correction folded into the step size
__global__ void adam_step(int64_t n, float *__restrict__ p, const float *__restrict__ g,
                          float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                          float b2, float eps, float c1, float c2)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gt = g[t];
    const float mt = b1 * m[t] + (1.f - b1) * gt;
    const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    p[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

inline unsigned nb256(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

extern "C" {

int pf3d_nn_gemm(int M, int N, int K, const float *A, int64_t sam, int64_t sak, const float *B,
                 int64_t sbn, int64_t sbk, float *C, int64_t ldc, int accumulate,
                 const float *bias, int relu, int splits, float *partial, void *stream)
{
    if (M < 0 || N < 0 || K < 0 || (M && N && (!A || !B || !C)))
        return pf3d_fail(PF3D_E_ARG, "pf3d_nn_gemm: bad arguments");
    if (!M || !N) return PF3D_OK;
    if (splits < 1 || (splits > 1 && !partial))
        return pf3d_fail(PF3D_E_ARG, "pf3d_nn_gemm: split-K needs a partial buffer");
    int kchunk = (K + splits - 1) / splits;
    kchunk = ((kchunk + TK - 1) / TK) * TK;
    if (kchunk == 0) kchunk = TK;
    const int z = (K + kchunk - 1) / kchunk;
    Gemm g{M, N, K, A, sam, sak, B, sbn, sbk, C, ldc, accumulate, bias, relu, kchunk,
           z > 1 ? partial : nullptr};
    dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM, z > 0 ? z : 1);
    if (grid.y > 65535) return pf3d_fail(PF3D_E_ARG, "pf3d_nn_gemm: M too large");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(gemm_bf16, grid, dim3(256), 0, s, g);
    if (z > 1)
        hipLaunchKernelGGL(splitk_reduce, dim3(nb256((int64_t)M * N)), dim3(256), 0, s, g, z);
    return pf3d_hip_check(hipGetLastError(), "pf3d_nn_gemm");
}

int pf3d_nn_im2col(int B, int H, int W, int C, int stride, const float *x, float *col,
                   void *stream)
{
    if (B <= 0 || H <= 0 || W <= 0 || C <= 0 || (stride != 1 && stride != 2) || !x || !col)
        return pf3d_fail(PF3D_E_ARG, "pf3d_nn_im2col: bad arguments");
    const int Ho = (H - 1) / stride + 1, Wo = (W - 1) / stride + 1;
    const int64_t n = (int64_t)B * Ho * Wo * 9 * C;
    hipLaunchKernelGGL(im2col3, dim3(nb256(n)), dim3(256), 0, (hipStream_t)stream, B, H, W, C,
                       stride, x, col);
    return pf3d_hip_check(hipGetLastError(), "pf3d_nn_im2col");
}

int pf3d_nn_col2im(int B, int H, int W, int C, int stride, const float *dcol, float *dx,
                   int accumulate, void *stream)
{
    if (B <= 0 || H <= 0 || W <= 0 || C <= 0 || (stride != 1 && stride != 2) || !dcol || !dx)
        return pf3d_fail(PF3D_E_ARG, "pf3d_nn_col2im: bad arguments");
    const int64_t n = (int64_t)B * H * W * C;
    hipLaunchKernelGGL(col2im3, dim3(nb256(n)), dim3(256), 0, (hipStream_t)stream, B, H, W, C,
                       stride, dcol, dx, accumulate);
    return pf3d_hip_check(hipGetLastError(), "pf3d_nn_col2im");
}

int pf3d_nn_relu_bwd(int64_t n, const float *y, float *dy, void *stream)
{
    if (n < 0 || (n && (!y || !dy))) return pf3d_fail(PF3D_E_ARG, "pf3d_nn_relu_bwd");
    if (!n) return PF3D_OK;
    hipLaunchKernelGGL(relu_bwd, dim3(nb256(n)), dim3(256), 0, (hipStream_t)stream, n, y, dy);
    return pf3d_hip_check(hipGetLastError(), "pf3d_nn_relu_bwd");
}

int pf3d_nn_bias_grad(int64_t rows, int C, const float *dy, float *db, float *work,
                      int64_t work_floats, void *stream)
{
    if (rows <= 0 || C <= 0 || !dy || !db || work_floats < 0 || (work_floats && !work))
        return pf3d_fail(PF3D_E_ARG, "pf3d_nn_bias_grad");
    if (work_floats < C)
        return pf3d_fail(PF3D_E_ARG, "pf3d_nn_bias_grad: work must hold at least C floats");
    // chunks of >= 256 rows, at most 1024 of them, as many as the work holds
    int64_t nchunk = (rows + 255) / 256;
    if (nchunk > 1024) nchunk = 1024;
    if (nchunk > work_floats / C) nchunk = work_floats / C;
    const int64_t chunk = (rows + nchunk - 1) / nchunk;
    nchunk = (rows + chunk - 1) / chunk;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(bias_part, dim3((unsigned)nchunk, (C + 63) / 64), dim3(256), 0, s, rows, C,
                       chunk, dy, work);
    hipLaunchKernelGGL(bias_sum, dim3((C + 3) / 4), dim3(256), 0, s, (int)nchunk, C, work, db);
    return pf3d_hip_check(hipGetLastError(), "pf3d_nn_bias_grad");
}

int pf3d_nn_up2_add(int B, int H, int W, int C, const float *lo, const float *skip, float *y,
                    void *stream)
{
    if (B <= 0 || H <= 0 || W <= 0 || C <= 0 || !lo || !skip || !y)
        return pf3d_fail(PF3D_E_ARG, "pf3d_nn_up2_add");
    const int64_t n = (int64_t)B * H * W * C;
    hipLaunchKernelGGL(up2_add, dim3(nb256(n)), dim3(256), 0, (hipStream_t)stream, B, H, W, C, lo,
                       skip, y);
    return pf3d_hip_check(hipGetLastError(), "pf3d_nn_up2_add");
}

int pf3d_nn_up2_bwd(int B, int H, int W, int C, const float *dy, float *dlo, void *stream)
{
    if (B <= 0 || H <= 0 || W <= 0 || C <= 0 || !dy || !dlo)
        return pf3d_fail(PF3D_E_ARG, "pf3d_nn_up2_bwd");
    const int64_t n = (int64_t)B * ((H + 1) / 2) * ((W + 1) / 2) * C;
    hipLaunchKernelGGL(up2_bwd, dim3(nb256(n)), dim3(256), 0, (hipStream_t)stream, B, H, W, C, dy,
                       dlo);
    return pf3d_hip_check(hipGetLastError(), "pf3d_nn_up2_bwd");
}

int pf3d_nn_node_avg(int B, int n, int C, const float *e, float *u, void *stream)
{
    if (B <= 0 || n <= 1 || C <= 0 || !e || !u) return pf3d_fail(PF3D_E_ARG, "pf3d_nn_node_avg");
    const int64_t tot = (int64_t)B * (n + 1) * (n + 1) * C;
    hipLaunchKernelGGL(node_avg, dim3(nb256(tot)), dim3(256), 0, (hipStream_t)stream, B, n, C, e,
                       u);
    return pf3d_hip_check(hipGetLastError(), "pf3d_nn_node_avg");
}

int pf3d_nn_node_loss_grad(int B, int n, int C, const float *u, const float *target,
                           const float *wc, float scale, float *de, void *stream)
{
    if (B <= 0 || n <= 1 || C <= 0 || !u || !target || !wc || !de)
        return pf3d_fail(PF3D_E_ARG, "pf3d_nn_node_loss_grad");
    const int64_t tot = (int64_t)B * n * n * C;
    hipLaunchKernelGGL(node_loss_grad, dim3(nb256(tot)), dim3(256), 0, (hipStream_t)stream, B, n,
                       C, u, target, wc, scale, de);
    return pf3d_hip_check(hipGetLastError(), "pf3d_nn_node_loss_grad");
}

int pf3d_nn_to_dofs(int64_t n, int C, const float *u, const float *scale, const uint8_t *fixed,
                    double *x, void *stream)
{
    if (n < 0 || C <= 0 || (n && (!u || !scale || !fixed || !x)))
        return pf3d_fail(PF3D_E_ARG, "pf3d_nn_to_dofs: bad arguments");
    if (!n) return PF3D_OK;
    hipLaunchKernelGGL(nn_to_dofs, dim3(nb256(n)), dim3(256), 0, (hipStream_t)stream, n, C, u,
                       scale, fixed, x);
    return pf3d_hip_check(hipGetLastError(), "pf3d_nn_to_dofs");
}

int pf3d_nn_energy_dots(int B, int64_t per, const double *x, const double *kx, const double *f,
                        const double *s, double *out, void *stream)
{
    if (B <= 0 || per <= 0 || !x || !kx || !f || !s || !out)
        return pf3d_fail(PF3D_E_ARG, "pf3d_nn_energy_dots: bad arguments");
    hipLaunchKernelGGL(energy_dots, dim3(B), dim3(256), 0, (hipStream_t)stream, per, x, kx, f, s,
                       out);
    return pf3d_hip_check(hipGetLastError(), "pf3d_nn_energy_dots");
}

int pf3d_nn_energy_grad(int B, int n, int C, const double *kx, const double *f,
                        const double *acl, const double *s, const float *scale,
                        const uint8_t *fixed, float gscale, float *de, void *stream)
{
    if (B <= 0 || n <= 1 || C <= 0 || !kx || !f || !acl || !s || !scale || !fixed || !de)
        return pf3d_fail(PF3D_E_ARG, "pf3d_nn_energy_grad: bad arguments");
    const int64_t tot = (int64_t)B * n * n * C;
    hipLaunchKernelGGL(energy_grad, dim3(nb256(tot)), dim3(256), 0, (hipStream_t)stream, B, n, C,
                       kx, f, acl, s, scale, fixed, gscale, de);
    return pf3d_hip_check(hipGetLastError(), "pf3d_nn_energy_grad");
}

int pf3d_nn_adam(int64_t n, float *p, const float *g, float *m, float *v, float lr, float b1,
                 float b2, float eps, int step, void *stream)
{
    if (n <= 0 || !p || !g || !m || !v || step < 1) return pf3d_fail(PF3D_E_ARG, "pf3d_nn_adam");
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_step, dim3(nb256(n)), dim3(256), 0, (hipStream_t)stream, n, p, g, m, v,
                       lr, b1, b2, eps, c1, c2);
    return pf3d_hip_check(hipGetLastError(), "pf3d_nn_adam");
}

}  // extern "C"

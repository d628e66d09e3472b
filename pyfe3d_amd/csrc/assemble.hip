// Element assembly kernels: KC0, KG (from displacements), KG (given stress)
// and M for the Quad4 / Quad4R / Tria3R shells, KC0 for BeamC.
//
// Work mapping (gfx950, wave64): one lane owns one 6x6 node-pair block
// (I, J) of one element, so a wave holds 4 quads (16 lanes each) or 7
// triangles (9 lanes each).  Each lane integrates its block in registers,
// rotates it to global axes (R6 K R6^T with R6 = blockdiag(R, R)) and parks
// it in an LDS tile; the wave then streams the element's slots out in the
// reference's slot order with unit-stride stores, so every store
// instruction writes 64 consecutive doubles of the COO value array.
#include "pf3d_device.hpp"

// unroll factor of the 2x2 Gauss loop of the Quad4 KC0 block integral
#ifndef PF3D_GP_UNROLL
#define PF3D_GP_UNROLL 4
#endif

namespace pf3d {

// ------------------------------------------------------------------------
// slot layouts: slot k of an element -> (local row, local col)
// ------------------------------------------------------------------------
template <int NN>
__device__ __forceinline__ void slot_layout(int mat, int mtype, int k, int &rl, int &cl)
{
    constexpr int NDOF = 6 * NN;
    if (mat == KC0) {  // quad4.pyx:1287-1293 (I, m, J, n)
        rl = k / NDOF;
        cl = k - rl * NDOF;
    } else if (mat == KG || mat == KGS) {  // quad4.pyx:1500-1932 (I, a, J, b), translations
        int row = k / (3 * NN), col = k - row * 3 * NN;
        rl = 6 * (row / 3) + row % 3;
        cl = 6 * (col / 3) + col % 3;
    } else if (mtype == 2) {  // quad4.pyx:8005-8872 lumped
        int I = k / (18 * NN), rem = k - I * 18 * NN;
        int g = rem / (9 * NN), rem2 = rem - g * 9 * NN;
        int d = rem2 / (3 * NN), col = rem2 - d * 3 * NN;
        rl = 6 * I + 3 * g + d;
        cl = 6 * (col / 3) + 3 * g + col % 3;
    } else {  // quad4.pyx:3157-4598: row d skips column (d+3) mod 6
        int row = k / (5 * NN), col = k - row * 5 * NN;
        int d = row % 6, t = col % 5, skip = (d + 3) % 6;
        rl = row;
        cl = 6 * (col / 5) + (t < skip ? t : t + 1);
    }
}

template <int NN>
__host__ __device__ constexpr int slot_count(int mat, int mtype)
{
    return mat == KC0 ? 36 * NN * NN
         : kg_like(mat) ? 9 * NN * NN
         : (mtype == 2 ? 18 * NN * NN : 30 * NN * NN);
}

// Select a[i] for a lane-varying i.  Written as a sum of products with 0/1
// weights (exact in IEEE arithmetic for finite values) so that the compiler
// keeps the operands in registers instead of spilling them to an indexed
// private array.
template <int N>
__device__ __forceinline__ double pick(const double (&a)[N], int i)
{
    double r = 0.;
#pragma unroll
    for (int k = 0; k < N; k++) r += a[k] * (double)(i == k);
    return r;
}

// bilinear shape values of the lane's node pair at (xi, eta); the Jacobian
// arithmetic is that of quad_shape (quad4.pyx:937-962)
struct Pair {
    double Ni, nxi, nyi, Nj, nxj, nyj, det;
};

__device__ __forceinline__ Pair quad_pair(const double *xe, double xi, double eta, double sxi,
                                          double syi, double sxj, double syj)
{
    const double sx[4] = {-1, 1, 1, -1}, sy[4] = {-1, -1, 1, 1};
    double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const double dxi = 0.25 * sx[i] * (1 + sy[i] * eta);
        const double deta = 0.25 * sy[i] * (1 + sx[i] * xi);
        J11 += dxi * xe[2 * i];
        J12 += dxi * xe[2 * i + 1];
        J21 += deta * xe[2 * i];
        J22 += deta * xe[2 * i + 1];
    }
    Pair p;
    p.det = J11 * J22 - J12 * J21;
    // one reciprocal instead of four divisions (within an ulp of J_ij / det)
    const double inv = 1. / p.det;
    const double j11 = J22 * inv, j12 = -J12 * inv, j21 = -J21 * inv, j22 = J11 * inv;
    const double dxi_i = 0.25 * sxi * (1 + syi * eta), deta_i = 0.25 * syi * (1 + sxi * xi);
    const double dxi_j = 0.25 * sxj * (1 + syj * eta), deta_j = 0.25 * syj * (1 + sxj * xi);
    p.Ni = 0.25 * (1 + sxi * xi) * (1 + syi * eta);
    p.Nj = 0.25 * (1 + sxj * xi) * (1 + syj * eta);
    p.nxi = j11 * dxi_i + j12 * deta_i;
    p.nyi = j21 * dxi_i + j22 * deta_i;
    p.nxj = j11 * dxi_j + j12 * deta_j;
    p.nyj = j21 * dxi_j + j22 * deta_j;
    return p;
}

__device__ __forceinline__ double node_sx(int I) { return (I == 1 || I == 2) ? 1. : -1.; }
__device__ __forceinline__ double node_sy(int I) { return (I >= 2) ? 1. : -1.; }

// ------------------------------------------------------------------------
// block integrands (lane owns block I, J; K is 6x6 in local axes)
// ------------------------------------------------------------------------

// membrane/bending B^T [[A,B],[B,D]] B over (u, v, rx, ry); quad4.pyx:1002-1030
__device__ __forceinline__ void add_abd(double K[6][6], const double C[6][6], double nxi,
                                        double nyi, double nxj, double nyj, double w)
{
    nxi *= w;
    nyi *= w;
    const int col[4] = {0, 1, 3, 4};
    // one column of C*B_J at a time (u, v, rx, ry), folded into K at once
#pragma unroll
    for (int t = 0; t < 4; t++) {
        double cb[6];
#pragma unroll
        for (int a = 0; a < 6; a++) {
            if (t == 0) cb[a] = C[a][0] * nxj + C[a][2] * nyj;
            else if (t == 1) cb[a] = C[a][1] * nyj + C[a][2] * nxj;
            else if (t == 2) cb[a] = -(C[a][4] * nyj + C[a][5] * nxj);
            else cb[a] = C[a][3] * nxj + C[a][5] * nyj;
        }
        K[0][col[t]] += nxi * cb[0] + nyi * cb[2];
        K[1][col[t]] += nyi * cb[1] + nxi * cb[2];
        K[3][col[t]] -= nyi * cb[4] + nxi * cb[5];
        K[4][col[t]] += nxi * cb[3] + nyi * cb[5];
    }
}

// transverse shear gyz = w,y - rx, gxz = w,x + ry split into gradient (w)
// and rotation (rx, ry) parts; quad4.pyx:1032-1051, 1107-1171
__device__ __forceinline__ void add_shear(double K[6][6], const double E[2][2], double Ni,
                                          double nxi, double nyi, double Nj, double nxj,
                                          double nyj, double w, bool gg, bool gr, bool rr)
{
    // E times the J columns: gradient (w) and rotations (rx, ry)
    const double egw0 = E[0][0] * nyj + E[0][1] * nxj, egw1 = E[1][0] * nyj + E[1][1] * nxj;
    const double erx0 = -E[0][0] * Nj, erx1 = -E[1][0] * Nj;
    const double ery0 = E[0][1] * Nj, ery1 = E[1][1] * Nj;
    const double gi0 = w * nyi, gi1 = w * nxi;  // row w of node I
    const double rxi = -w * Ni;                 // row rx: gyz component only
    const double ryi = w * Ni;                  // row ry: gxz component only
    if (gg) K[2][2] += gi0 * egw0 + gi1 * egw1;
    if (gr) {
        K[2][3] += gi0 * erx0 + gi1 * erx1;
        K[2][4] += gi0 * ery0 + gi1 * ery1;
        K[3][2] += rxi * egw0;
        K[4][2] += ryi * egw1;
    }
    if (rr) {
        K[3][3] += rxi * erx0;
        K[3][4] += rxi * ery0;
        K[4][3] += ryi * erx1;
        K[4][4] += ryi * ery1;
    }
}

// drilling rz - (v,x - u,y)/2 penalty; quad4.pyx:1053-1073
__device__ __forceinline__ void add_drill(double K[6][6], double Ni, double nxi, double nyi,
                                          double Nj, double nxj, double nyj, double w,
                                          bool rz_couple, bool rz_diag)
{
    const double bi[3] = {nyi / 2., -nxi / 2., Ni};
    const double bj[3] = {nyj / 2., -nxj / 2., Nj};
    const int dof[3] = {0, 1, 5};
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
            bool rz = (a == 2) || (b == 2);
            if (rz && !(rz_couple || (rz_diag && a == 2 && b == 2))) continue;
            K[dof[a]][dof[b]] += w * bi[a] * bj[b];
        }
}

// R6 K R6^T for one 6x6 node block (block-diagonal R6 = diag(R, R))
__device__ __forceinline__ void rotate_block(double K[6][6], const double *r)
{
    double T[6][6];
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int q = 0; q < 2; q++) {
            double S[3][3];  // R * K_pq
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int j = 0; j < 3; j++)
                    S[a][j] = r[a * 3 + 0] * K[3 * p + 0][3 * q + j] +
                              r[a * 3 + 1] * K[3 * p + 1][3 * q + j] +
                              r[a * 3 + 2] * K[3 * p + 2][3 * q + j];
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++)
                    T[3 * p + a][3 * q + b] =
                        S[a][0] * r[b * 3 + 0] + S[a][1] * r[b * 3 + 1] + S[a][2] * r[b * 3 + 2];
        }
#pragma unroll
    for (int a = 0; a < 6; a++)
#pragma unroll
        for (int b = 0; b < 6; b++) K[a][b] = T[a][b];
}

// ------------------------------------------------------------------------
// per-type local blocks
// ------------------------------------------------------------------------

// Gauss-point moments of the lane's node pair: with g = (N,x, N,y) and w
// the Gauss weight times detJ,
//   G[p][q] = sum w g_I[p] g_J[q],  NG[q] = sum w N_I g_J[q],
//   GN[p] = sum w g_I[p] N_J,       NN = sum w N_I N_J.
// Every 2x2-Gauss term of the Quad4 stiffness that has a constant material
// table (membrane/bending ABD, thick-shell shear gradients, drilling) is a
// fixed linear combination of these nine sums, so the per-point work is
// nine FMAs instead of the B^T C B products.
struct Moments {
    double G[2][2], NG[2], GN[2], NN;
};

__device__ __forceinline__ void add_moments(Moments &m, const Pair &n, double w)
{
    const double gi[2] = {w * n.nxi, w * n.nyi}, gj[2] = {n.nxj, n.nyj};
    const double wNi = w * n.Ni;
#pragma unroll
    for (int p = 0; p < 2; p++) {
#pragma unroll
        for (int q = 0; q < 2; q++) m.G[p][q] += gi[p] * gj[q];
        m.NG[p] += wNi * gj[p];
        m.GN[p] += gi[p] * n.Nj;
    }
    m.NN += wNi * n.Nj;
}

// membrane/bending part from G: the strains of DOF d of a node are
// s * g[p] in strain row r for its two (p, r, s) pairs (quad4.pyx:1002-1030):
//   u: (x, exx, +), (y, gxy, +)    v: (x, gxy, +), (y, eyy, +)
//   rx: (x, kxy, -), (y, kyy, -)   ry: (x, kxx, +), (y, kxy, +)
// so K[a][b] = sum_pq s_ap s_bq C[r_ap][r_bq] G[p][q].
__device__ __forceinline__ void abd_from_moments(double K[6][6], const double C[6][6],
                                                 const double G[2][2])
{
    constexpr int dof[4] = {0, 1, 3, 4};
    constexpr int row[4][2] = {{0, 2}, {2, 1}, {5, 4}, {3, 5}};
    constexpr double sg[4] = {1., 1., -1., 1.};
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            double t = 0.;
#pragma unroll
            for (int p = 0; p < 2; p++)
#pragma unroll
                for (int q = 0; q < 2; q++) t += C[row[a][p]][row[b][q]] * G[p][q];
            K[dof[a]][dof[b]] += (sg[a] * sg[b]) * t;
        }
}

// drilling (rz - (v,x - u,y)/2)^2 with b = (N,y/2, -N,x/2, N) on (u, v, rz)
// (quad4.pyx:1053-1073), from the moments
__device__ __forceinline__ void drill_from_moments(double K[6][6], const Moments &m, double f)
{
    K[0][0] += f * 0.25 * m.G[1][1];
    K[0][1] -= f * 0.25 * m.G[1][0];
    K[1][0] -= f * 0.25 * m.G[0][1];
    K[1][1] += f * 0.25 * m.G[0][0];
    K[0][5] += f * 0.5 * m.GN[1];
    K[1][5] -= f * 0.5 * m.GN[0];
    K[5][0] += f * 0.5 * m.NG[1];
    K[5][1] -= f * 0.5 * m.NG[0];
    K[5][5] += f * m.NN;
}

// Quad4 KC0: quad4.pyx:755-1172
__device__ __forceinline__ void quad4_kc0(double K[6][6], const Geom &g, const Shell &s, int I,
                                          int J)
{
    const bool thick = (s.h / sqrt(g.area)) >= 1.;
    const double sxi = node_sx(I), syi = node_sy(I), sxj = node_sx(J), syj = node_sy(J);
    Moments m = {{{0., 0.}, {0., 0.}}, {0., 0.}, {0., 0.}, 0.};
#pragma unroll PF3D_GP_UNROLL
    for (int p = 0; p < 4; p++) {
        const Pair n = quad_pair(g.xe, (p >> 1) ? GP2 : -GP2, (p & 1) ? GP2 : -GP2, sxi, syi,
                                 sxj, syj);
        add_moments(m, n, n.det);
    }
    abd_from_moments(K, s.C, m.G);
    // thick shells: shear gradient-gradient term at the 2x2 points
    // (quad4.pyx:1032-1051), rows/cols (gyz, gxz) <- (N,y, N,x)
    if (thick)
        K[2][2] += s.Es[0][0] * m.G[1][1] + s.Es[0][1] * m.G[1][0] + s.Es[1][0] * m.G[0][1] +
                   s.Es[1][1] * m.G[0][0];
    drill_from_moments(K, m, 1.);
    const Pair n = quad_pair(g.xe, 0., 0., sxi, syi, sxj, syj);
    add_shear(K, s.Es, n.Ni, n.nxi, n.nyi, n.Nj, n.nxj, n.nyj, 4. * n.det, !thick, true, true);
}

// Quad4R KC0: quad4r.pyx:1300-1312 (E1eq, E2eq), 3081-3462
__device__ __forceinline__ void quad4r_kc0(double K[6][6], const Geom &g, const Shell &s, int I,
                                           int J, double k6rot, const double *hgf)
{
    const double A11 = s.C[0][0], A12 = s.C[0][1], A16 = s.C[0][2];
    const double A22 = s.C[1][1], A26 = s.C[1][2], A66 = s.C[2][2];
    const double det3 = A11 * (A22 * A66 - A26 * A26) - A12 * (A12 * A66 - A26 * A16) +
                        A16 * (A12 * A26 - A22 * A16);
    const double a11 = (A22 * A66 - A26 * A26) / det3, a22 = (A11 * A66 - A16 * A16) / det3;
    const double E1eq = 1. / (s.h * a11), E2eq = 1. / (s.h * a22);
    const double damp = 1.0 + 1.0 / g.area, h3 = s.h * s.h * s.h;
    double Ehg[5];
    Ehg[0] = hgf[0] * 0.1 * E1eq * s.h / damp;
    Ehg[1] = hgf[1] * 0.1 * E2eq * s.h / damp;
    Ehg[3] = hgf[3] * 0.1 * E2eq * h3 / damp;
    Ehg[4] = hgf[4] * 0.1 * E1eq * h3 / damp;
    Ehg[2] = hgf[2] * 0.5 * (Ehg[3] + Ehg[4]);

    const double sxi = node_sx(I), syi = node_sy(I), sxj = node_sx(J), syj = node_sy(J);
    const Pair n = quad_pair(g.xe, 0., 0., sxi, syi, sxj, syj);
    const double w = 4. * n.det;
    add_abd(K, s.C, n.nxi, n.nyi, n.nxj, n.nyj, w);
    add_shear(K, s.Es, n.Ni, n.nxi, n.nyi, n.Nj, n.nxj, n.nyj, w, true, true, true);
    // hourglass modes gamma_I = d2N_I/dxdy at the centroid
    const double *xe = g.xe;
    const double J11 = 0.25 * (-xe[0] + xe[2] + xe[4] - xe[6]);
    const double J12 = 0.25 * (-xe[1] + xe[3] + xe[5] - xe[7]);
    const double J21 = 0.25 * (-xe[0] - xe[2] + xe[4] + xe[6]);
    const double J22 = 0.25 * (-xe[1] - xe[3] + xe[5] + xe[7]);
    const double dj = J11 * J22 - J12 * J21;
    const double j11 = J22 / dj, j12 = -J12 / dj, j21 = -J21 / dj, j22 = J11 / dj;
    const double gxy = 0.25 * (j11 * j22 + j12 * j21);
    const double gI = (I & 1) ? -gxy : gxy, gJ = (J & 1) ? -gxy : gxy;
#pragma unroll
    for (int d = 0; d < 5; d++) K[d][d] += w * Ehg[d] * gI * gJ;
    const double kd = 1e-6 * k6rot * A66;
#pragma unroll
    for (int p = 0; p < 4; p++) {
        const Pair m = quad_pair(g.xe, (p >> 1) ? GP2 : -GP2, (p & 1) ? GP2 : -GP2, sxi, syi,
                                 sxj, syj);
        add_drill(K, m.Ni, m.nxi, m.nyi, m.Nj, m.nxj, m.nyj, m.det * kd, true, true);
    }
}

// Tria3R shape-function gradients (constant): tria3r.pyx:2116-2121
__device__ __forceinline__ void tria_grad(const Geom &g, double Nx[3], double Ny[3])
{
    const double *p = g.xe;
    const double a2 = 2 * g.area;
    Nx[0] = (p[3] - p[5]) / a2; Nx[1] = (p[5] - p[1]) / a2; Nx[2] = (p[1] - p[3]) / a2;
    Ny[0] = (p[4] - p[2]) / a2; Ny[1] = (p[0] - p[4]) / a2; Ny[2] = (p[2] - p[0]) / a2;
}

// Tria3R KC0: tria3r.pyx:1105-1122 (shear factor), 2116-2324 (terms).
// Only the rz_I-rz_I drilling diagonal reaches the global matrix: the other
// rz couplings computed at :2230-2324 are not used by the scatter at :2325-2974.
// full_drill keeps every drilling coupling: the internal forces use them
// (tria3r.pyx:549-948) although the KC0 scatter drops all rz couplings but
// the rz_I-rz_I diagonal.
__device__ __forceinline__ void tria3r_kc0(double K[6][6], const Geom &g, const Shell &s, int I,
                                           int J, double k6rot, double alpha,
                                           bool full_drill = false)
{
    const double *p = g.xe;
    const double l12 = sqrt((p[0] - p[2]) * (p[0] - p[2]) + (p[1] - p[3]) * (p[1] - p[3]));
    const double l23 = sqrt((p[2] - p[4]) * (p[2] - p[4]) + (p[3] - p[5]) * (p[3] - p[5]));
    const double l31 = sqrt((p[4] - p[0]) * (p[4] - p[0]) + (p[5] - p[1]) * (p[5] - p[1]));
    double maxl = l12;
    if (l23 > maxl) maxl = l23;
    if (l31 > maxl) maxl = l31;
    const double f = 1 / (1 + alpha * maxl * maxl / (s.h * s.h));
    double E[2][2] = {{s.Es[0][0] * f, s.Es[0][1] * f}, {s.Es[1][0] * f, s.Es[1][1] * f}};
    double Nx[3], Ny[3];
    tria_grad(g, Nx, Ny);
    const double nxi = pick(Nx, I), nyi = pick(Ny, I), nxj = pick(Nx, J), nyj = pick(Ny, J);
    const double detJ = 2 * g.area, w = 0.5 * detJ, third = 1. / 3.;
    add_abd(K, s.C, nxi, nyi, nxj, nyj, w);
    add_shear(K, E, third, nxi, nyi, third, nxj, nyj, w, true, true, true);
    const double kd = 1e-6 * k6rot * s.C[2][2];
    const double P[3][3] = {{2. / 3, 1. / 6, 1. / 6}, {1. / 6, 1. / 6, 2. / 3}, {1. / 6, 2. / 3, 1. / 6}};
#pragma unroll
    for (int t = 0; t < 3; t++)
        add_drill(K, pick(P[t], I), nxi, nyi, pick(P[t], J), nxj, nyj, detJ * (0.5 / 3.) * kd,
                  full_drill, I == J);
}

// membrane stress resultants from local displacements (quad4.pyx:1965-1967)
template <int NN>
__device__ __forceinline__ void membrane_forces(const Shell &s, const double *Nx,
                                                const double *Ny, const double *ue,
                                                double Nf[3])
{
    double eps[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < NN; i++) {
        const double *u = ue + 6 * i;
        eps[0] += Nx[i] * u[0];
        eps[1] += Ny[i] * u[1];
        eps[2] += Ny[i] * u[0] + Nx[i] * u[1];
        eps[3] += Nx[i] * u[4];
        eps[4] -= Ny[i] * u[3];
        eps[5] += Ny[i] * u[4] - Nx[i] * u[3];
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double t = 0;
#pragma unroll
        for (int b = 0; b < 6; b++) t += s.C[a][b] * eps[b];
        Nf[a] = t;
    }
}

// ------------------------------------------------------------------------
// KG, KG given stress and M: the node-pair blocks of these matrices are a
// scalar per pair times one 6x6 pattern per element (KG: G_IJ n n^T on the
// translations, n the element normal; M: H_IJ R6 m R6^T), so the work is
// split in two phases.  Phase 1: one lane per element (64 elements per
// wave; 32 for M) computes the geometry, the per-pair scalars and the
// pattern into LDS.  Phase 2: the wave streams its elements' slots in the
// reference layout, consecutive lanes computing consecutive slots from LDS,
// so the stores are contiguous across lanes.
// ------------------------------------------------------------------------
// 16-byte vectors for the slot streaming; F_VEC (internal launch flag):
// every element's first slot is even and v, r, c are 16-byte aligned
using dbl2 = double __attribute__((ext_vector_type(2)));
using lng2 = long long __attribute__((ext_vector_type(2)));
constexpr int F_VEC = 4;

// the COO output is written once and not read again by this kernel:
// non-temporal stores keep it from evicting the inputs from the caches
// (measured on MI355X: 1.12 -> 0.96 ms for 1e6 Quad4 KC0)
#ifndef PF3D_NT_STORE
#define PF3D_NT_STORE 1
#endif
template <typename T>
__device__ __forceinline__ void store2(T *p, T v)
{
#if PF3D_NT_STORE
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}

// LDS layout of one shell_stream block (also sizes the mixed-type kernels)
template <int ET, int MAT>
struct StreamLds {
    static constexpr int NN = (ET == TRIA3R) ? 3 : 4;
    static constexpr int NP = NN * NN;
    static constexpr int W = (MAT == MASS) ? NP + 36 : NP + 3;  // per-element LDS doubles
    // elements per 64-lane block: the mass pattern needs 4x the LDS of KG,
    // so mass blocks take 32 elements to keep several waves per CU
    static constexpr int EPB = (MAT == MASS) ? 32 : 64;
    static constexpr int BYTES = EPB * (W * 8 + NN * 8 + 8);
};

template <int ET, int MAT>
__device__ __forceinline__ void shell_stream_body(const AsmArgs &a, int64_t blk, char *lds)
{
    using L = StreamLds<ET, MAT>;
    constexpr int NN = L::NN, NP = L::NP, W = L::W, EPB = L::EPB;
    double (*sd)[W] = reinterpret_cast<double (*)[W]>(lds);
    int64_t (*sconn)[NN] = reinterpret_cast<int64_t (*)[NN]>(lds + EPB * W * 8);
    int64_t *sk0 = reinterpret_cast<int64_t *>(lds + EPB * (W + NN) * 8);
    const int lane = threadIdx.x;
    const int64_t e0 = blk * EPB;
    const int64_t e = e0 + lane;
    const int mtype = a.mtype;
    const int nslot = slot_count<NN>(MAT, mtype);
    if (lane < EPB && e < a.n_elem) {
        V3 X[NN];
        int64_t cn[NN];
#pragma unroll
        for (int i = 0; i < NN; i++) {
            cn[i] = a.conn[e * NN + i];
            sconn[lane][i] = cn[i];
            const double *px = a.x + cn[i] / 2;
            X[i] = {px[0], px[1], px[2]};
        }
        sk0[lane] = a.init_k ? a.init_k[e] : (int64_t)nslot * e;
        V3 xm = {0., 0., 0.};
        if (a.evec) xm = {a.evec[3 * e], a.evec[3 * e + 1], a.evec[3 * e + 2]};
        Geom g;
        if (ET == TRIA3R) tria_geometry(X, xm, g);
        else quad_geometry(X, xm, g);
        Shell s;
        if (MAT == KG || MAT == MASS)  // given stress and piston theory read no property
            load_shell(a.props + (a.pid ? (int64_t)a.pid[e] : 0) * SHELL_STRIDE, g.m, s);
        double P[NP];
#pragma unroll
        for (int k = 0; k < NP; k++) P[k] = 0.;
        if (piston(MAT)) {
            // piston theory on the normal displacement, 2x2 Gauss
            // (quad4.pyx:9491-11917, quad4r.pyx:12789-15200): KA_beta
            // -int N_I dN_J/dX, KA_gamma int N_I N_J, CA -int N_I N_J, with
            // dN/dX = N,x r11 + N,y r21 (X = global x, the flow direction)
#pragma unroll
            for (int p = 0; p < 4; p++) {
                QShape q;
                quad_shape(g.xe, (p >> 1) ? GP2 : -GP2, (p & 1) ? GP2 : -GP2, q);
#pragma unroll
                for (int I = 0; I < 4; I++)
#pragma unroll
                    for (int J = 0; J < 4; J++) {
                        if (MAT == KAB)
                            P[I * 4 + J] -= q.N[I] * q.det * (q.Nx[J] * g.r[0] + q.Ny[J] * g.r[3]);
                        else if (MAT == KAG) P[I * 4 + J] += q.N[I] * q.N[J] * q.det;
                        else P[I * 4 + J] -= q.N[I] * q.N[J] * q.det;
                    }
            }
        }
        if (kg_like(MAT)) {
            if (MAT == KG || MAT == KGS) {
            // quad4.pyx:1365-2257 / 2259-3081, tria3r.pyx:3020-4062
            double ue[6 * NN];
            double Nf[3] = {0., 0., 0.};
            if (MAT == KG) {
#pragma unroll
                for (int i = 0; i < NN; i++)
#pragma unroll
                    for (int b = 0; b < 2; b++) {
                        const double *pu = a.u + cn[i] + 3 * b;
                        const double u0 = pu[0], u1 = pu[1], u2 = pu[2];
#pragma unroll
                        for (int t = 0; t < 3; t++)
                            ue[6 * i + 3 * b + t] =
                                g.r[t] * u0 + g.r[3 + t] * u1 + g.r[6 + t] * u2;
                    }
            } else {
                const double *ps = a.stress + (a.stress_per_elem ? 3 * e : 0);
                Nf[0] = ps[0]; Nf[1] = ps[1]; Nf[2] = ps[2];
            }
            if (ET == TRIA3R) {
                double Nx[3], Ny[3];
                tria_grad(g, Nx, Ny);
                if (MAT == KG) membrane_forces<3>(s, Nx, Ny, ue, Nf);
                const double w = 0.5 * 2 * g.area;
#pragma unroll
                for (int I = 0; I < 3; I++)
#pragma unroll
                    for (int J = 0; J < 3; J++)
                        P[I * 3 + J] = w * (Nx[J] * (Nx[I] * Nf[0] + Ny[I] * Nf[2]) +
                                            Ny[J] * (Nx[I] * Nf[2] + Ny[I] * Nf[1]));
            } else {
#pragma unroll
                for (int p = 0; p < 4; p++) {
                    QShape q;
                    quad_shape(g.xe, (p >> 1) ? GP2 : -GP2, (p & 1) ? GP2 : -GP2, q);
                    if (MAT == KG) membrane_forces<4>(s, q.Nx, q.Ny, ue, Nf);
#pragma unroll
                    for (int I = 0; I < 4; I++) {
                        const double fx = q.Nx[I] * Nf[0] + q.Ny[I] * Nf[2];
                        const double fy = q.Nx[I] * Nf[2] + q.Ny[I] * Nf[1];
#pragma unroll
                        for (int J = 0; J < 4; J++)
                            P[I * 4 + J] += q.det * (q.Nx[J] * fx + q.Ny[J] * fy);
                    }
                }
            }
            }
#pragma unroll
            for (int k = 0; k < NP; k++) sd[lane][k] = P[k];
            // element normal: third local axis in global coordinates
            sd[lane][NP + 0] = g.r[2];
            sd[lane][NP + 1] = g.r[5];
            sd[lane][NP + 2] = g.r[8];
        } else {
            // mass: quad4.pyx:3083-9489, tria3r.pyx:4063-7723
            if (ET == TRIA3R) {
                const double detJ = 2 * g.area;
                const double pc[3][3] = {{2. / 3, 1. / 6, 1. / 6}, {1. / 6, 1. / 6, 2. / 3},
                                         {1. / 6, 2. / 3, 1. / 6}};
                const double pl[3][3] = {{1, 0, 0}, {0, 0, 1}, {0, 1, 0}};
#pragma unroll
                for (int I = 0; I < 3; I++)
#pragma unroll
                    for (int J = 0; J < 3; J++) {
                        double H = 0.;
                        if (mtype == 1) H = detJ / 18.;
                        else
#pragma unroll
                            for (int t = 0; t < 3; t++) {
                                const double ni = mtype == 0 ? pc[t][I] : pl[t][I];
                                const double nj = mtype == 0 ? pc[t][J] : pl[t][J];
                                H += ni * nj * detJ * (0.5 / 3.);
                            }
                        P[I * 3 + J] = H;
                    }
            } else if (mtype == 1) {
#pragma unroll
                for (int k = 0; k < 16; k++) P[k] = 0.0625 * g.area;
            } else {
                const double pt = mtype == 0 ? GP2 : 1.0;
#pragma unroll
                for (int p = 0; p < 4; p++) {
                    QShape q;
                    quad_shape(g.xe, (p >> 1) ? pt : -pt, (p & 1) ? pt : -pt, q);
#pragma unroll
                    for (int I = 0; I < 4; I++)
#pragma unroll
                        for (int J = 0; J < 4; J++) P[I * 4 + J] += q.N[I] * q.N[J] * q.det;
                }
            }
            double M[6][6];
#pragma unroll
            for (int i = 0; i < 6; i++)
#pragma unroll
                for (int j = 0; j < 6; j++) M[i][j] = 0.;
            M[0][0] = M[1][1] = M[2][2] = s.intrho;
            M[3][3] = M[4][4] = s.intrhoz2;
            if (mtype != 2) {
                M[0][4] = M[4][0] = s.intrhoz;
                M[1][3] = M[3][1] = -s.intrhoz;
            }
            rotate_block(M, g.r);
#pragma unroll
            for (int k = 0; k < NP; k++) sd[lane][k] = P[k];
#pragma unroll
            for (int i = 0; i < 6; i++)
#pragma unroll
                for (int j = 0; j < 6; j++) sd[lane][NP + 6 * i + j] = M[i][j];
        }
    }
    __syncthreads();
    const int nel = (int)((a.n_elem - e0) < EPB ? (a.n_elem - e0) : EPB);
    const bool write_idx = !(a.flags & F_VALUES_ONLY);
    const bool overwrite = (a.flags & F_OVERWRITE);
    // slot k of an element -> node pair (I, J) and DOFs (dr, dc) in its 6x6 block
    auto slot_map = [&](int k, int &I, int &J, int &dr, int &dc) {
        if (kg_like(MAT)) {  // (I, a, J, b) over translations
            const int row = k / (3 * NN), col = k - row * 3 * NN;
            I = row / 3; dr = row - 3 * I;
            J = col / 3; dc = col - 3 * J;
        } else if (mtype == 2) {  // lumped: (I, g, d, J, t)
            const int Ib = k / (18 * NN), rem = k - Ib * 18 * NN;
            const int gg = rem / (9 * NN), rem2 = rem - gg * 9 * NN;
            const int d = rem2 / (3 * NN), col = rem2 - d * 3 * NN;
            I = Ib; dr = 3 * gg + d;
            J = col / 3; dc = 3 * gg + (col - 3 * J);
        } else {  // row d skips column (d + 3) mod 6
            const int row = k / (5 * NN), col = k - row * 5 * NN;
            I = row / 6; dr = row - 6 * I;
            J = col / 5;
            const int tt = col - 5 * J, skip = (dr + 3) % 6;
            dc = tt < skip ? tt : tt + 1;
        }
    };
    auto slot_val = [&](int el, int I, int J, int dr, int dc) {
        const double *d = sd[el];
        if (kg_like(MAT)) return (d[NP + dr] * d[I * NN + J]) * d[NP + dc];
        return d[I * NN + J] * d[NP + 6 * dr + dc];
    };
    const int total = nel * nslot;
    if (NN == 4 && (a.flags & F_VEC)) {
        // pairs of consecutive slots (every row of these layouts has an even
        // length, so a pair never crosses a row): one 16-byte store per pair;
        // F_VEC: contiguous 16-byte aligned blocks of an even slot count
        for (int t2 = lane; 2 * t2 < total; t2 += 64) {
            const int t = 2 * t2;
            const int el = t / nslot, k = t - el * nslot;
            int I0, J0, dr0, dc0, I1, J1, dr1, dc1;
            slot_map(k, I0, J0, dr0, dc0);
            slot_map(k + 1, I1, J1, dr1, dc1);
            const dbl2 val = {slot_val(el, I0, J0, dr0, dc0), slot_val(el, I1, J1, dr1, dc1)};
            const int64_t slot = sk0[el] + k;
            dbl2 *dst = reinterpret_cast<dbl2 *>(a.v + slot);
            if (overwrite) store2(dst, val);
            else *dst = *dst + val;
            if (write_idx) {
                store2(reinterpret_cast<lng2 *>(a.r + slot),
                       lng2{sconn[el][I0] + dr0, sconn[el][I1] + dr1});
                store2(reinterpret_cast<lng2 *>(a.c + slot),
                       lng2{sconn[el][J0] + dc0, sconn[el][J1] + dc1});
            }
        }
        return;
    }
    // consecutive lanes write consecutive slots of the block's elements
    for (int t = lane; t < total; t += 64) {
        const int el = t / nslot, k = t - el * nslot;
        int I, J, dr, dc;
        slot_map(k, I, J, dr, dc);
        const double val = slot_val(el, I, J, dr, dc);
        const int64_t slot = sk0[el] + k;
        if (overwrite) __builtin_nontemporal_store(val, a.v + slot);
        else a.v[slot] += val;
        if (write_idx) {
            a.r[slot] = sconn[el][I] + dr;
            a.c[slot] = sconn[el][J] + dc;
        }
    }
}

// ------------------------------------------------------------------------
// KC0 with block symmetry.  K_e is symmetric, so a lane owns one node-pair
// block (I, J) with I <= J: 10 lanes per quad (6 quads per wave), 6 lanes
// per triangle (10 per wave).  The element matrix leaves in four (three)
// phases, one per node row P: the lanes owning (P, J) write their block and
// the lanes owning (J, P), J < P, write its transpose into a 6 x 6nn LDS
// strip, and the wave streams the strip, which is one contiguous run of
// 6*6nn slots of the reference layout (quad4.pyx:1287-1293).
// ------------------------------------------------------------------------
template <int NN>
__device__ __forceinline__ void upper_pair(int sub, int &I, int &J)
{
    // sub -> (I, J), I <= J, row-major over the upper triangle
    I = 0;
#pragma unroll
    for (int t = 1; t < NN; t++) I += (sub >= t * NN - (t * (t - 1)) / 2) ? 1 : 0;
    J = I + sub - (I * NN - (I * (I - 1)) / 2);
}


template <int ET, int MAT>
__global__ void __launch_bounds__(64) shell_stream(AsmArgs a)
{
    __shared__ __align__(16) char lds[StreamLds<ET, MAT>::BYTES];
    shell_stream_body<ET, MAT>(a, blockIdx.x, lds);
}

// waves per SIMD the register allocation of the KC0 kernel targets
#ifndef PF3D_KC0_WAVES
#define PF3D_KC0_WAVES 3
#endif

template <int ET>
struct Kc0Lds {
    static constexpr int NN = (ET == TRIA3R) ? 3 : 4;
    static constexpr int NUP = NN * (NN + 1) / 2;
    static constexpr int EPW = 64 / NUP;
    static constexpr int NDOF = 6 * NN;
    static constexpr int RW = 6 * NDOF;  // slots per node row
    static constexpr int BYTES = EPW * (RW * 8 + NN * 8 + 8);
};

template <int ET, int FLAGS>
__device__ __forceinline__ void shell_kc0_body(const AsmArgs &a, int64_t blk, char *lds)
{
    using L = Kc0Lds<ET>;
    constexpr int NN = L::NN, NUP = L::NUP, EPW = L::EPW, NDOF = L::NDOF, RW = L::RW;
    double (*strip)[RW] = reinterpret_cast<double (*)[RW]>(lds);
    int64_t (*sconn)[NN] = reinterpret_cast<int64_t (*)[NN]>(lds + EPW * RW * 8);
    int64_t *sk0 = reinterpret_cast<int64_t *>(lds + EPW * (RW + NN) * 8);

    const int lane = threadIdx.x;
    const int slot = lane / NUP, sub = lane - slot * NUP;
    int I, J;
    upper_pair<NN>(sub, I, J);
    const int64_t e0 = blk * EPW;
    const int64_t e = e0 + slot;
    const bool active = slot < EPW && e < a.n_elem;

    double K[6][6];
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = 0; j < 6; j++) K[i][j] = 0.;
    if (active) {
        V3 X[NN];
        int64_t cn[NN];
#pragma unroll
        for (int i = 0; i < NN; i++) {
            cn[i] = a.conn[e * NN + i];
            const double *px = a.x + cn[i] / 2;
            X[i] = {px[0], px[1], px[2]};
        }
        if (sub == 0) {
#pragma unroll
            for (int i = 0; i < NN; i++) sconn[slot][i] = cn[i];
            sk0[slot] = a.init_k ? a.init_k[e] : (int64_t)(36 * NN * NN) * e;
        }
        V3 xm = {0., 0., 0.};
        if (a.evec) xm = {a.evec[3 * e], a.evec[3 * e + 1], a.evec[3 * e + 2]};
        Geom g;
        if (ET == TRIA3R) tria_geometry(X, xm, g);
        else quad_geometry(X, xm, g);
        Shell s;
        load_shell(a.props + (a.pid ? (int64_t)a.pid[e] : 0) * SHELL_STRIDE, g.m, s);
        const double k6 = a.k6rot ? a.k6rot[e] : 100.;
        if (ET == QUAD4) quad4_kc0(K, g, s, I, J);
        else if (ET == QUAD4R) {
            double hgf[5] = {1., 1., 1., 1., 1.};
            if (a.hg)
#pragma unroll
                for (int d = 0; d < 5; d++) hgf[d] = a.hg[5 * e + d];
            quad4r_kc0(K, g, s, I, J, k6, hgf);
        } else {
            tria3r_kc0(K, g, s, I, J, k6, a.alpha ? a.alpha[e] : 0.7);
        }
        rotate_block(K, g.r);
    }
    const int nel = (int)((a.n_elem - e0) < EPW ? (a.n_elem - e0) : EPW);
    constexpr bool write_idx = !(FLAGS & F_VALUES_ONLY);
    constexpr bool overwrite = (FLAGS & F_OVERWRITE);
    const bool vec_ok = (reinterpret_cast<uintptr_t>(a.v) & 15) == 0;
    for (int P = 0; P < NN; P++) {
        __syncthreads();  // the previous strip has been streamed out
        if (active && I == P) {
            double *t = &strip[slot][6 * J];
#pragma unroll
            for (int i = 0; i < 6; i++)
#pragma unroll
                for (int j = 0; j < 6; j++) t[i * NDOF + j] = K[i][j];
        } else if (active && J == P) {  // block (P, I) = K_IJ^T
            double *t = &strip[slot][6 * I];
#pragma unroll
            for (int i = 0; i < 6; i++)
#pragma unroll
                for (int j = 0; j < 6; j++) t[j * NDOF + i] = K[i][j];
        }
        __syncthreads();
        // pairs of consecutive slots (same row, same node column block):
        // one 16-byte LDS read and one 16-byte store per pair when the slot
        // is 16-byte aligned; unrolled so the LDS reads issue back to back
        constexpr int PR = RW / 2;                 // pairs per element strip
        constexpr int NIT = (EPW * PR + 63) / 64;  // pair loads per lane
#pragma unroll
        for (int it = 0; it < NIT; it++) {
            const int pi = lane + 64 * it;
            if (pi >= nel * PR) break;
            const int el = pi / PR, t = 2 * (pi - el * PR);
            const int64_t k = sk0[el] + P * RW + t;
            const dbl2 val = *reinterpret_cast<const dbl2 *>(&strip[el][t]);
            if (FLAGS & F_VEC) {
                // slots 16-byte aligned (init_k = size*e, aligned v)
                dbl2 *dst = reinterpret_cast<dbl2 *>(a.v + k);
                if (overwrite) store2(dst, val);
                else store2(dst, *dst + val);
            } else if (((k & 1) == 0) && vec_ok) {
                dbl2 *dst = reinterpret_cast<dbl2 *>(a.v + k);
                if (overwrite) *dst = val;
                else *dst = *dst + val;
            } else {
                if (overwrite) { a.v[k] = val.x; a.v[k + 1] = val.y; }
                else { a.v[k] += val.x; a.v[k + 1] += val.y; }
            }
            if (write_idx) {
                const int rl = t / NDOF, cl = t - rl * NDOF;
                const int64_t rr = sconn[el][P] + rl;
                const int64_t cc = sconn[el][cl / 6] + cl % 6;
                if (FLAGS & F_VEC) {
                    store2(reinterpret_cast<lng2 *>(a.r + k), lng2{rr, rr});
                    store2(reinterpret_cast<lng2 *>(a.c + k), lng2{cc, cc + 1});
                } else {
                    a.r[k] = rr; a.r[k + 1] = rr;
                    a.c[k] = cc; a.c[k + 1] = cc + 1;
                }
            }
        }
    }
}

template <int ET, int FLAGS>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PF3D_KC0_WAVES)))
shell_kc0_sym(AsmArgs a)
{
    __shared__ __align__(16) char lds[Kc0Lds<ET>::BYTES];
    shell_kc0_body<ET, FLAGS>(a, blockIdx.x, lds);
}

// ------------------------------------------------------------------------
// Element internal forces in global axes, fe[e][6nn] = R6 K_e ue with
// ue = blockdiag(R^T) u (update_probe_ue -> update_probe_finte -> rotation
// in update_fint: quad4.pyx:627-680, 1174-1202, 1316-1363; quad4r.pyx:549,
// 4620; tria3r.pyx:549, 2975).  Same lane mapping as shell_kc0_sym: the lane
// owning block (I, J), I <= J, adds K_IJ ue_J to node I and K_IJ^T ue_I to
// node J; the node sums are formed in a fixed lane order.
// ------------------------------------------------------------------------
template <int ET>
__global__ void __launch_bounds__(64) shell_fe(AsmArgs a, double *__restrict__ fe)
{
    constexpr int NN = (ET == TRIA3R) ? 3 : 4;
    constexpr int NUP = NN * (NN + 1) / 2;
    constexpr int EPW = 64 / NUP;
    __shared__ double part[EPW][NUP][2][6];
    __shared__ double sr[EPW][9];
    const int lane = threadIdx.x;
    const int slot = lane / NUP, sub = lane - slot * NUP;
    int I, J;
    upper_pair<NN>(sub, I, J);
    const int64_t e = (int64_t)blockIdx.x * EPW + slot;
    const bool active = slot < EPW && e < a.n_elem;
    if (active) {
        V3 X[NN];
        int64_t cn[NN];
#pragma unroll
        for (int i = 0; i < NN; i++) {
            cn[i] = a.conn[e * NN + i];
            const double *px = a.x + cn[i] / 2;
            X[i] = {px[0], px[1], px[2]};
        }
        V3 xm = {0., 0., 0.};
        if (a.evec) xm = {a.evec[3 * e], a.evec[3 * e + 1], a.evec[3 * e + 2]};
        Geom g;
        if (ET == TRIA3R) tria_geometry(X, xm, g);
        else quad_geometry(X, xm, g);
        Shell s;
        load_shell(a.props + (a.pid ? (int64_t)a.pid[e] : 0) * SHELL_STRIDE, g.m, s);
        double K[6][6];
#pragma unroll
        for (int i = 0; i < 6; i++)
#pragma unroll
            for (int j = 0; j < 6; j++) K[i][j] = 0.;
        const double k6 = a.k6rot ? a.k6rot[e] : 100.;
        if (ET == QUAD4) quad4_kc0(K, g, s, I, J);
        else if (ET == QUAD4R) {
            double hgf[5] = {1., 1., 1., 1., 1.};
            if (a.hg)
#pragma unroll
                for (int d = 0; d < 5; d++) hgf[d] = a.hg[5 * e + d];
            quad4r_kc0(K, g, s, I, J, k6, hgf);
        } else {
            tria3r_kc0(K, g, s, I, J, k6, a.alpha ? a.alpha[e] : 0.7, true);
        }
        // local displacements of nodes I and J
        int64_t ci = cn[0], cj = cn[0];
#pragma unroll
        for (int i = 1; i < NN; i++) {
            ci = (I == i) ? cn[i] : ci;
            cj = (J == i) ? cn[i] : cj;
        }
        double ui[6], uj[6];
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int t = 0; t < 3; t++) {
                double si = 0., sj = 0.;
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    si += g.r[3 * q + t] * a.u[ci + 3 * b + q];
                    sj += g.r[3 * q + t] * a.u[cj + 3 * b + q];
                }
                ui[3 * b + t] = si;
                uj[3 * b + t] = sj;
            }
#pragma unroll
        for (int i = 0; i < 6; i++) {
            double fi = 0., fj = 0.;
#pragma unroll
            for (int j = 0; j < 6; j++) {
                fi += K[i][j] * uj[j];
                fj += K[j][i] * ui[j];
            }
            part[slot][sub][0][i] = fi;
            part[slot][sub][1][i] = (I != J) ? fj : 0.;
        }
        if (sub == 0)
#pragma unroll
            for (int i = 0; i < 9; i++) sr[slot][i] = g.r[i];
    }
    __syncthreads();
    if (active && sub < NN) {
        const int P = sub;
        double f[6] = {0., 0., 0., 0., 0., 0.};
        for (int t = 0; t < NUP; t++) {
            int It, Jt;
            upper_pair<NN>(t, It, Jt);
            if (It == P)
                for (int i = 0; i < 6; i++) f[i] += part[slot][t][0][i];
            else if (Jt == P)
                for (int i = 0; i < 6; i++) f[i] += part[slot][t][1][i];
        }
        double *o = fe + (e * NN + P) * 6;
        for (int b = 0; b < 2; b++)
            for (int q = 0; q < 3; q++)
                o[3 * b + q] = sr[slot][3 * q] * f[3 * b] + sr[slot][3 * q + 1] * f[3 * b + 1] +
                               sr[slot][3 * q + 2] * f[3 * b + 2];
    }
}

// Geometry of each element: r (9), m (4), area/length, local x/y (2*nn).
template <int ET>
__global__ void element_geometry(int64_t n, const double *x, const int64_t *conn,
                                 const double *evec, double *geom, double *xe)
{
    constexpr int NN = ET == TRIA3R ? 3 : ((ET == BEAMC || ET == BEAMLR || ET == TRUSS || ET == SPRING) ? 2 : 4);
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    V3 X[NN];
#pragma unroll
    for (int i = 0; i < NN; i++) {
        const double *px = x + conn[e * NN + i] / 2;
        X[i] = {px[0], px[1], px[2]};
    }
    V3 ev = {0., 0., 0.};
    if (evec && ET != TRUSS && ET != SPRING) ev = {evec[3 * e], evec[3 * e + 1], evec[3 * e + 2]};
    Geom g;
    if (ET == TRIA3R) tria_geometry(X, ev, g);
    else if (ET == BEAMC || ET == BEAMLR) beam_geometry(X, ev, g);
    else if (ET == TRUSS || ET == SPRING) rod_geometry<ET>(X, evec ? evec + 6 * e : nullptr, g);
    else quad_geometry(X, ev, g);
    double *o = geom + 14 * e;
    for (int i = 0; i < 9; i++) o[i] = g.r[i];
    for (int i = 0; i < 4; i++) o[9 + i] = g.m[i];
    o[13] = g.area;
    if (xe) {
#pragma clang fp contract(off)
        for (int i = 0; i < NN; i++) {
            double xl = 0, yl = 0, zl = 0;
            xl = g.r[0] * X[i].x + g.r[3] * X[i].y + g.r[6] * X[i].z;
            yl = g.r[1] * X[i].x + g.r[4] * X[i].y + g.r[7] * X[i].z;
            zl = g.r[2] * X[i].x + g.r[5] * X[i].y + g.r[8] * X[i].z;
            xe[(3 * NN) * e + 3 * i] = xl;
            xe[(3 * NN) * e + 3 * i + 1] = yl;
            xe[(3 * NN) * e + 3 * i + 2] = zl;
        }
    }
}

// ------------------------------------------------------------------------
// BeamC: 16 beams per wave, 4 lanes per beam.  The first lane of a beam
// writes the closed-form 12x12 local matrix into LDS; each lane then rotates
// one 6x6 node block (R6 K R6^T) and the wave streams the slots out.
// ------------------------------------------------------------------------
struct BeamSec {
    double A, E, G, Iyy, Izz, Iyz, J, Ay, Az, ay, az, by, bz;
};

__device__ __forceinline__ BeamSec beam_section(const double *bp, double L)
{
    BeamSec s;
    s.A = bp[0]; s.E = bp[1]; s.G = bp[2]; s.Iyy = bp[3]; s.Izz = bp[4]; s.Iyz = bp[5];
    s.J = bp[6]; s.Ay = bp[7]; s.Az = bp[8];
    s.ay = 12 * s.E * s.Izz / (s.G * s.A * L * L);
    s.az = 12 * s.E * s.Iyy / (s.G * s.A * L * L);
    s.by = 1 / (1. - s.ay);
    s.bz = 1 / (1. - s.az);
    return s;
}

// fill the symmetric 12x12 T from its upper triangle U (row-major, U[i][j], j >= i)
__device__ __forceinline__ void beam_sym(double *T)
{
    for (int i = 1; i < 12; i++)
        for (int j = 0; j < i; j++) T[i * 12 + j] = T[j * 12 + i];
}

// KC0: beamc.pyx:543-624
__device__ void beam_kc0_local(double L, const BeamSec &s, double *U)
{
    const double A = s.A, E = s.E, G = s.G, Iyy = s.Iyy, Izz = s.Izz, Jt = s.J;
    const double Ay = s.Ay, Az = s.Az, ay = s.ay, az = s.az, by = s.by, bz = s.bz;
    const double L2 = L * L, L3 = L2 * L;
    const double kyy = by * by * (A * G * L2 * ay * ay + 12 * E * Izz);
    const double kzz = bz * bz * (A * G * L2 * az * az + 12 * E * Iyy);
    const double cyz = E * s.Iyz * by * bz;
    const double ta = Az * G * ay * by, tb = Ay * G * az * bz;
    for (int k = 0; k < 144; k++) U[k] = 0.;
#define B(i, j) U[(i) * 12 + (j)]
    B(0, 0) = A * E / L; B(0, 6) = -A * E / L; B(6, 6) = A * E / L;
    B(0, 4) = Az * E * bz * (1 - az) / L; B(0, 10) = Az * E * bz * (az - 1) / L;
    B(0, 5) = Ay * E * by * (ay - 1) / L; B(0, 11) = Ay * E * by * (1 - ay) / L;
    B(4, 6) = Az * E * bz * (az - 1) / L; B(6, 10) = Az * E * bz * (1 - az) / L;
    B(5, 6) = Ay * E * by * (1 - ay) / L; B(6, 11) = Ay * E * by * (ay - 1) / L;
    B(1, 1) = kyy / L3; B(1, 7) = -kyy / L3; B(7, 7) = kyy / L3;
    B(1, 5) = kyy / (2 * L2); B(1, 11) = kyy / (2 * L2);
    B(5, 7) = -kyy / (2 * L2); B(7, 11) = -kyy / (2 * L2);
    B(2, 2) = kzz / L3; B(2, 8) = -kzz / L3; B(8, 8) = kzz / L3;
    B(2, 4) = -kzz / (2 * L2); B(2, 10) = -kzz / (2 * L2);
    B(4, 8) = kzz / (2 * L2); B(8, 10) = kzz / (2 * L2);
    B(1, 2) = 12 * cyz / L3; B(1, 8) = -12 * cyz / L3;
    B(2, 7) = -12 * cyz / L3; B(7, 8) = 12 * cyz / L3;
    B(1, 4) = -6 * cyz / L2; B(1, 10) = -6 * cyz / L2;
    B(2, 5) = 6 * cyz / L2; B(2, 11) = 6 * cyz / L2;
    B(4, 7) = 6 * cyz / L2; B(5, 8) = -6 * cyz / L2;
    B(7, 10) = 6 * cyz / L2; B(8, 11) = -6 * cyz / L2;
    B(3, 3) = G * Jt / L; B(3, 9) = -G * Jt / L; B(9, 9) = G * Jt / L;
    B(1, 3) = ta / L; B(1, 9) = -ta / L; B(3, 7) = -ta / L; B(7, 9) = ta / L;
    B(2, 3) = -tb / L; B(2, 9) = tb / L; B(3, 8) = tb / L; B(8, 9) = -tb / L;
    B(3, 4) = tb / 2; B(3, 10) = tb / 2; B(4, 9) = -tb / 2; B(9, 10) = -tb / 2;
    B(3, 5) = ta / 2; B(3, 11) = ta / 2; B(5, 9) = -ta / 2; B(9, 11) = -ta / 2;
    const double ryy = bz * bz * (A * G * L2 * az * az / 4 + E * Iyy * az * az - 2 * E * Iyy * az + 4 * E * Iyy) / L;
    const double ryy2 = bz * bz * (A * G * L2 * az * az / 4 - E * Iyy * az * az + 2 * E * Iyy * az + 2 * E * Iyy) / L;
    const double rzz = by * by * (A * G * L2 * ay * ay / 4 + E * Izz * ay * ay - 2 * E * Izz * ay + 4 * E * Izz) / L;
    const double rzz2 = by * by * (A * G * L2 * ay * ay / 4 - E * Izz * ay * ay + 2 * E * Izz * ay + 2 * E * Izz) / L;
    B(4, 4) = ryy; B(10, 10) = ryy; B(4, 10) = ryy2;
    B(5, 5) = rzz; B(11, 11) = rzz; B(5, 11) = rzz2;
    const double c1 = cyz * (-ay * az + ay + az - 4) / L, c2 = cyz * (ay * az - ay - az - 2) / L;
    B(4, 5) = c1; B(10, 11) = c1; B(4, 11) = c2; B(5, 10) = c2;
#undef B
    beam_sym(U);
}

// KG for axial force N: beamc.pyx:1889-2180.  v, rz bend in y and w, ry in
// z; with s_rot(ry) = -1, s_rot(rz) = +1 and each DOF's own alpha/beta:
//   disp-disp +-(5 a_i a_j - 5 a_i - 5 a_j + 6)/(5L) (+ on the same node)
//   disp-rot  s_rot (+1 / -1 for the disp DOF on node 1 / 2) / 10
//   rot-rot   s_rot_i s_rot_j L ((q + 8)/60 same node, -(q + 2)/60 else)
// all times N beta_i beta_j.
__device__ void beam_kg_local(double L, const BeamSec &s, double N, double *U)
{
    for (int i = 0; i < 12; i++)
        for (int j = 0; j < 12; j++) {
            const int di = i % 6, dj = j % 6, ni = i / 6, nj = j / 6;
            const bool act_i = di == 1 || di == 2 || di == 4 || di == 5;
            const bool act_j = dj == 1 || dj == 2 || dj == 4 || dj == 5;
            double v = 0.;
            if (act_i && act_j) {
                const bool yi = (di == 1 || di == 5), yj = (dj == 1 || dj == 5);
                const double a_i = yi ? s.ay : s.az, a_j = yj ? s.ay : s.az;
                const double bb = N * (yi ? s.by : s.bz) * (yj ? s.by : s.bz);
                const bool ri = di >= 4, rj = dj >= 4;
                const double si = di == 4 ? -1. : 1., sj = dj == 4 ? -1. : 1.;
                const double q = 5 * a_i * a_j - 5 * a_i - 5 * a_j;
                double f;
                if (!ri && !rj) f = (ni == nj ? 1. : -1.) * (q + 6) / (5 * L);
                else if (!ri && rj) f = sj * (ni == 0 ? 1. : -1.) / 10;
                else if (ri && !rj) f = si * (nj == 0 ? 1. : -1.) / 10;
                else f = si * sj * L * (ni == nj ? (q + 8) / 60 : -(q + 2) / 60);
                v = bb * f;
            }
            U[i * 12 + j] = v;
        }
}

// M: consistent (mtype 0, beamc.pyx:2246-2968), lumped (mtype 1, :2969-3152)
__device__ void beam_m_local(double L, const BeamSec &s, const double *bp, int mtype, double *U)
{
    const double r = bp[9], ry = bp[10], rz = bp[11], ry2 = bp[12], rz2 = bp[13], ryz = bp[14];
    const double ay = s.ay, az = s.az, by = s.by, bz = s.bz, L2 = L * L;
    for (int k = 0; k < 144; k++) U[k] = 0.;
#define B(i, j) U[(i) * 12 + (j)]
    if (mtype == 1) {
        for (int n = 0; n < 2; n++) {
            const int o = 6 * n;
            B(o, o) = L * r / 2;
            B(o + 1, o + 1) = L * by * by * r * (ay * ay - 2 * ay + 1) / 2;
            B(o + 2, o + 2) = L * bz * bz * r * (az * az - 2 * az + 1) / 2;
            B(o + 3, o + 3) = L * (ry2 + rz2) / 2;
            B(o + 4, o + 4) = L * bz * bz * rz2 * (az * az - 2 * az + 1) / 2;
            B(o + 5, o + 5) = L * by * by * ry2 * (ay * ay - 2 * ay + 1) / 2;
        }
        return;
    }
    const double tyy = by * by * (70 * L2 * ay * ay * r - 147 * L2 * ay * r + 78 * L2 * r + 252 * ry2) / (210 * L);
    const double tzz = bz * bz * (70 * L2 * az * az * r - 147 * L2 * az * r + 78 * L2 * r + 252 * rz2) / (210 * L);
    const double tyy2 = by * by * (35 * L2 * ay * ay * r - 63 * L2 * ay * r + 27 * L2 * r - 252 * ry2) / (210 * L);
    const double tzz2 = bz * bz * (35 * L2 * az * az * r - 63 * L2 * az * r + 27 * L2 * r - 252 * rz2) / (210 * L);
    const double cy1 = by * by * (35 * L2 * ay * ay * r - 77 * L2 * ay * r + 44 * L2 * r + 420 * ay * ry2 + 84 * ry2) / 840;
    const double cy2 = by * by * (-35 * L2 * ay * ay * r + 63 * L2 * ay * r - 26 * L2 * r + 420 * ay * ry2 + 84 * ry2) / 840;
    const double cz1 = bz * bz * (-35 * L2 * az * az * r + 77 * L2 * az * r - 44 * L2 * r - 420 * az * rz2 - 84 * rz2) / 840;
    const double cz2 = bz * bz * (35 * L2 * az * az * r - 63 * L2 * az * r + 26 * L2 * r - 420 * az * rz2 - 84 * rz2) / 840;
    const double qzz = L * bz * bz * (7 * L2 * az * az * r - 14 * L2 * az * r + 8 * L2 * r + 280 * az * az * rz2 - 140 * az * rz2 + 112 * rz2) / 840;
    const double qyy = L * by * by * (7 * L2 * ay * ay * r - 14 * L2 * ay * r + 8 * L2 * r + 280 * ay * ay * ry2 - 140 * ay * ry2 + 112 * ry2) / 840;
    const double qzz2 = L * bz * bz * (-7 * L2 * az * az * r + 14 * L2 * az * r - 6 * L2 * r + 140 * az * az * rz2 + 140 * az * rz2 - 28 * rz2) / 840;
    const double qyy2 = L * by * by * (-7 * L2 * ay * ay * r + 14 * L2 * ay * r - 6 * L2 * r + 140 * ay * ay * ry2 + 140 * ay * ry2 - 28 * ry2) / 840;
    const double pyz = 6 * by * bz * ryz / (5 * L);
    const double sz1 = by * bz * ryz * (5 * az + 1) / 10, sy1 = by * bz * ryz * (5 * ay + 1) / 10;
    const double qyz = L * by * bz * ryz * (-20 * ay * az + 5 * ay + 5 * az - 8) / 60;
    const double qyz2 = L * by * bz * ryz * (-10 * ay * az - 5 * ay - 5 * az + 2) / 60;
    B(0, 0) = L * r / 3; B(0, 6) = L * r / 6; B(6, 6) = L * r / 3;
    B(0, 1) = by * ry / 2; B(0, 7) = -by * ry / 2; B(1, 6) = by * ry / 2; B(6, 7) = -by * ry / 2;
    B(0, 2) = bz * rz / 2; B(0, 8) = -bz * rz / 2; B(2, 6) = bz * rz / 2; B(6, 8) = -bz * rz / 2;
    B(0, 4) = L * bz * rz * (1 - 4 * az) / 12; B(6, 10) = B(0, 4);
    B(0, 5) = L * by * ry * (4 * ay - 1) / 12; B(6, 11) = B(0, 5);
    B(0, 10) = -L * bz * rz * (2 * az + 1) / 12; B(4, 6) = B(0, 10);
    B(0, 11) = L * by * ry * (2 * ay + 1) / 12; B(5, 6) = B(0, 11);
    B(1, 1) = tyy; B(7, 7) = tyy; B(1, 7) = tyy2;
    B(2, 2) = tzz; B(8, 8) = tzz; B(2, 8) = tzz2;
    B(1, 2) = pyz; B(7, 8) = pyz; B(1, 8) = -pyz; B(2, 7) = -pyz;
    B(1, 5) = cy1; B(7, 11) = -cy1; B(1, 11) = cy2; B(5, 7) = -cy2;
    B(2, 4) = cz1; B(8, 10) = -cz1; B(2, 10) = cz2; B(4, 8) = -cz2;
    B(1, 4) = -sz1; B(1, 10) = -sz1; B(4, 7) = sz1; B(7, 10) = sz1;
    B(2, 5) = sy1; B(2, 11) = sy1; B(5, 8) = -sy1; B(8, 11) = -sy1;
    B(1, 3) = L * by * rz * (20 * ay - 21) / 60; B(7, 9) = B(1, 3);
    B(1, 9) = L * by * rz * (10 * ay - 9) / 60; B(3, 7) = B(1, 9);
    B(2, 3) = L * bz * ry * (21 - 20 * az) / 60; B(8, 9) = B(2, 3);
    B(2, 9) = L * bz * ry * (9 - 10 * az) / 60; B(3, 8) = B(2, 9);
    B(3, 3) = L * (ry2 + rz2) / 3; B(9, 9) = B(3, 3); B(3, 9) = L * (ry2 + rz2) / 6;
    B(3, 4) = L2 * bz * ry * (5 * az - 6) / 120; B(9, 10) = -B(3, 4);
    B(3, 5) = L2 * by * rz * (5 * ay - 6) / 120; B(9, 11) = -B(3, 5);
    B(3, 10) = L2 * bz * ry * (4 - 5 * az) / 120; B(4, 9) = -B(3, 10);
    B(3, 11) = L2 * by * rz * (4 - 5 * ay) / 120; B(5, 9) = -B(3, 11);
    B(4, 4) = qzz; B(10, 10) = qzz; B(4, 10) = qzz2;
    B(5, 5) = qyy; B(11, 11) = qyy; B(5, 11) = qyy2;
    B(4, 5) = qyz; B(10, 11) = qyz; B(4, 11) = qyz2; B(5, 10) = qyz2;
#undef B
    beam_sym(U);
}

// BeamLR (beamlr.pyx): linear Timoshenko beam, all fields linear, one-point
// (reduced) integration.  Midpoint strains eps = u', ky = ry', kz = rz',
// gxy = v' - rz_m, gxz = w' + ry_m, phi = rx'; section stiffness
// D = E [[A, Az, -Ay], [Az, Iyy, -Iyz], [-Ay, -Iyz, Izz]] on (eps, ky, kz),
// S = G [[A, 0, -Az], [0, A, Ay], [-Az, Ay, J]] on (gxy, gxz, phi);
// K = L (Bm^T D Bm + Bs^T S Bs) (beamlr.pyx:354-406 internal forces K ue,
// 407-1187 the scattered K).  Each strain row has 2 or 4 terms.
__device__ void beamlr_kc0_local(double L, const double *bp, double *U)
{
    const double A = bp[0], E = bp[1], G = bp[2], Iyy = bp[3], Izz = bp[4], Iyz = bp[5];
    const double J = bp[6], Ay = bp[7], Az = bp[8];
    const double D[6][6] = {{E * A, E * Az, -E * Ay, 0., 0., 0.},
                            {E * Az, E * Iyy, -E * Iyz, 0., 0., 0.},
                            {-E * Ay, -E * Iyz, E * Izz, 0., 0., 0.},
                            {0., 0., 0., G * A, 0., -G * Az},
                            {0., 0., 0., 0., G * A, G * Ay},
                            {0., 0., 0., -G * Az, G * Ay, G * J}};
    const double il = 1. / L;
    // strain rows: DOF index and coefficient (4 terms, unused ones with 0)
    const int bi[6][4] = {{0, 6, 0, 6}, {4, 10, 4, 10}, {5, 11, 5, 11},
                          {1, 7, 5, 11}, {2, 8, 4, 10}, {3, 9, 3, 9}};
    const double bc[6][4] = {{-il, il, 0., 0.}, {-il, il, 0., 0.}, {-il, il, 0., 0.},
                             {-il, il, -0.5, -0.5}, {-il, il, 0.5, 0.5}, {-il, il, 0., 0.}};
    for (int k = 0; k < 144; k++) U[k] = 0.;
#pragma unroll
    for (int a = 0; a < 6; a++)
#pragma unroll
        for (int b = 0; b < 6; b++) {
            const double d = D[a][b];
            if (d == 0.) continue;
#pragma unroll
            for (int p = 0; p < 4; p++)
#pragma unroll
                for (int q = 0; q < 4; q++)
                    U[bi[a][p] * 12 + bi[b][q]] += L * d * bc[a][p] * bc[b][q];
        }
}

// BeamLR KG (beamlr.pyx:1226-1463): N int(N_I N_J) [[1, -1], [-1, 1]] on
// (ry, rz) of nodes I, J (L/3 on a node, L/6 across)
__device__ void beamlr_kg_local(double L, double N, double *U)
{
    for (int k = 0; k < 144; k++) U[k] = 0.;
    for (int I = 0; I < 2; I++)
        for (int J = 0; J < 2; J++) {
            const double f = N * L * (I == J ? 1. / 3. : 1. / 6.);
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++) U[(6 * I + 4 + a) * 12 + 6 * J + 4 + b] = a == b ? f : -f;
        }
}

// BeamLR mass (beamlr.pyx:1464-2424): section mass m of u = u0 + z ry - y rz,
// v = v0 - z rx, w = w0 + y rx; consistent: L/3, L/6 times m; lumped
// (Gauss-Lobatto at the nodes): L/2 times the diagonal of m
__device__ void beamlr_m_local(double L, const double *bp, int mtype, double *U)
{
    const double r = bp[9], ry = bp[10], rz = bp[11], ry2 = bp[12], rz2 = bp[13], ryz = bp[14];
    double m[6][6];
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) m[i][j] = 0.;
    m[0][0] = m[1][1] = m[2][2] = r;
    m[0][4] = m[4][0] = rz;
    m[0][5] = m[5][0] = -ry;
    m[1][3] = m[3][1] = -rz;
    m[2][3] = m[3][2] = ry;
    m[3][3] = ry2 + rz2;
    m[4][4] = rz2;
    m[5][5] = ry2;
    m[4][5] = m[5][4] = -ryz;
    for (int k = 0; k < 144; k++) U[k] = 0.;
    for (int I = 0; I < 2; I++)
        for (int J = 0; J < 2; J++) {
            if (mtype == 1 && I != J) continue;
            const double f = mtype == 1 ? L / 2 : L * (I == J ? 1. / 3. : 1. / 6.);
            for (int d = 0; d < 6; d++)
                for (int e = 0; e < 6; e++)
                    if (mtype != 1 || d == e) U[(6 * I + d) * 12 + 6 * J + e] = f * m[d][e];
        }
}

// slot -> (local row, local col) for beams: 12x12 row-major, or the lumped
// mass's node-diagonal 3x3 translation and rotation blocks (beamc.pyx:2969+)
__device__ __forceinline__ void beam_slot(int lumped, int k, int &rl, int &cl)
{
    if (lumped == 2) {
        // BeamLR KG: bending rotations (I, 3..5, J, 3..5), beamlr.pyx:1276-1386
        const int I = k / 18, d = (k % 18) / 6, J = (k % 6) / 3, e = k % 3;
        rl = 6 * I + 3 + d;
        cl = 6 * J + 3 + e;
    } else if (!lumped) {
        rl = k / 12;
        cl = k - rl * 12;
    } else {
        const int n = k / 18, g = (k % 18) / 9, t = k % 9;
        rl = 6 * n + 3 * g + t / 3;
        cl = 6 * n + 3 * g + t % 3;
    }
}

// beams per block BPB (4 lanes each, BPB <= 16)
template <int BPB>
struct BeamLds {
    static constexpr int BYTES = BPB * (144 * 8 + 2 * 8 + 8 + 9 * 8);
};

template <int ET, int MAT, int BPB>
__device__ __forceinline__ void beam_body(const AsmArgs &a, int64_t blk, char *lds)
{
    double (*tile)[144] = reinterpret_cast<double (*)[144]>(lds);
    int64_t (*sconn)[2] = reinterpret_cast<int64_t (*)[2]>(lds + BPB * 144 * 8);
    int64_t *sk0 = reinterpret_cast<int64_t *>(lds + BPB * 146 * 8);
    double (*sr)[9] = reinterpret_cast<double (*)[9]>(lds + BPB * 147 * 8);
    const int lane = threadIdx.x, slot = lane >> 2, I = (lane >> 1) & 1, J = lane & 1;
    const int64_t e0 = blk * BPB;
    const int64_t e = e0 + slot;
    const bool mine = slot < BPB && e < a.n_elem;
    // slot map: 0 full 12x12, 1 lumped node-diagonal, 2 BeamLR KG rotations
    const int lumped = ((MAT == MASS) && a.mtype == 1) ? 1 : ((ET == BEAMLR && MAT == KG) ? 2 : 0);
    const int nslot = lumped ? 36 : 144;
    if (mine && (lane & 3) == 0) {
        V3 X[2];
        int64_t cn[2];
        for (int i = 0; i < 2; i++) {
            cn[i] = a.conn[2 * e + i];
            const double *px = a.x + cn[i] / 2;
            X[i] = {px[0], px[1], px[2]};
        }
        sconn[slot][0] = cn[0];
        sconn[slot][1] = cn[1];
        sk0[slot] = a.init_k ? a.init_k[e] : (int64_t)nslot * e;
        V3 vxy = {a.evec[3 * e], a.evec[3 * e + 1], a.evec[3 * e + 2]};
        Geom g;
        beam_geometry(X, vxy, g);
        for (int i = 0; i < 9; i++) sr[slot][i] = g.r[i];
        const double *bp = a.props + (a.pid ? (int64_t)a.pid[e] : 0) * BEAM_STRIDE;
        double N = 0.;
        if (MAT == KG) {
            // axial force from the local axial displacements (beamc.pyx:1889,
            // beamlr.pyx:1388), without FMA contraction: N is a difference of
            // nearly equal projections, so the reference's rounding is kept
#pragma clang fp contract(off)
            double u1 = 0., u2 = 0.;
            for (int i = 0; i < 3; i++) {
                u1 += g.r[3 * i] * a.u[cn[0] + i];
                u2 += g.r[3 * i] * a.u[cn[1] + i];
            }
            N = bp[0] * bp[1] * (-u1 + u2) / g.area;
        }
        if (ET == BEAMLR) {
            if (MAT == KC0) beamlr_kc0_local(g.area, bp, tile[slot]);
            else if (MAT == KG) beamlr_kg_local(g.area, N, tile[slot]);
            else beamlr_m_local(g.area, bp, a.mtype, tile[slot]);
        } else {
            const BeamSec sec = beam_section(bp, g.area);
            if (MAT == KC0) beam_kc0_local(g.area, sec, tile[slot]);
            else if (MAT == KG) beam_kg_local(g.area, sec, N, tile[slot]);
            else beam_m_local(g.area, sec, bp, a.mtype, tile[slot]);
        }
    }
    __syncthreads();
    double K[6][6];
    if (mine) {
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) K[i][j] = tile[slot][(6 * I + i) * 12 + 6 * J + j];
        double r[9];
        for (int i = 0; i < 9; i++) r[i] = sr[slot][i];
        rotate_block(K, r);
    }
    __syncthreads();
    if (mine)
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) tile[slot][(6 * I + i) * 12 + 6 * J + j] = K[i][j];
    __syncthreads();
    const int nel = (int)((a.n_elem - e0) < BPB ? (a.n_elem - e0) : BPB);
    const bool write_idx = !(a.flags & F_VALUES_ONLY);
    const bool overwrite = (a.flags & F_OVERWRITE);
    for (int el = 0; el < nel; el++) {
        const int64_t k0 = sk0[el];
        for (int k = lane; k < nslot; k += 64) {
            int rl, cl;
            beam_slot(lumped, k, rl, cl);
            const double val = tile[el][rl * 12 + cl];
            if (overwrite) a.v[k0 + k] = val;
            else a.v[k0 + k] += val;
            if (write_idx) {
                a.r[k0 + k] = sconn[el][rl / 6] + rl % 6;
                a.c[k0 + k] = sconn[el][cl / 6] + cl % 6;
            }
        }
    }
}

template <int ET, int MAT>
__global__ void __launch_bounds__(64) beam_assemble(AsmArgs a)
{
    __shared__ __align__(16) char lds[BeamLds<16>::BYTES];
    beam_body<ET, MAT, 16>(a, blockIdx.x, lds);
}

// BeamC / BeamLR element internal forces (beamc.pyx:352-488, 1354-1391;
// beamlr.pyx:354-406, 1188-1225)
template <int ET>
__global__ void __launch_bounds__(64) beam_fe(AsmArgs a, double *__restrict__ fe)
{
    __shared__ double tile[16][144];
    __shared__ double sr[16][9];
    __shared__ double sue[16][12];
    const int lane = threadIdx.x, slot = lane >> 2, I = lane & 3;
    const int64_t e = (int64_t)blockIdx.x * 16 + slot;
    if (e < a.n_elem && I == 0) {
        V3 X[2];
        int64_t cn[2];
        for (int i = 0; i < 2; i++) {
            cn[i] = a.conn[2 * e + i];
            const double *px = a.x + cn[i] / 2;
            X[i] = {px[0], px[1], px[2]};
        }
        V3 vxy = {a.evec[3 * e], a.evec[3 * e + 1], a.evec[3 * e + 2]};
        Geom g;
        beam_geometry(X, vxy, g);
        for (int i = 0; i < 9; i++) sr[slot][i] = g.r[i];
        for (int n = 0; n < 2; n++)
            for (int b = 0; b < 2; b++)
                for (int t = 0; t < 3; t++) {
                    double s = 0.;
                    for (int q = 0; q < 3; q++) s += g.r[3 * q + t] * a.u[cn[n] + 3 * b + q];
                    sue[slot][6 * n + 3 * b + t] = s;
                }
        const double *bp = a.props + (a.pid ? (int64_t)a.pid[e] : 0) * BEAM_STRIDE;
        if (ET == BEAMLR) beamlr_kc0_local(g.area, bp, tile[slot]);
        else beam_kc0_local(g.area, beam_section(bp, g.area), tile[slot]);
    }
    __syncthreads();
    if (e < a.n_elem && I < 2) {
        double f[6];
        for (int i = 0; i < 6; i++) {
            double s = 0.;
            for (int j = 0; j < 12; j++) s += tile[slot][(6 * I + i) * 12 + j] * sue[slot][j];
            f[i] = s;
        }
        double *o = fe + (e * 2 + I) * 6;
        for (int b = 0; b < 2; b++)
            for (int q = 0; q < 3; q++)
                o[3 * b + q] = sr[slot][3 * q] * f[3 * b] + sr[slot][3 * q + 1] * f[3 * b + 1] +
                               sr[slot][3 * q + 2] * f[3 * b + 2];
    }
}

// ------------------------------------------------------------------------
// Truss and Spring: 64 rods per block, one wave.  Each lane forms the few
// numbers that fix its rod's global blocks (the rotated 3x3 stiffness blocks,
// or the mass scalars and axis products) in LDS; the wave then streams the
// block's slots, consecutive lanes on consecutive slots.
//   KC0 (72 slots, truss.pyx:386-801, spring.pyx:295-707): node pair (I, J),
//     group g (translations, rotations), slot 36 I + 18 g + 6 d + 3 J + e,
//     value (I == J ? 1 : -1) B_g[d][e] with B_g = R diag(k_3g..k_3g+2) R^T
//     (truss: k = (E A / L, 0, 0, G J / L, 0, 0)).
//   Truss consistent mass (144 slots, row-major 12x12, truss.pyx:894-1616):
//     node block (I, J) = f_IJ [[rho I3, w ex^T], [ex w^T, rt ex ex^T]],
//     f = L / 3 on the diagonal blocks, L / 6 across, w = ry ez - rz ey.
//   Truss lumped mass (36 slots, truss.pyx:1617-1800): node-diagonal
//     L rho / 2 I3 and L rt / 2 ex ex^T.
// ------------------------------------------------------------------------
constexpr int ROD_VALS = 19;

template <int ET>
__device__ __forceinline__ void rod_load(const AsmArgs &a, int64_t e, int64_t *cn, Geom &g)
{
    V3 X[2];
    for (int i = 0; i < 2; i++) {
        cn[i] = a.conn[2 * e + i];
        const double *px = a.x + cn[i] / 2;
        X[i] = {px[0], px[1], px[2]};
    }
    rod_geometry<ET>(X, ET == SPRING ? a.evec + 6 * e : nullptr, g);
}

// local stiffness per DOF of a rod: k[0..5] along x, y, z, rx, ry, rz
template <int ET>
__device__ __forceinline__ void rod_k(const AsmArgs &a, int64_t e, double L, double *k)
{
    const int64_t p = a.pid ? (int64_t)a.pid[e] : 0;
    if (ET == TRUSS) {
        const double *bp = a.props + p * BEAM_STRIDE;
        k[0] = bp[0] * bp[1] / L;
        k[1] = k[2] = 0.;
        k[3] = bp[2] * bp[6] / L;
        k[4] = k[5] = 0.;
    } else {
        const double *kp = a.props + p * SPRING_STRIDE;
        for (int i = 0; i < 6; i++) k[i] = kp[i];
    }
}

template <int ET, int MAT>
__global__ void __launch_bounds__(64) rod_assemble(AsmArgs a)
{
    __shared__ double sv[64][ROD_VALS + 2];
    __shared__ int64_t sconn[64][2];
    __shared__ int64_t sk0[64];
    const int lane = threadIdx.x;
    const int64_t e0 = (int64_t)blockIdx.x * 64, e = e0 + lane;
    const bool lumped = (MAT == MASS) && a.mtype == 1;
    const int nslot = MAT == KC0 ? 72 : (lumped ? 36 : 144);
    if (e < a.n_elem) {
        int64_t cn[2];
        Geom g;
        rod_load<ET>(a, e, cn, g);
        sconn[lane][0] = cn[0];
        sconn[lane][1] = cn[1];
        sk0[lane] = a.init_k ? a.init_k[e] : (int64_t)nslot * e;
        double *o = sv[lane];
        if (MAT == KC0) {
            double k[6];
            rod_k<ET>(a, e, g.area, k);
            for (int grp = 0; grp < 2; grp++)
                for (int d = 0; d < 3; d++)
                    for (int q = 0; q < 3; q++) {
                        double s = 0.;
                        for (int t = 0; t < 3; t++)
                            s += k[3 * grp + t] * g.r[3 * d + t] * g.r[3 * q + t];
                        o[9 * grp + 3 * d + q] = s;
                    }
        } else {
            const double *bp = a.props + (a.pid ? (int64_t)a.pid[e] : 0) * BEAM_STRIDE;
            const double L = g.area, rho = bp[9], ry = bp[10], rz = bp[11];
            const double rt = bp[12] + bp[13];
            const double f = lumped ? L / 2 : L / 3;
            o[0] = f * rho;
            for (int d = 0; d < 3; d++) {
                o[1 + d] = g.r[3 * d];                                          // ex
                o[4 + d] = f * (ry * g.r[3 * d + 2] - rz * g.r[3 * d + 1]);     // f w
            }
            for (int d = 0; d < 3; d++)
                for (int q = 0; q < 3; q++) o[7 + 3 * d + q] = f * rt * g.r[3 * d] * g.r[3 * q];
        }
    }
    __syncthreads();
    const int nel = (int)((a.n_elem - e0) < 64 ? (a.n_elem - e0) : 64);
    const bool write_idx = !(a.flags & F_VALUES_ONLY);
    const bool overwrite = (a.flags & F_OVERWRITE);
    const int total = nel * nslot;
    for (int idx = lane; idx < total; idx += 64) {
        const int el = idx / nslot, k = idx - el * nslot;
        const double *o = sv[el];
        int rl, cl;
        double val;
        if (MAT == KC0) {
            const int I = k / 36, grp = (k / 18) & 1, d = (k % 18) / 6, J = (k % 6) / 3,
                      q = k % 3;
            rl = 6 * I + 3 * grp + d;
            cl = 6 * J + 3 * grp + q;
            val = (I == J ? 1. : -1.) * o[9 * grp + 3 * d + q];
        } else if (lumped) {
            const int n = k / 18, grp = (k % 18) / 9, t = k % 9, d = t / 3, q = t % 3;
            rl = 6 * n + 3 * grp + d;
            cl = 6 * n + 3 * grp + q;
            val = grp ? o[7 + 3 * d + q] : (d == q ? o[0] : 0.);
        } else {
            rl = k / 12;
            cl = k - 12 * rl;
            const int I = rl / 6, J = cl / 6, p = rl % 6, q = cl % 6;
            const double f = I == J ? 1. : 0.5;
            if (p < 3 && q < 3) val = p == q ? f * o[0] : 0.;
            else if (p < 3) val = f * o[4 + p] * o[1 + q - 3];
            else if (q < 3) val = f * o[1 + p - 3] * o[4 + q];
            else val = f * o[7 + 3 * (p - 3) + q - 3];
        }
        const int64_t k0 = sk0[el] + k;
        if (overwrite) a.v[k0] = val;
        else a.v[k0] += val;
        if (write_idx) {
            a.r[k0] = sconn[el][rl / 6] + rl % 6;
            a.c[k0] = sconn[el][cl / 6] + cl % 6;
        }
    }
}

// Rod internal forces in global axes (truss.pyx:150-202 ue, :339-385 finte,
// :802-837 rotation; spring.pyx:208-294, 708-739): local forces
// k_t (ue_0 - ue_1) at node 0 and their negative at node 1.
template <int ET>
__global__ void __launch_bounds__(64) rod_fe(AsmArgs a, double *__restrict__ fe)
{
    const int64_t e = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (e >= a.n_elem) return;
    int64_t cn[2];
    Geom g;
    rod_load<ET>(a, e, cn, g);
    double k[6];
    rod_k<ET>(a, e, g.area, k);
    double fl[6];
    for (int b = 0; b < 2; b++)
        for (int t = 0; t < 3; t++) {
            double u0 = 0., u1 = 0.;
            for (int q = 0; q < 3; q++) {
                u0 += g.r[3 * q + t] * a.u[cn[0] + 3 * b + q];
                u1 += g.r[3 * q + t] * a.u[cn[1] + 3 * b + q];
            }
            fl[3 * b + t] = k[3 * b + t] * (u0 - u1);
        }
    for (int b = 0; b < 2; b++)
        for (int q = 0; q < 3; q++) {
            const double s = g.r[3 * q] * fl[3 * b] + g.r[3 * q + 1] * fl[3 * b + 1] +
                             g.r[3 * q + 2] * fl[3 * b + 2];
            fe[e * 12 + 3 * b + q] = s;
            fe[e * 12 + 6 + 3 * b + q] = -s;
        }
}

// ------------------------------------------------------------------------
// Mixed element types in one launch (e.g. a stiffened panel: Quad4 skin,
// Tria3R fill, BeamC stringers).  Blocks [blk0[t], blk0[t+1]) run the block
// body of element type t = Quad4, Quad4R, Tria3R, BeamC on group t, in one
// LDS buffer sized for the largest body.  The shell bodies keep their own
// block shapes; beams take 8 per block so that their LDS tile does not
// lower the occupancy of the shell blocks.
// ------------------------------------------------------------------------
struct MixedArgs {
    AsmArgs g[4];
    int64_t blk0[5];
};

constexpr int MIXED_BPB = 8;

constexpr int cmax(int a, int b) { return a > b ? a : b; }

constexpr int MIXED_KC0_LDS =
    cmax(cmax(Kc0Lds<QUAD4>::BYTES, Kc0Lds<TRIA3R>::BYTES), BeamLds<MIXED_BPB>::BYTES);

template <int MAT>
constexpr int mixed_stream_lds()
{
    return cmax(cmax(StreamLds<QUAD4, MAT>::BYTES, StreamLds<TRIA3R, MAT>::BYTES),
                BeamLds<MIXED_BPB>::BYTES);
}

template <int FLAGS>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PF3D_KC0_WAVES)))
mixed_kc0(MixedArgs m)
{
    __shared__ __align__(16) char lds[MIXED_KC0_LDS];
    const int64_t b = blockIdx.x;
    if (b < m.blk0[1]) shell_kc0_body<QUAD4, FLAGS>(m.g[0], b, lds);
    else if (b < m.blk0[2]) shell_kc0_body<QUAD4R, FLAGS>(m.g[1], b - m.blk0[1], lds);
    else if (b < m.blk0[3]) shell_kc0_body<TRIA3R, FLAGS>(m.g[2], b - m.blk0[2], lds);
    else beam_body<BEAMC, KC0, MIXED_BPB>(m.g[3], b - m.blk0[3], lds);
}

// KG from displacements or mass (MAT = KG or MASS)
template <int MAT>
__global__ void __launch_bounds__(64) mixed_stream(MixedArgs m)
{
    __shared__ __align__(16) char lds[mixed_stream_lds<MAT>()];
    const int64_t b = blockIdx.x;
    if (b < m.blk0[1]) shell_stream_body<QUAD4, MAT>(m.g[0], b, lds);
    else if (b < m.blk0[2]) shell_stream_body<QUAD4R, MAT>(m.g[1], b - m.blk0[1], lds);
    else if (b < m.blk0[3]) shell_stream_body<TRIA3R, MAT>(m.g[2], b - m.blk0[2], lds);
    else beam_body<BEAMC, MAT, MIXED_BPB>(m.g[3], b - m.blk0[3], lds);
}

}  // namespace pf3d

// ------------------------------------------------------------------------
// launchers (called from the C-ABI layer)
// ------------------------------------------------------------------------
using namespace pf3d;

template <int ET, int MAT>
static hipError_t launch_shell(const AsmArgs &a, hipStream_t s)
{
    constexpr int NN = (ET == TRIA3R) ? 3 : 4;
    if constexpr (MAT == KC0) {
        constexpr int EPW = 64 / (NN * (NN + 1) / 2);
        const int64_t blocks = (a.n_elem + EPW - 1) / EPW;
        if (blocks == 0) return hipSuccess;
        const bool vec = !a.init_k && (reinterpret_cast<uintptr_t>(a.v) & 15) == 0 &&
                         ((a.flags & F_VALUES_ONLY) ||
                          ((reinterpret_cast<uintptr_t>(a.r) & 15) == 0 &&
                           (reinterpret_cast<uintptr_t>(a.c) & 15) == 0));
        const dim3 grid((unsigned)blocks), blk(64);
        switch ((a.flags & (F_VALUES_ONLY | F_OVERWRITE)) | (vec ? F_VEC : 0)) {
        case 0: hipLaunchKernelGGL((shell_kc0_sym<ET, 0>), grid, blk, 0, s, a); break;
        case 1: hipLaunchKernelGGL((shell_kc0_sym<ET, 1>), grid, blk, 0, s, a); break;
        case 2: hipLaunchKernelGGL((shell_kc0_sym<ET, 2>), grid, blk, 0, s, a); break;
        case 3: hipLaunchKernelGGL((shell_kc0_sym<ET, 3>), grid, blk, 0, s, a); break;
        case 4: hipLaunchKernelGGL((shell_kc0_sym<ET, 4>), grid, blk, 0, s, a); break;
        case 5: hipLaunchKernelGGL((shell_kc0_sym<ET, 5>), grid, blk, 0, s, a); break;
        case 6: hipLaunchKernelGGL((shell_kc0_sym<ET, 6>), grid, blk, 0, s, a); break;
        default: hipLaunchKernelGGL((shell_kc0_sym<ET, 7>), grid, blk, 0, s, a);
        }
        return hipGetLastError();
    } else {
        constexpr int EPB = (MAT == MASS) ? 32 : 64;
        const int64_t blocks = (a.n_elem + EPB - 1) / EPB;
        if (blocks == 0) return hipSuccess;
        // paired 16-byte stores: contiguous blocks (no init_k) of an even
        // slot count (every Quad4/Quad4R layout) at 16-byte aligned arrays
        AsmArgs b = a;
        const bool vec = NN == 4 && !a.init_k && (reinterpret_cast<uintptr_t>(a.v) & 15) == 0 &&
                         ((a.flags & F_VALUES_ONLY) ||
                          ((reinterpret_cast<uintptr_t>(a.r) & 15) == 0 &&
                           (reinterpret_cast<uintptr_t>(a.c) & 15) == 0));
        if (vec) b.flags |= F_VEC;
        hipLaunchKernelGGL((shell_stream<ET, MAT>), dim3((unsigned)blocks), dim3(64), 0, s, b);
        return hipGetLastError();
    }
}

template <int ET>
static hipError_t launch_shell_mat(int mat, const AsmArgs &a, hipStream_t s)
{
    if constexpr (ET != TRIA3R) {
        switch (mat) {
        case KAB: return launch_shell<ET, KAB>(a, s);
        case KAG: return launch_shell<ET, KAG>(a, s);
        case CAM: return launch_shell<ET, CAM>(a, s);
        default: break;
        }
    }
    switch (mat) {
    case KC0: return launch_shell<ET, KC0>(a, s);
    case KG: return launch_shell<ET, KG>(a, s);
    case KGS: return launch_shell<ET, KGS>(a, s);
    default: return launch_shell<ET, MASS>(a, s);
    }
}

hipError_t pf3d_launch_shell(int etype, int mat, const AsmArgs &a, hipStream_t s)
{
    switch (etype) {
    case QUAD4: return launch_shell_mat<QUAD4>(mat, a, s);
    case QUAD4R: return launch_shell_mat<QUAD4R>(mat, a, s);
    default: return launch_shell_mat<TRIA3R>(mat, a, s);
    }
}

template <int ET>
static void launch_beam_et(int mat, const AsmArgs &a, dim3 grid, hipStream_t s)
{
    if (mat == KC0)
        hipLaunchKernelGGL((beam_assemble<ET, KC0>), grid, dim3(64), 0, s, a);
    else if (mat == KG)
        hipLaunchKernelGGL((beam_assemble<ET, KG>), grid, dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL((beam_assemble<ET, MASS>), grid, dim3(64), 0, s, a);
}

hipError_t pf3d_launch_beam(int etype, int mat, const AsmArgs &a, hipStream_t s)
{
    const int64_t blocks = (a.n_elem + 15) / 16;
    if (blocks == 0) return hipSuccess;
    if (etype == BEAMLR) launch_beam_et<BEAMLR>(mat, a, dim3((unsigned)blocks), s);
    else launch_beam_et<BEAMC>(mat, a, dim3((unsigned)blocks), s);
    return hipGetLastError();
}

hipError_t pf3d_launch_geometry(int etype, int64_t n, const double *x, const int64_t *conn,
                                const double *evec, double *geom, double *xe, hipStream_t s)
{
    const unsigned blocks = (unsigned)((n + 127) / 128);
    if (blocks == 0) return hipSuccess;
    switch (etype) {
    case QUAD4:
    case QUAD4R:
        hipLaunchKernelGGL(element_geometry<QUAD4>, dim3(blocks), dim3(128), 0, s, n, x, conn,
                           evec, geom, xe);
        break;
    case TRIA3R:
        hipLaunchKernelGGL(element_geometry<TRIA3R>, dim3(blocks), dim3(128), 0, s, n, x, conn,
                           evec, geom, xe);
        break;
    case TRUSS:
        hipLaunchKernelGGL(element_geometry<TRUSS>, dim3(blocks), dim3(128), 0, s, n, x, conn,
                           evec, geom, xe);
        break;
    case SPRING:
        hipLaunchKernelGGL(element_geometry<SPRING>, dim3(blocks), dim3(128), 0, s, n, x, conn,
                           evec, geom, xe);
        break;
    default:
        hipLaunchKernelGGL(element_geometry<BEAMC>, dim3(blocks), dim3(128), 0, s, n, x, conn,
                           evec, geom, xe);
    }
    return hipGetLastError();
}

// groups: per element type Quad4, Quad4R, Tria3R, BeamC (n_elem 0 when absent)
hipError_t pf3d_launch_mixed(int mat, const AsmArgs *groups, hipStream_t s)
{
    MixedArgs m;
    int64_t per_kc0[4] = {Kc0Lds<QUAD4>::EPW, Kc0Lds<QUAD4R>::EPW, Kc0Lds<TRIA3R>::EPW,
                          MIXED_BPB};
    const int epb = mat == MASS ? 32 : 64;
    int64_t per_stream[4] = {epb, epb, epb, MIXED_BPB};
    m.blk0[0] = 0;
    bool vec = true;
    for (int t = 0; t < 4; t++) {
        m.g[t] = groups[t];
        const int64_t per = mat == KC0 ? per_kc0[t] : per_stream[t];
        m.blk0[t + 1] = m.blk0[t] + (groups[t].n_elem + per - 1) / per;
        if (t < 3 && groups[t].n_elem > 0) {
            const AsmArgs &a = groups[t];
            vec = vec && !a.init_k && (reinterpret_cast<uintptr_t>(a.v) & 15) == 0 &&
                  ((a.flags & F_VALUES_ONLY) ||
                   ((reinterpret_cast<uintptr_t>(a.r) & 15) == 0 &&
                    (reinterpret_cast<uintptr_t>(a.c) & 15) == 0));
        }
    }
    const int64_t blocks = m.blk0[4];
    if (blocks == 0) return hipSuccess;
    if (blocks > 0xffffffffLL) return hipErrorInvalidValue;
    const dim3 grid((unsigned)blocks), blk(64);
    if (mat == KC0) {
        switch ((groups[0].flags & (F_VALUES_ONLY | F_OVERWRITE)) | (vec ? F_VEC : 0)) {
        case 0: hipLaunchKernelGGL(mixed_kc0<0>, grid, blk, 0, s, m); break;
        case 1: hipLaunchKernelGGL(mixed_kc0<1>, grid, blk, 0, s, m); break;
        case 2: hipLaunchKernelGGL(mixed_kc0<2>, grid, blk, 0, s, m); break;
        case 3: hipLaunchKernelGGL(mixed_kc0<3>, grid, blk, 0, s, m); break;
        case 4: hipLaunchKernelGGL(mixed_kc0<4>, grid, blk, 0, s, m); break;
        case 5: hipLaunchKernelGGL(mixed_kc0<5>, grid, blk, 0, s, m); break;
        case 6: hipLaunchKernelGGL(mixed_kc0<6>, grid, blk, 0, s, m); break;
        default: hipLaunchKernelGGL(mixed_kc0<7>, grid, blk, 0, s, m);
        }
    } else if (mat == KG) {
        hipLaunchKernelGGL(mixed_stream<KG>, grid, blk, 0, s, m);
    } else {
        hipLaunchKernelGGL(mixed_stream<MASS>, grid, blk, 0, s, m);
    }
    return hipGetLastError();
}

hipError_t pf3d_launch_rod(int etype, int mat, const AsmArgs &a, hipStream_t s)
{
    const int64_t blocks = (a.n_elem + 63) / 64;
    if (blocks == 0) return hipSuccess;
    const dim3 grid((unsigned)blocks), blk(64);
    if (etype == SPRING)
        hipLaunchKernelGGL((rod_assemble<SPRING, KC0>), grid, blk, 0, s, a);
    else if (mat == KC0)
        hipLaunchKernelGGL((rod_assemble<TRUSS, KC0>), grid, blk, 0, s, a);
    else
        hipLaunchKernelGGL((rod_assemble<TRUSS, MASS>), grid, blk, 0, s, a);
    return hipGetLastError();
}

hipError_t pf3d_launch_rod_fe(int etype, const AsmArgs &a, double *fe, hipStream_t s)
{
    const int64_t blocks = (a.n_elem + 63) / 64;
    if (blocks == 0) return hipSuccess;
    if (etype == SPRING)
        hipLaunchKernelGGL(rod_fe<SPRING>, dim3((unsigned)blocks), dim3(64), 0, s, a, fe);
    else
        hipLaunchKernelGGL(rod_fe<TRUSS>, dim3((unsigned)blocks), dim3(64), 0, s, a, fe);
    return hipGetLastError();
}

hipError_t pf3d_launch_shell_fe(int etype, const AsmArgs &a, double *fe, hipStream_t s)
{
    const int nup = etype == TRIA3R ? 6 : 10;
    const int64_t blocks = (a.n_elem + 64 / nup - 1) / (64 / nup);
    if (blocks == 0) return hipSuccess;
    switch (etype) {
    case QUAD4:
        hipLaunchKernelGGL(shell_fe<QUAD4>, dim3((unsigned)blocks), dim3(64), 0, s, a, fe);
        break;
    case QUAD4R:
        hipLaunchKernelGGL(shell_fe<QUAD4R>, dim3((unsigned)blocks), dim3(64), 0, s, a, fe);
        break;
    default:
        hipLaunchKernelGGL(shell_fe<TRIA3R>, dim3((unsigned)blocks), dim3(64), 0, s, a, fe);
    }
    return hipGetLastError();
}

hipError_t pf3d_launch_beam_fe(int etype, const AsmArgs &a, double *fe, hipStream_t s)
{
    const int64_t blocks = (a.n_elem + 15) / 16;
    if (blocks == 0) return hipSuccess;
    if (etype == BEAMLR)
        hipLaunchKernelGGL(beam_fe<BEAMLR>, dim3((unsigned)blocks), dim3(64), 0, s, a, fe);
    else
        hipLaunchKernelGGL(beam_fe<BEAMC>, dim3((unsigned)blocks), dim3(64), 0, s, a, fe);
    return hipGetLastError();
}

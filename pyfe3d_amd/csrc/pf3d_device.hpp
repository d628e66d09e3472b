// Device-side helpers shared by the element kernels: element axes, local
// coordinates, constitutive tables and bilinear/linear shape functions.
//
// Everything here runs in IEEE FP64.  Each function cites the reference
// (pyfe3d v0.7.0) lines whose arithmetic it reproduces.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pf3d {

constexpr int SHELL_STRIDE = 32;  // doubles per shell property record
constexpr int BEAM_STRIDE = 16;   // doubles per beam property record (BeamC, Truss)
constexpr int SPRING_STRIDE = 6;  // kxe kye kze krxe krye krze per spring record
constexpr double GP2 = 0.5773502691896257645092;

enum ElemType { QUAD4 = 0, QUAD4R = 1, TRIA3R = 2, BEAMC = 3, TRUSS = 4, SPRING = 5, BEAMLR = 6 };
enum Matrix { KC0 = 0, KG = 1, KGS = 2, MASS = 3, KAB = 4, KAG = 5, CAM = 6 };

// matrices on the translations (I, a, J, b) with value P_IJ n_a n_b: KG,
// KG given stress and the piston-theory KA_beta, KA_gamma, CA
__host__ __device__ constexpr bool kg_like(int mat)
{
    return mat == KG || mat == KGS || mat == KAB || mat == KAG || mat == CAM;
}
__host__ __device__ constexpr bool piston(int mat) { return mat == KAB || mat == KAG || mat == CAM; }

// flags of the assembly entry points
constexpr int F_VALUES_ONLY = 1;  // do not write row/col indices
constexpr int F_OVERWRITE = 2;    // v = value instead of v += value

// arguments of the assembly kernels (device pointers)
struct AsmArgs {
    int64_t n_elem;
    const double *x;        // [3*n_nodes]
    const int64_t *conn;    // [n_elem*nn] DOF offsets c_I (6*node position)
    const int64_t *init_k;  // [n_elem] first slot, or null for size*e
    const double *evec;     // [n_elem*3] xmat (shells) or vxy (beams), [n_elem*6] x axis +
                            // vxy (springs); may be null where optional
    const int32_t *pid;     // [n_elem] property index, or null for 0
    const double *props;    // property records
    const double *k6rot;    // [n_elem] or null (100.)
    const double *alpha;    // [n_elem] or null (0.7)
    const double *hg;       // [n_elem*5] or null (1.)
    const double *u;        // [6*n_nodes] displacements (KG)
    const double *stress;   // [n_elem*3] or [3] (KG given stress)
    int stress_per_elem;
    int mtype;
    int flags;
    int64_t *r;
    int64_t *c;
    double *v;
};

struct V3 {
    double x, y, z;
};

// The geometry prologue (axes, local coordinates, area) is evaluated without
// FMA contraction, in the reference's operation order: the shoelace area and
// the local coordinates of small elements far from the origin are
// ill-conditioned, and keeping the reference's rounding there keeps the
// element matrices within 1e-12 of the reference's.
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 cross(V3 a, V3 b)
{
#pragma clang fp contract(off)
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double dot(V3 a, V3 b)
{
#pragma clang fp contract(off)
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ V3 unit(V3 a, double *n = nullptr)
{
#pragma clang fp contract(off)
    double s = sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    if (n) *n = s;
    return {a.x / s, a.y / s, a.z / s};
}

// Element geometry: columns of r are the local axes in global coordinates
// (r[i*3+j] = r_{i+1,j+1}); m is the material-direction rotation.
struct Geom {
    double r[9];
    double m[4];
    double xe[8];  // local x, y of up to 4 nodes (local z is not used)
    double area;   // area (shells) or length (beams)
};

// quad4.pyx:577-624 (also quad4r.pyx, tria3r.pyx): projected material axis.
__device__ __forceinline__ void material_direction(V3 ex, V3 ey, V3 ez, V3 xm, double tol,
                                                   double m[4])
{
#pragma clang fp contract(off)
    m[0] = 1.; m[1] = 0.; m[2] = 0.; m[3] = 1.;
    double nrm = sqrt(xm.x * xm.x + xm.y * xm.y + xm.z * xm.z);
    if (!(nrm > tol)) return;
    xm = {xm.x / nrm, xm.y / nrm, xm.z / nrm};
    double t;
    V3 ym = unit(cross(ez, xm), &t);
    if (!(t > tol)) return;
    V3 xp = unit(cross(ym, ez));
    double c = dot(xp, ex);
    double s = sqrt(1. - c * c);
    m[0] = c;
    m[3] = c;
    if (dot(xp, ey) > 0) { m[1] = -s; m[2] = s; }
    else { m[1] = s; m[2] = -s; }
}

__device__ __forceinline__ void set_axes(Geom &g, V3 ex, V3 ey, V3 ez)
{
    g.r[0] = ex.x; g.r[3] = ex.y; g.r[6] = ex.z;
    g.r[1] = ey.x; g.r[4] = ey.y; g.r[7] = ey.z;
    g.r[2] = ez.x; g.r[5] = ez.y; g.r[8] = ez.z;
}

__device__ __forceinline__ void local_xy(Geom &g, const V3 *X, int nn)
{
#pragma clang fp contract(off)
    for (int i = 0; i < nn; i++) {
        g.xe[2 * i] = g.r[0] * X[i].x + g.r[3] * X[i].y + g.r[6] * X[i].z;
        g.xe[2 * i + 1] = g.r[1] * X[i].x + g.r[4] * X[i].y + g.r[7] * X[i].z;
    }
}

// Quad4/Quad4R: quad4.pyx:491-575 (axes), :682-752 (xe, area)
__device__ __forceinline__ void quad_geometry(const V3 *X, V3 xmat, Geom &g)
{
#pragma clang fp contract(off)
    V3 v13 = sub(X[2], X[0]), v42 = sub(X[1], X[3]);
    double t;
    V3 ez = unit(cross(v42, v13), &t);
    double tol = t / 1e10;
    V3 ex = unit({(v13.x + v42.x) / 2., (v13.y + v42.y) / 2., (v13.z + v42.z) / 2.});
    V3 ey = unit(cross(ez, ex));
    set_axes(g, ex, ey, ez);
    material_direction(ex, ey, ez, xmat, tol, g.m);
    local_xy(g, X, 4);
    const double *p = g.xe;
    g.area = 0.5 * fabs((p[0] * p[3] + p[2] * p[5] + p[4] * p[7] + p[6] * p[1]) -
                        (p[2] * p[1] + p[4] * p[3] + p[6] * p[5] + p[0] * p[7]));
}

// Tria3R: tria3r.pyx:294-425 (x axis along edge 1-2), :480-547
__device__ __forceinline__ void tria_geometry(const V3 *X, V3 xmat, Geom &g)
{
#pragma clang fp contract(off)
    V3 v12 = sub(X[1], X[0]), v13 = sub(X[2], X[0]);
    double t;
    V3 ez = unit(cross(v12, v13), &t);
    double tol = t / 1e10;
    V3 ex = unit(v12);
    V3 ey = unit(cross(ez, ex));
    set_axes(g, ex, ey, ez);
    material_direction(ex, ey, ez, xmat, tol, g.m);
    local_xy(g, X, 3);
    const double *p = g.xe;
    g.area = fabs(((p[2] - p[0]) * (p[5] - p[1]) - (p[4] - p[0]) * (p[3] - p[1])) / 2.);
}

// BeamC: beamc.pyx:158-232 (axes from vxy), :287-350 (length)
__device__ __forceinline__ void beam_geometry(const V3 *X, V3 vxy, Geom &g)
{
#pragma clang fp contract(off)
    V3 d = sub(X[1], X[0]);
    V3 ex = unit(d);
    V3 ez = unit(cross(ex, vxy));
    V3 ey = unit(cross(ez, ex));
    set_axes(g, ex, ey, ez);
    g.m[0] = 1.; g.m[1] = 0.; g.m[2] = 0.; g.m[3] = 1.;
    g.area = sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
}

// Truss: truss.pyx:203-272 (x along the element, off-axis vector
// (xj, xk, xi) for the plane of y), :323-342 (length).
// Spring: spring.pyx:147-206 (x axis and XY-plane vector given per element).
template <int ET>
__device__ __forceinline__ void rod_geometry(const V3 *X, const double *ev, Geom &g)
{
#pragma clang fp contract(off)
    V3 ex, v;
    if (ET == TRUSS) {
        const V3 d = sub(X[1], X[0]);
        ex = unit(d);
        v = {ex.y, ex.z, ex.x};
        g.area = sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
    } else {
        ex = unit(V3{ev[0], ev[1], ev[2]});
        v = {ev[3], ev[4], ev[5]};
        g.area = 0.;
    }
    const V3 ez = unit(cross(ex, v));
    const V3 ey = unit(cross(ez, ex));
    set_axes(g, ex, ey, ez);
    g.m[0] = 1.; g.m[1] = 0.; g.m[2] = 0.; g.m[3] = 1.;
}

// Laminate tables in element axes: A' = T A T^T when m12 != 0 (quad4.pyx:845-899);
// the transverse-shear terms are not rotated and take the shear correction
// factors (quad4.pyx:903-905).
struct Shell {
    double C[6][6];  // [[A, B], [B, D]]
    double Es[2][2]; // over (gyz, gxz)
    double h, intrho, intrhoz, intrhoz2;
};

__device__ __forceinline__ void rot3(double M[3][3], const double m[4])
{
    const double T[3][3] = {{m[0] * m[0], m[1] * m[1], 2 * m[0] * m[1]},
                            {m[2] * m[2], m[3] * m[3], 2 * m[2] * m[3]},
                            {m[0] * m[2], m[1] * m[3], m[0] * m[3] + m[1] * m[2]}};
    double TM[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            TM[i][j] = T[i][0] * M[0][j] + T[i][1] * M[1][j] + T[i][2] * M[2][j];
    // upper triangle only, mirrored: the rotated table stays exactly
    // symmetric (the reference computes its six entries once)
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = i; j < 3; j++) {
            M[i][j] = TM[i][0] * T[j][0] + TM[i][1] * T[j][1] + TM[i][2] * T[j][2];
            M[j][i] = M[i][j];
        }
}

__device__ __forceinline__ void load_sym3(double M[3][3], const double *p)
{
    M[0][0] = p[0]; M[0][1] = M[1][0] = p[1]; M[0][2] = M[2][0] = p[2];
    M[1][1] = p[3]; M[1][2] = M[2][1] = p[4]; M[2][2] = p[5];
}

__device__ __forceinline__ void load_shell(const double *__restrict__ p, const double m[4], Shell &s)
{
    double A[3][3], B[3][3], D[3][3];
    load_sym3(A, p);
    load_sym3(B, p + 6);
    load_sym3(D, p + 12);
    if (m[1] != 0.) {
        rot3(A, m);
        rot3(B, m);
        rot3(D, m);
    }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            s.C[i][j] = A[i][j];
            s.C[i][j + 3] = B[i][j];
            s.C[i + 3][j] = B[i][j];
            s.C[i + 3][j + 3] = D[i][j];
        }
    const double k13 = p[21], k23 = p[22];
    s.Es[0][0] = p[18] * k23;
    s.Es[0][1] = s.Es[1][0] = p[19] * 0.5 * (k13 + k23);
    s.Es[1][1] = p[20] * k13;
    s.h = p[23];
    s.intrho = p[25];
    s.intrhoz = p[26];
    s.intrhoz2 = p[27];
}

// Bilinear shape functions, Jacobian convention of quad4.pyx:937-962.
struct QShape {
    double N[4], Nx[4], Ny[4], det;
};

__device__ __forceinline__ void quad_shape(const double *xe, double xi, double eta, QShape &q)
{
    const double sx[4] = {-1, 1, 1, -1}, sy[4] = {-1, -1, 1, 1};
    double dxi[4], deta[4];
    double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        q.N[i] = 0.25 * (1 + sx[i] * xi) * (1 + sy[i] * eta);
        dxi[i] = 0.25 * sx[i] * (1 + sy[i] * eta);
        deta[i] = 0.25 * sy[i] * (1 + sx[i] * xi);
        J11 += dxi[i] * xe[2 * i];
        J12 += dxi[i] * xe[2 * i + 1];
        J21 += deta[i] * xe[2 * i];
        J22 += deta[i] * xe[2 * i + 1];
    }
    q.det = J11 * J22 - J12 * J21;
    const double j11 = J22 / q.det, j12 = -J12 / q.det, j21 = -J21 / q.det, j22 = J11 / q.det;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        q.Nx[i] = j11 * dxi[i] + j12 * deta[i];
        q.Ny[i] = j21 * dxi[i] + j22 * deta[i];
    }
}

}  // namespace pf3d

/*
 * CPU oracle for the stiffness-assembly hot path (TEST INFRASTRUCTURE ONLY).
 *
 * This file is a plain-C restatement of pyfe3d v0.7.0's element arithmetic:
 * rotation matrices, local coordinates, the local element matrices of Quad4,
 * Quad4R, Tria3R and BeamC, their rotation to global axes and the sparse
 * (row, col, value) slot layouts.  It exists to check the HIP kernels:
 * only tests/, the smoke check and the CPU-baseline leg of the benchmark
 * may load it.  The product path never calls it.
 *
 * It is written as dense B^T.D.B quadrature over sparse B rows (not as the
 * reference's expanded symbolic terms), so agreement with the golden vectors
 * is a real check of both.
 *
 * Parity pins (see tests/test_oracle.py):
 *   - golden vectors written by tests/golden/make_golden.py from the reference
 *     (KC0/KG/KG-given-stress/M slots, geometry, all four element types);
 *   - invariants: symmetry, rigid-body null space of KC0, mass sums, closed
 *     forms (Euler-Bernoulli beam limit, plane-stress Q4 membrane block).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define DOF 6

enum { ET_QUAD4 = 0, ET_QUAD4R = 1, ET_TRIA3R = 2, ET_BEAMC = 3, ET_TRUSS = 4, ET_SPRING = 5,
       ET_BEAMLR = 6 };
#define IS_BEAM(t) ((t) == ET_BEAMC || (t) == ET_BEAMLR)
#define IS_ROD(t) ((t) == ET_TRUSS || (t) == ET_SPRING)
#define SPR_STRIDE 8 /* spring record: kxe kye kze krxe krye krze, 2 unused */
enum { MAT_KC0 = 0, MAT_KG = 1, MAT_KGS = 2, MAT_M = 3, MAT_FINT = 4, MAT_KAB = 5, MAT_KAG = 6,
       MAT_CA = 7 };
#define IS_PISTON(m) ((m) >= MAT_KAB && (m) <= MAT_CA)

/* shell property record (doubles): A11 A12 A16 A22 A26 A66, B.., D.., E44 E45
 * E55, scf_k13, scf_k23, h, offset, intrho, intrhoz, intrhoz2 */
#define SP_STRIDE 32
/* beam property record: A E G Iyy Izz Iyz J Ay Az intrho intrhoy intrhoz
 * intrhoy2 intrhoz2 intrhoyz */
#define BP_STRIDE 16

static const double GP2 = 0.5773502691896257645092;

/* ------------------------------------------------------------------ */
/* geometry                                                             */
/* ------------------------------------------------------------------ */

static void normalize3(double *v, double *norm)
{
    double n = pow(v[0] * v[0] + v[1] * v[1] + v[2] * v[2], 0.5);
    v[0] /= n; v[1] /= n; v[2] /= n;
    if (norm) *norm = n;
}

static void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* Material direction: projects xmat onto the element plane and returns
 * cos/sin of the angle to the element x axis.
 * quad4.pyx:577-624 (same logic in quad4r.pyx and tria3r.pyx). */
static void material_direction(const double *ex, const double *ey, const double *ez,
                               const double *xmat_in, double tol, double m[4])
{
    double xm[3] = {xmat_in[0], xmat_in[1], xmat_in[2]};
    double nrm = pow(xm[0] * xm[0] + xm[1] * xm[1] + xm[2] * xm[2], 0.5);
    m[0] = 1.; m[1] = 0.; m[2] = 0.; m[3] = 1.;
    xm[0] /= nrm; xm[1] /= nrm; xm[2] /= nrm;
    if (!(nrm > tol)) return;
    double ym[3], t;
    cross3(ez, xm, ym);
    normalize3(ym, &t);
    if (!(t > tol)) return;
    double xp[3];
    cross3(ym, ez, xp);
    normalize3(xp, NULL);
    double c = xp[0] * ex[0] + xp[1] * ex[1] + xp[2] * ex[2];
    double s = pow(1 - c * c, 0.5);
    m[0] = c;
    m[3] = c;
    if (xp[0] * ey[0] + xp[1] * ey[1] + xp[2] * ey[2] > 0) {
        m[1] = -s; m[2] = s;
    } else {
        m[1] = s; m[2] = -s;
    }
}

/* r is row-major r[i*3+j] = r_{i+1,j+1}; columns are the local axes. */
static void set_r(double *r, const double *ex, const double *ey, const double *ez)
{
    for (int i = 0; i < 3; i++) {
        r[i * 3 + 0] = ex[i];
        r[i * 3 + 1] = ey[i];
        r[i * 3 + 2] = ez[i];
    }
}

/* Quad4/Quad4R element axes: quad4.pyx:526-575 */
static void quad_axes(const double *X, const double *xmat, double *r, double *m)
{
    double v13[3], v42[3], ez[3], ex[3], ey[3], tmp;
    for (int i = 0; i < 3; i++) {
        v13[i] = X[6 + i] - X[0 + i];
        v42[i] = X[3 + i] - X[9 + i];
    }
    cross3(v42, v13, ez);
    normalize3(ez, &tmp);
    double tol = tmp / 1e10;
    for (int i = 0; i < 3; i++) ex[i] = (v13[i] + v42[i]) / 2.;
    normalize3(ex, NULL);
    cross3(ez, ex, ey);
    normalize3(ey, NULL);
    set_r(r, ex, ey, ez);
    material_direction(ex, ey, ez, xmat, tol, m);
}

/* Tria3R element axes: tria3r.pyx:322-370 (x axis along edge 1-2) */
static void tria_axes(const double *X, const double *xmat, double *r, double *m)
{
    double v12[3], v13[3], ez[3], ex[3], ey[3], tmp;
    for (int i = 0; i < 3; i++) {
        v12[i] = X[3 + i] - X[0 + i];
        v13[i] = X[6 + i] - X[0 + i];
    }
    cross3(v12, v13, ez);
    normalize3(ez, &tmp);
    double tol = tmp / 1e10;
    for (int i = 0; i < 3; i++) ex[i] = v12[i];
    normalize3(ex, NULL);
    cross3(ez, ex, ey);
    normalize3(ey, NULL);
    set_r(r, ex, ey, ez);
    material_direction(ex, ey, ez, xmat, tol, m);
}

/* BeamC axes: beamc.pyx:190-232 */
static void beam_axes(const double *X, const double *vxy, double *r)
{
    double ex[3], ez[3], ey[3];
    for (int i = 0; i < 3; i++) ex[i] = X[3 + i] - X[i];
    normalize3(ex, NULL);
    cross3(ex, vxy, ez);
    normalize3(ez, NULL);
    cross3(ez, ex, ey);
    normalize3(ey, NULL);
    set_r(r, ex, ey, ez);
}

/* local node coordinates xe = R^T x (quad4.pyx:682-731) */
static void local_coords(int nn, const double *X, const double *r, double *xe)
{
    for (int j = 0; j < nn; j++)
        for (int a = 0; a < 3; a++) {
            double s = 0;
            for (int i = 0; i < 3; i++) s += r[i * 3 + a] * X[j * 3 + i];
            xe[j * 3 + a] = s;
        }
}

/* quad4.pyx:733-752 ; tria3r.pyx:530-547 ; beamc.pyx:336-350 */
static double quad_area(const double *xe)
{
    double x1 = xe[0], y1 = xe[1], x2 = xe[3], y2 = xe[4];
    double x3 = xe[6], y3 = xe[7], x4 = xe[9], y4 = xe[10];
    return 0.5 * fabs((x1 * y2 + x2 * y3 + x3 * y4 + x4 * y1) -
                      (x2 * y1 + x3 * y2 + x4 * y3 + x1 * y4));
}

static double tria_area(const double *xe)
{
    double x1 = xe[0], y1 = xe[1], x2 = xe[3], y2 = xe[4], x3 = xe[6], y3 = xe[7];
    return fabs(((x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1)) / 2.);
}

/* ------------------------------------------------------------------ */
/* constitutive                                                         */
/* ------------------------------------------------------------------ */

typedef struct {
    double A[3][3], B[3][3], D[3][3], Es[2][2];
    double h, intrho, intrhoz, intrhoz2;
} shell_t;

static void sym3(double M[3][3], const double *p)
{
    M[0][0] = p[0]; M[0][1] = M[1][0] = p[1]; M[0][2] = M[2][0] = p[2];
    M[1][1] = p[3]; M[1][2] = M[2][1] = p[4]; M[2][2] = p[5];
}

/* ABD rotated from the material direction to element axes, A' = T A T^T
 * (quad4.pyx:845-899; E44/E45/E55 are not rotated). */
static void rotate3(double M[3][3], const double m[4])
{
    double T[3][3] = {
        {m[0] * m[0], m[1] * m[1], 2 * m[0] * m[1]},
        {m[2] * m[2], m[3] * m[3], 2 * m[2] * m[3]},
        {m[0] * m[2], m[1] * m[3], m[0] * m[3] + m[1] * m[2]}};
    double TM[3][3], R[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += T[i][k] * M[k][j];
            TM[i][j] = s;
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += TM[i][k] * T[j][k];
            R[i][j] = s;
        }
    memcpy(M, R, sizeof(R));
}

static void load_shell(const double *p, const double m[4], shell_t *s)
{
    sym3(s->A, p + 0);
    sym3(s->B, p + 6);
    sym3(s->D, p + 12);
    if (m[1] != 0) { /* quad4.pyx:847 exact compare */
        rotate3(s->A, m);
        rotate3(s->B, m);
        rotate3(s->D, m);
    }
    double k13 = p[21], k23 = p[22];
    s->Es[0][0] = p[18] * k23;              /* gyz-gyz: E44 k23 */
    s->Es[0][1] = s->Es[1][0] = p[19] * 0.5 * (k13 + k23);
    s->Es[1][1] = p[20] * k13;              /* gxz-gxz: E55 k13 */
    s->h = p[23];
    s->intrho = p[25];
    s->intrhoz = p[26];
    s->intrhoz2 = p[27];
}

/* ------------------------------------------------------------------ */
/* shape functions                                                      */
/* ------------------------------------------------------------------ */

/* bilinear quad at (xi, eta): N, dN/dx, dN/dy, detJ; quad4.pyx:937-962 */
static double quad_shape(const double *xe, double xi, double eta,
                         double N[4], double Nx[4], double Ny[4])
{
    static const double sx[4] = {-1, 1, 1, -1}, sy[4] = {-1, -1, 1, 1};
    double dxi[4], deta[4];
    for (int i = 0; i < 4; i++) {
        N[i] = 0.25 * (1 + sx[i] * xi) * (1 + sy[i] * eta);
        dxi[i] = 0.25 * sx[i] * (1 + sy[i] * eta);
        deta[i] = 0.25 * sy[i] * (1 + sx[i] * xi);
    }
    double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
    for (int i = 0; i < 4; i++) {
        J11 += dxi[i] * xe[3 * i];
        J12 += dxi[i] * xe[3 * i + 1];
        J21 += deta[i] * xe[3 * i];
        J22 += deta[i] * xe[3 * i + 1];
    }
    double det = J11 * J22 - J12 * J21;
    double j11 = J22 / det, j12 = -J12 / det, j21 = -J21 / det, j22 = J11 / det;
    for (int i = 0; i < 4; i++) {
        Nx[i] = j11 * dxi[i] + j12 * deta[i];
        Ny[i] = j21 * dxi[i] + j22 * deta[i];
    }
    return det;
}

/* ------------------------------------------------------------------ */
/* B rows (local DOF order per node: u v w rx ry rz)                    */
/* ------------------------------------------------------------------ */

/* generalised membrane/bending strains: exx eyy gxy kxx kyy kxy (quad4.pyx:964-1000) */
static void b_abd(int nn, const double *Nx, const double *Ny, double B[6][24])
{
    memset(B, 0, sizeof(double) * 6 * 24);
    for (int I = 0; I < nn; I++) {
        int o = 6 * I;
        B[0][o + 0] = Nx[I];
        B[1][o + 1] = Ny[I];
        B[2][o + 0] = Ny[I];
        B[2][o + 1] = Nx[I];
        B[3][o + 4] = Nx[I];
        B[4][o + 3] = -Ny[I];
        B[5][o + 3] = -Nx[I];
        B[5][o + 4] = Ny[I];
    }
}

/* transverse shear, split in gradient and rotation parts:
 * gyz = w,y - rx ; gxz = w,x + ry (quad4.pyx:1107-1125) */
static void b_shear(int nn, const double *N, const double *Nx, const double *Ny,
                    double Bg[2][24], double Br[2][24])
{
    memset(Bg, 0, sizeof(double) * 2 * 24);
    memset(Br, 0, sizeof(double) * 2 * 24);
    for (int I = 0; I < nn; I++) {
        int o = 6 * I;
        Bg[0][o + 2] = Ny[I];
        Bg[1][o + 2] = Nx[I];
        Br[0][o + 3] = -N[I];
        Br[1][o + 4] = N[I];
    }
}

/* drilling: rz - (v,x - u,y)/2 (quad4.pyx:1053-1066) */
static void b_drill(int nn, const double *N, const double *Nx, const double *Ny, double *Bd)
{
    memset(Bd, 0, sizeof(double) * 24);
    for (int I = 0; I < nn; I++) {
        Bd[6 * I + 0] = Ny[I] / 2.;
        Bd[6 * I + 1] = -Nx[I] / 2.;
        Bd[6 * I + 5] = N[I];
    }
}

/* K += w * Ba^T C Bb, with Ba (na x n), C (na x nb), Bb (nb x n) */
static void add_btcb(int n, double *K, double w, int na, const double *Ba,
                     int nb, const double *C, const double *Bb)
{
    double CB[6][24];
    for (int a = 0; a < na; a++)
        for (int j = 0; j < n; j++) {
            double s = 0;
            for (int b = 0; b < nb; b++) s += C[a * nb + b] * Bb[b * 24 + j];
            CB[a][j] = s;
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double s = 0;
            for (int a = 0; a < na; a++) s += Ba[a * 24 + i] * CB[a][j];
            K[i * n + j] += w * s;
        }
}

static void abd6(const shell_t *s, double C[36])
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            C[i * 6 + j] = s->A[i][j];
            C[i * 6 + j + 3] = s->B[i][j];
            C[(i + 3) * 6 + j] = s->B[i][j];
            C[(i + 3) * 6 + j + 3] = s->D[i][j];
        }
}

/* ------------------------------------------------------------------ */
/* local KC0                                                             */
/* ------------------------------------------------------------------ */

/* Quad4: quad4.pyx:755-1172 */
static void quad4_kc0_local(const double *xe, double area, const shell_t *s, double *K)
{
    const int n = 24;
    double C[36], N[4], Nx[4], Ny[4], B[6][24], Bg[2][24], Br[2][24], Bd[24];
    abd6(s, C);
    memset(K, 0, sizeof(double) * n * n);
    int thick = (s->h / pow(area, 0.5)) >= 1.;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            double xi = a ? GP2 : -GP2, eta = b ? GP2 : -GP2;
            double det = quad_shape(xe, xi, eta, N, Nx, Ny);
            b_abd(4, Nx, Ny, B);
            add_btcb(n, K, det, 6, &B[0][0], 6, C, &B[0][0]);
            if (thick) {
                b_shear(4, N, Nx, Ny, Bg, Br);
                add_btcb(n, K, det, 2, &Bg[0][0], 2, &s->Es[0][0], &Bg[0][0]);
            }
            b_drill(4, N, Nx, Ny, Bd);
            double one = 1.0; /* unit drilling weight (quad4.pyx:1068-1073) */
            add_btcb(n, K, det, 1, Bd, 1, &one, Bd);
        }
    double det = quad_shape(xe, 0., 0., N, Nx, Ny);
    b_shear(4, N, Nx, Ny, Bg, Br);
    double w = 4. * det;
    if (!thick) add_btcb(n, K, w, 2, &Bg[0][0], 2, &s->Es[0][0], &Bg[0][0]);
    add_btcb(n, K, w, 2, &Bg[0][0], 2, &s->Es[0][0], &Br[0][0]);
    add_btcb(n, K, w, 2, &Br[0][0], 2, &s->Es[0][0], &Bg[0][0]);
    add_btcb(n, K, w, 2, &Br[0][0], 2, &s->Es[0][0], &Br[0][0]);
}

/* Quad4R: reduced 1-point + hourglass + 2x2 drilling, quad4r.pyx:1300-1312, 3081-3462 */
static void quad4r_kc0_local(const double *xe, double area, const shell_t *s, double k6rot,
                             const double *hg, double *K)
{
    const int n = 24;
    double C[36], N[4], Nx[4], Ny[4], B[6][24], Bg[2][24], Br[2][24], Bd[24];
    abd6(s, C);
    memset(K, 0, sizeof(double) * n * n);
    const double (*A)[3] = s->A;
    double det3 = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[1][2])
                - A[0][1] * (A[0][1] * A[2][2] - A[1][2] * A[0][2])
                + A[0][2] * (A[0][1] * A[1][2] - A[1][1] * A[0][2]);
    double a11 = (A[1][1] * A[2][2] - A[1][2] * A[1][2]) / det3;
    double a22 = (A[0][0] * A[2][2] - A[0][2] * A[0][2]) / det3;
    double E1eq = 1. / (s->h * a11), E2eq = 1. / (s->h * a22);
    double damp = 1.0 + 1.0 / area;
    double Eu = hg[0] * 0.1 * E1eq * s->h / damp;
    double Ev = hg[1] * 0.1 * E2eq * s->h / damp;
    double Erx = hg[3] * 0.1 * E2eq * s->h * s->h * s->h / damp;
    double Ery = hg[4] * 0.1 * E1eq * s->h * s->h * s->h / damp;
    double Ew = hg[2] * 0.5 * (Erx + Ery);

    double det = quad_shape(xe, 0., 0., N, Nx, Ny);
    double w = 4. * det;
    b_abd(4, Nx, Ny, B);
    add_btcb(n, K, w, 6, &B[0][0], 6, C, &B[0][0]);
    b_shear(4, N, Nx, Ny, Bg, Br);
    double Bs[2][24];
    for (int a = 0; a < 2; a++)
        for (int j = 0; j < 24; j++) Bs[a][j] = Bg[a][j] + Br[a][j];
    add_btcb(n, K, w, 2, &Bs[0][0], 2, &s->Es[0][0], &Bs[0][0]);
    /* hourglass: gamma_I = d2 N_I / dx dy at the centroid */
    static const double sxy[4] = {1, -1, 1, -1};
    double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
    static const double sx[4] = {-1, 1, 1, -1}, sy[4] = {-1, -1, 1, 1};
    for (int i = 0; i < 4; i++) {
        J11 += 0.25 * sx[i] * xe[3 * i];
        J12 += 0.25 * sx[i] * xe[3 * i + 1];
        J21 += 0.25 * sy[i] * xe[3 * i];
        J22 += 0.25 * sy[i] * xe[3 * i + 1];
    }
    double dj = J11 * J22 - J12 * J21;
    double j11 = J22 / dj, j12 = -J12 / dj, j21 = -J21 / dj, j22 = J11 / dj;
    double gam[4];
    for (int i = 0; i < 4; i++) gam[i] = 0.25 * sxy[i] * (j11 * j22 + j12 * j21);
    const double Ehg[5] = {Eu, Ev, Ew, Erx, Ery};
    for (int I = 0; I < 4; I++)
        for (int J = 0; J < 4; J++)
            for (int d = 0; d < 5; d++)
                K[(6 * I + d) * n + 6 * J + d] += w * Ehg[d] * gam[I] * gam[J];
    /* drilling, 2x2 Gauss, 1e-6 K6ROT A66 */
    double kd = 1e-6 * k6rot * A[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            double xi = a ? GP2 : -GP2, eta = b ? GP2 : -GP2;
            double d2 = quad_shape(xe, xi, eta, N, Nx, Ny);
            b_drill(4, N, Nx, Ny, Bd);
            add_btcb(n, K, d2, 1, Bd, 1, &kd, Bd);
        }
}

/* Tria3R: tria3r.pyx:1105-1122, 2116-2324 */
static void tria3r_kc0_local(const double *xe, double area, const shell_t *s0, double k6rot,
                             double alpha, int scatter_mask, double *K)
{
    const int n = 18;
    shell_t s = *s0;
    double x1 = xe[0], y1 = xe[1], x2 = xe[3], y2 = xe[4], x3 = xe[6], y3 = xe[7];
    double l12 = pow((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2), 0.5);
    double l23 = pow((x2 - x3) * (x2 - x3) + (y2 - y3) * (y2 - y3), 0.5);
    double l31 = pow((x3 - x1) * (x3 - x1) + (y3 - y1) * (y3 - y1), 0.5);
    double maxl = l12;
    if (l23 > maxl) maxl = l23;
    if (l31 > maxl) maxl = l31;
    double f = 1 / (1 + alpha * maxl * maxl / (s.h * s.h));
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) s.Es[i][j] *= f;
    double C[36], Nx[3], Ny[3], N[3], B[6][24], Bg[2][24], Br[2][24], Bd[24];
    abd6(&s, C);
    memset(K, 0, sizeof(double) * n * n);
    Nx[0] = (y2 - y3) / (2 * area); Nx[1] = (y3 - y1) / (2 * area); Nx[2] = (y1 - y2) / (2 * area);
    Ny[0] = (x3 - x2) / (2 * area); Ny[1] = (x1 - x3) / (2 * area); Ny[2] = (x2 - x1) / (2 * area);
    double detJ = 2 * area;
    double w = 0.5 * detJ;
    N[0] = N[1] = N[2] = 1. / 3.;
    b_abd(3, Nx, Ny, B);
    add_btcb(n, K, w, 6, &B[0][0], 6, C, &B[0][0]);
    b_shear(3, N, Nx, Ny, Bg, Br);
    double Bs[2][24];
    for (int a = 0; a < 2; a++)
        for (int j = 0; j < 24; j++) Bs[a][j] = Bg[a][j] + Br[a][j];
    add_btcb(n, K, w, 2, &Bs[0][0], 2, &s.Es[0][0], &Bs[0][0]);
    double kd = 1e-6 * k6rot * s.A[2][2];
    static const double P[3][3] = {{2. / 3, 1. / 6, 1. / 6}, {1. / 6, 1. / 6, 2. / 3},
                                   {1. / 6, 2. / 3, 1. / 6}};
    for (int q = 0; q < 3; q++) {
        b_drill(3, P[q], Nx, Ny, Bd);
        add_btcb(n, K, detJ * 0.5 / 3., 1, Bd, 1, &kd, Bd);
    }
    /* Only the rz_I-rz_I diagonal of the drilling rz couplings reaches the
     * global matrix: the u/v-rz and rz_I-rz_J (I != J) terms computed at
     * tria3r.pyx:2230-2324 are never used by the scatter at :2325-2974.
     * The internal forces (tria3r.pyx:549-948) do use them. */
    if (!scatter_mask) return;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            if ((i % 6 == 5 || j % 6 == 5) && i != j) K[i * n + j] = 0.;
}

/* BeamC: closed-form Timoshenko stiffness, beamc.pyx:543-624.
 * Local DOF order u v w rx ry rz per node; stored as upper triangle. */
static void beamc_kc0_local(double L, const double *bp, double *K)
{
    double A = bp[0], E = bp[1], G = bp[2], Iyy = bp[3], Izz = bp[4], Iyz = bp[5];
    double J = bp[6], Ay = bp[7], Az = bp[8];
    double ay = 12 * E * Izz / (G * A * L * L), az = 12 * E * Iyy / (G * A * L * L);
    double by = 1 / (1. - ay), bz = 1 / (1. - az);
    double L2 = L * L, L3 = L2 * L;
    /* bending-shear blocks */
    double kyy = by * by * (A * G * L2 * ay * ay + 12 * E * Izz);   /* v-v family */
    double kzz = bz * bz * (A * G * L2 * az * az + 12 * E * Iyy);   /* w-w family */
    double cyz = E * Iyz * by * bz;
    double U[12][12];
    memset(U, 0, sizeof(U));
    /* axial / axial-bending coupling through Ay, Az */
    U[0][0] = A * E / L;  U[0][6] = -A * E / L;  U[6][6] = A * E / L;
    U[0][4] = Az * E * bz * (1 - az) / L;   U[0][10] = Az * E * bz * (az - 1) / L;
    U[0][5] = Ay * E * by * (ay - 1) / L;   U[0][11] = Ay * E * by * (1 - ay) / L;
    U[4][6] = Az * E * bz * (az - 1) / L;   U[6][10] = Az * E * bz * (1 - az) / L;
    U[5][6] = Ay * E * by * (1 - ay) / L;   U[6][11] = Ay * E * by * (ay - 1) / L;
    /* v (1,7) with rz (5,11) */
    U[1][1] = kyy / L3;  U[1][7] = -kyy / L3;  U[7][7] = kyy / L3;
    U[1][5] = kyy / (2 * L2);  U[1][11] = kyy / (2 * L2);
    U[5][7] = -kyy / (2 * L2); U[7][11] = -kyy / (2 * L2);
    /* w (2,8) with ry (4,10) */
    U[2][2] = kzz / L3;  U[2][8] = -kzz / L3;  U[8][8] = kzz / L3;
    U[2][4] = -kzz / (2 * L2); U[2][10] = -kzz / (2 * L2);
    U[4][8] = kzz / (2 * L2);  U[8][10] = kzz / (2 * L2);
    /* v-w coupling through Iyz */
    U[1][2] = 12 * cyz / L3;  U[1][8] = -12 * cyz / L3;
    U[2][7] = -12 * cyz / L3; U[7][8] = 12 * cyz / L3;
    U[1][4] = -6 * cyz / L2;  U[1][10] = -6 * cyz / L2;
    U[2][5] = 6 * cyz / L2;   U[2][11] = 6 * cyz / L2;
    U[4][7] = 6 * cyz / L2;   U[5][8] = -6 * cyz / L2;
    U[7][10] = 6 * cyz / L2;  U[8][11] = -6 * cyz / L2;
    /* torsion and torsion-shear coupling through Ay, Az */
    U[3][3] = G * J / L;  U[3][9] = -G * J / L;  U[9][9] = G * J / L;
    U[1][3] = Az * G * ay * by / L;  U[1][9] = -Az * G * ay * by / L;
    U[3][7] = -Az * G * ay * by / L; U[7][9] = Az * G * ay * by / L;
    U[2][3] = -Ay * G * az * bz / L; U[2][9] = Ay * G * az * bz / L;
    U[3][8] = Ay * G * az * bz / L;  U[8][9] = -Ay * G * az * bz / L;
    U[3][4] = Ay * G * az * bz / 2;  U[3][10] = Ay * G * az * bz / 2;
    U[4][9] = -Ay * G * az * bz / 2; U[9][10] = -Ay * G * az * bz / 2;
    U[3][5] = Az * G * ay * by / 2;  U[3][11] = Az * G * ay * by / 2;
    U[5][9] = -Az * G * ay * by / 2; U[9][11] = -Az * G * ay * by / 2;
    /* rotations */
    double ryy = bz * bz * (A * G * L2 * az * az / 4 + E * Iyy * az * az - 2 * E * Iyy * az + 4 * E * Iyy) / L;
    double ryy2 = bz * bz * (A * G * L2 * az * az / 4 - E * Iyy * az * az + 2 * E * Iyy * az + 2 * E * Iyy) / L;
    double rzz = by * by * (A * G * L2 * ay * ay / 4 + E * Izz * ay * ay - 2 * E * Izz * ay + 4 * E * Izz) / L;
    double rzz2 = by * by * (A * G * L2 * ay * ay / 4 - E * Izz * ay * ay + 2 * E * Izz * ay + 2 * E * Izz) / L;
    U[4][4] = ryy;  U[10][10] = ryy;  U[4][10] = ryy2;
    U[5][5] = rzz;  U[11][11] = rzz;  U[5][11] = rzz2;
    double cyz1 = cyz * (-ay * az + ay + az - 4) / L;
    double cyz2 = cyz * (ay * az - ay - az - 2) / L;
    U[4][5] = cyz1;  U[10][11] = cyz1;
    U[4][11] = cyz2; U[5][10] = cyz2;
    for (int i = 0; i < 12; i++)
        for (int j = 0; j < 12; j++) K[i * 12 + j] = (j >= i) ? U[i][j] : U[j][i];
}

/* BeamC geometric stiffness for axial force N (beamc.pyx:1889-2180,
 * N = A E (u2 - u1)/L at :1889).  Only v, w, ry, rz take part; each DOF
 * belongs to a bending plane (v, rz: y; w, ry: z) with its own alpha/beta:
 *   disp-disp  +-(5 a_i a_j - 5 a_i - 5 a_j + 6)/(5L)   (+ same node)
 *   disp-rot   s_rot * (+1 node 1, -1 node 2 of the disp DOF) / 10
 *   rot-rot    s_rot_i s_rot_j L (5 a_i a_j - 5 a_i - 5 a_j + 8)/60 (same node)
 *              -s_rot_i s_rot_j L (5 a_i a_j - 5 a_i - 5 a_j + 2)/60 (other node)
 * all times N beta_i beta_j, with s_rot(ry) = -1, s_rot(rz) = +1. */
static void beamc_kg_local(double L, const double *bp, double N, double *K)
{
    double A = bp[0], E = bp[1], G = bp[2], Iyy = bp[3], Izz = bp[4];
    double ay = 12 * E * Izz / (G * A * L * L), az = 12 * E * Iyy / (G * A * L * L);
    double by = 1 / (1. - ay), bz = 1 / (1. - az);
    memset(K, 0, sizeof(double) * 144);
    for (int i = 0; i < 12; i++)
        for (int j = 0; j < 12; j++) {
            int di = i % 6, dj = j % 6, ni = i / 6, nj = j / 6;
            if (!(di == 1 || di == 2 || di == 4 || di == 5)) continue;
            if (!(dj == 1 || dj == 2 || dj == 4 || dj == 5)) continue;
            int yi = (di == 1 || di == 5), yj = (dj == 1 || dj == 5);
            double a_i = yi ? ay : az, a_j = yj ? ay : az;
            double b = N * (yi ? by : bz) * (yj ? by : bz);
            int ri = di >= 4, rj = dj >= 4;
            double si = di == 4 ? -1. : 1., sj = dj == 4 ? -1. : 1.;
            double q = 5 * a_i * a_j - 5 * a_i - 5 * a_j;
            double v;
            if (!ri && !rj) v = (ni == nj ? 1. : -1.) * (q + 6) / (5 * L);
            else if (!ri && rj) v = sj * (ni == 0 ? 1. : -1.) / 10;
            else if (ri && !rj) v = si * (nj == 0 ? 1. : -1.) / 10;
            else v = si * sj * L * (ni == nj ? (q + 8) / 60 : -(q + 2) / 60);
            K[i * 12 + j] = b * v;
        }
}

/* BeamC mass, consistent (mtype 0, beamc.pyx:2246-2968) and lumped
 * (mtype 1, beamc.pyx:2969-3152).  Local 12x12, symmetric; upper triangle
 * listed per row. */
static void beamc_m_local(double L, const double *bp, int mtype, double *K)
{
    double A = bp[0], E = bp[1], G = bp[2], Iyy = bp[3], Izz = bp[4];
    double r = bp[9], ry = bp[10], rz = bp[11], ry2 = bp[12], rz2 = bp[13], ryz = bp[14];
    double ay = 12 * E * Izz / (G * A * L * L), az = 12 * E * Iyy / (G * A * L * L);
    double by = 1 / (1. - ay), bz = 1 / (1. - az);
    double L2 = L * L;
    double U[12][12];
    memset(U, 0, sizeof(U));
    if (mtype == 1) {
        for (int n = 0; n < 2; n++) {
            int o = 6 * n;
            U[o][o] = L * r / 2;
            U[o + 1][o + 1] = L * by * by * r * (ay * ay - 2 * ay + 1) / 2;
            U[o + 2][o + 2] = L * bz * bz * r * (az * az - 2 * az + 1) / 2;
            U[o + 3][o + 3] = L * (ry2 + rz2) / 2;
            U[o + 4][o + 4] = L * bz * bz * rz2 * (az * az - 2 * az + 1) / 2;
            U[o + 5][o + 5] = L * by * by * ry2 * (ay * ay - 2 * ay + 1) / 2;
        }
    } else {
        /* translation-translation bending terms */
        double tyy = by * by * (70 * L2 * ay * ay * r - 147 * L2 * ay * r + 78 * L2 * r + 252 * ry2) / (210 * L);
        double tzz = bz * bz * (70 * L2 * az * az * r - 147 * L2 * az * r + 78 * L2 * r + 252 * rz2) / (210 * L);
        double tyy2 = by * by * (35 * L2 * ay * ay * r - 63 * L2 * ay * r + 27 * L2 * r - 252 * ry2) / (210 * L);
        double tzz2 = bz * bz * (35 * L2 * az * az * r - 63 * L2 * az * r + 27 * L2 * r - 252 * rz2) / (210 * L);
        /* translation-rotation bending terms */
        double cy1 = by * by * (35 * L2 * ay * ay * r - 77 * L2 * ay * r + 44 * L2 * r + 420 * ay * ry2 + 84 * ry2) / 840;
        double cy2 = by * by * (-35 * L2 * ay * ay * r + 63 * L2 * ay * r - 26 * L2 * r + 420 * ay * ry2 + 84 * ry2) / 840;
        double cz1 = bz * bz * (-35 * L2 * az * az * r + 77 * L2 * az * r - 44 * L2 * r - 420 * az * rz2 - 84 * rz2) / 840;
        double cz2 = bz * bz * (35 * L2 * az * az * r - 63 * L2 * az * r + 26 * L2 * r - 420 * az * rz2 - 84 * rz2) / 840;
        /* rotation-rotation bending terms */
        double qzz = L * bz * bz * (7 * L2 * az * az * r - 14 * L2 * az * r + 8 * L2 * r + 280 * az * az * rz2 - 140 * az * rz2 + 112 * rz2) / 840;
        double qyy = L * by * by * (7 * L2 * ay * ay * r - 14 * L2 * ay * r + 8 * L2 * r + 280 * ay * ay * ry2 - 140 * ay * ry2 + 112 * ry2) / 840;
        double qzz2 = L * bz * bz * (-7 * L2 * az * az * r + 14 * L2 * az * r - 6 * L2 * r + 140 * az * az * rz2 + 140 * az * rz2 - 28 * rz2) / 840;
        double qyy2 = L * by * by * (-7 * L2 * ay * ay * r + 14 * L2 * ay * r - 6 * L2 * r + 140 * ay * ay * ry2 + 140 * ay * ry2 - 28 * ry2) / 840;
        double pyz = 6 * by * bz * ryz / (5 * L);
        double sz1 = by * bz * ryz * (5 * az + 1) / 10, sy1 = by * bz * ryz * (5 * ay + 1) / 10;
        double qyz = L * by * bz * ryz * (-20 * ay * az + 5 * ay + 5 * az - 8) / 60;
        double qyz2 = L * by * bz * ryz * (-10 * ay * az - 5 * ay - 5 * az + 2) / 60;
        U[0][0] = L * r / 3;  U[0][6] = L * r / 6;  U[6][6] = L * r / 3;
        U[0][1] = by * ry / 2;  U[0][7] = -by * ry / 2;  U[1][6] = by * ry / 2;  U[6][7] = -by * ry / 2;
        U[0][2] = bz * rz / 2;  U[0][8] = -bz * rz / 2;  U[2][6] = bz * rz / 2;  U[6][8] = -bz * rz / 2;
        U[0][4] = L * bz * rz * (1 - 4 * az) / 12;   U[6][10] = U[0][4];
        U[0][5] = L * by * ry * (4 * ay - 1) / 12;   U[6][11] = U[0][5];
        U[0][10] = -L * bz * rz * (2 * az + 1) / 12; U[4][6] = U[0][10];
        U[0][11] = L * by * ry * (2 * ay + 1) / 12;  U[5][6] = U[0][11];
        U[1][1] = tyy;  U[7][7] = tyy;  U[1][7] = tyy2;
        U[2][2] = tzz;  U[8][8] = tzz;  U[2][8] = tzz2;
        U[1][2] = pyz;  U[7][8] = pyz;  U[1][8] = -pyz;  U[2][7] = -pyz;
        U[1][5] = cy1;  U[7][11] = -cy1;  U[1][11] = cy2;  U[5][7] = -cy2;
        U[2][4] = cz1;  U[8][10] = -cz1;  U[2][10] = cz2;  U[4][8] = -cz2;
        U[1][4] = -sz1; U[1][10] = -sz1; U[4][7] = sz1;  U[7][10] = sz1;
        U[2][5] = sy1;  U[2][11] = sy1;  U[5][8] = -sy1; U[8][11] = -sy1;
        U[1][3] = L * by * rz * (20 * ay - 21) / 60;  U[7][9] = U[1][3];
        U[1][9] = L * by * rz * (10 * ay - 9) / 60;   U[3][7] = U[1][9];
        U[2][3] = L * bz * ry * (21 - 20 * az) / 60;  U[8][9] = U[2][3];
        U[2][9] = L * bz * ry * (9 - 10 * az) / 60;   U[3][8] = U[2][9];
        U[3][3] = L * (ry2 + rz2) / 3;  U[9][9] = U[3][3];  U[3][9] = L * (ry2 + rz2) / 6;
        U[3][4] = L2 * bz * ry * (5 * az - 6) / 120;  U[9][10] = -U[3][4];
        U[3][5] = L2 * by * rz * (5 * ay - 6) / 120;  U[9][11] = -U[3][5];
        U[3][10] = L2 * bz * ry * (4 - 5 * az) / 120; U[4][9] = -U[3][10];
        U[3][11] = L2 * by * rz * (4 - 5 * ay) / 120; U[5][9] = -U[3][11];
        U[4][4] = qzz;  U[10][10] = qzz;  U[4][10] = qzz2;
        U[5][5] = qyy;  U[11][11] = qyy;  U[5][11] = qyy2;
        U[4][5] = qyz;  U[10][11] = qyz;  U[4][11] = qyz2;  U[5][10] = qyz2;
    }
    for (int i = 0; i < 12; i++)
        for (int j = 0; j < 12; j++) K[i * 12 + j] = (j >= i) ? U[i][j] : U[j][i];
}

/* ------------------------------------------------------------------ */
/* local KG (w-w block) and M                                            */
/* ------------------------------------------------------------------ */

/* membrane forces from local displacements: N = A eps + B kappa
 * (quad4.pyx:1965-1967) */
/* Truss (truss.pyx:386-801, 838-1800): axial E A / L and torsion G J / L;
 * consistent mass (two-point Gauss) and lumped mass (Gauss-Lobatto) with
 * the translations, the torsional inertia and the translation-torsion
 * coupling through the section's first moments. */
static void truss_kc0_local(double L, const double *bp, double *K)
{
    memset(K, 0, sizeof(double) * 144);
    const double ka = bp[0] * bp[1] / L, kt = bp[2] * bp[6] / L;
    for (int I = 0; I < 2; I++)
        for (int J = 0; J < 2; J++) {
            const double s = I == J ? 1. : -1.;
            K[(6 * I + 0) * 12 + 6 * J + 0] = s * ka;
            K[(6 * I + 3) * 12 + 6 * J + 3] = s * kt;
        }
}

static void truss_m_local(double L, const double *bp, int mtype, double *K)
{
    memset(K, 0, sizeof(double) * 144);
    const double rho = bp[9], ry = bp[10], rz = bp[11], rt = bp[12] + bp[13];
    if (mtype == 1) {
        for (int I = 0; I < 2; I++) {
            const int o = 6 * I;
            K[(o + 0) * 12 + o + 0] = K[(o + 1) * 12 + o + 1] = K[(o + 2) * 12 + o + 2] =
                L * rho / 2;
            K[(o + 3) * 12 + o + 3] = L * rt / 2;
        }
        return;
    }
    for (int I = 0; I < 2; I++)
        for (int J = 0; J < 2; J++) {
            const double f = L * (I == J ? 1. / 3. : 1. / 6.);
            const int a = 6 * I, b = 6 * J;
            K[(a + 0) * 12 + b + 0] = K[(a + 1) * 12 + b + 1] = K[(a + 2) * 12 + b + 2] = f * rho;
            K[(a + 1) * 12 + b + 3] = K[(a + 3) * 12 + b + 1] = -f * rz;
            K[(a + 2) * 12 + b + 3] = K[(a + 3) * 12 + b + 2] = f * ry;
            K[(a + 3) * 12 + b + 3] = f * rt;
        }
}

/* BeamLR (beamlr.pyx): linear Timoshenko beam, every field interpolated
 * linearly, one-point (reduced) integration.  At the midpoint
 *   eps = u', ky = ry', kz = rz', gxy = v' - rz_m, gxz = w' + ry_m, phi = rx'
 * with section stiffness D = E [[A, Az, -Ay], [Az, Iyy, -Iyz], [-Ay, -Iyz, Izz]]
 * on (eps, ky, kz) and S = G [[A, 0, -Az], [0, A, Ay], [-Az, Ay, J]] on
 * (gxy, gxz, phi); K = L (Bm^T D Bm + Bs^T S Bs) (beamlr.pyx:354-406, the
 * internal forces K ue; 407-1187 the scattered K). */
static void beamlr_kc0_local(double L, const double *bp, double *K)
{
    const double A = bp[0], E = bp[1], G = bp[2], Iyy = bp[3], Izz = bp[4], Iyz = bp[5];
    const double J = bp[6], Ay = bp[7], Az = bp[8];
    const double D[3][3] = {{E * A, E * Az, -E * Ay}, {E * Az, E * Iyy, -E * Iyz},
                            {-E * Ay, -E * Iyz, E * Izz}};
    const double S[3][3] = {{G * A, 0., -G * Az}, {0., G * A, G * Ay}, {-G * Az, G * Ay, G * J}};
    double Bm[3][12], Bs[3][12];
    memset(Bm, 0, sizeof(Bm));
    memset(Bs, 0, sizeof(Bs));
    Bm[0][0] = -1 / L; Bm[0][6] = 1 / L;
    Bm[1][4] = -1 / L; Bm[1][10] = 1 / L;
    Bm[2][5] = -1 / L; Bm[2][11] = 1 / L;
    Bs[0][1] = -1 / L; Bs[0][7] = 1 / L; Bs[0][5] = -0.5; Bs[0][11] = -0.5;
    Bs[1][2] = -1 / L; Bs[1][8] = 1 / L; Bs[1][4] = 0.5; Bs[1][10] = 0.5;
    Bs[2][3] = -1 / L; Bs[2][9] = 1 / L;
    for (int i = 0; i < 12; i++)
        for (int j = 0; j < 12; j++) {
            double s = 0.;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++)
                    s += Bm[a][i] * D[a][b] * Bm[b][j] + Bs[a][i] * S[a][b] * Bs[b][j];
            K[i * 12 + j] = L * s;
        }
}

/* BeamLR KG (beamlr.pyx:1226-1463): N = A E (u2 - u1) / L on the bending
 * rotations, N int(N_I N_J) [[1, -1], [-1, 1]] over (ry, rz) of nodes I, J. */
static void beamlr_kg_local(double L, double N, double *K)
{
    memset(K, 0, sizeof(double) * 144);
    for (int I = 0; I < 2; I++)
        for (int J = 0; J < 2; J++) {
            const double f = N * L * (I == J ? 1. / 3. : 1. / 6.);
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++)
                    K[(6 * I + 4 + a) * 12 + 6 * J + 4 + b] = (a == b ? f : -f);
        }
}

/* BeamLR mass (beamlr.pyx:1464-2424): section mass matrix m of
 * u = u0 + z ry - y rz, v = v0 - z rx, w = w0 + y rx; consistent: the
 * two-point Gauss integral of N_I N_J (L/3, L/6) times m; lumped (Gauss-
 * Lobatto at the nodes): L/2 times the diagonal of m. */
static void beamlr_m_local(double L, const double *bp, int mtype, double *K)
{
    const double r = bp[9], ry = bp[10], rz = bp[11], ry2 = bp[12], rz2 = bp[13], ryz = bp[14];
    double m[6][6];
    memset(m, 0, sizeof(m));
    m[0][0] = m[1][1] = m[2][2] = r;
    m[0][4] = m[4][0] = rz;
    m[0][5] = m[5][0] = -ry;
    m[1][3] = m[3][1] = -rz;
    m[2][3] = m[3][2] = ry;
    m[3][3] = ry2 + rz2;
    m[4][4] = rz2;
    m[5][5] = ry2;
    m[4][5] = m[5][4] = -ryz;
    memset(K, 0, sizeof(double) * 144);
    for (int I = 0; I < 2; I++)
        for (int J = 0; J < 2; J++) {
            if (mtype == 1 && I != J) continue;
            const double f = mtype == 1 ? L / 2 : L * (I == J ? 1. / 3. : 1. / 6.);
            for (int d = 0; d < 6; d++)
                for (int e = 0; e < 6; e++) {
                    if (mtype == 1 && d != e) continue;
                    K[(6 * I + d) * 12 + 6 * J + e] = f * m[d][e];
                }
        }
}

/* Spring (spring.pyx:295-707): k on each local DOF between the nodes */
static void spring_kc0_local(const double *kp, double *K)
{
    memset(K, 0, sizeof(double) * 144);
    for (int I = 0; I < 2; I++)
        for (int J = 0; J < 2; J++)
            for (int d = 0; d < 6; d++) K[(6 * I + d) * 12 + 6 * J + d] = (I == J ? 1. : -1.) * kp[d];
}

static void membrane_forces(int nn, const shell_t *s, const double *Nx, const double *Ny,
                            const double *ue, double Nf[3])
{
    double B[6][24], eps[6];
    b_abd(nn, Nx, Ny, B);
    for (int a = 0; a < 6; a++) {
        double t = 0;
        for (int j = 0; j < 6 * nn; j++) t += B[a][j] * ue[j];
        eps[a] = t;
    }
    for (int i = 0; i < 3; i++) {
        Nf[i] = 0;
        for (int j = 0; j < 3; j++) Nf[i] += s->A[i][j] * eps[j] + s->B[i][j] * eps[3 + j];
    }
}

/* G[I][J] = sum_gp grad N_I^T [[Nxx,Nxy],[Nxy,Nyy]] grad N_J detJ w
 * quad4.pyx:1933-2257 (from ue), quad4.pyx:2259-3081 (given stress),
 * tria3r.pyx:3020-3575 (1 point) */
static void shell_kg_ww(int etype, const double *xe, double area, const shell_t *s,
                        const double *ue, const double *stress, double G[4][4])
{
    memset(G, 0, sizeof(double) * 16);
    double N[4], Nx[4], Ny[4], Nf[3];
    if (etype == ET_TRIA3R) {
        double x1 = xe[0], y1 = xe[1], x2 = xe[3], y2 = xe[4], x3 = xe[6], y3 = xe[7];
        Nx[0] = (y2 - y3) / (2 * area); Nx[1] = (y3 - y1) / (2 * area); Nx[2] = (y1 - y2) / (2 * area);
        Ny[0] = (x3 - x2) / (2 * area); Ny[1] = (x1 - x3) / (2 * area); Ny[2] = (x2 - x1) / (2 * area);
        if (ue) membrane_forces(3, s, Nx, Ny, ue, Nf);
        else { Nf[0] = stress[0]; Nf[1] = stress[1]; Nf[2] = stress[2]; }
        double w = 0.5 * 2 * area;
        for (int I = 0; I < 3; I++)
            for (int J = 0; J < 3; J++)
                G[I][J] = w * (Nx[J] * (Nx[I] * Nf[0] + Ny[I] * Nf[2]) +
                               Ny[J] * (Nx[I] * Nf[2] + Ny[I] * Nf[1]));
        return;
    }
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            double xi = a ? GP2 : -GP2, eta = b ? GP2 : -GP2;
            double det = quad_shape(xe, xi, eta, N, Nx, Ny);
            if (ue) membrane_forces(4, s, Nx, Ny, ue, Nf);
            else { Nf[0] = stress[0]; Nf[1] = stress[1]; Nf[2] = stress[2]; }
            for (int I = 0; I < 4; I++)
                for (int J = 0; J < 4; J++)
                    G[I][J] += det * (Nx[J] * (Nx[I] * Nf[0] + Ny[I] * Nf[2]) +
                                      Ny[J] * (Nx[I] * Nf[2] + Ny[I] * Nf[1]));
        }
}

/* Piston-theory aerodynamic matrices of Quad4 / Quad4R on the normal
 * displacement, 2x2 Gauss (quad4.pyx:9491-11917, quad4r.pyx:12789-15200):
 *   KA_beta:  -int N_I dN_J/dX dA   (X = global x, the flow direction:
 *              dN/dX = N,x r11 + N,y r21 in element axes)
 *   KA_gamma:  int N_I N_J dA       CA: -int N_I N_J dA
 * G[I][J] multiplies n_a n_b of the element normal in the scatter. */
static void shell_piston_ww(int mat, const double *xe, const double *r, double G[4][4])
{
    memset(G, 0, sizeof(double) * 16);
    double N[4], Nx[4], Ny[4];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            double xi = a ? GP2 : -GP2, eta = b ? GP2 : -GP2;
            double det = quad_shape(xe, xi, eta, N, Nx, Ny);
            for (int I = 0; I < 4; I++)
                for (int J = 0; J < 4; J++) {
                    if (mat == MAT_KAB) G[I][J] -= N[I] * det * (Nx[J] * r[0] + Ny[J] * r[3]);
                    else if (mat == MAT_KAG) G[I][J] += N[I] * N[J] * det;
                    else G[I][J] -= N[I] * N[J] * det;
                }
        }
}

/* H[I][J] mass weights: quad4.pyx:3124, 4598-4640, 8873-8902;
 * tria3r.pyx:4107-4108 (valH1 = detJ/18), 3-point rule, vertex rule */
static void shell_mass_h(int etype, const double *xe, double area, int mtype, double H[4][4])
{
    memset(H, 0, sizeof(double) * 16);
    if (etype == ET_TRIA3R) {
        double detJ = 2 * area;
        if (mtype == 1) {
            for (int I = 0; I < 3; I++)
                for (int J = 0; J < 3; J++) H[I][J] = detJ / 18.;
            return;
        }
        double P[3][3];
        if (mtype == 0) {
            double a = 2. / 3, b = 1. / 6;
            double Q[3][3] = {{a, b, b}, {b, b, a}, {b, a, b}};
            memcpy(P, Q, sizeof(Q));
        } else {
            double Q[3][3] = {{1, 0, 0}, {0, 0, 1}, {0, 1, 0}};
            memcpy(P, Q, sizeof(Q));
        }
        double w = 0.5 / 3.;
        for (int q = 0; q < 3; q++)
            for (int I = 0; I < 3; I++)
                for (int J = 0; J < 3; J++) H[I][J] += P[q][I] * P[q][J] * detJ * w;
        return;
    }
    if (mtype == 1) {
        for (int I = 0; I < 4; I++)
            for (int J = 0; J < 4; J++) H[I][J] = 0.0625 * area;
        return;
    }
    double p = (mtype == 0) ? GP2 : 1.0;
    double N[4], Nx[4], Ny[4];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            double det = quad_shape(xe, a ? p : -p, b ? p : -p, N, Nx, Ny);
            for (int I = 0; I < 4; I++)
                for (int J = 0; J < 4; J++) H[I][J] += N[I] * N[J] * det;
        }
}

/* local 6x6 mass block pattern per unit weight (quad4.pyx:4642+):
 * intrho on u,v,w; +intrhoz on u-ry, -intrhoz on v-rx; intrhoz2 on rx,ry.
 * The lumped variant (mtype 2) drops the translation-rotation coupling. */
static void shell_mass_block(const shell_t *s, int lumped, double mb[6][6])
{
    memset(mb, 0, sizeof(double) * 36);
    mb[0][0] = mb[1][1] = mb[2][2] = s->intrho;
    mb[3][3] = mb[4][4] = s->intrhoz2;
    if (!lumped) {
        mb[0][4] = mb[4][0] = s->intrhoz;
        mb[1][3] = mb[3][1] = -s->intrhoz;
    }
}

/* ------------------------------------------------------------------ */
/* slot layouts (Appendix of the sparse patterns)                        */
/* ------------------------------------------------------------------ */

/* fills (row_local, col_local) for each slot; returns count */
static int layout(int etype, int mat, int mtype, int nn, int *rl, int *cl)
{
    int k = 0;
    if (IS_ROD(etype) && mat == MAT_KC0) {
        /* truss.pyx:432-800, spring.pyx:341-560: (I, group, d) rows of (J, e) */
        for (int I = 0; I < 2; I++)
            for (int g = 0; g < 2; g++)
                for (int d = 0; d < 3; d++)
                    for (int J = 0; J < 2; J++)
                        for (int e = 0; e < 3; e++) {
                            rl[k] = 6 * I + 3 * g + d; cl[k] = 6 * J + 3 * g + e; k++;
                        }
        return k;
    }
    if (etype == ET_TRUSS && mat == MAT_M && mtype == 0) {
        for (int i = 0; i < 12; i++)
            for (int j = 0; j < 12; j++) { rl[k] = i; cl[k] = j; k++; }
        return k;
    }
    if (etype == ET_BEAMLR && mat == MAT_KG) {
        /* beamlr.pyx:1276-1386: bending rotations (I, 3..5, J, 3..5) */
        for (int I = 0; I < 2; I++)
            for (int d = 3; d < 6; d++)
                for (int J = 0; J < 2; J++)
                    for (int e = 3; e < 6; e++) { rl[k] = 6 * I + d; cl[k] = 6 * J + e; k++; }
        return k;
    }
    if (mat == MAT_KC0 || (IS_BEAM(etype) && (mat == MAT_KG || (mat == MAT_M && mtype == 0)))) {
        for (int I = 0; I < nn; I++)
            for (int d = 0; d < 6; d++)
                for (int J = 0; J < nn; J++)
                    for (int e = 0; e < 6; e++) { rl[k] = 6 * I + d; cl[k] = 6 * J + e; k++; }
        return k;
    }
    if (mat == MAT_KG || mat == MAT_KGS || IS_PISTON(mat)) {
        for (int I = 0; I < nn; I++)
            for (int d = 0; d < 3; d++)
                for (int J = 0; J < nn; J++)
                    for (int e = 0; e < 3; e++) { rl[k] = 6 * I + d; cl[k] = 6 * J + e; k++; }
        return k;
    }
    /* mass */
    if (IS_BEAM(etype) || etype == ET_TRUSS) { /* lumped: node-diagonal translations + rotations */
        for (int I = 0; I < nn; I++)
            for (int g = 0; g < 2; g++)
                for (int d = 0; d < 3; d++)
                    for (int e = 0; e < 3; e++) {
                        rl[k] = 6 * I + 3 * g + d; cl[k] = 6 * I + 3 * g + e; k++;
                    }
        return k;
    }
    if (mtype == 2) {
        for (int I = 0; I < nn; I++)
            for (int g = 0; g < 2; g++)
                for (int d = 0; d < 3; d++)
                    for (int J = 0; J < nn; J++)
                        for (int e = 0; e < 3; e++) {
                            rl[k] = 6 * I + 3 * g + d; cl[k] = 6 * J + 3 * g + e; k++;
                        }
        return k;
    }
    for (int I = 0; I < nn; I++)
        for (int d = 0; d < 6; d++)
            for (int J = 0; J < nn; J++)
                for (int e = 0; e < 6; e++) {
                    if (e == (d + 3) % 6) continue;
                    rl[k] = 6 * I + d; cl[k] = 6 * J + e; k++;
                }
    return k;
}

int orc_slot_count(int etype, int mat, int mtype)
{
    int rl[576], cl[576];
    int nn = etype == ET_TRIA3R ? 3 : ((IS_BEAM(etype) || IS_ROD(etype)) ? 2 : 4);
    return layout(etype, mat, mtype, nn, rl, cl);
}

/* ------------------------------------------------------------------ */
/* per-element driver                                                    */
/* ------------------------------------------------------------------ */

static int nnodes(int etype)
{
    return etype == ET_TRIA3R ? 3 : ((IS_BEAM(etype) || IS_ROD(etype)) ? 2 : 4);
}

/* per-element orientation vectors: xmat / vxy (3), spring axis + vxy (6), truss none */
static int evec_stride(int etype) { return etype == ET_SPRING ? 6 : 3; }

/* geometry record: r11..r33 (9), m11 m12 m21 m22 (4), area or length (1) */
static void element_geometry(int etype, const double *x, const int64_t *cn,
                             const double *evec, double *g, double *xe)
{
    int nn = nnodes(etype);
    double X[12];
    for (int i = 0; i < nn; i++)
        for (int a = 0; a < 3; a++) X[3 * i + a] = x[3 * cn[i] + a];
    double m[4] = {1, 0, 0, 1};
    if (etype == ET_QUAD4 || etype == ET_QUAD4R) quad_axes(X, evec, g, m);
    else if (etype == ET_TRIA3R) tria_axes(X, evec, g, m);
    else if (etype == ET_TRUSS) {
        /* truss.pyx:203-272: x along the element, off-axis vector (xj, xk, xi) */
        double ex[3] = {X[3] - X[0], X[4] - X[1], X[5] - X[2]}, t;
        normalize3(ex, &t);
        const double v[3] = {ex[1], ex[2], ex[0]};
        double ez[3], ey[3];
        cross3(ex, v, ez);
        normalize3(ez, &t);
        cross3(ez, ex, ey);
        normalize3(ey, &t);
        set_r(g, ex, ey, ez);
    } else if (etype == ET_SPRING) {
        /* spring.pyx:147-206: x axis and XY-plane vector given */
        double ex[3] = {evec[0], evec[1], evec[2]}, t;
        normalize3(ex, &t);
        double ez[3], ey[3];
        cross3(ex, evec + 3, ez);
        normalize3(ez, &t);
        cross3(ez, ex, ey);
        normalize3(ey, &t);
        set_r(g, ex, ey, ez);
    } else beam_axes(X, evec, g);
    memcpy(g + 9, m, sizeof(m));
    local_coords(nn, X, g, xe);
    if (IS_BEAM(etype) || etype == ET_TRUSS) {
        double d[3] = {X[3] - X[0], X[4] - X[1], X[5] - X[2]};
        g[13] = pow(d[0] * d[0] + d[1] * d[1] + d[2] * d[2], 0.5);
    } else {
        g[13] = etype == ET_TRIA3R ? tria_area(xe) : quad_area(xe);
    }
}

int orc_geometry(int etype, int64_t E, const double *x, const int64_t *conn,
                 const double *evec, double *geom, double *xe_out)
{
    int nn = nnodes(etype);
    for (int64_t e = 0; e < E; e++) {
        double xe[12];
        element_geometry(etype, x, conn + e * nn, evec ? evec + evec_stride(etype) * e : NULL,
                         geom + 14 * e, xe);
        if (xe_out) memcpy(xe_out + 3 * nn * e, xe, sizeof(double) * 3 * nn);
    }
    return 0;
}

/* local displacement vector ue = blockdiag(R^T) u (quad4.pyx:627-680) */
static void local_displ(int nn, const double *u, const int64_t *cn, const double *r, double *ue)
{
    for (int j = 0; j < nn; j++)
        for (int blk = 0; blk < 2; blk++)
            for (int a = 0; a < 3; a++) {
                double s = 0;
                for (int i = 0; i < 3; i++) s += r[i * 3 + a] * u[6 * cn[j] + 3 * blk + i];
                ue[6 * j + 3 * blk + a] = s;
            }
}

/* dense local matrix of one element (6nn x 6nn) */
static int element_local(int etype, int mat, int mtype, const double *g, const double *xe,
                         const double *prop, double k6rot, double alpha, const double *hg,
                         const double *ue, const double *stress, double *K)
{
    int nn = nnodes(etype), n = 6 * nn;
    double area = g[13];
    if (etype == ET_TRUSS) {
        if (mat == MAT_KC0 || mat == MAT_FINT) truss_kc0_local(area, prop, K);
        else if (mat == MAT_M) truss_m_local(area, prop, mtype, K);
        else return -1;
        return 0;
    }
    if (etype == ET_SPRING) {
        if (mat == MAT_KC0 || mat == MAT_FINT) spring_kc0_local(prop, K);
        else return -1;
        return 0;
    }
    if (etype == ET_BEAMLR) {
        if (mat == MAT_KC0 || mat == MAT_FINT) beamlr_kc0_local(area, prop, K);
        else if (mat == MAT_KG) /* N = A E (u2 - u1) / L, beamlr.pyx:1388 */
            beamlr_kg_local(area, prop[0] * prop[1] * (-ue[0] + ue[6]) / area, K);
        else if (mat == MAT_M) beamlr_m_local(area, prop, mtype, K);
        else return -1;
        return 0;
    }
    if (etype == ET_BEAMC) {
        if (mat == MAT_KC0 || mat == MAT_FINT) beamc_kc0_local(area, prop, K);
        else if (mat == MAT_KG) /* N = A E (u2 - u1) / L, beamc.pyx:1889 */
            beamc_kg_local(area, prop, prop[0] * prop[1] * (-ue[0] + ue[6]) / area, K);
        else if (mat == MAT_M) beamc_m_local(area, prop, mtype, K);
        else return -1;
        return 0;
    }
    shell_t s;
    load_shell(prop, g + 9, &s);
    if (mat == MAT_KC0 || mat == MAT_FINT) {
        if (etype == ET_QUAD4) quad4_kc0_local(xe, area, &s, K);
        else if (etype == ET_QUAD4R) quad4r_kc0_local(xe, area, &s, k6rot, hg, K);
        else tria3r_kc0_local(xe, area, &s, k6rot, alpha, mat == MAT_KC0, K);
        return 0;
    }
    memset(K, 0, sizeof(double) * n * n);
    if (IS_PISTON(mat)) {
        if (etype == ET_TRIA3R) return -1;
        double G[4][4];
        shell_piston_ww(mat, xe, g, G);
        for (int I = 0; I < nn; I++)
            for (int J = 0; J < nn; J++) K[(6 * I + 2) * n + 6 * J + 2] = G[I][J];
        return 0;
    }
    if (mat == MAT_KG || mat == MAT_KGS) {
        double G[4][4];
        shell_kg_ww(etype, xe, area, &s, mat == MAT_KG ? ue : NULL, stress, G);
        for (int I = 0; I < nn; I++)
            for (int J = 0; J < nn; J++) K[(6 * I + 2) * n + 6 * J + 2] = G[I][J];
        return 0;
    }
    double H[4][4], mb[6][6];
    shell_mass_h(etype, xe, area, mtype, H);
    shell_mass_block(&s, mtype == 2, mb);
    for (int I = 0; I < nn; I++)
        for (int J = 0; J < nn; J++)
            for (int d = 0; d < 6; d++)
                for (int e = 0; e < 6; e++) K[(6 * I + d) * n + 6 * J + e] = H[I][J] * mb[d][e];
    return 0;
}

/*
 * Assemble one matrix for a batch of elements of one type.
 *   conn[E*nn]  node positions (DOF offset = 6*pos, coordinates at 3*pos)
 *   init_k[E]   first slot of each element
 *   evec[E*3]   xmat (shells) or vxy (beams)
 *   pid[E]      property index; props[npid*stride]
 * Values accumulate (+=); indices are written unless values_only.
 */
int orc_assemble(int etype, int mat, int mtype, int64_t E, const double *x,
                 const int64_t *conn, const int64_t *init_k, const double *evec,
                 const int64_t *pid, const double *props, const double *k6rot,
                 const double *alpha, const double *hg, const double *u,
                 const double *stress, int64_t *r, int64_t *c, double *v, int values_only)
{
    int nn = nnodes(etype), n = 6 * nn;
    int rl[576], cl[576];
    int ns = layout(etype, mat, mtype, nn, rl, cl);
    int stride = (IS_BEAM(etype) || etype == ET_TRUSS) ? BP_STRIDE
                 : (etype == ET_SPRING ? SPR_STRIDE : SP_STRIDE);
    static const double hg_default[5] = {1., 1., 1., 1., 1.};
    int err = 0;
    /* elements are independent and write disjoint slots: the loop may run
     * on several host threads (orc_set_threads) without changing a bit */
#pragma omp parallel for schedule(static) reduction(| : err)
    for (int64_t e = 0; e < E; e++) {
        const int64_t *cn = conn + e * nn;
        double g[14], xe[12], ue[24], K[576], Kg[576];
        element_geometry(etype, x, cn, evec ? evec + evec_stride(etype) * e : NULL, g, xe);
        if (mat == MAT_KG && u) local_displ(nn, u, cn, g, ue);
        int st = element_local(etype, mat, mtype, g, xe, props + pid[e] * stride,
                               k6rot ? k6rot[e] : 100., alpha ? alpha[e] : 0.7,
                               hg ? hg + 5 * e : hg_default, ue, stress, K);
        if (st) { err |= st; continue; }
        /* Kg = R6 K R6^T with R6 = blockdiag(R, R) per node */
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                int bi = i / 3 * 3, bj = j / 3 * 3, a = i % 3, b = j % 3;
                double s = 0;
                for (int p = 0; p < 3; p++)
                    for (int q = 0; q < 3; q++)
                        s += g[a * 3 + p] * K[(bi + p) * n + bj + q] * g[b * 3 + q];
                Kg[i * n + j] = s;
            }
        int64_t k0 = init_k[e];
        for (int k = 0; k < ns; k++) {
            if (!values_only) {
                r[k0 + k] = 6 * cn[rl[k] / 6] + rl[k] % 6;
                c[k0 + k] = 6 * cn[cl[k] / 6] + cl[k] % 6;
            }
            v[k0 + k] += Kg[rl[k] * n + cl[k]];
        }
    }
    return err;
}

/*
 * Internal forces: fint[c_I + d] += (R6 K_e ue)_I for every element, in
 * element order (update_probe_ue -> update_probe_xe -> update_fint:
 * quad4.pyx:1174-1202, 1316-1363; quad4r.pyx:549, 4620; tria3r.pyx:549,
 * 2975; beamc.pyx:352, 1354).
 */
int orc_fint(int etype, int64_t E, const double *x, const int64_t *conn, const double *evec,
             const int64_t *pid, const double *props, const double *k6rot, const double *alpha,
             const double *hg, const double *u, double *fint)
{
    int nn = nnodes(etype), n = 6 * nn;
    int stride = (IS_BEAM(etype) || etype == ET_TRUSS) ? BP_STRIDE
                 : (etype == ET_SPRING ? SPR_STRIDE : SP_STRIDE);
    static const double hg_default[5] = {1., 1., 1., 1., 1.};
    for (int64_t e = 0; e < E; e++) {
        const int64_t *cn = conn + e * nn;
        double g[14], xe[12], ue[24], K[576], fe[24];
        element_geometry(etype, x, cn, evec ? evec + evec_stride(etype) * e : NULL, g, xe);
        local_displ(nn, u, cn, g, ue);
        int st = element_local(etype, MAT_FINT, 0, g, xe, props + pid[e] * stride,
                               k6rot ? k6rot[e] : 100., alpha ? alpha[e] : 0.7,
                               hg ? hg + 5 * e : hg_default, ue, NULL, K);
        if (st) return st;
        for (int i = 0; i < n; i++) {
            double s = 0;
            for (int j = 0; j < n; j++) s += K[i * n + j] * ue[j];
            fe[i] = s;
        }
        for (int I = 0; I < nn; I++)
            for (int blk = 0; blk < 2; blk++)
                for (int a = 0; a < 3; a++) {
                    double s = 0;
                    for (int i = 0; i < 3; i++) s += g[a * 3 + i] * fe[6 * I + 3 * blk + i];
                    fint[6 * cn[I] + 3 * blk + a] += s;
                }
    }
    return 0;
}

/* threads of the element loop (1 = the reference's serial loop); returns
 * the count in use */
int orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

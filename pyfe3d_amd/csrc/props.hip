// Per-element shell property records from a thickness field, on the device.
//
// The reference's design-sweep callers set the laminate terms of one
// ShellProp per element from the design variables (isotropic_plate per
// thickness: shellprop_utils.py:182-206, ShellProp.calc_constitutive_matrix
// shellprop.pyx:568-621).  For a single isotropic layer without offset the
// terms scale with the thickness h: A, E44..E55 and intrho with h, B and
// intrhoz are zero, D and intrhoz2 with h^3; the shear correction factors
// do not depend on h.  The kernel scales the record of a unit-thickness
// plate, so a new thickness field costs one streaming pass instead of a
// host loop over ShellProp objects.  Likewise per-element ply stacks
// (laminate_records) and per-element lamination parameters (lp_records).
#include <stdint.h>

#include "../../include/pf3d.h"
#include "pf3d_device.hpp"

int pf3d_fail(int code, const char *fmt, ...);
int pf3d_hip_check(hipError_t e, const char *where);

namespace {

struct Unit {
    double f[pf3d::SHELL_STRIDE];
};

// field -> power of h (record order A(6) B(6) D(6) E44 E45 E55 k13 k23 h
// offset intrho intrhoz intrhoz2, 4 unused)
__device__ __forceinline__ int hpow(int k)
{
    if (k < 6) return 1;
    if (k < 12) return 2;
    if (k < 18) return 3;
    if (k < 21) return 1;
    if (k < 23) return 0;
    if (k == 23) return 1;  // h
    if (k == 24) return 1;  // offset
    if (k == 25) return 1;  // intrho
    if (k == 26) return 2;  // intrhoz
    if (k == 27) return 3;  // intrhoz2
    return 0;
}

__global__ void iso_records(int64_t n, const double *__restrict__ h, Unit u,
                            double *__restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * pf3d::SHELL_STRIDE) return;
    const int64_t e = t / pf3d::SHELL_STRIDE;
    const int k = (int)(t - e * pf3d::SHELL_STRIDE);
    const double he = h[e];
    const int pw = hpow(k);
    const double s = pw == 0 ? 1. : pw == 1 ? he : pw == 2 ? he * he : he * he * he;
    out[t] = u.f[k] * s;
}

// Laminate record of one element from its ply stack: ply stiffness in
// laminate axes (Lamina.rebuild, shellprop.pyx:278-337), through-thickness
// integration of A, B, D, E and the mass moments
// (ShellProp.calc_constitutive_matrix, shellprop.pyx:568-621) and the
// Vlachoutsis shear correction factors (ShellProp.calc_scf,
// shellprop.pyx:485-548: k = R^2 / (h sum_k D_k^2 int (s^2 - sb^2)^2 ds / 4)).  Material
// rows: e1 e2 nu12 g12 g13 g23 rho (8 doubles, last unused).  Plies of zero
// thickness are padding and skipped.
__global__ void __launch_bounds__(128)
    laminate_records(int64_t n, int np_, const double *__restrict__ theta,
                     const double *__restrict__ plyt, const int32_t *__restrict__ ply_mat,
                     const double *__restrict__ mats, const double *__restrict__ offset,
                     int calc_scf, double *__restrict__ out)
{
    // no FMA contraction: the stiffness sums round every product, as the
    // reference's ply loop does
#pragma clang fp contract(off)
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double *th = theta + e * np_, *tk = plyt + e * np_;
    const int32_t *pm = ply_mat ? ply_mat + e * np_ : nullptr;
    const double off = offset ? offset[e] : 0.;
    const double deg = 4. * atan(1.) / 180.;
    double h = 0.;
    for (int k = 0; k < np_; k++) h += tk[k];
    double A[6] = {0, 0, 0, 0, 0, 0}, B[6] = {0, 0, 0, 0, 0, 0}, D[6] = {0, 0, 0, 0, 0, 0};
    double E44 = 0., E45 = 0., E55 = 0., ir = 0., irz = 0., irz2 = 0.;
    // shear correction sums (calc_scf), accumulated in the same ply loop
    double D1 = 0., R1 = 0., den1 = 0., D2 = 0., R2 = 0., den2 = 0.;
    const double zb = -h / 2. + off;
    double h0 = zb;
    for (int k = 0; k < np_; k++) {
        if (tk[k] == 0.) continue;
        const double *m = mats + 8 * (pm ? pm[k] : 0);
        const double e1 = m[0], e2 = m[1], nu12 = m[2], g12 = m[3], g13 = m[4], g23 = m[5];
        const double nu21 = nu12 * e2 / e1;
        const double t = th[k] * deg;
        const double c = cos(t), s = sin(t);
        const double c2 = c * c, c3 = c2 * c, c4 = c2 * c2, s2 = s * s, s3 = s2 * s, s4 = s2 * s2;
        const double d = 1 - nu12 * nu21;
        const double q11 = e1 / d, q12 = nu12 * e2 / d, q22 = e2 / d;
        const double q44 = g23, q55 = g13, q66 = g12;
        double qL[6];
        qL[0] = q11 * c4 + 2 * (q12 + 2 * q66) * s2 * c2 + q22 * s4;
        qL[1] = (q11 + q22 - 4 * q66) * s2 * c2 + q12 * (s4 + c4);
        qL[2] = (q11 - q12 - 2 * q66) * s * c3 + (q12 - q22 + 2 * q66) * s3 * c;
        qL[3] = q11 * s4 + 2 * (q12 + 2 * q66) * s2 * c2 + q22 * c4;
        qL[4] = (q11 - q12 - 2 * q66) * s3 * c + (q12 - q22 + 2 * q66) * s * c3;
        qL[5] = (q11 + q22 - 2 * q12 - 2 * q66) * s2 * c2 + q66 * (s4 + c4);
        const double q44L = q44 * c2 + q55 * s2, q45L = (q55 - q44) * s * c;
        const double q55L = q55 * c2 + q44 * s2;
        const double hk_1 = h0;
        h0 += tk[k];
        const double hk = h0, rho = m[6];
        ir += rho * (hk - hk_1);
        irz += rho * (hk * hk / 2. - hk_1 * hk_1 / 2.);
        irz2 += rho * (hk * hk * hk / 3. - hk_1 * hk_1 * hk_1 / 3.);
        const double d1 = hk - hk_1;
        for (int q = 0; q < 6; q++) {
            A[q] += qL[q] * d1;
            B[q] += 1 / 2. * qL[q] * (hk * hk - hk_1 * hk_1);
            D[q] += 1 / 3. * qL[q] * (hk * hk * hk - hk_1 * hk_1 * hk_1);
        }
        E44 += q44L * d1;
        E45 += q45L * d1;
        E55 += q55L * d1;
        if (calc_scf) {
            // ply moduli along the laminate axes as calc_scf defines them
            // (linear in cos and sin), running sums D1, D2 over the plies;
            // int (s^2 - sb^2)^2 ds over the ply (s = z - offset, sb the
            // bottom surface) by 3-point Gauss-Legendre, exact for this
            // degree-4 integrand and free of the cancellation of its
            // expanded polynomial
            const double se1 = e1 * c + e2 * s, se2 = e2 * c + e1 * s;
            const double snu12 = nu12 * c + nu21 * s, snu21 = nu21 * c + nu12 * s;
            const double s1 = hk_1 - off, s2 = hk - off, sb = zb - off;
            const double half = (s2 - s1) / 2., mid = (s2 + s1) / 2.;
            const double xg = sqrt(0.6);
            const double ga = (mid - half * xg - sb) * (mid - half * xg + sb);
            const double gb = (mid - sb) * (mid + sb);
            const double gc = (mid + half * xg - sb) * (mid + half * xg + sb);
            const double F = half * (ga * ga * (5. / 9.) + gb * gb * (8. / 9.) + gc * gc * (5. / 9.));
            const double r3 = (s2 * s2 * s2 - s1 * s1 * s1) / 3.;
            const double dd = 1 - snu12 * snu21;
            D1 += se1 / dd;
            R1 += D1 * r3;
            den1 += D1 * D1 * F;
            D2 += se2 / dd;
            R2 += D2 * r3;
            den2 += D2 * D2 * F;
        }
    }
    double k13 = 5 / 6., k23 = 5 / 6.;
    if (calc_scf) {
        k13 = R1 * R1 / (h * den1 / 4.);
        k23 = R2 * R2 / (h * den2 / 4.);
    }
    double *o = out + e * pf3d::SHELL_STRIDE;
    for (int q = 0; q < 6; q++) {
        o[q] = A[q];
        o[6 + q] = B[q];
        o[12 + q] = D[q];
    }
    o[18] = E44; o[19] = E45; o[20] = E55;
    o[21] = k13; o[22] = k23;
    o[23] = h; o[24] = off;
    o[25] = ir; o[26] = irz; o[27] = irz2;
    o[28] = o[29] = o[30] = o[31] = 0.;
}

// Record of one element from its thickness h and lamination parameters
// xi = (xiA1..4, xiB1..4, xiD1..4, xiE1, xiE2) for one material given by its
// invariants (U1..U7, rho) (shellprop_from_LaminationParameters,
// shellprop.pyx:767-825): A = h (U0 + G xiA), B = h^2/4 G xiB,
// D = h^3/12 (U0 + G xiD), E = h (U6 + U7 xiE1, -U7 xiE2, U6 - U7 xiE1) with
// U0 = (U1, U4, 0, U1, 0, U5) over (11, 12, 16, 22, 26, 66) and G the
// gradient rows below.  One thread per record field: the 32 threads of an
// element read its 15 inputs from cache and write 256 contiguous bytes.
// Shear factors 5/6, offset 0; mass moments of a homogeneous plate.
__global__ void __launch_bounds__(256)
    lp_records(int64_t n, const double *__restrict__ thick, const double *__restrict__ xi,
               const int32_t *__restrict__ mat_id, const double *__restrict__ mats,
               double *__restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * pf3d::SHELL_STRIDE) return;
    const int64_t e = t / pf3d::SHELL_STRIDE;
    const int k = (int)(t - e * pf3d::SHELL_STRIDE);
    const double *u = mats + 8 * (mat_id ? mat_id[e] : 0);
    const double h = thick[e];
    const double *x = xi + 14 * e;
    double v = 0.;
    if (k < 18) {
        const int q = k % 6, grp = k / 6;
        const double *xg = x + 4 * grp;
        const double u1 = u[0], u2 = u[1], u3 = u[2], u4 = u[3], u5 = u[4];
        // U0[q] + sum_j G[q][j] xi_j, G rows: (u2 0 u3 0) (0 0 -u3 0)
        // (0 u2/2 0 u3) (-u2 0 u3 0) (0 u2/2 0 -u3) (0 0 -u3 0)
        double base, lin;
        switch (q) {
        case 0: base = u1; lin = u2 * xg[0] + u3 * xg[2]; break;
        case 1: base = u4; lin = -u3 * xg[2]; break;
        case 2: base = 0.; lin = 0.5 * u2 * xg[1] + u3 * xg[3]; break;
        case 3: base = u1; lin = -u2 * xg[0] + u3 * xg[2]; break;
        case 4: base = 0.; lin = 0.5 * u2 * xg[1] - u3 * xg[3]; break;
        default: base = u5; lin = -u3 * xg[2];
        }
        if (grp == 0) v = h * (base + lin);
        else if (grp == 1) v = h * h / 4. * lin;
        else v = h * h * h / 12. * (base + lin);
    } else if (k == 18) {
        v = h * (u[5] + u[6] * x[12]);
    } else if (k == 19) {
        v = h * (-u[6] * x[13]);
    } else if (k == 20) {
        v = h * (u[5] - u[6] * x[12]);
    } else if (k == 21 || k == 22) {
        v = 5. / 6.;
    } else if (k == 23) {
        v = h;
    } else if (k == 25) {
        v = u[7] * h;
    } else if (k == 27) {
        v = u[7] * h * h * h / 12.;
    }
    out[t] = v;
}

}  // namespace

extern "C" int pf3d_laminate_records(int64_t n, int n_plies, const double *theta_deg,
                                     const double *plyt, const int32_t *ply_mat, int n_mat,
                                     const double *mats, const double *offset, int calc_scf,
                                     double *out, void *stream)
{
    if (n < 0 || n_plies <= 0 || n_mat <= 0)
        return pf3d_fail(PF3D_E_ARG, "pf3d_laminate_records: bad sizes");
    if (n == 0) return PF3D_OK;
    if (!theta_deg || !plyt || !mats || !out)
        return pf3d_fail(PF3D_E_ARG, "pf3d_laminate_records: null argument");
    hipLaunchKernelGGL(laminate_records, dim3((unsigned)((n + 127) / 128)), dim3(128), 0,
                       (hipStream_t)stream, n, n_plies, theta_deg, plyt, ply_mat, mats, offset,
                       calc_scf, out);
    return pf3d_hip_check(hipGetLastError(), "pf3d_laminate_records");
}

extern "C" int pf3d_isotropic_records(int64_t n, const double *h, const double *unit_record,
                                      double *out, void *stream)
{
    if (n < 0) return pf3d_fail(PF3D_E_ARG, "negative n");
    if (n == 0) return PF3D_OK;
    if (!h || !unit_record || !out) return pf3d_fail(PF3D_E_ARG, "null argument");
    Unit u;
    for (int k = 0; k < pf3d::SHELL_STRIDE; k++) u.f[k] = unit_record[k];
    if (u.f[23] != 1.) return pf3d_fail(PF3D_E_ARG, "unit_record must have h = 1");
    const int64_t nt = n * pf3d::SHELL_STRIDE;
    hipLaunchKernelGGL(iso_records, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, n, h, u, out);
    return pf3d_hip_check(hipGetLastError(), "pf3d_isotropic_records");
}

extern "C" int pf3d_lp_records(int64_t n, const double *thickness, const double *xi,
                               const int32_t *mat_id, int n_mat, const double *mats,
                               double *out, void *stream)
{
    if (n < 0 || n_mat <= 0) return pf3d_fail(PF3D_E_ARG, "pf3d_lp_records: bad sizes");
    if (n == 0) return PF3D_OK;
    if (!thickness || !xi || !mats || !out)
        return pf3d_fail(PF3D_E_ARG, "pf3d_lp_records: null argument");
    const int64_t nt = n * pf3d::SHELL_STRIDE;
    hipLaunchKernelGGL(lp_records, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, n, thickness, xi, mat_id, mats, out);
    return pf3d_hip_check(hipGetLastError(), "pf3d_lp_records");
}

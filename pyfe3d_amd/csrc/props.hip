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
// host loop over ShellProp objects.
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

}  // namespace

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

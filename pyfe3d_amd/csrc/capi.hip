// C-ABI entry points: argument checks, error reporting and kernel launch.
#include <stdarg.h>
#include <stdio.h>
#include <string>

#include "../../include/pf3d.h"
#include "pf3d_device.hpp"

// launchers defined next to the kernels
hipError_t pf3d_launch_shell(int etype, int mat, const pf3d::AsmArgs &a, hipStream_t s);
hipError_t pf3d_launch_beam(int etype, int mat, const pf3d::AsmArgs &a, hipStream_t s);
hipError_t pf3d_launch_shell_fe(int etype, const pf3d::AsmArgs &a, double *fe, hipStream_t s);
hipError_t pf3d_launch_beam_fe(int etype, const pf3d::AsmArgs &a, double *fe, hipStream_t s);
hipError_t pf3d_launch_rod(int etype, int mat, const pf3d::AsmArgs &a, hipStream_t s);
hipError_t pf3d_launch_mixed(int mat, const pf3d::AsmArgs *groups, hipStream_t s);
hipError_t pf3d_launch_rod_fe(int etype, const pf3d::AsmArgs &a, double *fe, hipStream_t s);
hipError_t pf3d_launch_geometry(int etype, int64_t n, const double *x, const int64_t *conn,
                                const double *evec, double *geom, double *xe, hipStream_t s);

static thread_local std::string g_err;

namespace {
// One wave per element: the lanes read its `size` slots (coalesced) and
// the first lane counts the element once if any value is not finite
// (the reference divides by the area without a check, quad4.pyx:1-7, so a
// degenerate element writes inf/nan silently).
__global__ void __launch_bounds__(256) nonfinite_elements(int64_t n_elem, int size,
                                                          const int64_t *__restrict__ init_k,
                                                          const double *__restrict__ v,
                                                          unsigned long long *__restrict__ count)
{
    const int64_t e = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    bool bad = false;
    if (e < n_elem) {
        const int64_t k0 = init_k ? init_k[e] : (int64_t)size * e;
        for (int t = lane; t < size; t += 64) bad |= !isfinite(v[k0 + t]);
    }
    if (__any(bad) && lane == 0 && e < n_elem) atomicAdd(count, 1ULL);
}
}  // namespace

int pf3d_fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int pf3d_hip_check(hipError_t e, const char *where)
{
    if (e == hipSuccess) return PF3D_OK;
    return pf3d_fail(PF3D_E_HIP, "%s: %s", where, hipGetErrorString(e));
}

extern "C" {

int pf3d_version(void) { return 10000; }

int pf3d_nonfinite_elements(int64_t n_elem, int size, const int64_t *init_k, const double *v,
                            int64_t *count, void *stream)
{
    if (n_elem < 0 || size <= 0 || (n_elem && (!v || !count)))
        return pf3d_fail(PF3D_E_ARG, "pf3d_nonfinite_elements: bad arguments");
    if (!n_elem) return PF3D_OK;
    hipLaunchKernelGGL(nonfinite_elements, dim3((unsigned)((n_elem + 3) / 4)), dim3(256), 0,
                       (hipStream_t)stream, n_elem, size, init_k, v,
                       (unsigned long long *)count);
    return pf3d_hip_check(hipGetLastError(), "pf3d_nonfinite_elements");
}

const char *pf3d_last_error(void) { return g_err.c_str(); }

int pf3d_slot_count(int etype, int matrix, int mtype)
{
    if (etype == PF3D_TRUSS) {
        if (matrix == PF3D_KC0) return 72;
        if (matrix == PF3D_M) {
            if (mtype < 0 || mtype > 1) return pf3d_fail(PF3D_E_ARG, "Truss mtype must be 0 or 1");
            return mtype == 0 ? 144 : 36;
        }
        return pf3d_fail(PF3D_E_ARG, "Truss has KC0 and M only");
    }
    if (etype == PF3D_BEAMLR) {
        if (matrix == PF3D_KC0) return 144;
        if (matrix == PF3D_KG) return 36;
        if (matrix == PF3D_M) {
            if (mtype < 0 || mtype > 1) return pf3d_fail(PF3D_E_ARG, "BeamLR mtype must be 0 or 1");
            return mtype == 0 ? 144 : 36;
        }
        return pf3d_fail(PF3D_E_ARG, "BeamLR has no KG_given_stress");
    }
    if (etype == PF3D_SPRING) {
        if (matrix == PF3D_KC0) return 72;
        return pf3d_fail(PF3D_E_ARG, "Spring has KC0 only");
    }
    if (etype == PF3D_BEAMC) {
        if (matrix == PF3D_KC0 || matrix == PF3D_KG) return 144;
        if (matrix == PF3D_M) {
            if (mtype < 0 || mtype > 1) return pf3d_fail(PF3D_E_ARG, "BeamC mtype must be 0 or 1");
            return mtype == 0 ? 144 : 36;
        }
        return pf3d_fail(PF3D_E_ARG, "BeamC has no KG_given_stress");
    }
    if (etype < 0 || etype > PF3D_TRIA3R) return pf3d_fail(PF3D_E_ARG, "bad etype %d", etype);
    const int nn = etype == PF3D_TRIA3R ? 3 : 4;
    if (matrix >= PF3D_KA_BETA && matrix <= PF3D_CA) {
        if (etype == PF3D_TRIA3R) return pf3d_fail(PF3D_E_ARG, "Tria3R has no piston theory");
        return 144;
    }
    switch (matrix) {
    case PF3D_KC0: return 36 * nn * nn;
    case PF3D_KG:
    case PF3D_KG_STRESS: return 9 * nn * nn;
    case PF3D_M:
        if (mtype < 0 || mtype > 2) return pf3d_fail(PF3D_E_ARG, "bad mtype %d", mtype);
        return mtype == 2 ? 18 * nn * nn : 30 * nn * nn;
    }
    return pf3d_fail(PF3D_E_ARG, "bad matrix %d", matrix);
}

int pf3d_geometry(int etype, int64_t n_elem, const double *x, const int64_t *conn,
                  const double *evec, double *geom, double *xe, void *stream)
{
    if (etype < 0 || etype > PF3D_BEAMLR) return pf3d_fail(PF3D_E_ARG, "bad etype %d", etype);
    if (n_elem < 0) return pf3d_fail(PF3D_E_ARG, "negative n_elem");
    if (n_elem == 0) return PF3D_OK;
    if (!x || !conn || !geom) return pf3d_fail(PF3D_E_ARG, "null x/conn/geom");
    if ((etype == PF3D_BEAMC || etype == PF3D_BEAMLR) && !evec)
        return pf3d_fail(PF3D_E_ARG, "beams need vxy");
    if (etype == PF3D_SPRING && !evec) return pf3d_fail(PF3D_E_ARG, "Spring needs its axes");
    return pf3d_hip_check(pf3d_launch_geometry(etype, n_elem, x, conn, evec, geom, xe,
                                               (hipStream_t)stream),
                          "pf3d_geometry");
}

static int check_common(int etype, int64_t n_elem, const double *x, const int64_t *conn,
                        int64_t *r, int64_t *c, double *v, int flags)
{
    if (etype < 0 || etype > PF3D_TRIA3R) return pf3d_fail(PF3D_E_ARG, "bad shell etype %d", etype);
    if (n_elem < 0) return pf3d_fail(PF3D_E_ARG, "negative n_elem");
    if (n_elem > (int64_t)2147483647 * 16) return pf3d_fail(PF3D_E_ARG, "n_elem too large");
    if (n_elem == 0) return 1;
    if (!x || !conn || !v) return pf3d_fail(PF3D_E_ARG, "null x/conn/v");
    if (!(flags & PF3D_VALUES_ONLY) && (!r || !c))
        return pf3d_fail(PF3D_E_ARG, "null r/c without PF3D_VALUES_ONLY");
    if (flags & ~(PF3D_VALUES_ONLY | PF3D_OVERWRITE))
        return pf3d_fail(PF3D_E_ARG, "unknown flags 0x%x", flags);
    return 0;
}

int pf3d_shell_kc0(int etype, int64_t n_elem, const double *x, const int64_t *conn,
                   const int64_t *init_k, const double *xmat, const int32_t *pid,
                   const double *props, const double *k6rot, const double *alpha,
                   const double *hg, int64_t *r, int64_t *c, double *v, int flags,
                   void *stream)
{
    int st = check_common(etype, n_elem, x, conn, r, c, v, flags);
    if (st) return st > 0 ? PF3D_OK : st;
    if (!props) return pf3d_fail(PF3D_E_ARG, "null props");
    pf3d::AsmArgs a{};
    a.n_elem = n_elem; a.x = x; a.conn = conn; a.init_k = init_k; a.evec = xmat; a.pid = pid;
    a.props = props; a.k6rot = k6rot; a.alpha = alpha; a.hg = hg; a.flags = flags;
    a.r = r; a.c = c; a.v = v;
    return pf3d_hip_check(pf3d_launch_shell(etype, PF3D_KC0, a, (hipStream_t)stream),
                          "pf3d_shell_kc0");
}

int pf3d_shell_kg(int etype, int64_t n_elem, const double *x, const int64_t *conn,
                  const int64_t *init_k, const double *xmat, const int32_t *pid,
                  const double *props, const double *u, int64_t *r, int64_t *c, double *v,
                  int flags, void *stream)
{
    int st = check_common(etype, n_elem, x, conn, r, c, v, flags);
    if (st) return st > 0 ? PF3D_OK : st;
    if (!props || !u) return pf3d_fail(PF3D_E_ARG, "null props/u");
    pf3d::AsmArgs a{};
    a.n_elem = n_elem; a.x = x; a.conn = conn; a.init_k = init_k; a.evec = xmat; a.pid = pid;
    a.props = props; a.u = u; a.flags = flags; a.r = r; a.c = c; a.v = v;
    return pf3d_hip_check(pf3d_launch_shell(etype, PF3D_KG, a, (hipStream_t)stream),
                          "pf3d_shell_kg");
}

int pf3d_shell_kg_stress(int etype, int64_t n_elem, const double *x, const int64_t *conn,
                         const int64_t *init_k, const double *xmat, const double *stress,
                         int stress_per_elem, int64_t *r, int64_t *c, double *v, int flags,
                         void *stream)
{
    int st = check_common(etype, n_elem, x, conn, r, c, v, flags);
    if (st) return st > 0 ? PF3D_OK : st;
    if (!stress) return pf3d_fail(PF3D_E_ARG, "null stress");
    pf3d::AsmArgs a{};
    a.n_elem = n_elem; a.x = x; a.conn = conn; a.init_k = init_k; a.evec = xmat;
    a.pid = nullptr; a.props = nullptr; a.stress = stress;
    a.stress_per_elem = stress_per_elem; a.flags = flags; a.r = r; a.c = c; a.v = v;
    return pf3d_hip_check(pf3d_launch_shell(etype, PF3D_KG_STRESS, a, (hipStream_t)stream),
                          "pf3d_shell_kg_stress");
}

int pf3d_shell_m(int etype, int64_t n_elem, const double *x, const int64_t *conn,
                 const int64_t *init_k, const double *xmat, const int32_t *pid,
                 const double *props, int mtype, int64_t *r, int64_t *c, double *v, int flags,
                 void *stream)
{
    int st = check_common(etype, n_elem, x, conn, r, c, v, flags);
    if (st) return st > 0 ? PF3D_OK : st;
    if (!props) return pf3d_fail(PF3D_E_ARG, "null props");
    if (mtype < 0 || mtype > 2) return pf3d_fail(PF3D_E_ARG, "bad mtype %d", mtype);
    pf3d::AsmArgs a{};
    a.n_elem = n_elem; a.x = x; a.conn = conn; a.init_k = init_k; a.evec = xmat; a.pid = pid;
    a.props = props; a.mtype = mtype; a.flags = flags; a.r = r; a.c = c; a.v = v;
    return pf3d_hip_check(pf3d_launch_shell(etype, PF3D_M, a, (hipStream_t)stream),
                          "pf3d_shell_m");
}

int pf3d_beam_kc0(int64_t n_elem, const double *x, const int64_t *conn, const int64_t *init_k,
                  const double *vxy, const int32_t *pid, const double *props, int64_t *r,
                  int64_t *c, double *v, int flags, void *stream)
{
    if (n_elem < 0) return pf3d_fail(PF3D_E_ARG, "negative n_elem");
    if (n_elem == 0) return PF3D_OK;
    if (!x || !conn || !v || !vxy || !props) return pf3d_fail(PF3D_E_ARG, "null argument");
    if (!(flags & PF3D_VALUES_ONLY) && (!r || !c))
        return pf3d_fail(PF3D_E_ARG, "null r/c without PF3D_VALUES_ONLY");
    pf3d::AsmArgs a{};
    a.n_elem = n_elem; a.x = x; a.conn = conn; a.init_k = init_k; a.evec = vxy; a.pid = pid;
    a.props = props; a.flags = flags; a.r = r; a.c = c; a.v = v;
    return pf3d_hip_check(pf3d_launch_beam(PF3D_BEAMC, PF3D_KC0, a, (hipStream_t)stream), "pf3d_beam_kc0");
}

static int beam_check(int64_t n_elem, const double *x, const int64_t *conn, const double *vxy,
                      const double *props, int64_t *r, int64_t *c, double *v, int flags)
{
    if (n_elem < 0) return pf3d_fail(PF3D_E_ARG, "negative n_elem");
    if (n_elem == 0) return 1;
    if (!x || !conn || !v || !vxy || !props) return pf3d_fail(PF3D_E_ARG, "null argument");
    if (!(flags & PF3D_VALUES_ONLY) && (!r || !c))
        return pf3d_fail(PF3D_E_ARG, "null r/c without PF3D_VALUES_ONLY");
    if (flags & ~(PF3D_VALUES_ONLY | PF3D_OVERWRITE))
        return pf3d_fail(PF3D_E_ARG, "unknown flags 0x%x", flags);
    return 0;
}

int pf3d_beam_kg(int64_t n_elem, const double *x, const int64_t *conn, const int64_t *init_k,
                 const double *vxy, const int32_t *pid, const double *props, const double *u,
                 int64_t *r, int64_t *c, double *v, int flags, void *stream)
{
    int st = beam_check(n_elem, x, conn, vxy, props, r, c, v, flags);
    if (st) return st > 0 ? PF3D_OK : st;
    if (!u) return pf3d_fail(PF3D_E_ARG, "null u");
    pf3d::AsmArgs a{};
    a.n_elem = n_elem; a.x = x; a.conn = conn; a.init_k = init_k; a.evec = vxy; a.pid = pid;
    a.props = props; a.u = u; a.flags = flags; a.r = r; a.c = c; a.v = v;
    return pf3d_hip_check(pf3d_launch_beam(PF3D_BEAMC, PF3D_KG, a, (hipStream_t)stream), "pf3d_beam_kg");
}

int pf3d_beam_m(int64_t n_elem, const double *x, const int64_t *conn, const int64_t *init_k,
                const double *vxy, const int32_t *pid, const double *props, int mtype,
                int64_t *r, int64_t *c, double *v, int flags, void *stream)
{
    int st = beam_check(n_elem, x, conn, vxy, props, r, c, v, flags);
    if (st) return st > 0 ? PF3D_OK : st;
    if (mtype < 0 || mtype > 1) return pf3d_fail(PF3D_E_ARG, "BeamC mtype must be 0 or 1");
    pf3d::AsmArgs a{};
    a.n_elem = n_elem; a.x = x; a.conn = conn; a.init_k = init_k; a.evec = vxy; a.pid = pid;
    a.props = props; a.mtype = mtype; a.flags = flags; a.r = r; a.c = c; a.v = v;
    return pf3d_hip_check(pf3d_launch_beam(PF3D_BEAMC, PF3D_M, a, (hipStream_t)stream), "pf3d_beam_m");
}

int pf3d_shell_piston(int etype, int matrix, int64_t n_elem, const double *x,
                      const int64_t *conn, const int64_t *init_k, int64_t *r, int64_t *c,
                      double *v, int flags, void *stream)
{
    if (etype != PF3D_QUAD4 && etype != PF3D_QUAD4R)
        return pf3d_fail(PF3D_E_ARG, "piston theory: etype %d is not Quad4/Quad4R", etype);
    if (matrix < PF3D_KA_BETA || matrix > PF3D_CA)
        return pf3d_fail(PF3D_E_ARG, "piston theory: bad matrix %d", matrix);
    int st = check_common(etype, n_elem, x, conn, r, c, v, flags);
    if (st) return st > 0 ? PF3D_OK : st;
    pf3d::AsmArgs a{};
    a.n_elem = n_elem; a.x = x; a.conn = conn; a.init_k = init_k; a.flags = flags;
    a.r = r; a.c = c; a.v = v;
    return pf3d_hip_check(pf3d_launch_shell(etype, matrix, a, (hipStream_t)stream),
                          "pf3d_shell_piston");
}

int pf3d_beamlr_kc0(int64_t n_elem, const double *x, const int64_t *conn, const int64_t *init_k,
                    const double *vxy, const int32_t *pid, const double *props, int64_t *r,
                    int64_t *c, double *v, int flags, void *stream)
{
    int st = beam_check(n_elem, x, conn, vxy, props, r, c, v, flags);
    if (st) return st > 0 ? PF3D_OK : st;
    pf3d::AsmArgs a{};
    a.n_elem = n_elem; a.x = x; a.conn = conn; a.init_k = init_k; a.evec = vxy; a.pid = pid;
    a.props = props; a.flags = flags; a.r = r; a.c = c; a.v = v;
    return pf3d_hip_check(pf3d_launch_beam(PF3D_BEAMLR, PF3D_KC0, a, (hipStream_t)stream),
                          "pf3d_beamlr_kc0");
}

int pf3d_beamlr_kg(int64_t n_elem, const double *x, const int64_t *conn, const int64_t *init_k,
                   const double *vxy, const int32_t *pid, const double *props, const double *u,
                   int64_t *r, int64_t *c, double *v, int flags, void *stream)
{
    int st = beam_check(n_elem, x, conn, vxy, props, r, c, v, flags);
    if (st) return st > 0 ? PF3D_OK : st;
    if (!u) return pf3d_fail(PF3D_E_ARG, "null u");
    pf3d::AsmArgs a{};
    a.n_elem = n_elem; a.x = x; a.conn = conn; a.init_k = init_k; a.evec = vxy; a.pid = pid;
    a.props = props; a.u = u; a.flags = flags; a.r = r; a.c = c; a.v = v;
    return pf3d_hip_check(pf3d_launch_beam(PF3D_BEAMLR, PF3D_KG, a, (hipStream_t)stream),
                          "pf3d_beamlr_kg");
}

int pf3d_beamlr_m(int64_t n_elem, const double *x, const int64_t *conn, const int64_t *init_k,
                  const double *vxy, const int32_t *pid, const double *props, int mtype,
                  int64_t *r, int64_t *c, double *v, int flags, void *stream)
{
    int st = beam_check(n_elem, x, conn, vxy, props, r, c, v, flags);
    if (st) return st > 0 ? PF3D_OK : st;
    if (mtype < 0 || mtype > 1) return pf3d_fail(PF3D_E_ARG, "BeamLR mtype must be 0 or 1");
    pf3d::AsmArgs a{};
    a.n_elem = n_elem; a.x = x; a.conn = conn; a.init_k = init_k; a.evec = vxy; a.pid = pid;
    a.props = props; a.mtype = mtype; a.flags = flags; a.r = r; a.c = c; a.v = v;
    return pf3d_hip_check(pf3d_launch_beam(PF3D_BEAMLR, PF3D_M, a, (hipStream_t)stream),
                          "pf3d_beamlr_m");
}

int pf3d_beamlr_fe(int64_t n_elem, const double *x, const int64_t *conn, const double *vxy,
                   const int32_t *pid, const double *props, const double *u, double *fe,
                   void *stream)
{
    if (n_elem < 0) return pf3d_fail(PF3D_E_ARG, "negative n_elem");
    if (n_elem > (int64_t)2147483647 * 16) return pf3d_fail(PF3D_E_ARG, "n_elem too large");
    if (n_elem == 0) return PF3D_OK;
    if (!x || !conn || !vxy || !props || !u || !fe) return pf3d_fail(PF3D_E_ARG, "null argument");
    pf3d::AsmArgs a{};
    a.n_elem = n_elem; a.x = x; a.conn = conn; a.evec = vxy; a.pid = pid; a.props = props;
    a.u = u;
    return pf3d_hip_check(pf3d_launch_beam_fe(PF3D_BEAMLR, a, fe, (hipStream_t)stream),
                          "pf3d_beamlr_fe");
}

int pf3d_shell_fe(int etype, int64_t n_elem, const double *x, const int64_t *conn,
                  const double *xmat, const int32_t *pid, const double *props,
                  const double *k6rot, const double *alpha, const double *hg, const double *u,
                  double *fe, void *stream)
{
    if (etype < 0 || etype > PF3D_TRIA3R) return pf3d_fail(PF3D_E_ARG, "bad shell etype %d", etype);
    if (n_elem < 0) return pf3d_fail(PF3D_E_ARG, "negative n_elem");
    if (n_elem > (int64_t)2147483647 * 6) return pf3d_fail(PF3D_E_ARG, "n_elem too large");
    if (n_elem == 0) return PF3D_OK;
    if (!x || !conn || !props || !u || !fe) return pf3d_fail(PF3D_E_ARG, "null x/conn/props/u/fe");
    pf3d::AsmArgs a{};
    a.n_elem = n_elem; a.x = x; a.conn = conn; a.evec = xmat; a.pid = pid; a.props = props;
    a.k6rot = k6rot; a.alpha = alpha; a.hg = hg; a.u = u;
    return pf3d_hip_check(pf3d_launch_shell_fe(etype, a, fe, (hipStream_t)stream),
                          "pf3d_shell_fe");
}

int pf3d_beam_fe(int64_t n_elem, const double *x, const int64_t *conn, const double *vxy,
                 const int32_t *pid, const double *props, const double *u, double *fe,
                 void *stream)
{
    if (n_elem < 0) return pf3d_fail(PF3D_E_ARG, "negative n_elem");
    if (n_elem > (int64_t)2147483647 * 16) return pf3d_fail(PF3D_E_ARG, "n_elem too large");
    if (n_elem == 0) return PF3D_OK;
    if (!x || !conn || !vxy || !props || !u || !fe) return pf3d_fail(PF3D_E_ARG, "null argument");
    pf3d::AsmArgs a{};
    a.n_elem = n_elem; a.x = x; a.conn = conn; a.evec = vxy; a.pid = pid; a.props = props;
    a.u = u;
    return pf3d_hip_check(pf3d_launch_beam_fe(PF3D_BEAMC, a, fe, (hipStream_t)stream), "pf3d_beam_fe");
}

int pf3d_mixed_assemble(int matrix, int mtype, int n_groups, const pf3d_elem_group *groups,
                        const double *x, const double *u, int64_t *r, int64_t *c, double *v,
                        int flags, void *stream)
{
    if (matrix != PF3D_KC0 && matrix != PF3D_KG && matrix != PF3D_M)
        return pf3d_fail(PF3D_E_ARG, "mixed assembly takes KC0, KG or M (matrix %d)", matrix);
    if (n_groups < 0 || n_groups > 4) return pf3d_fail(PF3D_E_ARG, "n_groups %d", n_groups);
    if (n_groups > 0 && !groups) return pf3d_fail(PF3D_E_ARG, "null groups");
    if (!x || !v) return pf3d_fail(PF3D_E_ARG, "null x/v");
    if (!(flags & PF3D_VALUES_ONLY) && (!r || !c))
        return pf3d_fail(PF3D_E_ARG, "null r/c without PF3D_VALUES_ONLY");
    if (flags & ~(PF3D_VALUES_ONLY | PF3D_OVERWRITE))
        return pf3d_fail(PF3D_E_ARG, "unknown flags 0x%x", flags);
    if (matrix == PF3D_KG && !u) return pf3d_fail(PF3D_E_ARG, "null u");
    pf3d::AsmArgs a[4] = {};
    bool seen[4] = {false, false, false, false};
    for (int i = 0; i < n_groups; i++) {
        const pf3d_elem_group &g = groups[i];
        if (g.etype < PF3D_QUAD4 || g.etype > PF3D_BEAMC)
            return pf3d_fail(PF3D_E_ARG, "group %d: etype %d is not a shell or BeamC", i, g.etype);
        if (seen[g.etype]) return pf3d_fail(PF3D_E_ARG, "group %d: etype %d repeated", i, g.etype);
        seen[g.etype] = true;
        if (g.n_elem < 0 || g.k0 < 0) return pf3d_fail(PF3D_E_ARG, "group %d: negative size", i);
        if (g.n_elem == 0) continue;
        if (!g.conn || !g.props) return pf3d_fail(PF3D_E_ARG, "group %d: null conn/props", i);
        if (g.etype == PF3D_BEAMC && !g.evec)
            return pf3d_fail(PF3D_E_ARG, "group %d: BeamC needs vxy", i);
        if (matrix == PF3D_M && (mtype < 0 || mtype > (g.etype == PF3D_BEAMC ? 1 : 2)))
            return pf3d_fail(PF3D_E_ARG, "group %d: bad mtype %d", i, mtype);
        pf3d::AsmArgs &t = a[g.etype];
        t.n_elem = g.n_elem; t.x = x; t.conn = g.conn; t.init_k = g.init_k; t.evec = g.evec;
        t.pid = g.pid; t.props = g.props; t.k6rot = g.k6rot; t.alpha = g.alpha; t.hg = g.hg;
        t.u = u; t.mtype = mtype; t.flags = flags;
        t.v = v + g.k0;
        t.r = r ? r + g.k0 : nullptr;
        t.c = c ? c + g.k0 : nullptr;
    }
    for (int t = 0; t < 4; t++) a[t].flags = flags;
    return pf3d_hip_check(pf3d_launch_mixed(matrix, a, (hipStream_t)stream),
                          "pf3d_mixed_assemble");
}

static int rod_check(int64_t n_elem, const double *x, const int64_t *conn, const double *props,
                     int64_t *r, int64_t *c, double *v, int flags)
{
    if (n_elem < 0) return pf3d_fail(PF3D_E_ARG, "negative n_elem");
    if (n_elem > (int64_t)2147483647 * 64) return pf3d_fail(PF3D_E_ARG, "n_elem too large");
    if (n_elem == 0) return 1;
    if (!x || !conn || !v || !props) return pf3d_fail(PF3D_E_ARG, "null argument");
    if (!(flags & PF3D_VALUES_ONLY) && (!r || !c))
        return pf3d_fail(PF3D_E_ARG, "null r/c without PF3D_VALUES_ONLY");
    if (flags & ~(PF3D_VALUES_ONLY | PF3D_OVERWRITE))
        return pf3d_fail(PF3D_E_ARG, "unknown flags 0x%x", flags);
    return 0;
}

int pf3d_truss_kc0(int64_t n_elem, const double *x, const int64_t *conn, const int64_t *init_k,
                   const int32_t *pid, const double *props, int64_t *r, int64_t *c, double *v,
                   int flags, void *stream)
{
    int st = rod_check(n_elem, x, conn, props, r, c, v, flags);
    if (st) return st > 0 ? PF3D_OK : st;
    pf3d::AsmArgs a{};
    a.n_elem = n_elem; a.x = x; a.conn = conn; a.init_k = init_k; a.pid = pid; a.props = props;
    a.flags = flags; a.r = r; a.c = c; a.v = v;
    return pf3d_hip_check(pf3d_launch_rod(PF3D_TRUSS, PF3D_KC0, a, (hipStream_t)stream),
                          "pf3d_truss_kc0");
}

int pf3d_truss_m(int64_t n_elem, const double *x, const int64_t *conn, const int64_t *init_k,
                 const int32_t *pid, const double *props, int mtype, int64_t *r, int64_t *c,
                 double *v, int flags, void *stream)
{
    int st = rod_check(n_elem, x, conn, props, r, c, v, flags);
    if (st) return st > 0 ? PF3D_OK : st;
    if (mtype < 0 || mtype > 1) return pf3d_fail(PF3D_E_ARG, "Truss mtype must be 0 or 1");
    pf3d::AsmArgs a{};
    a.n_elem = n_elem; a.x = x; a.conn = conn; a.init_k = init_k; a.pid = pid; a.props = props;
    a.mtype = mtype; a.flags = flags; a.r = r; a.c = c; a.v = v;
    return pf3d_hip_check(pf3d_launch_rod(PF3D_TRUSS, PF3D_M, a, (hipStream_t)stream),
                          "pf3d_truss_m");
}

int pf3d_spring_kc0(int64_t n_elem, const double *x, const int64_t *conn, const int64_t *init_k,
                    const double *axes, const int32_t *pid, const double *k, int64_t *r,
                    int64_t *c, double *v, int flags, void *stream)
{
    int st = rod_check(n_elem, x, conn, k, r, c, v, flags);
    if (st) return st > 0 ? PF3D_OK : st;
    if (!axes) return pf3d_fail(PF3D_E_ARG, "null axes");
    pf3d::AsmArgs a{};
    a.n_elem = n_elem; a.x = x; a.conn = conn; a.init_k = init_k; a.evec = axes; a.pid = pid;
    a.props = k; a.flags = flags; a.r = r; a.c = c; a.v = v;
    return pf3d_hip_check(pf3d_launch_rod(PF3D_SPRING, PF3D_KC0, a, (hipStream_t)stream),
                          "pf3d_spring_kc0");
}

int pf3d_rod_fe(int etype, int64_t n_elem, const double *x, const int64_t *conn,
                const double *axes, const int32_t *pid, const double *props, const double *u,
                double *fe, void *stream)
{
    if (etype != PF3D_TRUSS && etype != PF3D_SPRING)
        return pf3d_fail(PF3D_E_ARG, "bad rod etype %d", etype);
    if (n_elem < 0) return pf3d_fail(PF3D_E_ARG, "negative n_elem");
    if (n_elem > (int64_t)2147483647 * 64) return pf3d_fail(PF3D_E_ARG, "n_elem too large");
    if (n_elem == 0) return PF3D_OK;
    if (!x || !conn || !props || !u || !fe) return pf3d_fail(PF3D_E_ARG, "null argument");
    if (etype == PF3D_SPRING && !axes) return pf3d_fail(PF3D_E_ARG, "null axes");
    pf3d::AsmArgs a{};
    a.n_elem = n_elem; a.x = x; a.conn = conn; a.evec = axes; a.pid = pid; a.props = props;
    a.u = u;
    return pf3d_hip_check(pf3d_launch_rod_fe(etype, a, fe, (hipStream_t)stream), "pf3d_rod_fe");
}

}  // extern "C"

"""Smoothed-aggregation algebraic multigrid (SA-AMG) preconditioned CG for
the Dirichlet-treated 6-DOF-per-node stiffness (a stronger preconditioner
than block Jacobi for the static solve that pyfe3d's callers run with
``spsolve``/``cg``, tests/test_quad4_static_point_load.py:84-102).

Setup (once per matrix):
  * host, O(nodes): geometric aggregation (nodes, then aggregates by
    centroid, grouped in boxes of ``box`` times the typical edge length);
    the near-null space (the six rigid-body modes of each node:
    translations, rotations about the centroid with the rotational DOFs
    themselves, zero on fixed DOFs); the tentative prolongator's blocks, by
    a rank-revealing orthonormal factorisation (SVD) of each aggregate's
    modes (coarse modes = its R);
  * GPU, 6x6-block sparse kernels: inverted diagonal blocks
    (pf3d_bjacobi), lambda_max(D^-1 A) by power iteration, A P0
    (pf3d_bsr_spgemm_*), the smoothed prolongator P = P0 - w D^-1 A P0 with
    w = 4/3 / lambda_max (pf3d_amg_smooth_prolongator), R = P^T
    (pf3d_bsr_transpose), the Galerkin operator R (A P), symmetrised with
    the coarse DOFs that carry no mode set to identity (pf3d_bsr_symmetrize).
Solve (GPU, this library's kernels only): PCG whose preconditioner is one
symmetric V-cycle with Chebyshev(D^-1 A) pre/post smoothing, restriction
P^T and prolongation P as 6x6-block sparse products (pf3d_bsr_spmv), and
a dense inverse on the coarsest level (pf3d_gemv_t: one workgroup per row of
the inverse); dot products with
pf3d_dot, vector updates with pf3d_axpby.  With precision 'fp32' the whole
V-cycle runs in single precision (pf3d_bsr_spmv_f32, pf3d_axpby_f32) and
the CG around it, in FP64, uses the flexible (Polak-Ribiere) beta.
"""
import ctypes
import time
import weakref

import numpy as np
import scipy.sparse as ss
import torch

from . import _lib


def _s():
    return torch.cuda.current_stream().cuda_stream


def rigid_body_modes(xyz, fixed):
    """(6 n, 6) near-null space: translations and small rotations about the
    centroid, rows of fixed DOFs zero."""
    n = xyz.shape[0]
    c = xyz - xyz.mean(axis=0)
    B = np.zeros((n, 6, 6))
    for d in range(3):
        B[:, d, d] = 1.
    x, y, z = c[:, 0], c[:, 1], c[:, 2]
    # rotation about x: (v, w) = (-z, y); about y: (u, w) = (z, -x); about z: (u, v) = (-y, x)
    B[:, 1, 3] = -z; B[:, 2, 3] = y; B[:, 3, 3] = 1.
    B[:, 0, 4] = z; B[:, 2, 4] = -x; B[:, 4, 4] = 1.
    B[:, 0, 5] = -y; B[:, 1, 5] = x; B[:, 5, 5] = 1.
    B = B.reshape(6 * n, 6)
    B[np.asarray(fixed, bool)] = 0.
    return B


def box_aggregates(xyz, size):
    """Aggregate id per point: boxes of edge `size` in space (compressed)."""
    lo = xyz.min(axis=0)
    k = np.floor((xyz - lo) / size).astype(np.int64)
    dims = k.max(axis=0) + 1
    key = (k[:, 0] * dims[1] + k[:, 1]) * dims[2] + k[:, 2]
    uniq, agg = np.unique(key, return_inverse=True)
    return agg.astype(np.int64), uniq.size


def tentative(agg, n_agg, B, tol=1e-10):
    """Blocks of the tentative prolongator P0 (block row i = Q[i] at block
    column agg[i]; orthonormal columns per aggregate), the coarse
    near-null space Bc (6 n_agg x 6) and the coarse DOFs without a mode
    (dead: bool, 6 n_agg)."""
    n = agg.size
    order = np.argsort(agg, kind='stable')
    counts = np.bincount(agg, minlength=n_agg)
    start = np.concatenate([[0], np.cumsum(counts)[:-1]])
    pos = np.empty(n, np.int64)
    pos[order] = np.arange(n) - start[agg[order]]
    m = counts.max()
    Bn = B.reshape(n, 6, 6)
    Bp = np.zeros((n_agg, m, 6, 6))
    Bp[agg, pos] = Bn
    U, S, Vt = np.linalg.svd(Bp.reshape(n_agg, 6 * m, 6), full_matrices=False)
    keep = S > tol * np.maximum(S[:, :1], 1e-300)
    U = U * keep[:, None, :]
    R = (S * keep)[:, :, None] * Vt                 # (n_agg, 6, 6): U R = Bp
    Q = U.reshape(n_agg, m, 6, 6)
    return np.ascontiguousarray(Q[agg, pos]), R.reshape(6 * n_agg, 6), ~keep.reshape(-1)


class _Dev:
    """The operator y = M x of a _Bsr matrix M on the GPU, with values and
    vectors in FP64 or FP32 (an FP32 copy of the values)."""

    def __init__(self, M, dtype=torch.float64):
        self.nb = M.nr
        self.row_ptr, self.col_idx = M.row_ptr, M.col_idx
        if dtype == torch.float64:
            self.val = M.val
            self.fn = 'pf3d_bsr_spmv'
        else:
            self.val = torch.empty(M.val.numel(), dtype=torch.float32, device=M.val.device)
            _lib.call('pf3d_to_f32', M.val.numel(), M.val.data_ptr(), self.val.data_ptr(), _s())
            self.fn = 'pf3d_bsr_spmv_f32'

    def __call__(self, x, y):
        _lib.call(self.fn, self.nb, self.row_ptr.data_ptr(), self.col_idx.data_ptr(),
                  self.val.data_ptr(), x.data_ptr(), y.data_ptr(), _s())
        return y


def _axpby(a, x, b, y):
    if y.dtype == torch.float32:
        _lib.call('pf3d_axpby_f32', y.numel(), float(a), x.data_ptr(), float(b), y.data_ptr(),
                  _s())
    else:
        _lib.call('pf3d_axpby', y.numel(), float(a), x.data_ptr(), float(b), y.data_ptr(), _s())


class _Bsr:
    """6x6 BSR matrix on the GPU: nr x nc blocks, int32 row pointers and
    ascending block columns, FP64 values (36 per block, row-major)."""

    def __init__(self, nr, nc, row_ptr, col_idx, val):
        self.nr, self.nc = int(nr), int(nc)
        self.row_ptr, self.col_idx, self.val = row_ptr, col_idx, val

    @property
    def nnzb(self):
        return self.col_idx.numel()

    def to_scipy(self):
        return ss.bsr_matrix((self.val.cpu().numpy().reshape(-1, 6, 6), self.col_idx.cpu().numpy(),
                              self.row_ptr.cpu().numpy()), shape=(6 * self.nr, 6 * self.nc))


def _work(name, dev, *args):
    nbytes = ctypes.c_int64()
    _lib.call(name, *args, ctypes.byref(nbytes))
    return torch.empty(nbytes.value, dtype=torch.uint8, device=dev), nbytes.value


def spgemm(X, Y):
    """C = X Y (pf3d_bsr_spgemm_symbolic, then pf3d_bsr_spgemm_numeric)."""
    if X.nc != Y.nr:
        raise ValueError('spgemm: inner sizes %d and %d differ' % (X.nc, Y.nr))
    dev = X.val.device
    work, nbytes = _work('pf3d_bsr_spgemm_bytes', dev, X.nr)
    cr = torch.empty(X.nr + 1, dtype=torch.int32, device=dev)
    nnz = ctypes.c_int64()
    _lib.call('pf3d_bsr_spgemm_symbolic', X.nr, X.row_ptr.data_ptr(), X.col_idx.data_ptr(),
              Y.row_ptr.data_ptr(), Y.col_idx.data_ptr(), cr.data_ptr(), work.data_ptr(), nbytes,
              ctypes.byref(nnz), _s())
    del work
    cc = torch.empty(nnz.value, dtype=torch.int32, device=dev)
    cv = torch.empty(nnz.value * 36, dtype=torch.float64, device=dev)
    _lib.call('pf3d_bsr_spgemm_numeric', X.nr, X.row_ptr.data_ptr(), X.col_idx.data_ptr(),
              X.val.data_ptr(), Y.row_ptr.data_ptr(), Y.col_idx.data_ptr(), Y.val.data_ptr(),
              cr.data_ptr(), cc.data_ptr(), cv.data_ptr(), nnz.value, _s())
    return _Bsr(X.nr, Y.nc, cr, cc, cv)


def transpose(X):
    """X^T (pf3d_bsr_transpose)."""
    dev = X.val.device
    work, nbytes = _work('pf3d_bsr_transpose_bytes', dev, X.nr, X.nc, X.nnzb)
    trp = torch.empty(X.nc + 1, dtype=torch.int32, device=dev)
    tci = torch.empty(X.nnzb, dtype=torch.int32, device=dev)
    tv = torch.empty_like(X.val)
    _lib.call('pf3d_bsr_transpose', X.nr, X.nc, X.nnzb, X.row_ptr.data_ptr(), X.col_idx.data_ptr(),
              X.val.data_ptr(), trp.data_ptr(), tci.data_ptr(), tv.data_ptr(), work.data_ptr(),
              nbytes, _s())
    return _Bsr(X.nc, X.nr, trp, tci, tv)


def _block_jacobi(A):
    """Inverted 6x6 diagonal blocks of A as a block-diagonal _Bsr."""
    dev = A.val.device
    dinv = torch.empty(A.nr * 36, dtype=torch.float64, device=dev)
    bad = torch.zeros(1, dtype=torch.int32, device=dev)
    _lib.call('pf3d_bjacobi', A.nr, A.row_ptr.data_ptr(), A.col_idx.data_ptr(), A.val.data_ptr(),
              dinv.data_ptr(), None, bad.data_ptr(), _s())
    idx = torch.arange(A.nr + 1, dtype=torch.int32, device=dev)
    return _Bsr(A.nr, A.nr, idx, idx[:A.nr], dinv)


def _lambda_max(A, D, iters=20, seed=0):
    """Power iteration on D^-1 A from a seeded uniform(-1, 1) start."""
    dev = A.val.device
    x = torch.as_tensor(np.random.default_rng(seed).uniform(-1, 1, 6 * A.nr), device=dev)
    t, y = torch.empty_like(x), torch.empty_like(x)
    s = torch.empty(1025, dtype=torch.float64, device=dev)
    opA, opD = _Dev(A), _Dev(D)
    lam = 1.
    for _ in range(iters):
        opA(x, t)
        opD(t, y)
        ny = _dot(y, y, s) ** 0.5
        lam = ny / _dot(x, x, s) ** 0.5
        _axpby(1. / ny, y, 0., x)
    return lam


def setup_hierarchy(row_ptr, col_idx, val, n, fixed, xyz, box=3.0, coarse_nodes=300,
                    max_levels=8, log=None):
    """Multigrid hierarchy of the n-node matrix (row_ptr, col_idx, val)
    (Dirichlet-treated; kept by reference as the finest level), built on
    the GPU except for the aggregation and the tentative prolongator's
    blocks.  Returns (levels, Ad): levels = list of dict(A, D, P, R, lam)
    of _Bsr matrices (D: inverted diagonal blocks of A) and the coarsest
    matrix, dense on the host."""
    t0 = time.perf_counter()
    dev = val.device
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    B = rigid_body_modes(xyz, fixed)
    rp, ci = row_ptr.cpu().numpy(), col_idx.cpu().numpy()
    rows = np.repeat(np.arange(n), np.diff(rp))
    off = rows != ci
    size = box * np.median(np.linalg.norm(xyz[rows[off]] - xyz[ci[off]], axis=1))
    del rows, off, rp, ci
    A = _Bsr(n, n, row_ptr, col_idx, val)
    levels = []
    flags = torch.zeros(2, dtype=torch.int32, device=dev)
    while A.nr > coarse_nodes and len(levels) < max_levels - 1:
        D = _block_jacobi(A)
        lam = _lambda_max(A, D)
        agg, n_agg = box_aggregates(xyz, size)
        if n_agg >= A.nr:
            break
        Q, Bc, dead = tentative(agg, n_agg, B)
        agg_d = torch.as_tensor(agg.astype(np.int32), device=dev)
        P0 = _Bsr(A.nr, n_agg, torch.arange(A.nr + 1, dtype=torch.int32, device=dev), agg_d,
                  torch.as_tensor(Q.reshape(-1), device=dev))
        T = spgemm(A, P0)                       # A P0, then P = P0 - w D^-1 A P0
        pv = torch.empty_like(T.val)
        _lib.call('pf3d_amg_smooth_prolongator', A.nr, T.row_ptr.data_ptr(), T.col_idx.data_ptr(),
                  T.val.data_ptr(), agg_d.data_ptr(), P0.val.data_ptr(), D.val.data_ptr(),
                  4. / 3. / lam, pv.data_ptr(), flags.data_ptr(), _s())
        P = _Bsr(A.nr, n_agg, T.row_ptr, T.col_idx, pv)
        del T, P0
        AP = spgemm(A, P)
        R = transpose(P)
        Ac = spgemm(R, AP)
        del AP
        acv = torch.empty_like(Ac.val)
        dead_d = torch.as_tensor(dead.astype(np.uint8), device=dev)
        _lib.call('pf3d_bsr_symmetrize', Ac.nr, Ac.nnzb, Ac.row_ptr.data_ptr(),
                  Ac.col_idx.data_ptr(), Ac.val.data_ptr(), dead_d.data_ptr(), acv.data_ptr(),
                  flags[1:].data_ptr(), _s())
        missing, unpaired = flags.tolist()
        if missing or unpaired:
            raise _lib.Pf3dError('multigrid setup: %d prolongator rows without their aggregate, '
                                 '%d coarse blocks without a transpose partner'
                                 % (missing, unpaired))
        Ac.val = acv
        levels.append(dict(A=A, D=D, P=P, R=R, lam=lam))
        cen = np.zeros((n_agg, 3))
        np.add.at(cen, agg, xyz)
        xyz = cen / np.bincount(agg, minlength=n_agg)[:, None]
        A, B = Ac, Bc
        size *= box
        if log:
            log('amg level %d: %d -> %d DOF (%.1f s)' % (len(levels) - 1, 6 * P.nr, 6 * P.nc,
                                                           time.perf_counter() - t0))
    return levels, A.to_scipy().toarray()


class SAAMG:
    """plan: BSR structure (row_ptr, col_idx, n_nodes) of A; val: A's BSR
    values after the Dirichlet treatment (identity rows/cols on fixed
    DOFs); fixed: bool per DOF; xyz: (n_nodes, 3) node coordinates."""

    def __init__(self, plan, val, fixed, xyz, box=3.0, coarse_nodes=300, max_levels=8,
                 cheb_degree=2, cheb_lower=1. / 30, precision='fp64', log=None, hierarchy=None):
        """precision 'fp32': every level's matrices, smoother and transfer
        operators in single precision (half the bytes of the V-cycle); the
        coarsest solve and the CG around it stay FP64.  hierarchy: the
        result of setup_hierarchy for this matrix, to share one setup
        between precisions."""
        t0 = time.perf_counter()
        dev = val.device
        if hierarchy is None:
            hierarchy = setup_hierarchy(plan.row_ptr, plan.col_idx, val, plan.n_nodes, fixed, xyz,
                                        box, coarse_nodes, max_levels, log)
        hier, Ad = hierarchy
        self.cheb_degree, self.cheb_lower = cheb_degree, cheb_lower
        if precision not in ('fp64', 'fp32'):
            raise ValueError(precision)
        self.precision = precision
        dt = torch.float64 if precision == 'fp64' else torch.float32
        self.levels = []
        for H in hier:
            self.levels.append(dict(A=_Dev(H['A'], dt), Dinv=_Dev(H['D'], dt), P=_Dev(H['P'], dt),
                                    R=_Dev(H['R'], dt), lam=H['lam'], n=6 * H['A'].nr))
        self.nc = Ad.shape[0]
        # row j of the inverse at coarse_inv[j * nc:]: pf3d_gemv_t forms each
        # entry of inv(Ad) b as one workgroup's dot product
        self.coarse_inv = torch.as_tensor(np.ascontiguousarray(np.linalg.inv(Ad)), device=dev)
        zeros = lambda m, t=dt: torch.zeros(m, dtype=t, device=dev)
        for i, L in enumerate(self.levels):
            m = L['n']
            L['r'], L['d'], L['t'] = zeros(m), zeros(m), zeros(m)
            # right-hand side and solution of the next level
            mc = self.levels[i + 1]['n'] if i + 1 < len(self.levels) else self.nc
            L['bc'], L['xc'] = zeros(mc), zeros(mc)
        # FP64 staging of the coarsest solve and of the fine-level vectors
        self.c64 = (zeros(self.nc, torch.float64), zeros(self.nc, torch.float64))
        if dt == torch.float32 and self.levels:
            m0 = self.levels[0]['n']
            self.f32 = (zeros(m0), zeros(m0))
        self.setup_s = time.perf_counter() - t0

    def _apply(self, r, z):
        if self.precision == 'fp64' or not self.levels:
            return self.vcycle(r, z)
        r32, z32 = self.f32
        _lib.call('pf3d_to_f32', r.numel(), r.data_ptr(), r32.data_ptr(), _s())
        self.vcycle(r32, z32)
        _lib.call('pf3d_to_f64', z.numel(), z32.data_ptr(), z.data_ptr(), _s())
        return z

    def capture(self):
        """Record one V-cycle (every launch with fixed buffers and
        coefficients; no host synchronisation inside) in a HIP graph that
        apply() replays: one graph launch instead of tens of kernel
        launches per CG iteration."""
        n = self.levels[0]['n'] if self.levels else self.nc
        dev = self.coarse_inv.device
        with _capture_guard():
            self.g_in = torch.zeros(n, dtype=torch.float64, device=dev)
            self.g_out = torch.zeros(n, dtype=torch.float64, device=dev)
            s = torch.cuda.Stream(device=dev)
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                self._apply(self.g_in, self.g_out)      # warm-up outside the graph
            torch.cuda.current_stream().wait_stream(s)
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._apply(self.g_in, self.g_out)
        self.graph = g
        return g

    def apply(self, r, z):
        """z = M^-1 r for FP64 vectors r, z (one V-cycle; the captured
        graph when capture() was called)."""
        g = getattr(self, 'graph', None)
        if g is None:
            return self._apply(r, z)
        _axpby(1., r, 0., self.g_in)
        g.replay()
        _axpby(1., self.g_out, 0., z)
        return z

    def _cheb(self, L, b, x, x_zero=False):
        """x += Chebyshev(D^-1 A) correction for A x = b (degree k, bounds
        [lower, 1.1] lambda_max); x_zero: x is known to be 0 (r = b, no
        product).  Each step is one fused pass (pf3d_bsr_fused): x += d,
        r -= A d and the next d = a D^-1 r + c d together."""
        lmax = 1.1 * L['lam']
        lmin = self.cheb_lower * L['lam']
        theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
        sigma = theta / delta
        rho_old = 1. / sigma
        A, D = L['A'], L['Dinv'].val
        r, cur, nxt = L['r'], L['d'], L['t']
        # r = b - A x (r = b from a zero x); d = D^-1 r / theta
        _fused(A, None if x_zero else x, b, r, dinv=D, dnew=cur, sa=1., sb=-1.,
               c1=1. / theta)
        for k in range(self.cheb_degree):
            if k == self.cheb_degree - 1:
                _axpby(1., cur, 1., x)             # x += d
                break
            rho = 1. / (2. * sigma - rho_old)
            # x += d; r -= A d; d = 2 rho / delta D^-1 r + rho rho_old d
            _fused(A, cur, None, r, dinv=D, dold=cur, dnew=nxt, xacc=x, sa=1., sb=-1.,
                   c1=2. * rho / delta, c2=rho * rho_old)
            cur, nxt = nxt, cur
            rho_old = rho

    def _mv_buffers(self, nv):
        cache = self.__dict__.setdefault('_mv', {})
        if nv in cache:
            return cache[nv]
        dev = self.coarse_inv.device
        dt = torch.float64 if self.precision == 'fp64' else torch.float32
        zeros = lambda m, t=dt: torch.zeros(m, dtype=t, device=dev)
        bufs = []
        for i, L in enumerate(self.levels):
            m = L['n']
            mc = self.levels[i + 1]['n'] if i + 1 < len(self.levels) else self.nc
            bufs.append(dict(r=zeros(nv * m), d=zeros(nv * m), t=zeros(nv * m),
                             bc=zeros(nv * mc), xc=zeros(nv * mc)))
        extra = dict(c64=(zeros(nv * self.nc, torch.float64), zeros(nv * self.nc, torch.float64)))
        if dt == torch.float32:
            n0 = self.levels[0]['n']
            extra['f32'] = (zeros(nv * n0), zeros(nv * n0))
        cache[nv] = (bufs, extra)
        return cache[nv]

    def _cheb_mv(self, L, B, nv, b, x, x_zero=False):
        """_cheb on nv vectors at once (pf3d_bsr_fused_mv)."""
        lmax = 1.1 * L['lam']
        lmin = self.cheb_lower * L['lam']
        theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
        sigma = theta / delta
        rho_old = 1. / sigma
        A, D, n = L['A'], L['Dinv'].val, L['n']
        r, cur, nxt = B['r'], B['d'], B['t']
        _fused_mv(A, nv, n, None if x_zero else x, n, b, r, dinv=D, dnew=cur, sa=1., sb=-1.,
                  c1=1. / theta)
        for k in range(self.cheb_degree):
            if k == self.cheb_degree - 1:
                _axpby(1., cur, 1., x)
                break
            rho = 1. / (2. * sigma - rho_old)
            _fused_mv(A, nv, n, cur, n, None, r, dinv=D, dold=cur, dnew=nxt, xacc=x, sa=1.,
                      sb=-1., c1=2. * rho / delta, c2=rho * rho_old)
            cur, nxt = nxt, cur
            rho_old = rho

    def _vcycle_mv(self, nv, bufs, extra, b, x, lvl=0):
        if lvl == len(self.levels):
            nc = self.nc
            if b.dtype == torch.float64:
                b64, x64 = b, x
            else:
                b64, x64 = extra['c64']
                _lib.call('pf3d_to_f64', nv * nc, b.data_ptr(), b64.data_ptr(), _s())
            for v in range(nv):
                _lib.call('pf3d_gemv_t', nc, nc, self.coarse_inv.data_ptr(), nc,
                          b64[v * nc:].data_ptr(), x64[v * nc:].data_ptr(), 1.0, 0.0, _s())
            if b.dtype != torch.float64:
                _lib.call('pf3d_to_f32', nv * nc, x64.data_ptr(), x.data_ptr(), _s())
            return x
        L, B = self.levels[lvl], bufs[lvl]
        n = L['n']
        mc = self.levels[lvl + 1]['n'] if lvl + 1 < len(self.levels) else self.nc
        _axpby(0., b, 0., x)
        self._cheb_mv(L, B, nv, b, x, x_zero=True)
        _fused_mv(L['A'], nv, n, x, n, b, B['r'], sa=1., sb=-1.)           # r = b - A x
        _fused_mv(L['R'], nv, n, B['r'], mc, None, B['bc'], sa=0., sb=1.)  # bc = R r
        self._vcycle_mv(nv, bufs, extra, B['bc'], B['xc'], lvl + 1)
        _fused_mv(L['P'], nv, mc, B['xc'], n, None, x, sa=1., sb=1.)       # x += P xc
        self._cheb_mv(L, B, nv, b, x)
        return x

    def apply_block(self, R, Z):
        """Z[v] = M^-1 R[v] for the rows of R (nv x n, FP64, contiguous):
        one V-cycle for all of them, each matrix block read once per pass
        (pf3d_bsr_fused_mv).  Z must be a contiguous nv x n FP64 tensor."""
        nv = R.shape[0]
        if not self.levels or nv == 1:
            for v in range(nv):
                self.apply(R[v], Z[v])
            return Z
        if nv > 8:
            for v0 in range(0, nv, 8):
                self.apply_block(R[v0:v0 + 8], Z[v0:v0 + 8])
            return Z
        bufs, extra = self._mv_buffers(nv)
        rf, zf = R.reshape(-1), Z.reshape(-1)
        if self.precision == 'fp64':
            self._vcycle_mv(nv, bufs, extra, rf, zf)
        else:
            r32, z32 = extra['f32']
            _lib.call('pf3d_to_f32', rf.numel(), rf.data_ptr(), r32.data_ptr(), _s())
            self._vcycle_mv(nv, bufs, extra, r32, z32)
            _lib.call('pf3d_to_f64', zf.numel(), z32.data_ptr(), zf.data_ptr(), _s())
        return Z

    def vcycle(self, b, x, lvl=0):
        """x = M^-1 b (x is overwritten)."""
        if lvl == len(self.levels):
            if b.dtype == torch.float64:
                _lib.call('pf3d_gemv_t', self.nc, self.nc, self.coarse_inv.data_ptr(), self.nc,
                          b.data_ptr(), x.data_ptr(), 1.0, 0.0, _s())
            else:
                b64, x64 = self.c64
                _lib.call('pf3d_to_f64', self.nc, b.data_ptr(), b64.data_ptr(), _s())
                _lib.call('pf3d_gemv_t', self.nc, self.nc, self.coarse_inv.data_ptr(), self.nc,
                          b64.data_ptr(), x64.data_ptr(), 1.0, 0.0, _s())
                _lib.call('pf3d_to_f32', self.nc, x64.data_ptr(), x.data_ptr(), _s())
            return x
        L = self.levels[lvl]
        _axpby(0., b, 0., x)                       # x = 0
        self._cheb(L, b, x, x_zero=True)
        r = L['r']
        _fused(L['A'], x, b, r, sa=1., sb=-1.)     # r = b - A x
        bc, xc = L['bc'], L['xc']
        L['R'](r, bc)
        self.vcycle(bc, xc, lvl + 1)
        _fused(L['P'], xc, None, x, sa=1., sb=1.)  # x += P xc
        self._cheb(L, b, x)
        return x


class _capture_guard:
    """Context for a graph capture: the device synchronised first, and
    Python's cyclic collector paused while the stream captures.  A CUDAGraph
    that became garbage inside a reference cycle would otherwise be
    destroyed by the collector at some allocation made during the capture,
    which HIP refuses while a stream captures; paused, such garbage is
    collected after the capture instead (and without the cost of a full
    collection before every capture)."""

    def __enter__(self):
        import gc
        torch.cuda.synchronize()
        self._was = gc.isenabled()
        gc.disable()
        return self

    def __exit__(self, *exc):
        import gc
        if self._was:
            gc.enable()
        return False


def _ptr(t):
    return None if t is None else t.data_ptr()


def _fused_mv(M, nv, ldx, xin, ld, b, r, dinv=None, dold=None, dnew=None, xacc=None, sa=1.,
              sb=1., c1=0., c2=0.):
    """pf3d_bsr_fused_mv(_f32) with the operator M (a _Dev) on nv vectors:
    xin's vectors ldx apart, the other operands' ld apart."""
    fn = 'pf3d_bsr_fused_mv' if M.fn == 'pf3d_bsr_spmv' else 'pf3d_bsr_fused_mv_f32'
    _lib.call(fn, M.nb, M.row_ptr.data_ptr(), M.col_idx.data_ptr(), M.val.data_ptr(), nv, ldx,
              _ptr(xin), ld, _ptr(b), r.data_ptr(), _ptr(dinv), _ptr(dold), _ptr(dnew),
              _ptr(xacc), float(sa), float(sb), float(c1), float(c2), _s())


def _fused(M, xin, b, r, dinv=None, dold=None, dnew=None, xacc=None, sa=1., sb=1., c1=0.,
           c2=0.):
    """One pf3d_bsr_fused(_f32) pass with the operator M (a _Dev):
    r = sa (b or r) + sb M xin; dnew = c1 Dinv r + c2 dold; xacc += dold."""
    fn = 'pf3d_bsr_fused' if M.fn == 'pf3d_bsr_spmv' else 'pf3d_bsr_fused_f32'
    _lib.call(fn, M.nb, M.row_ptr.data_ptr(), M.col_idx.data_ptr(), M.val.data_ptr(), _ptr(xin),
              _ptr(b), r.data_ptr(), _ptr(dinv), _ptr(dold), _ptr(dnew), _ptr(xacc), float(sa),
              float(sb), float(c1), float(c2), _s())


def _dot(a, b, out):
    """a . b through pf3d_dot; out holds the result (element 0) and the
    1024-double workspace after it."""
    _lib.call('pf3d_dot', a.numel(), a.data_ptr(), b.data_ptr(), out.data_ptr(),
              out[1:].data_ptr(), _s())
    return float(out[0].item())


class _CGIteration:
    """Buffers of one preconditioned-CG solve and one iteration captured in
    a HIP graph: q = A p, p.q, x and r updated, z = V-cycle(r), r.z (and
    r.z_old), r.r, beta, p updated -- scalars on the device
    (pf3d_cg_update_xr, pf3d_cg_beta, pf3d_cg_update_p), so replays need no
    host round trip.  Kept on the system per preconditioner and reused by
    later solves."""

    def __init__(self, system, amg):
        n = 6 * system.nb
        dev = system.val.device
        zeros = lambda m: torch.zeros(m, dtype=torch.float64, device=dev)
        # the system caches this object: a weak reference back keeps the
        # pair out of a reference cycle, so the captured graph is freed with
        # the system
        self._system = weakref.ref(system)
        self.amg, self.n = amg, n
        self.flexible = amg.precision != 'fp64'
        self.x, self.r, self.p, self.q, self.z = (zeros(n) for _ in range(5))
        self.zold = zeros(n) if self.flexible else None
        self.s = zeros(8)           # r.z, p.q, -, r.z new, beta, r.r, r.z_old
        self.work = zeros(1024)
        self.graph = None

    @property
    def system(self):
        return self._system()

    def _dot_into(self, a, b, k):
        _lib.call('pf3d_dot', self.n, a.data_ptr(), b.data_ptr(), self.s[k:].data_ptr(),
                  self.work.data_ptr(), _s())

    def step(self):
        """One iteration's launches (eager, or recorded by capture())."""
        self.system.matvec(self.p, self.q)
        self._dot_into(self.p, self.q, 1)
        _lib.call('pf3d_cg_update_xr', self.n, self.s.data_ptr(), self.p.data_ptr(),
                  self.q.data_ptr(), self.x.data_ptr(), self.r.data_ptr(), _s())
        if self.flexible:
            _axpby(1., self.z, 0., self.zold)
        self.amg._apply(self.r, self.z)
        self._dot_into(self.r, self.z, 3)
        if self.flexible:
            # beta = r.(z_new - z_old) / r_old.z_old keeps CG's rate with a
            # preconditioner that is linear only up to FP32 rounding
            self._dot_into(self.r, self.zold, 6)
        self._dot_into(self.r, self.r, 5)
        _lib.call('pf3d_cg_beta', self.s.data_ptr(), int(self.flexible), _s())
        _lib.call('pf3d_cg_update_p', self.n, self.s.data_ptr(), self.z.data_ptr(),
                  self.p.data_ptr(), _s())

    def capture(self):
        with _capture_guard():
            s = torch.cuda.Stream(device=self.x.device)
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                self.step()                 # warm-up outside the graph (zero buffers: no-op)
            torch.cuda.current_stream().wait_stream(s)
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=s):
                self.step()
            torch.cuda.synchronize()
        self.graph = g

    def restart(self):
        """p = z = V-cycle(r), s[0] = r.z (r holds the true residual)."""
        self.amg._apply(self.r, self.z)
        _axpby(1., self.z, 0., self.p)
        self._dot_into(self.r, self.z, 0)


def amg_pcg(system, amg, f, rtol=1e-8, maxiter=2000, x0=None, log=None, stall=400, check=None,
            graph=True):
    """PCG with one V-cycle per iteration on a StaticSystem's matrix.
    Returns (u, info) like StaticSystem.solve (compensated true residual).
    The iteration runs as one captured graph (graph=True) replayed `check`
    times between reads of the recursive residual (default: 8 below 1e6
    DOF, where a host round trip per iteration would cost more than the
    iteration's kernels, else 1); a run whose recursive residual has not
    halved in `stall` iterations restarts from the true residual (and ends
    when a restart does not gain 10 %).  The window is long because CG's
    residual norm is not monotone: on plates with thickness fields it rises
    up to ~10x and takes a few hundred iterations to come back (a 100-
    iteration window stopped 9 of 16 such plates, which then converge in
    240-760 iterations)."""
    b = system.rhs(f)
    if check is None:
        check = 8 if 6 * system.nb < 1000000 else 1
    cache = system.__dict__.setdefault('_cg_iterations', {})
    cg = cache.get(id(amg))
    if cg is None or cg.amg is not amg:
        cg = cache[id(amg)] = _CGIteration(system, amg)
    if graph and cg.graph is None:
        cg.capture()
    if x0 is None:
        cg.x.zero_()
    else:
        _axpby(1., x0, 0., cg.x)
    s = torch.empty(1025, dtype=torch.float64, device=b.device)
    bb = _dot(b, b, s) or 1.
    it = 0
    restarts = 0
    best = float('inf')
    while True:
        # (re)start from the true residual (compensated b - A x: plain FP64
        # row sums cancel to ~eps |A| |x|, above 1e-8 |b| on thin shells);
        # the recursive residual drifts on these ill-conditioned systems,
        # and each restart gains what the plain products of the inner CG
        # allow.  Stop when a restart gains less than 10 %.
        system.residual(b, cg.x, cg.r)
        rel = (_dot(cg.r, cg.r, s) / bb) ** 0.5
        if rel <= rtol or it >= maxiter or restarts > 12 or rel > 0.9 * best:
            break
        best = rel
        restarts += 1
        if log:
            log('amg pcg: restart %d at %d iterations, true relres %.3e' % (restarts, it, rel))
        cg.restart()
        low, since = rel, 0
        while it < maxiter:
            k = min(check, maxiter - it)
            for _ in range(k):
                if graph:
                    cg.graph.replay()
                else:
                    cg.step()
            it += k
            rr = (float(cg.s[5].item()) / bb) ** 0.5
            if rr <= rtol:
                break
            if rr <= 0.5 * low:
                low, since = rr, 0
            else:
                since += k
                if since >= stall:
                    break
    return cg.x.clone(), dict(iters=it, outer=restarts, relres=rel)


def residual_floor(system, x):
    """eps * || |A| |x| || / ||b||-style estimate of the smallest relative
    residual b - A x that FP64 arithmetic can show for the solution x (the
    rounding of the row sums of A x), with |b| taken as ||P f|| = ||A x||."""
    dev = x.device
    absval = system.val.abs()
    ax = x.abs()
    y = torch.empty_like(x)
    _lib.call('pf3d_bsr_spmv', system.nb, system.plan.row_ptr.data_ptr(),
              system.plan.col_idx.data_ptr(), absval.data_ptr(), ax.data_ptr(), y.data_ptr(), _s())
    s = torch.empty(1025, dtype=torch.float64, device=dev)
    big = _dot(y, y, s) ** 0.5
    system.matvec(x, y)
    return 2. ** -53 * big / (_dot(y, y, s) ** 0.5 or 1.)

"""The reference's CPU pipeline, restated with the oracle + scipy, for
comparisons in tests: per-element assembly -> coo_matrix.tocsc() ->
KC0[bu, :][:, bu] -> spsolve (tests/test_quad4_static_point_load.py:52-102)."""
import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.linalg import spsolve

from oracle import oracle as O


def assemble_cpu(etype, x, conn, props, pid=None, xmat=None, mat='KC0', **kw):
    E = conn.shape[0]
    size = O.slot_count(etype, mat, kw.get('mtype', 0))
    r = np.zeros(size * E, np.int64); c = np.zeros_like(r); v = np.zeros(size * E)
    O.assemble(etype, mat, x, conn, size * np.arange(E), np.zeros((E, 3)) if xmat is None else xmat,
               np.zeros(E, np.int64) if pid is None else pid, props, r, c, v, **kw)
    return r, c, v


def solve_cpu(r, c, v, n_dof, fixed, f):
    K = coo_matrix((v, (r, c)), shape=(n_dof, n_dof)).tocsc()
    bu = ~fixed
    u = np.zeros(n_dof)
    u[bu] = spsolve(K[bu, :][:, bu], f[bu])
    return u, K


def solve_cpu_refined(K, fixed, f, steps=6):
    """The exact solution of the sliced system to ~1e-14: scipy's sparse LU
    (spsolve's factorisation) with iterative refinement, the residuals
    f - K u accumulated in extended precision (np.longdouble, 64-bit
    significand) on u held as a pair hi + lo.  On thin laminate plates the
    drilling rotations rz are a near null space, so a single FP64 spsolve
    gets them only to ~1e-2 (two fill orderings can differ by ~1); this
    refined solution is the same to ~1e-14 from either start.  Returns
    (u, relres) with relres the extended-precision relative residual."""
    import scipy.sparse.linalg as sla
    bu = ~fixed
    Kuu = K.tocsr()[bu][:, bu].tocsc()
    fu = f[bu]
    lu = sla.splu(Kuu)
    Kr = Kuu.tocsr()
    rows = np.repeat(np.arange(Kr.shape[0]), np.diff(Kr.indptr))
    data = Kr.data.astype(np.longdouble)
    fl = fu.astype(np.longdouble)

    def resid(x):
        y = np.zeros(Kr.shape[0], np.longdouble)
        np.add.at(y, rows, data * x[Kr.indices])
        return fl - y

    x = lu.solve(fu).astype(np.longdouble)
    for _ in range(steps):
        x = x + lu.solve(resid(x).astype(np.float64))
    rel = float(np.linalg.norm(resid(x).astype(np.float64)) / np.linalg.norm(fu))
    u = np.zeros(f.size)
    u[bu] = x.astype(np.float64)
    return u, rel

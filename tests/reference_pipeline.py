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

"""Host (numpy/scipy) construction of the smoothed-aggregation hierarchy:
the checker for the GPU setup in pyfe3d_amd.amg.  Same steps, in scalar
CSR products: block-diagonal inverses, power iteration for
lambda_max(D^-1 A) from the same seeded start, geometric aggregates, the
tentative prolongator of the rigid-body modes, P = (I - 4/3 / lambda
D^-1 A) P0, A_c = P^T A P (+1 on coarse DOFs without a mode), symmetrised.
"""
import numpy as np
import scipy.sparse as ss

from pyfe3d_amd.amg import box_aggregates, rigid_body_modes, tentative


def block_diag_inv(A):
    n = A.shape[0] // 6
    A = A.tobsr(blocksize=(6, 6))
    A.sort_indices()
    D = np.zeros((n, 6, 6))
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    diag = rows == A.indices
    D[rows[diag]] = A.data[diag]
    return np.linalg.inv(D)


def lambda_max(A, Dinv, iters=20, seed=0):
    x = np.random.default_rng(seed).uniform(-1, 1, A.shape[0])
    lam = 1.
    for _ in range(iters):
        y = A @ x
        y = np.einsum('nij,nj->ni', Dinv, y.reshape(-1, 6)).ravel()
        lam = np.linalg.norm(y) / np.linalg.norm(x)
        x = y / np.linalg.norm(y)
    return lam


def setup_levels(A, fixed, xyz, box=3.0, coarse_nodes=300, max_levels=8):
    """List of dict(A, P, lam) (scipy CSR) and the coarsest matrix (dense)."""
    A = A.tocsr()
    n = A.shape[0] // 6
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    B = rigid_body_modes(xyz, fixed)
    Ab = A.tobsr(blocksize=(6, 6))
    rows = np.repeat(np.arange(n), np.diff(Ab.indptr))
    off = rows != Ab.indices
    size = box * np.median(np.linalg.norm(xyz[rows[off]] - xyz[Ab.indices[off]], axis=1))
    levels = []
    while A.shape[0] // 6 > coarse_nodes and len(levels) < max_levels - 1:
        nb = A.shape[0] // 6
        Dinv = block_diag_inv(A)
        lam = lambda_max(A, Dinv)
        agg, n_agg = box_aggregates(xyz, size)
        if n_agg >= nb:
            break
        Q, Bc, dead = tentative(agg, n_agg, B)
        P0 = ss.bsr_matrix((Q, agg, np.arange(nb + 1)), shape=(6 * nb, 6 * n_agg)).tocsr()
        Dm = ss.bsr_matrix((Dinv, np.arange(nb), np.arange(nb + 1)), shape=A.shape).tocsr()
        P = (P0 - (4. / 3. / lam) * (Dm @ (A @ P0))).tocsr()
        Ac = (P.T @ (A @ P)).tocsr()
        Ac = Ac + ss.diags(dead.astype(np.float64))
        Ac = (0.5 * (Ac + Ac.T)).tocsr()
        levels.append(dict(A=A, P=P, lam=lam))
        cen = np.zeros((n_agg, 3))
        np.add.at(cen, agg, xyz)
        xyz = cen / np.bincount(agg, minlength=n_agg)[:, None]
        A, B = Ac, Bc
        size *= box
    return levels, A.toarray()

"""The host multigrid construction (the checker of the GPU setup) on an
oracle-assembled plate: the tentative prolongator reproduces the rigid-body
modes with orthonormal columns per aggregate, the smoothed prolongator keeps
them in the interior (K annihilates rigid-body motion), and every coarse
operator is symmetric positive definite with the fine energy of its
prolongated vectors."""
import numpy as np
import scipy.sparse as ss

import amg_host
from oracle import oracle as O
from pyfe3d_amd.amg import box_aggregates, rigid_body_modes, tentative
from pyfe3d_amd.mesh import PlateMesh
from pyfe3d_amd.shellprop import isotropic_plate


def _plate():
    m = PlateMesh(25, 19, 1.0, 0.7, jitter=0.2, seed=4)
    rec = isotropic_plate(thickness=0.004, E=70e9, nu=0.33, calc_scf=True).record()
    E = m.quads.shape[0]
    r = np.zeros(576 * E, np.int64); c = np.zeros_like(r); v = np.zeros(576 * E)
    O.assemble('quad4', 'KC0', m.x, m.quads, 576 * np.arange(E), np.zeros((E, 3)),
               np.zeros(E, np.int64), rec[None], r, c, v)
    K = ss.coo_matrix((v, (r, c)), shape=(m.n_dof, m.n_dof)).tocsr()
    fixed = m.clamped()
    keep = ss.diags((~fixed).astype(float))
    A = (keep @ K @ keep + ss.diags(fixed.astype(float))).tocsr()
    return m, K, A, fixed


def test_tentative_reproduces_modes():
    m, K, A, fixed = _plate()
    B = rigid_body_modes(m.ncoords, fixed)
    agg, n_agg = box_aggregates(m.ncoords, 3 * 1.0 / 24)
    Q, Bc, dead = tentative(agg, n_agg, B)
    P0 = ss.bsr_matrix((Q, agg, np.arange(m.n_nodes + 1)), shape=(m.n_dof, 6 * n_agg))
    assert np.abs(P0 @ Bc - B).max() <= 1e-12 * np.abs(B).max()
    G = (P0.T @ P0).toarray()
    live = ~dead
    np.testing.assert_allclose(G[np.ix_(live, live)], np.eye(live.sum()), atol=1e-12)
    assert np.abs(G[dead]).max(initial=0.) == 0.
    # K annihilates the rigid-body modes of the free plate (the oracle's
    # drilling terms included)
    Bf = rigid_body_modes(m.ncoords, np.zeros(m.n_dof, bool))
    assert np.abs(K @ Bf).max() <= 1e-7 * abs(K).max() * np.abs(Bf).max()


def test_hierarchy_is_galerkin_and_spd():
    m, K, A, fixed = _plate()
    levels, Ad = amg_host.setup_levels(A, fixed, m.ncoords, coarse_nodes=20)
    assert len(levels) >= 2
    rng = np.random.default_rng(0)
    for i, L in enumerate(levels):
        P, Af = L['P'], L['A']
        Ac = levels[i + 1]['A'] if i + 1 < len(levels) else ss.csr_matrix(Ad)
        assert abs(Ac - Ac.T).max() <= 1e-14 * abs(Ac).max()
        x = rng.uniform(-1, 1, Ac.shape[0])
        Px = P @ x
        dead = np.isclose(Ac.diagonal(), 1.) & (abs(Ac).sum(axis=1).A.ravel() == 1.)
        x[dead] = 0.
        Px = P @ x
        assert abs(x @ (Ac @ x) - Px @ (Af @ Px)) <= 1e-10 * abs(Px @ (Af @ Px))
        assert L['lam'] > 1.
    w = np.linalg.eigvalsh(Ad)
    assert w.min() > 0.

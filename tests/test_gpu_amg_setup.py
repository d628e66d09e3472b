"""Multigrid setup on the GPU against scipy: the 6x6-block sparse product,
transpose, prolongator smoothing and symmetrisation kernels on random
matrices (structure exact, values to rounding), and the whole hierarchy
against the host construction in tests/amg_host.py (same aggregates,
lambda_max, prolongators and coarse operators to 1e-10)."""
import numpy as np
import pytest
import scipy.sparse as ss
import torch

from test_gpu_amg import _system

pytestmark = pytest.mark.gpu
DEV = 'cuda'


def _random_bsr(nr, nc, density, rng, empty_rows=()):
    M = ss.random(nr, nc, density=density, random_state=rng, format='csr')
    M = M.tolil()
    for i in empty_rows:
        M.rows[i] = []
        M.data[i] = []
    M = M.tocsr()
    M.sort_indices()
    data = rng.uniform(-1, 1, (M.nnz, 6, 6))
    return ss.bsr_matrix((data, M.indices, M.indptr), shape=(6 * nr, 6 * nc))


def _dev(M):
    from pyfe3d_amd.amg import _Bsr
    M = M.tobsr(blocksize=(6, 6))
    M.sort_indices()
    return _Bsr(M.shape[0] // 6, M.shape[1] // 6,
                torch.as_tensor(M.indptr.astype(np.int32), device=DEV),
                torch.as_tensor(M.indices.astype(np.int32), device=DEV),
                torch.as_tensor(np.ascontiguousarray(M.data.reshape(-1)), device=DEV))


def _check_same(G, ref, tol):
    """G: _Bsr from the GPU; ref: scipy matrix with the same block pattern
    (explicit zero blocks allowed on the GPU side)."""
    Gs = G.to_scipy()
    assert Gs.shape == ref.shape
    d = abs(Gs - ref).max()
    scale = abs(ref).max() or 1.
    assert d <= tol * scale, (d, scale)
    rb = ref.tobsr(blocksize=(6, 6))
    rb.sort_indices()
    return Gs, rb


def test_spgemm_matches_scipy():
    """Random X (137 x 91 blocks, empty rows) times Y (91 x 203): the
    product's block pattern equals scipy's, values to 1e-13."""
    rng = np.random.default_rng(3)
    from pyfe3d_amd.amg import spgemm
    X = _random_bsr(137, 91, 0.05, rng, empty_rows=(0, 5, 136))
    Y = _random_bsr(91, 203, 0.03, rng, empty_rows=(7,))
    C = spgemm(_dev(X), _dev(Y))
    ref = (X @ Y).tobsr(blocksize=(6, 6))
    ref.sort_indices()
    pattern = ss.csr_matrix((np.ones(X.indices.size), X.indices, X.indptr)) @ \
        ss.csr_matrix((np.ones(Y.indices.size), Y.indices, Y.indptr))
    pattern.sort_indices()
    assert (C.row_ptr.cpu().numpy() == pattern.indptr).all()
    assert (C.col_idx.cpu().numpy() == pattern.indices).all()
    _check_same(C, X @ Y, 1e-13)


def test_transpose_matches_scipy():
    rng = np.random.default_rng(4)
    from pyfe3d_amd.amg import transpose
    X = _random_bsr(150, 61, 0.04, rng, empty_rows=(2, 149))
    T = transpose(_dev(X))
    ref = X.T.tobsr(blocksize=(6, 6))
    ref.sort_indices()
    assert (T.row_ptr.cpu().numpy() == ref.indptr).all()
    assert (T.col_idx.cpu().numpy() == ref.indices).all()
    assert np.array_equal(T.val.cpu().numpy().reshape(-1, 6, 6), ref.data)


def test_symmetrize_and_smooth_kernels():
    """(V + V^T)/2 plus identity on flagged DOFs; P0 - w D^-1 T with P0's
    block at column agg[i] of row i."""
    from pyfe3d_amd import _lib
    from pyfe3d_amd.amg import spgemm, _Bsr
    rng = np.random.default_rng(5)
    S = _random_bsr(80, 80, 0.06, rng)
    S = (S + S.T + ss.identity(480)).tobsr(blocksize=(6, 6))   # symmetric pattern, diagonal
    S.sort_indices()
    G = _dev(S)
    dead = (rng.uniform(size=480) < 0.1).astype(np.uint8)
    out = torch.empty_like(G.val)
    flag = torch.zeros(1, dtype=torch.int32, device=DEV)
    dead_d = torch.as_tensor(dead, device=DEV)     # kept alive until the kernel has run
    _lib.call('pf3d_bsr_symmetrize', G.nr, G.nnzb, G.row_ptr.data_ptr(), G.col_idx.data_ptr(),
              G.val.data_ptr(), dead_d.data_ptr(), out.data_ptr(), flag.data_ptr(),
              torch.cuda.current_stream().cuda_stream)
    ref = 0.5 * (S + S.T) + ss.diags(dead.astype(float))
    assert int(flag.item()) == 0
    _check_same(_Bsr(G.nr, G.nc, G.row_ptr, G.col_idx, out), ref, 1e-15)
    # smoothing: A (60 x 60 blocks, diagonal present), 17 aggregates
    A = (_random_bsr(60, 60, 0.08, rng) + ss.identity(360) * 4.).tobsr(blocksize=(6, 6))
    A.sort_indices()
    agg = np.sort(rng.integers(0, 17, 60))
    agg[:17] = np.arange(17)
    agg.sort()
    Q = rng.uniform(-1, 1, (60, 6, 6))
    Dinv = np.linalg.inv(np.stack([A.toarray()[6 * i:6 * i + 6, 6 * i:6 * i + 6]
                                   for i in range(60)]))
    P0 = ss.bsr_matrix((Q, agg, np.arange(61)), shape=(360, 102))
    Ad = _dev(A)
    P0d = _dev(P0)
    T = spgemm(Ad, P0d)
    pv = torch.empty_like(T.val)
    w = 0.37
    agg_d = torch.as_tensor(agg.astype(np.int32), device=DEV)
    q_d = torch.as_tensor(Q.reshape(-1), device=DEV)
    dinv_d = torch.as_tensor(Dinv.reshape(-1), device=DEV)
    _lib.call('pf3d_amg_smooth_prolongator', 60, T.row_ptr.data_ptr(), T.col_idx.data_ptr(),
              T.val.data_ptr(), agg_d.data_ptr(), q_d.data_ptr(), dinv_d.data_ptr(), w,
              pv.data_ptr(), flag.data_ptr(), torch.cuda.current_stream().cuda_stream)
    Dm = ss.bsr_matrix((Dinv, np.arange(60), np.arange(61)), shape=(360, 360))
    ref = P0 - w * (Dm @ (A @ P0))
    assert int(flag.item()) == 0
    _check_same(_Bsr(60, 17, T.row_ptr, T.col_idx, pv), ref, 1e-13)


@pytest.mark.parametrize('case', ['plate', 'panel'])
def test_hierarchy_matches_host(case):
    """The GPU hierarchy equals the host (scipy) construction level by
    level: lambda_max to 1e-10, P and the coarse operators to 1e-10 of
    their largest entry."""
    import amg_host
    from pyfe3d_amd.amg import setup_hierarchy
    if case == 'plate':
        from pyfe3d_amd.mesh import PlateMesh
        from pyfe3d_amd.shellprop import isotropic_plate
        m = PlateMesh(41, 31, 1.0, 0.7, jitter=0.2, seed=4)
        rec = isotropic_plate(thickness=0.004, E=70e9, nu=0.33, calc_scf=True).record()
        fixed = m.clamped()
        sysm = _system(m, m.quads, 'quad4', rec, fixed)
        xyz = m.ncoords
    else:
        from pyfe3d_amd.solver import StaticSystem
        from pyfe3d_amd.sparse import BSRPlan
        from test_gpu_panel import _mixed, _panel
        p = _panel(n=48)
        fixed = p.fixed()
        _, r, c, v = _mixed(p, 'KC0')
        plan = BSRPlan(r, c, p.n_nodes)
        sysm = StaticSystem(plan, plan.apply(v), fixed)
        xyz = p.mesh.ncoords
    levels, Ad = setup_hierarchy(sysm.plan.row_ptr, sysm.plan.col_idx, sysm.val, sysm.nb, fixed,
                                 xyz, coarse_nodes=40)
    A = sysm.plan.to_scipy(sysm.val)
    hl, hAd = amg_host.setup_levels(A, fixed, xyz, coarse_nodes=40)
    assert len(levels) == len(hl) >= 2
    for L, H in zip(levels, hl):
        assert abs(L['lam'] - H['lam']) <= 1e-10 * H['lam']
        P = L['P'].to_scipy()
        assert abs(P - H['P']).max() <= 1e-10 * abs(H['P']).max()
        R = L['R'].to_scipy()
        assert abs(R - H['P'].T).max() <= 1e-10 * abs(H['P']).max()
        print(case, 'level', P.shape, 'lambda', L['lam'])
    assert np.abs(Ad - hAd).max() <= 1e-10 * np.abs(hAd).max()
    for L, H in zip(levels[1:], hl[1:]):
        Ac = L['A'].to_scipy()
        assert abs(Ac - H['A']).max() <= 1e-10 * abs(H['A']).max()

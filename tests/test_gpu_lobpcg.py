"""Lowest natural frequencies by multigrid-preconditioned LOBPCG on the GPU
against scipy's eigsh (shift-invert, as the reference tests call it,
tests/test_quad4_natural_freq.py:120-123) on the oracle's matrices: a
plate and the mixed-element stiffened panel."""
import numpy as np
import pytest
import scipy.sparse as ss
import scipy.sparse.linalg as sla
import torch

from test_gpu_eigen import _bsr, _cpu, _gpu_coo, _plate
from test_gpu_panel import _coo, _mixed, _oracle

pytestmark = pytest.mark.gpu


def _lobpcg(plan, K, M, fixed, xyz, k):
    from pyfe3d_amd.lobpcg import LOBPCG
    from pyfe3d_amd.solver import StaticSystem
    sysm = StaticSystem(plan, K, fixed)
    amg = sysm.setup_amg(xyz, coarse_nodes=80)
    return LOBPCG(plan, K, M, fixed, amg).solve(k=k, tol=1e-7, maxiter=400)


def test_lobpcg_plate_vs_eigsh():
    from pyfe3d_amd.shellprop import isotropic_plate
    a, b, E, nu, rho, h = 0.3, 0.5, 203.e9, 0.33, 7.83e3, 0.01
    m = _plate(17, 21, a, b)
    rec = isotropic_plate(thickness=h, E=E, nu=nu, calc_scf=True, rho=rho).record()
    x, y = m.ncoords[:, 0], m.ncoords[:, 1]
    edge = np.isclose(x, 0) | np.isclose(x, a) | np.isclose(y, 0) | np.isclose(y, b)
    fixed = np.zeros(m.n_dof, bool)
    for d in range(3):
        fixed[d::6] = edge
    rK, cK, vK = _gpu_coo('KC0', m, rec)
    rM, cM, vM = _gpu_coo('M', m, rec, 0)
    planK, K = _bsr(rK, cK, vK, m.n_nodes)
    _, M = _bsr(rM, cM, vM, m.n_nodes, planK)
    lam, X, info = _lobpcg(planK, K, M, fixed, m.ncoords, 6)
    bu = ~fixed
    Kc = _cpu('KC0', m, rec)[bu, :][:, bu]
    Mc = _cpu('M', m, rec, 0)[bu, :][:, bu]
    ref = np.sort(sla.eigsh(A=Kc, M=Mc, sigma=-1., which='LM', k=6, tol=1e-12)[0])
    print('lobpcg', lam, ref, info['iters'], info['residual'])
    np.testing.assert_allclose(lam, ref, rtol=1e-8)
    Xh = X.cpu().numpy()
    for i in range(3):
        xi = Xh[bu, i]
        r = Kc @ xi - lam[i] * (Mc @ xi)
        assert np.linalg.norm(r) <= 1e-6 * np.linalg.norm(Kc @ xi)


def test_lobpcg_stiffened_panel_vs_eigsh():
    """Quad4 skin, Tria3R cut-outs, BeamC stringers: K and M each from one
    mixed-type launch; the six lowest frequencies against eigsh."""
    from pyfe3d_amd.sparse import BSRPlan
    from pyfe3d_amd.panel import StiffenedPanel
    p = StiffenedPanel(n=32, a=0.6, b=0.5, spacing=16, n_cutouts=3, cutout=(3, 9), seed=9)
    fixed = p.fixed()
    _, r, c, v = _mixed(p, 'KC0')
    _, rm, cm, vm = _mixed(p, 'M', mtype=0)
    plan = BSRPlan(r, c, p.n_nodes)
    K = plan.apply(v)
    union = BSRPlan(torch.cat([r, rm]), torch.cat([c, cm]), p.n_nodes)
    assert torch.equal(union.row_ptr, plan.row_ptr) and torch.equal(union.col_idx, plan.col_idx)
    M = union.apply(torch.cat([torch.zeros_like(v), vm]))
    lam, X, info = _lobpcg(plan, K, M, fixed, p.mesh.ncoords, 6)
    bu = ~fixed
    parts, _ = _oracle(p, 'KC0')
    Kc = _coo(parts, p.n_dof)[bu, :][:, bu]
    partsM, _ = _oracle(p, 'M', mtype=0)
    Mc = _coo(partsM, p.n_dof)[bu, :][:, bu]
    ref = np.sort(sla.eigsh(A=Kc, M=Mc, sigma=-1., which='LM', k=6, tol=1e-12)[0])
    print('panel lobpcg', lam, ref, info['iters'], info['residual'])
    np.testing.assert_allclose(lam, ref, rtol=1e-8)


def test_lobpcg_buckling_stiffened_panel_vs_eigsh():
    """Linear buckling of the stiffened panel under end compression and
    pressure: KG from the static solution in one mixed-type launch, the
    three lowest load multipliers K x = lambda (-KG) x against eigsh on the
    oracle's matrices (as the reference's buckling tests call it)."""
    from pyfe3d_amd.lobpcg import LOBPCG
    from pyfe3d_amd.panel import StiffenedPanel
    from pyfe3d_amd.solver import StaticSystem
    from pyfe3d_amd.sparse import BSRPlan
    p = StiffenedPanel(n=32, a=0.6, b=0.5, spacing=16, n_cutouts=3, cutout=(3, 9), seed=9)
    fixed = p.fixed()
    f = p.loads(Nxx=1e3, p=2e2)
    _, r, c, v = _mixed(p, 'KC0')
    plan = BSRPlan(r, c, p.n_nodes)
    K = plan.apply(v)
    sysm = StaticSystem(plan, K, fixed)
    amg = sysm.setup_amg(p.mesh.ncoords, coarse_nodes=80)
    u, info = sysm.solve(f, method='direct', rtol=1e-10)
    uh = u.cpu().numpy()
    _, rg, cg, vg = _mixed(p, 'KG', u=uh)
    union = BSRPlan(torch.cat([r, rg]), torch.cat([c, cg]), p.n_nodes)
    assert torch.equal(union.row_ptr, plan.row_ptr) and torch.equal(union.col_idx, plan.col_idx)
    KG = union.apply(torch.cat([torch.zeros_like(v), vg]))
    lam, X, info = LOBPCG(plan, K, KG, fixed, amg, problem='buckling').solve(k=3, tol=1e-8,
                                                                           maxiter=400)
    bu = ~fixed
    parts, _ = _oracle(p, 'KC0')
    Kc = _coo(parts, p.n_dof)[bu, :][:, bu]
    partsG, _ = _oracle(p, 'KG', u=uh)
    KGc = _coo(partsG, p.n_dof)[bu, :][:, bu]
    mu = sla.eigsh(A=KGc, M=Kc, k=6, which='SA', tol=1e-12)[0]
    ref = np.sort(-1. / mu[mu < 0])[:3]
    print('panel buckling lobpcg', lam, ref, info['iters'], info['residual'])
    np.testing.assert_allclose(lam, ref, rtol=1e-8)
    Xh = X.cpu().numpy()
    for i in range(3):
        xi = Xh[bu, i]
        res = Kc @ xi + lam[i] * (KGc @ xi)
        assert np.linalg.norm(res) <= 1e-6 * np.linalg.norm(Kc @ xi)


def test_block_product_matches_single():
    """The multi-vector compensated product (3, 8 and 11 = 8 + 3 vectors)
    gives the single-vector compensated product's bits for every vector."""
    from pyfe3d_amd import _lib
    from pyfe3d_amd.shellprop import isotropic_plate
    m = _plate(13, 11, 0.3, 0.5)
    rec = isotropic_plate(thickness=0.01, E=203e9, nu=0.33, calc_scf=True, rho=7830.).record()
    rK, cK, vK = _gpu_coo('KC0', m, rec)
    plan, K = _bsr(rK, cK, vK, m.n_nodes)
    n = m.n_dof
    g = torch.Generator(device='cpu').manual_seed(7)
    s = torch.cuda.current_stream().cuda_stream
    for nv in (3, 8, 11):
        X = (torch.rand((nv, n), generator=g, dtype=torch.float64) - 0.5).to('cuda')
        Y = torch.empty_like(X)
        for v0 in range(0, nv, 8):
            k = min(8, nv - v0)
            _lib.call('pf3d_bsr_product_dd_mv', m.n_nodes, plan.row_ptr.data_ptr(),
                      plan.col_idx.data_ptr(), K.data_ptr(), k, n, X[v0].data_ptr(),
                      Y[v0].data_ptr(), s)
        for v in range(nv):
            y = torch.empty(n, dtype=torch.float64, device='cuda')
            _lib.call('pf3d_bsr_residual_dd', m.n_nodes, plan.row_ptr.data_ptr(),
                      plan.col_idx.data_ptr(), K.data_ptr(), X[v].data_ptr(), None,
                      y.data_ptr(), s)
            assert torch.equal(Y[v], y), (nv, v)

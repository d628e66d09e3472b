"""Natural frequencies and linear buckling on the GPU (shift-invert Lanczos
over the banded Cholesky factor) against scipy's eigsh on the
oracle-assembled matrices, and against the reference tests' closed forms
(tests/test_quad4_natural_freq.py:131-137 of pyfe3d)."""
import numpy as np
import pytest
import scipy.sparse as ss
import scipy.sparse.linalg as sla
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = 'cuda'


def _plate(nx=9, ny=11, a=0.3, b=0.5):
    from pyfe3d_amd.mesh import PlateMesh
    return PlateMesh(nx, ny, a, b)


def _gpu_coo(kind, m, rec, mtype=0, stress=None):
    from pyfe3d_amd import assembly as A
    E = m.quads.shape[0]
    size = {'KC0': 576, 'M': 480 if mtype != 2 else 288, 'KGs': 144}[kind]
    r = torch.zeros(size * E, dtype=torch.int64, device=DEV)
    c = torch.zeros_like(r)
    v = torch.zeros(size * E, dtype=torch.float64, device=DEV)
    x = torch.as_tensor(m.x, device=DEV)
    conn = torch.as_tensor(6 * m.quads, device=DEV)
    props = torch.as_tensor(rec[None], device=DEV)
    if kind == 'KC0':
        A.shell_kc0('quad4', x, conn, r, c, v, props)
    elif kind == 'M':
        A.shell_m('quad4', x, conn, r, c, v, props, mtype=mtype)
    else:
        A.shell_kg_stress('quad4', x, conn, torch.as_tensor(stress, dtype=torch.float64, device=DEV), r, c, v)
    return r, c, v


def _cpu(kind, m, rec, mtype=0, stress=None):
    E = m.quads.shape[0]
    mat, size = {'KC0': ('KC0', 576), 'M': ('M', 480 if mtype != 2 else 288),
                 'KGs': ('KGs', 144)}[kind]
    r = np.zeros(size * E, np.int64); c = np.zeros_like(r); v = np.zeros(size * E)
    O.assemble('quad4', mat, m.x, m.quads, size * np.arange(E), np.zeros((E, 3)),
               np.zeros(E, np.int64), rec[None], r, c, v, mtype=mtype,
               stress=None if stress is None else np.asarray(stress))
    n = m.n_dof
    return ss.coo_matrix((v, (r, c)), shape=(n, n)).tocsc()


def _bsr(r, c, v, n_nodes, ref_plan=None):
    from pyfe3d_amd.sparse import BSRPlan
    plan = BSRPlan(r, c, n_nodes)
    if ref_plan is not None:
        assert torch.equal(plan.row_ptr, ref_plan.row_ptr)
        assert torch.equal(plan.col_idx, ref_plan.col_idx)
    return plan, plan.apply(v)


@pytest.mark.parametrize('mtype', [0, 1, 2])
def test_natural_frequencies_vs_eigsh(mtype):
    from pyfe3d_amd.eigen import GeneralizedEigen
    from pyfe3d_amd.shellprop import isotropic_plate
    a, b, E, nu, rho, h = 0.3, 0.5, 203.e9, 0.33, 7.83e3, 0.01
    m = _plate(9, 11, a, b)
    rec = isotropic_plate(thickness=h, E=E, nu=nu, calc_scf=True, rho=rho).record()
    x, y = m.ncoords[:, 0], m.ncoords[:, 1]
    edge = np.isclose(x, 0) | np.isclose(x, a) | np.isclose(y, 0) | np.isclose(y, b)
    fixed = np.zeros(m.n_dof, bool)
    for d in range(3):
        fixed[d::6] = edge
    rK, cK, vK = _gpu_coo('KC0', m, rec)
    rM, cM, vM = _gpu_coo('M', m, rec, mtype)
    planK, K = _bsr(rK, cK, vK, m.n_nodes)
    _, M = _bsr(rM, cM, vM, m.n_nodes, planK)
    ev = GeneralizedEigen(planK, K, M, fixed, sigma=-1.)
    lam, X, info = ev.solve(k=6, tol=1e-10)
    # scipy on the oracle's matrices, as the reference test does
    bu = ~fixed
    Kc = _cpu('KC0', m, rec)[bu, :][:, bu]
    Mc = _cpu('M', m, rec, mtype)[bu, :][:, bu]
    ref = np.sort(sla.eigsh(A=Kc, M=Mc, sigma=-1., which='LM', k=6, tol=1e-12)[0])
    print(mtype, lam, ref, info['residual'])
    np.testing.assert_allclose(lam, ref, rtol=1e-8)
    # the reference test's closed form (5 %)
    D = 2 * h ** 3 * E / (3 * (1 - nu ** 2))
    wmn = (1 / a ** 2 + 1 / b ** 2) * np.sqrt(D * np.pi ** 4 / (2 * rho * h)) / 2
    assert np.isclose(wmn, lam[0] ** 0.5, rtol=0.05)
    # eigenvectors: K x = lam M x on the free DOFs
    Xh = X.cpu().numpy()
    for i in range(3):
        xi = Xh[bu, i]
        r = Kc @ xi - lam[i] * (Mc @ xi)
        assert np.linalg.norm(r) <= 1e-6 * np.linalg.norm(Kc @ xi)
    assert np.abs(Xh[fixed]).max() == 0.


def test_linear_buckling_vs_eigsh():
    """Simply supported plate under Nxx = -1 N/m (KG given stress): critical
    load factors of K x = lambda (-KG) x."""
    from pyfe3d_amd.eigen import GeneralizedEigen
    from pyfe3d_amd.shellprop import isotropic_plate
    a, b, E, nu, h = 0.3, 0.5, 70e9, 0.33, 0.002
    m = _plate(16, 21, a, b)
    rec = isotropic_plate(thickness=h, E=E, nu=nu, calc_scf=True).record()
    x, y = m.ncoords[:, 0], m.ncoords[:, 1]
    edge = np.isclose(x, 0) | np.isclose(x, a) | np.isclose(y, 0) | np.isclose(y, b)
    fixed = np.zeros(m.n_dof, bool)
    fixed[2::6] = edge
    fixed[0::6] = np.isclose(x, 0)
    fixed[1::6] = np.isclose(y, 0)
    stress = [-1., 0., 0.]
    rK, cK, vK = _gpu_coo('KC0', m, rec)
    rG, cG, vG = _gpu_coo('KGs', m, rec, stress=stress)
    planK, K = _bsr(rK, cK, vK, m.n_nodes)
    _, KG = _bsr(rG, cG, vG, m.n_nodes, planK)
    ev = GeneralizedEigen(planK, K, -KG, fixed, sigma=0.)
    lam, X, info = ev.solve(k=4, tol=1e-10)
    bu = ~fixed
    Kc = _cpu('KC0', m, rec)[bu, :][:, bu]
    KGc = _cpu('KGs', m, rec, stress=stress)[bu, :][:, bu]
    # scipy: eigenvalues mu of KG x = mu K x; lambda = -1/mu
    mu = sla.eigsh(A=KGc, M=Kc, k=4, which='SA', tol=1e-12)[0]
    ref = np.sort(-1. / mu[mu < 0])[:4]
    print(lam, ref, info['residual'])
    np.testing.assert_allclose(lam[:len(ref)], ref, rtol=1e-8)
    # classical plate buckling (simply supported, uniaxial): k pi^2 D / b^2
    D = E * h ** 3 / (12 * (1 - nu ** 2))
    kmin = min((mm * b / a + a / (mm * b)) ** 2 for mm in range(1, 6))
    Ncr = kmin * np.pi ** 2 * D / b ** 2
    assert np.isclose(lam[0], Ncr, rtol=0.05), (lam[0], Ncr)

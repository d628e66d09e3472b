"""COO->BSR, block-Jacobi PCG (FP64) and FP32/FP64 iterative refinement on
the GPU against the reference pipeline (oracle assembly + scipy) and the
reference test's known answer."""
import numpy as np
import pytest
import torch

from reference_pipeline import assemble_cpu, solve_cpu

pytestmark = pytest.mark.gpu
DEV = 'cuda'


def gpu_kc0(etype, mesh, conn, props, pid=None):
    from pyfe3d_amd import assembly as A
    E = conn.shape[0]
    size = A.slot_count(etype, 'KC0')
    r = torch.empty(size * E, dtype=torch.int64, device=DEV)
    c = torch.empty_like(r)
    v = torch.zeros(size * E, dtype=torch.float64, device=DEV)
    A.shell_kc0(etype, torch.as_tensor(mesh.x, device=DEV),
                torch.as_tensor(6 * conn, device=DEV), r, c, v,
                torch.as_tensor(np.atleast_2d(props), device=DEV),
                pid=None if pid is None else torch.as_tensor(pid, dtype=torch.int32, device=DEV))
    return r, c, v


def test_bsr_plan_matches_scipy():
    from pyfe3d_amd.mesh import PlateMesh
    from pyfe3d_amd.shellprop import laminated_plate
    from pyfe3d_amd.sparse import BSRPlan
    from scipy.sparse import coo_matrix
    m = PlateMesh(9, 7, 0.3, 0.2, jitter=0.2, seed=1)
    prop = laminated_plate([0, 45, -45, 90], 0.0002, (140e9, 10e9, 0.3, 5e9, 5e9, 5e9), 1600.)
    for etype, conn in (('quad4', m.quads), ('tria3r', m.trias)):
        r, c, v = gpu_kc0(etype, m, conn, prop.record())
        plan = BSRPlan(r, c, m.n_nodes)
        val = plan.apply(v)
        K = plan.to_scipy(val).toarray()
        Kref = coo_matrix((v.cpu().numpy(), (r.cpu().numpy(), c.cpu().numpy())),
                          shape=(m.n_dof, m.n_dof)).toarray()
        assert np.abs(K - Kref).max() <= 1e-14 * np.abs(Kref).max()
        # determinism: a second apply is bit-identical
        assert torch.equal(plan.apply(v), val)
        # structure: one 6x6 block per connected node pair
        pairs = set()
        for e in conn:
            for a in e:
                for b in e:
                    pairs.add((a, b))
        assert plan.nnzb == len(pairs)


def _reference_plate():
    """tests/test_quad4_static_point_load.py:12-115 of the reference."""
    from pyfe3d_amd.mesh import PlateMesh
    from pyfe3d_amd.shellprop import isotropic_plate
    nx, ny, a, b = 7, 11, 3., 7.
    m = PlateMesh(nx, ny, a, b)
    prop = isotropic_plate(thickness=0.005, E=200e9, nu=0.3, calc_scf=True)
    x, y = m.ncoords[:, 0], m.ncoords[:, 1]
    fixed = np.zeros(m.n_dof, dtype=bool)
    fixed[2::6] = np.isclose(x, 0.) | np.isclose(x, a) | np.isclose(y, 0) | np.isclose(y, b)
    fixed[0::6] = True
    fixed[1::6] = True
    f = np.zeros(m.n_dof)
    f[2::6][np.isclose(x, a / 2) & np.isclose(y, b / 2)] = 1.
    return m, prop, fixed, f


@pytest.mark.parametrize('method', ['fp64', 'mixed', 'ir', 'direct'])
def test_reference_point_load_known_answer(method):
    from pyfe3d_amd.sparse import BSRPlan
    from pyfe3d_amd.solver import StaticSystem
    m, prop, fixed, f = _reference_plate()
    r, c, v = gpu_kc0('quad4', m, m.quads, prop.record())
    plan = BSRPlan(r, c, m.n_nodes)
    sysm = StaticSystem(plan, plan.apply(v), fixed)
    u, info = sysm.solve(f, method=method, rtol=1e-10)
    u = u.cpu().numpy()
    assert info['relres'] <= 1e-10
    wmax_ref = 6.496928101916171e-05  # reference test, rtol 0.02
    assert np.isclose(u[2::6].max(), wmax_ref, rtol=0.02)
    rr, cc, vv = assemble_cpu('quad4', m.x, m.quads, prop.record()[None])
    uref, _ = solve_cpu(rr, cc, vv, m.n_dof, fixed, f)
    assert np.linalg.norm(u - uref) <= 1e-8 * np.linalg.norm(uref)
    assert u[fixed].max() == 0 and u[fixed].min() == 0


@pytest.mark.parametrize('method', ['fp64', 'mixed', 'ir', 'direct'])
@pytest.mark.parametrize('etype', ['quad4', 'quad4r', 'tria3r'])
def test_clamped_plate_point_load(etype, method):
    """20x20 plate, clamped, centre point load (the CPU-runnable case)."""
    from pyfe3d_amd.mesh import PlateMesh
    from pyfe3d_amd.shellprop import isotropic_plate
    from pyfe3d_amd.sparse import BSRPlan
    from pyfe3d_amd.solver import StaticSystem
    m = PlateMesh(21, 21, 1.0, 0.7)
    prop = isotropic_plate(thickness=0.005, E=70e9, nu=0.33, rho=2700., calc_scf=True)
    conn = m.trias if etype == 'tria3r' else m.quads
    fixed = m.clamped()
    f = np.zeros(m.n_dof)
    f[6 * m.centre_node() + 2] = -1000.
    r, c, v = gpu_kc0(etype, m, conn, prop.record())
    plan = BSRPlan(r, c, m.n_nodes)
    sysm = StaticSystem(plan, plan.apply(v), fixed)
    u, info = sysm.solve(f, method=method, rtol=1e-10)
    u = u.cpu().numpy()
    rr, cc, vv = assemble_cpu(etype, m.x, conn, prop.record()[None])
    uref, _ = solve_cpu(rr, cc, vv, m.n_dof, fixed, f)
    err = np.linalg.norm(u - uref) / np.linalg.norm(uref)
    print(etype, method, info, 'rel err vs spsolve %.2e' % err)
    assert err <= 1e-8

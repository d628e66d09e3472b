"""pf3d_pcg_static (no workspace, no allocation) against pf3d_pcg with a
caller workspace: same kernels, so the solutions must agree bit for bit.
Two calls in a row, the second on a smaller system, so the second runs on
the scratch array the first left behind."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _system(nx, ny, seed):
    from pyfe3d_amd import assembly as A
    from pyfe3d_amd.mesh import PlateMesh
    from pyfe3d_amd.shellprop import isotropic_plate
    from pyfe3d_amd.solver import StaticSystem
    from pyfe3d_amd.sparse import BSRPlan
    m = PlateMesh(nx, ny, 1.0, 0.7, jitter=0.2, seed=seed)
    prop = isotropic_plate(thickness=0.005, E=70e9, nu=0.33, rho=2700.)
    E = m.quads.shape[0]
    r = torch.empty(576 * E, dtype=torch.int64, device='cuda')
    c = torch.empty_like(r)
    v = torch.zeros(576 * E, dtype=torch.float64, device='cuda')
    A.shell_kc0('quad4', torch.as_tensor(m.x, device='cuda'),
                torch.as_tensor(6 * m.quads, device='cuda'), r, c, v,
                torch.as_tensor(prop.record()[None], device='cuda'))
    plan = BSRPlan(r, c, m.n_nodes)
    sysm = StaticSystem(plan, plan.apply(v), m.clamped())
    f = np.zeros(m.n_dof)
    f[6 * m.centre_node() + 2] = -1000.
    return sysm, sysm.rhs(f)


def _pcg(sysm, b, static):
    from pyfe3d_amd import _lib
    x = torch.full_like(b, float('nan'))
    iters = ctypes.c_int64()
    rel = ctypes.c_double()
    s = torch.cuda.current_stream().cuda_stream
    common = (sysm.nb, sysm.plan.row_ptr.data_ptr(), sysm.plan.col_idx.data_ptr(),
              sysm.val.data_ptr(), sysm.dinv.data_ptr(), b.data_ptr(), x.data_ptr(), 1e-8,
              100000, 0)
    if static:
        _lib.call('pf3d_pcg_static', *common, ctypes.byref(iters), ctypes.byref(rel), s)
    else:
        work = torch.full((sysm._w64.numel(),), 255, dtype=torch.uint8, device='cuda')
        _lib.call('pf3d_pcg', *common, work.data_ptr(), work.numel(), ctypes.byref(iters),
                  ctypes.byref(rel), s)
    torch.cuda.synchronize()
    return x, iters.value, rel.value


def test_static_workspace_pcg_matches_pcg():
    for nx, ny, seed in ((21, 21, 1), (13, 11, 2)):  # second call: smaller system
        print('pcg_static case', nx, ny, seed, flush=True)
        sysm, b = _system(nx, ny, seed)
        x0, it0, rel0 = _pcg(sysm, b, static=False)
        x1, it1, rel1 = _pcg(sysm, b, static=True)
        assert rel0 <= 1e-8 and rel1 <= 1e-8, (rel0, rel1)
        assert it0 == it1 and rel0 == rel1, (it0, it1, rel0, rel1)
        assert torch.equal(x0, x1)
        # the true residual of the static solve, compensated
        r = sysm.residual(b, x1)
        assert float(r.norm() / b.norm()) <= 1e-8


def test_static_workspace_rejects_large_systems():
    from pyfe3d_amd import _lib
    with pytest.raises(_lib.Pf3dError, match='static workspace'):
        t = torch.zeros(8, dtype=torch.float64, device='cuda')
        _lib.call('pf3d_pcg_static', 131073, t.data_ptr(), t.data_ptr(), t.data_ptr(),
                  t.data_ptr(), t.data_ptr(), t.data_ptr(), 1e-8, 10, 0,
                  ctypes.byref(ctypes.c_int64()), ctypes.byref(ctypes.c_double()),
                  torch.cuda.current_stream().cuda_stream)

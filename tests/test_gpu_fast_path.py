"""The contiguous-slot fast path of KC0 (no init_k, 16-byte aligned arrays:
paired 16-byte stores) entry by entry against the oracle, in the three
modes the benchmarks and the design sweep launch: values only + overwrite
(the headline assembly), indices + overwrite, values only + accumulate.
Output arrays are pre-filled with garbage, so every slot must be written
(overwrite) or added to exactly once (accumulate); indices bit-exact,
values within 1e-12 of each element's largest entry (quad4.pyx:1204-1313,
quad4r.pyx:1145-4619, tria3r.pyx:950-2974)."""
import numpy as np
import pytest
import torch

from golden_cases import TOL
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = 'cuda'


def T(a, dtype):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def _mesh(etype, nx, ny, seed):
    rng = np.random.default_rng(seed)
    xs, ys = np.meshgrid(np.linspace(0, 1.0, nx + 1), np.linspace(0, 0.7, ny + 1), indexing='ij')
    X, Y = xs.copy(), ys.copy()
    X[1:-1, 1:-1] += (1.0 / nx) * rng.uniform(-0.2, 0.2, (nx - 1, ny - 1))
    Y[1:-1, 1:-1] += (0.7 / ny) * rng.uniform(-0.2, 0.2, (nx - 1, ny - 1))
    Z = 0.02 * np.sin(3 * X) * np.cos(2 * Y)
    x = np.vstack([X.ravel(), Y.ravel(), Z.ravel()]).T.ravel()
    nid = np.arange((nx + 1) * (ny + 1)).reshape(nx + 1, ny + 1)
    q = np.stack([nid[:-1, :-1].ravel(), nid[1:, :-1].ravel(), nid[1:, 1:].ravel(),
                  nid[:-1, 1:].ravel()], axis=1)
    conn = q if etype != 'tria3r' else np.concatenate([q[:, [0, 1, 2]], q[:, [0, 2, 3]]])
    E = conn.shape[0]
    props = np.zeros((9, 32))
    for p in range(9):
        a = rng.uniform(0.5, 1.5, 6) * 1e8
        a[2] *= 0.1
        props[p, 0:6] = a
        props[p, 6:12] = rng.uniform(-1, 1, 6) * 1e3
        props[p, 12:18] = a * 1e-6
        props[p, 18:21] = [3e7, 2e6, 4e7]
        props[p, 21:24] = [0.83, 0.81, 0.002 * (1 + p % 3)]
    return dict(x=x, conn=conn, E=E, props=props, pid=rng.integers(0, 9, E),
                xmat=rng.normal(size=(E, 3)), k6rot=rng.uniform(1, 1000, E),
                alpha=rng.uniform(0.2, 1, E), hg=rng.uniform(0.5, 2, (E, 5)))


def _oracle(etype, d):
    E = d['E']
    size = O.slot_count(etype, 'KC0')
    r = np.zeros(size * E, np.int64)
    c = np.zeros_like(r)
    v = np.zeros(size * E)
    O.assemble(etype, 'KC0', d['x'], d['conn'], size * np.arange(E), d['xmat'], d['pid'],
               d['props'], r, c, v, k6rot=d['k6rot'], alpha=d['alpha'], hg=d['hg'])
    return r, c, v, size


def _worst(v, vref, size):
    dv = np.abs(v - vref).reshape(-1, size).max(axis=1)
    scale = np.abs(vref).reshape(-1, size).max(axis=1)
    return float((dv / scale).max())


@pytest.mark.parametrize('etype', ['quad4', 'quad4r', 'tria3r'])
@pytest.mark.parametrize('mode', ['values_overwrite', 'indices_overwrite', 'values_accumulate'])
def test_contiguous_fast_path_vs_oracle(etype, mode):
    from pyfe3d_amd import assembly as A
    d = _mesh(etype, 36, 25, {'quad4': 1, 'quad4r': 2, 'tria3r': 3}[etype])
    rr, cc, vv, size = _oracle(etype, d)
    E = d['E']
    values_only = mode != 'indices_overwrite'
    overwrite = mode != 'values_accumulate'
    base = 3.25
    v = torch.full((size * E,), float('nan') if overwrite else base, dtype=torch.float64,
                   device=DEV)
    r = torch.full((size * E,), -7, dtype=torch.int64, device=DEV)
    c = torch.full((size * E,), -7, dtype=torch.int64, device=DEV)
    # the fast path needs 16-byte aligned arrays (torch allocations are)
    assert v.data_ptr() % 16 == 0 and r.data_ptr() % 16 == 0 and c.data_ptr() % 16 == 0
    print('fast path', etype, mode, E, 'elements', flush=True)
    A.shell_kc0(etype, T(d['x'], torch.float64), T(6 * d['conn'], torch.int64),
                None if values_only else r, None if values_only else c, v,
                T(d['props'], torch.float64), pid=T(d['pid'], torch.int32),
                xmat=T(d['xmat'], torch.float64), k6rot=T(d['k6rot'], torch.float64),
                alpha=T(d['alpha'], torch.float64), hg=T(d['hg'], torch.float64),
                values_only=values_only, overwrite=overwrite)
    torch.cuda.synchronize()
    vg = v.cpu().numpy()
    if not overwrite:
        vg = vg - base
    assert np.isfinite(vg).all()
    err = _worst(vg, vv, size)
    print('max rel err %.2e' % err)
    assert err <= TOL, err
    if values_only:
        assert bool((r == -7).all()) and bool((c == -7).all())
    else:
        assert (r.cpu().numpy() == rr).all() and (c.cpu().numpy() == cc).all()


def test_fast_path_matches_permuted_slots_bitwise():
    """The same elements through the general path (permuted init_k) and the
    fast path give the same bits slot for slot."""
    from pyfe3d_amd import assembly as A
    d = _mesh('quad4', 30, 20, 9)
    E = d['E']
    perm = np.random.default_rng(4).permutation(E)
    args = (T(d['x'], torch.float64), T(6 * d['conn'], torch.int64))
    kw = dict(pid=T(d['pid'], torch.int32), xmat=T(d['xmat'], torch.float64))
    v_fast = torch.empty(576 * E, dtype=torch.float64, device=DEV)
    A.shell_kc0('quad4', *args, None, None, v_fast, T(d['props'], torch.float64),
                values_only=True, overwrite=True, **kw)
    v_perm = torch.zeros(576 * E, dtype=torch.float64, device=DEV)
    A.shell_kc0('quad4', *args, None, None, v_perm, T(d['props'], torch.float64),
                init_k=T(576 * perm, torch.int64), values_only=True, overwrite=True, **kw)
    back = v_perm.view(E, 576)[T(perm, torch.int64)]
    assert torch.equal(back, v_fast.view(E, 576))

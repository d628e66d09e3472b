"""Truss and Spring on the GPU through the C ABI: the reference's golden
vectors (truss.pyx, spring.pyx), and large random batches against the CPU
oracle (indices bit-exact; values within 1e-12 of each element's largest
entry; internal forces within 1e-12 of the largest force)."""
import numpy as np
import pytest
import torch

from golden_cases import TOL, cases, element_errors
from oracle import oracle as O
from test_gpu_assembly import T

pytestmark = pytest.mark.gpu
DEV = 'cuda'
FINT_TOL = 1e-12


def _axes(d):
    return np.concatenate([d['axes'], d['vxy']], axis=1)


def _gpu_rod(etype, d, key, mtype, size, init_k=True, values_only=False, overwrite=False,
             v0=None):
    from pyfe3d_amd import assembly as A
    E = d['conn'].shape[0]
    r = torch.zeros(size * E, dtype=torch.int64, device=DEV)
    c = torch.zeros_like(r)
    v = torch.zeros(size * E, dtype=torch.float64, device=DEV) if v0 is None else T(v0)
    x = T(d['x'], torch.float64)
    conn = T(6 * d['conn'], torch.int64)
    kw = dict(pid=T(d.get('pid', np.arange(E)), torch.int32), values_only=values_only,
              overwrite=overwrite)
    if init_k:
        kw['init_k'] = T(size * d['perm'], torch.int64)
    if etype == 'spring':
        A.spring_kc0(x, conn, T(_axes(d), torch.float64), r, c, v, T(d['k'], torch.float64), **kw)
    elif key == 'KC0':
        A.truss_kc0(x, conn, r, c, v, T(O.props_table(d['props'], 'truss'), torch.float64), **kw)
    else:
        A.truss_m(x, conn, r, c, v, T(O.props_table(d['props'], 'truss'), torch.float64),
                  mtype=mtype, **kw)
    torch.cuda.synchronize()
    return r.cpu().numpy(), c.cpu().numpy(), v.cpu().numpy()


ROD_KEYS = {'truss': (('KC0', 0), ('M0', 0), ('M1', 1)), 'spring': (('KC0', 0),)}


@pytest.mark.parametrize('etype', ['truss', 'spring'])
def test_rod_golden_gpu(etype):
    from pyfe3d_amd import assembly as A
    bad = []
    for name, d in cases(etype):
        E = d['conn'].shape[0]
        g, _ = A.geometry(etype, T(d['x'], torch.float64), T(6 * d['conn'], torch.int64),
                          T(_axes(d), torch.float64) if etype == 'spring' else None)
        g = g.cpu().numpy()
        np.testing.assert_allclose(g[:, :9], d['geom'][:, :9], rtol=0, atol=1e-14)
        if etype == 'truss':
            np.testing.assert_allclose(g[:, 13], d['geom'][:, 9], rtol=1e-14)
        for key, mtype in ROD_KEYS[etype]:
            size = d[key + 'v'].size // E
            r, c, v = _gpu_rod(etype, d, key, mtype, size)
            if not ((r == d[key + 'r']).all() and (c == d[key + 'c']).all()):
                bad.append((name, key, 'indices'))
            err = element_errors(v, d[key + 'v'], size, d['perm'])
            print(etype, name, key, 'max rel err %.2e' % err)
            if not err <= TOL:
                bad.append((name, key, err))
        # internal forces, then += semantics
        props = (T(O.props_table(d['props'], 'truss'), torch.float64) if etype == 'truss'
                 else T(d['k'], torch.float64))
        fint = torch.zeros(d['u'].size, dtype=torch.float64, device=DEV)
        for rep in (1, 2):
            A.update_fint(etype, T(d['x'], torch.float64), T(6 * d['conn'], torch.int64),
                          T(d['u'], torch.float64), fint, props,
                          axes=T(_axes(d), torch.float64) if etype == 'spring' else None,
                          pid=T(np.arange(E), torch.int32))
            torch.cuda.synchronize()
            scale = np.abs(d['fint']).max()
            err = np.abs(fint.cpu().numpy() - rep * d['fint']).max() / scale
            print(etype, name, 'fint x%d max rel err %.2e' % (rep, err))
            if not err <= rep * FINT_TOL:
                bad.append((name, 'fint', rep, err))
    assert not bad, bad


def _random_rods(etype, E, n_nodes, seed):
    """Random node cloud, random node pairs (no zero-length rod), mixed
    property records, shuffled slot blocks."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(-2, 2, 3 * n_nodes)
    a = rng.integers(0, n_nodes, E)
    b = (a + rng.integers(1, n_nodes, E)) % n_nodes
    conn = np.stack([a, b], axis=1).astype(np.int64)
    d = dict(x=x, conn=conn, perm=rng.permutation(E), pid=rng.integers(0, 7, E),
             u=rng.uniform(-1, 1, 6 * n_nodes) * 1e-3)
    if etype == 'truss':
        p = np.zeros((7, 15))
        p[:, 0] = rng.uniform(1e-5, 1e-3, 7)           # A
        p[:, 1] = rng.uniform(50e9, 210e9, 7)          # E
        p[:, 2] = p[:, 1] / 2.6                        # G
        p[:, 6] = rng.uniform(1e-10, 1e-7, 7)          # J
        p[:, 9] = 7850 * p[:, 0]                       # intrho
        p[:, 10:12] = rng.uniform(-1, 1, (7, 2)) * 1e-2
        p[:, 12:14] = rng.uniform(1e-6, 1e-4, (7, 2))
        d['props'] = p
    else:
        d['k'] = np.exp(rng.uniform(np.log(1e2), np.log(1e8), (7, 6)))
        d['axes'] = rng.normal(size=(E, 3))
        d['vxy'] = rng.normal(size=(E, 3))
    return d


@pytest.mark.parametrize('etype', ['truss', 'spring'])
def test_random_rods_vs_oracle(etype):
    """20,011 rods over 3,000 nodes: every slot of every matrix and the
    internal forces against the oracle; accumulate / overwrite / values-only
    with the default (contiguous) slot blocks."""
    from pyfe3d_amd import assembly as A
    E, N = 20011, 3000
    d = _random_rods(etype, E, N, seed=11 if etype == 'truss' else 12)
    ev = _axes(d) if etype == 'spring' else np.zeros((E, 3))
    props = d['props'] if etype == 'truss' else d['k']
    for key, mtype in ROD_KEYS[etype]:
        mat = 'KC0' if key == 'KC0' else 'M'
        size = O.slot_count(etype, mat, mtype)
        rr = np.zeros(size * E, np.int64); cc = np.zeros_like(rr); vv = np.zeros(size * E)
        O.assemble(etype, mat, d['x'], d['conn'], size * d['perm'], ev, d['pid'], props,
                   rr, cc, vv, mtype=mtype)
        r, c, v = _gpu_rod(etype, d, key, mtype, size)
        assert (r == rr).all() and (c == cc).all(), key
        err = element_errors(v, vv, size, d['perm'])
        print(etype, key, 'max rel err %.2e' % err)
        assert err <= TOL, (key, err)
        # contiguous blocks: v += twice from a base, then overwrite values only
        ident = np.arange(E)
        rr2 = np.zeros_like(rr); cc2 = np.zeros_like(rr); vv2 = np.zeros_like(vv)
        O.assemble(etype, mat, d['x'], d['conn'], size * ident, ev, d['pid'], props,
                   rr2, cc2, vv2, mtype=mtype)
        base = np.full(vv.size, 1.5)
        _, _, v = _gpu_rod(etype, d, key, mtype, size, init_k=False, v0=base)
        assert element_errors(v - 1.5, vv2, size, ident) <= TOL, key
        r, c, v = _gpu_rod(etype, d, key, mtype, size, init_k=False, v0=base, overwrite=True,
                           values_only=True)
        assert element_errors(v, vv2, size, ident) <= TOL, key
        assert (r == 0).all() and (c == 0).all()
    ref = O.fint(etype, d['x'], d['conn'], ev, d['pid'], props, d['u'])
    fint = torch.zeros(6 * N, dtype=torch.float64, device=DEV)
    A.update_fint(etype, T(d['x'], torch.float64), T(6 * d['conn'], torch.int64),
                  T(d['u'], torch.float64), fint,
                  T(O.props_table(props, 'truss') if etype == 'truss' else props, torch.float64),
                  axes=T(ev, torch.float64) if etype == 'spring' else None,
                  pid=T(d['pid'], torch.int32))
    torch.cuda.synchronize()
    err = np.abs(fint.cpu().numpy() - ref).max() / np.abs(ref).max()
    print(etype, 'fint max rel err %.2e' % err)
    assert err <= FINT_TOL


def test_rod_argument_errors():
    from pyfe3d_amd import _lib, assembly as A
    x = torch.zeros(6, dtype=torch.float64, device=DEV)
    conn = torch.tensor([[0, 6]], dtype=torch.int64, device=DEV)
    r = torch.zeros(144, dtype=torch.int64, device=DEV)
    v = torch.zeros(144, dtype=torch.float64, device=DEV)
    with pytest.raises(Exception):
        A.truss_m(x, conn, r, r, v, torch.zeros((1, 16), dtype=torch.float64, device=DEV), mtype=2)
    with pytest.raises(Exception):   # springs need their axes
        A.spring_kc0(x, conn, None, r, r, v, torch.zeros((1, 6), dtype=torch.float64, device=DEV))
    assert A.slot_count('truss', 'KC0') == 72 and A.slot_count('spring', 'KC0') == 72
    assert A.slot_count('truss', 'M', 0) == 144 and A.slot_count('truss', 'M', 1) == 36
    with pytest.raises(Exception):
        A.slot_count('spring', 'M')

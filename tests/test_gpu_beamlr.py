"""BeamLR (linear Timoshenko beam, reduced integration) on the GPU through
the C ABI: the reference's golden vectors (beamlr.pyx) and a large random
batch against the oracle (indices exact, values within 1e-12 of each
element's largest entry, internal forces within 1e-12 of the largest)."""
import numpy as np
import pytest
import torch

from golden_cases import TOL, cases, element_errors
from oracle import oracle as O
from test_gpu_assembly import T

pytestmark = pytest.mark.gpu
DEV = 'cuda'
KEYS = (('KC0', 'KC0', 0), ('KG', 'KG', 0), ('M0', 'M', 0), ('M1', 'M', 1))


def _gpu(d, key, mtype, size, init_k=True):
    from pyfe3d_amd import assembly as A
    E = d['conn'].shape[0]
    r = torch.zeros(size * E, dtype=torch.int64, device=DEV)
    c = torch.zeros_like(r)
    v = torch.zeros(size * E, dtype=torch.float64, device=DEV)
    args = (T(d['x'], torch.float64), T(6 * d['conn'], torch.int64), T(d['vxy'], torch.float64))
    kw = dict(pid=T(d['pid'], torch.int32), etype='beamlr')
    if init_k:
        kw['init_k'] = T(size * d['perm'], torch.int64)
    props = T(O.props_table(d['props'], 'beamlr'), torch.float64)
    if key == 'KC0':
        A.beam_kc0(*args, r, c, v, props, **kw)
    elif key == 'KG':
        A.beam_kg(*args, T(d['u'], torch.float64), r, c, v, props, **kw)
    else:
        A.beam_m(*args, r, c, v, props, mtype=mtype, **kw)
    torch.cuda.synchronize()
    return r.cpu().numpy(), c.cpu().numpy(), v.cpu().numpy()


def _fint(d):
    from pyfe3d_amd import assembly as A
    fint = torch.zeros(d['u'].size, dtype=torch.float64, device=DEV)
    A.update_fint('beamlr', T(d['x'], torch.float64), T(6 * d['conn'], torch.int64),
                  T(d['u'], torch.float64), fint,
                  T(O.props_table(d['props'], 'beamlr'), torch.float64),
                  vxy=T(d['vxy'], torch.float64), pid=T(d['pid'], torch.int32))
    torch.cuda.synchronize()
    return fint.cpu().numpy()


def test_beamlr_golden_gpu():
    bad = []
    for name, d in cases('beamlr'):
        E = d['conn'].shape[0]
        d = dict(d, pid=np.arange(E))
        for key, _, mtype in KEYS:
            size = d[key + 'v'].size // E
            r, c, v = _gpu(d, key, mtype, size)
            if not ((r == d[key + 'r']).all() and (c == d[key + 'c']).all()):
                bad.append((name, key, 'indices'))
            err = element_errors(v, d[key + 'v'], size, d['perm'])
            print('beamlr', name, key, 'max rel err %.2e' % err)
            if not err <= TOL:
                bad.append((name, key, err))
        f = _fint(d)
        err = np.abs(f - d['fint']).max() / np.abs(d['fint']).max()
        print('beamlr', name, 'fint max rel err %.2e' % err)
        if not err <= 1e-12:
            bad.append((name, 'fint', err))
    assert not bad, bad


def test_beamlr_random_vs_oracle():
    """10,007 beams over 2,000 random nodes, 5 property records with
    offsets (Ay, Az), Iyz and first mass moments."""
    rng = np.random.default_rng(21)
    E, N = 10007, 2000
    x = rng.uniform(-1, 1, 3 * N)
    a = rng.integers(0, N, E)
    conn = np.stack([a, (a + rng.integers(1, N, E)) % N], axis=1).astype(np.int64)
    props = np.zeros((5, 15))
    props[:, 0] = rng.uniform(1e-5, 1e-3, 5)
    props[:, 1] = 70e9
    props[:, 2] = 26e9
    props[:, 3:5] = rng.uniform(1e-9, 1e-7, (5, 2))
    props[:, 5] = rng.uniform(-1, 1, 5) * 1e-9
    props[:, 6] = props[:, 3] + props[:, 4]
    props[:, 7:9] = rng.uniform(-1, 1, (5, 2)) * 1e-6
    props[:, 9] = 2700 * props[:, 0]
    props[:, 10:12] = rng.uniform(-1, 1, (5, 2)) * 1e-3
    props[:, 12:14] = 2700 * props[:, 3:5]
    props[:, 14] = 2700 * props[:, 5]
    d = dict(x=x, conn=conn, perm=rng.permutation(E), pid=rng.integers(0, 5, E),
             vxy=rng.normal(size=(E, 3)), u=rng.uniform(-1, 1, 6 * N) * 1e-3, props=props)
    for key, mat, mtype in KEYS:
        size = O.slot_count('beamlr', mat, mtype)
        rr = np.zeros(size * E, np.int64); cc = np.zeros_like(rr); vv = np.zeros(size * E)
        O.assemble('beamlr', mat, x, conn, size * d['perm'], d['vxy'], d['pid'], props,
                   rr, cc, vv, mtype=mtype, u=d['u'])
        r, c, v = _gpu(d, key, mtype, size)
        assert (r == rr).all() and (c == cc).all(), key
        err = element_errors(v, vv, size, d['perm'])
        print('beamlr random', key, 'max rel err %.2e' % err)
        assert err <= TOL, (key, err)
    ref = O.fint('beamlr', x, conn, d['vxy'], d['pid'], props, d['u'])
    f = _fint(d)
    err = np.abs(f - ref).max() / np.abs(ref).max()
    print('beamlr random fint max rel err %.2e' % err)
    assert err <= 1e-12

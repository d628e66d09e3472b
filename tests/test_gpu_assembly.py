"""GPU element assembly through the C ABI against the reference's golden
vectors and against the CPU oracle (indices bit-exact; values within
1e-12 of each element's largest entry)."""
import numpy as np
import pytest
import torch

from golden_cases import SHELL_MATS, SHELLS, TOL, TOL_CASE, cases, element_errors
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = 'cuda'


def T(a, dtype=None):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def gpu_shell(etype, d, key, mat, mtype, values_only=False, overwrite=False, v0=None):
    from pyfe3d_amd import assembly as A
    E = d['conn'].shape[0]
    size = d[key + 'v'].size // E
    r = torch.zeros(size * E, dtype=torch.int64, device=DEV)
    c = torch.zeros_like(r)
    v = torch.zeros(size * E, dtype=torch.float64, device=DEV) if v0 is None else T(v0)
    x = T(d['x'], torch.float64)
    conn = T(6 * d['conn'], torch.int64)
    init_k = T(size * d['perm'], torch.int64)
    xmat = T(d['xmat'], torch.float64)
    props = T(O.props_table(d['props'], etype), torch.float64)
    pid = T(d['pid'], torch.int32)
    kw = dict(init_k=init_k, xmat=xmat, values_only=values_only, overwrite=overwrite)
    if mat == 'KC0':
        A.shell_kc0(etype, x, conn, r, c, v, props, pid=pid, k6rot=T(d['k6rot'], torch.float64),
                    alpha=T(d['alpha'], torch.float64), hg=T(d['hg'], torch.float64), **kw)
    elif mat == 'KG':
        A.shell_kg(etype, x, conn, T(d['u'], torch.float64), r, c, v, props, pid=pid, **kw)
    elif mat == 'KGs':
        A.shell_kg_stress(etype, x, conn, T(d['stress'], torch.float64), r, c, v, **kw)
    else:
        A.shell_m(etype, x, conn, r, c, v, props, mtype=mtype, pid=pid, **kw)
    torch.cuda.synchronize()
    return r.cpu().numpy(), c.cpu().numpy(), v.cpu().numpy(), size


@pytest.mark.parametrize('etype', SHELLS)
def test_shell_golden_gpu(etype):
    bad = []
    for name, d in cases(etype):
        for key, mat, mtype in SHELL_MATS:
            r, c, v, size = gpu_shell(etype, d, key, mat, mtype)
            if not ((r == d[key + 'r']).all() and (c == d[key + 'c']).all()):
                bad.append((name, key, 'indices'))
            err = element_errors(v, d[key + 'v'], size, d['perm'])
            print(etype, name, key, 'max rel err %.2e' % err)
            if not err <= TOL_CASE.get((etype, name), TOL):
                bad.append((name, key, err))
    assert not bad, bad


@pytest.mark.parametrize('etype', SHELLS)
def test_geometry_gpu(etype):
    from pyfe3d_amd import assembly as A
    for name, d in cases(etype):
        g, xe = A.geometry(etype, T(d['x'], torch.float64), T(6 * d['conn'], torch.int64),
                           T(d['xmat'], torch.float64))
        np.testing.assert_allclose(g.cpu().numpy(), d['geom'], rtol=0, atol=1e-14)
        np.testing.assert_allclose(xe.cpu().numpy(), d['xe'], rtol=0, atol=1e-14)


def test_beam_golden_gpu():
    from pyfe3d_amd import assembly as A
    bad = []
    for name, d in cases('beamc'):
        E = d['conn'].shape[0]
        for key, mtype in (('KC0', 0), ('KG', 0), ('M0', 0), ('M1', 1)):
            size = d[key + 'v'].size // E
            r = torch.zeros(size * E, dtype=torch.int64, device=DEV)
            c = torch.zeros_like(r)
            v = torch.zeros(size * E, dtype=torch.float64, device=DEV)
            args = (T(d['x'], torch.float64), T(6 * d['conn'], torch.int64),
                    T(d['vxy'], torch.float64))
            kw = dict(pid=T(np.arange(E), torch.int32), init_k=T(size * d['perm'], torch.int64))
            props = T(O.props_table(d['props'], 'beamc'), torch.float64)
            if key == 'KC0':
                A.beam_kc0(*args, r, c, v, props, **kw)
            elif key == 'KG':
                A.beam_kg(*args, T(d['u'], torch.float64), r, c, v, props, **kw)
            else:
                A.beam_m(*args, r, c, v, props, mtype=mtype, **kw)
            if not ((r.cpu().numpy() == d[key + 'r']).all() and
                    (c.cpu().numpy() == d[key + 'c']).all()):
                bad.append((name, key, 'indices'))
            err = element_errors(v.cpu().numpy(), d[key + 'v'], size, d['perm'])
            print('beamc', name, key, 'max rel err %.2e' % err)
            if not err <= TOL:
                bad.append((name, key, err))
    assert not bad, bad


def test_accumulate_overwrite_values_only():
    """v += (default), v = (overwrite), indices untouched with values_only
    (quad4.pyx:1278, 1313)."""
    name, d = next(cases('quad4'))
    E = d['conn'].shape[0]
    ref = d['KC0v']
    base = np.full(ref.size, 3.25)
    _, _, v, _ = gpu_shell('quad4', d, 'KC0', 'KC0', 0, v0=base)
    assert element_errors(v - 3.25, ref, 576, d['perm']) <= TOL
    r, c, v, _ = gpu_shell('quad4', d, 'KC0', 'KC0', 0, v0=base, overwrite=True,
                           values_only=True)
    assert element_errors(v, ref, 576, d['perm']) <= TOL
    assert (r == 0).all() and (c == 0).all()


def _random_mesh(nx, ny, seed):
    rng = np.random.default_rng(seed)
    xs, ys = np.meshgrid(np.linspace(0, 1.0, nx + 1), np.linspace(0, 0.7, ny + 1), indexing='ij')
    X = xs.copy(); Y = ys.copy()
    X[1:-1, 1:-1] += (1.0 / nx) * rng.uniform(-0.2, 0.2, (nx - 1, ny - 1))
    Y[1:-1, 1:-1] += (0.7 / ny) * rng.uniform(-0.2, 0.2, (nx - 1, ny - 1))
    Z = 0.01 * np.sin(3 * X) * np.cos(2 * Y)
    x = np.vstack([X.ravel(), Y.ravel(), Z.ravel()]).T.ravel()
    nid = np.arange((nx + 1) * (ny + 1)).reshape(nx + 1, ny + 1)
    conn = np.stack([nid[:-1, :-1].ravel(), nid[1:, :-1].ravel(), nid[1:, 1:].ravel(),
                     nid[:-1, 1:].ravel()], axis=1)
    return x, conn, rng


@pytest.mark.parametrize('etype', SHELLS)
def test_random_mesh_vs_oracle(etype):
    """A few thousand jittered, warped elements with per-element laminates."""
    from pyfe3d_amd import assembly as A
    x, quads, rng = _random_mesh(40, 30, 7)
    conn = quads if etype != 'tria3r' else np.concatenate(
        [quads[:, [0, 1, 2]], quads[:, [0, 2, 3]]])
    E = conn.shape[0]
    props = np.zeros((17, 32))
    for p in range(17):
        A_ = rng.uniform(0.5, 1.5, 6) * 1e8; A_[2] *= 0.1; A_[4] *= 0.1
        props[p, 0:6] = A_
        props[p, 6:12] = rng.uniform(-1, 1, 6) * 1e3
        props[p, 12:18] = A_ * 1e-6
        props[p, 18:21] = [3e7, 2e6, 4e7]
        props[p, 21:24] = [0.83, 0.81, 0.002 * (1 + p % 3)]
        props[p, 25:28] = [5., 1e-4, 1e-6]
    pid = rng.integers(0, 17, E)
    xmat = rng.normal(size=(E, 3))
    k6rot = rng.uniform(1, 1000, E)
    alpha = rng.uniform(0.2, 1, E)
    hg = rng.uniform(0.5, 2, (E, 5))
    u = rng.normal(size=x.size * 2) * 1e-3
    for mat, mtype in [('KC0', 0), ('KG', 0), ('M', 0)]:
        size = O.slot_count(etype, mat, mtype)
        rr = np.zeros(size * E, np.int64); cc = np.zeros_like(rr); vv = np.zeros(size * E)
        O.assemble(etype, mat, x, conn, size * np.arange(E), xmat, pid, props, rr, cc, vv,
                   mtype=mtype, k6rot=k6rot, alpha=alpha, hg=hg, u=u)
        r = torch.zeros(size * E, dtype=torch.int64, device=DEV)
        c = torch.zeros_like(r)
        v = torch.zeros(size * E, dtype=torch.float64, device=DEV)
        args = (T(x, torch.float64), T(6 * conn, torch.int64))
        kw = dict(pid=T(pid, torch.int32), xmat=T(xmat, torch.float64))
        if mat == 'KC0':
            A.shell_kc0(etype, *args, r, c, v, T(props, torch.float64), k6rot=T(k6rot, torch.float64),
                        alpha=T(alpha, torch.float64), hg=T(hg, torch.float64), **kw)
        elif mat == 'KG':
            A.shell_kg(etype, *args, T(u, torch.float64), r, c, v, T(props, torch.float64), **kw)
        else:
            A.shell_m(etype, *args, r, c, v, T(props, torch.float64), mtype=mtype, **kw)
        assert (r.cpu().numpy() == rr).all() and (c.cpu().numpy() == cc).all()
        err = element_errors(v.cpu().numpy(), vv, size, np.arange(E))
        assert err <= TOL, (etype, mat, err)


def test_empty_batch_is_a_no_op():
    from pyfe3d_amd import assembly as A
    x = torch.zeros(12, dtype=torch.float64, device=DEV)
    conn = torch.zeros((0, 4), dtype=torch.int64, device=DEV)
    v = torch.zeros(0, dtype=torch.float64, device=DEV)
    A.shell_kc0('quad4', x, conn, v.long(), v.long(), v, torch.zeros((1, 32), dtype=torch.float64,
                                                                          device=DEV))


@pytest.mark.parametrize('etype', ['quad4', 'quad4r'])
def test_piston_theory_golden_gpu(etype):
    """KA_beta, KA_gamma, CA against the reference's vectors (indices exact,
    values within 1e-12), and a random 3-D mesh against the oracle."""
    from pyfe3d_amd import assembly as A
    bad = []
    for name, d in cases(etype):
        E = d['conn'].shape[0]
        for key, mat in (('KAB', 'KA_beta'), ('KAG', 'KA_gamma'), ('CA', 'CA')):
            r = torch.zeros(144 * E, dtype=torch.int64, device=DEV)
            c = torch.zeros_like(r)
            v = torch.zeros(144 * E, dtype=torch.float64, device=DEV)
            A.shell_piston(etype, mat, T(d['x'], torch.float64), T(6 * d['conn'], torch.int64),
                           r, c, v, init_k=T(144 * d['perm'], torch.int64))
            torch.cuda.synchronize()
            if not ((r.cpu().numpy() == d[key + 'r']).all() and
                    (c.cpu().numpy() == d[key + 'c']).all()):
                bad.append((name, key, 'indices'))
            err = element_errors(v.cpu().numpy(), d[key + 'v'], 144, d['perm'])
            print(etype, name, key, 'max rel err %.2e' % err)
            if not err <= TOL_CASE.get((etype, name), TOL):
                bad.append((name, key, err))
    x, quads, rng = _random_mesh(23, 19, 8)
    E = quads.shape[0]
    for key, mat in (('KAB', 'KA_beta'), ('KAG', 'KA_gamma'), ('CA', 'CA')):
        r = torch.zeros(144 * E, dtype=torch.int64, device=DEV)
        c = torch.zeros_like(r)
        v = torch.zeros(144 * E, dtype=torch.float64, device=DEV)
        A.shell_piston(etype, mat, T(x, torch.float64), T(6 * quads, torch.int64), r, c, v)
        torch.cuda.synchronize()
        rr = np.zeros(144 * E, np.int64); cc = np.zeros_like(rr); vv = np.zeros(144 * E)
        O.assemble(etype, key, x, quads, 144 * np.arange(E), np.zeros((E, 3)),
                   np.zeros(E, np.int64), np.zeros((1, 32)), rr, cc, vv)
        assert (r.cpu().numpy() == rr).all() and (c.cpu().numpy() == cc).all(), key
        err = element_errors(v.cpu().numpy(), vv, 144, np.arange(E))
        print(etype, 'random', key, 'max rel err %.2e' % err)
        if not err <= TOL:
            bad.append(('random', key, err))
    assert not bad, bad


def test_nonfinite_elements_flags_degenerate_elements():
    """A plate with three collapsed quads (two nodes moved onto their
    neighbours: zero area, inf/nan in the reference's arithmetic) and a
    permuted init_k: the validation pass counts exactly those three, and
    none on the intact plate."""
    from pyfe3d_amd import assembly as A
    from pyfe3d_amd.mesh import PlateMesh
    from pyfe3d_amd.shellprop import isotropic_plate
    m = PlateMesh(11, 9, 1.0, 0.7)
    rec = isotropic_plate(thickness=0.005, E=70e9, nu=0.33).record()
    E = m.quads.shape[0]
    perm = np.random.default_rng(3).permutation(E)
    props = T(rec[None], torch.float64)
    conn = T(6 * m.quads, torch.int64)
    init_k = T(576 * perm, torch.int64)

    def count(x):
        v = torch.zeros(576 * E, dtype=torch.float64, device=DEV)
        A.shell_kc0('quad4', T(x, torch.float64), conn, None, None, v, props, init_k=init_k,
                    values_only=True)
        return A.nonfinite_elements(v, E, 576, init_k=init_k)

    assert count(m.x) == 0
    x = m.x.copy().reshape(-1, 3)
    bad = [5, 31, 70]
    for e in bad:
        n0, n1, n2, n3 = m.quads[e]
        x[n1] = x[n0]           # collapse the element to a line
        x[n2] = x[n3]
    # the oracle (the reference's arithmetic) decides which elements write
    # non-finite values: the collapsed ones, not their distorted neighbours
    r = np.zeros(576 * E, np.int64); c = np.zeros_like(r); v = np.zeros(576 * E)
    with np.errstate(all='ignore'):
        O.assemble('quad4', 'KC0', x.ravel(), m.quads, 576 * perm, np.zeros((E, 3)),
                   np.zeros(E, np.int64), rec[None], r, c, v)
    expect = sorted(set(np.where(~np.isfinite(v))[0] // 576))
    assert sorted(perm[bad]) == expect
    assert count(x.ravel()) == len(bad)


@pytest.mark.parametrize('etype', ['quad4', 'quad4r'])
def test_stream_paired_stores_match_slot_path(etype):
    """KG (from u and given stress), M (mtype 0, 1, 2) and piston theory
    with contiguous slot blocks (init_k NULL: 16-byte paired stores) give
    the same bits as with init_k = SIZE * e (one store per slot), in every
    mode: accumulate, overwrite, values only."""
    from pyfe3d_amd import assembly as A
    x, conn, rng = _random_mesh(23, 17, 9)
    E = conn.shape[0]
    xt, ct = T(x, torch.float64), T(6 * conn, torch.int64)
    rec = T(O.props_table(np.random.default_rng(2).uniform(1, 2, (1, 28)) * 1e6, etype),
            torch.float64)
    u = T(rng.uniform(-1, 1, x.size * 2) * 1e-3, torch.float64)
    stress = T([-1e3, 300., 75.], torch.float64)
    cases = [('KG', 0), ('KGs', 0), ('M', 0), ('M', 1), ('M', 2), ('KA_beta', 0),
             ('KA_gamma', 0), ('CA', 0)]
    for mat, mtype in cases:
        size = A.slot_count(etype, 'KG' if mat == 'KGs' else mat, mtype)
        for overwrite, values_only in ((False, False), (True, False), (True, True)):
            outs = []
            for init_k in (None, T(size * np.arange(E), torch.int64)):
                r = torch.zeros(size * E, dtype=torch.int64, device=DEV)
                c = torch.zeros_like(r)
                v = torch.full((size * E,), 0.5, dtype=torch.float64, device=DEV)
                kw = dict(init_k=init_k, overwrite=overwrite, values_only=values_only)
                if mat == 'KG':
                    A.shell_kg(etype, xt, ct, u, r, c, v, rec, **kw)
                elif mat == 'KGs':
                    A.shell_kg_stress(etype, xt, ct, stress, r, c, v, **kw)
                elif mat == 'M':
                    A.shell_m(etype, xt, ct, r, c, v, rec, mtype=mtype, **kw)
                else:
                    A.shell_piston(etype, mat, xt, ct, r, c, v, **kw)
                outs.append((r, c, v))
            (r0, c0, v0), (r1, c1, v1) = outs
            assert torch.equal(v0, v1) and torch.equal(r0, r1) and torch.equal(c0, c1), \
                (mat, mtype, overwrite, values_only)

"""The CPU oracle of the warm-start network, pinned by central finite
differences (float64) and by the adjoint identities of its building blocks.
The network has no reference counterpart, so these checks are what pin it."""
import numpy as np

from oracle import nn_oracle as N


def _setup(seed=0, B=2, n=16):
    rng = np.random.default_rng(seed)
    p = N.init_params(seed)
    # non-zero biases so that the bias gradients are exercised
    for k in p:
        if k.endswith('.b'):
            p[k] = rng.normal(size=p[k].shape) * 0.1
    x = rng.normal(size=(B, n, n, 2))
    t = rng.normal(size=(B, n + 1, n + 1, 6))
    wc = rng.uniform(0.5, 2.0, 6)
    return p, x, t, wc


def test_adjoints():
    rng = np.random.default_rng(1)
    for s in (1, 2):
        x = rng.normal(size=(2, 8, 8, 3))
        col = N.im2col(x, s)
        y = rng.normal(size=col.shape)
        assert np.isclose((col * y).sum(), (x * N.col2im(y, x.shape, s)).sum(), rtol=1e-12)
    lo = rng.normal(size=(2, 4, 4, 3))
    d = rng.normal(size=(2, 8, 8, 3))
    assert np.isclose((N.up2(lo, 8, 8) * d).sum(), (lo * N.up2_bwd(d, lo.shape)).sum(),
                      rtol=1e-12)
    e = rng.normal(size=(2, 6, 6, 3))
    du = rng.normal(size=(2, 7, 7, 3))
    assert np.isclose((N.node_avg(e) * du).sum(), (e * N.node_avg_bwd(du)).sum(), rtol=1e-12)


def test_im2col_is_a_convolution():
    """im2col + GEMM equals a direct 3x3 convolution sum (zero padding)."""
    rng = np.random.default_rng(2)
    x = rng.normal(size=(1, 5, 6, 2))
    w = rng.normal(size=(3, 9 * 2))
    for s in (1, 2):
        y = (N.im2col(x, s) @ w.T).reshape(1, (5 - 1) // s + 1, (6 - 1) // s + 1, 3)
        xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
        W4 = w.reshape(3, 3, 3, 2)
        for oy in range(y.shape[1]):
            for ox in range(y.shape[2]):
                ref = np.einsum('abc,oabc->o', xp[0, oy * s:oy * s + 3, ox * s:ox * s + 3], W4)
                assert np.allclose(y[0, oy, ox], ref, rtol=1e-12, atol=1e-12)


def test_gradients_match_finite_differences():
    p, x, t, wc = _setup()
    u, cache = N.forward(p, x)
    g = N.backward(p, cache, u, t, wc)
    rng = np.random.default_rng(3)
    h = 1e-6
    worst = 0.
    for name in sorted(p):
        flat = p[name].reshape(-1)
        for idx in rng.choice(flat.size, size=min(4, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp = N.loss(N.forward(p, x)[0], t, wc)
            flat[idx] = old - h
            lm = N.loss(N.forward(p, x)[0], t, wc)
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            an = g[name].reshape(-1)[idx]
            err = abs(fd - an) / max(1e-8, abs(fd) + abs(an))
            worst = max(worst, err)
            assert err < 1e-5, (name, idx, fd, an)
    print('worst relative gradient error %.2e' % worst)


def _energy_setup(seed=4, B=2, n=8):
    """Small clamped plates with random thickness fields: K (boundary
    conditions applied), loads and solution energies from the element
    oracle and scipy."""
    import scipy.sparse as ss
    import scipy.sparse.linalg as sla
    from oracle import oracle as O
    from pyfe3d_amd.mesh import PlateMesh
    from pyfe3d_amd.shellprop import isotropic_records
    rng = np.random.default_rng(seed)
    m = PlateMesh(n + 1, n + 1, 1.0, 1.0)
    E = m.quads.shape[0]
    fixed = m.clamped()
    P = ss.diags((~fixed).astype(float))
    Ks, fs, ss_, fx, us = [], [], [], [], []
    for b in range(B):
        rec = isotropic_records(rng.uniform(0.002, 0.004, E), 70e9, 0.33)
        r = np.zeros(576 * E, np.int64); c = np.zeros_like(r); v = np.zeros(576 * E)
        O.assemble('quad4', 'KC0', m.x, m.quads, 576 * np.arange(E), np.zeros((E, 3)),
                   np.arange(E), rec, r, c, v)
        K = ss.coo_matrix((v, (r, c)), shape=(m.n_dof, m.n_dof)).tocsr()
        K = (P @ K @ P + ss.diags(fixed.astype(float))).tocsr()
        f = np.zeros(m.n_dof)
        f[2::6] = -rng.uniform(0.5, 1.5, m.n_nodes)
        f[fixed] = 0.
        u = sla.spsolve(K.tocsc(), f)
        Ks.append(K); fs.append(f); ss_.append(float(u @ f)); fx.append(fixed); us.append(u)
    return Ks, fs, np.array(ss_), np.array(fx), us


def test_energy_loss_gradient_and_optimum():
    """The energy loss of the oracle: its node-value gradient against
    central differences, its parameter gradients (through backward_from)
    against central differences, L = 0 at the exact solution (any
    positive multiple) and 0 <= L <= 1 otherwise."""
    Ks, fs, ss_, fixed, us = _energy_setup()
    B, n, C = 2, 8, 6
    rng = np.random.default_rng(5)
    scale = np.array([0., 0., 1., 30., 30., 0.])
    # guesses near the solutions (L well inside (0, 1): a random guess has
    # L = 1 - O(1e-8), where differences of L cancel)
    ex = np.stack([u.reshape(n + 1, n + 1, C) / np.where(scale > 0, scale, 1.) for u in us])
    u = ex * (1 + 0.2 * rng.normal(size=ex.shape))
    L, du, _ = N.energy_loss(u, scale, fixed, Ks, fs, ss_)
    assert 0.05 < L < 1.
    for idx in [(0, 3, 4, 2), (1, 5, 2, 3), (0, 2, 6, 4)]:
        h = 1e-5 * abs(u[idx])
        up = u.copy(); up[idx] += h
        um = u.copy(); um[idx] -= h
        fd = (N.energy_loss(up, scale, fixed, Ks, fs, ss_)[0] -
              N.energy_loss(um, scale, fixed, Ks, fs, ss_)[0]) / (2 * h)
        assert abs(fd - du[idx]) <= 1e-6 * max(1e-8, abs(fd)), (idx, fd, du[idx])
    # the exact solutions, any multiple, have L = 0
    assert N.energy_loss(ex * 3.7, scale, fixed, Ks, fs, ss_)[0] < 1e-12
    assert N.energy_loss(-ex, scale, fixed, Ks, fs, ss_)[0] < 1e-12
    # parameter gradients through the network (the loss of a random
    # network is 1 - O(1e-8): differentiate around it with scales that make
    # its output comparable to the solutions)
    p = N.init_params(2)
    x = rng.normal(size=(B, n, n, 2))
    uu, cache = N.forward(p, x)
    L0, du, _ = N.energy_loss(uu, scale, fixed, Ks, fs, ss_)
    assert 0. <= L0 <= 1.
    g = N.backward_from(p, cache, du)
    h = 1e-6
    for name in ('c0.w', 'd3.w', 'head.b', 'u1.w'):
        flat = p[name].reshape(-1)
        for k in rng.choice(flat.size, 3, replace=False):
            old = flat[k]
            flat[k] = old + h
            # d loss = -d captured (the captured fraction is ~1e-8 here and
            # carries all its digits; the loss does not)
            lp = N.energy_loss(N.forward(p, x)[0], scale, fixed, Ks, fs, ss_)[2]
            flat[k] = old - h
            lm = N.energy_loss(N.forward(p, x)[0], scale, fixed, Ks, fs, ss_)[2]
            flat[k] = old
            fd = -(lp - lm) / (2 * h)
            an = g[name].reshape(-1)[k]
            assert abs(fd - an) <= 1e-5 * max(1e-9, abs(fd) + abs(an)), (name, k, fd, an)

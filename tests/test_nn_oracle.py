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

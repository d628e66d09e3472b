"""CPU oracle of the warm-start network (test infrastructure only).

A plain numpy restatement, in float64, of the encoder-decoder that
``pyfe3d_amd.warmstart`` runs with its own HIP kernels: same layer list,
same NHWC activations, same weight layout ``W[c_out][(ky, kx, c_in)]``,
same element-to-node averaging and loss.  Tests compare the GPU forward
pass and every parameter gradient with this restatement, and pin the
restatement itself against central finite differences (the network has no
counterpart in the reference, so there are no golden vectors for it).

Only tests/ and bench.py's CPU baseline may import this module.
"""
import numpy as np

# (name, c_in, c_out, stride, kernel, relu); the skip/upsample wiring is in
# forward() below
LAYERS = [
    ('c0', 2, 32, 1, 3, True),
    ('c1', 32, 32, 1, 3, True),
    ('d1', 32, 64, 2, 3, True),
    ('d2', 64, 128, 2, 3, True),
    ('d3', 128, 128, 2, 3, True),
    ('u3', 128, 64, 1, 3, True),
    ('u2', 64, 32, 1, 3, True),
    ('u1', 32, 32, 1, 3, True),
    ('head', 32, 6, 1, 1, False),
]


def init_params(seed=0, dtype=np.float64):
    """He-normal weights, zero biases."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, ci, co, _, k, _ in LAYERS:
        fan = k * k * ci
        p[name + '.w'] = (rng.standard_normal((co, fan)) * np.sqrt(2.0 / fan)).astype(dtype)
        p[name + '.b'] = np.zeros(co, dtype)
    return p


def im2col(x, s):
    """x (B, H, W, C) -> (B*Ho*Wo, 9*C), (ky, kx, c) fastest, zero padding 1."""
    B, H, W, C = x.shape
    Ho, Wo = (H - 1) // s + 1, (W - 1) // s + 1
    xp = np.zeros((B, H + 2, W + 2, C), x.dtype)
    xp[:, 1:-1, 1:-1] = x
    cols = np.empty((B, Ho, Wo, 3, 3, C), x.dtype)
    for ky in range(3):
        for kx in range(3):
            cols[:, :, :, ky, kx] = xp[:, ky:ky + s * Ho:s, kx:kx + s * Wo:s]
    return cols.reshape(B * Ho * Wo, 9 * C)


def col2im(dcol, shape, s):
    """Adjoint of im2col."""
    B, H, W, C = shape
    Ho, Wo = (H - 1) // s + 1, (W - 1) // s + 1
    d = dcol.reshape(B, Ho, Wo, 3, 3, C)
    dxp = np.zeros((B, H + 2, W + 2, C), dcol.dtype)
    for ky in range(3):
        for kx in range(3):
            dxp[:, ky:ky + s * Ho:s, kx:kx + s * Wo:s] += d[:, :, :, ky, kx]
    return dxp[:, 1:-1, 1:-1]


def up2(x, H, W):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)[:, :H, :W]


def up2_bwd(d, shape_lo):
    B, Hl, Wl, C = shape_lo
    out = np.zeros(shape_lo, d.dtype)
    H, W = d.shape[1], d.shape[2]
    for dy in range(2):
        for dx in range(2):
            part = d[:, dy::2, dx::2]
            out[:, :part.shape[1], :part.shape[2]] += part
    return out


def node_avg(e):
    """element outputs (B, n, n, C) -> node values (B, n+1, n+1, C), boundary 0."""
    B, n, _, C = e.shape
    u = np.zeros((B, n + 1, n + 1, C), e.dtype)
    u[:, 1:n, 1:n] = 0.25 * (e[:, :-1, :-1] + e[:, :-1, 1:] + e[:, 1:, :-1] + e[:, 1:, 1:])
    return u


def node_avg_bwd(du):
    B, nn, _, C = du.shape
    n = nn - 1
    g = np.zeros((B, n, n, C), du.dtype)
    inner = du[:, 1:n, 1:n] * 0.25
    g[:, :-1, :-1] += inner
    g[:, :-1, 1:] += inner
    g[:, 1:, :-1] += inner
    g[:, 1:, 1:] += inner
    return g


def _conv(p, name, x, s, k, relu):
    B, H, W, C = x.shape
    if k == 3:
        col = im2col(x, s)
        Ho, Wo = (H - 1) // s + 1, (W - 1) // s + 1
    else:
        col = x.reshape(B * H * W, C)
        Ho, Wo = H, W
    y = col @ p[name + '.w'].T + p[name + '.b']
    if relu:
        y = np.maximum(y, 0.)
    return y.reshape(B, Ho, Wo, -1), col


def forward(p, x):
    """x (B, 64, 64, 2) -> (node values (B, 65, 65, 6), cache)."""
    spec = {n: (s, k, r) for n, _, _, s, k, r in LAYERS}
    cache = {}

    def conv(name, inp):
        s, k, r = spec[name]
        y, col = _conv(p, name, inp, s, k, r)
        cache[name] = (inp.shape, col, y)
        return y

    a0 = conv('c0', x)
    a1 = conv('c1', a0)
    a2 = conv('d1', a1)
    a3 = conv('d2', a2)
    a4 = conv('d3', a3)
    z3 = up2(a4, a3.shape[1], a3.shape[2]) + a3
    a5 = conv('u3', z3)
    z2 = up2(a5, a2.shape[1], a2.shape[2]) + a2
    a6 = conv('u2', z2)
    z1 = up2(a6, a1.shape[1], a1.shape[2]) + a1
    a7 = conv('u1', z1)
    e = conv('head', a7)
    cache['shapes'] = dict(a4=a4.shape, a5=a5.shape, a6=a6.shape)
    return node_avg(e), cache


def loss(u, target, wc):
    """0.5 * mean over (sample, node) of sum_c wc[c] (u - t)^2 (interior nodes)."""
    B, nn, _, C = u.shape
    count = B * (nn - 2) ** 2
    d = (u - target)[:, 1:-1, 1:-1]
    return 0.5 * float((d * d * wc).sum()) / count


def energy_loss(u, scale, fixed, Ks, fs, ss):
    """Mean over the plates of the relative energy-norm error of the best
    multiple of the guess: x_b = u_b * scale (0 on fixed DOFs), a = x.f,
    c = x.K x, L_b = 1 - a^2 / (c s_b) with s_b = u*.f the solution's
    energy.  Ks: per-plate stiffness (boundary conditions applied, scipy or
    dense), fs: loads, ss: energies.  Returns (loss, du (B, nn, nn, C),
    captured) with captured = mean of a^2 / (c s) = 1 - loss, accurate where
    the loss is within rounding of 1."""
    B, nn, _, C = u.shape
    du = np.zeros_like(u)
    cap = 0.
    for b in range(B):
        x = (u[b] * scale).reshape(-1)
        x[fixed[b]] = 0.
        kx = Ks[b] @ x
        a, c = float(x @ fs[b]), float(x @ kx)
        cap += a * a / (c * ss[b])
        g = 2. * a / (c * ss[b]) * (a / c * kx - fs[b])
        g[fixed[b]] = 0.
        du[b] = g.reshape(nn, nn, C) * scale / B
    return 1. - cap / B, du, cap / B


def backward(p, cache, u, target, wc):
    """Gradients of loss() with respect to every parameter."""
    B, nn, _, C = u.shape
    count = B * (nn - 2) ** 2
    du = np.zeros_like(u)
    du[:, 1:-1, 1:-1] = (u - target)[:, 1:-1, 1:-1] * wc / count
    return backward_from(p, cache, du)


def backward_from(p, cache, du):
    """Parameter gradients given the gradient du of a loss with respect to
    the node values (B, n+1, n+1, C) of forward()."""
    spec = {n: (s, k, r) for n, _, _, s, k, r in LAYERS}
    g = {}

    def conv_bwd(name, dy):
        s, k, r = spec[name]
        in_shape, col, y = cache[name]
        dy = dy.reshape(-1, dy.shape[-1])
        if r:
            dy = dy * (y.reshape(-1, y.shape[-1]) > 0)
        g[name + '.w'] = dy.T @ col
        g[name + '.b'] = dy.sum(axis=0)
        dcol = dy @ p[name + '.w']
        if k == 3:
            return col2im(dcol, in_shape, s)
        return dcol.reshape(in_shape)

    de = node_avg_bwd(du)
    da7 = conv_bwd('head', de)
    dz1 = conv_bwd('u1', da7)
    da6 = up2_bwd(dz1, cache['shapes']['a6'])
    da1 = dz1.copy()
    dz2 = conv_bwd('u2', da6)
    da5 = up2_bwd(dz2, cache['shapes']['a5'])
    da2 = dz2.copy()
    dz3 = conv_bwd('u3', da5)
    da4 = up2_bwd(dz3, cache['shapes']['a4'])
    da3 = dz3.copy()
    da3 += conv_bwd('d3', da4)
    da2 += conv_bwd('d2', da3)
    da1 += conv_bwd('d1', da2)
    da0 = conv_bwd('c1', da1)
    conv_bwd('c0', da0)
    return g

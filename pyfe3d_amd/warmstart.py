"""Learned initial guess for the static solve of plates (warm start).

A small convolutional encoder-decoder (about 0.35M parameters, 4 levels,
32-128 channels) maps a plate's per-element thickness field and pressure
map (64 x 64 elements) to the shape of its deflection on the 65 x 65 node
grid.  The rotations are taken from that deflection (Kirchhoff: rx = w,y,
ry = -w,x), the amplitude from a one-dimensional Galerkin projection
(alpha = u0.f / u0.K u0, one sparse product), and the result is the PCG's
x0; a plate whose alpha is not positive starts from zero.  The solve
itself, and its FP64 residual criterion, are unchanged: the network can only
change how many iterations are needed.

Everything runs through the library's own HIP kernels (im2col / col2im,
bfloat16 GEMM on the matrix cores with FP32 accumulation, up-sampling,
losses and Adam); torch only holds the buffers.  Training data are solves
of the plain GPU solver (:mod:`pyfe3d_amd.sweep`).
"""
import numpy as np
import torch

from . import _lib

# (name, c_in, c_out, stride, kernel, relu).  Wiring: c0 -> c1 (skip1) ->
# d1 (skip2) -> d2 (skip3) -> d3; u3 = conv(up(d3) + skip3),
# u2 = conv(up(u3) + skip2), u1 = conv(up(u2) + skip1), head = 1x1 conv.
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
SPEC = {l[0]: l[1:] for l in LAYERS}


def _s(stream):
    return (stream if stream is not None else torch.cuda.current_stream()).cuda_stream


def n_params():
    return sum(co * k * k * ci + co for _, ci, co, _, k, _ in LAYERS)


class WarmStartNet:
    """Parameters, buffers and the forward/backward/Adam passes for a grid
    of n x n elements (n divisible by 8) and batches of up to ``batch``."""

    def __init__(self, n=64, batch=32, seed=0, device='cuda'):
        if n % 8:
            raise ValueError('n must be divisible by 8')
        self.n, self.batch, self.dev = n, batch, torch.device(device)
        rng = np.random.default_rng(seed)
        self.p, self.g, self.m, self.v = {}, {}, {}, {}
        for name, ci, co, _, k, _ in LAYERS:
            fan = k * k * ci
            w = rng.standard_normal((co, fan)) * np.sqrt(2.0 / fan)
            self.p[name + '.w'] = torch.as_tensor(w, dtype=torch.float32, device=self.dev)
            self.p[name + '.b'] = torch.zeros(co, dtype=torch.float32, device=self.dev)
        for k, t in self.p.items():
            self.g[k] = torch.zeros_like(t)
            self.m[k] = torch.zeros_like(t)
            self.v[k] = torch.zeros_like(t)
        self.step_count = 0
        self._alloc(batch)

    # ------------------------------------------------------------ buffers
    def _shape(self, name):
        """(H_in, H_out) of a layer on the n x n grid."""
        n = self.n
        return {'c0': (n, n), 'c1': (n, n), 'd1': (n, n // 2), 'd2': (n // 2, n // 4),
                'd3': (n // 4, n // 8), 'u3': (n // 4, n // 4), 'u2': (n // 2, n // 2),
                'u1': (n, n), 'head': (n, n)}[name]

    def _alloc(self, B):
        f = dict(dtype=torch.float32, device=self.dev)
        n = self.n
        self.B = B
        self.act, self.col = {}, {}
        self.dcol_size = 0
        for name, ci, co, s, k, _ in LAYERS:
            hi, ho = self._shape(name)
            self.act[name] = torch.empty((B, ho, ho, co), **f)
            if k == 3:
                self.col[name] = torch.empty((B * ho * ho, 9 * ci), **f)
                self.dcol_size = max(self.dcol_size, B * ho * ho * 9 * ci)
            else:
                self.dcol_size = max(self.dcol_size, B * ho * ho * ci)
        self.z = {3: torch.empty((B, n // 4, n // 4, 128), **f),
                  2: torch.empty((B, n // 2, n // 2, 64), **f),
                  1: torch.empty((B, n, n, 32), **f)}
        self.u = torch.empty((B, n + 1, n + 1, 6), **f)
        self.dcol = torch.empty(self.dcol_size, **f)
        # gradients of activations (same shapes), split-K partials
        self.dact = {k: torch.empty_like(t) for k, t in self.act.items()}
        self.dz = {k: torch.empty_like(t) for k, t in self.z.items()}
        self.splits = 64
        maxw = max(co * k * k * ci for _, ci, co, _, k, _ in LAYERS)
        self.partial = torch.empty(self.splits * maxw, **f)

    # -------------------------------------------------------------- pieces
    def _gemm(self, M, N, K, A, sam, sak, Bm, sbn, sbk, C, ldc, acc=0, bias=None, relu=0,
              splits=1, stream=None):
        _lib.call('pf3d_nn_gemm', M, N, K, A.data_ptr(), sam, sak, Bm.data_ptr(), sbn, sbk,
                  C.data_ptr(), ldc, acc, None if bias is None else bias.data_ptr(), relu,
                  splits, self.partial.data_ptr() if splits > 1 else None, _s(stream))

    def _conv(self, name, x, stream=None):
        ci, co, s, k, relu = SPEC[name]
        B = self.B
        hi, ho = self._shape(name)
        y = self.act[name]
        M = B * ho * ho
        if k == 3:
            col = self.col[name]
            _lib.call('pf3d_nn_im2col', B, hi, hi, ci, s, x.data_ptr(), col.data_ptr(),
                      _s(stream))
        else:
            col = x
        Kd = k * k * ci
        self._gemm(M, co, Kd, col, Kd, 1, self.p[name + '.w'], Kd, 1, y, co,
                   bias=self.p[name + '.b'], relu=int(relu), stream=stream)
        return y

    def _conv_bwd(self, name, x, dy, dx=None, dx_acc=0, stream=None):
        """dy: gradient of the layer output (masked here for ReLU);
        writes the weight/bias gradients and, if dx is given, (+)= dx."""
        ci, co, s, k, relu = SPEC[name]
        B = self.B
        hi, ho = self._shape(name)
        M = B * ho * ho
        Kd = k * k * ci
        y = self.act[name]
        st = _s(stream)
        if relu:
            _lib.call('pf3d_nn_relu_bwd', dy.numel(), y.data_ptr(), dy.data_ptr(), st)
        col = self.col[name] if k == 3 else x
        # dW[co][Kd] = sum_m dy[m][co] col[m][Kd]  (split-K over the pixels)
        self._gemm(co, Kd, M, dy, 1, co, col, 1, Kd, self.g[name + '.w'], Kd,
                   splits=min(self.splits, max(1, M // 256)), stream=stream)
        _lib.call('pf3d_nn_bias_grad', M, co, dy.data_ptr(), self.g[name + '.b'].data_ptr(), st)
        if dx is None:
            return
        if k == 3:
            dcol = self.dcol[:M * Kd]
            self._gemm(M, Kd, co, dy, co, 1, self.p[name + '.w'], 1, Kd, dcol, Kd,
                       stream=stream)
            _lib.call('pf3d_nn_col2im', B, hi, hi, ci, s, dcol.data_ptr(), dx.data_ptr(),
                      int(dx_acc), st)
        else:
            self._gemm(M, Kd, co, dy, co, 1, self.p[name + '.w'], 1, Kd, dx, Kd,
                       acc=int(dx_acc), stream=stream)

    # ------------------------------------------------------------- passes
    def forward(self, x, stream=None):
        """x: (B, n, n, 2) float32 contiguous -> node values (B, n+1, n+1, 6)."""
        if x.shape[0] != self.B:
            self._alloc(x.shape[0])
        x = x.contiguous()
        self._x = x
        n, B, st = self.n, self.B, _s(stream)
        a0 = self._conv('c0', x, stream)
        a1 = self._conv('c1', a0, stream)
        a2 = self._conv('d1', a1, stream)
        a3 = self._conv('d2', a2, stream)
        a4 = self._conv('d3', a3, stream)
        _lib.call('pf3d_nn_up2_add', B, n // 4, n // 4, 128, a4.data_ptr(), a3.data_ptr(),
                  self.z[3].data_ptr(), st)
        a5 = self._conv('u3', self.z[3], stream)
        _lib.call('pf3d_nn_up2_add', B, n // 2, n // 2, 64, a5.data_ptr(), a2.data_ptr(),
                  self.z[2].data_ptr(), st)
        a6 = self._conv('u2', self.z[2], stream)
        _lib.call('pf3d_nn_up2_add', B, n, n, 32, a6.data_ptr(), a1.data_ptr(),
                  self.z[1].data_ptr(), st)
        a7 = self._conv('u1', self.z[1], stream)
        e = self._conv('head', a7, stream)
        _lib.call('pf3d_nn_node_avg', B, n, 6, e.data_ptr(), self.u.data_ptr(), st)
        return self.u

    def backward(self, target, wc, stream=None):
        """Parameter gradients of 0.5 * mean over (sample, interior node) of
        sum_c wc[c] (u - target)^2 after forward()."""
        n, B, st = self.n, self.B, _s(stream)
        count = B * (n - 1) ** 2
        de = self.dact['head']
        _lib.call('pf3d_nn_node_loss_grad', B, n, 6, self.u.data_ptr(),
                  target.contiguous().data_ptr(), wc.data_ptr(), 1.0 / count, de.data_ptr(), st)
        A, D = self.act, self.dact
        self._conv_bwd('head', A['u1'], de, D['u1'], stream=stream)
        self._conv_bwd('u1', self.z[1], D['u1'], self.dz[1], stream=stream)
        # z1 = up(a6) + a1
        _lib.call('pf3d_nn_up2_bwd', B, n, n, 32, self.dz[1].data_ptr(), D['u2'].data_ptr(), st)
        D['c1'].copy_(self.dz[1])
        self._conv_bwd('u2', self.z[2], D['u2'], self.dz[2], stream=stream)
        _lib.call('pf3d_nn_up2_bwd', B, n // 2, n // 2, 64, self.dz[2].data_ptr(),
                  D['u3'].data_ptr(), st)
        D['d1'].copy_(self.dz[2])
        self._conv_bwd('u3', self.z[3], D['u3'], self.dz[3], stream=stream)
        _lib.call('pf3d_nn_up2_bwd', B, n // 4, n // 4, 128, self.dz[3].data_ptr(),
                  D['d3'].data_ptr(), st)
        D['d2'].copy_(self.dz[3])
        self._conv_bwd('d3', A['d2'], D['d3'], D['d2'], dx_acc=1, stream=stream)
        self._conv_bwd('d2', A['d1'], D['d2'], D['d1'], dx_acc=1, stream=stream)
        self._conv_bwd('d1', A['c1'], D['d1'], D['c1'], dx_acc=1, stream=stream)
        self._conv_bwd('c1', A['c0'], D['c1'], D['c0'], stream=stream)
        self._conv_bwd('c0', self._x, D['c0'], None, stream=stream)

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, stream=None):
        self.step_count += 1
        for k, p in self.p.items():
            _lib.call('pf3d_nn_adam', p.numel(), p.data_ptr(), self.g[k].data_ptr(),
                      self.m[k].data_ptr(), self.v[k].data_ptr(), lr, b1, b2, eps,
                      self.step_count, _s(stream))

    def params_numpy(self):
        return {k: t.detach().cpu().numpy().astype(np.float64) for k, t in self.p.items()}

    def grads_numpy(self):
        return {k: t.detach().cpu().numpy().astype(np.float64) for k, t in self.g.items()}


# ----------------------------------------------------------------------
# data: plates with smooth thickness fields under Gaussian pressure patches
# ----------------------------------------------------------------------

def pressure_map(nxe, nye, a, b, rng, p0=1e3):
    """Seeded sum of K ~ U{1..4} Gaussian patches on the element centroids."""
    xc = (np.arange(nxe) + 0.5) * a / nxe
    yc = (np.arange(nye) + 0.5) * b / nye
    X, Y = np.meshgrid(xc, yc, indexing='ij')
    p = np.zeros((nxe, nye))
    for _ in range(rng.integers(1, 5)):
        mx, my = rng.uniform(0.1 * a, 0.9 * a), rng.uniform(0.1 * b, 0.9 * b)
        sg = rng.uniform(0.05, 0.2) * a
        p += rng.uniform(0.5, 1.5) * p0 * np.exp(-((X - mx) ** 2 + (Y - my) ** 2) / (2 * sg ** 2))
    return p


def nodal_pressure_load(p, a, b):
    """Consistent nodal forces of a per-element constant pressure on a
    regular nxe x nye grid: f_z(node) = sum over its elements of p A / 4.
    Returns (nxe+1, nye+1) forces (positive along -z for p > 0)."""
    nxe, nye = p.shape
    q = p * (a / nxe) * (b / nye) / 4.
    f = np.zeros((nxe + 1, nye + 1))
    f[:-1, :-1] += q
    f[1:, :-1] += q
    f[:-1, 1:] += q
    f[1:, 1:] += q
    return -f


def features(h, p, h0=0.002, dh=0.001):
    """Network input (B, n, n, 2): (h - h0)/dh and p / max|p| per sample;
    also returns the per-sample load scale max|p|."""
    ps = np.abs(p).reshape(p.shape[0], -1).max(axis=1)
    ps = np.where(ps > 0, ps, 1.)
    x = np.stack([(h - h0) / dh, p / ps[:, None, None]], axis=-1).astype(np.float32)
    return x, ps


# ----------------------------------------------------------------------
# dataset, training, use as the PCG's initial guess
# ----------------------------------------------------------------------

class Scales:
    """Target normalisation.  The network learns the shape of the
    deflection w only: each training target is w / rms(w) of its plate
    (then divided by the dataset std), because the amplitude comes from the
    Galerkin projection at solve time and the rotations from w (Kirchhoff:
    for thin plates independent rotation guesses carry shear energy that
    is (L/h)^2 times their bending energy)."""

    def __init__(self, std=None):
        self.std = np.ones(6) if std is None else np.asarray(std, dtype=np.float64)


W_CHANNEL = 2  # w among (u, v, w, rx, ry, rz)


def _plates(n_var, n, seed, a=1.0, b=1.0):
    """Seeded designs: (h (V, n, n), p (V, n, n), f_z (V, n+1, n+1))."""
    from .sweep import smooth_field
    rng = np.random.default_rng(seed)
    h = np.empty((n_var, n, n))
    p = np.empty((n_var, n, n))
    f = np.empty((n_var, n + 1, n + 1))
    for v in range(n_var):
        h[v] = np.maximum(0.002 + 0.001 * smooth_field(n, n, a, b, 0.2 * a, rng), 0.0002)
        p[v] = pressure_map(n, n, a, b, rng)
        f[v] = nodal_pressure_load(p[v], a, b)
    return h, p, f


def solve_plates(h, f, mode='fp64', rtol=1e-8, sweep=None, x0=None):
    """Solve every plate (clamped edges) with the batched GPU solver.
    Returns (u (V, n+1, n+1, 6) float64 tensor, sweep)."""
    from .sweep import PlateSweep
    V, n = h.shape[0], h.shape[1]
    if sweep is None or sweep.n_var != V:
        sweep = PlateSweep([(n, n)] * V, a=1.0, b=1.0)
    sweep.set_thickness(h.reshape(-1))
    fd = np.zeros((V, (n + 1) * (n + 1), 6))
    fd[:, :, 2] = f.reshape(V, -1)
    sweep.set_loads(fd.reshape(-1))
    if x0 is not None:
        sweep.u.copy_(torch.as_tensor(x0, dtype=torch.float64).reshape(-1))
    sweep.step(mode=mode, rtol=rtol, x0=x0 is not None)
    return sweep.u.view(V, n + 1, n + 1, 6), sweep


def make_dataset(n_samples, seed, n=64, chunk=1024, rtol=1e-6, log=None):
    """Plates, their GPU solutions (to rtol: the network's own error is
    orders of magnitude above it) and the network inputs/targets (float32
    tensors on the GPU; targets normalised with per-channel std)."""
    X, U, PS = [], [], []
    sweep = None
    for c0 in range(0, n_samples, chunk):
        V = min(chunk, n_samples - c0)
        h, p, f = _plates(V, n, seed * 1000003 + c0)
        u, sweep = solve_plates(h, f, rtol=rtol,
                                sweep=sweep if sweep is not None and sweep.n_var == V else None)
        torch.cuda.synchronize()
        sweep.check()
        if float(sweep.relres.max()) > rtol:
            raise RuntimeError('training solve missed the tolerance')
        x, ps = features(h, p)
        X.append(torch.as_tensor(x, device='cuda'))
        U.append(u.to(torch.float32).clone())
        PS.append(ps)
        if log:
            log('dataset %d/%d' % (c0 + V, n_samples))
    X = torch.cat(X)
    U = torch.cat(U)
    w = U[:, 1:-1, 1:-1, W_CHANNEL]
    rms = w.pow(2).mean(dim=(1, 2)).sqrt().clamp_min(1e-30)
    T = U / rms[:, None, None, None]
    std = T[:, 1:-1, 1:-1].reshape(-1, 6).std(dim=0).double().cpu().numpy()
    std = np.where(std > 0, std, 1.)
    sc = Scales(std)
    T = T / torch.as_tensor(std, dtype=torch.float32, device='cuda')
    return X, T.contiguous(), sc


def train(net, X, T, epochs=10, batch=32, lr=1e-3, seed=0, log=None):
    """Adam on the standardised MSE of the deflection channel; returns the
    loss history per epoch."""
    wc = torch.zeros(6, dtype=torch.float32, device='cuda')
    wc[W_CHANNEL] = 1.
    S = X.shape[0]
    g = torch.Generator(device='cpu').manual_seed(seed)
    hist = []
    for ep in range(epochs):
        perm = torch.randperm(S, generator=g)
        tot, nb = 0., 0
        for b0 in range(0, S - batch + 1, batch):
            idx = perm[b0:b0 + batch].to('cuda')
            xb, tb = X[idx].contiguous(), T[idx].contiguous()
            u = net.forward(xb)
            net.backward(tb, wc)
            net.adam(lr=lr * (0.5 ** (ep // max(1, epochs // 3))))
            if b0 % (batch * 16) == 0:
                d = (u - tb)[:, 1:-1, 1:-1, W_CHANNEL]
                tot += 0.5 * float((d * d).sum()) / (batch * (net.n - 1) ** 2)
                nb += 1
        hist.append(tot / max(1, nb))
        if log:
            log('epoch %d loss %.4e' % (ep, hist[-1]))
    return hist


def kirchhoff_rotations(w, a=1.0, b=1.0):
    """(V, n+1, n+1) nodal deflections -> (V, n+1, n+1, 6) DOF vectors with
    rx = w,y and ry = -w,x (zero transverse shear for the element's
    gyz = w,y - rx, gxz = w,x + ry), central differences inside, one-sided
    at the edges; u, v, rz = 0."""
    V, n1 = w.shape[0], w.shape[1]
    dx, dy = a / (n1 - 1), b / (n1 - 1)
    wx = torch.empty_like(w)
    wy = torch.empty_like(w)
    wx[:, 1:-1] = (w[:, 2:] - w[:, :-2]) / (2 * dx)
    wx[:, 0] = (w[:, 1] - w[:, 0]) / dx
    wx[:, -1] = (w[:, -1] - w[:, -2]) / dx
    wy[:, :, 1:-1] = (w[:, :, 2:] - w[:, :, :-2]) / (2 * dy)
    wy[:, :, 0] = (w[:, :, 1] - w[:, :, 0]) / dy
    wy[:, :, -1] = (w[:, :, -1] - w[:, :, -2]) / dy
    u = torch.zeros((V, n1, n1, 6), dtype=w.dtype, device=w.device)
    u[..., 2] = w
    u[..., 3] = wy
    u[..., 4] = -wx
    return u


def predict(net, h, p, sc, batch=64):
    """Initial-guess shapes (V, n+1, n+1, 6) float64 on the GPU: the
    network's deflection with Kirchhoff rotations, before the Galerkin
    scaling (which sets the amplitude)."""
    x, _ = features(h, p)
    V = x.shape[0]
    w = torch.empty((V, net.n + 1, net.n + 1), dtype=torch.float64, device='cuda')
    for b0 in range(0, V, batch):
        xb = torch.as_tensor(x[b0:b0 + batch], device='cuda')
        u = net.forward(xb)
        w[b0:b0 + xb.shape[0]] = u[:xb.shape[0], :, :, W_CHANNEL].double() * sc.std[W_CHANNEL]
    return kirchhoff_rotations(w)


def galerkin_scale(sweep, u0, stats=None):
    """Per plate alpha = u0.f / u0.K u0 (K after the Dirichlet rows and
    columns, u0 with its fixed DOFs zeroed): the multiple of u0 closest to
    the solution in the energy norm, whose error ||u - alpha u0||_K^2 =
    ||u||_K^2 - (u0.f)^2 / u0.K u0 never exceeds that of a zero start (the
    norm in which CG converges).  Plates with alpha <= 0 (or no finite
    alpha) start from zero.  Returns (scaled u0, number of zero starts);
    ``stats`` (a dict), if given, receives the per-plate u0.f and u0.K u0."""
    V = sweep.n_var
    u = u0.reshape(-1).clone()
    u.mul_(1 - sweep.fixed.to(torch.float64))
    Ku = torch.empty_like(u)
    _lib.call('pf3d_bsr_spmv', sweep.n_nodes, sweep.plan.row_ptr.data_ptr(),
              sweep.plan.col_idx.data_ptr(), sweep.val.data_ptr(), u.data_ptr(), Ku.data_ptr(),
              torch.cuda.current_stream().cuda_stream)
    uv, kv, fv = u.view(V, -1), Ku.view(V, -1), sweep.f.view(V, -1)
    uf = (uv * fv).sum(dim=1)
    uku = (uv * kv).sum(dim=1)
    ok = (uku > 0) & (uf > 0) & torch.isfinite(uf) & torch.isfinite(uku)
    alpha = torch.where(ok, uf / torch.where(ok, uku, torch.ones_like(uku)),
                        torch.zeros_like(uku))
    uv.mul_(alpha[:, None])
    if stats is not None:
        stats['uf'], stats['uku'] = uf, uku
    return u, int((~ok).sum())


def save(net, sc, path):
    """Parameters and target scales as an .npz (no pickled objects)."""
    arrs = {k: v.astype(np.float32) for k, v in net.params_numpy().items()}
    arrs['scales.std'] = np.asarray(sc.std)
    np.savez(path, **arrs)


def load(path, batch=64, device='cuda'):
    z = np.load(path, allow_pickle=False)
    net = WarmStartNet(n=64, batch=batch, device=device)
    for k in net.p:
        net.p[k].copy_(torch.as_tensor(z[k], dtype=torch.float32))
    sc = Scales(std=z['scales.std'])
    return net, sc


def evaluate(net, sc, n_eval, seed, max_iter=60000, log=None):
    """Held-out plates: batched solves to 1e-8 from zero and from the
    network's (Galerkin-scaled) guess, FP64 PCG and FP32/FP64 refinement.
    Times are GPU event times of the whole batch (network inference and
    scaling included in the warm-start figures)."""
    from .sweep import PlateSweep
    n = net.n
    h, p, f = _plates(n_eval, n, seed)
    sw = PlateSweep([(n, n)] * n_eval, a=1.0, b=1.0)
    sw.set_thickness(h.reshape(-1))
    fd = np.zeros((n_eval, (n + 1) * (n + 1), 6))
    fd[:, :, 2] = f.reshape(n_eval, -1)
    sw.set_loads(fd.reshape(-1))
    sw.assemble()
    torch.cuda.synchronize()
    sw.check()
    out = dict(plates=n_eval)

    def timed(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        r = fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1), r

    if log:
        log('evaluate: %d plates assembled' % n_eval)
    for mode in ('fp64', 'mixed'):
        ms0, _ = timed(lambda: sw.solve(mode=mode, rtol=1e-8, max_iter=max_iter))
        if log:
            log('%s zero start: %.1f ms' % (mode, ms0))
        it0 = sw.iters.cpu().numpy().copy()
        rel0 = float(sw.relres.max())
        conv0 = int((sw.relres <= 1e-8).sum())

        # energy norm of the solution, ||u*||_K^2 = u*.f, for the error of
        # the warm start: ||u* - alpha u0||_K^2 / ||u*||_K^2 = 1 - (u0.f)^2 /
        # (u0.K u0 * u*.f)
        usf = (sw.u.view(n_eval, -1) * sw.f.view(n_eval, -1)).sum(dim=1)

        gstats = {}

        def warm():
            u0 = predict(net, h, p, sc)
            us, fallback = galerkin_scale(sw, u0, gstats)
            sw.u.copy_(us)
            sw.solve(mode=mode, rtol=1e-8, x0=True, max_iter=max_iter)
            return fallback
        ms1, fallback = timed(warm)
        uf, uku = gstats['uf'], gstats['uku']
        e2 = (1 - uf * uf / (uku * usf)).clamp_min(0).sqrt().cpu().numpy()
        it1 = sw.iters.cpu().numpy().copy()
        rel1 = float(sw.relres.max())
        conv1 = int((sw.relres <= 1e-8).sum())
        ms_net, _ = timed(lambda: predict(net, h, p, sc))
        out[mode] = dict(zero_ms=ms0, zero_solves_per_s=n_eval / (ms0 * 1e-3),
                         zero_iters_mean=float(it0.mean()), zero_max_relres=rel0,
                         zero_converged=conv0, warm_converged=conv1,
                         zero_iters_median=float(np.median(it0)),
                         warm_iters_median=float(np.median(it1)),
                         warm_ms=ms1, warm_solves_per_s=n_eval / (ms1 * 1e-3),
                         warm_iters_mean=float(it1.mean()), warm_max_relres=rel1,
                         network_ms=ms_net, fallback=fallback,
                         warm_energy_error_median=float(np.median(e2)),
                         warm_energy_error_mean=float(e2.mean()),
                         iteration_ratio=float(it0.mean() / max(1e-9, it1.mean())))
        if log:
            log('%s: zero %.1f ms (%.0f its), warm %.1f ms (%.0f its, net %.1f ms, %d fallback)'
                % (mode, ms0, it0.mean(), ms1, it1.mean(), ms_net, fallback))
    out['mixed_warm_vs_fp64_zero'] = out['fp64']['zero_ms'] / out['mixed']['warm_ms']
    return out

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
        _lib.call('pf3d_nn_bias_grad', M, co, dy.data_ptr(), self.g[name + '.b'].data_ptr(),
                  self.partial.data_ptr(), self.partial.numel(), st)
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
        self.backward_from(stream)

    def energy_loss(self, system, f, s_energy, scale, stream=None):
        """Energy-norm loss of the guesses of the last forward() on the
        plates of ``system`` (a PlateSweep of B plates of n x n elements,
        assembled): the mean over the plates of 1 - (x.f)^2 / ((x.K x) s),
        x the node values scaled per channel (0 on fixed DOFs), s = u*.f the
        solutions' energies.  Leaves the gradient with respect to the
        element outputs in place for backward_from(); returns the per-plate
        (x.f, x.K x, loss) on the device."""
        n, B, st = self.n, self.B, _s(stream)
        ndof = B * (n + 1) * (n + 1) * 6
        if not hasattr(self, '_x64') or self._x64.numel() != ndof:
            self._x64 = torch.empty(ndof, dtype=torch.float64, device=self.dev)
            self._kx = torch.empty_like(self._x64)
            self._acl = torch.empty(3 * B, dtype=torch.float64, device=self.dev)
        _lib.call('pf3d_nn_to_dofs', ndof, 6, self.u.data_ptr(), scale.data_ptr(),
                  system.fixed.data_ptr(), self._x64.data_ptr(), st)
        _lib.call('pf3d_bsr_spmv', system.n_nodes, system.plan.row_ptr.data_ptr(),
                  system.plan.col_idx.data_ptr(), system.val.data_ptr(), self._x64.data_ptr(),
                  self._kx.data_ptr(), st)
        _lib.call('pf3d_nn_energy_dots', B, ndof // B, self._x64.data_ptr(), self._kx.data_ptr(),
                  f.data_ptr(), s_energy.data_ptr(), self._acl.data_ptr(), st)
        _lib.call('pf3d_nn_energy_grad', B, n, 6, self._kx.data_ptr(), f.data_ptr(),
                  self._acl.data_ptr(), s_energy.data_ptr(), scale.data_ptr(),
                  system.fixed.data_ptr(), 1.0 / B, self.dact['head'].data_ptr(), st)
        return self._acl.view(B, 3)

    def backward_from(self, stream=None):
        """Parameter gradients from the gradient with respect to the element
        outputs held in ``dact['head']`` (set by backward() or
        energy_loss())."""
        n, B, st = self.n, self.B, _s(stream)
        de = self.dact['head']
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

    def gemm_flops(self):
        """(forward, backward) floating-point operations of the matrix-core
        GEMMs for the current batch: 2 M N K per product; backward computes
        dW for every layer and dX for every layer but the first."""
        fwd = bwd = 0
        for name, ci, co, _, k, _ in LAYERS:
            _, ho = self._shape(name)
            M, Kd = self.B * ho * ho, k * k * ci
            fwd += 2 * M * co * Kd
            bwd += 2 * M * co * Kd * (1 if name == 'c0' else 2)
        return fwd, bwd

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

# channel order of the node values: u, v, w, rx, ry, rz
W_CHANNEL = 2


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


def _dof_loads(f):
    """(V, n+1, n+1) vertical nodal forces -> (V * (n+1)^2 * 6) DOF vector."""
    V, n1, _ = f.shape
    fd = np.zeros((V, n1 * n1, 6))
    fd[:, :, 2] = f.reshape(V, -1)
    return fd.reshape(-1)


def solve_plates(h, f, mode='direct', rtol=1e-8, sweep=None, x0=None, max_iter=200000):
    """Solve every plate (clamped edges) with the batched GPU solvers
    (banded Cholesky for 'direct', block-Jacobi PCG for 'fp64' / 'mixed').
    Returns (u (V, n+1, n+1, 6) float64 tensor, sweep)."""
    from .sweep import PlateSweep
    V, n = h.shape[0], h.shape[1]
    if sweep is None or sweep.n_var != V:
        sweep = PlateSweep([(n, n)] * V, a=1.0, b=1.0)
    sweep.set_thickness(h.reshape(-1))
    sweep.set_loads(_dof_loads(f))
    if x0 is not None:
        sweep.u.copy_(torch.as_tensor(x0, dtype=torch.float64).reshape(-1))
    if mode == 'direct':
        sweep.step(mode='direct', rtol=rtol)
    else:
        sweep.step(mode=mode, rtol=rtol, x0=x0 is not None, max_iter=max_iter)
    return sweep.u.view(V, n + 1, n + 1, 6), sweep


class Scales:
    """Per-channel scale of the network's node values: the standard
    deviation over the training solutions, each normalised by the rms of
    its deflection (the amplitude of a guess comes from the Galerkin
    projection at solve time, so only shapes are learnt).  Channels that
    are zero in every solution (u, v, rz of these isotropic plates) get
    scale 0: the guess leaves them at zero."""

    def __init__(self, std):
        self.std = np.asarray(std, dtype=np.float64)

    def tensor(self, device='cuda'):
        return torch.as_tensor(self.std, dtype=torch.float32, device=device)


class EnergyData:
    """Training plates on the GPU: inputs X (S, n, n, 2), thicknesses H
    (S, n*n), loads F (S, (n+1)^2 * 6), solution energies s = u*.f (S) and
    the solutions' shapes T (S, n+1, n+1, 6) in the network's units
    (u / rms(w) / channel scale; for the nodal pre-training)."""

    def __init__(self, X, H, F, S, T, scales):
        self.X, self.H, self.F, self.S, self.T, self.scales = X, H, F, S, T, scales

    @property
    def size(self):
        return self.X.shape[0]


def make_energy_dataset(n_samples, seed, n=64, chunk=256, log=None):
    """Seeded plates solved by the batched banded Cholesky (every residual
    checked at 1e-8); keeps what the energy loss needs (no nodal targets):
    inputs, thicknesses, loads and u*.f, plus the channel scales."""
    X, H, F, S, U = [], [], [], [], []
    sweep = None
    ssum = np.zeros(6)
    count = 0
    for c0 in range(0, n_samples, chunk):
        V = min(chunk, n_samples - c0)
        h, p, f = _plates(V, n, seed * 1000003 + c0)
        u, sweep = solve_plates(h, f, mode='direct',
                                sweep=sweep if sweep is not None and sweep.n_var == V else None)
        torch.cuda.synchronize()
        sweep.check()
        if float(sweep.relres.max()) > 1e-8:
            raise RuntimeError('training solve missed the tolerance')
        x, _ = features(h, p)
        X.append(torch.as_tensor(x, device='cuda'))
        H.append(torch.as_tensor(h.reshape(V, -1), device='cuda'))
        F.append(sweep.f.view(V, -1).clone())
        S.append((sweep.u.view(V, -1) * sweep.f.view(V, -1)).sum(dim=1))
        w = u[:, 1:-1, 1:-1, W_CHANNEL]
        rms = w.pow(2).mean(dim=(1, 2)).sqrt().clamp_min(1e-300)
        Tn = (u / rms[:, None, None, None]).to(torch.float32)
        U.append(Tn)
        T = Tn[:, 1:-1, 1:-1].reshape(-1, 6).double()
        ssum += T.pow(2).sum(dim=0).cpu().numpy()
        count += T.shape[0]
        if log:
            log('dataset %d/%d' % (c0 + V, n_samples))
    std = np.sqrt(ssum / count)
    # channels that carry no displacement (relative to w) stay at zero
    std = np.where(std > 1e-9 * std[W_CHANNEL], std, 0.)
    T = torch.cat(U)
    T /= torch.as_tensor(np.where(std > 0, std, 1.), dtype=torch.float32, device='cuda')
    return EnergyData(torch.cat(X), torch.cat(H), torch.cat(F), torch.cat(S), T, Scales(std))


def train_energy(net, data, epochs=12, batch=32, lr=1e-3, seed=0, pretrain_epochs=3,
                 log=None):
    """Adam, first ``pretrain_epochs`` on the nodal MSE of the solution
    shapes (channels with a scale), then on the energy-norm loss (mean over
    the batch of the relative energy error of the best multiple of the
    guess), each batch's stiffness assembled on the fly (one KC0 launch and
    one block-plan apply for its plates).  Returns the mean energy loss of
    the training batches per epoch (also measured during the pre-training)."""
    from .sweep import PlateSweep
    n = net.n
    sweep = PlateSweep([(n, n)] * batch, a=1.0, b=1.0)
    scale = data.scales.tensor()
    wc = (scale > 0).to(torch.float32)
    S = data.size
    g = torch.Generator(device='cpu').manual_seed(seed)
    hist = []
    total = pretrain_epochs + epochs
    for ep in range(total):
        pre = ep < pretrain_epochs
        perm = torch.randperm(S, generator=g)
        tot = torch.zeros((), dtype=torch.float64, device='cuda')
        nb = 0
        k = ep if pre else ep - pretrain_epochs
        span = pretrain_epochs if pre else epochs
        lr_ep = lr * (0.5 ** (k // max(1, span // 3)))
        for b0 in range(0, S - batch + 1, batch):
            idx = perm[b0:b0 + batch].to('cuda')
            sweep.set_thickness(data.H[idx].reshape(-1))
            sweep.assemble()
            net.forward(data.X[idx].contiguous())
            acl = net.energy_loss(sweep, data.F[idx].contiguous().view(-1),
                                  data.S[idx].contiguous(), scale)
            tot += acl[:, 2].mean()
            nb += 1
            if pre:
                net.backward(data.T[idx].contiguous(), wc)
            else:
                net.backward_from()
            net.adam(lr=lr_ep)
        hist.append(float(tot) / max(1, nb))
        if log:
            log('epoch %d (%s) energy loss %.4e (lr %.1e)'
                % (ep, 'nodal MSE' if pre else 'energy', hist[-1], lr_ep))
    return hist


def predict(net, x, scales, fixed, batch=64):
    """FP64 DOF vectors (V * (n+1)^2 * 6) of the network's guesses for the
    inputs x (V, n, n, 2) (pf3d_nn_to_dofs: node values times the channel
    scales, fixed DOFs 0), before the Galerkin scaling."""
    V = x.shape[0]
    n1 = net.n + 1
    per = n1 * n1 * 6
    out = torch.empty(V * per, dtype=torch.float64, device='cuda')
    scale = scales.tensor()
    st = torch.cuda.current_stream().cuda_stream
    for b0 in range(0, V, batch):
        xb = torch.as_tensor(x[b0:b0 + batch], device='cuda')
        k = xb.shape[0]
        u = net.forward(xb)
        _lib.call('pf3d_nn_to_dofs', k * per, 6, u.data_ptr(), scale.data_ptr(),
                  fixed[b0 * per:(b0 + k) * per].data_ptr(), out[b0 * per:].data_ptr(), st)
    return out


def galerkin_scale(system, f, u0, n_var, stats=None):
    """Per plate alpha = u0.f / u0.K u0: the multiple of u0 closest to the
    solution in the energy norm, whose error ||u - alpha u0||_K^2 =
    ||u||_K^2 - (u0.f)^2 / u0.K u0 never exceeds that of a zero start (the
    norm in which CG converges).  Plates without a finite nonzero alpha
    start from zero (either sign is fine: the identity above does not
    depend on it).  ``system``: anything with plan, val, n_nodes (the
    plates' block-diagonal BSR, boundary conditions applied).  Returns
    (scaled u0, number of zero starts); ``stats`` receives u0.f, u0.K u0."""
    u = u0.clone()
    Ku = torch.empty_like(u)
    _lib.call('pf3d_bsr_spmv', system.n_nodes, system.plan.row_ptr.data_ptr(),
              system.plan.col_idx.data_ptr(), system.val.data_ptr(), u.data_ptr(), Ku.data_ptr(),
              torch.cuda.current_stream().cuda_stream)
    uv, kv, fv = u.view(n_var, -1), Ku.view(n_var, -1), f.view(n_var, -1)
    uf = (uv * fv).sum(dim=1)
    uku = (uv * kv).sum(dim=1)
    # the energy-norm identity holds for either sign of alpha (the energy
    # loss does not fix the sign of a shape)
    ok = (uku > 0) & (uf != 0) & torch.isfinite(uf) & torch.isfinite(uku)
    alpha = torch.where(ok, uf / torch.where(ok, uku, torch.ones_like(uku)),
                        torch.zeros_like(uku))
    uv.mul_(alpha[:, None])
    if stats is not None:
        stats['uf'], stats['uku'] = uf, uku
    return u, int((~ok).sum())


def save(net, sc, path, meta=None):
    """Parameters, channel scales and evaluation metadata (JSON text) as an
    .npz (no pickled objects)."""
    import json
    arrs = {k: v.astype(np.float32) for k, v in net.params_numpy().items()}
    arrs['scales.std'] = np.asarray(sc.std)
    arrs['meta.json'] = np.array(json.dumps(meta or {}))
    np.savez(path, **arrs)


def load(path, batch=64, device='cuda'):
    """(net, scales, meta) from save()."""
    import json
    z = np.load(path, allow_pickle=False)
    net = WarmStartNet(n=64, batch=batch, device=device)
    for k in net.p:
        net.p[k].copy_(torch.as_tensor(z[k], dtype=torch.float32))
    meta = json.loads(str(z['meta.json'])) if 'meta.json' in z.files else {}
    return net, Scales(z['scales.std']), meta


def _timed(fn):
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    r = fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1), r


def evaluate_pcg(net, sc, h, p, f, mode='fp64', max_iter=60000, log=None):
    """Held-out plates through the batched block-Jacobi PCG to 1e-8 (every
    plate), from zero and from the network's Galerkin-scaled guess (network
    and scaling inside the timed region).  Per-plate iteration counts."""
    from .sweep import PlateSweep
    V, n = h.shape[0], h.shape[1]
    sw = PlateSweep([(n, n)] * V, a=1.0, b=1.0)
    sw.set_thickness(h.reshape(-1))
    sw.set_loads(_dof_loads(f))
    sw.assemble()
    torch.cuda.synchronize()
    sw.check()
    x, _ = features(h, p)
    ms0, _ = _timed(lambda: sw.solve(mode=mode, rtol=1e-8, max_iter=max_iter))
    it0 = sw.iters.cpu().numpy().copy()
    ok0 = int((sw.relres <= 1e-8).sum())
    rel0 = float(sw.relres.max())
    usf = (sw.u.view(V, -1) * sw.f.view(V, -1)).sum(dim=1)
    gst = {}

    def warm():
        u0 = predict(net, x, sc, sw.fixed)
        us, fb = galerkin_scale(sw, sw.f, u0, V, gst)
        sw.u.copy_(us)
        sw.solve(mode=mode, rtol=1e-8, x0=True, max_iter=max_iter)
        return fb

    ms1, fallback = _timed(warm)
    it1 = sw.iters.cpu().numpy().copy()
    ok1 = int((sw.relres <= 1e-8).sum())
    rel1 = float(sw.relres.max())
    e2 = (1 - gst['uf'] ** 2 / (gst['uku'] * usf)).clamp_min(0).cpu().numpy()
    ms_net, _ = _timed(lambda: predict(net, x, sc, sw.fixed))
    out = dict(plates=V, zero_ms=ms0, warm_ms=ms1, network_ms=ms_net,
               zero_iters_mean=float(it0.mean()), warm_iters_mean=float(it1.mean()),
               iteration_ratio=float(it0.mean() / max(1e-9, it1.mean())),
               plates_fewer_iterations=int((it1 < it0).sum()),
               zero_converged=ok0, warm_converged=ok1, zero_max_relres=rel0,
               warm_max_relres=rel1, max_iter=max_iter, fallback=fallback,
               energy_error_median=float(np.median(np.sqrt(e2))),
               captured_energy_median=float(np.median(1 - e2)))
    if log:
        log('pcg %s, %d plates: zero %.0f ms (%.0f its), warm %.0f ms (%.0f its), ratio %.4f, '
            'energy error %.3f' % (mode, V, ms0, it0.mean(), ms1, it1.mean(),
                                   out['iteration_ratio'], out['energy_error_median']))
    return out, it0, it1


def evaluate_multigrid(net, sc, h, p, f, log=None):
    """The plates one by one through multigrid CG (FP64 V-cycle, setup per
    plate outside the timed region) to 1e-8, from zero and from the
    network's Galerkin-scaled guess (network and scaling timed with the
    warm solve).  Per-plate iteration counts."""
    from .solver import StaticSystem
    from .sweep import PlateSweep
    V, n = h.shape[0], h.shape[1]
    sw = PlateSweep([(n, n)], a=1.0, b=1.0)
    x, _ = features(h, p)
    it0, it1, ms0, ms1, conv, fb, rels = [], [], 0., 0., 0, 0, []
    for v in range(V):
        sw.set_thickness(h[v].reshape(-1))
        sw.set_loads(_dof_loads(f[v:v + 1]))
        sw.assemble()
        sysm = StaticSystem(sw.plan, sw.val, sw.fixed)
        sysm.setup_amg(sw.x.view(-1, 3).cpu().numpy(), precisions=('fp64',))
        sysm.solve(sw.f, method='amg', rtol=1e-8)   # warm-up: graph capture
        t0, (_, i0) = _timed(lambda: sysm.solve(sw.f, method='amg', rtol=1e-8))

        def warm():
            g = predict(net, x[v:v + 1], sc, sw.fixed)
            us, k = galerkin_scale(sw, sw.f, g, 1)
            return sysm.solve(sw.f, method='amg', rtol=1e-8, x0=us), k

        t1, ((_, i1), k) = _timed(warm)
        it0.append(i0['iters'])
        it1.append(i1['iters'])
        rels.append(max(i0['relres'], i1['relres']))
        ms0 += t0
        ms1 += t1
        fb += k
        conv += int(i0['converged'] and i1['converged'])
        # release this plate's graphs here, not inside the next capture
        sysm.__dict__.pop('_cg_iterations', None)
        del sysm
    it0, it1 = np.array(it0), np.array(it1)
    out = dict(plates=V, zero_ms=ms0, warm_ms=ms1, zero_iters_mean=float(it0.mean()),
               warm_iters_mean=float(it1.mean()),
               iteration_ratio=float(it0.mean() / max(1e-9, it1.mean())),
               plates_fewer_iterations=int((it1 < it0).sum()), all_converged=conv == V,
               converged=conv, max_relres=float(max(rels)), fallback=fb)
    if log:
        log('multigrid, %d plates: zero %.0f ms (%.1f its), warm %.0f ms (%.1f its), ratio %.4f'
            % (V, ms0, it0.mean(), ms1, it1.mean(), out['iteration_ratio']))
    return out


def evaluate(net, sc, n_eval, seed, chunk=500, multigrid_plates=100, max_iter=60000,
             log=None):
    """Held-out plates (seeds disjoint from training): block-Jacobi PCG on
    all of them in batches of ``chunk``, multigrid CG on the first
    ``multigrid_plates``, zero start against the warm start."""
    res = dict(plates=n_eval, pcg=[], multigrid=None)
    it0s, it1s = [], []
    for c0 in range(0, n_eval, chunk):
        V = min(chunk, n_eval - c0)
        h, p, f = _plates(V, 64, seed * 1000003 + c0)
        r, it0, it1 = evaluate_pcg(net, sc, h, p, f, max_iter=max_iter, log=log)
        res['pcg'].append(r)
        it0s.append(it0)
        it1s.append(it1)
        if c0 == 0 and multigrid_plates:
            res['multigrid'] = evaluate_multigrid(net, sc, h[:multigrid_plates],
                                                  p[:multigrid_plates], f[:multigrid_plates],
                                                  log=log)
    it0, it1 = np.concatenate(it0s), np.concatenate(it1s)
    pc = res['pcg']
    res['summary'] = dict(
        plates=int(it0.size),
        pcg_iteration_ratio=float(it0.mean() / it1.mean()),
        pcg_zero_ms=sum(r['zero_ms'] for r in pc), pcg_warm_ms=sum(r['warm_ms'] for r in pc),
        pcg_time_ratio=sum(r['zero_ms'] for r in pc) / sum(r['warm_ms'] for r in pc),
        pcg_plates_fewer_iterations=int((it1 < it0).sum()),
        pcg_all_converged=all(r['zero_converged'] == r['warm_converged'] == r['plates']
                              for r in pc),
        pcg_converged=(sum(r['zero_converged'] for r in pc),
                       sum(r['warm_converged'] for r in pc)),
        pcg_max_relres=max(max(r['zero_max_relres'], r['warm_max_relres']) for r in pc),
        fallback=sum(r['fallback'] for r in pc),
        energy_error_median=float(np.median([r['energy_error_median'] for r in pc])),
        multigrid_iteration_ratio=res['multigrid']['iteration_ratio']
        if res['multigrid'] else None,
        multigrid_time_ratio=res['multigrid']['zero_ms'] / res['multigrid']['warm_ms']
        if res['multigrid'] else None)
    return res


class WarmStart:
    """The trained network as the initial guess of the batched plate solves,
    with the policy recorded at training time: the guess is used only if it
    saved at least 1 % of the block-Jacobi PCG iterations on the held-out
    plates (``meta['use_guess']``); otherwise every plate starts from zero
    (the fallback).  Plates whose Galerkin multiple is not finite start from
    zero in any case."""

    def __init__(self, path=None, batch=64):
        import os
        path = path or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                    'warmstart_weights.npz')
        self.net, self.scales, self.meta = load(path, batch=batch)
        self.use_guess = bool(self.meta.get('use_guess', False))

    def start(self, sweep, h, p, force=False):
        """Write the initial guess of every plate of ``sweep`` (assembled,
        loads set; n x n plates with thickness h and pressure p, (V, n, n))
        into sweep.u and return (x0 flag for sweep.solve, zero starts)."""
        if not (self.use_guess or force):
            return False, sweep.n_var
        x, _ = features(h, p)
        g = predict(self.net, x, self.scales, sweep.fixed)
        us, fallback = galerkin_scale(sweep, sweep.f, g, sweep.n_var)
        sweep.u.copy_(us)
        return True, fallback

"""Warm-start network kernels on the GPU against the CPU oracle
(oracle/nn_oracle.py): the bfloat16 matrix-core GEMM against a float64
product of the same bfloat16-rounded operands, im2col/col2im and the
up-sampling exactly, and the whole forward/backward pass within the
accuracy that bfloat16 operands allow."""
import numpy as np
import pytest
import torch

from oracle import nn_oracle as N

pytestmark = pytest.mark.gpu
DEV = 'cuda'


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), back to float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7fff + ((u >> 16) & 1)) & 0xffff0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _call(name, *args):
    from pyfe3d_amd import _lib
    _lib.call(name, *args, torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize('M,N,K,ta,splits', [(64, 64, 32, False, 1), (77, 45, 50, False, 1),
                                              (130, 6, 288, False, 1), (33, 70, 1000, True, 7),
                                              (6, 288, 4096, True, 64)])
def test_gemm_matches_bf16_product(M, N, K, ta, splits):
    """Lane layout of the 16x16x16 bf16 MFMA, ragged edges, transposed A,
    bias + ReLU epilogue, accumulate and deterministic split-K."""
    rng = np.random.default_rng(M * 7 + N)
    A = rng.normal(size=(K, M) if ta else (M, K)).astype(np.float32)
    B = rng.normal(size=(N, K)).astype(np.float32)
    bias = rng.normal(size=N).astype(np.float32)
    C0 = rng.normal(size=(M, N)).astype(np.float32)
    Ab = bf16(A).T if ta else bf16(A)
    ref = Ab @ bf16(B).T
    tA, tB = torch.as_tensor(A, device=DEV), torch.as_tensor(B, device=DEV)
    tb = torch.as_tensor(bias, device=DEV)
    part = torch.empty(max(1, splits) * M * N, dtype=torch.float32, device=DEV)
    sam, sak = (1, M) if ta else (K, 1)
    outs = []
    for rep in range(2):
        C = torch.as_tensor(C0, device=DEV).clone()
        _call('pf3d_nn_gemm', M, N, K, tA.data_ptr(), sam, sak, tB.data_ptr(), K, 1,
              C.data_ptr(), N, 1, tb.data_ptr(), 1, splits,
              part.data_ptr() if splits > 1 else None)
        torch.cuda.synchronize()
        outs.append(C.cpu().numpy())
    want = np.maximum(C0 + ref + bias, 0)
    scale = np.abs(Ab).max() * np.abs(bf16(B)).max() * K
    assert np.abs(outs[0] - want).max() <= 1e-6 * scale + 1e-5 * np.abs(want).max()
    assert (outs[0] == outs[1]).all()


def test_im2col_col2im_up2_node_avg_exact():
    rng = np.random.default_rng(3)
    for s in (1, 2):
        x = rng.normal(size=(2, 8, 8, 5)).astype(np.float32)
        Ho = (8 - 1) // s + 1
        col = torch.empty((2 * Ho * Ho, 45), dtype=torch.float32, device=DEV)
        tx = torch.as_tensor(x, device=DEV)
        _call('pf3d_nn_im2col', 2, 8, 8, 5, s, tx.data_ptr(), col.data_ptr())
        assert np.array_equal(col.cpu().numpy(), N.im2col(x, s))
        d = rng.normal(size=col.shape).astype(np.float32)
        dx = torch.empty_like(tx)
        td = torch.as_tensor(d, device=DEV)
        _call('pf3d_nn_col2im', 2, 8, 8, 5, s, td.data_ptr(), dx.data_ptr(), 0)
        np.testing.assert_allclose(dx.cpu().numpy(), N.col2im(d.astype(np.float64), x.shape, s),
                                   rtol=1e-6, atol=1e-6)
    lo = rng.normal(size=(2, 4, 4, 3)).astype(np.float32)
    sk = rng.normal(size=(2, 8, 8, 3)).astype(np.float32)
    y = torch.empty((2, 8, 8, 3), dtype=torch.float32, device=DEV)
    tlo, tsk = torch.as_tensor(lo, device=DEV), torch.as_tensor(sk, device=DEV)
    _call('pf3d_nn_up2_add', 2, 8, 8, 3, tlo.data_ptr(), tsk.data_ptr(), y.data_ptr())
    assert np.array_equal(y.cpu().numpy(), (N.up2(lo, 8, 8) + sk).astype(np.float32))
    e = rng.normal(size=(2, 6, 6, 3)).astype(np.float32)
    u = torch.empty((2, 7, 7, 3), dtype=torch.float32, device=DEV)
    te = torch.as_tensor(e, device=DEV)
    _call('pf3d_nn_node_avg', 2, 6, 3, te.data_ptr(), u.data_ptr())
    np.testing.assert_allclose(u.cpu().numpy(), N.node_avg(e.astype(np.float64)), rtol=1e-6,
                               atol=1e-7)


def _cos(a, b):
    a, b = a.ravel(), b.ravel()
    return float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300))


def test_network_forward_backward_vs_oracle():
    """The GPU passes against the float64 oracle with the same parameters:
    outputs within bf16 accuracy, every gradient tensor aligned (cosine >=
    0.995) and of the same size (norm ratio within 3%)."""
    from pyfe3d_amd.warmstart import WarmStartNet
    rng = np.random.default_rng(5)
    net = WarmStartNet(n=16, batch=3, seed=1)
    for k in net.p:
        if k.endswith('.b'):
            net.p[k].copy_(torch.as_tensor(rng.normal(size=net.p[k].shape) * 0.05))
    p = net.params_numpy()
    x = rng.normal(size=(3, 16, 16, 2)).astype(np.float32)
    t = rng.normal(size=(3, 17, 17, 6)).astype(np.float32)
    wc = np.array([1., 2., 0.5, 1., 1.5, 1.], dtype=np.float32)
    u_ref, cache = N.forward(p, x.astype(np.float64))
    g_ref = N.backward(p, cache, u_ref, t.astype(np.float64), wc.astype(np.float64))
    u = net.forward(torch.as_tensor(x, device=DEV))
    net.backward(torch.as_tensor(t, device=DEV), torch.as_tensor(wc, device=DEV))
    torch.cuda.synchronize()
    ug = u.cpu().numpy()
    err = np.abs(ug - u_ref).max() / np.abs(u_ref).max()
    print('forward rel err %.2e' % err)
    assert err < 3e-2
    g = net.grads_numpy()
    for k in sorted(g_ref):
        c = _cos(g[k], g_ref[k])
        ratio = np.linalg.norm(g[k]) / np.linalg.norm(g_ref[k])
        print(k, 'cos %.5f ratio %.4f' % (c, ratio))
        assert c >= 0.995 and abs(ratio - 1) < 0.03, k


def test_adam_and_training_reduce_the_loss():
    from pyfe3d_amd.warmstart import WarmStartNet
    rng = np.random.default_rng(6)
    net = WarmStartNet(n=16, batch=8, seed=2)
    x = torch.as_tensor(rng.normal(size=(8, 16, 16, 2)).astype(np.float32), device=DEV)
    # a smooth target that depends on the input
    t = torch.zeros((8, 17, 17, 6), dtype=torch.float32, device=DEV)
    t[:, 1:-1, 1:-1, 2] = x[:, 1:, 1:, 0].mean(dim=(1, 2))[:, None, None]
    wc = torch.ones(6, dtype=torch.float32, device=DEV)
    losses = []
    for it in range(60):
        u = net.forward(x)
        d = (u - t)[:, 1:-1, 1:-1]
        losses.append(0.5 * float((d * d).sum()) / (8 * 15 * 15))
        net.backward(t, wc)
        net.adam(lr=2e-3)
    print('loss', losses[0], losses[-1])
    assert losses[-1] < 0.2 * losses[0]
    # Adam step against its formula on one tensor
    p = torch.tensor([1.0, -2.0], device=DEV)
    g = torch.tensor([0.5, 0.25], device=DEV)
    m = torch.zeros(2, device=DEV)
    v = torch.zeros(2, device=DEV)
    _call('pf3d_nn_adam', 2, p.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(), 0.1, 0.9,
          0.999, 1e-8, 1)
    mt = 0.1 * g.cpu().numpy()
    vt = 0.001 * g.cpu().numpy() ** 2
    want = np.array([1.0, -2.0]) - 0.1 * (mt / 0.1) / (np.sqrt(vt / 0.001) + 1e-8)
    np.testing.assert_allclose(p.cpu().numpy(), want, rtol=1e-6)


def test_energy_loss_gradients_vs_oracle():
    """The energy-norm loss on the GPU (pf3d_nn_to_dofs, the plates' BSR
    product, pf3d_nn_energy_dots / _grad, backward_from) against the
    float64 oracle on two 16 x 16 plates whose stiffness the oracle
    assembles itself: the per-plate losses within bf16 accuracy and every
    gradient tensor aligned (cosine >= 0.995, norm within 3%)."""
    import scipy.sparse as ss
    import scipy.sparse.linalg as sla
    from oracle import oracle as O
    from pyfe3d_amd.mesh import PlateMesh
    from pyfe3d_amd.shellprop import isotropic_records
    from pyfe3d_amd.sweep import PlateSweep
    from pyfe3d_amd.warmstart import WarmStartNet, _dof_loads, _plates, features
    n, B = 16, 2
    h, pr, fz = _plates(B, n, 21)
    sw = PlateSweep([(n, n)] * B, a=1.0, b=1.0)
    sw.set_thickness(h.reshape(-1))
    sw.set_loads(_dof_loads(fz))
    sw.assemble()
    m = PlateMesh(n + 1, n + 1, 1.0, 1.0)
    fixed = m.clamped()
    P = ss.diags((~fixed).astype(float))
    E = m.quads.shape[0]
    Ks, fs, s_ = [], [], []
    for b in range(B):
        rec = isotropic_records(h[b].reshape(-1), 70e9, 0.33, rho=2700.)
        r = np.zeros(576 * E, np.int64); c = np.zeros_like(r); v = np.zeros(576 * E)
        O.assemble('quad4', 'KC0', m.x, m.quads, 576 * np.arange(E), np.zeros((E, 3)),
                   np.arange(E), rec, r, c, v)
        K = ss.coo_matrix((v, (r, c)), shape=(m.n_dof, m.n_dof)).tocsr()
        K = (P @ K @ P + ss.diags(fixed.astype(float))).tocsr()
        f = _dof_loads(fz[b:b + 1])
        f[fixed] = 0.
        u = sla.spsolve(K.tocsc(), f)
        Ks.append(K); fs.append(f); s_.append(float(u @ f))
    x, _ = features(h, pr)
    net = WarmStartNet(n=n, batch=B, seed=3)
    rng = np.random.default_rng(4)
    for k in net.p:
        if k.endswith('.b'):
            net.p[k].copy_(torch.as_tensor(rng.normal(size=net.p[k].shape) * 0.05))
    scale = np.array([0., 0., 1e-3, 2e-2, 2e-2, 0.], dtype=np.float32)
    p = net.params_numpy()
    u_ref, cache = N.forward(p, x.astype(np.float64))
    fixed2 = np.stack([fixed, fixed])
    L_ref, du, cap = N.energy_loss(u_ref, scale.astype(np.float64), fixed2, Ks, fs, s_)
    g_ref = N.backward_from(p, cache, du)
    net.forward(torch.as_tensor(x, device=DEV))
    acl = net.energy_loss(sw, sw.f, torch.as_tensor(s_, dtype=torch.float64, device=DEV),
                          torch.as_tensor(scale, device=DEV))
    net.backward_from()
    torch.cuda.synchronize()
    cap_gpu = 1. - acl[:, 2].cpu().numpy()
    print('captured energy: gpu', cap_gpu, 'oracle mean', cap)
    assert abs(cap_gpu.mean() - cap) <= 3e-2 * cap
    g = net.grads_numpy()
    for k in sorted(g_ref):
        cs = _cos(g[k], g_ref[k])
        ratio = np.linalg.norm(g[k]) / np.linalg.norm(g_ref[k])
        print(k, 'cos %.5f ratio %.4f' % (cs, ratio))
        assert cs >= 0.995 and abs(ratio - 1) < 0.03, k

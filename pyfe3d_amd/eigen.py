"""Generalized symmetric eigenproblems A x = lambda B x on the GPU: natural
frequencies (A = KC0, B = M) and linear buckling (A = KC0, B = -KG).

Replaces the reference callers' ``eigsh(A=KC0uu, M=Muu, sigma=-1.,
which='LM')`` (tests/test_quad4_natural_freq.py:120-123) and
``eigsh(A=KGuu, M=KC0uu, sigma=1., which='SM', mode='cayley')``
(tests/test_quad4_linear_buckling_plate_mat_coord.py:184-185).

Shift-invert Lanczos: with C = A - sigma B (symmetric positive definite for
the shifts used here) the operator T = C^-1 B is self-adjoint in the
C-inner product and its largest eigenvalues theta = 1 / (lambda - sigma)
belong to the eigenvalues lambda closest above sigma.  Each step is one
block-sparse product with B, one banded Cholesky solve with C (factorised
once, on the GPU) and a full re-orthogonalisation against the basis in the
C-inner product (the basis V and U = C V are both kept, so no product with
C is needed).  Fixed DOFs are removed as in the static solve: identity rows
and columns in A and C, zero rows and columns in B.
"""
import numpy as np
import torch

from . import _lib
from .solver import StaticSystem


def _s():
    return torch.cuda.current_stream().cuda_stream


class GeneralizedEigen:
    """plan: BSR structure shared by A and B (e.g. BSRPlan of KC0);
    a_val, b_val: their BSR values; fixed: bool per DOF."""

    def __init__(self, plan, a_val, b_val, fixed, sigma=0.0):
        self.plan = plan
        nb = plan.n_nodes
        self.n = 6 * nb
        dev = a_val.device
        fixed_t = torch.as_tensor(np.asarray(fixed, dtype=np.uint8) if not torch.is_tensor(fixed)
                                  else fixed.to(torch.uint8), device=dev).contiguous()
        self.free = (1 - fixed_t.to(torch.float64))
        pr, pc = plan.row_ptr.data_ptr(), plan.col_idx.data_ptr()
        self.B = b_val.clone()
        _lib.call('pf3d_bsr_dirichlet_diag', nb, pr, pc, self.B.data_ptr(), fixed_t.data_ptr(),
                  0.0, _s())
        A = a_val.clone()
        _lib.call('pf3d_bsr_dirichlet', nb, pr, pc, A.data_ptr(), fixed_t.data_ptr(), _s())
        # C = A - sigma B
        _lib.call('pf3d_axpby', A.numel(), -float(sigma), self.B.data_ptr(), 1.0, A.data_ptr(),
                  _s())
        self.sigma = float(sigma)
        self.C = StaticSystem(plan, A, fixed_t)
        self.C.factor_direct()

    def _spmv(self, val, x, y):
        _lib.call('pf3d_bsr_spmv', self.plan.n_nodes, self.plan.row_ptr.data_ptr(),
                  self.plan.col_idx.data_ptr(), val.data_ptr(), x.data_ptr(), y.data_ptr(), _s())

    def _csolve(self, f, x):
        c = self.C
        _lib.call('pf3d_band_solve', 1, None, c._band_dof0.data_ptr(), c._band_bw.data_ptr(),
                  c._band_off.data_ptr(), c._band.data_ptr(), c._band_pnode.data_ptr(),
                  f.data_ptr(), x.data_ptr(), c._band_work.data_ptr(), 0, _s())

    def _dots(self, V, j, x, out):
        """out[0:j] = V[:, 0:j]^T x"""
        _lib.call('pf3d_gemv_t', self.n, j, V.data_ptr(), self.n, x.data_ptr(), out.data_ptr(),
                  1.0, 0.0, _s())

    def solve(self, k=6, m=None, tol=1e-10, seed=0, max_m=400):
        """The k eigenpairs with lambda closest above sigma, ascending.
        Returns (lambda (k,), x (n, k) on the GPU, info)."""
        n = self.n
        dev = self.B.device
        m = m or max(2 * k + 20, 40)
        while True:
            lam, X, res = self._lanczos(k, m, seed)
            if (res <= tol).all() or m >= max_m:
                return lam, X, dict(m=m, residual=res)
            m = min(2 * m, max_m)

    def _lanczos(self, k, m, seed):
        n = self.n
        dev = self.B.device
        V = torch.zeros((m + 1, n), dtype=torch.float64, device=dev)  # column-major n x (m+1)
        U = torch.zeros_like(V)                                          # U = C V
        g = torch.Generator(device='cpu').manual_seed(seed)
        v = torch.rand(n, generator=g, dtype=torch.float64).to(dev) - 0.5
        v.mul_(self.free)
        cv = torch.empty_like(v)
        self._spmv(self.C.val, v, cv)
        c1 = torch.empty(1, dtype=torch.float64, device=dev)
        self._dots(cv.view(1, -1), 1, v, c1)
        b0 = float(c1.item()) ** 0.5
        _lib.call('pf3d_axpby', n, 1.0 / b0, v.data_ptr(), 0.0, V[0].data_ptr(), _s())
        _lib.call('pf3d_axpby', n, 1.0 / b0, cv.data_ptr(), 0.0, U[0].data_ptr(), _s())
        alpha = np.zeros(m)
        beta = np.zeros(m)
        bv = torch.empty(n, dtype=torch.float64, device=dev)
        w = torch.empty_like(bv)
        cw = torch.empty_like(bv)
        coef = torch.empty(m + 1, dtype=torch.float64, device=dev)
        for j in range(m):
            self._spmv(self.B, V[j], bv)              # B v_j
            self._csolve(bv, w)                       # w = C^-1 B v_j
            cw.copy_(bv)                              # C w = B v_j
            # two passes of classical Gram-Schmidt in the C-inner product
            a_j = 0.
            for _ in range(2):
                self._dots(U, j + 1, w, coef)         # coef = U^T w = <v_i, w>_C
                _lib.call('pf3d_gemv_n', n, j + 1, V.data_ptr(), n, coef.data_ptr(), w.data_ptr(),
                          -1.0, 1.0, _s())
                _lib.call('pf3d_gemv_n', n, j + 1, U.data_ptr(), n, coef.data_ptr(),
                          cw.data_ptr(), -1.0, 1.0, _s())
                a_j += float(coef[j].item())
            alpha[j] = a_j
            self._dots(cw.view(1, -1), 1, w, c1)
            bj = max(float(c1.item()), 0.) ** 0.5
            beta[j] = bj
            if bj == 0.:
                m = j + 1
                break
            _lib.call('pf3d_axpby', n, 1.0 / bj, w.data_ptr(), 0.0, V[j + 1].data_ptr(), _s())
            _lib.call('pf3d_axpby', n, 1.0 / bj, cw.data_ptr(), 0.0, U[j + 1].data_ptr(), _s())
        T = np.diag(alpha[:m]) + np.diag(beta[:m - 1], 1) + np.diag(beta[:m - 1], -1)
        theta, S = np.linalg.eigh(T)
        order = np.argsort(-theta)[:k]
        theta, S = theta[order], S[:, order]
        res = np.abs(beta[m - 1] * S[m - 1, :]) / np.abs(theta)
        lam = self.sigma + 1.0 / theta
        Sg = torch.as_tensor(np.ascontiguousarray(S), device=dev)
        X = torch.empty((k, n), dtype=torch.float64, device=dev)
        for i in range(k):
            _lib.call('pf3d_gemv_n', n, m, V.data_ptr(), n, Sg[:, i].contiguous().data_ptr(),
                      X[i].data_ptr(), 1.0, 0.0, _s())
        idx = np.argsort(lam)
        return lam[idx], X[idx.tolist()].T.contiguous(), res[idx]

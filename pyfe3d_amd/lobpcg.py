"""Block eigensolver for the lowest modes of large models: LOBPCG
preconditioned by the smoothed-aggregation multigrid V-cycle.

Replaces the reference callers' ``eigsh(A=KC0uu, M=Muu, sigma=-1.,
which='LM')`` (tests/test_quad4_natural_freq.py:120-123) and
``eigsh(A=KGuu, M=KC0uu, which='SM')`` for linear buckling
(tests/test_quad4_linear_buckling_plate_mat_coord.py:184) where a banded
factorisation does not fit (the shift-invert Lanczos of
:mod:`pyfe3d_amd.eigen` factorises K on one workgroup), on the
Dirichlet-treated matrices (identity rows of K and zero rows of M or KG on
fixed DOFs; the basis is zero there):

* modal: K x = lambda M x, the k lowest lambda;
* buckling: K x = lambda (-KG) x, the k lowest positive lambda, solved as
  (-KG) x = mu K x for the k largest mu (lambda = 1 / mu), so that the
  B-matrix of the iteration is the positive definite K and the multigrid
  V-cycle of K is the preconditioner in both problems.

Each iteration (Knyazev's locally optimal block method): residuals
R = K X - M X Lambda, preconditioned W = V-cycle(R), Rayleigh-Ritz on
span[X, W, P] through the Gram matrices S K S^T and S M S^T
(pf3d_block_dot), whose small generalized problem is solved on the host
after dropping the M-Gram directions below 1e-13 of the largest (the basis
may become dependent near convergence); X, P and their K and M images are
updated by combination (pf3d_block_comb), so only W needs products with K
and M.
"""
import numpy as np
import scipy.linalg as sla
import torch

from . import _lib


def _s():
    return torch.cuda.current_stream().cuda_stream


class LOBPCG:
    """plan: BSR structure shared by K and the second matrix; k_val: K's
    BSR values; m_val: M's (problem 'modal') or KG's (problem 'buckling');
    fixed: bool per DOF; precond: object with apply(r, z) (e.g. an SAAMG
    built for the Dirichlet-treated K)."""

    def __init__(self, plan, k_val, m_val, fixed, precond, problem='modal'):
        if problem not in ('modal', 'buckling'):
            raise ValueError("problem must be 'modal' or 'buckling'")
        self.plan = plan
        self.problem = problem
        nb = plan.n_nodes
        self.n = 6 * nb
        dev = k_val.device
        fixed_t = torch.as_tensor(np.asarray(fixed, dtype=np.uint8), device=dev)
        pr, pc = plan.row_ptr.data_ptr(), plan.col_idx.data_ptr()
        K = k_val.clone()
        _lib.call('pf3d_bsr_dirichlet', nb, pr, pc, K.data_ptr(), fixed_t.data_ptr(), _s())
        G = m_val.clone() if problem == 'modal' else -m_val
        _lib.call('pf3d_bsr_dirichlet_diag', nb, pr, pc, G.data_ptr(), fixed_t.data_ptr(),
                  0.0, _s())
        # A x = theta B x: modal A = K, B = M (lowest theta); buckling
        # A = -KG, B = K (largest theta)
        self.A, self.B = (K, G) if problem == 'modal' else (G, K)
        self.largest = problem == 'buckling'
        self.free = (1 - fixed_t.to(torch.float64))
        self.precond = precond

    def _mul(self, val, x, y):
        """y = val x with compensated row sums: for the lowest modes K x is
        ~lambda M x, many digits below |K| |x|, and the Rayleigh quotients
        need it accurate to ~eps of itself."""
        _lib.call('pf3d_bsr_residual_dd', self.plan.n_nodes, self.plan.row_ptr.data_ptr(),
                  self.plan.col_idx.data_ptr(), val.data_ptr(), x.data_ptr(), None,
                  y.data_ptr(), _s())

    def _mul_block(self, val, X, Y):
        """Y[v] = val X[v] for the rows of X (compensated, each matrix block
        read once per 8 rows; the same bits as _mul per row)."""
        for v0 in range(0, X.shape[0], 8):
            nv = min(8, X.shape[0] - v0)
            _lib.call('pf3d_bsr_product_dd_mv', self.plan.n_nodes, self.plan.row_ptr.data_ptr(),
                      self.plan.col_idx.data_ptr(), val.data_ptr(), nv, self.n,
                      X[v0].data_ptr(), Y[v0].data_ptr(), _s())

    def _gram(self, A, B, out_host=True):
        p, q = A.shape[0], B.shape[0]
        C = torch.empty(p * q, dtype=torch.float64, device=A.device)
        _lib.call('pf3d_block_dot', self.n, p, A.data_ptr(), self.n, q, B.data_ptr(), self.n,
                  C.data_ptr(), self.work.data_ptr(), _s())
        return C.view(p, q).cpu().numpy()

    def _comb(self, S, C, Y):
        """Y (k, n) = C^T S for S (m, n), C (m, k) host array."""
        Cd = torch.as_tensor(np.ascontiguousarray(C), device=S.device)
        _lib.call('pf3d_block_comb', self.n, S.shape[0], S.data_ptr(), self.n, C.shape[1],
                  Cd.data_ptr(), Y.data_ptr(), self.n, _s())
        return Y

    def _orth(self, S, AS, BS, kb, m):
        """S[kb:m] -= X (X^T M S[kb:m]) (and the same combination on the K
        and M images), then unit M-norm per direction."""
        G = self._gram(BS[kb:m], S[:kb])          # (m-kb, kb) = S_i^T M X
        C = np.zeros((m, m - kb))
        C[:kb] = -G.T
        C[kb:] = np.eye(m - kb)
        n = self.n
        for arr in (S, AS, BS):
            tmp = self._comb(arr[:m], C, torch.empty((m - kb, n), dtype=torch.float64,
                                                     device=arr.device))
            arr[kb:m] = tmp
        nrm = np.sqrt(np.maximum(np.diag(self._gram(S[kb:m], BS[kb:m])), 1e-300))
        D = np.diag(1.0 / nrm)
        for arr in (S, AS, BS):
            tmp = self._comb(arr[kb:m], D, torch.empty((m - kb, n), dtype=torch.float64,
                                                       device=arr.device))
            arr[kb:m] = tmp

    def solve(self, k=6, block=None, tol=1e-7, maxiter=300, seed=0, log=None):
        """The k lowest eigenvalues (modal) or k lowest positive buckling
        load multipliers.  Converged when every one of the k residuals
        ||A x - theta B x|| <= tol * |theta| ||B x||.  Returns (lambda (k,)
        ascending, X (n, k) on the GPU, info)."""
        n = self.n
        dev = self.A.device
        kb = block or min(k + 2, 10)
        if 3 * kb > 32:
            raise ValueError('block size above 10')
        self.work = torch.empty(256 * 32 * 32, dtype=torch.float64, device=dev)
        g = torch.Generator(device='cpu').manual_seed(seed)
        X = (torch.rand((kb, n), generator=g, dtype=torch.float64) - 0.5).to(dev) * self.free
        S = torch.zeros((3 * kb, n), dtype=torch.float64, device=dev)   # [X, W, P]
        AS = torch.zeros_like(S)
        BS = torch.zeros_like(S)
        S[:kb] = X
        self._mul_block(self.A, S[:kb], AS[:kb])
        self._mul_block(self.B, S[:kb], BS[:kb])
        m = kb          # active vectors in S
        R = torch.empty((kb, n), dtype=torch.float64, device=dev)
        lam = None
        hist = []
        it = 0
        for it in range(maxiter + 1):
            # Rayleigh-Ritz on span S[:m]
            GA = self._gram(S[:m], AS[:m])
            GB = self._gram(S[:m], BS[:m])
            GA = 0.5 * (GA + GA.T)
            GB = 0.5 * (GB + GB.T)
            d, V = np.linalg.eigh(GB)
            keep = d > 1e-13 * d.max()
            T = V[:, keep] / np.sqrt(d[keep])
            theta, Y = np.linalg.eigh(T.T @ GA @ T)
            if self.largest:
                theta, Y = theta[::-1], Y[:, ::-1]
            C = T @ Y[:, :kb]                      # (m, kb), B-orthonormal Ritz vectors
            lam = theta[:kb]
            Xn = self._comb(S[:m], C, torch.empty((kb, n), dtype=torch.float64, device=dev))
            AXn = self._comb(AS[:m], C, torch.empty_like(Xn))
            BXn = self._comb(BS[:m], C, torch.empty_like(Xn))
            if m > kb:
                Cp = C.copy()
                Cp[:kb] = 0.                         # the W and P parts: new search directions
                Pn = self._comb(S[:m], Cp, torch.empty_like(Xn))
                APn = self._comb(AS[:m], Cp, torch.empty_like(Xn))
                BPn = self._comb(BS[:m], Cp, torch.empty_like(Xn))
            S[:kb], AS[:kb], BS[:kb] = Xn, AXn, BXn
            # residuals and their norms
            R.copy_(AXn)
            for i in range(kb):
                _lib.call('pf3d_axpby', n, -float(lam[i]), BXn[i].data_ptr(), 1.0,
                          R[i].data_ptr(), _s())
            rn = np.sqrt(np.diag(self._gram(R, R)))
            bn = np.sqrt(np.maximum(np.diag(self._gram(BXn, BXn)), 1e-300))
            rel = rn / (np.abs(lam) * bn)
            hist.append(float(rel[:k].max()))
            if log and it % 10 == 0:
                log('lobpcg %d: theta %s, max rel residual %.2e' % (it, lam[:3], rel[:k].max()))
            if rel[:k].max() <= tol or it == maxiter:
                break
            # W = T(R), then its images
            # zero on fixed DOFs: R is, and the V-cycle keeps them decoupled
            # (identity rows, zero prolongator rows); one multi-vector
            # V-cycle when the preconditioner has it
            if hasattr(self.precond, 'apply_block'):
                self.precond.apply_block(R, S[kb:2 * kb])
            else:
                for i in range(kb):
                    self.precond.apply(R[i], S[kb + i])
            self._mul_block(self.A, S[kb:2 * kb], AS[kb:2 * kb])
            self._mul_block(self.B, S[kb:2 * kb], BS[kb:2 * kb])
            if m > kb:
                S[2 * kb:], AS[2 * kb:], BS[2 * kb:] = Pn, APn, BPn
                m = 3 * kb
            else:
                m = 2 * kb
            # M-orthogonalise W and P against X (X^T M X = I) and scale each
            # direction to unit M-norm: keeps the Gram matrices well
            # conditioned when W and P shrink with the residuals
            self._orth(S, AS, BS, kb, m)
        vals = lam[:k] if not self.largest else 1. / lam[:k]
        order = np.argsort(vals)
        Xo = S[:kb][order.tolist()].T.contiguous()
        return vals[order], Xo, dict(iters=it, residual=rel[:k][order], history=hist,
                                     block=kb)

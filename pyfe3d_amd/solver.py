"""Static solve K u = f on the GPU: 6x6 block-Jacobi PCG in FP64, or FP32
PCG inside an FP64 iterative-refinement loop.

Replaces the reference callers' free-DOF slicing and scipy solve
(``KC0[bu, :][:, bu]`` then ``spsolve``/``cg``, e.g.
tests/test_quad4_static_point_load.py:84-102).  Fixed DOFs are removed by
P K P + (I - P) on the assembled BSR matrix, so u[fixed] = 0 and u[free]
is the solution of the sliced system.  Convergence is judged on the FP64
residual ||f - K u|| / ||f||.
"""
import ctypes

import numpy as np
import torch

from . import _lib


def _s(stream):
    return (stream if stream is not None else torch.cuda.current_stream()).cuda_stream


class StaticSystem:
    def __init__(self, plan, val, fixed, stream=None):
        """plan: BSRPlan; val: its BSR values (copied); fixed: bool per DOF."""
        self.plan = plan
        nb = plan.n_nodes
        self.nb = nb
        dev = val.device
        fixed = torch.as_tensor(np.asarray(fixed, dtype=np.uint8) if not torch.is_tensor(fixed)
                                else fixed.to(torch.uint8), device=dev).contiguous()
        if fixed.numel() != 6 * nb:
            raise ValueError('fixed must have one entry per DOF')
        self.fixed = fixed
        self.val = val.clone()
        s = _s(stream)
        _lib.call('pf3d_bsr_dirichlet', nb, plan.row_ptr.data_ptr(), plan.col_idx.data_ptr(),
                  self.val.data_ptr(), fixed.data_ptr(), s)
        self.dinv = torch.empty(nb * 36, dtype=torch.float64, device=dev)
        self.dinv32 = torch.empty(nb * 36, dtype=torch.float32, device=dev)
        bad = torch.zeros(1, dtype=torch.int32, device=dev)
        _lib.call('pf3d_bjacobi', nb, plan.row_ptr.data_ptr(), plan.col_idx.data_ptr(),
                  self.val.data_ptr(), self.dinv.data_ptr(), self.dinv32.data_ptr(),
                  bad.data_ptr(), s)
        self.val32 = torch.empty(self.val.numel(), dtype=torch.float32, device=dev)
        _lib.call('pf3d_to_f32', self.val.numel(), self.val.data_ptr(), self.val32.data_ptr(), s)
        if int(bad.item()):
            raise _lib.Pf3dError('%d singular 6x6 diagonal blocks' % int(bad.item()))
        nbytes = ctypes.c_int64()
        _lib.call('pf3d_pcg_bytes', nb, 0, ctypes.byref(nbytes))
        self._w64 = torch.empty(nbytes.value, dtype=torch.uint8, device=dev)
        _lib.call('pf3d_ir_bytes', nb, ctypes.byref(nbytes))
        self._wir = torch.empty(nbytes.value, dtype=torch.uint8, device=dev)

    def residual(self, b, x, r=None, stream=None):
        """r = b - K x with compensated accumulation (pf3d_bsr_residual_dd):
        accurate to ~eps of r itself, not ~eps |K| |x|."""
        if r is None:
            r = torch.empty_like(b)
        _lib.call('pf3d_bsr_residual_dd', self.nb, self.plan.row_ptr.data_ptr(),
                  self.plan.col_idx.data_ptr(), self.val.data_ptr(), x.data_ptr(), b.data_ptr(),
                  r.data_ptr(), _s(stream))
        return r

    def kmul_accurate(self, x, y=None, stream=None):
        """y = K x with compensated row sums (pf3d_bsr_residual_dd, b NULL)."""
        if y is None:
            y = torch.empty_like(x)
        _lib.call('pf3d_bsr_residual_dd', self.nb, self.plan.row_ptr.data_ptr(),
                  self.plan.col_idx.data_ptr(), self.val.data_ptr(), x.data_ptr(), None,
                  y.data_ptr(), _s(stream))
        return y

    def matvec(self, x, y=None, stream=None):
        if y is None:
            y = torch.empty_like(x)
        _lib.call('pf3d_bsr_spmv', self.nb, self.plan.row_ptr.data_ptr(),
                  self.plan.col_idx.data_ptr(), self.val.data_ptr(), x.data_ptr(), y.data_ptr(),
                  _s(stream))
        return y

    def rhs(self, f):
        """P f on the device (fixed DOFs zeroed)."""
        f = torch.as_tensor(f, dtype=torch.float64, device=self.val.device).clone()
        f[self.fixed.bool()] = 0.
        return f

    def factor_direct(self, stream=None):
        """Banded Cholesky of the system on the GPU (one workgroup), in the
        node numbering (given, or reverse Cuthill-McKee; host, once) with the
        narrower band.  Raises if the band is wider than the kernels accept."""
        import scipy.sparse as ss
        from scipy.sparse.csgraph import reverse_cuthill_mckee
        nb = self.nb
        if not hasattr(self, '_band'):
            rp = self.plan.row_ptr.cpu().numpy()
            ci = self.plan.col_idx.cpu().numpy()
            g = ss.csr_matrix((np.ones(ci.size), ci, rp), shape=(nb, nb))
            rows = np.repeat(np.arange(nb), np.diff(rp))
            best = None
            # the given numbering or reverse Cuthill-McKee, whichever band is narrower
            for pnode in (np.arange(nb, dtype=np.int64),
                          np.asarray(reverse_cuthill_mckee(g, symmetric_mode=True), np.int64)):
                pinv = np.empty_like(pnode)
                pinv[pnode] = np.arange(nb)
                w = 6 * int(np.abs(pinv[rows] - pinv[ci]).max()) + 5
                if best is None or w < best[0]:
                    best = (w, pnode, pinv)
            bw, pnode, pinv = best
            maxb = _lib.call('pf3d_band_max_width')
            if bw > maxb:
                raise ValueError('bandwidth %d after renumbering exceeds %d: use PCG' % (bw, maxb))
            dev = self.val.device
            self._band_bw = torch.tensor([bw], dtype=torch.int32, device=dev)
            self._band_off = torch.zeros(1, dtype=torch.int64, device=dev)
            self._band_dof0 = torch.tensor([0, 6 * nb], dtype=torch.int64, device=dev)
            self._band_var = torch.zeros(nb, dtype=torch.int32, device=dev)
            self._band_pinv = torch.as_tensor(pinv, device=dev)
            self._band_pnode = torch.as_tensor(pnode, device=dev)
            self._band_size = 6 * nb * (bw + 1)
            self._band = torch.empty(self._band_size, dtype=torch.float64, device=dev)
            self._band_work = torch.empty(6 * nb, dtype=torch.float64, device=dev)
            self._band_bad = torch.zeros(1, dtype=torch.int32, device=dev)
            self.bandwidth = bw
        _lib.call('pf3d_band_factor', 1, nb, None, self._band_dof0.data_ptr(),
                  self._band_bw.data_ptr(), self._band_off.data_ptr(), self._band_size,
                  self._band.data_ptr(), self.plan.row_ptr.data_ptr(),
                  self.plan.col_idx.data_ptr(), self.val.data_ptr(), self._band_var.data_ptr(),
                  self._band_pinv.data_ptr(), self._band_bad.data_ptr(), 0, 0, _s(stream))
        self._factored = True

    def setup_amg(self, xyz, log=None, precisions=('fp64',), graph=True, **kw):
        """Smoothed-aggregation multigrid for method='amg' (FP64 V-cycle)
        and/or 'amg32' (FP32 V-cycle): one setup from the node coordinates
        xyz (n_nodes, 3) (see pyfe3d_amd.amg; kept in self.amg_hierarchy),
        the V-cycle captured in a HIP graph unless graph=False.  Returns the
        last one built."""
        from .amg import SAAMG, setup_hierarchy
        if 'precision' in kw:
            raise TypeError("setup_amg takes precisions=('fp64',) / ('fp32',) / "
                            "('fp64', 'fp32'), not precision=")
        fixed = self.fixed.cpu().numpy().astype(bool)
        box = kw.pop('box', 3.0)
        coarse_nodes = kw.pop('coarse_nodes', 300)
        max_levels = kw.pop('max_levels', 8)
        hier = setup_hierarchy(self.plan.row_ptr, self.plan.col_idx, self.val, self.nb, fixed, xyz,
                               box, coarse_nodes, max_levels, log)
        self.amg_hierarchy = hier
        # captured CG iterations of an earlier preconditioner are stale
        self.__dict__.pop('_cg_iterations', None)
        m = None
        for prec in precisions:
            m = SAAMG(self.plan, self.val, fixed, xyz, precision=prec, hierarchy=hier, **kw)
            if graph:
                m.capture()
            if prec == 'fp64':
                self.amg = m
            else:
                self.amg32 = m
        return m

    def solve(self, f, method='fp64', rtol=1e-8, x0=None, maxiter=200000, inner_rtol=1e-4,
              max_outer=60, replace=50, refactor=False, fp32_maxiter=None, stream=None):
        """Returns (u, info) with info = dict(iters, outer, relres, converged);
        converged is relres <= rtol for the compensated FP64 relative
        residual ||f - K u|| / ||f|| (a solve that stops short of rtol is
        returned, flagged, not raised).

        method: 'fp64' (block-Jacobi PCG), 'mixed' (PCG on the FP32-stored
        matrix with FP64 vectors and a true FP64 residual replacing the
        recursive one every `replace` iterations; info['outer'] counts the
        replacements), 'ir' (FP64 iterative refinement around restarted
        FP32-matrix PCG solves to inner_rtol) or 'direct' (banded Cholesky,
        factor kept for later right-hand sides unless refactor), 'amg' (CG
        preconditioned by one smoothed-aggregation V-cycle per iteration,
        after setup_amg) or 'amg32' (the same with the V-cycle in FP32, after
        setup_amg(xyz, precisions=('fp32',)); when it stalls short of rtol, or
        after fp32_maxiter iterations if given, the FP64 V-cycle finishes from
        its u if setup_amg also built it, info['fp64_iters'] counting those
        iterations; maxiter bounds both phases together)."""
        b = self.rhs(f)
        x = torch.zeros_like(b) if x0 is None else x0.clone()
        pr, pc = self.plan.row_ptr.data_ptr(), self.plan.col_idx.data_ptr()
        relres = ctypes.c_double()
        if method == 'fp64':
            iters = ctypes.c_int64()
            _lib.call('pf3d_pcg', self.nb, pr, pc, self.val.data_ptr(), self.dinv.data_ptr(),
                      b.data_ptr(), x.data_ptr(), float(rtol), int(maxiter),
                      int(x0 is not None), self._w64.data_ptr(), self._w64.numel(),
                      ctypes.byref(iters), ctypes.byref(relres), _s(stream))
            info = dict(iters=iters.value, outer=0)
        elif method == 'mixed':
            stats = (ctypes.c_int64 * 3)()
            _lib.call('pf3d_pcg_mixed', self.nb, pr, pc, self.val.data_ptr(),
                      self.dinv.data_ptr(), self.val32.data_ptr(), self.dinv32.data_ptr(),
                      b.data_ptr(), x.data_ptr(), float(rtol), int(maxiter), int(replace),
                      int(x0 is not None), self._wir.data_ptr(), self._wir.numel(), stats,
                      ctypes.byref(relres), _s(stream))
            info = dict(iters=int(stats[0]) + int(stats[2]), outer=int(stats[1]),
                        fp64_iters=int(stats[2]))
        elif method == 'ir':
            stats = (ctypes.c_int64 * 2)()
            _lib.call('pf3d_ir_solve', self.nb, pr, pc, self.val.data_ptr(),
                      self.val32.data_ptr(), self.dinv32.data_ptr(), b.data_ptr(),
                      x.data_ptr(), float(rtol), float(inner_rtol), int(max_outer),
                      int(maxiter), int(x0 is not None), self._wir.data_ptr(),
                      self._wir.numel(), stats, ctypes.byref(relres), _s(stream))
            info = dict(iters=int(stats[1]), outer=int(stats[0]))
        elif method in ('amg', 'amg32'):
            from .amg import amg_pcg
            name = 'amg' if method == 'amg' else 'amg32'
            if getattr(self, name, None) is None:
                raise ValueError("method=%r needs setup_amg(xyz, precisions=(%r,)) first"
                                 % (method, 'fp64' if method == 'amg' else 'fp32'))
            budget = min(int(maxiter), 100000)
            first = budget if fp32_maxiter is None or method == 'amg' else \
                min(budget, int(fp32_maxiter))
            # the FP32 V-cycle gives out near the FP64 residual floor: leave
            # it after 25 iterations without halving the residual (the FP64
            # V-cycle then finishes), not the 400 the FP64 one is allowed
            u, info = amg_pcg(self, getattr(self, name), f, rtol=rtol, maxiter=first, x0=x0,
                              stall=25 if method == 'amg32' else 400)
            if method == 'amg32':
                info['fp64_iters'] = 0
                if (info['relres'] > rtol and info['iters'] < budget
                        and getattr(self, 'amg', None) is not None):
                    # the FP32 phase stalled or used its share: finish with
                    # the FP64 V-cycle from u, within what is left of maxiter
                    u, info64 = amg_pcg(self, self.amg, f, rtol=rtol,
                                        maxiter=budget - info['iters'], x0=u)
                    info = dict(iters=info['iters'] + info64['iters'],
                                outer=info['outer'] + info64['outer'],
                                relres=info64['relres'], fp64_iters=info64['iters'])
            info['converged'] = bool(info['relres'] <= rtol)
            return u, info
        elif method == 'direct':
            if refactor or not getattr(self, '_factored', False):
                self.factor_direct(stream=stream)
            _lib.call('pf3d_band_solve', 1, None, self._band_dof0.data_ptr(),
                      self._band_bw.data_ptr(), self._band_off.data_ptr(),
                      self._band.data_ptr(), self._band_pnode.data_ptr(), b.data_ptr(),
                      x.data_ptr(), self._band_work.data_ptr(), 0, _s(stream))
            if int(self._band_bad.item()):
                raise _lib.Pf3dError('band factorisation met a non-positive pivot')
            info = dict(iters=0, outer=0)
        else:
            raise ValueError(method)
        # true residual on the FP64 matrix (compensated products)
        r = self.residual(b, x, stream=stream)
        info['relres'] = float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(b))
        info['converged'] = bool(info['relres'] <= rtol)
        return x, info

    def solve_dd(self, f, method='amg', rtol=1e-8, u=None, inner_rtol=1e-3, max_steps=8,
                 maxiter=4000, log=None, stream=None):
        """Solution held as a double-double pair u = hi + lo (hi = fl(u)).

        On thin shells the rounding of an FP64 vector alone moves K u by
        ~eps |K| |u|, which can exceed rtol ||f||; hi + lo carries ~32
        significant digits.  Refinement: r = (f - K hi) - K lo, both products
        compensated (pf3d_bsr_residual_dd), K d = r solved to inner_rtol with
        `method`, lo += d and (hi, lo) renormalised (Fast2Sum), until the
        pair's relative residual is <= rtol.  ``u`` is a first FP64 solution
        (solved with `method` to rtol if None).

        Returns (hi, lo, info): info['relres'] is the pair's compensated
        relative residual, info['relres_fp64'] that of hi alone (the pair
        rounded to FP64), info['steps'] the refinement
        steps, info['iters'] all CG iterations, info['converged'] relres <=
        rtol."""
        b = self.rhs(f)
        nb_ = float(torch.linalg.vector_norm(b)) or 1.
        iters = 0
        if u is None:
            u, info0 = self.solve(f, method=method, rtol=rtol, maxiter=maxiter, stream=stream)
            iters += info0['iters']
        hi = u.clone()
        lo = torch.zeros_like(hi)
        r1 = torch.empty_like(hi)
        r = torch.empty_like(hi)
        steps = 0
        while True:
            self.residual(b, hi, r1, stream=stream)
            self.residual(r1, lo, r, stream=stream)
            rel = float(torch.linalg.vector_norm(r)) / nb_
            if log:
                log('solve_dd: step %d, pair relres %.3e' % (steps, rel))
            if rel <= rtol or steps >= max_steps:
                break
            d, info = self.solve(r, method=method, rtol=inner_rtol, maxiter=maxiter,
                                 stream=stream)
            iters += info['iters']
            steps += 1
            lo.add_(d)
            # TwoSum (exact for any magnitudes): hi + lo = s + e, s = fl(hi + lo)
            s = hi + lo
            bb = s - hi
            lo = (hi - (s - bb)) + (lo - bb)
            hi = s
        self.residual(b, hi, r1, stream=stream)
        rel64 = float(torch.linalg.vector_norm(r1)) / nb_
        return hi, lo, dict(relres=rel, relres_fp64=rel64, steps=steps, iters=iters,
                            converged=bool(rel <= rtol))

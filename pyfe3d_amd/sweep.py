"""Batched design sweep: many plate variants assembled and solved together.

The reference's design-optimisation tutorials re-assemble KC0 for every new
design (``update_KC0v_only=1``), convert with ``coo_matrix(...).tocsc()``,
slice the free DOFs and call ``spsolve``, one design at a time
(doc/source/tut-opt-lightweight-smeared-deflection-SLSQP.ipynb; the
static pattern of tests/test_quad4_static_point_load.py:40-102).  Here the
variants (structured Quad4 plates of uneven sizes, clamped edges, one
thickness per element) are laid out as one block-diagonal system:

* one KC0 launch assembles every element of every variant (values only);
* an element-block plan sums the COO into one BSR matrix (deterministic);
* Dirichlet rows/columns, block-Jacobi inverses and the FP32 copies are
  one launch each;
* one persistent launch solves every variant (a workgroup per variant,
  largest first), in FP64 PCG or FP32-inner / FP64 iterative refinement.

Nothing in :meth:`PlateSweep.step` synchronises with the host, so the whole
step can be captured in a graph and replayed (:meth:`PlateSweep.capture`).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from . import assembly as A
from .mesh import PlateMesh
from .shellprop import isotropic_plate
from .sparse import BlockPlan

# LDS panel width of the narrow factorisation shape: 74 KB of LDS, so two
# workgroups share a CU (the 448-wide panel takes one CU each)
NARROW_PANEL = 256


def split_order(bandwidths, sizes, narrow=NARROW_PANEL):
    """Processing order of the split banded factorisation: the systems
    whose bandwidth exceeds ``narrow`` first, then the others, largest
    (by ``sizes``) first within each group, ties in index order.  Returns
    (order as int32, number of wide systems)."""
    bw = np.asarray(bandwidths)
    by_size = np.argsort(-np.asarray(sizes), kind='stable')
    wide = bw[by_size] > narrow
    order = np.concatenate([by_size[wide], by_size[~wide]]).astype(np.int32)
    return order, int(wide.sum())


def _s(stream):
    return (stream if stream is not None else torch.cuda.current_stream()).cuda_stream


def smooth_field(nxe, nye, a, b, corr, rng):
    """Zero-mean, unit-variance smooth Gaussian random field on the element
    centroids of an nxe x nye plate: white noise filtered with a Gaussian
    kernel of standard deviation ``corr`` (length units)."""
    from scipy.ndimage import gaussian_filter
    w = rng.standard_normal((nxe, nye))
    g = gaussian_filter(w, sigma=(corr * nxe / a, corr * nye / b), mode='reflect')
    sd = g.std()
    return (g - g.mean()) / (sd if sd > 0 else 1.)


def _central(ne):
    """[lo, hi) of the interior node indices within the central 20% (at
    least one node) of a line of ne elements."""
    lo = max(1, int(np.floor(0.4 * ne)))
    hi = max(lo, min(ne - 1, int(np.ceil(0.6 * ne))))
    return lo, hi + 1


class PlateSweep:
    """n_var clamped Quad4 plates a x b with (nxe, nye) elements each."""

    def __init__(self, dims, a=1.0, b=0.7, E=70e9, nu=0.33, rho=2700., device='cuda',
                 stream=None):
        self.dims = [(int(p), int(q)) for p, q in dims]
        self.n_var = len(self.dims)
        self.a, self.b = a, b
        xs, conns, fixed, nodes, elems = [], [], [], [0], [0]
        for nxe, nye in self.dims:
            m = PlateMesh(nxe + 1, nye + 1, a, b)
            xs.append(m.x)
            conns.append(m.quads + nodes[-1])
            xx, yy = m.ncoords[:, 0], m.ncoords[:, 1]
            edge = (np.isclose(xx, 0) | np.isclose(xx, a) | np.isclose(yy, 0) | np.isclose(yy, b))
            fixed.append(np.repeat(edge, 6))
            nodes.append(nodes[-1] + m.n_nodes)
            elems.append(elems[-1] + m.quads.shape[0])
        self.node0 = np.array(nodes, dtype=np.int64)
        self.elem0 = np.array(elems, dtype=np.int64)
        self.n_nodes = int(self.node0[-1])
        self.n_elem = int(self.elem0[-1])
        dev = torch.device(device)
        self.device = dev
        self.x = torch.as_tensor(np.concatenate(xs), device=dev)
        self.conn = torch.as_tensor(6 * np.concatenate(conns), device=dev)
        self.fixed = torch.as_tensor(np.concatenate(fixed).astype(np.uint8), device=dev)
        self.row0 = torch.as_tensor(self.node0, device=dev)
        sizes = np.diff(self.node0)
        self.order = torch.as_tensor(np.argsort(-sizes, kind='stable').astype(np.int32),
                                     device=dev)
        self.plan = BlockPlan(self.conn, self.n_nodes, stream=stream)
        nnzb = self.plan.nnzb
        self.v = torch.empty(576 * self.n_elem, dtype=torch.float64, device=dev)
        self.val = torch.empty(nnzb * 36, dtype=torch.float64, device=dev)
        self.val32 = torch.empty(nnzb * 36, dtype=torch.float32, device=dev)
        self.dinv = torch.zeros(self.n_nodes * 36, dtype=torch.float64, device=dev)
        self.dinv32 = torch.zeros(self.n_nodes * 36, dtype=torch.float32, device=dev)
        self.singular = torch.zeros(1, dtype=torch.int32, device=dev)
        self.props = torch.empty((self.n_elem, A.SHELL_STRIDE), dtype=torch.float64, device=dev)
        unit = isotropic_plate(1.0, E, nu, rho=rho).record()
        self._unit = np.ascontiguousarray(unit, dtype=np.float64)
        self.pid = torch.arange(self.n_elem, dtype=torch.int32, device=dev)
        nbytes = ctypes.c_int64()
        _lib.call('pf3d_batch_bytes', self.n_nodes, ctypes.byref(nbytes))
        self.work = torch.empty(nbytes.value, dtype=torch.uint8, device=dev)
        self.f = torch.zeros(6 * self.n_nodes, dtype=torch.float64, device=dev)
        self.u = torch.zeros(6 * self.n_nodes, dtype=torch.float64, device=dev)
        self.iters = torch.zeros(self.n_var, dtype=torch.int32, device=dev)
        self.outer = torch.zeros(self.n_var, dtype=torch.int32, device=dev)
        self.relres = torch.zeros(self.n_var, dtype=torch.float64, device=dev)
        self.h = torch.full((self.n_elem,), 0.002, dtype=torch.float64, device=dev)
        self.graph = None

    # ---------------------------------------------------------------- inputs
    def random_designs(self, seed, h0=0.002, dh=0.001, hmin=0.0002, corr=0.2):
        """Seeded thickness fields h = h0 + dh * g (g: smooth Gaussian field,
        correlation length corr * a, clamped to h >= hmin) and one point
        load Fz = -1000 N at a random node of the central 20% of each plate.
        Returns the host arrays (h per element, f per DOF)."""
        rng = np.random.default_rng(seed)
        h = np.empty(self.n_elem)
        f = np.zeros(6 * self.n_nodes)
        for v, (nxe, nye) in enumerate(self.dims):
            g = smooth_field(nxe, nye, self.a, self.b, corr * self.a, rng)
            # element order of PlateMesh: i (x) major, j (y) minor
            h[self.elem0[v]:self.elem0[v + 1]] = np.maximum(h0 + dh * g.ravel(), hmin)
            i = rng.integers(*_central(nxe))
            j = rng.integers(*_central(nye))
            node = self.node0[v] + i * (nye + 1) + j
            f[6 * node + 2] = -1000.
        return h, f

    def set_thickness(self, h):
        """Per-element thickness (device or host array); the property
        records are rebuilt from it at the start of each step."""
        self.h.copy_(torch.as_tensor(h, dtype=torch.float64))

    def set_loads(self, f):
        """Nodal forces per DOF; fixed DOFs are zeroed (P f)."""
        self.f.copy_(torch.as_tensor(f, dtype=torch.float64))
        self.f.mul_(1 - self.fixed.to(torch.float64))

    # ------------------------------------------------------------- the step
    def assemble(self, stream=None, precond=True):
        """Property records from h -> KC0 of every variant -> BSR ->
        Dirichlet -> (precond: block Jacobi -> FP32 copies, which only the
        PCG modes read)."""
        s = _s(stream)
        _lib.call('pf3d_isotropic_records', self.n_elem, self.h.data_ptr(),
                  self._unit.ctypes.data, self.props.data_ptr(), s)
        A.shell_kc0('quad4', self.x, self.conn, None, None, self.v, self.props, pid=self.pid,
                    values_only=True, overwrite=True, stream=stream, check=False)
        pr, pc = self.plan.row_ptr.data_ptr(), self.plan.col_idx.data_ptr()
        self.plan.apply(self.v, out=self.val, stream=stream)
        _lib.call('pf3d_bsr_dirichlet', self.n_nodes, pr, pc, self.val.data_ptr(),
                  self.fixed.data_ptr(), s)
        if precond:
            self.precondition(stream=stream)

    def precondition(self, stream=None):
        """Block-Jacobi inverses (FP64 and FP32) and the FP32 copy of the
        assembled values, for the PCG modes."""
        s = _s(stream)
        pr, pc = self.plan.row_ptr.data_ptr(), self.plan.col_idx.data_ptr()
        _lib.call('pf3d_bjacobi', self.n_nodes, pr, pc, self.val.data_ptr(),
                  self.dinv.data_ptr(), self.dinv32.data_ptr(), self.singular.data_ptr(), s)
        _lib.call('pf3d_to_f32', self.val.numel(), self.val.data_ptr(), self.val32.data_ptr(), s)

    def solve(self, mode='mixed', rtol=1e-8, replace=50, max_iter=200000, x0=False,
              n_groups=0, threads=0, stream=None):
        """Solve every variant into self.u (x0: start from the current u).
        mode 'fp64': FP64 operator; 'mixed': FP32 operator and inverses with
        an FP64 true-residual replacement every ``replace`` iterations.
        Launch shape: n_groups persistent workgroups of ``threads`` threads;
        0, 0 takes the shape :meth:`autotune_pcg` chose for this mode, else
        the library's default (1024 threads, one workgroup per CU)."""
        mixed = mode == 'mixed'
        if mode not in ('mixed', 'fp64'):
            raise ValueError(mode)
        if not n_groups and not threads:
            n_groups, threads = getattr(self, 'pcg_config', {}).get(mode, (0, 0))
        _lib.call('pf3d_batch_solve', self.n_var, self.n_nodes, self.row0.data_ptr(),
                  self.order.data_ptr(), self.plan.row_ptr.data_ptr(),
                  self.plan.col_idx.data_ptr(), self.val.data_ptr(), self.dinv.data_ptr(),
                  self.val32.data_ptr() if mixed else None,
                  self.dinv32.data_ptr() if mixed else None, self.f.data_ptr(),
                  self.u.data_ptr(), 1 if x0 else 0, float(rtol), int(replace), int(max_iter),
                  self.work.data_ptr(), self.work.numel(), self.iters.data_ptr(),
                  self.outer.data_ptr(), self.relres.data_ptr(), int(n_groups), int(threads),
                  _s(stream))

    # ------------------------------------------------------- direct solves
    def _band_setup(self):
        """Node permutations (each plate numbered along its longer side, so
        the band spans the shorter one), bandwidths and band storage."""
        var, pinv, bws, sizes = [], [], [], []
        for v, (nxe, nye) in enumerate(self.dims):
            nx, ny = nxe + 1, nye + 1
            i, j = np.divmod(np.arange(nx * ny), ny)  # PlateMesh: node = i*ny + j
            new = i * ny + j if ny <= nx else j * nx + i
            pinv.append(self.node0[v] + new)
            var.append(np.full(nx * ny, v, np.int32))
            bws.append(6 * (min(nx, ny) + 1) + 5)
            sizes.append(6 * nx * ny * (bws[-1] + 1))
        bws = np.array(bws, np.int32)
        maxb = _lib.call('pf3d_band_max_width')
        if bws.max() > maxb:
            raise ValueError('plate too wide for the banded solver (bandwidth %d > %d)'
                             % (bws.max(), maxb))
        pinv = np.concatenate(pinv)
        pnode = np.empty_like(pinv)
        pnode[pinv] = np.arange(pinv.size)
        dev = self.device
        self.band_var = torch.as_tensor(np.concatenate(var), device=dev)
        self.band_pinv = torch.as_tensor(pinv, device=dev)
        self.band_pnode = torch.as_tensor(pnode, device=dev)
        self.band_bw = torch.as_tensor(bws, device=dev)
        self.band_n_host = 6 * np.array([(nxe + 1) * (nye + 1) for nxe, nye in self.dims], np.int64)
        self.band_bw_host = bws.astype(np.int64)
        boff = np.concatenate([[0], np.cumsum(sizes)])
        self.band_off = torch.as_tensor(boff[:-1].astype(np.int64), device=dev)
        self.band_size = int(boff[-1])
        self.band = torch.empty(self.band_size, dtype=torch.float64, device=dev)
        self.band_dof0 = torch.as_tensor(6 * self.node0, device=dev)
        self.band_work = torch.empty(6 * self.n_nodes, dtype=torch.float64, device=dev)
        self.band_bad = torch.zeros(1, dtype=torch.int32, device=dev)
        # processing order split in two groups, largest plates first in each:
        # bands too wide for the narrow LDS panel, then those that fit it
        split, self.band_n_wide = split_order(bws, np.diff(self.node0))
        self.band_order_split = torch.as_tensor(split, device=dev)

    def band_flops(self):
        """Floating-point operations of the banded Cholesky of every
        variant: column j updates the m_j = min(b, n - 1 - j) entries below
        it with m_j (m_j + 1) / 2 multiply-adds (square roots and divisions
        not counted)."""
        if not hasattr(self, 'band'):
            self._band_setup()
        n, b = self.band_n_host, np.minimum(self.band_bw_host, self.band_n_host - 1)
        full = np.maximum(n - b, 0) * b * (b + 1)
        tail = (b - 1) * b * (b + 1) // 3
        return int((full + tail).sum())

    def band_window_bytes(self, nf=32):
        """HBM bytes the 32-column panel factorisation moves: per panel, the
        panel's band entries read and written once and the trailing window's
        lower triangle (T (T + 1) / 2 entries, T = rows below the panel in
        reach) read and written once."""
        if not hasattr(self, 'band'):
            self._band_setup()
        total = 0
        for n, b in zip(self.band_n_host.tolist(), self.band_bw_host.tolist()):
            j0 = np.arange(0, n, nf)
            jn = np.minimum(nf, n - j0)
            rows = np.minimum(n - j0, b + jn)
            t = rows - jn
            total += int((t * (t + 1) // 2).sum()) * 16
        return total + 16 * int((self.band_n_host * (self.band_bw_host + 1)).sum())

    def _direct_config(self, config=None):
        """(factor threads, LDS panel width, solve threads, factor threads of
        the plates whose band fits the narrow panel or 0) of config, else
        of the chosen configuration, else the default."""
        c = tuple(config or getattr(self, 'direct_config', None) or (1024, 576, 1024))
        return c if len(c) == 4 else c + (0,)

    def factor_direct(self, config=None, stream=None, forward=False):
        """Band storage of every variant's stiffness matrix and its banded
        Cholesky factorisation in place (the first half of
        :meth:`solve_direct`).  With a narrow-band thread count in the
        configuration, the plates whose band fits the narrow LDS panel are
        factorised by a second launch of that shape (more workgroups per CU)
        at the same time as the wide ones.  forward: the factorisation also
        runs the forward substitution with the current loads (self.f),
        leaving y in self.band_work for :meth:`solve_direct`'s backward
        pass."""
        if not hasattr(self, 'band'):
            self._band_setup()
        ft, mb, _, sft = self._direct_config(config)
        s = _s(stream)
        _lib.call('pf3d_band_load', self.n_nodes, self.band_dof0.data_ptr(),
                  self.band_bw.data_ptr(), self.band_off.data_ptr(), self.band_size,
                  self.band.data_ptr(), self.plan.row_ptr.data_ptr(),
                  self.plan.col_idx.data_ptr(), self.val.data_ptr(), self.band_var.data_ptr(),
                  self.band_pinv.data_ptr(), s)
        fwd = ((self.band_pnode.data_ptr(), self.f.data_ptr(), self.band_work.data_ptr())
               if forward else (None, None, None))

        def chol(n_sys, ptr, threads, panel, st):
            _lib.call('pf3d_band_chol', n_sys, ptr, self.band_dof0.data_ptr(),
                      self.band_bw.data_ptr(), self.band_off.data_ptr(), self.band.data_ptr(),
                      self.band_bad.data_ptr(), int(threads), int(panel), *fwd, st)

        if not sft or mb <= NARROW_PANEL or self.band_n_wide == self.n_var:
            chol(self.n_var, self.order.data_ptr(), ft, mb, s)
            return
        nw = self.band_n_wide
        order = self.band_order_split
        # the two groups run concurrently (the narrow one on a side stream,
        # forked from and joined back to the caller's), so the narrow-band
        # plates fill the CUs the wide launch's tail leaves idle
        main = stream if stream is not None else torch.cuda.current_stream()
        if getattr(self, '_side', None) is None:
            self._side = torch.cuda.Stream(device=self.device)
        side = self._side
        side.wait_stream(main)
        chol(nw, order.data_ptr(), ft, mb, s)
        with torch.cuda.stream(side):
            chol(self.n_var - nw, order[nw:].data_ptr(), sft, NARROW_PANEL, side.cuda_stream)
        main.wait_stream(side)

    def solve_direct(self, rtol=1e-8, polish_iter=0, config=None, stream=None):
        """Banded Cholesky of every variant (one workgroup each) with the
        forward substitution fused into it, the backward substitution, then
        the FP64 true residual of each variant into self.relres (and, with
        polish_iter > 0, block-Jacobi PCG from the direct solution for any
        variant that misses rtol).  config: (factor threads, panel
        bandwidth, solve threads[, narrow-band factor threads]), default the
        one chosen by :meth:`autotune_direct`, else (1024, 576, 1024)."""
        if not hasattr(self, 'band'):
            self._band_setup()
        s = _s(stream)
        cfg = self._direct_config(config)
        st = cfg[2]
        self.factor_direct(config=cfg, stream=stream, forward=True)
        if polish_iter > 0:
            self.precondition(stream=stream)
        _lib.call('pf3d_band_backsolve', self.n_var, self.order.data_ptr(),
                  self.band_dof0.data_ptr(), self.band_bw.data_ptr(), self.band_off.data_ptr(),
                  self.band.data_ptr(), self.band_pnode.data_ptr(), self.band_work.data_ptr(),
                  self.u.data_ptr(), int(st), s)
        self.solve(mode='fp64', rtol=rtol, max_iter=polish_iter, x0=True, stream=stream)

    def autotune_key(self):
        """Autotune table key of this sweep's banded solver: Quad4 plates,
        panel class of the widest band, power-of-two buckets of the number
        of variants and of the mean DOF per variant, GPU architecture."""
        from . import autotune as AT
        if not hasattr(self, 'band'):
            self._band_setup()
        bmax = int(self.band_bw_host.max())
        panel = min(m for m in (256, 448, 576) if m >= bmax)
        bucket = (panel, AT.size_bucket(self.n_var),
                  AT.size_bucket(int(self.band_n_host.mean())))
        return AT.make_key('band_direct', 'quad4', bucket, AT.arch())

    def autotune_key_pcg(self, mode):
        """Autotune table key of this sweep's batched PCG in one mode:
        Quad4 plates, the mode, power-of-two buckets of the number of
        variants and of the mean DOF per variant, GPU architecture."""
        from . import autotune as AT
        bucket = (AT.size_bucket(self.n_var), AT.size_bucket(6 * self.n_nodes // self.n_var))
        return AT.make_key('batch_pcg_' + mode, 'quad4', bucket, AT.arch())

    def autotune_pcg(self, mode='fp64', log=None, repeats=1, use_table=True, probe_iters=200):
        """Launch shape of the persistent batched PCG (threads per workgroup
        x number of workgroups) for :meth:`solve` in this mode: the stored
        choice for the sweep's key if there is one (use_table), else every
        shape timed on probe_iters iterations of every variant (a residual
        target no system meets, so all do the same work per candidate) from
        the current designs, the fastest kept and recorded.  Needs the
        operator assembled and preconditioned (:meth:`assemble`).  Returns
        the table of timings (stored or measured)."""
        from . import autotune as AT
        key = self.autotune_key_pcg(mode)
        cfgs = getattr(self, 'pcg_config', {})
        self.pcg_config = cfgs
        if use_table:
            entry = AT.lookup(key)
            if entry is not None:
                c = entry['config']
                cfgs[mode] = (int(c['groups']), int(c['threads']))
                self.pcg_config_source = 'table'
                if log:
                    log('autotune %s: stored choice %s' % (key, cfgs[mode]))
                return [({k: v for k, v in t.items() if k != 'ms'}, t['ms'])
                        for t in entry.get('timings', [])]
        u0 = self.u.clone()

        def run(c):
            self.solve(mode=mode, rtol=1e-300, max_iter=probe_iters, n_groups=c['groups'],
                       threads=c['threads'])

        # one and two workgroups per CU at each width (the library's default
        # is one 1024-thread workgroup per CU)
        cands = [dict(threads=t, groups=k * 256 * (1024 // t)) for t in (256, 512, 1024)
                 for k in (1, 2)]
        best, table = AT.tune(cands, run, repeats=repeats, log=log)
        self.u.copy_(u0)
        cfgs[mode] = (best['groups'], best['threads'])
        self.pcg_config_source = 'measured'
        AT.record(key, dict(best), table)
        return table

    def autotune_direct(self, log=None, repeats=2, use_table=True):
        """Launch configuration of the banded solver (factor threads x LDS
        panel width, solve threads) for solve_direct / step('direct') /
        capture('direct'): the stored choice for this sweep's key if there is
        one (use_table), else every configuration that fits the sweep's
        widest band timed on the current designs, the fastest kept and
        recorded.  Returns the table of timings (stored or measured)."""
        from . import autotune as AT
        if not hasattr(self, 'band'):
            self._band_setup()
        key = self.autotune_key()
        if use_table:
            entry = AT.lookup(key)
            if entry is not None:
                c = entry['config']
                self.direct_config = (int(c['ft']), int(c['mb']), int(c['st']),
                                      int(c.get('sft', 0)))
                self.direct_config_source = 'table'
                if log:
                    log('autotune %s: stored choice %s' % (key, self.direct_config))
                return [({k: v for k, v in t.items() if k != 'ms'}, t['ms'])
                        for t in entry.get('timings', [])]
        bmax = int(self.band_bw.max())
        panels = [m for m in (256, 448, 576) if m >= bmax]

        def run(c):
            self.solve_direct(rtol=1.0, polish_iter=0,
                              config=(c['ft'], c['mb'], c['st'], c['sft']))

        # factorisation shape first (it dominates), then the solve's threads,
        # then (with plates on both sides of the narrow panel) a second
        # factorisation launch for the narrow-band plates
        cands = [dict(ft=ft, mb=mb, st=1024, sft=0) for mb in panels for ft in (256, 512, 1024)]
        best, table = AT.tune(cands, run, repeats=repeats, log=log)
        cands = [dict(best, st=st) for st in (256, 512)]
        best, table2 = AT.tune(cands + [best], run, repeats=repeats, log=log)
        table += table2
        if best['mb'] > NARROW_PANEL and 0 < self.band_n_wide < self.n_var:
            cands = [dict(best, sft=t) for t in (256, 512)]
            best, table3 = AT.tune(cands + [best], run, repeats=repeats, log=log)
            table += table3
        self.direct_config = (best['ft'], best['mb'], best['st'], best['sft'])
        self.direct_config_source = 'measured'
        AT.record(key, dict(best), table)
        return table

    def step(self, mode='mixed', rtol=1e-8, stream=None, **kw):
        self.assemble(stream=stream, precond=mode != 'direct')
        if mode == 'direct':
            self.solve_direct(rtol=rtol, stream=stream, **kw)
        else:
            self.solve(mode=mode, rtol=rtol, stream=stream, **kw)

    def capture(self, mode='mixed', rtol=1e-8, **kw):
        """Capture :meth:`step` in a graph; replay with ``self.graph.replay()``
        after new thicknesses (:meth:`set_thickness`) or loads (:meth:`set_loads`)."""
        from .amg import _capture_guard
        with _capture_guard():
            s = torch.cuda.Stream(device=self.device)
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                self.step(mode=mode, rtol=rtol, stream=s, **kw)  # warm-up outside the graph
            torch.cuda.current_stream().wait_stream(s)
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self.step(mode=mode, rtol=rtol, **kw)
        self.graph = g
        return g

    def check(self):
        """Raise if a diagonal block was singular or a band factorisation
        met a non-positive pivot."""
        if int(self.singular.item()):
            raise _lib.Pf3dError('%d singular diagonal blocks' % int(self.singular.item()))
        if hasattr(self, 'band_bad') and int(self.band_bad.item()):
            raise _lib.Pf3dError('%d band factorisations failed' % int(self.band_bad.item()))

    def variant(self, v):
        """(node range, element range) of variant v."""
        return (slice(int(self.node0[v]), int(self.node0[v + 1])),
                slice(int(self.elem0[v]), int(self.elem0[v + 1])))

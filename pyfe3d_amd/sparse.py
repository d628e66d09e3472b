"""COO slots -> 6x6 BSR matrix on the GPU (deterministic duplicate sum).

Replaces ``coo_matrix((KC0v, (KC0r, KC0c)), shape=(N, N)).tocsc()`` of the
reference's callers (tests/test_quad4_static_point_load.py:80).  The plan is
built once per mesh from the row/column slots; ``apply`` runs after every
re-assembly (e.g. the design-sweep pattern of the tutorials, where only the
values change).
"""
import numpy as np
import torch

from . import _lib


def _s(stream):
    return (stream if stream is not None else torch.cuda.current_stream()).cuda_stream


class BSRPlan:
    """Sorted-slot plan for one COO pattern; n_nodes = N_dof / 6."""

    def __init__(self, r, c, n_nodes, stream=None):
        assert r.is_cuda and c.is_cuda and r.dtype == torch.int64 and c.dtype == torch.int64
        n = r.numel()
        if c.numel() != n or n == 0:
            raise ValueError('r and c must be non-empty and of equal size')
        lo = min(int(r.min()), int(c.min()))
        hi = max(int(r.max()), int(c.max()))
        if lo < 0 or hi >= 6 * n_nodes:
            raise ValueError('row/col indices outside [0, %d)' % (6 * n_nodes))
        self.n_slots, self.n_nodes = n, int(n_nodes)
        L = _lib.load()
        import ctypes
        nbytes = ctypes.c_int64()
        _lib.call('pf3d_bsr_plan_bytes', n, self.n_nodes, ctypes.byref(nbytes))
        work = torch.empty(nbytes.value, dtype=torch.uint8, device=r.device)
        counts = (ctypes.c_int64 * 2)()
        _lib.call('pf3d_bsr_plan_sort', n, r.data_ptr(), c.data_ptr(), self.n_nodes,
                  work.data_ptr(), nbytes.value, counts, _s(stream))
        self.n_unique, self.nnzb = int(counts[0]), int(counts[1])
        dev = r.device
        self.perm = torch.empty(n, dtype=torch.int64, device=dev)
        self.seg = torch.empty(self.n_unique + 1, dtype=torch.int64, device=dev)
        self.dest = torch.empty(self.n_unique, dtype=torch.int64, device=dev)
        self.row_ptr = torch.empty(self.n_nodes + 1, dtype=torch.int32, device=dev)
        self.col_idx = torch.empty(self.nnzb, dtype=torch.int32, device=dev)
        _lib.call('pf3d_bsr_plan_emit', n, self.n_nodes, work.data_ptr(), nbytes.value,
                  self.n_unique, self.nnzb, self.perm.data_ptr(), self.seg.data_ptr(),
                  self.dest.data_ptr(), self.row_ptr.data_ptr(), self.col_idx.data_ptr(),
                  _s(stream))
        del work

    def apply(self, v, out=None, stream=None):
        """BSR values (nnzb*36, float64) = per-entry sums of the slot values."""
        assert v.is_cuda and v.dtype == torch.float64 and v.numel() == self.n_slots
        if out is None:
            out = torch.empty(self.nnzb * 36, dtype=torch.float64, device=v.device)
        _lib.call('pf3d_bsr_apply', self.n_unique, self.perm.data_ptr(), self.seg.data_ptr(),
                  self.dest.data_ptr(), v.data_ptr(), self.nnzb, out.data_ptr(), _s(stream))
        return out

    def to_scipy(self, val):
        """Host copy as scipy.sparse.bsr_matrix (for checks and plots)."""
        import scipy.sparse as ss
        data = val.detach().cpu().numpy().reshape(self.nnzb, 6, 6)
        return ss.bsr_matrix((data, self.col_idx.cpu().numpy(), self.row_ptr.cpu().numpy()),
                             shape=(6 * self.n_nodes, 6 * self.n_nodes))


class BlockPlan:
    """Element-block plan: the BSR structure from the element topology.

    For COO values written by the element entry points (block (I, J) of
    element e at ``init_k[e] + 36*nn*I + 6*J``, row stride ``6*nn``,
    quad4.pyx:1287-1293) the plan stores one slot offset per element block
    instead of one index per slot (16x fewer index bytes for Quad4) and
    ``apply`` reads 6 contiguous values per block row.  Sums run in
    ascending slot order, so the values are bit-identical to
    :class:`BSRPlan` on the same COO.  Exposes the same ``row_ptr``,
    ``col_idx``, ``nnzb``, ``n_nodes`` and ``apply``."""

    def __init__(self, conn, n_nodes, init_k=None, stream=None):
        import ctypes
        assert conn.is_cuda and conn.dtype == torch.int64 and conn.dim() == 2
        self.n_elem, self.nn = conn.shape
        self.n_nodes = int(n_nodes)
        if self.n_elem == 0:
            raise ValueError('empty connectivity')
        lo, hi = int(conn.min()), int(conn.max())
        if lo < 0 or hi >= 6 * self.n_nodes or bool((conn % 6 != 0).any()):
            raise ValueError('conn must hold DOF offsets 6*node with node < %d' % self.n_nodes)
        size = 36 * self.nn * self.nn
        if init_k is not None:
            assert init_k.is_cuda and init_k.dtype == torch.int64 and init_k.numel() == self.n_elem
            if int(init_k.min()) < 0:
                raise ValueError('negative init_k')
            self.n_slots = int(init_k.max()) + size  # slots v must hold
        else:
            self.n_slots = size * self.n_elem
        nbytes = ctypes.c_int64()
        _lib.call('pf3d_block_plan_bytes', self.n_elem, self.nn, self.n_nodes,
                  ctypes.byref(nbytes))
        work = torch.empty(nbytes.value, dtype=torch.uint8, device=conn.device)
        nnzb = ctypes.c_int64()
        pk = None if init_k is None else init_k.data_ptr()
        _lib.call('pf3d_block_plan_sort', self.n_elem, self.nn, conn.data_ptr(), pk,
                  self.n_nodes, work.data_ptr(), nbytes.value, ctypes.byref(nnzb), _s(stream))
        self.nnzb = int(nnzb.value)
        dev = conn.device
        self.src = torch.empty(self.n_elem * self.nn * self.nn, dtype=torch.int64, device=dev)
        self.src_ptr = torch.empty(self.nnzb + 1, dtype=torch.int64, device=dev)
        self.row_ptr = torch.empty(self.n_nodes + 1, dtype=torch.int32, device=dev)
        self.col_idx = torch.empty(self.nnzb, dtype=torch.int32, device=dev)
        _lib.call('pf3d_block_plan_emit', self.n_elem, self.nn, pk, self.n_nodes, work.data_ptr(),
                  nbytes.value, self.nnzb, self.src.data_ptr(), self.src_ptr.data_ptr(),
                  self.row_ptr.data_ptr(), self.col_idx.data_ptr(), _s(stream))
        (stream if stream is not None else torch.cuda.current_stream()).synchronize()
        del work

    def apply(self, v, out=None, accumulate=False, stream=None):
        """BSR values (nnzb*36, float64): sums of the element blocks of v."""
        assert v.is_cuda and v.dtype == torch.float64 and v.is_contiguous()
        if v.numel() < self.n_slots:
            raise ValueError('v has %d slots, the plan needs %d' % (v.numel(), self.n_slots))
        if out is None:
            if accumulate:
                raise ValueError('accumulate needs out')
            out = torch.empty(self.nnzb * 36, dtype=torch.float64, device=v.device)
        _lib.call('pf3d_block_apply', self.nnzb, self.src_ptr.data_ptr(), self.src.data_ptr(),
                  self.nn, v.data_ptr(), out.data_ptr(), 1 if accumulate else 0, _s(stream))
        return out

    to_scipy = BSRPlan.to_scipy

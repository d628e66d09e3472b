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

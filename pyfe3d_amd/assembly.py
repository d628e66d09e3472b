"""Batched element assembly on the GPU (one C-ABI call per element type and matrix).

These functions are the drop-in for the reference's per-element loops, e.g.
(tests/test_quad4_static_point_load.py:52-78 of pyfe3d)::

    for each element: quad.update_rotation_matrix(x); quad.update_probe_xe(x)
                      quad.update_KC0(KC0r, KC0c, KC0v, prop)

becomes one ``shell_kc0('quad4', x, conn, KC0r, KC0c, KC0v, props)`` call.
All array arguments are torch tensors on the GPU (torch is used only as the
device allocator); slot positions, index values and accumulation semantics
are those of the reference.  With ``check=True`` (default) the node offsets,
slot offsets and property indices are range-checked on the device before the
launch (a few small reductions and one synchronisation); callers that
re-assemble the same validated arrays in a loop pass ``check=False``.
"""
import torch

from . import _lib

ETYPE = {'quad4': 0, 'quad4r': 1, 'tria3r': 2, 'beamc': 3}
MATRIX = {'KC0': 0, 'KG': 1, 'KG_stress': 2, 'M': 3}
NNODES = {'quad4': 4, 'quad4r': 4, 'tria3r': 3, 'beamc': 2}
VALUES_ONLY = 1
OVERWRITE = 2
SHELL_STRIDE = 32
BEAM_STRIDE = 16


def _ptr(t, dtype, name, numel=None, optional=False):
    if t is None:
        if optional:
            return None
        raise ValueError('%s is required' % name)
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError('%s must be a torch tensor on the GPU' % name)
    if t.dtype != dtype:
        raise TypeError('%s must be %s, got %s' % (name, dtype, t.dtype))
    if not t.is_contiguous():
        raise ValueError('%s must be contiguous' % name)
    if numel is not None and t.numel() < numel:
        raise ValueError('%s has %d entries, needs at least %d' % (name, t.numel(), numel))
    return t.data_ptr()


def _stream(stream):
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def slot_count(etype, matrix='KC0', mtype=0):
    """Slots per element (the reference's *_SPARSE_SIZE, or the used prefix
    of M_SPARSE_SIZE for lumped masses)."""
    return _lib.call('pf3d_slot_count', ETYPE[etype], MATRIX[matrix], mtype)


def _flags(values_only, overwrite):
    return (VALUES_ONLY if values_only else 0) | (OVERWRITE if overwrite else 0)


def _common(etype, x, conn, init_k, evec, check=True):
    nn = NNODES[etype]
    if conn.dim() != 2 or conn.shape[1] != nn:
        raise ValueError('conn must have shape (n_elem, %d)' % nn)
    E = conn.shape[0]
    if check:
        _check_nodes(conn, x)
    return (E, _ptr(x, torch.float64, 'x'), _ptr(conn, torch.int64, 'conn'),
            _ptr(init_k, torch.int64, 'init_k', E, optional=True),
            _ptr(evec, torch.float64, 'xmat/vxy', 3 * E, optional=True))


def _outputs(r, c, v, need, values_only):
    pv = _ptr(v, torch.float64, 'v')
    if values_only:
        return None, None, pv
    return _ptr(r, torch.int64, 'r'), _ptr(c, torch.int64, 'c'), pv


def _check_span(n_slots_needed, r, c, v, values_only, init_k, check=True):
    """Every slot an element will write must lie inside r, c and v."""
    if not check:
        return
    outs = (('v', v),) + (() if values_only else (('r', r), ('c', c)))
    if init_k is None:
        for name, t in outs:
            if t.numel() < n_slots_needed:
                raise ValueError('%s has %d slots, needs %d' % (name, t.numel(), n_slots_needed))
    elif init_k.numel():
        size = n_slots_needed // max(1, init_k.numel())
        lo, hi = int(init_k.min()), int(init_k.max())
        for name, t in outs:
            if lo < 0 or hi + size > t.numel():
                raise ValueError('init_k spans slots [%d, %d) but %s has %d'
                                 % (lo, hi + size, name, t.numel()))


def _check_nodes(conn, x, u=None):
    """DOF offsets must address coordinates (x[c/2 .. c/2+2]) and u[c .. c+5]."""
    if conn.numel() == 0:
        return
    lo, hi = int(conn.min()), int(conn.max())
    if lo < 0 or hi // 2 + 3 > x.numel():
        raise ValueError('conn DOF offsets [%d, %d] outside x (%d coordinates)'
                         % (lo, hi, x.numel()))
    if u is not None and hi + 6 > u.numel():
        raise ValueError('conn DOF offsets up to %d outside u (%d entries)' % (hi, u.numel()))


def _check_pid(pid, props):
    if pid is not None and pid.numel():
        lo, hi = int(pid.min()), int(pid.max())
        if lo < 0 or hi >= props.shape[0]:
            raise ValueError('pid in [%d, %d] but %d property records' % (lo, hi, props.shape[0]))


def _props(props, pid, E, stride, check=True):
    if props.dim() != 2 or props.shape[1] != stride:
        raise ValueError('props must have shape (n_props, %d)' % stride)
    pp = _ptr(props, torch.float64, 'props')
    ppid = _ptr(pid, torch.int32, 'pid', E, optional=True)
    if check:
        _check_pid(pid, props)
    return pp, ppid


def geometry(etype, x, conn, evec=None, stream=None):
    """Element axes/material rotation/area (quad4.pyx:491-752 and peers).

    Returns (geom (E, 14): r11..r33, m11, m12, m21, m22, area|length;
             xe (E, 3*nn) local coordinates)."""
    E, px, pc, _, pe = _common(etype, x, conn, None, evec)
    geom = torch.empty((E, 14), dtype=torch.float64, device=x.device)
    xe = torch.empty((E, 3 * NNODES[etype]), dtype=torch.float64, device=x.device)
    _lib.call('pf3d_geometry', ETYPE[etype], E, px, pc, pe, geom.data_ptr(), xe.data_ptr(),
              _stream(stream))
    return geom, xe


def shell_kc0(etype, x, conn, r, c, v, props, pid=None, init_k=None, xmat=None, k6rot=None,
              alpha=None, hg=None, values_only=False, overwrite=False, stream=None, check=True):
    """KC0 of a batch of Quad4/Quad4R/Tria3R elements (update_KC0)."""
    E, px, pc, pk, pe = _common(etype, x, conn, init_k, xmat, check)
    _check_span(slot_count(etype, 'KC0') * E, r, c, v, values_only, init_k, check)
    pr, pcc, pv = _outputs(r, c, v, True, values_only)
    pp, ppid = _props(props, pid, E, SHELL_STRIDE, check)
    _lib.call('pf3d_shell_kc0', ETYPE[etype], E, px, pc, pk, pe, ppid, pp,
              _ptr(k6rot, torch.float64, 'k6rot', E, optional=True),
              _ptr(alpha, torch.float64, 'alpha', E, optional=True),
              _ptr(hg, torch.float64, 'hg', 5 * E, optional=True),
              pr, pcc, pv, _flags(values_only, overwrite), _stream(stream))


def shell_kg(etype, x, conn, u, r, c, v, props, pid=None, init_k=None, xmat=None,
             values_only=False, overwrite=False, stream=None, check=True):
    """KG from global displacements (update_probe_ue + update_KG)."""
    E, px, pc, pk, pe = _common(etype, x, conn, init_k, xmat, check)
    _check_span(slot_count(etype, 'KG') * E, r, c, v, values_only, init_k, check)
    pr, pcc, pv = _outputs(r, c, v, True, values_only)
    pp, ppid = _props(props, pid, E, SHELL_STRIDE, check)
    if check:
        _check_nodes(conn, x, u)
    _lib.call('pf3d_shell_kg', ETYPE[etype], E, px, pc, pk, pe, ppid, pp,
              _ptr(u, torch.float64, 'u'), pr, pcc, pv, _flags(values_only, overwrite),
              _stream(stream))


def shell_kg_stress(etype, x, conn, stress, r, c, v, init_k=None, xmat=None,
                    values_only=False, overwrite=False, stream=None, check=True):
    """KG for a given membrane stress state (update_KG_given_stress).

    ``stress`` has 3 entries (one state for all elements) or (E, 3)."""
    E, px, pc, pk, pe = _common(etype, x, conn, init_k, xmat, check)
    _check_span(slot_count(etype, 'KG_stress') * E, r, c, v, values_only, init_k, check)
    pr, pcc, pv = _outputs(r, c, v, True, values_only)
    if stress.numel() == 3:
        per = 0
    elif stress.numel() == 3 * E:
        per = 1
    else:
        raise ValueError('stress must have 3 or 3*n_elem entries')
    _lib.call('pf3d_shell_kg_stress', ETYPE[etype], E, px, pc, pk, pe,
              _ptr(stress, torch.float64, 'stress', 3), per, pr, pcc, pv,
              _flags(values_only, overwrite), _stream(stream))


def shell_m(etype, x, conn, r, c, v, props, mtype=0, pid=None, init_k=None, xmat=None,
            values_only=False, overwrite=False, stream=None, check=True):
    """Mass matrix (update_M), mtype 0 consistent, 1 reduced, 2 lumped."""
    E, px, pc, pk, pe = _common(etype, x, conn, init_k, xmat, check)
    _check_span(slot_count(etype, 'M', mtype) * E, r, c, v, values_only, init_k, check)
    pr, pcc, pv = _outputs(r, c, v, True, values_only)
    pp, ppid = _props(props, pid, E, SHELL_STRIDE, check)
    _lib.call('pf3d_shell_m', ETYPE[etype], E, px, pc, pk, pe, ppid, pp, int(mtype),
              pr, pcc, pv, _flags(values_only, overwrite), _stream(stream))


def beam_kc0(x, conn, vxy, r, c, v, props, pid=None, init_k=None, values_only=False,
             overwrite=False, stream=None, check=True):
    """BeamC KC0 of a batch (update_rotation_matrix(vxy, x) + update_KC0)."""
    E, px, pc, pk, pe = _common('beamc', x, conn, init_k, vxy, check)
    if pe is None:
        raise ValueError('vxy is required')
    _check_span(144 * E, r, c, v, values_only, init_k, check)
    pr, pcc, pv = _outputs(r, c, v, True, values_only)
    pp, ppid = _props(props, pid, E, BEAM_STRIDE, check)
    _lib.call('pf3d_beam_kc0', E, px, pc, pk, pe, ppid, pp, pr, pcc, pv,
              _flags(values_only, overwrite), _stream(stream))


def beam_kg(x, conn, vxy, u, r, c, v, props, pid=None, init_k=None, values_only=False,
            overwrite=False, stream=None, check=True):
    """BeamC KG from global displacements (update_probe_ue + update_KG)."""
    E, px, pc, pk, pe = _common('beamc', x, conn, init_k, vxy, check)
    if pe is None:
        raise ValueError('vxy is required')
    _check_span(144 * E, r, c, v, values_only, init_k, check)
    pr, pcc, pv = _outputs(r, c, v, True, values_only)
    pp, ppid = _props(props, pid, E, BEAM_STRIDE, check)
    if check:
        _check_nodes(conn, x, u)
    _lib.call('pf3d_beam_kg', E, px, pc, pk, pe, ppid, pp, _ptr(u, torch.float64, 'u'),
              pr, pcc, pv, _flags(values_only, overwrite), _stream(stream))


def beam_m(x, conn, vxy, r, c, v, props, mtype=0, pid=None, init_k=None, values_only=False,
           overwrite=False, stream=None, check=True):
    """BeamC mass matrix (update_M), mtype 0 consistent, 1 lumped."""
    E, px, pc, pk, pe = _common('beamc', x, conn, init_k, vxy, check)
    if pe is None:
        raise ValueError('vxy is required')
    _check_span(slot_count('beamc', 'M', mtype) * E, r, c, v, values_only, init_k, check)
    pr, pcc, pv = _outputs(r, c, v, True, values_only)
    pp, ppid = _props(props, pid, E, BEAM_STRIDE, check)
    _lib.call('pf3d_beam_m', E, px, pc, pk, pe, ppid, pp, int(mtype), pr, pcc, pv,
              _flags(values_only, overwrite), _stream(stream))

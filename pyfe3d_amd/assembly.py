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
import ctypes

import torch

from . import _lib

ETYPE = {'quad4': 0, 'quad4r': 1, 'tria3r': 2, 'beamc': 3, 'truss': 4, 'spring': 5,
         'beamlr': 6}
MATRIX = {'KC0': 0, 'KG': 1, 'KG_stress': 2, 'M': 3, 'KA_beta': 4, 'KA_gamma': 5, 'CA': 6}
NNODES = {'quad4': 4, 'quad4r': 4, 'tria3r': 3, 'beamc': 2, 'truss': 2, 'spring': 2,
          'beamlr': 2}
VALUES_ONLY = 1
OVERWRITE = 2
SHELL_STRIDE = 32
BEAM_STRIDE = 16
SPRING_STRIDE = 6


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


def nonfinite_elements(v, n_elem, size, init_k=None, stream=None):
    """Number of elements whose `size` slots (from init_k[e], or size*e)
    in the value array v hold inf or nan (pf3d_nonfinite_elements): the
    validation pass for degenerate elements, which the reference's
    elements write without a check."""
    if v.dtype != torch.float64 or not v.is_cuda:
        raise TypeError('v must be a float64 device tensor')
    if init_k is not None:
        if init_k.dtype != torch.int64 or init_k.numel() != n_elem:
            raise ValueError('init_k must hold one int64 per element')
        if n_elem and (int(init_k.min()) < 0 or int(init_k.max()) + size > v.numel()):
            raise ValueError('init_k slots outside v')
    elif size * n_elem > v.numel():
        raise ValueError('v holds fewer than size * n_elem slots')
    count = torch.zeros(1, dtype=torch.int64, device=v.device)
    s = (stream if stream is not None else torch.cuda.current_stream()).cuda_stream
    _lib.call('pf3d_nonfinite_elements', int(n_elem), int(size),
              None if init_k is None else init_k.data_ptr(), v.data_ptr(), count.data_ptr(), s)
    return int(count.item())


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
    """Element axes/material rotation/area (quad4.pyx:491-752 and peers);
    evec: xmat (shells), vxy (beams), None (trusses), (E, 6) axes (springs).

    Returns (geom (E, 14): r11..r33, m11, m12, m21, m22, area|length;
             xe (E, 3*nn) local coordinates)."""
    E, px, pc, _, pe = _common(etype, x, conn, None, evec)
    if etype == 'spring':
        pe = _spring_axes(evec, E)
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


def shell_piston(etype, matrix, x, conn, r, c, v, init_k=None, values_only=False,
                 overwrite=False, stream=None, check=True):
    """Piston-theory aerodynamic matrix of a batch of Quad4 / Quad4R
    elements: matrix 'KA_beta', 'KA_gamma' or 'CA' (update_KA_beta,
    update_KA_gamma, update_CA; quad4.pyx:9491-11917, quad4r.pyx:12789)."""
    if etype not in ('quad4', 'quad4r'):
        raise ValueError('piston theory is defined for quad4 and quad4r')
    if matrix not in ('KA_beta', 'KA_gamma', 'CA'):
        raise ValueError('matrix must be KA_beta, KA_gamma or CA')
    E, px, pc, pk, _ = _common(etype, x, conn, init_k, None, check)
    _check_span(144 * E, r, c, v, values_only, init_k, check)
    pr, pcc, pv = _outputs(r, c, v, True, values_only)
    _lib.call('pf3d_shell_piston', ETYPE[etype], MATRIX[matrix], E, px, pc, pk, pr, pcc, pv,
              _flags(values_only, overwrite), _stream(stream))


def _beam_entry(etype, name):
    if etype not in ('beamc', 'beamlr'):
        raise ValueError('beam etype must be beamc or beamlr, got %r' % etype)
    return ('pf3d_beam_' if etype == 'beamc' else 'pf3d_beamlr_') + name


def beam_kc0(x, conn, vxy, r, c, v, props, pid=None, init_k=None, values_only=False,
             overwrite=False, stream=None, check=True, etype='beamc'):
    """BeamC (or BeamLR, etype='beamlr') KC0 of a batch
    (update_rotation_matrix(vxy, x) + update_KC0)."""
    fn = _beam_entry(etype, 'kc0')
    E, px, pc, pk, pe = _common(etype, x, conn, init_k, vxy, check)
    if pe is None:
        raise ValueError('vxy is required')
    _check_span(144 * E, r, c, v, values_only, init_k, check)
    pr, pcc, pv = _outputs(r, c, v, True, values_only)
    pp, ppid = _props(props, pid, E, BEAM_STRIDE, check)
    _lib.call(fn, E, px, pc, pk, pe, ppid, pp, pr, pcc, pv,
              _flags(values_only, overwrite), _stream(stream))


def beam_kg(x, conn, vxy, u, r, c, v, props, pid=None, init_k=None, values_only=False,
            overwrite=False, stream=None, check=True, etype='beamc'):
    """BeamC (or BeamLR) KG from global displacements (update_probe_ue +
    update_KG): 144 slots for BeamC, 36 (rotations) for BeamLR."""
    fn = _beam_entry(etype, 'kg')
    E, px, pc, pk, pe = _common(etype, x, conn, init_k, vxy, check)
    if pe is None:
        raise ValueError('vxy is required')
    _check_span(slot_count(etype, 'KG') * E, r, c, v, values_only, init_k, check)
    pr, pcc, pv = _outputs(r, c, v, True, values_only)
    pp, ppid = _props(props, pid, E, BEAM_STRIDE, check)
    if check:
        _check_nodes(conn, x, u)
    _lib.call(fn, E, px, pc, pk, pe, ppid, pp, _ptr(u, torch.float64, 'u'),
              pr, pcc, pv, _flags(values_only, overwrite), _stream(stream))


def beam_m(x, conn, vxy, r, c, v, props, mtype=0, pid=None, init_k=None, values_only=False,
           overwrite=False, stream=None, check=True, etype='beamc'):
    """BeamC (or BeamLR) mass matrix (update_M), mtype 0 consistent, 1 lumped."""
    fn = _beam_entry(etype, 'm')
    E, px, pc, pk, pe = _common(etype, x, conn, init_k, vxy, check)
    if pe is None:
        raise ValueError('vxy is required')
    _check_span(slot_count(etype, 'M', mtype) * E, r, c, v, values_only, init_k, check)
    pr, pcc, pv = _outputs(r, c, v, True, values_only)
    pp, ppid = _props(props, pid, E, BEAM_STRIDE, check)
    _lib.call(fn, E, px, pc, pk, pe, ppid, pp, int(mtype), pr, pcc, pv,
              _flags(values_only, overwrite), _stream(stream))


def truss_kc0(x, conn, r, c, v, props, pid=None, init_k=None, values_only=False,
              overwrite=False, stream=None, check=True):
    """Truss KC0 of a batch (update_rotation_matrix(x) + update_probe_xe +
    update_length + update_KC0; truss.pyx:203-338, 386-801); props are beam
    records (A, E, G, J used)."""
    E, px, pc, pk, _ = _common('truss', x, conn, init_k, None, check)
    _check_span(72 * E, r, c, v, values_only, init_k, check)
    pr, pcc, pv = _outputs(r, c, v, True, values_only)
    pp, ppid = _props(props, pid, E, BEAM_STRIDE, check)
    _lib.call('pf3d_truss_kc0', E, px, pc, pk, ppid, pp, pr, pcc, pv,
              _flags(values_only, overwrite), _stream(stream))


def truss_m(x, conn, r, c, v, props, mtype=0, pid=None, init_k=None, values_only=False,
            overwrite=False, stream=None, check=True):
    """Truss mass matrix (update_M, truss.pyx:838-1800), mtype 0 consistent
    (144 slots), 1 lumped (36 slots)."""
    E, px, pc, pk, _ = _common('truss', x, conn, init_k, None, check)
    _check_span(slot_count('truss', 'M', mtype) * E, r, c, v, values_only, init_k, check)
    pr, pcc, pv = _outputs(r, c, v, True, values_only)
    pp, ppid = _props(props, pid, E, BEAM_STRIDE, check)
    _lib.call('pf3d_truss_m', E, px, pc, pk, ppid, pp, int(mtype), pr, pcc, pv,
              _flags(values_only, overwrite), _stream(stream))


def _spring_axes(axes, E):
    return _ptr(axes, torch.float64, 'axes', 6 * E)


def spring_kc0(x, conn, axes, r, c, v, k, pid=None, init_k=None, values_only=False,
               overwrite=False, stream=None, check=True):
    """Spring KC0 of a batch (update_rotation_matrix(xi, xj, xk, vxyi, vxyj,
    vxyk) + update_KC0; spring.pyx:147-206, 295-707).  axes (E, 6): element
    x axis and XY-plane vector; k: spring records (kxe, kye, kze, krxe, krye,
    krze), one per pid (or one shared record)."""
    E, px, pc, pk, _ = _common('spring', x, conn, init_k, None, check)
    pa = _spring_axes(axes, E)
    _check_span(72 * E, r, c, v, values_only, init_k, check)
    pr, pcc, pv = _outputs(r, c, v, True, values_only)
    pp, ppid = _props(k, pid, E, SPRING_STRIDE, check)
    _lib.call('pf3d_spring_kc0', E, px, pc, pk, pa, ppid, pp, pr, pcc, pv,
              _flags(values_only, overwrite), _stream(stream))


# ----------------------------------------------------------------------
# several element types in one launch
# ----------------------------------------------------------------------

class _CGroup(ctypes.Structure):
    """struct pf3d_elem_group of include/pf3d.h"""
    _fields_ = [('etype', ctypes.c_int), ('n_elem', ctypes.c_int64), ('conn', ctypes.c_void_p),
                ('k0', ctypes.c_int64), ('init_k', ctypes.c_void_p), ('evec', ctypes.c_void_p),
                ('pid', ctypes.c_void_p), ('props', ctypes.c_void_p), ('k6rot', ctypes.c_void_p),
                ('alpha', ctypes.c_void_p), ('hg', ctypes.c_void_p)]


class ElementGroup:
    """Elements of one type for :func:`mixed_assemble`: the arguments of the
    per-type entry point, plus ``k0``, the first slot of the group (the
    running ``init_k`` counter a caller keeps over its element lists)."""

    def __init__(self, etype, conn, props, pid=None, k0=0, init_k=None, evec=None, k6rot=None,
                 alpha=None, hg=None):
        if etype not in ('quad4', 'quad4r', 'tria3r', 'beamc'):
            raise ValueError('mixed groups take shells and beamc, got %r' % etype)
        self.etype, self.conn, self.props, self.pid = etype, conn, props, pid
        self.k0, self.init_k, self.evec = int(k0), init_k, evec
        self.k6rot, self.alpha, self.hg = k6rot, alpha, hg

    def slots(self, matrix='KC0', mtype=0):
        return slot_count(self.etype, matrix, mtype) * self.conn.shape[0]


def mixed_assemble(matrix, groups, x, r, c, v, u=None, mtype=0, values_only=False,
                   overwrite=False, stream=None, check=True):
    """KC0 / KG (from u) / M of several element types in ONE kernel launch
    (pf3d_mixed_assemble): each group writes the slots the per-type call
    would write, shifted by its k0."""
    mat = {'KC0': 0, 'KG': 1, 'M': 3}[matrix]
    cg = (_CGroup * max(1, len(groups)))()
    for i, g in enumerate(groups):
        beam = g.etype == 'beamc'
        E, px, pc, pk, pe = _common(g.etype, x, g.conn, g.init_k, g.evec, check)
        if beam and pe is None:
            raise ValueError('beamc group needs vxy (evec)')
        if check:
            if u is not None:
                _check_nodes(g.conn, x, u)
            size = slot_count(g.etype, matrix, mtype)
            outs = (('v', v),) + (() if values_only else (('r', r), ('c', c)))
            hi = g.k0 + (size * E if g.init_k is None else
                         (int(g.init_k.max()) + size if E else 0))
            lo = g.k0 + (0 if g.init_k is None or not E else int(g.init_k.min()))
            for name, t in outs:
                if lo < 0 or hi > t.numel():
                    raise ValueError('group %d writes slots [%d, %d) but %s has %d'
                                     % (i, lo, hi, name, t.numel()))
        pp, ppid = _props(g.props, g.pid, E, BEAM_STRIDE if beam else SHELL_STRIDE, check)
        cg[i] = _CGroup(ETYPE[g.etype], E, pc, g.k0, pk, pe, ppid, pp,
                        _ptr(g.k6rot, torch.float64, 'k6rot', E, optional=True),
                        _ptr(g.alpha, torch.float64, 'alpha', E, optional=True),
                        _ptr(g.hg, torch.float64, 'hg', 5 * E, optional=True))
    pr, pcc, pv = _outputs(r, c, v, True, values_only)
    pu = _ptr(u, torch.float64, 'u', optional=matrix != 'KG')
    _lib.call('pf3d_mixed_assemble', mat, int(mtype), len(groups), ctypes.cast(cg, ctypes.c_void_p),
              _ptr(x, torch.float64, 'x'), pu, pr, pcc, pv, _flags(values_only, overwrite),
              _stream(stream))


# ----------------------------------------------------------------------
# internal forces (update_fint)
# ----------------------------------------------------------------------

def shell_fe(etype, x, conn, u, props, pid=None, xmat=None, k6rot=None, alpha=None, hg=None,
             fe=None, stream=None, check=True):
    """Element internal forces in global axes, shape (E, nn, 6).

    The per-element part of ``update_fint`` (update_probe_ue +
    update_probe_finte + rotation; quad4.pyx:627-680, 1174-1202, 1316-1363;
    quad4r.pyx:549, 4620; tria3r.pyx:549, 2975)."""
    E, px, pc, _, pe = _common(etype, x, conn, None, xmat, check)
    if check:
        _check_nodes(conn, x, u)
    pp, ppid = _props(props, pid, E, SHELL_STRIDE, check)
    nn = NNODES[etype]
    if fe is None:
        fe = torch.empty((E, nn, 6), dtype=torch.float64, device=x.device)
    _lib.call('pf3d_shell_fe', ETYPE[etype], E, px, pc, pe, ppid, pp,
              _ptr(k6rot, torch.float64, 'k6rot', E, optional=True),
              _ptr(alpha, torch.float64, 'alpha', E, optional=True),
              _ptr(hg, torch.float64, 'hg', 5 * E, optional=True),
              _ptr(u, torch.float64, 'u'), _ptr(fe, torch.float64, 'fe', 6 * nn * E),
              _stream(stream))
    return fe


def beam_fe(x, conn, vxy, u, props, pid=None, fe=None, stream=None, check=True, etype='beamc'):
    """BeamC (or BeamLR) element internal forces in global axes, shape
    (E, 2, 6) (beamc.pyx:234-286, 352-488, 1354-1391; beamlr.pyx:236-288,
    354-406, 1188-1225)."""
    fn = _beam_entry(etype, 'fe')
    E, px, pc, _, pe = _common(etype, x, conn, None, vxy, check)
    if pe is None:
        raise ValueError('vxy is required')
    if check:
        _check_nodes(conn, x, u)
    pp, ppid = _props(props, pid, E, BEAM_STRIDE, check)
    if fe is None:
        fe = torch.empty((E, 2, 6), dtype=torch.float64, device=x.device)
    _lib.call(fn, E, px, pc, pe, ppid, pp, _ptr(u, torch.float64, 'u'),
              _ptr(fe, torch.float64, 'fe', 12 * E), _stream(stream))
    return fe


def rod_fe(etype, x, conn, u, props, axes=None, pid=None, fe=None, stream=None, check=True):
    """Truss or Spring element internal forces in global axes, shape
    (E, 2, 6) (truss.pyx:150-202, 339-385, 802-837; spring.pyx:208-294,
    708-739).  Truss: props are beam records; Spring: props are spring
    records and axes (E, 6) is required."""
    if etype not in ('truss', 'spring'):
        raise ValueError('rod_fe takes truss or spring, got %r' % etype)
    E, px, pc, _, _ = _common(etype, x, conn, None, None, check)
    pa = _spring_axes(axes, E) if etype == 'spring' else None
    if check:
        _check_nodes(conn, x, u)
    pp, ppid = _props(props, pid, E, BEAM_STRIDE if etype == 'truss' else SPRING_STRIDE, check)
    if fe is None:
        fe = torch.empty((E, 2, 6), dtype=torch.float64, device=x.device)
    _lib.call('pf3d_rod_fe', ETYPE[etype], E, px, pc, pa, ppid, pp,
              _ptr(u, torch.float64, 'u'), _ptr(fe, torch.float64, 'fe', 12 * E),
              _stream(stream))
    return fe


class NodeGather:
    """Deterministic element-to-node sum of element forces.

    Built once per mesh from ``conn`` (E, nn): a radix sort of the entries
    by node keeps, for every node, its (element, node) entries in ascending
    element order, so ``fint[6p + d] += sum fe[e, i, d]`` is summed in the
    order of the reference's element loop (quad4.pyx:1338-1361) and gives
    the same bits on every run."""

    def __init__(self, conn, n_nodes, stream=None):
        if conn.dim() != 2:
            raise ValueError('conn must have shape (n_elem, nn)')
        self.n_nodes = int(n_nodes)
        self.n_entries = conn.numel()
        self.nn = conn.shape[1]
        dev = conn.device
        self.offsets = torch.zeros(self.n_nodes + 1, dtype=torch.int64, device=dev)
        self.ids = torch.empty(max(1, self.n_entries), dtype=torch.int64, device=dev)
        if self.n_entries == 0:
            return
        nbytes = ctypes.c_int64()
        _lib.call('pf3d_incidence_bytes', self.n_entries, self.n_nodes, ctypes.byref(nbytes))
        work = torch.empty(nbytes.value, dtype=torch.uint8, device=dev)
        _lib.call('pf3d_incidence_plan', self.n_entries, _ptr(conn, torch.int64, 'conn'),
                  self.n_nodes, work.data_ptr(), nbytes.value, self.offsets.data_ptr(),
                  self.ids.data_ptr(), _stream(stream))
        # the workspace is released when this call returns
        (stream if stream is not None else torch.cuda.current_stream()).synchronize()

    def __call__(self, fe, fint, stream=None):
        """fint (6*n_nodes) += the node sums of fe (E, nn, 6)."""
        if fe.numel() != 6 * self.n_entries:
            raise ValueError('fe has %d entries, the plan %d' % (fe.numel(), 6 * self.n_entries))
        if self.n_entries == 0:
            return fint
        _lib.call('pf3d_gather_fint', self.n_nodes, self.offsets.data_ptr(), self.ids.data_ptr(),
                  _ptr(fe, torch.float64, 'fe'),
                  _ptr(fint, torch.float64, 'fint', 6 * self.n_nodes), _stream(stream))
        return fint


def update_fint(etype, x, conn, u, fint, props, gather=None, vxy=None, axes=None, stream=None,
                **kw):
    """fint += internal forces of a batch of elements, the loop
    ``for e in elems: e.update_probe_ue(u); e.update_fint(fint, prop)``.

    ``gather`` is a :class:`NodeGather` for ``conn`` (built here if None);
    vxy for beams, axes for springs."""
    if etype in ('beamc', 'beamlr'):
        fe = beam_fe(x, conn, vxy, u, props, stream=stream, etype=etype, **kw)
    elif etype in ('truss', 'spring'):
        fe = rod_fe(etype, x, conn, u, props, axes=axes, stream=stream, **kw)
    else:
        fe = shell_fe(etype, x, conn, u, props, stream=stream, **kw)
    if gather is None:
        gather = NodeGather(conn, fint.numel() // 6, stream=stream)
    return gather(fe, fint, stream=stream)

"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, the smoke check and the CPU-baseline leg of bench.py import
this module.  The library is a plain-C restatement of pyfe3d's element
arithmetic (see oracle.c for the file:line citations and for which
functions are pinned by golden vectors).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, 'liboracle.so')

ETYPES = {'quad4': 0, 'quad4r': 1, 'tria3r': 2, 'beamc': 3, 'truss': 4, 'spring': 5,
          'beamlr': 6}
MATS = {'KC0': 0, 'KG': 1, 'KGs': 2, 'M': 3, 'KAB': 5, 'KAG': 6, 'CA': 7}
NNODES = {'quad4': 4, 'quad4r': 4, 'tria3r': 3, 'beamc': 2, 'truss': 2, 'spring': 2,
          'beamlr': 2}
SHELL_STRIDE = 32
BEAM_STRIDE = 16
SPRING_STRIDE = 8

_lib = None


def build():
    subprocess.check_call(['make', '-s', '-C', HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        _lib.orc_assemble.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int64] + [P] * 14 + [ctypes.c_int]
        _lib.orc_assemble.restype = ctypes.c_int
        _lib.orc_geometry.argtypes = [ctypes.c_int, ctypes.c_int64, P, P, P, P, P]
        _lib.orc_slot_count.argtypes = [ctypes.c_int] * 3
        _lib.orc_set_threads.argtypes = [ctypes.c_int]
        _lib.orc_set_threads.restype = ctypes.c_int
        _lib.orc_set_threads(1)  # the reference's serial element loop by default
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def set_threads(n):
    """Host threads of the element loop of assemble(); returns the count."""
    return lib().orc_set_threads(int(n))


def slot_count(etype, mat, mtype=0):
    return lib().orc_slot_count(ETYPES[etype], MATS[mat], mtype)


def props_table(records, etype):
    """Pad (n, k) property records to the oracle's stride."""
    stride = {'beamc': BEAM_STRIDE, 'truss': BEAM_STRIDE, 'beamlr': BEAM_STRIDE,
              'spring': SPRING_STRIDE}.get(etype, SHELL_STRIDE)
    rec = np.atleast_2d(np.asarray(records, dtype=np.float64))
    out = np.zeros((rec.shape[0], stride))
    k = min(stride, rec.shape[1])
    out[:, :k] = rec[:, :k]
    return np.ascontiguousarray(out)


def geometry(etype, x, conn, evec):
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    E = conn.shape[0]
    g = np.zeros((E, 14))
    xe = np.zeros((E, 3 * NNODES[etype]))
    lib().orc_geometry(ETYPES[etype], E, _p(np.ascontiguousarray(x, np.float64)),
                       _p(conn), _p(np.ascontiguousarray(evec, np.float64)), _p(g), _p(xe))
    return g, xe


def assemble(etype, mat, x, conn, init_k, evec, pid, props, r, c, v, mtype=0,
             k6rot=None, alpha=None, hg=None, u=None, stress=None, values_only=False):
    """Accumulate one matrix of a batch of elements into (r, c, v)."""
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    E = conn.shape[0]
    f = lambda a, dt=np.float64: None if a is None else np.ascontiguousarray(a, dtype=dt)
    assert r.dtype == np.int64 and c.dtype == np.int64 and v.dtype == np.float64
    st = lib().orc_assemble(
        ETYPES[etype], MATS[mat], int(mtype), E, _p(f(x)), _p(conn),
        _p(f(init_k, np.int64)), _p(f(evec)), _p(f(pid, np.int64)),
        _p(props_table(props, etype)), _p(f(k6rot)), _p(f(alpha)), _p(f(hg)),
        _p(f(u)), _p(f(stress)), _p(r), _p(c), _p(v), int(bool(values_only)))
    if st:
        raise NotImplementedError('oracle has no restatement for %s %s' % (etype, mat))


def fint(etype, x, conn, evec, pid, props, u, k6rot=None, alpha=None, hg=None, out=None):
    """Accumulated internal forces of a batch (update_fint per element)."""
    L = lib()
    L.orc_fint.argtypes = [ctypes.c_int, ctypes.c_int64] + [ctypes.c_void_p] * 10
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    f = lambda a, dt=np.float64: None if a is None else np.ascontiguousarray(a, dtype=dt)
    if out is None:
        out = np.zeros(np.asarray(u).size)
    st = L.orc_fint(ETYPES[etype], conn.shape[0], _p(f(x)), _p(conn), _p(f(evec)),
                    _p(f(pid, np.int64)), _p(props_table(props, etype)), _p(f(k6rot)),
                    _p(f(alpha)), _p(f(hg)), _p(f(u)), _p(out))
    if st:
        raise NotImplementedError(etype)
    return out

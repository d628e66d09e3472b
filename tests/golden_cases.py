"""Iterate the golden fixtures written by tests/golden/make_golden.py."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, 'golden')
SHELLS = ['quad4', 'quad4r', 'tria3r']
# (fixture key, matrix, mtype)
SHELL_MATS = [('KC0', 'KC0', 0), ('KG', 'KG', 0), ('KGs', 'KGs', 0),
              ('M0', 'M', 0), ('M1', 'M', 1), ('M2', 'M', 2)]
# Per-element tolerance |dv| <= TOL * max|v_ref| over the element's slots.
TOL = 1e-12
# The reference's Quad4R computes detJ and the inverse Jacobian with expanded
# products of absolute local coordinates (quad4r.pyx:3094-3100, 4861-4866);
# on the 2 mm elements 0.2 m away from the origin of this case that loses
# about 4 digits, so the reference's own rounding error reaches ~2e-12.
TOL_CASE = {('quad4r', 'threshold_hfield'): 4e-12}


def load(etype):
    return np.load(os.path.join(GOLDEN, 'golden_%s.npz' % etype))


def cases(etype):
    z = load(etype)
    names = sorted(set(k.split('__')[0] for k in z.files))
    for name in names:
        yield name, {k.split('__', 1)[1]: z[k] for k in z.files if k.startswith(name + '__')}


def element_errors(v, vref, size, perm):
    """max over elements of max|dv| / max|vref| on that element's slots."""
    worst = 0.
    for e in range(len(perm)):
        s = slice(size * perm[e], size * perm[e] + size)
        scale = np.abs(vref[s]).max()
        d = np.abs(v[s] - vref[s]).max()
        worst = max(worst, d / scale if scale > 0 else d)
    return worst

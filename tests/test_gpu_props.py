"""Laminate property records built on the GPU (pf3d_laminate_records)
against the reference's golden ShellProp records (tests/golden) and against
the host builder on random per-element ply stacks with two materials,
offsets and padding plies."""
import json
import os

import numpy as np
import pytest
import torch

from pyfe3d_amd.shellprop import RECORD_FIELDS, laminated_plate, laminate_records_device, material_row

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _scf_cancellation(plyts, off):
    """Largest ratio sum|terms| / |sum| of the shear-correction polynomial
    over the plies (shellprop.pyx:485-548): one-ulp differences in its
    inputs move the factors by about eps times this ratio."""
    h = sum(plyts)
    zb = -h / 2 + off
    z1, amp = zb, 1.
    for t in plyts:
        z2 = z1 + t
        o = off
        terms = [15 * o * z1 ** 4, 30 * o * z1 ** 2 * zb * (2 * o - zb), -15 * o * z2 ** 4,
                 30 * o * z2 ** 2 * zb * (-2 * o + zb), -3 * z1 ** 5,
                 10 * z1 ** 3 * (-2 * o ** 2 - 2 * o * zb + zb ** 2),
                 -15 * z1 * zb ** 2 * (4 * o ** 2 - 4 * o * zb + zb ** 2), 3 * z2 ** 5,
                 10 * z2 ** 3 * (2 * o ** 2 + 2 * o * zb - zb ** 2),
                 15 * z2 * zb ** 2 * (4 * o ** 2 - 4 * o * zb + zb ** 2)]
        amp = max(amp, sum(abs(x) for x in terms) / abs(sum(terms)))
        z1 = z2
    return amp


def _close(got, ref, what, scf_amp=1.):
    """Fields within 1e-12 of themselves, and of their group's scale where
    a field can be rounding noise (B of a symmetric stack, E45 of a balanced
    one, intrhoz without offset: the device's cos/sin differ from the host
    libm's in the last bit): A, B, D entries within 1e-12 of the largest
    A/B/D entry, E entries of the largest E entry, intrhoz of intrho h."""
    ix = RECORD_FIELDS.index
    abd = np.abs(ref[:18]).max()
    es = max(abs(ref[ix('E44')]), abs(ref[ix('E55')]))
    for i, f in enumerate(RECORD_FIELDS):
        tol = 1e-12 * max(abs(ref[i]), 1e-30)
        if f[0] in 'ABD' and len(f) == 3:
            tol += 1e-12 * abd
        elif f in ('E44', 'E45', 'E55'):
            tol += 1e-12 * es
        elif f == 'intrhoz':
            tol += 1e-12 * abs(ref[ix('intrho')]) * ref[ix('h')]
        elif f in ('scf_k13', 'scf_k23'):
            tol = max(tol, 8 * 2. ** -53 * scf_amp * abs(ref[i]))
        assert abs(got[i] - ref[i]) <= tol, (what, f, got[i], ref[i], scf_amp)


def test_laminate_records_match_reference_golden():
    """Every record field of the reference-built ShellProps (the equivalent
    moduli e1, e2, g12, nu12, nu21 are not part of the record)."""
    z = np.load(os.path.join(HERE, 'golden', 'golden_shellprop.npz'))
    fields = [str(f) for f in z['fields']]
    specs = json.load(open(os.path.join(HERE, 'golden', 'golden_shellprop_specs.json')))
    assert len(specs) == len(z['records']) > 0
    for spec, ref in zip(specs, z['records']):
        stack = spec['stack']
        plyts = spec.get('plyts') or [spec['plyt']] * len(stack)
        mat = material_row(tuple(spec['laminaprop']), spec.get('rho', 0.))
        out = laminate_records_device([stack], [plyts], [mat],
                                      offset=[spec.get('offset', 0.)]).cpu().numpy()[0]
        full = np.array([ref[fields.index(f)] for f in RECORD_FIELDS])
        print(stack, 'max rel diff',
              np.max(np.abs(out[:len(RECORD_FIELDS)] - full) / np.maximum(np.abs(full), 1e-30)))
        _close(out, full, stack)


def test_laminate_records_random_stacks_vs_host():
    """400 elements, 1..8 plies (zero-thickness padding to 8), two
    materials, per-ply thicknesses and offsets up to a few thicknesses
    (where the shear-correction polynomial cancels by up to ~7 digits: its
    bar is 8 eps times the measured cancellation ratio)."""
    rng = np.random.default_rng(11)
    n, P = 400, 8
    lps = [(140e9, 10e9, 0.3, 5e9, 4e9, 3e9), (70e9, 0.33)]
    rhos = [1600., 2700.]
    mats = [material_row(lp, r) for lp, r in zip(lps, rhos)]
    theta = rng.choice([0., 15., 30., 45., -45., 60., 90., -72.5], (n, P))
    plyt = rng.uniform(1e-4, 5e-4, (n, P))
    pmat = rng.integers(0, 2, (n, P))
    nply = rng.integers(1, P + 1, n)
    for e in range(n):
        plyt[e, nply[e]:] = 0.
    off = rng.uniform(-1e-3, 1e-3, n)
    out = laminate_records_device(theta, plyt, mats, ply_mat=pmat, offset=off).cpu().numpy()
    for e in range(n):
        k = nply[e]
        ref = laminated_plate(stack=list(theta[e, :k]), plyts=list(plyt[e, :k]),
                              laminaprops=[lps[i] for i in pmat[e, :k]],
                              rhos=[rhos[i] for i in pmat[e, :k]], offset=off[e],
                              calc_scf=True).record()
        _close(out[e], ref, e, _scf_cancellation(list(plyt[e, :k]), off[e]))
    out2 = laminate_records_device(theta, plyt, mats, ply_mat=pmat, offset=off, calc_scf=False)
    assert torch.all(out2[:, RECORD_FIELDS.index('scf_k13')] == 5 / 6.)


def test_lp_records_match_host_builder():
    """pf3d_lp_records against shellprop_from_LaminationParameters on 300
    random designs (thickness, 14 lamination parameters, two materials, one
    of them tri-axial): A, B, D within 1e-13 of their group's scale, E of
    its own; and a stack's own lamination parameters give its A, B, D, E."""
    from pyfe3d_amd.shellprop import (LaminationParameters, lp_material_row, lp_records_device,
                                      read_laminaprop, shellprop_from_LaminationParameters)
    rng = np.random.default_rng(5)
    mats = [read_laminaprop((140e9, 10e9, 0.3, 5e9, 4e9, 3e9), 1600.),
            read_laminaprop((71e9, 7e9, 0.28, 7e9, 7e9, 7e9, 7e9, 0.28, 0.28), 1500.)]
    rows = [lp_material_row(m, m.rho) for m in mats]
    n = 300
    h = rng.uniform(5e-4, 5e-3, n)
    xi = rng.uniform(-1, 1, (n, 14))
    mid = rng.integers(0, 2, n)
    out = lp_records_device(h, xi, rows, mat_id=mid).cpu().numpy()
    ix = RECORD_FIELDS.index
    for e in range(n):
        lp = LaminationParameters()
        for k, val in zip(LaminationParameters.NAMES, xi[e]):
            setattr(lp, k, val)
        ref = shellprop_from_LaminationParameters(h[e], mats[mid[e]], lp).record()
        ref[ix('intrho')] = mats[mid[e]].rho * h[e]
        ref[ix('intrhoz2')] = mats[mid[e]].rho * h[e] ** 3 / 12.
        for X, p in (('A', 0), ('B', 1), ('D', 2)):
            sl = slice(6 * ('ABD'.index(X)), 6 * ('ABD'.index(X)) + 6)
            scale = np.abs(ref[0:6]).max() * h[e] ** p
            assert np.abs(out[e, sl] - ref[sl]).max() <= 1e-13 * scale, (e, X)
        es = slice(18, 21)
        assert np.abs(out[e, es] - ref[es]).max() <= 1e-13 * np.abs(ref[es]).max(), e
        assert np.allclose(out[e, 21:28], ref[21:28], rtol=1e-14, atol=0), e
        assert not out[e, 28:].any()
    # lamination parameters of real stacks through the device builder
    stacks = [[0, 45, -45, 90], [30, -60, 15, 90, 0, 45], [0]]
    for st in stacks:
        p = laminated_plate(st, 0.000125, (140e9, 10e9, 0.3, 5e9, 4e9, 3e9))
        lp = p.calc_lamination_parameters()
        r = lp_records_device([p.h], lp.vector()[None], [rows[0]]).cpu().numpy()[0]
        full = p.record()
        scale = np.abs(full[:6]).max()
        for X, pw in (('A', 0), ('B', 1), ('D', 2)):
            sl = slice(6 * 'ABD'.index(X), 6 * 'ABD'.index(X) + 6)
            assert np.abs(r[sl] - full[sl]).max() <= 1e-12 * scale * p.h ** pw, (st, X)

"""Host-side laminate tables against the reference's builders (golden
records written by tests/golden/make_golden.py) and closed forms."""
import json
import os

import numpy as np

from pyfe3d_amd.shellprop import RECORD_FIELDS, isotropic_plate, isotropic_records, laminated_plate

HERE = os.path.dirname(os.path.abspath(__file__))


def test_laminated_plate_matches_reference():
    z = np.load(os.path.join(HERE, 'golden', 'golden_shellprop.npz'))
    fields = [str(f) for f in z['fields']]
    specs = json.load(open(os.path.join(HERE, 'golden', 'golden_shellprop_specs.json')))
    for spec, ref in zip(specs, z['records']):
        spec['laminaprop'] = tuple(spec['laminaprop'])
        p = laminated_plate(calc_scf=True, **spec)
        got = np.array([getattr(p, f) for f in fields])
        scale = np.maximum(np.abs(ref), 1e-300)
        for f, a, b in zip(fields, got, ref):
            assert abs(a - b) <= 1e-12 * max(abs(b), 1e-30) + 1e-12 * np.abs(ref[:18]).max() * (
                f[0] in 'ABD' and len(f) == 3), (spec['stack'], f, a, b)


def test_isotropic_closed_form():
    h, E, nu, rho = 0.004, 70e9, 0.33, 2700.
    p = isotropic_plate(h, E, nu, rho=rho)
    Q = E / (1 - nu ** 2)
    assert np.isclose(p.A11, Q * h, rtol=1e-14)
    assert np.isclose(p.A12, nu * Q * h, rtol=1e-14)
    assert np.isclose(p.A66, E / 2 / (1 + nu) * h, rtol=1e-14)
    assert np.isclose(p.D11, Q * h ** 3 / 12, rtol=1e-13)
    assert abs(p.B11) < 1e-12 * p.A11 * h
    assert np.isclose(p.intrho, rho * h, rtol=1e-14)
    assert np.isclose(p.intrhoz2, rho * h ** 3 / 12, rtol=1e-13)
    assert np.isclose(p.scf_k13, 5 / 6., rtol=1e-12)


def test_isotropic_records_scale_with_thickness():
    hs = np.array([0.001, 0.002, 0.0035])
    rec = isotropic_records(hs, 70e9, 0.33, rho=2700.)
    for h, r in zip(hs, rec):
        ref = isotropic_plate(h, 70e9, 0.33, rho=2700.).record()
        assert np.allclose(r, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max()), (h, r - ref)

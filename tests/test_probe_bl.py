"""Quad4Probe.update_BL (quad4.pyx:272-395) on distorted quads: the
bilinear element reproduces linear displacement fields exactly, so each BL
row applied to the nodal values of a linear field gives that field's strain
at every natural point (a patch test; runs on the host)."""
import numpy as np

from pyfe3d_amd.compat import Quad4Probe


def test_update_bl_patch_test():
    rng = np.random.default_rng(2)
    for trial in range(20):
        corners = np.array([[0., 0.], [1., 0.], [1., 1.], [0., 1.]]) * rng.uniform(0.01, 3)
        corners += rng.uniform(-0.2, 0.2, (4, 2)) * corners.max()
        p = Quad4Probe()
        p.xe[:] = np.column_stack([corners, np.zeros(4)]).ravel()
        # linear fields of every DOF: f_d(x, y) = a_d + b_d x + c_d y
        a, b, cc = rng.normal(size=(3, 6))
        ue = np.concatenate([a + b * x + cc * y for x, y in corners])
        bu, cu = b[0], cc[0]
        bv, cv = b[1], cc[1]
        bw, cw = b[2], cc[2]
        for xi, eta in rng.uniform(-1, 1, (5, 2)):
            p.update_BL(xi, eta)
            # the field values at the natural point (bilinear interpolation)
            N = np.array([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta),
                          (1 - xi) * (1 + eta)]) / 4.
            x, y = N @ corners
            rx, ry, rz = (a[3:] + b[3:] * x + cc[3:] * y)
            expect = dict(BLexx=bu, BLeyy=cv, BLgxy=cu + bv, BLkxx=b[4], BLkyy=-cc[3],
                          BLkxy=cc[4] - b[3], BLgyz_grad=cw, BLgyz_rot=-rx, BLgxz_grad=bw,
                          BLgxz_rot=ry, BLdrilling=(cu - bv) / 2. + rz)
            for name, val in expect.items():
                got = getattr(p, name) @ ue
                assert abs(got - val) <= 1e-12 * (1 + abs(val)), (trial, name, got, val)

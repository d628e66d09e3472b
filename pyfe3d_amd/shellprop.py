"""Host-side laminate property tables (the input provider of the element kernels).

Mirrors pyfe3d's ``MatLamina``, ``Lamina``, ``ShellProp`` and the builders of
``shellprop_utils`` (read_laminaprop, laminated_plate, isotropic_plate) with
the same attribute names, and adds ``record()`` to pack a property into the
32-double record the GPU kernels read.  Runs once per property on the host
(microseconds), so it stays in numpy.

References: shellprop.pyx:128-196 (MatLamina.rebuild), :278-337
(Lamina.rebuild), :485-548 (calc_scf), :551-565 (calc_equivalent_properties),
:568-621 (calc_constitutive_matrix); shellprop_utils.py:13-206.
"""
import math

import numpy as np

RECORD_FIELDS = ['A11', 'A12', 'A16', 'A22', 'A26', 'A66',
                 'B11', 'B12', 'B16', 'B22', 'B26', 'B66',
                 'D11', 'D12', 'D16', 'D22', 'D26', 'D66',
                 'E44', 'E45', 'E55', 'scf_k13', 'scf_k23', 'h', 'offset',
                 'intrho', 'intrhoz', 'intrhoz2']
RECORD_STRIDE = 32


def deg2rad(thetadeg):
    return thetadeg * 4 * math.atan(1.) / 180.


class MatLamina:
    """Orthotropic lamina material (shellprop.pyx:128-196)."""

    def __init__(self):
        for k in ('e1 e2 e3 g12 g13 g23 nu12 nu21 nu13 nu31 nu23 nu32 rho a1 a2 a3 tref '
                  'q11 q12 q13 q21 q22 q23 q31 q32 q33 q44 q55 q66 '
                  'c11 c12 c13 c22 c23 c33 c44 c55 c66 u1 u2 u3 u4 u5 u6 u7').split():
            setattr(self, k, 0.)

    def rebuild(self):
        e1, e2, e3 = self.e1, self.e2, self.e3
        nu12, nu21, nu13, nu31, nu23, nu32 = (self.nu12, self.nu21, self.nu13, self.nu31,
                                              self.nu23, self.nu32)
        delta = (1 - nu12 * nu21 - nu23 * nu32 - nu31 * nu13 - 2 * nu21 * nu32 * nu13) / (e1 * e2)
        self.c11 = (1 - nu23 * nu23) / (delta * e2)
        self.c12 = (nu21 + nu31 * nu23) / (delta * e2)
        self.c13 = (nu31 + nu21 * nu32) / (delta * e2)
        self.c22 = (1 - nu13 * nu31) / (delta * e1)
        self.c23 = (nu32 + nu12 * nu31) / (delta * e1)
        self.c33 = e3 * (1 - nu12 * nu21) / (delta * e1 * e2)
        self.c44, self.c55, self.c66 = self.g23, self.g13, self.g12
        den = (1 - nu12 * nu21 - nu13 * nu31 - nu23 * nu32 - nu12 * nu23 * nu31
               - nu13 * nu21 * nu32)
        self.q11 = e1 * (1 - nu23 * nu32) / den
        self.q12 = e1 * (nu21 + nu23 * nu31) / den
        self.q13 = e1 * (nu31 + nu21 * nu32) / den
        self.q21 = e2 * (nu12 + nu13 * nu32) / den
        self.q22 = e2 * (1 - nu13 * nu31) / den
        self.q23 = e2 * (nu32 + nu12 * nu31) / den
        self.q31 = e3 * (nu13 + nu12 * nu32) / den
        self.q32 = e3 * (nu23 + nu13 * nu21) / den
        self.q33 = e3 * (1 - nu12 * nu21) / den
        self.q66, self.q44, self.q55 = self.g12, self.g23, self.g13
        self.u1 = (3 * self.q11 + 3 * self.q22 + 2 * self.q12 + 4 * self.q66) / 8.
        self.u2 = (self.q11 - self.q22) / 2.
        self.u3 = (self.q11 + self.q22 - 2 * self.q12 - 4 * self.q66) / 8.
        self.u4 = (self.q11 + self.q22 + 6 * self.q12 - 4 * self.q66) / 8.
        self.u5 = (self.u1 - self.u4) / 2.
        self.u6 = (self.q44 + self.q55) / 2.
        self.u7 = (self.q44 - self.q55) / 2.


class Lamina:
    """One ply with its stiffness in laminate axes (shellprop.pyx:278-337)."""

    def __init__(self):
        self.plyid = 0
        self.h = 0.
        self.thetadeg = 0.
        self.matlamina = None

    def rebuild(self):
        t = deg2rad(self.thetadeg)
        c, s = math.cos(t), math.sin(t)
        self.cost, self.sint = c, s
        self.cos2t, self.cos4t = math.cos(2 * t), math.cos(4 * t)
        self.sin2t, self.sin4t = math.sin(2 * t), math.sin(4 * t)
        c2, c3, c4 = c ** 2, c ** 3, c ** 4
        s2, s3, s4 = s ** 2, s ** 3, s ** 4
        m = self.matlamina
        d = 1 - m.nu12 * m.nu21
        q11, q12, q22 = m.e1 / d, m.nu12 * m.e2 / d, m.e2 / d
        q44, q55, q66 = m.g23, m.g13, m.g12
        self.q11L = q11 * c4 + 2 * (q12 + 2 * q66) * s2 * c2 + q22 * s4
        self.q12L = (q11 + q22 - 4 * q66) * s2 * c2 + q12 * (s4 + c4)
        self.q22L = q11 * s4 + 2 * (q12 + 2 * q66) * s2 * c2 + q22 * c4
        self.q16L = (q11 - q12 - 2 * q66) * s * c3 + (q12 - q22 + 2 * q66) * s3 * c
        self.q26L = (q11 - q12 - 2 * q66) * s3 * c + (q12 - q22 + 2 * q66) * s * c3
        self.q66L = (q11 + q22 - 2 * q12 - 2 * q66) * s2 * c2 + q66 * (s4 + c4)
        self.q44L = q44 * c2 + q55 * s2
        self.q45L = (q55 - q44) * s * c
        self.q55L = q55 * c2 + q44 * s2


class ShellProp:
    """Laminate ABD/E matrices, shear factors and mass integrals."""

    def __init__(self):
        for k in RECORD_FIELDS + ['e1', 'e2', 'g12', 'nu12', 'nu21']:
            setattr(self, k, 0.)
        self.scf_k13 = self.scf_k23 = 5 / 6.
        self.plies = []
        self.stack = []

    @property
    def A(self):
        return np.array([[self.A11, self.A12, self.A16], [self.A12, self.A22, self.A26],
                         [self.A16, self.A26, self.A66]])

    @property
    def B(self):
        return np.array([[self.B11, self.B12, self.B16], [self.B12, self.B22, self.B26],
                         [self.B16, self.B26, self.B66]])

    @property
    def D(self):
        return np.array([[self.D11, self.D12, self.D16], [self.D12, self.D22, self.D26],
                         [self.D16, self.D26, self.D66]])

    @property
    def E(self):
        return np.array([[self.E44, self.E45], [self.E45, self.E55]])

    @property
    def ABD(self):
        return np.block([[self.A, self.B], [self.B, self.D]])

    def calc_constitutive_matrix(self):
        """ABD/E by through-thickness integration (shellprop.pyx:568-621)."""
        self.h = sum(p.h for p in self.plies)
        for k in RECORD_FIELDS[:21] + ['intrho', 'intrhoz', 'intrhoz2']:
            setattr(self, k, 0.)
        h0 = -self.h / 2. + self.offset
        for ply in self.plies:
            hk_1 = h0
            h0 += ply.h
            hk = h0
            rho = ply.matlamina.rho
            self.intrho += rho * (hk - hk_1)
            self.intrhoz += rho * (hk * hk / 2. - hk_1 * hk_1 / 2.)
            self.intrhoz2 += rho * (hk * hk * hk / 3. - hk_1 * hk_1 * hk_1 / 3.)
            d1 = hk - hk_1
            d2 = 1 / 2. * (hk * hk - hk_1 * hk_1)
            d3 = 1 / 3. * (hk * hk * hk - hk_1 * hk_1 * hk_1)
            for ij in ('11', '12', '16', '22', '26', '66'):
                q = getattr(ply, 'q%sL' % ij)
                setattr(self, 'A' + ij, getattr(self, 'A' + ij) + q * d1)
                setattr(self, 'B' + ij, getattr(self, 'B' + ij) + 1 / 2. * q * (hk * hk - hk_1 * hk_1))
                setattr(self, 'D' + ij, getattr(self, 'D' + ij) + 1 / 3. * q * (hk * hk * hk - hk_1 * hk_1 * hk_1))
            self.E44 += ply.q44L * d1
            self.E45 += ply.q45L * d1
            self.E55 += ply.q55L * d1

    def calc_equivalent_properties(self):
        """Equivalent membrane moduli (shellprop.pyx:551-565)."""
        AI = np.linalg.inv(self.ABD)
        a11, a12, a22, a33 = AI[0, 0], AI[0, 1], AI[1, 1], AI[2, 2]
        self.e1 = 1. / (self.h * a11)
        self.e2 = 1. / (self.h * a22)
        self.g12 = 1. / (self.h * a33)
        self.nu12 = -a12 / a11
        self.nu21 = -a12 / a22

    def calc_scf(self):
        """Shear correction factors, Vlachoutsis 1992 (shellprop.pyx:485-548)."""
        D1 = R1 = den1 = D2 = R2 = den2 = 0.
        off = self.offset
        zb = -self.h / 2. + off
        z1 = zb
        for ply in self.plies:
            z2 = z1 + ply.h
            t = deg2rad(ply.thetadeg)
            m = ply.matlamina
            e1 = m.e1 * np.cos(t) + m.e2 * np.sin(t)
            e2 = m.e2 * np.cos(t) + m.e1 * np.sin(t)
            nu12 = m.nu12 * np.cos(t) + m.nu21 * np.sin(t)
            nu21 = m.nu21 * np.cos(t) + m.nu12 * np.sin(t)
            poly = (15 * off * z1 ** 4 + 30 * off * z1 ** 2 * zb * (2 * off - zb)
                    - 15 * off * z2 ** 4 + 30 * off * z2 ** 2 * zb * (-2 * off + zb)
                    - 3 * z1 ** 5 + 10 * z1 ** 3 * (-2 * off ** 2 - 2 * off * zb + zb ** 2)
                    - 15 * z1 * zb ** 2 * (4 * off ** 2 - 4 * off * zb + zb ** 2)
                    + 3 * z2 ** 5 + 10 * z2 ** 3 * (2 * off ** 2 + 2 * off * zb - zb ** 2)
                    + 15 * z2 * zb ** 2 * (4 * off ** 2 - 4 * off * zb + zb ** 2))
            D1 += e1 / (1 - nu12 * nu21)
            R1 += D1 * ((z2 - off) ** 3 / 3. - (z1 - off) ** 3 / 3.)
            den1 += m.g13 * ply.h * (self.h / ply.h) * D1 ** 2 * poly / (60 * m.g13)
            D2 += e2 / (1 - nu12 * nu21)
            R2 += D2 * ((z2 - off) ** 3 / 3. - (z1 - off) ** 3 / 3.)
            den2 += m.g23 * ply.h * (self.h / ply.h) * D2 ** 2 * poly / (60 * m.g23)
            z1 = z2
        self.scf_k13 = R1 ** 2 / den1
        self.scf_k23 = R2 ** 2 / den2

    def record(self):
        """The 32-double record read by the GPU kernels."""
        out = np.zeros(RECORD_STRIDE)
        out[:len(RECORD_FIELDS)] = [getattr(self, f) for f in RECORD_FIELDS]
        return out


def read_laminaprop(laminaprop, rho=0):
    """MatLamina from a 2-, 6- or 9-tuple (shellprop_utils.py:13-93)."""
    assert len(laminaprop) in (2, 6, 9), 'Invalid entry for laminaprop: ' + str(laminaprop)
    if len(laminaprop) == 2:
        e, nu = laminaprop
        g = e / (2 * (1 + nu))
        laminaprop = (e, e, nu, g, g, g, 0, 0, 0)
    elif len(laminaprop) == 6:
        laminaprop = tuple(laminaprop) + (0, 0, 0)
    m = MatLamina()
    m.e1, m.e2, m.e3 = laminaprop[0], laminaprop[1], laminaprop[6]
    m.nu12, m.nu13, m.nu23 = laminaprop[2], laminaprop[7], laminaprop[8]
    m.nu21 = m.nu12 * m.e2 / m.e1
    m.nu31 = m.nu13 * m.e3 / m.e1
    m.nu32 = m.nu23 * m.e3 / m.e2
    m.g12, m.g13, m.g23 = laminaprop[3], laminaprop[4], laminaprop[5]
    m.rho = rho
    m.rebuild()
    return m


def laminated_plate(stack, plyt=None, laminaprop=None, rho=0., plyts=None, laminaprops=None,
                    rhos=None, offset=0., calc_scf=True):
    """ShellProp of a ply stack (shellprop_utils.py:96-180)."""
    prop = ShellProp()
    prop.offset = offset
    prop.stack = list(stack)
    if plyts is None:
        if plyt is None:
            raise ValueError('plyt or plyts must be supplied')
        plyts = [plyt for _ in stack]
    if laminaprops is None:
        if laminaprop is None:
            raise ValueError('laminaprop or laminaprops must be supplied')
        laminaprops = [laminaprop for _ in stack]
    if rhos is None:
        rhos = [rho for _ in stack]
    if not (len(stack) == len(plyts) == len(laminaprops) == len(rhos)):
        raise ValueError('stack, plyts, laminaprops and rhos must have the same length')
    prop.h = 0.
    for t, lp, theta, r in zip(plyts, laminaprops, stack, rhos):
        ply = Lamina()
        ply.thetadeg = float(theta)
        ply.h = t
        prop.h += t
        ply.matlamina = read_laminaprop(lp, r)
        ply.rebuild()
        prop.plies.append(ply)
    prop.calc_constitutive_matrix()
    prop.calc_equivalent_properties()
    if calc_scf:
        prop.calc_scf()
    return prop


def isotropic_plate(thickness, E, nu, offset=0., calc_scf=True, rho=0.):
    """ShellProp of a single isotropic layer (shellprop_utils.py:182-206)."""
    return laminated_plate(plyt=thickness, stack=[0], laminaprop=(E, nu), rho=rho,
                           offset=offset, calc_scf=calc_scf)


def isotropic_records(h, E, nu, rho=0., calc_scf=True):
    """Records of isotropic plates for an array of thicknesses (one per
    element in a design sweep): A and E scale with h, D with h^3, B = 0; the
    shear factors and moduli are those of ``isotropic_plate``."""
    h = np.asarray(h, dtype=np.float64)
    base = isotropic_plate(1.0, E, nu, calc_scf=calc_scf, rho=rho)
    rec = np.zeros((h.size, RECORD_STRIDE))
    one = base.record()
    for i, f in enumerate(RECORD_FIELDS):
        if f[0] in 'AE' and f != 'E' and len(f) == 3:
            rec[:, i] = one[i] * h
        elif f[0] == 'D':
            rec[:, i] = one[i] * h ** 3
    rec[:, RECORD_FIELDS.index('scf_k13')] = base.scf_k13
    rec[:, RECORD_FIELDS.index('scf_k23')] = base.scf_k23
    rec[:, RECORD_FIELDS.index('h')] = h
    rec[:, RECORD_FIELDS.index('intrho')] = rho * h
    rec[:, RECORD_FIELDS.index('intrhoz2')] = rho * h ** 3 / 12.
    return rec


def material_row(laminaprop, rho=0.):
    """Material row (e1 e2 nu12 g12 g13 g23 rho 0) of the device laminate
    builder from a 2- or 6-tuple laminaprop (shellprop_utils.py:13-93)."""
    m = read_laminaprop(laminaprop, rho)
    return [m.e1, m.e2, m.nu12, m.g12, m.g13, m.g23, m.rho, 0.]


def laminate_records_device(theta_deg, plyt, mats, ply_mat=None, offset=None, calc_scf=True,
                            out=None, stream=None):
    """Records (n, 32) on the GPU of one laminate per element
    (pf3d_laminate_records): theta_deg, plyt (n, n_plies) ply angles and
    thicknesses (zero-thickness plies are padding), mats (n_mat, 8) rows of
    material_row, ply_mat (n, n_plies) material indices or None, offset (n)
    or None.  Arrays are torch tensors on the device or host arrays."""
    import torch
    from . import _lib
    dev = 'cuda'
    th = torch.as_tensor(np.asarray(theta_deg, dtype=np.float64) if not torch.is_tensor(theta_deg)
                         else theta_deg, dtype=torch.float64, device=dev).contiguous()
    tk = torch.as_tensor(np.asarray(plyt, dtype=np.float64) if not torch.is_tensor(plyt)
                         else plyt, dtype=torch.float64, device=dev).contiguous()
    if th.dim() != 2 or th.shape != tk.shape:
        raise ValueError('theta_deg and plyt must both be (n, n_plies)')
    n, n_plies = th.shape
    mt = torch.as_tensor(np.asarray(mats, dtype=np.float64), device=dev).reshape(-1, 8)
    pm = None
    if ply_mat is not None:
        pm = torch.as_tensor(np.asarray(ply_mat), dtype=torch.int32, device=dev).contiguous()
        if pm.shape != th.shape:
            raise ValueError('ply_mat must be (n, n_plies)')
        if int(pm.min()) < 0 or int(pm.max()) >= mt.shape[0]:
            raise ValueError('ply_mat outside [0, %d)' % mt.shape[0])
    off = None
    if offset is not None:
        off = torch.as_tensor(np.asarray(offset, dtype=np.float64), device=dev).reshape(n)
    if out is None:
        out = torch.empty((n, RECORD_STRIDE), dtype=torch.float64, device=dev)
    s = (stream if stream is not None else torch.cuda.current_stream()).cuda_stream
    _lib.call('pf3d_laminate_records', n, n_plies, th.data_ptr(), tk.data_ptr(),
              None if pm is None else pm.data_ptr(), mt.shape[0], mt.data_ptr(),
              None if off is None else off.data_ptr(), int(bool(calc_scf)), out.data_ptr(), s)
    return out

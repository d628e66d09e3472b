"""Shell property records: laminate stiffness, shear factors and mass
moments, on the host (one property at a time) and on the device (one per
element).

Written from classical laminate theory (Jones, *Mechanics of Composite
Materials*, ch. 2 and 4; Reddy, *Mechanics of Laminated Composite Plates
and Shells*, ch. 3):

* ply material: plane-stress reduced stiffness Q (with the through-thickness
  terms of a tri-axial material folded in, as the reference's material does)
  and the Tsai-Pagano invariants U1..U7;
* ply in laminate axes: Qbar(theta) written through the invariants
  (Qbar11 = U1 + U2 cos 2t + U3 cos 4t, ...), which is the textbook Q
  rotation in closed form;
* laminate: A, B, D, E = integrals of Qbar over z, z dz, z^2 dz, mass
  moments of rho likewise, all ply sums as dot products over the plies;
* shear correction (Vlachoutsis 1992): k = R^2 / (h int g^2 dz) with the
  integral of each ply evaluated exactly by 3-point Gauss-Legendre
  quadrature (the integrand is a degree-4 polynomial);
* lamination parameters (Tsai-Pagano / Fukunaga): the 14 xi of a stack and
  the stiffness of a (thickness, material, xi) design.

The public names, attributes and argument orders are pyfe3d's
(shellprop.pyx:128-865, shellprop_utils.py:13-206), the drop-in interface
of its property objects; ``record()`` packs a property into the 32-double
record the GPU kernels read.  Two quantities follow the reference's
definitions where they depart from the textbook, because the element
parity depends on them: the shear factors use ply moduli combined linearly
in cos and sin and a running sum over the plies of those moduli
(shellprop.pyx:507-546), and the 3-D material terms enter the invariants
(shellprop.pyx:128-196).
"""
import math

import numpy as np

RECORD_FIELDS = ['A11', 'A12', 'A16', 'A22', 'A26', 'A66',
                 'B11', 'B12', 'B16', 'B22', 'B26', 'B66',
                 'D11', 'D12', 'D16', 'D22', 'D26', 'D66',
                 'E44', 'E45', 'E55', 'scf_k13', 'scf_k23', 'h', 'offset',
                 'intrho', 'intrhoz', 'intrhoz2']
RECORD_STRIDE = 32
_VOIGT = ('11', '12', '16', '22', '26', '66')

# 3-point Gauss-Legendre nodes and weights on [-1, 1] (exact to degree 5)
_GAUSS3 = (np.array([-math.sqrt(0.6), 0., math.sqrt(0.6)]), np.array([5., 8., 5.]) / 9.)


def deg2rad(thetadeg):
    return np.asarray(thetadeg, dtype=np.float64) * (math.pi / 180.)


def _trig(theta_deg):
    """(cos 2t, sin 2t, cos 4t, sin 4t) of angles in degrees, stacked last."""
    t = deg2rad(theta_deg)
    return np.stack([np.cos(2 * t), np.sin(2 * t), np.cos(4 * t), np.sin(4 * t)], axis=-1)


def _membrane_terms(u, trig):
    """Stiffness terms (11, 12, 16, 22, 26, 66) from the invariants
    u = (U1..U5) and trigonometric coefficients (..., 4); trig may be ply
    angles' cos/sin (Qbar) or lamination parameters (xi).  Returns (..., 6)."""
    u1, u2, u3, u4, u5 = u[:5]
    c2, s2, c4, s4 = (trig[..., k] for k in range(4))
    return np.stack([u1 + u2 * c2 + u3 * c4,
                     u4 - u3 * c4,
                     0.5 * u2 * s2 + u3 * s4,
                     u1 - u2 * c2 + u3 * c4,
                     0.5 * u2 * s2 - u3 * s4,
                     u5 - u3 * c4], axis=-1)


def _membrane_gradient(u):
    """d(terms 11..66) / d(c2, s2, c4, s4): (6, 4), constant in the angles."""
    u2, u3 = u[1], u[2]
    return np.array([[u2, 0., u3, 0.], [0., 0., -u3, 0.], [0., 0.5 * u2, 0., u3],
                     [-u2, 0., u3, 0.], [0., 0.5 * u2, 0., -u3], [0., 0., -u3, 0.]])


def _shear_terms(u6, u7, c2, s2):
    """Transverse shear terms (44, 45, 55): Qbar44 = U6 + U7 cos 2t,
    Qbar45 = -U7 sin 2t, Qbar55 = U6 - U7 cos 2t."""
    return np.stack([u6 + u7 * c2, -u7 * s2, u6 - u7 * c2], axis=-1)


def _invariants(q11, q12, q22, q44, q55, q66):
    """Tsai-Pagano invariants U1..U7 of a reduced stiffness."""
    u1 = (3 * q11 + 3 * q22 + 2 * q12 + 4 * q66) / 8.
    u2 = (q11 - q22) / 2.
    u3 = (q11 + q22 - 2 * q12 - 4 * q66) / 8.
    u4 = (q11 + q22 + 6 * q12 - 4 * q66) / 8.
    u5 = (q11 + q22 - 2 * q12 + 4 * q66) / 8.
    return u1, u2, u3, u4, u5, (q44 + q55) / 2., (q44 - q55) / 2.


class MatLamina:
    """Orthotropic ply material (pyfe3d's MatLamina, shellprop.pyx:128-256).

    Engineering constants e1 e2 e3 g12 g13 g23 nu12 nu21 nu13 nu31 nu23 nu32
    rho; ``rebuild()`` derives the in-plane stiffness q11 q12 q22 (from the
    3-D orthotropic compliance: for e3 = 0 these are the plane-stress
    values E1/(1 - nu12 nu21), ...), q44 = g23, q55 = g13, q66 = g12 and
    the invariants u1..u7."""

    def __init__(self):
        for k in ('e1 e2 e3 g12 g13 g23 nu12 nu21 nu13 nu31 nu23 nu32 rho '
                  'q11 q12 q22 q44 q55 q66 u1 u2 u3 u4 u5 u6 u7').split():
            setattr(self, k, 0.)

    def rebuild(self):
        n12, n21, n13, n31, n23, n32 = (self.nu12, self.nu21, self.nu13, self.nu31,
                                        self.nu23, self.nu32)
        # determinant of the normalised 3x3 normal compliance
        det = 1. - n12 * n21 - n23 * n32 - n13 * n31 - 2. * n21 * n32 * n13
        self.q11 = self.e1 * (1. - n23 * n32) / det
        self.q12 = self.e1 * (n21 + n23 * n31) / det
        self.q22 = self.e2 * (1. - n13 * n31) / det
        self.q44, self.q55, self.q66 = self.g23, self.g13, self.g12
        (self.u1, self.u2, self.u3, self.u4, self.u5, self.u6,
         self.u7) = _invariants(self.q11, self.q12, self.q22, self.q44, self.q55, self.q66)

    def invariants(self):
        return np.array([self.u1, self.u2, self.u3, self.u4, self.u5, self.u6, self.u7])

    def trace_normalize_plane_stress(self):
        """Divide q and u by the trace Q11 + Q22 + 2 Q66 (Melo, Bi and Tsai
        2017, Eq. 1), as shellprop.pyx:198-232."""
        tr = self.q11 + self.q22 + 2 * self.q66
        for k in ('q11 q12 q22 q44 q55 q66 u1 u2 u3 u4 u5 u6 u7').split():
            setattr(self, k, getattr(self, k) / tr)

    def get_invariant_matrix(self):
        """Rows q11 q22 q12 q66 q16 q26 q44 q45 q55 over the columns
        (1, cos 2t, sin 2t, cos 4t, sin 4t) (shellprop.pyx:245-256)."""
        u1, u2, u3, u4, u5, u6, u7 = self.invariants()
        return np.array([[u1, u2, 0, u3, 0], [u1, -u2, 0, u3, 0], [u4, 0, 0, -u3, 0],
                         [u5, 0, 0, -u3, 0], [0, 0, u2 / 2., 0, u3], [0, 0, u2 / 2., 0, -u3],
                         [u6, u7, 0, 0, 0], [0, 0, -u7, 0, 0], [u6, -u7, 0, 0, 0]])


def _plane_stress_invariants(m):
    """Invariants of the ply's plane-stress stiffness (the in-plane
    constants only: the reference's plies use E1/(1 - nu12 nu21) whatever
    e3 is, shellprop.pyx:297-305)."""
    d = 1. - m.nu12 * m.nu21
    return np.array(_invariants(m.e1 / d, m.nu12 * m.e2 / d, m.e2 / d, m.g23, m.g13, m.g12))


class Lamina:
    """One ply (pyfe3d's Lamina, shellprop.pyx:259-386): plyid, h, thetadeg,
    matlamina; ``rebuild()`` sets the angle functions and the stiffness in
    laminate axes q11L q12L q16L q22L q26L q66L q44L q45L q55L."""

    def __init__(self):
        self.plyid = 0
        self.h = 0.
        self.thetadeg = 0.
        self.matlamina = None

    def rebuild(self):
        t = float(deg2rad(self.thetadeg))
        self.cost, self.sint = math.cos(t), math.sin(t)
        self.cos2t, self.sin2t = math.cos(2 * t), math.sin(2 * t)
        self.cos4t, self.sin4t = math.cos(4 * t), math.sin(4 * t)
        u = _plane_stress_invariants(self.matlamina)
        trig = np.array([self.cos2t, self.sin2t, self.cos4t, self.sin4t])
        for name, val in zip(_VOIGT, _membrane_terms(u, trig)):
            setattr(self, 'q%sL' % name, float(val))
        self.q44L, self.q45L, self.q55L = (float(x) for x in
                                           _shear_terms(u[5], u[6], self.cos2t, self.sin2t))

    def get_transf_matrix_displ_to_laminate(self):
        c, s = self.cost, self.sint
        return np.array([[c, s, 0.], [-s, c, 0.], [0., 0., 1.]])

    def get_constitutive_matrix(self):
        return np.array([[self.q11L, self.q12L, self.q16L, 0, 0],
                         [self.q12L, self.q22L, self.q26L, 0, 0],
                         [self.q16L, self.q26L, self.q66L, 0, 0],
                         [0, 0, 0, self.q44L, self.q45L],
                         [0, 0, 0, self.q45L, self.q55L]])

    def _stress_rotation(self, sign):
        c, s = self.cost, self.sint
        cs = s * c
        return np.array([[c * c, s * s, 0, 0, 0, sign * self.sin2t],
                         [s * s, c * c, 0, 0, 0, -sign * self.sin2t],
                         [0, 0, 1, 0, 0, 0],
                         [0, 0, 0, c, -sign * s, 0],
                         [0, 0, 0, sign * s, c, 0],
                         [-sign * cs, sign * cs, 0, 0, 0, c * c - s * s]])

    def get_transf_matrix_stress_to_lamina(self):
        return self._stress_rotation(1.)

    def get_transf_matrix_stress_to_laminate(self):
        return self._stress_rotation(-1.)


class LaminationParameters:
    """The 14 lamination parameters of first-order shear deformation
    theory: xiA1..4, xiB1..4, xiD1..4, xiE1..2 (shellprop.pyx:389-412)."""

    NAMES = ['xiA1', 'xiA2', 'xiA3', 'xiA4', 'xiB1', 'xiB2', 'xiB3', 'xiB4',
             'xiD1', 'xiD2', 'xiD3', 'xiD4', 'xiE1', 'xiE2']

    def __init__(self):
        for k in self.NAMES:
            setattr(self, k, 0.)

    def vector(self):
        return np.array([getattr(self, k) for k in self.NAMES])


class ShellProp:
    """Laminate stiffness, shear correction, mass moments (pyfe3d's
    ShellProp, shellprop.pyx:415-669)."""

    def __init__(self):
        for k in RECORD_FIELDS + ['e1', 'e2', 'g12', 'nu12', 'nu21']:
            setattr(self, k, 0.)
        self.scf_k13 = self.scf_k23 = 5 / 6.
        self.plies = []
        self.stack = []

    def _mat(self, X, sym=True):
        t = [getattr(self, X + k) for k in _VOIGT]
        return np.array([[t[0], t[1], t[2]], [t[1], t[3], t[4]], [t[2], t[4], t[5]]])

    @property
    def A(self):
        return self._mat('A')

    @property
    def B(self):
        return self._mat('B')

    @property
    def D(self):
        return self._mat('D')

    @property
    def E(self):
        return np.array([[self.E44, self.E45], [self.E45, self.E55]])

    @property
    def ABD(self):
        return np.block([[self.A, self.B], [self.B, self.D]])

    @property
    def ABDE(self):
        out = np.zeros((8, 8))
        out[:6, :6] = self.ABD
        out[6:, 6:] = self.E
        return out

    # -------------------------------------------------- ply integrals
    def _interfaces(self):
        t = np.array([p.h for p in self.plies], dtype=np.float64)
        if t.size == 0:
            raise ValueError('ShellProp with 0 plies!')
        h = float(t.sum())
        z = np.empty(t.size + 1)
        z[0] = -h / 2. + self.offset
        z[1:] = z[0] + np.cumsum(t)
        return h, z[:-1], z[1:]

    def calc_constitutive_matrix(self):
        """A = sum Qbar_k (z_k - z_k-1), B = sum Qbar_k (z_k^2 - z_k-1^2)/2,
        D = sum Qbar_k (z_k^3 - z_k-1^3)/3, E = sum Qbar_k^shear (z_k -
        z_k-1); intrho, intrhoz, intrhoz2 = integrals of rho, rho z, rho z^2."""
        h, z1, z2 = self._interfaces()
        self.h = h
        for p in self.plies:
            p.rebuild()
        Q = np.array([[getattr(p, 'q%sL' % k) for k in _VOIGT] for p in self.plies])
        S = np.array([[p.q44L, p.q45L, p.q55L] for p in self.plies])
        rho = np.array([p.matlamina.rho for p in self.plies])
        w1, w2, w3 = z2 - z1, (z2 * z2 - z1 * z1) / 2., (z2 ** 3 - z1 ** 3) / 3.
        for X, w in (('A', w1), ('B', w2), ('D', w3)):
            for k, val in zip(_VOIGT, Q.T @ w):
                setattr(self, X + k, float(val))
        self.E44, self.E45, self.E55 = (float(x) for x in S.T @ w1)
        self.intrho, self.intrhoz, self.intrhoz2 = float(rho @ w1), float(rho @ w2), \
            float(rho @ w3)

    def calc_equivalent_properties(self):
        """Apparent membrane moduli from the laminate compliance
        (ABD)^-1 (shellprop.pyx:551-565)."""
        a = np.linalg.inv(self.ABD)
        self.e1 = 1. / (self.h * a[0, 0])
        self.e2 = 1. / (self.h * a[1, 1])
        self.g12 = 1. / (self.h * a[2, 2])
        self.nu12 = -a[0, 1] / a[0, 0]
        self.nu21 = -a[0, 1] / a[1, 1]

    def calc_scf(self):
        """Shear correction factors k13, k23 (Vlachoutsis 1992): with
        s = z - offset, sb the bottom surface and D(z) the bending modulus,
        k = R^2 / (h sum_k D_k^2 int_ply (s^2 - sb^2)^2 ds / 4),
        R = sum_k D_k int_ply s^2 ds.  D_k as the reference defines it: the
        running sum over the plies up to k of E1'/(1 - nu12' nu21') with
        E1' = e1 cos t + e2 sin t, nu12' = nu12 cos t + nu21 sin t (and the
        1 <-> 2 counterparts for k23) (shellprop.pyx:485-548; the shear
        moduli there multiply and divide out)."""
        h, z1, z2 = self._interfaces()
        t = deg2rad([p.thetadeg for p in self.plies])
        c, s = np.cos(t), np.sin(t)
        e1 = np.array([p.matlamina.e1 for p in self.plies])
        e2 = np.array([p.matlamina.e2 for p in self.plies])
        n12 = np.array([p.matlamina.nu12 for p in self.plies])
        n21 = np.array([p.matlamina.nu21 for p in self.plies])
        d = 1. - (n12 * c + n21 * s) * (n21 * c + n12 * s)
        D1 = np.cumsum((e1 * c + e2 * s) / d)
        D2 = np.cumsum((e2 * c + e1 * s) / d)
        s1, s2 = z1 - self.offset, z2 - self.offset
        sb = s1[0]
        r = (s2 ** 3 - s1 ** 3) / 3.
        # int (s^2 - sb^2)^2 ds over each ply by 3-point Gauss (exact)
        xg, wg = _GAUSS3
        half, mid = (s2 - s1) / 2., (s2 + s1) / 2.
        sg = mid[:, None] + half[:, None] * xg[None, :]
        g = ((sg - sb) * (sg + sb)) ** 2
        F = half * (g @ wg)
        self.scf_k13 = float((D1 @ r) ** 2 / (h * ((D1 * D1) @ F) / 4.))
        self.scf_k23 = float((D2 @ r) ** 2 / (h * ((D2 * D2) @ F) / 4.))

    def force_balanced(self):
        """A16 = A26 = B16 = B26 = 0 (shellprop.pyx:623-635)."""
        if self.offset != 0.:
            raise RuntimeError('Laminates with offset cannot be forced balanced!')
        self.A16 = self.A26 = self.B16 = self.B26 = 0.

    def force_orthotropic(self):
        """A16 = A26 = B16 = B26 = D16 = D26 = 0 (shellprop.pyx:637-651)."""
        if self.offset != 0.:
            raise RuntimeError('Laminates with offset cannot be forced orthotropic!')
        self.A16 = self.A26 = self.B16 = self.B26 = self.D16 = self.D26 = 0.

    def force_symmetric(self):
        """B = 0 (shellprop.pyx:653-667)."""
        if self.offset != 0.:
            raise RuntimeError('Laminates with offset cannot be forced symmetric!')
        for k in _VOIGT:
            setattr(self, 'B' + k, 0.)

    def calc_lamination_parameters(self):
        """xiA = sum (zb_k - zb_k-1) f(t_k), xiB = 2 sum (zb_k^2 - zb_k-1^2) f,
        xiD = 4 sum (zb_k^3 - zb_k-1^3) f with zb = z/h and f = (cos 2t,
        sin 2t, cos 4t, sin 4t); xiE = sum (zb_k - zb_k-1) (cos 2t, sin 2t)
        (shellprop.pyx:669-720)."""
        h, z1, z2 = self._interfaces()
        for p in self.plies:
            p.rebuild()
        f = _trig([p.thetadeg for p in self.plies])
        a1, a2 = z1 / h, z2 / h
        lp = LaminationParameters()
        xi = np.concatenate([(a2 - a1) @ f, 2. * (a2 * a2 - a1 * a1) @ f,
                             4. * (a2 ** 3 - a1 ** 3) @ f, (a2 - a1) @ f[:, :2]])
        for k, val in zip(LaminationParameters.NAMES, xi):
            setattr(lp, k, float(val))
        return lp

    def record(self):
        """The 32-double record read by the GPU kernels."""
        out = np.zeros(RECORD_STRIDE)
        out[:len(RECORD_FIELDS)] = [getattr(self, f) for f in RECORD_FIELDS]
        return out


def force_balanced_LP(lp):
    """xiA2 = xiA4 = 0 (shellprop.pyx:722-732)."""
    lp.xiA2 = lp.xiA4 = 0.
    return lp


def force_symmetric_LP(lp):
    """xiB = 0 (shellprop.pyx:734-746)."""
    lp.xiB1 = lp.xiB2 = lp.xiB3 = lp.xiB4 = 0.
    return lp


def force_orthotropic_LP(lp):
    """xiA2 = xiA4 = xiB2 = xiB4 = xiD2 = xiD4 = 0 (shellprop.pyx:748-765)."""
    lp.xiA2 = lp.xiA4 = lp.xiB2 = lp.xiB4 = lp.xiD2 = lp.xiD4 = 0.
    return lp


def lp_stiffness(thickness, u, xi):
    """A, B, D (each (..., 6) over 11 12 16 22 26 66) and E ((..., 3) over
    44 45 55) of laminates of total thickness h made of one material with
    invariants u = (U1..U7) and lamination parameters xi (..., 14):
    A = h (U0 + G xiA), B = h^2/4 G xiB, D = h^3/12 (U0 + G xiD),
    E = h (U6 + U7 xiE1, -U7 xiE2, U6 - U7 xiE1), with U0 = (U1, U4, 0, U1,
    0, U5) and G the constant gradient of the terms in the xi."""
    h = np.asarray(thickness, dtype=np.float64)[..., None]
    xi = np.asarray(xi, dtype=np.float64)
    u0 = np.array([u[0], u[3], 0., u[0], 0., u[4]])
    G = _membrane_gradient(u)
    A = h * (u0 + xi[..., 0:4] @ G.T)
    B = h * h / 4. * (xi[..., 4:8] @ G.T)
    D = h ** 3 / 12. * (u0 + xi[..., 8:12] @ G.T)
    E = h * _shear_terms(u[5], u[6], xi[..., 12], xi[..., 13])
    return A, B, D, E


def shellprop_from_LaminationParameters(thickness, mat, lp):
    """ShellProp of thickness, MatLamina and LaminationParameters: A, B, D,
    E and h set; shear factors, offset and mass moments keep their defaults
    (shellprop.pyx:767-825)."""
    prop = ShellProp()
    prop.h = thickness
    A, B, D, E = lp_stiffness(thickness, mat.invariants(), lp.vector())
    for X, vals in (('A', A), ('B', B), ('D', D)):
        for k, val in zip(_VOIGT, vals):
            setattr(prop, X + k, float(val))
    prop.E44, prop.E45, prop.E55 = (float(x) for x in E)
    return prop


def shellprop_from_lamination_parameters(thickness, matlamina, xiA1, xiA2, xiA3, xiA4,
                                         xiB1, xiB2, xiB3, xiB4, xiD1, xiD2, xiD3, xiD4,
                                         xiE1=0., xiE2=0.):
    """The same from the 14 parameters as numbers (shellprop.pyx:828-865)."""
    lp = LaminationParameters()
    for k, val in zip(LaminationParameters.NAMES, (xiA1, xiA2, xiA3, xiA4, xiB1, xiB2, xiB3,
                                                   xiB4, xiD1, xiD2, xiD3, xiD4, xiE1, xiE2)):
        setattr(lp, k, float(val))
    return shellprop_from_LaminationParameters(thickness, matlamina, lp)


def read_laminaprop(laminaprop, rho=0):
    """MatLamina from (E, nu) (isotropic), (E1, E2, nu12, G12, G13, G23)
    (plane stress) or (E1, E2, nu12, G12, G13, G23, E3, nu13, nu23)
    (shellprop_utils.py:13-93); the minor Poisson ratios follow from
    reciprocity nu_ji = nu_ij E_j / E_i."""
    if len(laminaprop) not in (2, 6, 9):
        raise ValueError('Invalid entry for laminaprop: ' + str(laminaprop))
    if len(laminaprop) == 2:
        e, nu = laminaprop
        g = e / (2 * (1 + nu))
        laminaprop = (e, e, nu, g, g, g, 0, 0, 0)
    elif len(laminaprop) == 6:
        laminaprop = tuple(laminaprop) + (0, 0, 0)
    m = MatLamina()
    (m.e1, m.e2, m.nu12, m.g12, m.g13, m.g23, m.e3, m.nu13, m.nu23) = (float(x) for x in
                                                                       laminaprop)
    m.nu21 = m.nu12 * m.e2 / m.e1
    m.nu31 = m.nu13 * m.e3 / m.e1
    m.nu32 = m.nu23 * m.e3 / m.e2
    m.rho = float(rho)
    m.rebuild()
    return m


def laminated_plate(stack, plyt=None, laminaprop=None, rho=0., plyts=None, laminaprops=None,
                    rhos=None, offset=0., calc_scf=True):
    """ShellProp of a ply stack (angles in degrees, bottom to top) with one
    or per-ply thickness, material and density (shellprop_utils.py:96-180)."""
    n = len(stack)

    def per_ply(one, many, what):
        if many is None:
            if one is None:
                raise ValueError('%s or %ss must be supplied' % (what, what))
            return [one] * n
        return list(many)

    plyts = per_ply(plyt, plyts, 'plyt')
    laminaprops = per_ply(laminaprop, laminaprops, 'laminaprop')
    rhos = [rho] * n if rhos is None else list(rhos)
    if not (n == len(plyts) == len(laminaprops) == len(rhos)):
        raise ValueError('stack, plyts, laminaprops and rhos must have the same length')
    prop = ShellProp()
    prop.offset = offset
    prop.stack = list(stack)
    for k, (theta, t, lp, r) in enumerate(zip(stack, plyts, laminaprops, rhos)):
        ply = Lamina()
        ply.plyid = k
        ply.thetadeg = float(theta)
        ply.h = float(t)
        ply.matlamina = read_laminaprop(lp, r)
        prop.plies.append(ply)
    prop.calc_constitutive_matrix()
    prop.calc_equivalent_properties()
    if calc_scf:
        prop.calc_scf()
    return prop


def isotropic_plate(thickness, E, nu, offset=0., calc_scf=True, rho=0.):
    """ShellProp of one isotropic layer (shellprop_utils.py:182-206)."""
    return laminated_plate(plyt=thickness, stack=[0], laminaprop=(E, nu), rho=rho,
                           offset=offset, calc_scf=calc_scf)


def isotropic_records(h, E, nu, rho=0., calc_scf=True):
    """Records of isotropic plates for an array of thicknesses (one per
    element of a design sweep): A, E and intrho scale with h, D and
    intrhoz2 with h^3, B and intrhoz vanish; the shear factors do not
    depend on h."""
    h = np.asarray(h, dtype=np.float64)
    unit = isotropic_plate(1.0, E, nu, calc_scf=calc_scf, rho=rho).record()
    power = np.zeros(RECORD_STRIDE)
    power[0:6] = power[18:21] = power[23] = power[25] = 1
    power[12:18] = power[27] = 3
    rec = unit[None, :] * h[:, None] ** power[None, :]
    rec[:, 6:12] = 0.
    rec[:, 24] = rec[:, 26] = 0.
    return rec


def material_row(laminaprop, rho=0.):
    """Material row (e1 e2 nu12 g12 g13 g23 rho 0) of the device laminate
    builder from a 2- or 6-tuple laminaprop."""
    m = read_laminaprop(laminaprop, rho)
    return [m.e1, m.e2, m.nu12, m.g12, m.g13, m.g23, m.rho, 0.]


def _dev_f64(a, shape=None):
    import torch
    t = a if torch.is_tensor(a) else torch.as_tensor(np.asarray(a, dtype=np.float64))
    t = t.to(device='cuda', dtype=torch.float64).contiguous()
    return t if shape is None else t.reshape(shape)


def laminate_records_device(theta_deg, plyt, mats, ply_mat=None, offset=None, calc_scf=True,
                            out=None, stream=None):
    """Records (n, 32) on the GPU of one laminate per element
    (pf3d_laminate_records): theta_deg, plyt (n, n_plies) ply angles and
    thicknesses (zero-thickness plies are padding), mats (n_mat, 8) rows of
    material_row, ply_mat (n, n_plies) material indices or None, offset (n)
    or None.  Arrays are torch tensors on the device or host arrays."""
    import torch
    from . import _lib
    th, tk = _dev_f64(theta_deg), _dev_f64(plyt)
    if th.dim() != 2 or th.shape != tk.shape:
        raise ValueError('theta_deg and plyt must both be (n, n_plies)')
    n, n_plies = th.shape
    mt = _dev_f64(mats, (-1, 8))
    pm = None
    if ply_mat is not None:
        pm = torch.as_tensor(np.asarray(ply_mat) if not torch.is_tensor(ply_mat) else ply_mat,
                             dtype=torch.int32, device='cuda').contiguous()
        if pm.shape != th.shape:
            raise ValueError('ply_mat must be (n, n_plies)')
        if int(pm.min()) < 0 or int(pm.max()) >= mt.shape[0]:
            raise ValueError('ply_mat outside [0, %d)' % mt.shape[0])
    off = None if offset is None else _dev_f64(offset, (n,))
    if out is None:
        out = torch.empty((n, RECORD_STRIDE), dtype=torch.float64, device='cuda')
    s = (stream if stream is not None else torch.cuda.current_stream()).cuda_stream
    _lib.call('pf3d_laminate_records', n, n_plies, th.data_ptr(), tk.data_ptr(),
              None if pm is None else pm.data_ptr(), mt.shape[0], mt.data_ptr(),
              None if off is None else off.data_ptr(), int(bool(calc_scf)), out.data_ptr(), s)
    return out


def lp_material_row(mat, rho=0.):
    """Material row (U1..U7, rho) of the device lamination-parameter
    builder from a MatLamina."""
    return list(mat.invariants()) + [float(rho)]


def lp_records_device(thickness, xi, mats, mat_id=None, out=None, stream=None):
    """Records (n, 32) on the GPU from per-element lamination parameters
    (pf3d_lp_records; shellprop_from_LaminationParameters per element):
    thickness (n), xi (n, 14) in LaminationParameters.NAMES order, mats
    (n_mat, 8) rows of lp_material_row, mat_id (n) int or None (material
    0).  Shear factors 5/6, offset 0; with rho > 0 the mass moments of a
    homogeneous plate (intrho = rho h, intrhoz2 = rho h^3/12)."""
    import torch
    from . import _lib
    h = _dev_f64(thickness)
    n = h.numel()
    x = _dev_f64(xi, (n, 14))
    mt = _dev_f64(mats, (-1, 8))
    mid = None
    if mat_id is not None:
        mid = torch.as_tensor(np.asarray(mat_id) if not torch.is_tensor(mat_id) else mat_id,
                              dtype=torch.int32, device='cuda').contiguous().reshape(n)
        if n and (int(mid.min()) < 0 or int(mid.max()) >= mt.shape[0]):
            raise ValueError('mat_id outside [0, %d)' % mt.shape[0])
    if out is None:
        out = torch.empty((n, RECORD_STRIDE), dtype=torch.float64, device='cuda')
    s = (stream if stream is not None else torch.cuda.current_stream()).cuda_stream
    _lib.call('pf3d_lp_records', n, h.data_ptr(), x.data_ptr(),
              None if mid is None else mid.data_ptr(), mt.shape[0], mt.data_ptr(),
              out.data_ptr(), s)
    return out

"""The CPU oracle against the reference's golden vectors and against
invariants fixed by the mechanics (CPU only)."""
import numpy as np
import pytest

from golden_cases import SHELL_MATS, SHELLS, TOL, TOL_CASE, cases, element_errors
from oracle import oracle as O


def _run(etype, d, key, mat, mtype):
    E = d['conn'].shape[0]
    size = d[key + 'v'].size // E
    r = np.zeros(size * E, np.int64)
    c = np.zeros_like(r)
    v = np.zeros(size * E)
    beam = etype == 'beamc'
    O.assemble(etype, mat, d['x'], d['conn'], size * d['perm'],
               d['vxy'] if beam else d['xmat'], np.arange(E) if beam else d['pid'],
               d['props'], r, c, v, mtype=mtype,
               k6rot=None if beam else d['k6rot'], alpha=None if beam else d['alpha'],
               hg=None if beam else d['hg'], u=d['u'], stress=None if beam else d['stress'])
    return r, c, v, size


@pytest.mark.parametrize('etype', SHELLS)
def test_shell_golden(etype):
    for name, d in cases(etype):
        geom, _ = O.geometry(etype, d['x'], d['conn'], d['xmat'])
        np.testing.assert_allclose(geom, d['geom'], rtol=0, atol=1e-14)
        for key, mat, mtype in SHELL_MATS:
            r, c, v, size = _run(etype, d, key, mat, mtype)
            assert (r == d[key + 'r']).all() and (c == d[key + 'c']).all(), (name, key)
            err = element_errors(v, d[key + 'v'], size, d['perm'])
            assert err <= TOL_CASE.get((etype, name), TOL), (etype, name, key, err)


def test_beam_golden():
    for name, d in cases('beamc'):
        geom, _ = O.geometry('beamc', d['x'], d['conn'], d['vxy'])
        np.testing.assert_allclose(geom[:, :9], d['geom'][:, :9], rtol=0, atol=1e-14)
        np.testing.assert_allclose(geom[:, 13], d['geom'][:, 9], rtol=1e-14)
        for key, mat, mtype in (('KC0', 'KC0', 0), ('KG', 'KG', 0), ('M0', 'M', 0), ('M1', 'M', 1)):
            r, c, v, size = _run('beamc', d, key, mat, mtype)
            assert (r == d[key + 'r']).all() and (c == d[key + 'c']).all(), (name, key)
            assert element_errors(v, d[key + 'v'], size, d['perm']) <= TOL, (name, key)


def test_beam_mass_and_kg_invariants():
    """Lumped mass: each node carries half of intrho*L per translation;
    consistent mass has the same total translational mass; KG vanishes for a
    rigid translation (no axial force) and is linear in N."""
    L, E, A = 0.8, 70e9, 8e-5
    bp = np.zeros(16)
    bp[:9] = [A, E, E / 2.66 * 5 / 6, 1e-10, 4e-9, 0., 4.1e-9, 0., 0.]
    bp[9:15] = [0.216, 0., 0., 1e-6, 2e-7, 0.]
    X = np.array([[0.1, 0.2, 0.], [0.1 + L, 0.2, 0.]])
    vxy = np.array([0., 1., 0.])
    M1 = _dense_element('beamc', 'M', X, vxy, bp, mtype=1)
    M0 = _dense_element('beamc', 'M', X, vxy, bp, mtype=0)
    t = np.zeros(12); t[0::6] = 1
    assert np.isclose(t @ M1 @ t, bp[9] * L, rtol=1e-14)
    assert np.isclose(t @ M0 @ t, bp[9] * L, rtol=1e-14)
    u = np.zeros(12); u[6] = 1e-4  # stretch: N = E A 1e-4 / L
    K1 = _dense_element('beamc', 'KG', X, vxy, bp, u=u)
    K2 = _dense_element('beamc', 'KG', X, vxy, bp, u=2 * u)
    assert np.allclose(K2, 2 * K1, rtol=1e-14, atol=0)
    K0 = _dense_element('beamc', 'KG', X, vxy, bp, u=np.ones(12) * 1e-3 * t)
    assert np.abs(K0).max() == 0


def _dense_element(etype, mat, X, evec, prop, mtype=0, u=None, stress=None, k6rot=150.):
    nn = X.shape[0]
    conn = np.arange(nn)[None, :]
    size = O.slot_count(etype, mat, mtype)
    r = np.zeros(size, np.int64); c = np.zeros_like(r); v = np.zeros(size)
    O.assemble(etype, mat, X.ravel(), conn, np.zeros(1, np.int64), evec[None], np.zeros(1, np.int64),
               prop[None], r, c, v, mtype=mtype, k6rot=np.array([k6rot]), alpha=np.array([0.7]),
               hg=np.ones((1, 5)), u=u, stress=stress)
    K = np.zeros((6 * nn, 6 * nn))
    np.add.at(K, (r, c), v)
    return K


def _iso_prop(h=0.01, E=70e9, nu=0.3, rho=2700.):
    p = np.zeros(32)
    Q = E / (1 - nu ** 2)
    A = np.array([Q, nu * Q, 0, Q, 0, (1 - nu) / 2 * Q]) * h
    p[0:6] = A
    p[12:18] = A * h ** 2 / 12
    G = E / 2 / (1 + nu)
    p[18], p[20] = G * h, G * h
    p[21] = p[22] = 5 / 6.
    p[23] = h
    p[25], p[26], p[27] = rho * h, 0., rho * h ** 3 / 12
    return p


def _rigid_modes(X):
    """6 rigid-body modes of a node set in global DOFs (rotations about origin)."""
    nn = X.shape[0]
    modes = []
    for t in range(3):
        m = np.zeros(6 * nn)
        m[t::6] = 1
        modes.append(m)
    for ax in range(3):
        w = np.zeros(3); w[ax] = 1
        m = np.zeros(6 * nn)
        for i in range(nn):
            m[6 * i:6 * i + 3] = np.cross(w, X[i])
            m[6 * i + 3:6 * i + 6] = w
        modes.append(m)
    return np.array(modes)


@pytest.mark.parametrize('etype', ['quad4', 'quad4r', 'tria3r'])
def test_shell_invariants(etype):
    """KC0 is symmetric, PSD and annihilates the rigid-body modes of a flat
    element (drilling rz consistent with theta_z); mass sums to
    intrho*area per direction.  Quad4R's hourglass vector is orthogonal to
    linear fields only on parallelograms, and Tria3R keeps only the rz-rz
    diagonal of its drilling penalty (tria3r.pyx:2325-2974), so these two are
    checked on a parallelogram and without the drilling penalty."""
    from scipy.spatial.transform import Rotation
    R = Rotation.from_euler('zyx', [0.3, -0.5, 0.9]).as_matrix()
    if etype == 'quad4':
        P = np.array([[0, 0, 0], [0.11, 0.01, 0], [0.12, 0.09, 0], [-0.01, 0.1, 0]])
    else:
        P = np.array([[0, 0, 0], [0.11, 0.02, 0], [0.14, 0.12, 0], [0.03, 0.1, 0]])
    if etype == 'tria3r':
        P = P[:3]
    X = P @ R.T + np.array([0.3, -0.2, 0.5])
    prop = _iso_prop()
    K = _dense_element(etype, 'KC0', X, np.zeros(3), prop,
                       k6rot=0. if etype == 'tria3r' else 150.)
    scale = np.abs(K).max()
    assert np.abs(K - K.T).max() <= 1e-12 * scale
    ev = np.linalg.eigvalsh((K + K.T) / 2)
    assert ev.min() >= -1e-10 * scale
    modes = _rigid_modes(X)
    for m in modes:
        assert np.abs(K @ m).max() <= 1e-8 * scale * np.abs(m).max()
    M = _dense_element(etype, 'M', X, np.zeros(3), prop, mtype=0)
    area = 0.5 * np.linalg.norm(np.cross(P[2] - P[0], P[-1] - P[1] if etype != 'tria3r' else P[1] - P[0]))
    if etype == 'tria3r':
        area = 0.5 * np.linalg.norm(np.cross(P[1] - P[0], P[2] - P[0]))
    tr = modes[0]
    assert abs(tr @ M @ tr - prop[25] * area) <= 1e-12 * prop[25] * area


def test_quad4_membrane_is_plane_stress_q4():
    """Rectangular flat Quad4: the (u, v) block equals the textbook bilinear
    plane-stress element with 2x2 Gauss quadrature (plus drilling u/v terms,
    removed by setting the drilling-free reference below)."""
    a, b, h, E, nu = 0.2, 0.1, 0.01, 70e9, 0.3
    X = np.array([[0, 0, 0], [a, 0, 0], [a, b, 0], [0, b, 0]], float)
    prop = _iso_prop(h, E, nu)
    K = _dense_element('quad4', 'KC0', X, np.zeros(3), prop)
    idx = np.array([[6 * i, 6 * i + 1] for i in range(4)]).ravel()
    Kuv = K[np.ix_(idx, idx)]
    # textbook Q4 plane stress, independent implementation
    D = E / (1 - nu ** 2) * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]]) * h
    g = 1 / np.sqrt(3)
    Kt = np.zeros((8, 8)); Kd = np.zeros((8, 8))
    xy = X[:, :2]
    for xi in (-g, g):
        for eta in (-g, g):
            dN = 0.25 * np.array([[-(1 - eta), (1 - eta), (1 + eta), -(1 + eta)],
                                  [-(1 - xi), -(1 + xi), (1 + xi), (1 - xi)]])
            N = 0.25 * np.array([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta),
                                 (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)])
            J = dN @ xy
            dNx = np.linalg.solve(J, dN)
            B = np.zeros((3, 8))
            B[0, 0::2] = dNx[0]; B[1, 1::2] = dNx[1]
            B[2, 0::2] = dNx[1]; B[2, 1::2] = dNx[0]
            Kt += B.T @ D @ B * np.linalg.det(J)
            Bd = np.zeros(8)  # unit-weight drilling (u: N,y/2; v: -N,x/2)
            Bd[0::2] = dNx[1] / 2; Bd[1::2] = -dNx[0] / 2
            Kd += np.outer(Bd, Bd) * np.linalg.det(J)
    np.testing.assert_allclose(Kuv, Kt + Kd, rtol=0, atol=1e-12 * np.abs(Kt).max())


def test_beam_euler_bernoulli_limit():
    """BeamC with G -> infinity (alpha -> 0) is the Euler-Bernoulli frame element."""
    L, E, A, Iyy, Izz, J = 1.3, 200e9, 3e-4, 2e-8, 5e-8, 1e-8
    bp = np.zeros(16)
    bp[:7] = [A, E, 1e30, Iyy, Izz, 0., J]
    X = np.array([[0, 0, 0], [L, 0, 0]], float)
    K = _dense_element('beamc', 'KC0', X, np.array([0., 1., 0.]), bp)
    assert np.isclose(K[0, 0], E * A / L, rtol=1e-12)
    assert np.isclose(K[1, 1], 12 * E * Izz / L ** 3, rtol=1e-12)
    assert np.isclose(K[2, 2], 12 * E * Iyy / L ** 3, rtol=1e-12)
    assert np.isclose(K[1, 5], 6 * E * Izz / L ** 2, rtol=1e-12)
    assert np.isclose(K[5, 5], 4 * E * Izz / L, rtol=1e-12)
    assert np.isclose(K[5, 11], 2 * E * Izz / L, rtol=1e-12)
    assert np.isclose(K[4, 4], 4 * E * Iyy / L, rtol=1e-12)
    assert np.isclose(K[3, 3], 1e30 * J / L, rtol=1e-12)


FINT_TOL = 1e-12


def _fint_oracle(etype, d):
    beam = etype == 'beamc'
    E = d['conn'].shape[0]
    return O.fint(etype, d['x'], d['conn'], d['vxy'] if beam else d['xmat'],
                  np.arange(E) if beam else d['pid'], d['props'], d['u'],
                  k6rot=None if beam else d['k6rot'], alpha=None if beam else d['alpha'],
                  hg=None if beam else d['hg'])


@pytest.mark.parametrize('etype', SHELLS + ['beamc'])
def test_fint_golden(etype):
    """update_fint summed over every element (quad4.pyx:1316, beamc.pyx:1354)."""
    for name, d in cases(etype):
        f = _fint_oracle(etype, d)
        scale = np.abs(d['fint']).max()
        tol = TOL_CASE.get((etype, name), FINT_TOL)
        assert np.abs(f - d['fint']).max() <= tol * scale, (etype, name)


@pytest.mark.parametrize('etype', ['quad4', 'quad4r', 'beamc'])
def test_fint_is_kc0_times_u(etype):
    """For these elements the internal forces are KC0 u (linear statics);
    Tria3R is not: its KC0 scatter keeps only the rz-rz diagonal of the
    drilling term while its internal forces keep every coupling."""
    import scipy.sparse as ss
    for name, d in cases(etype):
        r, c, v, _ = _run(etype, d, 'KC0', 'KC0', 0)
        n = d['u'].size
        K = ss.coo_matrix((v, (r, c)), shape=(n, n)).tocsr()
        ku = K @ d['u']
        f = _fint_oracle(etype, d)
        assert np.abs(f - ku).max() <= 1e-11 * np.abs(ku).max(), (etype, name)


def _rod_evec(etype, d):
    """Truss: no orientation input (axes from the nodes); Spring: x axis and
    XY-plane vector per element."""
    E = d['conn'].shape[0]
    if etype == 'truss':
        return np.zeros((E, 3))
    return np.concatenate([d['axes'], d['vxy']], axis=1)


def _rod_props(etype, d):
    return d['props'] if etype == 'truss' else d['k']


ROD_MATS = {'truss': (('KC0', 'KC0', 0), ('M0', 'M', 0), ('M1', 'M', 1)),
            'spring': (('KC0', 'KC0', 0),)}


@pytest.mark.parametrize('etype', ['truss', 'spring'])
def test_rod_golden(etype):
    """Truss (truss.pyx:203-1800) and Spring (spring.pyx:147-707): rotation
    matrix, length, every KC0 / mass slot and the internal forces."""
    for name, d in cases(etype):
        E = d['conn'].shape[0]
        ev = _rod_evec(etype, d)
        geom, _ = O.geometry(etype, d['x'], d['conn'], ev)
        np.testing.assert_allclose(geom[:, :9], d['geom'][:, :9], rtol=0, atol=1e-14)
        if etype == 'truss':
            np.testing.assert_allclose(geom[:, 13], d['geom'][:, 9], rtol=1e-14)
        for key, mat, mtype in ROD_MATS[etype]:
            size = d[key + 'v'].size // E
            r = np.zeros(size * E, np.int64); c = np.zeros_like(r); v = np.zeros(size * E)
            O.assemble(etype, mat, d['x'], d['conn'], size * d['perm'], ev, np.arange(E),
                       _rod_props(etype, d), r, c, v, mtype=mtype)
            assert (r == d[key + 'r']).all() and (c == d[key + 'c']).all(), (name, key)
            assert element_errors(v, d[key + 'v'], size, d['perm']) <= TOL, (name, key)
        f = O.fint(etype, d['x'], d['conn'], ev, np.arange(E), _rod_props(etype, d), d['u'])
        assert np.abs(f - d['fint']).max() <= FINT_TOL * np.abs(d['fint']).max(), name
        # linear elements: the internal forces are KC0 u
        import scipy.sparse as ss
        size = d['KC0v'].size // E
        r = np.zeros(size * E, np.int64); c = np.zeros_like(r); v = np.zeros(size * E)
        O.assemble(etype, 'KC0', d['x'], d['conn'], size * d['perm'], ev, np.arange(E),
                   _rod_props(etype, d), r, c, v)
        n = d['u'].size
        ku = ss.coo_matrix((v, (r, c)), shape=(n, n)).tocsr() @ d['u']
        assert np.abs(f - ku).max() <= 1e-12 * np.abs(ku).max(), name


def test_rod_invariants():
    """Truss: a rigid translation produces no force, total translational
    mass is rho A L for both mass types, the axial stiffness of a rotated
    bar is E A / L along its axis; Spring: stiffness along the given axes."""
    L, A, Em = 1.3, 2e-4, 70e9
    bp = np.zeros(16)
    bp[:7] = [A, Em, Em / 2.6, 0, 0, 0, 3e-8]
    bp[9:14] = [2.7e3 * A, 1e-3, -2e-3, 4e-6, 5e-6]
    d = np.array([0.3, -0.5, 0.8]); d /= np.linalg.norm(d)
    X = np.array([[0.1, 0.2, 0.3], [0.1, 0.2, 0.3] + L * d])
    K = _dense_element('truss', 'KC0', X, np.zeros(3), bp)
    t = np.zeros(12); t[0::6] = 1
    assert np.abs(K @ t).max() <= 1e-9 * np.abs(K).max()
    s = np.zeros(12); s[6:9] = d  # stretch along the axis
    assert np.isclose(s @ K @ s, Em * A / L, rtol=1e-13)
    for mtype in (0, 1):
        M = _dense_element('truss', 'M', X, np.zeros(3), bp, mtype=mtype)
        assert np.isclose(t @ M @ t, bp[9] * L, rtol=1e-14)
        assert np.allclose(M, M.T, rtol=0, atol=1e-12 * np.abs(M).max())
    kk = np.array([1e3, 2e3, 3e3, 4e3, 5e3, 6e3, 0, 0])
    ax = np.array([0., 0., 1.]); vxy = np.array([1., 0., 0.])  # x_e = Z, y_e = X, z_e = Y
    Ks = _dense_element('spring', 'KC0', X, np.concatenate([ax, vxy]), kk)
    assert np.isclose(Ks[2, 2], 1e3) and np.isclose(Ks[0, 0], 2e3) and np.isclose(Ks[1, 1], 3e3)
    assert np.isclose(Ks[5, 5], 4e3) and np.isclose(Ks[3, 3], 5e3) and np.isclose(Ks[4, 4], 6e3)
    assert np.isclose(Ks[2, 8], -1e3)


def test_beamlr_golden():
    """BeamLR (beamlr.pyx): axes, length, KC0, KG (bending rotations only),
    consistent and lumped mass, internal forces."""
    for name, d in cases('beamlr'):
        E = d['conn'].shape[0]
        geom, _ = O.geometry('beamlr', d['x'], d['conn'], d['vxy'])
        np.testing.assert_allclose(geom[:, :9], d['geom'][:, :9], rtol=0, atol=1e-14)
        np.testing.assert_allclose(geom[:, 13], d['geom'][:, 9], rtol=1e-14)
        for key, mat, mtype in (('KC0', 'KC0', 0), ('KG', 'KG', 0), ('M0', 'M', 0), ('M1', 'M', 1)):
            size = d[key + 'v'].size // E
            r = np.zeros(size * E, np.int64); c = np.zeros_like(r); v = np.zeros(size * E)
            O.assemble('beamlr', mat, d['x'], d['conn'], size * d['perm'], d['vxy'], np.arange(E),
                       d['props'], r, c, v, mtype=mtype, u=d['u'])
            assert (r == d[key + 'r']).all() and (c == d[key + 'c']).all(), (name, key)
            assert element_errors(v, d[key + 'v'], size, d['perm']) <= TOL, (name, key)
        f = O.fint('beamlr', d['x'], d['conn'], d['vxy'], np.arange(E), d['props'], d['u'])
        assert np.abs(f - d['fint']).max() <= FINT_TOL * np.abs(d['fint']).max(), name


def test_beamlr_invariants():
    """Rigid-body modes give no force (the reduced integration keeps the
    shear terms consistent with rotations); the bending stiffness of a
    cantilever tip load matches Timoshenko's w = P L^3 / 3EI + P L / (G A)
    for one element with 1-point shear and linear fields (exactly: the
    element's flexibility L^3/(4EI) + L/(GA)); total mass rho A L."""
    L, A, Em, G = 0.9, 2e-4, 70e9, 26e9
    Iyy, Izz = 3e-9, 5e-9
    bp = np.zeros(16)
    bp[:9] = [A, Em, G, Iyy, Izz, 0., Iyy + Izz, 0., 0.]
    bp[9:15] = [2700 * A, 0., 0., 2700 * Izz, 2700 * Iyy, 0.]
    X = np.array([[0.1, -0.2, 0.3], [0.1 + L, -0.2, 0.3]])
    vxy = np.array([0., 1., 0.])
    K = _dense_element('beamlr', 'KC0', X, vxy, bp)
    # rigid translations and rotation about z through node 1
    for t in range(3):
        m = np.zeros(12); m[t::6] = 1.
        assert np.abs(K @ m).max() <= 1e-6 * np.abs(K).max()
    m = np.zeros(12); m[5] = m[11] = 1.; m[7] = L   # rz = 1, v = x
    assert np.abs(K @ m).max() <= 1e-6 * np.abs(K).max()
    # cantilever: node 1 clamped, tip force in y
    Kf = K[6:, 6:]
    f = np.zeros(6); f[1] = 1.
    w = np.linalg.solve(Kf, f)[1]
    assert np.isclose(w, L ** 3 / (4 * Em * Izz) + L / (G * A), rtol=1e-12)
    for mtype in (0, 1):
        M = _dense_element('beamlr', 'M', X, vxy, bp, mtype=mtype)
        t = np.zeros(12); t[0::6] = 1
        assert np.isclose(t @ M @ t, 2700 * A * L, rtol=1e-14)


@pytest.mark.parametrize('etype', ['quad4', 'quad4r'])
def test_piston_theory_golden(etype):
    """KA_beta, KA_gamma, CA (quad4.pyx:9491-11917, quad4r.pyx:12789-15200):
    translations (I, a, J, b), 144 slots, against the reference's vectors."""
    for name, d in cases(etype):
        E = d['conn'].shape[0]
        for key in ('KAB', 'KAG', 'CA'):
            r, c, v, size = _run(etype, d, key, key, 0)
            assert size == 144
            assert (r == d[key + 'r']).all() and (c == d[key + 'c']).all(), (name, key)
            err = element_errors(v, d[key + 'v'], size, d['perm'])
            assert err <= TOL_CASE.get((etype, name), TOL), (name, key, err)


def test_piston_theory_invariants():
    """On a flat rectangle in the XY plane: KA_gamma is the consistent
    'mass' pattern area*(4,2,1)/36 on w, CA = -KA_gamma, and KA_beta's rows
    sum to zero (a uniform w has no slope) while its columns applied to
    w = X give -int N_I dA = -area/4 per node."""
    a, b = 0.3, 0.2
    X = np.array([[0., 0., 0.], [a, 0., 0.], [a, b, 0.], [0., b, 0.]])
    p = _iso_prop()
    G = _dense_element('quad4', 'KAG', X, np.zeros(3), p)[2::6, 2::6]
    area = a * b
    pat = np.array([[4, 2, 1, 2], [2, 4, 2, 1], [1, 2, 4, 2], [2, 1, 2, 4]]) * area / 36
    np.testing.assert_allclose(G, pat, rtol=1e-13)
    C = _dense_element('quad4', 'CA', X, np.zeros(3), p)[2::6, 2::6]
    np.testing.assert_allclose(C, -pat, rtol=1e-13)
    B = _dense_element('quad4', 'KAB', X, np.zeros(3), p)[2::6, 2::6]
    assert np.abs(B.sum(axis=1)).max() <= 1e-15
    np.testing.assert_allclose(B @ X[:, 0], -np.full(4, area / 4), rtol=1e-13)
    # only the w DOFs of a flat element in the XY plane are touched
    full = _dense_element('quad4', 'KAB', X, np.zeros(3), p)
    mask = np.zeros((24, 24), bool); mask[2::6, 2::6] = True
    assert np.abs(full[~mask]).max() == 0

"""The pyfe3d-style element classes (pyfe3d_amd.compat) driven the way the
reference's own tests drive pyfe3d: the static point-load plate of
tests/test_quad4_static_point_load.py written against the façade (its
known answer and fint == fext check), the per-element slots against the
batched launch (same bits) and the oracle, KG/M/piston and BeamC through
the façade, and the one-launch batch helpers."""
import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.linalg import cg

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _plate(nx, ny, a, b):
    xtmp = np.linspace(0, a, nx)
    ytmp = np.linspace(0, b, ny)
    xmesh, ymesh = np.meshgrid(xtmp, ytmp)
    ncoords = np.vstack((xmesh.T.flatten(), ymesh.T.flatten(),
                         np.zeros_like(ymesh.T.flatten()))).T
    nids = 1 + np.arange(ncoords.shape[0])
    nid_pos = dict(zip(nids, np.arange(len(nids))))
    nids_mesh = nids.reshape(nx, ny)
    n1s = nids_mesh[:-1, :-1].flatten()
    n2s = nids_mesh[1:, :-1].flatten()
    n3s = nids_mesh[1:, 1:].flatten()
    n4s = nids_mesh[:-1, 1:].flatten()
    return ncoords, nid_pos, list(zip(n1s, n2s, n3s, n4s))


def test_static_plate_quad_point_load_through_facade():
    """tests/test_quad4_static_point_load.py:12-132 with the façade's
    classes: wmax within 2 % of 6.496928101916171e-05 and fint == fext
    (atol 1e-5); KC0 also equal to the oracle's slots (1e-12)."""
    from pyfe3d_amd.compat import DOF, DOUBLE, INT, Quad4, Quad4Data, Quad4Probe, isotropic_plate
    data = Quad4Data()
    probe = Quad4Probe()
    nx, ny, a, b = 7, 11, 3, 7
    ncoords, nid_pos, quads_n = _plate(nx, ny, a, b)
    x, y = ncoords[:, 0], ncoords[:, 1]
    ncoords_flatten = ncoords.flatten()
    num_elements = len(quads_n)
    KC0r = np.zeros(data.KC0_SPARSE_SIZE * num_elements, dtype=INT)
    KC0c = np.zeros(data.KC0_SPARSE_SIZE * num_elements, dtype=INT)
    KC0v = np.zeros(data.KC0_SPARSE_SIZE * num_elements, dtype=DOUBLE)
    N = DOF * nx * ny
    prop = isotropic_plate(thickness=0.005, E=200e9, nu=0.3, calc_scf=True)
    quads = []
    init_k_KC0 = 0
    for n1, n2, n3, n4 in quads_n:
        quad = Quad4(probe)
        quad.n1, quad.n2, quad.n3, quad.n4 = n1, n2, n3, n4
        quad.c1 = DOF * nid_pos[n1]
        quad.c2 = DOF * nid_pos[n2]
        quad.c3 = DOF * nid_pos[n3]
        quad.c4 = DOF * nid_pos[n4]
        quad.init_k_KC0 = init_k_KC0
        quad.K6ROT = 100.
        quad.update_rotation_matrix(ncoords_flatten)
        quad.update_probe_xe(ncoords_flatten)
        quad.update_KC0(KC0r, KC0c, KC0v, prop)
        quads.append(quad)
        init_k_KC0 += data.KC0_SPARSE_SIZE
    # the oracle's slots for the same elements
    conn = np.array([[nid_pos[n] for n in q] for q in quads_n], dtype=np.int64)
    rr = np.zeros_like(KC0r); cc = np.zeros_like(KC0c); vv = np.zeros_like(KC0v)
    O.assemble('quad4', 'KC0', ncoords_flatten, conn, 576 * np.arange(num_elements),
               np.zeros((num_elements, 3)), np.zeros(num_elements, np.int64),
               prop.record()[None], rr, cc, vv)
    assert (KC0r == rr).all() and (KC0c == cc).all()
    per = np.abs(KC0v - vv).reshape(num_elements, -1).max(axis=1)
    assert (per <= 1e-12 * np.abs(vv).reshape(num_elements, -1).max(axis=1)).all()

    KC0 = coo_matrix((KC0v, (KC0r, KC0c)), shape=(N, N)).tocsc()
    bk = np.zeros(N, dtype=bool)
    check = np.isclose(x, 0.) | np.isclose(x, a) | np.isclose(y, 0) | np.isclose(y, b)
    bk[2::DOF] = check
    bk[0::DOF] = True
    bk[1::DOF] = True
    bu = ~bk
    fext = np.zeros(N)
    fmid = 1.
    check = np.isclose(x, a / 2) & np.isclose(y, b / 2)
    fext[2::DOF][check] = fmid
    KC0uu = KC0[bu, :][:, bu]
    uu, info = cg(KC0uu, fext[bu], atol=1e-9)
    assert info == 0
    u = np.zeros(N)
    u[bu] = uu
    w = u[2::DOF].reshape(nx, ny).T
    assert np.isclose(6.496928101916171e-05, w.max(), rtol=0.02), w.max()
    fint = np.zeros(N)
    for quad in quads:
        quad.update_probe_xe(ncoords_flatten)
        quad.update_probe_ue(u)
        quad.update_fint(fint, prop)
    Kku = KC0[bk, :][:, bu]
    fext[bk] = Kku @ u[bu]
    assert np.allclose(fint, fext, atol=1e-5)


def _elements(cls_name, probe_name, ncoords, nid_pos, quads_n, tri=False):
    import pyfe3d_amd.compat as C
    probe = getattr(C, probe_name)()
    out = []
    for k, q in enumerate(quads_n):
        nodes = [q[0], q[1], q[2]] if tri else list(q)
        e = getattr(C, cls_name)(probe)
        for i, n in enumerate(nodes):
            setattr(e, 'n%d' % (i + 1), n)
            setattr(e, 'c%d' % (i + 1), 6 * nid_pos[n])
        out.append(e)
    return out


@pytest.mark.parametrize('cls_name,probe_name,etype', [('Quad4', 'Quad4Probe', 'quad4'),
                                                       ('Quad4R', 'Quad4RProbe', 'quad4r'),
                                                       ('Tria3R', 'Tria3RProbe', 'tria3r')])
def test_facade_matches_batch_launch(cls_name, probe_name, etype):
    """Per-element façade calls (material direction, KC0, KG from u, M
    mtype 0) write the same slots, bit for bit, as the one-launch batch
    helpers, on a warped plate with a material direction; KG given stress
    and fint through the façade against the oracle."""
    import pyfe3d_amd.compat as C
    ncoords, nid_pos, quads_n = _plate(6, 5, 1.0, 0.7)
    rng = np.random.default_rng(1)
    ncoords[:, 2] = 0.02 * rng.uniform(-1, 1, len(ncoords))
    xf = ncoords.flatten()
    u = rng.uniform(-1, 1, 6 * len(ncoords)) * 1e-3
    prop = C.laminated_plate([0, 45, -45, 90], plyt=0.0005,
                             laminaprop=(140e9, 10e9, 0.3, 5e9, 4e9, 3e9), rho=1600.)
    els = _elements(cls_name, probe_name, ncoords, nid_pos, quads_n, tri=(etype == 'tria3r'))
    data = getattr(C, cls_name + 'Data')()
    E = len(els)
    for k, e in enumerate(els):
        e.init_k_KC0 = data.KC0_SPARSE_SIZE * k
        e.init_k_KG = data.KG_SPARSE_SIZE * k
        e.init_k_M = data.M_SPARSE_SIZE * k
        e.update_rotation_matrix(xf, 1., 0.3, 0.)
        e.update_probe_xe(xf)

    def arrays(size):
        return (np.zeros(size * E, np.int64), np.zeros(size * E, np.int64), np.zeros(size * E))

    k1, k2 = arrays(data.KC0_SPARSE_SIZE), arrays(data.KC0_SPARSE_SIZE)
    for e in els:
        e.update_KC0(*k1, prop)
    C.update_KC0_batch(els, xf, *k2, prop)
    for a1, a2 in zip(k1, k2):
        assert np.array_equal(a1, a2)
    g1, g2 = arrays(data.KG_SPARSE_SIZE), arrays(data.KG_SPARSE_SIZE)
    for e in els:
        e.update_probe_ue(u)
        e.update_KG(*g1, prop)
    C.update_KG_batch(els, xf, u, *g2, prop)
    for a1, a2 in zip(g1, g2):
        assert np.array_equal(a1, a2)
    m1, m2 = arrays(data.M_SPARSE_SIZE), arrays(data.M_SPARSE_SIZE)
    for e in els:
        e.update_M(*m1, prop, mtype=0)
    C.update_M_batch(els, xf, *m2, prop, mtype=0)
    for a1, a2 in zip(m1, m2):
        assert np.array_equal(a1, a2)
    # KG given stress and the internal forces against the oracle
    s1 = arrays(data.KG_SPARSE_SIZE)
    for e in els:
        e.update_KG_given_stress(-1e3, 200., 50., *s1)
    nn = 3 if etype == 'tria3r' else 4
    conn = np.array([[nid_pos[n] for n in q[:nn]] for q in quads_n], dtype=np.int64)
    size = data.KG_SPARSE_SIZE
    rr = np.zeros(size * E, np.int64); cc = np.zeros_like(rr); vv = np.zeros(size * E)
    O.assemble(etype, 'KGs', xf, conn, size * np.arange(E), np.tile([1., 0.3, 0.], (E, 1)),
               np.zeros(E, np.int64), prop.record()[None], rr, cc, vv,
               stress=np.array([-1e3, 200., 50.]))
    assert (s1[0] == rr).all() and (s1[1] == cc).all()
    assert np.abs(s1[2] - vv).max() <= 1e-12 * np.abs(vv).max()
    fint = np.zeros(6 * len(ncoords))
    for e in els:
        e.update_probe_ue(u)
        e.update_fint(fint, prop)
    ref = O.fint(etype, xf, conn, np.tile([1., 0.3, 0.], (E, 1)), np.zeros(E, np.int64),
                 prop.record()[None], u)
    assert np.abs(fint - ref).max() <= 1e-12 * np.abs(ref).max()


def test_beamc_and_piston_through_facade():
    """BeamC (KC0, KG from u, lumped M, fint) and Quad4 piston theory
    through the façade against the oracle."""
    import pyfe3d_amd.compat as C
    rng = np.random.default_rng(5)
    n = 9
    ncoords = np.stack([np.linspace(0, 2, n), 0.1 * rng.uniform(-1, 1, n),
                        0.1 * rng.uniform(-1, 1, n)], axis=1)
    xf = ncoords.flatten()
    u = rng.uniform(-1, 1, 6 * n) * 1e-3
    prop = C.BeamProp()
    prop.A, prop.E, prop.G = 8e-5, 70e9, 26e9
    prop.Iyy, prop.Izz, prop.J = 4e-10, 1e-9, 1.4e-9
    prop.intrho, prop.intrhoy2, prop.intrhoz2 = 2700 * 8e-5, 2700 * 1e-9, 2700 * 4e-10
    probe = C.BeamCProbe()
    data = C.BeamCData()
    E = n - 1
    KC0 = [np.zeros(144 * E, np.int64), np.zeros(144 * E, np.int64), np.zeros(144 * E)]
    KG = [np.zeros(144 * E, np.int64), np.zeros(144 * E, np.int64), np.zeros(144 * E)]
    M = [np.zeros(36 * E, np.int64), np.zeros(36 * E, np.int64), np.zeros(36 * E)]
    fint = np.zeros(6 * n)
    for k in range(E):
        b = C.BeamC(probe)
        b.n1, b.n2, b.c1, b.c2 = k + 1, k + 2, 6 * k, 6 * (k + 1)
        b.init_k_KC0, b.init_k_KG, b.init_k_M = 144 * k, 144 * k, 36 * k
        b.update_rotation_matrix(0., 1., 0.2, xf)
        b.update_probe_xe(xf)
        b.update_KC0(*KC0, prop)
        b.update_probe_ue(u)
        b.update_KG(*KG, prop)
        b.update_M(*M, prop, mtype=1)
        b.update_fint(fint, prop)
    assert data.KC0_SPARSE_SIZE == 144
    conn = np.stack([np.arange(E), np.arange(1, n)], axis=1).astype(np.int64)
    vxy = np.tile([0., 1., 0.2], (E, 1))
    rec = O.props_table(prop.record()[None], 'beamc')
    for key, arrs, size, mtype in (('KC0', KC0, 144, 0), ('KG', KG, 144, 0), ('M', M, 36, 1)):
        rr = np.zeros(size * E, np.int64); cc = np.zeros_like(rr); vv = np.zeros(size * E)
        O.assemble('beamc', key, xf, conn, size * np.arange(E), vxy, np.zeros(E, np.int64), rec,
                   rr, cc, vv, mtype=mtype, u=u)
        assert (arrs[0] == rr).all() and (arrs[1] == cc).all(), key
        assert np.abs(arrs[2] - vv).max() <= 1e-12 * np.abs(vv).max(), key
    ref = O.fint('beamc', xf, conn, vxy, np.zeros(E, np.int64), rec, u)
    assert np.abs(fint - ref).max() <= 1e-12 * np.abs(ref).max()
    # Quad4 piston theory (update_KA_beta / update_KA_gamma / update_CA)
    ncoords, nid_pos, quads_n = _plate(4, 5, 0.5, 0.4)
    xq = ncoords.flatten()
    els = _elements('Quad4', 'Quad4Probe', ncoords, nid_pos, quads_n)
    Eq = len(els)
    conn = np.array([[nid_pos[m] for m in q] for q in quads_n], dtype=np.int64)
    for key, meth, attr in (('KAB', 'update_KA_beta', 'init_k_KA_beta'),
                            ('KAG', 'update_KA_gamma', 'init_k_KA_gamma'),
                            ('CA', 'update_CA', 'init_k_CA')):
        arrs = [np.zeros(144 * Eq, np.int64), np.zeros(144 * Eq, np.int64), np.zeros(144 * Eq)]
        for k, e in enumerate(els):
            setattr(e, attr, 144 * k)
            e.update_rotation_matrix(xq)
            e.update_probe_xe(xq)
            getattr(e, meth)(*arrs)
        rr = np.zeros(144 * Eq, np.int64); cc = np.zeros_like(rr); vv = np.zeros(144 * Eq)
        O.assemble('quad4', key, xq, conn, 144 * np.arange(Eq), np.zeros((Eq, 3)),
                   np.zeros(Eq, np.int64), np.zeros((1, 32)), rr, cc, vv)
        assert (arrs[0] == rr).all() and (arrs[1] == cc).all(), key
        assert np.abs(arrs[2] - vv).max() <= 1e-12 * np.abs(vv).max(), key


@pytest.mark.parametrize('cls_name', ['BeamLR', 'Truss', 'Spring'])
def test_beamlr_truss_spring_through_facade(cls_name):
    """BeamLR (KC0, KG, M 0/1, fint), Truss (KC0, M 0/1, fint) and Spring
    (KC0, fint) per element through the façade against the oracle on a
    rotated 3-D polyline; BeamLR's batch helper gives the same bits."""
    import pyfe3d_amd.compat as C
    rng = np.random.default_rng(8)
    n = 8
    ncoords = np.cumsum(rng.uniform(0.1, 0.3, (n, 3)), axis=0)
    xf = ncoords.flatten()
    u = rng.uniform(-1, 1, 6 * n) * 1e-3
    E = n - 1
    etype = cls_name.lower()
    conn = np.stack([np.arange(E), np.arange(1, n)], axis=1).astype(np.int64)
    probe = getattr(C, cls_name + 'Probe')()
    data = getattr(C, cls_name + 'Data')()
    if cls_name == 'Spring':
        k = np.array([1e5, 2e5, 3e5, 4e2, 5e2, 6e2])
        xdir = rng.normal(size=(E, 3))
        vxy = rng.normal(size=(E, 3))
        ev = np.concatenate([xdir, vxy], axis=1)
        props = k[None]
    else:
        prop = C.BeamProp()
        prop.A, prop.E, prop.G, prop.Iyy, prop.Izz, prop.J = 2e-4, 70e9, 26e9, 3e-9, 5e-9, 6e-9
        prop.intrho, prop.intrhoy2, prop.intrhoz2 = 2700 * 2e-4, 2700 * 5e-9, 2700 * 3e-9
        props = O.props_table(prop.record()[None], etype)
        ev = np.tile([0.2, 1., 0.1], (E, 1)) if cls_name == 'BeamLR' else np.zeros((E, 3))
    keys = {'BeamLR': [('KC0', 0), ('KG', 0), ('M', 0), ('M', 1)],
            'Truss': [('KC0', 0), ('M', 0), ('M', 1)], 'Spring': [('KC0', 0)]}[cls_name]
    for key, mtype in keys:
        size = O.slot_count(etype, key, mtype)
        out = [np.zeros(size * E, np.int64), np.zeros(size * E, np.int64), np.zeros(size * E)]
        fint = np.zeros(6 * n)
        els = []
        for i in range(E):
            e = getattr(C, cls_name)(probe)
            e.n1, e.n2, e.c1, e.c2 = i + 1, i + 2, 6 * i, 6 * (i + 1)
            e.init_k_KC0 = e.init_k_KG = e.init_k_M = size * i
            if cls_name == 'Spring':
                e.kxe, e.kye, e.kze, e.krxe, e.krye, e.krze = k
                e.update_rotation_matrix(*xdir[i], *vxy[i])
            elif cls_name == 'BeamLR':
                e.update_rotation_matrix(*ev[i], xf)
                e.update_probe_xe(xf)
            else:
                e.update_rotation_matrix(xf)
                e.update_probe_xe(xf)
            e.update_probe_ue(u)
            if key == 'KC0':
                if cls_name == 'Spring':
                    e.update_KC0(*out)
                    e.update_fint(fint)
                else:
                    e.update_KC0(*out, prop)
                    e.update_fint(fint, prop)
            elif key == 'KG':
                e.update_KG(*out, prop)
            else:
                e.update_M(*out, prop, mtype=mtype)
            els.append(e)
        rr = np.zeros(size * E, np.int64); cc = np.zeros_like(rr); vv = np.zeros(size * E)
        O.assemble(etype, key, xf, conn, size * np.arange(E), ev, np.zeros(E, np.int64), props,
                   rr, cc, vv, mtype=mtype, u=u)
        assert (out[0] == rr).all() and (out[1] == cc).all(), key
        assert np.abs(out[2] - vv).max() <= 1e-12 * np.abs(vv).max(), key
        if key == 'KC0':
            ref = O.fint(etype, xf, conn, ev, np.zeros(E, np.int64), props, u)
            assert np.abs(fint - ref).max() <= 1e-12 * np.abs(ref).max()
        if cls_name == 'BeamLR':
            out2 = [np.zeros(size * E, np.int64), np.zeros(size * E, np.int64),
                    np.zeros(size * E)]
            if key == 'KC0':
                C.update_KC0_batch(els, xf, *out2, prop)
            elif key == 'KG':
                C.update_KG_batch(els, xf, u, *out2, prop)
            else:
                C.update_M_batch(els, xf, *out2, prop, mtype=mtype)
            for a1, a2 in zip(out, out2):
                assert np.array_equal(a1, a2), key
    assert data.KC0_SPARSE_SIZE == O.slot_count(etype, 'KC0')

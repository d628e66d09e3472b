"""Write the golden element-matrix fixtures under ``tests/golden/``.

This script is the only place that runs the reference (pyfe3d v0.7.0,
Cython).  It compiles the reference's ``.pyx`` sources from
``/root/reference/pyfe3d`` into ``tests/golden/_refbuild/`` (ignored by git
and never sent anywhere else), imports them, runs the per-element calls on
small seeded meshes and stores inputs and outputs as ``.npz`` data.  Nothing
of the reference travels with the fixtures: only arrays of numbers.

Per-slot values of ``update_KC0`` / ``update_KG`` / ``update_KG_given_stress``
/ ``update_M`` are not fixed entry-wise by any reference test (the reference
tests pin solution-level quantities at 1e-4..2e-2), so these vectors are the
entry-wise anchor for the oracle and for the GPU kernels.

Usage:  python tests/golden/make_golden.py
"""
import glob
import importlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = '/root/reference'
BUILD = os.path.join(HERE, '_refbuild')
MODULES = ['shellprop', 'beamprop', 'quad4', 'quad4r', 'tria3r', 'beamc', 'truss', 'spring',
           'beamlr']


def build_reference():
    """Compile the reference modules needed here into ``_refbuild``."""
    libdir = os.path.join(BUILD, 'lib')
    pkgdir = os.path.join(libdir, 'pyfe3d')
    have = all(glob.glob(os.path.join(pkgdir, m + '*.so')) for m in MODULES)
    if not have:
        from Cython.Build import cythonize
        from setuptools import Extension
        from setuptools.dist import Distribution
        import setuptools.command.build_ext  # noqa: F401
        cwd = os.getcwd()
        os.chdir(REF)  # cythonize resolves the package from relative paths
        try:
            exts = [Extension('pyfe3d.' + m, sources=['pyfe3d/%s.pyx' % m],
                              language='c++', extra_compile_args=['-O2'])
                    for m in MODULES]
            exts = cythonize(exts, build_dir=os.path.join(BUILD, 'cy'),
                             language_level='3', quiet=True)
            dist = Distribution({'ext_modules': exts})
            cmd = dist.get_command_obj('build_ext')
            cmd.build_lib = libdir
            cmd.build_temp = os.path.join(BUILD, 'tmp')
            cmd.inplace = False
            cmd.ensure_finalized()
            cmd.run()
        finally:
            os.chdir(cwd)
        os.makedirs(pkgdir, exist_ok=True)
        open(os.path.join(pkgdir, '__init__.py'), 'w').close()
    sys.path.insert(0, libdir)
    return {m: importlib.import_module('pyfe3d.' + m) for m in MODULES}


def load_shellprop_utils(mods):
    """Load the reference's pure-Python ``shellprop_utils`` module."""
    spec = importlib.util.spec_from_file_location(
        'pyfe3d.shellprop_utils', os.path.join(REF, 'pyfe3d', 'shellprop_utils.py'))
    mod = importlib.util.module_from_spec(spec)
    sys.modules['pyfe3d.shellprop_utils'] = mod
    spec.loader.exec_module(mod)
    return mod


SHELL_FIELDS = ['A11', 'A12', 'A16', 'A22', 'A26', 'A66',
                'B11', 'B12', 'B16', 'B22', 'B26', 'B66',
                'D11', 'D12', 'D16', 'D22', 'D26', 'D66',
                'E44', 'E45', 'E55', 'scf_k13', 'scf_k23', 'h', 'offset',
                'intrho', 'intrhoz', 'intrhoz2', 'e1', 'e2', 'g12', 'nu12',
                'nu21']
BEAM_FIELDS = ['A', 'E', 'G', 'Iyy', 'Izz', 'Iyz', 'J', 'Ay', 'Az', 'intrho',
               'intrhoy', 'intrhoz', 'intrhoy2', 'intrhoz2', 'intrhoyz']


def shell_record(p):
    return np.array([getattr(p, f) for f in SHELL_FIELDS], dtype=np.float64)


def rot3(yaw, pitch, roll):
    cz, sz = np.cos(yaw), np.sin(yaw)
    cy, sy = np.cos(pitch), np.sin(pitch)
    cx, sx = np.cos(roll), np.sin(roll)
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    return Rz @ Ry @ Rx


def grid(nx, ny, a, b, rng, jitter=0.0, warp=0.0, R=None, shift=(0, 0, 0)):
    """Nodes x-major (pos = i*ny + j) as in the reference plate tests."""
    xs = np.linspace(0, a, nx)
    ys = np.linspace(0, b, ny)
    X, Y = np.meshgrid(xs, ys, indexing='ij')
    X = X.copy(); Y = Y.copy()
    Z = np.zeros_like(X)
    dx, dy = a / (nx - 1), b / (ny - 1)
    if jitter:
        X[1:-1, 1:-1] += dx * rng.uniform(-jitter, jitter, (nx - 2, ny - 2))
        Y[1:-1, 1:-1] += dy * rng.uniform(-jitter, jitter, (nx - 2, ny - 2))
    if warp:
        Z += warp * min(dx, dy) * rng.uniform(-1, 1, X.shape)
    P = np.vstack([X.ravel(), Y.ravel(), Z.ravel()]).T
    if R is not None:
        P = P @ R.T
    P = P + np.asarray(shift)
    nids = np.arange(nx * ny).reshape(nx, ny)
    quads = np.array([[nids[i, j], nids[i + 1, j], nids[i + 1, j + 1], nids[i, j + 1]]
                      for i in range(nx - 1) for j in range(ny - 1)], dtype=np.int64)
    return P.ravel().astype(np.float64), quads


def tri_split(quads):
    out = []
    for n1, n2, n3, n4 in quads:
        out.append([n1, n2, n3])
        out.append([n1, n3, n4])
    return np.array(out, dtype=np.int64)


def shell_props(spu, kind, n, rng):
    """Return (list of reference ShellProp, (n, nfields) records)."""
    props = []
    for e in range(n):
        if kind == 'iso':
            p = spu.isotropic_plate(thickness=0.005, E=70e9, nu=0.33, rho=2700.,
                                    calc_scf=True)
        elif kind == 'iso_thick':
            p = spu.isotropic_plate(thickness=0.6, E=70e9, nu=0.33, rho=2700.,
                                    calc_scf=True)
        elif kind == 'lam':
            stack = [float(rng.choice([0, 45, -45, 90, 30, -60])) for _ in range(5)]
            lp = (140e9, 10e9, 0.3, 5e9, 4e9, 3e9)
            p = spu.laminated_plate(stack=stack, plyt=0.000125 * (1 + 0.2 * e),
                                    laminaprop=lp, rho=1600., offset=0.0001 * (e % 3),
                                    calc_scf=True)
        elif kind == 'hfield':
            h = 0.002 + 0.001 * rng.uniform(-1, 1)
            p = spu.isotropic_plate(thickness=h, E=70e9, nu=0.33, rho=2700.,
                                    calc_scf=True)
        else:
            raise ValueError(kind)
        props.append(p)
    return props, np.array([shell_record(p) for p in props])


SHELL_CASES = [
    # name, mesh args, prop kind, xmat, per-element props
    dict(name='flat_iso', nx=3, ny=4, a=1.0, b=0.7, jitter=0.0, warp=0.0,
         rot=None, prop='iso', xmat=None, per_elem=False),
    dict(name='distorted_lam_xmat', nx=4, ny=3, a=0.3, b=0.2, jitter=0.3, warp=0.05,
         rot=(0.4, -0.3, 0.7), prop='lam', xmat=(1.0, 0.3, 0.2), per_elem=True),
    dict(name='rotated_thick', nx=3, ny=3, a=0.8, b=0.6, jitter=0.2, warp=0.0,
         rot=(1.1, 0.5, -0.2), prop='iso_thick', xmat=None, per_elem=False),
    dict(name='threshold_hfield', nx=4, ny=4, a=0.006, b=0.006, jitter=0.25, warp=0.0,
         rot=None, prop='hfield', xmat=(0.2, 1.0, 0.0), per_elem=True),
]


def run_shell_case(mods, spu, etype, case, rng):
    mod = mods[etype]
    cname = {'quad4': 'Quad4', 'quad4r': 'Quad4R', 'tria3r': 'Tria3R'}[etype]
    cls = getattr(mod, cname)
    probecls = getattr(mod, cname + 'Probe')
    data = getattr(mod, cname + 'Data')()
    R = rot3(*case['rot']) if case['rot'] is not None else None
    x, quads = grid(case['nx'], case['ny'], case['a'], case['b'], rng,
                    case['jitter'], case['warp'], R, shift=(0.1, -0.2, 0.05))
    conn = quads if etype != 'tria3r' else tri_split(quads)
    nn = conn.shape[1]
    E = conn.shape[0]
    N = x.size // 3
    nprops = E if case['per_elem'] else 1
    props, precs = shell_props(spu, case['prop'], nprops, rng)
    pid = np.arange(E) if case['per_elem'] else np.zeros(E, dtype=np.int64)
    xmat = np.zeros((E, 3))
    if case['xmat'] is not None:
        xmat[:] = case['xmat']
        xmat += 0.05 * rng.uniform(-1, 1, xmat.shape)
    # element order in the slot arrays is a permutation: tests init_k handling
    perm = rng.permutation(E)
    sizes = dict(KC0=data.KC0_SPARSE_SIZE, KG=data.KG_SPARSE_SIZE, M=data.M_SPARSE_SIZE)
    k6rot = rng.uniform(10., 1000., E) if etype != 'quad4' else np.full(E, 100.)
    alpha = rng.uniform(0.3, 1.0, E) if etype == 'tria3r' else np.zeros(E)
    hg = rng.uniform(0.5, 2.0, (E, 5)) if etype == 'quad4r' else np.ones((E, 5))
    u = rng.uniform(-1, 1, 6 * N) * 1e-3
    Nxx, Nyy, Nxy = -1.3e3, 0.4e3, 0.25e3

    out = {}
    keys = ['KC0', 'KG', 'KGs', 'M0', 'M1', 'M2']
    piston = etype != 'tria3r'   # piston theory: Quad4 / Quad4R only
    if piston:
        sizes.update(KAB=data.KA_BETA_SPARSE_SIZE, KAG=data.KA_GAMMA_SPARSE_SIZE,
                     CA=data.CA_SPARSE_SIZE)
        keys += ['KAB', 'KAG', 'CA']
    for key in keys:
        base = 'M' if key.startswith('M') else ('KG' if key.startswith('KG') else key)
        n = sizes[base] * E
        out[key + 'r'] = np.zeros(n, dtype=np.int64)
        out[key + 'c'] = np.zeros(n, dtype=np.int64)
        out[key + 'v'] = np.zeros(n, dtype=np.float64)
    fint = np.zeros(6 * N)
    geom = np.zeros((E, 14))
    xe = np.zeros((E, 3 * nn))
    probe = probecls()
    for e in range(E):
        el = cls(probe)
        for i in range(nn):
            setattr(el, 'n%d' % (i + 1), int(conn[e, i]))
            setattr(el, 'c%d' % (i + 1), int(6 * conn[e, i]))
        el.init_k_KC0 = int(sizes['KC0'] * perm[e])
        el.init_k_KG = int(sizes['KG'] * perm[e])
        el.init_k_M = int(sizes['M'] * perm[e])
        if etype != 'quad4':
            el.K6ROT = float(k6rot[e])
        if etype == 'tria3r':
            el.alpha_shear_locking = float(alpha[e])
        p = props[pid[e]]
        el.update_rotation_matrix(x, *xmat[e])
        el.update_probe_xe(x)
        geom[e] = [el.r11, el.r12, el.r13, el.r21, el.r22, el.r23, el.r31, el.r32,
                   el.r33, el.m11, el.m12, el.m21, el.m22, el.area]
        xe[e] = np.asarray(probe.xe)
        if etype == 'quad4r':
            el.update_KC0(out['KC0r'], out['KC0c'], out['KC0v'], p, 0, *hg[e])
        else:
            el.update_KC0(out['KC0r'], out['KC0c'], out['KC0v'], p)
        if piston:
            el.init_k_KA_beta = int(sizes['KAB'] * perm[e])
            el.init_k_KA_gamma = int(sizes['KAG'] * perm[e])
            el.init_k_CA = int(sizes['CA'] * perm[e])
            el.update_KA_beta(out['KABr'], out['KABc'], out['KABv'])
            el.update_KA_gamma(out['KAGr'], out['KAGc'], out['KAGv'])
            el.update_CA(out['CAr'], out['CAc'], out['CAv'])
        el.update_M(out['M0r'], out['M0c'], out['M0v'], p, 0)
        el.update_M(out['M1r'], out['M1c'], out['M1v'], p, 1)
        el.update_M(out['M2r'], out['M2c'], out['M2v'], p, 2)
        el.update_KG_given_stress(Nxx, Nyy, Nxy, out['KGsr'], out['KGsc'], out['KGsv'])
        el.update_probe_ue(u)
        el.update_probe_xe(x)
        el.update_KG(out['KGr'], out['KGc'], out['KGv'], p)
        if etype == 'quad4r':
            el.update_fint(fint, p, *hg[e])
        else:
            el.update_fint(fint, p)
    return dict(x=x, conn=conn, perm=perm, pid=pid, props=precs, xmat=xmat,
                k6rot=k6rot, alpha=alpha, hg=hg, u=u,
                stress=np.array([Nxx, Nyy, Nxy]), geom=geom, xe=xe, fint=fint, **out)


def beam_record(p):
    return np.array([getattr(p, f) for f in BEAM_FIELDS], dtype=np.float64)


def run_beam_case(mods, case, rng):
    bc = mods['beamc']
    BeamProp = mods['beamprop'].BeamProp
    n = case['n']
    # a curved 3D polyline of nodes
    t = np.linspace(0, 1, n + 1)
    P = np.vstack([case['L'] * t, 0.2 * np.sin(2 * t) + 0.01 * rng.uniform(-1, 1, t.size),
                   0.1 * t ** 2]).T
    if case['rot'] is not None:
        P = P @ rot3(*case['rot']).T
    x = P.ravel().astype(np.float64)
    N = n + 1
    conn = np.array([[i, i + 1] for i in range(n)], dtype=np.int64)
    E = n
    props, precs = [], []
    for e in range(E):
        p = BeamProp()
        b_, h_ = 0.02 * (1 + 0.1 * e), 0.004 * (1 + 0.05 * e)
        E_, nu = 70e9, 0.33
        p.A = b_ * h_
        p.E = E_
        p.G = E_ / 2 / (1 + nu) * 5 / 6
        p.Izz = b_ ** 3 * h_ / 12
        p.Iyy = b_ * h_ ** 3 / 12
        p.Iyz = case['Iyz'] * p.Iyy
        p.J = p.Izz + p.Iyy
        p.Ay = case['Ay'] * p.A
        p.Az = case['Az'] * p.A
        rho = 2700.
        p.intrho = rho * p.A
        p.intrhoy = rho * case['Ay'] * 1e-3
        p.intrhoz = rho * case['Az'] * 1e-3
        p.intrhoy2 = rho * p.Izz
        p.intrhoz2 = rho * p.Iyy
        p.intrhoyz = rho * p.Iyz
        props.append(p)
        precs.append(beam_record(p))
    precs = np.array(precs)
    vxy = np.tile(case['vxy'], (E, 1)) + 0.1 * rng.uniform(-1, 1, (E, 3))
    perm = rng.permutation(E)
    d = bc.BeamCData()
    sizes = dict(KC0=d.KC0_SPARSE_SIZE, KG=d.KG_SPARSE_SIZE, M=d.M_SPARSE_SIZE)
    u = rng.uniform(-1, 1, 6 * N) * 1e-3
    out = {}
    for key in ['KC0', 'KG', 'M0', 'M1']:
        base = 'M' if key.startswith('M') else key
        m = sizes[base] * E
        out[key + 'r'] = np.zeros(m, dtype=np.int64)
        out[key + 'c'] = np.zeros(m, dtype=np.int64)
        out[key + 'v'] = np.zeros(m, dtype=np.float64)
    fint = np.zeros(6 * N)
    geom = np.zeros((E, 10))
    probe = bc.BeamCProbe()
    for e in range(E):
        el = bc.BeamC(probe)
        el.n1, el.n2 = int(conn[e, 0]), int(conn[e, 1])
        el.c1, el.c2 = int(6 * conn[e, 0]), int(6 * conn[e, 1])
        el.init_k_KC0 = int(sizes['KC0'] * perm[e])
        el.init_k_KG = int(sizes['KG'] * perm[e])
        el.init_k_M = int(sizes['M'] * perm[e])
        el.update_rotation_matrix(*vxy[e], x)
        el.update_probe_xe(x)
        geom[e] = [el.r11, el.r12, el.r13, el.r21, el.r22, el.r23, el.r31, el.r32,
                   el.r33, el.length]
        p = props[e]
        el.update_KC0(out['KC0r'], out['KC0c'], out['KC0v'], p)
        el.update_M(out['M0r'], out['M0c'], out['M0v'], p, 0)
        el.update_M(out['M1r'], out['M1c'], out['M1v'], p, 1)
        el.update_probe_ue(u)
        el.update_probe_xe(x)
        el.update_KG(out['KGr'], out['KGc'], out['KGv'], p)
        el.update_fint(fint, p)
    return dict(x=x, conn=conn, perm=perm, props=precs, vxy=vxy, u=u, geom=geom,
                fint=fint, **out)


def run_beamlr_case(mods, case, rng):
    """BeamLR: the BeamC case's geometry and section properties through the
    reduced-integration Timoshenko element (beamlr.pyx)."""
    bl = mods['beamlr']
    BeamProp = mods['beamprop'].BeamProp
    n = case['n']
    t = np.linspace(0, 1, n + 1)
    P = np.vstack([case['L'] * t, 0.2 * np.sin(2 * t) + 0.01 * rng.uniform(-1, 1, t.size),
                   0.1 * t ** 2]).T
    if case['rot'] is not None:
        P = P @ rot3(*case['rot']).T
    x = P.ravel().astype(np.float64)
    N = n + 1
    conn = np.array([[i, i + 1] for i in range(n)], dtype=np.int64)
    E = n
    props, precs = [], []
    for e in range(E):
        p = BeamProp()
        b_, h_ = 0.02 * (1 + 0.1 * e), 0.004 * (1 + 0.05 * e)
        E_, nu = 70e9, 0.33
        p.A = b_ * h_
        p.E = E_
        p.G = E_ / 2 / (1 + nu) * 5 / 6
        p.Izz = b_ ** 3 * h_ / 12
        p.Iyy = b_ * h_ ** 3 / 12
        p.Iyz = case['Iyz'] * p.Iyy
        p.J = p.Izz + p.Iyy
        p.Ay = case['Ay'] * p.A
        p.Az = case['Az'] * p.A
        rho = 2700.
        p.intrho = rho * p.A
        p.intrhoy = rho * case['Ay'] * 1e-3
        p.intrhoz = rho * case['Az'] * 1e-3
        p.intrhoy2 = rho * p.Izz
        p.intrhoz2 = rho * p.Iyy
        p.intrhoyz = rho * p.Iyz
        props.append(p)
        precs.append(beam_record(p))
    precs = np.array(precs)
    vxy = np.tile(case['vxy'], (E, 1)) + 0.1 * rng.uniform(-1, 1, (E, 3))
    perm = rng.permutation(E)
    d = bl.BeamLRData()
    sizes = dict(KC0=d.KC0_SPARSE_SIZE, KG=d.KG_SPARSE_SIZE, M=d.M_SPARSE_SIZE)
    u = rng.uniform(-1, 1, 6 * N) * 1e-3
    out = {}
    for key in ['KC0', 'KG', 'M0', 'M1']:
        base = 'M' if key.startswith('M') else key
        m = sizes[base] * E
        out[key + 'r'] = np.zeros(m, dtype=np.int64)
        out[key + 'c'] = np.zeros(m, dtype=np.int64)
        out[key + 'v'] = np.zeros(m, dtype=np.float64)
    fint = np.zeros(6 * N)
    geom = np.zeros((E, 10))
    probe = bl.BeamLRProbe()
    for e in range(E):
        el = bl.BeamLR(probe)
        el.n1, el.n2 = int(conn[e, 0]), int(conn[e, 1])
        el.c1, el.c2 = int(6 * conn[e, 0]), int(6 * conn[e, 1])
        el.init_k_KC0 = int(sizes['KC0'] * perm[e])
        el.init_k_KG = int(sizes['KG'] * perm[e])
        el.init_k_M = int(sizes['M'] * perm[e])
        el.update_rotation_matrix(*vxy[e], x)
        el.update_probe_xe(x)
        el.update_length()
        geom[e] = [el.r11, el.r12, el.r13, el.r21, el.r22, el.r23, el.r31, el.r32,
                   el.r33, el.length]
        p = props[e]
        el.update_KC0(out['KC0r'], out['KC0c'], out['KC0v'], p)
        el.update_M(out['M0r'], out['M0c'], out['M0v'], p, 0)
        el.update_M(out['M1r'], out['M1c'], out['M1v'], p, 1)
        el.update_probe_ue(u)
        el.update_KG(out['KGr'], out['KGc'], out['KGv'], p)
        el.update_fint(fint, p)
    return dict(x=x, conn=conn, perm=perm, props=precs, vxy=vxy, u=u, geom=geom,
                fint=fint, **out)


BEAMLR_CASES = [
    dict(name='curved_3d', n=5, L=1.0, rot=(0.3, 0.2, -0.4), vxy=(0., 1., 0.3),
         Iyz=0.0, Ay=0.0, Az=0.0),
    dict(name='coupled', n=4, L=0.5, rot=(1.2, -0.7, 0.1), vxy=(0.1, 0.2, 1.0),
         Iyz=0.3, Ay=0.2, Az=-0.1),
]


BEAM_CASES = [
    dict(name='curved_3d', n=5, L=1.0, rot=(0.3, 0.2, -0.4), vxy=(0., 1., 0.3),
         Iyz=0.0, Ay=0.0, Az=0.0),
    dict(name='coupled', n=4, L=0.5, rot=(1.2, -0.7, 0.1), vxy=(0.1, 0.2, 1.0),
         Iyz=0.3, Ay=0.2, Az=-0.1),
]


def run_truss_case(mods, case, rng):
    tr = mods['truss']
    BeamProp = mods['beamprop'].BeamProp
    n = case['n']
    t = np.linspace(0, 1, n + 1)
    P = np.vstack([case['L'] * t, 0.3 * np.sin(3 * t) + 0.01 * rng.uniform(-1, 1, t.size),
                   0.2 * t ** 2 - 0.1 * t]).T
    P = P @ rot3(*case['rot']).T
    x = P.ravel().astype(np.float64)
    N = n + 1
    conn = np.array([[i, i + 1] for i in range(n)], dtype=np.int64)
    E = n
    props, precs = [], []
    for e in range(E):
        p = BeamProp()
        r_ = 0.01 * (1 + 0.2 * e)
        p.A = np.pi * r_ ** 2
        p.E = 200e9 * (1 + 0.1 * e)
        p.G = p.E / 2 / 1.3
        p.J = np.pi * r_ ** 4 / 2
        p.Iyy = p.Izz = p.J / 2
        rho = 7850.
        p.intrho = rho * p.A
        p.intrhoy = rho * p.A * 1e-3 * (e - 2)
        p.intrhoz = rho * p.A * 2e-3 * (1 - e)
        p.intrhoy2 = rho * p.Izz
        p.intrhoz2 = rho * p.Iyy * 1.1
        props.append(p)
        precs.append(beam_record(p))
    precs = np.array(precs)
    perm = rng.permutation(E)
    d = tr.TrussData()
    sizes = dict(KC0=d.KC0_SPARSE_SIZE, M=d.M_SPARSE_SIZE)
    u = rng.uniform(-1, 1, 6 * N) * 1e-3
    out = {}
    for key in ['KC0', 'M0', 'M1']:
        base = 'M' if key.startswith('M') else key
        m = sizes[base] * E
        out[key + 'r'] = np.zeros(m, dtype=np.int64)
        out[key + 'c'] = np.zeros(m, dtype=np.int64)
        out[key + 'v'] = np.zeros(m, dtype=np.float64)
    fint = np.zeros(6 * N)
    geom = np.zeros((E, 10))
    probe = tr.TrussProbe()
    for e in range(E):
        el = tr.Truss(probe)
        el.n1, el.n2 = int(conn[e, 0]), int(conn[e, 1])
        el.c1, el.c2 = int(6 * conn[e, 0]), int(6 * conn[e, 1])
        el.init_k_KC0 = int(sizes['KC0'] * perm[e])
        el.init_k_M = int(sizes['M'] * perm[e])
        el.update_rotation_matrix(x)
        el.update_probe_xe(x)
        el.update_length()
        geom[e] = [el.r11, el.r12, el.r13, el.r21, el.r22, el.r23, el.r31, el.r32,
                   el.r33, el.length]
        p = props[e]
        el.update_KC0(out['KC0r'], out['KC0c'], out['KC0v'], p)
        el.update_M(out['M0r'], out['M0c'], out['M0v'], p, 0)
        el.update_M(out['M1r'], out['M1c'], out['M1v'], p, 1)
        el.update_probe_ue(u)
        el.update_fint(fint, p)
    return dict(x=x, conn=conn, perm=perm, props=precs, u=u, geom=geom, fint=fint, **out)


def run_spring_case(mods, case, rng):
    sp_ = mods['spring']
    N = case['nodes']
    x = rng.uniform(-1, 1, 3 * N)
    conn = np.array(case['pairs'], dtype=np.int64)
    E = conn.shape[0]
    kk = np.exp(rng.uniform(np.log(1e2), np.log(1e7), (E, 6)))
    axes = rng.normal(size=(E, 3))
    vxy = rng.normal(size=(E, 3))
    perm = rng.permutation(E)
    d = sp_.SpringData()
    size = d.KC0_SPARSE_SIZE
    u = rng.uniform(-1, 1, 6 * N) * 1e-3
    out = dict(KC0r=np.zeros(size * E, np.int64), KC0c=np.zeros(size * E, np.int64),
               KC0v=np.zeros(size * E))
    fint = np.zeros(6 * N)
    geom = np.zeros((E, 9))
    probe = sp_.SpringProbe()
    for e in range(E):
        el = sp_.Spring(probe)
        el.n1, el.n2 = int(conn[e, 0]), int(conn[e, 1])
        el.c1, el.c2 = int(6 * conn[e, 0]), int(6 * conn[e, 1])
        el.init_k_KC0 = int(size * perm[e])
        el.kxe, el.kye, el.kze, el.krxe, el.krye, el.krze = [float(v) for v in kk[e]]
        el.update_rotation_matrix(*[float(v) for v in axes[e]], *[float(v) for v in vxy[e]])
        geom[e] = [el.r11, el.r12, el.r13, el.r21, el.r22, el.r23, el.r31, el.r32, el.r33]
        el.update_KC0(out['KC0r'], out['KC0c'], out['KC0v'])
        el.update_probe_ue(u)
        el.update_fint(fint)
    return dict(x=x, conn=conn, perm=perm, k=kk, axes=axes, vxy=vxy, u=u, geom=geom,
                fint=fint, **out)


TRUSS_CASES = [dict(name='polyline_3d', n=5, L=1.2, rot=(0.4, -0.3, 0.9)),
               dict(name='short_rotated', n=3, L=0.2, rot=(-1.1, 0.6, 2.0))]
SPRING_CASES = [dict(name='mixed', nodes=5, pairs=[[0, 1], [1, 2], [2, 3], [3, 4], [0, 4], [1, 3]]),
                dict(name='chain', nodes=4, pairs=[[0, 1], [1, 2], [2, 3]])]


def shellprop_cases(spu):
    """Laminate builder vectors (ABD, E, scf, integrals) for the host table builder."""
    cases = []
    lp = (140e9, 10e9, 0.3, 5e9, 4e9, 3e9)
    specs = [
        dict(stack=[0], plyt=0.005, laminaprop=(70e9, 0.33), rho=2700., offset=0.),
        dict(stack=[0, 45, -45, 90, 90, -45, 45, 0], plyt=0.000125, laminaprop=lp,
             rho=1600., offset=0.),
        dict(stack=[30, -60, 0, 45, 90], plyt=0.0002, laminaprop=lp, rho=1500.,
             offset=0.0003),
        dict(stack=[0, 90], plyt=0.001,
             laminaprop=(140e9, 10e9, 0.3, 5e9, 4e9, 3e9, 9e9, 0.3, 0.45), rho=1000.,
             offset=-0.0002),
    ]
    recs = []
    for s in specs:
        p = spu.laminated_plate(calc_scf=True, **s)
        recs.append(shell_record(p))
    return specs, np.array(recs)


def write_rods(mods):
    """Truss and Spring fixtures (their own seed: the other files are not
    touched when only these are regenerated)."""
    rng = np.random.default_rng(20261203)
    for name, cases, fn in (('truss', TRUSS_CASES, run_truss_case),
                            ('spring', SPRING_CASES, run_spring_case),
                            ('beamlr', BEAMLR_CASES, run_beamlr_case)):
        res = {}
        for case in cases:
            r = fn(mods, case, rng)
            for k, v in r.items():
                res[case['name'] + '__' + k] = v
        np.savez_compressed(os.path.join(HERE, 'golden_%s.npz' % name), **res)
        print(name, 'cases', [c['name'] for c in cases])


if __name__ == '__main__':
    mods = build_reference()
    write_rods(mods)
    if '--rods-only' in sys.argv:
        sys.exit(0)
    spu = load_shellprop_utils(mods)
    rng = np.random.default_rng(20261202)
    for etype in ['quad4', 'quad4r', 'tria3r']:
        res = {}
        for case in SHELL_CASES:
            r = run_shell_case(mods, spu, etype, case, rng)
            for k, v in r.items():
                res[case['name'] + '__' + k] = v
        np.savez_compressed(os.path.join(HERE, 'golden_%s.npz' % etype), **res)
        print(etype, 'cases', [c['name'] for c in SHELL_CASES])
    res = {}
    for case in BEAM_CASES:
        r = run_beam_case(mods, case, rng)
        for k, v in r.items():
            res[case['name'] + '__' + k] = v
    np.savez_compressed(os.path.join(HERE, 'golden_beamc.npz'), **res)
    specs, recs = shellprop_cases(spu)
    np.savez_compressed(os.path.join(HERE, 'golden_shellprop.npz'), records=recs,
                        fields=np.array(SHELL_FIELDS))
    import json
    with open(os.path.join(HERE, 'golden_shellprop_specs.json'), 'w') as f:
        json.dump(specs, f, indent=1)
    print('done')

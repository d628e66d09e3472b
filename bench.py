"""Benchmark of the stiffness-assembly-and-solve hot path on one MI355X.

Headline (``value``): KC0 assembly throughput, elements/s, on the
1000x1000 Quad4 laminate plate (1,002,001 nodes, 6,012,006 DOF; interior
nodes jittered +-0.2 element size, seeded; 8 plies of 0.125 mm with seeded
angles from {0, 45, -45, 90}; E1=140e9, E2=10e9, G12=G13=G23=5e9,
nu12=0.3; a x b = 1.0 x 0.7 m).  One step = one KC0 assembly of all 1e6
elements into the 576e6-slot value array (values only, overwrite: the
design-sweep re-assembly of the reference's tutorials), inputs resident in
HBM.  Alongside: measured fill and copy bandwidths on the same buffer, the
same assembly with int64 row/col indices and accumulating, KG (given
stress), M, internal forces, and static solves per second at 1e-8 relative
residual (plain FP64 block-Jacobi PCG against multigrid CG with FP64 and
FP32 V-cycles, the double-double refinement where an FP64 vector cannot
reach 1e-8, banded direct solves of a design sweep, the stiffened panel,
the 6M-DOF plate, the learned warm start), with the panel's lowest natural
frequencies; the CPU baseline times the reference-equivalent path on the
host.

Prints one JSON line.  ``--gpus N`` with N > 1 exits: this build measures
one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
HBM_PEAK_BPS = HBM_PEAK_GBS * 1e9
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix peak
# HBM bytes per launch of the values-only Quad4 KC0 kernel on the 1000x1000
# plate, from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of
# `python bench.py --profile-kc0` (KiB per dispatch, mean of 3 launches),
# each scaled by the reading of the same counter for a 1 GiB device copy in
# the same run (FETCH_SIZE reads 0.5 GiB for it: factor 2.0; WRITE_SIZE
# 1 GiB: 1.0): 60,062 KiB x 2 + 4,500,740 KiB = 4.732e9 B (algorithmic:
# 4.736e9 B); profiles/kc0_quad4_1000x1000_pmc.txt.  A counter run cannot
# be nested in the timed run, so the figure is carried here.
KC0_TRAFFIC_BYTES = (60062 * 2 + 4500740) * 1024
# algorithmic bytes of one Quad4 KC0 element, values only: 576 doubles out,
# 4 int64 DOF offsets and 12 coordinates in
KC0_BYTES_VALUES = 576 * 8 + 4 * 8 + 12 * 8
KC0_BYTES_INDICES = KC0_BYTES_VALUES + 576 * 16


def laminate_record(seed=2026):
    from pyfe3d_amd.shellprop import laminated_plate
    rng = np.random.default_rng(seed)
    stack = [float(a) for a in rng.choice([0., 45., -45., 90.], 8)]
    lp = (140e9, 10e9, 0.3, 5e9, 5e9, 5e9)
    prop = laminated_plate(stack, plyt=0.000125, laminaprop=lp, rho=1600., calc_scf=True)
    return prop, stack


def build_plate(n, seed=2026):
    from pyfe3d_amd.mesh import PlateMesh
    return PlateMesh(n + 1, n + 1, 1.0, 0.7, jitter=0.2, seed=seed)


def timed(fn, steps, warmup, stream):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(steps):
        fn()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    return ev0.elapsed_time(ev1) / steps, wall


def _oracle_kc0(mesh, prop, n_sample, threads, values_only):
    from oracle import oracle as O
    q = mesh.quads[:n_sample]
    E = q.shape[0]
    r = np.zeros(576 * E, np.int64)
    c = np.zeros_like(r)
    v = np.zeros(576 * E)
    used = O.set_threads(threads)
    t0 = time.perf_counter()
    O.assemble('quad4', 'KC0', mesh.x, q, 576 * np.arange(E), np.zeros((E, 3)),
               np.zeros(E, np.int64), prop.record()[None], r, c, v, values_only=values_only)
    dt = time.perf_counter() - t0
    O.set_threads(1)
    return E / dt, dt, used


def _reference_path(n, log):
    """The reference's CPU path on the n x n laminate plate under pressure,
    piece by piece (tests/test_quad4_static_point_load.py:52-102): the
    per-element KC0 loop (oracle, indices + values, one core), coo_matrix
    -> tocsc, the free-DOF slice and spsolve (SuperLU)."""
    import scipy.sparse as ss
    import scipy.sparse.linalg as sla
    from oracle import oracle as O
    mesh = build_plate(n)
    prop, _ = laminate_record()
    E = mesh.quads.shape[0]
    r = np.zeros(576 * E, np.int64)
    c = np.zeros_like(r)
    v = np.zeros(576 * E)
    O.set_threads(1)
    t0 = time.perf_counter()
    O.assemble('quad4', 'KC0', mesh.x, mesh.quads, 576 * np.arange(E), np.zeros((E, 3)),
               np.zeros(E, np.int64), prop.record()[None], r, c, v)
    t_kc0 = time.perf_counter() - t0
    t0 = time.perf_counter()
    K = ss.coo_matrix((v, (r, c)), shape=(mesh.n_dof, mesh.n_dof)).tocsc()
    t_csc = time.perf_counter() - t0
    bu = ~mesh.clamped()
    f = mesh.pressure_load(1e3)
    t0 = time.perf_counter()
    Kuu = K[bu, :][:, bu]
    t_slice = time.perf_counter() - t0
    t0 = time.perf_counter()
    uu = sla.spsolve(Kuu, f[bu])
    t_solve = time.perf_counter() - t0
    rel = float(np.linalg.norm(f[bu] - Kuu @ uu) / np.linalg.norm(f[bu]))
    out = dict(dof=mesh.n_dof, elements=E, kc0_s=t_kc0, tocsc_s=t_csc, slice_s=t_slice,
               spsolve_s=t_solve, total_s=t_kc0 + t_csc + t_slice + t_solve,
               spsolve_relres_fp64=rel, cores=1)
    log('cpu reference path %dx%d: %s' % (n, n, out))
    return out


def cpu_baseline(mesh, prop, n_sample, log, ladder=(20, 64, 128)):
    """The reference-equivalent CPU path timed on this host (the reference
    itself cannot travel with the benchmark): the oracle (a plain-C
    restatement of the per-element loop, pinned to the reference's golden
    vectors) on the first n_sample elements of the headline plate, values
    only (the GPU line's mode) on one core (the reference's loop is serial)
    and on all host cores, with indices on one core; and the whole
    reference path (KC0 loop, tocsc, slice, spsolve) on smaller plates."""
    ncpu = min(os.cpu_count() or 1, 16)
    v1, dt1, _ = _oracle_kc0(mesh, prop, n_sample, 1, True)
    vn, dtn, used = _oracle_kc0(mesh, prop, n_sample, ncpu, True)
    vi, dti, _ = _oracle_kc0(mesh, prop, n_sample // 2, 1, False)
    log('cpu baseline: KC0 values-only %.0f el/s (1 core), %.0f el/s (%d threads); '
        'with indices %.0f el/s' % (v1, vn, used, vi))
    legs = {'kc0_values_all_cores': dict(value=vn, cores=used, s=dtn),
            'kc0_indices_1_core': dict(value=vi, cores=1, s=dti, elements=n_sample // 2)}
    for n in ladder:
        try:
            legs['reference_path_%dx%d' % (n, n)] = _reference_path(n, log)
        except Exception as exc:  # reported, not hidden
            legs['reference_path_%dx%d' % (n, n)] = dict(error=str(exc))
    legs['reference_path_1000x1000'] = (
        'not run (DNF at 6,012,006 DOF): the host COO alone is 13.8 GB and SuperLU of '
        'the free-DOF system grows past the host memory cap; see the ladder above')
    return dict(value=v1, unit='elements/s', cores=1, kind='port',
                sample='KC0 values-only (overwrite) of the first %d elements of the '
                       '%dx%d laminate plate by the oracle, one core, %.1f s'
                       % (n_sample, mesh.nx - 1, mesh.ny - 1, dt1),
                legs=legs)


def solve_rates(n, repeats, log):
    """Static solves/s of an n x n clamped laminate plate under uniform
    pressure, to ||f - K u|| <= 1e-8 ||f|| (FP64 residual on FP64 K)."""
    import torch
    from pyfe3d_amd import assembly as A
    from pyfe3d_amd.sparse import BSRPlan
    from pyfe3d_amd.solver import StaticSystem
    mesh = build_plate(n)
    prop, _ = laminate_record()
    dev = 'cuda'
    E = mesh.quads.shape[0]
    r = torch.empty(576 * E, dtype=torch.int64, device=dev)
    c = torch.empty_like(r)
    v = torch.zeros(576 * E, dtype=torch.float64, device=dev)
    A.shell_kc0('quad4', torch.as_tensor(mesh.x, device=dev),
                torch.as_tensor(6 * mesh.quads, device=dev), r, c, v,
                torch.as_tensor(prop.record()[None], device=dev))
    plan = BSRPlan(r, c, mesh.n_nodes)
    sysm = StaticSystem(plan, plan.apply(v), mesh.clamped())
    f = mesh.pressure_load(1e3)
    out = {}
    # displacement reference: scipy's sparse LU of the same (Dirichlet-
    # treated) system on the host; the drilling rotations rz are nearly a
    # null space of these plates (sparse and dense LU differ there), so
    # their error is reported apart
    import scipy.sparse.linalg as sla
    uref = sla.spsolve(plan.to_scipy(sysm.val).tocsc(), sysm.rhs(f).cpu().numpy())
    keep = np.ones(mesh.n_dof, bool)
    keep[5::6] = False
    t0 = time.perf_counter()
    sysm.setup_amg(mesh.ncoords, precisions=('fp64', 'fp32'))
    amg_setup = time.perf_counter() - t0
    # the exact solution of the same system to ~1e-15: double-double
    # refinement (its residual far below the FP64 rounding of u)
    hi, lo, dd = sysm.solve_dd(f, method='amg', rtol=1e-13)
    uex = hi.cpu().numpy() + lo.cpu().numpy()
    out['exact_reference'] = dict(relres=dd['relres'], steps=dd['steps'])
    out['spsolve_u_err_vs_exact'] = float(np.linalg.norm(uref - uex) / np.linalg.norm(uex))
    out['spsolve_u_err_rz_vs_exact'] = float(np.linalg.norm((uref - uex)[~keep]) /
                                             np.linalg.norm(uex[~keep]))
    for method in ('direct', 'fp64', 'mixed', 'ir', 'amg', 'amg32'):
        u, info = sysm.solve(f, method=method, rtol=1e-8)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeats):
            # the direct solve is timed with its factorisation
            u, info = sysm.solve(f, method=method, rtol=1e-8, refactor=True)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / repeats
        ug = u.cpu().numpy()
        err = np.linalg.norm((ug - uref)[keep]) / np.linalg.norm(uref[keep])
        err_x = np.linalg.norm(ug - uex) / np.linalg.norm(uex)
        err_rz = np.linalg.norm((ug - uex)[~keep]) / np.linalg.norm(uex[~keep])
        out[method] = dict(ms=dt * 1e3, iters=info['iters'], outer=info['outer'],
                           relres=info['relres'], converged=info['converged'],
                           u_err_vs_spsolve_without_rz=err, u_err_vs_exact=err_x,
                           u_err_rz_vs_exact=err_rz)
        if info['converged']:
            out[method]['solves_per_s'] = 1. / dt
        log('solve %dx%d %s: %.2f ms, %s, u err vs exact %.2e (rz %.1e)'
            % (n, n, method, dt * 1e3, info, err_x, err_rz))
    out['amg']['setup_s'] = amg_setup
    return out


def sweep_dims(n_var, seed=2026):
    """Plate sizes of the design sweep: nx, ny ~ U{8..64} elements (seeded)."""
    rng = np.random.default_rng(seed)
    return [tuple(int(t) for t in rng.integers(8, 65, 2)) for _ in range(n_var)]


def sweep_rates(n_var, log, steps=2, modes=('direct', 'fp64', 'mixed')):
    """Batched design sweep: n_var clamped isotropic plates with smooth
    seeded thickness fields and a central point load, assembled and solved
    to 1e-8 (every variant) per step, as one captured graph per mode:
    banded Cholesky per plate ('direct'), FP64 block-Jacobi PCG ('fp64'),
    PCG reading the operator in FP32 ('mixed').  A step is the whole
    captured graph: property records, KC0, BSR, Dirichlet, block Jacobi,
    FP32 copies and the solve."""
    import torch
    from pyfe3d_amd.sweep import PlateSweep
    t0 = time.perf_counter()
    sw = PlateSweep(sweep_dims(n_var))
    h, f = sw.random_designs(7)
    sw.set_thickness(h)
    sw.set_loads(f)
    log('sweep: %d variants, %d elements, %d DOF, %d blocks (setup %.1f s)'
        % (n_var, sw.n_elem, 6 * sw.n_nodes, sw.plan.nnzb, time.perf_counter() - t0))
    out = dict(variants=n_var, elements=sw.n_elem, dof=6 * sw.n_nodes)
    tuned = None
    if 'direct' in modes:
        # launch configuration of the banded solver chosen by timing
        sw.assemble()
        t0 = time.perf_counter()
        table = sw.autotune_direct(log=log, repeats=1)
        tuned = dict(config=dict(zip(('factor_threads', 'panel_bandwidth', 'solve_threads',
                                      'narrow_band_factor_threads'), sw.direct_config)),
                     source=sw.direct_config_source, key=sw.autotune_key(),
                     tuning_s=time.perf_counter() - t0,
                     candidates=[dict(c, ms=round(ms, 3)) for c, ms in table])
        log('sweep direct: autotuned config %s' % (sw.direct_config,))
    pcg_tuned = {}
    for mode in modes:
        if mode != 'direct':
            # launch shape of the persistent batched PCG, chosen by timing
            sw.assemble()
            t0 = time.perf_counter()
            table = sw.autotune_pcg(mode, log=log)
            groups, threads = sw.pcg_config[mode]
            pcg_tuned[mode] = dict(config=dict(workgroups=groups, threads=threads),
                                   source=sw.pcg_config_source, key=sw.autotune_key_pcg(mode),
                                   tuning_s=time.perf_counter() - t0,
                                   candidates=[dict(c, ms=round(ms, 3)) for c, ms in table])
            log('sweep %s: autotuned PCG launch %s' % (mode, sw.pcg_config[mode]))
        t0 = time.perf_counter()
        g = sw.capture(mode=mode, rtol=1e-8)  # runs one eager step first
        torch.cuda.synchronize()
        first = time.perf_counter() - t0
        sw.check()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ev0.record()
        for _ in range(steps):
            g.replay()
        ev1.record()
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / steps
        rel = sw.relres.cpu().numpy()
        its = sw.iters.cpu().numpy()
        out[mode] = dict(ms_per_sweep=ms, solves_per_s=n_var / (ms * 1e-3),
                         max_relres=float(rel.max()), converged=int((rel <= 1e-8).sum()),
                         pcg_iters_total=int(its.sum()), pcg_iters_max=int(its.max()),
                         first_step_s=first)
        if mode == 'direct' and tuned:
            out[mode]['autotune'] = tuned
        if mode in pcg_tuned:
            out[mode]['autotune'] = pcg_tuned[mode]
        log('sweep %s: %.1f ms/sweep, %.1f solves/s, max relres %.2e, iters %d'
            % (mode, ms, n_var / (ms * 1e-3), rel.max(), its.sum()))
        del g
        sw.graph = None
    if 'direct' in modes:
        out['direct']['factor'] = factor_roofline(sw, log)
    return out


def factor_roofline(sw, log, reps=3):
    """The banded factorisation alone (band storage fill + Cholesky of all
    variants, autotuned shape), timed with HIP events on its stream, against
    the FP64 matrix-core peak (its bound: the band read and written once at
    8 TB/s takes less time than its flops at 78.6 TFLOP/s)."""
    import torch
    st = torch.cuda.current_stream()
    sw.factor_direct(stream=st)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(reps):
        sw.factor_direct(stream=st)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    flop = sw.band_flops()
    band_bytes = 16 * int((sw.band_n_host * (sw.band_bw_host + 1)).sum())
    window_bytes = sw.band_window_bytes()
    tf = flop / (ms * 1e-3) / 1e12
    out = dict(ms=ms, flop=flop, bound='mfma', achieved=tf, peak=FP64_PEAK_TFLOPS,
               unit='TFLOP/s', frac=tf / FP64_PEAK_TFLOPS, band_bytes=band_bytes,
               min_hbm_ms=band_bytes / HBM_PEAK_BPS * 1e3, panel_traffic_bytes=window_bytes,
               panel_traffic_gbps=window_bytes / (ms * 1e-3) / 1e9,
               note='band fill + banded Cholesky (band_zero, band_fill, band_chol); '
                    'panel_traffic = 32-column panels read/written once + trailing '
                    'window lower triangle read/written once per panel')
    log('sweep factor: %.1f ms, %.3g flop -> %.2f TFLOP/s = %.3f of FP64 peak; '
        'panel traffic %.0f GB/s' % (ms, flop, tf, out['frac'], out['panel_traffic_gbps']))
    return out


def modal_rates(n, log, k=6):
    """First k natural frequencies of an n x n simply supported isotropic
    plate (u, v, w fixed on the edges, consistent mass): assembly of KC0 and
    M, BSR, banded Cholesky of K + M and shift-invert Lanczos."""
    import torch
    from pyfe3d_amd import assembly as A
    from pyfe3d_amd.eigen import GeneralizedEigen
    from pyfe3d_amd.mesh import PlateMesh
    from pyfe3d_amd.shellprop import isotropic_plate
    from pyfe3d_amd.sparse import BSRPlan
    m = PlateMesh(n + 1, n + 1, 0.3, 0.5)
    rec = isotropic_plate(thickness=0.01, E=203e9, nu=0.33, rho=7.83e3).record()
    E = m.quads.shape[0]
    dev = 'cuda'
    x = torch.as_tensor(m.x, device=dev)
    conn = torch.as_tensor(6 * m.quads, device=dev)
    props = torch.as_tensor(rec[None], device=dev)
    xx, yy = m.ncoords[:, 0], m.ncoords[:, 1]
    edge = (np.isclose(xx, 0) | np.isclose(xx, 0.3) | np.isclose(yy, 0) | np.isclose(yy, 0.5))
    fixed = np.zeros(m.n_dof, bool)
    for d in range(3):
        fixed[d::6] = edge
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rk = torch.empty(576 * E, dtype=torch.int64, device=dev)
    ck = torch.empty_like(rk)
    vk = torch.zeros(576 * E, dtype=torch.float64, device=dev)
    A.shell_kc0('quad4', x, conn, rk, ck, vk, props)
    rm = torch.empty(480 * E, dtype=torch.int64, device=dev)
    cm = torch.empty_like(rm)
    vm = torch.zeros(480 * E, dtype=torch.float64, device=dev)
    A.shell_m('quad4', x, conn, rm, cm, vm, props, mtype=0)
    pk = BSRPlan(rk, ck, m.n_nodes)
    pm = BSRPlan(rm, cm, m.n_nodes)
    ev = GeneralizedEigen(pk, pk.apply(vk), pm.apply(vm), fixed, sigma=-1.)
    lam, X, info = ev.solve(k=k, tol=1e-10)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out = dict(ms=dt * 1e3, omega=[float(v) for v in np.sqrt(lam)], lanczos_steps=info['m'],
               max_residual=float(np.max(info['residual'])), dof=m.n_dof)
    log('modal %dx%d: %.1f ms, omega %s' % (n, n, dt * 1e3, np.sqrt(lam)[:3]))
    return out


# algorithmic bytes per element of the values-only KC0 launch (values out,
# int64 DOF offsets and coordinates in)
PANEL_KC0_BYTES = {'quad4': 576 * 8 + 4 * 8 + 12 * 8, 'tria3r': 324 * 8 + 3 * 8 + 9 * 8,
                   'beamc': 144 * 8 + 2 * 8 + 6 * 8 + 3 * 8}
PANEL_KG_BYTES = {'quad4': 144 * 8 + 4 * 8 + 12 * 8 + 24 * 8,
                  'tria3r': 81 * 8 + 3 * 8 + 9 * 8 + 18 * 8,
                  'beamc': 144 * 8 + 2 * 8 + 6 * 8 + 3 * 8 + 12 * 8}


def panel_rates(n, log, steps=10, maxiter=200000):
    """Stiffened panel (n x n Quad4 skin, Tria3R cut-out fill, BeamC
    stringers on every 16th node row): KC0 and KG of all three element
    types in one launch each, and the static solve to 1e-8 (plain FP64 PCG
    vs FP32/FP64 refinement)."""
    import torch
    from pyfe3d_amd import assembly as A
    from pyfe3d_amd.panel import StiffenedPanel
    from pyfe3d_amd.solver import StaticSystem
    from pyfe3d_amd.sparse import BSRPlan
    p = StiffenedPanel(n=n)
    stream = torch.cuda.current_stream()
    x = torch.as_tensor(p.mesh.x, device='cuda')
    groups, total = p.groups('KC0')
    counts = {g.etype: g.conn.shape[0] for g in groups}
    r = torch.empty(total, dtype=torch.int64, device='cuda')
    c = torch.empty_like(r)
    v = torch.empty(total, dtype=torch.float64, device='cuda')
    A.mixed_assemble('KC0', groups, x, r, c, v, overwrite=True)
    ms_k, _ = timed(lambda: A.mixed_assemble('KC0', groups, x, None, None, v, values_only=True,
                                             overwrite=True, check=False), steps, 2, stream)
    kbytes = sum(PANEL_KC0_BYTES[t] * e for t, e in counts.items())
    out = dict(elements=counts, dof=p.n_dof, slots=total,
               kc0_mixed=dict(ms=ms_k, elements_per_s=sum(counts.values()) / (ms_k * 1e-3),
                              gbs=kbytes / (ms_k * 1e-3) / 1e9))
    log('panel %dx%d: %s, KC0 one launch %.3f ms (%.0f GB/s)'
        % (n, n, counts, ms_k, kbytes / (ms_k * 1e-3) / 1e9))
    plan = BSRPlan(r, c, p.n_nodes)
    sysm = StaticSystem(plan, plan.apply(v), p.fixed())
    del r, c, v
    f = p.loads(Nxx=1e3, p=1e3)
    sol = {}
    t0 = time.perf_counter()
    sysm.setup_amg(p.mesh.ncoords, precisions=('fp64', 'fp32'))
    amg_setup = time.perf_counter() - t0
    log('panel amg setup %.1f s' % amg_setup)
    for method in ('amg', 'amg32', 'fp64'):
        # multigrid: the first solve also captures the CG iteration's graph
        # and pays first-use costs; the second (same matrix, same load) is
        # the repeated-solve time reported as ms
        first = None
        for rep in range(2 if method != 'fp64' else 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            u, info = sysm.solve(f, method=method, rtol=1e-8, maxiter=maxiter)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            log('panel solve %s: %.1f ms, %s' % (method, dt * 1e3, info))
            if first is None:
                first = dt
        sol[method] = dict(ms=dt * 1e3, first_ms=first * 1e3, iters=info['iters'],
                           outer=info['outer'], relres=info['relres'],
                           converged=info['converged'])
        if info['converged']:
            sol[method]['solves_per_s'] = 1. / dt
    sol['amg']['setup_s'] = amg_setup
    out['solve'] = sol
    gk, totk = p.groups('KG')
    vg = torch.empty(totk, dtype=torch.float64, device='cuda')
    ms_g, _ = timed(lambda: A.mixed_assemble('KG', gk, x, None, None, vg, u=u, values_only=True,
                                             overwrite=True, check=False), steps, 2, stream)
    gbytes = sum(PANEL_KG_BYTES[t] * e for t, e in counts.items())
    out['kg_mixed'] = dict(ms=ms_g, gbs=gbytes / (ms_g * 1e-3) / 1e9)
    log('panel KG(u) one launch %.3f ms' % ms_g)
    out['warm_start'] = ('not applicable: the network is trained on 64x64 isotropic plates '
                         'and has no input for stringers or cut-outs')
    # lowest natural frequencies: M from one mixed-type launch on K's
    # structure, LOBPCG preconditioned by the multigrid V-cycle
    from pyfe3d_amd.lobpcg import LOBPCG
    gm, totm = p.groups('M', 0)
    rm = torch.empty(totm, dtype=torch.int64, device='cuda')
    cm = torch.empty_like(rm)
    vm = torch.empty(totm, dtype=torch.float64, device='cuda')
    A.mixed_assemble('M', gm, x, rm, cm, vm, mtype=0, overwrite=True)
    gk, totk2 = p.groups('KC0')
    rk = torch.empty(totk2, dtype=torch.int64, device='cuda')
    ck = torch.empty_like(rk)
    vk = torch.empty(totk2, dtype=torch.float64, device='cuda')
    A.mixed_assemble('KC0', gk, x, rk, ck, vk, overwrite=True)
    gg, totg = p.groups('KG')
    rg = torch.empty(totg, dtype=torch.int64, device='cuda')
    cg = torch.empty_like(rg)
    vgu = torch.empty(totg, dtype=torch.float64, device='cuda')
    A.mixed_assemble('KG', gg, x, rg, cg, vgu, u=u, overwrite=True)
    # K, M and KG on one BSR structure (the union of their slot patterns)
    union = BSRPlan(torch.cat([rk, rm, rg]), torch.cat([ck, cm, cg]), p.n_nodes)
    zk, zm, zg = torch.zeros_like(vk), torch.zeros_like(vm), torch.zeros_like(vgu)
    Kb = union.apply(torch.cat([vk, zm, zg]))
    Mb = union.apply(torch.cat([zk, vm, zg]))
    KGb = union.apply(torch.cat([zk, zm, vgu]))
    del rm, cm, vm, rk, ck, vk, rg, cg, vgu, zk, zm, zg
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    # the FP32 V-cycle preconditions as well as the FP64 one (same 93
    # iterations) at 0.6x the cost per application
    lam, X, info = LOBPCG(union, Kb, Mb, p.fixed(), sysm.amg32).solve(k=6, tol=1e-7,
                                                                     maxiter=300, log=log)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out['modal_lobpcg'] = dict(ms=dt * 1e3, iters=info['iters'],
                               lam=[float(v) for v in lam],
                               max_residual=float(np.max(info['residual'])),
                               converged=bool(np.max(info['residual']) <= 1e-7))
    log('panel modes (LOBPCG + multigrid): %.0f ms, %d iterations, lambda %s, residual %.1e'
        % (dt * 1e3, info['iters'], lam, np.max(info['residual'])))
    # linear buckling under the static load case: K x = lambda (-KG(u)) x
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lam, X, info = LOBPCG(union, Kb, KGb, p.fixed(), sysm.amg32, problem='buckling').solve(
        k=3, tol=1e-7, maxiter=300, log=log)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out['buckling_lobpcg'] = dict(ms=dt * 1e3, iters=info['iters'],
                                  load_multipliers=[float(v) for v in lam],
                                  max_residual=float(np.max(info['residual'])),
                                  converged=bool(np.max(info['residual']) <= 1e-7))
    log('panel buckling (LOBPCG + multigrid): %.0f ms, %d iterations, lambda %s, residual %.1e'
        % (dt * 1e3, info['iters'], lam, np.max(info['residual'])))
    return out


def large_solve(n, log, plain_maxiter=10000):
    """Static solve of the n x n laminate plate (the headline mesh: KC0 of
    the jittered 8-ply plate, clamped edges, uniform pressure 1e3 Pa) to
    ||f - K u|| <= 1e-8 ||f||, residuals compensated (pf3d_bsr_residual_dd):

    * multigrid CG (smoothed-aggregation V-cycle, FP64 and FP32 V-cycles;
      setup on the GPU, timed apart): an FP64 vector stops near 1e-7 here;
    * the same solution refined as a double-double pair hi + lo
      (StaticSystem.solve_dd): the pair meets 1e-8; the residual of hi, the
      pair rounded to FP64, is reported (the FP64 rounding floor), with
      the distance of the plain FP64 solution from the pair;
    * plain FP64 block-Jacobi PCG, bounded at plain_maxiter iterations."""
    import torch
    from pyfe3d_amd import assembly as A
    from pyfe3d_amd.solver import StaticSystem
    from pyfe3d_amd.sparse import BlockPlan
    mesh = build_plate(n)
    prop, _ = laminate_record()
    E = mesh.quads.shape[0]
    dev = 'cuda'
    x = torch.as_tensor(mesh.x, device=dev)
    conn = torch.as_tensor(6 * mesh.quads, device=dev)
    v = torch.empty(576 * E, dtype=torch.float64, device=dev)
    A.shell_kc0('quad4', x, conn, None, None, v, torch.as_tensor(prop.record()[None], device=dev),
                values_only=True, overwrite=True)
    plan = BlockPlan(conn, mesh.n_nodes)
    sysm = StaticSystem(plan, plan.apply(v), mesh.clamped())
    del v
    torch.cuda.empty_cache()
    f = mesh.pressure_load(1e3)
    t0 = time.perf_counter()
    amg = sysm.setup_amg(mesh.ncoords, log=log, precisions=('fp64', 'fp32'))
    setup = time.perf_counter() - t0
    log('large solve %dx%d: amg setup %.1f s, %d levels' % (n, n, setup, len(amg.levels)))
    out = dict(dof=mesh.n_dof, amg_setup_s=setup, levels=len(amg.levels), rtol=1e-8)
    sols = {}
    for method in ('amg', 'amg32'):
        for rep in range(2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            u, info = sysm.solve(f, method=method, rtol=1e-8, maxiter=4000)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            log('large solve %dx%d %s: %.2f s, %s' % (n, n, method, dt, info))
        sols[method] = u
        out[method] = dict(s=dt, iters=info['iters'], relres=info['relres'],
                           converged=info['converged'], restarts=info['outer'],
                           ms_per_iter=dt * 1e3 / max(info['iters'], 1))
        if 'fp64_iters' in info:
            out[method]['fp64_iters'] = info['fp64_iters']
        if info['converged']:
            out[method]['solves_per_s'] = 1. / dt
    # double-double refinement from the FP64 multigrid solution
    for rep in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        hi, lo, dd = sysm.solve_dd(f, method='amg', rtol=1e-8, u=sols['amg'], maxiter=4000)
        torch.cuda.synchronize()
        dt_dd = time.perf_counter() - t0
    udd = hi + lo
    nrm = float(torch.linalg.vector_norm(udd))
    total = out['amg']['s'] + dt_dd
    out['amg_dd'] = dict(s=total, refine_s=dt_dd, steps=dd['steps'], refine_iters=dd['iters'],
                         relres=dd['relres'], converged=dd['converged'],
                         relres_fp64_rounding=dd['relres_fp64'])
    if dd['converged']:
        out['amg_dd']['solves_per_s'] = 1. / total
    for method in ('amg', 'amg32'):
        out[method]['u_err_vs_double_double'] = float(
            torch.linalg.vector_norm((sols[method] - hi) - lo)) / nrm
    log('large solve %dx%d double-double: %.2f s refine, %s, u err amg %.2e amg32 %.2e'
        % (n, n, dt_dd, dd, out['amg']['u_err_vs_double_double'],
           out['amg32']['u_err_vs_double_double']))
    # plain FP64 block-Jacobi PCG, bounded
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, info = sysm.solve(f, method='fp64', rtol=1e-8, maxiter=plain_maxiter)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out['plain_fp64_pcg'] = dict(s=dt, iters=info['iters'], relres=info['relres'],
                                 converged=info['converged'],
                                 ms_per_iter=dt * 1e3 / max(info['iters'], 1),
                                 note='bounded at %d iterations' % plain_maxiter)
    log('large solve %dx%d plain FP64 PCG: %.1f s, %s' % (n, n, dt, info))
    from pyfe3d_amd.amg import residual_floor
    out['fp64_residual_bound'] = residual_floor(sysm, sols['amg'])
    return out


def profile_kc0(n, launches=3):
    """`launches` values-only KC0 launches on the n x n plate (the bench's
    warm-up and timed steps for the kernel trace; 3 for the counter runs),
    then three copies of 1 GiB (the calibration of the HBM byte counters)."""
    import torch
    from pyfe3d_amd import assembly as A
    mesh = build_plate(n)
    prop, _ = laminate_record()
    E = mesh.quads.shape[0]
    x = torch.as_tensor(mesh.x, device='cuda')
    conn = torch.as_tensor(6 * mesh.quads, device='cuda')
    props = torch.as_tensor(prop.record()[None], device='cuda')
    v = torch.empty(576 * E, dtype=torch.float64, device='cuda')
    for _ in range(launches):
        A.shell_kc0('quad4', x, conn, None, None, v, props, values_only=True, overwrite=True,
                    check=False)
    torch.cuda.synchronize()
    src = torch.ones(1 << 27, dtype=torch.float64, device='cuda')
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print('profile-kc0: %d elements, %d KC0 launches, 3 x 1 GiB copies' % (E, launches),
          flush=True)


def profile_sweep(n_var, steps=3, config=(512, 448, 512, 256)):
    """`steps` direct-solve steps of the n_var-plate design sweep (records,
    KC0, BSR, Dirichlet, band fill, banded Cholesky, triangular solves) with
    the launch shape given (factor threads, panel bandwidth, solve threads,
    factor threads of the plates whose band fits the 256-wide panel; default
    the one the autotuner picks for the 4,096-plate sweep), for kernel
    traces and counter runs."""
    import torch
    from pyfe3d_amd.sweep import PlateSweep
    sw = PlateSweep(sweep_dims(n_var))
    h, f = sw.random_designs(7)
    sw.set_thickness(h)
    sw.set_loads(f)
    for _ in range(steps):
        sw.step(mode='direct', rtol=1e-8, config=config)
    torch.cuda.synchronize()
    sw.check()
    rel = float(sw.relres.max())
    print('profile-sweep: %d plates, %d steps, max relres %.2e' % (n_var, steps, rel), flush=True)


def profile_nn(steps, batch=64):
    """`steps` training steps (forward, nodal-loss backward, Adam) of the
    warm-start network on a batch of random 64 x 64 inputs, for kernel
    traces and counter runs; prints the GEMM flops per step and the step
    time from HIP events."""
    import torch
    from pyfe3d_amd import warmstart as W
    net = W.WarmStartNet(n=64, batch=batch, seed=3)
    g = torch.Generator(device='cuda').manual_seed(5)
    x = torch.randn((batch, 64, 64, 2), device='cuda', generator=g)
    target = torch.randn((batch, 65, 65, 6), device='cuda', generator=g)
    wc = torch.ones(6, device='cuda')
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    net.forward(x, st)
    net.backward(target, wc, st)
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(steps):
        net.forward(x, st)
        net.backward(target, wc, st)
        net.adam(stream=st)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    fwd, bwd = net.gemm_flops()
    print('profile-nn: batch %d, %d steps, %.3f ms per step, GEMM flops per step %d '
          '(forward %d, backward %d)' % (batch, steps, ms, fwd + bwd, fwd, bwd), flush=True)


def warm_start_rates(n_plates, log, multigrid_plates=16):
    """Held-out 64 x 64 plates (smooth thickness fields, Gaussian pressure
    patches): block-Jacobi PCG to 1e-8 per plate and multigrid CG on a
    subset, from zero and from the trained network's Galerkin-scaled guess,
    with the network's own held-out evaluation (2,000 plates, recorded when
    it was trained) alongside."""
    from pyfe3d_amd import warmstart as W
    path = os.path.join(ROOT, 'pyfe3d_amd', 'warmstart_weights.npz')
    net, sc, meta = W.load(path)
    res = W.evaluate(net, sc, n_plates, seed=12345, chunk=n_plates,
                     multigrid_plates=multigrid_plates, max_iter=60000, log=log)
    res['training'] = {k: meta.get(k) for k in ('samples', 'epochs', 'loss', 'train_s',
                                                'eval_seed', 'eval', 'eval_multigrid')}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    # the GPU reaches its sustained clocks after ~25 ms of back-to-back
    # launches: the default warm-up (50 x 0.8 ms) covers that ramp
    ap.add_argument('--steps', type=int, default=50)
    ap.add_argument('--warmup', type=int, default=50)
    ap.add_argument('--n', type=int, default=1000, help='elements per side')
    ap.add_argument('--cpu-sample', type=int, default=200000)
    ap.add_argument('--no-cpu-ladder', action='store_true',
                    help='skip the reference CPU path (tocsc + spsolve) on smaller plates')
    ap.add_argument('--no-solve', action='store_true')
    ap.add_argument('--sweep-variants', type=int, default=4096,
                    help='plates of the design sweep solved with the banded direct solver')
    ap.add_argument('--sweep-pcg-variants', type=int, default=256,
                    help='plates of the same sweep solved with PCG (FP64 and FP32 operator)')
    ap.add_argument('--warm-start-plates', type=int, default=64)
    ap.add_argument('--panel', type=int, default=512,
                    help='stiffened panel skin elements per side (0: skip)')
    ap.add_argument('--large-solve-n', type=int, default=1000,
                    help='plate size of the multigrid static solve in the default run (0: skip)')
    ap.add_argument('--large-solve', action='store_true',
                    help='FP64 PCG static solve of the n x n headline plate (long)')
    ap.add_argument('--panel-only', action='store_true',
                    help='run only the stiffened-panel lines (no headline)')
    ap.add_argument('--profile-kc0', action='store_true',
                    help='only: 3 KC0 launches and a 1 GiB device copy (counter runs)')
    ap.add_argument('--profile-sweep', type=int, default=0,
                    help='only: direct-solve steps of a design sweep of this many plates')
    ap.add_argument('--profile-launches', type=int, default=3,
                    help='KC0 launches of --profile-kc0 (the kernel trace uses warmup + steps)')
    ap.add_argument('--profile-nn', type=int, default=0,
                    help='run only this many forward + backward + Adam steps of the warm-start '
                         'network on a batch of 64 plates of 64 x 64 elements (for rocprof)')
    ap.add_argument('--quiet', action='store_true')
    args = ap.parse_args()
    log = (lambda *a: None) if args.quiet else (lambda *a: print(*a, file=sys.stderr, flush=True))
    if args.gpus != 1:
        print('bench.py measures one GPU: this build has no multi-GPU path '
              '(run with --gpus 1)', file=sys.stderr)
        sys.exit(2)

    import torch
    from pyfe3d_amd import assembly as A
    dev = 'cuda'
    stream = torch.cuda.current_stream()
    if args.profile_kc0:
        profile_kc0(args.n, args.profile_launches)
        return
    if args.profile_sweep:
        profile_sweep(args.profile_sweep)
        return
    if args.profile_nn:
        profile_nn(args.profile_nn)
        return
    if args.large_solve:
        print(json.dumps({'large_solve': large_solve(args.n, log)}), flush=True)
        return
    if args.panel_only:
        print(json.dumps({'stiffened_panel': panel_rates(args.panel or 512, log)}), flush=True)
        return

    t0 = time.perf_counter()
    mesh = build_plate(args.n)
    prop, stack = laminate_record()
    E = mesh.quads.shape[0]
    log('mesh %d elements, %d nodes (%.1f s)' % (E, mesh.n_nodes, time.perf_counter() - t0))
    x = torch.as_tensor(mesh.x, device=dev)
    conn = torch.as_tensor(6 * mesh.quads, device=dev)
    props = torch.as_tensor(prop.record()[None], device=dev)
    v = torch.empty(576 * E, dtype=torch.float64, device=dev)

    # validate the arrays once; the timed loops launch without host checks
    A.shell_kc0('quad4', x, conn, None, None, v, props, values_only=True, overwrite=True)

    def kc0_values():
        A.shell_kc0('quad4', x, conn, None, None, v, props, values_only=True, overwrite=True,
                    check=False)

    ms, wall = timed(kc0_values, args.steps, args.warmup, stream)
    log('KC0 values-only: %.3f ms/step (wall %.3f)' % (ms, wall))
    achieved = KC0_BYTES_VALUES * E / (ms * 1e-3) / 1e9

    extra = {}
    # measured HBM ceilings on the same 4.6 GB buffer: a write-only fill
    # (torch's vectorised fill kernel) and a device copy of half into half
    ms_fill, _ = timed(lambda: v.fill_(0.), 5, 1, stream)
    half = v.numel() // 2
    ms_copy, _ = timed(lambda: v[:half].copy_(v[half:2 * half]), 5, 1, stream)
    ms_sum, _ = timed(lambda: v.sum(), 5, 1, stream)
    extra['hbm_measured'] = dict(fill_write_gbs=v.numel() * 8 / (ms_fill * 1e-3) / 1e9,
                                 copy_read_write_gbs=2 * half * 8 / (ms_copy * 1e-3) / 1e9,
                                 sum_read_gbs=v.numel() * 8 / (ms_sum * 1e-3) / 1e9,
                                 bytes=v.numel() * 8)
    log('HBM: fill %.0f GB/s, copy %.0f GB/s, sum %.0f GB/s'
        % tuple(extra['hbm_measured'][k] for k in ('fill_write_gbs', 'copy_read_write_gbs',
                                                    'sum_read_gbs')))
    r = torch.empty(576 * E, dtype=torch.int64, device=dev)
    c = torch.empty_like(r)

    def kc0_indices():
        A.shell_kc0('quad4', x, conn, r, c, v, props, overwrite=True, check=False)

    ms_i, _ = timed(kc0_indices, max(3, args.steps // 4), 1, stream)
    extra['kc0_with_indices'] = dict(ms=ms_i, elements_per_s=E / (ms_i * 1e-3),
                                     gbs=KC0_BYTES_INDICES * E / (ms_i * 1e-3) / 1e9)
    log('KC0 with indices: %.3f ms' % ms_i)

    def kc0_accumulate():
        A.shell_kc0('quad4', x, conn, None, None, v, props, values_only=True, check=False)

    ms_a, _ = timed(kc0_accumulate, max(3, args.steps // 4), 1, stream)
    extra['kc0_accumulate'] = dict(ms=ms_a, elements_per_s=E / (ms_a * 1e-3),
                                   gbs=(KC0_BYTES_VALUES + 576 * 8) * E / (ms_a * 1e-3) / 1e9)
    del r, c
    vg = torch.empty(144 * E, dtype=torch.float64, device=dev)
    stress = torch.tensor([-1e3, 0., 0.], dtype=torch.float64, device=dev)
    ms_g, _ = timed(lambda: A.shell_kg_stress('quad4', x, conn, stress, None, None, vg,
                                              values_only=True, overwrite=True, check=False),
                    max(3, args.steps // 4), 1, stream)
    extra['kg_given_stress'] = dict(ms=ms_g, elements_per_s=E / (ms_g * 1e-3),
                                    gbs=(144 * 8 + 128) * E / (ms_g * 1e-3) / 1e9)
    del vg
    vm = torch.empty(480 * E, dtype=torch.float64, device=dev)
    ms_m, _ = timed(lambda: A.shell_m('quad4', x, conn, None, None, vm, props, mtype=0,
                                      values_only=True, overwrite=True, check=False),
                    max(3, args.steps // 4), 1, stream)
    extra['m_consistent'] = dict(ms=ms_m, elements_per_s=E / (ms_m * 1e-3),
                                 gbs=(480 * 8 + 128) * E / (ms_m * 1e-3) / 1e9)
    del vm
    log('KG %.3f ms, M %.3f ms' % (ms_g, ms_m))

    # per-element laminate records on the device (one 8-ply stack per element
    # with seeded angles: the property tables of a laminate design sweep)
    from pyfe3d_amd.shellprop import laminate_records_device, material_row
    g = torch.Generator(device=dev).manual_seed(5)
    angles = torch.tensor([0., 45., -45., 90.], dtype=torch.float64, device=dev)[
        torch.randint(0, 4, (E, 8), device=dev, generator=g)]
    plyt = torch.full((E, 8), 0.000125, dtype=torch.float64, device=dev)
    mats = [material_row((140e9, 10e9, 0.3, 5e9, 5e9, 5e9), 1600.)]
    recs = torch.empty((E, 32), dtype=torch.float64, device=dev)
    ms_r, _ = timed(lambda: laminate_records_device(angles, plyt, mats, out=recs),
                    max(3, args.steps // 4), 1, stream)
    extra['laminate_records'] = dict(ms=ms_r, elements_per_s=E / (ms_r * 1e-3), plies=8,
                                     gbs=(8 * 16 + 256) * E / (ms_r * 1e-3) / 1e9)
    log('laminate records (8 plies) %.3f ms' % ms_r)
    del angles, plyt, recs

    # internal forces (update_fint) of the same plate: element forces + ordered node gather
    u = torch.randn(mesh.n_dof, dtype=torch.float64, device=dev,
                    generator=torch.Generator(device=dev).manual_seed(3)) * 1e-4
    fint = torch.zeros(mesh.n_dof, dtype=torch.float64, device=dev)
    fe = torch.empty((E, 4, 6), dtype=torch.float64, device=dev)
    gather = A.NodeGather(conn, mesh.n_nodes)
    ms_f, _ = timed(lambda: A.update_fint('quad4', x, conn, u, fint, props, gather=gather,
                                          fe=fe, check=False),
                    max(3, args.steps // 4), 1, stream)
    extra['fint'] = dict(ms=ms_f, elements_per_s=E / (ms_f * 1e-3),
                         gbs=(24 * 8 + 4 * 8 + 12 * 8 + 24 * 8 * 2 + 24 * 8) * E / (ms_f * 1e-3) / 1e9)
    del fe, gather, fint
    log('fint %.3f ms' % ms_f)

    if not args.no_solve:
        try:
            extra['solve_64x64'] = solve_rates(64, 3, log)
        except Exception as exc:  # reported, not hidden
            extra['solve_64x64'] = dict(error=str(exc))
            log('solve failed: %s' % exc)

    if not args.no_solve and args.sweep_variants > 0:
        del v
        torch.cuda.empty_cache()
        try:
            extra['design_sweep'] = sweep_rates(args.sweep_variants, log, modes=('direct',))
        except Exception as exc:
            extra['design_sweep'] = dict(error=str(exc))
            log('sweep failed: %s' % exc)
        torch.cuda.empty_cache()
        if args.sweep_pcg_variants > 0:
            try:
                extra['design_sweep_pcg'] = sweep_rates(args.sweep_pcg_variants, log, steps=1,
                                                        modes=('direct', 'fp64', 'mixed'))
            except Exception as exc:
                extra['design_sweep_pcg'] = dict(error=str(exc))
                log('sweep (pcg) failed: %s' % exc)

    if not args.no_solve:
        try:
            extra['modal_64x64'] = modal_rates(64, log)
        except Exception as exc:
            extra['modal_64x64'] = dict(error=str(exc))
            log('modal failed: %s' % exc)

    if args.panel > 0:
        torch.cuda.empty_cache()
        try:
            extra['stiffened_panel'] = panel_rates(args.panel, log)
        except Exception as exc:
            extra['stiffened_panel'] = dict(error=str(exc))
            log('panel failed: %s' % exc)

    if not args.no_solve and args.large_solve_n > 0:
        torch.cuda.empty_cache()
        try:
            extra['large_solve_%dx%d' % (args.large_solve_n, args.large_solve_n)] = \
                large_solve(args.large_solve_n, log)
        except Exception as exc:
            extra['large_solve'] = dict(error=str(exc))
            log('large solve failed: %s' % exc)

    if not args.no_solve and args.warm_start_plates > 0:
        torch.cuda.empty_cache()
        try:
            extra['warm_start_64x64'] = warm_start_rates(args.warm_start_plates, log)
        except Exception as exc:
            extra['warm_start_64x64'] = dict(error=str(exc))
            log('warm start failed: %s' % exc)

    n_sample = min(args.cpu_sample, E)
    base = cpu_baseline(mesh, prop, n_sample, log,
                        ladder=() if args.no_cpu_ladder else (20, 64, 128))
    log('cpu baseline %.0f elements/s' % base['value'])

    line = {
        'metric': 'KC0 elements/s (%HBM roofline); solves/s at 1e-8 res, learned+mixed vs plain',
        'value': E / (ms * 1e-3),
        'unit': 'elements/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': ms,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic (seeded jittered plate and laminate, no external data)',
        'config': {'workload': 'Quad4 KC0 assembly, values only (overwrite), %dx%d jittered '
                               'plate, 8-ply laminate %s' % (args.n, args.n, stack),
                   'elements': E, 'nodes': mesh.n_nodes, 'dof': mesh.n_dof,
                   'slots': 576 * E},
        'roofline': {'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
                     'traffic': KC0_TRAFFIC_BYTES if args.n == 1000 else None,
                     'algorithmic_bytes': KC0_BYTES_VALUES * E,
                     'bytes_per_element': KC0_BYTES_VALUES,
                     'measured_fill_gbs': extra['hbm_measured']['fill_write_gbs'],
                     'frac_of_measured_fill':
                         achieved / extra['hbm_measured']['fill_write_gbs']},
        'cpu_baseline': base,
        'solves': solve_summary(extra),
        'extra': extra,
    }
    print(json.dumps(line), flush=True)


def solve_summary(extra):
    """Solves/s to 1e-8 (compensated FP64 residual on FP64 K) per workload,
    with the three effects apart: the preconditioner (plain FP64 block-
    Jacobi PCG against FP64 multigrid CG), the precision (FP64 against FP32
    V-cycle inside the same FP64 CG) and the learned warm start (zero start
    against the network's guess in the same solver).  Solves that miss the
    tolerance carry no rate."""
    out = {}

    def rate(d, key='ms'):
        return None if not d.get('converged', True) else 1e3 / d[key]

    s64 = extra.get('solve_64x64', {})
    if 'fp64' in s64 and 'amg' in s64 and 'amg32' in s64:
        out['plate_64x64'] = dict(
            plain_fp64_pcg_ms=s64['fp64']['ms'], multigrid_fp64_ms=s64['amg']['ms'],
            multigrid_fp32_vcycle_ms=s64['amg32']['ms'], direct_ms=s64['direct']['ms'],
            preconditioner_speedup=s64['fp64']['ms'] / s64['amg']['ms'],
            precision_speedup=s64['amg']['ms'] / s64['amg32']['ms'],
            mixed_multigrid_vs_plain=s64['fp64']['ms'] / s64['amg32']['ms'],
            all_converged=all(v.get('converged', False) for v in s64.values()
                              if isinstance(v, dict) and 'ms' in v))
    sp = extra.get('stiffened_panel', {}).get('solve', {})
    if 'fp64' in sp and 'amg32' in sp:
        out['stiffened_panel_512'] = dict(
            plain_fp64_pcg_ms=sp['fp64']['ms'], multigrid_fp64_ms=sp['amg']['ms'],
            multigrid_fp32_vcycle_ms=sp['amg32']['ms'],
            preconditioner_speedup=sp['fp64']['ms'] / sp['amg']['ms'],
            precision_speedup=sp['amg']['ms'] / sp['amg32']['ms'],
            all_converged=all(sp[k].get('converged', False) for k in ('fp64', 'amg', 'amg32')))
    ls = [v for k, v in extra.items() if k.startswith('large_solve_') and isinstance(v, dict)]
    if ls and 'amg' in ls[0]:
        L = ls[0]
        out['plate_1000x1000'] = dict(
            multigrid_fp64=dict(ms=L['amg']['s'] * 1e3, relres=L['amg']['relres'],
                                converged=L['amg']['converged']),
            multigrid_fp32_vcycle=dict(ms=L['amg32']['s'] * 1e3, relres=L['amg32']['relres'],
                                       converged=L['amg32']['converged']),
            multigrid_double_double=dict(ms=L['amg_dd']['s'] * 1e3,
                                         relres=L['amg_dd']['relres'],
                                         converged=L['amg_dd']['converged'],
                                         solves_per_s=L['amg_dd'].get('solves_per_s')),
            fp64_rounding_of_solution_relres=L['amg_dd']['relres_fp64_rounding'],
            u_err_fp64_multigrid_vs_double_double=L['amg']['u_err_vs_double_double'],
            plain_fp64_pcg=dict(ms=L['plain_fp64_pcg']['s'] * 1e3,
                                iters=L['plain_fp64_pcg']['iters'],
                                relres=L['plain_fp64_pcg']['relres'],
                                converged=L['plain_fp64_pcg']['converged']),
            multigrid_setup_s=L['amg_setup_s'])
    sw = extra.get('design_sweep', {}).get('direct')
    if sw:
        out['design_sweep_4096'] = dict(solves_per_s=sw['solves_per_s'],
                                        all_converged=sw['converged'] == sw.get('variants', 4096),
                                        method='batched banded Cholesky, autotuned launch shape')
    ws = extra.get('warm_start_64x64', {})
    if 'summary' in ws:
        out['learned_warm_start_64x64'] = ws['summary']
    return out


if __name__ == '__main__':
    main()

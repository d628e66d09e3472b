"""The C-ABI library loads and exports every entry point that
include/pf3d.h declares (no compute calls; runs without a GPU)."""
import ctypes
import os
import re

from pyfe3d_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, 'include')):
        if fn.endswith('.h'):
            text = open(os.path.join(ROOT, 'include', fn)).read()
            text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
            names |= set(re.findall(r'\b(pf3d_\w+)\s*\(', text))
    return names


def test_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared()
    assert len(names) >= 8
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_bound_signatures_are_declared():
    assert set(_lib.SIGNATURES) <= declared()


def test_host_only_queries():
    lib = _lib.load()
    assert lib.pf3d_version() > 0
    assert lib.pf3d_slot_count(0, 0, 0) == 576
    assert lib.pf3d_slot_count(2, 0, 0) == 324
    assert lib.pf3d_slot_count(3, 0, 0) == 144
    assert lib.pf3d_slot_count(0, 1, 0) == 144
    assert lib.pf3d_slot_count(0, 3, 2) == 288
    assert lib.pf3d_slot_count(2, 3, 0) == 270
    assert lib.pf3d_slot_count(4, 0, 0) == 72    # Truss KC0
    assert lib.pf3d_slot_count(4, 3, 0) == 144   # Truss consistent mass
    assert lib.pf3d_slot_count(4, 3, 1) == 36    # Truss lumped mass
    assert lib.pf3d_slot_count(5, 0, 0) == 72    # Spring KC0
    assert lib.pf3d_slot_count(5, 3, 0) < 0      # Spring has no mass
    assert lib.pf3d_slot_count(6, 0, 0) == 144   # BeamLR KC0
    assert lib.pf3d_slot_count(6, 1, 0) == 36    # BeamLR KG (rotations)
    assert lib.pf3d_slot_count(6, 3, 1) == 36    # BeamLR lumped mass
    assert lib.pf3d_slot_count(0, 4, 0) == 144   # Quad4 KA_beta
    assert lib.pf3d_slot_count(1, 6, 0) == 144   # Quad4R CA
    assert lib.pf3d_slot_count(2, 5, 0) < 0      # no piston theory for Tria3R
    assert lib.pf3d_slot_count(9, 0, 0) < 0
    assert b'etype' in lib.pf3d_last_error()


def test_argument_errors_need_no_device():
    lib = _lib.load()
    # negative size and null pointers are rejected before any HIP call
    st = lib.pf3d_shell_kc0(0, -1, None, None, None, None, None, None, None, None, None,
                            None, None, None, 0, None)
    assert st == -1
    st = lib.pf3d_shell_kc0(0, 10, None, None, None, None, None, None, None, None, None,
                            None, None, None, 0, None)
    assert st == -1 and b'null' in lib.pf3d_last_error()


def test_argument_errors_of_newer_entry_points():
    """Mixed-type, rod, beam, piston and solver entry points reject bad
    arguments before any device work (no GPU needed)."""
    import ctypes
    from pyfe3d_amd.assembly import _CGroup
    lib = _lib.load()
    call = lambda name, *a: getattr(lib, name)(*a)
    fake = ctypes.c_void_p(16)  # never dereferenced: the checks fail first
    # mixed: bad matrix, too many groups, repeated element type, BeamC without vxy
    g = (_CGroup * 2)()
    g[0] = _CGroup(0, 4, fake, 0, None, None, None, fake, None, None, None)
    g[1] = _CGroup(0, 4, fake, 0, None, None, None, fake, None, None, None)
    gp = ctypes.cast(g, ctypes.c_void_p)
    assert call('pf3d_mixed_assemble', 2, 0, 1, gp, fake, None, fake, fake, fake, 0, None) == -1
    assert b'KC0, KG or M' in lib.pf3d_last_error()
    assert call('pf3d_mixed_assemble', 0, 0, 5, gp, fake, None, fake, fake, fake, 0, None) == -1
    assert call('pf3d_mixed_assemble', 0, 0, 2, gp, fake, None, fake, fake, fake, 0, None) == -1
    assert b'repeated' in lib.pf3d_last_error()
    g[1] = _CGroup(3, 4, fake, 0, None, None, None, fake, None, None, None)
    assert call('pf3d_mixed_assemble', 0, 0, 2, gp, fake, None, fake, fake, fake, 0, None) == -1
    assert b'vxy' in lib.pf3d_last_error()
    assert call('pf3d_mixed_assemble', 1, 0, 1, gp, fake, None, fake, fake, fake, 0, None) == -1
    assert b'null u' in lib.pf3d_last_error()
    # piston theory is Quad4/Quad4R only
    assert call('pf3d_shell_piston', 2, 4, 4, fake, fake, None, fake, fake, fake, 0, None) == -1
    assert call('pf3d_shell_piston', 0, 3, 4, fake, fake, None, fake, fake, fake, 0, None) == -1
    # rods: springs need axes, Truss mass type range
    assert call('pf3d_spring_kc0', 4, fake, fake, None, None, None, fake, fake, fake, fake, 0,
                None) == -1
    assert call('pf3d_truss_m', 4, fake, fake, None, None, fake, 2, fake, fake, fake, 0,
                None) == -1
    assert call('pf3d_rod_fe', 3, 4, fake, fake, None, None, fake, fake, fake, None) == -1
    # BeamLR mass type range
    assert call('pf3d_beamlr_m', 4, fake, fake, None, fake, None, fake, 2, fake, fake, fake, 0,
                None) == -1
    # banded solver launch configuration
    assert call('pf3d_band_factor', 1, 1, None, fake, fake, fake, 10, fake, fake, fake, fake,
                fake, fake, fake, 128, 0, None) == -1
    assert b'threads' in lib.pf3d_last_error()
    assert call('pf3d_band_factor', 1, 1, None, fake, fake, fake, 10, fake, fake, fake, fake,
                fake, fake, fake, 1024, 300, None) == -1
    assert call('pf3d_band_solve', 1, None, fake, fake, fake, fake, fake, fake, fake, fake, 100,
                None) == -1
    assert call('pf3d_band_chol', 1, None, fake, fake, fake, fake, fake, 512, 300, None, None,
                None, None) == -1
    assert b'max_b' in lib.pf3d_last_error()
    assert call('pf3d_band_chol', 1, None, fake, fake, fake, fake, None, 512, 448, None, None,
                None, None) == -1
    assert call('pf3d_band_chol', 1, None, fake, fake, fake, fake, fake, 512, 448, None, None,
                fake, None) == -1
    assert b'forward substitution' in lib.pf3d_last_error()
    assert call('pf3d_band_backsolve', 1, None, fake, fake, fake, fake, fake, fake, fake, 128,
                None) == -1
    assert call('pf3d_band_load', 4, fake, fake, fake, -1, fake, fake, fake, fake, fake, fake,
                None) == -1
    # mixed-precision PCG needs every operand and a positive replacement period
    assert call('pf3d_pcg_mixed', 4, fake, fake, fake, fake, fake, fake, fake, fake, 1e-8, 10, 0,
                0, fake, 1 << 30, fake, fake, None) == -1
    # multigrid setup: sizes, null operands, in-place symmetrisation
    nb = ctypes.c_int64()
    assert call('pf3d_bsr_spgemm_bytes', 0, ctypes.byref(nb)) == -1
    assert call('pf3d_bsr_spgemm_symbolic', 4, fake, None, fake, fake, fake, fake, 1 << 20,
                ctypes.byref(nb), None) == -1
    assert call('pf3d_bsr_spgemm_numeric', 4, fake, fake, fake, fake, fake, fake, fake, None,
                None, 5, None) == -1
    assert call('pf3d_bsr_transpose_bytes', 4, 4, 0, ctypes.byref(nb)) == -1
    assert call('pf3d_bsr_transpose', 4, 4, 8, fake, fake, fake, fake, fake, fake, None, 0,
                None) == -1
    assert call('pf3d_amg_smooth_prolongator', 4, fake, fake, fake, None, fake, fake, 0.5, fake,
                fake, None) == -1
    assert call('pf3d_bsr_symmetrize', 4, 8, fake, fake, fake, None, fake, fake, None) == -1
    assert b'bad arguments' in lib.pf3d_last_error()
    # device-scalar CG updates
    assert call('pf3d_cg_update_xr', 0, fake, fake, fake, fake, fake, None) == -1
    assert call('pf3d_cg_beta', None, 0, None) == -1
    assert call('pf3d_cg_update_p', 4, fake, None, fake, None) == -1
    # laminate records: sizes and null operands
    assert call('pf3d_laminate_records', 4, 0, fake, fake, None, 1, fake, None, 1, fake,
                None) == -1
    assert call('pf3d_laminate_records', 4, 2, fake, None, None, 1, fake, None, 1, fake,
                None) == -1
    # validation pass
    assert call('pf3d_nonfinite_elements', 4, 0, None, fake, fake, None) == -1
    assert call('pf3d_nonfinite_elements', 4, 576, None, None, fake, None) == -1
    # fused multigrid step: aliasing and missing operands are refused
    assert call('pf3d_bsr_fused', 4, fake, fake, fake, fake, None, fake, None, None, None, None,
                1., -1., 0., 0., None) == -1      # xin == r
    assert call('pf3d_bsr_fused', 4, fake, fake, fake, None, None, fake, fake, None, None, None,
                1., -1., 1., 0., None) == -1      # dinv without dnew
    assert call('pf3d_bsr_fused_f32', 4, fake, fake, fake, None, None, None, None, None, None,
                None, 1., 1., 0., 0., None) == -1  # no r
    # multi-vector fused step: vector count and strides are checked
    assert call('pf3d_bsr_fused_mv', 4, fake, fake, fake, 9, 24, fake, 24, None, fake, None,
                None, None, None, 1., -1., 0., 0., None) == -1    # nv > 8
    assert call('pf3d_bsr_fused_mv_f32', 4, fake, fake, fake, 2, 24, None, 12, None, fake,
                None, None, None, None, 1., 1., 0., 0., None) == -1  # ld < 6 nb
    assert call('pf3d_bsr_product_dd_mv', 4, fake, fake, fake, 0, 24, fake, fake, None) == -1
    assert call('pf3d_bsr_product_dd_mv', 4, fake, fake, fake, 2, 24, fake, fake, None) == -1


def test_static_workspace_pcg_limits_need_no_device():
    """pf3d_pcg_static rejects null operands and systems beyond its fixed
    workspace before any device work."""
    import ctypes
    lib = _lib.load()
    fake = ctypes.c_void_p(16)  # never dereferenced: the checks fail first
    it, rel = ctypes.c_int64(), ctypes.c_double()
    args = (fake,) * 6 + (1e-8, 10, 0, ctypes.byref(it), ctypes.byref(rel), None)
    assert lib.pf3d_pcg_static(0, *args) == -1
    assert lib.pf3d_pcg_static(131073, *args) == -1
    assert b'static workspace' in lib.pf3d_last_error()
    assert lib.pf3d_pcg_static(10, None, *args[1:]) == -1

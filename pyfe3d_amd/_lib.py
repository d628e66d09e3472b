"""ctypes binding of ``libpf3d.so`` (the C ABI in ``include/pf3d.h``).

The library is built in-tree by ``pyfe3d_amd.build``.  There is no CPU
fallback: if the library is missing, every compute call raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libpf3d.so')

P = ctypes.c_void_p
I32 = ctypes.c_int
I64 = ctypes.c_int64
F64 = ctypes.c_double
F32 = ctypes.c_float

# name -> argtypes (all return int status except where noted)
SIGNATURES = {
    'pf3d_version': [],
    'pf3d_slot_count': [I32, I32, I32],
    'pf3d_geometry': [I32, I64, P, P, P, P, P, P],
    'pf3d_shell_kc0': [I32, I64, P, P, P, P, P, P, P, P, P, P, P, P, I32, P],
    'pf3d_shell_kg': [I32, I64, P, P, P, P, P, P, P, P, P, P, I32, P],
    'pf3d_shell_kg_stress': [I32, I64, P, P, P, P, P, I32, P, P, P, I32, P],
    'pf3d_shell_m': [I32, I64, P, P, P, P, P, P, I32, P, P, P, I32, P],
    'pf3d_beam_kc0': [I64, P, P, P, P, P, P, P, P, P, I32, P],
    'pf3d_beam_kg': [I64, P, P, P, P, P, P, P, P, P, P, I32, P],
    'pf3d_beam_m': [I64, P, P, P, P, P, P, I32, P, P, P, I32, P],
    'pf3d_isotropic_records': [I64, P, P, P, P],
    'pf3d_nonfinite_elements': [I64, I32, P, P, P, P],
    'pf3d_laminate_records': [I64, I32, P, P, P, I32, P, P, I32, P, P],
    'pf3d_lp_records': [I64, P, P, P, I32, P, P, P],
    'pf3d_shell_fe': [I32, I64, P, P, P, P, P, P, P, P, P, P, P],
    'pf3d_beam_fe': [I64, P, P, P, P, P, P, P, P],
    'pf3d_shell_piston': [I32, I32, I64, P, P, P, P, P, P, I32, P],
    'pf3d_beamlr_kc0': [I64, P, P, P, P, P, P, P, P, P, I32, P],
    'pf3d_beamlr_kg': [I64, P, P, P, P, P, P, P, P, P, P, I32, P],
    'pf3d_beamlr_m': [I64, P, P, P, P, P, P, I32, P, P, P, I32, P],
    'pf3d_beamlr_fe': [I64, P, P, P, P, P, P, P, P],
    'pf3d_truss_kc0': [I64, P, P, P, P, P, P, P, P, I32, P],
    'pf3d_truss_m': [I64, P, P, P, P, P, I32, P, P, P, I32, P],
    'pf3d_spring_kc0': [I64, P, P, P, P, P, P, P, P, P, I32, P],
    'pf3d_rod_fe': [I32, I64, P, P, P, P, P, P, P, P],
    'pf3d_mixed_assemble': [I32, I32, I32, P, P, P, P, P, P, I32, P],
    'pf3d_incidence_bytes': [I64, I64, P],
    'pf3d_incidence_plan': [I64, P, I64, P, I64, P, P, P],
    'pf3d_gather_fint': [I64, P, P, P, P, P],
    'pf3d_bsr_plan_bytes': [I64, I64, P],
    'pf3d_bsr_plan_sort': [I64, P, P, I64, P, I64, P, P],
    'pf3d_bsr_plan_emit': [I64, I64, P, I64, I64, I64, P, P, P, P, P, P],
    'pf3d_bsr_apply': [I64, P, P, P, P, I64, P, P],
    'pf3d_block_plan_bytes': [I64, I32, I64, P],
    'pf3d_block_plan_sort': [I64, I32, P, P, I64, P, I64, P, P],
    'pf3d_block_plan_emit': [I64, I32, P, I64, P, I64, I64, P, P, P, P, P],
    'pf3d_block_apply': [I64, P, P, I32, P, P, I32, P],
    'pf3d_bsr_spmv': [I64, P, P, P, P, P, P],
    'pf3d_bsr_spmv_f32': [I64, P, P, P, P, P, P],
    'pf3d_bsr_residual_dd': [I64, P, P, P, P, P, P, P],
    'pf3d_bsr_product_dd_mv': [I64, P, P, P, I32, I64, P, P, P],
    'pf3d_bsr_fused': [I64, P, P, P, P, P, P, P, P, P, P, F64, F64, F64, F64, P],
    'pf3d_bsr_fused_f32': [I64, P, P, P, P, P, P, P, P, P, P, F64, F64, F64, F64, P],
    'pf3d_bsr_fused_mv': [I64, P, P, P, I32, I64, P, I64, P, P, P, P, P, P, F64, F64, F64, F64,
                          P],
    'pf3d_bsr_fused_mv_f32': [I64, P, P, P, I32, I64, P, I64, P, P, P, P, P, P, F64, F64, F64,
                              F64, P],
    'pf3d_cg_update_xr': [I64, P, P, P, P, P, P],
    'pf3d_cg_beta': [P, I32, P],
    'pf3d_cg_update_p': [I64, P, P, P, P],
    'pf3d_bsr_spgemm_bytes': [I64, P],
    'pf3d_bsr_spgemm_symbolic': [I64, P, P, P, P, P, P, I64, P, P],
    'pf3d_bsr_spgemm_numeric': [I64, P, P, P, P, P, P, P, P, P, I64, P],
    'pf3d_bsr_transpose_bytes': [I64, I64, I64, P],
    'pf3d_bsr_transpose': [I64, I64, I64, P, P, P, P, P, P, P, I64, P],
    'pf3d_amg_smooth_prolongator': [I64, P, P, P, P, P, P, F64, P, P, P],
    'pf3d_bsr_symmetrize': [I64, I64, P, P, P, P, P, P, P],
    'pf3d_bsr_dirichlet': [I64, P, P, P, P, P],
    'pf3d_to_f32': [I64, P, P, P],
    'pf3d_bsr_dirichlet_diag': [I64, P, P, P, P, F64, P],
    'pf3d_gemv_t': [I64, I32, P, I64, P, P, F64, F64, P],
    'pf3d_gemv_n': [I64, I32, P, I64, P, P, F64, F64, P],
    'pf3d_axpby': [I64, F64, P, F64, P, P],
    'pf3d_dot': [I64, P, P, P, P, P],
    'pf3d_block_dot': [I64, I32, P, I64, I32, P, I64, P, P, P],
    'pf3d_block_comb': [I64, I32, P, I64, I32, P, P, I64, P],
    'pf3d_axpby_f32': [I64, F32, P, F32, P, P],
    'pf3d_to_f64': [I64, P, P, P],
    'pf3d_bjacobi': [I64, P, P, P, P, P, P, P],
    'pf3d_pcg_bytes': [I64, I32, P],
    'pf3d_pcg': [I64, P, P, P, P, P, P, F64, I32, I32, P, I64, P, P, P],
    'pf3d_pcg_static': [I64, P, P, P, P, P, P, F64, I32, I32, P, P, P],
    'pf3d_ir_bytes': [I64, P],
    'pf3d_ir_solve': [I64, P, P, P, P, P, P, P, F64, F64, I32, I32, I32, P, I64, P, P, P],
    'pf3d_pcg_mixed': [I64, P, P, P, P, P, P, P, P, F64, I64, I32, I32, P, I64, P, P, P],
    'pf3d_batch_bytes': [I64, P],
    'pf3d_band_max_width': [],
    'pf3d_band_factor': [I64, I64, P, P, P, P, I64, P, P, P, P, P, P, P, I32, I32, P],
    'pf3d_band_load': [I64, P, P, P, I64, P, P, P, P, P, P, P],
    'pf3d_band_chol': [I64, P, P, P, P, P, P, I32, I32, P, P, P, P],
    'pf3d_band_backsolve': [I64, P, P, P, P, P, P, P, P, I32, P],
    'pf3d_band_solve': [I64, P, P, P, P, P, P, P, P, P, I32, P],
    'pf3d_nn_gemm': [I32, I32, I32, P, I64, I64, P, I64, I64, P, I64, I32, P, I32, I32, P, P],
    'pf3d_nn_im2col': [I32, I32, I32, I32, I32, P, P, P],
    'pf3d_nn_col2im': [I32, I32, I32, I32, I32, P, P, I32, P],
    'pf3d_nn_relu_bwd': [I64, P, P, P],
    'pf3d_nn_bias_grad': [I64, I32, P, P, P, I64, P],
    'pf3d_nn_up2_add': [I32, I32, I32, I32, P, P, P, P],
    'pf3d_nn_up2_bwd': [I32, I32, I32, I32, P, P, P],
    'pf3d_nn_node_avg': [I32, I32, I32, P, P, P],
    'pf3d_nn_node_loss_grad': [I32, I32, I32, P, P, P, F32, P, P],
    'pf3d_nn_adam': [I64, P, P, P, P, F32, F32, F32, F32, I32, P],
    'pf3d_nn_to_dofs': [I64, I32, P, P, P, P, P],
    'pf3d_nn_energy_dots': [I32, I64, P, P, P, P, P, P],
    'pf3d_nn_energy_grad': [I32, I32, I32, P, P, P, P, P, P, F32, P, P],
    'pf3d_batch_solve': [I64, I64, P, P, P, P, P, P, P, P, P, P, I32, F64, I32, I32, P, I64,
                         P, P, P, I32, I32, P],
}

_lib = None


class Pf3dError(RuntimeError):
    pass


def load():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError('libpf3d.so not found at %s: run `python -m pyfe3d_amd.build` '
                          '(there is no CPU fallback)' % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.pf3d_last_error.argtypes = []
    lib.pf3d_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def call(name, *args):
    lib = load()
    st = getattr(lib, name)(*args)
    if st < 0:
        raise Pf3dError('%s failed (%d): %s' % (name, st, lib.pf3d_last_error().decode()))
    return st

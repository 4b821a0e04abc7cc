"""ctypes binding of libbemcs_hip.so (declarations in include/bemcs_hip.h).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbemcs_hip.so")

ABI_VERSION = 8
OK, EINVAL, EHIP, ENOCONV, EOVERFLOW = 0, -1, -2, -3, -4

vp = C.c_void_p
i32, i64, u32, u64, f32, f64 = C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_float, C.c_double


class BemcsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bemcs error {code}: {msg}")
        self.code = code


class Species(C.Structure):
    _fields_ = [("x", vp), ("y", vp), ("vx", vp), ("vy", vp), ("vz", vp), ("tag", vp),
                ("count", vp), ("cap", i64)]


class IonDiag(C.Structure):
    _fields_ = [("entered", u32), ("crossed", u32), ("exited", u32), ("lost_to_grid", u32),
                ("see_spawned", u32), ("see_kept", u32), ("mcc_events", u32),
                ("crossed_primary", u32), ("overflow", u32), ("surface_hits", u32),
                ("sum_vx", f64), ("sum_vx2", f64), ("sum_v", f64), ("sum_v2", f64),
                ("sum_E", f64), ("sum_v_primary", f64), ("sum_v_cex", f64)]


class IonStep(C.Structure):
    _fields_ = [("Exy", vp), ("is_bound", vp), ("n_sub", vp), ("cnt_next", vp), ("ppc", vp),
                ("crossed_angle", vp), ("diag", vp), ("e_cnt_next", vp),
                ("dt", f64), ("q_m", f64), ("Lx", f64), ("Ly", f64), ("m_ion", f64), ("q", f64),
                ("see_vth", f64), ("mcc_n0", f64), ("mcc_n0a", f64), ("mcc_rT", f64),
                ("mcc_vth", f64),
                ("dx32", f32), ("dy32", f32), ("x_entry32", f32), ("x_exit32", f32),
                ("x_plume32", f32), ("Lx32", f32), ("Ly32", f32), ("dt32", f32),
                ("nx", i32), ("ny", i32), ("periodic_y", i32), ("neutral_model", i32),
                ("cx_model", i32), ("mcc_enabled", i32), ("seed", u32), ("step", u32),
                ("heat_acc", vp), ("damage_acc", vp), ("see_stage", Species),
                ("half_m32", f32), ("w32", f32), ("ccell32", f32), ("accel32", f32), ("q32", f32),
                ("eth32", f32), ("ycoef32", f32), ("see_defer", i32),
                ("cx_t", vp), ("cx_c", vp), ("see_t", vp), ("see_c", vp),
                ("cx_lo", f64), ("cx_hi", f64), ("see_lo", f64), ("see_hi", f64),
                ("cx_n", i32), ("see_n", i32), ("count_window", i32), ("pad_", i32)]


class ElectronStep(C.Structure):
    _fields_ = [("Exy", vp), ("is_bound", vp), ("n_sub", vp), ("cnt_next", vp), ("overflow", vp),
                ("dt", f64), ("q_m", f64), ("Ly", f64),
                ("dx32", f32), ("dy32", f32), ("Lx32", f32), ("Ly32", f32),
                ("nx", i32), ("ny", i32), ("periodic_y", i32), ("sub_first", i32), ("sub_chunks", i32),
                ("count_window", i32)]


class Inject(C.Structure):
    _fields_ = [("is_bound", vp), ("cnt", vp), ("y_lo", f64 * 2), ("y_hi", f64 * 2), ("x0", f64),
                ("v_drift", f64), ("v_spread", f64), ("Lx", f64), ("Ly", f64),
                ("dx32", f32), ("dy32", f32), ("nx", i32), ("ny", i32), ("kind", i32),
                ("two_band", i32), ("band_f32", i32), ("seed", u32), ("step", u32),
                ("pid_base", u32), ("pad_", u32), ("overflow", vp), ("id_next", vp), ("target", i64)]


class Picard(C.Structure):
    _fields_ = [("cnt_ion", vp), ("cnt_e", vp), ("V_fixed", vp), ("V", vp), ("rho_out", vp),
                ("Exy", vp), ("cd_ion", f64), ("cd_e", f64), ("coeff", f64), ("V_plasma", f64),
                ("Te_up", f64), ("n0", f64), ("q", f64), ("omega", f64), ("dx_m", f64),
                ("dy_m", f64), ("rtol", f64), ("iterations", i32), ("max_iter", i32),
                ("w_ion", vp), ("w_e", vp)]


class StepBegin(C.Structure):
    _fields_ = [("ds_bits", vp), ("count_out_i", vp), ("count_out_e", vp), ("node_zero", vp * 3),
                ("n_nodes", i64), ("diag", vp), ("diag_bytes", i64), ("ppc_stats", vp),
                ("count_in_i", vp), ("count_in_e", vp), ("nin_out", vp), ("count_out_x", vp)]


class StepEnd(C.Structure):
    _fields_ = [("count_i", vp), ("count_e", vp), ("n_sub", vp), ("counts_out", vp), ("nsub_out", vp),
                ("V", vp), ("saddle_idx", vp), ("n_saddle", i32), ("pad_", i32), ("minpot_out", vp)]


class IonCase(C.Structure):
    _fields_ = [("in", Species), ("out", Species), ("e_out", Species), ("p", IonStep)]


class ElectronCase(C.Structure):
    _fields_ = [("in", Species), ("out", Species), ("p", ElectronStep)]


class SubstepsCase(C.Structure):
    _fields_ = [("x", vp), ("y", vp), ("vx", vp), ("vy", vp), ("vz", vp), ("n_dev", vp), ("Exy", vp),
                ("nx", i32), ("ny", i32), ("dx", f64), ("dy", f64), ("q_m", f64), ("dt", f64),
                ("dt2", f64), ("frac", f64), ("ds_max_bits", vp), ("n_sub_out", vp)]


class InjectCase(C.Structure):
    _fields_ = [("sp", Species), ("p", Inject), ("n", i64), ("n_bound_before", i64)]


class SortCase(C.Structure):
    _fields_ = [("in", Species), ("out", Species), ("n_bound", i64), ("nx", i32), ("ny", i32),
                ("dx", f64), ("dy", f64), ("scratch", vp), ("shift_x", i32), ("shift_y", i32)]


class PpcCase(C.Structure):
    _fields_ = [("ppc", vp), ("n_nodes", i64), ("accum", vp), ("stats", vp)]


class Erosion(C.Structure):
    _fields_ = [("damage_acc", vp), ("damage_map", vp), ("eroded_depth", vp), ("is_bound", vp),
                ("diag", vp), ("n_failed", vp), ("n_nodes", i64), ("thresh", f64), ("dy", f64)]


class Thermal(C.Structure):
    _fields_ = [("T_in", vp), ("T_out", vp), ("heat_acc", vp), ("heat_clear", vp), ("is_bound", vp),
                ("grid_bits", vp), ("workspace", vp), ("T_grids", vp), ("cool32", f32),
                ("t300_4_32", f32), ("fo_x32", f32), ("fo_y32", f32), ("nx", i32), ("ny", i32),
                ("n_grids", i32), ("sweeps", i32)]


class GridGeom(C.Structure):
    _fields_ = [("gstart", f64), ("gend", f64), ("V", f64), ("r", f64), ("tan_cham", f64),
                ("delta", f64)]


class Electrodes(C.Structure):
    _fields_ = [("is_bound", vp), ("V_fixed", vp), ("grid_bits", vp), ("grids", GridGeom * 8),
                ("hole_y", f64 * 2), ("Lx", f64), ("Ly", f64), ("screen_r", f64), ("pitch", f64),
                ("V_plasma", f64), ("n_grids", i32), ("n_holes", i32), ("nx", i32), ("ny", i32),
                ("geometry", i32), ("pad_", i32)]


class PercentileCase(C.Structure):
    _fields_ = [("angles", vp), ("n_dev", vp), ("q_percent", f64), ("min_count", u32), ("pad_", u32),
                ("out", vp)]


_SIGS = {
    "bemcs_abi_version": (i32, []),
    "bemcs_last_error": (C.c_char_p, []),
    "bemcs_push_boris": (i32, [vp, vp, vp, vp, vp, i64, vp, vp, vp, vp, vp, i32, i32, f64, f64,
                               f64, f64, i32, vp]),
    "bemcs_deposit_ngp_counts": (i32, [vp, vp, i64, f64, f64, i32, i32, vp, vp]),
    "bemcs_counts_to_rho": (i32, [vp, i64, f64, vp, vp]),
    "bemcs_deposit_cic": (i32, [vp, vp, i64, vp, f64, f64, i32, i32, vp, vp]),
    "bemcs_weights_to_rho": (i32, [vp, i64, f64, vp, vp]),
    "bemcs_substeps": (i32, [vp, vp, vp, vp, vp, vp, i64, vp, i32, i32, f64, f64, f64, f64, f64,
                             f64, vp, vp, vp]),
    "bemcs_poisson_count": (i64, [u32, u32, u32, f64]),
    "bemcs_det_exp": (f64, [f64]),
    "bemcs_det_log": (f64, [f64]),
    "bemcs_philox4x32_10": (None, [vp, vp, vp]),
    "bemcs_fastdiv_check": (i32, [C.c_float, C.c_float, vp, vp]),
    "bemcs_fastdiv_enabled": (i32, [C.c_float, vp]),
    "bemcs_ion_step": (i32, [vp, vp, vp, vp, i64, vp]),
    "bemcs_electron_step": (i32, [vp, vp, vp, i64, vp]),
    "bemcs_electron_push_chunk": (i32, [vp, vp, i32, i64, vp]),
    "bemcs_inject": (i32, [vp, vp, i64, i64, vp]),
    "bemcs_sort_by_cell": (i32, [vp, vp, i64, i32, i32, f64, f64, vp, vp]),
    "bemcs_sort_by_tile": (i32, [vp, vp, i64, i32, i32, f64, f64, i32, i32, vp, vp]),
    "bemcs_poisson_create": (i32, [vp, i32, i32, i32, vp]),
    "bemcs_poisson_create_opt": (i32, [vp, i32, i32, i32, i32, vp]),
    "bemcs_poisson_free": (i32, [vp]),
    "bemcs_poisson_levels": (i32, [vp]),
    "bemcs_poisson_set_level0_tile": (i32, [vp, i32]),
    "bemcs_poisson_update": (i32, [vp, vp, vp]),
    "bemcs_poisson_solve": (i32, [vp, vp, vp, f64, i32, vp, vp, vp]),
    "bemcs_poisson_picard": (i32, [vp, vp, vp, vp, vp]),
    "bemcs_efield": (i32, [vp, i32, i32, f64, f64, i32, vp, vp]),
    "bemcs_poisson_batch_create": (i32, [vp, i32, vp]),
    "bemcs_poisson_batch_free": (i32, [vp]),
    "bemcs_poisson_batch_size": (i32, [vp]),
    "bemcs_poisson_batch_picard": (i32, [vp, vp, vp, vp, vp]),
    "bemcs_percentile": (i32, [vp, vp, i64, f64, u32, vp, vp, vp]),
    "bemcs_ppc_stats": (i32, [vp, i64, vp, vp, vp]),
    "bemcs_step_begin": (i32, [vp, vp]),
    "bemcs_step_end": (i32, [vp, vp]),
    "bemcs_erosion": (i32, [vp, vp]),
    "bemcs_see_commit": (i32, [vp, vp, vp, vp, vp, i32, i32, C.c_float, C.c_float, C.c_float,
                               C.c_float, i64, vp]),
    "bemcs_thermal_workspace_bytes": (i64, [i32, i32, i32]),
    "bemcs_thermal": (i32, [vp, vp]),
    "bemcs_thermal_noalloc": (i32, [vp, vp]),
    "bemcs_rf_set": (i32, [vp, vp, i64, i32, vp, i32, f64, vp]),
    "bemcs_thermal_conduction": (i32, [vp, vp, vp, i32, i32, C.c_float, C.c_float, vp]),
    "bemcs_electrodes": (i32, [vp, vp]),
    "bemcs_grid_means": (i32, [vp, vp, i32, i32, i32, vp, vp, vp]),
    "bemcs_ion_step_batched": (i32, [vp, i32, i64, vp, vp]),
    "bemcs_electron_step_batched": (i32, [vp, i32, i64, vp, vp]),
    "bemcs_substeps_batched": (i32, [vp, i32, i64, vp]),
    "bemcs_inject_batched": (i32, [vp, i32, i64, vp]),
    "bemcs_sort_by_cell_batched": (i32, [vp, i32, i64, i64, vp]),
    "bemcs_step_begin_batched": (i32, [vp, i32, i64, vp]),
    "bemcs_step_end_batched": (i32, [vp, i32, vp]),
    "bemcs_ppc_stats_batched": (i32, [vp, i32, vp]),
    "bemcs_percentile_batched": (i32, [vp, i32, vp]),
    "bemcs_graph_capture_begin": (i32, [vp]),
    "bemcs_graph_capture_end": (i32, [vp, i32, vp]),
    "bemcs_graph_launch": (i32, [vp, vp]),
    "bemcs_graph_destroy": (i32, [vp]),
}

EXPORTED = tuple(_SIGS)

_lib = None


def load():
    """Load the shared library (once) and declare every entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with "
                          "`python -m thruster_cx_code_amd.build` (no CPU fallback exists)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.bemcs_abi_version() != ABI_VERSION:
        raise ImportError("libbemcs_hip.so ABI version mismatch")
    _lib = lib
    return lib


def check(rc, what=""):
    if rc != OK:
        msg = load().bemcs_last_error().decode(errors="replace")
        raise BemcsError(rc, f"{what}: {msg}" if what else msg)
    return rc

"""ctypes view of the C ABI in ``include/dsph_abi.h`` (libdsph.so).

The library is built in-tree by ``python -m dualsphysics_amd.build``. Loading
fails loudly when it is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import os

LIB_PATH = Path(os.environ.get("DSPH_LIB_PATH", Path(__file__).resolve().parent / "libdsph.so"))

DSPH_OK, DSPH_E_ARG, DSPH_E_LAUNCH, DSPH_E_NOMEM, DSPH_E_STATE, DSPH_E_NUMERIC = range(6)
KERNEL_WENDLAND = 2
VISCO_ARTIFICIAL, VISCO_LAMINARSPS, VISCO_LAMINAR = 1, 2, 3
DDT_NONE, DDT_DDT, DDT_DDT2, DDT_DDT2FULL = 0, 1, 2, 3
BC_DBC, BC_MDBC = 1, 2
SLIP_VEL0, SLIP_NOSLIP, SLIP_FREESLIP = 1, 2, 3
SHIFT_NONE, SHIFT_NOBOUND, SHIFT_NOFIXED, SHIFT_FULL = 0, 1, 2, 3
BMODE_DBC, BMODE_MDBC2, BMODE_MDBC2OFF = 0, 1, 2
OPT_FORCES_KERNEL = 1
FORCES_TILE, FORCES_ROWS = 0, 1
OPT_BOUND_KERNEL = 2
BOUND_AUTO, BOUND_TILE, BOUND_ROWS = 0, 1, 2
NP_LIMIT = 1 << 28              # DSPH_NP_LIMIT: particles per simulation / interaction
NOALLOC_NP_MAX = 1 << 20        # DSPH_NOALLOC_NP_MAX: particles per no-alloc interaction

u32p = C.POINTER(C.c_uint32)


class DsphCte(C.Structure):
    _fields_ = [(n, C.c_float) for n in (
        "massb", "massf", "kernelh", "kernelsize2", "poscellsize", "awen", "bwenh", "cs0",
        "eta2", "ddtkh", "ddtgz", "scell", "kernelsize", "dp", "cteb", "gamma", "rhopzero",
        "ovrhopzero", "movlimit")] + [
        ("tboundary", C.c_uint32), ("cellcode", C.c_uint32)] + [
        (n, C.c_double) for n in (
            "domposminx", "domposminy", "domposminz", "maprealposminx", "maprealposminy",
            "maprealposminz", "maprealsizex", "maprealsizey", "maprealsizez")] + [
        ("periactive", C.c_uint32), ("simulate2d", C.c_uint32),
        ("xperinc", C.c_double * 3), ("yperinc", C.c_double * 3), ("zperinc", C.c_double * 3)]


class DsphDivData(C.Structure):
    _fields_ = [("scelldiv", C.c_int32), ("ncx", C.c_int32), ("ncy", C.c_int32),
                ("ncz", C.c_int32), ("ncw", C.c_int32), ("cellfluid", C.c_uint32),
                ("cellzerox", C.c_int32), ("cellzeroy", C.c_int32), ("cellzeroz", C.c_int32),
                ("d_beginendcell", C.c_void_p), ("scell", C.c_float),
                ("domcellcode", C.c_uint32), ("domposminx", C.c_double),
                ("domposminy", C.c_double), ("domposminz", C.c_double),
                ("kernelsize2", C.c_float), ("poscellsize", C.c_float)]


class DsphInterParms(C.Structure):
    _fields_ = [("tkernel", C.c_int32), ("tvisco", C.c_int32), ("tdensity", C.c_int32),
                ("viscob", C.c_float), ("viscof", C.c_float), ("np", C.c_uint32),
                ("npb", C.c_uint32), ("npbok", C.c_uint32), ("divdata", DsphDivData),
                ("d_dcell", C.c_void_p), ("d_poscell", C.c_void_p), ("d_velrho", C.c_void_p),
                ("d_code", C.c_void_p), ("d_spstaurho2", C.c_void_p), ("d_viscdt", C.c_void_p),
                ("d_ar", C.c_void_p), ("d_ace", C.c_void_p), ("d_sps2strain", C.c_void_p),
                ("d_velgrad", C.c_void_p), ("d_boundmode", C.c_void_p),
                ("d_tangenvel", C.c_void_p), ("shiftmode", C.c_int32),
                ("d_shiftposfs", C.c_void_p), ("d_ftomassp", C.c_void_p),
                ("d_boundnor", C.c_void_p), ("d_nopenshift", C.c_void_p),
                ("d_motionvel", C.c_void_p)]


class DsphMotion(C.Structure):
    _fields_ = [("type", C.c_int32), ("idbegin", C.c_uint32), ("count", C.c_uint32),
                ("freq", C.c_double * 3), ("ampl", C.c_double * 3), ("phase", C.c_double * 3),
                ("start", C.c_double), ("finish", C.c_double), ("axisp1", C.c_double * 3),
                ("axisp2", C.c_double * 3), ("angfreq", C.c_double), ("angampl", C.c_double),
                ("angphase", C.c_double), ("angvel", C.c_double)]


class DsphFloating(C.Structure):
    _fields_ = [("begin", C.c_uint32), ("count", C.c_uint32), ("massbody", C.c_double),
                ("masspart", C.c_double), ("center", C.c_double * 3),
                ("inertia", C.c_double * 9), ("velini", C.c_double * 3),
                ("omegaini", C.c_double * 3)]


class DsphFloatingState(C.Structure):
    _fields_ = [("center", C.c_double * 3)] + [
        (n, C.c_float * 3) for n in ("angles", "fvel", "fomega", "facelin", "faceang")]


class DsphCase(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "dp", "kernelh", "cteb", "gamma", "rhopzero", "cs0", "massfluid", "massbound")] + [
        ("gravity", C.c_double * 3), ("cflnumber", C.c_double), ("dtini", C.c_double),
        ("dtmin", C.c_double)] + [(n, C.c_int32) for n in (
            "tstep", "verletsteps", "tkernel", "tvisco", "tdensity", "tboundary")] + [
        (n, C.c_double) for n in ("visco", "viscoboundfactor", "ddtvalue", "rhopoutmin",
                                  "rhopoutmax")] + [
        ("cellmode", C.c_int32), ("celldomfixed", C.c_int32),
        ("maprealposmin", C.c_double * 3), ("maprealposmax", C.c_double * 3)] + [
        (n, C.c_uint32) for n in ("casenp", "casenbound", "casenfixed", "casenmoving",
                                  "casenfloat", "casenfluid")] + [("motion", DsphMotion),
                                                                  ("slipmode", C.c_int32),
        ("shiftmode", C.c_int32), ("shiftcoef", C.c_double), ("shifttfs", C.c_double),
        ("periactive", C.c_int32), ("perinc", (C.c_double * 3) * 3),
        ("ftcount", C.c_uint32), ("nopenetration", C.c_int32),
        ("floatings", C.POINTER(DsphFloating)),
        ("simulate2d", C.c_int32), ("pad_2d", C.c_int32), ("simulate2dposy", C.c_double)]


class DsphMlp(C.Structure):
    _fields_ = [("d_params", C.c_void_p), ("d_wbf16", C.c_void_p), ("d_adam_m", C.c_void_p),
                ("d_adam_v", C.c_void_p), ("feat_mean", C.c_float * 11),
                ("feat_scale", C.c_float * 11), ("out_scale", C.c_float)]


class DsphEnsState(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("np", "npbok", "active", "nstep", "flags", "p0", "np0",
                                          "pad")] + [
        (n, C.c_double) for n in ("time", "dt_last", "dt_next")]


class DsphEnsShape(C.Structure):
    _fields_ = [("variant", C.c_int32)] + [
        (n, C.c_uint32) for n in ("persistent", "chunk", "waves_per_cu", "bound_block",
                                  "bound_tile")] + [
        (n, C.c_float) for n in ("ms_plain", "ms_persistent", "ms_bound")]


class DsphStats(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("np", "npb", "npbok", "nct", "nstep",
                                          "npf_out_total")] + [
        (n, C.c_double) for n in ("timestep", "dt_last", "velmax", "acemax", "viscdtmax")] + [
        ("ndiv", C.c_uint32), ("ndivfull", C.c_uint32), ("symplecticdtpre", C.c_double)]


# Every exported entry point of include/dsph_abi.h with its argument types.
_vp, _i, _u, _f, _d, _sz = C.c_void_p, C.c_int, C.c_uint32, C.c_float, C.c_double, C.c_size_t
SIGNATURES = {
    "dsph_ctx_create": [_i, C.POINTER(_vp)],
    "dsph_ctx_destroy": [_vp],
    "dsph_ctx_reserve": [_vp, _u, _u],
    "dsph_ctx_set_stream": [_vp, _vp],
    "dsph_ctx_get_stream": [_vp],
    "dsph_synchronize": [_vp],
    "dsph_last_error": [_vp],
    "dsph_malloc": [_vp, _sz, C.POINTER(_vp)],
    "dsph_free": [_vp, _vp],
    "dsph_memset_async": [_vp, _vp, _i, _sz],
    "dsph_memcpy_h2d": [_vp, _vp, _vp, _sz],
    "dsph_memcpy_d2h": [_vp, _vp, _vp, _sz],
    "dsph_memcpy_d2d_async": [_vp, _vp, _vp, _sz],
    "dsph_set_constants": [_vp, C.POINTER(DsphCte)],
    "dsph_ctx_set_option": [_vp, _i, _i],
    "dsph_cell_limits_sync": [_vp, _u, _u, _u, _vp, _vp, u32p, u32p],
    "dsph_presort_full": [_vp, _u, _u, _vp, _vp, u32p, u32p, _vp, _vp],
    "dsph_presort_fluid": [_vp, _u, _u, _u, _vp, _vp, u32p, u32p, _vp, _vp],
    "dsph_sort_by_cell": [_vp, _vp, _vp, _u, _u],
    "dsph_begin_end_cell": [_vp, _i, _u, _u, _u, _u, _vp, _vp],
    "dsph_cell_begin_end_sync": [_vp, _vp, _u, _u, _vp],
    "dsph_sort_data_basic": [_vp, _u, _u, _vp] + [_vp] * 12,
    "dsph_sort_data_basic_poscell": [_vp, _u, _u, _vp] + [_vp] * 12 + [C.POINTER(_d), _f, _vp],
    "dsph_sort_data_f4": [_vp, _u, _u, _vp, _vp, _vp],
    "dsph_sort_data_f3": [_vp, _u, _u, _vp, _vp, _vp],
    "dsph_sort_data_sym6": [_vp, _u, _u, _vp, _vp, _vp],
    "dsph_sort_data_d2_d_f4": [_vp, _u, _u, _vp] + [_vp] * 6,
    "dsph_update_poscell": [_vp, _u, C.POINTER(_d), _f, _vp, _vp, _vp],
    "dsph_interaction_forces": [_vp, C.POINTER(DsphInterParms)],
    "dsph_interaction_forces_noalloc": [_vp, C.POINTER(DsphInterParms)],
    "dsph_sps_tau": [_vp, _u, _u, _f, _f, _vp, _vp, _vp],
    "dsph_mdbc_correction": [_vp, _u, C.POINTER(DsphDivData), C.POINTER(_d)] + [_vp] * 6,
    "dsph_mdbc2_correction": [_vp, _u, C.POINTER(_f), C.POINTER(DsphDivData),
                              C.POINTER(_d)] + [_vp] * 10,
    "dsph_mdbc2_correction_slip": [_vp, _i, _u, C.POINTER(_f), C.POINTER(DsphDivData),
                                   C.POINTER(_d)] + [_vp] * 10,
    "dsph_mdbc2_nopen_step": [_vp, _u, _u, _vp, _vp, _d, _vp, _u, _vp, _vp, _vp],
    "dsph_mdbc2_motion_vel_ace": [_vp, _u, _u, _u, _d, _vp, _u, _vp, _vp, _vp, _vp],
    "dsph_velmax2": [_vp, _u, _u, _vp, _vp],
    "dsph_acemax2": [_vp, _u, _u, _vp, _vp],
    "dsph_acemax2_normal": [_vp, _u, _u, _vp, _vp, _vp, _vp],
    "dsph_periodic_ignore": [_vp, _u, _vp],
    "dsph_periodic_make_list": [_vp, _u, _u, _u, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "dsph_periodic_duplicate": [_vp, _u, _u, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                _vp, _vp, _vp, _vp, _vp],
    "dsph_redu_max_f32": [_vp, _u, _u, _vp, _vp],
    "dsph_redu_max_f32_sync": [_vp, _u, _u, _vp, C.POINTER(_f)],
    "dsph_dt_variable": [_vp, _vp, _d, _d, _d, _d, _vp, _u, _vp],
    "dsph_dt_advance": [_vp, _vp],
    "dsph_step_symplectic_pre": [_vp, _u, _u, _vp, _vp, _vp, _vp, _vp, _d, _vp, _u, _f, _f, _f,
                                 C.POINTER(_f), _vp, _vp, _vp, _vp],
    "dsph_step_symplectic_cor": [_vp, _u, _u, _vp, _vp, _vp, _vp, _vp, _d, _d, _vp, _u, _f, _f, _f,
                                 C.POINTER(_f), _vp, _vp, _vp, _vp],
    "dsph_step_verlet": [_vp, _u, _u, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _f, _f, _f, C.POINTER(_f),
                         _vp, _vp, _vp, _vp],
    "dsph_step_pos2": [_vp, _u, _u] + [_vp] * 8,
    "dsph_step_pos": [_vp, _u, _u] + [_vp] * 6,
    "dsph_move_lin_bound": [_vp, _u, _u, _u, C.POINTER(_d), C.POINTER(_f)] + [_vp] * 6,
    "dsph_ft_calc_ridp": [_vp, _u, _u, _u, _u, _vp, _vp],
    "dsph_ft_parts_sum_ace": [_vp, _u] + [_vp] * 7,
    "dsph_ft_parts_update": [_vp, _d, _u, _u, C.POINTER(_f), C.POINTER(_f), C.POINTER(_d)] +
                            [_vp] * 6,
    "dsph_move_motion_device": [_vp, _u, C.POINTER(DsphMotion), _d] + [_vp] * 8,
    "dsph_move_mat_bound": [_vp, _u, _u, _u, C.POINTER(_d), _d] + [_vp] * 7,
    "dsph_motion_matrix": [C.POINTER(DsphMotion), _d, _d, C.POINTER(_d)],
    "dsph_pips_count_sync": [_vp, _u, _u, _u, _u, C.POINTER(DsphDivData), _vp, _vp,
                             C.POINTER(C.c_uint64)],
    "dsph_pips_count_split_sync": [_vp, _u, _u, _u, _u, C.POINTER(DsphDivData), _vp, _vp,
                                   C.POINTER(C.c_uint64)],
    "dsph_ctx_set_timing": [_vp, _i],
    "dsph_ctx_timing_collect": [_vp],
    "dsph_ctx_timing_get": [_vp, _u, C.c_char_p, _sz, C.POINTER(C.c_uint64), C.POINTER(_d)],
    "dsph_ctx_timing_reset": [_vp],
    "dsph_sim_create": [_i, C.POINTER(DsphCase), _u, _vp, _vp, _vp, _vp, C.POINTER(_vp)],
    "dsph_sim_destroy": [_vp],
    "dsph_sim_set_time": [_vp, _d, _d, _u],
    "dsph_run_shifting": [_vp, _u, _u, _d, _vp, _u, _d, _d, _d, _f, _d, _vp, _vp],
    "dsph_sim_gauge_points": [_vp, _u, _vp, _vp, _vp],
    "dsph_sim_gauge_maxz": [_vp, _u, _vp, _vp, _vp],
    "dsph_sim_gauge_force": [_vp, _u, _u, _vp],
    "dsph_gauge_force": [_vp, _u, _u, _u, C.POINTER(DsphDivData)] + [_vp] * 6,
    "dsph_gauge_points": [_vp, _u, _vp, C.POINTER(DsphDivData), C.POINTER(_d), C.POINTER(_d)]
    + [_vp] * 6,
    "dsph_gauge_maxz": [_vp, _u, _vp, _vp, C.POINTER(DsphDivData), _d, _vp, _vp, _vp, _vp],
    "dsph_sim_last_error": [_vp],
    "dsph_sim_run": [_vp, _u, _i],
    "dsph_sim_stats": [_vp, C.POINTER(DsphStats)],
    "dsph_sim_download": [_vp, u32p, _vp, _vp, _vp, _vp],
    "dsph_sim_download_forces": [_vp, _vp, _vp],
    "dsph_sim_pips": [_vp, C.POINTER(C.c_uint64)],
    "dsph_sim_floating": [_vp, _u, C.POINTER(DsphFloatingState)],
    "dsph_sim_interact": [_vp],
    "dsph_sim_ctx": [_vp],
    "dsph_mlp_param_count": [],
    "dsph_mlp_bf16_count": [],
    "dsph_mlp_sync": [_vp, C.POINTER(DsphMlp)],
    "dsph_mlp_forward": [_vp, C.POINTER(DsphMlp), _u, _vp, _vp],
    "dsph_mlp_grad": [_vp, C.POINTER(DsphMlp), _u, _vp, _vp, _vp, _sz, _vp, _vp],
    "dsph_mlp_adam": [_vp, C.POINTER(DsphMlp), _vp, _f, _f, _f, _f, _u],
    "dsph_sps_tau_learned": [_vp, C.POINTER(DsphMlp), _u, _u, _f, _f] + [_vp] * 5,
    "dsph_filter_probes": [_vp, _u, _vp, _d, C.POINTER(DsphDivData), _f] + [_vp] * 5,
    "dsph_sim_set_closure": [_vp, C.POINTER(DsphMlp)],
    "dsph_sim_closure_nu": [_vp, _vp, _vp],
    "dsph_sim_filter_probes": [_vp, _u, _vp, _d, _vp],
    "dsph_ens_create": [_i, _u, C.POINTER(DsphCase), _vp, _vp, _vp, _vp, _vp, C.POINTER(_vp)],
    "dsph_ens_destroy": [_vp],
    "dsph_ens_last_error": [_vp],
    "dsph_ens_run": [_vp, _u, _u, _vp],
    "dsph_ens_set_graph": [_vp, _i],
    "dsph_ens_case_state": [_vp, _u, C.POINTER(DsphEnsState)],
    "dsph_ens_download": [_vp, _u, _vp, _vp, _vp, _vp, _vp],
    "dsph_ens_autotune": [_vp, _u, _vp, _vp],
    "dsph_ens_set_variant": [_vp, _u, C.c_int32],
    "dsph_ens_get_shape": [_vp, _u, C.POINTER(DsphEnsShape)],
    "dsph_ens_set_shape": [_vp, _u, C.POINTER(DsphEnsShape)],
    "dsph_ens_info": [_vp, _vp],
    "dsph_ens_set_timing": [_vp, _i],
    "dsph_ens_context": [_vp],
}
RESTYPES = {"dsph_last_error": C.c_char_p, "dsph_sim_last_error": C.c_char_p,
            "dsph_ctx_get_stream": _vp, "dsph_sim_ctx": _vp, "dsph_ens_last_error": C.c_char_p,
            "dsph_ens_context": _vp}

_lib = None


def load() -> C.CDLL:
    """Loads libdsph.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m dualsphysics_amd.build`"
                           " (there is no non-HIP fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = RESTYPES.get(name, C.c_int)
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return list(SIGNATURES)

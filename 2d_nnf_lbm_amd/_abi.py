"""ctypes binding of the C ABI declared in include/lbm_amd.h."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm_amd.so")

LBM_OK = 0
LBM_ERR_PARAM = -1
LBM_ERR_HIP = -2
LBM_ERR_UNSUPPORTED = -3
LBM_ERR_CAPACITY = -4
LBM_ERR_NAN = -5
LBM_ERR_IO = -6

LBM_NEWTONIAN, LBM_BINGHAM, LBM_CASSON = 0, 1, 2
LBM_BOUNCE_BACK, LBM_ZOU_HE, LBM_EQUILIBRIUM = 0, 1, 2
LBM_CONTACT_GLOWINSKI, LBM_CONTACT_WAN_TUREK, LBM_CONTACT_LJ = 0, 1, 2
LBM_RATE_LAGGED, LBM_RATE_LOCAL = 0, 1
LBM_MOTION_FREE, LBM_MOTION_OSC_X = 0, 1
LBM_FP64, LBM_FP32, LBM_FP16, LBM_BF16 = 0, 1, 2, 3
STORAGE = {"fp64": LBM_FP64, "fp32": LBM_FP32, "fp16": LBM_FP16, "bf16": LBM_BF16}

PF_NAMES = ["x", "y", "ang", "vx", "vy", "w", "diam", "fx", "fy", "tq", "fcx", "fcy",
            "rhop", "mp", "ip", "zeta", "vox", "voy", "vow", "fhx", "fhy", "fhw"]
PF = {n: i for i, n in enumerate(PF_NAMES)}
PF_COUNT = len(PF_NAMES)
H_NAMES = ["x", "y", "ang", "vx", "vy", "w", "fx", "fy", "tq", "fhx", "fhy", "fhw"]
H = {n: i for i, n in enumerate(H_NAMES)}
H_PCOUNT = len(H_NAMES)


class Params(ctypes.Structure):
    """lbm_params (include/lbm_amd.h)."""
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("rheology", ctypes.c_int32), ("rate_mode", ctypes.c_int32),
        ("q_mode", ctypes.c_int32), ("update_angle", ctypes.c_int32),
        ("storage", ctypes.c_int32), ("max_particles", ctypes.c_int32),
        ("motion", ctypes.c_int32), ("boundary", ctypes.c_int32),
        ("omega", ctypes.c_double),
        ("s0", ctypes.c_double), ("s1", ctypes.c_double), ("s2", ctypes.c_double),
        ("s3", ctypes.c_double), ("s5", ctypes.c_double), ("sq", ctypes.c_double),
        ("nu", ctypes.c_double),
        ("mu0", ctypes.c_double), ("tau_y", ctypes.c_double),
        ("m_pap", ctypes.c_double), ("gamma_min", ctypes.c_double),
        ("rho0", ctypes.c_double), ("gy", ctypes.c_double),
        ("wall", ctypes.c_double * 8), ("corner", ctypes.c_double * 2),
        ("zeta_c", ctypes.c_double), ("cij", ctypes.c_double),
        ("eps_p", ctypes.c_double), ("eps_w", ctypes.c_double),
        ("e_p", ctypes.c_double), ("e_w", ctypes.c_double),
        ("osc_u", ctypes.c_double), ("osc_period", ctypes.c_double),
        ("contact", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("wall_edge", ctypes.c_double * 4),
    ]


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_hp = ctypes.c_void_p

# name -> (argtypes after the handle)
SIGNATURES = {
    "lbm_destroy": [],
    "lbm_sync": [],
    "lbm_initialize": [],
    "lbm_set_f": [_dp], "lbm_get_f": [_dp],
    "lbm_set_fpost": [_dp], "lbm_get_fpost": [_dp],
    "lbm_set_fields": [_dp, _dp, _dp, _dp, _dp],
    "lbm_get_fields": [_dp, _dp, _dp, _dp, _dp],
    "lbm_get_stress": [_dp],
    "lbm_get_image": [_ip],
    "lbm_set_particles": [ctypes.c_int32, _dp],
    "lbm_get_particles": [_dp],
    "lbm_get_phi": [ctypes.c_int32, _dp],
    "lbm_restart_write": [ctypes.c_char_p], "lbm_restart_read": [ctypes.c_char_p],
    "lbm_snapshot_size": [ctypes.POINTER(ctypes.c_uint64)],
    "lbm_snapshot_save": [ctypes.c_void_p, ctypes.c_uint64],
    "lbm_snapshot_load": [ctypes.c_void_p, ctypes.c_uint64],
    "lbm_collide": [], "lbm_copy_post": [], "lbm_stream": [],
    "lbm_boundary": [ctypes.c_int32],
    "lbm_macros": [], "lbm_spm_phi": [], "lbm_spm_forcing": [],
    "lbm_hydro_force": [], "lbm_correct": [],
    "lbm_contact_force": [_ip, ctypes.c_int32, _ip],
    "lbm_particle_update": [ctypes.c_int32],
    "lbm_total_mass": [_dp],
    "lbm_step": [ctypes.c_int32, ctypes.c_int32, _dp],
    "lbm_last_pairs": [_ip, ctypes.c_int32, _ip],
    "lbm_step_timed": [ctypes.c_int32, ctypes.c_int32, _dp, _dp],
    "lbm_step_hist_static": [ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p],
    "lbm_adjoint_velocity": [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _dp, _dp, _dp],
    # batch entries (first argument: the lbm_batch pointer)
    "lbm_batch_destroy": [],
    "lbm_batch_size": [_ip],
    "lbm_batch_initialize": [],
    "lbm_batch_step": [ctypes.c_int32, ctypes.c_int32, _dp],
    "lbm_batch_step_timed": [ctypes.c_int32, ctypes.c_int32, _dp],
    "lbm_batch_set_graph": [ctypes.c_int32],
    "lbm_batch_set_tiles": [_ip],
    "lbm_batch_get_tiles": [_ip],
    "lbm_batch_tune": [ctypes.c_int32, _ip],
}

_lib = None


def load():
    """Load liblbm_amd.so; raises when the HIP library has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "liblbm_amd.so is missing: build it with python -c 'import __graft_entry__ as g; "
            "g.build()' (or python 2d_nnf_lbm_amd/build.py); there is no CPU fallback")
    lib = ctypes.CDLL(LIB_PATH)
    lib.lbm_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(_hp)]
    lib.lbm_create.restype = ctypes.c_int
    lib.lbm_last_error.argtypes = [_hp]
    lib.lbm_last_error.restype = ctypes.c_char_p
    lib.lbm_batch_create.argtypes = [ctypes.c_int32, ctypes.POINTER(Params), ctypes.POINTER(_hp)]
    lib.lbm_batch_create.restype = ctypes.c_int
    lib.lbm_batch_last_error.argtypes = [_hp]
    lib.lbm_batch_last_error.restype = ctypes.c_char_p
    lib.lbm_batch_sim.argtypes = [_hp, ctypes.c_int32]
    lib.lbm_batch_sim.restype = _hp
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = [_hp] + args
        fn.restype = ctypes.c_int
    _lib = lib
    return lib


def exported_names():
    return (["lbm_create", "lbm_last_error", "lbm_batch_create", "lbm_batch_last_error",
             "lbm_batch_sim"] + list(SIGNATURES))

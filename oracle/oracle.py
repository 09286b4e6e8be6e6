"""ctypes wrapper around the CPU oracle (lbm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm_oracle.so")

(PF_X, PF_Y, PF_ANG, PF_VX, PF_VY, PF_W, PF_DIAM, PF_FX, PF_FY, PF_TQ, PF_FCX,
 PF_FCY, PF_RHOP, PF_MP, PF_IP, PF_ZETA, PF_VOX, PF_VOY, PF_VOW, PF_FHX, PF_FHY,
 PF_FHW) = range(22)
PF_COUNT = 22
H_PCOUNT = 12
(H_X, H_Y, H_ANG, H_VX, H_VY, H_W, H_FX, H_FY, H_TQ, H_FHX, H_FHY, H_FHW) = range(12)


def build(force=False):
    """Compile the oracle into oracle/liblbm_oracle.so (gcc, no FMA contraction)."""
    src = os.path.join(HERE, "lbm_oracle.c")
    if (not force and os.path.exists(LIB_PATH)
            and os.path.getmtime(LIB_PATH) >= os.path.getmtime(src)):
        return LIB_PATH
    cmd = ["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
           src, "-o", LIB_PATH + ".tmp", "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(LIB_PATH + ".tmp", LIB_PATH)
    return LIB_PATH


class Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("rheology", ctypes.c_int32), ("rate_mode", ctypes.c_int32),
        ("q_mode", ctypes.c_int32), ("update_angle", ctypes.c_int32),
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
        ("contact", ctypes.c_int32), ("pad_", ctypes.c_int32),
        ("wall_edge", ctypes.c_double * 4),
    ]


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class _CState(ctypes.Structure):
    _fields_ = [
        ("f", _dp), ("fpost", _dp), ("rho", _dp), ("ux", _dp), ("uy", _dp),
        ("g", _dp), ("omega_bn", _dp), ("stress", _dp), ("image", _ip),
        ("phi", _dp), ("part", _dp), ("s_mrt", ctypes.c_double * 9),
        ("dpw_stale", ctypes.c_double), ("np", ctypes.c_int32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(Params)
        S = ctypes.POINTER(_CState)
        for name in ("orc_initialize", "orc_collide", "orc_copy_post", "orc_stream",
                     "orc_macros", "orc_phi", "orc_forcing", "orc_hydro",
                     "orc_correct"):
            getattr(L, name).argtypes = [P, S]
            getattr(L, name).restype = None
        L.orc_update.argtypes = [P, S, ctypes.c_int32]
        L.orc_update.restype = None
        L.orc_boundary.argtypes = [P, S, ctypes.c_int32]
        L.orc_contact.argtypes = [P, S, _ip, ctypes.c_int32]
        L.orc_contact.restype = ctypes.c_int32
        L.orc_total_mass.argtypes = [P, S]
        L.orc_total_mass.restype = ctypes.c_double
        L.orc_yielded_nodes.argtypes = [P, S]
        L.orc_yielded_nodes.restype = ctypes.c_int64
        L.orc_run.argtypes = [P, S, ctypes.c_int32, ctypes.c_int32, _dp]
        L.orc_feq.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_relaxation.argtypes = [P, ctypes.c_double, _dp]
        _lib = L
    return _lib


def make_params(cfg):
    """Params from a plain dict (keys as in the struct; wall/corner sequences)."""
    p = Params()
    for name, ctype in Params._fields_:
        if name in ("wall", "corner", "wall_edge"):
            vals = cfg.get(name, [0.0] * {"wall": 8, "corner": 2, "wall_edge": 4}[name])
            arr = getattr(p, name)
            for i, v in enumerate(vals):
                arr[i] = float(v)
        elif name in cfg:
            setattr(p, name, cfg[name])
    return p


class State:
    """Reference-layout state: F(0:8,0:NX,0:NY) etc. held in numpy arrays."""

    def __init__(self, params, np_particles=0):
        self.p = params
        nx, ny = params.nx, params.ny
        n = (nx + 1) * (ny + 1)
        self.nodes = n
        self.f = np.zeros((ny + 1, nx + 1, 9))
        self.fpost = np.zeros((ny + 1, nx + 1, 9))
        self.rho = np.zeros((ny + 1, nx + 1))
        self.ux = np.zeros((ny + 1, nx + 1))
        self.uy = np.zeros((ny + 1, nx + 1))
        self.g = np.zeros((ny + 1, nx + 1, 2))
        self.omega_bn = np.zeros((ny + 1, nx + 1))
        self.stress = np.zeros((ny + 1, nx + 1, 3))
        self.image = np.zeros((ny + 1, nx + 1), dtype=np.int32)
        self.np = np_particles
        self.phi = np.zeros((max(np_particles, 1), ny + 1, nx + 1))
        self.part = np.zeros((PF_COUNT, max(np_particles, 1)))
        self.c = _CState()
        self.c.dpw_stale = float("inf")
        self.c.np = np_particles
        self._bind()

    def _bind(self):
        c = self.c
        for name in ("f", "fpost", "rho", "ux", "uy", "g", "omega_bn", "stress",
                     "phi", "part"):
            arr = getattr(self, name)
            assert arr.flags.c_contiguous and arr.dtype == np.float64
            setattr(c, name, arr.ctypes.data_as(_dp))
        c.image = self.image.ctypes.data_as(_ip)

    def set_particles(self, part):
        """part: (PF_COUNT, np) array."""
        self.part[:, : self.np] = part
        self._bind()

    @property
    def s_mrt(self):
        return np.array(self.c.s_mrt[:])

    def call(self, name, *args):
        return getattr(lib(), name)(ctypes.byref(self.p), ctypes.byref(self.c), *args)

    def initialize(self):
        self.call("orc_initialize")

    def step_parts(self, kk):
        """One time step, routine by routine, in the reference order."""
        self.contact()
        self.call("orc_update", kk)
        self.call("orc_phi")
        self.call("orc_collide")
        self.call("orc_copy_post")
        self.call("orc_stream")
        self.call("orc_boundary", kk)
        self.call("orc_macros")
        self.call("orc_forcing")
        self.call("orc_hydro")
        self.call("orc_correct")

    def contact(self, max_pairs=0):
        if max_pairs:
            buf = np.zeros(2 * max_pairs, dtype=np.int32)
            n = self.call("orc_contact", buf.ctypes.data_as(_ip), max_pairs)
            return buf[: 2 * min(n, max_pairs)].reshape(-1, 2), n
        return self.call("orc_contact", None, 0)

    def run(self, kk0, nsteps, record=True):
        rec = self.np * H_PCOUNT + 2
        hist = np.zeros((nsteps, rec)) if record else None
        lib().orc_run(ctypes.byref(self.p), ctypes.byref(self.c), kk0, nsteps,
                      hist.ctypes.data_as(_dp) if record else None)
        return hist

    def total_mass(self):
        return self.call("orc_total_mass")

    def yielded_nodes(self):
        return self.call("orc_yielded_nodes")


def feq(ux, uy, rho):
    out = (ctypes.c_double * 9)()
    lib().orc_feq(ux, uy, rho, out)
    return np.array(out[:])

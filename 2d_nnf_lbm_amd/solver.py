"""Host-side mirror of the reference's per-step subroutine interface.

`D2Q9` wraps one lbm_handle.  Its methods carry the reference subroutine
names (F6 = 6_FreeFall_YieldStress/D2Q9_BinghamSediment_Validated_20191127.f90)
and the same argument meaning; the arrays they act on live on the GPU and are
exchanged in the reference layout (F(0:8,0:NX,0:NY) etc., returned here as
numpy arrays indexed [j, i, q]).
"""
import ctypes
import os

import numpy as np

from . import _abi


class LBMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("lbm error %d: %s" % (code, msg))
        self.code = code


def make_params(cfg):
    """lbm_params from a dict of field values (missing fields are zero)."""
    p = _abi.Params()
    for name, _ in _abi.Params._fields_:
        if name not in cfg:
            continue
        v = cfg[name]
        if name in ("wall", "corner", "wall_edge"):
            arr = getattr(p, name)
            for i, x in enumerate(v):
                arr[i] = float(x)
        elif name == "storage" and isinstance(v, str):
            p.storage = _abi.STORAGE[v]
        else:
            setattr(p, name, v)
    return p


def _ptr(a, ct=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ct)) if a is not None else None


class D2Q9:
    """One simulation on the GPU.  cfg: dict of lbm_params fields."""

    def __init__(self, cfg, _handle=None):
        self.lib = _abi.load()
        self.params = make_params(cfg)
        self.nx, self.ny = self.params.nx, self.params.ny
        self.shape = (self.ny + 1, self.nx + 1)
        self.np = 0
        self._owner = _handle is None
        if _handle is not None:
            # a simulation of a Batch: the batch owns the handle
            self._h = _handle
            return
        self._h = ctypes.c_void_p()
        rc = self.lib.lbm_create(ctypes.byref(self.params), ctypes.byref(self._h))
        if rc != 0:
            msg = self.lib.lbm_last_error(self._h if self._h else None).decode()
            if self._h:
                self.lib.lbm_destroy(self._h)
                self._h = None
            raise LBMError(rc, msg)
        self.np = 0

    def close(self):
        if getattr(self, "_h", None):
            if self._owner:
                self.lib.lbm_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _call(self, name, *args):
        rc = getattr(self.lib, name)(self._h, *args)
        if rc != 0:
            raise LBMError(rc, self.lib.lbm_last_error(self._h).decode())
        return rc

    # ------------------------------------------------------------ state
    def initialize(self):
        """INITIALIZE (F6:563-580)."""
        self._call("lbm_initialize")

    @property
    def f(self):
        out = np.empty(self.shape + (9,))
        self._call("lbm_get_f", _ptr(out))
        return out

    @f.setter
    def f(self, value):
        a = np.ascontiguousarray(value, dtype=np.float64).reshape(self.shape + (9,))
        self._call("lbm_set_f", _ptr(a))

    @property
    def fpost(self):
        out = np.empty(self.shape + (9,))
        self._call("lbm_get_fpost", _ptr(out))
        return out

    @fpost.setter
    def fpost(self, value):
        a = np.ascontiguousarray(value, dtype=np.float64).reshape(self.shape + (9,))
        self._call("lbm_set_fpost", _ptr(a))

    def get_fields(self):
        """(RHO, UX, UY, G, OMEGA_BN)"""
        rho, ux, uy, om = (np.empty(self.shape) for _ in range(4))
        g = np.empty(self.shape + (2,))
        self._call("lbm_get_fields", _ptr(rho), _ptr(ux), _ptr(uy), _ptr(g), _ptr(om))
        return rho, ux, uy, g, om

    def set_fields(self, rho=None, ux=None, uy=None, g=None, omega_bn=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
                for a in (rho, ux, uy, g, omega_bn)]
        self._call("lbm_set_fields", *[_ptr(a) for a in arrs])

    @property
    def stress(self):
        out = np.empty(self.shape + (3,))
        self._call("lbm_get_stress", _ptr(out))
        return out

    @property
    def image(self):
        out = np.empty(self.shape, dtype=np.int32)
        self._call("lbm_get_image", _ptr(out, ctypes.c_int32))
        return out

    def set_particles(self, part):
        """part: (PF_COUNT, n) array of particle records (LBM_PF_* order)."""
        part = np.ascontiguousarray(part, dtype=np.float64)
        assert part.shape[0] == _abi.PF_COUNT
        n = part.shape[1]
        self._call("lbm_set_particles", n, _ptr(part))
        self.np = n

    def get_particles(self):
        out = np.empty((_abi.PF_COUNT, self.np))
        if self.np:
            self._call("lbm_get_particles", _ptr(out))
        return out

    def get_phi(self, ip):
        out = np.empty(self.shape)
        self._call("lbm_get_phi", ip, _ptr(out))
        return out

    # ------------------------------------------------- checkpoint / restart
    def write_output(self, directory):
        """WRITE_OUTPUT (F6:377-405): the reference's text state files."""
        self._call("lbm_restart_write", os.fsencode(str(directory)))

    def read_input(self, directory):
        """READ_INPUT (F6:174-198); the particles must be set first."""
        self._call("lbm_restart_read", os.fsencode(str(directory)))

    def snapshot(self):
        """Binary snapshot of the whole run state (numpy uint8 array)."""
        n = ctypes.c_uint64(0)
        self._call("lbm_snapshot_size", ctypes.byref(n))
        buf = np.empty(n.value, dtype=np.uint8)
        self._call("lbm_snapshot_save", buf.ctypes.data_as(ctypes.c_void_p), n.value)
        return buf

    def load_snapshot(self, buf):
        """Continue the run saved by snapshot() (same parameters)."""
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        self._call("lbm_snapshot_load", buf.ctypes.data_as(ctypes.c_void_p), buf.size)
        # particle count: int32 after the 8-byte tag, version, sequence
        # position, nx, ny and maxp of the snapshot header
        self.np = int(buf[28:32].view(np.int32)[0])

    # ---------------------------------------------- reference subroutines
    def collision_mrt(self):
        """Collision_MRT (F6:867-962)."""
        self._call("lbm_collide")

    def copy_post(self):
        """F_POST = F (F6:329)."""
        self._call("lbm_copy_post")

    def streaming(self):
        """STREAMING (F6:1154-1190)."""
        self._call("lbm_stream")

    def boundarycondition(self, kk):
        """BOUNDARYCONDITION (F6:1195-1390)."""
        self._call("lbm_boundary", int(kk))

    def macroscopicvariables(self):
        """MACROSCOPICVARIABLES (F6:1394-1412)."""
        self._call("lbm_macros")

    def compute_concentration_function(self):
        """ComputeConcenterationFunction (F6:430-473)."""
        self._call("lbm_spm_phi")

    def compute_force_field_spm(self):
        """ComputeForceField_SPM (F6:477-530)."""
        self._call("lbm_spm_forcing")

    def compute_force_hydro(self):
        """ComputeForce_HYDRO (F6:534-558)."""
        self._call("lbm_hydro_force")

    def velocity_corrector(self):
        """VelocityCorrector (F6:1111-1125)."""
        self._call("lbm_correct")

    def compute_collision_force(self, max_pairs=0):
        """ComputeCollisionForce (F6:584-761, N-body F3:632-806).  Returns the
        in-range pairs (i<j, ascending) when max_pairs > 0."""
        n = ctypes.c_int32(0)
        if max_pairs:
            buf = np.zeros(2 * max_pairs, dtype=np.int32)
            self._call("lbm_contact_force", _ptr(buf, ctypes.c_int32), max_pairs,
                       ctypes.byref(n))
            return buf[: 2 * n.value].reshape(-1, 2)
        self._call("lbm_contact_force", None, 0, None)
        return None

    def particle_update(self, kk=0):
        """Particle update of loop iteration kk: inline in MAIN (F6:299-312),
        or the prescribed motion of F4:221-226."""
        self._call("lbm_particle_update", int(kk))

    def total_mass(self):
        """SUM(RHO) (F6:348)."""
        m = ctypes.c_double(0.0)
        self._call("lbm_total_mass", ctypes.byref(m))
        return m.value

    def step_parts(self, kk):
        """One time step through the per-subroutine entries, in the order of
        the reference's time loop (F6:293-345)."""
        self.compute_collision_force()
        self.particle_update(kk)
        self.compute_concentration_function()
        self.collision_mrt()
        self.copy_post()
        self.streaming()
        self.boundarycondition(kk)
        self.macroscopicvariables()
        self.compute_force_field_spm()
        self.compute_force_hydro()
        self.velocity_corrector()

    # ------------------------------------------------------------ fused
    def step(self, kk0, nsteps, record=True):
        """Fused time steps kk0+1..kk0+nsteps on the device.  Returns the
        per-step history (nsteps, np*12 + 2) when record."""
        if record:
            hist = np.zeros((nsteps, self.np * _abi.H_PCOUNT + 2))
            self._call("lbm_step", int(kk0), int(nsteps), _ptr(hist))
            return hist
        self._call("lbm_step", int(kk0), int(nsteps), None)
        return None

    def step_timed(self, kk0, nsteps):
        """Fused steps timed with HIP events: (ms over all launches of the call,
        ms over its nsteps back-to-back step launches)."""
        t, m = ctypes.c_double(0), ctypes.c_double(0)
        self._call("lbm_step_timed", int(kk0), int(nsteps), ctypes.byref(t), ctypes.byref(m))
        return t.value, m.value

    def adjoint_velocity(self, nsteps, window, checkpoint):
        """Loss = mean V_y of the particle over the last `window` of nsteps
        fused steps from the current state, and its gradient with respect to
        (tau_y, mu0, rho_p) by the discrete adjoint with a checkpoint every
        `checkpoint` steps.  Returns (loss, grad[3], device bytes used)."""
        loss = ctypes.c_double(0.0)
        mem = ctypes.c_double(0.0)
        grad = np.zeros(3)
        self._call("lbm_adjoint_velocity", int(nsteps), int(window), int(checkpoint),
                   ctypes.byref(loss), _ptr(grad), ctypes.byref(mem))
        return loss.value, grad, mem.value

    def last_pairs(self, max_pairs=4096):
        n = ctypes.c_int32(0)
        buf = np.zeros(2 * max_pairs, dtype=np.int32)
        self._call("lbm_last_pairs", _ptr(buf, ctypes.c_int32), max_pairs, ctypes.byref(n))
        return buf[: 2 * n.value].reshape(-1, 2)

    def sync(self):
        self._call("lbm_sync")


class Batch:
    """Independent simulations stepped together, one launch per time step
    (lbm_batch_* in include/lbm_amd.h).  cfgs: one lbm_params dict per
    simulation (same storage for all).  sims[i] is a D2Q9 view of simulation
    i for setting particles and reading state."""

    def __init__(self, cfgs):
        self.lib = _abi.load()
        n = len(cfgs)
        arr = (_abi.Params * n)(*[make_params(c) for c in cfgs])
        self._b = ctypes.c_void_p()
        rc = self.lib.lbm_batch_create(n, arr, ctypes.byref(self._b))
        if rc != 0:
            raise LBMError(rc, self.lib.lbm_last_error(None).decode())
        self.n = n
        self.sims = []
        for i, c in enumerate(cfgs):
            h = self.lib.lbm_batch_sim(self._b, i)
            self.sims.append(D2Q9(c, _handle=ctypes.c_void_p(h)))

    def close(self):
        if getattr(self, "_b", None):
            for s in self.sims:
                s.close()
            self.lib.lbm_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        self.close()

    def _call(self, name, *args):
        rc = getattr(self.lib, name)(self._b, *args)
        if rc != 0:
            raise LBMError(rc, self.lib.lbm_batch_last_error(self._b).decode())
        return rc

    def set_particles(self, i, part):
        self.sims[i].set_particles(part)

    def initialize(self):
        self._call("lbm_batch_initialize")

    def step(self, kk0, nsteps, record=True):
        """Fused steps of every simulation; with record, a list of per-
        simulation histories (nsteps, np_i*12 + 2)."""
        if not record:
            self._call("lbm_batch_step", int(kk0), int(nsteps), None)
            return None
        widths = [s.np * _abi.H_PCOUNT + 2 for s in self.sims]
        buf = np.zeros(nsteps * sum(widths))
        self._call("lbm_batch_step", int(kk0), int(nsteps), _ptr(buf))
        out, o = [], 0
        for w in widths:
            out.append(buf[o:o + nsteps * w].reshape(nsteps, w))
            o += nsteps * w
        return out

    def step_timed(self, kk0, nsteps):
        t = ctypes.c_double(0)
        self._call("lbm_batch_step_timed", int(kk0), int(nsteps), ctypes.byref(t))
        return t.value

    def set_graph(self, chunk):
        self._call("lbm_batch_set_graph", int(chunk))

    @property
    def tiles(self):
        t = np.zeros(self.n, dtype=np.int32)
        self._call("lbm_batch_get_tiles", _ptr(t, ctypes.c_int32))
        return t

    @tiles.setter
    def tiles(self, values):
        t = np.ascontiguousarray(values, dtype=np.int32)
        self._call("lbm_batch_set_tiles", _ptr(t, ctypes.c_int32))

    def tune(self, steps=10):
        t = np.zeros(self.n, dtype=np.int32)
        self._call("lbm_batch_tune", int(steps), _ptr(t, ctypes.c_int32))
        return t

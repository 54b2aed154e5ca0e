"""ctypes binding of liblfdem_amd.so (the gfx950 library) for tests and benchmarks.

The classes mirror LF_DEM's own interfaces: ``StokesSolver`` keeps the method
names and call protocol of StokesSolver.h:198-285; ``System`` exposes the
device-resident step loop of System.cpp:888-983.  There is no CPU fallback:
if the library is missing or no GPU is present the calls fail loudly.
"""
import atexit
import ctypes
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "liblfdem_amd.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_lib = None

STATUS = {0: "LFD_OK", -1: "LFD_ERR_ARG", -2: "LFD_ERR_HIP", -3: "LFD_ERR_NOT_SPD",
          -4: "LFD_ERR_NOT_CONVERGED", -5: "LFD_ERR_CAPACITY", -6: "LFD_ERR_BOX",
          -7: "LFD_ERR_UNSUPPORTED"}

# every function of include/lfdem_amd.h with its argument types
SIGNATURES = {
    "lfd_last_error": ([], ctypes.c_char_p),
    "lfd_default_params": ([_vp], None),
    "lfd_default_solve_opts": ([_vp], None),
    "lfd_device_count": ([], ctypes.c_int),
    "lfd_stokes_create": ([ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
    "lfd_stokes_destroy": ([_vp], None),
    "lfd_stokes_set_opts": ([_vp, _vp], ctypes.c_int),
    "lfd_stokes_reset": ([_vp, ctypes.c_int, _dp, ctypes.c_int], ctypes.c_int),
    "lfd_stokes_start_new_column": ([_vp], ctypes.c_int),
    "lfd_stokes_add_blocks": ([_vp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp], ctypes.c_int),
    "lfd_stokes_fill_done": ([_vp], ctypes.c_int),
    "lfd_stokes_solve_noalloc": ([ctypes.c_int, ctypes.c_int] + [_vp] * 9, ctypes.c_int),
    "lfd_stokes_reset_rhs": ([_vp], ctypes.c_int),
    "lfd_stokes_reset_rhs_torque": ([_vp], ctypes.c_int),
    "lfd_stokes_add_rhs_force": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
    "lfd_stokes_add_rhs_torque": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
    "lfd_stokes_set_rhs_force": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
    "lfd_stokes_set_rhs_torque": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
    "lfd_stokes_set_rhs": ([_vp, _dp], ctypes.c_int),
    "lfd_stokes_add_rhs": ([_vp, _dp], ctypes.c_int),
    "lfd_stokes_solve": ([_vp, _dp, ctypes.c_int, _vp], ctypes.c_int),
    "lfd_stokes_solving_done": ([_vp], ctypes.c_int),
    "lfd_stokes_multiply": ([_vp, _dp, _dp], ctypes.c_int),
    "lfd_system_create": ([ctypes.c_int, _dp, _dp, _dp, _dp, _vp, _vp, ctypes.POINTER(_vp)], ctypes.c_int),
    "lfd_system_destroy": ([_vp], None),
    "lfd_system_first_update": ([_vp], ctypes.c_int),
    "lfd_system_step": ([_vp, ctypes.c_int, ctypes.c_double], ctypes.c_int),
    "lfd_system_step_graph": ([_vp, ctypes.c_int, ctypes.c_double], ctypes.c_int),
    "lfd_system_graphs": ([_vp], ctypes.c_int),
    "lfd_system_sync": ([_vp], ctypes.c_int),
    "lfd_system_check_new_interactions": ([_vp], ctypes.c_int),
    "lfd_system_update_interactions": ([_vp, ctypes.c_int, ctypes.c_double], ctypes.c_int),
    "lfd_system_assemble": ([_vp], ctypes.c_int),
    "lfd_system_compute_velocities": ([_vp, ctypes.c_int], ctypes.c_int),
    "lfd_system_num_pairs": ([_vp], ctypes.c_int),
    "lfd_system_get_pairs": ([_vp, ctypes.c_int, _ip, _ip, _ip, _dp, _dp], ctypes.c_int),
    "lfd_system_get_contacts": ([_vp, ctypes.c_int, _ip, _ip, _dp, _dp], ctypes.c_int),
    "lfd_system_set_contacts": ([_vp, ctypes.c_int, _ip, _ip, _dp, _dp], ctypes.c_int),
    "lfd_system_set_clock": ([_vp, ctypes.c_double, ctypes.c_double], ctypes.c_int),
    "lfd_system_get_state": ([_vp, _dp, _dp, _dp, _dp], ctypes.c_int),
    "lfd_system_get_pair_output": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
    "lfd_system_get_na_disp": ([_vp, _dp], ctypes.c_int),
    "lfd_system_get_cells": ([_vp, _ip, _ip], ctypes.c_int),
    "lfd_system_get_blocks": ([_vp, ctypes.c_int, _dp, _dp, _ip, _ip], ctypes.c_int),
    "lfd_system_get_rhs": ([_vp, _dp], ctypes.c_int),
    "lfd_system_get_solve_stats": ([_vp, _vp, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
    "lfd_system_precond_waves": ([_vp, _ip], ctypes.c_int),
    "lfd_batch_precond_waves": ([_vp, ctypes.c_int, _ip], ctypes.c_int),
    "lfd_system_solver_slots": ([_vp, _ip], ctypes.c_int),
    "lfd_batch_solver_slots": ([_vp, ctypes.c_int, _ip], ctypes.c_int),
    "lfd_system_set_opts": ([_vp, _vp], ctypes.c_int),
    "lfd_system_profile": ([_vp, ctypes.c_int, _dp, _ip], ctypes.c_int),
    "lfd_system_set_control": ([_vp, ctypes.c_int, ctypes.c_double], ctypes.c_int),
    "lfd_system_set_shear_rate_limit": ([_vp, ctypes.c_double, ctypes.c_double], ctypes.c_int),
    "lfd_system_step_stress": ([_vp, ctypes.c_double], ctypes.c_int),
    "lfd_system_time_evolution": ([_vp, ctypes.c_double, _ip], ctypes.c_int),
    "lfd_system_compute_stress": ([_vp, ctypes.c_int], ctypes.c_int),
    "lfd_system_get_stress": ([_vp, _dp, _dp, _dp, _dp, _dp], ctypes.c_int),
    "lfd_system_get_parts": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
    "lfd_batch_create": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, ctypes.POINTER(_vp)],
                         ctypes.c_int),
    "lfd_batch_destroy": ([_vp], None),
    "lfd_batch_footprint": ([ctypes.c_int, _dp, _dp, _vp, _ip, _ip, ctypes.POINTER(ctypes.c_longlong)],
                            ctypes.c_int),
    "lfd_batch_room": ([_vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "lfd_batch_add": ([_vp, ctypes.c_int, _dp, _dp, _dp, _dp, _vp, ctypes.c_double, ctypes.c_double, _ip],
                      ctypes.c_int),
    "lfd_batch_remove": ([_vp, ctypes.c_int], ctypes.c_int),
    "lfd_batch_step": ([_vp, ctypes.c_int], ctypes.c_int),
    "lfd_batch_step_graph": ([_vp, ctypes.c_int], ctypes.c_int),
    "lfd_batch_graphs": ([_vp], ctypes.c_int),
    "lfd_batch_sync": ([_vp], ctypes.c_int),
    "lfd_batch_query": ([_vp, ctypes.c_int, _dp, _vp, _ip], ctypes.c_int),
    "lfd_batch_status": ([_vp, ctypes.c_int, _ip, _dp, _ip], ctypes.c_int),
    "lfd_batch_get_state": ([_vp, ctypes.c_int, _dp, _dp, _dp, _dp], ctypes.c_int),
    "lfd_batch_get_pair_output": ([_vp, ctypes.c_int, ctypes.c_int, _dp], ctypes.c_int),
    "lfd_batch_get_na_disp": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
    "lfd_batch_get_pairs": ([_vp, ctypes.c_int, ctypes.c_int, _ip, _ip, _ip, _dp, _dp], ctypes.c_int),
    "lfd_batch_num_pairs": ([_vp, ctypes.c_int], ctypes.c_int),
    "lfd_batch_info": ([_vp, _ip, _ip, _ip, _ip, _ip, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
    "lfd_batch_profile": ([_vp, ctypes.c_int, _dp, _ip], ctypes.c_int),
}


class Params(ctypes.Structure):
    """lfd_params (include/lfdem_amd.h)."""
    _fields_ = [(n, ctypes.c_double) for n in (
        "shear_rate", "theta_shear", "lub_max_gap", "lub_reduce", "kn", "kt",
        "mu_static", "mu_dynamic", "tau_n", "tau_t", "sd_coeff", "repulsion",
        "repulsive_length", "disp_max", "dt_max", "dt_min")] + [
        ("friction_model", ctypes.c_int), ("lubrication_model", ctypes.c_int), ("fixed_dt", ctypes.c_int)]


class SolveOpts(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("inner_tol", ctypes.c_double),
                ("max_outer", ctypes.c_int), ("max_inner", ctypes.c_int),
                ("lanes_per_row", ctypes.c_int), ("groups", ctypes.c_int), ("kind", ctypes.c_int),
                ("precond", ctypes.c_int), ("regroup_interval", ctypes.c_int), ("launch", ctypes.c_int),
                ("max_rhs", ctypes.c_int), ("agg_waves", ctypes.c_int)]


class SolveStats(ctypes.Structure):
    _fields_ = [("iters_inner", ctypes.c_int), ("iters_outer", ctypes.c_int),
                ("res_norm", ctypes.c_double), ("rhs_norm", ctypes.c_double)]


class LfdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (STATUS.get(code, "?"), code, msg))
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("liblfdem_amd.so is not built (run __graft_entry__.build()): %s" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


# Handles still open when the interpreter exits are destroyed by an exit hook
# of this module, while the HIP runtime is still up (before the runtime's own
# teardown and any profiler's finalisation run), newest first.
_LIVE = weakref.WeakValueDictionary()
_SERIAL = [0]


def _track(obj):
    _SERIAL[0] += 1
    _LIVE[_SERIAL[0]] = obj


@atexit.register
def _release_all():
    for key in sorted(list(_LIVE.keys()), reverse=True):
        obj = _LIVE.get(key)
        if obj is not None:
            obj.close()


def _check(rc):
    if rc < 0:
        raise LfdError(rc, lib().lfd_last_error().decode())
    return rc


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def default_params(**kw):
    p = Params()
    lib().lfd_default_params(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def material_functions(total, contact, shear_rate):
    """Normal-stress viscosities and particle pressures of a stress tensor
    (xx, xy, xz, yz, yy, zz) in internal units, as Simulation::outputData
    forms them (Simulation.cpp:694-699): N1 = (s_xx - s_zz) / rate,
    N2 = (s_zz - s_yy) / rate, pressure = -trace / 3 (also of the contact
    group).  LF_DEM prints viscosities times 6 pi and stresses times 6 pi
    times the internal / output force ratio (OutputData.h:118-127)."""
    t, c = np.asarray(total, dtype=float), np.asarray(contact, dtype=float)
    out = dict(particle_pressure=-(t[0]+t[4]+t[5])/3, particle_pressure_contact=-(c[0]+c[4]+c[5])/3)
    if shear_rate != 0:
        out["N1_viscosity"] = (t[0]-t[5])/shear_rate
        out["N2_viscosity"] = (t[5]-t[4])/shear_rate
    else:
        out["N1_viscosity"] = out["N2_viscosity"] = float("nan")
    return out


def solve_noalloc(np_, nblocks, dblocks, p0, p1, odblocks, rhs, x, opts=None, stream=None):
    """lfd_stokes_solve_noalloc on device arrays given by their addresses
    (integers, e.g. a device tensor's data_ptr()); returns the solve stats."""
    st = SolveStats()
    _check(lib().lfd_stokes_solve_noalloc(int(np_), int(nblocks), dblocks, p0, p1, odblocks, rhs, x,
                                          ctypes.byref(opts) if opts is not None else None, stream,
                                          ctypes.byref(st)))
    return {"iters_inner": st.iters_inner, "iters_outer": st.iters_outer, "res_norm": st.res_norm,
            "rhs_norm": st.rhs_norm}


def default_opts(**kw):
    o = SolveOpts()
    lib().lfd_default_solve_opts(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class StokesSolver:
    """GPU drop-in for LF_DEM's StokesSolver (StokesSolver.h:198-285)."""

    def __init__(self, np_total, np_mobile=None, opts=None):
        h = _vp()
        _check(lib().lfd_stokes_create(np_total, np_total if np_mobile is None else np_mobile,
                                       ctypes.byref(h)))
        self.h = h
        _track(self)
        self.np = np_total
        if opts is not None:
            _check(lib().lfd_stokes_set_opts(self.h, ctypes.byref(opts)))
        self.last_stats = SolveStats()

    def close(self):
        """Releases the device memory and stream (also done on deletion and at exit)."""
        if getattr(self, "h", None):
            lib().lfd_stokes_destroy(self.h)
            self.h = None

    __del__ = close

    def resetResistanceMatrix(self, nb_blocks_hint, drag_dblocks, pattern_changed=True):
        d = np.ascontiguousarray(drag_dblocks, dtype=np.float64)
        _check(lib().lfd_stokes_reset(self.h, int(nb_blocks_hint), _d(d), int(pattern_changed)))

    def startNewColumn(self):
        _check(lib().lfd_stokes_start_new_column(self.h))

    def addResistanceBlocks(self, i, j, dbi, dbj, odb):
        a = np.ascontiguousarray(dbi, dtype=np.float64)
        b = np.ascontiguousarray(dbj, dtype=np.float64)
        c = np.ascontiguousarray(odb, dtype=np.float64)
        _check(lib().lfd_stokes_add_blocks(self.h, int(i), int(j), _d(a), _d(b), _d(c)))

    def matrixFillingDone(self):
        _check(lib().lfd_stokes_fill_done(self.h))

    def resetRHS(self):
        _check(lib().lfd_stokes_reset_rhs(self.h))

    def resetRHStorque(self):
        _check(lib().lfd_stokes_reset_rhs_torque(self.h))

    def addToRHSForce(self, i, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        _check(lib().lfd_stokes_add_rhs_force(self.h, int(i), _d(f)))

    def addToRHSTorque(self, i, t):
        t = np.ascontiguousarray(t, dtype=np.float64)
        _check(lib().lfd_stokes_add_rhs_torque(self.h, int(i), _d(t)))

    def setRHSForce(self, i, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        _check(lib().lfd_stokes_set_rhs_force(self.h, int(i), _d(f)))

    def setRHSTorque(self, i, t):
        t = np.ascontiguousarray(t, dtype=np.float64)
        _check(lib().lfd_stokes_set_rhs_torque(self.h, int(i), _d(t)))

    def setRHS(self, ft):
        ft = np.ascontiguousarray(ft, dtype=np.float64).ravel()
        _check(lib().lfd_stokes_set_rhs(self.h, _d(ft)))

    def addToRHS(self, ft):
        ft = np.ascontiguousarray(ft, dtype=np.float64).ravel()
        _check(lib().lfd_stokes_add_rhs(self.h, _d(ft)))

    def solve(self, out_len=None):
        """Returns (velocity (np,3), ang_velocity (np,3))."""
        n = 6 * self.np if out_len is None else out_len
        uw = np.zeros(max(n, 1))
        _check(lib().lfd_stokes_solve(self.h, _d(uw), int(n), ctypes.byref(self.last_stats)))
        u = uw[:6 * self.np].reshape(self.np, 6)
        return u[:, :3].copy(), u[:, 3:].copy()

    def solvingIsDone(self):
        _check(lib().lfd_stokes_solving_done(self.h))

    def multiply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        y = np.zeros_like(x)
        _check(lib().lfd_stokes_multiply(self.h, _d(x), _d(y)))
        return y


class System:
    """Device-resident LF_DEM System (setupConfiguration + step loop)."""

    def __init__(self, pos, radius, L, shear_disp=(0.0, 0.0, 0.0), params=None, opts=None):
        self.params = params if params is not None else default_params()
        self.opts = opts if opts is not None else default_opts()
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        rad = np.ascontiguousarray(radius, dtype=np.float64)
        self.np = len(rad)
        self.radius = rad
        self.L = np.ascontiguousarray(L, dtype=np.float64)
        sd = np.ascontiguousarray(shear_disp, dtype=np.float64)
        h = _vp()
        _check(lib().lfd_system_create(self.np, _d(pos), _d(rad), _d(self.L), _d(sd),
                                       ctypes.byref(self.params), ctypes.byref(self.opts),
                                       ctypes.byref(h)))
        self.h = h
        _track(self)

    def close(self):
        """Releases the device memory and stream (also done on deletion and at exit)."""
        if getattr(self, "h", None):
            lib().lfd_system_destroy(self.h)
            self.h = None

    __del__ = close

    def set_opts(self, opts):
        self.opts = opts
        _check(lib().lfd_system_set_opts(self.h, ctypes.byref(opts)))

    def first_update(self):
        _check(lib().lfd_system_first_update(self.h))

    def step(self, nsteps=1, strain_end=-1.0, graph=False):
        """nsteps predictor-corrector steps; graph=True replays a captured step."""
        fn = lib().lfd_system_step_graph if graph else lib().lfd_system_step
        _check(fn(self.h, int(nsteps), float(strain_end)))

    def graphs(self):
        """Step graphs held (0-2)."""
        return lib().lfd_system_graphs(self.h)

    def sync(self):
        _check(lib().lfd_system_sync(self.h))

    def check_new_interactions(self):
        _check(lib().lfd_system_check_new_interactions(self.h))

    def update_interactions(self, in_predictor, dt):
        _check(lib().lfd_system_update_interactions(self.h, int(in_predictor), float(dt)))

    def assemble(self):
        _check(lib().lfd_system_assemble(self.h))

    def compute_velocities(self, in_predictor=True):
        _check(lib().lfd_system_compute_velocities(self.h, int(in_predictor)))

    @property
    def npairs(self):
        return _check(lib().lfd_system_num_pairs(self.h))

    def pairs(self):
        n = self.npairs
        cap = max(n, 1)
        p0 = np.zeros(cap, dtype=np.int32)
        p1 = np.zeros(cap, dtype=np.int32)
        st = np.zeros(cap, dtype=np.int32)
        geom = np.zeros((cap, 5))
        dt = np.zeros((cap, 3))
        n = _check(lib().lfd_system_get_pairs(self.h, cap, _i(p0), _i(p1), _i(st), _d(geom), _d(dt)))
        return dict(p0=p0[:n], p1=p1[:n], state=st[:n], nvec=geom[:n, :3], r=geom[:n, 3],
                    gap=geom[:n, 4], disp_tan=dt[:n])

    # columns of pair_output() (lfd_system_get_pair_output)
    PAIR_OUT = ("lub_normal", "lub_tan_x", "lub_tan_y", "lub_tan_z", "contact_normal", "contact_tan_x",
                "contact_tan_y", "contact_tan_z", "repulsion", "contact_xF_viscosity")

    def pair_output(self):
        """(P, 10) force columns of the interaction snapshot, in pairs() order."""
        cap = max(self.npairs, 1)
        out = np.zeros((cap, len(self.PAIR_OUT)))
        n = _check(lib().lfd_system_get_pair_output(self.h, cap, _d(out)))
        return out[:n].copy()

    def na_disp(self):
        """(N, 3) accumulated non-affine displacement (System::getNonAffineDisp)."""
        out = np.zeros((self.np, 3))
        _check(lib().lfd_system_get_na_disp(self.h, _d(out)))
        return out

    def contacts(self):
        """Active contacts (System::getContacts): p0, p1, disp_tan (n, 3), disp_rolling (n, 3)."""
        n = _check(lib().lfd_system_get_contacts(self.h, 0, None, None, None, None))
        cap = max(n, 1)
        p0 = np.zeros(cap, dtype=np.int32)
        p1 = np.zeros(cap, dtype=np.int32)
        dt = np.zeros((cap, 3))
        dr = np.zeros((cap, 3))
        n = _check(lib().lfd_system_get_contacts(self.h, cap, _i(p0), _i(p1), _d(dt), _d(dr)))
        return dict(p0=p0[:n], p1=p1[:n], disp_tan=dt[:n], disp_rolling=dr[:n])

    def set_contacts(self, p0, p1, disp_tan, disp_rolling=None):
        """System::setContacts, after first_update."""
        p0 = np.ascontiguousarray(p0, dtype=np.int32)
        p1 = np.ascontiguousarray(p1, dtype=np.int32)
        dt = np.ascontiguousarray(disp_tan, dtype=np.float64).reshape(-1, 3)
        dr = None if disp_rolling is None else np.ascontiguousarray(disp_rolling, dtype=np.float64).reshape(-1, 3)
        _check(lib().lfd_system_set_contacts(self.h, len(p0), _i(p0), _i(p1), _d(dt), None if dr is None else _d(dr)))

    def set_clock(self, time, strain):
        _check(lib().lfd_system_set_clock(self.h, float(time), float(strain)))

    def state(self):
        pos = np.zeros((self.np, 3))
        vel6 = np.zeros((self.np, 6))
        na6 = np.zeros((self.np, 6))
        sc = np.zeros(8)
        _check(lib().lfd_system_get_state(self.h, _d(pos), _d(vel6), _d(na6), _d(sc)))
        return dict(pos=pos, vel=vel6[:, :3].copy(), ang_vel=vel6[:, 3:].copy(), na_vel=na6[:, :3].copy(),
                    na_ang_vel=na6[:, 3:].copy(), shear_disp=sc[:3].copy(), dt=sc[3], time=sc[4],
                    strain=sc[5], max_velocity=sc[6], max_sliding_velocity=sc[7])

    def cells(self):
        dims = np.zeros(4, dtype=np.int32)
        labels = np.zeros(self.np, dtype=np.int32)
        nbox = _check(lib().lfd_system_get_cells(self.h, _i(dims), _i(labels)))
        return dict(nx=int(dims[0]), ny=int(dims[1]), nz=int(dims[2]), boxed=bool(dims[3]), nbox=nbox,
                    labels=labels)

    def blocks(self):
        n = self.npairs
        cap = max(n, 1)
        db = np.zeros((self.np, 15))
        od = np.zeros((cap, 18))
        p0 = np.zeros(cap, dtype=np.int32)
        p1 = np.zeros(cap, dtype=np.int32)
        n = _check(lib().lfd_system_get_blocks(self.h, cap, _d(db), _d(od), _i(p0), _i(p1)))
        return db, od[:n], p0[:n], p1[:n]

    def rhs(self):
        r = np.zeros(6 * self.np)
        _check(lib().lfd_system_get_rhs(self.h, _d(r)))
        return r

    def profile(self, enable):
        """(summed ms, launches) of solver launches since the last call; then (dis)ables recording."""
        ms = ctypes.c_double(0)
        n = ctypes.c_int(0)
        _check(lib().lfd_system_profile(self.h, int(enable), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    # ---- stress and stress control
    GROUPS = ("contact", "dashpot", "hydro", "repulsion")
    PARTS = ("contact", "dashpot", "hydro", "repulsion")

    def set_control(self, control, target_stress=0.0):
        """control: "rate" or "stress" (target_stress in internal units)."""
        _check(lib().lfd_system_set_control(self.h, {"rate": 0, "stress": 1}[control], float(target_stress)))

    def set_shear_rate_limit(self, init_strain_limit, init_rate_limit):
        _check(lib().lfd_system_set_shear_rate_limit(self.h, float(init_strain_limit), float(init_rate_limit)))

    def step_stress(self, strain_end=-1.0):
        _check(lib().lfd_system_step_stress(self.h, float(strain_end)))

    def time_evolution(self, strain_end):
        n = ctypes.c_int(0)
        _check(lib().lfd_system_time_evolution(self.h, float(strain_end), ctypes.byref(n)))
        return n.value

    def compute_stress(self, in_predictor=True):
        _check(lib().lfd_system_compute_stress(self.h, int(in_predictor)))

    def stress(self, particle=False):
        g = np.zeros((4, 6))
        t = np.zeros(6)
        rate = ctypes.c_double(0)
        visc = ctypes.c_double(0)
        ps = np.zeros((self.np, 6)) if particle else None
        _check(lib().lfd_system_get_stress(self.h, _d(g), _d(t), ctypes.byref(rate), ctypes.byref(visc),
                                           _d(ps) if particle else None))
        out = dict(groups=dict(zip(self.GROUPS, g)), total=t, shear_rate=rate.value, viscosity=visc.value)
        out.update(material_functions(t, out["groups"]["contact"], rate.value))
        if particle:
            out["particle"] = ps
        return out

    def part(self, name):
        v = np.zeros((self.np, 6))
        _check(lib().lfd_system_get_parts(self.h, self.PARTS.index(name), _d(v)))
        return v

    def precond_waves(self):
        """Waves of the aggregate preconditioner (0 under 6x6 block-Jacobi)."""
        w = ctypes.c_int(0)
        _check(lib().lfd_system_precond_waves(self.h, ctypes.byref(w)))
        return w.value

    def solver_slots(self):
        """Solver slot of each particle under the aggregate preconditioner
        (wave = slot // 8); all -1 under 6x6 block-Jacobi."""
        out = np.zeros(self.np, dtype=np.int32)
        _check(lib().lfd_system_solver_slots(self.h, _i(out)))
        return out

    def solve_stats(self):
        st = SolveStats()
        tot = ctypes.c_longlong(0)
        _check(lib().lfd_system_get_solve_stats(self.h, ctypes.byref(st), ctypes.byref(tot)))
        return dict(iters_inner=st.iters_inner, iters_outer=st.iters_outer, res_norm=st.res_norm,
                    rhs_norm=st.rhs_norm, total_inner=tot.value)


def _state_dict(pos, vel6, na6, sc):
    return dict(pos=pos, vel=vel6[:, :3].copy(), ang_vel=vel6[:, 3:].copy(), na_vel=na6[:, :3].copy(),
                na_ang_vel=na6[:, 3:].copy(), shear_disp=sc[:3].copy(), dt=sc[3], time=sc[4],
                strain=sc[5], max_velocity=sc[6], max_sliding_velocity=sc[7])


def footprint(pos_or_np, radius, L, params=None):
    """(slots, cells, device bytes) an instance needs in a Batch."""
    rad = np.ascontiguousarray(radius, dtype=np.float64)
    Lv = np.ascontiguousarray(L, dtype=np.float64)
    p = params if params is not None else default_params()
    sl = ctypes.c_int(0)
    ce = ctypes.c_int(0)
    by = ctypes.c_longlong(0)
    _check(lib().lfd_batch_footprint(len(rad), _d(rad), _d(Lv), ctypes.byref(p), ctypes.byref(sl),
                                     ctypes.byref(ce), ctypes.byref(by)))
    return sl.value, ce.value, by.value


class Batch:
    """Many independent Systems stepped together (lfd_batch_* of include/lfdem_amd.h)."""

    def __init__(self, slot_cap, cell_cap, inst_cap, control="rate", opts=None):
        self.opts = opts if opts is not None else default_opts()
        self.control = control
        h = _vp()
        _check(lib().lfd_batch_create(int(slot_cap), int(cell_cap), int(inst_cap), {"rate": 0, "stress": 1}[control],
                                      ctypes.byref(self.opts), ctypes.byref(h)))
        self.h = h
        _track(self)
        self.np = {}

    def close(self):
        """Releases the device memory and stream (also done on deletion and at exit)."""
        if getattr(self, "h", None):
            lib().lfd_batch_destroy(self.h)
            self.h = None

    __del__ = close

    def room(self, slots, cells):
        return bool(_check(lib().lfd_batch_room(self.h, int(slots), int(cells))))

    def add(self, pos, radius, L, shear_disp=(0.0, 0.0, 0.0), params=None, strain_end=-1.0, target_stress=0.0):
        """Admit one instance; returns its id (LfdError with LFD_ERR_CAPACITY when it does not fit)."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        rad = np.ascontiguousarray(radius, dtype=np.float64)
        Lv = np.ascontiguousarray(L, dtype=np.float64)
        sd = np.ascontiguousarray(shear_disp, dtype=np.float64)
        p = params if params is not None else default_params()
        i = ctypes.c_int(-1)
        _check(lib().lfd_batch_add(self.h, len(rad), _d(pos), _d(rad), _d(Lv), _d(sd), ctypes.byref(p),
                                   float(strain_end), float(target_stress), ctypes.byref(i)))
        self.np[i.value] = len(rad)
        return i.value

    def remove(self, iid):
        _check(lib().lfd_batch_remove(self.h, int(iid)))
        self.np.pop(iid, None)

    def step(self, nsteps=1, graph=False):
        fn = lib().lfd_batch_step_graph if graph else lib().lfd_batch_step
        _check(fn(self.h, int(nsteps)))

    def graphs(self):
        return lib().lfd_batch_graphs(self.h)

    def sync(self):
        _check(lib().lfd_batch_sync(self.h))

    def query(self, iid):
        sc = np.zeros(8)
        st = SolveStats()
        status = ctypes.c_int(0)
        _check(lib().lfd_batch_query(self.h, int(iid), _d(sc), ctypes.byref(st), ctypes.byref(status)))
        return dict(shear_disp=sc[:3].copy(), dt=sc[3], time=sc[4], strain=sc[5], status=status.value,
                    iters_inner=st.iters_inner, iters_outer=st.iters_outer)

    def status(self):
        """(ids, strains, statuses) of every live instance, one device read."""
        cap = max(len(self.np), 1)
        ids = np.zeros(cap, dtype=np.int32)
        strain = np.zeros(cap)
        st = np.zeros(cap, dtype=np.int32)
        m = _check(lib().lfd_batch_status(self.h, cap, _i(ids), _d(strain), _i(st)))
        return ids[:m], strain[:m], st[:m]

    def state(self, iid):
        n = self.np[iid]
        pos = np.zeros((n, 3))
        vel6 = np.zeros((n, 6))
        na6 = np.zeros((n, 6))
        sc = np.zeros(8)
        _check(lib().lfd_batch_get_state(self.h, int(iid), _d(pos), _d(vel6), _d(na6), _d(sc)))
        return _state_dict(pos, vel6, na6, sc)

    def pair_output(self, iid):
        """System.pair_output of instance iid."""
        cap = max(_check(lib().lfd_batch_num_pairs(self.h, int(iid))), 1)
        out = np.zeros((cap, len(System.PAIR_OUT)))
        n = _check(lib().lfd_batch_get_pair_output(self.h, int(iid), cap, _d(out)))
        return out[:n].copy()

    def solver_slots(self, iid):
        """System.solver_slots of instance iid."""
        out = np.zeros(self.np[iid], dtype=np.int32)
        _check(lib().lfd_batch_solver_slots(self.h, int(iid), _i(out)))
        return out

    def na_disp(self, iid):
        """System.na_disp of instance iid."""
        out = np.zeros((self.np[iid], 3))
        _check(lib().lfd_batch_get_na_disp(self.h, int(iid), _d(out)))
        return out

    def pairs(self, iid):
        n = _check(lib().lfd_batch_num_pairs(self.h, int(iid)))
        cap = max(n, 1)
        p0 = np.zeros(cap, dtype=np.int32)
        p1 = np.zeros(cap, dtype=np.int32)
        st = np.zeros(cap, dtype=np.int32)
        geom = np.zeros((cap, 5))
        dt = np.zeros((cap, 3))
        n = _check(lib().lfd_batch_get_pairs(self.h, int(iid), cap, _i(p0), _i(p1), _i(st), _d(geom), _d(dt)))
        return dict(p0=p0[:n], p1=p1[:n], state=st[:n], nvec=geom[:n, :3], r=geom[:n, 3],
                    gap=geom[:n, 4], disp_tan=dt[:n])

    def info(self):
        v = [ctypes.c_int(0) for _ in range(5)]
        by = ctypes.c_longlong(0)
        _check(lib().lfd_batch_info(self.h, *[ctypes.byref(x) for x in v], ctypes.byref(by)))
        return dict(live=v[0].value, slot_cap=v[1].value, cell_cap=v[2].value, inst_cap=v[3].value,
                    solver_launches=v[4].value, bytes=by.value)

    def profile(self, enable):
        ms = ctypes.c_double(0)
        n = ctypes.c_int(0)
        _check(lib().lfd_batch_profile(self.h, int(enable), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

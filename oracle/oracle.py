"""ctypes view of the oracle library (liblfdem_oracle.so).

TEST INFRASTRUCTURE ONLY: the CPU FP64 restatement of LF_DEM's Stokes-resistance
step path, used as the checker for the HIP product path.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblfdem_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class Params(ctypes.Structure):
    """Mirror of orc_params (lfdem_oracle.h)."""
    _fields_ = [(n, ctypes.c_double) for n in (
        "shear_rate", "theta_shear", "lub_max_gap", "lub_reduce", "kn", "kt",
        "mu_static", "mu_dynamic", "tau_n", "tau_t", "sd_coeff", "repulsion",
        "repulsive_length", "disp_max", "dt_max", "dt_min")] + [
        ("friction_model", ctypes.c_int), ("lubrication_model", ctypes.c_int), ("fixed_dt", ctypes.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.POINTER(Params)]
        for name in ("orc_destroy",):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        for name in ("orc_setup", "orc_first_update", "orc_boxset_update",
                     "orc_check_new_interaction", "orc_update_interactions",
                     "orc_compute_velocities", "orc_num_pairs"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
            getattr(L, name).restype = ctypes.c_int
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_last_error.argtypes = [ctypes.c_void_p]
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_chol_factor.restype = ctypes.c_void_p
        L.orc_chol_factor.argtypes = [ctypes.c_int, _dp, _ip, _ip, _dp, _ip]
        L.orc_chol_solve.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_chol_free.argtypes = [ctypes.c_void_p]
        L.orc_chol_flops.argtypes = [ctypes.c_void_p]
        L.orc_chol_flops.restype = ctypes.c_double
        L.orc_chol_nnz_blocks.argtypes = [ctypes.c_void_p]
        L.orc_chol_nnz_blocks.restype = ctypes.c_long
        L.orc_csc_nnz.restype = ctypes.c_long
        L.orc_lub_range.restype = ctypes.c_double
        L.orc_lub_range.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_set_dt.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_set_in_predictor.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_get_pairs.argtypes = [ctypes.c_void_p, _ip, _ip, _ip, _dp, _dp]
        L.orc_set_pair_history.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _ip, _ip, _dp, _dp]
        L.orc_get_state.argtypes = [ctypes.c_void_p] + [_dp] * 6
        L.orc_set_positions.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_set_velocities.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_get_box_info.argtypes = [ctypes.c_void_p, _ip, _ip]
        L.orc_get_neighbor_cells.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, ctypes.c_int]
        L.orc_build_blocks.argtypes = [ctypes.c_void_p, _dp, _ip, _ip, _dp, ctypes.c_int]
        L.orc_build_rhs.argtypes = [ctypes.c_void_p, _dp]
        L.orc_solve_current.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_periodize_diff.argtypes = [_dp, _dp, _dp]
        L.orc_periodize.argtypes = [_dp, _dp, _dp]
        L.orc_bruteforce_pairs.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_double,
                                           _ip, _ip, ctypes.c_int]
        L.orc_export_csc.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _ip, _ip, _dp, _ip, _ip, _dp]
        L.orc_csc_nnz.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_build_rhs_component.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_set_control.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        L.orc_set_shear_rate_limit.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        for name in ("orc_compute_velocities_components", "orc_calc_stress"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
            getattr(L, name).restype = ctypes.c_int
        L.orc_get_stress.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_get_components.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        for name in ("orc_shear_rate", "orc_viscosity"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
            getattr(L, name).restype = ctypes.c_double
        L.orc_gu_stress.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_step_ex.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.orc_time_evolution.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_get_na_disp.argtypes = [ctypes.c_void_p, _dp]
        L.orc_pair_output.argtypes = [ctypes.c_void_p, _dp]
        L.orc_pair_output.restype = ctypes.c_int
        _lib = L
    return _lib


def P(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype).ctypes.data_as(_dp if dtype == np.float64 else _ip)


def default_params(**kw):
    p = Params()
    lib().orc_default_params(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def periodize_diff(L, shear_disp, d):
    d = np.array(d, dtype=np.float64)
    z = lib().orc_periodize_diff(P(L), P(shear_disp), d.ctypes.data_as(_dp))
    return z, d


def lub_constants(a0, a1):
    out = np.zeros(36)
    lib().orc_lub_constants(ctypes.c_double(a0), ctypes.c_double(a1), out.ctypes.data_as(_dp))
    return out.reshape(9, 4)


def lub_blocks(a0, a1, gap, nvec, params):
    db0, db1, od = np.zeros(15), np.zeros(15), np.zeros(18)
    nv = np.ascontiguousarray(nvec, dtype=np.float64)
    lib().orc_lub_blocks(ctypes.c_double(a0), ctypes.c_double(a1), ctypes.c_double(gap),
                         nv.ctypes.data_as(_dp), ctypes.byref(params),
                         db0.ctypes.data_as(_dp), db1.ctypes.data_as(_dp), od.ctypes.data_as(_dp))
    return db0, db1, od


def dashpot_blocks(a0, a1, nvec, params):
    db0, db1, od = np.zeros(15), np.zeros(15), np.zeros(18)
    nv = np.ascontiguousarray(nvec, dtype=np.float64)
    lib().orc_dashpot_blocks(ctypes.c_double(a0), ctypes.c_double(a1), nv.ctypes.data_as(_dp),
                             ctypes.byref(params), db0.ctypes.data_as(_dp),
                             db1.ctypes.data_as(_dp), od.ctypes.data_as(_dp))
    return db0, db1, od


def dblock_dense(b):
    b = np.ascontiguousarray(b, dtype=np.float64)
    M = np.zeros(36)
    lib().orc_dblock_dense(b.ctypes.data_as(_dp), M.ctypes.data_as(_dp))
    return M.reshape(6, 6)


def odblock_dense(b):
    b = np.ascontiguousarray(b, dtype=np.float64)
    M = np.zeros(36)
    lib().orc_odblock_dense(b.ctypes.data_as(_dp), M.ctypes.data_as(_dp))
    return M.reshape(6, 6)


def bruteforce_pairs(pos, radius, L, shear_disp, lub_max_gap=0.5):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    rad = np.ascontiguousarray(radius, dtype=np.float64)
    np_ = len(rad)
    cap = 64 * np_
    p0 = np.zeros(cap, dtype=np.int32)
    p1 = np.zeros(cap, dtype=np.int32)
    n = lib().orc_bruteforce_pairs(np_, pos.ctypes.data_as(_dp), rad.ctypes.data_as(_dp),
                                   P(L), P(shear_disp), ctypes.c_double(lub_max_gap),
                                   p0.ctypes.data_as(_ip), p1.ctypes.data_as(_ip), cap)
    assert n <= cap
    return p0[:n].copy(), p1[:n].copy()


def assemble_dense(np_, dblocks, od_cp, od_rows, odblocks):
    """Full symmetric 6N x 6N matrix from the column-protocol blocks (tests, N <= ~600)."""
    A = np.zeros((6 * np_, 6 * np_))
    for i in range(np_):
        A[6 * i:6 * i + 6, 6 * i:6 * i + 6] = dblock_dense(dblocks[i])
        for k in range(od_cp[i], od_cp[i + 1]):
            j = od_rows[k]
            M = odblock_dense(odblocks[k])
            A[6 * j:6 * j + 6, 6 * i:6 * i + 6] = M
            A[6 * i:6 * i + 6, 6 * j:6 * j + 6] = M.T
    return A


class Chol:
    """Sparse 6x6-block Cholesky of the oracle (stand-in for CHOLMOD)."""

    def __init__(self, np_, dblocks, od_cp, od_rows, odblocks):
        st = ctypes.c_int(0)
        self._db = np.ascontiguousarray(dblocks, dtype=np.float64)
        self._cp = np.ascontiguousarray(od_cp, dtype=np.int32)
        self._rw = np.ascontiguousarray(od_rows, dtype=np.int32)
        self._od = np.ascontiguousarray(odblocks, dtype=np.float64)
        self.np = np_
        self.h = lib().orc_chol_factor(np_, self._db.ctypes.data_as(_dp), self._cp.ctypes.data_as(_ip),
                                       self._rw.ctypes.data_as(_ip), self._od.ctypes.data_as(_dp),
                                       ctypes.byref(st))
        self.status = st.value

    def solve(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        lib().orc_chol_solve(self.h, b.ctypes.data_as(_dp), x.ctypes.data_as(_dp))
        return x

    @property
    def flops(self):
        return lib().orc_chol_flops(self.h)

    @property
    def nnz_blocks(self):
        return lib().orc_chol_nnz_blocks(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_chol_free(self.h)
            self.h = None


class System:
    """Oracle System: positions (N,3), radii (N,), box L (3,), LE shear_disp (3,)."""

    def __init__(self, pos, radius, L, shear_disp=(0, 0, 0), params=None, setup=True):
        self.params = params if params is not None else default_params()
        self.np = len(radius)
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        self._rad = np.ascontiguousarray(radius, dtype=np.float64)
        self.L = np.array(L, dtype=np.float64)
        self.h = lib().orc_create(self.np, pos.ctypes.data_as(_dp), self._rad.ctypes.data_as(_dp),
                                  P(self.L), P(shear_disp), ctypes.byref(self.params))
        if setup:
            self._check(lib().orc_setup(self.h))

    def _check(self, rc):
        if rc < 0:
            raise RuntimeError(lib().orc_last_error(self.h).decode())
        return rc

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def first_update(self):
        return self._check(lib().orc_first_update(self.h))

    def step(self, strain_end=-1.0):
        return self._check(lib().orc_step(self.h, ctypes.c_double(strain_end)))

    def boxset_update(self):
        return self._check(lib().orc_boxset_update(self.h))

    def check_new_interaction(self):
        return lib().orc_check_new_interaction(self.h)

    def update_interactions(self):
        return lib().orc_update_interactions(self.h)

    def compute_velocities(self):
        return lib().orc_compute_velocities(self.h)

    def set_dt(self, dt):
        lib().orc_set_dt(self.h, dt)

    # ---- components, stress, stress control
    COMPONENTS = ("contact", "dashpot", "hydro", "repulsion")
    GROUPS = ("contact", "dashpot", "hydro", "repulsion")

    def rhs_component(self, comp):
        out = np.zeros(6*self.np)
        rc = lib().orc_build_rhs_component(self.h, self.COMPONENTS.index(comp), out.ctypes.data_as(_dp))
        return None if rc < 0 else out.reshape(self.np, 6)

    def set_control(self, control, target_stress=0.0):
        """control: "rate" or "stress" (target_stress in internal units)."""
        lib().orc_set_control(self.h, {"rate": 0, "stress": 1}[control], ctypes.c_double(target_stress))

    def set_shear_rate_limit(self, init_strain_limit, init_rate_limit):
        lib().orc_set_shear_rate_limit(self.h, ctypes.c_double(init_strain_limit), ctypes.c_double(init_rate_limit))

    def compute_velocities_components(self):
        return lib().orc_compute_velocities_components(self.h)

    def calc_stress(self):
        return lib().orc_calc_stress(self.h)

    def stress(self):
        g = np.zeros((4, 6))
        t = np.zeros(6)
        ps = np.zeros((self.np, 6))
        lib().orc_get_stress(self.h, g.ctypes.data_as(_dp), t.ctypes.data_as(_dp), ps.ctypes.data_as(_dp))
        return dict(groups=dict(zip(self.GROUPS, g)), total=t, particle=ps, viscosity=self.viscosity,
                    shear_rate=self.shear_rate)

    def components(self, comp):
        v = np.zeros((self.np, 3))
        w = np.zeros((self.np, 3))
        rc = lib().orc_get_components(self.h, self.COMPONENTS.index(comp), v.ctypes.data_as(_dp),
                                      w.ctypes.data_as(_dp))
        return None if rc < 0 else (v, w)

    def gu_stress(self, vel, ang):
        v = np.ascontiguousarray(vel, dtype=np.float64)
        w = np.ascontiguousarray(ang, dtype=np.float64)
        S = np.zeros((self.np, 6))
        lib().orc_gu_stress(self.h, v.ctypes.data_as(_dp), w.ctypes.data_as(_dp), S.ctypes.data_as(_dp))
        return S

    @property
    def shear_rate(self):
        return lib().orc_shear_rate(self.h)

    @property
    def viscosity(self):
        return lib().orc_viscosity(self.h)

    def step_ex(self, strain_end=-1.0, calc_stress=False, adapt_first=True):
        return self._check(lib().orc_step_ex(self.h, ctypes.c_double(strain_end), int(calc_stress), int(adapt_first)))

    def time_evolution(self, strain_end):
        return self._check(lib().orc_time_evolution(self.h, ctypes.c_double(strain_end)))

    def set_in_predictor(self, f):
        lib().orc_set_in_predictor(self.h, int(f))

    def set_positions(self, pos, shear_disp=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        sd = None if shear_disp is None else P(shear_disp)
        return self._check(lib().orc_set_positions(self.h, pos.ctypes.data_as(_dp), sd))

    def set_velocities(self, vel, ang):
        return lib().orc_set_velocities(self.h, P(vel), P(ang))

    @property
    def npairs(self):
        return lib().orc_num_pairs(self.h)

    def pairs(self):
        n = self.npairs
        p0 = np.zeros(n, dtype=np.int32)
        p1 = np.zeros(n, dtype=np.int32)
        cs = np.zeros(n, dtype=np.int32)
        geom = np.zeros((n, 5))
        dt = np.zeros((n, 3))
        lib().orc_get_pairs(self.h, p0.ctypes.data_as(_ip), p1.ctypes.data_as(_ip),
                            cs.ctypes.data_as(_ip), geom.ctypes.data_as(_dp), dt.ctypes.data_as(_dp))
        return dict(p0=p0, p1=p1, state=cs, nvec=geom[:, :3], r=geom[:, 3], gap=geom[:, 4],
                    disp_tan=dt)

    # columns of pair_output() (orc_pair_output)
    PAIR_OUT = ("lub_normal", "lub_tan_x", "lub_tan_y", "lub_tan_z", "contact_normal", "contact_tan_x",
                "contact_tan_y", "contact_tan_z", "repulsion", "contact_xF_viscosity")

    def pair_output(self):
        """(P, 10) per-interaction force columns of outputIntFileTxt, pair order."""
        out = np.zeros((max(self.npairs, 1), len(self.PAIR_OUT)))
        n = lib().orc_pair_output(self.h, out.ctypes.data_as(_dp))
        return out[:n].copy()

    def na_disp(self):
        """(N, 3) accumulated non-affine displacement (System::getNonAffineDisp)."""
        out = np.zeros((self.np, 3))
        lib().orc_get_na_disp(self.h, out.ctypes.data_as(_dp))
        return out

    def state(self):
        n = self.np
        arrs = [np.zeros((n, 3)) for _ in range(5)]
        sc = np.zeros(8)
        lib().orc_get_state(self.h, *[a.ctypes.data_as(_dp) for a in arrs], sc.ctypes.data_as(_dp))
        keys = ("pos", "vel", "ang_vel", "na_vel", "na_ang_vel")
        d = dict(zip(keys, arrs))
        d.update(shear_disp=sc[:3].copy(), dt=sc[3], time=sc[4], strain=sc[5],
                 max_velocity=sc[6], max_sliding_velocity=sc[7])
        return d

    def box_info(self):
        dims = np.zeros(4, dtype=np.int32)
        labels = np.zeros(self.np, dtype=np.int32)
        nbox = lib().orc_get_box_info(self.h, dims.ctypes.data_as(_ip), labels.ctypes.data_as(_ip))
        return dict(nx=int(dims[0]), ny=int(dims[1]), nz=int(dims[2]), boxed=bool(dims[3]),
                    nbox=nbox, labels=labels)

    def neighbor_cells(self, cell):
        out = np.zeros(64, dtype=np.int32)
        n = lib().orc_get_neighbor_cells(self.h, cell, out.ctypes.data_as(_ip), 64)
        return out[:n].copy()

    def blocks(self):
        n = self.npairs
        db = np.zeros((self.np, 15))
        cp = np.zeros(self.np + 1, dtype=np.int32)
        rows = np.zeros(max(n, 1), dtype=np.int32)
        od = np.zeros((max(n, 1), 18))
        nod = lib().orc_build_blocks(self.h, db.ctypes.data_as(_dp), cp.ctypes.data_as(_ip),
                                     rows.ctypes.data_as(_ip), od.ctypes.data_as(_dp), max(n, 1))
        return db, cp, rows[:nod].copy(), od[:nod].copy()

    def rhs(self):
        r = np.zeros(6 * self.np)
        lib().orc_build_rhs(self.h, r.ctypes.data_as(_dp))
        return r

    def set_pair_history(self, p0, p1, state=None, disp_tan=None, prev_disp_tan=None):
        p0 = np.ascontiguousarray(p0, dtype=np.int32)
        p1 = np.ascontiguousarray(p1, dtype=np.int32)
        st = None if state is None else np.ascontiguousarray(state, dtype=np.int32).ctypes.data_as(_ip)
        dt = None if disp_tan is None else P(disp_tan)
        pd = None if prev_disp_tan is None else P(prev_disp_tan)
        return lib().orc_set_pair_history(self.h, len(p0), p0.ctypes.data_as(_ip),
                                          p1.ctypes.data_as(_ip), st, dt, pd)

"""ctypes mirror of include/imex.h and a thin numpy-facing wrapper.

The shared library ``libimex_amd.so`` is built in-tree (``make -C imex_amd``).
Loading fails loudly if it is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

GC = 3
NVAR = 8

BC_TRANSMISSIVE, BC_REFLECTIVE, BC_PERIODIC, BC_DIRICHLET, BC_JET_INFLOW = 0, 1, 2, 3, 4
RECON_MINMOD, RECON_VANLEER, RECON_WENO5 = 1, 2, 4
EOS_IDEAL, EOS_RK = 0, 1
CT_CONSISTENT, CT_PLAIN = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
OP_MAG, OP_P_CARRIER, OP_P_JACOBIAN, OP_ENERGY, OP_POTENTIAL = 0, 1, 2, 3, 4
FORMULA_ENERGY, FORMULA_PRESSURE = 0, 1
MAGFIELD_STAGGER, MAGFIELD_UNSTAGGER, MAGFIELD_POTENTIAL_A, MAGFIELD_NODIVFREE = 0, 1, 2, 3
RECON_NONE, RECON_ENO3 = 0, 3
RECONSTATE_PRIM, RECONSTATE_CONS = 0, 1
CASE_VORTEX, CASE_ORSZAG_TANG, CASE_BLAST, CASE_JET = 0, 1, 2, 3
SCHED_GRAPH, SCHED_BUDGET, SCHED_HOST, SCHED_AUTO = 0, 1, 2, 3
TUNE_EXPLICIT, TUNE_KRYLOV = 1, 2
WARN_GMRES_MAXOUTER, WARN_NEWTON_MAXITER, WARN_POSITIVITY, WARN_EOS_MAXITER = 1, 2, 4, 8
ERR_ARG, ERR_HIP, ERR_NONFINITE, ERR_UNSUPPORTED, ERR_NOMEM = -1, -2, -3, -4, -5


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int),
        ("x0", C.c_double), ("x1", C.c_double), ("y0", C.c_double), ("y1", C.c_double),
        ("gamma", C.c_double),
        ("bc_x", C.c_int), ("bc_y", C.c_int),
        ("recon", C.c_int), ("minmod_central_y", C.c_int),
        ("eos", C.c_int),
        ("rk_alfa", C.c_double), ("rk_b", C.c_double), ("rk_R", C.c_double), ("rk_cv", C.c_double),
        ("cfl", C.c_double),
        ("floor_on", C.c_int),
        ("p_floor", C.c_double), ("theta", C.c_double), ("mbig_fac", C.c_double),
        ("newton_max", C.c_int), ("newton_tol", C.c_double),
        ("gmres_tol", C.c_double), ("gmres_restart", C.c_int), ("gmres_max_outer", C.c_int),
        ("shock_kappa1", C.c_double), ("time_stop", C.c_double),
        ("ct_mode", C.c_int), ("precision", C.c_int),
        ("mixed_inner_tol", C.c_double), ("mixed_outer_tol", C.c_double),
        ("schedule", C.c_int), ("budget_margin", C.c_int), ("mixed_pressure", C.c_int),
        ("formula", C.c_int), ("recon_state", C.c_int), ("magfield", C.c_int),
        ("jet_rho", C.c_double), ("jet_v", C.c_double), ("jet_B0", C.c_double), ("jet_p", C.c_double),
        ("jet_halfwidth", C.c_double), ("rho_amb", C.c_double), ("p_amb", C.c_double),
        ("mixed_flux", C.c_int),
        ("adjoint_recompute", C.c_int),
    ]


class Scalars(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("t", "dt", "adt", "a_max", "B_max_x", "B_max_y", "vx_max", "vy_max",
                                          "v_max", "ch", "ch_x", "ch_y", "prev_dt")] + \
               [("timestep", C.c_int), ("rks", C.c_int)]


class StageHost(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in ("u_con", "Qe", "Qs", "bx", "by", "bx_qs", "by_qs", "fx", "fy")]


class Stats(C.Structure):
    _fields_ = [(n, C.c_longlong) for n in ("steps", "mag_solves", "mag_gmres_iters", "mag_gmres_restarts",
                                            "p_solves", "p_gmres_iters", "p_gmres_restarts", "newton_iters",
                                            "op_applies_mag", "op_applies_p", "mixed_refinements")] + \
               [("warnings", C.c_int), ("max_floored", C.c_int), ("min_p", C.c_double), ("div_b", C.c_double)] + \
               [(n, C.c_longlong) for n in ("budget_steps", "budget_reruns", "graph_builds")]


class KernelProfile(C.Structure):
    _fields_ = [("name", C.c_char_p), ("launches", C.c_longlong), ("total_ms", C.c_double), ("bytes", C.c_double)]


class TuneInfo(C.Structure):
    _fields_ = [("explicit_tile_x", C.c_int), ("explicit_tile_y", C.c_int), ("krylov_blocks", C.c_int),
                ("explicit_us", C.c_double * 5), ("krylov_us", C.c_double * 3),
                ("krylov_p_blocks", C.c_int), ("krylov_p_us", C.c_double * 4)]


class AdjointInfo(C.Structure):
    _fields_ = [("steps", C.c_int), ("checkpoints", C.c_int), ("recomputed_steps", C.c_int),
                ("transposed_solves", C.c_longlong), ("mag_iters", C.c_longlong), ("p_iters", C.c_longlong),
                ("worst_solve_residual", C.c_double), ("backward_ms", C.c_double)]


class Diag(C.Structure):
    _fields_ = [("timestep", C.c_int)] + [(n, C.c_double) for n in ("t", "dt", "div_error", "B_inf", "dt_comp", "emin",
                                                                      "emin_x", "emin_y", "mass", "intE", "magE_max")]


class SolveInfo(C.Structure):
    _fields_ = [("iters", C.c_int), ("restarts", C.c_int), ("res0", C.c_double), ("res", C.c_double),
                ("status", C.c_int)]


def struct_dict(s):
    return {n: getattr(s, n) for n, _ in s._fields_}


# ---------------------------------------------------------------- problem presets (the reference's built-in cases)
def default_params():
    p = Params()
    lib().imex_default_params(C.byref(p))
    return p


def vortex_params(n=64, m=None, amp=1.0, eos=EOS_IDEAL):
    """MHD vortex on periodic [-5,5]^2, gamma 5/3, Minmod with the vortex central y-slope, StaggerCT."""
    p = default_params()
    p.nx, p.ny = n, (m or n)
    p.x0, p.x1, p.y0, p.y1 = -5.0, 5.0, -5.0, 5.0
    p.gamma = 5.0 / 3.0
    p.bc_x = p.bc_y = BC_PERIODIC
    p.recon, p.minmod_central_y = RECON_MINMOD, 1
    p.eos = eos
    p.floor_on, p.p_floor = 1, 1e-7
    return p


def orszag_tang_params(n=64):
    """Orszag-Tang on periodic [0,2pi]^2, gamma 5/3, WENO5, p_floor 1e-8."""
    p = default_params()
    p.nx = p.ny = n
    p.x0, p.x1, p.y0, p.y1 = 0.0, 2 * np.pi, 0.0, 2 * np.pi
    p.gamma = 5.0 / 3.0
    p.bc_x = p.bc_y = BC_PERIODIC
    p.recon, p.minmod_central_y = RECON_WENO5, 0
    p.floor_on, p.p_floor = 1, 1e-8
    return p


def blast_params(n=64):
    """Magnetic blast on [-0.5,0.5]^2, outflow (transmissive), gamma 1.4, VanLeer, p_floor 1e-10."""
    p = default_params()
    p.nx = p.ny = n
    p.x0, p.x1, p.y0, p.y1 = -0.5, 0.5, -0.5, 0.5
    p.gamma = 1.4
    p.bc_x = p.bc_y = BC_TRANSMISSIVE
    p.recon, p.minmod_central_y = RECON_VANLEER, 0
    p.floor_on, p.p_floor = 1, 1e-10
    return p


def jet_params(n=64, m=None):
    """MHD jet (testMHDJet, main.cpp:214-223) on [-0.5,0.5]x[0,1.5]: outflow on the sides and the top, the jet
    inflow strip |x| < 0.05 on the bottom (Mach 800 ... 10000 family; here the defaults of config.hpp:257-260 with
    jet_rho = gamma), gamma 1.4, Minmod."""
    p = default_params()
    p.nx, p.ny = n, (m or (3 * n) // 2)
    p.x0, p.x1, p.y0, p.y1 = -0.5, 0.5, 0.0, 1.5
    p.gamma = 1.4
    p.bc_x, p.bc_y = BC_TRANSMISSIVE, BC_JET_INFLOW
    p.recon, p.minmod_central_y = RECON_MINMOD, 0
    p.jet_rho, p.jet_p = p.gamma, 1.0
    p.time_stop = 0.00015
    return p


def case_of(name):
    """(params, init kwargs) of a named workload."""
    if name.startswith("vortex"):
        n = int(name[6:] or 64)
        return vortex_params(n), dict(case=CASE_VORTEX, amp=1.0, u0=1.0, v0=1.0)
    if name.startswith("ot"):
        n = int(name[2:] or 64)
        return orszag_tang_params(n), dict(case=CASE_ORSZAG_TANG)
    if name.startswith("blast"):
        n = int(name[5:] or 64)
        return blast_params(n), dict(case=CASE_BLAST, amp=100.0)
    if name.startswith("jet"):
        n = int(name[3:] or 64)
        return jet_params(n), dict(case=CASE_JET)
    raise ValueError(name)


# ---------------------------------------------------------------- library
_LIB = None
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libimex_amd.so")


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `make -C imex_amd` (no CPU fallback exists)")
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    sig = {
        "imex_default_params": (C.c_int, [C.POINTER(Params)]),
        "imex_init_case": (C.c_int, [C.POINTER(Params), C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, dp, dp, dp]),
        "imex_create": (C.c_int, [C.POINTER(Params), C.POINTER(vp)]),
        "imex_destroy": (C.c_int, [vp]),
        "imex_last_error": (C.c_char_p, []),
        "imex_device_bytes": (C.c_longlong, [vp]),
        "imex_set_state": (C.c_int, [vp, dp, dp, dp, C.c_double, C.c_int, C.c_double]),
        "imex_get_state": (C.c_int, [vp, dp, dp, dp, dp, C.POINTER(Scalars)]),
        "imex_step": (C.c_int, [vp, C.c_int, C.POINTER(Stats)]),
        "imex_step_async": (C.c_int, [vp, C.c_int]),
        "imex_sync": (C.c_int, [vp]),
        "imex_get_stats": (C.c_int, [vp, C.POINTER(Stats)]),
        "imex_reset_stats": (C.c_int, [vp]),
        "imex_div_b": (C.c_double, [vp]),
        "imex_div_b_noalloc": (C.c_int, [C.POINTER(Params), vp, vp, vp, vp]),
        "imex_set_stage": (C.c_int, [vp, C.POINTER(StageHost), C.POINTER(Scalars)]),
        "imex_get_stage": (C.c_int, [vp, C.POINTER(StageHost), C.POINTER(Scalars)]),
        "imex_cfl": (C.c_int, [vp, C.POINTER(Scalars)]),
        "imex_stage_combine": (C.c_int, [vp, C.c_int]),
        "imex_explicit_conv": (C.c_int, [vp]),
        "imex_mag_update_b": (C.c_int, [vp]),
        "imex_explicit_press": (C.c_int, [vp]),
        "imex_pre_subsystem_pressure": (C.c_int, [vp]),
        "imex_explicit_energy": (C.c_int, [vp]),
        "imex_pre_subsystem_energy": (C.c_int, [vp]),
        "imex_init_potential": (C.c_int, [C.POINTER(Params), C.c_int, C.c_double, dp]),
        "imex_set_potential": (C.c_int, [vp, dp]),
        "imex_get_potential": (C.c_int, [vp, dp]),
        "imex_batch_set_potential": (C.c_int, [vp, C.c_int, dp]),
        "imex_batch_get_potential": (C.c_int, [vp, C.c_int, dp]),
        "imex_set_stage_potential": (C.c_int, [vp, dp, dp]),
        "imex_get_stage_potential": (C.c_int, [vp, dp, dp]),
        "imex_mag_update_a": (C.c_int, [vp]),
        "imex_diag_enable": (C.c_int, [vp, C.c_int]),
        "imex_diag_read": (C.c_int, [vp, C.c_int, C.POINTER(Diag), C.c_int]),
        "imex_diagnose": (C.c_int, [vp, C.c_int, C.POINTER(Diag)]),
        "imex_diagnose_ct_emf": (C.c_int, [vp, C.c_int, dp]),
        "imex_write_vtk_host": (C.c_int, [C.POINTER(Params), dp, dp, C.c_double, C.c_int, C.c_char_p]),
        "imex_write_vtk": (C.c_int, [vp, C.c_int, C.c_char_p, C.c_int]),
        "imex_write_fields_host": (C.c_int, [C.POINTER(Params), dp, dp, C.c_char_p]),
        "imex_write_fields": (C.c_int, [vp, C.c_int, C.c_char_p]),
        "imex_write_history": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int), dp]),
        "imex_output_path": (C.c_int, [C.POINTER(Params), C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]),
        "imex_unstaggered_ct": (C.c_int, [vp]),
        "imex_ct_update": (C.c_int, [vp]),
        "imex_apply_op": (C.c_int, [vp, C.c_int, dp, dp, dp, C.c_double]),
        "imex_apply_op_T": (C.c_int, [vp, C.c_int, dp, dp, dp, C.c_double]),
        "imex_gmres": (C.c_int, [vp, C.c_int, dp, dp, C.c_int, dp, C.c_double, C.POINTER(SolveInfo)]),
        "imex_pressure_newton": (C.c_int, [vp, dp, dp, C.POINTER(SolveInfo)]),
        "imex_profile_enable": (C.c_int, [vp, C.c_int]),
        "imex_profile_read": (C.c_int, [vp, C.POINTER(KernelProfile), C.c_int]),
        "imex_batch_create": (C.c_int, [C.POINTER(Params), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(vp)]),
        "imex_batch_size": (C.c_int, [vp]),
        "imex_batch_set_state": (C.c_int, [vp, C.c_int, dp, dp, dp, C.c_double, C.c_int, C.c_double]),
        "imex_batch_get_state": (C.c_int, [vp, C.c_int, dp, dp, dp, dp, C.POINTER(Scalars)]),
        "imex_batch_div_b": (C.c_double, [vp, C.c_int]),
        "imex_autotune": (C.c_int, [vp, C.c_int, C.POINTER(TuneInfo)]),
        "imex_adjoint_forward": (C.c_int, [vp, C.c_int, C.c_int]),
        "imex_adjoint_backward": (C.c_int, [vp, dp, dp, dp, dp, dp, dp, C.POINTER(AdjointInfo)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def grid_shape(p):
    return (p.ny + 2 * GC, p.nx + 2 * GC, NVAR)


def array_shape(p):
    return (p.ny + 2 * GC, p.nx + 2 * GC)


def init_case(p, case, amp=1.0, u0=0.0, v0=0.0, smooth=0):
    """Primitive initial data (Grid layout [j][i][k]) and staggered faces (Array layout [j][i])."""
    u = np.zeros(grid_shape(p))
    bx = np.zeros(array_shape(p))
    by = np.zeros(array_shape(p))
    rc = lib().imex_init_case(C.byref(p), case, amp, u0, v0, smooth, ptr(u), ptr(bx), ptr(by))
    if rc < 0:
        raise RuntimeError(lib().imex_last_error().decode())
    return u, bx, by


def write_vtk_host(p, u_con, u_pri, t, nstep, file):
    """printAllVariablesToVTK's file from host Grids."""
    rc = lib().imex_write_vtk_host(C.byref(p), ptr(np.ascontiguousarray(u_con, dtype=np.float64)),
                                   ptr(np.ascontiguousarray(u_pri, dtype=np.float64)), t, nstep, str(file).encode())
    if rc < 0:
        raise ImexError(lib().imex_last_error().decode())


def write_fields_host(p, u_con, u_pri, directory):
    """The end-of-run RK_*.dat field files from host Grids."""
    rc = lib().imex_write_fields_host(C.byref(p), ptr(np.ascontiguousarray(u_con, dtype=np.float64)),
                                      ptr(np.ascontiguousarray(u_pri, dtype=np.float64)), str(directory).encode())
    if rc < 0:
        raise ImexError(lib().imex_last_error().decode())


def write_history(file, steps, values):
    """A "step value" history file (RK_div_error.dat, RK_B_inf.dat, RK_dt_comp.dat)."""
    st = np.ascontiguousarray(steps, dtype=np.int32)
    va = np.ascontiguousarray(values, dtype=np.float64)
    rc = lib().imex_write_history(str(file).encode(), len(st), st.ctypes.data_as(C.POINTER(C.c_int)), ptr(va))
    if rc < 0:
        raise ImexError(lib().imex_last_error().decode())


def output_path(p, root, testname):
    """The reference's output directory of a run with these parameters."""
    buf = C.create_string_buffer(4096)
    rc = lib().imex_output_path(C.byref(p), str(root).encode(), str(testname).encode(), buf, 4096)
    if rc < 0:
        raise ImexError(lib().imex_last_error().decode())
    return buf.value.decode()


def init_potential(p, case, amp=1.0):
    """Vector potential of a built-in case (Grid layout [j][i][k], A_x, A_y, A_z in slots 0..2)."""
    A = np.zeros(grid_shape(p))
    rc = lib().imex_init_potential(C.byref(p), case, amp, ptr(A))
    if rc < 0:
        raise RuntimeError(lib().imex_last_error().decode())
    return A


def new_stage(p):
    g, a = grid_shape(p), array_shape(p)
    return {k: np.zeros(g) for k in ("u_con", "Qe", "Qs")} | {k: np.zeros(a) for k in ("bx", "by", "bx_qs", "by_qs", "fx", "fy")}


def stage_struct(st):
    s = StageHost()
    for k, _ in StageHost._fields_:
        setattr(s, k, ptr(st[k]) if st.get(k) is not None else None)
    return s


class ImexError(RuntimeError):
    pass


class Solver:
    """One device context (imex_ctx) with numpy convenience methods."""

    def __init__(self, params):
        self.p = params
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.imex_create(C.byref(params), C.byref(h))
        if rc < 0:
            raise ImexError(f"imex_create: {self.L.imex_last_error().decode()}")
        self.h = h

    def _chk(self, rc, what):
        if rc < 0:
            raise ImexError(f"{what} failed ({rc}): {self.L.imex_last_error().decode()}")
        return rc

    def close(self):
        if self.h:
            self.L.imex_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, u_pri, bx, by, t=0.0, timestep=0, prev_dt=0.0):
        return self._chk(self.L.imex_set_state(self.h, ptr(np.ascontiguousarray(u_pri)), ptr(np.ascontiguousarray(bx)),
                                               ptr(np.ascontiguousarray(by)), t, timestep, prev_dt), "set_state")

    def diag_enable(self, capacity):
        return self._chk(self.L.imex_diag_enable(self.h, capacity), "diag_enable")

    def diag_read(self, i=0, max_records=100000):
        arr = (Diag * max_records)()
        n = self._chk(self.L.imex_diag_read(self.h, i, arr, max_records), "diag_read")
        return [struct_dict(arr[q]) for q in range(n)]

    def diagnose(self, i=0):
        d = Diag()
        self._chk(self.L.imex_diagnose(self.h, i, C.byref(d)), "diagnose")
        return struct_dict(d)

    def diagnose_ct_emf(self, i=0):
        out = np.zeros((4, 4))
        self._chk(self.L.imex_diagnose_ct_emf(self.h, i, ptr(out)), "diagnose_ct_emf")
        return out

    def write_vtk(self, file, nstep, i=0):
        return self._chk(self.L.imex_write_vtk(self.h, i, str(file).encode(), nstep), "write_vtk")

    def write_fields(self, directory, i=0):
        return self._chk(self.L.imex_write_fields(self.h, i, str(directory).encode()), "write_fields")

    def set_potential(self, A):
        return self._chk(self.L.imex_set_potential(self.h, ptr(np.ascontiguousarray(A, dtype=np.float64))), "set_potential")

    def get_potential(self):
        A = np.zeros(grid_shape(self.p))
        self._chk(self.L.imex_get_potential(self.h, ptr(A)), "get_potential")
        return A

    def set_stage_potential(self, A, As):
        return self._chk(self.L.imex_set_stage_potential(self.h, ptr(np.ascontiguousarray(A, dtype=np.float64)),
                                                         ptr(np.ascontiguousarray(As, dtype=np.float64))), "set_stage_potential")

    def get_stage_potential(self):
        A, As = np.zeros(grid_shape(self.p)), np.zeros(grid_shape(self.p))
        self._chk(self.L.imex_get_stage_potential(self.h, ptr(A), ptr(As)), "get_stage_potential")
        return A, As

    def get_state(self):
        p = self.p
        u_con, u_pri = np.zeros(grid_shape(p)), np.zeros(grid_shape(p))
        bx, by = np.zeros(array_shape(p)), np.zeros(array_shape(p))
        sc = Scalars()
        self._chk(self.L.imex_get_state(self.h, ptr(u_con), ptr(u_pri), ptr(bx), ptr(by), C.byref(sc)), "get_state")
        return u_con, u_pri, bx, by, struct_dict(sc)

    def step(self, n=1):
        st = Stats()
        rc = self._chk(self.L.imex_step(self.h, n, C.byref(st)), "step")
        return rc, struct_dict(st)

    def step_async(self, n=1):
        return self._chk(self.L.imex_step_async(self.h, n), "step_async")

    def sync(self):
        return self._chk(self.L.imex_sync(self.h), "sync")

    def stats(self):
        st = Stats()
        self._chk(self.L.imex_get_stats(self.h, C.byref(st)), "get_stats")
        return struct_dict(st)

    def reset_stats(self):
        self.L.imex_reset_stats(self.h)

    def div_b(self):
        return self.L.imex_div_b(self.h)

    def autotune(self, flags=TUNE_EXPLICIT | TUNE_KRYLOV):
        """Pick launch shapes by timing on the current state; returns the choice and the timings (us)."""
        ti = TuneInfo()
        self._chk(self.L.imex_autotune(self.h, flags, C.byref(ti)), "autotune")
        return dict(explicit_tile=(ti.explicit_tile_x, ti.explicit_tile_y), krylov_blocks=ti.krylov_blocks,
                    explicit_us=list(ti.explicit_us), krylov_us=list(ti.krylov_us),
                    krylov_p_blocks=ti.krylov_p_blocks, krylov_p_us=list(ti.krylov_p_us))

    def profile_enable(self, on=True):
        return self._chk(self.L.imex_profile_enable(self.h, 1 if on else 0), "profile_enable")

    def profile_read(self):
        arr = (KernelProfile * 64)()
        n = self._chk(self.L.imex_profile_read(self.h, arr, 64), "profile_read")
        return [dict(name=arr[i].name.decode(), launches=arr[i].launches, total_ms=arr[i].total_ms,
                     bytes=arr[i].bytes) for i in range(n)]

    def set_stage(self, st, sc):
        s = Scalars()
        for k, v in sc.items():
            setattr(s, k, v)
        return self._chk(self.L.imex_set_stage(self.h, C.byref(stage_struct(st)), C.byref(s)), "set_stage")

    def get_stage(self):
        st = new_stage(self.p)
        sc = Scalars()
        self._chk(self.L.imex_get_stage(self.h, C.byref(stage_struct(st)), C.byref(sc)), "get_stage")
        return st, struct_dict(sc)

    def sub(self, name):
        return self._chk(getattr(self.L, "imex_" + name)(self.h), name)

    def _n_unknowns(self, op, *vecs):
        n = self.p.nx * self.p.ny * (3 if op in (OP_MAG, OP_POTENTIAL) else 1)
        for v in vecs:
            if np.asarray(v).size != n:
                raise ValueError(f"operator {op} takes {n} unknowns, got {np.asarray(v).size}")
        return n

    def apply_op(self, op, x, p_freeze=None, mbig=0.0):
        self._n_unknowns(op, x)
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        pf = None if p_freeze is None else np.ascontiguousarray(p_freeze, dtype=np.float64)
        self._chk(self.L.imex_apply_op(self.h, op, ptr(x), ptr(y), ptr(pf), mbig), "apply_op")
        return y

    def gmres(self, op, b, x0, precond=0, p_freeze=None, mbig=0.0):
        self._n_unknowns(op, b, x0)
        x = np.array(x0, dtype=np.float64, copy=True)
        b = np.ascontiguousarray(b, dtype=np.float64)
        pf = None if p_freeze is None else np.ascontiguousarray(p_freeze, dtype=np.float64)
        info = SolveInfo()
        rc = self._chk(self.L.imex_gmres(self.h, op, ptr(b), ptr(x), precond, ptr(pf), mbig, C.byref(info)), "gmres")
        return x, struct_dict(info), rc

    def apply_op_T(self, op, x, p_freeze=None, mbig=0.0):
        self._n_unknowns(op, x)
        out = np.zeros_like(x)
        self._chk(self.L.imex_apply_op_T(self.h, op, ptr(np.ascontiguousarray(x)), ptr(out), ptr(p_freeze), mbig), "apply_op_T")
        return out

    def adjoint_forward(self, nsteps, every):
        """Run nsteps from the state given to set_state, keeping a checkpoint every `every` steps."""
        return self._chk(self.L.imex_adjoint_forward(self.h, nsteps, every), "adjoint_forward")

    def adjoint_backward(self, gu_con, gbx=None, gby=None):
        """Cotangent of the final (conservative state, faces) -> gradient w.r.t. the initial primitive state and faces."""
        p = self.p
        gu = np.ascontiguousarray(gu_con, dtype=np.float64)
        gbx = None if gbx is None else np.ascontiguousarray(gbx, dtype=np.float64)
        gby = None if gby is None else np.ascontiguousarray(gby, dtype=np.float64)
        au, abx, aby = np.zeros(grid_shape(p)), np.zeros(array_shape(p)), np.zeros(array_shape(p))
        info = AdjointInfo()
        self._chk(self.L.imex_adjoint_backward(self.h, ptr(gu), ptr(gbx), ptr(gby), ptr(au), ptr(abx), ptr(aby),
                                               C.byref(info)), "adjoint_backward")
        return au, abx, aby, struct_dict(info)

    def pressure_newton(self, b, p0):
        p = np.array(p0, dtype=np.float64, copy=True)
        b = np.ascontiguousarray(b, dtype=np.float64)
        info = SolveInfo()
        rc = self._chk(self.L.imex_pressure_newton(self.h, ptr(b), ptr(p), C.byref(info)), "pressure_newton")
        return p, struct_dict(info), rc


class Batch(Solver):
    """A batch context: independent instances sharing `params` except the grid size."""

    def __init__(self, params, sizes):
        self.p = params
        self.L = lib()
        self.sizes = [(int(a), int(b)) for a, b in sizes]
        n = len(self.sizes)
        nx = (C.c_int * n)(*[a for a, _ in self.sizes])
        ny = (C.c_int * n)(*[b for _, b in self.sizes])
        h = C.c_void_p()
        rc = self.L.imex_batch_create(C.byref(params), n, nx, ny, C.byref(h))
        if rc < 0:
            raise ImexError(f"imex_batch_create: {self.L.imex_last_error().decode()}")
        self.h = h

    def __len__(self):
        return self.L.imex_batch_size(self.h)

    def params_of(self, i):
        q = Params.from_buffer_copy(self.p)
        q.nx, q.ny = self.sizes[i]
        return q

    def set_instance(self, i, u_pri, bx, by, t=0.0, timestep=0, prev_dt=0.0):
        return self._chk(self.L.imex_batch_set_state(self.h, i, ptr(np.ascontiguousarray(u_pri)), ptr(np.ascontiguousarray(bx)),
                                                     ptr(np.ascontiguousarray(by)), t, timestep, prev_dt), "batch_set_state")

    def get_instance(self, i):
        p = self.params_of(i)
        u_con, u_pri = np.zeros(grid_shape(p)), np.zeros(grid_shape(p))
        bx, by = np.zeros(array_shape(p)), np.zeros(array_shape(p))
        sc = Scalars()
        self._chk(self.L.imex_batch_get_state(self.h, i, ptr(u_con), ptr(u_pri), ptr(bx), ptr(by), C.byref(sc)), "batch_get_state")
        return u_con, u_pri, bx, by, struct_dict(sc)

    def set_potential_of(self, i, A):
        return self._chk(self.L.imex_batch_set_potential(self.h, i, ptr(np.ascontiguousarray(A, dtype=np.float64))),
                         "batch_set_potential")

    def get_potential_of(self, i):
        A = np.zeros(grid_shape(self.params_of(i)))
        self._chk(self.L.imex_batch_get_potential(self.h, i, ptr(A)), "batch_get_potential")
        return A

    def div_b_of(self, i):
        return self.L.imex_batch_div_b(self.h, i)

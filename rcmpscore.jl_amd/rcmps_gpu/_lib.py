"""ctypes binding of librcmps_hip.so (C ABI in include/rcmps.h).

The library is the only compute path: when it is missing or no HIP device is present the
calls raise instead of falling back to anything else.
"""
from __future__ import annotations

import atexit
import ctypes
import threading
import weakref
from pathlib import Path

import os

# RCMPS_LIB selects an alternative build of the library (e.g. an instrumented one)
LIB_PATH = Path(os.environ.get("RCMPS_LIB") or Path(__file__).resolve().parent / "_lib" / "librcmps_hip.so")

RCMPS_OK = 0
STATUS = {
    0: "success", 1: "max_steps", 2: "step_underflow", 3: "nonfinite_error", 5: "missing_samples",
    7: "wait_limit", 10: "solver_failure", 11: "invalid_argument", 12: "device_error", 13: "unsupported",
    14: "scratch_limit",
}
SYSTEMS = ("kinetic forward", "kinetic adjoint", "kinetic-commutator forward",
           "kinetic-commutator adjoint", "phi4 forward", "phi4 adjoint")
FLAG_NO_BOUNDARY = 1
FLAG_SPLIT = 2
FLAG_KINETIC_KRYLOV = 4
FLAG_PHASE_TIMERS = 8
FLAG_DENSE = 16


class SolverFailure(RuntimeError):
    """A numerical solve failed (src/transfer.jl:28-31): line searches reject such trials."""


class DeviceControlFailure(SolverFailure):
    """The device DOP853 controller stopped (statuses of device_control_backend.py:23-27)."""

    def __init__(self, status, stats=None):
        self.status = status
        self.stats = stats
        super().__init__(f"device DOP853 failed: {STATUS.get(status, status)}"
                         + (f" in {SYSTEMS[stats.failed_system]}" if stats is not None
                            and 0 <= stats.failed_system < 6 else ""))


class Integrator(ctypes.Structure):
    _fields_ = [("n_quad", ctypes.c_int), ("max_point", ctypes.c_double), ("atol", ctypes.c_double),
                ("rtol", ctypes.c_double), ("maxiters", ctypes.c_longlong),
                ("initial_step", ctypes.c_double)]


class Stats(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("failed_system", ctypes.c_int),
                ("attempts", ctypes.c_longlong * 6), ("accepted", ctypes.c_longlong * 6),
                ("rejected", ctypes.c_longlong * 6), ("dense_steps", ctypes.c_longlong * 6),
                ("rhs_calls", ctypes.c_longlong * 6), ("term_values", ctypes.c_double * 3),
                ("trajectory_ms", ctypes.c_double), ("pullback_ms", ctypes.c_double),
                ("boundary_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("krylov_matvecs", ctypes.c_int), ("krylov_residual", ctypes.c_double),
                ("phase_us", (ctypes.c_double * 12) * 6),
                ("kinetic_nodes", ctypes.c_int), ("kinetic_basis", ctypes.c_int * 4),
                ("kinetic_actions", ctypes.c_int), ("kinetic_residual", ctypes.c_double),
                ("kinetic_ms", ctypes.c_double * 4)]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            if hasattr(v, "__len__"):
                out[name] = [list(x) if hasattr(x, "__len__") else x for x in v]
            else:
                out[name] = v
        out["status_name"] = STATUS.get(self.status, str(self.status))
        return out


class KineticOptions(ctypes.Structure):
    """rcmps_kinetic_options: KineticOptions (src/kinetic.jl:9-16) + KineticKrylovOptions
    (src/kinetic_krylov.jl:8-13)."""
    _fields_ = [("minnodes", ctypes.c_int), ("maxnodes", ctypes.c_int), ("atol", ctypes.c_double),
                ("rtol", ctypes.c_double), ("tolerance", ctypes.c_double), ("fallback_tolerance", ctypes.c_double),
                ("initialdim", ctypes.c_int), ("maxdim", ctypes.c_int)]


# C ABI symbols: name -> (restype, argtypes)
_P = ctypes.c_void_p
SIGNATURES = {
    "rcmps_context_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_P)]),
    "rcmps_context_destroy": (None, [_P]),
    "rcmps_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "rcmps_max_bond_dimension": (ctypes.c_int, []),
    "rcmps_state_create": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, _P, ctypes.c_double, ctypes.POINTER(_P),
                                          ctypes.POINTER(Stats)]),
    "rcmps_state_destroy": (None, [_P]),
    "rcmps_state_get": (ctypes.c_int, [_P, _P, ctypes.c_int, _P]),
    "rcmps_fg_phi4": (ctypes.c_int, [_P, _P, ctypes.c_double, ctypes.POINTER(Integrator), ctypes.c_double,
                                     ctypes.c_int, ctypes.POINTER(ctypes.c_double), _P, _P, _P,
                                     ctypes.POINTER(Stats)]),
    "rcmps_fg_phi4_batch": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(Integrator), ctypes.c_double, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_P), ctypes.POINTER(Stats)]),
    "rcmps_set_kinetic_options": (ctypes.c_int, [_P, ctypes.POINTER(KineticOptions)]),
    "rcmps_gemm_variant": (ctypes.c_int, [_P, ctypes.c_int]),
    "rcmps_set_nojump_threshold": (ctypes.c_int, [_P, ctypes.c_double]),
    "rcmps_set_layout": (ctypes.c_int, [_P, ctypes.c_int]),
    "rcmps_state_set_transfer_eigen": (ctypes.c_int, [_P, _P, ctypes.c_int, _P, _P, _P]),
    "rcmps_state_nojump_form": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]),
    "rcmps_gemm_tuning": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_int)]),
    "rcmps_kinetic": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_double), _P, _P, _P,
                                     ctypes.POINTER(Stats)]),
    "rcmps_fg_phi4_rk": (ctypes.c_int, [_P, _P, ctypes.c_double, ctypes.POINTER(Integrator), ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_double), _P, _P, ctypes.POINTER(Stats)]),
    "rcmps_boundary_environment": (ctypes.c_int, [_P, _P, _P, ctypes.c_double, _P, ctypes.POINTER(Stats)]),
    "rcmps_boundary_environment_noalloc": (ctypes.c_int, [_P, _P, _P, ctypes.c_double, _P, ctypes.POINTER(Stats)]),
    "rcmps_trajectory_samples": (ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.c_longlong]),
    "rcmps_excitation_create": (ctypes.c_int, [_P, _P, ctypes.c_double, ctypes.c_double, ctypes.POINTER(_P)]),
    "rcmps_excitation_destroy": (None, [_P]),
    "rcmps_excitation_poisson": (ctypes.c_int, [_P, _P, _P, ctypes.c_int, ctypes.c_int, _P]),
    "rcmps_excitation_action": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.POINTER(_P), ctypes.POINTER(_P),
                                               ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                               ctypes.POINTER(Integrator), ctypes.c_int, ctypes.c_int,
                                               ctypes.POINTER(Stats)]),
    "rcmps_excitation_action_multi": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(_P), ctypes.POINTER(_P),
                                                     ctypes.POINTER(_P), ctypes.c_double,
                                                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(Integrator),
                                                     ctypes.c_int, ctypes.c_int, ctypes.POINTER(Stats)]),
    "rcmps_excitation_element": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(_P), ctypes.POINTER(_P),
                                                ctypes.POINTER(_P), ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                                ctypes.POINTER(Integrator), ctypes.c_int, ctypes.c_int,
                                                ctypes.POINTER(ctypes.c_double), ctypes.POINTER(Stats)]),
}
EXC_KINETIC = 1
EXC_POTENTIAL = 2
EXC_DENSE = 4
EXC_GRAPH = 8
EXC_PERSISTENT = 16

_lock = threading.Lock()
_lib = None
_contexts = {}
_states = weakref.WeakSet()
_closed = False


def register_state(obj) -> None:
    """Track a State so that its device memory is released before the runtime shuts down."""
    _states.add(obj)


def closed() -> bool:
    return _closed


@atexit.register
def _shutdown() -> None:
    global _closed
    if _lib is None or _closed:
        return
    for s in list(_states):
        try:
            s.close()
        except Exception:
            pass
    for ctx in list(_contexts.values()):
        _lib.rcmps_context_destroy(ctx)
    _contexts.clear()
    _closed = True


def load():
    """Load the shared library (no device calls)."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                                   "(make -C rcmpscore.jl_amd/csrc)")
            lib = ctypes.CDLL(str(LIB_PATH))
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def context(device: int = 0):
    """The per-device context (one HIP stream and its workspaces)."""
    lib = load()
    with _lock:
        ctx = _contexts.get(device)
        if ctx is None:
            handle = _P()
            st = lib.rcmps_context_create(int(device), ctypes.byref(handle))
            if st != RCMPS_OK:
                raise RuntimeError(f"rcmps_context_create failed: {STATUS.get(st, st)}")
            ctx = handle
            _contexts[device] = ctx
    return ctx


def check(status: int, what: str, stats: Stats | None = None):
    if status == RCMPS_OK:
        return
    if status in (1, 2, 3, 5, 7):
        raise DeviceControlFailure(status, stats)
    if status == 10:
        raise SolverFailure(f"{what}: solver failure")
    if status in (11,):
        raise ValueError(f"{what}: invalid argument")
    if status == 13:
        raise NotImplementedError(f"{what}: configuration not supported by the built kernels")
    if status == 14:
        raise ValueError(f"{what}: input larger than the no-alloc entry point's fixed scratch")
    raise RuntimeError(f"{what}: {STATUS.get(status, status)}")


GEMM_VARIANTS = ("four waves per tile, k split", "four tiles per workgroup", "one tile per one-wave workgroup",
                 "32x32 block from LDS panels")


def gemm_variant(variant: int | None = None, device: int = 0):
    """Fix the batched-product tile variant of a device's context (0..3), or hand the choice back
    to the autotuner (None).  Every variant produces the same bits."""
    v = -1 if variant is None else int(variant)
    check(load().rcmps_gemm_variant(context(device), v), "gemm variant")


def nojump_threshold(kappa: float = 1e6, device: int = 0):
    """States created afterwards on the device apply the no-jump inverse in the Schur basis of Q when
    the condition estimate of its eigenvectors exceeds kappa (0: always; inf: never), else in the
    eigenbasis.  The default 1e6 leaves the eigenbasis to every diagonalisable Q met in practice and
    the Schur form to defective ones."""
    check(load().rcmps_set_nojump_threshold(context(device), float(kappa)), "nojump threshold")


def gemm_tuning(device: int = 0):
    """The autotuner's table: a list of dicts (padded_size, batch_class, variant, us per variant)."""
    lib = load()
    ctx = context(device)
    count = ctypes.c_int()
    check(lib.rcmps_gemm_tuning(ctx, 0, None, None, None, None, ctypes.byref(count)), "gemm tuning")
    n = count.value
    if n == 0:
        return []
    dp, cls, var = (ctypes.c_int * n)(), (ctypes.c_int * n)(), (ctypes.c_int * n)()
    nv = len(GEMM_VARIANTS)
    us = (ctypes.c_double * (nv * n))()
    check(lib.rcmps_gemm_tuning(ctx, n, dp, cls, var, us, ctypes.byref(count)), "gemm tuning")
    return [dict(padded_size=dp[i], batch_class=cls[i], variant=var[i], us=[us[nv * i + v] for v in range(nv)])
            for i in range(min(n, count.value))]

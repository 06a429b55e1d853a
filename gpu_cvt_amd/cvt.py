"""Python host mirror of the reference's Lloyd loop over the C ABI (ctypes).

The reference drives its loop from ``main()`` (main.cu:371-1073) with
positional arguments ``gpu_cvt <mesh.obj> [mode] [iters]``; this module
exposes the same loop as a :class:`Job` object over ``libcvt_hip.so``.  It
loads the in-tree library and fails loudly if it is missing -- there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# CVT_LIB selects another build of the library (the stage-clock profiling build)
LIB_PATH = os.environ.get("CVT_LIB") or os.path.join(_HERE, "libcvt_hip.so")

CVT_OK = 0
STATUS = {-1: "EINVAL", -2: "EHIP", -3: "ENOMEM", -4: "ESTATE", -5: "ENONFINITE"}

# Every entry point include/cvt.h declares (checked by the CPU tests).
EXPORTS = [
    "cvt_version", "cvt_last_error", "cvt_params_default", "cvt_job_create", "cvt_job_destroy",
    "cvt_job_info_get", "cvt_job_step", "cvt_job_step_timed", "cvt_job_get_sites", "cvt_job_get_knn",
    "cvt_job_get_knnv", "cvt_job_get_freeze", "cvt_job_get_normals", "cvt_job_set_state", "cvt_job_set_tiers",
    "cvt_job_get_debug", "cvt_job_energy", "cvt_job_point_mesh_d2", "cvt_job_profile",
    "cvt_bvh_create", "cvt_bvh_point_d2", "cvt_bvh_destroy", "cvt_job_remesh", "cvt_free", "cvt_job_stats",
    "cvt_batch_create", "cvt_batch_step", "cvt_batch_step_timed", "cvt_batch_destroy", "cvt_job_hausdorff",
    "cvt_server_config_default", "cvt_server_create", "cvt_server_submit", "cvt_server_poll",
    "cvt_server_get_sites", "cvt_server_wait_idle", "cvt_server_info_get", "cvt_server_destroy",
    "cvt_job_compute_tiers", "cvt_job_hausdorff_within", "cvt_bvh_point_d2_noalloc",
    "cvt_job_profile_stages", "cvt_mesh_quality", "cvt_job_quality", "cvt_mesh_compact",
]


class Params(C.Structure):
    _fields_ = [
        ("K", C.c_int), ("kproj", C.c_int), ("freeze_mode", C.c_int), ("freeze_need", C.c_int),
        ("freeze_disp_rel", C.c_double), ("freeze_disp_abs2", C.c_double),
        ("tier_streak", C.c_int * 6), ("tier_disp_abs2", C.c_double * 6),
        ("nan_policy", C.c_int), ("clip_R_mode", C.c_int), ("clip_R", C.c_double),
        ("site_cell_scale", C.c_double), ("vert_cell_scale", C.c_double),
        ("debug", C.c_int), ("exclude_self_vertex", C.c_int), ("remesh_R_mode", C.c_int),
        ("reserved", C.c_int * 5),
    ]


class IterStats(C.Structure):
    _fields_ = [
        ("count_same", C.c_int), ("count_low_disp", C.c_int), ("count_both", C.c_int),
        ("count_newly_frozen", C.c_int), ("frozen_total", C.c_int), ("knn_exact_rescans", C.c_int),
        ("knn_reranks", C.c_int), ("cell_failures", C.c_int), ("nonfinite_sites", C.c_int),
        ("deferred_sites", C.c_int), ("max_disp", C.c_float), ("deferred_projections", C.c_int),
        ("fp64_cells", C.c_int), ("far_walks", C.c_int),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class JobInfo(C.Structure):
    _fields_ = [
        ("nV", C.c_int), ("nF", C.c_int), ("nS", C.c_int), ("K", C.c_int), ("kproj", C.c_int),
        ("site_grid_dims", C.c_int * 3), ("vert_grid_dims", C.c_int * 3),
        ("site_table", C.c_int), ("vert_table", C.c_int),
        ("site_cell", C.c_double), ("vert_cell", C.c_double),
        ("bbox_lo", C.c_double * 3), ("bbox_hi", C.c_double * 3),
        ("bbox_diag", C.c_double), ("clip_R", C.c_double), ("area", C.c_double),
        ("device_bytes", C.c_ulonglong), ("iterations_done", C.c_int), ("reserved", C.c_int * 7),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_ulonglong), ("epoch_iters", C.c_int), ("min_check_iters", C.c_int),
        ("check_disp_rel", C.c_double), ("check_frozen_frac", C.c_double), ("ingest_threads", C.c_int),
        ("check_threads", C.c_int), ("batch_jobs0", C.c_int), ("batch_site_blocks0", C.c_int),
        ("reserved", C.c_int * 4),
    ]


class Result(C.Structure):
    _fields_ = [
        ("status", C.c_int), ("error", C.c_int), ("nS", C.c_int), ("iterations", C.c_int),
        ("hausdorff", C.c_double), ("bbox_diag", C.c_double),
        ("t_arrive", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double), ("reserved", C.c_int * 4),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class ServerInfo(C.Structure):
    _fields_ = [
        ("site_iters", C.c_double), ("busy_seconds", C.c_double),
        ("peak_bytes", C.c_ulonglong), ("used_bytes", C.c_ulonglong),
        ("epochs", C.c_longlong), ("rebuilds", C.c_longlong), ("hausdorff_checks", C.c_longlong),
        ("pending", C.c_int), ("active", C.c_int), ("failures", C.c_int),
        ("rebuilds_grow", C.c_int), ("compactions", C.c_int), ("extent_updates", C.c_int),
        ("reserved", C.c_int * 2),
        ("step_seconds", C.c_double), ("check_seconds", C.c_double), ("rebuild_seconds", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class StageMs(C.Structure):
    _fields_ = [("grid", C.c_double), ("frame", C.c_double), ("cell", C.c_double), ("project", C.c_double),
                ("precise", C.c_double), ("iter", C.c_double), ("reserved", C.c_double * 2)]


class Quality(C.Structure):
    _fields_ = [("Qmin", C.c_float), ("Qavg", C.c_float), ("theta_min", C.c_float), ("theta_min_avg", C.c_float),
                ("theta_lt_30_pct", C.c_float), ("theta_gt_90_pct", C.c_float), ("nF", C.c_int),
                ("reserved", C.c_int * 5)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class CVTError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed ({STATUS.get(code, code)}): {msg}")
        self.code = code


_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def _nullable(nd):
    """ndpointer that also accepts None."""

    class _P(nd):  # type: ignore[misc, valid-type]
        @classmethod
        def from_param(cls, obj):
            if obj is None:
                return None
            return super().from_param(obj)

    return _P


_dpn, _ipn, _upn = _nullable(_dp), _nullable(_ip), _nullable(_up)


def lib():
    """Load libcvt_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m gpu_cvt_amd.build` "
                           "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    L.cvt_version.restype = C.c_int
    L.cvt_last_error.restype = C.c_char_p
    L.cvt_params_default.argtypes = [C.POINTER(Params)]
    L.cvt_params_default.restype = None
    L.cvt_job_create.argtypes = [_dp, C.c_int, _ip, C.c_int, _dp, C.c_int, C.POINTER(Params), C.POINTER(C.c_void_p)]
    L.cvt_job_destroy.argtypes = [C.c_void_p]
    L.cvt_job_destroy.restype = None
    L.cvt_job_info_get.argtypes = [C.c_void_p, C.POINTER(JobInfo)]
    L.cvt_job_step.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    L.cvt_job_step_timed.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
    L.cvt_job_get_sites.argtypes = [C.c_void_p, _dp]
    L.cvt_job_get_knn.argtypes = [C.c_void_p, _ip]
    L.cvt_job_get_knnv.argtypes = [C.c_void_p, _ip]
    L.cvt_job_get_freeze.argtypes = [C.c_void_p, _upn, _upn]
    L.cvt_job_get_normals.argtypes = [C.c_void_p, _dp]
    L.cvt_job_set_state.argtypes = [C.c_void_p, _dpn, _ipn, _ipn, _upn, _upn, C.c_int]
    L.cvt_job_set_tiers.argtypes = [C.c_void_p, _up]
    L.cvt_job_compute_tiers.argtypes = [C.c_void_p, C.c_int, _upn, _dpn, _dpn]
    L.cvt_job_get_debug.argtypes = [C.c_void_p, _ipn, _dpn, _dpn, _dpn, _dpn, _dpn]
    L.cvt_job_energy.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.cvt_job_point_mesh_d2.argtypes = [C.c_void_p, _dp, C.c_int, _dp]
    L.cvt_job_profile.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.cvt_job_stats.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    L.cvt_job_profile_stages.argtypes = [C.c_void_p, C.c_int, C.POINTER(StageMs)]
    L.cvt_mesh_quality.argtypes = [_dp, C.c_int, _ip, C.c_int, C.POINTER(Quality)]
    L.cvt_mesh_compact.argtypes = [_dp, C.c_int, _ip, C.c_int, _dp, _ip, C.POINTER(C.c_int)]
    L.cvt_job_quality.argtypes = [C.c_void_p, _ip, C.c_int, C.POINTER(Quality)]
    L.cvt_job_hausdorff.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.cvt_job_hausdorff_within.argtypes = [C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_int),
                                           C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.cvt_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]
    L.cvt_batch_step.argtypes = [C.c_void_p, C.c_int]
    L.cvt_batch_step_timed.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
    L.cvt_batch_destroy.argtypes = [C.c_void_p]
    L.cvt_batch_destroy.restype = None
    L.cvt_job_remesh.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.c_int)]
    L.cvt_free.argtypes = [C.c_void_p]
    L.cvt_free.restype = None
    L.cvt_bvh_create.argtypes = [_dp, C.c_int, _ip, C.c_int, C.POINTER(C.c_void_p)]
    L.cvt_bvh_point_d2.argtypes = [C.c_void_p, _dp, C.c_int, _dp]
    L.cvt_bvh_destroy.argtypes = [C.c_void_p]
    L.cvt_bvh_point_d2_noalloc.argtypes = [C.c_void_p, _dp, C.c_int, _dp]
    L.cvt_server_config_default.argtypes = [C.POINTER(ServerConfig)]
    L.cvt_server_config_default.restype = None
    L.cvt_server_create.argtypes = [C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]
    L.cvt_server_submit.argtypes = [C.c_void_p, _dp, C.c_int, _ip, C.c_int, _dp, C.c_int, C.POINTER(Params),
                                    C.c_double, C.c_int, C.POINTER(C.c_ulonglong)]
    L.cvt_server_poll.argtypes = [C.c_void_p, C.c_ulonglong, C.POINTER(Result)]
    L.cvt_server_get_sites.argtypes = [C.c_void_p, C.c_ulonglong, _dp]
    L.cvt_server_wait_idle.argtypes = [C.c_void_p, C.c_double]
    L.cvt_server_info_get.argtypes = [C.c_void_p, C.POINTER(ServerInfo)]
    L.cvt_server_destroy.argtypes = [C.c_void_p]
    L.cvt_server_destroy.restype = None
    L.cvt_bvh_destroy.restype = None
    missing = [n for n in EXPORTS if not hasattr(L, n)]
    if missing:
        raise RuntimeError(f"{LIB_PATH} lacks exports {missing}: rebuild it")
    _lib = L
    return L


def _check(fn: str, rc: int) -> None:
    if rc != CVT_OK:
        msg = lib().cvt_last_error()
        raise CVTError(fn, rc, msg.decode() if msg else "")


def default_params(**kw) -> Params:
    p = Params()
    lib().cvt_params_default(C.byref(p))
    for k, v in kw.items():
        if k in ("tier_streak", "tier_disp_abs2"):
            arr = getattr(p, k)
            for t, x in enumerate(v):
                arr[t] = x
        else:
            setattr(p, k, v)
    return p


class Job:
    """One device-resident remeshing job (mesh + sites + Lloyd state)."""

    def __init__(self, V, F, S0, params: Params | None = None, **kw):
        L = lib()
        self.params = params if params is not None else default_params(**kw)
        V = np.ascontiguousarray(V, dtype=np.float64).reshape(-1, 3)
        F = np.ascontiguousarray(F, dtype=np.int32).reshape(-1, 3)
        S0 = np.ascontiguousarray(S0, dtype=np.float64).reshape(-1, 3)
        self.nV, self.nF, self.nS = len(V), len(F), len(S0)
        self.K = self.params.K
        self.kproj = self.params.kproj
        h = C.c_void_p()
        _check("cvt_job_create", L.cvt_job_create(V.reshape(-1), len(V), F.reshape(-1), len(F), S0.reshape(-1),
                                                   len(S0), C.byref(self.params), C.byref(h)))
        self._h = h

    # ------------------------------------------------------------- lifecycle
    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().cvt_job_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        if not self._h:
            raise CVTError("job", -4, "job is closed")
        return self._h

    def info(self) -> dict:
        o = JobInfo()
        _check("cvt_job_info_get", lib().cvt_job_info_get(self.handle, C.byref(o)))
        d = {}
        for k, _ in o._fields_:
            if k == "reserved":
                continue
            v = getattr(o, k)
            d[k] = list(v) if hasattr(v, "__len__") else v
        return d

    # --------------------------------------------------------------- running
    def step(self, iters: int = 1, stats: bool = False):
        if stats:
            arr = (IterStats * max(iters, 1))()
            _check("cvt_job_step", lib().cvt_job_step(self.handle, iters, C.cast(arr, C.c_void_p)))
            return [arr[i].as_dict() for i in range(iters)]
        _check("cvt_job_step", lib().cvt_job_step(self.handle, iters, None))
        return None

    def step_timed(self, iters: int) -> float:
        ms = C.c_double()
        _check("cvt_job_step_timed", lib().cvt_job_step_timed(self.handle, iters, C.byref(ms)))
        return ms.value

    def profile(self, iters: int) -> dict:
        """Mean device ms per iteration: site-grid rebuild, FP32 Lloyd kernel,
        FP64-resolving pass, whole iteration."""
        g, l, pr, t = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        _check("cvt_job_profile", lib().cvt_job_profile(self.handle, iters, C.byref(g), C.byref(l), C.byref(pr),
                                                        C.byref(t)))
        return {"grid_ms": g.value, "lloyd_ms": l.value, "precise_ms": pr.value, "iter_ms": t.value}

    def profile_stages(self, iters: int) -> dict:
        """Mean device ms per iteration of each kernel stage (grid, frame, cell,
        project, precise) and of the whole iteration."""
        o = StageMs()
        _check("cvt_job_profile_stages", lib().cvt_job_profile_stages(self.handle, iters, C.byref(o)))
        return {k: getattr(o, k) for k, _ in o._fields_ if k != "reserved"}

    def quality(self, F) -> dict:
        """Triangle quality (evaluate.cuh:388-462) of faces F over the current sites, on the device."""
        F = np.ascontiguousarray(F, dtype=np.int32).reshape(-1, 3)
        o = Quality()
        _check("cvt_job_quality", lib().cvt_job_quality(self.handle, F.reshape(-1), len(F), C.byref(o)))
        return o.as_dict()

    # ----------------------------------------------------------------- state
    def sites(self) -> np.ndarray:
        out = np.zeros((self.nS, 3))
        _check("cvt_job_get_sites", lib().cvt_job_get_sites(self.handle, out.reshape(-1)))
        return out

    def knn(self) -> np.ndarray:
        out = np.zeros((self.nS, self.K), np.int32)
        _check("cvt_job_get_knn", lib().cvt_job_get_knn(self.handle, out.reshape(-1)))
        return out

    def knnv(self) -> np.ndarray:
        out = np.zeros((self.nS, self.kproj), np.int32)
        _check("cvt_job_get_knnv", lib().cvt_job_get_knnv(self.handle, out.reshape(-1)))
        return out

    def freeze_state(self):
        fz = np.zeros(self.nS, np.uint8)
        st = np.zeros(self.nS, np.uint8)
        _check("cvt_job_get_freeze", lib().cvt_job_get_freeze(self.handle, fz, st))
        return fz, st

    def normals(self) -> np.ndarray:
        out = np.zeros((self.nV, 3))
        _check("cvt_job_get_normals", lib().cvt_job_get_normals(self.handle, out.reshape(-1)))
        return out

    def set_state(self, S=None, knn_prev=None, knnv_prev=None, frozen=None, streak=None, has_prev: int = -1):
        c = lambda a, t: None if a is None else np.ascontiguousarray(a, dtype=t).reshape(-1)  # noqa: E731
        _check("cvt_job_set_state", lib().cvt_job_set_state(
            self.handle, c(S, np.float64), c(knn_prev, np.int32), c(knnv_prev, np.int32), c(frozen, np.uint8),
            c(streak, np.uint8), int(has_prev)))

    def set_tiers(self, tier) -> None:
        _check("cvt_job_set_tiers", lib().cvt_job_set_tiers(self.handle, np.ascontiguousarray(tier, np.uint8)))

    def compute_tiers(self, scheme: int):
        """Device freeze tiers (scheme 1: 5-tier normal variation; 2: 6-tier with
        the covariance split); stored in the job for freeze_mode 2.  Returns
        (tier uint8, nv float64, L float64) per site."""
        tier = np.zeros(self.nS, np.uint8)
        nv = np.zeros(self.nS)
        L = np.zeros(self.nS)
        _check("cvt_job_compute_tiers", lib().cvt_job_compute_tiers(self.handle, int(scheme), tier, nv, L))
        return tier, nv, L

    def debug(self) -> dict:
        n = self.nS
        d = dict(nn_site=np.zeros(n, np.int32), U=np.zeros((n, 3)), W=np.zeros((n, 3)), cent=np.zeros((n, 3)),
                 area=np.zeros(n), moment=np.zeros(n))
        _check("cvt_job_get_debug", lib().cvt_job_get_debug(
            self.handle, d["nn_site"], d["U"].reshape(-1), d["W"].reshape(-1), d["cent"].reshape(-1), d["area"],
            d["moment"]))
        return d

    def energy(self) -> float:
        e = C.c_double()
        _check("cvt_job_energy", lib().cvt_job_energy(self.handle, C.byref(e)))
        return e.value

    def last_stats(self, n: int = 1):
        """The last n per-iteration records (oldest first)."""
        arr = (IterStats * max(n, 1))()
        _check("cvt_job_stats", lib().cvt_job_stats(self.handle, n, C.cast(arr, C.c_void_p)))
        return [arr[i].as_dict() for i in range(n)]

    def remesh(self, eps: float = 1e-6) -> np.ndarray:
        """Dual triangles of the current cells (int32 (m, 3)), vertices = sites()."""
        fp = C.POINTER(C.c_int)()
        nf = C.c_int()
        _check("cvt_job_remesh", lib().cvt_job_remesh(self.handle, float(eps), C.byref(fp), C.byref(nf)))
        try:
            F = np.ctypeslib.as_array(fp, shape=(nf.value * 3,)).copy().reshape(-1, 3) if nf.value else \
                np.zeros((0, 3), np.int32)
        finally:
            lib().cvt_free(fp)
        return F

    def hausdorff(self, eps: float = 1e-6):
        """(one-sided remesh->input Hausdorff, remeshed triangle count), on the device."""
        h = C.c_double()
        nf = C.c_int()
        _check("cvt_job_hausdorff", lib().cvt_job_hausdorff(self.handle, float(eps), C.byref(h), C.byref(nf)))
        return h.value, nf.value

    def hausdorff_within(self, tol: float, eps: float = 1e-6):
        """(converged, exact distance if converged else -1): whether the
        one-sided Hausdorff distance is <= tol, by a bounded device query."""
        w = C.c_int()
        h = C.c_double()
        _check("cvt_job_hausdorff_within", lib().cvt_job_hausdorff_within(self.handle, float(eps), float(tol),
                                                                          C.byref(w), C.byref(h), None))
        return bool(w.value), h.value

    def point_mesh_d2(self, P) -> np.ndarray:
        P = np.ascontiguousarray(P, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(len(P))
        _check("cvt_job_point_mesh_d2", lib().cvt_job_point_mesh_d2(self.handle, P.reshape(-1), len(P), out))
        return out


class BVH:
    """Device LBVH over a triangle mesh: squared point-to-surface distances."""

    def __init__(self, V, F):
        V = np.ascontiguousarray(V, dtype=np.float64).reshape(-1, 3)
        F = np.ascontiguousarray(F, dtype=np.int32).reshape(-1, 3)
        h = C.c_void_p()
        _check("cvt_bvh_create", lib().cvt_bvh_create(V.reshape(-1), len(V), F.reshape(-1), len(F), C.byref(h)))
        self._h = h

    def point_d2(self, P) -> np.ndarray:
        P = np.ascontiguousarray(P, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(len(P))
        _check("cvt_bvh_point_d2", lib().cvt_bvh_point_d2(self._h, P.reshape(-1), len(P), out))
        return out

    def point_d2_noalloc(self, P) -> np.ndarray:
        """point_d2 staged in the library's fixed device area (no device
        allocation; at most 1,048,576 points per call)."""
        P = np.ascontiguousarray(P, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(len(P))
        _check("cvt_bvh_point_d2_noalloc", lib().cvt_bvh_point_d2_noalloc(self._h, P.reshape(-1), len(P), out))
        return out

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().cvt_bvh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mesh_compact(V, F):
    """(V', F'): the vertices some face references, in their order, and the
    faces renumbered (compact_mesh, main.cu:240-268), on the device."""
    V = np.ascontiguousarray(V, dtype=np.float64).reshape(-1, 3)
    F = np.ascontiguousarray(F, dtype=np.int32).reshape(-1, 3)
    Vo = np.zeros_like(V)
    Fo = np.zeros_like(F)
    n = C.c_int()
    _check("cvt_mesh_compact", lib().cvt_mesh_compact(V.reshape(-1), len(V), F.reshape(-1), len(F), Vo.reshape(-1),
                                                      Fo.reshape(-1), C.byref(n)))
    return Vo[:n.value].copy(), Fo


def mesh_quality(V, F) -> dict:
    """Triangle quality (evaluate.cuh:388-462) of (V, F) on the device, V rounded to FP32."""
    V = np.ascontiguousarray(V, dtype=np.float64).reshape(-1, 3)
    F = np.ascontiguousarray(F, dtype=np.int32).reshape(-1, 3)
    o = Quality()
    _check("cvt_mesh_quality", lib().cvt_mesh_quality(V.reshape(-1), len(V), F.reshape(-1), len(F), C.byref(o)))
    return o.as_dict()


def hausdorff(V1, F1, V2, F2, n_samples: int, seed: int, symmetric: bool = True) -> float:
    """Hausdorff distance between two triangle meshes from area-uniform samples
    (n_samples on each surface).  The reference measures one side only
    (candidate -> reference, evaluate.cuh:359-383) over vertices, face
    centroids and edge midpoints; ``symmetric=False`` gives that direction
    (samples on mesh 1, distances to mesh 2)."""
    from .meshgen import sample_surface

    b2 = BVH(V2, F2)
    P1 = sample_surface(np.asarray(V1, np.float64), np.asarray(F1), n_samples, seed)
    h = float(np.sqrt(b2.point_d2(P1).max()))
    b2.close()
    if symmetric:
        b1 = BVH(V1, F1)
        P2 = sample_surface(np.asarray(V2, np.float64), np.asarray(F2), n_samples, seed + 1)
        h = max(h, float(np.sqrt(b1.point_d2(P2).max())))
        b1.close()
    return h


class Batch:
    """Jobs advanced together by one graph launch per iteration."""

    def __init__(self, jobs):
        self.jobs = list(jobs)
        arr = (C.c_void_p * len(self.jobs))(*[j.handle for j in self.jobs])
        h = C.c_void_p()
        _check("cvt_batch_create", lib().cvt_batch_create(arr, len(self.jobs), C.byref(h)))
        self._h = h

    def step(self, iters: int = 1) -> None:
        _check("cvt_batch_step", lib().cvt_batch_step(self._h, iters))

    def step_timed(self, iters: int) -> float:
        ms = C.c_double()
        _check("cvt_batch_step_timed", lib().cvt_batch_step_timed(self._h, iters, C.byref(ms)))
        return ms.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().cvt_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


STATUS_NAMES = {0: "queued", 1: "running", 2: "converged", 3: "max_iters", 4: "failed"}


class Server:
    """Continuous-batching request loop (cvt_server_*): submit meshes + sites
    at any time; admitted jobs advance together in epochs; each retires when
    its remeshed surface is within ``tol_rel * bbox_diag`` of the input
    (one-sided Hausdorff) or after ``max_iters`` iterations."""

    def __init__(self, hbm_budget_bytes: int | None = None, **kw):
        L = lib()
        cfg = ServerConfig()
        L.cvt_server_config_default(C.byref(cfg))
        if hbm_budget_bytes is not None:
            cfg.hbm_budget_bytes = int(hbm_budget_bytes)
        for k, v in kw.items():
            setattr(cfg, k, v)
        self.config = cfg
        h = C.c_void_p()
        _check("cvt_server_create", L.cvt_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._nS = {}

    def submit(self, V, F, S0, tol_rel: float = 1e-3, max_iters: int = 100, params: Params | None = None,
               **kw) -> int:
        p = params if params is not None else default_params(**kw)
        V = np.ascontiguousarray(V, dtype=np.float64).reshape(-1, 3)
        F = np.ascontiguousarray(F, dtype=np.int32).reshape(-1, 3)
        S0 = np.ascontiguousarray(S0, dtype=np.float64).reshape(-1, 3)
        t = C.c_ulonglong()
        _check("cvt_server_submit", lib().cvt_server_submit(self._h, V.reshape(-1), len(V), F.reshape(-1), len(F),
                                                            S0.reshape(-1), len(S0), C.byref(p), float(tol_rel),
                                                            int(max_iters), C.byref(t)))
        self._nS[t.value] = len(S0)
        return t.value

    def poll(self, ticket: int) -> dict:
        r = Result()
        _check("cvt_server_poll", lib().cvt_server_poll(self._h, ticket, C.byref(r)))
        d = r.as_dict()
        d["state"] = STATUS_NAMES.get(d["status"], "?")
        return d

    def sites(self, ticket: int) -> np.ndarray:
        out = np.zeros((self._nS[ticket], 3))
        _check("cvt_server_get_sites", lib().cvt_server_get_sites(self._h, ticket, out.reshape(-1)))
        return out

    def wait_idle(self, timeout_s: float = 600.0) -> None:
        _check("cvt_server_wait_idle", lib().cvt_server_wait_idle(self._h, float(timeout_s)))

    def info(self) -> dict:
        o = ServerInfo()
        _check("cvt_server_info_get", lib().cvt_server_info_get(self._h, C.byref(o)))
        return o.as_dict()

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().cvt_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

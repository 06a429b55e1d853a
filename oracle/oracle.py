"""ctypes wrapper over the FP64 CPU oracle (``oracle/liboracle_cvt.so``) and the
FP32 CPU baseline (``oracle/libcpu32_cvt.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s cpu_baseline leg, as the checker.  The product package
``gpu_cvt_amd`` never imports this module.

All arrays are float64 / int32 numpy arrays; geometry is (n, 3).  Inputs that
the GPU stores in FP32 must be rounded to FP32 by the caller first
(:func:`f32r`) so that both sides start from identical values.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_cvt.so")
_CPU32_PATH = os.path.join(_HERE, "libcpu32_cvt.so")
_lib = None
_cpu32 = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile the oracle and the FP32 CPU baseline with their Makefile (g++ only)."""
    stale = any(not os.path.exists(lib) or os.path.getmtime(lib) < os.path.getmtime(os.path.join(_HERE, src))
                for lib, src in ((_LIB_PATH, "cvt_oracle.cpp"), (_CPU32_PATH, "cvt_cpu32.cpp")))
    if force or stale:
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_hw_threads.restype = C.c_int
        L.orc_get_threads.restype = C.c_int
        L.orc_key_d2.argtypes = [_dp, _dp]
        L.orc_key_d2.restype = C.c_double
        L.orc_vf_csr.argtypes = [C.c_int, C.c_int, _ip, _ip, _ip]
        L.orc_vertex_normals.argtypes = [C.c_int, _dp, C.c_int, _ip, _ip, _ip, _dp]
        L.orc_frames.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_knn_brute.argtypes = [_dp, C.c_int, _dp, C.c_int, C.c_int, C.c_int, _ip]
        L.orc_knn_grid.argtypes = [_dp, C.c_int, _dp, C.c_int, C.c_int, C.c_int, _ip]
        L.orc_cells.argtypes = [_dp, C.c_int, _dp, _dp, _ip, C.c_int, C.c_double, _dp, _dp, _dp, _ip]
        L.orc_cell_poly.argtypes = [_dp, C.c_int, C.c_int, _dp, _dp, _ip, C.c_int, C.c_double, _dp, _dp, _dp, _dp]
        L.orc_cell_poly.restype = C.c_int
        L.orc_closest_point_tri.argtypes = [_dp, _dp, _dp, _dp, _dp]
        L.orc_project.argtypes = [C.c_int, _dp, _ip, C.c_int, _dp, C.c_int, _ip, C.c_int, _ip, _ip, _dp]
        L.orc_freeze.argtypes = [C.c_int, C.c_int, _dp, _dp, _ip, _ip, _up, _dp, _ip, C.c_int, C.c_int, _up, _up, _ip]
        L.orc_lloyd_step.argtypes = [
            C.c_int, _dp, C.c_int, _ip, _dp, _ip, _ip, C.c_int, _dp, C.c_int, C.c_int, C.c_double, C.c_int,
            _dp, _ip, _up, C.c_int, C.c_int, C.c_int, _ip, _ip, _up, _up, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _ip]
        L.orc_remesh.argtypes = [_dp, C.c_int, _dp, _dp, _ip, C.c_int, C.c_double, _dp, C.c_double, _ip]
        L.orc_remesh.restype = C.c_int
        L.orc_quality.argtypes = [np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS"), C.c_int, _ip,
                                  C.c_int, np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")]
        L.orc_point_mesh_d2.argtypes = [C.c_int, _dp, C.c_int, _dp, C.c_int, _ip, C.c_int, _dp]
        L.orc_tiers.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, C.c_int, C.c_int, _up, _dp, _dp]
        _lib = L
    return _lib


def f32r(a) -> np.ndarray:
    """Round to FP32 and return as contiguous float64 (the GPU's storage)."""
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).astype(np.float64))


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def hw_threads() -> int:
    return int(lib().orc_hw_threads())


def key_d2(q, p) -> float:
    return float(lib().orc_key_d2(_d(q), _d(p)))


def vf_csr(nV: int, F):
    F = _i(F)
    off = np.zeros(nV + 1, np.int32)
    faces = np.zeros(3 * len(F), np.int32)
    lib().orc_vf_csr(nV, len(F), F.reshape(-1), off, faces)
    return off, faces


def vertex_normals(V, F, vf=None):
    V = _d(V)
    F = _i(F)
    off, faces = vf if vf is not None else vf_csr(len(V), F)
    N = np.zeros_like(V)
    lib().orc_vertex_normals(len(V), V.reshape(-1), len(F), F.reshape(-1), off, faces, N.reshape(-1))
    return N


def frames(Nsel):
    Nsel = _d(Nsel)
    U = np.zeros_like(Nsel)
    W = np.zeros_like(Nsel)
    lib().orc_frames(len(Nsel), Nsel.reshape(-1), U.reshape(-1), W.reshape(-1))
    return U, W


def knn(P, Q, k: int, exclude_self: bool, brute: bool = False):
    P = _d(P)
    Q = _d(Q)
    out = np.zeros((len(Q), k), np.int32)
    fn = lib().orc_knn_brute if brute else lib().orc_knn_grid
    fn(P.reshape(-1), len(P), Q.reshape(-1), len(Q), int(k), int(bool(exclude_self)), out.reshape(-1))
    return out


def cells(S, U, W, knn_, R: float):
    S = _d(S)
    K = knn_.shape[1]
    n = len(S)
    cent = np.zeros((n, 3))
    area = np.zeros(n)
    mom = np.zeros(n)
    npoly = np.zeros(n, np.int32)
    lib().orc_cells(S.reshape(-1), n, _d(U).reshape(-1), _d(W).reshape(-1), _i(knn_).reshape(-1), K, float(R),
                    cent.reshape(-1), area, mom, npoly)
    return cent, area, mom, npoly


def cell_poly(S, i: int, u, v, nb, R: float):
    S = _d(S)
    poly = np.zeros(128)
    cent = np.zeros(3)
    area = np.zeros(1)
    mom = np.zeros(1)
    nb = _i(nb)
    n = lib().orc_cell_poly(S.reshape(-1), len(S), int(i), _d(u), _d(v), nb, len(nb), float(R), poly, cent, area, mom)
    return poly[: 2 * n].reshape(n, 2), cent, float(area[0]), float(mom[0])


def closest_point_tri(P, A, B, Cc):
    out = np.zeros(3)
    lib().orc_closest_point_tri(_d(P), _d(A), _d(B), _d(Cc), out)
    return out


def project(Cpts, knnV, V, F, vf=None):
    Cpts = _d(Cpts)
    V = _d(V)
    F = _i(F)
    knnV = _i(knnV)
    off, faces = vf if vf is not None else vf_csr(len(V), F)
    out = np.zeros_like(Cpts)
    lib().orc_project(len(Cpts), Cpts.reshape(-1), knnV.reshape(-1), knnV.shape[1], V.reshape(-1), len(V),
                      F.reshape(-1), len(F), off, faces, out.reshape(-1))
    return out


def freeze(S, Snew, knn_, prev_knn, tier, thresh2_tier, streak_tier, has_prev, apply, streak, frozen):
    """Returns (streak', frozen', counts[4]) without modifying the inputs."""
    n, K = knn_.shape
    st = np.ascontiguousarray(streak, np.uint8).copy()
    fz = np.ascontiguousarray(frozen, np.uint8).copy()
    counts = np.zeros(4, np.int32)
    tier = np.zeros(n, np.uint8) if tier is None else np.ascontiguousarray(tier, np.uint8)
    lib().orc_freeze(n, K, _d(S).reshape(-1), _d(Snew).reshape(-1), _i(knn_).reshape(-1), _i(prev_knn).reshape(-1),
                     tier, _d(thresh2_tier), _i(streak_tier), int(has_prev), int(apply), st, fz, counts)
    return st, fz, counts


class Mesh:
    """FP32-rounded mesh with CSR and FP64 vertex normals (oracle side)."""

    def __init__(self, V, F):
        self.V = f32r(V)
        self.F = _i(F)
        self.vf = vf_csr(len(self.V), self.F)
        self.N = vertex_normals(self.V, self.F, self.vf)

    @property
    def bbox_diag(self) -> float:
        e = self.V.max(0) - self.V.min(0)
        return float(np.sqrt((e * e).sum()))

    @property
    def bbox_max_edge(self) -> float:
        return float((self.V.max(0) - self.V.min(0)).max())


class LloydState:
    """Oracle-side job state mirroring the GPU job."""

    def __init__(self, S, K: int, kproj: int = 5):
        n = len(S)
        self.S = f32r(S)
        self.knn = np.zeros((n, K), np.int32)
        self.knnV = np.full((n, kproj), -1, np.int32)
        self.frozen = np.zeros(n, np.uint8)
        self.streak = np.zeros(n, np.uint8)
        self.tier = np.zeros(n, np.uint8)
        self.has_prev = 0


def lloyd_step(mesh: Mesh, st: LloydState, *, K: int, kproj: int = 5, R: float, freeze_mode: int = 0,
               thresh2_tier=None, streak_tier=None, nan_policy: int = 0, exclude_self: bool = False):
    """One FP64 Lloyd iteration from ``st`` (modified in place, except S).
    exclude_self: site i's centroid -> mesh kproj-NN skips mesh vertex i
    (main.cu:728, near_project.cuh:72-74).

    Returns a dict of intermediates and the new sites (float64, NOT rounded).
    """
    n = len(st.S)
    thresh2_tier = np.zeros(6) if thresh2_tier is None else _d(thresh2_tier)
    streak_tier = np.full(6, 5, np.int32) if streak_tier is None else _i(streak_tier)
    out = dict(
        nn_site=np.zeros(n, np.int32), U=np.zeros((n, 3)), W=np.zeros((n, 3)), cent=np.zeros((n, 3)),
        Snew=np.zeros((n, 3)), area=np.zeros(n), moment=np.zeros(n), counts=np.zeros(4, np.int32))
    off, faces = mesh.vf
    lib().orc_lloyd_step(
        len(mesh.V), mesh.V.reshape(-1), len(mesh.F), mesh.F.reshape(-1), mesh.N.reshape(-1), off, faces, n,
        _d(st.S).reshape(-1), K, kproj, float(R), int(freeze_mode), thresh2_tier, streak_tier, st.tier,
        int(st.has_prev), int(nan_policy), int(bool(exclude_self)), st.knn.reshape(-1), st.knnV.reshape(-1), st.frozen, st.streak,
        out["nn_site"], out["U"].reshape(-1), out["W"].reshape(-1), out["cent"].reshape(-1),
        out["Snew"].reshape(-1), out["area"], out["moment"], out["counts"])
    st.has_prev = 1
    out["knn"] = st.knn.copy()
    out["knnV"] = st.knnV.copy()
    return out


def point_mesh_d2(P, V, F, use_bvh: bool = True):
    P = _d(P)
    V = _d(V)
    F = _i(F)
    out = np.zeros(len(P))
    lib().orc_point_mesh_d2(len(P), P.reshape(-1), len(V), V.reshape(-1), len(F), F.reshape(-1), int(use_bvh), out)
    return out


QUALITY_KEYS = ("Qmin", "Qavg", "theta_min", "theta_min_avg", "theta_lt_30_pct", "theta_gt_90_pct")


def quality(P, F) -> dict:
    """Triangle quality in FP32 (evaluate.cuh:388-462): P rounded to float32."""
    P = np.ascontiguousarray(P, dtype=np.float32).reshape(-1, 3)
    F = _i(F).reshape(-1, 3)
    out = np.zeros(6, np.float32)
    lib().orc_quality(P.reshape(-1), len(P), F.reshape(-1), len(F), out)
    return dict(zip(QUALITY_KEYS, (float(x) for x in out)))


def remesh(S, U, W, knn_, R: float, Nsite, eps: float = 1e-6):
    """Dual triangles of the tangent-plane cells (int32 (m,3)), oriented."""
    S = _d(S)
    n, K = knn_.shape
    out = np.zeros(3 * n * (K + 4), np.int32)
    m = lib().orc_remesh(S.reshape(-1), n, _d(U).reshape(-1), _d(W).reshape(-1), _i(knn_).reshape(-1), K, float(R),
                         _d(Nsite).reshape(-1), float(eps), out)
    return out[: 3 * m].reshape(m, 3)


def tiers(knn_, nn, N, scheme: int):
    """Freeze tiers from neighbour lists knn_ (n, K), each site's vertex nn (n,)
    and vertex normals N (nV, 3): returns (tier uint8, nv, L)."""
    knn_ = _i(knn_)
    n, K = knn_.shape
    N = _d(N)
    tier = np.zeros(n, np.uint8)
    nv = np.zeros(n)
    L = np.zeros(n)
    lib().orc_tiers(n, K, knn_.reshape(-1), _i(nn), N.reshape(-1), len(N), int(scheme), tier, nv, L)
    return tier, nv, L


def cpu32_lloyd(mesh: Mesh, S, K: int, R: float, iters: int, threads: int) -> np.ndarray:
    """`iters` FP32 Lloyd iterations of the sites S on `threads` pinned host
    threads (cvt_cpu32.cpp, the CPU speed baseline); returns the new sites."""
    global _cpu32
    if _cpu32 is None:
        if not os.path.exists(_CPU32_PATH):
            build()
        L = C.CDLL(_CPU32_PATH)
        fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        L.cpu32_lloyd.argtypes = [C.c_int, fp, C.c_int, _ip, fp, _ip, _ip, C.c_double, C.c_int, fp, C.c_int,
                                  C.c_float, C.c_int, C.c_int]
        _cpu32 = L
    V = np.ascontiguousarray(mesh.V, np.float32)
    N = np.ascontiguousarray(mesh.N, np.float32)
    off, faces = mesh.vf
    S32 = np.ascontiguousarray(np.asarray(S, np.float32)).copy()
    e = mesh.V[mesh.F]
    area = float(0.5 * np.linalg.norm(np.cross(e[:, 1] - e[:, 0], e[:, 2] - e[:, 0]), axis=1).sum())
    rc = _cpu32.cpu32_lloyd(len(V), V.reshape(-1), len(mesh.F), mesh.F.reshape(-1), N.reshape(-1), off, faces, area,
                            len(S32), S32.reshape(-1), int(K), float(R), int(iters), int(threads))
    if rc != 0:
        raise RuntimeError(f"cpu32_lloyd failed ({rc})")
    return S32.astype(np.float64)


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), or '' when unknown."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""

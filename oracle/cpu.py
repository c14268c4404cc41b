"""ctypes wrapper of oracle/_build/libhpm_oracle.so (oracle/hpm_oracle.h).

TEST INFRASTRUCTURE: the CPU checker of the HIP path. Only tests/, the smoke
check and bench.py's CPU-baseline leg use it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional, Sequence

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "_build", "libhpm_oracle.so")
MAX_IMAGES = 33


def _structs():
    import importlib.util
    import sys

    name = "hpm_mvs_amd"
    if name not in sys.modules:
        d = os.path.join(os.path.dirname(ORACLE_DIR), "hpm-mvs_amd")
        spec = importlib.util.spec_from_file_location(name, os.path.join(d, "__init__.py"),
                                                      submodule_search_locations=[d])
        mod = importlib.util.module_from_spec(spec)
        sys.modules[name] = mod
        spec.loader.exec_module(mod)
    from hpm_mvs_amd import hpm  # noqa: E402

    return hpm


_hpm = _structs()
HpmCamera = _hpm.HpmCamera
HpmParams = _hpm.HpmParams

_FP = C.POINTER(C.c_float)
_UP = C.POINTER(C.c_uint32)
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)


COST_FN = C.CFUNCTYPE(C.c_float, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, _FP)


class OracleTrace(C.Structure):
    """hpm_oracle_trace: one pixel's propagation decisions."""

    _fields_ = [
        ("px", C.c_int32), ("py", C.c_int32), ("hit", C.c_int32),
        ("flag", C.c_int32 * 8), ("pos_scan", C.c_int32 * 8), ("judge", C.c_int32 * 3), ("pos", C.c_int32 * 8),
        ("ca", (C.c_float * 32) * 8), ("vprior", C.c_float * 32), ("pdf", C.c_float * 32),
        ("cdf", C.c_float * 32), ("draws", C.c_float * 15), ("vw", C.c_float * 32), ("tsel", C.c_uint32),
        ("wn", C.c_float), ("fc", C.c_float * 8), ("min_idx", C.c_int32), ("cost_now", C.c_float),
        ("rfc", C.c_float * 8), ("max_idx", C.c_int32), ("rcn", C.c_float), ("adopted", C.c_int32),
        ("nan_fallback", C.c_int32), ("restricted_cost", C.c_float), ("cand", (C.c_float * 4) * 5),
        ("cand_depth", C.c_float * 5), ("cand_cost", C.c_float * 5), ("cand_rcost", C.c_float * 5),
        ("cand_accepted", C.c_int32 * 5),
    ]


class OracleProblem(C.Structure):
    _fields_ = [
        ("num_images", C.c_int32),
        ("cams", HpmCamera * MAX_IMAGES),
        ("images", _FP * MAX_IMAGES),
        ("depths", _FP * MAX_IMAGES),
        ("params", HpmParams),
        ("prior_planes", _FP),
        ("prior_mask", _UP),
        ("call_index", C.c_uint32),
        ("salt", C.c_uint32),
        ("coarse", _FP),
        ("coarse_w", C.c_int32),
        ("coarse_h", C.c_int32),
        ("pre_cost", _FP),
        ("cost_hook", COST_FN),
        ("cost_hook_user", C.c_void_p),
        ("trace", C.POINTER(OracleTrace)),
    ]


_lib = None


def build():
    """Compiles the oracle with make (g++ -fopenmp)."""
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(OracleProblem)
        U64 = C.POINTER(C.c_uint64)
        L.hpm_oracle_run_patchmatch.argtypes = [P, _FP, _FP, _UP, U64, C.c_int]
        L.hpm_oracle_init.argtypes = [P, _FP, _FP, _UP, U64, C.c_int]
        L.hpm_oracle_half_iteration.argtypes = [P, C.c_int, C.c_int, _FP, _FP, _UP, U64, C.c_int]
        L.hpm_oracle_half_iteration_rows.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, _FP, _FP, _UP, U64,
                                                     C.c_int]
        L.hpm_oracle_finalize.argtypes = [P, _FP, C.c_int]
        L.hpm_oracle_filter.argtypes = [P, C.c_int, _FP, _FP, C.c_int]
        L.hpm_oracle_cost_vectors.argtypes = [P, _FP, _IP, C.c_int, C.c_int, _DP, C.c_int]
        L.hpm_oracle_jbu.argtypes = [_FP, C.c_int, C.c_int, _FP, C.c_int, C.c_int, _FP, _FP, C.c_int]
        for fn in (L.hpm_oracle_run_patchmatch, L.hpm_oracle_init, L.hpm_oracle_half_iteration,
                   L.hpm_oracle_finalize, L.hpm_oracle_filter, L.hpm_oracle_cost_vectors, L.hpm_oracle_jbu):
            fn.restype = C.c_int
        _lib = L
    return _lib


def _f(a):
    return a.ctypes.data_as(_FP)


def relaxed_params(cams: Sequence, base=None):
    """Depth range relaxed as in HPM::InuputInitialization (HPM.cpp:553-556)."""
    p = base if base is not None else _hpm.default_params()
    q = HpmParams()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(HpmParams))
    q.depth_min = float(np.float32(cams[0].depth_min) * np.float32(0.6))
    q.depth_max = float(np.float32(cams[0].depth_max) * np.float32(1.2))
    q.num_images = len(cams)
    return q


class Oracle:
    """One problem on the CPU, with the same state layout as the device."""

    def __init__(self, cams: Sequence, images: Sequence[np.ndarray], params,
                 depths: Optional[Sequence[np.ndarray]] = None, salt: int = 0, threads: int = 0):
        self.pb = OracleProblem()
        self.pb.num_images = len(cams)
        self._keep = []
        for i, (c, im) in enumerate(zip(cams, images)):
            self.pb.cams[i] = c
            a = np.ascontiguousarray(im, dtype=np.float32)
            self._keep.append(a)
            self.pb.images[i] = _f(a)
            if depths is not None:
                d = np.ascontiguousarray(depths[i], dtype=np.float32)
                self._keep.append(d)
                self.pb.depths[i] = _f(d)
        self.pb.params = params
        self.pb.salt = salt
        self.pb.call_index = 0
        self.threads = threads
        W, H = cams[0].width, cams[0].height
        self.W, self.H = W, H
        self.hyp = np.zeros((H, W, 4), np.float32)
        self.cost = np.zeros((H, W), np.float32)
        self.sel = np.zeros((H, W), np.uint32)
        self.n_eval = 0

    def set_prior(self, planes: np.ndarray, mask: np.ndarray):
        pl = np.ascontiguousarray(planes, dtype=np.float32).copy()
        mk = np.ascontiguousarray(mask, dtype=np.uint32)
        pl[mk == 0] = 0.0
        self._keep += [pl, mk]
        self.pb.prior_planes = _f(pl)
        self.pb.prior_mask = mk.ctypes.data_as(_UP)

    def set_hierarchy(self, coarse_world4: np.ndarray):
        """ACMM hierarchy: start from the coarser map (world normal xyz,
        depth per coarse pixel); pre_cost receives the pre-costs."""
        c = np.ascontiguousarray(coarse_world4, dtype=np.float32)
        self.pre_cost = np.zeros((self.H, self.W), np.float32)
        self._keep += [c, self.pre_cost]
        self.pb.coarse = _f(c)
        self.pb.coarse_h, self.pb.coarse_w = c.shape[:2]
        self.pb.pre_cost = _f(self.pre_cost)
        self.pb.params.hierarchy = 1

    def set_cost_hook(self, fn):
        """Replaces the NCC by fn(px, py, view, plane) -> float (micro-scenes;
        the oracle then runs on one thread)."""

        def cb(_user, px, py, view, plane):
            return float(fn(px, py, view, np.ctypeslib.as_array(plane, (4,)).copy()))

        self._hook = COST_FN(cb)
        self.pb.cost_hook = self._hook
        self.threads = 1

    def trace_pixel(self, x: int, y: int) -> OracleTrace:
        """Records the decisions of pixel (x, y) in the following
        half-iterations."""
        self._trace = OracleTrace()
        self._trace.px, self._trace.py = x, y
        self.pb.trace = C.pointer(self._trace)
        return self._trace

    def _state(self):
        return (_f(self.hyp), _f(self.cost), self.sel.ctypes.data_as(_UP))

    def run_patchmatch(self):
        n = C.c_uint64(0)
        rc = lib().hpm_oracle_run_patchmatch(C.byref(self.pb), *self._state(), C.byref(n), self.threads)
        if rc != 0:
            raise RuntimeError("oracle run_patchmatch failed")
        self.n_eval = n.value
        self.pb.call_index += 1

    def init(self):
        n = C.c_uint64(0)
        assert lib().hpm_oracle_init(C.byref(self.pb), *self._state(), C.byref(n), self.threads) == 0
        return n.value

    def half_iteration(self, it: int, colour: int):
        n = C.c_uint64(0)
        assert lib().hpm_oracle_half_iteration(C.byref(self.pb), it, colour, *self._state(), C.byref(n),
                                               self.threads) == 0
        return n.value

    def half_iteration_rows(self, it: int, colour: int, y0: int, y1: int):
        n = C.c_uint64(0)
        assert lib().hpm_oracle_half_iteration_rows(C.byref(self.pb), it, colour, y0, y1, *self._state(),
                                                    C.byref(n), self.threads) == 0
        return n.value

    def finalize(self):
        assert lib().hpm_oracle_finalize(C.byref(self.pb), _f(self.hyp), self.threads) == 0

    def filter(self, colour: int):
        assert lib().hpm_oracle_filter(C.byref(self.pb), colour, _f(self.hyp), _f(self.cost), self.threads) == 0

    def cost_vectors(self, planes: np.ndarray, pix: np.ndarray, fp64) -> np.ndarray:
        """fp64: False/0 the HIP path's FP32 arithmetic, True/1 the FP64
        reference formula, 2 the reference's FP32 order, 3 the FP32 warp."""
        pl = np.ascontiguousarray(planes, dtype=np.float32).reshape(-1, 4)
        px = np.ascontiguousarray(pix, dtype=np.int32).reshape(-1, 2)
        out = np.empty((len(pl), self.pb.num_images - 1), np.float64)
        rc = lib().hpm_oracle_cost_vectors(C.byref(self.pb), _f(pl), px.ctypes.data_as(_IP), len(pl), int(fp64),
                                           out.ctypes.data_as(_DP), self.threads)
        assert rc == 0
        return out


def fusion(cams, depths, normals, colors, problems, point_dtype, acmm=False):
    """RunFusion (main.cpp:502-633; with acmm, ACMM+NESP's, main.cpp:164-300),
    sequential: depths[i] (H, W), normals[i] (H, W, 3) world, colors[i]
    (H, W, 3) uint8 of view i; problems[i] = (i, [sources]). Returns the
    kept points as point_dtype."""
    L = lib()
    n = len(cams)
    max_src = 32
    ns = np.array([len(p[1]) for p in problems], np.int32)
    srcs = np.zeros((n, max_src), np.int32)
    for i, (_, s) in enumerate(problems):
        srcs[i, : len(s)] = s
    keep = [np.ascontiguousarray(d, dtype=np.float32) for d in depths]
    keep_n = [np.ascontiguousarray(x, dtype=np.float32) for x in normals]
    keep_c = [np.ascontiguousarray(x, dtype=np.uint8) for x in colors]
    dp = (_FP * n)(*[_f(a) for a in keep])
    np_ = (_FP * n)(*[_f(a) for a in keep_n])
    cp = (C.POINTER(C.c_uint8) * n)(*[a.ctypes.data_as(C.POINTER(C.c_uint8)) for a in keep_c])
    cam_arr = (type(cams[0]) * n)(*cams)
    cap = sum(int(d.size) for d in depths)
    out = np.empty(cap, point_dtype)
    fn = L.hpm_oracle_fusion_acmm if acmm else L.hpm_oracle_fusion
    fn.restype = C.c_int64
    fn.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, _IP, _IP, C.c_int, C.c_void_p,
                   C.c_int64]
    k = fn(n, C.cast(cam_arr, C.c_void_p), C.cast(dp, C.c_void_p), C.cast(np_, C.c_void_p),
                            C.cast(cp, C.c_void_p), ns.ctypes.data_as(_IP), srcs.ctypes.data_as(_IP), max_src,
                            C.c_void_p(out.ctypes.data), cap)
    return out[:k]


def jbu(guide: np.ndarray, src_nd: np.ndarray, threads: int = 0):
    g = np.ascontiguousarray(guide, dtype=np.float32)
    s = np.ascontiguousarray(src_nd, dtype=np.float32)
    h, w = g.shape
    sh, sw = s.shape[:2]
    od = np.empty((h, w), np.float32)
    on = np.empty((h, w, 4), np.float32)
    rc = lib().hpm_oracle_jbu(_f(g), w, h, _f(s), sw, sh, _f(od), _f(on), threads)
    return None if rc == 1 else (od, on)

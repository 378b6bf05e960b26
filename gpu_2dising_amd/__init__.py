"""MI355X-native replica-parallel Metropolis engine for the 2D Ising model.

Drop-in for GASP's GPU engine (gpu_method 0/1/2).  The engine is the C-ABI
library ``libgasp_amd.so`` (declared in ``include/gasp_amd.h``) with
hand-written gfx950 HIP kernels; this package holds its sources (``csrc/``),
its build script (``build.py``), the ``gasp`` command line program
(``bin/gasp``), the PyGASP-compatible extension module ``gasp`` and this
ctypes binding of the C ABI.

Nothing here falls back to a CPU path: if the library is missing, ``lib()``
raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libgasp_amd.so")
CLI_PATH = os.path.join(PKG_DIR, "bin", "gasp")

# Every symbol include/gasp_amd.h declares.
EXPORTS = (
    "gpuInitDevice", "gpuInitGrid", "gpuInitRand", "preComputeProbs_gpu", "select_gpu_method",
    "mc_driver_gpu", "init_grids_uniform", "init_grids_from_file", "init_grids_from_array",
    "init_fates", "read_input_grid", "write_ising_grids", "bootstrap_errbar",
    "init_genrand", "genrand_int32", "genrand_real3", "preComputeProbs_cpu", "mc_sweep_cpu",
    "compute_magnetisation_cpu", "compute_largest_cluster_cpu", "mc_driver_cpu",
    "gasp_last_error", "gasp_error_string", "gasp_clear_error", "gasp_init_rand_mt", "gasp_free",
    "gasp_set_tuning", "gasp_get_tuning", "gasp_get_run_stats", "gasp_set_conflict_filter",
    "gasp_engine_get_conflict_filter", "gasp_set_lattice_staging", "gasp_engine_get_lattice_staging",
    "gasp_set_dependence_tables",
    "gasp_engine_set_attempt",
    "gasp_engine_create", "gasp_engine_destroy", "gasp_engine_set_grids", "gasp_engine_get_grids",
    "gasp_engine_gather_grids", "gasp_engine_get_gridstates",
    "gasp_engine_set_inputs", "gasp_engine_set_freeze",
    "gasp_engine_sweep", "gasp_engine_sync", "gasp_engine_get_magnetisation",
    "gasp_engine_get_counts", "gasp_engine_set_fates", "gasp_engine_get_fates",
    "gasp_engine_largest_cluster", "gasp_engine_kernel_ms", "gasp_engine_stream", "gasp_engine_get_shape",
    "gasp_sweep_grids",
    "gasp_sweep_grids_noalloc",
    "gasp_graph_create", "gasp_graph_launch", "gasp_graph_ms", "gasp_graph_destroy", "gasp_autotune",
    "gasp_bootstrap_errbar_gpu", "gasp_bootstrap_errbar_libc",
    "gasp_cnn_num_params", "gasp_cnn_create", "gasp_cnn_destroy", "gasp_cnn_set_params",
    "gasp_cnn_get_params", "gasp_cnn_predict", "gasp_cnn_gradient", "gasp_cnn_train_step",
    "gasp_cnn_set_data", "gasp_cnn_train_epoch",
)

CHECK_NONE, CHECK_COUNT, CHECK_FATES = 0, 1, 2


class GaspError(RuntimeError):
    pass


# ------------------------------------------------------------------ C types
class McGrids(C.Structure):  # mc_grids_t
    _fields_ = [("L", C.c_int), ("ngrids", C.c_int), ("ising_grids", C.POINTER(C.c_int))]


class McSampler(C.Structure):  # mc_sampler_t
    _fields_ = [("tot_nsweeps", C.c_int), ("mag_output_int", C.c_int),
                ("grid_output_int", C.c_int)]


class McFunction(C.Structure):  # mc_function_t
    _fields_ = [("itask", C.c_int), ("dn_thr", C.c_double), ("up_thr", C.c_double),
                ("ninputs", C.c_int), ("initial_spin", C.c_int), ("cv", C.c_char_p),
                ("result", C.POINTER(C.c_float)), ("filename", C.c_char_p)]


class McGpuGrids(C.Structure):  # mc_gpu_grids_t
    _fields_ = [("L", C.c_int), ("ngrids", C.c_int), ("ising_grids", C.POINTER(C.c_int)),
                ("d_ising_grids", C.c_void_p), ("d_neighbour_list", C.c_void_p)]


class GpuRun(C.Structure):  # gpu_run_t
    _fields_ = [("d_state", C.c_void_p), ("threadsPerBlock", C.c_int), ("gpu_method", C.c_int)]


class RunStats(C.Structure):  # gasp_run_stats_t
    _fields_ = [("sweep_kernel_ms", C.c_double), ("attempts_per_replica", C.c_double),
                ("launches", C.c_int), ("slots_per_replica", C.c_int),
                ("waves_per_group", C.c_int), ("method", C.c_int)]


GridOutputFunc = C.CFUNCTYPE(C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int,
                             C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_char_p,
                             C.c_double, C.c_double, C.c_char_p)

_lib = None


def lib() -> C.CDLL:
    """The engine library with argument types set.  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GaspError(f"{LIB_PATH} is missing: build it with `python -m gpu_2dising_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, ip, fp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_float)
    L.gpuInitDevice.argtypes = [C.c_int, ip]
    L.gpuInitDevice.restype = C.c_int
    L.gpuInitGrid.argtypes = [C.c_int, C.c_int, C.c_int, ip, C.POINTER(vp), C.POINTER(vp)]
    L.gpuInitRand.argtypes = [C.c_int, C.c_int, C.c_ulong, C.POINTER(vp)]
    L.preComputeProbs_gpu.argtypes = [C.c_double, C.c_double]
    L.select_gpu_method.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    L.select_gpu_method.restype = C.c_int
    L.mc_driver_gpu.argtypes = [McGpuGrids, C.c_double, C.c_double, ip, McSampler, McFunction,
                                GpuRun, GridOutputFunc]
    L.init_grids_uniform.argtypes = [C.c_int, C.c_int, C.c_int]
    L.init_grids_uniform.restype = ip
    L.init_grids_from_file.argtypes = [C.c_int, C.c_int]
    L.init_grids_from_file.restype = ip
    L.init_grids_from_array.argtypes = [C.c_int, C.c_int, C.c_int, ip]
    L.init_grids_from_array.restype = ip
    L.init_fates.argtypes = [C.c_int]
    L.init_fates.restype = ip
    L.read_input_grid.argtypes = [C.c_int, C.c_int, ip]
    L.write_ising_grids.argtypes = [C.c_int, C.c_int, ip, C.c_int, fp, fp, C.c_char_p,
                                    C.c_double, C.c_double, C.c_char_p]
    L.write_ising_grids.restype = C.c_int
    L.bootstrap_errbar.argtypes = [C.c_int, C.c_int, ip]
    L.bootstrap_errbar.restype = C.c_double
    L.gasp_last_error.restype = C.c_int
    L.gasp_error_string.restype = C.c_char_p
    L.gasp_init_rand_mt.argtypes = [C.c_int, C.c_ulong, C.POINTER(vp)]
    L.gasp_init_rand_mt.restype = C.c_int
    L.gasp_free.argtypes = [vp]
    L.gasp_set_tuning.argtypes = [C.c_int, C.c_int]
    L.gasp_set_conflict_filter.argtypes = [C.c_int]
    L.gasp_set_lattice_staging.argtypes = [C.c_int]
    L.gasp_set_dependence_tables.argtypes = [C.c_int]
    L.gasp_get_tuning.argtypes = [C.c_int, C.c_int, C.c_int, ip, ip]
    L.gasp_get_run_stats.argtypes = [C.POINTER(RunStats)]
    L.gasp_engine_create.argtypes = [C.c_int, C.c_int, C.c_int, vp, C.c_double, C.c_double,
                                     C.POINTER(vp)]
    for name in ("gasp_engine_create", "gasp_engine_destroy", "gasp_engine_set_grids",
                 "gasp_engine_get_grids", "gasp_engine_sweep", "gasp_engine_sync",
                 "gasp_engine_get_magnetisation", "gasp_engine_get_counts",
                 "gasp_engine_set_fates", "gasp_engine_get_fates",
                 "gasp_engine_largest_cluster", "gasp_engine_set_inputs", "gasp_engine_set_freeze"):
        getattr(L, name).restype = C.c_int
    L.gasp_engine_destroy.argtypes = [vp]
    L.gasp_engine_set_grids.argtypes = [vp, vp, C.c_int]
    L.gasp_engine_get_grids.argtypes = [vp, vp, C.c_int]
    L.gasp_engine_gather_grids.argtypes = [vp, vp, C.c_int, vp]
    L.gasp_engine_gather_grids.restype = C.c_int
    L.gasp_engine_get_gridstates.argtypes = [vp, vp]
    L.gasp_engine_get_gridstates.restype = C.c_int
    L.gasp_engine_set_inputs.argtypes = [vp, vp, C.c_int]
    L.gasp_engine_set_freeze.argtypes = [vp, C.c_int]
    L.gasp_engine_sweep.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_double]
    L.gasp_engine_sync.argtypes = [vp]
    L.gasp_engine_get_magnetisation.argtypes = [vp, fp]
    L.gasp_engine_get_counts.argtypes = [vp, ip]
    L.gasp_engine_set_fates.argtypes = [vp, ip]
    L.gasp_engine_get_fates.argtypes = [vp, ip]
    L.gasp_engine_largest_cluster.argtypes = [vp, C.c_int, fp]
    L.gasp_engine_kernel_ms.argtypes = [vp]
    L.gasp_engine_kernel_ms.restype = C.c_double
    L.gasp_engine_get_shape.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.gasp_engine_get_conflict_filter.argtypes = [vp, C.POINTER(C.c_int)]
    L.gasp_engine_get_lattice_staging.argtypes = [vp, C.POINTER(C.c_int)]
    L.gasp_engine_set_attempt.argtypes = [vp, C.c_ulonglong]
    L.gasp_engine_stream.argtypes = [vp]
    L.gasp_engine_stream.restype = vp
    L.gasp_sweep_grids.argtypes = [C.c_int, C.c_int, C.c_int, vp, C.c_double, C.c_double, vp, C.c_int, vp, vp]
    L.gasp_sweep_grids.restype = C.c_int
    L.gasp_sweep_grids_noalloc.argtypes = L.gasp_sweep_grids.argtypes
    L.gasp_sweep_grids_noalloc.restype = C.c_int
    dp = C.POINTER(C.c_double)
    L.gasp_graph_create.argtypes = [C.POINTER(vp), C.c_int, ip, ip, ip, dp, dp, C.POINTER(vp)]
    L.gasp_graph_create.restype = C.c_int
    L.gasp_graph_launch.argtypes = [vp, C.c_int]
    L.gasp_graph_launch.restype = C.c_int
    L.gasp_graph_ms.argtypes = [vp]
    L.gasp_graph_ms.restype = C.c_double
    L.gasp_graph_destroy.argtypes = [vp]
    L.gasp_graph_destroy.restype = C.c_int
    L.gasp_autotune.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, ip, ip, ip,
                                dp]
    L.gasp_autotune.restype = C.c_int
    L.gasp_bootstrap_errbar_gpu.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong, vp]
    L.gasp_bootstrap_errbar_gpu.restype = C.c_int
    L.gasp_bootstrap_errbar_libc.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_uint, C.c_int, vp]
    L.gasp_bootstrap_errbar_libc.restype = C.c_int
    i8p = C.POINTER(C.c_int8)
    for name in ("gasp_cnn_num_params", "gasp_cnn_create", "gasp_cnn_destroy", "gasp_cnn_set_params",
                 "gasp_cnn_get_params", "gasp_cnn_predict", "gasp_cnn_gradient", "gasp_cnn_train_step",
                 "gasp_cnn_set_data", "gasp_cnn_train_epoch"):
        getattr(L, name).restype = C.c_int
    L.gasp_cnn_create.argtypes = [C.c_int, C.c_int, C.POINTER(vp)]
    L.gasp_cnn_destroy.argtypes = [vp]
    L.gasp_cnn_set_params.argtypes = [vp, fp]
    L.gasp_cnn_get_params.argtypes = [vp, fp]
    L.gasp_cnn_predict.argtypes = [vp, i8p, C.c_int, fp]
    L.gasp_cnn_gradient.argtypes = [vp, i8p, fp, C.c_int, fp, fp]
    L.gasp_cnn_train_step.argtypes = [vp, i8p, fp, C.c_int, C.c_float, fp]
    L.gasp_cnn_set_data.argtypes = [vp, i8p, fp, C.c_int]
    L.gasp_cnn_train_epoch.argtypes = [vp, ip, C.c_int, C.c_int, C.c_float, fp]
    _lib = L
    return L


def check(rc: int = None) -> None:
    """Raise GaspError if the library recorded an error."""
    code = lib().gasp_last_error() if rc is None else rc
    if code:
        raise GaspError(f"gasp error {code}: {lib().gasp_error_string().decode()}")


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


# ------------------------------------------------------------------ helpers
class RngState:
    """Owns a gasp_rng_state (Philox by default, MT19937 parity streams if mt=True)."""

    def __init__(self, ngrids: int, seed: int, mt: bool = False):
        self.ptr = C.c_void_p()
        if mt:
            check(lib().gasp_init_rand_mt(ngrids, seed, C.byref(self.ptr)))
        else:
            lib().gpuInitRand(ngrids, 64, seed, C.byref(self.ptr))
            check()

    def free(self):
        if self.ptr:
            lib().gasp_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Engine:
    """Stream-level engine over the C ABI (gasp_engine_*)."""

    def __init__(self, L: int, ngrids: int, method: int, rng: RngState, beta: float, h: float):
        self.L, self.R, self.method = L, ngrids, method
        self.rng = rng
        self.ptr = C.c_void_p()
        check(lib().gasp_engine_create(L, ngrids, method, rng.ptr, beta, h, C.byref(self.ptr)))

    def set_grids(self, grids: np.ndarray) -> None:
        g = np.ascontiguousarray(grids, dtype=np.int32).reshape(self.R, self.L * self.L)
        check(lib().gasp_engine_set_grids(self.ptr, g.ctypes.data, 0))

    def set_inputs(self, inputs: np.ndarray) -> None:
        """Input grids [n, L, L]; replica r starts from input r // (R // n)."""
        g = np.ascontiguousarray(inputs, dtype=np.int32).reshape(-1, self.L * self.L)
        check(lib().gasp_engine_set_inputs(self.ptr, g.ctypes.data, g.shape[0]))

    def set_freeze(self, on: bool) -> None:
        check(lib().gasp_engine_set_freeze(self.ptr, 1 if on else 0))

    def get_grids(self) -> np.ndarray:
        out = np.empty((self.R, self.L, self.L), dtype=np.int32)
        check(lib().gasp_engine_get_grids(self.ptr, out.ctypes.data, 0))
        return out

    def gridstates(self, out: np.ndarray = None) -> np.ndarray:
        """All grids as a gridstates.bin record body (uint8, ceil(R*L*L/8) bytes)."""
        n = (self.R * self.L * self.L + 7) // 8
        if out is None:
            out = np.empty(n, dtype=np.uint8)
        assert out.dtype == np.uint8 and out.size == n and out.flags.c_contiguous
        check(lib().gasp_engine_get_gridstates(self.ptr, out.ctypes.data))
        return out

    def gather(self, replicas) -> np.ndarray:
        """int8 grids [n, L, L] of the listed replicas."""
        idx = np.ascontiguousarray(replicas, dtype=np.int32)
        out = np.empty((idx.size, self.L, self.L), dtype=np.int8)
        check(lib().gasp_engine_gather_grids(self.ptr, idx.ctypes.data, idx.size, out.ctypes.data))
        return out

    def sweep(self, nsweeps: int, check_mode: int = CHECK_NONE, dn: float = -0.9,
              up: float = 0.9) -> None:
        check(lib().gasp_engine_sweep(self.ptr, nsweeps, check_mode, dn, up))

    def sync(self) -> None:
        check(lib().gasp_engine_sync(self.ptr))

    def magnetisation(self) -> np.ndarray:
        out = np.empty(self.R, dtype=np.float32)
        check(lib().gasp_engine_get_magnetisation(self.ptr, _fp(out)))
        return out

    def counts(self) -> np.ndarray:
        out = np.zeros(3, dtype=np.int32)
        check(lib().gasp_engine_get_counts(self.ptr, _ip(out)))
        return out

    def set_fates(self, fates: np.ndarray) -> None:
        f = np.ascontiguousarray(fates, dtype=np.int32)
        check(lib().gasp_engine_set_fates(self.ptr, _ip(f)))

    def fates(self) -> np.ndarray:
        out = np.empty(self.R, dtype=np.int32)
        check(lib().gasp_engine_get_fates(self.ptr, _ip(out)))
        return out

    def largest_cluster(self, spin: int) -> np.ndarray:
        out = np.empty(self.R, dtype=np.float32)
        check(lib().gasp_engine_largest_cluster(self.ptr, spin, _fp(out)))
        return out

    def kernel_ms(self) -> float:
        return lib().gasp_engine_kernel_ms(self.ptr)

    def set_attempt(self, attempt: int) -> None:
        """Move the replicas to attempt `attempt` of their Philox streams (rewind or skip)."""
        check(lib().gasp_engine_set_attempt(self.ptr, attempt))

    def shape(self) -> tuple:
        """(slots per replica, waves per workgroup or word split) of the engine's launches."""
        s, w = C.c_int(0), C.c_int(0)
        check(lib().gasp_engine_get_shape(self.ptr, C.byref(s), C.byref(w)))
        return s.value, w.value

    def lattice_staging(self) -> int:
        """The engine's lattice staging: 1 in HBM, 2 staged in LDS as bytes, 3 as bits (method 0 over several waves)."""
        f = C.c_int(0)
        check(lib().gasp_engine_get_lattice_staging(self.ptr, C.byref(f)))
        return f.value

    def conflict_filter(self) -> int:
        """The engine's conflict filter: 1 row/column masks, 2 mark plane."""
        f = C.c_int(0)
        check(lib().gasp_engine_get_conflict_filter(self.ptr, C.byref(f)))
        return f.value

    def close(self) -> None:
        if self.ptr:
            check(lib().gasp_engine_destroy(self.ptr))
            self.ptr = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Surrogate:
    """Committor-surrogate CNN over the C ABI (gasp_cnn_*)."""

    def __init__(self, L: int = 64, max_batch: int = 256):
        self.L, self.max_batch = L, max_batch
        self.ptr = C.c_void_p()
        check(lib().gasp_cnn_create(L, max_batch, C.byref(self.ptr)))
        self.nparam = lib().gasp_cnn_num_params()

    @staticmethod
    def _spins(spins):
        return np.ascontiguousarray(spins, dtype=np.int8)

    def set_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == self.nparam
        check(lib().gasp_cnn_set_params(self.ptr, _fp(p)))

    def params(self) -> np.ndarray:
        out = np.empty(self.nparam, dtype=np.float32)
        check(lib().gasp_cnn_get_params(self.ptr, _fp(out)))
        return out

    def predict(self, spins: np.ndarray) -> np.ndarray:
        s = self._spins(spins)
        out = np.empty(s.shape[0], dtype=np.float32)
        check(lib().gasp_cnn_predict(self.ptr, s.ctypes.data_as(C.POINTER(C.c_int8)), s.shape[0], _fp(out)))
        return out

    def gradient(self, spins: np.ndarray, labels: np.ndarray):
        s = self._spins(spins)
        y = np.ascontiguousarray(labels, dtype=np.float32)
        g = np.empty(self.nparam, dtype=np.float32)
        loss = C.c_float()
        check(lib().gasp_cnn_gradient(self.ptr, s.ctypes.data_as(C.POINTER(C.c_int8)), _fp(y), s.shape[0],
                                      _fp(g), C.byref(loss)))
        return loss.value, g

    def train_step(self, spins: np.ndarray, labels: np.ndarray, lr: float) -> float:
        s = self._spins(spins)
        y = np.ascontiguousarray(labels, dtype=np.float32)
        loss = C.c_float()
        check(lib().gasp_cnn_train_step(self.ptr, s.ctypes.data_as(C.POINTER(C.c_int8)), _fp(y), s.shape[0],
                                        lr, C.byref(loss)))
        return loss.value

    def set_data(self, spins: np.ndarray, labels: np.ndarray) -> None:
        s = self._spins(spins)
        y = np.ascontiguousarray(labels, dtype=np.float32)
        check(lib().gasp_cnn_set_data(self.ptr, s.ctypes.data_as(C.POINTER(C.c_int8)), _fp(y), s.shape[0]))

    def train_epoch(self, order: np.ndarray, batch: int, lr: float) -> float:
        o = np.ascontiguousarray(order, dtype=np.int32)
        loss = C.c_float()
        check(lib().gasp_cnn_train_epoch(self.ptr, _ip(o), o.size, batch, lr, C.byref(loss)))
        return loss.value

    def close(self) -> None:
        if self.ptr:
            lib().gasp_cnn_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Graph:
    """One captured hipGraph advancing several engines concurrently."""

    def __init__(self, engines, nsweeps, checks, dn, up, launches=None):
        n = len(engines)
        arr = (C.c_void_p * n)(*[e.ptr.value for e in engines])
        ns = np.ascontiguousarray(nsweeps, dtype=np.int32)
        ck = np.ascontiguousarray(checks, dtype=np.int32)
        lc = None if launches is None else np.ascontiguousarray(launches, dtype=np.int32)
        d = np.ascontiguousarray(dn, dtype=np.float64)
        u = np.ascontiguousarray(up, dtype=np.float64)
        self.engines = engines
        self.ptr = C.c_void_p()
        check(lib().gasp_graph_create(arr, n, _ip(ns), None if lc is None else _ip(lc), _ip(ck),
                                      d.ctypes.data_as(C.POINTER(C.c_double)),
                                      u.ctypes.data_as(C.POINTER(C.c_double)), C.byref(self.ptr)))

    def launch(self, replays: int = 1) -> None:
        check(lib().gasp_graph_launch(self.ptr, replays))

    def ms(self) -> float:
        return lib().gasp_graph_ms(self.ptr)

    def close(self) -> None:
        if self.ptr:
            lib().gasp_graph_destroy(self.ptr)
            self.ptr = C.c_void_p()


def autotune(method: int, L: int, ngrids: int, probe_sweeps: int = 5, beta: float = 0.54, h: float = 0.07,
             check_sweeps: int = 0):
    """gasp_autotune: (slots, waves or word split, sweeps per launch, attempts/s)."""
    s, w, k, r = C.c_int(), C.c_int(), C.c_int(), C.c_double()
    check(lib().gasp_autotune(method, L, ngrids, beta, h, probe_sweeps, check_sweeps, C.byref(s), C.byref(w),
                              C.byref(k), C.byref(r)))
    return s.value, w.value, k.value, r.value


def bootstrap_gpu(fates: np.ndarray, m: int, seed: int = 1) -> np.ndarray:
    """GPU bootstrap error bars of pB for fates [ninputs, n] (-1/0/1)."""
    f = np.ascontiguousarray(fates, dtype=np.int32)
    if f.ndim != 2:
        raise ValueError("fates must be [ninputs, n]")
    out = np.empty(f.shape[0], dtype=np.float64)
    check(lib().gasp_bootstrap_errbar_gpu(f.ctypes.data, 0, f.shape[0], f.shape[1], m, seed, out.ctypes.data))
    return out


def bootstrap_libc(fates: np.ndarray, m: int, seed: int = 1, nthreads: int = 0) -> np.ndarray:
    """bootstrap_errbar (src/bootstrap.c:6-63) for fates [ninputs, n], bit for
    bit as successive calls after srand(seed) with glibc's rand(), on host
    threads (the library jumps its own copy of the rand() sequence)."""
    f = np.ascontiguousarray(fates, dtype=np.int32)
    if f.ndim != 2:
        raise ValueError("fates must be [ninputs, n]")
    out = np.empty(f.shape[0], dtype=np.float64)
    check(lib().gasp_bootstrap_errbar_libc(f.ctypes.data, f.shape[0], f.shape[1], m, seed, nthreads,
                                           out.ctypes.data))
    return out


def set_tuning(slots: int = 0, waves: int = 0) -> None:
    lib().gasp_set_tuning(slots, waves)


def set_dependence_tables(max_entries: int = 0) -> None:
    """Method 0 over several waves: at most `max_entries` entries per LDS
    dependence table for engines created afterwards (0 = up to 16384)."""
    lib().gasp_set_dependence_tables(max_entries)


def set_lattice_staging(staging: int = 0) -> None:
    """Method 0 over several waves: 1 lattice in HBM, 2 staged in LDS as bytes, 3 as bits (0 = automatic)."""
    lib().gasp_set_lattice_staging(staging)


def set_conflict_filter(filter: int = 0) -> None:
    """Conflict filter of method 2 on split words: 1 row/column masks, 2 mark plane (0 = automatic)."""
    lib().gasp_set_conflict_filter(filter)


def get_tuning(method: int, L: int, ngrids: int) -> tuple[int, int]:
    s, w = C.c_int(), C.c_int()
    lib().gasp_get_tuning(method, L, ngrids, C.byref(s), C.byref(w))
    return s.value, w.value


def run_stats() -> RunStats:
    st = RunStats()
    lib().gasp_get_run_stats(C.byref(st))
    return st


def init_device(index: int = -1) -> int:
    """gpuInitDevice; returns the CU count.  Raises if no GPU is visible."""
    n = C.c_int(0)
    dev = lib().gpuInitDevice(index, C.byref(n))
    if dev < 0:
        raise GaspError("no GPU visible to the engine")
    return n.value


def run_driver(L: int, grids: np.ndarray, *, beta: float, h: float, rng: RngState, method: int,
               tot_nsweeps: int, mag_output_int: int, grid_output_int: int, itask: int = 0,
               dn_thr: float = -0.9, up_thr: float = 0.9, ninputs: int = 1,
               cv: str = "magnetisation", initial_spin: int = -1, threads_per_block: int = 64,
               fates: np.ndarray | None = None, outfunc=None) -> dict:
    """mc_driver_gpu through the C ABI (gpuInitGrid, preComputeProbs_gpu, mc_driver_gpu)."""
    Lb = lib()
    R = grids.shape[0]
    host = np.ascontiguousarray(grids, dtype=np.int32).reshape(R, L * L).copy()
    d_grids, d_nb = C.c_void_p(), C.c_void_p()
    Lb.gpuInitGrid(L, R, threads_per_block, _ip(host), C.byref(d_grids), C.byref(d_nb))
    check()
    Lb.preComputeProbs_gpu(beta, h)
    nres = tot_nsweeps // mag_output_int if itask == 0 else ninputs
    result = np.zeros(max(nres, 1), dtype=np.float32)
    if fates is None:
        fates = np.full(R, -1, dtype=np.int32)
    fates = np.ascontiguousarray(fates, dtype=np.int32)
    g = McGpuGrids(L, R, _ip(host), d_grids, d_nb)
    s = McSampler(tot_nsweeps, mag_output_int, grid_output_int)
    f = McFunction(itask, dn_thr, up_thr, ninputs, initial_spin, cv.encode(), _fp(result),
                   b"gridstates.bin")
    run = GpuRun(rng.ptr, threads_per_block, method)
    if outfunc == "write_ising_grids":  # the library's own gridstates.bin writer
        cb = GridOutputFunc(C.cast(Lb.write_ising_grids, C.c_void_p).value)
    else:
        cb = GridOutputFunc(outfunc) if outfunc is not None else GridOutputFunc(0)
    try:
        Lb.mc_driver_gpu(g, beta, h, _ip(fates), s, f, run, cb)
        check()
        # final device grids (the driver leaves the last state in d_ising_grids)
        final = np.empty((R, L * L), dtype=np.int32)
        _d2h(final, d_grids)
    finally:
        Lb.gasp_free(d_grids)
        Lb.gasp_free(d_nb)
    return {"result": result[:nres], "fates": fates, "host_grids": host.reshape(R, L, L),
            "final_grids": final.reshape(R, L, L), "stats": run_stats()}


_hip = None


def _d2h(dst: np.ndarray, dptr: C.c_void_p) -> None:
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemcpy.restype = C.c_int
    rc = _hip.hipMemcpy(dst.ctypes.data, dptr, dst.nbytes, 2)  # hipMemcpyDeviceToHost
    if rc:
        raise GaspError(f"hipMemcpy failed ({rc})")

"""ctypes binding of libhom.so (the C ABI in include/hom_abi.h).

This is the Python counterpart of the Julia ``ccall`` shim: host arrays go in
as plain pointers with Julia's layouts (column-major, interleaved complex), and
status codes come back as exceptions -- ``ValueError`` for HOM_ERR_ARGUMENT
(Julia ArgumentError), ``DimensionMismatch`` for HOM_ERR_DIMENSION and
``HomDeviceError`` for the rest.

The library must have been built (``build.py``); loading fails loudly when it
is missing, and nothing here falls back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from .types import GAMMA, BlochHighOrder, DimensionMismatch, HO_DiscreteSequence, HO_Phantom

LIB_PATH = Path(os.environ.get("HOM_LIBRARY") or Path(__file__).resolve().parent / "libhom.so")

HOM_OK, HOM_ERR_ARGUMENT, HOM_ERR_DIMENSION, HOM_ERR_DEVICE, HOM_ERR_MEMORY, HOM_ERR_STATE = range(6)
# Size of the per-device scratch array of hom_reduce_signal_noalloc (hom_abi.h).
HOM_REDUCE_SCRATCH_BYTES = 32 * 1024 * 1024
PRECISIONS = {"f32": 0, "f64": 1, "mixed": 2}
COIL_MODES = {"gemm": 0, "reduction": 1, "fused_warp": 2, "fused_tiled": 3}
KERNELS = {"plain_precession": 0, "plain_excitation": 1, "coil_partials": 2, "coil_finalize": 3,
           "mixed_segment": 4, "mixed_finalize": 5, "mixed_prep": 6, "fused_precession": 7}

# Every symbol declared in include/hom_abi.h.
ABI_SYMBOLS = (
    "hom_abi_version", "hom_device_count", "hom_last_error", "hom_plan_create", "hom_plan_destroy",
    "hom_upload_phantom", "hom_upload_csm", "hom_upload_sequence", "hom_reset_state",
    "hom_set_state", "hom_get_state", "hom_precession_block", "hom_excitation_block",
    "hom_reduce_signal", "hom_run", "hom_synchronize", "hom_signal_rows", "hom_download_signal",
    "hom_signal_element_size", "hom_buffer_alloc", "hom_buffer_free", "hom_buffer_upload",
    "hom_buffer_download", "hom_plan_set_timing", "hom_plan_kernel_time", "hom_kspace",
    "hom_kspace_rf_events", "hom_discretize", "hom_discrete_copy", "hom_discrete_free",
    "hom_plan_mixed_stats", "hom_reduce_signal_noalloc", "hom_check_workspace_options",
    "hom_plan_set_spin_tile", "hom_plan_set_spin_tile_auto", "hom_plan_get_spin_tile",
    "hom_plan_last_spin_tile",
)

# Every symbol declared in include/hom_serve.h.
SERVE_SYMBOLS = (
    "hom_server_create", "hom_server_destroy", "hom_server_last_error", "hom_server_get_config",
    "hom_server_put_phantom", "hom_server_put_csm", "hom_server_put_sequence",
    "hom_server_release", "hom_server_submit", "hom_server_footprint", "hom_server_step",
    "hom_server_drain", "hom_server_wait", "hom_server_status", "hom_server_result_shape",
    "hom_server_result", "hom_server_times", "hom_server_get_stats", "hom_server_autotune",
)
REQUEST_STATES = {0: "queued", 1: "active", 2: "done", 3: "failed"}
RESIDENT_KINDS = {"phantom": 0, "csm": 1, "sequence": 2}


class HomDeviceError(RuntimeError):
    """Device, memory or state failure reported by the library."""


_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_i64p = C.POINTER(C.c_int64)
_vp = C.c_void_p


class SequenceHost(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("t", _dp), ("dt", _dp), ("gx", _dp), ("gy", _dp),
                ("gz", _dp), ("b1", _dp), ("df", _dp), ("psi", _dp), ("adc", _u8p),
                ("excitation", _u8p), ("h", _dp * 16)]


class Wave(C.Structure):
    _fields_ = [("n", C.c_int64), ("t", _dp), ("a", _dp)]


class BlockDesc(C.Structure):
    _fields_ = [("duration", C.c_double), ("gx", Wave), ("gy", Wave), ("gz", Wave),
                ("rf_re", Wave), ("rf_im", Wave), ("rf_t0", C.c_double), ("rf_t1", C.c_double),
                ("rf_df", C.c_double), ("rf_phase", C.c_double), ("adc_n", C.c_int64),
                ("adc_t", _dp), ("h", Wave * 16)]


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("device_bytes", C.c_int64), ("host_bytes", C.c_int64),
                ("max_active", C.c_int), ("step_budget", C.c_int64), ("use_graphs", C.c_int),
                ("gpt", C.c_int), ("main_grid", C.c_int), ("threaded", C.c_int)]


class RequestDesc(C.Structure):
    _fields_ = [("phantom", C.c_int64), ("spin_first", C.c_int64), ("spin_count", C.c_int64),
                ("csm", C.c_int64), ("coil_count", C.c_int64), ("sequence", C.c_int64),
                ("node_first", C.c_int64), ("node_last", C.c_int64), ("flags", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("queued", C.c_int64), ("active", C.c_int64), ("steps", C.c_int64),
                ("graph_launches", C.c_int64), ("graphs_captured", C.c_int64),
                ("max_active_seen", C.c_int64), ("device_bytes_used", C.c_int64),
                ("device_bytes_peak", C.c_int64), ("busy_seconds", C.c_double),
                ("spin_samples", C.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_lib = None


def load_library():
    """Loads libhom.so and declares every ABI signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise HomDeviceError(f"{LIB_PATH} is missing: build it with build.py "
                             "(the HIP extension is required; there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    P = _vp
    sig = {
        "hom_abi_version": (C.c_int, []),
        "hom_device_count": (C.c_int, [C.POINTER(C.c_int)]),
        "hom_last_error": (C.c_char_p, [P]),
        "hom_plan_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, C.c_int64, C.c_int64,
                                      C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_int]),
        "hom_plan_destroy": (C.c_int, [P]),
        "hom_check_workspace_options": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int64,
                                                  C.c_int64]),
        "hom_upload_phantom": (C.c_int, [P, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "hom_upload_csm": (C.c_int, [P, _dp, C.c_int64]),
        "hom_upload_sequence": (C.c_int, [P, C.POINTER(SequenceHost)]),
        "hom_reset_state": (C.c_int, [P]),
        "hom_set_state": (C.c_int, [P, _dp, _dp]),
        "hom_get_state": (C.c_int, [P, _dp, _dp]),
        "hom_precession_block": (C.c_int, [P, C.c_int64, C.c_int64, C.c_int, _vp, C.c_int64]),
        "hom_excitation_block": (C.c_int, [P, C.c_int64, C.c_int64, C.c_int, _vp, C.c_int64]),
        "hom_reduce_signal": (C.c_int, [P, _vp, C.c_int64, C.c_int64, _vp, C.c_int64]),
        "hom_reduce_signal_noalloc": (C.c_int, [P, _vp, C.c_int64, C.c_int64, _vp, C.c_int64]),
        "hom_run": (C.c_int, [P, _i64p, _i64p, _u8p, C.c_int64, C.c_int]),
        "hom_synchronize": (C.c_int, [P]),
        "hom_signal_rows": (C.c_int, [P, _i64p]),
        "hom_download_signal": (C.c_int, [P, _vp, C.c_int64]),
        "hom_signal_element_size": (C.c_int, [P, _i64p]),
        "hom_buffer_alloc": (C.c_int, [P, C.c_int64, C.POINTER(_vp)]),
        "hom_buffer_free": (C.c_int, [P, _vp]),
        "hom_buffer_upload": (C.c_int, [P, _vp, _vp, C.c_int64]),
        "hom_buffer_download": (C.c_int, [P, _vp, _vp, C.c_int64]),
        "hom_plan_set_timing": (C.c_int, [P, C.c_int]),
        "hom_plan_kernel_time": (C.c_int, [P, C.c_int, _dp, _i64p]),
        "hom_plan_mixed_stats": (C.c_int, [P, _i64p, _i64p]),
        "hom_plan_set_spin_tile": (C.c_int, [P, C.c_int]),
        "hom_plan_set_spin_tile_auto": (C.c_int, [P, C.c_int]),
        "hom_plan_get_spin_tile": (C.c_int, [P, C.POINTER(C.c_int)]),
        "hom_plan_last_spin_tile": (C.c_int, [P, C.POINTER(C.c_int)]),
        "hom_kspace": (C.c_int, [C.POINTER(SequenceHost), C.c_double, _u8p, C.c_int64, _dp, _dp,
                                 _dp, _dp]),
        "hom_kspace_rf_events": (C.c_int, [C.POINTER(SequenceHost), C.c_double, _i64p, _dp,
                                           C.c_int64]),
        "hom_discretize": (C.c_int, [C.POINTER(BlockDesc), C.c_int64, C.c_double, C.c_double,
                                     C.c_int, C.POINTER(_vp), _i64p]),
        "hom_discrete_copy": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _u8p, _u8p, _dp]),
        "hom_discrete_free": (C.c_int, [_vp]),
        "hom_server_create": (C.c_int, [C.POINTER(_vp), C.POINTER(ServerConfig)]),
        "hom_server_destroy": (C.c_int, [P]),
        "hom_server_last_error": (C.c_char_p, [P]),
        "hom_server_get_config": (C.c_int, [P, C.POINTER(ServerConfig)]),
        "hom_server_put_phantom": (C.c_int, [P, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _i64p]),
        "hom_server_put_csm": (C.c_int, [P, C.c_int64, C.c_int64, _dp, C.c_int64, _i64p]),
        "hom_server_put_sequence": (C.c_int, [P, C.POINTER(SequenceHost), C.c_double, _i64p]),
        "hom_server_release": (C.c_int, [P, C.c_int, C.c_int64]),
        "hom_server_submit": (C.c_int, [P, C.POINTER(RequestDesc), _i64p]),
        "hom_server_footprint": (C.c_int, [P, C.POINTER(RequestDesc), _i64p]),
        "hom_server_step": (C.c_int, [P, C.POINTER(C.c_int)]),
        "hom_server_drain": (C.c_int, [P]),
        "hom_server_wait": (C.c_int, [P, C.c_int64, C.c_double, C.POINTER(C.c_int)]),
        "hom_server_status": (C.c_int, [P, C.c_int64, C.POINTER(C.c_int)]),
        "hom_server_result_shape": (C.c_int, [P, C.c_int64, _i64p, _i64p]),
        "hom_server_result": (C.c_int, [P, C.c_int64, _dp, C.c_int64]),
        "hom_server_times": (C.c_int, [P, C.c_int64, _dp, _dp, _dp]),
        "hom_server_get_stats": (C.c_int, [P, C.POINTER(ServerStats)]),
        "hom_server_autotune": (C.c_int, [P, C.POINTER(RequestDesc), C.POINTER(C.c_int), C.c_int,
                                          C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int),
                                          C.POINTER(C.c_int), _dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def device_count() -> int:
    lib = load_library()
    n = C.c_int(0)
    lib.hom_device_count(C.byref(n))
    return n.value


def check_workspace_options(precision: str, coil_combination: str, groupsize: int = 256,
                            adc_tile: int = 8, coil_tile: int = 1, device: int = -1):
    """The HighOrderGPUWorkspace keyword checks (HighOrderGPU.jl:24-35) through
    hom_check_workspace_options; raises ValueError (Julia ArgumentError)."""
    if precision not in PRECISIONS:
        raise ValueError("The precision must be f32, f64 or mixed.")
    if coil_combination not in COIL_MODES:
        raise ValueError(f"Unknown coil_combination: {coil_combination}.")
    for v in (groupsize, adc_tile, coil_tile):
        if not isinstance(v, (int, np.integer)) or isinstance(v, bool):
            raise ValueError("groupsize, adc_tile and coil_tile must be integers.")
    lib = load_library()
    rc = lib.hom_check_workspace_options(int(device), PRECISIONS[precision],
                                         COIL_MODES[coil_combination], int(groupsize),
                                         int(adc_tile), int(coil_tile))
    if rc != HOM_OK:
        _raise(rc, lib.hom_last_error(None).decode())


def _raise(code: int, msg: str):
    if code == HOM_ERR_ARGUMENT:
        raise ValueError(msg)
    if code == HOM_ERR_DIMENSION:
        raise DimensionMismatch(msg)
    raise HomDeviceError(msg)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _sequence_host(seq: HO_DiscreteSequence):
    """SequenceHost view of a sequence plus the arrays it points into."""
    keep = [_f64(seq.t), _f64(seq.dt if seq.dt.size else np.zeros(1)), _f64(seq.gx),
            _f64(seq.gy), _f64(seq.gz), np.ascontiguousarray(seq.b1, np.complex128),
            _f64(seq.df), _f64(seq.psi)]
    adc = np.ascontiguousarray(seq.adc, np.uint8)
    exc = np.ascontiguousarray(seq.excitation if seq.excitation.size else np.zeros(1), np.uint8)
    h = np.ascontiguousarray(seq.h, np.float64)
    hp = (_dp * 16)(*[C.cast(h.ctypes.data + j * h.strides[0], _dp) for j in range(16)])
    s = SequenceHost(seq.n_nodes, *[a.ctypes.data_as(_dp) for a in keep],
                     adc.ctypes.data_as(_u8p), exc.ctypes.data_as(_u8p), hp)
    return s, (keep, adc, exc, h, hp)


class DeviceBuffer:
    """Plan-owned device memory (for hosts without a GPU array library)."""

    def __init__(self, plan: "Plan", nbytes: int):
        self.plan = plan
        self.nbytes = int(nbytes)
        p = _vp()
        plan._check(plan.lib.hom_buffer_alloc(plan.handle, self.nbytes, C.byref(p)))
        self.ptr = p.value

    def upload(self, host: np.ndarray):
        host = np.ascontiguousarray(host) if host.flags.c_contiguous or not host.flags.f_contiguous \
            else host
        if host.nbytes > self.nbytes:
            raise DimensionMismatch("host array larger than the device buffer")
        self.plan._check(self.plan.lib.hom_buffer_upload(self.plan.handle, self.ptr,
                                                         host.ctypes.data, host.nbytes))

    def download(self, host: np.ndarray):
        if host.nbytes > self.nbytes:
            raise DimensionMismatch("host array larger than the device buffer")
        self.plan._check(self.plan.lib.hom_buffer_download(self.plan.handle, host.ctypes.data,
                                                           self.ptr, host.nbytes))
        return host

    def free(self):
        if self.ptr:
            self.plan._check(self.plan.lib.hom_buffer_free(self.plan.handle, self.ptr))
            self.ptr = None


class Plan:
    """One device-resident simulation plan: phantom, CSM, sequence, spin
    state, ADC-state workspace and signal buffer on one GPU and one stream."""

    def __init__(self, precision: str, n_spins: int, n_coils: int, n_nodes: int, max_adc: int,
                 max_rf_adc: int | None = None, gamma: float = GAMMA,
                 coil_combination: str = "gemm", device: int = 0, groupsize: int = 256,
                 adc_tile: int = 8, coil_tile: int = 1):
        """HighOrderGPUWorkspace's arguments (HighOrderGPU.jl:16-35), checked as
        there before the device is touched; groupsize, adc_tile and coil_tile
        are validated for the fused modes and select nothing else here."""
        if precision not in PRECISIONS:
            raise ValueError("The precision must be f32, f64 or mixed.")
        if coil_combination not in COIL_MODES:
            raise ValueError(f"Unknown coil_combination: {coil_combination}.")
        for name, v in (("max_adc", max_adc), ("max_rf_adc", max_rf_adc)):
            if v is not None and (isinstance(v, bool) or not isinstance(v, (int, np.integer))):
                raise ValueError("The maximum ADC count must be a nonnegative integer."
                                 if name == "max_adc" else
                                 "The RF ADC count must be an integer between zero and max_adc.")
        check_workspace_options(precision, coil_combination, groupsize, adc_tile, coil_tile)
        self.lib = load_library()
        self.precision = precision
        self.n_spins, self.n_coils, self.n_nodes = int(n_spins), int(n_coils), int(n_nodes)
        self.max_adc = int(max_adc)
        self.max_rf_adc = self.max_adc if max_rf_adc is None else int(max_rf_adc)
        self.gamma = float(gamma)
        h = _vp()
        rc = self.lib.hom_plan_create(C.byref(h), int(device), PRECISIONS[precision], self.n_spins,
                                      self.n_coils, self.n_nodes, self.max_adc, self.max_rf_adc,
                                      self.gamma, COIL_MODES[coil_combination])
        if rc != HOM_OK:
            _raise(rc, self.lib.hom_last_error(None).decode())
        self.handle = h.value
        self.cdtype = np.complex64 if precision == "f32" else np.complex128
        self._keep = None
        self._tile_explicit = False
        self._tile_applied = None

    # -- lifecycle ---------------------------------------------------------
    def close(self):
        if getattr(self, "handle", None):
            self.lib.hom_plan_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != HOM_OK:
            _raise(rc, self.lib.hom_last_error(self.handle).decode())

    # -- residency ---------------------------------------------------------
    def upload_phantom(self, obj: HO_Phantom):
        obj.validate()
        if len(obj) != self.n_spins or obj.n_coils != self.n_coils:
            raise DimensionMismatch("The phantom does not match the plan's spin and coil counts.")
        arrs = [_f64(a) for a in (obj.x, obj.y, obj.z, obj.rho, obj.T1, obj.T2, obj.dw)]
        self._check(self.lib.hom_upload_phantom(self.handle, *[a.ctypes.data_as(_dp) for a in arrs]))
        csm = np.asfortranarray(obj.csm, dtype=np.complex128)
        self._check(self.lib.hom_upload_csm(self.handle, csm.ctypes.data_as(_dp),
                                            max(csm.shape[0], 1)))

    def upload_sequence(self, seq: HO_DiscreteSequence):
        seq.validate()
        if seq.n_nodes != self.n_nodes:
            raise DimensionMismatch("The sequence node count does not match the plan.")
        s, _keep = _sequence_host(seq)
        self._check(self.lib.hom_upload_sequence(self.handle, C.byref(s)))

    def reset_state(self):
        self._check(self.lib.hom_reset_state(self.handle))

    def set_state(self, mxy, mz):
        mxy = np.ascontiguousarray(mxy, np.complex128)
        mz = _f64(mz)
        if mxy.size != self.n_spins or mz.size != self.n_spins:
            raise DimensionMismatch("State arrays must contain one value per spin.")
        self._check(self.lib.hom_set_state(self.handle, mxy.ctypes.data_as(_dp), mz.ctypes.data_as(_dp)))

    def get_state(self):
        mxy = np.zeros(self.n_spins, np.complex128)
        mz = np.zeros(self.n_spins, np.float64)
        self._check(self.lib.hom_get_state(self.handle, mxy.ctypes.data_as(_dp), mz.ctypes.data_as(_dp)))
        return mxy, mz

    # -- compute -------------------------------------------------------------
    @staticmethod
    def _flags(method) -> int:
        return method.flags if isinstance(method, BlochHighOrder) else int(method)

    def precession_block(self, first, last, method, out: DeviceBuffer | None, ld: int,
                         offset_elems: int = 0):
        self._apply_tuning(method)
        ptr = None if out is None else out.ptr + offset_elems * np.dtype(self.cdtype).itemsize
        self._check(self.lib.hom_precession_block(self.handle, int(first), int(last),
                                                  self._flags(method), ptr, int(ld)))

    def excitation_block(self, first, last, method, out: DeviceBuffer | None, ld: int,
                         offset_elems: int = 0):
        ptr = None if out is None else out.ptr + offset_elems * np.dtype(self.cdtype).itemsize
        self._check(self.lib.hom_excitation_block(self.handle, int(first), int(last),
                                                  self._flags(method), ptr, int(ld)))

    def reduce_signal(self, states: DeviceBuffer, ld_states: int, n_adc: int,
                      out: DeviceBuffer, ld: int, offset_elems: int = 0):
        ptr = out.ptr + offset_elems * np.dtype(self.cdtype).itemsize
        self._check(self.lib.hom_reduce_signal(self.handle, states.ptr if states else None,
                                               int(ld_states), int(n_adc), ptr, int(ld)))

    def reduce_signal_noalloc(self, states: DeviceBuffer, ld_states: int, n_adc: int,
                              out: DeviceBuffer, ld: int, offset_elems: int = 0):
        """hom_reduce_signal_noalloc: the same reduction with its partials in the
        library's per-device scratch array (no device allocation)."""
        ptr = out.ptr + offset_elems * np.dtype(self.cdtype).itemsize
        self._check(self.lib.hom_reduce_signal_noalloc(self.handle, states.ptr if states else None,
                                                       int(ld_states), int(n_adc), ptr, int(ld)))

    def _apply_tuning(self, method):
        """Mixed plans take the persisted spin tile of their shape class
        (tuning.json), or the library default when the file has none, as the
        upper bound of the library's per-segment choice, unless a tile was
        fixed explicitly."""
        if self.precision != "mixed" or self._tile_explicit:
            return
        from . import tuning
        tile = tuning.plan_spin_tile(self.n_coils, self._flags(method)) or tuning.DEFAULT_SPIN_TILE
        if tile != self._tile_applied:
            self._check(self.lib.hom_plan_set_spin_tile_auto(self.handle, int(tile)))
            self._tile_applied = tile

    def run(self, parts, method):
        self._apply_tuning(method)
        pf = np.ascontiguousarray(parts[0], np.int64)
        pl = np.ascontiguousarray(parts[1], np.int64)
        rf = np.ascontiguousarray(parts[2], np.uint8)
        self._keep = (pf, pl, rf)
        self._check(self.lib.hom_run(self.handle, pf.ctypes.data_as(_i64p), pl.ctypes.data_as(_i64p),
                                     rf.ctypes.data_as(_u8p), pf.size, self._flags(method)))

    def synchronize(self):
        self._check(self.lib.hom_synchronize(self.handle))

    def signal_rows(self) -> int:
        r = C.c_int64(0)
        self._check(self.lib.hom_signal_rows(self.handle, C.byref(r)))
        return r.value

    def download_signal(self) -> np.ndarray:
        rows = self.signal_rows()
        out = np.zeros((rows, self.n_coils), dtype=self.cdtype, order="F")
        if rows:
            self._check(self.lib.hom_download_signal(self.handle, out.ctypes.data, rows))
        return out

    def buffer(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    # -- timing --------------------------------------------------------------
    def set_timing(self, enable: bool):
        self._check(self.lib.hom_plan_set_timing(self.handle, 1 if enable else 0))

    def kernel_time(self, name: str):
        ms = C.c_double(0)
        n = C.c_int64(0)
        self._check(self.lib.hom_plan_kernel_time(self.handle, KERNELS[name], C.byref(ms), C.byref(n)))
        return ms.value, n.value

    @property
    def spin_tile(self) -> int:
        """16-spin groups per spin tile of the mixed kernels."""
        v = C.c_int(0)
        self._check(self.lib.hom_plan_get_spin_tile(self.handle, C.byref(v)))
        return v.value

    @spin_tile.setter
    def spin_tile(self, groups: int):
        """Fixes the tile for every segment (no per-segment choice)."""
        self._check(self.lib.hom_plan_set_spin_tile(self.handle, int(groups)))
        self._tile_explicit = True

    @property
    def last_spin_tile(self) -> int:
        """Spin tile the longest segment of the last mixed run used."""
        v = C.c_int(0)
        self._check(self.lib.hom_plan_last_spin_tile(self.handle, C.byref(v)))
        return v.value

    def mixed_stats(self):
        """(segments, fp64_segments) of the last mixed run: precession segments
        that acquired samples, and those on the all-FP64 phase body."""
        seg = C.c_int64(0)
        f64 = C.c_int64(0)
        self._check(self.lib.hom_plan_mixed_stats(self.handle, C.byref(seg), C.byref(f64)))
        return seg.value, f64.value


class Server:
    """ctypes layer over include/hom_serve.h (one method per entry point)."""

    def __init__(self, device: int = 0, device_bytes: int = 0, host_bytes: int = 0,
                 max_active: int = 0, step_budget: int = 0, use_graphs: bool = True,
                 gpt: int = 0, main_grid: int = 0, threaded: bool = False):
        self.lib = load_library()
        cfg = ServerConfig(int(device), int(device_bytes), int(host_bytes), int(max_active),
                           int(step_budget), 1 if use_graphs else 0, int(gpt), int(main_grid),
                           1 if threaded else 0)
        h = _vp()
        rc = self.lib.hom_server_create(C.byref(h), C.byref(cfg))
        if rc != HOM_OK:
            _raise(rc, self.lib.hom_server_last_error(None).decode())
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            self.lib.hom_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != HOM_OK:
            _raise(rc, self.lib.hom_server_last_error(self.handle).decode())

    def config(self) -> dict:
        c = ServerConfig()
        self._check(self.lib.hom_server_get_config(self.handle, C.byref(c)))
        return {name: getattr(c, name) for name, _ in c._fields_}

    def put_phantom(self, x, y, z, rho, T1, T2, dw) -> int:
        arrs = [_f64(a) for a in (x, y, z, rho, T1, T2, dw)]
        out = C.c_int64(0)
        self._check(self.lib.hom_server_put_phantom(self.handle, len(arrs[0]),
                                                    *[a.ctypes.data_as(_dp) for a in arrs],
                                                    C.byref(out)))
        return out.value

    def put_csm(self, phantom: int, csm) -> int:
        csm = np.asfortranarray(csm, dtype=np.complex128)
        out = C.c_int64(0)
        self._check(self.lib.hom_server_put_csm(self.handle, int(phantom), csm.shape[1],
                                                csm.ctypes.data_as(_dp), max(csm.shape[0], 1),
                                                C.byref(out)))
        return out.value

    def put_sequence(self, seq: HO_DiscreteSequence, gamma: float = GAMMA) -> int:
        seq.validate()
        s, _keep = _sequence_host(seq)
        out = C.c_int64(0)
        self._check(self.lib.hom_server_put_sequence(self.handle, C.byref(s), float(gamma),
                                                     C.byref(out)))
        return out.value

    def release(self, kind: str, ident: int):
        self._check(self.lib.hom_server_release(self.handle, RESIDENT_KINDS[kind], int(ident)))

    @staticmethod
    def request(phantom, spin_first, spin_count, csm, coil_count, sequence, node_first,
                node_last, flags) -> RequestDesc:
        return RequestDesc(int(phantom), int(spin_first), int(spin_count), int(csm),
                           int(coil_count), int(sequence), int(node_first), int(node_last),
                           int(flags))

    def submit(self, req: RequestDesc) -> int:
        out = C.c_int64(0)
        self._check(self.lib.hom_server_submit(self.handle, C.byref(req), C.byref(out)))
        return out.value

    def footprint(self, req: RequestDesc) -> int:
        out = C.c_int64(0)
        self._check(self.lib.hom_server_footprint(self.handle, C.byref(req), C.byref(out)))
        return out.value

    def step(self) -> bool:
        w = C.c_int(0)
        self._check(self.lib.hom_server_step(self.handle, C.byref(w)))
        return bool(w.value)

    def drain(self):
        self._check(self.lib.hom_server_drain(self.handle))

    def wait(self, rid: int, timeout: float = -1.0) -> str:
        st = C.c_int(0)
        self._check(self.lib.hom_server_wait(self.handle, int(rid), float(timeout), C.byref(st)))
        return REQUEST_STATES[st.value]

    def status(self, rid: int) -> str:
        st = C.c_int(0)
        self._check(self.lib.hom_server_status(self.handle, int(rid), C.byref(st)))
        return REQUEST_STATES[st.value]

    def times(self, rid: int):
        a, b, c = C.c_double(0), C.c_double(0), C.c_double(0)
        self._check(self.lib.hom_server_times(self.handle, int(rid), C.byref(a), C.byref(b),
                                              C.byref(c)))
        return a.value, b.value, c.value

    def result(self, rid: int) -> np.ndarray:
        rows, cols = C.c_int64(0), C.c_int64(0)
        self._check(self.lib.hom_server_result_shape(self.handle, int(rid), C.byref(rows),
                                                     C.byref(cols)))
        out = np.zeros((rows.value, cols.value), dtype=np.complex128, order="F")
        self._check(self.lib.hom_server_result(self.handle, int(rid), out.ctypes.data_as(_dp),
                                               max(rows.value, 1)))
        return out

    def stats(self) -> dict:
        st = ServerStats()
        self._check(self.lib.hom_server_get_stats(self.handle, C.byref(st)))
        return st.as_dict()

    def autotune(self, probe: RequestDesc, gpts=None, grids=None):
        gl = (C.c_int * len(gpts))(*gpts) if gpts else None
        rl = (C.c_int * len(grids))(*grids) if grids else None
        bg, bgrid, bs = C.c_int(0), C.c_int(0), C.c_double(0)
        self._check(self.lib.hom_server_autotune(self.handle, C.byref(probe), gl,
                                                 len(gpts) if gpts else 0, rl,
                                                 len(grids) if grids else 0, C.byref(bg),
                                                 C.byref(bgrid), C.byref(bs)))
        return bg.value, bgrid.value, bs.value


def kspace(seq: HO_DiscreteSequence, gamma: float = GAMMA, skip_rf=None):
    """hom_kspace: (K_nominal, K_nominal_adc, K_dfc, K_dfc_adc) of a discretized
    sequence (host computation in the library; no device needed)."""
    lib = load_library()
    seq.validate()
    if seq.n_nodes < 2:
        raise ValueError("Cannot integrate an empty sequence.")
    s, _keep = _sequence_host(seq)
    n_ev = C.c_int64(0)
    rc = lib.hom_kspace_rf_events(C.byref(s), float(gamma), C.byref(n_ev), None, 0)
    if rc != HOM_OK:
        _raise(rc, "invalid sequence")
    skip = None
    if skip_rf is not None:
        skip = np.ascontiguousarray(np.asarray(skip_rf, dtype=bool), dtype=np.uint8)
        if skip.size != n_ev.value:
            raise DimensionMismatch("skip_rf must contain one flag per RF event.")
    ni, nadc = seq.n_nodes - 1, int(np.count_nonzero(seq.adc))
    kn = np.zeros((ni, 3), order="F")
    kna = np.zeros((nadc, 3), order="F")
    kd = np.zeros((ni, 16), order="F")
    kda = np.zeros((nadc, 16), order="F")
    rc = lib.hom_kspace(C.byref(s), float(gamma),
                        skip.ctypes.data_as(_u8p) if skip is not None else None,
                        skip.size if skip is not None else 0, kn.ctypes.data_as(_dp),
                        kna.ctypes.data_as(_dp), kd.ctypes.data_as(_dp), kda.ctypes.data_as(_dp))
    if rc != HOM_OK:
        _raise(rc, "hom_kspace failed")
    return kn, kna, kd, kda


def rf_flip_angles(seq: HO_DiscreteSequence, gamma: float = GAMMA) -> np.ndarray:
    """Flip angle (degrees) of every RF event of a discretized sequence."""
    lib = load_library()
    s, _keep = _sequence_host(seq)
    n_ev = C.c_int64(0)
    _ok = lib.hom_kspace_rf_events(C.byref(s), float(gamma), C.byref(n_ev), None, 0)
    out = np.zeros(max(n_ev.value, 1))
    lib.hom_kspace_rf_events(C.byref(s), float(gamma), C.byref(n_ev), out.ctypes.data_as(_dp),
                             out.size)
    return out[:n_ev.value]


def discretize_blocks(descs, keep, dt_max: float, dt_rf: float, nominal_in_h: bool):
    """hom_discretize over a ctypes BlockDesc array; returns the node arrays."""
    lib = load_library()
    arr = (BlockDesc * max(len(descs), 1))(*descs)
    n = C.c_int64(0)
    handle = _vp()
    rc = lib.hom_discretize(arr, len(descs), float(dt_max), float(dt_rf), int(nominal_in_h),
                            C.byref(handle), C.byref(n))
    if rc != HOM_OK:
        _raise(rc, "invalid sequence block description")
    try:
        nn = n.value
        t, gx, gy, gz, df, psi = (np.zeros(nn) for _ in range(6))
        b1 = np.zeros(nn, np.complex128)
        adc = np.zeros(nn, np.uint8)
        exc = np.zeros(max(nn - 1, 1), np.uint8)
        h = np.zeros((16, nn))
        rc = lib.hom_discrete_copy(handle, t.ctypes.data_as(_dp), gx.ctypes.data_as(_dp),
                                   gy.ctypes.data_as(_dp), gz.ctypes.data_as(_dp),
                                   b1.ctypes.data_as(_dp), df.ctypes.data_as(_dp),
                                   psi.ctypes.data_as(_dp), adc.ctypes.data_as(_u8p),
                                   exc.ctypes.data_as(_u8p), h.ctypes.data_as(_dp))
        if rc != HOM_OK:
            _raise(rc, "hom_discrete_copy failed")
    finally:
        lib.hom_discrete_free(handle)
    return HO_DiscreteSequence(t=t, gx=gx, gy=gy, gz=gz, b1=b1, df=df, psi=psi,
                               adc=adc.astype(bool), excitation=exc[:max(nn - 1, 0)].astype(bool),
                               h=h)

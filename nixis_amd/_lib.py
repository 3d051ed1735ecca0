"""ctypes binding of libnixis_hip.so (C ABI in include/nixis.h).

The shared library is built in-tree by ``nixis_amd.build``.  There is no
fallback: if the library is missing, or no HIP device is present when a
device entry point is used, the call raises.
"""
import contextlib
import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnixis_hip.so")

NX_OK = 0
NX_EINVAL = -1
NX_ENOMEM = -2
NX_EHIP = -3
NX_ERANGE = -4
NX_EDIV0 = -5
NX_EBUSY = -6
NX_MAX_OCTAVES = 32
NX_OPT_GRAPH_CHUNK = 1
NX_OPT_PAIR_KERNEL = 2
NX_OPT_MULTI_KERNEL = 3
NX_OPT_CHAIN_PASSES = 4


class NixisError(RuntimeError):
    """A libnixis_hip call failed; ``code`` is the NX_E* status."""

    def __init__(self, code, msg):
        super().__init__(f"{msg} (status {code})")
        self.code = code


class NixisRangeError(NixisError, ValueError):
    pass


class NixisArgError(NixisError, ValueError):
    pass


class NixisBusyError(NixisError):
    pass


_c_void_p = ctypes.c_void_p


class NxTilePlan(ctypes.Structure):
    """Mirror of nx_tile_plan (include/nixis.h)."""
    _fields_ = [("blob", ctypes.c_void_p), ("bytes", ctypes.c_int64), ("V", ctypes.c_int64),
                ("ntiles", ctypes.c_int64), ("max_tile", ctypes.c_int32),
                ("max_load", ctypes.c_int32)] + [
        (f"off_{n}", ctypes.c_int64)
        for n in ("start", "cs", "load", "rbase", "ring", "cbase", "lcol", "lel")] + [
        ("steps", ctypes.c_int32), ("max_step1", ctypes.c_int32), ("off_pref", ctypes.c_int64),
        ("off_ebase", ctypes.c_int64)]


class NxRescaleOpts(ctypes.Structure):
    """Mirror of nx_rescale_opts (include/nixis.h)."""
    _fields_ = [("mid", ctypes.c_double), ("u_min", ctypes.c_double), ("u_max", ctypes.c_double),
                ("has_mid", ctypes.c_int), ("mode", ctypes.c_int), ("has_umin", ctypes.c_int),
                ("has_umax", ctypes.c_int)]


_c_i64 = ctypes.c_int64
_c_int = ctypes.c_int
_c_double = ctypes.c_double
_c_size = ctypes.c_size_t

# name -> (restype, argtypes); one entry per declaration in include/nixis.h
SIGNATURES = {
    "nx_last_error": (ctypes.c_char_p, []),
    "nx_abi_version": (_c_int, []),
    "nx_ctx_create": (_c_int, [_c_int, ctypes.POINTER(_c_void_p)]),
    "nx_ctx_destroy": (_c_int, [_c_void_p]),
    "nx_ctx_sync_check": (_c_int, [_c_void_p, _c_void_p]),
    "nx_ctx_set_option": (_c_int, [_c_void_p, _c_int, _c_i64]),
    "nx_init_tables": (_c_int, [_c_i64, _c_void_p, _c_void_p]),
    "nx_sample_octaves": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_void_p, _c_void_p, _c_int,
                                   _c_void_p, _c_void_p, _c_double, _c_void_p, _c_int,
                                   _c_void_p, _c_void_p, _c_void_p]),
    "nx_noise3d": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_void_p, _c_i64, _c_void_p,
                            _c_void_p, _c_void_p, _c_void_p]),
    "nx_rescale": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_double, _c_double, _c_void_p,
                            _c_void_p, _c_void_p]),
    "nx_rescale_ex": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_double, _c_double,
                               ctypes.POINTER(NxRescaleOpts), _c_void_p, _c_void_p, _c_void_p]),
    "nx_make_bool_elevation_mask": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_double,
                                             _c_void_p, _c_void_p]),
    "nx_power_rescale_workspace_bytes": (_c_size, [_c_i64]),
    "nx_power_rescale": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_i64, _c_int, _c_double,
                                  _c_void_p, _c_void_p, _c_void_p]),
    "nx_height_chain_workspace_bytes": (_c_size, [_c_i64]),
    "nx_terrain_heights": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_void_p, _c_void_p, _c_int,
                                    _c_void_p, _c_void_p, _c_double, _c_double, _c_double,
                                    _c_double, _c_void_p, _c_void_p, _c_void_p, _c_void_p,
                                    _c_void_p]),
    "nx_terrain_heights_noalloc": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_void_p, _c_void_p,
                                            _c_int, _c_void_p, _c_void_p, _c_double, _c_double,
                                            _c_double, _c_double, _c_void_p, _c_void_p, _c_void_p,
                                            _c_void_p]),
    "nx_height_chain": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_double, _c_double, _c_double,
                                 _c_void_p, _c_void_p, _c_void_p, _c_void_p]),
    "nx_climate_workspace_bytes": (_c_size, []),
    "nx_sample_insolation": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_double, _c_double,
                                      _c_double, _c_int, _c_double, _c_void_p, _c_void_p,
                                      _c_void_p]),
    "nx_insolation_slice": (_c_int, [_c_void_p, _c_double, _c_double, _c_void_p, _c_void_p,
                                     _c_void_p]),
    "nx_interpolate_insolation": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_void_p, _c_double,
                                           _c_void_p, _c_void_p]),
    "nx_assign_surface_temp": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_double, _c_double,
                                        _c_void_p, _c_void_p]),
    "nx_daily_insolation": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_double, _c_double,
                                     _c_void_p, _c_void_p, _c_void_p]),
    "nx_yearly_insolation": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_double, _c_double,
                                      _c_void_p, _c_void_p, _c_void_p]),
    "nx_knn_grid_bytes": (_c_size, [_c_i64]),
    "nx_knn_grid_build": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_void_p, _c_void_p]),
    "nx_knn_query": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_void_p, _c_i64, _c_int,
                              _c_void_p, _c_void_p, _c_void_p]),
    "nx_pixel_workspace_bytes": (_c_size, [_c_i64, _c_i64]),
    "nx_knn_query_pixels": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_i64, _c_i64, _c_double,
                                     _c_int, _c_void_p, _c_void_p, _c_void_p, _c_void_p]),
    "nx_make_ll_arr": (_c_int, [_c_void_p, _c_i64, _c_i64, _c_double, _c_void_p, _c_void_p,
                                _c_void_p]),
    "nx_make_gray_array": (_c_int, [_c_void_p, _c_i64, _c_int, _c_void_p, _c_void_p, _c_void_p,
                                    _c_i64, _c_void_p, _c_void_p]),
    "nx_erosion_state_bytes": (_c_size, [_c_i64]),
    "nx_erosion_load": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_void_p, _c_void_p,
                                 _c_void_p, _c_void_p]),
    "nx_erosion_run": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_void_p, _c_void_p, _c_int,
                                _c_double, _c_void_p]),
    "nx_icosphere_device": (_c_int, [_c_void_p, _c_i64, _c_void_p, _c_void_p, _c_void_p]),
    "nx_adjacency_workspace_bytes": (_c_size, [_c_i64]),
    "nx_build_adjacency_device": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_i64, _c_void_p,
                                           _c_void_p, _c_void_p]),
    "nx_sort_adjacency_device": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_i64, _c_void_p]),
    "nx_sfc_workspace_bytes": (_c_size, [_c_i64]),
    "nx_sfc_order_device": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_void_p, _c_void_p,
                                     _c_void_p]),
    "nx_ell_workspace_bytes": (_c_size, [_c_i64]),
    "nx_ell_from_adjacency_device": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_void_p, _c_void_p,
                                              _c_void_p, _c_void_p, _c_void_p, _c_void_p]),
    "nx_tile_plan_build": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_void_p, _c_i64,
                                    ctypes.POINTER(NxTilePlan)]),
    "nx_tile_plan_build_steps": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_int, _c_void_p, _c_i64,
                                          ctypes.POINTER(NxTilePlan)]),
    "nx_tile_plan_build_chunk": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_int, _c_int, _c_void_p,
                                          _c_i64, ctypes.POINTER(NxTilePlan)]),
    "nx_erosion_run_tiled": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_void_p, _c_void_p,
                                      ctypes.POINTER(NxTilePlan), _c_int, _c_double, _c_void_p]),
    "nx_erosion_run_rule": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_void_p, _c_void_p,
                                     ctypes.POINTER(NxTilePlan), _c_int, _c_double, _c_int,
                                     _c_void_p]),
    "nx_erosion1_run": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_i64, _c_void_p, _c_int,
                                 _c_void_p]),
    "nx_erosion_store": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_void_p, _c_void_p,
                                  _c_void_p, _c_void_p]),
    "nx_erosion_mass": (_c_int, [_c_void_p, _c_void_p, _c_i64, _c_void_p, _c_void_p]),
    "nx_icosphere_counts": (_c_int, [_c_i64, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i64)]),
    "nx_icosphere": (_c_int, [_c_i64, _c_void_p, _c_void_p]),
    "nx_build_adjacency": (_c_int, [_c_void_p, _c_i64, _c_i64, _c_void_p]),
    "nx_sort_adjacency": (_c_int, [_c_void_p, _c_i64]),
    "nx_sfc_order": (_c_int, [_c_void_p, _c_i64, _c_void_p]),
    "nx_ell_from_adjacency": (_c_int, [_c_void_p, _c_i64, _c_void_p, _c_void_p, _c_void_p,
                                       _c_void_p]),
}

_lib = None
_lock = threading.Lock()       # library loading
_ctx_lock = threading.Lock()   # the per-device context table


def lib():
    """Load libnixis_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                # torch first: its HIP runtime (libamdhip64.so.7) then also
                # satisfies this library.  Loaded the other way round, torch
                # pulls in a second HIP/HSA runtime whose streams and
                # ordering are not this library's.
                import torch  # noqa: F401
                if not os.path.exists(LIB_PATH):
                    raise ImportError(
                        f"{LIB_PATH} is missing: build it with `python -m nixis_amd.build` "
                        "(there is no CPU fallback)")
                L = ctypes.CDLL(LIB_PATH)
                for name, (res, args) in SIGNATURES.items():
                    fn = getattr(L, name)
                    fn.restype = res
                    fn.argtypes = args
                _lib = L
    return _lib


def check(rc):
    """Raise the matching exception for a non-zero status."""
    if rc == NX_OK:
        return
    msg = lib().nx_last_error().decode(errors="replace")
    if rc == NX_EDIV0:
        raise ZeroDivisionError(msg)
    if rc == NX_ERANGE:
        raise NixisRangeError(rc, msg)
    if rc == NX_EINVAL:
        raise NixisArgError(rc, msg)
    if rc == NX_ENOMEM:
        raise MemoryError(msg)
    if rc == NX_EBUSY:
        raise NixisBusyError(rc, msg)
    raise NixisError(rc, msg)


class Context:
    """One nx_ctx per HIP device (created on first use)."""

    _ctxs = {}

    def __init__(self, device):
        h = _c_void_p()
        check(lib().nx_ctx_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = device
        # the C context's option values (its defaults until set here)
        self.graph_chunk = 32
        self.pair_kernel = 0
        self.multi_kernel = 0
        self.chain_passes = 2

    @classmethod
    def get(cls, device=None):
        import torch
        if device is None:
            device = torch.cuda.current_device()
        if isinstance(device, torch.device):
            device = device.index if device.index is not None else torch.cuda.current_device()
        lib()  # load first: lib() takes _lock, which is not reentrant
        with _ctx_lock:
            c = cls._ctxs.get(device)
            if c is None:
                c = cls(device)
                cls._ctxs[device] = c
        return c

    def sync_check(self, stream=None):
        check(lib().nx_ctx_sync_check(self.handle, stream_ptr(stream)))

    def set_graph_chunk(self, steps):
        """Erosion steps per captured hipGraph (0: eager launches)."""
        check(lib().nx_ctx_set_option(self.handle, NX_OPT_GRAPH_CHUNK, int(steps)))
        self.graph_chunk = int(steps)

    def set_multi_kernel(self, variant):
        """Four-step erosion kernel variant 0..4 (0: default); same results."""
        check(lib().nx_ctx_set_option(self.handle, NX_OPT_MULTI_KERNEL, int(variant)))
        self.multi_kernel = int(variant)

    def set_pair_kernel(self, variant):
        """Two-step erosion kernel variant 0..6 (0: default); same results."""
        check(lib().nx_ctx_set_option(self.handle, NX_OPT_PAIR_KERNEL, int(variant)))
        self.pair_kernel = int(variant)

    def set_chain_passes(self, passes):
        """Height chain form: 2 (default) or 5 (a pass per step); same results."""
        check(lib().nx_ctx_set_option(self.handle, NX_OPT_CHAIN_PASSES, int(passes)))
        self.chain_passes = int(passes)

    @contextlib.contextmanager
    def options(self, graph_chunk=None, pair_kernel=None, multi_kernel=None):
        """Set the given erosion options for the body of a with-block and
        restore the previous values after it (None leaves one alone)."""
        setters = (("graph_chunk", self.set_graph_chunk, graph_chunk),
                   ("pair_kernel", self.set_pair_kernel, pair_kernel),
                   ("multi_kernel", self.set_multi_kernel, multi_kernel))
        old = {name: getattr(self, name) for name, _, _ in setters}
        try:
            for name, setter, value in setters:
                if value is not None and value != old[name]:
                    setter(value)
            yield self
        finally:
            for name, setter, _ in setters:
                if getattr(self, name) != old[name]:
                    setter(old[name])


def stream_ptr(stream=None):
    """hipStream_t of a torch stream (default: the current stream)."""
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return _c_void_p(stream.cuda_stream)


_side = {}


def side_stream(device=None):
    """A second stream on `device`: non-blocking, default priority, created
    once per device and kept for the life of the process, wrapped as a
    torch.cuda.ExternalStream.  torch's own stream pool is not used because
    creating it also creates its raised-priority streams."""
    import torch
    if device is None:
        device = torch.cuda.current_device()
    if isinstance(device, torch.device):
        device = device.index if device.index is not None else torch.cuda.current_device()
    lib()
    with _ctx_lock:
        s = _side.get(device)
        if s is None:
            hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime already loaded by torch
            fn = hip.hipStreamCreateWithFlags
            fn.restype = _c_int
            fn.argtypes = [ctypes.POINTER(_c_void_p), ctypes.c_uint]
            h = _c_void_p()
            with torch.cuda.device(device):
                rc = fn(ctypes.byref(h), 1)  # hipStreamNonBlocking
            if rc != 0 or not h.value:
                raise NixisError(NX_EHIP, f"hipStreamCreateWithFlags failed ({rc})")
            s = torch.cuda.ExternalStream(h.value, device=torch.device("cuda", device))
            _side[device] = s
    return s


_masked = {}


def masked_stream(n_cus, device=None, stride=1):
    """A stream on `device` whose kernels run on `n_cus` of its CUs (CU mask
    bits 0, stride, 2 stride, ...), non-blocking, default priority, created
    once per (device, n_cus, stride) and kept for the life of the process."""
    import torch
    if device is None:
        device = torch.cuda.current_device()
    if isinstance(device, torch.device):
        device = device.index if device.index is not None else torch.cuda.current_device()
    lib()
    key = (device, int(n_cus), int(stride))
    with _ctx_lock:
        s = _masked.get(key)
        if s is None:
            total = torch.cuda.get_device_properties(device).multi_processor_count
            bits = [i * stride for i in range(int(n_cus)) if i * stride < total]
            words = (total + 31) // 32
            mask = (ctypes.c_uint32 * words)()
            for b in bits:
                mask[b // 32] |= 1 << (b % 32)
            hip = ctypes.CDLL("libamdhip64.so.7")
            fn = hip.hipExtStreamCreateWithCUMask
            fn.restype = _c_int
            fn.argtypes = [ctypes.POINTER(_c_void_p), ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
            h = _c_void_p()
            with torch.cuda.device(device):
                rc = fn(ctypes.byref(h), words, mask)
            if rc != 0 or not h.value:
                raise NixisError(NX_EHIP, f"hipExtStreamCreateWithCUMask failed ({rc})")
            s = torch.cuda.ExternalStream(h.value, device=torch.device("cuda", device))
            _masked[key] = s
    return s


def ptr(t):
    """Device (or host) address of a tensor / ndarray, None for None."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return _c_void_p(t.data_ptr())
    return _c_void_p(t.ctypes.data)

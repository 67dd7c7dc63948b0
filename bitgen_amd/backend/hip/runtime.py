"""ctypes binding of libbitgen_hip.so (declared in include/bitgen_hip.h).

The product path has no fallback: if the library is missing or fails to
load, every entry point raises ``HipRuntimeError``.
"""
import ctypes
import os
import threading

from ...csrc import build as native_build


class HipRuntimeError(RuntimeError):
    pass


class GroupArgs(ctypes.Structure):
    """Mirror of ``bg_group_args_t`` (field order is ABI)."""

    _fields_ = [
        ("basis", ctypes.c_void_p),
        ("result", ctypes.c_void_p),
        ("status", ctypes.c_void_p),
        ("ctl", ctypes.c_void_p),
        ("n_words", ctypes.c_uint64),
        ("n_tiles", ctypes.c_uint32),
        ("n_queues", ctypes.c_uint32),
        ("n_carry", ctypes.c_uint32),
        ("take", ctypes.c_uint32),
        ("filter", ctypes.c_void_p),
    ]


class Job(ctypes.Structure):
    """Mirror of ``bg_job_t``."""

    _fields_ = [
        ("basis", ctypes.c_void_p),
        ("result", ctypes.c_void_p),
        ("status", ctypes.c_void_p),
        ("n_words", ctypes.c_uint64),
        ("group", ctypes.c_uint32),
        ("n_carry", ctypes.c_uint32),
        ("tile_start", ctypes.c_uint32),
        ("n_tiles", ctypes.c_uint32),
    ]


class BatchArgs(ctypes.Structure):
    """Mirror of ``bg_batch_args_t``."""

    _fields_ = [
        ("jobs", ctypes.c_void_p),
        ("tile_job", ctypes.c_void_p),
        ("ctl", ctypes.c_void_p),
        ("n_jobs", ctypes.c_uint32),
        ("total_tiles", ctypes.c_uint32),
        ("n_queues", ctypes.c_uint32),
        ("take", ctypes.c_uint32),
        ("filter", ctypes.c_void_p),
    ]


# Control block layout (BG_CTL_* in include/bitgen_hip.h): u32 word indices.
MAX_QUEUES = 256
CTL_STRIDE = 64
CTL_QDONE = 2 * MAX_QUEUES * CTL_STRIDE
CTL_ERROR = CTL_QDONE + CTL_STRIDE
CTL_EPOCH = CTL_ERROR + CTL_STRIDE
CTL_WORDS = CTL_EPOCH + CTL_STRIDE


class TDesc(ctypes.Structure):
    """Mirror of ``bg_tdesc_t``."""

    _fields_ = [(f, ctypes.c_uint64) for f in ("byte_off", "n_bytes", "basis_off", "n_words", "word_start")]


class KernelInfo(ctypes.Structure):
    _fields_ = [
        ("vgprs", ctypes.c_int32),
        ("sgprs", ctypes.c_int32),
        ("lds_bytes", ctypes.c_int32),
        ("scratch_bytes", ctypes.c_int32),
        ("max_threads", ctypes.c_int32),
    ]


EXPORTS = (
    "bg_init", "bg_last_error", "bg_transpose", "bg_module_load", "bg_module_load_data",
    "bg_module_unload", "bg_kernel_get", "bg_kernel_info", "bg_group_launch",
    "bg_to_reference_layout", "bg_popcount", "bg_stream_sync", "bg_batch_launch",
    "bg_graph_begin", "bg_graph_end", "bg_graph_launch", "bg_graph_destroy", "bg_transpose_batch",
    "bg_kernel_occupancy", "bg_ctl_words", "bg_group_launch_static", "bg_serve_pack", "bg_serve_enqueue",
    "bg_filter_launch", "bg_filter_batch_launch",
)

ENOSPACE = 1003


class ServeKernel(ctypes.Structure):
    _fields_ = [("kernel", ctypes.c_void_p), ("hr", ctypes.c_uint32), ("s", ctypes.c_uint32),
                ("status_slots", ctypes.c_uint32), ("n_out", ctypes.c_uint32), ("block_threads", ctypes.c_uint32),
                ("grid_cap", ctypes.c_uint32)]


class ServeSet(ctypes.Structure):
    _fields_ = [("first_kernel", ctypes.c_uint32), ("n_kernels", ctypes.c_uint32)]


class ServeLayout(ctypes.Structure):
    _fields_ = [("stage_bytes", ctypes.c_uint64), ("basis_words", ctypes.c_uint64),
                ("result_words", ctypes.c_uint64), ("status_words", ctypes.c_uint64),
                ("tdesc_off", ctypes.c_uint64), ("n_jobs", ctypes.c_uint32), ("n_queues", ctypes.c_uint32)]


class ServeQueue(ctypes.Structure):
    _fields_ = [("kernel", ctypes.c_uint32), ("n_jobs", ctypes.c_uint32), ("total_tiles", ctypes.c_uint32),
                ("blocks", ctypes.c_uint32), ("jobs_off", ctypes.c_uint64), ("tile_job_off", ctypes.c_uint64)]


# bg_serve_job_t as a numpy record: the job list is built vectorised
SERVE_JOB_DTYPE = [("data", "<u8"), ("n_bytes", "<u8"), ("set", "<u4"), ("reserved", "<u4")]

_lib = None
_lock = threading.Lock()
_initialised = set()


def library_path():
    return os.environ.get("BITGEN_HIP_LIB", native_build.LIB_PATH)


def load_library():
    """Load the C ABI (no GPU call is made)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        path = library_path()
        if not os.path.exists(path):
            raise HipRuntimeError(
                f"{path} is missing: build the hip backend first (python -c 'import __graft_entry__ as g; g.build()')")
        L = ctypes.CDLL(path)
        vp, u64, u32, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int
        L.bg_init.argtypes = [i32]
        L.bg_last_error.restype = ctypes.c_char_p
        L.bg_transpose.argtypes = [vp, u64, u64, vp, u64, u32, vp]
        L.bg_module_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
        L.bg_module_load_data.argtypes = [vp, ctypes.POINTER(vp)]
        L.bg_module_unload.argtypes = [vp]
        L.bg_kernel_get.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(vp)]
        L.bg_kernel_info.argtypes = [vp, ctypes.POINTER(KernelInfo)]
        L.bg_group_launch.argtypes = [vp, ctypes.POINTER(GroupArgs), u32, vp]
        L.bg_filter_launch.argtypes = [vp, vp, u64, u32, vp, u32, vp]
        L.bg_filter_batch_launch.argtypes = [vp, vp, vp, u32, vp, u32, vp]
        L.bg_kernel_occupancy.argtypes = [vp, u32, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
        L.bg_ctl_words.argtypes = []
        L.bg_to_reference_layout.argtypes = [vp, vp, u64, u64, vp]
        L.bg_popcount.argtypes = [vp, u64, vp, vp]
        L.bg_stream_sync.argtypes = [vp]
        L.bg_batch_launch.argtypes = [vp, ctypes.POINTER(BatchArgs), u32, u32, vp]
        L.bg_graph_begin.argtypes = [vp]
        L.bg_graph_end.argtypes = [vp, ctypes.POINTER(vp)]
        L.bg_graph_launch.argtypes = [vp, vp]
        L.bg_graph_destroy.argtypes = [vp]
        L.bg_transpose_batch.argtypes = [vp, vp, vp, u32, u64, vp]
        L.bg_group_launch_static.argtypes = [vp, vp, vp, u64, u32, u32, vp, vp]
        L.bg_serve_pack.argtypes = [vp, u32, vp, u32, vp, u32, vp, u64, u64, u64, u64, u64, u64, u64,
                                    ctypes.POINTER(ServeLayout), vp, u32, vp, u32]
        L.bg_serve_enqueue.argtypes = [ctypes.POINTER(ServeLayout), vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        for name in EXPORTS:
            if name not in ("bg_last_error", "bg_ctl_words"):
                getattr(L, name).restype = ctypes.c_int
        L.bg_ctl_words.restype = u32
        if L.bg_ctl_words() != CTL_WORDS:
            raise HipRuntimeError(f"{path}: control block of {L.bg_ctl_words()} words, expected {CTL_WORDS}")
        _lib = L
        return L


def _check(rc, what):
    if rc != 0:
        msg = load_library().bg_last_error().decode(errors="replace")
        raise HipRuntimeError(f"{what} failed ({rc}): {msg}")


def init(device=0):
    L = load_library()
    if device not in _initialised:
        _check(L.bg_init(device), "bg_init")
        _initialised.add(device)


def current_stream_ptr(stream=None):
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def transpose(bytes_ptr, n_bytes, in_stride, basis_ptr, n_words, n_inputs, stream):
    _check(load_library().bg_transpose(bytes_ptr, n_bytes, in_stride, basis_ptr, n_words, n_inputs, stream),
           "bg_transpose")


class Module:
    def __init__(self, path):
        self.path = path
        if not os.path.isfile(path):
            raise HipRuntimeError(f"code object {path} is missing (compile the pattern set first)")
        h = ctypes.c_void_p()
        _check(load_library().bg_module_load(path.encode(), ctypes.byref(h)), f"bg_module_load({path})")
        self.handle = h
        self.kernels = {}

    def kernel(self, name):
        if name not in self.kernels:
            k = ctypes.c_void_p()
            _check(load_library().bg_kernel_get(self.handle, name.encode(), ctypes.byref(k)), f"bg_kernel_get({name})")
            self.kernels[name] = k
        return self.kernels[name]

    def info(self, name):
        inf = KernelInfo()
        _check(load_library().bg_kernel_info(self.kernel(name), ctypes.byref(inf)), "bg_kernel_info")
        return {f: getattr(inf, f) for f, _ in KernelInfo._fields_}

    def free(self):
        if self.handle:
            _check(load_library().bg_module_unload(self.handle), "bg_module_unload")
            self.handle = None


# status words the static scratch of bg_group_launch_static holds
STATIC_STATUS_WORDS = 4 << 20


def group_launch_static(kernel, basis_ptr, result_ptr, n_words, n_tiles, n_carry, error_flag_ptr, stream):
    """bg_group_launch_static: one group kernel over one buffer with the
    library's static scratch (no workspace, no allocation); the launch's
    look-back error word lands in the device u32 at error_flag_ptr."""
    _check(load_library().bg_group_launch_static(kernel, basis_ptr, result_ptr, n_words, n_tiles, n_carry,
                                                 error_flag_ptr, stream), "bg_group_launch_static")


def filter_launch(kernel, basis_ptr, n_words, n_tiles, filter_ptr, stream, grid_blocks=0):
    """bg_filter_launch: a pattern set's byte-pair filter pre-pass over one
    buffer (answers of every sharing kernel, [n_tiles][row] u32 words)."""
    _check(load_library().bg_filter_launch(kernel, basis_ptr, n_words, n_tiles, filter_ptr, grid_blocks, stream),
           "bg_filter_launch")


def filter_batch_launch(kernel, jobs_ptr, tile_job_ptr, total_tiles, filter_ptr, stream, grid_blocks=0):
    """bg_filter_batch_launch: the set's filter pre-pass over a batch's
    global tiles (answers [total_tiles][row] u32 words)."""
    _check(load_library().bg_filter_batch_launch(kernel, jobs_ptr, tile_job_ptr, total_tiles, filter_ptr,
                                                 grid_blocks, stream), "bg_filter_batch_launch")


def group_launch(kernel, args, stream, grid_blocks=0):
    _check(load_library().bg_group_launch(kernel, ctypes.byref(args), grid_blocks, stream), "bg_group_launch")


def kernel_occupancy(kernel, block_threads):
    """(blocks per CU, CU count) of a loaded kernel at a block size."""
    n, cus = ctypes.c_int32(), ctypes.c_int32()
    _check(load_library().bg_kernel_occupancy(kernel, block_threads, ctypes.byref(n), ctypes.byref(cus)),
           "bg_kernel_occupancy")
    return n.value, cus.value


def new_ctl(device):
    """A zeroed control block (int32 tensor of CTL_WORDS words) on a device."""
    import torch

    return torch.zeros(CTL_WORDS, dtype=torch.int32, device=device)


def to_reference_layout(src_ptr, dst_ptr, n_words, n_streams, stream):
    _check(load_library().bg_to_reference_layout(src_ptr, dst_ptr, n_words, n_streams, stream),
           "bg_to_reference_layout")


def popcount(src_ptr, n_words, count_ptr, stream):
    _check(load_library().bg_popcount(src_ptr, n_words, count_ptr, stream), "bg_popcount")


def stream_sync(stream):
    _check(load_library().bg_stream_sync(stream), "bg_stream_sync")


def batch_launch(kernel, args, grid_blocks, block_threads, stream):
    _check(load_library().bg_batch_launch(kernel, ctypes.byref(args), grid_blocks, block_threads, stream),
           "bg_batch_launch")


class Graph:
    """A captured, instantiated graph of work enqueued on one stream."""

    def __init__(self, handle):
        self.handle = handle

    def launch(self, stream):
        _check(load_library().bg_graph_launch(self.handle, stream), "bg_graph_launch")

    def destroy(self):
        if self.handle:
            load_library().bg_graph_destroy(self.handle)
            self.handle = None


def graph_begin(stream):
    _check(load_library().bg_graph_begin(stream), "bg_graph_begin")


def graph_end(stream):
    h = ctypes.c_void_p()
    _check(load_library().bg_graph_end(stream, ctypes.byref(h)), "bg_graph_end")
    return Graph(h)


def transpose_batch(bytes_ptr, basis_ptr, descs_ptr, n_desc, total_words, stream):
    _check(load_library().bg_transpose_batch(bytes_ptr, basis_ptr, descs_ptr, n_desc, total_words, stream),
           "bg_transpose_batch")

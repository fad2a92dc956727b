"""MI355X-native JPEG 2000 (T.800 Part 1) decode path.

Python view of the C ABI in include/tc_device.h and include/tc_decode.h,
loaded from the in-tree libtc.so.  The library holds the HIP kernels (gfx950)
and the C++ host path; nothing here computes image data.  Importing works
without a GPU (parser entry points run on the host); decoding needs one, and
fails loudly when the library or the device is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtc.so")
# Packing values: 64 / 32 / 16 code-blocks per wavefront decoded in lockstep
# (TC_BPW_LOCKSTEP, TC_BPW_LOCKSTEP32, TC_BPW_LOCKSTEP16).
BPW_LOCKSTEP = -64
BPW_LOCKSTEP32 = -32
BPW_LOCKSTEP16 = -16

_lib = None


class TCError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{msg} (status {status})")
        self.status = status


class ImageInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "x0", "y0", "width", "height", "ncomp", "max_prec", "bytes_per_sample", "reversible",
        "nlevels", "ntx", "nty", "cbw", "cbh", "nlayers", "mct", "progression")] + [
        ("prec", ctypes.c_int32 * 4), ("sgnd", ctypes.c_int32 * 4), ("planar", ctypes.c_int32),
        ("comp_width", ctypes.c_int32 * 4), ("comp_height", ctypes.c_int32 * 4),
        ("jp2", ctypes.c_int32), ("colourspace", ctypes.c_int32)]


class CblkInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "tile", "comp", "res", "band", "x0", "y0", "x1", "y1", "npasses", "zero_bp", "mb",
        "nbytes")] + [("crc", ctypes.c_uint32)]


class DecoderStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("fill_ms", "t1_ms", "dwt_ms", "final_ms", "total_ms")] + [
        (n, ctypes.c_int64) for n in ("cblks", "codeword_bytes", "samples", "pixels", "dwt_bytes",
                                      "final_bytes")] + [
        ("bpw_small", ctypes.c_int32), ("bpw_large", ctypes.c_int32), ("images", ctypes.c_int32),
        ("device_bytes", ctypes.c_int64), ("t1_samples", ctypes.c_int64),
        ("t1_launches", ctypes.c_int32), ("t1_launch_class", ctypes.c_int32 * 4), ("t1_launch_bpw", ctypes.c_int32 * 4),
        ("t1_launch_blocks", ctypes.c_int64 * 4), ("t1_launch_bytes", ctypes.c_int64 * 4),
        ("t1_launch_ms", ctypes.c_float * 4), ("heavy_small", ctypes.c_int32), ("heavy_large", ctypes.c_int32)]


class DeviceProps(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("arch", ctypes.c_char * 32),
                ("compute_units", ctypes.c_int32), ("wavefront_size", ctypes.c_int32),
                ("xcds", ctypes.c_int32), ("lds_bytes_per_workgroup", ctypes.c_uint64),
                ("global_mem_bytes", ctypes.c_uint64), ("mem_base_align", ctypes.c_uint32),
                ("clock_khz", ctypes.c_int32), ("memory_clock_khz", ctypes.c_int32),
                ("memory_bus_bits", ctypes.c_int32)]


def lib():
    """The loaded libtc.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TCError(-1, f"{LIB_PATH} is missing: run thousandthchicken_amd/build.py (or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
    P = ctypes.POINTER
    sigs = {
        "tc_last_error": (ctypes.c_char_p, []),
        "tc_status_string": (ctypes.c_char_p, [i32]),
        "tc_device_count": (i32, [P(i32)]),
        "tc_device_open": (i32, [i32, P(vp)]),
        "tc_device_close": (i32, [vp]),
        "tc_device_get_props": (i32, [vp, P(DeviceProps)]),
        "tc_image_probe": (i32, [ctypes.c_char_p, sz, P(ImageInfo)]),
        "tc_image_probe_reduced": (i32, [ctypes.c_char_p, sz, i32, P(ImageInfo)]),
        "tc_decoder_set_reduce": (i32, [vp, i32]),
        "tc_decoder_set_region": (i32, [vp, i32, i32, i32, i32]),
        "tc_decoder_set_host_output": (i32, [vp, i32]),
        "tc_image_list_cblks": (i32, [ctypes.c_char_p, sz, vp, i32, P(i32)]),
        "tc_decoder_create": (i32, [vp, P(vp)]),
        "tc_decoder_destroy": (i32, [vp]),
        "tc_decoder_set_blocks_per_wave": (i32, [vp, i32, i32]),
        "tc_decoder_set_t1_kernel": (i32, [vp, i32]),
        "tc_decoder_autotune": (i32, [vp, i32, vp]),
        "tc_decoder_get_tuning": (i32, [vp, vp]),
        "tc_decoder_set_tuning": (i32, [vp, vp]),
        "tc_image_cblk_data": (i32, [ctypes.c_char_p, sz, i32, vp, sz, P(sz)]),
        "tc_decoder_decode": (i32, [vp, ctypes.c_char_p, sz, vp, sz, P(ImageInfo)]),
        "tc_decoder_prepare": (i32, [vp, i32, vp, vp]),
        "tc_decoder_run": (i32, [vp, i32]),
        "tc_decoder_sync": (i32, [vp]),
        "tc_decoder_fetch": (i32, [vp, i32, vp, sz]),
        "tc_decoder_output": (i32, [vp, i32, P(vp), P(sz)]),
        "tc_decoder_capture_graph": (i32, [vp]),
        "tc_decoder_launch_graph": (i32, [vp]),
        "tc_decoder_get_stats": (i32, [vp, P(DecoderStats)]),
        "tc_decoder_stream": (vp, [vp]),
        "tc_decode_staged": (i32, [vp, ctypes.c_char_p, sz, vp, sz]),
        "tc_decode_noalloc": (i32, [vp, ctypes.c_char_p, sz, vp, sz, P(ImageInfo)]),
        "tc_stream_create": (i32, [vp, P(vp)]),
        "tc_server_create": (i32, [vp, vp, P(vp)]),
        "tc_server_destroy": (i32, [vp]),
        "tc_server_submit": (i32, [vp, ctypes.c_char_p, sz, P(ctypes.c_uint64)]),
        "tc_server_submit_ref": (i32, [vp, ctypes.c_char_p, sz, P(ctypes.c_uint64)]),
        "tc_server_poll": (i32, [vp, ctypes.c_uint64, P(i32)]),
        "tc_server_wait": (i32, [vp, ctypes.c_uint64, ctypes.c_double]),
        "tc_server_result": (i32, [vp, ctypes.c_uint64, vp, sz, vp]),
        "tc_server_get_stats": (i32, [vp, vp]),
        "tc_server_release": (i32, [vp, ctypes.c_uint64]),
        "tc_server_reset_stats": (i32, [vp]),
        "tc_server_now": (ctypes.c_double, [vp]),
        "tc_stream_destroy": (i32, [vp]),
        "tc_stream_sync": (i32, [vp]),
        "tc_event_create": (i32, [P(vp)]),
        "tc_event_destroy": (i32, [vp]),
        "tc_event_record": (i32, [vp, vp]),
        "tc_event_sync": (i32, [vp]),
        "tc_event_query": (i32, [vp]),
        "tc_event_elapsed_ms": (i32, [vp, vp, P(ctypes.c_float)]),
        "tc_buf_alloc": (i32, [vp, sz, P(vp)]),
        "tc_buf_free": (i32, [vp]),
        "tc_buf_device_ptr": (vp, [vp]),
        "tc_buf_copy_h2d": (i32, [vp, vp, sz, vp, sz]),
        "tc_buf_copy_d2h": (i32, [vp, vp, vp, sz, sz]),
        "tc_buf_fill32": (i32, [vp, vp, sz, ctypes.c_uint32, sz]),
        "tc_buf_copy_d2d": (i32, [vp, vp, sz, vp, sz, sz]),
        "tc_launch_t1_decode_wave": (i32, [vp, i32, i32, vp, vp, i32, vp, sz, vp]),
        "tc_launch_t1_decode_stripe": (i32, [vp, i32, i32, vp, vp, i32, vp, sz, vp]),
        "tc_t1_lockstep_scratch_bytes": (sz, [i32, i32, i32, i32]),
        "tc_launch_t1_decode_lockstep": (i32, [vp, i32, i32, vp, vp, i32, i32, vp, sz, vp, sz, vp]),
        "tc_idwt_tile_shape": (i32, [i32, P(i32), P(i32), P(i32)]),
        "tc_launch_fdwt_level": (i32, [vp, i32, i32, vp, vp, i32, vp]),
        "tc_launch_fdwt_first": (i32, [vp, i32, i32, i32, i32, vp, vp, i32, vp, vp]),
        "tc_forward_create": (i32, [vp, P(vp)]),
        "tc_forward_destroy": (i32, [vp]),
        "tc_forward_prepare": (i32, [vp, i32, i32, i32, i32, i32, i32, i32]),
        "tc_forward_set_shape": (i32, [vp, i32]),
        "tc_forward_upload": (i32, [vp, vp]),
        "tc_forward_run": (i32, [vp]),
        "tc_forward_run_timed": (i32, [vp, P(ctypes.c_float), P(ctypes.c_float)]),
        "tc_forward_sync": (i32, [vp]),
        "tc_forward_download": (i32, [vp, vp]),
        "tc_forward_stream": (vp, [vp]),
        "tc_forward_image_ptr": (vp, [vp]),
        "tc_forward_planes_ptr": (vp, [vp]),
        "tc_dwt_plane": (i32, [vp, i32, i32, vp, vp, i32, i32, i32]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(status, what=""):
    if status != 0:
        msg = lib().tc_last_error().decode(errors="replace")
        raise TCError(status, f"{what}: {msg}" if what else msg)


def exported_symbols():
    """Names of every tc_* function declared in include/*.h."""
    import re
    inc = os.path.join(_HERE, "..", "include")
    names = set()
    for root, _, files in os.walk(inc):
        for f in files:
            if f.endswith(".h"):
                txt = open(os.path.join(root, f)).read()
                txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
                txt = re.sub(r"//[^\n]*", "", txt)
                names.update(re.findall(r"\b(tc_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def probe(cs: bytes, reduce=0) -> ImageInfo:
    """Main-header information; with reduce > 0, of the image decoded with its
    `reduce` finest DWT levels discarded."""
    info = ImageInfo()
    if reduce:
        check(lib().tc_image_probe_reduced(cs, len(cs), reduce, ctypes.byref(info)), "tc_image_probe_reduced")
    else:
        check(lib().tc_image_probe(cs, len(cs), ctypes.byref(info)), "tc_image_probe")
    return info


def list_cblks(cs: bytes):
    n = ctypes.c_int(0)
    check(lib().tc_image_list_cblks(cs, len(cs), None, 0, ctypes.byref(n)), "tc_image_list_cblks")
    arr = (CblkInfo * max(n.value, 1))()
    check(lib().tc_image_list_cblks(cs, len(cs), arr, n.value, ctypes.byref(n)), "tc_image_list_cblks")
    return [arr[i] for i in range(n.value)]


def cblk_data(cs: bytes, index: int) -> bytes:
    """Codeword bytes of one code-block (order of list_cblks)."""
    n = ctypes.c_size_t(0)
    check(lib().tc_image_cblk_data(cs, len(cs), index, None, 0, ctypes.byref(n)), "tc_image_cblk_data")
    buf = ctypes.create_string_buffer(max(n.value, 1))
    check(lib().tc_image_cblk_data(cs, len(cs), index, buf, n.value, ctypes.byref(n)), "tc_image_cblk_data")
    return buf.raw[:n.value]


class Device:
    def __init__(self, ordinal=0):
        h = ctypes.c_void_p()
        check(lib().tc_device_open(ordinal, ctypes.byref(h)), "tc_device_open")
        self.handle = h

    def props(self):
        p = DeviceProps()
        check(lib().tc_device_get_props(self.handle, ctypes.byref(p)), "tc_device_get_props")
        return p

    def close(self):
        if self.handle:
            lib().tc_device_close(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_device = None


def default_device():
    global _default_device
    if _default_device is None:
        _default_device = Device(0)
    return _default_device


def _shape(info: ImageInfo):
    dt = np.uint8 if info.bytes_per_sample == 1 else np.uint16
    if info.planar:
        return (sum(info.comp_width[c] * info.comp_height[c] for c in range(info.ncomp)),), dt
    return (info.height, info.width, info.ncomp), dt


def _unpack(out, info: ImageInfo):
    """Planar output (sub-sampled components) -> one 2-D array per component."""
    if not info.planar:
        return out
    res, off = [], 0
    for c in range(info.ncomp):
        w, h = info.comp_width[c], info.comp_height[c]
        res.append(out[off:off + w * h].reshape(h, w))
        off += w * h
    return res


def decode_staged(cs: bytes, device: "Device" = None) -> np.ndarray:
    """Decode through the drop-in stage classes (reference stage order)."""
    dev = device or default_device()
    info = probe(cs)
    shape, dt = _shape(info)
    out = np.empty(shape, dt)
    check(lib().tc_decode_staged(dev.handle, cs, len(cs), out.ctypes.data, out.nbytes), "tc_decode_staged")
    return out


CBLK_DESC = np.dtype([("data_off", "<u4"), ("data_len", "<u4"), ("out_off", "<u8"), ("out_pitch", "<u4"),
                      ("w", "<u2"), ("h", "<u2"), ("numbps", "u1"), ("npasses", "u1"), ("cls", "u1"),
                      ("mode", "u1"), ("delta", "<f4")])
assert CBLK_DESC.itemsize == 32


class Tuning(ctypes.Structure):
    _fields_ = [("bpw_small", ctypes.c_int32), ("bpw_large", ctypes.c_int32), ("t1_kernel", ctypes.c_int32),
                ("dwt_shape", ctypes.c_int32), ("t1_ms", ctypes.c_float), ("dwt_ms", ctypes.c_float),
                ("heavy_small", ctypes.c_int32), ("heavy_large", ctypes.c_int32), ("heavy_waves", ctypes.c_int32)]

    def as_dict(self):
        return {k: (round(getattr(self, k), 4) if k.endswith("_ms") else getattr(self, k)) for k, _ in self._fields_}


class ServerConfig(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int32), ("max_batch", ctypes.c_int32), ("memory_budget", ctypes.c_uint64),
                ("bpw_small", ctypes.c_int32), ("bpw_large", ctypes.c_int32), ("dwt_shape", ctypes.c_int32),
                ("max_batch_bytes", ctypes.c_uint64), ("overlap_kernels", ctypes.c_int32)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("ncomp", ctypes.c_int32), ("bytes_per_sample", ctypes.c_int32), ("submit_s", ctypes.c_double),
                ("start_s", ctypes.c_double), ("done_s", ctypes.c_double), ("batch", ctypes.c_uint64),
                ("batch_size", ctypes.c_int32), ("footprint", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("submitted", "completed", "failed", "batches", "pixels")] + [
        ("busy_s", ctypes.c_double), ("peak_bytes", ctypes.c_uint64), ("queued", ctypes.c_int32),
        ("in_flight", ctypes.c_int32), ("plan_s", ctypes.c_double), ("device_s", ctypes.c_double),
        ("copy_s", ctypes.c_double), ("slot_device_peak", ctypes.c_uint64)]


class Server:
    """Continuous-batching decode server over the C ABI (include/tc_server.h)."""

    def __init__(self, device: "Device" = None, slots=2, max_batch=256, memory_budget=32 << 30, bpw_small=0,
                 bpw_large=0, dwt_shape=0, max_batch_bytes=1 << 30, overlap_kernels=False):
        self.device = device or default_device()
        cfg = ServerConfig(slots, max_batch, memory_budget, bpw_small, bpw_large, dwt_shape, max_batch_bytes,
                           1 if overlap_kernels else 0)
        h = ctypes.c_void_p()
        check(lib().tc_server_create(self.device.handle, ctypes.byref(cfg), ctypes.byref(h)), "tc_server_create")
        self.handle = h
        self._hold = {}

    def submit(self, cs: bytes, copy=False) -> int:
        """Queues a codestream.  Without copy the server reads the bytes object
        itself, which this wrapper keeps alive until the request is read or released."""
        rid = ctypes.c_uint64()
        if copy:
            check(lib().tc_server_submit(self.handle, cs, len(cs), ctypes.byref(rid)), "submit")
        else:
            cs = bytes(cs)
            check(lib().tc_server_submit_ref(self.handle, cs, len(cs), ctypes.byref(rid)), "submit")
            self._hold[rid.value] = cs
        return rid.value

    def done(self, rid) -> bool:
        d = ctypes.c_int(0)
        check(lib().tc_server_poll(self.handle, rid, ctypes.byref(d)), "poll")
        return bool(d.value)

    def wait(self, rid, timeout=60.0) -> bool:
        st = lib().tc_server_wait(self.handle, rid, timeout)
        if st == 9:
            return False
        check(st, "wait")
        return True

    def info(self, rid) -> RequestInfo:
        i = RequestInfo()
        check(lib().tc_server_result(self.handle, rid, None, 0, ctypes.byref(i)), "result info")
        return i

    def result(self, rid):
        """(image array, RequestInfo); releases the request."""
        i = self.info(rid)
        if i.status != 0:
            self.release(rid)
            check(i.status, f"request {rid}")
        dt = np.uint8 if i.bytes_per_sample == 1 else np.uint16
        out = np.empty((i.height, i.width, i.ncomp), dt)
        check(lib().tc_server_result(self.handle, rid, out.ctypes.data, out.nbytes, None), "result")
        self._hold.pop(rid, None)
        return out, i

    def reset_stats(self):
        check(lib().tc_server_reset_stats(self.handle), "reset_stats")

    def release(self, rid):
        check(lib().tc_server_release(self.handle, rid), "release")
        self._hold.pop(rid, None)

    def now(self) -> float:
        return lib().tc_server_now(self.handle)

    def stats(self) -> ServerStats:
        s = ServerStats()
        check(lib().tc_server_get_stats(self.handle, ctypes.byref(s)), "stats")
        return s

    def close(self):
        if getattr(self, "handle", None):
            lib().tc_server_destroy(self.handle)  # drains the queue first
            self.handle = None
            self._hold.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """A device allocation through the C ABI (for tests of the typed launches)."""

    def __init__(self, device, nbytes):
        h = ctypes.c_void_p()
        check(lib().tc_buf_alloc(device.handle, nbytes, ctypes.byref(h)), "tc_buf_alloc")
        self.handle, self.nbytes = h, nbytes

    @property
    def ptr(self):
        return lib().tc_buf_device_ptr(self.handle)

    def upload(self, stream, arr):
        a = np.ascontiguousarray(arr)
        check(lib().tc_buf_copy_h2d(stream, self.handle, 0, a.ctypes.data, a.nbytes), "h2d")
        check(lib().tc_stream_sync(stream), "sync")

    def download(self, stream, dtype, count):
        out = np.empty(count, dtype)
        check(lib().tc_buf_copy_d2h(stream, out.ctypes.data, self.handle, 0, out.nbytes), "d2h")
        check(lib().tc_stream_sync(stream), "sync")
        return out

    def free(self):
        if self.handle:
            lib().tc_buf_free(self.handle)
            self.handle = None


class Decoder:
    """Decoder over the C ABI: per-image decode, and the resident/batched form."""

    def __init__(self, device: Device = None):
        self.device = device or default_device()
        h = ctypes.c_void_p()
        check(lib().tc_decoder_create(self.device.handle, ctypes.byref(h)), "tc_decoder_create")
        self.handle = h
        self._infos = []
        self._keep = None
        self._reduce = 0
        self._region = None

    def close(self):
        if getattr(self, "handle", None):
            lib().tc_decoder_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self, reps=3) -> "Tuning":
        """Times the prepared batch under candidate packings and DWT tile shapes; keeps the fastest."""
        t = Tuning()
        check(lib().tc_decoder_autotune(self.handle, reps, ctypes.byref(t)), "autotune")
        return t

    def tuning(self) -> "Tuning":
        t = Tuning()
        check(lib().tc_decoder_get_tuning(self.handle, ctypes.byref(t)), "get_tuning")
        return t

    def set_tuning(self, t: "Tuning"):
        check(lib().tc_decoder_set_tuning(self.handle, ctypes.byref(t)), "set_tuning")

    def set_t1_kernel(self, kind):
        check(lib().tc_decoder_set_t1_kernel(self.handle, kind), "set_t1_kernel")

    def set_reduce(self, reduce):
        """Decode with the `reduce` finest DWT levels discarded (0: full resolution)."""
        check(lib().tc_decoder_set_reduce(self.handle, reduce), "set_reduce")
        self._reduce = reduce

    def set_region(self, x=0, y=0, w=0, h=0):
        """Decode only the w x h window at (x, y) (clipped to the image); 0 sizes: the whole image."""
        check(lib().tc_decoder_set_region(self.handle, x, y, w, h), "set_region")
        self._region = (x, y, w, h) if w > 0 and h > 0 else None

    def _info(self, cs):
        info = probe(cs, self._reduce)
        if self._region:
            x, y, w, h = self._region
            info.width = max(0, min(w, info.width - x))
            info.height = max(0, min(h, info.height - y))
        return info

    def set_host_output(self, enable=True):
        """Output images in pinned host memory written directly by the final-level kernel."""
        check(lib().tc_decoder_set_host_output(self.handle, 1 if enable else 0), "set_host_output")

    def output_view(self, i) -> np.ndarray:
        """Image i in place (host output: a view of the pinned buffer; valid until the next prepare)."""
        p, n = ctypes.c_void_p(), ctypes.c_size_t()
        check(lib().tc_decoder_output(self.handle, i, ctypes.byref(p), ctypes.byref(n)), "output")
        shape, dt = _shape(self._infos[i])
        buf = (ctypes.c_uint8 * n.value).from_address(p.value)
        return _unpack(np.frombuffer(buf, dt).reshape(shape), self._infos[i])

    def set_blocks_per_wave(self, small, bpw):
        check(lib().tc_decoder_set_blocks_per_wave(self.handle, 1 if small else 0, bpw), "set_blocks_per_wave")

    def decode(self, cs: bytes) -> np.ndarray:
        info = self._info(cs)
        shape, dt = _shape(info)
        out = np.empty(shape, dt)
        check(lib().tc_decoder_decode(self.handle, cs, len(cs), out.ctypes.data, out.nbytes, None), "decode")
        return _unpack(out, info)

    def prepare(self, streams):
        streams = [bytes(s) for s in streams]
        self._infos = [self._info(s) for s in streams]
        n = len(streams)
        ptrs = (ctypes.c_char_p * n)(*streams)
        lens = (ctypes.c_size_t * n)(*[len(s) for s in streams])
        self._keep = (streams, ptrs, lens)
        check(lib().tc_decoder_prepare(self.handle, n, ctypes.cast(ptrs, ctypes.c_void_p),
                                       ctypes.cast(lens, ctypes.c_void_p)), "prepare")

    def run(self, timed=False):
        check(lib().tc_decoder_run(self.handle, 1 if timed else 0), "run")

    def sync(self):
        check(lib().tc_decoder_sync(self.handle), "sync")

    def capture_graph(self):
        check(lib().tc_decoder_capture_graph(self.handle), "capture_graph")

    def launch_graph(self):
        check(lib().tc_decoder_launch_graph(self.handle), "launch_graph")

    def fetch(self, i) -> np.ndarray:
        shape, dt = _shape(self._infos[i])
        out = np.empty(shape, dt)
        check(lib().tc_decoder_fetch(self.handle, i, out.ctypes.data, out.nbytes), "fetch")
        return _unpack(out, self._infos[i])

    def stats(self) -> DecoderStats:
        s = DecoderStats()
        check(lib().tc_decoder_get_stats(self.handle, ctypes.byref(s)), "stats")
        return s

    def stream_handle(self):
        return lib().tc_decoder_stream(self.handle)


class Forward:
    """GPU forward transform (encoder front end) over the C ABI: DC level shift,
    forward RCT/ICT and an L-level forward 5/3 or 9/7 DWT of one image, into one
    Mallat-layout plane per component (include/tc_forward.h)."""

    def __init__(self, device: "Device" = None):
        self.device = device or default_device()
        h = ctypes.c_void_p()
        check(lib().tc_forward_create(self.device.handle, ctypes.byref(h)), "tc_forward_create")
        self.handle = h
        self._geom = None

    def close(self):
        if getattr(self, "handle", None):
            lib().tc_forward_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_shape(self, shape):
        check(lib().tc_forward_set_shape(self.handle, shape), "set_shape")

    def prepare(self, w, h, ncomp, bits, levels, irreversible, mct):
        check(lib().tc_forward_prepare(self.handle, w, h, ncomp, bits, levels, 1 if irreversible else 0,
                                       1 if mct else 0), "forward prepare")
        self._geom = (w, h, ncomp, bits, irreversible)

    def upload(self, img: np.ndarray):
        w, h, nc, bits, _ = self._geom
        a = np.ascontiguousarray(img, np.uint16 if bits > 8 else np.uint8)
        assert a.size == w * h * nc
        check(lib().tc_forward_upload(self.handle, a.ctypes.data), "forward upload")

    def run(self):
        check(lib().tc_forward_run(self.handle), "forward run")

    def run_timed(self):
        """Runs once and returns (first-level ms, total ms) from stream events."""
        a, b = ctypes.c_float(), ctypes.c_float()
        check(lib().tc_forward_run_timed(self.handle, ctypes.byref(a), ctypes.byref(b)), "forward run_timed")
        return a.value, b.value

    def sync(self):
        check(lib().tc_forward_sync(self.handle), "forward sync")

    def download(self) -> np.ndarray:
        w, h, nc, _, irr = self._geom
        out = np.empty((nc, h, w), np.float32 if irr else np.int32)
        check(lib().tc_forward_download(self.handle, out.ctypes.data), "forward download")
        return out

    def stream_handle(self):
        return lib().tc_forward_stream(self.handle)


def forward_transform(img: np.ndarray, levels=5, irreversible=False, mct=None, bits=None,
                      device: "Device" = None) -> np.ndarray:
    """Encoder front end on the GPU: img [H, W] or [H, W, C] (uint8/uint16) ->
    [C, H, W] planes (int32 for 5/3 + RCT, float32 for 9/7 + ICT), each in the
    Mallat layout of an L-level forward DWT."""
    a = np.asarray(img)
    if a.ndim == 2:
        a = a[..., None]
    h, w, c = a.shape
    if bits is None:
        bits = 8 if a.dtype == np.uint8 else 16
    if mct is None:
        mct = c == 3
    f = Forward(device)
    try:
        f.prepare(w, h, c, bits, levels, irreversible, mct)
        f.upload(a)
        f.run()
        f.sync()
        return f.download()
    finally:
        f.close()


def dwt_plane(a: np.ndarray, levels, forward=True, device: "Device" = None) -> np.ndarray:
    """One plane through the DWT drop-in classes (DWTForward53/97 or
    DWTReverse53/97 by direction and dtype: int32 5/3, float32 9/7)."""
    a = np.ascontiguousarray(a)
    irr = a.dtype == np.float32
    if not irr:
        a = np.ascontiguousarray(a, np.int32)
    out = np.empty_like(a)
    dev = device or default_device()
    h, w = a.shape
    check(lib().tc_dwt_plane(dev.handle, 1 if forward else 0, 1 if irr else 0, a.ctypes.data, out.ctypes.data, w, h,
                             levels), "tc_dwt_plane")
    return out

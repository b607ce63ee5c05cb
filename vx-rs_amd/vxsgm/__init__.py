"""Python binding of the vxh SGM C ABI (include/vxh/sgm.h), shaped like the
reference's Rust API so callers and tests read the same way:

    Context()                       vx-vision/src/context.rs:26-36
    Context.texture_gray8           vx-vision/src/context.rs:39-41
    Context.texture_output_r32float vx-vision/src/context.rs:64-66
    Texture.read_r32float           vx-vision/src/texture.rs:404-427
    SGMStereoConfig                 vx-vision/src/kernels/sgm.rs:23-59
    SGMStereo(ctx)                  vx-vision/src/kernels/sgm.rs:73-105
    SGMStereo.compute_disparity     vx-vision/src/kernels/sgm.rs:108-265
    SGMStereo.compute               vx-vision/src/kernels/sgm.rs:269-288
    Error and its variants          vx-vision/src/error.rs:6-23

The library is the in-tree gfx950 build (vxsgm/libvxsgm.so).  Importing this
module without it raises ImportError: there is no CPU fallback.
"""
import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# VXSGM_LIB selects another build of the same library (same C ABI).
LIB_PATH = os.environ.get("VXSGM_LIB") or os.path.join(_HERE, "libvxsgm.so")
INCLUDE_DIR = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python vx-rs_amd/build_lib.py`")
lib = ctypes.CDLL(LIB_PATH)

_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_sz = ctypes.c_size_t


class vxh_sgm_config(ctypes.Structure):
    _fields_ = [("num_disparities", _u32), ("p1", ctypes.c_float),
                ("p2", ctypes.c_float), ("census_radius_x", _u32),
                ("census_radius_y", _u32), ("num_paths", _u32),
                ("lr_check", _u32), ("lr_tolerance", _u32)]


class vxh_camera_intrinsics(ctypes.Structure):
    _fields_ = [("fx", ctypes.c_float), ("fy", ctypes.c_float),
                ("cx", ctypes.c_float), ("cy", ctypes.c_float),
                ("width", _u32), ("height", _u32)]


class vxh_depth_map(ctypes.Structure):
    _fields_ = [("depth", _vp), ("intrinsics", vxh_camera_intrinsics),
                ("min_depth", ctypes.c_float), ("max_depth", ctypes.c_float)]


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_cfgp = ctypes.POINTER(vxh_sgm_config)
_sig("vxh_last_error", ctypes.c_char_p)
_sig("vxh_version", ctypes.c_char_p)
_sig("vxh_device_count", ctypes.c_int)
_sig("vxh_sgm_config_default", None, _cfgp)
_sig("vxh_sgm_config_new", None, _u32, _cfgp)
_sig("vxh_sgm_config_validate", ctypes.c_int, _cfgp, _u32, _u32)
_sig("vxh_context_create", ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp))
_sig("vxh_context_destroy", None, _vp)
_sig("vxh_texture_gray8", ctypes.c_int, _vp, _vp, _sz, _u32, _u32, ctypes.POINTER(_vp))
_sig("vxh_texture_output_r32float", ctypes.c_int, _vp, _u32, _u32, ctypes.POINTER(_vp))
_sig("vxh_texture_r32float", ctypes.c_int, _vp, _vp, _sz, _u32, _u32, ctypes.POINTER(_vp))
_sig("vxh_texture_format", _u32, _vp)
_sig("vxh_texture_output_gray8", ctypes.c_int, _vp, _u32, _u32, ctypes.POINTER(_vp))
_sig("vxh_texture_rgba8", ctypes.c_int, _vp, _vp, _sz, _u32, _u32, ctypes.POINTER(_vp))
_sig("vxh_texture_output_rgba8", ctypes.c_int, _vp, _u32, _u32, ctypes.POINTER(_vp))
_sig("vxh_texture_read_gray8", ctypes.c_int, _vp, _vp, _sz)
_sig("vxh_texture_read_rgba8", ctypes.c_int, _vp, _vp, _sz)
_sig("vxh_texture_device", ctypes.c_int, _vp)
_sig("vxh_context_device", ctypes.c_int, _vp)
_sig("vxh_texture_read_r32float", ctypes.c_int, _vp, _vp, _sz)
_sig("vxh_texture_write_gray8", ctypes.c_int, _vp, _vp, _sz)
_sig("vxh_texture_width", _u32, _vp)
_sig("vxh_texture_height", _u32, _vp)
_sig("vxh_texture_device_ptr", _vp, _vp)
_sig("vxh_texture_destroy", None, _vp)
_sig("vxh_sgm_create", ctypes.c_int, _vp, ctypes.POINTER(_vp))
_sig("vxh_sgm_destroy", None, _vp)
_sig("vxh_sgm_compute_disparity", ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _cfgp)
_sig("vxh_sgm_compute", ctypes.c_int, _vp, _vp, _vp, _vp, _cfgp,
     ctypes.POINTER(vxh_camera_intrinsics), ctypes.c_float,
     ctypes.POINTER(vxh_depth_map))
_sig("vxh_sgm_compute_host", ctypes.c_int, _vp, _vp, _vp, _u32, _u32, _cfgp, _vp)
_sig("vxh_sgm_workspace_bytes", _sz, _u32, _u32, _cfgp)
_sig("vxh_sgm_enqueue", ctypes.c_int, _vp, _vp, _vp, _u32, _u32, _cfgp, _vp, _vp,
     _sz, _vp)
_sig("vxh_sgm_enqueue_marked", ctypes.c_int, _vp, _vp, _vp, _u32, _u32, _cfgp, _vp, _vp,
     _sz, _vp, ctypes.POINTER(_vp))
_sig("vxh_sgm_workspace_status", ctypes.c_int, _vp, _vp)
_sig("vxh_sgm_enqueue_static", ctypes.c_int, _vp, _vp, _vp, _u32, _u32, _cfgp, _vp, _vp)
_sig("vxh_sgm_static_scratch_bytes", _sz)
_sig("vxh_sgm_set_split_wait", ctypes.c_int, _vp, _u32)
_sig("vxh_sgm_volume_format_in_effect", _u32, _vp, _u32, _cfgp)
_sig("vxh_sgm_workspace_error_offset", _sz)
_sig("vxh_texture_in_flight", ctypes.c_int, _vp)
_sig("vxh_sgm_census", ctypes.c_int, _vp, _u32, _u32, _u32, _u32, _vp, _vp)
_sig("vxh_sgm_aggregate", ctypes.c_int, _vp, _vp, _u32, _u32, _cfgp, _vp, _vp)
_sig("vxh_sgm_wta", ctypes.c_int, _vp, _u32, _u32, _cfgp, _vp, _vp)
_sig("vxh_sgm_volume_stride", _u32, _u32)
_sig("vxh_sgm_set_schedule", ctypes.c_int, _vp, _u32, _u32, _u32, _u32, _u32)
_sig("vxh_sgm_get_schedule", ctypes.c_int, _vp, _u32, ctypes.POINTER(_u32))
_sig("vxh_sgm_schedule_for", ctypes.c_int, _vp, _u32, _u32, _cfgp, ctypes.POINTER(_u32))
_sig("vxh_sgm_set_split", ctypes.c_int, _vp, _u32, _u32)
_sig("vxh_sgm_set_volume_format", ctypes.c_int, _vp, _u32, _u32)
_sig("vxh_util_copy", ctypes.c_int, _vp, _vp, _sz, _vp)
_sig("vxh_sgm_aggregate_layout", ctypes.c_int, _vp, _vp, _u32, _u32, _cfgp, _u32, _u32,
     _vp, _vp)
_sig("vxh_synth_pair", ctypes.c_int, _u32, _u32, _u32, _u32, _u32,
     ctypes.c_uint64, _vp, _vp, _vp)


# ---- errors (vx-vision/src/error.rs:6-23) ----
class Error(Exception):
    """Base of every failure the library reports."""
    code = -1


class DeviceNotFound(Error):
    code = 1


class ShaderMissing(Error):
    code = 2


class PipelineCompile(Error):
    code = 3


class BufferAlloc(Error):
    code = 4


class TextureSizeMismatch(Error):
    code = 5


class InvalidConfig(Error):
    code = 6


class Gpu(Error):
    code = 7


_ERRORS = {c.code: c for c in (DeviceNotFound, ShaderMissing, PipelineCompile,
                               BufferAlloc, TextureSizeMismatch, InvalidConfig, Gpu)}


def check(status):
    """Raises the Error subclass for a non-zero vxh_status."""
    if status != 0:
        msg = lib.vxh_last_error().decode(errors="replace")
        raise _ERRORS.get(status, Error)(msg)


def device_count():
    return lib.vxh_device_count()


# ---- config (vx-vision/src/kernels/sgm.rs:23-59) ----
@dataclass
class SGMStereoConfig:
    num_disparities: int = 128
    p1: float = 10.0
    p2: float = 120.0
    census_radius_x: int = 4
    census_radius_y: int = 3
    num_paths: int = 2
    lr_check: bool = False
    lr_tolerance: int = 1

    @classmethod
    def new(cls, num_disparities):
        return cls(num_disparities=num_disparities)

    def to_c(self):
        return vxh_sgm_config(self.num_disparities, self.p1, self.p2,
                              self.census_radius_x, self.census_radius_y,
                              self.num_paths, int(bool(self.lr_check)),
                              self.lr_tolerance)

    def validate(self, w, h):
        c = self.to_c()
        check(lib.vxh_sgm_config_validate(ctypes.byref(c), w, h))


def _as_u8(a):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    return a


class Texture:
    """Device texture owned by this object (R8 gray or R32 float)."""

    def __init__(self, handle):
        self._h = _vp(handle)

    @property
    def width(self):
        return lib.vxh_texture_width(self._h)

    @property
    def height(self):
        return lib.vxh_texture_height(self._h)

    @property
    def device_ptr(self):
        return lib.vxh_texture_device_ptr(self._h)

    @property
    def format(self):
        """TextureFormat value: GRAY8 (1), R32FLOAT (2) or RGBA8 (3)."""
        return lib.vxh_texture_format(self._h)

    def read_r32float(self):
        out = np.empty(self.width * self.height, dtype=np.float32)
        check(lib.vxh_texture_read_r32float(self._h, out.ctypes.data, out.size))
        return out

    def read_gray8(self):
        """Texture::read_gray8 (vx-vision/src/texture.rs:352-376)."""
        out = np.empty(self.width * self.height, dtype=np.uint8)
        check(lib.vxh_texture_read_gray8(self._h, out.ctypes.data, out.size))
        return out

    def read_rgba8(self):
        """Texture::read_rgba8 (vx-vision/src/texture.rs:378-402)."""
        out = np.empty(self.width * self.height * 4, dtype=np.uint8)
        check(lib.vxh_texture_read_rgba8(self._h, out.ctypes.data, out.size))
        return out

    def _release(self):
        """Gives up ownership of the handle (moved into a pipeline or a pool)."""
        h, self._h = self._h, _vp()
        if not h:
            raise InvalidConfig("invalid config: texture already moved")
        return h

    def write_gray8(self, pixels):
        """Refused (InvalidConfig) while a launched pipeline that uses the
        texture has not finished (UnifiedBuffer::write, vx-core/src/buffer.rs:68-72)."""
        p = _as_u8(pixels).ravel()
        check(lib.vxh_texture_write_gray8(self._h, p.ctypes.data, p.size))

    def is_in_flight(self):
        """UnifiedBuffer::is_in_flight (vx-core/src/buffer.rs:89-92)."""
        return bool(lib.vxh_texture_in_flight(self._h))

    def close(self):
        if self._h:
            lib.vxh_texture_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One HIP device (Context::new, vx-vision/src/context.rs:26-36)."""

    def __init__(self, device=0):
        h = _vp()
        check(lib.vxh_context_create(device, ctypes.byref(h)))
        self._h = h
        self.device = device

    def texture_gray8(self, pixels, width, height):
        p = _as_u8(pixels).ravel()
        h = _vp()
        check(lib.vxh_texture_gray8(self._h, p.ctypes.data, p.size, width, height,
                                    ctypes.byref(h)))
        return Texture(h.value)

    def texture_output_r32float(self, width, height):
        h = _vp()
        check(lib.vxh_texture_output_r32float(self._h, width, height, ctypes.byref(h)))
        return Texture(h.value)

    def texture_output_gray8(self, width, height):
        """Context::texture_output_gray8 (vx-vision/src/context.rs:49-51)."""
        h = _vp()
        check(lib.vxh_texture_output_gray8(self._h, width, height, ctypes.byref(h)))
        return Texture(h.value)

    # Context::texture_intermediate_* (context.rs:69-76): device memory either way.
    texture_intermediate_gray8 = texture_output_gray8

    def texture_intermediate_r32float(self, width, height):
        return self.texture_output_r32float(width, height)

    def texture_rgba8(self, pixels, width, height):
        """Context::texture_rgba8 (vx-vision/src/context.rs:54-56)."""
        p = _as_u8(pixels).ravel()
        h = _vp()
        check(lib.vxh_texture_rgba8(self._h, p.ctypes.data, p.size, width, height,
                                    ctypes.byref(h)))
        return Texture(h.value)

    def texture_output_rgba8(self, width, height):
        """Context::texture_output_rgba8 (vx-vision/src/context.rs:59-61)."""
        h = _vp()
        check(lib.vxh_texture_output_rgba8(self._h, width, height, ctypes.byref(h)))
        return Texture(h.value)

    def pool_stats(self):
        """(blocks, bytes held, bytes in use) of the context's scratch pool."""
        out = (_sz * 3)()
        check(lib.vxh_context_pool_stats(self._h, out))
        return tuple(out)

    def pool_trim(self):
        lib.vxh_context_pool_trim(self._h)

    def texture_r32float(self, data, width, height):
        """Context::texture_r32float (vx-vision/src/context.rs:44-46)."""
        d = np.ascontiguousarray(data, dtype=np.float32).ravel()
        h = _vp()
        check(lib.vxh_texture_r32float(self._h, d.ctypes.data, d.size, width, height,
                                       ctypes.byref(h)))
        return Texture(h.value)

    def close(self):
        if self._h:
            lib.vxh_context_destroy(self._h)
            self._h = _vp()


@dataclass
class CameraIntrinsics:
    fx: float
    fy: float
    cx: float
    cy: float
    width: int
    height: int


@dataclass
class DepthMap:
    depth: Texture
    intrinsics: CameraIntrinsics
    min_depth: float
    max_depth: float


class SGMStereo:
    """SGM kernel struct (vx-vision/src/kernels/sgm.rs:61-288)."""

    def __init__(self, ctx):
        h = _vp()
        check(lib.vxh_sgm_create(ctx._h, ctypes.byref(h)))
        self._h = h
        self._ctx = ctx

    def compute_disparity(self, ctx, left, right, output, config):
        c = config.to_c()
        check(lib.vxh_sgm_compute_disparity(self._h, ctx._h, left._h, right._h,
                                            output._h, ctypes.byref(c)))

    def encode(self, pipe, left, right, output, config):
        """The encode() form of compute_disparity: recorded into a Pipeline,
        run on its commit, same result."""
        c = config.to_c()
        check(lib.vxh_sgm_encode(self._h, pipe._h, left._h, right._h, output._h,
                                 ctypes.byref(c)))

    def compute(self, ctx, left, right, config, intrinsics, baseline):
        c = config.to_c()
        k = vxh_camera_intrinsics(intrinsics.fx, intrinsics.fy, intrinsics.cx,
                                  intrinsics.cy, intrinsics.width, intrinsics.height)
        out = vxh_depth_map()
        check(lib.vxh_sgm_compute(self._h, ctx._h, left._h, right._h, ctypes.byref(c),
                                  ctypes.byref(k), baseline, ctypes.byref(out)))
        return DepthMap(Texture(out.depth), intrinsics, out.min_depth, out.max_depth)

    def compute_host(self, left, right, width, height, config):
        """Host arrays in, host disparity map out (upload, compute, read back)."""
        c = config.to_c()
        l = _as_u8(left).ravel()
        r = _as_u8(right).ravel()
        out = np.empty(width * height, dtype=np.float32)
        check(lib.vxh_sgm_compute_host(self._h, l.ctypes.data, r.ctypes.data, width,
                                       height, ctypes.byref(c), out.ctypes.data))
        return out

    def set_schedule(self, num_disparities, schedule):
        """Lane layouts for pairs with this D: ((lanes_h, per_lane_h), (lanes,
        per_lane)) for the horizontal and the other directions, or one
        (lanes, per_lane) for all of them; None = default."""
        if not schedule:
            h = o = (0, 0)
        elif isinstance(schedule[0], (tuple, list)):
            h, o = schedule
        else:
            h = o = schedule
        check(lib.vxh_sgm_set_schedule(self._h, num_disparities, h[0], h[1], o[0], o[1]))

    def schedule(self, num_disparities):
        """((lanes_h, per_lane_h), (lanes, per_lane)) in effect for this D
        (for 1280x720 8-path pairs)."""
        out = (_u32 * 5)()
        check(lib.vxh_sgm_get_schedule(self._h, num_disparities, out))
        return (out[0], out[1]), (out[2], out[3])

    def schedule_for(self, width, height, config):
        """((lanes_h, per_lane_h), (lanes, per_lane), split) that a call on a
        width x height pair with `config` uses."""
        out = (_u32 * 5)()
        c = config.to_c()
        check(lib.vxh_sgm_schedule_for(self._h, width, height, ctypes.byref(c), out))
        return (out[0], out[1]), (out[2], out[3]), int(out[4])

    def split(self, num_disparities):
        """Split level for this D (for 1280x720 8-path pairs): 0 none, 1 the
        longest passes' scanlines run as two halves at once, 2 every long
        pass's."""
        out = (_u32 * 5)()
        check(lib.vxh_sgm_get_schedule(self._h, num_disparities, out))
        return int(out[4])

    def set_split(self, num_disparities, split):
        """split: False/0, True/1 (the longest passes), 2 (every long pass)
        or None (the default, which depends on size and path count)."""
        level = 0xFFFFFFFF if split is None else int(split)
        check(lib.vxh_sgm_set_split(self._h, num_disparities, level))

    def set_volume_format(self, num_disparities, fmt):
        """Per-direction volume storage for this D: "auto", "u8" or "u16"."""
        code = {"auto": 0, "u8": 1, "u16": 2}.get(fmt, 0xFFFFFFFF)   # the library rejects others
        check(lib.vxh_sgm_set_volume_format(self._h, num_disparities, code))

    def enqueue(self, d_left, d_right, width, height, config, d_out, d_workspace,
                workspace_bytes, stream=None):
        c = config.to_c()
        check(lib.vxh_sgm_enqueue(self._h, d_left, d_right, width, height,
                                  ctypes.byref(c), d_out, d_workspace,
                                  workspace_bytes, stream))

    def volume_format_in_effect(self, num_disparities, config):
        """1 (u8) or 2 (u16): the per-direction volume storage a call with this
        config uses on this handle."""
        c = config.to_c()
        return lib.vxh_sgm_volume_format_in_effect(self._h, num_disparities, ctypes.byref(c))

    def set_split_wait(self, polls=None):
        """Polls of a split head's bounded wait (None: the default)."""
        check(lib.vxh_sgm_set_split_wait(self._h, 0xFFFFFFFF if polls is None else polls))

    def workspace_status(self, d_workspace):
        """After the stream of an enqueue on d_workspace has drained: raises Gpu
        when a split scanline's bounded wait expired in that call."""
        check(lib.vxh_sgm_workspace_status(self._h, d_workspace))

    def enqueue_marked(self, d_left, d_right, width, height, config, d_out, d_workspace,
                       workspace_bytes, stream, events):
        """enqueue recording events[0..2] after census, aggregation and winner."""
        c = config.to_c()
        ev = (_vp * 3)(*[e.handle for e in events])
        check(lib.vxh_sgm_enqueue_marked(self._h, d_left, d_right, width, height,
                                         ctypes.byref(c), d_out, d_workspace,
                                         workspace_bytes, stream, ev))

    def close(self):
        if self._h:
            lib.vxh_sgm_destroy(self._h)
            self._h = _vp()


class vxh_sgm_pair_desc(ctypes.Structure):
    _fields_ = [("d_left", _vp), ("d_right", _vp), ("d_out", _vp), ("width", _u32),
                ("height", _u32), ("cfg", vxh_sgm_config), ("lanes", _u32),
                ("per_lane", _u32), ("lanes_h", _u32), ("per_lane_h", _u32),
                ("volume_format", _u32), ("split", _u32)]


class vxh_sgm_batch_options(ctypes.Structure):
    _fields_ = [("streams", _u32), ("form", _u32), ("blocks_per_cu", _u32),
                ("reserved", _u32), ("budget_bytes", ctypes.c_uint64)]


class vxh_sgm_batch_stats(ctypes.Structure):
    _fields_ = [("stream", _u32), ("lanes", _u32), ("per_lane", _u32), ("lanes_h", _u32),
                ("per_lane_h", _u32), ("volume_format", _u32), ("workspace_bytes", ctypes.c_uint64),
                ("chunk", _u32), ("form", _u32)]


BATCH_FORMS = {"persistent": 0, "streams": 1, "sweep": 2}


_sig("vxh_sgm_batch_create", ctypes.c_int, ctypes.c_int, ctypes.POINTER(vxh_sgm_pair_desc),
     _u32, ctypes.POINTER(_vp))
_sig("vxh_sgm_batch_create_ex", ctypes.c_int, ctypes.c_int,
     ctypes.POINTER(vxh_sgm_pair_desc), _u32, ctypes.POINTER(vxh_sgm_batch_options),
     ctypes.POINTER(_vp))
_sig("vxh_sgm_batch_launch", ctypes.c_int, _vp, _vp)
_sig("vxh_sgm_batch_run", ctypes.c_int, _vp)
_sig("vxh_sgm_batch_stats_get", ctypes.c_int, _vp, _u32, ctypes.POINTER(vxh_sgm_batch_stats))
_sig("vxh_sgm_batch_items", _u32, _vp)
_sig("vxh_sgm_batch_chunks", _u32, _vp)
_sig("vxh_sgm_batch_destroy", None, _vp)
_sig("vxh_sgm_batch_last_error", ctypes.c_char_p)


def _check_batch(status):
    if status != 0:
        msg = lib.vxh_sgm_batch_last_error().decode(errors="replace")
        raise _ERRORS.get(status, Error)(msg)


class SGMBatch:
    """Many pairs replayed from one hipGraph over several streams.

    `pairs` is a list of (d_left, d_right, d_out, width, height, config,
    schedule[, volume_format[, split]]) with device addresses; schedule is
    (lanes, per_lane) for every direction, ((lanes_h, per_lane_h),
    (lanes, per_lane)), or None; volume_format "auto", "u8" or "u16"; split
    a bool.  `options` is None (defaults) or a dict with any of "form"
    ("persistent", "streams" or "sweep"), "streams", "blocks_per_cu", "budget_bytes";
    a tuple (streams,) selects the streams form with that many streams."""

    def __init__(self, pairs, device=0, options=None):
        arr = (vxh_sgm_pair_desc * len(pairs))()
        for i, desc in enumerate(pairs):
            l, r, o, w, h, cfg, sched = desc[:7]
            fmt = {"auto": 0, "u8": 1, "u16": 2}[desc[7] if len(desc) > 7 else "auto"]
            split = int(desc[8]) if len(desc) > 8 else 0
            if not sched:
                hs = os = (0, 0)
            elif isinstance(sched[0], (tuple, list)):
                hs, os = sched
            else:
                hs = os = sched
            arr[i] = vxh_sgm_pair_desc(l, r, o, w, h, cfg.to_c(), os[0], os[1], hs[0],
                                       hs[1], fmt, split)
        h = _vp()
        opt = None
        if isinstance(options, dict):
            o = dict(options)
            o["form"] = BATCH_FORMS[o.get("form", "persistent")]
            opt = ctypes.byref(vxh_sgm_batch_options(
                o.get("streams", 0), o["form"], o.get("blocks_per_cu", 0), 0,
                o.get("budget_bytes", 0)))
        elif options:
            opt = ctypes.byref(vxh_sgm_batch_options(options[0], 1, 0, 0, 0))
        _check_batch(lib.vxh_sgm_batch_create_ex(device, arr, len(pairs), opt, ctypes.byref(h)))
        self._h = h
        self.n = len(pairs)

    def launch(self, stream=None):
        _check_batch(lib.vxh_sgm_batch_launch(self._h, stream))

    def run(self):
        _check_batch(lib.vxh_sgm_batch_run(self._h))

    def stats(self, i):
        s = vxh_sgm_batch_stats()
        _check_batch(lib.vxh_sgm_batch_stats_get(self._h, i, ctypes.byref(s)))
        return {"stream": s.stream, "schedule": ((s.lanes_h, s.per_lane_h), (s.lanes, s.per_lane)),
                "volume_format": {1: "u8", 2: "u16"}[s.volume_format],
                "workspace_bytes": s.workspace_bytes, "chunk": s.chunk,
                "form": {0: "persistent", 1: "streams", 2: "sweep"}[s.form]}

    def chunks(self):
        return lib.vxh_sgm_batch_chunks(self._h)

    @property
    def items(self):
        return lib.vxh_sgm_batch_items(self._h)

    def close(self):
        if self._h:
            lib.vxh_sgm_batch_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class vxh_server_config(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("max_inflight", _u32),
                ("num_streams", _u32), ("work_budget_pxd", ctypes.c_uint64),
                ("busy_work_budget_pxd", ctypes.c_uint64), ("backlog_requests", _u32),
                ("reserved", _u32)]


class vxh_server_stats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("rejected", ctypes.c_uint64), ("peak_admitted_bytes", ctypes.c_uint64),
                ("peak_inflight", ctypes.c_uint64), ("slot_hits", ctypes.c_uint64),
                ("slot_misses", ctypes.c_uint64), ("p50_ms", ctypes.c_double),
                ("p99_ms", ctypes.c_double), ("max_ms", ctypes.c_double),
                ("mean_queue_ms", ctypes.c_double), ("pxd_total", ctypes.c_double),
                ("first_submit_s", ctypes.c_double), ("last_complete_s", ctypes.c_double),
                ("peak_inflight_pxd", ctypes.c_double)]


_sig("vxh_server_request_bytes", ctypes.c_uint64, _u32, _u32, _cfgp)
_sig("vxh_server_create", ctypes.c_int, ctypes.c_int, ctypes.POINTER(vxh_server_config),
     ctypes.POINTER(_vp))
_sig("vxh_server_reserve", ctypes.c_int, _vp, _u32, _u32, _cfgp, _u32)
_sig("vxh_server_submit", ctypes.c_int, _vp, _vp, _vp, _u32, _u32, _cfgp, _vp,
     ctypes.POINTER(ctypes.c_uint64))
_sig("vxh_server_wait", ctypes.c_int, _vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double))
_sig("vxh_server_drain", ctypes.c_int, _vp)
_sig("vxh_server_get_stats", ctypes.c_int, _vp, ctypes.POINTER(vxh_server_stats))
_sig("vxh_server_reset_stats", None, _vp)
_sig("vxh_server_destroy", None, _vp)
_sig("vxh_server_last_error", ctypes.c_char_p)


def _check_server(status):
    if status != 0:
        msg = lib.vxh_server_last_error().decode(errors="replace")
        raise _ERRORS.get(status, Error)(msg)


def request_bytes(width, height, config):
    c = config.to_c()
    return lib.vxh_server_request_bytes(width, height, ctypes.byref(c))


class SGMServer:
    """Serving loop with HBM-budget admission (include/vxh/sgm_server.h)."""

    def __init__(self, hbm_budget_bytes=64 << 30, max_inflight=16, num_streams=8, device=0,
                 work_budget_pxd=0, busy_work_budget_pxd=0, backlog_requests=0):
        cfg = vxh_server_config(hbm_budget_bytes, max_inflight, num_streams, work_budget_pxd,
                                busy_work_budget_pxd, backlog_requests, 0)
        h = _vp()
        _check_server(lib.vxh_server_create(device, ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._keep = {}

    def reserve(self, width, height, config, count=1):
        """Pre-allocates `count` device slots for pairs of this shape (or smaller)."""
        c = config.to_c()
        _check_server(lib.vxh_server_reserve(self._h, width, height, ctypes.byref(c), count))

    def submit(self, left, right, width, height, config, out):
        """Queues a request; `out` (float32, w*h) is filled on completion."""
        l = _as_u8(left).ravel()
        r = _as_u8(right).ravel()
        assert out.dtype == np.float32 and out.flags.c_contiguous and out.size == width * height
        c = config.to_c()
        rid = ctypes.c_uint64()
        _check_server(lib.vxh_server_submit(self._h, l.ctypes.data, r.ctypes.data, width,
                                            height, ctypes.byref(c), out.ctypes.data,
                                            ctypes.byref(rid)))
        self._keep[rid.value] = out
        return rid.value

    def wait(self, rid):
        lat = ctypes.c_double()
        try:
            _check_server(lib.vxh_server_wait(self._h, rid, ctypes.byref(lat)))
        finally:
            self._keep.pop(rid, None)
        return lat.value

    def drain(self):
        _check_server(lib.vxh_server_drain(self._h))

    def wait_all(self):
        """Waits for (and so releases) every request submitted and not yet
        waited for; raises the first failure after releasing all."""
        err = None
        for rid in list(self._keep):
            try:
                self.wait(rid)
            except Error as e:
                err = err or e
        if err is not None:
            raise err

    def stats(self):
        s = vxh_server_stats()
        _check_server(lib.vxh_server_get_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in vxh_server_stats._fields_}

    def reset_stats(self):
        lib.vxh_server_reset_stats(self._h)

    def close(self):
        if self._h:
            lib.vxh_server_destroy(self._h)
            self._h = _vp()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class vxh_depth_filter_config(ctypes.Structure):
    _fields_ = [("radius", ctypes.c_int32), ("sigma_spatial", ctypes.c_float),
                ("sigma_depth", ctypes.c_float)]


class vxh_depth_median_config(ctypes.Structure):
    _fields_ = [("radius", ctypes.c_int32)]


_sig("vxh_depth_filter_config_default", None, ctypes.POINTER(vxh_depth_filter_config))
_sig("vxh_depth_median_config_default", None, ctypes.POINTER(vxh_depth_median_config))
_sig("vxh_depth_filter_create", ctypes.c_int, _vp, ctypes.POINTER(_vp))
_sig("vxh_depth_filter_destroy", None, _vp)
_sig("vxh_depth_filter_apply_bilateral", ctypes.c_int, _vp, _vp, _vp, _vp,
     ctypes.POINTER(vxh_depth_filter_config))
_sig("vxh_depth_filter_apply_median", ctypes.c_int, _vp, _vp, _vp, _vp,
     ctypes.POINTER(vxh_depth_median_config))
_sig("vxh_depth_bilateral_enqueue", ctypes.c_int, _vp, _vp, _u32, _u32,
     ctypes.POINTER(vxh_depth_filter_config), _vp)
_sig("vxh_depth_median_enqueue", ctypes.c_int, _vp, _vp, _u32, _u32,
     ctypes.POINTER(vxh_depth_median_config), _vp)
_sig("vxh_disparity_to_depth", ctypes.c_int, _vp, _vp, _u32, _u32, ctypes.c_float,
     ctypes.c_float, _vp)


@dataclass
class DepthFilterConfig:
    """DepthFilterConfig (vx-vision/src/kernels/depth_filter.rs:22-50)."""
    radius: int = 3
    sigma_spatial: float = 5.0
    sigma_depth: float = 0.05

    def to_c(self):
        return vxh_depth_filter_config(self.radius, self.sigma_spatial, self.sigma_depth)


@dataclass
class DepthMedianConfig:
    """DepthMedianConfig (vx-vision/src/kernels/depth_filter.rs:54-63)."""
    radius: int = 1

    def to_c(self):
        return vxh_depth_median_config(self.radius)


class DepthFilter:
    """DepthFilter (vx-vision/src/kernels/depth_filter.rs:66-249): bilateral
    and median filters of R32Float depth maps."""

    def __init__(self, ctx):
        h = _vp()
        check(lib.vxh_depth_filter_create(ctx._h, ctypes.byref(h)))
        self._h = h

    def apply_bilateral(self, ctx, input, output, config):
        c = config.to_c()
        check(lib.vxh_depth_filter_apply_bilateral(self._h, ctx._h, input._h, output._h,
                                                   ctypes.byref(c)))

    def apply_median(self, ctx, input, output, config):
        c = config.to_c()
        check(lib.vxh_depth_filter_apply_median(self._h, ctx._h, input._h, output._h,
                                                ctypes.byref(c)))

    def encode_bilateral(self, pipe, input, output, config):
        """DepthFilter::encode_bilateral (depth_filter.rs:183-214) into a Pipeline."""
        c = config.to_c()
        check(lib.vxh_depth_filter_encode_bilateral(self._h, pipe._h, input._h, output._h,
                                                    ctypes.byref(c)))

    def encode_median(self, pipe, input, output, config):
        """DepthFilter::encode_median (depth_filter.rs:217-246) into a Pipeline."""
        c = config.to_c()
        check(lib.vxh_depth_filter_encode_median(self._h, pipe._h, input._h, output._h,
                                                 ctypes.byref(c)))

    @staticmethod
    def enqueue_bilateral(d_in, d_out, width, height, config, stream=None):
        """Stream-ordered bilateral filter on device buffers of width*height floats."""
        c = config.to_c()
        check(lib.vxh_depth_bilateral_enqueue(d_in, d_out, width, height, ctypes.byref(c),
                                              stream))

    @staticmethod
    def enqueue_median(d_in, d_out, width, height, config, stream=None):
        """Stream-ordered median filter on device buffers of width*height floats."""
        c = config.to_c()
        check(lib.vxh_depth_median_enqueue(d_in, d_out, width, height, ctypes.byref(c), stream))

    def close(self):
        if self._h:
            lib.vxh_depth_filter_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def disparity_to_depth(d_disp, d_depth, width, height, fx, baseline, stream=None):
    """(fx * baseline) / d where d > 0, else 0, on device buffers."""
    check(lib.vxh_disparity_to_depth(d_disp, d_depth, width, height, fx, baseline, stream))


# ---- pipelines and texture pools (include/vxh/pipeline.h) ----
_texp = ctypes.POINTER(_vp)
_sig("vxh_pipeline_begin", ctypes.c_int, _vp, _texp)
_sig("vxh_pipeline_stream", _vp, _vp)
_sig("vxh_pipeline_retain", ctypes.c_int, _vp, _vp)
_sig("vxh_pipeline_commit", ctypes.c_int, _vp)
_sig("vxh_pipeline_wait", ctypes.c_int, _vp)
_sig("vxh_pipeline_commit_and_wait", ctypes.c_int, _vp)
_sig("vxh_pipeline_take_retained", _sz, _vp, _texp, _sz)
_sig("vxh_pipeline_is_committed", ctypes.c_int, _vp)
_sig("vxh_pipeline_node_count", _sz, _vp)
_sig("vxh_pipeline_relaunch", ctypes.c_int, _vp)
_sig("vxh_pipeline_destroy", None, _vp)
_sig("vxh_sgm_encode", ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _cfgp)
_sig("vxh_depth_filter_encode_bilateral", ctypes.c_int, _vp, _vp, _vp, _vp,
     ctypes.POINTER(vxh_depth_filter_config))
_sig("vxh_depth_filter_encode_median", ctypes.c_int, _vp, _vp, _vp, _vp,
     ctypes.POINTER(vxh_depth_median_config))
_sig("vxh_context_pool_stats", ctypes.c_int, _vp, ctypes.POINTER(_sz))
_sig("vxh_context_pool_trim", None, _vp)
_sig("vxh_texture_pool_create", ctypes.c_int, _sz, _texp)
_sig("vxh_texture_pool_destroy", None, _vp)
_sig("vxh_texture_pool_acquire", ctypes.c_int, _vp, _vp, _u32, _u32, _u32, _texp)
_sig("vxh_texture_pool_release", None, _vp, _vp)
_sig("vxh_texture_pool_clear", None, _vp)
_sig("vxh_texture_pool_cached_count", _sz, _vp)
_sig("vxh_texture_pool_total_acquires", ctypes.c_uint64, _vp)
_sig("vxh_texture_pool_total_hits", ctypes.c_uint64, _vp)
_sig("vxh_texture_pool_hit_rate", ctypes.c_double, _vp)


class TextureFormat:
    """TextureFormat (vx-vision/src/texture.rs:14-24)."""
    R8UNORM = 1
    R32FLOAT = 2
    RGBA8UNORM = 3


class Pipeline:
    """Pipeline (vx-vision/src/pipeline.rs:12-80): kernels encoded into one
    unit of GPU work, run on commit.  Encoding records into a HIP stream
    capture; commit launches the captured graph."""

    def __init__(self, handle):
        self._h = _vp(handle)

    @classmethod
    def begin(cls, ctx):
        h = _vp()
        check(lib.vxh_pipeline_begin(ctx._h, ctypes.byref(h)))
        return cls(h.value)

    def cmd_buf(self):
        """The capturing hipStream_t (None once committed)."""
        return lib.vxh_pipeline_stream(self._h)

    stream = property(cmd_buf)

    def retain(self, tex):
        """Moves `tex` into the pipeline until it completes (pipeline.rs:39-41)."""
        h = tex._release()
        st = lib.vxh_pipeline_retain(self._h, h)
        if st != 0:
            tex._h = _vp(h)
        check(st)

    def retain_all(self, textures):
        for t in textures:
            self.retain(t)

    def commit(self):
        check(lib.vxh_pipeline_commit(self._h))

    def wait(self):
        check(lib.vxh_pipeline_wait(self._h))

    def commit_and_wait(self):
        """Commits, waits, and returns the retained textures (pipeline.rs:49-54)."""
        check(lib.vxh_pipeline_commit_and_wait(self._h))
        return self.take_retained()

    def take_retained(self):
        n = lib.vxh_pipeline_take_retained(self._h, None, 0)
        arr = (_vp * max(n, 1))()
        got = lib.vxh_pipeline_take_retained(self._h, arr, n)
        return [Texture(arr[i]) for i in range(got)]

    def is_committed(self):
        return bool(lib.vxh_pipeline_is_committed(self._h))

    def node_count(self):
        return lib.vxh_pipeline_node_count(self._h)

    def relaunch(self):
        """Runs the committed graph again on the textures' current contents."""
        check(lib.vxh_pipeline_relaunch(self._h))

    def close(self):
        if self._h:
            lib.vxh_pipeline_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TexturePool:
    """TexturePool (vx-vision/src/pool.rs:17-147): textures cached per
    (width, height, format), at most max_per_bucket of each."""

    def __init__(self, max_per_bucket=8):
        h = _vp()
        check(lib.vxh_texture_pool_create(max_per_bucket, ctypes.byref(h)))
        self._h = h

    @classmethod
    def with_capacity(cls, max_per_bucket):
        return cls(max_per_bucket)

    def acquire(self, ctx, width, height, fmt):
        h = _vp()
        check(lib.vxh_texture_pool_acquire(self._h, ctx._h, width, height, fmt, ctypes.byref(h)))
        return Texture(h.value)

    def acquire_gray8(self, ctx, width, height):
        return self.acquire(ctx, width, height, TextureFormat.R8UNORM)

    def acquire_r32float(self, ctx, width, height):
        return self.acquire(ctx, width, height, TextureFormat.R32FLOAT)

    def acquire_rgba8(self, ctx, width, height):
        return self.acquire(ctx, width, height, TextureFormat.RGBA8UNORM)

    def release(self, tex):
        """Moves `tex` back into the pool (dropped if its bucket is full)."""
        lib.vxh_texture_pool_release(self._h, tex._release())

    def release_all(self, textures):
        for t in textures:
            self.release(t)

    def clear(self):
        lib.vxh_texture_pool_clear(self._h)

    def cached_count(self):
        return lib.vxh_texture_pool_cached_count(self._h)

    def total_acquires(self):
        return lib.vxh_texture_pool_total_acquires(self._h)

    def total_hits(self):
        return lib.vxh_texture_pool_total_hits(self._h)

    def hit_rate(self):
        return lib.vxh_texture_pool_hit_rate(self._h)

    def close(self):
        if self._h:
            lib.vxh_texture_pool_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- rectification (include/vxh/rectify.h) ----
class vxh_rectify_params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("fx", "fy", "cx", "cy", "k1", "k2", "p1", "p2",
                                                "k3")] + \
               [("R", ctypes.c_double * 9)] + \
               [(n, ctypes.c_double) for n in ("new_fx", "new_fy", "new_cx", "new_cy")]


_rpp = ctypes.POINTER(vxh_rectify_params)
_sig("vxh_undistorter_create", ctypes.c_int, _vp, _texp)
_sig("vxh_undistorter_destroy", None, _vp)
_sig("vxh_undistorter_apply", ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp)
_sig("vxh_undistorter_encode", ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp)
_sig("vxh_remap_enqueue", ctypes.c_int, _vp, _u32, _u32, _u32, _vp, _vp, _vp, _u32, _u32, _u32,
     _vp)
_sig("vxh_rectify_params_default", None, _rpp)
_sig("vxh_rectify_inverse", ctypes.c_int, _rpp, ctypes.POINTER(ctypes.c_float))
_sig("vxh_rectify_maps_enqueue", ctypes.c_int, _rpp, _u32, _u32, _vp, _vp, _vp)
_sig("vxh_rectify_maps", ctypes.c_int, _vp, _rpp, _vp, _vp)


@dataclass
class RectifyParams:
    """Camera model of the rectify maps: K and distortion of the input camera,
    the rectifying rotation R (row-major) and the output projection."""
    fx: float = 1.0
    fy: float = 1.0
    cx: float = 0.0
    cy: float = 0.0
    k1: float = 0.0
    k2: float = 0.0
    p1: float = 0.0
    p2: float = 0.0
    k3: float = 0.0
    R: tuple = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)
    new_fx: float = None
    new_fy: float = None
    new_cx: float = None
    new_cy: float = None

    def to_c(self):
        c = vxh_rectify_params()
        for n in ("fx", "fy", "cx", "cy", "k1", "k2", "p1", "p2", "k3"):
            setattr(c, n, getattr(self, n))
        c.R[:] = [float(v) for v in self.R]
        for n, base in (("new_fx", "fx"), ("new_fy", "fy"), ("new_cx", "cx"), ("new_cy", "cy")):
            v = getattr(self, n)
            setattr(c, n, getattr(self, base) if v is None else v)
        return c

    def as_dict(self):
        c = self.to_c()
        d = {n: getattr(c, n) for n in ("fx", "fy", "cx", "cy", "k1", "k2", "p1", "p2", "k3",
                                        "new_fx", "new_fy", "new_cx", "new_cy")}
        d["R"] = list(c.R)
        return d


def rectify_inverse(params):
    """inverse(Knew R) as the maps use it: 9 floats, row-major."""
    c = params.to_c()
    m = (ctypes.c_float * 9)()
    check(lib.vxh_rectify_inverse(ctypes.byref(c), m))
    return np.array(m, np.float32)


def rectify_maps(ctx, params, map_x, map_y):
    """Fills two R32Float textures with the rectify maps (synchronous)."""
    c = params.to_c()
    check(lib.vxh_rectify_maps(ctx._h, ctypes.byref(c), map_x._h, map_y._h))


def rectify_maps_enqueue(params, width, height, d_map_x, d_map_y, stream=None):
    c = params.to_c()
    check(lib.vxh_rectify_maps_enqueue(ctypes.byref(c), width, height, d_map_x, d_map_y, stream))


def remap_enqueue(d_in, in_format, in_w, in_h, d_map_x, d_map_y, d_out, out_format, out_w,
                  out_h, stream=None):
    """Stream-ordered remap on device buffers (formats 1 gray8, 2 R32Float)."""
    check(lib.vxh_remap_enqueue(d_in, in_format, in_w, in_h, d_map_x, d_map_y, d_out,
                                out_format, out_w, out_h, stream))


class Undistorter:
    """Undistorter (vx-vision/src/kernels/undistort.rs:14-137): remap through
    per-pixel coordinate maps, bilinear, clamp to edge."""

    def __init__(self, ctx):
        h = _vp()
        check(lib.vxh_undistorter_create(ctx._h, ctypes.byref(h)))
        self._h = h

    def apply(self, ctx, input, map_x, map_y, output):
        check(lib.vxh_undistorter_apply(self._h, ctx._h, input._h, map_x._h, map_y._h,
                                        output._h))

    def encode(self, pipe, input, map_x, map_y, output):
        check(lib.vxh_undistorter_encode(self._h, pipe._h, input._h, map_x._h, map_y._h,
                                         output._h))

    def close(self):
        if self._h:
            lib.vxh_undistorter_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- datasets (include/vxh/datasets.h) ----
class vxh_stereo_frame(ctypes.Structure):
    _fields_ = [("timestamp", ctypes.c_double), ("left_path", ctypes.c_char_p),
                ("right_path", ctypes.c_char_p), ("has_pose", ctypes.c_int32),
                ("rotation", ctypes.c_float * 9), ("translation", ctypes.c_float * 3)]


class vxh_stereo_calib(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("fx", "fy", "cx", "cy", "baseline")] + \
               [("valid", ctypes.c_int32)]


_sig("vxh_kitti_stereo_load", ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, _texp)
_sig("vxh_kitti_stereo_destroy", None, _vp)
_sig("vxh_kitti_stereo_len", _sz, _vp)
_sig("vxh_kitti_stereo_is_color", ctypes.c_int32, _vp)
_sig("vxh_kitti_stereo_frame", ctypes.c_int, _vp, _sz, ctypes.POINTER(vxh_stereo_frame))
_sig("vxh_kitti_stereo_calib", ctypes.c_int, _vp, ctypes.POINTER(vxh_stereo_calib))
_sig("vxh_kitti_stereo_read_pair", ctypes.c_int, _vp, _sz, _vp, _vp, _sz)
_sig("vxh_png_info", ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(_u32), ctypes.POINTER(_u32),
     ctypes.POINTER(_u32))
_sig("vxh_png_read", ctypes.c_int, ctypes.c_char_p, _vp, _sz)
_sig("vxh_png_read_luma8", ctypes.c_int, ctypes.c_char_p, _vp, _sz)


def _path(p):
    return os.fsencode(p)


def png_info(path):
    """(width, height, channels) of an 8-bit PNG."""
    w, h, c = _u32(), _u32(), _u32()
    check(lib.vxh_png_info(_path(path), ctypes.byref(w), ctypes.byref(h), ctypes.byref(c)))
    return w.value, h.value, c.value


def png_read(path):
    """Decoded samples as an (h, w) or (h, w, channels) uint8 array."""
    w, h, c = png_info(path)
    out = np.empty((h, w, c) if c > 1 else (h, w), np.uint8)
    check(lib.vxh_png_read(_path(path), out.ctypes.data, out.size))
    return out


def png_read_luma8(path):
    """DynamicImage::to_luma8 of the file, as the reference's examples load frames."""
    w, h, _ = png_info(path)
    out = np.empty((h, w), np.uint8)
    check(lib.vxh_png_read_luma8(_path(path), out.ctypes.data, out.size))
    return out


@dataclass
class StereoFrame:
    """DatasetFrame (vx-vision/src/datasets/mod.rs:12-24) for a camera pair."""
    timestamp: float
    left_path: str
    right_path: str
    pose: tuple = None     # (rotation 3x3 row-major, translation 3) or None


@dataclass
class StereoCalib:
    fx: float
    fy: float
    cx: float
    cy: float
    baseline: float


class KittiStereo:
    """KittiDataset (vx-vision/src/datasets/kitti.rs:11-107) for both cameras
    of a sequence: iterable over StereoFrame, len(), calib, read_pair."""

    def __init__(self, sequence_dir, poses_file=None):
        h = _vp()
        check(lib.vxh_kitti_stereo_load(_path(sequence_dir),
                                        None if poses_file is None else _path(poses_file),
                                        ctypes.byref(h)))
        self._h = h

    def __len__(self):
        return lib.vxh_kitti_stereo_len(self._h)

    @property
    def is_color(self):
        return bool(lib.vxh_kitti_stereo_is_color(self._h))

    def frame(self, i):
        f = vxh_stereo_frame()
        check(lib.vxh_kitti_stereo_frame(self._h, i, ctypes.byref(f)))
        pose = (tuple(f.rotation), tuple(f.translation)) if f.has_pose else None
        return StereoFrame(f.timestamp, os.fsdecode(f.left_path), os.fsdecode(f.right_path),
                           pose)

    def __iter__(self):
        return (self.frame(i) for i in range(len(self)))

    @property
    def calib(self):
        c = vxh_stereo_calib()
        check(lib.vxh_kitti_stereo_calib(self._h, ctypes.byref(c)))
        return StereoCalib(c.fx, c.fy, c.cx, c.cy, c.baseline) if c.valid else None

    def read_pair(self, i):
        """Frame i's images as (h, w) luma arrays, ready for SGM."""
        w, h, _ = png_info(self.frame(i).left_path)
        left = np.empty((h, w), np.uint8)
        right = np.empty((h, w), np.uint8)
        check(lib.vxh_kitti_stereo_read_pair(self._h, i, left.ctypes.data, right.ctypes.data,
                                             left.size))
        return left, right

    def close(self):
        if self._h:
            lib.vxh_kitti_stereo_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- sparse stereo (include/vxh/stereo_match.h) ----
CORNER_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("response", "<f4"),
                         ("pyramid_level", "<u4")])
ORB_DTYPE = np.dtype([("desc", "<u4", 8), ("angle", "<f4")])
MATCH_DTYPE = np.dtype([("left_idx", "<u4"), ("right_idx", "<u4"), ("disparity", "<f4"),
                        ("point_3d", "<f4", 3), ("pad", "<f4")])


class vxh_stereo_config(ctypes.Structure):
    _fields_ = [("max_epipolar", ctypes.c_float), ("min_disparity", ctypes.c_float),
                ("max_disparity", ctypes.c_float), ("max_hamming", ctypes.c_uint32),
                ("ratio_thresh", ctypes.c_float), ("fx", ctypes.c_float),
                ("fy", ctypes.c_float), ("cx", ctypes.c_float), ("cy", ctypes.c_float),
                ("baseline", ctypes.c_float)]


_sig("vxh_stereo_config_default", None, ctypes.POINTER(vxh_stereo_config))
_sig("vxh_stereo_matcher_create", ctypes.c_int, _vp, _texp)
_sig("vxh_stereo_matcher_destroy", None, _vp)
_sig("vxh_stereo_matcher_run", ctypes.c_int, _vp, _vp, _vp, _sz, _vp, _sz, _vp, _vp,
     ctypes.POINTER(vxh_stereo_config), _vp, ctypes.POINTER(_sz))
_sig("vxh_stereo_match_enqueue", ctypes.c_int, _vp, _vp, _vp, _vp, _u32, _u32,
     ctypes.POINTER(vxh_stereo_config), _vp, _vp, _vp)
_sig("vxh_stereo_hamming_matrix_enqueue", ctypes.c_int, _vp, _vp, _vp, _vp, _u32, _u32,
     ctypes.POINTER(vxh_stereo_config), _vp, _vp)


@dataclass
class StereoConfig:
    """StereoConfig (vx-vision/src/kernels/stereomatch.rs:19-69)."""
    max_epipolar: float = 2.0
    min_disparity: float = 1.0
    max_disparity: float = 120.0
    max_hamming: int = 50
    ratio_thresh: float = 0.8
    fx: float = 600.0
    fy: float = 600.0
    cx: float = 320.0
    cy: float = 240.0
    baseline: float = 0.12

    def to_c(self):
        return vxh_stereo_config(self.max_epipolar, self.min_disparity, self.max_disparity,
                                 self.max_hamming, self.ratio_thresh, self.fx, self.fy, self.cx,
                                 self.cy, self.baseline)

    def as_dict(self):
        return dict(self.__dict__)


class StereoMatcher:
    """StereoMatcher (vx-vision/src/kernels/stereomatch.rs:72-339)."""

    def __init__(self, ctx):
        h = _vp()
        check(lib.vxh_stereo_matcher_create(ctx._h, ctypes.byref(h)))
        self._h = h

    def run(self, ctx, left_kpts, right_kpts, left_descs, right_descs, config):
        """Keypoints as CORNER_DTYPE arrays, descriptors as ORB_DTYPE arrays;
        returns a MATCH_DTYPE array in ascending left index."""
        lk = np.ascontiguousarray(left_kpts, CORNER_DTYPE)
        rk = np.ascontiguousarray(right_kpts, CORNER_DTYPE)
        ld = np.ascontiguousarray(left_descs, ORB_DTYPE)
        rd = np.ascontiguousarray(right_descs, ORB_DTYPE)
        if len(lk) != len(ld) or len(rk) != len(rd):
            raise InvalidConfig("invalid config: keypoint and descriptor counts differ")
        out = np.zeros(max(len(lk), 1), MATCH_DTYPE)
        n = _sz()
        c = config.to_c()
        check(lib.vxh_stereo_matcher_run(self._h, ctx._h, lk.ctypes.data, len(lk),
                                         rk.ctypes.data, len(rk), ld.ctypes.data,
                                         rd.ctypes.data, ctypes.byref(c), out.ctypes.data,
                                         ctypes.byref(n)))
        return out[: n.value]

    def close(self):
        if self._h:
            lib.vxh_stereo_matcher_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stereo_match_enqueue(d_lk, d_rk, d_ld, d_rd, n_left, n_right, config, d_matches, d_count,
                         stream=None):
    c = config.to_c()
    check(lib.vxh_stereo_match_enqueue(d_lk, d_rk, d_ld, d_rd, n_left, n_right, ctypes.byref(c),
                                       d_matches, d_count, stream))


def stereo_hamming_matrix_enqueue(d_lk, d_rk, d_ld, d_rd, n_left, n_right, config, d_matrix,
                                  stream=None):
    c = config.to_c()
    check(lib.vxh_stereo_hamming_matrix_enqueue(d_lk, d_rk, d_ld, d_rd, n_left, n_right,
                                                ctypes.byref(c), d_matrix, stream))


MIX_SIZES = [(320, 240), (640, 480), (1280, 720), (1920, 1080)]
MIX_DISPARITIES = [64, 128, 256]


def mixture(n, seed=0x5347_0003, sizes=MIX_SIZES, disparities=MIX_DISPARITIES):
    """The batched / served workload mixture: n pairs, each a size and D drawn
    with a seeded generator, 8 paths.  Returns [(w, h, D, pair_seed)]."""
    import numpy as _np
    rng = _np.random.default_rng(seed)
    out = []
    for i in range(n):
        w, h = sizes[int(rng.integers(len(sizes)))]
        d = disparities[int(rng.integers(len(disparities)))]
        out.append((w, h, d, int(seed * 1000 + i)))
    return out


def mixture_pair(w, h, d, pair_seed):
    left, right, _ = synth_pair(w, h, d, 12 if w >= 1280 else 6, 2 if w >= 1280 else 0,
                                pair_seed)
    cfg = SGMStereoConfig(num_disparities=d, p1=10.0, p2=120.0, num_paths=8)
    return left, right, cfg


def workspace_bytes(width, height, config):
    c = config.to_c()
    return lib.vxh_sgm_workspace_bytes(width, height, ctypes.byref(c))


def synth_pair(width, height, num_disparities, n_rects, n_planes, seed):
    """Seeded integer-only synthetic pair (include/vxh/synth.h).

    Returns (left, right, truth) as (h, w) uint8, uint8 and int32 arrays."""
    left = np.empty((height, width), dtype=np.uint8)
    right = np.empty((height, width), dtype=np.uint8)
    truth = np.empty((height, width), dtype=np.int32)
    check(lib.vxh_synth_pair(width, height, num_disparities, n_rects, n_planes,
                             seed, left.ctypes.data, right.ctypes.data,
                             truth.ctypes.data))
    return left, right, truth


# Workloads used by the benchmark and the parity tests: (w, h, D, paths,
# rectangles, slanted planes, seed).
WORKLOADS = {
    "qvga_d64_4path": (320, 240, 64, 4, 6, 0, 0x53470000),
    "720p_d128_8path": (1280, 720, 128, 8, 12, 2, 0x53470001),
    "1080p_d256_8path": (1920, 1080, 256, 8, 12, 2, 0x53470002),
}


def workload_pair(name):
    w, h, d, paths, nr, npl, seed = WORKLOADS[name]
    left, right, truth = synth_pair(w, h, d, nr, npl, seed)
    cfg = SGMStereoConfig(num_disparities=d, p1=10.0, p2=120.0, num_paths=paths)
    return left, right, truth, cfg

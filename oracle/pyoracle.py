"""ctypes binding of the scalar CPU oracle (TEST INFRASTRUCTURE ONLY).

Only the test suite, the smoke check and the benchmark's CPU-baseline leg use
this module; it is the checker, never a code path of the product.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_sgm.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} missing: run `python oracle/build_oracle.py`")
lib = ctypes.CDLL(LIB_PATH)

_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)


_sig("oracle_census", None, _vp, _u32, _u32, _u32, _u32, _vp)
_sig("oracle_cost", _u32, _vp, _vp, _u32, _u32, _u32, _u32)
_sig("oracle_aggregate_dir", None, _vp, _vp, _u32, _u32, _u32, _u32, _u32,
     ctypes.c_int, ctypes.c_int, _vp)
_sig("oracle_wta", None, _vp, _u32, _u32, _u32, _vp)
_sig("oracle_wta_index", None, _vp, _u32, _u32, _u32, _vp)
_sig("oracle_lr_check", None, _vp, _u32, _u32, _u32, _u32, _vp)
_sig("oracle_direction", ctypes.c_int, _u32, ctypes.POINTER(ctypes.c_int),
     ctypes.POINTER(ctypes.c_int))
_sig("oracle_sgm", ctypes.c_int, _vp, _vp, _u32, _u32, _u32, ctypes.c_float,
     ctypes.c_float, _u32, _u32, _u32, _u32, _u32, _u32, _vp)
_sig("oracle_sgm_volume", ctypes.c_int, _vp, _vp, _u32, _u32, _u32, ctypes.c_float,
     ctypes.c_float, _u32, _u32, _u32, _u32, _vp)
_sig("oracle_scanline", None, _vp, _u32, _u32, _u32, _u32, _vp)


_sig("oracle_depth_bilateral", None, _vp, _u32, _u32, ctypes.c_int, ctypes.c_float,
     ctypes.c_float, _vp)
_sig("oracle_depth_median", None, _vp, _u32, _u32, ctypes.c_int, _vp)
_sig("oracle_disparity_to_depth", None, _vp, ctypes.c_size_t, ctypes.c_float, ctypes.c_float,
     _vp)

_sig("oracle_remap", None, _vp, _u32, _u32, _u32, _vp, _vp, _vp, _u32, _u32, _u32)
_sig("oracle_rectify_maps", ctypes.c_int, _vp, _u32, _u32, _vp, _vp, _vp)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def remap(img, map_x, map_y, out_gray8=None):
    """Undistorter remap (undistort.metal:4-20) of an (h_in, w_in) uint8 or
    float32 image through (h, w) float32 maps.  The output is gray8 when
    out_gray8 is True (default: same format as the input)."""
    img = np.ascontiguousarray(img)
    in_fmt = 1 if img.dtype == np.uint8 else 2
    if in_fmt == 2:
        img = _f32(img)
    mx, my = _f32(map_x), _f32(map_y)
    assert mx.shape == my.shape
    h, w = mx.shape
    gray_out = (in_fmt == 1) if out_gray8 is None else out_gray8
    out = np.empty((h, w), np.uint8 if gray_out else np.float32)
    lib.oracle_remap(img.ctypes.data, in_fmt, img.shape[1], img.shape[0], mx.ctypes.data,
                     my.ctypes.data, out.ctypes.data, 1 if gray_out else 2, w, h)
    return out


_sig("oracle_stereo_matrix", None, _vp, _vp, _vp, _vp, _u32, _u32, _vp, _vp)
_sig("oracle_stereo_match", _u32, _vp, _vp, _vp, _vp, _u32, _u32, _vp, _u32, _vp)

CORNER_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("response", "<f4"),
                         ("pyramid_level", "<u4")])
MATCH_DTYPE = np.dtype([("left_idx", "<u4"), ("right_idx", "<u4"), ("disparity", "<f4"),
                        ("point_3d", "<f4", 3), ("pad", "<f4")])


def _stereo_args(lk, rk, ld, rd, cfg):
    lk = np.ascontiguousarray(lk, CORNER_DTYPE)
    rk = np.ascontiguousarray(rk, CORNER_DTYPE)
    ld = np.ascontiguousarray(ld, np.uint32).reshape(-1, 8)
    rd = np.ascontiguousarray(rd, np.uint32).reshape(-1, 8)
    p = np.array([cfg["max_epipolar"], cfg["min_disparity"], cfg["max_disparity"],
                  cfg["ratio_thresh"], cfg["fx"], cfg["fy"], cfg["cx"], cfg["cy"],
                  cfg["baseline"]], np.float32)
    return lk, rk, ld, rd, p


def stereo_matrix(lk, rk, ld, rd, cfg):
    """StereoMatch.metal:32-62: the n_left x n_right u16 distance matrix."""
    lk, rk, ld, rd, p = _stereo_args(lk, rk, ld, rd, cfg)
    out = np.empty((len(lk), len(rk)), np.uint16)
    lib.oracle_stereo_matrix(lk.ctypes.data, rk.ctypes.data, ld.ctypes.data, rd.ctypes.data,
                             len(lk), len(rk), p.ctypes.data, out.ctypes.data)
    return out


def stereo_match(lk, rk, ld, rd, cfg):
    """StereoMatch.metal:68-115 in ascending left index: a MATCH_DTYPE array."""
    lk, rk, ld, rd, p = _stereo_args(lk, rk, ld, rd, cfg)
    out = np.zeros(max(len(lk), 1), MATCH_DTYPE)
    n = lib.oracle_stereo_match(lk.ctypes.data, rk.ctypes.data, ld.ctypes.data, rd.ctypes.data,
                                len(lk), len(rk), p.ctypes.data, int(cfg["max_hamming"]),
                                out.ctypes.data)
    return out[:n]


RECTIFY_FIELDS = ("fx", "fy", "cx", "cy", "k1", "k2", "p1", "p2", "k3")


def rectify_maps(params, w, h):
    """Rectify maps (rectify_oracle.cpp) for params with the keys of
    RECTIFY_FIELDS plus R (9 values) and new_fx, new_fy, new_cx, new_cy.
    Returns (map_x, map_y, m) or raises ValueError for a singular Knew R."""
    v = [float(params[k]) for k in RECTIFY_FIELDS] + [float(x) for x in params["R"]] + \
        [float(params[k]) for k in ("new_fx", "new_fy", "new_cx", "new_cy")]
    p = np.array(v, np.float64)
    mx = np.empty((h, w), np.float32)
    my = np.empty((h, w), np.float32)
    m = np.empty(9, np.float32)
    if lib.oracle_rectify_maps(p.ctypes.data, w, h, mx.ctypes.data, my.ctypes.data,
                               m.ctypes.data) != 0:
        raise ValueError("singular Knew * R")
    return mx, my, m


def depth_bilateral(depth, radius=3, sigma_spatial=5.0, sigma_depth=0.05):
    """DepthFilter::apply_bilateral (DepthFilter.metal:17-60) on an (h, w) map."""
    d = _f32(depth)
    h, w = d.shape
    out = np.empty_like(d)
    lib.oracle_depth_bilateral(d.ctypes.data, w, h, radius, sigma_spatial, sigma_depth,
                               out.ctypes.data)
    return out


def depth_median(depth, radius=1):
    """DepthFilter::apply_median (DepthFilter.metal:66-120) on an (h, w) map."""
    d = _f32(depth)
    h, w = d.shape
    out = np.empty_like(d)
    lib.oracle_depth_median(d.ctypes.data, w, h, radius, out.ctypes.data)
    return out


def disparity_to_depth(disp, fx, baseline):
    """(fx * baseline) / d for d > 0, else 0 (sgm.rs:284-286)."""
    d = _f32(disp)
    out = np.empty_like(d)
    lib.oracle_disparity_to_depth(d.ctypes.data, d.size, fx, baseline, out.ctypes.data)
    return out


def _u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)


def census(img, rx=4, ry=3):
    img = _u8(img)
    h, w = img.shape
    out = np.empty((h, w, 2), dtype=np.uint32)
    lib.oracle_census(img.ctypes.data, w, h, rx, ry, out.ctypes.data)
    return out


def direction(i):
    dx, dy = ctypes.c_int(), ctypes.c_int()
    if lib.oracle_direction(i, ctypes.byref(dx), ctypes.byref(dy)):
        raise IndexError(i)
    return dx.value, dy.value


def aggregate_dir(lc, rc, nd, p1, p2, dx, dy, S=None):
    h, w = lc.shape[:2]
    if S is None:
        S = np.zeros((h, w, nd), dtype=np.uint16)
    lib.oracle_aggregate_dir(np.ascontiguousarray(lc).ctypes.data,
                             np.ascontiguousarray(rc).ctypes.data, w, h, nd, p1, p2,
                             dx, dy, S.ctypes.data)
    return S


def wta(S):
    h, w, nd = S.shape
    out = np.empty((h, w), dtype=np.float32)
    lib.oracle_wta(np.ascontiguousarray(S).ctypes.data, w, h, nd, out.ctypes.data)
    return out


def wta_index(S):
    h, w, nd = S.shape
    out = np.empty((h, w), dtype=np.int32)
    lib.oracle_wta_index(np.ascontiguousarray(S).ctypes.data, w, h, nd, out.ctypes.data)
    return out


def lr_check(S, disp, tol=1):
    h, w, nd = S.shape
    disp = np.array(disp, dtype=np.float32, copy=True)
    lib.oracle_lr_check(np.ascontiguousarray(S).ctypes.data, w, h, nd, tol,
                        disp.ctypes.data)
    return disp


def sgm(left, right, nd=128, p1=10.0, p2=120.0, rx=4, ry=3, num_paths=2,
        lr_check=False, lr_tol=1, nthreads=1):
    left, right = _u8(left), _u8(right)
    h, w = left.shape
    out = np.empty((h, w), dtype=np.float32)
    st = lib.oracle_sgm(left.ctypes.data, right.ctypes.data, w, h, nd, p1, p2, rx, ry,
                        num_paths, int(bool(lr_check)), lr_tol, nthreads,
                        out.ctypes.data)
    if st:
        raise ValueError(f"oracle rejected the config (status {st})")
    return out


def sgm_volume(left, right, nd=128, p1=10.0, p2=120.0, rx=4, ry=3, num_paths=2,
               nthreads=1):
    left, right = _u8(left), _u8(right)
    h, w = left.shape
    S = np.empty((h, w, nd), dtype=np.uint16)
    st = lib.oracle_sgm_volume(left.ctypes.data, right.ctypes.data, w, h, nd, p1, p2,
                               rx, ry, num_paths, nthreads, S.ctypes.data)
    if st:
        raise ValueError(f"oracle rejected the config (status {st})")
    return S


def scanline(costs, p1=10, p2=120):
    costs = np.ascontiguousarray(costs, dtype=np.uint8)
    steps, nd = costs.shape
    out = np.empty((steps, nd), dtype=np.uint16)
    lib.oracle_scanline(costs.ctypes.data, steps, nd, p1, p2, out.ctypes.data)
    return out

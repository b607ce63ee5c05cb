"""Host-side checks of the C ABI that need no GPU: the library loads, exports
every function the public headers declare, and rejects bad configurations with
the reference's error variants (vx-vision/src/error.rs:6-23)."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import vxsgm

INCLUDE = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include")


def declared_functions():
    names = set()
    for path in glob.glob(os.path.join(INCLUDE, "**", "*.h"), recursive=True):
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(vxh_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_functions()
    assert "vxh_sgm_compute_disparity" in names and "vxh_synth_pair" in names
    assert len(names) >= 25


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports_declared_symbol(name):
    assert hasattr(vxsgm.lib, name), f"{name} declared in include/ but not exported"


def test_default_config_matches_reference():
    """SGMStereoConfig::default (sgm.rs:39-49)."""
    c = vxsgm.vxh_sgm_config()
    vxsgm.lib.vxh_sgm_config_default(ctypes.byref(c))
    assert (c.num_disparities, c.p1, c.p2, c.census_radius_x, c.census_radius_y) == \
        (128, 10.0, 120.0, 4, 3)
    assert (c.num_paths, c.lr_check, c.lr_tolerance) == (2, 0, 1)
    vxsgm.lib.vxh_sgm_config_new(32, ctypes.byref(c))
    assert c.num_disparities == 32 and c.p2 == 120.0
    py = vxsgm.SGMStereoConfig()
    assert (py.num_disparities, py.p1, py.p2, py.census_radius_x, py.census_radius_y,
            py.num_paths) == (128, 10.0, 120.0, 4, 3, 2)


@pytest.mark.parametrize("kw,w,h", [
    (dict(num_disparities=0), 64, 8),
    (dict(num_disparities=257), 300, 8),
    (dict(num_disparities=65), 64, 8),
    (dict(census_radius_x=5, census_radius_y=3), 64, 8),
    (dict(census_radius_x=33, census_radius_y=0), 64, 8),
    (dict(p1=-1.0), 64, 8),
    (dict(p2=65536.0), 64, 8),
    (dict(p1=float("nan")), 64, 8),
    (dict(num_paths=3), 64, 8),
    (dict(num_paths=16), 64, 8),
    (dict(), 0, 8),
])
def test_invalid_configs_rejected(kw, w, h):
    cfg = vxsgm.SGMStereoConfig(**{"num_disparities": 16, **kw})
    with pytest.raises(vxsgm.InvalidConfig):
        cfg.validate(w, h)


@pytest.mark.parametrize("kw", [dict(), dict(num_disparities=1), dict(num_disparities=256),
                                dict(census_radius_x=31, census_radius_y=0),
                                dict(census_radius_x=0, census_radius_y=0),
                                dict(p1=65535.5, p2=0.0), dict(num_paths=4),
                                dict(num_paths=8, lr_check=True)])
def test_valid_configs_accepted(kw):
    cfg = vxsgm.SGMStereoConfig(**{"num_disparities": 64, **kw})
    cfg.validate(256, 4)


def test_gray8_length_checked_before_device_use():
    """Texture::from_gray8 rejects a wrong byte count (texture.rs:44-53)."""
    out = ctypes.c_void_p()
    buf = np.zeros(10, np.uint8)
    st = vxsgm.lib.vxh_texture_gray8(None, buf.ctypes.data, 10, 3, 3, ctypes.byref(out))
    assert st == vxsgm.InvalidConfig.code
    assert b"expected 9 bytes" in vxsgm.lib.vxh_last_error()


def test_no_device_reports_device_not_found():
    if vxsgm.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(vxsgm.DeviceNotFound):
        vxsgm.Context(0)


def test_context_rejects_bad_ordinal():
    with pytest.raises(vxsgm.DeviceNotFound):
        vxsgm.Context(vxsgm.device_count() + 3)


def test_workspace_size_formula():
    cfg = vxsgm.SGMStereoConfig(num_disparities=128)
    n = 1280 * 720
    assert vxsgm.workspace_bytes(1280, 720, cfg) >= n * 8 * 2 + n * 128 * 2


def test_synth_pair_deterministic_and_sane():
    a = vxsgm.synth_pair(64, 48, 16, 3, 1, 7)
    b = vxsgm.synth_pair(64, 48, 16, 3, 1, 7)
    c = vxsgm.synth_pair(64, 48, 16, 3, 1, 8)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert not np.array_equal(a[0], c[0])
    assert a[2].min() >= 0 and a[2].max() <= 15


def align256(v):
    return (v + 255) // 256 * 256


@pytest.mark.parametrize("w,h,D,paths,p2,lr", [(640, 480, 100, 2, 120.0, False),
                                               (1920, 1080, 128, 8, 120.0, False),
                                               (320, 240, 64, 4, 400.0, False),
                                               (7, 5, 3, 8, 120.0, False),
                                               (1280, 720, 128, 8, 120.0, True),
                                               (33, 9, 40, 4, 300.0, True)])
def test_server_request_bytes_is_workspace(w, h, D, paths, p2, lr):
    """A served request is admitted against its whole pipeline workspace: census
    of both images, the u8 matching-cost volume (Dp bytes per pixel), the
    exception flags (8 bytes per pixel), the split launch's synchronisation
    words (error flag, a ticket counter per XCD, two flags per scanline
    group), and per direction a u8
    volume plus its high-byte volume, or a u16 volume; the larger of the two
    formats is reserved.  The LR check adds the integer winners (2 bytes per
    pixel): its right-view keys reuse the census area, and no summed volume
    is stored."""
    cfg = vxsgm.SGMStereoConfig(num_disparities=D, num_paths=paths, p2=p2, lr_check=lr)
    n, Dp = w * h, (D + 15) // 16 * 16
    sync = align256(4 * (10 + 16 * ((w + h) // 2 + 1)))
    common = 2 * align256(n * 8) + align256(n * Dp) + sync + (align256(n * 2) if lr else 0)
    u8 = common + align256(n * 8) + 2 * paths * align256(n * Dp)
    u16 = common + paths * align256(n * Dp * 2)
    assert vxsgm.workspace_bytes(w, h, cfg) == max(u8, u16)
    assert vxsgm.request_bytes(w, h, cfg) == max(u8, u16)


@pytest.mark.parametrize("kw", [dict(hbm_budget_bytes=0), dict(max_inflight=0),
                                dict(num_streams=0)])
def test_server_rejects_zero_limits(kw):
    with pytest.raises(vxsgm.InvalidConfig):
        vxsgm.SGMServer(**kw)


def test_server_without_device():
    if vxsgm.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(vxsgm.DeviceNotFound):
        vxsgm.SGMServer()


def test_batch_rejects_empty_and_bad_device():
    h = ctypes.c_void_p()
    st = vxsgm.lib.vxh_sgm_batch_create(0, None, 0, ctypes.byref(h))
    assert st == vxsgm.InvalidConfig.code

"""The KITTI stereo loader and PNG reader (include/vxh/datasets.h), on the
CPU with synthetic sequences (no dataset is available offline).

PNG decoding is pinned two ways: files written here by an encoder that
applies each of the specification's five row filters (so a wrong Sub/Up/
Average/Paeth predictor fails), and, when Pillow is importable, files Pillow
writes with its own adaptive filtering, decoded by Pillow as the reference.
The luma conversion restates the `image` crate 0.25's to_luma8, which the
reference's examples use (examples/feature_matching_demo.rs:43-48); the
crate is not vendored, so beyond its closed forms (gray stays gray, white
stays 255, weights sum to 10000) that conversion is parity unpinned.

The directory semantics follow vx-vision/src/datasets/kitti.rs:19-107: the
color camera is preferred, only *.png entries count, they are sorted by
name, timestamps are index / 10, and a pose line is any line with at least
12 parseable numbers."""
import os
import struct
import zlib

import numpy as np
import pytest

import vxsgm


def paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def encode_png(path, arr, filters=(0, 1, 2, 3, 4)):
    """8-bit PNG writer: row y uses filters[y % len(filters)]."""
    arr = np.asarray(arr, np.uint8)
    h, w = arr.shape[:2]
    c = 1 if arr.ndim == 2 else arr.shape[2]
    color = {1: 0, 2: 4, 3: 2, 4: 6}[c]
    rows = arr.reshape(h, w * c).astype(np.int64)
    out = bytearray()
    for y in range(h):
        f = filters[y % len(filters)]
        cur = rows[y]
        prev = rows[y - 1] if y else np.zeros_like(cur)
        enc = []
        for i in range(w * c):
            a = int(cur[i - c]) if i >= c else 0
            b = int(prev[i])
            cc = int(prev[i - c]) if i >= c else 0
            pred = [0, a, b, (a + b) // 2, paeth(a, b, cc)][f]
            enc.append((int(cur[i]) - pred) & 0xFF)
        out.append(f)
        out.extend(enc)

    def chunk(t, data):
        return (struct.pack(">I", len(data)) + t + data +
                struct.pack(">I", zlib.crc32(t + data) & 0xFFFFFFFF))

    ihdr = struct.pack(">IIBBBBB", w, h, 8, color, 0, 0, 0)
    with open(path, "wb") as fh:
        fh.write(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr))
        z = zlib.compress(bytes(out), 6)
        # split the stream over two IDAT chunks, as encoders may
        fh.write(chunk(b"IDAT", z[: len(z) // 2]) + chunk(b"IDAT", z[len(z) // 2:]))
        fh.write(chunk(b"tEXt", b"Comment\x00test") + chunk(b"IEND", b""))


@pytest.mark.parametrize("channels", [1, 2, 3, 4])
@pytest.mark.parametrize("filters", [(0,), (1,), (2,), (3,), (4,), (0, 1, 2, 3, 4)])
def test_png_filters_round_trip(tmp_path, channels, filters):
    rng = np.random.default_rng(channels * 10 + len(filters))
    shape = (9, 13) if channels == 1 else (9, 13, channels)
    arr = rng.integers(0, 256, shape, dtype=np.uint8)
    arr[3] = 255            # saturated row: predictor wraparound
    p = tmp_path / "x.png"
    encode_png(p, arr, filters)
    assert vxsgm.png_info(p) == (13, 9, channels)
    assert (vxsgm.png_read(p) == arr).all()


def test_png_against_pillow(tmp_path):
    Image = pytest.importorskip("PIL.Image")
    rng = np.random.default_rng(7)
    base = np.cumsum(rng.integers(0, 9, (47, 61, 4)), axis=1).astype(np.uint8)  # smooth: mixed filters
    for mode, arr in (("L", base[..., 0]), ("LA", base[..., :2]), ("RGB", base[..., :3]),
                      ("RGBA", base)):
        p = tmp_path / f"{mode}.png"
        Image.fromarray(arr, mode).save(p, optimize=True)
        want = np.asarray(Image.open(p))
        assert (vxsgm.png_read(p) == want).all(), mode


def luma_ref(rgb):
    rgb = rgb.astype(np.uint32)
    return ((2126 * rgb[..., 0] + 7152 * rgb[..., 1] + 722 * rgb[..., 2]) // 10000).astype(np.uint8)


def test_luma_conversion(tmp_path):
    gray = np.arange(256, dtype=np.uint8).reshape(16, 16)
    encode_png(tmp_path / "g.png", gray)
    assert (vxsgm.png_read_luma8(tmp_path / "g.png") == gray).all()
    encode_png(tmp_path / "ga.png", np.stack([gray, 255 - gray], -1))
    assert (vxsgm.png_read_luma8(tmp_path / "ga.png") == gray).all()   # alpha dropped
    rgb = np.random.default_rng(3).integers(0, 256, (11, 7, 3), dtype=np.uint8)
    rgb[0, :3] = [[255, 255, 255], [255, 0, 0], [0, 0, 255]]
    encode_png(tmp_path / "c.png", rgb)
    got = vxsgm.png_read_luma8(tmp_path / "c.png")
    assert list(got[0, :3]) == [255, 54, 18]   # 2126*255 // 10000, 722*255 // 10000
    assert (got == luma_ref(rgb)).all()


def test_png_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.png"
    p.write_bytes(b"not a png at all")
    with pytest.raises(vxsgm.InvalidConfig):
        vxsgm.png_read(p)
    encode_png(p, np.zeros((4, 4), np.uint8))
    raw = bytearray(p.read_bytes())
    raw[40] ^= 0xFF                 # inside IDAT: the CRC no longer matches
    p.write_bytes(bytes(raw))
    with pytest.raises(vxsgm.InvalidConfig, match="CRC"):
        vxsgm.png_read(p)
    with pytest.raises(vxsgm.InvalidConfig):
        vxsgm.png_read(tmp_path / "missing.png")
    # A header alone cannot make the reader allocate more than 1 GiB.
    ihdr = struct.pack(">IIBBBBB", 65535, 65535, 8, 6, 0, 0, 0)
    big = (b"\x89PNG\r\n\x1a\n" + struct.pack(">I", 13) + b"IHDR" + ihdr +
           struct.pack(">I", zlib.crc32(b"IHDR" + ihdr) & 0xFFFFFFFF))
    p.write_bytes(big)
    with pytest.raises(vxsgm.InvalidConfig, match="1 GiB"):
        vxsgm.png_read(p)


def make_sequence(root, n=4, color=False, size=(24, 10), calib=True, skip_right=None):
    left_dir = root / ("image_2" if color else "image_0")
    right_dir = root / ("image_3" if color else "image_1")
    left_dir.mkdir(parents=True)
    right_dir.mkdir(parents=True)
    rng = np.random.default_rng(5)
    w, h = size
    pairs = []
    for i in reversed(range(n)):      # written out of order: the loader sorts
        shape = (h, w, 3) if color else (h, w)
        lft = rng.integers(0, 256, shape, dtype=np.uint8)
        rgt = rng.integers(0, 256, shape, dtype=np.uint8)
        encode_png(left_dir / f"{i:06d}.png", lft)
        if i != skip_right:
            encode_png(right_dir / f"{i:06d}.png", rgt)
        pairs.append((lft, rgt))
    (left_dir / "notes.txt").write_text("ignored")
    (left_dir / ".png").write_bytes(b"")        # no extension: ignored
    (left_dir / "x.PNG").write_bytes(b"")       # other case: ignored
    if calib:
        P0 = "7.188560e+02 0 6.071928e+02 0 0 7.188560e+02 1.852157e+02 0 0 0 1 0"
        P1 = "7.188560e+02 0 6.071928e+02 -3.861448e+02 0 7.188560e+02 1.852157e+02 0 0 0 1 0"
        P2 = "7.188560e+02 0 6.071928e+02 4.538225e+01 0 7.188560e+02 1.852157e+02 -1.130887e-01 0 0 1 3.779761e-03"
        P3 = "7.188560e+02 0 6.071928e+02 -3.372877e+02 0 7.188560e+02 1.852157e+02 2.369057e+00 0 0 1 4.915215e-03"
        (root / "calib.txt").write_text(f"P0: {P0}\nP1: {P1}\nP2: {P2}\nP3: {P3}\nTr: 1 0 0 0 0 1 0 0 0 0 1 0\n")
    return list(reversed(pairs))


def test_gray_sequence(tmp_path):
    pairs = make_sequence(tmp_path, n=4)
    poses = tmp_path / "poses.txt"
    poses.write_text(
        "1 0 0 0.5 0 1 0 0.25 0 0 1 2\n"
        "short line 1 2 3\n"                                  # skipped: < 12 numbers
        "1 0 0 1.5 0 1 0 abc 0.25 0 0 1 3 9\r\n"              # 'abc' dropped, 13 numbers left
        "0 -1 0 7 1 0 0 8 0 0 1 9\n")
    ds = vxsgm.KittiStereo(tmp_path, poses)
    assert len(ds) == 4 and not ds.is_color
    frames = list(ds)
    assert [f.timestamp for f in frames] == [0.0, 0.1, 0.2, 0.3]
    assert [os.path.basename(f.left_path) for f in frames] == [f"{i:06d}.png" for i in range(4)]
    assert all("image_1" in f.right_path for f in frames)
    assert frames[0].pose == ((1, 0, 0, 0, 1, 0, 0, 0, 1), (0.5, 0.25, 2))
    # 'abc' dropped: the values after it shift left, so t = (v3, v7, v11) = (1.5, 0.25, 3)
    assert frames[1].pose == ((1, 0, 0, 0, 1, 0, 0, 0, 1), (1.5, 0.25, 3))
    assert frames[2].pose[0] == (0, -1, 0, 1, 0, 0, 0, 0, 1)
    assert frames[3].pose is None
    c = ds.calib
    assert c.fx == 718.856 and c.cx == 607.1928 and c.cy == 185.2157
    assert abs(c.baseline - 386.1448 / 718.856) < 1e-12
    for i, (lft, rgt) in enumerate(pairs):
        got_l, got_r = ds.read_pair(i)
        assert (got_l == lft).all() and (got_r == rgt).all()
    with pytest.raises(vxsgm.InvalidConfig):
        ds.frame(4)
    ds.close()


def test_color_sequence_preferred(tmp_path):
    make_sequence(tmp_path, n=2)                      # gray cameras present too
    pairs = make_sequence(tmp_path / "c", n=3, color=True)
    for name in ("image_2", "image_3"):
        os.rename(tmp_path / "c" / name, tmp_path / name)
    ds = vxsgm.KittiStereo(tmp_path)
    assert ds.is_color and len(ds) == 3
    assert all(f.pose is None for f in ds)
    left, right = ds.read_pair(1)
    assert (left == luma_ref(pairs[1][0])).all() and (right == luma_ref(pairs[1][1])).all()
    c = ds.calib
    assert abs(c.baseline - (45.38225 + 337.2877) / 718.856) < 1e-12   # P2 / P3


def test_sequence_errors(tmp_path):
    with pytest.raises(vxsgm.InvalidConfig, match="cannot read"):
        vxsgm.KittiStereo(tmp_path / "nothing")
    make_sequence(tmp_path / "s", n=3, skip_right=1)
    with pytest.raises(vxsgm.InvalidConfig, match="missing right image"):
        vxsgm.KittiStereo(tmp_path / "s")
    make_sequence(tmp_path / "t", n=2, calib=False)
    ds = vxsgm.KittiStereo(tmp_path / "t", tmp_path / "no_such_poses.txt")
    assert len(ds) == 2 and ds.calib is None and ds.frame(0).pose is None

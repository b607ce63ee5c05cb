"""Input generators shared by the tests: the reference's own test and benchmark
images, restated from their Rust sources, and small seeded random pairs."""
import numpy as np


def checkerboard_pair(w=64, h=64, shift=5):
    """vx-vision/tests/test_kernels.rs:786-816: 4-px checkerboard (200/50),
    right image = left shifted by `shift` with column 0 repeated."""
    x = np.arange(w)[None, :].repeat(h, 0)
    y = np.arange(h)[:, None].repeat(w, 1)
    left = np.where(((x // 4) + (y // 4)) % 2 == 0, 200, 50).astype(np.uint8)
    sx = np.where(x >= shift, x - shift, 0)
    right = np.where(((sx // 4) + (y // 4)) % 2 == 0, 200, 50).astype(np.uint8)
    return left, right


def bench_image(w, h):
    """vx-vision/benches/kernel_benchmarks.rs:14-24: gradient plus an xor
    pattern, computed in f32 and truncated to u8, wrapping add."""
    x = np.arange(w, dtype=np.uint32)[None, :].repeat(h, 0)
    y = np.arange(h, dtype=np.uint32)[:, None].repeat(w, 1)
    grad = ((x.astype(np.float32) / np.float32(w)) * np.float32(255.0)).astype(np.uint8)
    noise = (((x * 17) ^ (y * 31)) % 64).astype(np.uint8)
    return (grad.astype(np.uint16) + noise).astype(np.uint8)


def bench_pair(w, h, shift=8):
    """kernel_benchmarks.rs:199-208: right[i] = left[i - 8] when x >= 8."""
    left = bench_image(w, h)
    right = left.copy()
    right[:, shift:] = left[:, :-shift]
    return left, right


def random_pair(w, h, seed, smooth=True):
    rng = np.random.default_rng(seed)
    left = rng.integers(0, 256, (h, w), dtype=np.uint8)
    if smooth:
        p = np.pad(left.astype(np.int32), 1, mode="edge")
        left = ((sum(p[dy:dy + h, dx:dx + w] for dy in range(3) for dx in range(3)) + 4)
                // 9).astype(np.uint8)
    shift = int(rng.integers(0, max(1, w // 4)))
    right = np.roll(left, -shift, axis=1)
    noise = rng.integers(-3, 4, (h, w))
    right = np.clip(right.astype(np.int32) + noise, 0, 255).astype(np.uint8)
    return left, right

import sys
from pathlib import Path

import numpy as np
import pytest

# torch ships its own HIP runtime with the same soname as the system one;
# loading it before libvp8g.so makes both share one runtime in-process.
import torch  # noqa: F401

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT / "oxideav-vp8_amd", ROOT / "oracle", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle_ctypes import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def gpu():
    import vp8g
    ctx = vp8g.Context(0)
    yield ctx
    ctx.close()


class Lcg:
    """The reference tests' LCG (src/arnr.rs:685-694)."""

    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self) -> int:
        self.s = (self.s * 6364136223846793005 + 1) & 0xFFFFFFFFFFFFFFFF
        return self.s >> 33

    def noise(self, amp: int) -> int:
        return self.next() % (2 * amp + 1) - amp


W = H = 48


def clean_scene():
    """src/arnr.rs:697-708."""
    r, c = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    y = (64 + ((r * 3 + c * 2) & 0x7F)).astype(np.uint8).reshape(-1)
    u = np.full((W // 2) * (H // 2), 120, np.uint8)
    v = np.full((W // 2) * (H // 2), 136, np.uint8)
    return y, u, v


def noisy(plane, seed: int, amp: int) -> np.ndarray:
    """src/arnr.rs:711-717."""
    rng = Lcg(seed)
    out = np.empty(len(plane), np.uint8)
    for i, p in enumerate(plane):
        out[i] = min(255, max(0, int(p) + rng.noise(amp)))
    return out


def packed(w, h, y, u, v):
    return (w, h, y, u, v, w, (w + 1) // 2)


def assert_same(got, want, what="", context=None):
    """Byte-exact comparison with a short report (first differing rows)."""
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, f"{what}: shape {got.shape} != {want.shape}"
    if np.array_equal(got, want):
        return
    bad = np.argwhere(got != want)
    lines = [f"{what}: {len(bad)} of {got.size} elements differ"]
    for idx in bad[:6]:
        idx = tuple(int(i) for i in idx)
        extra = f" ctx={context(idx)}" if context else ""
        lines.append(f"  at {idx}: got {got[idx]} want {want[idx]}{extra}")
    raise AssertionError("\n".join(lines))


def mse(a, b) -> float:
    d = a.astype(np.int64) - b.astype(np.int64)
    return float((d * d).sum()) / len(a)


def translated(cy, dx):
    out = np.empty_like(cy)
    for r in range(H):
        for c in range(W):
            out[r * W + c] = cy[r * W + min(max(c - dx, 0), W - 1)]
    return out

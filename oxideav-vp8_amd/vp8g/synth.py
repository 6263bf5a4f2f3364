"""Deterministic synthetic I420 workloads for the ARNR / MC / SAD path.

All randomness comes from splitmix64 (Vigna), counter-based so it
vectorises: value k of stream `seed` is mix(seed + (k + 1) * GOLDEN).
Gaussians use Box-Muller in float64.  The bytes depend only on the seed,
so the CPU oracle and the GPU receive identical inputs.

Workloads
---------
* ``arnr_group_cif``: 352x288, 5 frames, centre 2.  Frame 0 is uniform
  noise smoothed by a 3x3 box blur ((sum + 4) // 9, edge replicated);
  frame t is frame t-1 shifted by an integer (dx, dy) in [-4, 4]^2 with
  edge replication (chroma by (dx/2, dy/2) truncated); every frame then
  gets Gaussian noise sigma 4, rounded and clipped.
* ``arnr_group_regions``: the 1080p form: an 8x8 grid of regions, each
  moving with its own integer velocity in [-vmax, vmax]^2 px/frame
  relative to frame 0, plus `occluders` rectangles per frame (sides
  16..96) replaced by fresh blurred noise, plus sigma-4 noise.
* ``batch_clips``: the uneven batch: `count` clips, resolution uniform
  over BATCH_RESOLUTIONS, group length uniform over {3, 5, 7, 9, 11},
  centre n // 2, content as ``arnr_group_regions``.
* ``mc_workload``: per-MB quarter-pel MVs uniform in [-64, 64], MB-aligned
  edge-padded reference planes, source MBs from the centre frame.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int) -> np.ndarray:
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        return z ^ (z >> np.uint64(31))


def uniform_int(seed: int, n: int, lo: int, hi: int) -> np.ndarray:
    """n integers uniform in [lo, hi] (modulo reduction of 64-bit draws)."""
    r = splitmix64(seed, n) % np.uint64(hi - lo + 1)
    return r.astype(np.int64) + lo


def gaussian(seed: int, n: int, sigma: float) -> np.ndarray:
    r = splitmix64(seed, 2 * n)
    u1 = ((r[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) / 9007199254740993.0
    u2 = (r[1::2] >> np.uint64(11)).astype(np.float64) / 9007199254740992.0
    return sigma * np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def box_blur3(p: np.ndarray) -> np.ndarray:
    """3x3 box blur with edge replication, (sum + 4) // 9."""
    q = np.pad(p.astype(np.int32), 1, mode="edge")
    s = sum(q[dy:dy + p.shape[0], dx:dx + p.shape[1]] for dy in range(3) for dx in range(3))
    return ((s + 4) // 9).astype(np.uint8)


def blurred_noise(seed: int, h: int, w: int) -> np.ndarray:
    return box_blur3((splitmix64(seed, h * w) >> np.uint64(56)).astype(np.uint8).reshape(h, w))


def shift_plane(p: np.ndarray, dx: int, dy: int) -> np.ndarray:
    """out(x, y) = p(clamp(x - dx), clamp(y - dy))."""
    h, w = p.shape
    ys = np.clip(np.arange(h) - dy, 0, h - 1)
    xs = np.clip(np.arange(w) - dx, 0, w - 1)
    return p[ys][:, xs]


def add_noise(p: np.ndarray, seed: int, sigma: float = 4.0) -> np.ndarray:
    g = gaussian(seed, p.size, sigma).reshape(p.shape)
    return np.clip(np.rint(p.astype(np.float64) + g), 0, 255).astype(np.uint8)


@dataclass
class Group:
    width: int
    height: int
    center: int
    y: list   # per frame (h, w) uint8
    u: list   # per frame (ch, cw)
    v: list
    seed: int
    clean: list | None = None  # per frame (y, u, v) before the noise stage

    @property
    def n(self) -> int:
        return len(self.y)

    def frames(self):
        """I420Frame-compatible tuples (width, height, y, u, v, y_stride, uv_stride)."""
        cw = (self.width + 1) // 2
        return [(self.width, self.height, self.y[i].reshape(-1), self.u[i].reshape(-1),
                 self.v[i].reshape(-1), self.width, cw) for i in range(self.n)]


def _seed(config: int, clip: int, plane: int) -> int:
    return 0x5EED0000 + 1000 * config + 10 * clip + plane


def arnr_group_cif(width: int = 352, height: int = 288, n: int = 5, clip: int = 0,
                   config: int = 0) -> Group:
    cw, ch = (width + 1) // 2, (height + 1) // 2
    base = _seed(config, clip, 0)
    cy = blurred_noise(base + 1, height, width)
    cu = blurred_noise(base + 2, ch, cw)
    cv = blurred_noise(base + 3, ch, cw)
    shifts = uniform_int(base + 4, 2 * n, -4, 4).reshape(n, 2)
    ys, us, vs = [], [], []
    for t in range(n):
        if t > 0:
            dx, dy = int(shifts[t, 0]), int(shifts[t, 1])
            cy = shift_plane(cy, dx, dy)
            cu = shift_plane(cu, int(dx / 2), int(dy / 2))
            cv = shift_plane(cv, int(dx / 2), int(dy / 2))
        ys.append(add_noise(cy, base + 100 + 3 * t))
        us.append(add_noise(cu, base + 101 + 3 * t))
        vs.append(add_noise(cv, base + 102 + 3 * t))
    return Group(width, height, n // 2, ys, us, vs, base)


def _region_warp(p: np.ndarray, t: int, vel: np.ndarray, grid: int, scale_div: int) -> np.ndarray:
    """Each of grid x grid regions reads frame 0 displaced by t * velocity."""
    h, w = p.shape
    ry = np.minimum(np.arange(h) * grid // h, grid - 1)
    rx = np.minimum(np.arange(w) * grid // w, grid - 1)
    reg = ry[:, None] * grid + rx[None, :]
    vx = np.trunc(vel[:, 0] / scale_div).astype(np.int64)[reg]
    vy = np.trunc(vel[:, 1] / scale_div).astype(np.int64)[reg]
    X = np.clip(np.arange(w)[None, :] - t * vx, 0, w - 1)
    Y = np.clip(np.arange(h)[:, None] - t * vy, 0, h - 1)
    return p[Y, X]


def _occlude(p: np.ndarray, seed: int, count: int, scale_div: int) -> np.ndarray:
    h, w = p.shape
    p = p.copy()
    r = uniform_int(seed, 4 * count, 0, 1 << 30).reshape(count, 4)
    for k in range(count):
        rw = max(1, (16 + int(r[k, 0]) % 81) // scale_div)
        rh = max(1, (16 + int(r[k, 1]) % 81) // scale_div)
        rw, rh = min(rw, w), min(rh, h)
        x0 = int(r[k, 2]) % (w - rw + 1)
        y0 = int(r[k, 3]) % (h - rh + 1)
        p[y0:y0 + rh, x0:x0 + rw] = blurred_noise(seed * 131 + k, rh, rw)
    return p


def arnr_group_regions(width: int = 1920, height: int = 1080, n: int = 7, vmax: int = 12,
                       occluders: int = 20, clip: int = 0, config: int = 1) -> Group:
    cw, ch = (width + 1) // 2, (height + 1) // 2
    base = _seed(config, clip, 0)
    y0 = blurred_noise(base + 1, height, width)
    u0 = blurred_noise(base + 2, ch, cw)
    v0 = blurred_noise(base + 3, ch, cw)
    vel = uniform_int(base + 4, 128, -vmax, vmax).reshape(64, 2)
    ys, us, vs, clean = [], [], [], []
    for t in range(n):
        fy = _region_warp(y0, t, vel, 8, 1)
        fu = _region_warp(u0, t, vel, 8, 2)
        fv = _region_warp(v0, t, vel, 8, 2)
        if occluders:
            fy = _occlude(fy, base + 500 + 7 * t, occluders, 1)
            fu = _occlude(fu, base + 501 + 7 * t, occluders, 2)
            fv = _occlude(fv, base + 502 + 7 * t, occluders, 2)
        clean.append((fy, fu, fv))
        ys.append(add_noise(fy, base + 100 + 3 * t))
        us.append(add_noise(fu, base + 101 + 3 * t))
        vs.append(add_noise(fv, base + 102 + 3 * t))
    return Group(width, height, n // 2, ys, us, vs, base, clean)


BATCH_RESOLUTIONS = ((176, 144), (352, 288), (640, 360), (1280, 720), (1920, 1080))
BATCH_LENGTHS = (3, 5, 7, 9, 11)


def batch_specs(count: int = 64, config: int = 2):
    """(width, height, n) of each clip of the uneven batch."""
    r = uniform_int(_seed(config, 0, 9), 2 * count, 0, 4).reshape(count, 2)
    return [(*BATCH_RESOLUTIONS[int(a)], BATCH_LENGTHS[int(b)]) for a, b in r]


def batch_clips(count: int = 64, config: int = 2, specs=None, workers: int = 8):
    """The uneven batch as a list of Groups (clip i seeded by (config, i)).
    Clips are generated on a thread pool; the bytes do not depend on it."""
    from concurrent.futures import ThreadPoolExecutor
    specs = specs if specs is not None else batch_specs(count, config)

    def one(i):
        w, h, n = specs[i]
        return arnr_group_regions(w, h, n, vmax=12, occluders=20, clip=i, config=config)

    with ThreadPoolExecutor(max(1, workers)) as ex:
        return list(ex.map(one, range(len(specs))))


def pad_to_mb(p: np.ndarray, mb: int) -> np.ndarray:
    """Edge-pad a plane on the right/bottom to a multiple of `mb`."""
    h, w = p.shape
    return np.pad(p, ((0, (-h) % mb), (0, (-w) % mb)), mode="edge")


@dataclass
class McWorkload:
    ref_y: np.ndarray   # (mb_rows*16, mb_cols*16)
    ref_u: np.ndarray   # (mb_rows*8, mb_cols*8)
    ref_v: np.ndarray
    mb_cols: int
    mb_rows: int
    mb_xy: np.ndarray   # (n, 2) uint32 (col, row)
    mvs: np.ndarray     # (n, 2) int16 (row, col)
    src: np.ndarray     # (n, 256) uint8 source MBs


def mc_workload(group: Group, clip: int = 0, config: int = 1, mv_range: int = 64) -> McWorkload:
    """Reference = frame centre+1 padded to whole MBs; one quarter-pel MV per
    MB uniform in [-mv_range, mv_range]; sources = the centre frame's MBs."""
    ref_i = min(group.center + 1, group.n - 1)
    ry = pad_to_mb(group.y[ref_i], 16)
    ru = pad_to_mb(group.u[ref_i], 8)
    rv = pad_to_mb(group.v[ref_i], 8)
    mb_rows, mb_cols = ry.shape[0] // 16, ry.shape[1] // 16
    ru = np.pad(ru, ((0, mb_rows * 8 - ru.shape[0]), (0, mb_cols * 8 - ru.shape[1])), mode="edge")
    rv = np.pad(rv, ((0, mb_rows * 8 - rv.shape[0]), (0, mb_cols * 8 - rv.shape[1])), mode="edge")
    n = mb_rows * mb_cols
    rr, cc = np.meshgrid(np.arange(mb_rows), np.arange(mb_cols), indexing="ij")
    mb_xy = np.stack([cc.reshape(-1), rr.reshape(-1)], 1).astype(np.uint32)
    mvs = uniform_int(_seed(config, clip, 7), 2 * n, -mv_range, mv_range).reshape(n, 2).astype(np.int16)
    cy = pad_to_mb(group.y[group.center], 16)
    src = cy.reshape(mb_rows, 16, mb_cols, 16).transpose(0, 2, 1, 3).reshape(n, 256)
    return McWorkload(np.ascontiguousarray(ry), np.ascontiguousarray(ru), np.ascontiguousarray(rv),
                      mb_cols, mb_rows, mb_xy, mvs, np.ascontiguousarray(src))

"""CPU restatement of the reference's host-side input preparation (TEST
INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
ribo_rasterize_fumaroles, ribo_vent_heat_src, ribo_seed_cells,
ribo_fit_matter_to_terrain and ribo_refit_matter_to_terrain).

Written from the reference's Rust, function by function, in binary32 numpy
arithmetic in the same operation order and with the same SplitMix64:
- rasterize_fumaroles        src/world/terrain.rs:316-336
- upload_heat (grid -> fluid) src/world/mod.rs:1366-1392
- seed_cells, fbm_field, bilinear  src/world/mod.rs:1990-2062
- fit_matter_to_terrain       src/world/mod.rs:1970-1983
- refit_matter_to_terrain     src/world/mod.rs:1870-1968
"""
from __future__ import annotations

import numpy as np

M64 = (1 << 64) - 1
F32 = np.float32
CHEM_CELL_CAP = 48
GAMMA_SOLID = 3
PIXEL = F32(5000.0)


class SplitMix:
    """splitmix64 (src/world/mod.rs:1842-1855)."""

    def __init__(self, s: int):
        self.s = s & M64

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & M64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        return z ^ (z >> 31)

    def f32(self) -> np.float32:
        return F32(self.next() >> 40) / F32(1 << 24)


def rasterize_fumaroles(grid: int, wupc: int, fumaroles) -> np.ndarray:
    """fumaroles: (x_frac, y_frac, strength, spread, enabled) tuples."""
    n = grid
    heat = np.zeros(n * n, F32)
    cell_world = F32(wupc)
    for (xf, yf, st, sp, en) in fumaroles:
        xf, yf, st, sp = F32(xf), F32(yf), F32(st), F32(sp)
        if en == 0 or not st > 0:
            continue
        cx, cy = xf * F32(n), yf * F32(n)
        r = max(sp / cell_world, F32(1.0))
        x0, x1 = int(max(np.floor(cx - r), F32(0))), min(int(np.ceil(cx + r)), n)
        y0, y1 = int(max(np.floor(cy - r), F32(0))), min(int(np.ceil(cy + r)), n)
        if x1 <= x0 or y1 <= y0:
            continue
        ys, xs = np.mgrid[y0:y1, x0:x1]
        dx = xs.astype(F32) + F32(0.5) - cx
        dy = ys.astype(F32) + F32(0.5) - cy
        d = np.sqrt(dx * dx + dy * dy)
        inside = d < r
        w0 = F32(1.0) - d / r
        add = st * w0 * w0 / PIXEL
        blk = heat.reshape(n, n)[y0:y1, x0:x1]
        blk[inside] = blk[inside] + add[inside]
    return heat


def vent_heat_src(grid: int, fluid: int, heat: np.ndarray, chem: np.ndarray, gain: float) -> np.ndarray:
    n, f = grid, fluid
    k = max(n // f, 1)
    scale = PIXEL * max(F32(gain), F32(0.0)) / F32(k * k)
    out = np.zeros((f, f, 2), F32)
    h, c = heat.reshape(n, n), chem.reshape(n, n)
    lim = min(n, f * k)
    # the reference adds row by row, then along the row: per fluid cell the
    # contributions arrive in (dy, dx) order
    for dy in range(k):
        for dx in range(k):
            out[..., 0] += h[dy:lim:k, dx:lim:k] * scale
            out[..., 1] += c[dy:lim:k, dx:lim:k] * scale
    return out.reshape(-1)


def fbm_field(m: int, seed: int) -> np.ndarray:
    def lattice(o, ix, iy):
        h = SplitMix(seed ^ ((o * 0x9E3779B97F4A7C15) & M64) ^ ((ix * 0xD6E8FEB86659FD93) & M64)
                     ^ ((iy * 0xCA5A826395121157) & M64))
        return h.f32()

    out = np.zeros(m * m, F32)
    amp, norm, freq = F32(0.5), F32(0.0), F32(4.0)
    ys, xs = np.divmod(np.arange(m * m), m)
    for o in range(6):
        px = xs.astype(F32) / F32(m) * freq
        py = ys.astype(F32) / F32(m) * freq
        fx, fy = np.floor(px), np.floor(py)
        tx, ty = px - fx, py - fy
        sx = tx * tx * (F32(3.0) - F32(2.0) * tx)
        sy = ty * ty * (F32(3.0) - F32(2.0) * ty)
        L = {}

        def lat(a, b):
            a, b = int(a), int(b)
            if (a, b) not in L:
                L[(a, b)] = lattice(o, a, b)
            return L[(a, b)]
        ix, iy = fx.astype(np.int64), fy.astype(np.int64)
        l00 = np.array([lat(a, b) for a, b in zip(ix, iy)], F32)
        l10 = np.array([lat(a + 1, b) for a, b in zip(ix, iy)], F32)
        l01 = np.array([lat(a, b + 1) for a, b in zip(ix, iy)], F32)
        l11 = np.array([lat(a + 1, b + 1) for a, b in zip(ix, iy)], F32)
        a = l00 + (l10 - l00) * sx
        b = l01 + (l11 - l01) * sx
        out = out + amp * (a + (b - a) * sy)
        norm = norm + amp
        amp = amp * F32(0.65)
        freq = freq * F32(2.0)
    return out / norm


def bilinear(f: np.ndarray, m: int, x: np.float32, y: np.float32) -> np.float32:
    x = min(max(x, F32(0.0)), F32(m - 1))
    y = min(max(y, F32(0.0)), F32(m - 1))
    x0, y0 = int(np.floor(x)), int(np.floor(y))
    x1, y1 = min(x0 + 1, m - 1), min(y0 + 1, m - 1)
    tx, ty = x - F32(x0), y - F32(y0)
    a = f[y0 * m + x0] + (f[y0 * m + x1] - f[y0 * m + x0]) * tx
    b = f[y1 * m + x0] + (f[y1 * m + x1] - f[y1 * m + x0]) * tx
    return a + (b - a) * ty


def seed_cells(grid: int, seed: int, density: float, active: float) -> np.ndarray:
    n = grid
    rng = SplitMix(seed)
    max_per_channel = CHEM_CELL_CAP // 4
    m = max(n // 4, 2)
    fields = [fbm_field(m, seed ^ (((ch + 1) * 0xA24BAED4963EE407) & M64)) for ch in range(4)]
    dens = min(max(F32(density), F32(0.0)), F32(1.0))
    act_p = min(max(F32(active), F32(0.0)), F32(1.0))
    cells = np.zeros(n * n * 4, np.uint32)
    for y in range(n):
        fy = (F32(y) + F32(0.5)) / F32(n) * F32(m) - F32(0.5)
        for x in range(n):
            fx = (F32(x) + F32(0.5)) / F32(n) * F32(m) - F32(0.5)
            for ch in range(4):
                d = min(max(F32(0.5) + F32(1.6) * (bilinear(fields[ch], m, fx, fy) - F32(0.5)), F32(0.0)), F32(1.0))
                expect = d * (F32(max_per_channel) - F32(2.0)) * dens
                count = min(int(expect + rng.f32()), max_per_channel)
                act = sum(1 for _ in range(count) if rng.f32() < act_p)
                cells[(y * n + x) * 4 + ch] = act | ((count - act) << 16)
    return cells


def capacity(g: int) -> int:
    return 0 if g >= GAMMA_SOLID else CHEM_CELL_CAP * (GAMMA_SOLID - g) // GAMMA_SOLID


def fit_matter_to_terrain(cells: np.ndarray, gamma: np.ndarray, seed: int) -> np.ndarray:
    cells = cells.copy()
    rng = SplitMix(seed ^ 0x0E170B0DE5EED5ED)
    for i, g in enumerate(gamma.tolist()):
        if g == 0:
            continue
        per_ch = 0 if g >= GAMMA_SOLID else CHEM_CELL_CAP * (GAMMA_SOLID - g) // GAMMA_SOLID // 4
        for c in range(4):
            act = sum(1 for _ in range(per_ch) if rng.f32() < F32(0.5))
            cells[i * 4 + c] = act | ((per_ch - act) << 16)
    return cells


def refit_matter_to_terrain(cells: np.ndarray, gamma: np.ndarray) -> np.ndarray:
    """Integer restatement; cells u32[4n] (not modified), returns the result."""
    n = len(gamma)
    kinds = np.zeros((n, 8), np.int64)  # kind k = channel * 2 + state
    c = cells.reshape(n, 4).astype(np.int64)
    kinds[:, 0::2] = c & 0xFFFF
    kinds[:, 1::2] = c >> 16
    g = gamma.tolist()
    caps = [capacity(x) for x in g]
    carry = [0] * 8
    room_total = 0
    for i in range(n):
        cap, tot = caps[i], int(kinds[i].sum())
        if tot <= cap:
            room_total += cap - tot
            continue
        over = tot - cap
        left = over
        for k in range(8):
            t = int(kinds[i, k]) * over // tot
            kinds[i, k] -= t
            carry[k] += t
            left -= t
        k = i % 8
        while left > 0:
            if kinds[i, k] > 0:
                kinds[i, k] -= 1
                carry[k] += 1
                left -= 1
            k = (k + 1) % 8
    pending = sum(carry)
    if pending == 0:
        return _pack(kinds)
    rest = [0] * 8
    if room_total > 0:
        give = [min(cv, cv * min(room_total, pending) // pending) for cv in carry]
        done = [0] * 8
        cum_room = 0
        for i in range(n):
            cap, tot = caps[i], int(kinds[i].sum())
            if tot >= cap:
                continue
            room = cap - tot
            cum_room += room
            for k in range(8):
                phase = k * room_total // 8
                target = (give[k] * cum_room + phase) // room_total
                want = target - done[k]
                done[k] = target
                d = min(want, room, 0xFFFF - int(kinds[i, k]))
                kinds[i, k] += d
                room -= d
                rest[k] += want - d
        for k in range(8):
            rest[k] += carry[k] - give[k]
    else:
        rest = list(carry)
    left = sum(rest)
    i = k = idle = 0
    while left > 0 and idle < n:
        cap, tot = caps[i], int(kinds[i].sum())
        if tot < cap:
            while rest[k] == 0:
                k = (k + 1) % 8
            if kinds[i, k] < 0xFFFF:
                kinds[i, k] += 1
                rest[k] -= 1
                left -= 1
                idle = 0
            k = (k + 1) % 8
        else:
            idle += 1
        i = (i + 1) % n
    if left > 0:
        j = next((j for j in range(n) if g[j] < GAMMA_SOLID), 0)
        for k in range(8):
            kinds[j, k] += rest[k]
    return _pack(kinds)


def _pack(kinds: np.ndarray) -> np.ndarray:
    w = (kinds[:, 0::2] + (kinds[:, 1::2] << 16)) & 0xFFFFFFFF
    return w.astype(np.uint32).reshape(-1)

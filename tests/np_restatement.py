"""A second, independent statement of the reference's CPU develop maths in numpy float32.

The C oracle (oracle/oracle.c) follows the reference's code structure: rows cached in strips,
AHD in 64-row bands of 192-column chunks, plane loops with masks. The functions here state the
same arithmetic per pixel over whole-frame planes, with no strips, chunks or masks, so that a
mistake in either statement (a dropped term, a wrong neighbour, a swapped candidate, a border
off by one) shows up as a difference between the two. numpy float32 arithmetic is IEEE single
precision with correct rounding and no contraction, and every expression below keeps the
reference's evaluation order, so the two statements must agree bit for bit.

Test infrastructure only (imported by tests/).
"""
from __future__ import annotations

import math

import numpy as np

F32 = np.float32
HALO = 6  # crates/flashraw/src/ahd.rs:37

XTRANS = "GGRGGB/GGBGGR/BRGRBG/GGBGGR/GGRGGB/RBGBRG"
_CODE = {"R": 0, "G": 1, "B": 2}
BAYER = {"RGGB": [[0, 1], [1, 2]], "BGGR": [[2, 1], [1, 0]], "GRBG": [[1, 0], [2, 1]], "GBRG": [[1, 2], [0, 1]]}


def parse_cfa(cfa):
    """CfaPattern::parse of rows separated by '/', or a Bayer name (types.rs:222-252)."""
    if isinstance(cfa, str) and cfa in BAYER:
        return np.asarray(BAYER[cfa], dtype=np.int64)
    if isinstance(cfa, str):
        return np.asarray([[_CODE[c] for c in row] for row in cfa.split("/")], dtype=np.int64)
    return np.asarray(cfa, dtype=np.int64)


def colours(cfa, w, h, origin=(0, 0)):
    """CFA colour of every frame pixel: the pattern at (x + origin.x, y + origin.y)."""
    p = parse_cfa(cfa)
    ys = (np.arange(h)[:, None] + origin[1]) % p.shape[0]
    xs = (np.arange(w)[None, :] + origin[0]) % p.shape[1]
    return p[ys, xs]


def normalise(mos, cfa, origin=(0, 0), black=(1, 1, 1, [0.0]), white=(16383,), row_deltas=None, col_deltas=None):
    """Calib::new + Calib::norm (process.rs:256-298) for every pixel of a one-sample mosaic:
    (v - black[pos] (+ row delta)(+ col delta)) * scale[pos], scale = 1 / (white - black) in f32."""
    h, w = mos.shape
    p = parse_cfa(cfa)
    brows, bcols, bsamples, bvals = black
    brows, bcols, bsamples = max(brows, 1), max(bcols, 1), max(bsamples, 1)
    bvals = np.asarray(bvals, dtype=F32)
    pw = min(math.lcm(p.shape[1], bcols), 64)
    ph = min(math.lcm(p.shape[0], brows), 64)
    Y = (np.arange(h)[:, None] + origin[1]) + np.zeros((1, w), dtype=np.int64)
    X = (np.arange(w)[None, :] + origin[0]) + np.zeros((h, 1), dtype=np.int64)
    yy, xx = Y % ph, X % pw
    idx = ((yy % brows) * bcols + xx % bcols) * bsamples
    bl = np.where(idx < bvals.size, bvals[np.minimum(idx, bvals.size - 1)], F32(0.0)).astype(F32)
    colour = p[yy % p.shape[0], xx % p.shape[1]]
    wv = np.asarray(white, dtype=np.int64)
    wsel = wv[np.minimum(colour, wv.size - 1)] if wv.size > 1 else np.full_like(colour, wv[0])
    wf = wsel.astype(F32)
    with np.errstate(divide="ignore"):
        scale = np.where(wf > bl, F32(1.0) / (wf - bl), F32(1.0)).astype(F32)
    b = bl.copy()
    if row_deltas is not None and len(row_deltas):
        rd = np.asarray(row_deltas, dtype=F32)
        b = b + np.where(Y < rd.size, rd[np.minimum(Y, rd.size - 1)], F32(0.0)).astype(F32)
    if col_deltas is not None and len(col_deltas):
        cd = np.asarray(col_deltas, dtype=F32)
        b = b + np.where(X < cd.size, cd[np.minimum(X, cd.size - 1)], F32(0.0)).astype(F32)
    return ((mos.astype(F32) - b) * scale).astype(F32)


def naive_bilinear(norm, col):
    """The reference test's naive demosaic (process.rs:2405-2446): own colour is the centre,
    the others the mean of the in-frame same-colour 3x3 samples, summed row-major, then s / n."""
    h, w = norm.shape
    out = np.zeros((h, w, 3), dtype=F32)
    for y in range(h):
        for x in range(w):
            own = col[y, x]
            for c in range(3):
                if c == own:
                    out[y, x, c] = norm[y, x]
                    continue
                s, n = F32(0.0), F32(0.0)
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        yy, xx = y + dy, x + dx
                        if 0 <= yy < h and 0 <= xx < w and col[yy, xx] == c:
                            s = F32(s + norm[yy, xx])
                            n = F32(n + F32(1.0))
                out[y, x, c] = F32(s / n) if n > 0 else F32(0.0)
    return out


def bilinear(norm, col, bayer=True):
    """The reference's bilinear demosaic exactly (Kernel::row_cols, process.rs:1045-1129):
    interior Bayer pixels by the fast-path formulas, other interior pixels by s * (1 / n) over
    the neighbour list, border pixels by the naive row-major mean."""
    h, w = norm.shape
    out = naive_bilinear(norm, col)  # border pixels (and a start for the interior)
    if h < 3 or w < 3:
        return out
    c = norm[1:-1, 1:-1]
    l, r = norm[1:-1, :-2], norm[1:-1, 2:]
    u, d = norm[:-2, 1:-1], norm[2:, 1:-1]
    ul, ur, dl, dr = norm[:-2, :-2], norm[:-2, 2:], norm[2:, :-2], norm[2:, 2:]
    cc = col[1:-1, 1:-1]
    if bayer:
        right = col[1:-1, 2:]
        half_lr = (l + r) * F32(0.5)
        half_ud = (u + d) * F32(0.5)
        orth = (((l + r) + u) + d) * F32(0.25)
        diag = (((ul + ur) + dl) + dr) * F32(0.25)
        R = np.where(cc == 0, c, np.where(cc == 2, diag, np.where(right == 0, half_lr, half_ud)))
        G = np.where(cc == 1, c, orth)
        B = np.where(cc == 2, c, np.where(cc == 0, diag, np.where(right == 2, half_lr, half_ud)))
        out[1:-1, 1:-1] = np.stack([R, G, B], axis=-1)
        return out
    for y in range(1, h - 1):
        for x in range(1, w - 1):
            for ch in range(3):
                if col[y, x] == ch:
                    out[y, x, ch] = norm[y, x]
                    continue
                s, n = F32(0.0), 0
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        if col[y + dy, x + dx] == ch:
                            s = F32(s + norm[y + dy, x + dx])
                            n += 1
                out[y, x, ch] = F32(s * (F32(1.0) / F32(n))) if n else F32(0.0)
    return out


def ahd(norm, col, wb):
    """The AHD demosaic (crates/flashraw/src/ahd.rs:140-392) stated per pixel. Returns camera RGB
    planes (h, w, 3) valid for HALO <= y < h - HALO, HALO <= x < w - HALO (NaN elsewhere)."""
    h, w = norm.shape
    wbv = np.asarray(wb, dtype=F32)
    m = np.minimum(norm * wbv[col], F32(1.0)).astype(F32)  # ahd.rs:149, no max(0)
    nan = F32(np.nan)

    def sh(a, dy, dx):
        """a shifted so that out[y, x] = a[y + dy, x + dx] (NaN outside)."""
        o = np.full_like(a, nan)
        ys = slice(max(0, -dy), h - max(0, dy))
        xs = slice(max(0, -dx), w - max(0, dx))
        yd = slice(max(0, dy), h - max(0, -dy))
        xd = slice(max(0, dx), w - max(0, -dx))
        o[ys, xs] = a[yd, xd]
        return o

    green = col == 1
    # 1. green along rows (H) and columns (V), limited to the range of the two nearest greens
    l, r, l2, r2 = sh(m, 0, -1), sh(m, 0, 1), sh(m, 0, -2), sh(m, 0, 2)
    u, d, u2, d2 = sh(m, -1, 0), sh(m, 1, 0), sh(m, -2, 0), sh(m, 2, 0)
    with np.errstate(invalid="ignore"):
        gh = np.minimum(np.maximum((l + r) * F32(0.5) + (F32(2.0) * m - l2 - r2) * F32(0.25), np.minimum(l, r)),
                        np.maximum(l, r))
        gv = np.minimum(np.maximum((u + d) * F32(0.5) + (F32(2.0) * m - u2 - d2) * F32(0.25), np.minimum(u, d)),
                        np.maximum(u, d))
    g = [np.where(green, m, gh), np.where(green, m, gv)]
    # 2. red and blue by colour differences; square-root lightness and chroma
    right = np.roll(col, -1, axis=1)  # colour of the right neighbour (wraps only outside the AHD region)
    hr = green & (right == 0)
    rgb, lab = [], []
    for dd in range(2):
        gd = g[dd]
        df = m - gd
        diag = gd + F32(0.25) * (((sh(df, -1, -1) + sh(df, -1, 1)) + sh(df, 1, -1)) + sh(df, 1, 1))
        row = gd + F32(0.5) * (sh(df, 0, -1) + sh(df, 0, 1))
        cl = gd + F32(0.5) * (sh(df, -1, 0) + sh(df, 1, 0))
        red = np.where(col == 0, m, np.where(green, np.where(hr, row, cl), diag))
        blue = np.where(col == 2, m, np.where(green, np.where(hr, cl, row), diag))
        with np.errstate(invalid="ignore"):
            sr, sg, sb = (np.sqrt(np.maximum(v, F32(0.0))) for v in (red, gd, blue))
        rgb.append((red, gd, blue))
        lab.append((((sr + F32(2.0) * sg) + sb) * F32(0.25), sr - sg, sb - sg))
    # 3. thresholds: the smaller of the largest row steps of H and column steps of V
    def lstep(L, dy, dx):
        return np.abs(sh(L, dy, dx) - L)

    def cstep(A, Bb, dy, dx):
        da, db = sh(A, dy, dx) - A, sh(Bb, dy, dx) - Bb
        return da * da + db * db

    LH, AH, BH = lab[0]
    LV, AV, BV = lab[1]
    with np.errstate(invalid="ignore"):
        el = np.minimum(np.maximum(lstep(LH, 0, -1), lstep(LH, 0, 1)), np.maximum(lstep(LV, -1, 0), lstep(LV, 1, 0)))
        ec = np.minimum(np.maximum(cstep(AH, BH, 0, -1), cstep(AH, BH, 0, 1)),
                        np.maximum(cstep(AV, BV, -1, 0), cstep(AV, BV, 1, 0)))
        # 4. homogeneity: neighbours within both thresholds, summed over 3x3
        hom = []
        for dd in range(2):
            L, A, Bb = lab[dd]
            cnt = np.zeros((h, w), dtype=F32)
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    if dy == 0 and dx == 0:
                        continue
                    ok = (lstep(L, dy, dx) <= el) & (cstep(A, Bb, dy, dx) <= ec)
                    cnt = cnt + ok.astype(F32)
            v = (sh(cnt, -1, 0) + cnt) + sh(cnt, 1, 0)
            hom.append((sh(v, 0, -1) + v) + sh(v, 0, 1))
    a, b = hom
    s = a + b
    with np.errstate(invalid="ignore", divide="ignore"):
        wgt = np.where(s > 0, a / s, F32(0.5)).astype(F32)
    out = np.stack([wgt * rgb[0][c] + (F32(1.0) - wgt) * rgb[1][c] for c in range(3)], axis=-1).astype(F32)
    valid = np.zeros((h, w), dtype=bool)
    valid[HALO:h - HALO, HALO:w - HALO] = True
    out[~valid] = nan
    return out


def finish(rgb, wb, matrix, levels=None):
    """finish_rgb with Highlights::Clip (process.rs:654-707): clamp(rgb * wb, 0, 1), then
    (m0 * r + m1 * g) + m2 * b per row, then v * scale - offset with levels."""
    wbv = np.asarray(wb, dtype=F32)
    c = [np.clip(rgb[..., i] * wbv[i], F32(0.0), F32(1.0)).astype(F32) for i in range(3)]
    m = np.asarray(matrix, dtype=F32)
    o = np.stack([(m[k, 0] * c[0] + m[k, 1] * c[1]) + m[k, 2] * c[2] for k in range(3)], axis=-1).astype(F32)
    if levels is not None:
        o = (o * F32(levels[0]) - F32(levels[1])).astype(F32)
    return o


def develop(mos, cfa, *, origin=(0, 0), black=(1, 1, 1, [0.0]), white=(16383,), wb=(1, 1, 1, 1),
            matrix=((1, 0, 0), (0, 1, 0), (0, 0, 1)), demosaic="bilinear", levels=None):
    """Linear f32 develop of the full frame (crop = frame, orientation 1): the demosaic, then
    finish. AHD pixels are already balanced (ahd.rs:149) and finish with unit WB
    (process.rs:1205-1212); border pixels of AHD use bilinear."""
    h, w = mos.shape
    col = colours(cfa, w, h, origin)
    norm = normalise(mos, cfa, origin, black, white)
    is_bayer = parse_cfa(cfa).shape == (2, 2)
    bil = bilinear(norm, col, bayer=is_bayer)
    out = finish(bil, wb, matrix, levels)
    if demosaic == "ahd" and h > 2 * HALO and w > 2 * HALO:
        a = ahd(norm, col, wb)
        fa = finish(a, (1.0, 1.0, 1.0), matrix, levels)
        out[HALO:h - HALO, HALO:w - HALO] = fa[HALO:h - HALO, HALO:w - HALO]
    return out


def rgba8(lin, lut):
    """Lut8 encode (process.rs:1679-1728): lut[min(trunc(clamp(v, 0, 1) * 4096), 4096)], alpha 255."""
    v = np.clip(lin, F32(0.0), F32(1.0)).astype(F32) * F32(4096.0)
    idx = np.minimum(v.astype(np.int64), 4096)
    px = lut[idx]
    return np.concatenate([px, np.full(px.shape[:-1] + (1,), 255, dtype=np.uint8)], axis=-1)


def linear16(lin):
    """(clamp(v, 0, 1) * 65535 + 0.5) as u16 (process.rs:1971-1973)."""
    v = np.clip(lin, F32(0.0), F32(1.0)).astype(F32) * F32(65535.0) + F32(0.5)
    return v.astype(np.uint16)


def orient(img, o):
    """EXIF orientation o (1..8) of an (h, w, c) image, stated with numpy flips and rotations."""
    return {
        1: lambda a: a,
        2: lambda a: a[:, ::-1],
        3: lambda a: a[::-1, ::-1],
        4: lambda a: a[::-1, :],
        5: lambda a: a.transpose(1, 0, 2),
        6: lambda a: np.rot90(a, -1),
        7: lambda a: a[::-1, ::-1].transpose(1, 0, 2),
        8: lambda a: np.rot90(a, 1),
    }[o](img)


def lcg_fill(w, h):
    """The reference tests' deterministic mosaic: sample i = (i * 7919) % 3900 + 100
    (process.rs:2331-2334, 2451-2453)."""
    i = np.arange(w * h, dtype=np.int64)
    return ((i * 7919) % 3900 + 100).astype(np.uint16).reshape(h, w)


def lcg_stream(w, h, seed=12345):
    """The reference tests' LCG mosaic: s = s * 1664525 + 1013904223, sample 100 + (s >> 20) % 3900
    (process.rs:2371-2375)."""
    out = np.empty(w * h, dtype=np.uint16)
    s = seed
    for k in range(w * h):
        s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
        out[k] = 100 + (s >> 20) % 3900
    return out.reshape(h, w)


def dct_matrix():
    """Orthonormal DCT-II (denoise/mod.rs:626-635) from its closed form, float64."""
    k = np.arange(8)[:, None]
    i = np.arange(8)[None, :]
    a = np.where(k == 0, math.sqrt(1 / 8), math.sqrt(2 / 8))
    return a * np.cos(math.pi * (2 * i + 1) * k / 16)


def dct_denoise(mos, black, sig_a, sig_b, thr=2.7, step=4):
    """The CFA-plane DCT block denoiser in float64 with scipy's DCT-II (an implementation that
    shares nothing with the oracle's matrix products): four 2x2 phase planes, 8x8 blocks at
    stride `step` plus the last position, AC coefficients with |c| < thr * sigma zeroed,
    sigma^2 = sig_a * max(mean, 0) + sig_b, uniform overlap-add. Returns the f32 mosaic in DN
    above black."""
    from scipy.fft import dctn, idctn
    h, w = mos.shape
    out = np.zeros((h, w), dtype=np.float64)
    for plane in range(4):
        dy, dx = plane >> 1, plane & 1
        p = mos[dy::2, dx::2].astype(np.float64) - float(np.float32(black[plane]))
        ph, pw = p.shape

        def pos(n):
            ps = list(range(0, n - 8 + 1, step))
            if ps[-1] != n - 8:
                ps.append(n - 8)
            return ps

        acc = np.zeros_like(p)
        cnt = np.zeros_like(p)
        for y0 in pos(ph):
            for x0 in pos(pw):
                c = dctn(p[y0:y0 + 8, x0:x0 + 8], type=2, norm="ortho")
                mean = c[0, 0] / 8.0
                t = float(np.float32(thr)) * math.sqrt(float(np.float32(sig_a)) * max(mean, 0.0)
                                                       + float(np.float32(sig_b)))
                keep = np.abs(c) >= t
                keep[0, 0] = True
                acc[y0:y0 + 8, x0:x0 + 8] += idctn(np.where(keep, c, 0.0), type=2, norm="ortho")
                cnt[y0:y0 + 8, x0:x0 + 8] += 1.0
        out[dy::2, dx::2] = acc / cnt
    return out


# ---------------------------------------------------------------- edge-aware (Bayer)

EDGE_HALO = 8  # crates/flashraw/src/edgeaware/mod.rs:66-71
CANON = (0, 1, 1, 2)  # the canonical Bayer pattern of the tables, RGGB


def edge_plan(cfa, origin):
    """Plan::new (edgeaware/mod.rs:204-230) for a Bayer pattern: the canonical position
    (cx, cy) of frame sample (0, 0)."""
    p = parse_cfa(cfa)
    for sy in range(2):
        for sx in range(2):
            if all(p[y, x] == CANON[((y + sy) % 2) * 2 + (x + sx) % 2] for y in range(2) for x in range(2)):
                return (sx + origin[0]) % 2, (sy + origin[1]) % 2
    raise ValueError("not a Bayer pattern")


def edge_tables_xtrans(seed):
    """Synthetic X-Trans tables (g0 [36][49], dir [36][3][4][49]) with the trained tables'
    layout: a 5x5 support inside the 7x7 (the outer ring zero), the own colour's directional
    filters zero, every filter non-negative and summing to 1."""
    rng = np.random.default_rng(seed)
    canon = parse_cfa(XTRANS).reshape(-1)
    ring = np.zeros((7, 7), bool)
    ring[1:6, 1:6] = True
    ring = ring.reshape(-1)

    def filt():
        f = np.where(ring, rng.random(49) ** 2, 0.0)
        f[24] += 0.5
        return (f / f.sum()).astype(F32)

    g0 = np.stack([filt() for _ in range(36)])
    d = np.zeros((36, 3, 4, 49), F32)
    for k in range(36):
        for c in range(3):
            if c != canon[k]:
                for dd in range(4):
                    d[k, c, dd] = filt()
    return g0, d


def edge_tables(seed, sparse=True):
    """Synthetic filter tables in the layout of bayer.bin (edgeaware/mod.rs:130-160): g0
    [4][49] and dir [4][3][4][49], each filter non-negative with zero taps and summing to 1 (so a
    flat field stays flat, mod.rs:13-14)."""
    rng = np.random.default_rng(seed)

    def filt():
        f = rng.random(49) ** 2
        if sparse:
            f[rng.random(49) < 0.35] = 0.0
        f[24] += 0.5
        return (f / f.sum()).astype(F32)

    g0 = np.stack([filt() for _ in range(4)])
    d = np.stack([np.stack([np.stack([filt() for _ in range(4)]) for _ in range(3)]) for _ in range(4)])
    return g0, d


def edge_plan_p(cfa, origin):
    """Plan::new (edgeaware/mod.rs:204-230) for a Bayer or X-Trans pattern: (cx, cy, canonical
    colours, period)."""
    p = parse_cfa(cfa)
    P = p.shape[0]
    canon = CANON if P == 2 else [c for c in parse_cfa(XTRANS).reshape(-1)]
    for sy in range(P):
        for sx in range(P):
            if all(p[y, x] == canon[((y + sy) % P) * P + (x + sx) % P] for y in range(P) for x in range(P)):
                return (sx + origin[0]) % P, (sy + origin[1]) % P, list(canon), P
    raise ValueError("neither Bayer nor X-Trans")


def edge_aware(norm, cfa, origin, wb, g0, dirt, beta):
    """The edge-aware demosaic (edgeaware/mod.rs:423-822) stated over whole-frame planes, one
    plane per canonical phase for the filters; Bayer or X-Trans (pass 1: the 3x3 median of the
    filtered differences; pass 2: the green re-estimate at red and blue sites). Returns camera
    RGB valid at distance >= 8 from the frame border."""
    h, w = norm.shape
    cx, cy, canon_list, P = edge_plan_p(cfa, origin)
    ys = (np.arange(h)[:, None] + cy) % P
    xs = (np.arange(w)[None, :] + cx) % P
    ph = ys * P + xs
    canon = np.asarray(canon_list)[ph]
    wbv = np.asarray(wb, dtype=F32)
    m = np.minimum(norm * wbv[canon], F32(1.0)).astype(F32)
    nan = F32(np.nan)

    def sh(a, dy, dx):
        o = np.full_like(a, nan)
        ys_ = slice(max(0, -dy), h - max(0, dy))
        xs_ = slice(max(0, -dx), w - max(0, dx))
        yd = slice(max(0, dy), h - max(0, -dy))
        xd = slice(max(0, dx), w - max(0, -dx))
        o[ys_, xs_] = a[yd, xd]
        return o

    shifted = {(ty, tx): sh(m, ty - 3, tx - 3) for ty in range(7) for tx in range(7)}
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        g0p = np.full((h, w), nan, dtype=F32)
        for k in range(P * P):
            acc = np.zeros((h, w), dtype=F32)
            for idx in range(49):
                c = g0[k, idx]
                if c != 0:
                    acc = (acc + F32(c) * shifted[(idx // 7, idx % 7)]).astype(F32)
            g0p = np.where(ph == k, acc, g0p)
        sg = np.sqrt(np.maximum(g0p, F32(0.0))).astype(F32)
        r2 = F32(0.70710677)
        dh = sh(sg, 0, 1) - sh(sg, 0, -1)
        dv = sh(sg, 1, 0) - sh(sg, -1, 0)
        d1 = (sh(sg, 1, 1) - sh(sg, -1, -1)) * r2
        d2 = (sh(sg, 1, -1) - sh(sg, -1, 1)) * r2
        wts = []
        for q in (dh * dh, dv * dv, d1 * d1, d2 * d2):
            hs = (sh(q, 0, -1) + q) + sh(q, 0, 1)
            u = ((sh(hs, -1, 0) + hs) + sh(hs, 1, 0)) + F32(1e-3)
            wts.append((F32(1.0) / (u * u)).astype(F32))
        inv = F32(1.0) / (((wts[0] + wts[1]) + wts[2]) + wts[3])
        wts = [(x * inv).astype(F32) for x in wts]
        planes = []
        for c in range(3):
            pl = np.full((h, w), nan, dtype=F32)
            for k in range(P * P):
                a = [np.zeros((h, w), dtype=F32) for _ in range(4)]
                for idx in range(49):
                    cs = dirt[k, c, :, idx]
                    if np.any(cs != 0):
                        v = shifted[(idx // 7, idx % 7)]
                        for d in range(4):
                            a[d] = (a[d] + F32(cs[d]) * v).astype(F32)
                est = (((wts[0] * a[0] + wts[1] * a[1]) + wts[2] * a[2]) + wts[3] * a[3]).astype(F32)
                pl = np.where(ph == k, est, pl)
            planes.append(np.where(canon == c, m, pl).astype(F32))
        R_, G_, B_ = planes
        # pass 1 (Bayer): gated mean of the measured differences of the 3x3 window
        inv1 = F32(1.0) / (F32(0.2) * np.maximum(G_, F32(0.0)) + F32(0.015))
        acc = {cc: [np.zeros((h, w), dtype=F32) for _ in range(4)] for cc in (0, 1)}
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                gk = sh(G_, dy, dx)
                e = (gk - G_) * inv1
                wt = F32(1.0) / (F32(1.0) + e * e)
                colk = sh(canon.astype(F32), dy, dx)
                for cc, ch, pl in ((0, 0, R_), (1, 2, B_)):
                    mk = np.where(colk == ch, F32(1.0), F32(0.0)).astype(F32)
                    d = sh(pl, dy, dx) - gk
                    sw, sd, su, n = acc[cc]
                    acc[cc] = [sw + mk * wt, sd + (mk * wt) * d, su + mk * d, n + mk]
        d1p = []
        for cc, ch, pl in ((0, 0, R_), (1, 2, B_)):
            if P == 2:
                sw, sd, su, n = acc[cc]
                mixed = np.where(sw > F32(1e-12), sd / sw, su / n)
            else:  # X-Trans: the 3x3 median of the filtered differences
                df = (pl - G_).astype(F32)
                mixed = np.median(np.stack([sh(df, dy, dx) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]), axis=0)
            d1p.append(np.where(canon == ch, pl - G_, mixed).astype(F32))
        # pass 2: the same gate over every pixel's difference
        inv2 = F32(1.0) / (F32(0.1) * np.maximum(G_, F32(0.0)) + F32(0.007))
        sw = np.zeros((h, w), dtype=F32)
        sdr = np.zeros((h, w), dtype=F32)
        sdb = np.zeros((h, w), dtype=F32)
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                e = (sh(G_, dy, dx) - G_) * inv2
                wt = F32(1.0) / (F32(1.0) + e * e)
                sw = sw + wt
                sdr = sdr + wt * sh(d1p[0], dy, dx)
                sdb = sdb + wt * sh(d1p[1], dy, dx)
        dr, db = sdr / sw, sdb / sw
        gv = G_
        if P == 6:
            xg = F32(min(max(F32(1.0) - F32(2.0) * F32(beta), F32(0.0)), F32(1.0)))
            if xg != 0:
                gv = np.where(canon == 0, G_ + xg * ((R_ - dr) - G_),
                              np.where(canon == 2, G_ + xg * ((B_ - db) - G_), G_)).astype(F32)
        v = np.stack([np.where(canon == 0, R_, gv + dr), gv, np.where(canon == 2, B_, gv + db)], axis=-1).astype(F32)
        if beta > 0:
            inv3 = F32(1.0) / (F32(0.2) * np.maximum(v[..., 1], F32(0.0)) + F32(0.015))
            sw = np.zeros((h, w), dtype=F32)
            s = [np.zeros((h, w), dtype=F32) for _ in range(3)]
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    q = [sh(v[..., c], dy, dx) for c in range(3)]
                    e = (q[1] - v[..., 1]) * inv3
                    wt = F32(1.0) / (F32(1.0) + e * e)
                    sw = sw + wt
                    s = [s[c] + wt * q[c] for c in range(3)]
            v = np.stack([v[..., c] + F32(beta) * (s[c] / sw - v[..., c]) for c in range(3)], axis=-1).astype(F32)
    out = np.full_like(v, nan)
    out[EDGE_HALO:h - EDGE_HALO, EDGE_HALO:w - EDGE_HALO] = v[EDGE_HALO:h - EDGE_HALO, EDGE_HALO:w - EDGE_HALO]
    return out


def develop_edge(mos, cfa, *, origin=(0, 0), black=(1, 1, 1, [0.0]), white=(16383,), wb=(1, 1, 1, 1),
                 matrix=((1, 0, 0), (0, 1, 0), (0, 0, 1)), g0=None, dirt=None, beta=0.0, levels=None):
    """Linear f32 develop with the edge-aware interior and the bilinear border (process.rs:1160-1225)."""
    h, w = mos.shape
    col = colours(cfa, w, h, origin)
    norm = normalise(mos, cfa, origin, black, white)
    out = finish(bilinear(norm, col, bayer=parse_cfa(cfa).shape == (2, 2)), wb, matrix, levels)
    if h > 2 * EDGE_HALO and w > 2 * EDGE_HALO:
        e = edge_aware(norm, cfa, origin, wb, g0, dirt, beta)
        fe = finish(e, (1.0, 1.0, 1.0), matrix, levels)
        s = (slice(EDGE_HALO, h - EDGE_HALO), slice(EDGE_HALO, w - EDGE_HALO))
        out[s] = fe[s]
    return out


# ---------------------------------------------------------------- inputs and layouts


def normalise_c(arr, cpp, cfa, origin=(0, 0), black=(1, 1, 1, [0.0]), white=(16383,), row_deltas=None,
                col_deltas=None):
    """Calib::new(info, c) + Calib::norm for every sample of a frame of `cpp` components
    (process.rs:256-298; kernels.rs:216-248): component c of pixel x uses black values
    [((y % rows) * cols + x % cols) * samples + c % samples] and, without a CFA, the white level
    of index min(c, n - 1). `arr` is (h, w * cpp), u16 or f32 raw values."""
    h, wc = arr.shape
    w = wc // cpp
    brows, bcols, bsamples, bvals = black
    brows, bcols, bsamples = max(brows, 1), max(bcols, 1), max(bsamples, 1)
    bvals = np.asarray(bvals, dtype=F32)
    p = parse_cfa(cfa) if cfa is not None else None
    cw, ch = (p.shape[1], p.shape[0]) if p is not None else (1, 1)
    pw, ph = min(math.lcm(cw, bcols), 64), min(math.lcm(ch, brows), 64)
    Y = (np.arange(h)[:, None] + origin[1]) + np.zeros((1, wc), dtype=np.int64)
    X = (np.arange(wc)[None, :] // cpp + origin[0]) + np.zeros((h, 1), dtype=np.int64)
    C = (np.arange(wc)[None, :] % cpp) + np.zeros((h, 1), dtype=np.int64)
    yy, xx = Y % ph, X % pw
    idx = ((yy % brows) * bcols + xx % bcols) * bsamples + C % bsamples
    bl = np.where(idx < bvals.size, bvals[np.minimum(idx, bvals.size - 1)], F32(0.0)).astype(F32)
    colour = p[yy % ch, xx % cw] if p is not None else C
    wv = np.asarray(white, dtype=np.int64)
    wsel = wv[np.minimum(colour, wv.size - 1)] if wv.size > 1 else np.full_like(colour, wv[0])
    wf = wsel.astype(F32)
    with np.errstate(divide="ignore"):
        scale = np.where(wf > bl, F32(1.0) / (wf - bl), F32(1.0)).astype(F32)
    b = bl.copy()
    if row_deltas is not None and len(row_deltas):
        rd = np.asarray(row_deltas, dtype=F32)
        b = b + np.where(Y < rd.size, rd[np.minimum(Y, rd.size - 1)], F32(0.0)).astype(F32)
    if col_deltas is not None and len(col_deltas):
        cd = np.asarray(col_deltas, dtype=F32)
        b = b + np.where(X < cd.size, cd[np.minimum(X, cd.size - 1)], F32(0.0)).astype(F32)
    return ((arr.astype(F32) - b) * scale).astype(F32)


def half_size(norm, cpp, col, crop):
    """Demosaic::HalfSize (process.rs:975-1036): output pixel (i, j) averages per colour the
    samples of the 2x2 block at (cx + 2i, cy + 2j) inside the frame, in the order (0,0), (1,0),
    (0,1), (1,1); `col` the colour plane of a CFA frame (cpp 1) or None. Returns camera RGB."""
    h, wc = norm.shape
    w = wc // cpp
    cx, cy, cw, ch = crop
    ow, oh = cw // 2, ch // 2
    acc = np.zeros((oh, ow, 4), F32)
    n = np.zeros((oh, ow, 4), F32)
    jj, ii = np.mgrid[0:oh, 0:ow]
    for dy in range(2):
        for dx in range(2):
            y = cy + 2 * jj + dy
            x = cx + 2 * ii + dx
            ok = (y < h) & (x < w)
            ys, xs = np.minimum(y, h - 1), np.minimum(x, w - 1)
            if cpp == 1 and col is not None:
                c = np.minimum(col[ys, xs], 3)
                v = norm[ys, xs]
                for k in range(4):
                    m = ok & (c == k)
                    acc[..., k] = np.where(m, acc[..., k] + v, acc[..., k]).astype(F32)
                    n[..., k] = np.where(m, n[..., k] + F32(1.0), n[..., k]).astype(F32)
            elif cpp == 1:
                v = norm[ys, xs]
                acc[..., 1] = np.where(ok, acc[..., 1] + v, acc[..., 1]).astype(F32)
                n[..., 1] = np.where(ok, n[..., 1] + F32(1.0), n[..., 1]).astype(F32)
            else:
                for k in range(min(cpp, 3)):
                    v = norm[ys, xs * cpp + k]
                    acc[..., k] = np.where(ok, acc[..., k] + v, acc[..., k]).astype(F32)
                    n[..., k] = np.where(ok, n[..., k] + F32(1.0), n[..., k]).astype(F32)
    with np.errstate(divide="ignore", invalid="ignore"):
        rgb = np.where(n[..., :3] > 0, acc[..., :3] / n[..., :3], F32(0.0)).astype(F32)
        if cpp == 1 and col is None:
            g = (acc[..., 1] / np.maximum(n[..., 1], F32(1.0))).astype(F32)
            rgb = np.stack([g, g, g], -1)
    return rgb


def packed_model(src, w, h, bits, order, stride):
    """The GPU kernels' statement of a packed sample (crates/flashraw-gpu/src/kernels.rs:333-357,
    common.wgsl packed_sample): the 3 logical bytes at bitpos / 8 of the row (word orders read the
    physical byte k ^ 1 / k ^ 3; bytes at or past row_bytes, or past the payload, read 0), then a
    shift and a mask."""
    src = np.asarray(src, dtype=np.uint32)
    row_bytes = (w * bits + 7) // 8
    ys, xs = np.mgrid[0:h, 0:w].astype(np.int64)
    bitpos = xs * bits
    k, sh = bitpos >> 3, bitpos & 7

    def byte(kk):
        p = kk ^ 1 if order == 2 else (kk ^ 3 if order == 3 else kk)
        a = ys * stride + p
        ok = (p < row_bytes) & (a < src.size)
        return np.where(ok, src[np.minimum(a, max(src.size - 1, 0))] if src.size else 0, 0).astype(np.uint32)

    b0, b1, b2 = byte(k), byte(k + 1), byte(k + 2)
    mask = np.uint32((1 << bits) - 1)
    if order == 1:
        v = ((b0 | (b1 << 8) | (b2 << 16)) >> sh.astype(np.uint32)) & mask
    else:
        v = (((b0 << 16) | (b1 << 8) | b2) >> (24 - sh - bits).astype(np.uint32)) & mask
    return v.astype(np.uint16)


def pack(samples, bits, order, stride):
    """Pack rows of samples (h, w) by the format definitions (decoder.rs:203-212): the MSB-first
    or LSB-first bit stream of each row, then for the word orders the MSB stream's bytes stored as
    little-endian 16- or 32-bit words (the row's byte count must then be a multiple of the word);
    rows start `stride` bytes apart, padding bytes 0xA5."""
    h, w = samples.shape
    row_bytes = (w * bits + 7) // 8
    wsz = 2 if order == 2 else (4 if order == 3 else 1)
    assert row_bytes % wsz == 0 and stride >= row_bytes
    out = bytearray([0xA5]) * (stride * h)
    for y in range(h):
        acc = 0
        if order == 1:
            for x in range(w):
                acc |= int(samples[y, x]) << (x * bits)
            row = bytearray(acc.to_bytes(row_bytes, "little"))
        else:
            for x in range(w):
                acc = (acc << bits) | int(samples[y, x])
            acc <<= row_bytes * 8 - w * bits
            row = bytearray(acc.to_bytes(row_bytes, "big"))
            for i in range(0, row_bytes, wsz):
                row[i:i + wsz] = row[i:i + wsz][::-1]
        out[y * stride:y * stride + row_bytes] = row
    return np.frombuffer(bytes(out), dtype=np.uint8)


# ---------------------------------------------------------------- display tone and curves


def tone_apply(lin, table, xmax, scale, m, white):
    """ToneMap::apply (crates/flashraw/src/tone.rs:148-207) per pixel of linear output RGB,
    in float32 with the reference's order; max/min are plain comparisons."""
    lin = lin.astype(F32)
    tab = np.asarray(table, F32).reshape(-1, 2)
    m = np.asarray(m, F32).reshape(3, 3)
    w = np.asarray(white, F32)
    p0, p1, p2 = lin[..., 0], lin[..., 1], lin[..., 2]

    def mx(a, b):
        return np.where(a > b, a, b).astype(F32)

    def mn(a, b):
        return np.where(a < b, a, b).astype(F32)

    c = [mx(((m[k, 0] * p0 + m[k, 1] * p1) + m[k, 2] * p2).astype(F32), F32(0.0)) for k in range(3)]
    hi, lo = mx(mx(c[0], c[1]), c[2]), mn(mn(c[0], c[1]), c[2])

    def curve(x):
        t = (np.sqrt(mn(x, F32(xmax))) * F32(scale)).astype(F32)
        t1 = mn(t, F32(len(tab) - 1))
        i = t1.astype(np.int64)
        return (tab[i, 0] + tab[i, 1] * (t - i.astype(F32))).astype(F32)

    fh, fl = curve(hi), curve(lo)
    d = (hi - lo).astype(F32)
    with np.errstate(divide="ignore", invalid="ignore"):
        k = np.where(d > 0, (fh - fl) / d, F32(0.0)).astype(F32)
    b = (fl - k * lo).astype(F32)
    return np.stack([(b * w[j] + k * lin[..., j]).astype(F32) for j in range(3)], -1)


def curve8(lin, lut):
    """Lut8 (process.rs:1679-1728): table[min(trunc(clamp(v, 0, 1) * (len - 1)), len - 1)]."""
    n = len(lut) - 1
    with np.errstate(invalid="ignore"):
        x = np.clip(lin, F32(0.0), F32(1.0)) * F32(n)
        i = np.where(np.isnan(x), 0, np.minimum(x, n)).astype(np.int64)
    return np.asarray(lut, np.uint8)[i]


def curve16(lin, lut):
    """srgb16 with a curve's table (process.rs:1942-1950): (a + (b - a) t + 0.5) as u16."""
    lut = np.asarray(lut, F32)
    x = (np.clip(lin, F32(0.0), F32(1.0)) * F32(16384.0)).astype(F32)
    i = np.minimum(np.where(np.isnan(x), 0, x).astype(np.int64), 16383)
    t = (x - i.astype(F32)).astype(F32)
    a, b = lut[i], lut[i + 1]
    v = (a + (b - a) * t + F32(0.5)).astype(F32)
    return np.clip(np.where(np.isnan(v), 0, v), 0, 65535).astype(np.uint16)


def synthetic_tone(ev=0.7, n=1024):
    """A tone in the GpuTone layout: an S-curve f over gamma-2.2 values tabulated at u = i / n as
    f(u^2), (value, step) pairs; the curve reaches 1 at xmax = 2^-ev; a ProPhoto-like matrix with
    unit row sums and its white."""
    u = np.arange(n + 1, dtype=np.float64) / n
    g = (u * u) ** (1 / 2.2)
    s = 3 * g ** 2 - 2 * g ** 3
    f = np.clip(0.6 * s + 0.4 * g, 0, 1) ** 2.2
    f32 = f.astype(F32)
    table = np.stack([f32[:-1], (f32[1:] - f32[:-1]).astype(F32)], -1)
    xmax = 2.0 ** -ev
    m = np.array([[0.53, 0.33, 0.14], [0.10, 0.82, 0.08], [0.02, 0.12, 0.86]])
    white = np.linalg.inv(m).sum(axis=1)
    return dict(table=table, xmax=F32(xmax), scale=F32(n / np.sqrt(xmax)), to_prophoto=m.astype(F32),
                white=white.astype(F32))


def synthetic_curve():
    """CurveTables (crates/flashraw/src/color.rs:356-366) of the gamma 1/1.8 encoding: 16385
    8-bit codes and 16385 16-bit codes."""
    x = np.arange(16385, dtype=np.float64) / 16384
    enc = x ** (1 / 1.8)
    lut8 = np.clip(enc * 255.0 + 0.5, 0, 255).astype(np.uint8)
    lut16 = (np.clip(enc, 0, 1) * 65535.0).astype(F32)
    return lut8, lut16

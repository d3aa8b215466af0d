"""CPU model of the fused gfx950 kernel (nanopyx_amd/csrc/npx_kernels.hip, esrrf_fused), block by block.

It repeats the kernel's index arithmetic statement for statement — tiles, halos, the row then column Catmull-Rom
passes of the lattice, the two column-parity lattice planes, the per-lane column sweep (u and v taken divided by
|dx|, the column's sign in the operands), and the edge normaliser — on the library's own tap tables and the ring
geometry each block forms from its pixel centres (block_ring), in float32, and
asserts that every LDS / table / global index the kernel forms stays inside its buffer and that every tap with a
non-zero weight reads inside its tile (clamped taps must carry weight 0). Tests run it on small stacks against the
oracle, so an indexing mistake shows up on the CPU before the kernel ever runs on a GPU.
"""
from __future__ import annotations

import numpy as np

from nanopyx_amd import _hip

f32 = np.float32
RING_SCALE = 65536.0  # kRingScale of the kernel: the lattice holds s G/|G|; indicator clamp(-s u / |dx|, 0, 1)


def _fma(a, b, c):
    """float32 a * b + c with one rounding of the exact product sum (float64 holds the product exactly)."""
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(np.float32)


def model_frames(stack, M, R, S, iw=True, Q=None):
    """Per-frame maps as the kernel computes them with Q input rows per lane (default: the per-frame kernel's)."""
    return model_multi(stack, M, [R], [S], iw=iw, Q=Q)[(float(R), float(S))]


def _pow_mode(Sf):
    return 0 if not Sf > 1 else {2.0: 2, 3.0: 3, 4.0: 4}.get(float(Sf), 1)


def block_ring(M, K, xr, yr, tSS, tSO):
    """The ring constants one block stages for half-width K, as the kernel forms them: each output pixel centre
    float32((r M + p + 0.5) / M) of the block's last column xr and last row yr
    (_c_sr_radial_gradient_convergence.c:21-22), window offsets float32((int(2 c) + o - K) / 2) - c (:31-37),
    d = sqrt(float32 dx^2 + dy^2) in double rounded to float32, w = (d exp(-d^2 / tSS))^4 in double (:4-6, :44),
    t = w / d; wsum is the float32 sum of w (j outer, i inner) and the prefix sums the double row sums, then column
    sums, of w. Returns dy[P, NW], dx[P, NW], d and t [P, NW(i), NW(j)], wsum[P], prefix[P, NW + 1, NW + 1]."""
    NW = 2 * K + 1
    o = np.arange(NW) - K

    def offsets(r):
        c = np.array([(M * r + p + 0.5) / M for p in range(M)], np.float64).astype(f32)
        v = ((np.floor(f32(2) * c)[:, None] + o[None, :]).astype(f32) / f32(2)).astype(f32)
        return (v - c[:, None]).astype(f32)                        # [phase index, o]

    offx, offy = offsets(xr), offsets(yr)
    ph = np.arange(M * M)
    dx, dy = offx[ph % M], offy[ph // M]                           # [P, NW]
    dd = ((dx * dx)[:, :, None] + (dy * dy)[:, None, :]).astype(f32)
    d = np.sqrt(dd.astype(np.float64)).astype(f32)
    ok = (d != 0) & (d <= f32(tSO))
    dn = np.where(ok, d, 1).astype(np.float64)
    y = dn * np.exp((-dn * dn) / np.float64(f32(tSS)))
    w = np.where(ok, ((y * y) * (y * y)).astype(f32), f32(0))
    t = np.where(ok, (w.astype(np.float64) / dn).astype(f32), f32(0))
    wj = np.transpose(w, (0, 2, 1))                                # [P, j, i]
    wsum = np.cumsum(wj.reshape(len(ph), -1), axis=1, dtype=f32)[:, -1]
    prefix = np.zeros((len(ph), NW + 1, NW + 1), f32)
    rows = np.cumsum(wj.astype(np.float64), axis=2)                # sum over i' <= i of w[j][i']
    prefix[:, 1:, 1:] = np.cumsum(rows, axis=1).astype(f32)
    return dict(dy=dy, dx=dx, d=d, t=t, wsum=wsum, prefix=prefix)


def model_multi(stack, M, radii, sens, iw=True, Q=None, KL=None, arrays=None):
    """Per-frame maps of every (radius, sensitivity), computed as one block of the batched kernel does: the lattice
    is built for half-width KL (default: the largest K of the radii) and every wave runs the ring of its own radius
    at offset KL - K into it. Returns {(R, S): (N, HM, WM) float32}.

    arrays = (gx, gy, img, pad_edges): the kernel's arrays mode (npx_rgc): the lattice is gathered from the caller's
    (N, 2HM, 2WM) gradients, the weight is the caller's (N, HM, WM) image and pad_edges sets the margin to 0; the
    stack then only gives (N, H, W)."""
    stack = np.asarray(stack)
    N, H, W = stack.shape
    tabs = [_hip.plan_tables(H, W, M, R) for R in radii]
    infos = [t["info"] for t in tabs]
    t = tabs[0]
    d = infos[0]
    Ks = [i["K"] for i in infos]
    KL = max(Ks) if KL is None else KL
    assert KL >= max(Ks)
    HG = KL // 2 + 3
    HI = HG + 1
    LP = 64 + KL
    Q = d["Q_frames"] if Q is None else Q
    TIR, TIC = Q + 2 * HI, 64 + 2 * HI
    TGR, TGC = Q + 2 * HG, 64 + 2 * HG
    LR = 2 * Q + 2 * KL
    HM, WM, nphase = d["HM"], d["WM"], d["nphase"]
    LPL = LR * LP
    outs = {(float(R), float(S)): np.zeros((N, HM, WM), np.float32) for R in radii for S in sens}
    nregx, nregy = (W + 63) // 64, (H + Q - 1) // Q
    lanes = np.arange(64)
    ring_cache = {}
    for f in range(N):
        img = stack[f].astype(np.float32)
        for by in range(nregy):
            for bx in range(nregx):
                x0, y0 = bx * 64, by * Q
                # (1) input tile
                s_in = np.zeros((TIR, TIC), np.float32)
                gy = y0 - HI + np.arange(TIR)
                gx = x0 - HI + np.arange(TIC)
                oky, okx = (gy >= 0) & (gy < H), (gx >= 0) & (gx < W)
                s_in[np.ix_(oky, okx)] = img[np.ix_(gy[oky], gx[okx])]
                if arrays is None:
                    # (2) Roberts
                    s_gc = np.zeros((TGR, TGC), np.float32)
                    s_gr = np.zeros((TGR, TGC), np.float32)
                    for r in range(TGR):
                        yy = y0 - HG + r
                        if yy < 0 or yy >= H:
                            continue
                        cc = np.arange(TGC)
                        xx = x0 - HG + cc
                        ok = (xx >= 0) & (xx < W)
                        r1 = yy - (y0 - HI)
                        r0 = (yy - 1 if yy > 0 else 0) - (y0 - HI)
                        c1 = xx - (x0 - HI)
                        c0 = np.where(xx > 0, xx - 1, 0) - (x0 - HI)
                        assert 0 <= r0 < TIR and 0 <= r1 < TIR
                        assert (c0[ok] >= 0).all() and (c1[ok] < TIC).all()
                        a_r0c1 = s_in[r0, c1[ok]]
                        a_r1c0 = s_in[r1, c0[ok]]
                        a_r0c0 = s_in[r0, c0[ok]]
                        a_r1c1 = s_in[r1, c1[ok]]
                        s_gc[r, ok] = a_r0c1 - a_r1c0 + a_r1c1 - a_r0c0
                        s_gr[r, ok] = -a_r0c1 + a_r1c0 + a_r1c1 - a_r0c0
                    # (3a) rows: Catmull-Rom of the gradient tile along rows for every lattice row
                    a_lo, b_lo = 2 * y0 - KL, 2 * x0 - KL
                    s_rx = np.zeros((LR, TGC), np.float32)
                    s_ry = np.zeros((LR, TGC), np.float32)
                    for la in range(LR):
                        aa = a_lo + la
                        if not (1 <= aa <= 2 * HM - 2):
                            continue
                        ai = min(aa, 2 * H)
                        wr = t["lrw"][ai]
                        rb = t["lrb"][ai] - (y0 - HG)
                        sx = np.zeros(TGC, np.float32)
                        sy = np.zeros(TGC, np.float32)
                        for u in range(4):
                            tr = min(max(rb + u, 0), TGR - 1)
                            if wr[u] != 0:
                                assert 0 <= rb + u < TGR, ("lattice row tap outside tile", aa)
                            sx = (sx + s_gc[tr] * wr[u]).astype(np.float32)
                            sy = (sy + s_gr[tr] * wr[u]).astype(np.float32)
                        s_rx[la], s_ry[la] = sx, sy
                    # (3b) columns + normalisation into the two column-parity planes: one lattice column per thread
                    # (plane, position), every rg_n-th lattice row of it
                    s_lat = np.zeros((2 * LPL, 2), np.float32)
                    seen = np.zeros(2 * LPL, bool)
                    NCOL = 2 * LP
                    NTHR = 64 * min(16, nphase * len(radii))  # threads of a block (up to 16 wave slots)
                    rg_n = max(1, NTHR // NCOL)
                    for tt in range(NCOL * rg_n):
                        col, rg = tt % NCOL, tt // NCOL
                        plane = 1 if col >= LP else 0
                        pos = col - plane * LP
                        bb = b_lo + 2 * pos + plane
                        for la in range(rg, LR, rg_n):
                            idx = plane * LPL + la * LP + pos
                            assert 0 <= idx < 2 * LPL and not seen[idx]
                            seen[idx] = True
                            if not (1 <= bb <= 2 * WM - 2):
                                continue
                            bi = min(bb, 2 * W)
                            wc = t["lcw"][bi]
                            cb = t["lcb"][bi] - (x0 - HG)
                            gxv = f32(0)
                            gyv = f32(0)
                            for v in range(4):
                                tc = min(max(cb + v, 0), TGC - 1)
                                if wc[v] != 0:
                                    assert 0 <= cb + v < TGC, ("lattice column tap outside tile", bb)
                                gxv = f32(gxv + s_rx[la, tc] * wc[v])
                                gyv = f32(gyv + s_ry[la, tc] * wc[v])
                            n2 = f32(gxv * gxv + gyv * gyv)
                            if n2 > 0:
                                inv = f32(1.0 / np.sqrt(np.float64(n2)))
                                s_lat[idx] = (gxv * inv, gyv * inv)
                    assert seen.all(), "lattice points left unwritten"
                else:
                    # lattice point (a, b): the caller's gradient at (min(a M, 2 HM - 1), min(b M, 2 WM - 1)),
                    # normalised; 0 outside a in [1, 2 HM - 2], b in [1, 2 WM - 2]
                    agx, agy = arrays[0], arrays[1]
                    s_lat = np.zeros((2 * LPL, 2), np.float32)
                    for idx in range(2 * LPL):
                        plane, rem = divmod(idx, LPL)
                        la, pos = divmod(rem, LP)
                        aa, bb = 2 * y0 - KL + la, 2 * x0 - KL + 2 * pos + plane
                        if not (1 <= aa <= 2 * HM - 2 and 1 <= bb <= 2 * WM - 2):
                            continue
                        ra, ca = min(aa * M, 2 * HM - 1), min(bb * M, 2 * WM - 1)
                        assert 0 <= ra < 2 * HM and 0 <= ca < 2 * WM
                        gxv, gyv = f32(agx[f, ra, ca]), f32(agy[f, ra, ca])
                        n2 = f32(gxv * gxv + gyv * gyv)
                        if n2 > 0:
                            inv = f32(1.0 / np.sqrt(np.float64(n2)))
                            s_lat[idx] = (gxv * inv, gyv * inv)
                # (4)+(5) waves: one per (radius, phase)
                for ri, R in enumerate(radii):
                    K = Ks[ri]
                    DK = KL - K
                    NW = 2 * K + 1
                    NR = 2 * Q + 2 * K - 1
                    margin = 0 if (arrays is not None and arrays[3]) else infos[ri]["margin"]
                    key = (ri, min(x0 + 63, W - 1), min(y0 + Q - 1, H - 1))
                    if key not in ring_cache:
                        ring_cache[key] = block_ring(M, K, key[1], key[2], infos[ri]["tSS"], infos[ri]["tSO"])
                    br = ring_cache[key]
                    dyt, dxt, dtab, ttab = br["dy"], br["dx"], br["d"], br["t"]
                    wsum = br["wsum"]
                    for ph in range(nphase):
                        py, px = divmod(ph, M)
                        ey, ex = (2 * py + 1) // M, (2 * px + 1) // M
                        xi = x0 + lanes
                        cM = M * xi + px
                        acc = np.zeros((Q, 64), np.float32)
                        for i in range(NW):
                            c = ex + i + DK
                            base = (c & 1) * LPL + (ey + DK) * LP + lanes + (c >> 1)
                            rows = base[None, :] + np.arange(NR)[:, None] * LP
                            assert rows.min() >= 0 and rows.max() < 2 * LPL
                            # each read stays inside its own column-parity plane
                            assert ((rows - (c & 1) * LPL) < LPL).all()
                            g = s_lat[rows]                       # (NR, 64, 2)
                            # the kernel stages -dy, 1 / dx (0 for dx = 0), t d and -t |dx| / s (s = RING_SCALE,
                            # |dx| taken as 1 for dx = 0); the lattice holds s g, and u, v enter divided by |dx|
                            dxv = np.float32(dxt[ph, i])
                            rdx = np.float32(np.float32(1.0) / dxv) if dxv != 0 else np.float32(0)
                            sg = 1 if rdx > 0 else (-1 if rdx < 0 else 0)
                            rs = np.float32(abs(rdx)) if sg else np.float32(1)
                            dxa = np.float32(abs(dxv)) if sg else np.float32(1)
                            gs = (g * np.float32(RING_SCALE)).astype(np.float32)  # exact: a power of two
                            for r in range(NR):
                                for jj in range(NW):
                                    rq = r - jj
                                    if rq < 0 or rq % 2 or rq // 2 >= Q:
                                        continue
                                    q = rq // 2
                                    dyn = np.float32(-dyt[ph, jj])
                                    cj = dyn if sg == 0 else np.float32(dyn * rs)
                                    td = np.float32(ttab[ph, i, jj] * dtab[ph, i, jj])
                                    ts = np.float32(np.float32(ttab[ph, i, jj] * np.float32(-1.0 / RING_SCALE)) * dxa)
                                    gx, gy = gs[r, :, 0], gs[r, :, 1]
                                    if sg > 0:
                                        z, vs = _fma(gy, cj, -gx), _fma(gx, cj, gy)
                                    elif sg < 0:
                                        z, vs = _fma(gy, cj, gx), _fma(gx, cj, -gy)
                                    else:
                                        z, vs = (gy * cj).astype(np.float32), (gx * cj).astype(np.float32)
                                    c = np.clip(z, 0, 1).astype(np.float32)
                                    x = _fma(np.abs(vs), ts, td)
                                    acc[q] = _fma(x, c, acc[q])
                        cMs = np.minimum(cM, WM - 1)
                        wcol = t["ccw"][cMs]
                        ccol = t["ccb"][cMs] - (x0 - HI)
                        for q in range(Q):
                            yi = y0 + q
                            rM = M * yi + py
                            inside = (xi < W) & (yi < H)
                            valid = inside & (rM >= margin) & (rM < HM - margin) & (cM >= margin) & (cM < WM - margin)
                            a0 = 2 * yi + ey
                            b0 = 2 * xi + ex
                            den = np.full(64, wsum[ph], np.float32)
                            # border pixels: normaliser from the 2D prefix sums of w over the valid window part
                            P = br["prefix"][ph]
                            j0 = max(0, 1 - (a0 - K))
                            j1 = max(0, min(NW, 2 * HM - 2 - (a0 - K) + 1))
                            i0 = np.maximum(0, 1 - (b0 - K))
                            i1 = np.maximum(0, np.minimum(NW, 2 * WM - 2 - (b0 - K) + 1))
                            assert 0 <= j0 <= NW and 0 <= j1 <= NW and (i0 >= 0).all() and (i0 <= NW).all()
                            edge = (j0 > 0) | (j1 < NW) | (i0 > 0) | (i1 < NW)
                            jl = min(j0, j1)
                            il = np.minimum(i0, i1)
                            pd = ((P[j1, i1] - P[jl, i1]) - (P[j1, il] - P[jl, il])).astype(np.float32)
                            den = np.where(edge, pd, den).astype(np.float32)
                            with np.errstate(invalid="ignore", divide="ignore"):
                                rgc0 = np.where(den > 0, acc[q] / den, 0).astype(np.float32)
                            inten = None
                            if iw and arrays is not None:
                                inten = arrays[2][f, min(rM, HM - 1), cMs].astype(np.float32)
                            elif iw:
                                rMs = min(rM, HM - 1)
                                wrow = t["crw"][rMs]
                                crow = t["crb"][rMs] - (y0 - HI)
                                sacc = np.zeros(64, np.float32)
                                for v in range(4):
                                    # every tap of every lane (also lanes past the image and zero-weight taps) reads
                                    # inside the tile: the kernel reads them without clamping
                                    assert ((ccol + v >= 0) & (ccol + v < TIC)).all(), "intensity column tap outside tile"
                                    tc = ccol + v
                                    p = np.zeros(64, np.float32)
                                    for u in range(4):
                                        assert 0 <= crow + u < TIR, "intensity row tap outside tile"
                                        tr = crow + u
                                        p = p + s_in[tr, tc] * wrow[u]
                                    sacc = sacc + p * wcol[:, v]
                                inten = sacc
                            for S in sens:
                                Sf = f32(S)
                                pow_mode = _pow_mode(Sf)
                                rgc = rgc0
                                if pow_mode == 2:
                                    rgc = rgc * rgc
                                elif pow_mode == 3:
                                    rgc = rgc * rgc * rgc
                                elif pow_mode == 4:
                                    rgc = (rgc * rgc) * (rgc * rgc)
                                elif pow_mode == 1:
                                    rgc = np.where(rgc > 0, np.exp2(Sf * np.log2(np.maximum(rgc, 1e-38))),
                                                   0).astype(np.float32)
                                if inten is not None:
                                    rgc = rgc * inten
                                v_out = np.where(valid, rgc, 0).astype(np.float32)
                                sel = inside
                                if sel.any():
                                    assert rM < HM and (cM[sel] < WM).all()
                                    outs[(float(R), float(S))][f, rM, cM[sel]] = v_out[sel]
    return outs

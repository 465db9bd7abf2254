"""CPU emulation of the paired multigrid cell-level kernels.

k_mg_down2 / k_mg_up2 (csrc/isph_mg.hip) give each block a tile of
MG_TX x MG_TY cells of the pair's coarser level and recompute on the two
finer levels the one-cell halos the tile needs, so that every value equals
the one-level-per-launch k_mg_down / k_mg_grid_up result bit for bit.  This
test restates both forms in FP32 with the kernels' expressions, loop
bounds, tile origins, clamped neighbours and write ownership, and checks on
odd and even grids with empty cells that
- the paired form reproduces the level-by-level results bit for bit;
- every LDS and global index stays inside its array (chk);
- every output cell is written (outputs start as NaN).
The tile shape is read from the kernel source, and a few other shapes are
checked too.  On the GPU the 1M multigrid solve's true residual is the same
to the last digit with either form; this emulation is the CPU evidence for
the index arithmetic, not a test of the compiled kernels.
"""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
f32 = np.float32
OM = f32(0.8)
OVER = f32(1.5)
TX, TY = 4, 8  # set per case


def shipped_tile():
    text = open(os.path.join(ROOT, "beach_slope_confined_amd", "csrc", "isph_mg.hip")).read()
    m = re.search(r"constexpr int MG_TX = (\d+), MG_TY = (\d+);", text)
    assert m, "tile shape not found in isph_mg.hip"
    return int(m.group(1)), int(m.group(2))


def clamp(e, n): return min(max(e, 0), n - 1)
def jac0(A, nc, c, b):
    dg = A[4 * nc + c]
    return f32(OM * b[c] / dg) if dg != 0 else f32(0)
def down(nx, ny, A, b, x, nx2, ny2, b2):
    nc = nx * ny
    for C in range(nx2 * ny2):
        X, Y = divmod(C, ny2); rs = []
        for k in range(4):
            cx, cy = 2 * X + (k >> 1), 2 * Y + (k & 1); r = f32(0)
            if cx < nx and cy < ny:
                c = cx * ny + cy; acc = f32(0)
                for d in range(9):
                    ex, ey = clamp(cx + d // 3 - 1, nx), clamp(cy + d % 3 - 1, ny)
                    acc = f32(acc + f32(A[d * nc + c] * jac0(A, nc, ex * ny + ey, b)))
                x[c] = jac0(A, nc, c, b); r = f32(b[c] - acc)
            rs.append(r)
        b2[C] = f32(f32(rs[0] + rs[1]) + f32(rs[2] + rs[3]))
def up(nx, ny, A, b, x, ny2, e, out):
    nc = nx * ny
    z = lambda ex, ey: f32(x[ex * ny + ey] + f32(OVER * e[(ex >> 1) * ny2 + (ey >> 1)]))
    for c in range(nc):
        cx, cy = divmod(c, ny); acc = f32(0)
        for d in range(9):
            ex, ey = clamp(cx + d // 3 - 1, nx), clamp(cy + d % 3 - 1, ny)
            acc = f32(acc + f32(A[d * nc + c] * z(ex, ey)))
        dg = A[4 * nc + c]; zc = z(cx, cy)
        out[c] = f32(zc + f32(OM * f32(b[c] - acc)) / dg) if dg != 0 else f32(0)
def chk(a, n):
    assert 0 <= a < n, (a, n); return a
def down2(nx0, ny0, A0, b0, x0, nx1, ny1, A1, b1, x1, nx2, ny2, b2):
    MW, MH = 2 * TX + 2, 2 * TY + 2; RW, RH = 2 * MW, 2 * MH; FW, FH = RW + 2, RH + 2
    nc0, nc1 = nx0 * ny0, nx1 * ny1; nty = (ny2 + TY - 1) // TY
    tiles = ((nx2 + TX - 1) // TX) * nty
    for blk in range(tiles):
        X0, Y0 = (blk // nty) * TX, (blk % nty) * TY
        mx0, my0 = 2 * X0 - 1, 2 * Y0 - 1; rx0, ry0 = 2 * mx0, 2 * my0; fx0, fy0 = rx0 - 1, ry0 - 1
        xf = [None] * (FW * FH); rf = [None] * (RW * RH); bm = [None] * (MW * MH); xm = [None] * (MW * MH); rm = [None] * (4 * TX * TY)
        for e in range(FW * FH):
            i, j = divmod(e, FH); ox, oy = fx0 + i, fy0 + j
            c = clamp(ox, nx0) * ny0 + clamp(oy, ny0); v = jac0(A0, nc0, c, b0); xf[e] = v
            if 4 * X0 <= ox < min(4 * (X0 + TX), nx0) and 4 * Y0 <= oy < min(4 * (Y0 + TY), ny0):
                x0[chk(c, nc0)] = v
        for e in range(RW * RH):
            i, j = divmod(e, RH); gx, gy = rx0 + i, ry0 + j; r = f32(0)
            if 0 <= gx < nx0 and 0 <= gy < ny0:
                c = gx * ny0 + gy; acc = f32(0)
                for d in range(9):
                    ex, ey = clamp(gx + d // 3 - 1, nx0), clamp(gy + d % 3 - 1, ny0)
                    acc = f32(acc + f32(A0[d * nc0 + c] * xf[chk((ex - fx0) * FH + (ey - fy0), FW * FH)]))
                r = f32(b0[c] - acc)
            rf[e] = r
        for e in range(MW * MH):
            p, q = divmod(e, MH); li, lj = 2 * p, 2 * q
            b = f32(f32(rf[chk(li * RH + lj, RW * RH)] + rf[chk(li * RH + lj + 1, RW * RH)]) + f32(rf[chk((li + 1) * RH + lj, RW * RH)] + rf[chk((li + 1) * RH + lj + 1, RW * RH)]))
            bm[e] = b; gx, gy = mx0 + p, my0 + q
            if 2 * X0 <= gx < min(2 * (X0 + TX), nx1) and 2 * Y0 <= gy < min(2 * (Y0 + TY), ny1):
                b1[chk(gx * ny1 + gy, nc1)] = b
        for e in range(MW * MH):
            p, q = divmod(e, MH); ox, oy = mx0 + p, my0 + q; gx, gy = clamp(ox, nx1), clamp(oy, ny1)
            dg = A1[4 * nc1 + gx * ny1 + gy]
            v = f32(OM * bm[chk((gx - mx0) * MH + (gy - my0), MW * MH)] / dg) if dg != 0 else f32(0)
            xm[e] = v
            if 2 * X0 <= ox < min(2 * (X0 + TX), nx1) and 2 * Y0 <= oy < min(2 * (Y0 + TY), ny1):
                x1[gx * ny1 + gy] = v
        for e in range(4 * TX * TY):
            p, q = divmod(e, 2 * TY); gx, gy = 2 * X0 + p, 2 * Y0 + q; r = f32(0)
            if gx < nx1 and gy < ny1:
                c = gx * ny1 + gy; acc = f32(0)
                for d in range(9):
                    ex, ey = clamp(gx + d // 3 - 1, nx1), clamp(gy + d % 3 - 1, ny1)
                    acc = f32(acc + f32(A1[d * nc1 + c] * xm[chk((ex - mx0) * MH + (ey - my0), MW * MH)]))
                r = f32(bm[chk((gx - mx0) * MH + (gy - my0), MW * MH)] - acc)
            rm[e] = r
        for t in range(TX * TY):
            P, Q = divmod(t, TY); X, Y = X0 + P, Y0 + Q
            if X < nx2 and Y < ny2:
                li, lj = 2 * P, 2 * Q
                b2[X * ny2 + Y] = f32(f32(rm[li * 2 * TY + lj] + rm[li * 2 * TY + lj + 1]) + f32(rm[(li + 1) * 2 * TY + lj] + rm[(li + 1) * 2 * TY + lj + 1]))
def up2(nx0, ny0, A0, b0, x0, nx1, ny1, A1, b1, x1, nx2, ny2, e2, out1, out0):
    FW, FH = 4 * TX + 2, 4 * TY + 2; MW, MH = 2 * TX + 2, 2 * TY + 2; ZW, ZH = MW + 2, MH + 2
    nc0, nc1 = nx0 * ny0, nx1 * ny1; nty = (ny2 + TY - 1) // TY
    tiles = ((nx2 + TX - 1) // TX) * nty
    for blk in range(tiles):
        X0, Y0 = (blk // nty) * TX, (blk % nty) * TY
        fx0, fy0 = 4 * X0 - 1, 4 * Y0 - 1; mx0, my0 = 2 * X0 - 1, 2 * Y0 - 1; zx0, zy0 = mx0 - 1, my0 - 1
        z1 = [None] * (ZW * ZH); o1 = [None] * (MW * MH); z0 = [None] * (FW * FH)
        for e in range(ZW * ZH):
            i, j = divmod(e, ZH); ex, ey = clamp(zx0 + i, nx1), clamp(zy0 + j, ny1)
            z1[e] = f32(x1[ex * ny1 + ey] + f32(OVER * e2[chk((ex >> 1) * ny2 + (ey >> 1), nx2 * ny2)]))
        for e in range(MW * MH):
            p, q = divmod(e, MH); gx, gy = mx0 + p, my0 + q; v = f32(0)
            if 0 <= gx < nx1 and 0 <= gy < ny1:
                c = gx * ny1 + gy; acc = f32(0)
                for d in range(9):
                    ex, ey = clamp(gx + d // 3 - 1, nx1), clamp(gy + d % 3 - 1, ny1)
                    acc = f32(acc + f32(A1[d * nc1 + c] * z1[chk((ex - zx0) * ZH + (ey - zy0), ZW * ZH)]))
                dg = A1[4 * nc1 + c]; zc = z1[chk((gx - zx0) * ZH + (gy - zy0), ZW * ZH)]
                v = f32(zc + f32(OM * f32(b1[c] - acc)) / dg) if dg != 0 else f32(0)
                if 2 * X0 <= gx < 2 * (X0 + TX) and 2 * Y0 <= gy < 2 * (Y0 + TY):
                    out1[c] = v
            o1[e] = v
        for e in range(FW * FH):
            i, j = divmod(e, FH); ex, ey = clamp(fx0 + i, nx0), clamp(fy0 + j, ny0)
            z0[e] = f32(x0[ex * ny0 + ey] + f32(OVER * o1[chk(((ex >> 1) - mx0) * MH + ((ey >> 1) - my0), MW * MH)]))
        for e in range(16 * TX * TY):
            p, q = divmod(e, 4 * TY); gx, gy = 4 * X0 + p, 4 * Y0 + q
            if gx >= nx0 or gy >= ny0: continue
            c = gx * ny0 + gy; acc = f32(0)
            for d in range(9):
                ex, ey = clamp(gx + d // 3 - 1, nx0), clamp(gy + d % 3 - 1, ny0)
                acc = f32(acc + f32(A0[d * nc0 + c] * z0[chk((ex - fx0) * FH + (ey - fy0), FW * FH)]))
            dg = A0[4 * nc0 + c]; zc = z0[chk((gx - fx0) * FH + (gy - fy0), FW * FH)]
            out0[c] = f32(zc + f32(OM * f32(b0[c] - acc)) / dg) if dg != 0 else f32(0)


def stencil(rng, nx, ny):
    nc = nx * ny
    A = rng.standard_normal(9 * nc).astype(f32) * f32(0.1)
    for c in range(nc):
        cx, cy = divmod(c, ny)
        for d in range(9):
            ex, ey = cx + d // 3 - 1, cy + d % 3 - 1
            if not (0 <= ex < nx and 0 <= ey < ny):
                A[d * nc + c] = 0
        A[4 * nc + c] = f32(1 + rng.random())
    for c in np.nonzero(rng.random(nc) < 0.1)[0]:  # empty cells: zero rows
        for d in range(9):
            A[d * nc + c] = 0
    return A


def same_bits(a, b):
    return np.array_equal(a.view(np.int32), b.view(np.int32))


def check_grid(nx0, ny0, seed):
    rng = np.random.default_rng(seed)
    nx1, ny1 = (nx0 + 1) // 2, (ny0 + 1) // 2
    nx2, ny2 = (nx1 + 1) // 2, (ny1 + 1) // 2
    A0, A1 = stencil(rng, nx0, ny0), stencil(rng, nx1, ny1)
    b0 = rng.standard_normal(nx0 * ny0).astype(f32)
    nan = lambda n: np.full(n, np.nan, f32)
    xa0, b1a, xa1, b2a = nan(nx0 * ny0), nan(nx1 * ny1), nan(nx1 * ny1), nan(nx2 * ny2)
    down(nx0, ny0, A0, b0, xa0, nx1, ny1, b1a)
    down(nx1, ny1, A1, b1a, xa1, nx2, ny2, b2a)
    xb0, b1b, xb1, b2b = nan(nx0 * ny0), nan(nx1 * ny1), nan(nx1 * ny1), nan(nx2 * ny2)
    down2(nx0, ny0, A0, b0, xb0, nx1, ny1, A1, b1b, xb1, nx2, ny2, b2b)
    for a, b, what in ((xa0, xb0, "x0"), (b1a, b1b, "b1"), (xa1, xb1, "x1"), (b2a, b2b, "b2")):
        assert not np.isnan(b).any(), (nx0, ny0, what, "unwritten")
        assert same_bits(a, b), (nx0, ny0, what)
    e2 = rng.standard_normal(nx2 * ny2).astype(f32)
    o1a, o0a = nan(nx1 * ny1), nan(nx0 * ny0)
    up(nx1, ny1, A1, b1a, xa1, ny2, e2, o1a)
    up(nx0, ny0, A0, b0, xa0, ny1, o1a, o0a)
    o1b, o0b = nan(nx1 * ny1), nan(nx0 * ny0)
    up2(nx0, ny0, A0, b0, xa0, nx1, ny1, A1, b1a, xa1, nx2, ny2, e2, o1b, o0b)
    for a, b, what in ((o1a, o1b, "out1"), (o0a, o0b, "out0")):
        assert not np.isnan(b).any(), (nx0, ny0, what, "unwritten")
        assert same_bits(a, b), (nx0, ny0, what)


GRIDS = ((7, 5), (16, 32), (21, 19), (40, 9), (3, 3), (33, 67))


@pytest.mark.parametrize("shape", ["shipped", (4, 8), (1, 16), (8, 4)])
def test_paired_levels_equal_level_by_level(shape):
    global TX, TY
    TX, TY = shipped_tile() if shape == "shipped" else shape
    for k, (nx0, ny0) in enumerate(GRIDS):
        check_grid(nx0, ny0, 3 + k)

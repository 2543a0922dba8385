"""Loop versions without FFTs, for tiny inputs (TEST INFRASTRUCTURE ONLY).

They restate linpatch_convolve (jubik/convolve.py:66-145) as explicit
index arithmetic: every patch window is weighted, placed at offset
``margin`` in an L x L frame, circularly convolved with the cut and
normalised kernel by direct summation, and added onto the canvas, which is
finally cropped.  They share no code with ``oracle.convolve``.
"""
import math


def _tent(w):
    half = w // 2
    if half == 1:
        ramp = [0.0]
    else:
        ramp = [t / (half - 1) for t in range(half)]
    return ramp + ramp[::-1]


def linpatch_loops(x, window, n, margin):
    """x: nested lists (H x W); window: per patch (2dx x 2dx) lists."""
    H, W = len(x), len(x[0])
    dx, dy = H // n, W // n
    wx, wy = 2 * dx, 2 * dy
    hx, hy = dx // 2, dy // 2
    Lx, Ly = wx + 2 * margin, wy + 2 * margin
    bx, by = _tent(wx), _tent(wy)

    def start(i, stride, padded, win):
        return min(i * stride, padded - win)

    out_canvas = [[0.0] * (W + 2 * margin + 2 * hy)
                  for _ in range(H + 2 * margin + 2 * hx)]
    for p in range(n * n):
        ix, iy = p % n, p // n
        sx = start(ix, dx, H + 2 * hx, wx)
        sy = start(iy, dy, W + 2 * hy, wy)
        frame = [[0.0] * Ly for _ in range(Lx)]
        for r in range(wx):
            for c in range(wy):
                i, j = sx + r - hx, sy + c - hy
                v = x[i][j] if (0 <= i < H and 0 <= j < W) else 0.0
                frame[margin + r][margin + c] = bx[r] * by[c] * v
        win = window[p]
        total = sum(sum(row) for row in win)
        taps = []
        for r in range(wx):
            for c in range(wy):
                if win[r][c] != 0.0:
                    taps.append((r - dx, c - dy, win[r][c] / total))
        cx0 = start(ix, dx, H + 2 * margin + 2 * hx, Lx)
        cy0 = start(iy, dy, W + 2 * margin + 2 * hy, Ly)
        for u in range(Lx):
            for v in range(Ly):
                acc = 0.0
                for a, b, k in taps:
                    acc += k * frame[(u - a) % Lx][(v - b) % Ly]
                out_canvas[cx0 + u][cy0 + v] += acc
    off_x, off_y = hx + margin, hy + margin
    return [[out_canvas[off_x + i][off_y + j] for j in range(W)]
            for i in range(H)]


def circular_convolve_loops(k, z):
    """Direct 2-D circular convolution sum_a k[a] z[u - a] (mod L)."""
    Lx, Ly = len(z), len(z[0])
    return [[sum(k[a][b] * z[(u - a) % Lx][(v - b) % Ly]
                 for a in range(Lx) for b in range(Ly))
             for v in range(Ly)] for u in range(Lx)]

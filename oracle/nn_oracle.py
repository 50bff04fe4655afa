"""CPU oracle of the learned coarse-correction network -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of the U-net that underworld2_amd's nn_kernels.hip evaluates and trains:
forward pass, loss, and a hand-derived backward pass.  The backward pass is pinned by central
finite differences of the forward pass (tests/test_nn_oracle.py); the HIP forward, loss and
parameter gradients are checked against this module on the GPU.

Parameter layout (flat float vector): for every convolution in order -- encoder (2 per level),
decoder (levels D-2 .. 0), head -- the weights [co][ci][27] (tap = (dz+1)*9 + (dy+1)*3 + (dx+1))
followed by the bias [co].
"""
import numpy as np


def coarser(n):
    return tuple((v + 1) // 2 for v in n)


def r1(n_f, n_c):
    """1D full-weighting restriction matrix (weights 1 at 2i, 1/2 at 2i+-1) = P1^T."""
    R = np.zeros((n_c, n_f))
    for i in range(n_c):
        for o, w in ((-1, 0.5), (0, 1.0), (1, 0.5)):
            I = 2 * i + o
            if 0 <= I < n_f:
                R[i, I] = w
    return R


def restrict(x, fine, coarse, s=1.0):
    """x[..., nz, ny, nx] on the fine node grid -> coarse grid (s * R x)."""
    Rz, Ry, Rx = r1(fine[2], coarse[2]), r1(fine[1], coarse[1]), r1(fine[0], coarse[0])
    y = np.einsum("kz,...zyx->...kyx", Rz, x)
    y = np.einsum("jy,...kyx->...kjx", Ry, y)
    y = np.einsum("ix,...kjx->...kji", Rx, y)
    return s * y


def prolong(x, fine, coarse, s=1.0):
    """coarse node grid -> fine (s * P x), P = R^T."""
    Rz, Ry, Rx = r1(fine[2], coarse[2]), r1(fine[1], coarse[1]), r1(fine[0], coarse[0])
    y = np.einsum("kz,...kji->...zji", Rz, x)
    y = np.einsum("jy,...zji->...zyi", Ry, y)
    y = np.einsum("ix,...zyi->...zyx", Rx, y)
    return s * y


def _shift(a, dz, dy, dx):
    """b[..., z, y, x] = a[..., z+dz, y+dy, x+dx] with zero padding."""
    out = np.zeros_like(a)
    nz, ny, nx = a.shape[-3:]
    zs, ze = max(0, -dz), min(nz, nz - dz)
    ys, ye = max(0, -dy), min(ny, ny - dy)
    xs, xe = max(0, -dx), min(nx, nx - dx)
    out[..., zs:ze, ys:ye, xs:xe] = a[..., zs + dz:ze + dz, ys + dy:ye + dy, xs + dx:xe + dx]
    return out


TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def conv(x, w, b):
    """x[B][ci][z][y][x], w[co][ci][27], b[co]."""
    out = np.zeros((x.shape[0], w.shape[0]) + x.shape[2:])
    for t, (dz, dy, dx) in enumerate(TAPS):
        out += np.einsum("oc,bczyx->bozyx", w[:, :, t], _shift(x, dz, dy, dx))
    return out + b[None, :, None, None, None]


def conv_bwd(x, w, gout):
    """gradients of conv(x, w, b) given dL/dout: (dx, dw, db)."""
    dw = np.zeros_like(w)
    dx = np.zeros_like(x)
    for t, (dz, dy, dx_) in enumerate(TAPS):
        xs = _shift(x, dz, dy, dx_)
        dw[:, :, t] = np.einsum("bozyx,bczyx->oc", gout, xs)
        # out[p] += w x[p+off]  ->  dx[q] += w^T gout[q-off]
        dx += _shift(np.einsum("oc,bozyx->bczyx", w[:, :, t], gout), -dz, -dy, -dx_)
    db = gout.sum(axis=(0, 2, 3, 4))
    return dx, dw, db


class Net:
    def __init__(self, grid, channels, depth, cin=4, cout=3):
        self.grid = [tuple(grid)]
        for _ in range(1, depth):
            self.grid.append(coarser(self.grid[-1]))
        self.C, self.D, self.cin, self.cout = channels, depth, cin, cout
        self.specs = []
        off = 0
        dims = [(cin, channels), (channels, channels)] + [(channels, channels)] * (2 * (depth - 1))
        dims += [(channels, channels)] * (depth - 1) + [(channels, cout)]
        for ci, co in dims:
            self.specs.append((ci, co, off, off + ci * co * 27))
            off += ci * co * 27 + co
        self.nparams = off

    def unpack(self, p):
        out = []
        for ci, co, wo, bo in self.specs:
            out.append((p[wo:bo].reshape(co, ci, 27), p[bo:bo + co]))
        return out

    def _shape(self, a, l):
        n = self.grid[l]
        return a.reshape(a.shape[0], a.shape[1], n[2], n[1], n[0])

    def forward(self, p, x):
        """x[B][cin][N] -> y[B][cout][N] and the cache for backward."""
        D = self.D
        W = self.unpack(p)
        relu = lambda a: np.maximum(a, 0.0)
        x = self._shape(x.astype(np.float64), 0)
        cache = {"x": x, "act": {}, "pool": {}, "sum": {}}
        inp = x
        for l in range(D):
            if l > 0:
                inp = restrict(cache["act"][2 * l - 1], self.grid[l - 1], self.grid[l], 0.125)
                cache["pool"][l] = inp
            a0 = relu(conv(inp, *W[2 * l]))
            a1 = relu(conv(a0, *W[2 * l + 1]))
            cache["act"][2 * l], cache["act"][2 * l + 1] = a0, a1
        up = cache["act"][2 * D - 1]
        for l in range(D - 2, -1, -1):
            c = 2 * D + (D - 2 - l)
            s = prolong(up, self.grid[l], self.grid[l + 1]) + cache["act"][2 * l + 1]
            cache["sum"][l] = s
            up = relu(conv(s, *W[c]))
            cache["act"][c] = up
        y = conv(up, *W[3 * D - 1])
        return y.reshape(y.shape[0], y.shape[1], -1), cache

    def loss(self, y, t, m):
        """mean over the batch of |m (y - t)|^2 / |m t|^2 and dL/dy."""
        B = y.shape[0]
        num = (m * (y - t) ** 2).reshape(B, -1).sum(1)
        den = (m * t * t).reshape(B, -1).sum(1) + 1e-30
        L = float(np.mean(num / den))
        g = 2.0 * m * (y - t) / den[:, None, None] / B
        return L, g

    def backward(self, p, cache, gy):
        D = self.D
        W = self.unpack(p)
        grads = [None] * len(W)
        act = cache["act"]
        gact = {}
        gy = self._shape(gy, 0)
        # head
        c = 3 * D - 1
        top = 3 * D - 2                        # last decoder output = input of the head
        gin, dw, db = conv_bwd(act[top], W[c][0], gy)
        grads[c] = (dw, db)
        gact[top] = gin
        gsum = {}
        for l in range(0, D - 1):
            c = 2 * D + (D - 2 - l)
            g = gact[c] * (act[c] > 0)
            gin, dw, db = conv_bwd(cache["sum"][l], W[c][0], g)
            grads[c] = (dw, db)
            gsum[l] = gin
            cup = 2 * D - 1 if l == D - 2 else 2 * D + (D - 2 - (l + 1))
            gact[cup] = restrict(gin, self.grid[l], self.grid[l + 1], 1.0)
        gpool = {}
        for l in range(D - 1, -1, -1):
            c1, c0 = 2 * l + 1, 2 * l
            if l <= D - 2:
                gact[c1] = gsum[l] + prolong(gpool[l + 1], self.grid[l], self.grid[l + 1], 0.125)
            g1 = gact[c1] * (act[c1] > 0)
            gin, dw, db = conv_bwd(act[c0], W[c1][0], g1)
            grads[c1] = (dw, db)
            g0 = gin * (act[c0] > 0)
            inp = cache["x"] if l == 0 else cache["pool"][l]
            gin0, dw, db = conv_bwd(inp, W[c0][0], g0)
            grads[c0] = (dw, db)
            if l > 0:
                gpool[l] = gin0
        flat = np.zeros(self.nparams)
        for (ci, co, wo, bo), (dw, db) in zip(self.specs, grads):
            flat[wo:bo] = dw.ravel()
            flat[bo:bo + co] = db
        return flat

    def loss_grad(self, p, x, t, m):
        y, cache = self.forward(p, x)
        L, gy = self.loss(y, t, m)
        return L, self.backward(p, cache, gy), y

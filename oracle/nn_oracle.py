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


# ------------------------------------------------------------------------------------------
# Learned Schur-complement stencil (nn_kernels.hip, SchurNet): 9 features per element (approxS
# log-ratios to the face neighbours, wall flags), L 3x3x3 convolutions with ReLU (9 -> C, then
# C -> C), a 1x1x1 head C -> 27;
# the stencil is identity + head; loss per sample |C r - t|^2 / |t|^2, mean over the batch.
# Parameter layout: per 3x3x3 layer w[C][ci][27] then b[C]; head w[27][C] then b[27].
# ------------------------------------------------------------------------------------------
def snn_features(grid, S):
    """feat[9][N] from approxS[N] on the element grid (nx, ny, nz), x fastest."""
    nx, ny, nz = grid
    t = -np.log(np.asarray(S, dtype=np.float64).reshape(nz, ny, nx))
    out = []
    for dz, dy, dx in ((0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)):
        valid = _shift(np.ones_like(t), dz, dy, dx) > 0
        out.append(np.where(valid, np.tanh(0.5 * (_shift(t, dz, dy, dx) - t)), 0.0))
    for ax, n in ((2, nx), (1, ny), (0, nz)):
        idx = np.arange(n)
        v = np.where(idx == 0, -1.0, np.where(idx == n - 1, 1.0, 0.0))
        shape = [1, 1, 1]
        shape[ax] = n
        out.append(np.broadcast_to(v.reshape(shape), t.shape))
    return np.stack(out, 0).reshape(9, -1)


class StencilNet:
    def __init__(self, grid, channels, layers):
        self.grid = tuple(grid)
        self.C, self.L = channels, layers
        self.specs = []
        off = 0
        for l in range(layers):
            ci = 9 if l == 0 else channels
            self.specs.append((ci, channels, off, off + ci * channels * 27, 27))
            off += ci * channels * 27 + channels
        self.specs.append((channels, 27, off, off + channels * 27, 1))
        off += channels * 27 + 27
        self.nparams = off

    def unpack(self, p):
        return [(p[wo:bo].reshape(co, ci, k) if k == 27 else p[wo:bo].reshape(co, ci), p[bo:bo + co])
                for ci, co, wo, bo, k in self.specs]

    def _shape(self, a):
        nx, ny, nz = self.grid
        return a.reshape(a.shape[0], a.shape[1], nz, ny, nx)

    def forward(self, p, x):
        """x[B][9][N] -> stencil offsets y[B][27][N] and the cache."""
        W = self.unpack(p)
        h = self._shape(np.asarray(x, dtype=np.float64))
        acts = []
        inp = h
        for l in range(self.L):
            a = np.maximum(conv(inp, *W[l]), 0.0)
            acts.append(a)
            inp = a
        wh, bh = W[self.L]
        y = np.einsum("oc,bczyx->bozyx", wh, inp) + bh[None, :, None, None, None]
        return y.reshape(y.shape[0], 27, -1), {"x": h, "acts": acts}

    def stencil(self, y, r):
        """out[b][e] = sum_s (delta_s,13 + y[b][s][e]) r[b][e + off(s)]."""
        yy = self._shape(y)
        rr = self._shape(np.asarray(r, dtype=np.float64)[:, None])[:, 0]
        out = np.zeros_like(rr)
        for s, (dz, dy, dx) in enumerate(TAPS):
            out += ((1.0 if s == 13 else 0.0) + yy[:, s]) * _shift(rr, dz, dy, dx)
        return out.reshape(out.shape[0], -1)

    def loss_grad(self, p, x, r, t):
        """mean over the batch of |C r - t|^2 / |t|^2, its parameter gradient, and C r."""
        W = self.unpack(p)
        y, cache = self.forward(p, x)
        out = self.stencil(y, r)
        t = np.asarray(t, dtype=np.float64)
        B = out.shape[0]
        den = (t * t).sum(1)
        L = float(np.mean(((out - t) ** 2).sum(1) / den))
        gout = 2.0 * (out - t) / den[:, None] / B
        rr = self._shape(np.asarray(r, dtype=np.float64)[:, None])[:, 0]
        go = self._shape(gout[:, None])[:, 0]
        gy = np.stack([go * _shift(rr, dz, dy, dx) for dz, dy, dx in TAPS], 1)  # [B][27][z][y][x]
        grads = [None] * (self.L + 1)
        top = cache["acts"][-1]
        wh, bh = W[self.L]
        grads[self.L] = (np.einsum("bozyx,bczyx->oc", gy, top), gy.sum(axis=(0, 2, 3, 4)))
        g = np.einsum("oc,bozyx->bczyx", wh, gy)
        for l in range(self.L - 1, -1, -1):
            g = g * (cache["acts"][l] > 0)
            inp = cache["x"] if l == 0 else cache["acts"][l - 1]
            gin, dw, db = conv_bwd(inp, W[l][0], g)
            grads[l] = (dw, db)
            g = gin
        flat = np.zeros(self.nparams)
        for (ci, co, wo, bo, k), (dw, db) in zip(self.specs, grads):
            flat[wo:bo] = dw.ravel()
            flat[bo:bo + co] = db
        return L, flat, out

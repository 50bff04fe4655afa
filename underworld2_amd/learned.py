"""Learned coarse-grid correction for the velocity multigrid cycle.

A small 3D convolutional U-net (nn_kernels.hip) acts on the node grid of one multigrid level.
It reads the residual reaching that level (three velocity channels, scaled by their RMS) plus
the level's log-viscosity relative to its geometric mean eta_ref, and returns a velocity
correction in units of rms(residual) / eta_ref.  Two ways to use it
inside the FP32 cycle of the coupled FGMRES solve:

* mode 0 ("replace"): the network stands in for the cycle below its level;
* mode 1 ("add"):     after the cycle below its level, the network predicts the error that
  the cycle left, from the residual it left (the (viscosity, residual, GMG error) triples).

Training data comes from the plain solver itself: ``collect`` records vectors reaching the
network's level during ordinary solves, and the level operator's own FGMRES/multigrid solve
(FP64, tight tolerance) provides the target correction.  Training runs on the GPU through the
library's forward/backward/Adam kernels (``ConvNet.fit``).  The true FP64 residual of the
outer FGMRES checks the result: when 20 iterations fail to reduce it by 1 % the solver drops
the network and continues with plain multigrid (``SolveStats.fallback_used``).

The reference has no learned component; this module has no counterpart there.
"""
import ctypes
import os

import numpy as np

from . import _lib as L
from .function import Function

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PATH = os.path.join(_HERE, "trained", "coarse_correction.npz")


# ------------------------------------------------------------------------------------------
# the network
# ------------------------------------------------------------------------------------------
class ConvNet:
    """The U-net on an (nx, ny, nz) node grid: inputs [B][4][N], outputs [B][3][N]."""

    def __init__(self, nodes, channels=32, depth=4, seed=0, params=None):
        self.nodes = tuple(int(v) for v in nodes)
        self.channels, self.depth = int(channels), int(depth)
        h = ctypes.c_void_p()
        L.check(L.lib().uwg_nn_create(*self.nodes, self.channels, self.depth, int(seed), ctypes.byref(h)), "ConvNet")
        self._h = h
        n = ctypes.c_long()
        L.check(L.lib().uwg_nn_param_count(self._h, ctypes.byref(n)))
        self.nparams = n.value
        if params is not None:
            self.set_params(params)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                L.lib().uwg_nn_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def N(self):
        return self.nodes[0] * self.nodes[1] * self.nodes[2]

    def get_params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        L.check(L.lib().uwg_nn_get_params(self._h, L.fptr(p)), "get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters, got {p.size}")
        L.check(L.lib().uwg_nn_set_params(self._h, L.fptr(p)), "set_params")

    def _f32(self, a, c):
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.ndim != 3 or a.shape[1] != c or a.shape[2] != self.N:
            raise ValueError(f"expected [B][{c}][{self.N}], got {a.shape}")
        return a

    def forward(self, x):
        x = self._f32(x, 4)
        y = np.zeros((x.shape[0], 3, self.N), dtype=np.float32)
        L.check(L.lib().uwg_nn_forward(self._h, x.shape[0], L.fptr(x), L.fptr(y)), "forward")
        return y

    def loss_grad(self, x, t, m):
        x, t, m = self._f32(x, 4), self._f32(t, 3), self._f32(m, 3)
        loss = ctypes.c_float()
        g = np.zeros(self.nparams, dtype=np.float32)
        L.check(L.lib().uwg_nn_loss_grad(self._h, x.shape[0], L.fptr(x), L.fptr(t), L.fptr(m), ctypes.byref(loss),
                                         L.fptr(g)), "loss_grad")
        return loss.value, g

    def set_data(self, x, t, m):
        x, t, m = self._f32(x, 4), self._f32(t, 3), self._f32(m, 3)
        L.check(L.lib().uwg_nn_set_data(self._h, x.shape[0], L.fptr(x), L.fptr(t), L.fptr(m)), "set_data")

    def fit(self, steps, batch=8, lr=1e-3, weight_decay=0.0, seed=0):
        """Adam on minibatches of the uploaded set; returns the per-step batch losses."""
        losses = np.zeros(steps, dtype=np.float32)
        L.check(L.lib().uwg_nn_train(self._h, int(steps), int(batch), float(lr), float(weight_decay), int(seed),
                                     L.fptr(losses)), "fit")
        return losses

    def evaluate(self, batch=8):
        loss = ctypes.c_float()
        L.check(L.lib().uwg_nn_eval(self._h, int(batch), ctypes.byref(loss)), "evaluate")
        return loss.value


# ------------------------------------------------------------------------------------------
# the preconditioner component
# ------------------------------------------------------------------------------------------
class LearnedCoarseCorrection:
    """Trained parameters bound to a multigrid level, identified by its element resolution."""

    MODES = {"replace": 0, "add": 1}

    def __init__(self, params, level_res, channels, depth, mode="add", meta=None):
        self.params = np.ascontiguousarray(params, dtype=np.float32)
        self.level_res = tuple(int(v) for v in level_res)
        self.channels, self.depth = int(channels), int(depth)
        if mode not in self.MODES:
            raise ValueError(f"mode must be one of {sorted(self.MODES)}")
        self.mode = mode
        self.meta = dict(meta or {})

    @property
    def nodes(self):
        return tuple(r + 1 for r in self.level_res)

    def attach(self, stokes, level_res=None):
        """Bind to the multigrid level of resolution ``level_res`` (default: the trained level;
        when the system has no such level, the level as many halvings below the finest grid as
        in training -- the inputs and output are non-dimensional, so the network transfers)."""
        if level_res is None:
            res = levels(stokes)
            if tuple(self.level_res) in res:
                level_res = self.level_res
            else:
                train_fine = int(self.meta.get("res", 0))
                if train_fine <= 0 or train_fine % self.level_res[0]:
                    raise ValueError(f"no multigrid level with resolution {self.level_res}; levels are {res}")
                factor = train_fine // self.level_res[0]
                want = tuple(r // factor for r in res[0])
                if any(r % factor for r in res[0]) or want not in res:
                    raise ValueError(f"no multigrid level {factor}x coarser than {res[0]}; levels are {res}")
                level_res = want
        lev = level_index(stokes, level_res)
        L.check(L.lib().uwg_stokes_load_nn(stokes._h, lev, self.params.size, L.fptr(self.params), self.channels,
                                           self.depth), "load learned correction")
        L.check(L.lib().uwg_stokes_set_nn_mode(stokes._h, self.MODES[self.mode]), "learned correction mode")
        stokes._nn_model = self
        return lev

    def save(self, path):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez_compressed(path, params=self.params, level_res=np.array(self.level_res), channels=self.channels,
                            depth=self.depth, mode=self.mode,
                            meta=np.array([f"{k}={v}" for k, v in sorted(self.meta.items())]))

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            meta = {}
            for item in z["meta"]:
                k, _, v = str(item).partition("=")
                meta[k] = v
            return cls(z["params"], tuple(z["level_res"]), int(z["channels"]), int(z["depth"]), str(z["mode"]),
                       meta)


def load_default(stokes=None):
    """The correction shipped with the package (or None if none has been trained)."""
    if not os.path.exists(DEFAULT_PATH):
        return None
    return LearnedCoarseCorrection.load(DEFAULT_PATH)


def levels(stokes):
    """Element resolution of every multigrid level of an assembled system, finest first."""
    n = ctypes.c_int()
    L.check(L.lib().uwg_stokes_levels(stokes._h, ctypes.byref(n), None))
    res = np.zeros((n.value, 3), dtype=np.int32)
    L.check(L.lib().uwg_stokes_levels(stokes._h, ctypes.byref(n), L.iptr(res)))
    return [tuple(int(v) for v in r) for r in res]


def level_index(stokes, level_res):
    res = levels(stokes)
    want = tuple(int(v) for v in level_res)
    for i, r in enumerate(res):
        if r == want:
            return i
    raise ValueError(f"no multigrid level with resolution {want}; levels are {res}")


# ------------------------------------------------------------------------------------------
# training data
# ------------------------------------------------------------------------------------------
def level_mask(stokes, lev, nnodes):
    bc = np.zeros(nnodes, dtype=np.uint8)
    L.check(L.lib().uwg_stokes_level_mask(stokes._h, lev, bc.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))))
    return np.stack([((bc >> d) & 1) == 0 for d in range(3)]).astype(np.float32)  # [3][N], 1 = free


def level_feature(stokes, lev, nnodes):
    """(log10(eta / eta_ref) per node, eta_ref) of a level."""
    f = np.zeros(nnodes, dtype=np.float32)
    er = ctypes.c_double()
    L.check(L.lib().uwg_stokes_level_feature(stokes._h, lev, L.fptr(f), ctypes.byref(er)))
    return f, er.value


def level_solve(stokes, lev, b, rtol=1e-8):
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    its = ctypes.c_int()
    rc = L.check(L.lib().uwg_stokes_level_solve(stokes._h, lev, L.dptr(b), L.dptr(x), float(rtol), ctypes.byref(its)),
                 "level_solve")
    return x, its.value, rc == 0


def make_sample(b, x, feat, mask, eta_ref=1.0):
    """One training triple from a level vector b (interleaved [N][3]) and the solution x of
    K_level x = b, non-dimensionalised exactly as the inference path does it: inputs b / rms(b),
    targets x * eta_ref / rms(b) (the network's output is multiplied by rms(b) / eta_ref)."""
    N = feat.size
    b = np.asarray(b, dtype=np.float64).reshape(N, 3)
    rms = np.sqrt(np.mean(b * b))
    s = 1.0 / rms if rms > 0 else 0.0
    xin = np.empty((4, N), dtype=np.float32)
    xin[:3] = (b * s).T
    xin[3] = feat
    t = (np.asarray(x, dtype=np.float64).reshape(N, 3) * (s * eta_ref)).T.astype(np.float32)
    return xin, t, mask.copy()


def collect(stokes, solver, level_res, kind, max_samples=16, every=1, target_rtol=1e-8):
    """Run one solve with vector capture on, then turn every captured vector into a training
    triple.  kind 0 records the right-hand side reaching the level (targets for mode 0),
    kind 1 the residual left by the cycle below it (targets for mode 1)."""
    lev = level_index(stokes, level_res)
    nodes = tuple(r + 1 for r in level_res)
    N = nodes[0] * nodes[1] * nodes[2]
    L.check(L.lib().uwg_stokes_capture(stokes._h, lev, int(kind), int(max_samples), int(every)), "capture")
    try:
        solver.solve()
        cnt = ctypes.c_int()
        buf = np.zeros((max_samples, N * 3), dtype=np.float32)
        L.check(L.lib().uwg_stokes_get_captured(stokes._h, L.fptr(buf), ctypes.byref(cnt)), "get_captured")
    finally:
        L.lib().uwg_stokes_capture(stokes._h, lev, 0, 0, 1)
    feat, eta_ref = level_feature(stokes, lev, N)
    mask = level_mask(stokes, lev, N)
    X, T, M = [], [], []
    for k in range(cnt.value):
        b = buf[k].astype(np.float64)
        if not np.any(b):
            continue
        x, _, ok = level_solve(stokes, lev, b, target_rtol)
        if not ok:
            continue
        xi, ti, mi = make_sample(b, x, feat, mask, eta_ref)
        X.append(xi)
        T.append(ti)
        M.append(mi)
    return np.array(X), np.array(T), np.array(M)


# ------------------------------------------------------------------------------------------
# random training problems: exp(Gaussian random field) viscosity, buoyant spheres
# ------------------------------------------------------------------------------------------
def gaussian_random_field(nodes, corr_len, rng):
    """Zero-mean, unit-variance stationary field on an (nx, ny, nz) node grid of the unit cube
    with a Gaussian covariance of the given correlation length (spectral synthesis)."""
    nx, ny, nz = nodes
    # periodic synthesis on a doubled box removes wrap-around correlation across the faces
    sx, sy, sz = 2 * nx, 2 * ny, 2 * nz
    h = 1.0 / (np.array(nodes) - 1)
    kx = 2 * np.pi * np.fft.fftfreq(sx, d=h[0])
    ky = 2 * np.pi * np.fft.fftfreq(sy, d=h[1])
    kz = 2 * np.pi * np.fft.rfftfreq(sz, d=h[2])
    K2 = kz[None, None, :] ** 2 + ky[None, :, None] ** 2 + kx[:, None, None] ** 2
    amp = np.exp(-0.25 * K2 * corr_len ** 2)
    noise = rng.standard_normal((sx, sy, sz))
    f = np.fft.irfftn(np.fft.rfftn(noise) * amp, s=(sx, sy, sz), axes=(0, 1, 2))[:nx, :ny, :nz]
    f = (f - f.mean()) / (f.std() + 1e-300)
    return np.ascontiguousarray(f.transpose(2, 1, 0))  # [z][y][x]


_G3 = np.array([-np.sqrt(0.6), 0.0, np.sqrt(0.6)])


def _gauss_offsets(mesh):
    """Physical offsets of the 27 Gauss points from the element centre [27][3] (q = qx + 3 qy +
    9 qz) and the element centres [E][3] (x fastest)."""
    res = np.array(mesh.elementRes)
    lo, hi = np.array(mesh.minCoord), np.array(mesh.maxCoord)
    h = (hi - lo) / res
    q = np.arange(27)
    off = np.stack([_G3[q % 3], _G3[(q // 3) % 3], _G3[q // 9]], 1) * (0.5 * h)[None, :]
    k, j, i = np.meshgrid(np.arange(res[2]), np.arange(res[1]), np.arange(res[0]), indexing="ij")
    cen = np.stack([lo[0] + (i.ravel() + 0.5) * h[0], lo[1] + (j.ravel() + 0.5) * h[1],
                    lo[2] + (k.ravel() + 0.5) * h[2]], 1)
    return off, cen


class NodalField(Function):
    """A scalar field given by its values on the node grid of the unit box, evaluated by
    trilinear interpolation (a host-side helper for building training problems)."""

    def __init__(self, values_zyx, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
        self.v = np.asarray(values_zyx, dtype=np.float64)
        self.lo, self.hi = np.asarray(lo, float), np.asarray(hi, float)

    def _gauss_values(self, mesh):
        """[E][27] values at the Gauss points of a 3D Q1 mesh whose nodes are this field's nodes:
        the trilinear interpolant at each point is a fixed combination of the element's corners."""
        nz, ny, nx = self.v.shape
        if mesh.dim != 3 or mesh._order != 1 or tuple(mesh.elementRes) != (nx - 1, ny - 1, nz - 1) or \
                not np.allclose(mesh.minCoord, self.lo) or not np.allclose(mesh.maxCoord, self.hi):
            return self._fn_eval(type("C", (), {"coords": _gauss_points_of(mesh)})())[:, 0]
        t = 0.5 * (1.0 + _G3)  # position of the Gauss points in [0, 1] along an element edge
        out = np.empty((nx - 1) * (ny - 1) * (nz - 1) * 27).reshape(-1, 27)
        v = self.v
        for q in range(27):
            tx, ty, tz = t[q % 3], t[(q // 3) % 3], t[q // 9]
            acc = 0.0
            for dz in (0, 1):
                wz = tz if dz else 1.0 - tz
                for dy in (0, 1):
                    wy = ty if dy else 1.0 - ty
                    for dx in (0, 1):
                        wx = tx if dx else 1.0 - tx
                        acc = acc + (wx * wy * wz) * v[dz:nz - 1 + dz, dy:ny - 1 + dy, dx:nx - 1 + dx]
            out[:, q] = acc.reshape(-1)
        return out

    def _fn_eval(self, ctx):
        x = ctx.coords
        nz, ny, nx = self.v.shape
        n = np.array([nx, ny, nz]) - 1
        u = (x - self.lo) / (self.hi - self.lo) * n
        i0 = np.clip(np.floor(u).astype(np.int64), 0, n - 1)
        t = u - i0
        out = np.zeros(x.shape[0])
        for dz in (0, 1):
            wz = t[:, 2] if dz else 1 - t[:, 2]
            for dy in (0, 1):
                wy = t[:, 1] if dy else 1 - t[:, 1]
                for dx in (0, 1):
                    wx = t[:, 0] if dx else 1 - t[:, 0]
                    out += wx * wy * wz * self.v[i0[:, 2] + dz, i0[:, 1] + dy, i0[:, 0] + dx]
        return out[:, None]


def _gauss_points_of(mesh):
    off, cen = _gauss_offsets(mesh)
    return (cen[:, None, :] + off[None, :, :]).reshape(-1, 3)


class _SpheresForce(Function):
    def __init__(self, centres, radii):
        self.c, self.r = np.asarray(centres, float), np.asarray(radii, float)

    def _gauss_values(self, mesh):
        off, cen = _gauss_offsets(mesh)
        out = np.zeros((cen.shape[0], 27, 3))
        for q in range(27):
            x = cen + off[q]
            rho = np.zeros(x.shape[0])
            for c, r in zip(self.c, self.r):
                rho = np.maximum(rho, (np.sum((x - c) ** 2, axis=1) < r * r).astype(float))
            out[:, q, 2] = -rho
        return out

    def _fn_eval(self, ctx):
        x = ctx.coords
        rho = np.zeros(x.shape[0])
        for c, r in zip(self.c, self.r):
            rho = np.maximum(rho, (np.sum((x - c) ** 2, axis=1) < r * r).astype(float))
        return np.stack([np.zeros_like(rho), np.zeros_like(rho), -rho], 1)


def random_problem(res, seed, log10_contrast=None, corr_len=None):
    """A free-slip unit-cube Stokes problem with eta = 10^(c * standardised GRF) (contrast
    10^c, c ~ U(0, 6)) and density 1 in 1-5 random spheres, assembled by Gauss quadrature.
    Returns (stokes, velocity, pressure, description)."""
    from . import mesh as _mesh, conditions as _cond, systems as _sys

    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 6.0) if log10_contrast is None else float(log10_contrast)
    ell = rng.uniform(0.1, 0.3) if corr_len is None else float(corr_len)
    g = gaussian_random_field((res + 1,) * 3, ell, rng)
    g = (g - g.min()) / (g.max() - g.min() + 1e-300)
    eta = NodalField(10.0 ** (c * (g - 0.5)))
    ns = int(rng.integers(1, 6))
    centres = rng.uniform(0.2, 0.8, size=(ns, 3))
    radii = rng.uniform(0.05, 0.15, size=ns)
    m = _mesh.FeMesh_Cartesian("Q1/dQ0", (res,) * 3, (0.0,) * 3, (1.0,) * 3)
    v = _mesh.MeshVariable(m, 3)
    p = _mesh.MeshVariable(m.subMesh, 1)
    walls = tuple(m.specialSets[f"Min{a}_VertexSet"] + m.specialSets[f"Max{a}_VertexSet"] for a in "IJK")
    bc = _cond.DirichletCondition(v, walls)
    st = _sys.Stokes(v, p, fn_viscosity=eta, fn_bodyforce=_SpheresForce(centres, radii), conditions=[bc])
    desc = {"seed": seed, "log10_contrast": c, "corr_len": ell, "spheres": ns}
    return st, v, p, desc


def build_training_set(res, level_res, kind, seeds, samples_per_problem=16, every=2, log=None):
    """Triples from the plain mixed-precision solver on random problems."""
    from . import systems as _sys

    X, T, M = [], [], []
    for s in seeds:
        st, v, p, desc = random_problem(res, s)
        solver = _sys.Solver(st)
        solver.set_method("block")
        solver.set_precision("mixed")
        solver.set_outer_rtol(1e-6)
        x, t, m = collect(st, solver, level_res, kind, samples_per_problem, every)
        if log:
            log(f"problem seed={s} c={desc['log10_contrast']:.2f} ell={desc['corr_len']:.2f} "
                f"its={solver.get_stats().outer_its} samples={len(x)}")
        if len(x):
            X.append(x)
            T.append(t)
            M.append(m)
        del solver, st
    return np.concatenate(X), np.concatenate(T), np.concatenate(M)


# ------------------------------------------------------------------------------------------
# Learned Schur-complement stencil
# ------------------------------------------------------------------------------------------
SCHUR_PATH = os.path.join(_HERE, "trained", "schur_stencil.npz")


class StencilNet:
    """The Schur-stencil network of the library (uwg_snn_*): 9 input features per element,
    ``layers`` 3x3x3 convolutions of ``channels`` channels with ReLU, a 1x1x1 head to the 27
    offsets of the element's stencil (identity + head).  Holds its training set on the device."""

    def __init__(self, channels=16, layers=2, seed=0, params=None):
        self.channels, self.layers = int(channels), int(layers)
        h = ctypes.c_void_p()
        L.check(L.lib().uwg_snn_create(self.channels, self.layers, int(seed), ctypes.byref(h)), "StencilNet")
        self._h = h
        n = ctypes.c_long()
        L.check(L.lib().uwg_snn_param_count(self._h, ctypes.byref(n)))
        self.nparams = n.value
        if params is not None:
            self.set_params(params)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                L.lib().uwg_snn_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        L.check(L.lib().uwg_snn_get_params(self._h, L.fptr(p)), "get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters, got {p.size}")
        L.check(L.lib().uwg_snn_set_params(self._h, L.fptr(p)), "set_params")

    NFEAT = 9

    @staticmethod
    def features(grid, approxS):
        """[9][N] input features of an element grid (nx, ny, nz) from approxS [N]."""
        N = int(np.prod(grid))
        S = np.ascontiguousarray(approxS, dtype=np.float64).reshape(N)
        f = np.zeros((9, N), dtype=np.float32)
        L.check(L.lib().uwg_snn_features(*[int(v) for v in grid], L.dptr(S), L.fptr(f)), "features")
        return f

    def forward(self, grid, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        N = int(np.prod(grid))
        y = np.zeros((x.shape[0], 27, N), dtype=np.float32)
        L.check(L.lib().uwg_snn_forward(self._h, *[int(v) for v in grid], x.shape[0], L.fptr(x), L.fptr(y)),
                "forward")
        return y

    def loss_grad(self, grid, x, r, t):
        x, r, t = (np.ascontiguousarray(a, dtype=np.float32) for a in (x, r, t))
        loss = ctypes.c_float()
        g = np.zeros(self.nparams, dtype=np.float32)
        L.check(L.lib().uwg_snn_loss_grad(self._h, *[int(v) for v in grid], x.shape[0], L.fptr(x), L.fptr(r),
                                          L.fptr(t), ctypes.byref(loss), L.fptr(g)), "loss_grad")
        return loss.value, g

    def add_samples(self, stokes, nsamples=2, seed=0, inner_rtol=1e-6):
        """(features, r^, e^) triples from an assembled 3D Stokes system, stored on the device:
        e^ random, r^ = D^-1/2 S D^-1/2 e^ applied by the plain solver (S = G^T K^-1 G)."""
        stokes._prepare_values()
        L.check(L.lib().uwg_snn_add_samples(self._h, stokes._h, int(nsamples), int(seed), float(inner_rtol)),
                "add_samples")

    @property
    def sample_count(self):
        n = ctypes.c_long()
        L.check(L.lib().uwg_snn_sample_count(self._h, ctypes.byref(n)))
        return n.value

    def clear_data(self):
        L.check(L.lib().uwg_snn_clear_data(self._h))

    def fit(self, steps, batch=4, lr=1e-3, weight_decay=0.0, seed=0, ntrain=None):
        """Adam on minibatches of the first ``ntrain`` stored samples; per-step batch losses."""
        ntrain = self.sample_count if ntrain is None else int(ntrain)
        losses = np.zeros(steps, dtype=np.float32)
        L.check(L.lib().uwg_snn_train(self._h, int(steps), int(batch), float(lr), float(weight_decay), int(seed),
                                      ntrain, L.fptr(losses)), "fit")
        return losses

    def evaluate(self, first, count):
        loss = ctypes.c_float()
        L.check(L.lib().uwg_snn_eval(self._h, int(first), int(count), ctypes.byref(loss)), "evaluate")
        return loss.value


class LearnedSchurStencil:
    """Trained parameters of the Schur-stencil network.  ``attach(stokes)`` makes the coupled
    solver's Schur block use the learned stencil whenever the solve runs with the learned
    option (``Solver.set_learned``); its coefficients are recomputed after every assembly."""

    def __init__(self, params, channels, layers, meta=None):
        self.params = np.ascontiguousarray(params, dtype=np.float32)
        self.channels, self.layers = int(channels), int(layers)
        self.meta = dict(meta or {})

    def attach(self, stokes):
        L.check(L.lib().uwg_stokes_load_schur_nn(stokes._h, self.channels, self.layers, self.params.size,
                                                 L.fptr(self.params)), "load learned Schur stencil")
        stokes._snn_model = self

    def save(self, path):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez_compressed(path, params=self.params, channels=self.channels, layers=self.layers,
                            meta=np.array([f"{k}={v}" for k, v in sorted(self.meta.items())]))

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            meta = {}
            for item in z["meta"]:
                k, _, v = str(item).partition("=")
                meta[k] = v
            return cls(z["params"], int(z["channels"]), int(z["layers"]), meta)


def load_default_schur():
    """The learned Schur stencil shipped with the package (or None if none has been trained)."""
    if not os.path.exists(SCHUR_PATH):
        return None
    return LearnedSchurStencil.load(SCHUR_PATH)


def inclusion_problem(res, seed):
    """A free-slip unit-cube problem with 1-5 spherical inclusions of viscosity 10^c
    (c ~ U(-4, 4) per inclusion) in a background of viscosity 1, density 1 in the inclusions,
    assembled by Gauss quadrature: sharp viscosity interfaces, as in sinker models."""
    from . import mesh as _mesh, conditions as _cond, systems as _sys

    rng = np.random.default_rng(seed)
    ns = int(rng.integers(1, 6))
    centres = rng.uniform(0.2, 0.8, size=(ns, 3))
    radii = rng.uniform(0.06, 0.2, size=ns)
    logeta = rng.uniform(-4.0, 4.0, size=ns)
    m = _mesh.FeMesh_Cartesian("Q1/dQ0", (res,) * 3, (0.0,) * 3, (1.0,) * 3)
    v = _mesh.MeshVariable(m, 3)
    p = _mesh.MeshVariable(m.subMesh, 1)
    walls = tuple(m.specialSets[f"Min{a}_VertexSet"] + m.specialSets[f"Max{a}_VertexSet"] for a in "IJK")
    bc = _cond.DirichletCondition(v, walls)
    st = _sys.Stokes(v, p, fn_viscosity=_SpheresViscosity(centres, radii, logeta),
                     fn_bodyforce=_SpheresForce(centres, radii), conditions=[bc])
    return st, v, p, {"seed": seed, "inclusions": ns}


class _SpheresViscosity(Function):
    def __init__(self, centres, radii, logeta):
        self.c, self.r, self.l = np.asarray(centres, float), np.asarray(radii, float), np.asarray(logeta, float)

    def _gauss_values(self, mesh):
        off, cen = _gauss_offsets(mesh)
        out = np.zeros((cen.shape[0], 27))
        for q in range(27):
            x = cen + off[q]
            le = np.zeros(x.shape[0])
            for c, r, lv in zip(self.c, self.r, self.l):
                le = np.where(np.sum((x - c) ** 2, axis=1) < r * r, lv, le)
            out[:, q] = 10.0 ** le
        return out

    def _fn_eval(self, ctx):
        x = ctx.coords
        le = np.zeros(x.shape[0])
        for c, r, lv in zip(self.c, self.r, self.l):
            le = np.where(np.sum((x - c) ** 2, axis=1) < r * r, lv, le)
        return (10.0 ** le)[:, None]


"""3D parity on grids the K apply covers with several tiles and z-chunks.

The 3D K apply tiles the element grid in 16 x 16 element columns that overlap by one column
(tile (bx, by) covers element columns 15 bx - 1 ... 15 bx + 14) and streams each tile up in
z-chunks of KZ node levels; every node has exactly one owner at the seams.  8^3 grids fit in
one tile and one chunk, so here:

* a ragged 31 x 17 x 23 grid on a non-cubic box (3 x 2 tiles), with the automatic chunk and a
  forced KZ = 3 (8 chunks): K (FP64), G, G^T against the oracle's assembled CSR (1e-12), the
  FP32-weight and the bfloat16-weight cycle applies against the oracle's K built from the same
  rounded weights (FP32 arithmetic: 1e-5 of the largest entry, max norm), and the residual
  kernels (b - K x) in FP64 and in the FP32 cycle;
* eqResiduals against the oracle's |f - K u - G p| and |G^T u| for a fixed random (u, p).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from sinker import K_from_W, build_oracle, build_uw

def _flush(msg):
    print(msg, flush=True)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


RAGGED = (31, 17, 23)
RAGGED_HI = (1.0, 0.6, 0.8)


@pytest.fixture(scope="module")
def ragged():
    _flush(f"ragged: 3D {RAGGED} sinker on [0,1]x[0,0.6]x[0,0.8], 6 ppc, seed 5, contrast 1e3")
    args = dict(eta=(1.0, 1000.0), rho=(0.0, 1.0), lo=(0.0, 0.0, 0.0), hi=RAGGED_HI)
    g = build_uw(RAGGED, 6, 5, (0.45, 0.3, 0.45), 0.2, **args)
    o = build_oracle(RAGGED, 6, 5, (0.45, 0.3, 0.45), 0.2, solve=False, **args)
    _flush("ragged: oracle assembled")
    g["stokes"]._prepare_values()
    return g, o


def _apply(st, which, x, nout=None):
    from underworld2_amd import _lib as L
    y = np.zeros(nout if nout is not None else x.size)
    L.check(L.lib().uwg_stokes_apply(st._h, which, 0, L.dptr(np.ascontiguousarray(x, dtype=np.float64)),
                                     L.dptr(y)), "apply")
    return y


def _weights(st, mesh):
    from underworld2_amd import _lib as L
    W = np.zeros((27, mesh.E))
    fe = np.zeros((mesh.E, 24))
    S = np.zeros(mesh.E)
    L.check(L.lib().uwg_stokes_get_assembly(st._h, 0, L.dptr(W), L.dptr(fe), L.dptr(S)))
    return W


def _bf16(W):
    """FP64 -> FP32 -> bfloat16 (round to nearest, ties to even), widened back."""
    u = W.astype(np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _system_from_W(o, W):
    from oracle import oracle as O
    mesh = o["mesh"]
    h = [(mesh.hi[i] - mesh.lo[i]) / mesh.res[i] for i in range(3)]
    Ke = K_from_W(W, 3, h)
    return O.System(mesh, mesh.freeslip_mask(), Ke, o["fe"], o["Ge"], o["S"])


@pytest.fixture(params=["auto", "kz3"])
def chunk(request):
    from underworld2_amd import _lib as L
    if request.param == "kz3":
        L.lib().uwg_set_apply_chunk(3, 0)
    yield request.param
    L.lib().uwg_set_apply_chunk(-1, 0)


def test_ragged_fp64_applies(ragged, chunk):
    g, o = ragged
    st = g["stokes"]
    sysm = o["system"]
    mesh = o["mesh"]
    _flush(f"ragged FP64 K/G/G^T, chunk {chunk}")
    rng = np.random.default_rng(21)
    x = np.zeros(mesh.N * 3)
    x[sysm.free] = rng.standard_normal(sysm.nfree)
    y = _apply(st, 0, x)
    ref = sysm.K @ x[sysm.free]
    assert np.abs(y[sysm.free] - ref).max() < 1e-12 * np.abs(ref).max()
    # Dirichlet rows are identity rows
    bcd = np.setdiff1d(np.arange(mesh.N * 3), sysm.free)
    assert np.array_equal(y[bcd], x[bcd])
    r = _apply(st, 6, x)
    assert np.abs(r[sysm.free] - (x[sysm.free] - ref)).max() < 1e-12 * np.abs(ref).max()
    pv = rng.standard_normal(mesh.E)
    yg = _apply(st, 1, pv, mesh.N * 3)
    refg = sysm.G @ pv
    assert np.abs(yg[sysm.free] - refg).max() < 1e-12 * np.abs(refg).max()
    q = _apply(st, 2, x, mesh.E)
    refq = sysm.G.T @ x[sysm.free]
    assert np.abs(q - refq).max() < 1e-12 * np.abs(refq).max()


def test_ragged_cycle_applies(ragged, chunk):
    """The FP32 cycle's applies against the oracle's K built from the same rounded weights."""
    from underworld2_amd import systems
    g, o = ragged
    st = g["stokes"]
    mesh = o["mesh"]
    _flush(f"ragged FP32 / bf16 cycle applies, chunk {chunk}")
    W = _weights(st, mesh)
    rng = np.random.default_rng(22)
    x = np.zeros(mesh.N * 3)
    free = o["system"].free
    x[free] = rng.standard_normal(free.size)
    x32 = x.astype(np.float32).astype(np.float64)
    for which, Wr, label in ((5, W.astype(np.float32).astype(np.float64), "fp32 weights"),
                             (4, _bf16(W), "bf16 weights")):
        sysr = _system_from_W(o, Wr)
        ref = sysr.K @ x32[free]
        y = _apply(st, which, x)
        err = np.abs(y[free] - ref).max() / np.abs(ref).max()
        _flush(f"  {label}: max error {err:.2e} of max |Kx|")
        assert err < 1e-5, label
        if which == 4:
            r = _apply(st, 7, x)
            err = np.abs(r[free] - (x32[free] - ref)).max() / np.abs(ref).max()
            assert err < 1e-5, "cycle residual kernel"
    systems.StokesSolver.set_cycle_weights("bf16")


def test_ragged_eq_residuals(ragged):
    """stokes.eqResiduals for a fixed (u, p) against the oracle's |f - K u - G p| and |G^T u|
    (Stokes_SLE.c:204-256 on the free dofs), and |f|."""
    from underworld2_amd import _lib as L
    g, o = ragged
    sysm = o["system"]
    mesh = o["mesh"]
    rng = np.random.default_rng(23)
    u = np.zeros((mesh.N, 3))
    uf = rng.standard_normal(sysm.nfree) * 1e-3
    u.reshape(-1)[sysm.free] = uf
    p = rng.standard_normal(mesh.E) * 1e-2
    g["vel"].data[:] = u
    g["p"].data[:, 0] = p
    r1, r2 = g["stokes"].eqResiduals
    o1, o2 = sysm.residuals(u, p)
    np.testing.assert_allclose(r1, o1, rtol=1e-12)
    np.testing.assert_allclose(r2, o2, rtol=1e-12)
    out = np.zeros(3)
    L.check(L.lib().uwg_stokes_residuals(g["stokes"]._h, L.dptr(out)))
    np.testing.assert_allclose(out[2], np.linalg.norm(sysm.f), rtol=1e-12)


def test_coarse_inverse_on_device_is_exact():
    """The coarsest level's dense inverse, assembled and inverted on the device: on a 6 x 6 x 5
    grid (odd in z, so the hierarchy is one level) one V-cycle is the coarse solve itself, and
    K applied to it returns the right-hand side (free dofs; identity rows elsewhere)."""
    g = build_uw((6, 6, 5), 8, 3, (0.5, 0.5, 0.5), 0.25, eta=(1.0, 1000.0), rho=(0.0, 1.0))
    st = g["stokes"]
    st._prepare_values()
    n = 7 * 7 * 6 * 3
    rng = np.random.default_rng(31)
    mask = np.zeros(n, dtype=bool)
    o = build_oracle((6, 6, 5), 8, 3, (0.5, 0.5, 0.5), 0.25, eta=(1.0, 1000.0), rho=(0.0, 1.0), solve=False)
    mask[o["system"].free] = True
    b = np.where(mask, rng.standard_normal(n), 0.0)
    x = _apply(st, 3, b)
    kx = _apply(st, 0, x)
    err = np.linalg.norm(kx - b) / np.linalg.norm(b)
    _flush(f"coarse inverse: |K V(b) - b| / |b| = {err:.2e}")
    assert err < 1e-10
    # and the direct solve agrees with the oracle's K on the free dofs
    ref = o["system"].K @ x[mask]
    assert np.abs(ref - b[mask]).max() < 1e-9 * np.abs(b).max()

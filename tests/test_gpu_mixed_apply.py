"""The FP32 multigrid cycle's K apply against the FP64 apply (the same operator).

* FP32 weights (``uwg_stokes_apply`` which = 5): FP32 rounding only, relative error < 2e-6;
* bfloat16 weights (which = 4, the cycle's default on 3D levels): each weight is rounded to
  8 significant bits (relative error <= 2^-9), the arithmetic stays FP32: relative error of
  K x below 2e-3 (measured ~2e-4), and measurably above the FP32 instance's (the packed
  weights are in use);
* ``set_cycle_weights("fp32")`` makes the cycle apply equal the FP32-weights apply.
Checked on every multigrid level of a 32^3 problem with a viscosity contrast of 1e3.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _apply(L, st, which, lev, x):
    y = np.zeros_like(x)
    L.check(L.lib().uwg_stokes_apply(st._h, which, lev, L.dptr(x), L.dptr(y)), "apply")
    return y


def test_cycle_apply_bf16_and_fp32_against_fp64():
    import underworld2_amd as uw
    from underworld2_amd import learned, _lib as L
    st, v, p, _ = learned.random_problem(32, 5, log10_contrast=3.0, corr_len=0.2)
    solver = uw.systems.Solver(st)
    solver.set_method("block")
    solver.set_precision("mixed")
    solver.set_outer_rtol(1e-6)
    solver.solve()
    assert solver.get_stats().converged
    import ctypes
    nlev = ctypes.c_int()
    res = np.zeros((16, 3), dtype=np.int32)
    L.check(L.lib().uwg_stokes_levels(st._h, ctypes.byref(nlev), res.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
    rng = np.random.default_rng(0)
    for lev in range(nlev.value):
        n = int(np.prod(res[lev] + 1)) * 3
        x = rng.standard_normal(n)
        y64 = _apply(L, st, 0, lev, x)
        y32 = _apply(L, st, 5, lev, x)
        ybf = _apply(L, st, 4, lev, x)
        nrm = np.linalg.norm(y64)
        e32, ebf = np.linalg.norm(y32 - y64) / nrm, np.linalg.norm(ybf - y64) / nrm
        print(f"level {lev} {tuple(res[lev])}: fp32 weights {e32:.2e}, bf16 weights {ebf:.2e}", flush=True)
        assert e32 < 2e-6
        assert ebf < 2e-3
        assert ebf > 10 * e32
    solver.set_cycle_weights("fp32")
    try:
        solver.solve()  # rebuilds the cycle's FP32 data without the bf16 copy
        x = rng.standard_normal(int(np.prod(res[0] + 1)) * 3)
        assert np.array_equal(_apply(L, st, 4, 0, x), _apply(L, st, 5, 0, x))
    finally:
        solver.set_cycle_weights("bf16")


def test_fp32_krylov_basis_against_fp64_basis():
    """The coupled mixed-precision solve with its Arnoldi basis in FP32 (the default) and in
    FP64: both stop on the FP64 true residual (relative < rtol, recomputed here from the
    returned statistics), and their velocities agree to the size the tolerance admits."""
    import underworld2_amd as uw
    from underworld2_amd import learned
    print("FP32 vs FP64 Krylov basis, 32^3 GRF problem, contrast 1e3", flush=True)
    sols, its = {}, {}
    for fmt in ("fp64", "fp32"):
        st, v, p, _ = learned.random_problem(32, 11, log10_contrast=3.0, corr_len=0.2)
        solver = uw.systems.Solver(st)
        solver.set_method("block")
        solver.set_precision("mixed")
        solver.set_outer_rtol(1e-6)
        solver.set_krylov_basis(fmt)
        try:
            solver.solve()
        finally:
            solver.set_krylov_basis("fp32")
        s = solver.get_stats()
        rel = float(np.hypot(s.res_momentum, s.res_continuity) / s.norm_f)
        assert s.converged and rel < 1e-6, (fmt, rel)
        sols[fmt] = np.array(v.data)
        its[fmt] = s.outer_its
    print(f"iterations: FP64 basis {its['fp64']}, FP32 basis {its['fp32']}", flush=True)
    assert abs(its["fp32"] - its["fp64"]) <= max(3, its["fp64"] // 10)
    d = np.linalg.norm(sols["fp32"] - sols["fp64"]) / np.linalg.norm(sols["fp64"])
    assert d < 1e-4, d

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

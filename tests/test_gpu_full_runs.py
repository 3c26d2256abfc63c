"""Whole runs on the GPU in FP64 against the reference's 4-rank solver, bit for bit: every dt
and the whole final state (through its SHA-256).
- C1, the reference's own CPU-runnable configuration (tests/golden/c1_reference.npz,
  make_golden.py's `c1` job): tanh-interface advection, 128x128 periodic, WENO-JS + LLF, no
  limiter, SSP-RK3, CFL 0.4, to t = 1 (6,919 steps); and its characteristic variant
  (RECONSTRUCTIONTYPE 1; c1_char_reference.npz, the `c1char` job);
- the reference's shipped config/settings.cfg (tests/golden/settings_reference.npz, the
  `settings` job): TESTCASE 0 (SMOOTH) 100x100, WENO-Z, characteristic, LLF, RK3, CFL 0.5,
  from the case's initial state to its final time."""
import hashlib

import numpy as np
import pytest

import wenofv
from conftest import golden

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("recon,fixture", [(0, "c1_reference.npz"), (1, "c1_char_reference.npz")])
def test_c1_bitwise_equal_to_reference(recon, fixture):
    ref = golden(fixture)
    U0, dts_ref = ref["U0"], ref["dts"]
    ny, nx = U0.shape[1:]
    print(f"c1 recon {recon} {nx}x{ny}, {int(ref['steps'])} reference steps", flush=True)
    sol = wenofv.Solver(nx, ny, domain=(0.0, 1.0, 0.0, 1.0), rho10=1.0, rho20=0.1, sound1=10.0, sound2=10.0,
                        preb=0.0, scheme=0, recon=recon, flux=0, bp=0, rk=2, cfl=0.4, bc=0, geometry_ranks=4,
                        precision=0)
    sol.set_state(U0)
    n, t, dts = sol.advance(t_end=1.0, max_dts=len(dts_ref) + 8)
    assert n == int(ref["steps"]) and sol.flags() == 0
    assert t == float(ref["t_final"])
    np.testing.assert_array_equal(dts, dts_ref)
    fin = np.ascontiguousarray(sol.get_state(), dtype="<f8")
    np.testing.assert_allclose(np.abs(fin.reshape(4, -1)).sum(axis=1), ref["final_l1"], rtol=0)
    assert hashlib.sha256(fin.tobytes()).hexdigest() == str(ref["final_sha256"])


def test_shipped_settings_bitwise_equal_to_reference():
    ref = golden("settings_reference.npz")
    kw, t_end = wenofv.reference_case(0)
    assert t_end == float(ref["t_end"])
    U0 = wenofv.reference_initial_state(0, 100, 100, geometry_ranks=4)
    np.testing.assert_array_equal(U0, ref["U0"])  # the reference's own initializeAve
    print(f"settings.cfg 100x100, {int(ref['steps'])} reference steps", flush=True)
    sol = wenofv.Solver(100, 100, scheme=1, recon=1, flux=0, bp=0, rk=2, cfl=0.5, geometry_ranks=4, precision=0,
                        **kw)
    sol.set_state(U0)
    n, t, dts = sol.advance(t_end=t_end, max_dts=len(ref["dts"]) + 8)
    assert n == int(ref["steps"]) and sol.flags() == 0
    assert t == float(ref["t_final"])
    np.testing.assert_array_equal(dts, ref["dts"])
    fin = np.ascontiguousarray(sol.get_state(), dtype="<f8")
    np.testing.assert_allclose(np.abs(fin.reshape(4, -1)).sum(axis=1), ref["final_l1"], rtol=0)
    assert hashlib.sha256(fin.tobytes()).hexdigest() == str(ref["final_sha256"])

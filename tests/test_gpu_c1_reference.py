"""The reference's own CPU-runnable configuration in full, on the GPU in FP64, against the
reference's 4-rank solver (tests/golden/c1_reference.npz, recorded by make_golden.py's `c1`
job): the tanh-interface advection case, 128x128 periodic, WENO-JS + LLF, no limiter, SSP-RK3,
CFL 0.4, run to t = 1 (about 6,900 steps).  Every dt and the whole final state must be
identical bit for bit (the state through its SHA-256)."""
import hashlib

import numpy as np
import pytest

import wenofv
from conftest import golden

pytestmark = pytest.mark.gpu


def test_full_run_bitwise_equal_to_reference():
    ref = golden("c1_reference.npz")
    U0, dts_ref = ref["U0"], ref["dts"]
    ny, nx = U0.shape[1:]
    print(f"c1 {nx}x{ny}, {int(ref['steps'])} reference steps", flush=True)
    sol = wenofv.Solver(nx, ny, domain=(0.0, 1.0, 0.0, 1.0), rho10=1.0, rho20=0.1, sound1=10.0, sound2=10.0,
                        preb=0.0, scheme=0, recon=0, flux=0, bp=0, rk=2, cfl=0.4, bc=0, geometry_ranks=4,
                        precision=0)
    sol.set_state(U0)
    n, t, dts = sol.advance(t_end=1.0, max_dts=len(dts_ref) + 8)
    assert n == int(ref["steps"]) and sol.flags() == 0
    assert t == float(ref["t_final"])
    np.testing.assert_array_equal(dts, dts_ref)
    fin = np.ascontiguousarray(sol.get_state(), dtype="<f8")
    np.testing.assert_allclose(np.abs(fin.reshape(4, -1)).sum(axis=1), ref["final_l1"], rtol=0)
    assert hashlib.sha256(fin.tobytes()).hexdigest() == str(ref["final_sha256"])

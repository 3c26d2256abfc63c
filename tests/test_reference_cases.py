"""Host-side reference test cases: parameters and initial cell averages equal the ones the
reference's own initializeSolver/initializeAve produced (the `_init` states recorded in
tests/golden/solver_runs.npz), bit for bit.  Runs without a GPU."""
import numpy as np
import pytest

import wenofv
from conftest import golden

RUNS = golden("solver_runs.npz")


@pytest.mark.parametrize("name", sorted(k[:-5] for k in RUNS.files if k.endswith("_opts")))
def test_initial_state_matches_reference(name):
    tc, nx, ny = (int(x) for x in RUNS[name + "_opts"][:3])
    U = wenofv.reference_initial_state(tc, nx, ny, geometry_ranks=4)
    np.testing.assert_array_equal(U, RUNS[name + "_init"])


def test_case_parameters():
    kw, t_end = wenofv.reference_case(0)
    assert kw["domain"] == (0.0, 1.0, 0.0, 1.0) and t_end == 0.05 and kw["bc"] == 0
    assert (kw["rho10"], kw["rho20"], kw["sound1"], kw["sound2"]) == (1.2, 1.0, 1.0, 1.0)
    kw, t_end = wenofv.reference_case(3)  # DAMBREAK: walls and gravity (no reference MPI counterpart)
    assert kw["bc"] == 1 and kw["gravity"] == 9.81

"""The WENO coefficient tables of csrc/weno_constants.h (hexfloat doubles).

Two checks per table:
- evaluation: the reference's expressions (include/Utility.hpp:12-39 for the linear weights,
  64-90 for the candidate coefficients) evaluated in IEEE double, operation by operation in
  the reference's order, equal the stored hexfloats bit for bit;
- mathematics: at each Gauss-Lobatto node x in {-1/2, -sqrt(5)/10, sqrt(5)/10, 1/2} (cell
  widths from the centre), candidate k applied to the cell averages of x^m over its three
  cells gives x^m for m <= 2, and sum_k d_k h_k gives x^m for m <= 4 (fifth order), so a
  wrong sign, index or node in either table fails here whatever the expressions say."""
import math
import os
import re

import numpy as np
import pytest

from conftest import REPO

HEADER = os.path.join(REPO, "two-phase-3equation-2d-pal-column_amd", "csrc", "weno_constants.h")
S5 = math.sqrt(5.0)
NODES = (-0.5, -S5 / 10.0, S5 / 10.0, 0.5)


def table(name, shape):
    text = open(HEADER).read()
    body = re.search(name + r"\[[^=]*=\s*\{(.*?)\};", text, re.S).group(1)
    vals = [float.fromhex(t) for t in re.findall(r"-?0x[0-9a-fA-Fp.+-]+", body)]
    return np.array(vals).reshape(shape)


def reference_d():
    """include/Utility.hpp:12-39."""
    return [[0.3, 0.6, 0.1],
            [(91.0 - 9.0 * S5) / 440.0, 129.0 / 220.0, (91.0 + 9.0 * S5) / 440.0],
            [(91.0 + 9.0 * S5) / 440.0, 129.0 / 220.0, (91.0 - 9.0 * S5) / 440.0],
            [0.1, 0.6, 0.3]]


def reference_h():
    """include/Utility.hpp:64-90: h_k's coefficients of its three cells."""
    return [
        [[-1.0 / 6.0, 5.0 / 6.0, 1.0 / 3.0], [1.0 / 3.0, 5.0 / 6.0, -1.0 / 6], [11.0 / 6.0, -7.0 / 6.0, 1.0 / 3.0]],
        [[-1.0 / 60.0 - S5 / 20.0, 1.0 / 30.0 + S5 / 5.0, 59.0 / 60.0 - 3 * S5 / 20.0],
         [-1.0 / 60.0 + S5 / 20.0, 31.0 / 30.0, -1.0 / 60 - S5 / 20.0],
         [59.0 / 60.0 + 3 * S5 / 20, 1.0 / 30.0 - S5 / 5.0, -1.0 / 60.0 + S5 / 20.0]],
        [[-1.0 / 60.0 + S5 / 20.0, 1.0 / 30.0 - S5 / 5.0, 59.0 / 60.0 + 3 * S5 / 20],
         [-1.0 / 60 - S5 / 20.0, 31.0 / 30.0, -1.0 / 60.0 + S5 / 20.0],
         [59.0 / 60.0 - 3 * S5 / 20.0, 1.0 / 30.0 + S5 / 5.0, -1.0 / 60.0 - S5 / 20.0]],
        [[1.0 / 3.0, -7.0 / 6.0, 11.0 / 6.0], [-1.0 / 6, 5.0 / 6.0, 1.0 / 3.0], [1.0 / 3.0, 5.0 / 6.0, -1.0 / 6.0]],
    ]


def cell_average(m, j):
    """Average of x^m over the unit cell centred at j."""
    return ((j + 0.5) ** (m + 1) - (j - 0.5) ** (m + 1)) / (m + 1)


def test_tables_equal_the_reference_expressions_bitwise():
    D, H = table("WENO_D", (4, 3)), table("WENO_H", (4, 3, 3))
    assert D.tobytes() == np.array(reference_d()).tobytes()
    assert H.tobytes() == np.array(reference_h()).tobytes()


@pytest.mark.parametrize("gp", range(4))
def test_candidates_are_exact_for_quadratics(gp):
    H = table("WENO_H", (4, 3, 3))
    x = NODES[gp]
    for k in range(3):
        cells = (k - 2, k - 1, k)  # candidate k's stencil {s[k], s[k+1], s[k+2]}
        for m in range(3):
            val = sum(H[gp, k, i] * cell_average(m, cells[i]) for i in range(3))
            assert val == pytest.approx(x ** m, abs=1e-14), (gp, k, m)
        # a cubic is not reproduced: the candidates are exactly second order
        cub = sum(H[gp, k, i] * cell_average(3, cells[i]) for i in range(3))
        assert abs(cub - x ** 3) > 1e-3


@pytest.mark.parametrize("gp", range(4))
def test_linear_weights_give_the_fifth_order_reconstruction(gp):
    D, H = table("WENO_D", (4, 3)), table("WENO_H", (4, 3, 3))
    x = NODES[gp]
    assert D[gp].sum() == pytest.approx(1.0, abs=1e-15) and (D[gp] > 0).all()
    for m in range(5):
        val = sum(D[gp, k] * sum(H[gp, k, i] * cell_average(m, (k - 2, k - 1, k)[i]) for i in range(3))
                  for k in range(3))
        assert val == pytest.approx(x ** m, abs=1e-13), (gp, m)
    quint = sum(D[gp, k] * sum(H[gp, k, i] * cell_average(5, (k - 2, k - 1, k)[i]) for i in range(3))
                for k in range(3))
    assert abs(quint - x ** 5) > 1e-4  # and not sixth order

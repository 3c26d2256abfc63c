"""End-to-end pin of the oracle on the reference's own recorded accuracy norms (SMOOTH case,
initial state from the host restatement of initializeAve, which matches the reference bit
for bit): the norms agree to 1e-12 relative."""
import numpy as np
import pytest

import oracle as orc
import wenofv
from known_answers import KNOWN, smooth_accuracy


@pytest.mark.parametrize("row", [k for k in KNOWN if k[0] <= 100], ids=lambda k: f"{k[0]}-s{k[1]}r{k[2]}f{k[3]}b{k[4]}")
def test_oracle_reproduces_reference_accuracy(row):
    n, scheme, recon, flux, bp, cfl, linf, l1, l2 = row
    U0 = wenofv.reference_initial_state(0, n, n, geometry_ranks=4)
    o = orc.Oracle(n, n, scheme=scheme, recon=recon, flux=flux, bp=bp, rk=2, cfl=cfl, rho10=1.2, rho20=1.0,
                   sound1=1.0, sound2=1.0)
    U, dts, t, err = o.run(U0, t_end=0.05)
    assert err == 0 and abs(t - 0.05) <= 1e-10
    got = smooth_accuracy(U, n, t)
    np.testing.assert_allclose(got, (linf, l1, l2), rtol=1e-12)

"""Parameter calibration on the GPU (the reference plans these helpers, README.md:80-87): the
chosen parameters meet the requested accuracy against the resolved FCM operator, on the
calibration sample and on a second suspension it has not seen, and a tighter tolerance is
never predicted cheaper."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_calibration_meets_its_tolerance():
    from ForceCouplingMethod import calibration as cal
    from ForceCouplingMethod.suspensions import smooth5_at_least

    model = cal.fit_cost_model()
    L, N = 60.0, 5000  # phi ~ 0.10
    loose = cal.calibrate(L, N, tol=1e-3, cost_model=model)
    tight = cal.calibrate(L, N, tol=1e-4, cost_model=model)
    for tol, r in ((1e-3, loose), (1e-4, tight)):
        cfg = r["config"]
        print(f"tol {tol:g}: ratio {cfg['kernel_widths_ratio']}, Sigma/h {r['sigma_per_h']}, M_G {cfg['M_G']}, "
              f"R_c/Sigma {r['cutoff_per_Sigma']}, grid {cfg['num_grid_points'][0]}, "
              f"predicted {r['predicted_ms']:.3f} ms, sample error {r['sample_error_measured']:.2e}")
        assert r["sample_error_measured"] <= tol
        M = cfg["num_grid_points"][0]
        assert M == smooth5_at_least(M)
    assert tight["predicted_ms"] >= loose["predicted_ms"]
    # a second suspension: the loose choice against its own resolved reference
    c = loose["config"]
    Ls = 30.0
    Y, F = cal._suspension(int(round(0.1 * Ls ** 3 / (4 / 3 * math.pi))), Ls, 0xBEE)
    Sigma = c["kernel_widths_ratio"] * cal.SIGMA_UNIT
    M = smooth5_at_least(math.ceil(Ls * loose["sigma_per_h"] / Sigma))
    V, _ = cal._apply(Y, F, Ls, c["kernel_widths_ratio"], M, c["M_G"], c["R_c"])
    Mr, MGr = cal.reference_grid(Ls)
    Vr, _ = cal._apply(Y, F, Ls, 1.0, Mr, MGr, 2.0, T=np.float64)
    e = cal._rel(V, Vr)
    print(f"unseen suspension: rel L2 {e:.2e}")
    assert e <= 3e-3

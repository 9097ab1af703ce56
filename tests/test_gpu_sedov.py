"""Dynamic block AMR Sedov blast on the GPU (leaves advanced by the batched
kernel between regrids), FP64 and FP32, with the checks of the reference's
Sedov test (test/test_sedov.jl): conservation in the closed box, positivity,
four-fold symmetry of the quadrant masses, and the self-similar growth
R ∝ t^(1/2).

The field itself is symmetric to ~1e-5 relative, not to round-off: the
representative fine neighbour used for coarse-side slopes across a coarse-fine
face is the one at the face's lo corner, which a mirror maps to the hi corner."""
import numpy as np
import pytest

from conftest import load_vespa

V = load_vespa()
from vespa_jl_amd import amr  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sedov_dynamic_amr(dtype):
    print(f"sedov {np.dtype(dtype).name}", flush=True)
    r1 = amr.run_sedov(tfinal=0.02, dtype=dtype)
    r2 = amr.run_sedov(tfinal=0.04, dtype=dtype)
    for r, tf in ((r1, 0.02), (r2, 0.04)):
        assert r["t"] == pytest.approx(tf, abs=1e-14)
        rel = 1e-12 if dtype == np.float64 else 1e-5
        assert r["totals"][0] == pytest.approx(r["totals0"][0], rel=rel)
        assert r["totals"][4] == pytest.approx(r["totals0"][4], rel=rel)
        U = r["U"].astype(np.float64)
        assert (U[0] > 0).all()
        assert {b[0] for b in r["blocks"]} == {0, 1, 2, 3}
        qm = amr.quadrant_masses(r["qb"], U[0])
        assert qm.max() - qm.min() < 1e-6 * qm.max()
    rho = r2["qb"].raster(r2["U"][:1].astype(np.float64))[0]
    sc = rho.max()
    for img in (rho[::-1, :], rho[:, ::-1], rho.T):
        assert np.abs(rho - img).max() / sc < 1e-4
    R1 = amr.shock_radius_mean(r1["qb"], r1["U"][0].astype(np.float64))
    R2 = amr.shock_radius_mean(r2["qb"], r2["U"][0].astype(np.float64))
    ratio, expected = R2 / R1, np.sqrt(2.0)
    assert abs(ratio / expected - 1.0) < 0.05, (R1, R2, ratio)

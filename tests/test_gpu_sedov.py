"""Dynamic block AMR Sedov blast on the GPU (leaves advanced by the batched
kernel between regrids): conservation in the closed box, four-fold symmetry
and the self-similar growth of the shock radius, FP64 and FP32."""
import numpy as np
import pytest

from conftest import load_vespa

V = load_vespa()
from vespa_jl_amd import amr  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sedov_dynamic_amr(dtype):
    print(f"sedov {np.dtype(dtype).name}", flush=True)
    times = (0.01, 0.02)
    r = amr.run_sedov(tfinal=0.02, dtype=dtype, snapshots=times)
    assert r["t"] == pytest.approx(0.02, abs=1e-14)
    rel = 1e-12 if dtype == np.float64 else 1e-5
    assert r["totals"][0] == pytest.approx(r["totals0"][0], rel=rel)
    assert r["totals"][4] == pytest.approx(r["totals0"][4], rel=rel)
    assert {b[0] for b in r["blocks"]} == {0, 1, 2, 3}
    rho = r["snapshots"][0.02][1]
    scale = rho.max()
    tol = 1e-9 if dtype == np.float64 else 1e-4
    assert np.abs(rho - rho[::-1, :]).max() / scale < tol
    assert np.abs(rho - rho[:, ::-1]).max() / scale < tol
    assert np.abs(rho - rho.T).max() / scale < tol
    R1 = amr.shock_radius(r["snapshots"][0.01][1])
    R2 = amr.shock_radius(rho)
    expo = np.log(R2 / R1) / np.log(2.0)
    assert 0.4 < expo < 0.6, (R1, R2, expo)

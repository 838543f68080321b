"""End-to-end learned warm start at a small scale: batched relaxation of
rotating dipoles, U-Net training by the project's kernels, evaluation on
held-out light-cylinder radii against the dipole start."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from coffeegpu_amd.warmstart import experiment, settle_step  # noqa: E402


def test_settle_step():
    h = [(10, 2.0), (20, 1.2), (30, 1.005), (40, 0.996), (50, 1.0)]
    assert settle_step(h, 1.0, 0.01) == 30
    assert settle_step([(10, 2.0), (20, 1.5)], 1.0, 0.01) is None


def test_small_warm_start_experiment():
    out = experiment(nr=64, nth=32, n_train=32, n_test=4, periods=3.0, rlc_lo=3.0, rlc_hi=6.0,
                     levels=3, base=16, epochs=40, batch=8, lr=3e-3, seed=1, eval_periods=5.0)
    print({k: v for k, v in out.items() if k != "runs"})
    for r in out["runs"]:
        print(r)
    assert out["loss_last"] < 0.5 * out["loss_first"]
    assert out["median_field_err_warm"] < out["median_field_err_dipole"]
    assert all(np.isfinite(r["L_ref"]) and r["L_ref"] > 0 for r in out["runs"])

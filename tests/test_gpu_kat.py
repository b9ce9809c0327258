"""The hand-derived known answers of tests/kat_cases.py on the GPU
(pg_propagate through the C ABI): the kernels' update rule, boundary
predicate and post-pass are checked against the values worked out from the
reference's rules, not only against the oracle.  One-sweep cases run with
tuning.max_iterations = 1 (the call then reports the cap, with the state of
that sweep written back)."""
import numpy as np
import pytest

import kat_cases

pytestmark = pytest.mark.gpu

CASES = kat_cases.all_cases()


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_known_answer_on_gpu(case, mode):
    import torch
    import py_propag as pp

    W, H = case["W"], case["H"]
    n = W * H
    dev = torch.device("cuda")
    f = [torch.from_numpy(case[k].copy()).to(dev) for k in ("speed", "az", "ecc")]
    t = torch.from_numpy(case["time"].copy()).to(dev)
    rx = torch.from_numpy(case["refs_x"].view(np.int16).copy()).to(dev)
    ry = torch.from_numpy(case["refs_y"].view(np.int16).copy()).to(dev)
    b = torch.zeros((n + 31) // 32, dtype=torch.int32, device=dev)
    g = pp.GeoReference(W, H, "+proj=utm +zone=30", [0.0, 1.0, 0.0, 0.0, 0.0, 1.0])
    tuning = {"mode": mode, "max_iterations": case["sweeps"]}
    try:
        pp.propagate_device(pp.Settings(g, float(case["max_time"])), *f, t, rx, ry, b, tuning=tuning)
    except pp.PropagNotConverged:
        assert case["sweeps"] > 0
    bad = kat_cases.mismatches(case, t.cpu().numpy(), rx.cpu().numpy().view(np.uint16),
                               ry.cpu().numpy().view(np.uint16), b.cpu().numpy().view(np.uint32))
    assert bad == []

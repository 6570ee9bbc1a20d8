"""Pins the float64 torch reimplementation used as the gradient oracle: its
forward must agree with the CPU oracle's periodogram on the same input."""
import numpy as np
import torch

from tests import fp64_periodogram as R


def test_fp64_forward_matches_oracle(oracle):
    rng = np.random.default_rng(0)
    n, tsamp = 2 ** 15, 256e-6
    x = rng.standard_normal(n).astype(np.float32)
    y = oracle.normalise(x)
    w = R.generate_width_trials(240)
    P, F, S = oracle.oracle().periodogram(y, tsamp, w, 0.5, 2.0, 240, 260)
    rows = R.periodogram_rows(torch.tensor(x, dtype=torch.float64), tsamp, w, 0.5, 2.0, 240, 260,
                              keep=lambda t: t["bins"] in (240, 251) and t["f"] < 12)
    assert rows
    for r0, s in rows.items():
        np.testing.assert_allclose(s.numpy(), S[r0:r0 + s.shape[0]], rtol=1e-5, atol=1e-4)


def test_fp64_plan_matches_oracle(oracle):
    trs = R.plan(2 ** 20, 256e-6, 0.5, 2.0, 240, 260)
    assert sum(t["rows_eval"] for t in trs) == oracle.oracle().periodogram_length(
        2 ** 20, 256e-6, 0.5, 2.0, 240, 260)


def test_fp64_downsample_matches_oracle(oracle):
    x = np.random.default_rng(1).standard_normal(10007).astype(np.float32)
    for f in (1.5, 6.51, 13.777):
        got = R.downsample(torch.tensor(x, dtype=torch.float64), f).numpy()
        np.testing.assert_allclose(got, oracle.oracle().downsample(x, f), rtol=0, atol=1e-5)


def test_fp64_ffa_matches_oracle(oracle):
    x = np.random.default_rng(2).integers(-9, 9, size=(37, 23)).astype(np.float32)
    got = R.ffa(torch.tensor(x, dtype=torch.float64)).numpy()
    assert np.array_equal(got, oracle.oracle().ffa2(x))

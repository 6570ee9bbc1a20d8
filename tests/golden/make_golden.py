"""Writes the golden fixtures of this directory from the compiled reference.

The reference's own C++ path (its headers under src/riptide/cpp, compiled by
``oracle/Makefile`` target ``ref`` into ``oracle/_ref/libriptide_ref.so``)
computes every expected output; inputs are seeded.  Run from the repository
root where that library is built:

    make -C oracle ref && python tests/golden/make_golden.py

Fixtures are data only (inputs and the reference's outputs), a few hundred KB.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libriptide_ref.so is not built (make -C oracle ref)")
    R = O.ref()
    rng = np.random.default_rng(20261203)
    fx = {}
    # ffa2: row counts across the merge-tree cases (powers of two, odd, 2^k +- 1)
    for i, (m, b) in enumerate([(8, 8), (37, 60), (150, 97), (257, 33), (513, 17)]):
        x = rng.standard_normal((m, b)).astype(np.float32)
        fx[f"ffa2_{i}_in"] = x
        fx[f"ffa2_{i}_out"] = R.ffa2(x)
    # snr2 (float64 prefix sums inside the reference)
    x = (rng.standard_normal((40, 250)) * 3).astype(np.float32)
    w = O.generate_width_trials(240).astype(np.uint64)
    fx["snr2_in"], fx["snr2_widths"] = x, w
    fx["snr2_out"] = R.snr2(x, w, 1.3)
    # downsample at integer and fractional factors
    x = rng.standard_normal(20011).astype(np.float32)
    fx["downsample_in"] = x
    for i, f in enumerate([2.0, 6.510416666666667, 13.777, 250.5]):
        fx[f"downsample_{i}_f"] = np.float64(f)
        fx[f"downsample_{i}_out"] = R.downsample(x, f)
    # running median (edge-padded window)
    x = rng.standard_normal(3001).astype(np.float32)
    fx["rmed_in"] = x
    for i, wd in enumerate([5, 101]):
        fx[f"rmed_{i}_width"] = np.int64(wd)
        fx[f"rmed_{i}_out"] = R.running_median(x, wd)
    # a whole periodogram of a normalised series with a pulse
    n, tsamp = 2 ** 13, 2e-3
    t = np.arange(n) * tsamp
    ph = (t / 0.7031) % 1.0
    sig = np.exp(-0.5 * ((ph - 0.5) / (0.03 / 2.355)) ** 2)
    x = (rng.standard_normal(n) + 15 * sig / np.sqrt((sig ** 2).sum())).astype(np.float32)
    x = O.normalise(x)
    w = O.generate_width_trials(240).astype(np.uint64)
    P, F, S = R.periodogram(x, tsamp, w, 0.5, 2.0, 240, 260)
    fx.update(pgram_in=x, pgram_tsamp=np.float64(tsamp), pgram_widths=w,
              pgram_args=np.array([0.5, 2.0, 240, 260], np.float64),
              pgram_periods=P, pgram_foldbins=F, pgram_snrs=S)
    path = os.path.join(HERE, "reference_outputs.npz")
    np.savez_compressed(path, **fx)
    print(f"wrote {path} ({os.path.getsize(path) / 1e3:.0f} kB, {len(fx)} arrays)")


if __name__ == "__main__":
    main()

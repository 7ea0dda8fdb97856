"""Writes tests/golden/corners.npz from the compiled reference.

Run in a container that holds the reference tree, after `make -C oracle ref`:

    python tests/golden/make_golden_corners.py

The inputs are the seeded corner cases of tests/test_oracle_corners.py
(regenerated there from tests/datagen.py, so only results are stored): the
infinity path of rdsum / rddot / rssum / rsdot at folds 2-4, an accumulate
into a top-index caller value, nrm2 over values spanning 2^-1000..2^1000 at
folds 2..52 (results and binned states), and dbdssq into an empty value that
carries a scale. Results come from the reference's scalar build
(libreproblas_ref_sisd.so); the tests also compare its AVX build.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import datagen as dg  # noqa: E402
import test_oracle_corners as tc  # noqa: E402
from cpu_libs import reference, reference_s  # noqa: E402


def main():
    R, RS = reference("sisd"), reference_s("sisd")
    if R is None or RS is None:
        raise SystemExit("oracle/_ref is not built: run `make -C oracle ref` where the reference tree is present")
    out = {}
    for key, op, dt, kind, n, seed, inc in tc.CASES:
        x, y = tc.make_input(op, dt, kind, n, seed, inc)
        lib = R if dt == np.float64 else RS
        for fold in (2, 3, 4):
            out[f"{key}_f{fold}"] = np.array([tc.run(lib, op, dt, fold, x, y, inc)], dtype=dt)
    for fold in (2, 3, 4):
        Y = np.zeros(2 * fold)
        R.dbdsum(fold, np.array([1.6e308, 1.0]), Y)
        x = dg.uniform(5000, 9)
        x[3000], x[3001] = 1.0e308, -np.inf
        R.dbdsum(fold, x, Y)
        out[f"acc_top_f{fold}"] = Y[:1].copy()
    for seed, fold in tc.nrm2_range_cases():
        x = tc.nrm2_range_input(seed)
        out[f"nrm2_{seed}_f{fold}"] = np.array([R.rdnrm2(fold, x)])
        Y = np.zeros(2 * fold)
        R.dbdssq(fold, x, 0.0, Y)
        out[f"ssq_{seed}_f{fold}"] = Y
    for fold in (2, 3, 4, 7):
        for j, (x, sy) in enumerate(((np.array([1.0, 2.0, 3.0]), R.dscale(1e60)),
                                     (dg.wide(3000, 530), R.dscale(1e-3)),
                                     (np.zeros(5), R.dscale(1e10)))):
            Y = np.zeros(2 * fold)
            out[f"ssq0_{j}_f{fold}"] = np.array([R.dbdssq(fold, x, sy, Y)])
            out[f"ssq0_{j}_state_f{fold}"] = Y
    path = os.path.join(HERE, "corners.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(out)} arrays")


if __name__ == "__main__":
    main()

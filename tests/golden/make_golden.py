"""Write the golden fixtures of the flat EZ path from the reference's own CPU build.

Runs in the build container only (it needs oracle/_ref, compiled from the
reference sources by `make -C oracle ref`).  For each case in cases.py and
each Scalar type it stores, in tests/golden/flat_<case>_<dtype>.npz:
  - sha256 of the initial state (the tests regenerate it from cases.py and
    check the hash, so the fixture pins the same input);
  - for every step 1..nsteps: interior sum of squares of the 7 fields;
  - at the DUMPS steps: the 7 field values at 512 seeded sample cells
    (interior and guard cells alike).
The 128^3 x 100-step wave (flat_c1_norms_f64.npz) stores the per-step norms
and 512 sampled cells after step 100.  flat_diag_<case>_<dtype>.npz (cases
DIAG_CASES) store the reference's per-cell energy bookkeeping (dU_KO,
dU_Epar, dU_EgtB and their _cum arrays, field_solver_EZ.cpp:447-453,588-623)
after steps 1 and DIAG_STEPS at 512 sampled interior cells.

downsample_<dtype>.npz hold the reference build's field-output sampling
(multi_array::downsample) of seeded arrays ("downsample" case name).

Usage: python tests/golden/make_golden.py [case ...]   (no case: every case and C1)
"""

from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from tests.golden.cases import CASES, DUMPS, case_wave  # noqa: E402

NSAMPLE = 512
DIAG_CASES = ("noise", "pulsar")
DIAG_DUMPS = (1, 3)
DIAG_NAMES = ("dU_KO", "dU_Epar", "dU_EgtB", "dU_KO_cum", "dU_Epar_cum", "dU_EgtB_cum")


def sample_index(shape, seed=12345):
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    return np.sort(rng.choice(n, size=min(NSAMPLE, n), replace=False))


def interior_sample_index(g, seed=54321):
    k, j, i = np.meshgrid(*[np.arange(gd, d - gd) for gd, d in
                            zip(g.guard[::-1], g.dims[::-1])], indexing="ij")
    flat = (i + (j + k * g.dims[1]) * g.dims[0]).ravel()
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(flat, size=min(NSAMPLE, flat.size), replace=False))


def diag_fixtures():
    from coffeegpu_amd.params import Grid

    for name in DIAG_CASES:
        for dtype in (np.float64, np.float32):
            sp, U0 = CASES[name](dtype)
            out, _, _ = O.run_reference(sp, U0, max(DIAG_DUMPS), dumps=DIAG_DUMPS)
            idx = interior_sample_index(Grid.from_params(sp, dtype))
            rec = {"sha256": np.array(digest(U0)), "index": idx, "dumps": np.array(DIAG_DUMPS),
                   "names": np.array(DIAG_NAMES)}
            for s in DIAG_DUMPS:
                rec[f"s{s}"] = out[("dg", s)].reshape(6, -1)[:, idx]
            path = os.path.join(HERE, f"flat_diag_{name}_{np.dtype(dtype).name}.npz")
            np.savez_compressed(path, **rec)
            print("wrote", path)


# every interior extent above the largest factor: where 1 < N <= d the
# reference steps beyond the array (multi_array.cpp:62-66)
DS_SHAPES = [((14, 12, 10), (3, 3, 3)), ((17, 11, 13), (3, 2, 3))]
DS_STAGGERS = [0b110, 0b101, 0b011, 0b001, 0b010, 0b100, 0b111]
DS_FACTORS = [1, 2, 3]


def downsample_fixtures():
    """downsample_<dtype>.npz: the reference build's multi_array::downsample
    (oracle/_ref/ref_downsample_*) of seeded arrays, for every E/B stagger
    and the cell centre, d = 1, 2, 3."""
    for dtype in (np.float64, np.float32):
        rng = np.random.default_rng(2024)
        rec = {}
        for s, (dims, guard) in enumerate(DS_SHAPES):
            a = rng.standard_normal((dims[2], dims[1], dims[0])).astype(dtype)
            rec[f"in{s}"] = a
            rec[f"guard{s}"] = np.array(guard)
            for st in DS_STAGGERS:
                for d in DS_FACTORS:
                    rec[f"out{s}_{st}_{d}"] = O.reference_downsample(a, guard, st, d)
        path = os.path.join(HERE, f"downsample_{np.dtype(dtype).name}.npz")
        np.savez_compressed(path, **rec)
        print("wrote", path)


def digest(U: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(U).tobytes()).hexdigest()


def main(names=None):
    if names and "diag" in names:
        diag_fixtures()
    if names and "downsample" in names:
        downsample_fixtures()
    for name, fn in CASES.items():
        if names and name not in names:
            continue
        for dtype in (np.float64, np.float32):
            sp, U0 = fn(dtype)
            nsteps = max(DUMPS)
            out, nrm, _ = O.run_reference(sp, U0, nsteps, dumps=DUMPS, norms=True)
            idx = sample_index(U0.shape[1:])
            rec = {"sha256": np.array(digest(U0)), "norms": nrm, "index": idx,
                   "dumps": np.array(DUMPS)}
            for s in DUMPS:
                rec[f"s{s}"] = out[s].reshape(7, -1)[:, idx]
            path = os.path.join(HERE, f"flat_{name}_{np.dtype(dtype).name}.npz")
            np.savez_compressed(path, **rec)
            print("wrote", path)
    if names:
        return
    # C1 at full size: 128^3, FP64, 100 steps
    sp, U0 = case_wave(np.float64, n=(128, 128, 128), seed=1)
    out, nrm, _ = O.run_reference(sp, U0, 100, dumps=(100,), norms=True)
    idx = sample_index(U0.shape[1:])
    path = os.path.join(HERE, "flat_c1_norms_f64.npz")
    np.savez_compressed(path, sha256=np.array(digest(U0)), norms=nrm, index=idx,
                        s100=out[100].reshape(7, -1)[:, idx])
    print("wrote", path)


if __name__ == "__main__":
    main(sys.argv[1:])

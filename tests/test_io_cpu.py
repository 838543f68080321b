"""Snapshot / field-output layer on the CPU.

- libcoffee_h5.so round trips arrays and scalars and writes the reference's
  layout: one dataset per array with the extent reversed (Fortran-order
  extent {x, y, z} -> dataset [z][y][x], hdf_wrapper_impl.hpp:29-35), scalar
  datasets for step (uint32) and time (Scalar); reads check rank, extent and
  element type.  h5dump (the HDF5 tools beside the library) reads the file
  independently.
- The oracle's field-output sampling (oracle.downsample) equals the reference
  build's multi_array::downsample bitwise on the committed fixtures
  (tests/golden/downsample_*.npz, written by oracle/_ref/ref_downsample_*),
  and closed forms pin it: a linear field is sampled at the centre of the
  cell, i.e. half a cell below along every staggered axis; a planted mistake
  (the stagger offset on the + side) fails the closed form.
"""

import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
from coffeegpu_amd import _h5

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump"
                                    if os.path.exists("/opt/conda/bin/h5dump") else None)

pytestmark = pytest.mark.skipif(not os.path.exists(_h5.LIB_PATH),
                                reason="libcoffee_h5.so not built (no HDF5 at build time)")


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.uint32, np.int32])
@pytest.mark.parametrize("shape", [(5, 4, 3), (6, 7), (9,)])
def test_array_round_trip(tmp_path, dtype, shape):
    rng = np.random.default_rng(1)
    a = (rng.standard_normal(shape) * 1000).astype(dtype)
    p = str(tmp_path / "a.h5")
    with _h5.H5File(p, "w") as f:
        f.write("A", a)
    with _h5.H5File(p) as f:
        b = f.read("A", shape, dtype)
    assert np.array_equal(a, b)


def test_scalars_and_has(tmp_path):
    p = str(tmp_path / "s.h5")
    with _h5.H5File(p, "w") as f:
        f.write_scalar("step", 4294967295, np.uint32)
        f.write_scalar("time", 0.1, np.float32)
        f.write_scalar("t64", 0.1, np.float64)
    with _h5.H5File(p) as f:
        assert f.read_scalar("step", np.uint32) == 4294967295
        assert f.read_scalar("time", np.float32) == np.float32(0.1)
        assert f.read_scalar("t64", np.float64) == 0.1
        assert f.has("step") and not f.has("nope")


def test_reads_check_shape_type_and_presence(tmp_path):
    p = str(tmp_path / "c.h5")
    with _h5.H5File(p, "w") as f:
        f.write("Ex", np.zeros((2, 3, 4), np.float32))
    with _h5.H5File(p) as f:
        with pytest.raises(_h5.H5Error, match="extent differs"):
            f.read("Ex", (2, 4, 3), np.float32)
        with pytest.raises(_h5.H5Error, match="rank differs"):
            f.read("Ex", (6, 4), np.float32)
        with pytest.raises(_h5.H5Error, match="element type differs"):
            f.read("Ex", (2, 3, 4), np.float64)
        with pytest.raises(_h5.H5Error, match="no dataset"):
            f.read("Ey", (2, 3, 4), np.float32)
    with pytest.raises(_h5.H5Error, match="cannot open"):
        _h5.H5File(str(tmp_path / "missing.h5"))


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not present")
def test_layout_seen_by_h5dump(tmp_path):
    """Dataset [z][y][x] for a Fortran-order {x, y, z} array, IEEE types,
    scalar step/time: what the reference's H5File::write produces."""
    p = str(tmp_path / "snap.h5")
    x, y, z = 4, 3, 2
    a = np.arange(x * y * z, dtype=np.float64).reshape(z, y, x)  # a[k, j, i] = i + x (j + y k)
    with _h5.H5File(p, "w") as f:
        f.write("Ex", a)
        f.write_scalar("step", 12, np.uint32)
        f.write_scalar("time", 1.5, np.float64)
    hdr = subprocess.run([H5DUMP, "-H", p], capture_output=True, text=True, check=True).stdout
    assert 'DATASET "Ex"' in hdr and "H5T_IEEE_F64LE" in hdr
    assert f"SIMPLE {{ ( {z}, {y}, {x} ) / ( {z}, {y}, {x} ) }}" in hdr
    assert "H5T_STD_U32LE" in hdr and "SCALAR" in hdr
    # element [k][j][i] of the dataset is the Fortran-order value i + x (j + y k)
    val = subprocess.run([H5DUMP, "-d", "Ex", "-s", "1,2,3", "-c", "1,1,1", p],
                         capture_output=True, text=True, check=True).stdout
    assert f"{3 + x * (2 + y * 1)}" in val.split("DATA {")[1]


# ---- field-output sampling ---------------------------------------------------

@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_oracle_downsample_matches_reference_fixtures(dtype):
    from tests.golden.make_golden import DS_FACTORS, DS_SHAPES, DS_STAGGERS

    z = np.load(os.path.join(GOLDEN, f"downsample_{np.dtype(dtype).name}.npz"))
    n = 0
    for s in range(len(DS_SHAPES)):
        a = z[f"in{s}"]
        guard = tuple(int(v) for v in z[f"guard{s}"])
        assert a.dtype == dtype
        for st in DS_STAGGERS:
            for d in DS_FACTORS:
                ref = z[f"out{s}_{st}_{d}"]
                got = O.downsample(a, guard, st, d)
                assert got.shape == ref.shape and np.array_equal(got, ref), (s, st, d)
                n += 1
    assert n == len(DS_SHAPES) * len(DS_STAGGERS) * len(DS_FACTORS)


def test_oracle_downsample_matches_reference_build_live():
    if not os.path.exists(os.path.join(O.REF_DIR, "ref_downsample_f64")):
        pytest.skip("reference build not present")
    rng = np.random.default_rng(5)
    a = rng.standard_normal((13, 15, 20))
    for st in (0b110, 0b001, 0b111):
        assert np.array_equal(O.downsample(a, (3, 3, 3), st, 2),
                              O.reference_downsample(a, (3, 3, 3), st, 2))


def _linear(dims, shift=(0, 0, 0)):
    k, j, i = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in dims[::-1]], indexing="ij")
    return 1.0 + 2.0 * (i + shift[0]) - 3.0 * (j + shift[1]) + 0.5 * (k + shift[2])


@pytest.mark.parametrize("st", [0b110, 0b101, 0b011, 0b001, 0b010, 0b100, 0b111, 0b000])
@pytest.mark.parametrize("d", [1, 2])
def test_downsample_closed_form_linear_field(st, d):
    """A linear field on nodes with the array's stagger is interpolated to the
    cell centre: along an axis whose stagger bit is 0 the value is the mean of
    nodes i-1 and i (position i - 1/2); along a bit-1 axis it is node i."""
    dims, guard = (14, 12, 10), (3, 3, 3)
    a = _linear(dims)
    got = O.downsample(a, guard, st, d).astype(np.float64)
    ex = O.downsample_extent(dims, guard, d)
    shift = tuple(-0.5 if not (st >> ax) & 1 else 0.0 for ax in range(3))
    idx = [np.arange(ex[ax]) * d + guard[ax] for ax in range(3)]
    k, j, i = np.meshgrid(idx[2], idx[1], idx[0], indexing="ij")
    want = 1.0 + 2.0 * (i + shift[0]) - 3.0 * (j + shift[1]) + 0.5 * (k + shift[2])
    assert np.array_equal(got, want.astype(np.float32))


def test_downsample_closed_form_rejects_plus_side_offset():
    """Planted mistake: taking the staggered neighbour at i+1 instead of i-1
    gives position i + 1/2 and fails the closed form."""
    dims, guard = (14, 12, 10), (3, 3, 3)
    a = _linear(dims)
    planted = O.downsample(a[:, :, ::-1].copy(), guard, 0b110, 1)[:, :, ::-1]
    right = O.downsample(a, guard, 0b110, 1)
    assert not np.array_equal(planted, right)


def test_downsample_2d_closed_form():
    """2D arrays (dims[2] == 1) take the x-y interpolation only."""
    dims, guard = (12, 10, 1), (3, 3, 0)
    k, j, i = np.meshgrid(np.arange(1.0), np.arange(10.0), np.arange(12.0), indexing="ij")
    a = 2.0 * i - 3.0 * j
    for st in (0b110, 0b101, 0b011, 0b001, 0b111):
        got = O.downsample(a, guard, st, 1)
        ex = O.downsample_extent(dims, guard, 1)
        sx = -0.5 if not st & 1 else 0.0
        sy = -0.5 if not st & 2 else 0.0
        jj, ii = np.meshgrid(np.arange(ex[1]) + 3, np.arange(ex[0]) + 3, indexing="ij")
        want = 2.0 * (ii + sx) - 3.0 * (jj + sy)
        assert got.shape == (1, ex[1], ex[0])
        assert np.array_equal(got[0], want.astype(np.float32)), st

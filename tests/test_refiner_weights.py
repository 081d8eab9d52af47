"""The refiner's weights file (CPU only): architecture tag, version, parameter
count and checksum are checked on load; a damaged or foreign file is refused."""
import numpy as np
import pytest

from spectrum_amd import capi
from spectrum_amd import refiner as R


def test_default_weights_are_valid():
    p = R.read_weights(R.DEFAULT_WEIGHTS)
    assert p.shape == (R.WEIGHTS_NPARAM,) and p.dtype == np.float32
    with np.load(R.DEFAULT_WEIGHTS, allow_pickle=False) as z:
        assert str(z["format"]) == R.WEIGHTS_FORMAT and int(z["version"]) == R.WEIGHTS_VERSION
        assert str(z["sha256"]) == R.params_digest(p)


def test_parameter_count_matches_the_library():
    assert capi.load().spx_mlp_param_count() == R.WEIGHTS_NPARAM


def test_round_trip_and_legacy_file(tmp_path):
    p = np.random.default_rng(1).standard_normal(R.WEIGHTS_NPARAM).astype(np.float32)
    f = tmp_path / "w.npz"
    R.write_weights(str(f), p, steps=3)
    assert np.array_equal(R.read_weights(str(f)), p)
    legacy = tmp_path / "legacy.npz"  # parameters only, as the first files were written
    np.savez(legacy, params=p)
    assert np.array_equal(R.read_weights(str(legacy)), p)


@pytest.mark.parametrize("damage", ["value", "format", "version", "shape", "dtype", "nan", "missing"])
def test_damaged_files_are_refused(tmp_path, damage):
    p = R.read_weights(R.DEFAULT_WEIGHTS).copy()
    fields = {"format": np.asarray(R.WEIGHTS_FORMAT), "version": np.asarray(R.WEIGHTS_VERSION),
              "sha256": np.asarray(R.params_digest(p))}
    if damage == "value":
        p[100] += 1e-3  # the checksum no longer matches
    elif damage == "format":
        fields["format"] = np.asarray("spx-refiner mlp 10-32-1 relu fp32")
    elif damage == "version":
        fields["version"] = np.asarray(R.WEIGHTS_VERSION + 1)
    elif damage == "shape":
        p = p[:-1]
        fields["sha256"] = np.asarray(R.params_digest(p))
    elif damage == "dtype":
        p = p.astype(np.float64)
    elif damage == "nan":
        p[7] = np.nan
        fields["sha256"] = np.asarray(R.params_digest(p))
    f = tmp_path / "bad.npz"
    if damage == "missing":
        np.savez(f, weights=p)
    else:
        np.savez(f, params=p, **fields)
    with pytest.raises(ValueError):
        R.read_weights(str(f))

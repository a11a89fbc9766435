"""simulation::save_results format (lib/simulation.cpp:38-63): five raw FP64
files per particle.  Pinned against the reference's own writer,
io::write_array (include/io.hpp:13-26), compiled into oracle/_ref, where
that library was built; otherwise against numpy's raw FP64 layout."""
import numpy as np
import pytest

from oracle import oracle as O

core = pytest.importorskip("magpy_amd.core")
from magpy_amd import results as R  # noqa: E402


def make_results(S=37, seed=0):
    rng = np.random.default_rng(seed)
    t = np.linspace(0, 1e-9, S)
    return R.Results(t, 1e4 * np.sin(t * 1e9), {0: rng.normal(size=S), 1: rng.normal(size=S)},
                     {0: rng.normal(size=S), 1: rng.normal(size=S)},
                     {0: rng.normal(size=S), 1: rng.normal(size=S)}, 2)


def test_files_match_the_reference_writer(tmp_path):
    res = make_results()
    prefix = str(tmp_path / "run")
    R.save_results(prefix, res, particle=1)
    arrays = {"mx": res.x[1], "my": res.y[1], "mz": res.z[1], "field": res.field,
              "time": res.time}
    ref = O.ref()
    for ext, a in arrays.items():
        got = (tmp_path / ("run." + ext)).read_bytes()
        if ref is not None:
            path = str(tmp_path / ("ref." + ext)).encode()
            assert ref.ref_write_array(path, O.ptr(O.f64(a)), a.size) == 0
            assert got == (tmp_path / ("ref." + ext)).read_bytes(), ext
        assert got == np.ascontiguousarray(a, dtype=np.float64).tobytes(), ext


def test_round_trip(tmp_path):
    res = make_results(S=5, seed=1)
    prefix = str(tmp_path / "x")
    R.save_results(prefix, res, particle=0)
    back = R.load_results(prefix)
    for c in "xyz":
        assert np.array_equal(getattr(back, c)[0], getattr(res, c)[0])
    assert np.array_equal(back.time, res.time) and np.array_equal(back.field, res.field)


def test_errors(tmp_path):
    with pytest.raises(ValueError):
        core.save_results(str(tmp_path / "no" / "such" / "dir"), [1.0], [1.0], [1.0], [1.0], [1.0])
    with pytest.raises(ValueError):
        core.save_results(str(tmp_path / "a"), [1.0, 2.0], [1.0], [1.0], [1.0], [1.0])

"""CPU checks of the C ABI library: it loads, exports every declared symbol,
and its host-only entry points (schedule, constants) agree with the oracle.
No device work is launched here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "magpy_amd.h")
LIB = os.path.join(ROOT, "magpy_amd", "libmagpy_amd.so")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?\s*(mp_\w+)\s*\(", text,
                                 re.M)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(LIB)
    names = declared_functions()
    assert len(names) >= 8
    for nm in names:
        assert hasattr(lib, nm), nm


def test_core_binding_imports_and_constants():
    import magpy_amd
    assert magpy_amd.core.get_KB() == 1.38064852e-23     # include/constants.hpp:10
    assert magpy_amd.core.get_mu0() == 1.25663706e-6     # include/constants.hpp:11
    assert magpy_amd.core.get_gamma() == 1.76086e11      # include/constants.hpp:12


@pytest.mark.parametrize("dt,end,S", [
    (2e-4, 20.0, 5), (2.0049e-4, 20.049, 101), (1e-2, 1.0, 7), (0.3, 1.0, 50), (0.1, 0.0, 3),
    (1.0025e-2, 3.3e2, 1001), (2.0049e-4, 3.34e3, 1000), (0.7, 0.7, 2), (1e-3, 1.0, 1),
])
def test_sample_schedule_matches_reference_loop(dt, end, S):
    import magpy_amd
    k_orc, n_orc = O.sample_schedule(dt, end, S)
    k, n = magpy_amd.core.sample_schedule(dt, end, S)
    assert np.array_equal(k, k_orc) and n == n_orc


def test_sample_schedule_random_grid():
    import magpy_amd
    rng = np.random.default_rng(5)
    for _ in range(300):
        dt = 10 ** rng.uniform(-4, 0)
        end = dt * rng.uniform(0, 3000)
        S = int(rng.integers(1, 200))
        k_orc, n_orc = O.sample_schedule(dt, end, S)
        k, n = magpy_amd.core.sample_schedule(dt, end, S)
        assert np.array_equal(k, k_orc) and n == n_orc, (dt, end, S)


def test_errors_are_reported_not_raised_through_the_abi():
    import magpy_amd
    with pytest.raises(ValueError):
        magpy_amd.core.sample_schedule(-1.0, 1.0, 4)
    # max_samples = -1 never terminates in the reference's loop: rejected up front
    with pytest.raises(ValueError):
        magpy_amd.core.simulate_ensemble([0, 1], [1e-8], [2.3e4], [[0, 0, 1]], [[0, 0, 1]],
                                         [[0, 0, 0]], 4e5, 0.1, 300, 'constant', 0.0, 0.0, [1],
                                         1e-14, 1e-12, -1)
    # a resumed run cannot start before step 0
    with pytest.raises(ValueError):
        magpy_amd.core.simulate_ensemble([0, 1], [1e-8], [2.3e4], [[0, 0, 1]], [[0, 0, 1]],
                                         [[0, 0, 0]], 4e5, 0.1, 300, 'constant', 0.0, 0.0, [1],
                                         1e-14, 1e-12, 4, start_step=-5)


def test_dom_arguments_are_checked_before_the_device():
    import magpy_amd
    bad = [dict(volume=[-1e-24]), dict(max_samples=1), dict(field_shape="triangle"),
           dict(field_shape="square_f", field_n_components=0), dict(time_step=0.0)]
    base = dict(volume=[1e-24], anisotropy=23e3, temperature=300.0, magnetisation=4e5,
                damping=0.1, initial_probabilities=[[1.0, 0.0]], time_step=1e-12,
                end_time=1e-9, max_samples=5, field_shape="sine", field_amplitude=1e4,
                field_frequency=3e5)
    for b in bad:
        kw = dict(base)
        kw.update(b)
        with pytest.raises(ValueError):
            magpy_amd.core.simulate_dom_ensemble(**kw)

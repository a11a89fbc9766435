"""Results of single-cluster and ensemble simulations.

Mirrors owlas/magpy's magpy/results.py (Results :6-91, EnsembleResults
:94-217) with the same methods and return values, with two fixes: the energy
integral uses this module's own trapezoid (scipy.integrate.trapz no longer
exists) and mu0 comes from the library (the reference calls an unimported
get_mu0, magpy/results.py:191).  EnsembleResults built from one batched run
keeps the batch arrays and creates per-member Results only on access.
"""
import numpy as np

from . import core


def trapezoid(y, x):
    """Trapezoidal integral of y dx in FP64 (lib/trap.cpp:16-29)."""
    y = np.asarray(y, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    if y.size < 2:
        return 0.0
    return float(np.sum(np.diff(x) * (y[1:] + y[:-1])) / 2.0)


class Results:
    """Results of a simulation of a single particle cluster (magpy/results.py:6-37).

    Args:
        time (np.ndarray): [M] sample times in seconds
        field (np.ndarray): [M] applied field (A/m) along z at each sample
        x, y, z (dict): {particle id: [M] array} of Ms*m components (A/m)
        N (int): number of particles in the cluster
    """

    def __init__(self, time, field, x, y, z, N):
        self.time = time
        self.field = field
        self.x = x
        self.y = y
        self.z = z
        self.N = N

    def plot(self):
        """One panel per particle showing its x, y and z magnetisation
        against time (magpy/results.py:39-59); returns the figure."""
        import matplotlib.pyplot as plt  # optional dependency, as in the reference
        fig, panels = plt.subplots(nrows=self.N, squeeze=False)
        for idx, ax in enumerate(panels[:, 0]):
            for comp in ('x', 'y', 'z'):
                ax.plot(self.time, getattr(self, comp)[idx], label=comp)
            ax.set_title('Particle %d' % idx)
            ax.set_xlabel('Time [s]')
            ax.set_ylabel('Magnetisation [A/m]')
            ax.legend()
        fig.tight_layout()
        return fig

    def magnetisation(self, direction='z'):
        """Sum over the cluster's particles of one component (magpy/results.py:61-75)."""
        return np.sum([vals for vals in getattr(self, direction).values()], axis=0)

    def final_state(self):
        """Last sample of every particle (magpy/results.py:77-91)."""
        return {
            'x': {k: v[-1] for k, v in self.x.items()},
            'y': {k: v[-1] for k, v in self.y.items()},
            'z': {k: v[-1] for k, v in self.z.items()},
        }


def save_results(prefix, results, particle=0):
    """Write one particle of a Results as simulation::save_results does
    (lib/simulation.cpp:38-63): raw FP64 prefix.mx, .my, .mz, .field, .time."""
    core.save_results(prefix, results.x[particle], results.y[particle], results.z[particle],
                      results.field, results.time)


def load_results(prefix):
    """Read the five files of save_results back into a one-particle Results."""
    arr = {ext: np.fromfile(prefix + "." + ext, dtype=np.float64)
           for ext in ("mx", "my", "mz", "field", "time")}
    return Results(arr["time"], arr["field"], {0: arr["mx"]}, {0: arr["my"]}, {0: arr["mz"]}, 1)


class _BatchResults:
    """Lazy sequence of per-member Results over one batched run."""

    def __init__(self, batch, offsets, member_time=None, member_field=None):
        self._b = batch
        self._o = np.asarray(offsets)
        self._mt = member_time
        self._mf = member_field

    def __len__(self):
        return len(self._o) - 1

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[j] for j in range(*i.indices(len(self)))]
        if i < 0:
            i += len(self)
        if not 0 <= i < len(self):
            raise IndexError(i)
        b, e = int(self._o[i]), int(self._o[i + 1])
        if 'x' not in self._b:
            raise ValueError("per-particle trajectories were not kept (trajectories=False)")
        time = self._mt[i] if self._mt is not None else self._b['time']
        field = self._mf[i] if self._mf is not None else self._b['field']
        return Results(time=time, field=field,
                       x={k: self._b['x'][b + k] for k in range(e - b)},
                       y={k: self._b['y'][b + k] for k in range(e - b)},
                       z={k: self._b['z'][b + k] for k in range(e - b)},
                       N=e - b)

    def __iter__(self):
        for i in range(len(self)):
            yield self[i]


class EnsembleResults:
    """Results of an ensemble of clusters (magpy/results.py:94-114).

    Args:
        results (list[Results]): results of each member; or the lazy batch
            sequence made by EnsembleResults.from_batch.
    """

    def __init__(self, results):
        self.results = results
        self.time = results[0].time if not isinstance(results, _BatchResults) else results._b['time']
        self.field = results[0].field if not isinstance(results, _BatchResults) else results._b['field']
        self._batch = results._b if isinstance(results, _BatchResults) else None

    @classmethod
    def from_batch(cls, batch, offsets):
        seq = _BatchResults(batch, offsets, batch.get('member_time'), batch.get('member_field'))
        return cls(seq)

    @property
    def stats(self):
        """Device-run diagnostics of a batched run (steps, Newton counts, kernel time)."""
        if self._batch is None:
            return {}
        return {k: self._batch[k] for k in ('n_steps', 'newton_iters', 'newton_failures',
                                            'kernel_ms', 'status')}

    def magnetisation(self, direction='z'):
        """Total magnetisation of each member (magpy/results.py:116-132)."""
        return [res.magnetisation(direction) for res in self.results]

    def ensemble_magnetisation(self, direction='z'):
        """Mean over members of each member's total magnetisation
        (magpy/results.py:134-151); reduced on the device in FP64 for
        batched runs."""
        if self._batch is not None and 'ensemble_m' in self._batch:
            col = {'x': 0, 'y': 1, 'z': 2}[direction]
            return self._batch['ensemble_m'][:, col] / len(self.results)
        return np.sum(self.magnetisation(direction), axis=0) / len(self.results)

    def final_state(self):
        """Final state of each member (magpy/results.py:153-165)."""
        return [res.final_state() for res in self.results]

    def energy_dissipated(self, start_time=None, end_time=None):
        """Hysteresis-loop energy -mu0 * integral of H dM over a time window
        (magpy/results.py:167-191), in J/m^3 per unit of the ensemble's mean
        cluster magnetisation."""
        before_mask = (self.time >= start_time) if start_time is not None else True
        after_mask = (self.time <= end_time) if end_time is not None else True
        mask = before_mask & after_mask
        if mask is True:
            mask = np.ones_like(self.time, dtype=bool)
        return -core.get_mu0() * trapezoid(self.field[mask], self.ensemble_magnetisation()[mask])

    def final_cycle_energy_dissipated(self, field_frequency):
        """Energy dissipated over the last field period (magpy/results.py:193-217)."""
        T = 1. / field_frequency
        return self.energy_dissipated(start_time=self.time[-1] - T)

"""Model and EnsembleModel with owlas/magpy's interface (magpy/model.py:16-208).

Model.simulate runs one cluster through the device integrator; the
EnsembleModel runs all of its members in one batched device call (the C++
simulation::ensemble_run behind core.simulate_ensemble) instead of one
joblib process per member.  ``n_jobs`` is accepted for compatibility and has
no effect.  Both keep the reference's defaults, including
``implicit_solve=True``; ``precision`` ('fp64', the reference's arithmetic,
or 'fp32') is an addition.
"""
import numpy as np

from . import core
from .results import EnsembleResults, Results


class Model:
    """A cluster of interacting magnetic nanoparticles (magpy/model.py:16-68).

    Args:
        radius (list of double): radius of each spherical particle (m)
        anisotropy (list of double): uniaxial anisotropy constant of each particle (J/m^3)
        anisotropy_axis (list of ndarray[double,3]): unit anisotropy axis of each particle
        magnetisation_direction (list of ndarray[double,3]): initial unit magnetisation
        location (list of ndarray[double,3]): particle positions (m)
        magnetisation (double): saturation magnetisation of all particles (A/m)
        damping (double): damping of all particles
        temperature (double): temperature (K)
        field_shape (str, optional): 'constant', 'square' or 'sine' (along z)
        field_frequency (double, optional): field frequency (Hz)
        field_amplitude (double, optional): field amplitude (A/m)
    """

    def __init__(self, radius, anisotropy, anisotropy_axis, magnetisation_direction, location,
                 magnetisation, damping, temperature, field_shape='constant',
                 field_frequency=0.0, field_amplitude=0.0):
        self.radius = np.array(radius)
        self.anisotropy = np.array(anisotropy)
        self.anisotropy_axis = np.array(anisotropy_axis)
        self.magnetisation_direction = np.array(magnetisation_direction)
        self.location = np.array(location)
        self.magnetisation = magnetisation
        self.damping = damping
        self.temperature = temperature
        self.field_shape = field_shape
        self.field_frequency = field_frequency
        self.field_amplitude = field_amplitude

    def simulate(self, end_time, time_step, max_samples, seed=1001, renorm=False,
                 interactions=True, implicit_solve=True, implicit_tol=1e-9, precision='fp64',
                 device=0):
        """Simulate the cluster (magpy/model.py:70-115).

        Args:
            end_time (float): simulated time (s)
            time_step (float): integrator step (s)
            max_samples (int): regularly spaced output samples
            seed (int): seed of the particle Wiener streams
            renorm (bool): rescale |m| to 1 after every step
            interactions (bool): dipolar interactions between the particles
            implicit_solve (bool): implicit midpoint (True) or Heun (False)
            implicit_tol (float): relative tolerance of the implicit Newton solve
            precision (str): 'fp64' or 'fp32' state arithmetic (Heun)
        Returns:
            Results
        """
        res = core.simulate(self.radius, self.anisotropy, self.anisotropy_axis,
                            self.magnetisation_direction, self.location, self.magnetisation,
                            self.damping, self.temperature, renorm, interactions,
                            implicit_solve, time_step, end_time, max_samples, seed,
                            self.field_shape, self.field_amplitude, self.field_frequency,
                            implicit_tol, precision=precision, device=device)
        return Results(**res)


class EnsembleModel:
    """Ensemble of particle clusters (magpy/model.py:118-156).

    Every member copies ``base_model``; each keyword argument is a list of N
    values that overrides that parameter for the i'th member.
    """

    def __init__(self, N, base_model, **kwargs):
        self.ensemble_size = N
        self.base_model_params = base_model.__dict__
        self.model_params = [
            {**self.base_model_params, **{key: kwargs[key][i] for key in kwargs}}
            for i in range(self.ensemble_size)
        ]
        self.models = [Model(**params) for params in self.model_params]

    def batch_arrays(self):
        """Member arrays concatenated for one batched call."""
        ms = self.models
        counts = [len(np.atleast_1d(m.radius)) for m in ms]
        offs = np.zeros(len(ms) + 1, dtype=np.int64)
        offs[1:] = np.cumsum(counts)
        cat = lambda name, shape: np.concatenate(  # noqa: E731
            [np.reshape(np.asarray(getattr(m, name), dtype=np.float64), shape) for m in ms])
        return dict(
            member_offsets=offs,
            radius=cat('radius', (-1,)),
            anisotropy=cat('anisotropy', (-1,)),
            anisotropy_axis=cat('anisotropy_axis', (-1, 3)),
            magnetisation_direction=cat('magnetisation_direction', (-1, 3)),
            location=cat('location', (-1, 3)),
            magnetisation=np.array([m.magnetisation for m in ms], dtype=np.float64),
            damping=np.array([m.damping for m in ms], dtype=np.float64),
            temperature=np.array([m.temperature for m in ms], dtype=np.float64),
            field_shape=[m.field_shape for m in ms],
            field_amplitude=np.array([m.field_amplitude for m in ms], dtype=np.float64),
            field_frequency=np.array([m.field_frequency for m in ms], dtype=np.float64),
        )

    def simulate(self, end_time, time_step, max_samples, random_state, renorm=False,
                 interactions=True, n_jobs=1, implicit_solve=True, implicit_tol=1e-9,
                 precision='fp64', trajectories=True, device=0):
        """Simulate every member (magpy/model.py:159-208).

        Member seeds are drawn as the reference draws them: numpy's global
        legacy generator seeded with ``random_state`` (magpy/model.py:202-203).
        ``trajectories=False`` keeps only the device-reduced ensemble
        magnetisation (what SAR needs) instead of every particle's samples.
        """
        np.random.seed(random_state)
        sim_seeds = np.random.randint(np.iinfo(np.int32).max, size=self.ensemble_size)
        arr = self.batch_arrays()
        res = core.simulate_ensemble(
            seeds=sim_seeds.astype(np.int64), time_step=time_step, end_time=end_time,
            max_samples=max_samples, renorm=renorm, interactions=interactions,
            use_implicit=implicit_solve, implicit_tol=implicit_tol, precision=precision,
            trajectories=trajectories, ensemble_sum=True, device=device,
            member_times=trajectories and self._members_differ(), **arr)
        return EnsembleResults.from_batch(res, arr['member_offsets'])

    def _members_differ(self):
        keys = ('anisotropy', 'magnetisation', 'damping', 'field_shape', 'field_frequency',
                'field_amplitude')
        first = self.model_params[0]
        for p in self.model_params[1:]:
            for k in keys:
                if not np.array_equal(np.asarray(p[k]), np.asarray(first[k])):
                    return True
        return False


class DOModel:
    """Two-state (up/down) probabilistic model of one particle with uniaxial
    anisotropy (magpy/model.py:211-300): the state probabilities obey a
    master equation with Neel-Brown rates in a field applied along the
    anisotropy axis ('constant', 'sine', 'square' or 'square_f' with
    field_n_components Fourier terms)."""

    def __init__(self, radius, anisotropy, initial_probabilities, magnetisation, damping,
                 temperature, field_shape='constant', field_frequency=0.0, field_amplitude=0.0,
                 field_n_components=1):
        self.radius = radius
        self.volume = 4. / 3 * np.pi * self.radius ** 3
        self.anisotropy = anisotropy
        self.initial_probabilities = np.array(initial_probabilities)
        self.magnetisation = magnetisation
        self.damping = damping
        self.temperature = temperature
        self.field_shape = field_shape
        self.field_frequency = field_frequency
        self.field_amplitude = field_amplitude
        self.field_n_components = field_n_components

    def simulate(self, end_time, time_step, max_samples):
        """Probabilities sampled at max_samples regular times (first-order
        hold) as a Results whose z holds p_up - p_down (magpy/model.py:262-300)."""
        results = core.simulate_dom(
            self.initial_probabilities, self.volume, self.anisotropy, self.temperature,
            self.magnetisation, self.damping, time_step, end_time, max_samples,
            self.field_shape, self.field_amplitude, self.field_frequency,
            self.field_n_components)
        return Results(**results)

"""Host-array particle filter: drop-in for ``gnss_gpu.particle_filter.ParticleFilter``
(python/gnss_gpu/particle_filter.py:10-222): same constructor, methods,
defaults, return types and errors.

Between calls the particles live in host memory, as the reference's legacy
filter keeps them: one float64 block of four rows (x, y, z, clock bias) and the
log-weights.  Every method hands those buffers to the in-place functions of
``_gnss_gpu_pf``, which run the step on the GPU and write the results back.
"""

from __future__ import annotations

import numpy as np

from . import _gnss_gpu_pf as _pf

# The reference wrapper resamples with 15 Megopolis iterations
# (python/gnss_gpu/particle_filter.py:168).
MEGOPOLIS_ITERATIONS = 15


def _megopolis(pf, rows):
    _pf.pf_resample_megopolis(*rows, pf._log_weights, pf.n_particles, MEGOPOLIS_ITERATIONS,
                              pf.seed)


def _systematic(pf, rows):
    _pf.pf_resample_systematic(*rows, pf._log_weights, pf.n_particles, pf.seed)


# resampling method -> routine; any other name resamples systematically, as the
# reference's if/else does (particle_filter.py:165-172)
_RESAMPLERS = {"megopolis": _megopolis}


class ParticleFilter:
    """GNSS particle filter over host arrays, ECEF coordinates.

    Parameters
    ----------
    n_particles : int
    sigma_pos, sigma_cb : float
        Predict noise standard deviations [m] (position per axis, clock bias).
    sigma_pr : float
        Pseudorange standard deviation [m] of the Gaussian weight.
    resampling : str
        "megopolis" (15 iterations of the reference rule) or "systematic".
    ess_threshold : float
        Resample when ESS < ess_threshold * n_particles.
    seed : int
    """

    def __init__(self, n_particles=1_000_000, sigma_pos=1.0, sigma_cb=300.0,
                 sigma_pr=5.0, resampling="megopolis", ess_threshold=0.5,
                 seed=42):
        self.n_particles = n_particles
        self.sigma_pos = sigma_pos
        self.sigma_cb = sigma_cb
        self.sigma_pr = sigma_pr
        self.resampling = resampling
        self.ess_threshold = ess_threshold
        self.seed = seed
        self._block = None          # float64 [4, N]: rows x, y, z, clock bias
        self._log_weights = None    # float64 [N]
        self._initialized = False
        self._step = 0

    def _rows(self):
        """The particle rows (x, y, z, cb) in the argument order of _gnss_gpu_pf;
        raises until initialize() has run."""
        if not self._initialized:
            raise RuntimeError("ParticleFilter not initialized. Call initialize() first.")
        return tuple(self._block)

    # Subclasses written against the reference (its ParticleFilter3D, for one)
    # pass these per-coordinate arrays to the in-place functions; they are views
    # of the block, so such writes land in the filter's state.
    def _row(k):
        return property(lambda self: None if self._block is None else self._block[k])

    _px, _py, _pz, _pcb = _row(0), _row(1), _row(2), _row(3)
    del _row

    def initialize(self, position_ecef, clock_bias=0.0, spread_pos=100.0, spread_cb=1000.0):
        """Scatter N particles around position_ecef (3,) [m] and clock_bias [m]."""
        x0 = np.asarray(position_ecef, dtype=np.float64).reshape(-1)[:3]
        n = self.n_particles
        self._block = np.empty((4, n), dtype=np.float64)
        self._log_weights = np.zeros(n, dtype=np.float64)
        _pf.pf_initialize(*self._block, float(x0[0]), float(x0[1]), float(x0[2]),
                          float(clock_bias), float(spread_pos), float(spread_cb), n, self.seed)
        self._initialized = True
        self._step = 0

    def predict(self, velocity=None, dt=1.0):
        """Move every particle by velocity * dt (ECEF, default 0) plus noise."""
        rows = self._rows()
        v = np.zeros(3) if velocity is None else np.asarray(velocity, dtype=np.float64).reshape(-1)
        self._step += 1
        _pf.pf_predict(*rows, *(np.array([c], dtype=np.float64) for c in v[:3]), float(dt),
                       float(self.sigma_pos), float(self.sigma_cb), self.n_particles, self.seed,
                       self._step)

    def update(self, sat_ecef, pseudoranges, weights=None):
        """Weight by the pseudoranges (overwrites the log-weights), then resample
        when the effective sample size falls below the threshold."""
        rows = self._rows()
        sat = np.asarray(sat_ecef, dtype=np.float64).reshape(-1, 3)
        pr = np.asarray(pseudoranges, dtype=np.float64).reshape(-1)
        w = (np.ones(pr.size, dtype=np.float64) if weights is None
             else np.asarray(weights, dtype=np.float64).reshape(-1))
        _pf.pf_weight(*rows, sat.reshape(-1), pr, w, self._log_weights, self.n_particles,
                      pr.size, float(self.sigma_pr))
        if self.get_ess() < self.ess_threshold * self.n_particles:
            self._resample()

    def _resample(self):
        _RESAMPLERS.get(self.resampling, _systematic)(self, self._rows())
        self._log_weights.fill(0.0)

    def estimate(self):
        """Weighted mean [x, y, z, clock_bias] as ndarray (4,)."""
        rows = self._rows()
        out = np.empty(4, dtype=np.float64)
        _pf.pf_estimate(*rows, self._log_weights, out, self.n_particles)
        return out

    def get_particles(self):
        """All particles as an (N, 4) array of [x, y, z, clock_bias]."""
        rows = self._rows()
        out = np.empty(4 * self.n_particles, dtype=np.float64)
        _pf.pf_get_particles(*rows, out, self.n_particles)
        return out.reshape(self.n_particles, 4)

    def get_ess(self):
        """Effective sample size of the current log-weights."""
        self._rows()
        return _pf.pf_compute_ess(self._log_weights, self.n_particles)

"""Device-resident particle filter for GNSS pseudorange positioning (MI355X).

Drop-in for ``gnss_gpu.particle_filter_device.ParticleFilterDevice``
(python/gnss_gpu/particle_filter_device.py:18-886): same constructor keywords
and defaults, same method names and return types.  The state stays on the GPU
between calls; only satellite rows (per update), the velocity guide (per
predict) and the requested results cross PCIe.

Additions (keyword-only, reference behaviour by default):
  n_isb / sigma_isb      inter-system-bias states for multi-GNSS pseudoranges
                         (``update(..., system_ids=...)``)
  megopolis_mode         "reference" (the reference kernel's rule) or "chesser"
                         (Megopolis of Chesser et al. 2021)
  megopolis_iterations   iterations per Megopolis resample (reference: 15)
  systematic_u0          "reference" (u0 = (seed % 1e6) / (1e6 N), identical
                         every call) or "fresh" (a new Philox u0 in [0,1))
  set_learned_likelihood a gnss_gpu_amd.learned.LearnedLikelihood; updates
                         given ``cn0=`` and ``los_flag=`` then use it in place
                         of the Gaussian / Student-t term
"""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np

from . import _gnss_gpu_pf_device as _dev


class ParticleFilterDevice:
    """High-performance particle filter with persistent GPU memory.

    Parameters
    ----------
    n_particles : int
        Number of particles (e.g., 1_000_000).
    sigma_pos : float
        Position noise standard deviation [m] for prediction.
    sigma_cb : float
        Clock bias noise standard deviation [m] for prediction.
    sigma_pr : float
        Pseudorange observation standard deviation [m] for weighting.
    nu : float
        Student's t degrees of freedom (0 = Gaussian, 1 = Cauchy).
    resampling : str
        "megopolis" or "systematic".
    ess_threshold : float
        Resample when ESS < ess_threshold * n_particles.
    seed : int
        Random seed.
    """

    def __init__(self, n_particles=1_000_000, sigma_pos=1.0, sigma_cb=300.0,
                 sigma_pr=5.0, nu=0.0, resampling="megopolis", ess_threshold=0.5,
                 seed=42, per_particle_nlos_gate=False,
                 per_particle_nlos_dd_pr_threshold_m=10.0,
                 per_particle_nlos_dd_carrier_threshold_cycles=0.5,
                 per_particle_nlos_undiff_pr_threshold_m=30.0,
                 per_particle_huber=False,
                 per_particle_huber_dd_pr_k=1.5,
                 per_particle_huber_dd_carrier_k=1.5,
                 per_particle_huber_undiff_pr_k=1.5,
                 sigma_vel=0.0,
                 velocity_guide_alpha=1.0,
                 rbpf_velocity_kf=False,
                 velocity_process_noise=0.0,
                 *, n_isb=0, sigma_isb=0.0, megopolis_mode="reference",
                 megopolis_iterations=15, systematic_u0="reference"):
        if megopolis_mode not in ("reference", "chesser"):
            raise ValueError("megopolis_mode must be 'reference' or 'chesser'")
        if systematic_u0 not in ("reference", "fresh"):
            raise ValueError("systematic_u0 must be 'reference' or 'fresh'")
        self.n_particles = n_particles
        self.sigma_pos = sigma_pos
        self.sigma_cb = sigma_cb
        self.sigma_pr = sigma_pr
        self.nu = nu
        self.resampling = resampling
        self.ess_threshold = ess_threshold
        self.seed = seed
        self.per_particle_nlos_gate = bool(per_particle_nlos_gate)
        self.per_particle_nlos_dd_pr_threshold_m = float(per_particle_nlos_dd_pr_threshold_m)
        self.per_particle_nlos_dd_carrier_threshold_cycles = float(
            per_particle_nlos_dd_carrier_threshold_cycles)
        self.per_particle_nlos_undiff_pr_threshold_m = float(
            per_particle_nlos_undiff_pr_threshold_m)
        self.per_particle_huber = bool(per_particle_huber)
        self.per_particle_huber_dd_pr_k = float(per_particle_huber_dd_pr_k)
        self.per_particle_huber_dd_carrier_k = float(per_particle_huber_dd_carrier_k)
        self.per_particle_huber_undiff_pr_k = float(per_particle_huber_undiff_pr_k)
        self.sigma_vel = float(sigma_vel)
        self.velocity_guide_alpha = float(velocity_guide_alpha)
        self.rbpf_velocity_kf = bool(rbpf_velocity_kf)
        self.velocity_process_noise = float(velocity_process_noise)
        self.n_isb = int(n_isb)
        self.sigma_isb = float(sigma_isb)
        self.megopolis_mode = megopolis_mode
        self.megopolis_iterations = int(megopolis_iterations)
        self.systematic_u0 = systematic_u0

        self._state = _dev.pf_device_create(n_particles, self.n_isb)
        self._initialized = False
        self._step = 0

    # -- helpers (particle_filter_device.py:138-146) --------------------------
    def _per_particle_threshold(self, value):
        if not self.per_particle_nlos_gate:
            return 0.0
        value = float(value)
        return value if np.isfinite(value) and value > 0.0 else 0.0

    def _per_particle_huber_k(self, value):
        value = float(value)
        return value if np.isfinite(value) and value > 0.0 else 1.5

    def _require_init(self):
        if not self._initialized:
            raise RuntimeError("ParticleFilterDevice not initialized. Call initialize() first.")

    # -- lifecycle -------------------------------------------------------------
    def initialize(self, position_ecef, clock_bias=0.0, spread_pos=100.0,
                   spread_cb=1000.0, velocity=None, spread_vel=0.0,
                   velocity_init_sigma=0.0, *, isb=None, spread_isb=0.0):
        """Scatter particles around an initial estimate (runs on the GPU)."""
        pos = np.asarray(position_ecef, dtype=np.float64).ravel()
        vel = (np.zeros(3, dtype=np.float64) if velocity is None
               else np.asarray(velocity, dtype=np.float64).ravel())
        _dev.pf_device_initialize(
            self._state, float(pos[0]), float(pos[1]), float(pos[2]), float(clock_bias),
            float(spread_pos), float(spread_cb), self.seed,
            float(vel[0]), float(vel[1]), float(vel[2]), float(spread_vel),
            float(velocity_init_sigma), init_isb=isb, spread_isb=spread_isb)
        self._initialized = True
        self._step = 0
        self._velocity_init_sigma = float(velocity_init_sigma)

    def predict(self, velocity=None, dt=1.0, sigma_pos=None, sigma_vel=None,
                velocity_guide_alpha=None, rbpf_velocity_kf=None,
                velocity_process_noise=None):
        """Predict step with optional velocity guide (24 bytes to the GPU)."""
        self._require_init()
        if velocity is None:
            vx, vy, vz = 0.0, 0.0, 0.0
        else:
            vel = np.asarray(velocity, dtype=np.float64).ravel()
            vx, vy, vz = float(vel[0]), float(vel[1]), float(vel[2])
        sp = float(self.sigma_pos if sigma_pos is None else sigma_pos)
        sv = float(self.sigma_vel if sigma_vel is None else sigma_vel)
        alpha = float(self.velocity_guide_alpha if velocity_guide_alpha is None
                      else velocity_guide_alpha)
        use_kf = self.rbpf_velocity_kf if rbpf_velocity_kf is None else bool(rbpf_velocity_kf)
        qv = float(self.velocity_process_noise if velocity_process_noise is None
                   else velocity_process_noise)
        self._step += 1
        _dev.pf_device_predict(self._state, vx, vy, vz, float(dt), sp, float(self.sigma_cb),
                               self.seed, self._step, sv, alpha, use_kf, qv,
                               sigma_isb=self.sigma_isb)

    def set_learned_likelihood(self, model, table=None):
        """Attach a LearnedLikelihood (None detaches).  The model must outlive
        the updates that use it.  table: evaluate the network per satellite
        over the bf16 codes of its input and look the particles up (same
        values); None chooses it when the particles outnumber the codes."""
        self._learned = model
        self._learned_table = (self.n_particles > 33602) if table is None else bool(table)
        _dev.pf_device_set_mlp(self._state, model)

    def update(self, sat_ecef, pseudoranges, weights=None, sigma_pr=None, resample=True,
               *, system_ids=None, cn0=None, los_flag=None):
        """Pseudorange weight update (overwrites the log-weights, like the reference).
        With a learned likelihood attached, ``cn0`` (dB-Hz) and ``los_flag``
        (line of sight at the particle mean) per satellite select it."""
        self._require_init()
        sat = np.asarray(sat_ecef, dtype=np.float64).reshape(-1, 3)
        pr = np.asarray(pseudoranges, dtype=np.float64).ravel()
        n_sat = len(pr)
        if weights is None:
            weights = np.ones(n_sat, dtype=np.float64)
        else:
            weights = np.asarray(weights, dtype=np.float64).ravel()
        if cn0 is not None or los_flag is not None:
            if getattr(self, "_learned", None) is None:
                raise RuntimeError("cn0/los_flag given but no learned likelihood is attached")
            if cn0 is None or los_flag is None:
                raise ValueError("the learned likelihood needs both cn0 and los_flag")
            _dev.pf_device_weight_learned(self._state, sat.ravel(), pr, weights, n_sat, cn0,
                                          los_flag, system_ids=system_ids,
                                          table=self._learned_table)
            if resample:
                _ = self.resample_if_needed()
            return
        sp = float(self.sigma_pr if sigma_pr is None else sigma_pr)
        _dev.pf_device_weight(
            self._state, sat.ravel(), pr, weights, n_sat, sp, float(self.nu),
            self._per_particle_threshold(self.per_particle_nlos_undiff_pr_threshold_m),
            self.per_particle_huber,
            self._per_particle_huber_k(self.per_particle_huber_undiff_pr_k),
            system_ids=system_ids)
        if resample:
            _ = self.resample_if_needed()

    def update_3d_bvh(self, bvh, sat_ecef, pseudoranges, weights=None, sigma_los=3.0,
                      sigma_nlos=30.0, nlos_bias=20.0, blocked_nlos_prob=1.0,
                      clear_nlos_prob=0.0, resample=True, *, system_ids=None):
        """Ray-traced NLOS likelihood on the device particles (the update of the
        reference's ParticleFilter3DBVH, particle_filter_3d_bvh.py:79-120, on this
        filter's resident state); overwrites the log-weights.  ``bvh`` is a
        ``BVHAccelerator`` over the building mesh (ECEF)."""
        self._require_init()
        sat = np.asarray(sat_ecef, dtype=np.float64).reshape(-1, 3)
        pr = np.asarray(pseudoranges, dtype=np.float64).ravel()
        n_sat = len(pr)
        w = (np.ones(n_sat, dtype=np.float64) if weights is None
             else np.asarray(weights, dtype=np.float64).ravel())
        _dev.pf_device_weight_3d_bvh(self._state, bvh, sat.ravel(), pr, w, n_sat, sigma_los,
                                     sigma_nlos, nlos_bias, blocked_nlos_prob, clear_nlos_prob,
                                     system_ids=system_ids)
        if resample:
            _ = self.resample_if_needed()

    def update_gmm(self, sat_ecef, pseudoranges, weights=None, sigma_pr=None, w_los=0.7,
                   mu_nlos=15.0, sigma_nlos=30.0, resample=True):
        """LOS / NLOS Gaussian-mixture likelihood (particle_filter_device.py:305-351):
        per satellite w_los N(e; 0, sigma_pr) + (1 - w_los) N(e; mu_nlos, sigma_nlos),
        ``weights`` scaling both exponents; overwrites the log-weights, then the
        ESS-based resampling unless ``resample`` is False."""
        self._require_init()
        sat = np.asarray(sat_ecef, dtype=np.float64).reshape(-1, 3)
        pr = np.asarray(pseudoranges, dtype=np.float64).ravel()
        n_sat = len(pr)
        w = (np.ones(n_sat, dtype=np.float64) if weights is None
             else np.asarray(weights, dtype=np.float64).ravel())
        sp = float(self.sigma_pr if sigma_pr is None else sigma_pr)
        _dev.pf_device_weight_gmm(self._state, sat.ravel(), pr, w, n_sat, sp, float(w_los),
                                  float(mu_nlos), float(sigma_nlos))
        if resample:
            _ = self.resample_if_needed()

    def update_dd_pseudorange(self, dd_result, sigma_pr=0.75, resample=True):
        """DD pseudorange likelihood (no clock bias); overwrites the log-weights
        (particle_filter_device.py:354-388)."""
        self._require_init()
        _dev.pf_device_weight_dd_pseudorange(
            self._state,
            np.asarray(dd_result.sat_ecef_k, dtype=np.float64).ravel(),
            np.asarray(dd_result.sat_ecef_ref, dtype=np.float64).ravel(),
            dd_result.dd_pseudorange_m, dd_result.base_range_k, dd_result.base_range_ref,
            dd_result.dd_weights, dd_result.n_dd, float(sigma_pr),
            self._per_particle_threshold(self.per_particle_nlos_dd_pr_threshold_m),
            self.per_particle_huber,
            self._per_particle_huber_k(self.per_particle_huber_dd_pr_k))
        if resample:
            _ = self.resample_if_needed()

    def update_carrier_afv(self, sat_ecef, carrier_phase_cycles, weights=None,
                           wavelength=0.190293673, sigma_cycles=0.05, resample=True):
        """Undifferenced carrier-phase AFV likelihood, added to the log-weights
        (particle_filter_device.py:390-432)."""
        self._require_init()
        sat = np.asarray(sat_ecef, dtype=np.float64).reshape(-1, 3)
        cp = np.asarray(carrier_phase_cycles, dtype=np.float64).ravel()
        n_sat = len(cp)
        w = (np.ones(n_sat, dtype=np.float64) if weights is None
             else np.asarray(weights, dtype=np.float64).ravel())
        _dev.pf_device_weight_carrier_afv(self._state, sat.ravel(), cp, w, n_sat,
                                          float(wavelength), float(sigma_cycles))
        if resample:
            _ = self.resample_if_needed()

    def update_dd_carrier_afv(self, dd_result, wavelength=None, sigma_cycles=0.05,
                              resample=True):
        """DD carrier-phase AFV likelihood, added to the log-weights; per-pair
        wavelengths from dd_result unless overridden (particle_filter_device.py:434-480)."""
        self._require_init()
        if wavelength is None:
            wl = getattr(dd_result, "wavelengths_m", None)
            wl = (np.full(dd_result.n_dd, 0.190293673, dtype=np.float64) if wl is None
                  else np.asarray(wl, dtype=np.float64).ravel())
        else:
            wl = np.full(dd_result.n_dd, float(wavelength), dtype=np.float64)
        _dev.pf_device_weight_dd_carrier_afv(
            self._state,
            np.asarray(dd_result.sat_ecef_k, dtype=np.float64).ravel(),
            np.asarray(dd_result.sat_ecef_ref, dtype=np.float64).ravel(),
            dd_result.dd_carrier_cycles, dd_result.base_range_k, dd_result.base_range_ref,
            dd_result.dd_weights, wl, dd_result.n_dd, float(sigma_cycles),
            self._per_particle_threshold(self.per_particle_nlos_dd_carrier_threshold_cycles),
            self.per_particle_huber,
            self._per_particle_huber_k(self.per_particle_huber_dd_carrier_k))
        if resample:
            _ = self.resample_if_needed()

    def _doppler_rows(self, sat_ecef, sat_vel, doppler_hz, weights):
        sat = np.asarray(sat_ecef, dtype=np.float64).reshape(-1, 3)
        sv = np.asarray(sat_vel, dtype=np.float64).reshape(-1, 3)
        dop = np.asarray(doppler_hz, dtype=np.float64).ravel()
        if sat.shape != sv.shape:
            raise ValueError("sat_ecef and sat_vel must have the same shape")
        if sat.shape[0] != dop.size:
            raise ValueError("doppler_hz length must match sat_ecef rows")
        n_sat = int(dop.size)
        w = (np.ones(n_sat, dtype=np.float64) if weights is None
             else np.asarray(weights, dtype=np.float64).ravel())
        if w.size != n_sat:
            raise ValueError("weights length must match doppler_hz")
        return sat, sv, dop, w, n_sat

    def update_doppler(self, sat_ecef, sat_vel, doppler_hz, weights=None,
                       wavelength=0.19029367279836488, sigma_mps=0.5,
                       velocity_update_gain=0.25, max_velocity_update_mps=10.0, resample=True):
        """Doppler likelihood and per-particle velocity nudge toward each
        particle's Doppler WLS solution (particle_filter_device.py:484-544)."""
        self._require_init()
        sat, sv, dop, w, n_sat = self._doppler_rows(sat_ecef, sat_vel, doppler_hz, weights)
        _dev.pf_device_weight_doppler(self._state, sat.ravel(), sv.ravel(), dop, w, n_sat,
                                      float(wavelength), float(sigma_mps),
                                      float(velocity_update_gain),
                                      float(max_velocity_update_mps))
        if resample:
            _ = self.resample_if_needed()

    def update_doppler_kf(self, sat_ecef, sat_vel, doppler_hz, weights=None,
                          wavelength=0.19029367279836488, sigma_mps=0.5, resample=True):
        """Per-particle velocity Kalman update (RBPF) from Doppler rows
        (particle_filter_device.py:546-582)."""
        self._require_init()
        sat, sv, dop, w, n_sat = self._doppler_rows(sat_ecef, sat_vel, doppler_hz, weights)
        _dev.pf_device_doppler_kf_update(self._state, sat.ravel(), sv.ravel(), dop, w, n_sat,
                                         float(wavelength), float(sigma_mps))
        if resample:
            _ = self.resample_if_needed()

    def position_update(self, ref_ecef, sigma_pos=30.0):
        """Gaussian position-domain likelihood added to the log-weights."""
        self._require_init()
        ref = np.asarray(ref_ecef, dtype=np.float64).ravel()
        _dev.pf_device_position_update(self._state, float(ref[0]), float(ref[1]),
                                       float(ref[2]), float(sigma_pos))

    def shift_clock_bias(self, shift: float):
        """Shift every particle's clock bias by a constant (O(1): moves the anchor)."""
        self._require_init()
        _dev.pf_device_shift_clock_bias(self._state, float(shift))

    def recenter_position(self, ref_ecef, max_shift_m=10000.0):
        """Shift the particle cloud centre to ``ref_ecef`` when within a cap."""
        self._require_init()
        ref = np.asarray(ref_ecef, dtype=np.float64).ravel()
        if ref.size != 3:
            raise ValueError("ref_ecef must have shape (3,)")
        if not np.all(np.isfinite(ref)):
            return float("inf"), False
        est = np.asarray(self.estimate(), dtype=np.float64)
        delta = ref[:3] - est[:3]
        shift_norm = float(np.linalg.norm(delta))
        cap = float(max_shift_m)
        if cap > 0.0 and shift_norm > cap:
            return shift_norm, False
        _dev.pf_device_shift_position(self._state, float(delta[0]), float(delta[1]),
                                      float(delta[2]))
        return shift_norm, True

    def correct_clock_bias(self, sat_ecef, pseudoranges, quantile=0.5):
        """Re-centre the clock bias on a quantile of pseudorange residuals."""
        self._require_init()
        est = np.asarray(self.estimate(), dtype=np.float64)
        pos = est[:3]
        current_cb = est[3]
        sat = np.asarray(sat_ecef, dtype=np.float64).reshape(-1, 3)
        pr = np.asarray(pseudoranges, dtype=np.float64).ravel()
        ranges = np.linalg.norm(sat - pos, axis=1)
        residuals = pr - ranges
        q = float(np.clip(quantile, 0.0, 1.0))
        expected_cb = float(np.quantile(residuals, q))
        _dev.pf_device_shift_clock_bias(self._state, expected_cb - current_cb)

    # -- resampling ------------------------------------------------------------
    def _resample(self):
        if self.resampling == "megopolis":
            if self.megopolis_mode == "chesser":
                _dev.pf_device_resample_megopolis_chesser(self._state, self.megopolis_iterations)
            else:
                _dev.pf_device_resample_megopolis(self._state, self.megopolis_iterations,
                                                  self.seed)
        else:
            _dev.pf_device_resample_systematic(
                self._state, self.seed, 0 if self.systematic_u0 == "reference" else 1)

    def resample_if_needed(self):
        """Resample when ESS < ess_threshold * n_particles; returns True if it ran."""
        ess = self.get_ess()
        if ess < self.ess_threshold * self.n_particles:
            self._resample()
            return True
        return False

    # -- results ---------------------------------------------------------------
    def estimate(self):
        """Weighted mean [x, y, z, clock_bias] (ECEF, metres)."""
        self._require_init()
        return _dev.pf_device_estimate(self._state)

    def estimate_full(self):
        """Weighted mean [x, y, z, clock_bias, isb...]."""
        self._require_init()
        return _dev.pf_device_estimate_full(self._state)

    def get_particles(self):
        """All particles as an (N, 4) array of [x, y, z, clock_bias]."""
        self._require_init()
        return _dev.pf_device_get_particles(self._state)

    def get_particle_states(self):
        """(N, 16) rows [x, y, z, cb, vx, vy, vz, Sigma_v (3x3)]."""
        self._require_init()
        return _dev.pf_device_get_particle_states(self._state)

    def set_particle_states(self, states):
        """Replace the (N, 16) states and reset the weights."""
        self._require_init()
        arr = np.asarray(states, dtype=np.float64)
        if arr.shape != (self.n_particles, 16):
            raise ValueError(f"states shape {arr.shape} != ({self.n_particles}, 16)")
        _dev.pf_device_set_particle_states(self._state, np.ascontiguousarray(arr))

    def get_log_weights(self):
        """Per-particle log-weights (synchronises the stream)."""
        self._require_init()
        out = np.empty(self.n_particles, dtype=np.float64)
        _dev.pf_device_get_log_weights(self._state, out)
        return out

    def set_log_weights(self, log_weights):
        """Replace the per-particle log-weights on the device."""
        self._require_init()
        lw = np.asarray(log_weights, dtype=np.float64).reshape(-1)
        if lw.size != self.n_particles:
            raise ValueError(f"log_weights size {lw.size} != {self.n_particles}")
        _dev.pf_device_set_log_weights(self._state, np.ascontiguousarray(lw))

    def position_mixture_update(self, refs_ecef, sigma_pos=1.0, mixture_weights=None):
        """Gaussian-mixture position likelihood (host-side, via get/set log-weights)."""
        self._require_init()
        refs = np.asarray(refs_ecef, dtype=np.float64).reshape(-1, 3)
        if refs.shape[0] == 0:
            return
        refs = refs[np.all(np.isfinite(refs), axis=1)]
        if refs.shape[0] == 0:
            return
        sigma = max(float(sigma_pos), 1.0e-6)
        if mixture_weights is None:
            mix = np.full(refs.shape[0], 1.0 / refs.shape[0], dtype=np.float64)
        else:
            mix = np.asarray(mixture_weights, dtype=np.float64).reshape(-1)
            if mix.size != refs.shape[0]:
                raise ValueError("mixture_weights length must match refs_ecef")
            mix = np.where(np.isfinite(mix) & (mix > 0.0), mix, 0.0)
            total = float(np.sum(mix))
            mix = (np.full(refs.shape[0], 1.0 / refs.shape[0], dtype=np.float64)
                   if total <= 0.0 else mix / total)
        particles = np.asarray(self.get_particles(), dtype=np.float64)
        log_weights = np.asarray(self.get_log_weights(), dtype=np.float64)
        diff = particles[:, np.newaxis, :3] - refs[np.newaxis, :, :]
        with np.errstate(divide="ignore"):
            log_components = (np.log(mix[np.newaxis, :])
                              - 0.5 * np.sum(diff * diff, axis=2) / (sigma * sigma))
        max_log = np.max(log_components, axis=1)
        finite = np.isfinite(max_log)
        increment = np.full(log_weights.shape, -np.inf, dtype=np.float64)
        increment[finite] = max_log[finite] + np.log(
            np.sum(np.exp(log_components[finite] - max_log[finite, np.newaxis]), axis=1))
        self.set_log_weights(log_weights + increment)

    def get_resample_ancestors(self):
        """Ancestors of the last systematic resample: out[j] = source slot of slot j."""
        self._require_init()
        if self.resampling != "systematic":
            raise RuntimeError("get_resample_ancestors requires resampling='systematic'")
        out = np.empty(self.n_particles, dtype=np.int32)
        _dev.pf_device_get_resample_ancestors(self._state, out)
        return out

    def get_ess(self):
        """Effective sample size (computed on the GPU; scalar to the host)."""
        self._require_init()
        return _dev.pf_device_ess(self._state)

    def get_position_spread(self, center=None):
        """Weighted RMS particle distance to ``center`` (default: the estimate)."""
        self._require_init()
        if center is None:
            center_arr = np.asarray(self.estimate()[:3], dtype=np.float64)
        else:
            center_arr = np.asarray(center, dtype=np.float64).ravel()
            if center_arr.size != 3:
                raise ValueError("center must have shape (3,)")
        return _dev.pf_device_position_spread(self._state, float(center_arr[0]),
                                              float(center_arr[1]), float(center_arr[2]))

    def sync(self):
        """Wait for all queued GPU work of this filter."""
        if getattr(self, "_state", None) is not None:
            _dev.pf_device_sync(self._state)

    def set_stream(self, stream=None):
        """Run this filter's kernels and copies on ``stream`` (a
        torch.cuda.Stream or a raw hipStream_t), in order with the caller's
        other work there; None returns to the filter's own stream.  Work
        already queued is finished first.  Methods that return host values
        synchronise the stream in use."""
        _dev.pf_device_set_stream(self._state, stream)

    # -- forward-backward smoother (particle_filter_device.py:892-1251) ---------
    # Host orchestration over this class's own device updates: the forward pass
    # records each epoch's inputs, smooth() runs a second filter backwards over
    # them and averages the two tracks.  Epoch records are dicts with the
    # reference's keys, which its tests build directly.
    def enable_smoothing(self):
        """Start recording epochs (store_epoch) for a later smooth()."""
        self._smooth_epochs = []
        self._smooth_enabled = True

    def store_epoch(self, sat_ecef, pseudoranges, weights, velocity, dt, spp_ref=None,
                    dd_pseudorange=None, dd_pseudorange_sigma=None, dd_pseudorange_source=None,
                    dd_carrier=None, dd_carrier_sigma=None, carrier_anchor_pseudorange=None,
                    carrier_anchor_sigma=None, carrier_afv=None, carrier_afv_sigma=None,
                    carrier_afv_wavelength=None, doppler_update=None, doppler_sigma_mps=None,
                    doppler_velocity_update_gain=None, doppler_max_velocity_update_mps=None,
                    *, system_ids=None):
        """Record this epoch's forward estimate and the observations used (after
        update/estimate); a no-op until enable_smoothing().  ``system_ids``
        (keyword-only) replays multi-GNSS pseudoranges for filters with ISB states."""
        if not getattr(self, "_smooth_enabled", False):
            return
        rec = {
            "estimate": np.asarray(self.estimate(), dtype=np.float64)[:3].copy(),
            "sat_ecef": _f64(sat_ecef),
            "pseudoranges": _f64(pseudoranges),
            "weights": _f64(weights),
            "velocity": None if velocity is None else _f64(velocity),
            "dt": float(dt),
            "spp_ref": None if spp_ref is None else _f64(spp_ref),
            "dd_pseudorange": _dd_record(dd_pseudorange, _DD_PR_FIELDS),
            "dd_pseudorange_sigma": _opt_float(dd_pseudorange_sigma),
            "dd_pseudorange_source": (None if dd_pseudorange_source is None
                                      else str(dd_pseudorange_source)),
            "dd_carrier": _dd_record(dd_carrier, _DD_CP_FIELDS),
            "dd_carrier_sigma": _opt_float(dd_carrier_sigma),
            "carrier_anchor_pseudorange": _rows_record(
                carrier_anchor_pseudorange, ("sat_ecef", "pseudoranges", "weights"),
                "pseudoranges"),
            "carrier_anchor_sigma": _opt_float(carrier_anchor_sigma),
            "carrier_afv": _rows_record(
                carrier_afv, ("sat_ecef", "carrier_phase_cycles", "weights"),
                "carrier_phase_cycles"),
            "carrier_afv_sigma": _opt_float(carrier_afv_sigma),
            "carrier_afv_wavelength": _opt_float(carrier_afv_wavelength),
            "doppler_update": _doppler_record(doppler_update),
            "doppler_sigma_mps": _opt_float(doppler_sigma_mps),
            "doppler_velocity_update_gain": _opt_float(doppler_velocity_update_gain),
            "doppler_max_velocity_update_mps": _opt_float(doppler_max_velocity_update_mps),
        }
        if system_ids is not None:
            rec["system_ids"] = np.asarray(system_ids, dtype=np.int32).ravel().copy()
        self._smooth_epochs.append(rec)

    def smooth(self, position_update_sigma=None, skip_widelane_dd_pseudorange=False):
        """Backward pass over the recorded epochs; returns (smoothed, forward),
        both (n_epochs, 3) ECEF, smoothed = (forward + backward) / 2.

        The backward filter has this filter's configuration and seed + 1, starts
        at the last forward estimate (clock bias: median pseudorange residual
        there; spreads 10 m / 100 m) and visits the epochs last to first with the
        velocity negated, replaying per epoch: the DD pseudorange update (or,
        without one or for a wide-lane one when ``skip_widelane_dd_pseudorange``,
        the clock-bias re-centring and the pseudorange update), the DD carrier
        AFV, the carrier-anchor pseudoranges, the undifferenced carrier AFV and,
        given ``position_update_sigma``, the SPP position update.  Doppler rows
        are not replayed (they constrain the forward-time velocity)."""
        epochs = getattr(self, "_smooth_epochs", None)
        if not getattr(self, "_smooth_enabled", False) or not epochs:
            raise RuntimeError("No stored epochs. Call enable_smoothing() before forward pass.")
        forward = np.array([ep["estimate"] for ep in epochs], dtype=np.float64)
        last = epochs[-1]
        start = forward[-1]
        sat_last = np.asarray(last["sat_ecef"], dtype=np.float64).reshape(-1, 3)
        cb0 = float(np.median(np.asarray(last["pseudoranges"], dtype=np.float64)
                              - np.linalg.norm(sat_last - start, axis=1)))
        # looked up at call time, so a replacement of the module's class is used
        bwd = ParticleFilterDevice(**_backward_config(self))
        bwd.initialize(start, clock_bias=cb0, spread_pos=10.0, spread_cb=100.0,
                       velocity_init_sigma=float(getattr(self, "_velocity_init_sigma", 0.0)))
        backward = np.zeros_like(forward)
        for i in range(len(epochs) - 1, -1, -1):
            _replay_backward(self, bwd, epochs[i], position_update_sigma,
                             skip_widelane_dd_pseudorange)
            backward[i] = np.asarray(bwd.estimate(), dtype=np.float64)[:3]
        self._smooth_enabled = False
        self._smooth_epochs = []
        return (forward + backward) / 2.0, forward


# -- smoother records and the backward replay ----------------------------------
_DD_PR_FIELDS = ("dd_pseudorange_m", "sat_ecef_k", "sat_ecef_ref", "base_range_k",
                 "base_range_ref", "dd_weights")
_DD_CP_FIELDS = ("dd_carrier_cycles", "sat_ecef_k", "sat_ecef_ref", "base_range_k",
                 "base_range_ref", "dd_weights", "wavelengths_m")
_DOPPLER_WAVELENGTH = 0.19029367279836488
_CARRIER_WAVELENGTH = 0.190293673
# constructor keywords copied into the backward filter (particle_filter_device.py:1117-1144)
_CONFIG_KEYS = ("n_particles", "sigma_pos", "sigma_cb", "sigma_pr", "nu", "resampling",
                "ess_threshold", "per_particle_nlos_gate", "per_particle_nlos_dd_pr_threshold_m",
                "per_particle_nlos_dd_carrier_threshold_cycles",
                "per_particle_nlos_undiff_pr_threshold_m", "per_particle_huber",
                "per_particle_huber_dd_pr_k", "per_particle_huber_dd_carrier_k",
                "per_particle_huber_undiff_pr_k", "sigma_vel", "velocity_guide_alpha",
                "rbpf_velocity_kf", "velocity_process_noise")
# this backend's keyword-only additions, carried when the filter has them
_EXTRA_KEYS = ("n_isb", "sigma_isb", "megopolis_mode", "megopolis_iterations", "systematic_u0")


def _f64(a):
    return np.asarray(a, dtype=np.float64).copy()


def _opt_float(v):
    return None if v is None else float(v)


def _dd_record(dd, fields):
    """Copy of a DD result's arrays, ids and count (None passes through)."""
    if dd is None:
        return None
    rec = {k: _f64(getattr(dd, k)) for k in fields}
    rec["ref_sat_ids"] = tuple(getattr(dd, "ref_sat_ids", ()))
    rec["n_dd"] = int(dd.n_dd)
    return rec


def _rows_record(obs, keys, count_key):
    """Copy of a dict of per-satellite rows plus their count."""
    if obs is None:
        return None
    rec = {k: _f64(obs[k]) for k in keys}
    rec["n_sat"] = int(np.asarray(obs[count_key]).size)
    return rec


def _doppler_record(obs):
    if obs is None:
        return None
    rec = {k: _f64(obs[k]) for k in ("sat_ecef", "sat_vel", "doppler_hz", "weights")}
    rec["wavelength_m"] = float(obs.get("wavelength_m", _DOPPLER_WAVELENGTH))
    rec["n_sat"] = int(np.asarray(obs["doppler_hz"]).size)
    return rec


def _backward_config(pf):
    cfg = {k: getattr(pf, k) for k in _CONFIG_KEYS}
    cfg["seed"] = pf.seed + 1
    cfg.update({k: getattr(pf, k) for k in _EXTRA_KEYS if hasattr(pf, k)})
    return cfg


def _or(v, default):
    return default if v is None else v


def _replay_backward(pf, bwd, ep, position_update_sigma, skip_widelane):
    """One epoch of the backward pass (particle_filter_device.py:1152-1243)."""
    vel = ep["velocity"]
    bwd.predict(velocity=None if vel is None else -np.asarray(vel, dtype=np.float64),
                dt=ep["dt"])
    sat = np.asarray(ep["sat_ecef"], dtype=np.float64).reshape(-1, 3)
    dd = ep.get("dd_pseudorange")
    if dd is not None and not (skip_widelane and ep.get("dd_pseudorange_source") == "widelane"):
        bwd.update_dd_pseudorange(SimpleNamespace(**dd),
                                  sigma_pr=float(_or(ep.get("dd_pseudorange_sigma"), pf.sigma_pr)))
    else:
        bwd.correct_clock_bias(sat, ep["pseudoranges"])
        sys_ids = ep.get("system_ids")
        if sys_ids is None:
            bwd.update(sat, ep["pseudoranges"], weights=ep["weights"])
        else:
            bwd.update(sat, ep["pseudoranges"], weights=ep["weights"], system_ids=sys_ids)
    dd_cp = ep.get("dd_carrier")
    if dd_cp is not None:
        bwd.resample_if_needed()
        bwd.update_dd_carrier_afv(SimpleNamespace(**dd_cp),
                                  sigma_cycles=float(_or(ep.get("dd_carrier_sigma"), 0.05)))
    anchor = ep.get("carrier_anchor_pseudorange")
    if anchor is not None:
        bwd.update(anchor["sat_ecef"], anchor["pseudoranges"], weights=anchor["weights"],
                   sigma_pr=float(_or(ep.get("carrier_anchor_sigma"), pf.sigma_pr)))
    afv = ep.get("carrier_afv")
    if afv is not None:
        bwd.resample_if_needed()
        bwd.update_carrier_afv(afv["sat_ecef"], afv["carrier_phase_cycles"],
                               weights=afv["weights"],
                               wavelength=float(_or(ep.get("carrier_afv_wavelength"),
                                                    _CARRIER_WAVELENGTH)),
                               sigma_cycles=float(_or(ep.get("carrier_afv_sigma"), 0.05)))
    if position_update_sigma is not None and ep.get("spp_ref") is not None:
        bwd.position_update(np.asarray(ep["spp_ref"], dtype=np.float64)[:3],
                            sigma_pos=position_update_sigma)

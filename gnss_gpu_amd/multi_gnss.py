"""Multi-GNSS WLS positioning with inter-system bias (ISB) estimation
(the ``MultiGNSSSolver`` of python/gnss_gpu/multi_gnss.py:31-240).

Same constants, constructor, methods and return types as the reference class.
``solve`` calls the single-epoch solver and ``solve_batch`` the batched one
(both on the GPU, see ``_gnss_gpu_multi_gnss``); there is no CPU fallback.
"""

from __future__ import annotations

import numpy as np

from ._gnss_gpu_multi_gnss import wls_multi_gnss as _wls_multi_gnss
from ._gnss_gpu_multi_gnss import wls_multi_gnss_batch as _wls_multi_gnss_batch

SYSTEM_GPS = 0
SYSTEM_GLONASS = 1
SYSTEM_GALILEO = 2
SYSTEM_BEIDOU = 3
SYSTEM_QZSS = 4

_SYSTEM_NAMES = {
    SYSTEM_GPS: "GPS",
    SYSTEM_GLONASS: "GLONASS",
    SYSTEM_GALILEO: "Galileo",
    SYSTEM_BEIDOU: "BeiDou",
    SYSTEM_QZSS: "QZSS",
}

_PRN_LETTERS = {"G": SYSTEM_GPS, "R": SYSTEM_GLONASS, "E": SYSTEM_GALILEO, "C": SYSTEM_BEIDOU,
                "J": SYSTEM_QZSS}


class MultiGNSSSolver:
    """Weighted least squares for position plus one clock bias per enabled
    constellation (multi_gnss.py:31-53).

    ``systems`` lists the enabled system ids (default ``[SYSTEM_GPS]``); they
    are sorted and numbered 0.. in that order, which is the clock order of the
    native solvers.  A satellite of a system that is not enabled is solved with
    the first enabled system's clock (multi_gnss.py:87, :138).
    """

    def __init__(self, systems: list[int] | None = None, max_iter: int = 10,
                 tol: float = 1e-4):
        self.systems = sorted(systems or [SYSTEM_GPS])
        self.max_iter = max_iter
        self.tol = tol
        self._sys_to_idx = {s: i for i, s in enumerate(self.systems)}
        self.n_systems = len(self.systems)

    def _map_ids(self, system_ids):
        ids = np.asarray(system_ids).astype(np.int64, copy=False)
        mapped = np.zeros(ids.shape, dtype=np.int32)
        for s, i in self._sys_to_idx.items():
            mapped[ids == s] = i
        return mapped

    def solve(self, sat_ecef: np.ndarray, pseudoranges: np.ndarray, system_ids: np.ndarray,
              weights: np.ndarray | None = None) -> tuple[np.ndarray, dict[int, float], int]:
        """Single epoch (multi_gnss.py:55-101): returns (position [3],
        {system_id: clock bias [m]}, iterations); iterations is -1 with fewer
        than 3 + n_systems satellites."""
        sat_ecef = np.ascontiguousarray(sat_ecef, dtype=np.float64)
        pseudoranges = np.ascontiguousarray(pseudoranges, dtype=np.float64)
        n_sat = len(pseudoranges)
        if weights is None:
            weights = np.ones(n_sat, dtype=np.float64)
        else:
            weights = np.ascontiguousarray(weights, dtype=np.float64)
        n_state = 3 + self.n_systems
        if n_sat < n_state:
            return np.zeros(3, dtype=np.float64), {s: 0.0 for s in self.systems}, -1
        mapped = self._map_ids(system_ids).ravel()
        result, n_iter = _wls_multi_gnss(sat_ecef.ravel(), pseudoranges, weights, mapped,
                                         self.n_systems, self.max_iter, self.tol)
        position = result[:3]
        clock_biases = {self.systems[k]: float(result[3 + k]) for k in range(self.n_systems)}
        return position, clock_biases, int(n_iter)

    def solve_batch(self, sat_ecef: np.ndarray, pseudoranges: np.ndarray,
                    system_ids: np.ndarray, weights: np.ndarray | None = None
                    ) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Epoch batch (multi_gnss.py:103-160): returns (positions [E, 3],
        clock biases [E, n_systems], iterations int32 [E])."""
        sat_ecef = np.ascontiguousarray(sat_ecef, dtype=np.float64)
        pseudoranges = np.ascontiguousarray(pseudoranges, dtype=np.float64)
        n_epoch, n_sat = pseudoranges.shape
        if weights is None:
            weights = np.ones_like(pseudoranges)
        else:
            weights = np.ascontiguousarray(weights, dtype=np.float64)
        n_state = 3 + self.n_systems
        if n_sat < n_state:
            return (np.zeros((n_epoch, 3), dtype=np.float64),
                    np.zeros((n_epoch, self.n_systems), dtype=np.float64),
                    np.full(n_epoch, -1, dtype=np.int32))
        mapped = self._map_ids(system_ids).reshape(n_epoch, n_sat)
        results, n_iters = _wls_multi_gnss_batch(sat_ecef, pseudoranges, weights, mapped,
                                                 self.n_systems, self.max_iter, self.tol)
        return results[:, :3], results[:, 3:], n_iters

    @staticmethod
    def prn_to_system(prn_str: str) -> tuple[int, int]:
        """'G01' -> (SYSTEM_GPS, 1); unknown letters map to GPS (multi_gnss.py:218-236)."""
        return _PRN_LETTERS.get(prn_str[0], SYSTEM_GPS), int(prn_str[1:])

    @staticmethod
    def system_name(system_id: int) -> str:
        """Human-readable system name (multi_gnss.py:237-240)."""
        return _SYSTEM_NAMES.get(system_id, f"Unknown({system_id})")


__all__ = ["MultiGNSSSolver", "SYSTEM_GPS", "SYSTEM_GLONASS", "SYSTEM_GALILEO",
           "SYSTEM_BEIDOU", "SYSTEM_QZSS"]

"""The multi-port solve result and the swept wrappers.

Same surface as the reference's ``PortSolution`` (reference
src/momwire/_port_solution.py:21-73) and ``_SweptPortSolutions``
(_port_solution.py:76-155): ``y`` (n_ports x n_ports admittance), ``coeffs``
(n_dof x n_ports, column j = the B-spline amplitudes for a 1 V drive at port
j, every other port shorted), ``port_currents`` (the same object as ``y``) and
an opaque per-solve ``basis`` handle.
"""

from contextlib import contextmanager
from dataclasses import dataclass
from typing import Any

import numpy as np


@dataclass(frozen=True)
class PortSolution:
    y: np.ndarray
    coeffs: np.ndarray
    port_currents: np.ndarray
    basis: Any

    @property
    def n_ports(self) -> int:
        return int(self.y.shape[0])


class _SweptPortSolutions:
    """Swept entry points over one per-k generator (_port_solution.py:76-155)."""

    def _port_solutions_swept(self, k_array):
        raise NotImplementedError

    def _port_count(self):
        raise NotImplementedError

    def _set_k(self, kk):
        self.k = float(kk)
        self.omega = self.k * self.c
        self.wavelength = self.c / (self.omega / (2 * np.pi))

    @contextmanager
    def _k_restored(self):
        saved = (self.k, self.wavelength, self.omega)
        try:
            yield
        finally:
            self.k, self.wavelength, self.omega = saved

    def compute_port_solution_swept(self, k_array):
        return list(self._port_solutions_swept(np.asarray(k_array, dtype=float)))

    def compute_y_matrix_swept(self, k_array) -> np.ndarray:
        k_array = np.asarray(k_array, dtype=float)
        n_p = self._port_count()
        out = np.zeros((k_array.shape[0], n_p, n_p), dtype=np.complex128)
        for i, sol in enumerate(self._port_solutions_swept(k_array)):
            out[i] = sol.y
        return out

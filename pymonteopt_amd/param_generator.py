"""Random tissue parameters for the virtual OBD data sets (pymopt/utils/param_generator.py:7-75).

`generate_variable_params` keeps the reference's keys, defaults and draw order, all from
NumPy's global generator (`np.random`), so a script that seeds `np.random` gets the same
parameters from either package:
1. dermis thickness, mu_a and mu_s' (each uniform over its range);
2. subcutis thickness, mu_a and mu_s' (the same);
3. (BV/TV, cortical thickness) from one bivariate normal. The first list entry is the
   mean and the second the standard deviation; `corr` is the correlation coefficient.
"""
from __future__ import annotations

import numpy as np

from .voxel import set_params

__all__ = ["generate_variable_params"]


class generate_variable_params:
    """pymopt/utils/param_generator.py:7-75."""

    def __init__(self):
        self.params = {
            "th_dermis": [1, 2],
            "ma_dermis": [0.00633, 0.08560],
            "msd_dermis": [1.420, 2.506],
            "th_subcutaneus": [1, 6],
            "ma_subcutaneus": [0.005, 0.012],
            "msd_subcutaneus": [0.83, 1.396],
            "bv_tv": [0.115, 0.02],
            "th_cortical": [0.669, 0.133],
            "corr": 0.54,
        }
        self.keys = list(self.params.keys())

    def set_params(self, *initial_data, **kwargs):
        set_params(self.params, self.keys, *initial_data, **kwargs)

    def generate(self, n):
        dermis = self._get_dermis_params(n)
        subcut = self._get_subcut_params(n)
        bone = self._get_bone_params(n)
        self.int_params = {
            "th_dermis": dermis[0], "ma_dermis": dermis[1], "msd_dermis": dermis[2],
            "th_subcutaneus": subcut[0], "ma_subcutaneus": subcut[1], "msd_subcutaneus": subcut[2],
            "bv_tv": bone[0], "th_cortical": bone[1],
        }
        return self.int_params

    def get_variable_params(self):
        return self.int_params

    @staticmethod
    def _uniform_dist(range_, n):
        lo, hi = range_
        return np.random.rand(n) * (hi - lo) + lo

    def _three_uniform(self, prefix, n):
        return tuple(self._uniform_dist(self.params[k + prefix], n) for k in ("th", "ma", "msd"))

    def _get_dermis_params(self, n):
        return self._three_uniform("_dermis", n)

    def _get_subcut_params(self, n):
        return self._three_uniform("_subcutaneus", n)

    def _get_bone_params(self, n):
        (m0, s0), (m1, s1) = self.params["bv_tv"], self.params["th_cortical"]
        c = s0 * s1 * self.params["corr"]   # in this order, so the covariance rounds the same
        bvtv, cth = np.random.multivariate_normal([m0, m1], [[s0 ** 2, c], [c, s1 ** 2]], n).T
        return bvtv, cth

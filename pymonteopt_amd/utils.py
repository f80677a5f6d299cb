"""The helpers of pymopt.utils that code around the walk imports (pymopt/utils/__init__.py).

* `set_params`: the parameter contract of `build(**params)` (utilities.py:22-32);
* `calTime`: prints a run's wall time (utilities.py:9-16);
* `check_folder`: creates a directory if missing (utilities.py:18-20);
* `ToJsonEncoder`: json encoder for NumPy scalars and arrays (utilities.py:34-43);
* `generate_variable_params`: the virtual OBD subject sampler (param_generator.py);
* `readDicom`, `reConstArray_8`, `reConstArray`: DICOM stacks (readDICOM.py:17-64).

The plotting helper `correlationLine` and the keyword-argument deprecation decorator are
not provided.
"""
from __future__ import annotations

import json
import os

import numpy as np

from .dicom import readDicom, reConstArray, reConstArray_8  # noqa: F401
from .voxel import set_params  # noqa: F401

__all__ = ["set_params", "calTime", "check_folder", "ToJsonEncoder", "generate_variable_params", "readDicom",
           "reConstArray_8", "reConstArray"]


def __getattr__(name):
    if name == "generate_variable_params":
        from .param_generator import generate_variable_params
        return generate_variable_params
    raise AttributeError(name)


def calTime(end, start):
    """Prints end - start (seconds) as minutes and seconds, or hours and minutes."""
    q, mod = divmod(end - start, 60)
    if q < 60:
        print("Calculation time: %d minutes %0.3f seconds." % (q, mod))
    else:
        q2, mod2 = divmod(q, 60)
        print("Calculation time: %d h %0.3f minutes." % (q2, mod2))


def check_folder(folder_dir):
    os.makedirs(folder_dir, exist_ok=True)


class ToJsonEncoder(json.JSONEncoder):
    """NumPy integers, floats and arrays as JSON numbers and lists. (The reference's
    fallback names an undefined class; here other objects go to the base encoder.)"""

    def default(self, obj):
        if isinstance(obj, np.integer):
            return int(obj)
        if isinstance(obj, np.floating):
            return float(obj)
        if isinstance(obj, np.ndarray):
            return obj.tolist()
        return super().default(obj)

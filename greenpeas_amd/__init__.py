"""greenpeas_amd -- MI355X-native detector-error-model compiler.

Drop-in for the DEM compile path of abbrazi/GreenPeas
(python/greenpeas/__init__.py:3-24): ``get_driver``, ``Driver``,
``CorrelationLevel`` with the same meaning, plus batched compilation,
differentiable probabilities and a serving loop.  All compiles run on the GPU
through libgreenpeas_amd.so (include/greenpeas_amd.h).
"""

from . import error_analysis
from .dem import DetectorErrorModel
from .error_analysis import CorrelationLevel, Driver, get_driver

__all__ = ["CorrelationLevel", "DetectorErrorModel", "Driver", "error_analysis", "get_driver"]
__version__ = "0.1.0"

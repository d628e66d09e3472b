"""MI355X-native (gfx950) stiffness assembly and solve for pyfe3d models.

Public constants mirror pyfe3d/__init__.py:18-27.
"""
import numpy as np

DOF = 6
INT = np.int64
DOUBLE = np.float64

__version__ = '0.1.0'

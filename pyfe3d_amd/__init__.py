"""MI355X-native (gfx950) stiffness assembly and solve for pyfe3d models.

Modules (all compute runs in ``libpf3d.so``, the C ABI of include/pf3d.h;
there is no CPU fallback):

* ``compat``    -- pyfe3d's element classes (Quad4, Quad4R, Tria3R, BeamC,
                   BeamLR, Truss, Spring, their Data/Probe classes, BeamProp)
                   over the GPU kernels, per element or in one launch;
* ``assembly``  -- batched KC0 / KG / M / piston-theory / internal-force
                   entry points, mixed element types in one launch;
* ``sparse``    -- COO slots to 6x6 block-sparse rows (slot and element-block
                   plans);
* ``solver``    -- the static solve (block-Jacobi PCG in FP64 or mixed
                   precision, refinement, banded direct, multigrid CG);
* ``amg``       -- smoothed-aggregation multigrid (setup on the GPU) and its
                   CG; ``lobpcg`` / ``eigen`` -- natural frequencies and
                   linear buckling;
* ``sweep``     -- batched design sweeps (one graph per step); ``autotune``;
* ``shellprop`` -- laminate property records (host builders and the device
                   builder); ``panel`` / ``mesh`` -- the benchmark models;
* ``warmstart`` -- the learned initial guess for the sweep's PCG.

Public constants mirror pyfe3d/__init__.py:18-27.
"""
import numpy as np

DOF = 6
INT = np.int64
DOUBLE = np.float64

__version__ = '0.1.0'

"""tomoj_amd -- MI355X-native parallel-beam projector path for TomoJ.

Hand-written HIP kernels for gfx950 behind the C ABI in ``include/tomoj_hip.h``
(``libtomoj_hip.so``): Joseph forward projection with fused residual,
voxel-driven backprojection, OS-SART/SIRT/ART and weighted-backprojection
loops, mirroring TomoJ's ``Projector`` plugin interface; the TVM projector's
end-of-iteration total-variation denoising, the Bayesian projector's
distance-volume update and the BgART projector's median and collapse filters.
"""
from . import _lib
from ._lib import TjError, TjCancelled, load as load_library
from .projector import (BayesianVoxelProjector3DHIP, BgARTVoxelProjector3DHIP, TVMProjector3DHIP,
                        VoxelProjector3DHIP, phantom_ellipsoids)
from .reconstruction import tilt_euler, tilt_eulers, order_shuffle, ossart, wbp
from . import workloads
from .parameters import (AdvancedReconstructionParameters, ReconstructionParameters,
                         ReconstructionThreadManager, TiltSeriesData)

__all__ = ["VoxelProjector3DHIP", "TVMProjector3DHIP", "BayesianVoxelProjector3DHIP", "BgARTVoxelProjector3DHIP",
           "AdvancedReconstructionParameters", "TjError", "TjCancelled", "load_library",
           "phantom_ellipsoids", "tilt_euler", "tilt_eulers", "order_shuffle", "ossart", "wbp",
           "workloads", "ReconstructionParameters", "ReconstructionThreadManager", "TiltSeriesData"]

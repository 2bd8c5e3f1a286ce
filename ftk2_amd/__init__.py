"""ftk2_amd — MI355X (gfx950) simplicial feature extraction behind FTK2's
SimplicialEngine API.  The compute path is libftk2h.so (HIP kernels + C ABI,
include/ftk2h.h); this package holds the Python mirror of the reference interface
and the synthetic input generators."""
from .engine import (ContourPredicate, CriticalPointPredicate, FeatureComplex, FiberPredicate,  # noqa: F401
                     RegularSimplicialMesh, SimplicialEngine, extract)

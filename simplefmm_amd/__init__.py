"""sfmm: MI355X-native Laplace FMM evaluator (drop-in for the icyspirit/SimpleFMM evaluation path)."""
from .api import (BOTH, CURL, FAR, FP64, MIXED, NEAR, POTENTIAL, POTGRAD, Box, FMM3D, OctreePartitioner,  # noqa: F401
                  SfmmError, build, lib)

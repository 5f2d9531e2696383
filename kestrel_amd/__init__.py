"""kestrel_amd: MI355X (gfx950) pairwise Jacquard / kinship maximum-likelihood fit.

The compute path is libkestrel_amd.so (HIP kernels + C ABI, include/kestrel_amd.h);
this package is its Python host binding plus the seeded synthetic cohort generator.
"""
from . import synth
from ._lib import OUTPUT_DTYPE, KestrelError, Tune
from .api import (Coefficients, Cohort, build_arrays, kinship, relatedness_coefficients_gl,
                  relatedness_coefficients_gt, run_cli, segment_bounds, tune_default)

__all__ = ["Coefficients", "Cohort", "KestrelError", "OUTPUT_DTYPE", "Tune", "build_arrays", "kinship",
           "relatedness_coefficients_gl", "relatedness_coefficients_gt", "run_cli", "segment_bounds", "synth",
           "tune_default"]

"""LMM association on the device (mirrors jamma.lmm exports)."""

from jamma_amd.lmm.eigen import eigendecompose_kinship
from jamma_amd.lmm.io import IncrementalAssocWriter, write_assoc_results
from jamma_amd.lmm.runner_streaming import run_lmm_association_jax, run_lmm_association_streaming
from jamma_amd.lmm.stats import AssocResult

__all__ = ["AssocResult", "IncrementalAssocWriter", "eigendecompose_kinship",
           "run_lmm_association_jax", "run_lmm_association_streaming", "write_assoc_results"]

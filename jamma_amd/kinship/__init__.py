"""Centred kinship (GEMMA -gk 1): exact integer GEMM on MFMA (mirrors jamma.kinship exports)."""

from jamma_amd.io.plink import get_chromosome_partitions
from jamma_amd.kinship.compute import (
    compute_centered_kinship,
    compute_kinship_streaming,
    compute_loco_kinship,
    compute_loco_kinship_streaming,
)
from jamma_amd.kinship.io import (
    read_kinship_matrix,
    write_kinship_matrix,
    write_loco_kinship_matrices,
)

__all__ = ["compute_centered_kinship", "compute_kinship_streaming", "compute_loco_kinship",
           "compute_loco_kinship_streaming", "get_chromosome_partitions", "read_kinship_matrix",
           "write_kinship_matrix", "write_loco_kinship_matrices"]

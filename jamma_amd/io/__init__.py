"""PLINK / covariate input (GEMMA formats)."""

from jamma_amd.io.covariate import read_covariate_file
from jamma_amd.io.plink import (
    BedFile,
    get_chromosome_partitions,
    get_plink_metadata,
    validate_plink_dimensions,
)

__all__ = ["BedFile", "get_plink_metadata", "get_chromosome_partitions",
           "validate_plink_dimensions", "read_covariate_file"]

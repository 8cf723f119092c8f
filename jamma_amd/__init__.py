"""jamma_amd: MI355X-native LMM association hot path of JAMMA.

The public entry points keep the reference's signatures:
``jamma_amd.kinship.compute_centered_kinship``,
``jamma_amd.lmm.run_lmm_association_streaming`` and ``jamma_amd.gwas``.
"""

__version__ = "0.1.0"

from jamma_amd.gwas import GWASResult, gwas  # noqa: E402

__all__ = ["gwas", "GWASResult", "__version__"]

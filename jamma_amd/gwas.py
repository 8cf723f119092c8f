"""One-call GWAS API (mirrors src/jamma/gwas.py:21-167)."""

from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

from jamma_amd.lmm.stats import AssocResult
from jamma_amd.pipeline import PipelineConfig, PipelineRunner


@dataclass
class GWASResult:
    associations: list[AssocResult]
    n_samples: int
    n_snps_tested: int
    timing: dict[str, float] = field(default_factory=dict)


def gwas(
    bfile: str | Path,
    *,
    kinship_file: str | Path | None = None,
    covariate_file: str | Path | None = None,
    lmm_mode: int = 1,
    maf: float = 0.01,
    miss: float = 0.05,
    output_dir: str | Path = "output",
    output_prefix: str = "result",
    save_kinship: bool = False,
    check_memory: bool = True,
    show_progress: bool = True,
    loco: bool = False,
    eigenvalue_file: str | Path | None = None,
    eigenvector_file: str | Path | None = None,
    write_eigen: bool = False,
    phenotype_column: int = 1,
    snps_file: str | Path | None = None,
    ksnps_file: str | Path | None = None,
    hwe: float = 0.0,
    precision: str | None = None,
) -> GWASResult:
    """Kinship (or kinship file) -> eigendecomposition -> LMM association -> .assoc.txt."""
    opt = lambda p: Path(p) if p is not None else None  # noqa: E731
    cfg = PipelineConfig(
        bfile=Path(bfile), kinship_file=opt(kinship_file), covariate_file=opt(covariate_file),
        lmm_mode=lmm_mode, maf=maf, miss=miss, output_dir=Path(output_dir),
        output_prefix=output_prefix, save_kinship=save_kinship, check_memory=check_memory,
        show_progress=show_progress, loco=loco, eigenvalue_file=opt(eigenvalue_file),
        eigenvector_file=opt(eigenvector_file), write_eigen=write_eigen,
        phenotype_column=phenotype_column, snps_file=opt(snps_file), ksnps_file=opt(ksnps_file),
        hwe_threshold=hwe, precision=precision)
    r = PipelineRunner(cfg).run()
    return GWASResult(associations=r.associations, n_samples=r.n_samples,
                      n_snps_tested=r.n_snps_tested,
                      timing={k: r.timing[k] for k in ("kinship_s", "lmm_s", "total_s")})

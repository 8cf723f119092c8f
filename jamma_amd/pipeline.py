"""GWAS pipeline orchestration (mirrors src/jamma/pipeline.py:39-609).

Validation, phenotype parsing, kinship (computed on the device or read from a
.cXX.txt file), covariates and the streaming association, with the kinship and
its eigendecomposition kept on the device between the two stages.
"""

from __future__ import annotations

import logging
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from jamma_amd.io.covariate import read_covariate_file
from jamma_amd.io.plink import get_plink_metadata, resident_bed, validate_plink_dimensions
from jamma_amd.kinship.compute import kinship_streaming_device
from jamma_amd.kinship.io import read_kinship_matrix, write_kinship_matrix
from jamma_amd.lmm.runner_streaming import run_lmm_association_streaming
from jamma_amd.lmm.stats import AssocResult

logger = logging.getLogger("jamma_amd")


@dataclass
class PipelineConfig:
    bfile: Path
    kinship_file: Path | None = None
    covariate_file: Path | None = None
    lmm_mode: int = 1
    maf: float = 0.01
    miss: float = 0.05
    output_dir: Path = Path("output")
    output_prefix: str = "result"
    save_kinship: bool = False
    check_memory: bool = True
    show_progress: bool = True
    loco: bool = False
    eigenvalue_file: Path | None = None
    eigenvector_file: Path | None = None
    write_eigen: bool = False
    phenotype_column: int = 1
    snps_file: Path | None = None
    ksnps_file: Path | None = None
    hwe_threshold: float = 0.0
    precision: str | None = None


@dataclass
class PipelineResult:
    associations: list[AssocResult]
    n_samples: int
    n_snps_tested: int
    assoc_path: Path
    timing: dict[str, float] = field(default_factory=dict)


def _read_id_list(path: Path) -> list[str]:
    out = []
    for line in open(path):
        t = line.split()
        if t:
            out.append(t[0])
    return out


def _resolve(ids, sid) -> np.ndarray:
    pos = {s: i for i, s in enumerate(sid)}
    return np.array(sorted(pos[i] for i in ids if i in pos), dtype=np.int64)


class PipelineRunner:
    def __init__(self, config: PipelineConfig) -> None:
        self.config = config

    def validate_inputs(self) -> None:
        c = self.config
        for ext in (".bed", ".bim", ".fam"):
            p = Path(f"{c.bfile}{ext}")
            if not p.exists():
                raise FileNotFoundError(f"PLINK {ext} file not found: {p}")
        validate_plink_dimensions(c.bfile)
        if c.phenotype_column < 1:
            raise ValueError(f"phenotype_column must be >= 1 (1-based), got {c.phenotype_column}")
        if c.lmm_mode not in (1, 2, 3, 4):
            raise ValueError(f"lmm_mode must be 1 (Wald), 2 (LRT), 3 (Score), or 4 (All), got "
                             f"{c.lmm_mode}")
        if c.loco and c.kinship_file is not None:
            raise ValueError("-k and -loco are mutually exclusive in this version. LOCO computes "
                             "kinship internally.")
        if (c.eigenvalue_file is None) != (c.eigenvector_file is None):
            raise ValueError("Both -d (eigenvalues) and -u (eigenvectors) must be provided together")
        for p, what in ((c.kinship_file, "Kinship matrix"), (c.covariate_file, "Covariate"),
                        (c.snps_file, "SNP list"), (c.ksnps_file, "Kinship SNP list"),
                        (c.eigenvalue_file, "Eigenvalue"), (c.eigenvector_file, "Eigenvector")):
            if p is not None and not Path(p).exists():
                raise FileNotFoundError(f"{what} file not found: {p}")
        if c.hwe_threshold < 0:
            raise ValueError(f"hwe_threshold must be >= 0, got {c.hwe_threshold}")
        if c.hwe_threshold > 1.0:
            raise ValueError(f"hwe_threshold must be in [0, 1] (p-value threshold), got "
                             f"{c.hwe_threshold}")
        if c.hwe_threshold > 0 and c.loco:  # pipeline.py:234-238
            raise ValueError("-hwe is not yet supported with -loco mode. Apply HWE filtering as a "
                             "pre-processing step.")

    def parse_phenotypes(self) -> tuple[np.ndarray, int]:
        """.fam column 5 + phenotype_column; "-9"/"NA" -> NaN (pipeline.py:240-293)."""
        col = 4 + self.config.phenotype_column
        rows = [line.split() for line in open(f"{self.config.bfile}.fam") if line.strip()]
        if rows and col >= len(rows[0]):
            raise ValueError(f"phenotype_column {self.config.phenotype_column} exceeds available "
                             f"columns in .fam file")
        vals = np.array([np.nan if r[col] in ("-9", "NA") else float(r[col]) for r in rows])
        n_valid = int(np.sum(~np.isnan(vals) & (vals != -9)))
        if n_valid == 0:
            raise ValueError("No samples with valid phenotypes")
        return vals, n_valid

    def load_covariates(self, n_samples: int):
        if self.config.covariate_file is None:
            return None
        cov, _ = read_covariate_file(self.config.covariate_file)
        if cov.shape[0] != n_samples:
            raise ValueError(f"Covariate file has {cov.shape[0]} rows but PLINK data has "
                             f"{n_samples} samples. Covariate rows must match sample count exactly.")
        return cov

    def run(self) -> PipelineResult:
        """The reference's run (pipeline.py:420-609) with the .bed payload resident in HBM
        for the whole run: every pass (kinship statistics and Gram, association
        statistics and tiles) decodes from that one copy."""
        self.validate_inputs()
        with resident_bed(self.config.bfile):
            return self._run()

    def _run(self) -> PipelineResult:
        c = self.config
        t0 = time.perf_counter()
        meta = get_plink_metadata(c.bfile)
        n_samples, n_snps = meta["n_samples"], meta["n_snps"]
        phenotypes, n_analyzed = self.parse_phenotypes()
        snps_idx = _resolve(_read_id_list(c.snps_file), meta["sid"]) if c.snps_file else None
        ksnps_idx = _resolve(_read_id_list(c.ksnps_file), meta["sid"]) if c.ksnps_file else None
        Path(c.output_dir).mkdir(parents=True, exist_ok=True)
        assoc_path = Path(c.output_dir) / f"{c.output_prefix}.assoc.txt"
        if c.loco:
            from jamma_amd.lmm.loco import run_lmm_loco

            cov = self.load_covariates(n_samples)
            t1 = time.perf_counter()
            res = run_lmm_loco(c.bfile, phenotypes, covariates=cov, maf_threshold=c.maf,
                               miss_threshold=c.miss, lmm_mode=c.lmm_mode, output_path=assoc_path,
                               show_progress=c.show_progress, save_kinship=c.save_kinship,
                               kinship_output_dir=c.output_dir,
                               kinship_output_prefix=c.output_prefix, snps_indices=snps_idx,
                               ksnps_indices=ksnps_idx, precision=c.precision)
            t2 = time.perf_counter()
            return PipelineResult(res, n_analyzed, n_snps, assoc_path,
                                  {"kinship_s": 0.0, "load_s": 0.0, "lmm_s": t2 - t1,
                                   "total_s": t2 - t0})
        t1 = time.perf_counter()
        eigenvalues = eigenvectors = None
        K = None
        if c.eigenvalue_file is not None:
            from jamma_amd.lmm.eigen_io import read_eigen_files

            eigenvalues, eigenvectors = read_eigen_files(c.eigenvalue_file, c.eigenvector_file,
                                                         n_samples=n_analyzed)
        elif c.kinship_file is not None:
            K = read_kinship_matrix(c.kinship_file, n_samples=n_samples)
        else:
            K = kinship_streaming_device(c.bfile, ksnps_indices=ksnps_idx).cpu().numpy()
        if K is not None and c.save_kinship:
            write_kinship_matrix(K, Path(c.output_dir) / f"{c.output_prefix}.cXX.txt")
        if K is not None and c.write_eigen:
            # pipeline.py:549-566: eigendecompose the valid-sample kinship, write the GEMMA
            # eigen files, and hand the decomposition to the association (no second one)
            from jamma_amd.lmm.eigen import eigh_device
            from jamma_amd.lmm.eigen_io import write_eigen_files

            valid = ~np.isnan(phenotypes) & (phenotypes != -9.0)
            Kv = K if np.all(valid) else K[np.ix_(valid, valid)]
            d, U = eigh_device(Kv)
            write_eigen_files(d, U, c.output_dir, c.output_prefix)
            cov0 = self.load_covariates(n_samples)
            if cov0 is None or not np.any(np.isnan(cov0[valid])):
                eigenvalues, eigenvectors = d, U
            # else the runner also drops samples with missing covariates and decomposes
            # its own subset; the files hold the phenotype-valid samples, as the reference's
        kinship_s = time.perf_counter() - t1
        cov = self.load_covariates(n_samples)
        t2 = time.perf_counter()
        res = run_lmm_association_streaming(
            c.bfile, phenotypes, kinship=K, covariates=cov, eigenvalues=eigenvalues,
            eigenvectors=eigenvectors, maf_threshold=c.maf, miss_threshold=c.miss,
            output_path=assoc_path, lmm_mode=c.lmm_mode, show_progress=c.show_progress,
            snps_indices=snps_idx, hwe_threshold=c.hwe_threshold, precision=c.precision)
        t3 = time.perf_counter()
        return PipelineResult(res, n_analyzed, n_snps, assoc_path,
                              {"kinship_s": kinship_s, "load_s": t2 - t0, "lmm_s": t3 - t2,
                               "total_s": t3 - t0})

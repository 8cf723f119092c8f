"""CPU FP64 oracle for the JAMMA LMM association hot path.

TEST INFRASTRUCTURE ONLY. Nothing in ``jamma_amd`` imports this package: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker / CPU baseline, never as the thing
measured or shipped.

It is a NumPy/SciPy restatement of the reference algorithm (michael-denyer/jamma
v2.0.1). Every function cites the reference file:line it follows (paths are
relative to the reference repository root). The reference itself needs JAX,
jaxlib and bed-reader, none of which are installed, so it cannot be imported or
run; the restatement is pinned instead by the reference's own GEMMA golden
fixtures (copied as data under ``tests/golden/``), its known-answer unit tests,
and closed forms (see ``tests/test_oracle.py``).

Modules:
  plink    -- PLINK .bed/.bim/.fam decoding (bed-reader's count-A1 semantics)
  kinship  -- centred kinship, streaming two-pass kinship, LOCO kinship
  lmm      -- Uab/Pab, REML/MLE, grid + golden-section lambda search,
              Wald/Score/LRT statistics, null model, the streaming runner
"""

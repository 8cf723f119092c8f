"""The lambda-search kernel variants that wide spectra and many covariates select.

jm_lmm_assoc picks, by the number of spectrum terms T and covariate columns c:
  T <= 256 and small LDS tables   16 lanes per SNP (k_lmm_assoc16)
  T <= 256 otherwise              one wavefront per SNP, KT <= 4 terms per lane,
                                  up to 8 moment rows in registers
  256 < T <= 384                  wavefront per SNP, KT = 6, 4 rows in registers
  384 < T <= 512                  KT = 8, 4 rows in registers
  T > 512                         KT = 8 in blocks of 512 terms (multi-block)
and rows beyond those held in registers (c + 2 > 4 or 8) are read from memory; the
shared covariate/phenotype rows sit in LDS only when they fit 48 KB. The serving
form jm_lmm_assoc_multi takes the same kernels with problems of different c (and
modes) in one launch: for T <= 256 the 16-lane kernel, which splits each block's 16
SNPs into runs of one problem (the test interleaves problems SNP by SNP, so every
run is one SNP long and every block reloads its rows 16 times).

Cohorts are built from chosen spectra (random orthogonal basis, eigenvalues drawn
uniformly or log-uniformly) whose compressed plans land in each range, with c = 1,
4 and 11 covariate columns (intercept included). Every SNP is compared with the
oracle's chunk compute (runner_streaming.py:471-554, likelihood_jax.py:419-883) on
the same rotated inputs: lambda rtol 5e-5, -log10 p within 1e-5, logL_H1 rtol 1e-9
(the golden-section bracket width bounds the lambda and p agreement; a comparison
that ties at rounding level can take the other branch).
"""

import numpy as np
import pytest
import torch

from oracle import lmm as ol

pytestmark = pytest.mark.gpu

# name: (n, eigenvalue draw, low, high, seed, T range the plan must land in)
SPECTRA = {
    "kt4": (400, "uniform", 0.01, 5.0, 31, (193, 256)),
    "kt6": (800, "uniform", 0.01, 5.0, 32, (257, 384)),
    "kt8": (1500, "uniform", 0.01, 5.0, 33, (385, 512)),
    "mb": (1000, "loguniform", 0.001, 10.0, 34, (513, 1024)),
}
N_CVT = (1, 4, 11)
P = 24  # SNPs per cohort

_cache: dict = {}


def _cohort(name):
    if name in _cache:
        return _cache[name]
    from jamma_amd.lmm.engine import Cohort

    n, kind, lo, hi, seed, t_range = SPECTRA[name]
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        d = np.sort(rng.uniform(lo, hi, n))
    else:
        d = np.sort(np.exp(rng.uniform(np.log(lo), np.log(hi), n)))
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    co = Cohort(d, torch.as_tensor(Q, device="cuda"), precision="f64")
    assert t_range[0] <= co.n_terms <= t_range[1], (name, co.n_terms)
    # genotype-like columns (0/1/2 with a few mean-imputed values)
    X = rng.binomial(2, rng.uniform(0.05, 0.5, P), size=(n, P)).astype(np.float64)
    X[rng.random((n, P)) < 0.01] = 0.7
    UT = Q.T
    UtX = UT @ X
    Z = torch.as_tensor(np.ascontiguousarray(UtX), device="cuda")
    _cache[name] = (d, UT, UtX, co, Z, rng)
    return _cache[name]


def _problem(name, n_cvt):
    key = (name, n_cvt)
    if key in _cache:
        return _cache[key]
    from jamma_amd.lmm.engine import AssocProblem

    d, UT, UtX, co, Z, _ = _cohort(name)
    n = d.shape[0]
    rng = np.random.default_rng(1000 * n_cvt + n)
    W = np.column_stack([np.ones(n)] + [rng.normal(size=n) for _ in range(n_cvt - 1)])
    y = UT.T @ (rng.normal(size=n) * np.sqrt(0.6 * d + 0.4)) + W @ rng.normal(size=n_cvt)
    pb = AssocProblem(co, y, W)
    xm = co.moments(Z, False, P, P, pb.V, pb.n_cvt + 1)
    lam0, ll0 = ol.null_model_mle(d, UT @ W, UT @ y, n_cvt)
    ref = ol.assoc_chunk(4, n_cvt, d, UT @ W, UT @ y, UtX, 1.0 / (lam0 * d + 1.0), ll0)
    np.testing.assert_allclose(pb.lambda_null, lam0, rtol=1e-6)
    _cache[key] = (pb, xm, ref)
    return _cache[key]


COLS = {"l_remle": 3, "logl_H1": 4, "l_mle": 5, "p_wald": 2, "p_lrt": 7, "p_score": 8}
MODE_KEYS = {1: ("l_remle", "logl_H1", "p_wald"), 2: ("l_mle", "p_lrt"), 3: ("p_score",),
             4: ("l_remle", "logl_H1", "l_mle", "p_wald", "p_lrt", "p_score")}


def _check(out, ref, mode, what):
    keys = MODE_KEYS[mode]
    for k in keys:
        got, want = out[:, COLS[k]], ref[k]
        if k.startswith("l_"):
            np.testing.assert_allclose(got, want, rtol=5e-5, err_msg=f"{what} {k}")
        elif k == "logl_H1":
            np.testing.assert_allclose(got, want, rtol=1e-9, err_msg=f"{what} {k}")
        else:
            np.testing.assert_allclose(-np.log10(got), -np.log10(want), atol=1e-5,
                                       err_msg=f"{what} {k}")
    # the columns a mode does not compute are NaN
    for k, col in COLS.items():
        if k not in keys:
            assert np.all(np.isnan(out[:, col])), f"{what}: {k} should be NaN in mode {mode}"
    if mode in (1, 4):
        ok = np.isfinite(ref["beta"])
        np.testing.assert_allclose(out[ok, 0], ref["beta"][ok], rtol=0,
                                   atol=1e-5 * np.max(np.abs(ref["se"][ok])), err_msg=f"{what} beta")
        np.testing.assert_allclose(out[ok, 1], ref["se"][ok], rtol=5e-5, err_msg=f"{what} se")


@pytest.mark.parametrize("n_cvt", N_CVT)
@pytest.mark.parametrize("name", list(SPECTRA))
def test_single_problem_kernels(name, n_cvt):
    """jm_lmm_assoc at each (T, c): the kernel variant the launcher picks."""
    d, UT, UtX, co, Z, _ = _cohort(name)
    pb, xm, ref = _problem(name, n_cvt)
    print(f"{name}: n = {d.shape[0]}, T = {co.n_terms}, c = {n_cvt}", flush=True)
    out = pb.assoc_moments(xm, P, 4).cpu().numpy()
    _check(out, ref, 4, f"{name} c={n_cvt}")


@pytest.mark.parametrize("name", list(SPECTRA))
def test_multi_problem_batch(name):
    """jm_lmm_assoc_multi with c = 1, 4, 11 and modes 4, 1, 2, 3 interleaved SNP by
    SNP in one launch: every SNP matches the oracle of its own problem, and its
    columns equal those of the single-problem launch."""
    from jamma_amd import _lib
    from jamma_amd.lmm.engine import assoc_multi

    d, UT, UtX, co, Z, _ = _cohort(name)
    specs = [(1, 4), (4, 1), (11, 2), (4, 3)]  # (c, mode) of each problem
    probs, modes, xms, refs = [], [], [], []
    for c, mode in specs:
        pb, xm, ref = _problem(name, c)
        probs.append(pb)
        modes.append(mode)
        xms.append(xm)
        refs.append(ref)
    snp_prob = np.arange(P, dtype=np.int32) % len(specs)
    T = co.n_terms
    snp_xmom = np.array([xms[k].data_ptr() + s * (probs[k].n_cvt + 2) * T * 8
                         for s, k in enumerate(snp_prob)], dtype=np.uint64)
    out_d = torch.full((P, _lib.JM_NOUT), -7.0, dtype=torch.float64, device="cuda")
    print(f"{name}: n = {d.shape[0]}, T = {T}, problems {specs}", flush=True)
    assoc_multi(probs, modes, snp_prob, snp_xmom, out_d)
    out = out_d.cpu().numpy()
    for k, (c, mode) in enumerate(specs):
        sel = snp_prob == k
        ref = {key: v[sel] for key, v in refs[k].items()}
        _check(out[sel], ref, mode, f"{name} multi c={c} mode={mode}")
        single = probs[k].assoc_moments(xms[k], P, 4).cpu().numpy()[sel]
        for key in MODE_KEYS[mode]:
            col = COLS[key]
            rtol = 5e-5 if key.startswith("l_") else 1e-9
            np.testing.assert_allclose(out[sel, col], single[:, col], rtol=rtol,
                                       err_msg=f"{name} multi vs single c={c} {key}")


@pytest.mark.parametrize("name", list(SPECTRA))
def test_multi_result_bits_do_not_depend_on_the_batch(name):
    """With one width_cvt, a problem's SNPs give the same bits launched alone and
    interleaved with wider problems (the LDS layout follows the batch, the kernel
    variant only the width)."""
    from jamma_amd import _lib
    from jamma_amd.lmm.engine import assoc_multi

    d, UT, UtX, co, Z, _ = _cohort(name)
    specs = [(1, 4), (4, 4), (11, 4), (3, 1)]
    probs, modes, xms = [], [], []
    for c, mode in specs:
        pb, xm, _ = _problem(name, c)
        probs.append(pb)
        modes.append(mode)
        xms.append(xm)
    T = co.n_terms
    snp_prob = np.arange(P, dtype=np.int32) % len(specs)
    addr = np.array([xms[k].data_ptr() + s * (probs[k].n_cvt + 2) * T * 8
                     for s, k in enumerate(snp_prob)], dtype=np.uint64)
    mix = torch.full((P, _lib.JM_NOUT), -7.0, dtype=torch.float64, device="cuda")
    assoc_multi(probs, modes, snp_prob, addr, mix, width_cvt=11)
    mix = mix.cpu().numpy()
    for k, (c, mode) in enumerate(specs):
        sel = np.where(snp_prob == k)[0]
        alone = torch.full((len(sel), _lib.JM_NOUT), -7.0, dtype=torch.float64, device="cuda")
        assoc_multi([probs[k]], [mode], np.zeros(len(sel), np.int32), addr[sel], alone,
                    width_cvt=11)
        assert np.array_equal(alone.cpu().numpy(), mix[sel], equal_nan=True), (name, c)

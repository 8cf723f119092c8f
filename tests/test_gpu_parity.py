"""Parity of the HIP path with the oracle (GPU). Every call goes through the C
ABI of libjamma_hip.so (via jamma_amd); the oracle is only the checker.

Tolerances: integer/byte work (decode, counts) is bit-exact; FP64 results
must match the FP64 oracle to rounding (kinship < 1e-12 abs; association
statistics rtol 1e-6, lambda rtol 5e-5 and -log10 p atol 1e-5, the width of
the golden-section bracket, see _assert_matches_oracle); the
mixed (BF16x3) rotation path is held to |delta -log10 p| <= 1e-4 and 100%
significance / effect-direction agreement, the acceptance bar of the mixed path.
"""

from __future__ import annotations

import os
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import kinship as ok
from oracle import lmm as ol
from oracle import plink as opl
from tests.util import GOLDEN, balding_nichols, fam_phenotypes, load_assoc, phenotype, synthetic_prefix

pytestmark = pytest.mark.gpu


def _dev():
    from jamma_amd.device import require_gpu

    return require_gpu()


# ------------------------------------------------------------------ decode
def test_decode_bed_exact():
    from jamma_amd.io.plink import decode

    rng = np.random.default_rng(1)
    n_tot, p = 203, 97
    G = rng.integers(0, 3, size=(n_tot, p)).astype(float)
    G[rng.random(G.shape) < 0.07] = np.nan
    raw = np.frombuffer(opl.encode_bed(G)[3:], dtype=np.uint8)
    bps = (n_tot + 3) // 4
    raw_d = torch.from_numpy(raw.copy()).to(_dev())
    sample_idx = np.sort(rng.choice(n_tot, size=150, replace=False)).astype(np.int32)
    snp_idx = np.array([0, 5, 96, 40, 41], dtype=np.int32)
    g, counts, ld = decode(raw_d, bps, len(snp_idx), sample_idx=sample_idx, snp_idx=snp_idx)
    g = g.cpu().numpy()
    counts = counts.cpu().numpy()
    ref = G[np.ix_(sample_idx, snp_idx)].T
    assert np.array_equal(g[:, :150], np.where(np.isnan(ref), -1, ref).astype(np.int8))
    assert np.all(g[:, 150:ld] == 0)
    exp = np.stack([(ref == 0).sum(1), (ref == 1).sum(1), (ref == 2).sum(1),
                    np.isnan(ref).sum(1)], 1)
    assert np.array_equal(counts, exp)


# ------------------------------------------------------------------ eigendecomposition
@pytest.mark.parametrize("n", [1, 7, 300, 1031])
def test_eigh_device_matches_lapack(n):
    """rocSOLVER dsytrd + dorgtr + dstedc on the device vs LAPACK dsyevd (numpy): same ascending
    spectrum, U diag(d) U^T reproduces K, U orthonormal."""
    from jamma_amd.lmm.eigen import eigh_device

    G = balding_nichols(n, 3 * n + 5, seed=n)
    K = ok.centered_kinship(G) if n > 1 else np.array([[0.7]])
    d, U = eigh_device(K)
    d_ref = np.linalg.eigvalsh(K)
    d_ref = np.where(np.abs(d_ref) < 1e-10, 0.0, d_ref)
    scale = max(1.0, float(np.abs(d_ref).max()))
    assert np.all(np.diff(d) >= -1e-12 * scale)
    assert np.max(np.abs(d - d_ref)) < 1e-11 * scale
    Uh = U.cpu().numpy()
    assert np.max(np.abs(Uh.T @ Uh - np.eye(n))) < 1e-11
    assert np.max(np.abs((Uh * d) @ Uh.T - K)) < 1e-11 * scale
    assert U.t().is_contiguous()


# ------------------------------------------------------------------ kinship
def test_kinship_gemma_fixture_streaming():
    from jamma_amd.kinship import compute_kinship_streaming

    K = compute_kinship_streaming(synthetic_prefix(), chunk_size=173)
    Kref = np.loadtxt(GOLDEN / "gemma_synthetic" / "gemma_kinship.cXX.txt")
    assert np.max(np.abs(K - Kref)) < 1e-10  # tests/test_gemma_integration.py:79


@pytest.mark.parametrize("n,p,miss,batch", [(300, 2500, 0.01, 1000), (129, 700, 0.0, 10000),
                                            (517, 1333, 0.05, 400)])
def test_kinship_int8_matches_oracle(n, p, miss, batch):
    from jamma_amd.kinship import compute_centered_kinship

    G = balding_nichols(n, p, miss=miss, seed=n)
    K = compute_centered_kinship(G, batch_size=batch)
    Kref = ok.centered_kinship(G)
    assert np.max(np.abs(K - Kref)) < 1e-12
    assert np.array_equal(K, K.T)


@pytest.mark.parametrize("form", ["2bit", "2bit-i8mfma", "int8"])
@pytest.mark.parametrize("n,tiles,stage_cols", [
    (600, [300, 300, 700, 1200, 37], 512),  # flush when full, stale columns, oversized tiles
    (5700, [3100, 1500], None),             # 276 tiles: a full round + K-split tail round
    (129, [70], None),                      # one tile, fewer K steps than a split needs
])
def test_syrk_int32_exact(n, tiles, stage_cols, form, monkeypatch):
    """S = X0 X0^T (missing calls as 0) is exact integer arithmetic: compare the lower
    triangle with a float64 product of the same integers (exact below 2^53), for the
    2-bit packed operand (tiles on 16-SNP boundaries, gaps of zero genotypes) on the FP4
    and on the int8 MFMA, and the int8 operand."""
    from jamma_amd.device import round_up, to_dev
    from jamma_amd.kinship.compute import KinshipAccumulator

    monkeypatch.setenv("JM_KINSHIP_INT8", "1" if form == "int8" else "0")
    monkeypatch.setenv("JM_KINSHIP_SYRK_I8MFMA", "1" if form == "2bit-i8mfma" else "0")
    rng = np.random.default_rng(n)
    p = sum(tiles)
    g = rng.integers(0, 3, size=(p, n)).astype(np.int8)
    g[rng.random(g.shape) < 0.01] = -1
    per_snp = 1.0 if form == "int8" else 0.25
    stage = 1 << 31 if stage_cols is None else int(stage_cols * round_up(n, 256) * per_snp)
    acc = KinshipAccumulator(n, stage_bytes=stage)
    assert acc.packed == (form != "int8")
    if stage_cols is not None:
        assert acc.stage_cols == stage_cols
    ld = round_up(n, 64)
    s0 = 0
    for t in tiles:
        tile = np.zeros((t, ld), np.int8)
        tile[:, :n] = g[s0:s0 + t]
        acc.add_tile(to_dev(tile, torch.int8), ld, t, np.zeros(t), False)
        s0 += t
    S = acc.S.cpu().numpy()
    x0 = np.where(g < 0, 0, g).astype(np.float64)
    ref = (x0.T @ x0).astype(np.int64)
    low = np.tril_indices(n)
    assert np.array_equal(S[low], ref[low])


@pytest.mark.parametrize("form", ["fp4", "i8mfma"])
def test_syrk_2bit_exact_past_2_24(form):
    """The FP4-MFMA product sums in FP32, exact while a launch covers at most 2^22 SNPs;
    over 2^22 + 768 SNPs of codes 2 (and 1 for half of the samples) every S entry is
    K c_i c_j, past 2^24 for the code-2 pairs, so the K split across launches has to be
    right. The int8-MFMA kernel on the same operand gives the same integers."""
    import subprocess
    import sys
    import textwrap

    # a child process with the kernel choice in its environment
    code = textwrap.dedent("""
        import numpy as np, torch
        from jamma_amd import _lib
        from jamma_amd.device import ptr, stream_ptr
        n, K = 300, (1 << 22) + 768
        ldp = (K + 255) // 256 * 64
        gP = torch.empty((512, ldp), dtype=torch.uint8, device="cuda")
        gP[:150] = 0xAA   # every 2-bit code 2
        gP[150:] = 0x55   # every 2-bit code 1
        S = torch.zeros((n, n), dtype=torch.int32, device="cuda")
        _lib.call("jm_kinship_syrk_2bit", ptr(gP), ldp, n, K, ptr(S), stream_ptr())
        torch.cuda.synchronize()
        c = np.where(np.arange(n) < 150, 2, 1).astype(np.int64)
        ref = K * np.outer(c, c)
        low = np.tril_indices(n)
        got = S.cpu().numpy().astype(np.int64)
        assert ref.max() > 1 << 24
        assert np.array_equal(got[low], ref[low]), np.abs(got[low] - ref[low]).max()
        print("ok")
    """)
    env = dict(os.environ, JM_KINSHIP_SYRK_I8MFMA="1" if form == "i8mfma" else "0")
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=300, cwd=str(Path(__file__).resolve().parents[1]))
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr


@pytest.mark.parametrize("n,n_snps,col0", [(70, 37, 48), (300, 600, 0), (257, 255, 16)])
def test_pack_gram_2bit_layout(n, n_snps, col0):
    """The packed operand holds SNP c's code (0/1/2, missing 0) of each sample in the
    interleaved word position the GEMM reads, zeros up to the 256-SNP group end, and
    leaves the columns before col0 and the padded rows untouched."""
    from jamma_amd import _lib
    from jamma_amd.device import ptr, round_up, stream_ptr, to_dev

    rng = np.random.default_rng(n_snps)
    ld_g = round_up(n, 64)
    g = np.zeros((n_snps, ld_g), np.int8)
    g[:, :n] = rng.integers(0, 3, size=(n_snps, n))
    g[:, :n][rng.random((n_snps, n)) < 0.05] = -1
    end = round_up(col0 + n_snps, 256)
    ldp = round_up(end, 256) // 4 + 64
    gP = torch.full((round_up(n, 256), ldp), 0xAB, dtype=torch.uint8, device="cuda")
    _lib.call("jm_pack_gram_2bit", ptr(to_dev(g, torch.int8)), ld_g, n_snps, n, ptr(gP), ldp,
              col0, stream_ptr())
    P = gP.cpu().numpy()
    words = P[:, : ldp // 4 * 4].view(np.uint32)  # (rows, ldp / 4) words
    codes = np.full((P.shape[0], end), -9, np.int64)
    for Q in range(col0 // 16, end // 16):
        off = (Q >> 4) * 64 + 16 * (Q & 3) + 4 * ((Q >> 2) & 3)
        w = words[:, off // 4].astype(np.int64)
        for e in range(16):
            codes[:, 16 * Q + e] = (w >> (2 * e)) & 3
    want = np.zeros((n, end - col0), np.int64)
    want[:, :n_snps] = np.where(g[:, :n] < 0, 0, g[:, :n]).T
    assert np.array_equal(codes[:n, col0:end], want)
    # bytes outside the written words (before col0, after the group end, padded rows)
    untouched = np.ones_like(P, dtype=bool)
    for Q in range(col0 // 16, end // 16):
        off = (Q >> 4) * 64 + 16 * (Q & 3) + 4 * ((Q >> 2) & 3)
        untouched[:n, off:off + 4] = False
    assert np.all(P[untouched] == 0xAB)


def test_loco_kinship_matches_oracle():
    from jamma_amd.kinship import compute_loco_kinship

    G = balding_nichols(150, 900, miss=0.02, seed=4)
    chrom = np.array([str(1 + i * 3 // 900) for i in range(900)])
    got = dict(compute_loco_kinship(G, chrom))
    ref = ok.loco_kinship(G, chrom)
    assert set(got) == set(ref)
    for c in ref:
        assert np.max(np.abs(got[c] - ref[c])) < 1e-12


# ------------------------------------------------------------------ rotation
@pytest.mark.parametrize("planes", [2, 3])
@pytest.mark.parametrize("n,p", [(100, 37), (333, 700), (520, 1100), (1000, 257)])
def test_rotation_f64_and_mixed(n, p, planes):
    from jamma_amd.lmm.engine import Cohort
    from jamma_amd.device import round_up, to_dev

    rng = np.random.default_rng(n)
    U = np.linalg.qr(rng.normal(size=(n, n)))[0]
    d = np.sort(rng.uniform(0.1, 3, n))
    G = rng.integers(0, 3, size=(n, p)).astype(float)
    G[rng.random(G.shape) < 0.02] = np.nan
    means = np.nanmean(G, axis=0).astype(np.float32).astype(np.float64)
    X = np.where(np.isnan(G), means, G)
    ref = U.T @ X
    co = Cohort(d, U, precision="both", bf16_planes=planes)
    ld = round_up(n, 64)
    g8 = np.zeros((p, ld), np.int8)
    g8[:, :n] = np.where(np.isnan(G), -1, G).astype(np.int8).T
    gd, md = to_dev(g8, torch.int8), to_dev(means, torch.float64)
    Z64, _, _ = co.rotate_genotypes(gd, ld, md, p, "f64")
    assert np.max(np.abs(Z64.cpu().numpy() - ref)) < 1e-12 * np.abs(ref).max()
    Z32, _, _ = co.rotate_genotypes(gd, ld, md, p, "mixed")
    err = np.abs(Z32.cpu().numpy().astype(np.float64) - ref)
    # Elementwise bound: hi + lo leaves |U - hi - lo| <= 2^-17 |U| (two bf16 roundings of
    # 8 stored bits each), hi + mid + lo 2^-25 |U|; the mean's split and the FP32
    # accumulation add rounding of the same scale, so |Z - ref| <= c (|U|^T |X|).
    scale = np.abs(U).T @ np.abs(X)
    ratio = (err / scale).max()
    print(f"planes={planes}: max |Z - ref| / (|U|^T|X|) = {ratio:.3e} (2^{np.log2(ratio):.1f})")
    assert ratio < (2.0 ** -16 if planes == 2 else 2.0 ** -18)
    # gathered columns (serving batches): repeated and permuted SNP rows
    idx = rng.integers(0, p, size=p + 45).astype(np.int32)
    cd = to_dev(idx, torch.int32)
    Zg, _, _ = co.rotate_genotypes(gd, ld, md, len(idx), "f64", col_idx=cd)
    assert np.max(np.abs(Zg.cpu().numpy() - ref[:, idx])) < 1e-12 * np.abs(ref).max()
    Zg, _, _ = co.rotate_genotypes(gd, ld, md, len(idx), "mixed", col_idx=cd)
    # the gather must reproduce the ungathered columns bit for bit
    assert np.array_equal(Zg.cpu().numpy(), Z32.cpu().numpy()[:, idx])


# ------------------------------------------------------------------ association
def _fixture_data():
    pre = synthetic_prefix()
    G = opl.read_bed(pre)
    y = fam_phenotypes(pre)
    K = np.loadtxt(GOLDEN / "gemma_synthetic" / "gemma_kinship.cXX.txt")
    C = np.loadtxt(GOLDEN / "gemma_covariate" / "covariates.txt")
    return G, y, K, C


def _snp_info(p):
    return [{"chr": "1", "rs": f"rs{j:04d}", "pos": 1000 * (j + 1), "a1": "A", "a0": "G"}
            for j in range(p)]


def _arrays(results, keys):
    return {k: np.array([getattr(r, k) for r in results], dtype=float) for k in keys}


def _assert_matches_oracle(got, ref, keys, rtol=1e-6, lam_rtol=5e-5):
    """FP64 device vs FP64 oracle. Evaluations agree to ~1e-14; a golden-section
    comparison that is a tie at that level may take the other branch, which moves
    lambda by at most one bracket step (rtol 5e-5, the reference's own JAX-vs-GEMMA
    bar, tests/test_lmm_validation.py:32). The final golden bracket is
    0.94 * 0.618^20 ~ 6e-5 wide in log(lambda), and the Wald statistic is not
    stationary in lambda, so such a branch moves -log10 p by up to ~1e-5: that is
    the p-value bar (observed: 3.1e-6 on one SNP of the covariate fixture). Such a
    step moves beta and se by (d/dlog lambda) x 6e-5, a few 1e-6 of a standard error:
    SNPs whose REML lambda differs by more than 1e-9 are held to 1e-5 se there, the
    others to 1e-7 se (plus rtol)."""
    flip = None
    if "l_remle" in keys and "se" in keys:
        flip = ~np.isclose(got["l_remle"], np.asarray(ref["l_remle"], dtype=float), rtol=1e-9,
                           atol=0.0)
        print(f"golden branches that differ: {int(flip.sum())} of {flip.size}", flush=True)
    for k in keys:
        a, e = got[k], np.asarray(ref[k], dtype=float)
        if k in ("l_remle", "l_mle"):
            np.testing.assert_allclose(a, e, rtol=lam_rtol, err_msg=k)
        elif k.startswith("p_"):
            dl = np.abs(np.log10(a) - np.log10(e))
            print(f"{k}: max |dlog10 p| = {np.nanmax(dl):.3e}", flush=True)
            np.testing.assert_allclose(-np.log10(a), -np.log10(e), atol=1e-5, err_msg=k)
        elif k in ("logl_H1",):
            np.testing.assert_allclose(a, e, rtol=1e-9, err_msg=k)
        elif k in ("beta", "se") and flip is not None:
            # beta = P_xy / P_xx with P_xy a covariate-projected (cancelling) sum: its
            # rounding is absolute on the standard-error scale, not relative to beta
            se = np.abs(np.asarray(ref["se"], dtype=float))
            bound = rtol * np.abs(e) + np.where(flip, 1e-5, 1e-7) * se
            dev = np.abs(a - e)
            assert np.all((dev <= bound) | (np.isnan(a) & np.isnan(e))), \
                (k, float(np.nanmax(dev / se)))
        else:
            np.testing.assert_allclose(a, e, rtol=rtol, atol=1e-12, err_msg=k)


ALL = ("beta", "se", "p_wald", "l_remle", "logl_H1", "l_mle", "p_lrt", "p_score")


@pytest.mark.parametrize("cov", [False, True])
def test_assoc_fixture_matches_oracle_and_gemma(cov):
    from jamma_amd.lmm import run_lmm_association_jax

    G, y, K, C = _fixture_data()
    res = run_lmm_association_jax(G.astype(np.float32), y, K, _snp_info(G.shape[1]),
                                  covariates=C if cov else None, lmm_mode=4, precision="f64")
    got = _arrays(res, ALL)
    ref = ol.run_lmm(G.astype(np.float32), y, kinship=K, covariates=C if cov else None, lmm_mode=4)
    assert len(res) == len(ref["snp_indices"])
    _assert_matches_oracle(got, ref, ALL)
    gem = load_assoc(GOLDEN / "gemma_all_tests" / ("gemma_all_covar.assoc.txt" if cov else
                                                   "gemma_all.assoc.txt"))
    np.testing.assert_allclose(got["p_wald"], gem["p_wald"], rtol=1e-4)
    np.testing.assert_allclose(got["p_score"], gem["p_score"], rtol=1e-4)
    np.testing.assert_allclose(got["p_lrt"], gem["p_lrt"], rtol=5e-3)
    np.testing.assert_allclose(got["se"], gem["se"], rtol=1e-5)


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_assoc_single_modes(mode):
    from jamma_amd.lmm import run_lmm_association_jax

    G, y, K, _ = _fixture_data()
    res = run_lmm_association_jax(G, y, K, _snp_info(G.shape[1]), lmm_mode=mode, precision="f64")
    ref = ol.run_lmm(G, y, kinship=K, lmm_mode=mode)
    keys = {1: ("beta", "se", "p_wald", "l_remle", "logl_H1"), 2: ("l_mle", "p_lrt"),
            3: ("beta", "se", "p_score")}[mode]
    _assert_matches_oracle(_arrays(res, keys), ref, keys)


@pytest.mark.parametrize("n,p,ncov,miss,seed", [(400, 300, 3, 0.01, 0), (257, 120, 0, 0.05, 1),
                                                (600, 200, 6, 0.0, 2)])
def test_assoc_random_matches_oracle(n, p, ncov, miss, seed):
    from jamma_amd.lmm import run_lmm_association_jax

    G = balding_nichols(n, p + 200, miss=miss, seed=seed)
    y, W = phenotype(G, n_cov=ncov, seed=seed)
    y[::37] = np.nan  # some missing phenotypes
    K = ok.centered_kinship(G[:, p:])
    Gt = G[:, :p]
    res = run_lmm_association_jax(Gt, y, K, _snp_info(p), covariates=W, lmm_mode=4,
                                  precision="f64")
    ref = ol.run_lmm(Gt, y, kinship=K, covariates=W, lmm_mode=4)
    assert [r.rs for r in res] == [f"rs{j:04d}" for j in ref["snp_indices"]]
    _assert_matches_oracle(_arrays(res, ALL), ref, ALL)


@pytest.mark.parametrize("n,p,ncov,seed", [(700, 203, 3, 3), (350, 97, 0, 4)])
def test_assoc_lane_groups_match_wave_kernel(monkeypatch, n, p, ncov, seed):
    """The single-problem search with 16 lanes per SNP against the wavefront-per-SNP
    kernel (JM_LMM_WAVE_PER_SNP=1) on the same inputs, and both against the oracle.
    SNP counts that are not multiples of four leave idle lane groups in the last
    wavefront, which must write nothing."""
    from jamma_amd.lmm import run_lmm_association_jax

    G = balding_nichols(n, p + 200, miss=0.01, seed=seed)
    y, W = phenotype(G, n_cov=ncov, seed=seed)
    K = ok.centered_kinship(G[:, p:])
    Gt = G[:, :p]

    def run():
        res = run_lmm_association_jax(Gt, y, K, _snp_info(p), covariates=W, lmm_mode=4,
                                      precision="f64")
        assert len(res) % 4 != 0
        return _arrays(res, ALL)

    monkeypatch.delenv("JM_LMM_WAVE_PER_SNP", raising=False)
    got16 = run()
    monkeypatch.setenv("JM_LMM_WAVE_PER_SNP", "1")
    got64 = run()
    ref = ol.run_lmm(Gt, y, kinship=K, covariates=W, lmm_mode=4)
    _assert_matches_oracle(got16, ref, ALL)
    _assert_matches_oracle(got64, ref, ALL)
    _assert_matches_oracle(got16, got64, ALL)


def test_mixed_precision_meets_bar():
    """BF16x3 rotation vs the FP64 oracle: |delta -log10 p| <= 1e-4, 100% significance and
    direction agreement (the mixed path's acceptance bar), at n = 1000."""
    from jamma_amd.lmm import run_lmm_association_jax

    G = balding_nichols(1000, 1400, miss=0.01, seed=7)
    y, W = phenotype(G[:, :400], n_causal=20, h2=0.6, n_cov=3, seed=7)
    K = ok.centered_kinship(G[:, 400:])
    Gt = G[:, :400]
    res = run_lmm_association_jax(Gt, y, K, _snp_info(400), covariates=W, lmm_mode=4,
                                  precision="mixed")
    got = _arrays(res, ALL)
    ref = ol.run_lmm(Gt, y, kinship=K, covariates=W, lmm_mode=4)
    for k in ("p_wald", "p_lrt", "p_score"):
        dl = np.abs(np.log10(got[k]) - np.log10(ref[k]))
        assert dl.max() <= 1e-4, (k, dl.max())
        for thr in (0.05, 0.01, 1e-3, 5e-8):
            assert np.array_equal(got[k] < thr, ref[k] < thr), (k, thr)
    assert np.array_equal(np.sign(got["beta"]), np.sign(ref["beta"]))


def test_streaming_runner_matches_gemma(tmp_path):
    from jamma_amd.lmm import run_lmm_association_streaming

    G, y, K, C = _fixture_data()
    res = run_lmm_association_streaming(synthetic_prefix(), y, kinship=K, covariates=C,
                                        lmm_mode=4, chunk_size=128, precision="f64")
    gem = load_assoc(GOLDEN / "gemma_all_tests" / "gemma_all_covar.assoc.txt")
    got = _arrays(res, ALL)
    assert [r.rs for r in res] == list(gem["rs"])
    np.testing.assert_allclose(got["p_wald"], gem["p_wald"], rtol=1e-4)
    np.testing.assert_allclose(got["beta"], gem["beta"], rtol=1e-2)
    # the text file path
    out = tmp_path / "a.assoc.txt"
    assert run_lmm_association_streaming(synthetic_prefix(), y, kinship=K, covariates=C,
                                         lmm_mode=1, output_path=out, precision="f64") == []
    lines = out.read_text().splitlines()
    assert lines[0].split("\t")[-1] == "p_wald" and len(lines) == 501


def test_gwas_end_to_end(tmp_path):
    import jamma_amd

    r = jamma_amd.gwas(synthetic_prefix(), lmm_mode=1, output_dir=tmp_path, output_prefix="t",
                       show_progress=False, precision="f64")
    assert r.n_samples == 100
    got = load_assoc(tmp_path / "t.assoc.txt")
    gem = load_assoc(GOLDEN / "gemma_synthetic" / "gemma_assoc.assoc.txt")
    np.testing.assert_allclose(got["p_wald"], gem["p_wald"], rtol=1e-4)
    np.testing.assert_allclose(got["se"], gem["se"], rtol=1e-5)


def test_loco_runner_matches_oracle(tmp_path):
    from jamma_amd.lmm.loco import run_lmm_loco, run_lmm_loco_multi

    G = balding_nichols(240, 600, miss=0.01, seed=11)
    chrom = np.array([str(1 + (j * 3) // 600) for j in range(600)])
    y, W = phenotype(G, n_cov=2, seed=11)
    y[5] = np.nan
    pre = tmp_path / "loco"
    opl.write_plink(pre, G, phenotypes=y, chromosomes=chrom)
    res = run_lmm_loco(pre, y, covariates=W, lmm_mode=4, precision="f64")
    ref = ol.run_lmm_loco(G, chrom, y, covariates=W, lmm_mode=4)
    order = np.concatenate([ref[c]["global_indices"] for c in sorted(ref)])
    assert [r.rs for r in res] == [f"rs{j:06d}" for j in order]
    cat = {k: np.concatenate([ref[c][k] for c in sorted(ref)]) for k in ALL}
    _assert_matches_oracle(_arrays(res, ALL), cat, ALL)
    # two phenotypes sharing the LOCO eigendecompositions
    y2 = y + 0.5 * np.nan_to_num(G[:, 3])
    multi = run_lmm_loco_multi(pre, np.column_stack([y, y2]), covariates=W, lmm_mode=1,
                               precision="f64")
    assert np.allclose(_arrays(multi[0], ("p_wald",))["p_wald"], cat["p_wald"], rtol=1e-6)
    ref2 = ol.run_lmm_loco(G, chrom, y2, covariates=W, lmm_mode=1)
    p2 = np.concatenate([ref2[c]["p_wald"] for c in sorted(ref2)])
    np.testing.assert_allclose(-np.log10(_arrays(multi[1], ("p_wald",))["p_wald"]), -np.log10(p2),
                               atol=1e-6)


def test_streaming_runner_hwe_filter(tmp_path):
    """-hwe in the streaming runner (runner_streaming.py:265-305): counts over the
    valid samples, SNPs with chi-square HWE p < threshold removed, the others'
    results unchanged. Expected set from the closed form with SciPy's chi2."""
    from scipy import stats as sst

    from jamma_amd.lmm import run_lmm_association_streaming

    rng = np.random.default_rng(21)
    n, p = 300, 120
    G = rng.binomial(2, rng.uniform(0.1, 0.5, p), size=(n, p)).astype(float)
    # HWE violations: heterozygote deficit / excess
    G[:, 10] = np.where(rng.random(n) < 0.5, 0.0, 2.0)
    G[:, 20] = np.where(rng.random(n) < 0.85, 1.0, 0.0)
    G[rng.random((n, p)) < 0.01] = np.nan
    y = rng.normal(size=n) + 0.2 * np.nan_to_num(G[:, 3])
    y[:7] = np.nan
    pre = tmp_path / "hwe"
    opl.write_plink(pre, G, phenotypes=y)
    K = ok.centered_kinship(G)
    base = run_lmm_association_streaming(pre, y, kinship=K, lmm_mode=1, precision="f64")
    thr = 1e-3
    filt = run_lmm_association_streaming(pre, y, kinship=K, lmm_mode=1, precision="f64",
                                         hwe_threshold=thr)
    Gv = G[~np.isnan(y)]
    aa, ab, bb = ((Gv == v).sum(0).astype(float) for v in (0, 1, 2))
    tot = aa + ab + bb
    with np.errstate(invalid="ignore", divide="ignore"):
        f = (2 * aa + ab) / (2 * tot)
        e = [tot * f ** 2, 2 * tot * f * (1 - f), tot * (1 - f) ** 2]
        chi = sum((o - x) ** 2 / x for o, x in zip((aa, ab, bb), e))
    hwe_p = sst.chi2.sf(np.nan_to_num(chi, nan=0.0), 1)
    rs_base = [r.rs for r in base]
    keep = {f"rs{j:06d}" for j in range(p) if hwe_p[j] >= thr}
    assert [r.rs for r in filt] == [r for r in rs_base if r in keep]
    assert "rs000010" not in [r.rs for r in filt] and "rs000020" not in [r.rs for r in filt]
    assert "rs000010" in rs_base and "rs000020" in rs_base
    by_rs = {r.rs: r for r in base}
    for r in filt:
        assert r.p_wald == by_rs[r.rs].p_wald and r.beta == by_rs[r.rs].beta


def test_runner_file_output_equals_per_result_lines(tmp_path):
    """The streaming runner's file path (native tile formatter on the writer thread)
    writes exactly the lines the per-result formatter gives for the results the
    in-memory path returns, for every mode (lmm/io.py:12-262)."""
    from jamma_amd.lmm import run_lmm_association_streaming
    from jamma_amd.lmm.io import _HEADERS, format_line

    G, y, K, C = _fixture_data()
    for mode, kind in ((1, "wald"), (2, "lrt"), (3, "score"), (4, "all")):
        res = run_lmm_association_streaming(synthetic_prefix(), y, kinship=K, covariates=C,
                                            lmm_mode=mode, chunk_size=96, precision="f64")
        out = tmp_path / f"m{mode}.assoc.txt"
        assert run_lmm_association_streaming(synthetic_prefix(), y, kinship=K, covariates=C,
                                             lmm_mode=mode, chunk_size=96, output_path=out,
                                             precision="f64") == []
        want = _HEADERS[kind] + "\n" + "".join(format_line(r, kind) + "\n" for r in res)
        assert out.read_text() == want, mode


def test_gwas_write_eigen_then_reuse(tmp_path):
    """gwas(write_eigen=True) writes the valid-sample eigendecomposition in GEMMA format
    (pipeline.py:549-566); a second gwas() reading those files (-d / -u) gives the same
    association to the files' 10 significant digits."""
    import jamma_amd
    from jamma_amd.lmm.eigen_io import read_eigen_files

    pre = synthetic_prefix()
    r1 = jamma_amd.gwas(pre, lmm_mode=4, output_dir=tmp_path / "a", output_prefix="t",
                        write_eigen=True, save_kinship=True, show_progress=False,
                        precision="f64")
    dp, up = tmp_path / "a" / "t.eigenD.txt", tmp_path / "a" / "t.eigenU.txt"
    d, U = read_eigen_files(dp, up, n_samples=r1.n_samples)
    K = np.loadtxt(tmp_path / "a" / "t.cXX.txt")
    gem = np.loadtxt(GOLDEN / "gemma_synthetic" / "gemma_kinship.cXX.txt")
    assert np.max(np.abs(K - gem)) < 1e-9
    y = fam_phenotypes(pre)
    valid = ~np.isnan(y)
    Kv = gem[np.ix_(valid, valid)]
    assert np.max(np.abs((U * d) @ U.T - Kv)) < 1e-8
    r2 = jamma_amd.gwas(pre, lmm_mode=4, output_dir=tmp_path / "b", output_prefix="t",
                        eigenvalue_file=dp, eigenvector_file=up, show_progress=False,
                        precision="f64")
    a = load_assoc(tmp_path / "a" / "t.assoc.txt")
    b = load_assoc(tmp_path / "b" / "t.assoc.txt")
    assert list(a["rs"]) == list(b["rs"]) and r1.n_samples == r2.n_samples
    for k in ("p_wald", "p_lrt", "p_score"):
        np.testing.assert_allclose(-np.log10(b[k]), -np.log10(a[k]), atol=2e-6, err_msg=k)


def test_gwas_loco_saves_kinships(tmp_path):
    """gwas(loco=True, save_kinship=True) forwards the kinship output to the LOCO runner
    (pipeline.py:491-506): one .loco.cXX.chr<c>.txt per chromosome, equal to the
    oracle's LOCO kinships at 10 significant digits."""
    import jamma_amd

    G = balding_nichols(150, 300, miss=0.01, seed=31)
    chrom = np.array([str(1 + (j * 3) // 300) for j in range(300)])
    y, _ = phenotype(G, n_cov=0, seed=31)
    pre = tmp_path / "lc"
    opl.write_plink(pre, G, phenotypes=y, chromosomes=chrom)
    jamma_amd.gwas(pre, lmm_mode=1, loco=True, save_kinship=True, output_dir=tmp_path / "o",
                   output_prefix="r", show_progress=False, precision="f64")
    ref = ok.loco_kinship(G, chrom, 0.01, 0.05)
    for c in ("1", "2", "3"):
        K = np.loadtxt(tmp_path / "o" / f"r.loco.cXX.chr{c}.txt")
        np.testing.assert_allclose(K, ref[c], rtol=1e-9, atol=1e-12)
    assert (tmp_path / "o" / "r.assoc.txt").exists()


def test_mixed_guard_band_recomputes_in_fp64(monkeypatch):
    """Mixed-precision p-values within the guard band of a significance threshold are
    recomputed on the FP64 path: with the band widened to a decade, every SNP with a
    p-value in [5e-9, 0.5] (any test) gets the FP64 path's row (to rounding), the
    others keep their mixed results (within the 1e-4 bar of the FP64 ones)."""
    from jamma_amd.lmm import engine, run_lmm_association_jax

    G = balding_nichols(400, 500, miss=0.01, seed=41)
    y, W = phenotype(G[:, :300], n_cov=2, seed=41, n_causal=30, h2=0.6)
    K = ok.centered_kinship(G[:, 300:])
    info = _snp_info(300)
    f64 = _arrays(run_lmm_association_jax(G[:, :300], y, K, info, covariates=W, lmm_mode=4,
                                          precision="f64"), ALL)
    monkeypatch.setattr(engine, "GUARD_BAND", 1.0)
    mixed = _arrays(run_lmm_association_jax(G[:, :300], y, K, info, covariates=W, lmm_mode=4,
                                            precision="mixed"), ALL)
    P = np.stack([f64[k] for k in ("p_wald", "p_lrt", "p_score")], 1)
    with np.errstate(divide="ignore"):
        lp = np.log10(P)
    dist = np.nanmin(np.abs(lp[:, :, None] - np.log10([0.05, 0.01, 1e-3, 5e-8])), axis=(1, 2))
    # SNPs within 1e-3 of the band's edge could be selected or not (the mixed p-value
    # decides): they are held to neither statement
    near, far = dist <= 1.0 - 1e-3, dist > 1.0 + 1e-3
    assert near.sum() > 50 and far.sum() > 0
    # the recomputed rows take the FP64 rotation of a few columns (another summation
    # order than the tile kernel's): the FP64 path's rows to rounding
    _assert_matches_oracle({k: mixed[k][near] for k in ALL}, {k: f64[k][near] for k in ALL}, ALL)
    for k in ("p_wald", "p_lrt", "p_score"):
        d = np.abs(np.log10(mixed[k][far]) - np.log10(f64[k][far]))
        assert np.nanmax(d) <= 1e-4, k


@pytest.mark.parametrize("n,ncols,gather", [(333, 1, False), (1031, 7, True), (5000, 16, True),
                                            (4099, 16, False)])
def test_rotation_f64_few_columns(n, ncols, gather):
    """jm_rotate_f64 with at most 16 SNP columns (the GEMV form: U^T streamed once,
    columns staged in LDS) against NumPy FP64, with missing calls mean-imputed, a
    row gather, and sizes that are not multiples of the 4-byte staging or the
    4096-index stages."""
    from jamma_amd import _lib
    from jamma_amd.device import ptr, round_up, stream_ptr

    rng = np.random.default_rng(n + ncols)
    UT = rng.normal(size=(n, n))
    p = 40
    ld = round_up(n, 64)
    g = np.zeros((p, ld), dtype=np.int8)
    g[:, :n] = rng.integers(0, 3, size=(p, n))
    g[:, :n][rng.random((p, n)) < 0.03] = -1
    mean = rng.uniform(0.1, 1.9, p)
    rows = rng.choice(p, ncols, replace=False).astype(np.int32) if gather else \
        np.arange(ncols, dtype=np.int32)
    X = np.where(g[rows, :n] < 0, mean[rows][:, None], g[rows, :n]).T  # (n, ncols)
    ref = UT @ X
    dev = _dev()
    Z = torch.full((n, ncols), np.nan, dtype=torch.float64, device=dev)
    UTd = torch.as_tensor(UT, device=dev)
    gd = torch.as_tensor(g, device=dev)
    md = torch.as_tensor(mean, device=dev)
    cd = torch.as_tensor(rows, device=dev) if gather else None
    _lib.call("jm_rotate_f64", ptr(UTd), n, ptr(gd), ld, ptr(md), ptr(cd), n, ncols, ptr(Z),
              ncols, stream_ptr())
    got = Z.cpu().numpy()
    scale = np.abs(UT) @ np.abs(X)
    assert np.all(np.abs(got - ref) <= 1e-13 * scale)


def _dosages(n, p, seed, miss=0.02):
    rng = np.random.default_rng(seed)
    af = rng.uniform(0.05, 0.5, p)
    G = rng.binomial(2, af, size=(n, p)).astype(np.float64)
    G += rng.normal(0, 0.15, size=(n, p))  # imputation-like dosages in [0, 2]
    G = np.clip(G, 0.0, 2.0)
    G[rng.random((n, p)) < miss] = np.nan
    G[:, 3] = 1.25  # monomorphic: filtered
    return G


@pytest.mark.parametrize("n,p,batch", [(211, 900, 256), (640, 1500, 10000)])
def test_dosage_kinship_matches_oracle(n, p, batch):
    """compute_centered_kinship on non-integer genotypes (kinship/compute.py:93-226):
    device centring (jm_center_columns_f64) + FP64 MFMA accumulation (jm_dgemm) vs the
    oracle's impute_and_center + Xc Xc^T / p, and the reference's filters."""
    from jamma_amd.kinship import compute_centered_kinship

    G = _dosages(n, p, n)
    K = compute_centered_kinship(G, batch_size=batch)
    ref = ok.centered_kinship(G)
    assert np.max(np.abs(K - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
    K2 = compute_centered_kinship(G, batch_size=batch, maf_threshold=0.1, miss_threshold=0.03)
    ref2 = ok.centered_kinship(G, maf_threshold=0.1, miss_threshold=0.03)
    assert np.max(np.abs(K2 - ref2)) < 1e-12 * max(1.0, np.max(np.abs(ref2)))


def test_dosage_association_matches_oracle():
    """run_lmm_association_jax on dosages: device mean imputation + FP64 rotation, then
    the same moments and search as hard calls, against the oracle run on the same
    float64 matrix (statistics in float64, the input dtype)."""
    from jamma_amd.lmm import run_lmm_association_jax

    G = _dosages(300, 400, 5)
    y, W = phenotype(np.nan_to_num(G[:, :150]), n_cov=2, seed=5)
    K = ok.centered_kinship(G[:, 150:])
    for precision in ("f64", "mixed"):  # dosages always take the FP64 rotation
        res = run_lmm_association_jax(G[:, :150], y, K, _snp_info(150), covariates=W,
                                      lmm_mode=4, precision=precision)
        ref = ol.run_lmm(G[:, :150], y, kinship=K, covariates=W, lmm_mode=4,
                         stat_dtype=np.float64)
        assert [r.rs for r in res] == [f"rs{j:04d}" for j in ref["snp_indices"]]
        assert "rs0003" not in [r.rs for r in res]
        _assert_matches_oracle(_arrays(res, ALL), ref, ALL)


def test_resident_bed_decodes_from_one_copy(tmp_path):
    """Inside resident_bed, every pass decodes from one HBM copy of the .bed payload and
    the association's pass-1 counts are the kinship pass's (all samples valid): kinship
    and association results are identical to the windowed-copy path's."""
    from jamma_amd.io.plink import BedFile, resident_bed
    from jamma_amd.kinship.compute import kinship_streaming_device
    from jamma_amd.lmm import run_lmm_association_streaming

    G = balding_nichols(160, 700, miss=0.02, seed=61)
    y, W = phenotype(G, n_cov=1, seed=61)
    pre = tmp_path / "res"
    opl.write_plink(pre, G, phenotypes=y)
    K0 = kinship_streaming_device(pre, chunk_size=256).cpu().numpy()
    r0 = run_lmm_association_streaming(pre, y, kinship=K0, covariates=W, lmm_mode=4,
                                       chunk_size=256, precision="f64")
    with resident_bed(pre):
        assert BedFile(pre).resident
        K1 = kinship_streaming_device(pre, chunk_size=256).cpu().numpy()
        assert BedFile(pre).cached_counts(None) is not None
        r1 = run_lmm_association_streaming(pre, y, kinship=K1, covariates=W, lmm_mode=4,
                                           chunk_size=256, precision="f64")
    assert not BedFile(pre).resident
    assert np.array_equal(K0, K1)
    assert [r.rs for r in r0] == [r.rs for r in r1]
    for k in ALL:
        assert np.array_equal(_arrays(r0, (k,))[k], _arrays(r1, (k,))[k], equal_nan=True), k

"""jm_spectrum_moments_noalloc: the allocation-free form of jm_spectrum_moments.

Called through the C ABI on two inputs in a row (the second with fewer SNP columns
and fewer partner rows, so stale partials of the first call sit in the shared scratch
array), each checked against jm_spectrum_moments with a caller-provided workspace:
the results must be bitwise equal (same kernels, same summation order). The
rejection of calls that need more partials than the scratch holds is a host check,
tested on the CPU (test_library_cpu.py).
"""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _moments_pair(co, Z, z_f32, p, V, nv):
    from jamma_amd import _lib
    from jamma_amd.device import ptr, stream_ptr

    T = co.n_terms
    ref = torch.full((p, nv + 1, T), np.nan, dtype=torch.float64, device="cuda")
    got = torch.full((p, nv + 1, T), np.nan, dtype=torch.float64, device="cuda")
    plan = ctypes.byref(co.plan_c)
    work = co.moments_work(p, nv)
    _lib.call("jm_spectrum_moments", ptr(Z), int(z_f32), p, p, ptr(V), V.shape[1], nv, plan,
              ptr(work), ptr(ref), T, stream_ptr())
    _lib.call("jm_spectrum_moments_noalloc", ptr(Z), int(z_f32), p, p, ptr(V), V.shape[1], nv,
              plan, ptr(got), T, stream_ptr())
    torch.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_moments_match_workspace_form():
    from jamma_amd import _lib
    from jamma_amd.lmm.engine import Cohort

    rng = np.random.default_rng(17)
    n = 700
    d = np.sort(rng.uniform(0.02, 4.0, n))
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    co = Cohort(d, torch.as_tensor(Q, device="cuda"), precision="f64")
    cap = int(_lib.lib.jm_moments_scratch_bytes())
    assert cap == 64 << 20
    for p, nv, z_f32 in ((1500, 5, True), (333, 2, False)):
        Zh = rng.normal(size=(n, p))
        Z = torch.as_tensor(Zh.astype(np.float32) if z_f32 else Zh, device="cuda")
        V = torch.as_tensor(rng.normal(size=(nv, n)), device="cuda")
        need = int(_lib.lib.jm_moments_work_bytes(ctypes.byref(co.plan_c), p, nv))
        assert 0 < need <= cap
        print(f"p = {p}, nv = {nv}, f32 = {z_f32}: {need} bytes of partials", flush=True)
        ref, got = _moments_pair(co, Z, z_f32, p, V, nv)
        assert np.all(np.isfinite(ref))
        assert np.array_equal(ref, got)

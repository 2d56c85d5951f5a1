"""letkf_core_batched_noalloc through the C ABI against letkf_core_batched
on the same problems: two calls in a row on the shared scratch, the second
with fewer problems (the scratch keeps the first call's contents), for
REAL(8) and REAL(4) arrays.  Both entry points run the same kernels, so the
outputs must agree bit for bit."""
import ctypes as C

import numpy as np
import pytest
import torch

from sbpom_letkf_v2_amd import _lib, engine, synth

pytestmark = pytest.mark.gpu


def _problems(rng, K, sizes, dtype):
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    L = int(offs[-1])
    Y = rng.standard_normal((max(L, 1), K))
    for b in range(len(sizes)):
        seg = slice(offs[b], offs[b + 1])
        Y[seg] -= Y[seg].mean(1, keepdims=True)
    rd = rng.uniform(0.3, 2.0, max(L, 1)) ** 2
    rl = rng.uniform(0.01, 1.0, max(L, 1))
    dp = rng.standard_normal(max(L, 1))
    return offs, [np.asarray(a, dtype) for a in (Y, rd, rl, dp)]


@pytest.mark.parametrize("K,real_kind", [(64, 8), (100, 4)])
def test_core_batched_noalloc_matches_core_batched(K, real_kind):
    L = _lib.lib()
    dt = np.float64 if real_kind == 8 else np.float32
    tdt = torch.float64 if real_kind == 8 else torch.float32
    pr = synth.make_problem(nlon=8, nlat=6, nlev=2, nmem=K, nobs=10, seed=1, real_kind=real_kind)
    prm = engine.make_params(pr)
    prm.cov_infl_mul = 1.1
    eng = engine.Engine(prm)
    vp = lambda t: C.c_void_p(t.data_ptr())
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a)).cuda()
    rng = np.random.default_rng(K)
    try:
        for call, sizes in enumerate(([0, 1, 3, K - 1, K, 2 * K + 5, 17, 0, 250, 40], [5, 0, K + 3])):
            print(f"case noalloc K={K} real_kind={real_kind} call {call} nb={len(sizes)}", flush=True)
            offs, arrs = _problems(rng, K, sizes, dt)
            tO = dev(offs)
            tY, trd, trl, tdp = (dev(a) for a in arrs)
            nb = len(sizes)
            w0 = torch.empty((nb, K), dtype=tdt, device="cuda")
            W0 = torch.empty((nb, K, K), dtype=tdt, device="cuda")
            _lib.check(L.letkf_core_batched(eng.h, nb, vp(tO), vp(tY), vp(trd), vp(trl), vp(tdp), vp(w0),
                                            vp(W0)))
            w1 = torch.full((nb, K), -1.0, dtype=tdt, device="cuda")
            W1 = torch.full((nb, K, K), -1.0, dtype=tdt, device="cuda")
            st = torch.full((1,), 99, dtype=torch.int32, device="cuda")
            s = torch.cuda.current_stream().cuda_stream
            _lib.check(L.letkf_core_batched_noalloc(C.byref(prm), nb, vp(tO), vp(tY), vp(trd), vp(trl), vp(tdp),
                                                    vp(w1), vp(W1), vp(st), C.c_void_p(s)))
            torch.cuda.synchronize()
            assert int(st.item()) == 0
            assert torch.equal(w1, w0) and torch.equal(W1, W0)
            # empty problems: letkf_core_noobs
            for b, n in enumerate(sizes):
                if n == 0:
                    assert torch.allclose(W1[b].double(), np.sqrt(1.1) * torch.eye(K, dtype=torch.float64,
                                                                                  device="cuda"), atol=1e-6)
    finally:
        eng.close()


def test_core_batched_noalloc_rejects_beyond_capacity():
    L = _lib.lib()
    pr = synth.make_problem(nlon=8, nlat=6, nlev=2, nmem=64, nobs=10, seed=1)
    prm = engine.make_params(pr)
    st = torch.zeros(1, dtype=torch.int32, device="cuda")
    rc = L.letkf_core_batched_noalloc(C.byref(prm), 1009, C.c_void_p(1), None, None, None, None, C.c_void_p(1),
                                      C.c_void_p(1), C.c_void_p(st.data_ptr()), None)
    assert rc == -4 and b"at most 1008" in L.letkf_last_error()

"""The analysis split into a geometry plan and a capturable enqueue
(letkf_da_main_plan / letkf_da_main_enqueue / letkf_da_main_status), and the
enqueue captured in a graph and replayed.

Bar: plan + enqueue reproduces letkf_da_main_device bit for bit (the same
kernels in the same order); every graph replay, with new forecast and
innovation values copied into the captured buffers and eager work between
replays, reproduces an eager analysis of the same values bit for bit.
"""
import numpy as np
import pytest
import torch

from sbpom_letkf_v2_amd import _lib, engine, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _problem(real_kind=8, K=24, seed=3):
    return synth.make_problem(nlon=40, nlat=32, nlev=6, nmem=K, nobs=2500, seed=seed, alpha_rtpp=0.7,
                              land_frac=0.1, real_kind=real_kind).to("cuda")


def test_enqueue_needs_a_plan():
    _say("case enqueue without a plan")
    pr = _problem()
    eng = engine.Engine(engine.make_params(pr, "fp64", "plain"))
    try:
        a3, a2 = torch.empty_like(pr.fcst3d), torch.empty_like(pr.fcst2d)
        with pytest.raises(_lib.EngineError):
            eng.enqueue_device(pr, a3, a2)
    finally:
        eng.close()


def test_enqueue_reads_current_parameters():
    """A new inflation factor between enqueues (same plan) gives the
    analysis of a fresh handle with that factor."""
    _say("case enqueue after set_params")
    pr = _problem()
    prm = engine.make_params(pr, "fp64", "batched")
    eng = engine.Engine(prm)
    try:
        eng.plan_device(pr)
        prm2 = engine.make_params(pr, "fp64", "batched")
        prm2.cov_infl_mul = 1.3
        prm2.alpha_rtpp = 0.5
        eng.set_params(prm2)
        a3, a2 = torch.empty_like(pr.fcst3d), torch.empty_like(pr.fcst2d)
        eng.enqueue_device(pr, a3, a2)
        eng.status()
    finally:
        eng.close()
    fresh = engine.Engine(prm2)
    try:
        w3, w2 = fresh.da_main_device(pr)
        torch.cuda.synchronize()
    finally:
        fresh.close()
    assert torch.equal(a3, w3) and torch.equal(a2, w2)


@pytest.mark.parametrize("precision,form", [("fp64", "plain"), ("fp32", "batched")])
def test_plan_enqueue_matches_da_main(precision, form):
    _say(f"case plan+enqueue {precision} {form}")
    pr = _problem()
    eng = engine.Engine(engine.make_params(pr, precision, form))
    try:
        want3, want2 = eng.da_main_device(pr)
        torch.cuda.synchronize()
        eng.plan_device(pr)
        a3, a2 = torch.empty_like(pr.fcst3d), torch.empty_like(pr.fcst2d)
        s = torch.cuda.current_stream()
        eng.enqueue_device(pr, a3, a2, stream=s.cuda_stream)
        eng.status(stream=s.cuda_stream)
        assert torch.equal(a3, want3) and torch.equal(a2, want2)
        # the plan survives a precision change (same geometry), not a new geometry
        other = _problem(seed=4)
        with pytest.raises(_lib.EngineError):
            eng.enqueue_device(other, a3, a2, stream=s.cuda_stream)
    finally:
        eng.close()


@pytest.mark.parametrize("real_kind", [8, 4])
def test_graph_replays_match_eager(real_kind):
    _say(f"case graph capture real_kind={real_kind}")
    pr = _problem(real_kind)
    ref = _problem(real_kind)   # same geometry, separate buffers for the eager runs
    prm = engine.make_params(pr, "fp32", "batched")
    eng = engine.Engine(prm)
    eager = engine.Engine(prm)
    try:
        eng.plan_device(pr)
        a3, a2 = torch.empty_like(pr.fcst3d), torch.empty_like(pr.fcst2d)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, capture_error_mode="thread_local"):
            eng.enqueue_device(pr, a3, a2, stream=torch.cuda.current_stream().cuda_stream)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(11)
        for rep in range(3):
            # new values in the captured buffers
            f3 = ref.fcst3d + 0.05 * rep * torch.randn(ref.fcst3d.shape, generator=gen, device="cuda",
                                                       dtype=torch.float64).to(ref.fcst3d.dtype)
            dep = ref.obsdep * (1.0 + 0.1 * rep)
            pr.fcst3d.copy_(f3)
            pr.obsdep.copy_(dep)
            g.replay()
            torch.cuda.synchronize()
            eng.status()
            ref.fcst3d.copy_(f3)
            ref.obsdep.copy_(dep)
            w3, w2 = eager.da_main_device(ref)
            torch.cuda.synchronize()
            d3 = (a3 - w3).abs().max().item()
            _say(f"  replay {rep}: max |graph - eager| = {d3:.3e}")
            assert torch.equal(a3, w3) and torch.equal(a2, w2)
    finally:
        eng.close()
        eager.close()


def test_c4_eight_cycles_in_one_graph():
    """Workload c4 at 1/16 of its horizontal size: the eight independent
    cycles (nmem 32/64/100/128, 0.5 .. 4 obs per column, REAL(4) I/O) each
    planned on its own handle (workspace cap 1 GB), their enqueues captured
    in ONE graph on one stream; three replays, with new forecast and
    innovation values and eager analyses between them, reproduce each
    cycle's eager analysis bit for bit."""
    cycles = synth.c4_cycles(scale=1.0 / 16.0, nlev=6)
    probs, engs, outs = [], [], []
    try:
        for c, kw in enumerate(cycles):
            pr = synth.make_problem(device="cuda", real_kind=4, sigma_h_cells=1.5, sigma_v=1.0, **kw)
            prm = engine.make_params(pr, "fp32" if c % 2 == 0 else "bf16x3", "batched")
            prm.workspace_bytes = 1 << 30
            eng = engine.Engine(prm)
            eng.plan_device(pr)
            probs.append(pr)
            engs.append(eng)
            outs.append((torch.empty_like(pr.fcst3d), torch.empty_like(pr.fcst2d)))
        _say(f"case c4 in one graph: {len(cycles)} cycles planned")
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, capture_error_mode="thread_local"):
            s = torch.cuda.current_stream().cuda_stream
            for eng, pr, (a3, a2) in zip(engs, probs, outs):
                eng.enqueue_device(pr, a3, a2, stream=s)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(21)
        for rep in range(3):
            for pr in probs:
                pr.fcst3d.add_((0.02 * torch.randn(pr.fcst3d.shape, generator=gen, device="cuda")).to(pr.fcst3d.dtype))
                pr.obsdep.mul_(1.05)
            g.replay()
            torch.cuda.synchronize()
            for c, (eng, pr, (a3, a2)) in enumerate(zip(engs, probs, outs)):
                eng.status()
                fresh = engine.Engine(engine.make_params(pr, "fp32" if c % 2 == 0 else "bf16x3", "batched"))
                try:
                    w3, w2 = fresh.da_main_device(pr)
                    torch.cuda.synchronize()
                finally:
                    fresh.close()
                same = torch.equal(a3, w3) and torch.equal(a2, w2)
                if not same:
                    _say(f"  replay {rep} cycle {c}: max |graph - eager| {(a3 - w3).abs().max().item():.3e}")
                assert same
            _say(f"  replay {rep}: all {len(cycles)} cycles bit-exact")
    finally:
        for eng in engs:
            eng.close()

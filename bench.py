"""Benchmark of the LETKF local analysis (da_main) on one MI355X.

Workload (default): 512 x 512 x 40 grid, K = 64 members, U/V/T/S + SSH,
2e5 observations (1.5e5 SST/SSH on 20-cell satellite-like stripes + 1250
full-depth T/S profiles), RTPP 0.9, synthetic Gaussian-random-field
ensemble generated on the GPU.  A step is one full analysis of every grid
point, inputs resident in HBM.

Forms:
  plain       FP64 throughout, K x K eigen-decomposition at every point with
              local obs
  batched     points with nobsl < K solved in the nobsl x nobsl form, the
              matrix functions from rational approximations and shifted
              tridiagonal solves; with --precision fp32 / bf16x3 the Gram
              products run on MFMA (the rest FP64), bf16x3t also runs the
              Householder reduction in FP32
  batched_ql  the batched form with the QL eigen-decomposition
Default: batched + bf16x3t (BF16 hi/lo splits with FP32 accumulation for the
Gram, FP32 Householder reduction), timed next to plain FP64; bf16x3 (FP64
reduction) and the batched_ql form are timed beside it.

Prints one JSON line (rank 0).  Progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

METRIC = "analysed grid pts/s at K=64: plain FP64 vs batched mixed-prec; max err vs FP64"
PEAK_FP64_TFLOPS = 78.6
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "fp32p": 157.3, "fp32x3": 157.3, "fp32d": 157.3, "bf16x3": 2500.0,
               "bf16x3t": 2500.0}   # dense, per operand type of the Gram
PEAK_HBM_GBS = 8000.0
EIG_KINDS = ["k_eigr_tridiag", "k_eigr_ql", "k_eigr_qvec", "k_eigr_vrot", "k_eigr_vout"]
# form "batched": kind 3 holds the rational functions' shifted solves (no QL, no replay)
EIG_KINDS_RAT = ["k_eigr_tridiag", "k_eigr_ql", "k_eigr_qvec", "k_eigr_rat", "k_eigr_vout"]


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--precision", default="bf16x3t",
                    choices=["fp64", "fp32", "fp32p", "fp32x3", "fp32d", "bf16x3", "bf16x3t"],
                    help="Gram arithmetic of the timed form (default: BF16 hi/lo splits on MFMA with FP32 "
                         "accumulation, as the K=64 configuration names, and the Householder reduction in FP32)")
    ap.add_argument("--form", default="batched", choices=["plain", "batched", "batched_ql"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target CPU time of the oracle baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the fixed vector-path rule instead of letkf_autotune on the chosen form")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the plan-once + graph-replay timing of the chosen form")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the BF16 hi/lo Gram variant, the c1 record (K=100 with the tracer, chosen "
                         "vs plain FP64) and the port-vs-reference CPU calibration of the default run")
    ap.add_argument("--c1-steps", type=int, default=2,
                    help="timed steps of each form in the c1 record of the default run")
    ap.add_argument("--adjoint", action="store_true",
                    help="differentiable workload (default c3): time forward + hand-written adjoint and "
                         "check it by central finite differences")
    ap.add_argument("--fd-coords", type=int, default=200)
    ap.add_argument("--xtoy", action="store_true",
                    help="time the observation operator H(x) of every member (default workload c1) on the "
                         "analysis arrays, with an oracle check and CPU baseline on a few members")
    ap.add_argument("--setobs", action="store_true",
                    help="time set_obs after H(x) (default workload c1): statistics, QC, compaction, cell "
                         "sort, AOEI over every obs, with a bit-exact oracle check and CPU baseline")
    ap.add_argument("--output", action="store_true",
                    help="time the device half of the output step (default workload c1): grid mean/spread "
                         "of analysis and forecast and the REAL(4) file images of every field, with a "
                         "bit-exact oracle check and CPU baseline on two levels")
    ap.add_argument("--lpfgm", action="store_true",
                    help="time the LPFGM core (letkf_lpfgm_core_batched) on local problems of nmem 100 and 84 "
                         "obs, with an oracle check and CPU baseline on a sample")
    ap.add_argument("--lpfgm-problems", type=int, default=50000)
    ap.add_argument("--lpfgm-da", action="store_true",
                    help="time da_main with the LPFGM core (iswitch_da 3) on the workload, with an oracle "
                         "check and CPU baseline on sampled points")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only an elementwise kernel that reads and writes 1 GiB (three launches), the "
                         "reference for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes")
    return ap.parse_args()


def run_xtoy(args):
    """H(x) (Trans_XtoY, letkf/run/src/common.f90:391-500) of all members of a
    workload's ensemble, read in place from the analysis arrays, for the
    workload's obs in a shuffled (file) order.  One step =
    one call over every (obs, member).  Roofline: HBM, algorithmic bytes
    = per (obs, member) the 4 x 6 column values an Akima obs reads (4 for
    SSH and surface-level obs) plus the hx store; depth columns are shared
    by all members and not counted."""
    import numpy as np
    import torch
    import pyoracle as po
    from sbpom_letkf_v2_amd import engine, synth
    wl = args.workload if args.workload != "c2" else "c1"
    log(f"generating workload {wl} on the GPU")
    pr = synth.config(wl, device="cuda", real_kind=8)
    meta = pr.meta
    nlon, nlat, nlev, K = meta["nlon"], meta["nlat"], meta["nlev"], meta["nmem"]
    nv3, nv2 = meta["nv3d"], meta["nv2d"]
    lon = pr.lon1.reshape(nlat, nlon).contiguous()
    lat = pr.lat1.reshape(nlat, nlon).contiguous()
    depth = pr.depth1.reshape(nlev, nlat, nlon).contiguous()
    fsm = torch.ones_like(lon)  # all sea: every obs does its full interpolation
    dlon = float(lon[0, 1] - lon[0, 0])
    dlat = float(lat[1, 0] - lat[0, 0])
    idx = (torch.floor((pr.obslon - lon[0, 0]) / dlon) + 1).to(torch.int32).contiguous()
    idy = (torch.floor((pr.obslat - lat[0, 0]) / dlat) + 1).to(torch.int32).contiguous()
    # obs in input (file) order: a fixed shuffle of the workload's sorted obs
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    sh = torch.randperm(int(pr.obselm.numel()), generator=g, device="cuda")
    elm, lono, lato, levo = (a[sh].contiguous() for a in (pr.obselm, pr.obslon, pr.obslat, pr.obslev))
    idx, idy = idx[sh].contiguous(), idy[sh].contiguous()
    v2 = pr.fcst2d if nv2 > 0 else torch.zeros((1, K, nlat * nlon), dtype=torch.float64, device="cuda")
    nobs = int(elm.numel())
    call = lambda: engine.trans_xtoy_device(lon, lat, depth, fsm, elm, idx, idy, lono, lato, levo, pr.fcst3d, v2,
                                            analysis_layout=True)
    for _ in range(args.warmup):
        call()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        hx, qc = call()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    units = nobs * K
    undef = 999.0
    ssh = (elm == synth.ID_Z).sum().item()
    per_obs = torch.where(elm == synth.ID_Z, 5, 25).to(torch.float64)
    bytes_step = float((per_obs.sum() * 8 * K).item())
    gbs = bytes_step / (ms * 1e-3) / 1e9
    defined = (hx != undef).float().mean().item()
    log(f"  {ms:.3f} ms per call, {units / (ms * 1e-3):.3e} obs x members/s, {gbs:.0f} GB/s algorithmic, "
        f"{defined:.3f} defined")
    # oracle on a few members: bit-exact check at full size and the CPU baseline
    nchk = 4
    cpu = None
    if not args.no_cpu:
        host = lambda t: t.cpu().numpy()
        g = po.OracleXtoY(nlon, nlat, nlev, nv3, max(nv2, 1), synth.ID_U, synth.ID_V, synth.ID_T, synth.ID_S,
                          synth.ID_Z, 1, 2, 3, 4, 1, 0.0, undef)
        cargs = [host(a) for a in (elm, idx, idy, lono, lato, levo)]
        lonh, lath, deph, fsmh = host(lon), host(lat), host(depth), host(fsm)
        mism, t_cpu = 0, 0.0
        for m in range(nchk):
            mm = (m * (K - 1)) // max(nchk - 1, 1)
            v3m = np.ascontiguousarray(host(pr.fcst3d[:, mm]).reshape(nv3, nlev, nlat, nlon))
            v2m = np.ascontiguousarray(host(v2[:, mm]).reshape(-1, nlat, nlon))
            t0 = time.perf_counter()
            want = po.trans_xtoy(g, *cargs, lonh.T, lath.T, deph.T, fsmh.T, v3m.T, v2m.T)
            t_cpu += time.perf_counter() - t0
            mism += int((host(hx[mm]) != want).sum())
        cpu = {"value": nchk * nobs / t_cpu, "unit": "obs x members/s", "cores": 1, "kind": "port",
               "sample": f"{nchk} members x {nobs} obs, FP64 C oracle (one thread), {t_cpu:.1f} s"}
        log(f"  oracle on {nchk} members: {mism} mismatching values; cpu {cpu['value']:.3e} obs x members/s")
        if mism:
            raise SystemExit("H(x) differs from the oracle")
    out = {"metric": "H(x) observations x members per second", "value": units / (ms * 1e-3),
           "unit": "obs x members/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": f"{wl}: {nlon}x{nlat}x{nlev} K={K} nobs={nobs} ({ssh} SSH), H(x) of every member "
                                  "from the analysis arrays"},
           "roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": gbs / PEAK_HBM_GBS, "traffic": None},
           "cpu_baseline": cpu, "defined_fraction": defined}
    print(json.dumps(out))


def run_setobs(args):
    """set_obs after the observation operator (common_obs.f90:139-402) on a
    workload's obs: H(x) of every member from the analysis arrays
    (untimed input), then one step = letkf_set_obs_device over all obs.
    Roofline: HBM, algorithmic bytes = hx read once and hdxf written once
    (nobs x nmem values each) plus 13 per-obs values read (elm, idx, idy,
    qc, lon, lat, lev, dat, err; hx mean/spread/dep/err staged) and 12
    written per kept obs."""
    import numpy as np
    import torch
    import pyoracle as po
    from sbpom_letkf_v2_amd import engine, synth
    wl = args.workload if args.workload != "c2" else "c1"
    log(f"generating workload {wl} on the GPU")
    pr = synth.config(wl, device="cuda", real_kind=8)
    meta = pr.meta
    nlon, nlat, nlev, K = meta["nlon"], meta["nlat"], meta["nlev"], meta["nmem"]
    lon = pr.lon1.reshape(nlat, nlon).contiguous()
    lat = pr.lat1.reshape(nlat, nlon).contiguous()
    depth = pr.depth1.reshape(nlev, nlat, nlon).contiguous()
    fsm = torch.ones_like(lon)
    dlon = float(lon[0, 1] - lon[0, 0])
    dlat = float(lat[1, 0] - lat[0, 0])
    # obs in input (file) order: a fixed shuffle of the workload's obs
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    nobs = int(pr.obselm.numel())
    sh = torch.randperm(nobs, generator=g, device="cuda")
    elm, olon, olat, olev, err = (a[sh].contiguous() for a in (pr.obselm, pr.obslon, pr.obslat, pr.obslev,
                                                               pr.obserr))
    idx = (torch.floor((olon - lon[0, 0]) / dlon) + 1).to(torch.int32).contiguous()
    idy = (torch.floor((olat - lat[0, 0]) / dlat) + 1).to(torch.int32).contiguous()
    v2 = pr.fcst2d if meta["nv2d"] > 0 else torch.zeros((1, K, nlat * nlon), dtype=torch.float64, device="cuda")
    hx, qc = engine.trans_xtoy_device(lon, lat, depth, fsm, elm, idx, idy, olon, olat, olev, pr.fcst3d, v2,
                                      analysis_layout=True)
    dat = (hx.mean(0) + err * torch.randn(nobs, generator=g, device="cuda", dtype=torch.float64)).contiguous()
    eng = engine.Engine(engine.make_params(pr))
    call = lambda: eng.set_obs_device(nlon, nlat, elm, olon, olat, olev, dat, err, idx, idy, hx, qc=qc)
    for _ in range(args.warmup):
        call()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        res = call()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    nk = res["nobs"]
    bytes_step = 8.0 * (nobs * K + nk * K + 13 * nobs + 12 * nk)
    gbs = bytes_step / (ms * 1e-3) / 1e9
    log(f"  {ms:.3f} ms per call (host sync included), {nobs / (ms * 1e-3):.3e} obs/s, {gbs:.0f} GB/s "
        f"algorithmic; kept {nk} of {nobs}")
    cpu = None
    if not args.no_cpu:
        host = lambda t: t.cpu().numpy()
        c = po.setobs_config(K, nlon, nlat)
        t0 = time.perf_counter()
        want = po.set_obs(c, host(elm), host(olon), host(olat), host(olev), host(dat), host(err), host(idx),
                          host(idy), host(hx).T, qc=host(qc))
        t_cpu = time.perf_counter() - t0
        ok = want["nobs"] == nk and np.array_equal(host(res["obshdxf"]).T, want["hdxf"]) and all(
            np.array_equal(host(res["obs" + k]), want[k]) for k in ("elm", "lon", "err", "dep", "hxfsprd")) \
            and np.array_equal(host(res["nobsgrd"]).T, want["nobsgrd"])
        cpu = {"value": nobs / t_cpu, "unit": "obs/s", "cores": 1, "kind": "port",
               "sample": f"all {nobs} obs x {K} members, FP64 C oracle (one thread), {t_cpu:.2f} s"}
        log(f"  oracle: bit-exact {ok}; cpu {cpu['value']:.3e} obs/s")
        if not ok:
            raise SystemExit("set_obs differs from the oracle")
    eng.close()
    out = {"metric": "set_obs observations per second", "value": nobs / (ms * 1e-3), "unit": "obs/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{wl}: {nlon}x{nlat} K={K} nobs={nobs} (kept {nk}), set_obs after H(x) of "
                                  "every member"},
           "roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": gbs / PEAK_HBM_GBS, "traffic": None},
           "cpu_baseline": cpu}
    print(json.dumps(out))


def run_output(args):
    """The device half of the output step after the analysis (write_ens_mpi,
    common_io.f90:863-1074) on a workload (default c1): one step = grid mean
    and spread of the analysis and of the forecast (ensemble_mesp,
    common.f90:505-596) + the REAL(4) file image of every written field
    (nmem analysis members, the two means and the two spreads: gather to the
    global grid, quasi-global wrap, file vertical order).  Roofline: HBM,
    algorithmic bytes = each ensemble read once and the mean/spread written
    (2 x (nmem + 2) x 8 B per grid value) plus 8 B read and 4 B written per
    value of each packed field."""
    import numpy as np
    import torch
    import pyoracle as po
    from sbpom_letkf_v2_amd import engine, synth
    wl = args.workload if args.workload != "c2" else "c1"
    log(f"generating workload {wl} on the GPU")
    pr = synth.config(wl, device="cuda", real_kind=8)
    meta = pr.meta
    nlon, nlat, nlev, K = meta["nlon"], meta["nlat"], meta["nlev"], meta["nmem"]
    nv3d, nv2d, nij = meta["nv3d"], meta["nv2d"], nlon * nlat
    f3, f2 = pr.fcst3d, pr.fcst2d
    if nv2d == 0:
        f2 = torch.zeros((1, K, nij), dtype=torch.float64, device="cuda")
        nv2d = 1
    # an analysis ensemble of the same shape (values do not change the cost)
    a3 = (f3 * 0.75).contiguous()
    a2 = (f2 * 0.75).contiguous()
    g3 = torch.empty((K + 4, nv3d, nlev, nlat, nlon), dtype=torch.float32, device="cuda")
    g2 = torch.empty((K + 4, nv2d, nlat, nlon), dtype=torch.float32, device="cuda")
    nval = nij * (nlev * nv3d + nv2d)
    log(f"  {nlon}x{nlat}x{nlev} nmem {K}: ensembles {2 * f3.numel() * 8 / 1e9:.1f} GB, file images "
        f"{(g3.numel() + g2.numel()) * 4 / 1e9:.1f} GB")
    ev = lambda: torch.cuda.Event(enable_timing=True)

    def step(t):
        t[0].record()
        am3, as3, am2, as2 = engine.ensemble_mesp_device(a3, a2)
        fm3, fs3, fm2, fs2 = engine.ensemble_mesp_device(f3, f2)
        t[1].record()
        engine.grid_pack_device(nlon, nlat, a3, a2, wrap=True, file_order=True, g3=g3[:K], g2=g2[:K])
        for j, (x3, x2) in enumerate(((am3, am2), (as3, as2), (fm3, fm2), (fs3, fs2))):
            engine.grid_pack_device(nlon, nlat, x3, x2, wrap=True, file_order=True, g3=g3[K + j:K + j + 1],
                                    g2=g2[K + j:K + j + 1])
        t[2].record()
        return am3, as3, am2, as2

    for _ in range(args.warmup):
        step([ev(), ev(), ev()])
    torch.cuda.synchronize()
    ts = []
    for _ in range(args.steps):
        t = [ev(), ev(), ev()]
        res = step(t)
        ts.append(t)
    torch.cuda.synchronize()
    ms_mesp = sum(t[0].elapsed_time(t[1]) for t in ts) / args.steps
    ms_pack = sum(t[1].elapsed_time(t[2]) for t in ts) / args.steps
    ms = ms_mesp + ms_pack
    b_mesp = 2.0 * nval * (K + 2) * 8
    b_pack = (K + 4.0) * nval * 12
    gbs = lambda b, t: b / (t * 1e-3) / 1e9
    log(f"  mesp x2 {ms_mesp:.2f} ms ({gbs(b_mesp, ms_mesp):.0f} GB/s algorithmic), pack of {K + 4} fields "
        f"{ms_pack:.2f} ms ({gbs(b_pack, ms_pack):.0f} GB/s), {nij * nlev / (ms * 1e-3):.3e} grid points/s")
    cpu = None
    if not args.no_cpu:
        # oracle on two levels (every member, every variable): mesp of both
        # ensembles and the file images of every field, bit for bit against
        # the GPU's output levels nlev-2, nlev-1 (the file order flips them)
        L = 2
        host3 = lambda t: np.asfortranarray(t[:, :, :L].cpu().numpy().transpose(3, 2, 1, 0))  # (nij,L,K,nv3d)
        host2 = lambda t: np.asfortranarray(t.cpu().numpy().transpose(2, 1, 0))
        ha3, hf3, ha2, hf2 = host3(a3), host3(f3), host2(a2), host2(f2)
        t0 = time.perf_counter()
        wa = po.ensemble_mesp(ha3, ha2)
        wf = po.ensemble_mesp(hf3, hf2)
        p3, p2 = po.grid_pack(nlon, nlat, ha3, ha2, wrap=True, file_order=True)
        x3 = np.asfortranarray(np.stack([wa[0], wa[1], wf[0], wf[1]], axis=2))
        x2 = np.asfortranarray(np.stack([wa[2], wa[3], wf[2], wf[3]], axis=1))
        q3, q2 = po.grid_pack(nlon, nlat, x3, x2, wrap=True, file_order=True)
        t_cpu = time.perf_counter() - t0
        am3, as3, am2, as2 = res
        ok = (np.array_equal(am3[:, :L].cpu().numpy(), wa[0].transpose(2, 1, 0))
              and np.array_equal(as3[:, :L].cpu().numpy(), wa[1].transpose(2, 1, 0))
              and np.array_equal(am2.cpu().numpy(), wa[2].T) and np.array_equal(as2.cpu().numpy(), wa[3].T)
              and np.array_equal(g3[:K, :, nlev - L:].cpu().numpy(), p3.transpose(4, 3, 2, 1, 0))
              and np.array_equal(g3[K:, :, nlev - L:].cpu().numpy(), q3.transpose(4, 3, 2, 1, 0))
              and np.array_equal(g2[:K].cpu().numpy(), p2.transpose(3, 2, 1, 0))
              and np.array_equal(g2[K:].cpu().numpy(), q2.transpose(3, 2, 1, 0)))
        # the 2-D variables are done in full on both sides; the CPU rate is
        # per 3-D grid point of the two levels
        cpu = {"value": nij * L / t_cpu, "unit": "grid points/s", "cores": 1, "kind": "port",
               "sample": f"{L} of {nlev} levels, all {K} members and variables, FP64 C oracle (one thread), "
                         f"{t_cpu:.2f} s"}
        log(f"  oracle on {L} levels: bit-exact {ok}; cpu {cpu['value']:.3e} grid points/s")
        if not ok:
            raise SystemExit("output step differs from the oracle")
    out = {"metric": "output step (grid mean/spread + file images) grid points per second",
           "value": nij * nlev / (ms * 1e-3), "unit": "grid points/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{wl}: {nlon}x{nlat}x{nlev} K={K} nv3d={nv3d} nv2d={nv2d}, mesp of analysis "
                                  f"and forecast + REAL(4) file images of {K + 4} fields"},
           "stage_ms": {"mesp_x2": ms_mesp, "pack": ms_pack},
           "roofline": {"bound": "hbm", "achieved": gbs(b_mesp + b_pack, ms), "peak": PEAK_HBM_GBS,
                        "unit": "GB/s", "frac": gbs(b_mesp + b_pack, ms) / PEAK_HBM_GBS, "traffic": None},
           "kernel_roofline": {"k_mesp": {"ms": ms_mesp, "achieved": gbs(b_mesp, ms_mesp), "unit": "GB/s"},
                               "k_grid_pack": {"ms": ms_pack, "achieved": gbs(b_pack, ms_pack), "unit": "GB/s"}},
           "cpu_baseline": cpu}
    print(json.dumps(out))


def run_lpfgm(args):
    """The local particle filter and its Gaussian-mixture extension
    (lpfgm_core, common_lpfgm.f90:444-494) batched over local problems of
    the north-star shape (nmem 100, 84 local obs each, the mean count of the
    c1 analysis), 100 resamplings per problem: one step = one
    letkf_lpfgm_core_batched call (mode 3, LPFGM).  Work figure per problem:
    the Gram 2 K^2 n, the Cholesky factor K^3 / 3, its inverse K^3 / 3,
    Pa = L^-T L^-1 K^3 / 3 and the product W_GM W_LPF 2 K^3 flop, FP64 on
    the vector units."""
    import ctypes as C
    import numpy as np
    import torch
    import pyoracle as po
    from sbpom_letkf_v2_amd import _lib, engine, synth
    K, nobsl, nmonte, nb = 100, 84, 100, args.lpfgm_problems
    rng = np.random.default_rng(5)
    offs = (np.arange(nb + 1) * nobsl).astype(np.int64)
    Y = rng.standard_normal((nb * nobsl, K)) * 0.6
    Y -= Y.mean(1, keepdims=True)
    rd = rng.uniform(0.5, 2.0, nb * nobsl) ** 2
    rl = rng.uniform(0.05, 1.0, nb * nobsl)
    dep = rng.standard_normal(nb * nobsl) * 0.6
    draws = np.ascontiguousarray(rng.uniform(size=(nmonte, K)))
    pr = synth.make_problem(nlon=8, nlat=6, nlev=2, nmem=K, nobs=10, seed=1)
    eng = engine.Engine(engine.make_params(pr))
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a)).cuda()
    tO, tY, trd, trl, tdp = dev(offs), dev(Y), dev(rd), dev(rl), dev(dep)
    w = torch.empty((nb, K), dtype=torch.float64, device="cuda")
    W = torch.empty((nb, K, K), dtype=torch.float64, device="cuda")
    vp = lambda t: C.c_void_p(t.data_ptr())
    rp = draws.ctypes.data_as(C.POINTER(C.c_double))
    call = lambda mode: _lib.check(_lib.lib().letkf_lpfgm_core_batched(
        eng.h, mode, 1.0, nmonte, 0, 1, rp, nb, vp(tO), vp(tY), vp(trd), vp(trl), vp(tdp), vp(w), vp(W)))
    call(2)   # LPF alone first (warm), then the timed LPFGM steps leave W
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    call(2)
    torch.cuda.synchronize()
    ms_lpf = (time.perf_counter() - t0) * 1e3
    for _ in range(args.warmup):
        call(3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        call(3)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    flop = nb * (2.0 * K * K * nobsl + 3.0 * K ** 3)
    tf = flop / (ms * 1e-3) / 1e12
    log(f"  LPFGM {ms:.1f} ms per {nb} problems ({nb / (ms * 1e-3):.3e} problems/s, {tf:.2f} TF/s FP64); "
        f"LPF alone {ms_lpf:.1f} ms")
    cpu = None
    if not args.no_cpu:
        Wg = W.cpu().numpy()
        ns = min(nb, 400)
        t0 = time.perf_counter()
        err = 0.0
        for b in range(ns):
            seg = slice(offs[b], offs[b + 1])
            rc, _, Wo = po.lpfgm_core(Y[seg], rd[seg], rl[seg], dep[seg], draws, mode=3, cov_infl_gm=1.0)
            err = max(err, float(np.abs(Wg[b] - Wo).max() / np.abs(Wo).max()))
        t_cpu = time.perf_counter() - t0
        cpu = {"value": ns / t_cpu, "unit": "problems/s", "cores": 1, "kind": "port",
               "sample": f"{ns} of {nb} problems, FP64 C oracle (one thread), {t_cpu:.2f} s"}
        log(f"  oracle on {ns} problems: max rel err {err:.2e}; cpu {cpu['value']:.1f} problems/s")
        if err > 1e-10:
            raise SystemExit("LPFGM differs from the oracle")
    eng.close()
    out = {"metric": "LPFGM local problems per second", "value": nb / (ms * 1e-3), "unit": "problems/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{nb} local problems, nmem {K}, {nobsl} obs each, nmonte {nmonte}, MR + DP"},
           "stage_ms": {"lpfgm": ms, "lpf_only": ms_lpf},
           "roofline": {"bound": "mfma", "achieved": tf, "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
                        "frac": tf / PEAK_TFLOPS["fp64"], "traffic": None},
           "cpu_baseline": cpu}
    print(json.dumps(out))


def run_lpfgm_da(args):
    """da_main with the LPFGM core (iswitch_da 3; letkf_da_main_lpfgm_device)
    on a workload (default c2), 100 resamplings per point, MR with diagonal
    priority, FP64; checked against the FP64 oracle with the same switch on a
    sample of points, which is also the CPU baseline."""
    import numpy as np
    import torch
    import pyoracle as po
    from sbpom_letkf_v2_amd import engine, synth
    wl = args.workload
    log(f"generating workload {wl} on the GPU")
    prd = synth.config(wl, device="cuda", real_kind=8, progress=log)
    meta = prd.meta
    K, nv3, nv2 = meta["nmem"], meta["nv3d"], meta["nv2d"]
    P = meta["nlon"] * meta["nlat"] * meta["nlev"]
    draws = np.random.default_rng(11).uniform(size=(100, K))
    lp = dict(mode=3, rand=draws, cov_infl_gm=1.0, rm="MR", dp=True)
    eng = engine.Engine(engine.make_params(prd))
    a3 = torch.empty_like(prd.fcst3d)
    a2 = torch.empty_like(prd.fcst2d)
    call = lambda: eng.da_main_lpfgm_device(prd, draws, mode=3, anal3d=a3, anal2d=a2)
    for _ in range(args.warmup):
        call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        call()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    eng.close()
    log(f"  LPFGM da_main {ms:.1f} ms per analysis, {P / (ms * 1e-3):.3e} grid points/s")
    cpu = parity = None
    if not args.no_cpu:
        pn_tabs = synth.Problem(
            params=dict(prd.params), **{k: getattr(prd, k) for k in synth.Problem.ARRAYS
                                        if k not in ("fcst3d", "fcst2d")},
            fcst3d=torch.zeros((nv3, K, 1, 1), dtype=torch.float64),
            fcst2d=torch.zeros((max(nv2, 1), K, 1), dtype=torch.float64)).numpy()
        opr = po.Problem(params=dict(pn_tabs.params), **{k: getattr(pn_tabs, k) for k in synth.Problem.ARRAYS})
        cpu, parity, npt = cpu_baseline(opr, prd, a3, a2, args.cpu_seconds, K, nv3, nv2, lpfgm=lp)
        log(f"  cpu {cpu['value']:.0f} pts/s on {cpu['cores']} cores; GPU vs oracle max rel err {parity:.2e} "
            f"over {npt} points")
        if parity > 1e-9:
            raise SystemExit("LPFGM analysis differs from the oracle")
    out = {"metric": "LPFGM analysed grid points per second", "value": P / (ms * 1e-3), "unit": "grid points/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{wl}: {meta['nlon']}x{meta['nlat']}x{meta['nlev']} K={K}, iswitch_da 3 "
                                  "(LPFGM), nmonte 100, MR + DP"},
           "parity_max_rel_err_vs_fp64_oracle_sample": parity, "cpu_baseline": cpu}
    print(json.dumps(out))


def _bits_sum(t):
    """Order-independent 64-bit checksum of a tensor's bit patterns (int32
    words weighted by their position), computed in slices: a replay whose
    output differs from the eager one in any element changes it."""
    import torch
    w = t.reshape(-1).view(torch.int32)
    h = 0
    step = 1 << 26
    for o in range(0, w.numel(), step):
        x = w[o:o + step].to(torch.int64)
        idx = torch.arange(o + 1, o + 1 + x.numel(), device=x.device, dtype=torch.int64)
        h = (h + int(((x & 0xFFFFFFFF) * (2 * idx + 1)).sum().item())) & ((1 << 64) - 1)
    return h


def run_c4(args):
    """Workload c4: eight independent analysis cycles (grids 128^2 .. 512^2 x
    40, nmem 32/64/100/128 round-robin, 0.5 .. 4 random obs per column),
    FP32 r_size I/O.  Eager: each cycle on a fresh handle in sequence,
    throughput = all cycles' points / all cycles' analysis time (chosen
    form), per-cycle error of the chosen form against the FP64 engine
    (batched, FP64 Gram) on the same cycle.  Graph: every cycle planned on
    its own handle (letkf_da_main_plan, workspaces capped at an equal share
    of the free memory), the eight
    enqueues captured in ONE hipGraph on one stream and replayed; a replay
    reproduces every cycle's eager analysis bit for bit (checksums of the
    output bits)."""
    import torch
    from sbpom_letkf_v2_amd import engine, synth
    tot_pts, tot_s, per, probs = 0, 0.0, [], []
    for c, kw in enumerate(synth.c4_cycles()):
        log(f"c4 cycle {c}: {kw['nlon']}x{kw['nlat']}x{kw['nlev']} nmem={kw['nmem']} nobs={kw['nobs']}")
        pr = synth.make_problem(device="cuda", real_kind=4, **kw)
        P = kw["nlon"] * kw["nlat"] * kw["nlev"]
        a3, a2, sec, stage, st = run_form(pr, args.precision, args.form, args.steps, args.warmup)
        r3, r2, _, _, _ = run_form(pr, "fp64", "batched", 1, 0)
        err = field_errors(a3, a2, r3, r2, kw["nv3d"], kw["nv2d"], kw["nmem"])["summary"]
        per.append({"cycle": c, "grid": f"{kw['nlon']}x{kw['nlat']}x{kw['nlev']}", "nmem": kw["nmem"],
                    "nobs": kw["nobs"], "ms": sec * 1e3, "points_per_s": P / sec, "active_points": st["active_points"],
                    "errors_vs_fp64": err})
        log(f"  {sec * 1e3:.1f} ms, {P / sec:.3e} pts/s, errors {err}")
        tot_pts += P
        tot_s += sec
        del r3, r2
        if args.no_graph:
            del pr, a3, a2
        else:
            probs.append((pr, a3, a2, _bits_sum(a3), _bits_sum(a2)))
        torch.cuda.empty_cache()
    graph = None
    if not args.no_graph:
        log("c4: every cycle planned on its own handle, the eight enqueues captured in one graph")
        engs = []
        # the eight handles hold their workspaces together: each gets an equal
        # share of three quarters of the free device memory
        share = int(0.75 * torch.cuda.mem_get_info()[0] / len(probs))
        try:
            t0 = time.perf_counter()
            for pr, a3, a2, _, _ in probs:
                prm = engine.make_params(pr, args.precision, args.form)
                prm.workspace_bytes = share
                eng = engine.Engine(prm)
                eng.plan_device(pr)
                engs.append(eng)
            plan_s = time.perf_counter() - t0
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, capture_error_mode="thread_local"):
                s = torch.cuda.current_stream().cuda_stream
                for eng, (pr, a3, a2, _, _) in zip(engs, probs):
                    eng.enqueue_device(pr, a3, a2, stream=s)
            g.replay()
            torch.cuda.synchronize()
            for eng in engs:
                eng.status()
            same = all(_bits_sum(a3) == h3 and _bits_sum(a2) == h2 for _, a3, a2, h3, h2 in probs)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(args.steps):
                g.replay()
            e1.record()
            torch.cuda.synchronize()
            for eng in engs:
                eng.status()
            ms = e0.elapsed_time(e1) / args.steps
            graph = {"ms_per_replay": ms, "points_per_s": tot_pts / (ms * 1e-3), "plan_s_all_cycles": plan_s,
                     "bit_exact_vs_eager": bool(same), "replays": args.steps,
                     "launch_sequence": "one hipGraph: the enqueues of all eight cycles on one stream"}
            log(f"  {graph}")
        finally:
            for eng in engs:
                eng.close()
    out = {"metric": "analysed grid pts/s over the eight c4 cycles", "value": tot_pts / tot_s,
           "unit": "grid points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": tot_s * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": f"{args.precision}+f64", "data": "synthetic",
           "config": {"workload": "c4: 8 cycles 128^2..512^2 x 40, nmem 32/64/100/128, 0.5..4 obs/column, "
                                  f"r_size FP32, {args.precision}/{args.form}"},
           "cycles": per, "graph": graph}
    print(json.dumps(out))


def calibrate_copy():
    import torch
    n = 1 << 28  # 2^28 float32 = 1 GiB read and 1 GiB written per launch
    x = torch.ones(n, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.mul(x, 1.0001, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "elementwise mul", "read_bytes_per_launch": 4 * n,
                      "write_bytes_per_launch": 4 * n, "launches": 3}))


def profile_traffic(args, kernel):
    """HBM bytes per launch of `kernel` (a kernel family) from the committed
    rocprofv3 summary of this workload: the FETCH_SIZE and WRITE_SIZE passes,
    each scaled by the counter's reading for the 1 GiB copy kernel."""
    import glob
    pat = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                       f"kernels_{args.workload}_*_{args.form}_{args.precision}.md")
    for path in sorted(glob.glob(pat)):
        for line in open(path):
            cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
            if len(cells) == 6 and cells[0] == kernel and cells[1] == "traffic":
                try:
                    return float(cells[5])
                except ValueError:
                    return None
    return None


def run_adjoint(args):
    """Loss = squared misfit of the analysis mean to the synthetic truth;
    gradient w.r.t. obs values, per-obs log sigma^2, RTPP alpha and rho."""
    import numpy as np
    import torch
    from sbpom_letkf_v2_amd import autograd as ag
    from sbpom_letkf_v2_amd import synth
    wl = args.workload if args.workload != "c2" else "c3"
    log(f"generating workload {wl} on the GPU")
    pr = synth.config(wl, device="cuda", real_kind=8, keep_truth=True, progress=log)
    meta = pr.meta
    K, nv3 = meta["nmem"], meta["nv3d"]
    P = meta["nlon"] * meta["nlat"] * meta["nlev"]
    t3 = pr.truth3d.reshape(nv3, meta["nlev"], -1)

    def loss(a3, a2):
        return ((a3.mean(1) - t3) ** 2).sum()

    st = ag.AnalysisState(pr, "fp64", "plain")
    theta = [pr.obsdep.clone(), torch.log(pr.obserr.clone() ** 2),
             torch.tensor(float(pr.params["alpha_rtpp"]), dtype=torch.float64, device="cuda"),
             torch.tensor(float(pr.params.get("cov_infl_mul", 1.0)), dtype=torch.float64, device="cuda")]

    def value(th):
        with torch.no_grad():
            a3, a2 = ag.analysis(st, *th)
            return loss(a3, a2).item()

    fwd, bwd = [], []
    for s in range(args.warmup + args.steps):
        th = [t.clone().requires_grad_(True) for t in theta]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        a3, a2 = ag.analysis(st, *th)
        L = loss(a3, a2)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        L.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if s >= args.warmup:
            fwd.append(t1 - t0)
            bwd.append(t2 - t1)
        log(f"  step {s}: forward {1e3 * (t1 - t0):.1f} ms, adjoint {1e3 * (t2 - t1):.1f} ms")
    g = [t.grad.detach() for t in th]
    gmax = max(float(x.abs().max()) for x in g)
    rng = np.random.default_rng(20261203)
    h = 1e-4
    errs = []
    nc = args.fd_coords
    picks = [(0, int(j)) for j in rng.choice(theta[0].numel(), nc // 2 - 1, False)] + \
            [(1, int(j)) for j in rng.choice(theta[1].numel(), nc - nc // 2 - 1, False)] + [(2, None), (3, None)]
    t0 = time.perf_counter()
    for which, j in picks:
        tp = [t.clone() for t in theta]
        tm = [t.clone() for t in theta]
        if j is None:
            tp[which] += h
            tm[which] -= h
            ad = float(g[which])
        else:
            tp[which][j] += h
            tm[which][j] -= h
            ad = float(g[which][j])
        fd = (value(tp) - value(tm)) / (2 * h)
        errs.append(abs(fd - ad) / max(abs(ad), 1e-3 * gmax))
    log(f"  FD check on {len(picks)} coordinates in {time.perf_counter() - t0:.0f}s: "
        f"max rel err {max(errs):.2e}")
    st.close()
    fm, bm = sum(fwd) / len(fwd), sum(bwd) / len(bwd)
    out = {"metric": "adjoint of the analysis: analysed grid pts/s (forward + adjoint)",
           "value": P / (fm + bm), "unit": "grid points/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": 1e3 * (fm + bm), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{wl}: {meta['nlon']}x{meta['nlat']}x{meta['nlev']} K={K} nv3d={nv3} "
                                  f"nobs={meta['nobs']}, loss = squared misfit of the analysis mean to the truth"},
           "adjoint": {"forward_ms": 1e3 * fm, "adjoint_ms": 1e3 * bm, "fd_coords": len(picks), "fd_h": h,
                       "fd_max_rel_err": max(errs), "fd_median_rel_err": float(np.median(errs)),
                       "grad_max": {"dep": float(g[0].abs().max()), "logvar": float(g[1].abs().max()),
                                    "alpha": float(g[2]), "rho": float(g[3])}}}
    print(json.dumps(out), flush=True)


def run_form(prd, precision, form, steps, warmup, autotune=False):
    import torch
    from sbpom_letkf_v2_amd import engine
    prm = engine.make_params(prd, precision, form)
    prm.stage_timing = 2        # stage marks + every eigensolve kernel launch
    eng = engine.Engine(prm)
    a3 = torch.empty_like(prd.fcst3d)
    a2 = torch.empty_like(prd.fcst2d)
    tune = None
    if autotune:
        t = time.perf_counter()
        _, _, times, choice = eng.autotune(prd, a3, a2)
        tune = {"seconds": time.perf_counter() - t, "bucket_ms_per_path": times, "paths": choice}
        log(f"  autotune: {tune}")
    for w in range(warmup):
        t = time.perf_counter()
        eng.da_main_device(prd, a3, a2)
        log(f"  warmup {w}: {time.perf_counter() - t:.2f}s {eng.stats()}")
    torch.cuda.synchronize()
    tot = {k: 0.0 for k in ("ms_search", "ms_gram", "ms_eig", "ms_transform", "ms_total")}
    kms = [0.0] * 25
    kn = [0] * 25
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        eng.da_main_device(prd, a3, a2)
        st = eng.stats()
        for k in tot:
            tot[k] += st[k]
        for q in range(25):
            kms[q] += st["ms_eigk"][q]
            kn[q] += st["n_eigk"][q]
        log(f"  step {s}: {st['ms_total']:.1f} ms")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = eng.stats()
    eng.close()
    stage = {k: v / steps for k, v in tot.items()}
    kinds = EIG_KINDS_RAT if form == "batched" else EIG_KINDS
    stage["eig_kernels"] = {kinds[q // 5] + f"<{(16, 32, 48, 64, 128)[q % 5]}>": {"ms": kms[q] / steps,
                                                                                    "launches": kn[q] / steps}
                            for q in range(25) if kn[q]}
    stage["autotune"] = tune
    return a3, a2, wall / steps, stage, st


def graph_replay(prd, precision, form, steps, a3_eager, a2_eager, paths=None):
    """The same analysis planned once (letkf_da_main_plan), its enqueue
    captured in a graph and replayed: ms per replay, and whether a replay
    reproduces the eager analysis bit for bit."""
    import torch
    from sbpom_letkf_v2_amd import engine
    eng = engine.Engine(engine.make_params(prd, precision, form))
    try:
        if paths:
            eng.set_tuning(paths)
        t = time.perf_counter()
        eng.plan_device(prd)
        plan_ms = (time.perf_counter() - t) * 1e3
        a3, a2 = torch.empty_like(prd.fcst3d), torch.empty_like(prd.fcst2d)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, capture_error_mode="thread_local"):
            eng.enqueue_device(prd, a3, a2, stream=torch.cuda.current_stream().cuda_stream)
        g.replay()
        torch.cuda.synchronize()
        eng.status()
        same = bool(torch.equal(a3, a3_eager) and torch.equal(a2, a2_eager))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            g.replay()
        e1.record()
        torch.cuda.synchronize()
        eng.status()
        ms = e0.elapsed_time(e1) / steps
    finally:
        eng.close()
    return {"ms_per_replay": ms, "plan_ms": plan_ms, "bit_exact_vs_eager": same}


def cpu_baseline(pr_host_tables, prd, a3_gpu, a2_gpu, target_s, K, nv3, nv2, lpfgm=None):
    """FP64 oracle (the C restatement, OpenMP) on a uniform random sample of
    points of the same workload; also returns the GPU-vs-oracle error at the
    sampled points (full-size parity)."""
    import numpy as np
    import torch
    import pyoracle as po
    nij1, nlev = pr_host_tables.depth1.shape
    P = nij1 * nlev
    cores, host = host_cores()
    rng = np.random.default_rng(7)

    def sample(npt):
        pts = np.sort(rng.choice(P, npt, replace=False))
        pk = (pts // nij1).astype(np.int32)
        pi = (pts % nij1).astype(np.int32)
        tp = torch.as_tensor(pts, device=prd.fcst3d.device)
        # fcst3d tensor memory is [v][m][p]
        x3 = prd.fcst3d.reshape(nv3, K, P)[:, :, tp].permute(2, 0, 1).contiguous().cpu().numpy()
        ti = torch.as_tensor(pi.astype(np.int64), device=prd.fcst3d.device)
        x2 = prd.fcst2d.reshape(nv2, K, nij1)[:, :, ti].permute(2, 0, 1).contiguous().cpu().numpy() \
            if nv2 else np.zeros((npt, 0, K))
        g3 = a3_gpu.reshape(nv3, K, P)[:, :, tp].permute(2, 0, 1).contiguous().cpu().numpy()
        g2 = a2_gpu.reshape(nv2, K, nij1)[:, :, ti].permute(2, 0, 1).contiguous().cpu().numpy() \
            if nv2 else None
        return pi, pk, x3, x2, g3, g2

    # calibrate on a small sample, then size the timed sample to ~target_s
    pi, pk, x3, x2, _, _ = sample(64 * cores)
    t0 = time.perf_counter()
    po.analyse_points(pr_host_tables, pi, pk, x3, x2, nthreads=cores, lpfgm=lpfgm)
    dt = time.perf_counter() - t0
    npt = int(min(P, max(64 * cores, 64 * cores * target_s / max(dt, 1e-3))))
    pi, pk, x3, x2, g3, g2 = sample(npt)
    t0 = time.perf_counter()
    rc, xa3, xa2, nl = po.analyse_points(pr_host_tables, pi, pk, x3, x2, nthreads=cores, lpfgm=lpfgm)
    dt = time.perf_counter() - t0
    assert rc == 0, rc
    err3 = float(np.abs(g3 - xa3).max() / np.abs(xa3).max())
    surf = pk == nlev - 1
    err2 = float(np.abs(g2[surf] - xa2[surf]).max() / np.abs(xa2[surf]).max()) if (nv2 and surf.any()) else 0.0
    return dict(value=npt / dt, unit="grid points/s", cores=cores, kind="port", host=host,
                sample=f"{npt} uniformly sampled (i,k) points of the same workload, "
                       f"FP64 C oracle with {cores} OpenMP threads, {dt:.1f} s"), max(err3, err2), npt


def host_cores():
    """Threads for the CPU baseline: every core this process may run on,
    capped by the host's per-GPU share when the environment sets one
    (OMP_NUM_THREADS), with the counts it was chosen from."""
    aff = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    cores = min(aff, share) if share > 0 else aff
    return cores, {"nproc": os.cpu_count(), "affinity": aff, "omp_num_threads": share or None,
                   "threads_used": cores}


def core_calibration(seconds=6.0):
    """The oracle's letkf_core against the reference's own letkf_core
    (common_letkf.f90:61-210, compiled unmodified: the shipped REAL(4) build
    at its fixed nmem = 128) on identical local problems, both in the
    shipped REAL(4) arithmetic, on the same threads: how fast the port that
    the CPU baseline times runs next to the reference.  None where the
    reference build is absent."""
    import concurrent.futures as cf
    import ctypes as C
    import numpy as np
    import pyoracle as po
    ref = po.ref()
    if ref is None:
        return None
    K = ref.ref_nmem()
    cores, _ = host_cores()
    rng = np.random.default_rng(11)
    sizes = [int(n) for n in rng.choice([40, 120, 300, 600], 64)]
    probs = []
    for n in sizes:
        Y = rng.standard_normal((n, K)).astype(np.float32)
        Y -= Y.mean(1, keepdims=True)
        probs.append((np.asfortranarray(Y), (rng.uniform(0.3, 2.0, n) ** 2).astype(np.float32),
                      rng.uniform(0.01, 1.0, n).astype(np.float32), rng.standard_normal(n).astype(np.float32)))
    F = C.POINTER(C.c_float)

    def run_ref(pb):
        Y, rd, rl, dp = pb
        w = np.zeros(K, np.float32)
        W = np.zeros((K, K), np.float32, order="F")
        ref.ref_letkf_core(Y.shape[0], Y.ctypes.data_as(F), rd.ctypes.data_as(F), rl.ctypes.data_as(F),
                           dp.ctypes.data_as(F), w.ctypes.data_as(F), W.ctypes.data_as(F))
        return W

    def run_port(pb):
        Y, rd, rl, dp = pb
        return po.letkf_core(Y, rd, rl, dp, 1.0, dtype=np.float32)[2]

    out = {}
    with cf.ThreadPoolExecutor(cores) as ex:
        for name, fn in (("reference", run_ref), ("port", run_port)):
            list(ex.map(fn, probs[:cores]))          # warm-up
            reps, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < seconds / 2 or reps == 0:
                res = list(ex.map(fn, probs))
                reps += 1
            out[name] = (time.perf_counter() - t0) / (reps * len(probs))
            out[name + "_W"] = res
    err = max(float(np.abs(a - b).max() / np.abs(b).max()) for a, b in zip(out["port_W"], out["reference_W"]))
    return {"nmem": K, "problems": len(probs), "local_obs": sorted(set(sizes)), "threads": cores,
            "reference_ms_per_problem": 1e3 * out["reference"], "port_ms_per_problem": 1e3 * out["port"],
            "port_speed_vs_reference": out["reference"] / out["port"], "port_vs_reference_max_rel_diff_W": err}


def field_errors(a3, a2, r3, r2, nv3, nv2, K):
    """Per-variable max and RMS error of the analysis ensemble mean and
    spread against the FP64 analysis, relative to the max |FP64 mean| and
    to the max FP64 spread of that variable."""
    import torch
    out = {}
    worst_mean = worst_sprd = 0.0
    names = ["U", "V", "T", "S", "TRACER"][:nv3] + ["SSH"][:nv2]
    for v in range(nv3 + nv2):
        if v < nv3:
            x, r = a3.reshape(nv3, K, -1)[v], r3.reshape(nv3, K, -1)[v]
        else:
            x, r = a2.reshape(nv2, K, -1)[v - nv3], r2.reshape(nv2, K, -1)[v - nv3]
        xm, rm = x.mean(0), r.mean(0)
        xs, rs = x.std(0), r.std(0)
        dm, ds = (xm - rm).abs(), (xs - rs).abs()
        sm, ss = rm.abs().max().item(), rs.max().item()
        e = {"mean_max": dm.max().item() / sm, "mean_rms": dm.pow(2).mean().sqrt().item() / sm,
             "sprd_max": ds.max().item() / ss, "sprd_rms": ds.pow(2).mean().sqrt().item() / ss}
        out[names[v] if v < len(names) else f"v{v}"] = e
        worst_mean = max(worst_mean, e["mean_max"])
        worst_sprd = max(worst_sprd, e["sprd_max"])
    out["summary"] = {"mean_max_rel": worst_mean, "sprd_max_rel": worst_sprd}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the analysis shards "
              "by grid points across GPUs with no exchange inside the path -- run N=1.",
              file=sys.stderr)
        sys.exit(2)
    if args.adjoint:
        run_adjoint(args)
        return
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.setobs:
        return run_setobs(args)
    if args.output:
        return run_output(args)
    if args.lpfgm:
        return run_lpfgm(args)
    if args.lpfgm_da:
        return run_lpfgm_da(args)
    if args.xtoy:
        run_xtoy(args)
        return
    if args.workload == "c4":
        run_c4(args)
        return
    import numpy as np
    import torch
    import pyoracle as po
    from sbpom_letkf_v2_amd import synth

    t0 = time.perf_counter()
    log(f"generating workload {args.workload} on the GPU")
    prd = synth.config(args.workload, device="cuda", real_kind=8, progress=log)
    torch.cuda.synchronize()
    meta = prd.meta
    K, nv3, nv2 = meta["nmem"], meta["nv3d"], meta["nv2d"]
    nij1 = meta["nlon"] * meta["nlat"]
    P = nij1 * meta["nlev"]
    log(f"generated in {time.perf_counter() - t0:.1f}s: P={P} K={K} nobs={meta['nobs']}")

    log(f"timing chosen form precision={args.precision} form={args.form}: "
        f"{args.warmup} warmup + {args.steps} steps")
    a3, a2, sec, stage, st = run_form(prd, args.precision, args.form, args.steps, args.warmup,
                                      autotune=not args.no_autotune)
    log(f"  {sec * 1e3:.1f} ms/step, stages {stage}")
    result_extra = {}
    ref3 = ref2 = None                      # FP64 analysis the errors are measured against
    plain_sec = None
    if (args.precision, args.form) != ("fp64", "plain") and not args.no_plain:
        log("timing plain FP64 form")
        ref3, ref2, psec, pstage, pst = run_form(prd, "fp64", "plain", max(1, args.steps // 2), 1)
        plain_sec = psec
        result_extra["plain_fp64"] = {"value": P / psec, "ms_per_step": psec * 1e3,
                                      "stage_ms": {k: v for k, v in pstage.items() if k.startswith("ms_")}}
        result_extra["speedup_vs_plain_fp64"] = psec / sec
        result_extra["errors_vs_fp64"] = field_errors(a3, a2, ref3, ref2, nv3, nv2, K)
        log(f"  speedup {psec / sec:.2f}x; errors {result_extra['errors_vs_fp64']['summary']}")
    elif args.precision == "fp64":
        ref3, ref2 = a3, a2
    if not args.no_graph:
        log("chosen form planned once, enqueue captured in a graph, replayed")
        result_extra["graph_replay"] = graph_replay(prd, args.precision, args.form, args.steps, a3, a2,
                                                    (stage.get("autotune") or {}).get("paths"))
        log(f"  {result_extra['graph_replay']}")
    # --- roofline of the dominant kernel: HIP events around its launches in the timed steps ---
    L = st["local_obs"]
    na = st["active_points"]
    nch = int(st["chunks"])
    ek = stage["eig_kernels"]

    def family(kind):
        ms = sum(v["ms"] for k, v in ek.items() if k.startswith(kind + "<"))
        n = sum(v["launches"] for k, v in ek.items() if k.startswith(kind + "<"))
        return ms, n

    tri_ms, tri_n = family("k_eigr_tridiag")
    ql_ms, ql_n = family("k_eigr_ql")
    vr_ms, vr_n = family("k_eigr_vrot")
    rot = float(st.get("rotations", 0))
    gram_kernel = "k_gram_fp64" if args.precision == "fp64" or K > 128 else "k_gram_mfma"
    kernels = {
        # Gram: 2 n K^2 (primal) or 2 n^2 K (dual) flops per point, one launch per chunk
        gram_kernel: dict(ms=stage["ms_gram"], launches=nch, work=st["gram_flops"], unit="TFLOP/s",
                          bound="valu" if gram_kernel == "k_gram_fp64" else "mfma",
                          peak=PEAK_TFLOPS[args.precision]),
        # Householder tridiagonalisation: (4/3) m^3 flops per point on the vector ALUs, FP64 (FP32 for
        # bf16x3t in the batched form: 157.3 TFLOP/s)
        "k_eigr_tridiag": dict(ms=tri_ms, launches=tri_n, work=st["eig_flops"], unit="TFLOP/s",
                               bound="valu", peak=157.3 if (args.precision == "bf16x3t" and args.form == "batched")
                               else PEAK_FP64_TFLOPS),
        # implicit QL: ~30 FP64 flops per logged rotation (Givens pair, refined rsqrt, shift update)
        "k_eigr_ql": dict(ms=ql_ms, launches=ql_n, work=30.0 * rot, unit="TFLOP/s",
                          bound="valu", peak=PEAK_FP64_TFLOPS),
        # rotation replay: every logged rotation on the 1+nv3d+nv2d vectors, forward and back, 6 flops each
        "k_eigr_vrot": dict(ms=vr_ms, launches=vr_n, work=12.0 * (1 + nv3 + nv2) * rot, unit="TFLOP/s",
                            bound="valu", peak=PEAK_FP64_TFLOPS),
        # tile update (points without local obs and overflowed eigensolves): read + write of the ensemble
        "k_transform": dict(ms=stage["ms_transform"], launches=nch, work=st["transform_bytes"],
                            unit="GB/s", bound="hbm", peak=PEAK_HBM_GBS),
    }
    kernels = {k: v for k, v in kernels.items() if v["launches"] and v["ms"] > 0 and v["work"] > 0}
    dom_name = max(kernels, key=lambda k: kernels[k]["ms"])
    dom = kernels[dom_name]
    scale = 1e12 if dom["unit"] == "TFLOP/s" else 1e9
    achieved = dom["work"] / (dom["ms"] * 1e-3) / scale
    # bound: "hbm" (8 TB/s), "mfma" (the matrix cores at the Gram's operand type) or "valu" (FP64 on the
    # vector ALUs: 78.6 TFLOP/s on MI355X, the same figure as its FP64 matrix peak)
    roofline = {"bound": dom["bound"], "achieved": achieved, "peak": dom["peak"], "unit": dom["unit"],
                "frac": achieved / dom["peak"], "traffic": profile_traffic(args, dom_name), "kernel": dom_name,
                "launches_per_step": dom["launches"], "avg_launch_ms": dom["ms"] / dom["launches"],
                "work_per_launch": dom["work"] / dom["launches"]}
    kernel_roofline = {}
    for k, v in kernels.items():
        sc = 1e12 if v["unit"] == "TFLOP/s" else 1e9
        a = v["work"] / max(v["ms"] * 1e-3, 1e-12) / sc
        kernel_roofline[k] = {"ms": v["ms"], "launches": v["launches"], "achieved": a, "unit": v["unit"],
                              "frac": a / v["peak"]}
    stage_ms = {k: v for k, v in stage.items() if k.startswith("ms_")}
    if stage.get("autotune"):
        result_extra["autotune"] = stage["autotune"]
    # --- CPU baseline + full-size parity at the sampled points ---
    cpu = None
    parity = None
    if not args.no_cpu:
        log("CPU baseline (FP64 oracle) on a bounded sample")
        pn_tabs = synth.Problem(
            params=dict(prd.params), **{k: getattr(prd, k) for k in synth.Problem.ARRAYS
                                        if k not in ("fcst3d", "fcst2d")},
            fcst3d=torch.zeros((nv3, K, 1, 1), dtype=torch.float64),
            fcst2d=torch.zeros((max(nv2, 1), K, 1), dtype=torch.float64)).numpy()
        opr = po.Problem(params=dict(pn_tabs.params),
                         **{k: getattr(pn_tabs, k) for k in synth.Problem.ARRAYS})
        if ref3 is not None:
            cpu, parity, npt = cpu_baseline(opr, prd, ref3, ref2, args.cpu_seconds, K, nv3, nv2)
            log(f"  cpu {cpu['value']:.0f} pts/s on {cpu['cores']} cores; GPU vs oracle max rel err "
                f"{parity:.2e} over {npt} points")
    if not args.no_variants and args.workload == "c2":
        alt = "fp32" if args.precision == "bf16x3" else "bf16x3"
        if ref3 is not None:
            log(f"timing the {alt} Gram (batched form) beside the chosen one")
            b3, b2, bsec, bstage, _ = run_form(prd, alt, "batched", 2, 1)
            result_extra[f"{alt}_batched"] = {
                "value": P / bsec, "ms_per_step": bsec * 1e3,
                "speedup_vs_plain_fp64": plain_sec / bsec if plain_sec else None,
                "errors_vs_fp64": field_errors(b3, b2, ref3, ref2, nv3, nv2, K)["summary"],
                "stage_ms": {k: v for k, v in bstage.items() if k.startswith("ms_")}}
            log(f"  {result_extra[f'{alt}_batched']}")
            del b3, b2
        if ref3 is not None and args.form == "batched":
            log("timing the batched form with the QL eigen-decomposition of T (batched_ql) beside it")
            # (the FP32 reduction of bf16x3t belongs to the rational path: the QL form runs it in FP64)
            qprec = "bf16x3" if args.precision == "bf16x3t" else args.precision
            q3, q2, qsec, qstage, _ = run_form(prd, qprec, "batched_ql", 2, 1)
            result_extra["batched_ql"] = {
                "precision": qprec, "value": P / qsec, "ms_per_step": qsec * 1e3,
                "speedup_vs_plain_fp64": plain_sec / qsec if plain_sec else None,
                "errors_vs_fp64": field_errors(q3, q2, ref3, ref2, nv3, nv2, K)["summary"],
                "stage_ms": {k: v for k, v in qstage.items() if k.startswith("ms_")}}
            log(f"  {result_extra['batched_ql']}")
            del q3, q2
        if not args.no_cpu:
            log("CPU: the oracle's letkf_core next to the reference's own (REAL(4), nmem 128)")
            result_extra["cpu_port_vs_reference_core"] = core_calibration()
            log(f"  {result_extra['cpu_port_vs_reference_core']}")
    out = {
        "metric": METRIC, "value": P / sec, "unit": "grid points/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": sec * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "bf16x3t": "bf16x3+f32+f64"}.get(args.precision, f"{args.precision}+f64"),
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {meta['nlon']}x{meta['nlat']}x{meta['nlev']} "
                               f"K={K} nv3d={nv3} nv2d={nv2} nobs={meta['nobs']} "
                               f"({meta['obs_layout']}), {args.precision}/{args.form}",
                   "points": P, "active_points": na, "local_obs": L,
                   "precision": args.precision, "form": args.form},
        "roofline": roofline, "kernel_roofline": kernel_roofline, "stage_ms": stage_ms,
        "eig_kernel_ms": {k: v["ms"] for k, v in ek.items()}, "cpu_baseline": cpu,
        "parity_max_rel_err_vs_fp64_oracle_sample": parity,
    }
    out.update(result_extra)
    if not args.no_variants and args.workload == "c2":
        del prd, a3, a2, ref3, ref2
        torch.cuda.empty_cache()
        out["c1"] = c1_record(args)
    print(json.dumps(out), flush=True)


def c1_record(args):
    """The target configuration (512x512x40, K=100, 5 3-D variables with
    the tracer + SSH, 2e5 obs): the chosen form next to plain FP64 on the
    same problem, the analysis errors between them, and the workload of
    both forms (sum of m^3 over the eigensolves, Gram flops)."""
    import torch
    from sbpom_letkf_v2_amd import synth
    log("c1 record: generating 512x512x40 K=100 with the tracer on the GPU")
    prd = synth.config("c1", device="cuda", real_kind=8)
    torch.cuda.synchronize()
    m = prd.meta
    K, nv3, nv2 = m["nmem"], m["nv3d"], m["nv2d"]
    P = m["nlon"] * m["nlat"] * m["nlev"]
    a3, a2, sec, stage, st = run_form(prd, args.precision, args.form, args.c1_steps, 1)
    log(f"  c1 chosen: {sec * 1e3:.1f} ms/step")
    r3, r2, psec, pstage, pst = run_form(prd, "fp64", "plain", args.c1_steps, 1)
    log(f"  c1 plain FP64: {psec * 1e3:.1f} ms/step")
    err = field_errors(a3, a2, r3, r2, nv3, nv2, K)
    rec = {
        "workload": f"c1: {m['nlon']}x{m['nlat']}x{m['nlev']} K={K} nv3d={nv3} (tracer) nv2d={nv2} "
                    f"nobs={m['nobs']} ({m['obs_layout']}), {args.precision}/{args.form}",
        "value": P / sec, "unit": "grid points/s", "steps": args.c1_steps, "ms_per_step": sec * 1e3,
        "plain_fp64_value": P / psec, "plain_fp64_ms_per_step": psec * 1e3,
        "speedup_vs_plain_fp64": psec / sec, "target_speedup": 4.0,
        "errors_vs_fp64": err,
        "stage_ms": {k: v for k, v in stage.items() if k.startswith("ms_")},
        "plain_stage_ms": {k: v for k, v in pstage.items() if k.startswith("ms_")},
        "eig_kernel_ms": {k: v["ms"] for k, v in stage["eig_kernels"].items()},
        "work": {"active_points": st["active_points"], "dual_points": st["dual_points"],
                     "eig_flops_chosen": st["eig_flops"], "eig_flops_plain": pst["eig_flops"],
                     "gram_flops_chosen": st["gram_flops"], "gram_flops_plain": pst["gram_flops"],
                     "eig_flops_ratio_plain_vs_chosen": pst["eig_flops"] / max(st["eig_flops"], 1.0)},
    }
    log(f"  c1 speedup {psec / sec:.2f}x; errors {err['summary']}")
    del a3, a2, r3, r2, prd
    torch.cuda.empty_cache()
    return rec


if __name__ == "__main__":
    main()

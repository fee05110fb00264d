"""Port-solve throughput of the MI355X B-spline path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload array32x32]

Default workload (the headline): the 32 x 32 dipole array, 15,360 unknowns
and 1,024 ports, solved by the mixed form (FP64 fill, FP32 factor and
all-port solve on MFMA, FP64-residual refinement, adjoint-corrected readout)
against the FP64 GPU path.  One step = one full port solution held on the GPU:
same-edge table, Galerkin fill of Z, factorisation, the all-port solve,
refinement and the Y readout, leaving Y and the coefficient columns in HBM.
Geometry tables are uploaded once before timing, as the reference caches
geometry and basis across solves.  Beside it: the FP64 rate, their ratio, the
Y error per correction, the graph-replay rate and the compute_port_solution()
rate (Y and coeffs copied to the host).

Other workloads: array8x8 (8 x 8 array, FP64, against the reference's golden
Y), sweep128, ground4x4, readout8x8, groundfield4x4.

Prints ONE JSON line: port-solves/s, the fill's Gentries/s, the per-stage
split measured with HIP events on the launch stream inside the timed region,
the roofline of the dominant kernel, the Y error, and a CPU baseline (the
reference's own compiled accelerator + LAPACK on this host's cores, on a
bounded sample).
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6     # MI355X dense FP64 (vector and matrix)
PEAK_FP32_TFLOPS = 157.3    # MI355X dense FP32 matrix
METRIC = "port-solves/s & fill Gentries/s; mixed/learned vs FP64 GPU at Y rel err ≤1e-8"

WORKLOADS = {
    "array8x8": dict(desc="8x8 array of z-directed 0.48-lambda dipoles, 41 segments each, "
                          "0.6-lambda pitch, +-0.02-lambda seeded jitter (seed 1), radius "
                          "1e-3 lambda, free space, 300 MHz, 64 centre feeds",
                     args=(8, 8, 41, 1), golden="c2_array8x8"),
    "array32x32": dict(desc="32x32 array of z-directed 0.48-lambda dipoles, 15 segments each "
                            "(15,360 unknowns), 0.6-lambda pitch, +-0.02-lambda seeded jitter "
                            "(seed 2), radius 1e-3 lambda, free space, 300 MHz, 1024 centre "
                            "feeds; mixed precision (FP32 LU on MFMA, FP64 residual refinement) "
                            "against the FP64 GPU path",
                       args=(32, 32, 15, 2), golden=None),
    "array32x32x41": dict(desc="32x32 array of z-directed 0.48-lambda dipoles, 41 segments each "
                               "(41,984 unknowns: the 8x8 array's element geometry), 0.6-lambda "
                               "pitch, +-0.02-lambda seeded jitter (seed 2), radius 1e-3 lambda, "
                               "free space, 300 MHz, 1024 centre feeds; mixed precision against "
                               "the FP64 GPU path",
                          args=(32, 32, 41, 2), golden=None),
    "ground4x4": dict(desc="4x4 array of horizontal x-directed 0.48-lambda dipoles, 21 segments "
                           "each, 0.5-lambda pitch, radius 1e-3 lambda, over lossy Sommerfeld "
                           "ground (seed 4: height 0.946 lambda, eps_r 17.8, sigma 0.0456 S/m, "
                           "5.18 MHz), 16 centre feeds; one step = one cold port solve (new "
                           "ground: grid cache cleared), Sommerfeld grid from the learned "
                           "surrogate with exact-engine fallback, FP64 fill and LU",
                      args=None, golden="c5_ground4x4"),
    "readout8x8": dict(desc="readout of all 64 port columns of the 8x8 array (the array8x8 "
                            "structure, 2,624 elements): element moments and the far-field "
                            "moments M_theta, M_phi on a 1-degree grid (theta 0-180, phi "
                            "0-359: 65,160 directions), free space, FP64",
                       args=(8, 8, 41, 1), golden=None),
    "groundfield4x4": dict(desc="Sommerfeld ground remainder of E (reflected_field_at) for all 16 "
                                "port columns of the grounded 4x4 array (the ground4x4 structure, "
                                "elements at 8 per segment: 2,688 sources) on a 64 x 64 grid of "
                                "points 0.1-1.1 wavelength above the ground, FP64",
                           args=None, golden=None),
    "hm16x16": dict(desc="16x16 array of z-directed 0.48-lambda dipoles, 15 segments each (3,840 "
                         "unknowns), 0.6-lambda pitch, +-0.02-lambda seeded jitter (seed 2), "
                         "radius 1e-3 lambda, free space, 300 MHz, 256 centre feeds; "
                         "block-low-rank operator (aca_eta 1, leaf 32, aca_tol 1e-4) and block "
                         "GMRES to solve_tol 1e-6, FP64",
                    args=(16, 16, 15, 2), golden="hm_array16x16"),
    "hm32x32": dict(desc="the array32x32 structure (15,360 unknowns, 1024 centre feeds) on the "
                         "block-low-rank operator (aca_eta 1, leaf 32, aca_tol 1e-4) and block "
                         "GMRES to solve_tol 1e-6, FP64",
                    args=(32, 32, 15, 2), golden=None),
    "sweep128": dict(desc="128 seeded random polyline structures (1-12 wires with shared "
                          "junction nodes, 43-827 unknowns) x 16 frequencies 0.8-1.2 x 300 MHz, "
                          "1-4 feeds, FP64; all fills and factorisations in one HIP graph",
                     args=None, golden=None),
}


def sweep_main(args, wl):
    """Batched sweep: 128 random structures x 16 frequencies, every fill and
    factorisation in one replayed HIP graph."""
    import torch

    from momwire_amd import workloads
    from momwire_amd.bspline import BSplineSolver
    from momwire_amd.sweep import BatchedSweep

    structs = workloads.sweep_structures(128, seed=3)
    ks = workloads.sweep_wavenumbers(16)
    t0 = time.perf_counter()
    bs = BatchedSweep(structs, ks)
    setup = time.perf_counter() - t0
    nbs = [D.nb for D in bs.devs]
    nprob = len(bs.problems)
    print(f"[bench] sweep: {nprob} problems, unknowns {min(nbs)}-{max(nbs)} (mean {np.mean(nbs):.0f}),"
          f" set-up {setup:.1f} s, autotune {bs.tuning} -> variant {bs.gemm_variant}",
          file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        bs.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bs.run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    ys = bs.y_matrices()
    errs = []
    for j in range(0, nprob, max(1, nprob // 12)):
        i, kk, y = ys[j]
        s = BSplineSolver(**structs[i])
        s._set_k(kk)
        ref = s.compute_port_solution().y
        errs.append(float(np.abs(y - ref).max() / np.abs(ref).max()))
    nk = len(ks)
    lu_fl = sum(nk * (lu_flops(n) + 8.0 * n * n * (bs.devs[i].ncols)) for i, n in enumerate(nbs))
    fill_fl = sum(nk * fill_flops(int(bs.devs[i].mw.n_seg), n) for i, n in enumerate(nbs))
    flops = lu_fl + fill_fl
    out = {
        "metric": METRIC, "value": nprob / dt, "unit": "port-solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * dt,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl["desc"], "problems": nprob, "unknowns_min": min(nbs),
                   "unknowns_max": max(nbs), "unknowns_mean": float(np.mean(nbs)),
                   "step": "one replay of the captured sweep graph (all problems)"},
        "autotune_s_per_sweep": {str(k): v for k, v in bs.tuning.items()},
        "gemm_variant": bs.gemm_variant,
        "solve_tflops_equiv": flops / dt / 1e12,
        "y_rel_err_vs_single_path_sampled": max(errs),
        "roofline": {"kernel": "the whole sweep graph (batched fills, panel factors, MFMA "
                               "updates and solves of every problem)",
                     "bound": "mfma", "achieved": flops / dt / 1e12, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": flops / dt / 1e12 / PEAK_FP64_TFLOPS,
                     "traffic": None,
                     "algorithmic": "sum over problems of the fill (16*53*N_seg^2 + "
                                    "726*n^2) and (8/3) n^3 + 8 n^2 p "
                                    f"= {fill_fl:.3e} + {lu_fl:.3e} flop per sweep"},
    }
    print(json.dumps(out))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    cores = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    return min(cores, int(omp)) if omp else cores


def stage_means(timers):
    return {k: float(np.mean([a.elapsed_time(b) for a, b in v])) for k, v in timers.items()}


def mixed_main(args, wl):
    """The headline: the 32x32 array solved by the mixed form against the FP64
    GPU path.  Timed region: K eager mixed solves, each stage bracketed by HIP
    events on the launch stream (the roofline's kernel durations come from
    these events, inside the timed region); then the same for FP64; then the
    graph-replay and the compute_port_solution() (Y and coeffs copied to the
    host) rates of both, and the mixed form's Y error against FP64 after
    every correction."""
    import torch

    from momwire_amd import workloads
    from momwire_amd.bspline import BSplineSolver

    kw = workloads.dipole_array(*wl["args"][:3], seed=wl["args"][3])
    s64 = BSplineSolver(**kw)
    smx = BSplineSolver(**kw, precision="mixed", refine_steps=args.refine_steps)
    mesh = s64._build()
    nb, ns = mesh.n_basis, mesh.n_seg
    npt = s64._port_count()
    print(f"[bench] {args.workload}: {nb} unknowns, {npt} ports; warmup {args.warmup}",
          file=sys.stderr, flush=True)
    res = {}
    for name, solver, fn in (("mixed", smx, smx._solve_device_mixed),
                             ("fp64", s64, s64._solve_device)):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        timers = {}
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn(timers=timers)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        stage = stage_means(timers)
        res[name] = dict(rate=args.steps / dt, ms=1e3 * dt / args.steps, stage_ms=stage)
        print(f"[bench] {name}: {1e3 * dt / args.steps:.1f} ms per solve over {args.steps} "
              f"steps {stage}", file=sys.stderr, flush=True)
    # graph replays and the drop-in call, a few each
    n_extra = args.extra_steps
    for name, solver in (("mixed", smx), ("fp64", s64)) if n_extra > 0 else ():
        solver.solve_graph()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n_extra):
            solver.solve_graph()
        torch.cuda.synchronize()
        res[name]["graph_rate"] = n_extra / (time.perf_counter() - t0)
        solver.compute_port_solution()
        t0 = time.perf_counter()
        for _ in range(n_extra):
            solver.compute_port_solution()
        res[name]["api_rate"] = n_extra / (time.perf_counter() - t0)
        print(f"[bench] {name}: graph {res[name]['graph_rate']:.2f}/s, compute_port_solution "
              f"{res[name]['api_rate']:.2f}/s", file=sys.stderr, flush=True)
    # the mixed form with the FP32 far-interaction fill (pairs beyond 3 segment
    # lengths), its fill time and its Y error per correction against FP64
    far = None
    if args.far_steps > 0:
        sfar = BSplineSolver(**kw, precision="mixed", refine_steps=args.refine_steps,
                             far_fill="fp32")
        sfar._solve_device_mixed()
        torch.cuda.synchronize()
        timers = {}
        t0 = time.perf_counter()
        for _ in range(args.far_steps):
            sfar._solve_device_mixed(timers=timers)
        torch.cuda.synchronize()
        far = {"port_solves_per_s": args.far_steps / (time.perf_counter() - t0),
               "stage_ms": stage_means(timers)}
        print(f"[bench] mixed + FP32 far fill: {far}", file=sys.stderr, flush=True)
    per_step, adjoint, final = [], [], None
    if not args.no_parity:
        print("[bench] Y error of the mixed form per correction", file=sys.stderr, flush=True)
        from momwire_amd.bspline import _to_host
        Y64, = _to_host(s64._solve_device()[0])
        hist = []
        smx._solve_device_mixed(history=hist)
        ynorm = np.abs(Y64).max()
        per_step = [float(np.abs(_to_host(h)[0] - Y64).max() / ynorm) for h in hist
                    if not isinstance(h, tuple)]
        adjoint = [float(np.abs(_to_host(h[1])[0] - Y64).max() / ynorm) for h in hist
                   if isinstance(h, tuple)]
        s64._check_info()
        smx._check_info()
        final = adjoint[-1] if adjoint else per_step[-1]
        if far is not None:
            hf = []
            sfar._solve_device_mixed(history=hf)
            far["y_rel_err_vs_fp64_per_step"] = [
                float(np.abs(_to_host(h)[0] - Y64).max() / ynorm) for h in hf
                if not isinstance(h, tuple)]
            far["y_rel_err_adjoint_corrected_per_step"] = [
                float(np.abs(_to_host(h[1])[0] - Y64).max() / ynorm) for h in hf
                if isinstance(h, tuple)]
    fill_s = res["fp64"]["stage_ms"]["fill"] * 1e-3
    # the factorisation each form ran: symmetric (mw_?sysv) when Z = Z^T
    sym32, sym64 = bool(smx._sym_used), bool(s64._sym_used)
    solve_flops = lu_flops(nb, sym32) + 8.0 * nb * nb * npt
    solve_flops64 = lu_flops(nb, sym64) + 8.0 * nb * nb * npt
    sm = res["mixed"]["stage_ms"]
    tf32 = solve_flops / (sm["solve32"] * 1e-3) / 1e12
    tf64 = solve_flops64 / (res["fp64"]["stage_ms"]["solve"] * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": res["mixed"]["rate"],
        "unit": "port-solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res["mixed"]["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128 fill and residual, c64 factor and solves", "data": "synthetic",
        "config": {"workload": wl["desc"], "unknowns": nb, "segments": ns, "ports": npt,
                   "precision": "mixed (FP64 fill, FP32 MFMA factor + all-port solve, one "
                                "FP64 residual, one FP32 correction, second-order readout "
                                "checked on the device)" if args.refine_steps is None else
                                f"mixed (FP64 fill, FP32 MFMA factor + all-port solve, "
                                f"{args.refine_steps} FP64-residual corrections, "
                                f"adjoint-corrected readout)",
                   "refine_steps": args.refine_steps, "launch": "eager, HIP events per stage"},
        "fp64_port_solves_per_s": res["fp64"]["rate"],
        "mixed_over_fp64": res["mixed"]["rate"] / res["fp64"]["rate"],
        "y_rel_err_vs_fp64": final,
        "y_rel_err_vs_fp64_per_step": per_step,
        "y_rel_err_adjoint_corrected_per_step": adjoint,
        "graph_replay_port_solves_per_s": {k: res[k].get("graph_rate") for k in res},
        "pcie_inclusive_port_solves_per_s": {k: res[k].get("api_rate") for k in res},
        "stage_ms": {"mixed": sm, "fp64": res["fp64"]["stage_ms"]},
        "fill_gentries_per_s": nb * nb / fill_s / 1e9,
        "fill_tflops": fill_flops(ns, nb) / fill_s / 1e12,
        "fp64_solve_tflops": tf64,
        "factorisation": {"mixed": "symmetric (mw_csysv)" if sym32 else "LU (mw_cgesv)",
                          "fp64": "symmetric (mw_zsysv)" if sym64 else "LU (mw_zgesv)"},
        "mixed_with_fp32_far_fill": far,
        "mixed_readout_check": getattr(smx, "mixed_readout", None),
    }
    # the mixed form's dominant stage: the FP32 factor + all-port solve
    fac = "(4/3) n^3" if sym32 else "(8/3) n^3"
    out["roofline"] = {"kernel": ("mw_csysv: FP32 symmetric factor + all-port solve (k_gemm_sub_pf "
                                  "K = 512 updates of the block-lower staircase on the FP32 "
                                  "32x32x2 MFMA, panels, back substitution)") if sym32 else
                                 ("mw_cgesv: FP32 factor + all-port solve (k_gemm_sub_pf K = 512 "
                                  "trailing updates on the FP32 32x32x2 MFMA, panels, back "
                                  "substitution)"),
                       "bound": "mfma", "achieved": tf32, "peak": PEAK_FP32_TFLOPS,
                       "unit": "TFLOP/s", "frac": tf32 / PEAK_FP32_TFLOPS, "traffic": None,
                       "algorithmic": f"{fac} + 8 n^2 p = {solve_flops:.3e} flop per solve "
                                      f"(n={nb}, p={npt})"}
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tpath):
        tr = json.load(open(tpath)).get(args.workload)
        if tr:
            out["roofline"]["traffic"] = tr["traffic_bytes_per_solve"]
            out["roofline"]["traffic_source"] = tr["source"]
            out["roofline"]["algorithmic_bytes"] = 2 * 8 * nb * (nb + npt)
    if not args.no_cpu and nb <= 20000:
        print("[bench] CPU baseline leg", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(kw, budget_s=0.0)
    elif not args.no_cpu:
        # one reference solve here is minutes of CPU work ((8/3) n^3 flop in LAPACK)
        out["cpu_baseline"] = None
    print(json.dumps(out))


def ground_fill_flops(n_seg, n_b):
    """Algorithmic flops of the grounded fill's extra pass: image moments at
    the free-space rate (16 x 53 per segment pair) and assembly (726 per
    entry), plus the remainder: 9 node pairs x ~400 flops per segment pair
    (4x4 Lagrange stencil of four complex surfaces, the eqs 143-147
    projection) and its (d+1)^2-moment contraction (9 x 8 x 9 per pair)."""
    return (16 * 53 + 9 * 400 + 648) * n_seg ** 2 + 726 * n_b ** 2


def ground_main(args, wl):
    """Cold port solves over Sommerfeld ground: the learned surrogate fills the
    grid (exact engine on the nodes outside its domain) against the exact
    engine; Y of both against the reference's golden solve."""
    import torch

    from momwire_amd import _ground, workloads
    from momwire_amd.bspline import BSplineSolver
    from momwire_amd.surrogate import SommerfeldSurrogate

    kw = workloads.grounded_array(4, 4, 21, seed=4)
    net = SommerfeldSurrogate.load()
    g = np.load(os.path.join(ROOT, "tests", "golden", f"{wl['golden']}.npz"))

    def cold(solver_kw, steps):
        s = BSplineSolver(**solver_kw)
        s._build()
        ts = []
        for _ in range(steps):
            _ground.clear_cache()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            Y, _W = s._solve_device()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return s, ts, Y.cpu().numpy()

    print(f"[bench] {args.workload}: warmup {args.warmup}", file=sys.stderr, flush=True)
    cold(kw, args.warmup)
    cold(dict(kw, somm_surrogate=net), args.warmup)
    se, t_exact, y_exact = cold(kw, args.steps)
    sl, t_learn, y_learn = cold(dict(kw, somm_surrogate=net), args.steps)
    print(f"[bench] cold solves: exact {np.median(t_exact)*1e3:.2f} ms, learned "
          f"{np.median(t_learn)*1e3:.2f} ms", file=sys.stderr, flush=True)
    m = sl._grid_master
    mesh = se._build()
    nb, ns = mesh.n_basis, mesh.n_seg

    # warm solves (grid cached, as in a sweep over one ground): graph replay
    for _ in range(args.warmup):
        se.solve_graph()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        se.solve_graph()
    torch.cuda.synchronize()
    warm = args.steps / (time.perf_counter() - t0)

    # stage split of warm eager solves: HIP events on the launch stream
    timers = {}
    for _ in range(max(3, args.steps // 2)):
        se._solve_device(timers=timers)
    torch.cuda.synchronize()
    stage_ms = {k: float(np.mean([a.elapsed_time(b) for a, b in v])) for k, v in timers.items()}

    # the grounded fill alone (same-edge table, free-space fill, ground pass),
    # 50 back to back between two HIP events on the launch stream, so the
    # figure is its kernels' time without launch gaps
    from momwire_amd import _lib
    lib = _lib.load()
    D = se._device()
    st = torch.cuda.current_stream().cuda_stream

    fill_only = se._fill_launcher(D, se.k, se.omega, D.Z.data_ptr(), D.ld, st)

    def grounded_fill():
        _lib.check(lib.mw_same_edge_table(D.mw, se.k, st), "same-edge")
        fill_only()

    grounded_fill()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(50):
        grounded_fill()
    ev[1].record()
    torch.cuda.synchronize()
    fill_kernels_ms = ev[0].elapsed_time(ev[1]) / 50

    # the image-only answer: what the remainder is worth
    s0 = BSplineSolver(**kw)
    s0._ground_grid = lambda k, omega: None
    y_img = s0._solve_device()[0].cpu().numpy()
    err = lambda a, b: float(np.abs(a - b).max() / np.abs(b).max())  # noqa: E731

    # the learned form across the whole C5 box: 16 draws, heights stratified
    # over 0.05-1 wavelength
    print("[bench] learned vs exact over 16 draws of the C5 box", file=sys.stderr, flush=True)
    box = []
    for params in workloads.ground_box_draws(16, seed=2026):
        kwb = workloads.grounded_array(4, 4, 21, params=params)
        _ground.clear_cache()
        ye = BSplineSolver(**kwb).compute_port_solution().y
        _ground.clear_cache()
        yl = BSplineSolver(**kwb, somm_surrogate=net).compute_port_solution().y
        sb = BSplineSolver(**kwb)
        sb._ground_grid = lambda k, omega: None
        box.append((params[0], err(yl, ye), err(sb.compute_port_solution().y, ye)))
    box_l = np.array([b[1] for b in box])
    box_0 = np.array([b[2] for b in box])
    fill_s = fill_kernels_ms * 1e-3
    flops = fill_flops(ns, nb) + ground_fill_flops(ns, nb)
    tf = flops / fill_s / 1e12
    out = {
        "metric": METRIC,
        "value": 1.0 / float(np.median(t_learn)),
        "unit": "port-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * float(np.median(t_learn)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl["desc"], "unknowns": nb, "segments": ns,
                   "ports": se._port_count(), "grid_nodes": m.n_nodes,
                   "grid_nodes_engine_fallback": m.engine_nodes},
        "cold_port_solves_per_s_learned": 1.0 / float(np.median(t_learn)),
        "cold_port_solves_per_s_exact_engine": 1.0 / float(np.median(t_exact)),
        "warm_port_solves_per_s": warm,
        "stage_ms_warm": stage_ms,
        "grounded_fill_ms": fill_kernels_ms,
        "y_rel_err_learned_vs_exact_engine": err(y_learn, y_exact),
        "y_rel_err_learned_across_c5_box": {
            "draws": len(box), "height_wl_range": [float(min(b[0] for b in box)),
                                                   float(max(b[0] for b in box))],
            "max": float(box_l.max()), "p99": float(np.percentile(box_l, 99)),
            "median": float(np.median(box_l)),
            "image_only_median": float(np.median(box_0)),
            "min_gain_over_image_only": float((box_0 / box_l).min())},
        "y_rel_err_image_only_vs_exact_engine": err(y_img, y_exact),
        "y_rel_err_exact_vs_reference": err(y_exact, g["y"]),
        "surrogate": getattr(net, "meta", ""),
        "roofline": {"kernel": "grounded fill (k_same_edge + k_fill + k_fill_ground)",
                     "bound": "valu", "achieved": tf, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": tf / PEAK_FP64_TFLOPS, "traffic": None,
                     "algorithmic": f"free fill 16*53*N^2 + 726*n^2 plus ground pass "
                                    f"(16*53 + 9*400 + 648)*N^2 + 726*n^2 = {flops:.3e} flop "
                                    f"(N={ns}, n={nb}) over 50 back-to-back grounded fills "
                                    "(k_same_edge + k_fill + k_fill_ground) between HIP events"},
    }
    if not args.no_cpu:
        print("[bench] CPU baseline leg", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(kw, cold_and_warm=True)
    print(json.dumps(out))


def far_flops(n_dir, n_el, p, ground=False):
    """Algorithmic flops of the far-field moments: one complex multiply-add
    (8 flops) per direction, element and product column (3 components x p
    port columns), twice over a ground (direct and image sums)."""
    return 8.0 * n_dir * n_el * 3 * p * (2 if ground else 1)


def readout_main(args, wl):
    """Every port column of the array8x8 solve read out as a far-field
    pattern: element moments (mw_element_moments) then the far-field moments
    (mw_far_moments) with the solution resident in HBM."""
    import torch

    from momwire_amd import readout, workloads
    from momwire_amd.bspline import BSplineSolver
    from oracle import readout_oracle as ro

    kw = workloads.dipole_array(*wl["args"][:3], seed=wl["args"][3])
    s = BSplineSolver(**kw)
    _Y, X = s._solve_device()
    torch.cuda.synchronize()
    rm = s.readout_mesh(1)
    mid = rm.device()["mid"]
    theta = np.radians(np.arange(0.0, 181.0))
    phi = np.radians(np.arange(0.0, 360.0))
    n_dir, p = len(theta) * len(phi), int(X.shape[1])
    print(f"[bench] {args.workload}: {rm.n_el} elements, {p} columns, {n_dir} directions; "
          f"warmup {args.warmup}", file=sys.stderr, flush=True)

    def step():
        mom, _delta = readout.element_moments(rm, X)
        return mom, readout.far_moments(mid, mom, s.k, theta, phi)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # the far-field kernel alone, HIP events on the launch stream
    mom, _d = readout.element_moments(rm, X)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(args.steps):
        mt, mp = readout.far_moments(mid, mom, s.k, theta, phi)
    ev[1].record()
    torch.cuda.synchronize()
    far_ms = ev[0].elapsed_time(ev[1]) / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _m, (mt, mp) = step()
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / args.steps
    # parity of a sample of directions and columns against the oracle
    cols = [0, 27, 63]
    rows = slice(0, 181, 30)
    mom_h = mom.cpu().numpy()
    mt_h, mp_h = mt.cpu().numpy(), mp.cpu().numpy()
    errs = []
    for j in cols:
        rt, rp = ro.far_moments(rm.mid, mom_h[..., j], s.k, theta[rows], phi)
        sc = max(np.abs(rt).max(), np.abs(rp).max())
        errs.append(max(np.abs(mt_h[rows, :, j] - rt).max(), np.abs(mp_h[rows, :, j] - rp).max())
                    / sc)
    flops = far_flops(n_dir, rm.n_el, p)
    tf = flops / (far_ms * 1e-3) / 1e12
    out = {
        "metric": "far-field pattern readout: port columns/s",
        "value": p / (ms * 1e-3),
        "unit": "port patterns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl["desc"], "elements": rm.n_el, "columns": p,
                   "directions": n_dir},
        "far_kernel_ms": far_ms,
        "pattern_rel_err_vs_oracle": float(max(errs)),
        "roofline": {"kernel": "k_far_moments", "bound": "mfma", "achieved": tf,
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": tf / PEAK_FP64_TFLOPS,
                     "traffic": None,
                     "algorithmic": f"8 flop x {n_dir} directions x {rm.n_el} elements x 3 x "
                                    f"{p} columns = {flops:.3e} flop per launch"},
    }
    if not args.no_cpu:
        # the reference's numpy far field (restated in oracle/readout_oracle.py), one
        # column over a band of theta rows, scaled to the whole grid and all columns
        print("[bench] CPU baseline leg", file=sys.stderr, flush=True)
        band = slice(0, 10)
        t0 = time.perf_counter()
        reps = 0
        while True:
            ro.far_moments(rm.mid, mom_h[..., reps % p], s.k, theta[band], phi)
            reps += 1
            if time.perf_counter() - t0 > 15.0:
                break
        per_band = (time.perf_counter() - t0) / reps
        per_pattern = per_band * len(theta) / 10
        cores = len(os.sched_getaffinity(0))
        if os.environ.get("OMP_NUM_THREADS"):
            cores = min(cores, int(os.environ["OMP_NUM_THREADS"]))
        out["cpu_baseline"] = {"value": 1.0 / per_pattern, "unit": "port patterns/s",
                               "cores": cores, "kind": "port",
                               "sample": f"{reps} single-column far fields over 10 theta rows "
                                         f"x 360 phi (numpy: the reference's _far_moments "
                                         f"restated; exp on one thread, the phase-moment "
                                         f"product on threaded BLAS), scaled to 181 rows"}
    print(json.dumps(out))


def groundfield_main(args, wl):
    """Every port column of the grounded 4x4 solve read out as the Sommerfeld
    remainder of E on a plane of points (mw_ground_fields), moments and grid
    resident in HBM."""
    import torch

    from momwire_amd import _ground, readout, workloads
    from momwire_amd.bspline import BSplineSolver
    from oracle import sommerfeld_oracle as so

    s = BSplineSolver(**workloads.grounded_array(4, 4, 21, seed=4))
    s.compute_port_solution()
    mid, mom, _n, _d = s.port_element_moments(subdiv=8)
    lam = s.wavelength
    gx = np.linspace(-0.5, 2.0, 64) * lam
    gy = np.linspace(-0.5, 2.0, 64) * lam
    xx, yy = np.meshgrid(gx, gy, indexing="ij")
    zz = (0.1 + (np.arange(64 * 64) % 11) / 10.0) * lam
    pts = np.column_stack([xx.ravel(), yy.ravel(), zz])
    eps_t = _ground.eps_tilde(s.ground_eps, s.omega, s.eps)
    grid = readout.ground_grid_for(pts, mid, eps_t, s.ground_z, s.k, s.omega, s.mu)
    pts_d = torch.from_numpy(pts).cuda()
    mid_d = torch.from_numpy(np.ascontiguousarray(mid)).cuda()
    p, n_src, n_pts = int(mom.shape[2]), int(mid.shape[0]), len(pts)
    print(f"[bench] {args.workload}: {n_pts} points, {n_src} sources, {p} columns; "
          f"warmup {args.warmup}", file=sys.stderr, flush=True)

    def step():
        return readout.reflected_field_at(pts_d, mid_d, mom, eps_t, s.ground_z, s.k, s.omega,
                                          mu=s.mu, grid=grid)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(args.steps):
        out = step()
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / args.steps
    # parity of a sample of points and columns against the oracle on the
    # device grid's own tables
    m = grid[0]
    vals = m.vals.cpu().numpy()
    tabs = [vals[o:o + a * b].reshape(a, b, 4).transpose(2, 0, 1)
            for (a, b), o in zip(m.shapes, m.off)]
    og = so.Grid(m.regions, (m.r1_max, m.r_break, m.r_near, m.r0_fill), tabs, so.K2_REF,
                 so.K2_REF * so.C_LIGHT, so.MU0_EXACT, m.eps_t).scaled(s.k, s.omega, s.mu)
    sel = np.arange(0, n_pts, 97)
    out_h = out.cpu().numpy()
    mom_h = mom.cpu().numpy()
    errs = []
    for c in (0, 7, 15):
        ref = so.reflected_field_at(pts[sel], mid, mom_h[:, :, c], s.ground_z, s.k, og)
        errs.append(np.abs(out_h[sel, :, c] - ref).max() / np.abs(ref).max())
    # interpolation 16 stencil nodes x 4 surfaces x complex FMA (256 flop) per
    # point-source pair, plus 40 flop per column for the combination
    flops = n_pts * n_src * (256 + 40 * p)
    tf = flops / (ms * 1e-3) / 1e12
    res = {
        "metric": "Sommerfeld ground-remainder near field: port field maps/s",
        "value": p / (ms * 1e-3),
        "unit": "port field maps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl["desc"], "points": n_pts, "sources": n_src, "columns": p},
        "field_rel_err_vs_oracle": float(max(errs)),
        # FP64 vector work, priced against the FP64 vector peak (78.6
        # TFLOP/s on MI355X, the same as the matrix peak)
        "roofline": {"kernel": "k_ground_fields", "bound": "valu", "achieved": tf,
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": tf / PEAK_FP64_TFLOPS,
                     "traffic": None,
                     "algorithmic": f"({n_pts} points x {n_src} sources) x (256 + 40 x {p}) "
                                    f"= {flops:.3e} flop per launch"},
    }
    if not args.no_cpu:
        print("[bench] CPU baseline leg", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        reps = 0
        chunk = pts[:64]
        while True:
            so.reflected_field_at(chunk, mid, mom_h[:, :, reps % p], s.ground_z, s.k, og)
            reps += 1
            if time.perf_counter() - t0 > 15.0:
                break
        per_chunk = (time.perf_counter() - t0) / reps
        cores = len(os.sched_getaffinity(0))
        if os.environ.get("OMP_NUM_THREADS"):
            cores = min(cores, int(os.environ["OMP_NUM_THREADS"]))
        res["cpu_baseline"] = {"value": 1.0 / (per_chunk * n_pts / 64), "unit": "port field maps/s",
                               "cores": cores, "kind": "port",
                               "sample": f"{reps} single-column evaluations at 64 points (numpy: "
                                         f"the reference's reflected_field_at restated), "
                                         f"scaled to {n_pts} points"}
    print(json.dumps(res))


def fill_flops(n_seg, n_b):
    """Algorithmic flops of the free-space fill: 16 quadrature points x 53 flops
    per segment pair plus 726 flops per Z entry (the reference's moment and
    assembly loops, _accelerators.cpp:464-537 and 1264-1303)."""
    return 16 * 53 * n_seg ** 2 + 726 * n_b ** 2


def lu_flops(n, symmetric=False):
    """Factorisation flops: (8/3) n^3 for LU; (4/3) n^3 for the symmetric
    factorisation (Z = Z^T, U = D L^T: half the trailing update)."""
    return (4.0 if symmetric else 8.0) / 3.0 * n ** 3


def cpu_baseline(kw, budget_s=20.0, cold_and_warm=False, max_runs=5):
    """The reference CPU path (its compiled accelerator + scipy LAPACK) on this
    host's cores: a bounded sample of whole port solves.  The k-independent
    set-up (geometry, basis, port columns) is built once outside the timed
    solves, as the reference caches it; the fill takes the reference's own
    route (tensor or chunked, by its swept_mem_mb budget).  budget_s <= 0:
    one timed solve.  cold_and_warm: the first solve fills the Sommerfeld
    grid (the reference's engine), the rest reuse it, as the reference's grid
    cache does; the value is the cold rate."""
    from oracle.ref_cpu import RefCpuSolver, load_ref_accel

    acc = load_ref_accel()
    if acc is None:
        raise RuntimeError("oracle/_ref is not built: the CPU baseline is the reference's own "
                           "compiled accelerator (make -C oracle ref)")
    kind = "reference"
    cores = host_cores()
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    where = f"OMP threads={cores}, CPU {cpu_model()}"
    s = RefCpuSolver(acc, **kw)
    s.prepared()
    route = ("chunked (assemble_Z_bspline_windowed)"
             if not s.tensor_fits_budget(len(s.prepared()[0]["seg_l"])) else "tensor")
    times = []
    if cold_and_warm:
        RefCpuSolver._GRIDS.clear()
        t0 = time.perf_counter()
        s.compute_port_solution()
        t_cold = time.perf_counter() - t0
        for _ in range(3):
            t0 = time.perf_counter()
            s.compute_port_solution()
            times.append(time.perf_counter() - t0)
        warm = float(np.median(times))
        return {"value": 1.0 / t_cold, "unit": "port-solves/s", "cores": cores, "kind": kind,
                "cpu_model": cpu_model(),
                "sample": f"1 cold port solve ({t_cold:.2f} s: grid filled by the reference's "
                          f"six-integral engine) and 3 warm ones (median {warm:.3f} s, "
                          f"{1.0 / warm:.2f}/s) of the same workload; {where}"}
    if budget_s > 0:
        s.compute_port_solution()  # warm-up
    t_start = time.perf_counter()
    parts = []
    while True:
        tm = {}
        t0 = time.perf_counter()
        s.compute_port_solution(timings=tm)
        times.append(time.perf_counter() - t0)
        parts.append(tm)
        if time.perf_counter() - t_start > budget_s or len(times) >= max_runs:
            break
    med = float(np.median(times))
    fill = float(np.median([p["fill_s"] for p in parts]))
    return {"value": 1.0 / med, "unit": "port-solves/s", "cores": cores, "kind": kind,
            "cpu_model": cpu_model(),
            "sample": f"{len(times)} full port solve(s) of the same workload (median {med:.2f} s: "
                      f"fill {fill:.2f} s on the reference's {route} route in its C++ "
                      f"accelerator, solve {med - fill:.2f} s in scipy LAPACK); {where}"}


ENTRY_FLOPS = 9 * 16 * 53  # one Z entry: 9 wing pairs x 16 point pairs x 53 flop


def hm_main(args, wl):
    """The block-low-rank port solve: one step = the operator build (near
    blocks, ACA of every far block, dense fallbacks) plus the block-GMRES
    solve of every port column and the readout, device-resident."""
    import torch

    from momwire_amd import _lib, workloads
    from momwire_amd.bspline import BSplineSolver
    from momwire_amd.hmatrix import HMatrixSolver

    kw = workloads.dipole_array(*wl["args"][:3], seed=wl["args"][3])
    s = HMatrixSolver(**kw)
    mesh = s._build()
    nb = mesh.n_basis
    t0 = time.perf_counter()
    part = s.build_partition()
    t_part = time.perf_counter() - t0
    print(f"[bench] {args.workload}: {nb} unknowns, {s._port_count()} ports, partition "
          f"{part.n_far} far + {part.n_near} near blocks in {t_part:.2f} s (host, once per "
          f"geometry)", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        s._solve()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        Y, W = s._solve()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    # stage split (HIP events on the launch stream): operator build, solve
    st = {"build": [], "solve": []}
    for _ in range(3):
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record()
        H = s.build_hmatrix()
        e1.record()
        D = s._dev
        B = torch.zeros((nb, D.ncols), dtype=torch.complex128, device="cuda")
        _lib.check(_lib.load().mw_zscatter_cols(nb, D.ncols, D.col_max, D.col_idx.data_ptr(),
                                                D.col_val.data_ptr(), B.data_ptr(), D.ncols,
                                                torch.cuda.current_stream().cuda_stream), "B")
        X, iters = H.solve(B, s.solve_tol)
        e2.record()
        torch.cuda.synchronize()
        st["build"].append(e0.elapsed_time(e1))
        st["solve"].append(e1.elapsed_time(e2))
    stage = {k: float(np.mean(v)) for k, v in st.items()}
    stats = H.stats()
    y = Y.cpu().numpy()
    # dense FP64 GPU solve of the same structure, for time and the approximation error
    d = BSplineSolver(**kw)
    d.solve_graph()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        Yd, _Wd = d.solve_graph(warm=False)
    torch.cuda.synchronize()
    dense_ms = 1e3 * (time.perf_counter() - t0) / 3
    yd = Yd.cpu().numpy()
    err_dense = float(np.abs(y - yd).max() / np.abs(yd).max())
    err_ref = None
    if wl["golden"]:
        g = np.load(os.path.join(ROOT, "tests", "golden", wl["golden"] + ".npz"))
        err_ref = float(np.abs(y - g["y"]).max() / np.abs(g["y"]).max())
    b = H.blocks
    far = part.blocks[:, 4] == 1
    lr = b["rank"] >= 0
    fl = far & lr
    samples = int((b["rank"][fl].astype(np.int64) * (b["m"][fl] + b["n"][fl])).sum())
    dense_entries = int((b["m"][~lr].astype(np.int64) * b["n"][~lr]).sum())
    entries = samples + dense_entries
    build_flops = entries * ENTRY_FLOPS
    achieved = build_flops / (stage["build"] * 1e-3) / 1e12
    out = {
        "metric": METRIC, "value": 1.0 / dt, "unit": "port-solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * dt,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl["desc"], "unknowns": nb, "ports": int(D.n_ports),
                   "step": "operator build + block GMRES + readout, device-resident"},
        "stage_ms": stage, "gmres_iterations": int(iters), "hmatrix_stats": stats,
        "operator_bytes": 16 * stats["storage"], "dense_z_bytes": 16 * nb * nb,
        "dense_fp64_gpu_ms": dense_ms,
        "y_rel_err_vs_dense_gpu": err_dense,
        "y_rel_err_vs_reference_hmatrix": err_ref,
        "roofline": {"kernel": "the operator build (k_hm_dense + k_hm_aca: Z entries at the "
                               "near blocks and the ACA samples)",
                     "bound": "valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "algorithmic": f"{entries} entries (dense blocks + ACA rows/columns "
                                    f"sum r (m + n)) x {ENTRY_FLOPS} flop"},
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="array32x32", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-parity", action="store_true",
                    help="array32x32: leave out the leg after the timed region that copies Y "
                         "to the host and measures the mixed form's error (the line then "
                         "carries no error figures); for hardware-counter runs")
    ap.add_argument("--far-steps", type=int, default=3,
                    help="array32x32: mixed solves timed with the FP32 far-interaction fill "
                         "(0: not measured)")
    ap.add_argument("--extra-steps", type=int, default=5,
                    help="array32x32: solves timed for each of the graph-replay and the "
                         "compute_port_solution() rates (0: not measured)")
    ap.add_argument("--eager", action="store_true",
                    help="launch kernels one by one instead of replaying the captured graph")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB three times (the "
                         "calibration kernel for HBM byte counters) and exit")
    ap.add_argument("--refine-steps", type=int, default=None,
                    help="FP64-residual corrections of the mixed form (array32x32); default: "
                         "one, plus the adjoint-corrected readout")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU layout is described, not built)", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        import torch

        n = args.copy_probe << 20
        a = torch.ones(n, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n, "launches": 3}))
        return

    wl = WORKLOADS[args.workload]
    if args.workload == "sweep128":
        return sweep_main(args, wl)
    if args.workload == "ground4x4":
        return ground_main(args, wl)
    if args.workload == "readout8x8":
        return readout_main(args, wl)
    if args.workload == "groundfield4x4":
        return groundfield_main(args, wl)
    if args.workload.startswith("hm"):
        return hm_main(args, wl)
    if wl["golden"] is None:
        return mixed_main(args, wl)

    import torch

    from momwire_amd import workloads
    from momwire_amd.bspline import BSplineSolver

    kw = workloads.dipole_array(*wl["args"][:3], seed=wl["args"][3])
    s = BSplineSolver(**kw)
    mesh = s._build()
    nb, ns = mesh.n_basis, mesh.n_seg
    n_ports = s._port_count()
    print(f"[bench] {args.workload}: {nb} unknowns, {ns} segments, {n_ports} ports; "
          f"warmup {args.warmup}", file=sys.stderr, flush=True)

    step = s._solve_device if args.eager else s.solve_graph
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(f"[bench] timed {args.steps} steps in {dt:.3f} s", file=sys.stderr, flush=True)

    # stage split: HIP events on the launch stream around each stage of eager
    # solves (a graph replay is one launch and cannot be split)
    timers = {}
    for _ in range(max(3, args.steps // 4)):
        s._solve_device(timers=timers)
    torch.cuda.synchronize()
    stage_ms = {k: float(np.mean([a.elapsed_time(b) for a, b in v])) for k, v in timers.items()}
    ms = 1e3 * dt / args.steps
    fill_s = stage_ms["fill"] * 1e-3
    solve_s = stage_ms["solve"] * 1e-3

    # parity of the timed path against the reference's golden Y
    Y, W = s._solve_device()
    y = Y.cpu().numpy()
    g = np.load(os.path.join(ROOT, "tests", "golden", f"{wl['golden']}.npz"))
    y_err = float(np.abs(y - g["y"]).max() / np.abs(g["y"]).max())
    s._check_info()

    # PCIe-inclusive rate: the drop-in compute_port_solution(), Y and coeffs
    # copied to the host every solve (it replays its captured graph from the
    # second solve at a wavenumber on)
    s.compute_port_solution()
    s.compute_port_solution()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    n_pcie = max(3, args.steps // 4)
    for _ in range(n_pcie):
        s.compute_port_solution()
    pcie_rate = n_pcie / (time.perf_counter() - t1)

    ncols = n_ports + len(s._port_plan()[1])
    sym = bool(s._sym_used)
    sv_flops = lu_flops(nb, sym) + 8.0 * nb * nb * ncols
    lu_tf = sv_flops / solve_s / 1e12
    fill_tf = fill_flops(ns, nb) / fill_s / 1e12
    dominant = "solve" if solve_s >= fill_s else "fill"
    if dominant == "solve":
        roof = {"kernel": ("factor + all-port solve (mw_zsysv at one level: Z = Z^T factored "
                           "without row exchanges; per 32-column panel, apply with the next "
                           "panel's update, then the next panel's factor in one launch with "
                           "the MFMA trailing update on the lower part and [B]; then back "
                           "substitution two panels per step)") if sym else
                          ("factor + all-port solve (mw_zgesv: per 32-column panel, threshold-"
                           "pivot fallback, apply with the next panel's update, then the next "
                           "panel's factor in one launch with the MFMA trailing update over "
                           "[Z | B]; then back substitution two panels per step)"),
                "bound": "mfma", "achieved": lu_tf, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": lu_tf / PEAK_FP64_TFLOPS, "traffic": None,
                "algorithmic": f"{'(4/3)' if sym else '(8/3)'} n^3 + 8 n^2 p = {sv_flops:.3e} "
                               f"flop per solve (n={nb}, p={ncols})"}
    else:
        roof = {"kernel": "fill (k_same_edge + k_fill)", "bound": "valu", "achieved": fill_tf,
                "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": fill_tf / PEAK_FP64_TFLOPS,
                "traffic": None,
                "algorithmic": f"16*53*N_seg^2 + 726*n_b^2 = {fill_flops(ns, nb):.3e} flop"}

    # HBM bytes per solve of the dominant kernels, measured by rocprofv3 PMC
    # runs (FETCH_SIZE, WRITE_SIZE) calibrated on a plain copy
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tpath) and dominant == "solve":
        tr = json.load(open(tpath)).get(args.workload)
        if tr:
            roof["traffic"] = tr["traffic_bytes_per_solve"]
            roof["traffic_source"] = (f"{tr['source']}: FETCH_SIZE x2.0 + WRITE_SIZE x1.0 "
                                      "(copy-kernel calibration), per solve")
            roof["algorithmic_bytes"] = 2 * 16 * nb * (nb + ncols)

    out = {
        "metric": METRIC,
        "value": args.steps / dt,
        "unit": "port-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl["desc"], "unknowns": nb, "segments": ns, "ports": n_ports,
                   "precision": "complex128 fill, LU and solve",
                   "launch": "eager" if args.eager else "hip graph replay"},
        "fill_gentries_per_s": nb * nb / fill_s / 1e9,
        "stage_ms": stage_ms,
        "fill_tflops": fill_tf,
        "lu_tflops": lu_tf,
        "y_rel_err_vs_reference": y_err,
        "pcie_inclusive_port_solves_per_s": pcie_rate,
        "roofline": roof,
    }
    if not args.no_cpu:
        print("[bench] CPU baseline leg", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(kw)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""Benchmark of the hot path on one MI355X.

Headline: zone-cycles/s of the FP64 hydro step (HLLC fluxes, conservative
update with floors, boundaries, CFL dt) on the 256^3 stellar-explosion state
(Lane-Emden n=3 star, lognormal k^-2 perturbations, 10|W| thermal deposit),
G = 0 so the step is the hydro sweeps alone.  Synthetic initial data, FP64
arithmetic.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-poisson] [--no-cpu]

Prints one JSON line.  The roofline block prices the fused step kernel at
80 algorithmic bytes per zone (read U 5x8 B, write U 5x8 B) against 8 TB/s;
its duration is the HIP-event average of that kernel on the library's stream.
The cpu_baseline block times the reference C solver (built from its own
sources, one core) on a bounded sample of the same workload.  Unless
--no-poisson, a "poisson" block reports 256^3 Poisson solves to a 1e-8
relative residual: plain FP64 SOR vs FP32 defect correction (inner solves by
red-black SOR sweeps, or by multigrid V-cycles of red-black sweeps) with and
without the learned guess, from a warm and from a cold start, and a
"selfgravity_step" block times whole steps with self-gravity.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "zone-cycles/s @256^3; Poisson ms to 1e-8 resid: plain SOR vs NN-guess+FP32"
# HBM bytes per launch of the step kernel from rocprofv3 FETCH_SIZE / WRITE_SIZE
# runs scaled by a copy-kernel calibration (written next to the summary)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hydro_step_256_traffic.json")


def measured_traffic(n):
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("cells") != n ** 3:
        return None
    return t


HBM_PEAK_GBS = 8000.0
BYTES_PER_ZONE = 80.0


def cpu_baseline(U, n, budget_s=20.0, gpu=None):
    """Reference C solver (or the oracle restatement when the reference build
    is absent) on one pinned core: hydro steps of the same 256^3 state, and,
    with the reference build, its red-black SOR on the same density (two
    capped solves: ms per iteration and the fixed cost of a solve).  `gpu`
    (the GPU's iteration counts) turns those into CPU estimates of the
    Poisson solve and of a warm self-gravity step."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as OL

    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except (AttributeError, OSError):
        pass
    g = OL.Grid(n, n, n, G=0.0, gamma=5.0 / 3.0, cfl=0.4, bc=(2,) * 6)
    kind = "reference" if OL.ref_lib() is not None else "port"
    s = OL.RefSolver(g) if kind == "reference" else OL.OracleSolver(g)
    s.set_state(U)
    t0 = time.perf_counter()
    s.step()  # warm-up (first touch of the arrays)
    warm = time.perf_counter() - t0
    steps = max(1, min(3, int(budget_s / max(warm, 1e-3)) - 1))
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step()
    el = time.perf_counter() - t0
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    del s
    out = {"value": n ** 3 * steps / el, "unit": "zone-cycles/s", "cores": 1, "kind": kind,
           "sample": f"{steps} hydro steps (G=0) of the same {n}^3 state after 1 warm-up step, "
                     f"single core of {cpu}",
           "ms_per_step": 1e3 * el / steps}
    lib = OL.ref_lib()
    if lib is not None:
        rho = np.ascontiguousarray(U[..., 0]).reshape(-1)
        h = lib.refg_create(n, n, n, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0, 1.0)
        times = {}
        for k in (2, 6):
            lib.refg_set_phi(h, OL.dptr(np.zeros(n ** 3)))
            lib.refg_set_params(h, 1.5, k, 1e-300)  # the cap ends the solve
            t0 = time.perf_counter()
            lib.refg_solve(h, OL.dptr(rho))
            times[k] = time.perf_counter() - t0
        lib.refg_destroy(h)
        per_it = (times[6] - times[2]) / 4 * 1e3
        fixed = max(0.0, times[2] * 1e3 - 2 * per_it)
        sor = {"ms_per_iteration": per_it, "ms_fixed_per_solve": fixed,
               "sample": f"reference gravity_solve (l = 2 moments, boundary, red-black SOR, "
                         f"acceleration) of the {n}^3 explosion density capped at 2 and at 6 "
                         f"iterations, same core"}
        if gpu:
            if gpu.get("plain_fp64_sor_iterations"):
                k = gpu["plain_fp64_sor_iterations"]
                sor["poisson_plain_fp64_warm_ms_estimate"] = fixed + k * per_it
                sor["poisson_estimate_note"] = (f"{k} iterations, the GPU's count for the warm "
                                                f"plain FP64 SOR solve to a 1e-8 relative "
                                                f"residual (the same iterates)")
            if gpu.get("selfgravity_warm_iterations"):
                k = gpu["selfgravity_warm_iterations"]
                sor["selfgravity_warm_step_ms_estimate"] = out["ms_per_step"] + fixed + k * per_it
                sor["selfgravity_estimate_note"] = (f"hydro step + fixed solve cost + {k} SOR "
                                                    f"iterations (the GPU's warm count)")
        out["poisson_sor"] = sor
    return out


def network_timing(H, model, n=128, reps=5):
    """Device time of the guess network at 128^3 (where it runs) and its
    MFMA use: flops = 2 * 27 * cin * cout per output voxel and layer; a
    training pass (forward + input and weight gradients) is 3x that."""
    import ctypes as C
    f, fb = C.c_double(0.0), C.c_double(0.0)
    if H.lib().nn_benchmark(model.h, n, reps, C.byref(f), C.byref(fb)) != 0:
        return {"timing_error": True}
    flops = H.lib().nn_forward_flops(n)
    peak = 2.5e15  # MI355X dense BF16 MFMA
    return {"grid": n, "forward_ms": f.value, "forward_backward_ms": fb.value,
            "forward_gflop": flops / 1e9,
            "forward_tflops": flops / (f.value * 1e-3) / 1e12,
            "forward_mfma_frac": flops / (f.value * 1e-3) / peak,
            "training_tflops": 3 * flops / (fb.value * 1e-3) / 1e12}


def sor_roofline(H, gr, rho, n, k1=100, k2=500):
    """Device time per red-black iteration from two solves capped at k1 and
    k2 iterations (the difference cancels the once-per-solve moments,
    boundary and set-up), priced at the algorithmic bytes of an iteration:
    FP64 SOR reads phi and rhs and writes phi (24 B per cell); the FP32
    inner sweeps of defect correction read the correction and the residual
    and write the correction (12 B per cell)."""
    import numpy as np

    out = {}
    for name, mode, bpc in (("sor_fp64", H.POISSON_FP64_SOR, 24.0),
                            ("sor_fp32_inner", H.POISSON_FP32_DC, 12.0)):
        gr.set_mode(mode)
        ms = {}
        for k in (k1, k2):
            gr.set_params(1.5, k, 1e-14)
            gr.set_phi(np.zeros((n, n, n)))
            gr.solve_potential(rho)
            ms[k] = (gr.s.last_solve_ms, gr.diagnostics()[0])
        per_it = (ms[k2][0] - ms[k1][0]) / (ms[k2][1] - ms[k1][1])
        achieved = bpc * n ** 3 / (per_it * 1e-3) / 1e9
        out[name] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "us_per_iteration": 1e3 * per_it,
                     "bytes_per_cell_per_iteration": bpc,
                     "iterations_timed": [ms[k1][1], ms[k2][1]]}
    gr.set_params(1.5, 200000, 1e-8)
    gr.set_mode(H.POISSON_FP64_SOR)
    return out


def rank_share(H, ics, n, ranks=(2, 4, 8), steps=30, k1=100, k2=300):
    """One rank's share of the n^3 explosion run under an x-slab
    decomposition over N GPUs, timed on this GPU at its real shape: the n/N
    planes around the centre (the hot core included) plus one ghost plane on
    each side, which is what a rank holds after its halo exchange.  The hydro
    step (G = 0) counts zone-cycles of the slab's own planes only; FP64 SOR
    and the FP32 inner sweeps are timed per red-black iteration from two
    capped solves on the slab's density.  The exchanges themselves are not
    run (one GPU): a hydro step sends one plane of 40 B cells to each
    neighbour, an SOR half-iteration one colour plane of 8 B cells, priced
    here at one xGMI link of 153 GB/s each way."""
    import numpy as np

    link = 153e9
    U = ics.explosion_n3(n)
    out = {"grid": n, "link_GB_per_s": link / 1e9, "hydro_steps": steps,
           "sor_iterations_timed": [k1, k2], "slabs": {}}
    for N in ranks:
        m = n // N
        lo = n // 2 - m // 2 - 1
        nx = m + 2
        slab = np.ascontiguousarray(U[lo:lo + nx])
        box = (lo / n, (lo + nx) / n, 0.0, 1.0, 0.0, 1.0)
        s = H.Solver(nx, n, n, box=box, G=0.0, gamma=5.0 / 3.0, cfl=0.4,
                     bc=(H.BC_OUTFLOW,) * 6)
        s.set_state(slab)
        s.benchmark(5)
        wall, step_ms, bc_ms = s.benchmark(steps)
        s.close()
        own = float(m) * n * n
        per_rank = own * steps / (wall * 1e-3)
        halo_us = 1e6 * n * n * 40.0 / link
        r = {"slab": [nx, n, n], "own_planes": m, "hydro_ms_per_step": wall / steps,
             "hydro_step_kernel_ms": step_ms, "hydro_boundary_ms": bc_ms,
             "zone_cycles_per_s_per_rank": per_rank,
             "zone_cycles_per_s_if_halo_hidden": N * per_rank,
             "hydro_halo_us_per_step_per_link": halo_us}
        rho = np.ascontiguousarray(slab[..., 0])
        g = H.Gravity(nx, n, n, box=box, G=1.0)
        g.set_lmax(6)
        g.set_criterion(H.SOR_CRIT_REL_RESID)
        for name, mode in (("sor_fp64", H.POISSON_FP64_SOR), ("sor_fp32_inner", H.POISSON_FP32_DC)):
            g.set_mode(mode)
            t = {}
            for k in (k1, k2):
                g.set_params(1.5, k, 1e-14)
                g.set_phi(np.zeros((nx, n, n)))
                g.solve_potential(rho)
                t[k] = (g.s.last_solve_ms, g.diagnostics()[0])
            r[name + "_us_per_iteration"] = 1e3 * (t[k2][0] - t[k1][0]) / (t[k2][1] - t[k1][1])
        g.close()
        r["sor_fp64_halo_us_per_iteration_per_link"] = 2 * 1e6 * (n * n / 2) * 8.0 / link
        out["slabs"][str(N)] = r
        print(f"[bench] rank share of {N}: {nx}x{n}x{n}, hydro {wall / steps:.3f} ms/step, "
              f"SOR {r['sor_fp64_us_per_iteration']:.1f} us/iteration", file=sys.stderr,
              flush=True)
    return out


def poisson_block(H, ics, n):
    """256^3 Poisson solves to ||f - L phi|| / ||f|| <= 1e-8 (interior).

    Warm start (the case of every step after the first): phi_0 = the
    converged potential of the initial state, one hydro step with gravity,
    then each arm starts from phi_0 on the new density: plain FP64 SOR, FP32
    defect correction, and the learned starts + FP32 defect correction: the
    potential guess and phi_0 plus the start correction's correction of it,
    each kept only when its residual beats phi_0's.  Cold start (the first
    step): plain FP64 SOR from zero vs the same arms from zero."""
    import numpy as np
    from hydrocode_astro1_amd import nn

    U = ics.explosion_n3(n)
    rho0 = np.ascontiguousarray(U[..., 0])
    s = H.Solver(n, n, n, G=1.0, gamma=5.0 / 3.0, cfl=0.4, bc=(2,) * 6)
    gr = s.gravity
    gr.set_lmax(6)
    gr.set_params(1.5, 200000, 1e-8)
    gr.set_criterion(H.SOR_CRIT_REL_RESID)
    s.set_state(U)
    t0 = time.perf_counter()
    s.step()  # cold solve of phi_0 (relative residual 1e-8) + one hydro step
    cold = {"iterations": gr.diagnostics()[0], "ms": gr.s.last_solve_ms,
            "wall_s": time.perf_counter() - t0}
    H.lib().gravity_sync_to_host(gr.h)
    phi0 = gr.phi()
    rho1 = s.state()[..., 0].copy()
    out = {"grid": n, "tolerance": 1e-8, "criterion": "relative residual, interior",
           "omega": 1.5, "multipole_lmax": 6}
    model = None
    if os.path.exists(nn.DEFAULT_WEIGHTS):
        model = nn.Model(path=nn.DEFAULT_WEIGHTS)
        gr.set_nn_model(model)
        out["network"] = {"weights": os.path.relpath(nn.DEFAULT_WEIGHTS, ROOT),
                          "parameters": model.num_params()}
        out["network"].update(network_timing(H, model))
    else:
        out["network"] = {"weights": None, "note": "no trained weights; network arms skipped"}
    if model is not None and os.path.exists(nn.CORRECTION_WEIGHTS):
        corr = nn.Model(path=nn.CORRECTION_WEIGHTS)
        gr.set_nn_correction(corr)
        out["network"]["correction_weights"] = os.path.relpath(nn.CORRECTION_WEIGHTS, ROOT)

    def arm(mode, phi_start, rho, inner=H.DC_INNER_SOR):
        gr.set_mode(mode)
        gr.set_dc_inner(inner)
        gr.set_phi(phi_start)
        st = gr.solve_potential(rho)
        it, _ = gr.diagnostics()
        r = {"status": st, "ms": gr.s.last_solve_ms, "rel_residual": gr.s.last_rel_residual}
        r["v_cycles" if inner == H.DC_INNER_MULTIGRID else "iterations"] = it
        if mode == H.POISSON_NN_FP32_DC:
            # start used: 0 the given one, 1 the potential guess, 2 the given
            # one plus the start correction
            r["start_used"] = gr.s.last_guess
            r["guess_used"] = bool(gr.s.last_guess in (1, 2))
            r["guess_rel_residual"] = gr.s.last_guess_rel_residual
            r["corrected_rel_residual"] = gr.s.last_corr_rel_residual
            r["start_rel_residual"] = gr.s.last_warm_rel_residual
        return r, gr.phi()

    warm, phis = {}, {}
    MG = H.DC_INNER_MULTIGRID
    arms = [("plain_fp64_sor", H.POISSON_FP64_SOR, H.DC_INNER_SOR),
            ("fp32_dc", H.POISSON_FP32_DC, H.DC_INNER_SOR),
            ("fp32_dc_multigrid", H.POISSON_FP32_DC, MG)]
    if model is not None:
        arms += [("nn_fp32_dc", H.POISSON_NN_FP32_DC, H.DC_INNER_SOR),
                 ("nn_fp32_dc_multigrid", H.POISSON_NN_FP32_DC, MG)]
    for name, mode, inner in arms:
        warm[name], phis[name] = arm(mode, phi0, rho1, inner)
    coldres, cphis = {"plain_fp64_sor": cold}, {}
    # without the network, so that the network's share is isolated
    coldres["fp32_dc"], cphis["fp32_dc"] = arm(H.POISSON_FP32_DC, np.zeros_like(phi0), rho0)
    coldres["fp32_dc_multigrid"], cphis["fp32_dc_multigrid"] = arm(
        H.POISSON_FP32_DC, np.zeros_like(phi0), rho0, MG)
    if model is not None:
        for name, inner in (("nn_fp32_dc", H.DC_INNER_SOR), ("nn_fp32_dc_multigrid", MG)):
            coldres[name], cphis[name] = arm(H.POISSON_NN_FP32_DC, np.zeros_like(phi0), rho0,
                                             inner)
    gr.set_dc_inner(H.DC_INNER_SOR)
    out["roofline"] = sor_roofline(H, gr, rho0, n)
    # tight FP64 solutions as the accuracy yardstick
    gr.set_mode(H.POISSON_FP64_SOR)
    gr.set_params(1.5, 400000, 1e-11)
    gr.set_phi(phis["plain_fp64_sor"])
    gr.solve_potential(rho1)
    ref1 = gr.phi()
    for name in warm:
        warm[name]["rel_l2_vs_tight_fp64"] = float(
            np.linalg.norm(phis[name] - ref1) / np.linalg.norm(ref1))
    if cphis:
        gr.set_phi(phi0)
        gr.solve_potential(rho0)
        ref0 = gr.phi()
        for name in cphis:
            coldres[name]["rel_l2_vs_tight_fp64"] = float(
                np.linalg.norm(cphis[name] - ref0) / np.linalg.norm(ref0))
    out["warm_start"] = warm
    out["cold_start"] = coldres
    plain = warm["plain_fp64_sor"]["ms"]
    out["speedup_warm_vs_plain"] = {k: plain / v["ms"] for k, v in warm.items()
                                    if k != "plain_fp64_sor" and v["ms"] > 0}
    out["speedup_cold_vs_plain"] = {k: cold["ms"] / v["ms"] for k, v in coldres.items()
                                    if k != "plain_fp64_sor" and v["ms"] > 0}
    gr.set_nn_model(None)
    gr.set_nn_correction(None)
    s.close()
    return out


def selfgravity_block(H, ics, n, steps=10, arms=None):
    """The full step with self-gravity on the same 256^3 state, l = 6
    multipole boundary.  "reference_sor": the reference's Poisson settings
    (red-black SOR, omega 1.5, max |update| < 1e-6, at most 10000
    iterations) with the default moment order, the reference's (terms added
    one after another: its bits).  "reference_sor_tree_moments": the same
    with the moments summed by the pairwise double-double tree (faster,
    within 1e-11 relative L1 of the reference's run after 200 steps).
    "nn_fp32_dc_multigrid": relative
    residual 1e-8, the learned guess (kept only when its residual beats the
    warm start) and FP32 defect correction with multigrid inner solves.
    "fp32_dc_multigrid": the same without a network.
    "nn_correction_fp32_dc_multigrid": the same with the start correction
    only (the warm start plus the network's correction of it, kept only when
    its residual beats the warm start's).  The first step solves cold, the
    next `steps` warm-start from the previous potential.  `arms` picks a
    subset by name."""
    from hydrocode_astro1_amd import nn

    U = ics.explosion_n3(n)

    def run(name, setup, steps=steps):
        s = H.Solver(n, n, n, G=1.0, gamma=5.0 / 3.0, cfl=0.4, bc=(H.BC_OUTFLOW,) * 6)
        gr = s.gravity
        gr.set_lmax(6)
        info = setup(gr)
        s.set_state(U)
        t0 = time.perf_counter()
        s.step()
        cold_ms = 1e3 * (time.perf_counter() - t0)
        cold_it = gr.diagnostics()[0]
        its, ms, starts = [], [], []
        for _ in range(steps):
            t0 = time.perf_counter()
            s.step()  # returns after reading dt back: the step is complete
            ms.append(1e3 * (time.perf_counter() - t0))
            its.append(gr.diagnostics()[0])
            starts.append(gr.s.last_guess)
            if len(its) % 20 == 0:
                print(f"[selfgravity] {name}: {len(its)}/{steps} warm steps", file=sys.stderr,
                      flush=True)
        gr.set_nn_model(None)
        gr.set_nn_correction(None)
        s.close()
        info.update({"first_step": {"ms": cold_ms, "iterations": cold_it},
                     "warm_ms_per_step": sum(ms) / len(ms), "warm_ms": ms,
                     "warm_iterations": its})
        if any(v >= 0 for v in starts):
            # 0 warm start, 1 potential guess, 2 corrected warm start
            info["warm_start_used"] = starts
        return info

    def reference(gr):
        gr.set_params(1.5, 10000, 1e-6)
        return {"sor": {"omega": 1.5, "tol_max_update": 1e-6, "max_iter": 10000},
                "moment_order": "reference (sequential)"}

    def reference_tree(gr):
        info = reference(gr)
        gr.set_moment_order(H.MOMENT_ORDER_TREE)
        info["moment_order"] = "tree"
        return info

    def learned_multigrid(gr):
        gr.set_params(1.5, 1000, 1e-8)
        gr.set_criterion(H.SOR_CRIT_REL_RESID)
        gr.set_mode(H.POISSON_NN_FP32_DC)
        gr.set_dc_inner(H.DC_INNER_MULTIGRID)
        if os.path.exists(nn.DEFAULT_WEIGHTS):
            gr.set_nn_model(nn.Model(path=nn.DEFAULT_WEIGHTS))
        return {"tol_rel_residual": 1e-8, "iterations_are": "V-cycles"}

    def multigrid(gr):
        gr.set_params(1.5, 1000, 1e-8)
        gr.set_criterion(H.SOR_CRIT_REL_RESID)
        gr.set_mode(H.POISSON_FP32_DC)
        gr.set_dc_inner(H.DC_INNER_MULTIGRID)
        return {"tol_rel_residual": 1e-8, "iterations_are": "V-cycles"}

    def corrected_multigrid(gr):
        info = multigrid(gr)
        gr.set_mode(H.POISSON_NN_FP32_DC)
        if os.path.exists(nn.CORRECTION_WEIGHTS):
            gr.set_nn_correction(nn.Model(path=nn.CORRECTION_WEIGHTS))
            info["correction_weights"] = os.path.relpath(nn.CORRECTION_WEIGHTS, ROOT)
        return info

    table = {"reference_sor": (reference, steps),
             "reference_sor_tree_moments": (reference_tree, steps),
             "nn_fp32_dc_multigrid": (learned_multigrid, steps),
             "fp32_dc_multigrid": (multigrid, steps),
             "nn_correction_fp32_dc_multigrid": (corrected_multigrid, steps)}
    # SOR inner sweeps (about a second per solve at 256^3): only when named
    def sor_inner(setup):
        def f(gr):
            info = setup(gr)
            gr.set_dc_inner(H.DC_INNER_SOR)
            gr.set_params(1.5, 400000, 1e-8)
            info["iterations_are"] = "SOR sweeps (omega 1.5)"
            return info
        return f

    named_only = {"fp32_dc": (sor_inner(multigrid), steps),
                  "nn_correction_fp32_dc": (sor_inner(corrected_multigrid), steps)}
    out = {"grid": n, "multipole_lmax": 6, "warm_steps": steps}
    for name, (setup, k) in table.items():
        if arms is None or name in arms:
            out[name] = run(name, setup, k)
    for name, (setup, k) in named_only.items():
        if arms is not None and name in arms:
            out[name] = run(name, setup, k)
    return out


def snapshot_study(H, ics, n, steps=200, every=20, big=(0, 100, 200)):
    """Mixed-precision and learned-guess Poisson solves on snapshots of the
    self-gravitating explosion run: the 256^3 state stepped `steps` times
    with the reference's Poisson settings (l = 6); at every `every`-th step s
    the density rho_s and the potential of step s-1 (zero for s = 0) are
    taken, and each arm solves rho_s to a 1e-8 relative residual from that
    potential: plain FP64 SOR, FP32 defect correction with SOR sweeps, with
    multigrid, the learned guess + multigrid, and the start correction with
    SOR sweeps and with multigrid.  Errors are relative L2
    against the plain FP64 SOR potential of the same solve.  At the steps in
    `big` the same snapshot is prolonged 2x2x2 piecewise constant (mass
    preserving) to 512^3 and solved by the three arms without the guess (the
    network covers grids up to 256^3)."""
    import numpy as np
    from hydrocode_astro1_amd import nn

    def log(msg):
        print(f"[study] {msg}", file=sys.stderr, flush=True)

    model = nn.Model(path=nn.DEFAULT_WEIGHTS) if os.path.exists(nn.DEFAULT_WEIGHTS) else None
    corr = (nn.Model(path=nn.CORRECTION_WEIGHTS) if os.path.exists(nn.CORRECTION_WEIGHTS)
            else None)
    arms = [("plain_fp64_sor", H.POISSON_FP64_SOR, H.DC_INNER_SOR),
            ("fp32_dc", H.POISSON_FP32_DC, H.DC_INNER_SOR),
            ("fp32_dc_multigrid", H.POISSON_FP32_DC, H.DC_INNER_MULTIGRID)]

    def solve_all(m, rho, phi_start, with_nn):
        g = H.Gravity(m, m, m, G=1.0)
        g.set_lmax(6)
        g.set_criterion(H.SOR_CRIT_REL_RESID)
        use = list(arms)
        if with_nn and model is not None:
            g.set_nn_model(model)
            use.append(("nn_fp32_dc_multigrid", H.POISSON_NN_FP32_DC, H.DC_INNER_MULTIGRID))
        if with_nn and corr is not None:
            use += [("nn_correction_fp32_dc", H.POISSON_NN_FP32_DC, H.DC_INNER_SOR),
                    ("nn_correction_fp32_dc_multigrid", H.POISSON_NN_FP32_DC,
                     H.DC_INNER_MULTIGRID)]
        res, ref = {}, None
        for name, mode, inner in use:
            # the potential guess alone, or the start correction alone
            g.set_nn_model(model if name.startswith("nn_fp32") else None)
            g.set_nn_correction(corr if name.startswith("nn_correction") else None)
            g.set_mode(mode)
            g.set_dc_inner(inner)
            g.set_params(1.5, 2000000, 1e-8)
            g.set_phi(phi_start)
            st = g.solve_potential(rho)
            phi = g.phi()
            r = {"status": st, "ms": g.s.last_solve_ms, "iterations": g.diagnostics()[0],
                 "rel_residual": g.s.last_rel_residual}
            if mode == H.POISSON_NN_FP32_DC:
                r["guess_used"] = bool(g.s.last_guess in (1, 2))
                r["start_used"] = g.s.last_guess
            if ref is None:
                ref = phi
            else:
                r["rel_l2_vs_plain_fp64"] = float(np.linalg.norm(phi - ref) / np.linalg.norm(ref))
            res[name] = r
            log(f"{m}^3 {name}: {r['ms']:.1f} ms, {r['iterations']} iterations")
        g.set_nn_model(None)
        g.set_nn_correction(None)
        g.close()
        plain = res["plain_fp64_sor"]["ms"]
        res["speedup_vs_plain"] = {k: plain / v["ms"] for k, v in res.items()
                                   if k != "plain_fp64_sor" and v["ms"] > 0}
        return res

    U = ics.explosion_n3(n)
    s = H.Solver(n, n, n, G=1.0, gamma=5.0 / 3.0, cfl=0.4, bc=(H.BC_OUTFLOW,) * 6)
    s.gravity.set_lmax(6)
    s.gravity.set_params(1.5, 10000, 1e-6)
    s.set_state(U)
    del U
    phi_prev = np.zeros((n, n, n))
    out = {"grid": n, "steps": steps, "every": every, "run_sor": {"omega": 1.5,
           "tol_max_update": 1e-6, "max_iter": 10000}, "snapshots": []}
    for step in range(steps + 1):
        if step % every == 0:
            rho = np.ascontiguousarray(s.state()[..., 0])
            log(f"snapshot at step {step}")
            snap = {"step": step, str(n): solve_all(n, rho, phi_prev, True)}
            if step in big:
                m = 2 * n
                rho2 = np.repeat(np.repeat(np.repeat(rho, 2, 0), 2, 1), 2, 2)
                phi2 = np.repeat(np.repeat(np.repeat(phi_prev, 2, 0), 2, 1), 2, 2)
                snap[str(m)] = solve_all(m, rho2, phi2, False)
                del rho2, phi2
            out["snapshots"].append(snap)
        if step == steps:
            break
        s.step()
        H.lib().gravity_sync_to_host(s.gravity.h)
        phi_prev = s.gravity.phi()
    s.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-poisson", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--poisson-probe", type=int, default=0, metavar="ITERS",
                    help="only time ITERS iterations of plain FP64 SOR and of FP32 defect "
                         "correction on the 256^3 explosion density (twice each), then exit")
    ap.add_argument("--multigrid-probe", type=int, default=0, metavar="REPS",
                    help="only run REPS cold FP32 defect-correction solves with multigrid inner "
                         "solves (relative residual 1e-8) on the 256^3 explosion density, then exit")
    ap.add_argument("--snapshot-no-512", action="store_true",
                    help="the snapshot study without its 512^3 prolongations")
    ap.add_argument("--snapshot-study", type=int, default=0, metavar="STEPS",
                    help="only run the Poisson snapshot study on STEPS steps of the "
                         "self-gravitating explosion (snapshots every 20 steps, 512^3 "
                         "prolongations at steps 0, STEPS/2 and STEPS), then exit")
    ap.add_argument("--selfgravity-probe", type=int, default=0, metavar="STEPS",
                    help="only run the self-gravity step block with STEPS warm steps, then exit")
    ap.add_argument("--selfgravity-arms", default="", metavar="NAMES",
                    help="comma-separated subset of the self-gravity arms (reference_sor, "
                         "reference_sor_tree_moments, nn_fp32_dc_multigrid)")
    ap.add_argument("--nn-probe", type=int, default=0, metavar="REPS",
                    help="only time REPS forward and REPS training passes of the guess network "
                         "on a 128^3 field, then exit")
    ap.add_argument("--rank-share", action="store_true",
                    help="only time one rank's share of the 256^3 run for 2, 4 and 8 x slabs "
                         "(hydro step and SOR iterations at the slab's shape), then exit")
    ap.add_argument("--calibrate", action="store_true",
                    help="only run the device layout-conversion copy (40 B read + 40 B written "
                         "per cell) used to scale HBM byte counters, then exit")
    args = ap.parse_args()
    # the library prints the reference's progress messages on stdout: send
    # them to stderr while the work runs, so that stdout carries one JSON line
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        out = run(args)
    finally:
        import ctypes
        ctypes.CDLL(None).fflush(None)
        sys.stdout.flush()
        os.dup2(saved, 1)
        os.close(saved)
    print(json.dumps(out))


def run(args):
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path "
              f"would shard as x slabs with halo planes and a dt all-reduce; that is not "
              f"built)", file=sys.stderr)
        sys.exit(2)

    import hydrocode_astro1_amd as H
    from hydrocode_astro1_amd import ics

    if H.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    n = args.n
    if args.selfgravity_probe > 0:
        arms = args.selfgravity_arms.split(",") if args.selfgravity_arms else None
        return {"selfgravity_step": selfgravity_block(H, ics, n, args.selfgravity_probe, arms)}
    if args.snapshot_study > 0:
        k = args.snapshot_study
        return {"snapshot_study": snapshot_study(H, ics, n, steps=k, every=max(1, k // 10),
                                                 big=() if args.snapshot_no_512 else
                                                 (0, k // 2, k))}
    if args.nn_probe > 0:
        from hydrocode_astro1_amd import nn
        return {"nn_probe": network_timing(H, nn.Model(path=nn.DEFAULT_WEIGHTS), 128,
                                           args.nn_probe)}
    if args.multigrid_probe > 0:
        import numpy as np
        rho = np.ascontiguousarray(ics.explosion_n3(n)[..., 0])
        g = H.Gravity(n, n, n, G=1.0)
        g.set_lmax(6)
        g.set_criterion(H.SOR_CRIT_REL_RESID)
        g.set_mode(H.POISSON_FP32_DC)
        g.set_dc_inner(H.DC_INNER_MULTIGRID)
        g.set_params(1.5, 1000, 1e-8)
        out = []
        for _ in range(args.multigrid_probe):
            g.set_phi(np.zeros((n, n, n)))
            st = g.solve_potential(rho)
            out.append({"status": st, "v_cycles": g.diagnostics()[0], "ms": g.s.last_solve_ms,
                        "rel_residual": g.s.last_rel_residual})
        return {"multigrid_probe": out, "grid": n}
    if args.poisson_probe > 0:
        import numpy as np
        rho = np.ascontiguousarray(ics.explosion_n3(n)[..., 0])
        g = H.Gravity(n, n, n, G=1.0)
        g.set_lmax(6)
        g.set_criterion(H.SOR_CRIT_REL_RESID)
        out = []
        for mode in (H.POISSON_FP64_SOR, H.POISSON_FP32_DC) * 2:
            g.set_mode(mode)
            g.set_params(1.5, args.poisson_probe, 1e-14)  # the cap ends each solve
            g.set_phi(np.zeros((n, n, n)))
            g.solve_potential(rho)
            it = g.diagnostics()[0]
            out.append({"mode": int(mode), "iterations": it, "ms": g.s.last_solve_ms,
                        "us_per_iteration": 1e3 * g.s.last_solve_ms / max(it, 1)})
        return {"poisson_probe": out, "grid": n}
    if args.rank_share:
        return {"rank_share": rank_share(H, ics, n)}
    if args.calibrate:
        import numpy as np
        s = H.Solver(n, n, n, G=0.0)
        s.set_state(np.ones((n, n, n, 5)))
        for _ in range(3):
            s.state()  # soa_to_aos_kernel: reads and writes 40 B per cell
        return {"calibration": "soa_to_aos_kernel", "cells": n ** 3,
                "bytes_read": 40 * n ** 3, "bytes_written": 40 * n ** 3}
    t0 = time.perf_counter()
    U = ics.explosion_n3(n)
    print(f"[bench] initial state {n}^3 built in {time.perf_counter() - t0:.1f}s", file=sys.stderr,
          flush=True)
    s = H.Solver(n, n, n, G=0.0, gamma=5.0 / 3.0, cfl=0.4, bc=(H.BC_OUTFLOW,) * 6)
    s.set_state(U)
    if args.warmup > 0:
        s.benchmark(args.warmup)
    wall_ms, step_ms, bc_ms = s.benchmark(args.steps)
    zones = float(n) ** 3
    value = zones * args.steps / (wall_ms * 1e-3)
    achieved = BYTES_PER_ZONE * zones / (step_ms * 1e-3) / 1e9
    print(f"[bench] hydro: {wall_ms / args.steps:.3f} ms/step, step kernel {step_ms:.3f} ms, "
          f"boundary kernel {bc_ms:.3f} ms", file=sys.stderr, flush=True)
    s.close()

    res = {
        "metric": METRIC, "value": value, "unit": "zone-cycles/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall_ms / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"stellar explosion (n=3 polytrope, lognormal k^-2 "
                               f"perturbations, 10|W| core deposit) {n}^3, hydro step G=0",
                   "grid": [n, n, n], "gamma": 5.0 / 3.0, "cfl": 0.4, "bc": "outflow"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "hydro_step_kernel", "kernel_ms": step_ms,
                     "boundary_kernel_ms": bc_ms, "bytes_per_zone": BYTES_PER_ZONE},
    }
    t = measured_traffic(n)
    if t is not None:
        rf = res["roofline"]
        rf["traffic"] = t["traffic_bytes"]
        rf["traffic_source"] = "profiles/hydro_step_256.md (rocprofv3 FETCH_SIZE + WRITE_SIZE)"
        if "valu_busy" in t:
            # the step kernel issues FP64 VALU work at this fraction of the
            # wave64 issue rate (one instruction per 4 cycles per SIMD)
            rf["valu_per_zone"] = t["valu_per_cell"]
            rf["valu_busy"] = t["valu_busy"]
    if not args.no_poisson:
        try:
            res["poisson"] = poisson_block(H, ics, n)
        except Exception as e:  # report, never hide the headline
            res["poisson"] = {"error": repr(e)}
        try:
            res["selfgravity_step"] = selfgravity_block(H, ics, n)
        except Exception as e:
            res["selfgravity_step"] = {"error": repr(e)}
    if not args.no_cpu:
        gpu = {}
        try:
            gpu["plain_fp64_sor_iterations"] = res["poisson"]["warm_start"]["plain_fp64_sor"][
                "iterations"]
        except (KeyError, TypeError):
            pass
        try:
            its = res["selfgravity_step"]["reference_sor"]["warm_iterations"]
            gpu["selfgravity_warm_iterations"] = round(sum(its) / len(its))
        except (KeyError, TypeError, ZeroDivisionError):
            pass
        res["cpu_baseline"] = cpu_baseline(U, n, gpu=gpu)
    return res


if __name__ == "__main__":
    main()

"""Throughput of MagLua's spin-dynamics step on one MI355X.

One step = one Heun step (PredictorCorrector.lua single corrector pass plus the
thermal kick) of the 256x256x8 film (periodic in-plane, open z; J=1, K=0.05 along z,
layered FFT dipole g=0.02 at truncation 1000, alpha=0.1, dt=0.002, T=0.5), all
524,288 spins resident in HBM, through the fused kernels of libmsd.so replayed as
a HIP graph; `value` is that plain-FFT throughput.  The `learned` object compares,
on the 512x512x4 film (g=0.1, T=0.2), the plain form with the surrogate dipole
field refreshed exactly every K steps (network trained first, outside the timing).
Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-learned]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X published HBM3E bandwidth



def kernel_bytes(spec, has_cls, comp):
    """Algorithmic HBM bytes per launch: every compulsory FP32 array read or written
    once by the kernel that consumes or produces it; neighbour re-reads are free.
    comp: the spin state and the corrector base carry FP32 low words (12 B per site)."""
    n = spec.n
    nxh = spec.nx // 2 + 1
    spec_bytes = 3 * spec.nz * nxh * spec.ny * 8          # one complex spectrum, 3 components
    tensor = 6 * spec.nz * nxh * spec.ny * 4               # one word per component, layer offsets 0..nz-1
    cls = n if has_cls else 0                              # one exchange-class byte per site
    lo = n * 24 if comp else 0                             # one low-word array read, one written
    return {
        # column pass: read M^ and the tensor, write H^
        "k_cols": 2 * spec_bytes + tensor,
        # predictor: read S (x,y,z,ms) and H^(S); write P, C and M^(P) (+ read S_lo, write C_lo)
        "k_step_pred": n * 16 + cls + spec_bytes + n * 24 + spec_bytes + lo,
        # corrector + thermal kick: read P, C, ms and H^(P); write S and M^(S) (+ read C_lo, write S_lo)
        "k_step_corr": n * 28 + cls + spec_bytes + n * 12 + spec_bytes + lo,
    }


def step_model_bytes(spec, comp):
    """Whole-step compulsory bytes per spin-step (field evaluation E counted twice
    plus predictor and corrector traffic): B = 2E + 96 (+ 48 with low words),
    E = 24 + 48 K/N + 24 K_T/N."""
    KN = (spec.nx // 2 + 1) / spec.nx
    KT = KN * (2 * spec.nz - 1) / spec.nz
    E = 12 + 24 * KN + 24 * KN + 24 * KT + 12
    return 2 * E + 96 + (48 if comp else 0)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(spec, budget_s=15.0, omp_budget_s=8.0):
    """FP64 oracle (this project's CPU restatement of the reference path) on the same
    workload for as many whole steps as fit the budget after one warm-up step: one
    thread (the reference's loops and FFTW plans are serial), then the same build with
    OpenMP over the host threads this process may use."""
    from maglua_amd import workloads as W
    from oracle import oracle as O

    def timed(threads, budget):
        o = W.build(O, spec, threads=threads)
        o["ss"].setSpins(*O.random_unit_spins(spec.seed, spec.n))
        sim = O.Sim(o["ss"], o["ex"], o["ani"], o["dip"], o["temp"], o["llg"])
        sim.heun(1)
        steps, t0 = 0, time.perf_counter()
        while True:
            sim.heun(1)
            steps += 1
            el = time.perf_counter() - t0
            if el > budget or steps >= 50:
                break
        return spec.n * steps / el, steps, el

    v1, s1, e1 = timed(False, budget_s)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    O.Sim.set_threads(threads)
    vt, st, et = timed(True, omp_budget_s)
    return {"value": v1, "unit": "spin-steps/s", "cores": 1, "kind": "port",
            "sample": f"{s1} Heun steps of the full {spec.nx}x{spec.ny}x{spec.nz} film after 1 warm-up step, "
                      f"FP64, single thread, {e1:.1f} s",
            "cpu_model": _cpu_model(),
            "openmp": {"value": vt, "unit": "spin-steps/s", "cores": threads,
                       "sample": f"{st} Heun steps, same build with OpenMP, {et:.1f} s"}}


def measure_traffic(film, schedule, kernel, comp, timeout_s=150):
    """HBM bytes per dispatch of `kernel`, measured in this run: two rocprofv3 counter
    passes (FETCH_SIZE, then WRITE_SIZE; the two do not fit one pass) over a child
    probe of the same workload and schedule (`--traffic-probe`: a few graph-replayed
    steps, then one device copy of the spins into a second system: 16 bytes per site,
    and 12 more of low words with the compensated state).  Each counter's KiB per
    dispatch is scaled by what it reads for those copies of known size.  Returns
    (bytes, detail) or (None, reason)."""
    from maglua_amd import workloads as W
    spec = W.FILM_256_T05 if film == "256x256x8" else W.FILM_512
    ncopy = 2 if comp else 1
    copy_kib = spec.n * (16 + (12 if comp else 0)) / 1024.0
    import csv
    import glob
    import shutil
    import signal
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3") or ("/opt/rocm/bin/rocprofv3" if os.path.exists("/opt/rocm/bin/rocprofv3") else None)
    if prof is None:
        return None, "rocprofv3 not found"
    env = dict(os.environ, TMPDIR="/tmp",
               MSD_ROW_TILE=f"{schedule['rows_per_tile']}:{schedule['threads']}:{schedule['split']}")
    got = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
            cmd = [prof, "--pmc", ctr, "--output-format", "csv", "-d", d, "-o", ctr, "--", sys.executable,
                   os.path.abspath(__file__), "--traffic-probe", "--film", film] + ([] if comp else ["--plain"])
            p = subprocess.Popen(cmd, cwd=ROOT, env=env, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                                 start_new_session=True)
            try:
                _, err = p.communicate(timeout=timeout_s)
            except subprocess.TimeoutExpired:
                os.killpg(p.pid, signal.SIGKILL)
                p.communicate()
                return None, f"{ctr} pass timed out"
            files = glob.glob(os.path.join(d, "**", f"{ctr}_counter_collection.csv"), recursive=True)
            if not files:
                return None, f"{ctr} pass gave no counters (rc {p.returncode}): " + err.decode(errors="replace")[-300:]
            kern, copies = [], []
            for fn in files:
                for r in csv.DictReader(open(fn)):
                    if r["Counter_Name"] != ctr:
                        continue
                    name = r["Kernel_Name"]
                    # the split y transform's column kernel is k_cols_split<...>
                    if (f"{kernel}<" in name or f"{kernel}_split<" in name) and "batch" not in name:
                        kern.append(float(r["Counter_Value"]))
                    elif "copyBuffer" in name:
                        copies.append(float(r["Counter_Value"]))
            if not kern or len(copies) < ncopy or sum(copies[-ncopy:]) <= 0:
                return None, f"{ctr}: {len(kern)} {kernel} dispatches, {len(copies)} copies in the counters"
            factor = copy_kib / sum(copies[-ncopy:])
            got[ctr] = (sum(kern) / len(kern), factor)
    fetch = got["FETCH_SIZE"][0] * got["FETCH_SIZE"][1] * 1024
    write = got["WRITE_SIZE"][0] * got["WRITE_SIZE"][1] * 1024
    return fetch + write, {"fetch_kib": round(got["FETCH_SIZE"][0], 1), "fetch_factor": round(got["FETCH_SIZE"][1], 4),
                           "write_kib": round(got["WRITE_SIZE"][0], 1), "write_factor": round(got["WRITE_SIZE"][1], 4)}


def traffic_probe(args):
    """The child run of measure_traffic: the workload in the schedule MSD_ROW_TILE names,
    a few graph-replayed steps, then the calibration copies."""
    import maglua_amd as M
    from maglua_amd import workloads as W
    spec = W.FILM_256_T05 if args.film == "256x256x8" else W.FILM_512
    g = W.build(M, spec)
    g["ss"].randomize(spec.seed)
    heun = M.Heun(g["ss"], g["ex"], g["ani"], g["dip"], g["temp"], g["llg"], compensated=not args.plain)
    other = M.SpinSystem(spec.nx, spec.ny, spec.nz)
    heun.step(6)
    M.synchronize()
    g["ss"].copySpinsTo(other)
    M.synchronize()


BF16_DENSE_PEAK = 2.5e15   # MI355X published dense BF16 MFMA FLOP/s
CNN_USEFUL_FLOP = 2 * 75 * (3 * 16 + 16 * 16 + 16 * 3)      # per site per evaluation (52,800)
CNN_EXEC_FLOP = 64 * 2 * 16 * 16 * 32 // 16                  # per site: 64 16x16x32 MFMAs per 16 sites (10 + 45 + 9)


def learned_vs_plain(M, W, train_pairs=20000, states=4, train_K=10, Ks=(10, 8, 6, 5, 4, 3, 2), gate=1e-4, tol=1e-3,
                     gate_steps=100, steps=500):
    """The 512x512x4 film (g=0.1, T=0.2): exact layered-FFT steps against the surrogate
    refreshed exactly every K steps, with the configured residual tolerance (1e-3
    relative: a window whose field check at its end exceeds it is run again exactly).
    Training: train_pairs (m - m_ref, H_dip(m) - H_dip(m_ref)) pairs from exact
    trajectories of `states` seeded random initial states.  For each K: the learned
    trajectory against the exact one (same state and noise) over gate_steps steps, the
    share of windows that fell back, and the learned rate; the reported K is the
    fastest whose trajectory stays within max|dm| <= gate."""
    import numpy as np
    from maglua_amd import surrogate as S

    spec = W.FILM_512
    g = W.build(M, spec)
    g["ss"].randomize(1)
    g["temp"].setSeed(1)
    heun = M.Heun(g["ss"], g["ex"], g["ani"], g["dip"], g["temp"], g["llg"])
    plain_schedule = heun.autotune(steps=20)      # the plain path at its best schedule
    heun.step(50)
    M.synchronize()
    t0 = time.perf_counter()
    heun.step(steps)
    M.synchronize()
    plain_s = (time.perf_counter() - t0) / steps
    plain = spec.n / plain_s
    # the local part of a step alone (exchange, anisotropy, thermal, no dipole): what a
    # learned step costs besides its two network evaluations
    loc = W.build(M, spec, with_dipole=False)
    loc["ss"].randomize(2)
    hl = M.Heun(loc["ss"], loc["ex"], loc["ani"], None, loc["temp"], loc["llg"])
    hl.step(20)
    M.synchronize()
    t0 = time.perf_counter()
    hl.step(steps)
    M.synchronize()
    local_s = (time.perf_counter() - t0) / steps
    del hl, loc

    sur = M.Surrogate(spec.nx, spec.ny, spec.nz)
    per = train_pairs // states
    t0 = time.perf_counter()
    losses = []
    for k in range(states):
        h = W.build(M, spec)
        h["ss"].randomize(100 + k)
        h["temp"].setSeed(100 + k)
        hk = M.Heun(h["ss"], h["ex"], h["ani"], h["dip"], h["temp"], h["llg"])
        hk.step(50)
        losses += S.train_on_trajectory(sur, h["ss"], hk, h["dip"], steps=per, K=train_K, lr=2e-3, seed=1,
                                        resume=k > 0)
        del hk, h
    M.synchronize()
    train_s = time.perf_counter() - t0

    # field error train_K steps after a refresh on a state the training never saw, beside
    # the frozen field and the floor of a linear network with this receptive field: the
    # exact kernel cut to |dx|, |dy| <= 6 (all four layers), setTruncation({1000, 6, 6})
    ref = g["ss"].copy()
    heun.step(train_K)
    err, frozen = sur.fieldError(g["dip"], g["ss"], ref)
    a, b = g["ss"].spins(), ref.spins()
    dm = [np.asarray(a[c]) - np.asarray(b[c]) for c in range(3)]
    sd = M.SpinSystem(spec.nx, spec.ny, spec.nz)
    sd.setSpins(*dm)
    full = M.Dipole(spec.nx, spec.ny, spec.nz)
    full.setStrength(spec.g)
    cut = M.Dipole(spec.nx, spec.ny, spec.nz)
    cut.setStrength(spec.g)
    cut.setTruncation((1000, 6, 6))   # l_settrunc's table form: radial 1000, x and y limits 6
    full.apply(sd)
    hf = np.stack(sd.field("Dipole"))
    cut.apply(sd)
    hc = np.stack(sd.field("Dipole"))
    g["dip"].apply(g["ss"])
    hmax = float(np.abs(np.stack(g["ss"].field("Dipole"))).max())
    floor = float(np.abs(hc - hf).max()) / hmax
    del sd, full, cut, hf, hc

    # one network evaluation, host-timed over back-to-back calls
    sur.setInputSystems(g["ss"], ref)
    sur.forward(download=False)
    M.synchronize()
    t0 = time.perf_counter()
    for _ in range(50):
        sur.forward(download=False)
    M.synchronize()
    cnn_s = (time.perf_counter() - t0) / 50
    del ref

    scan, best = [], None
    for K in Ks:
        a2, b2 = g["ss"].copy(), g["ss"].copy()
        ta2, tb2 = W.build(M, spec)["temp"], W.build(M, spec)["temp"]
        ta2.setSeed(9)
        tb2.setSeed(9)
        ea2 = M.Heun(a2, g["ex"], g["ani"], g["dip"], ta2, g["llg"])
        eb2 = M.Heun(b2, g["ex"], g["ani"], g["dip"], tb2, g["llg"])
        eb2.setSurrogate(sur, K=K, tol=tol)
        ea2.step(gate_steps)
        eb2.step(gate_steps)
        d = float(np.abs(np.stack(a2.spins()[:3]) - np.stack(b2.spins()[:3])).max())
        st = eb2.surrogateStats()
        # the learned rate at this K, fallbacks included
        eb2.step(2 * K)
        M.synchronize()
        t0 = time.perf_counter()
        eb2.step(steps)
        M.synchronize()
        rate = spec.n * steps / (time.perf_counter() - t0)
        st2 = eb2.surrogateStats()
        row = {"K": K, "traj_max_dm": d, "windows": st["windows"], "fallback_windows": st["fallbacks"],
               "max_residual": round(st["max_residual"], 6), "learned_value": round(rate, 1),
               "ratio": round(rate / plain, 4),
               "timed_fallback_share": round((st2["fallbacks"] - st["fallbacks"]) /
                                             max(1, st2["windows"] - st["windows"]), 4)}
        scan.append(row)
        if d <= gate and (best is None or rate > best["learned_value"]):
            best = row
        del a2, b2, ea2, eb2
    gate_ok = best is not None
    if best is None:
        best = scan[0]
    # Amdahl: with a free network, a learned step is the local part alone; with the
    # network at its measured cost it adds two evaluations per step
    return {
        "workload": spec.name() + ", plain layered FFT vs surrogate refreshed every K steps",
        "plain_value": round(plain, 1), "learned_value": best["learned_value"], "ratio": best["ratio"],
        "unit": "spin-steps/s", "K": best["K"], "residual_tol": tol,
        "gate": f"max|dm| <= {gate:g} over {gate_steps} steps against the exact trajectory (same state and "
                f"noise); every window kept has its field error at its end <= {tol:g} (others run again exactly)",
        "gate_met": gate_ok, "K_scan": scan,
        "plain_schedule": plain_schedule,
        "plain_step_us": round(1e6 * plain_s, 2), "local_step_us": round(1e6 * local_s, 2),
        "cnn_eval_us": round(1e6 * cnn_s, 2),
        "amdahl_bound_free_network": round(plain_s / local_s, 3),
        "bound_at_measured_network": round(plain_s / (local_s + 2 * cnn_s), 3),
        "field_error_after_train_K": round(err, 6), "frozen_field_error_after_train_K": round(frozen, 6),
        "truncated_kernel_floor_after_train_K": round(floor, 6),
        "train_pairs": per * states, "train_states": states, "train_K": train_K, "train_s": round(train_s, 2),
        "final_loss": round(float(np.mean(losses[-200:])), 6),
        "cnn_mfma_frac_useful": round(CNN_USEFUL_FLOP * spec.n / cnn_s / BF16_DENSE_PEAK, 4),
        "cnn_mfma_frac_executed": round(CNN_EXEC_FLOP * spec.n / cnn_s / BF16_DENSE_PEAK, 4),
        "network_dtype": "bf16 operands, fp32 accumulate",
    }


def conv3_pass_bytes(dims, open_, comp, site_axes=False):
    """Algorithmic HBM bytes per launch of each LongRange3D pass for one member
    (every compulsory array once per pass; FFT intermediates inside a pass free).
    The working spectrum W holds the nz stored planes (z >= nz of an open transform
    are zero and never stored), fxh = fx/2+1 words per row; the tensor one FP32 word
    per component per k with kz <= fz/2 and ky <= fy/2 stored (z and y parity of the
    lattice sums).
      y forward: read the ny row-pass rows, write fy rows; inverse y the reverse;
      z pass: read and write 3 nz fy fxh words, read 6 (fz/2+1) (fy/2+1) fxh tensor words;
      row passes: the site arrays of the predictor / corrector (as the layered
      kernels, plus a per-site axis and K when the easy axes are random) and the
      rows of W in and out."""
    nx, ny, nz = dims
    fx, fy, fz = (2 * nx, 2 * ny, 2 * nz) if open_ else (nx, ny, nz)
    fxh = fx // 2 + 1
    n = nx * ny * nz
    word = 3 * nz * fxh * 8                 # one y row of W, all stored planes and components
    rows_w = 2 * word * ny                  # row pass: W rows in and out
    lo = n * 24 if comp else 0
    cls = n if open_ else 0                 # exchange class byte per site (faces of an open lattice)
    axes = n * 16 if site_axes else 0
    return {"k3_ypass_fwd": word * (ny + fy), "k3_ypass_inv": word * (ny + fy),
            "k3_zconv": 2 * word * fy + 6 * (fz // 2 + 1) * (fy // 2 + 1) * fxh * 4,
            "k3_rows_pred": n * 16 + n * 24 + lo + cls + axes + rows_w,
            "k3_rows_corr": n * 28 + n * 12 + lo + cls + axes + rows_w}


def pass_roofline(prof, bytes_per_pass, step_ms, step_bytes):
    """roofline sub-object of a leg: every pass's achieved GB/s from its algorithmic
    bytes and event-timed duration, the dominant pass (most time per step), and the
    whole step (all passes' bytes over the driver-timed step)."""
    passes = {}
    for k, ms in prof.items():
        if k in bytes_per_pass and ms > 0:
            gbs = bytes_per_pass[k] / (ms * 1e-3) / 1e9
            passes[k] = {"avg_ms": round(ms, 4), "algorithmic_bytes": int(bytes_per_pass[k]),
                         "achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    dom = max(passes, key=lambda k: passes[k]["avg_ms"])
    step_gbs = step_bytes / (step_ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": dom, "achieved": passes[dom]["achieved"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": passes[dom]["frac"], "passes": passes,
            "step": {"bytes": int(step_bytes), "ms": round(step_ms, 4), "achieved": round(step_gbs, 1),
                     "frac": round(step_gbs / HBM_PEAK_GBS, 4)}}


def ensemble_leg(M, W, nrep=64, steps=50, R=16, tune_steps=10):
    """64 periodic replicas with sizes drawn (seed 0) from {32^3, 64^3, 128x128x16, 96x96x24},
    temperatures 0.1..2.0, each with its own seed, LongRange3D dipole (lattice sum to R)
    g=0.01, J=1, K=0.1: one set of launches per Heun step for the whole batch."""
    import numpy as np
    rng = np.random.default_rng(0)
    sizes = [(32, 32, 32), (64, 64, 64), (128, 128, 16), (96, 96, 24)]
    pick = rng.integers(0, 4, nrep)
    t0 = time.perf_counter()
    members, keep, nsites, temps = [], [], 0, []
    for i in range(nrep):
        dims = sizes[pick[i]]
        T = 0.1 + (2.0 - 0.1) * i / max(1, nrep - 1)
        spec = W.FilmSpec(*dims, J=1.0, K=0.1, g=0.0, pbc=(1, 1, 1), alpha=0.1, dt=0.002, T=T, seed=i)
        g = W.build(M, spec)
        g["ss"].randomize(i)
        g["temp"].setSeed(i)
        lr = M.LongRange3D(*dims)
        lr.setRadius(R)
        lr.setStrength(0.01)
        members.append(M.Heun(g["ss"], g["ex"], g["ani"], lr, g["temp"], g["llg"]))
        keep.append((g, lr))
        nsites += spec.n
        temps.append(T)
    E = M.Ensemble(members)
    # LongRange3D pass widths timed in the ensemble's own launches (states restored)
    widths = E.autotune(steps=tune_steps) if tune_steps > 0 else None
    E.step(2)
    M.synchronize()
    setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    E.step(steps)
    M.synchronize()
    el = time.perf_counter() - t0
    prof = E.profile(5)
    M.synchronize()
    tot = {}
    for i in range(nrep):
        for k, v in conv3_pass_bytes(sizes[pick[i]], False, True).items():
            tot[k] = tot.get(k, 0) + v
    roof = pass_roofline(prof, tot, 1e3 * el / steps,
                         2 * (tot["k3_ypass_fwd"] + tot["k3_zconv"] + tot["k3_ypass_inv"]) + tot["k3_rows_pred"]
                         + tot["k3_rows_corr"])
    # per-replica observables after the timed steps (outside the timing): |M| per
    # site from netMoment and the total field's dot with the spins per site, both
    # FP64 device reductions; the fields come from the operator API
    m_abs, hdot = [], []
    for (g, lr) in keep:
        ss = g["ss"]
        n = ss.nxyz()
        m_abs.append(round(ss.netMoment()[3] / n, 6))
        ss.resetFields()
        g["ex"].apply(ss)
        g["ani"].apply(ss)
        lr.apply(ss)
        ss.sumFields()
        hdot.append(round(ss.fieldDot("Total") / n, 6))
    return {"workload": f"{nrep} periodic replicas, sizes from 32^3/64^3/128x128x16/96x96x24 (seed 0), "
                        f"T 0.1..2.0, LongRange3D lattice sum R={R}, g=0.01, J=1, K=0.1",
            "value": round(nsites * steps / el, 1), "unit": "spin-steps/s", "sites": nsites, "steps": steps,
            "ms_per_step": round(1e3 * el / steps, 4), "setup_s": round(setup, 2),
            "roofline": roof, "pass_widths": widths,
            "temperatures": [round(t, 4) for t in temps], "m_abs_per_site": m_abs,
            "total_field_dot_spin_per_site": hdot}


def films_leg(M, W, nrep=8, steps=100, schedule=None):
    """nrep 256x256x8 films (the film's J, K, g, alpha, dt; temperatures 0.1..0.8, own
    seeds) in one Ensemble: each pass of a Heun step is one launch over all films.
    schedule: the row-tile shape and y-transform schedule the single film was tuned to
    (the members' Dipoles are created with it)."""
    from dataclasses import replace
    saved = os.environ.get("MSD_ROW_TILE")
    if schedule:
        os.environ["MSD_ROW_TILE"] = f"{schedule['rows_per_tile']}:{schedule['threads']}:{schedule['split']}"
    t0 = time.perf_counter()
    members, keep, nsites, temps = [], [], 0, []
    for i in range(nrep):
        T = 0.1 + (0.8 - 0.1) * i / max(1, nrep - 1)
        spec = replace(W.FILM_256_T05, T=T, seed=i, noise_seed=i)
        g = W.build(M, spec)
        g["ss"].randomize(i)
        members.append(M.Heun(g["ss"], g["ex"], g["ani"], g["dip"], g["temp"], g["llg"]))
        keep.append(g)
        nsites += spec.n
        temps.append(round(T, 4))
    if saved is None:
        os.environ.pop("MSD_ROW_TILE", None)
    else:
        os.environ["MSD_ROW_TILE"] = saved
    E = M.Ensemble(members)
    E.step(5)
    M.synchronize()
    setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    E.step(steps)
    M.synchronize()
    el = time.perf_counter() - t0
    return {"workload": f"{nrep} films 256x256x8 (J=1, K=0.05, g=0.02, alpha=0.1, dt=0.002), T 0.1..0.8, one "
                        "Ensemble: one launch per pass for all films",
            "value": round(nsites * steps / el, 1), "unit": "spin-steps/s", "sites": nsites, "steps": steps,
            "ms_per_step": round(1e3 * el / steps, 4), "setup_s": round(setup, 2), "temperatures": temps,
            "schedule": keep[0]["dip"].schedule()}


def cube_learned(M, W, spec, g, ani, lr, heun, plain, train_steps=100, train_K=8, Ks=(8, 4, 2, 1), gate=1e-4,
                 field_gate=1e-3, gate_steps=40, tol=1e-2, steps=40):
    """The 256^3 open cube with the surrogate (zero-padded in x and y) standing in for
    the 3D convolution between exact refreshes every K steps: trained on the cube's own
    trajectory, K is the largest of Ks passing both gates over gate_steps steps (the
    trajectory within max|dm| <= gate of the exact one from the same state and noise,
    the field at every window start within field_gate of the exact one)."""
    import numpy as np
    from maglua_amd import surrogate as S
    sur = M.Surrogate(spec.nx, spec.ny, spec.nz, periodic=False)
    t0 = time.perf_counter()
    losses = S.train_on_trajectory(sur, g["ss"], heun, lr, steps=train_steps, K=train_K, lr=2e-3, seed=1)
    M.synchronize()
    train_s = time.perf_counter() - t0
    ref = g["ss"].copy()
    heun.step(train_K)
    err, frozen = sur.fieldError(lr, g["ss"], ref)
    del ref

    def driver(ss, seed, learned_K=None):
        t = M.Thermal(spec.nx, spec.ny, spec.nz)
        t.setTemperature(spec.T)
        t.setSeed(seed)
        h = M.Heun(ss, g["ex"], ani, lr, t, g["llg"])
        if learned_K:
            h.setSurrogate(sur, K=learned_K, tol=tol)
        return h, t

    scan, K = [], None
    for k in Ks:
        a, b = g["ss"].copy(), g["ss"].copy()
        ea, _ta = driver(a, 9)
        eb, _tb = driver(b, 9, k)
        ea.step(gate_steps)
        eb.step(gate_steps)
        d = float(np.abs(np.stack(a.spins()[:3]) - np.stack(b.spins()[:3])).max())
        fe = float(eb.surrogateStats()["max_residual"])
        scan.append({"K": k, "traj_max_dm": d, "field_err_max": fe})
        del a, b, ea, eb
        if d <= gate and fe <= field_gate:
            K = k
            break
    gate_ok = K is not None
    if K is None:
        K = Ks[0]
    b = g["ss"].copy()
    eb, _tb = driver(b, 7, K)
    eb.step(2 * K)
    M.synchronize()
    t0 = time.perf_counter()
    eb.step(steps)
    M.synchronize()
    learned = spec.n * steps / (time.perf_counter() - t0)
    st = eb.surrogateStats()
    # the learned form's own rate: every window learned (no residual fallback), K = train_K
    eb.setSurrogate(sur, K=train_K, tol=1e30)
    eb.step(2 * train_K)
    M.synchronize()
    t0 = time.perf_counter()
    eb.step(steps)
    M.synchronize()
    raw = spec.n * steps / (time.perf_counter() - t0)
    return {"learned_value": round(learned, 1), "plain_value": round(plain, 1), "ratio": round(learned / plain, 4),
            "learned_steps_value_no_fallback": round(raw, 1), "learned_steps_ratio_no_fallback": round(raw / plain, 4),
            "K": K, "gate_met": gate_ok, "K_scan": scan, "residual_tol": tol,
            "gate": f"max|dm| <= {gate:g} over {gate_steps} steps against the exact trajectory (same state and "
                    f"noise) and dipole-field error <= {field_gate:g} at every window start",
            "timed_K": K, "timed_windows": st["windows"], "timed_fallback_windows": st["exact_windows"],
            "field_error_after_train_K": round(err, 6), "frozen_field_error_after_train_K": round(frozen, 6),
            "train_pairs": train_steps, "train_K": train_K, "train_s": round(train_s, 2),
            "final_loss": round(float(np.mean(losses[-10:])), 6)}


def cube_leg(M, W, n=256, steps=20, learned=True, tune_steps=10):
    """256^3 open cube (zero-padded 512^3 transforms), J=1, K=0.05 along seeded random axes,
    g=0.05, T=0.3, dt=0.002, and a 4-replica batch of 128^3 open cubes in one launch.  On
    the single cube, the surrogate form too (cube_learned)."""
    import numpy as np
    out = {}
    for label, dims, count in (("single", (n, n, n), 1), ("batch4", (n // 2, n // 2, n // 2), 4)):
        t0 = time.perf_counter()
        members, keep, nsites = [], [], 0
        for r in range(count):
            spec = W.FilmSpec(*dims, J=1.0, K=0.05, g=0.0, pbc=(0, 0, 0), alpha=0.1, dt=0.002, T=0.3, seed=r)
            g = W.build(M, spec, with_thermal=True)
            rng = np.random.default_rng(100 + r)
            ax = rng.standard_normal((spec.n, 3))
            ax /= np.linalg.norm(ax, axis=1, keepdims=True)
            ani = M.Anisotropy(*dims)
            ani.addMany(np.arange(spec.n, dtype=np.int32), ax, 0.05)
            g["ss"].randomize(r)
            g["temp"].setSeed(r)
            lr = M.LongRange3D(*dims, open=True)
            lr.setStrength(0.05)
            members.append(M.Heun(g["ss"], g["ex"], ani, lr, g["temp"], g["llg"]))
            keep.append((g, ani, lr))
            nsites += spec.n
        drv = members[0] if count == 1 else M.Ensemble(members)
        widths = drv.autotune(steps=tune_steps) if tune_steps > 0 else None
        drv.step(2)
        M.synchronize()
        setup = time.perf_counter() - t0
        t0 = time.perf_counter()
        drv.step(steps)
        M.synchronize()
        el = time.perf_counter() - t0
        prof = drv.profile(3)
        M.synchronize()
        pb = {k: v * count for k, v in conv3_pass_bytes(dims, True, True, site_axes=True).items()}
        step_b = 2 * (pb["k3_ypass_fwd"] + pb["k3_zconv"] + pb["k3_ypass_inv"]) + pb["k3_rows_pred"] + pb["k3_rows_corr"]
        out[label] = {"lattice": list(dims), "replicas": count, "value": round(nsites * steps / el, 1),
                      "ms_per_step": round(1e3 * el / steps, 3), "setup_s": round(setup, 2),
                      "pass_widths": widths, "roofline": pass_roofline(prof, pb, 1e3 * el / steps, step_b)}
        if count == 1 and learned:
            g, ani, lr = keep[0]
            out[label]["learned"] = cube_learned(M, W, spec, g, ani, lr, drv, nsites * steps / el)
        del drv, members, keep
    out["workload"] = "open cubes, LongRange3D zero-padded to twice the size, random easy axes, g=0.05, T=0.3"
    out["unit"] = "spin-steps/s"
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--profile-steps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-plain comparison on the 512x512x4 film")
    ap.add_argument("--train-pairs", type=int, default=20000, help="surrogate training pairs for the learned leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the batched-replica and 256^3 open-cube legs")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--tune-steps", type=int, default=100,
                    help="steps the autotuner times per candidate schedule before the run (0: default schedule)")
    ap.add_argument("--plain", action="store_true", help="plain FP32 spin state instead of the compensated pair")
    ap.add_argument("--film", choices=("256x256x8", "512x512x4"), default="256x256x8",
                    help="the plain workload: the 256x256x8 film (T=0.5) or the 512x512x4 film (g=0.1, T=0.2)")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-traffic", action="store_true", help="skip the two rocprofv3 counter passes behind roofline.traffic")
    ap.add_argument("--calibrate", action="store_true",
                    help="after timing, copy the spin system once on the device (two copy kernels of 8 MiB and "
                         "6 MiB) so a counter run can scale FETCH_SIZE / WRITE_SIZE by a copy of known size")
    args = ap.parse_args()
    if args.traffic_probe:
        traffic_probe(args)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "replicas would shard with no communication, see the design notes", file=sys.stderr)
        sys.exit(2)

    import maglua_amd as M
    from maglua_amd import workloads as W

    spec = W.FILM_256_T05 if args.film == "256x256x8" else W.FILM_512
    t_setup = time.perf_counter()
    g = W.build(M, spec)
    g["ss"].randomize(spec.seed)
    heun = M.Heun(g["ss"], g["ex"], g["ani"], g["dip"], g["temp"], g["llg"], graph=not args.no_graph,
                  compensated=not args.plain)
    # the autotuner picks the row-tile shape and the y-transform schedule (state restored)
    schedule = heun.autotune(steps=args.tune_steps) if args.tune_steps > 0 else g["dip"].schedule()
    heun.step(args.warmup)
    M.synchronize()
    print(f"# setup + {args.warmup} warm-up steps: {time.perf_counter() - t_setup:.2f} s", file=sys.stderr, flush=True)

    M.synchronize()
    t0 = time.perf_counter()
    heun.step(args.steps)
    M.synchronize()
    el = time.perf_counter() - t0
    value = spec.n * args.steps / el
    ms_step = 1e3 * el / args.steps
    print(f"# timed {args.steps} steps: {el:.3f} s", file=sys.stderr, flush=True)

    # kernel durations: each launch of one step repeated back to back between HIP
    # events on the library stream (the stream every kernel runs on)
    prof = heun.profile(args.profile_steps)
    M.synchronize()
    if args.calibrate:
        g["ss"].copy()
        M.synchronize()
    kb = kernel_bytes(spec, has_cls=True, comp=not args.plain)
    per_step = {k: v * (2 if k == "k_cols" else 1) for k, v in prof.items()}
    dom = max(per_step, key=per_step.get)
    achieved = kb[dom] / (prof[dom] * 1e-3) / 1e9
    traffic, tdetail = (None, "not measured (--no-traffic)")
    if not args.no_traffic:
        print("# roofline.traffic: two rocprofv3 counter passes over a probe run...", file=sys.stderr, flush=True)
        traffic, tdetail = measure_traffic(args.film, schedule, dom, not args.plain)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None if traffic is None else round(traffic),
            "traffic_counters": tdetail,
            "kernel": dom, "kernel_avg_ms": round(prof[dom], 5), "algorithmic_bytes": kb[dom],
            "step_model_bytes_per_spin": round(step_model_bytes(spec, not args.plain), 2),
            "step_model_GBs": round(step_model_bytes(spec, not args.plain) * value / 1e9, 1),
            "kernels_avg_ms": {k: round(v, 5) for k, v in prof.items()},
            "kernels": {k: {"avg_ms": round(v, 5), "algorithmic_bytes": kb[k], "launches_per_step": 2 if k == "k_cols" else 1,
                            "frac": round(kb[k] / (v * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)} for k, v in prof.items() if k in kb}}

    learned = None
    if not args.no_learned:
        print("# learned vs plain (512x512x4, training the surrogate first)...", file=sys.stderr, flush=True)
        learned = learned_vs_plain(M, W, train_pairs=args.train_pairs)

    ensemble = cube = None
    if not args.no_extra:
        print("# batched replicas and open cubes...", file=sys.stderr, flush=True)
        tune3 = min(args.tune_steps, 10)     # 3D pass widths: ten steps per candidate pair
        ensemble = ensemble_leg(M, W, tune_steps=tune3)
        ensemble["films"] = films_leg(M, W, schedule=schedule)
        cube = cube_leg(M, W, tune_steps=tune3)

    cpu = None
    if not args.no_cpu:
        print("# CPU baseline (FP64 oracle, 1 thread)...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(spec)

    out = {
        "metric": "spin-steps/sec on 1 MI355X, learned vs plain FFT dipolar, at max|\u0394m|\u22641e-4",
        "value": round(value, 1),
        "unit": "spin-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded sphere-uniform spins, counter-based thermal noise)",
        "config": {"workload": spec.name() + ", layered FFT dipole, Heun, plain form",
                   "lattice": [spec.nx, spec.ny, spec.nz], "spins": spec.n, "graph": not args.no_graph,
                   "spin_state": "fp32" if args.plain else "fp32 hi+lo (compensated)",
                   "schedule": {k: schedule[k] for k in ("rows_per_tile", "threads", "split")}},
        "roofline": roof,
        "cpu_baseline": cpu,
        "learned": learned,
        "ensemble": ensemble,
        "cube": cube,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

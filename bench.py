#!/usr/bin/env python3
"""Benchmark of the MI355X lattice-Boltzmann hot path.

Workload (default): plain FP64 D3Q27 thermal BGK (f + g), single level,
256^3 channel -- velocity inlet u = 0.05 at x-lo, zero-gradient outflow at
x-hi, periodic y/z, EB sphere r = 24 at (64,128,128) with halfway
bounce-back, tau = 0.55, 1000 steps.  Inputs are resident in HBM when the
timed region starts.

Prints ONE JSON line: MLUPS (cells x steps / s / 1e6), the HBM roofline of
the step's kernels (algorithmic bytes = read + write of every population
once), and the CPU oracle timed on this host's cores on a bounded sample of
the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "D3Q27 MLUPS & % HBM roofline, FP64 vs FP32-shifted at fixed error"
# the FP64-arithmetic collide of the shifted FP32 storage, run beside the
# storage's default FP32-arithmetic one (the autotuner then chooses among
# the block shapes of that arithmetic)
F32_FP64_KERNEL = "k_qpass_y2+k_collide_y2"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_info() -> tuple[int, str]:
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    return os.cpu_count() or 1, model


def cpu_baseline(cfg: dict, budget_s: float, threads: int) -> dict:
    """The CPU oracle (restatement of the reference step, OpenMP over z-slabs
    of each box) timed on this host: 1 warm-up step, then as many steps as fit
    in the budget (at least 1)."""
    from oracle import pyoracle

    pyoracle.set_threads(threads)
    t0 = time.perf_counter()
    O = pyoracle.build_oracle_level(cfg)
    t_init = time.perf_counter() - t0
    O.evolve(1)  # warm-up
    t0 = time.perf_counter()
    O.evolve(1)
    t1 = time.perf_counter() - t0
    n_more = max(0, min(19, int(budget_s / max(t1, 1e-9)) - 1))
    t0 = time.perf_counter()
    if n_more:
        O.evolve(n_more)
    t_rest = time.perf_counter() - t0
    steps = 1 + n_more
    secs = t1 + t_rest
    cells = cfg["ncell"][0] * cfg["ncell"][1] * cfg["ncell"][2]
    O.close()
    nproc, model = cpu_info()
    return {
        "value": cells * steps / secs / 1e6,
        "unit": "MLUPS",
        "cores": threads,
        "kind": "port",
        "sample": (f"{cfg['name']} ({len(cfg.get('boxes', [])) or 'max_box ' + str(cfg.get('max_box'))}), "
                   f"1 warm-up + {steps} timed steps of the CPU oracle; init {t_init:.1f}s; "
                   f"host {model} nproc={nproc}"),
    }


def copy_ceiling(torch, nbytes: int = 4 << 30, iters: int = 10) -> tuple[float, dict]:
    """Read + write bandwidth of the library's hand-written streaming copy
    kernel (mrb_stream_copy) on `nbytes` buffers, best over a few grid
    sizes: the measured HBM ceiling of this box.  Returns (GB/s, detail
    with every grid size's rate and torch's own copy for comparison)."""
    from marbles_amd import _lib
    from marbles_amd._lib import check

    lib = _lib.load()
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def rate(fn) -> float:
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        s.record(stream)
        for _ in range(iters):
            fn()
        e.record(stream)
        torch.cuda.synchronize()
        return 2.0 * nbytes * iters / (s.elapsed_time(e) * 1e-3) / 1e9

    detail = {}
    for blocks in (512, 1024, 2048, 4096):
        detail[f"blocks_{blocks}"] = round(rate(lambda: check(lib.mrb_stream_copy(b.data_ptr(), a.data_ptr(), nbytes,
                                                                                   blocks, sh))), 1)
    if not torch.equal(a, b):
        raise RuntimeError("mrb_stream_copy: copy differs from its source")
    detail["torch_copy"] = round(rate(lambda: b.copy_(a)), 1)
    del a, b
    best = max(v for k, v in detail.items() if k.startswith("blocks_"))
    return best, detail


def measured_traffic(storage: str, kernels) -> float | None:
    """HBM bytes per step of the given kernels from the committed rocprofv3
    summaries (calibrated FETCH_SIZE + WRITE_SIZE per dispatch, the `json`
    blocks of profiles/*.md), or None when no summary covers them."""
    import glob
    import re

    data = {}
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        try:
            text = open(path).read()
        except OSError:
            continue
        for block in re.findall(r"```json\n(.*?)\n```", text, re.S):
            try:
                data.update(json.loads(block).get(storage, {}).get("kernels", {}))
            except (ValueError, AttributeError):
                continue
    elem = {"f64": "double", "f32s": "float", "f16s": "_Float16"}[storage]
    total = 0.0
    for k in kernels:
        rec = data.get(f"{k}<{elem}>")
        if rec is None:
            return None
        total += rec["read_bytes"] + rec["write_bytes"]
    return total


def velocity_density(lvl, cfg) -> tuple:
    """(u, v, w, rho) of the level's current state, FP64 on the host."""
    from marbles_amd.stencil import velocity_density as vd

    return vd(lvl.valid("f").reshape(27, -1))


def time_level(lvl, steps: int, warmup: int, profile_steps: int, sh, torch, autotune: bool = False) -> dict:
    """Warm-up steps, then (optionally) the autotuner on the warmed-up state,
    then the timed steps and the per-kernel profile steps."""
    lvl.advance(warmup, stream=sh)
    tuned = lvl.autotune(reps=10, stream=sh) if autotune else {}
    if tuned:
        log(f"autotuned ({lvl.hot_kernel_name()}): " + ", ".join(f"{k}={v:.3f}" for k, v in tuned.items() if v > 0))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lvl.advance(steps, stream=sh)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    prof = lvl.advance_timed(profile_steps, stream=sh)
    return {"secs": secs, "per_kernel_ms": {k: v / profile_steps for k, v in prof.items()}}


def run_amr(args, torch) -> int:
    """Two-level workload: seeded uneven boxes (edges 8-64), fine level over
    20 % of the volume, thermal f+g FP64; MLUPS counts N_coarse + 2 N_fine
    updates per coarse step."""
    from marbles_amd import Hierarchy, configs

    n = args.n
    cfg = configs.amr_two_level(ncell=(n, n, n), nbox=(n ** 3) // 4096)
    t0 = time.perf_counter()
    H = Hierarchy(cfg)
    log(f"hierarchy built in {time.perf_counter() - t0:.1f}s: {len(cfg['coarse_boxes'])} coarse boxes, "
        f"{len(cfg['fine_boxes'])} fine boxes, fine fraction {cfg['fine_fraction']:.4f}")
    sh = torch.cuda.current_stream().cuda_stream
    ups = H.updates_per_step()
    H.advance(args.warmup, stream=sh, graph=args.graph)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    H.advance(args.steps, stream=sh, graph=args.graph)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    mlups = ups * args.steps / secs / 1e6
    log(f"timed {args.steps} coarse steps: {secs:.4f}s -> {mlups:.1f} MLUPS")
    H.close()
    # algorithmic bytes: every population of every update read once and
    # written once (864 B), as for the single-level step
    bpu = configs.bytes_per_update("f64", 2)
    achieved = bpu * ups * args.steps / secs / 1e9
    traffic = amr_traffic(n)
    print(json.dumps({
        "metric": "two-level D3Q27 MLUPS (N_coarse + 2 N_fine updates per coarse step)",
        "value": round(mlups, 2), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(secs / args.steps * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{cfg['name']}: {len(cfg['coarse_boxes'])} coarse boxes (edges 8-64), "
                               f"{len(cfg['fine_boxes'])} fine boxes over {cfg['fine_fraction']:.4f} of the "
                               f"volume, thermal f+g, T0={cfg['T0']}, tau={cfg['tau']}, Pr=0.71",
                   "updates_per_step": ups, "graph": bool(args.graph)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_update": bpu, "note": "whole coarse step (all kernels), wall clock"},
    }), flush=True)
    return 0


def amr_traffic(n: int) -> float | None:
    """HBM bytes per coarse step of the two-level workload from the committed
    rocprofv3 summary (calibrated FETCH_SIZE + WRITE_SIZE summed over the
    step's kernels), or None when the summary does not cover this size."""
    import re

    path = os.path.join(ROOT, "profiles", "amr_two_level_256.md")
    try:
        block = re.search(r"```json\n(.*?)\n```", open(path).read(), re.S).group(1)
        rec = json.loads(block)["amr_f64"]
    except (OSError, AttributeError, KeyError, ValueError):
        return None
    return rec["traffic_per_coarse_step"] if rec.get("n") == n else None


VANDV = {"cd": 3.096, "cl": 0.8396, "st": 0.2986}  # Docs/sphinx/VandV.rst:14


def run_cylinder(args, torch) -> int:
    """The reference's flow-past-a-cylinder validation: a cylinder deck run
    from its initial state, force coefficients and Strouhal number over the
    last 30 % of the run, beside the published values."""
    from marbles_amd import validation

    res, t, forces = validation.run_cylinder(args.deck, steps=args.steps, sample_every=25, reynolds=args.reynolds,
                                             log=log)
    out_dir = os.path.join(ROOT, "gpurun_out")
    if os.path.isdir(out_dir):
        import numpy as np

        np.savetxt(os.path.join(out_dir, f"forces_{res['cfg']}.txt"), np.column_stack([t, forces]),
                   header="step fx fy fz")
    log(json.dumps(res))
    rel = {}
    if res["strouhal"] is not None:
        rel = {"cd_max": res["cd_max"] / VANDV["cd"] - 1.0, "cl_max": res["cl_max"] / VANDV["cl"] - 1.0,
               "strouhal": res["strouhal"] / VANDV["st"] - 1.0}
    print(json.dumps({
        "metric": "EB cylinder drag / lift coefficients and Strouhal number",
        "value": res["cd_max"], "unit": "C_d (max over the window)", "n_gpus": 1, "steps": args.steps,
        "warmup": 0, "ms_per_step": round(res["seconds"] / args.steps * 1e3, 5), "higher_is_better": None,
        "vs_baseline": None, "dtype": "f64", "data": "deck",
        "config": {"workload": res["cfg"], "deck": args.deck, "kernel": res["kernel"],
                   "reynolds_umax": res["reynolds_umax"], "reynolds_umean": res["reynolds_umean"],
                   "nu": res["nu"]},
        "coefficients": {k: res[k] for k in ("cd_mean", "cd_max", "cd_min", "cl_max", "cl_min", "cl_amplitude",
                                             "strouhal", "periods_in_window", "window_steps")},
        "published": VANDV, "relative_to_published": rel, "mlups": round(res["mlups"], 1),
        "mass_change": res["mass_change"], "finite": res["finite"],
    }), flush=True)
    return 0 if res["finite"] else 1


def run_adjoint(args, torch) -> int:
    """Gradient of the flow objective (outlet-plane mean u_x + 1e-3 KE) with
    respect to n_s, tau and u0 on an n^3 box with an n_s field of seeded
    smoothed blobs (15 % mean), inflow 0.04: forward time, gradient time
    (recorded forward sweep + reverse sweep; mean of three calls after one
    untimed call) with ceil(sqrt(N)) checkpoints and with the memory-aware
    interval (value), and central differences (h = 1e-6) on 32 seeded
    cells, tau and 4 initial-velocity components."""
    import numpy as np

    from marbles_amd.adjoint import IsoSolver

    n = 128 if args.n == 256 else args.n
    nsteps = args.steps
    shape = (n, n, n)
    rng = np.random.default_rng(5)
    z, y, x = np.meshgrid(*[np.arange(n)] * 3, indexing="ij")
    s = np.zeros(shape)
    for _ in range(24):
        c = rng.uniform(0, n, 3)
        s += np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / (2 * (n / 16) ** 2))
    s = np.clip(s, 0, 1)
    s = np.clip(s * 0.15 / s.mean(), 0, 1)
    ns = torch.tensor(s, device="cuda")
    u0 = torch.zeros((3,) + shape, dtype=torch.float64, device="cuda")
    u0[0] = 0.04
    tau = 0.6
    S = IsoSolver((n, n, n), T0=1.0 / 3.0, u_in=0.04)
    reps = 3
    S.forward(ns, tau, u0, nsteps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        J = S.forward(ns, tau, u0, nsteps)
    torch.cuda.synchronize()
    t_fwd = (time.perf_counter() - t0) / reps

    def timed(every):
        """Gradient time (mean of three calls after one untimed call, which
        allocates the recording buffers), the recorded forward sweep and the
        reverse sweep separately, and the interval used."""
        S.gradient(ns, tau, u0, nsteps, every=every)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            out = S.gradient(ns, tau, u0, nsteps, every=every)
        torch.cuda.synchronize()
        t_grad = (time.perf_counter() - t0) / reps
        t_rec = t_rev = 0.0
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, token = S.record(ns, tau, u0, nsteps, every=every)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            S.reverse(token)
            torch.cuda.synchronize()
            t_rec += (t1 - t0) / reps
            t_rev += (time.perf_counter() - t1) / reps
        return out, {"every": S.record_interval(), "gradient_s": t_grad, "ratio": t_grad / t_fwd,
                     "recorded_forward_s": t_rec, "reverse_s": t_rev, "reverse_over_forward": t_rev / t_fwd}

    # sqrt(N) checkpoints (fewest buffers), then the memory-aware interval
    # (every state kept when they fit in HBM: no recomputation)
    (_, _, dns_sq, _), sq = timed(0)
    (J2, dtau, dns, du0), kept = timed(-1)
    identical = bool(torch.equal(dns, dns_sq))
    t_grad, t_rec, t_rev = kept["gradient_s"], kept["recorded_forward_s"], kept["reverse_s"]
    h = 1e-6
    checks = [(dtau, (S.forward(ns, tau + h, u0, nsteps) - S.forward(ns, tau - h, u0, nsteps)) / (2 * h))]
    flat = ns.view(-1)
    for c in rng.choice(ns.numel(), 32, replace=False):
        keep = flat[int(c)].item()
        flat[int(c)] = keep + h
        jp = S.forward(ns, tau, u0, nsteps)
        flat[int(c)] = keep - h
        jm = S.forward(ns, tau, u0, nsteps)
        flat[int(c)] = keep
        checks.append((dns.view(-1)[int(c)].item(), (jp - jm) / (2 * h)))
    uf = u0.view(-1)
    for c in rng.choice(u0.numel(), 4, replace=False):
        keep = uf[int(c)].item()
        uf[int(c)] = keep + h
        jp = S.forward(ns, tau, u0, nsteps)
        uf[int(c)] = keep - h
        jm = S.forward(ns, tau, u0, nsteps)
        uf[int(c)] = keep
        checks.append((du0.view(-1)[int(c)].item(), (jp - jm) / (2 * h)))
    scale = max(abs(g) for g, _ in checks)
    err = max(abs(g - d) for g, d in checks) / scale
    S.close()
    log(f"J={J:.12g} forward {t_fwd:.3f}s; checkpoints every {sq['every']}: gradient {sq['gradient_s']:.3f}s "
        f"({sq['ratio']:.2f}x); every {kept['every']}: gradient {t_grad:.3f}s ({t_grad / t_fwd:.2f}x; recorded "
        f"forward {t_rec / t_fwd:.2f}x + reverse {t_rev / t_fwd:.2f}x); same gradient: {identical}; "
        f"FD max rel err {err:.2e}")
    print(json.dumps({
        "metric": "adjoint gradient cost / forward cost (isothermal D3Q27, partial bounce-back; value = "
                  "recorded forward + reverse sweep with ceil(sqrt(N)) checkpoints and recomputation; "
                  "every state kept: memory_aware_ratio)",
        "value": round(sq["ratio"], 3), "unit": "x", "memory_aware_ratio": round(t_grad / t_fwd, 3),
        "n_gpus": 1, "steps": nsteps, "warmup": 0,
        "ms_per_step": round(t_fwd / nsteps * 1e3, 4), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{n}^3 box, n_s blobs (15 % mean), inflow 0.04, outflow, tau {tau}, "
                               f"{nsteps} steps, checkpoints every {sq['every']} steps (memory-aware "
                               f"interval: every {kept['every']})",
                   "J": J, "forward_s": t_fwd, "gradient_s": t_grad,
                   "recorded_forward_s": t_rec, "reverse_s": t_rev,
                   "reverse_over_forward": t_rev / t_fwd,
                   "sqrt_checkpoints": sq, "memory_aware": kept, "gradients_identical": identical,
                   "mlups_forward": n ** 3 * nsteps / t_fwd / 1e6,
                   "fd_max_rel_error": err, "fd_points": len(checks)},
    }), flush=True)
    return 0


def run_precision(args, torch) -> int:
    """Precision study: periodic n^3 box (default 512^3) with a seeded
    random-phase divergence-free velocity field (E(k) ~ k^-5/3, k <= 32,
    max |u| = 0.04), T0 = 0.2, run for the given steps with FP64,
    shifted-FP32 (FP32 arithmetic, the default, and FP64 arithmetic) and
    scaled-FP16 population storage (FP32 in its collide); velocity,
    density and kinetic energy of each compared with the
    FP64 run on the device (relative L2 and max errors).  tau = 0.55: at
    tau = 0.52 the thermal f+g model diverges on this field at 512^3
    between steps 200 and 400 (finite at 200, NaN at 400); tau = 0.55 stays
    finite through 2000 steps."""
    from marbles_amd import configs
    from marbles_amd.level import Level

    n = args.n if "--n" in sys.argv else 512
    nsteps = args.steps
    tau = 0.55
    cfg = configs.spectral_turbulence(n, tau=tau)
    t0 = time.perf_counter()
    fields = configs.compact_fields(cfg)
    log(f"spectral field {n}^3 built in {time.perf_counter() - t0:.1f}s")
    sh = torch.cuda.current_stream().cuda_stream
    N = n ** 3
    ref = None
    legs = {}
    chunk = 100
    for leg_name in ("f64", "f32s", "f32s_fp64", "f16s"):
        st = leg_name.split("_")[0]
        t0 = time.perf_counter()
        lvl = Level(cfg["ncell"], cfg["periodic"], cfg["prob_lo"], cfg["prob_hi"], layout="compact", storage=st)
        lvl.set_phys(cfg["nu"], cfg.get("alpha"))
        lvl.set_bc(cfg["bc_lo"], cfg["bc_hi"], cfg.get("velocity_bc"))
        if st != "f64":
            lvl.set_shift_state(*configs.reference_state(cfg))
        lvl.flags_all_regular()
        lvl.init_fields(*fields, adiabatic_exponent=cfg["ic"]["adiabatic_exponent"])
        torch.cuda.synchronize()
        log(f"{leg_name}: level initialised in {time.perf_counter() - t0:.1f}s")
        if leg_name == "f32s_fp64":  # the FP64-arithmetic collide of the same storage
            lvl.set_hot_kernel(F32_FP64_KERNEL)
        # the fastest step variant at this size among those of the level's
        # arithmetic, timed on the live state (which it leaves unchanged)
        lvl.autotune(reps=3, stream=sh)
        log(f"{leg_name}: kernel {lvl.hot_kernel_name()}")
        lvl.advance(args.warmup, stream=sh)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        done = 0
        while done < nsteps:
            k = min(chunk, nsteps - done)
            lvl.advance(k, stream=sh)
            torch.cuda.synchronize()
            done += k
            log(f"{leg_name}: {done}/{nsteps} steps, {time.perf_counter() - t0:.1f}s")
        secs = time.perf_counter() - t0
        mac = torch.empty(5 * N, dtype=torch.float64, device="cuda")
        lvl.macro_fields(mac, stream=sh)
        cons = lvl.conservation(stream=sh)
        sat = lvl.storage_saturations()
        kernel = lvl.hot_kernel_name()
        lvl.close()
        rho, u = mac[:N], mac[N:4 * N].view(3, N)
        ke = float(0.5 * (rho * (u * u).sum(0)).sum())
        leg = {"mlups": round(N * nsteps / secs / 1e6, 1), "seconds": round(secs, 2), "kernel": kernel,
               "kinetic_energy": ke, "mass": float(cons[0]), "energy_2rhoE": float(cons[4])}
        if st == "f64":
            ref = (rho, u, ke, cons)
        else:
            r0, u0, ke0, c0 = ref
            du = u - u0
            leg.update({
                "velocity_rel_l2": float(du.norm() / u0.norm()),
                "velocity_rel_max": float(du.abs().max() / u0.abs().max()),
                "density_rel_l2": float((rho - r0).norm() / r0.norm()),
                "density_rel_max": float((rho - r0).abs().max() / r0.abs().max()),
                "kinetic_energy_rel": abs(ke - ke0) / ke0,
                "mass_rel": abs(float(cons[0]) - float(c0[0])) / abs(float(c0[0])),
                "saturated_values": sat,
            })
        legs[leg_name] = leg
        log(f"{leg_name}: " + ", ".join(f"{k}={v:.3g}" if isinstance(v, float) else f"{k}={v}"
                                        for k, v in leg.items()))
        del mac
        torch.cuda.empty_cache()
    print(json.dumps({
        "metric": "D3Q27 MLUPS per population storage with errors against FP64 (precision study)",
        "value": legs["f32s"]["mlups"], "unit": "MLUPS", "n_gpus": 1, "steps": nsteps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * legs["f32s"]["seconds"] / nsteps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"spectral_{n}: periodic {n}^3, random-phase solenoidal field E(k)~k^-5/3 "
                               f"(k<=32), max|u|=0.04, tau {tau}, T0 0.2, thermal D3Q27 f+g; value = "
                               f"shifted-FP32 MLUPS",
                   "storages": legs},
    }), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--storage", default="f64", choices=["f64", "f32s", "f16s"])
    ap.add_argument("--n", type=int, default=256, help="domain edge (256 = the named workload)")
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--kernel", default=None,
                    help="channel: step variant to run (a name from Level.hot_kernel_variants(); default: the "
                         "level's own, the fastest for its storage at 256^3)")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU-oracle timing")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-f32", action="store_true", help="skip the shifted-FP32 and scaled-FP16 comparison runs")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep each level's default kernel instead of timing every variant on the live state")
    ap.add_argument("--workload", default="channel",
                    choices=["channel", "amr", "adjoint", "precision", "cylinder"],
                    help="channel: the single-level 256^3 channel (default); amr: the two-level workload; "
                         "adjoint: gradient cost and finite-difference check at 128^3; precision: FP64, "
                         "shifted FP32 and scaled FP16 storage on a 512^3 spectral field, errors vs FP64; "
                         "cylinder: force coefficients of an EB-cylinder deck against the published values")
    ap.add_argument("--deck", default="turek_2d2", choices=["turek_2d2", "single_cylinder"],
                    help="cylinder: the deck to run")
    ap.add_argument("--reynolds", type=float, default=None,
                    help="cylinder: Re on the mean inflow velocity (sets nu; default: the deck's nu)")
    ap.add_argument("--graph", action="store_true", help="amr: replay captured hipGraphs")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the benchmark runs on one GPU only (--gpus 1); sharding the level across GPUs is "
              "described, not implemented", file=sys.stderr)
        return 2

    import torch

    from marbles_amd import configs

    if not torch.cuda.is_available():
        log("bench.py: no GPU visible")
        return 1
    if args.workload == "precision":
        if "--steps" not in sys.argv:
            args.steps = 2000
        if "--warmup" not in sys.argv:
            args.warmup = 0
        return run_precision(args, torch)
    if args.workload == "cylinder":
        if "--steps" not in sys.argv:
            args.steps = 200000
        return run_cylinder(args, torch)
    if args.workload == "adjoint":
        if "--steps" not in sys.argv:
            args.steps = 200
        return run_adjoint(args, torch)
    if args.workload == "amr":
        if "--steps" not in sys.argv:
            args.steps = 10
        if "--warmup" not in sys.argv:
            args.warmup = 2
        return run_amr(args, torch)
    n = args.n
    scale = n / 256.0
    cfg = configs.channel_sphere((n, n, n), radius=24.0 * scale, center=(64.0 * scale, 128.0 * scale,
                                                                      128.0 * scale), tau=0.55, max_box=64)
    t0 = time.perf_counter()
    lvl = configs.build_level(cfg, layout="compact", storage=args.storage)
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    if args.kernel:
        lvl.set_hot_kernel(args.kernel)
    tune = not args.kernel and not args.no_autotune
    log(f"level built in {time.perf_counter() - t0:.2f}s; kernel: {lvl.hot_kernel_name()}"
        + (" (autotuned after the warm-up steps)" if tune else ""))

    cells = n ** 3
    run = time_level(lvl, args.steps, args.warmup, args.profile_steps, sh, torch, autotune=tune)
    secs = run["secs"]
    mlups = cells * args.steps / secs / 1e6
    log(f"timed {args.steps} steps: {secs:.4f}s -> {mlups:.1f} MLUPS")
    # per-kernel durations from HIP events on the launch stream
    per_kernel_ms = run["per_kernel_ms"]
    step_kernel_ms = sum(per_kernel_ms.values())
    bpu = configs.bytes_per_update(args.storage, 2)
    achieved = bpu * cells / (step_kernel_ms * 1e-3) / 1e9
    log("per-kernel ms/step: " + ", ".join(f"{k}={v:.4f}" for k, v in per_kernel_ms.items()))
    cons = lvl.conservation(stream=sh)
    log(f"conservation (mass, mom x/y/z, 2rhoE): {cons.tolist()}")

    # the same run with shifted FP32 and with scaled FP16 populations,
    # compared at equal step count
    legs = {"f32s": None, "f32s_fp64": None, "f16s": None}
    if args.storage == "f64" and not args.no_f32:
        import numpy as np

        (u64, rho64) = velocity_density(lvl, cfg)
        lvl.close()
        fluid = rho64 > 0
        umax = max(float(np.abs(c[fluid]).max()) for c in u64)
        for leg_name in legs:
            st = leg_name.split("_")[0]
            lr = configs.build_level(cfg, layout="compact", storage=st)
            if leg_name == "f32s_fp64":
                lr.set_hot_kernel(F32_FP64_KERNEL)
            rr = time_level(lr, args.steps, args.warmup, args.profile_steps, sh, torch, autotune=not args.no_autotune)
            log(f"{leg_name}: kernel {lr.hot_kernel_name()}")
            (ur, rhor) = velocity_density(lr, cfg)
            sat = lr.storage_saturations()
            lr.close()
            du = max(float(np.abs(a[fluid] - b[fluid]).max()) for a, b in zip(ur, u64)) / umax
            dr = float(np.abs(rhor[fluid] - rho64[fluid]).max() / np.abs(rho64[fluid]).max())
            kms = sum(rr["per_kernel_ms"].values())
            mr = cells * args.steps / rr["secs"] / 1e6
            br = configs.bytes_per_update(st, 2)
            legs[leg_name] = {
                "value": round(mr, 2),
                "unit": "MLUPS",
                "speedup_vs_f64": round(mr / mlups, 3),
                "max_rel_velocity_error": du,
                "max_rel_density_error": dr,
                "compared_after_steps": args.warmup + args.steps + args.profile_steps,
                "arithmetic": "f32" if any(("_h32" in k or "_f32" in k) for k in rr["per_kernel_ms"]) else "f64",
                "bytes_per_update": br,
                "achieved_gbs": round(br * cells / (kms * 1e-3) / 1e9, 1),
                "traffic": measured_traffic(st, rr["per_kernel_ms"]) if n == 256 else None,
                "kernel_ms_per_step": {k: round(v, 5) for k, v in rr["per_kernel_ms"].items()},
            }
            if st == "f16s":
                legs[leg_name]["saturated_values"] = sat
            log(f"{leg_name}: {mr:.1f} MLUPS ({mr / mlups:.2f}x), max rel velocity error {du:.2e}, "
                f"density {dr:.2e}")
    else:
        lvl.close()
    copy_gbs, copy_detail = copy_ceiling(torch)
    log(f"streaming-copy ceiling: {copy_gbs:.1f} GB/s ({copy_detail})")

    result = {
        "metric": METRIC,
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(secs / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f64": "f64", "f32s": "f32", "f16s": "f16"}[args.storage],
        "data": "synthetic",
        "config": {
            "workload": f"{cfg['name']}: thermal D3Q27 f+g, velocity inlet u=0.05, outflow, periodic y/z, "
                        f"EB sphere r={cfg['eb']['radius']:g} at {cfg['eb']['center']}, tau=0.55",
            "cells": cells,
            "storage": args.storage,
            "kernel": "+".join(per_kernel_ms),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(args.storage, per_kernel_ms) if n == 256 else None,
            "kernel": "+".join(per_kernel_ms),
            "bytes_per_update": bpu,
            "kernel_ms_per_step": {k: round(v, 5) for k, v in per_kernel_ms.items()},
            "copy_ceiling_gbs": round(copy_gbs, 1),
            "copy_ceiling_detail": copy_detail,
            "frac_of_copy_ceiling": round(achieved / copy_gbs, 4),
        },
        "cpu_baseline": None,
        "f32_shifted": legs["f32s"],
        "f32_shifted_fp64_arithmetic": legs["f32s_fp64"],
        "f16_scaled": legs["f16s"],
    }
    if not args.no_cpu_baseline:
        try:
            result["cpu_baseline"] = cpu_baseline(cfg, args.cpu_budget, min(args.cpu_threads, os.cpu_count() or 1))
        except Exception as e:  # report, do not hide
            log(f"cpu baseline failed: {e!r}")
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

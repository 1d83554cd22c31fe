#!/usr/bin/env python3
"""Throughput of the FP64 pseudo-spectral RK4 advance on one MI355X.

Workload (default): decaying isotropic turbulence LES at 128^3 modes (192^3 with
3/2 padding), FP64 throughout, Smagorinsky closure Cs = 0.16, nu = 1e-3, CFL-limited
dt (CFL = 0.5), seeded random-phase solenoidal field with E(k) ~ k^4 exp(-2 (k/4)^2),
u_rms = 1.  One "step" is one full advanceQ_RK4: the CFL dt on the stale velocity
and four low-storage stages, all on the GPU.

Prints one JSON line: Mpts*RK4-steps/s (N^3 * steps / wall second) with the
dominant kernel's HBM roofline figure (algorithmic bytes per launch / mean launch
time from HIP events) and the CPU baseline (the oracle's numpy/scipy restatement
of the reference on this host's cores, timed on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak
C16 = 16               # bytes per complex128


def kernel_bytes(N, closure):
    """Compulsory (algorithmic) bytes per launch of each kernel class at N^3, 3/2 rule:
    every input field read once, every output field written once."""
    M, H, Nc = 3 * N // 2, N // 2, N // 2 + 1
    smag = closure == "smagorinsky"
    nX, nY, nZ = (6, 9, 12) if smag else (3, 3, 6)
    state = N * N * Nc * C16
    return {
        "x_inverse": 3 * N * N * H * C16 + nX * M * N * H * C16,
        "y_inverse": nX * M * N * H * C16 + nY * M * M * H * C16,
        "z_pass": nY * M * M * H * C16 + nZ * M * M * H * C16,
        "y_forward": nZ * M * M * H * C16 + nZ * M * N * H * C16,
        "x_forward_epilogue": nZ * M * N * H * C16 + 3 * 3 * state,
        "cfl_x": 6 * state,
        "cfl_y": 6 * state,
        "cfl_z": 3 * state,
    }


def survey_bytes_per_step(N, closure):
    """The 3-pass pencil model's algorithmic bytes per RK4 step (4 stages):
    DNS 16[h(3N^2+18MN+18M^2)+9N^2Nc], Smagorinsky 16[h(3N^2+24MN+30M^2)+9N^2Nc]."""
    M, h, Nc = 1.5 * N, N / 2, N // 2 + 1
    if closure == "smagorinsky":
        b = 16 * (h * (3 * N * N + 24 * M * N + 30 * M * M) + 9 * N * N * Nc)
    else:
        b = 16 * (h * (3 * N * N + 18 * M * N + 18 * M * M) + 9 * N * N * Nc)
    return 4 * b


def measured_traffic(kernel, N, closure):
    """HBM bytes per launch of `kernel` as rocprofv3 FETCH_SIZE/WRITE_SIZE measured them (copy-calibrated),
    read from the committed counter summary for this workload; None when there is none."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"{'smag' if closure == 'smagorinsky' else 'dns'}{N}_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 6 and cells[0] == kernel:
            try:
                return int(float(cells[5]))
            except ValueError:
                return None
    return None


def cpu_baseline(N, closure, nu, cfl, steps, seed):
    """The oracle (numpy restatement of the reference's src/ path, scipy.fft with one
    worker per available core) advancing the same field; bounded sample of `steps`."""
    import numpy as np

    from oracle import spectral_oracle as so
    from pyspectral_amd import cases

    cores = so.default_workers()
    uh = cases.decaying_isotropic(N, seed=seed)
    g = so.Grid(N, N, N)
    st = so.State(g, *uh, nu=nu, turb_model=1 if closure == "smagorinsky" else 0, cfl=cfl)
    fft = so.FFTBackend(workers=cores)
    t0 = time.perf_counter()
    for _ in range(steps):
        so.advance_rk4(st, fft)
    el = time.perf_counter() - t0
    del np
    return {"value": N ** 3 * steps / el / 1e6, "unit": "Mpts*RK4-steps/s", "cores": cores, "kind": "port",
            "sample": f"{steps} RK4 steps of the same {N}^3 {closure} case (oracle, {fft.name}), "
                      f"{el:.1f} s wall", "seconds": el}


def bench_ensemble(args):
    """96 seeded decaying-turbulence members, 32 each at 32^3, 48^3 and 64^3 modes,
    nu ~ U[5e-4, 5e-3], k0 in {3, 4, 5}, Smagorinsky, CFL 0.5; every step is one
    replay of the captured ensemble graph.  Also times the same members stepped one
    SpectralSolver at a time (unbatched) over a few steps."""
    import numpy as np
    import torch

    from pyspectral_amd import cases
    from pyspectral_amd.ensemble import Ensemble
    from pyspectral_amd.solver import SpectralSolver

    rng = np.random.default_rng(args.seed)
    sizes = [32] * 32 + [48] * 32 + [64] * 32
    nus = list(rng.uniform(5e-4, 5e-3, len(sizes)))
    k0s = list(rng.choice([3.0, 4.0, 5.0], len(sizes)))
    ens = Ensemble([(n, n, n) for n in sizes], nus, model=1, cfl=0.5)
    q = ens.new_state()
    for i, n in enumerate(sizes):
        ens.member(q, i).copy_(torch.from_numpy(cases.decaying_isotropic(n, seed=args.seed + i, k0=k0s[i])))
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        qs, dt = q.clone(), ens.new_dt()
        for _ in range(args.warmup):
            ens.step_graph(q, qs, dt)
    stream.synchronize()
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        for _ in range(args.steps):
            ens.step_graph(q, qs, dt)
    stream.synchronize()
    el = time.perf_counter() - t0
    pts = sum(n ** 3 for n in sizes)
    value = pts * args.steps / el / 1e6
    E = ens.stats(q).cpu().numpy()
    # the same members, one solver each, eager launches
    singles = [SpectralSolver(n, n, n, 1, nu=nu, cfl=0.5) for n, nu in zip(sizes, nus)]
    uq = [ens.member(q, i).clone() for i in range(len(sizes))]
    us = [x.clone() for x in uq]
    ns = 10
    for i, s in enumerate(singles):
        s.step(uq[i], us[i])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(ns):
        for i, s in enumerate(singles):
            s.step(uq[i], us[i])
    torch.cuda.synchronize()
    el1 = time.perf_counter() - t1
    for s in singles:
        s.close()
    bytes_step = sum(survey_bytes_per_step(n, "smagorinsky") for n in sizes)
    return {
        "metric": "Mpts*RK4-steps/s, ensemble of 96 LES members (32x 32^3, 48^3, 64^3)",
        "value": round(value, 2), "unit": "Mpts*RK4-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded random-phase isotropic fields)",
        "config": {"workload": "batched ensemble, Smagorinsky, 96 members 32x(32^3, 48^3, 64^3), "
                               "nu ~ U[5e-4, 5e-3], k0 in {3,4,5}, CFL 0.5, one hipGraph replay per step",
                   "members": len(sizes), "points_per_step": pts, "parallelism": "single GPU"},
        "step_roofline": {"model_bytes_per_step": int(bytes_step),
                          "achieved": round(bytes_step / (el / args.steps) / 1e9, 1), "unit": "GB/s",
                          "frac": round(bytes_step / (el / args.steps) / 1e9 / HBM_PEAK_GBS, 4)},
        "unbatched": {"value": round(pts * ns / el1 / 1e6, 2), "ms_per_step": round(el1 / ns * 1e3, 3),
                      "how": "the same 96 members, one SpectralSolver.step each, eager launches"},
        "stats_after": {"E_mean": float(E[:, 0].mean()), "finite": bool(np.isfinite(E).all())},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--N", type=int, default=128)
    ap.add_argument("--closure", default="smagorinsky", choices=["smagorinsky", "dns"])
    ap.add_argument("--event-steps", type=int, default=50, help="steps timed per kernel with HIP events")
    ap.add_argument("--cpu-steps", type=int, default=2, help="oracle steps for the CPU baseline (0: skip)")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--no-autotune", action="store_true", help="keep the built-in kernel shapes")
    ap.add_argument("--workload", default="les128", choices=["les128", "ensemble"],
                    help="les128: the 128^3 LES metric (default); ensemble: 96 members of 32^3/48^3/64^3")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a known-size dispatch "
                         "that calibrates rocprof FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build advances one grid on one GPU; runs on more than one GPU are not part "
              "of it (see DESIGN.md for how the grid would shard).", file=sys.stderr)
        sys.exit(2)

    import ctypes as C

    import torch

    from pyspectral_amd import _lib, cases
    from pyspectral_amd.solver import SpectralSolver

    if args.workload == "ensemble":
        print(json.dumps(bench_ensemble(args)))
        return

    if args.copy_mib > 0:
        n = args.copy_mib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    N, nu, cfl = args.N, 1e-3, 0.5
    model = 1 if args.closure == "smagorinsky" else 0
    uh = cases.decaying_isotropic(N, seed=args.seed)
    solver = SpectralSolver(N, N, N, model, nu=nu, cfl=cfl, Cs=0.16)
    q = torch.from_numpy(uh).to("cuda")
    qs = q.clone()
    tuning = solver.tuning() if args.no_autotune else solver.autotune(q, reps=5)
    for _ in range(args.warmup):
        solver.step(q, qs)
    torch.cuda.synchronize()

    t0 = time.perf_counter()
    for _ in range(args.steps):
        solver.step(q, qs)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms_per_step = el / args.steps * 1e3
    value = N ** 3 * args.steps / el / 1e6

    # per-kernel durations from HIP events on the library's stream
    lib = solver.lib
    _lib.check(solver.ctx, lib.psx_timing_begin(solver.ctx), "psx_timing_begin")
    for _ in range(args.event_steps):
        solver.step(q, qs)
    ms = (C.c_double * len(_lib.KERNEL_NAMES))()
    cnt = (C.c_int * len(_lib.KERNEL_NAMES))()
    _lib.check(solver.ctx, lib.psx_timing_end(solver.ctx, ms, cnt), "psx_timing_end")
    kb = kernel_bytes(N, args.closure)
    kernels = {}
    for i, name in enumerate(_lib.KERNEL_NAMES):
        if cnt[i] == 0:
            continue
        avg_ms = ms[i] / cnt[i]
        entry = {"launches": cnt[i], "avg_us": round(avg_ms * 1e3, 2),
                 "share": round(ms[i] / max(sum(ms), 1e-30), 4)}
        if name in kb:
            entry["GBs"] = round(kb[name] / (avg_ms * 1e-3) / 1e9, 1)
        kernels[name] = entry
    dom = max(kernels, key=lambda k: kernels[k]["share"])
    ach = kb[dom] / (kernels[dom]["avg_us"] * 1e-6) / 1e9
    E = solver.stats(qs).cpu().tolist()
    finite = all(map(lambda x: x == x and abs(x) < 1e300, E))

    sb = survey_bytes_per_step(N, args.closure)
    out = {
        "metric": "Mpts*RK4-steps/s at 128^3 FP64 LES",
        "value": round(value, 2),
        "unit": "Mpts*RK4-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random-phase isotropic field)",
        "config": {"workload": f"decaying isotropic turbulence LES, {args.closure}, {N}^3 modes "
                               f"({3 * N // 2}^3 padded), nu={nu}, CFL={cfl}, RK4 (low-storage 4-stage)",
                   "grid": N, "closure": args.closure, "cfl": cfl, "nu": nu, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(dom, N, args.closure),
                     "bytes_per_launch": kb[dom]},
        "step_roofline": {"model_bytes_per_step": int(sb), "achieved": round(sb / (ms_per_step * 1e-3) / 1e9, 1),
                          "unit": "GB/s", "frac": round(sb / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "kernels": kernels,
        "tuning": tuning,
        "stats_after": {"E": E[0], "Omega": E[1], "finite": finite},
    }
    if args.cpu_steps > 0:
        out["cpu_baseline"] = cpu_baseline(N, args.closure, nu, cfl, args.cpu_steps, args.seed)
        out["speedup_vs_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

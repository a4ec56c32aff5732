"""Benchmark: D3Q19 GLUPS at 256^3 (Taylor-Green, periodic, tau 0.55, u0 0.04),
FP32 storage vs FP16 (f_i - w_i) storage, % of HBM roofline.

    python bench.py [--steps K] [--warmup W] [--size N] [--no-cpu] [--no-fp16]

A "step" is one lattice time step over the whole 256^3 lattice (one launch of
the fused stream-collide kernel).  Populations start resident in HBM (set
before the timed region).  Prints ONE JSON line.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "bit-physics_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E spec peak
Q = 19


def bytes_per_cell(storage: str) -> int:
    """Algorithmic HBM bytes per cell update: read 19 + write 19 populations."""
    return 2 * Q * {"f64": 8, "f32": 4, "f16": 2, "bf16": 2}[storage]


def time_lattice(amd, torch, shape, tau, storage, rho0, u0, steps, warmup):
    stream = torch.cuda.current_stream()
    lat = amd.Lattice(shape, tau, storage=storage)
    lat.set_equilibrium(rho0, u0)
    lat.step(warmup)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    t0.record(stream)
    lat.step(steps)
    t1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = t0.elapsed_time(t1)
    diag = lat.diagnostics()
    lat.close()
    return ms, wall, diag


def copy_peak_gbs(amd, torch, nbytes=2 << 30, reps=20):
    """Achievable HBM bandwidth: the library's plain 16 B/lane copy kernel."""
    import ctypes

    from lattice_boltzmann_d3q19_amd._lib import call

    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    for _ in range(3):
        call("d3q19_copy_kernel", ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()), nbytes, sp)
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for _ in range(reps):
        call("d3q19_copy_kernel", ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()), nbytes, sp)
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * nbytes / (ms * 1e-3) / 1e9


def cpu_baseline(shape, tau, rho0, u0, sample_steps):
    """The FP64 CPU oracle (C, OpenMP) on a bounded sample of the same workload."""
    import numpy as np

    import oracle

    oracle.build() if not (ROOT / "oracle" / "_build" / "libd3q19_oracle.so").exists() else None
    f = oracle.feq_field(rho0, u0)
    t = time.perf_counter()
    f = oracle.run_reference(f, tau, sample_steps)
    dt = time.perf_counter() - t
    ncell = int(np.prod(shape))
    return {
        "value": ncell * sample_steps / dt / 1e9,
        "unit": "GLUPS",
        "cores": oracle.num_threads(),
        "kind": "port",
        "sample": f"{sample_steps} steps of the same {shape[0]}x{shape[1]}x{shape[2]} Taylor-Green "
                  f"lattice, FP64 C restatement of bgk_step (collide+stream), OpenMP, {dt:.2f} s",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--tau", type=float, default=0.55)
    ap.add_argument("--u0", type=float, default=0.04)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp16", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this engine has no multi-GPU path "
              "(the sharding design is documented, not built). Run with --gpus 1.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import lattice_boltzmann_d3q19_amd as amd
    from lattice_boltzmann_d3q19_amd.initial import taylor_green, taylor_green_ke_decay

    torch.cuda.set_device(0)
    n = args.size
    shape = (n, n, n)
    ncell = n ** 3
    rho0, u0 = taylor_green(shape, args.u0)
    nu = (args.tau - 0.5) / 3.0

    copy_gbs = copy_peak_gbs(amd, torch)
    ms, wall, diag = time_lattice(amd, torch, shape, args.tau, "f32", rho0, u0, args.steps, args.warmup)
    glups = ncell * args.steps / (ms * 1e-3) / 1e9
    kernel_ms = ms / args.steps
    achieved = bytes_per_cell("f32") * ncell / (kernel_ms * 1e-3) / 1e9

    # KE after warmup + steps against the linear decay of the initial KE
    ke0 = 0.5 * float(np.sum(rho0 * np.sum(u0 * u0, axis=0)))
    t_total = args.warmup + args.steps
    ke_ratio = diag["kinetic_energy"] / ke0
    ke_lin = taylor_green_ke_decay(shape, nu, t_total)

    traffic = None
    tf = ROOT / "profiles" / f"hbm_traffic_k_step_f32_tg{n}.json"
    if tf.exists():
        traffic = json.loads(tf.read_text()).get("bytes_per_launch")

    result = {
        "metric": "D3Q19 GLUPS at 256^3 FP32 vs FP16-store, % HBM roofline; adjoint/forward ratio",
        "value": round(glups, 4),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"taylor-green-{n}^3-periodic",
            "nx": n, "ny": n, "nz": n, "tau": args.tau, "u0": args.u0, "storage": "f32",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "copy_peak_GBs": round(copy_gbs, 1),
        "frac_of_copy_peak": round(achieved / copy_gbs, 4),
        "kernel": "d3q19::k_step<TF32, BC_PERIODIC>",
        "algorithmic_bytes_per_cell": bytes_per_cell("f32"),
        "wall_s": round(wall, 3),
        "ke_ratio": ke_ratio,
        "ke_ratio_linear_theory": ke_lin,
        "mass_rel_drift": (diag["mass"] - float(np.sum(rho0))) / float(np.sum(rho0)),
    }

    if not args.no_fp16:
        ms16, _, diag16 = time_lattice(amd, torch, shape, args.tau, "f16", rho0, u0, args.steps, args.warmup)
        glups16 = ncell * args.steps / (ms16 * 1e-3) / 1e9
        ach16 = bytes_per_cell("f16") * ncell / (ms16 / args.steps * 1e-3) / 1e9
        result["fp16_store"] = {
            "value": round(glups16, 4),
            "unit": "GLUPS",
            "ratio_vs_fp32": round(glups16 / glups, 4),
            "achieved_GBs": round(ach16, 1),
            "frac": round(ach16 / HBM_PEAK_GBS, 4),
            "ke_rel_diff_vs_fp32": abs(diag16["kinetic_energy"] - diag["kinetic_energy"]) / diag["kinetic_energy"],
        }

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(shape, args.tau, rho0, u0, args.cpu_steps)

    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

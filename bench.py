"""Benchmark: D3Q19 GLUPS at 256^3 (Taylor-Green, periodic, tau 0.55, u0 0.04),
FP32 storage vs FP16 (f_i - w_i) storage, % of HBM roofline.

    python bench.py [--steps K] [--warmup W] [--size N] [--no-cpu] [--no-fp16] [--no-me16]

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
    """Algorithmic HBM bytes per cell update: read 19 + write 19 populations
    (the moment-encoded step reads and writes 19 16-bit codes; each code is
    read by about 4.8 pulls, served from cache)."""
    return 2 * Q * {"f64": 8, "f32": 4, "f16": 2, "bf16": 2, "bf16sr": 2, "me16": 2}[storage]


class ShaderClock:
    """The device's shader clock in MHz (libd3q19's d3q19_shader_clock: one
    wavefront counts shader-clock ticks against the 100 MHz real-time counter
    for ~20 us).  Read right before and after a timed window."""

    def __init__(self, torch):
        self.torch = torch
        self.buf = torch.zeros(2, dtype=torch.int64, device="cuda")

    def __call__(self) -> float:
        import ctypes

        from lattice_boltzmann_d3q19_amd._lib import call

        mhz = ctypes.c_double()
        call("d3q19_shader_clock", ctypes.c_void_p(self.buf.data_ptr()), ctypes.byref(mhz),
             ctypes.c_void_p(self.torch.cuda.current_stream().cuda_stream))
        return round(mhz.value, 1)


def settle(torch, lat, settle_ms, chunk=10, max_steps=None):
    """Run steps of `lat` until settle_ms of device time have passed (or
    max_steps steps); returns the number of steps.  At the start of a heavy
    launch sequence after host idle time the MI355X shader clock dips
    (2400 -> ~1700 MHz) for some tens of milliseconds and recovers; a
    VALU-heavier step (16-bit storage, moment codes) slows by up to 25% inside
    the dip, a memory-bound one barely moves."""
    stream = torch.cuda.current_stream()
    n, ms = 0, 0.0
    while ms < settle_ms and (max_steps is None or n + chunk <= max_steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        lat.step(chunk)
        e1.record(stream)
        torch.cuda.synchronize()
        ms += e0.elapsed_time(e1)
        n += chunk
    return n


# Untimed steps at most before the moment-encoded leg's settled window: its
# code ranges are calibrated over this many steps more than warm-up + timed.
ME16_SETTLE_CAP = 1000


def time_lattice(amd, torch, shape, tau, storage, rho0, u0, steps, warmup, settle_ms=0.0, clock=None,
                 me_ranges=None):
    """Time `steps` steps after `warmup` untimed ones (plus, with settle_ms > 0,
    untimed steps for that much device time before the warm-up).  A
    moment-encoded ("me16") lattice takes its code ranges from me_ranges
    (calibrated over warm-up + timed steps if not given)."""
    stream = torch.cuda.current_stream()
    settle_steps = 0
    if storage == "me16":
        from lattice_boltzmann_d3q19_amd.moment_encoded import moment_encoded_lattice

        if settle_ms > 0 and me_ranges is None:
            raise ValueError("a settled moment-encoded window needs ranges calibrated past the settle steps")
        lat = moment_encoded_lattice(shape, tau, warmup + steps, ranges=me_ranges, rho=rho0, u=u0)
        if settle_ms > 0:
            settle_steps = settle(torch, lat, settle_ms, max_steps=ME16_SETTLE_CAP)
    else:
        lat = amd.Lattice(shape, tau, storage=storage)
        lat.set_equilibrium(rho0, u0)
        if settle_ms > 0:
            settle_steps = settle(torch, lat, settle_ms)
    lat.step(warmup)
    torch.cuda.synchronize()
    clk0 = clock() if clock else None
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    t0.record(stream)
    lat.step(steps)
    t1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    clk1 = clock() if clock else None
    ms = t0.elapsed_time(t1)
    diag = lat.diagnostics()
    launches = lat.step_launches
    lat.close()
    info = {"settle_steps": settle_steps, "sclk_MHz_before": clk0, "sclk_MHz_after": clk1, "launches": launches}
    return ms, wall, diag, info


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


def adjoint_timing(torch, shape=(128, 128, 128), tau=0.6, steps=2000, K=50, storage="f32", seed=42):
    """Forward (with checkpoints) and backward (recompute + adjoint) wall times on
    the inverse-problem workload: 128^3 periodic, tau 0.6, 2000 steps, K = 50."""
    from lattice_boltzmann_d3q19_amd.autograd import DifferentiableLBM
    from lattice_boltzmann_d3q19_amd.initial import random_solenoidal

    u0 = torch.tensor(random_solenoidal(shape, 0.02, seed=seed), device="cuda", dtype=torch.float32)
    F = torch.zeros_like(u0)
    g = torch.tensor(random_solenoidal(shape, 1e-3, seed=seed + 1), device="cuda", dtype=torch.float32)
    solver = DifferentiableLBM(shape, tau, steps, checkpoint_every=K, storage=storage)
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    solver.forward_raw(u0, F)  # warm-up
    solver.backward_raw(g)
    torch.cuda.synchronize()
    ev[0].record(stream)
    solver.forward_raw(u0, F)
    ev[1].record(stream)
    solver.backward_raw(g)
    ev[2].record(stream)
    torch.cuda.synchronize()
    fwd = ev[0].elapsed_time(ev[1])
    bwd = ev[1].elapsed_time(ev[2])
    mem = solver.memory_bytes
    solver.close()
    return {
        "workload": f"inverse-{shape[0]}^3-periodic-{steps}steps-K{K}-{storage}",
        "forward_ms": round(fwd, 3),
        "backward_ms": round(bwd, 3),
        "adjoint_forward_ratio": round(bwd / fwd, 4),
        "forward_GLUPS": round(shape[0] * shape[1] * shape[2] * steps / (fwd * 1e-3) / 1e9, 3),
        "checkpoint_memory_GB": round(mem / 1e9, 3),
    }


def study_turbulence(amd, torch, n=384, steps=4000, every=250, tau=0.51, rms=0.03, seed=42):
    """Decaying turbulence: FP16 / BF16 (f - w) and FP32 storage against FP64 storage."""
    from lattice_boltzmann_d3q19_amd.initial import random_solenoidal, spectral_divergence

    shape = (n, n, n)
    u0 = random_solenoidal(shape, rms, seed=seed)
    div = spectral_divergence(u0)
    runs = {}
    for storage in ("f64", "f32", "f16", "bf16", "bf16sr", "me16"):
        if storage == "me16":
            from lattice_boltzmann_d3q19_amd.moment_encoded import moment_encoded_lattice

            lat = moment_encoded_lattice(shape, tau, steps, u=u0)
        else:
            lat = amd.Lattice(shape, tau, storage=storage)
            lat.set_equilibrium(None, u0)
        series = []
        total_ms = 0.0
        stream = torch.cuda.current_stream()
        for k in range(0, steps + 1, every):
            if k > 0:
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                lat.step(every)
                e1.record(stream)
                torch.cuda.synchronize()
                total_ms += e0.elapsed_time(e1)
            d = lat.diagnostics()
            series.append((k, d["kinetic_energy"], lat.enstrophy(), d["mass"]))
        lat.close()
        runs[storage] = {"series": series, "GLUPS": n ** 3 * steps / (total_ms * 1e-3) / 1e9}
        print(f"[study] turbulence {storage}: {runs[storage]['GLUPS']:.2f} GLUPS, "
              f"KE {series[0][1]:.6e} -> {series[-1][1]:.6e}", file=sys.stderr, flush=True)
    ref = runs["f64"]["series"]
    out = {"study": "turbulence", "workload": f"decaying-turbulence-{n}^3-periodic-tau{tau}-{steps}steps",
           "ic": {"rms": rms, "spectrum": "k^4 exp(-2 (k/8)^2)", "seed": seed, "spectral_divergence": div},
           "reference": "FP64 storage (FP64 collision) on the GPU"}
    for storage in ("f32", "f16", "bf16", "bf16sr", "me16"):
        ser = runs[storage]["series"]
        ke = max(abs(a[1] - b[1]) / b[1] for a, b in zip(ser, ref))
        en = max(abs(a[2] - b[2]) / b[2] for a, b in zip(ser, ref))
        ms = max(abs(a[3] - b[3]) / b[3] for a, b in zip(ser, ref))
        out[storage] = {"GLUPS": round(runs[storage]["GLUPS"], 3), "ke_max_rel_err": ke,
                        "enstrophy_max_rel_err": en, "mass_max_rel_err": ms,
                        "ke_final_rel_err": abs(ser[-1][1] - ref[-1][1]) / ref[-1][1]}
    out["f64"] = {"GLUPS": round(runs["f64"]["GLUPS"], 3),
                  "ke": [r[1] for r in ref], "enstrophy": [r[2] for r in ref], "mass": [r[3] for r in ref],
                  "steps": [r[0] for r in ref]}
    out["fp16_over_fp32_GLUPS"] = round(runs["f16"]["GLUPS"] / runs["f32"]["GLUPS"], 4)
    return out


def study_inverse(torch, n=128, steps=2000, K=50, tau=0.6, ndir=16, eps=1e-3, seed=42):
    """Inverse problem: adjoint gradients (FP32 and FP64) against central
    differences of the FP64 forward map along seeded random unit directions
    in (u0, F) (F components drawn 1e-3 smaller), step eps."""
    from lattice_boltzmann_d3q19_amd.autograd import DifferentiableLBM, velocity_matching_loss
    from lattice_boltzmann_d3q19_amd.initial import random_solenoidal

    shape = (n, n, n)
    dev = "cuda"
    u_true = torch.tensor(random_solenoidal(shape, 0.02, seed=seed), device=dev)
    zeros = torch.zeros_like(u_true)
    f64 = DifferentiableLBM(shape, tau, steps, checkpoint_every=K, storage="f64")
    target = f64.forward_raw(u_true, zeros).clone()
    u0 = torch.tensor(random_solenoidal(shape, 0.02, seed=seed + 7), device=dev)

    def grad(storage):
        s = f64 if storage == "f64" else DifferentiableLBM(shape, tau, steps, checkpoint_every=K, storage=storage)
        a = u0.clone().requires_grad_(True)
        b = zeros.clone().requires_grad_(True)
        loss = velocity_matching_loss(s(a, b), target)
        loss.backward()
        if s is not f64:
            s.close()
        return float(loss.detach()), a.grad.double(), b.grad.double()

    L32, gu32, gF32 = grad("f32")
    L64, gu64, gF64 = grad("f64")

    def loss_at(a, b):
        return float(velocity_matching_loss(f64.forward_raw(a, b), target))

    gen = torch.Generator(device="cpu").manual_seed(seed)
    fds, ad32, ad64 = [], [], []
    for k in range(ndir):
        du = torch.randn(u0.shape, generator=gen, dtype=torch.float64).to(dev)
        dF = 1e-3 * torch.randn(u0.shape, generator=gen, dtype=torch.float64).to(dev)
        nrm = (du.square().sum() + dF.square().sum()).sqrt()  # unit direction in (u0, F)
        du, dF = du / nrm, dF / nrm
        fds.append((loss_at(u0 + eps * du, zeros + eps * dF) - loss_at(u0 - eps * du, zeros - eps * dF)) / (2 * eps))
        ad32.append(float((gu32 * du).sum() + (gF32 * dF).sum()))
        ad64.append(float((gu64 * du).sum() + (gF64 * dF).sum()))
    scale = max(abs(v) for v in fds)
    rel = lambda a, b: float((a - b).norm() / b.norm())
    out = {"study": "inverse", "workload": f"inverse-{n}^3-periodic-tau{tau}-{steps}steps-K{K}",
           "loss_f32": L32, "loss_f64": L64, "fd_eps": eps, "directions": ndir,
           "fd_max_err_f32_adjoint": max(abs(a - f) for a, f in zip(ad32, fds)) / scale,
           "fd_max_err_f64_adjoint": max(abs(a - f) for a, f in zip(ad64, fds)) / scale,
           "f32_vs_f64_adjoint_grad_u0_rel": rel(gu32, gu64), "f32_vs_f64_adjoint_grad_F_rel": rel(gF32, gF64),
           "fd": fds, "adjoint_f32": ad32}
    out["timing"] = adjoint_timing(torch, shape, tau, steps, K, "f32", seed)
    f64.close()
    return out


def study_channel(amd, torch, shape=(256, 128, 128), tau=0.6, steps=4000, gx=1e-6, budget=64e9,
                  ndir=4, eps=1e-3, seed=42):
    """Channel with Brinkman obstacles: dJ/ds for J = mean outlet x-velocity,
    FP32 compute with FP16 (f - w) checkpoints vs FP32 checkpoints (K
    autotuned to the HBM budget), central differences of the FP64 forward."""
    from lattice_boltzmann_d3q19_amd.autograd import DifferentiableLBM, outlet_velocity
    from lattice_boltzmann_d3q19_amd.initial import random_spheres

    s_np = random_spheres(shape, 6.0, 12.0, 0.05, seed=seed, wall_axis=1)
    dev = "cuda"
    s = torch.tensor(s_np, device=dev)
    u0 = torch.zeros((3,) + shape, device=dev, dtype=torch.float64)
    F = torch.zeros_like(u0)
    F[0] = gx
    kw = dict(boundary="halfway", wall_faces=("-y", "+y"), brinkman=True, checkpoint_every=0,
              memory_budget=int(budget))
    out = {"study": "channel", "workload": f"channel-{shape[0]}x{shape[1]}x{shape[2]}-tau{tau}-{steps}steps",
           "solid_volume_fraction": float(s_np.mean()), "gx": gx, "memory_budget_GB": budget / 1e9}
    grads = {}
    for cks in ("f16", "f32"):
        solver = DifferentiableLBM(shape, tau, steps, storage="f32", checkpoint_storage=cks, **kw)
        stream = torch.cuda.current_stream()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        ev[0].record(stream)
        uT = solver.forward_raw(u0, F, s)
        ev[1].record(stream)
        g = torch.zeros_like(uT)
        g[0, -1] = 1.0 / (shape[1] * shape[2])
        _, _, gs = solver.backward_raw(g)
        ev[2].record(stream)
        torch.cuda.synchronize()
        fwd, bwd = ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2])
        grads[cks] = gs.double()
        out[f"checkpoints_{cks}"] = {"K": solver.checkpoint_every, "memory_GB": round(solver.memory_bytes / 1e9, 3),
                                     "J": float(outlet_velocity(uT)), "forward_ms": round(fwd, 1),
                                     "backward_ms": round(bwd, 1), "adjoint_forward_ratio": round(bwd / fwd, 3)}
        solver.close()
        torch.cuda.empty_cache()
        print(f"[study] channel {cks} checkpoints: {out[f'checkpoints_{cks}']}", file=sys.stderr, flush=True)
    a, b = grads["f16"], grads["f32"]
    out["grad_rel_err_f16_vs_f32_checkpoints"] = float((a - b).norm() / b.norm())

    # central differences of J(s) with the FP64 forward (stateful lattice)
    def J(sv):
        lat = amd.Lattice(shape, tau, storage="f64", boundary="halfway", wall_faces=("-y", "+y"),
                          force=(gx, 0.0, 0.0))
        lat.set_brinkman(sv).set_equilibrium()
        lat.step(steps)
        _, u = lat.moments()
        lat.close()
        return float(u[0, -1].mean())

    gen = torch.Generator(device="cpu").manual_seed(seed)
    fds, ads = [], []
    for _ in range(ndir):
        d = torch.randn(shape, generator=gen, dtype=torch.float64)
        d /= d.norm()  # unit direction: eps is the step length in parameter space
        fds.append((J((s.cpu() + eps * d).numpy()) - J((s.cpu() - eps * d).numpy())) / (2 * eps))
        ads.append(float((b.cpu() * d).sum()))
    out["fd"] = fds
    out["adjoint_f32_checkpoints"] = ads
    out["fd_max_rel_err"] = max(abs(x - y) for x, y in zip(ads, fds)) / max(abs(v) for v in fds)
    return out


def study_cavity(amd, torch, n=256, steps=1000, warmup=20, lid=0.05):
    """Step rates of the wall modes at n^3 (GLUPS, HIP events): the lid-driven
    cavity (half-way walls on all six faces, lid moving +x on the top z face, as
    the 64^3 cavity of the parity tests),
    half-way walls on z only, the reference's y-walls with a uniform force, and
    the periodic lattice, in FP32 and FP16 storage."""
    shape = (n, n, n)
    cases = {
        "cavity": dict(boundary="halfway", wall_faces=("-x", "+x", "-y", "+y", "-z", "+z"),
                       face_velocity={"+z": (lid, 0.0, 0.0)}),
        "zwalls": dict(boundary="halfway", wall_faces=("-z", "+z")),
        "ywalls": dict(boundary="ywalls", force=(1e-6, 0.0, 0.0)),
        "periodic": dict(),
    }
    out = {"study": "cavity", "workload": f"wall-modes-{n}^3-tau0.6-{steps}steps", "GLUPS": {}}
    stream = torch.cuda.current_stream()
    for name, kw in cases.items():
        for storage in ("f32", "f16"):
            lat = amd.Lattice(shape, 0.6, storage=storage, **kw).set_equilibrium()
            lat.step(warmup)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            lat.step(steps)
            e1.record(stream)
            torch.cuda.synchronize()
            out["GLUPS"][f"{name}_{storage}"] = round(n ** 3 * steps / (e0.elapsed_time(e1) * 1e-3) / 1e9, 2)
            lat.close()
            print(f"[study] cavity {name} {storage}: {out['GLUPS'][f'{name}_{storage}']} GLUPS",
                  file=sys.stderr, flush=True)
    return out


def study_transient(amd, torch, n=256, steps=300, tau=0.55, u0=0.04):
    """Per-step duration of the fused step right after set-up, FP32 and FP16
    storage, and the shader clock along the way.  For each storage: one
    lattice stepped back to back with an event after every step (durations),
    then a second one stepped in pairs with a clock reading after each pair."""
    from lattice_boltzmann_d3q19_amd.initial import taylor_green

    shape = (n, n, n)
    rho0, uu = taylor_green(shape, u0)
    clock = ShaderClock(torch)
    stream = torch.cuda.current_stream()
    out = {"study": "transient", "workload": f"taylor-green-{n}^3-periodic-tau{tau}", "steps": steps,
           "idle_sclk_MHz": clock()}
    copy_peak_gbs(amd, torch)  # the headline bench's first phase
    for storage in ("f32", "f16"):
        lat = amd.Lattice(shape, tau, storage=storage).set_equilibrium(rho0, uu)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        ev[0].record(stream)
        for k in range(steps):
            lat.step(1)
            ev[k + 1].record(stream)
        torch.cuda.synchronize()
        us = [round(ev[k].elapsed_time(ev[k + 1]) * 1e3, 1) for k in range(steps)]
        lat.close()
        lat = amd.Lattice(shape, tau, storage=storage).set_equilibrium(rho0, uu)
        torch.cuda.synchronize()
        clk = []
        for k in range(0, steps, 2):
            lat.step(2)
            clk.append(clock())
        lat.close()
        out[storage] = {"step_us": us, "sclk_MHz_every_2_steps": clk,
                        "mean_us_steps_1_10": sum(us[:10]) / 10, "mean_us_steps_11_40": sum(us[10:40]) / 30,
                        "mean_us_last_100": sum(us[-100:]) / 100, "min_sclk_MHz": min(clk)}
        print(f"[study] transient {storage}: steps 1-10 {out[storage]['mean_us_steps_1_10']:.1f} us, "
              f"11-40 {out[storage]['mean_us_steps_11_40']:.1f} us, last 100 {out[storage]['mean_us_last_100']:.1f} us, "
              f"min sclk {min(clk)} MHz", file=sys.stderr, flush=True)
    return out


def study_mms(levels=(32, 64, 128, 256, 512, 1024), storages=("f64", "f32")):
    """Forced Taylor-Green MMS ladder on the fused step (solution verification):
    RMS velocity error at t_final per level and storage, observed orders."""
    from lattice_boltzmann_d3q19_amd import verification as V

    out = {"study": "mms", "workload": "forced-taylor-green-NxNx3-periodic-t0.05",
           "known_answers_32_64_128": [V.KNOWN_ANSWERS[n] for n in (32, 64, 128)]}
    starts = {"reference_start": {}, "consistent_pressure": {"consistent_pressure": True},
              "consistent_pressure_and_stress": {"consistent_init": True}}
    for start, kw in starts.items():
        for storage in storages:
            if start == "consistent_pressure_and_stress" and storage != "f64":
                continue
            runs = []
            for n in levels:
                r = V.mms_error_lattice(n, storage, **kw)
                runs.append(r)
                print(f"[study] mms {start} {storage} N={n}: steps={r['n_steps']} err={r['err']:.4e} "
                      f"({r['seconds']:.2f} s)", file=sys.stderr, flush=True)
            errs = [r["err"] for r in runs]
            out[f"{start}_{storage}"] = {
                "levels": [{k: r[k] for k in ("N", "n_steps", "tau", "ma_max", "err_u", "err_v", "err", "seconds")}
                           for r in runs],
                "order_pairs": [math.log(errs[i] / errs[i + 1]) / math.log(2.0) for i in range(len(errs) - 1)],
                "order_fit_32_128": V.observed_order(levels[:3], errs[:3]),
                "order_fit_128_up": V.observed_order(levels[2:], errs[2:]),
            }
    for start in ("reference_start", "consistent_pressure"):
        a, b = out.get(f"{start}_f32"), out.get(f"{start}_f64")
        if a and b:
            out[f"{start}_f32_vs_f64_err_rel_diff"] = [abs(x["err"] - y["err"]) / y["err"]
                                                       for x, y in zip(a["levels"], b["levels"])]
    return out


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo)."""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(shape, tau, rho0, u0, sample_steps):
    """The FP64 CPU oracle (C, OpenMP) on a bounded sample of the same workload."""
    import numpy as np

    import oracle

    oracle.build() if not (ROOT / "oracle" / "_build" / "libd3q19_oracle.so").exists() else None
    f0 = oracle.feq_field(rho0, u0)
    ncell = int(np.prod(shape))
    threads = oracle.num_threads()
    t = time.perf_counter()
    oracle.run_reference(f0, tau, sample_steps)
    dt = time.perf_counter() - t
    # the reference's own posture is single-threaded: one step on one core
    oracle.set_num_threads(1)
    t = time.perf_counter()
    oracle.run_reference(f0, tau, 1)
    dt1 = time.perf_counter() - t
    oracle.set_num_threads(threads)
    return {
        "value": ncell * sample_steps / dt / 1e9,
        "unit": "GLUPS",
        "cores": threads,
        "cpu_model": cpu_model(),
        "nproc": os.cpu_count(),
        "affinity_cpus": len(os.sched_getaffinity(0)),
        "kind": "port",
        "sample": f"{sample_steps} steps of the same {shape[0]}x{shape[1]}x{shape[2]} Taylor-Green "
                  f"lattice, FP64 C restatement of bgk_step (collide+stream), OpenMP, {dt:.2f} s",
        "single_thread": {"value": ncell / dt1 / 1e9, "unit": "GLUPS", "cores": 1,
                          "sample": f"1 step of the same lattice on one thread, {dt1:.2f} s"},
    }


def config_baselines(amd, torch):
    """CPU (FP64 oracle, OpenMP) MLUPS beside the GPU GLUPS (FP32 storage) of the
    same workload, for the cavity (C1, full run) and turbulence (C3, 5 steps) configs."""
    import numpy as np

    import oracle
    from lattice_boltzmann_d3q19_amd.initial import random_solenoidal

    def gpu_rate(lat, steps):
        lat.step(5)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        lat.step(steps)
        e1.record()
        torch.cuda.synchronize()
        rate = math.prod(lat.shape) * steps / (e0.elapsed_time(e1) * 1e-3) / 1e9
        lat.close()
        return rate

    out = {}
    # C1: 64^3 lid-driven cavity, half-way walls on all six faces, lid (0.05, 0, 0) on z = Nz - 1, tau 0.6
    shape, tau, steps = (64, 64, 64), 0.6, 1000
    fu = np.zeros((6, 3))
    fu[5] = (0.05, 0.0, 0.0)
    f0 = oracle.feq_field(np.ones(shape), np.zeros((3,) + shape))
    t = time.perf_counter()
    oracle.run_halfway(f0, tau, steps, wall_mask=63, face_velocity=fu)
    dt = time.perf_counter() - t
    lat = amd.Lattice(shape, tau, storage="f32", boundary="halfway",
                      wall_faces=("-x", "+x", "-y", "+y", "-z", "+z"),
                      face_velocity={"+z": (0.05, 0.0, 0.0)}).set_equilibrium()
    out["C1-cavity-64^3-1000steps"] = {"cpu_MLUPS": math.prod(shape) * steps / dt / 1e6, "cpu_s": round(dt, 2),
                                       "gpu_GLUPS_f32": gpu_rate(lat, steps)}
    # C3: 384^3 decaying turbulence (random-phase solenoidal, rms 0.03), tau 0.51, 5 CPU steps
    shape, tau = (384, 384, 384), 0.51
    u = random_solenoidal(shape, 0.03, seed=42)
    f0 = oracle.feq_field(np.ones(shape), u)
    t = time.perf_counter()
    oracle.run_reference(f0, tau, 5)
    dt = time.perf_counter() - t
    del f0
    lat = amd.Lattice(shape, tau, storage="f32").set_equilibrium(None, torch.tensor(u, device="cuda",
                                                                                      dtype=torch.float32))
    del u
    out["C3-turbulence-384^3-5steps"] = {"cpu_MLUPS": math.prod(shape) * 5 / dt / 1e6, "cpu_s": round(dt, 2),
                                         "gpu_GLUPS_f32": gpu_rate(lat, 200)}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (headline default 5000; --study cavity: steps per case, default 1000)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--tau", type=float, default=0.55)
    ap.add_argument("--u0", type=float, default=0.04)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--settle-ms", type=float, default=200.0,
                    help="untimed device time before the FP16 and moment-encoded legs' settled windows")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--no-me16", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--study", choices=("turbulence", "inverse", "channel", "mms", "cavity", "transient"),
                    default=None,
                    help="run a validation study instead of the headline benchmark")
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
    if args.study == "turbulence":
        print(json.dumps(study_turbulence(amd, torch)), flush=True)
        return 0
    if args.study == "inverse":
        print(json.dumps(study_inverse(torch)), flush=True)
        return 0
    if args.study == "channel":
        print(json.dumps(study_channel(amd, torch)), flush=True)
        return 0
    if args.study == "mms":
        print(json.dumps(study_mms()), flush=True)
        return 0
    if args.study == "cavity":
        print(json.dumps(study_cavity(amd, torch, steps=args.steps or 1000)), flush=True)
        return 0
    if args.study == "transient":
        print(json.dumps(study_transient(amd, torch)), flush=True)
        return 0
    if args.steps is None:
        args.steps = 5000
    n = args.size
    shape = (n, n, n)
    ncell = n ** 3
    rho0, u0 = taylor_green(shape, args.u0)
    nu = (args.tau - 0.5) / 3.0

    clock = ShaderClock(torch)
    copy_gbs = copy_peak_gbs(amd, torch)
    ms, wall, diag, info = time_lattice(amd, torch, shape, args.tau, "f32", rho0, u0, args.steps, args.warmup,
                                        clock=clock)
    glups = ncell * args.steps / (ms * 1e-3) / 1e9
    # A step is `launches` launches of the fused kernel over equal runs of x
    # planes (plane chunks): algorithmic bytes per launch over the average
    # launch time, which is the per-step figure.
    launches = info["launches"]
    kernel_ms = ms / args.steps / launches
    alg_bytes_launch = bytes_per_cell("f32") * ncell / launches
    achieved = alg_bytes_launch / (kernel_ms * 1e-3) / 1e9

    # KE after warmup + steps against the linear decay of the initial KE
    ke0 = 0.5 * float(np.sum(rho0 * np.sum(u0 * u0, axis=0)))
    t_total = args.warmup + args.steps
    ke_ratio = diag["kinetic_energy"] / ke0
    ke_lin = taylor_green_ke_decay(shape, nu, t_total)

    # HBM bytes per launch from the committed rocprofv3 PMC summary (FETCH_SIZE
    # and WRITE_SIZE passes scaled by the copy kernel, per dispatch); named
    # with its source, used only if it was measured with this launch split
    traffic, traffic_source = None, None
    tf = ROOT / "profiles" / f"hbm_traffic_k_step_f32_tg{n}.json"
    if tf.exists():
        rec = json.loads(tf.read_text())
        if rec.get("launches_per_step", 1) == launches:
            traffic = rec.get("bytes_per_launch")
            traffic_source = {"file": str(tf.relative_to(ROOT)), "summary": rec.get("source"),
                              "measured_at_commit": rec.get("measured_at_commit"),
                              "counters": "rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, scaled by the copy kernel"}

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
            "traffic_source": traffic_source,
        },
        "copy_peak_GBs": round(copy_gbs, 1),
        "frac_of_copy_peak": round(achieved / copy_gbs, 4),
        "kernel": "d3q19::k_step<TF32, BC_PERIODIC>",
        "launches_per_step": launches,
        "algorithmic_bytes_per_cell": bytes_per_cell("f32"),
        "algorithmic_bytes_per_launch": int(alg_bytes_launch),
        "ms_per_launch": round(kernel_ms, 5),
        "wall_s": round(wall, 3),
        "ke_ratio": ke_ratio,
        "ke_ratio_linear_theory": ke_lin,
        "mass_rel_drift": (diag["mass"] - float(np.sum(rho0))) / float(np.sum(rho0)),
        "sclk_MHz": [info["sclk_MHz_before"], info["sclk_MHz_after"]],
    }

    if not args.no_fp16:
        # Two FP16 runs of K steps after W warm-up steps: the first exactly as the
        # FP32 leg (its kinetic energy is compared with FP32's at the same step),
        # the second after `settle_ms` of untimed steps, past the shader-clock dip
        # that follows the host-side set-up (the 16-bit step is VALU-heavier and
        # slows inside the dip; the rate reported is the settled one).
        ms16f, _, diag16, info16f = time_lattice(amd, torch, shape, args.tau, "f16", rho0, u0, args.steps,
                                                 args.warmup, clock=clock)
        ms16, _, _, info16 = time_lattice(amd, torch, shape, args.tau, "f16", rho0, u0, args.steps, args.warmup,
                                          settle_ms=args.settle_ms, clock=clock)
        glups16 = ncell * args.steps / (ms16 * 1e-3) / 1e9
        glups16f = ncell * args.steps / (ms16f * 1e-3) / 1e9
        ach16 = bytes_per_cell("f16") * ncell / (ms16 / args.steps * 1e-3) / 1e9
        result["fp16_store"] = {
            "value": round(glups16, 4),
            "unit": "GLUPS",
            "ratio_vs_fp32": round(glups16 / glups, 4),
            "achieved_GBs": round(ach16, 1),
            "frac": round(ach16 / HBM_PEAK_GBS, 4),
            "settle_ms": args.settle_ms,
            "settle_steps": info16["settle_steps"],
            "sclk_MHz": [info16["sclk_MHz_before"], info16["sclk_MHz_after"]],
            "first_window": {"value": round(glups16f, 4), "ratio_vs_fp32": round(glups16f / glups, 4),
                             "sclk_MHz": [info16f["sclk_MHz_before"], info16f["sclk_MHz_after"]],
                             "note": "K steps after only W warm-up steps, as the FP32 leg"},
            "ke_rel_diff_vs_fp32": abs(diag16["kinetic_energy"] - diag["kinetic_energy"]) / diag["kinetic_energy"],
        }

    if not args.no_me16:
        # Moment codes: ranges calibrated once (FP64 run on the GPU) over the
        # longest horizon either window runs; then, as for FP16, K steps after
        # W warm-up steps (the first window, whose kinetic energy is compared
        # with FP32's at the same step) and K steps after `settle_ms` of
        # untimed steps past the shader-clock dip (the rate reported).
        from lattice_boltzmann_d3q19_amd.moment_encoded import calibrate_moment_ranges

        me_ranges = calibrate_moment_ranges(shape, args.tau, args.warmup + args.steps + ME16_SETTLE_CAP,
                                            rho=rho0, u=u0)
        ms_mef, _, diag_me, info_mef = time_lattice(amd, torch, shape, args.tau, "me16", rho0, u0, args.steps,
                                                    args.warmup, clock=clock, me_ranges=me_ranges)
        ms_me, _, _, info_me = time_lattice(amd, torch, shape, args.tau, "me16", rho0, u0, args.steps, args.warmup,
                                            settle_ms=args.settle_ms, clock=clock, me_ranges=me_ranges)
        glups_me = ncell * args.steps / (ms_me * 1e-3) / 1e9
        glups_mef = ncell * args.steps / (ms_mef * 1e-3) / 1e9
        ach_me = bytes_per_cell("me16") * ncell / (ms_me / args.steps * 1e-3) / 1e9
        result["me16_store"] = {
            "value": round(glups_me, 4),
            "unit": "GLUPS",
            "ratio_vs_fp32": round(glups_me / glups, 4),
            "algorithmic_bytes_per_cell": bytes_per_cell("me16"),
            "achieved_GBs": round(ach_me, 1),
            "frac": round(ach_me / HBM_PEAK_GBS, 4),
            "settle_ms": args.settle_ms,
            "settle_steps": info_me["settle_steps"],
            "sclk_MHz": [info_me["sclk_MHz_before"], info_me["sclk_MHz_after"]],
            "first_window": {"value": round(glups_mef, 4), "ratio_vs_fp32": round(glups_mef / glups, 4),
                             "sclk_MHz": [info_mef["sclk_MHz_before"], info_mef["sclk_MHz_after"]],
                             "note": "K steps after only W warm-up steps, as the FP32 leg"},
            "ke_rel_diff_vs_fp32": abs(diag_me["kinetic_energy"] - diag["kinetic_energy"]) / diag["kinetic_energy"],
        }

    if not args.no_adjoint:
        result["adjoint"] = adjoint_timing(torch)
        result["adjoint_forward_ratio"] = result["adjoint"]["adjoint_forward_ratio"]

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(shape, args.tau, rho0, u0, args.cpu_steps)
        result["cpu_baseline"]["configs"] = config_baselines(amd, torch)
        result["cpu_baseline"]["reference_published_MLUPS"] = {
            "range": [1.1, 4.3], "posture": "single thread, FP64, capture included, i7-12700KF",
            "source": "docs/perf-ledger.md:27-42 of the reference"}

    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

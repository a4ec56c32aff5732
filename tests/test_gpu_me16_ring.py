"""GPU checks of the moment-coded storage's single ring of planes.

An "me16" lattice on the row grid keeps its codes in one ring of nx + gap x
planes: a step writes plane x of the new state into the slot of plane
x - gap of the old one, in the same launch, once the wavefronts pulling from
that plane have consumed their codes (per-plane counters; d3q19_kernels.hpp:
me_ring_sync).  The order in which wavefronts run then must not matter: a
step in one launch -- with the gap the library chooses, or a gap of 2 or 5
planes, where wavefronts do wait for each other -- must give the codes of
the same step launched one plane at a time (strictly ordered), bit for bit,
in every boundary mode; D3Q19_RING_GAP and D3Q19_STEP_CHUNKS (read when a
lattice is created) set the gap and the launches per step.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu

TAU = 0.6


@pytest.fixture(scope="module")
def amd():
    import lattice_boltzmann_d3q19_amd as m

    return m


def created(amd, shape, env, tau=TAU, **kw):
    saved = {k: os.environ.get(k) for k in ("D3Q19_RING_GAP", "D3Q19_STEP_CHUNKS", "D3Q19_RING_FAULT_TEST")}
    try:
        for k in saved:
            os.environ.pop(k, None)
        os.environ.update(env)
        return amd.Lattice(shape, tau, storage="me16", **kw)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


CASES = {
    "periodic": dict(),
    "forced-ywalls": dict(boundary="ywalls", force=(1e-5, 0.0, 2e-6), wall_velocity_top=(0.03, 0.0, 0.0)),
    "cavity": dict(boundary="halfway", wall_faces=("-x", "+x", "-y", "+y", "-z", "+z"),
                   face_velocity={"+z": (0.05, 0.0, 0.0)}),
}


@pytest.fixture(scope="module")
def initial():
    from lattice_boltzmann_d3q19_amd.initial import random_solenoidal

    shape = (96, 128, 128)
    return shape, random_solenoidal(shape, 0.02, seed=5)


@pytest.mark.parametrize("case", list(CASES))
def test_ring_step_in_one_launch_matches_plane_by_plane(amd, initial, case):
    from lattice_boltzmann_d3q19_amd.moment_encoded import calibrate_moment_ranges

    shape, u0 = initial
    steps = 12
    kw = CASES[case]
    lo, hi = calibrate_moment_ranges(shape, TAU, steps, u=u0, **kw)
    runs = {}
    for name, env in (("planes", {"D3Q19_STEP_CHUNKS": str(shape[0])}), ("auto", {}),
                      ("gap2", {"D3Q19_RING_GAP": "2"}), ("gap5", {"D3Q19_RING_GAP": "5"})):
        print_case("me16-ring", case=case, run=name, shape=shape, steps=steps)
        lat = created(amd, shape, env, **kw)
        lat.set_moment_ranges(lo, hi).set_equilibrium(None, u0).step(steps)
        runs[name] = (lat.raw_state()[0].copy(), lat.diagnostics(), lat.device_bytes)
        lat.close()
    ref_state, ref_diag, _ = runs["planes"]
    for name in ("auto", "gap2", "gap5"):
        state, diag, _ = runs[name]
        assert np.array_equal(state, ref_state), name
        assert diag == ref_diag, name
    # a tighter gap holds fewer planes
    assert runs["gap2"][2] < runs["gap5"][2] < runs["auto"][2]


RAGGED = {
    # nz = 100: rows padded to 128, blocks of 128 z x 4 rows; ny = 37: the last block has one live row
    "padded-rows-odd-ny": (40, 37, 100),
    # nx = 3: the automatic gap covers the whole lattice (no waits); gap 2 still writes in place
    "three-planes": (3, 20, 64),
    # nz = 192: three z segments of 64 per row, blocks of 8 rows of which 5 are live
    "short-blocks": (17, 5, 192),
}


@pytest.mark.parametrize("bc", ["forced", "cavity"])
@pytest.mark.parametrize("name", list(RAGGED))
def test_ring_on_ragged_lattices(amd, name, bc):
    """Partly live blocks (padded rows, ny not a multiple of the block's
    rows), a gap as large as the lattice, and plane chunks over a tight gap:
    the ring step equals the plane-by-plane step bit for bit."""
    from lattice_boltzmann_d3q19_amd.initial import random_solenoidal
    from lattice_boltzmann_d3q19_amd.moment_encoded import calibrate_moment_ranges

    shape = RAGGED[name]
    u0 = random_solenoidal(shape, 0.02, seed=11)
    kw = (dict(force=(2e-6, -1e-6, 3e-6)) if bc == "forced" else
          dict(boundary="halfway", wall_faces=("-x", "+x", "-y", "+y", "-z", "+z"),
               face_velocity={"+z": (0.05, 0.0, 0.0)}))
    steps = 9
    lo, hi = calibrate_moment_ranges(shape, TAU, steps, u=u0, **kw)
    runs = {}
    for run, env in (("planes", {"D3Q19_STEP_CHUNKS": str(shape[0])}), ("auto", {}), ("gap2", {"D3Q19_RING_GAP": "2"}),
                     ("gap2-chunks3", {"D3Q19_RING_GAP": "2", "D3Q19_STEP_CHUNKS": "3"})):
        print_case("me16-ring-ragged", shape=shape, bc=bc, run=run, steps=steps)
        lat = created(amd, shape, env, **kw)
        lat.set_moment_ranges(lo, hi).set_equilibrium(None, u0).step(steps)
        runs[run] = (lat.raw_state()[0].copy(), lat.populations(), lat.diagnostics())
        lat.close()
    ref = runs["planes"]
    assert np.isfinite(ref[1]).all()
    for run in ("auto", "gap2", "gap2-chunks3"):
        assert np.array_equal(runs[run][0], ref[0]), run
        assert np.array_equal(runs[run][1], ref[1]), run
        assert runs[run][2] == ref[2], run


def _sweep_cases(n=8, seed=20261204):
    """Seeded row-grid lattices: nx 3-40, ny 3-40, nz 64-300 (padded rows, 64- to
    256-cell z blocks), a boundary mode, a gap of 2 up to the automatic one,
    and 1-4 plane chunks."""
    rng = np.random.default_rng(seed)
    bcs = ("periodic", "forced", "forced-ywalls", "cavity")
    out = []
    for k in range(n):
        shape = (int(rng.integers(3, 41)), int(rng.integers(3, 41)), int(rng.integers(64, 301)))
        out.append((shape, bcs[k % len(bcs)], int(rng.integers(2, 9)), int(rng.integers(1, 5))))
    return out


SWEEP_KW = {
    "periodic": dict(),
    "forced": dict(force=(2e-6, 1e-6, -3e-6)),
    "forced-ywalls": dict(boundary="ywalls", force=(1e-5, 0.0, 0.0), wall_velocity_top=(0.02, 0.0, 0.01)),
    "cavity": dict(boundary="halfway", wall_faces=("-x", "+x", "-y", "+y", "-z", "+z"),
                   face_velocity={"+z": (0.05, 0.0, 0.0)}),
}


@pytest.mark.parametrize("shape,bc,gap,chunks", _sweep_cases())
def test_ring_on_seeded_lattices(amd, shape, bc, gap, chunks):
    """The ring step, at the automatic gap and at a seeded tighter gap with
    seeded plane chunks, equals the plane-by-plane step bit for bit."""
    from lattice_boltzmann_d3q19_amd.initial import random_solenoidal
    from lattice_boltzmann_d3q19_amd.moment_encoded import calibrate_moment_ranges

    kw = SWEEP_KW[bc]
    u0 = random_solenoidal(shape, 0.02, seed=shape[0] * 7 + shape[2])
    steps = 7
    lo, hi = calibrate_moment_ranges(shape, TAU, steps, u=u0, **kw)
    runs = {}
    for run, env in (("planes", {"D3Q19_STEP_CHUNKS": str(shape[0])}), ("auto", {}),
                     ("seeded", {"D3Q19_RING_GAP": str(gap), "D3Q19_STEP_CHUNKS": str(chunks)})):
        print_case("me16-ring-sweep", shape=shape, bc=bc, run=run, gap=gap, chunks=chunks)
        lat = created(amd, shape, env, **kw)
        lat.set_moment_ranges(lo, hi).set_equilibrium(None, u0).step(steps)
        runs[run] = (lat.raw_state()[0].copy(), lat.diagnostics())
        lat.close()
    for run in ("auto", "seeded"):
        assert np.array_equal(runs[run][0], runs["planes"][0]), run
        assert runs[run][1] == runs["planes"][1], run


def test_ring_at_the_benchmarked_size(amd):
    """256^3 Taylor-Green, the moment-coded leg's workload: the step in one
    launch with the automatic gap (18 planes, every CU's blocks resident)
    equals the plane-by-plane step bit for bit after 10 steps."""
    from lattice_boltzmann_d3q19_amd.initial import taylor_green
    from lattice_boltzmann_d3q19_amd.moment_encoded import calibrate_moment_ranges

    shape, tau, steps = (256, 256, 256), 0.55, 10
    rho0, u0 = taylor_green(shape, 0.04)
    lo, hi = calibrate_moment_ranges(shape, tau, steps, rho=rho0, u=u0)
    states = {}
    for run, env in (("planes", {"D3Q19_STEP_CHUNKS": "256"}), ("auto", {})):
        print_case("me16-ring-256", run=run, steps=steps)
        lat = created(amd, shape, env, tau=tau)
        lat.set_moment_ranges(lo, hi).set_equilibrium(rho0, u0).step(steps)
        assert lat.step_launches == (256 if run == "planes" else 1)
        states[run] = (lat.raw_state()[0], lat.diagnostics())
        lat.close()
    assert np.array_equal(states["auto"][0], states["planes"][0])
    assert states["auto"][1] == states["planes"][1]


def test_ring_holds_about_half_the_memory_of_fp16(amd):
    """256^3: the codes of one state and a gap of a few planes, against two
    FP16 lattices (38 B per cell each)."""
    shape = (256, 256, 256)
    print_case("me16-ring-memory", shape=shape)
    me = created(amd, shape, {})
    f16 = amd.Lattice(shape, TAU, storage="f16")
    ratio = me.device_bytes / f16.device_bytes
    per_cell = me.device_bytes / np.prod(shape)
    me.close()
    f16.close()
    print(f"me16 ring: {per_cell:.1f} B per cell, {ratio:.3f} of FP16", flush=True)
    assert ratio < 0.56
    assert per_cell < 42.0


def test_a_wait_that_gives_up_is_reported(amd, initial, tmp_path):
    """The waits' failure path, forced from the host: with
    D3Q19_RING_FAULT_TEST=1 every wait polls, once, a count that no block
    reaches and gives up, so every kernel still ends; the next synchronising
    call reports D3Q19_EHIP, and keeps reporting it until a state is loaded:
    a checkpoint restore or a new equilibrium makes the lattice readable again
    (its first step collides in place, without waits)."""
    from lattice_boltzmann_d3q19_amd._lib import D3Q19_EHIP, D3Q19Error
    from lattice_boltzmann_d3q19_amd.moment_encoded import calibrate_moment_ranges

    shape, u0 = initial
    print_case("me16-ring-wait-gives-up", shape=shape)
    lo, hi = calibrate_moment_ranges(shape, TAU, 4, u=u0)
    lat = created(amd, shape, {"D3Q19_RING_FAULT_TEST": "1"})
    lat.set_moment_ranges(lo, hi).set_equilibrium(None, u0)
    lat.step(1)
    good = lat.populations()  # the first step after a load collides in place: no waits
    ck = lat.checkpoint(tmp_path / "ring.npz")
    lat.step(2)
    for read in (lat.diagnostics, lat.populations):
        with pytest.raises(D3Q19Error) as e:
            read()
        assert e.value.status == D3Q19_EHIP
        assert "waited too long" in str(e.value)
    lat.restore(ck)
    assert np.array_equal(lat.populations(), good)
    lat.step(1)  # a fused step from the restored state: its waits give up again
    with pytest.raises(D3Q19Error):
        lat.diagnostics()
    lat.set_equilibrium(None, u0)
    lat.step(1).diagnostics()
    lat.close()

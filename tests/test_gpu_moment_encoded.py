"""GPU parity of the 16-bit moment-encoded storages.

"me16" holds the codes of the stored state (post-collision populations of
the last step; f^0 after a load) and steps them in one fused FP32 pass.  Its
FP64 model, composed from pinned oracle pieces, is
    f~^0 = Q_0(f^0),  f~^{n+1} = walls(stream(Q_n(collide(f~^n)))),
Q_n = decode o encode on the engine's grid (deviation moments, power-of-two
steps) with the stochastic rounding bits of step n, restated below from
their definition (a hash of cell and step).

* codec: a loaded state decodes to the codes the oracle's FP64 encoder picks
  (re-encoded: equal, or one code apart where FP32 and FP64 fall on opposite
  sides of a half-code boundary -- at most 2% of them); set_equilibrium
  stores the codes of loading feq_field's populations;
* moment ranges: the GPU calibration (FP64 lattice, per-frame moment range)
  equals the oracle's over the same trajectory, of f^n and of the stored
  states;
* trajectory: the quantized Poiseuille (and a periodic lattice with a force
  field) stays within the declared lbm-quantized tolerance of the FP64
  parent (err <= 1e-5 + 5e-2 |parent|, equivalence/tolerance.toml
  [defaults.lbm-quantized]) on every checked frame, and within three times
  the FP64 model's own distance from the parent of that model;
* the canonical moment-encoded capture (1001 frames, "me16ref") against the
  parent frame by frame, its measured max relative error over |u| > 1e-6
  near the published 1.98e-2, and the Poiseuille structure checks of
  ME/tests/python/test_frontier_equivalence.py:95-121;
* "me16ref" (the variant's own FP64 arithmetic: decode, division-form
  collision, stream + y-walls, encode): codes, decoded populations and
  moments equal the oracle's quantized trajectory bit for bit (y-wall channel
  with a uniform force, periodic lattice with a force field); set_equilibrium,
  checkpoint / restore, and the refused operations.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import print_case

pytestmark = pytest.mark.gpu

TAU, FX = 0.7, 1e-5
CANON = (64, 32, 3)


@pytest.fixture(scope="module")
def amd():
    import lattice_boltzmann_d3q19_amd as m

    return m


def oracle_ranges(frames, margin=0.25):
    lo = np.full(19, np.inf)
    hi = np.full(19, -np.inf)
    for f in frames:
        m = oracle.me_moments(f).reshape(19, -1)
        lo = np.minimum(lo, m.min(1))
        hi = np.maximum(hi, m.max(1))
    span = np.maximum(hi - lo, 1e-12)
    return lo - margin * span, hi + margin * span


def stored_frames(frames, F, tau=TAU):
    """The states an "me16" lattice stores along a run: f^0 after the load,
    then collide(f^n) after step n + 1."""
    return [frames[0]] + [oracle.collide(f, tau, F) for f in frames[:-1]]


def _mix32(h):
    h = h ^ (h >> np.uint32(16))
    h = h * np.uint32(0x7FEB352D)
    h = h ^ (h >> np.uint32(15))
    h = h * np.uint32(0x846CA68B)
    return h ^ (h >> np.uint32(16))


class Me16Grid:
    """The "me16" codec in FP64: deviation moments m' = M (f - w) on the grid
    lo'_k + code step_k, lo'_k = fp32(lo_k - (M w)_k), step_k the power of two
    at or above (hi_k - lo_k) / 65535; encode floor((m' - lo') / step + r_k),
    r_k the top 16 bits of key * (0x9e3779b1 (2k + 1)) over 2^16, key =
    mix32(cell ^ (n 0x9e3779b9 + 0x7f4a7c15)) for the encode of step n."""

    def __init__(self, lo, hi):
        m, minv, _ = oracle.me_basis()
        _, w, _, _ = oracle.constants()
        mw = np.array([sum(m[k, i] * w[i] for i in range(19)) for k in range(19)])
        self.m, self.minv, self.w = m, minv, w
        self.step = 2.0 ** np.ceil(np.log2((np.asarray(hi) - np.asarray(lo)) / 65535.0))
        self.lod = (np.asarray(lo) - mw).astype(np.float32).astype(np.float64)

    def encode(self, f, n):
        shape = f.shape[1:]
        d = (f - self.w.reshape((19,) + (1,) * len(shape))).reshape(19, -1)
        mp = self.m @ d
        cell = np.arange(d.shape[1], dtype=np.uint32)
        key = _mix32(cell ^ np.uint32((n * 0x9E3779B9 + 0x7F4A7C15) & 0xFFFFFFFF))
        k = np.arange(19, dtype=np.uint32)[:, None]
        r = ((key[None, :] * (np.uint32(0x9E3779B1) * (2 * k + 1)).astype(np.uint32)) >> np.uint32(16)) / 65536.0
        t = (mp - self.lod[:, None]) / self.step[:, None] + r
        return np.clip(np.floor(t), 0, 65535).astype(np.uint16).reshape((19,) + shape)

    def decode(self, codes):
        shape = codes.shape[1:]
        mp = self.lod[:, None] + codes.reshape(19, -1).astype(np.float64) * self.step[:, None]
        return (self.w[:, None] + self.minv @ mp).reshape((19,) + shape)


def me16_model(f0, steps, lo, hi, F, ywalls, tau=TAU):
    """FP64 frames f~^0 .. f~^steps of the "me16" quantization point."""
    g = Me16Grid(lo, hi)
    out = [g.decode(g.encode(f0, 0))]
    for n in range(steps):
        h = oracle.stream(g.decode(g.encode(oracle.collide(out[-1], tau, F), n)))
        out.append(oracle.apply_bounce_back_y_walls(h) if ywalls else h)
    return out


def parent_frames(steps, shape=CANON):
    F = np.zeros((3,) + shape)
    F[0] = FX
    frames = [oracle.feq_field(np.ones(shape), np.zeros((3,) + shape))]
    for _ in range(steps):
        frames.append(oracle.run_reference(frames[-1], TAU, 1, force_lattice=F, ywalls=True))
    return frames, F


def stored_codes(lat, shape):
    """The codes of an "me16" lattice as (19, Nx, Ny, Nz): moment-major for
    rows shorter than 64 cells, else row-blocked (19 segments of the padded
    row length per row x * Ny + y)."""
    raw = lat.raw_state()[0].view(np.uint16)
    nx, ny, nz = shape
    if nz < 64:
        return raw.reshape((19,) + shape)
    nzp = raw.size // (nx * ny * 19)
    return raw.reshape(nx, ny, 19, nzp)[..., :nz].transpose(2, 0, 1, 3)


@pytest.mark.parametrize("shape", [(6, 5, 4), (4, 3, 64), (3, 2, 130)], ids=lambda s: "x".join(map(str, s)))
def test_codec_matches_model(amd, shape):
    """A loaded state: every stored code is the FP64 model's (one apart where
    FP32 and FP64 fall on opposite sides of a code boundary, at most 1% of
    them), and the read-back is the FP64 decode of the stored codes to FP32
    rounding."""
    print_case("me16-codec", shape=shape)
    rng = np.random.default_rng(sum(shape))
    f = oracle.feq_field(1 + 0.01 * rng.standard_normal(shape), 0.02 * rng.standard_normal((3,) + shape))
    f *= 1 + 0.02 * rng.standard_normal(f.shape)
    lo, hi = oracle_ranges([f], margin=0.1)
    g = Me16Grid(lo, hi)
    lat = amd.Lattice(shape, TAU, storage="me16").set_moment_ranges(lo, hi)
    lat.set_populations(f)
    codes = stored_codes(lat, shape)
    diff = np.abs(codes.astype(np.int64) - g.encode(f, 0).astype(np.int64))
    assert diff.max() <= 1
    assert np.count_nonzero(diff) <= 0.01 * diff.size
    back = lat.populations()
    assert np.abs(back - g.decode(codes)).max() < 1e-6 * np.abs(f).max()
    lat.close()


def test_gpu_calibration_equals_oracle(amd):
    from lattice_boltzmann_d3q19_amd.moment_encoded import calibrate_moment_ranges

    print_case("me16-calibration")
    frames, F = parent_frames(100)
    kw = dict(warmup_steps=0, boundary="ywalls", force=(FX, 0.0, 0.0))
    for reference, states in ((True, frames), (False, stored_frames(frames, F))):
        lo0, hi0 = oracle_ranges(states)
        lo, hi = calibrate_moment_ranges(CANON, TAU, 100, reference_arithmetic=reference, **kw)
        # FP64 rounding: the engine sums w + (f - w), the oracle f (~40 ulp of
        # 1.0 measured on the density moment), and the engine's FP64 collision
        # is not bgk.py's operation order; 5e-14 is 1% of the narrowest
        # padded range
        span = hi0 - lo0
        tol = 1e-9 * span + 5e-14
        assert np.all(np.abs(lo - lo0) <= tol) and np.all(np.abs(hi - hi0) <= tol), reference
    # tau = 0.7 < 1 over-relaxes: the stored (post-collision) non-conserved
    # moments leave f^n's ranges
    lo_f, hi_f = oracle_ranges(frames, margin=0.0)
    lo_s, hi_s = oracle_ranges(stored_frames(frames, F), margin=0.0)
    assert np.any(lo_s < lo_f - 1e-9) or np.any(hi_s > hi_f + 1e-9)


@pytest.mark.parametrize("storage", ["f32", "f16", "f64"])
@pytest.mark.parametrize("shape", [(5, 4, 3), (3, 4, 64)], ids=lambda s: "x".join(map(str, s)))
def test_moment_range_any_storage(amd, storage, shape):
    print_case("moment-range", storage=storage, shape=shape)
    rng = np.random.default_rng(5)
    f = oracle.feq_field(1 + 0.01 * rng.standard_normal(shape), 0.02 * rng.standard_normal((3,) + shape))
    lat = amd.Lattice(shape, TAU, storage=storage).set_populations(f)
    lat.step(3)
    ref = oracle.me_moments(lat.populations()).reshape(19, -1)
    lo, hi = lat.moment_range()
    tol = {"f64": 1e-13, "f32": 1e-6, "f16": 1e-6}[storage]
    assert np.allclose(lo, ref.min(1), rtol=0, atol=tol) and np.allclose(hi, ref.max(1), rtol=0, atol=tol)


def _track(lat, frames, model, F, every):
    for n in range(len(frames)):
        if n:
            lat.step(1)
        if n % every and n != len(frames) - 1:
            continue
        _, u = lat.moments()
        up = oracle.macroscopic_velocity(frames[n], F)
        uq = oracle.macroscopic_velocity(model[n], F)
        assert np.all(np.abs(u - up) <= 1e-5 + 5e-2 * np.abs(up)), n
        # the FP32 engine stays within a few quantization errors of its FP64 model
        print(f"frame {n}: |u-uq| {np.abs(u - uq).max():.3e} |uq-up| {np.abs(uq - up).max():.3e}", flush=True)
        assert np.abs(u - uq).max() <= 3 * max(np.abs(uq - up).max(), 1e-9), n


def test_trajectory_tracks_parent_and_fp64_model(amd):
    print_case("me16-trajectory")
    steps = 200
    frames, F = parent_frames(steps)
    lo, hi = oracle_ranges(stored_frames(frames, F))
    lat = amd.Lattice(CANON, TAU, storage="me16", boundary="ywalls", force=(FX, 0.0, 0.0))
    lat.set_moment_ranges(lo, hi).set_equilibrium()
    _track(lat, frames, me16_model(frames[0], steps, lo, hi, F, True), F, 20)
    lat.close()


def test_periodic_force_field_tracks_parent_and_fp64_model(amd):
    print_case("me16-periodic-force-field")
    shape, steps, tau = (12, 10, 8), 60, 0.8
    rng = np.random.default_rng(17)
    f0 = oracle.feq_field(1.0 + 0.01 * rng.standard_normal(shape), 0.02 * rng.standard_normal((3,) + shape))
    F = 1e-4 * rng.standard_normal((3,) + shape)
    frames = [f0]
    for _ in range(steps):
        frames.append(oracle.run_reference(frames[-1], tau, 1, force_lattice=F))
    lo, hi = oracle_ranges(stored_frames(frames, F, tau))
    lat = amd.Lattice(shape, tau, storage="me16", force=F).set_moment_ranges(lo, hi).set_populations(f0)
    model = me16_model(f0, steps, lo, hi, F, False, tau)
    # the loaded state (rows shorter than 64 cells: moment-major codes):
    # every stored code is the model's, or one apart where FP32 and FP64 fall
    # on opposite sides of a code boundary
    raw, local, _ = lat.raw_state()
    assert local
    diff = np.abs(raw.view(np.uint16).reshape((19,) + shape).astype(np.int64)
                  - Me16Grid(lo, hi).encode(f0, 0).astype(np.int64))
    assert diff.max() <= 1 and np.count_nonzero(diff) <= 0.01 * diff.size
    _track(lat, frames, model, F, 15)
    lat.close()


def test_set_equilibrium_stores_the_codes_of_its_populations(amd):
    print_case("me16-set-equilibrium")
    shape = (9, 7, 66)
    rng = np.random.default_rng(5)
    rho = 1.0 + 0.01 * rng.standard_normal(shape)
    u = 0.03 * rng.standard_normal((3,) + shape)
    f = oracle.feq_field(rho, u)
    lo, hi = oracle_ranges([f], margin=0.1)
    a = amd.Lattice(shape, TAU, storage="me16").set_moment_ranges(lo, hi).set_equilibrium(rho, u)
    b = amd.Lattice(shape, TAU, storage="me16").set_moment_ranges(lo, hi).set_populations(f)
    assert np.array_equal(a.raw_state()[0], b.raw_state()[0])
    a.step(3)
    b.step(3)
    assert np.array_equal(a.populations(), b.populations())
    a.close()
    b.close()


def test_canonical_capture_within_lbm_quantized(amd, tmp_path):
    from lattice_boltzmann_d3q19_amd.capture import load_capture
    from lattice_boltzmann_d3q19_amd.sim import sim_runner_moment_encoded

    print_case("me16-canonical-capture")
    cap = load_capture(sim_runner_moment_encoded(42, tmp_path, payload_format="npz"))
    assert cap.manifest.sim["variant"] == "frontier-moment-encoded-16bit"
    frames, F = parent_frames(1000)
    max_rel = 0.0
    for n in range(1001):
        st = cap.step(n)
        rp = oracle.density_field(frames[n])
        up = oracle.macroscopic_velocity(frames[n], F)
        e = np.abs(st.state["u"] - up)
        assert np.all(e <= 1e-5 + 5e-2 * np.abs(up)), n
        assert np.all(np.abs(st.state["rho"] - rp) <= 1e-5 + 5e-2 * np.abs(rp)), n
        mask = np.abs(up) > 1e-6
        if mask.any():
            max_rel = max(max_rel, float((e[mask] / np.abs(up[mask])).max()))
        assert st.diagnostics["u_max_lat"] == float(np.abs(st.state["u"]).max())
    assert 1.0e-2 < max_rel < 3.0e-2
    u = cap.step(1000).state["u"]
    prof = u[0].mean(axis=(0, 2))
    ny = prof.shape[0]
    assert int(np.argmax(prof)) in {(ny - 1) // 2, ny // 2}
    assert prof[0] < 0.5 * prof[ny // 2] and prof[-1] < 0.5 * prof[ny // 2]
    assert np.abs(prof - prof[::-1]).max() <= 5e-2 * prof[ny // 2]


def test_me16_errors(amd):
    from lattice_boltzmann_d3q19_amd import D3Q19Error

    print_case("me16-errors")
    lo, hi = np.zeros(19), np.ones(19)
    lat = amd.Lattice((4, 4, 4), TAU, storage="me16")
    with pytest.raises(D3Q19Error, match="ESTATE"):
        lat.set_equilibrium()
    with pytest.raises(D3Q19Error, match="EINVAL"):
        lat.set_moment_ranges(hi, lo)
    lat.set_moment_ranges(lo - 1, hi).set_equilibrium()
    with pytest.raises(D3Q19Error, match="ESTATE"):
        lat.set_moment_ranges(lo - 1, hi)
    with pytest.raises(D3Q19Error, match="EINVAL"):
        amd.Lattice((4, 4, 4), TAU, storage="f32").set_moment_ranges(lo, hi)


def _me16ref_case(case):
    """(shape, f0, F (3, ...) or None, uniform force, ywalls, steps, ranges)."""
    if case == "poiseuille-ywalls":
        frames, F = parent_frames(200)
        return CANON, frames[0], F, (FX, 0.0, 0.0), True, 200, oracle_ranges(frames)
    shape = (12, 10, 8)
    rng = np.random.default_rng(7)
    f0 = oracle.feq_field(1.0 + 0.01 * rng.standard_normal(shape), 0.02 * rng.standard_normal((3,) + shape))
    F = 1e-4 * rng.standard_normal((3,) + shape)
    frames = [f0]
    for _ in range(60):
        frames.append(oracle.run_reference(frames[-1], TAU, 1, force_lattice=F))
    return shape, f0, F, None, False, 60, oracle_ranges(frames)


@pytest.mark.parametrize("case", ["poiseuille-ywalls", "periodic-force-field"])
def test_me16ref_equals_oracle_quantized_trajectory_bit_for_bit(amd, case):
    """'me16ref' steps the codes in the moment-encoded variant's FP64
    arithmetic (ME/src/lbm_me.cpp:386-409): on every checked frame the stored
    codes equal the oracle's quantized trajectory's codes, and the decoded
    populations and moments equal the oracle's to the bit.

    The oracle continues from its own codes between checked frames: decode
    then encode is not the identity for a moment whose range is narrower than
    FP64 resolves around its value (the channel's density, range about 1e-12
    wide around 1), so a trajectory restarted from decoded populations is a
    different trajectory."""
    print_case("me16ref-bit-exact", case=case)
    shape, f0, F, uniform, ywalls, steps, (lo, hi) = _me16ref_case(case)
    kw = dict(boundary="ywalls" if ywalls else "periodic")
    kw["force"] = uniform if uniform is not None else F
    lat = amd.Lattice(shape, TAU, storage="me16ref", **kw).set_moment_ranges(lo, hi).set_populations(f0)
    done = 0
    codes = oracle.me_encode(f0, lo, hi)
    for n in sorted(set(range(0, steps + 1, 20)) | {steps}):
        lat.step(n - done)
        q, codes = oracle.run_me(None, TAU, n - done, lo, hi, force_lattice=F, ywalls=ywalls, codes=codes,
                                 return_codes=True)
        done = n
        raw, local, count = lat.raw_state()
        assert local and count == n
        assert np.array_equal(raw.view(np.uint16).reshape(codes.shape), codes), n
        got = lat.populations()
        assert np.array_equal(got, q), (n, np.abs(got - q).max())
        rho, u = lat.moments()
        assert np.array_equal(rho, oracle.density_field(q)), n
        assert np.array_equal(u, oracle.macroscopic_velocity(q, F)), n
    lat.close()


def test_me16ref_equilibrium_checkpoint_and_refusals(amd, tmp_path):
    from lattice_boltzmann_d3q19_amd import D3Q19Error

    print_case("me16ref-equilibrium-refusals")
    shape = (8, 6, 4)
    frames, _ = parent_frames(20, shape)
    lo, hi = oracle_ranges(frames)
    rng = np.random.default_rng(3)
    rho = 1.0 + 0.01 * rng.standard_normal(shape)
    u = 0.02 * rng.standard_normal((3,) + shape)
    lat = amd.Lattice(shape, TAU, storage="me16ref")
    with pytest.raises(D3Q19Error, match="ESTATE"):
        lat.set_equilibrium(rho, u)
    lat.set_moment_ranges(lo, hi).set_equilibrium(rho, u)
    want = oracle.me_decode(oracle.me_encode(oracle.feq_field(rho, u), lo, hi), lo, hi)
    assert np.array_equal(lat.populations(), want)
    # a checkpoint of the codes continues bit for bit
    lat.step(5)
    ck = lat.checkpoint(tmp_path / "me16ref.npz")
    lat.step(5)
    other = amd.Lattice(shape, TAU, storage="me16ref")
    other.restore(ck)
    other.step(5)
    assert np.array_equal(other.populations(), lat.populations())
    other.close()
    for call in (lat.diagnostics, lat.moment_range, lat.enstrophy):
        with pytest.raises(D3Q19Error, match="EINVAL"):
            call()
    with pytest.raises(D3Q19Error, match="EINVAL"):
        lat.set_brinkman(np.zeros(shape))
    lat.close()
    with pytest.raises(D3Q19Error, match="EINVAL"):
        amd.Lattice(shape, TAU, storage="me16ref", boundary="halfway")

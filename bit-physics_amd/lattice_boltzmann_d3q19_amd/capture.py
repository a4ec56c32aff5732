"""Capture files of the canonical runs: manifest JSON + per-step payload.

Mirrors the reference testkit's capture-v1 contract
(tools/testkit/capture/writer.py:36-92, reader.py:22-98, manifest.py:30-73,
diff.py:17-110, tools/testkit/determinism/harness.py:97-140) so that output of
this engine can be diffed against the reference's own captures:

* the manifest is the capture-v1 JSON object (schema_version, sim, stack,
  config, run, payload, determinism), checked here by :func:`validate_manifest`
  against the rules of tools/testkit/schemas/capture-v1.json;
* the payload holds ``steps/{n}/state/{field}``, ``steps/{n}/diagnostics/{name}``
  and the ``metadata`` attributes.  It is HDF5 (written exactly as the
  reference writes it) when h5py is importable, else an ``.npz`` archive with
  the same keys -- ``payload.format`` says which.  The reference's readers
  need the HDF5 form.

No pickling anywhere: NPZ payloads are written and read with numpy's
``allow_pickle=False``.
"""

from __future__ import annotations

import hashlib
import json
import re
from dataclasses import asdict, dataclass, field
from pathlib import Path
from typing import Any, Callable, Iterable

import numpy as np

SCHEMA_VERSION = "1.0.0"
MAX_SUPPORTED_VERSION = "1.1.0"
FORMATS = ("hdf5", "npz")

# Cross-stack tolerance of the lbm category
# (tools/testkit/equivalence/tolerance.toml:38-40,146-152): relative 1e-5 of the
# field's max |value|, absolute 0.
LBM_RELATIVE_TOLERANCE = 1e-5


def _h5py():
    try:
        import h5py  # noqa: F401
    except ImportError:
        return None
    return h5py


@dataclass
class StepState:
    step: int
    state: dict = field(default_factory=dict)
    diagnostics: dict = field(default_factory=dict)


@dataclass
class CaptureManifest:
    schema_version: str
    sim: dict
    stack: dict
    config: dict
    run: dict
    payload: dict
    determinism: dict
    gradient_fields: list | None = None
    active_mask: dict | None = None

    def to_dict(self) -> dict:
        d = asdict(self)
        for k in ("gradient_fields", "active_mask"):
            if d.get(k) is None:
                d.pop(k, None)
        return d

    @classmethod
    def from_dict(cls, data: dict) -> "CaptureManifest":
        validate_manifest(data)
        return cls(**{k: data.get(k) for k in (
            "schema_version", "sim", "stack", "config", "run", "payload", "determinism",
            "gradient_fields", "active_mask")})


class ManifestError(ValueError):
    pass


def _require(obj, keys, where, allowed=None):
    if not isinstance(obj, dict):
        raise ManifestError(f"{where}: expected an object")
    missing = [k for k in keys if k not in obj]
    if missing:
        raise ManifestError(f"{where}: missing {missing}")
    extra = set(obj) - set(allowed if allowed is not None else keys)
    if extra:
        raise ManifestError(f"{where}: unexpected {sorted(extra)}")


def _nonempty_str(v, where):
    if not isinstance(v, str) or not v:
        raise ManifestError(f"{where}: expected a non-empty string")


def _is_int(v):
    return isinstance(v, int) and not isinstance(v, bool)


def validate_manifest(m: dict) -> None:
    """Check a manifest dict against the capture-v1 rules; raise ManifestError.

    Same required keys, closed objects, types, enums and patterns as
    tools/testkit/schemas/capture-v1.json, except that ``payload.format`` also
    admits ``"npz"`` (the payload written when h5py is absent).
    """
    top = ["schema_version", "sim", "stack", "config", "run", "payload", "determinism"]
    _require(m, top, "manifest", top + ["gradient_fields", "active_mask"])
    if not isinstance(m["schema_version"], str) or not re.fullmatch(r"\d+\.\d+\.\d+", m["schema_version"]):
        raise ManifestError("schema_version: expected MAJOR.MINOR.PATCH")
    _require(m["sim"], ["name", "category", "variant"], "sim")
    for k in ("name", "category", "variant"):
        _nonempty_str(m["sim"][k], f"sim.{k}")
    _require(m["stack"], ["name", "version", "build_id"], "stack")
    for k in ("name", "version", "build_id"):
        _nonempty_str(m["stack"][k], f"stack.{k}")
    c = m["config"]
    _require(c, ["tier", "dims", "dtype", "seed", "params"], "config")
    _nonempty_str(c["tier"], "config.tier")
    if not isinstance(c["dims"], list) or not c["dims"] or not all(_is_int(d) and d >= 1 for d in c["dims"]):
        raise ManifestError("config.dims: expected a non-empty list of positive integers")
    if c["dtype"] not in ("f32", "f64"):
        raise ManifestError("config.dtype: expected 'f32' or 'f64'")
    if not _is_int(c["seed"]):
        raise ManifestError("config.seed: expected an integer")
    if not isinstance(c["params"], dict):
        raise ManifestError("config.params: expected an object")
    r = m["run"]
    _require(r, ["step_count", "capture_interval", "wall_clock_seconds", "start_utc"], "run")
    if not _is_int(r["step_count"]) or r["step_count"] < 0:
        raise ManifestError("run.step_count: expected an integer >= 0")
    if not _is_int(r["capture_interval"]) or r["capture_interval"] < 1:
        raise ManifestError("run.capture_interval: expected an integer >= 1")
    if isinstance(r["wall_clock_seconds"], bool) or not isinstance(r["wall_clock_seconds"], (int, float)) \
            or r["wall_clock_seconds"] < 0:
        raise ManifestError("run.wall_clock_seconds: expected a number >= 0")
    _nonempty_str(r["start_utc"], "run.start_utc")
    p = m["payload"]
    _require(p, ["format", "path", "checksum"], "payload")
    if p["format"] not in FORMATS:
        raise ManifestError(f"payload.format: expected one of {FORMATS}")
    _nonempty_str(p["path"], "payload.path")
    if not isinstance(p["checksum"], str) or not re.fullmatch(r"sha256:[0-9a-f]{64}", p["checksum"]):
        raise ManifestError("payload.checksum: expected sha256:<64 hex>")
    d = m["determinism"]
    _require(d, ["claimed", "atomic_ops", "subgroup_ops"], "determinism")
    if d["claimed"] not in ("bit-exact-same-hw", "epsilon", "non-deterministic"):
        raise ManifestError("determinism.claimed: unknown value")
    if not isinstance(d["atomic_ops"], bool) or not isinstance(d["subgroup_ops"], bool):
        raise ManifestError("determinism.atomic_ops/subgroup_ops: expected booleans")


def _sha256(path: Path) -> str:
    h = hashlib.sha256()
    with path.open("rb") as fh:
        for chunk in iter(lambda: fh.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


def _metadata(manifest: CaptureManifest) -> dict:
    return {
        "schema_version": manifest.schema_version,
        "sim_name": manifest.sim.get("name", ""),
        "sim_category": manifest.sim.get("category", ""),
        "sim_variant": manifest.sim.get("variant", ""),
        "stack_name": manifest.stack.get("name", ""),
        "seed": int(manifest.config.get("seed", 0)),
    }


def _write_npz(path: Path, arrays: dict) -> None:
    """An .npz archive (a zip of .npy members, readable by numpy.load) whose
    bytes depend only on the arrays: members are stored uncompressed with a
    fixed timestamp, so equal captures have equal checksums."""
    import zipfile

    with zipfile.ZipFile(path, "w", compression=zipfile.ZIP_STORED, allowZip64=True) as zf:
        for key, arr in arrays.items():
            info = zipfile.ZipInfo(key + ".npy", date_time=(1980, 1, 1, 0, 0, 0))
            info.compress_type = zipfile.ZIP_STORED
            with zf.open(info, "w", force_zip64=True) as fh:
                np.lib.format.write_array(fh, np.asanyarray(arr), allow_pickle=False)


def write_capture(state_iter: Iterable[StepState], manifest: CaptureManifest, out_dir,
                  payload_format: str | None = None) -> Path:
    """Write the payload and the manifest JSON; return the manifest's path.

    ``payload_format``: "hdf5" (needs h5py), "npz", or None for HDF5 when h5py
    is importable and NPZ otherwise.  The payload file name is the stem of
    ``manifest.payload['path']`` with the format's suffix.
    """
    out_dir = Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    h5py = _h5py()
    fmt = payload_format or ("hdf5" if h5py is not None else "npz")
    if fmt not in FORMATS:
        raise ValueError(f"unknown payload format {fmt!r}")
    rel = manifest.payload.get("path")
    if not rel:
        raise ValueError("manifest.payload['path'] must be set before write_capture")
    stem = Path(rel).stem
    payload_path = out_dir / (stem + (".h5" if fmt == "hdf5" else ".npz"))
    manifest_path = out_dir / (stem + ".json")
    meta = _metadata(manifest)
    if fmt == "hdf5":
        if h5py is None:
            raise RuntimeError("payload format 'hdf5' needs h5py")
        with h5py.File(payload_path, "w", libver="earliest") as h:
            steps = h.create_group("steps", track_order=False)
            for st in state_iter:
                sg = steps.create_group(str(st.step), track_order=False)
                g = sg.create_group("state", track_order=False)
                for name, arr in st.state.items():
                    g.create_dataset(name, data=np.asarray(arr), track_times=False)
                g = sg.create_group("diagnostics", track_order=False)
                for name, v in st.diagnostics.items():
                    g.create_dataset(name, data=np.asarray(v), track_times=False)
            mg = h.create_group("metadata", track_order=False)
            for k, v in meta.items():
                mg.attrs[k] = v
    else:
        arrays = {}
        for st in state_iter:
            for name, arr in st.state.items():
                arrays[f"steps/{st.step}/state/{name}"] = np.asarray(arr)
            for name, v in st.diagnostics.items():
                arrays[f"steps/{st.step}/diagnostics/{name}"] = np.asarray(v)
        for k, v in meta.items():
            arrays[f"metadata/{k}"] = np.asarray(v)
        _write_npz(payload_path, arrays)
    d = manifest.to_dict()
    d["payload"] = dict(d["payload"], format=fmt, path=payload_path.name,
                        checksum="sha256:" + _sha256(payload_path))
    validate_manifest(d)
    with manifest_path.open("w", encoding="utf-8") as fh:
        json.dump(d, fh, indent=2, sort_keys=True)
    return manifest_path


class Capture:
    """Read-only view of a manifest + payload pair (reader.py:29-75)."""

    def __init__(self, manifest: CaptureManifest, payload_path: Path, metadata: dict):
        self.manifest = manifest
        self.metadata = metadata
        self._path = Path(payload_path)
        self._fmt = manifest.payload["format"]
        self._npz = None

    def _keys(self):
        if self._fmt == "npz":
            if self._npz is None:
                with np.load(self._path, allow_pickle=False) as z:
                    self._npz = {k: z[k] for k in z.files}
            return self._npz
        return None

    def _step_numbers(self):
        if self._fmt == "npz":
            return sorted({int(k.split("/")[1]) for k in self._keys() if k.startswith("steps/")})
        with _h5py().File(self._path, "r") as h:
            return sorted(int(k) for k in h["steps"]) if "steps" in h else []

    def steps(self):
        for n in self._step_numbers():
            yield self.step(n)

    def __iter__(self):
        return iter(self.steps())

    def step(self, n: int) -> StepState:
        state, diag = {}, {}
        if self._fmt == "npz":
            prefix = f"steps/{n}/"
            found = False
            for k, v in self._keys().items():
                if not k.startswith(prefix):
                    continue
                found = True
                kind, name = k[len(prefix):].split("/", 1)
                if kind == "state":
                    state[name] = np.asarray(v)
                else:
                    diag[name] = float(np.asarray(v).item())
            if not found:
                raise KeyError(f"steps/{n}")
        else:
            with _h5py().File(self._path, "r") as h:
                g = h[f"steps/{n}"]
                if "state" in g:
                    for name, ds in g["state"].items():
                        state[name] = np.asarray(ds[()])
                if "diagnostics" in g:
                    for name, ds in g["diagnostics"].items():
                        diag[name] = float(np.asarray(ds[()]).item())
        return StepState(step=n, state=state, diagnostics=diag)

    def field(self, step: int, name: str) -> np.ndarray:
        return self.step(step).state[name]


def load_capture(manifest_path) -> Capture:
    manifest_path = Path(manifest_path)
    with manifest_path.open("r", encoding="utf-8") as fh:
        data = json.load(fh)
    m = CaptureManifest.from_dict(data)
    p = Path(m.payload["path"])
    if not p.is_absolute():
        p = (manifest_path.parent / p).resolve()
    meta = {}
    if m.payload["format"] == "npz":
        with np.load(p, allow_pickle=False) as z:
            for k in z.files:
                if k.startswith("metadata/"):
                    v = z[k]
                    meta[k.split("/", 1)[1]] = v.item() if v.ndim == 0 else v
    else:
        with _h5py().File(p, "r") as h:
            if "metadata" in h:
                meta = dict(h["metadata"].attrs.items())
    return Capture(m, p, meta)


@dataclass
class CaptureDiff:
    bit_exact: bool
    max_abs_err: float
    max_rel_err: float
    mismatched_fields: list = field(default_factory=list)


def _field_errors(a, b):
    d = np.abs(a.astype(np.float64) - b.astype(np.float64))
    abs_err = float(d.max()) if d.size else 0.0
    den = np.maximum(np.abs(a.astype(np.float64)), np.abs(b.astype(np.float64)))
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.where(den > 0, d / den, 0.0)
    return abs_err, (float(rel.max()) if rel.size else 0.0)


def diff_captures(left, right, mode: str = "bit-exact", rtol: float = 0.0, atol: float = 0.0) -> CaptureDiff:
    """Step-by-step, field-by-field diff of two captures (diff.py:66-110).

    mode "bit-exact": a field matches when it is element-wise equal;
    mode "epsilon": when max |a - b| <= atol + rtol * max |b|.
    """
    lc, rc = load_capture(left), load_capture(right)
    lk, rk = set(lc._step_numbers()), set(rc._step_numbers())
    if lk != rk:
        return CaptureDiff(False, float("inf"), float("inf"),
                           [f"step:{n}:only-left" for n in sorted(lk - rk)] +
                           [f"step:{n}:only-right" for n in sorted(rk - lk)])
    max_abs = max_rel = 0.0
    bad = []
    for n in sorted(lk):
        ls, rs = lc.step(n), rc.step(n)
        for k in sorted(set(ls.state) | set(rs.state)):
            if k not in ls.state or k not in rs.state or ls.state[k].shape != rs.state[k].shape:
                bad.append(f"step:{n}:{k}")
                continue
            a, b = ls.state[k], rs.state[k]
            if a.dtype != b.dtype:
                raise TypeError(f"diff_captures: dtype mismatch on field {k!r}: {a.dtype} vs {b.dtype}")
            ea, er = _field_errors(a, b)
            max_abs, max_rel = max(max_abs, ea), max(max_rel, er)
            if mode == "bit-exact":
                ok = np.array_equal(a, b)
            else:
                ok = ea <= atol + rtol * (float(np.abs(b).max()) if b.size else 0.0)
            if not ok:
                bad.append(f"step:{n}:{k}")
    return CaptureDiff(not bad and max_abs == 0.0, max_abs, max_rel, bad)


def compare_captures(left, right, rtol: float = LBM_RELATIVE_TOLERANCE, atol: float = 0.0) -> CaptureDiff:
    """Cross-stack comparison at the lbm tolerance (equivalence/harness.py:86-185)."""
    lm, rm = load_capture(left).manifest, load_capture(right).manifest
    if (lm.sim.get("name"), lm.sim.get("category")) != (rm.sim.get("name"), rm.sim.get("category")):
        return CaptureDiff(False, float("inf"), float("inf"), ["sim:category-mismatch"])
    d = diff_captures(left, right, mode="epsilon", rtol=rtol, atol=atol)
    d.bit_exact = not d.mismatched_fields
    return d


@dataclass
class DeterminismVerdict:
    content_equivalent: bool
    detail: str


def run_twice_and_diff(runner: Callable[[int, Path], Path], seed: int = 42, tmp_dir=None) -> DeterminismVerdict:
    """Run ``runner(seed, out_dir)`` twice and diff the captures bit for bit
    (determinism/harness.py:97-140)."""
    import tempfile

    base = Path(tmp_dir) if tmp_dir is not None else Path(tempfile.mkdtemp(prefix="det-"))
    a, b = base / "run-a", base / "run-b"
    a.mkdir(parents=True, exist_ok=True)
    b.mkdir(parents=True, exist_ok=True)
    d = diff_captures(runner(seed, a), runner(seed, b), mode="bit-exact")
    if d.bit_exact:
        return DeterminismVerdict(True, "captures match exactly")
    head = ""
    if d.mismatched_fields:
        more = len(d.mismatched_fields) - 1
        head = f"; first mismatch at {d.mismatched_fields[0]}" + (f" (+{more} more)" if more else "")
    return DeterminismVerdict(False, f"max_abs_err={d.max_abs_err:g}, max_rel_err={d.max_rel_err:g}{head}")

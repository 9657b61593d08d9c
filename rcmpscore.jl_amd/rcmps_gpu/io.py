"""Portable state directories, "rcmps-portable-state" format 1, shared with the Julia package
(format specification: gpu/src/rcmps_gpu/io.py:1-104 of the reference; Julia side
gpu/julia/export_state.jl, import_state.jl).

A directory holds metadata.toml and headerless little-endian column-major complex128 D x D
matrices: R.bin and K.bin (required), rho.bin and gradient.bin (optional), each exactly 16 D^2
bytes.  read_portable / write_portable convert to and from a StateRecord (row-major numpy
arrays); rcmps_store / rcmps_load store and restore a State (src/storage.jl:4-38 semantics:
the stationary density is recomputed on load unless stored).
"""
from __future__ import annotations

import math
import os
import tempfile
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

try:  # Python >= 3.11
    import tomllib as _toml
except ModuleNotFoundError:  # Python 3.10
    import tomli as _toml

FORMAT = "rcmps-portable-state"
FORMAT_VERSION = 1
KEYS = ("format", "format_version", "D", "model", "coupling", "mass", "symmetry", "energy", "gradient_norm",
        "phi", "validation_passed", "producer", "source", "info", "integration", "manifold")
SYMMETRIES = ("Z2", "none")
MATRICES = ("R", "K", "rho", "gradient")


@dataclass(frozen=True)
class Integration:
    n_quad: int
    max_point: float
    atol: float
    rtol: float
    maxiters: int | None = None


@dataclass(frozen=True)
class Manifold:
    metric_reg: float
    precond_reg: float
    precond_reg_prefactor: float | None = None


@dataclass(frozen=True, eq=False)
class StateRecord:
    D: int
    model: str
    coupling: float
    R: np.ndarray | None = None
    K: np.ndarray | None = None
    mass: float = 1.0
    rho: np.ndarray | None = None
    gradient: np.ndarray | None = None
    energy: float | None = None
    gradient_norm: float | None = None
    symmetry: str | None = None
    phi: float | None = None
    validation_passed: bool | None = None
    integration: Integration | None = None
    manifold: Manifold | None = None
    producer: str | None = None
    source: str | None = None
    info: str | None = None
    files: tuple = ()
    metadata: dict = field(default_factory=dict)


def _int(v, name, minimum=1):
    if isinstance(v, bool) or not isinstance(v, (int, np.integer)) or v < minimum:
        raise ValueError(f"{name} must be an integer >= {minimum}, not {v!r}")
    return int(v)


def _real(v, name, positive=False):
    if isinstance(v, bool) or not isinstance(v, (int, float, np.integer, np.floating)):
        raise ValueError(f"{name} must be a real number, not {v!r}")
    v = float(v)
    if not math.isfinite(v) or (positive and v <= 0):
        raise ValueError(f"{name} must be finite{' and positive' if positive else ''}, not {v!r}")
    return v


def _text(v, name):
    if not isinstance(v, str):
        raise ValueError(f"{name} must be a string, not {v!r}")
    return v


def _bool(v, name):
    if not isinstance(v, bool):
        raise ValueError(f"{name} must be true or false, not {v!r}")
    return v


def read_matrix(path: Path, d: int) -> np.ndarray:
    path = Path(path)
    size = path.stat().st_size
    if size != 16 * d * d:
        raise ValueError(f"{path.name} holds {size} bytes, not 16 D^2 = {16 * d * d}")
    return np.ascontiguousarray(np.fromfile(path, dtype="<c16").reshape((d, d), order="F"))


def write_matrix(path: Path, a) -> None:
    a = np.asarray(a, dtype=np.complex128)
    if a.ndim != 2 or a.shape[0] != a.shape[1] or not np.isfinite(a).all():
        raise ValueError(f"{Path(path).name}: a finite square matrix is required")
    Path(path).write_bytes(np.asfortranarray(a).astype("<c16").tobytes(order="F"))


def read_portable(directory, *, arrays: bool = True) -> StateRecord:
    """Read a format-1 portable directory."""
    directory = Path(directory)
    meta_path = directory / "metadata.toml"
    if not meta_path.is_file():
        raise ValueError(f"{directory}: metadata.toml is missing")
    with meta_path.open("rb") as f:
        meta = _toml.load(f)
    if meta.get("format") != FORMAT:
        raise ValueError(f"{directory}: format must be {FORMAT!r}, not {meta.get('format')!r}")
    if _int(meta.get("format_version"), "format_version") != FORMAT_VERSION:
        raise ValueError(f"{directory}: unsupported format_version {meta['format_version']}")
    d = _int(meta.get("D"), "D")
    model = _text(meta.get("model"), "model")
    if not model:
        raise ValueError("the model name is empty")
    coupling = _real(meta.get("coupling"), "coupling")
    mass = _real(meta.get("mass", 1.0), "mass", positive=True)
    if model == "phi4" and mass != 1.0:
        raise ValueError("model phi4 has no mass parameter (its normal-ordering mass is 1)")
    opt = {}
    for key, conv in (("energy", _real), ("gradient_norm", _real), ("phi", _real)):
        if key in meta:
            opt[key] = conv(meta[key], key)
    if "symmetry" in meta:
        if meta["symmetry"] not in SYMMETRIES:
            raise ValueError(f"symmetry must be one of {SYMMETRIES}, not {meta['symmetry']!r}")
        opt["symmetry"] = meta["symmetry"]
    if "validation_passed" in meta:
        opt["validation_passed"] = _bool(meta["validation_passed"], "validation_passed")
    for key in ("producer", "source", "info"):
        if key in meta:
            opt[key] = _text(meta[key], key)
    if "integration" in meta:
        t = meta["integration"]
        opt["integration"] = Integration(
            _int(t.get("n_quad"), "integration.n_quad", 8), _real(t.get("max_point"), "integration.max_point", True),
            _real(t.get("atol"), "integration.atol", True), _real(t.get("rtol"), "integration.rtol", True),
            _int(t["maxiters"], "integration.maxiters") if "maxiters" in t else None)
    if "manifold" in meta:
        t = meta["manifold"]
        opt["manifold"] = Manifold(
            _real(t.get("metric_reg"), "manifold.metric_reg"), _real(t.get("precond_reg"), "manifold.precond_reg"),
            _real(t["precond_reg_prefactor"], "manifold.precond_reg_prefactor")
            if "precond_reg_prefactor" in t else None)
    files = tuple(sorted(p.name for p in directory.iterdir() if p.is_file()))
    mats = {}
    for name in MATRICES:
        p = directory / f"{name}.bin"
        if name in ("R", "K") and not p.is_file():
            raise ValueError(f"{directory}: {name}.bin is missing")
        if arrays and p.is_file():
            mats[name] = read_matrix(p, d)
        elif p.is_file():
            size = p.stat().st_size
            if size != 16 * d * d:
                raise ValueError(f"{p.name} holds {size} bytes, not {16 * d * d}")
    return StateRecord(D=d, model=model, coupling=coupling, mass=mass, files=files, metadata=meta, **mats, **opt)


def _toml_value(v):
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        v = float(v)
        if not math.isfinite(v):
            raise ValueError("metadata values must be finite")
        r = repr(v)
        return r if ("e" in r or "." in r) else r + ".0"
    if isinstance(v, str):
        esc = v.replace("\\", "\\\\").replace('"', '\\"').replace("\n", "\\n").replace("\t", "\\t")
        return f'"{esc}"'
    raise ValueError(f"unsupported metadata value {v!r}")


def write_portable(directory, record: StateRecord, *, overwrite: bool = False) -> Path:
    """Write `record` as a format-1 directory (atomically: a temporary sibling is renamed)."""
    directory = Path(directory)
    if directory.exists() and not overwrite:
        raise FileExistsError(f"{directory} exists")
    if record.R is None or record.K is None:
        raise ValueError("R and K are required")
    d = _int(record.D, "D")
    for name in MATRICES:
        a = getattr(record, name)
        if a is not None and np.asarray(a).shape != (d, d):
            raise ValueError(f"{name} must be {d} x {d}")
    lines = []
    top = {"format": FORMAT, "format_version": FORMAT_VERSION, "D": d, "model": record.model,
           "coupling": float(record.coupling), "mass": float(record.mass)}
    for key in ("symmetry", "energy", "gradient_norm", "phi", "validation_passed", "producer", "source", "info"):
        v = getattr(record, key)
        if v is not None:
            top[key] = v
    for key in KEYS:
        if key in top:
            lines.append(f"{key} = {_toml_value(top[key])}")
    if record.integration is not None:
        it = record.integration
        lines += ["", "[integration]", f"atol = {_toml_value(float(it.atol))}",
                  f"max_point = {_toml_value(float(it.max_point))}"]
        if it.maxiters is not None:
            lines.append(f"maxiters = {_toml_value(int(it.maxiters))}")
        lines += [f"n_quad = {_toml_value(int(it.n_quad))}", f"rtol = {_toml_value(float(it.rtol))}"]
    if record.manifold is not None:
        mf = record.manifold
        lines += ["", "[manifold]", f"metric_reg = {_toml_value(float(mf.metric_reg))}",
                  f"precond_reg = {_toml_value(float(mf.precond_reg))}"]
        if mf.precond_reg_prefactor is not None:
            lines.append(f"precond_reg_prefactor = {_toml_value(float(mf.precond_reg_prefactor))}")
    directory.parent.mkdir(parents=True, exist_ok=True)
    tmp = Path(tempfile.mkdtemp(prefix=directory.name + ".", dir=directory.parent))
    try:
        (tmp / "metadata.toml").write_text("\n".join(lines) + "\n", encoding="utf-8")
        for name in MATRICES:
            a = getattr(record, name)
            if a is not None:
                write_matrix(tmp / f"{name}.bin", a)
        if directory.exists():
            for p in directory.iterdir():
                p.unlink()
            directory.rmdir()
        os.replace(tmp, directory)
    except BaseException:
        for p in tmp.iterdir():
            p.unlink()
        tmp.rmdir()
        raise
    return directory


def record_of_state(s, *, energy=None, gradient=None, symmetry=None, producer="rcmps_gpu (MI355X)",
                    source=None) -> StateRecord:
    """The StateRecord of a State (src/storage.jl:4-24 fields)."""
    from .optimization import metric

    ip, mp = s.integrator_param, s.manifold_param
    e = s.e0 if energy is None else energy
    g = gradient if gradient is not None else (s.grad if np.any(s.grad) else None)
    gn = math.sqrt(metric(s, g, g)) if g is not None else None
    return StateRecord(
        D=s.D, model=s.model_param.name, coupling=float(np.real(s.model_param.coupling)), R=s.R, K=s.K,
        mass=float(getattr(s.model_param, "mass", 1.0)), rho=s.rho, gradient=g,
        energy=None if e is None or not math.isfinite(e) else float(e), gradient_norm=gn, symmetry=symmetry,
        integration=Integration(ip.n_quad, ip.max_point, ip.atol, ip.rtol, ip.maxiters),
        manifold=Manifold(mp.metric_reg, mp.precond_reg, mp.precond_reg_prefactor), producer=producer,
        source=source, info=s.info or None)


def jld2_fields(s) -> dict:
    """The schema-1 checkpoint fields of a State (src/storage.jl:10-16): e0 NaN and a zero gradient
    for a state never evaluated, as State(R, K) initialises them (src/state.jl:52).  An empty info
    string is stored as "rcmps_gpu" (the JLD2 encoding of an empty String is not reproduced)."""
    ip, mp = s.integrator_param, s.manifold_param
    e0 = s.e0 if s.e0 is not None and math.isfinite(s.e0) else float("nan")
    grad = s.grad if getattr(s, "grad", None) is not None else np.zeros((s.D, s.D), np.complex128)
    return dict(R=s.R, K=s.K, grad=grad, model=s.model_param.name, coupling=complex(s.model_param.coupling),
                e0=e0, info=s.info or "rcmps_gpu", integration=(ip.n_quad, ip.max_point, ip.atol, ip.rtol, ip.maxiters),
                manifold=(mp.metric_reg, mp.precond_reg_prefactor, mp.precond_reg))


def rcmps_store(s, path=None, *, directory="data", id=None, overwrite: bool = True, **kw) -> Path:
    """Store a State (src/storage.jl:4-26): a path ending in .jld2 (or no path: the reference's name
    "<model>_g<coupling>_D<D>_<id>.jld2" in `directory`) gets a JLD2 schema-1 checkpoint; any other
    path a portable directory (the format shared with the Julia package)."""
    if path is None:
        import secrets
        tag = id if id is not None else secrets.token_hex(5)
        path = Path(directory) / f"{s.model_param.name}_g{float(np.real(s.model_param.coupling))}_D{s.D}_{tag}.jld2"
    if str(path).endswith(".jld2"):
        from .jld2 import write_jld2
        return write_jld2(path, jld2_fields(s))
    return write_portable(path, record_of_state(s, **kw), overwrite=overwrite)


def rcmps_load(path, *, device: int = 0, recompute_rho: bool = False, steady_tol: float = 1e-11):
    """Load a State from a portable directory or a JLD2 checkpoint (schema 1, src/storage.jl:27-38);
    rho is checked if stored (recomputed otherwise; JLD2 checkpoints never store it)."""
    from .state import IntegratorParameters, ManifoldParameters, ModelParameters, State

    if str(path).endswith(".jld2"):
        from .jld2 import read_jld2
        rec = read_jld2(path)
    else:
        rec = read_portable(path)
    ip = IntegratorParameters()
    if rec.integration is not None:
        ip = IntegratorParameters(rec.integration.n_quad, rec.integration.max_point, rec.integration.atol,
                                  rec.integration.rtol, rec.integration.maxiters or 1_000_000)
    mp = ManifoldParameters()
    if rec.manifold is not None:
        mp = ManifoldParameters(rec.manifold.metric_reg,
                                rec.manifold.precond_reg_prefactor if rec.manifold.precond_reg_prefactor
                                is not None else 1e-2, rec.manifold.precond_reg)
    s = State(rec.R, rec.K, rho=None if recompute_rho else rec.rho, steady_tol=steady_tol, device=device,
              integrator_param=ip,
              model_param=ModelParameters(rec.model, rec.coupling, rec.mass), manifold_param=mp,
              info=rec.info or "")
    if rec.energy is not None:
        s.e0 = rec.energy
    if rec.gradient is not None:
        s.grad = rec.gradient
    return s

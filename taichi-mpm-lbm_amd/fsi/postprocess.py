"""Readers for the NPZ snapshots and validation summaries (reference:
fsi/postprocess.py). Host-side numpy only.

Snapshots are `snapshot_<step>.npz` files written by fsi.output. The time
series derived from them use the reference's keys and definitions
(fsi/postprocess.py:241-289), e.g. `lbm_total_mass_estimate` = sum of the
density field and `mpm_center_of_mass_*` = the unweighted mean particle
position, so tables and plots built on either implementation agree.
"""

from __future__ import annotations

import csv
import json
from dataclasses import dataclass
from pathlib import Path
from typing import Any

import numpy as np

# derived time-series keys, in the reference's column order
_ORDER = (
    "step", "time", "mpm_particle_count", "lbm_mean_density", "lbm_total_mass_estimate",
    "lbm_max_velocity_norm", "mpm_center_of_mass_x", "mpm_center_of_mass_y", "mpm_center_of_mass_z",
    "mpm_max_velocity_norm", "coupling_force_norm", "total_particle_coupling_force_norm",
    "total_fluid_coupling_force_norm", "coupling_unsupported_particle_count",
    "coupling_partial_support_particle_count", "coupling_clipped_particle_count", "ib_force_norm",
    "ib_total_force_norm", "ib_active_cell_count", "ib_clipped_cell_count", "contact_candidate_count",
    "contact_damped_particle_count",
)
_COUNTS = ("coupling_unsupported_particle_count", "coupling_partial_support_particle_count",
           "coupling_clipped_particle_count", "ib_active_cell_count", "ib_clipped_cell_count",
           "contact_candidate_count", "contact_damped_particle_count")
# snapshot field -> derived key of its Frobenius norm
_NORMS = (("coupling_force", "coupling_force_norm"), ("immersed_boundary_force", "ib_force_norm"),
          ("total_particle_coupling_force", "total_particle_coupling_force_norm"),
          ("total_fluid_coupling_force", "total_fluid_coupling_force_norm"),
          ("ib_total_force", "ib_total_force_norm"))
_SUMMARY_FIELDS = ("case_name", "case_passed", "metric_name", "value", "lower", "upper", "metric_passed",
                   "units", "description")


@dataclass(frozen=True)
class SnapshotInfo:
    path: Path
    step: int
    time: float
    particle_count: int
    lbm_shape: tuple[int, int, int]

    def to_dict(self) -> dict[str, Any]:
        return {"path": str(self.path), "step": self.step, "time": self.time,
                "particle_count": self.particle_count, "lbm_shape": self.lbm_shape}


def _scalar(v, kind):
    return kind(np.asarray(v).reshape(()).item())


def _step_of(path: Path):
    tail = path.stem[len("snapshot_"):] if path.stem.startswith("snapshot_") else ""
    return (0, int(tail)) if tail.isdigit() else (1, path.name)


def _plain(v: Any) -> Any:
    if isinstance(v, dict):
        return {k: _plain(x) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [_plain(x) for x in v]
    if isinstance(v, np.ndarray):
        return v.tolist()
    if isinstance(v, np.generic):
        return v.item()
    return v


def list_npz_snapshots(output_dir: str | Path) -> list[Path]:
    """snapshot_*.npz of a directory, ordered by step number."""
    d = Path(output_dir)
    return sorted(d.glob("snapshot_*.npz"), key=_step_of) if d.exists() else []


def load_npz_snapshot(path: str | Path) -> dict[str, np.ndarray]:
    with np.load(Path(path), allow_pickle=False) as data:
        return {k: np.array(data[k]) for k in data.files}


def snapshot_info(path: str | Path) -> SnapshotInfo:
    data = load_npz_snapshot(path)
    shape = tuple(int(v) for v in np.asarray(data["lbm_shape"]).reshape(-1))
    if len(shape) != 3:
        raise ValueError(f"lbm_shape must have 3 entries, got {shape}.")
    return SnapshotInfo(path=Path(path), step=_scalar(data["step"], int), time=_scalar(data["time"], float),
                        particle_count=_scalar(data["mpm_particle_count"], int), lbm_shape=shape)


def summarize_snapshots(output_dir: str | Path) -> list[SnapshotInfo]:
    return [snapshot_info(p) for p in list_npz_snapshots(output_dir)]


def _max_norm(v) -> float:
    a = np.asarray(v, dtype=np.float64)
    return 0.0 if a.size == 0 else float(np.linalg.norm(a.reshape(-1, a.shape[-1]), axis=1).max())


def _row(path: Path) -> dict[str, float | int]:
    d = load_npz_snapshot(path)
    row: dict[str, float | int] = {"step": _scalar(d["step"], int), "time": _scalar(d["time"], float),
                                   "mpm_particle_count": _scalar(d["mpm_particle_count"], int)}
    if "lbm_density" in d:
        row["lbm_mean_density"] = float(np.mean(d["lbm_density"]))
        row["lbm_total_mass_estimate"] = float(np.sum(d["lbm_density"]))
    if "lbm_velocity" in d:
        row["lbm_max_velocity_norm"] = _max_norm(d["lbm_velocity"])
    if "mpm_positions" in d and np.asarray(d["mpm_positions"]).size:
        c = np.mean(np.asarray(d["mpm_positions"], dtype=np.float64), axis=0)
        row.update(mpm_center_of_mass_x=float(c[0]), mpm_center_of_mass_y=float(c[1]),
                   mpm_center_of_mass_z=float(c[2]))
    if "mpm_velocities" in d:
        row["mpm_max_velocity_norm"] = _max_norm(d["mpm_velocities"])
    for field, key in _NORMS:
        if field in d:
            row[key] = float(np.linalg.norm(np.asarray(d[field])))
    for key in _COUNTS:
        if key in d:
            row[key] = _scalar(d[key], int)
    return row


def _ordered(keys) -> list[str]:
    keys = set(keys)
    return [k for k in _ORDER if k in keys] + sorted(keys.difference(_ORDER))


def extract_snapshot_timeseries(output_dir: str | Path) -> dict[str, np.ndarray]:
    """One array per derived key present in every snapshot, in step order."""
    rows = [_row(p) for p in list_npz_snapshots(output_dir)]
    if not rows:
        return {"step": np.array([], dtype=np.int64), "time": np.array([], dtype=np.float64),
                "mpm_particle_count": np.array([], dtype=np.int64)}
    common = set.intersection(*(set(r) for r in rows))
    return {k: np.array([r[k] for r in rows]) for k in _ordered(common)}


def summarize_coupling_diagnostics(output_dir: str | Path) -> dict[str, np.ndarray]:
    ts = extract_snapshot_timeseries(output_dir)
    keys = ("step", "time", "coupling_force_norm", "total_particle_coupling_force_norm",
            "total_fluid_coupling_force_norm", *_COUNTS[:3], "ib_force_norm", "ib_total_force_norm",
            *_COUNTS[3:])
    return {k: ts[k] for k in keys if k in ts}


def _checked_keys(ts: dict[str, np.ndarray]) -> list[str]:
    if not ts:
        raise ValueError("time series must not be empty.")
    keys = _ordered(ts)
    if len({len(np.asarray(ts[k])) for k in keys}) != 1:
        raise ValueError("all time series arrays must have the same length.")
    return keys


def write_timeseries_csv(timeseries: dict[str, np.ndarray], path: str | Path) -> Path:
    keys = _checked_keys(timeseries)
    out = Path(path)
    out.parent.mkdir(parents=True, exist_ok=True)
    with out.open("w", newline="", encoding="utf-8") as fh:
        w = csv.DictWriter(fh, fieldnames=keys)
        w.writeheader()
        for i in range(len(np.asarray(timeseries[keys[0]]))):
            w.writerow({k: _plain(np.asarray(timeseries[k])[i]) for k in keys})
    return out


def write_timeseries_json(timeseries: dict[str, np.ndarray], path: str | Path) -> Path:
    _checked_keys(timeseries)
    out = Path(path)
    out.parent.mkdir(parents=True, exist_ok=True)
    out.write_text(json.dumps(_plain(timeseries), indent=2), encoding="utf-8")
    return out


def load_validation_summary(path: str | Path) -> list[dict[str, Any]]:
    data = json.loads(Path(path).read_text(encoding="utf-8"))
    if not isinstance(data, list):
        raise ValueError("validation summary JSON must contain a list of reports.")
    return data


def validation_summary_table(reports: list[dict[str, Any]]) -> list[dict[str, Any]]:
    """One row per metric of each report dict (ValidationReport.to_dict())."""
    return [{"case_name": r.get("case_name", ""), "case_passed": bool(r.get("passed", False)),
             "metric_name": m.get("name", ""), "value": m.get("value"), "lower": m.get("lower"),
             "upper": m.get("upper"), "metric_passed": bool(m.get("passed", False)),
             "units": m.get("units", ""), "description": m.get("description", "")}
            for r in reports for m in r.get("metrics", [])]


def write_validation_summary_csv(rows: list[dict[str, Any]], path: str | Path) -> Path:
    out = Path(path)
    out.parent.mkdir(parents=True, exist_ok=True)
    with out.open("w", newline="", encoding="utf-8") as fh:
        w = csv.DictWriter(fh, fieldnames=list(_SUMMARY_FIELDS))
        w.writeheader()
        for r in rows:
            w.writerow({k: _plain(r.get(k)) for k in _SUMMARY_FIELDS})
    return out


def plot_timeseries(timeseries: dict[str, np.ndarray], y_key: str, output_path: str | Path, *,
                    x_key: str = "step") -> Path:
    """PNG line plot of one key (needs matplotlib)."""
    for k, role in ((x_key, "x_key"), (y_key, "y_key")):
        if k not in timeseries:
            raise KeyError(f"Missing {role} in time series: {k}.")
    import matplotlib.pyplot as plt

    out = Path(output_path)
    out.parent.mkdir(parents=True, exist_ok=True)
    fig, ax = plt.subplots()
    ax.plot(np.asarray(timeseries[x_key]), np.asarray(timeseries[y_key]), marker="o")
    ax.set_xlabel(x_key)
    ax.set_ylabel(y_key)
    ax.grid(True)
    fig.tight_layout()
    fig.savefig(out)
    plt.close(fig)
    return out

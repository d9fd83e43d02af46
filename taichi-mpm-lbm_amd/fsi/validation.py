"""Scalar validation metrics and reports (reference: fsi/validation.py).

A case reports named scalar metrics, each with optional lower/upper bounds; a
metric passes when it is finite and inside its bounds, a report when all its
metrics pass. Field names and dict layouts follow fsi/validation.py:8-97 so
reports serialise the same way.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any


@dataclass(frozen=True)
class ValidationMetric:
    """One scalar metric with optional pass bounds (fsi/validation.py:8-29)."""

    name: str
    value: float
    lower: float | None = None
    upper: float | None = None
    passed: bool = True
    units: str = ""
    description: str = ""

    def to_dict(self) -> dict[str, Any]:
        return {key: getattr(self, key)
                for key in ("name", "value", "lower", "upper", "passed", "units", "description")}


@dataclass(frozen=True)
class ValidationReport:
    """A named case and its metrics (fsi/validation.py:32-51)."""

    case_name: str
    metrics: tuple[ValidationMetric, ...]
    metadata: dict[str, Any] = field(default_factory=dict)

    @property
    def passed(self) -> bool:
        return all(m.passed for m in self.metrics)

    def to_dict(self) -> dict[str, Any]:
        return {"case_name": self.case_name, "passed": self.passed,
                "metrics": [m.to_dict() for m in self.metrics], "metadata": self.metadata}


def bounded_metric(name: str, value: float, *, lower: float | None = None, upper: float | None = None,
                   units: str = "", description: str = "") -> ValidationMetric:
    """Metric that passes when finite and within [lower, upper] (either bound
    optional; fsi/validation.py:54-76)."""
    v = float(value)
    ok = math.isfinite(v) and (lower is None or v >= float(lower)) and (upper is None or v <= float(upper))
    return ValidationMetric(name=name, value=v, lower=lower, upper=upper, passed=bool(ok), units=units,
                            description=description)


def finite_metric(name: str, value: float, *, upper: float | None = None, units: str = "",
                  description: str = "") -> ValidationMetric:
    """Finite metric with an optional upper bound (fsi/validation.py:79-93)."""
    return bounded_metric(name, value, upper=upper, units=units, description=description)


def relative_error(value: float, reference: float) -> float:
    """|value - reference| / |reference|, or the absolute difference when the
    reference is zero (fsi/validation.py:96-101)."""
    v, r = float(value), float(reference)
    return abs(v - r) if r == 0.0 else abs((v - r) / r)

"""Regression output in TDycore's format (src/tdyregression.c:141-273).

``TDyRegressionOutput`` writes ``<filename>.regression``: the liquid
pressure's max, min and mean over all cells, then the pressure at a few
sampled cells.  The sampled natural ids are ``floor(n / n_sample) * i``,
``i < n_sample`` (src/tdyregression.c:67-89, one process), with ``n_sample``
from ``-tdy_regression_test_num_cells_per_process`` (default 2, capped at
n).  The number formats are the reference's (``%21.13e``, ids ``%9d``), so the
files compare against its ``*.regression.gold`` files.
"""
from __future__ import annotations

import numpy as np


def _host(U) -> np.ndarray:
    if hasattr(U, "detach"):  # a torch tensor (device or host)
        return U.detach().cpu().numpy().astype(np.float64, copy=False)
    return np.asarray(U, dtype=np.float64)


def regression_samples(n: int, n_sample: int = 2):
    """Natural cell ids the reference samples for an n-cell vector."""
    n_sample = min(n_sample, n)
    step = n // n_sample
    return [step * i for i in range(n_sample)]


def regression_text(U, n_sample: int = 2) -> str:
    """The regression block for pressure U (one value per cell)."""
    u = _host(U)
    n = u.shape[0]
    lines = ["-- PRESSURE: Liquid Pressure --",
             "      Max: %21.13e" % u.max(),
             "      Min: %21.13e" % u.min(),
             "     Mean: %21.13e" % (u.sum() / n)]
    for c in regression_samples(n, n_sample):
        lines.append("%9d: %21.13e" % (c, u[c]))
    return "\n".join(lines) + "\n"


def write_regression(U, filename: str, n_sample: int = 2) -> str:
    """Write ``<filename>.regression`` and return its path."""
    path = f"{filename}.regression"
    with open(path, "w") as fh:
        fh.write(regression_text(U, n_sample))
    return path


def parse_regression(text: str) -> dict:
    """{'Max', 'Min', 'Mean', 'samples': {id: value}} from a regression block."""
    out = {"samples": {}}
    for line in text.splitlines():
        line = line.strip()
        if not line or line.startswith("--"):
            continue
        key, val = line.split(":")
        key = key.strip()
        if key in ("Max", "Min", "Mean"):
            out[key] = float(val)
        else:
            out["samples"][int(key)] = float(val)
    return out

"""Scans on disk for the file-driven driver (mcsart_recon): raw FP32 stacks in MATLAB
column-major order plus a JSON sidecar holding the gateway's GEO_PARA / ITER_PARA fields.

Layouts (src/entry_AUMISART.cu:163-190): a volume is (nx, ny, nz) column-major, i.e. the C-order
array (nz, ny, nx); a projection stack is (na, nb, n_views), i.e. C-order (n_views, nb, na).
"""
from __future__ import annotations

import json
import os
import subprocess
from typing import Mapping, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(HERE, "mcsart_recon")
FIELD_NAMES = ("alpha_x", "alpha_y", "alpha_z", "beta_x", "beta_y", "beta_z")


def write_raw(path: str, a: np.ndarray) -> None:
    np.ascontiguousarray(a, dtype=np.float32).tofile(path)


def read_raw(path: str, shape) -> np.ndarray:
    a = np.fromfile(path, dtype=np.float32)
    return a.reshape(shape)


def _jsonable(v):
    if isinstance(v, np.ndarray):
        return v.tolist()
    if isinstance(v, (np.floating, np.integer)):
        return v.item()
    return v


def write_scan(directory: str, geo: Mapping, it: Mapping, proj: np.ndarray,
               img0: Optional[np.ndarray] = None,
               fields: Optional[Sequence[np.ndarray]] = None) -> str:
    """Write SCAN.json and its raw files into `directory`; returns the JSON path."""
    os.makedirs(directory, exist_ok=True)
    doc = {"geo": {k: _jsonable(v) for k, v in geo.items()},
           "iter": {k: _jsonable(v) for k, v in it.items()}, "proj": "proj.raw"}
    write_raw(os.path.join(directory, "proj.raw"), proj)
    if img0 is not None:
        write_raw(os.path.join(directory, "img0.raw"), img0)
        doc["img0"] = "img0.raw"
    if fields is not None:
        doc["fields"] = {}
        for name, f in zip(FIELD_NAMES, fields):
            write_raw(os.path.join(directory, name + ".raw"), f)
            doc["fields"][name] = name + ".raw"
    path = os.path.join(directory, "scan.json")
    with open(path, "w") as f:
        json.dump(doc, f)
    return path


def run_cli(scan_json: str, outdir: str, dry_run: bool = False, timeout: float = 600):
    """Run mcsart_recon; returns the CompletedProcess (stdout holds the dry-run JSON line)."""
    if not os.path.exists(CLI):
        raise RuntimeError(f"{CLI} is not built (python -m mcsart_amd.build)")
    os.makedirs(outdir, exist_ok=True)
    cmd = [CLI, scan_json, outdir] + (["--dry-run"] if dry_run else [])
    return subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)


def read_outputs(outdir: str, nx: int, ny: int, nz: int, na: int, nb: int, n_views: int,
                 n_iter: int):
    """img (nz, ny, nx), nrm (n_iter*n_views,), alpha_x (nz, ny, nx), residual (V, nb, na)."""
    return (read_raw(os.path.join(outdir, "img.raw"), (nz, ny, nx)),
            read_raw(os.path.join(outdir, "nrm.raw"), (n_iter * n_views,)),
            read_raw(os.path.join(outdir, "alpha_x.raw"), (nz, ny, nx)),
            read_raw(os.path.join(outdir, "residual.raw"), (n_views, nb, na)))

"""Readers and writers of the reference's text formats (its data/ directory and
its CSV outputs), so that synthetic cases can drive the drop-in executable
exactly as files would drive the reference.

Formats (src/NmrRelaxationSolve.cpp of the reference):
  data/Nmr_start.dat   ik, jk, kk, one per line (cpp:68-93)
  data/Nmr_data.dat    61 lines "value<TAB>description"; keys are found by
                       substring in the description, first match wins
                       (cpp:189-723)
  data/Nmr_grad.dat    count, then one gradient per line (cpp:727-778)
  data/Poro.dat        header, then '0'/'1' cells, k outer, i inner (cpp:859-873)
  data/Vx.dat ...      header, then face velocities, i outer, k inner
                       (cpp:1466-1536); Vx has ik+1 faces along x, etc.
  results_*.csv        "Time; SumX; SumY; SumZ" then "%e; %e; %e; %e" rows
"""
from __future__ import annotations

import os

import numpy as np

# Line order and descriptions of the 61-line data file (the descriptions are
# what the reader matches on).
DATA_KEYS = [
    ("KEY", 1), ("IPRINT", 1), ("JPRINT", 1), ("KPRINT", 1), ("Printing on console", 500), ("NDWRIT", 1),
    ("NDVID", 1), ("NDAMPL", 1), ("NPOR - KEY FOR POROSITY", 1), ("NCOMP - KEY FOR COMPONENTS", 1),
    ("NFLUID - KEY FOR FLUID (1 - WATER, 2 - OIL)", 1), ("nSurfRelWater", 0), ("nSurfRelOil", 0),
    ("NVECTOR - KEY FOR VECTOR MULTIPLICATION", 0), ("NVELOC - KEY FOR VELOCITY", 0),
    ("NDIFFUS - KEY FOR DIFFUSION", 1), ("NDIRECT_FIELD (0 for  automatic calculation, 1 - X, 2 - Y, 3 - Z)", 3),
    ("NSEQUENCE - KEY FOR SEQUENCE", 0), ("NB1", 1), ("NB2", 1), ("NB3", 1), ("NB4", 1), ("NB5", 1), ("NB6", 1),
    ("DX", 1e-6), ("DY", 1e-6), ("DZ", 1e-6), ("BX0 - OUTER MAGNETIC FIELD (X)", 0.0),
    ("BY0 - OUTER MAGNETIC FIELD (Y)", 0.0), ("BZ0 - OUTER MAGNETIC FIELD (Z)", 1.0),
    ("BDX0 - ADDITIONAL MAGNETIC FIELD (X)", 0.0), ("BDY0 - ADDITIONAL MAGNETIC FIELD (Y)", 0.0),
    ("BDZ0 - ADDITIONAL MAGNETIC FIELD (Z)", 0.0), ("TIME1 - \"SMALL\" TIME", 0.0), ("TIME2 - \"BIG\" TIME", 1.0),
    ("GRAD_X - STATIONARY MAGNETIC FIELD GRADIENT (X) T/M", 0.0),
    ("GRAD_Y - STATIONARY MAGNETIC FIELD GRADIENT (Y) T/M", 0.0),
    ("GRAD_Z - STATIONARY MAGNETIC FIELD GRADIENT (Z) T/M", 0.0), ("AMPL_NOISE - NOISE AMPLITUDE", 0.0),
    ("VEL_MULT", 1.0), ("D_OIL", 2.1e-10), ("D_WAT", 2.3e-9), ("R1_OIL", 0.0), ("R1_WAT", 0.0), ("R2_OIL", 0.0),
    ("R2_WAT", 0.0), ("T1_OIL", 2.5e10), ("T1_WAT", 2.5e10), ("T2_OIL", 2.5e10), ("T2_WAT", 2.0), ("T0", 300.0),
    ("AN(1)", 1.0), ("AN(2)", 0.0), ("BN(1)", 0.0), ("BN(2)", 1.0), ("CFL", 1.0 / 30.0), ("CFL1", 1.0),
    ("TIMEW(1)", 1e-6), ("bufferX", 0), ("bufferY", 0), ("bufferZ", 0),
]
assert len(DATA_KEYS) == 61


def _fmt(v) -> str:
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return repr(float(v))


def write_inputs(root: str, shape, mask=None, vx=None, vy=None, vz=None, grad_table=None, n1=None, surf=None,
                 **overrides) -> None:
    """Write root/data/*.dat for a compact case.  overrides name a key by the
    start of its description (e.g. NSEQUENCE=2, R2_WAT=1e-5, Printing_on=10).
    n1: concentration (nz, ny, nx) for data/n1_init.dat (NCOMP=2); surf:
    {"SurfRel1WaterX": faces, ...} face relaxivity files, staggered like the
    velocities (X: (nz, ny, nx+1), Y: (nz, ny+1, nx), Z: (nz+1, ny, nx))."""
    nz, ny, nx = shape
    d = os.path.join(root, "data")
    os.makedirs(d, exist_ok=True)
    os.makedirs(os.path.join(root, "results_NMR"), exist_ok=True)
    with open(os.path.join(d, "Nmr_start.dat"), "w") as f:
        f.write(f"{nx}\n{ny}\n{nz}\n")
    ov = {k.replace("_", " ") if k.startswith("Printing") else k: v for k, v in overrides.items()}
    if mask is not None:
        ov.setdefault("NPOR", 1)
    if vx is not None:
        ov.setdefault("NVELOC", 1)
    lines = []
    for desc, val in DATA_KEYS:
        for k, v in ov.items():
            if desc.split(" ")[0] == k or (" " in k and desc.startswith(k)):
                val = v
        lines.append(f"{_fmt(val)}\t{desc}")
    with open(os.path.join(d, "Nmr_data.dat"), "w") as f:
        f.write("\n".join(lines) + "\n")
    table = np.zeros(399) if grad_table is None else np.asarray(grad_table, float)
    with open(os.path.join(d, "Nmr_grad.dat"), "w") as f:
        f.write(f"{len(table)}\n" + "".join(f"{repr(float(g))}\n" for g in table))
    if mask is not None:
        with open(os.path.join(d, "Poro.dat"), "w") as f:
            f.write("poro\n")
            for k in range(nz):
                f.write("\n".join(" ".join("1" if c else "0" for c in row) for row in mask[k]) + "\n")
    # i outer, j, k inner; X-staggered arrays have nx+1 faces, Y ny+1, Z nz+1
    staggered = []
    if vx is not None:
        staggered += [("Vx.dat", vx), ("Vy.dat", vy), ("Vz.dat", vz)]
    if n1 is not None:
        staggered.append(("n1_init.dat", n1))
    for key, v in (surf or {}).items():
        staggered.append((key + ".dat", v))
    for name, v in staggered:
        arr = np.transpose(np.asarray(v, float), (2, 1, 0))  # (i, j, k)
        with open(os.path.join(d, name), "w") as f:
            f.write(name + "\n")
            np.savetxt(f, arr.reshape(-1), fmt="%.17g")


def read_results(path: str) -> np.ndarray:
    """Rows of a results_*.csv: (n, 4) time, SumX, SumY, SumZ, or (n, 7) with
    the two-phase sums SumX1..SumZ1, SumX2..SumZ2."""
    rows = []
    with open(path) as f:
        next(f)
        for line in f:
            if line.strip():
                rows.append([float(x) for x in line.split(";")])
    return np.array(rows)

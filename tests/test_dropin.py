"""The drop-in NmrRelaxationSolve class (initializator_amd/host) driven by the
reference's own input files: parsing (CPU), refusal of configurations the GPU
solver does not implement (CPU), and results_gpu.csv / results_cpu.csv
against closed forms and the FP64 restatement of the reference's loops (GPU).
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from initializator_amd import generators, refio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "initializator_amd", "nmr_exe")
REFDATA = os.path.join(ROOT, "tests", "golden", "reference_data")


@pytest.fixture(scope="module")
def exe():
    if not os.path.exists(EXE):
        subprocess.run(["make", "-C", os.path.join(ROOT, "initializator_amd", "csrc")], check=True,
                       stdout=subprocess.DEVNULL)
    return EXE


def _run(exe, cwd, *args, env=None, check=True):
    e = dict(os.environ)
    e.update(env or {})
    r = subprocess.run([exe, *args], cwd=cwd, capture_output=True, text=True, env=e, timeout=600)
    if check and r.returncode != 0:
        raise AssertionError(f"nmr_exe failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    return r


def _config(exe, cwd, *args, full=False):
    r = _run(exe, cwd, *args, *([] if full else ["--print-config"]), env={"NMR_INIT_CONFIG": "1"} if full else None)
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_parses_the_reference_data_files(exe, tmp_path):
    """The shipped data/Nmr_data.dat (61 substring-keyed lines) and
    Nmr_grad.dat are read as the reference reads them (cpp:175-810)."""
    os.makedirs(tmp_path / "data")
    for f in ("Nmr_data.dat", "Nmr_grad.dat", "Nmr_start.dat"):
        shutil.copy(os.path.join(REFDATA, f), tmp_path / "data" / f)
    c = _config(exe, tmp_path, "200", "1", "399")
    assert (c["ik"], c["jk"], c["kk"]) == (10, 10, 10)
    assert c["ndPrint"] == 500 and c["nComponents"] == 2 and c["nSequence"] == 98
    assert c["nVelocity"] == 1 and c["nDirectionField"] == 1 and c["nVectorMult"] == 0
    assert c["dx"] == 2.67e-6 and c["bsx"] == 1.0 and c["time2"] == 70e-3 and c["time1"] == 0.0
    assert c["velMult"] == 1e-2 and c["dWater"] == 2.7e-9 and c["dOil"] == 2.1e-10
    assert c["cfl"] == 0.002 and c["cfl1"] == 0.002 and c["t2Water"] == 2.5e10
    assert c["pEq"] == 1e-3  # the 0.31 = 310 pores x 1/1000 of results_gpu.csv
    assert c["endTime"] == pytest.approx(70e-3)  # TIME2 + 2 TIME1 for sequence 98
    grad = [float(x) for x in open(os.path.join(REFDATA, "Nmr_grad.dat")).read().split()[1:]]
    assert c["gradientPulse"] == grad[199] == 0.0
    assert c["multipleGradients"] == [grad[199], grad[0], grad[398]]


def test_shipped_two_phase_config_needs_its_concentration_file(exe, tmp_path):
    """The shipped configuration is two-phase (NCOMP=2) but ships without
    data/n1_init.dat: init() says so and exits 1 before touching a device
    (the reference would run on with uninitialised concentrations)."""
    os.makedirs(tmp_path / "data")
    for f in ("Nmr_data.dat", "Nmr_grad.dat", "Nmr_start.dat"):
        shutil.copy(os.path.join(REFDATA, f), tmp_path / "data" / f)
    r = _run(exe, tmp_path, "1", "1", "1", check=False)
    assert r.returncode == 1 and "no file data/Poro.dat" in r.stderr  # NPOR=1: the mask is not shipped either
    with open(tmp_path / "data" / "Poro.dat", "w") as f:
        f.write("poro\n" + "1 " * 1000 + "\n")
    r = _run(exe, tmp_path, "1", "1", "1", check=False)
    assert r.returncode == 1 and "no file data/Vx.dat" in r.stderr  # NVELOC=1
    for name in ("Vx.dat", "Vy.dat", "Vz.dat"):
        with open(tmp_path / "data" / name, "w") as f:
            f.write(name + "\n" + "0 " * 1100 + "\n")
    r = _run(exe, tmp_path, "1", "1", "1", check=False)
    assert r.returncode == 1 and "no file data/n1_init.dat" in r.stderr


def _two_phase_case(tmp_path, shape=(6, 5, 7), seed=3, **ov):
    mask = generators.porosity(shape, 0.6, seed=seed, sigma=1.0)
    rng = np.random.default_rng(seed)
    n1 = np.clip(rng.random(shape) * 1.2 - 0.1, 0.0, 1.0)
    nz, ny, nx = shape
    surf = {"SurfRel1WaterX": rng.uniform(1e-6, 3e-5, (nz, ny, nx + 1)),
            "SurfRel1WaterY": rng.uniform(1e-6, 3e-5, (nz, ny + 1, nx)),
            "SurfRel1WaterZ": rng.uniform(1e-6, 3e-5, (nz + 1, ny, nx)),
            "SurfRel2WaterX": rng.uniform(1e-6, 3e-5, (nz, ny, nx + 1)),
            "SurfRel2WaterY": rng.uniform(1e-6, 3e-5, (nz, ny + 1, nx)),
            "SurfRel2WaterZ": rng.uniform(1e-6, 3e-5, (nz + 1, ny, nx))}
    base = dict(NCOMP=2, nSurfRelWater=1, R1_OIL=4e-6, R2_OIL=8e-6, T1_OIL=0.3, T2_OIL=0.05, T1_WAT=0.6,
                T2_WAT=0.2, D_OIL=2.1e-10, D_WAT=2.3e-9)
    base.update(ov)
    cfg = _write_case(tmp_path, shape, mask, None, n1=n1, surf=surf, **base)
    return cfg, mask, n1, surf


def _two_phase_oracle_coefficients(cfg, mask, n1, surf):
    from initializator_amd import layout

    p = oracle.params(mask.shape, cfg["DX"], 1.0, dy=cfg["DY"], dz=cfg["DZ"])
    act = layout.pad_mask(mask)
    fl = oracle.fluids(d=(cfg["D_WAT"], cfg["D_OIL"]), t1=(cfg["T1_WAT"], cfg["T1_OIL"]),
                       t2=(cfg["T2_WAT"], cfg["T2_OIL"]), r1=(cfg["R1_WAT"], cfg["R1_OIL"]),
                       r2=(cfg["R2_WAT"], cfg["R2_OIL"]), ncomp=2, an=(cfg["AN(1)"], cfg["AN(2)"]),
                       bn=(cfg["BN(1)"], cfg["BN(2)"]))
    n1f = n1.astype(np.float32).astype(np.float64)  # the reference reads the file into a float
    conc = (layout.pad_field(n1f), layout.pad_field(1.0 - n1f))
    faces = []
    for rel in ("1", "2"):
        for ax, stag in (("X", "x"), ("Y", "y"), ("Z", "z")):
            key = f"SurfRel{rel}Water{ax}"
            faces.append(np.ascontiguousarray(layout.pad_velocity(*[surf[f"SurfRel{rel}Water{a}"] for a in "XYZ"],
                                                                  mask.shape)["xyz".index(stag)]))
    faces += [None] * 6
    return p, act, fl, conc, faces


def test_two_phase_inputs_set_the_step_like_the_reference(exe, tmp_path):
    """n1_init.dat and face relaxivity files are read (i outer, k inner) and
    mixed as FillInDiffusinAndTime does: the step limits the drop-in derives
    equal those of the FP64 restatement's per-voxel D and T."""
    cfg, mask, n1, surf = _two_phase_case(tmp_path, NSEQUENCE=2, TIME1=2e-4, TIME2=5e-4)
    c = _config(exe, tmp_path, "5", "0", "0", full=True)
    p, act, fl, conc, faces = _two_phase_oracle_coefficients(cfg, mask, n1, surf)
    _, _, _, (dmax, tmin) = oracle.fill_coefficients_2f(p, act, fl, conc, faces)
    assert c["nComponents"] == 2
    assert c["dTimeDiff"] == pytest.approx(2.0 * cfg["CFL"] * cfg["DZ"] ** 2 / dmax, rel=1e-14)
    assert c["dTimeRel"] == pytest.approx(cfg["CFL1"] * tmin, rel=1e-14)


def test_written_inputs_round_trip(exe, tmp_path):
    """refio.write_inputs produces files the drop-in reads back, including the
    step size and step count of StepTime (cpp:2519-2534)."""
    shape = (6, 5, 7)
    mask = generators.porosity(shape, 0.5, seed=1, sigma=1.0)
    vx, vy, vz = generators.velocity(mask, seed=1, sigma=1.5, vmax=2e-3)
    cfg = _write_case(tmp_path, shape, mask, (vx, vy, vz), NSEQUENCE=2, TIME1=2e-4, TIME2=5e-4, R2_WAT=1e-5)
    c = _config(exe, tmp_path, "5", "0", "0", full=True)
    _, dt, nstep = _limits_only(cfg, mask, (vx, vy, vz))
    assert c["dTime"] == dt and c["nStep"] == nstep
    assert c["nPoro"] == 1 and c["nVelocity"] == 1 and c["nSequence"] == 2 and c["time1"] == 2e-4


def _limits_only(cfg, mask, vel):
    """dTime and nStep as the FP64 restatement computes them (no time loop)."""
    import math

    dx, dy, dz = cfg["DX"], cfg["DY"], cfg["DZ"]
    from initializator_amd import layout

    p = oracle.params(mask.shape, dx, 1.0, dy=dy, dz=dz)
    act = layout.pad_mask(mask)
    _, _, _, (dmax, tmin) = oracle.fill_coefficients(p, act, cfg["D_WAT"], cfg["T1_WAT"], cfg["T2_WAT"],
                                                     cfg["R1_WAT"], cfg["R2_WAT"])
    vmax = max(0.0, max(float(v.max()) for v in vel)) if cfg["NVELOC"] else 0.0
    vmin = min(0.0, min(float(v.min()) for v in vel)) if cfg["NVELOC"] else 0.0
    vm = vmax * cfg["VEL_MULT"] if vmax > abs(vmin) else abs(vmin) * cfg["VEL_MULT"]
    dt_vel = oracle._chain(dx, dy, dz, lambda h: cfg["CFL"] * h / (vm + oracle.SMALL)) if cfg["NVELOC"] else 1.0
    dt_diff = oracle._chain(dx, dy, dz, lambda h: 2.0 * cfg["CFL"] * h * h / dmax)
    dt = oracle.step_time(math.inf, cfg["CFL1"] / (oracle.G_REF * oracle.SMALL), dt_vel, dt_diff,
                          cfg["CFL1"] * tmin, cfg["TIME1"], cfg["TIME2"])
    end = {0: 2 * cfg["T2_WAT"], 1: 2 * cfg["T2_WAT"], 2: 4.0 * cfg["TIME1"] + cfg["TIME2"]}.get(
        cfg["NSEQUENCE"], cfg["TIME2"] + 2.0 * cfg["TIME1"])
    return None, dt, int(end / dt) + 1


def _write_case(root, shape, mask, vel, grad_table=None, n1=None, surf=None, **ov):
    cfg = dict(refio.DATA_KEYS)
    cfg = {k.split(" ")[0] if not k.startswith("Printing") else "Printing on": v for k, v in cfg.items()}
    cfg.update({k.replace("_", " ") if k.startswith("Printing") else k: v for k, v in ov.items()})
    if mask is not None:
        cfg["NPOR"] = 1
    if vel is not None:
        cfg["NVELOC"] = 1
    refio.write_inputs(str(root), shape, mask, *(vel if vel is not None else (None, None, None)),
                       grad_table=grad_table, n1=n1, surf=surf, **ov)
    return cfg


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", ["maccormack", "upwind"])
def test_fid_closed_form(exe, tmp_path, scheme):
    """FID on an all-pore box with bulk T2 only: the 90 deg pulse about x turns
    pEq*N = 1 along z into -1 along y, which then decays per update by
    a(dt/T2), a(x) = 1 - x + x^2/2 for the MacCormack update solve_gpu() uses
    by default, 1 - x for the upwind update (results_gpu.csv row i holds i+1
    updates and the t = 0 pulse that follows the first)."""
    shape = (10, 6, 8)
    cfg = _write_case(tmp_path, shape, None, None, NPOR=0, NSEQUENCE=0, T2_WAT=0.01, Printing_on=50)
    args = ["1", "0", "0"] + ([] if scheme == "maccormack" else ["--scheme", "upwind"])
    _run(exe, tmp_path, *args)
    rows = refio.read_results(tmp_path / "results_gpu.csv")
    c = _config(exe, tmp_path, "1", "0", "0", full=True)
    dt = c["dTime"]
    x = dt / 0.01
    a = 1 - x + x * x / 2 if scheme == "maccormack" else 1 - x
    i = np.arange(0, c["nStep"] + 1, 50)
    assert len(rows) == len(i)
    np.testing.assert_allclose(rows[:, 0], i * dt, rtol=1e-6)
    np.testing.assert_allclose(rows[:, 2], -a ** i, rtol=5e-5)  # FP32 rounding of a uniform field
    assert np.abs(rows[:, 1]).max() < 1e-6 and np.abs(rows[:, 3]).max() < 1e-6
    assert os.path.exists(tmp_path / "results_NMR" / "results_grad_1_0_0.csv")


@pytest.mark.gpu
def test_cpmg_echo_train_closed_form(exe, tmp_path):
    """CPMG (20 echoes on the GPU path): each 180 deg pulse about x flips My;
    between pulses My decays by the MacCormack factor 1 - x + x^2/2 per
    update, x = dt/T2."""
    shape = (6, 6, 6)
    T1s, T2 = 2e-4, 2e-3
    _write_case(tmp_path, shape, None, None, NPOR=0, NSEQUENCE=1, TIME1=T1s, T2_WAT=T2, Printing_on=7)
    _run(exe, tmp_path, "1", "0", "0")
    rows = refio.read_results(tmp_path / "results_gpu.csv")
    c = _config(exe, tmp_path, "1", "0", "0", full=True)
    dt = c["dTime"]
    # every step whose time lies in a window [(2e+1) T1, (2e+1) T1 + dt) gets a
    # 180 deg pulse (no flag gating in the CPMG branch): a window that rounds
    # to two steps pulses twice, as the reference does
    flips = [k for e in range(20) for k in range(c["nStep"] + 1)
             if (k * dt < (2 * e + 1) * T1s + dt) and (k * dt >= (2 * e + 1) * T1s)]
    assert len(flips) >= 5
    i = np.arange(0, c["nStep"] + 1, 7)
    sign = np.array([(-1) ** sum(1 for m in flips if m <= k) for k in i])
    x = dt / T2
    np.testing.assert_allclose(rows[:, 2], -sign * (1 - x + x * x / 2) ** i, rtol=5e-5, atol=1e-7)


@pytest.mark.gpu
@pytest.mark.parametrize("cadence,scheme,nb", [
    ("gpu", None, (1, 1, 1, 1, 1, 1)),
    ("cpu", None, (1, 1, 1, 1, 1, 1)),
    ("gpu", "upwind", (1, 1, 1, 1, 1, 1)),
    ("gpu", None, (3, 3, 3, 3, 1, 1)),
    ("cpu", None, (2, 3, 1, 3, 3, 3)),
])
def test_pfg_flow_against_fp64_loop(exe, tmp_path, cadence, scheme, nb):
    """Porous box with divergence-free flow, surface relaxation and a pulsed-
    gradient stimulated echo (sequence 2): every CSV row of solve_gpu()
    (results_gpu.csv, MacCormack unless --scheme upwind) or solve()
    (results_cpu.csv, upwind) matches the FP64 restatement of the reference's
    loop to the CSV's 7 digits, with zero-derivative, periodic or zero-value
    ghosts from the NB flags."""
    shape = (12, 10, 14)
    mask = generators.porosity(shape, 0.55, seed=6, sigma=1.2)
    vel = generators.velocity(mask, seed=6, sigma=1.5, vmax=5e-3)
    table = np.linspace(-0.03, 0.03, 399)
    nbk = {f"NB{i + 1}": v for i, v in enumerate(nb)}
    cfg = _write_case(tmp_path, shape, mask, vel, grad_table=table, NSEQUENCE=2, TIME1=2e-4, TIME2=4e-4,
                      R2_WAT=2e-5, R1_WAT=1e-5, T1_WAT=0.05, Printing_on=5, BX0=0.0, BZ0=1.0, **nbk)
    cfg["gradientPulse"] = float(table[349])
    args = ["350", "0", "0"] + (["--cpu-cadence"] if cadence == "cpu" else [])
    if scheme:
        args += ["--scheme", scheme]
    _run(exe, tmp_path, *args)
    rows = refio.read_results(tmp_path / f"results_{cadence}.csv")
    sch = {None: None, "upwind": oracle.UPWIND}[scheme]
    ref, dt, nstep = oracle.run_dropin(cfg, mask, vel, gpu=(cadence == "gpu"), scheme=sch)
    assert rows.shape == ref.shape, (rows.shape, ref.shape)
    np.testing.assert_allclose(rows[:, 0], ref[:, 0], rtol=2e-6)
    scale = np.abs(ref[:, 1:]).max()
    err = np.abs(rows[:, 1:] - ref[:, 1:]).max() / scale
    print(f"[dropin] sequence 2, {cadence} cadence, scheme {scheme or 'default'}, NB {nb}: {len(rows)} rows, "
          f"max error {err:.2e} of the signal scale")
    assert err < 5e-6


@pytest.mark.gpu
@pytest.mark.parametrize("cadence", ["gpu", "cpu"])
def test_noise_field_against_fp64_loop(exe, tmp_path, cadence):
    """AMPL_NOISE with NDIRECT_FIELD = 3: FillInMainField draws one glibc
    rand() per padded cell for the field along z (cpp:1973-1978), the field
    then dephases the spins every step; sequence 2 with flow on both loops,
    every CSV row against the FP64 restatement (the oracle draws the same
    sequence from its own restatement of the generator)."""
    shape = (9, 10, 12)
    mask = generators.porosity(shape, 0.55, seed=14, sigma=1.2)
    vel = generators.velocity(mask, seed=14, sigma=1.5, vmax=3e-3)
    table = np.linspace(-0.03, 0.03, 399)
    cfg = _write_case(tmp_path, shape, mask, vel, grad_table=table, NSEQUENCE=2, TIME1=2e-4, TIME2=4e-4,
                      R2_WAT=2e-5, R1_WAT=1e-5, T1_WAT=0.05, Printing_on=5, AMPL_NOISE=2e-5, NVECTOR=1)
    cfg["gradientPulse"] = float(table[349])
    args = ["350", "0", "0"] + (["--cpu-cadence"] if cadence == "cpu" else [])
    _run(exe, tmp_path, *args)
    rows = refio.read_results(tmp_path / f"results_{cadence}.csv")
    ref, dt, nstep = oracle.run_dropin(cfg, mask, vel, gpu=(cadence == "gpu"))
    assert rows.shape == ref.shape, (rows.shape, ref.shape)
    np.testing.assert_allclose(rows[:, 0], ref[:, 0], rtol=2e-6)
    scale = np.abs(ref[:, 1:]).max()
    err = np.abs(rows[:, 1:] - ref[:, 1:]).max() / scale
    quiet, _, _ = oracle.run_dropin(dict(cfg, AMPL_NOISE=0.0), mask, vel, gpu=(cadence == "gpu"))
    effect = np.abs(quiet[:, 1:] - ref[:, 1:]).max() / scale if quiet.shape == ref.shape else np.inf
    print(f"[dropin] noise field, {cadence} cadence: {len(rows)} rows, max error {err:.2e} of the signal scale "
          f"(the field moves the rows by {effect:.2e})", flush=True)
    assert err < 5e-6
    assert effect > 1e-3


@pytest.mark.parametrize("nb,modes", [
    ((1, 1, 1, 1, 1, 1), [0, 0, 0]),
    ((3, 1, 3, 1, 3, 3), [1, 0, 1]),
    ((2, 3, 1, 3, 1, 1), [2, 1, 0]),
    ((3, 1, 2, 1, 1, 1), [2, 0, 0]),
])
def test_boundary_flags_map_like_the_gpu_ghost_fill(exe, tmp_path, nb, modes):
    """NB1/NB3 (x), NB2/NB4 (y), NB5/NB6 (z) select the ghost fill as
    BoundaryConditionsForVectorField_gpu applies it (cu:1328-1396)."""
    _write_case(tmp_path, (4, 5, 6), None, None, NPOR=0, **{f"NB{i + 1}": v for i, v in enumerate(nb)})
    c = _config(exe, tmp_path, "1", "0", "0", full=True)
    assert c["bc"] == modes == [int(m) for m in oracle.nb_modes(nb)]


def test_refuses_boundary_flags_that_leave_ghosts_stale(exe, tmp_path):
    """x with NB1 = 1, NB3 = 3 matches no rule of the reference's ghost fill:
    init() says so and exits 1."""
    _write_case(tmp_path, (4, 5, 6), None, None, NPOR=0, NB1=1, NB3=3)
    r = _run(exe, tmp_path, "1", "0", "0", check=False)
    assert r.returncode == 1 and "ghosts undefined" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("cadence", ["gpu", "cpu"])
def test_two_phase_flow_against_fp64_loop(exe, tmp_path, cadence):
    """Two-phase mixture (NCOMP=2, n1_init.dat) with face-wise water
    relaxivities and flow, sequence 2: the 7-column results_*.csv rows
    (SumX1..SumZ1, SumX2..SumZ2) match the FP64 restatement of the
    reference's loop."""
    shape = (8, 9, 11)
    mask = generators.porosity(shape, 0.6, seed=9, sigma=1.2)
    vel = generators.velocity(mask, seed=9, sigma=1.5, vmax=4e-3)
    rng = np.random.default_rng(9)
    n1 = np.clip(generators._smooth_noise(shape, 9, 5, 1.5) * 4.0 + 0.5, 0.0, 1.0)
    nz, ny, nx = shape
    surf = {f"SurfRel{r}Water{a}": rng.uniform(1e-6, 3e-5, {"X": (nz, ny, nx + 1), "Y": (nz, ny + 1, nx),
                                                              "Z": (nz + 1, ny, nx)}[a])
            for r in "12" for a in "XYZ"}
    table = np.linspace(-0.03, 0.03, 399)
    cfg = _write_case(tmp_path, shape, mask, vel, grad_table=table, n1=n1, surf=surf, NCOMP=2, nSurfRelWater=1,
                      NSEQUENCE=2, TIME1=2e-4, TIME2=4e-4, R1_OIL=4e-6, R2_OIL=8e-6, T1_OIL=0.3, T2_OIL=0.05,
                      T1_WAT=0.6, T2_WAT=0.2, D_OIL=4e-10, Printing_on=5, BX0=0.0, BZ0=1.0)
    cfg["gradientPulse"] = float(table[349])
    args = ["350", "0", "0"] + (["--cpu-cadence"] if cadence == "cpu" else [])
    _run(exe, tmp_path, *args)
    rows = refio.read_results(tmp_path / f"results_{cadence}.csv")
    ref, dt, nstep = oracle.run_dropin(cfg, mask, vel, gpu=(cadence == "gpu"), n1=n1, surf=surf)
    assert rows.shape == ref.shape and rows.shape[1] == 7, (rows.shape, ref.shape)
    scale = np.abs(ref[:, 1:]).max()
    err = np.abs(rows[:, 1:] - ref[:, 1:]).max() / scale
    print(f"[dropin] two-phase sequence 2, {cadence} cadence: {len(rows)} rows, max error {err:.2e} of the scale")
    assert err < 5e-6


@pytest.mark.gpu
@pytest.mark.parametrize("nseq", [5, 9, 98, 99])
def test_other_gpu_sequences_against_fp64_loop(exe, tmp_path, nseq):
    """The remaining Sequence_gpu programs (cpp:2681-3146) on a porous box
    with flow: 5 (one 180 deg pulse), 9 (90 - gradient - 180 - gradient) and
    98 / 99 (the multiple-gradient pulses of flag 3); every results_gpu.csv
    row matches the FP64 restatement of the reference's GPU loop."""
    shape = (10, 9, 12)
    mask = generators.porosity(shape, 0.55, seed=11, sigma=1.2)
    vel = generators.velocity(mask, seed=11, sigma=1.5, vmax=3e-3)
    table = np.linspace(-0.02, 0.02, 399)
    cfg = _write_case(tmp_path, shape, mask, vel, grad_table=table, NSEQUENCE=nseq, TIME1=2e-4, TIME2=4e-4,
                      R2_WAT=2e-5, R1_WAT=1e-5, T1_WAT=0.05, Printing_on=4)
    args = ["350", "200", "100"]
    c = _config(exe, tmp_path, *args, full=True)
    cfg["gradientPulse"] = c["gradientPulse"]
    cfg["multipleGradients"] = tuple(c["multipleGradients"])
    _run(exe, tmp_path, *args)
    rows = refio.read_results(tmp_path / "results_gpu.csv")
    ref, dt, nstep = oracle.run_dropin(cfg, mask, vel, gpu=True)
    assert rows.shape == ref.shape, (rows.shape, ref.shape)
    np.testing.assert_allclose(rows[:, 0], ref[:, 0], rtol=2e-6)
    scale = np.abs(ref[:, 1:]).max()
    err = np.abs(rows[:, 1:] - ref[:, 1:]).max() / scale
    print(f"[dropin] sequence {nseq}: {len(rows)} rows, max error {err:.2e} of the signal scale", flush=True)
    assert err < 5e-6

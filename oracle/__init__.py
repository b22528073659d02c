"""FP64 CPU oracle of the reference's Bloch-Torrey update -- TEST
INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only as the checker (or the timed CPU baseline).  The
product (initializator_amd/, libnmr_hip.so) never imports or links it.

ctypes wrapper of oracle/libnmr_oracle.so (oracle/nmr_oracle.c), which
restates src/NmrRelaxationSolve.cpp of the reference in FP64 on its padded
layout.  The pinning status of every function is listed at the top of
nmr_oracle.c; the pinning tests live in tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_long

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnmr_oracle.so")

G_REF = 4.78941714e7


class OrParams(ctypes.Structure):
    _fields_ = [
        ("ik", c_int32),
        ("jk", c_int32),
        ("kk", c_int32),
        ("precession", c_int32),
        ("velocity", c_int32),
        ("diffusion", c_int32),
        ("dx", c_double),
        ("dy", c_double),
        ("dz", c_double),
        ("dt", c_double),
        ("gamma", c_double),
        ("p_eq", c_double),
        ("b0", c_double * 3),
        ("scheme", c_int32),
        ("bc", c_int32 * 3),
        ("advect_mask", c_int32),
    ]


class OrFluids(ctypes.Structure):
    """Water (index 0) and oil (1) properties and the two-phase projection."""

    _fields_ = [
        ("d", c_double * 2),
        ("t1", c_double * 2),
        ("t2", c_double * 2),
        ("r1", c_double * 2),
        ("r2", c_double * 2),
        ("ncomp", c_int32),
        ("reserved", c_int32),
        ("an1", c_double),
        ("an2", c_double),
        ("bn1", c_double),
        ("bn2", c_double),
    ]


def fluids(d=(2.3e-9, 2.1e-10), t1=(2.5e10, 2.5e10), t2=(2.0, 2.0), r1=(0.0, 0.0), r2=(0.0, 0.0), ncomp=2,
           an=(1.0, 0.0), bn=(0.0, 1.0)) -> OrFluids:
    f = OrFluids()
    for i in range(2):
        f.d[i], f.t1[i], f.t2[i], f.r1[i], f.r2[i] = d[i], t1[i], t2[i], r1[i], r2[i]
    f.ncomp = int(ncomp)
    f.an1, f.an2 = an
    f.bn1, f.bn2 = bn
    return f


UPWIND, MACCORMACK = 0, 1
BC_NEUMANN, BC_PERIODIC, BC_DIRICHLET = 0, 1, 2


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        dp = POINTER(c_double)
        cp = c_char_p
        P = POINTER(OrParams)
        L.or_num_cells.restype = c_long
        L.or_num_cells.argtypes = [P]
        L.or_bc_mask.argtypes = [P, cp]
        L.or_bc_field.argtypes = [P, dp]
        L.or_fill_coefficients.argtypes = [P, cp, c_double, c_double, c_double, c_double, c_double,
                                           dp, dp, dp, dp, dp, dp, dp, dp, dp, dp]
        L.or_fill_gradient.argtypes = [P, c_double, c_double, c_double, dp, dp, dp]
        L.or_glibc_rand.argtypes = [ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
        L.or_set_threads.restype = ctypes.c_int
        L.or_set_threads.argtypes = [ctypes.c_int]
        L.or_glibc_rand.restype = None
        L.or_fill_noise.argtypes = [P, c_double, dp]
        L.or_fill_noise.restype = None
        L.or_step_time.restype = c_double
        L.or_step_time.argtypes = [c_double] * 7
        L.or_explicit_simple.argtypes = [P, cp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp]
        L.or_output.restype = c_double
        L.or_output.argtypes = [P, cp, c_int, c_int, c_int, dp, POINTER(c_int32)]
        L.or_add_impulse.argtypes = [P, cp, c_int, c_double, c_double, c_double, c_double, dp, dp, dp]
        L.or_run.argtypes = [P, cp, dp, dp, dp, dp, dp, dp, c_int64, dp, dp, c_int, c_int, c_int,
                             dp, dp, dp, dp]
        L.or_maccormack.argtypes = [P, cp] + [dp] * 18
        F = POINTER(OrFluids)
        L.or_phase.argtypes = [F, c_double, c_double, dp, dp]
        L.or_fill_coefficients_2f.argtypes = [P, cp, F, dp, dp, POINTER(dp), dp, dp, dp, dp]
        L.or_output_2f.argtypes = [P, cp, c_int, c_int, c_int, F, dp, dp, dp, dp, dp]
        _lib = L
    return _lib


def set_threads(n: int = 0) -> int:
    """Threads of the oracle's OpenMP loops from now on (n > 0); returns the
    count in force."""
    return int(lib().or_set_threads(int(n)))


def _d(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(POINTER(c_double))


def _c(a):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return ctypes.cast(a.ctypes.data, c_char_p)


def params(shape, dx, dt, *, dy=None, dz=None, precession=False, velocity=False, diffusion=True,
           gamma=G_REF, p_eq=0.0, b0=(0.0, 0.0, 1.0), scheme=UPWIND, bc=(0, 0, 0),
           advect_mask=False) -> OrParams:
    nz, ny, nx = shape
    p = OrParams()
    p.scheme = int(scheme)
    p.advect_mask = int(bool(advect_mask))
    for i in range(3):
        p.bc[i] = int(bc[i])
    p.ik, p.jk, p.kk = nx, ny, nz
    p.precession, p.velocity, p.diffusion = int(precession), int(velocity), int(diffusion)
    p.dx = dx
    p.dy = dy if dy is not None else dx
    p.dz = dz if dz is not None else dx
    p.dt = dt
    p.gamma = gamma
    p.p_eq = p_eq
    for i in range(3):
        p.b0[i] = b0[i]
    return p


def padded_shape(p: OrParams):
    return (p.kk + 2, p.jk + 2, p.ik + 2)


def bc_mask(p, act):
    lib().or_bc_mask(ctypes.byref(p), _c(act))


def bc_field(p, f):
    lib().or_bc_field(ctypes.byref(p), _d(f))


def fill_coefficients(p, act, d_water, t1_water, t2_water, r1_water, r2_water, faces=None):
    """D, T1, T2 (padded) and (Dmax, Tmin) as FillInDiffusinAndTime leaves them."""
    shp = padded_shape(p)
    D = np.zeros(shp)
    t1 = np.zeros(shp)
    t2 = np.zeros(shp)
    out2 = np.zeros(2)
    f = faces or (None,) * 6
    lib().or_fill_coefficients(ctypes.byref(p), _c(act), d_water, t1_water, t2_water, r1_water, r2_water,
                               *[_d(x) for x in f], _d(D), _d(t1), _d(t2), _d(out2))
    return D, t1, t2, (out2[0], out2[1])


def phase(fl, x0, x1):
    a, b = c_double(), c_double()
    lib().or_phase(ctypes.byref(fl), float(x0), float(x1), ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def fill_coefficients_2f(p, act, fl, conc=None, faces=None):
    """Two-fluid D, T1, T2 (padded) and (Dmax, Tmin).  conc: (conc0, conc1)
    padded or None (water only); faces: 12 padded arrays or None entries
    (1Water x,y,z, 2Water x,y,z, 1Oil x,y,z, 2Oil x,y,z)."""
    shp = padded_shape(p)
    D, t1, t2, out2 = np.zeros(shp), np.zeros(shp), np.zeros(shp), np.zeros(2)
    fa = (POINTER(c_double) * 12)(*[_d(x) if x is not None else None for x in (faces or [None] * 12)])
    c0, c1 = (None, None) if conc is None else conc
    lib().or_fill_coefficients_2f(ctypes.byref(p), _c(act), ctypes.byref(fl), _d(c0), _d(c1), fa, _d(D), _d(t1),
                                  _d(t2), _d(out2))
    return D, t1, t2, (out2[0], out2[1])


def output_2f(p, act, buf, fl, conc, f):
    a, b = c_double(), c_double()
    lib().or_output_2f(ctypes.byref(p), _c(act), buf[0], buf[1], buf[2], ctypes.byref(fl), _d(conc[0]),
                       _d(conc[1]), _d(f), ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def fill_gradient(p, g):
    shp = padded_shape(p)
    nf = [np.zeros(shp) for _ in range(3)]
    lib().or_fill_gradient(ctypes.byref(p), g[0], g[1], g[2], *[_d(a) for a in nf])
    return nf


def glibc_rand(n):
    """The first n outputs of glibc's rand() from its default state."""
    out = np.empty(n, dtype=np.int32)
    lib().or_glibc_rand(int(n), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return out


def fill_noise(p, ampl, nf_axis):
    """Overwrite one padded field component with the reference's noise field."""
    lib().or_fill_noise(ctypes.byref(p), float(ampl), _d(nf_axis))


def step_time(dt_noise, dt_grad, dt_vel, dt_diff, dt_rel, time1, time2):
    return lib().or_step_time(dt_noise, dt_grad, dt_vel, dt_diff, dt_rel, time1, time2)


def explicit_simple(p, act, D, t1, t2, vel, nf, m, a):
    """One ExplicitSimple step in place on m (list of 3 padded arrays)."""
    lib().or_explicit_simple(ctypes.byref(p), _c(act), _d(D), _d(t1), _d(t2), *[_d(v) for v in vel],
                             *[_d(x) for x in nf], *[_d(x) for x in m], *[_d(x) for x in a])


def maccormack(p, act, D, t1, t2, vel, nf, m, pc):
    """One MacCormack step in place on m; pc: six zero-initialised padded
    scratch arrays (predictor, corrector) kept across steps."""
    lib().or_maccormack(ctypes.byref(p), _c(act), _d(D), _d(t1), _d(t2), *[_d(v) for v in vel],
                        *[_d(x) for x in nf], *[_d(x) for x in m], *[_d(x) for x in pc])


def add_impulse(p, act, flag, param, axis, m):
    """AddImpulse in place on m (list of 3 padded arrays)."""
    lib().or_add_impulse(ctypes.byref(p), _c(act), int(flag), float(param), float(axis[0]), float(axis[1]),
                         float(axis[2]), *[_d(x) for x in m])


def output(p, act, buf, f):
    n = c_int32()
    s = lib().or_output(ctypes.byref(p), _c(act), buf[0], buf[1], buf[2], _d(f), ctypes.byref(n))
    return s, n.value


def run(p, act, D, t1, t2, vel, nsteps, m, grad_sched=None, grad_const=None, buf=(0, 0, 0)):
    """nsteps of update + ghost fill + signal; m updated in place.
    Returns the (nsteps+1, 3) signal."""
    sig = np.zeros((nsteps + 1, 3))
    gs = None if grad_sched is None else np.ascontiguousarray(grad_sched, dtype=np.float64)
    gc = None if grad_const is None else np.ascontiguousarray(grad_const, dtype=np.float64)
    lib().or_run(ctypes.byref(p), _c(act), _d(D), _d(t1), _d(t2), *[_d(v) for v in vel], int(nsteps),
                 _d(gs), _d(gc), buf[0], buf[1], buf[2], *[_d(x) for x in m], _d(sig))
    return sig


def setup_case(case, p_eq=0.0, m0=None, scheme=UPWIND, bc=(0, 0, 0), advect_mask=False):
    """Padded oracle inputs for a generators.Case (one fluid, uniform
    relaxivity): params, mask, D, T1, T2, velocity, initial M."""
    from initializator_amd import layout  # conversions only

    has_v = case.vx is not None
    prec = case.grad_sched is not None
    p = params(case.shape, case.dx, case.dt, precession=prec, velocity=has_v, p_eq=p_eq, scheme=scheme, bc=bc,
               advect_mask=advect_mask)
    act = layout.pad_mask(case.mask)
    D, t1, t2, _ = fill_coefficients(p, act, case.diffusivity, case.t1_bulk, case.t2_bulk, case.rho, case.rho)
    if has_v:
        vel = [np.ascontiguousarray(v) for v in layout.pad_velocity(case.vx, case.vy, case.vz, case.shape)]
    else:
        vel = [np.zeros(padded_shape(p)) for _ in range(3)]
    if m0 is None:
        mx = np.where(case.mask != 0, 1.0, 0.0)
        m0 = [mx, np.zeros(case.shape), np.zeros(case.shape)]
    m = [np.ascontiguousarray(layout.pad_field(x)) for x in m0]
    return p, act, D, t1, t2, vel, m


def run_case(case, nsteps=None, p_eq=0.0, m0=None, buf=(0, 0, 0), scheme=UPWIND, bc=(0, 0, 0), advect_mask=False):
    """Oracle run of a Case: (signal (nsteps+1, 3), compact final fields)."""
    from initializator_amd import layout

    n = case.nsteps if nsteps is None else nsteps
    p, act, D, t1, t2, vel, m = setup_case(case, p_eq=p_eq, m0=m0, scheme=scheme, bc=bc, advect_mask=advect_mask)
    gs = None if case.grad_sched is None else case.grad_sched[:n]
    sig = run(p, act, D, t1, t2, vel, n, m, grad_sched=gs, buf=buf)
    return sig, [layout.unpad(x) for x in m]


# ---------------------------------------------------------------------------
# The reference's driver loops, restated for checking the drop-in executable.

PI = 3.1415926535897932384626433832795028841971
SMALL = 1e-50


def sequence_events(gpu, nseq, t, dt, T1, T2, d, gp, mg, flags):
    """Pulses (flag, param, axis) at time t: Sequence_gpu (cpp:2681-3146) when
    gpu, else the CPU Sequence (cpp:2610-2679).  flags: 5-element list."""
    ev = []

    def win(t0):
        return (t < t0 + dt) and (t >= t0)

    ninety = (2, 0.5 * PI, (d[2], d[0], d[1]))
    if nseq == 0:
        if win(0.0):
            ev.append(ninety)
    elif nseq == 1:
        if win(0.0):
            ev.append(ninety)
        for i in range(20 if gpu else 100):
            if win((2 * i + 1) * T1):
                ev.append((2, PI, (d[2], d[0], d[1])))
    elif nseq == 2:
        neg = (2, 0.5 * PI, (-d[2], d[0], d[1])) if gpu else ninety
        gpl = (1, gp, (d[0], d[1], d[2]))
        for n, (t0, e) in enumerate([(0.0, neg), (T1, gpl), (2.0 * T1, ninety), (2.0 * T1 + T2, neg),
                                     (3.0 * T1 + T2, gpl)]):
            if win(t0) and flags[n] == 0:
                ev.append(e)
                flags[n] = 1
    elif gpu and nseq == 9:
        gpl = (1, gp, (d[0], d[1], d[2]))
        for n, (t0, e) in enumerate([(0.0, ninety), (T1, gpl), (T1 + T2 / 2, (2, PI, (d[2], d[0], d[1]))),
                                     (T2 + T1, gpl)]):
            if win(t0) and flags[n] == 0:
                ev.append(e)
                flags[n] = 1
    elif gpu and nseq == 5:
        if win(0.0):
            ev.append((2, PI, (d[2], d[0], d[1])))
    elif gpu and nseq in (98, 99):
        seq = [(0.0, (2, 0.5 * PI, (0, -1, 0)) if nseq == 98 else None), (T1, (3, 1, tuple(mg))),
               (T1 + T2 / 2, (2, PI, (0, 0, 1)) if nseq == 98 else None),
               (T2 + T1, (3, 1, tuple(mg) if nseq == 98 else tuple(-x for x in mg)))]
        for n, (t0, e) in enumerate(seq):
            if win(t0) and flags[n] == 0:
                if e is not None:
                    ev.append(e)
                flags[n] = 1
    return ev


def _chain(dx, dy, dz, f):
    """The spacing selection of cpp:1585-1587 / 1853-1855 (last true wins)."""
    out = None
    if dx <= dy and dx <= dz:
        out = f(dx)
    if dy <= dy and dy <= dz:
        out = f(dy)
    if dz <= dy and dz <= dz:
        out = f(dz)
    return out


def nb_modes(nb):
    """Ghost fill per axis from NB1..NB6 by the rules of
    BoundaryConditionsForVectorField_gpu (cu:1328-1396): x zero derivative
    when NB1 = NB3 = 1, periodic when both are 3, zero value when either is 2
    (applied last); y (NB2, NB4) and z (NB5, NB6) zero derivative or periodic
    when both agree.  None where no rule refreshes the ghosts."""
    x = None
    if nb[0] == 1 and nb[2] == 1:
        x = BC_NEUMANN
    if nb[0] == 3 and nb[2] == 3:
        x = BC_PERIODIC
    if nb[0] == 2 or nb[2] == 2:
        x = BC_DIRICHLET

    def pair(lo, hi):
        return BC_NEUMANN if lo == hi == 1 else (BC_PERIODIC if lo == hi == 3 else None)

    return (x, pair(nb[1], nb[3]), pair(nb[4], nb[5]))


def run_dropin(cfg, mask, vel_raw, gpu=True, scheme=None, n1=None, surf=None):
    """The reference's solve_gpu (gpu=True, cu:1187-1227, MacCormack update)
    or solve (cpp:35-66, ExplicitSimple update) loop in FP64 for one fluid:
    returns (rows (n, 4), dTime, nStep).  scheme overrides the update.  The
    NB flags select the ghost fill in both loops.
    cfg: dict of the data-file values (see initializator_amd.refio)."""
    import math

    from initializator_amd import layout

    nz, ny, nx = mask.shape
    dx, dy, dz = cfg["DX"], cfg["DY"], cfg["DZ"]
    fluid = cfg.get("NFLUID", 1)
    c1, c2 = (1.0, 0.0) if fluid == 1 else (0.0, 1.0)
    Dw = c1 * cfg["D_WAT"] + c2 * cfg["D_OIL"]
    T1b = 1.0 / (c1 / cfg["T1_WAT"] + c2 / cfg["T1_OIL"])
    T2b = 1.0 / (c1 / cfg["T2_WAT"] + c2 / cfg["T2_OIL"])
    r1 = c1 * cfg["R1_WAT"] + c2 * cfg["R1_OIL"]
    r2 = c1 * cfg["R2_WAT"] + c2 * cfg["R2_OIL"]
    ndf = cfg["NDIRECT_FIELD"]
    bs = (cfg["BX0"], cfg["BY0"], cfg["BZ0"])
    b0 = [bs[i] if (ndf == i + 1 or ndf == 0) else 0.0 for i in range(3)]
    btot = math.sqrt(sum(x * x for x in bs))
    d = [(1.0 if ndf == i + 1 else 0.0) if ndf != 0 else bs[i] / btot for i in range(3)]
    buf = (cfg["bufferX"], cfg["bufferY"], cfg["bufferZ"])
    peq = 1.0 / ((nx - 2 * buf[0]) * (ny - 2 * buf[1]) * (nz - 2 * buf[2]))
    if scheme is None:
        scheme = MACCORMACK if gpu else UPWIND
    bc = nb_modes([cfg.get(f"NB{i}", 1) for i in range(1, 7)])
    p = params(mask.shape, dx, 1.0, dy=dy, dz=dz, precession=cfg["NVECTOR"] == 1, velocity=cfg["NVELOC"] == 1,
               diffusion=cfg["NDIFFUS"] == 1, p_eq=peq, b0=b0, scheme=scheme, bc=bc)
    act = layout.pad_mask(mask)
    two = cfg.get("NCOMP", 1) == 2
    conc = None
    if two or surf is not None:
        # mixture / face-wise relaxivity: FillInDiffusinAndTime with both fluids
        fprops = fluids(d=(cfg["D_WAT"], cfg["D_OIL"]), t1=(cfg["T1_WAT"], cfg["T1_OIL"]), t2=(cfg["T2_WAT"], cfg["T2_OIL"]),
                    r1=(cfg["R1_WAT"], cfg["R1_OIL"]), r2=(cfg["R2_WAT"], cfg["R2_OIL"]), ncomp=2 if two else 1,
                    an=(cfg["AN(1)"], cfg["AN(2)"]), bn=(cfg["BN(1)"], cfg["BN(2)"]))
        if two:
            n1f = np.asarray(n1, dtype=np.float32).astype(np.float64)  # read into a float (cpp:976-1006)
            conc = (np.ascontiguousarray(layout.pad_field(n1f)), np.ascontiguousarray(layout.pad_field(1.0 - n1f)))
            for cc in conc:  # ghosts hold 0 (cpp:935-944)
                cc[0, :, :] = cc[-1, :, :] = 0.0
                cc[:, 0, :] = cc[:, -1, :] = 0.0
                cc[:, :, 0] = cc[:, :, -1] = 0.0
        else:
            one = np.zeros(padded_shape(p))
            one[1:-1, 1:-1, 1:-1] = 1.0
            conc = (one, np.zeros(padded_shape(p))) if fluid == 1 else (np.zeros(padded_shape(p)), one)
        faces = [None] * 12
        for rel in (0, 1):
            for fi, fname in enumerate(("Water", "Oil")):
                keys = [f"SurfRel{rel + 1}{fname}{a}" for a in "XYZ"]
                if surf is not None and all(k in surf for k in keys):
                    arrs = layout.pad_velocity(*[surf[k] for k in keys], mask.shape)
                    for ax in range(3):
                        faces[fi * 6 + rel * 3 + ax] = np.ascontiguousarray(arrs[ax])
        D, t1, t2, (dmax, tmin) = fill_coefficients_2f(p, act, fprops, conc, faces)
    else:
        D, t1, t2, (dmax, tmin) = fill_coefficients(p, act, Dw, T1b, T2b, r1, r2)

    def row(t):
        if two:
            s1, s2 = zip(*[output_2f(p, act, buf, fprops, conc, m[q]) for q in range(3)])
            return [t] + list(s1) + list(s2)
        return [t] + [output(p, act, buf, m[q])[0] for q in range(3)]
    # step limits (cpp:1583-1587, 1849-1864, 1980-1988, 2519-2534)
    cfl, cfl1, g = cfg["CFL"], cfg["CFL1"], G_REF
    if cfg["NVELOC"] == 1:
        vmax = max(0.0, max(float(v.max()) for v in vel_raw))
        vmin = min(0.0, min(float(v.min()) for v in vel_raw))
        vm = vmax * cfg["VEL_MULT"] if vmax > abs(vmin) else abs(vmin) * cfg["VEL_MULT"]
        dt_vel = _chain(dx, dy, dz, lambda h: cfl * h / (vm + SMALL))
    else:
        dt_vel = 1.0
    dt_diff = _chain(dx, dy, dz, lambda h: 2.0 * cfl * h * h / dmax) if cfg["NDIFFUS"] == 1 else 1.0
    dt_rel = cfl1 * tmin
    grads = (cfg["GRAD_X"], cfg["GRAD_Y"], cfg["GRAD_Z"])
    ga = SMALL
    for gval, h, n in zip(grads, (dx, dy, dz), (nx, ny, nz)):
        if gval != 0:
            ga = abs(gval * h * n)
    noise = cfg["AMPL_NOISE"]
    dt_noise = cfl1 / (g * noise) if noise != 0 else math.inf
    dt_grad = cfl1 / (g * ga)
    T1s, T2s = cfg["TIME1"], cfg["TIME2"]
    dt = step_time(dt_noise, dt_grad, dt_vel, dt_diff, dt_rel, T1s, T2s)
    nseq = cfg["NSEQUENCE"]
    end = {0: 2 * cfg["T2_WAT"], 1: 2 * cfg["T2_WAT"], 2: 4.0 * T1s + T2s}.get(nseq, T2s + 2.0 * T1s)
    nstep = int(end / dt) + 1
    p.dt = dt
    if cfg["NVELOC"] == 1:
        vel = [np.ascontiguousarray(x * cfg["VEL_MULT"]) for x in layout.pad_velocity(*vel_raw, mask.shape)]
    else:
        vel = [np.zeros(padded_shape(p)) for _ in range(3)]
    pad_pore = act == ord("1")
    m = [np.where(pad_pore, peq * d[q], 0.0) for q in range(3)]
    nf = fill_gradient(p, grads if cfg["NVECTOR"] == 1 else (0.0, 0.0, 0.0))
    if noise != 0.0 and ndf != 0:  # FillInMainField's noise replaces that axis's gradient field
        fill_noise(p, noise, nf[ndf - 1])
    a = [x.copy() for x in m]
    pc = [np.zeros(padded_shape(p)) for _ in range(6)]
    for x in m:
        bc_field(p, x)
    flags = [0] * 5
    gp = cfg.get("gradientPulse", 0.0)
    mg = cfg.get("multipleGradients", (0.0, 0.0, 0.0))
    rows = []
    ndp = cfg["Printing on"]
    for i in range(nstep + 1):
        t = i * dt
        if not gpu and i % ndp == 0:
            rows.append(row(t))
        if scheme == MACCORMACK:
            maccormack(p, act, D, t1, t2, vel, nf, m, pc)
        else:
            explicit_simple(p, act, D, t1, t2, vel, nf, m, a)
        for fl, par, ax in sequence_events(gpu, nseq, t, dt, T1s, T2s, d, gp, mg, flags):
            add_impulse(p, act, fl, par, ax, m)
        for x in m:
            bc_field(p, x)
        if gpu and i % ndp == 0:
            rows.append(row(t))
    return np.array(rows), dt, nstep

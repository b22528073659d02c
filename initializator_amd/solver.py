"""Python host wrapper over the C ABI: one HipSolver = one nmr_ctx.

This mirrors what the reference's solve_gpu() does per run
(src/NmrRelaxationSolveKernels.cu:1056-1319): allocate and upload once,
step on the device, reduce the signal every step, download at the end.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import NmrProblemC, check
from .generators import G_REF, Case


def _dptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _addr(a):
    """Address of an array's data as a plain integer (for void* arguments):
    cheaper and steadier than a ctypes pointer object on a hot call."""
    return None if a is None else a.__array_interface__["data"][0]


class HipSolver:
    """Device-resident Bloch-Torrey solve of one problem."""

    def __init__(self, shape, ncomp: int, dx: float, dt: float, diffusivity: float, r2_bulk: float,
                 r1_bulk: float = 0.0, rho1: float = 0.0, rho2: float = 0.0, gamma: float = G_REF,
                 p_eq: float = 0.0, b0_dir=(0.0, 0.0, 1.0), grad=(0.0, 0.0, 0.0), diffusion: bool = True,
                 velocity: bool = False, precession: bool = False, phase_exact: bool = False,
                 buffer=(0, 0, 0), dy: float | None = None, dz: float | None = None,
                 scheme: int = _lib.NMR_SCHEME_UPWIND, bc=(0, 0, 0), advect_mask: bool = False):
        self.lib = _lib.load()
        nz, ny, nx = (int(s) for s in shape)
        self.shape = (nz, ny, nx)
        self.ncomp = int(ncomp)
        p = NmrProblemC()
        p.nx, p.ny, p.nz = nx, ny, nz
        p.ncomp = self.ncomp
        p.flags = ((_lib.NMR_F_DIFFUSION if diffusion else 0) | (_lib.NMR_F_VELOCITY if velocity else 0)
                   | (_lib.NMR_F_PRECESSION if precession else 0) | (_lib.NMR_F_ADVECT_MASK if advect_mask else 0))
        p.scheme = int(scheme)
        p.phase_mode = _lib.NMR_PHASE_EXACT if phase_exact else _lib.NMR_PHASE_LINEAR
        for a in range(3):
            p.buffer[a] = int(buffer[a])
            p.bc[a] = int(bc[a])
            p.b0_dir[a] = float(b0_dir[a])
            p.grad[a] = float(grad[a])
        p.dx = float(dx)
        p.dy = float(dy if dy is not None else dx)
        p.dz = float(dz if dz is not None else dx)
        p.dt = float(dt)
        p.diffusivity = float(diffusivity)
        p.r1_bulk = float(r1_bulk)
        p.r2_bulk = float(r2_bulk)
        p.rho1 = float(rho1)
        p.rho2 = float(rho2)
        p.gamma = float(gamma)
        p.p_eq = float(p_eq)
        self.problem = p
        self.velocity = velocity
        ctx = ctypes.c_void_p()
        check(self.lib.nmr_hip_create(ctypes.byref(p), ctypes.byref(ctx)), "nmr_hip_create")
        self.ctx = ctx

    @classmethod
    def from_case(cls, case: Case, precession: bool | None = None, **kw) -> "HipSolver":
        has_v = case.vx is not None
        prec = case.grad_sched is not None if precession is None else precession
        s = cls(case.shape, case.ncomp, case.dx, case.dt, case.diffusivity, 1.0 / case.t2_bulk,
                r1_bulk=1.0 / case.t1_bulk, rho1=case.rho, rho2=case.rho, velocity=has_v, precession=prec, **kw)
        s.upload_mask(case.mask)
        if has_v:
            s.upload_velocity(case.vx, case.vy, case.vz)
        return s

    # -- lifetime -------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "ctx", None) is not None and self.ctx.value:
            check(self.lib.nmr_hip_destroy(self.ctx), "nmr_hip_destroy")
            self.ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- data -----------------------------------------------------------------
    def upload_mask(self, mask: np.ndarray, layout: int = _lib.NMR_LAYOUT_COMPACT) -> None:
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        check(self.lib.nmr_hip_upload_mask(self.ctx, m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), layout),
              "nmr_hip_upload_mask")

    def upload_velocity(self, vx, vy, vz, layout: int = _lib.NMR_LAYOUT_COMPACT) -> None:
        arrs = [None if v is None else np.ascontiguousarray(v, dtype=np.float64) for v in (vx, vy, vz)]
        check(self.lib.nmr_hip_upload_velocity(self.ctx, *[_dptr(a) for a in arrs], layout),
              "nmr_hip_upload_velocity")

    def upload_magnetization(self, mx, my=None, mz=None, layout: int = _lib.NMR_LAYOUT_COMPACT) -> None:
        arrs = [None if v is None else np.ascontiguousarray(v, dtype=np.float64) for v in (mx, my, mz)]
        check(self.lib.nmr_hip_upload_magnetization(self.ctx, *[_dptr(a) for a in arrs], layout),
              "nmr_hip_upload_magnetization")

    def upload_coefficients(self, D, r2, r1=None, layout: int = _lib.NMR_LAYOUT_COMPACT) -> None:
        """Per-voxel diffusivity and relaxation rates 1/T1, 1/T2 (surface terms included)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None for a in (D, r1, r2)]
        check(self.lib.nmr_hip_upload_coefficients(self.ctx, *[_dptr(a) for a in arrs], layout),
              "nmr_hip_upload_coefficients")

    def upload_field(self, b, layout: int = _lib.NMR_LAYOUT_COMPACT) -> None:
        """Per-voxel field (T) along the main-field axis in place of that axis's
        gradient field (the reference's noise field); None removes it."""
        arr = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        check(self.lib.nmr_hip_upload_field(self.ctx, _dptr(arr), layout), "nmr_hip_upload_field")

    def upload_phase_weights(self, w1, w2, layout: int = _lib.NMR_LAYOUT_COMPACT) -> None:
        """Two-phase output weights; signal rows then hold 2*ncomp sums."""
        a, b = (np.ascontiguousarray(x, dtype=np.float64) for x in (w1, w2))
        check(self.lib.nmr_hip_upload_phase_weights(self.ctx, _dptr(a), _dptr(b), layout),
              "nmr_hip_upload_phase_weights")
        self.nsig = 2 * self.ncomp

    def download_magnetization(self, layout: int = _lib.NMR_LAYOUT_COMPACT):
        nz, ny, nx = self.shape
        shp = (nz, ny, nx) if layout == _lib.NMR_LAYOUT_COMPACT else (nz + 2, ny + 2, nx + 2)
        outs = [np.empty(shp) if q < self.ncomp else None for q in range(3)]
        check(self.lib.nmr_hip_download_magnetization(self.ctx, *[_dptr(a) for a in outs], layout),
              "nmr_hip_download_magnetization")
        return outs[: self.ncomp]

    # -- stepping -------------------------------------------------------------
    def set_tile(self, variant: int = -1, lz: int = 0) -> None:
        check(self.lib.nmr_hip_set_tile(self.ctx, int(variant), int(lz)), "nmr_hip_set_tile")

    def set_layout(self, brick: bool = True) -> None:
        """Sweep the state in 128-byte bricks of 8 x 4 voxels (default) or in the row layout."""
        check(self.lib.nmr_hip_set_layout(self.ctx, int(bool(brick))), "nmr_hip_set_layout")

    def run(self, nsteps: int, grad_sched: np.ndarray | None = None, want_signal: bool = True):
        """Advance nsteps; returns the (nsteps+1, ncomp) signal (or None)."""
        g = None
        if grad_sched is not None:
            g = np.ascontiguousarray(grad_sched, dtype=np.float64)
            if g.shape != (nsteps, 3):
                raise ValueError(f"grad_sched must have shape ({nsteps}, 3)")
        sig = np.empty((nsteps + 1, getattr(self, "nsig", self.ncomp))) if want_signal else None
        check(self.lib.nmr_hip_run(self.ctx, int(nsteps), _addr(g), _addr(sig)), "nmr_hip_run")
        return sig

    def pulse(self, flag: int, param: float, axis) -> None:
        """Rotate every pore voxel (AddImpulse flags 1-4)."""
        check(self.lib.nmr_hip_pulse(self.ctx, int(flag), float(param), float(axis[0]), float(axis[1]),
                                     float(axis[2])), "nmr_hip_pulse")

    def signal(self) -> np.ndarray:
        out = np.empty(getattr(self, "nsig", self.ncomp))
        check(self.lib.nmr_hip_signal(self.ctx, _dptr(out)), "nmr_hip_signal")
        return out

    def last_run_ms(self):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        check(self.lib.nmr_hip_last_run_ms(self.ctx, ctypes.byref(ms), ctypes.byref(n)), "nmr_hip_last_run_ms")
        return ms.value, n.value

    def device_bytes(self) -> int:
        b = ctypes.c_int64()
        check(self.lib.nmr_hip_device_bytes(self.ctx, ctypes.byref(b)), "nmr_hip_device_bytes")
        return b.value


def variants():
    """[(index, lanes_x, rows_per_thread, tile_x, tile_y, steps_per_pass)] of
    the sweep kernels."""
    lib = _lib.load()
    out = []
    for v in range(lib.nmr_hip_num_variants()):
        f = [ctypes.c_int32() for _ in range(5)]
        check(lib.nmr_hip_variant_info(v, *[ctypes.byref(x) for x in f]), "nmr_hip_variant_info")
        out.append((v,) + tuple(x.value for x in f))
    return out


class BatchSolver:
    """Many independent problems of uneven size advanced together: one
    persistent launch per step covers every instance, chunks of steps are
    replayed from a captured HIP graph.  The instances share the physics
    (ncomp, D, bulk relaxation, dx, dt, ...) and each has its own grid,
    surface relaxivity and pulsed-gradient spin echo along z (+G for the
    first round(delta/dt) steps, -G from step round(big_delta/dt) for as
    many)."""

    def __init__(self, instances, ncomp: int, dx: float, dt: float, diffusivity: float, r2_bulk: float,
                 r1_bulk: float = 0.0, gamma: float = G_REF, p_eq: float = 0.0, diffusion: bool = True,
                 velocity: bool = False, precession: bool = False, phase_exact: bool = False, buffer=(0, 0, 0)):
        """instances: iterable of dicts with shape (nz, ny, nx), rho1, rho2,
        grad_amp, delta, big_delta (missing keys default to 0)."""
        self.lib = _lib.load()
        inst = list(instances)
        self.ninst = len(inst)
        self.ncomp = int(ncomp)
        self.shapes = [tuple(int(x) for x in d["shape"]) for d in inst]
        arr = (_lib.NmrBatchInstanceC * max(1, self.ninst))()
        for i, d in enumerate(inst):
            nz, ny, nx = self.shapes[i]
            arr[i].nx, arr[i].ny, arr[i].nz = nx, ny, nz
            arr[i].rho1 = float(d.get("rho1", 0.0))
            arr[i].rho2 = float(d.get("rho2", d.get("rho1", 0.0)))
            arr[i].grad_amp = float(d.get("grad_amp", 0.0))
            arr[i].delta = float(d.get("delta", 0.0))
            arr[i].big_delta = float(d.get("big_delta", 0.0))
        p = NmrProblemC()
        p.nx, p.ny, p.nz = 1, 1, 1
        p.ncomp = self.ncomp
        p.flags = ((_lib.NMR_F_DIFFUSION if diffusion else 0) | (_lib.NMR_F_VELOCITY if velocity else 0)
                   | (_lib.NMR_F_PRECESSION if precession else 0))
        p.scheme = _lib.NMR_SCHEME_UPWIND
        p.phase_mode = _lib.NMR_PHASE_EXACT if phase_exact else _lib.NMR_PHASE_LINEAR
        for a in range(3):
            p.buffer[a] = int(buffer[a])
        p.b0_dir[2] = 1.0
        p.dx = p.dy = p.dz = float(dx)
        p.dt = float(dt)
        p.diffusivity = float(diffusivity)
        p.r1_bulk = float(r1_bulk)
        p.r2_bulk = float(r2_bulk)
        p.gamma = float(gamma)
        p.p_eq = float(p_eq)
        self.problem = p
        self.velocity = velocity
        h = ctypes.c_void_p()
        check(self.lib.nmr_hip_batch_create(ctypes.byref(p), self.ninst, arr, ctypes.byref(h)),
              "nmr_hip_batch_create")
        self.h = h

    def close(self) -> None:
        if getattr(self, "h", None) is not None and self.h.value:
            check(self.lib.nmr_hip_batch_destroy(self.h), "nmr_hip_batch_destroy")
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, i: int, mask, mag, vel=None) -> None:
        """mask (nz, ny, nx) 0/1; mag: list of ncomp arrays; vel: (vx, vy, vz)
        face arrays or None."""
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        v = [None] * 3 if vel is None else [np.ascontiguousarray(a, dtype=np.float64) for a in vel]
        g = [np.ascontiguousarray(a, dtype=np.float64) for a in mag] + [None] * (3 - len(mag))
        check(self.lib.nmr_hip_batch_upload(self.h, int(i), m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                            *[_dptr(a) for a in v], *[_dptr(a) for a in g]),
              "nmr_hip_batch_upload")

    def download(self, i: int):
        outs = [np.empty(self.shapes[i]) if q < self.ncomp else None for q in range(3)]
        check(self.lib.nmr_hip_batch_download(self.h, int(i), *[_dptr(a) for a in outs]), "nmr_hip_batch_download")
        return outs[: self.ncomp]

    def set_layout(self, brick: bool = True) -> None:
        """Keep the instances' state in 128-byte bricks of 8 x 4 voxels (default) or in rows."""
        check(self.lib.nmr_hip_batch_set_layout(self.h, int(bool(brick))), "nmr_hip_batch_set_layout")

    def set_tile(self, variant: int = -1, lz: int = 0, graph_steps: int = 64, persistent: bool = True) -> None:
        check(self.lib.nmr_hip_batch_set_tile(self.h, int(variant), int(lz), int(graph_steps), int(persistent)),
              "nmr_hip_batch_set_tile")

    def autotune(self, trial_steps: int = 5):
        """Returns (variant, lz, persistent) of the fastest launch."""
        v = ctypes.c_int32()
        lz = ctypes.c_int32()
        per = ctypes.c_int32()
        check(self.lib.nmr_hip_batch_autotune(self.h, int(trial_steps), ctypes.byref(v), ctypes.byref(lz),
                                              ctypes.byref(per)), "nmr_hip_batch_autotune")
        return v.value, lz.value, bool(per.value)

    def run(self, nsteps: int, want_signal: bool = True):
        """Advance every instance nsteps; returns (ninst, nsteps+1, ncomp)."""
        sig = np.empty((self.ninst, nsteps + 1, self.ncomp)) if want_signal else None
        check(self.lib.nmr_hip_batch_run(self.h, int(nsteps), _dptr(sig)), "nmr_hip_batch_run")
        return sig

    def last_run_ms(self):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        check(self.lib.nmr_hip_batch_last_run_ms(self.h, ctypes.byref(ms), ctypes.byref(n)),
              "nmr_hip_batch_last_run_ms")
        return ms.value, n.value


def _fptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Corrector:
    """The learned corrector (nmr_hip_corrector_*): a 3-layer BF16 3-D
    convolution net that corrects one coarse step of factor*dt toward
    factor plain steps; trained on samples of a BatchSolver, applied by
    run_learned on a HipSolver with a residual check and plain fallback."""

    NPARAMS = _lib.NMR_CORRECTOR_PARAMS

    def __init__(self, seed: int = 0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        check(self.lib.nmr_hip_corrector_create(int(seed), ctypes.byref(h)), "nmr_hip_corrector_create")
        self.h = h

    def close(self) -> None:
        if getattr(self, "h", None) is not None and self.h.value:
            check(self.lib.nmr_hip_corrector_destroy(self.h), "nmr_hip_corrector_destroy")
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        out = np.empty(self.NPARAMS, dtype=np.float32)
        check(self.lib.nmr_hip_corrector_get_params(self.h, _fptr(out)), "nmr_hip_corrector_get_params")
        return out

    @params.setter
    def params(self, p) -> None:
        a = np.ascontiguousarray(p, dtype=np.float32)
        if a.shape != (self.NPARAMS,):
            raise ValueError(f"expected {self.NPARAMS} parameters")
        check(self.lib.nmr_hip_corrector_set_params(self.h, _fptr(a)), "nmr_hip_corrector_set_params")

    def set_scales(self, vscale: float, out_scale: float) -> None:
        check(self.lib.nmr_hip_corrector_set_scales(self.h, float(vscale), float(out_scale)),
              "nmr_hip_corrector_set_scales")

    def set_conv(self, kind: int = 0, zc: int = 0) -> None:
        """Convolution kernel (0 z-marching columns, 1 tiles, 2 row gathers,
        3 z-marching with the learned solve's last layer on pore rows only);
        zc: output planes per z-marching column (0: 16).  Same bits for all."""
        check(self.lib.nmr_hip_corrector_set_conv(self.h, int(kind), int(zc)), "nmr_hip_corrector_set_conv")

    def scales(self):
        v, o = ctypes.c_double(), ctypes.c_double()
        check(self.lib.nmr_hip_corrector_get_scales(self.h, ctypes.byref(v), ctypes.byref(o)),
              "nmr_hip_corrector_get_scales")
        return v.value, o.value

    def forward(self, features: np.ndarray) -> np.ndarray:
        """features (nz, ny, nx, 6) -> output (nz, ny, nx, 2)."""
        f = np.ascontiguousarray(features, dtype=np.float32)
        nz, ny, nx, c = f.shape
        assert c == 6
        out = np.empty((nz, ny, nx, 2), dtype=np.float32)
        check(self.lib.nmr_hip_corrector_forward(self.h, nx, ny, nz, _fptr(f), _fptr(out)),
              "nmr_hip_corrector_forward")
        return out

    def backward(self, features: np.ndarray, grad_out: np.ndarray) -> np.ndarray:
        """Gradient of sum(grad_out * forward(features)) w.r.t. the parameters."""
        f = np.ascontiguousarray(features, dtype=np.float32)
        g = np.ascontiguousarray(grad_out, dtype=np.float32)
        nz, ny, nx, _ = f.shape
        out = np.empty(self.NPARAMS, dtype=np.float32)
        check(self.lib.nmr_hip_corrector_backward(self.h, nx, ny, nz, _fptr(f), _fptr(g), _fptr(out)),
              "nmr_hip_corrector_backward")
        return out

    def train(self, batch: "BatchSolver", factor: int = 8, rounds: int = 10, stride: int = 0, lr: float = 1e-3):
        """Returns the (rounds, ninst) losses."""
        o = _lib.NmrTrainOptsC()
        o.factor, o.rounds, o.stride, o.lr = int(factor), int(rounds), int(stride), float(lr)
        losses = np.empty((rounds, batch.ninst))
        check(self.lib.nmr_hip_corrector_train(self.h, batch.h, ctypes.byref(o), _dptr(losses)),
              "nmr_hip_corrector_train")
        return losses

    def run_learned(self, solver: "HipSolver", ncoarse: int, factor: int = 8, grad_sched=None, tol: float = 1e-3,
                    reasons: bool = False):
        """Returns (signal (ncoarse+1, 2), fallbacks) -- with reasons=True,
        fallbacks is (total, non-finite, energy, size)."""
        g = None if grad_sched is None else np.ascontiguousarray(grad_sched, dtype=np.float64)
        if g is not None and g.shape != (ncoarse * factor, 3):
            raise ValueError("grad_sched must have shape (ncoarse*factor, 3)")
        sig = np.empty((ncoarse + 1, 2))
        nf = (ctypes.c_int64 * 4)()
        check(self.lib.nmr_hip_run_learned(solver.ctx, self.h, int(ncoarse), int(factor), _dptr(g), float(tol),
                                           _dptr(sig), nf), "nmr_hip_run_learned")
        return sig, (tuple(nf) if reasons else nf[0])

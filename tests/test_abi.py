"""The C ABI library loads, exports every symbol include/nmr_hip.h declares,
and rejects bad arguments with status codes (no GPU needed: validation runs
before any device call)."""
import ctypes
import os
import re
import subprocess

import pytest

from initializator_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_lib.LIB_PATH):
        subprocess.run(["make", "-C", os.path.join(ROOT, "initializator_amd", "csrc")], check=True,
                       stdout=subprocess.DEVNULL)
    return _lib.load()


def _declared():
    text = open(os.path.join(ROOT, "include", "nmr_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nmr_hip_\w+)\s*\(", text)))


def test_exports_every_declared_symbol(lib):
    names = _declared()
    assert len(names) >= 16
    assert set(names) == set(_lib.EXPORTS)
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], check=True, capture_output=True,
                         text=True).stdout
    exported = set(re.findall(r"\bT\s+(\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_struct_layout_matches_header(lib, tmp_path):
    """The ctypes mirror has the offsets the C compiler gives nmr_problem."""
    fields = [f[0] for f in _lib.NmrProblemC._fields_]
    src = tmp_path / "layout.c"
    src.write_text(
        "#include <stdio.h>\n#include <stddef.h>\n#include \"nmr_hip.h\"\nint main(void){\n"
        + "".join(f'printf("%zu\\n", offsetof(nmr_problem, {f}));\n' for f in fields)
        + 'printf("%zu\\n", sizeof(nmr_problem));return 0;}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [getattr(_lib.NmrProblemC, f).offset for f in fields] + [ctypes.sizeof(_lib.NmrProblemC)]
    assert got == want


def test_batch_instance_layout_matches_header(lib, tmp_path):
    fields = [f[0] for f in _lib.NmrBatchInstanceC._fields_]
    src = tmp_path / "layout_b.c"
    src.write_text(
        "#include <stdio.h>\n#include <stddef.h>\n#include \"nmr_hip.h\"\nint main(void){\n"
        + "".join(f'printf("%zu\\n", offsetof(nmr_batch_instance, {f}));\n' for f in fields)
        + 'printf("%zu\\n", sizeof(nmr_batch_instance));return 0;}\n')
    exe = tmp_path / "layout_b"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [getattr(_lib.NmrBatchInstanceC, f).offset for f in fields] + [ctypes.sizeof(_lib.NmrBatchInstanceC)]
    assert got == want


def _problem(**kw):
    p = _lib.NmrProblemC()
    p.nx = p.ny = p.nz = 8
    p.ncomp = 2
    p.flags = _lib.NMR_F_DIFFUSION
    p.dx = p.dy = p.dz = 1e-6
    p.dt = 1e-5
    p.b0_dir[2] = 1.0
    for k, v in kw.items():
        if k == "b0":
            for i in range(3):
                p.b0_dir[i] = v[i]
        elif k == "grad":
            for i in range(3):
                p.grad[i] = v[i]
        elif k == "bc":
            for i in range(3):
                p.bc[i] = v[i]
        else:
            setattr(p, k, v)
    return p


@pytest.mark.parametrize("kw,code", [
    (dict(nx=0), _lib.NMR_ERR_INVALID),
    (dict(nz=-3), _lib.NMR_ERR_INVALID),
    (dict(ncomp=4), _lib.NMR_ERR_INVALID),
    (dict(scheme=7), _lib.NMR_ERR_UNSUPPORTED),
    (dict(phase_mode=9), _lib.NMR_ERR_INVALID),
    (dict(flags=64), _lib.NMR_ERR_INVALID),
    (dict(dt=0.0), _lib.NMR_ERR_INVALID),
    (dict(b0=(1.0, 0.0, 0.0)), _lib.NMR_ERR_UNSUPPORTED),
    (dict(b0=(0.0, 0.0, 0.0)), _lib.NMR_ERR_INVALID),
    (dict(ncomp=1, flags=_lib.NMR_F_PRECESSION), _lib.NMR_ERR_UNSUPPORTED),
    (dict(grad=(0.1, 0.0, 0.0)), _lib.NMR_ERR_UNSUPPORTED),
    (dict(phase_mode=_lib.NMR_PHASE_EXACT, ncomp=3), _lib.NMR_ERR_UNSUPPORTED),
    (dict(nx=4096, ny=4096, nz=4096), _lib.NMR_ERR_INVALID),
    # below 2^31 voxels, but a face array (Vx; Vy alone) passes the 4 GiB
    # that the sweep's 32-bit byte offsets address
    (dict(nx=1021, ny=1024, nz=1024), _lib.NMR_ERR_UNSUPPORTED),
    (dict(nx=1023, ny=1, nz=600000), _lib.NMR_ERR_UNSUPPORTED),
    (dict(bc=(0, 3, 0)), _lib.NMR_ERR_INVALID),
    (dict(scheme=_lib.NMR_SCHEME_MACCORMACK, phase_mode=_lib.NMR_PHASE_EXACT), _lib.NMR_ERR_UNSUPPORTED),
    (dict(bc=(1, 0, 0), phase_mode=_lib.NMR_PHASE_EXACT), _lib.NMR_ERR_UNSUPPORTED),
])
def test_create_rejects_bad_problems(lib, kw, code):
    ctx = ctypes.c_void_p(123)
    rc = lib.nmr_hip_create(ctypes.byref(_problem(**kw)), ctypes.byref(ctx))
    assert rc == code, lib.nmr_hip_last_error().decode()
    assert ctx.value is None
    assert lib.nmr_hip_last_error().decode()


def test_null_arguments(lib):
    assert lib.nmr_hip_create(None, None) == _lib.NMR_ERR_NULL
    assert lib.nmr_hip_destroy(None) == _lib.NMR_OK
    assert lib.nmr_hip_run(None, 1, None, None) == _lib.NMR_ERR_NULL
    assert lib.nmr_hip_signal(None, None) == _lib.NMR_ERR_NULL
    assert lib.nmr_hip_upload_mask(None, None, 0) == _lib.NMR_ERR_NULL


def test_strerror_and_variants(lib):
    for code in range(-6, 1):
        assert lib.nmr_hip_strerror(code)
    assert lib.nmr_hip_abi_version() == 2
    n = lib.nmr_hip_num_variants()
    assert n >= 4
    a = ctypes.c_int32()
    assert lib.nmr_hip_variant_info(n, None, None, None, None, None) == _lib.NMR_ERR_INVALID
    assert lib.nmr_hip_variant_info(0, ctypes.byref(a), None, None, None, None) == _lib.NMR_OK
    assert a.value in (8, 16, 32, 64)
    steps = []
    for v in range(n):
        st = ctypes.c_int32()
        assert lib.nmr_hip_variant_info(v, None, None, None, None, ctypes.byref(st)) == _lib.NMR_OK
        steps.append(st.value)
    assert set(steps) == {1, 2}


def _inst(n, **kw):
    arr = (_lib.NmrBatchInstanceC * max(n, 1))()
    for i in range(n):
        arr[i].nx, arr[i].ny, arr[i].nz = 8, 9, 10
        for k, v in kw.items():
            setattr(arr[i], k, v)
    return arr


@pytest.mark.parametrize("pkw,ikw,n,code", [
    (dict(), dict(), 0, _lib.NMR_ERR_INVALID),
    (dict(), dict(), 70000, _lib.NMR_ERR_INVALID),
    (dict(), dict(nx=0), 3, _lib.NMR_ERR_INVALID),
    (dict(), dict(nx=1023, ny=1, nz=600000), 2, _lib.NMR_ERR_UNSUPPORTED),
    (dict(), dict(delta=-1.0), 2, _lib.NMR_ERR_INVALID),
    (dict(scheme=3), dict(), 2, _lib.NMR_ERR_UNSUPPORTED),
    (dict(ncomp=1, flags=_lib.NMR_F_PRECESSION), dict(), 2, _lib.NMR_ERR_UNSUPPORTED),
    (dict(dt=-1.0), dict(), 2, _lib.NMR_ERR_INVALID),
    (dict(scheme=_lib.NMR_SCHEME_MACCORMACK), dict(), 2, _lib.NMR_ERR_UNSUPPORTED),
    (dict(bc=(1, 0, 0)), dict(), 2, _lib.NMR_ERR_UNSUPPORTED),
])
def test_batch_create_rejects_bad_arguments(lib, pkw, ikw, n, code):
    h = ctypes.c_void_p(77)
    rc = lib.nmr_hip_batch_create(ctypes.byref(_problem(**pkw)), n, _inst(min(n, 4), **ikw), ctypes.byref(h))
    assert rc == code, lib.nmr_hip_last_error().decode()
    assert h.value is None


def test_batch_null_arguments(lib):
    assert lib.nmr_hip_batch_create(None, 1, None, None) == _lib.NMR_ERR_NULL
    assert lib.nmr_hip_batch_destroy(None) == _lib.NMR_OK
    assert lib.nmr_hip_batch_run(None, 1, None) == _lib.NMR_ERR_NULL
    assert lib.nmr_hip_batch_upload(None, 0, None, None, None, None, None, None, None) == _lib.NMR_ERR_NULL
    assert lib.nmr_hip_batch_autotune(None, 1, None, None, None) == _lib.NMR_ERR_NULL
    assert lib.nmr_hip_batch_set_tile(None, 0, 0, 0, 1) == _lib.NMR_ERR_NULL


def test_measurement_entry_points_reject_bad_arguments(lib):
    out = ctypes.c_double()
    assert lib.nmr_hip_mfma_rate(16, 1, None) == _lib.NMR_ERR_NULL
    assert lib.nmr_hip_mfma_rate(0, 1, ctypes.byref(out)) == _lib.NMR_ERR_INVALID
    assert lib.nmr_hip_mfma_rate(1 << 25, 1, ctypes.byref(out)) == _lib.NMR_ERR_INVALID
    assert lib.nmr_hip_mfma_rate(16, 0, ctypes.byref(out)) == _lib.NMR_ERR_INVALID
    assert lib.nmr_hip_copy_bandwidth(1 << 20, 1, None) == _lib.NMR_ERR_NULL
    assert lib.nmr_hip_copy_bandwidth(1024, 1, ctypes.byref(out)) == _lib.NMR_ERR_INVALID

"""Batched solves (nmr_hip_batch_*): many uneven instances advanced by one
persistent launch per step, chunks replayed from a captured HIP graph.

* every instance's field equals, bit for bit, the single-instance solve of
  the same problem (same FP32 update), and its signal agrees to FP64
  summation-order rounding;
* graph replays equal direct launches bit for bit, across several replays
  with host synchronisation between them and across consecutive runs;
* autotuning leaves the state unchanged and the results tile-invariant;
* a direct check of two instances against the FP64 oracle at the parity
  bars of the single-instance tests (signal rel. L2 1e-5, voxel 1e-4).
"""
import numpy as np
import pytest

import oracle
from initializator_amd import generators
from initializator_amd.solver import BatchSolver, HipSolver

pytestmark = pytest.mark.gpu

DX, D = 1e-6, 2.3e-9
DT = generators.stable_dt(DX, D)
T2, T1 = 0.5, 2.5e10


def _lround(x):
    return int(np.floor(x + 0.5))


def _instances():
    """Uneven shapes (x not a multiple of 4, tiny, flat), per-instance
    relaxivity and PGSE timings that fall inside the run."""
    specs = [
        dict(shape=(20, 24, 29), rho1=5e-6, grad_amp=0.3, delta=10 * DT, big_delta=40 * DT, phi=0.35, seed=1),
        dict(shape=(33, 17, 40), rho1=30e-6, grad_amp=-0.8, delta=7 * DT, big_delta=20 * DT, phi=0.25, seed=2),
        dict(shape=(6, 41, 70), rho1=1e-6, grad_amp=0.05, delta=15 * DT, big_delta=60 * DT, phi=0.4, seed=3),
        dict(shape=(9, 9, 9), rho1=50e-6, grad_amp=1.2, delta=3 * DT, big_delta=5 * DT, phi=0.5, seed=4),
        dict(shape=(16, 16, 16), rho1=10e-6, grad_amp=0.0, delta=0.0, big_delta=0.0, phi=0.3, seed=5),
    ]
    for s in specs:
        s["rho2"] = s["rho1"]
        s["mask"] = generators.porosity(s["shape"], s["phi"], seed=s["seed"], sigma=1.5)
        s["vel"] = generators.velocity(s["mask"], seed=s["seed"], sigma=1.5, vmax=1e-3, d=(DX, DX, DX))
        rng = np.random.default_rng(s["seed"])
        s["m0"] = [s["mask"] * rng.random(s["shape"]), s["mask"] * (rng.random(s["shape"]) - 0.5)]
    return specs


def _schedule(s, nsteps, start=0):
    nd, nD = _lround(s["delta"] / DT), _lround(s["big_delta"] / DT)
    g = np.zeros((nsteps, 3))
    for k in range(nsteps):
        step = start + k
        if nD <= step < nD + nd:
            g[k, 2] = -s["grad_amp"]
        elif step < nd:
            g[k, 2] = s["grad_amp"]
    return g


def _batch(specs, graph_steps=16, variant=-1, lz=0, persistent=True):
    b = BatchSolver(specs, 2, DX, DT, D, 1.0 / T2, r1_bulk=1.0 / T1, velocity=True, precession=True)
    b.set_tile(variant, lz, graph_steps, persistent)
    for i, s in enumerate(specs):
        b.upload(i, s["mask"], s["m0"], s["vel"])
    return b


def _single(s, nsteps):
    h = HipSolver(s["shape"], 2, DX, DT, D, 1.0 / T2, r1_bulk=1.0 / T1, rho1=s["rho1"], rho2=s["rho2"],
                  velocity=True, precession=True)
    h.upload_mask(s["mask"])
    h.upload_velocity(*s["vel"])
    h.upload_magnetization(*s["m0"])
    sig = h.run(nsteps, _schedule(s, nsteps))
    f = h.download_magnetization()
    h.close()
    return sig, f


def test_batch_equals_single_instance_solves():
    specs = _instances()
    n = 150  # 9 graph replays of 16 steps + 6 direct steps
    b = _batch(specs, graph_steps=16)
    sig = b.run(n)
    ms, vox = b.last_run_ms()
    assert vox == n * sum(int(np.prod(s["shape"])) for s in specs) and ms > 0
    for i, s in enumerate(specs):
        ref_sig, ref_f = _single(s, n)
        f = b.download(i)
        for q in range(2):
            assert np.array_equal(f[q], ref_f[q]), (i, q, float(np.abs(f[q] - ref_f[q]).max()))
        np.testing.assert_allclose(sig[i], ref_sig, rtol=1e-11, atol=1e-12 * np.abs(ref_sig).max())
    print(f"[batch] 5 instances x {n} steps bit-identical to single solves", flush=True)
    b.close()


def test_graph_replays_equal_direct_launches():
    specs = _instances()
    direct = _batch(specs, graph_steps=0)
    graph = _batch(specs, graph_steps=8)
    sig_d, sig_g = [], []
    for n in (40, 37, 16):  # consecutive runs: the step count carries over
        sig_d.append(direct.run(n))
        sig_g.append(graph.run(n))
        for i in range(len(specs)):
            fd, fg = direct.download(i), graph.download(i)
            for q in range(2):
                assert np.array_equal(fd[q], fg[q]), (n, i, q)
    for a, c in zip(sig_d, sig_g):
        assert np.array_equal(a, c)
    # consecutive runs continue the step count (row 0 of a run = last row of the previous one)
    assert np.array_equal(sig_d[1][:, 0], sig_d[0][:, -1]) and np.array_equal(sig_d[2][:, 0], sig_d[1][:, -1])
    direct.close()
    graph.close()


def test_split_runs_equal_one_run():
    specs = _instances()
    one = _batch(specs, graph_steps=16)
    split = _batch(specs, graph_steps=16)
    s1 = one.run(90)
    a = split.run(33)
    c = split.run(57)
    for i in range(len(specs)):
        for q in range(2):
            assert np.array_equal(one.download(i)[q], split.download(i)[q])
    np.testing.assert_array_equal(s1[:, :34], a)
    np.testing.assert_array_equal(s1[:, 33:], c)
    one.close()
    split.close()


def test_autotune_keeps_state_and_tile_invariance():
    specs = _instances()
    tuned = _batch(specs)
    v, lz, per = tuned.autotune(3)
    assert (v == -3 or 0 <= v <= 6) and lz in (4, 8, 16, 32, 64)
    for i, s in enumerate(specs):
        f = tuned.download(i)
        for q in range(2):
            assert np.array_equal(f[q], s["m0"][q].astype(np.float32).astype(np.float64))
    plain = _batch(specs, variant=0, lz=5, persistent=not per)
    st, sp = tuned.run(60), plain.run(60)
    for i in range(len(specs)):
        for q in range(2):
            assert np.array_equal(tuned.download(i)[q], plain.download(i)[q])
    np.testing.assert_allclose(st, sp, rtol=1e-11, atol=1e-14)
    print(f"[batch] autotuned variant {v}, lz {lz}, persistent {per}", flush=True)
    tuned.close()
    plain.close()


@pytest.mark.parametrize("persistent", [True, False])
def test_per_instance_tile_shapes_equal_single_solves(persistent):
    """Variant -3: every instance takes its own tile shape (one launch per
    shape, several shapes in one captured graph); the fields equal the
    single-instance solves bit for bit."""
    specs = _instances()
    n = 40
    b = _batch(specs, graph_steps=8, variant=-3, lz=6, persistent=persistent)
    sig = b.run(n)
    for i, s in enumerate(specs):
        sig1, f1 = _single(s, n)
        f = b.download(i)
        for q in range(2):
            assert np.array_equal(f[q], f1[q]), (i, q)
        np.testing.assert_allclose(sig[i], sig1, rtol=1e-11, atol=1e-14)
    b.close()


def test_batch_against_fp64_oracle():
    specs = _instances()[:2]
    n = 120
    b = _batch(specs)
    sig = b.run(n)
    for i, s in enumerate(specs):
        case = generators.Case(s["shape"], s["mask"], *s["vel"], DX, DT, D, T2, T1, s["rho1"], n,
                               _schedule(s, n), 2)
        m0 = [s["m0"][0], s["m0"][1], np.zeros(s["shape"])]
        sig_ref, f_ref = oracle.run_case(case, m0=m0)
        f = b.download(i)
        e_sig = float(np.linalg.norm(sig[i] - sig_ref[:, :2]) / np.linalg.norm(sig_ref[:, :2]))
        e_vox = max(float(np.abs(f[q] - f_ref[q]).max()) for q in range(2))
        print(f"[batch] instance {i} {s['shape']}: signal relL2 {e_sig:.2e}, max voxel {e_vox:.2e}", flush=True)
        assert e_sig <= 1e-5 and e_vox <= 1e-4
    b.close()


def test_batch_state_errors():
    specs = _instances()[:2]
    b = BatchSolver(specs, 2, DX, DT, D, 1.0 / T2, velocity=True, precession=True)
    with pytest.raises(Exception, match="upload"):
        b.run(2)
    with pytest.raises(Exception):
        b.set_tile(0, 0, 3)  # odd graph chunk
    with pytest.raises(Exception):
        b.set_tile(0, 0, 2, 5)  # grid kind
    with pytest.raises(Exception):
        b.set_tile(8, 0, 0)  # a two-step variant
    b.close()


def test_batch_bricks_equal_rows():
    """The batch on bricks (the default: state and velocity of every
    instance in 8 x 4 bricks from run to run) against the row layout: the
    same fields bit for bit and the same signal rows, across runs of odd
    length (graph replays and direct steps), a download in the middle and a
    re-upload of one instance between runs."""
    specs = _instances()
    res = {}
    for brick in (True, False):
        b = _batch(specs, graph_steps=16)
        b.set_layout(brick)
        s1 = b.run(37)
        f1 = [b.download(i) for i in range(len(specs))]
        b.upload(2, specs[2]["mask"], specs[2]["m0"], specs[2]["vel"])
        s2 = b.run(20)
        f2 = [b.download(i) for i in range(len(specs))]
        res[brick] = (s1, f1, s2, f2)
        b.close()
    a, r = res[True], res[False]
    for fa, fr in ((a[1], r[1]), (a[3], r[3])):
        for i in range(len(specs)):
            for q in range(2):
                assert np.array_equal(fa[i][q], fr[i][q]), (i, q)
    np.testing.assert_allclose(a[0], r[0], rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(a[2], r[2], rtol=1e-12, atol=1e-300)
    # row 0 of the second run: instance 2 restarted from its upload, the
    # others continue from the first run's last row
    for i in range(len(specs)):
        if i != 2:
            assert np.array_equal(a[2][i][0], a[0][i][-1])
    print("[batch] bricks == rows across runs, downloads and a re-upload", flush=True)

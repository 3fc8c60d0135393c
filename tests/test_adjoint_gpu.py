"""The GPU adjoint against the oracle's.

- the forward run with a design box (per-cell E coefficients) and a DFT
  plane is bit-identical to the oracle in FP64;
- checkpoint / restore reproduces the same bits;
- the GPU transpose satisfies <L x, y> = <x, L^T y> and matches the oracle's
  transpose (a different implementation: operation-by-operation reverse on
  the reference's per-face arrays) to 1e-11 of the result's scale;
- dJ/d eps_r matches the oracle's to 1e-10 (FP64) for any checkpoint
  interval, and central differences on the GPU forward run;
- the FP32 gradient stays within 1e-4 of the FP64 one;
- the torch wrapper returns the engine's gradient, and a DesignProblem that
  keeps one engine across calls (new permittivity, rewind to step 0) gives
  the bits of a fresh engine for every new permittivity;
- all of the above with Drude cells, whose currents are adjoint state too.
"""
import numpy as np
import pytest

from adjoint_scenes import box_scene, design_scene, drude_box_scene, drude_design_scene
from oracle.oracle import OracleSim
from xfdtd_amd import scene as S

pytestmark = pytest.mark.gpu


def _sim(scene, steps, precision=64):
    from xfdtd_amd.sim import Simulation
    return Simulation(scene, steps, precision=precision)


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_design_forward_and_dft_bitwise(oracle_threads):
    sc, steps = design_scene(steps=120)
    g, o = _sim(sc, steps), OracleSim(sc, steps)
    g.step(steps)
    o.step(steps)
    for c in range(6):
        assert np.array_equal(g.get_field(c), o.get_field(c)), c
    assert np.array_equal(g.get_dft(), o.get_dft())


def test_checkpoint_restore_bitwise():
    sc, steps = design_scene(steps=80)
    g = _sim(sc, steps)
    g.step(20)
    g.checkpoint(3)
    g.step(30)
    a = [g.get_field(c) for c in range(6)] + [g.get_dft()]
    g.restore(3)
    assert g.current_step == 20
    g.step(30)
    b = [g.get_field(c) for c in range(6)] + [g.get_dft()]
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_gpu_transpose_identity_and_oracle(oracle_threads):
    steps = 25
    sc = box_scene()
    g = _sim(sc, steps)
    rng = np.random.default_rng(7)
    x0 = [rng.standard_normal(g.field_shape(c)) for c in range(6)]
    y = [rng.standard_normal(g.field_shape(c)) for c in range(6)]
    for c in range(6):
        g.set_field(c, x0[c])
    g.step(steps)
    lhs = sum(float(np.sum(g.get_field(c) * y[c])) for c in range(6))
    ly = _sim(sc, steps).adjoint_run(steps, y)
    rhs = sum(float(np.sum(x0[c] * ly[c])) for c in range(6))
    scale = sum(float(np.sum(np.abs(x0[c] * ly[c]))) for c in range(6))
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs)
    lo = OracleSim(sc, steps).adjoint_run(steps, y)
    for c in range(6):
        assert _rel(ly[c], lo[c]) <= 1e-11, c


@pytest.mark.parametrize("every", [7, 1000])
def test_gradient_matches_oracle(every, oracle_threads):
    sc, steps = design_scene()
    Jo, go = OracleSim(sc, steps).grad()
    g = _sim(sc, steps)
    Jg, gg = g.grad(checkpoint_every=every)
    assert Jg == pytest.approx(Jo, rel=1e-14)
    assert _rel(gg, go) <= 1e-10
    # the engine is left at the end of the forward run
    assert g.current_step == steps
    o = OracleSim(sc, steps)
    o.step(steps)
    for c in range(6):
        assert np.array_equal(g.get_field(c), o.get_field(c)), c


def test_gradient_checkpoint_interval_bitwise():
    sc, steps = design_scene()
    _, a = _sim(sc, steps).grad(checkpoint_every=13)
    _, b = _sim(sc, steps).grad(checkpoint_every=steps)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("tiling", [None, (64, 4, 8, 0), (64, 4, 8, 16), (128, 2, 8, 4), (256, 0, 4, 8)])
@pytest.mark.parametrize("scene", ["design", "drude", "tfsf"])
def test_replay_windows_bitwise(scene, tiling):
    """Replayed segments update only the tiles within reach of the design
    box by the segment's end; the gradient equals the no-recompute run's
    (no replays) bit for bit under every tile shape and order, with Drude
    cells outside the box and with TFSF faces."""
    if scene == "tfsf":
        from xfdtd_amd.workloads import batched_adjoint_set
        sc, steps = batched_adjoint_set(count=1, seed=0xC5, n=34, pml=5, box=(12, 18), drude=(8, 12),
                                        tfsf_gap=7)[0], 60
    else:
        sc, steps = (design_scene if scene == "design" else drude_design_scene)()
    out = []
    for every in (13, steps):
        g = _sim(sc, steps)
        if tiling is not None:
            g.set_tiling(*tiling)
        out.append(g.grad(checkpoint_every=every))
        g.close()
    assert out[0][0] == out[1][0]
    assert np.abs(out[1][1]).max() > 0
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("axis,comp", [(0, S.EZ), (1, S.EX), (2, S.EY)])
def test_reverse_windows_bitwise(axis, comp, oracle_threads):
    """The reverse steps launch only the planes the adjoint can have reached
    from the objective plane; the gradient equals the batched kernels' (which
    launch every cell) bit for bit, for planes across x, y and z, and
    matches the oracle's."""
    from xfdtd_amd.sim import Batch
    sc, steps = design_scene()
    n, pml = 26, 5
    sc.dft_plane = S.DftPlane(comp, axis, 18 if axis == 0 else 17, pml + 2, n - pml - 3, pml + 3, n - pml - 2,
                              sc.dft_plane.freq)
    J1, g1 = _sim(sc, steps).grad(checkpoint_every=13)
    b = Batch([_sim(sc, steps)])
    J2, g2 = b.grad(checkpoint_every=13)[0]
    b.close()
    assert J1 == J2 and J1 > 0
    assert np.array_equal(g1, g2)
    Jo, go = OracleSim(sc, steps).grad()
    assert J1 == pytest.approx(Jo, rel=1e-14)
    assert _rel(g1, go) <= 1e-10


@pytest.mark.parametrize("precision", [64, 32])
def test_gradient_schedules_bitwise(precision):
    """The adjoint sweeps overlapped with the segment replays on a second
    stream give the serial schedule's gradient bit for bit, also with the
    segments uneven (13 does not divide the run)."""
    sc, steps = design_scene()
    out = []
    for mode in (0, 1):
        g = _sim(sc, steps, precision)
        g.set_adjoint_overlap(mode)
        out.append(g.grad(checkpoint_every=13))
        g.close()
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])


def test_gradient_central_differences_on_gpu():
    sc, steps = design_scene()
    J, g = _sim(sc, steps).grad(checkpoint_every=40)
    eps = np.asarray(sc.design.eps_r, dtype=np.float32).copy()
    rng = np.random.default_rng(5)
    for flat in rng.choice(eps.size, 4, replace=False):
        e0 = float(eps[flat])
        hi, lo = np.float32(e0 + 1e-3), np.float32(e0 - 1e-3)
        vals = []
        for v in (hi, lo):
            e = eps.copy()
            e[flat] = v
            sc.design = S.Design(sc.design.box, e)
            s = _sim(sc, steps)
            s.step(steps)
            vals.append(float(np.sum(np.abs(s.get_dft()) ** 2)))
            s.close()
        sc.design = S.Design(sc.design.box, eps)
        fd = (vals[0] - vals[1]) / (float(hi) - float(lo))
        assert g.reshape(-1)[flat] == pytest.approx(fd, rel=2e-5, abs=1e-6 * np.abs(g).max())


def test_fp32_gradient_close_to_fp64():
    sc, steps = design_scene()
    J64, g64 = _sim(sc, steps, 64).grad(checkpoint_every=40)
    J32, g32 = _sim(sc, steps, 32).grad(checkpoint_every=40)
    print(f"FP32 J {abs(J32 - J64) / J64:.2e}, gradient {_rel(g32, g64):.2e}", flush=True)
    assert J32 == pytest.approx(J64, rel=1e-5)
    assert _rel(g32, g64) < 1e-4


def test_torch_autograd_function():
    import torch
    from xfdtd_amd.torch_adjoint import design_objective
    sc, steps = design_scene()
    b = sc.design.box
    eps = torch.tensor(np.asarray(sc.design.eps_r, dtype=np.float32).reshape(b[1] - b[0], b[3] - b[2], b[5] - b[4]),
                       dtype=torch.float64, requires_grad=True)
    J = design_objective(eps, sc, steps, checkpoint_every=50)
    (3.0 * J).backward()
    Jr, gr = _sim(sc, steps).grad(checkpoint_every=50)
    assert float(J.detach()) == Jr
    assert np.array_equal(eps.grad.numpy(), 3.0 * gr)
    with torch.no_grad():
        J2 = design_objective(eps, sc, steps)
    assert float(J2) == pytest.approx(Jr, rel=1e-12)


def test_torch_design_problem_reuses_one_engine():
    import copy

    import torch
    from xfdtd_amd.torch_adjoint import DesignProblem
    sc, steps = design_scene()
    b = sc.design.box
    shape = (b[1] - b[0], b[3] - b[2], b[5] - b[4])
    base = np.asarray(sc.design.eps_r, dtype=np.float32).reshape(shape)
    prob = DesignProblem(sc, steps, checkpoint_every=50)
    rng = np.random.default_rng(17)
    try:
        for trial in range(3):
            e = base if trial != 1 else (base + rng.uniform(-0.3, 0.3, shape)).astype(np.float32)
            print(f"DesignProblem call {trial}: eps_r in [{e.min():.3f}, {e.max():.3f}]", flush=True)
            eps = torch.tensor(e, dtype=torch.float64, requires_grad=True)
            J = prob(eps)
            J.backward()
            sc2 = copy.copy(sc)
            sc2.design = S.Design(b, e.ravel())
            fresh = _sim(sc2, steps)
            Jr, gr = fresh.grad(checkpoint_every=50)
            fresh.close()
            assert float(J.detach()) == Jr
            assert np.array_equal(eps.grad.numpy(), gr)
    finally:
        prob.close()


def test_gpu_transpose_identity_with_drude(oracle_threads):
    steps = 25
    sc = drude_box_scene()
    g = _sim(sc, steps)
    box = g.drude_box()
    sl = np.s_[box[0]:box[1], box[2]:box[3], box[4]:box[5]]
    rng = np.random.default_rng(9)
    shapes = [g.field_shape(c) for c in list(range(6)) + [0, 1, 2]]
    x = [rng.standard_normal(sh) for sh in shapes]
    y = [rng.standard_normal(sh) for sh in shapes]
    for c in range(3):       # J is state on the Drude box only
        for v in (x, y):
            keep = np.zeros(shapes[c], bool)
            keep[sl] = True
            v[6 + c] *= keep
    for c in range(6):
        g.set_field(c, x[c])
    for c in range(3):
        g.set_drude_j(c, x[6 + c][sl])
    g.step(steps)
    out_j = []
    for c in range(3):
        full = np.zeros(shapes[c])
        full[sl] = g.get_drude_j(c)
        out_j.append(full)
    out = [g.get_field(c) for c in range(6)] + out_j
    lhs = sum(float(np.sum(a * b)) for a, b in zip(out, y))
    ly = _sim(sc, steps).adjoint_run(steps, y)
    rhs = sum(float(np.sum(a * b)) for a, b in zip(x, ly))
    scale = sum(float(np.sum(np.abs(a * b))) for a, b in zip(x, ly))
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs)
    lo = OracleSim(sc, steps).adjoint_run(steps, y)
    for c in range(9):
        assert _rel(ly[c], lo[c]) <= 1e-11, c


@pytest.mark.parametrize("every", [7, 1000])
def test_gradient_with_drude_matches_oracle(every, oracle_threads):
    sc, steps = drude_design_scene()
    Jo, go = OracleSim(sc, steps).grad()
    g = _sim(sc, steps)
    Jg, gg = g.grad(checkpoint_every=every)
    assert Jg == pytest.approx(Jo, rel=1e-14)
    assert _rel(gg, go) <= 1e-10
    _, g32 = _sim(sc, steps, 32).grad(checkpoint_every=every)
    print(f"Drude FP32 gradient {_rel(g32, go):.2e}", flush=True)
    assert _rel(g32, go) < 1e-4

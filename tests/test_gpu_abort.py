"""Bounded device waits.  Every wait of a persistent kernel (the colouring kernel's grid barriers, the
barrier solver's batches, the dataflow / lane / tile solvers' readiness polls, the residual probe)
gives up after the world's spin limit, or at its first poll once the host has cancelled the step, and
every kernel still returns: the step reports JH_ERR_TIMEOUT instead of hanging the GPU.  The conditions
are forced from the host (a spin limit of one poll, or the abort flag raised before the step), so no
kernel faults or spins without end.  Afterwards the world is set from the last good state and steps on
bit-exact against the oracle replaying its solve order."""
import numpy as np
import pytest

from jitterphysics2_amd import scenes
from jitterphysics2_amd._abi import JH_ERR_TIMEOUT, JH_OK, PROBE_COLD, PROBE_PREVIOUS
from jitterphysics2_amd.gpu import GpuWorld
from tests.orc import OracleWorld
from tests.test_gpu_step import assert_arbiters_equal, assert_bodies_equal

pytestmark = pytest.mark.gpu

SOLVERS = ["flow", "barrier", "lanes", "tiles", "world"]


def _settled(solver, monkeypatch, n=900, steps=30):
    monkeypatch.setenv("JH_SOLVER", solver)
    scene = scenes.heap(n, seed=5)
    bodies, shapes, verts, cfg, dt = scene
    gw = GpuWorld(bodies, shapes, verts, cfg)
    for _ in range(steps):
        gw.step(dt)
    return gw, scene


def _recovers(gw, scene, good_bodies, good_arbiters, steps=4):
    _, shapes, verts, cfg, dt = scene
    gw.set_spin_limit(0)
    gw.set_bodies(good_bodies)
    gw.set_arbiters(good_arbiters)
    ow = OracleWorld(good_bodies, shapes, verts, cfg)
    ow.set_arbiters(good_arbiters)
    for s in range(steps):
        gw.step(dt)
        ow.step(dt, gw.solve_order()[0])
    assert_bodies_equal(gw.bodies(), ow.bodies(), "after the aborted step")
    assert_arbiters_equal(gw.arbiters(), ow.arbiters(), "after the aborted step")


@pytest.mark.parametrize("solver", SOLVERS)
def test_cancelled_step_returns_timeout(monkeypatch, solver):
    gw, scene = _settled(solver, monkeypatch)
    good_b, good_a = gw.bodies(), gw.arbiters()
    print(f"abort: solver {solver}, host cancel", flush=True)
    gw.cancel()
    assert gw.step_status(scene[4]) == JH_ERR_TIMEOUT
    assert b"gave up" in gw.lib.jh_last_error()
    # the flag was cleared when the step reported it: the next step runs normally
    assert gw.step_status(scene[4]) == JH_OK
    _recovers(gw, scene, good_b, good_a)


@pytest.mark.parametrize("solver", SOLVERS)
def test_spin_limit_gives_up(monkeypatch, solver):
    gw, scene = _settled(solver, monkeypatch)
    good_b, good_a = gw.bodies(), gw.arbiters()
    gw.set_spin_limit(1)  # any wait that is not satisfied at its first poll gives up
    print(f"abort: solver {solver}, spin limit 1", flush=True)
    codes = [gw.step_status(scene[4]) for _ in range(3)]
    assert set(codes) <= {JH_OK, JH_ERR_TIMEOUT} and JH_ERR_TIMEOUT in codes, codes
    _recovers(gw, scene, good_b, good_a)


def test_cancelled_probe_returns_timeout(monkeypatch):
    gw, scene = _settled("flow", monkeypatch)
    good_b, good_a = gw.bodies(), gw.arbiters()
    print("abort: residual probe, host cancel", flush=True)
    gw.cancel()
    with pytest.raises(RuntimeError, match="timed out"):
        gw.step_probe(scene[4], iterations=8, modes=PROBE_PREVIOUS | PROBE_COLD)
    _recovers(gw, scene, good_b, good_a)


def test_default_limit_never_trips_on_a_formed_heap(monkeypatch):
    # the default limit is far above any legitimate wait: a formed heap steps without a timeout
    gw, scene = _settled("flow", monkeypatch, n=3000, steps=90)
    assert gw.stats()["arbiters"] > 500
    for _ in range(5):
        assert gw.step_status(scene[4]) == JH_OK
    assert np.all(np.isfinite(gw.bodies()["position"]))

"""World.RawData exchange in the reference's 128-byte FP32 RigidBodyData layout (RigidBody.cs:107-160):
every field read back at its offset (_index@0, _lockFlag@4, Position@8, Velocity@20, AngularVelocity@32,
DeltaVelocity@44, DeltaAngularVelocity@56, Orientation@68, InverseInertiaWorld@84, InverseMassVector@108,
Flags@120), a byte-exact round trip through jh_world_set_rigidbodydata at two strides, stepping after an
upload bit-exact against the oracle, and the rejected writes leaving the state unchanged."""
import numpy as np
import pytest

from jitterphysics2_amd import scenes
from jitterphysics2_amd.gpu import GpuWorld
from tests.orc import OracleWorld
from tests.test_gpu_step import assert_bodies_equal

pytestmark = pytest.mark.gpu

F = {"position": (8, 3), "velocity": (20, 3), "angular_velocity": (32, 3), "delta_velocity": (44, 3),
     "delta_angular_velocity": (56, 3), "orientation": (68, 4)}


def field(rec, off, n, dtype=np.float32):
    return np.ascontiguousarray(rec[:, off:off + 4 * n]).view(dtype).reshape(len(rec), n)


def rot(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def sym(m6):
    a, b, c, d, e, f = m6
    return np.array([[a, b, c], [b, d, e], [c, e, f]], np.float64)


def settled(n=600, steps=40):
    bodies, shapes, verts, cfg, dt = scenes.heap(n, seed=8)
    gw = GpuWorld(bodies, shapes, verts, cfg)
    for _ in range(steps):
        gw.step(dt)
    return gw, (bodies, shapes, verts, cfg, dt)


@pytest.mark.parametrize("stride", [128, 160])
def test_rigidbodydata_layout(stride):
    gw, (bodies, *_rest) = settled()
    rec = gw.rigidbodydata(stride)
    b = gw.bodies()
    n = len(b)
    assert np.array_equal(field(rec, 0, 1, np.int32)[:, 0], np.arange(n))   # _index
    assert np.all(field(rec, 4, 1, np.int32) == 0)                          # _lockFlag
    for name, (off, k) in F.items():
        assert np.array_equal(field(rec, off, k), b[name]), name
    flags = field(rec, 120, 1, np.int32)[:, 0]
    assert np.array_equal(flags & 3, b["motion_type"]) and np.all(flags & 4)  # MotionType | IsActive
    # InverseInertiaWorld = R I^-1 R^T and InverseMassVector = (m^-1, m^-1, m^-1) for dynamic bodies
    iw, imv = field(rec, 84, 6).astype(np.float64), field(rec, 108, 3).astype(np.float64)
    for i in np.nonzero(b["motion_type"] == 0)[0][:200]:
        R = rot(b["orientation"][i].astype(np.float64))
        want = R @ sym(bodies["inverse_inertia"][i]) @ R.T
        np.testing.assert_allclose(sym(iw[i]), want, rtol=1e-5, atol=1e-6 * np.abs(want).max())
        np.testing.assert_allclose(imv[i], [bodies["inverse_mass"][i]] * 3, rtol=0)
    static = b["motion_type"] == 2
    assert np.all(iw[static] == 0) and np.all(imv[static] == 0)
    if stride > 128:
        assert np.all(rec[:, 128:] == 0)


@pytest.mark.parametrize("stride", [128, 160])
def test_rigidbodydata_round_trip_and_step(stride):
    gw, (bodies, shapes, verts, cfg, dt) = settled()
    rec = gw.rigidbodydata(stride).copy()
    n = len(rec)
    dyn = np.nonzero(gw.bodies()["motion_type"] == 0)[0]
    pos, vel = field(rec, 8, 3), field(rec, 20, 3)
    pos[dyn] += np.float32(0.01)
    vel[dyn, 1] += np.float32(-0.5)
    rec[:, 8:20] = pos.view(np.uint8).reshape(n, 12)
    rec[:, 20:32] = vel.view(np.uint8).reshape(n, 12)
    gw.set_rigidbodydata(rec)
    back = gw.rigidbodydata(stride)
    assert np.array_equal(back[:, :128], rec[:, :128])  # every byte of every record
    b = gw.bodies()
    assert np.array_equal(b["position"], pos) and np.array_equal(b["velocity"], vel)
    # stepping from the uploaded state equals the oracle stepping from it
    st = bodies.copy()
    for f in ("position", "orientation", "velocity", "angular_velocity", "delta_velocity", "delta_angular_velocity"):
        st[f] = b[f]
    ow = OracleWorld(st, shapes, verts, cfg)
    ow.set_arbiters(gw.arbiters())
    for _ in range(3):
        gw.step(dt)
        ow.step(dt, gw.solve_order()[0])
    assert_bodies_equal(gw.bodies(), ow.bodies(), "after RawData upload")


def test_rigidbodydata_rejected_writes_change_nothing():
    gw, _ = settled(n=200, steps=5)
    rec = gw.rigidbodydata(128)
    before = gw.bodies()
    flags = field(rec, 120, 1, np.int32)[:, 0]
    dyn = int(np.nonzero(gw.bodies()["motion_type"] == 0)[0][0])

    def try_write(mod, want):
        r = rec.copy()
        mod(r)
        status = int(gw.lib.jh_world_set_rigidbodydata(gw.handle, r.ctypes.data, len(r), 128))
        assert status == want, (status, gw.lib.jh_last_error())
        assert np.array_equal(gw.bodies()["position"], before["position"])

    def set_flags(r, value):
        r[dyn, 120:124] = np.array([value], np.int32).view(np.uint8)

    try_write(lambda r: set_flags(r, int(flags[dyn]) ^ 2), -5)        # motion type change
    try_write(lambda r: set_flags(r, int(flags[dyn]) | 8), -5)        # gyroscopic forces
    try_write(lambda r: set_flags(r, int(flags[dyn]) | 16), -5)       # a motion lock
    nan = np.array([np.nan], np.float32).view(np.uint8)
    try_write(lambda r: r.__setitem__((dyn, slice(24, 28)), nan), -2)  # Velocity.y
    assert int(gw.lib.jh_world_set_rigidbodydata(gw.handle, rec.ctypes.data, len(rec) - 1, 128)) == -1
    assert int(gw.lib.jh_world_set_rigidbodydata(gw.handle, rec.ctypes.data, len(rec), 100)) == -1

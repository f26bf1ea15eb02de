"""Islands and sleeping in the oracle (Collision/IslandHelper.cs, World.Step.cs:1230-1371, the sleep
timer of RigidBody.Update, RigidBody.cs:716-743), pinned by the reference's own outcome tests
(Behavior/SleepTests.cs, Robustness/IslandPoolingTests.cs) and by what the definitions fix: the
incrementally kept islands equal the connected components of the contact and constraint graph over
the non-static bodies (brute force), sleeping bodies do not move, contacts and constraints are solved
only while a non-static (dynamic) body of them is awake, and AllowDeactivation = false never lets a
body sleep.  CPU only.

Writing body records wakes each written body for the next step (a static one wakes the bodies it
touches), as the reference's Position setter does; the reference's tests that set a property use that.
Not restated here: body removal (StackedWakeOnRemoval), AllowedMotion (LockedKinematicMotionDoesNotWakeBody)
and the Island events."""
import numpy as np
import pytest

from jitterphysics2_amd import constraints as C
from jitterphysics2_amd import scenes
from jitterphysics2_amd._abi import MOTION_KINEMATIC, MOTION_STATIC, default_config
from tests.orc import OracleWorld

DT = 0.01


def advance(w, seconds, dt=DT):
    """Helper.AdvanceWorld (src/JitterTests/Helper.cs:5-10): (int)(seconds / dt) steps in FP32."""
    n = int(np.float32(seconds) / np.float32(dt))
    for _ in range(n):
        w.step(dt)


def world(sb, gravity=None, **cfg_over):
    bodies, shapes, verts = sb.arrays()
    cfg = default_config()
    if gravity is not None:
        cfg["gravity"] = gravity
    for k, v in cfg_over.items():
        cfg[k] = v
    return OracleWorld(bodies, shapes, verts, cfg), bodies


def active(w):
    return w.activation()[0]


def set_field(w, b, **fields):
    """Property setters on one body (RigidBody.Position, Velocity, MotionType, DeactivationTime ...)."""
    rec = w_bodies_full(w)[b:b + 1].copy()
    for k, v in fields.items():
        rec[k] = v
        w._records[k][b] = v
    w.update_bodies([b], rec)


def w_bodies_full(w):
    """The full records: the builder's inputs with the oracle's state merged in."""
    st = w.bodies()
    out = w._records.copy()
    for f in ("position", "orientation", "velocity", "angular_velocity", "delta_velocity", "delta_angular_velocity",
              "force", "torque"):
        out[f] = st[f]
    return out


def make(sb, gravity=None, **cfg):
    w, bodies = world(sb, gravity, **cfg)
    w._records = bodies.copy()
    return w


def test_deactivation_time():  # SleepTests.cs:38-87
    sb = scenes.SceneBuilder()
    b = sb.body((0.0, 0.0, 0.0))
    sb.bodies[b]["deactivation_time"] = 3.0
    w = make(sb, gravity=(0.0, 0.0, 0.0))
    assert active(w)[b]
    advance(w, 1)
    assert active(w)[b]
    advance(w, 3)
    assert not active(w)[b]
    set_field(w, b, deactivation_time=5.0)
    w.set_activation_state(b, True)
    assert not active(w)[b]  # takes effect in the next step
    advance(w, 1)
    assert active(w)[b]
    advance(w, 3)
    assert active(w)[b]
    advance(w, 2)
    assert not active(w)[b]


def test_kinematic_trigger_wake():  # SleepTests.cs:89-124
    sb = scenes.SceneBuilder()
    p = sb.body((0.0, 0.0, 0.0), motion=MOTION_KINEMATIC)
    sb.box(p, (2.0, 2.0, 2.0))
    b = sb.body((0.0, 2.0, 0.0))
    sb.box(b, (1.0, 1.0, 1.0))
    w = make(sb, gravity=(0.0, 0.0, 0.0))
    advance(w, 3)
    assert not active(w)[b]
    set_field(w, p, velocity=(0.0, 1.0, 0.0))
    assert not active(w)[b]
    advance(w, 1)
    assert active(w)[b], "a moving kinematic body that reaches a sleeping one wakes it"
    assert w.bodies()["position"][p][1] > 0


def test_stacked_sleeping_and_wake_propagation():  # SleepTests.cs:127-165
    sb = scenes.SceneBuilder()
    ids = []
    for i in range(3):
        b = sb.body((0.0, float(i), 0.0), motion=MOTION_STATIC if i == 0 else 0)
        sb.box(b, (1.0, 1.0, 1.0))
        ids.append(b)
    w = make(sb)
    advance(w, 3)
    assert not active(w).any()
    w.set_activation_state(ids[1], True)
    advance(w, 1)
    a = active(w)
    assert a[ids[1]] and a[ids[2]], "waking the middle box wakes the one it touches"
    assert not a[ids[0]]
    advance(w, 2)
    assert not active(w).any()


def test_static_base_movement_wakes_dynamic():  # SleepTests.cs:210-243
    sb = scenes.SceneBuilder()
    bottom = sb.body((0.0, 0.0, 0.0), motion=MOTION_STATIC)
    sb.box(bottom, (2.0, 2.0, 2.0))
    top = sb.body((0.0, 2.0, 0.0))
    sb.box(top, (1.0, 1.0, 1.0))
    w = make(sb)
    advance(w, 3)
    assert not active(w)[bottom] and not active(w)[top]
    pos = w.bodies()["position"][bottom] - np.float32([0.0, 0.5, 0.0])
    set_field(w, bottom, position=pos)
    advance(w, 1)
    assert active(w)[top], "moving a static body wakes what rests on it"
    assert not active(w)[bottom]


def test_static_body_is_never_active():  # SleepTests.cs:6-36
    sb = scenes.SceneBuilder()
    body = sb.body((0.0, 0.0, 0.0), motion=MOTION_STATIC)
    sb.box(body, (1.0, 1.0, 1.0))
    dyn = sb.body((0.0, 1e4, 0.0))  # far away until moved next to the static body
    sb.box(dyn, (1.0, 1.0, 1.0))
    w = make(sb)
    advance(w, 1)
    assert not active(w)[body]
    w.set_activation_state(body, True)
    assert not active(w)[body]
    advance(w, 1)
    assert not active(w)[body]
    set_field(w, body, position=(0.0, 10.0, 0.0))
    assert not active(w)[body]
    advance(w, 1)
    assert not active(w)[body]
    set_field(w, dyn, position=(0.0, 12.0, 0.0), velocity=(0.0, 0.0, 0.0))
    advance(w, 1)
    assert not active(w)[body]
    assert len(w.arbiters()) == 1  # the dynamic box landed on it


@pytest.mark.parametrize("angular", [False, True])
@pytest.mark.parametrize("speed", [0.2, 1e-7, 1e-25])
@pytest.mark.parametrize("deact", [0.0, 0.5])
def test_moving_kinematic_body_stays_awake(angular, speed, deact):  # SleepTests.cs:289-318
    sb = scenes.SceneBuilder()
    b = sb.body((0.0, 0.0, 0.0), motion=MOTION_KINEMATIC)
    sb.box(b, (1.0, 1.0, 1.0))
    sb.bodies[b]["deactivation_time"] = deact
    v = np.float32([0.0, speed, 0.0]) if angular else np.float32([speed, 0.0, 0.0])
    if angular:
        sb.bodies[b]["angular_velocity"] = v
    else:
        sb.bodies[b]["velocity"] = v
    w = make(sb, gravity=(0.0, 0.0, 0.0))
    advance(w, 3)
    assert active(w)[b]
    st = w.bodies()[b]
    assert np.array_equal(st["angular_velocity"] if angular else st["velocity"], v)
    if angular:
        x, y, z, qw = st["orientation"].astype(np.float64)
        fwd = np.array([1 - 2 * (y * y + z * z), 2 * (x * y + z * qw), 2 * (x * z - y * qw)])
        disp = np.arctan2(-fwd[2], fwd[0])
    else:
        disp = float(st["position"][0])
    assert abs(disp - 3 * speed) <= 0.05 * speed


@pytest.mark.parametrize("angular", [False, True])
@pytest.mark.parametrize("speed", [1e-7, 1e-9, 1e-25])
def test_nonzero_velocity_wakes_sleeping_kinematic_body(angular, speed):  # SleepTests.cs:320-343
    sb = scenes.SceneBuilder()
    b = sb.body((0.0, 0.0, 0.0), motion=MOTION_KINEMATIC)
    sb.box(b, (1.0, 1.0, 1.0))
    sb.bodies[b]["deactivation_time"] = 0.1
    w = make(sb, gravity=(0.0, 0.0, 0.0))
    advance(w, 1)
    assert not active(w)[b]
    v = np.float32([speed, 0.0, 0.0])
    set_field(w, b, **{"angular_velocity" if angular else "velocity": v})
    w.step(DT)
    assert active(w)[b]
    advance(w, 1)
    assert active(w)[b]
    assert np.array_equal(w.bodies()[b]["angular_velocity" if angular else "velocity"], v)


@pytest.mark.parametrize("angular", [False, True])
def test_kinematic_body_sleeps_after_motion_stops(angular):  # SleepTests.cs:345-362
    sb = scenes.SceneBuilder()
    b = sb.body((0.0, 0.0, 0.0), motion=MOTION_KINEMATIC)
    sb.box(b, (1.0, 1.0, 1.0))
    sb.bodies[b]["deactivation_time"] = 0.1
    f = "angular_velocity" if angular else "velocity"
    sb.bodies[b][f] = (0.0, 0.2, 0.0) if angular else (0.2, 0.0, 0.0)
    w = make(sb, gravity=(0.0, 0.0, 0.0))
    advance(w, 1)
    assert active(w)[b]
    set_field(w, b, **{f: (0.0, 0.0, 0.0)})
    advance(w, 1)
    assert not active(w)[b]


def test_switching_to_kinematic_preserves_motion_and_stays_awake():  # SleepTests.cs:364-381
    sb = scenes.SceneBuilder()
    b = sb.body((0.0, 0.0, 0.0), velocity=(0.2, 0.0, 0.0), angular_velocity=(0.0, 0.2, 0.0))
    sb.box(b, (1.0, 1.0, 1.0))
    sb.bodies[b]["deactivation_time"] = 0.1
    w = make(sb, gravity=(0.0, 0.0, 0.0))
    w.step(DT)
    st = w.bodies()[b]
    set_field(w, b, motion_type=MOTION_KINEMATIC)
    advance(w, 1)
    assert active(w)[b]
    assert np.array_equal(w.bodies()[b]["velocity"], st["velocity"])
    assert np.array_equal(w.bodies()[b]["angular_velocity"], st["angular_velocity"])


def test_slow_kinematic_platform_carries_its_rider():  # SleepTests.cs:402-423
    sb = scenes.SceneBuilder()
    p = sb.body((0.0, 0.0, 0.0), motion=MOTION_KINEMATIC)
    sb.box(p, (4.0, 1.0, 4.0))
    sb.bodies[p]["deactivation_time"] = 0.5
    r = sb.body((0.0, 1.0, 0.0))
    sb.box(r, (1.0, 1.0, 1.0))
    sb.bodies[r]["deactivation_time"] = 0.5
    w = make(sb)
    advance(w, 1)
    st = w.bodies()
    gap = st["position"][r][1] - st["position"][p][1]
    assert not active(w).any()  # both asleep before the platform starts
    set_field(w, p, velocity=(0.0, 0.1, 0.0))
    advance(w, 4)
    st = w.bodies()
    assert abs(st["position"][p][1] - 0.4) <= 0.02
    assert abs((st["position"][r][1] - st["position"][p][1]) - gap) <= 0.02


def test_single_body_sleeps_after_its_deactivation_time():  # SleepTests.cs:463-495 (without the event)
    sb = scenes.SceneBuilder()
    b = sb.body((0.0, 0.0, 0.0))
    sb.box(b, (1.0, 1.0, 1.0))
    w = make(sb, gravity=(0.0, 0.0, 0.0))
    n = 0
    while active(w)[b]:
        w.step(DT)
        n += 1
    # the timer reaches 1 s at the end of step 100 (float sum of 0.01); the next CheckDeactivation
    # finds the island unmarked
    t = np.float32(0)
    k = 0
    while t < np.float32(1.0):
        t = np.float32(t + np.float32(DT))
        k += 1
    assert n == k + 1
    _, sl = w.activation()
    assert sl[b] == t


def test_allow_deactivation_false_never_sleeps():
    sb = scenes.SceneBuilder()
    b = sb.body((0.0, 0.0, 0.0))
    sb.box(b, (1.0, 1.0, 1.0))
    w = make(sb, gravity=(0.0, 0.0, 0.0), allow_deactivation=0)
    advance(w, 3)
    act, sl = w.activation()
    assert act[b] and sl[b] > 2.9  # the timer runs, the island stays awake


def test_removing_bridge_splits_islands_and_keeps_both_awake():  # IslandPoolingTests.cs:9-38
    sb = scenes.SceneBuilder()
    ids = [sb.body((float(i), 0.0, 0.0)) for i in range(5)]
    bodies, shapes, verts = sb.arrays()
    left, right = ids[:3], ids[3:]
    cons = [C.ball_socket(bodies, left[0], left[1], (0.0, 0.0, 0.0)),
            C.ball_socket(bodies, left[1], left[2], (0.0, 0.0, 0.0)),
            C.ball_socket(bodies, right[0], right[1], (0.0, 0.0, 0.0)),
            C.ball_socket(bodies, left[0], right[0], (0.0, 0.0, 0.0))]
    cfg = default_config()
    cfg["gravity"] = (0.0, 0.0, 0.0)
    w = OracleWorld(bodies, shapes, verts, cfg)
    w.set_constraints(C.as_array(cons))
    isl = w.islands()
    assert len(set(isl.tolist())) == 1
    w.set_constraints(C.as_array(cons[:3]))
    isl = w.islands()
    assert set(isl[left].tolist()) == {left[0]} and set(isl[right].tolist()) == {right[0]}
    w.step(DT)
    assert active(w).all()


def components(n, motion, edges):
    """Connected components over non-static bodies (union-find), named by their smallest index;
    a static body is its own component."""
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for a, b in edges:
        if motion[a] == MOTION_STATIC or motion[b] == MOTION_STATIC:
            continue
        ra, rb = find(a), find(b)
        if ra != rb:
            parent[max(ra, rb)] = min(ra, rb)
    root = [find(i) for i in range(n)]
    lo = {}
    for i in range(n):
        lo.setdefault(root[i], i)
    return np.array([lo[root[i]] for i in range(n)], np.int32)


def test_islands_are_the_connected_components_while_a_heap_settles():
    """Merges and two-sided splits kept incrementally equal the components of the current contact
    graph at every step (arbiters appear and break as the heap falls and spreads); sleeping bodies
    keep their state exactly; a contact is solved iff a non-static body of it is awake."""
    bodies, shapes, verts, cfg, dt = scenes.heap(300, seed=3)
    cfg["allow_deactivation"] = 1
    bodies["deactivation_time"] = 0.3
    w = OracleWorld(bodies, shapes, verts, cfg)
    motion = bodies["motion_type"]
    slept = 0
    prev_state, prev_act = w.bodies(), active(w)
    for s in range(420):
        w.step(dt)
        arbs = w.all_arbiters()  # with those the next step removes: still connections until then
        edges = list(zip(arbs["body1"].tolist(), arbs["body2"].tolist()))
        want = components(len(bodies), motion, edges)
        got = w.islands()
        assert np.array_equal(got, want), f"step {s}: islands differ from the components"
        act = active(w)
        st = w.bodies()
        # asleep before and after this step: not moved at all
        still = ~prev_act & ~act
        for f in ("position", "orientation", "velocity", "angular_velocity"):
            assert np.array_equal(st[f][still], prev_state[f][still]), f"step {s}: a sleeping body changed"
        # an island is uniformly awake or asleep
        dyn = motion != MOTION_STATIC
        for root in np.unique(got[dyn]):
            members = act[(got == root) & dyn]
            assert members.all() or not members.any(), f"step {s}: island {root} mixed"
        slept = max(slept, int((~act & dyn).sum()))
        prev_state, prev_act = st, act
    assert slept > 100, "the settled heap should go to sleep"


def towers_scene():
    """Three towers of four unit boxes on a static ground and a loose box beside them."""
    sb = scenes.SceneBuilder()
    g = sb.body((0.0, -0.5, 0.0), motion=MOTION_STATIC, friction=0.6)
    sb.box(g, (60.0, 1.0, 60.0))
    towers = []
    for x in (-5.0, 0.0, 5.0):
        t = []
        for k in range(4):
            b = sb.body((x, 0.5 + k, 0.0), friction=0.6)
            sb.box(b, (1.0, 1.0, 1.0))
            t.append(b)
        towers.append(t)
    loose = sb.body((20.0, 0.5, 0.0), friction=0.6)
    sb.box(loose, (1.0, 1.0, 1.0))
    return sb, towers, loose


def test_a_landing_body_wakes_only_the_island_it_hits():
    sb, towers, loose = towers_scene()
    w = make(sb)
    advance(w, 2)
    assert not active(w).any(), "the towers and the loose box should be asleep"
    isl = w.islands()
    assert len({int(isl[t[0]]) for t in towers} | {int(isl[loose])}) == 4
    top = towers[1][-1]
    pos = w.bodies()["position"][top] + np.float32([0.0, 3.0, 0.0])
    set_field(w, loose, position=pos, velocity=(0.0, -5.0, 0.0))
    woke = False
    for _ in range(60):
        w.step(DT)
        act = active(w)
        assert not act[towers[0]].any() and not act[towers[2]].any(), "only the tower hit may wake"
        woke = woke or act[towers[1]].all()
    assert woke, "the landing box wakes the tower it hits"

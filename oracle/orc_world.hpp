// ORACLE — test infrastructure only (see orc_math.hpp).  Contact cache (ContactData/Contact),
// the projected Gauss-Seidel solver (PrepareForIterationAccelerated / IterateAccelerated),
// integration, end-of-step body update, and the step orchestration of World.StepInternal,
// restated from Jitter2.  The solve order is either given (a replay of an order produced
// elsewhere) or the key order (ArbiterKey ascending), which is the deterministic order of
// World.Deterministic.cs:142-203 for a world with a single island.
//
// Threads (OpenMP, orc_set_threads): the narrowphase of each pair, the shape boxes, the per-body
// sweeps and the contact refresh run in parallel; each writes only its own pair's or body's or
// arbiter's record, and contacts are registered into the cache afterwards in pair order, so results
// are identical for any thread count.  The Gauss-Seidel solve itself is sequential.
#pragma once
#include <map>
#include <memory>
#include <utility>
#include <vector>

#include "orc_collision.hpp"
#include "orc_constraints.hpp"

namespace orc {

// RigidBody.RestrictLinearMotion / RestrictAngularMotion (RigidBody.cs:651-713); `locks` holds the
// locked axes (bits 0-2 linear x, y, z; 3-5 angular x, y, z).
template <class R> inline void restrict_linear(V3<R>& v, int locks) {
    if (!(locks & 7)) return;
    if (locks & 1) v.x = R(0);
    if (locks & 2) v.y = R(0);
    if (locks & 4) v.z = R(0);
}
template <class R> inline void restrict_angular(V3<R>& v, int locks) {
    if (!(locks & 56)) return;
    if (locks & 8) v.x = R(0);
    if (locks & 16) v.y = R(0);
    if (locks & 32) v.z = R(0);
}
// The inverse inertia with locked rotations eliminated by Schur complements, axis by axis.
template <class R> inline void restrict_inertia(Sym<R>& I, int locks) {
    if ((locks & 56) == 56) {
        I = Sym<R>{R(0), R(0), R(0), R(0), R(0), R(0)};
        return;
    }
    if (locks & 8) {
        if (I.m11 > R(0)) {
            R xy = I.m12 / I.m11, xz = I.m13 / I.m11;
            I.m22 -= xy * I.m12;
            I.m23 -= xy * I.m13;
            I.m33 -= xz * I.m13;
        }
        I.m11 = I.m12 = I.m13 = R(0);
    }
    if (locks & 16) {
        if (I.m22 > R(0)) {
            R xy = I.m12 / I.m22, yz = I.m23 / I.m22;
            I.m11 -= xy * I.m12;
            I.m13 -= xy * I.m23;
            I.m33 -= yz * I.m23;
        }
        I.m12 = I.m22 = I.m23 = R(0);
    }
    if (locks & 32) {
        if (I.m33 > R(0)) {
            R xz = I.m13 / I.m33, yz = I.m23 / I.m33;
            I.m11 -= xz * I.m13;
            I.m12 -= xz * I.m23;
            I.m22 -= yz * I.m23;
        }
        I.m13 = I.m23 = I.m33 = R(0);
    }
}

template <class R> struct Body {
    // RigidBodyData (RigidBody.cs:107-250)
    V3<R> pos, vel, angvel, dvel, dangvel;
    Q4<R> q;
    Sym<R> inv_inertia_world;
    V3<R> inv_mass_vec;
    int motion = JH_MOTION_DYNAMIC;
    // managed RigidBody fields read by RigidBody.Update (RigidBody.cs:398-409, 716-764)
    Sym<R> inv_inertia;
    R inv_mass = R(1);
    R friction = R(0.2), restitution = R(0);
    R lin_damp = R(0.998), ang_damp = R(0.995);
    V3<R> force{}, torque{};
    bool gravity = true;
    bool speculative = false;  // RigidBody.EnableSpeculativeContacts (RigidBody.cs:877-900)
    int world = 0;
    // sleeping (RigidBody.cs:315, 398-400; RigidBodyData.IsActive): island, connections to other
    // non-static bodies (one entry per arbiter or constraint, InternalConnections)
    bool active = true;
    R sleep_time = R(0);
    int locks = 0;  // locked MotionAxes (RigidBodyData.Flags bits 4..9 shifted down): 0 = AllowedMotion.All
    R sleep_lin_sq = R(0.1), sleep_ang_sq = R(0.1), deact_time = R(1.0);
    int island = -1;
    std::vector<int> conn;

    // RigidBody.UpdateWorldInertia (RigidBody.cs:903-923) with the AllowedMotion restrictions.
    void update_world_inertia() {
        if (motion == JH_MOTION_DYNAMIC) {
            M3<R> m = mat_from_quat(q);
            inv_inertia_world = sym_transform(inv_inertia, m);
            if (locks & 56) restrict_inertia(inv_inertia_world, locks);
            inv_mass_vec = {inv_mass, inv_mass, inv_mass};
            restrict_linear(inv_mass_vec, locks);
        } else {
            inv_inertia_world = Sym<R>{R(0), R(0), R(0), R(0), R(0), R(0)};
            inv_mass_vec = {R(0), R(0), R(0)};
        }
    }
};

// Contact (Contact.cs:617-661); the Vector128 rows NormalTangentX/Y/Z hold (n, t1, t2) lanes.
template <class R> struct Contact {
    static constexpr R MaximumBias = R(100.0), BiasFactor = R(0.2), AllowedPenetration = R(0.01),
                       BreakThreshold = R(0.02);
    enum { NewContact = 2, SkipWarmStart = 4 };
    uint32_t flag = 0;
    R bias = 0, penalty = 0;
    V3<R> n{}, t1{}, t2{};
    R mass[3] = {0, 0, 0};
    R acc[3] = {0, 0, 0};
    V3<R> pos1{}, pos2{}, rel1{}, rel2{};
};

template <class R> struct Arbiter {
    uint32_t key1 = 0, key2 = 0;  // shape indices (ascending shape id)
    int b1 = 0, b2 = 0;
    uint32_t usage = 0;
    uint32_t mode = 0;
    R friction = 0, restitution = 0, specrel = R(0.9);
    Contact<R> c[4];
};

template <class R> struct StepTiming {
    R dt, substep_dt, inv_dt, inv_substep_dt;
    int substeps;
    // TimeStep(dt, substepCount) (TimeStep.cs:36-46)
    StepTiming(R d, int n) : dt(d), substeps(n) {
        substep_dt = d / R(n);
        inv_dt = d > R(0) ? R(1.0) / d : R(0);
        inv_substep_dt = substep_dt > R(0) ? R(1.0) / substep_dt : R(0);
    }
};

template <class R> struct ManifoldOut {
    uint32_t k1, k2;
    bool epa;
    bool spec;  // speculative contact: registered with SolveMode.Angular removed
    V3<R> normal, pa, pb;
    R pen;
    int count;
    V3<R> mA[4], mB[4];
};

template <class R> struct World {
    std::vector<Body<R>> bodies;
    std::vector<Shape<R>> shapes;
    std::vector<float> verts;
    std::map<std::pair<uint32_t, uint32_t>, Arbiter<R>> arbiters;  // ArbiterKey -> ContactData
    V3<R> gravity{R(0), R(-9.81), R(0)};
    int substeps = 1, iterations = 6, relaxations = 4;
    R specrel = R(0.9);
    R spec_threshold = R(10.0);  // World.SpeculativeVelocityThreshold (World.Detect.cs:122)
    R broadphase_dt = R(0.1);    // World.BroadphaseStepDt (World.cs:416-418): StepDt of the last step
    bool aux_contacts = true, persistent = true, triangle_edge_filter = false;
    std::vector<ManifoldOut<R>> last_manifolds;
    std::vector<std::pair<uint32_t, uint32_t>> last_pairs;
    // Arbiter lifecycle: arbiters left without intact contacts by UpdateContacts (brokenArbiters) stay
    // until the next step's RemoveBrokenArbiters, which frees those still empty (EndCollide); arbiters
    // created by Detect raise BeginCollide (World.Step.cs:839-914).  Keys of the last step's events.
    std::vector<std::pair<uint32_t, uint32_t>> broken, ev_begin, ev_end;
    // Constraints, solved after the contacts in each pass (World.Step.cs:1155-1191), in con_order (a
    // replay of an order produced elsewhere) or, when empty, in index order.
    std::vector<Constraint<R>> constraints;
    std::vector<int> con_order;

    ConstraintBody<R> cbody(int b) {
        Body<R>& x = bodies[b];
        return ConstraintBody<R>{&x.vel, &x.angvel, x.pos, x.inv_mass_vec, x.q, x.inv_inertia_world,
                                 x.motion == JH_MOTION_DYNAMIC};
    }
    // the enabled constraints in the solver's active partition (UpdateConstraintActivation)
    template <class F> void for_constraints(F f) {
        if (con_order.empty()) {
            for (auto& c : constraints)
                if (c.enabled && constraint_active(c)) f(c);
        } else {
            for (int i : con_order)
                if (constraints[i].enabled && constraint_active(constraints[i])) f(constraints[i]);
        }
    }
    long epa_calls = 0;

    // ---- Islands and sleeping (Collision/IslandHelper.cs, World.cs:696-720, 900-910,
    // World.Step.cs:1230-1371) ----------------------------------------------------------------
    // Islands: connected components of non-static bodies through arbiters and constraints, kept
    // incrementally (merge smaller into larger on a new connection, a two-sided search and split when
    // one goes).  MarkedAsActive / NeedsUpdate are ORed on merges and copied on splits; the active
    // set is the PartitionedSet's active partition.
    struct Island {
        std::vector<int> bodies;
        bool marked = true, needs = false, active = true, alive = true;
    };
    std::vector<Island> isl;
    std::vector<int> isl_free;
    bool allow_deactivation = true;     // World.AllowDeactivation (World.cs:333)
    std::vector<std::pair<uint32_t, uint32_t>> deferred;  // arbiters created by this step's Detect

    int island_new() {
        int i;
        if (!isl_free.empty()) {
            i = isl_free.back();
            isl_free.pop_back();
            isl[i] = Island{};
        } else {
            i = int(isl.size());
            isl.emplace_back();
        }
        return i;  // GetFromPool: MarkedAsActive = true, NeedsUpdate = false
    }
    void island_free(int i) {
        isl[i].alive = false;
        isl[i].bodies.clear();
        isl_free.push_back(i);
    }
    bool is_static(int b) const { return bodies[b].motion == JH_MOTION_STATIC; }
    // World.AddToActiveList (World.cs:900-910)
    void add_to_active(int i) {
        bool was = isl[i].active;
        isl[i].marked = true;
        if (!was) isl[i].needs = true;
        isl[i].active = true;
    }
    // IslandHelper.MergeIslands (IslandHelper.cs:236-277): both bodies non-static
    void merge_islands(int b1, int b2) {
        int i1 = bodies[b1].island, i2 = bodies[b2].island;
        if (i1 == i2) return;
        bool needs = isl[i1].needs || isl[i2].needs;
        bool marked = isl[i1].marked || isl[i2].marked;
        int small = isl[i1].bodies.size() > isl[i2].bodies.size() ? i2 : i1;
        int large = small == i1 ? i2 : i1;
        for (int b : isl[small].bodies) {
            bodies[b].island = large;
            isl[large].bodies.push_back(b);
        }
        isl[large].needs = isl[large].needs || needs;
        isl[large].marked = isl[large].marked || marked;
        island_free(small);
    }
    // IslandHelper.AddConnection (:103-121)
    void add_connection(int b1, int b2) {
        bool needs = !isl[bodies[b1].island].active || !isl[bodies[b2].island].active;
        if (!is_static(b1) && !is_static(b2)) {
            merge_islands(b1, b2);
            bodies[b1].conn.push_back(b2);
            bodies[b2].conn.push_back(b1);
        }
        if (needs) {
            if (!is_static(b1)) isl[bodies[b1].island].needs = true;
            if (!is_static(b2)) isl[bodies[b2].island].needs = true;
        }
    }
    // IslandHelper.SplitIslands (:143-233): searches from both bodies alternately; if they meet, the
    // island stays; otherwise the side that ran out first becomes a new island with the flags and the
    // partition of the source island.
    void split_islands(int b1, int b2) {
        int src = bodies[b1].island;
        std::vector<int> mark(bodies.size(), 0);  // InternalIslandMarker
        std::vector<int> vl{b1}, vr{b2};
        size_t hl = 0, hr = 0;  // queue heads (the queues are vl[hl..], vr[hr..])
        mark[b1] = 1;
        mark[b2] = 2;
        while (hl < vl.size() && hr < vr.size()) {
            int cur = vl[hl++];
            if (!is_static(cur))
                for (int c : bodies[cur].conn) {
                    if (mark[c] == 0) { vl.push_back(c); mark[c] = 1; }
                    else if (mark[c] == 2) return;
                }
            cur = vr[hr++];
            if (!is_static(cur))
                for (int c : bodies[cur].conn) {
                    if (mark[c] == 0) { vr.push_back(c); mark[c] = 2; }
                    else if (mark[c] == 1) return;
                }
        }
        int ni = island_new();
        isl[ni].needs = isl[src].needs;
        isl[ni].marked = isl[src].marked;
        isl[ni].active = isl[src].active;
        const std::vector<int>& moved = hl == vl.size() ? vl : vr;
        for (int b : moved) {
            auto& sb = isl[src].bodies;
            sb.erase(std::find(sb.begin(), sb.end(), b));
            isl[ni].bodies.push_back(b);
            bodies[b].island = ni;
        }
    }
    // IslandHelper.RemoveConnection (:123-141)
    void remove_connection(int b1, int b2) {
        auto rm = [](std::vector<int>& l, int b) {
            auto it = std::find(l.begin(), l.end(), b);
            if (it == l.end()) return;
            *it = l.back();
            l.pop_back();
        };
        rm(bodies[b1].conn, b2);
        rm(bodies[b2].conn, b1);
        if (bodies[b1].island == bodies[b2].island) split_islands(b1, b2);
    }
    // World.ActivateBodyNextStep (World.cs:696-720); with wake_static a static body's island is put on
    // the active list and the bodies it touches through arbiters and constraints are woken.
    void activate_body_next_step(int b, bool wake_static = false) {
        bodies[b].sleep_time = R(0);
        if (bodies[b].active) return;
        if (is_static(b) && !wake_static) return;
        add_to_active(bodies[b].island);
        if (is_static(b)) {
            for (auto& c : constraints)
                if (c.b1 == b || c.b2 == b) activate_body_next_step(c.b1 == b ? c.b2 : c.b1);
            for (auto& kv : arbiters)
                if (kv.second.b1 == b || kv.second.b2 == b)
                    activate_body_next_step(kv.second.b1 == b ? kv.second.b2 : kv.second.b1);
        }
        isl[bodies[b].island].needs = true;
    }
    // The rest of the AllowedMotion setter (RigidBody.cs:589-592): ClearContactCache (usage masks of
    // the body's arbiters to zero) and ResetWarmStart of its constraints (their accumulated impulses).
    void locks_changed(int b) {
        for (auto& kv : arbiters)
            if (kv.second.b1 == b || kv.second.b2 == b) kv.second.usage = 0;
        for (auto& c : constraints)
            if (c.b1 == b || c.b2 == b)
                for (int k = 0; k < constraint_impulse_count(c.type); ++k) c.s[k] = R(0);
    }
    // Islands rebuilt from scratch: every body awake in its own island (IslandHelper.BodyAdded and
    // AddToActiveList of World.CreateRigidBody), then every constraint and arbiter connected
    // (ConstraintCreated / ArbiterCreated, each followed by AddToActiveList of both islands).
    void islands_reset() {
        isl.clear();
        isl_free.clear();
        for (size_t b = 0; b < bodies.size(); ++b) {
            bodies[b].conn.clear();
            bodies[b].active = true;
            // World.CreateRigidBody leaves the timer at zero; MotionType = Static sets it to infinity
            // (DeactivateBodyNextStep, World.cs:793-800)
            bodies[b].sleep_time = is_static(int(b)) ? std::numeric_limits<R>::infinity() : R(0);
            int i = island_new();
            isl[i].bodies.push_back(int(b));
            bodies[b].island = i;
        }
        for (auto& c : constraints) {
            add_connection(c.b1, c.b2);
            add_to_active(bodies[c.b1].island);
            add_to_active(bodies[c.b2].island);
        }
        for (auto& kv : arbiters) {
            add_connection(kv.second.b1, kv.second.b2);
            add_to_active(bodies[kv.second.b1].island);
            add_to_active(bodies[kv.second.b2].island);
        }
    }
    // World.CheckDeactivation (World.Step.cs:1289-1371)
    void check_deactivation() {
        for (size_t i = 0; i < isl.size(); ++i) {
            Island& I = isl[i];
            if (!I.alive || !I.active) continue;
            bool deactivate = !I.marked;
            if (!allow_deactivation) deactivate = false;
            I.marked = false;
            bool needs = I.needs;
            I.needs = false;
            if (!deactivate && !needs) continue;
            for (int b : I.bodies) {
                Body<R>& x = bodies[b];
                if (x.active != deactivate) {
                    if (!needs) break;
                    continue;
                }
                if (deactivate) {
                    x.active = false;  // DeactivateBodyForSleep: velocities are kept
                } else {
                    if (x.motion == JH_MOTION_STATIC) continue;
                    x.active = true;
                    x.sleep_time = R(0);
                }
            }
            if (deactivate) I.active = false;
        }
    }
    // memContacts' active partition: a contact is active while a non-static body of it is awake
    // (contacts move with their non-static bodies' activation, World.Step.cs:1245-1264, 1325-1328)
    bool arbiter_active(const Arbiter<R>& a) const {
        const Body<R>& x = bodies[a.b1];
        const Body<R>& y = bodies[a.b2];
        return (x.motion != JH_MOTION_STATIC && x.active) || (y.motion != JH_MOTION_STATIC && y.active);
    }
    // World.UpdateConstraintActivation (World.cs:741-760)
    bool constraint_active(const Constraint<R>& c) const {
        const Body<R>& x = bodies[c.b1];
        const Body<R>& y = bodies[c.b2];
        return (x.motion == JH_MOTION_DYNAMIC && x.active) || (y.motion == JH_MOTION_DYNAMIC && y.active);
    }

    // RigidBodyShape.UpdateWorldBoundingBox(BroadphaseStepDt) (RigidBodyShape.cs:37-58): bodies with
    // speculative contacts get their box grown by dt * |velocity| on every side.
    void world_aabb(const Shape<R>& s, V3<R>& mn, V3<R>& mx) const {
        const Body<R>& b = bodies[s.body];
        s.aabb(b.q, b.pos, mn, mx);
        if (b.speculative) {
            R swept = broadphase_dt * length(b.vel);
            mn = mn - V3<R>{swept, swept, swept};
            mx = mx + V3<R>{swept, swept, swept};
        }
    }

    // The pair set visited by DynamicTree.EnumerateOverlaps (DynamicTree.cs:266-286) after the
    // World.DefaultDynamicTreeFilter (World.cs:447-455) and the Detect gating (World.Detect.cs:158-160):
    // shapes of different bodies in the same world, at least one body dynamic, world AABBs not
    // disjoint.  Sort-and-sweep along x; returns pairs (lower index, higher index) sorted.
    std::vector<std::pair<uint32_t, uint32_t>> compute_pairs(bool brute = false) const {
        size_t n = shapes.size();
        std::vector<V3<R>> mn(n), mx(n);
#pragma omp parallel for schedule(static) if (n > 4096)
        for (long i = 0; i < long(n); ++i) world_aabb(shapes[i], mn[i], mx[i]);
        std::vector<std::pair<uint32_t, uint32_t>> out;
        auto accept = [&](size_t i, size_t j) {
            const Shape<R>& a = shapes[i];
            const Shape<R>& b = shapes[j];
            if (a.body == b.body) return;
            const Body<R>& ba = bodies[a.body];
            const Body<R>& bb = bodies[b.body];
            if (ba.world != bb.world) return;
            if (!ba.active && !bb.active) return;  // World.Detect.cs:158
            if (ba.motion != JH_MOTION_DYNAMIC && bb.motion != JH_MOTION_DYNAMIC) return;
            if (disjoint(mn[i], mx[i], mn[j], mx[j])) return;
            out.emplace_back(uint32_t(std::min(i, j)), uint32_t(std::max(i, j)));
        };
        if (brute) {
            for (size_t i = 0; i < n; ++i)
                for (size_t j = i + 1; j < n; ++j) accept(i, j);
        } else {
            // sort by (world, min x) and sweep within each world: pairs never cross worlds
            std::vector<uint32_t> order(n);
            for (size_t i = 0; i < n; ++i) order[i] = uint32_t(i);
            auto wid = [&](uint32_t i) { return bodies[shapes[i].body].world; };
            std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
                int wa = wid(a), wb = wid(b);
                if (wa != wb) return wa < wb;
                return mn[a].x < mn[b].x || (mn[a].x == mn[b].x && a < b);
            });
            for (size_t oi = 0; oi < n; ++oi) {
                uint32_t i = order[oi];
                for (size_t oj = oi + 1; oj < n; ++oj) {
                    uint32_t j = order[oj];
                    if (wid(j) != wid(i) || mn[j].x > mx[i].x) break;
                    accept(i, j);
                }
            }
        }
        std::sort(out.begin(), out.end());
        return out;
    }

    // ContactData.Init (Contact.cs:271-284).
    void init_arbiter(Arbiter<R>& a, uint32_t k1, uint32_t k2) {
        a = Arbiter<R>{};
        a.key1 = k1;
        a.key2 = k2;
        a.b1 = shapes[k1].body;
        a.b2 = shapes[k2].body;
        a.friction = rmax(bodies[a.b1].friction, bodies[a.b2].friction);
        a.restitution = rmax(bodies[a.b1].restitution, bodies[a.b2].restitution);
        a.specrel = specrel;
        a.mode = 0;
        a.usage = 0;
    }
    // ContactData.ResetMode (Contact.cs:291-296).
    void reset_mode(Arbiter<R>& a, uint32_t remove = 0) {
        a.mode = (bodies[a.b1].motion == JH_MOTION_DYNAMIC) ? (JH_SOLVE_LINEAR1 | JH_SOLVE_ANGULAR1) : 0;
        a.mode |= (bodies[a.b2].motion == JH_MOTION_DYNAMIC) ? (JH_SOLVE_LINEAR2 | JH_SOLVE_ANGULAR2) : 0;
        a.mode &= ~remove;
    }

    // Contact.Initialize (Contact.cs:703-759).
    void contact_initialize(Arbiter<R>& cd, Contact<R>& c, V3<R> p1, V3<R> p2, V3<R> n, bool is_new) {
        const Body<R>& b1 = bodies[cd.b1];
        const Body<R>& b2 = bodies[cd.b2];
        c.rel1 = p1 - b1.pos;
        c.rel2 = p2 - b2.pos;
        c.pos1 = rotate_inv(c.rel1, b1.q);
        c.pos2 = rotate_inv(c.rel2, b2.q);
        if (!is_new) return;
        c.flag = Contact<R>::NewContact;
        c.acc[0] = c.acc[1] = c.acc[2] = R(0.0);
        V3<R> dv = b2.vel - b1.vel;
        if (cd.mode == 0 || (cd.mode & (JH_SOLVE_ANGULAR1 | JH_SOLVE_ANGULAR2)) != 0) {
            dv = b2.vel + cross(b2.angvel, c.rel2);
            dv = dv - (b1.vel + cross(b1.angvel, c.rel1));
        }
        R rel_nv = dot(dv, n);
        c.bias = 0;
        if (rel_nv < R(-1.0)) c.bias = -cd.restitution * rel_nv;
        V3<R> t1 = dv - n * rel_nv;
        R num = len_sq(t1);
        if (num > R(1e-12)) {
            num = R(1.0) / std::sqrt(num);
            t1 = t1 * num;
        } else {
            t1 = create_orthonormal(n);
        }
        V3<R> t2 = cross(t1, n);
        c.n = n;
        c.t1 = t1;
        c.t2 = t2;
    }

    // CalcArea4Points (Contact.cs:418-432) as evaluated lane-wise by the vectorized selector.
    static R area4(V3<R> p0, V3<R> p1, V3<R> p2, V3<R> p3) {
        V3<R> a0 = p0 - p1, a1 = p0 - p2, a2 = p0 - p3, b0 = p2 - p3, b1 = p1 - p3, b2 = p1 - p2;
        R c0 = len_sq(cross(a0, b0)), c1 = len_sq(cross(a1, b1)), c2 = len_sq(cross(a2, b2));
        return rmax(rmax(c0, c1), c2);
    }
    // SortCachedPoints (Contact.cs:529-553) with SelectCachedPointReplacementMask (:434-466).
    void sort_cached_points(Arbiter<R>& cd, V3<R> p1, V3<R> p2, V3<R> n) {
        V3<R> rp1 = p1 - bodies[cd.b1].pos;
        R a0 = area4(rp1, cd.c[1].rel1, cd.c[2].rel1, cd.c[3].rel1);
        R a1 = area4(rp1, cd.c[0].rel1, cd.c[2].rel1, cd.c[3].rel1);
        R a2 = area4(rp1, cd.c[0].rel1, cd.c[1].rel1, cd.c[3].rel1);
        R a3 = area4(rp1, cd.c[0].rel1, cd.c[1].rel1, cd.c[2].rel1);
        const R eps = -R(0.0001);
        R big = 0;
        uint32_t mask = 0;
        if (a0 > big + eps) { big = a0; mask = 1; }
        if (a1 > big + eps) { big = a1; mask = 2; }
        if (a2 > big + eps) { big = a2; mask = 4; }
        if (a3 > big + eps) { mask = 8; }
        int slot = mask == 2 ? 1 : mask == 4 ? 2 : mask == 8 ? 3 : 0;
        contact_initialize(cd, cd.c[slot], p1, p2, n, false);
        cd.usage |= mask;
    }
    // ContactData.AddContact (Contact.cs:329-416).
    void add_contact(Arbiter<R>& cd, V3<R> p1, V3<R> p2, V3<R> n) {
        if ((cd.usage & 15u) == 15u) { sort_cached_points(cd, p1, p2, n); return; }
        int closest = -1;
        R dsq = std::numeric_limits<R>::max();
        V3<R> rel = p1 - bodies[cd.b1].pos;
        for (int i = 0; i < 4; ++i) {
            if ((cd.usage & (1u << i)) == 0) continue;
            R d = len_sq(cd.c[i].rel1 - rel);
            if (d < dsq) { dsq = d; closest = i; }
        }
        if (dsq < Contact<R>::BreakThreshold * Contact<R>::BreakThreshold) {
            contact_initialize(cd, cd.c[closest], p1, p2, n, false);
            return;
        }
        for (int i = 0; i < 4; ++i) {
            if ((cd.usage & (1u << i)) == 0) {
                contact_initialize(cd, cd.c[i], p1, p2, n, true);
                cd.usage |= (1u << i);
                return;
            }
        }
    }

    // World.Detect (World.Detect.cs:125-218) for one broadphase pair (k1 < k2), default filters,
    // no speculative contacts; the default TriangleEdgeCollisionFilter only acts on TriangleShape.
    // ---- TriangleEdgeCollisionFilter (Collision/CollisionFilter/TriangleEdgeCollisionFilter.cs) ----
    // Triangle shapes carry their normal and the normals of NeighborA/B/C in p[0..11] (zero = none).
    enum { TefKeep = 0, TefDiscard = 1, TefEdge = 2 };
    static constexpr R kEdgeThreshold = R(0.01), kCosAngle = R(0.999), kProjectionThreshold = R(0.5);

    // EvaluateTriangleContact (:86-147)
    int tef_evaluate(const Shape<R>& t, const Body<R>& tb, bool tri_is_b, V3<R> cp, V3<R> normal, V3<R>& tri_n,
                     V3<R>& nb_n) const {
        tri_n = rotate(V3<R>{t.p[0], t.p[1], t.p[2]}, tb.q);
        nb_n = V3<R>{R(0), R(0), R(0)};
        if (tri_is_b) tri_n = -tri_n;
        if (dot(normal, tri_n) < -kCosAngle) return TefDiscard;
        V3<R> a = rotate(V3<R>{R(t.verts[0]), R(t.verts[1]), R(t.verts[2])}, tb.q);
        V3<R> b = rotate(V3<R>{R(t.verts[3]), R(t.verts[4]), R(t.verts[5])}, tb.q);
        V3<R> c = rotate(V3<R>{R(t.verts[6]), R(t.verts[7]), R(t.verts[8])}, tb.q);
        a = a + tb.pos;
        b = b + tb.pos;
        c = c + tb.pos;
        V3<R> nrm = cross(b - a, c - a);
        R inv = R(1.0) / std::sqrt(len_sq(nrm));
        nrm = V3<R>{nrm.x * inv, nrm.y * inv, nrm.z * inv};
        R dist = dot(cp - a, nrm);
        V3<R> pp = cp - nrm * dist;
        V3<R> n = b - a;
        V3<R> pma = pp - a;
        R d0 = len_sq(pma - (n * dot(pma, n)) * (R(1.0) / len_sq(n)));
        n = c - a;
        pma = pp - a;
        R d1 = len_sq(pma - (n * dot(pma, n)) * (R(1.0) / len_sq(n)));
        n = c - b;
        pma = pp - b;
        R d2 = len_sq(pma - (n * dot(pma, n)) * (R(1.0) / len_sq(n)));
        if (std::min(std::min(d0, d1), d2) > kEdgeThreshold * kEdgeThreshold) return TefKeep;
        int k = (d0 < d1 && d0 < d2) ? 3 : (d1 < d2 ? 2 : 1);  // NeighborC / NeighborB / NeighborA
        V3<R> nn{t.p[3 * k], t.p[3 * k + 1], t.p[3 * k + 2]};
        if (nn.x == R(0) && nn.y == R(0) && nn.z == R(0)) return TefKeep;
        nb_n = rotate(nn, tb.q);
        if (tri_is_b) nb_n = -nb_n;
        return TefEdge;
    }

    // Filter (:149-323); false = discard.  r may be replaced by the re-run with EPA threshold 0.
    bool tef_filter(const Shape<R>& sA, const Shape<R>& sB, const Body<R>& b1, const Body<R>& b2, MprResult<R>& r,
                    Polytope<R>* pt, long& epa) const {
        bool c1 = sA.type == JH_SHAPE_TRIANGLE, c2 = sB.type == JH_SHAPE_TRIANGLE;
        if (c1 == c2) return true;
        const Shape<R>& t = c1 ? sA : sB;
        const Body<R>& tb = c1 ? b1 : b2;
        if (tb.motion == JH_MOTION_DYNAMIC) return true;
        V3<R> tn, nn;
        int e = tef_evaluate(t, tb, c2, c1 ? r.pa : r.pb, r.normal, tn, nn);
        if (e == TefDiscard) return false;
        if (e == TefKeep) return true;
        bool spec = r.penetration < R(0);
        if (!spec) {
            MprResult<R> r2 = mpr_epa(sA, sB, b1.q, b2.q, b1.pos, b2.pos, pt, R(0));
            if (r2.epa) epa++;
            if (!r2.hit) return false;
            r = r2;
            e = tef_evaluate(t, tb, c2, c1 ? r.pa : r.pb, r.normal, tn, nn);
            if (e == TefDiscard) return false;
            if (e == TefKeep) return true;
        }
        V3<R> mid = (r.pa + r.pb) * R(0.5);
        if (dot(tn, nn) > kCosAngle) {
            R f5 = dot(r.normal, nn), f6 = dot(r.normal, tn);
            if (!spec) {
                r.penetration = R(0);
                r.pa = mid;
                r.pb = mid;
            }
            r.normal = f5 > f6 ? nn : tn;
            return true;
        }
        V3<R> cr = cross(nn, tn);
        R cl = len_sq(cr);
        V3<R> proj = r.normal - cr * (dot(cr, r.normal) / cl);
        if (len_sq(proj) < kProjectionThreshold * kProjectionThreshold) return false;
        R f1 = dot(cross(proj, nn), cr), f2 = dot(cross(proj, tn), cr);
        bool between = f1 <= R(0) && f2 >= R(0);
        if (!between) {
            R f3 = dot(r.normal, nn), f4 = dot(r.normal, tn);
            if (!spec) {
                r.penetration = R(0);
                r.pa = mid;
                r.pb = mid;
            }
            r.normal = f3 > f4 ? nn : tn;
        }
        return true;
    }

    // GetOrCreateArbiter (World.Detect.cs:417-430) with ContactData.Init (Contact.cs:271-284).
    Arbiter<R>& arbiter_for(uint32_t k1, uint32_t k2) {
        auto key = std::make_pair(k1, k2);
        auto it = arbiters.find(key);
        if (it == arbiters.end()) {
            Arbiter<R> a;
            init_arbiter(a, k1, k2);
            it = arbiters.emplace(key, a).first;
            ev_begin.push_back(key);  // HandleDeferredArbiters raises BeginCollide (World.Step.cs:881-911)
            deferred.push_back(key);
        }
        return it->second;
    }

    // Narrowphase result of one pair (World.Detect.cs:125-218): what RegisterContact will be given.
    struct Detected {
        int kind = 0;  // 0 nothing, 1 contact manifold, 2 speculative contact
        ManifoldOut<R> mo{};
        long epa = 0;  // EPA runs (including the edge filter's re-run)
    };

    // The speculative branch of World.Detect (World.Detect.cs:167-194) for a pair MPR found separated:
    // a linear sweep; a time of impact inside this step gives one contact point whose solve mode
    // drops the angular terms (RegisterContact(..., SolveMode.Angular), World.Detect.cs:331-337, 232-250).
    void detect_speculative(uint32_t k1, uint32_t k2, R step_dt, Polytope<R>* pt, Detected& d) const {
        const Shape<R>& sA = shapes[k1];
        const Shape<R>& sB = shapes[k2];
        const Body<R>& b1 = bodies[sA.body];
        const Body<R>& b2 = bodies[sB.body];
        if (!b1.speculative && !b2.speculative) return;
        V3<R> dv = b2.vel - b1.vel;
        if (len_sq(dv) < spec_threshold * spec_threshold) return;
        MprResult<R> r{};
        R toi;
        bool ok = sweep(sA, sB, b1.q, b2.q, b1.pos, b2.pos, b1.vel, b2.vel, r.pa, r.pb, r.normal, toi);
        if (!ok || toi > step_dt || toi == R(0.0)) return;
        r.hit = true;
        r.epa = false;
        r.penetration = dot(r.normal, r.pa - r.pb) * specrel;
        if (triangle_edge_filter && !tef_filter(sA, sB, b1, b2, r, pt, d.epa)) return;
        ManifoldOut<R>& mo = d.mo;
        mo.k1 = k1;
        mo.k2 = k2;
        mo.epa = false;
        mo.spec = true;
        mo.normal = r.normal;
        mo.pa = r.pa;
        mo.pb = r.pb;
        mo.pen = r.penetration;
        mo.count = 1;
        mo.mA[0] = r.pa;
        mo.mB[0] = r.pb;
        d.kind = 2;
    }

    // World.Detect (World.Detect.cs:125-218) up to the RegisterContact call: reads the world only.
    Detected detect_pair(uint32_t k1, uint32_t k2, R step_dt, Polytope<R>* pt) const {
        Detected d;
        const Shape<R>& sA = shapes[k1];
        const Shape<R>& sB = shapes[k2];
        const Body<R>& b1 = bodies[sA.body];
        const Body<R>& b2 = bodies[sB.body];
        if (b1.motion != JH_MOTION_DYNAMIC && b2.motion != JH_MOTION_DYNAMIC) return d;
        MprResult<R> r = mpr_epa(sA, sB, b1.q, b2.q, b1.pos, b2.pos, pt);
        if (r.epa) d.epa++;
        if (!r.hit) {
            detect_speculative(k1, k2, step_dt, pt, d);
            return d;
        }
        if (triangle_edge_filter && !tef_filter(sA, sB, b1, b2, r, pt, d.epa)) return d;
        ManifoldOut<R>& mo = d.mo;
        mo.k1 = k1;
        mo.k2 = k2;
        mo.epa = r.epa;
        mo.normal = r.normal;
        mo.pa = r.pa;
        mo.pb = r.pb;
        mo.pen = r.penetration;
        Manifold<R> m;
        if (aux_contacts) {
            m.build(sA, sB, b1.q, b2.q, b1.pos, b2.pos, r.pa, r.pb, r.normal);
        } else {
            m.count = 1;
            m.mA[0] = r.pa;
            m.mB[0] = r.pb;
        }
        mo.count = m.count;
        for (int i = 0; i < m.count; ++i) { mo.mA[i] = m.mA[i]; mo.mB[i] = m.mB[i]; }
        d.kind = 1;
        return d;
    }

    // RegisterContact (World.Detect.cs:294-324), in pair order.
    void register_detected(const Detected& d) {
        epa_calls += d.epa;
        if (d.kind == 0) return;
        const ManifoldOut<R>& mo = d.mo;
        last_manifolds.push_back(mo);
        Arbiter<R>& cd = arbiter_for(mo.k1, mo.k2);
        if (!persistent) cd.usage = 0;
        if (d.kind == 2) {
            reset_mode(cd, JH_SOLVE_ANGULAR1 | JH_SOLVE_ANGULAR2);
            add_contact(cd, mo.pa, mo.pb, mo.normal);
        } else {
            reset_mode(cd);
            for (int e = 0; e < mo.count; ++e) add_contact(cd, mo.mA[e], mo.mB[e], mo.normal);
        }
    }

    // Narrowphase over a pair list: parallel detection, registration in list order.
    void detect_all(const std::vector<std::pair<uint32_t, uint32_t>>& pairs, R step_dt) {
        std::vector<Detected> out(pairs.size());
        long n = long(pairs.size());
#pragma omp parallel if (n > 256)
        {
            std::unique_ptr<Polytope<R>> pt(new Polytope<R>());
#pragma omp for schedule(dynamic, 64)
            for (long i = 0; i < n; ++i) out[i] = detect_pair(pairs[i].first, pairs[i].second, step_dt, pt.get());
        }
        for (auto& d : out) register_detected(d);
    }

    static R sum3(const R* w, R a, R b, R c) { return w[0] * a + w[1] * b + w[2] * c; }

    // Contact.PrepareForIterationAccelerated (Contact.cs:1038-1180), no linear locks.
    void prepare_contact(Arbiter<R>& cd, Contact<R>& c, const StepTiming<R>& ts) {
        Body<R>& b1 = bodies[cd.b1];
        Body<R>& b2 = bodies[cd.b2];
        bool skip = (c.flag & Contact<R>::SkipWarmStart) != 0;
        c.flag &= ~uint32_t(Contact<R>::SkipWarmStart);
        R ws[3] = {skip ? R(0) : c.acc[0], skip ? R(0) : c.acc[1], skip ? R(0) : c.acc[2]};
        V3<R> cr1 = rotate(c.pos1, b1.q);
        V3<R> cr2 = rotate(c.pos2, b2.q);
        V3<R> imv{R(0), R(0), R(0)};
        if (cd.mode & JH_SOLVE_LINEAR1) imv = b1.inv_mass_vec;
        if (cd.mode & JH_SOLVE_LINEAR2) imv = imv + b2.inv_mass_vec;
        // Rows: X = (n.x, t1.x, t2.x), Y, Z.
        R NX[3] = {c.n.x, c.t1.x, c.t2.x}, NY[3] = {c.n.y, c.t1.y, c.t2.y}, NZ[3] = {c.n.z, c.t1.z, c.t2.z};
        R k[3] = {imv.x, imv.x, imv.x};
        if ((b1.locks & 7) || (b2.locks & 7))  // directional linear mass with linear locks
            for (int l = 0; l < 3; ++l) k[l] = (NX[l] * imv.x) * NX[l] + (NY[l] * imv.y) * NY[l] + (NZ[l] * imv.z) * NZ[l];
        if ((c.flag & Contact<R>::NewContact) == 0) c.bias = 0;
        V3<R> r1 = cr1 + b1.pos, r2 = cr2 + b2.pos;
        V3<R> dist = r1 - r2;
        R pen = dot(dist, c.n);
        if (pen < -Contact<R>::BreakThreshold) c.bias = pen * ts.inv_dt * cd.specrel;
        c.flag &= ~uint32_t(Contact<R>::NewContact);
        R ktx[3] = {0, 0, 0}, kty[3] = {0, 0, 0}, ktz[3] = {0, 0, 0};
        V3<R> lin{sum3(ws, NX[0], NX[1], NX[2]), sum3(ws, NY[0], NY[1], NY[2]), sum3(ws, NZ[0], NZ[1], NZ[2])};
        if (cd.mode & JH_SOLVE_LINEAR1) b1.vel = b1.vel - cmul(lin, b1.inv_mass_vec);
        if (cd.mode & JH_SOLVE_ANGULAR1) {
            V3<R> rp = c.rel1;
            const Sym<R>& I = b1.inv_inertia_world;
            R e1[3], e2[3], e3[3];
            for (int l = 0; l < 3; ++l) {
                R rrx = rp.y * NZ[l] - rp.z * NY[l];
                R rry = rp.z * NX[l] - rp.x * NZ[l];
                R rrz = rp.x * NY[l] - rp.y * NX[l];
                e1[l] = I.m11 * rrx + I.m12 * rry + I.m13 * rrz;
                e2[l] = I.m12 * rrx + I.m22 * rry + I.m23 * rrz;
                e3[l] = I.m13 * rrx + I.m23 * rry + I.m33 * rrz;
            }
            V3<R> ai{sum3(ws, e1[0], e1[1], e1[2]), sum3(ws, e2[0], e2[1], e2[2]), sum3(ws, e3[0], e3[1], e3[2])};
            b1.angvel = b1.angvel - ai;
            for (int l = 0; l < 3; ++l) {
                ktx[l] = e2[l] * rp.z - e3[l] * rp.y;
                kty[l] = e3[l] * rp.x - e1[l] * rp.z;
                ktz[l] = e1[l] * rp.y - e2[l] * rp.x;
            }
        }
        if (cd.mode & JH_SOLVE_LINEAR2) b2.vel = b2.vel + cmul(lin, b2.inv_mass_vec);
        if (cd.mode & JH_SOLVE_ANGULAR2) {
            V3<R> rp = c.rel2;
            const Sym<R>& I = b2.inv_inertia_world;
            R f1[3], f2[3], f3[3];
            for (int l = 0; l < 3; ++l) {
                R rrx = rp.y * NZ[l] - rp.z * NY[l];
                R rry = rp.z * NX[l] - rp.x * NZ[l];
                R rrz = rp.x * NY[l] - rp.y * NX[l];
                f1[l] = I.m11 * rrx + I.m12 * rry + I.m13 * rrz;
                f2[l] = I.m12 * rrx + I.m22 * rry + I.m23 * rrz;
                f3[l] = I.m13 * rrx + I.m23 * rry + I.m33 * rrz;
            }
            V3<R> ai{sum3(ws, f1[0], f1[1], f1[2]), sum3(ws, f2[0], f2[1], f2[2]), sum3(ws, f3[0], f3[1], f3[2])};
            b2.angvel = b2.angvel + ai;
            for (int l = 0; l < 3; ++l) {
                ktx[l] = ktx[l] + (f2[l] * rp.z - f3[l] * rp.y);
                kty[l] = kty[l] + (f3[l] * rp.x - f1[l] * rp.z);
                ktz[l] = ktz[l] + (f1[l] * rp.y - f2[l] * rp.x);
            }
        }
        for (int l = 0; l < 3; ++l) {
            R kres = NX[l] * ktx[l] + NY[l] * kty[l] + NZ[l] * ktz[l];
            R kk = k[l] + kres;
            c.mass[l] = kk > R(0.0) ? R(1.0) / kk : R(0.0);
        }
        c.penalty = Contact<R>::BiasFactor * ts.inv_dt * rmax(R(0.0), pen - Contact<R>::AllowedPenetration);
        c.penalty = rmin(c.penalty, Contact<R>::MaximumBias);
    }

    // Contact.IterateAccelerated (Contact.cs:1183-1248).
    void iterate_contact(Arbiter<R>& cd, Contact<R>& c, bool apply_bias) {
        Body<R>& b1 = bodies[cd.b1];
        Body<R>& b2 = bodies[cd.b2];
        V3<R> dv = b2.vel - b1.vel;
        if (cd.mode & (JH_SOLVE_ANGULAR1 | JH_SOLVE_ANGULAR2)) {
            dv = b2.vel + cross(b2.angvel, c.rel2);
            dv = dv - (b1.vel + cross(b1.angvel, c.rel1));
        }
        R bias = c.bias;
        if (apply_bias && c.penalty > 0) bias = rmax(c.penalty, c.bias);
        R NX[3] = {c.n.x, c.t1.x, c.t2.x}, NY[3] = {c.n.y, c.t1.y, c.t2.y}, NZ[3] = {c.n.z, c.t1.z, c.t2.z};
        R imp[3], old[3] = {c.acc[0], c.acc[1], c.acc[2]};
        R bv[3] = {bias, R(0), R(0)};
        for (int l = 0; l < 3; ++l) {
            R vd = NX[l] * dv.x + NY[l] * dv.y + NZ[l] * dv.z;
            imp[l] = c.mass[l] * (bv[l] - vd);
        }
        R maxT = cd.friction * c.acc[0];
        R lo[3] = {R(0), -maxT, -maxT}, hi[3] = {std::numeric_limits<R>::max(), maxT, maxT};
        for (int l = 0; l < 3; ++l) {
            R a = old[l] + imp[l];
            a = rmin(rmax(a, lo[l]), hi[l]);
            c.acc[l] = a;
            imp[l] = a - old[l];
        }
        V3<R> lin{sum3(imp, NX[0], NX[1], NX[2]), sum3(imp, NY[0], NY[1], NY[2]), sum3(imp, NZ[0], NZ[1], NZ[2])};
        if (cd.mode & JH_SOLVE_LINEAR1) b1.vel = b1.vel - cmul(lin, b1.inv_mass_vec);
        if (cd.mode & JH_SOLVE_ANGULAR1) b1.angvel = b1.angvel - sym_mul_contact(cross(c.rel1, lin), b1.inv_inertia_world);
        if (cd.mode & JH_SOLVE_LINEAR2) b2.vel = b2.vel + cmul(lin, b2.inv_mass_vec);
        if (cd.mode & JH_SOLVE_ANGULAR2) b2.angvel = b2.angvel + sym_mul_contact(cross(c.rel2, lin), b2.inv_inertia_world);
    }

    // Contact.UpdatePosition (Contact.cs:762-797).
    bool contact_update_position(Arbiter<R>& cd, Contact<R>& c) {
        const Body<R>& b1 = bodies[cd.b1];
        const Body<R>& b2 = bodies[cd.b2];
        c.rel1 = rotate(c.pos1, b1.q);
        V3<R> p1 = c.rel1 + b1.pos;
        c.rel2 = rotate(c.pos2, b2.q);
        V3<R> p2 = c.rel2 + b2.pos;
        V3<R> dist = p1 - p2;
        R pen = dot(dist, c.n);
        if (pen < -Contact<R>::BreakThreshold * R(0.1)) return false;
        dist = dist - pen * c.n;
        R tsq = len_sq(dist);
        if (tsq > Contact<R>::BreakThreshold * Contact<R>::BreakThreshold) return false;
        return true;
    }
    // ContactData.UpdatePosition (Contact.cs:238-264).
    void arbiter_update_position(Arbiter<R>& cd) {
        cd.usage &= 15u;
        cd.usage |= cd.usage << 4;
        for (int i = 0; i < 4; ++i)
            if ((cd.usage & (1u << i)) != 0 && !contact_update_position(cd, cd.c[i])) cd.usage &= ~(1u << i);
    }

    // World.Step (World.Step.cs:141-296) with a fixed solve order.
    // ---- Learned-warm-start features and residual probe --------------------------------------
    // These are definitions of this project (the reference runs fixed iteration counts and has no
    // residual); the arithmetic mirrors the device code operation for operation.
    struct Probe {
        int iterations = 0, modes = 0;
        long capacity = 0;
        double* residual = nullptr;  // [3][iterations + 1]
        float* features = nullptr;   // [capacity][24]
        float* targets = nullptr;    // [capacity][3]
        const float* learned = nullptr;  // [capacity][3]: the warm start of probe 1 (an external
                                         // network's impulses, replayed), rows in trace order
        uint32_t* flags = nullptr;       // [capacity]: the contact flags of each row at the warm start
        int32_t* row_bodies = nullptr;   // [capacity][2]: body1, body2 of each row
        long rows = 0, points = 0;
    };

    // 24 features of slot s at the warm start of the first substep (velocities after IntegrateForces).
    void contact_features(const Arbiter<R>& a, int s, float* f) const {
        const Contact<R>& c = a.c[s];
        const Body<R>& b1 = bodies[a.b1];
        const Body<R>& b2 = bodies[a.b2];
        V3<R> v1 = b1.vel, w1 = b1.angvel, v2 = b2.vel, w2 = b2.angvel;
        if (b1.motion == JH_MOTION_DYNAMIC) { w1 = w1 + b1.dangvel; v1 = v1 + b1.dvel; }
        if (b2.motion == JH_MOTION_DYNAMIC) { w2 = w2 + b2.dangvel; v2 = v2 + b2.dvel; }
        V3<R> dv = v2 + cross(w2, c.rel2);
        dv = dv - (v1 + cross(w1, c.rel1));
        V3<R> r1 = rotate(c.pos1, b1.q) + b1.pos;
        V3<R> r2 = rotate(c.pos2, b2.q) + b2.pos;
        R v[24] = {c.n.x, c.n.y, c.n.z, c.t1.x, c.t1.y, c.t1.z, c.t2.x, c.t2.y, c.t2.z,
                   dot(dv, c.n), dot(dv, c.t1), dot(dv, c.t2), dot(r1 - r2, c.n),
                   c.mass[0], c.mass[1], c.mass[2], c.acc[0], c.acc[1], c.acc[2], R(0), a.friction,
                   c.bias, c.penalty, (c.flag & Contact<R>::NewContact) ? R(1) : R(0)};
        for (int k = 0; k < 24; ++k) f[k] = float(v[k]);
    }

    // Sum over contact points of e_n^2 + e_t1^2 + e_t2^2 (see the device definition).
    // lam_old: normal impulses before the last iteration, per arbiter (4 slots), or empty.
    double residual(const std::vector<Arbiter<R>>& pa, const std::vector<R>& lam_old) const {
        double sum = 0.0;
        for (size_t ai = 0; ai < pa.size(); ++ai) {
            const Arbiter<R>& a = pa[ai];
            const Body<R>& b1 = bodies[a.b1];
            const Body<R>& b2 = bodies[a.b2];
            for (int k = 0; k < 4; ++k) {
                if (!(a.usage & (1u << k))) continue;
                const Contact<R>& c = a.c[k];
                V3<R> dv = b2.vel + cross(b2.angvel, c.rel2);
                dv = dv - (b1.vel + cross(b1.angvel, c.rel1));
                R vn = c.n.x * dv.x + c.n.y * dv.y + c.n.z * dv.z;
                R vt1 = c.t1.x * dv.x + c.t1.y * dv.y + c.t1.z * dv.z;
                R vt2 = c.t2.x * dv.x + c.t2.y * dv.y + c.t2.z * dv.z;
                R bias = c.penalty > R(0) ? std::max(c.penalty, c.bias) : c.bias;
                R ln = c.acc[0];
                R en = ln > R(0) ? bias - vn : std::max(R(0), bias - vn);
                R lim = a.friction * (lam_old.empty() ? ln : lam_old[4 * ai + k]);
                R et1 = std::fabs(c.acc[1]) < lim ? std::fabs(vt1) : R(0);
                R et2 = std::fabs(c.acc[2]) < lim ? std::fabs(vt2) : R(0);
                sum += double(en) * double(en) + double(et1) * double(et1) + double(et2) * double(et2);
            }
        }
        return sum;
    }

    // Probes of the first substep on copies (modes: bit 0 previous-frame warm start, bit 2 cold).
    void run_probe(const std::vector<Arbiter<R>*>& list, const StepTiming<R>& ts, Probe& pr) {
        pr.rows = pr.points = 0;
        for (auto& kv : arbiters)
            for (int s = 0; s < 4; ++s) {
                if (!(kv.second.usage & (1u << s)) || !arbiter_active(kv.second)) continue;
                if (pr.features && pr.rows < pr.capacity) contact_features(kv.second, s, pr.features + 24 * pr.rows);
                if (pr.flags && pr.rows < pr.capacity) pr.flags[pr.rows] = kv.second.c[s].flag;
                if (pr.row_bodies && pr.rows < pr.capacity) {
                    pr.row_bodies[2 * pr.rows] = int32_t(kv.second.b1);
                    pr.row_bodies[2 * pr.rows + 1] = int32_t(kv.second.b2);
                }
                if (pr.rows < pr.capacity) ++pr.rows;
                ++pr.points;
            }
        int K = pr.iterations;
        for (int m : {0, 1, 2}) {
            if (!(pr.modes & (1 << m))) continue;
            if (m == 1 && !pr.learned) continue;
            std::vector<Body<R>> saved = bodies;
            std::vector<Arbiter<R>> pa;
            pa.reserve(list.size());
            for (auto* a : list) pa.push_back(*a);
            if (m == 2)
                for (auto& a : pa)
                    for (auto& c : a.c) c.acc[0] = c.acc[1] = c.acc[2] = R(0);
            if (m == 1) {  // the given warm start, row by row (used slots in key order)
                std::map<std::pair<uint32_t, uint32_t>, Arbiter<R>*> by_key;
                for (auto& a : pa) by_key[{a.key1, a.key2}] = &a;
                long row = 0;
                for (auto& kv : arbiters)
                    for (int s = 0; s < 4; ++s) {
                        if (!(kv.second.usage & (1u << s)) || !arbiter_active(kv.second)) continue;
                        auto it = by_key.find(kv.first);
                        if (row < pr.capacity && it != by_key.end())
                            for (int r = 0; r < 3; ++r) it->second->c[s].acc[r] = R(pr.learned[3 * row + r]);
                        ++row;
                    }
            }
            for (auto& b : bodies) {
                if (b.motion != JH_MOTION_DYNAMIC || !b.active) continue;
                b.angvel = b.angvel + b.dangvel;
                b.vel = b.vel + b.dvel;
            }
            for (auto& a : pa)
                for (int i = 0; i < 4; ++i)
                    if (a.usage & (1u << i)) prepare_contact(a, a.c[i], ts);
            pr.residual[m * (K + 1)] = residual(pa, {});
            std::vector<R> lam(4 * pa.size());
            for (int k = 1; k <= K; ++k) {
                for (size_t ai = 0; ai < pa.size(); ++ai)
                    for (int i = 0; i < 4; ++i) lam[4 * ai + i] = pa[ai].c[i].acc[0];
                for (auto& a : pa)
                    for (int i = 0; i < 4; ++i)
                        if (a.usage & (1u << i)) iterate_contact(a, a.c[i], true);
                pr.residual[m * (K + 1) + k] = residual(pa, lam);
            }
            if (m == 0 && pr.targets) {
                std::map<std::pair<uint32_t, uint32_t>, const Arbiter<R>*> by_key;
                for (auto& a : pa) by_key[{a.key1, a.key2}] = &a;
                long row = 0;
                for (auto& kv : arbiters)
                    for (int s = 0; s < 4; ++s) {
                        if (!(kv.second.usage & (1u << s)) || !arbiter_active(kv.second)) continue;
                        if (row < pr.capacity) {
                            const Contact<R>& c = by_key.at(kv.first)->c[s];
                            for (int r = 0; r < 3; ++r) pr.targets[3 * row + r] = float(c.acc[r]);
                        }
                        ++row;
                    }
            }
            bodies = saved;
        }
    }

    void step(R dt, const std::vector<std::pair<uint32_t, uint32_t>>* order) { step_impl(dt, order, nullptr); }

    // World.StabilizeInternal (World.Step.cs:330-392), SolveMode.Regular, without CheckDeactivation
    // (no sleeping here): per substep SolveVelocities with `iters` iterations (contacts, then
    // constraints) and RelaxVelocities with `relax` iterations, on the existing contacts and
    // constraints; no detection, no IntegrateForces / IntegrateVelocities, no body update.  The order
    // is given (a replay) or the key order.
    void stabilize(R dt, int iters, int relax, const std::vector<std::pair<uint32_t, uint32_t>>* order) {
        if (dt < std::numeric_limits<R>::denorm_min()) return;
        StepTiming<R> ts(dt, substeps);
        check_deactivation();
        std::vector<Arbiter<R>*> list;
        if (order) {
            for (auto& k : *order) {
                auto it = arbiters.find(k);
                if (it != arbiters.end() && arbiter_active(it->second)) list.push_back(&it->second);
            }
        } else {
            for (auto& kv : arbiters)
                if (arbiter_active(kv.second)) list.push_back(&kv.second);
        }
        for (int s = 0; s < ts.substeps; ++s) {
            for (auto* a : list)
                for (int i = 0; i < 4; ++i)
                    if (a->usage & (1u << i)) prepare_contact(*a, a->c[i], ts);
            for_constraints([&](Constraint<R>& c) {
                ConstraintBody<R> A = cbody(c.b1), B = cbody(c.b2);
                constraint_prepare(c, A, B, ts.inv_dt, ts.inv_substep_dt, ts.substep_dt);
            });
            for (int it = 0; it < iters; ++it) {
                for (auto* a : list)
                    for (int i = 0; i < 4; ++i)
                        if (a->usage & (1u << i)) iterate_contact(*a, a->c[i], true);
                for_constraints([&](Constraint<R>& c) {
                    ConstraintBody<R> A = cbody(c.b1), B = cbody(c.b2);
                    constraint_iterate(c, A, B, ts.inv_substep_dt);
                });
            }
            for (int it = 0; it < relax; ++it)
                for (auto* a : list)
                    for (int i = 0; i < 4; ++i)
                        if (a->usage & (1u << i)) iterate_contact(*a, a->c[i], false);
        }
    }

    void step_impl(R dt, const std::vector<std::pair<uint32_t, uint32_t>>* order, Probe* probe) {
        if (!(dt >= R(0)) || !std::isfinite(dt)) return;
        if (dt < std::numeric_limits<R>::denorm_min()) return;
        StepTiming<R> ts(dt, substeps);
        ev_begin.clear();
        ev_end.clear();
        // Narrowphase over the broadphase pair set.
        last_manifolds.clear();
        epa_calls = 0;
        last_pairs = compute_pairs();
        deferred.clear();
        detect_all(last_pairs, ts.dt);
        // HandleDeferredArbiters (World.Step.cs:881-911): ArbiterCreated, AddToActiveList of both islands
        for (auto& k : deferred) {
            const Arbiter<R>& a = arbiters.at(k);
            add_connection(a.b1, a.b2);
            add_to_active(bodies[a.b1].island);
            add_to_active(bodies[a.b2].island);
        }
        deferred.clear();
        check_deactivation();
        // Solve order: given, else key order.
        std::vector<Arbiter<R>*> list;
        if (order) {
            for (auto& k : *order) {
                auto it = arbiters.find(k);
                if (it != arbiters.end() && arbiter_active(it->second)) list.push_back(&it->second);
            }
        } else {
            for (auto& kv : arbiters)
                if (arbiter_active(kv.second)) list.push_back(&kv.second);
        }
        if (probe) run_probe(list, ts, *probe);
        for (int s = 0; s < ts.substeps; ++s) {
            // IntegrateForces (World.Step.cs:1060-1072)
#pragma omp parallel for schedule(static) if (bodies.size() > 4096)
            for (long i = 0; i < long(bodies.size()); ++i) {
                Body<R>& b = bodies[i];
                if (b.motion != JH_MOTION_DYNAMIC || !b.active) continue;
                b.angvel = b.angvel + b.dangvel;
                b.vel = b.vel + b.dvel;
            }
            // SolveVelocities (World.Step.cs:1155-1191): prepare, then iterations
            for (auto* a : list)
                for (int i = 0; i < 4; ++i)
                    if (a->usage & (1u << i)) prepare_contact(*a, a->c[i], ts);
            for_constraints([&](Constraint<R>& c) {  // PrepareConstraints (World.Step.cs:621-668)
                ConstraintBody<R> A = cbody(c.b1), B = cbody(c.b2);
                constraint_prepare(c, A, B, ts.inv_dt, ts.inv_substep_dt, ts.substep_dt);
            });
            for (int it = 0; it < iterations; ++it) {
                for (auto* a : list)
                    for (int i = 0; i < 4; ++i)
                        if (a->usage & (1u << i)) iterate_contact(*a, a->c[i], true);
                for_constraints([&](Constraint<R>& c) {  // IterateConstraints (World.Step.cs:670-716)
                    ConstraintBody<R> A = cbody(c.b1), B = cbody(c.b2);
                    constraint_iterate(c, A, B, ts.inv_substep_dt);
                });
            }
            // IntegrateVelocities (World.Step.cs:1106-1133), no gyroscopic term
#pragma omp parallel for schedule(static) if (bodies.size() > 4096)
            for (long i = 0; i < long(bodies.size()); ++i) {
                Body<R>& b = bodies[i];
                if (b.motion == JH_MOTION_STATIC || !b.active) continue;
                b.pos = b.pos + b.vel * ts.substep_dt;
                Q4<R> rq = rotation_quaternion(b.angvel, ts.substep_dt);
                b.q = qnormalize(qmul(rq, b.q));
            }
            // RelaxVelocities (World.Step.cs:1135-1153)
            for (int it = 0; it < relaxations; ++it)
                for (auto* a : list)
                    for (int i = 0; i < 4; ++i)
                        if (a->usage & (1u << i)) iterate_contact(*a, a->c[i], false);
        }
        // RemoveBrokenArbiters (World.Step.cs:839-860): the arbiters broken by the previous step's
        // UpdateContacts that this step's Detect did not refill are freed (EndCollide).
        for (auto& k : broken) {
            auto it = arbiters.find(k);
            if (it == arbiters.end() || (it->second.usage & 15u) != 0) continue;
            int b1 = it->second.b1, b2 = it->second.b2;
            add_to_active(bodies[b1].island);
            add_to_active(bodies[b2].island);
            remove_connection(b1, b2);  // IslandHelper.ArbiterRemoved
            arbiters.erase(it);
            ev_end.push_back(k);
        }
        broken.clear();
        // UpdateContacts (World.Step.cs:862-878): arbiters left without intact contacts are queued for
        // removal; until then they carry no solver work.
        {
            std::vector<Arbiter<R>*> all;  // memContacts.Active
            all.reserve(arbiters.size());
            for (auto& kv : arbiters)
                if (arbiter_active(kv.second)) all.push_back(&kv.second);
#pragma omp parallel for schedule(static) if (all.size() > 4096)
            for (long i = 0; i < long(all.size()); ++i) arbiter_update_position(*all[i]);
            for (auto* a : all)
                if ((a->usage & 15u) == 0) broken.push_back({a->key1, a->key2});
        }
        std::sort(ev_begin.begin(), ev_begin.end());
        // ForeachActiveBody -> RigidBody.Update (RigidBody.cs:716-764): the sleep timer of every active
        // body (its island stays awake while the timer is short or a kinematic body moves), then the
        // dynamic bodies' velocity update
        std::vector<char> keep_awake(bodies.size(), 0);
#pragma omp parallel for schedule(static) if (bodies.size() > 4096)
        for (long i = 0; i < long(bodies.size()); ++i) {
            Body<R>& b = bodies[i];
            if (!b.active) continue;
            restrict_linear(b.vel, b.locks);  // RigidBody.cs:719-722
            restrict_angular(b.angvel, b.locks);
            bool kin = b.motion == JH_MOTION_KINEMATIC;
            auto zero = [](V3<R> v) { return v.x == R(0) && v.y == R(0) && v.z == R(0); };  // JVector ==
            bool resting = kin ? (zero(b.angvel) && zero(b.vel))
                               : (len_sq(b.angvel) < b.sleep_ang_sq && len_sq(b.vel) < b.sleep_lin_sq);
            b.sleep_time = resting ? b.sleep_time + ts.dt : R(0);
            if ((kin && !resting) || b.sleep_time < b.deact_time) keep_awake[i] = 1;
            if (b.motion != JH_MOTION_DYNAMIC) continue;
            b.update_world_inertia();
            b.angvel = b.angvel * b.ang_damp;
            b.vel = b.vel * b.lin_damp;
            b.dvel = cmul(b.force, b.inv_mass_vec) * ts.substep_dt;
            b.dangvel = sym_mul(b.torque, b.inv_inertia_world) * ts.substep_dt;
            if (b.gravity) {
                V3<R> g = gravity;
                restrict_linear(g, b.locks);
                b.dvel = b.dvel + g * ts.substep_dt;
            }
            b.force = {R(0), R(0), R(0)};
            b.torque = {R(0), R(0), R(0)};
        }
        for (size_t i = 0; i < bodies.size(); ++i)
            if (keep_awake[i]) isl[bodies[i].island].marked = true;  // InternalIsland.MarkedAsActive
        // DynamicTree.Update(multiThread, StepDt) (World.Step.cs:274-276): the next broadphase sweeps
        // speculative bodies' boxes over this step's length.
        broadphase_dt = ts.dt;
    }
};

}  // namespace orc

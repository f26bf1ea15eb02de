// ORACLE — test infrastructure only.  C entry points of the CPU restatement of Jitter2's step
// (liborc.so), called from tests/ through ctypes, from __graft_entry__.smoke() as the checker,
// and from bench.py's cpu_baseline leg.  fp64 != 0 selects the double instantiation
// (≙ Jitter2.Double).  Never part of the product path.
#include <chrono>
#include <cstdio>
#include <memory>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "orc_world.hpp"

using namespace orc;

namespace {
struct Handle {
    bool fp64;
    World<float>* wf = nullptr;
    World<double>* wd = nullptr;
    std::vector<float> verts;
};

template <class R> void set_config(World<R>& w, const jh_config* c) {
    w.gravity = {R(c->gravity[0]), R(c->gravity[1]), R(c->gravity[2])};
    w.substeps = c->substep_count;
    w.iterations = c->solver_iterations;
    w.relaxations = c->relaxation_iterations;
    w.specrel = R(c->speculative_relaxation_factor);
    w.spec_threshold = R(c->speculative_velocity_threshold);
    w.aux_contacts = c->enable_auxiliary_contact_points != 0;
    w.persistent = c->persistent_contact_manifold != 0;
    w.triangle_edge_filter = c->narrowphase_filter == JH_FILTER_TRIANGLE_EDGE;
    w.allow_deactivation = c->allow_deactivation != 0;
}

template <class R> V3<R> vin(const float* f) { return {R(f[0]), R(f[1]), R(f[2])}; }
template <class R> void vout(V3<R> v, float* f) { f[0] = float(v.x); f[1] = float(v.y); f[2] = float(v.z); }

template <class R> void body_in(Body<R>& o, const jh_body& r) {
    o.pos = vin<R>(r.position);
    o.q = {R(r.orientation[0]), R(r.orientation[1]), R(r.orientation[2]), R(r.orientation[3])};
    o.vel = vin<R>(r.velocity);
    o.angvel = vin<R>(r.angular_velocity);
    o.dvel = vin<R>(r.delta_velocity);
    o.dangvel = vin<R>(r.delta_angular_velocity);
    const float* I = r.inverse_inertia;
    o.inv_inertia = {R(I[0]), R(I[1]), R(I[2]), R(I[3]), R(I[4]), R(I[5])};
    o.inv_mass = R(r.inverse_mass);
    o.friction = R(r.friction);
    o.restitution = R(r.restitution);
    o.lin_damp = R(r.linear_damping_multiplier);
    o.ang_damp = R(r.angular_damping_multiplier);
    o.force = vin<R>(r.force);
    o.torque = vin<R>(r.torque);
    o.motion = r.motion_type;
    o.gravity = r.affected_by_gravity != 0;
    o.world = r.world_id;
    o.speculative = r.enable_speculative_contacts != 0;
    o.sleep_lin_sq = R(r.deactivation_threshold_linear_sq);
    o.sleep_ang_sq = R(r.deactivation_threshold_angular_sq);
    o.deact_time = R(r.deactivation_time);
    o.locks = r.locked_axes & 63;
    // the velocity setters (and the AllowedMotion setter) clear locked components (RigidBody.cs:584-588)
    restrict_linear(o.vel, o.locks);
    restrict_linear(o.dvel, o.locks);
    restrict_angular(o.angvel, o.locks);
    restrict_angular(o.dangvel, o.locks);
    o.update_world_inertia();
}
// Replacing the records is every property setter at once: a new body count or a motion-type change
// rebuilds the islands with every body awake (World.CreateRigidBody); otherwise each body is
// ActivateBodyNextStep(body, wakeUpStatic = true) as the Position setter does (RigidBody.cs:593).
template <class R> void set_bodies(World<R>& w, const jh_body* b, int n) {
    bool rebuild = (int)w.bodies.size() != n;
    for (int i = 0; i < n && !rebuild; ++i) rebuild = w.bodies[i].motion != b[i].motion_type;
    std::vector<int> relocked;
    for (int i = 0; i < n && i < (int)w.bodies.size(); ++i)
        if (w.bodies[i].locks != (b[i].locked_axes & 63)) relocked.push_back(i);
    w.bodies.resize(n);
    for (int i = 0; i < n; ++i) body_in(w.bodies[i], b[i]);
    for (int i : relocked) w.locks_changed(i);
    if (rebuild) w.islands_reset();
    else
        for (int i = 0; i < n; ++i) w.activate_body_next_step(i, true);
}
// The listed bodies' records replaced, the same rules per listed body; -1 for an index out of range.
template <class R> int update_bodies(World<R>& w, const int* idx, const jh_body* b, int count) {
    for (int k = 0; k < count; ++k)
        if (idx[k] < 0 || idx[k] >= (int)w.bodies.size()) return -1;
    bool rebuild = false;
    for (int k = 0; k < count; ++k) {
        rebuild = rebuild || w.bodies[idx[k]].motion != b[k].motion_type;
        bool relock = w.bodies[idx[k]].locks != (b[k].locked_axes & 63);
        body_in(w.bodies[idx[k]], b[k]);
        if (relock) w.locks_changed(idx[k]);
    }
    if (rebuild) w.islands_reset();
    else
        for (int k = 0; k < count; ++k) w.activate_body_next_step(idx[k], true);
    return 0;
}

template <class R> void get_bodies(const World<R>& w, jh_body* b, int n) {
    for (int i = 0; i < n && i < (int)w.bodies.size(); ++i) {
        const Body<R>& o = w.bodies[i];
        vout(o.pos, b[i].position);
        b[i].orientation[0] = float(o.q.x);
        b[i].orientation[1] = float(o.q.y);
        b[i].orientation[2] = float(o.q.z);
        b[i].orientation[3] = float(o.q.w);
        vout(o.vel, b[i].velocity);
        vout(o.angvel, b[i].angular_velocity);
        vout(o.dvel, b[i].delta_velocity);
        vout(o.dangvel, b[i].delta_angular_velocity);
        vout(o.force, b[i].force);
        vout(o.torque, b[i].torque);
    }
}

template <class R> void get_state64(const World<R>& w, double* out, int n) {
    for (int i = 0; i < n && i < (int)w.bodies.size(); ++i) {
        const Body<R>& o = w.bodies[i];
        double* d = out + 13 * i;
        d[0] = o.pos.x; d[1] = o.pos.y; d[2] = o.pos.z;
        d[3] = o.q.x; d[4] = o.q.y; d[5] = o.q.z; d[6] = o.q.w;
        d[7] = o.vel.x; d[8] = o.vel.y; d[9] = o.vel.z;
        d[10] = o.angvel.x; d[11] = o.angvel.y; d[12] = o.angvel.z;
    }
}

template <class R> void set_shapes(World<R>& w, const jh_shape* s, int n, const float* verts) {
    w.shapes.resize(n);
    for (int i = 0; i < n; ++i) {
        Shape<R>& o = w.shapes[i];
        o.type = s[i].type;
        o.body = s[i].body;
        o.id = s[i].shape_id;
        for (int k = 0; k < 12; ++k) o.p[k] = R(s[i].params[k]);
        o.verts = verts ? verts + 3 * s[i].vertex_offset : nullptr;
        o.nverts = s[i].vertex_count;
    }
    w.arbiters.clear();
    w.broken.clear();
    w.islands_reset();
}

template <class R> void contact_out(const Contact<R>& c, jh_contact& o) {
    o.flag = c.flag;
    o.bias = float(c.bias);
    o.penalty_bias = float(c.penalty);
    vout(c.n, o.normal);
    vout(c.t1, o.tangent1);
    vout(c.t2, o.tangent2);
    for (int k = 0; k < 3; ++k) { o.mass_normal_tangent[k] = float(c.mass[k]); o.accumulated[k] = float(c.acc[k]); }
    vout(c.pos1, o.position1);
    vout(c.pos2, o.position2);
    vout(c.rel1, o.relative_position1);
    vout(c.rel2, o.relative_position2);
}
template <class R> void contact_in(Contact<R>& c, const jh_contact& o) {
    c.flag = o.flag;
    c.bias = R(o.bias);
    c.penalty = R(o.penalty_bias);
    c.n = vin<R>(o.normal);
    c.t1 = vin<R>(o.tangent1);
    c.t2 = vin<R>(o.tangent2);
    for (int k = 0; k < 3; ++k) { c.mass[k] = R(o.mass_normal_tangent[k]); c.acc[k] = R(o.accumulated[k]); }
    c.pos1 = vin<R>(o.position1);
    c.pos2 = vin<R>(o.position2);
    c.rel1 = vin<R>(o.relative_position1);
    c.rel2 = vin<R>(o.relative_position2);
}

// Arbiters with intact contacts; those waiting for RemoveBrokenArbiters are listed only with all.
template <class R> int get_arbiters(const World<R>& w, jh_arbiter* out, int cap, bool all = false) {
    int n = 0;
    for (auto& kv : w.arbiters) {
        if (!all && (kv.second.usage & 15u) == 0) continue;
        if (n < cap && out) {
            const Arbiter<R>& a = kv.second;
            jh_arbiter& o = out[n];
            o.shape1 = a.key1;
            o.shape2 = a.key2;
            o.body1 = a.b1;
            o.body2 = a.b2;
            o.usage_mask = a.usage;
            o.mode = a.mode;
            o.friction = float(a.friction);
            o.restitution = float(a.restitution);
            o.speculative_relaxation_factor = float(a.specrel);
            o.reserved = 0;
            for (int i = 0; i < 4; ++i) contact_out(a.c[i], o.contacts[i]);
        }
        n++;
    }
    return n;
}
template <class R> void set_arbiters(World<R>& w, const jh_arbiter* in, int n) {
    w.arbiters.clear();
    w.broken.clear();
    for (int i = 0; i < n; ++i) {
        Arbiter<R> a;
        a.key1 = in[i].shape1;
        a.key2 = in[i].shape2;
        a.b1 = in[i].body1;
        a.b2 = in[i].body2;
        a.usage = in[i].usage_mask;
        a.mode = in[i].mode;
        a.friction = R(in[i].friction);
        a.restitution = R(in[i].restitution);
        a.specrel = R(in[i].speculative_relaxation_factor);
        for (int k = 0; k < 4; ++k) contact_in(a.c[k], in[i].contacts[k]);
        w.arbiters[std::make_pair(a.key1, a.key2)] = a;
    }
    w.islands_reset();
}

template <class R> int get_manifolds(const World<R>& w, jh_manifold* out, int cap) {
    int n = 0;
    for (auto& m : w.last_manifolds) {
        if (n < cap && out) {
            jh_manifold& o = out[n];
            o.shape1 = m.k1;
            o.shape2 = m.k2;
            o.count = m.count;
            o.flags = (m.epa ? 1 : 0) | (m.spec ? 2 : 0);
            vout(m.normal, o.normal);
            o.penetration = float(m.pen);
            vout(m.pa, o.point_a);
            vout(m.pb, o.point_b);
            for (int i = 0; i < 4; ++i) {
                if (i < m.count) { vout(m.mA[i], o.manifold_a[i]); vout(m.mB[i], o.manifold_b[i]); }
                else { for (int k = 0; k < 3; ++k) { o.manifold_a[i][k] = 0; o.manifold_b[i][k] = 0; } }
            }
        }
        n++;
    }
    return n;
}

template <class R> void set_constraints(World<R>& w, const jh_constraint* in, int n) {
    w.constraints.assign(n, Constraint<R>());
    for (int i = 0; i < n; ++i) {
        Constraint<R>& c = w.constraints[i];
        c.type = in[i].type;
        c.b1 = in[i].body1;
        c.b2 = in[i].body2;
        c.enabled = in[i].enabled;
        for (int k = 0; k < 16; ++k) c.p[k] = R(in[i].params[k]);
        for (int k = 0; k < 8; ++k) c.s[k] = R(in[i].state[k]);
    }
    w.con_order.clear();
    w.islands_reset();
}
template <class R> int get_constraints(const World<R>& w, jh_constraint* out, int cap) {
    int n = (int)w.constraints.size();
    for (int i = 0; i < n && i < cap && out; ++i) {
        const Constraint<R>& c = w.constraints[i];
        out[i].type = c.type;
        out[i].body1 = c.b1;
        out[i].body2 = c.b2;
        out[i].enabled = c.enabled;
        for (int k = 0; k < 16; ++k) out[i].params[k] = float(c.p[k]);
        for (int k = 0; k < 8; ++k) out[i].state[k] = float(c.s[k]);
    }
    return n;
}

Shape<float> shape_f(const jh_shape& s, const float* verts) {
    Shape<float> o;
    o.type = s.type;
    o.body = s.body;
    for (int k = 0; k < 12; ++k) o.p[k] = s.params[k];
    o.verts = verts ? verts + 3 * s.vertex_offset : nullptr;
    o.nverts = s.vertex_count;
    return o;
}
Shape<double> shape_d(const jh_shape& s, const float* verts) {
    Shape<double> o;
    o.type = s.type;
    o.body = s.body;
    for (int k = 0; k < 12; ++k) o.p[k] = s.params[k];
    o.verts = verts ? verts + 3 * s.vertex_offset : nullptr;
    o.nverts = s.vertex_count;
    return o;
}

// Contact-level restatement for known-answer tests: bodies[0..1] and one arbiter (body1 = 0,
// body2 = 1).  op 0: AddContact(p1, p2, n) after ResetMode; op 1: PrepareForIteration; op 2:
// Iterate(applyBias = true); op 3: Iterate(applyBias = false); op 4: UpdatePosition.  The bodies and
// the arbiter are updated in place (FP32 or FP64 evaluation, stored back as float).
template <class R> int contact_op_t(jh_body* bodies, jh_arbiter* arb, int op, float dt, int substeps,
                                     const float* p1, const float* p2, const float* n) {
    World<R> w;
    set_bodies(w, bodies, 2);
    jh_arbiter a0 = *arb;
    a0.shape1 = 0;
    a0.shape2 = 1;
    set_arbiters(w, &a0, 1);
    Arbiter<R>& a = w.arbiters.begin()->second;
    a.b1 = arb->body1;
    a.b2 = arb->body2;
    StepTiming<R> ts(R(dt), substeps);
    switch (op) {
        case 0:
            w.reset_mode(a);
            w.add_contact(a, vin<R>(p1), vin<R>(p2), vin<R>(n));
            break;
        case 1:
            for (int i = 0; i < 4; ++i)
                if (a.usage & (1u << i)) w.prepare_contact(a, a.c[i], ts);
            break;
        case 2:
        case 3:
            for (int i = 0; i < 4; ++i)
                if (a.usage & (1u << i)) w.iterate_contact(a, a.c[i], op == 2);
            break;
        case 4:
            w.arbiter_update_position(a);
            break;
        default:
            return -1;
    }
    get_bodies(w, bodies, 2);
    jh_arbiter out;
    get_arbiters(w, &out, 1, true);
    out.body1 = arb->body1;
    out.body2 = arb->body2;
    out.shape1 = arb->shape1;
    out.shape2 = arb->shape2;
    *arb = out;
    return 0;
}
}  // namespace

#define DISPATCH(h, expr) ((h)->fp64 ? ([&](World<double>& W) { return expr; })(*(h)->wd) \
                                     : ([&](World<float>& W) { return expr; })(*(h)->wf))

namespace {
template <class R>
int mpr_manifold_t(const jh_shape* a, const jh_shape* b, const float* verts, const float* qa, const float* pa,
                   const float* qb, const float* pb, float epa_threshold, float* out,
                   Shape<R> (*mk)(const jh_shape&, const float*)) {
    Shape<R> sa = mk(*a, verts), sb = mk(*b, verts);
    std::unique_ptr<Polytope<R>> poly(new Polytope<R>());
    Q4<R> QA{R(qa[0]), R(qa[1]), R(qa[2]), R(qa[3])}, QB{R(qb[0]), R(qb[1]), R(qb[2]), R(qb[3])};
    MprResult<R> r = mpr_epa(sa, sb, QA, QB, vin<R>(pa), vin<R>(pb), poly.get(), R(epa_threshold));
    vout(r.pa, out);
    vout(r.pb, out + 3);
    vout(r.normal, out + 6);
    out[9] = float(r.penetration);
    out[10] = r.epa ? 1.f : 0.f;
    out[11] = 0;
    if (r.hit) {
        Manifold<R> m;
        m.build(sa, sb, QA, QB, vin<R>(pa), vin<R>(pb), r.pa, r.pb, r.normal);
        out[11] = float(m.count);
        for (int i = 0; i < m.count && i < 4; ++i) { vout(m.mA[i], out + 12 + 3 * i); vout(m.mB[i], out + 24 + 3 * i); }
    }
    return r.hit ? 1 : 0;
}
}  // namespace

extern "C" {

void* orc_create(int fp64, const jh_config* cfg) {
    Handle* h = new Handle();
    h->fp64 = fp64 != 0;
    if (h->fp64) h->wd = new World<double>();
    else h->wf = new World<float>();
    if (cfg) DISPATCH(h, (set_config(W, cfg), 0));
    return h;
}
void orc_destroy(void* p) {
    Handle* h = (Handle*)p;
    delete h->wf;
    delete h->wd;
    delete h;
}
int orc_set_config(void* p, const jh_config* cfg) {
    Handle* h = (Handle*)p;
    DISPATCH(h, (set_config(W, cfg), 0));
    return 0;
}
int orc_set_bodies(void* p, const jh_body* b, int n) {
    Handle* h = (Handle*)p;
    DISPATCH(h, (set_bodies(W, b, n), 0));
    return 0;
}
int orc_get_bodies(void* p, jh_body* b, int n) {
    Handle* h = (Handle*)p;
    DISPATCH(h, (get_bodies(W, b, n), 0));
    return 0;
}
// 13 doubles per body: position, orientation (x,y,z,w), velocity, angular velocity.
int orc_get_state64(void* p, double* out, int n) {
    Handle* h = (Handle*)p;
    DISPATCH(h, (get_state64(W, out, n), 0));
    return 0;
}
int orc_set_shapes(void* p, const jh_shape* s, int n, const float* verts, int nverts) {
    Handle* h = (Handle*)p;
    h->verts.assign(verts ? verts : (const float*)nullptr, verts ? verts + 3 * (size_t)nverts : nullptr);
    const float* v = h->verts.empty() ? nullptr : h->verts.data();
    DISPATCH(h, (set_shapes(W, s, n, v), 0));
    return 0;
}
int orc_set_arbiters(void* p, const jh_arbiter* a, int n) {
    Handle* h = (Handle*)p;
    DISPATCH(h, (set_arbiters(W, a, n), 0));
    return 0;
}
int orc_get_arbiters(void* p, jh_arbiter* a, int cap) {
    Handle* h = (Handle*)p;
    return DISPATCH(h, get_arbiters(W, a, cap));
}
// Every held arbiter, also those whose contacts all broke and that the next step removes.
int orc_get_arbiters_all(void* p, jh_arbiter* a, int cap) {
    Handle* h = (Handle*)p;
    return DISPATCH(h, get_arbiters(W, a, cap, true));
}
int orc_get_manifolds(void* p, jh_manifold* m, int cap) {
    Handle* h = (Handle*)p;
    return DISPATCH(h, get_manifolds(W, m, cap));
}
// Pair set on the current state (sorted); brute != 0 uses the O(n^2) all-pairs scan.
int orc_compute_pairs(void* p, uint32_t* out, int cap, int brute) {
    Handle* h = (Handle*)p;
    return DISPATCH(h, ([&]() {
        auto pairs = W.compute_pairs(brute != 0);
        int n = (int)pairs.size();
        for (int i = 0; i < n && i < cap; ++i) { out[2 * i] = pairs[i].first; out[2 * i + 1] = pairs[i].second; }
        return n;
    })());
}
// Solver mass properties per body, 9 doubles: InverseMassVector, InverseInertiaWorld (M11 M12 M13 M22
// M23 M33), as RigidBody.UpdateWorldInertia left them.
int orc_get_mass(void* p, double* out, int n) {
    Handle* h = (Handle*)p;
    return DISPATCH(h, ([&]() {
        int m = std::min(n, (int)W.bodies.size());
        for (int i = 0; i < m; ++i) {
            const auto& b = W.bodies[i];
            double* d = out + 9 * i;
            d[0] = b.inv_mass_vec.x; d[1] = b.inv_mass_vec.y; d[2] = b.inv_mass_vec.z;
            d[3] = b.inv_inertia_world.m11; d[4] = b.inv_inertia_world.m12; d[5] = b.inv_inertia_world.m13;
            d[6] = b.inv_inertia_world.m22; d[7] = b.inv_inertia_world.m23; d[8] = b.inv_inertia_world.m33;
        }
        return m;
    })());
}
// Activation state per body: active[i] = RigidBody.IsActive, sleep[i] = InternalSleepTime.
int orc_get_activation(void* p, uint8_t* active, float* sleep, int n) {
    Handle* h = (Handle*)p;
    return DISPATCH(h, ([&]() {
        int m = std::min(n, (int)W.bodies.size());
        for (int i = 0; i < m; ++i) {
            if (active) active[i] = W.bodies[i].active ? 1 : 0;
            if (sleep) sleep[i] = float(W.bodies[i].sleep_time);
        }
        return m;
    })());
}
int orc_update_bodies(void* p, const int* idx, const jh_body* b, int count) {
    Handle* h = (Handle*)p;
    return DISPATCH(h, update_bodies(W, idx, b, count));
}
// Island of each body: the index of its body with the smallest index (-1 never occurs).
int orc_get_islands(void* p, int* out, int n) {
    Handle* h = (Handle*)p;
    return DISPATCH(h, ([&]() {
        int m = std::min(n, (int)W.bodies.size());
        std::vector<int> lo(W.isl.size(), -1);
        for (int i = (int)W.bodies.size() - 1; i >= 0; --i) lo[W.bodies[i].island] = i;
        for (int i = 0; i < m; ++i) out[i] = lo[W.bodies[i].island];
        return m;
    })());
}
// RigidBody.SetActivationState (RigidBody.cs:1025-1038): active != 0 -> World.ActivateBodyNextStep,
// else World.DeactivateBodyNextStep (the sleep timer set to infinity).
int orc_set_activation_state(void* p, int body, int active) {
    Handle* h = (Handle*)p;
    return DISPATCH(h, ([&]() {
        if (body < 0 || body >= (int)W.bodies.size()) return -1;
        if (active) W.activate_body_next_step(body);
        else W.bodies[body].sleep_time = std::numeric_limits<decltype(W.gravity.x)>::infinity();
        return 0;
    })());
}
// One World.Step; order = nkeys (key1, key2) pairs to replay, or NULL for key order.
int orc_step(void* p, float dt, const uint32_t* order, int nkeys) {
    Handle* h = (Handle*)p;
    std::vector<std::pair<uint32_t, uint32_t>> ord;
    if (order) {
        ord.resize(nkeys);
        for (int i = 0; i < nkeys; ++i) ord[i] = {order[2 * i], order[2 * i + 1]};
    }
    DISPATCH(h, (W.step(decltype(W.gravity.x)(dt), order ? &ord : nullptr), 0));
    return 0;
}
// World.Stabilize(dt, iterations, relaxations); order = nkeys (key1, key2) pairs or NULL (key order).
// Argument checks as StabilizeInternal (World.Step.cs:332-350): -1 bad argument, -2 non-finite dt.
int orc_stabilize(void* p, float dt, int iters, int relax, const uint32_t* order, int nkeys) {
    Handle* h = (Handle*)p;
    if (!std::isfinite(dt)) return -2;
    if (dt < 0.0f || iters < 1 || relax < 0) return -1;
    std::vector<std::pair<uint32_t, uint32_t>> ord;
    if (order) {
        ord.resize(nkeys);
        for (int i = 0; i < nkeys; ++i) ord[i] = {order[2 * i], order[2 * i + 1]};
    }
    DISPATCH(h, (W.stabilize(decltype(W.gravity.x)(dt), iters, relax, order ? &ord : nullptr), 0));
    return 0;
}
// Step with residual probes of the first substep (modes: bit 0 previous-frame warm start, bit 2
// cold); residual[3][iterations + 1], features [capacity][24], targets [capacity][3], flags and the
// two bodies of each row (optional); counts[0] = rows written, counts[1] = contact points.
int orc_step_probe(void* p, float dt, const uint32_t* order, int nkeys, int iterations, int modes, double* residual,
                   float* features, float* targets, long capacity, long* counts, const float* learned,
                   uint32_t* flags, int32_t* row_bodies) {
    Handle* h = (Handle*)p;
    std::vector<std::pair<uint32_t, uint32_t>> ord(nkeys);
    for (int i = 0; i < nkeys; ++i) ord[i] = {order[2 * i], order[2 * i + 1]};
    for (int i = 0; i < 3 * (iterations + 1); ++i) residual[i] = 0.0;
    auto run = [&](auto& W) {
        typename std::remove_reference<decltype(W)>::type::Probe pr;
        pr.iterations = iterations;
        pr.modes = modes;
        pr.capacity = capacity;
        pr.residual = residual;
        pr.features = features;
        pr.targets = targets;
        pr.learned = learned;
        pr.flags = flags;
        pr.row_bodies = row_bodies;
        W.step_impl(decltype(W.gravity.x)(dt), &ord, &pr);
        counts[0] = pr.rows;
        counts[1] = pr.points;
        return 0;
    };
    DISPATCH(h, run(W));
    return 0;
}
long orc_epa_calls(void* p) {
    Handle* h = (Handle*)p;
    return DISPATCH(h, W.epa_calls);
}
int orc_arbiter_count(void* p) {
    Handle* h = (Handle*)p;
    return DISPATCH(h, get_arbiters(W, nullptr, 0));
}
int orc_set_constraints(void* p, const jh_constraint* c, int n) {
    Handle* h = (Handle*)p;
    DISPATCH(h, (set_constraints(W, c, n), 0));
    return 0;
}
int orc_get_constraints(void* p, jh_constraint* c, int cap) {
    Handle* h = (Handle*)p;
    return DISPATCH(h, get_constraints(W, c, cap));
}
// Order in which the next steps visit the constraints (indices); n = 0 restores index order.
int orc_set_constraint_order(void* p, const int* order, int n) {
    Handle* h = (Handle*)p;
    auto set = [&](auto& W) {
        W.con_order.assign(order, order + n);
        for (int i : W.con_order)
            if (i < 0 || i >= (int)W.constraints.size()) return -1;
        return 0;
    };
    return h->fp64 ? set(*h->wd) : set(*h->wf);
}
// BeginCollide / EndCollide of the last step as (Key1, Key2) pairs, each list sorted by key.
// counts[0] / counts[1] receive the full list lengths.
int orc_get_events(void* p, uint32_t* begin, int begin_cap, uint32_t* end, int end_cap, int* counts) {
    Handle* h = (Handle*)p;
    auto put = [](const std::vector<std::pair<uint32_t, uint32_t>>& v, uint32_t* out, int cap) {
        for (int i = 0; i < (int)v.size() && i < cap && out; ++i) { out[2 * i] = v[i].first; out[2 * i + 1] = v[i].second; }
        return (int)v.size();
    };
    counts[0] = DISPATCH(h, put(W.ev_begin, begin, begin_cap));
    counts[1] = DISPATCH(h, put(W.ev_end, end, end_cap));
    return 0;
}

// ---- pure functions for known-answer tests ------------------------------------------------
// Shape support map in the local frame (FP32 or FP64 evaluation, result rounded to float).
int orc_support(int fp64, const jh_shape* s, const float* verts, const float* dir, float* out) {
    if (fp64) {
        Shape<double> sh = shape_d(*s, verts);
        vout(sh.support(vin<double>(dir)), out);
    } else {
        Shape<float> sh = shape_f(*s, verts);
        vout(sh.support(vin<float>(dir)), out);
    }
    return 0;
}
// World AABB of a shape at pose (q, p): out = min xyz, max xyz.
int orc_aabb(int fp64, const jh_shape* s, const float* verts, const float* q, const float* p, float* out) {
    if (fp64) {
        Shape<double> sh = shape_d(*s, verts);
        V3<double> mn, mx;
        sh.aabb({q[0], q[1], q[2], q[3]}, vin<double>(p), mn, mx);
        vout(mn, out);
        vout(mx, out + 3);
    } else {
        Shape<float> sh = shape_f(*s, verts);
        V3<float> mn, mx;
        sh.aabb({q[0], q[1], q[2], q[3]}, vin<float>(p), mn, mx);
        vout(mn, out);
        vout(mx, out + 3);
    }
    return 0;
}
// NarrowPhase.MprEpa + optional BuildManifold for two posed shapes.  out: [0..2] pA, [3..5] pB,
// [6..8] normal, [9] penetration, [10] epa flag, [11] manifold count, [12..23] manifold A,
// [24..35] manifold B.  Returns 1 on hit.
int orc_mpr_manifold(int fp64, const jh_shape* a, const jh_shape* b, const float* verts, const float* qa,
                     const float* pa, const float* qb, const float* pb, float epa_threshold, float* out) {
    if (fp64) return mpr_manifold_t<double>(a, b, verts, qa, pa, qb, pb, epa_threshold, out, shape_d);
    return mpr_manifold_t<float>(a, b, verts, qa, pa, qb, pb, epa_threshold, out, shape_f);
}
// BuildManifold alone (CollisionManifold.cs:393-566) from a given contact.  out as above from [11].
int orc_build_manifold(int fp64, const jh_shape* a, const jh_shape* b, const float* verts, const float* qa,
                       const float* pa, const float* qb, const float* pb, const float* pA, const float* pB,
                       const float* n, float* out) {
    if (fp64) {
        Shape<double> sa = shape_d(*a, verts), sb = shape_d(*b, verts);
        Manifold<double> m;
        m.build(sa, sb, {qa[0], qa[1], qa[2], qa[3]}, {qb[0], qb[1], qb[2], qb[3]}, vin<double>(pa), vin<double>(pb),
                vin<double>(pA), vin<double>(pB), vin<double>(n));
        out[0] = float(m.count);
        for (int i = 0; i < m.count; ++i) { vout(m.mA[i], out + 1 + 3 * i); vout(m.mB[i], out + 13 + 3 * i); }
        return m.count;
    }
    Shape<float> sa = shape_f(*a, verts), sb = shape_f(*b, verts);
    Manifold<float> m;
    m.build(sa, sb, {qa[0], qa[1], qa[2], qa[3]}, {qb[0], qb[1], qb[2], qb[3]}, vin<float>(pa), vin<float>(pb),
            vin<float>(pA), vin<float>(pB), vin<float>(n));
    out[0] = float(m.count);
    for (int i = 0; i < m.count; ++i) { vout(m.mA[i], out + 1 + 3 * i); vout(m.mB[i], out + 13 + 3 * i); }
    return m.count;
}
int orc_contact_op(int fp64, jh_body* bodies, jh_arbiter* arb, int op, float dt, int substeps, const float* p1,
                   const float* p2, const float* n) {
    if (fp64) return contact_op_t<double>(bodies, arb, op, dt, substeps, p1, p2, n);
    return contact_op_t<float>(bodies, arb, op, dt, substeps, p1, p2, n);
}
// SimplexSolverAB: AddVertex for n vertices (V, A, B: [n][3]) in order, then GetClosest.  results[i] =
// AddVertex's return for vertex i; out: [0..2] the last closest point, [3..5] pointA, [6..8] pointB,
// [9] the used-slot mask.
}  // extern "C"
namespace {
template <class R> int simplex_ab_t(int n, const double* V, const double* A, const double* B, int* results, double* out) {
    SimplexAB<R> s;
    V3<R> closest{};
    for (int i = 0; i < n; ++i) {
        Vertex<R> x;
        x.V = {R(V[3 * i]), R(V[3 * i + 1]), R(V[3 * i + 2])};
        x.A = {R(A[3 * i]), R(A[3 * i + 1]), R(A[3 * i + 2])};
        x.B = {R(B[3 * i]), R(B[3 * i + 1]), R(B[3 * i + 2])};
        results[i] = s.add_vertex(x, closest) ? 1 : 0;
    }
    V3<R> pa, pb;
    s.get_closest(pa, pb);
    const V3<R> o[3] = {closest, pa, pb};
    for (int k = 0; k < 3; ++k) { out[3 * k] = double(o[k].x); out[3 * k + 1] = double(o[k].y); out[3 * k + 2] = double(o[k].z); }
    out[9] = double(s.mask);
    return 0;
}
}  // namespace
extern "C" {
int orc_simplex_ab(int fp64, int n, const double* V, const double* A, const double* B, int* results, double* out) {
    if (n < 0) return -1;
    if (fp64) return simplex_ab_t<double>(n, V, A, B, results, out);
    return simplex_ab_t<float>(n, V, A, B, results, out);
}
// NarrowPhase.Sweep of two posed shapes with linear velocities va, vb.  out: [0..2] pointA, [3..5]
// pointB, [6..8] normal, [9] lambda.  Returns 1 if the sweep succeeded.
}  // extern "C"
namespace {
template <class R>
int sweep_t(const jh_shape* a, const jh_shape* b, const float* verts, const float* qa, const float* pa, const float* va,
            const float* qb, const float* pb, const float* vb, double* out, Shape<R> (*mk)(const jh_shape&, const float*)) {
    Shape<R> sa = mk(*a, verts), sb = mk(*b, verts);
    Q4<R> QA{R(qa[0]), R(qa[1]), R(qa[2]), R(qa[3])}, QB{R(qb[0]), R(qb[1]), R(qb[2]), R(qb[3])};
    V3<R> A, B, N;
    R lambda;
    bool ok = sweep(sa, sb, QA, QB, vin<R>(pa), vin<R>(pb), vin<R>(va), vin<R>(vb), A, B, N, lambda);
    const V3<R> o[3] = {A, B, N};
    for (int k = 0; k < 3; ++k) { out[3 * k] = double(o[k].x); out[3 * k + 1] = double(o[k].y); out[3 * k + 2] = double(o[k].z); }
    out[9] = double(lambda);
    return ok ? 1 : 0;
}
}  // namespace
extern "C" {
int orc_sweep(int fp64, const jh_shape* a, const jh_shape* b, const float* verts, const float* qa, const float* pa,
              const float* va, const float* qb, const float* pb, const float* vb, double* out) {
    if (fp64) return sweep_t<double>(a, b, verts, qa, pa, va, qb, pb, vb, out, shape_d);
    return sweep_t<float>(a, b, verts, qa, pa, va, qb, pb, vb, out, shape_f);
}
// StableMath.SinCos / Atan2 in FP32 (float) for the trigonometry checks.
void orc_sincos_f(float a, float* s, float* c) { StableMath<float>::sincos(a, *s, *c); }
float orc_atan2_f(float y, float x) { return StableMath<float>::atan2(y, x); }
double orc_atan2_d(double y, double x) { return StableMath<double>::atan2(y, x); }
void orc_sincos_d(double a, double* s, double* c) { StableMath<double>::sincos(a, *s, *c); }
// MathHelper.RotationQuaternion (FP32).
void orc_rotation_quaternion_f(const float* w, float dt, float* q) {
    Q4<float> r = rotation_quaternion(vin<float>(w), dt);
    q[0] = r.x; q[1] = r.y; q[2] = r.z; q[3] = r.w;
}
// ExpansionArithmetic.Side on FP32 coordinates (exact orientation sign).
int orc_exact_side(const float* a, const float* b, const float* c, const float* p) {
    return exact::side(vin<float>(a), vin<float>(b), vin<float>(c), vin<float>(p));
}
// Threads of the parallel parts of the step (OpenMP); results do not depend on it.
int orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
    return omp_get_max_threads();
#else
    return 1;
#endif
}
// Wall-clock time of k steps in key order (CPU baseline).
double orc_time_steps(void* p, float dt, int k) {
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < k; ++i) orc_step(p, dt, nullptr, 0);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"

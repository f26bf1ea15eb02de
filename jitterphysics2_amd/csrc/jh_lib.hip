// libjitterhip: C-ABI (include/jitterhip.h) over the gfx950 kernels.  One jh_world owns its device
// buffers and one HIP stream; jh_world_step enqueues the whole step and synchronises once in the
// middle (pair count, to size the following buffers) and once at the end (counters, timings).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "jh_kernels.h"
#include "jh_islands.h"
#include "jh_lws.h"
#include "jh_scan.h"
#include "jh_sort.h"

using namespace jh;

namespace {

// Last error text of the calling thread (jh_last_error).  A plain array: nothing to destroy when a
// thread or the process exits.
thread_local char g_err[512] = "";

jh_status fail(jh_status s, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return s;
}

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) return fail(JH_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_));         \
    } while (0)

// Device buffer owned by a world: freed by release() or, at the latest, with the world.
template <class T> struct DBuf {
    T* p = nullptr;
    size_t cap = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    hipError_t ensure(size_t n, bool keep = false) {
        if (n <= cap && p) return hipSuccess;
        size_t nc = std::max<size_t>(n + n / 4, 64);
        T* q = nullptr;
        hipError_t e = hipMalloc((void**)&q, nc * sizeof(T));
        if (e != hipSuccess) return e;
        if (keep && p && cap) {
            e = hipMemcpy(q, p, cap * sizeof(T), hipMemcpyDeviceToDevice);
            if (e != hipSuccess) return e;
        }
        if (p) (void)hipFree(p);
        p = q;
        cap = nc;
        return hipSuccess;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Device counters (ints).
enum {
    C_PAIRS = 0,
    C_DEEP = 1,
    C_COLL = 2,
    C_NEW = 3,
    C_A0 = 4,  // arbiter count of buffer 0
    C_A1 = 5,  // arbiter count of buffer 1
    C_NCOL = 6,
    C_REM = 7,   // 2 entries
    C_TERR = 9,  // traversal stack overflow flag
    C_NS = 10,   // shape count
    C_NB = 11,   // body count
    C_CONTACTS = 12,  // contact points solved
    C_EPAQ = 13,      // k_epa_wave: next deep pair to hand out
    C_HULL = 14,      // pairs left to k_narrow_hulls
    C_SPEC = 15,      // speculative sweep queue
    C_NMOV = 16,      // tile solver: non-static bodies with shapes
    C_TCELLS = 17,    // tile solver: (tile, colour) cells + 1
    C_MAXDEG = 19,    // most solved arbiters on one non-static body
    C_ZOMB = 20,      // arbiters removed at the end of the last step (keys in zkey); kept across steps
    C_EVB = 21,       // BeginCollide events of this step
    C_EVE = 22,       // EndCollide events of this step
    C_ASLEEP = 23,    // non-static bodies asleep after the last CheckDeactivation
    C_COUNT = 25      // (the host mirror's last entry is scratch)
};

// Solver kernels (jh_stats.solver, jh_tuning.solver).
enum { SOLVER_BARRIER = 0, SOLVER_FLOW = 1, SOLVER_WORLDS = 2, SOLVER_TILES = 3, SOLVER_LANES = 4 };
constexpr int kSmallWorld = 2048;  // bodies: worlds up to this size default to the world-local solver

constexpr int kEpaMaxThreads = 32768;  // polytopes resident at once (epa::kBytesPerThread each)

int bits_for(uint32_t v) {
    int b = 0;
    while (v) { b++; v >>= 1; }
    return b;
}
int grid_for(int n, int block = 256, int cap = 4096) {
    int g = (n + block - 1) / block;
    return std::max(1, std::min(g, cap));
}

// Device copy of a warm-start model (BF16 weights in both layouts, FP32 biases, normalisation).
struct LwsWeights {
    DBuf<uint16_t> w1, w2, w3, w2t, w3t;
    DBuf<float> b1, b2, b3, mean, istd, tscale;
    bool ready = false;
    hipError_t alloc() {
        hipError_t e;
        if ((e = w1.ensure(lws::kHid * lws::kInPad)) != hipSuccess) return e;
        if ((e = w2.ensure(lws::kHid * lws::kHid)) != hipSuccess) return e;
        if ((e = w3.ensure(lws::kOutPad * lws::kHid)) != hipSuccess) return e;
        if ((e = w2t.ensure(lws::kHid * lws::kHid)) != hipSuccess) return e;
        if ((e = w3t.ensure(lws::kHid * lws::kOutPad)) != hipSuccess) return e;
        if ((e = b1.ensure(lws::kHid)) != hipSuccess) return e;
        if ((e = b2.ensure(lws::kHid)) != hipSuccess) return e;
        if ((e = b3.ensure(lws::kOutPad)) != hipSuccess) return e;
        if ((e = mean.ensure(lws::kIn)) != hipSuccess) return e;
        if ((e = istd.ensure(lws::kIn)) != hipSuccess) return e;
        return tscale.ensure(3);
    }
    lws::Model model() const {
        lws::Model M;
        M.w1 = w1.p; M.w2 = w2.p; M.w3 = w3.p; M.w2t = w2t.p; M.w3t = w3t.p;
        M.b1 = b1.p; M.b2 = b2.p; M.b3 = b3.p; M.mean = mean.p; M.istd = istd.p; M.tscale = tscale.p;
        return M;
    }
    void release() {
        w1.release(); w2.release(); w3.release(); w2t.release(); w3t.release();
        b1.release(); b2.release(); b3.release(); mean.release(); istd.release(); tscale.release();
        ready = false;
    }
};

}  // namespace

struct jh_world {
    jh_config cfg;
    int device = 0;
    int cu_count = 0;
    hipStream_t st = nullptr;
    int nb = 0, ns = 0, nv = 0;
    int max_world = 0;
    int n_spec = 0;              // bodies with speculative contacts
    float broadphase_dt = 0.1f;  // World.BroadphaseStepDt (World.cs:416-418): dt of the last step
    std::vector<jh_body> host_bodies;
    // bodies
    DBuf<float4> pos, quat, vel, angvel, dvel, dangvel, imv, iw0, il0, mat, mat2, force, torque;
    DBuf<float2> iw1, il1;
    DBuf<int> motion, gravity, world, locks;
    DBuf<jh_body> body_io;
    // islands and sleeping (jh_islands.h): flags and timers per body, the solver's motion types, and
    // the union-find of each CheckDeactivation
    DBuf<float4> slp;
    DBuf<uint32_t> sflags, hfl, gfl, anew;
    DBuf<int> smotion, uf, hroot, groot, uidx;
    int n_asleep = 0;        // non-static bodies asleep after the last check
    bool isl_dirty = false;  // flags for the next check may be set
    DBuf<char> rbd_io;
    // shapes
    DBuf<ShapeRec> shapes;
    DBuf<float> verts;
    // broadphase
    DBuf<float4> amin, amax, lmin, lmax, nmin, nmax;
    DBuf<uint64_t> mk0, mk1;
    DBuf<uint32_t> mv0, mv1;
    DBuf<int> left, right, parent, leaf_parent, nflags;
    // broadphase launch sequence captured as a HIP graph (broadphase_signature), replayed every step
    int use_graphs = 1;
    hipGraph_t bp_graph = nullptr;
    hipGraphExec_t bp_exec = nullptr;
    std::vector<uint64_t> bp_sig;
    uint64_t* bp_keys = nullptr;
    uint32_t* bp_order = nullptr;
    long stats_graph_captures = 0;
    DBuf<float4> nbox;           // LBVH child records (k_lbvh_refit, k_traverse_wave): 4 per node
    DBuf<int4> nrng;
    int traverse_wave = 1;       // pair traversal: 1 k_traverse_wave, 0 k_traverse (JH_TRAVERSE=thread)
    int traverse_qpw = 32;       // k_traverse_wave: leaves per wave (JH_TRAVERSE_QPW=64|32|16)
    DBuf<int2> range;
    DBuf<uint32_t> bounds, hist;
    DBuf<uint64_t> pk0, pk1;
    int pair_cap = 0;
    int pairs_in_k1 = 0;  // sorted pairs ended in pk1
    int n_pairs = 0;
    // counters
    DBuf<int> cnt;
    int* hcnt = nullptr;  // pinned host mirror
    // narrowphase
    DBuf<jh_manifold> man;
    DBuf<DeepRec> deep;
    DBuf<char> epa_work, epa_side;
    DBuf<EpaRes> epa_res;
    DBuf<int> specq;  // pairs queued for the speculative sweep
    DBuf<int> hullq;  // pairs with point clouds, narrowphased by k_narrow_hulls
    bool has_clouds = false;  // some shape is a point cloud of at most kLdsCloud vertices
    int narrow_lds = 0;       // stage those clouds in LDS (JH_NARROW_LDS=1; measured slower, off by default)
    int narrow_sort = 1;      // narrowphase over the pairs grouped by shape types (JH_NARROW_SORT=0: key order)
    DBuf<int> pperm, pcls;    // that grouping: pair permutation; class counts and bases
    DBuf<uint64_t> zkey, evb, eve;  // keys of arbiters removed by the last step; collision events
    // constraints (solved by the barrier solver after the contacts of each pass)
    std::vector<jh_constraint> host_con;
    DBuf<jh_constraint> con;
    DBuf<con::Work> con_work;
    DBuf<int> con_order, con_offs;
    std::vector<int> con_order_h, con_offs_h;  // batch order of the enabled constraints
    bool con_dirty = false;                    // recolour before the next step (constraints or bodies changed)
    int epa_threads = 0;
    bool epa_wave = true;      // wave-cooperative EPA (k_epa_wave); false: one thread per pair (k_epa)
    int epa_wave_blocks = 0;   // resident single-wave workgroups of k_epa_wave
    // contact cache
    DBuf<uint32_t> flags, offs, sums, cnew, noffs;
    DBuf<uint64_t> ckey, nkey;
    DBuf<int> cpair, cfound;
    DBuf<jh_arbiter> arb[2];
    DBuf<uint64_t> akey[2];
    int cur = 0;
    int a_host = 0;  // arbiters in arb[cur] (after the last step / upload)
    // colouring + order
    DBuf<int> colour, chist, coffs;
    DBuf<unsigned long long> bmax, mlo, mhi;  // bmax: 2 x body capacity
    DBuf<int> cwork;                          // colouring worklists: 2 x arbiter capacity
    // colour-ordered contact rows of the dataflow solver (ContactSoA), capacity soa_cap arbiters
    DBuf<int2> soa_bodies;
    DBuf<uint2> soa_um;
    DBuf<float2> soa_fs;
    DBuf<int4> soa_nd;
    DBuf<uint32_t> soa_flag;
    DBuf<float> soa_f;
    int soa_cap = 0;
    // world-local solver (many independent worlds)
    int n_worlds = 0, world_blocks = 0, largest_world = 0;
    int world_occ[4] = {0, 0, 0, 0};  // resident k_solve_worlds workgroups per CU at 64, 128, 256, 512 threads
    DBuf<int> wbody, wboff, wsched, wcells, wqueue;
    int world_threads = 512;     // k_solve_worlds workgroup size (autotuned: 64 .. 512)
    DBuf<uint32_t> whist, woff, wcur, wsums;
    DBuf<uint64_t> wlist;
    DBuf<float4> vwpack;  // dataflow solver: packed velocities, 2 per body
    // solver kernel: SOLVER_FLOW unless chosen otherwise (environment, world count, autotuner)
    int kind = SOLVER_FLOW;
    // lane solver (k_solve_lanes): workgroups and arbiter records
    int lane_blocks = 0, lane_max = 0;
    int flow_qmode = 1;                // k_solve_flow chunk ownership (SolveArgs.qmode)
    DBuf<float4> soa_mv;               // versioned mutable contact fields of k_solve_flow
    DBuf<int> fqueue;
    DBuf<float4> rec;
    // tile solver (one workgroup per Morton range of bodies, private bodies in LDS)
    int tiles = 0;            // tiles = workgroups of k_solve_tiles
    int n_moving = 0;         // non-static bodies (upper bound of the tiled bodies)
    int tile_occ_lds = -1, tile_occ = 0;  // occupancy cache: LDS bytes -> resident workgroups per CU
    int sorted_in1 = 0;       // broadphase order ended in mv1
    DBuf<int> tfirst, tbody, tile_of, tslot;
    DBuf<uint32_t> tflags, trank, tshared, thist, toff, tcur, tsums;
    DBuf<uint64_t> tlist;
    DBuf<uint64_t> ok0, ok1, solve_keys;
    // dataflow solver
    DBuf<int2> rank;
    DBuf<int> deg;
    DBuf<int> scnt;  // per body: arbiters in the serial colour batch (k_rank_serial)
    DBuf<unsigned> bar;       // grid barrier state and abort flag of the persistent kernels (64 words)
    unsigned spin_limit = 0;  // polls before a device wait gives up; 0: kSpinLimitDefault
    unsigned* habort = nullptr;  // pinned host mirror of the abort flag
    bool recolour = true;  // next colouring ignores the colours stored in the arbiters
    int solve_count = 0, solve_batches = 0;
    std::vector<int> solve_offsets;
    // cooperative occupancy
    int coop_colour = 0, coop_solve = 0, flow_blocks = 0, colour_blocks = 64, flow_max = 0;
    hipEvent_t ev[10];  // 0 start, 1 broadphase, 2 narrowphase, 3 cache, 4 colouring, 5 solve, 6 end, 7 EPA start,
                        // 8 / 9 around the solver kernel itself
    // learned warm start and residual probe
    LwsWeights lws;
    DBuf<float4> pvel, pang;
    DBuf<jh_arbiter> parb;
    DBuf<double> ppart;
    DBuf<float> pfeat, ptarg, plam, plearn;
    DBuf<uint8_t> pvalid;
    int coop_probe = 0;
    jh_stats stats;
    bool have_step = false;
    int ev_begin = 0, ev_end = 0;  // collision events of the last step (in evb / eve)

    unsigned spin() const { return spin_limit ? spin_limit : kSpinLimitDefault; }
    Bodies bodies() {
        Bodies B;
        B.pos = pos.p; B.quat = quat.p; B.vel = vel.p; B.angvel = angvel.p; B.dvel = dvel.p; B.dangvel = dangvel.p;
        B.imv = imv.p; B.iw0 = iw0.p; B.iw1 = iw1.p; B.il0 = il0.p; B.il1 = il1.p; B.mat = mat.p; B.mat2 = mat2.p;
        B.force = force.p; B.torque = torque.p; B.motion = motion.p; B.gravity = gravity.p; B.world = world.p;
        B.slp = slp.p; B.sflags = sflags.p; B.locks = locks.p;
        return B;
    }
    // Islands are checked while deactivation is allowed or a body sleeps; otherwise every body is awake.
    bool islands_on() const { return cfg.allow_deactivation != 0 || n_asleep > 0; }
    // The bodies as the solver sees them: a sleeping body is static.
    Bodies solver_bodies() {
        Bodies B = bodies();
        if (islands_on()) B.motion = smotion.p;
        return B;
    }
};

namespace {

__global__ void k_reset_counters(int* cnt, uint32_t* bounds) {
    int t = threadIdx.x;
    if (t == C_PAIRS || t == C_DEEP || t == C_TERR || t == C_COLL || t == C_NEW || t == C_CONTACTS || t == C_HULL ||
        t == C_MAXDEG || t == C_SPEC || t == C_EVB || t == C_EVE || t == C_EPAQ)
        cnt[t] = 0;
    if (t < 3) bounds[t] = 0xffffffffu;
    else if (t < 6) bounds[t] = 0u;
}
__global__ void k_set_int(int* p, int v) { *p = v; }
__global__ void k_mark_listed(const int* __restrict__ idx, int n, uint32_t* __restrict__ mark) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) mark[idx[i]] = 1u;
}
__global__ void k_locks_changed(jh_arbiter* __restrict__ arb, const int* __restrict__ a_count,
                                jh_constraint* __restrict__ con, int ncon, const uint32_t* __restrict__ mark) {
    int A = *a_count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < A + ncon; i += gridDim.x * blockDim.x) {
        if (i < A) {
            if (mark[arb[i].body1] || mark[arb[i].body2]) arb[i].usage_mask = 0u;
        } else {
            jh_constraint& c = con[i - A];
            if (!mark[c.body1] && !mark[c.body2]) continue;
            int n = con::impulse_count(c.type);
            for (int k = 0; k < n; ++k) c.state[k] = 0.0f;
        }
    }
}
__global__ void k_solver_motion(int nb, const int* __restrict__ motion, const uint32_t* __restrict__ sflags,
                                int* __restrict__ smotion) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x)
        smotion[b] = (motion[b] != JH_MOTION_STATIC && !(sflags[b] & isl::kAwake)) ? JH_MOTION_STATIC : motion[b];
}
__global__ void k_solve_keys(const uint64_t* __restrict__ order, const uint64_t* __restrict__ akey,
                             const int* __restrict__ a_count, uint64_t* __restrict__ out) {
    int A = *a_count;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < A; k += gridDim.x * blockDim.x)
        out[k] = akey[uint32_t(order[k])];
}

jh_status check_config(const jh_config* c) {
    if (!c) return fail(JH_ERR_INVALID_ARG, "config is NULL");
    if (c->substep_count < 1) return fail(JH_ERR_INVALID_ARG, "substep_count must be >= 1 (World.SubstepCount)");
    if (c->solver_iterations < 1) return fail(JH_ERR_INVALID_ARG, "solver_iterations must be >= 1");
    if (c->relaxation_iterations < 0) return fail(JH_ERR_INVALID_ARG, "relaxation_iterations must be >= 0");
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(c->gravity[k])) return fail(JH_ERR_NON_FINITE, "gravity must be finite");
    if (c->narrowphase_filter != JH_FILTER_NONE && c->narrowphase_filter != JH_FILTER_TRIANGLE_EDGE)
        return fail(JH_ERR_INVALID_ARG, "narrowphase_filter must be JH_FILTER_NONE or JH_FILTER_TRIANGLE_EDGE");
    if (c->warm_start != JH_WARM_START_PREVIOUS && c->warm_start != JH_WARM_START_LEARNED)
        return fail(JH_ERR_INVALID_ARG, "warm_start must be JH_WARM_START_PREVIOUS or JH_WARM_START_LEARNED");
    return JH_OK;
}

// RigidBody.DeactivationTime / DeactivationThreshold setters (ArgumentCheck.NonNegative,
// RigidBody.cs:476-507): finite and not negative.
jh_status check_sleep_fields(const jh_body& b, int i) {
    const float v[3] = {b.deactivation_threshold_linear_sq, b.deactivation_threshold_angular_sq, b.deactivation_time};
    for (float x : v)
        if (!std::isfinite(x) || x < 0.0f)
            return fail(JH_ERR_INVALID_ARG, "body %d: deactivation thresholds and time must be finite and >= 0", i);
    return JH_OK;
}

// The solver's motion types after a host write (a sleeping body stays static to the solver).
void refresh_solver_motion(jh_world* w) {
    if (w->nb == 0) return;
    k_solver_motion<<<grid_for(w->nb, 256, 1 << 20), 256, 0, w->st>>>(w->nb, w->motion.p, w->sflags.p, w->smotion.p);
}

// ActivateBodyNextStep (active != 0; wake_static: wakeUpStatic, a static body wakes its partners) or
// DeactivateBodyNextStep for the listed bodies (device indices).
void wake_listed(jh_world* w, const int* idx, int n, int active, int wake_static) {
    hipStream_t st = w->st;
    isl::k_activate<<<grid_for(n, 256, 1 << 20), 256, 0, st>>>(idx, n, active, wake_static, w->motion.p, w->sflags.p,
                                                               w->slp.p);
    if (active && wake_static) {
        int ncon = int(w->host_con.size());
        int* cnt = w->cnt.p;
        int total = w->a_host + int(w->zkey.cap) + ncon;
        isl::k_wake_partners<<<grid_for(std::max(total, 1)), 256, 0, st>>>(
            w->arb[w->cur].p, cnt + (w->cur == 0 ? C_A0 : C_A1), w->zkey.p, w->zkey.cap ? cnt + C_ZOMB : nullptr,
            w->shapes.p, w->con.p, ncon, w->motion.p, w->sflags.p, w->slp.p);
        isl::k_clear_bits<<<grid_for(std::max(w->nb, 1), 256, 1 << 20), 256, 0, st>>>(w->nb, w->sflags.p, isl::kSwake);
    }
    w->isl_dirty = true;
}

// The rest of the AllowedMotion setter for bodies whose locks changed (RigidBody.cs:589-592):
// ClearContactCache (the usage masks of their arbiters to zero) and ResetWarmStart of their
// constraints (the accumulated impulses).
jh_status locks_changed(jh_world* w, const std::vector<int>& bodies) {
    hipStream_t st = w->st;
    HIPCHK(w->hfl.ensure(std::max(w->nb, 1)));
    k_fill_u32<<<grid_for(std::max(w->nb, 1), 256, 1 << 20), 256, 0, st>>>(w->hfl.p, size_t(std::max(w->nb, 1)), 0u);
    HIPCHK(w->uidx.ensure(bodies.size()));
    HIPCHK(hipMemcpyAsync(w->uidx.p, bodies.data(), bodies.size() * sizeof(int), hipMemcpyHostToDevice, st));
    k_mark_listed<<<grid_for(int(bodies.size())), 256, 0, st>>>(w->uidx.p, int(bodies.size()), w->hfl.p);
    int ncon = int(w->host_con.size());
    int total = w->a_host + ncon;
    k_locks_changed<<<grid_for(std::max(total, 1)), 256, 0, st>>>(
        w->arb[w->cur].p, w->cnt.p + (w->cur == 0 ? C_A0 : C_A1), w->con.p, ncon, w->hfl.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return JH_OK;
}

// Islands rebuilt with every body awake and on the active list (new contact cache or constraint set).
void reset_islands(jh_world* w) {
    if (w->nb == 0) return;
    isl::k_upload<<<grid_for(w->nb, 256, 1 << 20), 256, 0, w->st>>>(w->nb, w->motion.p, w->sflags.p, w->slp.p, 1);
    refresh_solver_motion(w);
    w->n_asleep = 0;
    w->isl_dirty = true;
}

// A convex hull's vertex block (JH_SHAPE_CONVEXHULL): V vertices, then (first, end) per vertex into
// the neighbour list, contiguous from 0, then the list of neighbour indices below V, all within the
// vertex array.  Only a well-formed graph keeps every hill climb inside the block.
jh_status check_hull(const jh_shape& s, const float* vertices, int nv, int i) {
    long V = s.vertex_count, off = s.vertex_offset;
    if (V < 1 || V > 65535 || off < 0 || off + V > nv)
        return fail(JH_ERR_INVALID_ARG, "shape %d: bad convex hull vertex range", i);
    const float* base = vertices + 3 * off;
    size_t avail = 3 * size_t(nv - off);
    if (5 * size_t(V) > avail) return fail(JH_ERR_INVALID_ARG, "shape %d: convex hull adjacency out of bounds", i);
    const float* adj = base + 3 * V;
    auto as_index = [](float f, long& out) {
        if (!(f >= 0.0f) || f > 16777216.0f || float(long(f)) != f) return false;
        out = long(f);
        return true;
    };
    long prev_end = 0, n_list = 0;
    for (long k = 0; k < V; ++k) {
        long lo, hi;
        if (!as_index(adj[2 * k], lo) || !as_index(adj[2 * k + 1], hi) || lo != prev_end || hi < lo)
            return fail(JH_ERR_INVALID_ARG, "shape %d: convex hull adjacency ranges are not contiguous", i);
        prev_end = hi;
    }
    n_list = prev_end;
    if (5 * size_t(V) + size_t(n_list) > avail)
        return fail(JH_ERR_INVALID_ARG, "shape %d: convex hull neighbour list out of bounds", i);
    const float* nbl = adj + 2 * V;
    for (long k = 0; k < n_list; ++k) {
        long x;
        if (!as_index(nbl[k], x) || x >= V)
            return fail(JH_ERR_INVALID_ARG, "shape %d: convex hull neighbour index out of range", i);
    }
    for (long k = 0; k < 3 * V; ++k)
        if (!std::isfinite(base[k])) return fail(JH_ERR_NON_FINITE, "shape %d: non-finite hull vertex", i);
    return JH_OK;
}

jh_status ensure_pair_buffers(jh_world* w, int n) {
    HIPCHK(w->pk0.ensure(n));
    HIPCHK(w->pk1.ensure(n));
    w->pair_cap = int(std::min(w->pk0.cap, w->pk1.cap));
    return JH_OK;
}

// Broadphase: AABBs, Morton keys, sort, LBVH, refit, traversal, pair sort.  Syncs once to read the
// pair count (growing the pair buffer and re-traversing on overflow).
void traverse_launch(jh_world* w, const uint64_t* keys, const uint32_t* order);

// The broadphase launches up to the first traversal: the same sequence every step while the shapes,
// the buffers and the broadphase step length stay, so it is captured once into a HIP graph and
// replayed (JH_GRAPH=0: launched directly).  Returns where the sorted keys and shape order ended.
void broadphase_launches(jh_world* w, uint64_t*& keys, uint32_t*& order) {
    hipStream_t st = w->st;
    Bodies B = w->bodies();
    int ns = w->ns;
    int* cnt = w->cnt.p;
    k_reset_counters<<<1, 64, 0, st>>>(cnt, w->bounds.p);
    int g = grid_for(ns, 256, 1 << 20);
    k_aabb<<<g, 256, 0, st>>>(w->shapes.p, ns, w->verts.p, B, w->broadphase_dt, w->amin.p, w->amax.p, w->bounds.p);
    k_morton<<<g, 256, 0, st>>>(w->shapes.p, ns, w->world.p, w->amin.p, w->amax.p, w->bounds.p, w->mk0.p, w->mv0.p);
    int shifts[8];
    int nsh = 0;
    for (int s = 0; s < 32; s += 8) shifts[nsh++] = s;
    int wb = bits_for(uint32_t(w->max_world));
    for (int s = 32; s < 32 + wb; s += 8) shifts[nsh++] = s;
    bool in1 = rsort::sort_pairs(w->mk0.p, w->mk1.p, w->mv0.p, w->mv1.p, cnt + C_NS, ns, shifts, nsh, w->hist.p, st);
    keys = in1 ? w->mk1.p : w->mk0.p;
    order = in1 ? w->mv1.p : w->mv0.p;
    int gi = grid_for(ns - 1, 256, 1 << 20);
    k_lbvh_build<<<gi, 256, 0, st>>>(keys, ns, w->left.p, w->right.p, w->range.p, w->parent.p, w->leaf_parent.p,
                                     w->nflags.p);
    const uint32_t* sfl = w->islands_on() ? w->sflags.p : nullptr;
    k_lbvh_refit<<<g, 256, 0, st>>>(ns, order, w->amin.p, w->amax.p, w->lmin.p, w->lmax.p, w->left.p, w->right.p,
                                    w->parent.p, w->leaf_parent.p, w->nmin.p, w->nmax.p, w->nflags.p, w->shapes.p,
                                    w->motion.p, sfl, w->range.p, w->traverse_wave ? w->nbox.p : nullptr,
                                    w->nrng.p);
    traverse_launch(w, keys, order);
}

void traverse_launch(jh_world* w, const uint64_t* keys, const uint32_t* order) {
    hipStream_t st = w->st;
    int ns = w->ns;
    int* cnt = w->cnt.p;
    int g = grid_for(ns, 256, 1 << 20);
    const uint32_t* sfl = w->islands_on() ? w->sflags.p : nullptr;
    k_set_int<<<1, 1, 0, st>>>(cnt + C_PAIRS, 0);
    if (w->traverse_wave) {
        // one thread per leaf at 64 leaves per wave; fewer leaves per wave only while the grid stays
        // within the launch cap (4 threads per leaf at 16)
        int q = int64_t(ns) * 4 <= (int64_t(1) << 28) ? w->traverse_qpw : 64;
        int gw = grid_for(int((int64_t(ns) + q - 1) / q * 64), 256, 1 << 20);
        k_traverse_wave<<<gw, 256, 0, st>>>(ns, keys, w->lmin.p, w->lmax.p, w->nbox.p, w->nrng.p, w->pk0.p,
                                            cnt + C_PAIRS, w->pair_cap, cnt + C_TERR, q);
    } else
        k_traverse<<<g, 256, 0, st>>>(ns, keys, order, w->shapes.p, w->motion.p, w->lmin.p, w->lmax.p, w->left.p,
                                      w->right.p, w->range.p, w->nmin.p, w->nmax.p, w->pk0.p, cnt + C_PAIRS,
                                      w->pair_cap, cnt + C_TERR, sfl);
}

// What the broadphase graph's launches depend on: a graph is replayed only while this is unchanged.
std::vector<uint64_t> broadphase_signature(jh_world* w) {
    std::vector<uint64_t> v;
    Bodies B = w->bodies();
    const uint64_t* bw = reinterpret_cast<const uint64_t*>(&B);
    v.assign(bw, bw + (sizeof(Bodies) + 7) / 8);
    uint32_t dtb;
    memcpy(&dtb, &w->broadphase_dt, 4);
    const void* ptrs[] = {w->cnt.p, w->bounds.p, w->shapes.p, w->verts.p, w->world.p, w->amin.p, w->amax.p,
                          w->mk0.p, w->mk1.p, w->mv0.p, w->mv1.p, w->hist.p, w->left.p, w->right.p, w->range.p,
                          w->parent.p, w->leaf_parent.p, w->nflags.p, w->lmin.p, w->lmax.p, w->nmin.p, w->nmax.p,
                          w->nbox.p, w->nrng.p, w->motion.p, w->sflags.p, w->pk0.p};
    for (const void* p : ptrs) v.push_back(uint64_t(uintptr_t(p)));
    v.push_back(uint64_t(uint32_t(w->ns)) | uint64_t(uint32_t(w->max_world)) << 32);
    v.push_back(uint64_t(uint32_t(w->pair_cap)) | uint64_t(dtb) << 32);
    v.push_back(uint64_t(w->traverse_wave) | uint64_t(w->islands_on() ? 1 : 0) << 8 | uint64_t(w->traverse_qpw) << 16);
    return v;
}

jh_status broadphase(jh_world* w) {
    hipStream_t st = w->st;
    int ns = w->ns;
    int* cnt = w->cnt.p;
    w->n_pairs = 0;
    if (ns < 2) {
        k_reset_counters<<<1, 64, 0, st>>>(cnt, w->bounds.p);
        HIPCHK(hipGetLastError());
        return JH_OK;
    }
    uint64_t* keys = nullptr;
    uint32_t* order = nullptr;
    if (w->use_graphs) {
        std::vector<uint64_t> sig = broadphase_signature(w);
        if (!w->bp_exec || sig != w->bp_sig) {  // capture the launches once for this configuration
            if (w->bp_exec) HIPCHK(hipGraphExecDestroy(w->bp_exec));
            if (w->bp_graph) HIPCHK(hipGraphDestroy(w->bp_graph));
            w->bp_exec = nullptr;
            w->bp_graph = nullptr;
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            broadphase_launches(w, w->bp_keys, w->bp_order);
            hipError_t e = hipGetLastError();
            hipGraph_t gr = nullptr;
            hipError_t e2 = hipStreamEndCapture(st, &gr);
            if (e != hipSuccess) return fail(JH_ERR_HIP, "broadphase launch: %s", hipGetErrorString(e));
            if (e2 != hipSuccess) return fail(JH_ERR_HIP, "broadphase capture: %s", hipGetErrorString(e2));
            w->bp_graph = gr;
            HIPCHK(hipGraphInstantiate(&w->bp_exec, gr, nullptr, nullptr, 0));
            w->bp_sig = std::move(sig);
            w->stats_graph_captures++;
        }
        HIPCHK(hipGraphLaunch(w->bp_exec, st));
        keys = w->bp_keys;
        order = w->bp_order;
    } else {
        broadphase_launches(w, keys, order);
    }
    w->sorted_in1 = keys == w->mk1.p;
    for (int attempt = 0; attempt < 3; ++attempt) {
        if (attempt > 0) traverse_launch(w, keys, order);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(w->hcnt, cnt, C_COUNT * sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (w->hcnt[C_TERR]) return fail(JH_ERR_CAPACITY, "LBVH traversal stack overflow");
        int np = w->hcnt[C_PAIRS];
        if (np <= w->pair_cap) {
            w->n_pairs = np;
            break;
        }
        w->stats.pair_overflow++;
        jh_status s = ensure_pair_buffers(w, np + np / 2);
        if (s != JH_OK) return s;
    }
    int np = w->n_pairs;
    if (np > 1) {
        int pshift[8];
        int npsh = 0;
        int sb = bits_for(uint32_t(ns - 1));
        for (int s = 0; s < sb; s += 8) pshift[npsh++] = s;
        for (int s = 32; s < 32 + sb; s += 8) pshift[npsh++] = s;
        HIPCHK(w->hist.ensure(rsort::hist_entries(size_t(w->pair_cap))));
        w->pairs_in_k1 = rsort::sort_pairs(w->pk0.p, w->pk1.p, nullptr, nullptr, cnt + C_PAIRS, np, pshift, npsh,
                                           w->hist.p, st);
    } else {
        w->pairs_in_k1 = 0;
    }
    HIPCHK(hipGetLastError());
    return JH_OK;
}

jh_status ensure_step_buffers(jh_world* w) {
    size_t np = std::max(w->n_pairs, 1);
    HIPCHK(w->man.ensure(np));
    HIPCHK(w->deep.ensure(np));
    HIPCHK(w->flags.ensure(np));
    HIPCHK(w->offs.ensure(np));
    HIPCHK(w->sums.ensure(np / 256 + 2));
    HIPCHK(w->cnew.ensure(np));
    HIPCHK(w->noffs.ensure(np));
    HIPCHK(w->ckey.ensure(np));
    HIPCHK(w->nkey.ensure(np));
    HIPCHK(w->cpair.ensure(np));
    HIPCHK(w->cfound.ensure(np));
    size_t na = size_t(w->a_host) + np;
    HIPCHK(w->arb[w->cur].ensure(na, true));
    HIPCHK(w->akey[w->cur].ensure(na, true));
    HIPCHK(w->arb[1 - w->cur].ensure(na));
    HIPCHK(w->akey[1 - w->cur].ensure(na));
    HIPCHK(w->colour.ensure(na));
    HIPCHK(w->cwork.ensure(2 * na));
    HIPCHK(w->zkey.ensure(na, true));  // holds the last step's removed keys until k_events has read them
    HIPCHK(w->evb.ensure(np));
    HIPCHK(w->eve.ensure(w->zkey.cap));
    HIPCHK(w->anew.ensure(na));
    if (w->kind != SOLVER_BARRIER && int(na) > w->soa_cap) {
        size_t cap = na + na / 4;
        HIPCHK(w->soa_bodies.ensure(cap));
        HIPCHK(w->soa_um.ensure(cap));
        HIPCHK(w->soa_fs.ensure(cap));
        HIPCHK(w->soa_nd.ensure(cap));
        HIPCHK(w->soa_flag.ensure(4 * cap));
        HIPCHK(w->soa_f.ensure(4 * kSoaFields * cap));
        HIPCHK(w->soa_mv.ensure(4 * 3 * cap));
        w->soa_cap = int(cap);
    }
    HIPCHK(w->ok0.ensure(na));
    HIPCHK(w->ok1.ensure(na));
    HIPCHK(w->solve_keys.ensure(na));
    HIPCHK(w->rank.ensure(na));
    HIPCHK(w->flags.ensure(na));
    HIPCHK(w->offs.ensure(na));
    HIPCHK(w->sums.ensure(na / 256 + 2));
    HIPCHK(w->hist.ensure(rsort::hist_entries(na)));
    if (w->kind == SOLVER_TILES) HIPCHK(w->tlist.ensure(na));
    if (w->kind == SOLVER_LANES) HIPCHK(w->rec.ensure(na * kRecVec));
    return JH_OK;
}

// Batches of the enabled constraints: first fit in index order over the colours of their dynamic
// bodies (static and kinematic bodies are only read by constraints, so they never conflict), then
// the indices grouped colour by colour.  Recomputed when constraints or body motion types change.
jh_status colour_constraints(jh_world* w) {
    int n = int(w->host_con.size());
    std::vector<int> colour(n, -1);
    std::vector<std::vector<uint64_t>> used(w->nb);
    int ncol = 0;
    auto is_dyn = [&](int b) { return w->host_bodies[b].motion_type == JH_MOTION_DYNAMIC; };
    for (int i = 0; i < n; ++i) {
        const jh_constraint& c = w->host_con[i];
        if (c.body1 >= w->nb || c.body2 >= w->nb)
            return fail(JH_ERR_STATE, "constraint %d refers to a body beyond the %d bodies set", i, w->nb);
        if (!c.enabled) continue;
        int b[2] = {c.body1, c.body2};
        for (int col = 0;; ++col) {
            bool free = true;
            for (int e = 0; e < 2 && free; ++e) {
                if (!is_dyn(b[e])) continue;
                const auto& u = used[b[e]];
                size_t word = size_t(col) >> 6;
                if (word < u.size() && (u[word] >> (col & 63) & 1)) free = false;
            }
            if (!free) continue;
            for (int e = 0; e < 2; ++e) {
                if (!is_dyn(b[e])) continue;
                auto& u = used[b[e]];
                size_t word = size_t(col) >> 6;
                if (u.size() <= word) u.resize(word + 1, 0);
                u[word] |= 1ull << (col & 63);
            }
            colour[i] = col;
            ncol = std::max(ncol, col + 1);
            break;
        }
    }
    std::vector<int> count(ncol + 1, 0);
    for (int i = 0; i < n; ++i)
        if (colour[i] >= 0) count[colour[i] + 1]++;
    for (int k = 0; k < ncol; ++k) count[k + 1] += count[k];
    w->con_offs_h.assign(count.begin(), count.end());
    w->con_order_h.assign(count[ncol], 0);
    std::vector<int> fill(count.begin(), count.end() - 1);
    for (int i = 0; i < n; ++i)
        if (colour[i] >= 0) w->con_order_h[fill[colour[i]]++] = i;
    // Within a batch the constraints touch disjoint dynamic bodies, so their order is free: group them
    // by kind, so that a wavefront runs one kind's code instead of diverging over several.
    for (int k = 0; k < ncol; ++k)
        std::stable_sort(w->con_order_h.begin() + count[k], w->con_order_h.begin() + count[k + 1],
                         [&](int a, int b) { return w->host_con[a].type < w->host_con[b].type; });
    HIPCHK(w->con_order.ensure(std::max<size_t>(w->con_order_h.size(), 1)));
    HIPCHK(w->con_offs.ensure(w->con_offs_h.size()));
    if (!w->con_order_h.empty())
        HIPCHK(hipMemcpyAsync(w->con_order.p, w->con_order_h.data(), w->con_order_h.size() * sizeof(int),
                              hipMemcpyHostToDevice, w->st));
    HIPCHK(hipMemcpyAsync(w->con_offs.p, w->con_offs_h.data(), w->con_offs_h.size() * sizeof(int),
                          hipMemcpyHostToDevice, w->st));
    w->con_dirty = false;
    return JH_OK;
}

int constraint_batches(const jh_world* w) { return w->con_offs_h.empty() ? 0 : int(w->con_offs_h.size()) - 1; }

jh_status launch_coop(const void* fn, int grid, int block, void** args, hipStream_t st) {
    HIPCHK(hipLaunchCooperativeKernel(fn, dim3(grid), dim3(block), args, 0, st));
    return JH_OK;
}

// Tile solver: dynamic LDS of k_solve_tiles for T tiles (24 bytes per tile body), 0 if the tiles do
// not fit (more than 64 KiB per tile, or fewer co-resident workgroups than tiles).
size_t tile_lds_bytes(jh_world* w, int T) {
    if (T < 1) return 0;
    size_t per = (size_t(std::max(w->n_moving, 1)) + T - 1) / T;
    size_t lds = std::max<size_t>(per, 1) * 24;
    if (lds > 65536) return 0;
    if (int(lds) != w->tile_occ_lds) {
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_solve_tiles, 256, lds) != hipSuccess) occ = 0;
        w->tile_occ_lds = int(lds);
        w->tile_occ = occ;
    }
    return size_t(w->tile_occ) * w->cu_count >= size_t(T) ? lds : 0;
}

// Tile solver set-up: tiles from the broadphase order, the shared flags, and the (tile, colour)
// ordered arbiter list.
jh_status tile_setup(jh_world* w, const uint64_t* order, const int* a_cur, int amax_) {
    hipStream_t st = w->st;
    int* cnt = w->cnt.p;
    int ns = w->ns, nb = std::max(w->nb, 1), T = w->tiles;
    // tiles hold the bodies the solver moves: sleeping bodies (static to the solver) stay out
    const int* smotion = w->solver_bodies().motion;
    size_t cells = size_t(T) * 128 + 1;
    HIPCHK(w->tfirst.ensure(nb));
    HIPCHK(w->tile_of.ensure(nb));
    HIPCHK(w->tslot.ensure(nb));
    HIPCHK(w->tbody.ensure(nb));
    HIPCHK(w->tshared.ensure(nb));
    HIPCHK(w->tflags.ensure(std::max(ns, 1)));
    HIPCHK(w->trank.ensure(std::max(ns, 1)));
    HIPCHK(w->tsums.ensure(std::max<size_t>(size_t(std::max(ns, 1)), cells) / 256 + 2));
    HIPCHK(w->thist.ensure(cells));
    HIPCHK(w->toff.ensure(cells));
    HIPCHK(w->tcur.ensure(cells));
    int gb = grid_for(nb);
    k_fill_u32<<<gb, 256, 0, st>>>((uint32_t*)w->tfirst.p, size_t(nb), 0x7fffffffu);
    k_fill_u32<<<gb, 256, 0, st>>>((uint32_t*)w->tile_of.p, size_t(nb), 0xffffffffu);
    k_fill_u32<<<gb, 256, 0, st>>>(w->tshared.p, size_t(nb), 0u);
    int gc = grid_for(int(std::min<size_t>(cells, 1 << 30)));
    k_fill_u32<<<gc, 256, 0, st>>>(w->thist.p, cells, 0u);
    k_fill_u32<<<gc, 256, 0, st>>>(w->tcur.p, cells, 0u);
    if (ns >= 2) {  // the broadphase sorted the shapes
        const uint32_t* sorted = w->sorted_in1 ? w->mv1.p : w->mv0.p;
        int gs = grid_for(ns);
        k_body_first<<<gs, 256, 0, st>>>(sorted, cnt + C_NS, w->shapes.p, w->tfirst.p);
        k_tile_flags<<<gs, 256, 0, st>>>(sorted, cnt + C_NS, w->shapes.p, smotion, w->tfirst.p, w->tflags.p);
        scan::exclusive_scan(w->tflags.p, w->trank.p, cnt + C_NS, ns, w->tsums.p, cnt + C_NMOV, st);
        k_tile_assign<<<gs, 256, 0, st>>>(sorted, cnt + C_NS, w->shapes.p, w->tflags.p, w->trank.p, cnt + C_NMOV, T,
                                          w->tbody.p, w->tile_of.p, w->tslot.p);
    } else {
        k_set_int<<<1, 1, 0, st>>>(cnt + C_NMOV, 0);
    }
    int ga = grid_for(std::max(amax_, 1));
    jh_arbiter* arb = w->arb[w->cur].p;
    k_tile_hist<<<ga, 256, 0, st>>>(order, w->coffs.p, arb, w->colour.p, smotion, w->tile_of.p, w->thist.p,
                                    w->tshared.p);
    k_set_int<<<1, 1, 0, st>>>(cnt + C_TCELLS, int(cells));
    scan::exclusive_scan(w->thist.p, w->toff.p, cnt + C_TCELLS, int(cells), w->tsums.p, nullptr, st);
    k_tile_scatter<<<ga, 256, 0, st>>>(order, w->coffs.p, arb, w->colour.p, smotion, w->tile_of.p, w->toff.p,
                                       w->tcur.p, w->tlist.p);
    k_tile_serial<<<1, 64, 0, st>>>(order, w->coffs.p, arb, smotion, w->tile_of.p, w->toff.p, w->tcur.p,
                                    w->tlist.p);
    (void)a_cur;
    HIPCHK(hipGetLastError());
    return JH_OK;
}

// Residual probes and traces of the first substep (see jh_world_step_probe).  S describes the solve
// about to run; every probe works on scratch copies of the velocities and the arbiters.
jh_status run_probe(jh_world* w, const SolveArgs& S, const int* a_count, int amax_, jh_probe* pr) {
    hipStream_t st = w->st;
    Bodies B = w->bodies();
    int K = pr->iterations;
    HIPCHK(hipMemcpyAsync(w->hcnt + C_COUNT - 1, a_count, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    int A = w->hcnt[C_COUNT - 1];
    int grid = std::min(w->coop_probe, grid_for(std::max(std::max(amax_, w->nb), 256), 256, 1 << 20));
    HIPCHK(w->pvel.ensure(std::max(w->nb, 1)));
    HIPCHK(w->pang.ensure(std::max(w->nb, 1)));
    HIPCHK(w->parb.ensure(std::max(A, 1)));
    HIPCHK(w->ppart.ensure(size_t(K + 1) * grid));
    HIPCHK(w->plam.ensure(4 * size_t(std::max(A, 1))));
    bool trace = pr->capacity > 0;
    int ga = grid_for(std::max(4 * A, 1));
    if (trace) {
        HIPCHK(w->pfeat.ensure(size_t(std::max(4 * A, 1)) * lws::kIn));
        HIPCHK(w->ptarg.ensure(size_t(std::max(4 * A, 1)) * 3));
        HIPCHK(w->pvalid.ensure(std::max(4 * A, 1)));
        lws::k_features<<<ga, 256, 0, st>>>(S.arb, a_count, S.B, w->pfeat.p, w->pvalid.p);
    }
    std::vector<double> part(size_t(K + 1) * grid);
    for (int m = 0; m < 3; ++m) {
        for (int k = 0; k <= JH_PROBE_MAX_ITERATIONS; ++k) pr->residual[m][k] = 0.0;
        if (!(pr->modes & (1 << m))) continue;
        if (w->nb > 0) {
            HIPCHK(hipMemcpyAsync(w->pvel.p, w->vel.p, size_t(w->nb) * sizeof(float4), hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpyAsync(w->pang.p, w->angvel.p, size_t(w->nb) * sizeof(float4), hipMemcpyDeviceToDevice, st));
        }
        if (A > 0)
            HIPCHK(hipMemcpyAsync(w->parb.p, S.arb, size_t(A) * sizeof(jh_arbiter), hipMemcpyDeviceToDevice, st));
        SolveArgs P = S;
        P.arb = w->parb.p;
        P.B.vel = w->pvel.p;
        P.B.angvel = w->pang.p;
        if (m == 2) lws::k_zero_acc<<<ga, 256, 0, st>>>(w->parb.p, a_count);
        if (m == 1) {
            lws::k_infer<<<grid_for(std::max(4 * A, 1), 256, 2048), 256, 0, st>>>(w->parb.p, a_count, P.B,
                                                                                 w->lws.model());
            if (trace && pr->learned) {
                HIPCHK(w->plearn.ensure(size_t(std::max(4 * A, 1)) * 3));
                lws::k_gather_acc<<<ga, 256, 0, st>>>(w->parb.p, a_count, w->plearn.p);
            }
        }
        HIPCHK(hipGetLastError());
        void* args[] = {&P, &K, &w->ppart.p, &w->plam.p};
        HIPCHK(hipLaunchCooperativeKernel((const void*)lws::k_probe, dim3(grid), dim3(256), args, 0, st));
        if (m == 0 && trace) lws::k_gather_acc<<<ga, 256, 0, st>>>(w->parb.p, a_count, w->ptarg.p);
        HIPCHK(hipMemcpyAsync(part.data(), w->ppart.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int k = 0; k <= K; ++k) {
            double sum = 0.0;
            for (int b = 0; b < grid; ++b) sum += part[size_t(k) * grid + b];
            pr->residual[m][k] = sum;
        }
    }
    pr->points = 0;
    pr->rows = 0;
    if (A > 0) {
        std::vector<uint8_t> valid(4 * size_t(A));
        std::vector<float> feat, targ;
        if (trace) {
            feat.resize(4 * size_t(A) * lws::kIn);
            targ.resize(4 * size_t(A) * 3);
            bool learned = pr->learned && (pr->modes & 2);
            std::vector<float> lrn(learned ? 4 * size_t(A) * 3 : 0);
            if (learned)
                HIPCHK(hipMemcpyAsync(lrn.data(), w->plearn.p, lrn.size() * sizeof(float), hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(valid.data(), w->pvalid.p, valid.size(), hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(feat.data(), w->pfeat.p, feat.size() * sizeof(float), hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(targ.data(), w->ptarg.p, targ.size() * sizeof(float), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            int64_t rows = 0;
            for (size_t i = 0; i < valid.size(); ++i) {
                if (!valid[i]) continue;
                if (rows < pr->capacity) {
                    memcpy(pr->features + rows * lws::kIn, &feat[i * lws::kIn], lws::kIn * sizeof(float));
                    memcpy(pr->targets + rows * 3, &targ[i * 3], 3 * sizeof(float));
                    if (learned) memcpy(pr->learned + rows * 3, &lrn[i * 3], 3 * sizeof(float));
                }
                ++rows;
            }
            pr->points = rows;
            pr->rows = std::min<int64_t>(rows, pr->capacity);
        } else {
            // count of contact points from the arbiter usage masks
            std::vector<jh_arbiter> a(A);
            HIPCHK(hipMemcpyAsync(a.data(), S.arb, size_t(A) * sizeof(jh_arbiter), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            for (auto& x : a) pr->points += __builtin_popcount(x.usage_mask & 15u);
        }
    }
    return JH_OK;
}

// Graph colouring of the arbiters in arb[cur] (edges between non-static bodies) and the colour-sorted
// solve order (*order_out, device keys (colour << 34 | contacts << 32) | arbiter); colour count, chunk
// count and the contact points to solve land in the device counters.
jh_status colour_arbiters(jh_world* w, int cur, int* a_cur, int amax_, const uint64_t** order_out) {
    hipStream_t st = w->st;
    int* cnt = w->cnt.p;
    ColourArgs P;
    P.arb = w->arb[cur].p;
    P.a_count = a_cur;
    P.motion = w->islands_on() ? w->smotion.p : w->motion.p;
    P.colour = w->colour.p;
    P.body_max = w->bmax.p;
    P.nb = int(w->bmax.cap / 2);
    P.mask_lo = w->mlo.p;
    P.mask_hi = w->mhi.p;
    P.work = w->cwork.p;
    P.wcount = cnt + C_REM;
    P.ncolours = cnt + C_NCOL;
    P.hist = w->chist.p;
    P.bar = w->bar.p;
    P.spin = w->spin();
    // colours persist across steps; recolour everything after body changes or when the palette
    // has spread (first-fit over many steps can leave more colours than a fresh colouring)
    P.keep = (w->recolour || w->solve_batches > 48) ? 0 : 1;
    w->recolour = false;
    int grid = std::min(w->colour_blocks, grid_for(std::max(amax_, 128), 256, 1 << 20));
    void* args[] = {&P};
    jh_status s = launch_coop((const void*)k_colour, grid, 256, args, st);
    if (s != JH_OK) return s;
    k_colour_keys<<<grid_for(std::max(amax_, 1)), 256, 0, st>>>(w->colour.p, w->arb[cur].p, a_cur, w->ok0.p,
                                                                 cnt + C_CONTACTS);
    int csh[2] = {32, 40};  // (colour, contacts) in bits 32..41
    bool in1 = rsort::sort_pairs(w->ok0.p, w->ok1.p, nullptr, nullptr, a_cur, std::max(amax_, 1), csh, 2,
                                 w->hist.p, st);
    k_colour_offsets<<<1, 1, 0, st>>>(w->chist.p, w->coffs.p);
    const uint64_t* order = in1 ? w->ok1.p : w->ok0.p;
    k_solve_keys<<<grid_for(std::max(amax_, 1)), 256, 0, st>>>(order, w->akey[cur].p, a_cur, w->solve_keys.p);
    HIPCHK(hipGetLastError());
    *order_out = order;
    return JH_OK;
}

// World.CheckDeactivation on the components of the contact and constraint graph (jh_islands.h).
// In a step (stabilize false): after the contact cache, arb[cur] holds this step's arbiters, evb the
// keys of those Detect created, eve the broken ones RemoveBrokenArbiters frees in this step (edges of
// this check's graph until then).  For World.Stabilize: arb[cur] and the broken arbiters waiting in
// zkey, no new ones.
jh_status check_islands(jh_world* w, int cur, int* a_cur, int amax_, bool stabilize) {
    hipStream_t st = w->st;
    int* cnt = w->cnt.p;
    int nb = w->nb;
    if (nb == 0) return JH_OK;
    int gb = grid_for(nb, 256, 1 << 20);
    int ga = grid_for(std::max(amax_, 1));
    const int* motion = w->motion.p;
    const jh_arbiter* arb = w->arb[cur].p;
    int ncon = int(w->host_con.size());
    isl::k_reset<<<gb, 256, 0, st>>>(nb, w->uf.p, w->hfl.p, w->gfl.p);
    if (!stabilize) {
        k_fill_u32<<<ga, 256, 0, st>>>(w->anew.p, size_t(std::max(amax_, 1)), 0u);
        isl::k_mark_new<<<ga, 256, 0, st>>>(w->evb.p, cnt + C_EVB, w->akey[cur].p, a_cur, w->anew.p);
    }
    // H: the graph before this step's new arbiters
    isl::k_union_arbiters<<<ga, 256, 0, st>>>(arb, a_cur, stabilize ? nullptr : w->anew.p, 0u, motion, w->uf.p);
    const uint64_t* ekeys = stabilize ? w->zkey.p : w->eve.p;
    const int* ecount = cnt + (stabilize ? C_ZOMB : C_EVE);
    isl::k_union_keys<<<grid_for(std::max(int(w->zkey.cap), 1)), 256, 0, st>>>(ekeys, ecount, w->shapes.p, motion,
                                                                              w->uf.p);
    if (ncon > 0) isl::k_union_constraints<<<grid_for(ncon), 256, 0, st>>>(w->con.p, ncon, motion, w->uf.p);
    isl::k_roots<<<gb, 256, 0, st>>>(nb, w->uf.p, w->hroot.p);
    isl::k_hflags<<<gb, 256, 0, st>>>(nb, motion, w->sflags.p, w->hroot.p, w->hfl.p);
    // G: with the new arbiters
    const int* groot = w->hroot.p;
    if (!stabilize) {
        isl::k_union_arbiters<<<ga, 256, 0, st>>>(arb, a_cur, w->anew.p, 1u, motion, w->uf.p);
        isl::k_roots<<<gb, 256, 0, st>>>(nb, w->uf.p, w->groot.p);
        groot = w->groot.p;
    }
    isl::k_h_to_g<<<gb, 256, 0, st>>>(nb, motion, w->hroot.p, groot, w->hfl.p, w->gfl.p);
    if (!stabilize)
        isl::k_new_edges<<<ga, 256, 0, st>>>(arb, a_cur, w->anew.p, motion, w->sflags.p, w->hroot.p, w->hfl.p, groot,
                                             w->gfl.p);
    k_set_int<<<1, 1, 0, st>>>(cnt + C_ASLEEP, 0);
    isl::k_decide<<<gb, 256, 0, st>>>(nb, motion, groot, w->gfl.p, w->sflags.p, w->slp.p,
                                      w->cfg.allow_deactivation, w->smotion.p, cnt + C_ASLEEP);
    if (!stabilize) {
        // this step's RemoveBrokenArbiters, seen by the next check
        isl::k_e_marks<<<grid_for(std::max(int(w->zkey.cap), 1)), 256, 0, st>>>(w->eve.p, cnt + C_EVE, w->shapes.p,
                                                                               motion, groot, w->gfl.p, w->sflags.p);
        isl::k_e_apply<<<gb, 256, 0, st>>>(nb, motion, groot, w->gfl.p, w->sflags.p);
    }
    HIPCHK(hipGetLastError());
    w->isl_dirty = true;
    return JH_OK;
}

jh_status do_step(jh_world* w, float dt, jh_probe* probe) {
    hipStream_t st = w->st;
    Bodies B = w->bodies();
    int* cnt = w->cnt.p;
    const jh_config& cfg = w->cfg;
    // TimeStep (TimeStep.cs:36-46)
    StepTime ts;
    ts.dt = dt;
    ts.substeps = cfg.substep_count;
    ts.substep_dt = dt / float(cfg.substep_count);
    ts.inv_dt = dt > 0.0f ? 1.0f / dt : 0.0f;
    ts.inv_substep_dt = ts.substep_dt > 0.0f ? 1.0f / ts.substep_dt : 0.0f;

    HIPCHK(hipEventRecord(w->ev[0], st));
    jh_status s = broadphase(w);
    if (s != JH_OK) return s;
    HIPCHK(hipEventRecord(w->ev[1], st));
    s = ensure_step_buffers(w);
    if (s != JH_OK) return s;
    int np = w->n_pairs;
    const uint64_t* pairs = w->pairs_in_k1 ? w->pk1.p : w->pk0.p;
    int aux = cfg.enable_auxiliary_contact_points;
    // narrowphase
    if (np > 0) {
        int spec = w->n_spec > 0;
        if (spec) HIPCHK(w->specq.ensure(size_t(np)));
        float thr = cfg.speculative_velocity_threshold;
        bool hulls = w->has_clouds && w->narrow_lds;
        if (hulls) HIPCHK(w->hullq.ensure(size_t(np)));
        const int* perm = nullptr;
        if (w->narrow_sort) {  // pairs grouped by their shapes' types (one launch per pair over np)
            HIPCHK(w->pperm.ensure(size_t(np)));
            if (!w->pcls.p) {
                HIPCHK(w->pcls.ensure(2 * kPairClasses));
                HIPCHK(hipMemsetAsync(w->pcls.p, 0, 2 * kPairClasses * sizeof(int), st));
            }
            int gp = (np + 255) / 256;
            k_class_count<<<gp, 256, 0, st>>>(pairs, cnt + C_PAIRS, w->shapes.p, w->pcls.p);
            k_class_scan<<<1, 64, 0, st>>>(w->pcls.p, w->pcls.p + kPairClasses);
            k_class_scatter<<<gp, 256, 0, st>>>(pairs, cnt + C_PAIRS, w->shapes.p, w->pcls.p + kPairClasses, w->pperm.p);
            perm = w->pperm.p;
        }
        k_narrow<<<grid_for(np, 128, 1 << 20), 128, 0, st>>>(pairs, cnt + C_PAIRS, w->shapes.p, w->verts.p, B, aux,
                                                             cfg.narrowphase_filter, w->man.p, w->deep.p,
                                                             cnt + C_DEEP, spec, thr * thr, w->specq.p,
                                                             cnt + C_SPEC, hulls ? w->hullq.p : nullptr,
                                                             cnt + C_HULL, perm);
        if (hulls)  // the pairs with point clouds, their vertices staged in LDS
            k_narrow_hulls<<<grid_for(np, 64, 1 << 20), 64, 0, st>>>(pairs, w->hullq.p, cnt + C_HULL, w->shapes.p,
                                                                      w->verts.p, B, aux, cfg.narrowphase_filter,
                                                                      w->man.p, w->deep.p, cnt + C_DEEP, spec,
                                                                      thr * thr, w->specq.p, cnt + C_SPEC);
        if (spec)
            k_sweep<<<grid_for(np, 64, 4096), 64, 0, st>>>(pairs, w->shapes.p, w->verts.p, B, cfg.narrowphase_filter,
                                                           dt, cfg.speculative_relaxation_factor, w->specq.p,
                                                           cnt + C_SPEC, w->man.p);
        HIPCHK(hipEventRecord(w->ev[7], st));
        if (w->epa_wave) {
            int blocks = std::max(1, std::min(w->epa_wave_blocks, np));
            HIPCHK(w->epa_side.ensure(size_t(w->epa_wave_blocks) * epw::kSideBytes));
            HIPCHK(w->epa_res.ensure(std::max(np, 1)));
            k_epa_wave<<<blocks, 64, 0, st>>>(pairs, w->shapes.p, w->verts.p, B, w->deep.p, cnt + C_DEEP,
                                              w->epa_side.p, w->epa_res.p, cnt + C_EPAQ);
            k_epa_finish<<<grid_for(np, 128, 1 << 20), 128, 0, st>>>(pairs, w->shapes.p, w->verts.p, B, aux, w->man.p,
                                                                  w->deep.p, cnt + C_DEEP, w->epa_res.p);
        } else {
            int et = std::min(kEpaMaxThreads, ((np + 63) / 64) * 64);
            if (et > w->epa_threads) {
                HIPCHK(w->epa_work.ensure(size_t(et) * epa::kBytesPerThread));
                w->epa_threads = et;
            }
            k_epa<<<w->epa_threads / 64, 64, 0, st>>>(pairs, w->shapes.p, w->verts.p, B, aux, w->man.p, w->deep.p,
                                                      cnt + C_DEEP, w->epa_work.p);
        }
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(w->ev[2], st));
    // contact cache: colliding pairs -> existing/new arbiters -> merged key-sorted cache -> AddContact
    int cur = w->cur, nxt = 1 - w->cur;
    int* a_cur = cnt + (cur == 0 ? C_A0 : C_A1);
    int* a_nxt = cnt + (nxt == 0 ? C_A0 : C_A1);
    if (np > 0) {
        int g = grid_for(np);
        k_collide_flags<<<g, 256, 0, st>>>(w->man.p, cnt + C_PAIRS, w->flags.p);
        scan::exclusive_scan(w->flags.p, w->offs.p, cnt + C_PAIRS, np, w->sums.p, cnt + C_COLL, st);
        k_collide_compact<<<g, 256, 0, st>>>(pairs, cnt + C_PAIRS, w->flags.p, w->offs.p, w->akey[cur].p, a_cur,
                                             w->ckey.p, w->cpair.p, w->cfound.p, w->cnew.p);
        scan::exclusive_scan(w->cnew.p, w->noffs.p, cnt + C_COLL, np, w->sums.p, cnt + C_NEW, st);
        k_new_compact<<<g, 256, 0, st>>>(w->ckey.p, cnt + C_COLL, w->cnew.p, w->noffs.p, w->nkey.p);
    }
    k_events<<<grid_for(std::max(np, w->a_host), 256, 256), 256, 0, st>>>(
        w->nkey.p, cnt + C_NEW, w->zkey.p, cnt + C_ZOMB, w->ckey.p, cnt + C_COLL, w->evb.p, cnt + C_EVB, w->eve.p,
        cnt + C_EVE);
    int amax_ = w->a_host + np;
    k_merge_existing<<<grid_for(16 * std::max(w->a_host, 1)), 256, 0, st>>>(w->arb[cur].p, w->akey[cur].p, a_cur, w->nkey.p,
                                                                       cnt + C_NEW, w->arb[nxt].p, w->akey[nxt].p);
    k_merge_new<<<grid_for(std::max(np, 1)), 256, 0, st>>>(w->akey[cur].p, a_cur, w->nkey.p, cnt + C_NEW, w->shapes.p, B,
                                                           cfg.speculative_relaxation_factor, w->arb[nxt].p,
                                                           w->akey[nxt].p, a_nxt);
    if (np > 0)
        k_register<<<grid_for(np), 256, 0, st>>>(w->ckey.p, w->cpair.p, w->cfound.p, w->noffs.p, cnt + C_COLL,
                                                 w->akey[cur].p, a_cur, w->nkey.p, cnt + C_NEW, w->man.p, B,
                                                 cfg.persistent_contact_manifold, w->arb[nxt].p);
    HIPCHK(hipGetLastError());
    cur = nxt;
    w->cur = cur;
    a_cur = a_nxt;
    if (w->islands_on()) {
        s = check_islands(w, cur, a_cur, amax_, false);
        if (s != JH_OK) return s;
    } else if (w->isl_dirty) {
        isl::k_clear_bits<<<grid_for(std::max(w->nb, 1), 256, 1 << 20), 256, 0, st>>>(
            w->nb, w->sflags.p, isl::kTouch | isl::kM1 | isl::kN1 | isl::kBnew | isl::kSwake);
        w->isl_dirty = false;
    }
    HIPCHK(hipEventRecord(w->ev[3], st));
    // colouring of the arbiters (graph edges between dynamic bodies)
    // the tile solver needs its tiles to fit; otherwise this step runs the dataflow solver
    size_t tile_lds = w->kind == SOLVER_TILES ? tile_lds_bytes(w, w->tiles) : 0;
    int kind = (w->kind == SOLVER_TILES && tile_lds == 0) ? SOLVER_FLOW : w->kind;
    if (!w->host_con.empty()) {
        if (w->con_dirty) {
            s = colour_constraints(w);
            if (s != JH_OK) return s;
        }
        if (constraint_batches(w) > 0) kind = SOLVER_BARRIER;  // constraints run in the barrier solver
    }
    bool use_tiles = kind == SOLVER_TILES;
    {
        const uint64_t* order = nullptr;
        s = colour_arbiters(w, cur, a_cur, amax_, &order);
        if (s != JH_OK) return s;
        if (kind == SOLVER_FLOW || kind == SOLVER_TILES || kind == SOLVER_LANES) {
            // per-body update sequence: rank of each arbiter visit and body degree, from the
            // colouring's per-body colour masks (the serial batch ranked by one workgroup)
            int ga = grid_for(std::max(amax_, 1));
            const int* smo = w->solver_bodies().motion;
            k_rank_clear<<<ga, 256, 0, st>>>(w->arb[cur].p, order, w->coffs.p, w->scnt.p);
            k_rank_serial<<<1, 256, 0, st>>>(w->arb[cur].p, order, w->coffs.p, smo, w->rank.p, w->scnt.p);
            k_rank_masks<<<ga, 256, 0, st>>>(w->arb[cur].p, order, w->coffs.p, smo, w->mlo.p, w->mhi.p, w->scnt.p,
                                             w->rank.p, w->deg.p, cnt + C_MAXDEG);
            if (use_tiles) {
                s = tile_setup(w, order, a_cur, amax_);
                if (s != JH_OK) return s;
            }
        }
        HIPCHK(hipEventRecord(w->ev[4], st));
        // persistent solver over all substeps
        SolveArgs S;
        S.arb = w->arb[cur].p;
        S.order = order;
        S.offs = w->coffs.p;
        S.ncolours = cnt + C_NCOL;
        S.rank = w->rank.p;
        S.deg = w->deg.p;
        S.bar = w->bar.p;
        S.spin = w->spin();
        S.B = w->solver_bodies();
        S.nb = w->nb;
        S.ts = ts;
        S.iterations = cfg.solver_iterations;
        S.relaxations = cfg.relaxation_iterations;
        S.soa.bodies = w->soa_bodies.p;
        S.soa.um = w->soa_um.p;
        S.soa.fs = w->soa_fs.p;
        S.soa.flag = w->soa_flag.p;
        S.soa.f = w->soa_f.p;
        S.soa.cap = w->soa_cap;
        S.vw = w->vwpack.p;
        HIPCHK(w->fqueue.ensure(2 * size_t(cfg.substep_count)));
        S.queue = w->fqueue.p;
        S.qmode = w->flow_qmode;
        S.soa.mv = kind == SOLVER_FLOW ? w->soa_mv.p : nullptr;
        S.soa.nd = kind == SOLVER_FLOW ? w->soa_nd.p : nullptr;
        S.con.c = w->con.p;
        S.con.work = w->con_work.p;
        S.con.order = w->con_order.p;
        S.con.offs = w->con_offs.p;
        S.con.nbatch = kind == SOLVER_BARRIER ? constraint_batches(w) : 0;
        S.stabilize = 0;
        if (probe) {
            s = run_probe(w, S, a_cur, amax_, probe);
            if (s != JH_OK) return s;
        }
        if (cfg.warm_start == JH_WARM_START_LEARNED) {
            lws::k_infer<<<grid_for(std::max(4 * amax_, 1), 256, 2048), 256, 0, st>>>(S.arb, a_cur, S.B,
                                                                                     w->lws.model());
            HIPCHK(hipGetLastError());
        }
        void* sargs[] = {&S};
        if (kind == SOLVER_WORLDS) {
            // world-ordered arbiter list: (world, colour) histogram -> offsets -> scatter
            int gsoa = grid_for(std::max(amax_, 1));
            size_t cells = size_t(w->n_worlds) * 128 + 1;
            HIPCHK(w->wlist.ensure(std::max(amax_, 1)));
            k_fill_u32<<<grid_for(int(std::min<size_t>(cells, 1 << 30))), 256, 0, st>>>(w->whist.p, cells, 0u);
            k_fill_u32<<<grid_for(int(std::min<size_t>(cells, 1 << 30))), 256, 0, st>>>(w->wcur.p, cells, 0u);
            k_world_hist<<<gsoa, 256, 0, st>>>(order, w->coffs.p, S.arb, w->colour.p, w->world.p, w->whist.p);
            scan::exclusive_scan(w->whist.p, w->woff.p, w->wcells.p, int(cells), w->wsums.p, nullptr, st);
            k_world_scatter<<<gsoa, 256, 0, st>>>(order, w->coffs.p, S.arb, w->colour.p, w->world.p, w->woff.p,
                                                  w->wcur.p, w->wlist.p);
            k_world_serial<<<1, 64, 0, st>>>(order, w->coffs.p, S.arb, w->world.p, w->woff.p, w->wcur.p, w->wlist.p);
            k_soa_gather<<<gsoa, 256, 0, st>>>(S.arb, w->wlist.p, w->coffs.p, S.soa);
            k_set_int<<<1, 1, 0, st>>>(w->wqueue.p, 0);
            WorldArgs Wd;
            Wd.wbody = w->wbody.p;
            Wd.wboff = w->wboff.p;
            Wd.wsched = w->wsched.p;
            Wd.woff = w->woff.p;
            Wd.W = w->n_worlds;
            Wd.queue = w->wqueue.p;
            HIPCHK(hipEventRecord(w->ev[8], st));
            k_solve_worlds<<<std::max(1, std::min(w->world_blocks, w->n_worlds)), w->world_threads, 0, st>>>(S, Wd);
            HIPCHK(hipEventRecord(w->ev[9], st));
            k_soa_scatter<<<gsoa, 256, 0, st>>>(S.arb, w->wlist.p, w->coffs.p, S.soa);
            HIPCHK(hipGetLastError());
        } else if (use_tiles) {
            int gsoa = grid_for(std::max(amax_, 1));
            k_soa_gather<<<gsoa, 256, 0, st>>>(S.arb, w->tlist.p, w->coffs.p, S.soa);
            TileArgs Tl;
            Tl.tbody = w->tbody.p;
            Tl.nmov = cnt + C_NMOV;
            Tl.tile_of = w->tile_of.p;
            Tl.tslot = w->tslot.p;
            Tl.shared = w->tshared.p;
            Tl.toff = w->toff.p;
            Tl.tlist = w->tlist.p;
            Tl.T = w->tiles;
            void* targs[] = {&S, &Tl};
            HIPCHK(hipEventRecord(w->ev[8], st));
            HIPCHK(hipLaunchCooperativeKernel((const void*)k_solve_tiles, dim3(w->tiles), dim3(256), targs,
                                              unsigned(tile_lds), st));
            HIPCHK(hipEventRecord(w->ev[9], st));
            k_soa_scatter<<<gsoa, 256, 0, st>>>(S.arb, w->tlist.p, w->coffs.p, S.soa);
            HIPCHK(hipGetLastError());
        } else if (kind == SOLVER_LANES) {
            int gsoa = grid_for(std::max(amax_, 1));
            k_rec_gather<<<gsoa, 256, 0, st>>>(S.arb, order, w->coffs.p, w->rank.p, w->deg.p, w->motion.p, w->rec.p);
            void* largs[] = {&S, &w->rec.p};
            HIPCHK(hipEventRecord(w->ev[8], st));
            s = launch_coop((const void*)k_solve_lanes, w->lane_blocks, 256, largs, st);
            if (s != JH_OK) return s;
            HIPCHK(hipEventRecord(w->ev[9], st));
            k_rec_scatter<<<gsoa, 256, 0, st>>>(S.arb, order, w->coffs.p, w->rec.p);
        } else if (kind == SOLVER_FLOW) {
            int gsoa = grid_for(std::max(amax_, 1));
            k_soa_gather<<<gsoa, 256, 0, st>>>(S.arb, order, w->coffs.p, S.soa, S.rank, S.deg, S.B.motion);
            HIPCHK(hipEventRecord(w->ev[8], st));
            s = launch_coop((const void*)k_solve_flow, w->flow_blocks, 256, sargs, st);
            if (s != JH_OK) return s;
            HIPCHK(hipEventRecord(w->ev[9], st));
            k_soa_scatter<<<gsoa, 256, 0, st>>>(S.arb, order, w->coffs.p, S.soa);
        } else {
            int sgrid = std::min(w->coop_solve, grid_for(std::max(std::max(std::max(amax_, w->nb), 256),
                                                                  int(w->con_order_h.size())), 256, 1 << 20));
            HIPCHK(hipEventRecord(w->ev[8], st));
            s = launch_coop((const void*)k_solve, sgrid, 256, sargs, st);
            if (s != JH_OK) return s;
            HIPCHK(hipEventRecord(w->ev[9], st));
        }
        if (s != JH_OK) return s;
    }
    HIPCHK(hipEventRecord(w->ev[5], st));
    // UpdateContacts + removal of arbiters without contacts, then RigidBody.Update
    {
        int g = grid_for(std::max(amax_, 1));
        k_update_contacts<<<g, 256, 0, st>>>(w->arb[cur].p, a_cur, B, w->flags.p, w->solver_bodies().motion);
        scan::exclusive_scan(w->flags.p, w->offs.p, a_cur, std::max(amax_, 1), w->sums.p,
                             cnt + (cur == 0 ? C_A1 : C_A0), st);
        k_compact_arbiters<<<grid_for(16 * std::max(amax_, 1)), 256, 0, st>>>(
            w->arb[cur].p, w->akey[cur].p, a_cur, w->flags.p, w->offs.p,
                                              w->arb[1 - cur].p, w->akey[1 - cur].p, w->zkey.p, cnt + C_ZOMB);
        w->cur = 1 - cur;
        if (w->nb > 0)
            k_update_bodies<<<grid_for(w->nb, 256, 1 << 20), 256, 0, st>>>(B, w->nb, ts, cfg.gravity[0], cfg.gravity[1],
                                                                         cfg.gravity[2]);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(w->ev[6], st));
    HIPCHK(hipMemcpyAsync(w->hcnt, cnt, C_COUNT * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(w->habort, w->bar.p + kBarAbort, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    // DynamicTree.Update(multiThread, StepDt) (World.Step.cs:274-276)
    w->broadphase_dt = dt;
    if (*w->habort) {
        // every persistent kernel of the step has drained; their barrier counts are left mid-way
        HIPCHK(hipMemsetAsync(w->bar.p, 0, 64 * sizeof(unsigned), st));
        HIPCHK(hipStreamSynchronize(st));
        *w->habort = 0;
        w->have_step = false;
        w->recolour = true;
        return fail(JH_ERR_TIMEOUT,
                    "a persistent kernel's wait gave up after %u polls or the step was cancelled; the world's device "
                    "state is unspecified until bodies and arbiters are set again",
                    w->spin());
    }
    // solve-order bookkeeping and stats
    if (w->islands_on()) w->n_asleep = w->hcnt[C_ASLEEP];
    w->a_host = w->hcnt[w->cur == 0 ? C_A0 : C_A1];
    w->solve_count = w->hcnt[cur == 0 ? C_A0 : C_A1];
    w->solve_batches = w->hcnt[C_NCOL];
    jh_stats& S = w->stats;
    S.bodies = w->nb;
    S.shapes = w->ns;
    S.pairs = np;
    S.colliding_pairs = w->hcnt[C_COLL];
    S.arbiters = w->a_host;
    S.arbiters_solved = w->solve_count;
    S.contacts = w->hcnt[C_CONTACTS];
    S.colours = w->hcnt[C_NCOL];
    S.epa_calls = w->hcnt[C_DEEP];
    S.max_degree = w->hcnt[C_MAXDEG];
    w->ev_begin = w->hcnt[C_EVB];
    w->ev_end = w->hcnt[C_EVE];
    S.solver = kind;
    S.bodies_asleep = w->n_asleep;
    S.steps++;
    float ms[7];
    for (int k = 0; k < 6; ++k) HIPCHK(hipEventElapsedTime(&ms[k], w->ev[k], w->ev[k + 1]));
    S.ms_broadphase = ms[0];
    S.ms_narrowphase = ms[1];
    S.ms_contact_cache = ms[2];
    S.ms_colouring = ms[3];
    S.ms_solve = ms[4];
    S.ms_update = ms[5];
    float epa_ms = 0;
    if (np > 0) HIPCHK(hipEventElapsedTime(&epa_ms, w->ev[7], w->ev[2]));
    S.ms_epa = np > 0 ? epa_ms : 0.0;
    float sk_ms;
    HIPCHK(hipEventElapsedTime(&sk_ms, w->ev[8], w->ev[9]));
    S.ms_solve_kernel = sk_ms;
    float tot;
    HIPCHK(hipEventElapsedTime(&tot, w->ev[0], w->ev[6]));
    S.ms_step = tot;
    w->have_step = true;
    return JH_OK;
}

}  // namespace

extern "C" {

const char* jh_last_error(void) { return g_err; }
int32_t jh_version(void) { return 1; }

// Sizes of the ABI structs (checked by bindings): 0 config, 1 body, 2 shape, 3 contact, 4 arbiter,
// 5 manifold, 6 stats.
int32_t jh_struct_size(int32_t which) {
    switch (which) {
        case 0: return sizeof(jh_config);
        case 1: return sizeof(jh_body);
        case 2: return sizeof(jh_shape);
        case 3: return sizeof(jh_contact);
        case 4: return sizeof(jh_arbiter);
        case 5: return sizeof(jh_manifold);
        case 6: return sizeof(jh_stats);
        case 7: return sizeof(jh_probe);
        case 8: return sizeof(jh_tuning);
        case 9: return sizeof(jh_constraint);
        default: return -1;
    }
}

jh_status jh_config_defaults(jh_config* c) {
    if (!c) return fail(JH_ERR_INVALID_ARG, "config is NULL");
    memset(c, 0, sizeof(*c));
    c->gravity[1] = -9.81f;
    c->substep_count = 1;
    c->solver_iterations = 6;
    c->relaxation_iterations = 4;
    c->speculative_relaxation_factor = 0.9f;
    c->enable_auxiliary_contact_points = 1;
    c->persistent_contact_manifold = 1;
    c->speculative_velocity_threshold = 10.0f;
    c->allow_deactivation = 1;
    return JH_OK;
}

jh_status jh_world_create(const jh_config* config, jh_world** out) {
    if (!out) return fail(JH_ERR_INVALID_ARG, "out_world is NULL");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(JH_ERR_NO_DEVICE, "no HIP device visible");
    jh_config c;
    if (config) c = *config;
    else jh_config_defaults(&c);
    jh_status s = check_config(&c);
    if (s != JH_OK) return s;
    jh_world* w = new jh_world();
    w->cfg = c;
    memset(&w->stats, 0, sizeof(w->stats));
    HIPCHK(hipGetDevice(&w->device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, w->device));
    w->cu_count = prop.multiProcessorCount;
    HIPCHK(hipStreamCreateWithFlags(&w->st, hipStreamNonBlocking));
    for (auto& e : w->ev) HIPCHK(hipEventCreate(&e));
    HIPCHK(hipHostMalloc((void**)&w->hcnt, (C_COUNT + 1) * sizeof(int), hipHostMallocDefault));
    w->habort = reinterpret_cast<unsigned*>(w->hcnt + C_COUNT);
    *w->habort = 0;
    HIPCHK(w->cnt.ensure(C_COUNT));
    HIPCHK(hipMemset(w->cnt.p, 0, C_COUNT * sizeof(int)));
    HIPCHK(w->bounds.ensure(8));
    HIPCHK(w->chist.ensure(128));
    HIPCHK(w->coffs.ensure(129));
    int occ = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_colour, 256, 0));
    w->coop_colour = std::max(1, occ) * w->cu_count;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_solve, 256, 0));
    w->coop_solve = std::max(1, occ) * w->cu_count;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_solve_worlds, 512, 0));
    w->world_blocks = std::max(1, occ) * w->cu_count;
    for (int k = 0; k < 4; ++k) {
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_solve_worlds, 64 << k, 0));
        w->world_occ[k] = std::max(1, occ);
    }
    if (const char* e = getenv("JH_WORLD_THREADS")) {
        int t = atoi(e);
        for (int k = 0; k < 4; ++k)
            if (t == (64 << k)) {
                w->world_threads = t;
                w->world_blocks = w->world_occ[k] * w->cu_count;
            }
    }
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_epa_wave, 64, 0));
    w->epa_wave_blocks = std::max(1, occ) * w->cu_count;
    if (const char* e = getenv("JH_EPA")) w->epa_wave = strcmp(e, "thread") != 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)lws::k_probe, 256, 0));
    w->coop_probe = std::max(1, occ) * w->cu_count;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_solve_flow, 256, 0));
    w->flow_max = std::max(1, occ) * w->cu_count;
    if (getenv("JH_AUTOTUNE_VERBOSE")) {
        hipFuncAttributes fa;
        if (hipFuncGetAttributes(&fa, (const void*)k_solve_flow) == hipSuccess)
            fprintf(stderr, "k_solve_flow: %d registers, %zu B static LDS, %zu B scratch, %d max threads; %d blocks per CU\n",
                    fa.numRegs, fa.sharedSizeBytes, fa.localSizeBytes, fa.maxThreadsPerBlock, occ);
    }
    w->flow_blocks = std::min(w->flow_max, w->cu_count);
    HIPCHK(w->bar.ensure(64));
    HIPCHK(hipMemset(w->bar.p, 0, 64 * sizeof(unsigned)));
    // launch-shape overrides for tuning sweeps (clamped to the co-resident maximum)
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_solve_lanes, 256, 0));
    w->lane_max = std::max(1, occ) * w->cu_count;
    w->lane_blocks = std::min(w->lane_max, w->cu_count);
    if (const char* e = getenv("JH_SOLVE_BLOCKS")) {
        w->coop_solve = std::max(1, std::min(atoi(e), w->coop_solve));
        w->flow_blocks = std::max(1, std::min(atoi(e), w->flow_max));
        w->lane_blocks = std::max(1, std::min(atoi(e), w->lane_max));
    }
    if (const char* e = getenv("JH_FLOW_QUEUE")) w->flow_qmode = std::max(0, std::min(atoi(e), 2));
    if (const char* e = getenv("JH_TRAVERSE")) w->traverse_wave = strcmp(e, "thread") != 0;
    if (const char* e = getenv("JH_NARROW_LDS")) w->narrow_lds = atoi(e) != 0;
    if (const char* e = getenv("JH_NARROW_SORT")) w->narrow_sort = atoi(e) != 0;
    if (const char* e = getenv("JH_GRAPH")) w->use_graphs = atoi(e) != 0;
    if (const char* e = getenv("JH_TRAVERSE_QPW")) {
        int q = atoi(e);
        if (q == 16 || q == 32 || q == 64) w->traverse_qpw = q;
    }
    w->tiles = w->cu_count;
    if (const char* e = getenv("JH_TILES")) w->tiles = std::max(1, std::min(atoi(e), 8 * w->cu_count));
    w->colour_blocks = std::min(64, w->coop_colour);
    if (const char* e = getenv("JH_COLOUR_BLOCKS")) w->colour_blocks = std::max(1, std::min(atoi(e), w->coop_colour));
    *out = w;
    return JH_OK;
}

jh_status jh_world_destroy(jh_world* w) {
    if (!w) return fail(JH_ERR_INVALID_ARG, "world is NULL");
    (void)hipStreamSynchronize(w->st);
    DBuf<float4>* f4[] = {&w->pos, &w->quat, &w->vel, &w->angvel, &w->dvel, &w->dangvel, &w->imv, &w->iw0,
                          &w->il0, &w->mat, &w->mat2, &w->force, &w->torque, &w->amin, &w->amax, &w->lmin,
                          &w->lmax, &w->nmin, &w->nmax};
    for (auto* b : f4) b->release();
    w->iw1.release(); w->il1.release();
    w->motion.release(); w->gravity.release(); w->world.release();
    w->body_io.release(); w->rbd_io.release(); w->shapes.release(); w->verts.release();
    w->mk0.release(); w->mk1.release(); w->mv0.release(); w->mv1.release();
    w->left.release(); w->right.release(); w->parent.release(); w->leaf_parent.release(); w->nflags.release();
    w->nbox.release(); w->nrng.release();
    w->range.release(); w->bounds.release(); w->hist.release(); w->pk0.release(); w->pk1.release();
    w->cnt.release(); w->man.release(); w->deep.release(); w->epa_work.release(); w->epa_side.release(); w->epa_res.release(); w->fqueue.release(); w->soa_mv.release();
    w->flags.release(); w->offs.release(); w->sums.release(); w->cnew.release(); w->noffs.release();
    w->ckey.release(); w->nkey.release(); w->cpair.release(); w->cfound.release();
    for (int k = 0; k < 2; ++k) { w->arb[k].release(); w->akey[k].release(); }
    w->colour.release(); w->cwork.release(); w->soa_bodies.release(); w->soa_um.release();
    w->soa_fs.release(); w->soa_nd.release(); w->soa_flag.release(); w->soa_f.release();
    w->wbody.release(); w->wboff.release(); w->wsched.release(); w->wcells.release(); w->wqueue.release();
    w->whist.release(); w->woff.release(); w->wcur.release(); w->wsums.release(); w->wlist.release();
    w->tfirst.release(); w->tbody.release(); w->tile_of.release(); w->tslot.release(); w->tflags.release();
    w->trank.release(); w->tshared.release(); w->thist.release(); w->toff.release(); w->tcur.release();
    w->tsums.release(); w->tlist.release(); w->rec.release();
    w->vwpack.release(); w->chist.release(); w->coffs.release(); w->bmax.release(); w->mlo.release(); w->mhi.release();
    w->ok0.release(); w->ok1.release(); w->solve_keys.release();
    w->rank.release(); w->deg.release(); w->scnt.release();
    w->bar.release();
    w->lws.release(); w->pvel.release(); w->pang.release(); w->parb.release(); w->ppart.release();
    w->pfeat.release(); w->ptarg.release(); w->pvalid.release(); w->plam.release(); w->plearn.release();
    if (w->bp_exec) (void)hipGraphExecDestroy(w->bp_exec);
    if (w->bp_graph) (void)hipGraphDestroy(w->bp_graph);
    for (auto& e : w->ev) (void)hipEventDestroy(e);
    if (w->hcnt) (void)hipHostFree(w->hcnt);
    (void)hipStreamDestroy(w->st);
    delete w;
    return JH_OK;
}

jh_status jh_world_set_config(jh_world* w, const jh_config* c) {
    if (!w) return fail(JH_ERR_INVALID_ARG, "world is NULL");
    jh_status s = check_config(c);
    if (s != JH_OK) return s;
    // Deactivation switched on while every body is awake: the islands were not tracked, so every
    // island is put on the active list for the next check (AddToActiveList: none sleeps before then).
    bool touch_all = !w->cfg.allow_deactivation && c->allow_deactivation && w->n_asleep == 0 && w->nb > 0;
    w->cfg = *c;
    if (touch_all) {
        isl::k_set_bits<<<grid_for(w->nb, 256, 1 << 20), 256, 0, w->st>>>(w->nb, w->sflags.p, isl::kTouch);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(w->st));
        w->isl_dirty = true;
    }
    return JH_OK;
}

jh_status jh_world_set_bodies(jh_world* w, const jh_body* bodies, int32_t n) {
    if (!w || (n > 0 && !bodies) || n < 0) return fail(JH_ERR_INVALID_ARG, "bad arguments");
    // the broadphase packs a body index into 30 bits of its leaf records (kLeafBody)
    if (n > int(kLeafBody)) return fail(JH_ERR_CAPACITY, "at most %u bodies per world", kLeafBody);
    int maxw = 0;
    for (int i = 0; i < n; ++i) {
        const jh_body& b = bodies[i];
        if (b.motion_type < 0 || b.motion_type > 2) return fail(JH_ERR_INVALID_ARG, "body %d: bad motion_type", i);
        if (b.world_id < 0) return fail(JH_ERR_INVALID_ARG, "body %d: negative world_id", i);
        const float* f = (const float*)&b;
        for (int k = 0; k < 36; ++k)
            if (!std::isfinite(f[k])) return fail(JH_ERR_NON_FINITE, "body %d: non-finite field", i);
        jh_status sv = check_sleep_fields(b, i);
        if (sv != JH_OK) return sv;
        maxw = std::max(maxw, b.world_id);
    }
    // a new body count or a motion-type change rebuilds the islands with every body awake
    bool rebuild = n != w->nb || int(w->host_bodies.size()) != n;
    for (int i = 0; i < n && !rebuild; ++i) rebuild = w->host_bodies[i].motion_type != bodies[i].motion_type;
    std::vector<int> relocked;  // bodies whose motion locks change
    for (int i = 0; i < n && i < int(w->host_bodies.size()); ++i)
        if ((w->host_bodies[i].locked_axes & 63) != (bodies[i].locked_axes & 63)) relocked.push_back(i);
    w->nb = n;
    w->max_world = maxw;
    w->n_spec = 0;
    for (int i = 0; i < n; ++i) w->n_spec += bodies[i].enable_speculative_contacts != 0;
    w->con_dirty = true;  // constraint batches depend on which bodies are dynamic
    w->host_bodies.assign(bodies, bodies + n);
    {
        // bodies grouped by world (counting sort), worlds scheduled largest first
        int W = n > 0 ? maxw + 1 : 0;
        std::vector<int> off(W + 1, 0), list(std::max(n, 1)), sched(std::max(W, 1));
        for (int i = 0; i < n; ++i) off[bodies[i].world_id + 1]++;
        for (int k = 0; k < W; ++k) off[k + 1] += off[k];
        std::vector<int> fill(off.begin(), off.end() - (W > 0 ? 1 : 0));
        for (int i = 0; i < n; ++i) list[fill[bodies[i].world_id]++] = i;
        for (int k = 0; k < W; ++k) sched[k] = k;
        std::stable_sort(sched.begin(), sched.begin() + W,
                         [&](int a, int b) { return off[a + 1] - off[a] > off[b + 1] - off[b]; });
        w->n_worlds = W;
        w->largest_world = 0;
        for (int k = 0; k < W; ++k) w->largest_world = std::max(w->largest_world, off[k + 1] - off[k]);
        const char* e = getenv("JH_SOLVER");
        if (e && strcmp(e, "barrier") == 0) w->kind = SOLVER_BARRIER;
        else if (e && strcmp(e, "world") == 0) w->kind = SOLVER_WORLDS;
        else if (e && strcmp(e, "tiles") == 0) w->kind = SOLVER_TILES;
        else if (e && strcmp(e, "lanes") == 0) w->kind = SOLVER_LANES;
        else if (e && strcmp(e, "flow") == 0) w->kind = SOLVER_FLOW;
        else {
            // many worlds, or only small ones (one workgroup per world beats a device-wide launch
            // whose grid barriers and dependency round trips dominate a few hundred arbiters)
            w->kind = (W >= 64 || w->largest_world <= kSmallWorld) ? SOLVER_WORLDS : SOLVER_FLOW;
        }
        w->n_moving = 0;
        for (int i = 0; i < n; ++i) w->n_moving += bodies[i].motion_type != JH_MOTION_STATIC;
        if (W > 0) {
            HIPCHK(w->wbody.ensure(std::max(n, 1)));
            HIPCHK(w->wboff.ensure(W + 1));
            HIPCHK(w->wsched.ensure(W));
            HIPCHK(w->wcells.ensure(1));
            HIPCHK(w->wqueue.ensure(1));
            size_t cells = size_t(W) * 128 + 1;
            HIPCHK(w->whist.ensure(cells));
            HIPCHK(w->woff.ensure(cells));
            HIPCHK(w->wcur.ensure(cells));
            HIPCHK(w->wsums.ensure(cells / 256 + 2));
            HIPCHK(hipMemcpyAsync(w->wbody.p, list.data(), size_t(n) * sizeof(int), hipMemcpyHostToDevice, w->st));
            HIPCHK(hipMemcpyAsync(w->wboff.p, off.data(), size_t(W + 1) * sizeof(int), hipMemcpyHostToDevice, w->st));
            HIPCHK(hipMemcpyAsync(w->wsched.p, sched.data(), size_t(W) * sizeof(int), hipMemcpyHostToDevice, w->st));
            k_set_int<<<1, 1, 0, w->st>>>(w->wcells.p, int(cells));
            HIPCHK(hipStreamSynchronize(w->st));
        }
    }
    size_t nn = std::max(n, 1);
    DBuf<float4>* f4[] = {&w->pos, &w->quat, &w->vel, &w->angvel, &w->dvel, &w->dangvel, &w->imv,
                          &w->iw0, &w->il0, &w->mat, &w->mat2, &w->force, &w->torque};
    for (auto* b : f4) HIPCHK(b->ensure(nn));
    HIPCHK(w->iw1.ensure(nn));
    HIPCHK(w->il1.ensure(nn));
    HIPCHK(w->motion.ensure(nn));
    HIPCHK(w->gravity.ensure(nn));
    HIPCHK(w->world.ensure(nn));
    HIPCHK(w->locks.ensure(nn));
    HIPCHK(w->body_io.ensure(nn));
    HIPCHK(w->bmax.ensure(2 * nn));
    HIPCHK(w->mlo.ensure(nn));
    HIPCHK(w->mhi.ensure(nn));
    HIPCHK(w->deg.ensure(nn));
    HIPCHK(w->scnt.ensure(nn));
    HIPCHK(w->vwpack.ensure(2 * nn));
    HIPCHK(w->slp.ensure(nn));
    HIPCHK(w->sflags.ensure(nn));
    HIPCHK(w->hfl.ensure(nn));
    HIPCHK(w->gfl.ensure(nn));
    HIPCHK(w->smotion.ensure(nn));
    HIPCHK(w->uf.ensure(nn));
    HIPCHK(w->hroot.ensure(nn));
    HIPCHK(w->groot.ensure(nn));
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(w->body_io.p, bodies, n * sizeof(jh_body), hipMemcpyHostToDevice, w->st));
        k_unpack_bodies<<<grid_for(n, 256, 1 << 20), 256, 0, w->st>>>(w->body_io.p, n, w->bodies());
        isl::k_upload<<<grid_for(n, 256, 1 << 20), 256, 0, w->st>>>(n, w->motion.p, w->sflags.p, w->slp.p,
                                                                    rebuild ? 1 : 0);
        HIPCHK(hipGetLastError());
        refresh_solver_motion(w);
        w->isl_dirty = true;
    }
    if (rebuild) w->n_asleep = 0;
    if (!relocked.empty()) {
        jh_status s = locks_changed(w, relocked);
        if (s != JH_OK) return s;
    }
    k_set_int<<<1, 1, 0, w->st>>>(w->cnt.p + C_NB, n);
    w->recolour = true;  // motion types may have changed
    HIPCHK(hipStreamSynchronize(w->st));
    return JH_OK;
}

jh_status jh_world_get_bodies(jh_world* w, jh_body* bodies, int32_t n) {
    if (!w || !bodies || n < 0 || n > w->nb) return fail(JH_ERR_INVALID_ARG, "bad arguments");
    if (n == 0) return JH_OK;
    k_pack_bodies<<<grid_for(n, 256, 1 << 20), 256, 0, w->st>>>(w->bodies(), n, w->body_io.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(bodies, w->body_io.p, n * sizeof(jh_body), hipMemcpyDeviceToHost, w->st));
    HIPCHK(hipStreamSynchronize(w->st));
    return JH_OK;
}

jh_status jh_world_update_bodies(jh_world* w, const int32_t* index, const jh_body* records, int32_t count) {
    if (!w || count < 0 || (count > 0 && (!index || !records))) return fail(JH_ERR_INVALID_ARG, "bad arguments");
    if (count == 0) return JH_OK;
    std::vector<char> seen(w->nb, 0);
    bool motion_change = false;
    for (int k = 0; k < count; ++k) {
        int i = index[k];
        if (i < 0 || i >= w->nb) return fail(JH_ERR_INVALID_ARG, "record %d: body index %d out of range", k, i);
        if (seen[i]) return fail(JH_ERR_INVALID_ARG, "record %d: body %d listed twice", k, i);
        seen[i] = 1;
        const jh_body& b = records[k];
        if (b.motion_type < 0 || b.motion_type > 2) return fail(JH_ERR_INVALID_ARG, "record %d: bad motion_type", k);
        if (b.world_id != w->host_bodies[i].world_id)
            return fail(JH_ERR_UNSUPPORTED, "record %d: a world_id change needs jh_world_set_bodies", k);
        if ((b.enable_speculative_contacts != 0) != (w->host_bodies[i].enable_speculative_contacts != 0))
            return fail(JH_ERR_UNSUPPORTED, "record %d: EnableSpeculativeContacts changes need jh_world_set_bodies", k);
        const float* f = (const float*)&b;
        for (int q = 0; q < 36; ++q)
            if (!std::isfinite(f[q])) return fail(JH_ERR_NON_FINITE, "record %d: non-finite field", k);
        jh_status sv = check_sleep_fields(b, k);
        if (sv != JH_OK) return sv;
        motion_change = motion_change || b.motion_type != w->host_bodies[i].motion_type;
    }
    if (motion_change) {
        // a motion-type change rebuilds the islands and the world's bookkeeping: the whole set
        std::vector<jh_body> all(w->nb);
        HIPCHK(w->body_io.ensure(std::max(w->nb, 1)));
        jh_status s = jh_world_get_bodies(w, all.data(), w->nb);
        if (s != JH_OK) return s;
        for (int k = 0; k < count; ++k) all[index[k]] = records[k];
        return jh_world_set_bodies(w, all.data(), w->nb);
    }
    HIPCHK(w->body_io.ensure(std::max<size_t>(size_t(count), 1)));
    HIPCHK(w->uidx.ensure(std::max<size_t>(size_t(count), 1)));
    HIPCHK(hipMemcpyAsync(w->body_io.p, records, size_t(count) * sizeof(jh_body), hipMemcpyHostToDevice, w->st));
    HIPCHK(hipMemcpyAsync(w->uidx.p, index, size_t(count) * sizeof(int), hipMemcpyHostToDevice, w->st));
    k_unpack_listed<<<grid_for(count, 256, 1 << 20), 256, 0, w->st>>>(w->body_io.p, w->uidx.p, count, w->bodies());
    std::vector<int> relocked;
    for (int k = 0; k < count; ++k)
        if ((w->host_bodies[index[k]].locked_axes & 63) != (records[k].locked_axes & 63)) relocked.push_back(index[k]);
    for (int k = 0; k < count; ++k) w->host_bodies[index[k]] = records[k];
    if (!relocked.empty()) {
        jh_status s = locks_changed(w, relocked);
        if (s != JH_OK) return s;
    }
    // each listed body: ActivateBodyNextStep(body, wakeUpStatic = true)
    wake_listed(w, w->uidx.p, count, 1, 1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(w->st));
    return JH_OK;
}

jh_status jh_world_set_activation_state(jh_world* w, int32_t body, int32_t active) {
    if (!w) return fail(JH_ERR_INVALID_ARG, "world is NULL");
    if (body < 0 || body >= w->nb) return fail(JH_ERR_INVALID_ARG, "body index %d out of range", body);
    HIPCHK(w->uidx.ensure(1));
    HIPCHK(hipMemcpyAsync(w->uidx.p, &body, sizeof(int), hipMemcpyHostToDevice, w->st));
    wake_listed(w, w->uidx.p, 1, active ? 1 : 0, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(w->st));
    return JH_OK;
}

jh_status jh_world_get_activation(jh_world* w, uint8_t* active, float* sleep_time, int32_t n) {
    if (!w || n < 0 || n > w->nb) return fail(JH_ERR_INVALID_ARG, "bad arguments");
    if (n == 0) return JH_OK;
    std::vector<uint32_t> f(n);
    std::vector<float4> t(n);
    HIPCHK(hipMemcpyAsync(f.data(), w->sflags.p, size_t(n) * sizeof(uint32_t), hipMemcpyDeviceToHost, w->st));
    HIPCHK(hipMemcpyAsync(t.data(), w->slp.p, size_t(n) * sizeof(float4), hipMemcpyDeviceToHost, w->st));
    HIPCHK(hipStreamSynchronize(w->st));
    for (int i = 0; i < n; ++i) {
        if (active) active[i] = (f[i] & isl::kAwake) ? 1 : 0;
        if (sleep_time) sleep_time[i] = t[i].w;
    }
    return JH_OK;
}

jh_status jh_world_set_rigidbodydata(jh_world* w, const void* records, int32_t n, int32_t stride) {
    if (!w || (n > 0 && !records) || n < 0 || stride < 128) return fail(JH_ERR_INVALID_ARG, "bad arguments");
    if (n != w->nb) return fail(JH_ERR_INVALID_ARG, "%d records for %d bodies: RawData covers every body", n, w->nb);
    if (n == 0) return JH_OK;
    const char* in = static_cast<const char*>(records);
    std::vector<uint32_t> sf(n);
    HIPCHK(hipMemcpyAsync(sf.data(), w->sflags.p, size_t(n) * sizeof(uint32_t), hipMemcpyDeviceToHost, w->st));
    HIPCHK(hipStreamSynchronize(w->st));
    for (int i = 0; i < n; ++i) {
        const char* r = in + size_t(i) * stride;
        int32_t flags;
        memcpy(&flags, r + 120, 4);
        if ((flags & 3) != w->host_bodies[i].motion_type)
            return fail(JH_ERR_UNSUPPORTED, "body %d: the motion type differs from the body's; set it with jh_world_set_bodies", i);
        if (((flags & 4) != 0) != ((sf[i] & isl::kAwake) != 0))
            return fail(JH_ERR_UNSUPPORTED, "body %d: IsActive differs from the body's; use jh_world_set_activation_state", i);
        if (flags & 8) return fail(JH_ERR_UNSUPPORTED, "body %d: gyroscopic integration is not on the GPU path", i);
        if (((flags >> 4) & 63) != (w->host_bodies[i].locked_axes & 63))
            return fail(JH_ERR_UNSUPPORTED, "body %d: the motion locks differ from the body's; set them with jh_world_update_bodies", i);
        if (flags & ~0x3ff) return fail(JH_ERR_UNSUPPORTED, "body %d: undefined flag bits", i);
        float f[28];
        memcpy(f, r + 8, sizeof(f));
        for (int k = 0; k < 28; ++k)
            if (!std::isfinite(f[k])) return fail(JH_ERR_NON_FINITE, "body %d: non-finite field at offset %d", i, 8 + 4 * k);
    }
    HIPCHK(w->rbd_io.ensure(size_t(n) * stride));
    HIPCHK(hipMemcpyAsync(w->rbd_io.p, records, size_t(n) * stride, hipMemcpyHostToDevice, w->st));
    k_unpack_rigidbodydata<<<grid_for(n, 256, 1 << 20), 256, 0, w->st>>>(w->rbd_io.p, n, stride, w->bodies());
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(w->st));
    return JH_OK;
}

jh_status jh_world_get_rigidbodydata(jh_world* w, void* records, int32_t n, int32_t stride) {
    if (!w || !records || n < 0 || n > w->nb || stride < 128) return fail(JH_ERR_INVALID_ARG, "bad arguments");
    if (n == 0) return JH_OK;
    HIPCHK(w->rbd_io.ensure(size_t(n) * stride));
    HIPCHK(hipMemsetAsync(w->rbd_io.p, 0, size_t(n) * stride, w->st));
    k_pack_rigidbodydata<<<grid_for(n, 256, 1 << 20), 256, 0, w->st>>>(w->bodies(), n, w->rbd_io.p, stride);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(records, w->rbd_io.p, size_t(n) * stride, hipMemcpyDeviceToHost, w->st));
    HIPCHK(hipStreamSynchronize(w->st));
    return JH_OK;
}

jh_status jh_world_set_shapes(jh_world* w, const jh_shape* shapes, int32_t n, const float* vertices, int32_t nv) {
    if (!w || n < 0 || (n > 0 && !shapes) || nv < 0 || (nv > 0 && !vertices))
        return fail(JH_ERR_INVALID_ARG, "bad arguments");
    std::vector<ShapeRec> recs(n);
    for (int i = 0; i < n; ++i) {
        const jh_shape& s = shapes[i];
        if (i > 0 && !(shapes[i - 1].shape_id < s.shape_id))
            return fail(JH_ERR_INVALID_ARG, "shape %d: shape ids must be strictly ascending", i);
        if (s.body < 0 || s.body >= w->nb) return fail(JH_ERR_INVALID_ARG, "shape %d: body index out of range", i);
        if (s.type < JH_SHAPE_BOX || s.type > JH_SHAPE_CONVEXHULL)
            return fail(JH_ERR_INVALID_ARG, "shape %d: unknown type", i);
        if (s.type == JH_SHAPE_CONVEXHULL) {
            jh_status hs = check_hull(s, vertices, nv, i);
            if (hs != JH_OK) return hs;
        }
        if ((s.type == JH_SHAPE_POINTCLOUD && (s.vertex_count < 1)) || (s.type == JH_SHAPE_TRIANGLE && s.vertex_count != 3))
            return fail(JH_ERR_INVALID_ARG, "shape %d: bad vertex count", i);
        if ((s.type == JH_SHAPE_POINTCLOUD || s.type == JH_SHAPE_TRIANGLE) &&
            (s.vertex_offset < 0 || s.vertex_offset + s.vertex_count > nv))
            return fail(JH_ERR_INVALID_ARG, "shape %d: vertex range out of bounds", i);
        for (int k = 0; k < 12; ++k)
            if (!std::isfinite(s.params[k])) return fail(JH_ERR_NON_FINITE, "shape %d: non-finite params", i);
        ShapeRec& r = recs[i];
        r.type = s.type;
        r.body = s.body;
        for (int k = 0; k < 12; ++k) r.p[k] = s.params[k];
        r.voff = s.vertex_offset;
        r.vcnt = s.vertex_count;
    }
    w->ns = n;
    w->nv = nv;
    w->has_clouds = false;
    for (int i = 0; i < n; ++i) w->has_clouds = w->has_clouds || lds_cloud(recs[i]);
    size_t nn = std::max(n, 2);
    HIPCHK(w->shapes.ensure(nn));
    HIPCHK(w->verts.ensure(std::max<size_t>(3 * size_t(nv), 3)));
    DBuf<float4>* f4[] = {&w->amin, &w->amax, &w->lmin, &w->lmax, &w->nmin, &w->nmax};
    for (auto* b : f4) HIPCHK(b->ensure(nn));
    HIPCHK(w->mk0.ensure(nn)); HIPCHK(w->mk1.ensure(nn)); HIPCHK(w->mv0.ensure(nn)); HIPCHK(w->mv1.ensure(nn));
    HIPCHK(w->left.ensure(nn)); HIPCHK(w->right.ensure(nn)); HIPCHK(w->parent.ensure(nn));
    HIPCHK(w->leaf_parent.ensure(nn)); HIPCHK(w->nflags.ensure(nn)); HIPCHK(w->range.ensure(nn));
    HIPCHK(w->nbox.ensure(4 * nn)); HIPCHK(w->nrng.ensure(nn));
    HIPCHK(w->hist.ensure(rsort::hist_entries(size_t(nn))));
    int want = w->cfg.pair_capacity > 0 ? w->cfg.pair_capacity : 24 * n + 1024;
    if (w->pair_cap < want) {
        jh_status s = ensure_pair_buffers(w, want);
        if (s != JH_OK) return s;
    }
    if (n > 0) HIPCHK(hipMemcpyAsync(w->shapes.p, recs.data(), n * sizeof(ShapeRec), hipMemcpyHostToDevice, w->st));
    if (nv > 0) HIPCHK(hipMemcpyAsync(w->verts.p, vertices, 3 * size_t(nv) * sizeof(float), hipMemcpyHostToDevice, w->st));
    k_set_int<<<1, 1, 0, w->st>>>(w->cnt.p + C_NS, n);
    // a new shape set invalidates the contact cache (RigidBody.ClearContactCache on shape changes)
    k_set_int<<<1, 1, 0, w->st>>>(w->cnt.p + C_A0, 0);
    k_set_int<<<1, 1, 0, w->st>>>(w->cnt.p + C_A1, 0);
    k_set_int<<<1, 1, 0, w->st>>>(w->cnt.p + C_ZOMB, 0);
    w->a_host = 0;
    w->cur = 0;
    reset_islands(w);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(w->st));
    return JH_OK;
}

jh_status jh_world_stabilize(jh_world* w, float dt, int32_t iterations, int32_t relaxations) {
    if (!w) return fail(JH_ERR_STATE, "world is NULL");
    // StabilizeInternal argument checks, in its order (World.Step.cs:332-350)
    if (!std::isfinite(dt)) return fail(JH_ERR_NON_FINITE, "dt must be finite");
    if (dt < 0.0f) return fail(JH_ERR_INVALID_ARG, "Time step cannot be negative.");
    if (iterations < 1) return fail(JH_ERR_INVALID_ARG, "Solver iterations cannot be smaller than one.");
    if (relaxations < 0) return fail(JH_ERR_INVALID_ARG, "Relaxation iterations cannot be smaller than zero.");
    if (dt < 1.401298464e-45f || w->nb == 0) return JH_OK;  // nothing to do
    hipStream_t st = w->st;
    int* cnt = w->cnt.p;
    jh_status s = ensure_step_buffers(w);
    if (s != JH_OK) return s;
    if (!w->host_con.empty() && w->con_dirty) {
        s = colour_constraints(w);
        if (s != JH_OK) return s;
    }
    const jh_config& cfg = w->cfg;
    StepTime ts;  // TimeStep (TimeStep.cs:36-46)
    ts.dt = dt;
    ts.substeps = cfg.substep_count;
    ts.substep_dt = dt / float(cfg.substep_count);
    ts.inv_dt = 1.0f / dt;
    ts.inv_substep_dt = ts.substep_dt > 0.0f ? 1.0f / ts.substep_dt : 0.0f;
    int cur = w->cur;
    int* a_cur = cnt + (cur == 0 ? C_A0 : C_A1);
    int amax_ = w->a_host;
    k_set_int<<<1, 1, 0, st>>>(cnt + C_CONTACTS, 0);
    if (w->islands_on()) {  // StabilizeInternal starts with CheckDeactivation (World.Step.cs:362)
        s = check_islands(w, cur, a_cur, amax_, true);
        if (s != JH_OK) return s;
    }
    const uint64_t* order = nullptr;
    s = colour_arbiters(w, cur, a_cur, amax_, &order);
    if (s != JH_OK) return s;
    // the colour-batch solver without IntegrateForces / IntegrateVelocities (SolveArgs.stabilize)
    SolveArgs S{};
    S.arb = w->arb[cur].p;
    S.order = order;
    S.offs = w->coffs.p;
    S.ncolours = cnt + C_NCOL;
    S.bar = w->bar.p;
    S.spin = w->spin();
    S.B = w->solver_bodies();
    S.nb = w->nb;
    S.ts = ts;
    S.iterations = iterations;
    S.relaxations = relaxations;
    S.con.c = w->con.p;
    S.con.work = w->con_work.p;
    S.con.order = w->con_order.p;
    S.con.offs = w->con_offs.p;
    S.con.nbatch = constraint_batches(w);
    S.stabilize = 1;
    int sgrid = std::min(w->coop_solve, grid_for(std::max(std::max(std::max(amax_, w->nb), 256),
                                                          int(w->con_order_h.size())), 256, 1 << 20));
    void* sargs[] = {&S};
    s = launch_coop((const void*)k_solve, sgrid, 256, sargs, st);
    if (s != JH_OK) return s;
    HIPCHK(hipMemcpyAsync(w->hcnt, cnt, C_COUNT * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(w->habort, w->bar.p + kBarAbort, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (*w->habort) {
        HIPCHK(hipMemsetAsync(w->bar.p, 0, 64 * sizeof(unsigned), st));
        HIPCHK(hipStreamSynchronize(st));
        *w->habort = 0;
        w->have_step = false;
        w->recolour = true;
        return fail(JH_ERR_TIMEOUT, "jh_world_stabilize: a wait of the solver gave up after %u polls or it was cancelled",
                    w->spin());
    }
    // jh_world_get_solve_order now describes this solve
    if (w->islands_on()) w->n_asleep = w->hcnt[C_ASLEEP];
    w->solve_count = w->hcnt[cur == 0 ? C_A0 : C_A1];
    w->solve_batches = w->hcnt[C_NCOL];
    w->have_step = true;
    return JH_OK;
}

jh_status jh_world_set_spin_limit(jh_world* w, uint32_t polls) {
    if (!w) return fail(JH_ERR_STATE, "world is NULL");
    w->spin_limit = polls;
    return JH_OK;
}

jh_status jh_world_cancel(jh_world* w) {
    if (!w) return fail(JH_ERR_STATE, "world is NULL");
    unsigned one = 1u;
    HIPCHK(hipMemcpyAsync(w->bar.p + kBarAbort, &one, sizeof(unsigned), hipMemcpyHostToDevice, w->st));
    HIPCHK(hipStreamSynchronize(w->st));
    return JH_OK;
}

jh_status jh_world_step(jh_world* w, float dt) {
    if (!w) return fail(JH_ERR_STATE, "world is NULL");
    // World.StepInternal argument checks (World.Step.cs:155-165)
    if (!std::isfinite(dt)) return fail(JH_ERR_NON_FINITE, "dt must be finite");
    if (dt < 0.0f) return fail(JH_ERR_INVALID_ARG, "Time step cannot be negative.");
    if (dt < 1.401298464e-45f) return JH_OK;
    if (w->nb == 0) return JH_OK;
    if (w->cfg.warm_start == JH_WARM_START_LEARNED && !w->lws.ready)
        return fail(JH_ERR_STATE, "learned warm start selected but no model was set (jh_world_set_warm_start_model)");
    return do_step(w, dt, nullptr);
}

jh_status jh_world_compute_pairs(jh_world* w, uint32_t* pairs, int32_t capacity, int32_t* out_count) {
    if (!w || !out_count || capacity < 0 || (capacity > 0 && !pairs)) return fail(JH_ERR_INVALID_ARG, "bad arguments");
    jh_status s = broadphase(w);
    if (s != JH_OK) return s;
    int np = w->n_pairs;
    *out_count = np;
    if (np > 0 && capacity > 0) {
        std::vector<uint64_t> tmp(std::min(np, capacity));
        HIPCHK(hipMemcpyAsync(tmp.data(), w->pairs_in_k1 ? w->pk1.p : w->pk0.p, tmp.size() * 8, hipMemcpyDeviceToHost,
                              w->st));
        HIPCHK(hipStreamSynchronize(w->st));
        for (size_t i = 0; i < tmp.size(); ++i) {
            pairs[2 * i] = uint32_t(tmp[i] >> 32);
            pairs[2 * i + 1] = uint32_t(tmp[i] & 0xffffffffu);
        }
    }
    if (np > capacity) return fail(JH_ERR_CAPACITY, "pair count %d exceeds capacity %d", np, capacity);
    return JH_OK;
}

jh_status jh_world_get_manifolds(jh_world* w, jh_manifold* out, int32_t capacity, int32_t* out_count) {
    if (!w || !out_count || capacity < 0 || (capacity > 0 && !out)) return fail(JH_ERR_INVALID_ARG, "bad arguments");
    *out_count = 0;
    if (!w->have_step || w->n_pairs == 0) return JH_OK;
    std::vector<jh_manifold> all(w->n_pairs);
    HIPCHK(hipMemcpyAsync(all.data(), w->man.p, all.size() * sizeof(jh_manifold), hipMemcpyDeviceToHost, w->st));
    HIPCHK(hipStreamSynchronize(w->st));
    int n = 0;
    for (auto& m : all) {
        if (m.count <= 0) continue;
        if (n < capacity) out[n] = m;
        n++;
    }
    *out_count = n;
    if (n > capacity) return fail(JH_ERR_CAPACITY, "manifold count %d exceeds capacity %d", n, capacity);
    return JH_OK;
}

jh_status jh_world_set_constraints(jh_world* w, const jh_constraint* in, int32_t n) {
    if (!w || n < 0 || (n > 0 && !in)) return fail(JH_ERR_INVALID_ARG, "bad arguments");
    for (int i = 0; i < n; ++i) {
        const jh_constraint& c = in[i];
        if (c.type < JH_CONSTRAINT_BALL_SOCKET || c.type > JH_CONSTRAINT_ELLIPTICAL_CONE_LIMIT)
            return fail(JH_ERR_INVALID_ARG, "constraint %d: unknown type %d", i, c.type);
        if (c.body1 < 0 || c.body1 >= w->nb || c.body2 < 0 || c.body2 >= w->nb)
            return fail(JH_ERR_INVALID_ARG, "constraint %d: body index out of range", i);
        if (c.body1 == c.body2)  // SameBodyException (World.CreateConstraint, World.cs:860-861)
            return fail(JH_ERR_INVALID_ARG, "constraint %d: body1 and body2 are the same body", i);
        for (int k = 0; k < 16; ++k)
            if (std::isnan(c.params[k])) return fail(JH_ERR_NON_FINITE, "constraint %d: NaN parameter", i);
        // ServoMath.CheckMaximum (Internal/ServoMath.cs:15-20): a servo's force or torque limit is >= 0
        int imax = c.type == JH_CONSTRAINT_ANGULAR_SERVO       ? 4
                   : c.type == JH_CONSTRAINT_LINEAR_SERVO      ? 9
                   : c.type == JH_CONSTRAINT_LINEAR_AXIS_SERVO ? 10
                   : c.type == JH_CONSTRAINT_TWIST_SERVO       ? 8
                                                               : -1;
        if (imax >= 0 && c.params[imax] < 0.0f)
            return fail(JH_ERR_INVALID_ARG, "constraint %d: negative servo maximum", i);
        // EllipticalConeLimit.ValidateLimits (EllipticalConeLimit.cs:168-191): 0 <= min <= max < pi, max > 0,
        // inner semiaxes both zero or both positive
        if (c.type == JH_CONSTRAINT_ELLIPTICAL_CONE_LIMIT) {
            const float* p = c.params;
            bool ok = true;
            for (int r = 0; r < 2; ++r) {
                float mn = p[9 + 2 * r], mx = p[10 + 2 * r];
                ok = ok && std::isfinite(mn) && std::isfinite(mx) && mn >= 0.0f && mx > 0.0f && mx < 3.14159265f &&
                     mn <= mx;
            }
            if (!ok || ((p[9] == 0.0f) != (p[11] == 0.0f)))
                return fail(JH_ERR_INVALID_ARG, "constraint %d: elliptical cone limits out of range", i);
        }
    }
    w->host_con.assign(in, in + n);
    size_t cap = std::max(n, 1);
    HIPCHK(w->con.ensure(cap));
    HIPCHK(w->con_work.ensure(cap));
    if (n > 0) HIPCHK(hipMemcpyAsync(w->con.p, in, n * sizeof(jh_constraint), hipMemcpyHostToDevice, w->st));
    HIPCHK(hipMemsetAsync(w->con_work.p, 0, cap * sizeof(con::Work), w->st));
    w->con_dirty = true;
    jh_status s = colour_constraints(w);
    if (s != JH_OK) return s;
    reset_islands(w);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(w->st));
    return JH_OK;
}

jh_status jh_world_get_constraints(jh_world* w, jh_constraint* out, int32_t capacity, int32_t* out_count) {
    if (!w || !out_count || capacity < 0 || (capacity > 0 && !out)) return fail(JH_ERR_INVALID_ARG, "bad arguments");
    int n = int(w->host_con.size());
    *out_count = n;
    int k = std::min(n, capacity);
    if (k > 0) {
        HIPCHK(hipMemcpyAsync(out, w->con.p, k * sizeof(jh_constraint), hipMemcpyDeviceToHost, w->st));
        HIPCHK(hipStreamSynchronize(w->st));
    }
    if (n > capacity) return fail(JH_ERR_CAPACITY, "constraint count %d exceeds capacity %d", n, capacity);
    return JH_OK;
}

jh_status jh_world_get_constraint_order(jh_world* w, int32_t* order, int32_t capacity, int32_t* batch_offsets,
                                        int32_t batch_capacity, int32_t* out_count, int32_t* out_batch_count) {
    if (!w || !out_count || !out_batch_count || capacity < 0 || batch_capacity < 0 || (capacity > 0 && !order) ||
        (batch_capacity > 0 && !batch_offsets))
        return fail(JH_ERR_INVALID_ARG, "bad arguments");
    if (w->con_dirty && !w->host_con.empty()) {
        jh_status s = colour_constraints(w);
        if (s != JH_OK) return s;
        HIPCHK(hipStreamSynchronize(w->st));
    }
    int n = int(w->con_order_h.size()), nb = constraint_batches(w);
    *out_count = n;
    *out_batch_count = nb;
    for (int i = 0; i < std::min(n, capacity); ++i) order[i] = w->con_order_h[i];
    for (int i = 0; i < std::min(nb + 1, batch_capacity); ++i) batch_offsets[i] = w->con_offs_h[i];
    if (n > capacity || nb + 1 > batch_capacity)
        return fail(JH_ERR_CAPACITY, "constraint order (%d, %d batches) exceeds capacity", n, nb);
    return JH_OK;
}

jh_status jh_world_get_collision_events(jh_world* w, uint32_t* begin_pairs, int32_t begin_capacity,
                                        int32_t* out_begin_count, uint32_t* end_pairs, int32_t end_capacity,
                                        int32_t* out_end_count) {
    if (!w || !out_begin_count || !out_end_count || begin_capacity < 0 || end_capacity < 0 ||
        (begin_capacity > 0 && !begin_pairs) || (end_capacity > 0 && !end_pairs))
        return fail(JH_ERR_INVALID_ARG, "bad arguments");
    *out_begin_count = *out_end_count = 0;
    if (!w->have_step) return JH_OK;
    int nb = w->ev_begin, ne = w->ev_end;
    std::vector<uint64_t> kb(nb), ke(ne);
    if (nb > 0) HIPCHK(hipMemcpyAsync(kb.data(), w->evb.p, nb * sizeof(uint64_t), hipMemcpyDeviceToHost, w->st));
    if (ne > 0) HIPCHK(hipMemcpyAsync(ke.data(), w->eve.p, ne * sizeof(uint64_t), hipMemcpyDeviceToHost, w->st));
    HIPCHK(hipStreamSynchronize(w->st));
    std::sort(kb.begin(), kb.end());  // appended by many threads: sort into key order
    std::sort(ke.begin(), ke.end());
    auto put = [](const std::vector<uint64_t>& k, uint32_t* out, int cap) {
        for (int i = 0; i < int(k.size()) && i < cap; ++i) {
            out[2 * i] = uint32_t(k[i] >> 32);
            out[2 * i + 1] = uint32_t(k[i] & 0xffffffffu);
        }
    };
    put(kb, begin_pairs, begin_capacity);
    put(ke, end_pairs, end_capacity);
    *out_begin_count = nb;
    *out_end_count = ne;
    if (nb > begin_capacity || ne > end_capacity)
        return fail(JH_ERR_CAPACITY, "%d begin / %d end events exceed capacities %d / %d", nb, ne, begin_capacity,
                    end_capacity);
    return JH_OK;
}

jh_status jh_world_get_arbiters(jh_world* w, jh_arbiter* out, int32_t capacity, int32_t* out_count) {
    if (!w || !out_count || capacity < 0 || (capacity > 0 && !out)) return fail(JH_ERR_INVALID_ARG, "bad arguments");
    int n = w->a_host;
    *out_count = n;
    int k = std::min(n, capacity);
    if (k > 0) {
        HIPCHK(hipMemcpyAsync(out, w->arb[w->cur].p, k * sizeof(jh_arbiter), hipMemcpyDeviceToHost, w->st));
        HIPCHK(hipStreamSynchronize(w->st));
    }
    if (n > capacity) return fail(JH_ERR_CAPACITY, "arbiter count %d exceeds capacity %d", n, capacity);
    return JH_OK;
}

jh_status jh_world_set_arbiters(jh_world* w, const jh_arbiter* in, int32_t n) {
    if (!w || n < 0 || (n > 0 && !in)) return fail(JH_ERR_INVALID_ARG, "bad arguments");
    std::vector<jh_arbiter> a(in, in + n);
    for (auto& r : a) {
        if (r.shape1 >= uint32_t(w->ns) || r.shape2 >= uint32_t(w->ns) || !(r.shape1 < r.shape2))
            return fail(JH_ERR_INVALID_ARG, "arbiter keys must be shape indices with shape1 < shape2");
        if (r.body1 < 0 || r.body1 >= w->nb || r.body2 < 0 || r.body2 >= w->nb)
            return fail(JH_ERR_INVALID_ARG, "arbiter body index out of range");
    }
    std::sort(a.begin(), a.end(), [](const jh_arbiter& x, const jh_arbiter& y) {
        return x.shape1 < y.shape1 || (x.shape1 == y.shape1 && x.shape2 < y.shape2);
    });
    std::vector<uint64_t> keys(n);
    for (int i = 0; i < n; ++i) {
        keys[i] = (uint64_t(a[i].shape1) << 32) | a[i].shape2;
        if (i > 0 && keys[i] == keys[i - 1]) return fail(JH_ERR_INVALID_ARG, "duplicate arbiter key");
    }
    size_t nn = std::max(n, 1);
    HIPCHK(w->arb[w->cur].ensure(nn));
    HIPCHK(w->akey[w->cur].ensure(nn));
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(w->arb[w->cur].p, a.data(), n * sizeof(jh_arbiter), hipMemcpyHostToDevice, w->st));
        HIPCHK(hipMemcpyAsync(w->akey[w->cur].p, keys.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice, w->st));
    }
    k_set_int<<<1, 1, 0, w->st>>>(w->cnt.p + (w->cur == 0 ? C_A0 : C_A1), n);
    k_set_int<<<1, 1, 0, w->st>>>(w->cnt.p + C_ZOMB, 0);  // the restored cache has no pending removals
    reset_islands(w);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(w->st));
    w->a_host = n;
    w->recolour = true;  // colours in uploaded arbiters are not trusted
    return JH_OK;
}

jh_status jh_world_get_solve_order(jh_world* w, uint32_t* keys, int32_t key_capacity, int32_t* batch_offsets,
                                   int32_t batch_capacity, int32_t* out_key_count, int32_t* out_batch_count) {
    if (!w || !out_key_count || !out_batch_count) return fail(JH_ERR_INVALID_ARG, "bad arguments");
    *out_key_count = 0;
    *out_batch_count = 0;
    if (!w->have_step) return JH_OK;
    std::vector<int> offs(129);
    HIPCHK(hipMemcpyAsync(offs.data(), w->coffs.p, 129 * sizeof(int), hipMemcpyDeviceToHost, w->st));
    HIPCHK(hipStreamSynchronize(w->st));
    int nc = w->solve_batches;
    // batches: colours 0..nc-1, then the serial overflow batch if it is not empty
    std::vector<int> bo;
    for (int c = 0; c <= nc; ++c) bo.push_back(offs[c]);
    int serial = offs[kMaxColours + 1] - offs[kMaxColours];
    if (serial > 0) bo.push_back(offs[nc] + serial);
    int nkeys = bo.back();
    int nbatch = int(bo.size()) - 1;
    *out_key_count = nkeys;
    *out_batch_count = nbatch;
    if (keys && key_capacity > 0 && nkeys > 0) {
        std::vector<uint64_t> k(nkeys);
        HIPCHK(hipMemcpyAsync(k.data(), w->solve_keys.p, nkeys * 8, hipMemcpyDeviceToHost, w->st));
        HIPCHK(hipStreamSynchronize(w->st));
        // colour-sorted keys: colours 0..nc-1 are contiguous, the serial batch (colour 127) follows them
        for (int i = 0; i < std::min(nkeys, key_capacity); ++i) {
            keys[2 * i] = uint32_t(k[i] >> 32);
            keys[2 * i + 1] = uint32_t(k[i] & 0xffffffffu);
        }
    }
    if (batch_offsets && batch_capacity > 0)
        for (int i = 0; i < std::min(nbatch + 1, batch_capacity); ++i) batch_offsets[i] = bo[i];
    if (nkeys > key_capacity || nbatch + 1 > batch_capacity)
        return fail(JH_ERR_CAPACITY, "solve order needs %d keys and %d offsets", nkeys, nbatch + 1);
    return JH_OK;
}

jh_status jh_world_get_stats(jh_world* w, jh_stats* s) {
    if (!w || !s) return fail(JH_ERR_INVALID_ARG, "bad arguments");
    *s = w->stats;
    return JH_OK;
}

}  // extern "C"

struct jh_lws {
    hipStream_t st = nullptr;
    int device = 0;
    uint16_t* scratch = nullptr;  // this GPU's staging array of jh_lws_predict_device (from lws_scratch_for)
    LwsWeights W;
    DBuf<float> params, grad, m, v;
    float hmean[lws::kIn], histd[lws::kIn], htscale[3];
    // resident data
    DBuf<uint16_t> xb;
    DBuf<float> tn;
    int64_t n = 0;
    // workspace
    DBuf<uint16_t> xt, h1t, h2t, dz1t, dz2t, dyt;
    DBuf<float> bias_part, gw1, gw2, gw3, tmp_f, tmp_t, out;
    DBuf<double> loss_part, loss_acc;
    int64_t adam_t = 0;
};

// ---- Staging without allocation (jh_lws_predict_device) ---------------------------------------
// The normalised BF16 inputs of up to JH_LWS_PREDICT_MAX_ROWS rows, one copy per GPU for the life of
// the process.  In host code `scratch` only names the symbol: each device's copy is found once, with
// that device current, by hipGetSymbolAddress (lws_scratch_for), and only that address is used.
static_assert(size_t(JH_LWS_PREDICT_MAX_ROWS) * lws::kInPad * sizeof(uint16_t) == (4u << 20), "4 MiB");
static __device__ __align__(256) uint16_t scratch[size_t(JH_LWS_PREDICT_MAX_ROWS) * lws::kInPad];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint16_t* g_scratch_ptr[kMaxDevices] = {};

// Address of the current device's staging array (looked up once per device).
jh_status lws_scratch_for(int dev, uint16_t** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(JH_ERR_UNSUPPORTED, "device %d: at most %d devices", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(JH_ERR_HIP, "hipGetSymbolAddress returned NULL");
        g_scratch_ptr[dev] = static_cast<uint16_t*>(p);
    }
    *out = g_scratch_ptr[dev];
    return JH_OK;
}

// True when p is device memory of GPU dev (allocated memory, not host or registered memory).
bool is_device_memory(const void* p, int dev) {
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    (void)hipGetLastError();  // an unknown host pointer is reported as an error, which is not sticky
    return e == hipSuccess && a.type == hipMemoryTypeDevice && a.device == dev;
}

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

jh_status lws_refresh(jh_lws* L) {
    lws::CopyArgs C;
    C.params = L->params.p;
    C.w1 = L->W.w1.p; C.w2 = L->W.w2.p; C.w3 = L->W.w3.p; C.w2t = L->W.w2t.p; C.w3t = L->W.w3t.p;
    C.b1 = L->W.b1.p; C.b2 = L->W.b2.p; C.b3 = L->W.b3.p;
    lws::k_refresh<<<(lws::kHid * lws::kHid + 255) / 256, 256, 0, L->st>>>(C);
    HIPCHK(hipGetLastError());
    return JH_OK;
}

jh_status lws_upload_norm(jh_lws* L) {
    HIPCHK(hipMemcpyAsync(L->W.mean.p, L->hmean, sizeof(L->hmean), hipMemcpyHostToDevice, L->st));
    HIPCHK(hipMemcpyAsync(L->W.istd.p, L->histd, sizeof(L->histd), hipMemcpyHostToDevice, L->st));
    HIPCHK(hipMemcpyAsync(L->W.tscale.p, L->htscale, sizeof(L->htscale), hipMemcpyHostToDevice, L->st));
    return JH_OK;
}

// One minibatch of resident samples [begin, begin + count): forward/backward, gradients, and an Adam
// update when `adam`; the batch's sum of squared errors is added to loss_acc[slot].
jh_status lws_batch(jh_lws* L, int64_t begin, int count, bool grads, bool adam, float lr, int slot) {
    hipStream_t st = L->st;
    int ld = (count + 15) / 16 * 16;
    int waves = ld / 16;
    HIPCHK(L->xt.ensure(size_t(lws::kInPad) * ld));
    HIPCHK(L->h1t.ensure(size_t(lws::kHid) * ld));
    HIPCHK(L->h2t.ensure(size_t(lws::kHid) * ld));
    HIPCHK(L->dz1t.ensure(size_t(lws::kHid) * ld));
    HIPCHK(L->dz2t.ensure(size_t(lws::kHid) * ld));
    HIPCHK(L->dyt.ensure(size_t(lws::kOutPad) * ld));
    HIPCHK(L->bias_part.ensure(size_t(waves) * lws::kBiasCols));
    HIPCHK(L->loss_part.ensure(waves));
    int ksplit = std::min(ld, 512);
    int splits = (ld + ksplit - 1) / ksplit;
    HIPCHK(L->gw1.ensure(size_t(splits) * lws::kHid * lws::kInPad));
    HIPCHK(L->gw2.ensure(size_t(splits) * lws::kHid * lws::kHid));
    HIPCHK(L->gw3.ensure(size_t(splits) * lws::kOutPad * lws::kHid));
    lws::Batch T;
    T.xb = L->xb.p;
    T.tn = L->tn.p;
    T.begin = int(begin);
    T.count = count;
    T.ld = ld;
    T.gscale = 2.0f / (3.0f * float(count));
    T.xt = L->xt.p; T.h1t = L->h1t.p; T.h2t = L->h2t.p; T.dz1t = L->dz1t.p; T.dz2t = L->dz2t.p; T.dyt = L->dyt.p;
    T.bias_part = L->bias_part.p;
    T.loss_part = L->loss_part.p;
    lws::Model M = L->W.model();
    lws::k_fwd_bwd<<<(waves + 3) / 4, 256, 0, st>>>(M, T);
    lws::k_sum_loss<<<1, 64, 0, st>>>(L->loss_part.p, waves, L->loss_acc.p + slot);
    if (grads) {
        lws::GemmArgs G;
        G.job[0] = {L->dz1t.p, L->xt.p, L->gw1.p, lws::kHid, lws::kInPad};
        G.job[1] = {L->dz2t.p, L->h1t.p, L->gw2.p, lws::kHid, lws::kHid};
        G.job[2] = {L->dyt.p, L->h2t.p, L->gw3.p, lws::kOutPad, lws::kHid};
        for (int j = 0; j < 3; ++j) G.tiles[j] = (G.job[j].M / 16) * (G.job[j].N / 16);
        G.splits = splits;
        G.ksplit = ksplit;
        G.ld = ld;
        int gw = (G.tiles[0] + G.tiles[1] + G.tiles[2]) * splits;
        lws::k_grad_gemm<<<(gw + 3) / 4, 256, 0, st>>>(G);
        lws::OptArgs O;
        O.params = L->params.p; O.grad = L->grad.p; O.m = L->m.p; O.v = L->v.p;
        O.gw1 = L->gw1.p; O.gw2 = L->gw2.p; O.gw3 = L->gw3.p; O.bias_part = L->bias_part.p;
        O.splits = splits;
        O.waves = waves;
        O.adam = adam ? 1 : 0;
        O.lr = lr; O.beta1 = 0.9f; O.beta2 = 0.999f; O.eps = 1e-8f;
        if (adam) L->adam_t++;
        O.bc1 = float(1.0 - std::pow(0.9, double(std::max<int64_t>(L->adam_t, 1))));
        O.bc2 = float(1.0 - std::pow(0.999, double(std::max<int64_t>(L->adam_t, 1))));
        lws::k_optimizer<<<(lws::kParams + 255) / 256, 256, 0, st>>>(O);
        if (adam) {
            jh_status s = lws_refresh(L);
            if (s != JH_OK) return s;
        }
    }
    HIPCHK(hipGetLastError());
    return JH_OK;
}

jh_status lws_check_range(jh_lws* L, int64_t begin, int64_t count) {
    if (begin < 0 || count < 1 || begin + count > L->n)
        return fail(JH_ERR_INVALID_ARG, "sample range [%lld, %lld) outside the %lld resident samples", (long long)begin,
                    (long long)(begin + count), (long long)L->n);
    return JH_OK;
}

}  // namespace

extern "C" {

jh_status jh_lws_create(uint64_t seed, jh_lws** out) {
    if (!out) return fail(JH_ERR_INVALID_ARG, "out_model is NULL");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(JH_ERR_NO_DEVICE, "no HIP device visible");
    jh_lws* L = new jh_lws();
    HIPCHK(hipGetDevice(&L->device));
    {
        jh_status sc = lws_scratch_for(L->device, &L->scratch);
        if (sc != JH_OK) {
            delete L;
            return sc;
        }
    }
    HIPCHK(hipStreamCreateWithFlags(&L->st, hipStreamNonBlocking));
    HIPCHK(L->W.alloc());
    HIPCHK(L->params.ensure(lws::kParams));
    HIPCHK(L->grad.ensure(lws::kParams));
    HIPCHK(L->m.ensure(lws::kParams));
    HIPCHK(L->v.ensure(lws::kParams));
    HIPCHK(L->loss_acc.ensure(64));
    HIPCHK(hipMemset(L->m.p, 0, lws::kParams * sizeof(float)));
    HIPCHK(hipMemset(L->v.p, 0, lws::kParams * sizeof(float)));
    // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) for every weight and bias of a layer
    std::vector<float> p(lws::kParams);
    uint64_t st = seed;
    auto fill = [&](int off, int count, int fan_in) {
        float b = 1.0f / std::sqrt(float(fan_in));
        for (int i = 0; i < count; ++i) {
            double u = double(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
            p[off + i] = float((2.0 * u - 1.0) * b);
        }
    };
    fill(lws::oW1, lws::kHid * lws::kIn + lws::kHid, lws::kIn);
    fill(lws::oW2, lws::kHid * lws::kHid + lws::kHid, lws::kHid);
    fill(lws::oW3, lws::kOut * lws::kHid + lws::kOut, lws::kHid);
    HIPCHK(hipMemcpy(L->params.p, p.data(), p.size() * sizeof(float), hipMemcpyHostToDevice));
    for (int k = 0; k < lws::kIn; ++k) { L->hmean[k] = 0.0f; L->histd[k] = 1.0f; }
    for (int k = 0; k < 3; ++k) L->htscale[k] = 1.0f;
    jh_status s = lws_upload_norm(L);
    if (s == JH_OK) s = lws_refresh(L);
    if (s != JH_OK) {
        delete L;
        return s;
    }
    HIPCHK(hipStreamSynchronize(L->st));
    L->W.ready = true;
    *out = L;
    return JH_OK;
}

jh_status jh_lws_destroy(jh_lws* L) {
    if (!L) return fail(JH_ERR_INVALID_ARG, "model is NULL");
    (void)hipStreamSynchronize(L->st);
    L->W.release();
    DBuf<float>* f[] = {&L->params, &L->grad, &L->m, &L->v, &L->tn, &L->bias_part, &L->gw1, &L->gw2, &L->gw3,
                        &L->tmp_f, &L->tmp_t, &L->out};
    for (auto* b : f) b->release();
    DBuf<uint16_t>* h[] = {&L->xb, &L->xt, &L->h1t, &L->h2t, &L->dz1t, &L->dz2t, &L->dyt};
    for (auto* b : h) b->release();
    L->loss_part.release();
    L->loss_acc.release();
    (void)hipStreamDestroy(L->st);
    delete L;
    return JH_OK;
}

jh_status jh_lws_set_params(jh_lws* L, const float* params, int32_t count) {
    if (!L || !params || count != lws::kParams)
        return fail(JH_ERR_INVALID_ARG, "expected %d parameters", lws::kParams);
    for (int i = 0; i < count; ++i)
        if (!std::isfinite(params[i])) return fail(JH_ERR_NON_FINITE, "parameter %d is not finite", i);
    HIPCHK(hipMemcpyAsync(L->params.p, params, count * sizeof(float), hipMemcpyHostToDevice, L->st));
    HIPCHK(hipMemsetAsync(L->m.p, 0, lws::kParams * sizeof(float), L->st));
    HIPCHK(hipMemsetAsync(L->v.p, 0, lws::kParams * sizeof(float), L->st));
    L->adam_t = 0;
    jh_status s = lws_refresh(L);
    if (s != JH_OK) return s;
    HIPCHK(hipStreamSynchronize(L->st));
    return JH_OK;
}

jh_status jh_lws_get_params(jh_lws* L, float* params, int32_t count) {
    if (!L || !params || count != lws::kParams)
        return fail(JH_ERR_INVALID_ARG, "expected %d parameters", lws::kParams);
    HIPCHK(hipMemcpyAsync(params, L->params.p, count * sizeof(float), hipMemcpyDeviceToHost, L->st));
    HIPCHK(hipStreamSynchronize(L->st));
    return JH_OK;
}

jh_status jh_lws_set_normalisation(jh_lws* L, const float* mean, const float* inv_std, const float* tscale) {
    if (!L || !mean || !inv_std || !tscale) return fail(JH_ERR_INVALID_ARG, "bad arguments");
    for (int k = 0; k < lws::kIn; ++k)
        if (!std::isfinite(mean[k]) || !std::isfinite(inv_std[k])) return fail(JH_ERR_NON_FINITE, "normalisation");
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(tscale[k]) || !(tscale[k] >= 0.0f))
            return fail(JH_ERR_INVALID_ARG, "target_scale must be finite and >= 0");
    memcpy(L->hmean, mean, sizeof(L->hmean));
    memcpy(L->histd, inv_std, sizeof(L->histd));
    memcpy(L->htscale, tscale, sizeof(L->htscale));
    jh_status s = lws_upload_norm(L);
    if (s != JH_OK) return s;
    HIPCHK(hipStreamSynchronize(L->st));
    return JH_OK;
}

jh_status jh_lws_set_data(jh_lws* L, const float* features, const float* targets, int64_t count) {
    if (!L || !features || !targets || count < 1 || count > (int64_t(1) << 28))
        return fail(JH_ERR_INVALID_ARG, "bad arguments");
    HIPCHK(L->tmp_f.ensure(size_t(count) * lws::kIn));
    HIPCHK(L->tmp_t.ensure(size_t(count) * 3));
    HIPCHK(L->xb.ensure(size_t(count) * lws::kInPad));
    HIPCHK(L->tn.ensure(size_t(count) * 3));
    HIPCHK(hipMemcpyAsync(L->tmp_f.p, features, size_t(count) * lws::kIn * sizeof(float), hipMemcpyHostToDevice, L->st));
    HIPCHK(hipMemcpyAsync(L->tmp_t.p, targets, size_t(count) * 3 * sizeof(float), hipMemcpyHostToDevice, L->st));
    lws::k_prepare_data<<<int((count + 255) / 256), 256, 0, L->st>>>(L->tmp_f.p, L->tmp_t.p, int(count), L->W.mean.p,
                                                                      L->W.istd.p, L->W.tscale.p, L->xb.p, L->tn.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(L->st));
    L->tmp_f.release();
    L->tmp_t.release();
    L->n = count;
    return JH_OK;
}

jh_status jh_lws_train(jh_lws* L, int64_t begin, int64_t count, int32_t epochs, int32_t batch, float lr,
                       double* epoch_loss) {
    if (!L) return fail(JH_ERR_INVALID_ARG, "model is NULL");
    jh_status s = lws_check_range(L, begin, count);
    if (s != JH_OK) return s;
    if (epochs < 1 || epochs > 64 || batch < 16 || batch > (1 << 16) || !(lr > 0.0f) || !std::isfinite(lr))
        return fail(JH_ERR_INVALID_ARG, "epochs 1..64, batch 16..65536, lr > 0");
    HIPCHK(hipMemsetAsync(L->loss_acc.p, 0, 64 * sizeof(double), L->st));
    for (int e = 0; e < epochs; ++e)
        for (int64_t b = begin; b < begin + count; b += batch) {
            int c = int(std::min<int64_t>(batch, begin + count - b));
            s = lws_batch(L, b, c, true, true, lr, e);
            if (s != JH_OK) return s;
        }
    std::vector<double> acc(64);
    HIPCHK(hipMemcpyAsync(acc.data(), L->loss_acc.p, 64 * sizeof(double), hipMemcpyDeviceToHost, L->st));
    HIPCHK(hipStreamSynchronize(L->st));
    if (epoch_loss)
        for (int e = 0; e < epochs; ++e) epoch_loss[e] = acc[e] / (3.0 * double(count));
    return JH_OK;
}

jh_status jh_lws_evaluate(jh_lws* L, int64_t begin, int64_t count, double* mse) {
    if (!L || !mse) return fail(JH_ERR_INVALID_ARG, "bad arguments");
    jh_status s = lws_check_range(L, begin, count);
    if (s != JH_OK) return s;
    HIPCHK(hipMemsetAsync(L->loss_acc.p, 0, sizeof(double), L->st));
    for (int64_t b = begin; b < begin + count; b += 65536) {
        s = lws_batch(L, b, int(std::min<int64_t>(65536, begin + count - b)), false, false, 0.0f, 0);
        if (s != JH_OK) return s;
    }
    double acc = 0;
    HIPCHK(hipMemcpyAsync(&acc, L->loss_acc.p, sizeof(double), hipMemcpyDeviceToHost, L->st));
    HIPCHK(hipStreamSynchronize(L->st));
    *mse = acc / (3.0 * double(count));
    return JH_OK;
}

jh_status jh_lws_gradient(jh_lws* L, int64_t begin, int64_t count, float* grad, double* loss) {
    if (!L || !grad) return fail(JH_ERR_INVALID_ARG, "bad arguments");
    jh_status s = lws_check_range(L, begin, count);
    if (s != JH_OK) return s;
    if (count > (1 << 20)) return fail(JH_ERR_INVALID_ARG, "gradient over at most 2^20 samples");
    HIPCHK(hipMemsetAsync(L->loss_acc.p, 0, sizeof(double), L->st));
    s = lws_batch(L, begin, int(count), true, false, 0.0f, 0);
    if (s != JH_OK) return s;
    double acc = 0;
    HIPCHK(hipMemcpyAsync(grad, L->grad.p, lws::kParams * sizeof(float), hipMemcpyDeviceToHost, L->st));
    HIPCHK(hipMemcpyAsync(&acc, L->loss_acc.p, sizeof(double), hipMemcpyDeviceToHost, L->st));
    HIPCHK(hipStreamSynchronize(L->st));
    if (loss) *loss = acc / (3.0 * double(count));
    return JH_OK;
}

jh_status jh_lws_predict(jh_lws* L, const float* features, int64_t count, float* outputs) {
    if (!L || !features || !outputs || count < 1 || count > (int64_t(1) << 26))
        return fail(JH_ERR_INVALID_ARG, "bad arguments");
    HIPCHK(L->tmp_f.ensure(size_t(count) * lws::kIn));
    HIPCHK(L->out.ensure(size_t(count) * 3));
    DBuf<uint16_t> xb;
    HIPCHK(xb.ensure(size_t(count) * lws::kInPad));
    HIPCHK(hipMemcpyAsync(L->tmp_f.p, features, size_t(count) * lws::kIn * sizeof(float), hipMemcpyHostToDevice, L->st));
    lws::k_prepare_data<<<int((count + 255) / 256), 256, 0, L->st>>>(L->tmp_f.p, nullptr, int(count), L->W.mean.p,
                                                                      L->W.istd.p, L->W.tscale.p, xb.p, nullptr);
    int waves = int((count + 15) / 16);
    lws::k_predict<<<(waves + 3) / 4, 256, 0, L->st>>>(L->W.model(), xb.p, int(count), L->out.p);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(outputs, L->out.p, size_t(count) * 3 * sizeof(float), hipMemcpyDeviceToHost, L->st);
    if (e == hipSuccess) e = hipStreamSynchronize(L->st);
    xb.release();
    if (e != hipSuccess) return fail(JH_ERR_HIP, "jh_lws_predict: %s", hipGetErrorString(e));
    return JH_OK;
}

jh_status jh_lws_predict_device(jh_lws* L, const float* features, int64_t count, float* outputs) {
    if (!L || !features || !outputs || count < 1) return fail(JH_ERR_INVALID_ARG, "bad arguments");
    if (count > JH_LWS_PREDICT_MAX_ROWS)
        return fail(JH_ERR_CAPACITY, "jh_lws_predict_device: %lld rows exceed the staging array's %d rows",
                    (long long)count, JH_LWS_PREDICT_MAX_ROWS);
    if (!L->scratch) return fail(JH_ERR_STATE, "jh_lws_predict_device: no staging array on this device");
    if (!is_device_memory(features, L->device) || !is_device_memory(outputs, L->device))
        return fail(JH_ERR_INVALID_ARG, "jh_lws_predict_device: features and outputs must be device memory of GPU %d",
                    L->device);
    int n = int(count);
    lws::k_prepare_data<<<(n + 255) / 256, 256, 0, L->st>>>(features, nullptr, n, L->W.mean.p, L->W.istd.p,
                                                            L->W.tscale.p, L->scratch, nullptr);
    int waves = (n + 15) / 16;
    lws::k_predict<<<(waves + 3) / 4, 256, 0, L->st>>>(L->W.model(), L->scratch, n, outputs);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(L->st));
    return JH_OK;
}

jh_status jh_world_set_warm_start_model(jh_world* w, jh_lws* L) {
    if (!w || !L) return fail(JH_ERR_INVALID_ARG, "bad arguments");
    HIPCHK(hipStreamSynchronize(L->st));
    HIPCHK(w->lws.alloc());
    struct { void* d; const void* s; size_t n; } cp[] = {
        {w->lws.w1.p, L->W.w1.p, size_t(lws::kHid) * lws::kInPad * 2}, {w->lws.w2.p, L->W.w2.p, size_t(lws::kHid) * lws::kHid * 2},
        {w->lws.w3.p, L->W.w3.p, size_t(lws::kOutPad) * lws::kHid * 2}, {w->lws.w2t.p, L->W.w2t.p, size_t(lws::kHid) * lws::kHid * 2},
        {w->lws.w3t.p, L->W.w3t.p, size_t(lws::kHid) * lws::kOutPad * 2}, {w->lws.b1.p, L->W.b1.p, lws::kHid * 4},
        {w->lws.b2.p, L->W.b2.p, lws::kHid * 4}, {w->lws.b3.p, L->W.b3.p, lws::kOutPad * 4},
        {w->lws.mean.p, L->W.mean.p, lws::kIn * 4}, {w->lws.istd.p, L->W.istd.p, lws::kIn * 4},
        {w->lws.tscale.p, L->W.tscale.p, 3 * 4}};
    for (auto& c : cp) HIPCHK(hipMemcpyAsync(c.d, c.s, c.n, hipMemcpyDeviceToDevice, w->st));
    HIPCHK(hipStreamSynchronize(w->st));
    w->lws.ready = true;
    return JH_OK;
}

jh_status jh_world_step_probe(jh_world* w, float dt, jh_probe* probe) {
    if (!w || !probe) return fail(JH_ERR_INVALID_ARG, "bad arguments");
    if (probe->iterations < 1 || probe->iterations > JH_PROBE_MAX_ITERATIONS)
        return fail(JH_ERR_INVALID_ARG, "iterations must be in 1..%d", JH_PROBE_MAX_ITERATIONS);
    if (probe->modes < 1 || probe->modes > 7) return fail(JH_ERR_INVALID_ARG, "modes must be a non-empty subset of 0..2");
    if (probe->capacity < 0 || (probe->capacity > 0 && (!probe->features || !probe->targets)))
        return fail(JH_ERR_INVALID_ARG, "capacity > 0 needs features and targets");
    if ((probe->modes & 2) && !w->lws.ready) return fail(JH_ERR_STATE, "learned probe needs a warm-start model");
    if (!std::isfinite(dt)) return fail(JH_ERR_NON_FINITE, "dt must be finite");
    if (!(dt >= 1.401298464e-45f)) return fail(JH_ERR_INVALID_ARG, "probe needs dt > 0");
    if (w->nb == 0) return fail(JH_ERR_STATE, "no bodies");
    if (w->cfg.warm_start == JH_WARM_START_LEARNED && !w->lws.ready)
        return fail(JH_ERR_STATE, "learned warm start selected but no model was set");
    return do_step(w, dt, probe);
}

jh_status jh_world_autotune(jh_world* w, float dt, int32_t rounds, jh_tuning* out) {
    if (!w || !out || rounds < 1 || rounds > 64) return fail(JH_ERR_INVALID_ARG, "bad arguments (rounds 1..64)");
    if (!std::isfinite(dt) || !(dt >= 1.401298464e-45f)) return fail(JH_ERR_INVALID_ARG, "autotune needs dt > 0");
    if (w->nb == 0) return fail(JH_ERR_STATE, "no bodies");
    if (w->cfg.warm_start == JH_WARM_START_LEARNED && !w->lws.ready)
        return fail(JH_ERR_STATE, "learned warm start selected but no model was set");
    memset(out, 0, sizeof(*out));
    int cu = w->cu_count;
    // candidates, de-duplicated after clamping to what can be co-resident
    auto uniq = [](std::vector<int> v, int hi) {
        for (int& x : v) x = std::max(1, std::min(x, hi));
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        return v;
    };
    // solver candidates (kind, workgroups); the barrier solver only tunes its own grid
    struct Cand { int kind, n, threads; };
    std::vector<Cand> sc;
    if (w->kind == SOLVER_BARRIER || !w->host_con.empty()) {  // constraints run in the barrier solver
        if (!w->host_con.empty()) w->kind = SOLVER_BARRIER;
        // narrow grids too: each grid barrier's cache write-back and invalidate grow with the workgroups
        for (int n : uniq({cu / 16, cu / 8, cu / 4, cu / 2, cu, 2 * cu}, w->coop_solve)) sc.push_back({SOLVER_BARRIER, n, 0});
    } else {
        for (int n : uniq({cu / 16, cu / 8, cu / 4, cu / 2, 3 * cu / 4, cu, 3 * cu / 2, 2 * cu}, w->flow_max))
            sc.push_back({SOLVER_FLOW, n, 0});
        for (int n : uniq({cu / 4, cu / 2, cu, 2 * cu}, w->lane_max)) sc.push_back({SOLVER_LANES, n, 0});
        // colour batches with a relaxed grid barrier each: wins when the batches are large (many worlds)
        for (int n : uniq({cu / 2, cu, 2 * cu}, w->coop_solve)) sc.push_back({SOLVER_BARRIER, n, 0});
        for (int T : {cu / 2, cu, 2 * cu})
            if (T >= 1 && tile_lds_bytes(w, T) > 0) sc.push_back({SOLVER_TILES, T, 0});
        // one workgroup per world: hopeless for a few large worlds (a 131k-body world takes ~0.5 s)
        if (w->n_worlds >= 64 || (w->n_worlds >= 1 && w->largest_world <= 16 * kSmallWorld))
            for (int k = 0; k < 4; ++k) sc.push_back({SOLVER_WORLDS, w->world_occ[k] * cu, 64 << k});
    }
    auto width = [&](int kind) -> int& {
        return kind == SOLVER_BARRIER ? w->coop_solve : kind == SOLVER_LANES ? w->lane_blocks
             : kind == SOLVER_TILES ? w->tiles : kind == SOLVER_WORLDS ? w->world_blocks : w->flow_blocks;
    };
    Cand def{w->kind, width(w->kind), w->world_threads};
    int keep[5];
    for (int k = 0; k < 5; ++k) keep[k] = width(k);
    int keep_threads = w->world_threads;
    auto apply = [&](Cand c) {
        w->kind = c.kind;
        width(c.kind) = c.n;
        if (c.kind == SOLVER_WORLDS) w->world_threads = c.threads;
    };
    // round-robin over the candidates so that the evolving scene weighs them equally
    {
        std::vector<double> tot(sc.size(), 0.0);
        for (int r = 0; r < rounds; ++r)
            for (size_t i = 0; i < sc.size(); ++i) {
                apply(sc[i]);
                jh_status st = do_step(w, dt, nullptr);
                if (st != JH_OK) return st;
                tot[i] += w->stats.ms_solve;
            }
        size_t bi = 0;
        for (size_t i = 1; i < sc.size(); ++i)
            if (tot[i] < tot[bi]) bi = i;
        if (getenv("JH_AUTOTUNE_VERBOSE"))
            for (size_t i = 0; i < sc.size(); ++i)
                fprintf(stderr, "autotune: solver %d x %d (%d threads): %.3f ms solve\n", sc[i].kind, sc[i].n,
                        sc[i].threads, tot[i] / rounds);
        for (int k = 0; k < 5; ++k) width(k) = keep[k];
        w->world_threads = keep_threads;
        apply(sc[bi]);
        out->ms_solve_best = tot[bi] / rounds;
        out->ms_solve_default = out->ms_solve_best;
        for (size_t i = 0; i < sc.size(); ++i)
            if (sc[i].kind == def.kind && sc[i].n == def.n && (def.kind != SOLVER_WORLDS || sc[i].threads == def.threads))
                out->ms_solve_default = tot[i] / rounds;
    }
    std::vector<int> cc = uniq({16, 32, 64, 128, 256, 512}, w->coop_colour);
    int colour_default = w->colour_blocks;
    {
        std::vector<double> tot(cc.size(), 0.0);
        for (int r = 0; r < rounds; ++r)
            for (size_t i = 0; i < cc.size(); ++i) {
                w->colour_blocks = cc[i];
                jh_status st = do_step(w, dt, nullptr);
                if (st != JH_OK) return st;
                tot[i] += w->stats.ms_colouring;
            }
        size_t bi = 0;
        for (size_t i = 1; i < cc.size(); ++i)
            if (tot[i] < tot[bi]) bi = i;
        w->colour_blocks = cc[bi];
        out->ms_colour_best = tot[bi] / rounds;
        out->ms_colour_default = out->ms_colour_best;
        for (size_t i = 0; i < cc.size(); ++i)
            if (cc[i] == colour_default) out->ms_colour_default = tot[i] / rounds;
    }
    // pair traversal: wavefront at 32, 64 or 16 leaves per wave, or one thread per leaf; same pair set
    // either way (candidate 0 is the setting the world had)
    {
        const int tw[4] = {w->traverse_wave, 1, 1, 1 - w->traverse_wave};
        int q0 = w->traverse_qpw;
        const int tq[4] = {q0, q0 == 64 ? 32 : 64, q0 == 16 ? 32 : 16, q0};
        double tot[4] = {0.0, 0.0, 0.0, 0.0};
        for (int r = 0; r < rounds; ++r)
            for (int m = 0; m < 4; ++m) {
                w->traverse_wave = tw[m];
                w->traverse_qpw = tq[m];
                jh_status st = do_step(w, dt, nullptr);
                if (st != JH_OK) return st;
                tot[m] += w->stats.ms_broadphase;
            }
        int best = 0;
        for (int m = 1; m < 4; ++m)
            if (tot[m] < tot[best]) best = m;
        w->traverse_wave = tw[best];
        w->traverse_qpw = tq[best];
        out->ms_broadphase_best = tot[best] / rounds;
        out->ms_broadphase_default = tot[0] / rounds;
    }
    out->traversal = w->traverse_wave ? w->traverse_qpw : 0;
    out->solver = w->kind;
    out->solve_blocks = width(w->kind);
    out->colour_blocks = w->colour_blocks;
    out->candidates = int(sc.size() + cc.size() + 4);
    out->steps = rounds * out->candidates;
    return JH_OK;
}

}  // extern "C"

// Islands and sleeping (Collision/IslandHelper.cs, World.CheckDeactivation World.Step.cs:1289-1371,
// the sleep timer of RigidBody.Update RigidBody.cs:724-743) without an incremental island structure.
//
// The reference keeps islands incrementally: a new contact merges two islands (flags ORed), a removed
// one splits an island with a two-sided search (flags copied to both halves), and CheckDeactivation
// visits the islands on the active list.  Here the islands of each CheckDeactivation are recomputed
// as the connected components of the contact and constraint graph over the non-static bodies
// (lock-free union-find, root = smallest body index), and the island flags are rebuilt from what
// happened since the previous CheckDeactivation, on the components of the graph at the moment each
// event happened:
//   G  graph at this check (arbiters after this step's Detect, the broken ones still waiting for
//      RemoveBrokenArbiters, constraints);  B  arbiters created by this step's Detect;
//   H = G - B  graph after the last RemoveBrokenArbiters (the islands RigidBody.Update marked).
// Marks (MarkedAsActive): an H component holding a body that RigidBody.Update marked (kMark2), that
// the host woke (kTouch) or whose previous-G component lost an arbiter in RemoveBrokenArbiters (kM1:
// AddToActiveList before the split marks the whole component); a G component holding an endpoint
// of a B arbiter (AddToActiveList in HandleDeferredArbiters).  NeedsUpdate: kN1 (the previous-G
// component was asleep when RemoveBrokenArbiters put it on the active list), kTouch, and a B arbiter
// with an endpoint whose island was off the active list (AddConnection).  A G component is visited
// when it has a B endpoint or one of its H components was on the active list (awake, kM1 or
// kTouch).  Static bodies are their own islands; kSact is whether a static body's island is on the
// active list (a static body's IsActive, kAwake, is set only on creation).
#pragma once

#include "jh_kernels.h"

namespace jh {
namespace isl {

// per-body flags (sflags)
constexpr uint32_t kAwake = 1u;   // RigidBodyData.IsActive
constexpr uint32_t kSact = 2u;    // static body: its island is on the active list
constexpr uint32_t kMark2 = 4u;   // RigidBody.Update marked the body's island (for the next check)
constexpr uint32_t kTouch = 8u;   // ActivateBodyNextStep on a sleeping body (mark, NeedsUpdate, active list)
constexpr uint32_t kM1 = 16u;     // the body's component lost an arbiter in RemoveBrokenArbiters
constexpr uint32_t kN1 = 32u;     // ... while asleep (NeedsUpdate)
constexpr uint32_t kBnew = 64u;   // static body: endpoint of an arbiter created by this step's Detect
constexpr uint32_t kSwake = 128u; // static body woken by the host in this call: wake its partners
// per-root flags of H components (hfl) and G components (gfl)
constexpr uint32_t kHm = 1u, kHn = 2u, kHact = 4u;
constexpr uint32_t kGm = 1u, kGn = 2u, kGproc = 4u, kGe = 8u;

__device__ __forceinline__ int uf_load(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving; parents only ever move to smaller indices.
__device__ __forceinline__ int uf_find(int* parent, int x) {
    int p = uf_load(&parent[x]);
    while (p != x) {
        int gp = uf_load(&parent[p]);
        if (gp != p) atomicCAS(&parent[x], p, gp);
        x = p;
        p = gp;
    }
    return x;
}

// Hooks the larger root under the smaller one until both ends share a root.
__device__ __forceinline__ void uf_unite(int* parent, int a, int b) {
    while (true) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a < b) { int t = a; a = b; b = t; }
        int old = atomicCAS(&parent[a], a, b);
        if (old == a) return;
    }
}

__global__ void k_reset(int nb, int* __restrict__ parent, uint32_t* __restrict__ hfl, uint32_t* __restrict__ gfl) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        parent[b] = b;
        hfl[b] = 0u;
        gfl[b] = 0u;
    }
}

// anew[position of the arbiter] = 1 for each BeginCollide key of this step (akey is sorted).
__global__ void k_mark_new(const uint64_t* __restrict__ evb, const int* __restrict__ n_evb,
                           const uint64_t* __restrict__ akey, const int* __restrict__ a_count,
                           uint32_t* __restrict__ anew) {
    int n = *n_evb, A = *a_count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int f = lower_bound_u64(akey, A, evb[i]);
        if (f < A && akey[f] == evb[i]) anew[f] = 1u;
    }
}

// Arbiters as edges: want_new 0 the H edges (arbiters not created this step), 1 the B edges.
__global__ void k_union_arbiters(const jh_arbiter* __restrict__ arb, const int* __restrict__ a_count,
                                 const uint32_t* __restrict__ anew, uint32_t want_new, const int* __restrict__ motion,
                                 int* __restrict__ parent) {
    int A = *a_count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < A; i += gridDim.x * blockDim.x) {
        if ((anew ? anew[i] : 0u) != want_new) continue;
        int u = arb[i].body1, v = arb[i].body2;
        if (motion[u] == JH_MOTION_STATIC || motion[v] == JH_MOTION_STATIC) continue;
        uf_unite(parent, u, v);
    }
}

// Arbiter keys (shape index pairs) as edges: the broken arbiters that RemoveBrokenArbiters frees.
__global__ void k_union_keys(const uint64_t* __restrict__ keys, const int* __restrict__ n_ptr,
                             const ShapeRec* __restrict__ shapes, const int* __restrict__ motion,
                             int* __restrict__ parent) {
    int n = *n_ptr;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int u = shapes[uint32_t(keys[i] >> 32)].body, v = shapes[uint32_t(keys[i])].body;
        if (motion[u] == JH_MOTION_STATIC || motion[v] == JH_MOTION_STATIC) continue;
        uf_unite(parent, u, v);
    }
}

// Constraints connect their bodies whether enabled or not (IslandHelper.ConstraintCreated).
__global__ void k_union_constraints(const jh_constraint* __restrict__ con, int n, const int* __restrict__ motion,
                                    int* __restrict__ parent) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int u = con[i].body1, v = con[i].body2;
        if (motion[u] == JH_MOTION_STATIC || motion[v] == JH_MOTION_STATIC) continue;
        uf_unite(parent, u, v);
    }
}

__global__ void k_roots(int nb, int* __restrict__ parent, int* __restrict__ root) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) root[b] = uf_find(parent, b);
}

// Flags of the H components from their bodies.
__global__ void k_hflags(int nb, const int* __restrict__ motion, const uint32_t* __restrict__ sflags,
                         const int* __restrict__ hroot, uint32_t* __restrict__ hfl) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        if (motion[b] == JH_MOTION_STATIC) continue;
        uint32_t f = sflags[b];
        uint32_t h = ((f & (kMark2 | kTouch | kM1)) ? kHm : 0u) | ((f & (kTouch | kN1)) ? kHn : 0u) |
                     ((f & (kAwake | kTouch | kM1)) ? kHact : 0u);
        if (h) atomicOr(&hfl[hroot[b]], h);
    }
}

// Each H component's flags onto the G component holding it.
__global__ void k_h_to_g(int nb, const int* __restrict__ motion, const int* __restrict__ hroot,
                         const int* __restrict__ groot, const uint32_t* __restrict__ hfl, uint32_t* __restrict__ gfl) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        if (motion[b] == JH_MOTION_STATIC || hroot[b] != b) continue;
        uint32_t h = hfl[b];
        uint32_t g = ((h & kHm) ? kGm : 0u) | ((h & kHn) ? kGn : 0u) | ((h & kHact) ? kGproc : 0u);
        if (g) atomicOr(&gfl[groot[b]], g);
    }
}

// HandleDeferredArbiters (World.Step.cs:881-911) for the B arbiters: IslandHelper.AddConnection
// (NeedsUpdate when an endpoint's island is off the active list) and AddToActiveList of both islands.
__global__ void k_new_edges(const jh_arbiter* __restrict__ arb, const int* __restrict__ a_count,
                            const uint32_t* __restrict__ anew, const int* __restrict__ motion,
                            uint32_t* __restrict__ sflags, const int* __restrict__ hroot,
                            const uint32_t* __restrict__ hfl, const int* __restrict__ groot,
                            uint32_t* __restrict__ gfl) {
    int A = *a_count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < A; i += gridDim.x * blockDim.x) {
        if (!anew[i]) continue;
        int e[2] = {arb[i].body1, arb[i].body2};
        bool off = false;
        for (int k = 0; k < 2; ++k) {
            int x = e[k];
            bool on = motion[x] == JH_MOTION_STATIC ? (sflags[x] & (kSact | kM1 | kTouch)) != 0u
                                                    : (hfl[hroot[x]] & kHact) != 0u;
            off = off || !on;
        }
        for (int k = 0; k < 2; ++k) {
            int x = e[k];
            if (motion[x] == JH_MOTION_STATIC) atomicOr(&sflags[x], kBnew);
            else atomicOr(&gfl[groot[x]], kGm | kGproc | (off ? kGn : 0u));
        }
    }
}

// CheckDeactivation on the G components (and the static bodies' own islands), the solver's motion
// types (a sleeping body is static to the solver: its contacts and constraints are inactive, its
// integration skipped) and the count of sleeping non-static bodies.  Clears the flags consumed.
__global__ void k_decide(int nb, const int* __restrict__ motion, const int* __restrict__ groot,
                         const uint32_t* __restrict__ gfl, uint32_t* __restrict__ sflags, float4* __restrict__ slp,
                         int allow, int* __restrict__ smotion, int* __restrict__ asleep) {
    int local = 0;
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        uint32_t f = sflags[b];
        int m = motion[b];
        bool awake = (f & kAwake) != 0u;
        bool sact = (f & kSact) != 0u;
        if (m == JH_MOTION_STATIC) {
            if (sact || (f & (kM1 | kTouch | kBnew))) {
                bool marked = (f & (kM1 | kMark2 | kTouch | kBnew)) != 0u;
                if (allow && !marked) awake = sact = false;  // DeactivateBodyForSleep; island inactive
                else sact = true;
            }
        } else {
            uint32_t g = gfl[groot[b]];
            if (g & kGproc) {
                if (allow && !(g & kGm)) {
                    awake = false;  // velocities are kept (DeactivateBodyForSleep)
                } else if ((g & kGn) && !awake) {
                    awake = true;
                    slp[b].w = 0.0f;
                }
            }
            if (!awake) local++;
        }
        sflags[b] = (awake ? kAwake : 0u) | (sact ? kSact : 0u);
        smotion[b] = awake ? m : JH_MOTION_STATIC;
    }
    if (local) atomicAdd(asleep, local);
}

// The broken arbiters this step's RemoveBrokenArbiters frees, as marks for the next check: the
// component holding an endpoint is put on the active list before it splits (kM1; kN1 if asleep).
__global__ void k_e_marks(const uint64_t* __restrict__ keys, const int* __restrict__ n_ptr,
                          const ShapeRec* __restrict__ shapes, const int* __restrict__ motion,
                          const int* __restrict__ groot, uint32_t* __restrict__ gfl, uint32_t* __restrict__ sflags) {
    int n = *n_ptr;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int e[2] = {shapes[uint32_t(keys[i] >> 32)].body, shapes[uint32_t(keys[i])].body};
        for (int k = 0; k < 2; ++k) {
            if (motion[e[k]] == JH_MOTION_STATIC) atomicOr(&sflags[e[k]], kM1);
            else atomicOr(&gfl[groot[e[k]]], kGe);
        }
    }
}

__global__ void k_e_apply(int nb, const int* __restrict__ motion, const int* __restrict__ groot,
                          const uint32_t* __restrict__ gfl, uint32_t* __restrict__ sflags) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        if (motion[b] == JH_MOTION_STATIC || !(gfl[groot[b]] & kGe)) continue;
        uint32_t f = sflags[b];
        sflags[b] = f | kM1 | ((f & kAwake) ? 0u : kN1);
    }
}

// Host writes.  rebuild: the islands are rebuilt with every body awake and on the active list
// (World.CreateRigidBody; a static body's timer infinite, as MotionType = Static leaves it);
// otherwise every body is ActivateBodyNextStep(body, wakeUpStatic = true), as a Position write.
__global__ void k_upload(int nb, const int* __restrict__ motion, uint32_t* __restrict__ sflags,
                         float4* __restrict__ slp, int rebuild) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        bool st = motion[b] == JH_MOTION_STATIC;
        if (rebuild) {
            sflags[b] = kAwake | kTouch | (st ? kSact : 0u);
            slp[b].w = st ? __builtin_inff() : 0.0f;
        } else {
            slp[b].w = 0.0f;
            if (!(sflags[b] & kAwake)) sflags[b] |= kTouch;
        }
    }
}

// ActivateBodyNextStep for listed bodies (World.cs:696-720); a static body also wakes its partners
// (wakeUpStatic, k_wake_partners).  active == 0: DeactivateBodyNextStep (the timer to infinity).
__global__ void k_activate(const int* __restrict__ idx, int n, int active, int wake_static,
                           const int* __restrict__ motion, uint32_t* __restrict__ sflags, float4* __restrict__ slp) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int b = idx[i];
        if (!active) {
            slp[b].w = __builtin_inff();
            continue;
        }
        slp[b].w = 0.0f;
        uint32_t f = sflags[b];
        if (f & kAwake) continue;
        bool st = motion[b] == JH_MOTION_STATIC;
        if (st && !wake_static) continue;
        sflags[b] = f | kTouch | (st ? kSwake : 0u);
    }
}

__device__ __forceinline__ void wake_partner(int x, int y, const int* motion, uint32_t* sflags, float4* slp) {
    if (!(sflags[x] & kSwake)) return;
    slp[y].w = 0.0f;
    if (motion[y] != JH_MOTION_STATIC && !(sflags[y] & kAwake)) atomicOr(&sflags[y], kTouch);
}

// Bodies touching a woken static body through arbiters (held, and broken ones not yet freed) and
// constraints: ActivateBodyNextStep of each partner.
__global__ void k_wake_partners(const jh_arbiter* __restrict__ arb, const int* __restrict__ a_count,
                                const uint64_t* __restrict__ zkey, const int* __restrict__ z_count,
                                const ShapeRec* __restrict__ shapes, const jh_constraint* __restrict__ con, int ncon,
                                const int* __restrict__ motion, uint32_t* __restrict__ sflags,
                                float4* __restrict__ slp) {
    int A = a_count ? *a_count : 0, Z = z_count ? *z_count : 0;
    int n = A + Z + ncon;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int u, v;
        if (i < A) {
            u = arb[i].body1;
            v = arb[i].body2;
        } else if (i < A + Z) {
            uint64_t k = zkey[i - A];
            u = shapes[uint32_t(k >> 32)].body;
            v = shapes[uint32_t(k)].body;
        } else {
            u = con[i - A - Z].body1;
            v = con[i - A - Z].body2;
        }
        wake_partner(u, v, motion, sflags, slp);
        wake_partner(v, u, motion, sflags, slp);
    }
}

__global__ void k_clear_bits(int nb, uint32_t* __restrict__ sflags, uint32_t bits) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) sflags[b] &= ~bits;
}

__global__ void k_set_bits(int nb, uint32_t* __restrict__ sflags, uint32_t bits) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) sflags[b] |= bits;
}

}  // namespace isl
}  // namespace jh

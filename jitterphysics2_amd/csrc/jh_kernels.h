// Per-step kernels of the gfx950 path: body pack/unpack, broadphase (AABB, Morton, LBVH build,
// refit, traversal), narrowphase (MPR + manifold, EPA on the deep queue), contact cache
// (merge-join against the key-sorted arbiter array, AddContact), graph colouring, the persistent
// projected Gauss-Seidel solver, contact refresh and the end-of-step body update.
#pragma once
#include <hip/hip_cooperative_groups.h>

#include "jh_collide.h"
#include "jh_constraints.h"
#include "jh_epa.h"
#include "jh_epa_wave.h"
#include "jh_filter.h"
#include "jh_sweep.h"

namespace cg = cooperative_groups;

namespace jh {

constexpr uint32_t kNoColour = 0xffffffffu;  // arbiter without a colour from a previous step

// ---- Device state ---------------------------------------------------------------------------
struct Bodies {
    float4 *pos, *quat, *vel, *angvel, *dvel, *dangvel;
    float4* imv;          // InverseMassVector (xyz); w = 1 if the body has linear locks
    float4* iw0;          // InverseInertiaWorld M11 M12 M13 M22
    float2* iw1;          // InverseInertiaWorld M23 M33
    float4* il0;          // body-frame inverse inertia M11 M12 M13 M22
    float2* il1;          // M23 M33
    float4* mat;          // inverse mass, friction, restitution, linear damping multiplier
    float4* mat2;         // angular damping multiplier, EnableSpeculativeContacts (1 or 0), -, -
    float4 *force, *torque;
    int* motion;          // MotionType
    int* gravity;         // AffectedByGravity
    int* world;           // sub-world id
    int* locks;           // locked MotionAxes (the complement of AllowedMotion), bits 0-5
    float4* slp;          // DeactivationThreshold linear^2, angular^2, DeactivationTime, sleep timer
    uint32_t* sflags;     // island flags (jh_islands.h); bit 0 = RigidBodyData.IsActive
};

struct StepTime {
    float dt, substep_dt, inv_dt, inv_substep_dt;
    int substeps;
};

// ---- Bounded waits ---------------------------------------------------------------------------
// Barrier state of a world (64 words, zeroed at creation): [0] arrivals, [32] generation, [33] abort.
// Generation and abort flag are read together with one 64-bit load, so a waiting workgroup sees an
// abort at its next poll.  Every wait of a persistent kernel (grid barrier, dataflow readiness poll)
// gives up after `spin` polls: it raises the abort flag, and every wave that sees the flag returns
// from its kernel, so a launch always drains.  The host reads the flag after the step and reports
// JH_ERR_TIMEOUT; the barrier state is then reset.
constexpr unsigned kSpinLimitDefault = 1u << 23;  // polls per wait (each >= 1 device-coherent round trip)
constexpr int kBarGen = 32, kBarAbort = 33;

__device__ __forceinline__ void raise_abort(unsigned* bar) {
    __hip_atomic_store(bar + kBarAbort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool abort_raised(unsigned* bar) {
    return __hip_atomic_load(bar + kBarAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
}
// Bookkeeping of one poll loop (n: polls so far, uniform across the wave): false when the wait must
// end, either because this wait exceeded the limit (the abort is raised here) or because another
// wave raised it.  The flag itself is read every 64 polls.
__device__ __forceinline__ bool keep_polling(unsigned* bar, unsigned& n, unsigned spin) {
    ++n;
    if (n >= spin) {
        raise_abort(bar);
        return false;
    }
    if ((n & 63u) == 0u) return !abort_raised(bar);
    return true;
}
__device__ __forceinline__ unsigned long long gen_and_abort(unsigned* bar, int order) {
    return order == __ATOMIC_ACQUIRE
               ? __hip_atomic_load((unsigned long long*)(bar + kBarGen), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)
               : __hip_atomic_load((unsigned long long*)(bar + kBarGen), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid barrier for a cooperative launch: one arrival per workgroup, sense by generation counter.
// Returns false (uniformly over the workgroup) when the launch is aborting: the abort flag is set, a
// wave of this workgroup arrives with local_ok false (it gave up a wait, and the flag is raised), or
// this wait gives up.  Every wave of the workgroup must call it, also one that gave up.  The
// generation snapshot decides whether to arrive at all, so it is complete before the arrival is
// issued; the last arriver resets the count before it publishes the next generation (release order).
__device__ __forceinline__ bool grid_barrier(unsigned* bar, unsigned spin, bool local_ok = true) {
    __shared__ unsigned bar_ok;
    bool failed = __syncthreads_or(!local_ok);
    if (threadIdx.x == 0) {
        unsigned* count = bar;
        unsigned* gen = bar + kBarGen;
        unsigned long long ga = gen_and_abort(bar, __ATOMIC_RELAXED);
        bool ok = !failed && (ga >> 32) == 0ull;
        if (ok) {
            unsigned g = unsigned(ga);
            unsigned arrived = __hip_atomic_fetch_add(count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (arrived == gridDim.x - 1) {
                __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(gen, g + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                unsigned n = 0;
                while (true) {
                    unsigned long long x = gen_and_abort(bar, __ATOMIC_ACQUIRE);
                    if (unsigned(x) != g) break;
                    if ((x >> 32) != 0ull || ++n >= spin) {
                        if (n >= spin) raise_abort(bar);
                        ok = false;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
        bar_ok = ok ? 1u : 0u;
    }
    __syncthreads();
    return bar_ok != 0u;
}

// Device-coherent 4/8-byte accesses (agent-scope relaxed atomics: sc1 loads and stores, served past
// the L2 of the issuing XCD), for data that workgroups share between relaxed grid barriers.
template <class T> __device__ __forceinline__ T cld(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ void cst(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Grid barrier for device-coherent accesses (cld / cst, coh_ld / coh_st, device-scope atomics): each wave
// waits for its own loads and stores to complete (vmcnt 0; the workgroup barrier alone does not wait for
// them), then one relaxed arrival per workgroup — no L2 write-back or invalidate.
// The generation snapshot feeds the abort branch, so it has been performed before the arrival is
// issued (a snapshot taken after another workgroup completed the barrier would never be left); the
// last arriver waits for its count reset to complete (vmcnt 0) before it publishes the next
// generation, so a released workgroup's next arrival cannot be wiped by a late reset.
__device__ __forceinline__ bool grid_barrier_coh(unsigned* bar, unsigned spin, bool local_ok = true) {
    __shared__ unsigned bar_ok;
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0), expcnt and lgkmcnt left at their maxima
    bool failed = __syncthreads_or(!local_ok);
    if (threadIdx.x == 0) {
        unsigned* count = bar;
        unsigned* gen = bar + kBarGen;
        unsigned long long ga = gen_and_abort(bar, __ATOMIC_RELAXED);
        bool ok = !failed && (ga >> 32) == 0ull;
        if (ok) {
            unsigned g = unsigned(ga);
            unsigned arrived = __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (arrived == gridDim.x - 1) {
                __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                __builtin_amdgcn_s_waitcnt(0x0F70);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                __hip_atomic_store(gen, g + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                unsigned n = 0;
                while (true) {
                    unsigned long long x = gen_and_abort(bar, __ATOMIC_RELAXED);
                    if (unsigned(x) != g) break;
                    if ((x >> 32) != 0ull || ++n >= spin) {
                        if (n >= spin) raise_abort(bar);
                        ok = false;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
        bar_ok = ok ? 1u : 0u;
    }
    __syncthreads();
    return bar_ok != 0u;
}
// Slot in a list appended to by many threads: one global atomic per wave (ballot + prefix count)
// instead of one per element.  Call with the lanes that may append active; returns -1 for !pred.
__device__ __forceinline__ int wave_append(int* counter, bool pred) {
    unsigned long long m = __ballot(pred);
    if (!m) return -1;
    int lane = threadIdx.x & 63;
    int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, leader, 64);
    unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    return pred ? base + __popcll(m & below) : -1;
}

__device__ __forceinline__ sym3 ld_iw(const Bodies& B, int b) {
    float4 a = B.iw0[b];
    float2 c = B.iw1[b];
    return {a.x, a.y, a.z, a.w, c.x, c.y};
}

// ---- Bodies ---------------------------------------------------------------------------------
// RigidBody.RestrictLinearMotion / RestrictAngularMotion (RigidBody.cs:651-713) for locked axes
// `locks` (bits 0-2 linear x, y, z; 3-5 angular x, y, z).
__device__ __forceinline__ vec3 restrict_linear(vec3 v, int locks) {
    if (locks & 1) v.x = 0.0f;
    if (locks & 2) v.y = 0.0f;
    if (locks & 4) v.z = 0.0f;
    return v;
}
__device__ __forceinline__ vec3 restrict_angular(vec3 v, int locks) {
    if (locks & 8) v.x = 0.0f;
    if (locks & 16) v.y = 0.0f;
    if (locks & 32) v.z = 0.0f;
    return v;
}
// The inverse inertia with locked rotations eliminated by Schur complements, axis by axis.
__device__ __forceinline__ sym3 restrict_inertia(sym3 I, int locks) {
    if ((locks & 56) == 56) return sym3{0, 0, 0, 0, 0, 0};
    if (locks & 8) {
        if (I.m11 > 0.0f) {
            float xy = I.m12 / I.m11, xz = I.m13 / I.m11;
            I.m22 -= xy * I.m12;
            I.m23 -= xy * I.m13;
            I.m33 -= xz * I.m13;
        }
        I.m11 = I.m12 = I.m13 = 0.0f;
    }
    if (locks & 16) {
        if (I.m22 > 0.0f) {
            float xy = I.m12 / I.m22, yz = I.m23 / I.m22;
            I.m11 -= xy * I.m12;
            I.m13 -= xy * I.m23;
            I.m33 -= yz * I.m23;
        }
        I.m12 = I.m22 = I.m23 = 0.0f;
    }
    if (locks & 32) {
        if (I.m33 > 0.0f) {
            float xz = I.m13 / I.m33, yz = I.m23 / I.m33;
            I.m11 -= xz * I.m13;
            I.m12 -= xz * I.m23;
            I.m22 -= yz * I.m23;
        }
        I.m13 = I.m23 = I.m33 = 0.0f;
    }
    return I;
}

// RigidBody.UpdateWorldInertia (RigidBody.cs:903-923) with the AllowedMotion restrictions.
__device__ __forceinline__ void update_world_inertia(const Bodies& B, int b) {
    int locks = B.locks[b];
    float ll = (locks & 7) ? 1.0f : 0.0f;
    if (B.motion[b] == JH_MOTION_DYNAMIC) {
        float4 l0 = B.il0[b];
        float2 l1 = B.il1[b];
        sym3 il{l0.x, l0.y, l0.z, l0.w, l1.x, l1.y};
        sym3 w = sym_rotate(il, mat_from_quat(ldq(B.quat[b])));
        if (locks & 56) w = restrict_inertia(w, locks);
        B.iw0[b] = make_float4(w.m11, w.m12, w.m13, w.m22);
        B.iw1[b] = make_float2(w.m23, w.m33);
        float im = B.mat[b].x;
        vec3 imv = restrict_linear(vec3{im, im, im}, locks);
        B.imv[b] = make_float4(imv.x, imv.y, imv.z, ll);
    } else {
        B.iw0[b] = make_float4(0, 0, 0, 0);
        B.iw1[b] = make_float2(0, 0);
        B.imv[b] = make_float4(0, 0, 0, ll);
    }
}

__device__ __forceinline__ void unpack_body(const jh_body& s, int b, const Bodies& B) {
    B.pos[b] = make_float4(s.position[0], s.position[1], s.position[2], 0);
    B.quat[b] = make_float4(s.orientation[0], s.orientation[1], s.orientation[2], s.orientation[3]);
    B.vel[b] = make_float4(s.velocity[0], s.velocity[1], s.velocity[2], 0);
    B.angvel[b] = make_float4(s.angular_velocity[0], s.angular_velocity[1], s.angular_velocity[2], 0);
    B.dvel[b] = make_float4(s.delta_velocity[0], s.delta_velocity[1], s.delta_velocity[2], 0);
    B.dangvel[b] = make_float4(s.delta_angular_velocity[0], s.delta_angular_velocity[1], s.delta_angular_velocity[2], 0);
    B.il0[b] = make_float4(s.inverse_inertia[0], s.inverse_inertia[1], s.inverse_inertia[2], s.inverse_inertia[3]);
    B.il1[b] = make_float2(s.inverse_inertia[4], s.inverse_inertia[5]);
    B.mat[b] = make_float4(s.inverse_mass, s.friction, s.restitution, s.linear_damping_multiplier);
    B.mat2[b] = make_float4(s.angular_damping_multiplier, s.enable_speculative_contacts ? 1.0f : 0.0f, 0, 0);
    B.force[b] = make_float4(s.force[0], s.force[1], s.force[2], 0);
    B.torque[b] = make_float4(s.torque[0], s.torque[1], s.torque[2], 0);
    B.motion[b] = s.motion_type;
    B.gravity[b] = s.affected_by_gravity;
    B.world[b] = s.world_id;
    int locks = s.locked_axes & 63;
    B.locks[b] = locks;
    if (locks) {  // the velocity and AllowedMotion setters clear locked components (RigidBody.cs:584-588)
        B.vel[b] = st3(restrict_linear(ld3(B.vel[b]), locks));
        B.dvel[b] = st3(restrict_linear(ld3(B.dvel[b]), locks));
        B.angvel[b] = st3(restrict_angular(ld3(B.angvel[b]), locks));
        B.dangvel[b] = st3(restrict_angular(ld3(B.dangvel[b]), locks));
    }
    B.slp[b] = make_float4(s.deactivation_threshold_linear_sq, s.deactivation_threshold_angular_sq,
                           s.deactivation_time, 0.0f);
    update_world_inertia(B, b);
}

__global__ void k_unpack_bodies(const jh_body* __restrict__ in, int n, Bodies B) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < n) unpack_body(in[b], b, B);
}

// Records in[k] for bodies idx[k] (distinct indices).
__global__ void k_unpack_listed(const jh_body* __restrict__ in, const int* __restrict__ idx, int n, Bodies B) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) unpack_body(in[k], idx[k], B);
}

__global__ void k_pack_bodies(Bodies B, int n, jh_body* __restrict__ out) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    jh_body& o = out[b];
    float4 v;
    v = B.pos[b]; o.position[0] = v.x; o.position[1] = v.y; o.position[2] = v.z;
    v = B.quat[b]; o.orientation[0] = v.x; o.orientation[1] = v.y; o.orientation[2] = v.z; o.orientation[3] = v.w;
    v = B.vel[b]; o.velocity[0] = v.x; o.velocity[1] = v.y; o.velocity[2] = v.z;
    v = B.angvel[b]; o.angular_velocity[0] = v.x; o.angular_velocity[1] = v.y; o.angular_velocity[2] = v.z;
    v = B.dvel[b]; o.delta_velocity[0] = v.x; o.delta_velocity[1] = v.y; o.delta_velocity[2] = v.z;
    v = B.dangvel[b]; o.delta_angular_velocity[0] = v.x; o.delta_angular_velocity[1] = v.y; o.delta_angular_velocity[2] = v.z;
    v = B.il0[b]; o.inverse_inertia[0] = v.x; o.inverse_inertia[1] = v.y; o.inverse_inertia[2] = v.z; o.inverse_inertia[3] = v.w;
    float2 w = B.il1[b]; o.inverse_inertia[4] = w.x; o.inverse_inertia[5] = w.y;
    v = B.mat[b]; o.inverse_mass = v.x; o.friction = v.y; o.restitution = v.z; o.linear_damping_multiplier = v.w;
    float2 m2 = make_float2(B.mat2[b].x, B.mat2[b].y);
    o.angular_damping_multiplier = m2.x;
    v = B.force[b]; o.force[0] = v.x; o.force[1] = v.y; o.force[2] = v.z;
    v = B.torque[b]; o.torque[0] = v.x; o.torque[1] = v.y; o.torque[2] = v.z;
    o.motion_type = B.motion[b];
    o.affected_by_gravity = B.gravity[b];
    o.world_id = B.world[b];
    o.enable_speculative_contacts = m2.y != 0.0f;
    v = B.slp[b];
    o.deactivation_threshold_linear_sq = v.x;
    o.deactivation_threshold_angular_sq = v.y;
    o.deactivation_time = v.z;
    o.locked_axes = B.locks[b];
}

// RigidBodyData in the reference's 128-byte FP32 layout (RigidBody.cs:107-160).
__global__ void k_pack_rigidbodydata(Bodies B, int n, char* __restrict__ out, int stride) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    char* r = out + size_t(b) * stride;
    int* ri = (int*)r;
    float* rf = (float*)(r + 8);
    ri[0] = b;
    ri[1] = 0;
    float4 v;
    v = B.pos[b]; rf[0] = v.x; rf[1] = v.y; rf[2] = v.z;
    v = B.vel[b]; rf[3] = v.x; rf[4] = v.y; rf[5] = v.z;
    v = B.angvel[b]; rf[6] = v.x; rf[7] = v.y; rf[8] = v.z;
    v = B.dvel[b]; rf[9] = v.x; rf[10] = v.y; rf[11] = v.z;
    v = B.dangvel[b]; rf[12] = v.x; rf[13] = v.y; rf[14] = v.z;
    v = B.quat[b]; rf[15] = v.x; rf[16] = v.y; rf[17] = v.z; rf[18] = v.w;
    v = B.iw0[b]; rf[19] = v.x; rf[20] = v.y; rf[21] = v.z; rf[22] = v.w;
    float2 w = B.iw1[b]; rf[23] = w.x; rf[24] = w.y;
    v = B.imv[b]; rf[25] = v.x; rf[26] = v.y; rf[27] = v.z;
    // MotionType | IsActive | the locked axes (AllowedMotion's complement, bits 4..9)
    ((int*)(rf + 28))[0] = (B.motion[b] & 3) | ((B.sflags[b] & 1u) ? 4 : 0) | (B.locks[b] << 4);
}

// The body state of the reference's 128-byte FP32 RigidBodyData records (World.RawData): position,
// velocities, delta velocities, orientation, InverseInertiaWorld and InverseMassVector; the flags
// were checked on the host.
__global__ void k_unpack_rigidbodydata(const char* __restrict__ in, int n, int stride, Bodies B) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const float* rf = (const float*)(in + size_t(b) * stride + 8);
    B.pos[b] = make_float4(rf[0], rf[1], rf[2], 0.0f);
    B.vel[b] = make_float4(rf[3], rf[4], rf[5], 0.0f);
    B.angvel[b] = make_float4(rf[6], rf[7], rf[8], 0.0f);
    B.dvel[b] = make_float4(rf[9], rf[10], rf[11], 0.0f);
    B.dangvel[b] = make_float4(rf[12], rf[13], rf[14], 0.0f);
    B.quat[b] = make_float4(rf[15], rf[16], rf[17], rf[18]);
    B.iw0[b] = make_float4(rf[19], rf[20], rf[21], rf[22]);
    B.iw1[b] = make_float2(rf[23], rf[24]);
    B.imv[b] = make_float4(rf[25], rf[26], rf[27], 0.0f);
}

// ---- Broadphase -----------------------------------------------------------------------------
__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// World AABB per shape; bounds[0..2] = min centre (ordered), bounds[3..5] = max centre.  Bodies with
// speculative contacts get the box grown by bdt * |velocity| on every side, bdt = the length of the
// previous step (RigidBodyShape.UpdateWorldBoundingBox(BroadphaseStepDt), RigidBodyShape.cs:37-58).
__global__ void k_aabb(const ShapeRec* __restrict__ shapes, int ns, const float* __restrict__ verts, Bodies B,
                       float bdt, float4* __restrict__ amin, float4* __restrict__ amax,
                       uint32_t* __restrict__ bounds) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    vec3 c{0, 0, 0};
    bool valid = i < ns;
    if (valid) {
        ShapeRec s = shapes[i];
        vec3 mn, mx;
        shape_aabb(s, verts, ldq(B.quat[s.body]), ld3(B.pos[s.body]), mn, mx);
        if (B.mat2[s.body].y != 0.0f) {
            float swept = bdt * sqrtf(len_sq(ld3(B.vel[s.body])));
            mn = mn - vec3{swept, swept, swept};
            mx = mx + vec3{swept, swept, swept};
        }
        amin[i] = st3(mn);
        amax[i] = st3(mx);
        c = 0.5f * (mn + mx);
    }
    // block reduction of the centre bounds
    __shared__ uint32_t red[6];
    if (threadIdx.x < 6) red[threadIdx.x] = threadIdx.x < 3 ? 0xffffffffu : 0u;
    __syncthreads();
    if (valid) {
        atomicMin(&red[0], f2ord(c.x));
        atomicMin(&red[1], f2ord(c.y));
        atomicMin(&red[2], f2ord(c.z));
        atomicMax(&red[3], f2ord(c.x));
        atomicMax(&red[4], f2ord(c.y));
        atomicMax(&red[5], f2ord(c.z));
    }
    __syncthreads();
    if (threadIdx.x < 3) atomicMin(&bounds[threadIdx.x], red[threadIdx.x]);
    else if (threadIdx.x < 6) atomicMax(&bounds[threadIdx.x], red[threadIdx.x]);
}

__device__ __forceinline__ uint32_t expand_bits10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// 64-bit sort key: world id in the high word, 30-bit Morton code of the AABB centre in the low word.
__global__ void k_morton(const ShapeRec* __restrict__ shapes, int ns, const int* __restrict__ body_world,
                         const float4* __restrict__ amin, const float4* __restrict__ amax,
                         const uint32_t* __restrict__ bounds, uint64_t* __restrict__ keys,
                         uint32_t* __restrict__ vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    float4 a = amin[i], b = amax[i];
    float lo[3] = {ord2f(bounds[0]), ord2f(bounds[1]), ord2f(bounds[2])};
    float hi[3] = {ord2f(bounds[3]), ord2f(bounds[4]), ord2f(bounds[5])};
    float c[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        float ext = hi[k] - lo[k];
        float t = ext > 0.0f ? (c[k] - lo[k]) / ext : 0.0f;
        t = fminf(fmaxf(t * 1024.0f, 0.0f), 1023.0f);
        q[k] = uint32_t(t);
    }
    uint32_t m = (expand_bits10(q[0]) << 2) | (expand_bits10(q[1]) << 1) | expand_bits10(q[2]);
    keys[i] = (uint64_t(uint32_t(body_world[shapes[i].body])) << 32) | m;
    vals[i] = uint32_t(i);
}

// Karras (2012) LBVH over sorted keys; equal keys are split by index.  Children < 0 are leaves (~j).
__device__ __forceinline__ int lbvh_delta(const uint64_t* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = k[i], b = k[j];
    if (a == b) return 64 + __clz(uint32_t(i ^ j));
    return __clzll(a ^ b);
}

__global__ void k_lbvh_build(const uint64_t* __restrict__ keys, int n, int* __restrict__ left, int* __restrict__ right,
                             int2* __restrict__ range, int* __restrict__ parent, int* __restrict__ leaf_parent,
                             int* __restrict__ flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (lbvh_delta(keys, n, i, i + 1) - lbvh_delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = lbvh_delta(keys, n, i, i - d);
    int lmax = 2;
    while (lbvh_delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (lbvh_delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = lbvh_delta(keys, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (lbvh_delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + min(d, 0);
    int lo = min(i, j), hi = max(i, j);
    int lc = (lo == gamma) ? ~gamma : gamma;
    int rc = (hi == gamma + 1) ? ~(gamma + 1) : (gamma + 1);
    left[i] = lc;
    right[i] = rc;
    range[i] = make_int2(lo, hi);
    flags[i] = 0;
    if (lc < 0) leaf_parent[~lc] = i; else parent[lc] = i;
    if (rc < 0) leaf_parent[~rc] = i; else parent[rc] = i;
    if (i == 0) parent[0] = -1;
}

// Boxes shared between waves of the refit go through agent-coherent 64-bit accesses (relaxed
// atomics: they bypass the per-XCD L2), so no device-wide cache write-back or invalidate is needed.
__device__ __forceinline__ void coh_st4(float4* p, float4 v) {
    unsigned long long* q = reinterpret_cast<unsigned long long*>(p);
    __hip_atomic_store(q, (unsigned long long)__float_as_uint(v.x) | ((unsigned long long)__float_as_uint(v.y) << 32),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 1, (unsigned long long)__float_as_uint(v.z) | ((unsigned long long)__float_as_uint(v.w) << 32),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float4 coh_ld4(const float4* p) {
    unsigned long long* q = (unsigned long long*)(p);
    unsigned long long a = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned long long b = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return make_float4(__uint_as_float(uint32_t(a)), __uint_as_float(uint32_t(a >> 32)), __uint_as_float(uint32_t(b)),
                       __uint_as_float(uint32_t(b >> 32)));
}

// Leaf boxes in sorted order, then bottom-up union of child boxes (second arrival continues).  A
// thread publishes its box (coherent stores, waited for with s_waitcnt) before it bumps the node's
// arrival counter; the second arrival reads both children with coherent loads.
// Leaf j's boxes carry the leaf's identity in their w components (box tests read x, y, z only):
// lmin[j].w = the shape index, lmax[j].w = its body | dynamic << 30 | awake << 31, so a traversal
// tests a leaf with the two loads of its box and no dependent loads.
//
// Child records for k_traverse_wave: internal node c's two children in one 64-byte record plus 16
// bytes, so a traversal step is one round of loads:
//   nbox[4c + 0] = (left min xyz, left code)   nbox[4c + 1] = (left max xyz, 0)
//   nbox[4c + 2] = (right min xyz, right code) nbox[4c + 3] = (right max xyz, 0)
//   nrng[c] = (left a, left b, right a, right b)
// code: internal node index, or ~j for leaf j; (a, b): an internal child's leaf range, a leaf
// child's (shape index, packed body/dynamic/awake word).
constexpr uint32_t kLeafBody = 0x3fffffffu, kLeafDyn = 1u << 30, kLeafAwake = 1u << 31;

__global__ void k_lbvh_refit(int n, const uint32_t* __restrict__ order, const float4* __restrict__ amin,
                             const float4* __restrict__ amax, float4* __restrict__ lmin, float4* __restrict__ lmax,
                             const int* __restrict__ left, const int* __restrict__ right, const int* __restrict__ parent,
                             const int* __restrict__ leaf_parent, float4* __restrict__ nmin, float4* __restrict__ nmax,
                             int* __restrict__ flags, const ShapeRec* __restrict__ shapes,
                             const int* __restrict__ motion, const uint32_t* __restrict__ sflags,
                             const int2* __restrict__ range, float4* __restrict__ nbox, int4* __restrict__ nrng) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t s = order[j];
    int b = shapes[s].body;
    uint32_t info = uint32_t(b) | (motion[b] == JH_MOTION_DYNAMIC ? kLeafDyn : 0u) |
                    ((sflags == nullptr || (sflags[b] & 1u) != 0u) ? kLeafAwake : 0u);
    float4 mn = amin[s], mx = amax[s];
    mn.w = __uint_as_float(s);
    mx.w = __uint_as_float(info);
    coh_st4(&lmin[j], mn);
    coh_st4(&lmax[j], mx);
    if (n == 1) return;
    int node = leaf_parent[j];
    while (node >= 0) {
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        // the first child to arrive stops; the second sees both child boxes
        if (__hip_atomic_fetch_add(&flags[node], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) return;
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        int l = left[node], r = right[node];
        float4 a0 = coh_ld4(l < 0 ? &lmin[~l] : &nmin[l]);
        float4 a1 = coh_ld4(l < 0 ? &lmax[~l] : &nmax[l]);
        float4 b0 = coh_ld4(r < 0 ? &lmin[~r] : &nmin[r]);
        float4 b1 = coh_ld4(r < 0 ? &lmax[~r] : &nmax[r]);
        coh_st4(&nmin[node], make_float4(fminf(a0.x, b0.x), fminf(a0.y, b0.y), fminf(a0.z, b0.z), 0));
        coh_st4(&nmax[node], make_float4(fmaxf(a1.x, b1.x), fmaxf(a1.y, b1.y), fmaxf(a1.z, b1.z), 0));
        if (nbox) {  // child record (read only by a later launch)
            int2 gl = l < 0 ? make_int2(__float_as_int(a0.w), __float_as_int(a1.w)) : range[l];
            int2 gr = r < 0 ? make_int2(__float_as_int(b0.w), __float_as_int(b1.w)) : range[r];
            float4* rec = nbox + 4 * size_t(node);
            rec[0] = make_float4(a0.x, a0.y, a0.z, __int_as_float(l));
            rec[1] = make_float4(a1.x, a1.y, a1.z, 0.0f);
            rec[2] = make_float4(b0.x, b0.y, b0.z, __int_as_float(r));
            rec[3] = make_float4(b1.x, b1.y, b1.z, 0.0f);
            nrng[node] = make_int4(gl.x, gl.y, gr.x, gr.y);
        }
        node = parent[node];
    }
}

__device__ __forceinline__ bool box_disjoint(float4 amn, float4 amx, float4 bmn, float4 bmx) {
    // JBoundingBox.Disjoint (JBoundingBox.cs:314-318)
    return amx.x < bmn.x || amn.x > bmx.x || amx.y < bmn.y || amn.y > bmx.y || amx.z < bmn.z || amn.z > bmx.z;
}

__device__ __forceinline__ int lower_bound_u64(const uint64_t* __restrict__ a, int n, uint64_t key) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

constexpr int kTraverseStack = 96;

// One thread per sorted leaf whose body is dynamic: walks the LBVH and emits each pair of the
// reference's pair set once (DynamicTree.cs:266-286 + World.cs:447-455 + World.Detect.cs:158-160):
// different bodies of the same world, at least one dynamic, world AABBs not disjoint.  Pairs are
// appended to a per-wave LDS buffer (LDS atomics); the wave flushes it to the global list with one
// global atomic whenever it holds kPairFlush or more pairs (each loop turn adds at most 2 per lane,
// so 256 slots never overflow) and once more after the walk.  The list is sorted afterwards.
constexpr int kPairBuf = 256;
constexpr int kPairFlush = kPairBuf - 128;

__device__ __forceinline__ void flush_pairs(uint64_t* wbuf, int* wcnt, uint64_t* __restrict__ pairs,
                                            int* __restrict__ pair_count, int cap) {
    // called by the currently active lanes of one wave, all together
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    int c = __hip_atomic_load(wcnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    uint64_t act = __ballot(1);
    int lane = threadIdx.x & 63;
    int me = __popcll(act & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
    int nact = __popcll(act);
    int leader = __ffsll((long long)act) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(pair_count, c);
    base = __shfl(base, leader, 64);
    for (int k = me; k < c; k += nact)
        if (base + k < cap) pairs[base + k] = wbuf[k];
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (lane == leader) __hip_atomic_store(wcnt, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__global__ void __launch_bounds__(256) k_traverse(int n, const uint64_t* __restrict__ keys,
                                                  const uint32_t* __restrict__ order,
                                                  const ShapeRec* __restrict__ shapes, const int* __restrict__ motion,
                                                  const float4* __restrict__ lmin, const float4* __restrict__ lmax,
                                                  const int* __restrict__ left, const int* __restrict__ right,
                                                  const int2* __restrict__ range, const float4* __restrict__ nmin,
                                                  const float4* __restrict__ nmax, uint64_t* __restrict__ pairs,
                                                  int* __restrict__ pair_count, int cap, int* __restrict__ err,
                                                  const uint32_t* __restrict__ sflags) {
    __shared__ uint64_t sbuf[4][kPairBuf];
    __shared__ int scnt[4];
    int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) scnt[wv] = 0;
    __builtin_amdgcn_wave_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    uint64_t* wbuf = sbuf[wv];
    int* wcnt = &scnt[wv];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool active = i < n && n >= 2;
    uint32_t si = 0;
    int bi = 0;
    bool awake_i = true;
    if (active) {
        si = order[i];
        bi = shapes[si].body;
        active = motion[bi] == JH_MOTION_DYNAMIC;
        if (sflags) awake_i = (sflags[bi] & 1u) != 0u;
    }
    if (active) {
        uint64_t w = keys[i] >> 32;
        int lo = lower_bound_u64(keys, n, w << 32);
        int hi = lower_bound_u64(keys, n, (w + 1) << 32);
        float4 qmn = lmin[i], qmx = lmax[i];
        int stack[kTraverseStack];
        int sp = 0;
        int node = 0;
        while (true) {
            int ch[2] = {left[node], right[node]};
            for (int k = 0; k < 2; ++k) {
                int c = ch[k];
                if (c < 0) {
                    int j = ~c;
                    if (j < lo || j >= hi || j == i) continue;
                    float4 mn = lmin[j], mx = lmax[j];
                    if (box_disjoint(qmn, qmx, mn, mx)) continue;
                    uint32_t sj = __float_as_uint(mn.w), inf = __float_as_uint(mx.w);  // packed leaf record
                    if (int(inf & kLeafBody) == bi) continue;
                    if ((inf & kLeafDyn) != 0u && j < i) continue;
                    if (!awake_i && !(inf & kLeafAwake)) continue;  // both asleep (World.Detect.cs:158)
                    uint32_t a = min(si, sj), b = max(si, sj);
                    int slot = __hip_atomic_fetch_add(wcnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                    wbuf[slot] = (uint64_t(a) << 32) | b;
                } else {
                    int2 r = range[c];
                    if (r.y < lo || r.x >= hi) continue;
                    if (box_disjoint(qmn, qmx, nmin[c], nmax[c])) continue;
                    if (sp < kTraverseStack) stack[sp++] = c;
                    else atomicOr(err, 1);
                }
            }
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            if (__hip_atomic_load(wcnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) >= kPairFlush)
                flush_pairs(wbuf, wcnt, pairs, pair_count, cap);
            if (sp == 0) break;
            node = stack[--sp];
        }
    }
    // all lanes of the wave are back together here
    flush_pairs(wbuf, wcnt, pairs, pair_count, cap);
}

// Wavefront traversal: the 64 lanes of a wave hold 64 consecutive leaves of the Morton order
// (neighbouring boxes) and walk the LBVH together.  At each internal node the wave loads the node's
// child record (scalar loads: the node is the same for the whole wave) and every lane tests its query
// against both children; a leaf child is settled at once by each lane's pair rules, an internal
// child is entered when any lane's query overlaps it (a ballot).  When both children are entered the
// right one waits on a wave-shared stack in LDS.  No lane diverges and no lane keeps a stack in
// private memory.  Same pair rules as k_traverse (the list is sorted afterwards).
constexpr int kWaveStack = 128;

__global__ void __launch_bounds__(256) k_traverse_wave(int n, const uint64_t* __restrict__ keys,
                                                       const float4* __restrict__ lmin,
                                                       const float4* __restrict__ lmax,
                                                       const float4* __restrict__ nbox,
                                                       const int4* __restrict__ nrng, uint64_t* __restrict__ pairs,
                                                       int* __restrict__ pair_count, int cap, int* __restrict__ err,
                                                       int qpw) {
    __shared__ uint64_t sbuf[4][kPairBuf];
    __shared__ int scnt[4];
    __shared__ int sstk[4][kWaveStack];
    int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) scnt[wv] = 0;
    __builtin_amdgcn_wave_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    uint64_t* wbuf = sbuf[wv];
    int* wcnt = &scnt[wv];
    int* stk = sstk[wv];
    // qpw consecutive leaves per wave (64, 32 or 16): fewer leaves per wave give shorter walks and more
    // waves to hide the node loads' latency
    int lane = threadIdx.x & 63;
    int i = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * qpw + lane;
    bool active = lane < qpw && i < n && n >= 2;
    float4 qmn = make_float4(0, 0, 0, 0), qmx = qmn;
    uint32_t si = 0, bi = 0;
    bool awake_i = true;
    int lo = 0, hi = 0;
    if (active) {
        qmn = lmin[i];
        qmx = lmax[i];
        uint32_t inf = __float_as_uint(qmx.w);
        si = __float_as_uint(qmn.w);
        bi = inf & kLeafBody;
        awake_i = (inf & kLeafAwake) != 0u;
        active = (inf & kLeafDyn) != 0u;
    }
    if (active) {
        uint64_t w = keys[i] >> 32;
        lo = lower_bound_u64(keys, n, w << 32);
        hi = lower_bound_u64(keys, n, (w + 1) << 32);
    }
    if (__ballot(active) == 0ull) return;  // wave-uniform
    // one child: a leaf is settled here (per lane), an internal node returns whether the wave enters it
    auto child = [&](int code, float4 bmn, float4 bmx, int a, int b) -> bool {
        if (code < 0) {
            int j = ~code;
            bool hit = active && j >= lo && j < hi && j != i && !box_disjoint(qmn, qmx, bmn, bmx);
            uint32_t inf = uint32_t(b);
            // different bodies; two dynamic shapes once (from the lower leaf); not both asleep
            if (hit && (inf & kLeafBody) != bi && !((inf & kLeafDyn) != 0u && j < i) &&
                (awake_i || (inf & kLeafAwake) != 0u)) {
                uint32_t sj = uint32_t(a);
                uint32_t x = min(si, sj), y = max(si, sj);
                int slot = __hip_atomic_fetch_add(wcnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                wbuf[slot] = (uint64_t(x) << 32) | y;
            }
            return false;
        }
        bool hit = active && b >= lo && a < hi && !box_disjoint(qmn, qmx, bmn, bmx);
        return __ballot(hit) != 0ull;
    };
    int sp = 0;
    int node = 0;
    while (true) {
        const float4* rec = nbox + 4 * size_t(node);
        float4 r0 = rec[0], r1 = rec[1], r2 = rec[2], r3 = rec[3];
        int4 g = nrng[node];
        int cl = __float_as_int(r0.w), cr = __float_as_int(r2.w);
        bool dl = child(cl, r0, r1, g.x, g.y);
        bool dr = child(cr, r2, r3, g.z, g.w);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        // a step adds at most two pairs per lane: 127 + 128 < kPairBuf
        if (__hip_atomic_load(wcnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) >= kPairFlush)
            flush_pairs(wbuf, wcnt, pairs, pair_count, cap);
        int next;
        if (dl) {
            if (dr) {
                if (sp < kWaveStack) stk[sp++] = cr;
                else if ((threadIdx.x & 63) == 0) atomicOr(err, 1);
            }
            next = cl;
        } else if (dr) {
            next = cr;
        } else {
            if (sp == 0) break;
            next = stk[--sp];
        }
        node = __builtin_amdgcn_readfirstlane(next);
    }
    flush_pairs(wbuf, wcnt, pairs, pair_count, cap);
}

// ---- Narrowphase ----------------------------------------------------------------------------
struct DeepRec {
    int pair;
    int kind;  // kDeepPlain / kDeepFilter / kDeepRerun
    MVertex portal[4];
};

__device__ __forceinline__ void pair_frame(const ShapeRec* shapes, const float* verts, const Bodies& B, uint64_t key,
                                           MinkowskiPair& mk, quat& qa, quat& qb, vec3& pa, vec3& pb) {
    uint32_t s1 = uint32_t(key >> 32), s2 = uint32_t(key & 0xffffffffu);
    mk.a = ShapeView{&shapes[s1], verts};
    mk.b = ShapeView{&shapes[s2], verts};
    int b1 = shapes[s1].body, b2 = shapes[s2].body;
    qa = ldq(B.quat[b1]);
    qb = ldq(B.quat[b2]);
    pa = ld3(B.pos[b1]);
    pb = ld3(B.pos[b2]);
    // NarrowPhase.MprEpa (NarrowPhase.cs:910-912): B in A's frame
    mk.qb = qconjmul(qa, qb);
    vec3 pos = pb - pa;
    mk.pb = rot_inv(pos, qa);
}

__device__ __forceinline__ void write_manifold(jh_manifold& o, uint64_t key, const MprOut& r, quat qa, vec3 posA,
                                               const ManifoldBuf* m, int epa) {
    o.shape1 = uint32_t(key >> 32);
    o.shape2 = uint32_t(key & 0xffffffffu);
    o.flags = epa;
    // back to world space (NarrowPhase.cs:920-924)
    vec3 pa = rot(r.pa, qa) + posA;
    vec3 pb = rot(r.pb, qa) + posA;
    vec3 n = rot(r.normal, qa);
    o.normal[0] = n.x; o.normal[1] = n.y; o.normal[2] = n.z;
    o.penetration = r.pen;
    o.point_a[0] = pa.x; o.point_a[1] = pa.y; o.point_a[2] = pa.z;
    o.point_b[0] = pb.x; o.point_b[1] = pb.y; o.point_b[2] = pb.z;
    if (!m) {
        o.count = 0;
        return;
    }
    o.count = m->count;
    for (int i = 0; i < 4; ++i) {
        vec3 a = i < m->count ? m->mA[i] : vec3{0, 0, 0};
        vec3 b = i < m->count ? m->mB[i] : vec3{0, 0, 0};
        o.manifold_a[i][0] = a.x; o.manifold_a[i][1] = a.y; o.manifold_a[i][2] = a.z;
        o.manifold_b[i][0] = b.x; o.manifold_b[i][1] = b.y; o.manifold_b[i][2] = b.z;
    }
}

__device__ __forceinline__ void manifold_for(const MinkowskiPair& mk, quat qa, quat qb, vec3 posA, vec3 posB,
                                             const MprOut& r, int aux, ManifoldBuf& m) {
    vec3 pa = rot(r.pa, qa) + posA;
    vec3 pb = rot(r.pb, qa) + posA;
    vec3 n = rot(r.normal, qa);
    if (aux) {
        build_manifold(mk.a, mk.b, qa, qb, posA, posB, pa, pb, n, m);
    } else {
        m.count = 1;
        m.mA[0] = pa;
        m.mB[0] = pb;
    }
}

// Filtered pairs (TriangleEdgeCollisionFilter): the world-space answer is filtered before the
// manifold is built (World.Detect.cs:197-215).
__device__ __forceinline__ tef::WRes to_world(const MprOut& r, quat qa, vec3 posA) {
    tef::WRes w;
    w.pa = rot(r.pa, qa) + posA;
    w.pb = rot(r.pb, qa) + posA;
    w.n = rot(r.normal, qa);
    w.pen = r.pen;
    return w;
}

__device__ __forceinline__ void write_filtered(jh_manifold& o, uint64_t key, const tef::WRes& w, bool keep,
                                               const MinkowskiPair& mk, quat qa, quat qb, vec3 posA, vec3 posB,
                                               int aux, int epa) {
    o.shape1 = uint32_t(key >> 32);
    o.shape2 = uint32_t(key & 0xffffffffu);
    o.flags = epa;
    o.normal[0] = w.n.x; o.normal[1] = w.n.y; o.normal[2] = w.n.z;
    o.penetration = w.pen;
    o.point_a[0] = w.pa.x; o.point_a[1] = w.pa.y; o.point_a[2] = w.pa.z;
    o.point_b[0] = w.pb.x; o.point_b[1] = w.pb.y; o.point_b[2] = w.pb.z;
    ManifoldBuf m;
    m.count = 0;
    if (keep) {
        if (aux) {
            build_manifold(mk.a, mk.b, qa, qb, posA, posB, w.pa, w.pb, w.n, m);
        } else {
            m.count = 1;
            m.mA[0] = w.pa;
            m.mB[0] = w.pb;
        }
    }
    o.count = m.count;
    for (int i = 0; i < 4; ++i) {
        vec3 a = i < m.count ? m.mA[i] : vec3{0, 0, 0};
        vec3 b = i < m.count ? m.mB[i] : vec3{0, 0, 0};
        o.manifold_a[i][0] = a.x; o.manifold_a[i][1] = a.y; o.manifold_a[i][2] = a.z;
        o.manifold_b[i][0] = b.x; o.manifold_b[i][1] = b.y; o.manifold_b[i][2] = b.z;
    }
}

__device__ __forceinline__ tef::Ctx filter_ctx(int side, const ShapeRec* shapes, uint64_t key, const float* verts,
                                               quat qa, quat qb, vec3 pa, vec3 pb) {
    tef::Ctx c;
    c.tri = side == 1 ? &shapes[uint32_t(key >> 32)] : &shapes[uint32_t(key & 0xffffffffu)];
    c.verts = verts;
    c.q = side == 1 ? qa : qb;
    c.pos = side == 1 ? pa : pb;
    c.tri_is_b = side == 2;
    return c;
}

// Deep-queue entry kinds (DeepRec.kind): plain EPA refinement; EPA then the whole filter; EPA as the
// filter's re-run with threshold 0 (TriangleEdgeCollisionFilter.cs:186-210), then its second half.
enum { kDeepPlain = 0, kDeepFilter = 1, kDeepRerun = 2 };

// World.Detect (World.Detect.cs:125-218) per broadphase pair: MPR, then the manifold; pairs deeper
// than EpaPenetrationThreshold (NarrowPhase.cs:877) go to the EPA queue.  With the triangle edge
// filter on, filtered pairs run MPR with threshold 0 so the portal is kept for the filter's re-run;
// the MPR answer is the same for any threshold.
// Separated pairs with a body that has speculative contacts and a relative speed of at least the
// speculative velocity threshold (spec_thr_sq = its square) go to the sweep queue (World.Detect.cs:
// 167-176); spec_any = 0 when no body has speculative contacts.
// Narrowphase of one pair (World.Detect.cs:125-218): MPR, then the manifold, the EPA queue, the
// triangle edge filter or the speculative sweep queue.  mk may carry LDS copies of point clouds.
__device__ __forceinline__ void narrow_pair(int p, uint64_t key, MinkowskiPair& mk, quat qa, quat qb, vec3 pa, vec3 pb,
                                            const ShapeRec* __restrict__ shapes, const float* __restrict__ verts,
                                            const Bodies& B, int aux, int filter, jh_manifold* __restrict__ out,
                                            DeepRec* __restrict__ deep, int* __restrict__ deep_count, int spec_any,
                                            float spec_thr_sq, int* __restrict__ spec_queue,
                                            int* __restrict__ spec_count) {
    int side = filter ? tef::applies(*mk.a.s, *mk.b.s, B.motion) : 0;
    MprOut r;
    solve_mpr(mk, side ? 0.0f : 0.02f, r);
    if (!r.hit) {
        write_manifold(out[p], key, r, qa, pa, nullptr, 0);
        if (spec_any) {
            int b1 = mk.a.s->body, b2 = mk.b.s->body;
            if (B.mat2[b1].y != 0.0f || B.mat2[b2].y != 0.0f) {
                vec3 dv = ld3(B.vel[b2]) - ld3(B.vel[b1]);
                if (!(len_sq(dv) < spec_thr_sq)) spec_queue[atomicAdd(spec_count, 1)] = p;
            }
        }
        return;
    }
    if (!side) {
        if (r.deep) {
            int slot = atomicAdd(deep_count, 1);
            DeepRec& d = deep[slot];
            d.pair = p;
            d.kind = kDeepPlain;
            for (int k = 0; k < 4; ++k) d.portal[k] = r.portal[k];
        }
        ManifoldBuf m;
        manifold_for(mk, qa, qb, pa, pb, r, aux, m);
        write_manifold(out[p], key, r, qa, pa, &m, 0);
        return;
    }
    tef::Ctx c = filter_ctx(side, shapes, key, verts, qa, qb, pa, pb);
    tef::WRes w = to_world(r, qa, pa);
    bool keep;
    int kind = -1;
    if (r.pen > 0.02f) {  // MprEpa ran EPA already: the re-run repeats the same answer
        kind = kDeepFilter;
        keep = tef::filter_same(c, w);
    } else {
        vec3 tn, nn;
        int e = tef::eval_res(c, w, tn, nn);
        if (e == tef::Discard) {
            keep = false;
        } else if (e == tef::Keep) {
            keep = true;
        } else if (w.pen < 0.0f) {
            keep = tef::adjust(w, tn, nn);
        } else {
            if (r.deep) kind = kDeepRerun;  // the re-run refines with EPA (penetration > 0)
            keep = tef::after_rerun(c, w);  // answer if EPA fails: the MPR one again
        }
    }
    if (kind >= 0) {
        int slot = atomicAdd(deep_count, 1);
        DeepRec& d = deep[slot];
        d.pair = p;
        d.kind = kind;
        for (int k = 0; k < 4; ++k) d.portal[k] = r.portal[k];
    }
    write_filtered(out[p], key, w, keep, mk, qa, qb, pa, pb, aux, 0);
}

// Optionally (JH_NARROW_LDS=1), point clouds of up to kLdsCloud vertices are narrowphased by
// k_narrow_hulls: k_narrow leaves their pairs to it (hull_list), and it stages each pair's clouds in
// LDS once, so the support scans of MPR and the manifold read LDS instead of the L2 (a 32-point cloud
// is 384 bytes, scanned whole by every support call).  Bit-exact, but slower on the full form
// (narrowphase 8.4 against 6.0 ms): 768 bytes of LDS per lane leave 3 waves per CU, and the L2
// already serves the clouds, so it is off by default.
constexpr int kLdsCloud = 32;
__host__ __device__ __forceinline__ bool lds_cloud(const ShapeRec& s) {
    return s.type == JH_SHAPE_POINTCLOUD && s.vcnt <= kLdsCloud;
}

// Pairs grouped by the types of their two shapes, so a wave of the narrowphase runs one pair of
// support maps instead of every type's branch: class = 8 * type1 + type2; counts per class (LDS
// histogram per workgroup), an exclusive scan, then each pair's position in its class's range.  The
// order within a class is arbitrary: every result is written at the pair's own index.
constexpr int kPairClasses = 64;
__device__ __forceinline__ int pair_class(const ShapeRec* shapes, uint64_t key) {
    return (shapes[uint32_t(key >> 32)].type & 7) * 8 + (shapes[uint32_t(key & 0xffffffffu)].type & 7);
}
__global__ void __launch_bounds__(256) k_class_count(const uint64_t* __restrict__ pairs, const int* __restrict__ n_ptr,
                                                     const ShapeRec* __restrict__ shapes, int* __restrict__ counts) {
    __shared__ int h[kPairClasses];
    if (threadIdx.x < kPairClasses) h[threadIdx.x] = 0;
    __syncthreads();
    int n = *n_ptr;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) atomicAdd(&h[pair_class(shapes, pairs[p])], 1);
    __syncthreads();
    if (threadIdx.x < kPairClasses && h[threadIdx.x] > 0) atomicAdd(&counts[threadIdx.x], h[threadIdx.x]);
}
// One wave: base = exclusive scan of the counts; the counts are cleared for the next step.
__global__ void __launch_bounds__(64) k_class_scan(int* __restrict__ counts, int* __restrict__ base) {
    int l = threadIdx.x;
    int v = counts[l];
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (l >= o) x += y;
    }
    base[l] = x - v;
    counts[l] = 0;
}
__global__ void __launch_bounds__(256) k_class_scatter(const uint64_t* __restrict__ pairs, const int* __restrict__ n_ptr,
                                                       const ShapeRec* __restrict__ shapes, int* __restrict__ base,
                                                       int* __restrict__ perm) {
    __shared__ int h[kPairClasses], off[kPairClasses];
    if (threadIdx.x < kPairClasses) h[threadIdx.x] = 0;
    __syncthreads();
    int n = *n_ptr;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    int c = 0, r = 0;
    if (p < n) {
        c = pair_class(shapes, pairs[p]);
        r = atomicAdd(&h[c], 1);
    }
    __syncthreads();
    if (threadIdx.x < kPairClasses && h[threadIdx.x] > 0) off[threadIdx.x] = atomicAdd(&base[threadIdx.x], h[threadIdx.x]);
    __syncthreads();
    if (p < n) perm[off[c] + r] = p;
}

__global__ void __launch_bounds__(128) k_narrow(const uint64_t* __restrict__ pairs, const int* __restrict__ n_ptr,
                                                const ShapeRec* __restrict__ shapes, const float* __restrict__ verts,
                                                Bodies B, int aux, int filter, jh_manifold* __restrict__ out,
                                                DeepRec* __restrict__ deep, int* __restrict__ deep_count,
                                                int spec_any, float spec_thr_sq, int* __restrict__ spec_queue,
                                                int* __restrict__ spec_count, int* __restrict__ hull_list,
                                                int* __restrict__ hull_count, const int* __restrict__ perm) {
    int n = *n_ptr;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        int p = perm ? perm[t] : t;  // pairs grouped by shape types, or in key order
        uint64_t key = pairs[p];
        if (hull_list) {
            uint32_t s1 = uint32_t(key >> 32), s2 = uint32_t(key & 0xffffffffu);
            if (lds_cloud(shapes[s1]) || lds_cloud(shapes[s2])) {
                hull_list[atomicAdd(hull_count, 1)] = p;
                continue;
            }
        }
        MinkowskiPair mk;
        quat qa, qb;
        vec3 pa, pb;
        pair_frame(shapes, verts, B, key, mk, qa, qb, pa, pb);
        narrow_pair(p, key, mk, qa, qb, pa, pb, shapes, verts, B, aux, filter, out, deep, deep_count, spec_any,
                    spec_thr_sq, spec_queue, spec_count);
    }
}

// The pairs k_narrow left: one wavefront per workgroup, each lane's pair's point clouds copied to its
// own 768 bytes of LDS (two clouds of up to kLdsCloud vertices), then the same narrowphase.
__global__ void __launch_bounds__(64) k_narrow_hulls(const uint64_t* __restrict__ pairs,
                                                     const int* __restrict__ hull_list,
                                                     const int* __restrict__ hull_count,
                                                     const ShapeRec* __restrict__ shapes,
                                                     const float* __restrict__ verts, Bodies B, int aux, int filter,
                                                     jh_manifold* __restrict__ out, DeepRec* __restrict__ deep,
                                                     int* __restrict__ deep_count, int spec_any, float spec_thr_sq,
                                                     int* __restrict__ spec_queue, int* __restrict__ spec_count) {
    __shared__ float cloud[64][2][3 * kLdsCloud];
    int n = *hull_count;
    int t = threadIdx.x;
    for (int i = blockIdx.x * blockDim.x + t; i < n; i += gridDim.x * blockDim.x) {
        int p = hull_list[i];
        uint64_t key = pairs[p];
        MinkowskiPair mk;
        quat qa, qb;
        vec3 pa, pb;
        pair_frame(shapes, verts, B, key, mk, qa, qb, pa, pb);
        if (lds_cloud(*mk.a.s)) {
            const float* src = verts + 3 * size_t(mk.a.s->voff);
            for (int k = 0; k < 3 * mk.a.s->vcnt; ++k) cloud[t][0][k] = src[k];
            mk.a.lv = cloud[t][0];
        }
        if (lds_cloud(*mk.b.s)) {
            const float* src = verts + 3 * size_t(mk.b.s->voff);
            for (int k = 0; k < 3 * mk.b.s->vcnt; ++k) cloud[t][1][k] = src[k];
            mk.b.lv = cloud[t][1];
        }
        narrow_pair(p, key, mk, qa, qb, pa, pb, shapes, verts, B, aux, filter, out, deep, deep_count, spec_any,
                    spec_thr_sq, spec_queue, spec_count);
    }
}

// A successful EPA answer replaces the MPR one for the pair (filtered pairs go through the triangle
// edge filter first).
__device__ __forceinline__ void epa_finish(const DeepRec& d, uint64_t key, const ShapeRec* shapes, const float* verts,
                                           const Bodies& B, const MinkowskiPair& mk, quat qa, quat qb, vec3 pa,
                                           vec3 pb, const MprOut& r, int aux, jh_manifold* out) {
    if (d.kind == kDeepPlain) {
        ManifoldBuf m;
        manifold_for(mk, qa, qb, pa, pb, r, aux, m);
        write_manifold(out[d.pair], key, r, qa, pa, &m, 1);
        return;
    }
    int side = tef::applies(*mk.a.s, *mk.b.s, B.motion);
    tef::Ctx c = filter_ctx(side, shapes, key, verts, qa, qb, pa, pb);
    tef::WRes wr = to_world(r, qa, pa);
    bool keep = d.kind == kDeepFilter ? tef::filter_same(c, wr) : tef::after_rerun(c, wr);
    write_filtered(out[d.pair], key, wr, keep, mk, qa, qb, pa, pb, aux, 1);
}

// EPA over the deep queue: a successful EPA replaces the MPR answer and its manifold.  `work`
// holds gridDim.x * blockDim.x interleaved polytopes (epa::kBytesPerThread each).
__global__ void __launch_bounds__(64) k_epa(const uint64_t* __restrict__ pairs, const ShapeRec* __restrict__ shapes,
                                            const float* __restrict__ verts, Bodies B, int aux,
                                            jh_manifold* __restrict__ out, const DeepRec* __restrict__ deep,
                                            const int* __restrict__ deep_count, char* __restrict__ work) {
    int n = *deep_count;
    int T = gridDim.x * blockDim.x;
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    epa::Work w;
    w.bind(work, T, tid);
    for (int q = tid; q < n; q += T) {
        const DeepRec& d = deep[q];
        uint64_t key = pairs[d.pair];
        MinkowskiPair mk;
        quat qa, qb;
        vec3 pa, pb;
        pair_frame(shapes, verts, B, key, mk, qa, qb, pa, pb);
        for (int k = 0; k < 4; ++k) w.set_vertex(k, d.portal[k]);
        MprOut r;
        r.hit = 1;
        if (!epa::solve(w, mk, r.pa, r.pb, r.normal, r.pen)) continue;
        epa_finish(d, key, shapes, verts, B, mk, qa, qb, pa, pb, r, aux, out);
    }
}

// The wave-cooperative form: one wavefront (workgroup of 64) per deep pair, polytope in LDS, the
// per-wave HBM area of jh_epa_wave.h in `side`.  Same results as k_epa bit for bit.
// Result of the wave EPA for deep entry q; k_epa_finish turns it into the pair's manifold.
struct EpaRes {
    int ok;
    float pen;
    vec3 pa, pb, normal;
};

__global__ void __launch_bounds__(64, 6) k_epa_wave(const uint64_t* __restrict__ pairs, const ShapeRec* __restrict__ shapes,
                                                 const float* __restrict__ verts, Bodies B,
                                                 const DeepRec* __restrict__ deep, const int* __restrict__ deep_count,
                                                 char* __restrict__ side, EpaRes* __restrict__ res,
                                                 int* __restrict__ next) {
    __shared__ epw::Lds L;
    int lane = threadIdx.x;
    epw::Poly w;
    w.L = &L;
    w.bind(side + size_t(blockIdx.x) * epw::kSideBytes);
    int n = *deep_count;
    // the first pair of each wave by its index, then the next unclaimed pair from a counter: pairs take
    // 1 to 85 iterations, so a fixed round-robin leaves a long tail behind the unlucky waves
    for (int q = blockIdx.x; q < n;) {
        const DeepRec& d = deep[q];
        uint64_t key = pairs[d.pair];
        MinkowskiPair mk;
        quat qa, qb;
        vec3 pa, pb;
        pair_frame(shapes, verts, B, key, mk, qa, qb, pa, pb);
        for (int k = 0; k < 4; ++k) epw::set_vertex(w, k, d.portal[k], lane);
        EpaRes r;
        r.ok = epw::solve(w, mk, lane, r.pa, r.pb, r.normal, r.pen) ? 1 : 0;
        if (lane == 0) res[q] = r;
        epw::wave_sync();
        int t = 0;
        if (lane == 0) t = atomicAdd(next, 1);
        q = __builtin_amdgcn_readfirstlane(t) + int(gridDim.x);
    }
}

// Manifolds of the deep pairs from the wave EPA's results, one thread per pair (the manifold and the
// triangle filter are serial per pair; here every lane has a pair of its own).
__global__ void __launch_bounds__(128) k_epa_finish(const uint64_t* __restrict__ pairs,
                                                    const ShapeRec* __restrict__ shapes,
                                                    const float* __restrict__ verts, Bodies B, int aux,
                                                    jh_manifold* __restrict__ out, const DeepRec* __restrict__ deep,
                                                    const int* __restrict__ deep_count,
                                                    const EpaRes* __restrict__ res) {
    int n = *deep_count;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        EpaRes e = res[q];
        if (!e.ok) continue;  // EPA failed: the MPR answer stays
        const DeepRec& d = deep[q];
        uint64_t key = pairs[d.pair];
        MinkowskiPair mk;
        quat qa, qb;
        vec3 pa, pb;
        pair_frame(shapes, verts, B, key, mk, qa, qb, pa, pb);
        MprOut r;
        r.hit = 1;
        r.pa = e.pa;
        r.pb = e.pb;
        r.normal = e.normal;
        r.pen = e.pen;
        epa_finish(d, key, shapes, verts, B, mk, qa, qb, pa, pb, r, aux, out);
    }
}

// The speculative branch of World.Detect (World.Detect.cs:167-194) for the queued pairs, one thread
// per pair: a linear sweep (NarrowPhase.Sweep, NarrowPhase.cs:1100-1137); a time of impact inside the
// step (0 < toi <= dt) gives one contact point with penetration normal . (pA - pB) * specrel, after
// the triangle edge filter when it applies.  Flags bit 1 marks it for RegisterContact with
// SolveMode.Angular removed.
__global__ void __launch_bounds__(64) k_sweep(const uint64_t* __restrict__ pairs, const ShapeRec* __restrict__ shapes,
                                              const float* __restrict__ verts, Bodies B, int filter, float dt,
                                              float specrel, const int* __restrict__ spec_queue,
                                              const int* __restrict__ spec_count, jh_manifold* __restrict__ out) {
    int n = *spec_count;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        int p = spec_queue[q];
        uint64_t key = pairs[p];
        MinkowskiPair mk;
        quat qa, qb;
        vec3 pa, pb;
        pair_frame(shapes, verts, B, key, mk, qa, qb, pa, pb);
        vec3 va = ld3(B.vel[mk.a.s->body]), vb = ld3(B.vel[mk.b.s->body]);
        tef::WRes w;
        float toi;
        if (!sweep_pair(mk, qa, pa, va, vb, w.pa, w.pb, w.n, toi)) continue;
        if (toi > dt || toi == 0.0f) continue;
        w.pen = dot(w.n, w.pa - w.pb) * specrel;
        int side = filter ? tef::applies(*mk.a.s, *mk.b.s, B.motion) : 0;
        if (side) {
            // TriangleEdgeCollisionFilter.Filter (:149-323): a speculative answer (penetration < 0)
            // skips the re-run; otherwise the re-run's MPR finds the pair separated again: discard.
            tef::Ctx c = filter_ctx(side, shapes, key, verts, qa, qb, pa, pb);
            vec3 tn, nn;
            int e = tef::eval_res(c, w, tn, nn);
            if (e == tef::Discard) continue;
            if (e == tef::Edge && !(w.pen < 0.0f && tef::adjust(w, tn, nn))) continue;
        }
        jh_manifold& o = out[p];
        o.flags = 2;
        o.count = 1;
        o.normal[0] = w.n.x; o.normal[1] = w.n.y; o.normal[2] = w.n.z;
        o.penetration = w.pen;
        o.point_a[0] = w.pa.x; o.point_a[1] = w.pa.y; o.point_a[2] = w.pa.z;
        o.point_b[0] = w.pb.x; o.point_b[1] = w.pb.y; o.point_b[2] = w.pb.z;
        for (int i = 0; i < 4; ++i) {
            vec3 a = i == 0 ? w.pa : vec3{0, 0, 0};
            vec3 b = i == 0 ? w.pb : vec3{0, 0, 0};
            o.manifold_a[i][0] = a.x; o.manifold_a[i][1] = a.y; o.manifold_a[i][2] = a.z;
            o.manifold_b[i][0] = b.x; o.manifold_b[i][1] = b.y; o.manifold_b[i][2] = b.z;
        }
    }
}

// ---- Contact cache --------------------------------------------------------------------------
__global__ void k_collide_flags(const jh_manifold* __restrict__ man, const int* __restrict__ n_ptr,
                                uint32_t* __restrict__ flags) {
    int n = *n_ptr;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x)
        flags[p] = man[p].count > 0 ? 1u : 0u;
}

// Colliding pairs (in key order) and whether their arbiter exists (binary search of the sorted cache).
__global__ void k_collide_compact(const uint64_t* __restrict__ pairs, const int* __restrict__ n_ptr,
                                  const uint32_t* __restrict__ flags, const uint32_t* __restrict__ offs,
                                  const uint64_t* __restrict__ akey, const int* __restrict__ a_count,
                                  uint64_t* __restrict__ ckey, int* __restrict__ cpair, int* __restrict__ cfound,
                                  uint32_t* __restrict__ cnew) {
    int n = *n_ptr;
    int A = *a_count;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        if (!flags[p]) continue;
        int c = offs[p];
        uint64_t k = pairs[p];
        ckey[c] = k;
        cpair[c] = p;
        int f = lower_bound_u64(akey, A, k);
        bool found = f < A && akey[f] == k;
        cfound[c] = found ? f : -1;
        cnew[c] = found ? 0u : 1u;
    }
}

__global__ void k_new_compact(const uint64_t* __restrict__ ckey, const int* __restrict__ c_count,
                              const uint32_t* __restrict__ cnew, const uint32_t* __restrict__ noffs,
                              uint64_t* __restrict__ nkey) {
    int n = *c_count;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
        if (cnew[c]) nkey[noffs[c]] = ckey[c];
}

// Merge the key-sorted cache with the key-sorted new keys into the other buffer.
// Arbiter records are moved whole by groups of 16 lanes, as 8-byte words (a record is 65 of them),
// so a wave's loads and stores are contiguous.  Unused slots travel too: like the reference's
// contact memory they keep what their last contact left (a slot reused by AddContact starts from
// it), and the learned warm start's features read some of those fields before PrepareForIteration.
static_assert(sizeof(jh_arbiter) == 520 && sizeof(jh_contact) == 120, "arbiter record layout");
constexpr int kArbWords = 65;
__device__ __forceinline__ void copy_arbiter16(const jh_arbiter* src, jh_arbiter* dst, int sub) {
    const uint64_t* s = reinterpret_cast<const uint64_t*>(src);
    uint64_t* d = reinterpret_cast<uint64_t*>(dst);
    for (int q = sub; q < kArbWords; q += 16) d[q] = s[q];
}

__global__ void k_merge_existing(const jh_arbiter* __restrict__ ain, const uint64_t* __restrict__ akey,
                                 const int* __restrict__ a_count, const uint64_t* __restrict__ nkey,
                                 const int* __restrict__ n_new, jh_arbiter* __restrict__ aout,
                                 uint64_t* __restrict__ okey) {
    int A = *a_count, N = *n_new;
    int sub = threadIdx.x & 15;
    for (int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 4; i < A; i += (gridDim.x * blockDim.x) >> 4) {
        int pos = 0;
        if (sub == 0) {
            uint64_t k = akey[i];
            pos = i + lower_bound_u64(nkey, N, k);
            okey[pos] = k;
        }
        pos = __shfl(pos, 0, 16);
        copy_arbiter16(&ain[i], &aout[pos], sub);
    }
}

// ContactData.Init (Contact.cs:271-284) for each new arbiter; memContacts.Allocate(clear: true).
__global__ void k_merge_new(const uint64_t* __restrict__ akey, const int* __restrict__ a_count,
                            const uint64_t* __restrict__ nkey, const int* __restrict__ n_new,
                            const ShapeRec* __restrict__ shapes, Bodies B, float specrel,
                            jh_arbiter* __restrict__ aout, uint64_t* __restrict__ okey, int* __restrict__ out_count) {
    int A = *a_count, N = *n_new;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < N; j += gridDim.x * blockDim.x) {
        uint64_t k = nkey[j];
        int pos = j + lower_bound_u64(akey, A, k);
        jh_arbiter& a = aout[pos];
        a.shape1 = uint32_t(k >> 32);
        a.shape2 = uint32_t(k & 0xffffffffu);
        a.body1 = shapes[a.shape1].body;
        a.body2 = shapes[a.shape2].body;
        float4 m1 = B.mat[a.body1], m2 = B.mat[a.body2];
        a.usage_mask = 0u;
        a.mode = 0u;
        a.friction = fmaxr(m1.y, m2.y);
        a.restitution = fmaxr(m1.z, m2.z);
        a.speculative_relaxation_factor = specrel;
        a.reserved = kNoColour;
        uint64_t* cw = reinterpret_cast<uint64_t*>(&a.contacts[0]);
        for (int q = 0; q < int(sizeof(a.contacts) / 8); ++q) cw[q] = 0ull;
        okey[pos] = k;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *out_count = A + N;
}

// Contact.Initialize (Contact.cs:703-759)
__device__ __forceinline__ void contact_initialize(jh_arbiter& cd, jh_contact& c, const Bodies& B, vec3 p1, vec3 p2,
                                                   vec3 n, bool is_new) {
    vec3 x1 = ld3(B.pos[cd.body1]), x2 = ld3(B.pos[cd.body2]);
    quat q1 = ldq(B.quat[cd.body1]), q2 = ldq(B.quat[cd.body2]);
    vec3 rel1 = p1 - x1, rel2 = p2 - x2;
    vec3 pos1 = rot_inv(rel1, q1), pos2 = rot_inv(rel2, q2);
    c.relative_position1[0] = rel1.x; c.relative_position1[1] = rel1.y; c.relative_position1[2] = rel1.z;
    c.relative_position2[0] = rel2.x; c.relative_position2[1] = rel2.y; c.relative_position2[2] = rel2.z;
    c.position1[0] = pos1.x; c.position1[1] = pos1.y; c.position1[2] = pos1.z;
    c.position2[0] = pos2.x; c.position2[1] = pos2.y; c.position2[2] = pos2.z;
    if (!is_new) return;
    c.flag = 2u;  // NewContact
    c.accumulated[0] = c.accumulated[1] = c.accumulated[2] = 0.0f;
    vec3 v1 = ld3(B.vel[cd.body1]), v2 = ld3(B.vel[cd.body2]);
    vec3 dv = v2 - v1;
    if (cd.mode == 0 || (cd.mode & (JH_SOLVE_ANGULAR1 | JH_SOLVE_ANGULAR2)) != 0) {
        vec3 w1 = ld3(B.angvel[cd.body1]), w2 = ld3(B.angvel[cd.body2]);
        dv = v2 + cross(w2, rel2);
        dv = dv - (v1 + cross(w1, rel1));
    }
    float rel_nv = dot(dv, n);
    c.bias = 0.0f;
    if (rel_nv < -1.0f) c.bias = -cd.restitution * rel_nv;
    vec3 t1 = dv - n * rel_nv;
    float num = len_sq(t1);
    if (num > 1e-12f) {
        num = 1.0f / sqrtf(num);
        t1 = t1 * num;
    } else {
        t1 = create_orthonormal(n);
    }
    vec3 t2 = cross(t1, n);
    c.normal[0] = n.x; c.normal[1] = n.y; c.normal[2] = n.z;
    c.tangent1[0] = t1.x; c.tangent1[1] = t1.y; c.tangent1[2] = t1.z;
    c.tangent2[0] = t2.x; c.tangent2[1] = t2.y; c.tangent2[2] = t2.z;
}

__device__ __forceinline__ vec3 rel1_of(const jh_contact& c) {
    return {c.relative_position1[0], c.relative_position1[1], c.relative_position1[2]};
}

// CalcArea4Points (Contact.cs:418-432) as evaluated lane-wise by the vectorized selector (:469-527).
__device__ __forceinline__ float area4(vec3 p0, vec3 p1, vec3 p2, vec3 p3) {
    vec3 a0 = p0 - p1, a1 = p0 - p2, a2 = p0 - p3, b0 = p2 - p3, b1 = p1 - p3, b2 = p1 - p2;
    float c0 = len_sq(cross(a0, b0)), c1 = len_sq(cross(a1, b1)), c2 = len_sq(cross(a2, b2));
    return fmaxr(fmaxr(c0, c1), c2);
}

// ContactData.AddContact (Contact.cs:329-416) with SortCachedPoints (:529-553).
__device__ __forceinline__ void add_contact(jh_arbiter& cd, const Bodies& B, vec3 p1, vec3 p2, vec3 n) {
    if ((cd.usage_mask & 15u) == 15u) {
        vec3 rp1 = p1 - ld3(B.pos[cd.body1]);
        vec3 r0 = rel1_of(cd.contacts[0]), r1 = rel1_of(cd.contacts[1]), r2 = rel1_of(cd.contacts[2]),
             r3 = rel1_of(cd.contacts[3]);
        float a0 = area4(rp1, r1, r2, r3), a1 = area4(rp1, r0, r2, r3), a2 = area4(rp1, r0, r1, r3),
              a3 = area4(rp1, r0, r1, r2);
        const float eps = -0.0001f;
        float big = 0.0f;
        uint32_t mask = 0;
        if (a0 > big + eps) { big = a0; mask = 1; }
        if (a1 > big + eps) { big = a1; mask = 2; }
        if (a2 > big + eps) { big = a2; mask = 4; }
        if (a3 > big + eps) { mask = 8; }
        int slot = mask == 2 ? 1 : mask == 4 ? 2 : mask == 8 ? 3 : 0;
        contact_initialize(cd, cd.contacts[slot], B, p1, p2, n, false);
        cd.usage_mask |= mask;
        return;
    }
    int closest = -1;
    float dsq = 3.402823466e+38f;
    vec3 rel = p1 - ld3(B.pos[cd.body1]);
    for (int i = 0; i < 4; ++i) {
        if ((cd.usage_mask & (1u << i)) == 0) continue;
        float d = len_sq(rel1_of(cd.contacts[i]) - rel);
        if (d < dsq) {
            dsq = d;
            closest = i;
        }
    }
    if (dsq < 0.02f * 0.02f) {
        contact_initialize(cd, cd.contacts[closest], B, p1, p2, n, false);
        return;
    }
    for (int i = 0; i < 4; ++i) {
        if ((cd.usage_mask & (1u << i)) == 0) {
            contact_initialize(cd, cd.contacts[i], B, p1, p2, n, true);
            cd.usage_mask |= (1u << i);
            return;
        }
    }
}

// World.RegisterContact (World.Detect.cs:294-324) for each colliding pair.
__global__ void k_register(const uint64_t* __restrict__ ckey, const int* __restrict__ cpair,
                           const int* __restrict__ cfound, const uint32_t* __restrict__ noffs,
                           const int* __restrict__ c_count, const uint64_t* __restrict__ akey_old,
                           const int* __restrict__ a_count_old, const uint64_t* __restrict__ nkey,
                           const int* __restrict__ n_new, const jh_manifold* __restrict__ man, Bodies B,
                           int persistent, jh_arbiter* __restrict__ arb) {
    int n = *c_count, A = *a_count_old, N = *n_new;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        uint64_t k = ckey[c];
        int f = cfound[c];
        int pos = f >= 0 ? f + lower_bound_u64(nkey, N, k) : int(noffs[c]) + lower_bound_u64(akey_old, A, k);
        jh_arbiter& cd = arb[pos];  // in place: one colliding pair per arbiter
        const jh_manifold& m = man[cpair[c]];
        if (!persistent) cd.usage_mask = 0;
        // ContactData.ResetMode (Contact.cs:291-296); speculative contacts remove SolveMode.Angular
        // (World.Detect.cs:191-192)
        cd.mode = (B.motion[cd.body1] == JH_MOTION_DYNAMIC) ? (JH_SOLVE_LINEAR1 | JH_SOLVE_ANGULAR1) : 0;
        cd.mode |= (B.motion[cd.body2] == JH_MOTION_DYNAMIC) ? (JH_SOLVE_LINEAR2 | JH_SOLVE_ANGULAR2) : 0;
        if (m.flags & 2) cd.mode &= ~uint32_t(JH_SOLVE_ANGULAR1 | JH_SOLVE_ANGULAR2);
        vec3 n3{m.normal[0], m.normal[1], m.normal[2]};
        for (int e = 0; e < m.count; ++e) {
            vec3 a{m.manifold_a[e][0], m.manifold_a[e][1], m.manifold_a[e][2]};
            vec3 b{m.manifold_b[e][0], m.manifold_b[e][1], m.manifold_b[e][2]};
            add_contact(cd, B, a, b, n3);
        }
    }
}

// ---- Graph colouring (Jones-Plassmann order, first-fit colours) -----------------------------
constexpr int kMaxColours = 127;  // colour 127 = serial overflow batch

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Jones-Plassmann priority: hashed key in the high 40 bits, arbiter index in the low 24 (unique).
__device__ __forceinline__ unsigned long long colour_prio(const jh_arbiter& r, int a) {
    uint64_t h = mix64((uint64_t(r.shape1) << 32) | r.shape2);
    return (h & 0xffffffffff000000ull) | (uint64_t(a) & 0xffffffull);
}

struct ColourArgs {
    jh_arbiter* arb;     // jh_arbiter.reserved carries the arbiter's colour from step to step
    const int* a_count;
    const int* motion;   // the solver's motion types: a sleeping body counts as static
    int* colour;
    unsigned long long* body_max;  // [2][nb]: Jones-Plassmann maxima, double-buffered by round parity
    int nb;
    unsigned long long* mask_lo;
    unsigned long long* mask_hi;
    int* work;           // [2][A]: uncoloured arbiters, double-buffered by round parity
    int* wcount;         // [2] worklist sizes
    int* ncolours;
    int* hist;           // [128]
    unsigned* bar;       // grid barrier state and abort flag
    unsigned spin;       // polls before a wait gives up (raises the abort flag)
    int keep;            // reuse the colours of arbiters that persist from the previous step
};

// Graph colouring of the arbiters: edges are arbiters, nodes are non-static bodies (static bodies
// never change, so arbiters through them never conflict).  Arbiters keep their colour between
// steps: the bodies of an arbiter never change, so persisting colours stay a proper colouring, and
// only the others are coloured, in Jones-Plassmann rounds over a worklist (local priority maxima take
// the first colour free on both bodies; the losers form the next round's worklist).  Two grid
// barriers per round: the priority maxima of round r live in body_max[r & 1], and the decide sweep of
// round r clears body_max[(r + 1) & 1] for its bodies, which are the only ones round r + 1 touches.
// Colours >= 127 fall into a serial batch.
__global__ void __launch_bounds__(256) k_colour(ColourArgs P) {
    __shared__ int lhist[128];  // per-workgroup colour counts, added to P.hist once at the end
    int A = *P.a_count;
    int stride = gridDim.x * blockDim.x;
    int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (threadIdx.x < 128) lhist[threadIdx.x] = 0;
    unsigned long long* bmax[2] = {P.body_max, P.body_max + P.nb};
    int* work[2] = {P.work, P.work + A};
    for (int a = t0; a < A; a += stride) {
        const jh_arbiter& r = P.arb[a];
        bool active = (r.usage_mask & 15u) != 0 &&
                      (P.motion[r.body1] != JH_MOTION_STATIC || P.motion[r.body2] != JH_MOTION_STATIC);
        cst(&P.colour[a], active ? -1 : -2);
        cst(&bmax[0][r.body1], 0ull);
        cst(&bmax[0][r.body2], 0ull);
        cst(&P.mask_lo[r.body1], 0ull);
        cst(&P.mask_lo[r.body2], 0ull);
        cst(&P.mask_hi[r.body1], 0ull);
        cst(&P.mask_hi[r.body2], 0ull);
    }
    if (t0 < 128) cst(&P.hist[t0], 0);
    if (t0 == 0) { cst(&P.wcount[0], 0); cst(&P.wcount[1], 0); *P.ncolours = 0; }
    if (!grid_barrier_coh(P.bar, P.spin)) return;
    // persisting colours, and the initial worklist
    for (int a = t0; a < A; a += stride) {
        bool uncoloured = false;
        if (cld(&P.colour[a]) == -1) {
            const jh_arbiter& r = P.arb[a];
            uint32_t c = r.reserved;
            if (P.keep && c < uint32_t(kMaxColours)) {
                cst(&P.colour[a], int(c));
                unsigned long long bit = 1ull << (c & 63);
                unsigned long long* m1 = c < 64 ? P.mask_lo : P.mask_hi;
                if (P.motion[r.body1] != JH_MOTION_STATIC) atomicOr(&m1[r.body1], bit);
                if (P.motion[r.body2] != JH_MOTION_STATIC) atomicOr(&m1[r.body2], bit);
                atomicAdd(&lhist[c], 1);
            } else {
                uncoloured = true;
            }
        }
        int slot = wave_append(&P.wcount[0], uncoloured);
        if (uncoloured) cst(&work[0][slot], a);
    }
    if (!grid_barrier_coh(P.bar, P.spin)) return;
    for (int round = 0;; ++round) {
        int cur = round & 1, nxt = cur ^ 1;
        int n = __hip_atomic_load(&P.wcount[cur], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (n == 0) break;
        unsigned long long* bm = bmax[cur];
        unsigned long long* bn = bmax[nxt];
        for (int i = t0; i < n; i += stride) {
            int a = cld(&work[cur][i]);
            const jh_arbiter& r = P.arb[a];
            unsigned long long prio = colour_prio(r, a);
            if (P.motion[r.body1] != JH_MOTION_STATIC) atomicMax(&bm[r.body1], prio);
            if (P.motion[r.body2] != JH_MOTION_STATIC) atomicMax(&bm[r.body2], prio);
        }
        if (t0 == 0) cst(&P.wcount[nxt], 0);
        if (!grid_barrier_coh(P.bar, P.spin)) return;
        for (int i = t0; i < n; i += stride) {
            int a = cld(&work[cur][i]);
            const jh_arbiter& r = P.arb[a];
            unsigned long long prio = colour_prio(r, a);
            bool d1 = P.motion[r.body1] != JH_MOTION_STATIC, d2 = P.motion[r.body2] != JH_MOTION_STATIC;
            bool win = (!d1 || cld(&bm[r.body1]) == prio) && (!d2 || cld(&bm[r.body2]) == prio);
            cst(&bn[r.body1], 0ull);
            cst(&bn[r.body2], 0ull);
            int slot = wave_append(&P.wcount[nxt], !win);
            if (!win) {
                cst(&work[nxt][slot], a);
                continue;
            }
            unsigned long long lo = 0, hi = 0;
            if (d1) { lo |= cld(&P.mask_lo[r.body1]); hi |= cld(&P.mask_hi[r.body1]); }
            if (d2) { lo |= cld(&P.mask_lo[r.body2]); hi |= cld(&P.mask_hi[r.body2]); }
            int c;
            if (~lo) c = __ffsll((long long)~lo) - 1;
            else if (~hi) c = 64 + __ffsll((long long)~hi) - 1;
            else c = kMaxColours;
            if (c >= kMaxColours) c = kMaxColours;
            else if (c < 64) {  // the round's only winner on these bodies: plain read-modify-write
                if (d1) cst(&P.mask_lo[r.body1], cld(&P.mask_lo[r.body1]) | (1ull << c));
                if (d2) cst(&P.mask_lo[r.body2], cld(&P.mask_lo[r.body2]) | (1ull << c));
            } else {
                if (d1) cst(&P.mask_hi[r.body1], cld(&P.mask_hi[r.body1]) | (1ull << (c - 64)));
                if (d2) cst(&P.mask_hi[r.body2], cld(&P.mask_hi[r.body2]) | (1ull << (c - 64)));
            }
            cst(&P.colour[a], c);  // read at the end by the workgroup that owns index a
            atomicAdd(&lhist[c], 1);
        }
        if (!grid_barrier_coh(P.bar, P.spin)) return;
    }
    for (int a = t0; a < A; a += stride) {
        int c = cld(&P.colour[a]);
        P.arb[a].reserved = (c >= 0 && c < kMaxColours) ? uint32_t(c) : kNoColour;
    }
    __syncthreads();
    if (threadIdx.x < 128 && lhist[threadIdx.x]) atomicAdd(&P.hist[threadIdx.x], lhist[threadIdx.x]);
    if (!grid_barrier_coh(P.bar, P.spin)) return;
    if (t0 == 0) {
        int nc = 0;
        for (int c = 0; c < kMaxColours; ++c)
            if (cld(&P.hist[c]) > 0) nc = c + 1;
        *P.ncolours = nc;
    }
}

// Sort keys for the colour batches: (colour << 32) | arbiter index; inactive arbiters sort last.
// Also counts the contact points that the solve will visit.
// Solve-order keys: (colour << 2 | contacts - 1) << 32 | arbiter.  Within a colour the order is free
// (no two arbiters of one colour share a non-static body), so arbiters are grouped by contact count:
// the dataflow solver's 64-arbiter chunks then run the same number of contact rows in every lane.
// Uncoloured (unsolved) arbiters sort last.
__global__ void k_colour_keys(const int* __restrict__ colour, const jh_arbiter* __restrict__ arb,
                              const int* __restrict__ a_count, uint64_t* __restrict__ keys, int* __restrict__ contacts) {
    int A = *a_count;
    int local = 0;
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < A; a += gridDim.x * blockDim.x) {
        int c = colour[a];
        int n = __popc(arb[a].usage_mask & 15u);
        uint32_t cc = c < 0 ? 1023u : (uint32_t(c) << 2) | uint32_t(max(n, 1) - 1);
        keys[a] = (uint64_t(cc) << 32) | uint32_t(a);
        if (c >= 0) local += n;
    }
    if (local) atomicAdd(contacts, local);
}

// Colour offsets of the colour-sorted order.
__global__ void k_colour_offsets(const int* __restrict__ hist, int* __restrict__ offs) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        int s = 0;
        for (int c = 0; c < 128; ++c) {
            offs[c] = s;
            s += hist[c];
        }
        offs[128] = s;
    }
}

// ---- Solver ---------------------------------------------------------------------------------
// Body state a contact needs, held in registers while one arbiter's slots are processed (the
// reference updates the same RigidBodyData slot by slot; values are identical).
struct BodyReg {
    vec3 v, w, imv;
    sym3 I;
    bool ll;  // linear locks: the contact's linear mass is directional (Contact.cs:1055-1061)
};

__device__ __forceinline__ void load_static_part(const Bodies& B, int b, BodyReg& r) {
    float4 iv = B.imv[b];
    r.imv = vec3{iv.x, iv.y, iv.z};
    r.ll = iv.w != 0.0f;
    r.I = ld_iw(B, b);
}

// Contact.PrepareForIterationAccelerated (Contact.cs:1038-1180), no linear locks.
__device__ __forceinline__ void prepare_contact(const jh_arbiter& cd, jh_contact& c, BodyReg& B1, BodyReg& B2,
                                                vec3 x1, quat q1, vec3 x2, quat q2, const StepTime& ts) {
    uint32_t mode = cd.mode;
    bool skip = (c.flag & 4u) != 0;
    c.flag &= ~4u;
    float ws[3] = {skip ? 0.0f : c.accumulated[0], skip ? 0.0f : c.accumulated[1], skip ? 0.0f : c.accumulated[2]};
    vec3 pos1{c.position1[0], c.position1[1], c.position1[2]};
    vec3 pos2{c.position2[0], c.position2[1], c.position2[2]};
    vec3 cr1 = rot(pos1, q1);
    vec3 cr2 = rot(pos2, q2);
    vec3 imv{0.0f, 0.0f, 0.0f};
    if (mode & JH_SOLVE_LINEAR1) imv = B1.imv;
    if (mode & JH_SOLVE_LINEAR2) imv = imv + B2.imv;
    if ((c.flag & 2u) == 0) c.bias = 0.0f;
    vec3 r1 = cr1 + x1, r2 = cr2 + x2;
    vec3 dist = r1 - r2;
    float NX[3] = {c.normal[0], c.tangent1[0], c.tangent2[0]};
    float NY[3] = {c.normal[1], c.tangent1[1], c.tangent2[1]};
    float NZ[3] = {c.normal[2], c.tangent1[2], c.tangent2[2]};
    float k[3] = {imv.x, imv.x, imv.x};
    if (B1.ll || B2.ll)
        for (int l = 0; l < 3; ++l) k[l] = (NX[l] * imv.x) * NX[l] + (NY[l] * imv.y) * NY[l] + (NZ[l] * imv.z) * NZ[l];
    vec3 n{NX[0], NY[0], NZ[0]};
    float pen = dot(dist, n);
    if (pen < -0.02f) c.bias = pen * ts.inv_dt * cd.speculative_relaxation_factor;
    c.flag &= ~2u;
    float ktx[3] = {0, 0, 0}, kty[3] = {0, 0, 0}, ktz[3] = {0, 0, 0};
    vec3 lin{ws[0] * NX[0] + ws[1] * NX[1] + ws[2] * NX[2], ws[0] * NY[0] + ws[1] * NY[1] + ws[2] * NY[2],
             ws[0] * NZ[0] + ws[1] * NZ[1] + ws[2] * NZ[2]};
    if (mode & JH_SOLVE_LINEAR1) B1.v = B1.v - cmul(lin, B1.imv);
    if (mode & JH_SOLVE_ANGULAR1) {
        vec3 rp{c.relative_position1[0], c.relative_position1[1], c.relative_position1[2]};
        const sym3& I = B1.I;
        float e1[3], e2[3], e3[3];
        for (int l = 0; l < 3; ++l) {
            float rrx = rp.y * NZ[l] - rp.z * NY[l];
            float rry = rp.z * NX[l] - rp.x * NZ[l];
            float rrz = rp.x * NY[l] - rp.y * NX[l];
            e1[l] = I.m11 * rrx + I.m12 * rry + I.m13 * rrz;
            e2[l] = I.m12 * rrx + I.m22 * rry + I.m23 * rrz;
            e3[l] = I.m13 * rrx + I.m23 * rry + I.m33 * rrz;
        }
        vec3 ai{ws[0] * e1[0] + ws[1] * e1[1] + ws[2] * e1[2], ws[0] * e2[0] + ws[1] * e2[1] + ws[2] * e2[2],
                ws[0] * e3[0] + ws[1] * e3[1] + ws[2] * e3[2]};
        B1.w = B1.w - ai;
        for (int l = 0; l < 3; ++l) {
            ktx[l] = e2[l] * rp.z - e3[l] * rp.y;
            kty[l] = e3[l] * rp.x - e1[l] * rp.z;
            ktz[l] = e1[l] * rp.y - e2[l] * rp.x;
        }
    }
    if (mode & JH_SOLVE_LINEAR2) B2.v = B2.v + cmul(lin, B2.imv);
    if (mode & JH_SOLVE_ANGULAR2) {
        vec3 rp{c.relative_position2[0], c.relative_position2[1], c.relative_position2[2]};
        const sym3& I = B2.I;
        float f1[3], f2[3], f3[3];
        for (int l = 0; l < 3; ++l) {
            float rrx = rp.y * NZ[l] - rp.z * NY[l];
            float rry = rp.z * NX[l] - rp.x * NZ[l];
            float rrz = rp.x * NY[l] - rp.y * NX[l];
            f1[l] = I.m11 * rrx + I.m12 * rry + I.m13 * rrz;
            f2[l] = I.m12 * rrx + I.m22 * rry + I.m23 * rrz;
            f3[l] = I.m13 * rrx + I.m23 * rry + I.m33 * rrz;
        }
        vec3 ai{ws[0] * f1[0] + ws[1] * f1[1] + ws[2] * f1[2], ws[0] * f2[0] + ws[1] * f2[1] + ws[2] * f2[2],
                ws[0] * f3[0] + ws[1] * f3[1] + ws[2] * f3[2]};
        B2.w = B2.w + ai;
        for (int l = 0; l < 3; ++l) {
            ktx[l] = ktx[l] + (f2[l] * rp.z - f3[l] * rp.y);
            kty[l] = kty[l] + (f3[l] * rp.x - f1[l] * rp.z);
            ktz[l] = ktz[l] + (f1[l] * rp.y - f2[l] * rp.x);
        }
    }
    for (int l = 0; l < 3; ++l) {
        float kres = NX[l] * ktx[l] + NY[l] * kty[l] + NZ[l] * ktz[l];
        float kk = k[l] + kres;
        c.mass_normal_tangent[l] = kk > 0.0f ? 1.0f / kk : 0.0f;
    }
    float pb = 0.2f * ts.inv_dt * fmaxr(0.0f, pen - 0.01f);
    c.penalty_bias = fminr(pb, 100.0f);
}

// Contact.IterateAccelerated (Contact.cs:1183-1248)
__device__ __forceinline__ void iterate_contact(const jh_arbiter& cd, jh_contact& c, BodyReg& B1, BodyReg& B2,
                                                bool apply_bias) {
    uint32_t mode = cd.mode;
    vec3 rel1{c.relative_position1[0], c.relative_position1[1], c.relative_position1[2]};
    vec3 rel2{c.relative_position2[0], c.relative_position2[1], c.relative_position2[2]};
    vec3 dv = B2.v - B1.v;
    if (mode & (JH_SOLVE_ANGULAR1 | JH_SOLVE_ANGULAR2)) {
        dv = B2.v + cross(B2.w, rel2);
        dv = dv - (B1.v + cross(B1.w, rel1));
    }
    float bias = c.bias;
    if (apply_bias && c.penalty_bias > 0) bias = fmaxr(c.penalty_bias, c.bias);
    float NX[3] = {c.normal[0], c.tangent1[0], c.tangent2[0]};
    float NY[3] = {c.normal[1], c.tangent1[1], c.tangent2[1]};
    float NZ[3] = {c.normal[2], c.tangent1[2], c.tangent2[2]};
    float old[3] = {c.accumulated[0], c.accumulated[1], c.accumulated[2]};
    float bv[3] = {bias, 0.0f, 0.0f};
    float imp[3];
    for (int l = 0; l < 3; ++l) {
        float vd = NX[l] * dv.x + NY[l] * dv.y + NZ[l] * dv.z;
        imp[l] = c.mass_normal_tangent[l] * (bv[l] - vd);
    }
    float maxT = cd.friction * old[0];
    float lo[3] = {0.0f, -maxT, -maxT}, hi[3] = {3.402823466e+38f, maxT, maxT};
    for (int l = 0; l < 3; ++l) {
        float a = old[l] + imp[l];
        a = fminr(fmaxr(a, lo[l]), hi[l]);
        c.accumulated[l] = a;
        imp[l] = a - old[l];
    }
    vec3 lin{imp[0] * NX[0] + imp[1] * NX[1] + imp[2] * NX[2], imp[0] * NY[0] + imp[1] * NY[1] + imp[2] * NY[2],
             imp[0] * NZ[0] + imp[1] * NZ[1] + imp[2] * NZ[2]};
    if (mode & JH_SOLVE_LINEAR1) B1.v = B1.v - cmul(lin, B1.imv);
    if (mode & JH_SOLVE_ANGULAR1) B1.w = B1.w - sym_mul(cross(rel1, lin), B1.I);
    if (mode & JH_SOLVE_LINEAR2) B2.v = B2.v + cmul(lin, B2.imv);
    if (mode & JH_SOLVE_ANGULAR2) B2.w = B2.w + sym_mul(cross(rel2, lin), B2.I);
}

constexpr int kBufferDword3 = 0x00020000;  // raw buffer resource word 3 for gfx9-family targets
constexpr int kCohSc1 = 16;                // cache-policy operand: sc1

// Coherent body state for the barrier solver.  Within one launch, workgroups share the body velocities
// and the poses that integrate_velocities writes; with every access to them a device-coherent (sc1)
// 16-byte buffer access, served past the L2 of the issuing XCD, the grid barrier between batches only
// has to see each workgroup's accesses completed — no L2 write-back on arrival, no L2 invalidate on
// departure, which an agent-scope release/acquire barrier pays once per workgroup.  Coh = false keeps
// plain accesses behind the fenced barrier.
struct CohBodies {
    __amdgpu_buffer_rsrc_t vel, angvel, pos, quat;
};
__device__ __forceinline__ CohBodies coh_bodies(const Bodies& B, int nb) {
    return {__builtin_amdgcn_make_buffer_rsrc((void*)B.vel, 0, nb * 16, kBufferDword3),
            __builtin_amdgcn_make_buffer_rsrc((void*)B.angvel, 0, nb * 16, kBufferDword3),
            __builtin_amdgcn_make_buffer_rsrc((void*)B.pos, 0, nb * 16, kBufferDword3),
            __builtin_amdgcn_make_buffer_rsrc((void*)B.quat, 0, nb * 16, kBufferDword3)};
}
__device__ __forceinline__ float4 coh_ld(__amdgpu_buffer_rsrc_t r, int i) {
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, i * 16, 0, kCohSc1));
}
__device__ __forceinline__ void coh_st(__amdgpu_buffer_rsrc_t r, int i, float4 v) {
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, i * 16, 0, kCohSc1);
}
template <bool Coh> __device__ __forceinline__ float4 bld(const float4* a, __amdgpu_buffer_rsrc_t r, int i) {
    if constexpr (Coh) return coh_ld(r, i);
    else return a[i];
}
template <bool Coh> __device__ __forceinline__ void bst(float4* a, __amdgpu_buffer_rsrc_t r, int i, float4 v) {
    if constexpr (Coh) coh_st(r, i, v);
    else a[i] = v;
}
template <bool Coh> __device__ __forceinline__ bool sync_grid(unsigned* bar, unsigned spin) {
    if constexpr (Coh) return grid_barrier_coh(bar, spin);
    else return grid_barrier(bar, spin);
}

// phase 0: prepare, 1: iterate with bias, 2: iterate without bias (relaxation)
template <bool Coh = false>
__device__ __forceinline__ void arbiter_pass(jh_arbiter& a, BodyReg& B1, BodyReg& B2, const Bodies& B,
                                             const CohBodies& cb, int phase, const StepTime& ts) {
    if (phase == 0) {
        vec3 x1 = ld3(bld<Coh>(B.pos, cb.pos, a.body1)), x2 = ld3(bld<Coh>(B.pos, cb.pos, a.body2));
        quat q1 = ldq(bld<Coh>(B.quat, cb.quat, a.body1)), q2 = ldq(bld<Coh>(B.quat, cb.quat, a.body2));
        for (int s = 0; s < 4; ++s)
            if (a.usage_mask & (1u << s)) prepare_contact(a, a.contacts[s], B1, B2, x1, q1, x2, q2, ts);
    } else {
        for (int s = 0; s < 4; ++s)
            if (a.usage_mask & (1u << s)) iterate_contact(a, a.contacts[s], B1, B2, phase == 1);
    }
}

// Colour-ordered structure-of-arrays copy of the contact rows for the dataflow solver: position k of
// the solve order holds arbiter order[k]; row r is its r-th used contact slot (used slots compacted in
// slot order; usage mask (1 << n) - 1), and field f of row r lives at f_[(r * kSoaFields + f) * cap + k],
// so the 64 lanes of a chunk (consecutive positions) read each field with one contiguous access.
// Gathered from the arbiters after colouring, scattered back (the fields the solve writes) after it.
constexpr int kSoaFields = 29;  // bias, penalty, normal 3, tangent1 3, tangent2 3, mass 3, acc 3, rel1 3, rel2 3, pos1 3, pos2 3
enum { kFBias = 0, kFPen = 1, kFN = 2, kFT1 = 5, kFT2 = 8, kFMass = 11, kFAcc = 14, kFRel1 = 17, kFRel2 = 20,
       kFPos1 = 23, kFPos2 = 26 };
struct ContactSoA {
    int2* bodies;    // [cap] body1, body2
    uint2* um;       // [cap] usage mask, mode
    float2* fs;      // [cap] friction, speculative relaxation factor
    uint32_t* flag;  // [4][cap]
    float* f;        // [4][kSoaFields][cap]
    float4* mv;      // [4][3][cap] versioned mutable fields (k_solve_flow), see load_mutable
    int4* nd;        // [cap] k_solve_flow: update counts of body1 and body2 (-1: static to the solver) and the
                     // arbiter's rank in each body's update sequence
    int cap;
    __device__ __forceinline__ float& at(int s, int fld, int k) const {
        return f[(size_t(s) * kSoaFields + fld) * cap + k];
    }
};

// Constraints of the barrier solver: enabled constraints in batches with disjoint dynamic bodies.
struct ConArgs {
    jh_constraint* c;
    con::Work* work;     // what each constraint's prepare leaves for its iterations
    const int* order;    // constraint indices, batch by batch
    const int* offs;     // nbatch + 1 batch offsets into order
    int nbatch;
};

struct SolveArgs {
    jh_arbiter* arb;
    const uint64_t* order;   // (colour << 32) | arbiter, sorted
    const int* offs;         // 129 colour offsets
    const int* ncolours;
    const int2* rank;        // per order position: rank of the arbiter in each body's update sequence
    const int* deg;          // per body: arbiters touching it
    int* queue;              // k_solve_flow: chunk counters, 2 per substep
    int qmode;               // k_solve_flow chunk ownership: 0 round-robin, 1 first pass of each phase,
                             // 2 taken afresh in every pass
    unsigned* bar;           // grid barrier state and abort flag
    unsigned spin;           // polls before a wait gives up (raises the abort flag)
    Bodies B;
    int nb;
    StepTime ts;
    int iterations, relaxations;
    ContactSoA soa;          // dataflow solver: colour-ordered contact rows
    float4* vw;              // dataflow solvers: per body (v.x v.y v.z n)(w.x w.y w.z n), 32 bytes
    ConArgs con;             // barrier solver: constraints (nbatch 0: none)
    int stabilize;           // barrier solver: 1 = World.Stabilize (no IntegrateForces / IntegrateVelocities)
};


template <bool Coh = false>
__device__ __forceinline__ void integrate_forces(const SolveArgs& P, int t0, int stride) {
    const Bodies& B = P.B;
    CohBodies cb = coh_bodies(B, P.nb);
    for (int b = t0; b < P.nb; b += stride) {  // IntegrateForces (World.Step.cs:1060-1072)
        if (B.motion[b] != JH_MOTION_DYNAMIC) continue;
        bst<Coh>(B.angvel, cb.angvel, b, st3(ld3(bld<Coh>(B.angvel, cb.angvel, b)) + ld3(B.dangvel[b])));
        bst<Coh>(B.vel, cb.vel, b, st3(ld3(bld<Coh>(B.vel, cb.vel, b)) + ld3(B.dvel[b])));
    }
}
template <bool Coh = false>
__device__ __forceinline__ void integrate_velocities(const SolveArgs& P, int t0, int stride) {
    const Bodies& B = P.B;
    CohBodies cb = coh_bodies(B, P.nb);
    for (int b = t0; b < P.nb; b += stride) {  // IntegrateVelocities (World.Step.cs:1106-1133)
        if (B.motion[b] == JH_MOTION_STATIC) continue;
        vec3 p = ld3(bld<Coh>(B.pos, cb.pos, b)) + ld3(bld<Coh>(B.vel, cb.vel, b)) * P.ts.substep_dt;
        bst<Coh>(B.pos, cb.pos, b, st3(p));
        quat rq = rotation_quaternion(ld3(bld<Coh>(B.angvel, cb.angvel, b)), P.ts.substep_dt);
        bst<Coh>(B.quat, cb.quat, b, stq(qnormalize(qmul(rq, ldq(bld<Coh>(B.quat, cb.quat, b))))));
    }
}

// Barrier solver: one grid barrier per colour batch per pass.  False when the launch is aborting.
template <bool Coh = false>
__device__ __forceinline__ bool solve_batches(const SolveArgs& P, int phase, int nc) {
    int stride = gridDim.x * blockDim.x;
    int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    const Bodies& B = P.B;
    CohBodies cb = coh_bodies(B, P.nb);
    for (int c = 0; c <= nc; ++c) {
        int col = c < nc ? c : kMaxColours;
        int beg = P.offs[col], end = P.offs[col + 1];
        if (beg == end) continue;
        int first = col == kMaxColours ? 0 : t0;
        int step = col == kMaxColours ? 1 : stride;
        if (col != kMaxColours || t0 == 0)
            for (int k = beg + first; k < end; k += step) {
                jh_arbiter& a = P.arb[uint32_t(P.order[k])];
                BodyReg b1, b2;
                b1.v = ld3(bld<Coh>(B.vel, cb.vel, a.body1));
                b1.w = ld3(bld<Coh>(B.angvel, cb.angvel, a.body1));
                load_static_part(B, a.body1, b1);
                b2.v = ld3(bld<Coh>(B.vel, cb.vel, a.body2));
                b2.w = ld3(bld<Coh>(B.angvel, cb.angvel, a.body2));
                load_static_part(B, a.body2, b2);
                arbiter_pass<Coh>(a, b1, b2, B, cb, phase, P.ts);
                if (a.mode & (JH_SOLVE_LINEAR1 | JH_SOLVE_ANGULAR1)) {
                    bst<Coh>(B.vel, cb.vel, a.body1, st3(b1.v));
                    bst<Coh>(B.angvel, cb.angvel, a.body1, st3(b1.w));
                }
                if (a.mode & (JH_SOLVE_LINEAR2 | JH_SOLVE_ANGULAR2)) {
                    bst<Coh>(B.vel, cb.vel, a.body2, st3(b2.v));
                    bst<Coh>(B.angvel, cb.angvel, a.body2, st3(b2.w));
                }
            }
        if (!sync_grid<Coh>(P.bar, P.spin)) return false;
    }
    return true;
}

// PrepareConstraints / IterateConstraints (World.Step.cs:621-716) after the contacts of the same pass:
// phase 0 prepares, phase 1 iterates.  One grid barrier per batch.
template <bool Coh>
__device__ __forceinline__ void con_load(const Bodies& B, const CohBodies& cb, int b, con::CBody& r) {
    r.v = ld3(bld<Coh>(B.vel, cb.vel, b));
    r.w = ld3(bld<Coh>(B.angvel, cb.angvel, b));
    r.pos = ld3(bld<Coh>(B.pos, cb.pos, b));
    r.q = ldq(bld<Coh>(B.quat, cb.quat, b));
    r.imv = ld3(B.imv[b]);
    r.I = ld_iw(B, b);
    r.dyn = B.motion[b] == JH_MOTION_DYNAMIC;
}
template <bool Coh = false>
__device__ __forceinline__ bool solve_constraint_batches(const SolveArgs& P, int phase) {
    const ConArgs& C = P.con;
    int stride = gridDim.x * blockDim.x;
    int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    const Bodies& B = P.B;
    CohBodies cb = coh_bodies(B, P.nb);
    for (int b = 0; b < C.nbatch; ++b) {
        int beg = C.offs[b], end = C.offs[b + 1];
        for (int k = beg + t0; k < end; k += stride) {
            int i = C.order[k];
            jh_constraint c = C.c[i];
            con::Work w = C.work[i];
            con::CBody A, Bb;
            con_load<Coh>(B, cb, c.body1, A);
            con_load<Coh>(B, cb, c.body2, Bb);
            if (!A.dyn && !Bb.dyn) continue;  // suspended (World.UpdateConstraintActivation, World.cs:741-760)
            if (phase == 0) con::prepare(c, w, A, Bb, P.ts.inv_dt, P.ts.inv_substep_dt, P.ts.substep_dt);
            else con::iterate(c, w, A, Bb, P.ts.inv_substep_dt);
            for (int q = 0; q < 8; ++q) C.c[i].state[q] = c.state[q];
            if (phase == 0) C.work[i] = w;
            if (A.dyn) {
                bst<Coh>(B.vel, cb.vel, c.body1, st3(A.v));
                bst<Coh>(B.angvel, cb.angvel, c.body1, st3(A.w));
            }
            if (Bb.dyn) {
                bst<Coh>(B.vel, cb.vel, c.body2, st3(Bb.v));
                bst<Coh>(B.angvel, cb.angvel, c.body2, st3(Bb.w));
            }
        }
        if (!sync_grid<Coh>(P.bar, P.spin)) return false;
    }
    return true;
}

__global__ void __launch_bounds__(256) k_solve(SolveArgs P) {
    int stride = gridDim.x * blockDim.x;
    int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    int nc = *P.ncolours;
    for (int s = 0; s < P.ts.substeps; ++s) {
        if (!P.stabilize) {
            integrate_forces<true>(P, t0, stride);
            if (!grid_barrier_coh(P.bar, P.spin)) return;
        }
        if (!solve_batches<true>(P, 0, nc)) return;
        if (!solve_constraint_batches<true>(P, 0)) return;
        for (int it = 0; it < P.iterations; ++it) {
            if (!solve_batches<true>(P, 1, nc)) return;
            if (!solve_constraint_batches<true>(P, 1)) return;
        }
        if (!P.stabilize) {
            integrate_velocities<true>(P, t0, stride);
            if (!grid_barrier_coh(P.bar, P.spin)) return;
        }
        for (int it = 0; it < P.relaxations; ++it)
            if (!solve_batches<true>(P, 2, nc)) return;
    }
}

// ---- Dataflow solver ------------------------------------------------------------------------
// Every non-static body counts the arbiter visits applied to it in this step.  In global pass g
// (prepare, iterations and relaxations of every substep, in order), the arbiter at colour-order
// position k waits until each of its bodies has seen g * deg(b) + rank(k, b) visits: exactly the
// visits that precede it in the colour-ordered Gauss-Seidel sequence.  Each body therefore sees its
// updates in the same order as the barrier solver (bit-identical results) without grid barriers
// between colours.  The count travels inside the body's packed velocity record (see coh_poll_vw).
// Waves visit 64-position chunks of the colour order (handed out by flow_passes_queue,
// flow_passes_own or flow_passes), each wave its chunks in increasing order, so the earliest
// unfinished chunk always has its dependencies met: the loop cannot deadlock on a cooperative (fully
// resident) grid.  k_soa_gather lays the arbiters out by colour-order position.
__global__ void __launch_bounds__(256) k_soa_gather(const jh_arbiter* __restrict__ arb, const uint64_t* __restrict__ order,
                                                    const int* __restrict__ offs, ContactSoA S,
                                                    const int2* __restrict__ rank = nullptr,
                                                    const int* __restrict__ deg = nullptr,
                                                    const int* __restrict__ motion = nullptr) {
    int A = offs[128];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < A; k += gridDim.x * blockDim.x) {
        const jh_arbiter& a = arb[uint32_t(order[k])];
        int n = __popc(a.usage_mask & 15u);
        S.bodies[k] = make_int2(a.body1, a.body2);
        if (S.nd) {
            int2 r = rank[k];
            S.nd[k] = make_int4(motion[a.body1] != JH_MOTION_STATIC ? deg[a.body1] : -1,
                                motion[a.body2] != JH_MOTION_STATIC ? deg[a.body2] : -1, r.x, r.y);
        }
        S.um[k] = make_uint2((1u << n) - 1u, a.mode);  // used slots compacted to rows 0 .. n-1, in slot order
        S.fs[k] = make_float2(a.friction, a.speculative_relaxation_factor);
        for (int j = 0, s = 0; s < 4; ++s) {
            if (!(a.usage_mask & (1u << s))) continue;
            const jh_contact& c = a.contacts[s];
            int r = j++;
            if (S.mv) {
                float4* m = S.mv + size_t(r) * 3 * S.cap + k;
                float v0 = __int_as_float(-1);
                m[0] = make_float4(c.accumulated[0], c.accumulated[1], c.accumulated[2], v0);
                m[S.cap] = make_float4(c.bias, c.penalty_bias, c.mass_normal_tangent[0], v0);
                m[2 * size_t(S.cap)] = make_float4(c.mass_normal_tangent[1], c.mass_normal_tangent[2], __uint_as_float(c.flag), v0);
            }
            S.flag[size_t(r) * S.cap + k] = c.flag;
            S.at(r, kFBias, k) = c.bias;
            S.at(r, kFPen, k) = c.penalty_bias;
            for (int l = 0; l < 3; ++l) {
                S.at(r, kFN + l, k) = c.normal[l];
                S.at(r, kFT1 + l, k) = c.tangent1[l];
                S.at(r, kFT2 + l, k) = c.tangent2[l];
                S.at(r, kFMass + l, k) = c.mass_normal_tangent[l];
                S.at(r, kFAcc + l, k) = c.accumulated[l];
                S.at(r, kFRel1 + l, k) = c.relative_position1[l];
                S.at(r, kFRel2 + l, k) = c.relative_position2[l];
                S.at(r, kFPos1 + l, k) = c.position1[l];
                S.at(r, kFPos2 + l, k) = c.position2[l];
            }
        }
    }
}

__global__ void __launch_bounds__(256) k_soa_scatter(jh_arbiter* __restrict__ arb, const uint64_t* __restrict__ order,
                                                     const int* __restrict__ offs, ContactSoA S) {
    int A = offs[128];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < A; k += gridDim.x * blockDim.x) {
        jh_arbiter& a = arb[uint32_t(order[k])];
        uint32_t usage = a.usage_mask & 15u;  // row r holds the r-th used slot
        for (int j = 0, s = 0; s < 4; ++s) {
            if (!(usage & (1u << s))) continue;
            int r = j++;
            jh_contact& c = a.contacts[s];
            if (S.mv) {
                const float4* m = S.mv + size_t(r) * 3 * S.cap + k;
                float4 m0 = m[0], m1 = m[S.cap], m2 = m[2 * size_t(S.cap)];
                c.accumulated[0] = m0.x; c.accumulated[1] = m0.y; c.accumulated[2] = m0.z;
                c.bias = m1.x; c.penalty_bias = m1.y; c.mass_normal_tangent[0] = m1.z;
                c.mass_normal_tangent[1] = m2.x; c.mass_normal_tangent[2] = m2.y; c.flag = __float_as_uint(m2.z);
                continue;
            }
            c.flag = S.flag[size_t(r) * S.cap + k];
            c.bias = S.at(r, kFBias, k);
            c.penalty_bias = S.at(r, kFPen, k);
            for (int l = 0; l < 3; ++l) {
                c.mass_normal_tangent[l] = S.at(r, kFMass + l, k);
                c.accumulated[l] = S.at(r, kFAcc + l, k);
            }
        }
    }
}

// Packed velocity record of the dataflow solvers, 32 bytes per body:
//   (v.x v.y v.z n)(w.x w.y w.z n),  n = visits applied to the body so far in this launch,
// the count repeated in both 16-byte halves.  A visit is published by storing both halves with the
// new count; a waiting visit takes the record as soon as both halves carry the count it waits for,
// so one pair of loads is both the readiness poll and the data, and publishing needs no separate
// counter update nor a wait for the data stores.  Each half is one aligned 16-byte buffer access
// (single-copy atomic within a cache line), so a half with the expected count holds that visit's
// values.  The accesses carry the sc1 bit (device coherence: served past the L2 of the issuing
// XCD), like the compiler's agent-scope atomics.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t vw_rsrc(const SolveArgs& P) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)P.vw, 0, P.nb * 32, kBufferDword3);
}
// True (and v, w set) when body b's record shows exactly `need` visits; `seen` gets the smaller of the
// two counts (how many visits the body still lacks is need - seen).
__device__ __forceinline__ bool coh_poll_vw(__amdgpu_buffer_rsrc_t r, int b, int need, vec3& v, vec3& w, int& seen) {
    float4 a = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, b * 32, 0, kCohSc1));
    float4 c = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, b * 32 + 16, 0, kCohSc1));
    v = {a.x, a.y, a.z};
    w = {c.x, c.y, c.z};
    seen = min(__float_as_int(a.w), __float_as_int(c.w));
    return __float_as_int(a.w) == need && __float_as_int(c.w) == need;
}
__device__ __forceinline__ bool coh_poll_vw(__amdgpu_buffer_rsrc_t r, int b, int need, vec3& v, vec3& w) {
    int seen;
    return coh_poll_vw(r, b, need, v, w, seen);
}
__device__ __forceinline__ void coh_publish_vw(__amdgpu_buffer_rsrc_t r, int b, vec3 v, vec3 w, int n) {
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    float fn = __int_as_float(n);
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, make_float4(v.x, v.y, v.z, fn)), r, b * 32, 0,
                                           kCohSc1);
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, make_float4(w.x, w.y, w.z, fn)), r, b * 32 + 16,
                                           0, kCohSc1);
}
// Device-coherent read of a whole record (the integration loops of k_solve_flow, which sits between
// relaxed grid barriers): returns the count, v and w.
__device__ __forceinline__ int coh_ld_vw(__amdgpu_buffer_rsrc_t r, int b, vec3& v, vec3& w) {
    float4 a = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, b * 32, 0, kCohSc1));
    float4 c = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, b * 32 + 16, 0, kCohSc1));
    v = {a.x, a.y, a.z};
    w = {c.x, c.y, c.z};
    return __float_as_int(a.w);
}
// Plain accesses, for phases separated from the dataflow passes by grid barriers.
__device__ __forceinline__ int ld_vw(const SolveArgs& P, int b, vec3& v, vec3& w) {
    float4 a = P.vw[2 * b], c = P.vw[2 * b + 1];
    v = {a.x, a.y, a.z};
    w = {c.x, c.y, c.z};
    return __float_as_int(a.w);
}
__device__ __forceinline__ void st_vw(const SolveArgs& P, int b, vec3 v, vec3 w, int n) {
    P.vw[2 * b] = make_float4(v.x, v.y, v.z, __int_as_float(n));
    P.vw[2 * b + 1] = make_float4(w.x, w.y, w.z, __int_as_float(n));
}

// Contact fields the solver reads/writes, held in registers for one arbiter visit.
struct SlotReg {
    uint32_t flag;
    float bias, penalty;
    float nx[3], ny[3], nz[3];  // lanes: normal, tangent1, tangent2
    float mass[3], acc[3];
    vec3 rel1, rel2, pos1, pos2;
};

__device__ __forceinline__ void load_slot(const jh_contact& c, SlotReg& r, bool prepare) {
    r.flag = c.flag;
    r.bias = c.bias;
    r.penalty = c.penalty_bias;
    r.nx[0] = c.normal[0]; r.ny[0] = c.normal[1]; r.nz[0] = c.normal[2];
    r.nx[1] = c.tangent1[0]; r.ny[1] = c.tangent1[1]; r.nz[1] = c.tangent1[2];
    r.nx[2] = c.tangent2[0]; r.ny[2] = c.tangent2[1]; r.nz[2] = c.tangent2[2];
    for (int l = 0; l < 3; ++l) { r.mass[l] = c.mass_normal_tangent[l]; r.acc[l] = c.accumulated[l]; }
    r.rel1 = vec3{c.relative_position1[0], c.relative_position1[1], c.relative_position1[2]};
    r.rel2 = vec3{c.relative_position2[0], c.relative_position2[1], c.relative_position2[2]};
    if (prepare) {
        r.pos1 = vec3{c.position1[0], c.position1[1], c.position1[2]};
        r.pos2 = vec3{c.position2[0], c.position2[1], c.position2[2]};
    }
}
__device__ __forceinline__ void load_slot_soa(const ContactSoA& S, int s, int k, SlotReg& r, bool prepare) {
    r.flag = S.flag[size_t(s) * S.cap + k];
    r.bias = S.at(s, kFBias, k);
    r.penalty = S.at(s, kFPen, k);
    for (int l = 0; l < 3; ++l) {
        r.nx[l] = S.at(s, (l == 0 ? kFN : l == 1 ? kFT1 : kFT2) + 0, k);
        r.ny[l] = S.at(s, (l == 0 ? kFN : l == 1 ? kFT1 : kFT2) + 1, k);
        r.nz[l] = S.at(s, (l == 0 ? kFN : l == 1 ? kFT1 : kFT2) + 2, k);
        r.mass[l] = S.at(s, kFMass + l, k);
        r.acc[l] = S.at(s, kFAcc + l, k);
    }
    r.rel1 = vec3{S.at(s, kFRel1, k), S.at(s, kFRel1 + 1, k), S.at(s, kFRel1 + 2, k)};
    r.rel2 = vec3{S.at(s, kFRel2, k), S.at(s, kFRel2 + 1, k), S.at(s, kFRel2 + 2, k)};
    if (prepare) {
        r.pos1 = vec3{S.at(s, kFPos1, k), S.at(s, kFPos1 + 1, k), S.at(s, kFPos1 + 2, k)};
        r.pos2 = vec3{S.at(s, kFPos2, k), S.at(s, kFPos2 + 1, k), S.at(s, kFPos2 + 2, k)};
    }
}
__device__ __forceinline__ void store_slot_soa(const ContactSoA& S, int s, int k, const SlotReg& r, bool prepare) {
    if (prepare) {
        S.flag[size_t(s) * S.cap + k] = r.flag;
        S.at(s, kFBias, k) = r.bias;
        S.at(s, kFPen, k) = r.penalty;
        for (int l = 0; l < 3; ++l) S.at(s, kFMass + l, k) = r.mass[l];
    }
    for (int l = 0; l < 3; ++l) S.at(s, kFAcc + l, k) = r.acc[l];
}

__device__ __forceinline__ void store_slot(jh_contact& c, const SlotReg& r, bool prepare) {
    if (prepare) {
        c.flag = r.flag;
        c.bias = r.bias;
        c.penalty_bias = r.penalty;
        for (int l = 0; l < 3; ++l) c.mass_normal_tangent[l] = r.mass[l];
    }
    for (int l = 0; l < 3; ++l) c.accumulated[l] = r.acc[l];
}

// Same arithmetic as prepare_contact, on register copies, in two parts.  prepare_fixed computes what
// depends only on the poses, the inertia and the contact frame (penetration, effective masses,
// penalty bias), which no other visit of the solve changes; prepare_warm applies the part that
// depends on the mutable fields (flags, bias, accumulated impulses) and on the velocities.  The
// dataflow solver runs the first before it waits for its inputs.
struct SlotPrep {
    float pen, penalty, mass[3];
};
__device__ __forceinline__ void prepare_fixed(uint32_t mode, const SlotReg& c, const BodyReg& B1, const BodyReg& B2,
                                              vec3 x1, quat q1, vec3 x2, quat q2, const StepTime& ts, SlotPrep& sp) {
    vec3 cr1 = rot(c.pos1, q1);
    vec3 cr2 = rot(c.pos2, q2);
    vec3 imv{0.0f, 0.0f, 0.0f};
    if (mode & JH_SOLVE_LINEAR1) imv = B1.imv;
    if (mode & JH_SOLVE_LINEAR2) imv = imv + B2.imv;
    vec3 r1 = cr1 + x1, r2 = cr2 + x2;
    vec3 dist = r1 - r2;
    const float* NX = c.nx;
    const float* NY = c.ny;
    const float* NZ = c.nz;
    float k[3] = {imv.x, imv.x, imv.x};
    if (B1.ll || B2.ll)
        for (int l = 0; l < 3; ++l) k[l] = (NX[l] * imv.x) * NX[l] + (NY[l] * imv.y) * NY[l] + (NZ[l] * imv.z) * NZ[l];
    vec3 n{NX[0], NY[0], NZ[0]};
    float pen = dot(dist, n);
    sp.pen = pen;
    float ktx[3] = {0, 0, 0}, kty[3] = {0, 0, 0}, ktz[3] = {0, 0, 0};
    if (mode & JH_SOLVE_ANGULAR1) {
        vec3 rp = c.rel1;
        const sym3& I = B1.I;
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            float rrx = rp.y * NZ[l] - rp.z * NY[l];
            float rry = rp.z * NX[l] - rp.x * NZ[l];
            float rrz = rp.x * NY[l] - rp.y * NX[l];
            float e1 = I.m11 * rrx + I.m12 * rry + I.m13 * rrz;
            float e2 = I.m12 * rrx + I.m22 * rry + I.m23 * rrz;
            float e3 = I.m13 * rrx + I.m23 * rry + I.m33 * rrz;
            ktx[l] = e2 * rp.z - e3 * rp.y;
            kty[l] = e3 * rp.x - e1 * rp.z;
            ktz[l] = e1 * rp.y - e2 * rp.x;
        }
    }
    if (mode & JH_SOLVE_ANGULAR2) {
        vec3 rp = c.rel2;
        const sym3& I = B2.I;
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            float rrx = rp.y * NZ[l] - rp.z * NY[l];
            float rry = rp.z * NX[l] - rp.x * NZ[l];
            float rrz = rp.x * NY[l] - rp.y * NX[l];
            float f1 = I.m11 * rrx + I.m12 * rry + I.m13 * rrz;
            float f2 = I.m12 * rrx + I.m22 * rry + I.m23 * rrz;
            float f3 = I.m13 * rrx + I.m23 * rry + I.m33 * rrz;
            ktx[l] = ktx[l] + (f2 * rp.z - f3 * rp.y);
            kty[l] = kty[l] + (f3 * rp.x - f1 * rp.z);
            ktz[l] = ktz[l] + (f1 * rp.y - f2 * rp.x);
        }
    }
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        float kres = NX[l] * ktx[l] + NY[l] * kty[l] + NZ[l] * ktz[l];
        float kk = k[l] + kres;
        sp.mass[l] = kk > 0.0f ? 1.0f / kk : 0.0f;
    }
    float pb = 0.2f * ts.inv_dt * fmaxr(0.0f, pen - 0.01f);
    sp.penalty = fminr(pb, 100.0f);
}
// I (rp x n_l) for the three lanes; the inputs pass through an empty asm so that a visit recomputes
// these after its wait instead of keeping the copies prepare_fixed made alive across it.
__device__ __forceinline__ void lever_inertia(vec3 rp, sym3 I, const float* NX, const float* NY, const float* NZ,
                                              float e1[3], float e2[3], float e3[3]) {
    asm volatile("" : "+v"(rp.x), "+v"(rp.y), "+v"(rp.z));
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        float rrx = rp.y * NZ[l] - rp.z * NY[l];
        float rry = rp.z * NX[l] - rp.x * NZ[l];
        float rrz = rp.x * NY[l] - rp.y * NX[l];
        e1[l] = I.m11 * rrx + I.m12 * rry + I.m13 * rrz;
        e2[l] = I.m12 * rrx + I.m22 * rry + I.m23 * rrz;
        e3[l] = I.m13 * rrx + I.m23 * rry + I.m33 * rrz;
    }
}
__device__ __forceinline__ void prepare_warm(uint32_t mode, float specrel, SlotReg& c, BodyReg& B1, BodyReg& B2,
                                             const SlotPrep& sp, const StepTime& ts) {
    bool skip = (c.flag & 4u) != 0;
    c.flag &= ~4u;
    float ws[3] = {skip ? 0.0f : c.acc[0], skip ? 0.0f : c.acc[1], skip ? 0.0f : c.acc[2]};
    if ((c.flag & 2u) == 0) c.bias = 0.0f;
    if (sp.pen < -0.02f) c.bias = sp.pen * ts.inv_dt * specrel;
    c.flag &= ~2u;
    c.penalty = sp.penalty;
    for (int l = 0; l < 3; ++l) c.mass[l] = sp.mass[l];
    const float* NX = c.nx;
    const float* NY = c.ny;
    const float* NZ = c.nz;
    vec3 lin{ws[0] * NX[0] + ws[1] * NX[1] + ws[2] * NX[2], ws[0] * NY[0] + ws[1] * NY[1] + ws[2] * NY[2],
             ws[0] * NZ[0] + ws[1] * NZ[1] + ws[2] * NZ[2]};
    if (mode & JH_SOLVE_LINEAR1) B1.v = B1.v - cmul(lin, B1.imv);
    if (mode & JH_SOLVE_ANGULAR1) {
        float e1[3], e2[3], e3[3];
        lever_inertia(c.rel1, B1.I, NX, NY, NZ, e1, e2, e3);
        vec3 ai{ws[0] * e1[0] + ws[1] * e1[1] + ws[2] * e1[2], ws[0] * e2[0] + ws[1] * e2[1] + ws[2] * e2[2],
                ws[0] * e3[0] + ws[1] * e3[1] + ws[2] * e3[2]};
        B1.w = B1.w - ai;
    }
    if (mode & JH_SOLVE_LINEAR2) B2.v = B2.v + cmul(lin, B2.imv);
    if (mode & JH_SOLVE_ANGULAR2) {
        float f1[3], f2[3], f3[3];
        lever_inertia(c.rel2, B2.I, NX, NY, NZ, f1, f2, f3);
        vec3 ai{ws[0] * f1[0] + ws[1] * f1[1] + ws[2] * f1[2], ws[0] * f2[0] + ws[1] * f2[1] + ws[2] * f2[2],
                ws[0] * f3[0] + ws[1] * f3[1] + ws[2] * f3[2]};
        B2.w = B2.w + ai;
    }
}
__device__ __forceinline__ void prepare_reg(uint32_t mode, float specrel, SlotReg& c, BodyReg& B1, BodyReg& B2,
                                            vec3 x1, quat q1, vec3 x2, quat q2, const StepTime& ts) {
    SlotPrep sp;
    prepare_fixed(mode, c, B1, B2, x1, q1, x2, q2, ts, sp);
    prepare_warm(mode, specrel, c, B1, B2, sp, ts);
}

// Same arithmetic as iterate_contact, on register copies.
__device__ __forceinline__ void iterate_reg(uint32_t mode, float friction, SlotReg& c, BodyReg& B1, BodyReg& B2,
                                            bool apply_bias) {
    vec3 dv = B2.v - B1.v;
    if (mode & (JH_SOLVE_ANGULAR1 | JH_SOLVE_ANGULAR2)) {
        dv = B2.v + cross(B2.w, c.rel2);
        dv = dv - (B1.v + cross(B1.w, c.rel1));
    }
    float bias = c.bias;
    if (apply_bias && c.penalty > 0) bias = fmaxr(c.penalty, c.bias);
    float old[3] = {c.acc[0], c.acc[1], c.acc[2]};
    float bv[3] = {bias, 0.0f, 0.0f};
    float imp[3];
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        float vd = c.nx[l] * dv.x + c.ny[l] * dv.y + c.nz[l] * dv.z;
        imp[l] = c.mass[l] * (bv[l] - vd);
    }
    float maxT = friction * old[0];
    float lo[3] = {0.0f, -maxT, -maxT}, hi[3] = {3.402823466e+38f, maxT, maxT};
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        float a = old[l] + imp[l];
        a = fminr(fmaxr(a, lo[l]), hi[l]);
        c.acc[l] = a;
        imp[l] = a - old[l];
    }
    vec3 lin{imp[0] * c.nx[0] + imp[1] * c.nx[1] + imp[2] * c.nx[2], imp[0] * c.ny[0] + imp[1] * c.ny[1] + imp[2] * c.ny[2],
             imp[0] * c.nz[0] + imp[1] * c.nz[1] + imp[2] * c.nz[2]};
    if (mode & JH_SOLVE_LINEAR1) B1.v = B1.v - cmul(lin, B1.imv);
    if (mode & JH_SOLVE_ANGULAR1) B1.w = B1.w - sym_mul(cross(c.rel1, lin), B1.I);
    if (mode & JH_SOLVE_LINEAR2) B2.v = B2.v + cmul(lin, B2.imv);
    if (mode & JH_SOLVE_ANGULAR2) B2.w = B2.w + sym_mul(cross(c.rel2, lin), B2.I);
}

// Versioned mutable contact fields of the dataflow solver.  Row r of position k has three 16-byte
// records, each carrying the pass that wrote it in its last word:
//   m0 (acc.n, acc.t1, acc.t2, pass of the arbiter's last visit)
//   m1 (bias, penalty, mass.n, pass of its last prepare)   m2 (mass.t1, mass.t2, flag, same)
// A visit in pass g needs m0 from pass g - 1 and m1/m2 from the last prepare, so whichever wave
// visits the arbiter reads them with device-coherent loads, polled together with its two bodies'
// records until the versions match (usually at once: the previous visit of the arbiter lies a whole
// pass back).  This keeps every chunk-to-wave assignment exact, including a fresh one in every pass.
// The gather writes version -1.
// One buffer resource per row keeps the byte offsets (< 3 x 16 x cap) in 32 bits.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t mv_rsrc(const ContactSoA& S, int s) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)(S.mv + size_t(s) * 3 * S.cap), 0, 3 * 16 * S.cap, kBufferDword3);
}
__device__ __forceinline__ bool load_mutable(const ContactSoA& S, int s, int k, SlotReg& c, int want_visit,
                                             int want_prepare, bool prepare) {
    __amdgpu_buffer_rsrc_t r = mv_rsrc(S, s);
    int o = k * 16;
    float4 m0 = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, o, 0, kCohSc1));
    float4 m1 = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, o + S.cap * 16, 0, kCohSc1));
    float4 m2 = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, o + 2 * S.cap * 16, 0, kCohSc1));
    c.acc[0] = m0.x; c.acc[1] = m0.y; c.acc[2] = m0.z;
    c.bias = m1.x; c.flag = __float_as_uint(m2.z);
    if (!prepare) { c.penalty = m1.y; c.mass[0] = m1.z; c.mass[1] = m2.x; c.mass[2] = m2.y; }  // a prepare recomputes them
    return __float_as_int(m0.w) == want_visit && __float_as_int(m1.w) == want_prepare &&
           __float_as_int(m2.w) == want_prepare;
}
// The fields travel by value: through a reference the compiler widens the three accumulated-impulse
// loads into one 16-byte load that reaches into the next field, which keeps the caller's register copy
// of the contact in scratch memory.
__device__ __forceinline__ void store_mutable(const ContactSoA& S, int s, int k, int g, bool prepare, float a0, float a1,
                                              float a2, float bias, float penalty, float m0, float m1, float m2,
                                              uint32_t flag) {
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    __amdgpu_buffer_rsrc_t r = mv_rsrc(S, s);
    int o = k * 16;
    float vg = __int_as_float(g);
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, make_float4(a0, a1, a2, vg)), r, o, 0, kCohSc1);
    if (prepare) {
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, make_float4(bias, penalty, m0, vg)), r,
                                               o + S.cap * 16, 0, kCohSc1);
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, make_float4(m1, m2, __uint_as_float(flag), vg)),
                                               r, o + 2 * S.cap * 16, 0, kCohSc1);
    }
}
// Only the fields no visit changes (contact frame, anchors).
__device__ __forceinline__ void load_slot_fixed(const ContactSoA& S, int s, int k, SlotReg& r, bool prepare) {
    for (int l = 0; l < 3; ++l) {
        r.nx[l] = S.at(s, (l == 0 ? kFN : l == 1 ? kFT1 : kFT2) + 0, k);
        r.ny[l] = S.at(s, (l == 0 ? kFN : l == 1 ? kFT1 : kFT2) + 1, k);
        r.nz[l] = S.at(s, (l == 0 ? kFN : l == 1 ? kFT1 : kFT2) + 2, k);
    }
    r.rel1 = vec3{S.at(s, kFRel1, k), S.at(s, kFRel1 + 1, k), S.at(s, kFRel1 + 2, k)};
    r.rel2 = vec3{S.at(s, kFRel2, k), S.at(s, kFRel2 + 1, k), S.at(s, kFRel2 + 2, k)};
    if (prepare) {
        r.pos1 = vec3{S.at(s, kFPos1, k), S.at(s, kFPos1 + 1, k), S.at(s, kFPos1 + 2, k)};
        r.pos2 = vec3{S.at(s, kFPos2, k), S.at(s, kFPos2 + 1, k), S.at(s, kFPos2 + 2, k)};
    }
}

// One dataflow visit of the arbiter at colour-order position k in global pass g.  Everything that
// does not depend on other visits (contact frames and anchors, body inertia, poses) is fetched before
// the dependency wait; the bodies' velocity records and the arbiter's versioned mutable fields are
// polled until they show the awaited visit counts / passes, then the visit computes in registers and
// publishes both.
// False when the launch is aborting (a wait gave up here or elsewhere).
// What a visit needs first: the bodies, slots, mode, friction and update counts of the arbiter at
// colour-order position k (inactive lanes of a wave's last chunk read position 0 and do nothing).
struct FlowHead {
    int2 bb;
    uint2 um;
    float2 fs;
    int4 nd;
};
__device__ __forceinline__ int flow_pos(int k, bool active) { return active ? k : 0; }
__device__ __forceinline__ FlowHead load_head(const ContactSoA& S, int k) {
    return FlowHead{S.bodies[k], S.um[k], S.fs[k], S.nd[k]};
}
template <bool PREPARE>
__device__ __forceinline__ bool flow_visit(const SolveArgs& P, const FlowHead& h, int k, int g, bool apply_bias,
                                           bool active) {
    const Bodies& B = P.B;
    k = flow_pos(k, active);
    const ContactSoA& S = P.soa;
    int b1 = h.bb.x, b2 = h.bb.y;
    uint32_t usage = h.um.x, mode = h.um.y;
    float friction = h.fs.x, specrel = h.fs.y;
    bool m1 = h.nd.x >= 0, m2 = h.nd.y >= 0;
    int need1 = m1 ? g * h.nd.x + h.nd.z : 0;
    int need2 = m2 ? g * h.nd.y + h.nd.w : 0;
    // passes the mutable fields must come from: the arbiter's previous visit (-1: the gather) and the
    // last prepare (for a prepare: the previous substep's, or the gather)
    int per_sub = 1 + P.iterations + P.relaxations;
    int want_visit = g - 1;
    int want_prep = PREPARE ? (g >= per_sub ? g - per_sub : -1) : (g / per_sub) * per_sub;
    SlotReg sr[4];
#pragma unroll
    for (int s = 0; s < 4; ++s)
        if (usage & (1u << s)) load_slot_fixed(S, s, k, sr[s], PREPARE);
    BodyReg R1, R2;
    load_static_part(B, b1, R1);
    load_static_part(B, b2, R2);
    vec3 x1, x2;
    quat q1, q2;
    if (PREPARE) {  // poses written by the previous substep's IntegrateVelocities, in this launch
        CohBodies cb = coh_bodies(B, P.nb);
        x1 = ld3(coh_ld(cb.pos, b1));
        x2 = ld3(coh_ld(cb.pos, b2));
        q1 = ldq(coh_ld(cb.quat, b1));
        q2 = ldq(coh_ld(cb.quat, b2));
    }
    SlotPrep sp[4];
    if (PREPARE) {
#pragma unroll
        for (int s = 0; s < 4; ++s)
            if (usage & (1u << s)) prepare_fixed(mode, sr[s], R1, R2, x1, q1, x2, q2, P.ts, sp[s]);
    }
    __amdgpu_buffer_rsrc_t vr = vw_rsrc(P);
    if (!m1) coh_ld_vw(vr, b1, R1.v, R1.w);
    if (!m2) coh_ld_vw(vr, b2, R2.v, R2.w);
    // Each lane computes and publishes as soon as its own inputs are ready; the wave leaves when all
    // of its lanes have (a lane never waits for an unrelated lane's dependencies).  An input that
    // showed the awaited version is not polled again (no other visit can change it before this one).
    bool done = !active;
    bool ok1 = !m1, ok2 = !m2, rows = false;
    unsigned polls = 0;
    while (true) {
        asm volatile("" ::: "memory");
        if (!done) {
            if (!rows) {
                rows = true;
#pragma unroll
                for (int s = 0; s < 4; ++s)
                    if (usage & (1u << s)) rows = load_mutable(S, s, k, sr[s], want_visit, want_prep, PREPARE) && rows;
            }
            if (!ok1) ok1 = coh_poll_vw(vr, b1, need1, R1.v, R1.w);
            if (!ok2) ok2 = coh_poll_vw(vr, b2, need2, R2.v, R2.w);
            if (ok1 && ok2 && rows) {
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    if (!(usage & (1u << s))) continue;
                    if (PREPARE) prepare_warm(mode, specrel, sr[s], R1, R2, sp[s], P.ts);
                    else iterate_reg(mode, friction, sr[s], R1, R2, apply_bias);
                }
                if (m1) coh_publish_vw(vr, b1, R1.v, R1.w, need1 + 1);
                if (m2) coh_publish_vw(vr, b2, R2.v, R2.w, need2 + 1);
#pragma unroll
                for (int s = 0; s < 4; ++s)
                    if (usage & (1u << s))
                        store_mutable(S, s, k, g, PREPARE, sr[s].acc[0], sr[s].acc[1], sr[s].acc[2], sr[s].bias,
                                      sr[s].penalty, sr[s].mass[0], sr[s].mass[1], sr[s].mass[2], sr[s].flag);
                done = true;
            }
        }
        if (__all(done)) return true;
        if (!keep_polling(P.bar, polls, P.spin)) return false;
        __builtin_amdgcn_s_sleep(2);
    }
}

// Chunks of a phase handed out in global order from a counter, in every pass: a wave that frees up
// takes the oldest chunk nobody has taken, so a ready chunk never waits behind a slow chunk of the
// same wave.  Chunks are the 64-position blocks of the colour order, so a chunk's bounds follow from
// the counter without a table load; a block that straddles two colours may hold an arbiter that
// waits for another lane of the same wave, which the lane-wise wait loop of flow_visit serves.
// Deadlock-free: the oldest unfinished chunk has been taken, and its predecessors (lower positions,
// possibly in the same chunk) are done or progress.  Exact because an arbiter's mutable fields are
// versioned (load_mutable), whichever wave visits it.
__device__ __forceinline__ bool flow_passes_queue(const SolveArgs& P, int g_first, int npass, int phase_first,
                                                  int phase_rest, int* queue) {
    int lane = threadIdx.x & 63;
    int A = P.offs[128];
    int J = (A + 63) >> 6;
    int total = J * npass;
    while (true) {
        int t = 0;
        if (lane == 0) t = atomicAdd(queue, 1);
        t = __shfl(t, 0, 64);
        if (t >= total) return true;
        int p = t / J, j = t - p * J;
        int g = g_first + p;
        int phase = p == 0 ? phase_first : phase_rest;
        int k = (j << 6) + lane;
        bool active = k < A;
        FlowHead h = load_head(P.soa, flow_pos(k, active));
        bool ok = phase == 0 ? flow_visit<true>(P, h, k, g, false, active)
                             : flow_visit<false>(P, h, k, g, phase == 1, active);
        if (!ok) return false;
    }
}

// Chunk ownership decided in the first pass of a phase: waves take chunks from a counter in global
// order, record them, and replay the same list in the phase's later passes.  Deadlock-free: the first pass takes chunks in global order, the replays walk each list
// in increasing order, so the earliest unfinished visit always has its predecessors done.  A wave
// takes at most kOwnMax chunks per phase; with more chunks than all waves can own, the phase uses the
// round-robin ownership of flow_passes instead.
constexpr int kOwnMax = 64;
__device__ __forceinline__ bool flow_passes(const SolveArgs& P, int g_first, int npass, int phase_first,
                                            int phase_rest);
__device__ __forceinline__ bool flow_passes_own(const SolveArgs& P, int g_first, int npass, int phase_first,
                                                int phase_rest, int* queue, int* own) {
    int lane = threadIdx.x & 63;
    int A = P.offs[128];
    int J = (A + 63) >> 6;  // 64-position chunks of the colour order
    if (J > ((gridDim.x * blockDim.x) >> 6) * kOwnMax) return flow_passes(P, g_first, npass, phase_first, phase_rest);
    int n = 0;
    for (int p = 0; p < npass; ++p) {
        int g = g_first + p;
        int phase = p == 0 ? phase_first : phase_rest;
        for (int i = 0;; ++i) {
            int j;
            if (p == 0) {
                if (i == kOwnMax) break;
                int t = 0;
                if (lane == 0) t = atomicAdd(queue, 1);
                j = __shfl(t, 0, 64);
                if (j >= J) break;
                if (lane == 0) own[i] = j;
                n = i + 1;
            } else {
                if (i == n) break;
                j = own[i];
            }
            int k = (j << 6) + lane;
            bool active = k < A;
            FlowHead h = load_head(P.soa, flow_pos(k, active));
            bool ok = phase == 0 ? flow_visit<true>(P, h, k, g, false, active)
                                 : flow_visit<false>(P, h, k, g, phase == 1, active);
            if (!ok) return false;
        }
    }
    return true;
}

__device__ __forceinline__ bool flow_passes(const SolveArgs& P, int g_first, int npass, int phase_first,
                                            int phase_rest) {
    int lane = threadIdx.x & 63;
    int W = (gridDim.x * blockDim.x) >> 6;
    int gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int A = P.offs[128];
    int J = (A + 63) >> 6;  // 64-position chunks of the colour order
    for (int p = 0; p < npass; ++p) {
        int g = g_first + p;
        int phase = p == 0 ? phase_first : phase_rest;
        for (int j = gw; j < J; j += W) {
            int k = (j << 6) + lane;
            bool active = k < A;
            FlowHead h = load_head(P.soa, flow_pos(k, active));
            bool ok = phase == 0 ? flow_visit<true>(P, h, k, g, false, active)
                                 : flow_visit<false>(P, h, k, g, phase == 1, active);
            if (!ok) return false;
        }
    }
    return true;
}

// One launch per step (World.Step.cs:234-253): per substep IntegrateForces | prepare + iterations
// (dataflow) | IntegrateVelocities | relaxations (dataflow), grid barriers only between these phases.
// Everything workgroups share within the launch — the velocity records, the poses IntegrateVelocities
// writes for the next substep's prepare, the chunk counters, the versioned contact fields — goes
// through device-coherent (sc1) accesses, so the phase barriers are relaxed ones (each wave's accesses
// complete, one relaxed arrival per workgroup; no L2 write-back or invalidate).
__global__ void __launch_bounds__(256) k_solve_flow(SolveArgs P) {
    __shared__ int own_all[4][kOwnMax];  // chunks owned by each wave in the current phase
    int* own = own_all[threadIdx.x >> 6];
    int stride = gridDim.x * blockDim.x;
    int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    int per_sub = 1 + P.iterations + P.relaxations;
    const Bodies& B = P.B;
    __amdgpu_buffer_rsrc_t vr = vw_rsrc(P);
    CohBodies cb = coh_bodies(B, P.nb);
    for (int b = t0; b < P.nb; b += stride) coh_publish_vw(vr, b, ld3(B.vel[b]), ld3(B.angvel[b]), 0);
    if (P.qmode)  // visible after the first grid barrier
        for (int i = t0; i < 2 * P.ts.substeps; i += stride)
            __hip_atomic_store(&P.queue[i], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int s = 0; s < P.ts.substeps; ++s) {
        for (int b = t0; b < P.nb; b += stride) {  // IntegrateForces (World.Step.cs:1060-1072)
            if (B.motion[b] != JH_MOTION_DYNAMIC) continue;
            vec3 v, w;
            int n = coh_ld_vw(vr, b, v, w);
            w = w + ld3(B.dangvel[b]);
            v = v + ld3(B.dvel[b]);
            coh_publish_vw(vr, b, v, w, n);
        }
        if (!grid_barrier_coh(P.bar, P.spin)) return;
        bool ok;
        if (P.qmode == 2) ok = flow_passes_queue(P, s * per_sub, 1 + P.iterations, 0, 1, P.queue + 2 * s);
        else if (P.qmode == 1) ok = flow_passes_own(P, s * per_sub, 1 + P.iterations, 0, 1, P.queue + 2 * s, own);
        else ok = flow_passes(P, s * per_sub, 1 + P.iterations, 0, 1);
        // a wave that gave up still takes part in the barrier, which then fails for the workgroup
        if (!grid_barrier_coh(P.bar, P.spin, ok)) return;
        for (int b = t0; b < P.nb; b += stride) {  // IntegrateVelocities (World.Step.cs:1106-1133)
            if (B.motion[b] == JH_MOTION_STATIC) continue;
            vec3 v, w;
            coh_ld_vw(vr, b, v, w);
            vec3 p = ld3(coh_ld(cb.pos, b)) + v * P.ts.substep_dt;
            coh_st(cb.pos, b, st3(p));
            quat rq = rotation_quaternion(w, P.ts.substep_dt);
            coh_st(cb.quat, b, stq(qnormalize(qmul(rq, ldq(coh_ld(cb.quat, b))))));
        }
        if (!grid_barrier_coh(P.bar, P.spin)) return;
        if (P.relaxations > 0) {
            int g0 = s * per_sub + 1 + P.iterations;
            if (P.qmode == 2) ok = flow_passes_queue(P, g0, P.relaxations, 2, 2, P.queue + 2 * s + 1);
            else if (P.qmode == 1) ok = flow_passes_own(P, g0, P.relaxations, 2, 2, P.queue + 2 * s + 1, own);
            else ok = flow_passes(P, g0, P.relaxations, 2, 2);
            if (!grid_barrier_coh(P.bar, P.spin, ok)) return;
        }
    }
    for (int b = t0; b < P.nb; b += stride) {
        vec3 v, w;
        coh_ld_vw(vr, b, v, w);
        B.vel[b] = st3(v);
        B.angvel[b] = st3(w);
    }
}

// ---- Lane solver ----------------------------------------------------------------------------
// The dataflow protocol of k_solve_flow with one sequence per lane instead of one per wave: lane L
// of the grid visits colour-order positions L, L + NL, L + 2 NL, ... (NL lanes in the grid), pass
// after pass, each as soon as its two bodies show the awaited counts, independently of the other
// lanes of its wave.  Deadlock-free on a cooperative grid for the same reason as k_solve_flow: the
// earliest unfinished visit is at the head of its lane's sequence with all its predecessors done.
// Each position's data is one contiguous record of 16-byte vectors, so a lane's loads stay
// vectorised whatever position its neighbours are at.
//   header h0 (b1, b2, flags, friction)  flags = usage | mode << 4 | non-static1 << 8 | non-static2 << 9
//          h1 (deg1, rank1, deg2, rank2)  h2 (speculative relaxation factor, -, -, -)
//   slot s  s0 (flag, bias, penalty, mass.n)  s1 (mass.t1, mass.t2, acc.n, acc.t1)
//           s2 (acc.t2, n.x, n.y, n.z)        s3 (t1.x, t1.y, t1.z, t2.x)
//           s4 (t2.y, t2.z, rel1.x, rel1.y)   s5 (rel1.z, rel2.x, rel2.y, rel2.z)
//           s6 (pos1.x, pos1.y, pos1.z, pos2.x)  s7 (pos2.y, pos2.z, -, -)
constexpr int kRecHdr = 3, kRecSlot = 8, kRecVec = kRecHdr + 4 * kRecSlot;

__global__ void __launch_bounds__(256) k_rec_gather(const jh_arbiter* __restrict__ arb, const uint64_t* __restrict__ order,
                                                    const int* __restrict__ offs, const int2* __restrict__ rank,
                                                    const int* __restrict__ deg, const int* __restrict__ motion,
                                                    float4* __restrict__ rec) {
    int A = offs[128];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < A; k += gridDim.x * blockDim.x) {
        const jh_arbiter& a = arb[uint32_t(order[k])];
        float4* r = rec + size_t(k) * kRecVec;
        bool m1 = motion[a.body1] != JH_MOTION_STATIC, m2 = motion[a.body2] != JH_MOTION_STATIC;
        uint32_t flags = (a.usage_mask & 15u) | ((a.mode & 15u) << 4) | (m1 ? 256u : 0u) | (m2 ? 512u : 0u);
        int2 rk = rank[k];
        r[0] = make_float4(__int_as_float(a.body1), __int_as_float(a.body2), __uint_as_float(flags), a.friction);
        r[1] = make_float4(__int_as_float(m1 ? deg[a.body1] : 0), __int_as_float(rk.x),
                           __int_as_float(m2 ? deg[a.body2] : 0), __int_as_float(rk.y));
        r[2] = make_float4(a.speculative_relaxation_factor, 0.0f, 0.0f, 0.0f);
        for (int s = 0; s < 4; ++s) {
            if (!(a.usage_mask & (1u << s))) continue;
            const jh_contact& c = a.contacts[s];
            float4* q = r + kRecHdr + s * kRecSlot;
            q[0] = make_float4(__uint_as_float(c.flag), c.bias, c.penalty_bias, c.mass_normal_tangent[0]);
            q[1] = make_float4(c.mass_normal_tangent[1], c.mass_normal_tangent[2], c.accumulated[0], c.accumulated[1]);
            q[2] = make_float4(c.accumulated[2], c.normal[0], c.normal[1], c.normal[2]);
            q[3] = make_float4(c.tangent1[0], c.tangent1[1], c.tangent1[2], c.tangent2[0]);
            q[4] = make_float4(c.tangent2[1], c.tangent2[2], c.relative_position1[0], c.relative_position1[1]);
            q[5] = make_float4(c.relative_position1[2], c.relative_position2[0], c.relative_position2[1],
                               c.relative_position2[2]);
            q[6] = make_float4(c.position1[0], c.position1[1], c.position1[2], c.position2[0]);
            q[7] = make_float4(c.position2[1], c.position2[2], 0.0f, 0.0f);
        }
    }
}
__global__ void __launch_bounds__(256) k_rec_scatter(jh_arbiter* __restrict__ arb, const uint64_t* __restrict__ order,
                                                     const int* __restrict__ offs, const float4* __restrict__ rec) {
    int A = offs[128];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < A; k += gridDim.x * blockDim.x) {
        jh_arbiter& a = arb[uint32_t(order[k])];
        const float4* r = rec + size_t(k) * kRecVec;
        for (int s = 0; s < 4; ++s) {
            if (!(a.usage_mask & (1u << s))) continue;
            jh_contact& c = a.contacts[s];
            const float4* q = r + kRecHdr + s * kRecSlot;
            float4 q0 = q[0], q1 = q[1], q2 = q[2];
            c.flag = __float_as_uint(q0.x);
            c.bias = q0.y;
            c.penalty_bias = q0.z;
            c.mass_normal_tangent[0] = q0.w;
            c.mass_normal_tangent[1] = q1.x;
            c.mass_normal_tangent[2] = q1.y;
            c.accumulated[0] = q1.z;
            c.accumulated[1] = q1.w;
            c.accumulated[2] = q2.x;
        }
    }
}

__device__ __forceinline__ void load_slot_rec(const float4* q, SlotReg& r, bool prepare) {
    float4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4], q5 = q[5];
    r.flag = __float_as_uint(q0.x);
    r.bias = q0.y;
    r.penalty = q0.z;
    r.mass[0] = q0.w; r.mass[1] = q1.x; r.mass[2] = q1.y;
    r.acc[0] = q1.z; r.acc[1] = q1.w; r.acc[2] = q2.x;
    r.nx[0] = q2.y; r.ny[0] = q2.z; r.nz[0] = q2.w;
    r.nx[1] = q3.x; r.ny[1] = q3.y; r.nz[1] = q3.z;
    r.nx[2] = q3.w; r.ny[2] = q4.x; r.nz[2] = q4.y;
    r.rel1 = vec3{q4.z, q4.w, q5.x};
    r.rel2 = vec3{q5.y, q5.z, q5.w};
    if (prepare) {
        float4 q6 = q[6], q7 = q[7];
        r.pos1 = vec3{q6.x, q6.y, q6.z};
        r.pos2 = vec3{q6.w, q7.x, q7.y};
    }
}
__device__ __forceinline__ void store_slot_rec(float4* q, const SlotReg& r, bool prepare) {
    if (prepare) q[0] = make_float4(__uint_as_float(r.flag), r.bias, r.penalty, r.mass[0]);
    q[1] = make_float4(r.mass[1], r.mass[2], r.acc[0], r.acc[1]);
    q[2] = make_float4(r.acc[2], r.nx[0], r.ny[0], r.nz[0]);
}

// This lane's sequence over passes [g_first, g_first + npass) of one phase: the first pass prepares
// when prep_first, the others iterate (with the penalty bias when `bias`).  False when the launch is
// aborting.
__device__ __forceinline__ bool lane_passes(const SolveArgs& P, float4* __restrict__ rec, int A, int g_first,
                                            int npass, bool prep_first, bool bias) {
    const Bodies& B = P.B;
    int NL = gridDim.x * blockDim.x;
    int L = blockIdx.x * blockDim.x + threadIdx.x;
    int per = A > L ? (A - L + NL - 1) / NL : 0;  // positions of this lane per pass
    int p = 0, i = 0;                              // current pass and position index
    bool finished = per == 0 || npass == 0;
    bool loaded = false, ok1 = true, ok2 = true, prep = false, m1 = false, m2 = false;
    int k = 0, b1 = 0, b2 = 0, need1 = 0, need2 = 0;
    uint32_t usage = 0, mode = 0;
    float friction = 0.0f, specrel = 0.0f;
    SlotReg sr[4];
    BodyReg R1, R2;
    vec3 x1{0, 0, 0}, x2{0, 0, 0};
    quat q1{0, 0, 0, 1}, q2{0, 0, 0, 1};
    __amdgpu_buffer_rsrc_t vr = vw_rsrc(P);
    unsigned polls = 0;
    while (true) {
        asm volatile("" ::: "memory");
        if (!finished && !loaded) {
            k = L + i * NL;
            prep = prep_first && p == 0;
            const float4* r = rec + size_t(k) * kRecVec;
            float4 h0 = r[0], h1 = r[1];
            b1 = __float_as_int(h0.x);
            b2 = __float_as_int(h0.y);
            uint32_t fl = __float_as_uint(h0.z);
            friction = h0.w;
            usage = fl & 15u;
            mode = (fl >> 4) & 15u;
            m1 = (fl & 256u) != 0;
            m2 = (fl & 512u) != 0;
            int g = g_first + p;
            need1 = g * __float_as_int(h1.x) + __float_as_int(h1.y);
            need2 = g * __float_as_int(h1.z) + __float_as_int(h1.w);
#pragma unroll
            for (int s = 0; s < 4; ++s)
                if (usage & (1u << s)) load_slot_rec(r + kRecHdr + s * kRecSlot, sr[s], prep);
            load_static_part(B, b1, R1);
            load_static_part(B, b2, R2);
            if (prep) {
                specrel = r[2].x;
                x1 = ld3(B.pos[b1]);
                x2 = ld3(B.pos[b2]);
                q1 = ldq(B.quat[b1]);
                q2 = ldq(B.quat[b2]);
            }
            if (!m1) ld_vw(P, b1, R1.v, R1.w);
            if (!m2) ld_vw(P, b2, R2.v, R2.w);
            ok1 = !m1;
            ok2 = !m2;
            loaded = true;
        }
        bool progressed = false;
        if (loaded) {
            if (!ok1) ok1 = coh_poll_vw(vr, b1, need1, R1.v, R1.w);
            if (!ok2) ok2 = coh_poll_vw(vr, b2, need2, R2.v, R2.w);
            if (ok1 && ok2) {
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    if (!(usage & (1u << s))) continue;
                    if (prep) prepare_reg(mode, specrel, sr[s], R1, R2, x1, q1, x2, q2, P.ts);
                    else iterate_reg(mode, friction, sr[s], R1, R2, bias);
                }
                if (m1) coh_publish_vw(vr, b1, R1.v, R1.w, need1 + 1);
                if (m2) coh_publish_vw(vr, b2, R2.v, R2.w, need2 + 1);
                float4* r = rec + size_t(k) * kRecVec;
#pragma unroll
                for (int s = 0; s < 4; ++s)
                    if (usage & (1u << s)) store_slot_rec(r + kRecHdr + s * kRecSlot, sr[s], prep);
                loaded = false;
                progressed = true;
                if (++i == per) {
                    i = 0;
                    if (++p == npass) finished = true;
                }
            }
        }
        if (__all(finished)) return true;
        if (__any(progressed)) {
            polls = 0;
        } else {
            if (!keep_polling(P.bar, polls, P.spin)) return false;
            __builtin_amdgcn_s_sleep(2);
        }
    }
}

// One launch per step, like k_solve_flow, with per-lane sequences in the dataflow phases.
__global__ void __launch_bounds__(256) k_solve_lanes(SolveArgs P, float4* __restrict__ rec) {
    int stride = gridDim.x * blockDim.x;
    int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    int per_sub = 1 + P.iterations + P.relaxations;
    int A = P.offs[128];
    const Bodies& B = P.B;
    for (int b = t0; b < P.nb; b += stride) st_vw(P, b, ld3(B.vel[b]), ld3(B.angvel[b]), 0);
    for (int s = 0; s < P.ts.substeps; ++s) {
        for (int b = t0; b < P.nb; b += stride) {  // IntegrateForces (World.Step.cs:1060-1072)
            if (B.motion[b] != JH_MOTION_DYNAMIC) continue;
            vec3 v, w;
            int n = ld_vw(P, b, v, w);
            w = w + ld3(B.dangvel[b]);
            v = v + ld3(B.dvel[b]);
            st_vw(P, b, v, w, n);
        }
        if (!grid_barrier(P.bar, P.spin)) return;
        bool ok = lane_passes(P, rec, A, s * per_sub, 1 + P.iterations, true, true);
        if (!grid_barrier(P.bar, P.spin, ok)) return;
        for (int b = t0; b < P.nb; b += stride) {  // IntegrateVelocities (World.Step.cs:1106-1133)
            if (B.motion[b] == JH_MOTION_STATIC) continue;
            vec3 v, w;
            ld_vw(P, b, v, w);
            B.pos[b] = st3(ld3(B.pos[b]) + v * P.ts.substep_dt);
            quat rq = rotation_quaternion(w, P.ts.substep_dt);
            B.quat[b] = stq(qnormalize(qmul(rq, ldq(B.quat[b]))));
        }
        if (!grid_barrier(P.bar, P.spin)) return;
        if (P.relaxations > 0) {
            ok = lane_passes(P, rec, A, s * per_sub + 1 + P.iterations, P.relaxations, false, false);
            if (!grid_barrier(P.bar, P.spin, ok)) return;
        }
    }
    for (int b = t0; b < P.nb; b += stride) {
        vec3 v, w;
        ld_vw(P, b, v, w);
        B.vel[b] = st3(v);
        B.angvel[b] = st3(w);
    }
}

// ---- World-local solver --------------------------------------------------------------------
// For batches of many independent worlds: one workgroup solves one world at a time (worlds taken from
// a queue, largest first), walking that world's arbiters colour by colour with a workgroup barrier
// between colours.  Arbiters of one colour touch disjoint non-static bodies, and each body sees its
// arbiters in colour order (the global solve order restricted to its world), so the results equal
// the colour-batch solver's bit for bit.  All accesses of a world come from one workgroup on one CU,
// so plain loads and stores ordered by the barriers are coherent: no device-scope atomics.
struct WorldArgs {
    const int* wbody;     // body indices grouped by world
    const int* wboff;     // [W + 1]
    const int* wsched;    // worlds, largest first
    const uint32_t* woff; // [W * 128 + 1]: start of (world, colour) in the world-ordered arbiter list
    int W;
    int* queue;           // next world to take
};

// What a world-local visit reads first (bodies, used slots and mode, friction); the next batch's is
// fetched before the batch barrier, so a visit starts with one round of loads, not two.
struct WorldHead {
    int2 bb;
    uint2 um;
    float2 fs;
};
__device__ __forceinline__ WorldHead world_head(const ContactSoA& S, int p) { return {S.bodies[p], S.um[p], S.fs[p]}; }

template <bool PREPARE>
__device__ __forceinline__ void world_visit(const SolveArgs& P, int p, bool apply_bias, const WorldHead& h) {
    const Bodies& B = P.B;
    const ContactSoA& S = P.soa;
    int2 bb = h.bb;
    uint2 um = h.um;
    float2 fs = h.fs;
    int b1 = bb.x, b2 = bb.y;
    uint32_t usage = um.x, mode = um.y;
    SlotReg sr[4];
#pragma unroll
    for (int s = 0; s < 4; ++s)
        if (usage & (1u << s)) load_slot_soa(S, s, p, sr[s], PREPARE);
    BodyReg R1, R2;
    load_static_part(B, b1, R1);
    load_static_part(B, b2, R2);
    R1.v = ld3(B.vel[b1]); R1.w = ld3(B.angvel[b1]);
    R2.v = ld3(B.vel[b2]); R2.w = ld3(B.angvel[b2]);
    vec3 x1, x2;
    quat q1, q2;
    if (PREPARE) {
        x1 = ld3(B.pos[b1]);
        x2 = ld3(B.pos[b2]);
        q1 = ldq(B.quat[b1]);
        q2 = ldq(B.quat[b2]);
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (!(usage & (1u << s))) continue;
        if (PREPARE) prepare_reg(mode, fs.y, sr[s], R1, R2, x1, q1, x2, q2, P.ts);
        else iterate_reg(mode, fs.x, sr[s], R1, R2, apply_bias);
    }
    if (mode & (JH_SOLVE_LINEAR1 | JH_SOLVE_ANGULAR1)) { B.vel[b1] = st3(R1.v); B.angvel[b1] = st3(R1.w); }
    if (mode & (JH_SOLVE_LINEAR2 | JH_SOLVE_ANGULAR2)) { B.vel[b2] = st3(R2.v); B.angvel[b2] = st3(R2.w); }
#pragma unroll
    for (int s = 0; s < 4; ++s)
        if (usage & (1u << s)) store_slot_soa(S, s, p, sr[s], PREPARE);
}

// One pass over a world's colour batches.  segs lists the non-empty batches' bounds (nseg batches,
// the serial batch last when present); h holds this thread's first visit's head for batch 0 and
// returns the one for batch 0 of the next pass.
__device__ __forceinline__ void world_pass(const SolveArgs& P, const uint32_t* segs, int nseg, bool serial,
                                           int phase, WorldHead& h) {
    const ContactSoA& S = P.soa;
    int t = int(threadIdx.x);
    for (int j = 0; j < nseg; ++j) {
        int beg = int(segs[j]), end = int(segs[j + 1]);
        bool ser = serial && j == nseg - 1;
        if (!ser) {
            for (int p = beg + t; p < end; p += blockDim.x) {
                WorldHead hp = p == beg + t ? h : world_head(S, p);
                if (phase == 0) world_visit<true>(P, p, false, hp);
                else world_visit<false>(P, p, phase == 1, hp);
            }
        } else if (t == 0) {  // the serial batch, in solve order
            for (int p = beg; p < end; ++p) {
                WorldHead hp = world_head(S, p);
                if (phase == 0) world_visit<true>(P, p, false, hp);
                else world_visit<false>(P, p, phase == 1, hp);
            }
        }
        // the next batch's first head (batch 0 again after the last one): no visit writes heads
        int jn = j + 1 < nseg ? j + 1 : 0;
        int pn = int(segs[jn]) + t;
        if (pn < int(segs[jn + 1])) h = world_head(S, pn);
        __syncthreads();
    }
}

__global__ void __launch_bounds__(512) k_solve_worlds(SolveArgs P, WorldArgs Wd) {
    __shared__ uint32_t segs[kMaxColours + 2];
    __shared__ int take, nseg_s, serial_s;
    const Bodies& B = P.B;
    while (true) {
        if (threadIdx.x == 0) take = atomicAdd(Wd.queue, 1);
        __syncthreads();
        int q = take;
        if (q >= Wd.W) break;
        int w = Wd.wsched[q];
        int bo = Wd.wboff[w], be = Wd.wboff[w + 1];
        if (threadIdx.x == 0) {  // bounds of the world's non-empty batches, serial batch last
            const uint32_t* wo = Wd.woff + size_t(w) * 128;
            int n = 0;
            uint32_t prev = wo[0];
            for (int c = 0; c <= kMaxColours; ++c) {
                uint32_t nx = wo[c + 1];
                if (nx != prev) segs[n++] = prev;
                prev = nx;
            }
            segs[n] = prev;
            nseg_s = n;
            serial_s = wo[kMaxColours + 1] != wo[kMaxColours];
        }
        __syncthreads();
        int nseg = nseg_s;
        bool serial = serial_s != 0;
        WorldHead h{};
        if (nseg > 0 && int(segs[0]) + int(threadIdx.x) < int(segs[1])) h = world_head(P.soa, int(segs[0]) + threadIdx.x);
        for (int s = 0; s < P.ts.substeps; ++s) {
            for (int i = bo + int(threadIdx.x); i < be; i += blockDim.x) {  // IntegrateForces
                int b = Wd.wbody[i];
                if (B.motion[b] != JH_MOTION_DYNAMIC) continue;
                B.angvel[b] = st3(ld3(B.angvel[b]) + ld3(B.dangvel[b]));
                B.vel[b] = st3(ld3(B.vel[b]) + ld3(B.dvel[b]));
            }
            __syncthreads();
            for (int it = 0; it <= P.iterations; ++it) world_pass(P, segs, nseg, serial, it == 0 ? 0 : 1, h);
            for (int i = bo + int(threadIdx.x); i < be; i += blockDim.x) {  // IntegrateVelocities
                int b = Wd.wbody[i];
                if (B.motion[b] == JH_MOTION_STATIC) continue;
                vec3 pnew = ld3(B.pos[b]) + ld3(B.vel[b]) * P.ts.substep_dt;
                B.pos[b] = st3(pnew);
                quat rq = rotation_quaternion(ld3(B.angvel[b]), P.ts.substep_dt);
                B.quat[b] = stq(qnormalize(qmul(rq, ldq(B.quat[b]))));
            }
            __syncthreads();
            for (int it = 0; it < P.relaxations; ++it) world_pass(P, segs, nseg, serial, 2, h);
        }
        __syncthreads();
    }
}

// (world, colour) histogram of the solved arbiters; world of an arbiter = world of its first body.
__global__ void k_world_hist(const uint64_t* __restrict__ order, const int* __restrict__ offs,
                             const jh_arbiter* __restrict__ arb, const int* __restrict__ colour,
                             const int* __restrict__ body_world, uint32_t* __restrict__ hist) {
    int A = offs[128];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < A; k += gridDim.x * blockDim.x) {
        uint32_t a = uint32_t(order[k]);
        int c = colour[a];
        atomicAdd(&hist[size_t(body_world[arb[a].body1]) * 128 + c], 1u);
    }
}
// World-ordered arbiter list; within one (world, colour) the slot order is free (the arbiters are
// independent) except in the serial batch, which k_world_serial fills in solve order.
__global__ void k_world_scatter(const uint64_t* __restrict__ order, const int* __restrict__ offs,
                                const jh_arbiter* __restrict__ arb, const int* __restrict__ colour,
                                const int* __restrict__ body_world, const uint32_t* __restrict__ woff,
                                uint32_t* __restrict__ cursor, uint64_t* __restrict__ wlist) {
    int A = offs[kMaxColours];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < A; k += gridDim.x * blockDim.x) {
        uint32_t a = uint32_t(order[k]);
        size_t cell = size_t(body_world[arb[a].body1]) * 128 + colour[a];
        wlist[woff[cell] + atomicAdd(&cursor[cell], 1u)] = a;
    }
}
__global__ void k_world_serial(const uint64_t* __restrict__ order, const int* __restrict__ offs,
                               const jh_arbiter* __restrict__ arb, const int* __restrict__ body_world,
                               const uint32_t* __restrict__ woff, uint32_t* __restrict__ cursor,
                               uint64_t* __restrict__ wlist) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int k = offs[kMaxColours]; k < offs[kMaxColours + 1]; ++k) {
        uint32_t a = uint32_t(order[k]);
        size_t cell = size_t(body_world[arb[a].body1]) * 128 + kMaxColours;
        wlist[woff[cell] + cursor[cell]++] = a;
    }
}
__global__ void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = v;
}

// Per-step set-up of the dataflow solver.
// Ranks and degrees from the colouring's per-body colour masks, without sorting.  After k_colour,
// mask_lo/mask_hi of every solver-non-static body that an active arbiter touches hold exactly the
// colours 0..126 of its arbiters, and no two of them share a colour, so in the colour order
// (colours ascending, the serial batch 127 last) the arbiters of body b before one of colour c are
// those of b's colours below c: rank = popcount of b's mask below c.  The serial batch may hold
// several arbiters of one body; their ranks continue after all of b's coloured arbiters, in
// position order (k_rank_serial).  deg = b's coloured arbiters + its serial ones.  Same values as
// sorting the (body, position) keys.
__device__ __forceinline__ int colour_of_position(const uint64_t* order, int k) {
    return int((order[k] >> 34) & 0xffu);  // solve-order key: (colour << 2 | contacts - 1) << 32 | arbiter
}
__device__ __forceinline__ int mask_rank(const unsigned long long* lo, const unsigned long long* hi, int b, int c) {
    unsigned long long l = lo[b], h = hi[b];
    if (c >= kMaxColours) return __popcll(l) + __popcll(h);
    if (c < 64) return __popcll(l & ((1ull << c) - 1ull));
    return __popcll(l) + __popcll(h & ((1ull << (c - 64)) - 1ull));
}
// Serial counts of the bodies of the solve order cleared.
__global__ void k_rank_clear(const jh_arbiter* __restrict__ arb, const uint64_t* __restrict__ order,
                             const int* __restrict__ offs, int* __restrict__ scnt) {
    int A = offs[128];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < A; k += gridDim.x * blockDim.x) {
        const jh_arbiter& a = arb[uint32_t(order[k])];
        scnt[a.body1] = 0;
        scnt[a.body2] = 0;
    }
}
// One workgroup walks the serial batch [offs[127], offs[128]) in 256-position chunks: a position's
// rank among its body's serial arbiters = the body's count from earlier chunks + the earlier
// positions of this chunk touching the body.  Writes that rank to rank[k] (completed by
// k_rank_masks) and leaves each body's serial count in scnt.
__global__ void __launch_bounds__(256) k_rank_serial(const jh_arbiter* __restrict__ arb,
                                                     const uint64_t* __restrict__ order, const int* __restrict__ offs,
                                                     const int* __restrict__ motion, int2* __restrict__ rank,
                                                     int* __restrict__ scnt) {
    __shared__ int sb[2][256];
    int s0 = offs[kMaxColours], s1 = offs[128];
    int t = threadIdx.x;
    for (int base = s0; base < s1; base += 256) {
        int k = base + t;
        int b1 = -1, b2 = -1;
        if (k < s1) {
            const jh_arbiter& a = arb[uint32_t(order[k])];
            if (motion[a.body1] != JH_MOTION_STATIC) b1 = a.body1;
            if (motion[a.body2] != JH_MOTION_STATIC) b2 = a.body2;
        }
        sb[0][t] = b1;
        sb[1][t] = b2;
        __syncthreads();
        int r1 = b1 >= 0 ? __hip_atomic_load(&scnt[b1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
        int r2 = b2 >= 0 ? __hip_atomic_load(&scnt[b2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
        for (int u = 0; u < t; ++u) {
            int c1 = sb[0][u], c2 = sb[1][u];
            if (b1 >= 0) r1 += int(c1 == b1) + int(c2 == b1);
            if (b2 >= 0) r2 += int(c1 == b2) + int(c2 == b2);
        }
        if (k < s1) rank[k] = make_int2(r1, r2);
        __syncthreads();  // every count of this chunk is read before any is raised
        if (b1 >= 0) __hip_atomic_fetch_add(&scnt[b1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (b2 >= 0) __hip_atomic_fetch_add(&scnt[b2], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_s_waitcnt(0);  // the additions are performed before the next chunk reads
        __syncthreads();
    }
}
__global__ void k_rank_masks(const jh_arbiter* __restrict__ arb, const uint64_t* __restrict__ order,
                             const int* __restrict__ offs, const int* __restrict__ motion,
                             const unsigned long long* __restrict__ mask_lo,
                             const unsigned long long* __restrict__ mask_hi, const int* __restrict__ scnt,
                             int2* __restrict__ rank, int* __restrict__ deg, int* __restrict__ max_deg) {
    int A = offs[128];
    int m = 0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < A; k += gridDim.x * blockDim.x) {
        const jh_arbiter& a = arb[uint32_t(order[k])];
        int c = colour_of_position(order, k);
        int2 r = c >= kMaxColours ? rank[k] : make_int2(0, 0);
        int b[2] = {a.body1, a.body2};
        int rr[2] = {r.x, r.y};
        for (int e = 0; e < 2; ++e) {
            if (motion[b[e]] == JH_MOTION_STATIC) continue;
            rr[e] += mask_rank(mask_lo, mask_hi, b[e], c);
            int d = mask_rank(mask_lo, mask_hi, b[e], kMaxColours) + scnt[b[e]];
            deg[b[e]] = d;  // the same value from each arbiter of the body
            m = max(m, d);
        }
        rank[k] = make_int2(rr[0], rr[1]);
    }
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0 && m > 0) atomicMax(max_deg, m);
}

// ---- Tile solver ----------------------------------------------------------------------------
// One workgroup per tile: a contiguous range of the non-static bodies in broadphase (Morton) order.
// Each arbiter belongs to the tile of its first non-static body.  A body is *private* when every
// arbiter touching it belongs to its own tile: its v/w live in LDS for the whole launch and its
// visits are ordered by workgroup barriers alone (the tile walks its arbiters colour by colour).
// Other non-static bodies are *shared*: their v/w live in the packed global records and their
// visits are ordered by the dataflow counters (need = g * deg + rank, as in k_solve_flow).  Every
// body therefore sees its visits in colour order, so results equal the colour-batch solver's bit
// for bit.  No deadlock on a cooperative grid: the tile whose current (pass, colour) is smallest
// only waits for visits at strictly earlier positions, all finished.
struct TileArgs {
    const int* tbody;       // non-static bodies with shapes, Morton order; tile t holds [tile_begin(t), tile_begin(t+1))
    const int* nmov;        // number of entries of tbody
    const int* tile_of;     // per body: tile, -1 for static bodies and bodies without shapes
    const int* tslot;       // per body: index within its tile
    const uint32_t* shared; // per body: 1 if an arbiter of another tile touches it
    const uint32_t* toff;   // [T * 128 + 1]: start of (tile, colour) in tlist
    const uint64_t* tlist;  // (colour-order position << 32) | arbiter, tile-major, colour within tile
    int T;
};
__device__ __forceinline__ int tile_begin(int t, int M, int T) { return int((int64_t(t) * M + T - 1) / T); }

// Body -> first position of one of its shapes in the sorted broadphase order.
__global__ void k_body_first(const uint32_t* __restrict__ sorted, const int* __restrict__ ns_ptr,
                             const ShapeRec* __restrict__ shapes, int* __restrict__ first) {
    int ns = *ns_ptr;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += gridDim.x * blockDim.x)
        atomicMin(&first[shapes[sorted[i]].body], i);
}
__global__ void k_tile_flags(const uint32_t* __restrict__ sorted, const int* __restrict__ ns_ptr,
                             const ShapeRec* __restrict__ shapes, const int* __restrict__ motion,
                             const int* __restrict__ first, uint32_t* __restrict__ flags) {
    int ns = *ns_ptr;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += gridDim.x * blockDim.x) {
        int b = shapes[sorted[i]].body;
        flags[i] = (motion[b] != JH_MOTION_STATIC && first[b] == i) ? 1u : 0u;
    }
}
__global__ void k_tile_assign(const uint32_t* __restrict__ sorted, const int* __restrict__ ns_ptr,
                              const ShapeRec* __restrict__ shapes, const uint32_t* __restrict__ flags,
                              const uint32_t* __restrict__ rank, const int* __restrict__ nmov, int T,
                              int* __restrict__ tbody, int* __restrict__ tile_of, int* __restrict__ tslot) {
    int ns = *ns_ptr, M = *nmov;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += gridDim.x * blockDim.x) {
        if (!flags[i]) continue;
        int b = shapes[sorted[i]].body;
        int r = int(rank[i]);
        int t = int(int64_t(r) * T / M);
        tbody[r] = b;
        tile_of[b] = t;
        tslot[b] = r - tile_begin(t, M, T);
    }
}
__device__ __forceinline__ int arbiter_tile(const jh_arbiter& a, const int* motion, const int* tile_of) {
    return motion[a.body1] != JH_MOTION_STATIC ? tile_of[a.body1] : tile_of[a.body2];
}
// (tile, colour) histogram of the solved arbiters, and the shared flags of their bodies.
__global__ void k_tile_hist(const uint64_t* __restrict__ order, const int* __restrict__ offs,
                            const jh_arbiter* __restrict__ arb, const int* __restrict__ colour,
                            const int* __restrict__ motion, const int* __restrict__ tile_of,
                            uint32_t* __restrict__ hist, uint32_t* __restrict__ shared) {
    int A = offs[128];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < A; k += gridDim.x * blockDim.x) {
        uint32_t ai = uint32_t(order[k]);
        const jh_arbiter& a = arb[ai];
        int t = arbiter_tile(a, motion, tile_of);
        atomicAdd(&hist[size_t(t) * 128 + colour[ai]], 1u);
        if (motion[a.body1] != JH_MOTION_STATIC && tile_of[a.body1] != t) shared[a.body1] = 1u;
        if (motion[a.body2] != JH_MOTION_STATIC && tile_of[a.body2] != t) shared[a.body2] = 1u;
    }
}
// Tile-ordered list; within one (tile, colour) the order is free (the arbiters are independent)
// except in the serial batch, which k_tile_serial fills in solve order.
__global__ void k_tile_scatter(const uint64_t* __restrict__ order, const int* __restrict__ offs,
                               const jh_arbiter* __restrict__ arb, const int* __restrict__ colour,
                               const int* __restrict__ motion, const int* __restrict__ tile_of,
                               const uint32_t* __restrict__ toff, uint32_t* __restrict__ cursor,
                               uint64_t* __restrict__ tlist) {
    int A = offs[kMaxColours];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < A; k += gridDim.x * blockDim.x) {
        uint32_t ai = uint32_t(order[k]);
        size_t cell = size_t(arbiter_tile(arb[ai], motion, tile_of)) * 128 + colour[ai];
        tlist[toff[cell] + atomicAdd(&cursor[cell], 1u)] = (uint64_t(uint32_t(k)) << 32) | ai;
    }
}
__global__ void k_tile_serial(const uint64_t* __restrict__ order, const int* __restrict__ offs,
                              const jh_arbiter* __restrict__ arb, const int* __restrict__ motion,
                              const int* __restrict__ tile_of, const uint32_t* __restrict__ toff,
                              uint32_t* __restrict__ cursor, uint64_t* __restrict__ tlist) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int k = offs[kMaxColours]; k < offs[kMaxColours + 1]; ++k) {
        uint32_t ai = uint32_t(order[k]);
        size_t cell = size_t(arbiter_tile(arb[ai], motion, tile_of)) * 128 + kMaxColours;
        tlist[toff[cell] + cursor[cell]++] = (uint64_t(uint32_t(k)) << 32) | ai;
    }
}

// One visit of the arbiter at tile-list position p in global pass g.  Private bodies are read and
// written in LDS; shared bodies wait for their counters and go through the coherent records.  False
// when the launch is aborting (then nothing is computed).
template <bool PREPARE>
__device__ __forceinline__ bool tile_visit(const SolveArgs& P, const TileArgs& Tl, float4* lv, float2* lw, int p,
                                           int g, bool apply_bias) {
    const Bodies& B = P.B;
    const ContactSoA& S = P.soa;
    int2 bb = S.bodies[p];
    uint2 um = S.um[p];
    float2 fs = S.fs[p];
    int b1 = bb.x, b2 = bb.y;
    uint32_t usage = um.x, mode = um.y;
    bool m1 = B.motion[b1] != JH_MOTION_STATIC, m2 = B.motion[b2] != JH_MOTION_STATIC;
    bool sh1 = m1 && Tl.shared[b1] != 0u, sh2 = m2 && Tl.shared[b2] != 0u;
    int need1 = 0, need2 = 0;
    if (sh1 || sh2) {
        int2 r = P.rank[int(Tl.tlist[p] >> 32)];
        if (sh1) need1 = g * P.deg[b1] + r.x;
        if (sh2) need2 = g * P.deg[b2] + r.y;
    }
    SlotReg sr[4];
#pragma unroll
    for (int s = 0; s < 4; ++s)
        if (usage & (1u << s)) load_slot_soa(S, s, p, sr[s], PREPARE);
    BodyReg R1, R2;
    load_static_part(B, b1, R1);
    load_static_part(B, b2, R2);
    vec3 x1, x2;
    quat q1, q2;
    if (PREPARE) {
        x1 = ld3(B.pos[b1]);
        x2 = ld3(B.pos[b2]);
        q1 = ldq(B.quat[b1]);
        q2 = ldq(B.quat[b2]);
    }
    int s1 = (m1 && !sh1) ? Tl.tslot[b1] : 0, s2 = (m2 && !sh2) ? Tl.tslot[b2] : 0;
    if (!m1) { R1.v = ld3(B.vel[b1]); R1.w = ld3(B.angvel[b1]); }
    else if (!sh1) { float4 a = lv[s1]; float2 c = lw[s1]; R1.v = {a.x, a.y, a.z}; R1.w = {a.w, c.x, c.y}; }
    if (!m2) { R2.v = ld3(B.vel[b2]); R2.w = ld3(B.angvel[b2]); }
    else if (!sh2) { float4 a = lv[s2]; float2 c = lw[s2]; R2.v = {a.x, a.y, a.z}; R2.w = {a.w, c.x, c.y}; }
    __amdgpu_buffer_rsrc_t vr = vw_rsrc(P);
    bool done = false, ok1 = !sh1, ok2 = !sh2, gave_up = false;
    unsigned polls = 0;
    while (true) {
        asm volatile("" ::: "memory");
        if (!done && !(ok1 && ok2) && !keep_polling(P.bar, polls, P.spin)) done = gave_up = true;
        if (!done) {
            if (!ok1) ok1 = coh_poll_vw(vr, b1, need1, R1.v, R1.w);
            if (!ok2) ok2 = coh_poll_vw(vr, b2, need2, R2.v, R2.w);
            if (ok1 && ok2) {
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    if (!(usage & (1u << s))) continue;
                    if (PREPARE) prepare_reg(mode, fs.y, sr[s], R1, R2, x1, q1, x2, q2, P.ts);
                    else iterate_reg(mode, fs.x, sr[s], R1, R2, apply_bias);
                }
                bool w1 = (mode & (JH_SOLVE_LINEAR1 | JH_SOLVE_ANGULAR1)) != 0;
                bool w2 = (mode & (JH_SOLVE_LINEAR2 | JH_SOLVE_ANGULAR2)) != 0;
                if (sh1) coh_publish_vw(vr, b1, R1.v, R1.w, need1 + 1);
                else if (w1) { lv[s1] = make_float4(R1.v.x, R1.v.y, R1.v.z, R1.w.x); lw[s1] = make_float2(R1.w.y, R1.w.z); }
                if (sh2) coh_publish_vw(vr, b2, R2.v, R2.w, need2 + 1);
                else if (w2) { lv[s2] = make_float4(R2.v.x, R2.v.y, R2.v.z, R2.w.x); lw[s2] = make_float2(R2.w.y, R2.w.z); }
#pragma unroll
                for (int s = 0; s < 4; ++s)
                    if (usage & (1u << s)) store_slot_soa(S, s, p, sr[s], PREPARE);
                done = true;
            }
        }
        if (__all(done)) return !gave_up;
        __builtin_amdgcn_s_sleep(2);
    }
}

// One pass of a tile: its arbiters colour by colour, a workgroup barrier after each non-empty
// colour; the serial batch by one thread in solve order.  False (for the whole workgroup) when the
// launch is aborting.
__device__ __forceinline__ bool tile_pass(const SolveArgs& P, const TileArgs& Tl, float4* lv, float2* lw,
                                          const uint32_t* segs, int nc, int g, int phase) {
    bool ok = true;
    for (int c = 0; c < nc; ++c) {
        int beg = int(segs[c]), end = int(segs[c + 1]);
        if (beg == end) continue;
        for (int p = beg + int(threadIdx.x); p < end && ok; p += blockDim.x)
            ok = phase == 0 ? tile_visit<true>(P, Tl, lv, lw, p, g, false)
                            : tile_visit<false>(P, Tl, lv, lw, p, g, phase == 1);
        if (__syncthreads_or(!ok)) return false;
    }
    int beg = int(segs[kMaxColours]), end = int(segs[kMaxColours + 1]);
    if (beg < end) {
        if (threadIdx.x == 0)
            for (int p = beg; p < end && ok; ++p)
                ok = phase == 0 ? tile_visit<true>(P, Tl, lv, lw, p, g, false)
                                : tile_visit<false>(P, Tl, lv, lw, p, g, phase == 1);
        if (__syncthreads_or(!ok)) return false;
    }
    return true;
}

// One launch per step, grid = T tiles (cooperative); dynamic LDS holds 24 bytes per tile body.
__global__ void __launch_bounds__(256) k_solve_tiles(SolveArgs P, TileArgs Tl) {
    extern __shared__ float4 tile_lds[];
    __shared__ uint32_t segs[kMaxColours + 2];
    const Bodies& B = P.B;
    int t = blockIdx.x;
    int M = *Tl.nmov;
    int tb = tile_begin(t, M, Tl.T), te = tile_begin(t + 1, M, Tl.T);
    int cap = te - tb;
    float4* lv = tile_lds;
    float2* lw = (float2*)(tile_lds + cap);
    int nc = *P.ncolours;
    for (int c = threadIdx.x; c <= kMaxColours + 1; c += blockDim.x) segs[c] = Tl.toff[size_t(t) * 128 + c];
    int per_sub = 1 + P.iterations + P.relaxations;
    int stride = gridDim.x * blockDim.x;
    int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    // v/w: private bodies into LDS, shared bodies into the packed records
    for (int i = tb + int(threadIdx.x); i < te; i += blockDim.x) {
        int b = Tl.tbody[i];
        vec3 v = ld3(B.vel[b]), w = ld3(B.angvel[b]);
        if (Tl.shared[b]) st_vw(P, b, v, w, 0);
        else { lv[i - tb] = make_float4(v.x, v.y, v.z, w.x); lw[i - tb] = make_float2(w.y, w.z); }
    }
    for (int s = 0; s < P.ts.substeps; ++s) {
        // IntegrateForces (World.Step.cs:1060-1072)
        for (int i = tb + int(threadIdx.x); i < te; i += blockDim.x) {
            int b = Tl.tbody[i];
            if (B.motion[b] != JH_MOTION_DYNAMIC) continue;
            vec3 v, w;
            bool sh = Tl.shared[b] != 0u;
            int l = i - tb, n = 0;
            if (sh) n = ld_vw(P, b, v, w);
            else { float4 a = lv[l]; float2 c = lw[l]; v = {a.x, a.y, a.z}; w = {a.w, c.x, c.y}; }
            w = w + ld3(B.dangvel[b]);
            v = v + ld3(B.dvel[b]);
            if (sh) st_vw(P, b, v, w, n);
            else { lv[l] = make_float4(v.x, v.y, v.z, w.x); lw[l] = make_float2(w.y, w.z); }
        }
        for (int b = t0; b < P.nb; b += stride) {  // dynamic bodies without shapes touch no arbiter
            if (B.motion[b] != JH_MOTION_DYNAMIC || Tl.tile_of[b] >= 0) continue;
            B.angvel[b] = st3(ld3(B.angvel[b]) + ld3(B.dangvel[b]));
            B.vel[b] = st3(ld3(B.vel[b]) + ld3(B.dvel[b]));
        }
        if (!grid_barrier(P.bar, P.spin)) return;
        for (int it = 0; it <= P.iterations; ++it)
            if (!tile_pass(P, Tl, lv, lw, segs, nc, s * per_sub + it, it == 0 ? 0 : 1)) return;
        if (!grid_barrier(P.bar, P.spin)) return;
        // IntegrateVelocities (World.Step.cs:1106-1133)
        for (int i = tb + int(threadIdx.x); i < te; i += blockDim.x) {
            int b = Tl.tbody[i];
            vec3 v, w;
            if (Tl.shared[b]) ld_vw(P, b, v, w);
            else { float4 a = lv[i - tb]; float2 c = lw[i - tb]; v = {a.x, a.y, a.z}; w = {a.w, c.x, c.y}; }
            B.pos[b] = st3(ld3(B.pos[b]) + v * P.ts.substep_dt);
            quat rq = rotation_quaternion(w, P.ts.substep_dt);
            B.quat[b] = stq(qnormalize(qmul(rq, ldq(B.quat[b]))));
        }
        for (int b = t0; b < P.nb; b += stride) {
            if (B.motion[b] == JH_MOTION_STATIC || Tl.tile_of[b] >= 0) continue;
            B.pos[b] = st3(ld3(B.pos[b]) + ld3(B.vel[b]) * P.ts.substep_dt);
            quat rq = rotation_quaternion(ld3(B.angvel[b]), P.ts.substep_dt);
            B.quat[b] = stq(qnormalize(qmul(rq, ldq(B.quat[b]))));
        }
        if (!grid_barrier(P.bar, P.spin)) return;
        if (P.relaxations > 0) {
            for (int it = 0; it < P.relaxations; ++it)
                if (!tile_pass(P, Tl, lv, lw, segs, nc, s * per_sub + 1 + P.iterations + it, 2)) return;
            if (!grid_barrier(P.bar, P.spin)) return;
        }
    }
    for (int i = tb + int(threadIdx.x); i < te; i += blockDim.x) {
        int b = Tl.tbody[i];
        vec3 v, w;
        if (Tl.shared[b]) ld_vw(P, b, v, w);
        else { float4 a = lv[i - tb]; float2 c = lw[i - tb]; v = {a.x, a.y, a.z}; w = {a.w, c.x, c.y}; }
        B.vel[b] = st3(v);
        B.angvel[b] = st3(w);
    }
}

// ---- End of step ----------------------------------------------------------------------------
// ContactData.UpdatePosition (Contact.cs:238-264, 762-797); flags arbiters that keep contacts.
// Only the active contacts (memContacts.Active: a non-static body awake; smotion = the solver's
// motion types, a sleeping body static); the others keep their data.
__global__ void k_update_contacts(jh_arbiter* __restrict__ arb, const int* __restrict__ a_count, Bodies B,
                                  uint32_t* __restrict__ keep, const int* __restrict__ smotion) {
    int A = *a_count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < A; i += gridDim.x * blockDim.x) {
        jh_arbiter& cd = arb[i];
        if (smotion[cd.body1] == JH_MOTION_STATIC && smotion[cd.body2] == JH_MOTION_STATIC) {
            keep[i] = 1u;
            continue;
        }
        uint32_t usage = cd.usage_mask & 15u;
        usage |= usage << 4;
        vec3 x1 = ld3(B.pos[cd.body1]), x2 = ld3(B.pos[cd.body2]);
        quat q1 = ldq(B.quat[cd.body1]), q2 = ldq(B.quat[cd.body2]);
        for (int s = 0; s < 4; ++s) {
            if ((usage & (1u << s)) == 0) continue;
            jh_contact& c = cd.contacts[s];
            vec3 rel1 = rot(vec3{c.position1[0], c.position1[1], c.position1[2]}, q1);
            vec3 p1 = rel1 + x1;
            vec3 rel2 = rot(vec3{c.position2[0], c.position2[1], c.position2[2]}, q2);
            vec3 p2 = rel2 + x2;
            c.relative_position1[0] = rel1.x; c.relative_position1[1] = rel1.y; c.relative_position1[2] = rel1.z;
            c.relative_position2[0] = rel2.x; c.relative_position2[1] = rel2.y; c.relative_position2[2] = rel2.z;
            vec3 dist = p1 - p2;
            vec3 n{c.normal[0], c.normal[1], c.normal[2]};
            float pen = dot(dist, n);
            bool ok = true;
            if (pen < -0.02f * 0.1f) ok = false;
            else {
                dist = dist - pen * n;
                if (len_sq(dist) > 0.02f * 0.02f) ok = false;
            }
            if (!ok) usage &= ~(1u << s);
        }
        cd.usage_mask = usage;
        keep[i] = (usage & 15u) ? 1u : 0u;
    }
}

// Arbiters with intact contacts move to the other buffer; the keys of the others (brokenArbiters,
// World.Step.cs:862-878) go to zkey, in key order, for the next step's EndCollide test.
__global__ void k_compact_arbiters(const jh_arbiter* __restrict__ ain, const uint64_t* __restrict__ akey,
                                   const int* __restrict__ a_count, const uint32_t* __restrict__ keep,
                                   const uint32_t* __restrict__ offs, jh_arbiter* __restrict__ aout,
                                   uint64_t* __restrict__ okey, uint64_t* __restrict__ zkey,
                                   int* __restrict__ zcount) {
    int A = *a_count;
    if (A == 0 && blockIdx.x == 0 && threadIdx.x == 0) *zcount = 0;
    int sub = threadIdx.x & 15;
    for (int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 4; i < A; i += (gridDim.x * blockDim.x) >> 4) {
        // the 16 lanes of a group share i (and so keep[i])
        uint32_t o = offs[i];
        if (!keep[i]) {
            if (sub == 0) {
                zkey[i - o] = akey[i];
                if (i == A - 1) *zcount = A - int(o);
            }
            continue;
        }
        if (sub == 0) {
            okey[o] = akey[i];
            if (i == A - 1) *zcount = A - int(o + 1);
        }
        copy_arbiter16(&ain[i], &aout[o], sub);
    }
}

// Collision events of a step, with the reference's arbiter lifetime (World.Step.cs:839-911): an
// arbiter broken by the previous step's UpdateContacts (zkey) is freed only if this step's
// narrowphase did not refill it (EndCollide); a pair registered now with no arbiter from the previous
// step, broken ones included, creates one (BeginCollide).  nkey: pairs new to the cache (sorted), ckey:
// all registered pairs (sorted).
__global__ void k_events(const uint64_t* __restrict__ nkey, const int* __restrict__ n_new,
                         const uint64_t* __restrict__ zkey, const int* __restrict__ zcount,
                         const uint64_t* __restrict__ ckey, const int* __restrict__ c_count,
                         uint64_t* __restrict__ begin, int* __restrict__ n_begin, uint64_t* __restrict__ end,
                         int* __restrict__ n_end) {
    int N = *n_new, Z = *zcount, C = *c_count;
    int t0 = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    for (int i = t0; i < N; i += stride) {
        uint64_t k = nkey[i];
        int f = lower_bound_u64(zkey, Z, k);
        if (!(f < Z && zkey[f] == k)) begin[atomicAdd(n_begin, 1)] = k;
    }
    for (int i = t0; i < Z; i += stride) {
        uint64_t k = zkey[i];
        int f = lower_bound_u64(ckey, C, k);
        if (!(f < C && ckey[f] == k)) end[atomicAdd(n_end, 1)] = k;
    }
}

// RigidBody.Update (RigidBody.cs:716-764) for the awake bodies (ForeachActiveBody), no motion locks:
// the sleep timer (a kinematic body rests only with both velocities exactly zero) and the island mark
// for the next CheckDeactivation (bit 2 of sflags), then the dynamic bodies' damping, forces and
// gravity.
__global__ void k_update_bodies(Bodies B, int nb, StepTime ts, float gx, float gy, float gz) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    uint32_t fl = B.sflags[b];
    if (!(fl & 1u)) return;
    int mt = B.motion[b];
    int locks = B.locks[b];
    if (locks) {  // RigidBody.cs:719-722
        B.vel[b] = st3(restrict_linear(ld3(B.vel[b]), locks));
        B.angvel[b] = st3(restrict_angular(ld3(B.angvel[b]), locks));
    }
    {
        float4 sl = B.slp[b];
        vec3 v = ld3(B.vel[b]), w = ld3(B.angvel[b]);
        bool kin = mt == JH_MOTION_KINEMATIC;
        bool resting = kin ? (w.x == 0.0f && w.y == 0.0f && w.z == 0.0f && v.x == 0.0f && v.y == 0.0f && v.z == 0.0f)
                           : (len_sq(w) < sl.y && len_sq(v) < sl.x);
        float t = resting ? sl.w + ts.dt : 0.0f;
        B.slp[b].w = t;
        bool mark = (kin && !resting) || t < sl.z;
        B.sflags[b] = (fl & ~4u) | (mark ? 4u : 0u);
    }
    if (mt != JH_MOTION_DYNAMIC) return;
    update_world_inertia(B, b);
    float4 m = B.mat[b];
    B.angvel[b] = st3(ld3(B.angvel[b]) * B.mat2[b].x);
    B.vel[b] = st3(ld3(B.vel[b]) * m.w);
    vec3 dv = cmul(ld3(B.force[b]), ld3(B.imv[b])) * ts.substep_dt;
    vec3 dw = sym_mul(ld3(B.torque[b]), ld_iw(B, b)) * ts.substep_dt;
    if (B.gravity[b]) dv = dv + restrict_linear(vec3{gx, gy, gz}, locks) * ts.substep_dt;
    B.dvel[b] = st3(dv);
    B.dangvel[b] = st3(dw);
    B.force[b] = make_float4(0, 0, 0, 0);
    B.torque[b] = make_float4(0, 0, 0, 0);
}

}  // namespace jh

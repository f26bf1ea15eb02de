/*
 * jitterhip — C ABI of the MI355X (gfx950) implementation of Jitter2's per-step hot path:
 * broadphase pair finding -> MPR/EPA narrowphase with one-shot manifolds -> contact cache ->
 * substepped projected Gauss-Seidel (sequential impulse) solver -> integration.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and returns a jh_status
 * (0 = ok, negative = error; jh_last_error() gives a thread-local message).  No exceptions
 * cross this boundary.  A world handle is not thread-safe; distinct handles may be stepped
 * from distinct host threads (each handle owns one HIP stream).  Device memory is owned by the
 * handle and released by jh_world_destroy.  All arrays passed in are copied before return;
 * all arrays passed out are caller-allocated.
 *
 * Reference interfaces replaced (paths relative to the Jitter2 source tree):
 *   World.Step(dt, multiThread)              src/Jitter2/World.Step.cs:141-151, StepInternal :153-296
 *   World.CreateRigidBody / RigidBody.AddShape src/Jitter2/World.cs:943-956, Dynamics/RigidBody.cs:1163-1209
 *   RigidBodyData (body state)               src/Jitter2/Dynamics/RigidBody.cs:107-250
 *   ContactData / Contact (contact cache)    src/Jitter2/Dynamics/Contact.cs:30-118, 617-661
 *   World.RawData spans                      src/Jitter2/World.cs:95-141
 *   DynamicTree.EnumerateOverlaps (pairs)    src/Jitter2/Collision/DynamicTree/DynamicTree.cs:266-324
 *   ISupportMappable shapes                  src/Jitter2/Collision/NarrowPhase/ISupportMappable.cs:19-33
 */
#ifndef JITTERHIP_H
#define JITTERHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t jh_status;

enum {
    JH_OK = 0,
    JH_ERR_INVALID_ARG = -1,  /* ArgumentException in World.Step.cs:155-165 and property setters   */
    JH_ERR_NON_FINITE = -2,   /* ArgumentCheck.Finite(dt) in World.Step.cs:157                       */
    JH_ERR_CAPACITY = -3,     /* MaximumSizeException, Unmanaged/PartitionedBuffer.cs:134-139         */
    JH_ERR_HIP = -4,          /* a HIP runtime call failed                                            */
    JH_ERR_UNSUPPORTED = -5,  /* a feature outside the GPU path (motion locks, gyroscopic forces ...)  */
    JH_ERR_STATE = -6,        /* ObjectDisposedException / call order errors                          */
    JH_ERR_NO_DEVICE = -7,    /* no HIP device visible                                                */
    JH_ERR_TIMEOUT = -8       /* a persistent kernel's bounded wait gave up, or jh_world_cancel: no
                                 counterpart in the reference (its solver runs on host threads)       */
};

/* MotionType, Dynamics/RigidBody.cs:75-95 (bits 0-1 of RigidBodyData.Flags). */
enum { JH_MOTION_DYNAMIC = 0, JH_MOTION_KINEMATIC = 1, JH_MOTION_STATIC = 2 };

/* Shape kinds of the hot path (files under Collision/Shapes). */
enum {
    JH_SHAPE_BOX = 0,        /* BoxShape.cs: params[0..2] = half extents                                  */
    JH_SHAPE_SPHERE = 1,     /* SphereShape.cs: params[0] = radius                                         */
    JH_SHAPE_CAPSULE = 2,    /* CapsuleShape.cs: params[0] = radius, params[1] = half length              */
    JH_SHAPE_POINTCLOUD = 3, /* PointCloudShape.cs / VertexSupportMap.cs: params[0..2] = center (GetCenter),
                                params[3..5] = shift, params[6..8] = cached local box min,
                                params[9..11] = cached local box max; vertices[vertex_offset ..]           */
    JH_SHAPE_TRIANGLE = 4,   /* TriangleShape.cs (static meshes): vertices[vertex_offset .. +3] = A, B, C;
                              * params: triangle normal(3), normals of NeighborA, B, C (3 each, zero =
                              * none), body frame (TriangleMesh.cs:56-63, 300-319)                     */
    JH_SHAPE_CONVEXHULL = 5  /* ConvexHullShape.cs: params as POINTCLOUD (center = the hull's centre of
                              * mass, shift, cached local box); vertex_count = V distinct hull vertices
                              * (first-occurrence order of the triangle corners), vertices[vertex_offset
                              * .. +V]; the vertex array continues with the adjacency as small integers
                              * in floats: per vertex (first, end) into the neighbour list, then the
                              * neighbour list (each vertex's neighbours ascending), padded to whole
                              * vertices (3 floats).  SupportMap: the hill climb from vertex 0 with the
                              * plateau pass (:378-450).                                               */
};

/* ContactData.SolveMode bits, Dynamics/Contact.cs:36-58. */
enum { JH_SOLVE_LINEAR1 = 1, JH_SOLVE_ANGULAR1 = 2, JH_SOLVE_LINEAR2 = 4, JH_SOLVE_ANGULAR2 = 8 };

/* World-level settings; names and defaults follow World.cs:306-412 and World.Detect.cs:88-122. */
typedef struct jh_config {
    float gravity[3];                    /* World.Gravity, default (0, -9.81, 0)                 */
    int32_t substep_count;               /* World.SubstepCount, default 1                        */
    int32_t solver_iterations;           /* World.SolverIterations.solver, default 6             */
    int32_t relaxation_iterations;       /* World.SolverIterations.relaxation, default 4         */
    float speculative_relaxation_factor; /* World.SpeculativeRelaxationFactor, default 0.9       */
    int32_t enable_auxiliary_contact_points; /* World.EnableAuxiliaryContactPoints, default 1    */
    int32_t persistent_contact_manifold; /* World.PersistentContactManifold, default 1           */
    int32_t pair_capacity;               /* broadphase pair buffer size; 0 = 24 per shape         */
    int32_t warm_start;                  /* JH_WARM_START_*; no counterpart in the reference      */
    int32_t narrowphase_filter;          /* World.NarrowPhaseFilter: JH_FILTER_*, default none     */
    float speculative_velocity_threshold; /* World.SpeculativeVelocityThreshold, default 10       */
    int32_t allow_deactivation;          /* World.AllowDeactivation, default 1 (World.cs:333): islands
                                            whose bodies rest for their DeactivationTime sleep       */
    int32_t reserved[2];
} jh_config;

/* World.NarrowPhaseFilter (World.Detect.cs:197-203).  TRIANGLE_EDGE = TriangleEdgeCollisionFilter
 * (Collision/CollisionFilter/TriangleEdgeCollisionFilter.cs) with its defaults: EdgeThreshold 0.01,
 * AngleThreshold acos(0.999), ProjectionThreshold 0.5, FilterDynamicBodies false. */
enum { JH_FILTER_NONE = 0, JH_FILTER_TRIANGLE_EDGE = 1 };

/* Warm start of the contact impulses at the start of a step.  PREVIOUS is the reference's
 * (PrepareForIteration replays the previous frame's Accumulated, Contact.cs:802-817); LEARNED first
 * replaces Accumulated with the prediction of the model given to jh_world_set_warm_start_model. */
enum { JH_WARM_START_PREVIOUS = 0, JH_WARM_START_LEARNED = 1 };

/* One rigid body.  Position/orientation/velocities and the delta velocities are the solver state
 * of RigidBodyData (RigidBody.cs:107-250); the rest are the managed RigidBody fields that
 * RigidBody.Update (RigidBody.cs:716-764) reads.  InverseInertiaWorld and InverseMassVector are
 * derived on upload from orientation, inverse_inertia and inverse_mass (UpdateWorldInertia,
 * RigidBody.cs:903-923). */
typedef struct jh_body {
    float position[3];
    float orientation[4];            /* quaternion x, y, z, w                                      */
    float velocity[3];
    float angular_velocity[3];
    float delta_velocity[3];         /* RigidBodyData.DeltaVelocity, applied each substep          */
    float delta_angular_velocity[3];
    float inverse_inertia[6];        /* body frame M11 M12 M13 M22 M23 M33                         */
    float inverse_mass;
    float friction;                  /* RigidBody.Friction, default 0.2                            */
    float restitution;               /* RigidBody.Restitution, default 0                           */
    float linear_damping_multiplier; /* 1 - RigidBody.Damping.linear, default 0.998                */
    float angular_damping_multiplier;/* 1 - RigidBody.Damping.angular, default 0.995               */
    float force[3];                  /* RigidBody.Force, cleared by each step                      */
    float torque[3];
    int32_t motion_type;             /* JH_MOTION_*                                                 */
    int32_t affected_by_gravity;     /* RigidBody.AffectedByGravity, default 1                     */
    int32_t world_id;                /* independent sub-world (batched worlds); 0 for one world    */
    int32_t enable_speculative_contacts; /* RigidBody.EnableSpeculativeContacts, default 0:
                                        the body's broadphase box is swept over dt * |velocity|
                                        and pairs MPR finds separated are swept for a time of
                                        impact (World.Detect.cs:162-194)                          */
    /* RigidBody.DeactivationThreshold and DeactivationTime (RigidBody.cs:398-400, 476-507): the body
     * rests while |v|^2 < linear_sq and |w|^2 < angular_sq (the squares, as the reference stores them;
     * defaults 0.1, 0.1) and may sleep after resting for deactivation_time seconds (default 1). */
    float deactivation_threshold_linear_sq;
    float deactivation_threshold_angular_sq;
    float deactivation_time;
    /* The complement of RigidBody.AllowedMotion (RigidBody.cs:571-594, RigidBodyData.Flags bits 4..9):
     * JH_LOCK_* bits of the world axes the body may not move along or rotate about; 0 = free.
     * Locked components of the velocities are cleared when the record is written; the solver sees a
     * zero inverse mass along locked translations and the Schur complement of the inverse inertia for
     * locked rotations (RigidBody.UpdateWorldInertia RigidBody.cs:903-923).  A change of a body's locks
     * also clears its contact cache and the warm start of its constraints, as the setter does. */
    int32_t locked_axes;
} jh_body;

/* MotionAxes (RigidBody.cs:26-53) as lock bits of jh_body.locked_axes. */
enum {
    JH_LOCK_LINEAR_X = 1, JH_LOCK_LINEAR_Y = 2, JH_LOCK_LINEAR_Z = 4,
    JH_LOCK_ANGULAR_X = 8, JH_LOCK_ANGULAR_Y = 16, JH_LOCK_ANGULAR_Z = 32
};

/* One shape attached to a body.  shape_id is Shape.ShapeId (Collision/Shapes/Shape.cs:28); shapes
 * must be passed in strictly ascending shape_id order so that array order equals key order. */
typedef struct jh_shape {
    uint64_t shape_id;
    int32_t body;            /* index into the body array                                          */
    int32_t type;            /* JH_SHAPE_*                                                          */
    float params[12];
    int32_t vertex_offset;   /* point clouds / triangles: first vertex in the vertex array         */
    int32_t vertex_count;
} jh_shape;

/* One contact slot, field for field the Contact struct (Contact.cs:617-661). */
typedef struct jh_contact {
    uint32_t flag;                 /* Contact.Flags: 2 = NewContact, 4 = SkipWarmStart   */
    float bias;
    float penalty_bias;
    float normal[3];
    float tangent1[3];
    float tangent2[3];
    float mass_normal_tangent[3];
    float accumulated[3];          /* normal, tangent1, tangent2 accumulated impulse    */
    float position1[3];            /* body-1 local anchor                               */
    float position2[3];            /* body-2 local anchor                               */
    float relative_position1[3];   /* world offset from body 1                          */
    float relative_position2[3];   /* world offset from body 2                          */
} jh_contact;

/* One arbiter (ContactData, Contact.cs:30-118) keyed by the ordered shape pair (Arbiter.cs:67-103). */
typedef struct jh_arbiter {
    uint32_t shape1;       /* index of the shape with the smaller shape_id (ArbiterKey.Key1) */
    uint32_t shape2;       /* ArbiterKey.Key2                                                */
    int32_t body1;
    int32_t body2;
    uint32_t usage_mask;   /* low nibble = intact slots                                      */
    uint32_t mode;         /* JH_SOLVE_* bits                                                */
    float friction;        /* max of the two bodies (Contact.cs:276)                         */
    float restitution;     /* max of the two bodies (Contact.cs:277)                         */
    float speculative_relaxation_factor;
    uint32_t reserved;     /* solver colour kept between steps (internal; ignored on upload) */
    jh_contact contacts[4];
} jh_arbiter;

/* Constraints solved with the contacts (World.Step.cs:1155-1191: each substep prepares contacts, then
 * constraints; each iteration solves contacts, then constraints; relaxation solves contacts only).
 * params / state per kind, the Data fields of the reference's constraint (set by its Initialize and
 * properties; state carries over between steps):
 *   BALL_SOCKET (Dynamics/Constraints/BallSocket.cs): params LocalAnchor1[3], LocalAnchor2[3] (body
 *       frames), BiasFactor, Softness; state AccumulatedImpulse[3].
 *   HINGE_ANGLE (HingeAngle.cs): params MinAngle, MaxAngle, FixedAngle (half angles as Initialize stores
 *       them), BiasFactor, LimitBias, LimitSoftness, Softness, Axis[3] (body-2 frame), Q0[4] (x y z w);
 *       state AccumulatedImpulse[3], Hemisphere (0, +1, -1).
 *   FIXED_ANGLE (FixedAngle.cs): params BiasFactor, Softness, Q0[4]; state AccumulatedImpulse[3],
 *       Hemisphere.
 *   DISTANCE_LIMIT (DistanceLimit.cs): params LocalAnchor1[3], LocalAnchor2[3], BiasFactor, Softness,
 *       Distance, LimitMin, LimitMax; state AccumulatedImpulse.
 *   TWIST_ANGLE (TwistAngle.cs): params B[3] (body-2 frame), Q0[4], Angle1, Angle2, FixedAngle, BiasFactor,
 *       Softness; state AccumulatedImpulse, Hemisphere.
 *   CONE_LIMIT (ConeLimit.cs): params LocalAxis1[3], LocalAxis2[3], BiasFactor, Softness, MinAngle,
 *       MaxAngle; state AccumulatedImpulse, the last Jacobian J0[3].
 *   ANGULAR_MOTOR (AngularMotor.cs), LINEAR_MOTOR (LinearMotor.cs): params LocalAxis1[3], LocalAxis2[3],
 *       Velocity (TargetVelocity), MaxForce (MaximumForce / MaximumTorque); state AccumulatedImpulse.
 *   POINT_ON_LINE (PointOnLine.cs): params LocalAxis[3] (body-1 frame), LocalAnchor1[3], LocalAnchor2[3],
 *       BiasFactor, LimitBias, Softness, LimitSoftness, Min, Max; state AccumulatedImpulse[3].
 *   POINT_ON_PLANE (PointOnPlane.cs): params LocalAxis[3] (body-1 frame), LocalAnchor1[3], LocalAnchor2[3],
 *       BiasFactor, Softness, Min, Max; state AccumulatedImpulse.
 *   ANGULAR_SERVO (AngularServo.cs): params TargetRelativeOrientation[4] (x y z w), MaximumTorque,
 *       Softness, BiasFactor; state AccumulatedImpulse[3], Hemisphere.
 *   LINEAR_SERVO (LinearServo.cs): params LocalAnchor1[3], LocalAnchor2[3], TargetPosition[3] (body-1
 *       frame), MaximumForce, Softness, BiasFactor; state AccumulatedImpulse[3].
 *   LINEAR_AXIS_SERVO (LinearAxisServo.cs): params LocalAxis[3] (body-1 frame), LocalAnchor1[3],
 *       LocalAnchor2[3], TargetPosition, MaximumForce, Softness, BiasFactor; state AccumulatedImpulse.
 *   TWIST_SERVO (TwistServo.cs): params LocalAxis2[3] (body-2 frame), ReferenceOrientation[4],
 *       TargetAngle, MaximumTorque, Softness, BiasFactor; state AccumulatedImpulse, Hemisphere.
 *   ELLIPTICAL_CONE_LIMIT (EllipticalConeLimit.cs): params LocalAxis1[3], LocalAxis2[3], LocalReference1[3]
 *       (body-1 frame, perpendicular to LocalAxis1), MinAngle1, MaxAngle1 (semiaxes along the reference),
 *       MinAngle2, MaxAngle2 (perpendicular), Softness, BiasFactor; state AccumulatedImpulse, Jacobian[3],
 *       Clamp (0 inside, 1 outer, 2 inner, 3 fixed).
 * A servo's MaximumForce / MaximumTorque may be +inf (unlimited); it must not be negative or NaN. */
enum { JH_CONSTRAINT_BALL_SOCKET = 0, JH_CONSTRAINT_HINGE_ANGLE = 1, JH_CONSTRAINT_FIXED_ANGLE = 2,
       JH_CONSTRAINT_DISTANCE_LIMIT = 3, JH_CONSTRAINT_TWIST_ANGLE = 4, JH_CONSTRAINT_CONE_LIMIT = 5,
       JH_CONSTRAINT_ANGULAR_MOTOR = 6, JH_CONSTRAINT_LINEAR_MOTOR = 7, JH_CONSTRAINT_POINT_ON_LINE = 8,
       JH_CONSTRAINT_POINT_ON_PLANE = 9, JH_CONSTRAINT_ANGULAR_SERVO = 10, JH_CONSTRAINT_LINEAR_SERVO = 11,
       JH_CONSTRAINT_LINEAR_AXIS_SERVO = 12, JH_CONSTRAINT_TWIST_SERVO = 13, JH_CONSTRAINT_ELLIPTICAL_CONE_LIMIT = 14 };

typedef struct jh_constraint {
    int32_t type;           /* JH_CONSTRAINT_*                                                  */
    int32_t body1, body2;   /* body indices (Constraint.Body1 / Body2)                           */
    int32_t enabled;        /* Constraint.IsEnabled                                              */
    float params[16];
    float state[8];
} jh_constraint;

/* Narrowphase result of one colliding broadphase pair in the last step (World.Detect.cs:165-217). */
typedef struct jh_manifold {
    uint32_t shape1, shape2;
    int32_t count;          /* manifold points (CollisionManifold.Count), 1..4             */
    int32_t flags;          /* bit 0: EPA refined the MPR answer; bit 1: speculative contact
                               (NarrowPhase.Sweep, one point, SolveMode.Angular removed)      */
    float normal[3];
    float penetration;
    float point_a[3];       /* MPR/EPA deepest point on shape 1                            */
    float point_b[3];
    float manifold_a[4][3];
    float manifold_b[4][3];
} jh_manifold;

/* Per-step counters and stage timings (World.DebugTimings buckets, World.Step.cs:44-81). */
typedef struct jh_stats {
    int64_t bodies, shapes, pairs, colliding_pairs, arbiters, contacts, colours, epa_calls;
    int64_t steps;
    int64_t pair_overflow;
    int64_t arbiters_solved;  /* arbiters in the last solve (before contact refresh removed broken ones) */
    double ms_broadphase, ms_narrowphase, ms_contact_cache, ms_colouring, ms_solve, ms_update;
    double ms_step;
    double ms_epa;            /* part of ms_narrowphase spent in the EPA kernel */
    int64_t solver;           /* solver kernel of the last step: 0 colour batches with grid barriers,
                                 1 dataflow (wave chunks), 2 world-local, 3 tiles, 4 dataflow (lanes) */
    int64_t max_degree;       /* most solved arbiters on one non-static body (a lower bound of the
                                 colour count); dataflow and tile solvers only, else 0 */
    double ms_solve_kernel;   /* part of ms_solve spent in the solver kernel itself */
    int64_t bodies_asleep;    /* non-static bodies asleep after the last CheckDeactivation */
} jh_stats;

typedef struct jh_world jh_world;

/* Error text of the last failing call on this thread (never NULL). */
const char* jh_last_error(void);
/* ABI version: 1. */
int32_t jh_version(void);
/* sizeof of the ABI structs, for binding checks: 0 jh_config, 1 jh_body, 2 jh_shape, 3 jh_contact,
 * 4 jh_arbiter, 5 jh_manifold, 6 jh_stats, 7 jh_probe, 8 jh_tuning, 9 jh_constraint; -1 otherwise. */
int32_t jh_struct_size(int32_t which);

/* World lifetime (World ctor World.cs:423-441, Dispose World.cs:977-990).  config may be NULL for
 * the reference defaults. */
jh_status jh_world_create(const jh_config* config, jh_world** out_world);
jh_status jh_world_destroy(jh_world* world);
jh_status jh_world_set_config(jh_world* world, const jh_config* config);
jh_status jh_config_defaults(jh_config* out_config);

/* Scene upload/download.  set_bodies replaces all bodies; set_shapes replaces all shapes and their
 * vertex data (vertices = 3 floats each) and clears the contact cache when the shape set changes.
 * Sleeping: writing a body's record is the reference's property setters at once, so each body is
 * World.ActivateBodyNextStep(body, wakeUpStatic: true) as the Position setter does (RigidBody.cs:593:
 * its sleep timer restarts, a sleeping body wakes at the next step, a static one wakes the bodies it
 * touches); a new body count or a motion-type change rebuilds the islands with every body awake
 * (World.CreateRigidBody).  set_shapes, set_arbiters and set_constraints also rebuild the islands. */
jh_status jh_world_set_bodies(jh_world* world, const jh_body* bodies, int32_t count);
jh_status jh_world_get_bodies(jh_world* world, jh_body* bodies, int32_t count);
jh_status jh_world_set_shapes(jh_world* world, const jh_shape* shapes, int32_t count,
                              const float* vertices, int32_t vertex_count);
/* The records of the listed bodies only (RigidBody property setters on those bodies), with
 * set_bodies' rules for each listed body; indices distinct.  world_id and EnableSpeculativeContacts
 * must stay (JH_ERR_UNSUPPORTED: use jh_world_set_bodies). */
jh_status jh_world_update_bodies(jh_world* world, const int32_t* index, const jh_body* records, int32_t count);

/* Sleeping (Collision/IslandHelper.cs, World.CheckDeactivation World.Step.cs:1289-1371, RigidBody.Update
 * RigidBody.cs:716-743).  get_activation: RigidBody.IsActive (1 awake) and the sleep timer
 * (InternalSleepTime, seconds; infinite after DeactivateBodyNextStep) of bodies 0..count-1.
 * set_activation_state: RigidBody.SetActivationState (RigidBody.cs:1025-1041), effective at the next
 * step: active != 0 restarts the timer and wakes a sleeping non-static body's island; 0 sets the
 * timer to infinity, so the island sleeps once no body in it moves.  jh_stats.bodies_asleep counts
 * the sleeping non-static bodies. */
jh_status jh_world_get_activation(jh_world* world, uint8_t* active, float* sleep_time, int32_t count);
jh_status jh_world_set_activation_state(jh_world* world, int32_t body, int32_t active);

/* The reference's 128-byte RigidBodyData layout (FP32), for zero-copy exchange with World.RawData
 * (World.cs:95-141).  stride = bytes between records (>= 128). */
jh_status jh_world_get_rigidbodydata(jh_world* world, void* records, int32_t count, int32_t stride);
/* Writes of World.RawData back to the world: count = the body count, records in body order.  Takes
 * Position, Velocity, AngularVelocity, DeltaVelocity, DeltaAngularVelocity, Orientation,
 * InverseInertiaWorld and InverseMassVector as they are (the solver uses them until the end-of-step
 * RigidBody.Update recomputes the world inertia, as in the reference); _index and _lockFlag are
 * ignored.  Flags: the MotionType bits must equal the body's (JH_ERR_UNSUPPORTED otherwise: change
 * it with jh_world_set_bodies), the IsActive bit the body's activation (JH_ERR_UNSUPPORTED: use
 * jh_world_set_activation_state; a RawData write wakes nothing), the gyroscopic and motion-lock bits
 * (3..9) must be clear (JH_ERR_UNSUPPORTED); non-finite values give JH_ERR_NON_FINITE.  Nothing is
 * written on an error. */
jh_status jh_world_set_rigidbodydata(jh_world* world, const void* records, int32_t count, int32_t stride);

/* World.Step(dt): one full step (narrowphase over the pair set, contact cache, substepped
 * solve, contact refresh, body update, broadphase).  Returns after the step finished on the GPU.
 * JH_ERR_TIMEOUT: a grid barrier or dependency wait of a persistent kernel gave up (see
 * jh_world_set_spin_limit) or the step was cancelled; every kernel of the step has returned and the
 * world stays usable, but its device state is unspecified until bodies and arbiters are set again. */
jh_status jh_world_step(jh_world* world, float dt);

/* World.Stabilize(dt, solverIterations, relaxationIterations) (World.Step.cs:318-392): per substep,
 * SolveVelocities with `iterations` iterations and RelaxVelocities with `relaxations`, on the contacts
 * and constraints the world holds (from the last step or jh_world_set_arbiters / set_constraints):
 * no collision detection, no force or position integration, no body update; warm-starts from the
 * accumulated impulses.  Argument checks and errors as StabilizeInternal; dt below float.Epsilon is a
 * no-op.  Solved in colour batches (the colours of the current arbiters); jh_world_get_solve_order then
 * returns this solve's order.  Starts with CheckDeactivation, as the reference does; only the contacts
 * and constraints with an awake dynamic body are solved. */
jh_status jh_world_stabilize(jh_world* world, float dt, int32_t iterations, int32_t relaxations);

/* Bounded device waits (no counterpart in the reference, whose solver runs on host threads; for callers
 * that must never hang the GPU).  Each wait of a persistent kernel gives up after `polls` polls
 * (0: the default, 2^23, several seconds) and ends the launch; the step then returns JH_ERR_TIMEOUT.
 * jh_world_cancel raises the device abort flag, so the next step's persistent kernels end at their
 * first wait and it returns JH_ERR_TIMEOUT (the flag is cleared when the step reports it). */
jh_status jh_world_set_spin_limit(jh_world* world, uint32_t polls);
jh_status jh_world_cancel(jh_world* world);

/* Broadphase only, on the current state: the pair set {(Key1, Key2)} that
 * DynamicTree.EnumerateOverlaps + World.Detect gating would visit, sorted by key.
 * pairs holds 2 * capacity uint32 (shape indices).  *out_count is the full count even when it
 * exceeds capacity (then JH_ERR_CAPACITY is returned). */
jh_status jh_world_compute_pairs(jh_world* world, uint32_t* pairs, int32_t capacity, int32_t* out_count);

/* Narrowphase results of the last step, sorted by key. */
jh_status jh_world_get_manifolds(jh_world* world, jh_manifold* manifolds, int32_t capacity, int32_t* out_count);

/* Collision events of the last step: RigidBody.BeginCollide / EndCollide (Dynamics/RigidBody.cs:330-350),
 * raised for both bodies of an arbiter at World.Step.cs:881-911 (HandleDeferredArbiters) and :839-860
 * (RemoveBrokenArbiters).  begin = arbiters the narrowphase created in this step; end = arbiters freed in
 * this step: left without intact contacts by the previous step's UpdateContacts and not refilled by this
 * step's narrowphase.  An arbiter that breaks and is refilled in the next step raises neither, as in the
 * reference.  Pairs are (Key1, Key2) shape indices (2 uint32 each), each list in key order (the
 * reference raises them in narrowphase order); the counts are the full list lengths, and
 * JH_ERR_CAPACITY is returned when a capacity is short. */
jh_status jh_world_get_collision_events(jh_world* world, uint32_t* begin_pairs, int32_t begin_capacity,
                                        int32_t* out_begin_count, uint32_t* end_pairs, int32_t end_capacity,
                                        int32_t* out_end_count);

/* Constraints (World.CreateConstraint, World.cs:850-888, and Constraint.Initialize): set replaces all
 * constraints (state included); get returns them with their current state.  With constraints the
 * step solves with colour batches and grid barriers (jh_stats.solver 0). */
jh_status jh_world_set_constraints(jh_world* world, const jh_constraint* constraints, int32_t count);
jh_status jh_world_get_constraints(jh_world* world, jh_constraint* constraints, int32_t capacity,
                                   int32_t* out_count);
/* Constraint solve order: constraint indices in the order the solver visits them, in batches of
 * constraints with disjoint dynamic bodies (batch_offsets: batch_count + 1 entries); replaying the
 * indices sequentially after the contacts of each pass reproduces the GPU's order. */
jh_status jh_world_get_constraint_order(jh_world* world, int32_t* order, int32_t capacity, int32_t* batch_offsets,
                                        int32_t batch_capacity, int32_t* out_count, int32_t* out_batch_count);

/* Contact cache (arbiters sorted by key) after the last step, and restore (checkpoint/resume).  Arbiters
 * left without intact contacts are dropped at the end of the step that broke them (the reference frees
 * them in the next step; only the events above keep its timing), so they are never listed. */
jh_status jh_world_get_arbiters(jh_world* world, jh_arbiter* arbiters, int32_t capacity, int32_t* out_count);
jh_status jh_world_set_arbiters(jh_world* world, const jh_arbiter* arbiters, int32_t count);

/* Solve order of the last step: arbiter keys (2 uint32 each) in the order the solver visited them,
 * grouped into batches of body-disjoint arbiters (graph colours); batch_offsets has batch_count+1
 * entries.  Replaying the keys sequentially on a CPU reproduces the GPU's Gauss-Seidel order. */
jh_status jh_world_get_solve_order(jh_world* world, uint32_t* keys, int32_t key_capacity,
                                   int32_t* batch_offsets, int32_t batch_capacity,
                                   int32_t* out_key_count, int32_t* out_batch_count);

jh_status jh_world_get_stats(jh_world* world, jh_stats* stats);

/* ---- Learned warm start and residual probe ---------------------------------------------------
 * No counterpart in the reference (it runs fixed iteration counts, World.cs:346-368, and warm-starts
 * from the previous frame).  A per-contact-point MLP 24 -> 128 -> 128 -> 3 (GELU; BF16 weights and
 * activations, FP32 accumulation on MFMA) predicts the accumulated impulse the solver converges to.
 *
 * Features (24 floats per contact point, at the warm start of the first substep, velocities after
 * IntegrateForces): normal(3), tangent1(3), tangent2(3), relative contact velocity along normal,
 * tangent1, tangent2 (3), penetration(1), MassNormalTangent(3), previous Accumulated(3), substep
 * index(1), friction(1), Bias(1), PenaltyBias(1), NewContact(1).
 * Network output y (3 floats): predicted impulse = previous Accumulated + clamp(y, +-3) * target_scale *
 * s, with the per-contact scale s = |MassNormal| (|v.n| + |v.t1| + |v.t2| + |PenaltyBias| + 0.05).
 * Parameters (JH_LWS_PARAMS floats): W1[128][24], b1[128], W2[128][128], b2[128], W3[3][128], b3[3]. */
#define JH_LWS_FEATURES 24
#define JH_LWS_PARAMS 20099
#define JH_PROBE_MAX_ITERATIONS 256

typedef struct jh_lws jh_lws;

/* Model with seeded initial weights (uniform +-1/sqrt(fan_in), SplitMix64 stream of `seed`),
 * zero mean / unit scale normalisation. */
jh_status jh_lws_create(uint64_t seed, jh_lws** out_model);
jh_status jh_lws_destroy(jh_lws* model);
jh_status jh_lws_set_params(jh_lws* model, const float* params, int32_t count);
jh_status jh_lws_get_params(jh_lws* model, float* params, int32_t count);
/* mean[24], inv_std[24]: feature normalisation (f - mean) * inv_std; target_scale[3] >= 0 (0 turns
 * the correction off: the prediction is then the previous-frame impulse). */
jh_status jh_lws_set_normalisation(jh_lws* model, const float* mean, const float* inv_std, const float* target_scale);
/* Training data made resident on the device: features [n][24] and targets [n][3] (the accumulated
 * impulse after many plain iterations); normalised with the current normalisation. */
jh_status jh_lws_set_data(jh_lws* model, const float* features, const float* targets, int64_t count);
/* Adam (beta 0.9 / 0.999, eps 1e-8) on the mean squared error of the normalised outputs over the
 * resident samples [begin, begin + count), in minibatches of `batch` taken in order; epoch_loss
 * (may be NULL) receives the mean loss of each epoch. */
jh_status jh_lws_train(jh_lws* model, int64_t begin, int64_t count, int32_t epochs, int32_t batch, float lr,
                       double* epoch_loss);
/* Mean squared error of the normalised outputs over resident samples [begin, begin + count). */
jh_status jh_lws_evaluate(jh_lws* model, int64_t begin, int64_t count, double* mse);
/* Gradient of that loss with respect to every parameter (JH_LWS_PARAMS floats), no update. */
jh_status jh_lws_gradient(jh_lws* model, int64_t begin, int64_t count, float* grad, double* loss);
/* Network outputs y [n][3] for raw features [n][24]. */
jh_status jh_lws_predict(jh_lws* model, const float* features, int64_t count, float* outputs);

/* jh_lws_predict without any allocation.  No counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  features [count][24] and outputs [count][3] are device buffers on the model's GPU; the
 * normalised BF16 inputs are staged in one fixed device array per GPU of JH_LWS_PREDICT_MAX_ROWS rows
 * (4 MiB, held until the process ends), which every call on that GPU shares: calls that overlap on one
 * GPU race.  count > JH_LWS_PREDICT_MAX_ROWS returns JH_ERR_CAPACITY; a pointer that is not device
 * memory of the model's GPU returns JH_ERR_INVALID_ARG.  Returns after the outputs are written. */
#define JH_LWS_PREDICT_MAX_ROWS 65536
jh_status jh_lws_predict_device(jh_lws* model, const float* features, int64_t count, float* outputs);

/* Copies the model (weights and normalisation) into the world for JH_WARM_START_LEARNED. */
jh_status jh_world_set_warm_start_model(jh_world* world, jh_lws* model);

/* Residual probe: iters-to-residual of the first substep of this step, per warm start, on identical
 * states.  Per contact point, after PGS iteration k: e_n = bias - v_n if lambda_n > 0 else
 * max(0, bias - v_n); e_t = |v_t| if |lambda_t| < friction * lambda_n(old) else 0 (both tangents),
 * lambda_n(old) = the normal impulse before iteration k (the one its friction clamp used).
 * residual[m][k] = sum of squares over all points after iteration k (k = 0: after the warm start),
 * for m = 0 previous-frame warm start, 1 learned, 2 cold (zero impulses); R_k = sqrt(residual[m][k] /
 * residual[2][0]).  Each probe runs on a scratch copy; the step itself then runs as jh_world_step.
 * With capacity > 0, the features of every contact point and its accumulated impulse after
 * `iterations` previous-frame-warm-started iterations are written (training traces); rows are the used
 * contact slots in arbiter key order. */
typedef struct jh_probe {
    int32_t iterations;      /* K, 1 .. JH_PROBE_MAX_ITERATIONS                             */
    int32_t modes;           /* bit m set: run probe m (m as above)                         */
    int32_t capacity;        /* rows available in features / targets                        */
    int32_t reserved;
    float* features;         /* [capacity][24]                                              */
    float* targets;          /* [capacity][3]                                               */
    float* learned;          /* [capacity][3] or NULL: with probe 1, the impulse each row is
                                warm-started with by the network (the in-step inference)      */
    int64_t points;          /* out: contact points                                          */
    int64_t rows;            /* out: rows written (min(points, capacity))                   */
    double residual[3][JH_PROBE_MAX_ITERATIONS + 1];  /* out                                 */
} jh_probe;

jh_status jh_world_step_probe(jh_world* world, float dt, jh_probe* probe);

/* ---- Launch-shape autotuner (no counterpart in the reference) --------------------------------
 * Times candidate workgroup counts of the persistent solver (1/16 .. 2x the CU count, clamped to what
 * can be co-resident), with several worlds also the world-local solver (one workgroup per world), the
 * workgroup count of the colouring kernel (16 .. 512) and the pair traversal (wavefront at 64, 32 or
 * 16 leaves per wave, or per-thread)
 * on real steps, round-robin over the candidates for `rounds` rounds each, and keeps the fastest.  The
 * world is stepped rounds * candidates times (the simulation advances); results are deterministic
 * whatever the launch shape. */
typedef struct jh_tuning {
    int32_t solver;            /* chosen solver kernel (as jh_stats.solver; 4 = lanes)           */
    int32_t solve_blocks;      /* its workgroups (tiles for the tile solver)                      */
    int32_t colour_blocks;     /* chosen colouring workgroups          */
    int32_t candidates;
    int32_t steps;             /* steps taken                          */
    int32_t traversal;         /* chosen pair traversal: 0 per thread, else the wavefront traversal
                                  (k_traverse_wave) with this many leaves per wave (64, 32 or 16) */
    double ms_solve_best, ms_solve_default;    /* mean solver time: chosen / default shape      */
    double ms_colour_best, ms_colour_default;  /* mean colouring time: chosen / default shape   */
    double ms_broadphase_best, ms_broadphase_default;  /* mean broadphase time: chosen / wavefront */
} jh_tuning;

jh_status jh_world_autotune(jh_world* world, float dt, int32_t rounds, jh_tuning* out);

#ifdef __cplusplus
}
#endif

#endif /* JITTERHIP_H */

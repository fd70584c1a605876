/*
 * vfd.h - C ABI of the MI355X-native DFSPH step (HIP, gfx950).
 *
 * This is the drop-in boundary for VFD's DFSPH simulation component. The reference exposes it as a C++ class,
 * DFSPHSimulation (VFD/Source/Simulation/DFSPH/DFSPHSimulator.h:10-56), forwarding to DFSPHImplementation
 * (VFD/Source/Simulation/DFSPH/DFSPHImplementation.h:23-137). Every entry point below names the reference member it
 * replaces. The ABI uses plain pointers and sizes only; every call returns an int status (0 = VFD_OK) and stores a
 * message for vfd_last_error().
 *
 * Data conventions
 *   - vec3 arrays are tightly packed float[3*n] (x0,y0,z0,x1,...), in the caller's ORIGINAL particle order. The
 *     library keeps particles internally in Morton (cell) order and maps back on every read.
 *   - A vfd_sim holds one or more scenes ("batch"). The plain functions act on scene 0; vfd_batch_* take a scene index.
 *   - Threads: one thread at a time drives a vfd_sim (every call not listed below). While it runs a call - for
 *     instance vfd_simulate, or while vfd_simulate_async's worker runs - any other thread may poll with
 *     vfd_get_simulation_state, vfd_get_debug_info, vfd_get_description, vfd_get_info, vfd_get_particle_count,
 *     vfd_scene_count, vfd_get_frame_count and vfd_get_frame (the editor's pattern, ComponentPanel.cpp:446-705,
 *     Scene.cpp:323-399); a second driving call meanwhile returns VFD_ERROR_STATE instead of racing. Work is
 *     enqueued on the stream set with vfd_set_stream (default: a non-blocking stream the simulation creates on the
 *     current device).
 */
#ifndef VFD_H
#define VFD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VFD_ABI_VERSION 3

/* Status codes. */
enum {
    VFD_OK = 0,
    VFD_ERROR_INVALID_ARGUMENT = 1,
    VFD_ERROR_HIP = 2,
    VFD_ERROR_STATE = 3,
    VFD_ERROR_OUT_OF_MEMORY = 4,
    VFD_ERROR_NOT_SUPPORTED = 5,
    VFD_ERROR_NONFINITE = 6
};

/* Mirrors DFSPHSimulationDescription (VFD/Source/Simulation/DFSPH/Structures/DFSPHSimulationDescription.h:9-50),
 * all 26 fields, same names and meaning. Booleans are int32 (0/1). Tolerances are percentages, as in the reference. */
typedef struct vfd_desc {
    /* Time step */
    float TimeStepSize;
    float MinTimeStepSize;
    float MaxTimeStepSize;
    float FrameLength;
    uint32_t FrameCount;
    /* Pressure solver */
    uint32_t MinPressureSolverIterations;
    uint32_t MaxPressureSolverIterations;
    float MaxPressureSolverError;
    /* Divergence solver */
    int32_t EnableDivergenceSolverError;
    uint32_t MinDivergenceSolverIterations;
    uint32_t MaxDivergenceSolverIterations;
    float MaxDivergenceSolverError;
    /* Viscosity solver */
    int32_t EnableViscositySolver;
    uint32_t MinViscositySolverIterations;
    uint32_t MaxViscositySolverIterations;
    float MaxViscositySolverError;
    float Viscosity;
    float BoundaryViscosity;
    float TangentialDistanceFactor;
    /* Surface tension solver */
    int32_t EnableSurfaceTensionSolver;
    uint32_t SurfaceTensionSmoothPassCount;
    float SurfaceTension;
    int32_t TemporalSmoothing;
    int32_t CSDFix;
    int32_t CSD;
    /* Scene */
    float ParticleRadius;
    float Gravity[3];
} vfd_desc;

/* Mirrors DFSPHSimulationInfo (VFD/Source/Simulation/DFSPH/Structures/DFSPHSimulationInfo.h:5-44): the constants
 * derived by SetDescription (DFSPHImplementation.cu:318-368) and ComputeTimeStepSize (:395-428). */
typedef struct vfd_info {
    uint32_t ParticleCount;
    uint32_t RigidBodyCount;
    float SupportRadius;
    float SupportRadius2;
    float ParticleRadius;
    float ParticleDiameter;
    float TimeStepSize;
    float TimeStepSize2;
    float TimeStepSizeInverse;
    float TimeStepSize2Inverse;
    float Volume;
    float Density0;
    float ParticleMass;
    float ParticleMassInverse;
    float Viscosity;
    float BoundaryViscosity;
    float DynamicViscosity;
    float DynamicBoundaryViscosity;
    float TangentialDistanceFactor;
    float TangentialDistance;
    float SurfaceTension;
    uint32_t SurfaceTensionSampleCount;
    float ClassifierSlope;
    float ClassifierConstant;
    int32_t TemporalSmoothing;
    float SmoothingFactor;
    float Factor;
    float NeighborParticleRadius;
    float MonteCarloFactor;
    float Gravity[3];
} vfd_info;

/* Mirrors DFSPHDebugInfo (VFD/Source/Simulation/DFSPH/Structures/DFSPHDebugInfo.h:8-28). Phase timers are wall
 * milliseconds of the last step measured with HIP events (the reference uses steady_clock around synchronised
 * launches). The last four members are additions. */
typedef struct vfd_debug_info {
    uint32_t IterationCount;
    uint32_t DivergenceSolverIterationCount;
    uint32_t PressureSolverIterationCount;
    uint32_t ViscositySolverIterationCount;
    float DivergenceSolverError;
    float PressureSolverError;
    float ViscositySolverError;
    float FrameTime;
    uint32_t FrameIndex;
    float NeighborhoodSearchTime;
    float BaseSolverTime;
    float DivergenceSolverTime;
    float SurfaceTensionSolverTime;
    float ViscositySolverTime;
    float PressureSolverTime;
    /* additions */
    float TimeStepSize;
    float MaxVelocityMagnitude;
    uint32_t MaxNeighborCount;
    uint64_t TotalNeighborCount;
    uint32_t StagedGroupCount;   /* neighbour groups of the whole simulation whose halo was staged in LDS */
    uint32_t GatheredGroupCount; /* groups whose halo exceeded the limit (neighbour data gathered from HBM) */
    uint32_t MaxHaloSize;        /* distinct neighbours of the largest staged group */
} vfd_debug_info;

/* Boundary shapes. The reference builds its volume map from a triangle mesh (RigidBodyDescription,
 * VFD/Source/Simulation/DFSPH/RigidBody/RigidBody.cuh:12-21; mesh distance in Utility/SDF/MeshDistance.cpp); this
 * build takes either a mesh (vfd_add_boundary_mesh) or the exact signed distance of an analytic shape
 * (vfd_add_boundary). d(x) < 0 inside the shape. */
enum {
    VFD_SHAPE_BOX = 0,      /* Center, HalfExtents */
    VFD_SHAPE_SPHERE = 1,   /* Center, Radius */
    VFD_SHAPE_PLANE = 2,    /* Center = a point on the plane, Normal (unit, pointing out of the solid); needs a domain */
    VFD_SHAPE_CYLINDER = 3  /* Center, Radius, HalfExtents[1] = half height; axis along y */
};

/* Mirrors RigidBodyDescription (RigidBody.cuh:12-21): Inverted, Padding, CollisionMapResolution, with the mesh and
 * transform replaced by an analytic shape. The volume-map domain is the shape's bounds grown by 8*SupportRadius +
 * (Padding - ParticleRadius), as in RigidBody.cu:35-37, unless HasDomain is set. */
typedef struct vfd_boundary_desc {
    int32_t Shape;
    float Center[3];
    float HalfExtents[3];
    float Radius;
    float Normal[3];
    int32_t Inverted;
    float Padding;
    uint32_t CollisionMapResolution[3];
    int32_t HasDomain;
    float DomainMin[3];
    float DomainMax[3];
} vfd_boundary_desc;

/* A triangle mesh as TriangleMesh::GetVertices / GetTriangles hold it (Renderer/Mesh/TriangleMesh.cpp:39-128):
 * Vertices is float[3*VertexCount], Triangles is uint32[3*TriangleCount] of zero-based vertex indices, wound
 * counter-clockwise seen from outside (the face normal cross(x1 - x0, x2 - x0) points out of the solid). */
typedef struct vfd_mesh {
    const float* Vertices;
    uint32_t VertexCount;
    const uint32_t* Triangles;
    uint32_t TriangleCount;
} vfd_mesh;

/* RigidBodyDescription (RigidBody.cuh:15-24) without its mesh pointer: Transform is the column-major 4x4 matrix
 * (glm::mat4, Transform[4*c + r] = row r of column c) applied to every vertex as Transform * (v, 1). */
typedef struct vfd_mesh_boundary_desc {
    float Transform[16];
    int32_t Inverted;
    float Padding;
    uint32_t CollisionMapResolution[3];
} vfd_mesh_boundary_desc;

/* SampleMode (Utility/Sampler/ParticleSampler.h:7-12). */
enum {
    VFD_SAMPLE_MIN_DENSITY = 0,
    VFD_SAMPLE_MEDIUM_DENSITY = 1,
    VFD_SAMPLE_MAX_DENSITY = 2
};

/* Per-particle fields readable with vfd_get_field (original order). Components per particle in brackets. The names
 * are the DFSPHParticle members (VFD/Source/Simulation/DFSPH/Structures/DFSPHParticle.h:8-32) plus solver vectors. */
enum {
    VFD_FIELD_POSITION = 0,                  /* [3] */
    VFD_FIELD_VELOCITY = 1,                  /* [3] */
    VFD_FIELD_ACCELERATION = 2,              /* [3] */
    VFD_FIELD_PRESSURE_ACCELERATION = 3,     /* [3] */
    VFD_FIELD_PRESSURE_RESIDUUM = 4,         /* [1] */
    VFD_FIELD_DENSITY = 5,                   /* [1] */
    VFD_FIELD_DENSITY_ADVECTION = 6,         /* [1] */
    VFD_FIELD_PRESSURE_RHO2 = 7,             /* [1] kappa of the density solve */
    VFD_FIELD_PRESSURE_RHO2_V = 8,           /* [1] kappa of the divergence solve */
    VFD_FIELD_FACTOR = 9,                    /* [1] */
    VFD_FIELD_VELOCITY_DIFFERENCE = 10,      /* [3] */
    VFD_FIELD_MC_SURFACE_NORMAL = 11,        /* [3] */
    VFD_FIELD_MC_SURFACE_NORMAL_SMOOTH = 12, /* [3] */
    VFD_FIELD_MC_SURFACE_CURVATURE = 13,     /* [1] */
    VFD_FIELD_MC_SURFACE_CURVATURE_SMOOTH = 14, /* [1] */
    VFD_FIELD_DELTA_FINAL_CURVATURE = 15,    /* [1] */
    VFD_FIELD_BOUNDARY_XJ = 16,              /* [3] per body */
    VFD_FIELD_BOUNDARY_VOLUME = 17,          /* [1] per body */
    VFD_FIELD_VISCOSITY_SOLUTION = 18,       /* [3] x of the viscosity solve (ViscosityGradientG) */
    VFD_FIELD_VISCOSITY_PRECONDITIONER = 19, /* [9] row-major inverse diagonal block */
    VFD_FIELD_COUNT = 20
};

/* Step phases, in the order of DFSPHImplementation::OnUpdate (DFSPHImplementation.cu:63-170). vfd_step_until runs
 * the step up to and including the named phase (a testing aid; the next vfd_step starts a fresh step). */
enum {
    VFD_PHASE_SEARCH = 0,
    VFD_PHASE_BOUNDARY = 1,
    VFD_PHASE_DENSITY = 2,
    VFD_PHASE_FACTOR = 3,
    VFD_PHASE_DIVERGENCE = 4,
    VFD_PHASE_CLEAR_ACCELERATION = 5,
    VFD_PHASE_SURFACE_TENSION = 6,
    VFD_PHASE_VISCOSITY = 7,
    VFD_PHASE_TIME_STEP = 8,
    VFD_PHASE_VELOCITY = 9,
    VFD_PHASE_PRESSURE = 10,
    VFD_PHASE_POSITION = 11,
    VFD_PHASE_COUNT = 12
};

typedef struct vfd_sim vfd_sim;

/* Library */
int vfd_abi_version(void);
const char* vfd_last_error(void);
/* Fills the reference defaults of DFSPHSimulationDescription.h:11-49. */
int vfd_desc_default(vfd_desc* desc);

/* Scene-description I/O, as VFD's scene files store a simulation component: serialize(Archive&,
 * DFSPHSimulationDescription&) and DFSPHSimulationComponent::save/load (Scene/Components/
 * DFSPHSimulationComponent.h:7-39, 57-70) write the 26 fields under cereal names ("timeStepSize", ...,
 * "gravity": {"x", "y", "z"}) inside a "description" object. vfd_read_descriptions parses a JSON text (e.g. a whole
 * scene file such as Resources/Scenes/New/DFSPH/default.json) and fills out[k] for the k-th "description" object
 * found, k < capacity, starting from the reference defaults for fields the object does not name; *count receives the
 * number of objects found (out may be NULL to count). vfd_write_description writes {"description": {...}} in the
 * same form (floats with 17 significant digits, so a read gives back the same values); buffer may be NULL to query
 * *length (the terminating NUL needs one more byte). */
int vfd_read_descriptions(const char* json, size_t length, vfd_desc* out, uint32_t capacity, uint32_t* count);
int vfd_write_description(const vfd_desc* desc, char* buffer, size_t capacity, size_t* length);

/* FluidObject ctor (FluidObject.cpp:6-24) with ParticleSampler::SampleMeshVolume (ParticleSampler.cpp:7-91): the
 * mesh's vertices are transformed by `transform` (column-major 4x4 as in vfd_mesh_boundary_desc; NULL = identity),
 * a one-field signed-distance map of the mesh (SDF.cu:16-35: bounding box grown by 0.001 of its diagonal, negated
 * when `inverted`) is built with `resolution` cells, and the lattice of `mode` (VFD_SAMPLE_*) over the bounding box
 * is kept where the interpolated distance is negative, in the reference's z, y, x loop order. Runs on the current
 * HIP device. out (float[3*capacity], may be NULL) receives the first min(capacity, *count) samples; *count receives
 * the number of samples. */
int vfd_sample_mesh_volume(const vfd_mesh* mesh, const float* transform, float radius, const uint32_t resolution[3],
                           int32_t inverted, int32_t mode, float* out, uint32_t capacity, uint32_t* count);

/* DFSPHSimulation(desc) (DFSPHSimulator.cpp:10-18) / DFSPHImplementation ctor (DFSPHImplementation.cu:14-21). */
int vfd_create(const vfd_desc* desc, vfd_sim** out);
/* Batch of scene_count independent scenes advanced by one set of launches. Every scene must use the same
 * ParticleRadius (one kernel table). No reference counterpart. */
int vfd_batch_create(const vfd_desc* descs, uint32_t scene_count, vfd_sim** out);
/* ~DFSPHImplementation (DFSPHImplementation.cu:23-31). */
int vfd_destroy(vfd_sim* sim);
/* hipStream_t passed as void*; NULL selects the simulation's own stream again. No reference counterpart (the
 * reference uses the default stream). Solver-loop graphs are captured on this stream, so it must not be the legacy
 * null stream for them to be used. */
int vfd_set_stream(vfd_sim* sim, void* stream);
int vfd_scene_count(const vfd_sim* sim, uint32_t* count);

/* SetDescription / GetDescription / GetInfo (DFSPHImplementation.cu:313-373). */
int vfd_set_description(vfd_sim* sim, const vfd_desc* desc);
int vfd_get_description(const vfd_sim* sim, vfd_desc* desc);
int vfd_get_info(const vfd_sim* sim, vfd_info* info);
int vfd_batch_set_description(vfd_sim* sim, uint32_t scene, const vfd_desc* desc);
int vfd_batch_get_info(const vfd_sim* sim, uint32_t scene, vfd_info* info);

/* SetFluidObjects (DFSPHImplementation.cu:172-253): positions of every fluid object concatenated, one velocity per
 * particle (the reference holds one per object, FluidObject.h:21-31; expand it per particle). A simulation holds at
 * most 2^28 particle slots (every scene's count rounded up to a multiple of 256); beyond that the call that allocates
 * the particle arrays returns VFD_ERROR_NOT_SUPPORTED. */
int vfd_set_fluid(vfd_sim* sim, const float* positions, const float* velocities, uint32_t count);
int vfd_batch_set_fluid(vfd_sim* sim, uint32_t scene, const float* positions, const float* velocities,
                        uint32_t count);

/* RigidBody ctor + SetRigidBodies (RigidBody.cu:10-73, DFSPHImplementation.cu:255-271): builds the two-field volume
 * map (signed distance, boundary volume) on the GPU. Call after vfd_set_fluid, as in the reference. */
int vfd_add_boundary(vfd_sim* sim, const vfd_boundary_desc* body);
int vfd_batch_add_boundary(vfd_sim* sim, uint32_t scene, const vfd_boundary_desc* body);
/* RigidBody ctor for a mesh body (RigidBody.cu:10-73): the vertices are transformed, the map domain is their bounding
 * box (which, as BoundingBox(vertices) starts from a zero box, always contains the origin) grown by 8*SupportRadius +
 * (Padding - ParticleRadius), and field 0 is sign * (signed mesh distance - (Padding - ParticleRadius)) with the
 * signed distance of MeshDistance::SignedDistance (MeshDistance.cpp:186-225: closest triangle, sign from the
 * angle-weighted vertex, edge or face normal of the closest feature). Both fields are built on the GPU. */
int vfd_add_boundary_mesh(vfd_sim* sim, const vfd_mesh_boundary_desc* body, const vfd_mesh* mesh);
int vfd_batch_add_boundary_mesh(vfd_sim* sim, uint32_t scene, const vfd_mesh_boundary_desc* body,
                                const vfd_mesh* mesh);
int vfd_clear_boundaries(vfd_sim* sim, uint32_t scene);
/* Number of map nodes and the node coefficients of one field (0 = distance, 1 = volume), node layout of
 * SDF::AddFunction (Utility/SDF/SDF.cu:45-139). nodes may be NULL to query the count. */
int vfd_get_boundary_map(vfd_sim* sim, uint32_t scene, uint32_t body, uint32_t field, float* nodes,
                         uint32_t* node_count);

/* DFSPHImplementation::OnUpdate (DFSPHImplementation.cu:63-170): one step of every scene. Returns
 * VFD_ERROR_NONFINITE (after completing the step; the state keeps its values) when a particle's new position or
 * velocity is not finite; the reference has no such check and lets NaN propagate. */
int vfd_step(vfd_sim* sim);
/* No counterpart in the reference. vfd_step for callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc): the same step with the same results, which calls no hipMalloc, hipFree,
 * hipHostMalloc or graph instantiation. The neighbour grid's cell arrays and the surface-tension curvature tables
 * come from a static scratch array of the library (32 MiB + 272 KiB on each device the library runs on, held until
 * the process ends); every other buffer is the simulation's own and must already exist: they are allocated by the
 * first vfd_step (or vfd_get_particles) after vfd_set_fluid, and reused for any smaller particle count later.
 * Limits: at most 4,194,304 grid cells in all scenes together (VFD_ERROR_NOT_SUPPORTED, checked before the step
 * changes any state), at most 65,536 curvature-table floats over the surface-tension scenes at the largest time step
 * the descriptions allow and 4,096 scenes (VFD_ERROR_NOT_SUPPORTED). Returns VFD_ERROR_STATE, without stepping, when
 * a buffer would have to be allocated or grown, or when frame capture (vfd_set_frame_capture), adjoint recording
 * (vfd_tape) or profiling is on. Iteration loops are launched directly, not through HIP graphs. The scratch is shared
 * by every no-alloc step on one device and holds the last call's contents: calls that overlap on one device (two
 * threads, two simulations) race; calls one after another are safe, since the step waits for its work before it
 * returns. */
int vfd_step_no_alloc(vfd_sim* sim);
/* Runs `steps` steps. */
int vfd_advance(vfd_sim* sim, uint32_t steps);
/* DFSPHImplementation::Simulate (DFSPHImplementation.cu:33-61): restores the baked particles, resets dt and runs
 * until FrameCount frames are captured, on the calling thread (vfd_simulate_async runs it on a worker thread, as the
 * reference's DFSPHSimulation::Simulate does). */
int vfd_simulate(vfd_sim* sim);
/* DFSPHImplementation::SimulationState (DFSPHImplementation.h:27-32): None after creation, vfd_set_fluid and the
 * boundary calls (SetFluidObjects :177, SetRigidBodies :259), Simulating while vfd_simulate runs, Ready once it has
 * captured FrameCount frames (Simulate :48, :58). A failed vfd_simulate leaves None. */
enum { VFD_STATE_NONE = 0, VFD_STATE_SIMULATING = 1, VFD_STATE_READY = 2 };
/* GetSimulationState (DFSPHImplementation.h:45, .cu:273-276). Any thread, any time. */
int vfd_get_simulation_state(const vfd_sim* sim, int32_t* state);
/* DFSPHSimulation::Simulate (DFSPHSimulator.cpp:156-166), which the editor calls from its UI thread: restores the
 * baked particles and releases the previous frames on the calling thread, then runs the frame loop of vfd_simulate on
 * a worker thread of the library (on the calling thread's HIP device) and returns at once. Poll it with the calls the
 * header lists as safe from other threads; vfd_wait joins the worker and returns vfd_simulate's status (with its
 * message in vfd_last_error). Until vfd_wait, driving calls return VFD_ERROR_STATE; vfd_destroy waits for the worker. */
int vfd_simulate_async(vfd_sim* sim);
int vfd_wait(vfd_sim* sim);
/* Testing aid: runs the current step up to and including `phase` (VFD_PHASE_*). */
int vfd_step_until(vfd_sim* sim, int32_t phase);
/* Enables/disables frame capture in vfd_step (default on, as in OnUpdate :148-167). A captured frame is packed in
 * original order on the GPU and copied into pinned host memory on a copy stream of its own, overlapping the
 * following steps (36 bytes per particle per frame stay pinned until vfd_simulate, vfd_set_fluid or vfd_destroy
 * releases the frames to a pool the next frames reuse). */
int vfd_set_frame_capture(vfd_sim* sim, int32_t enabled);

/* Reads and writes of particle state, original order. vfd_get_particles = ConvertParticlesToBuffer + copy
 * (DFSPHKernels.cu:6-22); any output pointer may be NULL. */
int vfd_get_particles(vfd_sim* sim, float* positions, float* velocities, float* accelerations);
int vfd_batch_get_particles(vfd_sim* sim, uint32_t scene, float* positions, float* velocities,
                            float* accelerations);
int vfd_get_field(vfd_sim* sim, uint32_t scene, int32_t field, uint32_t body, float* out);
/* Writes one persistent field (POSITION, VELOCITY, VELOCITY_DIFFERENCE, DELTA_FINAL_CURVATURE). */
int vfd_set_field(vfd_sim* sim, uint32_t scene, int32_t field, const float* values);
int vfd_get_particle_count(const vfd_sim* sim, uint32_t scene, uint32_t* count);

/* NeighborSet (ParticleSearch/NeighborSet.h:8-30) in original ids: counts[n], offsets[n] (exclusive scan) and
 * ids[total]. ids may be NULL to query total only. */
int vfd_get_neighbors(vfd_sim* sim, uint32_t scene, uint32_t* counts, uint32_t* offsets, uint32_t* ids,
                      uint64_t* total);

/* GetDebugInfo (DFSPHImplementation.cu:390-393), GetMaxVelocityMagnitude, GetCurrentTimeStepSize. */
int vfd_get_debug_info(vfd_sim* sim, uint32_t scene, vfd_debug_info* info);

/* DFSPHParticleBuffer (ParticleBuffer/DFSPHParticleBuffer.h:13-45): number of captured frames and one frame's
 * {x, v, a} (36 bytes per particle, DFSPHParticleSimple) with its MaxVelocityMagnitude and CurrentTimeStep.
 * vfd_get_frame waits for that frame's copy to complete. */
int vfd_get_frame_count(vfd_sim* sim, uint32_t scene, uint32_t* count);
int vfd_get_frame(vfd_sim* sim, uint32_t scene, uint32_t frame, float* xva, float* max_velocity, float* time_step);

/* Per-kernel timing with HIP events on the simulation's stream (no reference counterpart; the reference times
 * phases with steady_clock around synchronised launches, Debug/Timer.h:8-33). Enabling resets the statistics. */
typedef struct vfd_kernel_stat {
    char name[48];
    uint32_t launches;
    uint32_t pad;
    double total_ms;
} vfd_kernel_stat;
int vfd_set_profiling(vfd_sim* sim, int32_t enabled);
int vfd_get_profile(vfd_sim* sim, vfd_kernel_stat* out, uint32_t capacity, uint32_t* count);

/* Adjoint of the step (no reference counterpart). vfd_tape(sim, 1) starts recording the following steps (clearing
 * earlier records), vfd_tape(sim, 0) stops. vfd_adjoint maps cotangents of the last recorded step's positions and
 * velocities (lx_out, lv_out: 3n floats in original particle order; lv_out may be NULL for zero) to those of the first
 * recorded step's input (lx_in, lv_in) and d/dmu (mu = dynamic viscosity = Viscosity * rho0). Lists, time steps,
 * solver iteration counts and clamp masks are held at their recorded values; kernel derivatives are the spline's.
 * Boundary contact samples (with their boundary-viscosity samples) and push-outs are differentiated through the
 * volume map (their branches held as recorded); surface tension through its smoothing pass, with the Monte-Carlo
 * normals and curvatures held, and with temporal smoothing through the blended curvature carried from step to step
 * (the one carried into the first recorded step is held; the loss reads positions and velocities only). d/dmu is the
 * fluid viscosity's (BoundaryViscosity is held). Scope: no particle pushed out twice in a step or after sampling
 * another body, no particle at the neighbour cap (else VFD_ERROR_NOT_SUPPORTED, from vfd_adjoint). vfd_set_fluid
 * discards the records; their device buffers are kept for the next recording until vfd_destroy. In a batch every
 * scene is recorded; vfd_adjoint reverses scene 0 and vfd_batch_adjoint scene `scene` (its own time steps, solver
 * iteration counts, viscosity, bodies and surface tension; cotangents of that scene's particles only, its d/dmu). */
int vfd_tape(vfd_sim* sim, int32_t on);
int vfd_tape_steps(vfd_sim* sim, uint32_t* steps);
int vfd_adjoint(vfd_sim* sim, const float* lx_out, const float* lv_out, float* lx_in, float* lv_in, double* lmu);
int vfd_batch_adjoint(vfd_sim* sim, uint32_t scene, const float* lx_out, const float* lv_out, float* lx_in,
                      float* lv_in, double* lmu);

/* Launch and solver-control knobs; no counterpart in the reference (which synchronises after every iteration).
 * Results do not depend on them: a converged scene's extra iterations are skipped on the device, and a staged halo
 * hands the same values to the same arithmetic as a gather from HBM. */
typedef struct vfd_tuning {
    uint32_t SolverCheckInterval; /* Jacobi / CG iterations enqueued between two convergence checks, >= 1 */
    uint32_t HaloLimit;           /* largest group halo staged in LDS (<= 3808); 0: every neighbour from HBM */
    uint32_t UseGraphs;           /* 1: each chunk of iterations is one launch of a captured HIP graph */
    uint32_t PredictIterations;   /* 1: a solver's first chunk is one iteration short of what it ran in the
                                     previous step, so that few empty iterations follow its convergence (default 0) */
    uint32_t AdjointGroupLoops;   /* vfd_adjoint's transposed neighbour loops (the pressure solves' and the
                                     viscosity operator's): 0 one thread per particle gathering
                                     from HBM; 1 per neighbour group with the step's halo staged in LDS (the halos
                                     are rebuilt from the recorded lists); 2 (default) per group once there are
                                     as many groups of 1024 particles as compute units. Results are the same bits
                                     either way. */
} vfd_tuning;
int vfd_get_tuning(vfd_sim* sim, vfd_tuning* out);
int vfd_set_tuning(vfd_sim* sim, const vfd_tuning* tuning);

/* PrecomputedDFSPHCubicKernel (Kernel/DFSPHKernels.h:7-143): W[10000], GradientW[10001] and
 * {Radius, Radius2, InvStepSize, WZero, K, L}. */
int vfd_get_kernel_table(vfd_sim* sim, float* w, float* gradient_w, float* scalars);

#ifdef __cplusplus
}
#endif

#endif /* VFD_H */

/*
 * dsph.h — C ABI of the MI355X-native DEM–SPH per-timestep hot path.
 *
 * The reference solver (kaiqideng/DEM-SPHSource) has no FFI layer: its host
 * classes call five phase functions that take a `DeviceData&` and launch CUDA
 * kernels.  This header replaces that downward seam with plain C entry points
 * (opaque context, pointers + sizes, int status codes) so that the
 * solverBase/dataBase host API, a ctypes binding or any other FFI can drive
 * the HIP kernels.  Each entry point names the reference interface it
 * replaces (file:line in the reference tree).
 *
 * Conventions
 *  - every function returns DSPH_OK (0) or a negative DSPH_E* code; the text
 *    of the last error is returned by dsph_last_error(ctx);
 *  - one host thread per context; no global mutable state; the library owns
 *    every device buffer; host arrays passed in are copied during the call;
 *  - particle indices crossing this boundary are always the caller's
 *    ("original") indices: the order in which particles were uploaded, as in
 *    the reference's HostData vectors.  The device keeps particles in its
 *    own (cell-sorted) order internally;
 *  - vectors are interleaved xyz doubles (the reference's double3 AoS).
 */
#ifndef DSPH_H
#define DSPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
/* Scenes with a NaN/inf position, radius or smoothing length are rejected
 * at upload (DSPH_E_INVALID).                                              */
#define DSPH_OK             0
#define DSPH_E_INVALID    (-1)  /* bad argument / inconsistent scene       */
#define DSPH_E_HIP        (-2)  /* HIP runtime error                         */
#define DSPH_E_CAPACITY   (-3)  /* a neighbor/contact list overflowed        */
#define DSPH_E_STATE      (-4)  /* call out of order (e.g. step before init) */
#define DSPH_E_NOMEM      (-5)  /* device or host allocation failed          */
#define DSPH_E_NODEVICE   (-6)  /* no HIP device / bad device index          */
#define DSPH_E_NONFINITE  (-7)  /* a particle position became NaN or inf
                                   (DSPH_STATUS_NONFINITE)                   */
#define DSPH_E_IO         (-8)  /* a file could not be written               */

/* ---- precision of the fluid pair loops ------------------------------- */
#define DSPH_PREC_FP64   0  /* everything FP64 (the reference's arithmetic)   */
#define DSPH_PREC_MIXED  1  /* FP32 pair terms accumulated in FP64,
                               FP64 integration, springs and quaternions      */

/* ---- behaviour flags (reference quirks) ------------------------------ */
/* default 0 = the reference's own semantics                                 */
#define DSPH_FLAG_WALL_CONTINUE   (1u << 0) /* wall loop skips a non-touching
                                               wall instead of returning
                                               (integrate.cu:535-541)        */
#define DSPH_FLAG_CLUMP_QUATERNION (1u << 1)/* integrate clump orientation
                                               q <- normalize(q + 1/2 q*w dt)
                                               (myQua.h:46-55); the reference
                                               keeps q frozen                */
#define DSPH_FLAG_ART_VISCOSITY   (1u << 2) /* add the artificial-viscosity
                                               term of SPHKernelDEMContactModel
                                               .h:74-77 (never called
                                               upstream) with the particle's
                                               alpha (dsph_scene.fluid_alpha)*/

/* ---- list kinds for dsph_download_list -------------------------------- */
#define DSPH_LIST_FLUID_FLUID  0 /* fluid2Fluid forward half list            */
#define DSPH_LIST_FLUID_SOLID  1 /* fluid2Solid list                         */
#define DSPH_LIST_SOLID_SOLID  2 /* solid2Solid contact list                 */
#define DSPH_LIST_BOND         3 /* solidBond2Solid (pointed, pointing)      */

/* ---- status words (dsph_status) -------------------------------------- */
#define DSPH_STATUS_FF_OVERFLOW   0 /* max fluid-fluid neighbors seen > cap */
#define DSPH_STATUS_FS_OVERFLOW   1 /* fluid-solid pairs needed > capacity   */
#define DSPH_STATUS_SS_OVERFLOW   2 /* solid-solid pairs needed > capacity   */
#define DSPH_STATUS_OUT_OF_DOMAIN 3 /* particles hashed to the spare cell    */
#define DSPH_STATUS_FF_MAX        4 /* max fluid-fluid neighbors of a particle*/
#define DSPH_STATUS_FS_TOTAL      5 /* fluid-solid pairs found               */
#define DSPH_STATUS_SS_TOTAL      6 /* solid-solid contacts found            */
#define DSPH_STATUS_SUP_OVERFLOW  7 /* a fluid Verlet superset row overflowed */
#define DSPH_STATUS_REBUILDS      8 /* superset rebuilds since upload        */
#define DSPH_STATUS_SUP_MAX       9 /* longest superset row of the last build */
#define DSPH_STATUS_NONFINITE    10 /* fluid particles + movable solids with a
                                       NaN/inf position at the last search
                                       (such a position never recovers: the
                                       state stays poisoned until reloaded) */
#define DSPH_STATUS_WORDS        12

/*
 * Scene upload format: the reference's HostData SoA (hostStructs.h:7-399)
 * flattened to plain arrays.  NULL optional arrays mean "zeros" (or the
 * reference constructor default noted).  n_comb = n_material*(n_material+1)/2
 * and the contact tables are field-major (all E, then all G, ...).
 */
typedef struct dsph_scene {
    /* dataBase::setDomain / setGravity (dataBase.h:93-109) */
    double domain_origin[3];
    double domain_size[3];
    double gravity[3];

    /* HostFluid (hostStructs.h:49-77), filled by dataBase::addFluid */
    int32_t n_fluid;
    const double* fluid_pos;    /* [3n] */
    const double* fluid_vel;    /* [3n] */
    const double* fluid_h;      /* [n]  smoothing length (effectiveRadii)  */
    const double* fluid_rho0;   /* [n]                                      */
    const double* fluid_drho;   /* [n]  optional                            */
    const double* fluid_c;      /* [n]  speed of sound                      */
    const double* fluid_nu;     /* [n]  kinematic viscosity                 */

    /* HostSolid (hostStructs.h:79-118): walls, spheres, pebbles */
    int32_t n_solid;
    const double* solid_pos;        /* [3n] */
    const double* solid_vel;        /* [3n] optional */
    const double* solid_angvel;     /* [3n] optional */
    const double* solid_normal;     /* [3n] optional: kept when dsph_initialize
                                       is told not to compute normals */
    const double* solid_radius;     /* [n]  */
    const double* solid_eff_radius; /* [n]  search radius (2.6r wall, 1.1r DEM) */
    const double* solid_inv_mass;   /* [n]  0 = fixed */
    const int32_t* solid_material;  /* [n]  optional (0) */
    const int32_t* solid_cluster;   /* [n]  optional (-1) */
    const int32_t* solid_clump;     /* [n]  optional (-1) */
    const int32_t* solid_dummy;     /* [n]  optional (0) */

    /* HostClump (hostStructs.h:120-150) */
    int32_t n_clump;
    const double* clump_centroid;     /* [3n] */
    const double* clump_vel;          /* [3n] optional */
    const double* clump_angvel;       /* [3n] optional */
    const double* clump_orientation;  /* [4n] optional (1,0,0,0) */
    const double* clump_inv_inertia;  /* [6n] xx,yy,zz,xy,xz,yz (body frame) */
    const double* clump_inv_mass;     /* [n]  */
    const int32_t* clump_pebble_start;/* [n]  first solid index */
    const int32_t* clump_pebble_end;  /* [n]  one past the last */

    /* HostInfiniteWall (hostStructs.h:152-171) */
    int32_t n_wall;
    const double* wall_point;   /* [3n] */
    const double* wall_normal;  /* [3n] */
    const int32_t* wall_material;/* [n] optional (0) */

    /* HostSolidContactModel (hostStructs.h:339-375) */
    int32_t n_material;         /* reference default 10 (solverBase.cpp:16) */
    const double* hertz;        /* [8*n_comb]  E,G,res,kr/ks,kt/ks,mu_s,mu_r,mu_t */
    const double* linear;       /* [11*n_comb] kn,ks,kr,kt,dn,ds,dr,dt,mu_s,mu_r,mu_t */
    const double* bonded;       /* [6*n_comb]  gamma,E,kn/ks,sigma_s,C,mu */

    /* artificial viscosity coefficient per fluid particle, used with
     * DSPH_FLAG_ART_VISCOSITY; optional (0).  A pair uses the alpha of the
     * particle being updated: alpha is uniform within a scene.  No
     * reference counterpart (the reference's helper takes one alpha). */
    const double* fluid_alpha;  /* [n_fluid] */

    /* HostCylinderWall (hostStructs.h:176-200), filled by
     * dataBase::addCylinderWall: a point on the axis, the axis direction
     * (normalised on upload), the wall's angular velocity and radius.  The
     * reference never copies these walls to the device (dataBase.cpp:
     * 434-449) and never stores the radius (:69-82); here they act on the
     * solids as calCylinderWallForceTorque (integrate.cu:612-715) describes,
     * after the infinite walls.                                            */
    int32_t n_cylinder;
    const double* cylinder_point;    /* [3n] */
    const double* cylinder_axis;     /* [3n] */
    const double* cylinder_angvel;   /* [3n] optional (0) */
    const double* cylinder_radius;   /* [n]  */
    const int32_t* cylinder_material;/* [n]  optional (0) */
} dsph_scene;

typedef struct dsph_ctx dsph_ctx;

/* ---- lifecycle ---------------------------------------------------------
 * replaces: cudaSetDevice in solverBase::initialize (solverBase.cpp:19-26)
 * and DeviceData ownership in dataBase (dataBase.h:26-29, 239-240). */
int         dsph_create(int device_index, int precision, dsph_ctx** out);
int         dsph_destroy(dsph_ctx* ctx);
const char* dsph_last_error(const dsph_ctx* ctx);
const char* dsph_version(void);

/* ---- set-up ------------------------------------------------------------
 * replaces: dataBase::buildDeviceData + setSpatialGrids
 * (dataBase.cpp:415-449).  Copies the scene, sizes the hashed grid and
 * allocates list storage (fluid-fluid neighbors per particle, fluid-solid and
 * solid-solid totals; pass 0 for the reference defaults 40*N_f-derived).   */
int dsph_upload_scene(dsph_ctx* ctx, const dsph_scene* scene);
/* Batched ensemble: n_scenes independent scenes in one context, advanced
 * together by every phase (one kernel launch per phase for all of them) with
 * one grid per scene (segmented hash grids), so each scene's hashes, lists
 * and fields are those of the scene run alone.  Particles are numbered in
 * scene order (the caller indices of scene k follow those of scene k-1);
 * cluster and clump ids and pebble ranges are shifted accordingly.  Walls
 * act only on their own scene's solids.  The scenes share n_material, the
 * contact tables, gravity (those of scenes[0]) and the time step.  No
 * reference counterpart: the reference runs one scene per process.       */
int dsph_upload_batch(dsph_ctx* ctx, const dsph_scene* scenes,
                      int32_t n_scenes);
/* per-scene offsets of a batch (arrays of n_scenes + 1, any may be NULL):
 * caller indices of fluid, solids, clumps and walls, and global cell
 * indices (hashes read back are global: scene k's cells start at
 * cell_offset[k]) */
int dsph_batch_info(dsph_ctx* ctx, int32_t* n_scenes, int32_t* fluid_offset,
                    int32_t* solid_offset, int32_t* clump_offset,
                    int32_t* wall_offset, int32_t* cell_offset);
int dsph_set_capacity(dsph_ctx* ctx, int32_t ff_per_particle,
                      int64_t fs_total, int64_t ss_total);
/* per-particle rows of the fluid–fluid and fluid–solid Verlet supersets */
int dsph_set_superset_capacity(dsph_ctx* ctx, int32_t ff_per_particle,
                               int32_t fs_per_particle);
int dsph_set_flags(dsph_ctx* ctx, uint32_t flags);
int dsph_set_gravity(dsph_ctx* ctx, const double g[3]); /* dataBase.h:105-109 */
/* 0 disables the sorted-order reorder of fluid particles that follows every
 * superset rebuild (1, the default, keeps it) */
int dsph_set_reorder_period(dsph_ctx* ctx, int32_t period);
/* Verlet skin of the fluid neighbour superset in units of the largest fluid
 * smoothing length (default 0.3; 0 rebuilds the superset every search).
 * Lists stay exact for any value; the skin only trades superset size for
 * rebuild frequency. */
int dsph_set_skin(dsph_ctx* ctx, double skin_over_h);

/* Launch-shape tunables (no reference counterpart; results do not depend on
 * them, only the speed does):
 *   DSPH_TUNE_PAIR_BLOCK   workgroup size of the fluid pair loops, a
 *                          multiple of 64 in [64, 512] (default 256);
 *   DSPH_TUNE_BUILD_BLOCK  particles per wavefront in the Verlet superset
 *                          build, [1, 64] (default 64);
 *   DSPH_TUNE_GROUP        consecutive (cell- and sub-cell-sorted) fluid
 *                          particles that share one superset row, the union
 *                          of their supersets: 1, 2, 4 or 8 (default 1).  A
 *                          shared row costs more candidates per particle but
 *                          lets the group's gathers coalesce.
 *   DSPH_TUNE_TILE         mixed precision only: 1 runs the pair loops over
 *                          tiles of 512 consecutive sorted particles, each
 *                          staging its neighbours' records in LDS once and
 *                          walking its rows there (default); 0 gathers every
 *                          candidate's record from HBM.  Results are bitwise
 *                          the same either way.  Needs the reorder (period
 *                          > 0); a tile whose neighbours do not fit in LDS
 *                          gathers from HBM.
 * The Verlet skin (dsph_set_skin) and the reorder period complete the
 * autotuner's search space (dsph.py autotune, dataBase::autotune).        */
#define DSPH_TUNE_PAIR_BLOCK   1
#define DSPH_TUNE_BUILD_BLOCK  2
#define DSPH_TUNE_GROUP        3
#define DSPH_TUNE_TILE         4
int dsph_set_tuning(dsph_ctx* ctx, int32_t key, int32_t value);
int dsph_get_tuning(dsph_ctx* ctx, int32_t key, int32_t* value);

/* replaces: solverBase::initialize lines 34-37 (solverBase.cpp:81-84):
 * neighborSearch(dev,0,1), addBondData (dataBase.cpp:270-346) and
 * setSolidNormal (dataBase.cpp:348-413).  Normals passed in the scene are
 * kept when compute_normals == 0. */
int dsph_initialize(dsph_ctx* ctx, int compute_normals);

/* ---- the five phases (asynchronous on the context's stream) ------------ */
/* neighborSearch (neighborSearch.h:28, neighborSearch.cu:314-381); lists
 * that overflow are grown and the search re-run (one status read), as the
 * reference's setNum reallocation does                                      */
int dsph_neighbor_search(dsph_ctx* ctx, int32_t i_step, int32_t gap);
/* solidIntegrateBeforeContact (integrate.h:43, integrate.cu:964-980)       */
int dsph_solid_before_contact(dsph_ctx* ctx, double dt);
/* fluidIntegrate (integrate.h:19, integrate.cu:385-421)                    */
int dsph_fluid_integrate(dsph_ctx* ctx, double dt, int32_t i_step, int32_t gap);
/* calSolidContactAfterFluidIntegrate (integrate.h:41, integrate.cu:940-962)*/
int dsph_contact_forces(dsph_ctx* ctx, double dt);
/* solidIntegrateAfterContact (integrate.h:45, integrate.cu:982-998)        */
int dsph_solid_after_contact(dsph_ctx* ctx, double dt);

/* solverBase::update without the user hook and output
 * (solverBase.cpp:47-59): the five phases in order.  check_capacity != 0
 * synchronises after the search and grows lists that overflowed, re-running
 * the search (the reference's setNum realloc, deviceStructs.h:483-491); it
 * returns DSPH_E_NONFINITE when that search met a NaN/inf position.        */
int dsph_step(dsph_ctx* ctx, int32_t i_step, double dt, int32_t gap,
              int check_capacity);
/* n consecutive steps i_step0 .. i_step0+n-1 replayed from captured HIP
 * graphs (no host synchronisation inside).  Replayed searches cannot grow
 * a list: the call first copies the state on the device, and when a list
 * overflowed at some step of the run it returns to that copy and runs the
 * n steps again eagerly with checked searches (as dsph_step with
 * check_capacity), which grow what overflows.  The copy costs one
 * device-to-device pass over the state per call and device memory of the
 * state's size.  A NaN/inf position met by the last search of the run is
 * returned as DSPH_E_NONFINITE (nothing is rolled back).  No reference
 * counterpart (the reference syncs three times per step,
 * neighborSearch.cu:339-340,373). */
int dsph_step_n(dsph_ctx* ctx, int32_t i_step0, int32_t n, double dt,
                int32_t gap);

/* Per-scene time steps of a batch (dsph_upload_batch): scene k advances by
 * dt[k] (n_scenes values, each > 0) per step, as it does run alone, instead
 * of every scene by the one dt of the call.  While they are set, the dt
 * argument of dsph_step, dsph_step_n and the phase functions must be 0
 * (DSPH_E_INVALID otherwise); dt = NULL clears them.  A context of one
 * scene returns DSPH_E_INVALID.  No reference counterpart (the reference
 * runs one scene at a time). */
int dsph_set_scene_time_steps(dsph_ctx* ctx, const double* dt);
int dsph_synchronize(dsph_ctx* ctx);

/* ---- full-state checkpoint ---------------------------------------------
 * No reference counterpart: the reference writes VTU frames only
 * (solverBase.cpp:69-333) and keeps the spring history on the device
 * (deviceStructs.h:506-509, 693-711), so a run cannot be resumed.  The blob
 * holds every device array a later step reads: the fluid in its device
 * order with sumW0 and both record buffers, the hashed grids, the Verlet
 * supersets and their build positions, the fluid-solid and contact lists
 * with their forces and spring history, solids, clumps, wall springs,
 * bonds, the contact tables and the status words, plus the capacities,
 * launch tuning, Verlet skin, flags and gravity of the run.
 * dsph_save_state needs an initialised context; with buffer == NULL it
 * only stores the blob size in *size_out.  dsph_load_state needs a context
 * with the same scene uploaded (same precision and counts); it takes the
 * blob's capacities, tuning, skin, flags and gravity and leaves the
 * context initialised, so the next step continues the saved run bit for
 * bit.  The blob is in this library's own layout (host byte order).       */
int dsph_save_state(dsph_ctx* ctx, void* buffer, int64_t cap, int64_t* size_out);
int dsph_load_state(dsph_ctx* ctx, const void* buffer, int64_t size);

/* ---- runtime parameter changes ------------------------------------------
 * replaces: dataBase::setHertzianContactModel / setLinearContactModel /
 * setBondedContactModel after the run started (dataBase.h:111-167, which
 * re-copy HostSolidContactModel to the device); the tables keep the upload's
 * n_material and field-major layout; NULL leaves a table unchanged.        */
int dsph_set_contact_models(dsph_ctx* ctx, const double* hertz,
                            const double* linear, const double* bonded);
/* replaces: dataBase::setClumpVelocity during the run (dataBase.h:169-182):
 * the clump's translational velocity; its pebbles follow at the next clump
 * kinematics update. */
int dsph_set_clump_velocity(dsph_ctx* ctx, int32_t clump_index,
                            const double velocity[3]);

/* ---- runtime mutation (dataBase.cpp:179-224) ---------------------------- */
int dsph_add_external_force(dsph_ctx* ctx, int32_t solid_index,
                            const double force[3]);
int dsph_add_global_damping(dsph_ctx* ctx, int32_t solid_index, double c_d);

/* ---- read-back (synchronous), caller's original order ------------------
 * replaces: fluid/solid/clump uploadState (deviceStructs.h:199-205,
 * 287-297, 357-365).  Any pointer may be NULL to skip that field.          */
int dsph_download_fluid(dsph_ctx* ctx, double* pos, double* vel,
                        double* acc, double* drho, double* sumw0);
int dsph_download_solid(dsph_ctx* ctx, double* pos, double* vel,
                        double* acc, double* angvel, double* torque,
                        double* normal);
int dsph_download_clump(dsph_ctx* ctx, double* centroid, double* vel,
                        double* acc, double* angvel, double* torque,
                        double* orientation);
/* fluid2Solid per-pair force (integrate.cu:282-330), aligned with the
 * DSPH_LIST_FLUID_SOLID list order; n = number of pairs                    */
int dsph_download_fluid_solid_force(dsph_ctx* ctx, double* force, int64_t cap,
                                    int64_t* n_out);
/* ---- particles added after the run started ----------------------------
 * replaces: dataBase::addFluid / addSolid / addCluster / addClump called
 * after step 0 (dataBase.cpp:113-117, 226-268: uploadState, insertData and
 * copy of the whole device data, setSpatialGrids, neighborSearch(dev, 0, 1))
 * and addBondData's later branch (dataBase.cpp:309-343).  The fluid,
 * solids and clumps of extra go after the existing ones of each species
 * (caller indices continue; extra's clump pebble ranges and solid clump
 * indices count within extra; cluster ids are taken as given).  Every
 * existing particle keeps its state: positions, velocities, accelerations,
 * dRho, sumW0, angular state, normals, contact springs (carried over by
 * (i, j)), wall spring histories and bonds.  The lists are searched again
 * at step 0.  With solids added, a contact of the new list becomes a bond
 * when its pointed solid comes after the last bonded one, the two touch
 * and share a cluster with a bonded modulus.  New solids keep extra's
 * normals (none are estimated after the start).  Differences from the
 * reference: sumW0 is kept, and a new fluid particle's is its kernel sum
 * at the append (the reference reallocates sumW0 as zeros, so its next
 * re-initialisation divides by zero); wall spring histories are kept (the
 * reference rebuilds them empty); with no bond yet any contact may become
 * one.  extra's walls must be empty (DSPH_E_INVALID); its domain, gravity
 * and contact tables are ignored.  Needs an initialised context of one
 * scene (DSPH_E_STATE otherwise, and after dsph_upload_batch).  Graphs are
 * captured again on the next dsph_step_n.                                 */
int dsph_append_scene(dsph_ctx* ctx, const dsph_scene* extra);

/* ---- frames: the VTU output of the step loop -------------------------
 * replaces: solverBase::outputFluidVTU, outputSolidVTU and
 * outputFluid2SolidVTU (solverBase.cpp:69-333), which copy the device state
 * to the host mirrors and format every value as ASCII while the GPU waits.
 * dsph_frame_begin packs the species asked for on the device (FP32 values,
 * caller order, the reference's arrays) on the context's stream and queues
 * the copy of the fluid and solid arrays into pinned host memory on a copy
 * stream of its own: it returns at once and the steps that follow run
 * beside the copy.  dsph_frame_write waits for that copy, copies the
 * fluid–solid pairs (exactly the pairs of the list, a count the device
 * knows) and writes the files as binary appended VTU (raw encoding, UInt64
 * block headers) with the reference's names and arrays:
 *   fluid_NNNN.vtu        points, "smooth length", "velocity", "density"
 *                         (rho0 + dRho), "pressure" (c^2 dRho);
 *   solid_NNNN.vtu        points, "radius", "clusterID", "clumpID",
 *                         "velocity", "angularVelocity", "normal";
 *   fluid2Solid_NNNN.vtu  per fluid–solid pair: the midpoint of the pair,
 *                         "force" (as dsph_download_fluid_solid_force), and
 *                         "fluidID", "solidID" (caller indices), in the
 *                         device list's order (the reference writes its own
 *                         list order, which has no further meaning);
 * each with the FieldData TIME and STEP given to dsph_frame_begin.  NNNN is
 * frame, four digits.  species: a nonzero OR of DSPH_FRAME_FLUID,
 * DSPH_FRAME_SOLID, DSPH_FRAME_FLUID_SOLID.  A frame lives in one of
 * DSPH_FRAME_SLOTS slots until it is written: dsph_frame_write uses only its
 * slot, so it may run on another thread while the caller steps and begins
 * frames in the other slot (not while dsph_frame_begin fills the same slot,
 * and not across dsph_upload_*, dsph_append_scene or dsph_destroy).  It reports by return code
 * only (dsph_last_error is not updated): DSPH_E_STATE when the slot holds no
 * frame, DSPH_E_IO when a file cannot be written.  No reference counterpart
 * for the split (the reference's output is synchronous).                 */
#define DSPH_FRAME_SLOTS       2
#define DSPH_FRAME_FLUID       1
#define DSPH_FRAME_SOLID       2
#define DSPH_FRAME_FLUID_SOLID 4
#define DSPH_FRAME_ALL         7
int dsph_frame_begin(dsph_ctx* ctx, int32_t slot, int32_t i_step, double time,
                     uint32_t species);
int dsph_frame_write(dsph_ctx* ctx, int32_t slot, const char* dir, int32_t frame);
/* solid2Solid per-contact force/torque/springs, aligned with the
 * DSPH_LIST_SOLID_SOLID order; any pointer may be NULL                    */
int dsph_download_contacts(dsph_ctx* ctx, double* force, double* torque,
                           double* sliding, double* rolling, double* torsion,
                           int64_t cap, int64_t* n_out);
/* bonds: normal force, torsion torque, shear force, bending torque, state  */
int dsph_download_bonds(dsph_ctx* ctx, double* normal_force, double* torsion,
                        double* shear, double* bending, int32_t* is_bonded,
                        int64_t cap, int64_t* n_out);

/* Lists in the reference's own order and index space: CSR by the pointed
 * particle ascending, each segment in stencil (z,y,x) order with cell
 * members in ascending original index (neighborSearch.cu:115-312).
 * count[] (may be NULL) receives the per-pointed-particle counts.          */
int dsph_download_list(dsph_ctx* ctx, int kind, int32_t* pointed,
                       int32_t* pointing, int64_t cap, int64_t* n_out,
                       int32_t* count);

/* Hashed grid of the last search (neighborSearch.cu:19-113):
 * hash[] per particle in original order; perm[] = stable sort permutation
 * (sorted slot -> original index, objectHash::index); cell_start/cell_end
 * per cell (-1 = empty), n_cells = Gx*Gy*Gz + 1.  species 0 fluid, 1 solid.
 * For a batch, grid_info describes scene 0 and the arrays span the cells of
 * all scenes (dsph_batch_info gives each scene's cell range).            */
int dsph_grid_info(dsph_ctx* ctx, int32_t* grid_size3, double* cell_size3,
                   int32_t* n_cells);
int dsph_download_grid(dsph_ctx* ctx, int species, int32_t* hash,
                       int32_t* perm, int32_t* cell_start, int32_t* cell_end);

int dsph_status(dsph_ctx* ctx, int64_t words[DSPH_STATUS_WORDS]);

/* Candidates the fluid pair loops walk, from the rows of the last search
 * (synchronous, diagnostic): out[0] fluid particles, out[1] candidates summed
 * over particles (a particle walks its group's whole superset row, or its
 * exact row when walked on the reference stencil), out[2] row entries
 * stored (a shared row once), out[3] the longest row.  Dividing out[1] by
 * out[0] gives the candidates per particle that the pair loops evaluate,
 * against the neighbours per particle of the fluid-fluid list.  No
 * reference counterpart (the reference walks its 27 cells every pass). */
int dsph_superset_stats(dsph_ctx* ctx, int64_t out[4]);

/* The pair-loop tiles of the last superset build (DSPH_TUNE_TILE):
 * out[0] tiles (0 when the tiles are off), out[1] tiles staged in LDS (the
 * others' neighbours did not fit and they gather from HBM), out[2] records
 * staged over those tiles, out[3] the largest tile union.  No reference
 * counterpart. */
int dsph_tile_stats(dsph_ctx* ctx, int64_t out[4]);

/* ---- the hashed grid of caller-held particles --------------------------
 * replaces: updateGridCellStartEnd for one species (neighborSearch.cu:89-113)
 * — calculatePointHash (:19-36), the stable key/value sort of the hashes
 * with values 0..n-1 (sortKeyValuePairs, :3-9) and findStartAndEnd (:46-87)
 * — on arrays the caller holds in device memory, on the context's device
 * and stream (the context needs no scene).  The grid is the spatialGrid of
 * dataBase::setSpatialGrids (dataBase.cpp:415-432): bounds origin ..
 * origin + size, cells[] cells per axis, cell size size / cells, and a
 * spare cell n_cells - 1 (n_cells = cells[0]*cells[1]*cells[2] + 1) for
 * points outside the bounds.
 *   pos                   device, [3n] xyz doubles
 *   hash                  device out, [n]: each point's cell
 *   perm                  device out, [n]: sorted position -> point index
 *                         (ascending cell, ascending index within a cell)
 *   cell_start, cell_end  device out, [n_cells]: the sorted positions
 *                         [start, end) of each cell, -1 for empty cells
 * Asynchronous: the results are complete after dsph_synchronize.  The work
 * buffer (about 5 words per point) is the context's, grown on demand.    */
int dsph_hash_sort(dsph_ctx* ctx, const double* pos, int32_t n, const double origin[3],
                   const double size[3], const int32_t cells[3], int32_t* hash,
                   int32_t* perm, int32_t* cell_start, int32_t* cell_end);
/* dsph_hash_sort for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * no reference counterpart.  Same arguments and results; the work buffer is
 * a fixed 24 MiB device array of the library (each GPU it runs on holds its
 * own copy until the process ends), so it takes at most 1,048,576 points
 * (larger n: DSPH_E_INVALID) and calls no allocating routine.  Calls on one
 * GPU share that array: calls that overlap on one GPU (two streams or
 * contexts at once) race.                                                  */
int dsph_hash_sort_noalloc(dsph_ctx* ctx, const double* pos, int32_t n,
                           const double origin[3], const double size[3],
                           const int32_t cells[3], int32_t* hash, int32_t* perm,
                           int32_t* cell_start, int32_t* cell_end);

/* profiling aid: per kernel family, the total device time (ms) and the
 * launch count since the last call or dsph_set_timing, measured with HIP
 * events on the context stream (ms[2k], ms[2k+1]); the families include
 * "frame_pack" (dsph_frame_begin's packing) and "frame_copy" (its copy, on
 * the frames' copy stream).  names is a NUL-separated list.  Returns the
 * number of entries written. */
int dsph_kernel_times(dsph_ctx* ctx, char* names, int64_t names_cap,
                      double* ms, int32_t cap);
int dsph_set_timing(dsph_ctx* ctx, int enabled);

#ifdef __cplusplus
}
#endif
#endif /* DSPH_H */

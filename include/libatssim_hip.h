/*
 * libatssim_hip.h -- C ABI of the MI355X-native VBD cloth backend.
 *
 * This is the GPU device slot of the reference's heterogeneous task
 * scheduler (ChengzhuUwU/libAtsSim).  In the reference the slot is the
 * lambda registered as `Launcher::Implementation(DeviceTypeGpu, ...)`
 * (src/heterogeneous/launcher.h:556-579) that forwards each scheduled task to
 * `GpuSolver::fn_dispatch(const Launcher::LaunchParam&)`
 * (examples/example3_hybrid_cpu_gpu.cpp:1744-1920), which encodes Metal
 * kernels into a command list.  Here the same tasks are HIP kernels on one
 * MI355X (gfx950), reached through plain C entry points: plain pointers and
 * sizes, integer status codes, a thread-local last-error string.
 *
 * All positions/velocities crossing this boundary are float[3 * num_verts]
 * in vertex-id order (the reference's AoS Float3 with 12-byte stride).
 * Device pointers are only exposed by the lats_vbd_device_* accessors.
 */
#ifndef LIBATSSIM_HIP_H
#define LIBATSSIM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference prints and continues; here every call
 *      returns one of these and sets lats_last_error()) ------------------- */
#define LATS_OK 0
#define LATS_ERR_INVALID_ARGUMENT 1
#define LATS_ERR_HIP 2
#define LATS_ERR_NO_DEVICE 3
#define LATS_ERR_UNSUPPORTED 4
#define LATS_ERR_OUT_OF_MEMORY 5
#define LATS_ERR_TIMEOUT 6          /* a bounded host wait gave up */

/* Message of the last failing call on this thread ("" if none). */
const char* lats_last_error(void);

/* ======================================================================
 * Topology: replaces SimMesh::read_mesh_file's edge extraction
 * (src/Utils/mesh_reader.h:206-298) and the adjacency / rest-state /
 * colouring set-up of init_mesh + CpuSolver::init_xpbd_system
 * (examples/example2_async_iter.cpp:229-374, 572-646).
 * ====================================================================== */
typedef struct lats_topology lats_topology;

/* which-array selectors for lats_topology_size / lats_topology_copy */
#define LATS_TOPO_EDGES 0            /* int32[2E]  sorted (min,max) edges            */
#define LATS_TOPO_BENDING 1          /* int32[4B]  (e0,e1,opp_lowfid,opp_highfid)    */
#define LATS_TOPO_REST_LENGTH 2      /* float[E]                                     */
#define LATS_TOPO_BENDING_Q 3        /* float[16B] Q[b][col][row]                    */
#define LATS_TOPO_ADJ_EDGES_CSR 4    /* uint32 packed CSR (V+1 offsets, then ids)    */
#define LATS_TOPO_ADJ_BENDING_CSR 5  /* uint32 packed CSR                            */
#define LATS_TOPO_CLUSTERS_CSR 6     /* uint32 packed CSR of colour clusters         */
#define LATS_TOPO_COLOUR 7           /* uint8[V] colour id per vertex                */
#define LATS_TOPO_FACES 8            /* int32[3F] triangles                          */

int lats_topology_build(const float* x, uint32_t num_verts, const int32_t* faces, uint32_t num_faces,
                        lats_topology** out);
void lats_topology_destroy(lats_topology* t);
/* The topology on disk (replaces save_to_binary / read_from_binary,
 * src/Utils/mesh_reader.h:321-343, without their hard-coded path): one file
 * per array in directory `dir` (which must exist), "<name>.dat" holding the
 * array's raw little-endian elements, the element count given by the file
 * size.  Files: faces (int32 3F), edges (int32 2E), bending_edges (int32 4B),
 * edges_rest_state_length (float E), bending_edges_Q (float 16B),
 * vert_adj_edges, vert_adj_bending_edges, clusterd_per_vertex_bending (uint32
 * packed CSR as upload_2d_csr_from, examples/example2_async_iter.cpp:99-127),
 * per_vertex_bending_cluster_id (uint8 V).  Loading checks the arrays against
 * each other (CSR headers and offsets, id ranges, every vertex in exactly one
 * cluster, cluster ids) and fails with LATS_ERR_INVALID_ARGUMENT otherwise. */
int lats_topology_save(const lats_topology* t, const char* dir);
int lats_topology_load(const char* dir, lats_topology** out);

/* ---- Wavefront OBJ (mesh input and frame export) ----------------------
 * Reading keeps what SimMesh::read_mesh_file keeps of a file
 * (src/Utils/mesh_reader.cpp:11-200, tinyobjloader defaults): `v` positions
 * in file order; faces fan-triangulated from their first corner, 1-based or
 * negative indices, `v/vt/vn` corners; a triangle repeating a vertex is
 * dropped (mesh_reader.cpp:172-182).  Writing is the format of
 * SolverInterface::save_mesh_to_obj (examples/example2_async_iter.cpp:1649-1710). */
typedef struct lats_obj lats_obj;
int lats_obj_read(const char* path, lats_obj** out);
int lats_obj_parse(const char* text, lats_obj** out);
/* counts: vertices, triangles, triangles dropped */
int lats_obj_counts(const lats_obj* obj, uint32_t counts[3]);
/* x: float[3 * vertices], faces: int32[3 * triangles] (0-based); either may be null */
int lats_obj_copy(const lats_obj* obj, float* x, int32_t* faces);
void lats_obj_destroy(lats_obj* obj);
int lats_obj_write(const char* path, const float* x, uint32_t num_verts, const int32_t* faces, uint32_t num_faces,
                   uint32_t mesh_id);
/* counts[0..3] = num_verts, num_edges, num_bending_edges, num_clusters */
int lats_topology_counts(const lats_topology* t, uint32_t counts[4]);
/* number of elements (not bytes) of array `which` */
int lats_topology_size(const lats_topology* t, int which, uint64_t* num_elements);
int lats_topology_copy(const lats_topology* t, int which, void* dst);

/* ======================================================================
 * Seeded synthetic scenes (grid cloths and per-frame wind).
 * ====================================================================== */
/* x: float[3*nx*ny]; faces: int32[6*(nx-1)*(ny-1)] */
int lats_scene_grid(uint32_t nx, uint32_t ny, float spacing, float jitter, uint64_t seed, float* x,
                    int32_t* faces);
/* Icosphere (icosahedron subdivided `level` times): call with x == NULL to
 * get the sizes, then with arrays of 3*num_verts floats and 3*num_faces ints. */
int lats_scene_icosphere(uint32_t level, float radius, const float centre[3], float* x, int32_t* faces,
                         uint32_t* num_verts, uint32_t* num_faces);
/* out: float[3*num_frames], force per vertex for each frame */
int lats_scene_wind(uint32_t num_frames, uint32_t num_verts, float max_total_force, uint64_t seed, float* out);

/* ======================================================================
 * Single-cloth VBD context: the GPU slot's state (the reference's
 * BasicMeshData + XpbdData SharedArrays, example3_hybrid_cpu_gpu.cpp:16-92)
 * held in HBM.
 * ====================================================================== */

/* Subset of SceneParams (src/SharedDefine/scene_params.h:9-163) read by the
 * VBD path. */
typedef struct lats_scene_params {
    float implicit_dt;                 /* scene_params.h:94 (1/60)                       */
    uint32_t num_substep;              /* scene_params.h:68                               */
    uint32_t constraint_iter_count;    /* scene_params.h:71                               */
    float stiffness_stretch_spring;    /* scene_params.h:101                              */
    float stiffness_quadratic_bending; /* get_stiffness_quadratic_bending() (k * bending_ui), :160 */
    float damping_cloth;               /* scene_params.h:115                              */
    float gravity[3];                  /* scene_params.h:145; predict uses (0,-9.8,0) as the reference */
} lats_scene_params;

/* Per-mesh arrays, consumed as-is (colouring and clusters are not
 * recomputed).  Formats are those of lats_topology_copy. */
typedef struct lats_mesh_desc {
    uint32_t num_verts, num_edges, num_bending_edges, num_clusters;
    const float* rest_x;               /* float[3V]: initial state, v = 0                 */
    const int32_t* edges;              /* int32[2E]                                       */
    const int32_t* bending_edges;      /* int32[4B]                                       */
    const float* edges_rest_length;    /* float[E]                                        */
    const float* bending_Q;            /* float[16B]                                      */
    const uint32_t* vert_adj_edges_csr;
    const uint32_t* vert_adj_bending_csr;
    const uint32_t* clusters_csr;
    const float* vert_mass;            /* float[V]                                        */
    const uint8_t* is_fixed;           /* uint8[V]                                        */
} lats_mesh_desc;

typedef struct lats_vbd lats_vbd;

/* Frame-execution strategies of lats_vbd_physics_step / lats_vbd_run_frames. */
#define LATS_STRATEGY_AUTO 0        /* pick the fastest for the mesh                    */
#define LATS_STRATEGY_GRAPH 1       /* one kernel per colour, frame captured in a hipGraph */
#define LATS_STRATEGY_PERSISTENT 2  /* one cooperative kernel per frame, grid barrier per colour */
#define LATS_STRATEGY_FLOW 3        /* one cooperative kernel per run of frames, tiles of vertices hand
                                       each colour pass on to their neighbouring tiles (no grid barrier) */
#define LATS_STRATEGY_PATCH 4       /* several colour passes per launch: each workgroup holds one patch of
                                       the cloth and its halo in LDS and recomputes the halo's updates
                                       (identical values), frame captured in a hipGraph */

/* Creates the context on the current HIP device; uploads everything. */
int lats_vbd_create(const lats_mesh_desc* mesh, const lats_scene_params* params, lats_vbd** out);
int lats_vbd_destroy(lats_vbd* ctx);
int lats_vbd_set_strategy(lats_vbd* ctx, int strategy);
/* Patch strategy geometry.  patches: a multiple of 8 (0: about 256 vertices
 * per patch, 64..1024 patches); colours_per_launch: colour passes one launch
 * runs (0 or >= the colour count: a whole sweep per launch); threads per
 * workgroup: 256, 512 or 1024.  Patch updates run on 8 lanes per vertex
 * whatever lats_vbd_set_lanes_per_vertex sets (the other strategies use
 * that).  The layout is built on first use; a region that does not fit one
 * workgroup's LDS gives
 * LATS_ERR_UNSUPPORTED.  No counterpart in the reference (its colour loop
 * dispatches one pass per colour, example3_hybrid_cpu_gpu.cpp:1217-1348). */
typedef struct lats_patch_info {
    uint32_t patches, colours_per_launch, launches_per_sweep, threads, lanes;
    uint32_t max_region;                /* largest region of one workgroup (LDS slots, 16 B each) */
    uint64_t updates_per_sweep;         /* vertex updates of one sweep, halo updates included */
    uint64_t region_slots_per_sweep;    /* positions staged into LDS per sweep */
    uint64_t ell_entries_per_sweep;     /* local ELL entries (springs + pairs) read per sweep */
    double build_ms;                    /* host time of the last layout build */
} lats_patch_info;
int lats_vbd_set_patches(lats_vbd* ctx, uint32_t patches, uint32_t colours_per_launch, uint32_t threads);
/* Builds the layout if needed, then describes it. */
int lats_vbd_get_patches(lats_vbd* ctx, lats_patch_info* info);
/* Solver of physics_step / run_frames / dispatch: VBD (default) or XPBD
 * (CpuSolver::physics_step_xpbd, example2_async_iter.cpp:1017-1081: predict,
 * multipliers reset, per iteration every spring cluster then every bending
 * cluster, update velocity; constraint clusters from the reference's greedy
 * constraint colouring, built on the first switch).  XPBD runs the graph
 * strategy, ignores external forces (the reference's XPBD has none) and
 * rejects obstacles and the asynchronous buffers. */
#define LATS_SOLVER_VBD 0
#define LATS_SOLVER_XPBD 1
int lats_vbd_set_solver(lats_vbd* ctx, int solver);
int lats_vbd_get_solver(lats_vbd* ctx, int* solver);
/* XPBD constraint clusters as a packed CSR (n+1 offsets, then constraint
 * ids in the reference's edge / bending-pair numbering); which: 0 springs,
 * 1 bending pairs.  *size receives the element count; copies min(cap, size). */
int lats_vbd_xpbd_clusters(lats_vbd* ctx, int which, uint32_t* csr, uint64_t cap, uint64_t* size);
/* Lanes per vertex of the colour passes (1, 4, 8, 16): with G > 1 the
 * incident springs and bending pairs of a vertex are gathered by G lanes in
 * parallel and summed in the reference's order (results are identical). */
int lats_vbd_set_lanes_per_vertex(lats_vbd* ctx, uint32_t lanes);
int lats_vbd_get_lanes_per_vertex(lats_vbd* ctx, uint32_t* lanes);
/* Block mapping of the colour passes (graph strategy): every colour is cut
 * into `regions` (1..8) bands of consecutive vertex ids, of equal vertex
 * count, and block b of every colour pass works on band b % regions.  With 8, each XCD updates the same
 * band in every pass (workgroups go to the XCDs round-robin).  Results are
 * identical for every value; default 1. */
int lats_vbd_set_regions(lats_vbd* ctx, uint32_t regions);
int lats_vbd_get_regions(lats_vbd* ctx, uint32_t* regions);
/* Threads per block of the colour passes (graph strategy): 64, 128 or 256
 * (default).  A colour pass of a small cloth has few waves; smaller blocks
 * spread them more evenly over the CUs.  Results are identical. */
int lats_vbd_set_colour_block(lats_vbd* ctx, uint32_t threads);
int lats_vbd_get_colour_block(lats_vbd* ctx, uint32_t* threads);
/* Grid size of the persistent strategy (0: one thread per vertex of the
 * largest colour, capped at the co-resident limit); for the flow strategy
 * the number of vertex tiles, one block each (0: min(256, V / 128)). */
int lats_vbd_set_persistent_blocks(lats_vbd* ctx, uint32_t blocks);
int lats_vbd_get_strategy(lats_vbd* ctx, int* strategy, uint32_t* persistent_blocks);
/* Autotuner: times `frames` frames with the graph strategy (1 or 8 cloth
 * bands, 64 / 128 / 256-thread blocks), the persistent strategy at several grid sizes and the flow
 * strategy at several tile counts, for 1/4/8/16 lanes per vertex, in two
 * passes (a candidate's best time counts), keeps the fastest for later
 * lats_vbd_physics_step / lats_vbd_run_frames calls, and restores the state. */
int lats_vbd_autotune(lats_vbd* ctx, uint32_t frames, int* chosen_strategy, uint32_t* chosen_blocks);
/* What the last lats_vbd_autotune measured: *count candidates; out (6 floats
 * per candidate, min(cap, count) copied) = {strategy, persistent blocks or
 * flow tiles, lanes per vertex, bands, best milliseconds for `frames` frames,
 * colour-pass block threads}. */
int lats_vbd_autotune_results(lats_vbd* ctx, float* out, uint32_t cap, uint32_t* count);
/* Include the bending term in the recorded energy (the reference's GPU
 * energy does, example3.metal:369-395; its CPU energy drops it,
 * example2_async_iter.cpp:832-846).  Default 1. */
int lats_vbd_set_energy_bending(lats_vbd* ctx, int include_bending);

/* Frame state in vertex-id order (≙ sa_x_frame_end / sa_v_frame_end). */
int lats_vbd_set_state(lats_vbd* ctx, const float* x, const float* v);
int lats_vbd_get_state(lats_vbd* ctx, float* x, float* v);
/* Uniform external force per vertex, added to m*g in the inertia term. */
int lats_vbd_set_external_force(lats_vbd* ctx, const float force[3]);

/* One frame (≙ SolverInterface::physics_step(SolverTypeVBD) +
 * GpuSolver::physics_step_vbd, example3_hybrid_cpu_gpu.cpp:1510-1576).
 * record_energy != 0: one substep and the energy history of
 * constraint_iter_count + 1 values (before iteration 0, then after each). */
int lats_vbd_physics_step(lats_vbd* ctx, int record_energy);
int lats_vbd_get_energy(lats_vbd* ctx, float* out, uint32_t capacity, uint32_t* count);

/* num_frames frames back to back on the device, external force per frame
 * from forces[3*f] (NULL: keep the current one).  No host synchronisation
 * until the call returns. */
int lats_vbd_run_frames(lats_vbd* ctx, uint32_t num_frames, const float* forces);
/* GPU time (HIP events on the context's stream) of the last
 * lats_vbd_physics_step / lats_vbd_run_frames call, in ms. */
int lats_vbd_last_elapsed_ms(lats_vbd* ctx, float* ms);
int lats_vbd_synchronize(lats_vbd* ctx);

/* ---- obstacle contact (VBD::compute_obstacle_collision,
 *      xpbd_constraints.h:4238-4284; narrow phase NarrowPhase::template_vf
 *      :2742-2820; hash SpatialHashing :2333-2393) ------------------------
 * A static triangle obstacle (face normals = normalize(cross(b-a, c-a)),
 * i.e. counter-clockwise faces point outwards).  After predict, every
 * vertex takes the (up to 4) nearest faces within `radius` from the hash
 * cell that contains it (cells of size `cell` >= 2*radius hold every face
 * whose box grown by `radius` overlaps them); each contact adds the
 * penalty k*(thickness - (x - xs).n)^2/2 while (x - xs).n < thickness.
 * nf == 0 removes the obstacle. */
int lats_vbd_set_obstacle(lats_vbd* ctx, const float* x, uint32_t num_verts, const int32_t* faces, uint32_t num_faces,
                          float thickness, float stiffness, float radius, float cell);
/* Contacts of the last detection, vertex-id order: counts[V], face ids
 * fids[4*V] (ascending, -1 padded). */
int lats_vbd_get_contacts(lats_vbd* ctx, uint32_t* counts, int32_t* fids);

/* ---- frame pieces (physics_step_vbd, example2_async_iter.cpp:1082-1143,
 *      in stages; all enqueue on the context's stream) ------------------- */
int lats_vbd_frame_begin(lats_vbd* ctx);              /* predict (+ contact detection)      */
int lats_vbd_iterate(lats_vbd* ctx, uint32_t sweeps); /* colour sweeps on the positions     */
int lats_vbd_frame_end(lats_vbd* ctx);                /* update velocity, advance the frame */
/* Energy (incremental potential) of the current positions; synchronises. */
int lats_vbd_energy_now(lats_vbd* ctx, float* energy);
/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops).  The energy of
 * lats_vbd_energy_now, by the same kernels, of the positions x_dev (device
 * float[3V] in slot order, see lats_vbd_device_positions; NULL: the
 * context's current positions), with the context's frame start, velocities
 * and contacts; one float is written to energy_dev (device memory).  Enqueued
 * on the context's stream; returns without waiting.  Allocates nothing: the
 * FP64 block partials use one fixed array per GPU of 32768 doubles (256 KiB),
 * so vertices + springs + bending pairs must be <= 8,388,608
 * (LATS_ERR_INVALID_ARGUMENT otherwise).  Every call on one GPU uses that
 * array: calls on different contexts of one GPU whose executions overlap
 * (different streams) race and must be ordered by the caller. */
int lats_vbd_energy_noalloc(lats_vbd* ctx, const float* x_dev, float* energy_dev);

/* ---- scheduler slot: GpuSolver::fn_dispatch ------------------------- */
/* Function ids of the tasks the VBD DAG registers
 * (example3_hybrid_cpu_gpu.cpp:1921-1988; FunctionID launcher.h:15-221). */
#define LATS_FN_ADDITIONAL_ROOT 0
#define LATS_FN_ADDITIONAL_TERMINAL 1
#define LATS_FN_PREDICT_POSITION 2      /* id_xpbd_predict_position     */
#define LATS_FN_VBD_ALL_IN_ONE 3        /* id_vbd_all_in_one            */
#define LATS_FN_CONSTRAINT_LAST_NODE 4  /* id_xpbd_constraint_last_node */
#define LATS_FN_UPDATE_VELOCITY 5       /* id_xpbd_update_velocity      */
#define LATS_FN_RESET_CONSTRAINS 6      /* id_xpbd_reset_constrains (XPBD multipliers; no state in VBD) */
#define LATS_FN_XPBD_STRETCH 7          /* id_xpbd_constraint_stretch_mass_spring, cluster_idx */
#define LATS_FN_XPBD_BENDING 8          /* id_xpbd_constraint_bending, cluster_idx */

#define LATS_BUFFER_DEFAULT 0x0FFFFFFFu /* default_buffer_mask / input_buffer_mask, launcher.h:836-837 */
#define LATS_BUFFER_NONE 0xFFFFFFFFu    /* -1u                                   */
#define LATS_NUM_ASYNC_BUFFERS 32       /* example3_hybrid_cpu_gpu.cpp:1747        */

/* Mirror of Launcher::LaunchParam (launcher.h:435-539) fields the slot reads. */
typedef struct lats_launch_param {
    uint32_t function_id;
    uint32_t iter_idx;
    uint32_t cluster_idx;
    uint32_t buffer_idx;
    uint32_t left_buffer_idx;
    uint32_t right_buffer_idx;
    uint32_t num_input_buffers;
    const uint32_t* input_buffer_idxs;
    uint32_t is_allocated_to_main_device;
} lats_launch_param;

/* Enqueues the task's kernels on the context's stream (asynchronous). */
int lats_vbd_dispatch(lats_vbd* ctx, const lats_launch_param* param);
/* Energy recording for dispatched tasks (≙ SceneParams::print_xpbd_convergence):
 * on != 0 clears the history; energies are then appended at the points the
 * reference's prev/post hooks evaluate them (example3_hybrid_cpu_gpu.cpp:1819-1855).
 * Read them with lats_vbd_get_energy. */
int lats_vbd_set_record_energy(lats_vbd* ctx, int on);

/* ---- CPU<->GPU signalling for the heterogeneous launch mode ---------
 * Replaces the MTLSharedEvent pair (command_list.h:474-480 wait_cpu,
 * launcher.cpp:1017-1021/1092-1095 signal) and waitUntilCompleted.  The
 * host flag lives in pinned host memory. */
/* Enqueue: the stream waits until the host flag >= value. */
int lats_vbd_stream_wait_host(lats_vbd* ctx, uint32_t value);
/* Host side: set the host flag (CPU signals the GPU). */
int lats_vbd_host_signal(lats_vbd* ctx, uint32_t value);
/* Enqueue a GPU->CPU completion marker; lats_vbd_wait_marker blocks the
 * host until the marker has passed. */
int lats_vbd_record_marker(lats_vbd* ctx, uint32_t marker);
int lats_vbd_wait_marker(lats_vbd* ctx, uint32_t marker);
/* Non-blocking: *done = 1 once the marker has passed (for bounded waits). */
int lats_vbd_query_marker(lats_vbd* ctx, uint32_t marker, int* done);
/* Elapsed GPU ms between two markers (≙ GPUEndTime - GPUStartTime). */
int lats_vbd_marker_elapsed_ms(lats_vbd* ctx, uint32_t m0, uint32_t m1, float* ms);

/* ---- position buffers across processors ------------------------------
 * The reference shares every buffer between CPU and GPU (unified memory,
 * SharedArray, example3_hybrid_cpu_gpu.cpp:16-92); here the context owns
 * them in device memory and a heterogeneous launch moves a buffer across at
 * a cross-processor edge.  Both calls are enqueued on the context's stream
 * (order them with markers / host waits), use float[3V] in vertex order on
 * the host side and need pinned host memory (lats_host_alloc) to run
 * asynchronously. */
#define LATS_BUF_ITER 0     /* iterate buffer `idx` (sa_async_iter_positions) */
#define LATS_BUF_BEGIN 1    /* begin buffer `idx` (sa_async_begin_positions)  */
#define LATS_BUF_X 2        /* sa_x (idx ignored)                             */
#define LATS_BUF_X_START 3  /* sa_x_start                                     */
#define LATS_BUF_V 4        /* sa_v                                           */
int lats_vbd_buffer_download(lats_vbd* ctx, int which, uint32_t idx, float* host);
int lats_vbd_buffer_upload(lats_vbd* ctx, int which, uint32_t idx, const float* host);
/* Pinned (page-locked) host memory for the copies above. */
int lats_host_alloc(size_t bytes, void** out);
int lats_host_free(void* p);

/* ======================================================================
 * Batched cloths: many independent cloths of uneven sizes, one persistent
 * launch per frame (a block owns a whole cloth; cloths are pulled largest
 * first from a work counter), the frame captured in a hipGraph.  No
 * reference counterpart (the reference runs one cloth); every cloth follows
 * the single-cloth arithmetic above bit for bit.
 * ====================================================================== */
typedef struct lats_batch lats_batch;
/* meshes[k] as for lats_vbd_create; stretch[k] / bending[k] override the
 * stiffnesses of params per cloth (NULL: use params). */
int lats_batch_create(uint32_t num_cloths, const lats_mesh_desc* meshes, const float* stretch, const float* bending,
                      const lats_scene_params* params, lats_batch** out);
int lats_batch_destroy(lats_batch* b);
int lats_batch_counts(lats_batch* b, uint32_t* num_cloths, uint64_t* total_verts);
int lats_batch_set_state(lats_batch* b, uint32_t cloth, const float* x, const float* v);
int lats_batch_get_state(lats_batch* b, uint32_t cloth, float* x, float* v);
/* Threads per cloth-block (multiple of 64, <= 1024) and resident blocks
 * (0: as many as fit). */
int lats_batch_set_launch(lats_batch* b, uint32_t block_threads, uint32_t blocks);
int lats_batch_get_launch(lats_batch* b, uint32_t* block_threads, uint32_t* blocks);
/* Lanes per vertex of the batch kernel (1, 4 or 8; bit-identical results). */
int lats_batch_set_lanes_per_vertex(lats_batch* b, uint32_t lanes);
int lats_batch_get_lanes_per_vertex(lats_batch* b, uint32_t* lanes);
/* LDS staging: a cloth of at most `slots` vertices keeps its positions in
 * its block's LDS (12 bytes per vertex) for the whole frame and writes them
 * back at the frame's end; larger cloths gather from HBM.  0 turns staging
 * off; more than max_slots (what one block's LDS holds on this device) is
 * rejected.  The default stages every cloth that fits.  Results are
 * bit-identical either way. */
int lats_batch_set_lds_slots(lats_batch* b, uint32_t slots);
int lats_batch_get_lds_slots(lats_batch* b, uint32_t* slots, uint32_t* max_slots);
/* Autotuner of the cloth-to-block mapping (LDS staging off / on x lanes per
 * vertex x block size x resident blocks): times `frames` frames for each
 * candidate, keeps the fastest, restores the state. */
int lats_batch_autotune(lats_batch* b, uint32_t frames, uint32_t* block_threads, uint32_t* blocks);
int lats_batch_run_frames(lats_batch* b, uint32_t num_frames);
int lats_batch_last_elapsed_ms(lats_batch* b, float* ms);

/* ---- device accessors (benchmarks, zero-copy interop) ---------------- */
/* Device pointer of the colour-ordered position buffer, float[3V]; slot s
 * holds vertex perm[s] (lats_vbd_copy_permutation). */
int lats_vbd_device_positions(lats_vbd* ctx, void** dptr);
int lats_vbd_copy_permutation(lats_vbd* ctx, uint32_t* perm);
/* The context's HIP stream (a hipStream_t): a caller's own kernels and copies
 * enqueued on it run in order with the context's work, host-value waits and
 * markers (≙ adding tasks to the reference's command list,
 * command_list.h:474-480).  Owned by the context. */
int lats_vbd_get_stream(lats_vbd* ctx, void** stream);
/* Measurement: captures `iterations` colour sweeps (one VBD kernel launch
 * per colour) in a graph, replays it once to warm up and once between two
 * HIP events on the context's stream, then restores the positions.
 * mean_ms = event time / launches (kernels run back to back, as in a frame
 * graph); launches = number of launches timed; vertex_updates = vertex
 * updates they performed. */
int lats_vbd_time_colour_kernels(lats_vbd* ctx, uint32_t iterations, float* mean_ms, uint32_t* launches,
                                 uint64_t* vertex_updates);
/* Measurement, XPBD solver: `iterations` passes over every spring cluster
 * (one kernel per cluster) captured in one graph and timed as above, then
 * the same over every bending cluster; the positions are restored (the
 * multipliers restart at zero with the next frame).  *_ms = mean
 * milliseconds per launch, *_launches = launches timed. */
int lats_vbd_time_xpbd_kernels(lats_vbd* ctx, uint32_t iterations, float* stretch_ms, float* bending_ms,
                               uint32_t* stretch_launches, uint32_t* bending_launches);
/* Bytes of device memory held by the context. */
int lats_vbd_device_bytes(lats_vbd* ctx, uint64_t* bytes);

/* ---- learned warm start (no reference counterpart) -------------------
 * The reference starts every VBD frame from the inertial guess
 * (Core::predict_position, src/Solver/xpbd_constraints.h:321-384).  This
 * optional module predicts the end-of-frame correction x_end - x_pred of
 * every vertex with a per-vertex MLP (48 features -> 128 -> 128 -> 128 -> 3,
 * SiLU; BF16 weights on the matrix cores, FP32 accumulation and FP32 master
 * weights), trained with Adam on frames the plain solver produces, and
 * keeps the prediction only when it lowers the frame's VBD energy.
 * Parameters are float[43408]: W1 [64][128] (rows 48..63 zero), b1[128],
 * W2 [128][128], b2, W3 [128][128], b3, W4 [128][16] (columns 3..15 zero),
 * b4[16]; W row-major [in][out]. */
typedef struct lats_mlp lats_mlp;
int lats_mlp_create(uint64_t seed, lats_mlp** out);
int lats_mlp_destroy(lats_mlp* m);
int lats_mlp_num_params(uint32_t* n);
int lats_mlp_get_params(lats_mlp* m, float* params);
int lats_mlp_set_params(lats_mlp* m, const float* params);
/* Features are divided by these 48 scales before the network sees them. */
int lats_mlp_get_feature_scale(lats_mlp* m, float* scale48);
int lats_mlp_set_feature_scale(lats_mlp* m, const float* scale48);
/* Checks on caller rows (float[M][48] -> float[M][3]; inputs rounded to BF16). */
int lats_mlp_forward_host(lats_mlp* m, const float* X, uint32_t M, float* Y);
/* Gradient of mean((Y - T)^2) over the M x 3 outputs (no update). */
int lats_mlp_gradient_host(lats_mlp* m, const float* X, const float* T, uint32_t M, float* grad, double* loss);
/* Binds a solver context (allocates per-vertex buffers, clears the
 * correction history).  The frame calls below need the bound context and
 * fail on any other one (also on a new context created where a destroyed
 * bound one was).  Every frame call returns with the context's stream idle. */
int lats_mlp_bind(lats_mlp* m, lats_vbd* ctx);
int lats_mlp_reset_history(lats_mlp* m);
/* The learned guess: the prediction smoothed `smooth_passes` times over the
 * spring 1-ring ((own + ring mean) / 2, pinned neighbours as zero), and, with
 * half_step != 0, retried at half length before the inertial fallback. */
int lats_mlp_set_guess_options(lats_mlp* m, uint32_t smooth_passes, int half_step);
/* Stopping rule of a warm-started frame that is given no target (target <= 0
 * in lats_mlp_warm_start_frame, use_model != 0): from a kept guess, sweep
 * until the energy decrease per sweep since the previous check (the first
 * check compares with the guess's energy; a rise counts as no decrease) is
 * <= rel_decrease * |energy|, at most max_sweeps; from the inertial guess,
 * the plain iteration count (the plain frame).  No reference solve is
 * involved, so a frame may stop above the energy the plain solve would
 * reach.  0 turns it off (then such a frame runs max_sweeps); default 0. */
int lats_mlp_set_stop_rule(lats_mlp* m, float rel_decrease);
/* Training rows: one plain frame (predict, all iterations, update) of the
 * bound context, recording its features and its correction as rows. */
int lats_mlp_reserve_rows(lats_mlp* m, uint64_t rows);
int lats_mlp_record_frame(lats_mlp* m, lats_vbd* ctx);
/* Appends caller rows (float[M][48] features, already scaled; float[M][3] targets). */
int lats_mlp_dataset_append(lats_mlp* m, const float* X, const float* T, uint32_t M);
int lats_mlp_dataset_rows(lats_mlp* m, uint64_t* rows);
int lats_mlp_dataset_read(lats_mlp* m, uint64_t first, uint32_t count, float* X, float* T);
int lats_mlp_dataset_clear(lats_mlp* m);
/* Sets the scales to the recorded features' RMS and rescales the rows. */
int lats_mlp_fit_feature_scale(lats_mlp* m, float* scale48);
/* `steps` Adam steps (lr, beta 0.9 / 0.999) on batches of batch_rows rows
 * drawn with replacement from splitmix64(seed + step). */
int lats_mlp_train(lats_mlp* m, uint32_t steps, uint32_t batch_rows, float lr, uint64_t seed, double* mean_loss);
/* Measurement: mean duration (HIP events on the network's stream) of one
 * inference forward pass over the first `rows` recorded rows, and of one
 * whole training step on them (forward, backward, reduction, Adam at
 * lr 1e-4, BF16 weight packing: the parameters change). */
int lats_mlp_time_kernels(lats_mlp* m, uint32_t rows, uint32_t iters, float* forward_ms, float* step_ms);
/* One frame of the bound context: predict, learned guess (use_model != 0)
 * kept only if its energy is <= the inertial guess's, then colour sweeps
 * until the energy is <= target (target <= 0: the stopping rule of
 * lats_mlp_set_stop_rule, or max_sweeps when it is off), at most
 * max_sweeps; a kept guess gets at least first_check sweeps, then a check
 * every check_every; from the inertial guess the first check comes after
 * the plain iteration count; update.  out[6] = {sweeps, kept,
 * E(inertial guess), E(learned guess), E(final), target}. */
int lats_mlp_warm_start_frame(lats_mlp* m, lats_vbd* ctx, float target, uint32_t max_sweeps, uint32_t first_check,
                              uint32_t check_every, int use_model, float* out);
/* The same frame, with the target set to the energy that `ref` (a second
 * context of the same mesh) reaches after its full iteration count from
 * ctx's frame start, recording ctx's features and ref's correction as
 * training rows (rows from the states the warm-started solver visits). */
int lats_mlp_record_guided_frame(lats_mlp* m, lats_vbd* ctx, lats_vbd* ref, uint32_t max_sweeps, uint32_t first_check,
                                 uint32_t check_every, int use_model, float* out);

#ifdef __cplusplus
}
#endif
#endif /* LIBATSSIM_HIP_H */

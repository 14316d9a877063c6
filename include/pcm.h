/*
 * pcm.h — C-ABI of the MI355X (gfx950) mechanics-timestep library `libpcm.so`.
 *
 * This is the drop-in boundary for paraPhysiCell's mechanics solver (src/solver/).
 * Each entry point replaces one reference entry point; the replaced symbol is cited
 * beside it.  All pointers are plain host pointers, all sizes are plain integers, and
 * every function returns a status code (PCM_OK == 0).  pcm_last_error() gives the
 * message of the most recent failure on the calling thread.
 *
 * Data contract (mirrors src/cell_data.h:14-196):
 *   - agent fields are laid out exactly like the reference's std::vector members:
 *     one array per field, `count * width` elements, agent-major (positions of agent i
 *     at [i*dims, i*dims+dims)); 2D simulations store exactly two coordinates;
 *   - real-valued fields hold `real_bytes`-wide IEEE values (4 = float, 8 = double),
 *     index fields are int32 (biofvm::index_t), flags are uint8;
 *   - per-agent lists (springs, attached cells, neighbors) travel as CSR
 *     (offsets[count+1], indices[offsets[count]]).
 *
 * Residency: the handle owns device-resident state.  Host buffers are borrowed only for
 * the duration of a call.  A handle is not thread-safe; use one handle per thread.
 */
#ifndef PCM_H
#define PCM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes */
#define PCM_OK 0
#define PCM_ERR_ARGUMENT 1    /* bad argument (null pointer, size out of range, ...)     */
#define PCM_ERR_DEVICE 2      /* a HIP runtime call failed                                */
#define PCM_ERR_CAPACITY 3    /* agent capacity, list capacity or budget exceeded         */
#define PCM_ERR_UNSUPPORTED 4 /* configuration the device path does not implement        */
#define PCM_ERR_STATE 5       /* the device status word reports an aborted step           */

/* ---------------------------------------------------------------- field table */
/* element types */
#define PCM_T_REAL 0 /* real_bytes wide */
#define PCM_T_I32 1
#define PCM_T_U8 2
/* width kinds: elements per agent */
#define PCM_W_ONE 0  /* 1                      */
#define PCM_W_DIMS 1 /* mesh dims (1, 2 or 3)   */
#define PCM_W_DEFS 2 /* cell definitions count  */
#define PCM_W_SUBS 3 /* substrates count        */

/* X(name, element type, width kind) — order defines the field ids. */
#define PCM_FIELD_LIST(X)                                                                                              \
	X(POSITIONS, PCM_T_REAL, PCM_W_DIMS)                          /* agent_data.positions            */            \
	X(VELOCITIES, PCM_T_REAL, PCM_W_DIMS)                         /* velocities                      */            \
	X(PREVIOUS_VELOCITIES, PCM_T_REAL, PCM_W_DIMS)                /* previous_velocities             */            \
	X(RADIUS, PCM_T_REAL, PCM_W_ONE)                              /* geometries.radius               */            \
	X(NUCLEAR_RADIUS, PCM_T_REAL, PCM_W_ONE)                      /* geometries.nuclear_radius       */            \
	X(SURFACE_AREA, PCM_T_REAL, PCM_W_ONE)                        /* geometries.surface_area         */            \
	X(VOLUME, PCM_T_REAL, PCM_W_ONE)                              /* agent_data.volumes (total)      */            \
	X(FLUID, PCM_T_REAL, PCM_W_ONE)                               /* volumes.fluid                   */            \
	X(SOLID, PCM_T_REAL, PCM_W_ONE)                               /* volumes.solid                   */            \
	X(NUCLEAR, PCM_T_REAL, PCM_W_ONE)                             /* volumes.nuclear                 */            \
	X(NUCLEAR_FLUID, PCM_T_REAL, PCM_W_ONE)                       /* volumes.nuclear_fluid           */            \
	X(NUCLEAR_SOLID, PCM_T_REAL, PCM_W_ONE)                       /* volumes.nuclear_solid           */            \
	X(CYTOPLASMIC, PCM_T_REAL, PCM_W_ONE)                         /* volumes.cytoplasmic             */            \
	X(CYTOPLASMIC_FLUID, PCM_T_REAL, PCM_W_ONE)                   /* volumes.cytoplasmic_fluid       */            \
	X(CYTOPLASMIC_SOLID, PCM_T_REAL, PCM_W_ONE)                   /* volumes.cytoplasmic_solid       */            \
	X(FLUID_FRACTION, PCM_T_REAL, PCM_W_ONE)                      /* volumes.fluid_fraction          */            \
	X(CYTOPLASMIC_TO_NUCLEAR_RATIO, PCM_T_REAL, PCM_W_ONE)        /* volumes.cytoplasmic_to_nuclear_ratio */       \
	X(TARGET_SOLID_CYTOPLASMIC, PCM_T_REAL, PCM_W_ONE)            /* volumes.target_solid_cytoplasmic */           \
	X(TARGET_SOLID_NUCLEAR, PCM_T_REAL, PCM_W_ONE)                /* volumes.target_solid_nuclear    */            \
	X(CELL_CELL_ADHESION_STRENGTH, PCM_T_REAL, PCM_W_ONE)         /* mechanics.*                     */            \
	X(CELL_CELL_REPULSION_STRENGTH, PCM_T_REAL, PCM_W_ONE)                                                         \
	X(CELL_BM_REPULSION_STRENGTH, PCM_T_REAL, PCM_W_ONE)                                                           \
	X(RELATIVE_MAXIMUM_ADHESION_DISTANCE, PCM_T_REAL, PCM_W_ONE)                                                   \
	X(CELL_ADHESION_AFFINITIES, PCM_T_REAL, PCM_W_DEFS)                                                            \
	X(MAXIMUM_NUMBER_OF_ATTACHMENTS, PCM_T_I32, PCM_W_ONE)                                                         \
	X(ATTACHMENT_ELASTIC_CONSTANT, PCM_T_REAL, PCM_W_ONE)                                                          \
	X(ATTACHMENT_RATE, PCM_T_REAL, PCM_W_ONE)                                                                      \
	X(DETACHMENT_RATE, PCM_T_REAL, PCM_W_ONE)                                                                      \
	X(IS_MOTILE, PCM_T_U8, PCM_W_ONE)                             /* motilities.*                    */            \
	X(PERSISTENCE_TIME, PCM_T_REAL, PCM_W_ONE)                                                                     \
	X(MIGRATION_SPEED, PCM_T_REAL, PCM_W_ONE)                                                                      \
	X(MIGRATION_BIAS_DIRECTION, PCM_T_REAL, PCM_W_DIMS)                                                            \
	X(MIGRATION_BIAS, PCM_T_REAL, PCM_W_ONE)                                                                       \
	X(MOTILITY_VECTOR, PCM_T_REAL, PCM_W_DIMS)                                                                     \
	X(RESTRICT_TO_2D, PCM_T_U8, PCM_W_ONE)                                                                         \
	X(DEAD, PCM_T_U8, PCM_W_ONE)                                  /* deaths.dead                     */            \
	X(DEAD_PHAGOCYTOSIS_RATE, PCM_T_REAL, PCM_W_ONE)              /* interactions.*                  */            \
	X(LIVE_PHAGOCYTOSIS_RATES, PCM_T_REAL, PCM_W_DEFS)                                                             \
	X(DAMAGE_RATE, PCM_T_REAL, PCM_W_ONE)                                                                          \
	X(ATTACK_RATES, PCM_T_REAL, PCM_W_DEFS)                                                                        \
	X(IMMUNOGENICITIES, PCM_T_REAL, PCM_W_DEFS)                                                                    \
	X(FUSION_RATES, PCM_T_REAL, PCM_W_DEFS)                                                                        \
	X(SIMPLE_PRESSURE, PCM_T_REAL, PCM_W_ONE)                     /* states.*                        */            \
	X(NUMBER_OF_NUCLEI, PCM_T_I32, PCM_W_ONE)                                                                      \
	X(DAMAGE, PCM_T_REAL, PCM_W_ONE)                                                                               \
	X(TOTAL_ATTACK_TIME, PCM_T_REAL, PCM_W_ONE)                                                                    \
	X(CELL_DEFINITION_INDEX, PCM_T_I32, PCM_W_ONE)                /* cell_definition_indices         */            \
	X(IS_MOVABLE, PCM_T_U8, PCM_W_ONE)                            /* is_movable                      */            \
	X(FLAGS, PCM_T_U8, PCM_W_ONE)                                 /* flags (cell_state_flag)         */            \
	X(INTERNALIZED_SUBSTRATES, PCM_T_REAL, PCM_W_SUBS)            /* agent_data.internalized_substrates */         \
	X(FRACTION_TRANSFERRED_WHEN_INGESTED, PCM_T_REAL, PCM_W_SUBS) /* agent_data.fraction_transferred_... */        \
	X(ORIGIN, PCM_T_I32, PCM_W_ONE) /* index of the agent in the last host upload; moves with the agent */          \
	X(CONTACT_FUNCTION, PCM_T_U8, PCM_W_ONE) /* functions.contact_function as a PCM_CONTACT_* code */              \
	X(CHEMOTAXIS_INDEX, PCM_T_I32, PCM_W_ONE)                     /* motilities.chemotaxis_index     */            \
	X(CHEMOTAXIS_DIRECTION, PCM_T_I32, PCM_W_ONE)                 /* motilities.chemotaxis_direction */            \
	X(CHEMOTACTIC_SENSITIVITIES, PCM_T_REAL, PCM_W_SUBS)          /* motilities.chemotactic_sensitivities */       \
	X(MIGRATION_BIAS_DIRECTION_FUNCTION, PCM_T_U8, PCM_W_ONE) /* motilities.update_migration_bias_direction, PCM_BIAS_* */ \
	X(ORIENTATION, PCM_T_REAL, PCM_W_DIMS)                        /* states.orientation              */            \
	X(POLARITY, PCM_T_REAL, PCM_W_ONE)                            /* geometries.polarity             */

#define PCM_FIELD_ENUM_(name, type, width) PCM_F_##name,
enum pcm_field
{
	PCM_FIELD_LIST(PCM_FIELD_ENUM_) PCM_FIELD_COUNT
};
#undef PCM_FIELD_ENUM_

/* cell_state_flag values (src/cell_data.h:155-160) */
#define PCM_FLAG_NONE 0
#define PCM_FLAG_TO_REMOVE 1
#define PCM_FLAG_TO_DIVIDE 2

/* PCM_F_CONTACT_FUNCTION codes.  The reference keeps a std::function per cell
 * (src/cell_functions.h:26) that evaluate_interactions (src/simulator.cpp:65-82) calls for
 * every attached cell; the device knows the one model the reference ships. */
#define PCM_CONTACT_NONE 0             /* contact_function == nullptr */
#define PCM_CONTACT_STANDARD_ELASTIC 1 /* standard_elastic_contact_function (standard_models.cpp:643-654) */
#define PCM_CONTACT_OTHER 255          /* any other function: pcm_set_field refuses it (PCM_ERR_UNSUPPORTED),
										  so such a model keeps its host solver */

/* PCM_F_MIGRATION_BIAS_DIRECTION_FUNCTION codes: the direction update that
 * update_motility_single (position_solver.cpp:132-136) calls after a new random walk, as
 * builder.cpp:1184-1233 selects it.  The chemotaxis models read the substrate gradients of
 * pcm_set_gradients. */
#define PCM_BIAS_NONE 0                              /* no update function                     */
#define PCM_BIAS_CHEMOTAXIS 1                        /* chemotaxis_function (standard_models.cpp:542-555) */
#define PCM_BIAS_ADVANCED_CHEMOTAXIS 2               /* advanced_chemotaxis_function<false> (557-583) */
#define PCM_BIAS_ADVANCED_CHEMOTAXIS_NORMALIZED 3    /* advanced_chemotaxis_function<true>  (557-583) */

/* link lists */
#define PCM_LINKS_SPRINGS 0  /* states.springs        */
#define PCM_LINKS_ATTACHED 1 /* states.attached_cells */

/* ---------------------------------------------------------------- configuration */
typedef struct pcm_mesh_desc
{
	int32_t dims;        /* 1, 2 or 3 (microenvironment mesh dims)                         */
	int32_t bbox_min[3]; /* microenvironment bounding box, integer microns                  */
	int32_t bbox_max[3];
	int32_t voxel[3]; /* mechanics voxel shape (default {30,30,30}, src/environment.h:10)     */
} pcm_mesh_desc;

typedef struct pcm_config
{
	pcm_mesh_desc mesh;
	int32_t real_bytes;  /* 4 (FP32 real_t) or 8 (FP64 real_t)                          */
	int32_t cell_defs;   /* environment::cell_definitions_count                           */
	int32_t substrates;  /* microenvironment::substrates_count                             */
	int32_t capacity;    /* maximum agents the handle will hold                          */
	uint64_t seed;       /* key of the counter-based random stream                       */
} pcm_config;

/* The environment scalars a mechanics phase reads (src/environment.h:30-41). */
typedef struct pcm_step_params
{
	double mechanics_time_step;        /* environment::mechanics_time_step             */
	int32_t virtual_wall_at_domain_edges; /* environment::virtual_wall_at_domain_edges */
	int32_t automated_spring_adhesion; /* environment::automated_spring_adhesion        */
} pcm_step_params;

typedef struct pcm_sim pcm_sim; /* one simulation's device-resident mechanics state */

/* ---------------------------------------------------------------- handle lifetime */
/* solver::initialize (src/solver/solver.cpp:5) + containers_solver::initialize
 * (src/solver/host/containers_solver.cpp:10-13): allocates the device SoA, the mechanics
 * voxel grid and the list storage. */
int pcm_create(const pcm_config* config, pcm_sim** out);
int pcm_destroy(pcm_sim* sim);
const char* pcm_last_error(void);
/* Field metadata; needs no GPU. */
int pcm_field_info(int32_t field, int32_t* type, int32_t* width_kind, const char** name);
int pcm_version(void);

/* ---------------------------------------------------------------- data movement */
/* cell_data::agents_count.  Setting the count resets ORIGIN to the identity, empties the
 * spring, attached-cell and neighbor lists and clears CONTACT_FUNCTION of agents
 * [0, count); upload fields and lists after it. */
int pcm_set_count(pcm_sim* sim, int32_t count);
int pcm_get_count(pcm_sim* sim, int32_t* count);
/* One field, `count * width` elements, where count is the current agent count
 * (pcm_get_count: below the last pcm_set_count once the container update has removed
 * agents).  Setting reads, and getting writes, exactly that many elements. */
int pcm_set_field(pcm_sim* sim, int32_t field, const void* host);
int pcm_get_field(pcm_sim* sim, int32_t field, void* host);
/* pcm_set_field for `nfields` fields in one call (fields[k] read from hosts[k]): every
 * field is checked before any is copied, the copies are queued together and the call waits
 * once, so uploading a whole cell_data costs one round trip instead of one per field. */
int pcm_set_fields(pcm_sim* sim, int32_t nfields, const int32_t* fields, const void* const* hosts);
/* pcm_get_field for `nfields` fields in one call (fields[k] written to hosts[k]): one read
 * of the count, the copies queued together through page-locked staging, one wait (the
 * read-back of a finished request in a serving loop). */
int pcm_get_fields(pcm_sim* sim, int32_t nfields, const int32_t* fields, void* const* hosts);
/* springs / attached_cells as CSR (offsets has count+1 entries) */
int pcm_set_links(pcm_sim* sim, int32_t which, const int32_t* offsets, const int32_t* indices);
int pcm_get_links_size(pcm_sim* sim, int32_t which, int64_t* total);
int pcm_get_links(pcm_sim* sim, int32_t which, int32_t* offsets, int32_t* indices);
/* states.neighbors as built by update_cell_neighbors (empty for non-movable agents) */
int pcm_set_neighbors(pcm_sim* sim, const int32_t* offsets, const int32_t* indices);
int pcm_get_neighbors_size(pcm_sim* sim, int64_t* total);
int pcm_get_neighbors(pcm_sim* sim, int32_t* offsets, int32_t* indices);
/* environment::cells_in_mechanics_voxels: counts[voxel_count], agents concatenated in
 * voxel order (each bucket in the reference's bucket order). */
int pcm_get_voxel_count(pcm_sim* sim, int64_t* voxel_count);
int pcm_get_voxel_buckets(pcm_sim* sim, int32_t* counts, int32_t* agents);
/* environment::deaths_count increments made on the device since creation */
int pcm_get_deaths_count(pcm_sim* sim, int64_t* deaths);
/* Output packing, for the per-cell matrix of the MultiCellDS writer
 * (PhysiCell_MultiCellDS.cpp:1735-1790 writes each cell as a column of doubles): for every
 * agent, fields[0..nfields) in order, each as widths[k] doubles — the field's own elements,
 * zero-padded when widths[k] is larger (fill(v, dims) of the writer pads 2D vectors to 3),
 * widths[k] == 0 meaning the field's own width.  `out` receives count * row doubles,
 * agent-major, row = the sum of the widths; the packing runs on the device, one copy comes
 * back.  At most 256 columns. */
int pcm_pack_fields(pcm_sim* sim, int32_t nfields, const int32_t* fields, const int32_t* widths, double* out);
/* The removals made on the device since the last call, as the sequence of positions the
 * reference's container.remove_at(i) received (containers_solver.cpp:81-123): replaying
 * remove_at on a host container in this order reproduces the device's agent order, so host-
 * only SoA members follow.  With positions == NULL only the length is returned; otherwise
 * the log is copied (capacity >= length) and restarts. */
int pcm_removal_log(pcm_sim* sim, int32_t* positions, int64_t capacity, int64_t* length);
/* The substrate gradients that cell::nearest_gradient reads (paraBioFVM's gradient solver,
 * run at simulator.cpp:163), for the chemotaxis direction updates.  `diffusion_voxel` is
 * the microenvironment's voxel shape; the diffusion grid spans the mesh bounding box with
 * ceil(extent / voxel) voxels per used dimension, x fastest.  `gradients` holds
 * voxels * substrates * dims reals, index ((voxel * substrates) + substrate) * dims + d.
 * The device keeps its own copy until the next call. */
int pcm_set_gradients(pcm_sim* sim, const int32_t diffusion_voxel[3], const void* gradients);
/* mechanics-step counter that keys the random stream (incremented by
 * pcm_update_cell_forces and by every pcm_step iteration) */
int pcm_set_step_index(pcm_sim* sim, uint64_t step);
int pcm_get_step_index(pcm_sim* sim, uint64_t* step);

/* ---------------------------------------------------------------- mechanics-mesh lookups (host) */
/* common_solver::get_mesh_position (src/solver/host/common_solver.cpp:18-38): the mechanics
 * voxel coordinates of a position of mesh->dims reals (real_bytes wide, host memory), by
 * the convention the device bins with: (int)((p - bbox_min) / voxel) in real_t, clamped to
 * the grid; coordinates at and beyond dims are 0.  Needs no GPU. */
int pcm_get_mesh_position(const pcm_mesh_desc* mesh, int32_t real_bytes, const void* position, int32_t voxel[3]);
/* common_solver::get_mesh_index (common_solver.cpp:10-16): the linear index
 * x + gx * (y + gy * z) of voxel coordinates, the index of cells_in_mechanics_voxels and of
 * pcm_get_voxel_buckets.  Needs no GPU. */
int pcm_get_mesh_index(const pcm_mesh_desc* mesh, const int32_t voxel[3], int32_t* index);

/* ---------------------------------------------------------------- reference phases */
/* containers_solver::update_mechanics_mesh (src/solver/host/containers_solver.cpp:15-56) */
int pcm_update_mechanics_mesh(pcm_sim* sim);
/* position_solver::update_cell_neighbors (src/solver/host/position_solver.cpp:262-286) */
int pcm_update_cell_neighbors(pcm_sim* sim);
/* evaluate_interactions (src/simulator.cpp:65-82): every agent whose contact function is
 * PCM_CONTACT_STANDARD_ELASTIC adds sqrt(k_i k_j a_i[t_j] a_j[t_i]) (x_j - x_i) to its
 * velocity for each attached cell j, in list order (standard_models.cpp:643-654). */
int pcm_evaluate_interactions(pcm_sim* sim, const pcm_step_params* params);
/* position_solver::update_cell_forces (src/solver/host/position_solver.cpp:233-260) */
int pcm_update_cell_forces(pcm_sim* sim, const pcm_step_params* params);
/* position_solver::update_motility (src/solver/host/position_solver.cpp:288-314) */
int pcm_update_motility(pcm_sim* sim, const pcm_step_params* params);
/* position_solver::update_basement_membrane_interactions (position_solver.cpp:316-335) */
int pcm_update_basement_membrane_interactions(pcm_sim* sim, const pcm_step_params* params);
/* position_solver::update_spring_attachments (src/solver/host/position_solver.cpp:451-479) */
int pcm_update_spring_attachments(pcm_sim* sim, const pcm_step_params* params);
/* position_solver::update_positions (src/solver/host/position_solver.cpp:506-520) */
int pcm_update_positions(pcm_sim* sim, const pcm_step_params* params);
/* interactions_solver::update_cell_cell_interactions (interactions_solver.cpp:255-265) */
int pcm_update_cell_cell_interactions(pcm_sim* sim, const pcm_step_params* params);
/* containers_solver::update_cell_container_for_mechanics (containers_solver.cpp:140-147) */
int pcm_update_cell_container_for_mechanics(pcm_sim* sim, const pcm_step_params* params);
/* containers_solver::update_cell_container_for_phenotype (containers_solver.cpp:149-187) in
 * the reference's single-thread order: every agent of [0, count) flagged
 * PCM_FLAG_TO_DIVIDE divides (cell::divide, cell.cpp:108-163) in ascending index order, its
 * daughter appended at count + rank; then the removal loop visits [0, count) (removals and
 * the removal log as in pcm_update_cell_container_for_mechanics).  Daughters copy the
 * parent's phenotype fields (phenotype.cpp:18-30, after the parent's volume halving) and
 * start with zero velocities, pressure and damage, one nucleus, movable, ORIGIN -1 and a
 * random ORIENTATION (cell::assign_orientation without set_orientation).  The parent's
 * INTERNALIZED_SUBSTRATES halve when compute_internalized_substrates != 0 (the
 * microenvironment option of cell.cpp:120).  release_internalized_substrates and custom
 * data (paraBioFVM, host members) are the caller's: see pcm_division_log.  As in
 * simulate_phenotype (simulator.cpp:208-216), pcm_update_mechanics_mesh and
 * pcm_update_cell_neighbors follow before the next mechanics step: the neighbor lists
 * still name the agents of before. */
int pcm_update_cell_container_for_phenotype(pcm_sim* sim, const pcm_step_params* params,
											int32_t compute_internalized_substrates);
/* environment::divisions_count increments made on the device */
int pcm_get_divisions_count(pcm_sim* sim, int64_t* divisions);
/* The parents of the last pcm_update_cell_container_for_phenotype's divisions, in division
 * order; division r created the agent at (count before the call) + r.  With parents == NULL
 * only the length is returned. */
int pcm_division_log(pcm_sim* sim, int32_t* parents, int64_t capacity, int64_t* length);

/* The mechanics step of simulator::simulate_mechanics (src/simulator.cpp:171-192), run
 * `steps` times without host round trips: contact functions, forces, motility, basement
 * membrane, springs, positions, cell-cell interactions, container compaction, mechanics
 * mesh, neighbors. */
int pcm_step(pcm_sim* sim, const pcm_step_params* params, int32_t steps);
/* Wait for queued device work; reports PCM_ERR_STATE if a step aborted. */
int pcm_synchronize(pcm_sim* sim);

/* ---------------------------------------------------------------- launch shape
 * No counterpart in the reference (whose loops have no launch shape).  The lane-parallel
 * kernels of the step can run with several shapes, all giving bit-identical results:
 * lanes per agent of the force/velocity kernel and of the attachment and interaction
 * kernels (4, 8 or 16), and the neighbor search as 1, 2 or 4 wavefronts per occupied voxel
 * sharing an LDS copy of its stencil.  Until a shape is set or tuned, each pcm_step takes
 * {4, 4, 2} from 65536 agents up and {8, 8, 2} below; pcm_get_tuning reports the shape in
 * use.  A batch group's default is {8, 8, 2}. */
typedef struct pcm_tuning
{
	int32_t velocity_lanes;
	int32_t event_lanes;
	int32_t neighbor_waves;
} pcm_tuning;
int pcm_set_tuning(pcm_sim* sim, const pcm_tuning* tuning);
int pcm_get_tuning(pcm_sim* sim, pcm_tuning* tuning);
/* Autotuner: times each candidate shape on the simulation's own steps (results are the
 * same for every shape), twice in opposite orders, and keeps, per kernel family, the
 * fastest.  Advances the simulation by 1 + 6 * steps_per_candidate steps; writes the
 * chosen shape if `chosen`. */
int pcm_autotune(pcm_sim* sim, const pcm_step_params* params, int32_t steps_per_candidate, pcm_tuning* chosen);

/* ---------------------------------------------------------------- ensembles
 * No counterpart in the reference, which advances one simulation per process: a batch
 * advances many simulations together for ensemble and serving workloads.  Members of
 * one dimensionality share every kernel launch of the step (one launch covers all of
 * them, simulations of different sizes included); the 1D, 2D and 3D groups run on
 * separate streams.  A member keeps its handle and device state: its own calls keep
 * working (they run on the group's stream) and pcm_destroy retires it first.  Results
 * are the same as stepping each simulation alone. */
typedef struct pcm_batch pcm_batch;
int pcm_batch_create(int32_t real_bytes, pcm_batch** out);
int pcm_batch_destroy(pcm_batch* batch);
int pcm_batch_admit(pcm_batch* batch, pcm_sim* sim);
int pcm_batch_retire(pcm_batch* batch, pcm_sim* sim);
int pcm_batch_size(pcm_batch* batch, int32_t* sims);
/* Every member advances `steps` mechanics steps.  With params != NULL all members step under
 * those parameters (and keep them); with params == NULL each member steps under its own:
 * those of its last pcm_set_step_params, pcm_step or phase call. */
int pcm_batch_step(pcm_batch* batch, const pcm_step_params* params, int32_t steps);
/* The environment scalars a simulation steps under when its batch is stepped with
 * params == NULL (environment::mechanics_time_step, virtual_wall_at_domain_edges,
 * automated_spring_adhesion: each simulation of the reference owns its environment). */
int pcm_set_step_params(pcm_sim* sim, const pcm_step_params* params);
int pcm_batch_synchronize(pcm_batch* batch);
/* enable != 0: each group's step is captured once as a hipGraph over all member slots and
 * replayed per step (joining and leaving rewrite slots without a new capture).  A capture
 * happens inside pcm_batch_step when none matches; no other host thread may make
 * synchronous device calls (pcm_create, uploads, reads) meanwhile, since those invalidate
 * it: disable graphs for steps that overlap such work. */
int pcm_batch_use_graphs(pcm_batch* batch, int32_t enable);
/* enable != 0 (the default): each kernel of a group's step is one launch over exactly the
 * blocks its members need, back to back (a block map gives each block its member), instead
 * of one block row per member sized for the largest; results are the same either way. */
int pcm_batch_flatten(pcm_batch* batch, int32_t enable);
/* pcm_autotune for every dimensionality group of the batch, on the batch's own steps (one
 * shape per group, used by its segmented launches); members advance 1 + 6 * steps. */
int pcm_batch_autotune(pcm_batch* batch, const pcm_step_params* params, int32_t steps_per_candidate);
/* The launch shape of the batch's `dims` group (1, 2 or 3), as pcm_set_tuning sets it for
 * one simulation: kept for members that join later, so a server can apply a shape tuned
 * earlier on the same kind of workload instead of tuning on live requests. */
int pcm_batch_set_tuning(pcm_batch* batch, int32_t dims, const pcm_tuning* tuning);
int pcm_batch_get_tuning(pcm_batch* batch, int32_t dims, pcm_tuning* tuning);
/* The agent count of each of `count` members as the batch's last pcm_batch_step left it
 * (every step call gathers all members' scalars at its end): no device access and no wait,
 * unlike pcm_get_count.  Fails with PCM_ERR_ARGUMENT for a simulation that is not a member. */
int pcm_batch_counts(pcm_batch* batch, int32_t count, pcm_sim* const* sims, int32_t* counts);

/* ---------------------------------------------------------------- kernel timing */
/* Counterpart of the reference's per-phase timers (simulator.cpp:14-29,
 * simulator_measurer.h:5-33), at kernel granularity: while a profile is open, every launch
 * of a kernel whose bit is set in `kernel_mask` is bracketed by HIP events on the handle's
 * stream; pcm_profile_end waits for the work and returns, per kernel id, the summed event
 * time in milliseconds and the launch count (arrays of PCM_KERNEL_COUNT entries).
 * INTERACT_CHECK is no longer launched (the interaction pass detects its own conflicts);
 * the id is kept so that the other ids do not change. */
#define PCM_KERNEL_LIST(X)                                                                                             \
	X(BEGIN_STEP) X(BIN_CLEAR) X(BIN_COUNT) X(BIN_CLAIM) X(BIN_SCATTER) X(BIN_ORDER) X(NEIGHBORS) X(FORCES)              \
	X(MOTILITY) X(MEMBRANE) X(CONTRACT) X(POSITIONS) X(VELOCITY_FUSED) X(SPRING_FIRE) X(SPRING_ERASE)                   \
	X(SPRING_COMPACT) X(SPRING_COLLECT) X(SPRING_RESOLVE) X(SPRING_APPEND) X(INTERACT_TENTATIVE) X(INTERACT_CHECK)       \
	X(INTERACT_APPLY) X(INTERACT_SERIAL) X(COMPACT_FLAG) X(COMPACT_PLAN) X(COMPACT_APPLY) X(COMPACT_RESTORE)            \
	X(COMPACT_DONE) X(FORCE_RECORDS) X(CONTACT) X(DIVIDE_PLAN) X(DIVIDE_APPLY) X(DIVIDE_LINKS) X(PACK)             \
	X(SPRING_RESOLVE_FIRST)
#define PCM_KERNEL_ENUM_(name) PCM_K_##name,
enum pcm_kernel
{
	PCM_KERNEL_LIST(PCM_KERNEL_ENUM_) PCM_KERNEL_COUNT
};
#undef PCM_KERNEL_ENUM_
int pcm_kernel_name(int32_t kernel, const char** name);
/* Work counters of the most recent step, 8 entries: spring attachment events, resolution
 * rounds, attaching agents, interaction candidates, interaction replay used (0/1), agents
 * removed, occupied mechanics voxels, reserved (0). */
int pcm_get_stats(pcm_sim* sim, int32_t* stats);
int pcm_profile_begin(pcm_sim* sim, uint64_t kernel_mask);
/* Time one launch in `every` of each selected kernel (default 1): the event pair around a
 * launch costs the stream a few microseconds, so a sampled profile disturbs the throughput
 * it is taken from less.  pcm_profile_end reports the timed launches only. */
int pcm_profile_every(pcm_sim* sim, int32_t every);
int pcm_profile_end(pcm_sim* sim, double* total_ms, int64_t* launches);

/* ---------------------------------------------------------------- allocation-free neighbor build
 * No counterpart in the reference.  For callers that cannot allocate device memory: code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.  It is
 * pcm_update_mechanics_mesh followed by pcm_update_cell_neighbors (containers_solver.cpp:15-56,
 * position_solver.cpp:262-286) on caller-owned device arrays, with no handle: the mesh
 * scratch (voxel counts and starts, buckets, the voxel-sorted copy) lives in one fixed
 * library area of PCM_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.  The call
 * allocates nothing, copies nothing between host and device and does not wait: it enqueues
 * kernels on `stream` (a hipStream_t; NULL is the null stream) and returns.
 *   positions     count * mesh->dims reals (real_bytes wide), device memory;
 *   radius, relative_maximum_adhesion_distance   count reals each, device memory;
 *   is_movable    count bytes, device memory, or NULL (every agent movable);
 *   neighbor_counts[count], neighbor_indices[count * max_neighbors]   device outputs: row i
 *                 holds agent i's list in the reference's order; non-movable agents get an
 *                 empty list, as in pcm_get_neighbors;
 *   needed[1]     device output: 0 when every list fit in max_neighbors, else the longest
 *                 list, and rows then hold their first max_neighbors entries.
 * The call needs at most 8 KiB + 8 * voxels + (20 + 4 * real_bytes) * count bytes of the
 * area (about 22 MB for 88,700 3D agents on a 134^3 grid); a larger call fails with PCM_ERR_CAPACITY
 * before enqueueing anything.  Every call on one GPU shares that GPU's area and leaves its
 * contents behind, so calls that overlap on one GPU (two streams, two threads) race. */
#define PCM_NOALLOC_SCRATCH_BYTES (96u << 20)
int pcm_neighbors_noalloc(const pcm_mesh_desc* mesh, int32_t real_bytes, int32_t count, const void* positions,
						  const void* radius, const void* relative_maximum_adhesion_distance, const uint8_t* is_movable,
						  int32_t max_neighbors, int32_t* neighbor_counts, int32_t* neighbor_indices, int32_t* needed,
						  void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PCM_H */

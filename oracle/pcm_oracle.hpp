// pcm_oracle.hpp — ORACLE: test infrastructure only.
//
// A CPU restatement of paraPhysiCell's mechanics timestep (src/solver/host/*), written
// for this repository so that results of the MI355X library can be checked.  Only the
// tests, the benchmark's cpu_baseline leg and the smoke check may link, call or run it;
// the product library never does.
//
// It follows the reference solver phase by phase (each function cites the file:line it
// restates) and keeps the reference's floating-point operation order, its draw order and
// its OpenMP structure (orphaned `omp for` worksharing inside one enclosing parallel
// region, atomics for the voxel histogram, critical sections / mutexes for springs,
// interactions and removal).  Called outside a parallel region it runs single-threaded,
// which is the canonical "reference order".
//
// The reference needs paraBioFVM (absent: empty submodule) for a handful of symbols.  The
// stand-ins below supply them:
//   - cartesian_mesh grid shape: ceil((max - min) / voxel) per used dimension, 1 otherwise;
//   - cartesian_mesh::voxel_position: (index_t)((p - min) / voxel) in real_t, clamped to the
//     grid;                                                    (parity unpinned, see header)
//   - real_t = float or double, index_t = int32;               (parity unpinned)
//   - agent_data::remove / remove_at: swap-with-last, pinned by test/test.cpp:52-115.
//
// Random numbers: two modes.  MT mode draws from a thread_local std::mt19937 through
// std::uniform_real_distribution exactly as src/random.cpp:8-49 (single-thread stream of
// the reference).  KEYED mode draws each number from the counter-based stream of
// paraphysicell_amd/csrc/pcm_rng.h keyed by (step, stream, agent, slot, draw); that is the
// mode the device library is compared in.
#pragma once

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <numbers>
#include <random>
#include <vector>

#ifdef _OPENMP
	#include <omp.h>
#endif

#include "../include/pcm.h"
#include "../paraphysicell_amd/csrc/pcm_rng.h"

namespace pcm_oracle {

using index_t = std::int32_t;

enum class cell_state_flag : std::uint8_t
{
	none = 0,
	to_remove = 1,
	to_divide = 2
};

// ------------------------------------------------------------------ mesh stand-in
struct cartesian_mesh
{
	index_t dims = 3;
	index_t bounding_box_mins[3] = { 0, 0, 0 };
	index_t bounding_box_maxs[3] = { 0, 0, 0 };
	index_t voxel_shape[3] = { 30, 30, 30 };
	index_t grid_shape[3] = { 1, 1, 1 };

	cartesian_mesh() = default;
	cartesian_mesh(index_t d, const index_t* mins, const index_t* maxs, const index_t* voxel) : dims(d)
	{
		for (int k = 0; k < 3; k++)
		{
			bounding_box_mins[k] = mins[k];
			bounding_box_maxs[k] = maxs[k];
			voxel_shape[k] = voxel[k];
			if (k < d)
			{
				const index_t extent = maxs[k] - mins[k];
				grid_shape[k] = std::max<index_t>(1, (extent + voxel[k] - 1) / voxel[k]);
			}
			else
				grid_shape[k] = 1;
		}
	}

	std::size_t voxel_count() const { return (std::size_t)grid_shape[0] * grid_shape[1] * grid_shape[2]; }

	template <class R>
	void voxel_position(const R* p3, index_t out[3]) const
	{
		for (int k = 0; k < 3; k++)
		{
			if (k >= dims)
			{
				out[k] = 0;
				continue;
			}
			const R q = (p3[k] - (R)bounding_box_mins[k]) / (R)voxel_shape[k];
			index_t c = (index_t)q;
			c = c < 0 ? 0 : (c >= grid_shape[k] ? grid_shape[k] - 1 : c);
			out[k] = c;
		}
	}

	// voxel centre, as used by test/test.cpp:361 to place agents
	template <class R>
	void voxel_center(const index_t c[3], R out[3]) const
	{
		for (int k = 0; k < 3; k++)
			out[k] = (R)bounding_box_mins[k] + (R)voxel_shape[k] * (R)c[k] + (R)voxel_shape[k] / (R)2;
	}
};

// ------------------------------------------------------------------ SoA (src/cell_data.h)
template <class R>
struct agent_data_t
{
	index_t agents_count = 0;
	std::vector<R> positions, volumes, internalized_substrates, fraction_transferred_when_ingested;
};

template <class R>
struct volume_data_t
{
	std::vector<R> solid, fluid, fluid_fraction, nuclear, nuclear_fluid, nuclear_solid, cytoplasmic,
		cytoplasmic_fluid, cytoplasmic_solid, cytoplasmic_to_nuclear_ratio, target_solid_cytoplasmic,
		target_solid_nuclear;
};

template <class R>
struct geometry_data_t
{
	std::vector<R> radius, nuclear_radius, surface_area, polarity;
};

template <class R>
struct mechanics_data_t
{
	std::vector<R> cell_cell_adhesion_strength, cell_cell_repulsion_strength, cell_BM_repulsion_strength,
		cell_adhesion_affinities, relative_maximum_adhesion_distance;
	std::vector<index_t> maximum_number_of_attachments;
	std::vector<R> attachment_elastic_constant, attachment_rate, detachment_rate;
};

template <class R>
struct motility_data_t
{
	std::vector<std::uint8_t> is_motile;
	std::vector<R> persistence_time, migration_speed, migration_bias_direction, migration_bias, motility_vector;
	std::vector<std::uint8_t> restrict_to_2d;
	std::vector<index_t> chemotaxis_index, chemotaxis_direction; // cell_data.h:96-97
	std::vector<R> chemotactic_sensitivities;                     // cell_data.h:98
	// cell_data.h:100 update_migration_bias_direction, as a PCM_BIAS_* code per agent
	std::vector<std::uint8_t> bias_direction_function;
};

struct death_data_t
{
	std::vector<std::uint8_t> dead;
};

template <class R>
struct interactions_data_t
{
	std::vector<R> dead_phagocytosis_rate, live_phagocytosis_rates, damage_rate, attack_rates, immunogenicities,
		fusion_rates;
};

template <class R>
struct cell_state_data_t
{
	std::vector<std::vector<index_t>> neighbors, springs, attached_cells;
	std::vector<R> orientation; // cell_data.h:144
	std::vector<R> simple_pressure;
	std::vector<index_t> number_of_nuclei;
	std::vector<R> damage, total_attack_time;
};

template <class R>
struct cell_data
{
	agent_data_t<R> agent_data;
	volume_data_t<R> volumes;
	geometry_data_t<R> geometries;
	mechanics_data_t<R> mechanics;
	motility_data_t<R> motilities;
	death_data_t deaths;
	interactions_data_t<R> interactions;
	cell_state_data_t<R> states;
	std::vector<R> previous_velocities, velocities;
	std::vector<index_t> cell_definition_indices;
	std::vector<std::uint8_t> is_movable;
	std::vector<cell_state_flag> flags;
	std::vector<index_t> origin; // identity in the last upload (bookkeeping of the harness)
	// cell_functions.h:26 contact_function, as a PCM_CONTACT_* code per agent
	std::vector<std::uint8_t> contact_function;

	index_t& agents_count = agent_data.agents_count;

	cell_data() = default;
	cell_data(const cell_data&) = delete;
};

// Visits every per-agent array with its element width; the order is the pcm.h field order.
template <class R, class F>
void for_each_field(cell_data<R>& d, index_t dims, index_t defs, index_t subs, F&& f)
{
	f(PCM_F_POSITIONS, d.agent_data.positions, dims);
	f(PCM_F_VELOCITIES, d.velocities, dims);
	f(PCM_F_PREVIOUS_VELOCITIES, d.previous_velocities, dims);
	f(PCM_F_RADIUS, d.geometries.radius, 1);
	f(PCM_F_NUCLEAR_RADIUS, d.geometries.nuclear_radius, 1);
	f(PCM_F_SURFACE_AREA, d.geometries.surface_area, 1);
	f(PCM_F_VOLUME, d.agent_data.volumes, 1);
	f(PCM_F_FLUID, d.volumes.fluid, 1);
	f(PCM_F_SOLID, d.volumes.solid, 1);
	f(PCM_F_NUCLEAR, d.volumes.nuclear, 1);
	f(PCM_F_NUCLEAR_FLUID, d.volumes.nuclear_fluid, 1);
	f(PCM_F_NUCLEAR_SOLID, d.volumes.nuclear_solid, 1);
	f(PCM_F_CYTOPLASMIC, d.volumes.cytoplasmic, 1);
	f(PCM_F_CYTOPLASMIC_FLUID, d.volumes.cytoplasmic_fluid, 1);
	f(PCM_F_CYTOPLASMIC_SOLID, d.volumes.cytoplasmic_solid, 1);
	f(PCM_F_FLUID_FRACTION, d.volumes.fluid_fraction, 1);
	f(PCM_F_CYTOPLASMIC_TO_NUCLEAR_RATIO, d.volumes.cytoplasmic_to_nuclear_ratio, 1);
	f(PCM_F_TARGET_SOLID_CYTOPLASMIC, d.volumes.target_solid_cytoplasmic, 1);
	f(PCM_F_TARGET_SOLID_NUCLEAR, d.volumes.target_solid_nuclear, 1);
	f(PCM_F_CELL_CELL_ADHESION_STRENGTH, d.mechanics.cell_cell_adhesion_strength, 1);
	f(PCM_F_CELL_CELL_REPULSION_STRENGTH, d.mechanics.cell_cell_repulsion_strength, 1);
	f(PCM_F_CELL_BM_REPULSION_STRENGTH, d.mechanics.cell_BM_repulsion_strength, 1);
	f(PCM_F_RELATIVE_MAXIMUM_ADHESION_DISTANCE, d.mechanics.relative_maximum_adhesion_distance, 1);
	f(PCM_F_CELL_ADHESION_AFFINITIES, d.mechanics.cell_adhesion_affinities, defs);
	f(PCM_F_MAXIMUM_NUMBER_OF_ATTACHMENTS, d.mechanics.maximum_number_of_attachments, 1);
	f(PCM_F_ATTACHMENT_ELASTIC_CONSTANT, d.mechanics.attachment_elastic_constant, 1);
	f(PCM_F_ATTACHMENT_RATE, d.mechanics.attachment_rate, 1);
	f(PCM_F_DETACHMENT_RATE, d.mechanics.detachment_rate, 1);
	f(PCM_F_IS_MOTILE, d.motilities.is_motile, 1);
	f(PCM_F_PERSISTENCE_TIME, d.motilities.persistence_time, 1);
	f(PCM_F_MIGRATION_SPEED, d.motilities.migration_speed, 1);
	f(PCM_F_MIGRATION_BIAS_DIRECTION, d.motilities.migration_bias_direction, dims);
	f(PCM_F_MIGRATION_BIAS, d.motilities.migration_bias, 1);
	f(PCM_F_MOTILITY_VECTOR, d.motilities.motility_vector, dims);
	f(PCM_F_RESTRICT_TO_2D, d.motilities.restrict_to_2d, 1);
	f(PCM_F_DEAD, d.deaths.dead, 1);
	f(PCM_F_DEAD_PHAGOCYTOSIS_RATE, d.interactions.dead_phagocytosis_rate, 1);
	f(PCM_F_LIVE_PHAGOCYTOSIS_RATES, d.interactions.live_phagocytosis_rates, defs);
	f(PCM_F_DAMAGE_RATE, d.interactions.damage_rate, 1);
	f(PCM_F_ATTACK_RATES, d.interactions.attack_rates, defs);
	f(PCM_F_IMMUNOGENICITIES, d.interactions.immunogenicities, defs);
	f(PCM_F_FUSION_RATES, d.interactions.fusion_rates, defs);
	f(PCM_F_SIMPLE_PRESSURE, d.states.simple_pressure, 1);
	f(PCM_F_NUMBER_OF_NUCLEI, d.states.number_of_nuclei, 1);
	f(PCM_F_DAMAGE, d.states.damage, 1);
	f(PCM_F_TOTAL_ATTACK_TIME, d.states.total_attack_time, 1);
	f(PCM_F_CELL_DEFINITION_INDEX, d.cell_definition_indices, 1);
	f(PCM_F_IS_MOVABLE, d.is_movable, 1);
	f(PCM_F_FLAGS, d.flags, 1);
	f(PCM_F_INTERNALIZED_SUBSTRATES, d.agent_data.internalized_substrates, subs);
	f(PCM_F_FRACTION_TRANSFERRED_WHEN_INGESTED, d.agent_data.fraction_transferred_when_ingested, subs);
	f(PCM_F_ORIGIN, d.origin, 1);
	f(PCM_F_CONTACT_FUNCTION, d.contact_function, 1);
	f(PCM_F_CHEMOTAXIS_INDEX, d.motilities.chemotaxis_index, 1);
	f(PCM_F_CHEMOTAXIS_DIRECTION, d.motilities.chemotaxis_direction, 1);
	f(PCM_F_CHEMOTACTIC_SENSITIVITIES, d.motilities.chemotactic_sensitivities, subs);
	f(PCM_F_MIGRATION_BIAS_DIRECTION_FUNCTION, d.motilities.bias_direction_function, 1);
	f(PCM_F_ORIENTATION, d.states.orientation, dims);
	f(PCM_F_POLARITY, d.geometries.polarity, 1);
}

// ------------------------------------------------------------------ random service
enum class rng_mode
{
	keyed = 0, // counter-based stream shared with the device library
	mt = 1     // thread_local std::mt19937, src/random.cpp:8-49
};

struct random_service
{
	rng_mode mode = rng_mode::keyed;
	std::uint64_t seed = 0;
	std::uint64_t step = 0;

	static std::mt19937& generator()
	{
		static thread_local std::mt19937 g;
		return g;
	}

	// src/random.cpp:27-46 for one thread: seed_seq over {seed + thread id}.
	void seed_mt(unsigned int s)
	{
		seed = s;
		auto body = [s]() {
#ifdef _OPENMP
			const unsigned tid = (unsigned)omp_get_thread_num();
#else
			const unsigned tid = 0;
#endif
			std::vector<unsigned> init { s + tid };
			std::seed_seq seq(init.begin(), init.end());
			std::vector<unsigned> out(1);
			seq.generate(out.begin(), out.end());
			generator().seed(out[0]);
		};
#ifdef _OPENMP
		if (omp_in_parallel())
			body();
		else
	#pragma omp parallel
			body();
#else
		body();
#endif
	}

	// One uniform(lo, hi) draw.  In keyed mode (stream, agent, slot, draw) name the number.
	template <class R>
	R uniform(pcm_stream stream, index_t agent, index_t slot, std::uint32_t draw, R lo = 0, R hi = 1)
	{
		return uniform_at<R>(step, stream, agent, slot, draw, lo, hi);
	}

	// The same with an explicit counter in place of the mechanics step.
	template <class R>
	R uniform_at(std::uint64_t counter, pcm_stream stream, index_t agent, index_t slot, std::uint32_t draw, R lo = 0,
				 R hi = 1)
	{
		if (mode == rng_mode::mt)
		{
			std::uniform_real_distribution<R> distribution(lo, hi);
			return distribution(generator());
		}
		const R u = pcm_uniform01<R>(seed, counter, stream, (std::uint32_t)agent, (std::uint32_t)slot, draw);
		return pcm_scale_uniform<R>(u, lo, hi);
	}
};

// ------------------------------------------------------------------ environment stand-in
// Field names follow src/environment.h:12-60 so the drop-in shim compiles against it.
template <class R>
struct environment
{
	struct microenvironment_t
	{
		cartesian_mesh mesh;
		index_t substrates_count = 1;
		// the gradient solver's output (simulator.cpp:163): per diffusion voxel, substrate and
		// dimension, ((voxel * substrates) + substrate) * dims + d
		std::vector<R> gradients;

		// cell::nearest_gradient(s) (cell.h, paraBioFVM's agent): the gradient of substrate s at
		// the diffusion voxel that contains the position
		const R* nearest_gradient(const R* position, index_t s) const
		{
			R p3[3] = { 0, 0, 0 };
			for (index_t k = 0; k < mesh.dims; k++)
				p3[k] = position[k];
			index_t c[3];
			mesh.voxel_position(p3, c);
			const std::size_t v = (std::size_t)c[0] + mesh.grid_shape[0] * ((std::size_t)c[1] + mesh.grid_shape[1] * c[2]);
			return gradients.data() + (v * substrates_count + s) * mesh.dims;
		}
	} m;

	bool virtual_wall_at_domain_edges = false;
	bool automated_spring_adhesion = true;
	index_t divisions_count = 0;
	index_t deaths_count = 0;
	// calls of update_cell_container_for_phenotype: keys the division draws
	std::uint64_t division_epoch = 0;
	// parents of the last phenotype update's divisions, in division order
	std::vector<index_t> division_log;
	cartesian_mesh mechanics_mesh;
	R mechanics_time_step = (R)0.1;
	index_t cell_definitions_count = 1;
	std::vector<std::vector<index_t>> cells_in_mechanics_voxels;

	cell_data<R> data;
	random_service rng;

	environment(const pcm_mesh_desc& md, index_t defs, index_t subs)
	{
		const index_t diffusion_voxel[3] = { 20, 20, 20 };
		m.mesh = cartesian_mesh(md.dims, md.bbox_min, md.bbox_max, diffusion_voxel);
		m.substrates_count = subs;
		mechanics_mesh = cartesian_mesh(md.dims, md.bbox_min, md.bbox_max, md.voxel); // environment.cpp:16
		cell_definitions_count = defs;
		cells_in_mechanics_voxels.assign(mechanics_mesh.voxel_count(), {});
	}

	void resize(index_t n)
	{
		const index_t dims = m.mesh.dims;
		data.agents_count = n;
		for_each_field(data, dims, cell_definitions_count, m.substrates_count,
					   [n](int, auto& v, index_t w) { v.resize((std::size_t)n * w); });
		data.states.neighbors.resize(n);
		data.states.springs.resize(n);
		data.states.attached_cells.resize(n);
		for (index_t i = 0; i < n; i++)
			data.origin[i] = i;
	}

	// cell_data::remove (src/cell_data.cpp:285-313): the last agent moves into `index`.
	void remove_at(index_t index)
	{
		auto& d = data;
		d.agents_count--;
		const index_t last = d.agents_count;
		if (index == last)
			return;
		for_each_field(d, m.mesh.dims, cell_definitions_count, m.substrates_count, [index, last](int, auto& v, index_t w) {
			for (index_t k = 0; k < w; k++)
				v[(std::size_t)index * w + k] = v[(std::size_t)last * w + k];
		});
		d.states.neighbors[index] = std::move(d.states.neighbors[last]);
		d.states.springs[index] = std::move(d.states.springs[last]);
		d.states.attached_cells[index] = std::move(d.states.attached_cells[last]);
	}
};

// ------------------------------------------------------------------ helpers (solver_helper.h)
template <class R>
constexpr R zero_threshold = (R)1e-16; // solver_helper.h:12

template <class R>
constexpr R simple_pressure_coefficient = (R)36.64504274775163; // position_solver.cpp:13

// position_helper<d>::distance / difference_and_distance (solver_helper.h:54-67,127-141,212-228)
template <int D, class R>
inline R pair_distance(const R* a, const R* b)
{
	if constexpr (D == 1)
		return std::abs(a[0] - b[0]);
	else if constexpr (D == 2)
		return std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]));
	else
		return std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
}

template <int D, class R>
inline R pair_difference(const R* a, const R* b, R* diff)
{
	R sq = 0;
	for (int k = 0; k < D; k++)
		diff[k] = a[k] - b[k];
	if constexpr (D == 1)
		return std::abs(diff[0]);
	else if constexpr (D == 2)
		sq = diff[0] * diff[0] + diff[1] * diff[1];
	else
		sq = diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2];
	return std::sqrt(sq);
}

// update_membrane_velocity (solver_helper.h:14-28)
template <class R>
inline void membrane_push(R position, R wall, R sign, R radius, R strength, R& velocity)
{
	R distance = std::abs(wall - position);
	distance = std::max<R>(distance, (R)0.00001);
	R repulsion = 1 - distance / radius;
	repulsion = repulsion < 0 ? 0 : repulsion;
	repulsion *= repulsion * strength * sign;
	velocity += repulsion * distance;
}

// update_geometry (interactions_solver.cpp:89-96).  The reference divides real_t volumes by
// a double expression, takes a double cube root and stores real_t; kept as written.
template <class R>
inline void update_geometry(index_t i, R* radius, R* nuclear_radius, R* surface_area, const R* total_volume,
							const R* nuclear_volume)
{
	const double sphere = (double)std::numbers::pi_v<R> * 4.0 / 3.0;
	radius[i] = (R)std::cbrt((double)total_volume[i] / sphere);
	nuclear_radius[i] = (R)std::cbrt((double)nuclear_volume[i] / sphere);
	surface_area[i] = (R)(4.0 * (double)std::numbers::pi_v<R> * (double)radius[i] * (double)radius[i]);
}

// ------------------------------------------------------------------ the solver
template <class R>
class solver
{
public:
	using env_t = environment<R>;

	std::unique_ptr<std::atomic<index_t>[]> cells_in_voxels_sizes;
	std::mutex removal_mtx, interactions_mtx;

	void initialize(env_t& e) // containers_solver.cpp:10-13
	{
		cells_in_voxels_sizes = std::make_unique<std::atomic<index_t>[]>(e.mechanics_mesh.voxel_count());
	}

	static index_t mesh_index(const index_t c[3], const cartesian_mesh& mesh) // common_solver.cpp:10-16
	{
		return c[0] + mesh.grid_shape[0] * (c[1] + mesh.grid_shape[1] * c[2]);
	}

	static void mesh_position(const R* p, const cartesian_mesh& mesh, index_t out[3]) // common_solver.cpp:18-38
	{
		R p3[3] = { 0, 0, 0 };
		for (index_t k = 0; k < mesh.dims; k++)
			p3[k] = p[k];
		mesh.voxel_position(p3, out);
		for (index_t k = mesh.dims; k < 3; k++)
			out[k] = 0;
	}

	// containers_solver::update_mechanics_mesh (containers_solver.cpp:15-56)
	void update_mechanics_mesh(env_t& e)
	{
		const std::size_t voxels = e.mechanics_mesh.voxel_count();
#pragma omp for
		for (std::size_t v = 0; v < voxels; v++)
		{
			e.cells_in_mechanics_voxels[v].clear();
			cells_in_voxels_sizes[v].store(0, std::memory_order_relaxed);
		}
		auto& d = e.data;
		const index_t dims = e.mechanics_mesh.dims;
#pragma omp for
		for (index_t i = 0; i < d.agents_count; i++)
		{
			index_t c[3];
			mesh_position(d.agent_data.positions.data() + (std::size_t)i * dims, e.mechanics_mesh, c);
			cells_in_voxels_sizes[mesh_index(c, e.mechanics_mesh)].fetch_add(1, std::memory_order_relaxed);
		}
#pragma omp for
		for (std::size_t v = 0; v < voxels; v++)
			e.cells_in_mechanics_voxels[v].resize(cells_in_voxels_sizes[v].load(std::memory_order_relaxed));
#pragma omp for
		for (index_t i = 0; i < d.agents_count; i++)
		{
			index_t c[3];
			mesh_position(d.agent_data.positions.data() + (std::size_t)i * dims, e.mechanics_mesh, c);
			const index_t v = mesh_index(c, e.mechanics_mesh);
			const index_t slot = cells_in_voxels_sizes[v].fetch_sub(1, std::memory_order_relaxed) - 1;
			e.cells_in_mechanics_voxels[v][slot] = i;
		}
	}

	// common_solver::for_each_in_mech_neighborhood (common_solver.h:77-105)
	template <class F>
	static void for_each_in_neighborhood(const env_t& e, const index_t c[3], index_t i, F&& f)
	{
		const auto& g = e.mechanics_mesh.grid_shape;
		for (index_t z = -1; z <= 1; z++)
		{
			if (c[2] + z >= g[2] || c[2] + z < 0)
				continue;
			for (index_t y = -1; y <= 1; y++)
			{
				if (c[1] + y >= g[1] || c[1] + y < 0)
					continue;
				for (index_t x = -1; x <= 1; x++)
				{
					if (c[0] + x >= g[0] || c[0] + x < 0)
						continue;
					const index_t nc[3] = { c[0] + x, c[1] + y, c[2] + z };
					for (index_t j : e.cells_in_mechanics_voxels[mesh_index(nc, e.mechanics_mesh)])
						if (j != i)
							f(j);
				}
			}
		}
	}

	// position_solver::update_cell_neighbors (position_solver.cpp:79-97, 180-195, 262-286)
	template <int D>
	void neighbors_internal(env_t& e)
	{
		auto& d = e.data;
		const R* pos = d.agent_data.positions.data();
		const R* radius = d.geometries.radius.data();
		const R* rel = d.mechanics.relative_maximum_adhesion_distance.data();
#pragma omp for
		for (index_t i = 0; i < d.agents_count; i++)
		{
			if (d.is_movable[i] == 0)
				continue;
			index_t c[3];
			mesh_position(pos + (std::size_t)i * D, e.mechanics_mesh, c);
			auto& out = d.states.neighbors[i];
			for_each_in_neighborhood(e, c, i, [&](index_t j) {
				const R reach = rel[i] * radius[i] + rel[j] * radius[j];
				const R dist = pair_distance<D>(pos + (std::size_t)i * D, pos + (std::size_t)j * D);
				if (dist <= reach)
					out.push_back(j);
			});
		}
	}

	void update_cell_neighbors(env_t& e)
	{
		auto& d = e.data;
#pragma omp for
		for (index_t i = 0; i < d.agents_count; i++)
			d.states.neighbors[i].clear();
		dispatch(e, [&]<int D>() { neighbors_internal<D>(e); });
	}

	// solve_pair (position_solver.cpp:24-77)
	template <int D>
	static void solve_pair(index_t lhs, index_t rhs, index_t defs, R* velocity, R* pressure, const R* position,
						   const R* radius, const R* repulsion_strength, const R* adhesion_strength, const R* rel,
						   const R* affinity, const index_t* type)
	{
		R diff[D];
		const R distance = std::max<R>(
			pair_difference<D>(position + (std::size_t)lhs * D, position + (std::size_t)rhs * D, diff), (R)0.00001);

		R repulsion = 1 - distance / (radius[lhs] + radius[rhs]);
		repulsion = repulsion < 0 ? 0 : repulsion;
		repulsion *= repulsion;
		pressure[lhs] += repulsion * simple_pressure_coefficient<R>;
		pressure[rhs] += repulsion * simple_pressure_coefficient<R>;
		repulsion *= std::sqrt(repulsion_strength[lhs] * repulsion_strength[rhs]);

		const R adhesion_distance = rel[lhs] * radius[lhs] + rel[rhs] * radius[rhs];
		R adhesion = 1 - distance / adhesion_distance;
		adhesion *= adhesion;
		adhesion *= std::sqrt(adhesion_strength[lhs] * adhesion_strength[rhs]
							  * affinity[(std::size_t)lhs * defs + type[rhs]] * affinity[(std::size_t)rhs * defs + type[lhs]]);

		const R force = (repulsion - adhesion) / distance;
		for (int k = 0; k < D; k++)
			velocity[(std::size_t)lhs * D + k] += force * diff[k];
	}

	// position_solver::update_cell_forces (position_solver.cpp:15-22, 99-113, 158-178, 233-260)
	template <int D>
	void forces_internal(env_t& e)
	{
		auto& d = e.data;
#pragma omp for
		for (index_t i = 0; i < d.agents_count; i++)
		{
			if (d.is_movable[i] == 0)
				continue;
			for (const index_t j : d.states.neighbors[i])
				solve_pair<D>(i, j, e.cell_definitions_count, d.velocities.data(), d.states.simple_pressure.data(),
							  d.agent_data.positions.data(), d.geometries.radius.data(),
							  d.mechanics.cell_cell_repulsion_strength.data(),
							  d.mechanics.cell_cell_adhesion_strength.data(),
							  d.mechanics.relative_maximum_adhesion_distance.data(),
							  d.mechanics.cell_adhesion_affinities.data(), d.cell_definition_indices.data());
		}
	}

	// evaluate_interactions (src/simulator.cpp:65-82): for an agent with a contact function,
	// the function runs once per attached cell, in list order.  The one model the reference
	// ships, standard_elastic_contact_function (standard_models.cpp:643-654), pulls the agent
	// towards the other with the spring constant of the automated springs.
	void evaluate_interactions(env_t& e)
	{
		auto& d = e.data;
		const index_t dims = e.mechanics_mesh.dims;
		const index_t defs = e.cell_definitions_count;
		const R* k = d.mechanics.attachment_elastic_constant.data();
		const R* aff = d.mechanics.cell_adhesion_affinities.data();
		const index_t* type = d.cell_definition_indices.data();
		const R* x = d.agent_data.positions.data();
		R* v = d.velocities.data();
#pragma omp for
		for (index_t i = 0; i < d.agents_count; i++)
		{
			if (d.contact_function[i] != PCM_CONTACT_STANDARD_ELASTIC)
				continue;
			for (const index_t j : d.states.attached_cells[i])
			{
				const R adhesion = std::sqrt(k[i] * k[j] * aff[i * defs + type[j]] * aff[j * defs + type[i]]);
				for (index_t c = 0; c < dims; c++)
					v[i * dims + c] += adhesion * (x[j * dims + c] - x[i * dims + c]);
			}
		}
	}

	void update_cell_forces(env_t& e)
	{
		// the mechanics step counter that keys the random stream advances here
#pragma omp single
		e.rng.step++;
		auto& d = e.data;
#pragma omp for
		for (index_t i = 0; i < d.agents_count; i++)
			d.states.simple_pressure[i] = 0;
		dispatch(e, [&]<int D>() { forces_internal<D>(e); });
	}

	// position_solver::update_motility (position_solver.cpp:115-145, 197-212, 288-314) with
	// position_helper<d>::random_walk / update_motility_vector / normalize_and_scale / add.
	// The reference hands the restrict_to_2d *array pointer* to random_walk(bool, ...)
	// (position_solver.cpp:131), so in 3D the walk always takes the planar branch
	// (solver_helper.h:252-256); kept as the reference behaves.
	// normalize (solver_helper.h:30-45)
	template <int D>
	static void normalize(R* v)
	{
		R length = 0;
		for (int d = 0; d < D; d++)
			length += v[d] * v[d];
		length = std::sqrt(length);
		for (int d = 0; d < D; d++)
			v[d] = length > zero_threshold<R> ? v[d] / length : 0;
	}

	// The direction updates the reference ships: chemotaxis_function and
	// advanced_chemotaxis_function<normalize_each_gradient> (standard_models.cpp:542-583).
	// They draw no random numbers.
	template <int D>
	void update_migration_bias_direction(env_t& e, index_t i)
	{
		auto& d = e.data;
		auto& mot = d.motilities;
		const R* x = d.agent_data.positions.data() + (std::size_t)i * D;
		R* dir = mot.migration_bias_direction.data() + (std::size_t)i * D;
		const std::uint8_t code = mot.bias_direction_function[i];
		if (code == PCM_BIAS_CHEMOTAXIS)
		{
			const R* g = e.m.nearest_gradient(x, mot.chemotaxis_index[i]);
			for (int k = 0; k < D; k++)
				dir[k] = g[k] * (R)mot.chemotaxis_direction[i];
		}
		else
		{
			const index_t subs = e.m.substrates_count;
			for (int k = 0; k < D; k++)
				dir[k] = 0;
			for (index_t s = 0; s < subs; s++)
			{
				R g[D];
				const R* src = e.m.nearest_gradient(x, s);
				for (int k = 0; k < D; k++)
					g[k] = src[k];
				if (code == PCM_BIAS_ADVANCED_CHEMOTAXIS_NORMALIZED)
					normalize<D>(g);
				for (int k = 0; k < D; k++)
					dir[k] += g[k] * mot.chemotactic_sensitivities[(std::size_t)i * subs + s];
			}
		}
		normalize<D>(dir);
	}

	template <int D>
	void motility_internal(env_t& e)
	{
		auto& d = e.data;
		const R dt = e.mechanics_time_step;
		auto& mot = d.motilities;
#pragma omp for
		for (index_t i = 0; i < d.agents_count; i++)
		{
			if (mot.is_motile[i] == 0)
				continue;
			R* mv = mot.motility_vector.data() + (std::size_t)i * D;
			if (e.rng.template uniform<R>(PCM_STREAM_MOTILITY, i, 0, 0) < dt / mot.persistence_time[i])
			{
				R walk[D];
				if constexpr (D == 1)
				{
					const R u = e.rng.template uniform<R>(PCM_STREAM_MOTILITY, i, 0, 1);
					walk[0] = u < (R)0.5 ? -1 : 1;
				}
				else
				{
					const R theta = e.rng.template uniform<R>(PCM_STREAM_MOTILITY, i, 0, 1, (R)0,
															  2 * std::numbers::pi_v<R>);
					walk[0] = std::cos(theta);
					walk[1] = std::sin(theta);
					if constexpr (D == 3)
						walk[2] = 0;
				}
				// update_migration_bias_direction_f[i](cell) (position_solver.cpp:132-136)
				if (mot.bias_direction_function[i] != PCM_BIAS_NONE)
					update_migration_bias_direction<D>(e, i);
				const R b = mot.migration_bias[i];
				const R* dir = mot.migration_bias_direction.data() + (std::size_t)i * D;
				for (int k = 0; k < D; k++)
					mv[k] = (1 - b) * walk[k] + b * dir[k];
				R length;
				if constexpr (D == 1)
					length = std::abs(mv[0]);
				else if constexpr (D == 2)
					length = std::sqrt(mv[0] * mv[0] + mv[1] * mv[1]);
				else
					length = std::sqrt(mv[0] * mv[0] + mv[1] * mv[1] + mv[2] * mv[2]);
				const R speed = mot.migration_speed[i];
				for (int k = 0; k < D; k++)
					mv[k] = length > zero_threshold<R> ? mv[k] * speed / length : 0;
			}
			for (int k = 0; k < D; k++)
				d.velocities[(std::size_t)i * D + k] += mv[k];
		}
	}

	void update_motility(env_t& e)
	{
		dispatch(e, [&]<int D>() { motility_internal<D>(e); });
	}

	// position_solver::update_basement_membrane_interactions (position_solver.cpp:147-156,
	// 214-231, 316-335) with position_helper<d>::update_membrane_velocities.  The 1D
	// variant reads bounding_box_maxs[1] for the upper wall (solver_helper.h:109); kept.
	template <int D>
	void membrane_internal(env_t& e)
	{
		auto& d = e.data;
		const auto& mesh = e.m.mesh;
#pragma omp for
		for (index_t i = 0; i < d.agents_count; i++)
		{
			if (d.is_movable[i] == 0)
				continue;
			R* v = d.velocities.data() + (std::size_t)i * D;
			const R* p = d.agent_data.positions.data() + (std::size_t)i * D;
			const R r = d.geometries.radius[i];
			const R s = d.mechanics.cell_BM_repulsion_strength[i];
			for (int k = 0; k < D; k++)
			{
				const index_t upper = (D == 1) ? mesh.bounding_box_maxs[1] : mesh.bounding_box_maxs[k];
				membrane_push<R>(p[k], (R)mesh.bounding_box_mins[k], (R)1, r, s, v[k]);
				membrane_push<R>(p[k], (R)upper, (R)-1, r, s, v[k]);
			}
		}
	}

	void update_basement_membrane_interactions(env_t& e)
	{
		if (!e.virtual_wall_at_domain_edges)
			return;
		dispatch(e, [&]<int D>() { membrane_internal<D>(e); });
	}

	// update_spring_attachments_internal (position_solver.cpp:374-449)
	void springs_internal(env_t& e)
	{
		constexpr index_t erased = -1;
		auto& d = e.data;
		auto& springs = d.states.springs;
		const R dt = e.mechanics_time_step;
		const index_t defs = e.cell_definitions_count;
		const index_t n = d.agents_count;

#pragma omp for
		for (index_t i = 0; i < n; i++)
		{
			for (index_t s = 0; s < (index_t)springs[i].size(); s++)
			{
				if (e.rng.template uniform<R>(PCM_STREAM_DETACH, i, s, 0) <= d.mechanics.detachment_rate[i] * dt)
				{
#pragma omp critical
					{
						const index_t other = springs[i][s];
						if (other != erased)
						{
							springs[i][s] = erased;
							*std::find(springs[other].begin(), springs[other].end(), i) = erased;
						}
					}
				}
			}
		}

#pragma omp for
		for (index_t i = 0; i < n; i++)
			springs[i].erase(std::remove(springs[i].begin(), springs[i].end(), erased), springs[i].end());

#pragma omp for
		for (index_t i = 0; i < n; i++)
		{
			const auto& nb = d.states.neighbors[i];
			for (std::size_t s = 0; s < nb.size(); s++)
			{
				const index_t j = nb[s];
				if (j < i)
					continue;
				const R p_l =
					d.mechanics.attachment_rate[i] * dt * d.mechanics.cell_adhesion_affinities[(std::size_t)i * defs + d.cell_definition_indices[j]];
				const R p_r =
					d.mechanics.attachment_rate[j] * dt * d.mechanics.cell_adhesion_affinities[(std::size_t)j * defs + d.cell_definition_indices[i]];
				if (e.rng.template uniform<R>(PCM_STREAM_ATTACH, i, (index_t)s, 0) <= p_l
					|| e.rng.template uniform<R>(PCM_STREAM_ATTACH, i, (index_t)s, 1) <= p_r)
				{
#pragma omp critical
					{
						if ((index_t)springs[i].size() < d.mechanics.maximum_number_of_attachments[i]
							&& (index_t)springs[j].size() < d.mechanics.maximum_number_of_attachments[j])
						{
							springs[i].push_back(j);
							springs[j].push_back(i);
						}
					}
				}
			}
		}
	}

	// spring_contract_function (position_solver.cpp:337-370)
	template <int D>
	void contract_internal(env_t& e)
	{
		auto& d = e.data;
		const index_t defs = e.cell_definitions_count;
#pragma omp for
		for (index_t i = 0; i < d.agents_count; i++)
		{
			if (d.is_movable[i] == 0)
				continue;
			for (const index_t j : d.states.springs[i])
			{
				const index_t ti = d.cell_definition_indices[i], tj = d.cell_definition_indices[j];
				const R k_ij = d.mechanics.attachment_elastic_constant[i] * d.mechanics.attachment_elastic_constant[j]
							   * d.mechanics.cell_adhesion_affinities[(std::size_t)i * defs + tj]
							   * d.mechanics.cell_adhesion_affinities[(std::size_t)j * defs + ti];
				const R adhesion = (R)std::sqrt(k_ij);
				for (int k = 0; k < D; k++)
				{
					const R diff = d.agent_data.positions[(std::size_t)j * D + k] - d.agent_data.positions[(std::size_t)i * D + k];
					d.velocities[(std::size_t)i * D + k] += adhesion * diff;
				}
			}
		}
	}

	void update_spring_attachments(env_t& e) // position_solver.cpp:451-479
	{
		if (!e.automated_spring_adhesion)
			return;
		springs_internal(e);
		dispatch(e, [&]<int D>() { contract_internal<D>(e); });
	}

	// update_positions_internal (position_solver.cpp:481-520)
	template <int D>
	void positions_internal(env_t& e)
	{
		auto& d = e.data;
		const R dt = e.mechanics_time_step;
#pragma omp for
		for (index_t i = 0; i < d.agents_count; i++)
		{
			if (!d.is_movable[i])
				continue;
			const R factor = (R)(dt * 1.5);
			const R previous_factor = (R)(dt * -0.5);
			for (int k = 0; k < D; k++)
			{
				const std::size_t o = (std::size_t)i * D + k;
				d.agent_data.positions[o] += d.velocities[o] * factor + d.previous_velocities[o] * previous_factor;
				d.previous_velocities[o] = d.velocities[o];
				d.velocities[o] = 0;
			}
		}
	}

	void update_positions(env_t& e)
	{
		dispatch(e, [&]<int D>() { positions_internal<D>(e); });
	}

	// ingest / fuse / attack (interactions_solver.cpp:8-141)
	static void ingest(env_t& e, index_t lhs, index_t rhs)
	{
		auto& d = e.data;
		auto& v = d.volumes;
		R* total = d.agent_data.volumes.data();
		v.cytoplasmic_fluid[lhs] += v.fluid[rhs];
		v.cytoplasmic_solid[lhs] += v.solid[rhs];
		v.fluid[lhs] += v.fluid[rhs];
		v.solid[lhs] += v.solid[rhs];
		v.cytoplasmic[lhs] += v.fluid[rhs] + v.solid[rhs];
		total[lhs] += v.fluid[rhs] + v.solid[rhs];
		v.fluid_fraction[lhs] = v.fluid[lhs] / (total[lhs] + zero_threshold<R>);
		v.cytoplasmic_to_nuclear_ratio[lhs] = v.cytoplasmic_solid[lhs] / (v.nuclear_solid[lhs] + zero_threshold<R>);

		const index_t S = e.m.substrates_count;
		for (index_t k = 0; k < S; k++)
			d.agent_data.internalized_substrates[(std::size_t)lhs * S + k] +=
				d.agent_data.internalized_substrates[(std::size_t)rhs * S + k]
				* d.agent_data.fraction_transferred_when_ingested[(std::size_t)rhs * S + k];

		update_geometry<R>(lhs, d.geometries.radius.data(), d.geometries.nuclear_radius.data(),
						   d.geometries.surface_area.data(), total, v.nuclear.data());
		d.flags[rhs] = cell_state_flag::to_remove;
	}

	static void fuse(env_t& e, index_t lhs, index_t rhs)
	{
		auto& d = e.data;
		auto& v = d.volumes;
		R* total = d.agent_data.volumes.data();
		const index_t D = e.mechanics_mesh.dims;
		const R both = total[lhs] + total[rhs];
		for (index_t k = 0; k < D; k++)
		{
			R* p = d.agent_data.positions.data();
			p[(std::size_t)lhs * D + k] =
				(p[(std::size_t)lhs * D + k] * total[lhs] + p[(std::size_t)rhs * D + k] * total[rhs]) / both;
		}

		v.cytoplasmic_fluid[lhs] += v.cytoplasmic_fluid[rhs];
		v.nuclear_fluid[lhs] += v.nuclear_fluid[rhs];
		v.cytoplasmic_solid[lhs] += v.cytoplasmic_solid[rhs];
		v.nuclear_solid[lhs] += v.nuclear_solid[rhs];
		v.target_solid_cytoplasmic[lhs] += v.target_solid_cytoplasmic[rhs];
		v.target_solid_nuclear[lhs] += v.target_solid_nuclear[rhs];
		v.fluid[lhs] = v.nuclear_fluid[lhs] + v.cytoplasmic_fluid[lhs];
		v.solid[lhs] = v.nuclear_solid[lhs] + v.cytoplasmic_solid[lhs];
		v.nuclear[lhs] = v.nuclear_fluid[lhs] + v.nuclear_solid[lhs];
		v.cytoplasmic[lhs] = v.cytoplasmic_fluid[lhs] + v.cytoplasmic_solid[lhs];
		total[lhs] = v.nuclear[lhs] + v.cytoplasmic[lhs];
		v.fluid_fraction[lhs] = v.fluid[lhs] / (total[lhs] + zero_threshold<R>);
		v.cytoplasmic_to_nuclear_ratio[lhs] = v.cytoplasmic_solid[lhs] / (v.nuclear_solid[lhs] + zero_threshold<R>);

		const index_t S = e.m.substrates_count;
		for (index_t k = 0; k < S; k++)
			d.agent_data.internalized_substrates[(std::size_t)lhs * S + k] +=
				d.agent_data.internalized_substrates[(std::size_t)rhs * S + k];

		update_geometry<R>(lhs, d.geometries.radius.data(), d.geometries.nuclear_radius.data(),
						   d.geometries.surface_area.data(), total, v.nuclear.data());
		d.states.number_of_nuclei[lhs] += d.states.number_of_nuclei[rhs];
		d.flags[rhs] = cell_state_flag::to_remove;
	}

	// update_cell_cell_interactions_internal (interactions_solver.cpp:143-253)
	void update_cell_cell_interactions(env_t& e)
	{
		auto& d = e.data;
		const index_t defs = e.cell_definitions_count;
		const R dt = e.mechanics_time_step;
		auto& it = d.interactions;
		auto& flag = d.flags;
#pragma omp for
		for (index_t i = 0; i < d.agents_count; i++)
		{
			bool attacked_once = false, phagocytosed_once = false, fused_once = false;
			if (d.deaths.dead[i] == 1)
				continue;
			const auto& nb = d.states.neighbors[i];
			for (std::size_t s = 0; s < nb.size(); s++)
			{
				const index_t j = nb[s];
				const index_t tj = d.cell_definition_indices[j];
				const R u = e.rng.template uniform<R>(PCM_STREAM_INTERACT, i, (index_t)s, 0);

				if (d.deaths.dead[j] == 1)
				{
					if (u < it.dead_phagocytosis_rate[i] * dt)
					{
						std::lock_guard<std::mutex> l(interactions_mtx);
						if (flag[i] == cell_state_flag::to_remove)
							break;
						if (flag[j] == cell_state_flag::to_remove)
							continue;
						ingest(e, i, j);
					}
					continue;
				}

				if (!phagocytosed_once && u < it.live_phagocytosis_rates[(std::size_t)i * defs + tj] * dt)
				{
					{
						std::lock_guard<std::mutex> l(interactions_mtx);
						if (flag[i] == cell_state_flag::to_remove)
							break;
						if (flag[j] == cell_state_flag::to_remove)
							continue;
						ingest(e, i, j);
					}
					phagocytosed_once = true;
					continue;
				}

				const R attack_r = it.attack_rates[(std::size_t)i * defs + tj];
				const R immuno_r = it.immunogenicities[(std::size_t)j * defs + d.cell_definition_indices[i]];
				if (!attacked_once && u < attack_r * immuno_r * dt)
				{
					{
						std::lock_guard<std::mutex> l(interactions_mtx);
						if (flag[i] == cell_state_flag::to_remove)
							break;
						if (flag[j] == cell_state_flag::to_remove)
							continue;
						d.states.damage[j] += it.damage_rate[i] * dt;
						d.states.total_attack_time[j] += dt;
					}
					attacked_once = true;
					continue;
				}

				if (!fused_once && u < it.fusion_rates[(std::size_t)i * defs + tj] * dt)
				{
					{
						std::lock_guard<std::mutex> l(interactions_mtx);
						if (flag[i] == cell_state_flag::to_remove)
							break;
						if (flag[j] == cell_state_flag::to_remove)
							continue;
						fuse(e, i, j);
					}
					fused_once = true;
					continue;
				}
			}
		}
	}

	// remove_attached / rename_attached (containers_solver.cpp:58-79)
	static void strip_links(index_t gone, std::vector<std::vector<index_t>>& lists)
	{
		for (index_t partner : lists[gone])
		{
			auto& l = lists[partner];
			auto pos = std::find(l.begin(), l.end(), gone);
			*pos = l.back();
			l.pop_back();
		}
	}

	static void rename_links(index_t from, index_t to, std::vector<std::vector<index_t>>& lists)
	{
		for (index_t partner : lists[from])
		{
			auto& l = lists[partner];
			*std::find(l.begin(), l.end(), from) = to;
		}
	}

	// remove_single (containers_solver.cpp:81-123)
	void remove_single(env_t& e, index_t i)
	{
		auto& d = e.data;
		const auto& mesh = e.m.mesh;
		while (true)
		{
			bool out_of_bounds = false;
			for (index_t k = 0; k < mesh.dims; k++)
			{
				const R p = d.agent_data.positions[(std::size_t)i * mesh.dims + k];
				if (p < (R)mesh.bounding_box_mins[k] || p > (R)mesh.bounding_box_maxs[k])
					out_of_bounds = true;
			}
			if (!out_of_bounds && d.flags[i] != cell_state_flag::to_remove)
				return;

			std::lock_guard<std::mutex> l(removal_mtx);
			if (i >= d.agents_count)
				return;
			if (d.flags[i] == cell_state_flag::to_remove)
				e.deaths_count++;
			const index_t last = d.agents_count - 1;
			strip_links(i, d.states.springs);
			rename_links(last, i, d.states.springs);
			strip_links(i, d.states.attached_cells);
			rename_links(last, i, d.states.attached_cells);
			e.remove_at(i);
			if (i == d.agents_count)
				return;
		}
	}

	// position_helper<D>::random_walk(false, walk) (solver_helper.h:82-86, 160-165, 250-267),
	// keyed by the division epoch, the parent and a slot
	template <int D>
	static void free_walk(env_t& e, std::uint64_t epoch, index_t parent, index_t slot, R* walk)
	{
		auto u = [&](std::uint32_t draw, R lo, R hi) {
			return e.rng.template uniform_at<R>(epoch, PCM_STREAM_DIVISION, parent, slot, draw, lo, hi);
		};
		if constexpr (D == 1)
			walk[0] = u(0, 0, 1) < (R)0.5 ? -1 : 1;
		else if constexpr (D == 2)
		{
			const R theta = u(0, 0, 2 * std::numbers::pi_v<R>);
			walk[0] = std::cos(theta);
			walk[1] = std::sin(theta);
		}
		else
		{
			const R theta = u(0, 0, 2 * std::numbers::pi_v<R>);
			const R z = u(1, -1, 1);
			const R r = std::sqrt(1 - z * z);
			walk[0] = std::cos(theta) * r;
			walk[1] = std::sin(theta) * r;
			walk[2] = z;
		}
	}

	// cell_data::add + the cell constructor (cell_data.cpp:262-283, cell.cpp:33-60): a new
	// agent at the end with zero velocities, no flag, movable, default state.  Its phenotype
	// fields are overwritten by copy_from.
	index_t create(env_t& e)
	{
		auto& d = e.data;
		const index_t k = d.agents_count;
		const index_t dims = e.mechanics_mesh.dims;
		for_each_field(d, dims, e.cell_definitions_count, e.m.substrates_count,
					   [k](int, auto& v, index_t w) { v.resize((std::size_t)(k + 1) * w); });
		d.agents_count = k + 1;
		d.states.neighbors.resize(k + 1);
		d.states.springs.resize(k + 1);
		d.states.attached_cells.resize(k + 1);
		d.states.neighbors[k].clear();
		d.states.springs[k].clear();
		d.states.attached_cells[k].clear();
		for (index_t c = 0; c < dims; c++)
		{
			d.velocities[k * dims + c] = 0;
			d.previous_velocities[k * dims + c] = 0;
			d.states.orientation[k * dims + c] = 0;
		}
		d.flags[k] = cell_state_flag::none;
		d.is_movable[k] = 1;
		d.states.simple_pressure[k] = 0;
		d.states.number_of_nuclei[k] = 1;
		d.states.damage[k] = 0;
		d.states.total_attack_time[k] = 0;
		d.origin[k] = -1; // not part of any upload
		return k;
	}

	// The fields cell::copy_from hands from the parent to the daughter: the phenotype
	// (phenotype.cpp:18-30 over data_storage_phenotype.cpp:388-527 and Death::copy,
	// implicit_storage_phenotype.cpp:473-480), the cell definition index and the cell
	// functions (cell.cpp:92-106).  Positions, velocities, state, flags and movability are
	// the daughter's own.
	static bool copied_on_division(int f)
	{
		switch (f)
		{
			case PCM_F_POSITIONS:
			case PCM_F_VELOCITIES:
			case PCM_F_PREVIOUS_VELOCITIES:
			case PCM_F_SIMPLE_PRESSURE:
			case PCM_F_NUMBER_OF_NUCLEI:
			case PCM_F_DAMAGE:
			case PCM_F_TOTAL_ATTACK_TIME:
			case PCM_F_IS_MOVABLE:
			case PCM_F_FLAGS:
			case PCM_F_ORIGIN:
			case PCM_F_ORIENTATION:
				return false;
			default:
				return true;
		}
	}

	// cell::divide (cell.cpp:108-163) of `parent` into the fresh agent `daughter`
	template <int D>
	void divide(env_t& e, index_t parent, index_t daughter, std::uint64_t epoch, bool compute_internalized)
	{
		auto& d = e.data;
		const index_t subs = e.m.substrates_count;
		strip_links(parent, d.states.springs);
		strip_links(parent, d.states.attached_cells);
		d.states.springs[parent].clear();
		d.states.attached_cells[parent].clear();
		if (compute_internalized) // molecular_t::divide (data_storage_phenotype.cpp:552-558)
			for (index_t s = 0; s < subs; s++)
				d.agent_data.internalized_substrates[parent * subs + s] *= (R)0.5;
		d.states.damage[parent] = 0;
		d.states.total_attack_time[parent] = 0;

		R walk[D];
		free_walk<D>(e, epoch, parent, 0, walk);
		const R* o = d.states.orientation.data() + parent * D;
		R tmp = 0;
		for (int k = 0; k < D; k++)
			tmp += walk[k] * o[k];
		for (int k = 0; k < D; k++)
			walk[k] = (walk[k] - tmp * o[k] * d.geometries.polarity[parent]) * d.geometries.radius[parent];

		// volume_t::divide = multiply_by_factor(0.5) (data_storage_phenotype.cpp:530-550)
		for (auto* v : { &d.agent_data.volumes, &d.volumes.solid, &d.volumes.fluid, &d.volumes.nuclear,
						 &d.volumes.nuclear_fluid, &d.volumes.nuclear_solid, &d.volumes.cytoplasmic,
						 &d.volumes.cytoplasmic_fluid, &d.volumes.cytoplasmic_solid, &d.volumes.target_solid_cytoplasmic,
						 &d.volumes.target_solid_nuclear })
			(*v)[parent] *= (R)0.5;
		update_geometry(parent, d.geometries.radius.data(), d.geometries.nuclear_radius.data(),
						d.geometries.surface_area.data(), d.agent_data.volumes.data(), d.volumes.nuclear.data());

		for_each_field(d, D, e.cell_definitions_count, subs, [&](int f, auto& v, index_t w) {
			if (copied_on_division(f))
				for (index_t c = 0; c < w; c++)
					v[daughter * w + c] = v[parent * w + c];
		});
		free_walk<D>(e, epoch, parent, 1, d.states.orientation.data() + daughter * D); // assign_orientation

		R* x = d.agent_data.positions.data();
		for (int k = 0; k < D; k++)
		{
			x[daughter * D + k] = x[parent * D + k] + walk[k];
			x[parent * D + k] -= walk[k] * (R)0.5;
		}
	}

	// containers_solver::update_cell_container_for_phenotype (containers_solver.cpp:149-187),
	// single-threaded: divisions in ascending parent order, then the removal loop over the
	// agents that existed before.  release_internalized_substrates (paraBioFVM) is not part
	// of the restatement.
	void update_cell_container_for_phenotype(env_t& e, bool compute_internalized)
	{
		auto& d = e.data;
		const index_t n = d.agents_count;
		const std::uint64_t epoch = e.division_epoch++;
		e.division_log.clear();
		for (index_t i = 0; i < n; i++)
		{
			if (d.flags[i] != cell_state_flag::to_divide)
				continue;
			d.flags[i] = cell_state_flag::none;
			const index_t k = create(e);
			e.divisions_count++;
			e.division_log.push_back(i);
			dispatch(e, [&]<int D>() { divide<D>(e, i, k, epoch, compute_internalized); });
		}
		for (index_t i = 0; i < n; i++)
			remove_single(e, i);
	}

	// containers_solver::update_cell_container_for_mechanics (containers_solver.cpp:125-147)
	void update_cell_container_for_mechanics(env_t& e)
	{
		const index_t n = e.data.agents_count;
#pragma omp barrier
#pragma omp for
		for (index_t i = 0; i < n; i++)
			remove_single(e, i);
	}

	// simulator::simulate_mechanics, mechanics part (src/simulator.cpp:171-192)
	void step(env_t& e)
	{
		evaluate_interactions(e);
		update_cell_forces(e);
		update_motility(e);
		update_basement_membrane_interactions(e);
		update_spring_attachments(e);
		update_positions(e);
		update_cell_cell_interactions(e);
		update_cell_container_for_mechanics(e);
		update_mechanics_mesh(e);
		update_cell_neighbors(e);
	}

private:
	template <class F>
	static void dispatch(env_t& e, F&& f)
	{
		switch (e.mechanics_mesh.dims)
		{
			case 1:
				f.template operator()<1>();
				break;
			case 2:
				f.template operator()<2>();
				break;
			default:
				f.template operator()<3>();
				break;
		}
	}
};

} // namespace pcm_oracle

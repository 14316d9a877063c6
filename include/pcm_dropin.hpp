// pcm_dropin.hpp — C++ drop-in for paraPhysiCell's mechanics solver (src/solver/).
//
// Provides the reference's solver classes by name — position_solver, containers_solver,
// interactions_solver and the aggregate solver (src/solver/solver.h:9-17,
// src/solver/host/position_solver.h:7-21, containers_solver.h:13-27,
// interactions_solver.h:9-15) — templated on the environment type, each entry point
// forwarding to libpcm.so through include/pcm.h.
//
// Calling convention, as in the reference: an entry point may be called by every thread of
// an enclosing OpenMP team (the reference's orphaned `omp for`) or outside any team.  Inside
// a team one thread makes the device call and all threads leave together (`omp single` and
// its barrier).  Errors, which the reference never reports, throw std::runtime_error.
//
// Residency: the agent SoA lives on the GPU between mechanics phases.  Host code that reads
// cell data calls pcm_dropin::sync_to_host(e) first; host code that writes it calls
// pcm_dropin::host_modified(e) afterwards (the next device phase re-uploads).  Both mirror
// simulator::sync_data_host for the diffusion solver (src/simulator.cpp:95-109).
//
// The integrating code supplies two functions for its environment type, found by
// argument-dependent lookup:
//   auto& pcm_cell_data(Env& e);              // the cell_data SoA (cell_data.h:162-196)
//   void  pcm_host_remove_at(Env& e, int i);  // the container's remove_at (swap with last)
// and, to divide on the device (update_cell_container_for_phenotype), a third:
//   void  pcm_host_divide(Env& e, int parent); // append the daughter's host object
// and optionally, when cells carry contact functions (cell_functions.h:26) or chemotaxis
// direction updates (cell_data.h:100):
//   void  pcm_contact_functions(Env& e, std::vector<uint8_t>& codes);        // PCM_CONTACT_* per agent
//   void  pcm_bias_direction_functions(Env& e, std::vector<uint8_t>& codes); // PCM_BIAS_* per agent
//   void  pcm_gradients(Env& e, int32_t diffusion_voxel[3], std::vector<real_t>& gradients);
// Without them every agent uploads PCM_CONTACT_NONE / PCM_BIAS_NONE.  A contact function
// other than standard_elastic_contact_function is reported as PCM_CONTACT_OTHER, which the
// device refuses (the upload throws), so such a model keeps the host solver.
//
// The solver's free functions that host code calls on host data keep their names here:
// remove_attached, update_geometry and common_solver::get_mesh_index / get_mesh_position.  pcm_gradients fills
// the layout of pcm_set_gradients; pcm_dropin::gradients_modified(e) after each gradient
// solve (simulator.cpp:163) makes the next device phase upload them.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <numbers>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
	#include <omp.h>
#endif

#include "pcm.h"

namespace pcm_dropin {

inline void check(int rc, const char* what)
{
	if (rc != PCM_OK)
		throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + pcm_last_error());
}

// Run `f` on one thread of the current team (or directly outside a team); every thread
// returns after it finished and every thread sees its exception.
template <class F>
void on_one_thread(F&& f)
{
#ifdef _OPENMP
	if (omp_in_parallel())
	{
		static std::exception_ptr error;
	#pragma omp single
		{
			error = nullptr;
			try
			{
				f();
			}
			catch (...)
			{
				error = std::current_exception();
			}
		}
		if (error)
			std::rethrow_exception(error);
		return;
	}
#endif
	f();
}

// ---------------------------------------------------------------- SoA <-> field ids
// Every device field with the reference's member that holds it (cell_data.h:14-196).
template <class CellData, class F>
void for_each_field(CellData& d, F&& f)
{
	f(PCM_F_POSITIONS, d.agent_data.positions);
	f(PCM_F_VELOCITIES, d.velocities);
	f(PCM_F_PREVIOUS_VELOCITIES, d.previous_velocities);
	f(PCM_F_RADIUS, d.geometries.radius);
	f(PCM_F_NUCLEAR_RADIUS, d.geometries.nuclear_radius);
	f(PCM_F_SURFACE_AREA, d.geometries.surface_area);
	f(PCM_F_VOLUME, d.agent_data.volumes);
	f(PCM_F_FLUID, d.volumes.fluid);
	f(PCM_F_SOLID, d.volumes.solid);
	f(PCM_F_NUCLEAR, d.volumes.nuclear);
	f(PCM_F_NUCLEAR_FLUID, d.volumes.nuclear_fluid);
	f(PCM_F_NUCLEAR_SOLID, d.volumes.nuclear_solid);
	f(PCM_F_CYTOPLASMIC, d.volumes.cytoplasmic);
	f(PCM_F_CYTOPLASMIC_FLUID, d.volumes.cytoplasmic_fluid);
	f(PCM_F_CYTOPLASMIC_SOLID, d.volumes.cytoplasmic_solid);
	f(PCM_F_FLUID_FRACTION, d.volumes.fluid_fraction);
	f(PCM_F_CYTOPLASMIC_TO_NUCLEAR_RATIO, d.volumes.cytoplasmic_to_nuclear_ratio);
	f(PCM_F_TARGET_SOLID_CYTOPLASMIC, d.volumes.target_solid_cytoplasmic);
	f(PCM_F_TARGET_SOLID_NUCLEAR, d.volumes.target_solid_nuclear);
	f(PCM_F_CELL_CELL_ADHESION_STRENGTH, d.mechanics.cell_cell_adhesion_strength);
	f(PCM_F_CELL_CELL_REPULSION_STRENGTH, d.mechanics.cell_cell_repulsion_strength);
	f(PCM_F_CELL_BM_REPULSION_STRENGTH, d.mechanics.cell_BM_repulsion_strength);
	f(PCM_F_RELATIVE_MAXIMUM_ADHESION_DISTANCE, d.mechanics.relative_maximum_adhesion_distance);
	f(PCM_F_CELL_ADHESION_AFFINITIES, d.mechanics.cell_adhesion_affinities);
	f(PCM_F_MAXIMUM_NUMBER_OF_ATTACHMENTS, d.mechanics.maximum_number_of_attachments);
	f(PCM_F_ATTACHMENT_ELASTIC_CONSTANT, d.mechanics.attachment_elastic_constant);
	f(PCM_F_ATTACHMENT_RATE, d.mechanics.attachment_rate);
	f(PCM_F_DETACHMENT_RATE, d.mechanics.detachment_rate);
	f(PCM_F_IS_MOTILE, d.motilities.is_motile);
	f(PCM_F_PERSISTENCE_TIME, d.motilities.persistence_time);
	f(PCM_F_MIGRATION_SPEED, d.motilities.migration_speed);
	f(PCM_F_MIGRATION_BIAS_DIRECTION, d.motilities.migration_bias_direction);
	f(PCM_F_MIGRATION_BIAS, d.motilities.migration_bias);
	f(PCM_F_MOTILITY_VECTOR, d.motilities.motility_vector);
	f(PCM_F_RESTRICT_TO_2D, d.motilities.restrict_to_2d);
	f(PCM_F_DEAD, d.deaths.dead);
	f(PCM_F_DEAD_PHAGOCYTOSIS_RATE, d.interactions.dead_phagocytosis_rate);
	f(PCM_F_LIVE_PHAGOCYTOSIS_RATES, d.interactions.live_phagocytosis_rates);
	f(PCM_F_DAMAGE_RATE, d.interactions.damage_rate);
	f(PCM_F_ATTACK_RATES, d.interactions.attack_rates);
	f(PCM_F_IMMUNOGENICITIES, d.interactions.immunogenicities);
	f(PCM_F_FUSION_RATES, d.interactions.fusion_rates);
	f(PCM_F_SIMPLE_PRESSURE, d.states.simple_pressure);
	f(PCM_F_NUMBER_OF_NUCLEI, d.states.number_of_nuclei);
	f(PCM_F_DAMAGE, d.states.damage);
	f(PCM_F_TOTAL_ATTACK_TIME, d.states.total_attack_time);
	f(PCM_F_CELL_DEFINITION_INDEX, d.cell_definition_indices);
	f(PCM_F_IS_MOVABLE, d.is_movable);
	f(PCM_F_FLAGS, d.flags);
	f(PCM_F_INTERNALIZED_SUBSTRATES, d.agent_data.internalized_substrates);
	f(PCM_F_FRACTION_TRANSFERRED_WHEN_INGESTED, d.agent_data.fraction_transferred_when_ingested);
	f(PCM_F_CHEMOTAXIS_INDEX, d.motilities.chemotaxis_index);
	f(PCM_F_CHEMOTAXIS_DIRECTION, d.motilities.chemotaxis_direction);
	f(PCM_F_CHEMOTACTIC_SENSITIVITIES, d.motilities.chemotactic_sensitivities);
	f(PCM_F_ORIENTATION, d.states.orientation);
	f(PCM_F_POLARITY, d.geometries.polarity);
}

// vector<vector<index_t>> <-> CSR
template <class Lists>
void to_csr(const Lists& lists, int n, std::vector<int32_t>& off, std::vector<int32_t>& idx)
{
	off.assign(n + 1, 0);
	idx.clear();
	for (int i = 0; i < n; i++)
	{
		off[i] = (int32_t)idx.size();
		for (auto j : lists[i])
			idx.push_back((int32_t)j);
	}
	off[n] = (int32_t)idx.size();
	if (idx.empty())
		idx.push_back(0);
}

template <class Lists>
void from_csr(Lists& lists, int n, const std::vector<int32_t>& off, const std::vector<int32_t>& idx)
{
	for (int i = 0; i < n; i++)
		lists[i].assign(idx.begin() + off[i], idx.begin() + off[i + 1]);
}

// ---------------------------------------------------------------- residency
struct residency
{
	pcm_sim* sim = nullptr;
	int capacity = 0;
	bool host_newer = true;    // the device copy must be refreshed before the next phase
	bool device_newer = false; // the host copy must be refreshed before host code reads it
	bool gradients_newer = true; // the substrate gradients must be uploaded before the next phase
	int min_capacity = 0;        // agent capacity the next handle needs at least
	unsigned stale = 0;          // sync_parts the host copy lacks
	int64_t deaths_seen = 0;
	uint64_t seed = 0;

	~residency()
	{
		if (sim)
			pcm_destroy(sim);
	}
};

// What sync_to_host brings back.  sync_fields: the removals replayed on the host container,
// every device field, springs, attached cells and deaths_count — what phenotype code
// reads.  The neighbor lists and the voxel buckets (read by signal_behavior.cpp:645-661,
// 757-765, 870, 893-895, cell.cpp:261-267 and the output writers) come only when asked for.
enum sync_parts : unsigned
{
	sync_fields = 1,
	sync_neighbors = 2,
	sync_buckets = 4,
	sync_all = 7
};

template <class Env>
void download(Env& e, residency& r, unsigned parts = sync_all);
template <class Env>
void download_fields(Env& e, residency& r);

inline std::map<const void*, std::unique_ptr<residency>>& registry()
{
	static std::map<const void*, std::unique_ptr<residency>> r;
	return r;
}

inline std::mutex& registry_mutex()
{
	static std::mutex m;
	return m;
}

template <class Env>
residency& state_of(Env& e)
{
	std::lock_guard<std::mutex> l(registry_mutex());
	auto& slot = registry()[(const void*)&e];
	if (!slot)
		slot = std::make_unique<residency>();
	return *slot;
}

// the random stream key of the device draws (the reference seeds std::mt19937,
// src/random.cpp:27-46; the device stream is counter based)
template <class Env>
void set_seed(Env& e, uint64_t seed)
{
	state_of(e).seed = seed;
}

template <class Env>
pcm_step_params params_of(Env& e)
{
	pcm_step_params p;
	p.mechanics_time_step = (double)e.mechanics_time_step;
	p.virtual_wall_at_domain_edges = e.virtual_wall_at_domain_edges ? 1 : 0;
	p.automated_spring_adhesion = e.automated_spring_adhesion ? 1 : 0;
	return p;
}

template <class Env>
concept has_contact_functions = requires(Env& e, std::vector<uint8_t>& codes) { pcm_contact_functions(e, codes); };
template <class Env>
concept has_bias_direction_functions =
	requires(Env& e, std::vector<uint8_t>& codes) { pcm_bias_direction_functions(e, codes); };
template <class Env>
concept has_gradients = requires(Env& e, int32_t* voxel) {
	pcm_gradients(e, voxel, pcm_cell_data(e).agent_data.positions);
};

template <class Env>
void upload(Env& e, residency& r)
{
	auto& d = pcm_cell_data(e);
	const int n = (int)d.agents_count;
	using real = std::remove_reference_t<decltype(d.agent_data.positions[0])>;
	if (!r.sim || n > r.capacity)
	{
		if (r.sim)
			check(pcm_destroy(r.sim), "pcm_destroy");
		pcm_config c {};
		c.mesh.dims = e.m.mesh.dims;
		for (int k = 0; k < 3; k++)
		{
			c.mesh.bbox_min[k] = e.m.mesh.bounding_box_mins[k];
			c.mesh.bbox_max[k] = e.m.mesh.bounding_box_maxs[k];
			c.mesh.voxel[k] = e.mechanics_mesh.voxel_shape[k];
		}
		c.real_bytes = (int)sizeof(real);
		c.cell_defs = e.cell_definitions_count;
		c.substrates = e.m.substrates_count;
		c.capacity = std::max(n < 1024 ? 2048 : 2 * n, r.min_capacity);
		c.seed = r.seed;
		check(pcm_create(&c, &r.sim), "pcm_create");
		r.capacity = c.capacity;
		r.deaths_seen = 0;
		r.gradients_newer = true; // a new handle holds no gradients
	}
	check(pcm_set_count(r.sim, n), "pcm_set_count");
	for_each_field(d, [&](int id, auto& v) { check(pcm_set_field(r.sim, id, v.data()), "pcm_set_field"); });
	if constexpr (has_contact_functions<Env>)
	{
		std::vector<uint8_t> codes(std::max(n, 1), PCM_CONTACT_NONE);
		pcm_contact_functions(e, codes);
		check(pcm_set_field(r.sim, PCM_F_CONTACT_FUNCTION, codes.data()), "pcm_set_field");
	}
	if constexpr (has_bias_direction_functions<Env>)
	{
		std::vector<uint8_t> codes(std::max(n, 1), PCM_BIAS_NONE);
		pcm_bias_direction_functions(e, codes);
		check(pcm_set_field(r.sim, PCM_F_MIGRATION_BIAS_DIRECTION_FUNCTION, codes.data()), "pcm_set_field");
	}
	std::vector<int32_t> off, idx;
	to_csr(d.states.springs, n, off, idx);
	check(pcm_set_links(r.sim, PCM_LINKS_SPRINGS, off.data(), idx.data()), "pcm_set_links");
	to_csr(d.states.attached_cells, n, off, idx);
	check(pcm_set_links(r.sim, PCM_LINKS_ATTACHED, off.data(), idx.data()), "pcm_set_links");
	to_csr(d.states.neighbors, n, off, idx);
	check(pcm_set_neighbors(r.sim, off.data(), idx.data()), "pcm_set_neighbors");
	int64_t log_len = 0;
	check(pcm_removal_log(r.sim, nullptr, 0, &log_len), "pcm_removal_log");
	std::vector<int32_t> discard(log_len + 1);
	check(pcm_removal_log(r.sim, discard.data(), log_len + 1, &log_len), "pcm_removal_log");
	r.host_newer = false;
	r.stale = 0;
}

template <class Env>
residency& on_device(Env& e)
{
	auto& r = state_of(e);
	if (r.host_newer && r.sim && r.stale)
	{
		// the upload resends the lists: fetch the parts the host never synced
		if (r.stale & sync_fields)
			throw std::runtime_error("pcm_dropin: host_modified without a prior sync_to_host of the fields");
		download(e, r, r.stale);
	}
	if (r.host_newer || !r.sim)
		upload(e, r);
	if constexpr (has_gradients<Env>)
		if (r.gradients_newer)
		{
			int32_t voxel[3] = { 0, 0, 0 };
			std::remove_reference_t<decltype(pcm_cell_data(e).agent_data.positions)> g;
			pcm_gradients(e, voxel, g);
			if (!g.empty()) // an environment without a gradient solve supplies none
				check(pcm_set_gradients(r.sim, voxel, g.data()), "pcm_set_gradients");
			r.gradients_newer = false;
		}
	return r;
}

// The gradient solver ran (simulator.cpp:163): upload the gradients before the next phase.
template <class Env>
void gradients_modified(Env& e)
{
	on_one_thread([&] { state_of(e).gradients_newer = true; });
}

// Bring the requested parts of the host copy up to date.  Lists and buckets follow the
// device's agent order, so they imply the field part (and its removal replay).
template <class Env>
void download(Env& e, residency& r, unsigned parts)
{
	auto& d = pcm_cell_data(e);
	parts &= r.stale;
	if (parts & (sync_neighbors | sync_buckets))
		parts |= r.stale & sync_fields;
	const int32_t n_host = (int32_t)d.agents_count;
	if (parts & sync_fields)
		download_fields(e, r);
	const int32_t n = (int32_t)d.agents_count;
	(void)n_host;
	if (parts & sync_neighbors)
	{
		int64_t total = 0;
		check(pcm_get_neighbors_size(r.sim, &total), "pcm_get_neighbors_size");
		std::vector<int32_t> off(n + 1), idx(total + 1);
		check(pcm_get_neighbors(r.sim, off.data(), idx.data()), "pcm_get_neighbors");
		from_csr(d.states.neighbors, n, off, idx);
	}
	if (parts & sync_buckets)
	{
		int64_t voxels = 0;
		check(pcm_get_voxel_count(r.sim, &voxels), "pcm_get_voxel_count");
		std::vector<int32_t> counts(voxels), agents(std::max<int64_t>(r.capacity, 1));
		check(pcm_get_voxel_buckets(r.sim, counts.data(), agents.data()), "pcm_get_voxel_buckets");
		int64_t t = 0;
		for (int64_t v = 0; v < voxels; v++)
		{
			auto& b = e.cells_in_mechanics_voxels[v];
			b.assign(agents.begin() + t, agents.begin() + t + counts[v]);
			t += counts[v];
		}
	}
	r.stale &= ~parts;
	r.device_newer = r.stale != 0;
}

// The field part: replay the device's removals on the host container so host-only members
// follow, then copy every device field, springs, attached cells and the deaths count.
template <class Env>
void download_fields(Env& e, residency& r)
{
	auto& d = pcm_cell_data(e);
	int64_t log_len = 0;
	check(pcm_removal_log(r.sim, nullptr, 0, &log_len), "pcm_removal_log");
	std::vector<int32_t> log(log_len + 1);
	check(pcm_removal_log(r.sim, log.data(), log_len + 1, &log_len), "pcm_removal_log");
	for (int64_t k = 0; k < log_len; k++)
		pcm_host_remove_at(e, log[k]);
	int32_t n = 0;
	check(pcm_get_count(r.sim, &n), "pcm_get_count");
	if (n != (int32_t)d.agents_count)
		throw std::runtime_error("pcm_dropin: host container and device disagree on the agent count");
	for_each_field(d, [&](int id, auto& v) { check(pcm_get_field(r.sim, id, v.data()), "pcm_get_field"); });
	auto fetch = [&](int which, auto& lists) {
		int64_t total = 0;
		check(pcm_get_links_size(r.sim, which, &total), "pcm_get_links_size");
		std::vector<int32_t> off(n + 1), idx(total + 1);
		check(pcm_get_links(r.sim, which, off.data(), idx.data()), "pcm_get_links");
		from_csr(lists, n, off, idx);
	};
	fetch(PCM_LINKS_SPRINGS, d.states.springs);
	fetch(PCM_LINKS_ATTACHED, d.states.attached_cells);
	int64_t deaths = 0;
	check(pcm_get_deaths_count(r.sim, &deaths), "pcm_get_deaths_count");
	e.deaths_count += (decltype(e.deaths_count))(deaths - r.deaths_seen);
	r.deaths_seen = deaths;
}

// sync_to_host(e) brings everything back; sync_to_host(e, sync_fields) leaves the neighbor
// lists and buckets on the device until a later call asks for them.
template <class Env>
void sync_to_host(Env& e, unsigned parts = sync_all)
{
	on_one_thread([&] {
		auto& r = state_of(e);
		if (r.sim && r.device_newer)
			download(e, r, parts);
	});
}

template <class Env>
void host_modified(Env& e)
{
	on_one_thread([&] { state_of(e).host_newer = true; });
}

template <class Env>
void release(Env& e)
{
	std::lock_guard<std::mutex> l(registry_mutex());
	registry().erase((const void*)&e);
}

template <class Env, class Call>
void run(Env& e, Call&& call)
{
	on_one_thread([&] {
		auto& r = on_device(e);
		call(r.sim);
		r.device_newer = true;
		r.stale = sync_all;
	});
}

// ---------------------------------------------------------------- output formats
// Optional: the agent id the writers print (agent::id), by index; without it the index.
//   long long pcm_agent_id(Env& e, int i);
template <class Env>
concept has_agent_ids = requires(Env& e) { pcm_agent_id(e, 0); };

template <class Env>
long long agent_id(Env& e, int i)
{
	if constexpr (has_agent_ids<Env>)
		return (long long)pcm_agent_id(e, i);
	else
		return i;
}

// write_neighbor_graph / write_attached_cells_graph (PhysiCell_MultiCellDS.cpp:2051-2137)
// straight from the device's lists, without building the host's vector-of-vectors: one
// line "id: id,id,..." per agent, lines separated by '\n', none after the last.
template <class Env>
void write_graph(const std::string& filename, Env& e, int which)
{
	on_one_thread([&] {
		auto& r = state_of(e);
		if (!r.sim)
			throw std::runtime_error("pcm_dropin: no device state to write");
		if (r.device_newer)
			download(e, r, sync_fields); // the host's agent order, for the ids
		int32_t n = 0;
		check(pcm_get_count(r.sim, &n), "pcm_get_count");
		int64_t total = 0;
		std::vector<int32_t> off(n + 1);
		if (which == PCM_LINKS_ATTACHED)
			check(pcm_get_links_size(r.sim, which, &total), "pcm_get_links_size");
		else
			check(pcm_get_neighbors_size(r.sim, &total), "pcm_get_neighbors_size");
		std::vector<int32_t> idx(total + 1);
		if (which == PCM_LINKS_ATTACHED)
			check(pcm_get_links(r.sim, which, off.data(), idx.data()), "pcm_get_links");
		else
			check(pcm_get_neighbors(r.sim, off.data(), idx.data()), "pcm_get_neighbors");
		std::string text;
		for (int32_t i = 0; i < n; i++)
		{
			text += std::to_string(agent_id(e, i));
			text += ": ";
			for (int32_t k = off[i]; k < off[i + 1]; k++)
			{
				text += std::to_string(agent_id(e, idx[k]));
				if (k != off[i + 1] - 1)
					text += ',';
			}
			if (i != n - 1)
				text += '\n';
		}
		std::ofstream of(filename, std::ios::out | std::ios::binary);
		of << text;
		if (!of)
			throw std::runtime_error("pcm_dropin: cannot write " + filename);
	});
}

template <class Env>
void write_neighbor_graph(const std::string& filename, Env& e)
{
	write_graph(filename, e, -1);
}

template <class Env>
void write_attached_cells_graph(const std::string& filename, Env& e)
{
	write_graph(filename, e, PCM_LINKS_ATTACHED);
}

// Columns of the MultiCellDS per-cell matrix that live on the device
// (PhysiCell_MultiCellDS.cpp:1735-1790), packed there: see pcm_pack_fields.  Returns
// count * sum(widths) doubles, agent-major.
template <class Env>
std::vector<double> pack_fields(Env& e, const std::vector<int32_t>& fields, const std::vector<int32_t>& widths)
{
	std::vector<double> out;
	on_one_thread([&] {
		auto& r = on_device(e);
		int32_t n = 0;
		check(pcm_get_count(r.sim, &n), "pcm_get_count");
		int64_t row = 0;
		for (std::size_t k = 0; k < fields.size(); k++)
		{
			int32_t type = 0, kind = 0;
			const char* name = nullptr;
			check(pcm_field_info(fields[k], &type, &kind, &name), "pcm_field_info");
			const int own = kind == PCM_W_DIMS ? e.m.mesh.dims
							: kind == PCM_W_DEFS ? e.cell_definitions_count
							: kind == PCM_W_SUBS ? e.m.substrates_count
												 : 1;
			row += widths[k] ? widths[k] : own;
		}
		out.assign((std::size_t)n * row, 0.0);
		if (n > 0)
			check(pcm_pack_fields(r.sim, (int32_t)fields.size(), fields.data(), widths.data(), out.data()),
				  "pcm_pack_fields");
	});
	return out;
}

// ---------------------------------------------------------------- reference classes
// evaluate_interactions (simulator.cpp:65-82): the contact functions over attached cells,
// which the reference runs right before update_cell_forces.  simulate_mechanics includes it.
template <class Env>
void evaluate_interactions(Env& e)
{
	run(e, [&](pcm_sim* s) {
		const auto p = params_of(e);
		check(pcm_evaluate_interactions(s, &p), "pcm_evaluate_interactions");
	});
}

template <class Env>
class position_solver
{
public:
	static void update_cell_forces(Env& e)
	{
		run(e, [&](pcm_sim* s) {
			const auto p = params_of(e);
			check(pcm_update_cell_forces(s, &p), "pcm_update_cell_forces");
		});
	}

	static void update_cell_neighbors(Env& e)
	{
		run(e, [&](pcm_sim* s) { check(pcm_update_cell_neighbors(s), "pcm_update_cell_neighbors"); });
	}

	static void update_motility(Env& e)
	{
		run(e, [&](pcm_sim* s) {
			const auto p = params_of(e);
			check(pcm_update_motility(s, &p), "pcm_update_motility");
		});
	}

	static void update_basement_membrane_interactions(Env& e)
	{
		run(e, [&](pcm_sim* s) {
			const auto p = params_of(e);
			check(pcm_update_basement_membrane_interactions(s, &p), "pcm_update_basement_membrane_interactions");
		});
	}

	static void update_spring_attachments(Env& e)
	{
		run(e, [&](pcm_sim* s) {
			const auto p = params_of(e);
			check(pcm_update_spring_attachments(s, &p), "pcm_update_spring_attachments");
		});
	}

	static void update_positions(Env& e)
	{
		run(e, [&](pcm_sim* s) {
			const auto p = params_of(e);
			check(pcm_update_positions(s, &p), "pcm_update_positions");
		});
	}
};

template <class Env>
class containers_solver
{
public:
	// containers_solver::initialize: a fresh device handle for this environment
	void initialize(Env& e)
	{
		on_one_thread([&] {
			auto& r = state_of(e);
			if (r.sim)
			{
				check(pcm_destroy(r.sim), "pcm_destroy");
				r.sim = nullptr;
			}
			upload(e, r);
		});
	}

	void update_mechanics_mesh(Env& e)
	{
		run(e, [&](pcm_sim* s) { check(pcm_update_mechanics_mesh(s), "pcm_update_mechanics_mesh"); });
	}

	void update_cell_container_for_mechanics(Env& e)
	{
		run(e, [&](pcm_sim* s) {
			const auto p = params_of(e);
			check(pcm_update_cell_container_for_mechanics(s, &p), "pcm_update_cell_container_for_mechanics");
		});
	}

	// containers_solver::update_cell_container_for_phenotype (containers_solver.cpp:149-187):
	// divisions and the removal loop on the device, then the host container appends each
	// daughter's host object through pcm_host_divide(e, parent) (c.create() and the host
	// members of cell::copy_from: type, custom data with divide_conserved_quantities,
	// parameters, functions, phenotype members outside the device fields); the removals
	// replay at the next sync_to_host.  `cell_solver` is the reference's BioFVM argument:
	// release_internalized_substrates is not run here.
	template <class CellSolver>
	void update_cell_container_for_phenotype(Env& e, CellSolver& cell_solver)
	{
		(void)cell_solver;
		on_one_thread([&] {
			auto& r0 = state_of(e);
			if (r0.sim && r0.device_newer && !r0.host_newer)
				download(e, r0, sync_fields); // earlier removals replay before the daughters are appended
			int32_t internalized = 0;
			if constexpr (requires { e.m.compute_internalized_substrates; })
				internalized = e.m.compute_internalized_substrates ? 1 : 0;
			const auto p = params_of(e);
			auto& r = on_device(e);
			int rc = pcm_update_cell_container_for_phenotype(r.sim, &p, internalized);
			if (rc == PCM_ERR_CAPACITY)
			{
				// divisions at most double the population: a larger handle, then again
				int32_t n = 0;
				check(pcm_get_count(r.sim, &n), "pcm_get_count");
				if (r.device_newer)
					download(e, r);
				r.min_capacity = 2 * n + 1024;
				check(pcm_destroy(r.sim), "pcm_destroy");
				r.sim = nullptr;
				upload(e, r);
				rc = pcm_update_cell_container_for_phenotype(r.sim, &p, internalized);
			}
			check(rc, "pcm_update_cell_container_for_phenotype");
			int64_t len = 0;
			check(pcm_division_log(r.sim, nullptr, 0, &len), "pcm_division_log");
			std::vector<int32_t> parents(len + 1);
			check(pcm_division_log(r.sim, parents.data(), len + 1, &len), "pcm_division_log");
			for (int64_t k = 0; k < len; k++)
				pcm_host_divide(e, parents[k]);
			e.divisions_count += (decltype(e.divisions_count))len;
			r.device_newer = true;
			r.stale = sync_all;
		});
	}
};

template <class Env>
class interactions_solver
{
public:
	void update_cell_cell_interactions(Env& e)
	{
		run(e, [&](pcm_sim* s) {
			const auto p = params_of(e);
			check(pcm_update_cell_cell_interactions(s, &p), "pcm_update_cell_cell_interactions");
		});
	}
};

template <class Env>
class solver
{
public:
	position_solver<Env> position;
	containers_solver<Env> containers;
	interactions_solver<Env> interactions;

	void initialize(Env& e) { containers.initialize(e); }

	// `steps` whole mechanics steps (simulator.cpp:171-192, contact functions included) in
	// one device submission
	void simulate_mechanics(Env& e, int steps = 1)
	{
		run(e, [&](pcm_sim* s) {
			const auto p = params_of(e);
			check(pcm_step(s, &p, steps), "pcm_step");
		});
	}
};

// ---------------------------------------------------------------- free functions of the solver
// Host code outside the solver calls these on host data (the agent data must be on the host:
// sync_to_host(e) before, host_modified(e) after a write).

// remove_attached (containers_solver.h:29; containers_solver.cpp:58-69), called by
// cell::divide and cell removal (cell.cpp:111-112, 233): every partner listed in
// attached[to_remove] loses one entry naming to_remove, replaced by its list's last entry.
// The lists must be symmetric (as cell::attach_cells makes them); a partner whose list does
// not name to_remove is left unchanged.
template <class Index>
void remove_attached(Index to_remove, std::vector<Index>* __restrict__ attached)
{
	for (const Index partner : attached[to_remove])
	{
		std::vector<Index>& list = attached[partner];
		for (std::size_t k = 0; k < list.size(); k++)
			if (list[k] == to_remove)
			{
				list[k] = list.back();
				list.pop_back();
				break;
			}
	}
}

// update_geometry (interactions_solver.h:17-20; interactions_solver.cpp:89-96), called by
// the phenotype's volume update (data_storage_phenotype.cpp:113-118): sphere radii of the
// total and nuclear volumes and the surface area, computed in double (the reference's
// literals are double) and stored in real_t, as the device's ingestion and fusion do.
template <class Index, class Real>
void update_geometry(Index i, Real* __restrict__ radius, Real* __restrict__ nuclear_radius,
					 Real* __restrict__ surface_area, const Real* __restrict__ total_volume,
					 const Real* __restrict__ nuclear_volume)
{
	const double sphere = (double)std::numbers::pi_v<Real> * 4.0 / 3.0;
	radius[i] = (Real)std::cbrt((double)total_volume[i] / sphere);
	nuclear_radius[i] = (Real)std::cbrt((double)nuclear_volume[i] / sphere);
	surface_area[i] = (Real)(4.0 * (double)std::numbers::pi_v<Real> * (double)radius[i] * (double)radius[i]);
}

// common_solver's public lookups (common_solver.h:15-19), used by cell.cpp:261-267 to find an
// agent's bucket in cells_in_mechanics_voxels.  `Mesh` is the mechanics mesh type (dims,
// bounding_box_mins/maxs, voxel_shape); the voxel of a position follows the device's
// binning (pcm_get_mesh_position), so host lookups agree with the device-built buckets.
struct common_solver
{
	template <class Mesh>
	static pcm_mesh_desc mesh_desc(const Mesh& mesh)
	{
		pcm_mesh_desc d {};
		d.dims = (int32_t)mesh.dims;
		for (int k = 0; k < 3; k++)
		{
			d.bbox_min[k] = (int32_t)mesh.bounding_box_mins[k];
			d.bbox_max[k] = (int32_t)mesh.bounding_box_maxs[k];
			d.voxel[k] = (int32_t)mesh.voxel_shape[k];
		}
		return d;
	}

	// common_solver::get_mesh_index (common_solver.cpp:10-16): x + gx * (y + gy * z)
	template <class Point, class Mesh>
	static int32_t get_mesh_index(const Point& position, const Mesh& mesh)
	{
		const pcm_mesh_desc d = mesh_desc(mesh);
		const int32_t c[3] = { (int32_t)position[0], (int32_t)position[1], (int32_t)position[2] };
		int32_t index = 0;
		check(pcm_get_mesh_index(&d, c, &index), "pcm_get_mesh_index");
		return index;
	}

	// common_solver::get_mesh_position (common_solver.cpp:18-38): the voxel coordinates of a
	// position of mesh.dims reals; coordinates at and beyond dims are 0
	template <class Real, class Mesh>
	static std::array<int32_t, 3> get_mesh_position(const Real* position, const Mesh& mesh)
	{
		const pcm_mesh_desc d = mesh_desc(mesh);
		std::array<int32_t, 3> c {};
		check(pcm_get_mesh_position(&d, (int32_t)sizeof(Real), position, c.data()), "pcm_get_mesh_position");
		return c;
	}
};

} // namespace pcm_dropin

// Drop-in test: the reference's solver classes replaced by pcm_dropin (libpcm.so on the GPU),
// driven like test/test.cpp drives the reference (calls made by every thread of an OpenMP
// team), compared with the oracle solver on an identical environment.
//
// Exit status 0 when every case passes.  Run by tests/test_dropin.py on a GPU box.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <random>
#include <sstream>

#include "../../include/pcm_dropin.hpp"
#include "../../oracle/pcm_oracle.hpp"

namespace pcm_oracle {
// customization points of the drop-in for the oracle's environment stand-in
template <class R>
cell_data<R>& pcm_cell_data(environment<R>& e)
{
	return e.data;
}

template <class R>
void pcm_host_remove_at(environment<R>& e, int i)
{
	e.remove_at(i);
}

template <class R>
void pcm_contact_functions(environment<R>& e, std::vector<uint8_t>& codes)
{
	for (index_t i = 0; i < e.data.agents_count; i++)
		codes[i] = e.data.contact_function[i];
}

// c.create() for the stand-in: one more agent; the device fields arrive with the download
template <class R>
void pcm_host_divide(environment<R>& e, int parent)
{
	(void)parent;
	e.resize(e.data.agents_count + 1);
}

template <class R>
void pcm_bias_direction_functions(environment<R>& e, std::vector<uint8_t>& codes)
{
	for (index_t i = 0; i < e.data.agents_count; i++)
		codes[i] = e.data.motilities.bias_direction_function[i];
}

template <class R>
void pcm_gradients(environment<R>& e, int32_t voxel[3], std::vector<R>& gradients)
{
	for (int k = 0; k < 3; k++)
		voxel[k] = e.m.mesh.voxel_shape[k];
	gradients = e.m.gradients;
}
} // namespace pcm_oracle

using namespace pcm_oracle;

static int failures = 0;

#define EXPECT(cond, ...)                                                                                              \
	do                                                                                                                 \
	{                                                                                                                  \
		if (!(cond))                                                                                                   \
		{                                                                                                              \
			failures++;                                                                                                \
			std::printf("FAIL %s:%d: ", __FILE__, __LINE__);                                                           \
			std::printf(__VA_ARGS__);                                                                                  \
			std::printf("\n");                                                                                         \
		}                                                                                                              \
	} while (0)

static bool float_eq(double a, double b)
{
	const float fa = (float)a, fb = (float)b;
	if (fa == fb)
		return true;
	int32_t ia, ib;
	std::memcpy(&ia, &fa, 4);
	std::memcpy(&ib, &fb, 4);
	if (ia < 0)
		ia = (int32_t)0x80000000 - ia;
	if (ib < 0)
		ib = (int32_t)0x80000000 - ib;
	return std::llabs((long long)ia - ib) <= 4;
}

// host_velocity_solver.simple (test/test.cpp:141-191) through the drop-in
template <class R>
void velocity_simple(int dims)
{
	pcm_mesh_desc md {};
	md.dims = dims;
	for (int k = 0; k < 3; k++)
	{
		md.bbox_min[k] = 0;
		md.bbox_max[k] = 80;
		md.voxel[k] = 30;
	}
	environment<R> e(md, 3, 2);
	e.resize(3);
	auto& d = e.data;
	for (int i = 0; i < 3; i++)
	{
		d.is_movable[i] = 1;
		d.states.number_of_nuclei[i] = 1;
	}
	const int offsets[2] = { 0, 10 };
	for (int c = 0; c < 2; c++)
	{
		const int o = offsets[c];
		for (int k = 0; k < dims; k++)
			d.agent_data.positions[c * dims + k] = (R)(1 + o + k);
		d.cell_definition_indices[c] = c;
		d.geometries.radius[c] = (R)(4 + o);
		d.mechanics.cell_cell_adhesion_strength[c] = (R)(5 + o);
		d.mechanics.cell_cell_repulsion_strength[c] = (R)(6 + o);
		d.mechanics.relative_maximum_adhesion_distance[c] = (R)(7 + o);
		for (int t = 0; t < 3; t++)
			d.mechanics.cell_adhesion_affinities[c * 3 + t] = (R)(8 + o + t);
	}
	for (int k = 0; k < dims; k++)
		d.agent_data.positions[2 * dims + k] = 70;

	pcm_dropin::containers_solver<environment<R>> cs;
	cs.initialize(e);
#pragma omp parallel
	{
		cs.update_mechanics_mesh(e);
		pcm_dropin::position_solver<environment<R>>::update_cell_neighbors(e);
		pcm_dropin::position_solver<environment<R>>::update_cell_forces(e);
	}
	pcm_dropin::sync_to_host(e);
	const double expected = dims == 3 ? 55.61362893236258 : (dims == 2 ? 69.556601146394243 : 100.15967659502964);
	for (int k = 0; k < dims; k++)
	{
		EXPECT(float_eq(d.velocities[k], expected), "simple dims %d: v0 %g", dims, (double)d.velocities[k]);
		EXPECT(float_eq(d.velocities[dims + k], -expected), "simple dims %d: v1 %g", dims,
			   (double)d.velocities[dims + k]);
		EXPECT(d.velocities[2 * dims + k] == 0, "simple dims %d: v2", dims);
	}
	pcm_dropin::release(e);
}

// a dense random population with springs, interactions and removals
template <class R>
void fill_population(environment<R>& e, int n, unsigned seed)
{
	std::mt19937 g(seed);
	std::uniform_real_distribution<double> u(0.0, 1.0);
	const int D = e.m.mesh.dims;
	e.resize(n);
	auto& d = e.data;
	for (int i = 0; i < n; i++)
	{
		for (int k = 0; k < D; k++)
		{
			d.agent_data.positions[i * D + k] = (R)(-80 + 160 * u(g));
			d.velocities[i * D + k] = (R)(u(g) - 0.5);
			d.previous_velocities[i * D + k] = (R)(u(g) - 0.5);
			d.motilities.migration_bias_direction[i * D + k] = (R)(u(g) - 0.5);
		}
		d.geometries.radius[i] = (R)(7 + 3 * u(g));
		d.geometries.nuclear_radius[i] = 5;
		d.agent_data.volumes[i] = 2494;
		d.volumes.fluid[i] = 1870;
		d.volumes.solid[i] = 624;
		d.volumes.nuclear[i] = 540;
		d.volumes.nuclear_fluid[i] = 405;
		d.volumes.nuclear_solid[i] = 135;
		d.volumes.cytoplasmic[i] = 1954;
		d.volumes.cytoplasmic_fluid[i] = 1465;
		d.volumes.cytoplasmic_solid[i] = 489;
		d.mechanics.cell_cell_adhesion_strength[i] = (R)0.4;
		d.mechanics.cell_cell_repulsion_strength[i] = 10;
		d.mechanics.cell_BM_repulsion_strength[i] = 100;
		d.mechanics.relative_maximum_adhesion_distance[i] = (R)1.25;
		for (int t = 0; t < e.cell_definitions_count; t++)
		{
			d.mechanics.cell_adhesion_affinities[i * e.cell_definitions_count + t] = (R)(0.5 + u(g));
			d.interactions.attack_rates[i * e.cell_definitions_count + t] = (R)(u(g) < 0.3 ? 1.0 : 0.0);
			d.interactions.immunogenicities[i * e.cell_definitions_count + t] = 1;
			d.interactions.fusion_rates[i * e.cell_definitions_count + t] = (R)(u(g) < 0.1 ? 0.5 : 0.0);
		}
		d.mechanics.maximum_number_of_attachments[i] = 4;
		d.mechanics.attachment_elastic_constant[i] = (R)0.02;
		d.mechanics.attachment_rate[i] = (R)(2 * u(g));
		d.mechanics.detachment_rate[i] = (R)(0.5 * u(g));
		d.motilities.is_motile[i] = u(g) < 0.8;
		d.motilities.persistence_time[i] = (R)(0.2 + u(g));
		d.motilities.migration_speed[i] = 1;
		d.motilities.migration_bias[i] = (R)(u(g) < 0.3 ? u(g) : 0.0);
		d.deaths.dead[i] = u(g) < 0.05;
		d.interactions.dead_phagocytosis_rate[i] = (R)(2 * u(g));
		d.interactions.damage_rate[i] = 1;
		d.states.number_of_nuclei[i] = 1;
		d.cell_definition_indices[i] = (int)(u(g) * e.cell_definitions_count);
		d.is_movable[i] = u(g) < 0.9;
		d.flags[i] = cell_state_flag::none;
	}
	// symmetric attached cells (cell::attach_cells, cell.cpp:219-223) on some agents, and
	// standard elastic contact functions on half of them
	for (int i = 0; i < n; i++)
	{
		d.contact_function[i] = u(g) < 0.5 ? PCM_CONTACT_STANDARD_ELASTIC : PCM_CONTACT_NONE;
		if (u(g) < 0.3)
		{
			const int links = 1 + (int)(3 * u(g));
			for (int k = 0; k < links; k++)
			{
				const int j = (int)(u(g) * n) % n;
				if (j == i)
					continue;
				d.states.attached_cells[i].push_back(j);
				d.states.attached_cells[j].push_back(i);
			}
		}
	}
	// chemotaxis models over a random gradient field on a 25 x 20 x 15 um diffusion grid
	for (int i = 0; i < n; i++)
	{
		d.motilities.bias_direction_function[i] = (uint8_t)(4 * u(g)) % 4;
		d.motilities.chemotaxis_index[i] = 0;
		d.motilities.chemotaxis_direction[i] = u(g) < 0.5 ? -1 : 1;
		for (int s = 0; s < e.m.substrates_count; s++)
			d.motilities.chemotactic_sensitivities[i * e.m.substrates_count + s] = (R)(2 * u(g) - 1);
	}
	{
		const int32_t dvox[3] = { 25, 20, 15 };
		const auto& mm = e.mechanics_mesh;
		e.m.mesh = cartesian_mesh(mm.dims, mm.bounding_box_mins, mm.bounding_box_maxs, dvox);
		e.m.gradients.resize(e.m.mesh.voxel_count() * e.m.substrates_count * D);
		for (auto& v : e.m.gradients)
			v = (R)(2 * u(g) - 1);
	}
	// a few agents leave the box during the first step
	for (int i = 0; i < n / 100 + 1; i++)
	{
		d.agent_data.positions[i * D] = (R)(e.m.mesh.bounding_box_maxs[0] - 0.5);
		d.velocities[i * D] = 50;
	}
}

template <class R>
double rel_err(const std::vector<R>& a, const std::vector<R>& b, size_t count)
{
	double worst = 0;
	for (size_t k = 0; k < count; k++)
		worst = std::max(worst, std::abs((double)a[k] - (double)b[k]) / std::max(std::abs((double)b[k]), 1.0));
	return worst;
}

template <class R>
void population_steps(int dims, int n, int steps, bool fused)
{
	const double tol = sizeof(R) == 4 ? 1e-5 : 1e-12;
	pcm_mesh_desc md {};
	md.dims = dims;
	for (int k = 0; k < 3; k++)
	{
		md.bbox_min[k] = -150;
		md.bbox_max[k] = 150;
		md.voxel[k] = 30;
	}
	environment<R> ref(md, 2, 1), dev(md, 2, 1);
	fill_population(ref, n, 42 + dims);
	fill_population(dev, n, 42 + dims);
	for (auto* e : { &ref, &dev })
	{
		e->mechanics_time_step = (R)0.1;
		e->virtual_wall_at_domain_edges = true;
		e->automated_spring_adhesion = true;
	}
	ref.rng.mode = rng_mode::keyed;
	ref.rng.seed = 99;
	pcm_dropin::set_seed(dev, 99);

	solver<R> oracle;
	oracle.initialize(ref);
	oracle.update_mechanics_mesh(ref);
	oracle.update_cell_neighbors(ref);

	pcm_dropin::solver<environment<R>> s;
	s.initialize(dev);
	s.containers.update_mechanics_mesh(dev);
	s.position.update_cell_neighbors(dev);

	for (int step = 0; step < steps; step++)
	{
		oracle.step(ref);
		if (fused)
			s.simulate_mechanics(dev, 1);
		else
		{
#pragma omp parallel
			{
				pcm_dropin::evaluate_interactions(dev);
				s.position.update_cell_forces(dev);
				s.position.update_motility(dev);
				s.position.update_basement_membrane_interactions(dev);
				s.position.update_spring_attachments(dev);
				s.position.update_positions(dev);
				s.interactions.update_cell_cell_interactions(dev);
				s.containers.update_cell_container_for_mechanics(dev);
				s.containers.update_mechanics_mesh(dev);
				s.position.update_cell_neighbors(dev);
			}
		}
		pcm_dropin::sync_to_host(dev);
		auto& a = dev.data;
		auto& b = ref.data;
		EXPECT(a.agents_count == b.agents_count, "dims %d step %d: count %d vs %d", dims, step, a.agents_count,
			   b.agents_count);
		if (a.agents_count != b.agents_count)
		{
			pcm_dropin::release(dev);
			return;
		}
		const size_t nn = (size_t)b.agents_count;
		EXPECT(rel_err(a.agent_data.positions, b.agent_data.positions, nn * dims) <= tol, "dims %d step %d positions",
			   dims, step);
		EXPECT(rel_err(a.previous_velocities, b.previous_velocities, nn * dims) <= tol, "dims %d step %d velocities",
			   dims, step);
		EXPECT(rel_err(a.states.simple_pressure, b.states.simple_pressure, nn) <= tol, "dims %d step %d pressure", dims,
			   step);
		EXPECT(rel_err(a.states.damage, b.states.damage, nn) <= tol, "dims %d step %d damage", dims, step);
		EXPECT(rel_err(a.agent_data.volumes, b.agent_data.volumes, nn) <= tol, "dims %d step %d volumes", dims, step);
		bool lists_equal = true;
		for (size_t i = 0; i < nn; i++)
			lists_equal = lists_equal && a.states.neighbors[i] == b.states.neighbors[i]
						  && a.states.springs[i] == b.states.springs[i]
						  && a.states.attached_cells[i] == b.states.attached_cells[i];
		EXPECT(lists_equal, "dims %d step %d: neighbor/spring lists differ", dims, step);
		bool buckets_equal = dev.cells_in_mechanics_voxels == ref.cells_in_mechanics_voxels;
		EXPECT(buckets_equal, "dims %d step %d: voxel buckets differ", dims, step);
		EXPECT(dev.deaths_count == ref.deaths_count, "dims %d step %d: deaths %d vs %d", dims, step, dev.deaths_count,
			   ref.deaths_count);
		// continue from identical states: the oracle's state is the next step's input
		std::vector<const void*> src(PCM_FIELD_COUNT, nullptr);
		std::vector<size_t> bytes(PCM_FIELD_COUNT, 0);
		for_each_field(b, dims, ref.cell_definitions_count, ref.m.substrates_count, [&](int id, auto& v, index_t) {
			src[id] = v.data();
			bytes[id] = v.size() * sizeof(v[0]);
		});
		for_each_field(a, dims, ref.cell_definitions_count, ref.m.substrates_count, [&](int id, auto& v, index_t) {
			std::memcpy((void*)v.data(), src[id], std::min(bytes[id], v.size() * sizeof(v[0])));
		});
		a.states.springs = b.states.springs;
		a.states.attached_cells = b.states.attached_cells;
		pcm_dropin::host_modified(dev);
	}
	std::printf("population dims %d %s %s: %d agents after %d steps, %d deaths\n", dims,
				sizeof(R) == 4 ? "FP32" : "FP64", fused ? "fused" : "phases", dev.data.agents_count, steps,
				dev.deaths_count);
	pcm_dropin::release(dev);
}

// update_cell_container_for_phenotype through the drop-in against the oracle's, then steps
template <class R>
void phenotype_update(int dims, int n)
{
	const double tol = sizeof(R) == 4 ? 1e-5 : 1e-12;
	pcm_mesh_desc md {};
	md.dims = dims;
	for (int k = 0; k < 3; k++)
	{
		md.bbox_min[k] = -150;
		md.bbox_max[k] = 150;
		md.voxel[k] = 30;
	}
	environment<R> ref(md, 2, 1), dev(md, 2, 1);
	for (auto* e : { &ref, &dev })
	{
		fill_population(*e, n, 77 + dims);
		for (int i = 0; i < n; i += 20)
			e->data.flags[i] = cell_state_flag::to_divide;
		for (int i = 0; i < n; i++)
		{
			e->data.geometries.polarity[i] = (R)(i % 3 == 0);
			for (int k = 0; k < dims; k++)
				e->data.states.orientation[i * dims + k] = (R)(k == i % dims);
		}
		e->mechanics_time_step = (R)0.1;
		e->virtual_wall_at_domain_edges = true;
		e->automated_spring_adhesion = true;
	}
	ref.rng.mode = rng_mode::keyed;
	ref.rng.seed = 5;
	pcm_dropin::set_seed(dev, 5);

	solver<R> oracle;
	oracle.initialize(ref);
	oracle.update_mechanics_mesh(ref);
	oracle.update_cell_neighbors(ref);
	pcm_dropin::solver<environment<R>> s;
	s.initialize(dev);
	s.containers.update_mechanics_mesh(dev);
	s.position.update_cell_neighbors(dev);

	oracle.update_cell_container_for_phenotype(ref, false);
	int cell_solver_stub = 0;
	oracle.update_mechanics_mesh(ref); // simulate_phenotype's housekeeping (simulator.cpp:208-216)
	oracle.update_cell_neighbors(ref);
#pragma omp parallel
	{
		s.containers.update_cell_container_for_phenotype(dev, cell_solver_stub);
		s.containers.update_mechanics_mesh(dev);
		s.position.update_cell_neighbors(dev);
	}
	for (int step = 0; step < 3; step++)
	{
		pcm_dropin::sync_to_host(dev);
		auto& a = dev.data;
		auto& b = ref.data;
		EXPECT(a.agents_count == b.agents_count, "phenotype dims %d step %d: count %d vs %d", dims, step,
			   a.agents_count, b.agents_count);
		EXPECT(dev.divisions_count == ref.divisions_count, "phenotype dims %d: divisions %d vs %d", dims,
			   dev.divisions_count, ref.divisions_count);
		if (a.agents_count != b.agents_count)
			break;
		const size_t nn = (size_t)b.agents_count;
		EXPECT(rel_err(a.agent_data.positions, b.agent_data.positions, nn * dims) <= tol,
			   "phenotype dims %d step %d positions", dims, step);
		EXPECT(rel_err(a.agent_data.volumes, b.agent_data.volumes, nn) <= tol, "phenotype dims %d step %d volumes",
			   dims, step);
		EXPECT(rel_err(a.states.orientation, b.states.orientation, nn * dims) <= tol,
			   "phenotype dims %d step %d orientation", dims, step);
		bool lists_equal = true;
		for (size_t i = 0; i < nn; i++)
			lists_equal = lists_equal && a.states.springs[i] == b.states.springs[i]
						  && a.states.attached_cells[i] == b.states.attached_cells[i];
		EXPECT(lists_equal, "phenotype dims %d step %d: link lists differ", dims, step);
		oracle.step(ref);
		s.simulate_mechanics(dev, 1);
	}
	std::printf("phenotype dims %d %s: %d agents, %d divisions\n", dims, sizeof(R) == 4 ? "FP32" : "FP64",
				dev.data.agents_count, dev.divisions_count);
	pcm_dropin::release(dev);
}

// write_neighbor_graph's text (PhysiCell_MultiCellDS.cpp:2051-2093) from host lists, with
// the agent index as id: a stringstream drained into the output after every agent
static std::string graph_text(const std::vector<std::vector<index_t>>& lists, int n)
{
	std::stringstream buffer;
	std::ostringstream of;
	for (int i = 0; i < n; i++)
	{
		buffer << i << ": ";
		const int size = (int)lists[i].size();
		for (int j = 0; j < size; j++)
		{
			buffer << lists[i][j];
			if (j != size - 1)
				buffer << ",";
		}
		if (i != n - 1)
			buffer << std::endl;
		of << buffer.rdbuf();
	}
	return of.str();
}

static std::string read_file(const std::string& path)
{
	std::ifstream f(path, std::ios::binary);
	std::stringstream ss;
	ss << f.rdbuf();
	return ss.str();
}

// lazy host formats: partial sync, graph dumps from the device lists, packed output columns
template <class R>
void formats(int dims, int n)
{
	pcm_mesh_desc md {};
	md.dims = dims;
	for (int k = 0; k < 3; k++)
	{
		md.bbox_min[k] = -150;
		md.bbox_max[k] = 150;
		md.voxel[k] = 30;
	}
	environment<R> ref(md, 2, 1), dev(md, 2, 1);
	fill_population(ref, n, 91 + dims);
	fill_population(dev, n, 91 + dims);
	for (auto* e : { &ref, &dev })
	{
		e->mechanics_time_step = (R)0.1;
		e->virtual_wall_at_domain_edges = true;
		e->automated_spring_adhesion = true;
	}
	ref.rng.mode = rng_mode::keyed;
	ref.rng.seed = 3;
	pcm_dropin::set_seed(dev, 3);
	solver<R> oracle;
	oracle.initialize(ref);
	oracle.update_mechanics_mesh(ref);
	oracle.update_cell_neighbors(ref);
	pcm_dropin::solver<environment<R>> s;
	s.initialize(dev);
	s.containers.update_mechanics_mesh(dev);
	s.position.update_cell_neighbors(dev);
	for (int step = 0; step < 2; step++)
	{
		oracle.step(ref);
		s.simulate_mechanics(dev, 1);
	}
	// the field part only: counts and fields current, neighbor lists still the old ones
	pcm_dropin::sync_to_host(dev, pcm_dropin::sync_fields);
	EXPECT(dev.data.agents_count == ref.data.agents_count, "formats dims %d: count", dims);
	EXPECT(dev.data.states.springs == ref.data.states.springs, "formats dims %d: springs after sync_fields", dims);
	EXPECT((pcm_dropin::state_of(dev).stale & (pcm_dropin::sync_neighbors | pcm_dropin::sync_buckets)) != 0,
		   "formats dims %d: neighbor lists or buckets synced too early", dims);
	const int nn = dev.data.agents_count;

	const std::string dir = "/tmp/pcm_dropin_formats_" + std::to_string(dims);
	pcm_dropin::write_neighbor_graph(dir + "_neighbors.txt", dev);
	pcm_dropin::write_attached_cells_graph(dir + "_attached.txt", dev);
	EXPECT(read_file(dir + "_neighbors.txt") == graph_text(ref.data.states.neighbors, nn),
		   "formats dims %d: neighbor graph text", dims);
	EXPECT(read_file(dir + "_attached.txt") == graph_text(ref.data.states.attached_cells, nn),
		   "formats dims %d: attached graph text", dims);

	const std::vector<int32_t> fields = { PCM_F_POSITIONS, PCM_F_VOLUME, PCM_F_ORIENTATION, PCM_F_VELOCITIES,
										  PCM_F_SIMPLE_PRESSURE, PCM_F_NUMBER_OF_NUCLEI, PCM_F_CELL_DEFINITION_INDEX,
										  PCM_F_DEAD };
	const std::vector<int32_t> widths = { 3, 0, 3, 3, 0, 0, 0, 0 };
	const auto packed = pcm_dropin::pack_fields(dev, fields, widths);
	const int row = 3 + 1 + 3 + 3 + 1 + 1 + 1 + 1;
	bool same = packed.size() == (size_t)nn * row;
	const auto& d = dev.data;
	for (int i = 0; same && i < nn; i++)
	{
		const double* p = packed.data() + (size_t)i * row;
		for (int k = 0; k < 3; k++)
		{
			same = same && p[k] == (k < dims ? (double)d.agent_data.positions[i * dims + k] : 0.0);
			same = same && p[4 + k] == (k < dims ? (double)d.states.orientation[i * dims + k] : 0.0);
			same = same && p[7 + k] == (k < dims ? (double)d.velocities[i * dims + k] : 0.0);
		}
		same = same && p[3] == (double)d.agent_data.volumes[i] && p[10] == (double)d.states.simple_pressure[i]
			   && p[11] == (double)d.states.number_of_nuclei[i] && p[12] == (double)d.cell_definition_indices[i]
			   && p[13] == (double)d.deaths.dead[i];
	}
	EXPECT(same, "formats dims %d: packed output columns", dims);
	// everything: the neighbor lists arrive
	pcm_dropin::sync_to_host(dev);
	EXPECT(dev.data.states.neighbors == ref.data.states.neighbors, "formats dims %d: neighbors after full sync", dims);
	EXPECT(dev.cells_in_mechanics_voxels == ref.cells_in_mechanics_voxels, "formats dims %d: buckets", dims);
	EXPECT(pcm_dropin::state_of(dev).stale == 0, "formats dims %d: parts left stale", dims);
	std::printf("formats dims %d %s: %d agents\n", dims, sizeof(R) == 4 ? "FP32" : "FP64", nn);
	pcm_dropin::release(dev);
}

// The solver's free functions through the drop-in, on the host lists and fields that host
// code hands them: remove_attached as cell::divide and cell removal call it
// (cell.cpp:111-112, 233) against the oracle's restatement, then device steps from the
// edited state; update_geometry against the oracle's; and the bucket lookup of
// cell.cpp:261-267 (get_mesh_position + get_mesh_index) into the device-built buckets.
template <class R>
void free_functions(int dims, int n)
{
	const double tol = sizeof(R) == 4 ? 1e-5 : 1e-12;
	pcm_mesh_desc md {};
	md.dims = dims;
	for (int k = 0; k < 3; k++)
	{
		md.bbox_min[k] = -150;
		md.bbox_max[k] = 150;
		md.voxel[k] = 30;
	}
	environment<R> ref(md, 2, 1), dev(md, 2, 1);
	fill_population(ref, n, 7 + dims);
	fill_population(dev, n, 7 + dims);
	for (auto* e : { &ref, &dev })
	{
		e->mechanics_time_step = (R)0.1;
		e->virtual_wall_at_domain_edges = true;
		e->automated_spring_adhesion = true;
	}
	ref.rng.mode = rng_mode::keyed;
	ref.rng.seed = 5;
	pcm_dropin::set_seed(dev, 5);
	solver<R> oracle;
	oracle.initialize(ref);
	oracle.update_mechanics_mesh(ref);
	oracle.update_cell_neighbors(ref);
	pcm_dropin::solver<environment<R>> s;
	s.initialize(dev);
	s.containers.update_mechanics_mesh(dev);
	s.position.update_cell_neighbors(dev);
	// a few steps, so that springs form
	for (int step = 0; step < 3; step++)
	{
		oracle.step(ref);
		s.simulate_mechanics(dev, 1);
	}
	pcm_dropin::sync_to_host(dev);
	auto& a = dev.data;
	auto& b = ref.data;
	EXPECT(a.agents_count == b.agents_count, "free functions dims %d: count", dims);
	if (a.agents_count != b.agents_count)
	{
		pcm_dropin::release(dev);
		return;
	}
	const int nn = a.agents_count;

	// the bucket lookup of cell.cpp:261-267, on the device's buckets and against the oracle's
	bool found_all = true, same_voxel = true;
	for (int i = 0; i < nn; i++)
	{
		const R* p = a.agent_data.positions.data() + (size_t)i * dims;
		const auto c = pcm_dropin::common_solver::get_mesh_position(p, dev.mechanics_mesh);
		const int32_t v = pcm_dropin::common_solver::get_mesh_index(c, dev.mechanics_mesh);
		index_t co[3];
		solver<R>::mesh_position(p, dev.mechanics_mesh, co);
		same_voxel = same_voxel && co[0] == c[0] && co[1] == c[1] && co[2] == c[2]
					 && solver<R>::mesh_index(co, dev.mechanics_mesh) == v;
		const auto& bucket = dev.cells_in_mechanics_voxels[v];
		found_all = found_all && std::find(bucket.begin(), bucket.end(), i) != bucket.end();
	}
	EXPECT(same_voxel, "dims %d: get_mesh_position / get_mesh_index differ from the oracle's", dims);
	EXPECT(found_all, "dims %d: an agent is missing from the bucket its position looks up", dims);

	// remove_attached on both lists of every fifth agent with springs, as cell::divide detaches
	// a cell; the oracle's restatement edits the oracle's copy
	int removed = 0;
	for (int i = 0; i < nn; i += 5)
	{
		if (b.states.springs[i].empty())
			continue;
		pcm_dropin::remove_attached((index_t)i, a.states.springs.data());
		pcm_dropin::remove_attached((index_t)i, a.states.attached_cells.data());
		a.states.springs[i].clear();
		a.states.attached_cells[i].clear();
		solver<R>::strip_links((index_t)i, b.states.springs);
		solver<R>::strip_links((index_t)i, b.states.attached_cells);
		b.states.springs[i].clear();
		b.states.attached_cells[i].clear();
		removed++;
	}
	EXPECT(removed > 0, "dims %d: no agent with springs", dims);
	bool lists_equal = true;
	for (int i = 0; i < nn; i++)
		lists_equal = lists_equal && a.states.springs[i] == b.states.springs[i]
					  && a.states.attached_cells[i] == b.states.attached_cells[i];
	EXPECT(lists_equal, "dims %d: remove_attached differs from the oracle's", dims);

	// update_geometry on new volumes (as the phenotype's volume update calls it)
	std::mt19937 g(11 + dims);
	std::uniform_real_distribution<double> u(500.0, 5000.0);
	bool geometry_equal = true;
	for (int i = 0; i < nn; i++)
	{
		const R total = (R)u(g), nuclear = (R)(0.2 * u(g));
		a.agent_data.volumes[i] = b.agent_data.volumes[i] = total;
		a.volumes.nuclear[i] = b.volumes.nuclear[i] = nuclear;
		pcm_dropin::update_geometry((index_t)i, a.geometries.radius.data(), a.geometries.nuclear_radius.data(),
									a.geometries.surface_area.data(), a.agent_data.volumes.data(), a.volumes.nuclear.data());
		pcm_oracle::update_geometry((index_t)i, b.geometries.radius.data(), b.geometries.nuclear_radius.data(),
									b.geometries.surface_area.data(), b.agent_data.volumes.data(), b.volumes.nuclear.data());
		geometry_equal = geometry_equal && a.geometries.radius[i] == b.geometries.radius[i]
						 && a.geometries.nuclear_radius[i] == b.geometries.nuclear_radius[i]
						 && a.geometries.surface_area[i] == b.geometries.surface_area[i];
	}
	EXPECT(geometry_equal, "dims %d: update_geometry differs from the oracle's", dims);

	// continue on the device from the oracle's edited state: one step each, compared
	std::vector<const void*> src(PCM_FIELD_COUNT, nullptr);
	std::vector<size_t> bytes(PCM_FIELD_COUNT, 0);
	for_each_field(b, dims, ref.cell_definitions_count, ref.m.substrates_count, [&](int id, auto& v, index_t) {
		src[id] = v.data();
		bytes[id] = v.size() * sizeof(v[0]);
	});
	for_each_field(a, dims, ref.cell_definitions_count, ref.m.substrates_count, [&](int id, auto& v, index_t) {
		std::memcpy((void*)v.data(), src[id], std::min(bytes[id], v.size() * sizeof(v[0])));
	});
	pcm_dropin::host_modified(dev);
	oracle.step(ref);
	s.simulate_mechanics(dev, 1);
	pcm_dropin::sync_to_host(dev);
	EXPECT(a.agents_count == b.agents_count, "dims %d: count after the edit", dims);
	if (a.agents_count == b.agents_count)
	{
		const size_t m = (size_t)b.agents_count;
		EXPECT(rel_err(a.agent_data.positions, b.agent_data.positions, m * dims) <= tol,
			   "dims %d: positions after the edit", dims);
		EXPECT(rel_err(a.geometries.radius, b.geometries.radius, m) <= tol, "dims %d: radii after the edit", dims);
		bool after = true;
		for (size_t i = 0; i < m; i++)
			after = after && a.states.springs[i] == b.states.springs[i] && a.states.neighbors[i] == b.states.neighbors[i];
		EXPECT(after, "dims %d: lists after the edit", dims);
	}
	std::printf("free functions dims %d %s: %d agents, %d detached\n", dims, sizeof(R) == 4 ? "FP32" : "FP64", nn,
				removed);
	pcm_dropin::release(dev);
}

int main()
{
	for (int dims = 1; dims <= 3; dims++)
	{
		free_functions<float>(dims, dims == 3 ? 1000 : 400);
		free_functions<double>(dims, dims == 3 ? 1000 : 400);
	}
	for (int dims = 1; dims <= 3; dims++)
	{
		velocity_simple<float>(dims);
		velocity_simple<double>(dims);
	}
	for (int dims = 2; dims <= 3; dims++)
	{
		population_steps<float>(dims, dims == 3 ? 1500 : 700, 4, false);
		population_steps<float>(dims, dims == 3 ? 1500 : 700, 4, true);
		population_steps<double>(dims, dims == 3 ? 1500 : 700, 3, false);
	}
	for (int dims = 2; dims <= 3; dims++)
		formats<float>(dims, dims == 3 ? 1200 : 600);
	for (int dims = 1; dims <= 3; dims++)
	{
		phenotype_update<float>(dims, dims == 3 ? 1200 : 500);
		phenotype_update<double>(dims, dims == 3 ? 1200 : 500);
	}
	std::printf("%s (%d failures)\n", failures ? "FAILED" : "PASSED", failures);
	return failures ? 1 : 0;
}

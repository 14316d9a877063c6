// pcm_oracle_capi.cpp — ORACLE: test infrastructure only.
//
// C entry points over the CPU restatement in pcm_oracle.hpp.  They mirror the product's
// C-ABI (include/pcm.h) name for name with an `oracle_` prefix so that a test can drive
// the oracle and the device library with the same host arrays.  Only tests/, bench.py's
// cpu_baseline leg and __graft_entry__.smoke() load this library.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <memory>
#include <string>

#include "pcm_oracle.hpp"

using namespace pcm_oracle;

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg)
{
	g_error = msg;
	return code;
}

struct oracle_base
{
	virtual ~oracle_base() = default;
	int threads = 1;
	virtual int real_bytes() const = 0;
	virtual int set_count(index_t n) = 0;
	virtual index_t count() const = 0;
	virtual int set_field(int f, const void* src) = 0;
	virtual int get_field(int f, void* dst) = 0;
	virtual std::vector<std::vector<index_t>>& lists(int which) = 0;
	virtual std::vector<std::vector<index_t>>& buckets() = 0;
	virtual index_t& deaths() = 0;
	virtual random_service& rng() = 0;
	virtual void set_params(const pcm_step_params* p) = 0;
	virtual int set_gradients(const int32_t* voxel, const void* g) = 0;
	virtual int ready() = 0;
	virtual void phenotype(bool compute_internalized) = 0;
	virtual index_t divisions() const = 0;
	virtual const std::vector<index_t>& division_log() const = 0;
	virtual void phase(int which) = 0;
	virtual void step(int steps) = 0;
};

enum phase_id
{
	PH_MESH,
	PH_NEIGHBORS,
	PH_FORCES,
	PH_MOTILITY,
	PH_MEMBRANE,
	PH_SPRINGS,
	PH_POSITIONS,
	PH_INTERACTIONS,
	PH_CONTAINER,
	PH_CONTACT
};

template <class R>
struct oracle_impl final : oracle_base
{
	environment<R> env;
	solver<R> s;

	oracle_impl(const pcm_config& c) : env(c.mesh, c.cell_defs, c.substrates)
	{
		s.initialize(env);
		env.rng.seed = c.seed;
	}

	int real_bytes() const override { return sizeof(R); }
	index_t count() const override { return env.data.agents_count; }

	int set_count(index_t n) override
	{
		env.resize(n);
		return PCM_OK;
	}

	template <class G>
	int with_field(int f, G&& g)
	{
		bool found = false;
		for_each_field(env.data, env.m.mesh.dims, env.cell_definitions_count, env.m.substrates_count,
					   [&](int id, auto& v, index_t w) {
						   if (id == f)
						   {
							   found = true;
							   g(v, w);
						   }
					   });
		return found ? PCM_OK : fail(PCM_ERR_ARGUMENT, "unknown field id");
	}

	int set_field(int f, const void* src) override
	{
		const index_t n = env.data.agents_count;
		const index_t subs = env.m.substrates_count;
		if (f == PCM_F_CHEMOTAXIS_INDEX)
			for (index_t i = 0; i < n; i++)
				if (((const int32_t*)src)[i] < 0 || ((const int32_t*)src)[i] >= std::max<index_t>(subs, 1))
					return fail(PCM_ERR_ARGUMENT, "chemotaxis_index out of range");
		if (f == PCM_F_MIGRATION_BIAS_DIRECTION_FUNCTION)
			for (index_t i = 0; i < n; i++)
				if (((const uint8_t*)src)[i] > PCM_BIAS_ADVANCED_CHEMOTAXIS_NORMALIZED
					|| (((const uint8_t*)src)[i] != PCM_BIAS_NONE && subs < 1))
					return fail(PCM_ERR_UNSUPPORTED, "unknown direction update code, or chemotaxis without substrates");
		if (f == PCM_F_CONTACT_FUNCTION)
			for (index_t i = 0; i < n; i++)
				if (((const uint8_t*)src)[i] > PCM_CONTACT_STANDARD_ELASTIC)
					return fail(PCM_ERR_UNSUPPORTED, "unknown contact function code");
		return with_field(f, [&](auto& v, index_t w) {
			using T = typename std::remove_reference_t<decltype(v)>::value_type;
			std::memcpy(v.data(), src, sizeof(T) * (std::size_t)env.data.agents_count * w);
		});
	}

	int get_field(int f, void* dst) override
	{
		return with_field(f, [&](auto& v, index_t w) {
			using T = typename std::remove_reference_t<decltype(v)>::value_type;
			std::memcpy(dst, v.data(), sizeof(T) * (std::size_t)env.data.agents_count * w);
		});
	}

	std::vector<std::vector<index_t>>& lists(int which) override
	{
		if (which == PCM_LINKS_SPRINGS)
			return env.data.states.springs;
		if (which == PCM_LINKS_ATTACHED)
			return env.data.states.attached_cells;
		return env.data.states.neighbors;
	}

	std::vector<std::vector<index_t>>& buckets() override { return env.cells_in_mechanics_voxels; }
	index_t& deaths() override { return env.deaths_count; }
	random_service& rng() override { return env.rng; }

	int set_gradients(const int32_t* voxel, const void* g) override
	{
		const auto& mm = env.mechanics_mesh;
		for (int k = 0; k < 3; k++)
			if (voxel[k] < 1)
				return fail(PCM_ERR_ARGUMENT, "diffusion voxel shape must be positive");
		env.m.mesh = cartesian_mesh(mm.dims, mm.bounding_box_mins, mm.bounding_box_maxs, voxel);
		env.m.gradients.resize(env.m.mesh.voxel_count() * env.m.substrates_count * mm.dims);
		std::memcpy(env.m.gradients.data(), g, sizeof(R) * env.m.gradients.size());
		return PCM_OK;
	}

	void phenotype(bool compute_internalized) override
	{
		s.update_cell_container_for_phenotype(env, compute_internalized);
	}
	index_t divisions() const override { return env.divisions_count; }
	const std::vector<index_t>& division_log() const override { return env.division_log; }

	// chemotaxis reads gradients: refuse to step without them
	int ready() override
	{
		if (!env.m.gradients.empty())
			return PCM_OK;
		const auto& code = env.data.motilities.bias_direction_function;
		for (index_t i = 0; i < env.data.agents_count; i++)
			if (code[i] != PCM_BIAS_NONE)
				return fail(PCM_ERR_STATE, "chemotaxis without gradients (pcm_set_gradients)");
		return PCM_OK;
	}

	void set_params(const pcm_step_params* p) override
	{
		env.mechanics_time_step = (R)p->mechanics_time_step;
		env.virtual_wall_at_domain_edges = p->virtual_wall_at_domain_edges != 0;
		env.automated_spring_adhesion = p->automated_spring_adhesion != 0;
	}

	void run_phase(int which)
	{
		switch (which)
		{
			case PH_MESH:
				s.update_mechanics_mesh(env);
				break;
			case PH_NEIGHBORS:
				s.update_cell_neighbors(env);
				break;
			case PH_FORCES:
				s.update_cell_forces(env);
				break;
			case PH_MOTILITY:
				s.update_motility(env);
				break;
			case PH_MEMBRANE:
				s.update_basement_membrane_interactions(env);
				break;
			case PH_SPRINGS:
				s.update_spring_attachments(env);
				break;
			case PH_POSITIONS:
				s.update_positions(env);
				break;
			case PH_INTERACTIONS:
				s.update_cell_cell_interactions(env);
				break;
			case PH_CONTAINER:
				s.update_cell_container_for_mechanics(env);
				break;
			case PH_CONTACT:
				s.evaluate_interactions(env);
				break;
		}
	}

	void phase(int which) override
	{
		if (threads <= 1)
		{
			run_phase(which);
			return;
		}
#pragma omp parallel num_threads(threads)
		run_phase(which);
	}

	// src/simulator.cpp:302-340: the whole time loop runs inside one parallel region.
	void step(int steps) override
	{
		if (threads <= 1)
		{
			for (int k = 0; k < steps; k++)
				s.step(env);
			return;
		}
#pragma omp parallel num_threads(threads)
		{
			for (int k = 0; k < steps; k++)
				s.step(env);
		}
	}
};

oracle_base* cast(void* h) { return static_cast<oracle_base*>(h); }

} // namespace

extern "C" {

const char* oracle_last_error(void) { return g_error.c_str(); }

int oracle_create(const pcm_config* c, void** out)
{
	if (!c || !out)
		return fail(PCM_ERR_ARGUMENT, "null argument");
	if (c->mesh.dims < 1 || c->mesh.dims > 3 || c->cell_defs < 1 || c->substrates < 0)
		return fail(PCM_ERR_ARGUMENT, "bad configuration");
	try
	{
		if (c->real_bytes == 4)
			*out = new oracle_impl<float>(*c);
		else if (c->real_bytes == 8)
			*out = new oracle_impl<double>(*c);
		else
			return fail(PCM_ERR_ARGUMENT, "real_bytes must be 4 or 8");
	}
	catch (const std::exception& ex)
	{
		return fail(PCM_ERR_CAPACITY, ex.what());
	}
	return PCM_OK;
}

int oracle_destroy(void* h)
{
	delete cast(h);
	return PCM_OK;
}

int oracle_set_threads(void* h, int threads)
{
	cast(h)->threads = threads < 1 ? 1 : threads;
	return PCM_OK;
}

// mode 0: keyed counter-based stream (as the device); mode 1: std::mt19937 (as the reference)
int oracle_set_rng(void* h, int mode, uint64_t seed)
{
	auto& r = cast(h)->rng();
	if (mode == 1)
	{
		r.mode = rng_mode::mt;
		r.seed_mt((unsigned)seed);
	}
	else
	{
		r.mode = rng_mode::keyed;
		r.seed = seed;
	}
	return PCM_OK;
}

int oracle_set_count(void* h, int32_t n)
{
	if (n < 0)
		return fail(PCM_ERR_ARGUMENT, "negative count");
	return cast(h)->set_count(n);
}

int oracle_get_count(void* h, int32_t* n)
{
	*n = cast(h)->count();
	return PCM_OK;
}

int oracle_set_field(void* h, int32_t f, const void* src) { return cast(h)->set_field(f, src); }
int oracle_get_field(void* h, int32_t f, void* dst) { return cast(h)->get_field(f, dst); }

int oracle_set_links(void* h, int32_t which, const int32_t* off, const int32_t* idx)
{
	auto* o = cast(h);
	if (which != PCM_LINKS_SPRINGS && which != PCM_LINKS_ATTACHED)
		return fail(PCM_ERR_ARGUMENT, "bad link list id");
	// The reference's removal (remove_attached, containers_solver.cpp:58-69) finds the removed
	// agent in each partner's list and writes through the iterator without checking it: lists
	// that are not symmetric (j appears in i's list as often as i in j's, never i in its own)
	// are undefined behaviour there, so the oracle refuses them.
	const index_t n = o->count();
	for (index_t i = 0; i < n; i++)
		for (int32_t s = off[i]; s < off[i + 1]; s++)
		{
			const index_t j = idx[s];
			if (j < 0 || j >= n || j == i)
				return fail(PCM_ERR_ARGUMENT, "link list entry out of range or naming its own agent");
			const auto mine = std::count(idx + off[i], idx + off[i + 1], j);
			const auto theirs = std::count(idx + off[j], idx + off[j + 1], i);
			if (mine != theirs)
				return fail(PCM_ERR_ARGUMENT, "link lists are not symmetric");
		}
	auto& l = o->lists(which);
	for (index_t i = 0; i < n; i++)
		l[i].assign(idx + off[i], idx + off[i + 1]);
	return PCM_OK;
}

static int lists_size(std::vector<std::vector<index_t>>& l, index_t n, int64_t* total)
{
	int64_t t = 0;
	for (index_t i = 0; i < n; i++)
		t += (int64_t)l[i].size();
	*total = t;
	return PCM_OK;
}

static int lists_get(std::vector<std::vector<index_t>>& l, index_t n, int32_t* off, int32_t* idx)
{
	int64_t t = 0;
	for (index_t i = 0; i < n; i++)
	{
		off[i] = (int32_t)t;
		for (index_t j : l[i])
			idx[t++] = j;
	}
	off[n] = (int32_t)t;
	return PCM_OK;
}

int oracle_get_links_size(void* h, int32_t which, int64_t* total)
{
	return lists_size(cast(h)->lists(which), cast(h)->count(), total);
}

int oracle_get_links(void* h, int32_t which, int32_t* off, int32_t* idx)
{
	return lists_get(cast(h)->lists(which), cast(h)->count(), off, idx);
}

int oracle_set_neighbors(void* h, const int32_t* off, const int32_t* idx)
{
	auto* o = cast(h);
	auto& l = o->lists(2);
	for (index_t i = 0; i < o->count(); i++)
		l[i].assign(idx + off[i], idx + off[i + 1]);
	return PCM_OK;
}

int oracle_get_neighbors_size(void* h, int64_t* total) { return lists_size(cast(h)->lists(2), cast(h)->count(), total); }
int oracle_get_neighbors(void* h, int32_t* off, int32_t* idx) { return lists_get(cast(h)->lists(2), cast(h)->count(), off, idx); }

int oracle_get_voxel_count(void* h, int64_t* v)
{
	*v = (int64_t)cast(h)->buckets().size();
	return PCM_OK;
}

int oracle_get_voxel_buckets(void* h, int32_t* counts, int32_t* agents)
{
	auto& b = cast(h)->buckets();
	int64_t t = 0;
	for (std::size_t v = 0; v < b.size(); v++)
	{
		counts[v] = (int32_t)b[v].size();
		for (index_t j : b[v])
			agents[t++] = j;
	}
	return PCM_OK;
}

int oracle_get_deaths_count(void* h, int64_t* d)
{
	*d = cast(h)->deaths();
	return PCM_OK;
}

int oracle_set_step_index(void* h, uint64_t s)
{
	cast(h)->rng().step = s;
	return PCM_OK;
}

int oracle_get_step_index(void* h, uint64_t* s)
{
	*s = cast(h)->rng().step;
	return PCM_OK;
}

#define ORACLE_PHASE(NAME, ID)                                                                                         \
	int oracle_##NAME(void* h, const pcm_step_params* p)                                                               \
	{                                                                                                                  \
		if (p)                                                                                                         \
			cast(h)->set_params(p);                                                                                    \
		if (int rc = cast(h)->ready())                                                                                 \
			return rc;                                                                                                 \
		cast(h)->phase(ID);                                                                                            \
		return PCM_OK;                                                                                                 \
	}

ORACLE_PHASE(update_mechanics_mesh, PH_MESH)
ORACLE_PHASE(update_cell_neighbors, PH_NEIGHBORS)
ORACLE_PHASE(update_cell_forces, PH_FORCES)
ORACLE_PHASE(update_motility, PH_MOTILITY)
ORACLE_PHASE(update_basement_membrane_interactions, PH_MEMBRANE)
ORACLE_PHASE(update_spring_attachments, PH_SPRINGS)
ORACLE_PHASE(update_positions, PH_POSITIONS)
ORACLE_PHASE(update_cell_cell_interactions, PH_INTERACTIONS)
ORACLE_PHASE(update_cell_container_for_mechanics, PH_CONTAINER)
ORACLE_PHASE(evaluate_interactions, PH_CONTACT)

int oracle_update_cell_container_for_phenotype(void* h, const pcm_step_params* p, int32_t compute_internalized)
{
	if (p)
		cast(h)->set_params(p);
	cast(h)->phenotype(compute_internalized != 0);
	return PCM_OK;
}

int oracle_get_divisions_count(void* h, int64_t* d)
{
	*d = cast(h)->divisions();
	return PCM_OK;
}

int oracle_division_log(void* h, int32_t* parents, int64_t capacity, int64_t* length)
{
	const auto& log = cast(h)->division_log();
	*length = (int64_t)log.size();
	if (parents)
	{
		if (capacity < (int64_t)log.size())
			return fail(PCM_ERR_CAPACITY, "division log longer than the buffer");
		for (std::size_t k = 0; k < log.size(); k++)
			parents[k] = log[k];
	}
	return PCM_OK;
}

int oracle_set_gradients(void* h, const int32_t* voxel, const void* g)
{
	if (!voxel || !g)
		return fail(PCM_ERR_ARGUMENT, "null argument");
	return cast(h)->set_gradients(voxel, g);
}

int oracle_step(void* h, const pcm_step_params* p, int32_t steps)
{
	cast(h)->set_params(p);
	if (int rc = cast(h)->ready())
		return rc;
	cast(h)->step(steps);
	return PCM_OK;
}

int oracle_set_step_params(void* h, const pcm_step_params* p)
{
	if (!p)
		return fail(PCM_ERR_ARGUMENT, "null params");
	cast(h)->set_params(p);
	return PCM_OK;
}

// Wall time of `steps` mechanics steps (after `warmup` untimed ones), in milliseconds.
int oracle_time_steps(void* h, const pcm_step_params* p, int32_t warmup, int32_t steps, double* ms)
{
	cast(h)->set_params(p);
	if (int rc = cast(h)->ready())
		return rc;
	cast(h)->step(warmup);
	const auto t0 = std::chrono::steady_clock::now();
	cast(h)->step(steps);
	const auto t1 = std::chrono::steady_clock::now();
	*ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
	return PCM_OK;
}

// One Philox4x32-10 block, for known-answer checks of the shared stream.
int oracle_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out)
{
	const pcm_u32x4 b = pcm_philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
	for (int k = 0; k < 4; k++)
		out[k] = b.v[k];
	return PCM_OK;
}

int oracle_max_threads(void)
{
#ifdef _OPENMP
	return omp_get_max_threads();
#else
	return 1;
#endif
}

} // extern "C"

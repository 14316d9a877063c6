// pcm_device.h — device-side state of one simulation and the per-agent arithmetic of the
// mechanics step, shared by all kernels of libpcm.so.
//
// Layout in HBM (all arrays sized by the handle's agent capacity):
//   - every agent field of include/pcm.h in the reference's own layout (agent-major,
//     `width` elements per agent), so uploads and downloads are plain copies;
//   - neighbor lists: fixed stride K per agent (nbr_idx[i*K + s], nbr_cnt[i]); the list
//     of every agent is built (non-movable ones too, they feed simple_pressure), the
//     reference-visible list of a non-movable agent is empty;
//   - springs / attached cells: fixed stride L / LA per agent with counts;
//   - mechanics voxel grid: vcnt[V], vstart[V] (only voxels touched by agents are ever
//     written), bucket[] holding each voxel's agents contiguously in the reference's
//     bucket order, and a voxel-sorted copy of (x, y, z, rel*r) for the neighbor search.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pcm.h"
#include "pcm_rng.h"

namespace pcm {

// device status bits (Scalars::status); any set bit makes later kernels of the step return
enum : int
{
	ST_NEIGHBOR_OVERFLOW = 1, // some agent has more than K neighbors: grow K, rebuild
	ST_LIST_OVERFLOW = 2,     // a link list would exceed its fixed capacity
	ST_CAPACITY = 4,          // a scratch buffer capacity was exceeded
	ST_BUCKET_OVERFLOW = 8,   // a voxel holds more agents than its unordered bucket table: grow B, rebuild
};

struct Scalars
{
	int n;                 // agents_count
	int status;            // ST_* bits
	unsigned long long step; // mechanics step index (keys the random stream)
	unsigned long long claim; // voxel claims of the mesh update: agents << 32 | voxels (zero between updates)
	int n_binned;          // agents binned by the last mesh update (for the sparse clear)
	int div_count;         // divisions planned by the last phenotype-cadence update
	int cand_count;        // interaction candidate slots collected this step
	int conflict;          // interaction speculation found a cross-agent dependency
	int bad_count;         // agents the container update removes this step
	int max_neighbors;     // largest neighbor count seen by the last neighbor build
	int moves;             // agents moved by the container update
	long long deaths;      // environment::deaths_count increments
	long long divisions;   // environment::divisions_count increments
	int n_occ;             // occupied mechanics voxels of the last mesh update
	int log_len;           // entries in the removal log (may exceed log_cap: log lost)
	int compact_limit;     // > 0: the removal loop visits agents below it only
	int res_open;          // actors the first resolve round left with undecided events
	int ticket;            // blocks of the current launch that have finished (last-block hand-off)
	int max_bucket;        // largest voxel population seen by an overflowing fused mesh update
	unsigned res_done;     // low word of the step whose later attach-resolve rounds have finished
	// diagnostics of the last step (pcm_get_stats): attachment events, resolver rounds,
	// resolver chunks, interaction candidates, speculation conflict, removals, occupied voxels
	int stat[8];
};

// element type (PCM_T_*) of a field id
__host__ __device__ constexpr int pcm_field_type(int f)
{
	constexpr int types[PCM_FIELD_COUNT] = {
#define PCM_FIELD_TYPE_(name, type, width) type,
		PCM_FIELD_LIST(PCM_FIELD_TYPE_)
#undef PCM_FIELD_TYPE_
	};
	return types[f];
}

// Pointers into device memory.  In device code they carry the global address space, so
// accesses compile to global_* instructions; generic (flat) ones would also count against
// the LDS counter and make every LDS or cross-lane wait drain outstanding memory traffic.
#if defined(__HIP_DEVICE_COMPILE__)
	#define PCM_GLOBAL __attribute__((address_space(1)))
#else
	#define PCM_GLOBAL
#endif
template <class T>
using gptr = T PCM_GLOBAL*;

template <class R>
struct SimCtx
{
	// configuration
	int dims, T, S, cap;
	int bmin[3], bmax[3], vox[3], grid[3];
	long long vcount;
	int K;       // neighbor slots per agent
	int B;       // slots per voxel of the unordered bucket table (3D fused mesh update)
	int L;       // spring slots per agent
	int LA;      // attached-cell slots per agent
	int cand_cap; // interaction candidate buffer
	int bad_cap;  // removal list buffer
	unsigned long long seed;
	// environment scalars of the current phase call
	R dt;
	int wall;
	int springs;
	int contacts;   // some uploaded agent has a contact function
	unsigned long long div_epoch; // phenotype-update call: keys the division draws
	int div_internal;             // divisions halve internalized substrates
	int chemotaxis; // some uploaded agent has a migration bias direction function
	// substrate gradients on the diffusion grid (pcm_set_gradients), or null
	gptr<const R> grad;
	int dvox[3], dgrid[3];
	gptr<Scalars> sc;

	// agent fields (include/pcm.h order); fptr/fbytes give the generic view
	gptr<void> fptr[PCM_FIELD_COUNT];
	int fbytes[PCM_FIELD_COUNT]; // bytes per agent
	gptr<R> pos, vel, pvel, radius, nuc_radius, surface, volume, fluid, solid, nuclear, nuc_fluid, nuc_solid, cyto, cyto_fluid, cyto_solid, fluid_frac, c2n, target_cyto, target_nuc;
	gptr<R> adhesion, repulsion, bm_repulsion, rel_adh, affinity;
	gptr<int> max_att;
	gptr<R> elastic, att_rate, det_rate;
	gptr<uint8_t> motile;
	gptr<R> persistence, speed, bias_dir, bias, motility;
	gptr<uint8_t> restrict2d;
	gptr<uint8_t> dead;
	gptr<R> dead_phago, live_phago, damage_rate, attack, immuno, fusion;
	gptr<R> pressure;
	gptr<int> nuclei;
	gptr<R> damage, attack_time;
	gptr<int> type;
	gptr<uint8_t> movable;
	gptr<uint8_t> flags;
	gptr<R> internal, frac_ingest;
	gptr<int> origin;
	gptr<uint8_t> contact; // PCM_CONTACT_* code per agent
	gptr<int> chemo_index, chemo_dir;
	gptr<R> sensitivities;
	gptr<uint8_t> bias_fn; // PCM_BIAS_* code per agent
	gptr<R> orientation, polarity;
	gptr<int> div_list; // dividing agents of the last phenotype update, ascending

	// link lists
	gptr<int> spr_cnt, spr_idx, atc_cnt, atc_idx;
	// neighbor lists
	gptr<int> nbr_cnt, nbr_idx;
	// mechanics voxel grid
	gptr<int> vcnt, vstart, vox_of, vrank, btmp, bucket, occ;
	gptr<int> vmem; // 3D: B slots per voxel, the voxel's agents in arrival order (fused mesh update)
	gptr<R> sorted; // 4 reals per bucket slot: x, y, z, rel*radius
	gptr<R> frec;   // 8 reals per agent: x, y, z, radius, repulsion, adhesion, rel, type|movable<<30 bits

	// spring scratch
	gptr<uint32_t> fire, erase;  // per agent slot masks (L <= 32)
	gptr<int> own_tgt;            // per actor, K slots: partners of its attachment events, in slot order
	gptr<uint8_t> own_st;         // per actor, K slots: 0 undecided, 1 success, 2 failure
	gptr<int> own_cnt;            // attachment events of an actor this step
	gptr<int> inbox;              // per agent, K (actor, k) pairs: events that target it
	gptr<int> in_cnt;             // inbox fill (zero outside the spring phase)
	gptr<int> att_list;           // actors with attachment events, a fixed segment per collect block
	gptr<int> blk_cnt;            // actors listed in each collect block's segment
	gptr<int> blk_ev;             // attachment events of each segment (counted by the first resolve round)
	gptr<int> att_c;              // the actors of all segments, compacted by the resolver
	gptr<int> att_c2;             // the resolver's next-round actors (beyond its LDS lists)
	// interaction scratch
	gptr<int> cand;               // candidate (actor, slot) pairs
	gptr<unsigned long long> ev_t; // interaction events of an agent (EVW_* fields), raised atomically
	gptr<int> rem_cnt;            // tentative removal events of an actor (<= 4), -1 = none recorded
	gptr<int> rem_ev;             // 4 per actor: slot | kind << 24
	gptr<int> atk_tgt;            // attacked agent of an actor, -1 = none
	// compaction scratch
	gptr<int> bad, remap, seq, move_src, move_dst;
	gptr<int> rm_log; // positions passed to remove_at, in the reference's order, since the last read
	int log_cap;
	// compaction copy plan: one entry per 4-byte word (or byte) of an agent's fields
	int mv_n;
	uint16_t mv_field[192];
	uint16_t mv_unit[192];
	uint8_t mv_bytes[192];
};

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// exclusive prefix sum over the 64 lanes of a wavefront; `total` receives the wave sum
__device__ __forceinline__ int wave_exclusive_scan(int x, int& total)
{
	int v = x;
	const int lane = lane_id();
	for (int o = 1; o < 64; o <<= 1)
	{
		const int y = __shfl_up(v, o, 64);
		if (lane >= o)
			v += y;
	}
	total = __shfl(v, 63, 64);
	return v - x;
}

// exclusive prefix sum over a block of up to 1024 threads (all threads must call it)
__device__ __forceinline__ int block_exclusive_scan(int x, int& total)
{
	__shared__ int s_wave[16];
	int wt;
	const int pre = wave_exclusive_scan(x, wt);
	const int w = threadIdx.x >> 6;
	if (lane_id() == 63)
		s_wave[w] = wt;
	__syncthreads();
	int off = 0, tot = 0;
	const int nw = (blockDim.x + 63) >> 6;
	for (int k = 0; k < nw; k++)
	{
		if (k < w)
			off += s_wave[k];
		tot += s_wave[k];
	}
	__syncthreads();
	total = tot;
	return off + pre;
}

// Reserve `count` consecutive slots of a global counter for each lane with one atomic per
// wavefront; returns the lane's first slot.  All lanes of the wavefront must call it.
__device__ __forceinline__ int wave_reserve(int* counter, int count)
{
	int total;
	const int pre = wave_exclusive_scan(count, total);
	int base = 0;
	if (lane_id() == 0 && total > 0)
		base = atomicAdd(counter, total);
	base = __shfl(base, 0, 64);
	return base + pre;
}

// An agent's own row of a per-cell-type table (affinities, rates), read by the type of a
// partner: with at most TYPE_ROW_REGS types the row is loaded once into registers, so
// the lookup costs no load after the partner's type arrives.
constexpr int TYPE_ROW_REGS = 4;
template <class R>
struct type_row
{
	R v0, v1, v2, v3; // named, not an array: a selected array element would go to scratch
	gptr<const R> row;
	int T;
	__device__ __forceinline__ R at(int t) const
	{
		if (T > TYPE_ROW_REGS)
			return row[t];
		return t == 0 ? v0 : (t == 1 ? v1 : (t == 2 ? v2 : v3));
	}
};

template <class R>
__device__ __forceinline__ type_row<R> load_type_row(gptr<const R> row, int T)
{
	type_row<R> r;
	r.row = row;
	r.T = T;
	r.v0 = T > 0 ? row[0] : (R)0;
	r.v1 = T > 1 ? row[1] : (R)0;
	r.v2 = T > 2 ? row[2] : (R)0;
	r.v3 = T > 3 ? row[3] : (R)0;
	return r;
}

// -------------------------------------------------------------------------- force records
__device__ __forceinline__ float bits_to_real(int b, float) { return __int_as_float(b); }
__device__ __forceinline__ double bits_to_real(int b, double) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ int real_to_bits(float r) { return __float_as_int(r); }
__device__ __forceinline__ int real_to_bits(double r) { return (int)__double_as_longlong(r); }

template <class R>
struct force_rec
{
	R p[3];
	R r, rep, adh, rel;
	int type;
	bool movable;
};

template <class R, int D>
__device__ __forceinline__ void write_force_record(const SimCtx<R>& c, int i)
{
	R* f = c.frec + (size_t)i * 8;
	for (int k = 0; k < 3; k++)
		f[k] = k < D ? c.pos[(size_t)i * D + k] : (R)0;
	f[3] = c.radius[i];
	f[4] = c.repulsion[i];
	f[5] = c.adhesion[i];
	f[6] = c.rel_adh[i];
	f[7] = bits_to_real(c.type[i] | (c.movable[i] ? (1 << 30) : 0), R(0));
}

template <class R>
__device__ __forceinline__ force_rec<R> read_force_record(const SimCtx<R>& c, int j)
{
	force_rec<R> o;
	const R* f = c.frec + (size_t)j * 8;
	R v[8];
	if constexpr (sizeof(R) == 4)
	{
		const float4 a = *(const float4*)f;
		const float4 b = *(const float4*)(f + 4);
		v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
	}
	else
	{
		for (int k = 0; k < 8; k += 2)
		{
			const double2 a = *(const double2*)(f + k);
			v[k] = a.x, v[k + 1] = a.y;
		}
	}
	o.p[0] = v[0], o.p[1] = v[1], o.p[2] = v[2];
	o.r = v[3], o.rep = v[4], o.adh = v[5], o.rel = v[6];
	const int bits = real_to_bits(v[7]);
	o.type = bits & ((1 << 30) - 1);
	o.movable = (bits >> 30) & 1;
	return o;
}

// -------------------------------------------------------------------------- helpers
template <class R>
__device__ __forceinline__ R real_max(R a, R b)
{
	return a < b ? b : a; // std::max(a, b)
}

template <class R>
__device__ __forceinline__ R zero_threshold()
{
	return (R)1e-16;
}

template <class R>
__device__ __forceinline__ R pressure_coefficient()
{
	return (R)36.64504274775163;
}

// position_helper<D>::distance (solver_helper.h:54-58, 127-131, 212-217)
template <int D, class R>
__device__ __forceinline__ R pair_distance(const R* a, const R* b)
{
	if constexpr (D == 1)
		return fabs(a[0] - b[0]);
	else if constexpr (D == 2)
	{
		const R s = (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]);
		return sqrt(s);
	}
	else
	{
		const R s = (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]);
		return sqrt(s);
	}
}

// sqrt(s) <= reach with the correctly rounded square root of the reference, decided
// without the square root when s is clearly on one side: if s > reach^2 (1 + e) then
// sqrt_rn(s) >= sqrt(s)(1 - u) > reach, and if s < reach^2 (1 - e) then
// sqrt_rn(s) <= sqrt(s)(1 + u) < reach, for e = 1e-5 (float) / 1e-12 (double) far above the
// rounding errors u of sqrt and of the products that form the thresholds.
template <class R>
__device__ __forceinline__ bool within_reach(R s, R reach)
{
	const R e = sizeof(R) == 4 ? (R)1e-5 : (R)1e-12;
	const R r2 = reach * reach;
	if (!(r2 >= (sizeof(R) == 4 ? (R)1.2e-38 : (R)2.3e-308))) // zero, subnormal or NaN threshold
		return sqrt(s) <= reach;
	if (s > r2 * ((R)1 + e))
		return false;
	if (s < r2 * ((R)1 - e))
		return true;
	return sqrt(s) <= reach;
}

// position_helper<D>::distance(a, b) <= reach, bit-exact with the reference's test
template <int D, class R>
__device__ __forceinline__ bool pair_within(const R* a, const R* b, R reach)
{
	if constexpr (D == 1)
		return fabs(a[0] - b[0]) <= reach;
	else if constexpr (D == 2)
		return within_reach<R>((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]), reach);
	else
		return within_reach<R>(
			(a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]), reach);
}

// voxel coordinate of a position: (index_t)((p - min) / voxel), clamped to the grid
template <class R>
__device__ __forceinline__ int voxel_index(const SimCtx<R>& c, const R* p)
{
	int co[3] = { 0, 0, 0 };
	for (int k = 0; k < c.dims; k++)
	{
		const R q = (p[k] - (R)c.bmin[k]) / (R)c.vox[k];
		int v = (int)q;
		v = v < 0 ? 0 : (v >= c.grid[k] ? c.grid[k] - 1 : v);
		co[k] = v;
	}
	return co[0] + c.grid[0] * (co[1] + c.grid[1] * co[2]);
}

template <class R>
__device__ __forceinline__ void voxel_coords(const SimCtx<R>& c, int v, int co[3])
{
	co[0] = v % c.grid[0];
	co[1] = (v / c.grid[0]) % c.grid[1];
	co[2] = v / (c.grid[0] * c.grid[1]);
}

// solve_pair (position_solver.cpp:24-77) for one neighbor: adds the pair force to v and
// returns the repulsion term that feeds simple_pressure.
template <int D, class R>
__device__ __forceinline__ R pair_force(const SimCtx<R>& c, int i, int j, const R* pi, R* v)
{
	const R* pj = c.pos + (size_t)j * D;
	R diff[D];
	R sq = 0;
	for (int k = 0; k < D; k++)
		diff[k] = pi[k] - pj[k];
	R distance;
	if constexpr (D == 1)
		distance = fabs(diff[0]);
	else if constexpr (D == 2)
	{
		sq = diff[0] * diff[0] + diff[1] * diff[1];
		distance = sqrt(sq);
	}
	else
	{
		sq = diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2];
		distance = sqrt(sq);
	}
	distance = real_max<R>(distance, (R)0.00001);

	R rep = 1 - distance / (c.radius[i] + c.radius[j]);
	rep = rep < 0 ? (R)0 : rep;
	rep *= rep;
	const R pressure_term = rep;
	rep *= sqrt(c.repulsion[i] * c.repulsion[j]);

	const R reach = c.rel_adh[i] * c.radius[i] + c.rel_adh[j] * c.radius[j];
	R adh = 1 - distance / reach;
	adh *= adh;
	const int ti = c.type[i], tj = c.type[j];
	adh *= sqrt(c.adhesion[i] * c.adhesion[j] * c.affinity[(size_t)i * c.T + tj] * c.affinity[(size_t)j * c.T + ti]);

	const R force = (rep - adh) / distance;
	for (int k = 0; k < D; k++)
		v[k] += force * diff[k];
	return pressure_term;
}

// repulsion term of a pair only (the simple_pressure contribution, position_solver.cpp:41-51)
template <int D, class R>
__device__ __forceinline__ R pair_pressure_term(const SimCtx<R>& c, int i, int j, const R* pi)
{
	const R* pj = c.pos + (size_t)j * D;
	R distance = pair_distance<D>(pi, pj);
	distance = real_max<R>(distance, (R)0.00001);
	R rep = 1 - distance / (c.radius[i] + c.radius[j]);
	rep = rep < 0 ? (R)0 : rep;
	return rep * rep;
}

// update_membrane_velocity (solver_helper.h:14-28)
template <class R>
__device__ __forceinline__ void membrane_push(R position, R wall, R sign, R radius, R strength, R& velocity)
{
	R distance = fabs(wall - position);
	distance = real_max<R>(distance, (R)0.00001);
	R rep = 1 - distance / radius;
	rep = rep < 0 ? (R)0 : rep;
	rep *= rep * strength * sign;
	velocity += rep * distance;
}

// Inputs of update_motility_single for one agent, loaded ahead of use.
template <int D, class R>
struct motility_in
{
	bool motile;
	R m[D], dir[D];
	R persistence, bias, speed;
};

template <int D, class R>
__device__ __forceinline__ motility_in<D, R> load_motility(const SimCtx<R>& c, int i)
{
	motility_in<D, R> in;
	in.motile = c.motile[i] != 0;
	for (int k = 0; k < D; k++)
	{
		in.m[k] = c.motility[(size_t)i * D + k];
		in.dir[k] = c.bias_dir[(size_t)i * D + k];
	}
	in.persistence = c.persistence[i];
	in.bias = c.bias[i];
	in.speed = c.speed[i];
	return in;
}

// normalize (solver_helper.h:30-45)
template <int D, class R>
__device__ __forceinline__ void normalize_dir(R* v)
{
	R length = 0;
	for (int k = 0; k < D; k++)
		length += v[k] * v[k];
	length = sqrt(length);
	for (int k = 0; k < D; k++)
		v[k] = length > zero_threshold<R>() ? v[k] / length : (R)0;
}

// cell::nearest_gradient (paraBioFVM agent): the gradients of the diffusion voxel that
// holds position p, for every substrate
template <int D, class R>
__device__ __forceinline__ const R* nearest_gradients(const SimCtx<R>& c, const R* p)
{
	int co[3] = { 0, 0, 0 };
	for (int k = 0; k < D; k++)
	{
		const R q = (p[k] - (R)c.bmin[k]) / (R)c.dvox[k];
		int v = (int)q;
		v = v < 0 ? 0 : (v >= c.dgrid[k] ? c.dgrid[k] - 1 : v);
		co[k] = v;
	}
	const size_t v = (size_t)co[0] + (size_t)c.dgrid[0] * ((size_t)co[1] + (size_t)c.dgrid[1] * co[2]);
	return c.grad + v * c.S * D;
}

// The migration bias direction updates the reference ships (standard_models.cpp:542-583):
// chemotaxis_function and advanced_chemotaxis_function<normalize_each_gradient>.
template <int D, class R>
__device__ __forceinline__ void update_bias_direction(const SimCtx<R>& c, int i, int code, R* dir)
{
	R p[D];
	for (int k = 0; k < D; k++)
		p[k] = c.pos[(size_t)i * D + k];
	const R* g = nearest_gradients<D, R>(c, p);
	if (code == PCM_BIAS_CHEMOTAXIS)
	{
		const R sign = (R)c.chemo_dir[i];
		const R* gs = g + (size_t)c.chemo_index[i] * D;
		for (int k = 0; k < D; k++)
			dir[k] = gs[k] * sign;
	}
	else
	{
		for (int k = 0; k < D; k++)
			dir[k] = 0;
		for (int s = 0; s < c.S; s++)
		{
			R gs[D];
			for (int k = 0; k < D; k++)
				gs[k] = g[(size_t)s * D + k];
			if (code == PCM_BIAS_ADVANCED_CHEMOTAXIS_NORMALIZED)
				normalize_dir<D, R>(gs);
			const R w = c.sensitivities[(size_t)i * c.S + s];
			for (int k = 0; k < D; k++)
				dir[k] += gs[k] * w;
		}
	}
	normalize_dir<D, R>(dir);
}

// update_motility_single (position_solver.cpp:115-145) for one agent.  The 3D walk is
// planar because the reference passes the restrict_to_2d array pointer as the bool.
template <int D, class R>
__device__ __forceinline__ void motility_apply(const SimCtx<R>& c, int i, unsigned long long step,
											   const motility_in<D, R>& in, R* v, bool writer)
{
	if (!in.motile)
		return;
	R m[D];
	for (int k = 0; k < D; k++)
		m[k] = in.m[k];
	const pcm_u32x4 blk = pcm_block(c.seed, step, PCM_STREAM_MOTILITY, (uint32_t)i, 0u, 0u);
	const R u0 = pcm_draw<R>::from(blk, 0);
	if (u0 < c.dt / in.persistence)
	{
		R walk[D];
		const R u1 = pcm_draw<R>::from(blk, 1);
		if constexpr (D == 1)
			walk[0] = u1 < (R)0.5 ? (R)-1 : (R)1;
		else
		{
			const R two_pi = (R)2 * (R)3.14159265358979323846;
			const R theta = pcm_scale_uniform<R>(u1, (R)0, two_pi);
			walk[0] = cos(theta);
			walk[1] = sin(theta);
			if constexpr (D == 3)
				walk[2] = 0;
		}
		R dir[D];
		for (int k = 0; k < D; k++)
			dir[k] = in.dir[k];
		if (c.chemotaxis)
		{
			// update_migration_bias_direction_f[i] (position_solver.cpp:132-136)
			const int code = c.bias_fn[i];
			if (code != PCM_BIAS_NONE)
			{
				update_bias_direction<D, R>(c, i, code, dir);
				if (writer)
					for (int k = 0; k < D; k++)
						c.bias_dir[(size_t)i * D + k] = dir[k];
			}
		}
		const R b = in.bias;
		for (int k = 0; k < D; k++)
			m[k] = (1 - b) * walk[k] + b * dir[k];
		R length;
		if constexpr (D == 1)
			length = fabs(m[0]);
		else if constexpr (D == 2)
			length = sqrt(m[0] * m[0] + m[1] * m[1]);
		else
			length = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
		for (int k = 0; k < D; k++)
		{
			m[k] = length > zero_threshold<R>() ? m[k] * in.speed / length : (R)0;
			if (writer)
				c.motility[(size_t)i * D + k] = m[k];
		}
	}
	for (int k = 0; k < D; k++)
		v[k] += m[k];
}

template <int D, class R>
__device__ __forceinline__ void motility_agent(const SimCtx<R>& c, int i, unsigned long long step, R* v)
{
	if (c.motile[i] == 0)
		return;
	motility_apply<D, R>(c, i, step, load_motility<D, R>(c, i), v, true);
}

// position_helper<D>::update_membrane_velocities (solver_helper.h:103-110, 184-193, 288-299)
template <int D, class R>
__device__ __forceinline__ void membrane_apply(const SimCtx<R>& c, R r, R s, const R* p, R* v)
{
	for (int k = 0; k < D; k++)
	{
		const int upper = (D == 1) ? c.bmax[1] : c.bmax[k];
		membrane_push<R>(p[k], (R)c.bmin[k], (R)1, r, s, v[k]);
		membrane_push<R>(p[k], (R)upper, (R)-1, r, s, v[k]);
	}
}

template <int D, class R>
__device__ __forceinline__ void membrane_agent(const SimCtx<R>& c, int i, const R* p, R* v)
{
	const R r = c.radius[i];
	const R s = c.bm_repulsion[i];
	for (int k = 0; k < D; k++)
	{
		const int upper = (D == 1) ? c.bmax[1] : c.bmax[k];
		membrane_push<R>(p[k], (R)c.bmin[k], (R)1, r, s, v[k]);
		membrane_push<R>(p[k], (R)upper, (R)-1, r, s, v[k]);
	}
}

// standard_elastic_contact_function (standard_models.cpp:643-654) of agent i towards j:
// the pull sqrt(k_i k_j a_i[t_j] a_j[t_i]) (x_j - x_i), the product taken left to right.
template <int D, class R>
__device__ __forceinline__ R contact_strength(const SimCtx<R>& c, R ki, const R* aff_row_i, int ti, int j, int tj)
{
	const R kij = ki * c.elastic[j] * aff_row_i[tj] * c.affinity[(size_t)j * c.T + ti];
	return (R)sqrt((double)kij); // correctly rounded in R either way
}

// evaluate_interactions (simulator.cpp:65-82) for one agent, one thread: the attached cells
// in list order
template <int D, class R>
__device__ __forceinline__ void contact_agent(const SimCtx<R>& c, int i, R* v)
{
	const int na = c.atc_cnt[i];
	const int* row = c.atc_idx + (size_t)i * c.LA;
	const R ki = c.elastic[i];
	const int ti = c.type[i];
	const R* aff_row = c.affinity + (size_t)i * c.T;
	R pi[D];
	for (int k = 0; k < D; k++)
		pi[k] = c.pos[(size_t)i * D + k];
	for (int s = 0; s < na; s++)
	{
		const int j = row[s];
		const R a = contact_strength<D, R>(c, ki, aff_row, ti, j, c.type[j]);
		for (int k = 0; k < D; k++)
			v[k] += a * (c.pos[(size_t)j * D + k] - pi[k]);
	}
}

// spring_contract_function (position_solver.cpp:337-370) for one agent
template <int D, class R>
__device__ __forceinline__ void contract_agent(const SimCtx<R>& c, int i, const R* p, R* v)
{
	const int cnt = c.spr_cnt[i];
	const int ti = c.type[i];
	const R ki = c.elastic[i];
	const R ai_row = 0;
	(void)ai_row;
	for (int s = 0; s < cnt; s++)
	{
		const int j = c.spr_idx[(size_t)i * c.L + s];
		const int tj = c.type[j];
		const R kij = ki * c.elastic[j] * c.affinity[(size_t)i * c.T + tj] * c.affinity[(size_t)j * c.T + ti];
		const R adhesion = (R)sqrt((double)kij);
		const R* pj = c.pos + (size_t)j * D;
		for (int k = 0; k < D; k++)
			v[k] += adhesion * (pj[k] - p[k]);
	}
}

// update_geometry (interactions_solver.cpp:89-96): double arithmetic as in the reference
template <class R>
__device__ __forceinline__ void update_geometry(const SimCtx<R>& c, int i)
{
	const double pi_r = (double)(R)3.14159265358979323846;
	const double sphere = pi_r * 4.0 / 3.0;
	c.radius[i] = (R)cbrt((double)c.volume[i] / sphere);
	c.nuc_radius[i] = (R)cbrt((double)c.nuclear[i] / sphere);
	c.surface[i] = (R)(4.0 * pi_r * (double)c.radius[i] * (double)c.radius[i]);
}

// ingest (interactions_solver.cpp:8-27, 57-67, 98-112)
template <class R>
__device__ __forceinline__ void ingest(const SimCtx<R>& c, int lhs, int rhs)
{
	c.cyto_fluid[lhs] += c.fluid[rhs];
	c.cyto_solid[lhs] += c.solid[rhs];
	c.fluid[lhs] += c.fluid[rhs];
	c.solid[lhs] += c.solid[rhs];
	c.cyto[lhs] += c.fluid[rhs] + c.solid[rhs];
	c.volume[lhs] += c.fluid[rhs] + c.solid[rhs];
	c.fluid_frac[lhs] = c.fluid[lhs] / (c.volume[lhs] + zero_threshold<R>());
	c.c2n[lhs] = c.cyto_solid[lhs] / (c.nuc_solid[lhs] + zero_threshold<R>());
	for (int k = 0; k < c.S; k++)
		c.internal[(size_t)lhs * c.S + k] += c.internal[(size_t)rhs * c.S + k] * c.frac_ingest[(size_t)rhs * c.S + k];
	update_geometry<R>(c, lhs);
	c.flags[rhs] = PCM_FLAG_TO_REMOVE;
}

// fuse (interactions_solver.cpp:29-55, 69-87, 114-134)
template <class R>
__device__ __forceinline__ void fuse(const SimCtx<R>& c, int lhs, int rhs)
{
	const int D = c.dims;
	const R both = c.volume[lhs] + c.volume[rhs];
	for (int k = 0; k < D; k++)
		c.pos[(size_t)lhs * D + k] =
			(c.pos[(size_t)lhs * D + k] * c.volume[lhs] + c.pos[(size_t)rhs * D + k] * c.volume[rhs]) / both;
	c.cyto_fluid[lhs] += c.cyto_fluid[rhs];
	c.nuc_fluid[lhs] += c.nuc_fluid[rhs];
	c.cyto_solid[lhs] += c.cyto_solid[rhs];
	c.nuc_solid[lhs] += c.nuc_solid[rhs];
	c.target_cyto[lhs] += c.target_cyto[rhs];
	c.target_nuc[lhs] += c.target_nuc[rhs];
	c.fluid[lhs] = c.nuc_fluid[lhs] + c.cyto_fluid[lhs];
	c.solid[lhs] = c.nuc_solid[lhs] + c.cyto_solid[lhs];
	c.nuclear[lhs] = c.nuc_fluid[lhs] + c.nuc_solid[lhs];
	c.cyto[lhs] = c.cyto_fluid[lhs] + c.cyto_solid[lhs];
	c.volume[lhs] = c.nuclear[lhs] + c.cyto[lhs];
	c.fluid_frac[lhs] = c.fluid[lhs] / (c.volume[lhs] + zero_threshold<R>());
	c.c2n[lhs] = c.cyto_solid[lhs] / (c.nuc_solid[lhs] + zero_threshold<R>());
	for (int k = 0; k < c.S; k++)
		c.internal[(size_t)lhs * c.S + k] += c.internal[(size_t)rhs * c.S + k];
	update_geometry<R>(c, lhs);
	c.nuclei[lhs] += c.nuclei[rhs];
	c.flags[rhs] = PCM_FLAG_TO_REMOVE;
}

} // namespace pcm

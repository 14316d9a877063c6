// pcm_kernels.h — gfx950 kernels of the mechanics step (included once, by pcm_lib.hip).
//
// Every per-agent kernel takes the simulation contexts in device memory and uses
// blockIdx.y as the simulation index, so one launch can advance a batch of simulations;
// blocks past a simulation's agent count return at once.  Every kernel returns when the
// simulation's status word is non-zero (an aborted step).
#pragma once

#include <type_traits>

#include "pcm_device.h"

namespace pcm {

constexpr int BLOCK = 256;
constexpr int RESOLVE_THREADS = 1024;
constexpr int CAND_CHUNK = 4096;    // interaction candidates ordered per LDS pass
constexpr int BAD_CHUNK = 8192;     // removed agents ordered in LDS
constexpr int LANE_BUF = 8;         // appends a lane buffers before one wavefront-wide reservation
constexpr int VELOCITY_ROUNDS = 1;  // list slots per lane gathered together
constexpr int MAX_EVENT_LANES = 16; // largest lane count of the event-drawing kernels
constexpr int CLAIM_THREADS = 1024; // block of the voxel claim kernel

__device__ __forceinline__ int gid() { return blockIdx.x * blockDim.x + threadIdx.x; }



// Launch geometry.  A plain launch gives each simulation one grid row: slot blockIdx.y, its
// blocks along x.  A batch's flattened launch (bmap set) has a single row holding the
// blocks of all members back to back, so no block is spent on a member smaller than the
// largest: entry b of bmap is block b's {member slot, block index within the member, the
// member's block count, 0}, and {-1, ...} pads the grid (such a block returns at once).
// A block entry is the same for every thread of the block, so the exit is block-uniform.
#define PCM_GEOMETRY                                                                                                   \
	int slot_ = (int)blockIdx.y, bid_ = (int)blockIdx.x, nblk_ = (int)gridDim.x;                                      \
	if (bmap)                                                                                                          \
	{                                                                                                                  \
		const int4 e_ = bmap[blockIdx.x];                                                                              \
		if (e_.x < 0)                                                                                                  \
			return;                                                                                                    \
		slot_ = e_.x;                                                                                                  \
		bid_ = e_.y;                                                                                                   \
		nblk_ = e_.z;                                                                                                  \
	}                                                                                                                  \
	(void)nblk_;

// global thread index within the block's simulation
#define GID (bid_ * (int)blockDim.x + (int)threadIdx.x)

// Each wavefront reads the status word and returns on its own.  That is only safe where the
// word cannot change during the launch, or where the block's wavefronts share nothing
// (no LDS, no barrier): a kernel that can raise the status itself while its blocks are
// still being dispatched uses PCM_PROLOGUE_BLOCK instead.
#define PCM_PROLOGUE(c)                                                                                                \
	PCM_GEOMETRY                                                                                                       \
	const SimCtx<R>& c = ctxs[slot_];                                                                                 \
	if (c.sc->status)                                                                                                  \
		return;

// Block-uniform exit, for kernels whose wavefronts cooperate through LDS or barriers and
// whose own blocks can raise the status word.  With a per-wavefront test, another block's
// raise landing between two wavefronts' reads lets one wavefront leave while the other
// stays: the one that stays then reads LDS the other was to fill (stencil offsets, scan
// partials) and indexes global memory with it.  Thread 0 reads the word once and every
// wavefront decides on that one value.
#define PCM_PROLOGUE_BLOCK(c)                                                                                          \
	PCM_GEOMETRY                                                                                                       \
	const SimCtx<R>& c = ctxs[slot_];                                                                                 \
	{                                                                                                                  \
		__shared__ int s_status_;                                                                                      \
		if (threadIdx.x == 0)                                                                                          \
			s_status_ = c.sc->status;                                                                                  \
		__syncthreads();                                                                                               \
		if (s_status_)                                                                                                 \
			return;                                                                                                    \
	}

// The neighbor builds ignore the one bit they raise themselves (neighbor overflow), so their
// exit decision cannot change during their launch and a per-wavefront test is uniform over
// each block.  Every block then finishes its lists (rows hold their first K entries) and
// max_neighbors is the longest list of the whole population, which the host's regrowth
// needs.  After an overflow the rest of the step's kernels return, so a neighbor build of a
// later step in the same call recomputes the same lists from the same mesh, in bounds.
#define PCM_PROLOGUE_NEIGHBORS(c)                                                                                      \
	PCM_GEOMETRY                                                                                                       \
	const SimCtx<R>& c = ctxs[slot_];                                                                                 \
	if (c.sc->status & ~ST_NEIGHBOR_OVERFLOW)                                                                          \
		return;

// ------------------------------------------------------------------ step bookkeeping
template <class R>
__global__ void k_begin_step(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	if (GID == 0)
		c.sc->step++;
}

// every member's scalars into one array (a batch reads all of them with one copy)
template <class R>
__global__ void k_gather_scalars(const SimCtx<R>* __restrict__ ctxs, Scalars* __restrict__ out)
{
	if (threadIdx.x == 0)
		out[blockIdx.y] = *ctxs[blockIdx.y].sc;
}

// ------------------------------------------------------------------ counting sort into voxels
// containers_solver::update_mechanics_mesh (containers_solver.cpp:15-56).  The reference's
// single-thread scatter (fetch_sub from the voxel count) leaves every bucket in descending
// agent order; the device reproduces that order exactly: histogram -> bucket allocation ->
// unordered scatter -> rank by "number of bucket members with a larger index".

// One agent's detachment draws for the spring update keyed by `step` (one per list slot)
// and its cleared erase mask.  A batch launches the spring kernels when any member has
// springs on; a member with them off fires nothing and draws nothing (its erase pass then
// only stores the step counter).
template <class R>
__device__ __forceinline__ void spring_fire_agent(const SimCtx<R>& c, int i, unsigned long long step)
{
	const int cnt = c.springs ? c.spr_cnt[i] : 0;
	const R rate = c.det_rate[i] * c.dt;
	uint32_t mask = 0;
	for (int s = 0; s < cnt; s++)
	{
		const R u = pcm_uniform01<R>(c.seed, step, PCM_STREAM_DETACH, (uint32_t)i, (uint32_t)s, 0u);
		if (u <= rate)
			mask |= 1u << s;
	}
	c.fire[i] = mask;
	c.erase[i] = 0;
}

// Whether agent i's non-fired slot s (partner p, fired masks fi and fp, p's list lp of
// length cp) is erased by the partner's firing.
template <class R>
__device__ __forceinline__ bool erased_by_partner(const int* li, int cnt, int s, int i, int p, uint32_t fi, const int* lp,
												  int cp, uint32_t fp)
{
	if (i < p)
	{
		int f_i = 0, nf_rank = 0;
		for (int u = 0; u < cnt; u++)
			if (li[u] == p)
			{
				const bool fu = (fi >> u) & 1u;
				f_i += fu;
				if (!fu && u < s)
					nf_rank++;
			}
		int g = 0, occ = 0;
		for (int u = 0; u < cp; u++)
			if (lp[u] == i)
			{
				occ++;
				if (occ > f_i && ((fp >> u) & 1u))
					g++;
			}
		return nf_rank < g;
	}
	int rank = 0;
	for (int u = 0; u < s; u++)
		rank += li[u] == p;
	int f_p = 0;
	for (int u = 0; u < cp; u++)
		f_p += lp[u] == i && ((fp >> u) & 1u);
	return rank < f_p;
}

// k_spring_erase's decisions for agent i's slots lane, lane + ERASE_LANES, ..., as the
// fused step's mesh update makes them for the next step (after the fired slots are drawn):
// the lane's partners, then their fired masks, are requested together; the list rules run
// only for the rare slots whose partner fired
constexpr int ERASE_LANES = 4;

template <class R>
__device__ __forceinline__ void spring_erase_lanes(const SimCtx<R>& c, int i, int lane, int n)
{
	constexpr int U = 32 / ERASE_LANES; // slots per lane at the largest list width
	const int cnt = c.springs ? c.spr_cnt[i] : 0;
	if (lane >= cnt)
		return; // the erase mask was cleared with the fired masks
	const int* li = c.spr_idx + (size_t)i * c.L;
	const uint32_t fi = c.fire[i];
	int ps[U];
	uint32_t fps[U];
#pragma unroll
	for (int u = 0; u < U; u++)
	{
		const int s = lane + u * ERASE_LANES;
		ps[u] = s < cnt ? li[s] : -1;
	}
#pragma unroll
	for (int u = 0; u < U; u++)
		fps[u] = ps[u] >= 0 && ps[u] < n ? c.fire[ps[u]] : 0u;
	uint32_t erase = 0;
#pragma unroll
	for (int u = 0; u < U; u++)
	{
		const int s = lane + u * ERASE_LANES, p = ps[u];
		if (s >= cnt || p < 0 || p >= n)
			continue;
		if ((fi >> s) & 1u)
			erase |= 1u << s;
		else if (fps[u] != 0
				 && erased_by_partner<R>(li, cnt, s, i, p, fi, c.spr_idx + (size_t)p * c.L, c.spr_cnt[p], fps[u]))
			erase |= 1u << s;
	}
	if (erase)
		atomicOr(&c.erase[i], erase);
}

template <class R>
__global__ void k_bin_clear(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	const int i = GID;
	if (i < c.sc->n_binned)
		c.vcnt[c.vox_of[i]] = 0;
	if (i == 0)
		c.sc->n_occ = 0;
}

template <class R, int D>
__global__ void k_bin_count(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	const int i = GID;
	const int n = c.sc->n;
	if (i == 0)
		c.sc->n_binned = n;
	if (i >= n)
		return;
	const int v = voxel_index<R>(c, c.pos + (size_t)i * D);
	c.vox_of[i] = v;
	c.vrank[i] = atomicAdd(&c.vcnt[v], 1);
}

// Each occupied voxel's first agent claims the voxel's bucket range.  Claims are summed
// per block of CLAIM_THREADS so that a block issues one 64-bit atomic on a packed counter
// (agents in the high word, occupied voxels in the low word): device-scope atomics on one
// cache line serialise across the XCDs, so there are as few of them as possible.
// RESTORE: also return the container update's remap entries to the identity (the link
// edits that read them have finished; the next step's interaction pass zeroes the update's
// counters)
template <class R, bool RESTORE>
__global__ void __launch_bounds__(CLAIM_THREADS) k_bin_claim(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	const int i = GID;
	const int n = c.sc->n;
	if (RESTORE)
	{
		if (i < c.sc->bad_count)
			c.remap[c.bad[i]] = c.bad[i];
		if (i < c.sc->moves)
			c.remap[c.move_src[i]] = c.move_src[i];
	}
	// blocks past the agents skip the block-wide scans (block 0 always claims, for n = 0)
	if (bid_ > 0 && (int)bid_ * CLAIM_THREADS >= n)
		return;
	int amount = 0, v = -1;
	if (i < n && c.vrank[i] == 0)
	{
		v = c.vox_of[i];
		amount = c.vcnt[v];
	}
	int total_amount, total_voxels;
	const int pre_amount = block_exclusive_scan(amount, total_amount);
	const int pre_voxels = block_exclusive_scan(amount > 0 ? 1 : 0, total_voxels);
	__shared__ unsigned long long s_base;
	if (threadIdx.x == 0)
		s_base = total_voxels
					 ? atomicAdd(&c.sc->claim, ((unsigned long long)total_amount << 32) | (unsigned)total_voxels)
					 : 0ull;
	__syncthreads();
	if (amount > 0)
	{
		c.vstart[v] = (int)(s_base >> 32) + pre_amount;
		c.occ[(int)(s_base & 0xffffffffull) + pre_voxels] = v;
	}
}

template <class R>
__global__ void k_bin_scatter(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	const int i = GID;
	if (i >= c.sc->n)
		return;
	if (i == 0)
	{
		// the claims have finished
		c.sc->n_occ = (int)(c.sc->claim & 0xffffffffull);
		c.sc->claim = 0;
	}
	c.btmp[c.vstart[c.vox_of[i]] + c.vrank[i]] = i;
}

// RECORDS: also write the agents' force records for the step's gather kernels (a mesh
// update for the neighbor lists alone has no force-record buffer).  DETACH: the mesh update
// of a fused step also decides the next step's erased spring slots (k_spring_erase's work;
// k_bin_scatter drew the fired slots)
// With DETACH the launch has the member's ordering blocks followed by its erase blocks
// (ERASE_LANES lanes per agent).  FIXED (the fused mesh update with voxel tables, after k_compact_count
// with FIXED): the voxel's members come from its unordered table, the occupied-voxel count
// is taken from the claim counter, and the compaction's remap entries are returned to the
// identity (the ordering blocks then cover the removal list too).
template <class R, int D, bool RECORDS = true, bool DETACH = false, bool FIXED = false>
__global__ void k_bin_order(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	const int n = c.sc->n;
	const int restore = FIXED ? c.sc->bad_count : 0; // moves <= removals
	if (DETACH)
	{
		const int order_blocks = (max(n, restore) + BLOCK - 1) / BLOCK;
		if (bid_ >= order_blocks)
		{
			const long long t = (long long)(bid_ - order_blocks) * blockDim.x + threadIdx.x;
			const int a = (int)(t / ERASE_LANES);
			if (a < n)
				spring_erase_lanes<R>(c, a, (int)(t % ERASE_LANES), n);
			return;
		}
	}
	const int i = GID;
	if (FIXED)
	{
		if (i == 0)
		{
			c.sc->n_occ = (int)(c.sc->claim & 0xffffffffull);
			c.sc->claim = 0;
		}
		if (i < restore)
		{
			const int x = c.bad[i];
			c.remap[x] = x;
		}
		if (i < c.sc->moves)
		{
			const int x = c.move_src[i];
			c.remap[x] = x;
		}
	}
	if (i >= n)
		return;
	// the agent's own data first: its loads overlap the rank chain below
	if (RECORDS)
		write_force_record<R, D>(c, i);
	R rec[4];
	for (int k = 0; k < 3; k++)
		rec[k] = k < D ? c.pos[(size_t)i * D + k] : (R)0;
	rec[3] = c.rel_adh[i] * c.radius[i];
	const int v = c.vox_of[i];
	const int s = c.vstart[v], m = c.vcnt[v];
	const int* members = FIXED ? c.vmem + (size_t)v * c.B : c.btmp + s;
	int rank = 0;
#pragma unroll 8
	for (int k = 0; k < m; k++)
		rank += members[k] > i;
	const int slot = s + rank;
	c.bucket[slot] = i;
	R* out = c.sorted + (size_t)slot * 4;
	for (int k = 0; k < 4; k++)
		out[k] = rec[k];
}

// force records of the current agent state, when the mesh update's copy is stale
template <class R, int D>
__global__ void k_force_records(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	const int i = GID;
	if (i < c.sc->n)
		write_force_record<R, D>(c, i);
}

// ------------------------------------------------------------------ neighbor lists
// position_solver::update_cell_neighbors (position_solver.cpp:79-97, 180-195, 262-286)
// with the 3^d stencil of common_solver::for_each_in_mech_neighborhood
// (common_solver.h:77-105): z outer, y, x inner, each bucket in bucket order, self
// skipped.  Lists are built for every agent; a non-movable agent's list only feeds
// simple_pressure and is reported empty.
//
// One wavefront per occupied voxel: the voxel's 3^d stencil buckets are staged once in
// LDS in stencil order, then every agent of the voxel tests 64 candidates per iteration;
// a ballot and a prefix popcount write the passing candidates in candidate order.
template <class R>
constexpr int nbr_tile()
{
	// candidates staged per block (a multiple of 64): at tissue densities a 3D stencil holds
	// about 240 candidates and up to ~370 (a 30 um voxel, radii up to 11 um); larger ones take
	// the per-lane path from HBM
	return 448;
}

// far-away records staged after the candidates (one 64-candidate chunk)
constexpr int NBR_PAD = 64;

// staged candidates per block for sparse populations (a multiple of 64): at a few agents per
// occupied voxel a 3D stencil holds a few dozen candidates
constexpr int NBR_TILE_SPARSE = 128;

// a coordinate whose square overflows: no staged padding record is within reach
template <class R>
constexpr R far_coordinate()
{
	return sizeof(R) == 4 ? (R)1e30 : (R)1e300;
}

// pair_within for a staged candidate record (position, rel_adh * radius) against the
// agent's: the same rounding, with within_reach's decision written so that the common
// outcomes need no branch (the square root runs only for lanes within the rounding band)
template <class R>
__device__ __forceinline__ bool within_reach_lanes(R s, R reach)
{
	// each comparison is used in one polarity only (no negated twins); for a NaN s or r2
	// the result is false, as within_reach's is
	const R e = sizeof(R) == 4 ? (R)1e-5 : (R)1e-12;
	const R r2 = reach * reach;
	const bool small = r2 < (sizeof(R) == 4 ? (R)1.2e-38 : (R)2.3e-308); // zero or subnormal threshold
	const bool below_hi = s <= r2 * ((R)1 + e);
	const bool above_lo = s >= r2 * ((R)1 - e);
	bool in = below_hi;
	if (small || (below_hi && above_lo))
		in = sqrt(s) <= reach;
	return in;
}

template <int D, class R>
__device__ __forceinline__ bool candidate_within(const R* pi, const R* q)
{
	if constexpr (D == 1)
		return fabs(pi[0] - q[0]) <= pi[3] + q[3];
	else
	{
		const R d0 = pi[0] - q[0], d1 = pi[1] - q[1];
		R s = d0 * d0 + d1 * d1;
		if constexpr (D == 3)
		{
			const R d2 = pi[2] - q[2];
			s = s + d2 * d2;
		}
		return within_reach_lanes<R>(s, pi[3] + q[3]);
	}
}

// The reach test of a staged candidate without a branch in the common case.  d is
// |pi - q|^2 - reach^2 with the reference's rounded coordinate differences and reach, the
// squares summed and the difference taken with fused multiply-adds.  Against the
// reference's decision, sqrt(s) <= reach with s rounded in its own order: |s_fma - s| is at
// most 6u*s (u the unit roundoff), and sqrt(s) <= reach holds exactly when s is below
// (reach + ulp/2)^2, within reach^2 * (1 + 2u); the rounding of d adds u*|d|.  Near the
// boundary s <= 2 reach^2, so the sign of d is the reference's decision whenever |d| > 16u
// reach^2.  `band_limit` is 64u times an upper bound of reach^2 for every candidate of the
// agent, plus a floor far above the subnormal range: lanes with |d| <= band_limit (rare)
// take the exact test instead.  A padding record's coordinate overflows s to +inf, so d is
// +inf: out, and not in the band.
template <class R>
__device__ __forceinline__ R fused_mul_add(R a, R b, R c)
{
	if constexpr (sizeof(R) == 4)
		return __builtin_fmaf(a, b, c);
	else
		return __builtin_fma(a, b, c);
}

template <int D, class R>
__device__ __forceinline__ R reach_excess(const R* pi, const R* q, R band_limit, bool& band)
{
	const R d0 = pi[0] - q[0], d1 = pi[1] - q[1];
	R s;
	if constexpr (D == 3)
	{
		const R d2 = pi[2] - q[2];
		s = fused_mul_add<R>(d0, d0, fused_mul_add<R>(d1, d1, d2 * d2));
	}
	else
		s = fused_mul_add<R>(d0, d0, d1 * d1);
	const R reach = pi[3] + q[3];
	const R d = fused_mul_add<R>(-reach, reach, s);
	band = (d < (R)0 ? -d : d) <= band_limit;
	return d;
}

// the band limit of an agent whose reach term is mi, among candidates whose largest reach
// term is rmax (rmax >= mi, the agent being among them)
template <class R>
__device__ __forceinline__ R reach_band_limit(R mi, R rmax)
{
	const R u = sizeof(R) == 4 ? (R)5.9604644775390625e-8 : (R)1.1102230246251565e-16;
	const R floor_ = sizeof(R) == 4 ? (R)1e-30 : (R)1e-290;
	const R r = mi + rmax;
	return (R)64 * u * (r * r) * (R)1.001 + floor_;
}

// One agent's list from its 3^d stencil, in stencil order and bucket order: position pi and
// reach mi = rel_adh * radius of agent i, voxel coordinates co.  DIRECT: the buckets may
// predate the current positions, so candidates are tested with their current positions and
// radii (what the reference does with stale buckets); otherwise the voxel-sorted copy made
// by the mesh update is used.
template <class R, int D, bool DIRECT>
__device__ __forceinline__ void neighbors_from(const SimCtx<R>& c, int i, const R* pi, R mi, const int* co)
{
	int* out = c.nbr_idx + (size_t)i * c.K;
	int cnt = 0;
	for (int z = -1; z <= 1; z++)
	{
		if (co[2] + z >= c.grid[2] || co[2] + z < 0)
			continue;
		for (int y = -1; y <= 1; y++)
		{
			if (co[1] + y >= c.grid[1] || co[1] + y < 0)
				continue;
			for (int x = -1; x <= 1; x++)
			{
				if (co[0] + x >= c.grid[0] || co[0] + x < 0)
					continue;
				const int nv = (co[0] + x) + c.grid[0] * ((co[1] + y) + c.grid[1] * (co[2] + z));
				const int m = c.vcnt[nv];
				if (m == 0)
					continue;
				const int s = c.vstart[nv];
				for (int k = s; k < s + m; k++)
				{
					const int j = c.bucket[k];
					if (j == i)
						continue;
					R q[4] = { 0, 0, 0, 0 };
					if (DIRECT)
					{
						for (int d = 0; d < D; d++)
							q[d] = c.pos[(size_t)j * D + d];
						q[3] = c.rel_adh[j] * c.radius[j];
					}
					else
						for (int d = 0; d < 4; d++)
							q[d] = c.sorted[(size_t)k * 4 + d];
					if (pair_within<D, R>(pi, q, mi + q[3]))
					{
						if (cnt < c.K)
							out[cnt] = j;
						cnt++;
					}
				}
			}
		}
	}
	if (cnt > c.K)
	{
		atomicOr(&c.sc->status, ST_NEIGHBOR_OVERFLOW);
		atomicMax(&c.sc->max_neighbors, cnt);
		cnt = c.K;
	}
	c.nbr_cnt[i] = cnt;
}

template <class R, int D, bool DIRECT>
__device__ __forceinline__ void neighbors_one_agent(const SimCtx<R>& c, int i)
{
	R pi[3] = { 0, 0, 0 };
	for (int k = 0; k < D; k++)
		pi[k] = c.pos[(size_t)i * D + k];
	const R mi = c.rel_adh[i] * c.radius[i];
	int co[3];
	voxel_coords<R>(c, DIRECT ? voxel_index<R>(c, c.pos + (size_t)i * D) : c.vox_of[i], co);
	neighbors_from<R, D, DIRECT>(c, i, pi, mi, co);
}

// the path for stencils larger than the LDS tile, out of line so that its registers do
// not count against the tiled path; with 7 wavefronts per SIMD (72 VGPRs) the tiled path
// keeps its loop-invariant lane values in registers instead of scratch
template <class R, int D>
__device__ __noinline__ void neighbors_one_agent_large_stencil(const SimCtx<R>& c, int i)
{
	neighbors_one_agent<R, D, false>(c, i);
}

// TILE: candidates staged per block (nbr_tile, or NBR_TILE_SPARSE for populations with a
// few agents per occupied voxel, whose stencils are short: a smaller tile lets more blocks
// share a CU's LDS)
template <class R, int D, int W, int TILE = nbr_tile<R>()>
__global__ void __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 7 : 6))) k_neighbors(
	const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE_NEIGHBORS(c)
	__shared__ R s_rec[(TILE + NBR_PAD) * 4]; // candidates, then far-away records
	__shared__ int s_idx[TILE];
	__shared__ int s_pre[28];
	__shared__ int s_vs[27];
	__shared__ R s_rmax[W]; // largest reach (rel_adh * radius) among the staged candidates, per wavefront
	const int tid = threadIdx.x;
	const int lane = tid & 63;
	const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
	const int n_occ = c.sc->n_occ;
	const unsigned long long below = (1ull << lane) - 1ull;
	for (int w = bid_; w < n_occ; w += nblk_)
	{
		const int v = c.occ[w];
		int co[3];
		voxel_coords<R>(c, v, co);
		if (wave == 0)
		{
			int cnt = 0, vs = 0;
			if (lane < 27)
			{
				// the lane's stencil offsets are recomputed per voxel (the empty asm hides that
				// they are loop-invariant): kept live across the voxel loop they cost a register
				// spill per wavefront
				int l = lane;
				asm volatile("" : "+v"(l));
				const int x = co[0] + l % 3 - 1, y = co[1] + (l / 3) % 3 - 1, z = co[2] + l / 9 - 1;
				if (x >= 0 && x < c.grid[0] && y >= 0 && y < c.grid[1] && z >= 0 && z < c.grid[2])
				{
					const int nv = x + c.grid[0] * (y + c.grid[1] * z);
					cnt = c.vcnt[nv];
					vs = c.vstart[nv];
				}
			}
			int C;
			const int pre = wave_exclusive_scan(cnt, C);
			if (lane < 27)
			{
				s_pre[lane] = pre;
				s_vs[lane] = vs;
			}
			if (lane == 0)
				s_pre[27] = C;
		}
		__syncthreads();
		const int C = __builtin_amdgcn_readfirstlane(s_pre[27]);
		const int m = __builtin_amdgcn_readfirstlane(s_pre[14] - s_pre[13]); // the centre voxel is stencil entry 13
		if (C <= TILE)
		{
			// stage the stencil in stencil order.  A candidate's stencil entry is the last
			// entry starting at or below its position: first its row of three entries (the
			// number of row starts at or below it; lane k of every wavefront holds the start
			// of entry k, read back with readlane), then the entry within the row
			const int lane_pre = s_pre[lane < 28 ? lane : 27];
			R rmax = 0;
			for (int t = tid; t < C; t += 64 * W)
			{
				int row = 0;
#pragma unroll
				for (int r = 1; r < 9; r++)
					row += __builtin_amdgcn_readlane(lane_pre, 3 * r) <= t;
				const int e0 = 3 * row;
				const int sv = e0 + (s_pre[e0 + 1] <= t) + (s_pre[e0 + 2] <= t);
				const int g = s_vs[sv] + t - s_pre[sv];
				const int j = c.bucket[g];
				R rec[4];
				for (int k = 0; k < 4; k++)
					rec[k] = c.sorted[(size_t)g * 4 + k];
				s_idx[t] = j;
				for (int k = 0; k < 4; k++)
					s_rec[t * 4 + k] = rec[k];
				rmax = rec[3] > rmax ? rec[3] : rmax;
			}
			for (int o = 32; o > 0; o >>= 1)
			{
				const R other = __shfl_xor(rmax, o, 64);
				rmax = other > rmax ? other : rmax;
			}
			if (lane == 0)
				s_rmax[wave] = rmax;
			// far-away records after the candidates, so that the test below runs on every lane
			// of a 64-candidate chunk starting anywhere before C with no bounds branch (their
			// lanes are masked too)
			for (int t = C + tid; t < C + NBR_PAD; t += 64 * W)
			{
				for (int k = 0; k < 3; k++)
					s_rec[t * 4 + k] = far_coordinate<R>();
				s_rec[t * 4 + 3] = 0;
			}
			__syncthreads();
			const int own0 = __builtin_amdgcn_readlane(lane_pre, 13);
			R block_rmax = 0;
			for (int k = 0; k < W; k++)
				block_rmax = s_rmax[k] > block_rmax ? s_rmax[k] : block_rmax;
			// the voxel's z faces (exact in real_t: integers of at most 24 bits)
			const R z_lo = (R)c.bmin[2] + (R)co[2] * (R)c.vox[2];
			const R z_hi = z_lo + (R)c.vox[2];
			const int plane0 = __builtin_amdgcn_readlane(lane_pre, 9);  // first entry of the dz = 0 plane
			const int plane1 = __builtin_amdgcn_readlane(lane_pre, 18); // first entry of the dz = +1 plane
			for (int a = wave; a < m; a += W)
			{
				const int self_t = own0 + a;
				const int i = __builtin_amdgcn_readfirstlane(s_idx[self_t]);
				R pi[4];
				for (int k = 0; k < 4; k++)
					pi[k] = s_rec[self_t * 4 + k];
				int* out = c.nbr_idx + (size_t)i * c.K;
				int base = 0;
				// Stencil entries run z-major, so the dz = -1 and dz = +1 planes are the
				// first and last runs of candidates.  A plane farther from the agent than its
				// reach plus the largest candidate reach (plus a margin far above the
				// rounding of the voxel assignment) holds no neighbor and is skipped; the
				// candidates tested keep their order, so lists are unchanged.
				int t_begin = 0, t_end = C;
				if (D == 3)
				{
					const R bound = pi[3] + block_rmax + (R)0.01;
					if (pi[2] - z_lo > bound)
						t_begin = plane0;
					if (z_hi - pi[2] > bound)
						t_end = plane1;
					t_begin = __builtin_amdgcn_readfirstlane(t_begin);
					t_end = __builtin_amdgcn_readfirstlane(t_end);
				}
				const R band_limit = reach_band_limit<R>(pi[3], block_rmax);
				// the chunk of the agent's own slot (t_begin <= self_t < t_end: the pruned
				// planes never hold the agent's own voxel) and its bit there
				const int self_chunk = __builtin_amdgcn_readfirstlane(t_begin + ((self_t - t_begin) & ~63));
				const unsigned long long self_bit = 1ull << __builtin_amdgcn_readfirstlane(self_t - self_chunk);
				for (int t0 = t_begin; t0 < t_end; t0 += 64)
				{
					const int t = t0 + lane;
					// decision value: the candidate is within reach when it is <= 0
					R dec;
					if constexpr (D == 1)
						dec = candidate_within<D, R>(pi, s_rec + t * 4) ? (R)-1 : (R)1;
					else
					{
						bool band;
						dec = reach_excess<D, R>(pi, s_rec + t * 4, band_limit, band);
						if (__ballot(band)) // rare (uniform branch): lanes in the rounding band decide exactly
						{
							const bool exact = candidate_within<D, R>(pi, s_rec + t * 4);
							dec = band ? (exact ? (R)-1 : (R)1) : dec;
						}
					}
					// The agent's own slot is cleared in the one chunk that holds it.  Lanes
					// past t_end read the dz = +1 plane that was pruned (out of reach, for the
					// reason it was pruned: its face is farther than any reach, with a margin
					// above the voxel assignment's rounding) or the far-away padding, so the
					// test rejects them itself.
					unsigned long long mask = __ballot(dec <= (R)0);
					if (t0 == self_chunk)
						mask &= ~self_bit;
					if (__builtin_amdgcn_inverse_ballot_w64(mask))
					{
						// rank among the passing lanes below, plus base
						const unsigned slot = __builtin_amdgcn_mbcnt_hi(
							(unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, (unsigned)base));
						if (slot < (unsigned)c.K)
							out[slot] = s_idx[t];
					}
					base += __popcll(mask);
				}
				if (lane == 0)
				{
					if (base > c.K)
					{
						atomicOr(&c.sc->status, ST_NEIGHBOR_OVERFLOW);
						atomicMax(&c.sc->max_neighbors, base);
						base = c.K;
					}
					c.nbr_cnt[i] = base;
				}
			}
		}
		else
		{
			// stencil larger than the LDS tile: one lane per agent, straight from HBM
			const int s0 = c.vstart[v];
			for (int a = tid; a < m; a += 64 * W)
				neighbors_one_agent_large_stencil<R, D>(c, c.bucket[s0 + a]);
		}
		__syncthreads();
	}
}

// Stencils of a few dozen candidates (2D and 1D populations at tissue densities) leave a
// wavefront per agent mostly idle and spend a block per voxel on staging: one lane per agent
// instead, agents taken in bucket order, so that the lanes of a wavefront share stencil
// voxels and their candidate loads hit the same lines; the agent's own position and reach
// come from its voxel-sorted record.  Candidates, order and test are those of k_neighbors.
template <class R, int D>
__global__ void k_neighbors_lanes(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE_NEIGHBORS(c)
	static_assert(D < 3, "3D stencils use k_neighbors");
	constexpr int E = D == 1 ? 3 : 9; // stencil voxels, x fastest (the z row of a 2D grid is 0)
	const int k0 = GID;
	if (k0 >= c.sc->n_binned)
		return;
	const int i = c.bucket[k0];
	R pi[4];
	for (int d = 0; d < 4; d++)
		pi[d] = c.sorted[(size_t)k0 * 4 + d];
	int co[3];
	// the record holds the binned position exactly, so this is the agent's bucket voxel
	voxel_coords<R>(c, voxel_index<R>(c, pi), co);
	// every stencil voxel's count and start requested at once
	int cnt_e[E], start_e[E];
#pragma unroll
	for (int e = 0; e < E; e++)
	{
		const int x = co[0] + e % 3 - 1, y = D == 2 ? co[1] + e / 3 - 1 : 0;
		const bool in = x >= 0 && x < c.grid[0] && y >= 0 && y < c.grid[1];
		const int nv = in ? x + c.grid[0] * y : 0;
		cnt_e[e] = in ? c.vcnt[nv] : 0;
		start_e[e] = in ? c.vstart[nv] : 0;
	}
	int* out = c.nbr_idx + (size_t)i * c.K;
	int cnt = 0;
	const R mi = pi[3];
#pragma unroll
	for (int e = 0; e < E; e++)
	{
		const int s = start_e[e], m = cnt_e[e];
		for (int k = s; k < s + m; k++)
		{
			const int j = c.bucket[k];
			R q[4];
			for (int d = 0; d < 4; d++)
				q[d] = c.sorted[(size_t)k * 4 + d];
			if (j != i && pair_within<D, R>(pi, q, mi + q[3]))
			{
				if (cnt < c.K)
					out[cnt] = j;
				cnt++;
			}
		}
	}
	if (cnt > c.K)
	{
		atomicOr(&c.sc->status, ST_NEIGHBOR_OVERFLOW);
		atomicMax(&c.sc->max_neighbors, cnt);
		cnt = c.K;
	}
	c.nbr_cnt[i] = cnt;
}

template <class R, int D>
__global__ void k_neighbors_direct(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE_NEIGHBORS(c)
	const int i = GID;
	if (i >= c.sc->n)
		return;
	neighbors_one_agent<R, D, true>(c, i);
}

// ------------------------------------------------------------------ neighbor build without allocation
// The mesh and neighbor kernels above, run on caller-owned agent arrays with the mesh
// scratch in a fixed device area.  The context arrives as a kernel argument and is stored
// into that area for the kernels that follow; the step scalars restart and the voxel counts
// of the whole grid are cleared (the area keeps whatever the previous call left).
template <class R>
__global__ void k_noalloc_setup(const SimCtx<R> h, SimCtx<R>* __restrict__ dst, int* __restrict__ needed)
{
	const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (blockIdx.x == 0)
	{
		constexpr int words = (int)(sizeof(SimCtx<R>) / sizeof(int));
		static_assert(sizeof(SimCtx<R>) % sizeof(int) == 0, "context copied in 4-byte words");
		const int* src = (const int*)&h;
		int* out = (int*)dst;
		for (int w = threadIdx.x; w < words; w += blockDim.x)
			out[w] = src[w];
		if (threadIdx.x == 0)
		{
			Scalars s = {};
			s.n = h.cap;
			*h.sc = s;
			*needed = 0;
		}
	}
	for (long long v = t; v < h.vcount; v += (long long)gridDim.x * blockDim.x)
		h.vcnt[v] = 0;
}

// Lists of non-movable agents are reported empty (update_cell_neighbors_single lists movable
// agents only, position_solver.cpp:190-191); `needed` receives the longest list when some
// list did not fit.
template <class R>
__global__ void k_noalloc_finish(const SimCtx<R>* __restrict__ ctxs, int* __restrict__ needed)
{
	const SimCtx<R>& c = ctxs[0];
	const int i = gid();
	if (i == 0)
		*needed = (c.sc->status & ST_NEIGHBOR_OVERFLOW) ? c.sc->max_neighbors : 0;
	if (i < c.sc->n && c.movable && !c.movable[i])
		c.nbr_cnt[i] = 0;
}

// number of neighbors of i the reference would list (empty for non-movable agents)
template <class R>
__device__ __forceinline__ int visible_neighbors(const SimCtx<R>& c, int i)
{
	return c.movable[i] ? c.nbr_cnt[i] : 0;
}

// ------------------------------------------------------------------ velocity phases
// update_cell_forces (position_solver.cpp:15-22, 99-113, 158-178, 233-260).  Velocity is
// summed in neighbor-list order; simple_pressure is gathered: agent k receives the
// repulsion term of every pair (k, j) once for each movable member of the pair, which is
// what the reference's two-sided `simple_pressure[lhs/rhs] +=` accumulates.
template <int D, class R>
__device__ __forceinline__ R forces_agent(const SimCtx<R>& c, int i, const R* pi, R* v)
{
	const bool mov = c.movable[i] != 0;
	const int cnt = c.nbr_cnt[i];
	const int* nb = c.nbr_idx + (size_t)i * c.K;
	R p = 0;
	for (int s = 0; s < cnt; s++)
	{
		const int j = nb[s];
		const R term = mov ? pair_force<D, R>(c, i, j, pi, v) : pair_pressure_term<D, R>(c, i, j, pi);
		const R t = term * pressure_coefficient<R>();
		if (mov)
			p += t;
		if (c.movable[j])
			p += t;
	}
	return p;
}

template <class R, int D>
__global__ void k_forces(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	const int g = GID;
	if (g >= c.sc->n)
		return;
	const int i = g;
	R pi[D], v[D];
	for (int k = 0; k < D; k++)
	{
		pi[k] = c.pos[(size_t)i * D + k];
		v[k] = c.vel[(size_t)i * D + k];
	}
	c.pressure[i] = forces_agent<D, R>(c, i, pi, v);
	if (c.movable[i])
		for (int k = 0; k < D; k++)
			c.vel[(size_t)i * D + k] = v[k];
}

// evaluate_interactions (simulator.cpp:65-82): agents whose contact function is the standard
// elastic one; there is no movable gating, as in the reference
template <class R, int D>
__global__ void k_contact(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	const int i = GID;
	if (i >= c.sc->n || c.contact[i] != PCM_CONTACT_STANDARD_ELASTIC)
		return;
	R v[D];
	for (int k = 0; k < D; k++)
		v[k] = c.vel[(size_t)i * D + k];
	contact_agent<D, R>(c, i, v);
	for (int k = 0; k < D; k++)
		c.vel[(size_t)i * D + k] = v[k];
}

template <class R, int D>
__global__ void k_motility(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	const int i = GID;
	if (i >= c.sc->n || c.motile[i] == 0)
		return;
	R v[D];
	for (int k = 0; k < D; k++)
		v[k] = c.vel[(size_t)i * D + k];
	motility_agent<D, R>(c, i, c.sc->step, v);
	for (int k = 0; k < D; k++)
		c.vel[(size_t)i * D + k] = v[k];
}

template <class R, int D>
__global__ void k_membrane(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	const int i = GID;
	if (i >= c.sc->n || !c.movable[i])
		return;
	R p[D], v[D];
	for (int k = 0; k < D; k++)
	{
		p[k] = c.pos[(size_t)i * D + k];
		v[k] = c.vel[(size_t)i * D + k];
	}
	membrane_agent<D, R>(c, i, p, v);
	for (int k = 0; k < D; k++)
		c.vel[(size_t)i * D + k] = v[k];
}

template <class R, int D>
__global__ void k_contract(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	const int i = GID;
	if (i >= c.sc->n || !c.movable[i])
		return;
	R p[D], v[D];
	for (int k = 0; k < D; k++)
	{
		p[k] = c.pos[(size_t)i * D + k];
		v[k] = c.vel[(size_t)i * D + k];
	}
	contract_agent<D, R>(c, i, p, v);
	for (int k = 0; k < D; k++)
		c.vel[(size_t)i * D + k] = v[k];
}

// update_positions_internal (position_solver.cpp:481-504): Adams-Bashforth 2
template <class R, int D>
__global__ void k_positions(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	const int i = GID;
	if (i >= c.sc->n || !c.movable[i])
		return;
	const R factor = (R)((double)c.dt * 1.5);
	const R previous_factor = (R)((double)c.dt * -0.5);
	for (int k = 0; k < D; k++)
	{
		const size_t o = (size_t)i * D + k;
		const R v = c.vel[o];
		c.pos[o] += v * factor + c.pvel[o] * previous_factor;
		c.pvel[o] = v;
		c.vel[o] = 0;
	}
}

// solve_pair (position_solver.cpp:24-77) on force records; same operation order as
// pair_force.  Returns the repulsion term that feeds simple_pressure.
template <int D, class R>
__device__ __forceinline__ R pair_force_rec(const force_rec<R>& a, const force_rec<R>& b, R aff_ab, R aff_ba,
											  bool with_force, R* v)
{
	R diff[D];
	for (int k = 0; k < D; k++)
		diff[k] = a.p[k] - b.p[k];
	R distance;
	if constexpr (D == 1)
		distance = fabs(diff[0]);
	else if constexpr (D == 2)
		distance = sqrt(diff[0] * diff[0] + diff[1] * diff[1]);
	else
		distance = sqrt(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]);
	distance = real_max<R>(distance, (R)0.00001);
	R rep = 1 - distance / (a.r + b.r);
	rep = rep < 0 ? (R)0 : rep;
	rep *= rep;
	const R term = rep;
	if (with_force)
	{
		rep *= sqrt(a.rep * b.rep);
		const R reach = a.rel * a.r + b.rel * b.r;
		R adh = 1 - distance / reach;
		adh *= adh;
		adh *= sqrt(a.adh * b.adh * aff_ab * aff_ba);
		const R force = (rep - adh) / distance;
		for (int k = 0; k < D; k++)
			v[k] += force * diff[k];
	}
	return term;
}

// Lane SRC of each aligned group of G lanes, broadcast to the group (G <= 32): ds_swizzle
// in bit mode with and-mask ~(G - 1) and or-mask SRC, so no address register is needed
template <int G, int SRC>
__device__ __forceinline__ float group_bcast(float x)
{
	static_assert(G <= 32 && (G & (G - 1)) == 0 && SRC < G, "groups of a power of two lanes within 32");
	constexpr int pattern = (32 - G) | (SRC << 5);
	return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(x), pattern));
}

template <int G, int SRC>
__device__ __forceinline__ double group_bcast(double x)
{
	static_assert(G <= 32 && (G & (G - 1)) == 0 && SRC < G, "groups of a power of two lanes within 32");
	constexpr int pattern = (32 - G) | (SRC << 5);
	const long long b = __double_as_longlong(x);
	const int lo = __builtin_amdgcn_ds_swizzle((int)(b & 0xffffffffll), pattern);
	const int hi = __builtin_amdgcn_ds_swizzle((int)(b >> 32), pattern);
	return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// f(integral_constant<int, S>) for S = 0 .. G-1, unrolled at compile time
template <int G, int S = 0, class F>
__device__ __forceinline__ void for_each_lane(F&& f)
{
	if constexpr (S < G)
	{
		f(std::integral_constant<int, S> {});
		for_each_lane<G, S + 1>(f);
	}
}

// v[k] += the group's f[k] in lane order (the ordered sum of one slot chunk).  Lanes past
// the list hold -0 in f, so their terms leave v exactly as it is; the loop over the group's
// lanes is unrolled at compile time, so the broadcasts are requested ahead of the adds
// instead of one shuffle round trip per lane.
template <int G, int D, class R>
__device__ __forceinline__ void add_group_terms(R* v, const R* f)
{
	for_each_lane<G>([&](auto s) {
		for (int k = 0; k < D; k++)
			v[k] += group_bcast<G, decltype(s)::value>(f[k]);
	});
}

// Accepted attachments appended to agent a's spring list (the append step of
// update_spring_attachments_internal, position_solver.cpp:430-449): incoming partners in
// ascending actor order, then the agent's own accepted events in event order.  The G
// lanes of a group run the same loops; entry e is stored by lane e % G, which is the lane
// that reads it back in the contraction loop, and lane 0 stores the counters.  Returns the
// list length.  in_cnt, na and cnt are the agent's in_cnt, own_cnt and spr_cnt.
template <class R, int G>
__device__ __forceinline__ int spring_append_group(const SimCtx<R>& c, int a, int lane, int in_cnt, int na, int cnt,
												  bool coherent)
{
	const int nin = min(in_cnt, c.K);
	if (nin + na == 0)
		return cnt;
	int* out = c.spr_idx + (size_t)a * c.L;
	bool overflow = false;
	int last = -1;
	for (int found = 0; found < nin; found++)
	{
		int best = 0x7fffffff, best_k = 0;
		for (int q = 0; q < nin; q++)
		{
			const int x = c.inbox[2 * ((size_t)a * c.K + q)];
			if (x > last && x < best)
			{
				best = x;
				best_k = c.inbox[2 * ((size_t)a * c.K + q) + 1];
			}
		}
		last = best;
		if (event_state<R>(c, (size_t)best * c.K + best_k, coherent) == 1)
		{
			if (cnt < c.L)
			{
				if (cnt % G == lane)
					out[cnt] = best;
			}
			else
				overflow = true;
			cnt++;
		}
	}
	for (int k = 0; k < na; k++)
		if (event_state<R>(c, (size_t)a * c.K + k, coherent) == 1)
		{
			const int j = c.own_tgt[(size_t)a * c.K + k];
			if (cnt < c.L)
			{
				if (cnt % G == lane)
					out[cnt] = j;
			}
			else
				overflow = true;
			cnt++;
		}
	cnt = min(cnt, c.L);
	if (lane == 0)
	{
		if (overflow)
			atomicOr(&c.sc->status, ST_LIST_OVERFLOW);
		c.spr_cnt[a] = cnt;
		c.in_cnt[a] = 0;
		c.own_cnt[a] = 0;
	}
	return cnt;
}

// Forces, motility, basement membrane and spring contraction of one agent in the
// reference's summation order (simulator.cpp:173-176), written to the velocity; neighbor
// and spring-partner data come from the force records.  G lanes share an agent: lane s of
// a group evaluates list slots s, s+G, ... (gather + pair arithmetic), then every lane of
// the group adds the slot terms in list order, read through lane shuffles, so sums are
// rounded exactly as the serial reference sums them.  Motility and the membrane run redundantly on the group's lanes
// (lane 0 writes), spring contraction is spread over lanes like the neighbor loop.
// CLEAR: also clear the voxel counts of the last mesh update (nothing reads them until
// this step's mesh update)
// The step's later attach-resolve rounds (over the actors the first round left open) run
// first in the member's first block, which then marks them done; only the wavefronts whose
// agents have attachment events wait for the mark, just before the append.  The first
// block is dispatched before the member's other blocks, so waiting cannot starve it.
template <class R, int D, bool CLEAR, int G>
__global__ void __launch_bounds__(BLOCK) k_velocity_fused(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_GEOMETRY
	const SimCtx<R>& c = ctxs[slot_];
	const bool later_rounds = c.springs && c.sc->res_open > 0;
	const unsigned step_tag = (unsigned)c.sc->step;
	if (bid_ == 0)
	{
		__shared__ int s_go, s_next;
		if (threadIdx.x == 0)
			s_go = c.sc->status == 0 && c.springs ? (c.sc->res_open > 0 ? 1 : 2) : 0;
		__syncthreads();
		if (s_go == 1)
		{
			resolve_rounds<R>(c, &s_next, false);
			// publish the decisions before the done mark (release: the L2 written back; the
			// waiting wavefronts read the states with agent-scope loads)
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
			if (threadIdx.x == 0)
				__hip_atomic_store(&c.sc->res_done, step_tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		}
		else if (s_go == 2 && threadIdx.x == 0)
			c.sc->stat[1] = 1;
	}
	if (c.sc->status)
		return;
	constexpr int B = VELOCITY_ROUNDS;
	const int t = GID;
	if (CLEAR && t < c.sc->n_binned)
		c.vcnt[c.vox_of[t]] = 0;
	const int g = t / G;
	const int lane = t & (G - 1);
	if (g >= c.sc->n)
		return;
	const int i = g;
	// the agent's record, velocity, count and first list slots are requested together;
	// the motility, membrane and spring inputs are loaded where they are used, which keeps
	// the register count (and so the occupancy) of the neighbor loop low
	const force_rec<R> me = read_force_record<R>(c, i);
	const int cnt = c.nbr_cnt[i];
	const int* nb = c.nbr_idx + (size_t)i * c.K;
	int jn[B];
#pragma unroll
	for (int u = 0; u < B; u++)
		jn[u] = u * G + lane < c.K ? nb[u * G + lane] : -1; // in the row; used only below cnt
	R v[D];
	for (int k = 0; k < D; k++)
		v[k] = c.vel[(size_t)i * D + k];

	const bool mov = me.movable;
	R p[D];
	for (int k = 0; k < D; k++)
		p[k] = me.p[k];
	if (c.contacts && c.contact[i] == PCM_CONTACT_STANDARD_ELASTIC)
	{
		// evaluate_interactions (simulator.cpp:65-82) runs before the forces: one attached
		// cell per lane, the terms summed in list order
		const int na = c.atc_cnt[i];
		const int* row = c.atc_idx + (size_t)i * c.LA;
		const R ki = c.elastic[i];
		const R* aff_i = c.affinity + (size_t)i * c.T;
		for (int base = 0; base < na; base += G)
		{
			R f[D];
			for (int k = 0; k < D; k++)
				f[k] = -(R)0; // -0 + x is exactly x for every x
			if (base + lane < na)
			{
				const force_rec<R> o = read_force_record<R>(c, row[base + lane]);
				const R a = contact_strength<D, R>(c, ki, aff_i, me.type, row[base + lane], o.type);
				for (int k = 0; k < D; k++)
					f[k] = a * (o.p[k] - p[k]);
			}
			add_group_terms<G, D, R>(v, f);
		}
	}
	const type_row<R> aff_row = load_type_row<R>(c.affinity + (size_t)i * c.T, c.T);
	const R coefficient = pressure_coefficient<R>();
	R pressure = 0;
	for (int base = 0; base < cnt; base += B * G)
	{
		R f[B][D], mine[B], theirs[B];
		// the next iteration's slots are requested now, ahead of this iteration's gathers
		int jn_next[B];
#pragma unroll
		for (int u = 0; u < B; u++)
			jn_next[u] = base + B * G + u * G + lane < cnt ? nb[base + B * G + u * G + lane] : -1;
		force_rec<R> other[B];
		R aff_ba[B];
#pragma unroll
		for (int u = 0; u < B; u++)
			if (base + u * G + lane < cnt)
			{
				other[u] = read_force_record<R>(c, jn[u]);
				aff_ba[u] = c.affinity[(size_t)jn[u] * c.T + me.type];
			}
#pragma unroll
		for (int u = 0; u < B; u++)
		{
			for (int k = 0; k < D; k++)
				f[u][k] = -(R)0; // -0 + x is exactly x for every x
			mine[u] = 0;
			theirs[u] = 0;
			if (base + u * G + lane < cnt)
			{
				const R term = pair_force_rec<D, R>(me, other[u], aff_row.at(other[u].type), aff_ba[u], mov, f[u]);
				const R tp = term * coefficient;
				mine[u] = mov ? tp : (R)0;
				theirs[u] = other[u].movable ? tp : (R)0;
			}
		}
		// the slot terms in list order
		// (lanes past the list hold -0 forces and +0 pressure terms, and the pressure sum is
		// never -0, so adding them changes nothing)
#pragma unroll
		for (int u = 0; u < B; u++)
		{
			if (base + u * G >= cnt)
				continue;
			if (mov)
				add_group_terms<G, D, R>(v, f[u]);
			for_each_lane<G>([&](auto s) {
				pressure += group_bcast<G, decltype(s)::value>(mine[u]);
				pressure += group_bcast<G, decltype(s)::value>(theirs[u]);
			});
		}
#pragma unroll
		for (int u = 0; u < B; u++)
			jn[u] = jn_next[u];
	}
	if (lane == 0)
		c.pressure[i] = pressure;
	// the per-agent inputs of motility, membrane, spring append and contraction are
	// requested together and awaited once (the empty asm needs their values), instead of
	// one round trip per stage
	const motility_in<D, R> mot = load_motility<D, R>(c, i);
	const R bm_rep = c.wall ? c.bm_repulsion[i] : (R)0;
	int in_cnt = 0, own_cnt = 0, ns = 0;
	R ki = 0;
	if (c.springs)
	{
		in_cnt = c.in_cnt[i];
		own_cnt = c.own_cnt[i];
		ns = c.spr_cnt[i];
		ki = c.elastic[i];
	}
	asm volatile("" ::"v"(mot.speed), "v"(mot.persistence), "v"(mot.bias), "v"(bm_rep), "v"(in_cnt), "v"(own_cnt),
				 "v"(ns), "v"(ki));
	if (mot.motile)
		motility_apply<D, R>(c, i, c.sc->step, mot, v, lane == 0);
	if (mov && c.wall)
		membrane_apply<D, R>(c, me.r, bm_rep, p, v);
	// the step's attachments join the list first (no separate append launch in a step)
	if (c.springs)
	{
		const bool wait = later_rounds && __any(in_cnt + own_cnt > 0);
		if (wait)
			wait_for_resolve_rounds<R>(c, step_tag);
		ns = spring_append_group<R, G>(c, i, lane, in_cnt, own_cnt, ns, wait);
	}
	if (mov && c.springs)
	{
		// spring_contract_function (position_solver.cpp:337-370)
		const int* sp = c.spr_idx + (size_t)i * c.L;
		// each round's slots are requested one round ahead (after the append above: an
		// appended entry is read back by the lane that stored it)
		int js_next[B];
#pragma unroll
		for (int u = 0; u < B; u++)
			js_next[u] = u * G + lane < ns ? sp[u * G + lane] : -1;
		for (int base = 0; base < ns; base += B * G)
		{
			int js[B];
#pragma unroll
			for (int u = 0; u < B; u++)
			{
				js[u] = js_next[u];
				js_next[u] = base + B * G + u * G + lane < ns ? sp[base + B * G + u * G + lane] : -1;
			}
			R f[B][D];
			force_rec<R> other[B];
			R kj[B], aff_ba[B];
#pragma unroll
			for (int u = 0; u < B; u++)
				if (base + u * G + lane < ns)
				{
					other[u] = read_force_record<R>(c, js[u]);
					kj[u] = c.elastic[js[u]];
					aff_ba[u] = c.affinity[(size_t)js[u] * c.T + me.type];
				}
#pragma unroll
			for (int u = 0; u < B; u++)
			{
				for (int k = 0; k < D; k++)
					f[u][k] = -(R)0; // -0 + x is exactly x for every x
				if (base + u * G + lane < ns)
				{
					const R kij = ki * kj[u] * aff_row.at(other[u].type) * aff_ba[u];
					const R adhesion = (R)sqrt((double)kij);
					for (int k = 0; k < D; k++)
						f[u][k] = adhesion * (other[u].p[k] - p[k]);
				}
			}
#pragma unroll
			for (int u = 0; u < B; u++)
				if (base + u * G < ns)
					add_group_terms<G, D, R>(v, f[u]);
		}
	}
	if (lane == 0)
		for (int k = 0; k < D; k++)
			c.vel[(size_t)i * D + k] = v[k];
}

// ------------------------------------------------------------------ springs
// update_spring_attachments_internal (position_solver.cpp:374-449).
//
// Detachment: one draw per (agent, slot).  The single-thread reference visits agents in
// ascending order; a fired slot of agent a erases its own entry and the first remaining
// entry of a in the partner's list.  For a slot s of a holding partner p that gives:
//   a < p: erased iff s fired, or s is among the first g non-fired entries of p in a's
//          list, where g counts p's fired entries of a after the first f_a (f_a = fired
//          entries of p in a's list);
//   a > p: erased iff s fired, or s is among the first f_p entries of p in a's list
//          (f_p = fired entries of a in p's list).
// STARTS_STEP: the kernel opens a mechanics step, whose counter is one above the stored
// one until k_spring_resolve_first stores it (k_begin_step's work, without its launch)
template <class R, bool STARTS_STEP>
__global__ void k_spring_fire(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	const int i = GID;
	if (i == 0)
		c.sc->res_open = 0; // the resolver's open-actor list starts empty every update
	if (i >= c.sc->n)
		return;
	spring_fire_agent<R>(c, i, c.sc->step + (STARTS_STEP ? 1 : 0));
}

// one thread per (agent, slot)
template <class R>
__global__ void k_spring_erase(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	const long long t = (long long)bid_ * blockDim.x + threadIdx.x;
	const int i = (int)(t / c.L), s = (int)(t % c.L);
	const int n = c.sc->n;
	if (i >= n)
		return;
	// count, slot and fire mask in one round trip (the slot is inside the row's storage
	// and is used only below the count)
	const int* li = c.spr_idx + (size_t)i * c.L;
	const int cnt = c.spr_cnt[i];
	const int p = li[s];
	const uint32_t fi = c.fire[i];
	if (s >= cnt || p < 0 || p >= n)
		return;
	if ((fi >> s) & 1u)
	{
		atomicOr(&c.erase[i], 1u << s);
		return;
	}
	const uint32_t fp = c.fire[p];
	if (fp == 0)
		return; // both rules need a fired entry in p's list
	if (erased_by_partner<R>(li, cnt, s, i, p, fi, c.spr_idx + (size_t)p * c.L, c.spr_cnt[p], fp))
		atomicOr(&c.erase[i], 1u << s);
}

// std::remove of the erased entries (position_solver.cpp:404-411); runs in the attachment
// collection kernel, which reads neighbor lists only
template <class R>
__device__ __forceinline__ void spring_compact_agent(const SimCtx<R>& c, int i)
{
	const uint32_t e = c.erase[i];
	if (e == 0)
		return;
	int* li = c.spr_idx + (size_t)i * c.L;
	const int cnt = c.spr_cnt[i];
	int w = 0;
	for (int s = 0; s < cnt; s++)
		if (!((e >> s) & 1u))
			li[w++] = li[s];
	c.spr_cnt[i] = w;
}

// Attachment events: every (i, slot) with j = neighbor > i whose draws succeed.  An actor
// keeps its events' partners in slot order (own_tgt); each target collects (actor, k) references in
// an unordered inbox; actors with events are listed for the resolver.  G lanes share an
// agent: lane l draws for slots l, l+G, ...; a ballot ranks the events in slot order.
// STARTS_STEP: the step counter is one above the stored one (k_spring_resolve_first stores it)
template <class R, int G, bool STARTS_STEP>
__global__ void __launch_bounds__(BLOCK) k_spring_collect(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE_BLOCK(c)
	// a block past the agents (a batch's grid covers its largest member) owns no segment
	// and skips the block-wide scan below
	if ((int)bid_ * (BLOCK / G) >= c.sc->n)
		return;
	const int t = GID;
	const int g = t / G;
	const int lane = t & (G - 1);
	const int group_shift = (int)(threadIdx.x & 63) & ~(G - 1);
	const unsigned long long group_mask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
	int n_ev = 0;
	int i = -1;
	if (g < c.sc->n)
	{
		i = g;
		if (lane == 0)
			spring_compact_agent<R>(c, i);
		const int cnt = c.springs ? visible_neighbors<R>(c, i) : 0; // springs off: no attachment draws
		const int* nb = c.nbr_idx + (size_t)i * c.K;
		const int ti = c.type[i];
		const R rate_i = c.att_rate[i] * c.dt;
		const type_row<R> aff_i = load_type_row<R>(c.affinity + (size_t)i * c.T, c.T);
		const unsigned long long step = c.sc->step + (STARTS_STEP ? 1 : 0);
		int j_next = lane < cnt ? nb[lane] : -1; // each slot is requested one iteration ahead
		for (int base = 0; base < cnt; base += G)
		{
			const int s = base + lane;
			bool event = false;
			int j = j_next;
			j_next = s + G < cnt ? nb[s + G] : -1;
			if (s < cnt)
			{
				if (j > i)
				{
					const R p_l = rate_i * aff_i.at(c.type[j]);
					const R p_r = c.att_rate[j] * c.dt * c.affinity[(size_t)j * c.T + ti];
					const pcm_u32x4 b = pcm_block(c.seed, step, PCM_STREAM_ATTACH, (uint32_t)i, (uint32_t)s, 0u);
					event = pcm_draw<R>::from(b, 0) <= p_l || pcm_draw<R>::from(b, 1) <= p_r;
				}
			}
			const unsigned long long m = (__ballot(event) >> group_shift) & group_mask;
			if (event)
			{
				const int k = n_ev + __popcll(m & ((1ull << lane) - 1ull));
				c.own_tgt[(size_t)i * c.K + k] = j;
				c.own_st[(size_t)i * c.K + k] = 0;
				const int slot = atomicAdd(&c.in_cnt[j], 1);
				if (slot < c.K)
				{
					c.inbox[2 * ((size_t)j * c.K + slot)] = i;
					c.inbox[2 * ((size_t)j * c.K + slot) + 1] = k;
				}
				else
					atomicOr(&c.sc->status, ST_CAPACITY);
			}
			n_ev += __popcll(m);
		}
		if (lane == 0)
			c.own_cnt[i] = n_ev;
	}
	// actors are listed per block, with no device-wide counter (same-address atomics from
	// every XCD serialise): block b's actors fill att_list[b * APB, b * APB + blk_cnt[b])
	constexpr int APB = BLOCK / G;
	const bool lead = lane == 0 && n_ev > 0;
	int total;
	const int rank = block_exclusive_scan(lead ? 1 : 0, total);
	if (lead)
		c.att_list[bid_ * APB + rank] = i;
	// only blocks that cover agents own a segment (a batch's grid covers its largest member)
	if (threadIdx.x == 0 && (int)bid_ * APB < c.sc->n)
		c.blk_cnt[bid_] = total;
}

// One actor's undecided attachment events, in its event order, against the states decided
// so far (0 undecided, 1 success, 2 failure).  Returns whether an event stays undecided.
template <class R>
__device__ __forceinline__ bool resolve_actor(const SimCtx<R>& c, volatile uint8_t* st, int a, int na)
{
	const int c0a = c.spr_cnt[a], maxa = c.max_att[a];
	int in_lo = 0, in_hi = 0;
	const int nin = min(c.in_cnt[a], c.K);
	for (int q = 0; q < nin; q++)
	{
		const int x = c.inbox[2 * ((size_t)a * c.K + q)], kx = c.inbox[2 * ((size_t)a * c.K + q) + 1];
		const int sx = st[(size_t)x * c.K + kx];
		in_lo += sx == 1;
		in_hi += sx != 2;
	}
	bool open = false;
	int own_lo = 0, own_hi = 0;
	for (int k = 0; k < na; k++)
	{
		int sk = st[(size_t)a * c.K + k];
		if (sk == 0)
		{
			const int lo_a = c0a + in_lo + own_lo, hi_a = c0a + in_hi + own_hi;
			const int cond_a = lo_a >= maxa ? 2 : (hi_a < maxa ? 1 : 0);
			const int j = c.own_tgt[(size_t)a * c.K + k];
			int lo_j = c.spr_cnt[j], hi_j = lo_j;
			const int maxj = c.max_att[j];
			const int njin = min(c.in_cnt[j], c.K);
			for (int q = 0; q < njin; q++)
			{
				const int x = c.inbox[2 * ((size_t)j * c.K + q)];
				if (x >= a)
					continue;
				const int sx = st[(size_t)x * c.K + c.inbox[2 * ((size_t)j * c.K + q) + 1]];
				lo_j += sx == 1;
				hi_j += sx != 2;
			}
			const int cond_j = lo_j >= maxj ? 2 : (hi_j < maxj ? 1 : 0);
			if (cond_a == 2 || cond_j == 2)
				sk = 2;
			else if (cond_a == 1 && cond_j == 1)
				sk = 1;
			if (sk != 0)
				st[(size_t)a * c.K + k] = (uint8_t)sk;
			else
				open = true;
		}
		own_lo += sk == 1;
		own_hi += sk != 2;
	}
	return open;
}

// Attachment resolution in the reference's order: event (i, j) succeeds iff, when the
// single-thread loop reaches it (i ascending, neighbor order), both endpoints hold fewer
// springs than their maximum.  The events before it at endpoint i are all of i's inbox
// (their actors are smaller) and i's earlier own events; at endpoint j they are j's inbox
// entries with a smaller actor.  Each endpoint's count is bounded below by the successes
// and above by the successes plus the undecided among those events; rounds repeat over
// the actors that still hold undecided events until none is left (the earliest undecided
// event is always decidable).  Decisions only ever turn undecided into final, so a round
// may read states of the same round or older ones and stay correct.
//
// First round: every actor at once, one block per collect block's actor segment; actors
// left with undecided events go to att_c for the single-block rounds that follow.
// STARTS_STEP: also store the step counter the step's earlier kernels used (no thread of
// this kernel or of k_spring_resolve reads it)
template <class R, int G, bool STARTS_STEP>
__global__ void __launch_bounds__(64) k_spring_resolve_first(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	constexpr int APB = BLOCK / G;
	const int b = bid_;
	const int n = c.sc->n;
	if (b == 0 && threadIdx.x == 0)
	{
		c.sc->stat[7] = (n + APB - 1) / APB; // collect segments, for the event and actor counts
		if (STARTS_STEP)
			c.sc->step++;
	}
	if (b * APB >= n)
		return;
	const int m = c.blk_cnt[b];
	volatile uint8_t* st = c.own_st;
	int events = 0;
	for (int k = threadIdx.x; k < m; k += 64)
	{
		const int a = c.att_list[b * APB + k];
		const int na = c.own_cnt[a];
		events += na;
		if (resolve_actor<R>(c, st, a, na))
			c.att_c[atomicAdd(&c.sc->res_open, 1)] = a;
	}
	for (int o = 32; o > 0; o >>= 1)
		events += __shfl_xor(events, o);
	if (threadIdx.x == 0)
		c.blk_ev[b] = events;
}

// Later rounds, on one block of any size, over the actors the first round left open.
// clear_open: reset the open count when done (the fused step's velocity kernel keeps it:
// its wavefronts read it to know whether to wait for these rounds).
template <class R>
__device__ void resolve_rounds(const SimCtx<R>& c, int* s_next, bool clear_open)
{
	int count = c.sc->res_open;
	int* cur = c.att_c;
	int* nxt = c.att_c2;
	if (threadIdx.x == 0)
		*s_next = 0;
	__syncthreads();
	volatile uint8_t* st = c.own_st;
	int rounds = 1;
	while (count > 0)
	{
		if (rounds > (long long)count * c.K + 1)
		{
			if (threadIdx.x == 0)
				atomicOr(&c.sc->status, ST_CAPACITY);
			return;
		}
		for (int t = threadIdx.x; t < count; t += blockDim.x)
		{
			const int a = cur[t];
			if (resolve_actor<R>(c, st, a, c.own_cnt[a]))
				nxt[atomicAdd(s_next, 1)] = a;
		}
		__threadfence();
		__syncthreads();
		count = *s_next;
		rounds++;
		__syncthreads();
		if (threadIdx.x == 0)
			*s_next = 0;
		int* t = cur;
		cur = nxt;
		nxt = t;
		__syncthreads();
	}
	if (threadIdx.x == 0)
	{
		c.sc->stat[1] = rounds;
		if (clear_open)
			c.sc->res_open = 0;
	}
}

// The later rounds as a launch of their own (the springs phase; a fused step runs them in
// the velocity kernel's first block)
template <class R>
__global__ void __launch_bounds__(RESOLVE_THREADS) k_spring_resolve(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	(void)bmap; // one block per simulation: never flattened
	const SimCtx<R>& c = ctxs[blockIdx.y];
	__shared__ int s_next;
	if (c.sc->status)
		return;
	resolve_rounds<R>(c, &s_next, true);
}

// A wavefront whose agents have attachment events, in a step whose later resolve rounds
// run in the velocity kernel's first block, waits for that block's done mark (the step's
// counter), or for the step being abandoned (a raised status word); a bound on the wait
// raises ST_CAPACITY instead of hanging.
template <class R>
__device__ __forceinline__ void wait_for_resolve_rounds(const SimCtx<R>& c, unsigned tag)
{
	for (int spin = 0;; spin++)
	{
		if (__hip_atomic_load(&c.sc->res_done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == tag)
			break;
		if (__hip_atomic_load(&c.sc->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
			break;
		if (spin > (1 << 22))
		{
			atomicOr(&c.sc->status, ST_CAPACITY);
			break;
		}
		__builtin_amdgcn_s_sleep(8);
	}
}

// an attachment event's decided state; `coherent` reads past this XCD's L2 (states the
// velocity kernel's first block decided in this launch)
template <class R>
__device__ __forceinline__ int event_state(const SimCtx<R>& c, size_t k, bool coherent)
{
	if (coherent)
		return __hip_atomic_load(&c.own_st[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	return c.own_st[k];
}

// Successful springs are appended to each endpoint's list in the order the reference
// pushes them: events from smaller actors first (by actor), then the agent's own events in
// neighbor order.
template <class R>
__global__ void k_spring_append(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	const int a = GID;
	if (a >= c.sc->n)
		return;
	const int nin = min(c.in_cnt[a], c.K);
	const int na = c.own_cnt[a];
	if (nin + na == 0)
		return;
	int cnt = c.spr_cnt[a];
	int* out = c.spr_idx + (size_t)a * c.L;
	int last = -1;
	for (int found = 0; found < nin; found++)
	{
		int best = 0x7fffffff, best_k = 0;
		for (int q = 0; q < nin; q++)
		{
			const int x = c.inbox[2 * ((size_t)a * c.K + q)];
			if (x > last && x < best)
			{
				best = x;
				best_k = c.inbox[2 * ((size_t)a * c.K + q) + 1];
			}
		}
		last = best;
		if (c.own_st[(size_t)best * c.K + best_k] == 1)
		{
			if (cnt < c.L)
				out[cnt] = best;
			else
				atomicOr(&c.sc->status, ST_LIST_OVERFLOW);
			cnt++;
		}
	}
	for (int k = 0; k < na; k++)
		if (c.own_st[(size_t)a * c.K + k] == 1)
		{
			const int j = c.own_tgt[(size_t)a * c.K + k];
			if (cnt < c.L)
				out[cnt] = j;
			else
				atomicOr(&c.sc->status, ST_LIST_OVERFLOW);
			cnt++;
		}
	c.spr_cnt[a] = min(cnt, c.L);
	c.in_cnt[a] = 0;
	c.own_cnt[a] = 0;
}

// Block-wide bitonic sort of `count` 64-bit keys (with 32-bit payloads) in shared memory.
__device__ inline void block_sort_pairs(unsigned long long* key, int* val, int count)
{
	int P = 1;
	while (P < count)
		P <<= 1;
	for (int t = count + threadIdx.x; t < P; t += blockDim.x)
	{
		key[t] = ~0ull;
		val[t] = -1;
	}
	__syncthreads();
	for (int k = 2; k <= P; k <<= 1)
		for (int j = k >> 1; j > 0; j >>= 1)
		{
			for (int t = threadIdx.x; t < P; t += blockDim.x)
			{
				const int u = t ^ j;
				if (u > t)
				{
					const bool up = (t & k) == 0;
					const unsigned long long a = key[t], b = key[u];
					if ((a > b) == up)
					{
						key[t] = b;
						key[u] = a;
						const int tv = val[t];
						val[t] = val[u];
						val[u] = tv;
					}
				}
			}
			__syncthreads();
		}
}

// Block-wide count of records (stride ints apart, actor first) whose actor lies in [lo, hi).
__device__ inline int block_count_range(const int* rec, int stride, int count, int lo, int hi, int* red)
{
	int local = 0;
	for (int e = threadIdx.x; e < count; e += blockDim.x)
	{
		const int a = rec[stride * e];
		local += a >= lo && a < hi;
	}
	__syncthreads();
	if (threadIdx.x == 0)
		*red = 0;
	__syncthreads();
	atomicAdd(red, local);
	__syncthreads();
	const int r = *red;
	__syncthreads();
	return r;
}

// Largest hi in (lo, n] such that at most `chunk` records have their actor in [lo, hi).
__device__ inline int block_chunk_end(const int* rec, int stride, int count, int lo, int n, int chunk, int* red)
{
	if (block_count_range(rec, stride, count, lo, n, red) <= chunk)
		return n;
	int a = lo + 1, b = n;
	while (a < b)
	{
		const int mid = a + (b - a + 1) / 2;
		if (block_count_range(rec, stride, count, lo, mid, red) <= chunk)
			a = mid;
		else
			b = mid - 1;
	}
	return a;
}

// lane k's value of x (k wave-uniform)
template <class R>
__device__ __forceinline__ R lane_real(R x, int k)
{
	if constexpr (sizeof(R) == 4)
		return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), k));
	else
	{
		const long long b = __double_as_longlong(x);
		const int lo = __builtin_amdgcn_readlane((int)b, k);
		const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
		return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
	}
}

// ------------------------------------------------------------------ cell-cell interactions
// update_cell_cell_interactions_internal (interactions_solver.cpp:143-253).
//
// k_interact_tentative: every live actor walks its neighbor list, draws one number per
// neighbor, records the slots where any branch can fire (candidates) and plays the
// reference's branch logic against the flags as they stand before the phase.  The
// outcome equals the reference's order-dependent result unless agents interact through
// more than one event (an agent removed and also acting, attacked or removed twice).
// Every event is an atomic read-modify-write of one word per agent (removals targeting
// it, attacks targeting it, whether it acted): of two events on one agent the later one
// sees the earlier, so the pass itself raises the conflict flag, and then
// k_interact_serial replays all candidates in the reference's order on one thread
// instead of k_interact_apply.
enum : int
{
	EV_INGEST = 0,
	EV_FUSE = 2
};
constexpr unsigned long long EVW_REMOVE = 1ull;                 // bits 0-31: removals targeting the agent
constexpr unsigned long long EVW_REMOVE_MASK = 0xffffffffull;
constexpr unsigned long long EVW_ATTACK = 1ull << 32;           // bits 32-62: attacks targeting the agent
constexpr unsigned long long EVW_ACTED = 1ull << 63;            // the agent acted
__device__ __forceinline__ int evw_removals(unsigned long long w) { return (int)(w & EVW_REMOVE_MASK); }
__device__ __forceinline__ int evw_attacks(unsigned long long w) { return (int)((w >> 32) & 0x7fffffffull); }

// AB2: first apply the agent's position update (update_positions precedes the interactions
// and this pass reads no positions).  G lanes share an agent: lane l loads and draws for
// slots l, l+G, ...; the rare successful draws are then walked in slot order by the whole
// group (the per-actor state is group-uniform), reading the slot data through shuffles.
template <class R, int D, bool AB2, int G>
__global__ void __launch_bounds__(BLOCK) k_interact_tentative(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	const int t = GID;
	const int g = t / G;
	const int lane = t & (G - 1);
	const int group_shift = (int)(threadIdx.x & 63) & ~(G - 1);
	const unsigned long long group_mask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
	if (AB2 && t == 0)
	{
		// the previous step's container counters: their last readers (the remap restore of
		// the mesh update) have finished, and this step's apply pass counts from zero
		c.sc->bad_count = 0;
		c.sc->moves = 0;
	}
	if (g >= c.sc->n)
		return;
	const int i = g;
	// the actor's inputs are requested first: the compiler cannot move loads past the
	// position update's stores, so they would cost a second round trip after it.
	// Every branch compares a draw u in [0,1) with `<` against rate * dt, so an actor whose
	// rates are all <= 0 never acts and its draws need not be made (bitwise ors, not
	// short-circuits: the rates are requested together)
	bool any_rate = c.dead_phago[i] > 0;
	for (int k = 0; k < c.T; k++)
		any_rate |= (c.live_phago[(size_t)i * c.T + k] > 0) | (c.attack[(size_t)i * c.T + k] > 0)
					| (c.fusion[(size_t)i * c.T + k] > 0);
	any_rate &= c.dead[i] != 1;
	const bool movable = c.movable[i];
	const int listed = c.nbr_cnt[i];
	const int ti = c.type[i];
	const bool self_removed = c.flags[i] == PCM_FLAG_TO_REMOVE;
	if (AB2 && lane < D && movable)
	{
		const R factor = (R)((double)c.dt * 1.5);
		const R previous_factor = (R)((double)c.dt * -0.5);
		const size_t o = (size_t)i * D + lane;
		const R v = c.vel[o];
		c.pos[o] += v * factor + c.pvel[o] * previous_factor;
		c.pvel[o] = v;
		c.vel[o] = 0;
	}
	if (lane == 0)
		c.atk_tgt[i] = -1;
	int n_rem = 0;
	bool acted = false;
	if (any_rate)
	{
		const int cnt = movable ? listed : 0; // visible_neighbors
		const int* nb = c.nbr_idx + (size_t)i * c.K;
		const R dt = c.dt;
		const unsigned long long step = c.sc->step;
		const R p_dead = c.dead_phago[i] * dt;
		bool attacked_once = false, phagocytosed_once = false, fused_once = false, stop = false;
		int j_next = lane < cnt ? nb[lane] : -1; // each slot is requested one iteration ahead
		for (int base = 0; base < cnt && !stop; base += G)
		{
			const int s = base + lane;
			int j = j_next;
			j_next = s + G < cnt ? nb[s + G] : -1;
			// bit 0 candidate, 1 dead target, 2 u < p_live, 3 u < p_attack, 4 u < p_fuse,
			// 5 target flagged for removal
			int bits = 0;
			if (s < cnt)
			{
				const int tj = c.type[j];
				const bool dead_j = c.dead[j] == 1;
				const R p_live = c.live_phago[(size_t)i * c.T + tj] * dt;
				const R p_attack = c.attack[(size_t)i * c.T + tj] * c.immuno[(size_t)j * c.T + ti] * dt;
				const R p_fuse = c.fusion[(size_t)i * c.T + tj] * dt;
				// no branch can fire when its threshold is <= 0: u < threshold is false
				if (dead_j ? p_dead > 0 : ((p_live > 0) | (p_attack > 0) | (p_fuse > 0)))
				{
					const R u = pcm_uniform01<R>(c.seed, step, PCM_STREAM_INTERACT, (uint32_t)i, (uint32_t)s, 0u);
					const bool lt_live = u < p_live, lt_attack = u < p_attack, lt_fuse = u < p_fuse;
					const bool candidate = dead_j ? (u < p_dead) : (lt_live || lt_attack || lt_fuse);
					if (candidate)
						bits = 1 | (dead_j ? 2 : 0) | (lt_live ? 4 : 0) | (lt_attack ? 8 : 0) | (lt_fuse ? 16 : 0)
							   | (c.flags[j] == PCM_FLAG_TO_REMOVE ? 32 : 0);
				}
			}
			unsigned long long m = (__ballot(bits != 0) >> group_shift) & group_mask;
			while (m)
			{
				const int src = __ffsll((long long)m) - 1;
				m &= m - 1;
				const int b = __shfl(bits, src, G);
				const int jj = __shfl(j, src, G);
				const int ss = base + src;
				if (lane == 0)
				{
					const int k = atomicAdd(&c.sc->cand_count, 1);
					if (k < c.cand_cap)
					{
						c.cand[2 * k] = i;
						c.cand[2 * k + 1] = ss;
					}
					else
						atomicOr(&c.sc->status, ST_CAPACITY);
				}
				int kind = -1; // -1 none, 0 ingest, 1 attack, 2 fuse
				if (b & 2)
					kind = 0;
				else if (!phagocytosed_once && (b & 4))
					kind = 0;
				else if (!attacked_once && (b & 8))
					kind = 1;
				else if (!fused_once && (b & 16))
					kind = 2;
				if (kind < 0)
					continue;
				acted = true;
				if (self_removed)
				{
					stop = true;
					break;
				}
				if (b & 32)
					continue;
				if (kind == 1)
				{
					if (lane == 0)
					{
						c.atk_tgt[i] = jj;
						// an attack conflicts with a removal of the same target
						if (atomicAdd(&c.ev_t[jj], EVW_ATTACK) & EVW_REMOVE_MASK)
							c.sc->conflict = 1;
					}
					attacked_once = true;
					continue;
				}
				if (!(b & 2))
				{
					if (kind == 0)
						phagocytosed_once = true;
					else
						fused_once = true;
				}
				if (lane == 0)
				{
					// a removal conflicts with any other event on its target
					if (atomicAdd(&c.ev_t[jj], EVW_REMOVE) != 0)
						c.sc->conflict = 1;
					if (n_rem < 4)
						c.rem_ev[4 * i + n_rem] = ss | ((kind == 2 ? EV_FUSE : EV_INGEST) << 24);
					else
						c.sc->conflict = 1;
				}
				n_rem++;
			}
		}
	}
	if (lane == 0)
	{
		c.rem_cnt[i] = n_rem;
		// an actor conflicts with a removal targeting it
		if (acted && (atomicOr(&c.ev_t[i], EVW_ACTED) & EVW_REMOVE_MASK))
			c.sc->conflict = 1;
	}
}

// out of the box of the microenvironment mesh (containers_solver.cpp:89-96), with no
// early exit so that the coordinates are requested together
template <class R, int D>
__device__ __forceinline__ bool outside_box(const SimCtx<R>& c, const R* p)
{
	bool out = false;
	for (int k = 0; k < D; k++)
		out |= (p[k] < (R)c.bmin[k]) | (p[k] > (R)c.bmax[k]);
	return out;
}

// One agent's events (i = its global thread).  FLAG: also decide, from the agent's final
// state, whether the container update removes it, and append it to the removal list (the
// speculation is exact here, so an agent with a removal event targeting it is removed).
// Every lane of a wavefront calls it; returns whether the lane stored what the container
// plan reads: a removal entry, or the removal flag of an agent it ingested or fused.
template <class R, int D, bool FLAG>
__device__ __forceinline__ bool interact_apply_agent(const SimCtx<R>& c, int i)
{
	if ((i & ~63) >= c.sc->n)
		return false; // the whole wavefront is past the agents: nothing to reserve
	const bool live_thread = i < c.sc->n;
	bool removed = false, flagged = false;
	if (live_thread)
	{
	const int nr = c.rem_cnt[i];
	flagged = nr > 0;
	const unsigned long long w = c.ev_t[i];
	// the removal test's inputs are requested with the counters, in one round trip; only
	// the actor's own fusion moves it, so then its position is read again after its events
	// (another actor's event on it makes it a removal target, removed whatever its flag)
	R p[D];
	uint8_t fl = 0;
	if (FLAG)
	{
		fl = c.flags[i];
		for (int k = 0; k < D; k++)
			p[k] = c.pos[(size_t)i * D + k];
	}
	for (int k = 0; k < nr; k++)
	{
		const int e = c.rem_ev[4 * i + k];
		const int j = c.nbr_idx[(size_t)i * c.K + (e & 0xFFFFFF)];
		if ((e >> 24) == EV_FUSE)
			fuse<R>(c, i, j);
		else
			ingest<R>(c, i, j);
	}
	const int na = evw_attacks(w);
	if (na > 0)
	{
		// attackers, ascending: the reference applies attacks in actor order.  Each pass
		// is branch-free and unrolled, so eight list entries and their targets are
		// requested per round trip (a serial scan made the attacked agents' wavefronts the
		// kernel's tail)
		const int cnt = c.nbr_cnt[i];
		const int* nb = c.nbr_idx + (size_t)i * c.K;
		int last = -1;
		for (int found = 0; found < na; found++)
		{
			int best = 0x7fffffff;
#pragma unroll 8
			for (int s = 0; s < cnt; s++)
			{
				const int k = nb[s];
				const bool hit = (k > last) & (k < best) & (c.atk_tgt[k] == i);
				best = hit ? k : best;
			}
			if (best == 0x7fffffff)
				break;
			c.damage[i] += c.damage_rate[best] * c.dt;
			c.attack_time[i] += c.dt;
			last = best;
		}
	}
	if (FLAG)
	{
		if (nr > 0)
			for (int k = 0; k < D; k++)
				p[k] = c.pos[(size_t)i * D + k];
		removed = (evw_removals(w) > 0) | (fl == PCM_FLAG_TO_REMOVE) | outside_box<R, D>(c, p);
	}
	if (w != 0) // written by few agents: the others' lines stay clean
		c.ev_t[i] = 0;
	}
	if (FLAG)
	{
		const int k = wave_reserve(&c.sc->bad_count, removed ? 1 : 0);
		if (removed)
			c.bad[k] = i;
	}
	return removed || flagged;
}

// Agent i's removal entry when the speculative pass found a conflict: the apply pass does
// not run, so the entry reflects the state before the serial replay (removal flag set, or
// outside the box); the replay then settles the agents its ingestions and fusions touched.
// Every lane of a wavefront calls it.
template <class R, int D>
__device__ __forceinline__ bool interact_flag_agent(const SimCtx<R>& c, int i)
{
	if ((i & ~63) >= c.sc->n)
		return false;
	bool removed = false;
	if (i < c.sc->n)
	{
		R p[D];
		for (int k = 0; k < D; k++)
			p[k] = c.pos[(size_t)i * D + k];
		removed = (c.flags[i] == PCM_FLAG_TO_REMOVE) | outside_box<R, D>(c, p);
	}
	const int k = wave_reserve(&c.sc->bad_count, removed ? 1 : 0);
	if (removed)
		c.bad[k] = i;
	return removed;
}

template <class R, int D, bool FLAG>
__global__ void k_interact_apply(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	if (c.sc->conflict)
		return;
	interact_apply_agent<R, D, FLAG>(c, GID);
}

// Exact replay of the reference loop over the candidates, in (actor, slot) order, when the
// speculative pass found a dependency.  Candidates are ordered CAND_CHUNK at a time.
template <class R>
__device__ void compact_plan_block(const SimCtx<R>& c, int* sb, int sb_cap, int* s_red, int b, int n0, int limit, int log);

// LDS of the serial pass: candidates ordered per pass, removed agents ordered at once
template <int CAND, int BADC, bool PLAN>
struct serial_lds
{
	unsigned long long key[CAND];
	int val[CAND];
	int count, red;
	int rescan; // the replay touched more agents than sb lists: the plan rescans every agent
	int sb[PLAN ? BADC : 1];
};

// PLAN: then build the removal list when the replay ran (the apply pass built it otherwise)
// and plan the container update in the same block.  Runs on one block of any size.
template <class R, int D, bool PLAN, int CAND, int BADC>
__device__ void interact_serial_block(const SimCtx<R>& c, serial_lds<CAND, BADC, PLAN>& sh)
{
	// every scalar the pass reads, requested together
	const int status = c.sc->status, cand_count = c.sc->cand_count, conflict = c.sc->conflict;
	const int bad0 = *(volatile int*)&c.sc->bad_count, n0 = c.sc->n;
	const int limit = c.sc->compact_limit, log0 = c.sc->log_len;
	if (status)
		return;
	const int C = min(cand_count, c.cand_cap);
	const bool replay = conflict != 0;
	// PLAN: the agents the replay's ingestions and fusions touch, listed in sh.sb (wavefront 0)
	int touched = 0;
	bool touch_overflow = false;
	if (replay && C > 0)
	{
		const int n = c.sc->n;
		const R dt = c.dt;
		const unsigned long long step = c.sc->step;
		// per-actor state survives chunk boundaries
		int cur = -1;
		bool broken = false, attacked_once = false, phagocytosed_once = false, fused_once = false;
		int lo = 0;
		while (lo < n)
		{
			const int hi = block_chunk_end(c.cand, 2, C, lo, n, CAND, &sh.red);
			if (threadIdx.x == 0)
				sh.count = 0;
			__syncthreads();
			for (int e = threadIdx.x; e < C; e += blockDim.x)
			{
				const int ac = c.cand[2 * e];
				if (ac >= lo && ac < hi)
				{
					const int k = atomicAdd(&sh.count, 1);
					if (k < CAND)
					{
						sh.key[k] = ((unsigned long long)ac << 32) | (unsigned)c.cand[2 * e + 1];
						sh.val[k] = e;
					}
				}
			}
			__syncthreads();
			// one actor's candidates (at most one per list slot) always fit a pass unless the
			// rows are longer than the pass
			if (sh.count > CAND)
			{
				if (threadIdx.x == 0)
					atomicOr(&c.sc->status, ST_CAPACITY);
				return;
			}
			const int m = sh.count;
			block_sort_pairs(sh.key, sh.val, m);
			// Wavefront 0 walks the ordered candidates 64 at a time.  Each lane first loads one
			// candidate's inputs (target, draw, the four branch thresholds, the damage
			// increment, both removal flags, the target's damage and attack time as they stand
			// now), so that the walk runs on registers in uniform control flow instead of a few
			// dependent global loads per candidate.  Of what the decisions read, only the
			// removal flags change during the replay (ingestion and fusion flag their target);
			// attacks add to their target's damage and attack time.  After each such event every
			// lane that names the same agent takes the new value.  Thresholds are computed as
			// the reference's expressions are (rate * dt, attack * immunogenicity * dt).
			if (threadIdx.x < 64)
			{
				const int lane = threadIdx.x;
				const volatile uint8_t* vflags = c.flags;
				const volatile R* vdamage = c.damage;
				const volatile R* vattack_time = c.attack_time;
				for (int g0 = 0; g0 < m; g0 += 64)
				{
					const int q = g0 + lane;
					int ai = -1, aj = -1, dead_j = 0, fl_i = 0, fl_j = 0;
					R u = 1, p_dead = 0, p_live = 0, p_attack = 0, p_fuse = 0, dr = 0, dmg = 0, atk = 0;
					if (q < m)
					{
						ai = (int)(sh.key[q] >> 32);
						const int slot = (int)(sh.key[q] & 0xffffffffu);
						aj = c.nbr_idx[(size_t)ai * c.K + slot];
						const int ti = c.type[ai], tj = c.type[aj];
						u = pcm_uniform01<R>(c.seed, step, PCM_STREAM_INTERACT, (uint32_t)ai, (uint32_t)slot, 0u);
						dead_j = c.dead[aj];
						p_dead = c.dead_phago[ai] * dt;
						p_live = c.live_phago[(size_t)ai * c.T + tj] * dt;
						p_attack = c.attack[(size_t)ai * c.T + tj] * c.immuno[(size_t)aj * c.T + ti] * dt;
						p_fuse = c.fusion[(size_t)ai * c.T + tj] * dt;
						dr = c.damage_rate[ai] * dt;
						fl_i = vflags[ai];
						fl_j = vflags[aj];
						dmg = vdamage[aj];
						atk = vattack_time[aj];
					}
					const int qn = min(64, m - g0);
					for (int k = 0; k < qn; k++)
					{
						const int i = __builtin_amdgcn_readlane(ai, k);
						if (i != cur)
						{
							cur = i;
							broken = attacked_once = phagocytosed_once = fused_once = false;
						}
						if (broken)
							continue;
						const int j = __builtin_amdgcn_readlane(aj, k);
						const R uk = lane_real<R>(u, k);
						const bool removed_i = __builtin_amdgcn_readlane(fl_i, k) == PCM_FLAG_TO_REMOVE;
						const bool removed_j = __builtin_amdgcn_readlane(fl_j, k) == PCM_FLAG_TO_REMOVE;
						// 0: ingest, 1: attack, 2: fuse
						int event = -1;
						if (__builtin_amdgcn_readlane(dead_j, k) == 1)
						{
							if (!(uk < lane_real<R>(p_dead, k)))
								continue;
							event = 0;
						}
						else if (!phagocytosed_once && uk < lane_real<R>(p_live, k))
						{
							event = 0;
							if (!removed_i && !removed_j)
								phagocytosed_once = true;
						}
						else if (!attacked_once && uk < lane_real<R>(p_attack, k))
						{
							event = 1;
							if (!removed_i && !removed_j)
								attacked_once = true;
						}
						else if (!fused_once && uk < lane_real<R>(p_fuse, k))
						{
							event = 2;
							if (!removed_i && !removed_j)
								fused_once = true;
						}
						else
							continue;
						if (removed_i)
						{
							broken = true;
							continue;
						}
						if (removed_j)
							continue;
						if (event == 1)
						{
							const R nd = lane_real<R>(dmg, k) + lane_real<R>(dr, k);
							const R na = lane_real<R>(atk, k) + dt;
							if (lane == 0)
							{
								c.damage[j] = nd;
								c.attack_time[j] = na;
							}
							dmg = aj == j ? nd : dmg;
							atk = aj == j ? na : atk;
						}
						else
						{
							if (lane == 0)
							{
								if (event == 0)
									ingest<R>(c, i, j);
								else
									fuse<R>(c, i, j);
							}
							fl_i = ai == j ? (int)PCM_FLAG_TO_REMOVE : fl_i;
							fl_j = aj == j ? (int)PCM_FLAG_TO_REMOVE : fl_j;
							if (PLAN)
							{
								if (touched + 2 <= BADC)
								{
									if (lane == 0)
									{
										sh.sb[touched] = i;
										sh.sb[touched + 1] = j;
									}
									touched += 2;
								}
								else
									touch_overflow = true;
							}
						}
					}
					// the next lanes' loads read what lane 0 stored
					__threadfence_block();
				}
			}
			__threadfence();
			__syncthreads();
			lo = hi;
		}
		// clear the event words the tentative pass raised (on targets and on actors)
		for (int e = threadIdx.x; e < C; e += blockDim.x)
		{
			const int i = c.cand[2 * e];
			const int j = c.nbr_idx[(size_t)i * c.K + c.cand[2 * e + 1]];
			c.ev_t[j] = 0;
			c.ev_t[i] = 0;
		}
	}
	if (threadIdx.x == 0)
		sh.rescan = touch_overflow ? 1 : 0; // thread 0 walked the candidates
	__syncthreads();
	const bool rescan = sh.rescan != 0;
	if (threadIdx.x == 0)
	{
		c.sc->stat[3] = C;
		c.sc->stat[4] = conflict;
		c.sc->conflict = 0;
		c.sc->cand_count = 0;
	}
	if (PLAN)
	{
		// The apply launch's other blocks listed the removals as they stood before the replay
		// (removal flag set, or outside the box).  Of the replay's effects only ingestion and
		// fusion change that (their target is flagged; a fusion moves its actor), so the
		// agents they touched are settled one by one: listed if now removed, taken off the
		// list if a fusion brought them back into the box.  When more were touched than the
		// LDS list holds, the whole population is scanned again instead.
		if (replay && rescan)
		{
			__syncthreads();
			if (threadIdx.x == 0)
				c.sc->bad_count = 0;
			__syncthreads();
			// the replay's writes are read back by other wavefronts: bypass the L1
			const int n = c.sc->n;
			const volatile uint8_t* flags = c.flags;
			const volatile R* pos = c.pos;
			for (int i = threadIdx.x; i < n; i += blockDim.x)
			{
				bool out = flags[i] == PCM_FLAG_TO_REMOVE;
				for (int k = 0; k < D; k++)
				{
					const R p = pos[(size_t)i * D + k];
					if (p < (R)c.bmin[k] || p > (R)c.bmax[k])
						out = true;
				}
				if (out)
					c.bad[atomicAdd(&c.sc->bad_count, 1)] = i;
			}
			// the list entries other wavefronts appended are read by the plan below
			__threadfence();
		}
		else if (replay && threadIdx.x == 0)
		{
			// the other blocks' list entries were acquired at this block's start; the flags and
			// positions read here changed, if at all, by this thread's own replay
			int nb = c.sc->bad_count;
			for (int t = 0; t < touched; t++)
			{
				const int x = sh.sb[t];
				bool seen = false;
				for (int u = 0; u < t; u++)
					seen |= sh.sb[u] == x;
				if (seen)
					continue;
				R p[D];
				for (int k = 0; k < D; k++)
					p[k] = c.pos[(size_t)x * D + k];
				const bool out = (c.flags[x] == PCM_FLAG_TO_REMOVE) | outside_box<R, D>(c, p);
				int at = -1;
				for (int u = 0; u < nb; u++)
					at = c.bad[u] == x ? u : at;
				if (out && at < 0)
					c.bad[nb++] = x;
				else if (!out && at >= 0)
					c.bad[at] = c.bad[--nb];
			}
			c.sc->bad_count = nb;
			// the list is read by the plan's other wavefronts below
			__threadfence();
		}
		__syncthreads();
		// the replay appends to the removal list: its count is read again
		const int b = replay ? *(volatile int*)&c.sc->bad_count : bad0;
		compact_plan_block<R>(c, sh.sb, BADC, &sh.red, b, n0, limit, log0);
	}
}

template <class R, int D, bool PLAN>
__global__ void __launch_bounds__(RESOLVE_THREADS) k_interact_serial(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	(void)bmap; // one block per simulation: never flattened
	__shared__ serial_lds<CAND_CHUNK, BAD_CHUNK, PLAN> sh;
	interact_serial_block<R, D, PLAN, CAND_CHUNK, BAD_CHUNK>(ctxs[blockIdx.y], sh);
}

// LDS passes of the serial pass when it runs in the last block of the apply launch (a
// smaller footprint than the stand-alone kernel's, which every apply block would carry)
constexpr int TAIL_CAND_CHUNK = 1024;
constexpr int TAIL_BAD_CHUNK = 1024;

// The step's apply pass (FLAG mode), then the serial pass with the container-update plan in
// the block that finishes last, so that the pair costs one launch.  Every block takes a
// ticket from the scalars' counter whatever the status word says (the counter returns to
// zero for the next launch); the block that takes the last ticket resets it and runs the
// serial pass after all other blocks' work.  The serial pass reads two things this launch
// writes: removal-list entries, and the removal flags that ingestion and fusion set on
// their targets (the plan counts flagged removals as deaths).  The wavefronts that wrote
// either, and their block's ticket taker, publish them with a device-scope fence first.
template <class R, int D>
__global__ void __launch_bounds__(BLOCK) k_interact_apply_plan(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_GEOMETRY
	const SimCtx<R>& c = ctxs[slot_];
	__shared__ serial_lds<TAIL_CAND_CHUNK, TAIL_BAD_CHUNK, true> sh;
	__shared__ int s_state;
	if (threadIdx.x == 0)
		s_state = c.sc->status ? 2 : (c.sc->conflict ? 1 : 0); // 2: stop, 1: replay follows
	__syncthreads();
	bool wrote = false;
	if (s_state < 2)
	{
		wrote = s_state == 0 ? interact_apply_agent<R, D, true>(c, GID) : interact_flag_agent<R, D>(c, GID);
		// publish (release: the XCD's L2 written back; no invalidation needed)
		if (__any(wrote))
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
	}
	const int written = __syncthreads_or(wrote);
	if (threadIdx.x == 0)
	{
		if (written)
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
		s_state = atomicAdd(&c.sc->ticket, 1) == nblk_ - 1;
	}
	__syncthreads();
	if (!s_state)
		return;
	// the last block reads what the others published: acquire only (its L2 invalidated,
	// nothing written back)
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
	if (threadIdx.x == 0)
		c.sc->ticket = 0;
	interact_serial_block<R, D, true, TAIL_CAND_CHUNK, TAIL_BAD_CHUNK>(c, sh);
}

// ------------------------------------------------------------------ container compaction
// update_cell_container_for_mechanics (containers_solver.cpp:58-147).  In the reference's
// single-thread order removal i -> (last agent moves into i) leaves: kept agents below
// N' = N - removed stay; holes below N' (ascending) receive the kept agents of [N', N) in
// descending order.  Link lists of survivors are edited as the reference's sequence of
// remove_attached (swap with back) and rename_attached calls would edit them.
template <class R, int D>
__global__ void k_compact_flag(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	const int i = GID;
	if (i >= c.sc->n)
		return;
	bool bad = c.flags[i] == PCM_FLAG_TO_REMOVE;
	for (int k = 0; k < D; k++)
	{
		const R p = c.pos[(size_t)i * D + k];
		if (p < (R)c.bmin[k] || p > (R)c.bmax[k])
			bad = true;
	}
	if (bad)
	{
		const int k = atomicAdd(&c.sc->bad_count, 1);
		c.bad[k] = i;
	}
}

__device__ inline void block_sort_ints(int* a, int count)
{
	int P = 1;
	while (P < count)
		P <<= 1;
	for (int t = count + threadIdx.x; t < P; t += blockDim.x)
		a[t] = 0x7fffffff;
	__syncthreads();
	for (int k = 2; k <= P; k <<= 1)
		for (int j = k >> 1; j > 0; j >>= 1)
		{
			for (int t = threadIdx.x; t < P; t += blockDim.x)
			{
				const int u = t ^ j;
				if (u > t)
				{
					const bool up = (t & k) == 0;
					const int x = a[t], y = a[u];
					if ((x > y) == up)
					{
						a[t] = y;
						a[u] = x;
					}
				}
			}
			__syncthreads();
		}
}

// The removal loop of update_cell_container_for_phenotype (containers_solver.cpp:182-186)
// visits only i < limit: a bad agent at or above it is removed only when, as the last
// agent, it is moved into a visited hole; otherwise it stays.  `list` holds the bad agents
// ascending.  Single thread.
template <class R>
__device__ void compact_plan_limited(const SimCtx<R>& c, const int* list, int b, int limit, int n0, int log)
{
	int top = n0, tb = b - 1, q = 0, moves = 0, removed = 0;
	long long deaths = 0;
	for (int h = 0; h < b; h++)
	{
		const int hole = list[h];
		if (hole >= limit || hole >= top)
			break;
		int victim = hole;
		for (;;)
		{
			c.remap[victim] = -1 - q++;
			deaths += c.flags[victim] == PCM_FLAG_TO_REMOVE;
			removed++;
			if (log < c.log_cap)
				c.rm_log[log] = hole;
			log++;
			const int t = --top;
			if (t == hole)
				break; // the hole held the last agent
			while (tb >= 0 && list[tb] > t)
				tb--;
			if (tb >= 0 && list[tb] == t)
			{
				victim = t; // the last agent is bad too: it moves into the hole and goes
				tb--;
				continue;
			}
			c.move_src[moves] = t;
			c.move_dst[moves] = hole;
			c.remap[t] = hole;
			moves++;
			break;
		}
		if (top == hole)
			break; // the count fell to the visited index: the loop ends
	}
	c.sc->log_len = log;
	c.sc->moves = moves;
	c.sc->stat[5] = removed;
	c.sc->n = top;
	atomicAdd((unsigned long long*)&c.sc->deaths, (unsigned long long)deaths);
}

// The container update's plan on one block.  b, n0, limit and log are the removal count,
// agent count, compact_limit and log_len as the caller read them (read together, so the
// plan's chain of dependent loads is short); s_red is an LDS int.
template <class R>
__device__ void compact_plan_block(const SimCtx<R>& c, int* sb, int sb_cap, int* s_red, int b, int n0, int limit, int log)
{
	if (b == 0)
	{
		if (threadIdx.x == 0)
			c.sc->stat[5] = 0;
		return;
	}
	// ascending list of removed agents: in LDS when it fits, else sorted in place in HBM
	// (the list may have been filled earlier in the same kernel: read past the L1)
	const volatile int* bad_list = c.bad;
	int* list = sb;
	if (b <= sb_cap)
	{
		for (int t = threadIdx.x; t < b; t += blockDim.x)
			sb[t] = bad_list[t];
		__syncthreads();
		block_sort_ints(sb, b);
	}
	else
	{
		// the removal list buffer holds cap + padding entries
		block_sort_ints(c.bad, b);
		list = c.bad;
	}
	__syncthreads();
	if (limit != 0)
	{
		if (threadIdx.x == 0)
			compact_plan_limited<R>(c, list, b, limit < 0 ? 0 : limit, n0, log);
		return;
	}
	// removed agents flagged for removal are deaths: their flags are read in parallel
	if (threadIdx.x == 0)
		*s_red = 0;
	__syncthreads();
	int d = 0;
	for (int t = threadIdx.x; t < b; t += blockDim.x)
		d += c.flags[list[t]] == PCM_FLAG_TO_REMOVE;
	if (d)
		atomicAdd(s_red, d);
	__syncthreads();
	if (threadIdx.x != 0)
		return;
	const long long deaths = *s_red;
	const int n1 = n0 - b;
	int holes = 0;
	while (holes < b && list[holes] < n1)
		holes++;
	// walk the tail [n1, n0) downwards: good agents fill holes, bad ones get their
	// position in the removal sequence
	int q = 0;
	auto log_removal = [&](int position) {
		if (log < c.log_cap)
			c.rm_log[log] = position;
		log++;
	};
	int tb = b - 1; // index of the largest not-yet-visited bad tail agent
	int top = n0;   // agents >= top have been consumed
	int moves = 0;
	for (int h = 0; h < holes; h++)
	{
		const int hole = list[h];
		c.remap[hole] = -1 - q++;
		log_removal(hole);
		// consume tail agents from top-1 down until a good one fills the hole
		for (;;)
		{
			const int t = --top;
			if (t < n1)
			{
				atomicOr(&c.sc->status, ST_CAPACITY); // inconsistent removal list
				return;
			}
			if (tb >= holes && list[tb] == t)
			{
				c.remap[t] = -1 - q++;
				log_removal(hole); // moved into the hole, then removed there
				tb--;
				continue;
			}
			c.move_src[moves] = t;
			c.move_dst[moves] = hole;
			c.remap[t] = hole;
			moves++;
			break;
		}
	}
	// remaining bad agents in [n1, top): n1 first, then top-1 down to n1+1
	if (top > n1)
	{
		c.remap[n1] = -1 - q++;
		log_removal(n1);
		for (int t = top - 1; t > n1; t--)
		{
			c.remap[t] = -1 - q++;
			log_removal(n1); // each in turn becomes the last agent, moves into n1, is removed
		}
	}
	c.sc->log_len = log;
	c.sc->moves = moves;
	c.sc->stat[5] = b;
	c.sc->n = n1;
	atomicAdd((unsigned long long*)&c.sc->deaths, (unsigned long long)deaths);
}

template <class R>
__global__ void __launch_bounds__(RESOLVE_THREADS) k_compact_plan(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	(void)bmap; // one block per simulation: never flattened
	const SimCtx<R>& c = ctxs[blockIdx.y];
	__shared__ int sb[BAD_CHUNK];
	__shared__ int s_red;
	const int status = c.sc->status, b = *(volatile int*)&c.sc->bad_count, n0 = c.sc->n;
	const int limit = c.sc->compact_limit, log = c.sc->log_len;
	if (status)
		return;
	compact_plan_block<R>(c, sb, BAD_CHUNK, &s_red, b, n0, limit, log);
}

// remove_attached (swap with back) for removed partners in removal order, then
// rename_attached for moved partners, on one survivor's list
__device__ inline void edit_links(int* list, int* cnt_ptr, const int* remap)
{
	int cnt = *cnt_ptr;
	if (cnt == 0)
		return;
	// no early exit: eight entries and their remap values are requested at a time, so an
	// untouched list (the common case) costs two round trips per eight entries
	bool touched = false;
#pragma unroll 8
	for (int s = 0; s < cnt; s++)
	{
		const int e = list[s];
		touched |= remap[e] != e;
	}
	if (!touched)
		return;
	for (;;)
	{
		// next removal op touching this list: smallest sequence number among removed entries
		// (branch-free and unrolled: eight entries and their remap values per round trip)
		int best_seq = 0x7fffffff, best_id = -1;
#pragma unroll 8
		for (int s = 0; s < cnt; s++)
		{
			const int e = list[s];
			const int r = remap[e];
			const bool better = (r < 0) & (-1 - r < best_seq);
			best_seq = better ? -1 - r : best_seq;
			best_id = better ? e : best_id;
		}
		if (best_id < 0)
			break;
		// the removed agent strips every entry of itself, each time the first one
		for (;;)
		{
			int pos = -1;
			for (int s = 0; s < cnt; s++)
				if (list[s] == best_id)
				{
					pos = s;
					break;
				}
			if (pos < 0)
				break;
			list[pos] = list[cnt - 1];
			cnt--;
		}
	}
	// renames: eight entries and their new indices are read before the eight stores
	for (int s0 = 0; s0 < cnt; s0 += 8)
	{
		int v[8];
#pragma unroll
		for (int u = 0; u < 8; u++)
			v[u] = s0 + u < cnt ? list[s0 + u] : 0;
#pragma unroll
		for (int u = 0; u < 8; u++)
			v[u] = s0 + u < cnt ? remap[v[u]] : 0;
#pragma unroll
		for (int u = 0; u < 8; u++)
			if (s0 + u < cnt)
				list[s0 + u] = v[u];
	}
	*cnt_ptr = cnt;
}

// The container update's data movement, in one launch: the first MOVE_BLOCKS blocks move
// every field and list row of each surviving tail agent into its hole (one wavefront per
// move, lanes over the words of the copy plan; the moved lists are edited by the mover),
// the other blocks edit the link lists of the survivors that stay in place.
constexpr int MOVE_BLOCKS = 16;

// one surviving tail agent's move into its hole, by one wavefront
template <class R>
__device__ __forceinline__ void compact_move(const SimCtx<R>& c, int src, int dst, int lane)
{
	// the copy plan's words (at most MOVE_UNITS per lane): every lane reads its plan
	// entries and source words first, then stores, so the words cost one round trip
	// instead of one per 64 (the stores could alias the plan for the compiler)
	constexpr int MOVE_UNITS = (int)(sizeof(c.mv_field) / sizeof(c.mv_field[0]) + 63) / 64;
	uint8_t* dst_at[MOVE_UNITS];
	uint32_t word[MOVE_UNITS];
	int bytes[MOVE_UNITS];
#pragma unroll
	for (int r = 0; r < MOVE_UNITS; r++)
	{
		const int u = lane + 64 * r;
		bytes[r] = 0;
		if (u < c.mv_n)
		{
			const int f = c.mv_field[u];
			const size_t nb = (size_t)c.fbytes[f];
			bytes[r] = c.mv_bytes[u];
			const size_t off = (size_t)c.mv_unit[u] * bytes[r];
			const uint8_t* base = (const uint8_t*)c.fptr[f];
			dst_at[r] = (uint8_t*)base + (size_t)dst * nb + off;
			word[r] = bytes[r] == 4 ? *(const uint32_t*)(base + (size_t)src * nb + off)
									: (uint32_t)base[(size_t)src * nb + off];
		}
	}
#pragma unroll
	for (int r = 0; r < MOVE_UNITS; r++)
	{
		if (bytes[r] == 4)
			*(uint32_t*)dst_at[r] = word[r];
		else if (bytes[r] == 1)
			*dst_at[r] = (uint8_t)word[r];
	}
	// list rows: lanes over the entries (a serial copy costs one round trip per entry)
	for (int t = lane; t < c.K; t += 64)
		c.nbr_idx[(size_t)dst * c.K + t] = c.nbr_idx[(size_t)src * c.K + t];
	for (int t = lane; t < c.L; t += 64)
		c.spr_idx[(size_t)dst * c.L + t] = c.spr_idx[(size_t)src * c.L + t];
	for (int t = lane; t < c.LA; t += 64)
		c.atc_idx[(size_t)dst * c.LA + t] = c.atc_idx[(size_t)src * c.LA + t];
	if (lane == 0)
	{
		c.nbr_cnt[dst] = c.nbr_cnt[src];
		c.spr_cnt[dst] = c.spr_cnt[src];
		c.atc_cnt[dst] = c.atc_cnt[src];
	}
	// the lane-0 edits below read rows the other lanes wrote
	__threadfence_block();
	if (lane == 0)
	{
		edit_links(c.spr_idx + (size_t)dst * c.L, c.spr_cnt + dst, c.remap);
		edit_links(c.atc_idx + (size_t)dst * c.LA, c.atc_cnt + dst, c.remap);
	}
}

template <class R>
__global__ void k_compact_apply(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	if (c.sc->bad_count == 0)
		return;
	if (bid_ < MOVE_BLOCKS)
	{
		const int waves = MOVE_BLOCKS * (blockDim.x >> 6);
		const int moves = c.sc->moves;
		for (int k = bid_ * (blockDim.x >> 6) + (threadIdx.x >> 6); k < moves; k += waves)
			compact_move<R>(c, c.move_src[k], c.move_dst[k], lane_id());
		return;
	}
	const int i = (bid_ - MOVE_BLOCKS) * blockDim.x + threadIdx.x;
	if (i >= c.sc->n || c.remap[i] < 0) // holes receive their moved agent's edited lists
		return;
	edit_links(c.spr_idx + (size_t)i * c.L, c.spr_cnt + i, c.remap);
	edit_links(c.atc_idx + (size_t)i * c.LA, c.atc_cnt + i, c.remap);
}

// The container update's data movement and the next mesh update's voxel histogram in one
// launch.  Each agent of the compacted population is counted into the voxel of its final
// position: an agent that stays by its own thread, a moved one by the wavefront that moves
// it (with the position it moves), so the count never reads a slot a mover writes.  The
// remap entries the link edits read are returned to the identity by a later kernel (the
// voxel claim, or with FIXED the ordering kernel).  The thread that finishes an agent's
// spring list also draws the next step's detachments for it (k_spring_fire's work, keyed
// by the stored step counter plus one; spring lists, counts, rates and indices are final
// here).
// FIXED (voxel tables): the agent also takes its slot in the voxel's unordered table (B slots per
// voxel, `vmem`), and the voxel's first agent gives the voxel the next range of B bucket
// slots (one counter atomic per block) and lists it as occupied: no claim or scatter launch
// follows.  A voxel with more than B agents raises ST_BUCKET_OVERFLOW (the host grows B
// and rebuilds the mesh the dense way).
template <class R, int D, bool FIXED>
__device__ __forceinline__ int count_into_voxel(const SimCtx<R>& c, int a, int src)
{
	const int v = voxel_index<R>(c, c.pos + (size_t)src * D);
	const int r = atomicAdd(&c.vcnt[v], 1);
	c.vox_of[a] = v;
	c.vrank[a] = r;
	if (FIXED)
	{
		if (r < c.B)
			c.vmem[(size_t)v * c.B + r] = a;
		else
		{
			atomicOr(&c.sc->status, ST_BUCKET_OVERFLOW);
			atomicMax(&c.sc->max_bucket, r + 1);
		}
	}
	return r == 0 ? v : -1;
}

template <class R, int D, bool FIXED>
__global__ void k_compact_count(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE_BLOCK(c)
	const bool compacting = c.sc->bad_count != 0;
	const int n = c.sc->n;
	const unsigned long long next_step = c.sc->step + 1;
	if (bid_ < MOVE_BLOCKS)
	{
		if (!compacting)
			return;
		const int waves = MOVE_BLOCKS * (blockDim.x >> 6);
		const int moves = c.sc->moves;
		const int lane = lane_id();
		for (int k = bid_ * (blockDim.x >> 6) + (threadIdx.x >> 6); k < moves; k += waves)
		{
			const int src = c.move_src[k], dst = c.move_dst[k];
			compact_move<R>(c, src, dst, lane);
			if (lane == 0)
			{
				// the moved fields and the edited list were written by this wavefront
				spring_fire_agent<R>(c, dst, next_step);
				const int v = count_into_voxel<R, D, FIXED>(c, dst, src);
				if (FIXED && v >= 0)
				{
					const int w = (int)atomicAdd(&c.sc->claim, 1ull);
					c.occ[w] = v;
					c.vstart[v] = w * c.B;
				}
			}
		}
		return;
	}
	const int i = (bid_ - MOVE_BLOCKS) * blockDim.x + threadIdx.x;
	if (i == 0)
	{
		c.sc->n_binned = n;
		c.sc->res_open = 0; // the next step's resolver starts with no open actors
	}
	int first = -1; // the voxel this thread's agent opened (FIXED)
	// a hole receives its agent, counted and drawn for by the mover
	if (i < n && !(compacting && c.remap[i] < 0))
	{
		if (compacting)
		{
			edit_links(c.spr_idx + (size_t)i * c.L, c.spr_cnt + i, c.remap);
			edit_links(c.atc_idx + (size_t)i * c.LA, c.atc_cnt + i, c.remap);
		}
		spring_fire_agent<R>(c, i, next_step);
		first = count_into_voxel<R, D, FIXED>(c, i, i);
	}
	if (FIXED)
	{
		__shared__ int s_base;
		int total;
		const int pre = block_exclusive_scan(first >= 0 ? 1 : 0, total);
		if (threadIdx.x == 0)
			s_base = total ? (int)atomicAdd(&c.sc->claim, (unsigned long long)total) : 0;
		__syncthreads();
		if (first >= 0)
		{
			const int w = s_base + pre;
			c.occ[w] = first;
			c.vstart[first] = w * c.B;
		}
	}
}

template <class R>
__global__ void k_compact_restore(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	const int k = GID;
	const int b = c.sc->bad_count;
	if (k < b)
	{
		const int x = c.bad[k];
		c.remap[x] = x;
	}
	if (k < c.sc->moves)
	{
		const int x = c.move_src[k];
		c.remap[x] = x;
	}
}

template <class R>
__global__ void k_compact_done(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	if (GID == 0)
	{
		c.sc->bad_count = 0;
		c.sc->moves = 0;
		c.sc->compact_limit = 0;
	}
}

// ------------------------------------------------------------------ output packing
// One output double per thread: element (agent, column) of the agent-major matrix; column k
// reads component comp[k] of field fld[k] (a component beyond the field's width is 0).
template <class R>
__global__ void k_pack(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap, const int* __restrict__ fld, const int* __restrict__ comp,
					   int row, double* __restrict__ out)
{
	PCM_PROLOGUE(c)
	const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	const long long total = (long long)c.sc->n * row;
	if (e >= total)
		return;
	const int i = (int)(e / row), k = (int)(e % row);
	const int f = fld[k], m = comp[k];
	const int nb = c.fbytes[f];
	const uint8_t* base = (const uint8_t*)c.fptr[f] + (size_t)i * nb;
	double v = 0;
	switch (pcm_field_type(f))
	{
		case PCM_T_REAL:
			if (m < nb / (int)sizeof(R))
				v = (double)((const R*)base)[m];
			break;
		case PCM_T_I32:
			if (m < nb / 4)
				v = (double)((const int*)base)[m];
			break;
		default:
			if (m < nb)
				v = (double)base[m];
	}
	out[e] = v;
}

// ------------------------------------------------------------------ phenotype-cadence update
// update_cell_container_for_phenotype (containers_solver.cpp:149-187) in the reference's
// single-thread order: divisions in ascending parent order (cell::divide, cell.cpp:108-163),
// then the removal loop over the agents that existed before (compact_plan_block with
// compact_limit).

// the agents flagged to_divide, ascending, and their count; one block of 1024
template <class R>
__global__ void __launch_bounds__(1024) k_divide_plan(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	(void)bmap; // one block per simulation: never flattened
	const SimCtx<R>& c = ctxs[blockIdx.y];
	__shared__ int wave_total[16];
	__shared__ int base;
	if (c.sc->status)
		return;
	const int n = c.sc->n;
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	if (threadIdx.x == 0)
		base = 0;
	__syncthreads();
	for (int start = 0; start < n; start += 1024)
	{
		const int i = start + threadIdx.x;
		const bool div = i < n && c.flags[i] == PCM_FLAG_TO_DIVIDE;
		const unsigned long long m = __ballot(div);
		if (lane == 0)
			wave_total[w] = __popcll(m);
		__syncthreads();
		int off = base;
		for (int k = 0; k < w; k++)
			off += wave_total[k];
		if (div)
			c.div_list[off + __popcll(m & ((1ull << lane) - 1ull))] = i;
		__syncthreads();
		if (threadIdx.x == 0)
		{
			int t = 0;
			for (int k = 0; k < 16; k++)
				t += wave_total[k];
			base += t;
		}
		__syncthreads();
	}
	if (threadIdx.x == 0)
		c.sc->div_count = base;
}

// position_helper<D>::random_walk(false, walk) (solver_helper.h:82-86, 160-165, 250-267),
// keyed by the phenotype-update call, the parent and a slot
template <int D, class R>
__device__ __forceinline__ void division_walk(const SimCtx<R>& c, int parent, uint32_t slot, R* walk)
{
	const pcm_u32x4 blk = pcm_block(c.seed, c.div_epoch, PCM_STREAM_DIVISION, (uint32_t)parent, slot, 0u);
	if constexpr (D == 1)
		walk[0] = pcm_draw<R>::from(blk, 0) < (R)0.5 ? (R)-1 : (R)1;
	else
	{
		const R two_pi = (R)2 * (R)3.14159265358979323846;
		const R theta = pcm_scale_uniform<R>(pcm_draw<R>::from(blk, 0), (R)0, two_pi);
		if constexpr (D == 2)
		{
			walk[0] = cos(theta);
			walk[1] = sin(theta);
		}
		else
		{
			const R z = pcm_scale_uniform<R>(pcm_draw<R>::from(blk, 1), (R)-1, (R)1);
			const R r = sqrt(1 - z * z);
			walk[0] = cos(theta) * r;
			walk[1] = sin(theta) * r;
			walk[2] = z;
		}
	}
}

// the fields cell::copy_from hands to the daughter: phenotype (phenotype.cpp:18-30), cell
// definition index and cell functions (cell.cpp:92-106)
__device__ __forceinline__ bool copied_on_division(int f)
{
	return !(f == PCM_F_POSITIONS || f == PCM_F_VELOCITIES || f == PCM_F_PREVIOUS_VELOCITIES
			 || f == PCM_F_SIMPLE_PRESSURE || f == PCM_F_NUMBER_OF_NUCLEI || f == PCM_F_DAMAGE
			 || f == PCM_F_TOTAL_ATTACK_TIME || f == PCM_F_IS_MOVABLE || f == PCM_F_FLAGS || f == PCM_F_ORIGIN
			 || f == PCM_F_ORIENTATION);
}

// cell::divide of division r: parent div_list[r], daughter at n + r
template <class R, int D>
__global__ void k_divide_apply(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	const int r = GID;
	if (r >= c.sc->div_count)
		return;
	const int p = c.div_list[r];
	const int d = c.sc->n + r;
	c.flags[p] = PCM_FLAG_NONE;
	// the parent leaves its partners' lists in division order (cell.cpp:111-112)
	c.remap[p] = -1 - r;
	c.spr_cnt[p] = 0;
	c.atc_cnt[p] = 0;
	if (c.div_internal)
		for (int s = 0; s < c.S; s++)
			c.internal[(size_t)p * c.S + s] *= (R)0.5;
	c.damage[p] = 0;
	c.attack_time[p] = 0;

	R walk[D];
	division_walk<D, R>(c, p, 0u, walk);
	R tmp = 0;
	for (int k = 0; k < D; k++)
		tmp += walk[k] * c.orientation[(size_t)p * D + k];
	const R pol = c.polarity[p], rad = c.radius[p];
	for (int k = 0; k < D; k++)
		walk[k] = (walk[k] - tmp * c.orientation[(size_t)p * D + k] * pol) * rad;

	// volume_t::divide (data_storage_phenotype.cpp:530-550), geometry_t::update
	R* halves[11] = { c.volume, c.solid, c.fluid, c.nuclear, c.nuc_fluid, c.nuc_solid,
					  c.cyto, c.cyto_fluid, c.cyto_solid, c.target_cyto, c.target_nuc };
	for (int k = 0; k < 11; k++)
		halves[k][p] *= (R)0.5;
	update_geometry<R>(c, p);

	for (int f = 0; f < PCM_FIELD_COUNT; f++)
	{
		if (!copied_on_division(f))
			continue;
		const int nb = c.fbytes[f];
		const uint8_t* src = (const uint8_t*)c.fptr[f] + (size_t)p * nb;
		uint8_t* dst = (uint8_t*)c.fptr[f] + (size_t)d * nb;
		if ((nb & 3) == 0)
			for (int k = 0; k < nb / 4; k++)
				((uint32_t*)dst)[k] = ((const uint32_t*)src)[k];
		else
			for (int k = 0; k < nb; k++)
				dst[k] = src[k];
	}
	for (int k = 0; k < D; k++)
	{
		c.vel[(size_t)d * D + k] = 0;
		c.pvel[(size_t)d * D + k] = 0;
	}
	c.pressure[d] = 0;
	c.nuclei[d] = 1;
	c.damage[d] = 0;
	c.attack_time[d] = 0;
	c.movable[d] = 1;
	c.flags[d] = PCM_FLAG_NONE;
	c.origin[d] = -1;
	c.spr_cnt[d] = 0;
	c.atc_cnt[d] = 0;
	c.nbr_cnt[d] = 0;
	division_walk<D, R>(c, p, 1u, c.orientation + (size_t)d * D); // cell::assign_orientation
	for (int k = 0; k < D; k++)
	{
		const R x = c.pos[(size_t)p * D + k];
		c.pos[(size_t)d * D + k] = x + walk[k];
		c.pos[(size_t)p * D + k] = x - walk[k] * (R)0.5;
	}
}

// partners of dividing agents drop them, in division order (remove_attached, swap with back)
template <class R>
__global__ void k_divide_links(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	const int i = GID;
	if (i >= c.sc->n || c.sc->div_count == 0 || c.remap[i] < 0)
		return;
	edit_links(c.spr_idx + (size_t)i * c.L, c.spr_cnt + i, c.remap);
	edit_links(c.atc_idx + (size_t)i * c.LA, c.atc_cnt + i, c.remap);
}

// restore the parents' remap, count the daughters in, and bound the removal loop that
// follows to the agents that existed before
template <class R>
__global__ void k_divide_done(const SimCtx<R>* __restrict__ ctxs, const int4* __restrict__ bmap)
{
	PCM_PROLOGUE(c)
	const int r = GID;
	const int count = c.sc->div_count;
	if (r < count)
		c.remap[c.div_list[r]] = c.div_list[r];
	if (r == 0)
	{
		const int n = c.sc->n;
		c.sc->compact_limit = n > 0 ? n : -1;
		c.sc->n = n + count;
		c.sc->divisions += count;
	}
}

} // namespace pcm

// pcm_lib.hip — libpcm.so: the C-ABI of include/pcm.h over device-resident state.
//
// One pcm_sim owns one simulation: its agent SoA in HBM, list storage, the mechanics
// voxel grid, scratch buffers and a HIP stream of default priority.  Phase entry points
// enqueue kernels and return; pcm_step enqueues whole mechanics steps back to back and
// synchronizes once at the end (re-running the neighbor build when the fixed per-agent
// neighbor capacity had to grow).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <vector>

#include "pcm_kernels.h"

using namespace pcm;

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg)
{
	g_error = msg;
	return code;
}

#define PCM_HIP(expr)                                                                                                  \
	do                                                                                                                 \
	{                                                                                                                  \
		hipError_t e_ = (expr);                                                                                        \
		if (e_ != hipSuccess)                                                                                          \
			return fail(PCM_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));                           \
	} while (0)

#define PCM_TRY(expr)                                                                                                  \
	do                                                                                                                 \
	{                                                                                                                  \
		int rc_ = (expr);                                                                                              \
		if (rc_ != PCM_OK)                                                                                             \
			return rc_;                                                                                                \
	} while (0)

struct field_meta
{
	int type, width_kind;
	const char* name;
};

#define PCM_FIELD_META_(name, type, width) { type, width, #name },
const field_meta k_fields[PCM_FIELD_COUNT] = { PCM_FIELD_LIST(PCM_FIELD_META_) };
#undef PCM_FIELD_META_

int next_pow2(long long x)
{
	int p = 1;
	while (p < x)
		p <<= 1;
	return p;
}

// ---------------------------------------------------------------------------- device memory pool
// Device memory of destroyed simulations is kept for later handles on the same device
// instead of going back to the runtime: a serving loop creates and destroys a handle per
// request, and hipMalloc/hipFree per request would cost milliseconds each, hipFree waiting
// for the whole device (the other simulations' queued steps included).  A cached block
// serves any request from half its size up to its size.  Blocks return to the runtime only
// when the cache exceeds POOL_CACHE_LIMIT or a hipMalloc fails for lack of memory.  A block
// is given back only once the stream that used it is idle (every caller synchronizes its
// stream first), so its next owner cannot race its last one.
constexpr size_t POOL_CACHE_LIMIT = 64ull << 30;
constexpr int POOL_DEVICES = 64;

class device_pool
{
public:
	int take(size_t bytes, void** out, size_t* got, int* device)
	{
		int dev = 0;
		PCM_HIP(hipGetDevice(&dev));
		if (dev < 0 || dev >= POOL_DEVICES)
			return fail(PCM_ERR_UNSUPPORTED, "device ordinal beyond the memory pool's table");
		const size_t want = (std::max<size_t>(bytes, 16) + 255) & ~(size_t)255;
		{
			std::lock_guard<std::mutex> lock(mu_);
			auto& m = free_[dev];
			auto it = m.lower_bound(want);
			if (it != m.end() && it->first / 2 <= want)
			{
				*out = it->second;
				*got = it->first;
				*device = dev;
				cached_ -= it->first;
				m.erase(it);
				return PCM_OK;
			}
		}
		void* p = nullptr;
		hipError_t e = hipMalloc(&p, want);
		if (e == hipErrorOutOfMemory)
		{
			(void)hipGetLastError();
			trim(dev);
			e = hipMalloc(&p, want);
		}
		if (e != hipSuccess)
			return fail(PCM_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
		*out = p;
		*got = want;
		*device = dev;
		return PCM_OK;
	}

	void give(void* p, size_t bytes, int dev)
	{
		std::lock_guard<std::mutex> lock(mu_);
		free_[dev].emplace(bytes, p);
		cached_ += bytes;
		while (cached_ > POOL_CACHE_LIMIT)
		{
			// the largest cached block of any device goes back to the runtime
			int big = -1;
			for (int d = 0; d < POOL_DEVICES; d++)
				if (!free_[d].empty() && (big < 0 || free_[d].rbegin()->first > free_[big].rbegin()->first))
					big = d;
			auto it = std::prev(free_[big].end());
			(void)hipFree(it->second);
			cached_ -= it->first;
			free_[big].erase(it);
		}
	}

private:
	void trim(int dev)
	{
		std::lock_guard<std::mutex> lock(mu_);
		for (auto& [bytes, p] : free_[dev])
		{
			(void)hipFree(p);
			cached_ -= bytes;
		}
		free_[dev].clear();
	}

	std::mutex mu_;
	std::multimap<size_t, void*> free_[POOL_DEVICES];
	size_t cached_ = 0;
};

// A stream capture and calls on other threads: the runtime invalidates a batch's graph
// capture when another thread allocates, copies synchronously or waits on a stream during
// it (a serving loop sets requests up on worker threads while it steps the batch).  Every
// simulation-level entry point holds this lock shared; a batch holds it exclusively only
// between the begin and the end of a capture.
std::shared_mutex g_capture_mu;

device_pool& pool()
{
	static device_pool* p = new device_pool(); // never destroyed: blocks outlive static destructors
	return *p;
}

// ---------------------------------------------------------------------------- pinned staging and streams
// Reads of device arrays into caller (pageable) memory go device -> page-locked block (DMA,
// every array of a call queued together) -> one wait -> host copies out.  A staged pageable
// copy instead waits once per array and shares the runtime's staging buffer with every
// other thread's copies, which serialises the read-back of concurrent tear-downs.  Blocks
// are powers of two from 64 KiB, kept for reuse up to PINNED_CACHE_LIMIT.
constexpr size_t PINNED_CACHE_LIMIT = 1ull << 30;

class pinned_pool
{
public:
	int take(size_t bytes, void** out, size_t* got)
	{
		size_t want = 64u << 10;
		while (want < bytes)
			want <<= 1;
		{
			std::lock_guard<std::mutex> lock(mu_);
			auto it = free_.find(want);
			if (it != free_.end())
			{
				*out = it->second;
				*got = want;
				cached_ -= want;
				free_.erase(it);
				return PCM_OK;
			}
		}
		void* p = nullptr;
		PCM_HIP(hipHostMalloc(&p, want, hipHostMallocDefault));
		*out = p;
		*got = want;
		return PCM_OK;
	}

	void give(void* p, size_t bytes)
	{
		std::lock_guard<std::mutex> lock(mu_);
		if (cached_ + bytes > PINNED_CACHE_LIMIT)
		{
			(void)hipHostFree(p);
			return;
		}
		free_.emplace(bytes, p);
		cached_ += bytes;
	}

private:
	std::mutex mu_;
	std::multimap<size_t, void*> free_;
	size_t cached_ = 0;
};

pinned_pool& pinned()
{
	static pinned_pool* p = new pinned_pool(); // never destroyed, like the device pool
	return *p;
}

// Streams of destroyed handles are kept for later handles of the same device: creating and
// destroying a stream per request costs the serving loop's set-up and tear-down threads
// a runtime round trip each.  A stream is given back only once it is idle.  All streams
// have the default priority.
constexpr size_t STREAM_CACHE_LIMIT = 256;

class stream_pool
{
public:
	int take(hipStream_t* out)
	{
		int dev = 0;
		PCM_HIP(hipGetDevice(&dev));
		{
			std::lock_guard<std::mutex> lock(mu_);
			for (auto it = free_.begin(); it != free_.end(); ++it)
				if (it->first == dev)
				{
					*out = it->second;
					free_.erase(it);
					return PCM_OK;
				}
		}
		PCM_HIP(hipStreamCreateWithFlags(out, hipStreamNonBlocking));
		return PCM_OK;
	}

	// `dev`: the device the stream was created on
	void give(hipStream_t s, int dev)
	{
		std::lock_guard<std::mutex> lock(mu_);
		if (free_.size() >= STREAM_CACHE_LIMIT)
		{
			(void)hipStreamDestroy(s);
			return;
		}
		free_.emplace_back(dev, s);
	}

private:
	std::mutex mu_;
	std::vector<std::pair<int, hipStream_t>> free_;
};

stream_pool& streams()
{
	static stream_pool* p = new stream_pool();
	return *p;
}

} // namespace

struct pcm_sim
{
	virtual ~pcm_sim() = default;
	virtual int init() = 0;
	virtual int set_count(int n) = 0;
	virtual int get_count(int* n) = 0;
	virtual int set_field(int f, const void* src) = 0;
	virtual int set_fields(int nfields, const int32_t* ids, const void* const* srcs) = 0;
	virtual int get_field(int f, void* dst) = 0;
	virtual int get_fields(int nfields, const int32_t* ids, void* const* dsts) = 0;
	virtual int set_links(int which, const int32_t* off, const int32_t* idx) = 0;
	virtual int get_links(int which, int32_t* off, int32_t* idx, int64_t* total) = 0;
	virtual int get_neighbors(int32_t* off, int32_t* idx, int64_t* total) = 0;
	virtual int set_neighbors(const int32_t* off, const int32_t* idx) = 0;
	virtual int get_buckets(int32_t* counts, int32_t* agents) = 0;
	virtual long long voxel_count() const = 0;
	virtual int get_deaths(int64_t* d) = 0;
	virtual int get_divisions(int64_t* d) = 0;
	virtual int pack_fields(int nfields, const int32_t* fields, const int32_t* widths, double* out) = 0;
	virtual int division_log(int32_t* parents, int64_t capacity, int64_t* length) = 0;
	virtual int phenotype_update(const pcm_step_params* p, int compute_internalized) = 0;
	virtual int set_step(uint64_t s) = 0;
	virtual int set_step_params(const pcm_step_params& p) = 0;
	virtual int get_step(uint64_t* s) = 0;
	virtual int set_gradients(const int32_t* voxel, const void* g) = 0;
	virtual int phase(int which, const pcm_step_params* p) = 0;
	virtual int step(const pcm_step_params* p, int steps) = 0;
	virtual int synchronize() = 0;
	virtual int removal_log(int32_t* positions, int64_t capacity, int64_t* length) = 0;
	virtual int stats(int32_t* out) = 0;
	virtual int profile_begin(uint64_t mask) = 0;
	virtual int profile_every(int every) = 0;
	virtual int profile_end(double* total_ms, int64_t* launches) = 0;
	virtual int set_tuning(const pcm_tuning* t) = 0;
	virtual int get_tuning(pcm_tuning* t) = 0;
	virtual int autotune(const pcm_step_params* p, int steps, pcm_tuning* chosen) = 0;
};

struct pcm_batch
{
	virtual ~pcm_batch() = default;
	virtual int admit(pcm_sim* sim) = 0;
	virtual int retire(pcm_sim* sim) = 0;
	virtual int size(int32_t* sims) = 0;
	virtual int step(const pcm_step_params* p, int steps) = 0;
	virtual int synchronize() = 0;
	virtual int autotune(const pcm_step_params* p, int steps) = 0;
	virtual int counts(int n, pcm_sim* const* sims, int32_t* out) = 0;
	virtual int set_tuning(int dims, const pcm_tuning* t) = 0;
	virtual int get_tuning(int dims, pcm_tuning* t) = 0;
};

namespace {

// Block maps of a batch's flattened launches, one per grid shape the step uses: per agent
// (BLOCK threads), 4, 8 or 16 lanes per agent, per spring slot (erase), per CLAIM_THREADS
// agents (claim), per occupied voxel (neighbor search, at most 16384 blocks a member), the
// container update's move blocks plus one block per BLOCK agents, and the fused mesh
// update's ordering blocks plus its erase blocks (ERASE_LANES lanes per agent).
enum map_kind
{
	MAP_AGENT,
	MAP_LANES4,
	MAP_LANES8,
	MAP_LANES16,
	MAP_ERASE,
	MAP_CLAIM,
	MAP_NEIGHBORS,
	MAP_COMPACT,
	MAP_ORDER,
	MAP_KINDS,
	MAP_NONE = -1
};

inline int map_of_lanes(int lanes) { return lanes == 4 ? MAP_LANES4 : (lanes == 16 ? MAP_LANES16 : MAP_LANES8); }

// x rounded up to four significant bits (at most 1/8 above it): eight values per octave, so
// grid sizes and block counts derived from slowly drifting counts change rarely
inline int round_up_sig4(int x)
{
	if (x <= 8)
		return std::max(x, 1);
	int s = 0;
	while ((x >> s) >= 16)
		s++;
	const long long step = 1ll << s;
	return (int)(((x + step - 1) / step) * step);
}

// fewer than four agents per occupied voxel (3D stencils of a few dozen candidates)
inline bool sparse_population(long long agents, long long occupied)
{
	return occupied > 0 && agents < 4 * occupied;
}

// blocks of one member in a launch of `kind`, for n agents, L spring slots and `nbr` blocks
// of the neighbor search (from the member's occupied voxels)
inline int member_blocks(int kind, long long n, long long L, int nbr)
{
	auto ceil_div = [](long long a, long long b) { return (a + b - 1) / b; };
	long long b = 1;
	switch (kind)
	{
		case MAP_AGENT:
			b = ceil_div(n, BLOCK);
			break;
		case MAP_LANES4:
			b = ceil_div(n * 4, BLOCK);
			break;
		case MAP_LANES8:
			b = ceil_div(n * 8, BLOCK);
			break;
		case MAP_LANES16:
			b = ceil_div(n * 16, BLOCK);
			break;
		case MAP_ERASE:
			b = ceil_div(n * L, BLOCK);
			break;
		case MAP_CLAIM:
			b = ceil_div(n, CLAIM_THREADS);
			break;
		case MAP_NEIGHBORS:
			b = std::min<long long>(n, nbr);
			break;
		case MAP_COMPACT:
			return MOVE_BLOCKS + (int)std::max<long long>(1, ceil_div(n, BLOCK));
		case MAP_ORDER:
			b = ceil_div(n, BLOCK) + ceil_div(n * ERASE_LANES, BLOCK);
			break;
	}
	return (int)std::max<long long>(1, b);
}

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
class sim_impl final : public pcm_sim
{
public:
	explicit sim_impl(const pcm_config& cfg) : cfg_(cfg) {}

	~sim_impl() override
	{
		if (stream_)
		{
			// the handle's own stream goes back to the stream cache once idle
			(void)hipStreamSynchronize(stream_);
			streams().give(stream_, stream_device_);
		}
		for (auto& r : prof_records_)
		{
			(void)hipEventDestroy(r.start);
			(void)hipEventDestroy(r.stop);
		}
		for (hipEvent_t e : event_pool_)
			(void)hipEventDestroy(e);
		for (const auto& a : allocations_)
			pool().give(a.ptr, a.bytes, a.device);
	}

	int init() override
	{
		auto& h = h_;
		std::memset(&h, 0, sizeof(h));
		h.dims = cfg_.mesh.dims;
		h.T = cfg_.cell_defs;
		h.S = cfg_.substrates;
		h.cap = cfg_.capacity;
		for (int k = 0; k < 3; k++)
		{
			h.bmin[k] = cfg_.mesh.bbox_min[k];
			h.bmax[k] = cfg_.mesh.bbox_max[k];
			h.vox[k] = cfg_.mesh.voxel[k];
			if (k < h.dims)
			{
				if (h.vox[k] <= 0 || h.bmax[k] < h.bmin[k])
					return fail(PCM_ERR_ARGUMENT, "bad mesh box or voxel shape");
				h.grid[k] = std::max(1, (h.bmax[k] - h.bmin[k] + h.vox[k] - 1) / h.vox[k]);
			}
			else
				h.grid[k] = 1;
		}
		h.vcount = (long long)h.grid[0] * h.grid[1] * h.grid[2];
		if (h.vcount > 0x7fffffffLL)
			return fail(PCM_ERR_CAPACITY, "mechanics voxel grid larger than 2^31 voxels");
		h.K = h.dims == 3 ? 32 : (h.dims == 2 ? 16 : 8);
		// voxel tables (B slots per voxel) for the fused mesh update: every 3D simulation, and 1D
		// and 2D ones whose capacity keeps them below the lane-per-agent neighbor search (which
		// walks the dense bucket layout the claim and scatter launches build)
		h.B = h.dims == 3 || h.cap < LANES_NEIGHBOR_MIN_AGENTS ? 32 : 0;
		h.L = 12;
		h.LA = 1;
		h.seed = cfg_.seed;
		h.dt = (R)0.1;
		h.springs = 1;

		PCM_HIP(hipGetDevice(&stream_device_));
		PCM_TRY(streams().take(&stream_));
		in_init_ = true;
		const int rc = init_storage();
		in_init_ = false;
		PCM_TRY(rc);
		PCM_TRY(push_ctx()); // waits for the zeroing and the copies of init_storage
		return PCM_OK;
	}

	int init_storage()
	{
		auto& h = h_;
		const int cap = std::max(1, h.cap);
		const int dims = h.dims;

		for (int f = 0; f < PCM_FIELD_COUNT; f++)
		{
			const int w = width(k_fields[f].width_kind);
			const int eb = k_fields[f].type == PCM_T_REAL ? (int)sizeof(R) : (k_fields[f].type == PCM_T_I32 ? 4 : 1);
			h.fbytes[f] = w * eb;
			PCM_TRY(alloc(&h.fptr[f], (size_t)cap * std::max(1, h.fbytes[f])));
		}
		bind_fields();
		// copy plan of the compaction move: every 4-byte word of the word-sized fields, every
		// byte of the byte-sized ones
		h.mv_n = 0;
		for (int f = 0; f < PCM_FIELD_COUNT; f++)
		{
			const bool bytes = k_fields[f].type == PCM_T_U8;
			const int units = bytes ? h.fbytes[f] : h.fbytes[f] / 4;
			for (int u = 0; u < units; u++)
			{
				if (h.mv_n >= (int)(sizeof(h.mv_field) / sizeof(h.mv_field[0])))
					return fail(PCM_ERR_UNSUPPORTED, "too many cell definitions or substrates for the copy plan");
				h.mv_field[h.mv_n] = (uint16_t)f;
				h.mv_unit[h.mv_n] = (uint16_t)u;
				h.mv_bytes[h.mv_n] = bytes ? 1 : 4;
				h.mv_n++;
			}
		}

		PCM_TRY(alloc_t(&h.sc, 1));
		PCM_TRY(alloc_t(&h.spr_cnt, cap));
		PCM_TRY(alloc_t(&h.spr_idx, (size_t)cap * h.L));
		PCM_TRY(alloc_t(&h.atc_cnt, cap));
		PCM_TRY(alloc_t(&h.atc_idx, (size_t)cap * h.LA));
		PCM_TRY(alloc_t(&h.nbr_cnt, cap));
		PCM_TRY(alloc_neighbor_storage());
		PCM_TRY(alloc_t(&h.vcnt, (size_t)h.vcount));
		PCM_TRY(alloc_t(&h.vstart, (size_t)h.vcount));
		PCM_TRY(alloc_t(&h.vox_of, cap));
		PCM_TRY(alloc_t(&h.vrank, cap));
		PCM_TRY(alloc_t(&h.btmp, cap));
		PCM_TRY(alloc_t(&h.occ, cap));
		PCM_TRY(alloc_bucket_storage());
		PCM_TRY(alloc_t(&h.frec, (size_t)cap * 8));
		PCM_TRY(alloc_t(&h.own_cnt, cap));
		PCM_TRY(alloc_t(&h.in_cnt, cap));
		// one segment of BLOCK / lanes actors per collect block (lanes <= MAX_EVENT_LANES)
		PCM_TRY(alloc_t(&h.att_list, (size_t)cap + BLOCK));
		PCM_TRY(alloc_t(&h.blk_cnt, (size_t)cap / (BLOCK / MAX_EVENT_LANES) + 2));
		PCM_TRY(alloc_t(&h.blk_ev, (size_t)cap / (BLOCK / MAX_EVENT_LANES) + 2));
		PCM_TRY(alloc_t(&h.att_c, cap));
		PCM_TRY(alloc_t(&h.att_c2, cap));
		PCM_TRY(alloc_t(&h.fire, cap));
		PCM_TRY(alloc_t(&h.erase, cap));
		PCM_TRY(alloc_t(&h.ev_t, cap));
		PCM_TRY(alloc_t(&h.rem_cnt, cap));
		PCM_TRY(alloc_t(&h.rem_ev, (size_t)cap * 4));
		PCM_TRY(alloc_t(&h.atk_tgt, cap));
		h.bad_cap = next_pow2(cap);
		PCM_TRY(alloc_t(&h.bad, (size_t)h.bad_cap));
		PCM_TRY(alloc_t(&h.remap, cap));
		PCM_TRY(alloc_t(&h.move_src, cap));
		PCM_TRY(alloc_t(&h.move_dst, cap));
		PCM_TRY(alloc_t(&h.div_list, cap));
		h.log_cap = (int)std::min<long long>(0x7ffffff0LL, 4LL * cap + 64);
		PCM_TRY(alloc_t(&h.rm_log, (size_t)h.log_cap));
		(void)dims;

		std::vector<int> ident(cap);
		for (int i = 0; i < cap; i++)
			ident[i] = i;
		// ordered after the zeroing on the same stream; `ident` lives until init waits
		remap_ident_ = std::move(ident);
		PCM_HIP(hipMemcpyAsync(h.remap, remap_ident_.data(), sizeof(int) * cap, hipMemcpyHostToDevice, stream_));
		PCM_TRY(alloc_t(&d_, 1));
		return PCM_OK;
	}

	// ---------------------------------------------------------------- data movement
	int set_count(int n) override
	{
		if (n < 0 || n > h_.cap)
			return fail(PCM_ERR_CAPACITY, "agent count exceeds the handle's capacity");
		PCM_TRY(read_scalars());
		// the voxel counts of the last binning are cleared here, over the agents it binned:
		// later clears cover the new launch bound only, which may be smaller
		if (sc_.n_binned > 0)
		{
			PCM_TRY(launch(PCM_K_BIN_CLEAR, k_bin_clear<R>, (sc_.n_binned + BLOCK - 1) / BLOCK, BLOCK));
			sc_.n_binned = 0;
			sc_.n_occ = 0;
		}
		binned_launch_ = 0;
		sc_.n = n;
		sc_.status = 0;
		sc_.cand_count = 0;
		sc_.conflict = 0;
		sc_.bad_count = 0;
		sc_.moves = 0;
		sc_.claim = 0;
		sc_.log_len = 0;
		sc_.res_done = ~0u;
		// stream-ordered copies and fills, one wait at the end (the sources are members)
		PCM_HIP(hipMemcpyAsync(h_.sc, &sc_, sizeof(Scalars), hipMemcpyHostToDevice, stream_));
		n_launch_ = n;
		mesh_fresh_ = false;
		records_fresh_ = false;
		detach_ready_ = false;
		if ((int)origin_ident_.size() < n)
		{
			const int from = (int)origin_ident_.size();
			origin_ident_.resize(n);
			for (int i = from; i < n; i++)
				origin_ident_[i] = i;
		}
		if (n > 0)
		{
			PCM_HIP(hipMemcpyAsync(h_.origin, origin_ident_.data(), sizeof(int) * n, hipMemcpyHostToDevice, stream_));
			PCM_HIP(hipMemsetAsync(h_.spr_cnt, 0, sizeof(int) * n, stream_));
			PCM_HIP(hipMemsetAsync(h_.atc_cnt, 0, sizeof(int) * n, stream_));
			PCM_HIP(hipMemsetAsync(h_.nbr_cnt, 0, sizeof(int) * n, stream_));
			PCM_HIP(hipMemsetAsync(h_.contact, 0, n, stream_));
			PCM_HIP(hipMemsetAsync(h_.bias_fn, 0, n, stream_));
		}
		PCM_HIP(hipStreamSynchronize(stream_));
		if (h_.contacts || h_.chemotaxis)
		{
			h_.contacts = 0;
			h_.chemotaxis = 0;
			PCM_TRY(push_ctx());
		}
		return PCM_OK;
	}

	int get_count(int* n) override
	{
		PCM_TRY(read_scalars());
		*n = sc_.n;
		return PCM_OK;
	}

	int set_field(int f, const void* src) override
	{
		if (f < 0 || f >= PCM_FIELD_COUNT || !src)
			return fail(PCM_ERR_ARGUMENT, "bad field id or null buffer");
		// the current count: removals since the last upload leave it below n_launch_
		PCM_TRY(read_scalars());
		PCM_TRY(check_field(f, src, sc_.n));
		PCM_HIP(hipMemcpyAsync(h_.fptr[f], src, (size_t)sc_.n * h_.fbytes[f], hipMemcpyHostToDevice, stream_));
		PCM_HIP(hipStreamSynchronize(stream_));
		return PCM_OK;
	}

	// Several fields in one call: one read of the count, every check first (nothing is
	// copied when one fails), then the copies queued on the stream and one wait.
	int set_fields(int nfields, const int32_t* ids, const void* const* srcs) override
	{
		if (nfields < 0 || (nfields > 0 && (!ids || !srcs)))
			return fail(PCM_ERR_ARGUMENT, "null field table");
		for (int k = 0; k < nfields; k++)
			if (ids[k] < 0 || ids[k] >= PCM_FIELD_COUNT || !srcs[k])
				return fail(PCM_ERR_ARGUMENT, "bad field id or null buffer at entry " + std::to_string(k));
		PCM_TRY(read_scalars());
		const int n = sc_.n;
		for (int k = 0; k < nfields; k++)
			PCM_TRY(check_field(ids[k], srcs[k], n));
		for (int k = 0; k < nfields; k++)
			PCM_HIP(hipMemcpyAsync(h_.fptr[ids[k]], srcs[k], (size_t)n * h_.fbytes[ids[k]], hipMemcpyHostToDevice, stream_));
		PCM_HIP(hipStreamSynchronize(stream_));
		return PCM_OK;
	}

	// The checks and host-side consequences of uploading field f (n agents from src):
	// value ranges the kernels index with, spring-row growth, the contact and chemotaxis
	// switches of the context, and which derived device data goes stale.
	int check_field(int f, const void* src, int n)
	{
		if (f == PCM_F_MAXIMUM_NUMBER_OF_ATTACHMENTS && n > 0)
		{
			const int* m = (const int*)src;
			int mx = 0;
			for (int i = 0; i < n; i++)
				mx = std::max(mx, m[i]);
			if (mx > 32)
				return fail(PCM_ERR_UNSUPPORTED, "maximum_number_of_attachments above 32 is not supported");
			if (mx > h_.L)
				PCM_TRY(grow_springs(mx));
		}
		if (f == PCM_F_CHEMOTAXIS_INDEX)
		{
			// gradient rows are read at this index
			for (int i = 0; i < n; i++)
				if (((const int32_t*)src)[i] < 0 || ((const int32_t*)src)[i] >= std::max(h_.S, 1))
					return fail(PCM_ERR_ARGUMENT, "chemotaxis_index out of range");
		}
		if (f == PCM_F_CELL_DEFINITION_INDEX)
		{
			// the per-type rows (affinities, rates, immunogenicities) are read at this index
			for (int i = 0; i < n; i++)
				if (((const int32_t*)src)[i] < 0 || ((const int32_t*)src)[i] >= std::max(h_.T, 1))
					return fail(PCM_ERR_ARGUMENT, "cell_definition_index outside [0, cell_defs)");
		}
		if (f == PCM_F_MIGRATION_BIAS_DIRECTION_FUNCTION)
		{
			const uint8_t* code = (const uint8_t*)src;
			int any = 0;
			for (int i = 0; i < n; i++)
			{
				if (code[i] > PCM_BIAS_ADVANCED_CHEMOTAXIS_NORMALIZED || (code[i] != PCM_BIAS_NONE && h_.S < 1))
					return fail(PCM_ERR_UNSUPPORTED, "unknown direction update code, or chemotaxis without substrates");
				any |= code[i] != PCM_BIAS_NONE;
			}
			if (any != h_.chemotaxis)
			{
				h_.chemotaxis = any;
				PCM_TRY(push_ctx());
			}
		}
		if (f == PCM_F_CONTACT_FUNCTION)
		{
			const uint8_t* code = (const uint8_t*)src;
			int any = 0;
			for (int i = 0; i < n; i++)
			{
				if (code[i] > PCM_CONTACT_STANDARD_ELASTIC)
					return fail(PCM_ERR_UNSUPPORTED, "unknown contact function code");
				any |= code[i];
			}
			if (any != h_.contacts)
			{
				h_.contacts = any;
				PCM_TRY(push_ctx());
			}
		}
		if (f == PCM_F_POSITIONS || f == PCM_F_RADIUS || f == PCM_F_RELATIVE_MAXIMUM_ADHESION_DISTANCE)
			mesh_fresh_ = false;
		records_fresh_ = false;
		detach_ready_ = false;
		return PCM_OK;
	}

	int get_field(int f, void* dst) override
	{
		return get_fields(1, &f, &dst);
	}

	// Several fields in one call: one read of the count, the copies queued together into a
	// pinned block, one wait.
	int get_fields(int nfields, const int32_t* ids, void* const* dsts) override
	{
		if (nfields < 0 || (nfields > 0 && (!ids || !dsts)))
			return fail(PCM_ERR_ARGUMENT, "null field table");
		for (int k = 0; k < nfields; k++)
			if (ids[k] < 0 || ids[k] >= PCM_FIELD_COUNT || !dsts[k])
				return fail(PCM_ERR_ARGUMENT, "bad field id or null buffer at entry " + std::to_string(k));
		PCM_TRY(read_scalars());
		std::vector<staged_read> reads(nfields);
		for (int k = 0; k < nfields; k++)
			reads[k] = { dsts[k], h_.fptr[ids[k]], (size_t)sc_.n * h_.fbytes[ids[k]] };
		return read_staged(reads.data(), nfields);
	}

	int set_links(int which, const int32_t* off, const int32_t* idx) override
	{
		detach_ready_ = false;
		if (which != PCM_LINKS_SPRINGS && which != PCM_LINKS_ATTACHED)
			return fail(PCM_ERR_ARGUMENT, "bad link list id");
		if (!off)
			return fail(PCM_ERR_ARGUMENT, "null offsets");
		PCM_TRY(read_scalars());
		const int n = sc_.n;
		int longest = 0;
		for (int i = 0; i < n; i++)
		{
			if (off[i + 1] < off[i])
				return fail(PCM_ERR_ARGUMENT, "offsets must be non-decreasing");
			longest = std::max(longest, off[i + 1] - off[i]);
		}
		if (which == PCM_LINKS_SPRINGS)
		{
			if (longest > 32)
				return fail(PCM_ERR_UNSUPPORTED, "more than 32 springs on one agent");
			if (longest > h_.L)
				PCM_TRY(grow_springs(longest));
		}
		else if (longest > h_.LA)
			PCM_TRY(grow_attached(longest));
		for (int i = 0; i < n; i++)
			for (int s = off[i]; s < off[i + 1]; s++)
				if (!idx || idx[s] < 0 || idx[s] >= n || idx[s] == i)
					return fail(PCM_ERR_ARGUMENT, "link list entry is not another agent's index");
		const int W = which == PCM_LINKS_SPRINGS ? h_.L : h_.LA;
		std::vector<int> cnt(std::max(n, 1)), rows((size_t)std::max(n, 1) * W, -1);
		for (int i = 0; i < n; i++)
		{
			cnt[i] = off[i + 1] - off[i];
			for (int s = 0; s < cnt[i]; s++)
				rows[(size_t)i * W + s] = idx[off[i] + s];
		}
		int* dc = which == PCM_LINKS_SPRINGS ? h_.spr_cnt : h_.atc_cnt;
		int* di = which == PCM_LINKS_SPRINGS ? h_.spr_idx : h_.atc_idx;
		if (n > 0)
		{
			PCM_TRY(copy_sync(dc, cnt.data(), sizeof(int) * n, hipMemcpyHostToDevice));
			PCM_TRY(copy_sync(di, rows.data(), sizeof(int) * (size_t)n * W, hipMemcpyHostToDevice));
		}
		return PCM_OK;
	}

	int get_links(int which, int32_t* off, int32_t* idx, int64_t* total) override
	{
		if (which != PCM_LINKS_SPRINGS && which != PCM_LINKS_ATTACHED)
			return fail(PCM_ERR_ARGUMENT, "bad link list id");
		PCM_TRY(read_scalars());
		const int n = sc_.n;
		const int W = which == PCM_LINKS_SPRINGS ? h_.L : h_.LA;
		// the rows only when entries are wanted (a size query needs the counts)
		std::vector<int> cnt(std::max(n, 1)), rows(idx ? (size_t)std::max(n, 1) * W : 0);
		if (n > 0)
		{
			const staged_read reads[2] = {
				{ cnt.data(), which == PCM_LINKS_SPRINGS ? h_.spr_cnt : h_.atc_cnt, sizeof(int) * (size_t)n },
				{ rows.data(), which == PCM_LINKS_SPRINGS ? h_.spr_idx : h_.atc_idx, idx ? sizeof(int) * (size_t)n * W : 0 },
			};
			PCM_TRY(read_staged(reads, 2));
		}
		int64_t t = 0;
		for (int i = 0; i < n; i++)
		{
			if (off)
				off[i] = (int32_t)t;
			for (int s = 0; s < cnt[i]; s++, t++)
				if (idx)
					idx[t] = rows[(size_t)i * W + s];
		}
		if (off)
			off[n] = (int32_t)t;
		*total = t;
		return PCM_OK;
	}

	int get_neighbors(int32_t* off, int32_t* idx, int64_t* total) override
	{
		PCM_TRY(read_scalars());
		const int n = sc_.n;
		const int K = h_.K;
		std::vector<int> cnt(std::max(n, 1));
		std::vector<uint8_t> mov(std::max(n, 1));
		// the rows only when entries are wanted (a size query needs the counts)
		const bool rows_wanted = off || idx;
		std::unique_ptr<int[]> rows(new int[rows_wanted ? (size_t)std::max(n, 1) * K : 1]);
		if (n > 0)
		{
			const staged_read reads[3] = {
				{ cnt.data(), h_.nbr_cnt, sizeof(int) * (size_t)n },
				{ mov.data(), h_.movable, (size_t)n },
				{ rows.get(), h_.nbr_idx, rows_wanted ? sizeof(int) * (size_t)n * K : 0 },
			};
			PCM_TRY(read_staged(reads, 3));
		}
		int64_t t = 0;
		for (int i = 0; i < n; i++)
			t += mov[i] ? cnt[i] : 0;
		*total = t;
		if (!rows_wanted)
			return PCM_OK;
		t = 0;
		for (int i = 0; i < n; i++)
		{
			if (off)
				off[i] = (int32_t)t;
			if (!mov[i])
				continue;
			if (idx)
				std::copy(rows.get() + (size_t)i * K, rows.get() + (size_t)i * K + cnt[i], idx + t);
			t += cnt[i];
		}
		if (off)
			off[n] = (int32_t)t;
		return PCM_OK;
	}

	// Host neighbor lists (states.neighbors).  A non-movable agent's reference list is
	// empty; its device list (which feeds simple_pressure) is rebuilt as the movable agents
	// that list it, in ascending order.
	int set_neighbors(const int32_t* off, const int32_t* idx) override
	{
		detach_ready_ = false;
		if (!off || !idx)
			return fail(PCM_ERR_ARGUMENT, "null neighbor lists");
		PCM_TRY(read_scalars());
		const int n = sc_.n;
		std::vector<uint8_t> mov(std::max(n, 1));
		if (n > 0)
			PCM_TRY(copy_sync(mov.data(), h_.movable, n, hipMemcpyDeviceToHost));
		// two passes over the flat lists: count (a movable agent's own entries, a
		// non-movable agent's listers), then fill the fixed-stride rows
		std::vector<int> cnt(std::max(n, 1), 0);
		for (int i = 0; i < n; i++)
		{
			if (off[i + 1] < off[i])
				return fail(PCM_ERR_ARGUMENT, "offsets must be non-decreasing");
			if (!mov[i])
				continue;
			cnt[i] += off[i + 1] - off[i];
			for (int s = off[i]; s < off[i + 1]; s++)
			{
				const int j = idx[s];
				if (j < 0 || j >= n || j == i)
					return fail(PCM_ERR_ARGUMENT, "neighbor entry is not another agent's index");
				if (!mov[j])
					cnt[j]++;
			}
		}
		int longest = 0;
		for (int i = 0; i < n; i++)
			longest = std::max(longest, cnt[i]);
		if (longest > h_.K)
			PCM_TRY(grow_neighbors(longest));
		const int K = h_.K;
		std::unique_ptr<int[]> rows(new int[(size_t)std::max(n, 1) * K]);
		std::vector<int> fill(std::max(n, 1), 0);
		for (int i = 0; i < n; i++) // ascending: a non-movable agent's listers in order
		{
			if (!mov[i])
				continue;
			std::copy(idx + off[i], idx + off[i + 1], rows.get() + (size_t)i * K);
			for (int s = off[i]; s < off[i + 1]; s++)
				if (!mov[idx[s]])
					rows[(size_t)idx[s] * K + fill[idx[s]]++] = i;
		}
		for (int i = 0; i < n; i++)
			std::fill(rows.get() + (size_t)i * K + cnt[i], rows.get() + (size_t)(i + 1) * K, -1);
		if (n > 0)
		{
			PCM_TRY(copy_sync(h_.nbr_cnt, cnt.data(), sizeof(int) * n, hipMemcpyHostToDevice));
			PCM_TRY(copy_sync(h_.nbr_idx, rows.get(), sizeof(int) * (size_t)n * K, hipMemcpyHostToDevice));
		}
		return PCM_OK;
	}

	long long voxel_count() const override { return h_.vcount; }

	int get_buckets(int32_t* counts, int32_t* agents) override
	{
		PCM_TRY(read_scalars());
		const size_t V = (size_t)h_.vcount;
		std::vector<int> vc(V), vs(V);
		PCM_TRY(copy_sync(vc.data(), h_.vcnt, sizeof(int) * V, hipMemcpyDeviceToHost));
		PCM_TRY(copy_sync(vs.data(), h_.vstart, sizeof(int) * V, hipMemcpyDeviceToHost));
		// bucket slots in use: one range per occupied voxel (dense, or B apart with voxel tables)
		size_t span = 0;
		for (size_t v = 0; v < V; v++)
			if (vc[v] > 0)
				span = std::max(span, (size_t)vs[v] + (size_t)vc[v]);
		std::vector<int> b(std::max<size_t>(span, 1));
		if (span > 0)
			PCM_TRY(copy_sync(b.data(), h_.bucket, sizeof(int) * span, hipMemcpyDeviceToHost));
		int64_t t = 0;
		for (size_t v = 0; v < V; v++)
		{
			counts[v] = vc[v];
			for (int k = 0; k < vc[v]; k++)
				agents[t++] = b[vs[v] + k];
		}
		return PCM_OK;
	}

	int get_deaths(int64_t* d) override
	{
		PCM_TRY(read_scalars());
		*d = sc_.deaths;
		return PCM_OK;
	}

	// the per-cell output matrix, packed on the device (PhysiCell_MultiCellDS.cpp:1735-1790)
	int pack_fields(int nfields, const int32_t* fields, const int32_t* widths, double* out) override
	{
		std::vector<int> fld, comp;
		for (int k = 0; k < nfields; k++)
		{
			if (fields[k] < 0 || fields[k] >= PCM_FIELD_COUNT || widths[k] < 0)
				return fail(PCM_ERR_ARGUMENT, "bad field id or width");
			const int own = width(k_fields[fields[k]].width_kind);
			const int w = widths[k] == 0 ? own : widths[k];
			for (int m = 0; m < w; m++)
			{
				fld.push_back(fields[k]);
				comp.push_back(m);
			}
		}
		const int row = (int)fld.size();
		if (row == 0 || row > 256)
			return fail(PCM_ERR_ARGUMENT, "between 1 and 256 output columns");
		PCM_TRY(read_scalars());
		const long long total = (long long)sc_.n * row;
		if (total == 0)
			return PCM_OK;
		if (!pack_cols_)
			PCM_TRY(alloc_t(&pack_cols_, 512));
		if ((size_t)total > pack_cap_)
		{
			release(pack_out_);
			pack_out_ = nullptr;
			PCM_TRY(alloc_t(&pack_out_, (size_t)total));
			pack_cap_ = (size_t)total;
		}
		std::vector<int> cols(512, 0);
		std::copy(fld.begin(), fld.end(), cols.begin());
		std::copy(comp.begin(), comp.end(), cols.begin() + 256);
		PCM_TRY(copy_sync(pack_cols_, cols.data(), sizeof(int) * 512, hipMemcpyHostToDevice));
		const int blocks = (int)((total + BLOCK - 1) / BLOCK);
		hipLaunchKernelGGL(k_pack<R>, dim3(blocks), dim3(BLOCK), 0, stream_, (const SimCtx<R>*)d_, (const int4*)nullptr,
						   (const int*)pack_cols_, (const int*)(pack_cols_ + 256), row, pack_out_);
		PCM_HIP(hipGetLastError());
		PCM_HIP(hipStreamSynchronize(stream_));
		PCM_TRY(copy_sync(out, pack_out_, sizeof(double) * (size_t)total, hipMemcpyDeviceToHost));
		return PCM_OK;
	}

	int get_divisions(int64_t* d) override
	{
		PCM_TRY(read_scalars());
		*d = sc_.divisions;
		return PCM_OK;
	}

	int division_log(int32_t* parents, int64_t capacity, int64_t* length) override
	{
		*length = div_log_len_;
		if (!parents || div_log_len_ == 0)
			return PCM_OK;
		if (capacity < div_log_len_)
			return fail(PCM_ERR_CAPACITY, "division log longer than the buffer");
		PCM_HIP(hipStreamSynchronize(stream_));
		PCM_TRY(copy_sync(parents, h_.div_list, sizeof(int32_t) * div_log_len_, hipMemcpyDeviceToHost));
		return PCM_OK;
	}

	// update_cell_container_for_phenotype (containers_solver.cpp:149-187): plan the
	// divisions, check the capacity before anything changes, divide, then the removal loop
	// over the agents that existed before
	int phenotype_update(const pcm_step_params* p, int compute_internalized) override
	{
		if (p)
			PCM_TRY(set_params(*p));
		PCM_TRY(read_scalars());
		if (sc_.status)
			return fail(PCM_ERR_STATE, "device status word is set");
		h_.div_epoch = division_epoch_;
		h_.div_internal = compute_internalized != 0;
		PCM_TRY(push_ctx());
		PCM_TRY(launch(PCM_K_DIVIDE_PLAN, k_divide_plan<R>, 1, 1024));
		PCM_TRY(read_scalars());
		const long long total = (long long)sc_.n + sc_.div_count;
		if (total > h_.cap)
			return fail(PCM_ERR_CAPACITY, "divisions would exceed the handle's agent capacity");
		division_epoch_++;
		div_log_len_ = sc_.div_count;
		n_launch_ = (int)total;
		mesh_fresh_ = false;
		records_fresh_ = false;
		detach_ready_ = false;
		const int blocks = std::max(1, (sc_.div_count + BLOCK - 1) / BLOCK);
		switch (h_.dims)
		{
			case 1:
				PCM_TRY(launch(PCM_K_DIVIDE_APPLY, k_divide_apply<R, 1>, blocks, BLOCK));
				break;
			case 2:
				PCM_TRY(launch(PCM_K_DIVIDE_APPLY, k_divide_apply<R, 2>, blocks, BLOCK));
				break;
			default:
				PCM_TRY(launch(PCM_K_DIVIDE_APPLY, k_divide_apply<R, 3>, blocks, BLOCK));
		}
		PCM_TRY(per_agent(PCM_K_DIVIDE_LINKS, k_divide_links<R>));
		PCM_TRY(launch(PCM_K_COMPACT_DONE, k_divide_done<R>, blocks, BLOCK));
		switch (h_.dims)
		{
			case 1:
				return container_kernels<1>();
			case 2:
				return container_kernels<2>();
			default:
				return container_kernels<3>();
		}
	}

	int set_step_params(const pcm_step_params& p) override { return set_params(p); }

	int set_step(uint64_t s) override
	{
		detach_ready_ = false;
		PCM_TRY(read_scalars());
		sc_.step = s;
		sc_.res_done = ~0u; // no step's later resolve rounds are marked done
		PCM_TRY(copy_sync(h_.sc, &sc_, sizeof(Scalars), hipMemcpyHostToDevice));
		return PCM_OK;
	}

	int get_step(uint64_t* s) override
	{
		PCM_TRY(read_scalars());
		*s = sc_.step;
		return PCM_OK;
	}

	int set_gradients(const int32_t* voxel, const void* g) override
	{
		detach_ready_ = false;
		int grid[3];
		size_t voxels = 1;
		for (int k = 0; k < 3; k++)
		{
			if (voxel[k] < 1)
				return fail(PCM_ERR_ARGUMENT, "diffusion voxel shape must be positive");
			const int extent = cfg_.mesh.bbox_max[k] - cfg_.mesh.bbox_min[k];
			grid[k] = k < h_.dims ? std::max(1, (extent + voxel[k] - 1) / voxel[k]) : 1;
			voxels *= (size_t)grid[k];
		}
		const size_t count = voxels * (size_t)h_.S * (size_t)h_.dims;
		PCM_HIP(hipStreamSynchronize(stream_));
		if (count != grad_count_)
		{
			release((void*)h_.grad);
			R* buf = nullptr;
			PCM_TRY(alloc_t(&buf, std::max<size_t>(count, 1)));
			h_.grad = (gptr<const R>)buf;
			grad_count_ = count;
		}
		if (count > 0)
			PCM_TRY(copy_sync((void*)h_.grad, g, sizeof(R) * count, hipMemcpyHostToDevice));
		for (int k = 0; k < 3; k++)
		{
			h_.dvox[k] = voxel[k];
			h_.dgrid[k] = grid[k];
		}
		return push_ctx();
	}

	// chemotaxis reads the gradients: no step or phase runs without them
	int ready() const
	{
		if (h_.chemotaxis && !h_.grad)
			return fail(PCM_ERR_STATE, "chemotaxis without gradients (pcm_set_gradients)");
		return PCM_OK;
	}

	// ---------------------------------------------------------------- phases
	int phase(int which, const pcm_step_params* p) override
	{
		PCM_TRY(ready());
		detach_ready_ = false;
		if (p)
			PCM_TRY(set_params(*p));
		switch (h_.dims)
		{
			case 1:
				return phase_d<1>(which);
			case 2:
				return phase_d<2>(which);
			default:
				return phase_d<3>(which);
		}
	}

	int step(const pcm_step_params* p, int steps) override
	{
		if (!p || steps < 0)
			return fail(PCM_ERR_ARGUMENT, "null params or negative step count");
		PCM_TRY(ready());
		PCM_TRY(set_params(*p));
		PCM_TRY(read_scalars());
		if (sc_.status)
			return fail(PCM_ERR_STATE, "device status word is set");
		if (!tune_chosen_)
			tune_ = default_tuning(sc_.n);
		tables_ = h_.B > 0;
		// alone, a small population's launch is bound by its latency, not by the blocks a CU
		// can hold: below a few thousand agents the default shape stays
		neighbors_sparse_ = sc_.n >= LANES_NEIGHBOR_MIN_AGENTS && sparse_population(sc_.n, sc_.n_occ);
		const unsigned long long target = sc_.step + (unsigned long long)steps;
		for (int guard = 0; guard < 64; guard++)
		{
			const long long remaining = (long long)(target - sc_.step);
			for (long long s = 0; s < remaining; s++)
				PCM_TRY(dispatch_step());
			PCM_TRY(read_scalars());
			if (sc_.status == 0)
				return PCM_OK;
			PCM_TRY(recover());
			if (sc_.step >= target)
				return PCM_OK;
		}
		return fail(PCM_ERR_STATE, "mechanics step did not converge after repeated neighbor-capacity growth");
	}

	// positions of the removals since the last read, then the log restarts
	int removal_log(int32_t* positions, int64_t capacity, int64_t* length) override
	{
		PCM_TRY(read_scalars());
		const long long len = sc_.log_len;
		*length = len;
		if (!positions)
			return PCM_OK;
		if (len > h_.log_cap)
			return fail(PCM_ERR_CAPACITY, "removal log overflowed: read it at least every " +
											   std::to_string(h_.log_cap) + " removals");
		if (capacity < len)
			return fail(PCM_ERR_ARGUMENT, "output buffer smaller than the removal log");
		if (len > 0)
			PCM_TRY(copy_sync(positions, h_.rm_log, sizeof(int) * len, hipMemcpyDeviceToHost));
		sc_.log_len = 0;
		PCM_TRY(copy_sync(&h_.sc->log_len, &sc_.log_len, sizeof(int), hipMemcpyHostToDevice));
		return PCM_OK;
	}

	int stats(int32_t* out) override
	{
		PCM_TRY(read_scalars());
		for (int k = 0; k < 8; k++)
			out[k] = sc_.stat[k];
		out[6] = sc_.n_occ; // occupied voxels, as the step's position update counted them
		// attachment events and actors: summed over the collect segments of the last step
		const int segments = sc_.stat[7];
		out[0] = out[2] = out[7] = 0;
		if (segments > 0)
		{
			std::vector<int> cnt(segments), ev(segments);
			PCM_TRY(copy_sync(cnt.data(), h_.blk_cnt, sizeof(int) * segments, hipMemcpyDeviceToHost));
			PCM_TRY(copy_sync(ev.data(), h_.blk_ev, sizeof(int) * segments, hipMemcpyDeviceToHost));
			for (int b = 0; b < segments; b++)
			{
				out[0] += ev[b];
				out[2] += cnt[b];
			}
		}
		return PCM_OK;
	}

	int synchronize() override
	{
		PCM_TRY(read_scalars());
		if (sc_.status)
			return fail(PCM_ERR_STATE, "device status word is set: " + std::to_string(sc_.status));
		return PCM_OK;
	}

private:
	pcm_config cfg_;
	SimCtx<R> h_ {};
	SimCtx<R>* d_ = nullptr;
	Scalars sc_ {};
	hipStream_t stream_ = nullptr;
	int stream_device_ = 0; // the device stream_ was created on
	struct allocation
	{
		void* ptr;
		size_t bytes; // the pool block's size
		int device;
	};
	std::vector<allocation> allocations_;
	bool in_init_ = false; // allocations zero their memory asynchronously; init waits once at its end
	int n_launch_ = 0; // agents at the last upload: an upper bound of the device count
	// launch target: a batch points these at its context array for one segmented launch
	// over all its simulations (grid.y = simulation); grid bounds then cover every member
	const SimCtx<R>* launch_ctx_ = nullptr; // nullptr: this simulation's own context
	const int4* flat_maps_ = nullptr;       // a batch's block maps of a flattened launch, or nullptr
	int flat_off_[MAP_KINDS] = {};          // first entry of each map kind in flat_maps_
	int flat_blocks_[MAP_KINDS] = {};       // grid size of each map kind (padded)
	int launch_y_ = 1;
	int grid_n_ = 0; // agent bound of grid sizes (0: n_launch_)
	int grid_L_ = 0; // spring-slot bound of grid sizes (0: h_.L)
	uint64_t ctx_version_ = 0; // increments whenever the device context changes
	size_t grad_count_ = 0;    // reals in h_.grad
	uint64_t division_epoch_ = 0; // phenotype-update calls so far
	int* pack_cols_ = nullptr;     // column table of pcm_pack_fields: 256 field ids, 256 components
	double* pack_out_ = nullptr;   // its output matrix
	size_t pack_cap_ = 0;
	int div_log_len_ = 0;         // divisions of the last phenotype update
	std::vector<int> remap_ident_; // host source of init's identity copy
	std::vector<int> origin_ident_; // host source of set_count's ORIGIN identity
	pcm_tuning tune_ = { 8, 8, 2 }; // launch shape of the lane-parallel kernels
	bool tune_chosen_ = false;      // set by pcm_set_tuning or the autotuner; else by size
	// the fused mesh update fills the voxel tables (no claim or scatter launch): set per step,
	// or by a batch for its group (only when every member has tables)
	bool tables_ = false;

	// the launch shape when none was chosen: from 65536 agents up four lanes per agent (more
	// agents per wavefront: 147 against 149 us per step at 88.7k agents), below that eight
	// (more lanes spread fewer agents: 1-3% faster from 900 to 15k agents)
	static pcm_tuning default_tuning(int n)
	{
		return n >= 65536 ? pcm_tuning { 4, 4, 2 } : pcm_tuning { 8, 8, 2 };
	}
	void* batch_ = nullptr;    // the batch this simulation belongs to, if any
	bool mesh_fresh_ = false; // voxel buckets describe the current positions and radii
	// a 3D population with few agents per occupied voxel: its neighbor search stages short
	// stencils, one wavefront per voxel with a small tile (set from the last scalars read, or
	// by a batch for its group)
	bool neighbors_sparse_ = false;
	bool records_fresh_ = false; // force records describe the current agent fields
	// the last work enqueued was a fused step, whose mesh update decided the next step's
	// detachments (fired and erased spring slots), and nothing has changed the agents, their
	// lists, the step counter or the step parameters since
	bool detach_ready_ = false;
	int binned_launch_ = 1 << 30; // launch bound of the last mesh update (>= agents it binned)

	int width(int kind) const
	{
		switch (kind)
		{
			case PCM_W_DIMS:
				return h_.dims;
			case PCM_W_DEFS:
				return h_.T;
			case PCM_W_SUBS:
				return h_.S;
			default:
				return 1;
		}
	}

	// zeroed device memory into any pointer type (global-address-space pointers included)
	template <class P>
	int alloc(P* p, size_t bytes)
	{
		void* v = nullptr;
		size_t got = 0;
		int dev = 0;
		PCM_TRY(pool().take(bytes, &v, &got, &dev));
		allocations_.push_back({ v, got, dev });
		PCM_HIP(hipMemsetAsync(v, 0, std::max<size_t>(bytes, 16), stream_));
		// outside init the caller may follow with copies on the null stream
		if (!in_init_)
			PCM_HIP(hipStreamSynchronize(stream_));
		*p = (P)v;
		return PCM_OK;
	}

	template <class T>
	int alloc_t(T** p, size_t n)
	{
		void* v = nullptr;
		PCM_TRY(alloc(&v, sizeof(T) * n));
		*p = (T*)v;
		return PCM_OK;
	}

	// the caller has synchronized the stream: the block is idle
	void release(void* p)
	{
		auto it = std::find_if(allocations_.begin(), allocations_.end(), [p](const allocation& a) { return a.ptr == p; });
		if (it != allocations_.end())
		{
			pool().give(it->ptr, it->bytes, it->device);
			allocations_.erase(it);
		}
	}

	void bind_fields()
	{
		auto& h = h_;
		auto F = [&](int f) { return h.fptr[f]; };
		h.pos = (gptr<R>)F(PCM_F_POSITIONS);
		h.vel = (gptr<R>)F(PCM_F_VELOCITIES);
		h.pvel = (gptr<R>)F(PCM_F_PREVIOUS_VELOCITIES);
		h.radius = (gptr<R>)F(PCM_F_RADIUS);
		h.nuc_radius = (gptr<R>)F(PCM_F_NUCLEAR_RADIUS);
		h.surface = (gptr<R>)F(PCM_F_SURFACE_AREA);
		h.volume = (gptr<R>)F(PCM_F_VOLUME);
		h.fluid = (gptr<R>)F(PCM_F_FLUID);
		h.solid = (gptr<R>)F(PCM_F_SOLID);
		h.nuclear = (gptr<R>)F(PCM_F_NUCLEAR);
		h.nuc_fluid = (gptr<R>)F(PCM_F_NUCLEAR_FLUID);
		h.nuc_solid = (gptr<R>)F(PCM_F_NUCLEAR_SOLID);
		h.cyto = (gptr<R>)F(PCM_F_CYTOPLASMIC);
		h.cyto_fluid = (gptr<R>)F(PCM_F_CYTOPLASMIC_FLUID);
		h.cyto_solid = (gptr<R>)F(PCM_F_CYTOPLASMIC_SOLID);
		h.fluid_frac = (gptr<R>)F(PCM_F_FLUID_FRACTION);
		h.c2n = (gptr<R>)F(PCM_F_CYTOPLASMIC_TO_NUCLEAR_RATIO);
		h.target_cyto = (gptr<R>)F(PCM_F_TARGET_SOLID_CYTOPLASMIC);
		h.target_nuc = (gptr<R>)F(PCM_F_TARGET_SOLID_NUCLEAR);
		h.adhesion = (gptr<R>)F(PCM_F_CELL_CELL_ADHESION_STRENGTH);
		h.repulsion = (gptr<R>)F(PCM_F_CELL_CELL_REPULSION_STRENGTH);
		h.bm_repulsion = (gptr<R>)F(PCM_F_CELL_BM_REPULSION_STRENGTH);
		h.rel_adh = (gptr<R>)F(PCM_F_RELATIVE_MAXIMUM_ADHESION_DISTANCE);
		h.affinity = (gptr<R>)F(PCM_F_CELL_ADHESION_AFFINITIES);
		h.max_att = (gptr<int>)F(PCM_F_MAXIMUM_NUMBER_OF_ATTACHMENTS);
		h.elastic = (gptr<R>)F(PCM_F_ATTACHMENT_ELASTIC_CONSTANT);
		h.att_rate = (gptr<R>)F(PCM_F_ATTACHMENT_RATE);
		h.det_rate = (gptr<R>)F(PCM_F_DETACHMENT_RATE);
		h.motile = (gptr<uint8_t>)F(PCM_F_IS_MOTILE);
		h.persistence = (gptr<R>)F(PCM_F_PERSISTENCE_TIME);
		h.speed = (gptr<R>)F(PCM_F_MIGRATION_SPEED);
		h.bias_dir = (gptr<R>)F(PCM_F_MIGRATION_BIAS_DIRECTION);
		h.bias = (gptr<R>)F(PCM_F_MIGRATION_BIAS);
		h.motility = (gptr<R>)F(PCM_F_MOTILITY_VECTOR);
		h.restrict2d = (gptr<uint8_t>)F(PCM_F_RESTRICT_TO_2D);
		h.dead = (gptr<uint8_t>)F(PCM_F_DEAD);
		h.dead_phago = (gptr<R>)F(PCM_F_DEAD_PHAGOCYTOSIS_RATE);
		h.live_phago = (gptr<R>)F(PCM_F_LIVE_PHAGOCYTOSIS_RATES);
		h.damage_rate = (gptr<R>)F(PCM_F_DAMAGE_RATE);
		h.attack = (gptr<R>)F(PCM_F_ATTACK_RATES);
		h.immuno = (gptr<R>)F(PCM_F_IMMUNOGENICITIES);
		h.fusion = (gptr<R>)F(PCM_F_FUSION_RATES);
		h.pressure = (gptr<R>)F(PCM_F_SIMPLE_PRESSURE);
		h.nuclei = (gptr<int>)F(PCM_F_NUMBER_OF_NUCLEI);
		h.damage = (gptr<R>)F(PCM_F_DAMAGE);
		h.attack_time = (gptr<R>)F(PCM_F_TOTAL_ATTACK_TIME);
		h.type = (gptr<int>)F(PCM_F_CELL_DEFINITION_INDEX);
		h.movable = (gptr<uint8_t>)F(PCM_F_IS_MOVABLE);
		h.flags = (gptr<uint8_t>)F(PCM_F_FLAGS);
		h.internal = (gptr<R>)F(PCM_F_INTERNALIZED_SUBSTRATES);
		h.frac_ingest = (gptr<R>)F(PCM_F_FRACTION_TRANSFERRED_WHEN_INGESTED);
		h.origin = (gptr<int>)F(PCM_F_ORIGIN);
		h.contact = (gptr<uint8_t>)F(PCM_F_CONTACT_FUNCTION);
		h.chemo_index = (gptr<int>)F(PCM_F_CHEMOTAXIS_INDEX);
		h.chemo_dir = (gptr<int>)F(PCM_F_CHEMOTAXIS_DIRECTION);
		h.sensitivities = (gptr<R>)F(PCM_F_CHEMOTACTIC_SENSITIVITIES);
		h.bias_fn = (gptr<uint8_t>)F(PCM_F_MIGRATION_BIAS_DIRECTION_FUNCTION);
		h.orientation = (gptr<R>)F(PCM_F_ORIENTATION);
		h.polarity = (gptr<R>)F(PCM_F_POLARITY);
	}

	int alloc_neighbor_storage()
	{
		const size_t cap = std::max(1, h_.cap);
		const size_t K = h_.K;
		PCM_TRY(alloc_t(&h_.nbr_idx, cap * K));
		PCM_TRY(alloc_t(&h_.own_tgt, cap * K));
		PCM_TRY(alloc_t(&h_.own_st, cap * K));
		PCM_TRY(alloc_t(&h_.inbox, cap * K * 2));
		h_.cand_cap = (int)std::min<size_t>(0x7ffffff0, cap * K + 16);
		PCM_TRY(alloc_t(&h_.cand, (size_t)h_.cand_cap * 2));
		return PCM_OK;
	}

	// bucket slots: with voxel tables (3D, and 1D and 2D below the lane search), B per occupied voxel
	// (at most one voxel per agent), else one per agent
	int alloc_bucket_storage()
	{
		const size_t cap = std::max(1, h_.cap);
		const size_t slots = cap * (size_t)std::max(1, h_.B);
		PCM_TRY(alloc_t(&h_.bucket, slots));
		PCM_TRY(alloc_t(&h_.sorted, slots * 4));
		if (h_.B > 0)
			PCM_TRY(alloc_t(&h_.vmem, (size_t)h_.vcount * h_.B));
		return PCM_OK;
	}

	// a voxel held more agents than its unordered table: grow the tables (nothing in them
	// survives a mesh update)
	int grow_buckets(int needed)
	{
		PCM_HIP(hipStreamSynchronize(stream_));
		release(h_.bucket);
		release(h_.sorted);
		release(h_.vmem);
		h_.B = next_pow2(std::max(needed, 2 * h_.B));
		PCM_TRY(alloc_bucket_storage());
		return push_ctx();
	}

	int grow_neighbors(int needed)
	{
		PCM_HIP(hipStreamSynchronize(stream_));
		release(h_.nbr_idx);
		release(h_.own_tgt);
		release(h_.own_st);
		release(h_.inbox);
		release(h_.cand);
		h_.K = next_pow2(std::max(needed, 2 * h_.K));
		PCM_TRY(alloc_neighbor_storage());
		return push_ctx();
	}

	template <class P>
	int grow_rows(P* idx, int oldW, int newW)
	{
		const size_t cap = std::max(1, h_.cap);
		std::vector<int> rows(cap * oldW), grown(cap * newW, -1);
		PCM_TRY(copy_sync(rows.data(), (void*)*idx, sizeof(int) * cap * oldW, hipMemcpyDeviceToHost));
		for (size_t i = 0; i < cap; i++)
			for (int s = 0; s < oldW; s++)
				grown[i * newW + s] = rows[i * oldW + s];
		release((void*)*idx);
		PCM_TRY(alloc(idx, sizeof(int) * cap * newW));
		PCM_TRY(copy_sync((void*)*idx, grown.data(), sizeof(int) * cap * newW, hipMemcpyHostToDevice));
		return PCM_OK;
	}

	int grow_springs(int L)
	{
		PCM_HIP(hipStreamSynchronize(stream_));
		PCM_TRY(grow_rows(&h_.spr_idx, h_.L, L));
		h_.L = L;
		return push_ctx();
	}

	int grow_attached(int LA)
	{
		PCM_HIP(hipStreamSynchronize(stream_));
		PCM_TRY(grow_rows(&h_.atc_idx, h_.LA, LA));
		h_.LA = LA;
		return push_ctx();
	}

	int push_ctx()
	{
		ctx_version_++;
		PCM_HIP(hipMemcpyAsync(d_, &h_, sizeof(h_), hipMemcpyHostToDevice, stream_));
		// the host copy may change before the async copy runs
		PCM_HIP(hipStreamSynchronize(stream_));
		return PCM_OK;
	}

	int set_params(const pcm_step_params& p)
	{
		const R dt = (R)p.mechanics_time_step;
		const int wall = p.virtual_wall_at_domain_edges != 0;
		const int springs = p.automated_spring_adhesion != 0;
		if (dt == h_.dt && wall == h_.wall && springs == h_.springs)
			return PCM_OK;
		detach_ready_ = false; // the fired slots were drawn against the old rate * dt
		h_.dt = dt;
		h_.wall = wall;
		h_.springs = springs;
		return push_ctx();
	}

	// a copy ordered on this simulation's stream, waited for before returning: copies of
	// different simulations (a serving loop's worker threads) do not queue behind each
	// other on the null stream
	struct staged_read
	{
		void* dst;       // host destination (any memory)
		const void* src; // device source
		size_t bytes;
	};

	// device -> pinned block (all reads queued together) -> one wait -> host copies out
	int read_staged(const staged_read* reads, int count)
	{
		size_t total = 0;
		for (int k = 0; k < count; k++)
			total += (reads[k].bytes + 255) & ~(size_t)255;
		if (total == 0)
			return PCM_OK;
		void* block = nullptr;
		size_t got = 0;
		PCM_TRY(pinned().take(total, &block, &got));
		size_t off = 0;
		hipError_t e = hipSuccess;
		for (int k = 0; k < count && e == hipSuccess; k++)
		{
			if (reads[k].bytes)
				e = hipMemcpyAsync((char*)block + off, reads[k].src, reads[k].bytes, hipMemcpyDeviceToHost, stream_);
			off += (reads[k].bytes + 255) & ~(size_t)255;
		}
		// wait even after a failed enqueue: the block must be idle before it is reused
		const hipError_t w = hipStreamSynchronize(stream_);
		if (e == hipSuccess && w == hipSuccess)
		{
			off = 0;
			for (int k = 0; k < count; k++)
			{
				if (reads[k].bytes)
					std::memcpy(reads[k].dst, (const char*)block + off, reads[k].bytes);
				off += (reads[k].bytes + 255) & ~(size_t)255;
			}
		}
		pinned().give(block, got);
		PCM_HIP(e);
		PCM_HIP(w);
		return PCM_OK;
	}

	int copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind)
	{
		if (bytes == 0)
			return PCM_OK;
		PCM_HIP(hipMemcpyAsync(dst, src, bytes, kind, stream_));
		PCM_HIP(hipStreamSynchronize(stream_));
		return PCM_OK;
	}

	int read_scalars()
	{
		PCM_HIP(hipMemcpyAsync(&sc_, h_.sc, sizeof(Scalars), hipMemcpyDeviceToHost, stream_));
		PCM_HIP(hipStreamSynchronize(stream_));
		return PCM_OK;
	}

	// neighbor overflow: grow K, rebuild the lists of the aborted step, clear the status
	int recover()
	{
		detach_ready_ = false;
		if (sc_.status & ST_BUCKET_OVERFLOW)
		{
			// the fused 3D mesh update stopped after the compaction kernel (its histogram and
			// voxel assignment are complete): finish the step's container update and mesh
			// update the dense way with larger tables, then its neighbor lists
			if (sc_.status & ~(ST_BUCKET_OVERFLOW | ST_NEIGHBOR_OVERFLOW))
				return fail(PCM_ERR_STATE, "device status word is set: " + std::to_string(sc_.status));
			PCM_TRY(grow_buckets(sc_.max_bucket));
			sc_.status = 0;
			sc_.max_bucket = 0;
			sc_.claim = 0; // the occupied voxels the fused update counted are claimed again
			PCM_TRY(copy_sync(h_.sc, &sc_, sizeof(Scalars), hipMemcpyHostToDevice));
			PCM_TRY(per_agent(PCM_K_COMPACT_RESTORE, k_compact_restore<R>));
			PCM_TRY(launch(PCM_K_COMPACT_DONE, k_compact_done<R>, 1, 64));
			switch (h_.dims)
			{
				case 1:
					PCM_TRY(mesh_kernels<1>());
					break;
				case 2:
					PCM_TRY(mesh_kernels<2>());
					break;
				default:
					PCM_TRY(mesh_kernels<3>());
			}
			PCM_TRY(launch_neighbors());
			PCM_TRY(read_scalars());
			if (!(sc_.status & ST_NEIGHBOR_OVERFLOW))
				return sc_.status ? fail(PCM_ERR_STATE, "mesh rebuild failed after bucket growth") : PCM_OK;
		}
		if (sc_.status & ~ST_NEIGHBOR_OVERFLOW)
			return fail(PCM_ERR_STATE, "device status word is set: " + std::to_string(sc_.status));
		PCM_TRY(grow_neighbors(sc_.max_neighbors));
		sc_.status = 0;
		sc_.max_neighbors = 0;
		PCM_TRY(copy_sync(h_.sc, &sc_, sizeof(Scalars), hipMemcpyHostToDevice));
		PCM_TRY(launch_neighbors());
		PCM_TRY(read_scalars());
		if (sc_.status)
			return fail(PCM_ERR_STATE, "neighbor rebuild failed after growth");
		return PCM_OK;
	}

	int gn() const { return grid_n_ ? grid_n_ : n_launch_; }
	int gl() const { return grid_L_ ? grid_L_ : h_.L; }
	int grid_agents() const { return std::max(1, (gn() + BLOCK - 1) / BLOCK); }

	// `kind`: the launch's block map in a batch's flattened launch (MAP_NONE: one block row
	// per simulation, as for single-block kernels)
	template <class Kern>
	int launch(int tag, Kern kernel, int blocks, int threads, int kind = MAP_NONE)
	{
		const bool timed = ((prof_mask_ >> tag) & 1ull) && prof_seen_[tag]++ % prof_every_ == 0;
		hipEvent_t ev[2] = { nullptr, nullptr };
		if (timed)
		{
			PCM_TRY(take_events(ev));
			PCM_HIP(hipEventRecord(ev[0], stream_));
		}
		const int4* map = nullptr;
		int gy = launch_y_;
		if (flat_maps_ && kind != MAP_NONE)
		{
			map = flat_maps_ + flat_off_[kind];
			blocks = flat_blocks_[kind];
			gy = 1;
		}
		hipLaunchKernelGGL(kernel, dim3(blocks, gy), dim3(threads), 0, stream_,
						   launch_ctx_ ? launch_ctx_ : (const SimCtx<R>*)d_, map);
		PCM_HIP(hipGetLastError());
		if (timed)
		{
			PCM_HIP(hipEventRecord(ev[1], stream_));
			prof_records_.push_back({ tag, ev[0], ev[1] });
		}
		return PCM_OK;
	}

	template <class Kern>
	int per_agent(int tag, Kern kernel)
	{
		return launch(tag, kernel, grid_agents(), BLOCK, MAP_AGENT);
	}

	// `lanes` threads per agent
	template <class Kern>
	int per_lanes(int tag, Kern kernel, int lanes)
	{
		return launch(tag, kernel, std::max(1, (int)(((long long)gn() * lanes + BLOCK - 1) / BLOCK)), BLOCK, map_of_lanes(lanes));
	}

	struct prof_record
	{
		int tag;
		hipEvent_t start, stop;
	};
	uint64_t prof_mask_ = 0;
	int prof_every_ = 1;                    // time one launch in this many of each kernel
	long long prof_seen_[PCM_KERNEL_COUNT] = {}; // launches of each kernel while profiling
	std::vector<prof_record> prof_records_;
	std::vector<hipEvent_t> event_pool_;

	int take_events(hipEvent_t* ev)
	{
		for (int k = 0; k < 2; k++)
		{
			if (!event_pool_.empty())
			{
				ev[k] = event_pool_.back();
				event_pool_.pop_back();
			}
			else
				PCM_HIP(hipEventCreate(&ev[k]));
		}
		return PCM_OK;
	}

public:
	int profile_begin(uint64_t mask) override
	{
		PCM_HIP(hipStreamSynchronize(stream_));
		for (auto& r : prof_records_)
		{
			event_pool_.push_back(r.start);
			event_pool_.push_back(r.stop);
		}
		prof_records_.clear();
		prof_mask_ = mask;
		for (auto& k : prof_seen_)
			k = 0;
		return PCM_OK;
	}

	int profile_every(int every) override
	{
		if (every < 1)
			return fail(PCM_ERR_ARGUMENT, "profile sampling stride must be >= 1");
		prof_every_ = every;
		return PCM_OK;
	}

	int profile_end(double* total_ms, int64_t* launches) override
	{
		PCM_HIP(hipStreamSynchronize(stream_));
		for (int k = 0; k < PCM_KERNEL_COUNT; k++)
		{
			if (total_ms)
				total_ms[k] = 0;
			if (launches)
				launches[k] = 0;
		}
		for (auto& r : prof_records_)
		{
			float ms = 0;
			PCM_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
			if (total_ms)
				total_ms[r.tag] += ms;
			if (launches)
				launches[r.tag]++;
			event_pool_.push_back(r.start);
			event_pool_.push_back(r.stop);
		}
		prof_records_.clear();
		prof_mask_ = 0;
		return PCM_OK;
	}

private:

	static constexpr int LANES_NEIGHBOR_MIN_AGENTS = 4096;

	// one block per occupied voxel (the kernel strides over more): sized from the occupied
	// voxels the host saw last, with headroom, so that few blocks start with no voxel
	int neighbor_blocks() const
	{
		int b = std::min(gn(), 16384);
		if (!grid_n_ && sc_.n_occ > 0)
			b = std::min(b, sc_.n_occ + sc_.n_occ / 4 + 64);
		return std::max(1, b);
	}

	template <int D>
	int neighbors_d()
	{
		// 1D and 2D stencils hold a few dozen candidates: above a few thousand agents one
		// lane per agent beats a block per occupied voxel (far fewer blocks); below, the
		// block per voxel has more parallelism to hide its loads
		if (mesh_fresh_ && D < 3 && !tables_ && gn() >= LANES_NEIGHBOR_MIN_AGENTS)
			return per_agent(PCM_K_NEIGHBORS, k_neighbors_lanes<R, D < 3 ? D : 2>);
		if (mesh_fresh_ && D == 3 && neighbors_sparse_)
			return launch(PCM_K_NEIGHBORS, k_neighbors<R, D, 1, NBR_TILE_SPARSE>, neighbor_blocks(), 64, MAP_NEIGHBORS);
		if (mesh_fresh_)
		{
			switch (tune_.neighbor_waves)
			{
				case 1:
					return launch(PCM_K_NEIGHBORS, k_neighbors<R, D, 1>, neighbor_blocks(), 64, MAP_NEIGHBORS);
				case 4:
					return launch(PCM_K_NEIGHBORS, k_neighbors<R, D, 4>, neighbor_blocks(), 256, MAP_NEIGHBORS);
				default:
					return launch(PCM_K_NEIGHBORS, k_neighbors<R, D, 2>, neighbor_blocks(), 128, MAP_NEIGHBORS);
			}
		}
		// the voxel buckets predate position changes: search each agent's stencil around
		// its current position, as the reference does with stale buckets
		return per_agent(PCM_K_NEIGHBORS, k_neighbors_direct<R, D>);
	}

	int launch_neighbors()
	{
		switch (h_.dims)
		{
			case 1:
				return neighbors_d<1>();
			case 2:
				return neighbors_d<2>();
			default:
				return neighbors_d<3>();
		}
	}

	// FUSED: the histogram ran in the step's compaction kernel (after the clear).  With voxel
	// tables (tables_: 3D, and 1D and 2D below the lane search) that kernel also filled the voxels'
	// unordered tables and gave each occupied voxel its bucket range, so the ordering kernel
	// follows at once; otherwise the claim and scatter kernels lay the buckets out densely
	// (the lane-per-agent neighbor search walks them).  The ordering kernel also decides the
	// next step's erased spring slots.
	template <int D, bool FUSED = false>
	int mesh_kernels()
	{
		mesh_fresh_ = true;
		records_fresh_ = true;
		binned_launch_ = n_launch_;
		if (!FUSED)
		{
			PCM_TRY(per_agent(PCM_K_BIN_CLEAR, k_bin_clear<R>));
			PCM_TRY(per_agent(PCM_K_BIN_COUNT, k_bin_count<R, D>));
		}
		const bool fixed = FUSED && tables_;
		if (!fixed)
		{
			PCM_TRY(launch(PCM_K_BIN_CLAIM, k_bin_claim<R, FUSED>, std::max(1, (gn() + CLAIM_THREADS - 1) / CLAIM_THREADS),
						   CLAIM_THREADS, MAP_CLAIM));
			PCM_TRY(per_agent(PCM_K_BIN_SCATTER, k_bin_scatter<R>));
		}
		if (!FUSED)
			return per_agent(PCM_K_BIN_ORDER, k_bin_order<R, D>);
		// the fused step's ordering blocks, then the next step's erase decisions
		const int erase_blocks = (int)(((long long)gn() * ERASE_LANES + BLOCK - 1) / BLOCK);
		if (fixed)
			return launch(PCM_K_BIN_ORDER, k_bin_order<R, D, true, true, true>, grid_agents() + erase_blocks, BLOCK,
						  MAP_ORDER);
		return launch(PCM_K_BIN_ORDER, k_bin_order<R, D, true, true, false>, grid_agents() + erase_blocks, BLOCK,
					  MAP_ORDER);
	}

	template <int D, bool CLEAR>
	int launch_velocity()
	{
		switch (tune_.velocity_lanes)
		{
			case 4:
				return per_lanes(PCM_K_VELOCITY_FUSED, k_velocity_fused<R, D, CLEAR, 4>, 4);
			case 16:
				return per_lanes(PCM_K_VELOCITY_FUSED, k_velocity_fused<R, D, CLEAR, 16>, 16);
			default:
				return per_lanes(PCM_K_VELOCITY_FUSED, k_velocity_fused<R, D, CLEAR, 8>, 8);
		}
	}

	template <int D, bool AB2>
	int launch_tentative()
	{
		switch (tune_.event_lanes)
		{
			case 4:
				return per_lanes(PCM_K_INTERACT_TENTATIVE, k_interact_tentative<R, D, AB2, 4>, 4);
			case 16:
				return per_lanes(PCM_K_INTERACT_TENTATIVE, k_interact_tentative<R, D, AB2, 16>, 16);
			default:
				return per_lanes(PCM_K_INTERACT_TENTATIVE, k_interact_tentative<R, D, AB2, 8>, 8);
		}
	}

	// IN_STEP: the kernels open a fused mechanics step, so they also advance the step
	// counter (the detachment draws use the new value; the erase kernel stores it), and the
	// step's velocity kernel appends the accepted attachments instead of k_spring_append
	// IN_STEP: the kernels open a fused mechanics step, so the step counter they key their
	// draws with is one above the stored one until the first resolve round stores it, and
	// the step's velocity kernel appends the accepted attachments instead of k_spring_append.
	// detach_ready: the previous fused step's mesh update decided the fired and erased slots
	template <bool IN_STEP>
	int spring_update_kernels(bool detach_ready = false)
	{
		if (!detach_ready)
		{
			PCM_TRY(per_agent(PCM_K_SPRING_FIRE, k_spring_fire<R, IN_STEP>));
			PCM_TRY(launch(PCM_K_SPRING_ERASE, k_spring_erase<R>,
						   std::max(1, (int)(((long long)gn() * gl() + BLOCK - 1) / BLOCK)), BLOCK, MAP_ERASE));
		}
		// the first resolve round runs one block per collect block (segments of BLOCK / G)
		const int segments = std::max(1, (int)(((long long)gn() * tune_.event_lanes + BLOCK - 1) / BLOCK));
		switch (tune_.event_lanes)
		{
			case 4:
				PCM_TRY(per_lanes(PCM_K_SPRING_COLLECT, k_spring_collect<R, 4, IN_STEP>, 4));
				PCM_TRY(launch(PCM_K_SPRING_RESOLVE_FIRST, k_spring_resolve_first<R, 4, IN_STEP>, segments, 64,
							   MAP_LANES4));
				break;
			case 16:
				PCM_TRY(per_lanes(PCM_K_SPRING_COLLECT, k_spring_collect<R, 16, IN_STEP>, 16));
				PCM_TRY(launch(PCM_K_SPRING_RESOLVE_FIRST, k_spring_resolve_first<R, 16, IN_STEP>, segments, 64,
							   MAP_LANES16));
				break;
			default:
				PCM_TRY(per_lanes(PCM_K_SPRING_COLLECT, k_spring_collect<R, 8, IN_STEP>, 8));
				PCM_TRY(launch(PCM_K_SPRING_RESOLVE_FIRST, k_spring_resolve_first<R, 8, IN_STEP>, segments, 64,
							   MAP_LANES8));
		}
		// a fused step runs the later rounds in its velocity kernel's first block, whose
		// append waits for them
		if (IN_STEP)
			return PCM_OK;
		PCM_TRY(launch(PCM_K_SPRING_RESOLVE, k_spring_resolve<R>, 1, RESOLVE_THREADS));
		return per_agent(PCM_K_SPRING_APPEND, k_spring_append<R>);
	}

	template <int D>
	int interaction_kernels()
	{
		PCM_TRY((launch_tentative<D, false>()));
		PCM_TRY(per_agent(PCM_K_INTERACT_APPLY, k_interact_apply<R, D, false>));
		return launch(PCM_K_INTERACT_SERIAL, k_interact_serial<R, D, false>, 1, RESOLVE_THREADS);
	}

	int compact_apply_kernel()
	{
		return launch(PCM_K_COMPACT_APPLY, k_compact_apply<R>, MOVE_BLOCKS + grid_agents(), BLOCK, MAP_COMPACT);
	}

	template <int D>
	int container_kernels()
	{
		PCM_TRY(per_agent(PCM_K_COMPACT_FLAG, k_compact_flag<R, D>));
		PCM_TRY(launch(PCM_K_COMPACT_PLAN, k_compact_plan<R>, 1, RESOLVE_THREADS));
		PCM_TRY(compact_apply_kernel());
		PCM_TRY(per_agent(PCM_K_COMPACT_RESTORE, k_compact_restore<R>));
		return launch(PCM_K_COMPACT_DONE, k_compact_done<R>, 1, 64);
	}

	template <int D>
	int phase_d(int which)
	{
		if (which != PH_MESH && which != PH_NEIGHBORS)
			records_fresh_ = false;
		switch (which)
		{
			case PH_MESH:
				return mesh_kernels<D>();
			case PH_NEIGHBORS:
			{
				PCM_TRY(neighbors_d<D>());
				PCM_TRY(read_scalars());
				if (sc_.status & ST_NEIGHBOR_OVERFLOW)
					PCM_TRY(recover());
				return PCM_OK;
			}
			case PH_CONTACT:
				if (!h_.contacts)
					return PCM_OK;
				return per_agent(PCM_K_CONTACT, k_contact<R, D>);
			case PH_FORCES:
				PCM_TRY(launch(PCM_K_BEGIN_STEP, k_begin_step<R>, 1, 64));
				return per_agent(PCM_K_FORCES, k_forces<R, D>);
			case PH_MOTILITY:
				return per_agent(PCM_K_MOTILITY, k_motility<R, D>);
			case PH_MEMBRANE:
				if (!h_.wall)
					return PCM_OK;
				return per_agent(PCM_K_MEMBRANE, k_membrane<R, D>);
			case PH_SPRINGS:
				if (!h_.springs)
					return PCM_OK;
				PCM_TRY(spring_update_kernels<false>());
				return per_agent(PCM_K_CONTRACT, k_contract<R, D>);
			case PH_POSITIONS:
				mesh_fresh_ = false;
				return per_agent(PCM_K_POSITIONS, k_positions<R, D>);
			case PH_INTERACTIONS:
				mesh_fresh_ = false; // fusion moves agents, ingestion changes radii
				return interaction_kernels<D>();
			case PH_CONTAINER:
				mesh_fresh_ = false;
				return container_kernels<D>();
		}
		return fail(PCM_ERR_ARGUMENT, "unknown phase");
	}

	// one mechanics step (simulator.cpp:173-192): the spring-list update runs first since
	// forces, motility and the membrane do not read springs; velocity contributions are
	// then summed in the reference's order by the fused velocity kernel.
	// detach_ready: the previous fused step's mesh update decided this step's detachments
	template <int D>
	int step_d(bool records_fresh, bool fold_clear, bool detach_ready)
	{
		if (h_.springs)
			PCM_TRY(spring_update_kernels<true>(detach_ready));
		else
			PCM_TRY(launch(PCM_K_BEGIN_STEP, k_begin_step<R>, 1, 64));
		if (!records_fresh)
			PCM_TRY(per_agent(PCM_K_FORCE_RECORDS, k_force_records<R, D>));
		// fold_clear: the previous mesh update's voxel counts are cleared by the velocity
		// kernel, whose grid covers every agent that update binned
		if (fold_clear)
			PCM_TRY((launch_velocity<D, true>()));
		else
			PCM_TRY((launch_velocity<D, false>()));
		PCM_TRY((launch_tentative<D, true>()));
		PCM_TRY(per_agent(PCM_K_INTERACT_APPLY, k_interact_apply_plan<R, D>));
		if (!fold_clear)
			PCM_TRY(per_agent(PCM_K_BIN_CLEAR, k_bin_clear<R>));
		// the container update's moves and the mesh update's histogram
		if (tables_)
			PCM_TRY(launch(PCM_K_COMPACT_APPLY, k_compact_count<R, D, true>, MOVE_BLOCKS + grid_agents(), BLOCK,
						   MAP_COMPACT));
		else
			PCM_TRY(launch(PCM_K_COMPACT_APPLY, k_compact_count<R, D, false>, MOVE_BLOCKS + grid_agents(), BLOCK,
						   MAP_COMPACT));
		PCM_TRY((mesh_kernels<D, true>()));
		return neighbors_d<D>();
	}

	int dispatch_step()
	{
		const bool ready = detach_ready_;
		detach_ready_ = false;
		PCM_TRY(dispatch_step_flags(records_fresh_, binned_launch_ <= n_launch_, ready));
		detach_ready_ = true;
		return PCM_OK;
	}

	int dispatch_step_flags(bool records_fresh, bool fold_clear, bool detach_ready)
	{
		switch (h_.dims)
		{
			case 1:
				return step_d<1>(records_fresh, fold_clear, detach_ready);
			case 2:
				return step_d<2>(records_fresh, fold_clear, detach_ready);
			default:
				return step_d<3>(records_fresh, fold_clear, detach_ready);
		}
	}

public:
	// ---------------------------------------------------------------- batch membership
	hipStream_t stream() const { return stream_; }
	void set_stream(hipStream_t s) { stream_ = s; }
	const SimCtx<R>& host_ctx() const { return h_; }
	uint64_t ctx_version() const { return ctx_version_; }
	int dims() const { return h_.dims; }
	int launch_bound() const { return n_launch_; }
	void set_neighbors_sparse(bool on) { neighbors_sparse_ = on; }
	void set_tables(bool on) { tables_ = on; }
	bool has_tables() const { return h_.B > 0; }
	int spring_slots() const { return h_.L; }
	// neighbor-search blocks of this member in a batch's flattened launch: its occupied
	// voxels at the last step seen, with headroom, rounded to four significant bits so that
	// the block maps (and the captured graph) change only when the count drifts by about
	// 1/8; the kernel strides over voxels beyond its blocks, so any count is correct
	int neighbor_map_blocks() const
	{
		int b = std::min(n_launch_, 16384);
		if (sc_.n_occ > 0)
			b = std::min(b, round_up_sig4(sc_.n_occ + sc_.n_occ / 4 + 64));
		return std::max(1, b);
	}
	bool records_fresh() const { return records_fresh_; }
	bool detach_ready() const { return detach_ready_; }
	bool binned_within(int grid_n) const { return binned_launch_ <= grid_n; }
	void*& batch() { return batch_; }
	int apply_params(const pcm_step_params& p) { return set_params(p); }
	void set_scalars(const Scalars& sc) { sc_ = sc; }
	const Scalars& scalars() const { return sc_; }

	pcm_tuning tuning() const { return tune_; }

	static bool valid_tuning(const pcm_tuning& t)
	{
		auto one_of = [](int v, int a, int b, int c) { return v == a || v == b || v == c; };
		return one_of(t.velocity_lanes, 4, 8, 16) && one_of(t.event_lanes, 4, 8, 16) && one_of(t.neighbor_waves, 1, 2, 4);
	}

	int set_tuning(const pcm_tuning* t) override
	{
		if (!t || !valid_tuning(*t))
			return fail(PCM_ERR_ARGUMENT, "tuning values must be 4, 8 or 16 lanes and 1, 2 or 4 wavefronts");
		tune_ = *t;
		tune_chosen_ = true;
		return PCM_OK;
	}

	int get_tuning(pcm_tuning* t) override
	{
		if (!t)
			return fail(PCM_ERR_ARGUMENT, "null tuning");
		*t = tune_chosen_ ? tune_ : default_tuning(sc_.n);
		return PCM_OK;
	}

	// The launch shapes give bit-identical results, so the tuner times real steps: each
	// candidate shape runs `steps` steps with its kernels bracketed by events, in two
	// rounds of opposite order (the better round counts), and each kernel family keeps its
	// fastest shape.  The simulation advances 1 + 6 * steps steps.  `run` advances this
	// simulation (or the batch it leads) by some steps.
	template <class Run>
	int autotune_with(int steps, Run&& run, pcm_tuning* chosen)
	{
		static const pcm_tuning candidates[3] = { { 4, 4, 1 }, { 8, 8, 2 }, { 16, 16, 4 } };
		const uint64_t mask = (1ull << PCM_K_VELOCITY_FUSED) | (1ull << PCM_K_SPRING_COLLECT)
							  | (1ull << PCM_K_SPRING_RESOLVE) | (1ull << PCM_K_SPRING_RESOLVE_FIRST)
							  | (1ull << PCM_K_INTERACT_TENTATIVE)
							  | (1ull << PCM_K_NEIGHBORS);
		tune_chosen_ = true; // the candidates' steps keep the shape they are given
		PCM_TRY(run(1));     // warm-up with the current shape
		double vel[3] = { 1e300, 1e300, 1e300 }, ev[3] = { 1e300, 1e300, 1e300 }, nb[3] = { 1e300, 1e300, 1e300 };
		std::vector<double> ms(PCM_KERNEL_COUNT);
		std::vector<int64_t> cnt(PCM_KERNEL_COUNT);
		static const int order[6] = { 0, 1, 2, 2, 1, 0 };
		for (int r = 0; r < 6; r++)
		{
			const int k = order[r];
			tune_ = candidates[k];
			PCM_TRY(profile_begin(mask));
			PCM_TRY(run(steps));
			PCM_TRY(profile_end(ms.data(), cnt.data()));
			auto avg = [&](int tag) { return cnt[tag] ? ms[tag] / (double)cnt[tag] : 0.0; };
			vel[k] = std::min(vel[k], avg(PCM_K_VELOCITY_FUSED));
			ev[k] = std::min(ev[k], avg(PCM_K_SPRING_COLLECT) + avg(PCM_K_SPRING_RESOLVE_FIRST)
										+ avg(PCM_K_SPRING_RESOLVE) + avg(PCM_K_INTERACT_TENTATIVE));
			nb[k] = std::min(nb[k], avg(PCM_K_NEIGHBORS));
		}
		auto best = [](const double* t) {
			int b = 1; // the default shape wins ties
			for (int k = 0; k < 3; k++)
				if (t[k] < t[b])
					b = k;
			return b;
		};
		tune_.velocity_lanes = candidates[best(vel)].velocity_lanes;
		tune_.event_lanes = candidates[best(ev)].event_lanes;
		tune_.neighbor_waves = candidates[best(nb)].neighbor_waves;
		if (chosen)
			*chosen = tune_;
		return PCM_OK;
	}

	int autotune(const pcm_step_params* p, int steps, pcm_tuning* chosen) override
	{
		if (!p || steps < 1)
			return fail(PCM_ERR_ARGUMENT, "null params or no steps per candidate");
		return autotune_with(steps, [&](int k) { return step(p, k); }, chosen);
	}

	// `steps` fused steps of every simulation in `ctxs` (this one is the batch's first
	// member), as one launch per kernel with grid.y = m
	// maps: the batch's block maps (device, offsets and padded totals per map kind) for
	// flattened launches, or nullptr for one block row per member
	int run_segmented(const SimCtx<R>* ctxs, int m, int grid_n, int grid_L, bool records_fresh, bool fold_clear,
					  bool detach_ready, int steps, const int4* maps = nullptr, const int* offs = nullptr,
					  const int* totals = nullptr)
	{
		launch_ctx_ = ctxs;
		launch_y_ = m;
		grid_n_ = grid_n;
		grid_L_ = grid_L;
		flat_maps_ = maps;
		for (int k = 0; k < MAP_KINDS; k++)
		{
			flat_off_[k] = maps ? offs[k] : 0;
			flat_blocks_[k] = maps ? totals[k] : 0;
		}
		int rc = PCM_OK;
		for (int k = 0; k < steps && rc == PCM_OK; k++)
		{
			rc = dispatch_step_flags(records_fresh, fold_clear, detach_ready);
			records_fresh = true;
			fold_clear = true;
			detach_ready = true;
		}
		launch_ctx_ = nullptr;
		launch_y_ = 1;
		grid_n_ = 0;
		grid_L_ = 0;
		flat_maps_ = nullptr;
		return rc;
	}

	// after segmented steps: the mesh, records and voxel counts are this step's
	void mark_stepped()
	{
		mesh_fresh_ = true;
		records_fresh_ = true;
		detach_ready_ = true;
		binned_launch_ = n_launch_;
	}

	// a member whose segmented steps stopped on neighbor overflow: grow, rebuild, then
	// finish its remaining steps alone
	int finish_steps(unsigned long long target)
	{
		PCM_TRY(read_scalars());
		for (int guard = 0; guard < 64 && sc_.step < target; guard++)
		{
			if (sc_.status)
				PCM_TRY(recover());
			const long long remaining = (long long)(target - sc_.step);
			for (long long s = 0; s < remaining; s++)
				PCM_TRY(dispatch_step());
			PCM_TRY(read_scalars());
		}
		if (sc_.status)
			PCM_TRY(recover());
		if (sc_.step < target)
			return fail(PCM_ERR_STATE, "mechanics step did not converge after repeated neighbor-capacity growth");
		return PCM_OK;
	}
};

// ---------------------------------------------------------------------------- batches
// Simulations of one real type advance together.  Members with the same dimensionality
// form a group whose kernels are launched once for all of them: the group's context
// array has one slot per member (grid.y = slot, each block reads its slot's context),
// and free slots hold an empty placeholder simulation whose blocks exit at once.  The
// 1D, 2D and 3D groups run on their own streams, so their launches overlap.  With graphs
// on, one step of a group is captured once (a hipGraph of the step's launches over all
// slots) and replayed each step; joining or leaving only rewrites context slots, and a
// new capture is needed only when the slot count, the grid bound or the step parameters
// change.  Members keep their device state and handles; while admitted, their own calls
// run on the group's stream.
template <class R>
class batch_impl final : public pcm_batch
{
public:
	~batch_impl() override
	{
		for (auto& g : groups_)
		{
			if (g.stream)
				(void)hipStreamSynchronize(g.stream);
			for (auto& mb : g.slots)
				if (mb.sim)
				{
					mb.sim->set_stream(mb.own);
					mb.sim->batch() = nullptr;
				}
			if (g.exec)
				(void)hipGraphExecDestroy(g.exec);
			if (g.placeholder)
				g.placeholder->set_stream(g.placeholder_own);
			g.placeholder.reset();
			if (g.stream)
				(void)hipStreamDestroy(g.stream);
			if (g.d_ctx)
				(void)hipFree(g.d_ctx);
			if (g.d_sc)
				(void)hipFree(g.d_sc);
			if (g.d_map)
				(void)hipFree(g.d_map);
		}
	}

	int admit(pcm_sim* base) override
	{
		auto* sim = dynamic_cast<sim_impl<R>*>(base);
		if (!sim)
			return fail(PCM_ERR_ARGUMENT, "simulation of another real type than the batch");
		if (sim->batch())
			return fail(PCM_ERR_STATE, "simulation already belongs to a batch");
		group& g = groups_[sim->dims() - 1];
		PCM_TRY(ensure_group(g, sim->dims()));
		PCM_HIP(hipStreamSynchronize(sim->stream()));
		PCM_HIP(hipStreamSynchronize(g.stream));
		member m;
		m.sim = sim;
		m.own = sim->stream();
		sim->set_stream(g.stream);
		sim->batch() = static_cast<pcm_batch*>(this);
		for (auto& slot : g.slots)
			if (!slot.sim)
			{
				slot = m;
				return PCM_OK;
			}
		g.slots.push_back(m);
		return PCM_OK;
	}

	int retire(pcm_sim* base) override
	{
		for (auto& g : groups_)
			for (auto& slot : g.slots)
				if (slot.sim == base)
				{
					PCM_HIP(hipStreamSynchronize(g.stream));
					slot.sim->set_stream(slot.own);
					slot.sim->batch() = nullptr;
					slot = member {};
					return PCM_OK;
				}
		return fail(PCM_ERR_ARGUMENT, "simulation is not a member of this batch");
	}

	int size(int32_t* sims) override
	{
		int32_t total = 0;
		for (auto& g : groups_)
			for (auto& slot : g.slots)
				total += slot.sim != nullptr;
		*sims = total;
		return PCM_OK;
	}

	// p == nullptr: every member steps under its own parameters (its last pcm_step,
	// phase call or pcm_set_step_params); each kernel reads them from its slot's context
	int step(const pcm_step_params* p, int steps) override
	{
		if (steps < 0)
			return fail(PCM_ERR_ARGUMENT, "negative step count");
		// launch every group, then collect: the groups' streams run concurrently
		for (auto& g : groups_)
			PCM_TRY(launch_group(g, p, steps));
		for (auto& g : groups_)
			PCM_TRY(collect_group(g, steps));
		return PCM_OK;
	}

	int synchronize() override
	{
		for (auto& g : groups_)
			if (g.stream)
				PCM_HIP(hipStreamSynchronize(g.stream));
		return PCM_OK;
	}

	void use_graphs(bool on) { graphs_ = on; }
	void use_flattened_grids(bool on) { flatten_ = on; }

	// agent counts of members as the batch's last step left them: the scalars each step call
	// gathers for all members at its end (no device access)
	int counts(int n, pcm_sim* const* sims, int32_t* out) override
	{
		for (int k = 0; k < n; k++)
		{
			auto* sim = dynamic_cast<sim_impl<R>*>(sims[k]);
			if (!sim || sim->batch() != static_cast<pcm_batch*>(this))
				return fail(PCM_ERR_ARGUMENT, "entry " + std::to_string(k) + " is not a member of this batch");
			out[k] = sim->scalars().n;
		}
		return PCM_OK;
	}

	// the launch shape of a dimensionality group (kept for members that join later)
	int set_tuning(int dims, const pcm_tuning* t) override
	{
		if (dims < 1 || dims > 3)
			return fail(PCM_ERR_ARGUMENT, "dims must be 1, 2 or 3");
		group& g = groups_[dims - 1];
		PCM_TRY(ensure_group(g, dims));
		return g.placeholder->set_tuning(t);
	}

	int get_tuning(int dims, pcm_tuning* t) override
	{
		if (dims < 1 || dims > 3)
			return fail(PCM_ERR_ARGUMENT, "dims must be 1, 2 or 3");
		group& g = groups_[dims - 1];
		PCM_TRY(ensure_group(g, dims));
		return g.placeholder->get_tuning(t);
	}

	// tune each group's launch shape on its real steps (graphs are off while tuning)
	int autotune(const pcm_step_params* p, int steps) override
	{
		if (steps < 1)
			return fail(PCM_ERR_ARGUMENT, "no steps per candidate");
		const bool graphs = graphs_;
		graphs_ = false;
		int rc = PCM_OK;
		for (auto& g : groups_)
		{
			bool any = false;
			for (auto& slot : g.slots)
				any = any || slot.sim != nullptr;
			if (!any || rc != PCM_OK)
				continue;
			rc = g.placeholder->autotune_with(
				steps,
				[&](int k) {
					PCM_TRY(launch_group(g, p, k));
					return collect_group(g, k);
				},
				nullptr);
		}
		graphs_ = graphs;
		return rc;
	}

private:
	struct member
	{
		sim_impl<R>* sim = nullptr;
		hipStream_t own = nullptr;
		unsigned long long start = 0;
	};
	struct capture_key
	{
		int slots = 0, grid_n = 0, grid_L = 0, springs = -1;
		unsigned long long maps = 0; // map version of a flattened capture (0: none)
		int velocity_lanes = 0, event_lanes = 0, neighbor_waves = 0;
		bool sparse = false; // the neighbor search's sparse-population kernel
		bool tables = false; // the mesh update fills voxel tables
		bool operator==(const capture_key&) const = default;
	};
	struct group
	{
		std::vector<member> slots;
		std::unique_ptr<sim_impl<R>> placeholder; // the empty simulation of free slots; launches the group
		hipStream_t placeholder_own = nullptr;    // its own stream, restored before it is destroyed
		hipStream_t stream = nullptr;
		SimCtx<R>* d_ctx = nullptr;
		Scalars* d_sc = nullptr;
		int cap = 0;
		int xcds = 0; // XCDs of the device (flattened maps deal blocks over them)
		hipGraphExec_t exec = nullptr;
		capture_key key;
		std::vector<SimCtx<R>> h_ctx;
		std::vector<Scalars> h_sc;
		// block maps of flattened launches (map_kind order, concatenated), the layout they
		// describe (per slot: launch bound and spring slots) and a version for graph keys
		int4* d_map = nullptr;
		size_t map_cap = 0;
		std::vector<int4> h_map;
		std::vector<long long> map_sig;
		int map_off[MAP_KINDS] = {};
		int map_tot[MAP_KINDS] = {};
		unsigned long long map_version = 0;
		bool active = false; // launched by the current step call
	};
	group groups_[3];
	bool graphs_ = false;
	bool flatten_ = true;

	// The block maps of the group's members (slots 0..m-1), rebuilt and uploaded only when
	// a slot's simulation, launch bound or spring-slot count changed.  Each map's grid is
	// padded to four significant bits, so small changes keep the graph's grid sizes.
	int build_maps(group& g, int m)
	{
		std::vector<long long> sig;
		sig.reserve(2 * (size_t)m);
		for (int k = 0; k < m; k++)
		{
			auto* sim = g.slots[k].sim;
			sig.push_back(sim ? sim->launch_bound() : -1);
			sig.push_back(sim ? sim->spring_slots() : 0);
			sig.push_back(sim ? sim->neighbor_map_blocks() : 0);
		}
		if (g.d_map && sig == g.map_sig)
			return PCM_OK;
		// Workgroups go to the device's XCDs in turn (grid block b to XCD b mod X), and each XCD
		// has its own L2.  The members' blocks, in member order, are cut into X equal shares,
		// one per XCD, so that a member's gathers stay within the one or few XCDs its share
		// covers instead of all of them.  A member spanning several XCDs deals its blocks among
		// them in proportion to their parts, so that blocks of unequal cost (a member's ordering
		// blocks before its erase blocks, its voxel blocks before its spare ones) spread evenly.
		// Each member's block 0, which its other blocks may wait for in the velocity kernel,
		// comes first in its XCD's list, ahead of every block that could wait.  Grid block
		// i * X + x holds entry i of XCD x's list.
		if (g.xcds < 1)
		{
			int dev = 0, x = 1;
			PCM_HIP(hipGetDevice(&dev));
			if (hipDeviceGetAttribute(&x, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess || x < 1)
				x = 1;
			g.xcds = std::min(x, 64);
		}
		const int X = g.xcds;
		g.h_map.clear();
		std::vector<std::vector<int4>> heads(X), rest(X);
		std::vector<long long> part(X), dealt(X);
		for (int kind = 0; kind < MAP_KINDS; kind++)
		{
			g.map_off[kind] = (int)g.h_map.size();
			long long total = 0;
			for (int k = 0; k < m; k++)
				if (auto* sim = g.slots[k].sim)
					total += member_blocks(kind, sim->launch_bound(), sim->spring_slots(), sim->neighbor_map_blocks());
			if (total > 0x7fffffffLL / 4)
				return fail(PCM_ERR_CAPACITY, "flattened batch grid too large");
			if (kind == MAP_NEIGHBORS)
			{
				// the neighbor search keeps one member's blocks dealt over all XCDs in turn
				// (its voxel blocks cost far more than its spare ones, and per-XCD lists of
				// unequal cost at equal positions held back the in-order dispatch)
				for (int k = 0; k < m; k++)
				{
					auto* sim = g.slots[k].sim;
					if (!sim)
						continue;
					const int b = member_blocks(kind, sim->launch_bound(), sim->spring_slots(), sim->neighbor_map_blocks());
					for (int j = 0; j < b; j++)
						g.h_map.push_back(make_int4(k, j, b, 0));
				}
				const int padded = round_up_grid((int)std::max<long long>(total, 1));
				for (long long x = total; x < padded; x++)
					g.h_map.push_back(make_int4(-1, 0, 0, 0));
				g.map_tot[kind] = padded;
				continue;
			}
			const long long share = std::max<long long>(1, (total + X - 1) / X);
			for (int x = 0; x < X; x++)
			{
				heads[x].clear();
				rest[x].clear();
				rest[x].reserve((size_t)share);
			}
			long long pos = 0;
			for (int k = 0; k < m; k++)
			{
				auto* sim = g.slots[k].sim;
				if (!sim)
					continue;
				const int b = member_blocks(kind, sim->launch_bound(), sim->spring_slots(), sim->neighbor_map_blocks());
				const long long start = pos;
				const int x0 = (int)std::min<long long>(start / share, X - 1);
				const int x1 = (int)std::min<long long>((start + b - 1) / share, X - 1);
				pos += b;
				if (x0 == x1)
				{
					heads[x0].push_back(make_int4(k, 0, b, 0));
					for (int j = 1; j < b; j++)
						rest[x0].push_back(make_int4(k, j, b, 0));
					continue;
				}
				for (int x = x0; x <= x1; x++)
				{
					const long long lo = std::max<long long>(start, (long long)x * share);
					const long long hi = x == X - 1 ? start + b : std::min<long long>(start + b, (long long)(x + 1) * share);
					part[x] = std::max<long long>(hi - lo, 0);
					dealt[x] = 0;
				}
				for (int j = 0; j < b; j++)
				{
					// the XCD furthest behind its part (block j of b due part * (j + 1) / b)
					int best = x0;
					long long best_lag = LLONG_MIN;
					for (int x = x0; x <= x1; x++)
					{
						if (dealt[x] >= part[x])
							continue;
						const long long lag = part[x] * (j + 1) - dealt[x] * (long long)b;
						if (lag > best_lag)
						{
							best_lag = lag;
							best = x;
						}
					}
					dealt[best]++;
					(j == 0 ? heads[best] : rest[best]).push_back(make_int4(k, j, b, 0));
				}
			}
			size_t longest = 1;
			for (int x = 0; x < X; x++)
				longest = std::max(longest, heads[x].size() + rest[x].size());
			const int per = round_up_grid((int)longest);
			for (int i = 0; i < per; i++)
				for (int x = 0; x < X; x++)
				{
					const size_t h = heads[x].size();
					if ((size_t)i < h)
						g.h_map.push_back(heads[x][i]);
					else if ((size_t)i - h < rest[x].size())
						g.h_map.push_back(rest[x][i - h]);
					else
						g.h_map.push_back(make_int4(-1, 0, 0, 0));
				}
			g.map_tot[kind] = per * X;
		}
		// the previous maps may still be read by queued launches
		PCM_HIP(hipStreamSynchronize(g.stream));
		if (g.h_map.size() > g.map_cap)
		{
			if (g.d_map)
				PCM_HIP(hipFree(g.d_map));
			g.d_map = nullptr;
			const size_t cap = std::max(g.h_map.size(), 2 * g.map_cap);
			PCM_HIP(hipMalloc((void**)&g.d_map, sizeof(int4) * cap));
			g.map_cap = cap;
		}
		PCM_HIP(hipMemcpyAsync(g.d_map, g.h_map.data(), sizeof(int4) * g.h_map.size(), hipMemcpyHostToDevice, g.stream));
		g.map_sig = std::move(sig);
		g.map_version++;
		return PCM_OK;
	}

	int ensure_group(group& g, int dims)
	{
		if (g.placeholder)
			return PCM_OK;
		PCM_HIP(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
		pcm_config c {};
		c.mesh.dims = dims;
		for (int k = 0; k < 3; k++)
		{
			c.mesh.bbox_min[k] = 0;
			c.mesh.bbox_max[k] = 1;
			c.mesh.voxel[k] = 1;
		}
		c.real_bytes = (int)sizeof(R);
		c.cell_defs = 1;
		c.substrates = 0;
		c.capacity = 1;
		auto ph = std::make_unique<sim_impl<R>>(c);
		PCM_TRY(ph->init());
		PCM_TRY(ph->set_count(0));
		PCM_HIP(hipStreamSynchronize(ph->stream()));
		g.placeholder_own = ph->stream();
		ph->set_stream(g.stream);
		g.placeholder = std::move(ph);
		return PCM_OK;
	}

	int reserve(group& g, int m)
	{
		if (m <= g.cap)
			return PCM_OK;
		const int cap = std::max(m, 2 * g.cap);
		PCM_HIP(hipStreamSynchronize(g.stream));
		if (g.d_ctx)
			PCM_HIP(hipFree(g.d_ctx));
		if (g.d_sc)
			PCM_HIP(hipFree(g.d_sc));
		g.d_ctx = nullptr;
		g.d_sc = nullptr;
		PCM_HIP(hipMalloc((void**)&g.d_ctx, sizeof(SimCtx<R>) * cap));
		PCM_HIP(hipMalloc((void**)&g.d_sc, sizeof(Scalars) * cap));
		g.cap = cap;
		if (g.exec)
		{
			PCM_HIP(hipGraphExecDestroy(g.exec));
			g.exec = nullptr;
		}
		return PCM_OK;
	}

	// every slot's scalars in one copy
	int gather(group& g, int m)
	{
		hipLaunchKernelGGL(k_gather_scalars<R>, dim3(1, m), dim3(64), 0, g.stream, (const SimCtx<R>*)g.d_ctx, g.d_sc);
		PCM_HIP(hipGetLastError());
		g.h_sc.resize(m);
		PCM_HIP(hipMemcpyAsync(g.h_sc.data(), g.d_sc, sizeof(Scalars) * m, hipMemcpyDeviceToHost, g.stream));
		PCM_HIP(hipStreamSynchronize(g.stream));
		return PCM_OK;
	}

	// a group's grid bound: its largest member rounded up to four significant bits, at most
	// 1/8 above it (a power of two wasted up to half of the blocks); eight bounds per octave
	// keep graph recaptures rare as members join and leave
	static int round_up_grid(int x) { return round_up_sig4(x); }

	int launch_group(group& g, const pcm_step_params* p, int steps)
	{
		g.active = false;
		int members = 0;
		for (auto& slot : g.slots)
			members += slot.sim != nullptr;
		if (members == 0 || steps == 0)
			return PCM_OK;
		g.active = true;
		const int m = (int)g.slots.size();
		PCM_TRY(reserve(g, m));
		int grid_n = 1, grid_L = 1;
		bool records_fresh = true, detach_ready = true;
		// the leader launches the spring kernels when any member has springs on (members with
		// them off skip the work inside those kernels)
		int any_springs = 0;
		for (auto& slot : g.slots)
			if (slot.sim)
			{
				PCM_TRY(slot.sim->ready());
				if (p)
					PCM_TRY(slot.sim->apply_params(*p));
				any_springs |= slot.sim->host_ctx().springs;
			}
		pcm_step_params lp = p ? *p : pcm_step_params { 0.1, 0, 0 };
		lp.automated_spring_adhesion = any_springs;
		PCM_TRY(g.placeholder->apply_params(lp));
		for (auto& slot : g.slots)
			if (slot.sim)
			{
				grid_n = std::max(grid_n, slot.sim->launch_bound());
				grid_L = std::max(grid_L, slot.sim->spring_slots());
				records_fresh = records_fresh && slot.sim->records_fresh();
				detach_ready = detach_ready && slot.sim->detach_ready();
			}
		grid_n = round_up_grid(grid_n);
		// the group's neighbor search kernel, from its members' agents per occupied voxel
		long long agents = 0, occupied = 0;
		for (auto& slot : g.slots)
			if (slot.sim)
			{
				agents += slot.sim->scalars().n;
				occupied += slot.sim->scalars().n_occ;
			}
		const bool sparse = sparse_population(agents, occupied);
		g.placeholder->set_neighbors_sparse(sparse);
		// the group's mesh update fills voxel tables only when every member has them
		bool tables = true;
		for (auto& slot : g.slots)
			if (slot.sim)
				tables = tables && slot.sim->has_tables();
		g.placeholder->set_tables(tables);
		bool fold_clear = true;
		for (auto& slot : g.slots)
			if (slot.sim)
				fold_clear = fold_clear && slot.sim->binned_within(flatten_ ? slot.sim->launch_bound() : grid_n);
		g.h_ctx.resize(m);
		for (int k = 0; k < m; k++)
			g.h_ctx[k] = g.slots[k].sim ? g.slots[k].sim->host_ctx() : g.placeholder->host_ctx();
		PCM_HIP(hipMemcpyAsync(g.d_ctx, g.h_ctx.data(), sizeof(SimCtx<R>) * m, hipMemcpyHostToDevice, g.stream));
		PCM_TRY(gather(g, m));
		for (int k = 0; k < m; k++)
			if (g.slots[k].sim)
			{
				if (g.h_sc[k].status)
					return fail(PCM_ERR_STATE, "a batch member's device status word is set");
				g.slots[k].start = g.h_sc[k].step;
			}
		auto* leader = g.placeholder.get();
		const int4* maps = nullptr;
		if (flatten_)
		{
			PCM_TRY(build_maps(g, m));
			maps = g.d_map;
		}
		auto run = [&](bool fresh, bool clear, bool detach, int k) {
			return leader->run_segmented(g.d_ctx, m, grid_n, grid_L, fresh, clear, detach, k, maps, g.map_off,
										 g.map_tot);
		};
		int done = 0;
		if (!records_fresh || !fold_clear || !detach_ready)
		{
			// the first step refreshes records, clears counts or decides the detachments for
			// some member: no graph
			PCM_TRY(run(records_fresh, fold_clear, detach_ready, 1));
			done = 1;
		}
		if (!graphs_)
			return run(true, true, true, steps - done);
		const pcm_tuning t = leader->tuning();
		const capture_key key { m, grid_n, grid_L, leader->host_ctx().springs, maps ? g.map_version : 0,
								t.velocity_lanes, t.event_lanes, t.neighbor_waves, sparse, tables };
		if (!g.exec || !(key == g.key))
		{
			if (g.exec)
			{
				PCM_HIP(hipGraphExecDestroy(g.exec));
				g.exec = nullptr;
			}
			// while other threads are inside the library, this call's steps run as plain
			// launches instead of waiting for them; a later call captures
			std::unique_lock<std::shared_mutex> no_other_calls(g_capture_mu, std::try_to_lock);
			if (!no_other_calls.owns_lock())
				return run(true, true, true, steps - done);
			hipGraph_t graph = nullptr;
			PCM_HIP(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
			const int rc = run(true, true, true, 1);
			const hipError_t end = hipStreamEndCapture(g.stream, &graph);
			no_other_calls.unlock();
			if (end != hipSuccess)
			{
				// a capture the runtime invalidated (its launches then fail too) ran nothing:
				// this call's steps run as plain launches, the next call captures again
				(void)hipGetLastError();
				if (graph)
					(void)hipGraphDestroy(graph);
				return run(true, true, true, steps - done);
			}
			PCM_TRY(rc);
			const hipError_t inst = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
			PCM_HIP(hipGraphDestroy(graph));
			PCM_HIP(inst);
			g.key = key;
		}
		for (int k = done; k < steps; k++)
			PCM_HIP(hipGraphLaunch(g.exec, g.stream));
		return PCM_OK;
	}

	int collect_group(group& g, int steps)
	{
		if (!g.active)
			return PCM_OK;
		const int m = (int)g.slots.size();
		PCM_TRY(gather(g, m));
		for (int k = 0; k < m; k++)
		{
			auto* sim = g.slots[k].sim;
			if (!sim)
				continue;
			sim->mark_stepped();
			sim->set_scalars(g.h_sc[k]);
			const unsigned long long target = g.slots[k].start + (unsigned long long)steps;
			if (g.h_sc[k].status || g.h_sc[k].step != target)
				PCM_TRY(sim->finish_steps(target));
		}
		return PCM_OK;
	}
};

pcm_sim* checked(pcm_sim* s) { return s; }

// the mechanics grid of a mesh description, as pcm_create sizes it (ceil of extent / voxel)
int grid_of(const pcm_mesh_desc& m, int grid[3])
{
	long long cells = 1;
	for (int k = 0; k < 3; k++)
	{
		if (k < m.dims)
		{
			if (m.voxel[k] <= 0 || m.bbox_max[k] < m.bbox_min[k])
				return fail(PCM_ERR_ARGUMENT, "bad mesh box or voxel shape");
			grid[k] = std::max(1, (m.bbox_max[k] - m.bbox_min[k] + m.voxel[k] - 1) / m.voxel[k]);
		}
		else
			grid[k] = 1;
		cells *= grid[k];
	}
	if (cells > 0x7fffffffLL)
		return fail(PCM_ERR_CAPACITY, "mechanics voxel grid larger than 2^31 voxels");
	return PCM_OK;
}

// voxel_index of pcm_device.h for one coordinate, on the host: the same real_t arithmetic
template <class R>
int host_voxel_coord(R p, const pcm_mesh_desc& m, const int grid[3], int k)
{
	const R q = (p - (R)m.bbox_min[k]) / (R)m.voxel[k];
	const int v = (int)q;
	return v < 0 ? 0 : (v >= grid[k] ? grid[k] - 1 : v);
}

// ---------------------------------------------------------------------------- allocation-free neighbor build
// The mesh scratch of pcm_neighbors_noalloc: one fixed area per GPU, held until the process
// ends.  In host code `scratch` names a host-side stand-in, not device memory: the device
// address comes from hipGetSymbolAddress, taken once per device with that device current,
// and only that address is ever handed to kernels.
static __device__ __align__(256) unsigned char scratch[PCM_NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

int scratch_of_current_device(unsigned char** out)
{
	int dev = -1;
	PCM_HIP(hipGetDevice(&dev));
	if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
		return fail(PCM_ERR_UNSUPPORTED, "device ordinal beyond the scratch table");
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	if (!g_scratch_ptr[dev])
	{
		void* p = nullptr;
		PCM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
		if (!p)
			return fail(PCM_ERR_DEVICE, "hipGetSymbolAddress returned no address for the scratch area");
		g_scratch_ptr[dev] = p;
	}
	*out = (unsigned char*)g_scratch_ptr[dev];
	return PCM_OK;
}

// bytes of the scratch area a call needs (the layout of noalloc_build, each part aligned)
constexpr size_t noalloc_align(size_t b) { return (b + 255) & ~(size_t)255; }
size_t noalloc_bytes(long long voxels, long long count, int real_bytes)
{
	const size_t n = (size_t)std::max<long long>(count, 1);
	return noalloc_align(sizeof(SimCtx<double>)) + noalloc_align(sizeof(Scalars)) + 2 * noalloc_align(4 * (size_t)voxels) +
		   5 * noalloc_align(4 * n) + noalloc_align(4 * (size_t)real_bytes * n);
}

template <class R, int D>
int noalloc_build(const SimCtx<R>& proto, unsigned char* area, int* needed, hipStream_t stream)
{
	SimCtx<R> h = proto;
	const size_t n = (size_t)std::max(h.cap, 1);
	unsigned char* p = area;
	auto take = [&p](size_t bytes) {
		unsigned char* q = p;
		p += noalloc_align(bytes);
		return q;
	};
	SimCtx<R>* d_ctx = (SimCtx<R>*)take(sizeof(SimCtx<double>));
	h.sc = (gptr<Scalars>)take(sizeof(Scalars));
	h.vcnt = (gptr<int>)take(4 * (size_t)h.vcount);
	h.vstart = (gptr<int>)take(4 * (size_t)h.vcount);
	h.vox_of = (gptr<int>)take(4 * n);
	h.vrank = (gptr<int>)take(4 * n);
	h.btmp = (gptr<int>)take(4 * n);
	h.bucket = (gptr<int>)take(4 * n);
	h.occ = (gptr<int>)take(4 * n);
	h.sorted = (gptr<R>)take(4 * sizeof(R) * n);
	if ((size_t)(p - area) > PCM_NOALLOC_SCRATCH_BYTES)
		return fail(PCM_ERR_CAPACITY, "scratch layout exceeds the area"); // checked by the caller first

	const int agents = h.cap;
	const int blocks = std::max(1, (agents + BLOCK - 1) / BLOCK);
	const int clear_blocks = (int)std::min<long long>(std::max<long long>(1, (h.vcount + BLOCK - 1) / BLOCK), 8192);
	hipLaunchKernelGGL(k_noalloc_setup<R>, dim3(clear_blocks), dim3(BLOCK), 0, stream, h, d_ctx, needed);
	PCM_HIP(hipGetLastError());
	const SimCtx<R>* ctx = d_ctx;
	hipLaunchKernelGGL((k_bin_count<R, D>), dim3(blocks), dim3(BLOCK), 0, stream, ctx, (const int4*)nullptr);
	hipLaunchKernelGGL((k_bin_claim<R, false>), dim3(std::max(1, (agents + CLAIM_THREADS - 1) / CLAIM_THREADS)),
					   dim3(CLAIM_THREADS), 0, stream, ctx, (const int4*)nullptr);
	hipLaunchKernelGGL(k_bin_scatter<R>, dim3(blocks), dim3(BLOCK), 0, stream, ctx, (const int4*)nullptr);
	hipLaunchKernelGGL((k_bin_order<R, D, false>), dim3(blocks), dim3(BLOCK), 0, stream, ctx, (const int4*)nullptr);
	hipLaunchKernelGGL((k_neighbors<R, D, 2>), dim3(std::max(1, std::min(agents, 16384))), dim3(128), 0, stream, ctx, (const int4*)nullptr);
	hipLaunchKernelGGL(k_noalloc_finish<R>, dim3(blocks), dim3(BLOCK), 0, stream, ctx, needed);
	PCM_HIP(hipGetLastError());
	return PCM_OK;
}

template <class R>
int noalloc_dispatch(const SimCtx<R>& h, unsigned char* area, int* needed, hipStream_t stream)
{
	switch (h.dims)
	{
		case 1:
			return noalloc_build<R, 1>(h, area, needed, stream);
		case 2:
			return noalloc_build<R, 2>(h, area, needed, stream);
		default:
			return noalloc_build<R, 3>(h, area, needed, stream);
	}
}

template <class R>
int noalloc_entry(const pcm_mesh_desc& mesh, int count, const void* positions, const void* radius, const void* rel_adh,
				  const uint8_t* movable, int K, int32_t* counts, int32_t* indices, int32_t* needed, hipStream_t stream)
{
	SimCtx<R> h;
	std::memset(&h, 0, sizeof(h));
	h.dims = mesh.dims;
	for (int k = 0; k < 3; k++)
	{
		h.bmin[k] = mesh.bbox_min[k];
		h.bmax[k] = mesh.bbox_max[k];
		h.vox[k] = mesh.voxel[k];
		if (k < h.dims)
		{
			if (h.vox[k] <= 0 || h.bmax[k] < h.bmin[k])
				return fail(PCM_ERR_ARGUMENT, "bad mesh box or voxel shape");
			h.grid[k] = std::max(1, (h.bmax[k] - h.bmin[k] + h.vox[k] - 1) / h.vox[k]);
		}
		else
			h.grid[k] = 1;
	}
	h.vcount = (long long)h.grid[0] * h.grid[1] * h.grid[2];
	const size_t need = noalloc_bytes(h.vcount, count, (int)sizeof(R));
	if (need > PCM_NOALLOC_SCRATCH_BYTES)
		return fail(PCM_ERR_CAPACITY, "pcm_neighbors_noalloc: " + std::to_string(h.vcount) + " voxels and " +
										  std::to_string(count) + " agents need " + std::to_string(need) +
										  " bytes of scratch, above PCM_NOALLOC_SCRATCH_BYTES (" +
										  std::to_string((size_t)PCM_NOALLOC_SCRATCH_BYTES) + ")");
	h.cap = count;
	h.K = K;
	h.pos = (gptr<R>)(R*)positions;
	h.radius = (gptr<R>)(R*)radius;
	h.rel_adh = (gptr<R>)(R*)rel_adh;
	h.movable = (gptr<uint8_t>)(uint8_t*)movable;
	h.nbr_cnt = (gptr<int>)counts;
	h.nbr_idx = (gptr<int>)indices;
	unsigned char* area = nullptr;
	PCM_TRY(scratch_of_current_device(&area));
	return noalloc_dispatch<R>(h, area, needed, stream);
}

} // namespace

// ============================================================================ C-ABI
extern "C" {

int pcm_version(void) { return 1; }

const char* pcm_last_error(void) { return g_error.c_str(); }

int pcm_field_info(int32_t field, int32_t* type, int32_t* width_kind, const char** name)
{
	if (field < 0 || field >= PCM_FIELD_COUNT)
		return fail(PCM_ERR_ARGUMENT, "bad field id");
	if (type)
		*type = k_fields[field].type;
	if (width_kind)
		*width_kind = k_fields[field].width_kind;
	if (name)
		*name = k_fields[field].name;
	return PCM_OK;
}

int pcm_create(const pcm_config* config, pcm_sim** out)
{
	if (!config || !out)
		return fail(PCM_ERR_ARGUMENT, "null argument");
	*out = nullptr;
	const auto& c = *config;
	if (c.mesh.dims < 1 || c.mesh.dims > 3)
		return fail(PCM_ERR_ARGUMENT, "mesh dims must be 1, 2 or 3");
	if (c.cell_defs < 1 || c.substrates < 0 || c.capacity < 0)
		return fail(PCM_ERR_ARGUMENT, "bad cell definition, substrate or capacity count");
	int devices = 0;
	if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0)
		return fail(PCM_ERR_DEVICE, "no HIP device is available");
	std::shared_lock<std::shared_mutex> capture_guard(g_capture_mu);
	std::unique_ptr<pcm_sim> s;
	if (c.real_bytes == 4)
		s.reset(new sim_impl<float>(c));
	else if (c.real_bytes == 8)
		s.reset(new sim_impl<double>(c));
	else
		return fail(PCM_ERR_ARGUMENT, "real_bytes must be 4 or 8");
	const int rc = s->init();
	if (rc != PCM_OK)
		return rc;
	*out = s.release();
	return PCM_OK;
}

int pcm_destroy(pcm_sim* sim)
{
	if (!sim)
		return PCM_OK;
	std::shared_lock<std::shared_mutex> capture_guard(g_capture_mu);
	if (auto* s4 = dynamic_cast<sim_impl<float>*>(sim); s4 && s4->batch())
		PCM_TRY(((pcm_batch*)s4->batch())->retire(sim));
	if (auto* s8 = dynamic_cast<sim_impl<double>*>(sim); s8 && s8->batch())
		PCM_TRY(((pcm_batch*)s8->batch())->retire(sim));
	delete sim;
	return PCM_OK;
}

int pcm_batch_create(int32_t real_bytes, pcm_batch** out)
{
	if (!out)
		return fail(PCM_ERR_ARGUMENT, "null argument");
	*out = nullptr;
	if (real_bytes == 4)
		*out = new batch_impl<float>();
	else if (real_bytes == 8)
		*out = new batch_impl<double>();
	else
		return fail(PCM_ERR_ARGUMENT, "real_bytes must be 4 or 8");
	return PCM_OK;
}

int pcm_batch_destroy(pcm_batch* batch)
{
	delete batch;
	return PCM_OK;
}

int pcm_batch_admit(pcm_batch* batch, pcm_sim* sim)
{
	if (!batch || !sim)
		return fail(PCM_ERR_ARGUMENT, "null handle");
	return batch->admit(sim);
}

int pcm_batch_retire(pcm_batch* batch, pcm_sim* sim)
{
	if (!batch || !sim)
		return fail(PCM_ERR_ARGUMENT, "null handle");
	return batch->retire(sim);
}

int pcm_batch_size(pcm_batch* batch, int32_t* sims)
{
	if (!batch || !sims)
		return fail(PCM_ERR_ARGUMENT, "null argument");
	return batch->size(sims);
}

int pcm_set_step_params(pcm_sim* sim, const pcm_step_params* params)
{
	if (!sim || !params)
		return fail(PCM_ERR_ARGUMENT, "null handle or params");
	return sim->set_step_params(*params);
}

int pcm_batch_step(pcm_batch* batch, const pcm_step_params* params, int32_t steps)
{
	if (!batch)
		return fail(PCM_ERR_ARGUMENT, "null handle");
	return batch->step(params, steps);
}

int pcm_batch_synchronize(pcm_batch* batch)
{
	if (!batch)
		return fail(PCM_ERR_ARGUMENT, "null handle");
	return batch->synchronize();
}

int pcm_set_tuning(pcm_sim* sim, const pcm_tuning* tuning)
{
	if (!sim)
		return fail(PCM_ERR_ARGUMENT, "null handle");
	return sim->set_tuning(tuning);
}

int pcm_get_tuning(pcm_sim* sim, pcm_tuning* tuning)
{
	if (!sim)
		return fail(PCM_ERR_ARGUMENT, "null handle");
	return sim->get_tuning(tuning);
}

int pcm_autotune(pcm_sim* sim, const pcm_step_params* params, int32_t steps_per_candidate, pcm_tuning* chosen)
{
	if (!sim)
		return fail(PCM_ERR_ARGUMENT, "null handle");
	return sim->autotune(params, steps_per_candidate, chosen);
}

int pcm_batch_flatten(pcm_batch* batch, int32_t enable)
{
	if (!batch)
		return fail(PCM_ERR_ARGUMENT, "null handle");
	if (auto* b4 = dynamic_cast<batch_impl<float>*>(batch))
		b4->use_flattened_grids(enable != 0);
	else if (auto* b8 = dynamic_cast<batch_impl<double>*>(batch))
		b8->use_flattened_grids(enable != 0);
	return PCM_OK;
}

int pcm_batch_autotune(pcm_batch* batch, const pcm_step_params* params, int32_t steps_per_candidate)
{
	if (!batch)
		return fail(PCM_ERR_ARGUMENT, "null handle");
	return batch->autotune(params, steps_per_candidate);
}

int pcm_batch_counts(pcm_batch* batch, int32_t count, pcm_sim* const* sims, int32_t* counts)
{
	if (!batch || count < 0 || (count > 0 && (!sims || !counts)))
		return fail(PCM_ERR_ARGUMENT, "null argument");
	return batch->counts(count, sims, counts);
}

int pcm_batch_set_tuning(pcm_batch* batch, int32_t dims, const pcm_tuning* tuning)
{
	if (!batch)
		return fail(PCM_ERR_ARGUMENT, "null handle");
	return batch->set_tuning(dims, tuning);
}

int pcm_batch_get_tuning(pcm_batch* batch, int32_t dims, pcm_tuning* tuning)
{
	if (!batch || !tuning)
		return fail(PCM_ERR_ARGUMENT, "null argument");
	return batch->get_tuning(dims, tuning);
}

int pcm_batch_use_graphs(pcm_batch* batch, int32_t enable)
{
	if (!batch)
		return fail(PCM_ERR_ARGUMENT, "null handle");
	if (auto* b4 = dynamic_cast<batch_impl<float>*>(batch))
		b4->use_graphs(enable != 0);
	else if (auto* b8 = dynamic_cast<batch_impl<double>*>(batch))
		b8->use_graphs(enable != 0);
	return PCM_OK;
}

#define PCM_NEED(sim)                                                                                                  \
	if (!(sim))                                                                                                        \
		return fail(PCM_ERR_ARGUMENT, "null handle");                                                                  \
	std::shared_lock<std::shared_mutex> capture_guard_(g_capture_mu)

int pcm_set_count(pcm_sim* sim, int32_t count)
{
	PCM_NEED(sim);
	return sim->set_count(count);
}

int pcm_get_count(pcm_sim* sim, int32_t* count)
{
	PCM_NEED(sim);
	if (!count)
		return fail(PCM_ERR_ARGUMENT, "null output");
	return sim->get_count(count);
}

int pcm_set_field(pcm_sim* sim, int32_t field, const void* host)
{
	PCM_NEED(sim);
	return sim->set_field(field, host);
}

int pcm_get_fields(pcm_sim* sim, int32_t nfields, const int32_t* fields, void* const* hosts)
{
	PCM_NEED(sim);
	return sim->get_fields(nfields, fields, hosts);
}

int pcm_set_fields(pcm_sim* sim, int32_t nfields, const int32_t* fields, const void* const* hosts)
{
	PCM_NEED(sim);
	return sim->set_fields(nfields, fields, hosts);
}

int pcm_get_field(pcm_sim* sim, int32_t field, void* host)
{
	PCM_NEED(sim);
	return sim->get_field(field, host);
}

int pcm_set_links(pcm_sim* sim, int32_t which, const int32_t* offsets, const int32_t* indices)
{
	PCM_NEED(sim);
	return sim->set_links(which, offsets, indices);
}

int pcm_get_links_size(pcm_sim* sim, int32_t which, int64_t* total)
{
	PCM_NEED(sim);
	if (!total)
		return fail(PCM_ERR_ARGUMENT, "null output");
	return sim->get_links(which, nullptr, nullptr, total);
}

int pcm_get_links(pcm_sim* sim, int32_t which, int32_t* offsets, int32_t* indices)
{
	PCM_NEED(sim);
	int64_t t = 0;
	if (!offsets || !indices)
		return fail(PCM_ERR_ARGUMENT, "null output");
	return sim->get_links(which, offsets, indices, &t);
}

int pcm_set_neighbors(pcm_sim* sim, const int32_t* offsets, const int32_t* indices)
{
	PCM_NEED(sim);
	return sim->set_neighbors(offsets, indices);
}

int pcm_get_neighbors_size(pcm_sim* sim, int64_t* total)
{
	PCM_NEED(sim);
	if (!total)
		return fail(PCM_ERR_ARGUMENT, "null output");
	return sim->get_neighbors(nullptr, nullptr, total);
}

int pcm_get_neighbors(pcm_sim* sim, int32_t* offsets, int32_t* indices)
{
	PCM_NEED(sim);
	int64_t t = 0;
	if (!offsets || !indices)
		return fail(PCM_ERR_ARGUMENT, "null output");
	return sim->get_neighbors(offsets, indices, &t);
}

int pcm_get_voxel_count(pcm_sim* sim, int64_t* voxel_count)
{
	PCM_NEED(sim);
	if (!voxel_count)
		return fail(PCM_ERR_ARGUMENT, "null output");
	*voxel_count = sim->voxel_count();
	return PCM_OK;
}

int pcm_get_voxel_buckets(pcm_sim* sim, int32_t* counts, int32_t* agents)
{
	PCM_NEED(sim);
	if (!counts || !agents)
		return fail(PCM_ERR_ARGUMENT, "null output");
	return sim->get_buckets(counts, agents);
}

int pcm_pack_fields(pcm_sim* sim, int32_t nfields, const int32_t* fields, const int32_t* widths, double* out)
{
	PCM_NEED(sim);
	if (nfields < 1 || !fields || !widths || !out)
		return fail(PCM_ERR_ARGUMENT, "null argument or no fields");
	return sim->pack_fields(nfields, fields, widths, out);
}

int pcm_update_cell_container_for_phenotype(pcm_sim* sim, const pcm_step_params* params,
											int32_t compute_internalized_substrates)
{
	PCM_NEED(sim);
	if (!params)
		return fail(PCM_ERR_ARGUMENT, "null params");
	return sim->phenotype_update(params, compute_internalized_substrates);
}

int pcm_get_divisions_count(pcm_sim* sim, int64_t* divisions)
{
	PCM_NEED(sim);
	if (!divisions)
		return fail(PCM_ERR_ARGUMENT, "null output");
	return sim->get_divisions(divisions);
}

int pcm_division_log(pcm_sim* sim, int32_t* parents, int64_t capacity, int64_t* length)
{
	PCM_NEED(sim);
	if (!length)
		return fail(PCM_ERR_ARGUMENT, "null length");
	return sim->division_log(parents, capacity, length);
}

int pcm_get_deaths_count(pcm_sim* sim, int64_t* deaths)
{
	PCM_NEED(sim);
	if (!deaths)
		return fail(PCM_ERR_ARGUMENT, "null output");
	return sim->get_deaths(deaths);
}

int pcm_set_gradients(pcm_sim* sim, const int32_t diffusion_voxel[3], const void* gradients)
{
	PCM_NEED(sim);
	if (!diffusion_voxel || !gradients)
		return fail(PCM_ERR_ARGUMENT, "null diffusion voxel or gradients");
	return sim->set_gradients(diffusion_voxel, gradients);
}

int pcm_set_step_index(pcm_sim* sim, uint64_t step)
{
	PCM_NEED(sim);
	return sim->set_step(step);
}

int pcm_get_step_index(pcm_sim* sim, uint64_t* step)
{
	PCM_NEED(sim);
	if (!step)
		return fail(PCM_ERR_ARGUMENT, "null output");
	return sim->get_step(step);
}

int pcm_update_mechanics_mesh(pcm_sim* sim)
{
	PCM_NEED(sim);
	return sim->phase(PH_MESH, nullptr);
}

int pcm_update_cell_neighbors(pcm_sim* sim)
{
	PCM_NEED(sim);
	return sim->phase(PH_NEIGHBORS, nullptr);
}

#define PCM_PHASE_ENTRY(NAME, ID)                                                                                      \
	int NAME(pcm_sim* sim, const pcm_step_params* params)                                                              \
	{                                                                                                                  \
		PCM_NEED(sim);                                                                                                 \
		if (!params)                                                                                                   \
			return fail(PCM_ERR_ARGUMENT, "null params");                                                              \
		return sim->phase(ID, params);                                                                                 \
	}

PCM_PHASE_ENTRY(pcm_evaluate_interactions, PH_CONTACT)
PCM_PHASE_ENTRY(pcm_update_cell_forces, PH_FORCES)
PCM_PHASE_ENTRY(pcm_update_motility, PH_MOTILITY)
PCM_PHASE_ENTRY(pcm_update_basement_membrane_interactions, PH_MEMBRANE)
PCM_PHASE_ENTRY(pcm_update_spring_attachments, PH_SPRINGS)
PCM_PHASE_ENTRY(pcm_update_positions, PH_POSITIONS)
PCM_PHASE_ENTRY(pcm_update_cell_cell_interactions, PH_INTERACTIONS)
PCM_PHASE_ENTRY(pcm_update_cell_container_for_mechanics, PH_CONTAINER)

int pcm_step(pcm_sim* sim, const pcm_step_params* params, int32_t steps)
{
	PCM_NEED(sim);
	return sim->step(params, steps);
}

int pcm_synchronize(pcm_sim* sim)
{
	PCM_NEED(sim);
	return sim->synchronize();
}

int pcm_removal_log(pcm_sim* sim, int32_t* positions, int64_t capacity, int64_t* length)
{
	PCM_NEED(sim);
	if (!length)
		return fail(PCM_ERR_ARGUMENT, "null output");
	return sim->removal_log(positions, capacity, length);
}

int pcm_get_stats(pcm_sim* sim, int32_t* stats)
{
	PCM_NEED(sim);
	if (!stats)
		return fail(PCM_ERR_ARGUMENT, "null output");
	return sim->stats(stats);
}

int pcm_kernel_name(int32_t kernel, const char** name)
{
#define PCM_KERNEL_NAME_(n) #n,
	static const char* names[PCM_KERNEL_COUNT] = { PCM_KERNEL_LIST(PCM_KERNEL_NAME_) };
#undef PCM_KERNEL_NAME_
	if (kernel < 0 || kernel >= PCM_KERNEL_COUNT || !name)
		return fail(PCM_ERR_ARGUMENT, "bad kernel id");
	*name = names[kernel];
	return PCM_OK;
}

int pcm_profile_every(pcm_sim* sim, int32_t every)
{
	PCM_NEED(sim);
	return sim->profile_every(every);
}

int pcm_profile_begin(pcm_sim* sim, uint64_t kernel_mask)
{
	PCM_NEED(sim);
	return sim->profile_begin(kernel_mask);
}

int pcm_profile_end(pcm_sim* sim, double* total_ms, int64_t* launches)
{
	PCM_NEED(sim);
	return sim->profile_end(total_ms, launches);
}

int pcm_get_mesh_position(const pcm_mesh_desc* mesh, int32_t real_bytes, const void* position, int32_t voxel[3])
{
	if (!mesh || !position || !voxel || mesh->dims < 1 || mesh->dims > 3 || (real_bytes != 4 && real_bytes != 8))
		return fail(PCM_ERR_ARGUMENT, "pcm_get_mesh_position: null pointer, dims or real_bytes out of range");
	int grid[3];
	PCM_TRY(grid_of(*mesh, grid));
	for (int k = 0; k < 3; k++)
		voxel[k] = k < mesh->dims
					   ? (real_bytes == 4 ? host_voxel_coord<float>(((const float*)position)[k], *mesh, grid, k)
										  : host_voxel_coord<double>(((const double*)position)[k], *mesh, grid, k))
					   : 0;
	return PCM_OK;
}

int pcm_get_mesh_index(const pcm_mesh_desc* mesh, const int32_t voxel[3], int32_t* index)
{
	if (!mesh || !voxel || !index || mesh->dims < 1 || mesh->dims > 3)
		return fail(PCM_ERR_ARGUMENT, "pcm_get_mesh_index: null pointer or dims out of range");
	int grid[3];
	PCM_TRY(grid_of(*mesh, grid));
	for (int k = 0; k < 3; k++)
		if (voxel[k] < 0 || voxel[k] >= grid[k])
			return fail(PCM_ERR_ARGUMENT, "pcm_get_mesh_index: voxel coordinates outside the grid");
	*index = voxel[0] + grid[0] * (voxel[1] + grid[1] * voxel[2]);
	return PCM_OK;
}

int pcm_neighbors_noalloc(const pcm_mesh_desc* mesh, int32_t real_bytes, int32_t count, const void* positions,
						  const void* radius, const void* relative_maximum_adhesion_distance, const uint8_t* is_movable,
						  int32_t max_neighbors, int32_t* neighbor_counts, int32_t* neighbor_indices, int32_t* needed,
						  void* stream)
{
	if (!mesh || mesh->dims < 1 || mesh->dims > 3)
		return fail(PCM_ERR_ARGUMENT, "pcm_neighbors_noalloc: mesh dims must be 1, 2 or 3");
	if (real_bytes != 4 && real_bytes != 8)
		return fail(PCM_ERR_ARGUMENT, "pcm_neighbors_noalloc: real_bytes must be 4 or 8");
	if (count < 0 || max_neighbors < 1)
		return fail(PCM_ERR_ARGUMENT, "pcm_neighbors_noalloc: negative count or max_neighbors below 1");
	if (!needed || (count > 0 && (!positions || !radius || !relative_maximum_adhesion_distance || !neighbor_counts ||
								  !neighbor_indices)))
		return fail(PCM_ERR_ARGUMENT, "pcm_neighbors_noalloc: null device buffer");
	if ((long long)count * max_neighbors > 0x7fffffffLL)
		return fail(PCM_ERR_ARGUMENT, "pcm_neighbors_noalloc: count * max_neighbors exceeds 2^31 entries");
	hipStream_t s = (hipStream_t)stream;
	if (real_bytes == 4)
		return noalloc_entry<float>(*mesh, count, positions, radius, relative_maximum_adhesion_distance, is_movable,
									max_neighbors, neighbor_counts, neighbor_indices, needed, s);
	return noalloc_entry<double>(*mesh, count, positions, radius, relative_maximum_adhesion_distance, is_movable,
								 max_neighbors, neighbor_counts, neighbor_indices, needed, s);
}

} // extern "C"

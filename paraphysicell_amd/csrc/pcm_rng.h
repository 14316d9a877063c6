// pcm_rng.h — counter-based random stream of the mechanics step.
//
// The reference draws from a thread_local std::mt19937 (src/random.cpp:8-49), whose
// stream depends on the OpenMP schedule.  The device path instead keys every draw by
// (simulation seed, mechanics step, stream, agent, slot, draw number) through
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
// 1, 2, 3", SC'11), so a draw is the same whatever thread, launch or batch computes it.
//
// Draw sites (reference file:line -> stream id):
//   update_motility_single     position_solver.cpp:127, solver_helper.h:84,162 -> MOTILITY
//   spring detachment          position_solver.cpp:386                         -> DETACH
//   spring attachment          position_solver.cpp:435                         -> ATTACH
//   cell-cell interactions     interactions_solver.cpp:167                     -> INTERACT
//   cell::divide, assign_orientation  cell.cpp:134-139,172-184                 -> DIVISION
//     (keyed by the phenotype-update call in place of the mechanics step)
//
// Header-only; compiles as host C++ (g++) and as HIP device code (hipcc).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
	#define PCM_HD __host__ __device__ __forceinline__
#else
	#define PCM_HD inline
#endif

enum pcm_stream : uint32_t
{
	PCM_STREAM_MOTILITY = 1,
	PCM_STREAM_DETACH = 2,
	PCM_STREAM_ATTACH = 3,
	PCM_STREAM_INTERACT = 4,
	PCM_STREAM_DIVISION = 5,
};

struct pcm_u32x4
{
	uint32_t v[4];
};

PCM_HD void pcm_mulhilo32(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo)
{
	const uint64_t p = (uint64_t)a * (uint64_t)b;
	hi = (uint32_t)(p >> 32);
	lo = (uint32_t)p;
}

// Philox4x32 with 10 rounds: multipliers and Weyl key increments from the SC'11 paper.
PCM_HD pcm_u32x4 pcm_philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1)
{
	const uint32_t m0 = 0xD2511F53u, m1 = 0xCD9E8D57u;
	const uint32_t w0 = 0x9E3779B9u, w1 = 0xBB67AE85u;
#if defined(__HIPCC__)
	#pragma unroll
#endif
	for (int r = 0; r < 10; ++r)
	{
		uint32_t hi0, lo0, hi1, lo1;
		pcm_mulhilo32(m0, c0, hi0, lo0);
		pcm_mulhilo32(m1, c2, hi1, lo1);
		const uint32_t n0 = hi1 ^ c1 ^ k0;
		const uint32_t n2 = hi0 ^ c3 ^ k1;
		c0 = n0;
		c1 = lo1;
		c2 = n2;
		c3 = lo0;
		k0 += w0;
		k1 += w1;
	}
	pcm_u32x4 out;
	out.v[0] = c0;
	out.v[1] = c1;
	out.v[2] = c2;
	out.v[3] = c3;
	return out;
}

// [0,1) with 24 random bits (float) or 53 random bits (double) — the same range as
// std::generate_canonical, which std::uniform_real_distribution uses in the reference.
PCM_HD float pcm_u01f(uint32_t x) { return (float)(x >> 8) * 5.9604644775390625e-08f; }
PCM_HD double pcm_u01d(uint32_t a, uint32_t b)
{
	return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * 1.1102230246251565404e-16;
}

// One Philox block serves 4 float draws or 2 double draws of the same (stream, agent, slot).
PCM_HD pcm_u32x4 pcm_block(uint64_t seed, uint64_t step, uint32_t stream, uint32_t agent, uint32_t slot,
						   uint32_t block)
{
	return pcm_philox4x32_10(agent, slot, (uint32_t)step, (stream << 16) | (block & 0xFFFFu), (uint32_t)seed,
							 (uint32_t)(seed >> 32));
}

template <class R>
struct pcm_draw;

template <>
struct pcm_draw<float>
{
	static constexpr uint32_t per_block = 4;
	PCM_HD static float from(const pcm_u32x4& b, uint32_t k) { return pcm_u01f(b.v[k]); }
};

template <>
struct pcm_draw<double>
{
	static constexpr uint32_t per_block = 2;
	PCM_HD static double from(const pcm_u32x4& b, uint32_t k) { return pcm_u01d(b.v[2 * k], b.v[2 * k + 1]); }
};

// The `draw`-th uniform [0,1) number of (seed, step, stream, agent, slot).
template <class R>
PCM_HD R pcm_uniform01(uint64_t seed, uint64_t step, uint32_t stream, uint32_t agent, uint32_t slot, uint32_t draw)
{
	const uint32_t per = pcm_draw<R>::per_block;
	const pcm_u32x4 b = pcm_block(seed, step, stream, agent, slot, draw / per);
	return pcm_draw<R>::from(b, draw % per);
}

// std::uniform_real_distribution<R>(lo, hi) maps a canonical draw u to u * (hi - lo) + lo.
template <class R>
PCM_HD R pcm_scale_uniform(R u, R lo, R hi)
{
	return u * (hi - lo) + lo;
}

// Device-side helpers shared by the kernel sources.
//
// Amplitude types: double2 (complex128) and float2 (complex64).  Arithmetic
// runs in the amplitude's own real type; every norm, probability and other
// reduction accumulates in FP64.
#pragma once

#include "dk_internal.h"

namespace dkd {

constexpr int TPB = 256;          // threads per block (4 wave64)
constexpr int WAVES = TPB / 64;

template <typename T> struct real_of;
template <> struct real_of<double2> { typedef double type; };
template <> struct real_of<float2> { typedef float type; };

template <typename T> __device__ __forceinline__ T czero()
{
	T v;
	v.x = 0;
	v.y = 0;
	return v;
}

template <typename T, typename R> __device__ __forceinline__ T mk(R x, R y)
{
	T v;
	v.x = x;
	v.y = y;
	return v;
}

// acc + m * v
template <typename T> __device__ __forceinline__ T cmac(T acc, T m, T v)
{
	acc.x = fma(m.x, v.x, acc.x);
	acc.x = fma(-m.y, v.y, acc.x);
	acc.y = fma(m.x, v.y, acc.y);
	acc.y = fma(m.y, v.x, acc.y);
	return acc;
}

template <typename T> __device__ __forceinline__ T cmul(T a, T b)
{
	T r;
	r.x = fma(a.x, b.x, -a.y * b.y);
	r.y = fma(a.x, b.y, a.y * b.x);
	return r;
}

template <typename T, typename R> __device__ __forceinline__ T cscale(T v, R s)
{
	v.x *= s;
	v.y *= s;
	return v;
}

// |v|^2 in FP64 whatever the amplitude type
template <typename T> __device__ __forceinline__ double cabs2d(T v)
{
	const double x = v.x, y = v.y;
	return fma(x, x, y * y);
}

template <typename T> __device__ __forceinline__ void st_stream(T *p, T v)
{
	__builtin_nontemporal_store(v.x, &p->x);
	__builtin_nontemporal_store(v.y, &p->y);
}

// Streaming (non-temporal) load of an amplitude read once per pass: on
// gfx950 this lifts full-state passes by 3-6% over cached loads (measured
// at 30 qubits: 1-qubit gate 6.48 -> 6.70 TB/s, copy 6.26 -> 6.40 TB/s).
template <typename T> __device__ __forceinline__ T ld_stream(const T *p)
{
	T v;
	v.x = __builtin_nontemporal_load(&p->x);
	v.y = __builtin_nontemporal_load(&p->y);
	return v;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1)
		v += __shfl_xor(v, o);
	return v;
}

// Sum over a TPB-thread block; valid in thread 0; ends synchronised.
__device__ __forceinline__ double block_sum(double v, double *red)
{
	v = wave_sum(v);
	if ((threadIdx.x & 63) == 0)
		red[threadIdx.x >> 6] = v;
	__syncthreads();
	double t = 0.0;
	if (threadIdx.x == 0) {
#pragma unroll
		for (int w = 0; w < WAVES; ++w)
			t += red[w];
	}
	__syncthreads();
	return t;
}

// Insert a zero bit at position p.
__device__ __forceinline__ u64 ins0(u64 x, unsigned p)
{
	const u64 lo = (1ull << p) - 1;
	return ((x & ~lo) << 1) | (x & lo);
}

// LDS tile description for k-qubit gates: the 2^npos amplitudes whose index
// bits vary over pos[] (targets plus the lowest non-target bits), in slot
// order; the other index bits are fixed by the tile number.
constexpr int TILE_BITS = 12;
constexpr int TILE = 1 << TILE_BITS;

struct TileArgs {
	u64 dep_lo[64];            // slot bits 0..5 -> index bits
	u64 dep_hi[64];            // slot bits 6..11 -> index bits
	unsigned short colslot[64];// gate column j -> slot offset of its targets
	unsigned char pos[TILE_BITS]; // tile positions, ascending
	unsigned char tslot[8];    // slot positions of the targets, ascending
	int npos;
	// index bits known to be zero in the input (dense gates): a tile whose
	// fixed bits meet zmask holds only zeros, and so does its output,
	// written as zeros when zwrite (out of place), else left as it is
	int zwrite;
	u64 zmask;
};

__device__ __forceinline__ u64 tile_base(u64 tile, const TileArgs &ta)
{
	u64 base = tile;
	for (int b = 0; b < ta.npos; ++b)
		base = ins0(base, ta.pos[b]);
	return base;
}

__device__ __forceinline__ u64 tile_index(u64 base, int s, const TileArgs &ta)
{
	return base | ta.dep_lo[s & 63] | ta.dep_hi[s >> 6];
}

// LDS position of tile slot s.  Gate kernels read a tile along strides of
// 16, 256, ... slots (one group per lane); XOR-folding the higher slot bits
// into the low ones spreads such reads over all LDS banks.  A bijection on
// every aligned run of 16 (complex128) or 32 (complex64) slots.
template <typename T> __device__ __forceinline__ int lds_pos(int s);
template <> __device__ __forceinline__ int lds_pos<double2>(int s)
{
	return s ^ (((s >> 4) ^ (s >> 8)) & 15);
}
template <> __device__ __forceinline__ int lds_pos<float2>(int s)
{
	return s ^ (((s >> 5) ^ (s >> 10)) & 31);
}

// Slot of group g: its bits deposited around the (sorted) target slots.
template <int K>
__device__ __forceinline__ int group_slot(int g, const TileArgs &ta)
{
#pragma unroll
	for (int b = 0; b < K; ++b)
		g = (int)ins0((u64)g, ta.tslot[b]);
	return g;
}

} // namespace dkd

// Host helper: tile description for a gate on `targets` in an n-qubit state;
// the tile spans min(tile_bits, n) index bits (tile_bits <= 12).
void dk_tile_args(unsigned n, unsigned k, const unsigned *targets,
		  dkd::TileArgs &ta, int tile_bits = dkd::TILE_BITS);

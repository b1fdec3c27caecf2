// Batched small registers: one persistent launch runs many independent
// registers of uneven size, each with its own gate list, an optional reduced
// density matrix and a full measurement with injected rolls.
//
// Optionally every register also reports the marginal P(qubit q = 1) of its
// final state for every q (registry_prob, qops.c:72-95), in FP64.
//
// Work distribution: workgroups take registers from an atomic counter,
// largest first.  A register of up to 2^LB amplitudes (LB = 12 for
// complex128, 13 for complex64) lives entirely in LDS.  Larger registers live
// in a device pool and run as cache-blocked passes: a pass is a run of
// consecutive gates whose targets fit one 2^LB-amplitude tile; the workgroup
// streams the register through LDS tile by tile, applying the whole run to
// each tile (controls outside the tile are decided per tile).  Registers of
// at least "split_qubits" qubits (14 by default) would hold one workgroup
// for long (a 2^20-amplitude register streams at about one CU's share of
// HBM), so their gates run before the persistent launch as rounds of
// cache-blocked passes, planned per register as a circuit's passes are
// (dk_circuit.cpp: tiles of qubits 0..5 plus six chosen qubits held in
// registers, gates merged and applied in registers, LDS only for
// exchanges; every gate in some pass): round r is one launch of the pass
// kernel's round form (dk_pass.hip) over the tiles of the r-th pass of
// every such register, the launch boundary ordering the passes; the
// persistent launch then only reads their final states.  Results agree
// with the one-workgroup form to rounding (merged gates, another order).
//
// Gate semantics are registry_apply's (qops.c:176-261); since every gate is
// unitary, renormalising once after the last gate gives the reference's
// per-gate renormalised values to rounding.  Measurement visits qubits from
// the most significant down (doki.c:1097-1121) with outcome P1 > roll
// (qops.c:132); P1 of each step is the squared norm of the upper half of the
// still-possible index range over that of the whole range, in FP64.
#include <algorithm>
#include <cstring>
#include <vector>

#include "dk_device.h"
#include "dk_pass.h"

using namespace dkd;

int dk_plan_register_passes(unsigned nq, size_t count, const double *mats,
			    const unsigned *targets, const uint64_t *cms,
			    const uint64_t *ams, std::vector<dkp::PassDesc> &descs,
			    std::vector<dkp::PassOp> &ops,
			    std::vector<uint32_t> &op_begin);
int dk_launch_pass_round(int dtype, void *pool, const dkp::PassEntry *ents,
			 unsigned nents, const dkp::PassOp *ops, u64 nitems,
			 int init, unsigned grid, hipStream_t st);

namespace {

constexpr int BT = 512;            // threads per workgroup
constexpr int BW = BT / 64;        // waves
constexpr int LDS_BYTES = 65536;   // tile storage per workgroup

// Chunk sums of the split registers' final states (their measurement,
// marginals and norm read these instead of streaming the state through one
// workgroup).  A chunk is CHUNK consecutive amplitudes of one register; its
// record holds CREC doubles: [0] the sum of |psi_i|^2 over the chunk,
// [1 + b] the part of it over indices with bit b set (b < 12).  Sums are
// taken in a fixed order.
constexpr int CHUNK_BITS = 12, CHUNK = 1 << CHUNK_BITS, CREC = 1 + CHUNK_BITS;
constexpr int CT = 256;            // threads of the chunk-sum kernel

struct BReg {
	u64 off;        // amplitude offset of its pool slot (pooled registers)
	u64 coff;       // split registers: first of its chunks' sum records
	u64 keep;       // density: kept qubits (0 = no density)
	u64 zmask;      // split registers: index bits outside the final state's
	                // support (amplitudes there are zeros the pool may not
	                // hold: they are taken as zero, never read)
	unsigned nq;
	unsigned pass_begin, pass_end;
	unsigned roll_off;  // first roll; also first outcome slot
	unsigned dens_off;  // first element of its density output
	unsigned flags;     // 1 measure, 2 density, 4 export, 8 marginal probs
};

struct BPass {
	u64 pmask;                 // index bits covered by the tile
	unsigned gate_begin, gate_end;
	unsigned npos;
	unsigned char pos[16];     // tile positions, ascending
	unsigned char slot_of[64]; // qubit -> tile slot bit (positions only)
};

// A gate of a pass: the host splits its controls at the pass's tile, so a
// tile only tests the outside part against its fixed index bits.
struct BGate {
	double m[8];        // 2x2 row-major, interleaved
	u64 cm, am;         // controls / anticontrols outside the tile (index bits)
	unsigned cin, ain;  // controls / anticontrols inside the tile (slot bits)
	unsigned t;         // target qubit
	unsigned sb;        // target slot
};

enum { BF_MEASURE = 1, BF_DENSITY = 2, BF_EXPORT = 4, BF_PROBS = 8, BF_SPLIT = 16 };

__device__ __forceinline__ double bsum(double v, double *red)
{
	v = wave_sum(v);
	if ((threadIdx.x & 63) == 0)
		red[threadIdx.x >> 6] = v;
	__syncthreads();
	double t = 0.0;
#pragma unroll
	for (int w = 0; w < BW; ++w)
		t += red[w];
	__syncthreads();
	return t; // valid in every thread
}

__device__ __forceinline__ u64 deposit(u64 x, const unsigned char *pos, int n)
{
	u64 r = 0;
	for (int b = 0; b < n; ++b)
		r |= ((x >> b) & 1ull) << pos[b];
	return r;
}

// Index of slot s = threadIdx.x + k * BT of a pooled register's tile: the
// deposit is linear over disjoint slot bits, so a thread deposits its own
// part once and ORs in the part of k * BT (the same for the whole block).
template <int PER>
struct TileIndex {
	u64 own;
	u64 part[PER];
	__device__ __forceinline__ TileIndex(u64 base, const BPass &ps)
	{
		own = base | deposit((u64)threadIdx.x, ps.pos, ps.npos);
#pragma unroll
		for (int k = 0; k < PER; ++k)
			part[k] = deposit((u64)k * BT, ps.pos, ps.npos);
	}
	__device__ __forceinline__ u64 operator[](int k) const { return own | part[k]; }
};

// A gate as one tile sees it: its matrix, the in-tile control masks (slot
// bits; `ctl` when there are any) and whether the out-of-tile controls
// leave it on for this tile.  Slot arithmetic is 32-bit (a tile holds at
// most 2^13 slots).
template <typename T> struct TileGate {
	T m00, m01, m10, m11;
	unsigned cm_in, am_in;
	int sb;
	bool on, ctl;
};

template <typename T>
__device__ __forceinline__ TileGate<T> tile_gate_of(const BGate &g, u64 base)
{
	typedef typename real_of<T>::type R;
	TileGate<T> c;
	c.cm_in = g.cin;
	c.am_in = g.ain;
	c.on = (base & g.cm) == g.cm && (base & g.am) == 0;
	c.ctl = (g.cin | g.ain) != 0;
	c.sb = (int)g.sb;
	c.m00 = mk<T>((R)g.m[0], (R)g.m[1]);
	c.m01 = mk<T>((R)g.m[2], (R)g.m[3]);
	c.m10 = mk<T>((R)g.m[4], (R)g.m[5]);
	c.m11 = mk<T>((R)g.m[6], (R)g.m[7]);
	return c;
}

__device__ __forceinline__ int ins0i(int x, int b)
{
	return ((x >> b) << (b + 1)) | (x & ((1 << b) - 1));
}

// K consecutive gates on pairwise distinct slot bits in one sweep: a thread
// owns the 2^K amplitudes at s | (any subset of the K target slots), reads
// them from LDS once, applies the gates in order in registers and writes
// them back once -- each amplitude sees the same operations in the same
// order as with one sweep per gate.  Bit i of a member's number is gate i's
// target slot.
template <int K, typename T>
__device__ void tile_group(T *tile, int tb, const TileGate<T> *c)
{
	constexpr int N = 1 << K;
	int sl[K];
#pragma unroll
	for (int i = 0; i < K; ++i)
		sl[i] = c[i].sb;
	// ascending, for the zero insertions
#pragma unroll
	for (int i = 0; i < K; ++i)
#pragma unroll
		for (int j = 0; j + 1 < K - i; ++j)
			if (sl[j] > sl[j + 1]) {
				const int t = sl[j];
				sl[j] = sl[j + 1];
				sl[j + 1] = t;
			}
	int off[N];
#pragma unroll
	for (int m = 0; m < N; ++m) {
		int o = 0;
#pragma unroll
		for (int i = 0; i < K; ++i)
			if ((m >> i) & 1)
				o |= 1 << c[i].sb;
		off[m] = o;
	}
	const int groups = 1 << (tb - K);
	for (int q = threadIdx.x; q < groups; q += BT) {
		int s = q;
#pragma unroll
		for (int i = 0; i < K; ++i)
			s = ins0i(s, sl[i]);
		T v[N];
#pragma unroll
		for (int m = 0; m < N; ++m)
			v[m] = tile[lds_pos<T>(s | off[m])];
#pragma unroll
		for (int i = 0; i < K; ++i) {
			const TileGate<T> &g = c[i];
			if (!g.on)
				continue;
#pragma unroll
			for (int m = 0; m < N; ++m) {
				if ((m >> i) & 1)
					continue;
				const unsigned s0 = (unsigned)(s | off[m]);
				if (g.ctl && ((s0 & g.cm_in) != g.cm_in || (s0 & g.am_in) != 0))
					continue;
				const int m1 = m | (1 << i);
				const T x = v[m], y = v[m1];
				v[m] = cmac(cmac(czero<T>(), g.m00, x), g.m01, y);
				v[m1] = cmac(cmac(czero<T>(), g.m10, x), g.m11, y);
			}
		}
#pragma unroll
		for (int m = 0; m < N; ++m)
			tile[lds_pos<T>(s | off[m])] = v[m];
	}
}

// Every gate of pass ps on a tile of 2^tb slots (all threads; ends with a
// barrier).  Gate records are staged in shared memory GCH at a time, and
// consecutive gates on distinct slot bits go two per sweep.
constexpr int GCH = 32;

template <typename T>
__device__ void tile_pass(T *tile, int tb, const BPass &ps, u64 base,
			  const BGate *__restrict__ gates, BGate *gsh)
{
	// the gate range in registers: after the last barrier below, thread 0
	// may already be writing the next pass into `ps`
	const unsigned gb = ps.gate_begin, ge = ps.gate_end;
	for (unsigned g0 = gb; g0 < ge; g0 += GCH) {
		const int cnt = (int)min((unsigned)GCH, ge - g0);
		for (int i = threadIdx.x; i < cnt; i += BT)
			gsh[i] = gates[g0 + i];
		__syncthreads();
		for (int i = 0; i < cnt;) {
			// the next one or two gates on distinct slots (groups of
			// three measured slower on gfx950: their registers cost
			// occupancy, 24.6 against 18.1 ms for 4096 registers)
			const TileGate<T> c0 = tile_gate_of<T>(gsh[i], base);
			TileGate<T> c1 = c0;
			int k = 1;
			if (i + 1 < cnt && tb >= 2) {
				c1 = tile_gate_of<T>(gsh[i + 1], base);
				if (c1.sb != c0.sb)
					k = 2;
			}
			if (k == 2) {
				const TileGate<T> c[2] = {c0, c1};
				tile_group<2>(tile, tb, c);
			} else {
				tile_group<1>(tile, tb, &c0);
			}
			__syncthreads();
			i += k;
		}
	}
}

// Value of amplitude i of the register (LDS-resident or pooled).
template <typename T>
__device__ __forceinline__ T amp_at(const T *tile, const T *st, bool resident,
				    u64 i)
{
	return resident ? tile[lds_pos<T>((int)i)] : st[i];
}

template <typename T>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(4))) void k_batch(
	const BReg *__restrict__ regs, const unsigned *__restrict__ order,
	unsigned nregs, const BPass *__restrict__ passes,
	const BGate *__restrict__ gates, const double *__restrict__ rolls,
	T *__restrict__ pool, signed char *__restrict__ outcomes,
	double *__restrict__ p1s, double2 *__restrict__ dens,
	double *__restrict__ probs, unsigned *__restrict__ counter,
	const double *__restrict__ csums)
{
	constexpr int LT = LDS_BYTES / (int)sizeof(T);
	constexpr int LB = sizeof(T) == 16 ? 12 : 13;
	__shared__ T tile[LT];
	__shared__ double red[BW];
	__shared__ unsigned item;
	__shared__ BPass ps;
	__shared__ BGate gsh[GCH];

	for (;;) {
		if (threadIdx.x == 0)
			item = atomicAdd(counter, 1u);
		__syncthreads();
		const unsigned k = item;
		__syncthreads();
		if (k >= nregs)
			break;
		const BReg R = regs[order[k]];
		const unsigned n = R.nq;
		const u64 len = 1ull << n;
		const bool resident = len <= (u64)LT;
		T *st = pool + R.off;

		if (resident) {
			for (int s = threadIdx.x; s < (int)len; s += BT) {
				T v = czero<T>();
				if (s == 0)
					v.x = 1;
				tile[lds_pos<T>(s)] = v;
			}
			__syncthreads();
			for (unsigned pi = R.pass_begin; pi < R.pass_end; ++pi) {
				if (threadIdx.x == 0)
					ps = passes[pi];
				__syncthreads();
				tile_pass(tile, (int)n, ps, 0, gates, gsh);
			}
		} else if (!(R.flags & BF_SPLIT)) {
			for (u64 i = threadIdx.x; i < len; i += BT) {
				T v = czero<T>();
				if (i == 0)
					v.x = 1;
				st[i] = v;
			}
			__syncthreads();
			for (unsigned pi = R.pass_begin; pi < R.pass_end; ++pi) {
				if (threadIdx.x == 0)
					ps = passes[pi];
				__syncthreads();
				const u64 ntiles = len >> LB;
				for (u64 tl = 0; tl < ntiles; ++tl) {
					u64 base = tl;
					for (int b = 0; b < (int)ps.npos; ++b)
						base = ins0(base, ps.pos[b]);
					const TileIndex<LT / BT> ix(base, ps);
#pragma unroll
					for (int k = 0; k < LT / BT; ++k)
						tile[lds_pos<T>((int)threadIdx.x + k * BT)] = st[ix[k]];
					__syncthreads();
					tile_pass(tile, LB, ps, base, gates, gsh);
#pragma unroll
					for (int k = 0; k < LT / BT; ++k)
						st[ix[k]] = tile[lds_pos<T>((int)threadIdx.x + k * BT)];
					__syncthreads();
				}
			}
		}

		// squared norm of the final state (marginals and densities divide
		// by it; the measurement's P1 is a ratio of its own sums)
		const bool split = (R.flags & BF_SPLIT) != 0;
		const double *cs = csums + R.coff * CREC;   // split: chunk records
		const u64 nch = split ? len >> CHUNK_BITS : 0;
		double norm2 = 1.0;
		if (R.flags & (BF_PROBS | BF_DENSITY)) {
			double acc = 0.0;
			if (split)
				for (u64 c = threadIdx.x; c < nch; c += BT)
					acc += cs[c * CREC];
			else
				for (u64 i = threadIdx.x; i < len; i += BT)
					acc += cabs2d(amp_at(tile, st, resident, i));
			norm2 = bsum(acc, red);
		}

		if ((R.flags & BF_PROBS) && split) {
			for (unsigned q = 0; q < n; ++q) {
				double acc1 = 0.0;
				for (u64 c = threadIdx.x; c < nch; c += BT) {
					if (q < (unsigned)CHUNK_BITS)
						acc1 += cs[c * CREC + 1 + q];
					else if ((c >> (q - CHUNK_BITS)) & 1)
						acc1 += cs[c * CREC];
				}
				const double s1 = bsum(acc1, red);
				if (threadIdx.x == 0)
					probs[R.roll_off + q] = s1 / norm2;
			}
		} else if (R.flags & BF_PROBS) {
			for (unsigned q = 0; q < n; ++q) {
				const u64 half = len >> 1;
				double acc1 = 0.0;
				for (u64 i = threadIdx.x; i < half; i += BT)
					acc1 += cabs2d(amp_at(tile, st, resident,
							      ins0(i, q) | (1ull << q)));
				const double s1 = bsum(acc1, red);
				if (threadIdx.x == 0)
					probs[R.roll_off + q] = s1 / norm2;
			}
		}

		if ((R.flags & BF_EXPORT) && resident) {
			for (u64 i = threadIdx.x; i < len; i += BT)
				st[i] = tile[lds_pos<T>((int)i)];
			__syncthreads();
		}

		if (R.flags & BF_DENSITY) {
			// rho[i][j] = sum_t psi(i,t) conj(psi(j,t)) / norm^2 over the
			// kept qubits (ascending = reduced index bit order)
			unsigned char kp[8];
			int a = 0;
			for (unsigned q = 0; q < n; ++q)
				if ((R.keep >> q) & 1)
					kp[a++] = (unsigned char)q;
			const int m = (int)n - a;
			const int dim = 1 << a;
			const u64 T_ = 1ull << m;
			if (resident) {
				for (int e = threadIdx.x; e < dim * dim; e += BT) {
					const u64 ki = deposit((u64)(e / dim), kp, a);
					const u64 kj = deposit((u64)(e % dim), kp, a);
					double2 s = make_double2(0.0, 0.0);
					for (u64 t = 0; t < T_; ++t) {
						u64 td = t;
						for (int b = 0; b < a; ++b)
							td = ins0(td, kp[b]);
						const T x = tile[lds_pos<T>((int)(ki | td))];
						const T y = tile[lds_pos<T>((int)(kj | td))];
						const double xr = x.x, xi = x.y, yr = y.x, yi = y.y;
						s.x = fma(xr, yr, fma(xi, yi, s.x));
						s.y = fma(xi, yr, fma(-xr, yi, s.y));
					}
					dens[R.dens_off + e] = make_double2(s.x / norm2, s.y / norm2);
				}
			} else {
				// a pooled state: its t-vectors (the dim kept-qubit
				// amplitudes of one value t of the other bits) gathered
				// into LDS a chunk at a time and read there by every entry
				// (one thread per entry reading the pool took ~1.7 ms for
				// a 14-qubit register, latency-bound); each entry still
				// sums over t in ascending order, with the same operations.
				// Up to 8 entries per thread (dim <= 64), four at a time:
				// the state is gathered once per four
				constexpr int EPR = 4;
				const int tc = LT / dim;         // t-vectors per chunk
				for (int e0 = 0; e0 < dim * dim; e0 += EPR * BT) {
					double2 acc[EPR];
#pragma unroll
					for (int k = 0; k < EPR; ++k)
						acc[k] = make_double2(0.0, 0.0);
					for (u64 t0 = 0; t0 < T_; t0 += (u64)tc) {
						const int nt = T_ - t0 < (u64)tc ? (int)(T_ - t0) : tc;
						for (int s = threadIdx.x; s < nt * dim; s += BT) {
							u64 td = t0 + (u64)(s / dim);
							for (int b = 0; b < a; ++b)
								td = ins0(td, kp[b]);
							const u64 ix = deposit((u64)(s % dim), kp, a) | td;
							tile[lds_pos<T>(s)] =
								(ix & R.zmask) ? czero<T>() : st[ix];
						}
						__syncthreads();
						// a thread's entries share their column j (BT is a
						// multiple of dim): y is read once per t
						const int j = (int)threadIdx.x & (dim - 1);
						for (int tt = 0; tt < nt; ++tt) {
							const T y = tile[lds_pos<T>(tt * dim + j)];
							const double yr = y.x, yi = y.y;
#pragma unroll
							for (int k = 0; k < EPR; ++k) {
								const int e = e0 + (int)threadIdx.x + k * BT;
								if (e < dim * dim) {
									const T x = tile[lds_pos<T>(tt * dim + (e >> a))];
									const double xr = x.x, xi = x.y;
									acc[k].x = fma(xr, yr, fma(xi, yi, acc[k].x));
									acc[k].y = fma(xi, yr, fma(-xr, yi, acc[k].y));
								}
							}
						}
						__syncthreads();
					}
#pragma unroll
					for (int k = 0; k < EPR; ++k) {
						const int e = e0 + (int)threadIdx.x + k * BT;
						if (e < dim * dim)
							dens[R.dens_off + e] = make_double2(
								acc[k].x / norm2, acc[k].y / norm2);
					}
				}
			}
			__syncthreads();
		}

		if ((R.flags & BF_MEASURE) && split) {
			// steps above the chunk bits from the chunk totals, then the
			// chosen chunk in LDS for the last CHUNK_BITS steps
			u64 prefix = 0;
			unsigned step = 0;
			for (; step < n - CHUNK_BITS; ++step) {
				const unsigned w = n - 1 - step - CHUNK_BITS;
				const u64 h = 1ull << w, b0 = prefix << (w + 1);
				double lo = 0.0, hi = 0.0;
				for (u64 i = threadIdx.x; i < h; i += BT) {
					lo += cs[(b0 + i) * CREC];
					hi += cs[(b0 + h + i) * CREC];
				}
				const double slo = bsum(lo, red);
				const double shi = bsum(hi, red);
				const double p1 = shi / (slo + shi);
				const bool v = p1 > rolls[R.roll_off + step];
				if (threadIdx.x == 0) {
					outcomes[R.roll_off + step] = v ? 1 : 0;
					p1s[R.roll_off + step] = p1;
				}
				prefix = (prefix << 1) | (v ? 1u : 0u);
			}
			const T *chunk = st + (prefix << CHUNK_BITS);
			for (int i = threadIdx.x; i < CHUNK; i += BT)
				tile[lds_pos<T>(i)] = (((prefix << CHUNK_BITS) | (u64)i) & R.zmask)
							      ? czero<T>() : chunk[i];
			__syncthreads();
			u64 lp = 0;
			for (; step < n; ++step) {
				const unsigned q = n - 1 - step;
				const int start = (int)(lp << (q + 1)), half = 1 << q;
				double lo = 0.0, hi = 0.0;
				for (int i = threadIdx.x; i < half; i += BT) {
					lo += cabs2d(tile[lds_pos<T>(start + i)]);
					hi += cabs2d(tile[lds_pos<T>(start + half + i)]);
				}
				const double slo = bsum(lo, red);
				const double shi = bsum(hi, red);
				const double p1 = shi / (slo + shi);
				const bool v = p1 > rolls[R.roll_off + step];
				if (threadIdx.x == 0) {
					outcomes[R.roll_off + step] = v ? 1 : 0;
					p1s[R.roll_off + step] = p1;
				}
				lp = (lp << 1) | (v ? 1u : 0u);
			}
		} else if (R.flags & BF_MEASURE) {
			u64 prefix = 0;
			for (unsigned step = 0; step < n; ++step) {
				const unsigned q = n - 1 - step;
				const u64 start = prefix << (q + 1), half = 1ull << q;
				double lo = 0.0, hi = 0.0;
				// four index pairs per thread per step: more loads in
				// flight for pooled registers
				for (u64 i0 = threadIdx.x; i0 < half; i0 += 4 * BT) {
					T x[4], y[4];
#pragma unroll
					for (int u = 0; u < 4; ++u) {
						const u64 i = i0 + (u64)u * BT;
						x[u] = y[u] = czero<T>();
						if (i < half) {
							x[u] = amp_at(tile, st, resident, start + i);
							y[u] = amp_at(tile, st, resident, start + half + i);
						}
					}
#pragma unroll
					for (int u = 0; u < 4; ++u) {
						lo += cabs2d(x[u]);
						hi += cabs2d(y[u]);
					}
				}
				const double slo = bsum(lo, red);
				const double shi = bsum(hi, red);
				const double p1 = shi / (slo + shi);
				const bool v = p1 > rolls[R.roll_off + step];
				if (threadIdx.x == 0) {
					outcomes[R.roll_off + step] = v ? 1 : 0;
					p1s[R.roll_off + step] = p1;
				}
				prefix = (prefix << 1) | (v ? 1u : 0u);
			}
		}
		__syncthreads();
	}
}


struct ChunkEntry {
	u64 first;      // first item (chunk) of this register
	u64 off;        // its amplitude offset in the pool
	u64 coff;       // its first chunk record
	u64 zmask;      // index bits outside its final support (zeros, not read)
};

template <typename T>
__global__ __launch_bounds__(CT) void k_chunk_sums(
	const T *__restrict__ pool, const ChunkEntry *__restrict__ ents,
	unsigned nents, u64 nitems, double *__restrict__ sums)
{
	__shared__ double red[CT / 64][6];
	__shared__ unsigned ent_sh;
	const unsigned lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	for (u64 it = blockIdx.x; it < nitems; it += gridDim.x) {
		if (threadIdx.x == 0) {
			unsigned lo = 0, hi = nents - 1;
			while (lo < hi) {
				const unsigned mid = (lo + hi + 1) / 2;
				if (ents[mid].first <= it)
					lo = mid;
				else
					hi = mid - 1;
			}
			ent_sh = lo;
		}
		__syncthreads();
		const unsigned e = __builtin_amdgcn_readfirstlane(ent_sh);
		const u64 c = it - ents[e].first;
		const u64 zm = ents[e].zmask;
		if ((c << CHUNK_BITS) & zm) {
			// a chunk outside the support: zeros, not read (the pool
			// may not hold them)
			if (threadIdx.x < CREC)
				sums[(ents[e].coff + c) * CREC + threadIdx.x] = 0.0;
			__syncthreads();   // ent_sh is rewritten next
			continue;
		}
		const unsigned zl = (unsigned)zm & (CHUNK - 1);
		const T *src = pool + ents[e].off + (c << CHUNK_BITS);
		// amplitude k * CT + threadIdx.x: bits 0..7 from the thread, bits
		// 8..11 from k
		double tot = 0.0, hb[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
		for (int k = 0; k < CHUNK / CT; ++k) {
			const unsigned j = k * CT + threadIdx.x;
			const double a = (j & zl) ? 0.0 : cabs2d(src[j]);
			tot += a;
#pragma unroll
			for (int b = 0; b < 4; ++b)
				if ((k >> b) & 1)
					hb[b] += a;
		}
		// block sums: the total, bits 8..11, and bits 0..7 (a thread's
		// total where its own bit is set)
		double v[5] = {tot, hb[0], hb[1], hb[2], hb[3]};
		double lowb[8];
#pragma unroll
		for (int b = 0; b < 8; ++b)
			lowb[b] = ((threadIdx.x >> b) & 1) ? tot : 0.0;
		double *rec = sums + (ents[e].coff + c) * CREC;
#pragma unroll
		for (int j = 0; j < 5; ++j) {
			const double t = wave_sum(v[j]);
			if (lane == 0)
				red[w][j] = t;
		}
		__syncthreads();
		if (threadIdx.x == 0) {
			double t[5] = {0, 0, 0, 0, 0};
			for (int ww = 0; ww < CT / 64; ++ww)
				for (int j = 0; j < 5; ++j)
					t[j] += red[ww][j];
			rec[0] = t[0];
			for (int b = 0; b < 4; ++b)
				rec[1 + 8 + b] = t[1 + b];
		}
		__syncthreads();
#pragma unroll
		for (int b = 0; b < 8; ++b) {
			const double t = wave_sum(lowb[b]);
			if (lane == 0)
				red[w][b & 3] = t;
			if ((b & 3) == 3) {
				__syncthreads();
				if (threadIdx.x == 0)
					for (int j = 0; j < 4; ++j) {
						double x = 0.0;
						for (int ww = 0; ww < CT / 64; ++ww)
							x += red[ww][j];
						rec[1 + (b - 3) + j] = x;
					}
				__syncthreads();
			}
		}
	}
}

// Zeros over the first `len` amplitudes of the pool (before the rounds).
template <typename T>
__global__ __launch_bounds__(256) void k_zero_pool(T *__restrict__ pool, u64 len)
{
	for (u64 i = (u64)blockIdx.x * 256 + threadIdx.x; i < len;
	     i += (u64)gridDim.x * 256)
		st_stream(pool + i, czero<T>());
}

// NaN over the whole pool (the "poison_pool" setting: a run that read an
// amplitude no pass wrote would carry the NaN into its results)
template <typename T>
__global__ __launch_bounds__(256) void k_poison_pool(T *__restrict__ pool, u64 len)
{
	const u64 i = (u64)blockIdx.x * 256 + threadIdx.x;
	if (i < len)
		pool[i] = mk<T>((typename real_of<T>::type)__builtin_nan(""),
				(typename real_of<T>::type)__builtin_nan(""));
}

__global__ void k_reset_counter(unsigned *c)
{
	if (threadIdx.x == 0)
		*c = 0u;
}

} // namespace

// ===========================================================================

struct dk_batch {
	int dtype;
	std::vector<BReg> regs;
	std::vector<BPass> passes;
	std::vector<BGate> gates;
	std::vector<double> rolls;
	unsigned dens_total = 0;
	int export_states = 0;
	int use_graph = 0;
	int pass_qubits = 0; // 0 = tile size
	int probs = 0;       // marginal probabilities of every register
	int split_qubits = 14; // pooled registers this large run their passes
			       // as grid-wide rounds (0 = never)
	int poison_pool = 0;   // NaN over the pool before each run (unless
			       // states are exported, which needs it cleared)
	std::vector<BGate> orig;          // every register's gates as added
	std::vector<size_t> orig_begin;   // first gate of register i in orig
	std::vector<dkp::PassEntry> rounds; // every round's entries, in order
	std::vector<dkp::PassOp> pops;    // ops of the split registers' passes
	std::vector<size_t> round_off;    // first entry of round r
	std::vector<u64> round_items;     // tiles of round r
	dkp::PassEntry *d_rounds = nullptr;
	dkp::PassOp *d_pops = nullptr;
	std::vector<ChunkEntry> chunks;   // split registers' chunk-sum entries
	u64 chunk_items = 0;
	ChunkEntry *d_chunks = nullptr;
	double *d_csums = nullptr;
	// device side; buffers are kept (and grown) across uploads
	bool uploaded = false;
	char *dbuf = nullptr;
	u64 dbuf_bytes = 0;
	void *pool = nullptr;
	u64 pool_bytes = 0;
	u64 pool_amps = 0;   // amplitudes of the pool the registers use
	char *hbuf = nullptr;
	u64 hbuf_bytes = 0;
	BReg *d_regs = nullptr;
	unsigned *d_order = nullptr;
	BPass *d_passes = nullptr;
	BGate *d_gates = nullptr;
	double *d_rolls = nullptr;
	signed char *d_out = nullptr;
	double *d_p1 = nullptr;
	double *d_probs = nullptr;
	double2 *d_dens = nullptr;
	unsigned *d_counter = nullptr;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	signed char *h_out = nullptr;  // pinned, so graph replays can copy
	double *h_p1 = nullptr;
	double *h_probs = nullptr;
	double2 *h_dens = nullptr;
	bool ran = false;
};

static int batch_lb(int dtype)
{
	return dtype == DK_C64 ? 13 : 12;
}

// The uploaded layout (and a captured graph) no longer match the batch.
static void batch_invalidate(dk_batch *b)
{
	if (b->exec)
		(void)hipGraphExecDestroy(b->exec);
	if (b->graph)
		(void)hipGraphDestroy(b->graph);
	b->exec = nullptr;
	b->graph = nullptr;
	b->uploaded = false;
}

static void batch_release_device(dk_batch *b)
{
	batch_invalidate(b);
	if (b->dbuf)
		dk_device_free(b->dbuf, b->dbuf_bytes);
	if (b->pool)
		dk_device_free(b->pool, b->pool_bytes);
	if (b->hbuf)
		(void)hipHostFree(b->hbuf);
	b->hbuf = nullptr;
	b->hbuf_bytes = 0;
	b->dbuf = nullptr;
	b->dbuf_bytes = 0;
	b->pool = nullptr;
	b->pool_bytes = 0;
}

extern "C" int dk_batch_new(int dtype, dk_batch **out)
{
	if (!out || (dtype != DK_C128 && dtype != DK_C64)) {
		dk_set_error("unknown amplitude type");
		return DK_E_ARG;
	}
	dk_batch *b = new dk_batch;
	b->dtype = dtype;
	*out = b;
	return DK_OK;
}

extern "C" void dk_batch_free(dk_batch *b)
{
	if (!b)
		return;
	batch_release_device(b);
	delete b;
}

extern "C" int dk_batch_set(dk_batch *b, const char *key, int64_t value)
{
	if (!b || !key) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	if (!strcmp(key, "export_states"))
		b->export_states = value != 0;
	else if (!strcmp(key, "use_graph"))
		b->use_graph = value != 0;
	else if (!strcmp(key, "pass_qubits"))
		b->pass_qubits = (int)value;
	else if (!strcmp(key, "probs"))
		b->probs = value != 0;
	else if (!strcmp(key, "split_qubits"))
		b->split_qubits = (int)std::max<int64_t>(0, value);
	else if (!strcmp(key, "poison_pool"))
		b->poison_pool = value != 0;
	else {
		dk_set_error("unknown batch setting %s", key);
		return DK_E_ARG;
	}
	batch_invalidate(b);
	// results of an earlier run were made under the old settings
	b->ran = false;
	return DK_OK;
}

// Drops every register; device and pinned buffers stay for the next fill.
extern "C" int dk_batch_clear(dk_batch *b)
{
	if (!b) {
		dk_set_error("NULL batch");
		return DK_E_ARG;
	}
	b->regs.clear();
	b->orig.clear();
	b->orig_begin.clear();
	b->passes.clear();
	b->gates.clear();
	b->rolls.clear();
	b->dens_total = 0;
	b->ran = false;
	batch_invalidate(b);
	return DK_OK;
}

// Gate record of a 1-qubit gate (registry_apply's argument meaning).
extern "C" int dk_batch_add(dk_batch *b, unsigned nq, size_t num_gates,
			    const double *matrices, const unsigned *targets,
			    const uint64_t *ctrl_masks, const uint64_t *actrl_masks,
			    const double *rolls, uint64_t density_keep,
			    size_t *index)
{
	if (!b || nq < 1 || nq > 26 || (num_gates && (!matrices || !targets))) {
		dk_set_error("batched registers hold 1..26 qubits");
		return DK_E_ARG;
	}
	const u64 valid = (1ull << nq) - 1;
	if ((density_keep & ~valid) || __builtin_popcountll(density_keep) > 6) {
		dk_set_error("density keeps at most 6 of the register's qubits");
		return DK_E_ARG;
	}
	if (rolls)
		for (unsigned i = 0; i < nq; ++i)
			if (rolls[i] < 0.0 || rolls[i] >= 1.0) {  // NaN passes
				dk_set_error("roll not in interval [0, 1)!");
				return DK_E_ROLL;
			}
	std::vector<BGate> gl(num_gates);
	for (size_t g = 0; g < num_gates; ++g) {
		BGate &x = gl[g];
		memcpy(x.m, matrices + 8 * g, sizeof(x.m));
		x.t = targets[g];
		x.cm = ctrl_masks ? ctrl_masks[g] : 0;
		x.am = actrl_masks ? actrl_masks[g] : 0;
		x.cin = x.ain = x.sb = 0;
		if (x.t >= nq || ((x.cm | x.am) & ~valid) || (x.cm & x.am) ||
		    (((x.cm | x.am) >> x.t) & 1)) {
			dk_set_error("invalid gate %zu for a %u-qubit register", g, nq);
			return DK_E_ARG;
		}
	}
	BReg R;
	memset(&R, 0, sizeof(R));
	R.nq = nq;
	R.keep = density_keep;
	R.roll_off = (unsigned)b->rolls.size();
	R.flags = (rolls ? BF_MEASURE : 0) | (density_keep ? BF_DENSITY : 0);
	R.dens_off = b->dens_total;
	if (density_keep) {
		const unsigned d = 1u << __builtin_popcountll(density_keep);
		b->dens_total += d * d;
	}
	for (unsigned i = 0; i < nq; ++i)
		b->rolls.push_back(rolls ? rolls[i] : 0.0);
	// passes: runs of gates whose targets fit one tile
	const int lb = batch_lb(b->dtype);
	const int tb = (int)std::min<unsigned>(nq, (unsigned)lb);
	int budget = b->pass_qubits > 0 ? std::min(b->pass_qubits, tb) : tb;
	R.pass_begin = (unsigned)b->passes.size();
	size_t g = 0;
	do {
		BPass P;
		memset(&P, 0, sizeof(P));
		P.gate_begin = (unsigned)b->gates.size();
		u64 tmask = 0;
		while (g < num_gates) {
			const u64 nm = tmask | (1ull << gl[g].t);
			if (__builtin_popcountll(nm) > budget)
				break;
			tmask = nm;
			b->gates.push_back(gl[g]);
			++g;
		}
		P.gate_end = (unsigned)b->gates.size();
		// positions: the targets plus the lowest other qubits
		u64 pm = tmask;
		for (unsigned q = 0; q < nq && __builtin_popcountll(pm) < tb; ++q)
			pm |= 1ull << q;
		int c = 0;
		for (unsigned q = 0; q < nq; ++q)
			if ((pm >> q) & 1) {
				P.slot_of[q] = (unsigned char)c;
				P.pos[c++] = (unsigned char)q;
			}
		P.npos = (unsigned)c;
		P.pmask = pm;
		// split each gate's controls at the tile; target slot
		for (unsigned gi = P.gate_begin; gi < P.gate_end; ++gi) {
			BGate &x = b->gates[gi];
			for (u64 m = x.cm & pm; m; m &= m - 1)
				x.cin |= 1u << P.slot_of[__builtin_ctzll(m)];
			for (u64 m = x.am & pm; m; m &= m - 1)
				x.ain |= 1u << P.slot_of[__builtin_ctzll(m)];
			x.cm &= ~pm;
			x.am &= ~pm;
			x.sb = P.slot_of[x.t];
		}
		b->passes.push_back(P);
	} while (g < num_gates);
	R.pass_end = (unsigned)b->passes.size();
	if (index)
		*index = b->regs.size();
	b->orig_begin.push_back(b->orig.size());
	b->orig.insert(b->orig.end(), gl.begin(), gl.end());
	b->regs.push_back(R);
	batch_invalidate(b);
	b->ran = false;
	return DK_OK;
}

static int batch_upload(dk_batch *b)
{
	const int lb = batch_lb(b->dtype);
	const size_t asz = dk_amp_size(b->dtype);
	// pool slots: registers larger than a tile, and every register when the
	// final states are exported
	u64 pool_amps = 0;
	for (auto &R : b->regs) {
		const bool split = b->split_qubits > 0 && R.nq > (unsigned)lb &&
				   R.nq >= (unsigned)b->split_qubits;
		R.flags = (R.flags & ~(BF_EXPORT | BF_PROBS | BF_SPLIT)) |
			  (b->export_states ? BF_EXPORT : 0) | (b->probs ? BF_PROBS : 0) |
			  (split ? BF_SPLIT : 0);
		R.zmask = 0;   // set below for split registers
		const u64 len = 1ull << R.nq;
		if (R.nq > (unsigned)lb || b->export_states) {
			R.off = pool_amps;
			pool_amps += len;
		} else {
			R.off = 0;
		}
	}
	std::vector<unsigned> order(b->regs.size());
	for (size_t i = 0; i < order.size(); ++i)
		order[i] = (unsigned)i;
	std::stable_sort(order.begin(), order.end(), [&](unsigned x, unsigned y) {
		return b->regs[x].nq > b->regs[y].nq;
	});
	// rounds of the split registers' passes: each register's gates planned
	// as cache-blocked passes; round r holds pass r of every split
	// register with more than r passes
	b->rounds.clear();
	b->pops.clear();
	b->round_off.clear();
	b->round_items.clear();
	{
		std::vector<std::vector<dkp::PassDesc>> descs(b->regs.size());
		std::vector<std::vector<uint32_t>> obeg(b->regs.size());
		std::vector<double> mats;
		std::vector<unsigned> tg;
		std::vector<u64> cm, am;
		size_t maxp = 0;
		for (size_t i = 0; i < b->regs.size(); ++i) {
			const BReg &R = b->regs[i];
			if (!(R.flags & BF_SPLIT))
				continue;
			const size_t g0 = b->orig_begin[i];
			const size_t g1 = i + 1 < b->regs.size() ? b->orig_begin[i + 1]
								 : b->orig.size();
			mats.resize(8 * (g1 - g0));
			tg.resize(g1 - g0);
			cm.resize(g1 - g0);
			am.resize(g1 - g0);
			for (size_t g = g0; g < g1; ++g) {
				const BGate &x = b->orig[g];
				memcpy(&mats[8 * (g - g0)], x.m, sizeof(x.m));
				tg[g - g0] = x.t;
				cm[g - g0] = x.cm;
				am[g - g0] = x.am;
			}
			int rc = dk_plan_register_passes(R.nq, g1 - g0, mats.data(), tg.data(),
							 cm.data(), am.data(), descs[i],
							 b->pops, obeg[i]);
			if (rc)
				return rc;
			if (descs[i].empty()) {
				// no gates: one empty pass still writes |0...0>
				dkp::PassDesc d;
				memset(&d, 0, sizeof(d));
				for (int h = 0; h < dkp::HIGH_BITS; ++h)
					d.hpos[h] = d.hsorted[h] = (uint8_t)(dkp::LANE_BITS + h);
				descs[i].push_back(d);
				obeg[i].push_back((uint32_t)b->pops.size());
			}
			maxp = std::max(maxp, descs[i].size());
		}
		// index bits each register's state may have set so far: every
		// register starts from |0...0>, and a pass widens it by its tile
		std::vector<u64> support(b->regs.size(), 0);
		for (size_t r = 0; r < maxp; ++r) {
			b->round_off.push_back(b->rounds.size());
			u64 items = 0;
			for (unsigned i : order) {
				const BReg &R = b->regs[i];
				if (!(R.flags & BF_SPLIT) || r >= descs[i].size())
					continue;
				dkp::PassEntry e;
				memset(&e, 0, sizeof(e));
				e.first = items;
				e.off = R.off;
				e.pd = descs[i][r];
				// tiles outside the support are zeros: skipped, neither
				// read nor written (the pool is not cleared, so no
				// later reader may read them either: the final support
				// goes to the chunk sums and the persistent launch).  A
				// finite first pass therefore runs only the tile holding
				// amplitude 0 (tile 0); inside a tile that runs, the
				// amplitudes outside the support are taken as zero
				// without a read
				const u64 valid = (1ull << R.nq) - 1;
				e.pd.zmask = e.pd.finite ? (~support[i] & valid) : 0;
				e.pd.zwrite = 0;
				e.pd.kmask = ~support[i] & valid;
				u64 tmask = (1ull << dkp::LANE_BITS) - 1;
				for (int h = 0; h < dkp::HIGH_BITS; ++h)
					tmask |= 1ull << e.pd.hpos[h];
				// (after a non-finite entry nothing is known any more)
				support[i] = e.pd.finite ? (support[i] | tmask) : ~0ull;
				e.op_begin = obeg[i][r];
				e.nq = R.nq;
				b->rounds.push_back(e);
				items += r == 0 && e.pd.finite
						 ? 1
						 : 1ull << (R.nq - (unsigned)dkp::PASS_TILE_BITS);
			}
			b->round_items.push_back(items);
		}
		for (size_t i = 0; i < b->regs.size(); ++i)
			if (b->regs[i].flags & BF_SPLIT)
				b->regs[i].zmask = ~support[i] & ((1ull << b->regs[i].nq) - 1);
	}
	// chunk sums of every split register's final state
	b->chunks.clear();
	b->chunk_items = 0;
	for (unsigned i : order) {
		BReg &R = b->regs[i];
		R.coff = 0;
		if (!(R.flags & BF_SPLIT))
			continue;
		R.coff = b->chunk_items;
		b->chunks.push_back(ChunkEntry{b->chunk_items, R.off, R.coff, R.zmask});
		b->chunk_items += 1ull << (R.nq - (unsigned)CHUNK_BITS);
	}
	auto al = [](u64 x) { return (x + 255) & ~255ull; };
	const u64 nr = b->regs.size(), nroll = b->rolls.size();
	u64 o_regs = 0, o_order = al(o_regs + nr * sizeof(BReg));
	u64 o_passes = al(o_order + nr * sizeof(unsigned));
	u64 o_gates = al(o_passes + b->passes.size() * sizeof(BPass));
	u64 o_rolls = al(o_gates + std::max<size_t>(1, b->gates.size()) * sizeof(BGate));
	u64 o_out = al(o_rolls + std::max<u64>(1, nroll) * sizeof(double));
	u64 o_p1 = al(o_out + std::max<u64>(1, nroll));
	u64 o_probs = al(o_p1 + std::max<u64>(1, nroll) * sizeof(double));
	u64 o_dens = al(o_probs + std::max<u64>(1, nroll) * sizeof(double));
	u64 o_cnt = al(o_dens + std::max<u64>(1, b->dens_total) * sizeof(double2));
	u64 o_rounds = al(o_cnt + 256);
	u64 o_pops = al(o_rounds + std::max<size_t>(1, b->rounds.size()) *
				   sizeof(dkp::PassEntry));
	u64 o_chunks = al(o_pops + std::max<size_t>(1, b->pops.size()) * sizeof(dkp::PassOp));
	u64 o_csums = al(o_chunks + std::max<size_t>(1, b->chunks.size()) * sizeof(ChunkEntry));
	const u64 total = o_csums + std::max<u64>(1, b->chunk_items) * CREC * sizeof(double);
	if (b->dbuf_bytes < total) {
		if (b->dbuf)
			dk_device_free(b->dbuf, b->dbuf_bytes);
		b->dbuf = nullptr;
		b->dbuf_bytes = 0;
		int rc = dk_device_alloc(total, (void **)&b->dbuf);
		if (rc)
			return rc;
		b->dbuf_bytes = total;
	}
	b->pool_amps = pool_amps;
	const u64 pool_need = std::max<u64>(256, pool_amps * asz);
	if (b->pool_bytes < pool_need) {
		if (b->pool)
			dk_device_free(b->pool, b->pool_bytes);
		b->pool = nullptr;
		b->pool_bytes = 0;
		int rc = dk_device_alloc(pool_need, &b->pool);
		if (rc)
			return rc;
		b->pool_bytes = pool_need;
	}
	b->d_regs = (BReg *)(b->dbuf + o_regs);
	b->d_order = (unsigned *)(b->dbuf + o_order);
	b->d_passes = (BPass *)(b->dbuf + o_passes);
	b->d_gates = (BGate *)(b->dbuf + o_gates);
	b->d_rolls = (double *)(b->dbuf + o_rolls);
	b->d_out = (signed char *)(b->dbuf + o_out);
	b->d_p1 = (double *)(b->dbuf + o_p1);
	b->d_probs = (double *)(b->dbuf + o_probs);
	b->d_dens = (double2 *)(b->dbuf + o_dens);
	b->d_counter = (unsigned *)(b->dbuf + o_cnt);
	b->d_rounds = (dkp::PassEntry *)(b->dbuf + o_rounds);
	b->d_pops = (dkp::PassOp *)(b->dbuf + o_pops);
	b->d_chunks = (ChunkEntry *)(b->dbuf + o_chunks);
	b->d_csums = (double *)(b->dbuf + o_csums);
	hipStream_t st = DK_STREAM;
	DK_HIP(hipMemcpyAsync(b->d_regs, b->regs.data(), nr * sizeof(BReg),
			      hipMemcpyHostToDevice, st));
	DK_HIP(hipMemcpyAsync(b->d_order, order.data(), nr * sizeof(unsigned),
			      hipMemcpyHostToDevice, st));
	DK_HIP(hipMemcpyAsync(b->d_passes, b->passes.data(),
			      b->passes.size() * sizeof(BPass),
			      hipMemcpyHostToDevice, st));
	if (!b->gates.empty())
		DK_HIP(hipMemcpyAsync(b->d_gates, b->gates.data(),
				      b->gates.size() * sizeof(BGate),
				      hipMemcpyHostToDevice, st));
	if (nroll)
		DK_HIP(hipMemcpyAsync(b->d_rolls, b->rolls.data(), nroll * sizeof(double),
				      hipMemcpyHostToDevice, st));
	if (!b->rounds.empty())
		DK_HIP(hipMemcpyAsync(b->d_rounds, b->rounds.data(),
				      b->rounds.size() * sizeof(dkp::PassEntry),
				      hipMemcpyHostToDevice, st));
	if (!b->pops.empty())
		DK_HIP(hipMemcpyAsync(b->d_pops, b->pops.data(),
				      b->pops.size() * sizeof(dkp::PassOp),
				      hipMemcpyHostToDevice, st));
	if (!b->chunks.empty())
		DK_HIP(hipMemcpyAsync(b->d_chunks, b->chunks.data(),
				      b->chunks.size() * sizeof(ChunkEntry),
				      hipMemcpyHostToDevice, st));
	DK_HIP(hipStreamSynchronize(st));
	// one pinned block: p1 and probs (doubles), density (double2),
	// outcomes (bytes)
	const u64 h_roll = al(std::max<u64>(1, nroll) * sizeof(double));
	const u64 h_dens = al(std::max<u64>(1, b->dens_total) * sizeof(double2));
	const u64 hb = 2 * h_roll + h_dens + al(std::max<u64>(1, nroll));
	if (b->hbuf_bytes < hb) {
		// grown geometrically from 8 MiB: hipHostFree waits for the whole
		// device, every lane's work included
		const u64 want = std::max<u64>(std::max<u64>(hb, 2 * b->hbuf_bytes),
					       8ull << 20);
		if (b->hbuf)
			(void)hipHostFree(b->hbuf);
		b->hbuf = nullptr;
		b->hbuf_bytes = 0;
		DK_HIP(hipHostMalloc((void **)&b->hbuf, want, hipHostMallocDefault));
		b->hbuf_bytes = want;
	}
	char *hp = b->hbuf;
	memset(hp, 0, hb);
	b->h_p1 = (double *)hp;
	b->h_probs = (double *)(hp + h_roll);
	b->h_dens = (double2 *)(hp + 2 * h_roll);
	b->h_out = (signed char *)(hp + 2 * h_roll + h_dens);
	b->uploaded = true;
	return DK_OK;
}

static int batch_enqueue(dk_batch *b)
{
	hipStream_t st = DK_STREAM;
	const unsigned nregs = (unsigned)b->regs.size();
	const unsigned grid = (unsigned)std::max(1, g_ctx.num_cus) * 2;
	if (!b->round_items.empty() && (b->export_states || b->poison_pool)) {
		// the rounds leave every amplitude outside a register's support
		// unwritten; exported states need those zeros in the pool.  One
		// amplitude per thread over the whole pool (a kernel, not a
		// memset, so that the sequence can be captured in a graph; a
		// capped grid-stride loop streams at two thirds of the rate)
		const unsigned zg = (unsigned)std::min<u64>(
			std::max<u64>(1, (b->pool_amps + 255) / 256), (1ull << 31) / 256);
		const bool zero = b->export_states != 0;
		if (b->dtype == DK_C64)
			hipLaunchKernelGGL(zero ? k_zero_pool<float2> : k_poison_pool<float2>,
					   dim3(zg), dim3(256), 0, st, (float2 *)b->pool,
					   b->pool_amps);
		else
			hipLaunchKernelGGL(zero ? k_zero_pool<double2> : k_poison_pool<double2>,
					   dim3(zg), dim3(256), 0, st, (double2 *)b->pool,
					   b->pool_amps);
		DK_LAUNCHED();
	}
	for (size_t r = 0; r < b->round_items.size(); ++r) {
		const size_t e1 = r + 1 < b->round_off.size() ? b->round_off[r + 1]
							       : b->rounds.size();
		const unsigned nents = (unsigned)(e1 - b->round_off[r]);
		const u64 items = b->round_items[r];
		// one workgroup per tile, up to 2^20 (the kernel loops beyond)
		const unsigned rg = (unsigned)std::min<u64>(items, 1ull << 20);
		DK_TRY(dk_launch_pass_round(b->dtype, b->pool, b->d_rounds + b->round_off[r],
					    nents, b->d_pops, items, r == 0, rg, st));
	}
	if (b->chunk_items) {
		const unsigned cg = (unsigned)std::min<u64>(b->chunk_items, 1ull << 20);
		if (b->dtype == DK_C64)
			hipLaunchKernelGGL(k_chunk_sums<float2>, dim3(cg), dim3(CT), 0, st,
					   (const float2 *)b->pool, b->d_chunks,
					   (unsigned)b->chunks.size(), b->chunk_items, b->d_csums);
		else
			hipLaunchKernelGGL(k_chunk_sums<double2>, dim3(cg), dim3(CT), 0, st,
					   (const double2 *)b->pool, b->d_chunks,
					   (unsigned)b->chunks.size(), b->chunk_items, b->d_csums);
		DK_LAUNCHED();
	}
	hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, st, b->d_counter);
	if (b->dtype == DK_C64)
		hipLaunchKernelGGL(k_batch<float2>, dim3(grid), dim3(BT), 0, st,
				   b->d_regs, b->d_order, nregs, b->d_passes, b->d_gates,
				   b->d_rolls, (float2 *)b->pool, b->d_out, b->d_p1,
				   b->d_dens, b->d_probs, b->d_counter, b->d_csums);
	else
		hipLaunchKernelGGL(k_batch<double2>, dim3(grid), dim3(BT), 0, st,
				   b->d_regs, b->d_order, nregs, b->d_passes, b->d_gates,
				   b->d_rolls, (double2 *)b->pool, b->d_out, b->d_p1,
				   b->d_dens, b->d_probs, b->d_counter, b->d_csums);
	DK_LAUNCHED();
	const u64 nroll = b->rolls.size();
	if (nroll) {
		DK_HIP(hipMemcpyAsync(b->h_out, b->d_out, nroll,
				      hipMemcpyDeviceToHost, st));
		DK_HIP(hipMemcpyAsync(b->h_p1, b->d_p1, nroll * sizeof(double),
				      hipMemcpyDeviceToHost, st));
	}
	if (nroll && b->probs)
		DK_HIP(hipMemcpyAsync(b->h_probs, b->d_probs, nroll * sizeof(double),
				      hipMemcpyDeviceToHost, st));
	if (b->dens_total)
		DK_HIP(hipMemcpyAsync(b->h_dens, b->d_dens,
				      b->dens_total * sizeof(double2),
				      hipMemcpyDeviceToHost, st));
	return DK_OK;
}

// Runs every register of the batch in one persistent launch.  With
// "use_graph" set, the launch sequence is captured once as a HIP graph and
// replayed on later runs.
extern "C" int dk_batch_run(dk_batch *b)
{
	if (!b) {
		dk_set_error("NULL batch");
		return DK_E_ARG;
	}
	int rc = dk_ensure_init();
	if (rc)
		return rc;
	if (b->regs.empty()) {
		b->ran = true;
		return DK_OK;
	}
	if (!b->uploaded) {
		rc = batch_upload(b);
		if (rc)
			return rc;
	}
	hipStream_t st = DK_STREAM;
	if (b->use_graph) {
		if (!b->exec) {
			DK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
			rc = batch_enqueue(b);
			hipGraph_t gr = nullptr;
			hipError_t e = hipStreamEndCapture(st, &gr);
			if (rc)
				return rc;
			if (e != hipSuccess) {
				dk_set_error("graph capture failed: %s", hipGetErrorString(e));
				return DK_E_DEVICE;
			}
			b->graph = gr;
			DK_HIP(hipGraphInstantiate(&b->exec, gr, nullptr, nullptr, 0));
		}
		DK_HIP(hipGraphLaunch(b->exec, st));
	} else {
		rc = batch_enqueue(b);
		if (rc)
			return rc;
	}
	DK_HIP(hipStreamSynchronize(st));
	b->ran = true;
	return DK_OK;
}

extern "C" int dk_batch_size(const dk_batch *b, size_t *n)
{
	if (!b || !n) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	*n = b->regs.size();
	return DK_OK;
}

// Outcomes of register i in the order of registry_measure's list (slot s is
// qubit nq-1-s) and the P1 each was decided on.
extern "C" int dk_batch_outcomes(const dk_batch *b, size_t i, int8_t *results,
				 double *p1)
{
	if (!b || i >= b->regs.size() || !b->ran) {
		dk_set_error("batch not run or index out of range");
		return DK_E_ARG;
	}
	const BReg &R = b->regs[i];
	for (unsigned s = 0; s < R.nq; ++s) {
		if (results)
			results[s] = (R.flags & BF_MEASURE) ? b->h_out[R.roll_off + s] : -1;
		if (p1)
			p1[s] = b->h_p1[R.roll_off + s];
	}
	return DK_OK;
}

// Marginal P(qubit q = 1) of register i's final state, q = 0..nq-1.  Needs
// "probs" set before the run.
extern "C" int dk_batch_probs(const dk_batch *b, size_t i, double *p)
{
	if (!b || i >= b->regs.size() || !b->ran || !p || !b->probs) {
		dk_set_error("batch not run with probs, or bad index");
		return DK_E_ARG;
	}
	const BReg &R = b->regs[i];
	memcpy(p, b->h_probs + R.roll_off, R.nq * sizeof(double));
	return DK_OK;
}

extern "C" int dk_batch_density(const dk_batch *b, size_t i, double *rho)
{
	if (!b || i >= b->regs.size() || !b->ran || !rho) {
		dk_set_error("batch not run or index out of range");
		return DK_E_ARG;
	}
	const BReg &R = b->regs[i];
	if (!(R.flags & BF_DENSITY)) {
		dk_set_error("register %zu has no density request", i);
		return DK_E_ARG;
	}
	const unsigned d = 1u << __builtin_popcountll(R.keep);
	memcpy(rho, b->h_dens + R.dens_off, (size_t)d * d * sizeof(double2));
	return DK_OK;
}

// Final (pre-measurement) state of register i, normalised, as complex128.
// Needs "export_states" set before the run.
extern "C" int dk_batch_state(const dk_batch *b, size_t i, double *values)
{
	if (!b || i >= b->regs.size() || !b->ran || !values || !b->export_states) {
		dk_set_error("batch not run with export_states, or bad index");
		return DK_E_ARG;
	}
	const BReg &R = b->regs[i];
	const u64 len = 1ull << R.nq;
	const size_t asz = dk_amp_size(b->dtype);
	std::vector<char> raw(len * asz);
	DK_HIP(hipMemcpy(raw.data(), (char *)b->pool + R.off * asz, len * asz,
			 hipMemcpyDeviceToHost));
	double n2 = 0.0;
	for (u64 k = 0; k < len; ++k) {
		double re, im;
		if (b->dtype == DK_C64) {
			const float *f = (const float *)raw.data();
			re = f[2 * k];
			im = f[2 * k + 1];
		} else {
			const double *f = (const double *)raw.data();
			re = f[2 * k];
			im = f[2 * k + 1];
		}
		values[2 * k] = re;
		values[2 * k + 1] = im;
		n2 += re * re + im * im;
	}
	const double inv = 1.0 / sqrt(n2);
	for (u64 k = 0; k < 2 * len; ++k)
		values[k] *= inv;
	return DK_OK;
}

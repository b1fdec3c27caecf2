// Dense k-qubit gates (k = 4, 5, 6) on the matrix cores.
//
// An LDS tile holds 64 groups of D = 2^k amplitudes.  The tile is the
// complex matrix Psi (D rows x 64 columns); the gate computes O = M * Psi,
// with M the row-major gate (row = output basis state).  Complex products use
// the four-real-product form
//     Re O = Re M Re Psi - Im M Im Psi,   Im O = Re M Im Psi + Im M Re Psi,
// each a sequence of 16x16x4 MFMAs accumulating in the amplitude's precision
// (FP64: v_mfma_f64_16x16x4f64, FP32: v_mfma_f32_16x16x4f32); the streaming
// form for k = 6 uses three real products instead (see k_dense_stream).
//
// A block owns one tile of 64 * CW groups (64 * CW * D amplitudes: the
// gate's targets plus the 6 + log2(CW) lowest other index bits).  Its waves
// form RT row tiles times CW column sets: wave (rt, cw) computes rows
// 16rt..16rt+15 of the 64 columns of set cw (four 16-column tiles), its A
// fragments in registers.  The tile is small enough that several blocks
// share a CU, so one block's loads overlap another's matrix work; CW > 1
// gives small gates (k = 4, one row tile) more bytes per block.
// Operand layouts (16x16x4, wave64): lane l holds A[l%16][l/16] and
// B[l/16][l%16]; accumulator element r of lane l holds
// D[out_row(l/16, r)][l%16] (see mfma16 below).
// Input scaling by 1/norm is applied to the outputs (the product is linear).
#include <algorithm>
#include <cstring>

#include "dk_device.h"

using namespace dkd;

namespace {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename R> struct mfma16;
// out_row(g, r): row of the 16x16 result held in element r of a lane of row
// group g = lane / 16 (measured on gfx950: FP64 interleaves the groups,
// FP32 keeps four consecutive rows per group).
template <> struct mfma16<double> {
	typedef d4_t acc;
	static __device__ __forceinline__ acc op(double a, double b, acc c)
	{
		return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
	}
	static __device__ __forceinline__ int out_row(int g, int r)
	{
		return g + 4 * r;
	}
};
template <> struct mfma16<float> {
	typedef f4_t acc;
	static __device__ __forceinline__ acc op(float a, float b, acc c)
	{
		return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
	}
	static __device__ __forceinline__ int out_row(int g, int r)
	{
		return 4 * g + r;
	}
};

// One tile: load it, multiply on the matrix cores, store it, leave its
// partial |out|^2 sum.
template <typename T, int K, int CW>
__device__ __forceinline__ void dense_tile(
	const T *in, T *out, const double *__restrict__ nrm,
	const T *__restrict__ M, const TileArgs &ta, double *__restrict__ partials,
	u64 tile, T *lds, double *red)
{
	typedef typename real_of<T>::type R;
	typedef mfma16<R> MF;
	typedef typename MF::acc acc_t;
	constexpr int D = 1 << K;       // 16, 32, 64
	constexpr int KS = D / 4;       // k-steps of the product
	constexpr int RT = D / 16;      // row tiles
	constexpr int NW = RT * CW;     // waves per block
	constexpr int NT = 64 * NW;     // threads per block
	constexpr int TS = 64 * CW * D; // tile: 64 * CW groups of D amplitudes

	const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	const int rt = wv % RT, cw = wv / RT;
	const int li = lane & 15, lk = lane >> 4;
	const u64 base = tile_base(tile, ta);
	constexpr int PER = TS / NT; // 16 slots per thread
	if (base & ta.zmask) {
		// a tile of zeros stays zeros: no read, no product
		if (ta.zwrite) {
			const u64 tp = base | ta.dep_lo[threadIdx.x & 63] |
				       ta.dep_hi[threadIdx.x >> 6];
#pragma unroll
			for (int j = 0; j < PER; ++j)
				st_stream(out + (tp | ta.dep_hi[(j * NT) >> 6]), czero<T>());
		}
		if (threadIdx.x == 0)
			partials[tile] = 0.0;
		return;
	}
	// complex128 k = 4 tiles load straight into LDS (measured on gfx950 at
	// 30 qubits: 6.39 against 6.67 ms; k = 5 and 6 ran 4% and 23% slower)
	constexpr bool LDS_DIRECT = sizeof(T) == 16 && K == 4;

	// Slot s = threadIdx.x + j * NT: its index bits are this thread's part
	// (computed once) OR the part of j * NT (the same for the whole block).
	const u64 tpart = base | ta.dep_lo[threadIdx.x & 63] |
			  ta.dep_hi[threadIdx.x >> 6];
	// issue the tile's loads first, then set up fragments while they fly
	if constexpr (LDS_DIRECT) {
		// Straight into LDS (global_load_lds_dwordx4): a wave fills 64
		// consecutive positions per step, lane l position p = its slot
		// index.  Position p holds slot lds_pos(p) (lds_pos is its own
		// inverse and keeps bits >= 4), whose index differs from slot
		// p's by the deposit of the 4-bit correction x; the deposit is
		// linear, so it is XORed in from the deposits of bits 0..3.
		const u64 d1 = ta.dep_lo[1], d2 = ta.dep_lo[2], d4 = ta.dep_lo[4],
			  d8 = ta.dep_lo[8];
#pragma unroll
		for (int j = 0; j < PER; ++j) {
			const int p = threadIdx.x + j * NT;
			const int x = ((p >> 4) ^ (p >> 8)) & 15;
			const u64 fix = ((x & 1) ? d1 : 0) ^ ((x & 2) ? d2 : 0) ^
					((x & 4) ? d4 : 0) ^ ((x & 8) ? d8 : 0);
			const u64 gi = (tpart | ta.dep_hi[(j * NT) >> 6]) ^ fix;
			__builtin_amdgcn_global_load_lds(
				(__attribute__((address_space(1))) void *)(in + gi),
				(__attribute__((address_space(3))) void *)(lds + (p & ~63)),
				16, 0, 0);
		}
	} else {
#pragma unroll
		for (int j = 0; j < PER; ++j) {
			const int s = threadIdx.x + j * NT;
			lds[lds_pos<T>(s)] = in[tpart | ta.dep_hi[(j * NT) >> 6]];
		}
	}

	// A fragments: this wave's 16 rows of M, all k-steps
	R ar[KS], ai[KS];
#pragma unroll
	for (int ks = 0; ks < KS; ++ks) {
		const T m = M[(size_t)(rt * 16 + li) * D + ks * 4 + lk];
		ar[ks] = m.x;
		ai[ks] = m.y;
	}
	// B slots: this lane's four columns (groups) and its row of each k-step
	int gs[4];
#pragma unroll
	for (int ct = 0; ct < 4; ++ct)
		gs[ct] = group_slot<K>(cw * 64 + ct * 16 + li, ta);
	int ks_slot[KS];
#pragma unroll
	for (int ks = 0; ks < KS; ++ks)
		ks_slot[ks] = ta.colslot[ks * 4 + lk];
	const R inv = (R)(1.0 / nrm[DK_SCAL_NORM]);
	if constexpr (LDS_DIRECT)
		__builtin_amdgcn_s_waitcnt(0); // this wave's LDS loads have landed
	__syncthreads();

	acc_t cr[4], ci[4];
#pragma unroll
	for (int ct = 0; ct < 4; ++ct) {
		cr[ct] = acc_t{0, 0, 0, 0};
		ci[ct] = acc_t{0, 0, 0, 0};
	}
#pragma unroll
	for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
		for (int ct = 0; ct < 4; ++ct) {
			const T b = lds[lds_pos<T>(gs[ct] | ks_slot[ks])];
			cr[ct] = MF::op(ar[ks], b.x, cr[ct]);
			cr[ct] = MF::op(ai[ks], -b.y, cr[ct]);
			ci[ct] = MF::op(ar[ks], b.y, ci[ct]);
			ci[ct] = MF::op(ai[ks], b.x, ci[ct]);
		}
	}
	__syncthreads(); // every wave has read the tile
#pragma unroll
	for (int r = 0; r < 4; ++r) {
		const int os = ta.colslot[rt * 16 + MF::out_row(lk, r)];
#pragma unroll
		for (int ct = 0; ct < 4; ++ct)
			lds[lds_pos<T>(gs[ct] | os)] = mk<T>(cr[ct][r] * inv, ci[ct][r] * inv);
	}
	__syncthreads();
	double acc = 0.0;
#pragma unroll
	for (int j = 0; j < PER; ++j) {
		const int s = threadIdx.x + j * NT;
		const T o = lds[lds_pos<T>(s)];
		st_stream(out + (tpart | ta.dep_hi[(j * NT) >> 6]), o);
		acc += cabs2d(o);
	}
	acc = wave_sum(acc);
	if (lane == 0)
		red[wv] = acc;
	__syncthreads();
	if (threadIdx.x == 0) {
		double t = 0.0;
#pragma unroll
		for (int w = 0; w < NW; ++w)
			t += red[w];
		partials[tile] = t;
	}
}

// LOOP: tiles beyond the grid (launches of more than 2^31 threads); the
// single-tile form keeps the fragment set-up after the tile's loads.
template <typename T, int K, int CW, bool LOOP>
__global__ __launch_bounds__((4 << K) * CW) void k_apply_dense(
	const T *in, T *out, const double *__restrict__ nrm,
	const T *__restrict__ M, TileArgs ta, double *__restrict__ partials,
	u64 ntiles)
{
	__shared__ T lds[64 * CW * (1 << K)];
	__shared__ double red[(1 << K) / 16 * CW];
	if constexpr (LOOP) {
		for (u64 tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
			dense_tile<T, K, CW>(in, out, nrm, M, ta, partials, tile, lds, red);
			__syncthreads();
		}
	} else {
		dense_tile<T, K, CW>(in, out, nrm, M, ta, partials, blockIdx.x, lds, red);
	}
}

// ---------------------------------------------------------------------------
// Streaming form.  Persistent waves; each block first stages the gate in LDS
// in fragment order (entry (rt * KS + ks) * 64 + lane holds A[rt * 16 +
// lane % 16][ks * 4 + lane / 16], so fragment reads are conflict free).  The
// groups (2^k amplitudes differing in the target bits) are numbered by
// their other index bits; a column block is 16 consecutive groups.  A wave
// takes column blocks gw, gw + W, ... (W waves in the grid); it reads its
// block straight from HBM into registers in the B layout of the 16x16x4
// MFMA (lane l: group l % 16, gate column ks * 4 + l / 16), reads the next
// block while the matrix cores work on this one, and stores the D rows of
// results straight from the accumulators.  No tile barriers; a wave never
// reads what another writes, so in == out is safe.  One |out|^2 partial per
// wave.
// ---------------------------------------------------------------------------
struct StreamArgs {
	u64 rowoff[64];            // gate row / column j -> its target index bits
	unsigned char tsorted[8];  // target bit positions, ascending
	// index bits known to be zero in the input: a column block whose base
	// meets zmask holds only zeros, and so does its output (written as
	// zeros when zwrite, else left as it is)
	int zwrite;
	u64 zmask;
};

constexpr int SW = 4;              // waves per streaming block

// One amplitude as one vector load / store, non-temporal.
typedef double dv2 __attribute__((ext_vector_type(2)));
typedef float fv2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ldv(const double2 *p)
{
	const dv2 v = __builtin_nontemporal_load((const dv2 *)p);
	return make_double2(v.x, v.y);
}
__device__ __forceinline__ float2 ldv(const float2 *p)
{
	const fv2 v = __builtin_nontemporal_load((const fv2 *)p);
	return make_float2(v.x, v.y);
}
__device__ __forceinline__ void stv(double2 *p, double2 v)
{
	dv2 x;
	x.x = v.x;
	x.y = v.y;
	__builtin_nontemporal_store(x, (dv2 *)p);
}
__device__ __forceinline__ void stv(float2 *p, float2 v)
{
	fv2 x;
	x.x = v.x;
	x.y = v.y;
	__builtin_nontemporal_store(x, (fv2 *)p);
}

template <typename T, int K, int CT, bool G3, int RH>
__global__ __launch_bounds__(SW * 64) void k_dense_stream(
	const T *in, T *out, const double *__restrict__ nrm,
	const T *__restrict__ M, StreamArgs sa, double *__restrict__ partials,
	u64 ncg)
{
	typedef typename real_of<T>::type R;
	typedef mfma16<R> MF;
	typedef typename MF::acc acc_t;
	constexpr int D = 1 << K, KS = D / 4, RT = D / 16;
	__shared__ T afr[D * D];
	__shared__ u64 roff[D];
	const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
	for (int e = threadIdx.x; e < D * D; e += SW * 64) {
		const int l = e & 63, f = e >> 6;
		const int rt = f / KS, ks = f % KS;
		afr[e] = M[(size_t)(rt * 16 + (l & 15)) * D + ks * 4 + (l >> 4)];
	}
	for (int e = threadIdx.x; e < D; e += SW * 64)
		roff[e] = sa.rowoff[e];
	__syncthreads();
	// group li + 16 c of a column block (c < CT), deposited around the
	// targets: the deposit is linear over disjoint bits
	u64 col[CT];
#pragma unroll
	for (int c = 0; c < CT; ++c) {
		u64 x = (u64)(li + 16 * c);
#pragma unroll
		for (int b = 0; b < K; ++b)
			x = ins0(x, sa.tsorted[b]);
		col[c] = x;
	}
	const R inv = (R)(1.0 / nrm[DK_SCAL_NORM]);
	const u64 W = (u64)gridDim.x * SW;
	const u64 gw = (u64)blockIdx.x * SW + (threadIdx.x >> 6);
	auto block_base = [&](u64 cg) {
		u64 x = cg * (16 * CT);
#pragma unroll
		for (int b = 0; b < K; ++b)
			x = ins0(x, sa.tsorted[b]);
		return x;
	};
	double nacc = 0.0;
	T bn[CT][KS];
	if (gw < ncg && !(block_base(gw) & sa.zmask)) {
		const u64 nb = block_base(gw);
#pragma unroll
		for (int ks = 0; ks < KS; ++ks)
#pragma unroll
			for (int c = 0; c < CT; ++c)
				bn[c][ks] = ldv(in + (nb | col[c] | roff[ks * 4 + lk]));
	}
	for (u64 cg = gw; cg < ncg; cg += W) {
		// the LDS tables are read inside the loop: hoisted out of it, the
		// gate's fragments alone would take 256 registers (k = 6)
		int fl = lane;
		asm volatile("" : "+v"(fl));
		const int flk = fl >> 4;
		T bc[CT][KS];
#pragma unroll
		for (int c = 0; c < CT; ++c)
#pragma unroll
			for (int ks = 0; ks < KS; ++ks)
				bc[c][ks] = bn[c][ks];
		const u64 base = block_base(cg);
		if (cg + W < ncg && !(block_base(cg + W) & sa.zmask)) {
			const u64 nb = block_base(cg + W);
#pragma unroll
			for (int ks = 0; ks < KS; ++ks)
#pragma unroll
				for (int c = 0; c < CT; ++c)
					bn[c][ks] = ldv(in + (nb | col[c] | roff[ks * 4 + flk]));
		}
		if (base & sa.zmask) {
			// a block of zeros stays zeros: no product (the wave's rows
			// j = lk + 4m of its groups, when written)
			if (sa.zwrite) {
#pragma unroll
				for (int m = 0; m < D / 4; ++m)
#pragma unroll
					for (int c = 0; c < CT; ++c)
						stv(out + (base | col[c] | roff[4 * m + flk]),
						    czero<T>());
			}
			continue;
		}
		if constexpr (G3) {
			// three real products per complex one (Gauss):
			//   P = Re A Re B, Q = Im A Im B, S = (Re A + Im A)(Re B + Im B),
			//   Re O = P - Q,  Im O = S - P - Q;
			// RH row tiles at a time keep the three accumulator sets in
			// the register budget of the four-product form (a rolled
			// loop: unrolled, the sets of consecutive row tiles overlap)
#pragma nounroll
			for (int rh = 0; rh < RT; rh += RH) {
				acc_t P[CT][RH], Q[CT][RH], S[CT][RH];
#pragma unroll
				for (int c = 0; c < CT; ++c)
#pragma unroll
					for (int j = 0; j < RH; ++j) {
						P[c][j] = acc_t{0, 0, 0, 0};
						Q[c][j] = acc_t{0, 0, 0, 0};
						S[c][j] = acc_t{0, 0, 0, 0};
					}
#pragma unroll
				for (int ks = 0; ks < KS; ++ks) {
					T a[RH];
					R bs[CT];
#pragma unroll
					for (int j = 0; j < RH; ++j)
						a[j] = afr[((rh + j) * KS + ks) * 64 + fl];
#pragma unroll
					for (int c = 0; c < CT; ++c) {
						// per row tile pass: hoisted out of the rolled
						// loop, the sums would hold KS more registers
						bs[c] = bc[c][ks].x + bc[c][ks].y;
						asm volatile("" : "+v"(bs[c]));
					}
#pragma unroll
					for (int j = 0; j < RH; ++j)
#pragma unroll
						for (int c = 0; c < CT; ++c)
							P[c][j] = MF::op(a[j].x, bc[c][ks].x, P[c][j]);
#pragma unroll
					for (int j = 0; j < RH; ++j)
#pragma unroll
						for (int c = 0; c < CT; ++c)
							Q[c][j] = MF::op(a[j].y, bc[c][ks].y, Q[c][j]);
#pragma unroll
					for (int j = 0; j < RH; ++j)
#pragma unroll
						for (int c = 0; c < CT; ++c)
							S[c][j] = MF::op(a[j].x + a[j].y, bs[c], S[c][j]);
				}
#pragma unroll
				for (int j = 0; j < RH; ++j)
#pragma unroll
					for (int r = 0; r < 4; ++r) {
						const u64 ro = roff[(rh + j) * 16 + MF::out_row(flk, r)];
#pragma unroll
						for (int c = 0; c < CT; ++c) {
							const R re = P[c][j][r] - Q[c][j][r];
							const R im = (S[c][j][r] - P[c][j][r]) - Q[c][j][r];
							const T o = mk<T>(re * inv, im * inv);
							stv(out + (base | col[c] | ro), o);
							nacc += cabs2d(o);
						}
					}
			}
			continue;
		}
		acc_t cr[CT][RT], ci[CT][RT];
#pragma unroll
		for (int c = 0; c < CT; ++c)
#pragma unroll
			for (int rt = 0; rt < RT; ++rt) {
				cr[c][rt] = acc_t{0, 0, 0, 0};
				ci[c][rt] = acc_t{0, 0, 0, 0};
			}
		// per k-step, every accumulator's real-part product first and its
		// imaginary-part product second: an accumulator's two MFMAs are
		// 2 * RT * CT instructions apart (back to back they stall the
		// matrix pipe on the dependency); each element still sums its
		// terms in the same order
#pragma unroll
		for (int ks = 0; ks < KS; ++ks) {
			T a[RT];
#pragma unroll
			for (int rt = 0; rt < RT; ++rt)
				a[rt] = afr[(rt * KS + ks) * 64 + fl];
#pragma unroll
			for (int rt = 0; rt < RT; ++rt)
#pragma unroll
				for (int c = 0; c < CT; ++c) {
					cr[c][rt] = MF::op(a[rt].x, bc[c][ks].x, cr[c][rt]);
					ci[c][rt] = MF::op(a[rt].x, bc[c][ks].y, ci[c][rt]);
				}
#pragma unroll
			for (int rt = 0; rt < RT; ++rt)
#pragma unroll
				for (int c = 0; c < CT; ++c) {
					cr[c][rt] = MF::op(a[rt].y, -bc[c][ks].y, cr[c][rt]);
					ci[c][rt] = MF::op(a[rt].y, bc[c][ks].x, ci[c][rt]);
				}
		}
#pragma unroll
		for (int rt = 0; rt < RT; ++rt)
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				const u64 ro = roff[rt * 16 + MF::out_row(flk, r)];
#pragma unroll
				for (int c = 0; c < CT; ++c) {
					const T o = mk<T>(cr[c][rt][r] * inv, ci[c][rt][r] * inv);
					stv(out + (base | col[c] | ro), o);
					nacc += cabs2d(o);
				}
			}
	}
	nacc = wave_sum(nacc);
	if (lane == 0)
		partials[gw] = nacc;
}

// Column tiles (16 groups each) a wave takes per step: enough bytes in
// flight for the small gates, few enough registers for the large ones.
template <typename T, int K> constexpr int stream_ct()
{
	return sizeof(T) == 16 ? (K == 4 ? 2 : 1) : (K == 4 ? 4 : K == 5 ? 2 : 1);
}

// The three-product form where the matrix cores bound the kernel (k = 6;
// measured at 30 qubits: complex128 8.02 -> 7.10 ms, complex64 4.54 ->
// 3.87-4.02 ms; k = 5 moved by 1% or less, and keeps the four-product form,
// whose sums match the tiled kernel's bit for bit), and the row tiles it
// takes at a time.
template <typename T, int K> constexpr bool stream_g3() { return K == 6; }
template <typename T, int K> constexpr int stream_rh() { return 1; }

template <typename T, int K>
constexpr auto stream_kernel = k_dense_stream<T, K, stream_ct<T, K>(), stream_g3<T, K>(),
					      stream_rh<T, K>()>;

// Grid of the streaming form: as many blocks as are resident at once (the
// blocks are persistent and take equal shares), fewer for small states.
template <typename T, int K> unsigned stream_grid(u64 ncg)
{
	static int bpc = 0;
	if (!bpc && hipOccupancyMaxActiveBlocksPerMultiprocessor(
			    &bpc, stream_kernel<T, K>, SW * 64, 0) !=
			    hipSuccess)
		bpc = 1;
	const u64 want = (u64)std::max(1, g_ctx.num_cus) * (u64)std::max(1, bpc);
	return (unsigned)std::max<u64>(1, std::min<u64>(want, (ncg + SW - 1) / SW));
}

template <typename T, int K>
void launch_stream(const T *src, T *dst, const double *scal, const T *M,
		   const StreamArgs &sa, double *parts, u64 ncg, unsigned grid)
{
	hipLaunchKernelGGL((stream_kernel<T, K>), dim3(grid),
			   dim3(SW * 64), 0, DK_STREAM, src, dst, scal, M, sa, parts, ncg);
}

// Column sets per block for k = 4 (tunable "dense_k4_cw": 1, 2 or 4; 0
// picks 1, the fastest on gfx950 at 30 qubits averaged over six target
// positions: complex128 6.43 / 6.62 / 7.36 ms for 1 / 2 / 4 sets, complex64
// 3.36 / 3.45 / 3.58 ms.  Before complex128 tiles were loaded straight into
// LDS, two sets were the faster there: 6.2 against 10.7 ms.)
int k4_cw(int dtype)
{
	(void)dtype;
	if (g_ctx.dense_k4_cw)
		return g_ctx.dense_k4_cw;
	return 1;
}

template <typename T, int K, int CW>
void launch_shape(const T *src, T *dst, const double *scal, const T *M,
		  const TileArgs &ta, double *parts, u64 nblocks)
{
	constexpr unsigned NT = (4u << K) * CW;
	const dim3 grid(dk_grid_cap(nblocks, NT));
	if (grid.x < nblocks)
		hipLaunchKernelGGL((k_apply_dense<T, K, CW, true>), grid, dim3(NT), 0,
				   DK_STREAM, src, dst, scal, M, ta, parts, nblocks);
	else
		hipLaunchKernelGGL((k_apply_dense<T, K, CW, false>), grid, dim3(NT), 0,
				   DK_STREAM, src, dst, scal, M, ta, parts, nblocks);
}

template <typename T>
int launch_dense(const dk_state *in, dk_state *out, const dk_gate *g,
		 const TileArgs &ta, u64 nblocks, double *parts)
{
	const T *M = std::is_same<T, double2>::value ? (const T *)g->dmat
						      : (const T *)g->dmat32;
	const T *src = (const T *)in->amp;
	T *dst = (T *)out->amp;
	switch (g->k) {
	case 4:
		if (k4_cw(in->dtype) == 4)
			launch_shape<T, 4, 4>(src, dst, in->scal, M, ta, parts, nblocks);
		else if (k4_cw(in->dtype) == 2)
			launch_shape<T, 4, 2>(src, dst, in->scal, M, ta, parts, nblocks);
		else
			launch_shape<T, 4, 1>(src, dst, in->scal, M, ta, parts, nblocks);
		break;
	case 5:
		launch_shape<T, 5, 1>(src, dst, in->scal, M, ta, parts, nblocks);
		break;
	case 6:
		launch_shape<T, 6, 1>(src, dst, in->scal, M, ta, parts, nblocks);
		break;
	default:
		dk_set_error("dense gates act on 4..6 qubits");
		return DK_E_ARG;
	}
	DK_LAUNCHED();
	return DK_OK;
}

} // namespace

unsigned dk_dense_tile_bits(unsigned k, int dtype)
{
	const int cw = k4_cw(dtype);
	const unsigned cwb = k == 4 ? (cw == 4 ? 2 : cw == 2 ? 1 : 0) : 0;
	return k + 6 + cwb;
}

// Index bits known to be zero in the input (none when the gate has a
// non-finite entry: 0 * inf is not zero).
static u64 dense_zmask(const dk_state *in, const dk_gate *g)
{
	return g->finite ? (~in->support & (in->len - 1)) : 0;
}

// Streaming dense launch: grid, partials (one per wave) and the gate's
// index tables.
static int launch_dense_stream(const dk_state *in, dk_state *out, const dk_gate *g,
			       const unsigned *targets, double **parts, u64 *nparts,
			       double *scratch, u64 scratch_cap)
{
	const unsigned k = g->k, D = 1u << k;
	StreamArgs sa;
	memset(&sa, 0, sizeof(sa));
	for (unsigned j = 0; j < D; ++j) {
		u64 o = 0;
		for (unsigned b = 0; b < k; ++b)
			if ((j >> b) & 1)
				o |= 1ull << targets[b];
		sa.rowoff[j] = o;
	}
	unsigned ts[8] = {0};
	for (unsigned b = 0; b < k; ++b)
		ts[b] = targets[b];
	std::sort(ts, ts + k);
	for (unsigned b = 0; b < k; ++b)
		sa.tsorted[b] = (unsigned char)ts[b];
	sa.zmask = dense_zmask(in, g);
	sa.zwrite = in != out;
	const bool c128 = in->dtype != DK_C64;
	int ct = 1;
	if (k == 4)
		ct = c128 ? stream_ct<double2, 4>() : stream_ct<float2, 4>();
	else if (k == 5)
		ct = c128 ? stream_ct<double2, 5>() : stream_ct<float2, 5>();
	else
		ct = c128 ? stream_ct<double2, 6>() : stream_ct<float2, 6>();
	const u64 ngroups = in->len >> (k + 4);            // 16-group units
	if (ngroups == 0 || ngroups % (u64)ct) {
		dk_set_error("dense gate of %u qubits: state of %u qubits too small "
			     "for the streaming form", k, in->nq);
		return DK_E_ARG;
	}
	const u64 ncg = ngroups / (u64)ct;  // blocks of 16 ct groups
	unsigned grid = 1;
	if (k == 4)
		grid = c128 ? stream_grid<double2, 4>(ncg) : stream_grid<float2, 4>(ncg);
	else if (k == 5)
		grid = c128 ? stream_grid<double2, 5>(ncg) : stream_grid<float2, 5>(ncg);
	else
		grid = c128 ? stream_grid<double2, 6>(ncg) : stream_grid<float2, 6>(ncg);
	const u64 nw = (u64)grid * SW;
	int rc = scratch ? dk_partials_in(scratch, scratch_cap, nw, parts)
			 : dk_partials_reserve(nw, parts);
	if (rc)
		return rc;
	*nparts = nw;
	const double2 *M128 = g->dmat;
	const float2 *M64 = g->dmat32;
#define DK_STREAM_K(KK)                                                        \
	case KK:                                                               \
		if (c128)                                                      \
			launch_stream<double2, KK>((const double2 *)in->amp,   \
						   (double2 *)out->amp, in->scal, M128, \
						   sa, *parts, ncg, grid);      \
		else                                                           \
			launch_stream<float2, KK>((const float2 *)in->amp,     \
						  (float2 *)out->amp, in->scal, M64, \
						  sa, *parts, ncg, grid);       \
		break;
	switch (k) {
		DK_STREAM_K(4)
		DK_STREAM_K(5)
		DK_STREAM_K(6)
	default:
		dk_set_error("dense gates act on 4..6 qubits");
		return DK_E_ARG;
	}
#undef DK_STREAM_K
	DK_LAUNCHED();
	return DK_OK;
}

int dk_launch_apply_dense(const dk_state *in, dk_state *out, const dk_gate *g,
			  const unsigned *targets, double **parts, u64 *nparts,
			  double *scratch, u64 scratch_cap)
{
	// the streaming form where it measured faster (30 qubits, gfx950):
	// complex128 k = 5 6.3-7.6 against 7.8-8.2 ms, k = 6 7.1 against 11.3
	// ms (77 TF/s), whatever the targets; the tiles stay ahead for k = 4.
	// complex64 streams well only while the 16 groups a wave reads are
	// long runs: with no target among qubits 0..2, k = 5 takes 3.2-3.4
	// against 4.1 ms and k = 6 3.9-4.0 against 6.35 ms (137-142 TF/s).
	// With low targets it falls to 5-13 ms (a wave's 8-byte reads touch a
	// quarter of each line) and the tiles run.
	bool stream = false;
	if (g_ctx.dense_stream == 2) {
		stream = true;
	} else if (g_ctx.dense_stream == 1 && g->k >= 5) {
		stream = in->dtype == DK_C128;
		if (!stream) {
			stream = true;
			for (unsigned b = 0; b < g->k; ++b)
				if (targets[b] < 3)
					stream = false;
		}
	}
	if (stream)
		return launch_dense_stream(in, out, g, targets, parts, nparts, scratch,
					   scratch_cap);
	// one block per tile of 64 * CW groups: targets plus the 6 + log2(CW)
	// lowest other bits
	const unsigned tb = dk_dense_tile_bits(g->k, in->dtype);
	if (in->nq < tb) {
		dk_set_error("dense gate of %u qubits needs %u qubits of state", g->k, tb);
		return DK_E_ARG;
	}
	TileArgs ta;
	dk_tile_args(in->nq, g->k, targets, ta, (int)tb);
	ta.zmask = dense_zmask(in, g);
	ta.zwrite = in != out;
	const u64 nblocks = in->len >> ta.npos;
	int rc = scratch ? dk_partials_in(scratch, scratch_cap, nblocks, parts)
			 : dk_partials_reserve(nblocks, parts);
	if (rc)
		return rc;
	*nparts = nblocks;
	return in->dtype == DK_C64
		       ? launch_dense<float2>(in, out, g, ta, nblocks, *parts)
		       : launch_dense<double2>(in, out, g, ta, nblocks, *parts);
}

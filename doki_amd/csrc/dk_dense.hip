// Dense k-qubit gates (k = 4, 5, 6) on the matrix cores.
//
// An LDS tile holds 64 groups of D = 2^k amplitudes.  The tile is the
// complex matrix Psi (D rows x 64 columns); the gate computes O = M * Psi,
// with M the row-major gate (row = output basis state).  Complex products use
// the exact four-real-product form
//     Re O = Re M Re Psi - Im M Im Psi,   Im O = Re M Im Psi + Im M Re Psi,
// each a sequence of 16x16x4 MFMAs accumulating in the amplitude's precision
// (FP64: v_mfma_f64_16x16x4f64, FP32: v_mfma_f32_16x16x4f32).
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
	// complex128 k = 4 tiles load straight into LDS (measured on gfx950 at
	// 30 qubits: 6.39 against 6.67 ms; k = 5 and 6 ran 4% and 23% slower)
	constexpr bool LDS_DIRECT = sizeof(T) == 16 && K == 4;

	// Slot s = threadIdx.x + j * NT: its index bits are this thread's part
	// (computed once) OR the part of j * NT (the same for the whole block).
	constexpr int PER = TS / NT; // 16 slots per thread
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

int dk_launch_apply_dense(const dk_state *in, dk_state *out, const dk_gate *g,
			  const unsigned *targets, double **parts, u64 *nparts,
			  double *scratch, u64 scratch_cap)
{
	// one block per tile of 64 * CW groups: targets plus the 6 + log2(CW)
	// lowest other bits
	const unsigned tb = dk_dense_tile_bits(g->k, in->dtype);
	if (in->nq < tb) {
		dk_set_error("dense gate of %u qubits needs %u qubits of state", g->k, tb);
		return DK_E_ARG;
	}
	TileArgs ta;
	dk_tile_args(in->nq, g->k, targets, ta, (int)tb);
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

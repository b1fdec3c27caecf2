// Dense k-qubit gates (k = 4, 5, 6) on the matrix cores.
//
// A 4096-amplitude LDS tile holds G = 4096 / D groups of D = 2^k amplitudes
// (the gate's targets plus the lowest other index bits).  The tile is the
// complex matrix Psi (D rows x G columns); the gate computes O = M * Psi,
// with M the row-major gate (row = output basis state).  Complex products use
// the exact four-real-product form
//     Re O = Re M Re Psi - Im M Im Psi,   Im O = Re M Im Psi + Im M Re Psi,
// each a sequence of 16x16x4 MFMAs accumulating in the amplitude's precision
// (FP64: v_mfma_f64_16x16x4f64, FP32: v_mfma_f32_16x16x4f32).
//
// Wave w owns a 16-row tile of M (rows 16*rt..) against 64 consecutive groups
// (four 16-column tiles), so its A fragments — 16 rows x all D columns of M —
// stay in registers for the whole launch while tiles stream through LDS.
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

template <typename T, int K>
__global__ __launch_bounds__(TPB) void k_apply_dense(
	const T *in, T *out, const double *__restrict__ nrm,
	const T *__restrict__ M, TileArgs ta, u64 ntiles,
	double *__restrict__ partials)
{
	typedef typename real_of<T>::type R;
	typedef mfma16<R> MF;
	typedef typename MF::acc acc_t;
	constexpr int D = 1 << K;   // 16, 32, 64
	constexpr int KS = D / 4;   // k-steps of the product
	constexpr int RT = D / 16;  // row tiles
	__shared__ T lds[TILE];
	__shared__ double red[WAVES];

	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	const int rt = w % RT, cb = w / RT; // row tile, 64-column block
	const int li = lane & 15, lk = lane >> 4;

	// A fragments of this wave's 16 rows of M, all k-steps
	R ar[KS], ai[KS], nai[KS];
#pragma unroll
	for (int ks = 0; ks < KS; ++ks) {
		const T m = M[(size_t)(rt * 16 + li) * D + ks * 4 + lk];
		ar[ks] = m.x;
		ai[ks] = m.y;
		nai[ks] = -m.y;
	}
	// B slots: this lane's four columns (groups) and its row of each k-step
	int gs[4];
#pragma unroll
	for (int ct = 0; ct < 4; ++ct)
		gs[ct] = group_slot<K>(cb * 64 + ct * 16 + li, ta);
	int ks_slot[KS];
#pragma unroll
	for (int ks = 0; ks < KS; ++ks)
		ks_slot[ks] = ta.colslot[ks * 4 + lk];
	int out_slot[4];
#pragma unroll
	for (int r = 0; r < 4; ++r)
		out_slot[r] = ta.colslot[rt * 16 + MF::out_row(lk, r)];

	const R inv = (R)(1.0 / nrm[DK_SCAL_NORM]);
	double acc = 0.0;
	for (u64 tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
		const u64 base = tile_base(tile, ta);
		for (int s = threadIdx.x; s < TILE; s += TPB)
			lds[s] = in[tile_index(base, s, ta)];
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
				const T b = lds[gs[ct] | ks_slot[ks]];
				cr[ct] = MF::op(ar[ks], b.x, cr[ct]);
				cr[ct] = MF::op(nai[ks], b.y, cr[ct]);
				ci[ct] = MF::op(ar[ks], b.y, ci[ct]);
				ci[ct] = MF::op(ai[ks], b.x, ci[ct]);
			}
		}
		__syncthreads(); // every wave has read the tile
#pragma unroll
		for (int ct = 0; ct < 4; ++ct)
#pragma unroll
			for (int r = 0; r < 4; ++r)
				lds[gs[ct] | out_slot[r]] =
					mk<T>(cr[ct][r] * inv, ci[ct][r] * inv);
		__syncthreads();
		for (int s = threadIdx.x; s < TILE; s += TPB) {
			const T o = lds[s];
			st_stream(out + tile_index(base, s, ta), o);
			acc += cabs2d(o);
		}
		__syncthreads();
	}
	const double tot = block_sum(acc, red);
	if (threadIdx.x == 0)
		partials[blockIdx.x] = tot;
}

template <typename T>
int launch_dense(const dk_state *in, dk_state *out, const dk_gate *g,
		 const TileArgs &ta, u64 ntiles, unsigned nblocks, double *parts)
{
	const T *M = std::is_same<T, double2>::value ? (const T *)g->dmat
						      : (const T *)g->dmat32;
	const T *src = (const T *)in->amp;
	T *dst = (T *)out->amp;
	hipStream_t st = g_ctx.stream;
	switch (g->k) {
	case 4:
		hipLaunchKernelGGL((k_apply_dense<T, 4>), dim3(nblocks), dim3(TPB), 0,
				   st, src, dst, in->scal, M, ta, ntiles, parts);
		break;
	case 5:
		hipLaunchKernelGGL((k_apply_dense<T, 5>), dim3(nblocks), dim3(TPB), 0,
				   st, src, dst, in->scal, M, ta, ntiles, parts);
		break;
	case 6:
		hipLaunchKernelGGL((k_apply_dense<T, 6>), dim3(nblocks), dim3(TPB), 0,
				   st, src, dst, in->scal, M, ta, ntiles, parts);
		break;
	default:
		dk_set_error("dense gates act on 4..6 qubits");
		return DK_E_ARG;
	}
	DK_LAUNCHED();
	return DK_OK;
}

} // namespace

int dk_launch_apply_dense(const dk_state *in, dk_state *out, const dk_gate *g,
			  const unsigned *targets, double **parts, u64 *nparts)
{
	TileArgs ta;
	dk_tile_args(in->nq, g->k, targets, ta);
	const u64 ntiles = in->len >> TILE_BITS;
	// persistent-style grid: the gate's fragments are loaded once per block
	const u64 resident = (u64)std::max(1, g_ctx.num_cus) *
			     (in->dtype == DK_C64 ? 4 : 2);
	const unsigned nblocks = (unsigned)std::min<u64>(ntiles, resident);
	int rc = dk_partials_reserve(nblocks, parts);
	if (rc)
		return rc;
	*nparts = nblocks;
	return in->dtype == DK_C64
		       ? launch_dense<float2>(in, out, g, ta, ntiles, nblocks, *parts)
		       : launch_dense<double2>(in, out, g, ta, ntiles, nblocks, *parts);
}

// gfx950 kernels of the state-vector core (everything except the MFMA dense
// gate kernel, which lives in dk_dense.hip).
//
// Layout: a state of n qubits is 2^n amplitudes — complex128 (double2) or
// complex64 (float2), re then im — in one contiguous HBM allocation; index
// bit q is qubit q.  Every kernel reads a state through its normalisation
// constant (value = raw * (1/norm)), as the reference's state_get does
// (src/doki/qstate.h:86).  Reductions are FP64 and deterministic for a given
// state size: fixed grids, per-block partials summed in block order.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "dk_device.h"

using namespace dkd;

namespace {

template <typename T> struct Mat2 {
	T m00, m01, m10, m11;
};

template <typename T> Mat2<T> mat2_of(const dk_gate *g)
{
	Mat2<T> M;
	T *dst[4] = {&M.m00, &M.m01, &M.m10, &M.m11};
	for (int i = 0; i < 4; ++i) {
		dst[i]->x = (decltype(dst[i]->x))g->hmat[i].x;
		dst[i]->y = (decltype(dst[i]->y))g->hmat[i].y;
	}
	return M;
}

// ---------------------------------------------------------------------------
// One-qubit gates.  Full-grid forms: one launch covers the state exactly
// once, one amplitude per thread, blocks dispatched in address order so the
// amplitudes in flight form a sliding window over HBM.  Each block leaves one
// partial |out|^2 sum.
//
// t >= 6: a block holds G pair groups of two waves each.  In a group, wave
// 0 reads the lower amplitudes of 64 consecutive pairs (one contiguous run),
// wave 1 the partner run 2^t amplitudes above; partners are exchanged
// through LDS.  One partial per block.
// ---------------------------------------------------------------------------
template <typename T, int G>
__device__ __forceinline__ void pairs_block(
	const T *in, T *out, typename real_of<T>::type inv, const Mat2<T> &M,
	unsigned t, u64 cm, u64 acm, double *__restrict__ partials, u64 vb,
	T *buf, double *red)
{
	const u64 bit = 1ull << t, lo = bit - 1;
	const u64 p = (vb * G + (threadIdx.x >> 7)) * 64 + (threadIdx.x & 63);
	const u64 i0 = ((p & ~lo) << 1) | (p & lo);
	const bool upper = (threadIdx.x >> 6) & 1;
	const u64 my = upper ? (i0 | bit) : i0;
	const T v = cscale(ld_stream(in + my), inv);
	buf[threadIdx.x] = v;
	__syncthreads();
	const T w = buf[threadIdx.x ^ 64];
	const T x = upper ? w : v, y = upper ? v : w;
	T o = v;
	if ((i0 & cm) == cm && (i0 & acm) == 0) {
		o = upper ? cmac(cmac(czero<T>(), M.m10, x), M.m11, y)
			  : cmac(cmac(czero<T>(), M.m00, x), M.m01, y);
	}
	st_stream(out + my, o);
	const double s = wave_sum(cabs2d(o));
	if ((threadIdx.x & 63) == 0)
		red[threadIdx.x >> 6] = s;
	__syncthreads();
	if (threadIdx.x == 0) {
		double tot = 0.0;
#pragma unroll
		for (int w = 0; w < 2 * G; ++w)
			tot += red[w];
		partials[vb] = tot;
	}
}

// LOOP: virtual blocks beyond the grid (launches of more than 2^31 threads).
template <typename T, int G, bool LOOP>
__global__ __launch_bounds__(128 * G) void k_apply1_pairs(
	const T *in, T *out, const double *__restrict__ nrm, Mat2<T> M,
	unsigned t, u64 cm, u64 acm, double *__restrict__ partials, u64 nvb)
{
	typedef typename real_of<T>::type R;
	__shared__ T buf[128 * G];
	__shared__ double red[2 * G];
	const R inv = (R)(1.0 / nrm[DK_SCAL_NORM]);
	if constexpr (LOOP) {
		for (u64 vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
			pairs_block<T, G>(in, out, inv, M, t, cm, acm, partials, vb, buf, red);
			__syncthreads();
		}
	} else {
		pairs_block<T, G>(in, out, inv, M, t, cm, acm, partials, blockIdx.x, buf,
				  red);
	}
}

// t = 6..17: one wave per block holds 32 pairs, lanes 0..31 the lower
// amplitudes (a contiguous 512-byte run) and lanes 32..63 their partners
// 2^t above; the partner comes from lane ^ 32 (no LDS buffer, no barrier).
// Measured at 30 qubits: 1-5% faster than the two-wave LDS exchange for
// these targets (6.5-6.8 TB/s), slower from t = 18 on (5.9 TB/s at 20-21).
template <typename T>
__device__ __forceinline__ void halves_block(
	const T *in, T *out, typename real_of<T>::type inv, const Mat2<T> &M,
	unsigned t, u64 cm, u64 acm, double *__restrict__ partials, u64 vb)
{
	const u64 bit = 1ull << t, lo = bit - 1;
	const unsigned lane = threadIdx.x & 63;
	const u64 p = vb * 32 + (lane & 31);
	const u64 i0 = ((p & ~lo) << 1) | (p & lo);
	const bool upper = lane >= 32;
	const u64 my = upper ? (i0 | bit) : i0;
	const T v = cscale(ld_stream(in + my), inv);
	T w;
	w.x = __shfl_xor(v.x, 32);
	w.y = __shfl_xor(v.y, 32);
	const T x = upper ? w : v, y = upper ? v : w;
	T o = v;
	if ((i0 & cm) == cm && (i0 & acm) == 0) {
		o = upper ? cmac(cmac(czero<T>(), M.m10, x), M.m11, y)
			  : cmac(cmac(czero<T>(), M.m00, x), M.m01, y);
	}
	st_stream(out + my, o);
	const double s = wave_sum(cabs2d(o));
	if (threadIdx.x == 0)
		partials[vb] = s;
}

template <typename T>
__global__ __launch_bounds__(64) void k_apply1_halves(
	const T *in, T *out, const double *__restrict__ nrm, Mat2<T> M,
	unsigned t, u64 cm, u64 acm, double *__restrict__ partials)
{
	typedef typename real_of<T>::type R;
	const R inv = (R)(1.0 / nrm[DK_SCAL_NORM]);
	halves_block<T>(in, out, inv, M, t, cm, acm, partials, blockIdx.x);
}

// t < 6: W waves per block, each over 64 consecutive amplitudes; the
// partner amplitude comes from lane ^ 2^t.  One partial per block.
template <typename T, int W>
__device__ __forceinline__ void lanes_block(
	const T *in, T *out, typename real_of<T>::type inv, const Mat2<T> &M,
	unsigned t, u64 cm, u64 acm, double *__restrict__ partials, u64 vb,
	double *red)
{
	const u64 i = vb * 64 * W + threadIdx.x;
	const T v = cscale(ld_stream(in + i), inv);
	T w;
	w.x = __shfl_xor(v.x, 1 << t);
	w.y = __shfl_xor(v.y, 1 << t);
	const bool upper = (i >> t) & 1;
	const T x = upper ? w : v, y = upper ? v : w;
	T o = v;
	if ((i & cm) == cm && (i & acm) == 0) {
		o = upper ? cmac(cmac(czero<T>(), M.m10, x), M.m11, y)
			  : cmac(cmac(czero<T>(), M.m00, x), M.m01, y);
	}
	st_stream(out + i, o);
	const double s = wave_sum(cabs2d(o));
	if (W == 1) {
		if (threadIdx.x == 0)
			partials[vb] = s;
		return;
	}
	if ((threadIdx.x & 63) == 0)
		red[threadIdx.x >> 6] = s;
	__syncthreads();
	if (threadIdx.x == 0) {
		double tot = 0.0;
#pragma unroll
		for (int w = 0; w < W; ++w)
			tot += red[w];
		partials[vb] = tot;
	}
}

template <typename T, int W, bool LOOP>
__global__ __launch_bounds__(64 * W) void k_apply1_lanes(
	const T *in, T *out, const double *__restrict__ nrm, Mat2<T> M,
	unsigned t, u64 cm, u64 acm, double *__restrict__ partials, u64 nvb)
{
	typedef typename real_of<T>::type R;
	__shared__ double red[W];
	const R inv = (R)(1.0 / nrm[DK_SCAL_NORM]);
	if constexpr (LOOP) {
		for (u64 vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
			lanes_block<T, W>(in, out, inv, M, t, cm, acm, partials, vb, red);
			__syncthreads();
		}
	} else {
		lanes_block<T, W>(in, out, inv, M, t, cm, acm, partials, blockIdx.x, red);
	}
}

// Small registers (fewer than 128 amplitudes): pairs, grid-stride, guarded.
template <typename T>
__global__ __launch_bounds__(TPB) void k_apply1_small(
	const T *in, T *out, const double *__restrict__ nrm, Mat2<T> M,
	unsigned t, u64 cm, u64 acm, u64 npairs, double *__restrict__ partials)
{
	typedef typename real_of<T>::type R;
	__shared__ double red[WAVES];
	const R inv = (R)(1.0 / nrm[DK_SCAL_NORM]);
	const u64 bit = 1ull << t, lo = bit - 1;
	double acc = 0.0;
	for (u64 p = (u64)blockIdx.x * TPB + threadIdx.x; p < npairs;
	     p += (u64)gridDim.x * TPB) {
		const u64 i0 = ((p & ~lo) << 1) | (p & lo);
		const T x = cscale(in[i0], inv), y = cscale(in[i0 | bit], inv);
		T o0 = x, o1 = y;
		if ((i0 & cm) == cm && (i0 & acm) == 0) {
			o0 = cmac(cmac(czero<T>(), M.m00, x), M.m01, y);
			o1 = cmac(cmac(czero<T>(), M.m10, x), M.m11, y);
		}
		out[i0] = o0;
		out[i0 | bit] = o1;
		acc += cabs2d(o0) + cabs2d(o1);
	}
	const double tot = block_sum(acc, red);
	if (threadIdx.x == 0)
		partials[blockIdx.x] = tot;
}

// ---------------------------------------------------------------------------
// k-qubit gates (k = 1..6, with controls) through an LDS tile on the vector
// ALUs.  Loads and stores walk the tile in slot order, so consecutive lanes
// touch consecutive addresses whenever the low positions are contiguous.
// ---------------------------------------------------------------------------
template <typename T, int K>
__global__ __launch_bounds__(TPB) void k_apply_tile(
	const T *in, T *out, const double *__restrict__ nrm,
	const T *__restrict__ M, TileArgs ta, u64 cm, u64 acm, u64 ntiles,
	double *__restrict__ partials)
{
	typedef typename real_of<T>::type R;
	constexpr int D = 1 << K;
	constexpr int RR = D < 16 ? D : 16;  // rows per work item
	constexpr int NRB = D / RR;          // row blocks per group
	constexpr int OPT = TILE / TPB;      // outputs per thread (16)
	constexpr int IPT = OPT / RR;        // work items per thread
	__shared__ T lds[TILE];
	__shared__ double red[WAVES];

	const int tsz = 1 << ta.npos;
	const int G = tsz >> K;
	const int nwork = G * NRB;
	const R inv = (R)(1.0 / nrm[DK_SCAL_NORM]);
	double acc = 0.0;

	// slot s = threadIdx.x + j * TPB: index bits = thread part | part of j*TPB
	const u64 tdep = ta.dep_lo[threadIdx.x & 63] | ta.dep_hi[threadIdx.x >> 6];
	for (u64 tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
		const u64 base = tile_base(tile, ta) | tdep;
#pragma unroll
		for (int j = 0; j < OPT; ++j) {
			const int s = threadIdx.x + j * TPB;
			if (s < tsz)
				lds[lds_pos<T>(s)] =
					cscale(in[base | ta.dep_hi[(j * TPB) >> 6]], inv);
		}
		__syncthreads();

		T res[IPT][RR];
		int gsl[IPT], rbs[IPT];
#pragma unroll
		for (int it = 0; it < IPT; ++it) {
			const int w = threadIdx.x + it * TPB;
			gsl[it] = -1;
			rbs[it] = 0;
			if (w >= nwork)
				continue;
			const int g = w % G;
			const int rb = NRB == 1 ? 0 : w / G;
			rbs[it] = rb;
			const int gs = group_slot<K>(g, ta);
			gsl[it] = gs;
			const u64 gi = tile_index(base & ~tdep, gs, ta);
			if ((gi & cm) == cm && (gi & acm) == 0) {
#pragma unroll
				for (int r = 0; r < RR; ++r)
					res[it][r] = czero<T>();
				for (int j = 0; j < D; ++j) {
					const T x = lds[lds_pos<T>(gs | ta.colslot[j])];
					const T *mrow = M + (size_t)(rb * RR) * D + j;
#pragma unroll
					for (int r = 0; r < RR; ++r)
						res[it][r] = cmac(res[it][r], mrow[(size_t)r * D], x);
				}
			} else {
#pragma unroll
				for (int r = 0; r < RR; ++r)
					res[it][r] = lds[lds_pos<T>(gs | ta.colslot[rb * RR + r])];
			}
		}
		__syncthreads();
#pragma unroll
		for (int it = 0; it < IPT; ++it) {
			if (gsl[it] < 0)
				continue;
#pragma unroll
			for (int r = 0; r < RR; ++r)
				lds[lds_pos<T>(gsl[it] | ta.colslot[rbs[it] * RR + r])] = res[it][r];
		}
		__syncthreads();
#pragma unroll
		for (int j = 0; j < OPT; ++j) {
			const int s = threadIdx.x + j * TPB;
			if (s < tsz) {
				const T o = lds[lds_pos<T>(s)];
				out[base | ta.dep_hi[(j * TPB) >> 6]] = o;
				acc += cabs2d(o);
			}
		}
		__syncthreads();
	}
	const double tot = block_sum(acc, red);
	if (threadIdx.x == 0)
		partials[blockIdx.x] = tot;
}

// ---------------------------------------------------------------------------
// Reductions.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_finish_sum(
	const double *__restrict__ partials, int count, double *dst, int take_sqrt)
{
	__shared__ double red[WAVES];
	double s = 0.0;
	for (int i = threadIdx.x; i < count; i += TPB)
		s += partials[i];
	s = block_sum(s, red);
	if (threadIdx.x == 0)
		dst[0] = take_sqrt ? sqrt(s) : s;
}

// First level of a long reduction: out[b] = sum of in[b*4096 .. b*4096+4095].
constexpr int RCHUNK = 4096;
__global__ __launch_bounds__(TPB) void k_reduce_chunks(
	const double *__restrict__ in, u64 count, double *__restrict__ out)
{
	__shared__ double red[WAVES];
	const u64 base = (u64)blockIdx.x * RCHUNK;
	double s = 0.0;
	for (int k = threadIdx.x; k < RCHUNK; k += TPB) {
		const u64 i = base + k;
		if (i < count)
			s += in[i];
	}
	s = block_sum(s, red);
	if (threadIdx.x == 0)
		out[blockIdx.x] = s;
}

// One block per column: dst[c] = sum_b partials[b * cols + c].
__global__ __launch_bounds__(TPB) void k_finish_cols(
	const double *__restrict__ partials, int rows, int cols, double *dst)
{
	__shared__ double red[WAVES];
	const int c = blockIdx.x;
	double s = 0.0;
	for (int b = threadIdx.x; b < rows; b += TPB)
		s += partials[(size_t)b * cols + c];
	s = block_sum(s, red);
	if (threadIdx.x == 0)
		dst[c] = s;
}

// ---------------------------------------------------------------------------
// Probabilities.
// ---------------------------------------------------------------------------
// Half-index form: block of 64 threads per 64 indices with bit q set.
template <typename T>
__global__ __launch_bounds__(TPB) void k_prob(
	const T *__restrict__ in, const double *__restrict__ nrm, unsigned q,
	u64 half, double *__restrict__ partials)
{
	__shared__ double red[WAVES];
	const double inv = 1.0 / nrm[DK_SCAL_NORM];
	const u64 bit = 1ull << q, lo = bit - 1;
	double acc = 0.0;
	for (u64 j = (u64)blockIdx.x * TPB + threadIdx.x; j < half;
	     j += (u64)gridDim.x * TPB) {
		const T a = ld_stream(in + (((j & ~lo) << 1) | bit | (j & lo)));
		const double x = a.x * inv, y = a.y * inv;
		acc += fma(x, x, y * y);
	}
	const double tot = block_sum(acc, red);
	if (threadIdx.x == 0)
		partials[blockIdx.x] = tot;
}

// All single-qubit probabilities in one pass.  A wave takes chunks of 64*U
// consecutive amplitudes: index bits 0..5 are the lane, bits 6..6+LU-1 the
// unroll slot, higher bits the chunk number.  The amplitudes come from a
// source: a stored state (ReadAmps), or the product amplitudes of a join,
// stored as they are made (JoinAmps).
constexpr int PQ = 40; // qubits covered by the one-pass kernel

template <typename T> struct ReadAmps {
	const T *in;
	const double *nrm;
	struct Ready {
		const T *__restrict__ in;
		double inv;
		__device__ double2 operator()(u64 i) const
		{
			const T a = ld_stream(in + i);
			return make_double2(a.x * inv, a.y * inv);
		}
	};
	__device__ Ready ready() const { return {in, 1.0 / nrm[DK_SCAL_NORM]}; }
};

// r[i] = (a[i >> bbits] / |a|) (b[i & bmask] / |b|), written to o; the value
// returned is the stored one (the joined state has norm 1).
template <typename T> struct JoinAmps {
	typedef typename real_of<T>::type R;
	const T *a;
	const double *na;
	const T *b;
	const double *nb;
	T *o;
	unsigned bbits;
	struct Ready {
		const T *__restrict__ a;
		const T *__restrict__ b;
		T *__restrict__ o;
		R ia, ib;
		unsigned bbits;
		u64 bmask;
		__device__ double2 operator()(u64 i) const
		{
			const T r = cmul(cscale(a[i >> bbits], ia),
					 cscale(b[i & bmask], ib));
			st_stream(o + i, r);
			return make_double2(r.x, r.y);
		}
	};
	__device__ Ready ready() const
	{
		return {a, b, o, (R)(1.0 / na[DK_SCAL_NORM]),
			(R)(1.0 / nb[DK_SCAL_NORM]), bbits, (1ull << bbits) - 1};
	}
};

// The grid holds 2^wbits waves and wave w takes chunks c = w + k 2^wbits,
// k = 0, 1, ...: the chunk bits below wbits are the wave's own and settle
// at the end; the loop keeps sums only for the bits of k (fewer than KB, as
// the launch guarantees).
constexpr int PROB_KB = 10;

template <int LU, typename Src>
__global__ __launch_bounds__(TPB) void k_prob_all(
	Src src, u64 len, unsigned wbits, double *__restrict__ partials,
	double *__restrict__ set_norm)
{
	constexpr int U = 1 << LU;
	constexpr int KB = PROB_KB;
	constexpr int NH = PQ - 6 - LU;
	__shared__ double red[WAVES][PQ];
	const auto amp = src.ready();
	const int lane = threadIdx.x & 63;
	const u64 wid = (u64)blockIdx.x * WAVES + (threadIdx.x >> 6);
	const u64 nch = len >> (6 + LU);
	double tot = 0.0, accu[LU], acck[KB];
#pragma unroll
	for (int b = 0; b < LU; ++b)
		accu[b] = 0.0;
#pragma unroll
	for (int b = 0; b < KB; ++b)
		acck[b] = 0.0;
	unsigned k = 0;
	for (u64 c = wid; c < nch; c += 1ull << wbits, ++k) {
		double x[U];
		const u64 first = (c << (6 + LU)) + lane;
#pragma unroll
		for (int u = 0; u < U; ++u) {
			const double2 a = amp(first + (u64)u * 64);
			x[u] = fma(a.x, a.x, a.y * a.y);
		}
		double s = 0.0;
#pragma unroll
		for (int u = 0; u < U; ++u)
			s += x[u];
#pragma unroll
		for (int b = 0; b < LU; ++b) {
			double t = 0.0;
#pragma unroll
			for (int u = 0; u < U; ++u)
				if ((u >> b) & 1)
					t += x[u];
			accu[b] += t;
		}
		tot += s;
#pragma unroll
		for (int b = 0; b < KB; ++b)
			if ((k >> b) & 1) // the same for the whole wave
				acck[b] += s;
	}
	const int w = threadIdx.x >> 6;
	const double all = wave_sum(tot);
#pragma unroll
	for (int q = 0; q < 6; ++q) {
		const double v = wave_sum(((lane >> q) & 1) ? tot : 0.0);
		if (lane == 0)
			red[w][q] = v;
	}
#pragma unroll
	for (int b = 0; b < LU; ++b) {
		const double v = wave_sum(accu[b]);
		if (lane == 0)
			red[w][6 + b] = v;
	}
	// chunk bits: below wbits from the wave index, then k's bits, then 0
	for (int b = lane; b < NH; b += 64)
		if (b < (int)wbits)
			red[w][6 + LU + b] = ((wid >> b) & 1) ? all : 0.0;
		else if (b >= (int)wbits + KB)
			red[w][6 + LU + b] = 0.0;
#pragma unroll
	for (int j = 0; j < KB; ++j) {
		const double v = wave_sum(acck[j]);
		if (lane == 0 && (int)wbits + j < NH)
			red[w][6 + LU + wbits + j] = v;
	}
	__syncthreads();
	if (threadIdx.x < PQ) {
		double s = 0.0;
#pragma unroll
		for (int ww = 0; ww < WAVES; ++ww)
			s += red[ww][threadIdx.x];
		partials[(size_t)blockIdx.x * PQ + threadIdx.x] = s;
	}
	if (set_norm && blockIdx.x == 0 && threadIdx.x == 0)
		set_norm[DK_SCAL_NORM] = 1.0;
}

// ---------------------------------------------------------------------------
// Measurement: decision and collapse.
// ---------------------------------------------------------------------------
// dec[0] = outcome (0/1), dec[1] = P1.  When next_norm is given it receives
// sqrt(P) with P = P1 (outcome 1) or 1 - P1 (outcome 0).
__global__ __launch_bounds__(TPB) void k_decide(
	const double *__restrict__ partials, int count, double roll, double *dec,
	double *next_norm)
{
	__shared__ double red[WAVES];
	double s = 0.0;
	for (int i = threadIdx.x; i < count; i += TPB)
		s += partials[i];
	s = block_sum(s, red);
	if (threadIdx.x == 0) {
		const double p1 = s;
		const bool v = p1 > roll;
		const double p = v ? p1 : 1.0 - p1;
		dec[0] = v ? 1.0 : 0.0;
		dec[1] = p1;
		if (next_norm)
			next_norm[DK_SCAL_NORM] = sqrt(p);
	}
}

// o[j] = s[j with bit q inserted = outcome] / s.norm; when next_q >= 0 also
// the block partials of sum_{j: bit next_q set} |o_j / o.norm|^2.
template <typename T>
__global__ __launch_bounds__(TPB) void k_collapse(
	const T *__restrict__ in, const double *__restrict__ nrm_in, unsigned q,
	const double *__restrict__ dec, T *__restrict__ out,
	const double *__restrict__ nrm_out, int next_q, u64 olen,
	double *__restrict__ partials)
{
	typedef typename real_of<T>::type R;
	__shared__ double red[WAVES];
	const R inv = (R)(1.0 / nrm_in[DK_SCAL_NORM]);
	const double inv_o = 1.0 / nrm_out[DK_SCAL_NORM];
	const u64 bit = 1ull << q, lo = bit - 1;
	const u64 off = dec[0] != 0.0 ? bit : 0;
	double acc = 0.0;
	for (u64 j = (u64)blockIdx.x * TPB + threadIdx.x; j < olen;
	     j += (u64)gridDim.x * TPB) {
		const T o = cscale(ld_stream(in + (((j & ~lo) << 1) | off | (j & lo))), inv);
		out[j] = o;
		if (next_q >= 0 && ((j >> next_q) & 1)) {
			const double x = o.x * inv_o, y = o.y * inv_o;
			acc += fma(x, x, y * y);
		}
	}
	const double tot = block_sum(acc, red);
	if (threadIdx.x == 0)
		partials[blockIdx.x] = tot;
}

// ---------------------------------------------------------------------------
// Copies, join, fills, read-out.
// ---------------------------------------------------------------------------
template <typename TI, typename TO>
__global__ __launch_bounds__(TPB) void k_normalised_copy(
	const TI *__restrict__ in, const double *__restrict__ nrm,
	TO *__restrict__ out, double *out_scal, u64 len)
{
	const double inv = 1.0 / nrm[DK_SCAL_NORM];
	for (u64 i = (u64)blockIdx.x * TPB + threadIdx.x; i < len;
	     i += (u64)gridDim.x * TPB) {
		const TI a = ld_stream(in + i);
		TO o;
		o.x = (decltype(o.x))(a.x * inv);
		o.y = (decltype(o.y))(a.y * inv);
		st_stream(out + i, o);
	}
	if (blockIdx.x == 0 && threadIdx.x == 0)
		out_scal[DK_SCAL_NORM] = 1.0;
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_widen(const T *__restrict__ in,
					       u64 count, double2 *__restrict__ out)
{
	for (u64 i = (u64)blockIdx.x * TPB + threadIdx.x; i < count;
	     i += (u64)gridDim.x * TPB)
		out[i] = make_double2(in[i].x, in[i].y);
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_join(
	const T *__restrict__ a, const double *__restrict__ na,
	const T *__restrict__ b, const double *__restrict__ nb, unsigned bbits,
	u64 olen, T *__restrict__ o, double *o_scal)
{
	typedef typename real_of<T>::type R;
	const R ia = (R)(1.0 / na[DK_SCAL_NORM]), ib = (R)(1.0 / nb[DK_SCAL_NORM]);
	const u64 bmask = (1ull << bbits) - 1;
	for (u64 idx = (u64)blockIdx.x * TPB + threadIdx.x; idx < olen;
	     idx += (u64)gridDim.x * TPB) {
		const T x = cscale(a[idx >> bbits], ia);
		const T y = cscale(b[idx & bmask], ib);
		st_stream(o + idx, cmul(x, y));
	}
	if (blockIdx.x == 0 && threadIdx.x == 0)
		o_scal[DK_SCAL_NORM] = 1.0;
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_fill_zero_state(T *__restrict__ amp,
							 double *scal, u64 len)
{
	for (u64 i = (u64)blockIdx.x * TPB + threadIdx.x; i < len;
	     i += (u64)gridDim.x * TPB) {
		T v = czero<T>();
		if (i == 0)
			v.x = 1;
		st_stream(amp + i, v);
	}
	if (blockIdx.x == 0 && threadIdx.x == 0)
		scal[DK_SCAL_NORM] = 1.0;
}

__global__ void k_set_scalar(double *dst, double v)
{
	if (threadIdx.x == 0)
		dst[0] = v;
}

__global__ void k_set_u64(unsigned long long *dst, u64 v)
{
	if (threadIdx.x == 0)
		dst[0] = v;
}

// Smallest index >= start holding a nonzero raw amplitude, via atomicMin.
template <typename T>
__global__ __launch_bounds__(TPB) void k_first_nonzero(
	const T *__restrict__ in, u64 start, u64 len, unsigned long long *dst)
{
	__shared__ unsigned long long red[WAVES];
	u64 best = len;
	for (u64 i = start + (u64)blockIdx.x * TPB + threadIdx.x; i < len;
	     i += (u64)gridDim.x * TPB) {
		const T v = in[i];
		if (v.x != 0 || v.y != 0) {
			best = i;
			break;
		}
	}
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) {
		const u64 other = __shfl_xor((unsigned long long)best, o);
		best = other < best ? other : best;
	}
	if ((threadIdx.x & 63) == 0)
		red[threadIdx.x >> 6] = best;
	__syncthreads();
	if (threadIdx.x == 0) {
		u64 m = red[0];
		for (int w = 1; w < WAVES; ++w)
			m = red[w] < m ? red[w] : m;
		if (m < len)
			atomicMin(dst, (unsigned long long)m);
	}
}

// ---------------------------------------------------------------------------
// Reduced density matrix: rho[i][j] = sum_t A[i][t] conj(A[j][t]) with
// A[i][t] = amp[keep(i) | traced(t)] / norm, accumulated in FP64.  16x16
// output tiles, t staged through LDS 16 at a time; each block covers one
// t-chunk and writes its partial tile; k_density_sum adds chunks in order.
// ---------------------------------------------------------------------------
struct DensArgs {
	u64 keep_lo[64];    // kept index bits 0..5 -> state index bits
	u64 keep_hi[64];    // kept index bits 6..11 -> state index bits
	unsigned char kpos[TILE_BITS]; // kept positions, ascending
	int a;              // kept qubits
	int m;              // traced qubits
	u64 tmask, tval;    // only amplitudes with (index & tmask) == tval
	                    // count (tmask: traced bits; 0 = all)
};

// Amplitude idx of the state as the density reads it: zero where the traced
// bits fail the constraint (a projection folded into the density).
template <typename T>
__device__ __forceinline__ T dens_load(const T *in, u64 idx, const DensArgs &da)
{
	return (idx & da.tmask) == da.tval ? in[idx] : czero<T>();
}

__device__ __forceinline__ u64 traced_index(u64 t, const DensArgs &da)
{
	for (int b = 0; b < da.a; ++b)
		t = ins0(t, da.kpos[b]);
	return t;
}

__device__ __forceinline__ u64 keep_index(u64 i, const DensArgs &da)
{
	return da.keep_lo[i & 63] | da.keep_hi[(i >> 6) & 63];
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_density_tile(
	const T *__restrict__ in, const double *__restrict__ nrm, DensArgs da,
	u64 tchunk, int tiles_per_row, double2 *__restrict__ part)
{
	__shared__ double2 ai[16][17];
	__shared__ double2 aj[16][17];
	const double inv = 1.0 / nrm[DK_SCAL_NORM];
	const u64 dim = 1ull << da.a;
	const int ti = blockIdx.x / tiles_per_row, tj = blockIdx.x % tiles_per_row;
	const int r = threadIdx.x >> 4, c = threadIdx.x & 15;
	const u64 gi = (u64)ti * 16 + r, gj = (u64)tj * 16 + c;
	const u64 t0 = (u64)blockIdx.y * tchunk;
	const u64 T_ = 1ull << da.m;
	const u64 t1 = t0 + tchunk < T_ ? t0 + tchunk : T_;
	double2 acc = make_double2(0.0, 0.0);
	for (u64 tb = t0; tb < t1; tb += 16) {
		const u64 t = tb + c;
		const u64 ri = (u64)ti * 16 + r, rj = (u64)tj * 16 + r;
		double2 vi = make_double2(0.0, 0.0), vj = vi;
		if (t < t1) {
			const u64 tix = traced_index(t, da);
			if (ri < dim) {
				const T a = dens_load(in, keep_index(ri, da) | tix, da);
				vi = make_double2(a.x * inv, a.y * inv);
			}
			if (rj < dim) {
				const T a = dens_load(in, keep_index(rj, da) | tix, da);
				vj = make_double2(a.x * inv, a.y * inv);
			}
		}
		ai[r][c] = vi;
		aj[r][c] = vj;
		__syncthreads();
#pragma unroll
		for (int k = 0; k < 16; ++k) {
			const double2 x = ai[r][k], y = aj[c][k];
			acc.x = fma(x.x, y.x, acc.x);
			acc.x = fma(x.y, y.y, acc.x);
			acc.y = fma(x.y, y.x, acc.y);
			acc.y = fma(-x.x, y.y, acc.y);
		}
		__syncthreads();
	}
	if (gi < dim && gj < dim)
		part[(size_t)blockIdx.y * dim * dim + gi * dim + gj] = acc;
}

__global__ __launch_bounds__(TPB) void k_density_sum(
	const double2 *__restrict__ part, int nchunks, u64 count,
	double2 *__restrict__ rho)
{
	for (u64 e = (u64)blockIdx.x * TPB + threadIdx.x; e < count;
	     e += (u64)gridDim.x * TPB) {
		double2 s = make_double2(0.0, 0.0);
		for (int k = 0; k < nchunks; ++k) {
			const double2 v = part[(size_t)k * count + e];
			s.x += v.x;
			s.y += v.y;
		}
		rho[e] = s;
	}
}

// Reduced density of at most 6 kept qubits as a Gram matrix: with A the
// 2^a x 2^m matrix A[i][t] = psi(kept bits = i, traced bits = t),
// rho = A A^H.  Split over t: a block stages chunks of CH amplitudes (COLS
// = CH / DIM columns of A, consecutive t, so loads run along the state's
// low traced bits) in LDS, and each thread accumulates an RB x RB block of
// rho over every S-th column of the chunk, S = threads per rho block.  The
// block then adds its S replicas and leaves one DIM x DIM partial.
constexpr int DENS_CH = 2048;

template <typename T, int DIM, int RB>
__global__ __launch_bounds__(TPB) void k_density_gram(
	const T *__restrict__ in, const double *__restrict__ nrm, DensArgs da,
	u64 nchunks, int cols, double2 *__restrict__ part)
{
	constexpr int NBR = DIM / RB;          // rho blocks per row
	constexpr int NB = NBR * NBR;          // rho blocks
	constexpr int S = TPB / NB;            // threads per rho block
	static_assert(NB <= TPB && TPB % NB == 0, "block shape");
	__shared__ double2 lds[DENS_CH];
	const double inv = 1.0 / nrm[DK_SCAL_NORM];
	const int blk = threadIdx.x % NB, sub = threadIdx.x / NB;
	const int i0 = (blk / NBR) * RB, j0 = (blk % NBR) * RB;
	const int ch = cols * DIM;             // amplitudes per chunk
	double2 acc[RB][RB];
#pragma unroll
	for (int a = 0; a < RB; ++a)
#pragma unroll
		for (int b = 0; b < RB; ++b)
			acc[a][b] = make_double2(0.0, 0.0);
	for (u64 c = blockIdx.x; c < nchunks; c += gridDim.x) {
		const u64 t0 = c * (u64)cols;
		for (int e = threadIdx.x; e < ch; e += TPB) {
			const int i = e / cols, t = e - i * cols;
			const T v = dens_load(in, keep_index((u64)i, da) | traced_index(t0 + t, da), da);
			lds[e] = make_double2(v.x * inv, v.y * inv);
		}
		__syncthreads();
		for (int t = sub; t < cols; t += S) {
			double2 x[RB], y[RB];
#pragma unroll
			for (int a = 0; a < RB; ++a) {
				x[a] = lds[(i0 + a) * cols + t];
				y[a] = lds[(j0 + a) * cols + t];
			}
#pragma unroll
			for (int a = 0; a < RB; ++a)
#pragma unroll
				for (int b = 0; b < RB; ++b) {
					// x * conj(y)
					acc[a][b].x = fma(x[a].x, y[b].x, acc[a][b].x);
					acc[a][b].x = fma(x[a].y, y[b].y, acc[a][b].x);
					acc[a][b].y = fma(x[a].y, y[b].x, acc[a][b].y);
					acc[a][b].y = fma(-x[a].x, y[b].y, acc[a][b].y);
				}
		}
		__syncthreads();
	}
	// add the S replicas of each rho block (in replica order) through LDS,
	// one row a of the RB x RB blocks at a time (RB * TPB <= DENS_CH
	// entries); replica r of block bb is thread r * NB + bb
	static_assert(RB * TPB <= DENS_CH, "reduction buffer");
	double2 *red = lds;
#pragma unroll
	for (int a = 0; a < RB; ++a) {
#pragma unroll
		for (int b = 0; b < RB; ++b)
			red[b * TPB + threadIdx.x] = acc[a][b];
		__syncthreads();
		// rho rows i = ib * RB + a, all columns j
		for (int e = threadIdx.x; e < NBR * DIM; e += TPB) {
			const int ib = e / DIM, j = e % DIM;
			const int i = ib * RB + a, bb = ib * NBR + j / RB, b = j % RB;
			double2 sum = make_double2(0.0, 0.0);
			for (int r = 0; r < S; ++r) {
				const double2 v = red[b * TPB + r * NB + bb];
				sum.x += v.x;
				sum.y += v.y;
			}
			part[(size_t)blockIdx.x * DIM * DIM + (size_t)i * DIM + j] = sum;
		}
		__syncthreads();
	}
}

// One block per requested element.
template <typename T>
__global__ __launch_bounds__(TPB) void k_density_elements(
	const T *__restrict__ in, const double *__restrict__ nrm, DensArgs da,
	const u64 *__restrict__ rows, const u64 *__restrict__ cols,
	double2 *__restrict__ out)
{
	__shared__ double red[WAVES];
	const double inv = 1.0 / nrm[DK_SCAL_NORM];
	const u64 ki = keep_index(rows[blockIdx.x], da);
	const u64 kj = keep_index(cols[blockIdx.x], da);
	const u64 T_ = 1ull << da.m;
	double2 acc = make_double2(0.0, 0.0);
	for (u64 t = threadIdx.x; t < T_; t += TPB) {
		const u64 tix = traced_index(t, da);
		const T a = dens_load(in, ki | tix, da), b = dens_load(in, kj | tix, da);
		const double2 x = make_double2(a.x * inv, a.y * inv);
		const double2 y = make_double2(b.x * inv, -b.y * inv);
		acc = cmac(acc, x, y);
	}
	const double re = block_sum(acc.x, red);
	const double im = block_sum(acc.y, red);
	if (threadIdx.x == 0)
		out[blockIdx.x] = make_double2(re, im);
}

// ---------------------------------------------------------------------------
// Overlap and distance of two states of one size (normalised amplitudes):
// per-block FP64 partials of sum conj(a_i) b_i and max |a_i - b_i|, then one
// block combines them in block order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1)
		v = fmax(v, __shfl_xor(v, o));
	return v;
}

template <typename Ta, typename Tb>
__global__ __launch_bounds__(TPB) void k_compare(const Ta *a, const Tb *b,
						 const double *__restrict__ na,
						 const double *__restrict__ nb, u64 len,
						 double *__restrict__ parts)
{
	__shared__ double red[3][WAVES];
	const double ia = 1.0 / na[DK_SCAL_NORM], ib = 1.0 / nb[DK_SCAL_NORM];
	double sr = 0.0, si = 0.0, mx = 0.0;
	for (u64 i = (u64)blockIdx.x * TPB + threadIdx.x; i < len;
	     i += (u64)gridDim.x * TPB) {
		const Ta x = ld_stream(a + i);
		const Tb y = ld_stream(b + i);
		const double xr = (double)x.x * ia, xi = (double)x.y * ia;
		const double yr = (double)y.x * ib, yi = (double)y.y * ib;
		sr = fma(xr, yr, fma(xi, yi, sr));
		si = fma(xr, yi, fma(-xi, yr, si));
		const double dr = xr - yr, di = xi - yi;
		mx = fmax(mx, sqrt(fma(dr, dr, di * di)));
	}
	sr = wave_sum(sr);
	si = wave_sum(si);
	mx = wave_max(mx);
	if ((threadIdx.x & 63) == 0) {
		red[0][threadIdx.x >> 6] = sr;
		red[1][threadIdx.x >> 6] = si;
		red[2][threadIdx.x >> 6] = mx;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		double r = 0.0, im = 0.0, m = 0.0;
#pragma unroll
		for (int w = 0; w < WAVES; ++w) {
			r += red[0][w];
			im += red[1][w];
			m = fmax(m, red[2][w]);
		}
		parts[3 * (u64)blockIdx.x] = r;
		parts[3 * (u64)blockIdx.x + 1] = im;
		parts[3 * (u64)blockIdx.x + 2] = m;
	}
}

__global__ __launch_bounds__(TPB) void k_compare_finish(const double *__restrict__ parts,
							int nblocks, double *__restrict__ out)
{
	__shared__ double red[3][WAVES];
	double sr = 0.0, si = 0.0, mx = 0.0;
	for (int k = threadIdx.x; k < nblocks; k += TPB) {
		sr += parts[3 * k];
		si += parts[3 * k + 1];
		mx = fmax(mx, parts[3 * k + 2]);
	}
	sr = wave_sum(sr);
	si = wave_sum(si);
	mx = wave_max(mx);
	if ((threadIdx.x & 63) == 0) {
		red[0][threadIdx.x >> 6] = sr;
		red[1][threadIdx.x >> 6] = si;
		red[2][threadIdx.x >> 6] = mx;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		double r = 0.0, im = 0.0, m = 0.0;
#pragma unroll
		for (int w = 0; w < WAVES; ++w) {
			r += red[0][w];
			im += red[1][w];
			m = fmax(m, red[2][w]);
		}
		out[0] = r;
		out[1] = im;
		out[2] = m;
	}
}

// ---------------------------------------------------------------------------
// Host-side launch helpers
// ---------------------------------------------------------------------------
// One thread per unit (element-wise kernels): every load of the launch is in
// flight at once, which a capped grid-stride loop cannot match on HBM.  A
// launch holds fewer than 2^32 threads, so beyond 2^31 units the kernels'
// loops take over.
unsigned grid_full(u64 units, u64 per_block)
{
	const u64 b = (units + per_block - 1) / per_block;
	return (unsigned)std::max<u64>(1, std::min<u64>(b, (1ull << 31) / per_block));
}


unsigned grid_for(u64 units, u64 per_block)
{
	u64 b = (units + per_block - 1) / per_block;
	u64 cap = g_ctx.apply_blocks > 0 ? (u64)g_ctx.apply_blocks
					 : (u64)std::max(1, g_ctx.num_cus) * 8;
	// keep the grid a power of two so grid strides divide 2^n sizes
	u64 p = 1;
	while (p * 2 <= cap)
		p *= 2;
	b = std::min(b, p);
	return (unsigned)std::max<u64>(1, b);
}

template <typename T>
int launch_apply_typed(const dk_state *in, dk_state *out, const dk_gate *g,
		       const unsigned *targets, u64 cm, u64 acm, double *parts,
		       u64 *nblocks_out)
{
	hipStream_t st = DK_STREAM;
	const T *src = (const T *)in->amp;
	T *dst = (T *)out->amp;
	u64 nblocks = 0;
	if (g->k == 1) {
		const Mat2<T> M = mat2_of<T>(g);
		const unsigned t = targets[0];
		// pair groups / waves per block: 1, 2 or 4 (fewer, larger blocks
		// leave fewer partials), never more than the state holds
		int G = g_ctx.apply1_group;
		while (G > 1 && in->len < 128ull * G)
			G >>= 1;
		if (in->len >= 128 && t >= 6 && t <= 17 &&
		    dk_grid_cap(in->len / 64, 64) == in->len / 64) {
			// one launch covers the state (beyond 2^31 threads the
			// looping pair kernel measured faster)
			nblocks = in->len / 64;
			hipLaunchKernelGGL((k_apply1_halves<T>), dim3((unsigned)nblocks), dim3(64),
					   0, st, src, dst, in->scal, M, t, cm, acm, parts);
		} else if (in->len >= 128 && t >= 6) {
			nblocks = in->len / (128ull * G);
#define DK_PAIRS(GG)                                                           \
	do {                                                                   \
		const unsigned gr = dk_grid_cap(nblocks, 128 * GG);            \
		if (gr < nblocks)                                              \
			hipLaunchKernelGGL((k_apply1_pairs<T, GG, true>), dim3(gr), \
					   dim3(128 * GG), 0, st, src, dst, in->scal, \
					   M, t, cm, acm, parts, nblocks);         \
		else                                                           \
			hipLaunchKernelGGL((k_apply1_pairs<T, GG, false>), dim3(gr), \
					   dim3(128 * GG), 0, st, src, dst, in->scal, \
					   M, t, cm, acm, parts, nblocks);         \
	} while (0)
			if (G == 4)
				DK_PAIRS(4);
			else if (G == 2)
				DK_PAIRS(2);
			else
				DK_PAIRS(1);
#undef DK_PAIRS
		} else if (in->len >= 128) {
			nblocks = in->len / (64ull * G);
#define DK_LANES(GG)                                                           \
	do {                                                                   \
		const unsigned gr = dk_grid_cap(nblocks, 64 * GG);             \
		if (gr < nblocks)                                              \
			hipLaunchKernelGGL((k_apply1_lanes<T, GG, true>), dim3(gr), \
					   dim3(64 * GG), 0, st, src, dst, in->scal, \
					   M, t, cm, acm, parts, nblocks);         \
		else                                                           \
			hipLaunchKernelGGL((k_apply1_lanes<T, GG, false>), dim3(gr), \
					   dim3(64 * GG), 0, st, src, dst, in->scal, \
					   M, t, cm, acm, parts, nblocks);         \
	} while (0)
			if (G == 4)
				DK_LANES(4);
			else if (G == 2)
				DK_LANES(2);
			else
				DK_LANES(1);
#undef DK_LANES
		} else {
			nblocks = 1;
			hipLaunchKernelGGL(k_apply1_small<T>, dim3(1), dim3(TPB), 0, st,
					   src, dst, in->scal, M, t, cm, acm, in->len >> 1,
					   parts);
		}
	} else {
		const unsigned k = g->k;
		TileArgs ta;
		dk_tile_args(in->nq, k, targets, ta);
		const u64 ntiles = in->len >> ta.npos;
		nblocks = grid_for(ntiles, 1);
		const T *M = std::is_same<T, double2>::value ? (const T *)g->dmat
							      : (const T *)g->dmat32;
#define DK_TILE_CASE(KK)                                                    \
	case KK:                                                            \
		hipLaunchKernelGGL((k_apply_tile<T, KK>), dim3((unsigned)nblocks), \
				   dim3(TPB), 0, st, src, dst, in->scal, M, ta, \
				   cm, acm, ntiles, parts);                 \
		break;
		switch (k) {
			DK_TILE_CASE(2)
			DK_TILE_CASE(3)
			DK_TILE_CASE(4)
			DK_TILE_CASE(5)
			DK_TILE_CASE(6)
		default:
			dk_set_error("gate of %u qubits is not supported", k);
			return DK_E_ARG;
		}
#undef DK_TILE_CASE
	}
	DK_LAUNCHED();
	*nblocks_out = nblocks;
	return DK_OK;
}

} // namespace

// ===========================================================================

void dk_tile_args(unsigned n, unsigned k, const unsigned *targets, TileArgs &ta,
		  int tile_bits)
{
	memset(&ta, 0, sizeof(ta));
	const int npos = (int)std::min<unsigned>(
		(unsigned)std::min(tile_bits, (int)TILE_BITS), n);
	bool is_t[64] = {false};
	for (unsigned b = 0; b < k; ++b)
		is_t[targets[b]] = true;
	int cnt = 0, extra = npos - (int)k;
	for (unsigned p = 0; p < n && cnt < npos; ++p) {
		if (is_t[p]) {
			ta.pos[cnt++] = (unsigned char)p;
		} else if (extra > 0) {
			ta.pos[cnt++] = (unsigned char)p;
			--extra;
		}
	}
	ta.npos = npos;
	int slot_of[64] = {0};
	for (int b = 0; b < npos; ++b)
		slot_of[ta.pos[b]] = b;
	for (int s = 0; s < 64; ++s) {
		u64 lo = 0, hi = 0;
		for (int b = 0; b < 6; ++b) {
			if (!((s >> b) & 1))
				continue;
			if (b < npos)
				lo |= 1ull << ta.pos[b];
			if (6 + b < npos)
				hi |= 1ull << ta.pos[6 + b];
		}
		ta.dep_lo[s] = lo;
		ta.dep_hi[s] = hi;
	}
	for (unsigned j = 0; j < (1u << k); ++j) {
		unsigned off = 0;
		for (unsigned b = 0; b < k; ++b)
			if ((j >> b) & 1)
				off |= 1u << slot_of[targets[b]];
		ta.colslot[j] = (unsigned short)off;
	}
	int ts[8] = {0};
	for (unsigned b = 0; b < k; ++b)
		ts[b] = slot_of[targets[b]];
	std::sort(ts, ts + k);
	for (unsigned b = 0; b < k; ++b)
		ta.tslot[b] = (unsigned char)ts[b];
}

int dk_launch_apply(const dk_state *in, dk_state *out, const dk_gate *g,
		    const unsigned *targets, u64 cm, u64 acm, double *scratch,
		    u64 scratch_cap)
{
	double *parts = nullptr;
	u64 nblocks = 0;
	int rc;
	// a gate uploaded on another lane: this lane's stream waits for the
	// upload on the device (the host does not block)
	if (g->dmat && g->lane != dk_cur())
		DK_HIP(hipStreamWaitEvent(DK_STREAM, g->uploaded, 0));
	DK_TRY(dk_time_mark(0));
	if (g_ctx.dense_mfma && g->k >= 4 && cm == 0 && acm == 0 &&
	    in->nq >= dk_dense_tile_bits(g->k, in->dtype)) {
		rc = dk_launch_apply_dense(in, out, g, targets, &parts, &nblocks,
					   scratch, scratch_cap);
	} else {
		// every non-dense form leaves at most one partial per 64 amplitudes
		const u64 bound = std::max<u64>(in->len / 64, 1);
		rc = scratch ? dk_partials_in(scratch, scratch_cap, bound, &parts)
			     : dk_partials_reserve(bound, &parts);
		if (rc)
			return rc;
		rc = in->dtype == DK_C64
			     ? launch_apply_typed<float2>(in, out, g, targets, cm, acm,
							  parts, &nblocks)
			     : launch_apply_typed<double2>(in, out, g, targets, cm, acm,
							   parts, &nblocks);
	}
	if (rc)
		return rc;
	// the gate mixes amplitudes only along its targets (a non-finite entry
	// times a zero amplitude is NaN: then nothing is known any more)
	u64 tmask = 0;
	for (unsigned b = 0; b < g->k; ++b)
		tmask |= 1ull << targets[b];
	out->support = g->finite ? (in->support | tmask) : ~0ull;
	DK_TRY(dk_time_mark(1));
	if (g->dmat)
		DK_TRY(dk_gate_mark_use(g));
	return dk_finish_partials(parts, nblocks, out->scal + DK_SCAL_NORM, 1);
}

// Records, on the calling lane, that its queued work has read gate g (the
// gate's release waits for the last such record of every lane).
int dk_gate_mark_use(const dk_gate *g)
{
	dk_lane *l = dk_cur();
	std::lock_guard<std::mutex> lk(g->uses_mu);
	hipEvent_t ev = nullptr;
	for (auto &u : g->uses)
		if (u.first == l)
			ev = u.second;
	if (!ev) {
		DK_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
		g->uses.emplace_back(l, ev);
	}
	DK_HIP(hipEventRecord(ev, l->stream));
	return DK_OK;
}

// Deterministic sum of `count` device partials into *dst (sqrt when asked):
// one finishing block for short lists, otherwise 4096-wide chunks first.
// The partials buffer holds room for the chunk sums after `count`.
int dk_finish_partials(double *parts, u64 count, double *dst, int take_sqrt)
{
	hipStream_t st = DK_STREAM;
	if (count <= 4 * RCHUNK) {
		hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(TPB), 0, st, parts,
				   (int)count, dst, take_sqrt);
		DK_LAUNCHED();
		return DK_OK;
	}
	const u64 nb = (count + RCHUNK - 1) / RCHUNK;
	double *second = parts + count;
	hipLaunchKernelGGL(k_reduce_chunks, dim3((unsigned)nb), dim3(TPB), 0, st,
			   parts, count, second);
	DK_LAUNCHED();
	return dk_finish_partials(second, nb, dst, take_sqrt);
}

u64 dk_partials_need(u64 count)
{
	u64 need = count + 16;
	for (u64 c = count; c > 4 * RCHUNK; c = (c + RCHUNK - 1) / RCHUNK)
		need += (c + RCHUNK - 1) / RCHUNK;
	return need;
}

int dk_partials_in(double *scratch, u64 cap, u64 count, double **out)
{
	const u64 need = dk_partials_need(count);
	if (need > cap) {
		dk_set_error("the launch needs %llu scratch doubles; the fixed "
			     "scratch holds %llu",
			     (unsigned long long)need, (unsigned long long)cap);
		return DK_E_ARG;
	}
	*out = scratch;
	return DK_OK;
}

// Ensures the shared partials buffer holds `count` values plus the room the
// chunked finishing pass needs after them.  Growth is stream-ordered.
int dk_partials_reserve(u64 count, double **out)
{
	const u64 need = dk_partials_need(count);
	if (need > dk_cur()->partials_cap) {
		double *p = nullptr;
		DK_TRY(dk_device_alloc(need * sizeof(double), (void **)&p));
		if (dk_cur()->partials)
			dk_device_free(dk_cur()->partials,
				       dk_cur()->partials_cap * sizeof(double));
		dk_cur()->partials = p;
		dk_cur()->partials_cap = need;
	}
	*out = dk_cur()->partials;
	return DK_OK;
}

#define DK_BY_TYPE(s, CALL)                                                  \
	do {                                                                 \
		if ((s)->dtype == DK_C64) {                                  \
			CALL(float2);                                        \
		} else {                                                     \
			CALL(double2);                                       \
		}                                                            \
	} while (0)

int dk_launch_prob_partials(const dk_state *s, unsigned q, int *nparts_out)
{
	const u64 half = s->len >> 1;
	const unsigned nb = grid_for(half, (u64)TPB * 4);
#define L(T)                                                                \
	hipLaunchKernelGGL(k_prob<T>, dim3(nb), dim3(TPB), 0, DK_STREAM,  \
			    (const T *)s->amp, s->scal, q, half, dk_cur()->partials)
	DK_BY_TYPE(s, L);
#undef L
	DK_LAUNCHED();
	*nparts_out = (int)nb;
	return DK_OK;
}

int dk_launch_prob(const dk_state *s, unsigned q, double *dst)
{
	int nb = 0;
	int rc = dk_launch_prob_partials(s, q, &nb);
	if (rc)
		return rc;
	hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(TPB), 0, DK_STREAM,
			   dk_cur()->partials, nb, dst, 0);
	DK_LAUNCHED();
	return DK_OK;
}

// Grid of the one-pass marginals kernel over 2^nq amplitudes: four times the
// usual grid cap, for more chunk loads in flight, and at least enough waves
// that none takes 2^PROB_KB chunks; a power of two, 2^wbits waves.  The
// join-and-marginals launch uses the same grid, so its sums are made in the
// same order.
// cap: most blocks (a power of two), 0 for four times the grid-stride cap.
static unsigned prob_all_grid(u64 len, int lu, unsigned *wbits, u64 cap = 0)
{
	static_assert((WAVES & (WAVES - 1)) == 0, "waves per block: a power of two");
	const u64 nch = len >> (6 + lu);
	u64 nb = std::min<u64>(std::max<u64>(1, nch / WAVES),
			       cap ? cap : 4ull * grid_for(~0ull >> 1, 1));
	nb = std::max<u64>(nb, (nch >> PROB_KB) / WAVES);
	*wbits = (unsigned)__builtin_ctzll(nb * WAVES);
	return (unsigned)nb;
}

static void launch_finish_marginals(const double *parts, unsigned nq,
				    unsigned nb, double *dst)
{
	hipLaunchKernelGGL(k_finish_cols, dim3(nq), dim3(TPB), 0, DK_STREAM,
			   parts, (int)nb, PQ, dst);
}

// Unroll bits of the marginals kernel (chunks of 64 * 2^LU amplitudes per
// wave step).  Reading a state, eight amplitudes per lane and step stream
// fastest (30 qubits: 2.48 ms, against 2.62 / 2.72 ms with four / two);
// writing a join, two do (32 qubits: 10.7 ms, against 12.4 ms with eight:
// fewer stores per lane between loop steps).  The two forms' sums then take
// different orders and agree to rounding.
constexpr int PROB_LU = 3;
constexpr int JOIN_PROB_LU = 1;
// Blocks of the one-pass join: with the marginals kernel's 8192 its write
// rate depended on where the state sat in physical memory (32 qubits: 13.0
// ms in a fresh process, 10.9 ms after the arena gave its chunks back and
// made new ones); with 2^17 blocks (64 loop steps per wave) 10.7-10.8 ms
// either way.  Partials: 2^17 x 40 doubles.
constexpr u64 JOIN_PROB_BLOCKS = 1ull << 17;

int dk_launch_prob_all(const dk_state *s, double *dst)
{
	constexpr int LU = PROB_LU;
	if (s->nq < 6 + LU || s->nq > PQ) {
		for (unsigned q = 0; q < s->nq; ++q) {
			int rc = dk_launch_prob(s, q, dst + q);
			if (rc)
				return rc;
		}
		return DK_OK;
	}
	unsigned wbits;
	const unsigned nb = prob_all_grid(s->len, LU, &wbits);
	double *parts;
	DK_TRY(dk_partials_reserve((u64)nb * PQ, &parts));
#define L(T)                                                                  \
	hipLaunchKernelGGL((k_prob_all<LU, ReadAmps<T>>), dim3(nb), dim3(TPB), \
			   0, DK_STREAM, ReadAmps<T>{(const T *)s->amp, s->scal}, \
			   s->len, wbits, parts, (double *)nullptr)
	DK_BY_TYPE(s, L);
#undef L
	DK_LAUNCHED();
	launch_finish_marginals(parts, s->nq, nb, dst);
	DK_LAUNCHED();
	return DK_OK;
}

int dk_launch_join_prob_all(const dk_state *a, const dk_state *b, dk_state *o,
			    double *dst)
{
	constexpr int LU = JOIN_PROB_LU;
	if (o->nq < 6 + LU || o->nq > PQ) {
		int rc = dk_launch_join(a, b, o);
		return rc ? rc : dk_launch_prob_all(o, dst);
	}
	unsigned wbits;
	const unsigned nb = prob_all_grid(o->len, LU, &wbits, JOIN_PROB_BLOCKS);
	double *parts;
	DK_TRY(dk_partials_reserve((u64)nb * PQ, &parts));
#define L(T)                                                                  \
	hipLaunchKernelGGL((k_prob_all<LU, JoinAmps<T>>), dim3(nb), dim3(TPB), \
			   0, DK_STREAM,                                       \
			   JoinAmps<T>{(const T *)a->amp, a->scal,              \
				       (const T *)b->amp, b->scal, (T *)o->amp, \
				       b->nq},                                  \
			   o->len, wbits, parts, o->scal)
	DK_BY_TYPE(o, L);
#undef L
	DK_LAUNCHED();
	launch_finish_marginals(parts, o->nq, nb, dst);
	DK_LAUNCHED();
	return DK_OK;
}

// out = in / norm per component (an IEEE division, as state_get divides,
// qstate.h:86), into a complex128 state whose norm becomes 1.
template <typename TI>
__global__ __launch_bounds__(TPB) void k_values_copy(
	const TI *__restrict__ in, const double *__restrict__ nrm,
	double2 *__restrict__ out, double *out_scal, u64 len)
{
	const double n = nrm[DK_SCAL_NORM];
	for (u64 i = (u64)blockIdx.x * TPB + threadIdx.x; i < len;
	     i += (u64)gridDim.x * TPB) {
		const TI a = ld_stream(in + i);
		double2 o;
		o.x = (double)a.x / n;
		o.y = (double)a.y / n;
		st_stream(out + i, o);
	}
	if (blockIdx.x == 0 && threadIdx.x == 0)
		out_scal[DK_SCAL_NORM] = 1.0;
}

int dk_launch_values_copy(const dk_state *src, dk_state *dst)
{
	const unsigned nb = grid_full(src->len, TPB);
	if (src->dtype == DK_C128)
		hipLaunchKernelGGL((k_values_copy<double2>), dim3(nb), dim3(TPB), 0,
				   DK_STREAM, src->a128(), src->scal, dst->a128(),
				   dst->scal, src->len);
	else
		hipLaunchKernelGGL((k_values_copy<float2>), dim3(nb), dim3(TPB), 0,
				   DK_STREAM, src->a64(), src->scal, dst->a128(),
				   dst->scal, src->len);
	DK_LAUNCHED();
	return DK_OK;
}

int dk_launch_normalised_copy(const dk_state *src, dk_state *dst)
{
	const unsigned nb = grid_full(src->len, TPB);
	hipStream_t st = DK_STREAM;
	if (src->dtype == DK_C128 && dst->dtype == DK_C128)
		hipLaunchKernelGGL((k_normalised_copy<double2, double2>), dim3(nb),
				   dim3(TPB), 0, st, src->a128(), src->scal,
				   dst->a128(), dst->scal, src->len);
	else if (src->dtype == DK_C128)
		hipLaunchKernelGGL((k_normalised_copy<double2, float2>), dim3(nb),
				   dim3(TPB), 0, st, src->a128(), src->scal,
				   dst->a64(), dst->scal, src->len);
	else if (dst->dtype == DK_C128)
		hipLaunchKernelGGL((k_normalised_copy<float2, double2>), dim3(nb),
				   dim3(TPB), 0, st, src->a64(), src->scal,
				   dst->a128(), dst->scal, src->len);
	else
		hipLaunchKernelGGL((k_normalised_copy<float2, float2>), dim3(nb),
				   dim3(TPB), 0, st, src->a64(), src->scal,
				   dst->a64(), dst->scal, src->len);
	DK_LAUNCHED();
	return DK_OK;
}

int dk_launch_widen(const dk_state *s, u64 first, u64 count, double2 *dst)
{
	const unsigned nb = grid_for(count, TPB * 4);
	hipLaunchKernelGGL(k_widen<float2>, dim3(nb), dim3(TPB), 0, DK_STREAM,
			   s->a64() + first, count, dst);
	DK_LAUNCHED();
	return DK_OK;
}

int dk_launch_decide(const double *parts, int nparts, double roll,
		     double *dec_slot, dk_state *next)
{
	hipLaunchKernelGGL(k_decide, dim3(1), dim3(TPB), 0, DK_STREAM, parts,
			   nparts, roll, dec_slot, next ? next->scal : nullptr);
	DK_LAUNCHED();
	return DK_OK;
}

int dk_launch_collapse(const dk_state *s, unsigned q, const double *dec_slot,
		       dk_state *o, int next_q, int *nparts_out)
{
	const u64 olen = s->len >> 1;
	const unsigned nb = grid_for(olen, (u64)TPB * 4);
#define L(T)                                                                  \
	hipLaunchKernelGGL(k_collapse<T>, dim3(nb), dim3(TPB), 0, DK_STREAM, \
			    (const T *)s->amp, s->scal, q, dec_slot, (T *)o->amp,  \
			    o->scal, next_q, olen, dk_cur()->partials)
	DK_BY_TYPE(s, L);
#undef L
	DK_LAUNCHED();
	*nparts_out = (int)nb;
	return DK_OK;
}

// A whole circuit of one-qubit gates on a small state in one workgroup: the
// state in LDS (read through its norm), one sweep over the index pairs per
// gate with a barrier between gates (the arithmetic of the one-qubit
// kernels), then written back with the norm of the result.  One launch
// instead of two (gate and norm) per gate.
constexpr int SMALL_THREADS = 1024;
constexpr int SMALL_LDS = 65536;

template <typename T>
__global__ __launch_bounds__(SMALL_THREADS) void k_small_run(
	const T *in, const double *__restrict__ in_scal, T *out, double *out_scal,
	const dk_small_gate *__restrict__ gates, unsigned ngates, unsigned nq)
{
	typedef typename real_of<T>::type R;
	__shared__ T st[SMALL_LDS / sizeof(T)];
	__shared__ double red[SMALL_THREADS / 64];
	const unsigned len = 1u << nq, half = len >> 1;
	const R inv = (R)(1.0 / in_scal[DK_SCAL_NORM]);
	for (unsigned i = threadIdx.x; i < len; i += SMALL_THREADS)
		st[i] = cscale(in[i], inv);
	__syncthreads();
	for (unsigned gi = 0; gi < ngates; ++gi) {
		const dk_small_gate &g = gates[gi];
		const unsigned t = g.t;
		const u64 cm = g.cm, am = g.am;
		const T m00 = mk<T>((R)g.m[0], (R)g.m[1]), m01 = mk<T>((R)g.m[2], (R)g.m[3]);
		const T m10 = mk<T>((R)g.m[4], (R)g.m[5]), m11 = mk<T>((R)g.m[6], (R)g.m[7]);
		for (unsigned p = threadIdx.x; p < half; p += SMALL_THREADS) {
			const unsigned i0 = ((p >> t) << (t + 1)) | (p & ((1u << t) - 1));
			const unsigned i1 = i0 | (1u << t);
			if (((u64)i0 & cm) == cm && ((u64)i0 & am) == 0) {
				const T x = st[i0], y = st[i1];
				st[i0] = cmac(cmac(czero<T>(), m00, x), m01, y);
				st[i1] = cmac(cmac(czero<T>(), m10, x), m11, y);
			}
		}
		__syncthreads();
	}
	double acc = 0.0;
	for (unsigned i = threadIdx.x; i < len; i += SMALL_THREADS) {
		out[i] = st[i];
		acc += cabs2d(st[i]);
	}
	acc = wave_sum(acc);
	if ((threadIdx.x & 63) == 0)
		red[threadIdx.x >> 6] = acc;
	__syncthreads();
	if (threadIdx.x == 0) {
		double s = 0.0;
		for (int w = 0; w < SMALL_THREADS / 64; ++w)
			s += red[w];
		out_scal[DK_SCAL_NORM] = sqrt(s);
	}
}

unsigned dk_small_qubits(int dtype)
{
	return dtype == DK_C64 ? 13 : 12;
}

int dk_launch_small_run(const dk_state *in, dk_state *out,
			const dk_small_gate *gates, unsigned ngates)
{
	if (in->nq > dk_small_qubits(in->dtype) || out->nq != in->nq ||
	    out->dtype != in->dtype) {
		dk_set_error("one-workgroup circuit on a state of %u qubits", in->nq);
		return DK_E_ARG;
	}
#define L(T)                                                                  \
	hipLaunchKernelGGL(k_small_run<T>, dim3(1), dim3(SMALL_THREADS), 0,       \
			   DK_STREAM, (const T *)in->amp, in->scal, (T *)out->amp, \
			   out->scal, gates, ngates, in->nq)
	DK_BY_TYPE(in, L);
#undef L
	DK_LAUNCHED();
	return DK_OK;
}

int dk_launch_join(const dk_state *a, const dk_state *b, dk_state *o)
{
	const unsigned nb = grid_full(o->len, TPB);
#define L(T)                                                                  \
	hipLaunchKernelGGL(k_join<T>, dim3(nb), dim3(TPB), 0, DK_STREAM,     \
			    (const T *)a->amp, a->scal, (const T *)b->amp, b->scal, \
			    b->nq, o->len, (T *)o->amp, o->scal)
	DK_BY_TYPE(o, L);
#undef L
	DK_LAUNCHED();
	return DK_OK;
}

int dk_launch_first_nonzero(const dk_state *s, u64 start, u64 *dst_dev)
{
	hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, DK_STREAM,
			   (unsigned long long *)dst_dev, s->len);
	DK_LAUNCHED();
	if (start >= s->len)
		return DK_OK;
	const unsigned nb = grid_for(s->len - start, TPB * 4);
#define L(T)                                                                  \
	hipLaunchKernelGGL(k_first_nonzero<T>, dim3(nb), dim3(TPB), 0,         \
			    DK_STREAM, (const T *)s->amp, start, s->len,         \
			    (unsigned long long *)dst_dev)
	DK_BY_TYPE(s, L);
#undef L
	DK_LAUNCHED();
	return DK_OK;
}

static void dens_args(const dk_state *s, u64 keep_mask, DensArgs &da, u64 tmask,
		      u64 tval)
{
	memset(&da, 0, sizeof(da));
	da.tmask = tmask;
	da.tval = tval & tmask;
	int a = 0;
	for (unsigned p = 0; p < s->nq; ++p)
		if ((keep_mask >> p) & 1)
			da.kpos[a++] = (unsigned char)p;
	da.a = a;
	da.m = (int)s->nq - a;
	for (int v = 0; v < 64; ++v) {
		u64 lo = 0, hi = 0;
		for (int b = 0; b < 6; ++b) {
			if (!((v >> b) & 1))
				continue;
			if (b < a)
				lo |= 1ull << da.kpos[b];
			if (6 + b < a)
				hi |= 1ull << da.kpos[6 + b];
		}
		da.keep_lo[v] = lo;
		da.keep_hi[v] = hi;
	}
}

// Gram-matrix form for a <= 6 kept qubits (see k_density_gram).
static int launch_density_gram(const dk_state *s, const DensArgs &da,
			       double2 *rho_dev)
{
	const int dim = 1 << da.a;
	const u64 ch = std::min<u64>(DENS_CH, s->len);
	const int cols = (int)(ch / (u64)dim);
	const u64 nchunks = s->len / ch;
	const u64 nblk = std::min<u64>(nchunks, (u64)std::max(1, g_ctx.num_cus) * 4);
	double2 *part = nullptr;
	const u64 part_bytes = nblk * (u64)dim * dim * sizeof(double2);
	DK_TRY(dk_device_alloc(part_bytes, (void **)&part));
#define G(T_, D_, R_)                                                          \
	hipLaunchKernelGGL((k_density_gram<T_, D_, R_>), dim3((unsigned)nblk),      \
			   dim3(TPB), 0, DK_STREAM, (const T_ *)s->amp, s->scal, da, \
			   nchunks, cols, part)
#define L(T_)                                                                  \
	switch (dim) {                                                         \
	case 1: G(T_, 1, 1); break;                                            \
	case 2: G(T_, 2, 1); break;                                            \
	case 4: G(T_, 4, 1); break;                                            \
	case 8: G(T_, 8, 2); break;                                            \
	case 16: G(T_, 16, 4); break;                                          \
	case 32: G(T_, 32, 4); break;                                          \
	default: G(T_, 64, 4); break;                                          \
	}
	DK_BY_TYPE(s, L);
#undef L
#undef G
	DK_LAUNCHED();
	const unsigned nb = grid_for((u64)dim * dim, TPB);
	hipLaunchKernelGGL(k_density_sum, dim3(nb), dim3(TPB), 0, DK_STREAM,
			   part, (int)nblk, (u64)dim * dim, rho_dev);
	DK_LAUNCHED();
	dk_device_free(part, part_bytes);
	return DK_OK;
}

int dk_launch_density(const dk_state *s, u64 keep_mask, double2 *rho_dev,
		      u64 tmask, u64 tval)
{
	DensArgs da;
	dens_args(s, keep_mask, da, tmask, tval);
	if (da.a <= 6)
		return launch_density_gram(s, da, rho_dev);
	const u64 dim = 1ull << da.a;
	const u64 T = 1ull << da.m;
	const int tpr = (int)((dim + 15) / 16);
	const u64 ntiles = (u64)tpr * tpr;
	// t-chunks: enough blocks to fill the device, partial buffer <= 256 MiB
	u64 want = std::max<u64>(1, (u64)std::max(1, g_ctx.num_cus) * 8 / ntiles);
	u64 maxc = std::max<u64>(1, (256ull << 20) / (dim * dim * 16));
	u64 nch = std::min<u64>(std::min<u64>(want, maxc), std::max<u64>(1, T / 16));
	u64 p = 1;
	while (p * 2 <= nch)
		p *= 2;
	nch = std::min<u64>(p, 65535);
	const u64 tchunk = T / nch;
	double2 *part = nullptr;
	const u64 part_bytes = nch * dim * dim * sizeof(double2);
	DK_TRY(dk_device_alloc(part_bytes, (void **)&part));
#define L(T_)                                                                  \
	hipLaunchKernelGGL(k_density_tile<T_>, dim3((unsigned)ntiles, (unsigned)nch), \
			    dim3(TPB), 0, DK_STREAM, (const T_ *)s->amp, s->scal, \
			    da, tchunk, tpr, part)
	DK_BY_TYPE(s, L);
#undef L
	DK_LAUNCHED();
	const unsigned nb = grid_for(dim * dim, TPB);
	hipLaunchKernelGGL(k_density_sum, dim3(nb), dim3(TPB), 0, DK_STREAM,
			   part, (int)nch, dim * dim, rho_dev);
	DK_LAUNCHED();
	dk_device_free(part, part_bytes);
	return DK_OK;
}

int dk_launch_density_elements(const dk_state *s, u64 keep_mask, size_t count,
			       const u64 *rows_dev, const u64 *cols_dev,
			       double2 *out_dev, u64 tmask, u64 tval)
{
	DensArgs da;
	dens_args(s, keep_mask, da, tmask, tval);
	if (count == 0)
		return DK_OK;
#define L(T)                                                                  \
	hipLaunchKernelGGL(k_density_elements<T>, dim3((unsigned)count),       \
			    dim3(TPB), 0, DK_STREAM, (const T *)s->amp, s->scal, \
			    da, rows_dev, cols_dev, out_dev)
	DK_BY_TYPE(s, L);
#undef L
	DK_LAUNCHED();
	return DK_OK;
}

int dk_launch_fill_zero_state(dk_state *s)
{
	const unsigned nb = grid_full(s->len, TPB);
#define L(T)                                                                  \
	hipLaunchKernelGGL(k_fill_zero_state<T>, dim3(nb), dim3(TPB), 0,       \
			    DK_STREAM, (T *)s->amp, s->scal, s->len)
	DK_BY_TYPE(s, L);
#undef L
	DK_LAUNCHED();
	return DK_OK;
}

// All-ones bytes (a NaN in every double and float) over a new allocation:
// the "poison_alloc" setting, which makes a read of memory no kernel wrote
// show in the results.  Words first, then the bytes of a ragged tail.
__global__ __launch_bounds__(TPB) void k_poison_bytes(unsigned *__restrict__ w, u64 words,
						      unsigned char *__restrict__ tail,
						      unsigned ntail)
{
	for (u64 i = (u64)blockIdx.x * TPB + threadIdx.x; i < words;
	     i += (u64)gridDim.x * TPB)
		w[i] = 0xffffffffu;
	if (blockIdx.x == 0 && threadIdx.x < ntail)
		tail[threadIdx.x] = 0xff;
}

int dk_launch_poison(void *p, u64 bytes)
{
	const u64 words = bytes / 4;
	const unsigned nb = grid_full(std::max<u64>(words, 1), TPB);
	hipLaunchKernelGGL(k_poison_bytes, dim3(nb), dim3(TPB), 0, DK_STREAM,
			   (unsigned *)p, words, (unsigned char *)p + 4 * words,
			   (unsigned)(bytes & 3));
	DK_LAUNCHED();
	return DK_OK;
}

// Guard bands around a block (the "guard_alloc" setting): a known word
// pattern before and after it, counted back when the block is released.
constexpr unsigned GUARD_WORD = 0xa5c3e1f7u;

__global__ __launch_bounds__(TPB) void k_guard_fill(unsigned *__restrict__ g, u64 words)
{
	for (u64 i = (u64)blockIdx.x * TPB + threadIdx.x; i < words;
	     i += (u64)gridDim.x * TPB)
		g[i] = GUARD_WORD;
}

__global__ __launch_bounds__(TPB) void k_guard_check(const unsigned *__restrict__ g,
						     u64 words, unsigned *__restrict__ bad)
{
	unsigned n = 0;
	for (u64 i = (u64)blockIdx.x * TPB + threadIdx.x; i < words;
	     i += (u64)gridDim.x * TPB)
		n += g[i] != GUARD_WORD;
	if (n)
		atomicAdd(bad, n);
}

int dk_launch_guard_fill(void *g, u64 bytes)
{
	hipLaunchKernelGGL(k_guard_fill, dim3(64), dim3(TPB), 0, DK_STREAM,
			   (unsigned *)g, bytes / 4);
	DK_LAUNCHED();
	return DK_OK;
}

int dk_launch_guard_check(const void *g, u64 bytes, unsigned *bad)
{
	hipLaunchKernelGGL(k_guard_check, dim3(64), dim3(TPB), 0, DK_STREAM,
			   (const unsigned *)g, bytes / 4, bad);
	DK_LAUNCHED();
	return DK_OK;
}

int dk_launch_set_norm(dk_state *s, double v)
{
	hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(64), 0, DK_STREAM,
			   s->scal + DK_SCAL_NORM, v);
	DK_LAUNCHED();
	return DK_OK;
}

// dst[0..2] (device) <- re and im of sum conj(a_i) b_i, max |a_i - b_i|, over
// normalised amplitudes of two states of one size.
int dk_launch_compare(const dk_state *a, const dk_state *b, double *dst)
{
	const int nblocks = (int)std::min<u64>(std::max(1, g_ctx.num_cus) * 8ull,
					       std::max<u64>(1, a->len / TPB));
	double *parts = nullptr;
	DK_TRY(dk_partials_reserve(3ull * (u64)nblocks, &parts));
	hipStream_t st = DK_STREAM;
#define DK_CMP(TA, TB)                                                          \
	hipLaunchKernelGGL((k_compare<TA, TB>), dim3(nblocks), dim3(TPB), 0, st,  \
			   (const TA *)a->amp, (const TB *)b->amp, a->scal, b->scal, \
			   a->len, parts)
	if (a->dtype == DK_C64 && b->dtype == DK_C64)
		DK_CMP(float2, float2);
	else if (a->dtype == DK_C64)
		DK_CMP(float2, double2);
	else if (b->dtype == DK_C64)
		DK_CMP(double2, float2);
	else
		DK_CMP(double2, double2);
#undef DK_CMP
	DK_LAUNCHED();
	hipLaunchKernelGGL(k_compare_finish, dim3(1), dim3(TPB), 0, st, parts, nblocks, dst);
	DK_LAUNCHED();
	return DK_OK;
}

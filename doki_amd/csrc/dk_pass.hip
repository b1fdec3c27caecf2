// Cache-blocked multi-gate pass kernel (see dk_pass.h for the tile, the
// mapping and the op encoding).
//
// Per tile: every thread loads its 16 amplitudes (scaled by 1/norm, the
// reference's state_get, qstate.h:86), runs the pass's ops on them in
// registers, and stores them back with one partial |out|^2 sum per tile.
// Each gate keeps registry_apply's semantics (qops.c:176-261): amplitudes
// whose controls are off keep their value.  The state is renormalised once
// per pass instead of once per gate; all gates are unitary, so the values
// agree with per-gate renormalisation to rounding.
//
// In place is safe: a workgroup has read its whole tile before it writes.
#include <algorithm>

#include "dk_device.h"
#include "dk_pass.h"

using namespace dkd;
using namespace dkp;

namespace {

// A complex value as two reals (register arrays of these promote to
// registers reliably; the HIP vector types do not always).
template <typename R> struct Cx {
	R re, im;
};

template <typename R> struct M22 {
	Cx<R> m00, m01, m10, m11;
};

template <typename R>
__device__ __forceinline__ Cx<R> cx_sel(bool c, const Cx<R> &a, const Cx<R> &b)
{
	return {c ? a.re : b.re, c ? a.im : b.im};
}

// m * (xr + i xi)
template <typename R>
__device__ __forceinline__ void cx_mul(const Cx<R> &m, R xr, R xi, R &orr, R &oi)
{
	orr = fma(m.re, xr, -m.im * xi);
	oi = fma(m.re, xi, m.im * xr);
}

// m0 * x + m1 * y
template <typename R>
__device__ __forceinline__ void row2(const Cx<R> &m0, R xr, R xi, const Cx<R> &m1,
				     R yr, R yi, R &orr, R &oi)
{
	R ar, ai;
	cx_mul(m0, xr, xi, ar, ai);
	ar = fma(m1.re, yr, ar);
	ar = fma(-m1.im, yi, ar);
	ai = fma(m1.re, yi, ai);
	ai = fma(m1.im, yr, ai);
	orr = ar;
	oi = ai;
}

// Whether register r passes the register-bit part of a control condition
// (uniform: r is a compile-time index, rmask an op constant).
__device__ __forceinline__ bool reg_ok(int r, unsigned rmask)
{
	const unsigned rcm = rmask & 0xff, ram = rmask >> 8;
	return ((unsigned)r & rcm) == rcm && ((unsigned)r & ram) == 0;
}

// Control modes of a gate op:
//   CM_NONE: no controls, every amplitude takes the new value;
//   CM_REG: the thread-word part of the condition holds for the whole wave
//     (checked by the caller) and registers whose register-bit part fails
//     are skipped (a uniform branch);
//   CM_LANE: the condition varies across lanes: selects per lane, and
//     registers whose register-bit part fails are skipped.
enum { CM_NONE = 0, CM_REG = 1, CM_LANE = 2 };

// Body of a register that a control may skip: the empty volatile asm keeps
// the uniform skip a branch (if-converted, every register would be computed
// and the old values selected back).
#define DK_KEEP_BRANCH() asm volatile("")

// Dense 2x2 gate on register bit J.
template <int J, int CM, typename R>
__device__ __forceinline__ void gate_reg(R (&vr)[REGS], R (&vi)[REGS],
					 const M22<R> &M, bool ok, unsigned rmask)
{
#pragma unroll
	for (int r = 0; r < REGS; ++r) {
		if (r & (1 << J))
			continue;
		if (CM != CM_NONE && !reg_ok(r, rmask))
			continue;
		constexpr int d = 1 << J;
		R o0r, o0i, o1r, o1i;
		row2(M.m00, vr[r], vi[r], M.m01, vr[r | d], vi[r | d], o0r, o0i);
		row2(M.m10, vr[r], vi[r], M.m11, vr[r | d], vi[r | d], o1r, o1i);
		if (CM == CM_LANE) {
			o0r = ok ? o0r : vr[r];
			o0i = ok ? o0i : vi[r];
			o1r = ok ? o1r : vr[r | d];
			o1i = ok ? o1i : vi[r | d];
		}
		vr[r] = o0r;
		vi[r] = o0i;
		vr[r | d] = o1r;
		vi[r | d] = o1i;
		if (CM != CM_NONE)
			DK_KEEP_BRANCH();
	}
}

template <int CM, typename R>
__device__ __forceinline__ void gate_reg_any(R (&vr)[REGS], R (&vi)[REGS],
					     const M22<R> &M, bool ok, unsigned rmask,
					     unsigned j)
{
	// the empty asm markers keep the compiler from merging the four bodies
	// into one with a variable register index (which would move the tile to
	// scratch memory)
	switch (j) {
	case 0:
		gate_reg<0, CM>(vr, vi, M, ok, rmask);
		asm volatile("; register bit 0");
		break;
	case 1:
		gate_reg<1, CM>(vr, vi, M, ok, rmask);
		asm volatile("; register bit 1");
		break;
	case 2:
		gate_reg<2, CM>(vr, vi, M, ok, rmask);
		asm volatile("; register bit 2");
		break;
	default:
		gate_reg<3, CM>(vr, vi, M, ok, rmask);
		asm volatile("; register bit 3");
		break;
	}
}

// Dense 2x2 gate on register bit J whose matrix is M0 where register bit K
// is 0 and M1 where it is 1 (a phase on |11> absorbed into a gate on one of
// its qubits).
template <int J, int K, typename R>
__device__ __forceinline__ void gate_reg_mux(R (&vr)[REGS], R (&vi)[REGS],
					     const M22<R> &M0, const M22<R> &M1)
{
#pragma unroll
	for (int r = 0; r < REGS; ++r) {
		if (r & (1 << J))
			continue;
		constexpr int d = 1 << J;
		const M22<R> &M = ((r >> K) & 1) ? M1 : M0;
		R o0r, o0i, o1r, o1i;
		row2(M.m00, vr[r], vi[r], M.m01, vr[r | d], vi[r | d], o0r, o0i);
		row2(M.m10, vr[r], vi[r], M.m11, vr[r | d], vi[r | d], o1r, o1i);
		vr[r] = o0r;
		vi[r] = o0i;
		vr[r | d] = o1r;
		vi[r | d] = o1i;
	}
}

template <int J, typename R>
__device__ __forceinline__ void gate_reg_mux_k(R (&vr)[REGS], R (&vi)[REGS],
					       const M22<R> &M0, const M22<R> &M1,
					       unsigned k)
{
	// k != J (checked by the planner's validator); distinct asm markers:
	// see gate_reg_any
	switch (k) {
	case 0:
		if constexpr (J != 0)
			gate_reg_mux<J, 0>(vr, vi, M0, M1);
		asm volatile("; mux selector bit 0");
		break;
	case 1:
		if constexpr (J != 1)
			gate_reg_mux<J, 1>(vr, vi, M0, M1);
		asm volatile("; mux selector bit 1");
		break;
	case 2:
		if constexpr (J != 2)
			gate_reg_mux<J, 2>(vr, vi, M0, M1);
		asm volatile("; mux selector bit 2");
		break;
	default:
		if constexpr (J != 3)
			gate_reg_mux<J, 3>(vr, vi, M0, M1);
		asm volatile("; mux selector bit 3");
		break;
	}
}

template <typename R>
__device__ __forceinline__ void gate_reg_mux_any(R (&vr)[REGS], R (&vi)[REGS],
						 const M22<R> &M0, const M22<R> &M1,
						 unsigned j, unsigned k)
{
	switch (j) {
	case 0:
		gate_reg_mux_k<0>(vr, vi, M0, M1, k);
		asm volatile("; mux register bit 0");
		break;
	case 1:
		gate_reg_mux_k<1>(vr, vi, M0, M1, k);
		asm volatile("; mux register bit 1");
		break;
	case 2:
		gate_reg_mux_k<2>(vr, vi, M0, M1, k);
		asm volatile("; mux register bit 2");
		break;
	default:
		gate_reg_mux_k<3>(vr, vi, M0, M1, k);
		asm volatile("; mux register bit 3");
		break;
	}
}

// The value of lane ^ 2^B.  B < 2: one DPP quad permutation, no LDS
// traffic.  B >= 2: ds_bpermute, which measured faster on gfx950 than DPP
// mirror pairs (two dependent moves with hazard waits) and than the
// permlane swaps (copies, a select and hazard waits per dword).
template <int B>
__device__ __forceinline__ int lane_xor_i(int v, unsigned lane)
{
	if constexpr (B == 0)
		return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false); // [1,0,3,2]
	else if constexpr (B == 1)
		return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false); // [2,3,0,1]
	else
		return __builtin_amdgcn_ds_bpermute((int)((lane ^ (1u << B)) << 2), v);
}

template <int B>
__device__ __forceinline__ double lane_xor(double v, unsigned lane)
{
	const int lo = lane_xor_i<B>(__double2loint(v), lane);
	const int hi = lane_xor_i<B>(__double2hiint(v), lane);
	return __hiloint2double(hi, lo);
}

template <int B>
__device__ __forceinline__ float lane_xor(float v, unsigned lane)
{
	return __int_as_float(lane_xor_i<B>(__float_as_int(v), lane));
}

// Dense 2x2 gate on lane bit B: the partner amplitude is in lane ^ 2^B.
// Every lane takes part in the exchange; CM_LANE selects afterwards.
template <int B, int CM, typename R>
__device__ __forceinline__ void gate_lane(R (&vr)[REGS], R (&vi)[REGS],
					  const M22<R> &M, bool ok, unsigned rmask,
					  unsigned lane)
{
	const bool upper = (lane >> B) & 1;
	const Cx<R> a0 = cx_sel(upper, M.m11, M.m00); // own amplitude
	const Cx<R> a1 = cx_sel(upper, M.m10, M.m01); // partner's
#pragma unroll
	for (int r = 0; r < REGS; ++r) {
		if (CM != CM_NONE && !reg_ok(r, rmask))
			continue;
		const R pr = lane_xor<B>(vr[r], lane), pi = lane_xor<B>(vi[r], lane);
		R orr, oi;
		row2(a0, vr[r], vi[r], a1, pr, pi, orr, oi);
		if (CM == CM_LANE) {
			orr = ok ? orr : vr[r];
			oi = ok ? oi : vi[r];
		}
		vr[r] = orr;
		vi[r] = oi;
		if (CM != CM_NONE)
			DK_KEEP_BRANCH();
	}
}

template <int CM, typename R>
__device__ __forceinline__ void gate_lane_any(R (&vr)[REGS], R (&vi)[REGS],
					      const M22<R> &M, bool ok, unsigned rmask,
					      unsigned b, unsigned lane)
{
	// distinct asm markers: see gate_reg_any
	switch (b) {
	case 0:
		gate_lane<0, CM>(vr, vi, M, ok, rmask, lane);
		asm volatile("; lane bit 0");
		break;
	case 1:
		gate_lane<1, CM>(vr, vi, M, ok, rmask, lane);
		asm volatile("; lane bit 1");
		break;
	case 2:
		gate_lane<2, CM>(vr, vi, M, ok, rmask, lane);
		asm volatile("; lane bit 2");
		break;
	case 3:
		gate_lane<3, CM>(vr, vi, M, ok, rmask, lane);
		asm volatile("; lane bit 3");
		break;
	case 4:
		gate_lane<4, CM>(vr, vi, M, ok, rmask, lane);
		asm volatile("; lane bit 4");
		break;
	default:
		gate_lane<5, CM>(vr, vi, M, ok, rmask, lane);
		asm volatile("; lane bit 5");
		break;
	}
}

// One diagonal op: a sign change (NEG) or a complex factor per register.
template <int CM, bool REG, bool NEG, typename R>
__device__ __forceinline__ void gate_diag_impl(R (&vr)[REGS], R (&vi)[REGS],
					       const M22<R> &M, bool ok, unsigned rmask,
					       unsigned a, u64 tb)
{
	const Cx<R> ft = cx_sel((tb >> a) & 1, M.m11, M.m00);
#pragma unroll
	for (int r = 0; r < REGS; ++r) {
		if (CM != CM_NONE && !reg_ok(r, rmask))
			continue;
		R orr, oi;
		if (NEG) {
			orr = -vr[r];
			oi = -vi[r];
		} else {
			const Cx<R> f = REG ? cx_sel((r >> a) & 1, M.m11, M.m00) : ft;
			cx_mul(f, vr[r], vi[r], orr, oi);
		}
		if (CM == CM_LANE) {
			orr = ok ? orr : vr[r];
			oi = ok ? oi : vi[r];
		}
		vr[r] = orr;
		vi[r] = oi;
		if (CM != CM_NONE)
			DK_KEEP_BRANCH();
	}
}

// Diagonal gate: factor m00 where the qubit is 0, m11 where it is 1.  The
// qubit is register bit a (REG) or bit a of the thread word.  (The planner
// sends a phase gate, m00 = 1, as the factor m11 controlled on its own
// target, so the control modes skip the amplitudes it leaves alone; a
// factor of -1 on both sides, as Z and CZ become, is a sign change.)
template <int CM, bool REG, typename R>
__device__ __forceinline__ void gate_diag(R (&vr)[REGS], R (&vi)[REGS],
					  const M22<R> &M, bool ok, unsigned rmask,
					  unsigned a, u64 tb)
{
	const bool neg = M.m00.re == (R)-1 && M.m00.im == (R)0 && M.m11.re == (R)-1 &&
			 M.m11.im == (R)0;
	if (neg)
		gate_diag_impl<CM, REG, true>(vr, vi, M, ok, rmask, a, tb);
	else
		gate_diag_impl<CM, REG, false>(vr, vi, M, ok, rmask, a, tb);
}

template <typename R>
__device__ __forceinline__ M22<R> op_matrix(const PassOp *__restrict__ op)
{
	M22<R> M;
	M.m00 = {(R)op->m[0], (R)op->m[1]};
	M.m01 = {(R)op->m[2], (R)op->m[3]};
	M.m10 = {(R)op->m[4], (R)op->m[5]};
	M.m11 = {(R)op->m[6], (R)op->m[7]};
	return M;
}

// One gate op.
template <int CM, typename R>
__device__ __forceinline__ void run_gate(R (&vr)[REGS], R (&vi)[REGS], unsigned kind,
					 unsigned a, unsigned rmask, const M22<R> &M,
					 bool ok, u64 tb, unsigned lane)
{
	switch (kind) {
	case OP_REG:
		gate_reg_any<CM>(vr, vi, M, ok, rmask, a);
		break;
	case OP_LANE:
		gate_lane_any<CM>(vr, vi, M, ok, rmask, a, lane);
		break;
	case OP_DIAG_REG:
		gate_diag<CM, true>(vr, vi, M, ok, rmask, a, tb);
		break;
	default: // OP_DIAG_THREAD
		gate_diag<CM, false>(vr, vi, M, ok, rmask, a, tb);
		break;
	}
}

// High slot bits held by register r / by wave w under a mapping.
__device__ __forceinline__ unsigned reg_slots(int r, const unsigned (&rs)[REG_BITS])
{
	unsigned s = 0;
#pragma unroll
	for (int j = 0; j < REG_BITS; ++j)
		if ((r >> j) & 1)
			s |= 1u << rs[j];
	return s;
}

__device__ __forceinline__ unsigned wave_slots(unsigned w, const unsigned (&ws)[WAVE_BITS])
{
	unsigned s = 0;
#pragma unroll
	for (int b = 0; b < WAVE_BITS; ++b)
		if ((w >> b) & 1)
			s |= 1u << ws[b];
	return s;
}

// A 6-bit slot set without slot bit h (higher slot bits shift down).
__device__ __forceinline__ unsigned squeeze(unsigned slots, unsigned h)
{
	return (slots & ((1u << h) - 1)) | ((slots >> (h + 1)) << h);
}

// Index bits of the high slot bits in `slots`.
__device__ __forceinline__ u64 slot_index(unsigned slots, const u64 (&hb)[HIGH_BITS])
{
	u64 x = 0;
#pragma unroll
	for (int h = 0; h < HIGH_BITS; ++h)
		if ((slots >> h) & 1)
			x |= hb[h];
	return x;
}

// Exchange buffer.  complex128: half a tile (32 KiB; exchanges run in two
// rounds), so that LDS admits the three workgroups per CU the kernel's
// registers allow; complex64: the whole tile (32 KiB; one round), since its
// registers already cap the CU at four workgroups.  32 KiB exactly: any
// other LDS variable would cost a workgroup per CU.
template <typename T> constexpr bool exch_halves() { return sizeof(T) == 16; }
template <typename T> constexpr int exch_entries()
{
	return 1 << (PASS_TILE_BITS - (exch_halves<T>() ? 1 : 0));
}

// Index bits of a tile's fixed part: the tile number spread over the index
// bits outside the tile.
__device__ __forceinline__ u64 tile_base(u64 tile, const PassDesc &pd)
{
	u64 base = tile << LANE_BITS;
#pragma unroll
	for (int h = 0; h < HIGH_BITS; ++h)
		base = ins0(base, pd.hsorted[h]);
	return base;
}

// A thread's 16 amplitudes of the tile at `base` in the load mapping
// (register bit j = high slot j, wave bit b = slot 4 + b), as stored.
template <typename T, typename R>
__device__ __forceinline__ void tile_fetch(const T *in, u64 base,
					   const u64 (&hb)[HIGH_BITS], unsigned lane,
					   unsigned w, R (&xr)[REGS], R (&xi)[REGS])
{
	const u64 tb = base | lane | slot_index(w << REG_BITS, hb);
#pragma unroll
	for (int r = 0; r < REGS; ++r) {
		const T x = ld_stream(in + (tb | slot_index((unsigned)r, hb)));
		xr[r] = x.x;
		xi[r] = x.y;
	}
}

// tile_fetch taking as zero every amplitude whose index meets kmask (index
// bits known to be zero, whose memory may never have been written): such
// amplitudes are not read.  The register part of an index is the same for
// the whole workgroup, so the per-register test is uniform.
template <typename T, typename R>
__device__ __forceinline__ void tile_fetch_known(const T *in, u64 base,
						 const u64 (&hb)[HIGH_BITS], unsigned lane,
						 unsigned w, u64 kmask, R (&xr)[REGS],
						 R (&xi)[REGS])
{
	const u64 tb = base | lane | slot_index(w << REG_BITS, hb);
	const bool none = (tb & kmask) != 0;
#pragma unroll
	for (int r = 0; r < REGS; ++r) {
		const u64 ri = slot_index((unsigned)r, hb);
		xr[r] = xi[r] = (R)0;
		if (!none && !(ri & kmask)) {
			const T x = ld_stream(in + (tb | ri));
			xr[r] = x.x;
			xi[r] = x.y;
		}
	}
}

// Zeros into a thread's 16 amplitudes of the tile at `base` (load mapping).
template <typename T>
__device__ __forceinline__ void tile_zero(T *out, u64 base, const u64 (&hb)[HIGH_BITS],
					  unsigned lane, unsigned w)
{
	const u64 tb = base | lane | slot_index(w << REG_BITS, hb);
#pragma unroll
	for (int r = 0; r < REGS; ++r)
		st_stream(out + (tb | slot_index((unsigned)r, hb)), czero<T>());
}

// The pass's ops on one tile held in vr/vi (normalised values, load
// mapping), then the store and the tile's partial |out|^2 sum.
template <typename T, typename R>
__device__ __forceinline__ void tile_run(
	T *out, const PassDesc &pd, const PassOp *__restrict__ ops,
	double *__restrict__ partials, u64 tile, u64 base,
	const u64 (&hb)[HIGH_BITS], unsigned lane, unsigned w, R (&vr)[REGS],
	R (&vi)[REGS], T *lds, double *red)
{
	constexpr bool HALVES = exch_halves<T>();
	// mapping at load: register bit j = high slot j, wave bit b = slot 4 + b
	unsigned rs[REG_BITS], ws[WAVE_BITS];
#pragma unroll
	for (int j = 0; j < REG_BITS; ++j)
		rs[j] = (unsigned)j;
#pragma unroll
	for (int b = 0; b < WAVE_BITS; ++b)
		ws[b] = (unsigned)(REG_BITS + b);

	// the thread word: index bits shared by all of the thread's registers
	u64 tb = base | lane | slot_index(wave_slots(w, ws), hb);

	for (unsigned i = 0; i < pd.nops; ++i) {
		const PassOp *__restrict__ op = ops + i;
		const unsigned kind = op->kind, a = op->a, rmask = op->rmask;
		const u64 tcm = op->tcm, tam = op->tam;
		if (kind == OP_EXCH) {
			const unsigned b = op->b;
			unsigned nrs[REG_BITS], nws[WAVE_BITS];
#pragma unroll
			for (int j = 0; j < REG_BITS; ++j)
				nrs[j] = (a >> (3 * j)) & 7;
#pragma unroll
			for (int k = 0; k < WAVE_BITS; ++k)
				nws[k] = (b >> (3 * k)) & 7;
			const unsigned wsl0 = wave_slots(w, ws), wsl1 = wave_slots(w, nws);
			if constexpr (HALVES) {
				// Register bit js holds the same high slot hs before
				// and after (the planner guarantees one): the
				// amplitudes with slot hs = 0 and those with hs = 1
				// are exchanged in two rounds, each by the 8
				// registers whose bit js is that value, before and
				// after.
				unsigned js = 0;
#pragma unroll
				for (int j = REG_BITS - 1; j >= 0; --j)
					if (rs[j] == nrs[j])
						js = (unsigned)j;
				const unsigned hs = rs[js];
#pragma unroll
				for (unsigned half = 0; half < 2; ++half) {
					__syncthreads(); // earlier reads of the buffer are done
#pragma unroll
					for (int r = 0; r < REGS; ++r)
						if (((r >> js) & 1) == half)
							lds[lane | (squeeze(wsl0 | reg_slots(r, rs), hs)
								    << LANE_BITS)] = mk<T>(vr[r], vi[r]);
					__syncthreads();
#pragma unroll
					for (int r = 0; r < REGS; ++r)
						if (((r >> js) & 1) == half) {
							const T x = lds[lane | (squeeze(wsl1 | reg_slots(r, nrs), hs)
										<< LANE_BITS)];
							vr[r] = x.x;
							vi[r] = x.y;
						}
				}
			} else {
				__syncthreads(); // the previous exchange's reads are done
#pragma unroll
				for (int r = 0; r < REGS; ++r)
					lds[lane | ((wsl0 | reg_slots(r, rs)) << LANE_BITS)] =
						mk<T>(vr[r], vi[r]);
				__syncthreads();
#pragma unroll
				for (int r = 0; r < REGS; ++r) {
					const T x = lds[lane | ((wsl1 | reg_slots(r, nrs)) << LANE_BITS)];
					vr[r] = x.x;
					vi[r] = x.y;
				}
			}
#pragma unroll
			for (int j = 0; j < REG_BITS; ++j)
				rs[j] = nrs[j];
#pragma unroll
			for (int k = 0; k < WAVE_BITS; ++k)
				ws[k] = nws[k];
			tb = base | lane | slot_index(wave_slots(w, ws), hb);
			continue;
		}
		if (kind == OP_REG_MUX || kind == OP_LANE_MUX) {
			// two matrices (this record and the next), chosen by a
			// selector qubit; no controls
			const M22<R> M0 = op_matrix<R>(op), M1 = op_matrix<R>(op + 1);
			const unsigned sel = op->b;
			++i;
			if (sel & SEL_REG) {
				gate_reg_mux_any(vr, vi, M0, M1, a, sel & (REG_BITS - 1));
			} else {
				// a thread-word bit >= 6: one value for the whole wave
				const bool one = __builtin_amdgcn_readfirstlane((int)((tb >> sel) & 1));
				const M22<R> M = one ? M1 : M0;
				if (kind == OP_REG_MUX)
					gate_reg_any<CM_NONE>(vr, vi, M, true, 0, a);
				else
					gate_lane_any<CM_NONE>(vr, vi, M, true, 0, a, lane);
			}
			continue;
		}
		const M22<R> M = op_matrix<R>(op);
		const bool ok = ((tb & tcm) == tcm) & ((tb & tam) == 0);
		if ((tcm | tam | rmask) == 0) {
			run_gate<CM_NONE>(vr, vi, kind, a, rmask, M, true, tb, lane);
		} else if (((tcm | tam) & ((1u << LANE_BITS) - 1)) == 0) {
			// the thread-word condition reads only index bits >= 6: the
			// same for the whole wave
			if (__builtin_amdgcn_readfirstlane((int)ok))
				run_gate<CM_REG>(vr, vi, kind, a, rmask, M, true, tb, lane);
		} else {
			run_gate<CM_LANE>(vr, vi, kind, a, rmask, M, ok, tb, lane);
		}
	}

	double s = 0.0;
#pragma unroll
	for (int r = 0; r < REGS; ++r) {
		st_stream(out + (tb | slot_index(reg_slots(r, rs), hb)), mk<T>(vr[r], vi[r]));
		s += fma((double)vr[r], (double)vr[r], (double)vi[r] * (double)vi[r]);
	}
	if (!partials)
		return;
	s = wave_sum(s);
	__syncthreads(); // red aliases the exchange buffer: its reads are done
	if (lane == 0)
		red[w] = s;
	__syncthreads();
	if (threadIdx.x == 0) {
		double t = 0.0;
#pragma unroll
		for (int ww = 0; ww < THREADS / 64; ++ww)
			t += red[ww];
		partials[tile] = t;
	}
}

// One tile, loaded on the spot.
template <typename T>
__device__ __forceinline__ void pass_tile(
	const T *in, T *out, typename real_of<T>::type inv, const PassDesc &pd,
	const PassOp *__restrict__ ops, double *__restrict__ partials, u64 tile,
	T *lds, double *red)
{
	typedef typename real_of<T>::type R;
	const unsigned lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	u64 hb[HIGH_BITS];
#pragma unroll
	for (int h = 0; h < HIGH_BITS; ++h)
		hb[h] = 1ull << pd.hpos[h];
	const u64 base = tile_base(tile, pd);
	if (base & pd.zmask) {
		// a tile of zeros stays zeros: no read, no ops
		if (pd.zwrite)
			tile_zero(out, base, hb, lane, w);
		if (partials && threadIdx.x == 0)
			partials[tile] = 0.0;
		return;
	}
	R vr[REGS], vi[REGS];
	tile_fetch(in, base, hb, lane, w, vr, vi);
#pragma unroll
	for (int r = 0; r < REGS; ++r) {
		vr[r] *= inv;
		vi[r] *= inv;
	}
	tile_run<T>(out, pd, ops, partials, tile, base, hb, lane, w, vr, vi, lds, red);
}

// Occupancy: the exchange buffer is the kernel's only LDS (32 KiB: five
// workgroups per CU), so registers set the waves per SIMD.  complex128 asks
// for three: the op code then fits 168 VGPRs without spilling (left alone,
// the round kernel took 201 and ran at two; four waves spill the tile).
// complex64 runs at four in about 110 VGPRs; five spill (measured slower).
template <typename T> constexpr int pass_waves() { return sizeof(T) == 16 ? 3 : 4; }

// LOOP: tiles beyond the grid (launches of more than 2^31 threads).
template <typename T, bool LOOP>
__global__ __launch_bounds__(THREADS)
__attribute__((amdgpu_waves_per_eu(pass_waves<T>()))) void k_pass(
	const T *in, T *out, const double *__restrict__ nrm, PassDesc pd,
	const PassOp *__restrict__ ops, double *__restrict__ partials, u64 ntiles)
{
	typedef typename real_of<T>::type R;
	__shared__ T lds[exch_entries<T>()];
	// the per-wave partial sums reuse the exchange buffer
	double *red = reinterpret_cast<double *>(lds);
	const R inv = (R)(1.0 / nrm[DK_SCAL_NORM]);
	if constexpr (LOOP) {
		for (u64 t = blockIdx.x; t < ntiles; t += gridDim.x) {
			pass_tile<T>(in, out, inv, pd, ops, partials, t, lds, red);
			__syncthreads();
		}
	} else {
		pass_tile<T>(in, out, inv, pd, ops, partials, blockIdx.x, lds, red);
	}
}

// One round of passes over many registers held in a pool: item `it` is
// tile it - first of the entry holding it (entries ascend by `first`);
// workgroups take items blockIdx.x, blockIdx.x + gridDim.x, ...  The launch
// boundary orders one round's passes before the next round's.  `init`: the
// first round, whose tiles start from |0...0> instead of the pool.  An
// entry's kmask marks the amplitudes its register has not written yet (the
// pool is not cleared): they are zeros, taken as such without a read.  No
// renormalisation between passes (the caller normalises at the end).
template <typename T>
__global__ __launch_bounds__(THREADS)
__attribute__((amdgpu_waves_per_eu(pass_waves<T>()))) void k_pass_round(
	T *pool, const PassEntry *__restrict__ ents, unsigned nents,
	const PassOp *__restrict__ ops, u64 nitems, int init)
{
	typedef typename real_of<T>::type R;
	__shared__ T lds[exch_entries<T>()];
	const unsigned lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	for (u64 it = blockIdx.x; it < nitems; it += gridDim.x) {
		// the entry holding item `it`, found by every wave on its own
		// (uniform values: scalar loads, no LDS, no barrier)
		unsigned lo = 0, hi = nents - 1;
		while (lo < hi) {
			const unsigned mid = (lo + hi + 1) / 2;
			if (ents[mid].first <= it)
				lo = mid;
			else
				hi = mid - 1;
		}
		const unsigned e = __builtin_amdgcn_readfirstlane(lo);
		const PassDesc pd = ents[e].pd;
		T *st = pool + ents[e].off;
		const PassOp *eops = ops + ents[e].op_begin;
		const u64 tile = it - ents[e].first;
		u64 hb[HIGH_BITS];
#pragma unroll
		for (int h = 0; h < HIGH_BITS; ++h)
			hb[h] = 1ull << pd.hpos[h];
		const u64 base = tile_base(tile, pd);
		if (base & pd.zmask) {
			// uniform for the workgroup: every thread skips the barrier
			if (pd.zwrite)
				tile_zero(st, base, hb, lane, w);
			continue;
		}
		R vr[REGS], vi[REGS];
		if (init) {
#pragma unroll
			for (int r = 0; r < REGS; ++r)
				vr[r] = vi[r] = (R)0;
			// amplitude 0: tile 0, lane 0, wave 0, register 0
			if (base == 0 && lane == 0 && w == 0)
				vr[0] = (R)1;
		} else if (pd.kmask) {
			tile_fetch_known(st, base, hb, lane, w, pd.kmask, vr, vi);
		} else {
			tile_fetch(st, base, hb, lane, w, vr, vi);
		}
		tile_run<T>(st, pd, eops, nullptr, tile, base, hb, lane, w, vr, vi,
			    lds, nullptr);
		__syncthreads();
	}
}

template <typename T>
int launch_pass_typed(const dk_state *in, dk_state *out, const PassDesc &pd,
		      const PassOp *ops_dev, double *parts, u64 ntiles)
{
	const unsigned gr = dk_grid_cap(ntiles, THREADS);
	if (gr < ntiles)
		hipLaunchKernelGGL((k_pass<T, true>), dim3(gr), dim3(THREADS), 0,
				   DK_STREAM, (const T *)in->amp, (T *)out->amp,
				   in->scal, pd, ops_dev, parts, ntiles);
	else
		hipLaunchKernelGGL((k_pass<T, false>), dim3(gr), dim3(THREADS), 0,
				   DK_STREAM, (const T *)in->amp, (T *)out->amp,
				   in->scal, pd, ops_dev, parts, ntiles);
	DK_LAUNCHED();
	return DK_OK;
}

} // namespace

// One pass over `in` into `out` (which may be `in`): the ops at ops_dev
// (device memory, pd.nops of them), then out's norm from the tile partials.
int dk_launch_pass(const dk_state *in, dk_state *out, const PassDesc &pd,
		   const PassOp *ops_dev)
{
	if (in->nq < (unsigned)PASS_TILE_BITS || out->nq != in->nq ||
	    out->dtype != in->dtype) {
		dk_set_error("a pass needs matching states of at least %d qubits",
			     PASS_TILE_BITS);
		return DK_E_ARG;
	}
	u64 seen = 0, sorted = 0;
	for (int h = 0; h < HIGH_BITS; ++h) {
		const unsigned q = pd.hpos[h];
		if (q < (unsigned)LANE_BITS || q >= in->nq || ((seen >> q) & 1)) {
			dk_set_error("bad pass tile (qubit %u)", q);
			return DK_E_ARG;
		}
		seen |= 1ull << q;
		if (pd.hsorted[h] >= in->nq ||
		    (h > 0 && pd.hsorted[h] <= pd.hsorted[h - 1])) {
			dk_set_error("bad pass tile order");
			return DK_E_ARG;
		}
		sorted |= 1ull << pd.hsorted[h];
	}
	if (sorted != seen) {
		dk_set_error("bad pass tile order");
		return DK_E_ARG;
	}
	const u64 ntiles = in->len >> PASS_TILE_BITS;
	// tiles outside the input's support hold zeros and keep them (only
	// when every matrix entry is finite: 0 * inf would not be zero)
	PassDesc d = pd;
	const u64 valid = in->len - 1;
	d.zmask = pd.finite ? (~in->support & valid) : 0;
	// in place, the zeros are already there
	d.zwrite = in != out;
	// a state's memory holds its zeros: every amplitude of a tile is read
	d.kmask = 0;
	const u64 tmask = ((1ull << LANE_BITS) - 1) | seen;
	double *parts = nullptr;
	DK_TRY(dk_partials_reserve(ntiles, &parts));
	DK_TRY(dk_time_mark(0));
	int rc = in->dtype == DK_C64
			 ? launch_pass_typed<float2>(in, out, d, ops_dev, parts, ntiles)
			 : launch_pass_typed<double2>(in, out, d, ops_dev, parts, ntiles);
	if (rc)
		return rc;
	// (a non-finite entry times a zero amplitude is NaN: then nothing is
	// known any more)
	out->support = pd.finite ? (in->support | tmask) : ~0ull;
	DK_TRY(dk_time_mark(1));
	return dk_finish_partials(parts, ntiles, out->scal + DK_SCAL_NORM, 1);
}

// One round of batched passes (see k_pass_round): `ents` and `ops` are device
// arrays; `grid` workgroups walk the round's items.
int dk_launch_pass_round(int dtype, void *pool, const PassEntry *ents,
			 unsigned nents, const PassOp *ops, u64 nitems, int init,
			 unsigned grid, hipStream_t st)
{
	if (!nitems)
		return DK_OK;
	if (dtype == DK_C64)
		hipLaunchKernelGGL(k_pass_round<float2>, dim3(grid), dim3(THREADS), 0, st,
				   (float2 *)pool, ents, nents, ops, nitems, init);
	else
		hipLaunchKernelGGL(k_pass_round<double2>, dim3(grid), dim3(THREADS), 0, st,
				   (double2 *)pool, ents, nents, ops, nitems, init);
	DK_LAUNCHED();
	return DK_OK;
}

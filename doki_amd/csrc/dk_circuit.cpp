// Gate lists and gate fusion (host side of libdoki_hip.so).
//
// A dk_circuit records gates with registry_apply's argument meaning.  Fusion
// groups them into blocks of at most K qubits and multiplies each block into
// one dense 2^m x 2^m unitary (FP64, on the host), so a circuit costs one
// state pass per block instead of one per gate.
//
// Placement rule: a gate on qubit set Q may join block b when no block after
// b touches Q (so it commutes past every block it moves over) and b's qubit
// set stays within K.  Among admissible blocks the one gaining the fewest new
// qubits wins (ties: the earliest); otherwise a new block opens at the end.
//
// Cache-blocked passes (dk_circuit_passes) are the other plan: each pass
// streams the state once and applies every one-qubit gate whose target lies
// in the pass's 12-qubit tile (qubits 0..5 and six chosen ones), plus every
// diagonal gate wherever its target is; controls may sit anywhere.  Gates
// move ahead of gates left for later passes only when they commute with
// them (checked on the qubits each gate uses non-diagonally).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "dk_internal.h"
#include "dk_pass.h"

int dk_launch_pass(const dk_state *in, dk_state *out, const dkp::PassDesc &pd,
		   const dkp::PassOp *ops_dev);

typedef std::complex<double> cd;

namespace {

struct CGate {
	unsigned k;
	std::vector<cd> m;           // 2^k x 2^k row-major
	unsigned t[DK_MAX_GATE_QUBITS];
	u64 cm, am, qmask;
	dk_gate *dev = nullptr;      // device copy for unfused runs
};

struct CBlock {
	u64 qmask = 0;
	std::vector<int> gates;
	std::vector<unsigned> qubits; // ascending; local bit i = qubits[i]
	std::vector<cd> U;            // 2^m x 2^m row-major
	dk_gate *dev = nullptr;
};

int popc(u64 x)
{
	return __builtin_popcountll(x);
}

// Applies gate g to the local vector v (2^m amplitudes over `qubits`), with
// registry_apply's semantics (qops.c:176-261): controlled-off entries copy.
void apply_local(const CGate &g, const std::vector<unsigned> &qubits,
		 std::vector<cd> &v)
{
	const size_t dim = v.size();
	int pos[64];
	for (size_t i = 0; i < qubits.size(); ++i)
		pos[qubits[i]] = (int)i;
	u64 lcm = 0, lam = 0;
	for (unsigned q = 0; q < 64; ++q) {
		if ((g.cm >> q) & 1)
			lcm |= 1ull << pos[q];
		if ((g.am >> q) & 1)
			lam |= 1ull << pos[q];
	}
	unsigned lt[DK_MAX_GATE_QUBITS];
	for (unsigned b = 0; b < g.k; ++b)
		lt[b] = (unsigned)pos[g.t[b]];
	const unsigned gd = 1u << g.k;
	std::vector<cd> out(dim);
	for (size_t i = 0; i < dim; ++i) {
		if ((i & lcm) != lcm || (i & lam) != 0) {
			out[i] = v[i];
			continue;
		}
		unsigned row = 0;
		for (unsigned b = 0; b < g.k; ++b)
			row |= (unsigned)((i >> lt[b]) & 1) << b;
		cd acc = 0;
		for (unsigned c = 0; c < gd; ++c) {
			size_t j = i;
			for (unsigned b = 0; b < g.k; ++b) {
				if ((c >> b) & 1)
					j |= (size_t)1 << lt[b];
				else
					j &= ~((size_t)1 << lt[b]);
			}
			acc += g.m[(size_t)row * gd + c] * v[j];
		}
		out[i] = acc;
	}
	v.swap(out);
}

} // namespace

// A step of a pass plan: a pass (index into passes) or one gate run alone.
struct PlanItem {
	bool pass;
	int idx;
};

struct PassPlan {
	dkp::PassDesc desc;
	size_t first_op;
	unsigned ngates;
};

struct dk_circuit {
	unsigned nq;
	std::vector<CGate> gates;
	std::vector<CBlock> blocks;
	unsigned fuse_k = 0;
	// cache-blocked pass plan
	bool passes_planned = false;
	std::vector<PlanItem> plan;
	std::vector<PassPlan> passes;
	std::vector<dkp::PassOp> pops;
	dkp::PassOp *pops_dev = nullptr;
	size_t pops_dev_bytes = 0;
	dk_lane *pops_lane = nullptr;
	// one-qubit gates of a planned circuit on a small state, for the
	// one-workgroup kernel (device copy made on the first run)
	std::vector<dk_small_gate> sgates;
	dk_small_gate *sgates_dev = nullptr;
	size_t sgates_dev_bytes = 0;
	dk_lane *sgates_lane = nullptr;
};

extern "C" int dk_circuit_new(unsigned num_qubits, dk_circuit **out)
{
	if (!out || num_qubits < 1 || num_qubits > DK_MAX_QUBITS) {
		dk_set_error("Number of qubits exceeds maximum");
		return DK_E_QUBITS;
	}
	dk_circuit *c = new dk_circuit;
	c->nq = num_qubits;
	*out = c;
	return DK_OK;
}

static void drop_device_gates(dk_circuit *c)
{
	if (c->pops_dev) {
		if (!dk_closed()) {
			(void)hipStreamSynchronize(c->pops_lane->stream);
			dk_device_free(c->pops_dev, c->pops_dev_bytes);
		}
		c->pops_dev = nullptr;
		c->pops_dev_bytes = 0;
	}
	if (c->sgates_dev) {
		if (!dk_closed()) {
			(void)hipStreamSynchronize(c->sgates_lane->stream);
			dk_device_free(c->sgates_dev, c->sgates_dev_bytes);
		}
		c->sgates_dev = nullptr;
		c->sgates_dev_bytes = 0;
	}
	c->sgates.clear();
	for (auto &b : c->blocks)
		if (b.dev) {
			dk_gate_free(b.dev);
			b.dev = nullptr;
		}
	for (auto &g : c->gates)
		if (g.dev) {
			dk_gate_free(g.dev);
			g.dev = nullptr;
		}
}

extern "C" void dk_circuit_free(dk_circuit *c)
{
	if (!c)
		return;
	drop_device_gates(c);
	delete c;
}

// Checks one gate against the circuit and builds its record.
static int make_gate(const dk_circuit *c, unsigned k, const double *matrix,
		     const unsigned *targets, uint64_t ctrl_mask,
		     uint64_t actrl_mask, CGate &g)
{
	if (!matrix || !targets || k < 1 || k > DK_MAX_GATE_QUBITS) {
		dk_set_error("gates act on 1..%d qubits", DK_MAX_GATE_QUBITS);
		return DK_E_ARG;
	}
	const u64 valid = c->nq >= 64 ? ~0ull : ((1ull << c->nq) - 1);
	g.k = k;
	u64 tm = 0;
	for (unsigned b = 0; b < k; ++b) {
		if (targets[b] >= c->nq || ((tm >> targets[b]) & 1)) {
			dk_set_error("Target qubit out of range");
			return DK_E_ARG;
		}
		g.t[b] = targets[b];
		tm |= 1ull << targets[b];
	}
	if ((ctrl_mask | actrl_mask) & ~valid) {
		dk_set_error("Control qubit out of range");
		return DK_E_ARG;
	}
	if ((ctrl_mask & actrl_mask) || (tm & (ctrl_mask | actrl_mask))) {
		dk_set_error("A target cannot also be a control or an anticontrol");
		return DK_E_ARG;
	}
	g.cm = ctrl_mask;
	g.am = actrl_mask;
	g.qmask = tm | ctrl_mask | actrl_mask;
	const size_t n = (size_t)1 << (2 * k);
	g.m.resize(n);
	for (size_t i = 0; i < n; ++i)
		g.m[i] = cd(matrix[2 * i], matrix[2 * i + 1]);
	return DK_OK;
}

// New gates invalidate any previous plan and its device copies.
static void invalidate_plan(dk_circuit *c)
{
	drop_device_gates(c);
	c->blocks.clear();
	c->fuse_k = 0;
	c->passes_planned = false;
}

extern "C" int dk_circuit_add(dk_circuit *c, unsigned k, const double *matrix,
			      const unsigned *targets, uint64_t ctrl_mask,
			      uint64_t actrl_mask)
{
	if (!c) {
		dk_set_error("NULL circuit");
		return DK_E_ARG;
	}
	CGate g;
	const int rc = make_gate(c, k, matrix, targets, ctrl_mask, actrl_mask, g);
	if (rc)
		return rc;
	c->gates.push_back(std::move(g));
	invalidate_plan(c);
	return DK_OK;
}

// All or nothing: every gate is checked before any is appended, and the
// plan is invalidated once per call.
extern "C" int dk_circuit_add_many(dk_circuit *c, size_t count,
				   const double *matrices, const unsigned *targets,
				   const uint64_t *ctrl_masks,
				   const uint64_t *actrl_masks)
{
	if (!c || (count && (!matrices || !targets))) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	std::vector<CGate> add(count);
	for (size_t i = 0; i < count; ++i) {
		const int rc = make_gate(c, 1, matrices + 8 * i, targets + i,
					 ctrl_masks ? ctrl_masks[i] : 0,
					 actrl_masks ? actrl_masks[i] : 0, add[i]);
		if (rc)
			return rc;
	}
	if (!count)
		return DK_OK;
	c->gates.reserve(c->gates.size() + count);
	for (auto &g : add)
		c->gates.push_back(std::move(g));
	invalidate_plan(c);
	return DK_OK;
}

extern "C" int dk_circuit_fuse(dk_circuit *c, unsigned max_qubits,
			       size_t *num_blocks)
{
	if (!c) {
		dk_set_error("NULL circuit");
		return DK_E_ARG;
	}
	drop_device_gates(c);
	c->blocks.clear();
	c->passes_planned = false;
	c->fuse_k = max_qubits;
	if (max_qubits == 0) {
		if (num_blocks)
			*num_blocks = 0;
		return DK_OK;
	}
	if (max_qubits > DK_MAX_GATE_QUBITS)
		max_qubits = c->fuse_k = DK_MAX_GATE_QUBITS;
	std::vector<int> last(64, -1);
	for (size_t gi = 0; gi < c->gates.size(); ++gi) {
		const CGate &g = c->gates[gi];
		if ((unsigned)popc(g.qmask) > max_qubits) {
			dk_set_error("gate on %d qubits (with controls) exceeds the "
				     "fusion width %u", popc(g.qmask), max_qubits);
			c->blocks.clear();
			c->fuse_k = 0;
			return DK_E_ARG;
		}
		int b0 = -1;
		for (unsigned q = 0; q < c->nq; ++q)
			if ((g.qmask >> q) & 1)
				b0 = std::max(b0, last[q]);
		int best = -1, best_new = 1 << 30;
		const int nb = (int)c->blocks.size();
		for (int b = std::max(b0, std::max(0, nb - 64)); b < nb; ++b) {
			const u64 u = c->blocks[b].qmask | g.qmask;
			if ((unsigned)popc(u) > max_qubits)
				continue;
			const int added = popc(g.qmask & ~c->blocks[b].qmask);
			if (added < best_new) {
				best_new = added;
				best = b;
			}
		}
		if (best < 0) {
			c->blocks.emplace_back();
			best = (int)c->blocks.size() - 1;
		}
		CBlock &blk = c->blocks[best];
		blk.gates.push_back((int)gi);
		blk.qmask |= g.qmask;
		for (unsigned q = 0; q < c->nq; ++q)
			if ((g.qmask >> q) & 1)
				last[q] = best;
	}
	// block unitaries: U = G_last ... G_first on the block's local qubits
	for (auto &blk : c->blocks) {
		blk.qubits.clear();
		for (unsigned q = 0; q < c->nq; ++q)
			if ((blk.qmask >> q) & 1)
				blk.qubits.push_back(q);
		const size_t dim = (size_t)1 << blk.qubits.size();
		blk.U.assign(dim * dim, cd(0, 0));
		std::vector<cd> col(dim);
		for (size_t j = 0; j < dim; ++j) {
			std::fill(col.begin(), col.end(), cd(0, 0));
			col[j] = 1;
			for (int gi : blk.gates)
				apply_local(c->gates[gi], blk.qubits, col);
			for (size_t i = 0; i < dim; ++i)
				blk.U[i * dim + j] = col[i];
		}
	}
	if (num_blocks)
		*num_blocks = c->blocks.size();
	return DK_OK;
}

extern "C" int dk_circuit_num_gates(const dk_circuit *c, size_t *n)
{
	if (!c || !n) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	*n = c->gates.size();
	return DK_OK;
}

extern "C" int dk_circuit_block(const dk_circuit *c, size_t i,
				unsigned *num_qubits, unsigned *qubits,
				double *matrix)
{
	if (!c || i >= c->blocks.size()) {
		dk_set_error("block index out of range");
		return DK_E_ARG;
	}
	const CBlock &b = c->blocks[i];
	if (num_qubits)
		*num_qubits = (unsigned)b.qubits.size();
	if (qubits)
		for (size_t k = 0; k < b.qubits.size(); ++k)
			qubits[k] = b.qubits[k];
	if (matrix)
		for (size_t k = 0; k < b.U.size(); ++k) {
			matrix[2 * k] = b.U[k].real();
			matrix[2 * k + 1] = b.U[k].imag();
		}
	return DK_OK;
}

// ---- cache-blocked pass plans ---------------------------------------------------

namespace {

using dkp::PassOp;

// A one-qubit gate is diagonal when its off-diagonal entries are exactly 0.
bool is_diag(const CGate &g)
{
	return g.k == 1 && g.m[1] == cd(0, 0) && g.m[2] == cd(0, 0);
}

// Qubits a gate uses non-diagonally (its targets, unless it is a diagonal
// one-qubit gate) and diagonally (controls, and a diagonal gate's target).
// Two gates commute when neither's non-diagonal qubits meet the other's
// qubits: on every shared qubit both act diagonally.
void gate_sets(const CGate &g, u64 &N, u64 &D)
{
	u64 tm = 0;
	for (unsigned b = 0; b < g.k; ++b)
		tm |= 1ull << g.t[b];
	if (is_diag(g)) {
		N = 0;
		D = tm | g.cm | g.am;
	} else {
		N = tm;
		D = g.cm | g.am;
	}
}

// 2x2 product a * b (row-major).
void mul22(const cd *a, const cd *b, cd *r)
{
	cd t[4];
	t[0] = a[0] * b[0] + a[1] * b[2];
	t[1] = a[0] * b[1] + a[1] * b[3];
	t[2] = a[2] * b[0] + a[3] * b[2];
	t[3] = a[2] * b[1] + a[3] * b[3];
	for (int i = 0; i < 4; ++i)
		r[i] = t[i];
}

// Builds the op list of one pass: tile qubits `high` (six of them, qubits
// >= 6), gates `gs` in an order that respects their commutation.
struct PassBuilder {
	const dk_circuit *c;
	std::vector<PassOp> &ops;
	int slot_of[64];
	unsigned rsl[dkp::REG_BITS], wsl[dkp::WAVE_BITS];
	// the last gate op since the last exchange, for merging
	long last = -1;
	unsigned last_q = 0;
	u64 last_cm = 0, last_am = 0;
	// Per op of this pass (ops[first + i] <-> meta[i]): the qubits it acts
	// on non-diagonally, and for a phase on |11> of two qubits (a diagonal
	// diag(1, phase) with one control) those qubits, which a later dense
	// op on either of them may absorb.
	struct OpMeta {
		u64 N = 0;
		bool phase2 = false;
		bool diag1 = false;  // an uncontrolled diagonal diag(d0, d1) on qa
		unsigned qa = 0, qb = 0;
		cd phase = 1.0, d0 = 1.0, d1 = 1.0;
	};
	size_t first;
	std::vector<OpMeta> meta;

	PassBuilder(const dk_circuit *c_, std::vector<PassOp> &ops_)
		: c(c_), ops(ops_), first(ops_.size())
	{
		for (int q = 0; q < 64; ++q)
			slot_of[q] = -1;
	}

	// Location of qubit q under the current mapping: 0 lane (*pos = bit),
	// 1 register (*pos = j), 2 wave (*pos = b), 3 fixed (*pos = q).
	int where(unsigned q, unsigned *pos) const
	{
		if (q < (unsigned)dkp::LANE_BITS) {
			*pos = q;
			return 0;
		}
		const int h = slot_of[q];
		if (h < 0) {
			*pos = q;
			return 3;
		}
		for (unsigned j = 0; j < (unsigned)dkp::REG_BITS; ++j)
			if (rsl[j] == (unsigned)h) {
				*pos = j;
				return 1;
			}
		for (unsigned b = 0; b < (unsigned)dkp::WAVE_BITS; ++b)
			if (wsl[b] == (unsigned)h) {
				*pos = b;
				return 2;
			}
		*pos = q;
		return 3; // not reached: every high slot is a register or wave slot
	}

	bool executable(const CGate &g) const
	{
		if (is_diag(g))
			return true;
		unsigned pos;
		const int w = where(g.t[0], &pos);
		return w == 0 || w == 1;
	}

	// Splits a control mask into register bits (returned, 4 bits) and the
	// qubits of the thread word (*thread).
	unsigned split_mask(u64 m, uint64_t *thread) const
	{
		unsigned reg = 0;
		*thread = 0;
		for (; m; m &= m - 1) {
			const unsigned q = (unsigned)__builtin_ctzll(m);
			unsigned pos;
			if (where(q, &pos) == 1)
				reg |= 1u << pos;
			else
				*thread |= 1ull << q;
		}
		return reg;
	}

	// Dense uncontrolled gate m on q absorbing an earlier diagonal: an
	// uncontrolled one on q (m becomes m * diag(d0, d1); returns FOLD_MORE,
	// the caller tries again and emits m), or a phase on |11> of (q, s): m
	// where s is 0, m * diag(1, phase) where s is 1, as one mux op (emitted
	// here; returns FOLD_DONE).  FOLD_NONE: nothing to absorb.  The phase moves past the ops in between, which act on neither
	// q nor s non-diagonally.  The selector s must be a register bit (for a
	// register target) or a thread-word bit >= 6 (one value per wave).
	enum { FOLD_NONE, FOLD_MORE, FOLD_DONE };
	int try_fold(unsigned q, cd (&m)[4])
	{
		unsigned tpos, spos;
		const int tw = where(q, &tpos);
		u64 B = 0; // qubits acted on non-diagonally after the candidate
		for (size_t i = ops.size(); i-- > first && ops.size() - i <= 64;) {
			const OpMeta &mt = meta[i - first];
			if (mt.diag1 && mt.qa == q && !((B >> q) & 1)) {
				// an uncontrolled diagonal on q: m * diag(d0, d1),
				// emitted as a plain op by the caller
				m[0] *= mt.d0;
				m[2] *= mt.d0;
				m[1] *= mt.d1;
				m[3] *= mt.d1;
				ops.erase(ops.begin() + (long)i);
				meta.erase(meta.begin() + (long)(i - first));
				last = -1;
				return FOLD_MORE;
			}
			if (mt.phase2 && (mt.qa == q || mt.qb == q)) {
				const unsigned sq = mt.qa == q ? mt.qb : mt.qa;
				const int sw = where(sq, &spos);
				const bool sel_ok = (sw == 1 && tw == 1) || sw == 2 || sw == 3;
				if (sel_ok && !(B & ((1ull << q) | (1ull << sq)))) {
					const cd ph = mt.phase;
					ops.erase(ops.begin() + (long)i);
					meta.erase(meta.begin() + (long)(i - first));
					PassOp o, d;
					memset(&o, 0, sizeof(o));
					memset(&d, 0, sizeof(d));
					const cd m1[4] = {m[0], m[1] * ph, m[2], m[3] * ph};
					for (int k = 0; k < 4; ++k) {
						o.m[2 * k] = m[k].real();
						o.m[2 * k + 1] = m[k].imag();
						d.m[2 * k] = m1[k].real();
						d.m[2 * k + 1] = m1[k].imag();
					}
					o.kind = tw == 1 ? dkp::OP_REG_MUX : dkp::OP_LANE_MUX;
					o.a = tpos;
					o.b = sw == 1 ? (dkp::SEL_REG | spos) : sq;
					d.kind = dkp::OP_MDATA;
					ops.push_back(o);
					ops.push_back(d);
					OpMeta mo;
					mo.N = 1ull << q;
					meta.push_back(mo);
					meta.push_back(OpMeta());
					last = -1;
					return FOLD_DONE;
				}
			}
			B |= mt.N;
			if ((B >> q) & 1)
				break;
		}
		return FOLD_NONE;
	}

	void emit_gate(const CGate &g)
	{
		const unsigned q = g.t[0];
		cd m[4] = {g.m[0], g.m[1], g.m[2], g.m[3]};
		if (last >= 0 && last_q == q && last_cm == g.cm && last_am == g.am) {
			// same target and controls as the previous op: one product
			const PassOp &o = ops[(size_t)last];
			cd prev[4];
			for (int i = 0; i < 4; ++i)
				prev[i] = cd(o.m[2 * i], o.m[2 * i + 1]);
			mul22(m, prev, m);
			ops.pop_back(); // the previous op is rebuilt below
			meta.pop_back();
		} else if (g.cm == 0 && g.am == 0 && !(m[1] == cd(0, 0) && m[2] == cd(0, 0))) {
			int f;
			while ((f = try_fold(q, m)) == FOLD_MORE) {
			}
			if (f == FOLD_DONE)
				return;
		}
		PassOp o;
		memset(&o, 0, sizeof(o));
		for (int i = 0; i < 4; ++i) {
			o.m[2 * i] = m[i].real();
			o.m[2 * i + 1] = m[i].imag();
		}
		const bool diag = m[1] == cd(0, 0) && m[2] == cd(0, 0);
		unsigned pos;
		const int w = where(q, &pos);
		if (diag) {
			o.kind = w == 1 ? dkp::OP_DIAG_REG : dkp::OP_DIAG_THREAD;
			o.a = w == 1 ? pos : q;
		} else {
			o.kind = w == 0 ? dkp::OP_LANE : dkp::OP_REG;
			o.a = pos;
		}
		o.rmask = split_mask(g.cm, &o.tcm) | split_mask(g.am, &o.tam) << 8;
		ops.push_back(o);
		OpMeta mt;
		if (!diag) {
			mt.N = 1ull << q;
		} else if (g.cm == 0 && g.am == 0) {
			mt.diag1 = true;
			mt.qa = q;
			mt.d0 = m[0];
			mt.d1 = m[3];
		} else if (m[0] == cd(1, 0) && g.am == 0 && g.cm && !(g.cm & (g.cm - 1))) {
			mt.phase2 = true;
			mt.qa = q;
			mt.qb = (unsigned)__builtin_ctzll(g.cm);
			mt.phase = m[3];
		}
		meta.push_back(mt);
		last = (long)ops.size() - 1;
		last_q = q;
		last_cm = g.cm;
		last_am = g.am;
	}

	// New mapping with register slots `want` (any order).  A slot that
	// stays in registers keeps its register bit: with 4 of 6 slots in
	// registers before and after, at least two do, and the kernel splits
	// the exchange into two halves on one of them.
	void emit_exchange(const unsigned (&want)[dkp::REG_BITS])
	{
		bool used[dkp::HIGH_BITS] = {false}, placed[dkp::HIGH_BITS] = {false};
		for (int j = 0; j < dkp::REG_BITS; ++j)
			used[want[j]] = true;
		unsigned nr[dkp::REG_BITS];
		bool filled[dkp::REG_BITS] = {false};
		for (int j = 0; j < dkp::REG_BITS; ++j)
			if (used[rsl[j]]) {
				nr[j] = rsl[j];
				filled[j] = placed[rsl[j]] = true;
			}
		for (int j = 0, k = 0; j < dkp::REG_BITS; ++j) {
			if (filled[j])
				continue;
			while (placed[want[k]])
				++k;
			nr[j] = want[k];
			placed[want[k]] = true;
		}
		unsigned nw[dkp::WAVE_BITS], k = 0;
		for (unsigned h = 0; h < (unsigned)dkp::HIGH_BITS; ++h)
			if (!used[h])
				nw[k++] = h;
		PassOp o;
		memset(&o, 0, sizeof(o));
		o.kind = dkp::OP_EXCH;
		for (int j = 0; j < dkp::REG_BITS; ++j) {
			rsl[j] = nr[j];
			o.a |= nr[j] << (3 * j);
		}
		for (int b = 0; b < dkp::WAVE_BITS; ++b) {
			wsl[b] = nw[b];
			o.b |= nw[b] << (3 * b);
		}
		ops.push_back(o);
		meta.push_back(OpMeta());
		last = -1;
	}
};

// Whether ops [first, first+n) are well formed for the pass kernel: every
// exchange names a permutation of the high slots, every bit index is in
// range.  The kernel indexes LDS and registers by these fields.
bool pass_ops_valid(const PassOp *ops, size_t n, unsigned nq)
{
	unsigned rs[dkp::REG_BITS]; // register slots of the current mapping
	for (int j = 0; j < dkp::REG_BITS; ++j)
		rs[j] = (unsigned)j;
	for (size_t i = 0; i < n; ++i) {
		const PassOp &o = ops[i];
		switch (o.kind) {
		case dkp::OP_EXCH: {
			unsigned seen = 0;
			bool kept = false; // a register bit keeps its slot
			for (int j = 0; j < dkp::REG_BITS; ++j) {
				const unsigned s = (o.a >> (3 * j)) & 7;
				seen |= 1u << s;
				kept |= s == rs[j];
				rs[j] = s;
			}
			for (int b = 0; b < dkp::WAVE_BITS; ++b)
				seen |= 1u << ((o.b >> (3 * b)) & 7);
			if (seen != (1u << dkp::HIGH_BITS) - 1 || !kept ||
			    (o.a >> (3 * dkp::REG_BITS)) || (o.b >> (3 * dkp::WAVE_BITS)))
				return false;
			continue;
		}
		case dkp::OP_REG_MUX:
		case dkp::OP_LANE_MUX: {
			// a second-matrix record follows; no controls; the selector
			// is another register bit or a thread-word bit >= 6
			const bool reg = o.kind == dkp::OP_REG_MUX;
			if (i + 1 >= n || ops[i + 1].kind != dkp::OP_MDATA ||
			    o.a >= (unsigned)(reg ? dkp::REG_BITS : dkp::LANE_BITS) ||
			    o.rmask || o.tcm || o.tam)
				return false;
			if (o.b & dkp::SEL_REG) {
				const unsigned j = o.b & ~dkp::SEL_REG;
				if (!reg || j >= (unsigned)dkp::REG_BITS || j == o.a)
					return false;
			} else if (o.b < (unsigned)dkp::LANE_BITS || o.b >= nq) {
				return false;
			}
			++i; // the data record
			continue;
		}
		case dkp::OP_REG:
		case dkp::OP_DIAG_REG:
			if (o.a >= (unsigned)dkp::REG_BITS)
				return false;
			break;
		case dkp::OP_LANE:
			if (o.a >= (unsigned)dkp::LANE_BITS)
				return false;
			break;
		case dkp::OP_DIAG_THREAD:
			if (o.a >= nq)
				return false;
			break;
		default:
			return false;
		}
		const u64 valid = nq >= 64 ? ~0ull : (1ull << nq) - 1;
		if ((o.rmask & ~0x0f0fu) || (o.tcm & ~valid) || (o.tam & ~valid) || o.b)
			return false;
	}
	return true;
}

// One pass over tile qubits `chosen` (high qubits, first-used first) and the
// gates `gs` (in circuit order, all absorbable).
void build_pass(dk_circuit *c, std::vector<unsigned> chosen,
		const std::vector<int> &gs)
{
	// fill the tile to six high qubits
	u64 cmask = 0;
	for (unsigned q : chosen)
		cmask |= 1ull << q;
	for (unsigned q = dkp::LANE_BITS; chosen.size() < (size_t)dkp::HIGH_BITS; ++q)
		if (!((cmask >> q) & 1)) {
			chosen.push_back(q);
			cmask |= 1ull << q;
		}
	PassPlan pp;
	memset(&pp.desc, 0, sizeof(pp.desc));
	pp.first_op = c->pops.size();
	pp.ngates = (unsigned)gs.size();
	PassBuilder B(c, c->pops);
	// high slot h <- chosen[h]: the first four used start in registers
	for (int h = 0; h < dkp::HIGH_BITS; ++h) {
		pp.desc.hpos[h] = (uint8_t)chosen[h];
		B.slot_of[chosen[h]] = h;
	}
	std::vector<unsigned> srt(chosen.begin(), chosen.end());
	std::sort(srt.begin(), srt.end());
	for (int h = 0; h < dkp::HIGH_BITS; ++h)
		pp.desc.hsorted[h] = (uint8_t)srt[h];
	for (int j = 0; j < dkp::REG_BITS; ++j)
		B.rsl[j] = (unsigned)j;
	for (int b = 0; b < dkp::WAVE_BITS; ++b)
		B.wsl[b] = (unsigned)(dkp::REG_BITS + b);

	std::vector<int> pending(gs.begin(), gs.end());
	while (!pending.empty()) {
		// the first gate that can run now and commutes with every pending
		// gate before it; a gate that merges into the previous op (same
		// target and controls) goes first
		u64 bN = 0, bD = 0;
		size_t pick = pending.size(), merge = pending.size();
		for (size_t i = 0; i < pending.size(); ++i) {
			const CGate &g = c->gates[pending[i]];
			u64 N, D;
			gate_sets(g, N, D);
			if (!(N & (bN | bD)) && !(D & bN) && B.executable(g)) {
				if (pick == pending.size())
					pick = i;
				if (B.last >= 0 && g.t[0] == B.last_q && g.cm == B.last_cm &&
				    g.am == B.last_am) {
					merge = i;
					break;
				}
				if (B.last < 0)
					break;
			}
			bN |= N;
			bD |= D;
		}
		if (merge < pending.size())
			pick = merge;
		if (pick < pending.size()) {
			B.emit_gate(c->gates[pending[pick]]);
			pending.erase(pending.begin() + (long)pick);
			continue;
		}
		// nothing runs under this mapping: bring the high targets of the
		// next pending gates into registers (the first pending gate's
		// target among them)
		unsigned nr[dkp::REG_BITS];
		int cnt = 0;
		bool in_nr[dkp::HIGH_BITS] = {false};
		for (size_t i = 0; i < pending.size() && cnt < dkp::REG_BITS; ++i) {
			const CGate &g = c->gates[pending[i]];
			if (is_diag(g) || g.t[0] < (unsigned)dkp::LANE_BITS)
				continue;
			const int h = B.slot_of[g.t[0]];
			if (!in_nr[h]) {
				in_nr[h] = true;
				nr[cnt++] = (unsigned)h;
			}
		}
		// keep current register slots in the remaining places
		for (int j = 0; j < dkp::REG_BITS && cnt < dkp::REG_BITS; ++j)
			if (!in_nr[B.rsl[j]]) {
				in_nr[B.rsl[j]] = true;
				nr[cnt++] = B.rsl[j];
			}
		B.emit_exchange(nr);
	}
	// a phase (diagonal with m00 = 1) becomes the factor m11 controlled on
	// its own target: the kernel then skips the amplitudes it leaves alone
	// (whole waves or registers where the target is held per wave or per
	// register)
	for (size_t i = pp.first_op; i < c->pops.size(); ++i) {
		PassOp &o = c->pops[i];
		if ((o.kind != dkp::OP_DIAG_REG && o.kind != dkp::OP_DIAG_THREAD) ||
		    o.m[0] != 1.0 || o.m[1] != 0.0)
			continue;
		o.m[0] = o.m[6];
		o.m[1] = o.m[7];
		if (o.kind == dkp::OP_DIAG_REG)
			o.rmask |= 1u << o.a;
		else
			o.tcm |= 1ull << o.a;
	}
	pp.desc.nops = (uint32_t)(c->pops.size() - pp.first_op);
	pp.desc.finite = 1;
	for (size_t i = pp.first_op; i < c->pops.size(); ++i)
		for (int j = 0; j < 8; ++j)
			if (!std::isfinite(c->pops[i].m[j]))
				pp.desc.finite = 0;
	pp.desc.zmask = 0;
	pp.desc.kmask = 0;
	c->passes.push_back(pp);
	c->plan.push_back({true, (int)c->passes.size() - 1});
}

} // namespace

namespace {

// A planned step before op building: a pass (tile qubits, absorbed gates)
// or a gate run alone.
struct PlanSpec {
	bool pass;
	int gate;                      // lone gate
	std::vector<unsigned> chosen;  // pass: high tile qubits, first-used first
	std::vector<int> gates;        // pass: absorbed gates, circuit order
};

// Splits `rem` into the gates a pass over high tile qubits `hmask` absorbs
// and the rest (deferred, in order).  A gate is absorbed when it is a
// one-qubit gate on a tile qubit or diagonal, and commutes with every
// deferred gate before it.
void absorb(const dk_circuit *c, const std::vector<int> &rem, u64 hmask,
	    std::vector<int> *absorbed, std::vector<int> *deferred)
{
	u64 bN = 0, bD = 0;
	for (int gi : rem) {
		const CGate &g = c->gates[gi];
		u64 N, D;
		gate_sets(g, N, D);
		const bool take = g.k == 1 && !(N & (bN | bD)) && !(D & bN) &&
				  (is_diag(g) || g.t[0] < (unsigned)dkp::LANE_BITS ||
				   ((hmask >> g.t[0]) & 1));
		if (take) {
			if (absorbed)
				absorbed->push_back(gi);
		} else {
			if (deferred)
				deferred->push_back(gi);
			bN |= N;
			bD |= D;
		}
	}
}

// High tile qubits of the next pass.  First-seen: the targets of the
// absorbable gates in circuit order, until six are chosen.  Look-ahead:
// six times, the qubit whose addition absorbs the most gates.
u64 choose_tile(const dk_circuit *c, const std::vector<int> &rem, bool lookahead)
{
	u64 hmask = 0;
	int nh = 0;
	if (!lookahead) {
		u64 bN = 0, bD = 0;
		for (int gi : rem) {
			const CGate &g = c->gates[gi];
			u64 N, D;
			gate_sets(g, N, D);
			bool take = false;
			if (g.k == 1 && !(N & (bN | bD)) && !(D & bN)) {
				const unsigned t = g.t[0];
				if (is_diag(g) || t < (unsigned)dkp::LANE_BITS || ((hmask >> t) & 1)) {
					take = true;
				} else if (nh < dkp::HIGH_BITS) {
					hmask |= 1ull << t;
					++nh;
					take = true;
				}
			}
			if (!take) {
				bN |= N;
				bD |= D;
			}
		}
		return hmask;
	}
	u64 cand = 0;
	for (int gi : rem) {
		const CGate &g = c->gates[gi];
		if (g.k == 1 && !is_diag(g) && g.t[0] >= (unsigned)dkp::LANE_BITS)
			cand |= 1ull << g.t[0];
	}
	std::vector<int> ab;
	for (; nh < dkp::HIGH_BITS && cand; ++nh) {
		size_t best = 0;
		unsigned bq = 64;
		for (u64 m = cand; m; m &= m - 1) {
			const unsigned q = (unsigned)__builtin_ctzll(m);
			ab.clear();
			absorb(c, rem, hmask | (1ull << q), &ab, nullptr);
			if (bq == 64 || ab.size() > best) {
				best = ab.size();
				bq = q;
			}
		}
		hmask |= 1ull << bq;
		cand &= ~(1ull << bq);
	}
	return hmask;
}

// Plans the whole circuit with one tile rule.  lone: a pass that would
// absorb a single dense gate leaves it to the one-qubit kernel instead.
std::vector<PlanSpec> plan_specs(const dk_circuit *c, bool lookahead, bool lone = true)
{
	std::vector<PlanSpec> out;
	std::vector<int> rem(c->gates.size());
	for (size_t i = 0; i < rem.size(); ++i)
		rem[i] = (int)i;
	if (c->nq < (unsigned)dkp::PASS_TILE_BITS) {
		// too small for a tile: every gate runs alone
		for (int gi : rem)
			out.push_back({false, gi, {}, {}});
		return out;
	}
	while (!rem.empty()) {
		if (c->gates[rem[0]].k != 1) {
			out.push_back({false, rem[0], {}, {}});
			rem.erase(rem.begin());
			continue;
		}
		const u64 hmask = choose_tile(c, rem, lookahead);
		std::vector<int> absorbed, deferred;
		absorb(c, rem, hmask, &absorbed, &deferred);
		rem.swap(deferred);
		if (absorbed.empty()) {
			// cannot happen with one-qubit gates first in `rem`; kept as
			// a guard against an endless loop: the first gate runs alone
			out.push_back({false, rem[0], {}, {}});
			rem.erase(rem.begin());
			continue;
		}
		if (lone && absorbed.size() == 1 && !is_diag(c->gates[absorbed[0]])) {
			// one dense gate: the one-qubit kernel does it alone
			out.push_back({false, absorbed[0], {}, {}});
			continue;
		}
		// high tile qubits in order of first use as a dense target
		PlanSpec ps{true, -1, {}, std::move(absorbed)};
		u64 seen = 0;
		for (int gi : ps.gates) {
			const CGate &g = c->gates[gi];
			const unsigned t = g.t[0];
			if (!is_diag(g) && ((hmask >> t) & 1) && !((seen >> t) & 1)) {
				seen |= 1ull << t;
				ps.chosen.push_back(t);
			}
		}
		for (u64 m = hmask & ~seen; m; m &= m - 1)
			ps.chosen.push_back((unsigned)__builtin_ctzll(m));
		out.push_back(std::move(ps));
	}
	return out;
}

} // namespace

extern "C" int dk_circuit_passes(dk_circuit *c, size_t *num_passes)
{
	if (!c) {
		dk_set_error("NULL circuit");
		return DK_E_ARG;
	}
	drop_device_gates(c);
	c->blocks.clear();
	c->fuse_k = 0;
	c->plan.clear();
	c->passes.clear();
	c->pops.clear();
	c->passes_planned = true;
	// both tile rules; the plan with fewer state passes (passes plus lone
	// gates) wins, first-seen on ties
	std::vector<PlanSpec> a = plan_specs(c, false);
	std::vector<PlanSpec> b = plan_specs(c, true);
	const std::vector<PlanSpec> &best = b.size() < a.size() ? b : a;
	for (const PlanSpec &ps : best) {
		if (ps.pass)
			build_pass(c, ps.chosen, ps.gates);
		else
			c->plan.push_back({false, ps.gate});
	}
	for (const PassPlan &pp : c->passes)
		if (!pass_ops_valid(&c->pops[pp.first_op], pp.desc.nops, c->nq)) {
			dk_set_error("internal error: malformed pass plan");
			c->passes_planned = false;
			return DK_E_ARG;
		}
	if (num_passes)
		*num_passes = c->passes.size();
	return DK_OK;
}

extern "C" int dk_circuit_pass(const dk_circuit *c, size_t i, unsigned *num_gates,
			       unsigned *num_ops, unsigned *tile_qubits)
{
	if (!c || !c->passes_planned || i >= c->passes.size()) {
		dk_set_error("pass index out of range");
		return DK_E_ARG;
	}
	const PassPlan &pp = c->passes[i];
	if (num_gates)
		*num_gates = pp.ngates;
	if (num_ops)
		*num_ops = pp.desc.nops;
	if (tile_qubits) {
		for (int b = 0; b < dkp::LANE_BITS; ++b)
			tile_qubits[b] = (unsigned)b;
		for (int h = 0; h < dkp::HIGH_BITS; ++h)
			tile_qubits[dkp::LANE_BITS + h] = pp.desc.hpos[h];
	}
	return DK_OK;
}

extern "C" int dk_circuit_pass_ops(const dk_circuit *c, size_t i, void *ops,
				   size_t capacity, size_t *count)
{
	static_assert(sizeof(PassOp) == 96, "PassOp records are 96 bytes");
	if (!c || !c->passes_planned || i >= c->passes.size()) {
		dk_set_error("pass index out of range");
		return DK_E_ARG;
	}
	const PassPlan &pp = c->passes[i];
	if (count)
		*count = pp.desc.nops;
	if (ops) {
		if (capacity < pp.desc.nops) {
			dk_set_error("room for %zu ops, the pass has %u", capacity,
				     (unsigned)pp.desc.nops);
			return DK_E_ARG;
		}
		memcpy(ops, &c->pops[pp.first_op], pp.desc.nops * sizeof(PassOp));
	}
	return DK_OK;
}

extern "C" int dk_circuit_plan_step(const dk_circuit *c, size_t i, int *is_pass,
				    size_t *index)
{
	if (!c || !c->passes_planned || i >= c->plan.size()) {
		dk_set_error("plan step out of range");
		return DK_E_ARG;
	}
	if (is_pass)
		*is_pass = c->plan[i].pass ? 1 : 0;
	if (index)
		*index = (size_t)c->plan[i].idx;
	return DK_OK;
}

static int gate_to_device(unsigned k, const std::vector<cd> &m, dk_gate **out)
{
	std::vector<double> flat(2 * m.size());
	for (size_t i = 0; i < m.size(); ++i) {
		flat[2 * i] = m[i].real();
		flat[2 * i + 1] = m[i].imag();
	}
	return dk_gate_new(k, flat.data(), out);
}

// Runs the circuit (its fused blocks when fused, else gate by gate).  The
// first step writes a new state unless in_place; the rest run in place.
extern "C" int dk_circuit_run(dk_circuit *c, dk_state *st, int in_place,
			      dk_state **out)
{
	if (!c || !st || !out) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	if (st->nq != c->nq) {
		dk_set_error("circuit of %u qubits on a state of %u", c->nq, st->nq);
		return DK_E_ARG;
	}
	int rc = dk_ensure_init();
	if (rc)
		return rc;
	DK_SETTLE(st);
	dk_state *cur = st;
	bool first = !in_place;
	auto step = [&](const dk_gate *g, const unsigned *t, u64 cm, u64 am) {
		dk_state *o = cur;
		if (first) {
			int r = dk_alloc_state(cur->nq, cur->dtype, &o);
			if (r)
				return r;
		}
		int r = dk_launch_apply(cur, o, g, t, cm, am);
		if (r) {
			if (o != cur)
				dk_release_state(o);
			return r;
		}
		o->lane = dk_cur();
		cur = o;
		first = false;
		return (int)DK_OK;
	};
	auto step_pass = [&](const dkp::PassDesc &pd, const PassOp *ops_dev) {
		dk_state *o = cur;
		if (first) {
			int r = dk_alloc_state(cur->nq, cur->dtype, &o);
			if (r)
				return r;
		}
		int r = dk_launch_pass(cur, o, pd, ops_dev);
		if (r) {
			if (o != cur)
				dk_release_state(o);
			return r;
		}
		o->lane = dk_cur();
		cur = o;
		first = false;
		return (int)DK_OK;
	};
	// A planned circuit on a state smaller than a pass tile (where the plan
	// runs every gate alone): every run of one-qubit gates in one launch of
	// one workgroup holding the state in LDS; multi-qubit gates alone
	// between.  (At 12 qubits the one-tile passes measured faster: 200
	// gates in 0.34 against 0.40 ms through the drop-in module.)
	auto small_run = [&]() {
		if (c->sgates.empty()) {
			for (const CGate &g : c->gates) {
				if (g.k != 1)
					continue;
				dk_small_gate sg;
				memset(&sg, 0, sizeof(sg));
				for (int j = 0; j < 4; ++j) {
					sg.m[2 * j] = g.m[(size_t)j].real();
					sg.m[2 * j + 1] = g.m[(size_t)j].imag();
				}
				sg.cm = g.cm;
				sg.am = g.am;
				sg.t = g.t[0];
				c->sgates.push_back(sg);
			}
		}
		if (!c->sgates.empty() && !c->sgates_dev) {
			const size_t bytes = c->sgates.size() * sizeof(dk_small_gate);
			int r = dk_device_alloc(bytes, (void **)&c->sgates_dev);
			if (r)
				return r;
			c->sgates_dev_bytes = bytes;
			c->sgates_lane = dk_cur();
			// the host records stay until the device copy is dropped,
			// which waits for this lane first
			hipError_t e = hipMemcpyAsync(c->sgates_dev, c->sgates.data(), bytes,
						      hipMemcpyHostToDevice, DK_STREAM);
			if (e != hipSuccess) {
				dk_set_error("gate upload failed: %s", hipGetErrorString(e));
				return (int)DK_E_DEVICE;
			}
		} else if (c->sgates_dev && c->sgates_lane != dk_cur()) {
			DK_HIP(hipStreamSynchronize(c->sgates_lane->stream));
			c->sgates_lane = dk_cur();
		}
		size_t si = 0;
		for (size_t gi = 0; gi < c->gates.size();) {
			CGate &g = c->gates[gi];
			if (g.k != 1) {
				if (!g.dev) {
					int r = gate_to_device(g.k, g.m, &g.dev);
					if (r)
						return r;
				}
				int r = step(g.dev, g.t, g.cm, g.am);
				if (r)
					return r;
				++gi;
				continue;
			}
			size_t n1 = 0;
			while (gi + n1 < c->gates.size() && c->gates[gi + n1].k == 1)
				++n1;
			dk_state *o = cur;
			if (first) {
				int r = dk_alloc_state(cur->nq, cur->dtype, &o);
				if (r)
					return r;
			}
			int r = dk_time_mark(0);
			if (!r)
				r = dk_launch_small_run(cur, o, c->sgates_dev + si, (unsigned)n1);
			if (!r)
				r = dk_time_mark(1);
			if (r) {
				if (o != cur)
					dk_release_state(o);
				return r;
			}
			o->lane = dk_cur();
			o->support = ~0ull;
			cur = o;
			first = false;
			si += n1;
			gi += n1;
		}
		return (int)DK_OK;
	};
	if (c->passes_planned && c->nq < (unsigned)dkp::PASS_TILE_BITS &&
	    c->nq <= dk_small_qubits(st->dtype)) {
		rc = small_run();
	} else if (c->passes_planned) {
		if (!c->pops.empty() && !c->pops_dev) {
			const size_t bytes = c->pops.size() * sizeof(PassOp);
			rc = dk_device_alloc(bytes, (void **)&c->pops_dev);
			if (rc)
				return rc;
			c->pops_dev_bytes = bytes;
			c->pops_lane = dk_cur();
			// the host ops stay untouched until the plan is dropped, which
			// waits for this lane first
			hipError_t e = hipMemcpyAsync(c->pops_dev, c->pops.data(), bytes,
						      hipMemcpyHostToDevice, DK_STREAM);
			if (e != hipSuccess) {
				dk_set_error("pass upload failed: %s", hipGetErrorString(e));
				return DK_E_DEVICE;
			}
		} else if (c->pops_dev && c->pops_lane != dk_cur()) {
			// uploaded (or last read) on another lane
			DK_HIP(hipStreamSynchronize(c->pops_lane->stream));
			c->pops_lane = dk_cur();
		}
		for (const PlanItem &it : c->plan) {
			if (it.pass) {
				const PassPlan &pp = c->passes[(size_t)it.idx];
				rc = step_pass(pp.desc, c->pops_dev + pp.first_op);
			} else {
				CGate &g = c->gates[(size_t)it.idx];
				if (!g.dev) {
					rc = gate_to_device(g.k, g.m, &g.dev);
					if (rc)
						break;
				}
				rc = step(g.dev, g.t, g.cm, g.am);
			}
			if (rc)
				break;
		}
	} else if (c->fuse_k > 0) {
		for (auto &b : c->blocks) {
			if (!b.dev) {
				rc = gate_to_device((unsigned)b.qubits.size(), b.U, &b.dev);
				if (rc)
					break;
			}
			rc = step(b.dev, b.qubits.data(), 0, 0);
			if (rc)
				break;
		}
	} else {
		for (auto &g : c->gates) {
			if (!g.dev) {
				rc = gate_to_device(g.k, g.m, &g.dev);
				if (rc)
					break;
			}
			rc = step(g.dev, g.t, g.cm, g.am);
			if (rc)
				break;
		}
	}
	if (rc) {
		if (cur != st)
			dk_release_state(cur);
		return rc;
	}
	if (first) {
		// empty circuit, out of place: a normalised copy
		return dk_state_clone(st, out);
	}
	*out = cur;
	return DK_OK;
}

// Pass plan of a register's one-qubit gates (count gates: 2x2 matrices at
// mats + 8 i, targets, control and anticontrol masks) on nq >= 12 qubits with
// every gate inside a pass: appends pass i's description to descs and its
// ops to ops, from op_begin[i].  The better of the two tile rules, as
// dk_circuit_passes picks it.  DK_E_ARG if some gate would have to run
// alone.  For the batched engine's grid-wide rounds.
int dk_plan_register_passes(unsigned nq, size_t count, const double *mats,
			    const unsigned *targets, const uint64_t *cms,
			    const uint64_t *ams, std::vector<dkp::PassDesc> &descs,
			    std::vector<dkp::PassOp> &ops,
			    std::vector<uint32_t> &op_begin)
{
	if (nq < (unsigned)dkp::PASS_TILE_BITS) {
		dk_set_error("a pass needs a register of at least %d qubits",
			     dkp::PASS_TILE_BITS);
		return DK_E_ARG;
	}
	dk_circuit c;
	c.nq = nq;
	c.gates.reserve(count);
	for (size_t i = 0; i < count; ++i) {
		CGate g;
		g.k = 1;
		g.t[0] = targets[i];
		g.cm = cms ? cms[i] : 0;
		g.am = ams ? ams[i] : 0;
		g.qmask = (1ull << g.t[0]) | g.cm | g.am;
		g.m.resize(4);
		for (int e = 0; e < 4; ++e)
			g.m[e] = cd(mats[8 * i + 2 * e], mats[8 * i + 2 * e + 1]);
		c.gates.push_back(std::move(g));
	}
	// first-seen tiles always take the first remaining gate's target, so
	// that plan has passes only; the look-ahead plan is used when it has
	// too and is shorter
	std::vector<PlanSpec> a = plan_specs(&c, false, false);
	std::vector<PlanSpec> b = plan_specs(&c, true, false);
	auto passes_only = [](const std::vector<PlanSpec> &v) {
		for (const PlanSpec &ps : v)
			if (!ps.pass)
				return false;
		return true;
	};
	const bool bok = passes_only(b);
	if (!passes_only(a) && !bok) {
		dk_set_error("internal error: a one-qubit gate left out of the passes");
		return DK_E_ARG;
	}
	const std::vector<PlanSpec> &best =
		bok && (b.size() < a.size() || !passes_only(a)) ? b : a;
	for (const PlanSpec &ps : best)
		build_pass(&c, ps.chosen, ps.gates);
	for (const PassPlan &pp : c.passes) {
		if (!pass_ops_valid(&c.pops[pp.first_op], pp.desc.nops, nq)) {
			dk_set_error("internal error: malformed pass plan");
			return DK_E_ARG;
		}
		descs.push_back(pp.desc);
		op_begin.push_back((uint32_t)(ops.size() + pp.first_op));
	}
	ops.insert(ops.end(), c.pops.begin(), c.pops.end());
	return DK_OK;
}

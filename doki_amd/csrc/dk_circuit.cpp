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
#include <algorithm>
#include <complex>
#include <cstring>
#include <vector>

#include "dk_internal.h"

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

struct dk_circuit {
	unsigned nq;
	std::vector<CGate> gates;
	std::vector<CBlock> blocks;
	unsigned fuse_k = 0;
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

extern "C" int dk_circuit_add(dk_circuit *c, unsigned k, const double *matrix,
			      const unsigned *targets, uint64_t ctrl_mask,
			      uint64_t actrl_mask)
{
	if (!c || !matrix || !targets || k < 1 || k > DK_MAX_GATE_QUBITS) {
		dk_set_error("gates act on 1..%d qubits", DK_MAX_GATE_QUBITS);
		return DK_E_ARG;
	}
	const u64 valid = c->nq >= 64 ? ~0ull : ((1ull << c->nq) - 1);
	CGate g;
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
	c->gates.push_back(std::move(g));
	// a new gate invalidates any previous plan
	drop_device_gates(c);
	c->blocks.clear();
	c->fuse_k = 0;
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
	if (c->fuse_k > 0) {
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

// Request serving over the state-vector core: continuous batching on one
// GPU, behind the C ABI (dk_server_*, dk_submit_circuit, dk_poll).
//
// A request builds one or two registers (from |0...0> through a list of
// one-qubit gates, or from a session register that stays resident between
// requests, with more gates applied to it) and ends in one operation: the
// marginals of every qubit, a full measurement with injected rolls, a
// reduced density, the join of its two registers with the marginals of the
// result, or nothing (the register is kept as a session).  Requests are
// admitted by the device memory their states need, first come first
// served, except that small requests (one new register of at most
// small_qubits qubits and a read-out the batched engine covers) pass a
// large request that does not fit yet.
//
// A batcher thread runs every admitted small request queued since its
// previous launch in one dk_batch launch, so arrivals during a launch join
// the next one; worker threads, each on an execution lane of its own, run
// the large requests as cache-blocked passes in place, several at once on
// the device.  A session register is used by one request at a time (in
// order of execution); its gates are applied to it in place.
//
// Every result depends only on its request: the batched engine runs each
// register in a fixed way whatever else is in the launch, and the lane path
// launches the same kernels with the same grids whatever runs beside it.
// The reference has no counterpart (its callers run registry_* calls one at
// a time, doki.c:302-1192).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "doki_hip.h"

typedef uint64_t u64;

namespace {

double now_s()
{
	// CLOCK_MONOTONIC on Linux, the clock of Python's time.perf_counter
	return std::chrono::duration<double>(
		       std::chrono::steady_clock::now().time_since_epoch())
		.count();
}

u64 state_bytes(unsigned nq, int dtype)
{
	return (u64)(dtype == DK_C64 ? 8 : 16) << nq;
}

struct SReg {
	unsigned nq = 0;
	u64 session = 0;
	std::vector<double> mats;
	std::vector<unsigned> tg;
	std::vector<u64> cm, am;
};

struct SReq {
	u64 ticket = 0;
	int op = 0;
	std::vector<SReg> regs;
	std::vector<double> rolls;
	u64 keep_mask = 0;
	int keep_result = 0;
	u64 need = 0;       // device bytes held while it runs
	u64 kept = 0;       // of them, bytes that stay with a new session
	bool small = false;
	// result
	int rc = DK_OK;
	std::string err;
	std::vector<double> probs, p1, rho;
	std::vector<int8_t> outcomes;
	u64 session_out = 0;
	double t_submit = 0, t_start = 0, t_done = 0;
	bool done = false;
};

struct Session {
	dk_state *st = nullptr;
	unsigned nq = 0;
	u64 bytes = 0;
	bool busy = false;
	bool dead = false;  // freed while a request still held it
};

} // namespace

struct dk_server {
	int dtype = DK_C128;
	unsigned small_qubits = 14;
	int plan = -1;               // -1 passes (12+ qubits), 0 gates, 3..6 fusion
	size_t max_batch = 2048;
	u64 budget = 0, used = 0;
	std::mutex mu;
	std::condition_variable cv_small, cv_large, cv_done, cv_idle;
	std::deque<SReq *> pending, small, large;
	std::map<u64, SReq *> reqs;  // every submitted request until polled
	std::map<u64, Session> sessions;
	u64 next_ticket = 1, next_session = 1;
	int running = 0;             // requests taken by the batcher or a worker
	bool closing = false, exit = false;
	std::vector<std::thread> threads;
	u64 batches = 0, batched = 0, completed = 0, rejected = 0;
};

namespace {

void finish(dk_server *S, SReq *q, int rc, const char *msg)
{
	// caller holds S->mu
	q->rc = rc;
	if (rc && msg)
		q->err = msg;
	q->t_done = now_s();
	q->done = true;
	S->completed += rc == DK_OK;
	S->cv_done.notify_all();
}

// Admission under S->mu: FIFO; a small request passes a waiting large one
// when it fits.
void admit_locked(dk_server *S)
{
	bool blocked = false;
	std::deque<SReq *> keep;
	while (!S->pending.empty()) {
		SReq *q = S->pending.front();
		S->pending.pop_front();
		const bool fits = S->used + q->need <= S->budget;
		if (fits && (!blocked || q->small)) {
			S->used += q->need;
			if (q->small) {
				S->small.push_back(q);
				S->cv_small.notify_one();
			} else {
				S->large.push_back(q);
				S->cv_large.notify_one();
			}
		} else {
			blocked = blocked || !q->small;
			keep.push_back(q);
		}
	}
	S->pending.swap(keep);
}

void release(dk_server *S, SReq *q)
{
	// caller holds S->mu; kept bytes move to the new session
	S->used -= q->need - q->kept;
	admit_locked(S);
}

int run_circuit(dk_server *S, dk_state *st, const SReg &r)
{
	const size_t ng = r.tg.size();
	if (!ng)
		return DK_OK;
	dk_circuit *c = nullptr;
	int rc = dk_circuit_new(dk_state_num_qubits(st), &c);
	if (!rc)
		rc = dk_circuit_add_many(c, ng, r.mats.data(), r.tg.data(), r.cm.data(),
					 r.am.data());
	if (!rc) {
		const unsigned n = dk_state_num_qubits(st);
		// (below 12 qubits the plan runs in one workgroup, one launch)
		if (S->plan < 0)
			rc = dk_circuit_passes(c, nullptr);
		else if (S->plan >= 3 && n >= (unsigned)S->plan + 6)
			rc = dk_circuit_fuse(c, (unsigned)S->plan, nullptr);
	}
	dk_state *out = nullptr;
	if (!rc)
		rc = dk_circuit_run(c, st, 1, &out);
	dk_circuit_free(c);
	return rc;
}

// Takes session `id` for the calling request (waits while another holds it).
Session *take_session(dk_server *S, u64 id)
{
	std::unique_lock<std::mutex> lk(S->mu);
	for (;;) {
		auto it = S->sessions.find(id);
		if (it == S->sessions.end() || it->second.dead)
			return nullptr;
		if (!it->second.busy) {
			it->second.busy = true;
			return &it->second;
		}
		S->cv_idle.wait(lk);
	}
}

void give_session(dk_server *S, u64 id)
{
	std::lock_guard<std::mutex> lk(S->mu);
	auto it = S->sessions.find(id);
	if (it == S->sessions.end())
		return;
	it->second.busy = false;
	if (it->second.dead) {
		dk_state_free(it->second.st);
		S->used -= it->second.bytes;
		S->sessions.erase(it);
		admit_locked(S);
	}
	S->cv_idle.notify_all();
}

// A large request on the calling thread's lane.  Its session registers
// are taken first, in ascending id order (two requests never wait on each
// other), and given back at the end.
int run_large(dk_server *S, SReq *q)
{
	dk_state *st[2] = {nullptr, nullptr};
	bool own[2] = {false, false};
	const unsigned nr = (unsigned)q->regs.size();
	std::vector<u64> ids;
	for (unsigned i = 0; i < nr; ++i)
		if (q->regs[i].session)
			ids.push_back(q->regs[i].session);
	std::sort(ids.begin(), ids.end());
	ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
	std::map<u64, dk_state *> held;
	int rc = DK_OK;
	for (u64 id : ids) {
		Session *s = take_session(S, id);
		if (!s) {
			rc = DK_E_ARG;
			q->err = "no such session register";
			break;
		}
		held[id] = s->st;
	}
	for (unsigned i = 0; i < nr && !rc; ++i) {
		const SReg &r = q->regs[i];
		if (r.session) {
			st[i] = held[r.session];
		} else {
			rc = dk_state_new_typed(r.nq, 1, S->dtype, &st[i]);
			own[i] = !rc;
		}
		if (!rc)
			rc = run_circuit(S, st[i], r);
	}
	if (!rc) {
		const unsigned n = dk_state_num_qubits(st[0]);
		switch (q->op) {
		case DK_OP_PROB:
			q->probs.resize(n);
			rc = dk_prob_all(st[0], q->probs.data());
			break;
		case DK_OP_MEASURE: {
			q->outcomes.resize(n);
			q->p1.resize(n);
			dk_state *left = nullptr;
			rc = dk_measure(st[0], n >= 64 ? ~0ull : ((1ull << n) - 1),
					q->rolls.data(), q->rolls.size(), q->outcomes.data(),
					q->p1.data(), &left);
			if (left)
				dk_state_free(left);
			break;
		}
		case DK_OP_DENSITY: {
			const unsigned a = (unsigned)__builtin_popcountll(q->keep_mask);
			q->rho.resize((size_t)2 << (2 * a));
			rc = dk_reduced_density(st[0], q->keep_mask, q->rho.data());
			break;
		}
		case DK_OP_JOIN: {
			const unsigned m = n + dk_state_num_qubits(st[1]);
			q->probs.resize(m);
			dk_state *j = nullptr;
			rc = dk_join_prob_all(st[0], st[1], &j, q->probs.data());
			if (!rc && q->keep_result) {
				std::lock_guard<std::mutex> lk(S->mu);
				const u64 id = S->next_session++;
				S->sessions[id] = Session{j, m, state_bytes(m, S->dtype), false, false};
				q->session_out = id;
			} else if (j) {
				dk_state_free(j);
			}
			break;
		}
		default:
			break;
		}
	}
	if (rc && q->err.empty())
		q->err = dk_last_error();
	// a new register kept as a session (for joins, the joined state is)
	if (!rc && q->keep_result && q->op != DK_OP_JOIN) {
		if (own[0]) {
			std::lock_guard<std::mutex> lk(S->mu);
			const u64 id = S->next_session++;
			const unsigned n = dk_state_num_qubits(st[0]);
			S->sessions[id] = Session{st[0], n, state_bytes(n, S->dtype), false, false};
			q->session_out = id;
			own[0] = false;
		} else {
			q->session_out = q->regs[0].session;
		}
	}
	(void)dk_synchronize();
	for (unsigned i = 0; i < nr; ++i)
		if (own[i])
			dk_state_free(st[i]);
	for (auto &kv : held)
		give_session(S, kv.first);
	return rc;
}

// Every thread of a server runs on a lane of its own, made by dk_server_new
// (a thread without one would share the default lane, its stream and its
// scratch, with any other such thread).
void worker(dk_server *S, dk_lane *lane)
{
	dk_lane_bind(lane);
	for (;;) {
		SReq *q;
		{
			std::unique_lock<std::mutex> lk(S->mu);
			S->cv_large.wait(lk, [&] { return S->exit || !S->large.empty(); });
			if (S->large.empty())
				break;
			q = S->large.front();
			S->large.pop_front();
			++S->running;
		}
		q->t_start = now_s();
		const int rc = run_large(S, q);
		std::lock_guard<std::mutex> lk(S->mu);
		if (rc)
			q->kept = 0;
		release(S, q);
		--S->running;
		finish(S, q, rc, q->err.c_str());
		S->cv_idle.notify_all();
	}
	if (lane)
		dk_lane_free(lane);
}

void batcher(dk_server *S, dk_lane *lane)
{
	dk_lane_bind(lane);
	dk_batch *b = nullptr;
	int brc = dk_batch_new(S->dtype, &b);
	if (!brc)
		brc = dk_batch_set(b, "probs", 1);
	for (;;) {
		std::vector<SReq *> items;
		{
			std::unique_lock<std::mutex> lk(S->mu);
			S->cv_small.wait(lk, [&] { return S->exit || !S->small.empty(); });
			if (S->small.empty())
				break;
			while (!S->small.empty() && items.size() < S->max_batch) {
				items.push_back(S->small.front());
				S->small.pop_front();
			}
			S->running += (int)items.size();
		}
		const double t0 = now_s();
		int rc = brc;
		if (!rc)
			rc = dk_batch_clear(b);
		// a register the engine refuses fails its own request only (the
		// batch is left as it was); idx: the register's place in the batch
		std::vector<size_t> idx(items.size(), SIZE_MAX);
		size_t added = 0;
		for (size_t i = 0; i < items.size() && !rc; ++i) {
			SReq *q = items[i];
			q->t_start = t0;
			const SReg &r = q->regs[0];
			const int ra = dk_batch_add(b, r.nq, r.tg.size(), r.mats.data(),
						    r.tg.data(), r.cm.data(), r.am.data(),
						    q->op == DK_OP_MEASURE ? q->rolls.data() : nullptr,
						    q->op == DK_OP_DENSITY ? q->keep_mask : 0, &idx[i]);
			if (ra) {
				q->rc = ra;
				q->err = dk_last_error();
				idx[i] = SIZE_MAX;
			} else {
				++added;
			}
		}
		if (!rc && added)
			rc = dk_batch_run(b);
		std::string err = rc ? dk_last_error() : "";
		for (size_t i = 0; i < items.size() && !rc; ++i) {
			SReq *q = items[i];
			if (idx[i] == SIZE_MAX)
				continue;
			const unsigned n = q->regs[0].nq;
			int r2 = DK_OK;
			if (q->op == DK_OP_PROB) {
				q->probs.resize(n);
				r2 = dk_batch_probs(b, idx[i], q->probs.data());
			} else if (q->op == DK_OP_MEASURE) {
				q->outcomes.resize(n);
				q->p1.resize(n);
				r2 = dk_batch_outcomes(b, idx[i], q->outcomes.data(), q->p1.data());
			} else {
				const unsigned a = (unsigned)__builtin_popcountll(q->keep_mask);
				q->rho.resize((size_t)2 << (2 * a));
				r2 = dk_batch_density(b, idx[i], q->rho.data());
			}
			if (r2) {
				q->rc = r2;
				q->err = dk_last_error();
			}
		}
		std::lock_guard<std::mutex> lk(S->mu);
		if (!rc && added) {
			S->batches++;
			S->batched += added;
		}
		for (SReq *q : items) {
			release(S, q);
			finish(S, q, rc ? rc : q->rc, rc ? err.c_str() : q->err.c_str());
		}
		S->running -= (int)items.size();
		S->cv_idle.notify_all();
	}
	if (b)
		dk_batch_free(b);
	if (lane)
		dk_lane_free(lane);
}

int copy_reg(const dk_reg_spec &in, SReg &out)
{
	out.nq = in.num_qubits;
	out.session = in.session;
	const size_t ng = in.num_gates;
	if (ng && (!in.matrices || !in.targets))
		return DK_E_ARG;
	out.mats.assign(in.matrices, in.matrices + 8 * ng);
	out.tg.assign(in.targets, in.targets + ng);
	out.cm.assign(ng, 0);
	out.am.assign(ng, 0);
	if (in.ctrl_masks)
		std::copy(in.ctrl_masks, in.ctrl_masks + ng, out.cm.begin());
	if (in.actrl_masks)
		std::copy(in.actrl_masks, in.actrl_masks + ng, out.am.begin());
	return DK_OK;
}

} // namespace

extern "C" int dk_server_new(int workers, unsigned small_qubits, int dtype,
			     uint64_t budget, dk_server **out)
{
	if (!out || workers < 1 || workers > 64 ||
	    (dtype != DK_C128 && dtype != DK_C64)) {
		return DK_E_ARG;
	}
	int rc = dk_init(0);
	if (rc)
		return rc;
	uint64_t lib_budget = 0, lib_used = 0;
	rc = dk_memory_budget(&lib_budget, &lib_used);
	if (rc)
		return rc;
	std::vector<dk_lane *> lanes((size_t)workers + 1, nullptr);
	for (auto &l : lanes) {
		rc = dk_lane_new(&l);
		if (rc) {
			for (auto &m : lanes)
				if (m)
					dk_lane_free(m);
			return rc;
		}
	}
	dk_server *S = new dk_server;
	S->dtype = dtype;
	S->small_qubits = std::min(small_qubits, 26u);
	S->budget = budget ? budget : lib_budget - std::min(lib_budget, lib_used);
	S->threads.emplace_back(batcher, S, lanes[0]);
	for (int i = 0; i < workers; ++i)
		S->threads.emplace_back(worker, S, lanes[(size_t)i + 1]);
	*out = S;
	return DK_OK;
}

extern "C" int dk_server_set(dk_server *S, const char *key, int64_t value)
{
	if (!S || !key)
		return DK_E_ARG;
	std::lock_guard<std::mutex> lk(S->mu);
	if (!strcmp(key, "plan")) {
		if (!(value == -1 || value == 0 || (value >= 3 && value <= 6)))
			return DK_E_ARG;
		S->plan = (int)value;
	} else if (!strcmp(key, "max_batch")) {
		if (value < 1)
			return DK_E_ARG;
		S->max_batch = (size_t)value;
	} else {
		return DK_E_ARG;
	}
	return DK_OK;
}

extern "C" int dk_submit_circuit(dk_server *S, const dk_request *r, uint64_t *ticket)
{
	if (!S || !r || !ticket)
		return DK_E_ARG;
	const unsigned want = r->op == DK_OP_JOIN ? 2u : 1u;
	if (r->op < DK_OP_PROB || r->op > DK_OP_KEEP || r->num_regs != want)
		return DK_E_ARG;
	SReq *q = new SReq;
	q->op = r->op;
	q->keep_mask = r->keep_mask;
	q->keep_result = r->keep_result || r->op == DK_OP_KEEP;
	q->regs.resize(want);
	int rc = DK_OK;
	for (unsigned i = 0; i < want && !rc; ++i)
		rc = copy_reg(r->regs[i], q->regs[i]);
	std::unique_lock<std::mutex> lk(S->mu);
	// register sizes (sessions: their own) and the request's memory
	unsigned nq[2] = {0, 0};
	for (unsigned i = 0; i < want && !rc; ++i) {
		SReg &g = q->regs[i];
		if (g.session) {
			auto it = S->sessions.find(g.session);
			if (it == S->sessions.end() || it->second.dead)
				rc = DK_E_ARG;
			else
				g.nq = it->second.nq;
		} else if (g.nq < 1 || g.nq > DK_MAX_QUBITS) {
			rc = DK_E_QUBITS;
		} else {
			q->need += state_bytes(g.nq, S->dtype);
		}
		nq[i] = g.nq;
		// every gate as the batched engine and the circuits take it, so
		// that an admitted request cannot fail another one's launch
		const u64 valid = g.nq >= 64 ? ~0ull : (1ull << g.nq) - 1;
		for (size_t k = 0; k < g.tg.size() && !rc; ++k) {
			const unsigned t = g.tg[k];
			const u64 cm = g.cm[k], am = g.am[k];
			if (t >= g.nq || ((cm | am) & ~valid) || (cm & am) ||
			    (((cm | am) >> t) & 1))
				rc = DK_E_ARG;
		}
	}
	if (!rc && q->op == DK_OP_MEASURE) {
		if (!r->rolls) {
			rc = DK_E_ROLL;
		} else {
			q->rolls.assign(r->rolls, r->rolls + nq[0]);
			for (double v : q->rolls)
				if (v < 0.0 || v >= 1.0)  // NaN passes, as the reference takes it
					rc = DK_E_ROLL;
		}
		q->need += state_bytes(nq[0], S->dtype); // the collapse chain
	}
	if (!rc && q->op == DK_OP_DENSITY &&
	    (q->keep_mask == 0 || (nq[0] < 64 && (q->keep_mask >> nq[0])) ||
	     __builtin_popcountll(q->keep_mask) > 12))
		rc = DK_E_ARG;
	if (!rc && q->op == DK_OP_JOIN && q->regs[0].session &&
	    q->regs[0].session == q->regs[1].session &&
	    (!q->regs[0].tg.empty() || !q->regs[1].tg.empty()))
		rc = DK_E_ARG;  // which of the two would the gates act on first?
	if (!rc && q->op == DK_OP_JOIN) {
		if (nq[0] + nq[1] > DK_MAX_QUBITS)
			rc = DK_E_QUBITS;
		q->need += state_bytes(nq[0] + nq[1], S->dtype);
		if (q->keep_result)
			q->kept = state_bytes(nq[0] + nq[1], S->dtype);
	} else if (!rc && q->keep_result && !q->regs[0].session) {
		q->kept = state_bytes(nq[0], S->dtype);
	}
	if (rc) {
		delete q;
		return rc;
	}
	q->small = (q->op == DK_OP_PROB || q->op == DK_OP_MEASURE ||
		    q->op == DK_OP_DENSITY) &&
		   !q->regs[0].session && !q->keep_result &&
		   nq[0] <= S->small_qubits &&
		   __builtin_popcountll(q->keep_mask) <= 6;
	q->ticket = S->next_ticket++;
	q->t_submit = now_s();
	S->reqs[q->ticket] = q;
	*ticket = q->ticket;
	if (q->need > S->budget) {
		S->rejected++;
		finish(S, q, DK_E_BUDGET, "the request's states exceed the serving budget");
		return DK_OK;
	}
	S->pending.push_back(q);
	admit_locked(S);
	return DK_OK;
}

extern "C" int dk_poll(dk_server *S, uint64_t ticket, int timeout_ms,
		       dk_result *res)
{
	if (!S)
		return DK_E_ARG;
	std::unique_lock<std::mutex> lk(S->mu);
	auto it = S->reqs.find(ticket);
	if (it == S->reqs.end())
		return DK_E_ARG;
	SReq *q = it->second;
	auto ready = [&] { return q->done; };
	if (timeout_ms < 0)
		S->cv_done.wait(lk, ready);
	else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
		return DK_E_PENDING;
	if (res) {
		if (res->probs && !q->probs.empty())
			memcpy(res->probs, q->probs.data(), q->probs.size() * sizeof(double));
		if (res->outcomes && !q->outcomes.empty())
			memcpy(res->outcomes, q->outcomes.data(), q->outcomes.size());
		if (res->p1 && !q->p1.empty())
			memcpy(res->p1, q->p1.data(), q->p1.size() * sizeof(double));
		if (res->rho && !q->rho.empty())
			memcpy(res->rho, q->rho.data(), q->rho.size() * sizeof(double));
		res->session = q->session_out;
		res->t_submit = q->t_submit;
		res->t_start = q->t_start;
		res->t_done = q->t_done;
		snprintf(res->error, sizeof(res->error), "%s", q->err.c_str());
	}
	const int rc = q->rc;
	S->reqs.erase(it);
	delete q;
	return rc;
}

extern "C" int dk_session_free(dk_server *S, uint64_t session)
{
	if (!S)
		return DK_E_ARG;
	std::lock_guard<std::mutex> lk(S->mu);
	auto it = S->sessions.find(session);
	if (it == S->sessions.end() || it->second.dead)
		return DK_E_ARG;
	if (it->second.busy) {
		it->second.dead = true;   // freed by the request holding it
		return DK_OK;
	}
	dk_state_free(it->second.st);
	S->used -= it->second.bytes;
	S->sessions.erase(it);
	admit_locked(S);
	return DK_OK;
}

extern "C" int dk_server_stats(dk_server *S, uint64_t *batches, uint64_t *batched,
			       uint64_t *completed, uint64_t *rejected)
{
	if (!S)
		return DK_E_ARG;
	std::lock_guard<std::mutex> lk(S->mu);
	if (batches)
		*batches = S->batches;
	if (batched)
		*batched = S->batched;
	if (completed)
		*completed = S->completed;
	if (rejected)
		*rejected = S->rejected;
	return DK_OK;
}

extern "C" void dk_server_free(dk_server *S)
{
	if (!S)
		return;
	{
		// every admitted request runs to its end; then the threads leave
		std::unique_lock<std::mutex> lk(S->mu);
		S->closing = true;
		// (pending requests that still do not fit once nothing runs wait
		// for session memory; they are dropped)
		S->cv_idle.wait(lk, [&] {
			return S->small.empty() && S->large.empty() && S->running == 0;
		});
		S->exit = true;
		S->cv_small.notify_all();
		S->cv_large.notify_all();
	}
	for (auto &t : S->threads)
		t.join();
	for (auto &kv : S->sessions)
		dk_state_free(kv.second.st);
	for (auto &kv : S->reqs)
		delete kv.second;
	delete S;
}

/*
 * doki_hip.h — C ABI of libdoki_hip.so, the MI355X (gfx950) state-vector core.
 *
 * Plain C: opaque handles, pointers and sizes only.  States live in device
 * memory (HBM) for their whole life; only scalars (probabilities, outcomes,
 * norms) and explicit read-outs cross to the host.  All work is enqueued on
 * the calling thread's lane (a HIP stream; see dk_lane_new); calls that
 * return host values synchronise it.
 *
 * Each entry point replaces one function of the reference's internal C API
 * (Mowstyl/Doki 1.6.1, src/doki/qops.h, qstate.h) or one step of its Python
 * wrappers (src/doki/doki.c); the citation is given per declaration.
 *
 * Value semantics follow the reference exactly: a state holds raw amplitudes
 * and a normalisation constant, and every read returns raw / norm
 * (qstate.h:84-86).  Amplitude index bit q is qubit q.  Complex numbers are
 * interleaved (re, im) doubles.
 *
 * Return codes (int): 0 ok; the nonzero values mirror the reference's exit
 * codes where one exists (qstate.c:23-88, qops.c:176-261).
 */
#ifndef DOKI_HIP_H_
#define DOKI_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DK_OK 0
#define DK_E_ALLOC 2        /* device allocation failed (qstate.c:44-82 code 2) */
#define DK_E_QUBITS 3       /* qubit count outside 1..62 (qstate.c:29-31) */
#define DK_E_ARG 5          /* invalid argument (ids, sizes, null handles) */
#define DK_E_ROLL 6         /* measurement roll missing or outside [0, 1) */
#define DK_E_ZERO_NORM 7    /* post-collapse normalisation is 0 */
#define DK_E_DEVICE 8       /* HIP runtime / device failure */
#define DK_E_BUDGET 9       /* state would exceed the device memory budget */

#define DK_MAX_QUBITS 62    /* MAX_NUM_QUBITS = log2_64(INTMAX_MAX), platform.h:270 */
#define DK_MAX_GATE_QUBITS 6

/* Amplitude storage types.  DK_C128 is the reference's COMPLEX_TYPE
 * (double _Complex, platform.h:199-216).  DK_C64 (float2, FP32 arithmetic)
 * is an extra mode with no reference counterpart; its norms, probabilities
 * and renormalisations are still computed in FP64. */
#define DK_C128 0
#define DK_C64 1

typedef struct dk_state dk_state;
typedef struct dk_gate dk_gate;
typedef struct dk_circuit dk_circuit;
typedef struct dk_lane dk_lane;

/* Library set-up on a HIP device (first call wins; later calls are no-ops).
 * No reference counterpart: the reference has no device. */
int dk_init(int device);

/* Message describing the last nonzero return code on this thread.
 * Replaces the PyErr_SetString texts of doki.c at the C level. */
const char *dk_last_error(void);

/* Wait for all queued work.  No reference counterpart (the CPU path is
 * synchronous). */
int dk_synchronize(void);

/* Execution lanes.  A lane is a HIP stream with its own reduction and
 * measurement scratch; each host thread issues its work on the lane it bound
 * last (the library's default lane until it binds one; dk_lane_bind(NULL)
 * returns to it), so threads on different lanes run concurrently on the
 * device.  An operation on a state last written on another lane first waits
 * for that lane.  Reads of one state issued concurrently from several lanes
 * are ordered by the caller.  dk_synchronize waits for the calling thread's
 * lane.  A lane serves one thread at a time: threads that call the library
 * concurrently each bind a lane of their own (the default lane included).
 * dk_lane_free waits for the lane's work and keeps it for a later
 * dk_lane_new; a thread still bound to it returns to the default lane.
 * No reference counterpart: they serve concurrent request serving. */
int dk_lane_new(dk_lane **out);
int dk_lane_bind(dk_lane *lane);
void dk_lane_free(dk_lane *lane);

/* Device-memory budget in bytes for live states (default: 92% of free HBM at
 * init); allocations beyond it fail with DK_E_BUDGET.  `used` may be NULL. */
int dk_memory_budget(uint64_t *budget, uint64_t *used);
int dk_set_memory_budget(uint64_t budget);

/* Device memory held by the library beyond live states: 2 GiB chunks of the
 * state arena (all, and unmapped ones) and bytes of released blocks kept for
 * reuse.  Any pointer may be NULL.  No reference counterpart. */
int dk_memory_stats(uint64_t *arena_chunks, uint64_t *arena_free_chunks,
		    uint64_t *cached_bytes);

/* Returns cached blocks and every arena chunk no live state needs to the
 * driver (which wipes memory before handing it out again, at tens of GB/s).
 * No reference counterpart. */
int dk_memory_trim(void);

/* ---- states: qstate.c / qstate.h ---------------------------------------- */

/* |0...0> when zero_state != 0, else uninitialised; norm 1.
 * Replaces state_init, qstate.c:23-88 (via registry_new, doki.c:302-341). */
int dk_state_new(unsigned num_qubits, int zero_state, dk_state **out);

/* State holding `values` (2^n interleaved complex) as given, norm 1 — not
 * normalised.  Replaces registry_new_data + custom_state_init_*,
 * doki.c:597-720. */
int dk_state_from_host(unsigned num_qubits, const double *values,
		       dk_state **out);

/* Normalised copy (raw / norm, norm 1).  Replaces state_clone,
 * qstate.c:90-105 (registry_clone, doki.c:343-398). */
int dk_state_clone(const dk_state *src, dk_state **out);

/* A complex128 state holding 2^n values that are already in device memory
 * of the library's device (another library's tensor, for example), copied as
 * given, norm 1 (as dk_state_from_host).  Waits for all work queued on the
 * device first, whatever stream queued it.  Replaces registry_new_data,
 * doki.c:625-720, for device-resident input (no host round trip). */
int dk_state_from_device(unsigned num_qubits, const void *values,
			 dk_state **out);

/* A new complex128 state holding src's values (each raw component divided
 * by the norm, as state_get reads them, qstate.h:86), norm 1: its raw
 * amplitudes are the values, equal bit for bit to what registry_get
 * returns.  For handing a state to another library.  No reference
 * counterpart. */
int dk_state_values(const dk_state *src, dk_state **out);

/* Device address of s's raw amplitudes (2^n interleaved complex of s's
 * type) and the HIP device ordinal holding them, once the work queued on s
 * is complete (the calling thread waits).  Valid until s is freed.  No
 * reference counterpart (the reference's states never leave the host). */
int dk_state_device_address(const dk_state *s, void **amp, int *device);

/* dk_state_new with a storage type (DK_C128 or DK_C64).  No reference
 * counterpart for DK_C64 (the reference fixes PRECISION 2, platform.h:199). */
int dk_state_new_typed(unsigned num_qubits, int zero_state, int dtype,
		       dk_state **out);

/* Storage type of a state (DK_C128 or DK_C64). */
int dk_state_dtype(const dk_state *s);

/* Normalised copy into another storage type (state_clone with a type
 * change; qstate.c:90-105 for DK_C128 -> DK_C128). */
int dk_state_convert(const dk_state *src, int dtype, dk_state **out);

/* Releases the state's device memory.  Replaces state_clear,
 * qstate.c:107-121 (registry_del / capsule destructor, doki.c:258-270). */
void dk_state_free(dk_state *s);

unsigned dk_state_num_qubits(const dk_state *s);

/* Bytes the reference would report: 64 + 16 * 2^n on LP64.  Replaces
 * state_mem_size, qstate.c:123-135 (registry_mem, doki.c:1272-1295). */
uint64_t dk_state_mem_size(const dk_state *s);

/* Bytes of device memory this state holds. */
uint64_t dk_state_device_bytes(const dk_state *s);

/* The state's normalisation constant (reads return raw / norm). */
int dk_state_norm(const dk_state *s, double *norm);

/* ---- gates: qgate.h ----------------------------------------------------- */

/* A k-qubit gate (1 <= k <= 6): 2^k x 2^k complex matrix, row-major,
 * interleaved; row = output basis state.  Copied to the device once.
 * Replaces struct qgate, qgate.h:25-32, built by gate_new, doki.c:417-509. */
int dk_gate_new(unsigned num_qubits, const double *matrix, dk_gate **out);
void dk_gate_free(dk_gate *g);

/* ---- operations: qops.c ------------------------------------------------- */

/* out = gate applied to in: for every index whose control bits (ctrl_mask)
 * are all 1 and anticontrol bits (actrl_mask) all 0,
 *   out_i = sum_j M[row(i)][j] * in(i with target bits <- j),
 * row(i) = sum_b bit(i, targets[b]) << b; other indices copy in_i.
 * out's norm = sqrt(sum |out_i|^2).  Input values are read normalised.
 * When in_place != 0 the result overwrites `in` and *out is set to `in`.
 * Replaces apply_gate, qops.c:176-261 (registry_apply, doki.c:722-941). */
int dk_apply(dk_state *in, const dk_gate *gate, const unsigned *targets,
	     uint64_t ctrl_mask, uint64_t actrl_mask, int in_place,
	     dk_state **out);

/* dk_apply writing into `out`, a live state the caller owns with in's qubit
 * count and amplitude type (out == in applies in place), that allocates no
 * device memory: the block partials of the norm live in a fixed device array
 * of DK_NOALLOC_SCRATCH_BYTES, so states of at most DK_NOALLOC_MAX_QUBITS
 * qubits are accepted (DK_E_ARG beyond).  Values and norm equal dk_apply's.
 * Every call on the device shares that array: calls that overlap on the
 * device (issued from different lanes) race and must be ordered by the
 * caller.  The library must be initialised (dk_init, or any earlier call).
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that forbid it). */
#define DK_NOALLOC_SCRATCH_BYTES (16u << 20)
#define DK_NOALLOC_MAX_QUBITS 26
int dk_apply_noalloc(dk_state *in, const dk_gate *gate,
		     const unsigned *targets, uint64_t ctrl_mask,
		     uint64_t actrl_mask, dk_state *out);

/* P(qubit q = 1) = sum over indices with bit q set of |raw/norm|^2.
 * Replaces probability, qops.c:72-95 (registry_prob, doki.c:1194-1228). */
int dk_prob(const dk_state *s, unsigned q, double *p);

/* All n single-qubit probabilities in one pass over the state.  Same values
 * as n calls of dk_prob (summation order differs).  No single reference
 * counterpart: batches n calls of probability, qops.c:72-95. */
int dk_prob_all(const dk_state *s, double *p);

/* Measurement of the qubits in `mask`, most significant first, consuming
 * one roll per measured qubit: outcome = P1 > roll (a NaN roll passes the
 * [0, 1) check and gives 0, as in the reference); the kept half is
 * renormalised by sqrt(P), P = P1 or 1 - P1.  results[i] describes qubit
 * n-1-i: -1 not measured, 0 or 1.  p1_out (may be NULL) receives P1 per
 * results[] slot.  *out is the remaining state, or NULL when every qubit was
 * measured.  Replaces the loop of registry_measure, doki.c:1023-1192, with
 * measure / collapse, qops.c:124-174. */
int dk_measure(const dk_state *s, uint64_t mask, const double *rolls,
	       size_t num_rolls, int8_t *results, double *p1_out,
	       dk_state **out);

/* Kronecker join, `most` in the high qubits: r[i * 2^n2 + j] = a_i * b_j,
 * norm 1.  Replaces join, qops.c:97-122 (registry_join, doki.c:943-1021). */
int dk_join(const dk_state *most, const dk_state *least, dk_state **out);

/* dk_join and the marginal probabilities of every qubit of the joined state
 * in one pass over it: *out as dk_join gives it, p[q] = P(qubit q = 1) as
 * dk_prob_all gives it for *out (p holds most.nq + least.nq doubles).  The
 * state is written once and not read back.  No counterpart in the reference,
 * where a caller joins (registry_join, doki.c:943-1021) and then calls
 * registry_prob (doki.c:1194-1228) once per qubit; for servers and circuit
 * runners that join registers to read their marginals. */
int dk_join_prob_all(const dk_state *most, const dk_state *least,
		     dk_state **out, double *p);

/* Index of the first amplitude whose raw value is nonzero (2^n if none).
 * Device half of get_global_phase, qops.c:46-70. */
int dk_first_nonzero(const dk_state *s, uint64_t *index);

/* Overlap and distance of two states of the same size (either amplitude
 * type on each side), over normalised amplitudes: inner[0..1] = re, im of
 * sum_i conj(a_i) * b_i (|inner|^2 is the fidelity) and *max_abs =
 * max_i |a_i - b_i|.  One pass over both states, FP64 partials combined in a
 * fixed order.  No counterpart in the reference; for checking fused and
 * complex64 results against complex128 ones.  Either output may be NULL. */
int dk_state_compare(const dk_state *a, const dk_state *b, double *inner,
		     double *max_abs);

/* Raw stored amplitudes [first, first+count) and the norm constant, so the
 * caller divides exactly as state_get does (qstate.h:86).  Read-out of
 * registry_get, doki.c:546-595. */
int dk_read_raw(const dk_state *s, uint64_t first, uint64_t count,
		double *values, double *norm);

/* Normalised values [first, first+count) (raw * (1/norm) on the device). */
int dk_read(const dk_state *s, uint64_t first, uint64_t count,
	    double *values);

/* Reduced density matrix of the qubits in keep_mask (ascending qubit order =
 * reduced index bit order), tracing out the rest:
 *   rho[i][j] = sum_t psi(i, t) * conj(psi(j, t)),
 * written row-major interleaved, dim = 2^popcount(keep_mask) <= 2^12.
 * Replaces the lazy chain density_matrix / _densityFun (qops.c:266-293,
 * 580-601) read through partial_trace / _PartialTFunct (funmatrix.c:905-1037)
 * and funmatrix_get (doki.c:1499-1579). */
int dk_reduced_density(const dk_state *s, uint64_t keep_mask, double *rho);

/* Selected elements of the same reduced density matrix, for shapes too large
 * to materialise: out[e] = rho[rows[e]][cols[e]]. */
int dk_reduced_density_elements(const dk_state *s, uint64_t keep_mask,
				size_t count, const uint64_t *rows,
				const uint64_t *cols, double *out);

/* The same two with projections folded in: only the terms of the traced
 * sum whose state index satisfies (index & traced_mask) == traced_value are
 * added (traced_mask names traced qubits only).  This is the partial trace
 * of a projected density: a projection keeps the rows whose bit has the
 * given value (funmatrix.c:837-918), so tracing that qubit out afterwards
 * (funmatrix.c:920-1037) keeps only the term with that value. */
int dk_reduced_density_masked(const dk_state *s, uint64_t keep_mask,
			      uint64_t traced_mask, uint64_t traced_value,
			      double *rho);
int dk_reduced_density_elements_masked(const dk_state *s, uint64_t keep_mask,
				       uint64_t traced_mask,
				       uint64_t traced_value, size_t count,
				       const uint64_t *rows, const uint64_t *cols,
				       double *out);

/* ---- gate lists and fusion ------------------------------------------------
 * No single reference counterpart: a caller (QSimov's circuit layer) issues
 * the same gates one registry_apply at a time (doki.c:722-941).  A circuit
 * records them with registry_apply's argument meaning; fusion groups them
 * into dense blocks of at most max_block_qubits qubits (controls absorbed),
 * each applied as one gate, so a run costs one state pass per block. */

int dk_circuit_new(unsigned num_qubits, dk_circuit **out);
void dk_circuit_free(dk_circuit *c);

/* Appends a gate: as dk_apply's gate/targets/ctrl_mask/actrl_mask. */
int dk_circuit_add(dk_circuit *c, unsigned k, const double *matrix,
		   const unsigned *targets, uint64_t ctrl_mask,
		   uint64_t actrl_mask);

/* Appends count one-qubit gates in one call: gate i is the 2x2 matrix at
 * matrices + 8 i (row-major, interleaved), on targets[i], with ctrl_masks[i]
 * and actrl_masks[i] (either array may be NULL: no controls), checked as
 * dk_circuit_add checks it.  All or nothing: on an invalid gate none of the
 * call's gates is kept.  No counterpart in the reference (a circuit layer
 * above registry_apply, doki.c:722-941, submitting its gate list at once). */
int dk_circuit_add_many(dk_circuit *c, size_t count, const double *matrices,
			const unsigned *targets, const uint64_t *ctrl_masks,
			const uint64_t *actrl_masks);

/* Plans fused blocks of at most max_block_qubits (1..6) qubits; 0 runs the
 * gates one by one.  *num_blocks (may be NULL) receives the block count. */
int dk_circuit_fuse(dk_circuit *c, unsigned max_block_qubits,
		    size_t *num_blocks);

int dk_circuit_num_gates(const dk_circuit *c, size_t *num_gates);

/* Block i of the plan: its qubits (ascending; local bit b = qubits[b]) and
 * its dense 2^m x 2^m unitary (row-major, interleaved).  Either output may
 * be NULL. */
int dk_circuit_block(const dk_circuit *c, size_t i, unsigned *num_qubits,
		     unsigned *qubits, double *matrix);

/* Plans cache-blocked passes instead of fused blocks: each pass streams the
 * state through registers once and applies every one-qubit gate whose target
 * lies in the pass's tile (qubits 0..5 plus six chosen qubits), and every
 * diagonal one-qubit gate wherever its target is; controls may be anywhere.
 * Gates are moved ahead of gates left for a later pass only where they
 * commute.  Gates on several qubits run alone between passes.  States of
 * fewer than 12 qubits run gate by gate.  *num_passes (may be NULL)
 * receives the pass count.  Results equal gate-by-gate application up to
 * rounding (the state is renormalised once per pass). */
int dk_circuit_passes(dk_circuit *c, size_t *num_passes);

/* Pass i of a pass plan: the gates it absorbs, its op count (gates after
 * merging same-target neighbours, plus register exchanges) and its 12 tile
 * qubits (tile_qubits[0..5] = 0..5).  Any output may be NULL. */
int dk_circuit_pass(const dk_circuit *c, size_t i, unsigned *num_gates,
		    unsigned *num_ops, unsigned *tile_qubits);

/* Step i of a pass plan, in execution order: *is_pass = 1 and *index = the
 * pass, or *is_pass = 0 and *index = the gate (in order of dk_circuit_add)
 * that runs alone.  DK_E_ARG past the last step. */
int dk_circuit_plan_step(const dk_circuit *c, size_t i, int *is_pass,
			 size_t *index);

/* Copies the op records of pass i (96 bytes each, the PassOp layout of
 * doki_amd/csrc/dk_pass.h) into ops (room for `capacity` records); for
 * inspection and host-side checking of plans.  *count receives the number
 * of records. */
int dk_circuit_pass_ops(const dk_circuit *c, size_t i, void *ops,
			size_t capacity, size_t *count);

/* Runs the circuit on st: the first step writes a new state (*out) unless
 * in_place, later steps run in place.  Works on DK_C128 and DK_C64 states. */
int dk_circuit_run(dk_circuit *c, dk_state *st, int in_place, dk_state **out);

/* ---- batches of small registers ---------------------------------------------
 * No single reference counterpart: the reference runs each register through
 * registry_new / registry_apply / registry_density / registry_measure
 * (doki.c:302-1192) one call at a time.  A batch holds many independent
 * registers (1..26 qubits, each starting at |0...0>), each with a list of
 * (anti)controlled 1-qubit gates, an optional reduced density matrix of at
 * most 6 kept qubits (taken after the gates) and an optional measurement of
 * every qubit with injected rolls; dk_batch_run executes all of them in one
 * persistent kernel launch. */
typedef struct dk_batch dk_batch;

int dk_batch_new(int dtype, dk_batch **out);
void dk_batch_free(dk_batch *b);

/* Settings: "export_states" (keep final states readable by dk_batch_state),
 * "use_graph" (capture the launch as a HIP graph, replay on later runs),
 * "pass_qubits" (target qubits per cache-blocked pass, 0 = tile width),
 * "probs" (every register reports its marginal qubit probabilities),
 * "split_qubits" (registers of at least this many qubits, and larger than
 * one LDS tile, run their passes as grid-wide rounds before the persistent
 * launch; 14 by default, 0 = never; results agree to rounding either way,
 * the rounds merging neighbouring gates), "poison_pool" (NaN over the
 * device pool before each run; the pool is cleared only when states are
 * exported, and no result reads an amplitude the run did not write, which
 * this setting lets a test show). */
int dk_batch_set(dk_batch *b, const char *key, int64_t value);

/* Drops every register and keeps the device and pinned buffers for the next
 * fill (a serving loop refills one batch per launch). */
int dk_batch_clear(dk_batch *b);

/* Adds a register.  matrices: num_gates 2x2 gates, 8 doubles each
 * (row-major, interleaved); targets/ctrl_masks/actrl_masks per gate, with
 * registry_apply's meaning (masks may be NULL).  rolls: num_qubits rolls in
 * [0, 1) consumed from the most significant qubit down, or NULL for no
 * measurement.  density_keep: kept qubits of the reduced density (0 = none).
 * *index (may be NULL) receives the register's position. */
int dk_batch_add(dk_batch *b, unsigned num_qubits, size_t num_gates,
		 const double *matrices, const unsigned *targets,
		 const uint64_t *ctrl_masks, const uint64_t *actrl_masks,
		 const double *rolls, uint64_t density_keep, size_t *index);

/* Runs every register of the batch in one persistent launch. */
int dk_batch_run(dk_batch *b);

int dk_batch_size(const dk_batch *b, size_t *num_registers);

/* Outcomes of register i in registry_measure's list order (slot s is qubit
 * num_qubits-1-s; -1 when not measured) and the P1 each was decided on. */
int dk_batch_outcomes(const dk_batch *b, size_t i, int8_t *results,
		      double *p1);

/* Marginal P(qubit q = 1) of register i's final state for q = 0..nq-1, as
 * registry_prob (doki.c:1194-1228) would return them (needs "probs"). */
int dk_batch_probs(const dk_batch *b, size_t i, double *p);

/* Reduced density of register i: 2^a x 2^a (a = kept qubits), row-major,
 * interleaved; kept qubits in ascending order are the index bits. */
int dk_batch_density(const dk_batch *b, size_t i, double *rho);

/* Final normalised state of register i as complex128 (needs
 * "export_states"). */
int dk_batch_state(const dk_batch *b, size_t i, double *values);

/* ---- serving -----------------------------------------------------------------
 * Continuous batching of circuit requests on the library's device.  A
 * request builds one register (two for a join) -- from |0...0> through a
 * list of one-qubit gates, or from a session register that stays resident
 * on the device between requests, with more gates applied to it in place --
 * and ends in one read-out.  Requests are admitted by the device memory
 * their states need, first come first served (a small request passes a
 * large one that does not fit yet); small ones (one new register of at most
 * small_qubits qubits read by DK_OP_PROB, _MEASURE or _DENSITY) run together
 * in one batched launch per round, the others on worker threads with
 * execution lanes of their own, as cache-blocked passes.  Results equal
 * those of the same request run alone.  No single reference counterpart:
 * the reference's callers issue registry_new / registry_apply /
 * registry_prob / registry_measure / registry_density / registry_join
 * (doki.c:302-1228) one call at a time. */
typedef struct dk_server dk_server;

#define DK_E_PENDING 10     /* dk_poll: the request has not finished yet */

#define DK_OP_PROB 0        /* P(qubit q = 1) of every qubit -> probs[n] */
#define DK_OP_MEASURE 1     /* measurement of every qubit, MSB first, with
			       rolls[n] -> outcomes[n] (registry_measure's
			       list order), p1[n] */
#define DK_OP_DENSITY 2     /* reduced density over keep_mask -> rho */
#define DK_OP_JOIN 3        /* join(regs[0], regs[1]), regs[0] the high
			       qubits, and the marginals of the result ->
			       probs[n0 + n1] */
#define DK_OP_KEEP 4        /* no read-out: the register becomes a session */

/* A register of a request: `session` nonzero names a live session register
 * (num_qubits is then ignored), else a new |0...0> of num_qubits qubits.
 * The gates (2x2, as dk_circuit_add_many takes them) are applied to it in
 * order; for a session register they change the session. */
typedef struct {
	unsigned num_qubits;
	uint64_t session;
	size_t num_gates;
	const double *matrices;
	const unsigned *targets;
	const uint64_t *ctrl_masks;   /* may be NULL */
	const uint64_t *actrl_masks;  /* may be NULL */
} dk_reg_spec;

typedef struct {
	int op;                  /* DK_OP_* */
	unsigned num_regs;       /* 2 for DK_OP_JOIN, else 1 */
	dk_reg_spec regs[2];
	const double *rolls;     /* DK_OP_MEASURE */
	uint64_t keep_mask;      /* DK_OP_DENSITY: kept qubits (at most 12) */
	int keep_result;         /* nonzero: the request's register (the joined
				    one for a join) stays resident as a
				    session; DK_OP_KEEP implies it */
} dk_request;

typedef struct {
	double *probs;           /* each may be NULL; sizes as the op says */
	int8_t *outcomes;
	double *p1;
	double *rho;
	uint64_t session;        /* the kept session register, or 0 */
	double t_submit, t_start, t_done; /* CLOCK_MONOTONIC seconds */
	char error[256];         /* message of a failed request */
} dk_result;

/* A server with `workers` worker threads (1..64) and one batcher; budget 0
 * takes the library's free memory budget. */
int dk_server_new(int workers, unsigned small_qubits, int dtype,
		  uint64_t budget, dk_server **out);
/* Settings: "plan" (-1 cache-blocked passes from 12 qubits on, the default;
 * 0 gate by gate; 3..6 dense fusion width), "max_batch" (requests per
 * batched launch, 2048 by default). */
int dk_server_set(dk_server *s, const char *key, int64_t value);
/* Copies the request (the caller's arrays may go at once) and queues it;
 * *ticket names it for dk_poll.  Requests whose states exceed the budget
 * complete at once with DK_E_BUDGET. */
int dk_submit_circuit(dk_server *s, const dk_request *r, uint64_t *ticket);
/* Waits up to timeout_ms (< 0: no limit) for the request: DK_E_PENDING if it
 * has not finished, else its own return code, with the results copied to
 * *res (may be NULL); the ticket is then consumed. */
int dk_poll(dk_server *s, uint64_t ticket, int timeout_ms, dk_result *res);
/* Releases a session register (after the request using it, if one does). */
int dk_session_free(dk_server *s, uint64_t session);
int dk_server_stats(dk_server *s, uint64_t *batches, uint64_t *batched,
		    uint64_t *completed, uint64_t *rejected);
/* Runs every admitted request to its end, stops the threads and frees the
 * session registers; requests not polled by then are dropped. */
void dk_server_free(dk_server *s);

/* ---- timing, tuning, device ------------------------------------------------
 * No reference counterparts. */

/* Records event `slot` (0..63) on the library stream. */
int dk_event_record(int slot);
/* Milliseconds between two recorded events (waits for the second). */
int dk_event_elapsed(int first_slot, int second_slot, float *ms);

/* Durations (ms) of the gate-kernel launches recorded in the calling lane's
 * timing ring since the last read (dk_tune "time_ring" = capacity in
 * launches; every gate, dense-block and pass launch records a start and end
 * event on its own stream, without synchronising).  Waits for the last
 * recorded launch, then empties the ring.  *count (may be NULL) receives the
 * number of durations; ms may be NULL to query it.  No reference
 * counterpart: measurement. */
int dk_time_ring_read(float *ms, size_t capacity, size_t *count);
/* Tunables: "apply_blocks" (grid cap of grid-stride kernels, 0 = auto),
 * "time_apply" (events 62/63 bracket every gate kernel), "dense_mfma"
 * (dense k >= 4 gates on MFMA, else on the vector ALUs), "apply1_group"
 * (1, 2 or 4 pair groups per block of the 1-qubit gate kernels),
 * "dense_stream" (dense gates on MFMA: 0 LDS tiles, 1 the faster form per
 * amplitude type, k and low target qubits (default), 2 the streaming
 * form), "dense_k4_cw"
 * (1, 2 or 4 sets of 64 amplitude groups per dense 4-qubit block, 0 =
 * by amplitude type), "arena_max_chunks"
 * (cap on the 2 GiB chunks of the state arena; 0 restores the default, the
 * free memory at init less 6 GiB), "arena_sub_mib" (chunks created from
 * then on are made of physical pieces of this many MiB, a power of two,
 * mapped in a scrambled order; 0, the default, one piece per chunk),
 * "arena_reuse" (1, the default: whole-chunk blocks are carved from larger
 * cached blocks and neighbouring cached blocks merge; 0: only exact sizes
 * are reused), "poison_alloc" (1: every device block the library hands out,
 * states, scratch and gate matrices, is first filled with all-ones bytes, a
 * NaN in every amplitude, so that a read of memory nothing wrote shows in
 * the results; a testing aid, 0 by default), "guard_alloc" (1: every
 * block below 64 MiB gets 64 KiB guard bands of a known pattern before and
 * after it; its release synchronises the lane's stream, counts changed
 * guard words and reports them on stderr with the block's size; a testing
 * aid, 0 by default). */
int dk_tune(const char *key, int64_t value);
/* Compute units and architecture name of the library's device. */
int dk_device_info(int *num_cus, char *arch, size_t arch_len);
/* The HIP device ordinal the library runs on (initialising it). */
int dk_device_ordinal(int *device);

#ifdef __cplusplus
}
#endif

#endif /* DOKI_HIP_H_ */

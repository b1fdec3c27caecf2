// Internal declarations shared by the HIP sources of libdoki_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <utility>
#include <vector>

#include "doki_hip.h"

typedef uint64_t u64;

// ---- error reporting -------------------------------------------------------

void dk_set_error(const char *fmt, ...);

#define DK_HIP(call)                                                        \
	do {                                                                \
		hipError_t e_ = (call);                                     \
		if (e_ != hipSuccess) {                                     \
			dk_set_error("%s failed: %s", #call,                \
				     hipGetErrorString(e_));                \
			return DK_E_DEVICE;                                 \
		}                                                           \
	} while (0)

#define DK_TRY(call)                                                        \
	do {                                                                \
		int rc_ = (call);                                           \
		if (rc_)                                                    \
			return rc_;                                         \
	} while (0)

#define DK_LAUNCHED()                                                       \
	do {                                                                \
		hipError_t e_ = hipGetLastError();                          \
		if (e_ != hipSuccess) {                                     \
			dk_set_error("kernel launch failed: %s",            \
				     hipGetErrorString(e_));                \
			return DK_E_DEVICE;                                 \
		}                                                           \
	} while (0)

// ---- device-resident objects -------------------------------------------------

// Per-state device scalars, stored after the amplitudes.
//   [0] norm constant (reads return raw / norm)
enum { DK_SCAL_NORM = 0, DK_SCAL_BYTES = 256 };

struct dk_lane;

struct dk_state {
	unsigned nq;
	int dtype;      // DK_C128 or DK_C64
	dk_lane *lane;  // lane whose stream last wrote the state
	u64 len;        // 2^nq amplitudes
	char *base;     // device allocation: amplitudes (then scalars, unless
	                // sbase holds them)
	char *sbase;    // separate scalars block of arena-sized states, or null
	double *scal;   // device scalars
	void *amp;      // device amplitudes (raw), double2 or float2
	u64 bytes;      // bytes of the allocation
	// index bits that may be set in a nonzero amplitude: every amplitude
	// whose index has a bit outside `support` is zero (~0: no knowledge;
	// a new |0...0> state has 0, gates and passes widen it)
	u64 support;
	double2 *a128() const { return (double2 *)amp; }
	float2 *a64() const { return (float2 *)amp; }
};

static inline size_t dk_amp_size(int dtype)
{
	return dtype == DK_C64 ? sizeof(float2) : sizeof(double2);
}

struct dk_gate {
	unsigned k;
	unsigned dim;       // 2^k
	double2 *dmat;      // device FP64 copy, row-major (k > 1)
	float2 *dmat32;     // device FP32 copy, row-major (k > 1)
	dk_lane *lane;      // lane that uploaded dmat
	hipEvent_t uploaded = nullptr;  // recorded after the upload (k > 1)
	// per lane, an event recorded after its last launch reading dmat (the
	// gate's release waits for them); launches on several lanes may read
	// the gate concurrently
	mutable std::mutex uses_mu;
	mutable std::vector<std::pair<dk_lane *, hipEvent_t>> uses;
	double2 hmat[64 * 64];
	float2 hmat32[64 * 64]; // source of the FP32 upload
	bool finite = true;     // every matrix entry is finite
};

// ---- library context ---------------------------------------------------------

enum { DK_PARTIALS = 1 << 20 }; // initial doubles of reduction scratch

// An execution lane: a HIP stream with its own reduction scratch.  Each host
// thread works on its bound lane (the default lane unless dk_lane_bind chose
// another), so threads on different lanes run concurrently on the device.
struct dk_lane {
	hipStream_t stream = nullptr;
	double *partials = nullptr;   // reduction scratch, grows on demand
	u64 partials_cap = 0;         // doubles held by `partials`
	double *dec = nullptr;        // measurement decisions scratch (device)
	double *host_pinned = nullptr;// pinned staging (DK_PINNED doubles)
	hipEvent_t events[64] = {};
	// timing ring: event pairs recorded around every gate-kernel launch on
	// this lane while ring_cap > 0 (dk_tune "time_ring")
	hipEvent_t *ring = nullptr;   // 2 * ring_cap events
	size_t ring_cap = 0, ring_used = 0;
};

struct dk_ctx {
	int device = -1;
	int num_cus = 0;
	bool ready = false;
	dk_lane lane0;                // the default lane
	u64 budget = 0, used = 0;
	std::mutex mu;
	// tunables
	int apply_blocks = 0;   // 0 = auto
	int time_apply = 0;     // record events[62]/[63] around the gate kernel
	int dense_mfma = 1;     // dense k>=4 gates on MFMA (else VALU tiles)
	int apply1_group = 1;   // pair groups / waves per 1-qubit gate block
	int dense_k4_cw = 0;    // column sets per dense k = 4 block (0 = auto)
	int dense_stream = 1;   // dense form: 0 tiles, 1 by type and k, 2 streaming
	u64 arena_max_chunks = 0;     // cap on 2 GiB state-arena chunks
	u64 arena_sub_bytes = 0;  // physical pieces of an arena chunk (0: one piece)
	int arena_reuse = 1;      // whole-chunk blocks: carve and merge (0: exact sizes)
	int poison_alloc = 0;     // all-ones bytes over every block handed out
	int guard_alloc = 0;      // guard bands around pool blocks, checked at release
	u64 arena_default_chunks = 0; // the cap at init: free memory less 6 GiB
	// device address of the file-scope scratch array of dk_runtime.hip
	// (block partials of dk_apply_noalloc), taken once at init with the
	// device current; the library binds one device per process
	double *noalloc_scratch = nullptr;
};

enum { DK_PINNED = 1 << 16 };

extern dk_ctx g_ctx;

// Marks the start (end = 0) or end (end = 1) of a gate-kernel launch on the
// calling lane: events[62] / [63] when time_apply is set, and the next pair
// of the lane's timing ring while it has room.
int dk_time_mark(int end);

// The calling thread's lane, and its stream.
dk_lane *dk_cur();
#define DK_STREAM (dk_cur()->stream)
int dk_lane_setup(dk_lane *l);

// Waits for the lane that last wrote `s` when that is not the calling
// thread's lane.  Work on one state from several lanes is ordered this way;
// the caller orders reads it issues concurrently from different lanes.
int dk_settle(const dk_state *s);
#define DK_SETTLE(s)                                                        \
	do {                                                                \
		int rc_ = dk_settle(s);                                     \
		if (rc_)                                                    \
			return rc_;                                         \
	} while (0)

int dk_ensure_init();
int dk_alloc_state(unsigned nq, int dtype, dk_state **out);
void dk_release_state(dk_state *s);
// Device memory (dk_memory.cpp).  dk_device_free releases in stream order
// on the calling lane and takes the size given to dk_device_alloc.  Blocks of
// DK_ARENA_MIN bytes and more come from the chunk arena, rounded up to
// dk_block_bytes(bytes) (a power of two up to 1 GiB, whole 2 GiB chunks
// beyond).
enum : u64 { DK_ARENA_MIN = 64ull << 20 };
int dk_device_alloc(u64 bytes, void **out);
void dk_device_free(void *p, u64 bytes);
u64 dk_block_bytes(u64 bytes);
bool dk_closed();

// ---- kernel launchers (dk_kernels.hip, dk_dense.hip) --------------------------

// Grid for `nvb` virtual blocks of `threads` threads: a launch holds fewer
// than 2^32 threads, so beyond 2^31 the kernels loop over virtual blocks.
static inline unsigned dk_grid_cap(u64 nvb, unsigned threads)
{
	const u64 cap = (1ull << 31) / threads;
	return (unsigned)(nvb < 1 ? 1 : nvb < cap ? nvb : cap);
}

// Gate application; block partial sums of |out|^2 finish into out's norm on
// the device.  The partials go to the calling lane's growing buffer, or, when
// `scratch` is given, into those `scratch_cap` doubles (nothing is allocated;
// DK_E_ARG when the launch needs more).
int dk_launch_apply(const dk_state *in, dk_state *out, const dk_gate *g,
		    const unsigned *targets, u64 cm, u64 acm,
		    double *scratch = nullptr, u64 scratch_cap = 0);

// Records that the calling lane's queued work reads gate g's device copy.
int dk_gate_mark_use(const dk_gate *g);

// A one-qubit gate of a circuit run whole in one workgroup's LDS.
struct dk_small_gate {
	double m[8];      // 2x2 row-major, interleaved re/im
	u64 cm, am;       // control / anticontrol index bits
	unsigned t;       // target qubit
	unsigned pad;
};

// Most qubits of a state one workgroup holds in LDS (64 KiB of amplitudes).
unsigned dk_small_qubits(int dtype);

// `ngates` one-qubit gates (device array) on `in`, into `out` (may be `in`),
// in one launch of one workgroup: the state is read through its norm, every
// gate applied in order in LDS, the result written with its norm.  For
// states of at most dk_small_qubits(dtype) qubits.
int dk_launch_small_run(const dk_state *in, dk_state *out,
			const dk_small_gate *gates, unsigned ngates);

// Dense k-qubit gate (4 <= k <= 6, no controls) on MFMA, for states of at
// least dk_dense_tile_bits(k) qubits (the index bits one block covers).
unsigned dk_dense_tile_bits(unsigned k, int dtype);
int dk_launch_apply_dense(const dk_state *in, dk_state *out, const dk_gate *g,
			  const unsigned *targets, double **parts, u64 *nparts,
			  double *scratch = nullptr, u64 scratch_cap = 0);

// Shared partials buffer with room for `count` block partials plus the
// chunked finishing pass; deterministic finishing sum of such partials.
int dk_partials_reserve(u64 count, double **out);
// Doubles a partials buffer needs for `count` partials and their finishing.
u64 dk_partials_need(u64 count);
// `count` partials in caller-given scratch of `cap` doubles (no allocation).
int dk_partials_in(double *scratch, u64 cap, u64 count, double **out);
int dk_finish_partials(double *parts, u64 count, double *dst, int take_sqrt);

// sum_{i: bit q of i = 1} |raw_i / norm|^2 -> *dst (device scalar).
int dk_launch_prob(const dk_state *s, unsigned q, double *dst);
int dk_launch_prob_all(const dk_state *s, double *dst /* device, nq */);
int dk_launch_prob_partials(const dk_state *s, unsigned q, int *nparts_out);

// Scaled copy (and type conversion): dst = src * (1 / src.norm), norm 1.
int dk_launch_normalised_copy(const dk_state *src, dk_state *dst);
int dk_launch_values_copy(const dk_state *src, dk_state *dst);

// Measurement step: decide outcome from device partials, write the norm of
// the collapsed state; collapse (with fused partials for the next qubit).
int dk_launch_decide(const double *parts, int nparts, double roll,
		     double *dec_slot, dk_state *next);
int dk_launch_collapse(const dk_state *s, unsigned q, const double *dec_slot,
		       dk_state *o, int next_q, int *nparts_out);

int dk_launch_join(const dk_state *a, const dk_state *b, dk_state *o);
// join into o and the marginal probabilities of o's qubits into dst (device,
// o->nq doubles), in one pass where 9 <= o->nq <= 40
int dk_launch_join_prob_all(const dk_state *a, const dk_state *b, dk_state *o,
			    double *dst);

int dk_launch_first_nonzero(const dk_state *s, u64 start, u64 *dst_dev);

// Reduced densities; amplitudes whose index fails (idx & tmask) == tval
// (tmask on traced bits only) are left out of the traced sum.
int dk_launch_density(const dk_state *s, u64 keep_mask, double2 *rho_dev,
		      u64 tmask = 0, u64 tval = 0);
int dk_launch_density_elements(const dk_state *s, u64 keep_mask, size_t count,
			       const u64 *rows_dev, const u64 *cols_dev,
			       double2 *out_dev, u64 tmask = 0, u64 tval = 0);

int dk_launch_fill_zero_state(dk_state *s);
int dk_launch_set_norm(dk_state *s, double v);
int dk_launch_poison(void *p, u64 bytes);
int dk_launch_guard_fill(void *g, u64 bytes);
int dk_launch_guard_check(const void *g, u64 bytes, unsigned *bad);
int dk_launch_widen(const dk_state *s, u64 first, u64 count, double2 *dst);

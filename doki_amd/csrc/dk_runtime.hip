// C ABI of libdoki_hip.so: device context, state lifetime and the entry
// points declared in include/doki_hip.h.  Kernels live in dk_kernels.hip.
#include <algorithm>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "dk_internal.h"

dk_ctx g_ctx;
static thread_local char t_err[512];

// Block partials of dk_apply_noalloc: a fixed device array, so that entry
// point allocates nothing.  Host code never passes `scratch` itself to a
// kernel (on the host the name is only a stand-in); it uses the device
// address taken at init (g_ctx.noalloc_scratch).
enum : u64 { DK_NOALLOC_DOUBLES = DK_NOALLOC_SCRATCH_BYTES / sizeof(double) };
static __device__ __align__(256) double scratch[DK_NOALLOC_DOUBLES];
static bool g_closed = false;

void dk_set_error(const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(t_err, sizeof(t_err), fmt, ap);
	va_end(ap);
}

extern "C" const char *dk_last_error(void)
{
	return t_err;
}

// Device slots inside dk_cur()->dec
enum {
	DEC_SLOTS = 64,            // measurement decisions: 2 doubles each
	DEC_SCALAR = 2 * DEC_SLOTS,// one device scalar result
	DEC_U64 = DEC_SCALAR + 1,  // one u64 result
	DEC_PROBS = DEC_SCALAR + 2,// up to 64 probabilities
	DEC_TOTAL = DEC_PROBS + 64
};

static thread_local dk_lane *t_lane = nullptr;

dk_lane *dk_cur()
{
	return t_lane ? t_lane : &g_ctx.lane0;
}

// Stream, scratch, pinned staging and timing events of a lane; on failure
// nothing stays allocated.
static int lane_setup_once(dk_lane *l)
{
	hipStream_t st = nullptr;
	hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
	// DK_PARTIALS doubles: a pool-sized block
	if (e == hipSuccess)
		e = hipMallocAsync((void **)&l->partials, DK_PARTIALS * sizeof(double), st);
	if (e == hipSuccess)
		e = hipMalloc((void **)&l->dec, 2 * 64 * sizeof(double) + 2048);
	if (e == hipSuccess)
		e = hipHostMalloc((void **)&l->host_pinned, DK_PINNED * sizeof(double),
				  hipHostMallocDefault);
	for (auto &ev : l->events)
		if (e == hipSuccess)
			e = hipEventCreate(&ev);
	if (e == hipSuccess)
		e = hipStreamSynchronize(st);
	if (e == hipSuccess) {
		l->stream = st;
		l->partials_cap = DK_PARTIALS;
		return DK_OK;
	}
	dk_set_error("lane setup failed: %s", hipGetErrorString(e));
	(void)hipGetLastError();
	for (auto &ev : l->events)
		if (ev) {
			(void)hipEventDestroy(ev);
			ev = nullptr;
		}
	if (l->host_pinned)
		(void)hipHostFree(l->host_pinned);
	if (l->dec)
		(void)hipFree(l->dec);
	if (l->partials && st)
		(void)hipFreeAsync(l->partials, st);
	if (st) {
		(void)hipStreamSynchronize(st);
		(void)hipStreamDestroy(st);
	}
	l->host_pinned = nullptr;
	l->dec = l->partials = nullptr;
	return DK_E_DEVICE;
}

// The arena may hold nearly all device memory in cached blocks: when a
// lane's scratch does not fit, they go back to the driver and setup runs
// once more.
int dk_lane_setup(dk_lane *l)
{
	if (lane_setup_once(l) == DK_OK)
		return DK_OK;
	(void)dk_memory_trim();
	return lane_setup_once(l);
}

extern "C" int dk_init(int device)
{
	std::lock_guard<std::mutex> lk(g_ctx.mu);
	if (g_ctx.ready)
		return DK_OK;
	int count = 0;
	hipError_t e = hipGetDeviceCount(&count);
	if (e != hipSuccess || count <= 0) {
		dk_set_error("no HIP device available: %s",
			     e != hipSuccess ? hipGetErrorString(e) : "0 devices");
		return DK_E_DEVICE;
	}
	if (device < 0 || device >= count) {
		dk_set_error("device %d out of range (%d devices)", device, count);
		return DK_E_ARG;
	}
	DK_HIP(hipSetDevice(device));
	hipDeviceProp_t prop;
	DK_HIP(hipGetDeviceProperties(&prop, device));
	if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
		dk_set_error("libdoki_hip is built for gfx950 only; device %d is %s",
			     device, prop.gcnArchName);
		return DK_E_DEVICE;
	}
	g_ctx.device = device;
	g_ctx.num_cus = prop.multiProcessorCount;
	// the no-alloc scratch's address on this device (the device is current)
	double *sp = nullptr;
	DK_HIP(hipGetSymbolAddress((void **)&sp, HIP_SYMBOL(scratch)));
	g_ctx.noalloc_scratch = sp;
	// (nothing to hand back yet: the arena is empty, and trimming it would
	// take this lock again)
	int rc = lane_setup_once(&g_ctx.lane0);
	if (rc)
		return rc;
	hipMemPool_t pool;
	DK_HIP(hipDeviceGetDefaultMemPool(&pool, device));
	uint64_t thr = UINT64_MAX;
	DK_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr));
	// A block freed on one lane's stream comes back to that stream, or to
	// one that has waited for an event recorded after the free: no reuse
	// on the runtime's own view of completion, and no waits it would add
	// by itself between the lanes' streams.
	int no = 0;
	DK_HIP(hipMemPoolSetAttribute(pool, hipMemPoolReuseAllowOpportunistic, &no));
	DK_HIP(hipMemPoolSetAttribute(pool, hipMemPoolReuseAllowInternalDependencies, &no));
	size_t fr = 0, tot = 0;
	DK_HIP(hipMemGetInfo(&fr, &tot));
	g_ctx.budget = (u64)((double)fr * 0.92);
	g_ctx.used = 0;
	// the state arena leaves 6 GiB of the free memory to the pool
	const u64 headroom = 6ull << 30, chunk = 2ull << 30;
	g_ctx.arena_default_chunks = fr > headroom + chunk ? (fr - headroom) / chunk : 1;
	g_ctx.arena_max_chunks = g_ctx.arena_default_chunks;
	g_ctx.ready = true;
	return DK_OK;
}

int dk_ensure_init()
{
	if (g_ctx.ready)
		return DK_OK;
	const char *dev = getenv("DOKI_HIP_DEVICE");
	return dk_init(dev ? atoi(dev) : 0);
}

#define DK_INIT()                                                           \
	do {                                                                \
		int rc_ = dk_ensure_init();                                 \
		if (rc_)                                                    \
			return rc_;                                         \
	} while (0)

extern "C" int dk_synchronize(void)
{
	DK_INIT();
	DK_HIP(hipStreamSynchronize(DK_STREAM));
	return DK_OK;
}

// Marks the library closed: later frees only drop host bookkeeping.  The
// extension calls it when the interpreter finalises, so that any state freed
// after that point never reaches a HIP runtime that may be tearing down.
bool dk_closed()
{
	return g_closed;
}

extern "C" void dk_shutdown(void)
{
	if (g_ctx.ready)
		(void)hipDeviceSynchronize();
	g_closed = true;
}

// ---- lanes ------------------------------------------------------------------

// Freed lanes are kept for reuse rather than destroyed: states remember the
// lane that last wrote them, and that pointer must stay valid.
static std::vector<dk_lane *> g_spare_lanes;

extern "C" int dk_lane_new(dk_lane **out)
{
	DK_INIT();
	if (!out) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	{
		std::lock_guard<std::mutex> lk(g_ctx.mu);
		if (!g_spare_lanes.empty()) {
			*out = g_spare_lanes.back();
			g_spare_lanes.pop_back();
			return DK_OK;
		}
	}
	dk_lane *l = new dk_lane;
	int rc = dk_lane_setup(l);
	if (rc) {
		delete l;
		return rc;
	}
	*out = l;
	return DK_OK;
}

extern "C" int dk_lane_bind(dk_lane *l)
{
	DK_INIT();
	t_lane = l;
	return DK_OK;
}

extern "C" void dk_lane_free(dk_lane *l)
{
	if (!l || l == &g_ctx.lane0)
		return;
	if (t_lane == l)
		t_lane = nullptr;
	if (!g_closed)
		(void)hipStreamSynchronize(l->stream);
	std::lock_guard<std::mutex> lk(g_ctx.mu);
	g_spare_lanes.push_back(l);
}

// A state last written on another lane: wait for that lane's work first.
int dk_settle(const dk_state *s)
{
	if (s && s->lane && s->lane != dk_cur())
		DK_HIP(hipStreamSynchronize(s->lane->stream));
	return DK_OK;
}

extern "C" int dk_memory_budget(uint64_t *budget, uint64_t *used)
{
	DK_INIT();
	std::lock_guard<std::mutex> lk(g_ctx.mu);
	if (budget)
		*budget = g_ctx.budget;
	if (used)
		*used = g_ctx.used;
	return DK_OK;
}

extern "C" int dk_set_memory_budget(uint64_t budget)
{
	DK_INIT();
	std::lock_guard<std::mutex> lk(g_ctx.mu);
	g_ctx.budget = budget;
	return DK_OK;
}

// ---- states -----------------------------------------------------------------

int dk_alloc_state(unsigned nq, int dtype, dk_state **out)
{
	*out = nullptr;
	if (nq < 1 || nq > DK_MAX_QUBITS) {
		dk_set_error("Number of qubits exceeds maximum");
		return DK_E_QUBITS;
	}
	if (dtype != DK_C128 && dtype != DK_C64) {
		dk_set_error("unknown amplitude type %d", dtype);
		return DK_E_ARG;
	}
	DK_INIT();
	if (nq > 44) {
		dk_set_error("Failed to allocate state chunk: 2^%u amplitudes "
			     "exceed device memory", nq);
		return DK_E_BUDGET;
	}
	const u64 len = 1ull << nq;
	const u64 abytes = len * dk_amp_size(dtype);
	const u64 bytes = DK_SCAL_BYTES + abytes;
	{
		std::lock_guard<std::mutex> lk(g_ctx.mu);
		if (g_ctx.used + bytes > g_ctx.budget) {
			dk_set_error("Failed to allocate state chunk: %llu bytes "
				     "requested, %llu of %llu budget bytes in use",
				     (unsigned long long)bytes,
				     (unsigned long long)g_ctx.used,
				     (unsigned long long)g_ctx.budget);
			return DK_E_BUDGET;
		}
		g_ctx.used += bytes;
	}
	// amplitudes first, so they start on the allocation's own alignment;
	// arena blocks (power-of-two sizes) hold only the amplitudes, and the
	// scalars take a small block of their own
	const bool split = abytes >= DK_ARENA_MIN;
	char *base = nullptr, *sbase = nullptr;
	int rc = dk_device_alloc(split ? abytes : bytes, (void **)&base);
	if (!rc && split) {
		rc = dk_device_alloc(DK_SCAL_BYTES, (void **)&sbase);
		if (rc)
			dk_device_free(base, abytes);
	}
	if (rc) {
		std::lock_guard<std::mutex> lk(g_ctx.mu);
		g_ctx.used -= bytes;
		return rc;
	}
	dk_state *s = new dk_state;
	s->nq = nq;
	s->dtype = dtype;
	s->lane = dk_cur();
	s->len = len;
	s->base = base;
	s->sbase = sbase;
	s->amp = base;
	s->scal = (double *)(split ? sbase : base + abytes);
	s->bytes = bytes;
	s->support = ~0ull;
	*out = s;
	return DK_OK;
}

void dk_release_state(dk_state *s)
{
	if (!s)
		return;
	if (!g_closed && g_ctx.ready) {
		(void)dk_settle(s);
		if (s->sbase) {
			dk_device_free(s->base, s->bytes - DK_SCAL_BYTES);
			dk_device_free(s->sbase, DK_SCAL_BYTES);
		} else {
			dk_device_free(s->base, s->bytes);
		}
		std::lock_guard<std::mutex> lk(g_ctx.mu);
		g_ctx.used -= s->bytes;
	}
	delete s;
}

extern "C" void dk_state_free(dk_state *s)
{
	dk_release_state(s);
}

extern "C" int dk_state_new_typed(unsigned num_qubits, int zero_state,
				  int dtype, dk_state **out)
{
	dk_state *s;
	int rc = dk_alloc_state(num_qubits, dtype, &s);
	if (rc)
		return rc;
	rc = zero_state ? dk_launch_fill_zero_state(s) : dk_launch_set_norm(s, 1.0);
	if (rc) {
		dk_release_state(s);
		return rc;
	}
	if (zero_state)
		s->support = 0;
	*out = s;
	return DK_OK;
}

extern "C" int dk_state_dtype(const dk_state *s)
{
	return s ? s->dtype : -1;
}

extern "C" int dk_state_convert(const dk_state *src, int dtype, dk_state **out)
{
	if (!src || !out) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	DK_SETTLE(src);
	dk_state *s;
	int rc = dk_alloc_state(src->nq, dtype, &s);
	if (rc)
		return rc;
	rc = dk_launch_normalised_copy(src, s);
	if (rc) {
		dk_release_state(s);
		return rc;
	}
	*out = s;
	return DK_OK;
}

extern "C" int dk_state_new(unsigned num_qubits, int zero_state, dk_state **out)
{
	dk_state *s;
	int rc = dk_alloc_state(num_qubits, DK_C128, &s);
	if (rc)
		return rc;
	rc = zero_state ? dk_launch_fill_zero_state(s) : dk_launch_set_norm(s, 1.0);
	if (rc) {
		dk_release_state(s);
		return rc;
	}
	if (zero_state)
		s->support = 0;
	*out = s;
	return DK_OK;
}

extern "C" int dk_state_from_host(unsigned num_qubits, const double *values,
				  dk_state **out)
{
	if (!values) {
		dk_set_error("values is NULL");
		return DK_E_ARG;
	}
	dk_state *s;
	int rc = dk_alloc_state(num_qubits, DK_C128, &s);
	if (rc)
		return rc;
	hipError_t e = hipMemcpyAsync(s->amp, values, s->len * sizeof(double2),
				      hipMemcpyHostToDevice, DK_STREAM);
	if (e == hipSuccess)
		e = hipStreamSynchronize(DK_STREAM);
	if (e != hipSuccess) {
		dk_set_error("copy to device failed: %s", hipGetErrorString(e));
		dk_release_state(s);
		return DK_E_DEVICE;
	}
	rc = dk_launch_set_norm(s, 1.0);
	if (rc) {
		dk_release_state(s);
		return rc;
	}
	*out = s;
	return DK_OK;
}

extern "C" int dk_state_from_device(unsigned num_qubits, const void *values,
				    dk_state **out)
{
	DK_INIT();
	if (!values || !out) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	dk_state *s;
	int rc = dk_alloc_state(num_qubits, DK_C128, &s);
	if (rc)
		return rc;
	// the values may come from work on any stream of the device (another
	// library's): wait for all of it, then copy on this lane
	hipError_t e = hipDeviceSynchronize();
	if (e == hipSuccess)
		e = hipMemcpyAsync(s->amp, values, s->len * sizeof(double2),
				   hipMemcpyDeviceToDevice, DK_STREAM);
	if (e != hipSuccess) {
		dk_set_error("device copy failed: %s", hipGetErrorString(e));
		dk_release_state(s);
		return DK_E_DEVICE;
	}
	rc = dk_launch_set_norm(s, 1.0);
	if (rc) {
		dk_release_state(s);
		return rc;
	}
	*out = s;
	return DK_OK;
}

extern "C" int dk_state_values(const dk_state *src, dk_state **out)
{
	if (!src || !out) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	DK_SETTLE(src);
	dk_state *s;
	int rc = dk_alloc_state(src->nq, DK_C128, &s);
	if (rc)
		return rc;
	rc = dk_launch_values_copy(src, s);
	if (rc) {
		dk_release_state(s);
		return rc;
	}
	*out = s;
	return DK_OK;
}

extern "C" int dk_state_device_address(const dk_state *s, void **amp, int *device)
{
	if (!s || !amp) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	DK_SETTLE(s);
	// the work queued on the state completes before its address is handed
	// to a caller that reads it on streams of its own
	DK_HIP(hipStreamSynchronize(DK_STREAM));
	*amp = s->amp;
	if (device)
		*device = g_ctx.device;
	return DK_OK;
}

extern "C" int dk_state_clone(const dk_state *src, dk_state **out)
{
	if (!src) {
		dk_set_error("NULL state");
		return DK_E_ARG;
	}
	DK_SETTLE(src);
	dk_state *s;
	int rc = dk_alloc_state(src->nq, src->dtype, &s);
	if (rc)
		return rc;
	rc = dk_launch_normalised_copy(src, s);
	if (rc) {
		dk_release_state(s);
		return rc;
	}
	*out = s;
	return DK_OK;
}

extern "C" unsigned dk_state_num_qubits(const dk_state *s)
{
	return s ? s->nq : 0;
}

extern "C" uint64_t dk_state_mem_size(const dk_state *s)
{
	return s ? 64ull + 16ull * s->len : 0;
}

extern "C" uint64_t dk_state_device_bytes(const dk_state *s)
{
	return s ? s->bytes : 0;
}

// Copies `count` doubles from device to host through the pinned buffer.
static int copy_to_host(double *dst, const void *src, size_t count)
{
	const char *p = (const char *)src;
	while (count > 0) {
		const size_t n = std::min<size_t>(count, DK_PINNED);
		DK_HIP(hipMemcpyAsync(dk_cur()->host_pinned, p, n * sizeof(double),
				      hipMemcpyDeviceToHost, DK_STREAM));
		DK_HIP(hipStreamSynchronize(DK_STREAM));
		memcpy(dst, dk_cur()->host_pinned, n * sizeof(double));
		dst += n;
		p += n * sizeof(double);
		count -= n;
	}
	return DK_OK;
}

// Raw amplitudes [first, first+count) as interleaved doubles (complex64
// states are widened on the device first).
static int read_amplitudes(const dk_state *s, u64 first, u64 count,
			   double *values)
{
	if (s->dtype == DK_C128)
		return copy_to_host(values, s->a128() + first, 2 * count);
	const u64 chunk = DK_PINNED / 2;
	double2 *tmp = nullptr;
	const u64 tmp_bytes = std::min(count, chunk) * sizeof(double2);
	DK_TRY(dk_device_alloc(tmp_bytes, (void **)&tmp));
	int rc = DK_OK;
	for (u64 done = 0; !rc && done < count; done += chunk) {
		const u64 n = std::min(chunk, count - done);
		rc = dk_launch_widen(s, first + done, n, tmp);
		if (!rc)
			rc = copy_to_host(values + 2 * done, tmp, 2 * n);
	}
	dk_device_free(tmp, tmp_bytes);
	return rc;
}

extern "C" int dk_state_norm(const dk_state *s, double *norm)
{
	DK_INIT();
	if (!s || !norm) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	DK_SETTLE(s);
	return copy_to_host(norm, s->scal + DK_SCAL_NORM, 1);
}

// ---- gates -------------------------------------------------------------------

extern "C" int dk_gate_new(unsigned num_qubits, const double *matrix,
			   dk_gate **out)
{
	*out = nullptr;
	if (num_qubits < 1 || num_qubits > DK_MAX_GATE_QUBITS || !matrix) {
		dk_set_error("gates act on 1..%d qubits", DK_MAX_GATE_QUBITS);
		return DK_E_ARG;
	}
	DK_INIT();
	dk_gate *g = new dk_gate;
	g->k = num_qubits;
	g->dim = 1u << num_qubits;
	g->dmat = nullptr;
	g->dmat32 = nullptr;
	g->lane = dk_cur();
	const size_t n = (size_t)g->dim * g->dim;
	memcpy(g->hmat, matrix, n * sizeof(double2));
	for (size_t i = 0; i < n; ++i)
		if (!std::isfinite(g->hmat[i].x) || !std::isfinite(g->hmat[i].y))
			g->finite = false;
	if (num_qubits > 1) {
		// FP64 and FP32 copies in one stream-ordered allocation of the
		// calling lane; the host sources live in the gate itself
		for (size_t i = 0; i < n; ++i)
			g->hmat32[i] = make_float2((float)g->hmat[i].x, (float)g->hmat[i].y);
		char *buf = nullptr;
		const size_t b64 = n * sizeof(double2), b32 = n * sizeof(float2);
		hipStream_t st = DK_STREAM;
		int rc = dk_device_alloc(b64 + b32, (void **)&buf);
		if (rc) {
			delete g;
			return rc;
		}
		hipError_t e = hipMemcpyAsync(buf, g->hmat, b64, hipMemcpyHostToDevice, st);
		if (e == hipSuccess)
			e = hipMemcpyAsync(buf + b64, g->hmat32, b32, hipMemcpyHostToDevice,
					   st);
		if (e != hipSuccess) {
			dk_set_error("gate upload failed: %s", hipGetErrorString(e));
			dk_device_free(buf, b64 + b32);
			(void)hipStreamSynchronize(st);
			delete g;
			return DK_E_DEVICE;
		}
		g->dmat = (double2 *)buf;
		g->dmat32 = (float2 *)(buf + b64);
		if (hipEventCreateWithFlags(&g->uploaded, hipEventDisableTiming) !=
			    hipSuccess ||
		    hipEventRecord(g->uploaded, st) != hipSuccess) {
			dk_set_error("gate upload event failed");
			(void)hipStreamSynchronize(st);
			dk_device_free(buf, b64 + b32);
			delete g;
			return DK_E_DEVICE;
		}
	}
	*out = g;
	return DK_OK;
}

extern "C" void dk_gate_free(dk_gate *g)
{
	if (!g)
		return;
	if (g->dmat && !g_closed) {
		// the upload and every lane's last launch reading the gate finish
		// first; the release is then ordered on the calling lane
		(void)hipEventSynchronize(g->uploaded);
		for (auto &u : g->uses)
			(void)hipEventSynchronize(u.second);
		const size_t n = (size_t)g->dim * g->dim;
		dk_device_free(g->dmat, n * (sizeof(double2) + sizeof(float2)));
	}
	if (!g_closed) {
		if (g->uploaded)
			(void)hipEventDestroy(g->uploaded);
		for (auto &u : g->uses)
			(void)hipEventDestroy(u.second);
	}
	delete g;
}

// ---- operations ----------------------------------------------------------------

// Argument checks shared by dk_apply and dk_apply_noalloc.
static int check_apply_args(const dk_state *in, const dk_gate *gate,
			    const unsigned *targets, u64 ctrl_mask,
			    u64 actrl_mask)
{
	const unsigned n = in->nq;
	const u64 valid = n >= 64 ? ~0ull : ((1ull << n) - 1);
	u64 tmask = 0;
	for (unsigned b = 0; b < gate->k; ++b) {
		if (targets[b] >= n) {
			dk_set_error("Target qubit out of range");
			return DK_E_ARG;
		}
		if ((tmask >> targets[b]) & 1) {
			dk_set_error("Repeated target qubit");
			return DK_E_ARG;
		}
		tmask |= 1ull << targets[b];
	}
	if ((ctrl_mask & ~valid) || (actrl_mask & ~valid)) {
		dk_set_error("Control qubit out of range");
		return DK_E_ARG;
	}
	if (ctrl_mask & actrl_mask) {
		dk_set_error("A control cannot also be an anticontrol");
		return DK_E_ARG;
	}
	if (tmask & (ctrl_mask | actrl_mask)) {
		dk_set_error("A target cannot also be a control or an anticontrol");
		return DK_E_ARG;
	}
	return DK_OK;
}

extern "C" int dk_apply(dk_state *in, const dk_gate *gate,
			const unsigned *targets, uint64_t ctrl_mask,
			uint64_t actrl_mask, int in_place, dk_state **out)
{
	DK_INIT();
	if (!in || !gate || !targets || !out) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	DK_TRY(check_apply_args(in, gate, targets, ctrl_mask, actrl_mask));
	const unsigned n = in->nq;
	DK_SETTLE(in);
	dk_state *o = in;
	if (!in_place) {
		int rc = dk_alloc_state(n, in->dtype, &o);
		if (rc)
			return rc;
	}
	int rc = dk_launch_apply(in, o, gate, targets, ctrl_mask, actrl_mask);
	if (rc) {
		if (o != in)
			dk_release_state(o);
		return rc;
	}
	o->lane = dk_cur();
	*out = o;
	return DK_OK;
}

// dk_apply into a state the caller owns, with the block partials of the
// norm in the fixed scratch array: no device memory is allocated or freed.
// Calls on one device share the array, so calls that overlap on the device
// (from different lanes) race; every partial is written before it is read,
// so nothing in the array needs clearing between calls.
extern "C" int dk_apply_noalloc(dk_state *in, const dk_gate *gate,
				const unsigned *targets, uint64_t ctrl_mask,
				uint64_t actrl_mask, dk_state *out)
{
	if (!in || !gate || !targets || !out) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	if (!g_ctx.ready || !g_ctx.noalloc_scratch) {
		dk_set_error("the library is not initialised (dk_init)");
		return DK_E_DEVICE;
	}
	if (in->nq > DK_NOALLOC_MAX_QUBITS) {
		dk_set_error("dk_apply_noalloc: a state of %u qubits exceeds the "
			     "fixed scratch (%u qubits at most)", in->nq,
			     (unsigned)DK_NOALLOC_MAX_QUBITS);
		return DK_E_ARG;
	}
	if (out != in && (out->nq != in->nq || out->dtype != in->dtype)) {
		dk_set_error("dk_apply_noalloc: the output state must have the "
			     "input's qubit count and amplitude type");
		return DK_E_ARG;
	}
	DK_TRY(check_apply_args(in, gate, targets, ctrl_mask, actrl_mask));
	DK_SETTLE(in);
	DK_SETTLE(out);
	DK_TRY(dk_launch_apply(in, out, gate, targets, ctrl_mask, actrl_mask,
			       g_ctx.noalloc_scratch, DK_NOALLOC_DOUBLES));
	out->lane = dk_cur();
	return DK_OK;
}

extern "C" int dk_prob(const dk_state *s, unsigned q, double *p)
{
	DK_INIT();
	if (!s || !p) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	if (q >= s->nq) {
		dk_set_error("Qubit id out of range");
		return DK_E_ARG;
	}
	DK_SETTLE(s);
	double *dst = dk_cur()->dec + DEC_SCALAR;
	int rc = dk_launch_prob(s, q, dst);
	if (rc)
		return rc;
	return copy_to_host(p, dst, 1);
}

int dk_launch_compare(const dk_state *a, const dk_state *b, double *dst);

extern "C" int dk_state_compare(const dk_state *a, const dk_state *b, double *inner,
				double *max_abs)
{
	DK_INIT();
	if (!a || !b) {
		dk_set_error("NULL state");
		return DK_E_ARG;
	}
	if (a->len != b->len) {
		dk_set_error("states of different sizes");
		return DK_E_ARG;
	}
	DK_SETTLE(a);
	DK_SETTLE(b);
	double *dst = dk_cur()->dec + DEC_PROBS;
	DK_TRY(dk_launch_compare(a, b, dst));
	double r[3];
	DK_TRY(copy_to_host(r, dst, 3));
	if (inner) {
		inner[0] = r[0];
		inner[1] = r[1];
	}
	if (max_abs)
		*max_abs = r[2];
	return DK_OK;
}

extern "C" int dk_prob_all(const dk_state *s, double *p)
{
	DK_INIT();
	if (!s || !p) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	DK_SETTLE(s);
	double *dst = dk_cur()->dec + DEC_PROBS;
	int rc = dk_launch_prob_all(s, dst);
	if (rc)
		return rc;
	return copy_to_host(p, dst, s->nq);
}

int dk_launch_prob_partials(const dk_state *s, unsigned q, int *nparts_out);

extern "C" int dk_measure(const dk_state *s, uint64_t mask, const double *rolls,
			  size_t num_rolls, int8_t *results, double *p1_out,
			  dk_state **out)
{
	DK_INIT();
	if (!s || !results || !out) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	*out = nullptr;
	DK_SETTLE(s);
	const unsigned n = s->nq;
	// qubits to measure, most significant first (bits >= n are ignored)
	std::vector<unsigned> qs;
	for (unsigned i = 0; i < n; ++i) {
		const unsigned q = n - 1 - i;
		results[i] = -1;
		if (p1_out)
			p1_out[i] = 0.0;
		if ((mask >> q) & 1)
			qs.push_back(q);
	}
	for (size_t k = 0; k < qs.size(); ++k) {
		if (k >= num_rolls || !rolls || rolls[k] < 0.0 || rolls[k] >= 1.0) {
			dk_set_error("roll not in interval [0, 1)!");
			return DK_E_ROLL;
		}
	}
	if (qs.empty())
		return dk_state_clone(s, out);
	if (qs.size() > DEC_SLOTS) {
		dk_set_error("too many measured qubits");
		return DK_E_ARG;
	}
	const dk_state *cur = s;
	dk_state *owned = nullptr; // cur when it is a temporary
	dk_state *next = nullptr;
	int rc = DK_OK;
	int nparts = 0;
	// P1 of the first measured qubit
	rc = dk_launch_prob_partials(cur, qs[0], &nparts);
	if (!rc && cur->nq > 1)
		rc = dk_alloc_state(cur->nq - 1, s->dtype, &next);
	if (!rc)
		rc = dk_launch_decide(dk_cur()->partials, nparts, rolls[0],
				      dk_cur()->dec + 0, next);
	for (size_t k = 0; !rc && k < qs.size(); ++k) {
		if (cur->nq == 1)
			break; // last qubit: nothing left to collapse into
		const int next_q = k + 1 < qs.size() ? (int)qs[k + 1] : -1;
		rc = dk_launch_collapse(cur, qs[k], dk_cur()->dec + 2 * k, next,
					next_q, &nparts);
		if (rc)
			break;
		if (owned)
			dk_release_state(owned);
		owned = next;
		cur = next;
		next = nullptr;
		if (next_q >= 0) {
			if (cur->nq > 1)
				rc = dk_alloc_state(cur->nq - 1, s->dtype, &next);
			if (!rc)
				rc = dk_launch_decide(dk_cur()->partials, nparts,
						      rolls[k + 1],
						      dk_cur()->dec + 2 * (k + 1), next);
		}
	}
	if (rc) {
		if (next)
			dk_release_state(next);
		if (owned)
			dk_release_state(owned);
		return rc;
	}
	std::vector<double> dec(2 * qs.size());
	rc = copy_to_host(dec.data(), dk_cur()->dec, dec.size());
	if (rc) {
		if (owned)
			dk_release_state(owned);
		return rc;
	}
	for (size_t k = 0; k < qs.size(); ++k) {
		const unsigned slot = n - 1 - qs[k];
		results[slot] = dec[2 * k] != 0.0 ? 1 : 0;
		if (p1_out)
			p1_out[slot] = dec[2 * k + 1];
		const double p = dec[2 * k] != 0.0 ? dec[2 * k + 1]
						   : 1.0 - dec[2 * k + 1];
		const bool leaves_qubits = (n - 1 - k) > 0;
		if (leaves_qubits && sqrt(p) == 0.0) {
			if (owned)
				dk_release_state(owned);
			dk_set_error("New normalization constant is 0. Please report "
				     "this error with the steps to reproduce it.");
			return DK_E_ZERO_NORM;
		}
	}
	if (n - qs.size() > 0) {
		*out = owned;
	} else {
		if (owned)
			dk_release_state(owned);
		*out = nullptr;
	}
	return DK_OK;
}

extern "C" int dk_join(const dk_state *most, const dk_state *least,
		       dk_state **out)
{
	DK_INIT();
	if (!most || !least || !out) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	if (most->dtype != least->dtype) {
		dk_set_error("joined states must share an amplitude type");
		return DK_E_ARG;
	}
	DK_SETTLE(most);
	DK_SETTLE(least);
	dk_state *o;
	int rc = dk_alloc_state(most->nq + least->nq, most->dtype, &o);
	if (rc)
		return rc;
	rc = dk_launch_join(most, least, o);
	if (rc) {
		dk_release_state(o);
		return rc;
	}
	*out = o;
	return DK_OK;
}

extern "C" int dk_join_prob_all(const dk_state *most, const dk_state *least,
				dk_state **out, double *p)
{
	DK_INIT();
	if (!most || !least || !out || !p) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	if (most->dtype != least->dtype) {
		dk_set_error("joined states must share an amplitude type");
		return DK_E_ARG;
	}
	DK_SETTLE(most);
	DK_SETTLE(least);
	dk_state *o;
	int rc = dk_alloc_state(most->nq + least->nq, most->dtype, &o);
	if (rc)
		return rc;
	double *dst = dk_cur()->dec + DEC_PROBS;
	rc = dk_launch_join_prob_all(most, least, o, dst);
	if (!rc)
		rc = copy_to_host(p, dst, o->nq);
	if (rc) {
		dk_release_state(o);
		return rc;
	}
	*out = o;
	return DK_OK;
}

extern "C" int dk_read_raw(const dk_state *s, uint64_t first, uint64_t count,
			   double *values, double *norm)
{
	DK_INIT();
	if (!s) {
		dk_set_error("NULL state");
		return DK_E_ARG;
	}
	if (first > s->len || count > s->len - first) {
		dk_set_error("Out of bounds");
		return DK_E_ARG;
	}
	DK_SETTLE(s);
	int rc = DK_OK;
	if (count > 0 && values)
		rc = read_amplitudes(s, first, count, values);
	if (!rc && norm)
		rc = copy_to_host(norm, s->scal + DK_SCAL_NORM, 1);
	return rc;
}

extern "C" int dk_read(const dk_state *s, uint64_t first, uint64_t count,
		       double *values)
{
	double nrm = 1.0;
	int rc = dk_read_raw(s, first, count, values, &nrm);
	if (rc)
		return rc;
	const double inv = 1.0 / nrm;
	for (uint64_t i = 0; i < 2 * count; ++i)
		values[i] *= inv;
	return DK_OK;
}

extern "C" int dk_first_nonzero(const dk_state *s, uint64_t *index)
{
	DK_INIT();
	if (!s || !index) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	DK_SETTLE(s);
	// most states have a nonzero amplitude near the start: look there first
	const u64 head = std::min<u64>(s->len, 2048);
	std::vector<double> buf(2 * head);
	int rc = read_amplitudes(s, 0, head, buf.data());
	if (rc)
		return rc;
	for (u64 i = 0; i < head; ++i)
		if (buf[2 * i] != 0.0 || buf[2 * i + 1] != 0.0) {
			*index = i;
			return DK_OK;
		}
	u64 *dst = (u64 *)(dk_cur()->dec + DEC_U64);
	rc = dk_launch_first_nonzero(s, head, dst);
	if (rc)
		return rc;
	double tmp;
	rc = copy_to_host(&tmp, dst, 1);
	if (rc)
		return rc;
	memcpy(index, &tmp, sizeof(tmp));
	return DK_OK;
}

static int popcount64(u64 x)
{
	return __builtin_popcountll(x);
}

extern "C" int dk_reduced_density(const dk_state *s, uint64_t keep_mask,
				  double *rho)
{
	return dk_reduced_density_masked(s, keep_mask, 0, 0, rho);
}

extern "C" int dk_reduced_density_masked(const dk_state *s, uint64_t keep_mask,
					 uint64_t traced_mask,
					 uint64_t traced_value, double *rho)
{
	DK_INIT();
	if (!s || !rho) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	const u64 valid = (1ull << s->nq) - 1;
	if ((keep_mask & ~valid) || (traced_mask & ~valid) ||
	    (traced_mask & keep_mask)) {
		dk_set_error("kept or constrained qubit out of range");
		return DK_E_ARG;
	}
	DK_SETTLE(s);
	const int a = popcount64(keep_mask);
	if (a > 12) {
		dk_set_error("reduced density of %d qubits is too large to "
			     "materialise (limit 12)", a);
		return DK_E_ARG;
	}
	const u64 dim = 1ull << a;
	double2 *dev = nullptr;
	DK_TRY(dk_device_alloc(dim * dim * sizeof(double2), (void **)&dev));
	int rc = dk_launch_density(s, keep_mask, dev, traced_mask, traced_value);
	if (!rc)
		rc = copy_to_host(rho, dev, 2 * dim * dim);
	dk_device_free(dev, dim * dim * sizeof(double2));
	return rc;
}

extern "C" int dk_reduced_density_elements(const dk_state *s, uint64_t keep_mask,
					   size_t count, const uint64_t *rows,
					   const uint64_t *cols, double *out)
{
	return dk_reduced_density_elements_masked(s, keep_mask, 0, 0, count, rows,
						  cols, out);
}

extern "C" int dk_reduced_density_elements_masked(
	const dk_state *s, uint64_t keep_mask, uint64_t traced_mask,
	uint64_t traced_value, size_t count, const uint64_t *rows,
	const uint64_t *cols, double *out)
{
	DK_INIT();
	if (!s || (count && (!rows || !cols || !out))) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	const u64 valid = (1ull << s->nq) - 1;
	if ((keep_mask & ~valid) || (traced_mask & ~valid) ||
	    (traced_mask & keep_mask)) {
		dk_set_error("kept or constrained qubit out of range");
		return DK_E_ARG;
	}
	DK_SETTLE(s);
	const u64 dim = 1ull << popcount64(keep_mask);
	for (size_t e = 0; e < count; ++e)
		if (rows[e] >= dim || cols[e] >= dim) {
			dk_set_error("Out of bounds");
			return DK_E_ARG;
		}
	if (count == 0)
		return DK_OK;
	char *buf = nullptr;
	const size_t bytes = count * (2 * sizeof(u64) + sizeof(double2));
	DK_TRY(dk_device_alloc(bytes, (void **)&buf));
	u64 *drows = (u64 *)buf, *dcols = drows + count;
	double2 *dout = (double2 *)(dcols + count);
	int rc = DK_OK;
	hipError_t e = hipMemcpyAsync(drows, rows, count * sizeof(u64),
				      hipMemcpyHostToDevice, DK_STREAM);
	if (e == hipSuccess)
		e = hipMemcpyAsync(dcols, cols, count * sizeof(u64),
				   hipMemcpyHostToDevice, DK_STREAM);
	if (e == hipSuccess)
		e = hipStreamSynchronize(DK_STREAM);
	if (e != hipSuccess) {
		dk_set_error("copy failed: %s", hipGetErrorString(e));
		rc = DK_E_DEVICE;
	}
	if (!rc)
		rc = dk_launch_density_elements(s, keep_mask, count, drows, dcols,
						dout, traced_mask, traced_value);
	if (!rc)
		rc = copy_to_host(out, dout, 2 * count);
	dk_device_free(buf, bytes);
	return rc;
}

// ---- timing and tuning ------------------------------------------------------------

extern "C" int dk_event_record(int slot)
{
	DK_INIT();
	if (slot < 0 || slot >= 64) {
		dk_set_error("event slot out of range");
		return DK_E_ARG;
	}
	DK_HIP(hipEventRecord(dk_cur()->events[slot], DK_STREAM));
	return DK_OK;
}

extern "C" int dk_event_elapsed(int a, int b, float *ms)
{
	DK_INIT();
	if (a < 0 || a >= 64 || b < 0 || b >= 64 || !ms) {
		dk_set_error("event slot out of range");
		return DK_E_ARG;
	}
	DK_HIP(hipEventSynchronize(dk_cur()->events[b]));
	DK_HIP(hipEventElapsedTime(ms, dk_cur()->events[a], dk_cur()->events[b]));
	return DK_OK;
}

int dk_time_mark(int end)
{
	dk_lane *l = dk_cur();
	if (g_ctx.time_apply)
		DK_HIP(hipEventRecord(l->events[end ? 63 : 62], l->stream));
	if (l->ring_used < l->ring_cap) {
		DK_HIP(hipEventRecord(l->ring[2 * l->ring_used + (end ? 1 : 0)], l->stream));
		if (end)
			++l->ring_used;
	}
	return DK_OK;
}

// Sizes the calling lane's timing ring (pairs of events); 0 turns it off.
static int time_ring_setup(size_t pairs)
{
	dk_lane *l = dk_cur();
	if (pairs > (1u << 16)) {
		dk_set_error("time_ring holds at most 65536 launches");
		return DK_E_ARG;
	}
	if (l->ring) {
		DK_HIP(hipStreamSynchronize(l->stream));
		for (size_t i = 0; i < 2 * l->ring_cap; ++i)
			(void)hipEventDestroy(l->ring[i]);
		delete[] l->ring;
		l->ring = nullptr;
	}
	l->ring_cap = l->ring_used = 0;
	if (!pairs)
		return DK_OK;
	l->ring = new hipEvent_t[2 * pairs];
	for (size_t i = 0; i < 2 * pairs; ++i)
		DK_HIP(hipEventCreate(&l->ring[i]));
	l->ring_cap = pairs;
	return DK_OK;
}

extern "C" int dk_time_ring_read(float *ms, size_t capacity, size_t *count)
{
	DK_INIT();
	dk_lane *l = dk_cur();
	const size_t n = l->ring_used;
	if (count)
		*count = n;
	if (!ms)
		return DK_OK;
	if (capacity < n) {
		dk_set_error("room for %zu durations, the ring holds %zu", capacity, n);
		return DK_E_ARG;
	}
	if (n)
		DK_HIP(hipEventSynchronize(l->ring[2 * n - 1]));
	for (size_t i = 0; i < n; ++i)
		DK_HIP(hipEventElapsedTime(&ms[i], l->ring[2 * i], l->ring[2 * i + 1]));
	l->ring_used = 0;
	return DK_OK;
}

extern "C" int dk_tune(const char *key, int64_t value)
{
	DK_INIT();
	if (!key) {
		dk_set_error("NULL key");
		return DK_E_ARG;
	}
	if (!strcmp(key, "apply_blocks")) {
		g_ctx.apply_blocks = (int)value;
		return DK_OK;
	}
	if (!strcmp(key, "time_apply")) {
		g_ctx.time_apply = value != 0;
		return DK_OK;
	}
	if (!strcmp(key, "time_ring"))
		return time_ring_setup(value > 0 ? (size_t)value : 0);
	if (!strcmp(key, "dense_mfma")) {
		g_ctx.dense_mfma = value != 0;
		return DK_OK;
	}
	if (!strcmp(key, "apply1_group")) {
		if (value != 1 && value != 2 && value != 4) {
			dk_set_error("apply1_group is 1, 2 or 4");
			return DK_E_ARG;
		}
		g_ctx.apply1_group = (int)value;
		return DK_OK;
	}
	if (!strcmp(key, "dense_stream")) {
		if (value < 0 || value > 2) {
			dk_set_error("dense_stream is 0 (tiles), 1 (auto) or 2 (streaming)");
			return DK_E_ARG;
		}
		g_ctx.dense_stream = (int)value;
		return DK_OK;
	}
	if (!strcmp(key, "dense_k4_cw")) {
		if (value != 0 && value != 1 && value != 2 && value != 4) {
			dk_set_error("dense_k4_cw is 0 (auto), 1, 2 or 4");
			return DK_E_ARG;
		}
		g_ctx.dense_k4_cw = (int)value;
		return DK_OK;
	}
	if (!strcmp(key, "arena_sub_mib")) {
		if (value < 0 || value > 2048 || (value & (value - 1))) {
			dk_set_error("arena_sub_mib is 0 (whole chunks) or a power of two up to 2048");
			return DK_E_ARG;
		}
		g_ctx.arena_sub_bytes = (u64)value << 20;
		return DK_OK;
	}
	if (!strcmp(key, "arena_max_chunks")) {
		g_ctx.arena_max_chunks = value > 0 ? (u64)value
						   : g_ctx.arena_default_chunks;
		return DK_OK;
	}
	if (!strcmp(key, "guard_alloc")) {
		g_ctx.guard_alloc = value != 0;
		return DK_OK;
	}
	if (!strcmp(key, "poison_alloc")) {
		g_ctx.poison_alloc = value != 0;
		return DK_OK;
	}
	if (!strcmp(key, "arena_reuse")) {
		if (value < 0 || value > 1) {
			dk_set_error("arena_reuse is 0 (exact sizes) or 1 (carve and merge)");
			return DK_E_ARG;
		}
		g_ctx.arena_reuse = (int)value;
		return DK_OK;
	}
	dk_set_error("unknown tunable %s", key);
	return DK_E_ARG;
}

extern "C" int dk_device_ordinal(int *device)
{
	DK_INIT();
	if (!device) {
		dk_set_error("NULL argument");
		return DK_E_ARG;
	}
	*device = g_ctx.device;
	return DK_OK;
}

extern "C" int dk_device_info(int *num_cus, char *arch, size_t arch_len)
{
	DK_INIT();
	hipDeviceProp_t prop;
	DK_HIP(hipGetDeviceProperties(&prop, g_ctx.device));
	if (num_cus)
		*num_cus = prop.multiProcessorCount;
	if (arch && arch_len)
		snprintf(arch, arch_len, "%s", prop.gcnArchName);
	return DK_OK;
}

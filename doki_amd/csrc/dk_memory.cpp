// Device memory of libdoki_hip.so: state vectors, batch buffers and the like.
//
// Physical memory for everything of 64 MiB and more comes from an arena of
// 2 GiB chunks (hipMemCreate) that stay with the process:
//  * blocks of 2 GiB and more (state vectors of 2^27 and more complex128
//    amplitudes) are virtual ranges mapped onto as many whole chunks as they
//    need, sizes rounded up to whole chunks;
//  * blocks of 64 MiB up to 1 GiB are power-of-two pieces of "split" chunks,
//    each split chunk mapped at an address of its own and divided by a buddy
//    system (sizes rounded up to a power of two);
//  * blocks below 64 MiB come from the device's stream-ordered pool
//    (hipMallocAsync, unlimited release threshold).
// Released blocks of the first two kinds stay in size-keyed caches and are
// handed out again.  A whole-chunk request that no cached block fits exactly
// takes the front of the smallest larger one (the rest stays cached, still
// mapped), and released whole-chunk blocks merge with cached neighbours of
// the same address reservation, so that sizes can change without mapping
// anew (mapping and access setting cost about 0.14 ms per GiB, unmapping
// 0.13).  A chunk shortage is met by creating chunks while the device has
// memory, then by unmapping cached blocks and merging cached pieces back
// into whole chunks.  Only when the pool itself runs dry do free chunks go
// back to the driver, one at a time.
//
// The arena exists because memory returned to the driver is wiped before the
// driver hands it out again, at a few tens of GB/s: freeing a 275 GB state
// with hipFree and allocating 34 GB afterwards stalls for seconds, while
// unmapping and remapping the same chunks takes about 0.3 ms per GiB.
//
// Every released block carries an event recorded on the releasing lane's
// stream after the work queued there.  A lane reusing it waits for that event
// on the device; unmapping and merging wait for it on the host.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>
#include <mutex>
#include <set>
#include <unordered_map>
#include <vector>

#include "dk_buddy.h"
#include "dk_internal.h"

namespace {

enum : u64 {
	MID_BLOCK = dkb::PIECE,    // smallest arena piece
	CHUNK = dkb::CHUNK,        // physical chunk
};
static_assert(dkb::PIECE == DK_ARENA_MIN, "arena pieces start at DK_ARENA_MIN");
using dkb::order_of;

// A chunk is CHUNK bytes of physical memory made of `subs.size()` equal
// handles.  They are mapped in a scrambled order (virtual piece j onto handle
// (j * SCRAMBLE) mod n), so that amplitudes a power-of-two distance of one
// piece or more apart do not sit at the same power-of-two distance in
// physical memory.
struct Handle {
	std::vector<hipMemGenericAllocationHandle_t> subs;
};
constexpr unsigned SCRAMBLE = 77;  // odd: a bijection mod any power of two

// An event recorded on a releasing lane's stream after the work it queued
// on the block.
struct Done {
	hipEvent_t ev;
	dk_lane *lane;
};

struct Cached {
	void *p;
	std::vector<Done> done;  // one per release merged into the block
};

struct Mapped {          // a block over whole chunks
	u64 bytes;
	std::vector<Handle> chunks;
	void *resv;          // the address reservation it lies in
};

struct Resv {            // an address reservation of whole-chunk blocks
	u64 bytes;
	unsigned parts;      // mapped blocks inside it
};

std::mutex g_mem_mu;
std::vector<Handle> g_free_chunks;             // created, unmapped
u64 g_chunks_total = 0;
std::multimap<u64, Cached> g_whole_free;        // released whole-chunk blocks
std::map<u64, u64> g_whole_at;                  // the same by address -> bytes
std::unordered_map<void *, Mapped> g_mapped;    // every whole-chunk block
std::unordered_map<void *, Resv> g_resv;        // their address reservations
std::multimap<u64, Cached> g_piece_free;        // released buddy pieces
struct Split {
	Handle h;
	void *resv;  // its address reservation (two chunks, for alignment)
};
std::unordered_map<u64, Split> g_split;         // split chunk base -> chunk
dkb::Buddy g_buddy;                             // free pieces of split chunks

hipMemAllocationProp chunk_prop()
{
	hipMemAllocationProp prop = {};
	prop.type = hipMemAllocationTypePinned;
	prop.location.type = hipMemLocationTypeDevice;
	prop.location.id = g_ctx.device;
	return prop;
}

void chunk_release(Handle &h)
{
	for (auto s : h.subs)
		(void)hipMemRelease(s);
	h.subs.clear();
}

hipError_t chunk_create(Handle *out)
{
	const hipMemAllocationProp prop = chunk_prop();
	u64 sub = g_ctx.arena_sub_bytes;
	if (sub == 0 || sub > CHUNK || CHUNK % sub)
		sub = CHUNK;
	Handle h;
	for (u64 off = 0; off < CHUNK; off += sub) {
		hipMemGenericAllocationHandle_t s;
		hipError_t e = hipMemCreate(&s, sub, &prop, 0);
		if (e != hipSuccess) {
			chunk_release(h);
			return e;
		}
		h.subs.push_back(s);
	}
	*out = std::move(h);
	return hipSuccess;
}

// Maps a chunk at va (CHUNK bytes of reserved addresses); on failure nothing
// stays mapped.
hipError_t chunk_map(void *va, const Handle &h)
{
	const u64 n = h.subs.size(), sub = CHUNK / n;
	for (u64 j = 0; j < n; ++j) {
		hipError_t e = hipMemMap((char *)va + j * sub, sub, 0,
					 h.subs[(j * SCRAMBLE) & (n - 1)], 0);
		if (e != hipSuccess) {
			for (u64 i = 0; i < j; ++i)
				(void)hipMemUnmap((char *)va + i * sub, sub);
			return e;
		}
	}
	return hipSuccess;
}

void chunk_unmap(void *va, const Handle &h)
{
	const u64 n = h.subs.size(), sub = CHUNK / n;
	for (u64 j = 0; j < n; ++j)
		(void)hipMemUnmap((char *)va + j * sub, sub);
}

Cached released_here(void *p)
{
	Cached c{p, {}};
	Done d{nullptr, dk_cur()};
	if (hipEventCreateWithFlags(&d.ev, hipEventDisableTiming) == hipSuccess) {
		(void)hipEventRecord(d.ev, DK_STREAM);
		c.done.push_back(d);
	}
	return c;
}

// The calling lane's stream waits for the previous owners' work where that
// ran on other lanes.
void wait_for(const Cached &c)
{
	for (const Done &d : c.done)
		if (d.lane != dk_cur())
			(void)hipStreamWaitEvent(DK_STREAM, d.ev, 0);
}

// Hands a cached block to the calling lane.
void *take(Cached &c)
{
	wait_for(c);
	for (const Done &d : c.done)
		(void)hipEventDestroy(d.ev);
	c.done.clear();
	return c.p;
}

void settle_host(Cached &c)
{
	for (const Done &d : c.done) {
		(void)hipEventSynchronize(d.ev);
		(void)hipEventDestroy(d.ev);
	}
	c.done.clear();
}

int oom(u64 bytes, hipError_t e)
{
	(void)hipGetLastError();
	dk_set_error("Failed to allocate state chunk: %llu bytes: %s",
		     (unsigned long long)bytes, hipGetErrorString(e));
	return DK_E_ALLOC;
}

// ---- buddy pieces of split chunks -------------------------------------------

// Cached pieces go back into the buddy system (merging where they can).
bool flush_pieces()
{
	if (g_piece_free.empty())
		return false;
	for (auto &kv : g_piece_free) {
		settle_host(kv.second);
		g_buddy.release((u64)kv.second.p, order_of(kv.first));
	}
	g_piece_free.clear();
	return true;
}

// Split chunks whose pieces are all free become whole free chunks again.
bool unsplit_free_chunks()
{
	return g_buddy.drain_whole([](u64 base) {
		auto sp = g_split.find(base);
		chunk_unmap((void *)base, sp->second.h);
		(void)hipMemAddressFree(sp->second.resv, 2 * CHUNK);
		g_free_chunks.push_back(std::move(sp->second.h));
		g_split.erase(sp);
	});
}

// ---- whole chunks ---------------------------------------------------------------

void unmap_whole(void *va)
{
	auto it = g_mapped.find(va);
	if (it == g_mapped.end())
		return;
	for (size_t i = 0; i < it->second.chunks.size(); ++i)
		chunk_unmap((char *)va + i * CHUNK, it->second.chunks[i]);
	auto r = g_resv.find(it->second.resv);
	if (r != g_resv.end() && --r->second.parts == 0) {
		(void)hipMemAddressFree(r->first, r->second.bytes);
		g_resv.erase(r);
	}
	for (auto &h : it->second.chunks)
		g_free_chunks.push_back(std::move(h));
	g_mapped.erase(it);
}

// Removes the cached whole-chunk block at p (bytes long) from the caches.
Cached uncache_whole(u64 p, u64 bytes)
{
	auto range = g_whole_free.equal_range(bytes);
	for (auto it = range.first; it != range.second; ++it)
		if ((u64)it->second.p == p) {
			Cached c = std::move(it->second);
			g_whole_free.erase(it);
			g_whole_at.erase(p);
			return c;
		}
	return Cached{nullptr, {}};
}

// Block b (right after block a, in the same reservation) joins block a.
void merge_mapped(void *a, void *b)
{
	auto ia = g_mapped.find(a), ib = g_mapped.find(b);
	Mapped &ma = ia->second;
	for (auto &h : ib->second.chunks)
		ma.chunks.push_back(std::move(h));
	ma.bytes += ib->second.bytes;
	--g_resv[ma.resv].parts;
	g_mapped.erase(ib);
}

// A released (or split-off) whole-chunk block goes to the cache, merged with
// cached neighbours of its reservation; the merged block waits for every
// part's release.
void cache_whole(Cached c, u64 bytes)
{
	if (!g_ctx.arena_reuse) {
		g_whole_at[(u64)c.p] = bytes;
		g_whole_free.emplace(bytes, std::move(c));
		return;
	}
	void *resv = g_mapped.at(c.p).resv;
	auto left = g_whole_at.lower_bound((u64)c.p);
	if (left != g_whole_at.begin()) {
		--left;
		const u64 lp = left->first, lb = left->second;
		if (lp + lb == (u64)c.p && g_mapped.at((void *)lp).resv == resv) {
			Cached lc = uncache_whole(lp, lb);
			merge_mapped(lc.p, c.p);
			for (const Done &d : c.done)
				lc.done.push_back(d);
			c = std::move(lc);
			bytes += lb;
		}
	}
	auto right = g_whole_at.find((u64)c.p + bytes);
	if (right != g_whole_at.end() && g_mapped.at((void *)right->first).resv == resv) {
		const u64 rb = right->second;
		Cached rc = uncache_whole(right->first, rb);
		merge_mapped(c.p, rc.p);
		for (const Done &d : rc.done)
			c.done.push_back(d);
		bytes += rb;
	}
	g_whole_at[(u64)c.p] = bytes;
	g_whole_free.emplace(bytes, std::move(c));
}

// Unmaps one cached whole-chunk block (the largest) and keeps its chunks.
bool harvest_whole()
{
	if (g_whole_free.empty())
		return false;
	auto it = std::prev(g_whole_free.end());
	Cached c = std::move(it->second);
	g_whole_at.erase((u64)c.p);
	g_whole_free.erase(it);
	settle_host(c);
	unmap_whole(c.p);
	return true;
}

// Makes `need` chunks free: new ones while the device (and the cap) allow,
// then chunks of cached blocks and pieces.
int gather_chunks(size_t need, u64 bytes)
{
	bool trimmed = false;
	while (g_free_chunks.size() < need) {
		const bool capped = g_chunks_total >= g_ctx.arena_max_chunks;
		hipError_t e = hipErrorOutOfMemory;
		if (!capped) {
			Handle h;
			e = chunk_create(&h);
			if (e == hipSuccess) {
				g_free_chunks.push_back(std::move(h));
				++g_chunks_total;
				continue;
			}
			(void)hipGetLastError();
		}
		if (harvest_whole())
			continue;
		if (flush_pieces() && unsplit_free_chunks())
			continue;
		if (unsplit_free_chunks())
			continue;
		if (!capped && !trimmed) {
			// memory parked in the stream-ordered pool
			hipMemPool_t pool;
			if (hipDeviceGetDefaultMemPool(&pool, g_ctx.device) == hipSuccess) {
				(void)hipStreamSynchronize(DK_STREAM);
				(void)hipMemPoolTrimTo(pool, 0);
			}
			trimmed = true;
			continue;
		}
		return oom(bytes, e);
	}
	return DK_OK;
}

int alloc_whole(u64 bytes, void **out)
{
	auto it = g_whole_free.find(bytes);
	if (it != g_whole_free.end()) {
		*out = take(it->second);
		g_whole_at.erase((u64)it->second.p);
		g_whole_free.erase(it);
		return DK_OK;
	}
	// the front of the smallest larger cached block, without mapping: the
	// rest stays cached and still waits for the block's last users (this
	// lane's stream waits for them too)
	auto big = g_ctx.arena_reuse ? g_whole_free.upper_bound(bytes) : g_whole_free.end();
	if (big != g_whole_free.end()) {
		const u64 total = big->first;
		Cached c = std::move(big->second);
		g_whole_at.erase((u64)c.p);
		g_whole_free.erase(big);
		wait_for(c);
		auto im = g_mapped.find(c.p);
		Mapped rest;
		rest.bytes = total - bytes;
		rest.resv = im->second.resv;
		const size_t k = bytes / CHUNK;
		for (size_t i = k; i < im->second.chunks.size(); ++i)
			rest.chunks.push_back(std::move(im->second.chunks[i]));
		im->second.chunks.resize(k);
		im->second.bytes = bytes;
		++g_resv[rest.resv].parts;
		void *rp = (char *)c.p + bytes;
		g_mapped.emplace(rp, std::move(rest));
		*out = c.p;
		cache_whole(Cached{rp, std::move(c.done)}, total - bytes);
		return DK_OK;
	}
	const size_t need = bytes / CHUNK;
	int rc = gather_chunks(need, bytes);
	if (rc)
		return rc;
	void *va = nullptr;
	hipError_t e = hipMemAddressReserve(&va, bytes, CHUNK, nullptr, 0);
	if (e != hipSuccess)
		return oom(bytes, e);
	Mapped m;
	m.bytes = bytes;
	m.resv = va;
	g_resv[va] = Resv{bytes, 1};
	for (size_t i = 0; i < need && e == hipSuccess; ++i) {
		e = chunk_map((char *)va + i * CHUNK, g_free_chunks.back());
		if (e == hipSuccess) {
			m.chunks.push_back(std::move(g_free_chunks.back()));
			g_free_chunks.pop_back();
		}
	}
	if (e == hipSuccess) {
		const hipMemAllocationProp prop = chunk_prop();
		hipMemAccessDesc ad = {};
		ad.location = prop.location;
		ad.flags = hipMemAccessFlagsProtReadWrite;
		e = hipMemSetAccess(va, bytes, &ad, 1);
	}
	g_mapped.emplace(va, std::move(m));
	if (e != hipSuccess) {
		unmap_whole(va);
		return oom(bytes, e);
	}
	*out = va;
	return DK_OK;
}

int alloc_piece(u64 bytes, void **out)
{
	auto it = g_piece_free.find(bytes);
	if (it != g_piece_free.end()) {
		*out = take(it->second);
		g_piece_free.erase(it);
		return DK_OK;
	}
	const int o = order_of(bytes);
	u64 addr;
	if (!g_buddy.take(o, &addr) && !(flush_pieces() && g_buddy.take(o, &addr))) {
		// split a chunk, mapped at an address of its own
		int rc = gather_chunks(1, bytes);
		if (rc)
			return rc;
		// the buddy arithmetic needs a chunk-aligned base: reserve two
		// chunks of addresses and map at the aligned one inside
		void *resv = nullptr;
		hipError_t e = hipMemAddressReserve(&resv, 2 * CHUNK, 0, nullptr, 0);
		if (e != hipSuccess)
			return oom(bytes, e);
		void *va = (void *)(((u64)resv + CHUNK - 1) & ~(CHUNK - 1));
		e = chunk_map(va, g_free_chunks.back());
		if (e == hipSuccess) {
			const hipMemAllocationProp prop = chunk_prop();
			hipMemAccessDesc ad = {};
			ad.location = prop.location;
			ad.flags = hipMemAccessFlagsProtReadWrite;
			e = hipMemSetAccess(va, CHUNK, &ad, 1);
			if (e != hipSuccess)
				chunk_unmap(va, g_free_chunks.back());
		}
		if (e != hipSuccess) {
			(void)hipMemAddressFree(resv, 2 * CHUNK);
			return oom(bytes, e);
		}
		g_split.emplace((u64)va, Split{std::move(g_free_chunks.back()), resv});
		g_free_chunks.pop_back();
		g_buddy.release((u64)va, dkb::ORDERS - 1);
		g_buddy.take(o, &addr);
	}
	*out = (void *)addr;
	return DK_OK;
}

// Pool shortage: cached blocks and pieces go back to chunks, and free chunks
// go back to the driver one at a time until `retry` succeeds.
template <typename F>
hipError_t relieve_pool(F retry)
{
	hipError_t e = retry();
	while (e != hipSuccess) {
		(void)hipGetLastError();
		if (g_free_chunks.empty()) {
			if (!harvest_whole() && !(flush_pieces(), unsplit_free_chunks()))
				break;
			continue;
		}
		chunk_release(g_free_chunks.back());
		g_free_chunks.pop_back();
		--g_chunks_total;
		e = retry();
	}
	return e;
}

} // namespace

u64 dk_block_bytes(u64 bytes)
{
	if (bytes < MID_BLOCK)
		return bytes;
	if (bytes <= CHUNK / 2)
		return MID_BLOCK << order_of(bytes);
	return (bytes + CHUNK - 1) / CHUNK * CHUNK;
}

namespace {

int device_alloc(u64 bytes, void **out)
{
	*out = nullptr;
	if (bytes < MID_BLOCK) {
		hipError_t e = hipMallocAsync(out, bytes, DK_STREAM);
		if (e != hipSuccess) {
			(void)hipGetLastError();
			std::lock_guard<std::mutex> lk(g_mem_mu);
			e = relieve_pool([&] { return hipMallocAsync(out, bytes, DK_STREAM); });
		}
		if (e != hipSuccess) {
			*out = nullptr;
			return oom(bytes, e);
		}
		return DK_OK;
	}
	const u64 b = dk_block_bytes(bytes);
	std::lock_guard<std::mutex> lk(g_mem_mu);
	return b >= CHUNK ? alloc_whole(b, out) : alloc_piece(b, out);
}

// Guard bands ("guard_alloc"): a pool block is handed out GUARD bytes into
// a larger one, with GUARD bytes of a known pattern before it and after its
// size rounded up to 256; its release counts the changed guard words (the
// stream is synchronised) and reports them with the block's size.
constexpr u64 GUARD = 64 << 10;
struct Guarded {
	char *raw;
	u64 span;
};
std::mutex g_guard_mu;
std::unordered_map<u64, Guarded> g_guarded;
unsigned *g_guard_bad = nullptr;

int guarded_alloc(u64 bytes, void **out)
{
	const u64 body = (bytes + 255) / 256 * 256;
	const u64 span = body + 2 * GUARD;
	if (span >= MID_BLOCK)
		return device_alloc(bytes, out);
	char *raw = nullptr;
	int rc = device_alloc(span, (void **)&raw);
	if (rc)
		return rc;
	(void)dk_launch_guard_fill(raw, GUARD);
	(void)dk_launch_guard_fill(raw + GUARD + body, GUARD);
	*out = raw + GUARD;
	std::lock_guard<std::mutex> lk(g_guard_mu);
	g_guarded[(u64)*out] = Guarded{raw, span};
	return DK_OK;
}

// True when p was a guarded block (then released here).
bool guarded_free(void *p, u64 bytes)
{
	Guarded g;
	{
		std::lock_guard<std::mutex> lk(g_guard_mu);
		auto it = g_guarded.find((u64)p);
		if (it == g_guarded.end())
			return false;
		g = it->second;
		g_guarded.erase(it);
		if (!g_guard_bad && hipMalloc((void **)&g_guard_bad, 2 * sizeof(unsigned)) != hipSuccess)
			g_guard_bad = nullptr;
	}
	if (g_guard_bad) {
		unsigned h[2] = {0, 0};
		unsigned *bad = nullptr;
		if (hipMallocAsync((void **)&bad, 2 * sizeof(unsigned), DK_STREAM) == hipSuccess) {
			(void)hipMemcpyAsync(bad, h, sizeof(h), hipMemcpyHostToDevice, DK_STREAM);
			(void)dk_launch_guard_check(g.raw, GUARD, bad);
			(void)dk_launch_guard_check(g.raw + g.span - GUARD, GUARD, bad + 1);
			(void)hipMemcpyAsync(h, bad, sizeof(h), hipMemcpyDeviceToHost, DK_STREAM);
			(void)hipStreamSynchronize(DK_STREAM);
			(void)hipFreeAsync(bad, DK_STREAM);
			if (h[0] || h[1])
				fprintf(stderr, "[guard] block of %llu bytes at %p: %u words "
					"before and %u after it changed\n",
					(unsigned long long)bytes, p, h[0], h[1]);
		}
	}
	(void)hipFreeAsync(g.raw, DK_STREAM);
	return true;
}

} // namespace

int dk_device_alloc(u64 bytes, void **out)
{
	const int rc = g_ctx.guard_alloc && bytes < MID_BLOCK ? guarded_alloc(bytes, out)
							     : device_alloc(bytes, out);
	if (!rc && g_ctx.poison_alloc)
		return dk_launch_poison(*out, bytes);
	return rc;
}

void dk_device_free(void *p, u64 bytes)
{
	if (!p || !g_ctx.ready || dk_closed())
		return;
	if (bytes < MID_BLOCK) {
		if (!guarded_free(p, bytes))
			(void)hipFreeAsync(p, DK_STREAM);
		return;
	}
	const u64 b = dk_block_bytes(bytes);
	Cached c = released_here(p);
	std::lock_guard<std::mutex> lk(g_mem_mu);
	if (b >= CHUNK)
		cache_whole(std::move(c), b);
	else
		g_piece_free.emplace(b, std::move(c));
}

// Cached blocks and pieces, and every chunk no live block needs, go back to
// the driver; the pool is trimmed.
extern "C" int dk_memory_trim(void)
{
	int rc = dk_ensure_init();
	if (rc)
		return rc;
	std::lock_guard<std::mutex> lk(g_mem_mu);
	while (harvest_whole()) {
	}
	flush_pieces();
	unsplit_free_chunks();
	for (auto &h : g_free_chunks)
		chunk_release(h);
	g_chunks_total -= g_free_chunks.size();
	g_free_chunks.clear();
	hipMemPool_t pool;
	if (hipDeviceGetDefaultMemPool(&pool, g_ctx.device) == hipSuccess) {
		(void)hipDeviceSynchronize();
		(void)hipMemPoolTrimTo(pool, 0);
	}
	return DK_OK;
}

extern "C" int dk_memory_stats(uint64_t *arena_chunks, uint64_t *arena_free_chunks,
			       uint64_t *cached_bytes)
{
	std::lock_guard<std::mutex> lk(g_mem_mu);
	u64 cached = 0;
	for (auto &kv : g_whole_free)
		cached += kv.first;
	for (auto &kv : g_piece_free)
		cached += kv.first;
	if (arena_chunks)
		*arena_chunks = g_chunks_total;
	if (arena_free_chunks)
		*arena_free_chunks = g_free_chunks.size();
	if (cached_bytes)
		*cached_bytes = cached;
	return DK_OK;
}

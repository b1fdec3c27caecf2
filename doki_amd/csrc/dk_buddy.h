// Buddy division of 2 GiB chunks into power-of-two pieces of 64 MiB to
// 1 GiB (host bookkeeping only; chunks are identified by their base address,
// which is a multiple of the chunk size).
#pragma once

#include <cstdint>
#include <set>

namespace dkb {

constexpr uint64_t PIECE = 64ull << 20;
constexpr uint64_t CHUNK = 2ull << 30;
constexpr int ORDERS = 6;  // pieces of PIECE << o; order 5 is a whole chunk

inline int order_of(uint64_t bytes)
{
	int o = 0;
	while ((PIECE << o) < bytes)
		++o;
	return o;
}

struct Buddy {
	std::set<uint64_t> free_[ORDERS];

	// A piece of order o at addr becomes free, merging with its free buddy
	// up to whole chunks.
	void release(uint64_t addr, int o)
	{
		while (o < ORDERS - 1) {
			const uint64_t base = addr & ~(CHUNK - 1);
			const uint64_t buddy = base + ((addr - base) ^ (PIECE << o));
			auto it = free_[o].find(buddy);
			if (it == free_[o].end())
				break;
			free_[o].erase(it);
			addr = addr < buddy ? addr : buddy;
			++o;
		}
		free_[o].insert(addr);
	}

	// Takes a free piece of order o, splitting a larger one if needed.
	bool take(int o, uint64_t *addr)
	{
		int k = o;
		while (k < ORDERS && free_[k].empty())
			++k;
		if (k == ORDERS)
			return false;
		const uint64_t a = *free_[k].begin();
		free_[k].erase(free_[k].begin());
		while (k > o) {
			--k;
			free_[k].insert(a + (PIECE << k));
		}
		*addr = a;
		return true;
	}

	// Whole free chunks (order 5), removed from the buddy system.
	template <typename F>
	bool drain_whole(F on_chunk)
	{
		bool any = false;
		auto &top = free_[ORDERS - 1];
		for (auto it = top.begin(); it != top.end(); it = top.erase(it)) {
			on_chunk(*it);
			any = true;
		}
		return any;
	}
};

} // namespace dkb

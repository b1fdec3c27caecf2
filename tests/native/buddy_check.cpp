// Randomised check of the chunk buddy bookkeeping (doki_amd/csrc/dk_buddy.h):
// pieces handed out never overlap, stay inside their chunk, are aligned to
// their size, and releasing everything merges back into whole chunks.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "dk_buddy.h"

using dkb::CHUNK;
using dkb::PIECE;

int main(int argc, char **argv)
{
	const unsigned seed = argc > 1 ? (unsigned)atoi(argv[1]) : 1u;
	std::mt19937_64 rng(seed);
	dkb::Buddy b;
	const int chunks = 5;
	for (int c = 0; c < chunks; ++c)
		b.release((uint64_t)(c + 3) * CHUNK, dkb::ORDERS - 1);
	std::map<uint64_t, uint64_t> live;  // addr -> bytes
	for (int step = 0; step < 200000; ++step) {
		if (live.empty() || rng() % 2) {
			const int o = (int)(rng() % (dkb::ORDERS - 1));
			uint64_t a;
			if (!b.take(o, &a))
				continue;
			const uint64_t sz = PIECE << o;
			if (a % sz || a / CHUNK != (a + sz - 1) / CHUNK) {
				printf("misplaced piece %llx size %llx\n",
				       (unsigned long long)a, (unsigned long long)sz);
				return 1;
			}
			auto nx = live.lower_bound(a);
			if ((nx != live.end() && nx->first < a + sz) ||
			    (nx != live.begin() && std::prev(nx)->first +
							   std::prev(nx)->second > a)) {
				printf("overlap at %llx\n", (unsigned long long)a);
				return 1;
			}
			live[a] = sz;
		} else {
			auto it = live.begin();
			std::advance(it, (long)(rng() % live.size()));
			b.release(it->first, dkb::order_of(it->second));
			live.erase(it);
		}
	}
	for (auto &kv : live)
		b.release(kv.first, dkb::order_of(kv.second));
	int whole = 0;
	b.drain_whole([&](uint64_t) { ++whole; });
	for (int o = 0; o < dkb::ORDERS; ++o)
		if (!b.free_[o].empty()) {
			printf("order %d not merged (%zu left)\n", o, b.free_[o].size());
			return 1;
		}
	if (whole != chunks) {
		printf("%d whole chunks back, want %d\n", whole, chunks);
		return 1;
	}
	printf("ok\n");
	return 0;
}

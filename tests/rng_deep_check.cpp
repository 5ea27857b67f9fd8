// Host check of the device RNG engines against libstdc++ std::mt19937_64: every output, random
// patterns of draws and discards, and replay copies (forks) taken at random points, for streams
// long enough to cross the first, second and third twists; and streams whose spill slot is
// prepared ahead of output 624 (mt_prepare_third, at a random point, forks carrying it along).
// Prints "ok" or the first mismatch.
#include <cstdio>
#include <random>
#include <vector>

#include "../hdrestir_amd/csrc/mt64.h"

using namespace hdr;

// wavefront.h: rng_fork
static Mt64Deep fork_of(const Mt64Deep &r, uint64_t *slot2) {
    Mt64Deep c = r;
    if (r.k > 624u || r.ready) mt_copy_slot(slot2, r.slot);
    c.slot = slot2;
    return c;
}

int main() {
    const int kDraws = 1400;
    std::vector<uint64_t> slots(2 * kMtN);
    for (uint32_t seed : {0u, 1u, 12345u, 0xdeadbeefu, 4294967295u, 777u, 31337u, 2024u}) {
        std::mt19937_64 ref(seed);
        std::vector<uint64_t> want(kDraws);
        for (auto &w : want) w = ref();
        // every output (both engines)
        Mt64Deep d;
        d.init(seed, slots.data());
        Mt64 m;
        std::vector<uint64_t> mslot(kMtN);
        m.init(seed, mslot.data());
        for (int i = 0; i < kDraws; ++i) {
            const uint64_t a = d.next_raw(), b = m.next_raw();
            if (a != want[i] || b != want[i]) {
                std::printf("seed %u output %d: deep %llx mt64 %llx want %llx\n", seed, i, (unsigned long long)a,
                            (unsigned long long)b, (unsigned long long)want[i]);
                return 1;
            }
        }
        // the prepared slot equals the lazily built one at output 624
        {
            std::vector<uint64_t> lazy(2 * kMtN), prep(2 * kMtN);
            Mt64Deep a;
            a.init(seed, lazy.data());
            for (int i = 0; i <= 624; ++i) (void)a.next_raw();
            mt_prepare_third(prep.data(), prep.data() + kMtN, seed);
            for (int q = 0; q < kMtN; ++q)
                if (prep[size_t(q)] != lazy[size_t(q)]) {
                    std::printf("seed %u prepared slot word %d differs\n", seed, q);
                    return 1;
                }
        }
        // draws and discards in a pseudo-random pattern, with forks replayed against the reference;
        // odd repetitions prepare the slot at a random point before output 624
        uint64_t lcg = seed * 2862933555777941757ULL + 3037000493ULL;
        for (int rep = 0; rep < 20; ++rep) {
            Mt64Deep r;
            r.init(seed, slots.data());
            lcg = lcg * 6364136223846793005ULL + 1442695040888963407ULL;
            const int prepAt = (rep & 1) ? 300 + int((lcg >> 33) % 324u) : -1;
            for (int i = 0; i < kDraws; ++i) {
                if (i == prepAt) {
                    mt_prepare_third(r.slot, slots.data() + kMtN, seed);
                    r.ready = true;
                }
                lcg = lcg * 6364136223846793005ULL + 1442695040888963407ULL;
                const int op = int((lcg >> 33) % 7u);
                if (op < 3) {
                    r.discard();
                } else if (op == 3 && i + 40 < kDraws) {
                    Mt64Deep f = fork_of(r, slots.data() + kMtN);
                    for (int q = 0; q < 40; ++q) {
                        const uint64_t v = f.next_raw();
                        if (v != want[i + q]) {
                            std::printf("seed %u fork at %d draw %d mismatch\n", seed, i, q);
                            return 1;
                        }
                    }
                    const uint64_t v = r.next_raw();
                    if (v != want[i]) {
                        std::printf("seed %u after fork at %d mismatch\n", seed, i);
                        return 1;
                    }
                } else {
                    const uint64_t v = r.next_raw();
                    if (v != want[i]) {
                        std::printf("seed %u rep %d output %d: %llx want %llx\n", seed, rep, i, (unsigned long long)v,
                                    (unsigned long long)want[i]);
                        return 1;
                    }
                }
            }
        }
    }
    std::printf("ok\n");
    return 0;
}

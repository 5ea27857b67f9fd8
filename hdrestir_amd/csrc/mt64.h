// mt64.h -- per-pixel std::mt19937_64 stream in O(1) registers.
//
// The reference seeds one std::mt19937_64 per pixel per frame (source/core/rng.h:10-56,
// source/passes/integration_pass.cpp:210-211) and draws floats through
// std::uniform_real_distribution<float> (libstdc++ generate_canonical<float, 24>: one 64-bit
// output x -> float(x) * 2^-64, clamped below 1). Materialising the 312-word state per pixel
// would cost 2.5 KB per pixel, so this engine evaluates the first twist lazily:
//
//   seeding        s[0] = seed, s[i] = F * (s[i-1] ^ (s[i-1] >> 62)) + i      (i < 312)
//   first twist    s'[k] = s[k+156] ^ T(s[k], s[k+1])                         (k < 156)
//                  s'[k] = s[k] ^ T(s[k-156], s[k-155]) ^ T(s[k], s[k+1])     (156 <= k < 311)
//                  s'[311] = s'[155] ^ T(s[311], s'[0])
//   output k       temper(s'[k])
//
// with T(a, b) = (((a & UPPER) | (b & LOWER)) >> 1) ^ (b & 1 ? A : 0). Two cursors walk the
// seeding recurrence 156 words apart (lo at j, hi at j + 156), so output k costs two
// recurrence steps. Outputs from 312 on (second twist) are produced by materialising the
// full state in a caller-provided 312-word global slot; paths that can draw that many are
// rare (deep paths with large candidate counts) and the host sizes the slot pool for them.
#pragma once

#include "gfm.h"

namespace hdr {

constexpr uint64_t kMtF = 6364136223846793005ULL;
constexpr uint64_t kMtA = 0xB5026F5AA96619E9ULL;
constexpr uint64_t kMtLower = (1ULL << 31) - 1ULL;
constexpr uint64_t kMtUpper = ~kMtLower;
constexpr int kMtN = 312;
constexpr int kMtM = 156;

HD uint64_t mt_step(uint64_t prev, uint32_t i) { return kMtF * (prev ^ (prev >> 62)) + uint64_t(i); }
HD uint64_t mt_twist(uint64_t a, uint64_t b) {
    const uint64_t y = (a & kMtUpper) | (b & kMtLower);
    return (y >> 1) ^ ((y & 1ULL) ? kMtA : 0ULL);
}
HD uint64_t mt_temper(uint64_t z) {
    z ^= (z >> 29) & 0x5555555555555555ULL;
    z ^= (z << 17) & 0x71D67FFFEDA60000ULL;
    z ^= (z << 37) & 0xFFF7EEE000000000ULL;
    z ^= (z >> 43);
    return z;
}
// libstdc++ generate_canonical<float, 24> with one 64-bit draw, then uniform_real_distribution
// (b - a = 1, a = 0): float(x) / 2^64, or nextafter(1, 0) if that rounds to 1.
HD float mt_canonical(uint64_t x) {
    float r = float(x) * 0x1p-64f;  // exact scaling by 2^-64
    return (r >= 1.0f) ? 0.99999994f : r;
}

struct Mt64 {
    uint64_t lo, lo1, hi, hi1;  // s[j], s[j+1], s[j+156], s[j+157]
    uint64_t s156;
    uint64_t *slot;             // 312-word spill slot for outputs >= 312 (may be null)
    uint32_t seed;
    uint32_t k;                 // outputs produced so far
    uint32_t j;                 // cursor position

    HD void init(uint32_t sd, uint64_t *spill) {
        seed = sd;
        slot = spill;
        k = 0;
        j = 0;
        lo = uint64_t(sd);
        lo1 = mt_step(lo, 1);
        uint64_t x = lo1;
#pragma unroll 4
        for (uint32_t i = 2; i <= 156; ++i) x = mt_step(x, i);
        s156 = x;
        hi = x;
        hi1 = mt_step(x, 157);
    }
    HD void advance() {
        lo = lo1;
        lo1 = mt_step(lo1, j + 2);
        hi = hi1;
        hi1 = mt_step(hi1, j + 158);
        ++j;
    }
    HD void restart() {
        j = 0;
        lo = uint64_t(seed);
        lo1 = mt_step(lo, 1);
        hi = s156;
        hi1 = mt_step(s156, 157);
    }
    // Fills the spill slot with the state after the second twist (outputs 312..623 untempered).
    HD void materialise_second() {
        uint64_t *s = slot;
        s[0] = uint64_t(seed);
        for (int i = 1; i < kMtN; ++i) s[i] = mt_step(s[i - 1], uint32_t(i));
        for (int pass = 0; pass < 2; ++pass) twist_in_place(s);
    }
    HD static void twist_in_place(uint64_t *s) {
        for (int q = 0; q < kMtN - kMtM; ++q) s[q] = s[q + kMtM] ^ mt_twist(s[q], s[q + 1]);
        for (int q = kMtN - kMtM; q < kMtN - 1; ++q) s[q] = s[q + kMtM - kMtN] ^ mt_twist(s[q], s[q + 1]);
        s[kMtN - 1] = s[kMtM - 1] ^ mt_twist(s[kMtN - 1], s[0]);
    }
    HD uint64_t next_raw() {
        uint64_t y;
        if (k < 156u) {
            y = hi ^ mt_twist(lo, lo1);
            advance();
        } else if (k < 311u) {
            if (k == 156u) restart();
            y = hi ^ mt_twist(lo, lo1) ^ mt_twist(hi, hi1);
            advance();
        } else if (k == 311u) {
            const uint64_t t155 = hi ^ mt_twist(lo, lo1);  // s'[155]
            const uint64_t s1 = mt_step(uint64_t(seed), 1);
            const uint64_t t0 = s156 ^ mt_twist(uint64_t(seed), s1);  // s'[0]
            y = t155 ^ mt_twist(hi, t0);
        } else {
            if (!slot) {  // the host allocates slots whenever a pixel can draw this far
                ++k;
                return 0;
            }
            const uint32_t idx = (k - 312u) % uint32_t(kMtN);
            if (idx == 0u) {
                if (k == 312u) materialise_second();
                else twist_in_place(slot);
            }
            y = slot[idx];
        }
        ++k;
        return mt_temper(y);
    }
    HD float next_float() { return mt_canonical(next_raw()); }
    HD void discard() { (void)next_raw(); }
};

// The stream for callers that provably never draw past output 310 (the host checks a per-pixel
// draw bound): the second-twist branches are dropped and a discarded draw only moves the cursors.
struct Mt64Lean : Mt64 {
    HD uint64_t next_raw() {
        uint64_t y;
        if (k < 156u) {
            y = hi ^ mt_twist(lo, lo1);
        } else {
            if (k == 156u) restart();
            y = hi ^ mt_twist(lo, lo1) ^ mt_twist(hi, hi1);
        }
        advance();
        ++k;
        return mt_temper(y);
    }
    HD float next_float() { return mt_canonical(next_raw()); }
    HD void discard() {
        if (k == 156u) restart();
        advance();
        ++k;
    }
};

}  // namespace hdr

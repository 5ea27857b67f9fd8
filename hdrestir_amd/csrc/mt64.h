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
// Mt64Deep (below) also evaluates the second twist lazily and needs the slot from output 624 on.
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
    bool ready;                 // Mt64Deep: the slot already holds the third twist (mt_prepare_third)

    HD void init(uint32_t sd, uint64_t *spill) {
        seed = sd;
        slot = spill;
        k = 0;
        j = 0;
        ready = false;
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

// The stream for paths that may draw past output 310 (the wavefront path's deep paths).
// Outputs 0 .. 310 as Mt64 (first twist from two seeding cursors). Outputs 311 .. 623 -- the
// second twist -- are evaluated lazily as well, from the same two cursors restarted at the seed:
// with X the engine's word sequence (X[i] = seeding word S(i) for i < 312, X[n] = X[n-156] ^
// T(X[n-312], X[n-311]) after) and B1(j) = X[312+j], B2(j) = X[624+j] (output 312+j is B2(j)
// tempered):
//   B1(u)       = S(u+156) ^ T(S(u), S(u+1))                      u < 156
//   B1(u+156)   = B1(u) ^ T(S(u+156), S(u+157))                   u < 155
//   B1(311)     = B1(155) ^ T(S(311), B1(0))
//   B2(u)       = B1(u+156) ^ T(B1(u), B1(u+1))                   u < 156    (pass A, u = k - 312)
//   B2(u+156)   = B2(u) ^ T(B1(u+156), B1(u+157))                 u < 155    (pass B, u = k - 468)
//   B2(311)     = B2(155) ^ T(B1(311), B2(0))
// so each output costs a few seeding steps and twists and no memory. B1(155) and B1(311) are kept
// from pass A for pass B. From output 624 on (third twist) the state is materialised in the spill
// slot; those outputs go through one out-of-line function, so the inlined draw sites stay small.
struct Mt64Deep : Mt64 {
    uint64_t b155, b311;  // B1(155), B1(311), set during pass A

    HD uint64_t b1_0() const { return s156 ^ mt_twist(uint64_t(seed), mt_step(uint64_t(seed), 1)); }
    HD uint64_t b1_156(uint64_t c0) const { return c0 ^ mt_twist(s156, mt_step(s156, 157)); }

    // Outputs 311 .. 623 (y != nullptr) or their discard (only the cursors move).
    HD void second(uint64_t *y) {
        if (k == 311u) {  // B1(311): cursors at j = 155 (lo = S(155), lo1 = S(156), hi = S(311))
            if (y) {
                const uint64_t t155 = hi ^ mt_twist(lo, lo1);
                *y = t155 ^ mt_twist(hi, b1_0());
            }
            return;
        }
        if (k == 312u || k == 468u) restart();
        if (k < 468u) {  // pass A, u = j = k - 312
            if (j < 155u) {
                if (y) {
                    const uint64_t P = hi ^ mt_twist(lo, lo1);  // B1(u)
                    const uint64_t lo2 = mt_step(lo1, j + 2);
                    const uint64_t Q = hi1 ^ mt_twist(lo1, lo2);  // B1(u+1)
                    *y = P ^ mt_twist(hi, hi1) ^ mt_twist(P, Q);
                }
                advance();
            } else {  // u = 155: hi = S(311)
                const uint64_t P = hi ^ mt_twist(lo, lo1);  // B1(155)
                const uint64_t c0 = b1_0();
                b155 = P;
                b311 = P ^ mt_twist(hi, c0);
                if (y) *y = b311 ^ mt_twist(P, b1_156(c0));
            }
        } else if (k < 623u) {  // pass B, u = j = k - 468 <= 154
            if (y) {
                const uint64_t P = hi ^ mt_twist(lo, lo1);                // B1(u)
                const uint64_t lo2 = mt_step(lo1, j + 2);
                const uint64_t Q = hi1 ^ mt_twist(lo1, lo2);              // B1(u+1)
                const uint64_t R = P ^ mt_twist(hi, hi1);                 // B1(u+156)
                const uint64_t B2u = R ^ mt_twist(P, Q);                  // B2(u)
                const uint64_t S = j < 154u ? Q ^ mt_twist(hi1, mt_step(hi1, j + 158)) : b311;  // B1(u+157)
                *y = B2u ^ mt_twist(R, S);
            }
            advance();
        } else if (y) {  // k = 623: B2(311) = B2(155) ^ T(B1(311), B2(0))
            const uint64_t c0 = b1_0(), c156 = b1_156(c0);
            const uint64_t s1 = mt_step(uint64_t(seed), 1), s2 = mt_step(s1, 2);
            const uint64_t b1 = mt_step(s156, 157) ^ mt_twist(s1, s2);  // B1(1)
            const uint64_t b2_0 = c156 ^ mt_twist(c0, b1);
            const uint64_t b2_155 = b311 ^ mt_twist(b155, c156);
            *y = b2_155 ^ mt_twist(b311, b2_0);
        }
    }
    // Output 624 and beyond: the spill slot holds the state after the third twist, twisted again
    // every 312 outputs. A slot prepared ahead of output 624 (ready) is used as it is.
    HD uint64_t third() {
        if (!slot) return 0;  // the host allocates slots whenever a pixel can draw this far
        const uint32_t idx = (k - 624u) % uint32_t(kMtN);
        if (idx == 0u) {
            if (k != 624u) {
                twist_in_place(slot);
            } else if (!ready) {
                materialise_second();
                twist_in_place(slot);
            }
        }
        return slot[idx];
    }

    HD uint64_t next_raw();
    HD float next_float() { return mt_canonical(next_raw()); }
    HD void discard();
};

static __host__ __device__ __noinline__ uint64_t mt_deep_next(Mt64Deep &r) {
    uint64_t y = 0;
    if (r.k < 624u) r.second(&y);
    else y = r.third();
    ++r.k;
    return mt_temper(y);
}
static __host__ __device__ __noinline__ void mt_deep_discard(Mt64Deep &r) {
    if (r.k < 624u) r.second(nullptr);
    else (void)r.third();
    ++r.k;
}
static __host__ __device__ __noinline__ void mt_copy_slot(uint64_t *__restrict__ dst,
                                                          const uint64_t *__restrict__ src) {
#pragma unroll 8
    for (int q = 0; q < kMtN; ++q) dst[q] = src[q];
}

// One twist from `in` to `out` (two distinct slots): the in-place twist's recurrence with every
// old word read from `in`, so the loads of a chunk do not wait on the chunk's stores. Chunks of 8
// words: loads first, then the words.
HD void mt_twist_into(uint64_t *__restrict__ out, const uint64_t *__restrict__ in) {
    constexpr int kC = 8;
    for (int q0 = 0; q0 < kMtN - kMtM; q0 += kC) {  // 156 = 19 chunks + 4
        uint64_t a[kC + 1], b[kC];
        const int n = (kMtN - kMtM - q0) < kC ? (kMtN - kMtM - q0) : kC;
#pragma unroll
        for (int c = 0; c <= kC; ++c)
            if (c <= n) a[c] = in[q0 + c];
#pragma unroll
        for (int c = 0; c < kC; ++c)
            if (c < n) b[c] = in[q0 + c + kMtM];
#pragma unroll
        for (int c = 0; c < kC; ++c)
            if (c < n) out[q0 + c] = b[c] ^ mt_twist(a[c], a[c + 1]);
    }
    for (int q0 = kMtN - kMtM; q0 < kMtN - 1; q0 += kC) {  // 155 = 19 chunks + 3
        uint64_t a[kC + 1], b[kC];
        const int n = (kMtN - 1 - q0) < kC ? (kMtN - 1 - q0) : kC;
#pragma unroll
        for (int c = 0; c <= kC; ++c)
            if (c <= n) a[c] = in[q0 + c];
#pragma unroll
        for (int c = 0; c < kC; ++c)
            if (c < n) b[c] = out[q0 + c + kMtM - kMtN];  // new words, written by the first half
#pragma unroll
        for (int c = 0; c < kC; ++c)
            if (c < n) out[q0 + c] = b[c] ^ mt_twist(a[c], a[c + 1]);
    }
    out[kMtN - 1] = out[kMtM - 1] ^ mt_twist(in[kMtN - 1], out[0]);
}

// The spill slot's state at output 624 (the third twist) computed ahead of time: the seeding words
// into `scratch`, then three twists ping-ponging scratch -> slot -> scratch -> slot. Equal to what
// Mt64Deep::third computes lazily at output 624 (materialise_second, then twist_in_place).
HD void mt_prepare_third(uint64_t *__restrict__ slot, uint64_t *__restrict__ scratch, uint32_t seed) {
    uint64_t x = uint64_t(seed);
    scratch[0] = x;
    for (int i = 1; i < kMtN; ++i) {
        x = mt_step(x, uint32_t(i));
        scratch[i] = x;
    }
    mt_twist_into(slot, scratch);
    mt_twist_into(scratch, slot);
    mt_twist_into(slot, scratch);
}

HD uint64_t Mt64Deep::next_raw() {
    if (k >= 311u) return mt_deep_next(*this);
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
HD void Mt64Deep::discard() {
    if (k >= 311u) {
        mt_deep_discard(*this);
        return;
    }
    if (k == 156u) restart();
    advance();
    ++k;
}

}  // namespace hdr

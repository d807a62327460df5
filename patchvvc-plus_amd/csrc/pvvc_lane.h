/*
 * pvvc_lane.h -- RLGR colour decoding with one lane per slice.
 *
 * The scalar decoder (rlgr_decode_wave) runs one slice per wavefront on the
 * scalar unit, and a CU issues one SIMD's scalar instructions per cycle in
 * turn, so slices compete for that unit.  This decoder runs one slice per
 * lane on the vector unit instead: 64 slices per wavefront, no scalar work
 * per symbol.  Each lane is far slower than the scalar loop, so it takes the
 * short slices of a large batch while the scalar decoder takes the long ones,
 * on the same CUs at the same time.
 *
 * Semantics are those of rlgr_decode_wave (entropy_codec.cc:207-283 with the
 * statuses of this decoder): MSB-first bits, zero bits past the end of the
 * colour bytes, k_P from 0 and k_RP from 8 (L = 4), the 32-ones escape,
 * SYMBOL_RANGE for a symbol outside 32 bits, BITSTREAM for a run exponent
 * above 30 or more bits consumed than the bytes hold.  Dequantisation as in
 * invert_raht_octree.cc:63-69: coef[i] = float(int64(Unsign2Sign(u)) * qp).
 *
 * One iteration decodes one event: in run mode the run prefix (its one-bits
 * in closed form) and the zero count, then the Golomb-Rice code of the
 * symbol that ends the run; in k = 0 mode the code alone.  Events that do
 * not fit the fast form (escapes, codes over 32 bits, long prefixes, runs
 * reaching the end) take a bit-serial form of the same state machine.
 *
 * Only non-zero coefficients are stored: the caller zeroes coef first.
 *
 * The functions are host callable too, so that the decoder can be checked
 * against the oracle on the CPU.
 */
#ifndef PVVC_LANE_H_
#define PVVC_LANE_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pvvc_hip.h"

namespace pvvc {

#define PVVC_LANE_FN __host__ __device__ __forceinline__

/* Bit source of one lane: a 64-bit window W (left aligned, zeros below its
 * nb valid bits) filled from big-endian dwords of the colour bytes, read
 * through an aligned-down base with `skip` bits before the stream's start;
 * `pre` is the next dword, loaded one refill ahead. */
struct LaneBits {
    const uint32_t* dw;  /* 4-byte aligned base at or before the first byte */
    uint32_t end;        /* bit position of the stream's end, from dw[0] bit 0 */
    uint32_t nd;         /* index of the next dword to append */
    uint32_t pre;        /* dword nd, masked, already in registers */
    uint64_t W;
    uint32_t nb;
    uint32_t skip;
};

/* Dword i of the stream view, big endian, with the bits at or past `end`
 * cleared; 0 entirely past the end.  A dword that holds the end is read byte
 * by byte, so nothing past the colour bytes is ever read. */
PVVC_LANE_FN uint32_t lane_dword(const LaneBits& b, uint32_t i) {
    const uint32_t lo = 32u * i;
    if (lo >= b.end) return 0u;
    if (lo + 32u <= b.end) return __builtin_bswap32(b.dw[i]);
    const uint8_t* p = reinterpret_cast<const uint8_t*>(b.dw + i);
    uint32_t d = 0;
    for (uint32_t j = 0; j < 4u; ++j) d = (d << 8) | (lo + 8u * j < b.end ? static_cast<uint32_t>(p[j]) : 0u);
    return d;
}

PVVC_LANE_FN void lane_init(LaneBits& b, const uint8_t* src, uint32_t nbytes) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(src);
    b.dw = reinterpret_cast<const uint32_t*>(a & ~uintptr_t(3));
    b.skip = 8u * static_cast<uint32_t>(a & 3u);
    b.end = b.skip + 8u * nbytes;
    const uint32_t d0 = lane_dword(b, 0) << b.skip;
    b.W = static_cast<uint64_t>(d0) << 32;
    b.nb = 32u - b.skip;
    b.nd = 1;
    b.pre = lane_dword(b, 1);
}

/* Appends 32 bits when at most 32 are valid, so nb >= 33 afterwards. */
PVVC_LANE_FN void lane_refill(LaneBits& b) {
    if (b.nb <= 32u) {
        b.W |= static_cast<uint64_t>(b.pre) << (32u - b.nb);
        b.nb += 32u;
        b.nd += 1u;
        b.pre = lane_dword(b, b.nd);
    }
}

PVVC_LANE_FN void lane_skip(LaneBits& b, uint32_t n) {
    b.W = n >= 64u ? 0ull : (b.W << n);
    b.nb -= n;
}

/* Read(bits) for bits <= 32 (entropy_codec.cc:186-194). */
PVVC_LANE_FN uint32_t lane_get(LaneBits& b, uint32_t n) {
    if (n == 0) return 0u;
    if (b.nb < n) lane_refill(b);
    const uint32_t v = static_cast<uint32_t>(b.W >> (64u - n));
    lane_skip(b, n);
    return v;
}

PVVC_LANE_FN uint32_t lane_ones(uint64_t w) {
    const uint64_t inv = ~w;
    return inv ? static_cast<uint32_t>(__builtin_clzll(inv)) : 64u;
}

/* GRRead (entropy_codec.cc:196-205), escape included. */
PVVC_LANE_FN uint64_t lane_gr(LaneBits& b, uint32_t kR) {
    lane_refill(b);
    const uint32_t ones = lane_ones(b.W);
    if (ones >= 32u) {
        lane_skip(b, 32u);
        return lane_get(b, 32u);
    }
    lane_skip(b, ones + 1u);
    return (static_cast<uint64_t>(ones) << kR) | lane_get(b, kR);
}

PVVC_LANE_FN uint32_t lane_adapt_krp(uint32_t kRP, uint64_t p) {
    if (p) {
        const uint32_t pp = p > 200 ? 200u : static_cast<uint32_t>(p);
        const uint32_t v = kRP + pp - 1;
        return v > 128 ? 128u : v;
    }
    return kRP < 2 ? 0u : kRP - 2;
}

PVVC_LANE_FN float lane_dequant(uint32_t u, uint32_t qp) {
    const int32_t s = static_cast<int32_t>((u >> 1) ^ (0u - (u & 1u)));
    return static_cast<float>(static_cast<long long>(s) * static_cast<long long>(qp));
}

/* S(x) = sum over i < x of 2^(i >> 1), for x <= 9 */
PVVC_LANE_FN uint32_t lane_s(uint32_t x) {
    const uint32_t h = x >> 1;
    return (x & 1u) ? 3u * (1u << h) - 2u : 2u * ((1u << h) - 1u);
}

/* Decodes total symbols into coef (zeroed by the caller); returns the status. */
PVVC_LANE_FN int rlgr_decode_lane(const uint8_t* src, uint32_t nbytes, uint32_t total, uint32_t qp,
                                                float* coef) {
    LaneBits b;
    lane_init(b, src, nbytes);
    lane_refill(b);
    uint32_t kP = 0, kRP = 8, n = 0;
    int st = PVVC_OK;
    while (n < total) {
        const bool run = kP >= 4u;
        const uint32_t kR = kRP >> 2;
        bool fast = true;
        if (run) {
            /* fast run part: at most 6 one-bits, exponent at most 24, all of
             * it in the first 32 bits of the window */
            const uint32_t t = lane_ones(b.W);
            const uint32_t k2 = (kP + 2u * t) >> 2;
            const uint32_t bits = t + 1u + k2;
            fast = t <= 6u && k2 <= 24u && bits <= 32u;
            if (fast) {
                /* the one-bits add sum over j < t of 2^((kP + 2j) >> 2)
                 * = 2^(kP >> 2) (S(c + t) - S(c)), c = (kP >> 1) & 1 */
                const uint32_t c = (kP >> 1) & 1u;
                uint32_t m = (lane_s(c + t) - lane_s(c)) << (kP >> 2);
                if (k2) m += static_cast<uint32_t>((b.W << (t + 1u)) >> (64u - k2));
                lane_skip(b, bits);
                kP += 2u * t;
                if (m >= total - n) { /* the run reaches the end */
                    n = total;
                    break;
                }
                n += m;
                lane_refill(b);
            }
        }
        if (fast) {
            const uint32_t p = lane_ones(b.W);
            const uint32_t len = p + 1u + kR;
            uint64_t u;
            if (p < 32u && len <= 32u) { /* within the 33+ valid bits */
                const uint32_t rem = kR ? static_cast<uint32_t>((b.W << (p + 1u)) >> (64u - kR)) : 0u;
                u = (static_cast<uint64_t>(p) << kR) | rem;
                lane_skip(b, len);
            } else {
                u = lane_gr(b, kR);
            }
            const uint64_t v = run ? u + 1u : u;
            if (v >= (1ull << 32)) {
                st = PVVC_ERR_SYMBOL_RANGE;
                break;
            }
            if (v) coef[n] = lane_dequant(static_cast<uint32_t>(v), qp);
            ++n;
            kRP = lane_adapt_krp(kRP, u >> kR);
            kP = run ? kP - 1u : (u ? (kP ? kP - 1u : 0u) : kP + 3u);
            lane_refill(b);
            continue;
        }
        /* ---- run part in the bit-serial form of the state machine ---- */
        uint32_t k = kP >> 2;
        uint64_t mm = 0;
        bool bad = false;
        for (;;) {
            lane_refill(b);
            const uint32_t ones = lane_ones(b.W);
            for (uint32_t i = 0; i < ones; ++i) {
                mm += 1ull << k;
                kP += 2u;
                k = kP >> 2;
                if (k > 30u) {
                    bad = true;
                    break;
                }
            }
            if (bad) break;
            if (ones < b.nb) {
                lane_skip(b, ones + 1u);
                break;
            }
            lane_skip(b, ones);
        }
        if (bad) {
            st = PVVC_ERR_BITSTREAM;
            break;
        }
        mm += lane_get(b, k);
        if (mm >= static_cast<uint64_t>(total - n)) {
            n = total;
            break;
        }
        n += static_cast<uint32_t>(mm);
        const uint64_t u = lane_gr(b, kR);
        if (u + 1 >= (1ull << 32)) {
            st = PVVC_ERR_SYMBOL_RANGE;
            break;
        }
        coef[n] = lane_dequant(static_cast<uint32_t>(u + 1), qp);
        ++n;
        kRP = lane_adapt_krp(kRP, u >> kR);
        kP = kP - 1u;
        lane_refill(b);
    }
    const int64_t consumed = static_cast<int64_t>(b.nd) * 32 - b.skip - b.nb;
    if (st == PVVC_OK && consumed > 8ll * nbytes) st = PVVC_ERR_BITSTREAM;
    return st;
}

}  // namespace pvvc

#endif

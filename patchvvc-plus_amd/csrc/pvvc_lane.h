/*
 * pvvc_lane.h -- RLGR colour decoding with one lane per slice.
 *
 * The scalar decoder (rlgr_decode_wave) runs one slice per wavefront on the
 * scalar unit, and a SIMD issues one scalar instruction every ~4.4 cycles
 * for all its wavefronts together, so slices compete for that unit.  This
 * decoder runs one slice per lane on the vector unit instead: 64 slices per
 * wavefront, a few scalar instructions per event for the whole wavefront.  A
 * lane is slower than the scalar loop (several times more vector
 * instructions per event than the loop has scalar ones), so it takes the
 * short slices of a batch while the scalar decoder takes the long ones, on
 * the same CUs at the same time.
 *
 * Semantics are those of rlgr_decode_wave (entropy_codec.cc:207-283 with the
 * statuses of this decoder): MSB-first bits, zero bits past the end of the
 * colour bytes, k_P from 0 and k_RP from 8 (L = 4), the 32-ones escape,
 * SYMBOL_RANGE for a symbol outside 32 bits, BITSTREAM for a run exponent
 * above 30 or more bits consumed than the bytes hold.  Dequantisation as in
 * invert_raht_octree.cc:63-69: coef[i] = float(int64(Unsign2Sign(u)) * qp).
 *
 * Bits come through a two-dword window d0:d1 (big endian) whose next unread
 * bit is at MSB index 32 - sh (sh in [0, 31]), so the next 32 bits are one
 * funnel shift; `q` holds the dword after d1, loaded when the window last
 * moved (one move per 32 bits consumed), so its load has had that long.
 * Dwords are read whole through the 4-byte aligned base: the aligned dword
 * that holds the last colour byte must be readable (in a batch image every
 * frame's bytes are padded to 256).
 *
 * One iteration decodes one event.  The common forms, taken by all lanes
 * together: in run mode (k_P >= 4) a run prefix of at most 6 one-bits (their
 * sum in closed form) with an exponent of at most 24, then the zero count;
 * then the Golomb-Rice code of the symbol (prefix and remainder within 32
 * bits).  Anything else (escapes, longer codes or prefixes) takes a
 * bit-serial form of the same state machine, which a wavefront only enters
 * when one of its lanes needs it.
 *
 * Zeros inside runs are not stored: the caller zeroes coef first.  The
 * functions are host callable too, so that the decoder can be checked
 * against the oracle on the CPU.
 */
#ifndef PVVC_LANE_H_
#define PVVC_LANE_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pvvc_hip.h"

namespace pvvc {

#define PVVC_LANE_FN __host__ __device__ __forceinline__

/* Stream dwords read straight from memory: dword i of the 4-byte aligned
 * span that holds the bytes (i past the end reads the last one, whose bits
 * past the end lane_fix clears); with no bytes, the aligned dword at the
 * stream's address, which must be readable. */
struct LaneMem {
    const uint32_t* dw; /* aligned base */
    uint32_t last;      /* index of the dword that holds the last bit */
    PVVC_LANE_FN uint32_t raw(uint32_t i) const {
#if defined(__HIP_DEVICE_COMPILE__)
        const auto* g = (const __attribute__((address_space(1))) uint32_t*)(dw); /* a global load */
        return g[i < last ? i : last];
#else
        return dw[i < last ? i : last];
#endif
    }
    PVVC_LANE_FN void ensure(uint32_t) {}
};

/* Stream dwords staged through LDS: a ring of 128 dwords per lane (row
 * i mod 128, column = lane) and a spare row for the writes a lane does not
 * need.  ensure(nd) tops every lane of the wavefront up to dwords
 * [nd, nd + 128) when one of them holds fewer than 16 past nd, in rounds of
 * 16 unconditional loads (indices clamped to the span) per lane, so that the
 * wavefront waits once per 16 dwords of a round rather than, at every window
 * move, on whichever lane's read missed the caches. */
constexpr uint32_t kLaneRing = 128;

/* true when pred holds in any active lane of the wavefront (on the host: a
 * single lane) */
PVVC_LANE_FN bool lane_any(bool pred) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __ballot(pred) != 0;
#else
    return pred;
#endif
}

struct LaneRing {
    const uint32_t* dw; /* global, aligned base */
    uint32_t last;      /* index of the dword that holds the last bit */
    uint32_t fill;      /* dwords staged so far */
    uint32_t* col;      /* this lane's column of the ring */
    PVVC_LANE_FN uint32_t raw(uint32_t i) const { return col[(i & (kLaneRing - 1u)) * 64u]; }
    PVVC_LANE_FN void ensure(uint32_t nd) {
        if (!lane_any(fill - nd < 16u)) return;
#if defined(__HIP_DEVICE_COMPILE__)
        const auto* g = (const __attribute__((address_space(1))) uint32_t*)(dw); /* global loads */
#else
        const uint32_t* g = dw;
#endif
        const uint32_t target = nd + kLaneRing;
        while (lane_any(fill < target)) {
            uint32_t v[16];
#pragma unroll
            for (uint32_t r = 0; r < 16u; ++r) {
                const uint32_t i = fill + r;
                v[r] = g[i < last ? i : last];
            }
            const uint32_t k = target - fill < 16u ? target - fill : 16u;
#pragma unroll
            for (uint32_t r = 0; r < 16u; ++r)
                col[(r < k ? ((fill + r) & (kLaneRing - 1u)) : kLaneRing) * 64u] = v[r];
            fill += k;
        }
    }
};

template <class Src>
struct LaneBits {
    Src src;
    uint32_t end;       /* bit position of the stream's end, from dword 0 bit 0 */
    uint32_t d0, d1;    /* window */
    uint32_t q;         /* dword nd as stored (byte order and end not applied) */
    uint32_t sh;        /* next unread bit: MSB index 32 - sh of d0:d1 */
    uint32_t nd;
};

PVVC_LANE_FN uint32_t lane_clz(uint32_t x) { return x ? static_cast<uint32_t>(__builtin_clz(x)) : 32u; }

/* (x >> off) & (2^w - 1) for w < 32 (0 when w = 0) */
PVVC_LANE_FN uint32_t lane_bfe(uint32_t x, uint32_t off, uint32_t w) { return (x >> off) & ((1u << w) - 1u); }

/* Big-endian value of stored dword `raw` at index i, with the bits at or past
 * `end` cleared (0 entirely past it). */
PVVC_LANE_FN uint32_t lane_fix(uint32_t end, uint32_t raw, uint32_t i) {
    const int32_t vb = static_cast<int32_t>(end) - static_cast<int32_t>(32u * i);
    const uint32_t mask = vb >= 32 ? 0xFFFFFFFFu : (vb <= 0 ? 0u : 0xFFFFFFFFu << (32 - vb));
    return __builtin_bswap32(raw) & mask;
}

template <class B>
PVVC_LANE_FN uint32_t lane_peek(const B& b) {
    return static_cast<uint32_t>(((static_cast<uint64_t>(b.d0) << 32) | b.d1) >> b.sh);
}

/* Consumes c <= 32 bits.  The read of the next dword is issued when the
 * window moves and first used when it moves again. */
template <class B>
PVVC_LANE_FN void lane_consume(B& b, uint32_t c) {
    const int32_t s = static_cast<int32_t>(b.sh) - static_cast<int32_t>(c);
    if (s < 0) {
        b.d0 = b.d1;
        b.d1 = lane_fix(b.end, b.q, b.nd);
        b.q = b.src.raw(++b.nd);
    }
    b.sh = static_cast<uint32_t>(s) & 31u;
}

/* Bits consumed since dword 0 bit 0. */
template <class B>
PVVC_LANE_FN uint32_t lane_pos(const B& b) { return 32u * (b.nd - 1u) - b.sh; }

/* skip: bits before the stream's first byte in dword 0 */
template <class B>
PVVC_LANE_FN void lane_init(B& b, uint32_t skip, uint32_t nbytes) {
    b.end = skip + 8u * nbytes;
    b.src.ensure(0);
    b.d0 = 0;
    b.d1 = lane_fix(b.end, b.src.raw(0), 0);
    b.q = b.src.raw(1);
    b.nd = 1;
    b.sh = 0;
    lane_consume(b, skip);
}

/* Read(bits) for bits <= 32 (entropy_codec.cc:186-194). */
template <class B>
PVVC_LANE_FN uint32_t lane_get(B& b, uint32_t n) {
    if (n == 0) return 0u;
    const uint32_t v = lane_peek(b) >> (32u - n);
    lane_consume(b, n);
    return v;
}

/* GRRead (entropy_codec.cc:196-205), escape included. */
template <class B>
PVVC_LANE_FN uint64_t lane_gr(B& b, uint32_t kR) {
    const uint32_t ones = lane_clz(~lane_peek(b));
    if (ones >= 32u) {
        lane_consume(b, 32u);
        return lane_get(b, 32u);
    }
    lane_consume(b, ones + 1u);
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

/* float(int64(Unsign2Sign(v)) * qp) for v < 2^32; products within int32
 * convert from int32 (the same rounding) */
PVVC_LANE_FN float lane_dequant(uint32_t v, uint32_t qp) {
    const int32_t s = static_cast<int32_t>((v >> 1) ^ (0u - (v & 1u)));
    const int64_t prod = static_cast<int64_t>(s) * static_cast<int64_t>(qp);
    const int32_t lo = static_cast<int32_t>(prod);
    if (static_cast<int64_t>(lo) == prod) return static_cast<float>(lo);
    return static_cast<float>(static_cast<long long>(prod));
}

/* sum over j < t of 2^((kP + 2j) >> 2), t <= 6: with e = kP >> 2 and
 * c = (kP >> 1) & 1 it is 2^e (S(c + t) - S(c)), where S(x) = sum over
 * i < x of 2^(i >> 1) = (2 + (x & 1)) 2^(x >> 1) - 2 and S(c) = c. */
PVVC_LANE_FN uint32_t lane_run_ones(uint32_t kP, uint32_t t) {
    const uint32_t c = (kP >> 1) & 1u;
    const uint32_t y = c + t;
    return (((2u + (y & 1u)) << (y >> 1)) - 2u - c) << (kP >> 2);
}

/* Decodes total symbols into coef (zeroed by the caller) from the stream
 * behind `src` (nbytes bytes from bit `skip` of its dword 0); returns the
 * status.  Each event first has src hold at least 16 dwords past the window:
 * an event reads at most 6. */
template <class Src>
PVVC_LANE_FN int rlgr_decode_lane_from(const Src& src, uint32_t skip, uint32_t nbytes, uint32_t total, uint32_t qp,
                                       float* coef) {
    LaneBits<Src> b;
    b.src = src;
    lane_init(b, skip, nbytes);
    uint32_t kP = 0, kRP = 8, n = 0;
    int st = PVVC_OK;
    while (n < total) {
        b.src.ensure(b.nd);
        const uint32_t run = kP >= 4u ? 1u : 0u;
        uint32_t x = lane_peek(b);
        /* ---- run prefix and zero count (run mode) ---- */
        const uint32_t t = lane_clz(~x);
        const uint32_t k2 = (kP + 2u * t) >> 2;
        const uint32_t b1 = t + 1u + k2;
        const bool frun = t <= 6u && k2 <= 24u && b1 <= 32u;
        if (run && frun) {
            const uint32_t m = lane_run_ones(kP, t) + lane_bfe(x, 32u - b1, k2);
            lane_consume(b, b1);
            kP += 2u * t;
            if (m >= total - n) { /* the run reaches the end */
                n = total;
                break;
            }
            n += m;
            x = lane_peek(b);
        }
        /* ---- Golomb-Rice code (the prefix p is u >> kR here) ---- */
        const uint32_t kR = kRP >> 2;
        const uint32_t p = lane_clz(~x);
        const uint32_t len = p + 1u + kR;
        if ((!run || frun) && len <= 32u) {
            const uint32_t u = (p << kR) | lane_bfe(x, 32u - len, kR);
            lane_consume(b, len);
            coef[n] = lane_dequant(u + run, qp);
            ++n;
            /* k_RP + p - 1 stays below 128: p + kR <= 31 and kRP <= 4 kR + 3 */
            kRP = p ? kRP + p - 1u : (kRP < 2u ? 0u : kRP - 2u);
            kP = run ? kP - 1u : (u ? (kP ? kP - 1u : 0u) : kP + 3u);
            continue;
        }
        /* ---- the same event in bit-serial form ---- */
        if (!run) {
            const uint64_t u = lane_gr(b, kR);
            if (u >= (1ull << 32)) {
                st = PVVC_ERR_SYMBOL_RANGE;
                break;
            }
            if (u) coef[n] = lane_dequant(static_cast<uint32_t>(u), qp);
            ++n;
            kRP = lane_adapt_krp(kRP, u >> kR);
            kP = u ? (kP ? kP - 1u : 0u) : kP + 3u;
            continue;
        }
        if (!frun) { /* the run prefix one bit at a time */
            uint32_t k = kP >> 2;
            uint64_t mm = 0;
            bool bad = false;
            for (;;) {
                const uint32_t ones = lane_clz(~lane_peek(b));
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
                if (ones < 32u) {
                    lane_consume(b, ones + 1u);
                    break;
                }
                lane_consume(b, 32u);
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
        }
        const uint64_t u = lane_gr(b, kR);
        if (u + 1 >= (1ull << 32)) {
            st = PVVC_ERR_SYMBOL_RANGE;
            break;
        }
        coef[n] = lane_dequant(static_cast<uint32_t>(u + 1), qp);
        ++n;
        kRP = lane_adapt_krp(kRP, u >> kR);
        kP = kP - 1u;
    }
    if (st == PVVC_OK && lane_pos(b) - skip > 8u * nbytes) st = PVVC_ERR_BITSTREAM;
    return st;
}

/* The decoder over bytes in memory; src must be readable (its aligned dword)
 * even when nbytes is 0. */
PVVC_LANE_FN int rlgr_decode_lane(const uint8_t* src, uint32_t nbytes, uint32_t total, uint32_t qp, float* coef) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(src);
    const uint32_t skip = 8u * static_cast<uint32_t>(a & 3u);
    const uint32_t end = skip + 8u * nbytes;
    const LaneMem m{reinterpret_cast<const uint32_t*>(a & ~uintptr_t(3)), end ? (end - 1u) / 32u : 0u};
    return rlgr_decode_lane_from(m, skip, nbytes, total, qp, coef);
}

}  // namespace pvvc

#endif

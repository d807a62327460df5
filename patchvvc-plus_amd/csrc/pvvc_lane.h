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
 * Bits come through a two-dword window d0:d1 whose next unread bit is at MSB
 * index 32 - sh (sh in [0, 31]), so the next 32 bits are one funnel shift.
 * The two dwords after it are held in q and q2.  A consume moves the window
 * by at most one dword, by selects (no branch), and reads the dword after q2
 * every time, so that the read is used one consume later.  The dwords come
 * from a source (LaneMem: straight from memory; LaneRing: staged through
 * LDS) already in big-endian order with the bits past the end cleared.
 * Dwords are read whole through the 4-byte aligned base: the aligned dword
 * that holds the last colour byte must be readable (in a batch image every
 * frame's bytes are padded to 256).
 *
 * One iteration decodes one event in every lane that is still decoding,
 * with no branch that depends on the lane: a lane that is finished, or whose
 * event takes another form, consumes nothing and keeps its state.  The
 * common forms: in run mode (k_P >= 4) a run prefix of at most 6 one-bits
 * (their sum in closed form) with an exponent of at most 24, then the zero
 * count; then the Golomb-Rice code of the symbol (prefix and remainder within
 * 32 bits).  Anything else (escapes, longer codes or prefixes) takes a
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

/* true when pred holds in any active lane of the wavefront (on the host: a
 * single lane) */
PVVC_LANE_FN bool lane_any(bool pred) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __ballot(pred) != 0;
#else
    return pred;
#endif
}

/* Stored dword i of the 4-byte aligned span that holds the bytes (i past
 * the end reads the last one; with no bytes, the aligned dword at the
 * stream's address, which must be readable). */
PVVC_LANE_FN uint32_t lane_load(const uint32_t* dw, uint32_t last, uint32_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
    const auto* g = (const __attribute__((address_space(1))) uint32_t*)(dw); /* a global load */
    return g[i < last ? i : last];
#else
    return dw[i < last ? i : last];
#endif
}

/* Stream dwords straight from memory. */
struct LaneMem {
    const uint32_t* dw; /* aligned base */
    uint32_t end;       /* bit position of the stream's end, from dword 0 bit 0 */
    uint32_t last;      /* index of the dword that holds the last bit */
    PVVC_LANE_FN uint32_t raw(uint32_t i) const { return lane_fix(end, lane_load(dw, last, i), i); }
    PVVC_LANE_FN void ensure(uint32_t) {}
};

/* Stream dwords staged through LDS: a ring of 128 dwords per lane (row
 * i mod 128, column = lane) and a spare row for the writes a lane does not
 * need.  ensure(j) tops every lane of the wavefront up to dwords
 * [j, j + 128) when one of them holds fewer than 12 past j (the window,
 * q, q2 and the 6 dwords at most that one event reads), in rounds of 16
 * unconditional loads (indices clamped to the span) per lane, so that the
 * wavefront waits once per round rather than, at every window move, on
 * whichever lane's read missed the caches.  The ring holds dwords in
 * big-endian order with the bits past the end cleared. */
constexpr uint32_t kLaneRing = 128;
constexpr uint32_t kLaneLow = 12;

struct LaneRing {
    const uint32_t* dw; /* global, aligned base */
    uint32_t end;
    uint32_t last;
    uint32_t fill;      /* dwords staged so far */
    uint32_t* col;      /* this lane's column of the ring */
    PVVC_LANE_FN uint32_t raw(uint32_t i) const { return col[(i & (kLaneRing - 1u)) * 64u]; }
    PVVC_LANE_FN void ensure(uint32_t j) {
        if (!lane_any(fill - j < kLaneLow)) return;
        const uint32_t target = j + kLaneRing;
        while (lane_any(fill < target)) {
            uint32_t v[16];
#pragma unroll
            for (uint32_t r = 0; r < 16u; ++r) v[r] = lane_load(dw, last, fill + r);
            const uint32_t k = target - fill < 16u ? target - fill : 16u;
#pragma unroll
            for (uint32_t r = 0; r < 16u; ++r)
                col[(r < k ? ((fill + r) & (kLaneRing - 1u)) : kLaneRing) * 64u] = lane_fix(end, v[r], fill + r);
            fill += k;
        }
    }
};

/* Bit reader over a source: d1 holds dword J, d0 dword J - 1, q dword
 * J + 1, q2 dword J + 2. */
template <class Src>
struct LaneBits {
    Src src;
    uint32_t d0, d1, q, q2;
    uint32_t sh; /* next unread bit: MSB index 32 - sh of d0:d1 */
    uint32_t J;
};

template <class B>
PVVC_LANE_FN uint32_t lane_peek(const B& b) {
    return static_cast<uint32_t>(((static_cast<uint64_t>(b.d0) << 32) | b.d1) >> b.sh);
}

/* Consumes c <= 32 bits, without a branch. */
template <class B>
PVVC_LANE_FN void lane_consume(B& b, uint32_t c) {
    const int32_t s = static_cast<int32_t>(b.sh) - static_cast<int32_t>(c);
    const bool adv = s < 0;
    b.d0 = adv ? b.d1 : b.d0;
    b.d1 = adv ? b.q : b.d1;
    b.q = adv ? b.q2 : b.q;
    b.J += adv ? 1u : 0u;
    b.q2 = b.src.raw(b.J + 2u); /* first used at the next consume */
    b.sh = static_cast<uint32_t>(s) & 31u;
}

/* Bits consumed since dword 0 bit 0. */
template <class B>
PVVC_LANE_FN uint32_t lane_pos(const B& b) { return 32u * b.J - b.sh; }

/* skip: bits before the stream's first byte in dword 0 */
template <class B>
PVVC_LANE_FN void lane_init(B& b, uint32_t skip) {
    b.src.ensure(0);
    b.d0 = 0;
    b.d1 = b.src.raw(0);
    b.q = b.src.raw(1);
    b.q2 = b.src.raw(2);
    b.J = 0;
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

/* float(int64(Unsign2Sign(v)) * qp) for v < 2^32 and qp < 256: the product
 * is below 2^40 in magnitude, so it is exact in FP64, and one FP64 -> FP32
 * conversion rounds it as the int64 -> FP32 conversion does (to nearest,
 * ties to even) */
PVVC_LANE_FN float lane_dequant(uint32_t v, uint32_t qp) {
    const int32_t s = static_cast<int32_t>((v >> 1) ^ (0u - (v & 1u)));
    return static_cast<float>(static_cast<double>(s) * static_cast<double>(qp));
}

/* sum over j < t of 2^((kP + 2j) >> 2), t <= 6: with e = kP >> 2 and
 * c = (kP >> 1) & 1 it is 2^e (S(c + t) - S(c)), where S(x) = sum over
 * i < x of 2^(i >> 1) = (2 + (x & 1)) 2^(x >> 1) - 2 and S(c) = c. */
PVVC_LANE_FN uint32_t lane_run_ones(uint32_t kP, uint32_t t) {
    const uint32_t c = (kP >> 1) & 1u;
    const uint32_t y = c + t;
    return (((2u + (y & 1u)) << (y >> 1)) - 2u - c) << ((kP >> 2) & 31u); /* the mask only matters off the fast path */
}

/* Decodes total symbols into coef (zeroed by the caller) from the stream
 * behind `src` (nbytes bytes from bit `skip` of its dword 0); returns the
 * status. */
template <class Src>
PVVC_LANE_FN int rlgr_decode_lane_from(const Src& src, uint32_t skip, uint32_t nbytes, uint32_t total, uint32_t qp,
                                       float* coef) {
    LaneBits<Src> b;
    b.src = src;
    lane_init(b, skip);
    uint32_t kP = 0, kRP = 8, n = 0;
    int st = PVVC_OK;
    bool act = n < total;
    while (lane_any(act)) {
        b.src.ensure(b.J);
        const bool run = kP >= 4u;
        /* ---- run prefix and zero count (run mode) ---- */
        const uint32_t x = lane_peek(b);
        const uint32_t t = lane_clz(~x);
        const uint32_t k2 = (kP + 2u * t) >> 2;
        const uint32_t b1 = t + 1u + k2;
        const bool frun = t <= 6u && k2 <= 24u && b1 <= 32u;
        const bool rp = act && run && frun;
        const uint32_t m = lane_run_ones(kP, t) + lane_bfe(x, 32u - (b1 <= 32u ? b1 : 32u), k2 <= 24u ? k2 : 0u);
        lane_consume(b, rp ? b1 : 0u);
        kP += rp ? 2u * t : 0u;
        const bool ends = rp && m >= total - n; /* the run reaches the end */
        n = rp ? (ends ? total : n + m) : n;
        const bool act2 = act && !ends;
        /* ---- Golomb-Rice code (the prefix p is u >> kR here) ---- */
        const uint32_t x2 = lane_peek(b);
        const uint32_t kR = kRP >> 2;
        const uint32_t p = lane_clz(~x2);
        const uint32_t len = p + 1u + kR;
        const bool fast = act2 && (!run || frun) && len <= 32u;
        const uint32_t u = (p << (kR & 31u)) | lane_bfe(x2, 32u - (len <= 32u ? len : 32u), kR < 32u ? kR : 0u);
        lane_consume(b, fast ? len : 0u);
        if (fast) coef[n] = lane_dequant(u + (run ? 1u : 0u), qp);
        n += fast ? 1u : 0u;
        /* k_RP + p - 1 stays below 128: p + kR <= 31 and kRP <= 4 kR + 3 */
        const uint32_t krp_f = p ? kRP + p - 1u : (kRP < 2u ? 0u : kRP - 2u);
        const uint32_t kp_f = run ? kP - 1u : (u ? (kP ? kP - 1u : 0u) : kP + 3u);
        kRP = fast ? krp_f : kRP;
        kP = fast ? kp_f : kP;
        /* ---- the same event in bit-serial form ---- */
        const bool slow = act2 && !fast;
        if (lane_any(slow) && slow) {
            do {
                if (!run) {
                    const uint64_t us = lane_gr(b, kR);
                    if (us >= (1ull << 32)) {
                        st = PVVC_ERR_SYMBOL_RANGE;
                        break;
                    }
                    if (us) coef[n] = lane_dequant(static_cast<uint32_t>(us), qp);
                    ++n;
                    kRP = lane_adapt_krp(kRP, us >> kR);
                    kP = us ? (kP ? kP - 1u : 0u) : kP + 3u;
                    break;
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
                const uint64_t us = lane_gr(b, kR);
                if (us + 1 >= (1ull << 32)) {
                    st = PVVC_ERR_SYMBOL_RANGE;
                    break;
                }
                coef[n] = lane_dequant(static_cast<uint32_t>(us + 1), qp);
                ++n;
                kRP = lane_adapt_krp(kRP, us >> kR);
                kP = kP - 1u;
            } while (false);
        }
        act = st == PVVC_OK && n < total;
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
    const LaneMem m{reinterpret_cast<const uint32_t*>(a & ~uintptr_t(3)), end, end ? (end - 1u) / 32u : 0u};
    return rlgr_decode_lane_from(m, skip, nbytes, total, qp, coef);
}

}  // namespace pvvc

#endif

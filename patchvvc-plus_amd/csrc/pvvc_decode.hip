/*
 * pvvc_decode.hip -- HIP kernels and C ABI of the per-patch GoP decoder.
 *
 * Device pipeline per slice: see pvvc_kernels.h.  Host API: include/pvvc_hip.h.
 */
#include "pvvc_internal.h"
#include "pvvc_kernels.h"
#include "pvvc_lane.h"
#include "pvvc_runtime.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <thread>
#include <tuple>
#include <vector>

namespace pvvc {

/* ======================================================================== */
/* Device: RLGR decoder (src/common/entropy_codec.cc:163-283)               */
/* ======================================================================== */

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

/* Big-endian dword at byte `off` of the colour stream; bytes past the end read 0. */
__device__ __forceinline__ uint32_t fetch_be(const uint8_t* __restrict__ src, uint32_t nbytes, uint32_t off,
                                             uint32_t aligned) {
    if (aligned && off + 4u <= nbytes) return __builtin_bswap32(*reinterpret_cast<const uint32_t*>(src + off));
    uint32_t v = 0;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) v = (v << 8) | ((off + i) < nbytes ? static_cast<uint32_t>(src[off + i]) : 0u);
    return v;
}

/* Wave-uniform MSB-first bit reader.  `cur`/`nxt` are per-lane: dword `lane`
 * of the current and the next 256-byte chunk. */
struct BitReader {
    const uint8_t* src;
    uint32_t nbytes;
    uint32_t aligned;
    uint32_t cur, nxt;
    uint32_t chunk;
    uint32_t wpos;   /* next dword index to append to the window */
    uint64_t win;    /* left-aligned window, nb valid bits, zero below */
    int32_t nb;
};

__device__ __forceinline__ uint32_t br_chunk(const BitReader& br, uint32_t chunk) {
    return fetch_be(br.src, br.nbytes, (chunk * 64u + lane_id()) * 4u, br.aligned);
}

__device__ __forceinline__ void br_init(BitReader& br, const uint8_t* src, uint32_t nbytes) {
    br.src = src;
    br.nbytes = nbytes;
    br.aligned = (reinterpret_cast<uintptr_t>(src) & 3u) == 0;
    br.cur = br_chunk(br, 0);
    br.nxt = br_chunk(br, 1);
    br.chunk = 0;
    br.wpos = 0;
    br.win = 0;
    br.nb = 0;
}

/* Makes `cur` the 64-dword chunk that holds dword wpos. */
__device__ __forceinline__ void br_sync(BitReader& br) {
    if ((br.wpos >> 6) != br.chunk) {
        br.cur = br.nxt;
        br.chunk += 1;
        br.nxt = br_chunk(br, br.chunk + 1);
    }
}

/* Appends 32 bits when at most 32 are left, so nb > 32 afterwards. */
__device__ __forceinline__ void br_refill(BitReader& br) {
    if (br.nb <= 32) {
        const uint32_t l = br.wpos & 63u;
        br_sync(br);
        const uint32_t w = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(br.cur), static_cast<int>(l)));
        br.win |= static_cast<uint64_t>(w) << (32 - br.nb);
        br.nb += 32;
        br.wpos += 1;
    }
}

__device__ __forceinline__ void br_skip(BitReader& br, uint32_t b) {
    br.win = b >= 64 ? 0ull : (br.win << b);
    br.nb -= static_cast<int32_t>(b);
}

/* Read(bits) for 0 <= bits <= 32 (entropy_codec.cc:186-194). */
__device__ __forceinline__ uint32_t br_get(BitReader& br, uint32_t b) {
    if (b == 0) return 0;
    if (br.nb < static_cast<int32_t>(b)) br_refill(br);
    const uint32_t v = static_cast<uint32_t>(br.win >> (64 - b));
    br_skip(br, b);
    return v;
}

__device__ __forceinline__ uint32_t leading_ones(uint64_t w) {
    const uint64_t inv = ~w;
    return inv ? static_cast<uint32_t>(__builtin_clzll(inv)) : 64u;
}

/* GRRead (entropy_codec.cc:196-205): unary prefix counted with one
 * leading-ones scan; 32 ones is the escape to a raw 32-bit value. */
__device__ __forceinline__ uint64_t br_gr(BitReader& br, uint32_t kR) {
    br_refill(br);
    const uint32_t ones = leading_ones(br.win);
    if (ones >= 32) {
        br_skip(br, 32);
        return br_get(br, 32);
    }
    br_skip(br, ones + 1);
    return (static_cast<uint64_t>(ones) << kR) | br_get(br, kR);
}

/* k_RP adaptation (entropy_codec.cc:240-248, 259-267). */
__device__ __forceinline__ uint32_t adapt_krp(uint32_t kRP, uint64_t p) {
    if (p) {
        const uint32_t pp = p > 200 ? 200u : static_cast<uint32_t>(p);
        const uint32_t v = kRP + pp - 1;
        return v > 128 ? 128u : v;
    }
    return kRP < 2 ? 0u : kRP - 2;
}

/* Unsign2Sign (entropy_codec.h:73-81) for u < 2^32. */
__device__ __forceinline__ int32_t unsign2sign(uint64_t u) {
    const uint32_t h = static_cast<uint32_t>(u >> 1);
    return static_cast<int32_t>(h ^ (0u - static_cast<uint32_t>(u & 1u)));
}

/* Common-case RLGR symbols as a tight loop on the scalar unit
 * (entropy_codec.cc:207-283).  State in SGPRs: the 64-bit window `win`
 * (left aligned, zero below its nb valid bits; s[94:95] inside the loop, so
 * that its top dword s95 can be named) with nn = 32 - nb, the next dword
 * `wpos` of the chunk held one dword per lane in `cur`, k_RP, k_P and
 * neg = nrel - lim (mod 2^32; its carry out marks the end of the 64-symbol
 * chunk).  Inside the loop neg lives in M0, the lane select v_writelane
 * takes beside a scalar data operand.  Per symbol:
 *  - k = 0: Golomb-Rice code (unary prefix by one leading-ones count, code
 *    length = prefix + 1 + k_R), k_P / k_RP adaptation (:255-276);
 *  - k > 0: up to 6 one-bits of the run prefix (each adds 2^k, k_P += 2),
 *    the run length in k bits, the zero run (only advances the symbol index),
 *    then the Golomb-Rice coded symbol + 1 and the adaptation of :240-251;
 *  - the scalar unit only finds where each code ends and adapts the state:
 *    it does not assemble the symbol's value.  v_writelane into lane
 *    (nrel - lim) mod 64 (selected by the low bits of `neg`) keeps what the
 *    value is made from: aT = the 32 window bits that start with the
 *    Golomb-Rice code, aK = k_R, and aF = 1 after a zero run (that symbol is
 *    stored as u + 1).  When the chunk is stored, every lane decodes its
 *    own value at once in vector code (lane_symbol: leading ones, a bit-field
 *    extract), then Unsign2Sign and the dequantisation.  The writes are
 *    vector instructions, which issue beside the scalar ones; they replace
 *    five scalar instructions per symbol (two shifts, the OR of quotient and
 *    remainder, the remainder's shift count, the + 1), at the price of one
 *    compare for the p != 0 carry of the k_RP step.  A k = 0 symbol is 0
 *    when the top k_R + 1 bits of the window are all zero; that test (one
 *    shift) chooses the zero block;
 *  - the window takes 32 more bits when fewer than 33 remain.  The code's
 *    length is added to nn as soon as it is known: no carry (the common
 *    case) means the code lies in the window's first 32 bits and no refill
 *    follows, so that one test stands for both; a carry leaves the line for
 *    a block that tests for a code over 32 bits (an exit, nn restored), then
 *    finishes the symbol and refills.  After a carry nn is 32 - nb of the
 *    shorter window, the shift that places the next dword below it.
 * k_P is not held in a register: it is the position in the code.  One
 * block per k_P value decodes one symbol with k = k_P / 4 and the shift
 * counts of its run entry as constants: G (k_P >= 12, the only one that
 * keeps k_P in its register), R11..R8 (k = 2), R7..R4 (k = 1), S3..S1 and
 * four copies of S0 (k = 0), laid out in that order.  A symbol that lowers
 * k_P by one (a run symbol; a non-zero k = 0 symbol, floor 0) falls through
 * into the next block, so the common path takes no branch: a taken scalar
 * branch costs a wavefront ~30 cycles against ~4 for an instruction, and a
 * single slice's decode (the drop-in launch) is bound by that latency; S0d
 * jumps back to S0a.  A zero k = 0 symbol (u = 0) leaves the line for its
 * state's zero block (560-563), which shifts the window, lowers k_RP and
 * jumps to the block of k_P + 3 (S3, or R4..R6: run mode); its lanes decode
 * to 0.  A run block reads the run length with an arithmetic shift of the
 * window: when the prefix starts with a one-bit, that value is negative and
 * its add to neg always carries, so the one branch for a zero run past the
 * chunk end also catches prefixes with one-bits; the block it leads to
 * (680-691) undoes the add and tells the two apart by the sign word s91.
 * A run prefix of exactly one one-bit (82 % of the prefixes with
 * one-bits on c1 streams) has its own block per run state (734-741): the
 * one-bit stands for 2^k zeros, k_P + 2 fixes the width k' = (k_P + 2) / 4
 * of the length that follows at a known place (one s_bfe), and the event
 * ends in state k_P + 1, a known block.  Longer prefixes take the general
 * path (36), which ends with k_P in its register and enters its block
 * through the dispatch (91) that also starts the loop.  Every exit first
 * writes k_P (the stubs 600-723).  Per symbol this saves the k_P step and
 * its mode test, on the run path the exponent and two shift counts, and
 * most of the general prefix path: 14.2 -> 11.2 scalar instructions per
 * symbol on c1 colour streams.  k_RP's step (p ? p - 1 : -2,
 * clamped to [0, 128]) is kRP + p + [p != 0] - 2, the bracket taken as the
 * carry-in of s_addc from the SCC of s_cmp_lg p, 0.  The clamp to 128 never binds here: a code
 * is decoded only when p + 1 + k_R <= 32 (run paths: p + k_R <= 29), and
 * k_RP <= 4 k_R + 3, so the new k_RP is at most 126 - 3p.  Nor does the
 * clamp to 0 on a non-zero k = 0 symbol: p = 0 with u != 0 needs k_R >= 1,
 * i.e. k_RP >= 4.  Only the zero symbols and the run paths (where u = 0 is a
 * valid symbol, 1) keep it.  A run entry whose prefix has no one-bits
 * (the common case) is the run block itself: its zero count is the k-bit
 * suffix alone and `neg` advances in place;
 * prefixes of two or more one-bits take the one general path, laid out as
 * entry (36), one-bits loop (31/35) and symbol (32) in a line.  The
 * leading-ones count of a Golomb code is s_flbit of s_orn2(1, window): that
 * value is never 0, so the count stays below 64 and s_orn2 sets SCC, which
 * s_flbit leaves for the s_addc that adds the code's terminating zero.  A
 * run's zero count is read straight from the window (its top bit is the
 * prefix's terminating zero).  Symbols that end in a refill finish out of
 * line, one block per state copy (block + 400, 580-583, 38).
 * A code is decoded here only when it lies entirely in the first 32 bits of
 * the window; every exit leaves the state of the last whole symbol (the
 * early nn and m0 updates are undone on the way out: 19, 97, 98 and 722; a
 * chunk rotation also takes back the wpos step, 7), except a zero run of a
 * run block or a one-bit prefix block that reaches the chunk end: that exit
 * (754-761, 774-781) keeps the run's prefix and length consumed, leaves in
 * neg bit 30 and the count of its zeros past the chunk end (other exits
 * leave 0 or a negative count there), and leaves only the run's symbol to
 * the caller.  The loop returns
 * to the caller, at a symbol boundary, on: chunk complete, a chunk rotation
 * of `cur`, a unary prefix of 32+ or a code longer than 32 bits, more than 6
 * one-bits in a run prefix or k > 24 after them, or a zero run reaching the
 * chunk end.  s80-s97 are scratch.  An SGPR written by a scalar
 * instruction and then read as a v_readlane / v_writelane lane select gets
 * 4 wait states (here at least four instructions on every path;
 * tests/test_rlgr_asm.py checks the distances on the loop text). */
__device__ __forceinline__ void rlgr_fast_run(uint64_t& win, int32_t& nn, uint32_t& wpos, uint32_t& kRP,
                                               uint32_t& kP, uint32_t& neg, uint32_t cur, uint32_t pre, uint32_t& aT,
                                               uint32_t& aK, uint32_t& aF) {
    asm volatile(
        "s_mov_b64 s[94:95], %[win]\n\t"
        "s_mov_b32 s96, %[pre]\n\t"
        "s_mov_b32 s97, 0\n\t"
        "s_mov_b32 m0, %[neg]\n\t"
        "s_cmp_lt_u32 %[kP], 4\n\t"
        "s_cbranch_scc1 93f\n\t"
        "91:\n\t"
        "s_cmp_lt_u32 %[kP], 8\n\t"
        "s_cbranch_scc0 92f\n\t"
        "s_cmp_eq_u32 %[kP], 5\n\t"
        "s_cbranch_scc1 105f\n\t"
        "s_cmp_eq_u32 %[kP], 6\n\t"
        "s_cbranch_scc1 106f\n\t"
        "s_cmp_eq_u32 %[kP], 7\n\t"
        "s_cbranch_scc1 107f\n\t"
        "s_branch 104f\n\t"
        "92:\n\t"
        "s_cmp_lt_u32 %[kP], 12\n\t"
        "s_cbranch_scc0 100f\n\t"
        "s_cmp_eq_u32 %[kP], 8\n\t"
        "s_cbranch_scc1 108f\n\t"
        "s_cmp_eq_u32 %[kP], 9\n\t"
        "s_cbranch_scc1 109f\n\t"
        "s_cmp_eq_u32 %[kP], 10\n\t"
        "s_cbranch_scc1 110f\n\t"
        "s_branch 111f\n\t"
        "93:\n\t"
        "s_cmp_eq_u32 %[kP], 0\n\t"
        "s_cbranch_scc1 120f\n\t"
        "s_cmp_eq_u32 %[kP], 1\n\t"
        "s_cbranch_scc1 121f\n\t"
        "s_cmp_eq_u32 %[kP], 2\n\t"
        "s_cbranch_scc1 122f\n\t"
        "s_branch 123f\n\t"
        "100:\n\t"
        /* ---- G: k >= 3, k_P in its register ---- */
        "s_bitcmp1_b64 s[94:95], 63\n\t"
        "s_cbranch_scc1 723f\n\t"
        "s_lshr_b32 s85, %[kP], 2\n\t"
        "s_sub_u32 s88, 63, s85\n\t"
        "s_lshr_b64 s[90:91], s[94:95], s88\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_cbranch_scc1 722f\n\t"
        "s_add_u32 s86, s85, 1\n\t"
        "s_lshl_b64 s[92:93], s[94:95], s86\n\t"
        "s_lshr_b32 s89, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[92:93]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s86, s86, s84\n\t"
        "s_add_u32 s86, s86, s89\n\t"
        "s_add_u32 %[nn], %[nn], s86\n\t"
        "v_writelane_b32 %[aT], s93, m0\n\t"
        "s_cbranch_scc1 500f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "300:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 720f\n\t"
        "s_sub_u32 %[kP], %[kP], 1\n\t"
        "s_cmp_lt_u32 %[kP], 12\n\t"
        "s_cbranch_scc0 100b\n\t"
        "111:\n\t"
        /* ---- R11: k_P = 11, run entry without one-bits ---- */
        "s_ashr_i64 s[90:91], s[94:95], 61\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_cbranch_scc1 691f\n\t"
        "s_lshl_b64 s[92:93], s[94:95], 3\n\t"
        "s_lshr_b32 s89, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[92:93]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s86, s84, 3\n\t"
        "s_add_u32 s86, s86, s89\n\t"
        "s_add_u32 %[nn], %[nn], s86\n\t"
        "v_writelane_b32 %[aT], s93, m0\n\t"
        "s_cbranch_scc1 511f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "311:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 610f\n\t"
        "110:\n\t"
        /* ---- R10: k_P = 10, run entry without one-bits ---- */
        "s_ashr_i64 s[90:91], s[94:95], 61\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_cbranch_scc1 690f\n\t"
        "s_lshl_b64 s[92:93], s[94:95], 3\n\t"
        "s_lshr_b32 s89, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[92:93]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s86, s84, 3\n\t"
        "s_add_u32 s86, s86, s89\n\t"
        "s_add_u32 %[nn], %[nn], s86\n\t"
        "v_writelane_b32 %[aT], s93, m0\n\t"
        "s_cbranch_scc1 510f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "310:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 609f\n\t"
        "109:\n\t"
        /* ---- R9: k_P = 9, run entry without one-bits ---- */
        "s_ashr_i64 s[90:91], s[94:95], 61\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_cbranch_scc1 689f\n\t"
        "s_lshl_b64 s[92:93], s[94:95], 3\n\t"
        "s_lshr_b32 s89, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[92:93]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s86, s84, 3\n\t"
        "s_add_u32 s86, s86, s89\n\t"
        "s_add_u32 %[nn], %[nn], s86\n\t"
        "v_writelane_b32 %[aT], s93, m0\n\t"
        "s_cbranch_scc1 509f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "309:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 608f\n\t"
        "108:\n\t"
        /* ---- R8: k_P = 8, run entry without one-bits ---- */
        "s_ashr_i64 s[90:91], s[94:95], 61\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_cbranch_scc1 688f\n\t"
        "s_lshl_b64 s[92:93], s[94:95], 3\n\t"
        "s_lshr_b32 s89, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[92:93]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s86, s84, 3\n\t"
        "s_add_u32 s86, s86, s89\n\t"
        "s_add_u32 %[nn], %[nn], s86\n\t"
        "v_writelane_b32 %[aT], s93, m0\n\t"
        "s_cbranch_scc1 508f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "308:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 607f\n\t"
        "107:\n\t"
        /* ---- R7: k_P = 7, run entry without one-bits ---- */
        "s_ashr_i64 s[90:91], s[94:95], 62\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_cbranch_scc1 687f\n\t"
        "s_lshl_b64 s[92:93], s[94:95], 2\n\t"
        "s_lshr_b32 s89, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[92:93]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s86, s84, 2\n\t"
        "s_add_u32 s86, s86, s89\n\t"
        "s_add_u32 %[nn], %[nn], s86\n\t"
        "v_writelane_b32 %[aT], s93, m0\n\t"
        "s_cbranch_scc1 507f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "307:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 606f\n\t"
        "106:\n\t"
        /* ---- R6: k_P = 6, run entry without one-bits ---- */
        "s_ashr_i64 s[90:91], s[94:95], 62\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_cbranch_scc1 686f\n\t"
        "s_lshl_b64 s[92:93], s[94:95], 2\n\t"
        "s_lshr_b32 s89, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[92:93]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s86, s84, 2\n\t"
        "s_add_u32 s86, s86, s89\n\t"
        "s_add_u32 %[nn], %[nn], s86\n\t"
        "v_writelane_b32 %[aT], s93, m0\n\t"
        "s_cbranch_scc1 506f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "306:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 605f\n\t"
        "105:\n\t"
        /* ---- R5: k_P = 5, run entry without one-bits ---- */
        "s_ashr_i64 s[90:91], s[94:95], 62\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_cbranch_scc1 685f\n\t"
        "s_lshl_b64 s[92:93], s[94:95], 2\n\t"
        "s_lshr_b32 s89, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[92:93]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s86, s84, 2\n\t"
        "s_add_u32 s86, s86, s89\n\t"
        "s_add_u32 %[nn], %[nn], s86\n\t"
        "v_writelane_b32 %[aT], s93, m0\n\t"
        "s_cbranch_scc1 505f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "305:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 604f\n\t"
        "104:\n\t"
        /* ---- R4: k_P = 4, run entry without one-bits ---- */
        "s_ashr_i64 s[90:91], s[94:95], 62\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_cbranch_scc1 684f\n\t"
        "s_lshl_b64 s[92:93], s[94:95], 2\n\t"
        "s_lshr_b32 s89, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[92:93]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s86, s84, 2\n\t"
        "s_add_u32 s86, s86, s89\n\t"
        "s_add_u32 %[nn], %[nn], s86\n\t"
        "v_writelane_b32 %[aT], s93, m0\n\t"
        "s_cbranch_scc1 504f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "304:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 603f\n\t"
        "123:\n\t"
        /* ---- S3: k_P = 3, k = 0 symbol ---- */
        "s_lshr_b32 s85, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[94:95]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s87, s84, s85\n\t"
        "s_add_u32 %[nn], %[nn], s87\n\t"
        "v_writelane_b32 %[aT], s95, m0\n\t"
        "s_cbranch_scc1 523f\n\t"
        "s_sub_u32 s88, 63, s85\n\t"
        "s_lshr_b64 s[82:83], s[94:95], s88\n\t"
        "v_writelane_b32 %[aK], s85, m0\n\t"
        "s_cbranch_scc0 563f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "s_addc_u32 s90, %[kRP], s84\n\t"
        "s_sub_i32 %[kRP], s90, 2\n\t"
        "323:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 602f\n\t"
        "122:\n\t"
        /* ---- S2: k_P = 2, k = 0 symbol ---- */
        "s_lshr_b32 s85, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[94:95]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s87, s84, s85\n\t"
        "s_add_u32 %[nn], %[nn], s87\n\t"
        "v_writelane_b32 %[aT], s95, m0\n\t"
        "s_cbranch_scc1 522f\n\t"
        "s_sub_u32 s88, 63, s85\n\t"
        "s_lshr_b64 s[82:83], s[94:95], s88\n\t"
        "v_writelane_b32 %[aK], s85, m0\n\t"
        "s_cbranch_scc0 562f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "s_addc_u32 s90, %[kRP], s84\n\t"
        "s_sub_i32 %[kRP], s90, 2\n\t"
        "322:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 601f\n\t"
        "121:\n\t"
        /* ---- S1: k_P = 1, k = 0 symbol ---- */
        "s_lshr_b32 s85, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[94:95]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s87, s84, s85\n\t"
        "s_add_u32 %[nn], %[nn], s87\n\t"
        "v_writelane_b32 %[aT], s95, m0\n\t"
        "s_cbranch_scc1 521f\n\t"
        "s_sub_u32 s88, 63, s85\n\t"
        "s_lshr_b64 s[82:83], s[94:95], s88\n\t"
        "v_writelane_b32 %[aK], s85, m0\n\t"
        "s_cbranch_scc0 561f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "s_addc_u32 s90, %[kRP], s84\n\t"
        "s_sub_i32 %[kRP], s90, 2\n\t"
        "321:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 600f\n\t"
        "120:\n\t"
        /* ---- S0: k_P = 0, k = 0 symbol ---- */
        "s_lshr_b32 s85, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[94:95]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s87, s84, s85\n\t"
        "s_add_u32 %[nn], %[nn], s87\n\t"
        "v_writelane_b32 %[aT], s95, m0\n\t"
        "s_cbranch_scc1 520f\n\t"
        "s_sub_u32 s88, 63, s85\n\t"
        "s_lshr_b64 s[82:83], s[94:95], s88\n\t"
        "v_writelane_b32 %[aK], s85, m0\n\t"
        "s_cbranch_scc0 560f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "s_addc_u32 s90, %[kRP], s84\n\t"
        "s_sub_i32 %[kRP], s90, 2\n\t"
        "320:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 600f\n\t"
        "130:\n\t"
        /* ---- S0: k_P = 0, k = 0 symbol ---- */
        "s_lshr_b32 s85, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[94:95]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s87, s84, s85\n\t"
        "s_add_u32 %[nn], %[nn], s87\n\t"
        "v_writelane_b32 %[aT], s95, m0\n\t"
        "s_cbranch_scc1 530f\n\t"
        "s_sub_u32 s88, 63, s85\n\t"
        "s_lshr_b64 s[82:83], s[94:95], s88\n\t"
        "v_writelane_b32 %[aK], s85, m0\n\t"
        "s_cbranch_scc0 560f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "s_addc_u32 s90, %[kRP], s84\n\t"
        "s_sub_i32 %[kRP], s90, 2\n\t"
        "330:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 600f\n\t"
        "140:\n\t"
        /* ---- S0: k_P = 0, k = 0 symbol ---- */
        "s_lshr_b32 s85, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[94:95]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s87, s84, s85\n\t"
        "s_add_u32 %[nn], %[nn], s87\n\t"
        "v_writelane_b32 %[aT], s95, m0\n\t"
        "s_cbranch_scc1 540f\n\t"
        "s_sub_u32 s88, 63, s85\n\t"
        "s_lshr_b64 s[82:83], s[94:95], s88\n\t"
        "v_writelane_b32 %[aK], s85, m0\n\t"
        "s_cbranch_scc0 560f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "s_addc_u32 s90, %[kRP], s84\n\t"
        "s_sub_i32 %[kRP], s90, 2\n\t"
        "340:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 600f\n\t"
        "150:\n\t"
        /* ---- S0: k_P = 0, k = 0 symbol ---- */
        "s_lshr_b32 s85, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[94:95]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s87, s84, s85\n\t"
        "s_add_u32 %[nn], %[nn], s87\n\t"
        "v_writelane_b32 %[aT], s95, m0\n\t"
        "s_cbranch_scc1 550f\n\t"
        "s_sub_u32 s88, 63, s85\n\t"
        "s_lshr_b64 s[82:83], s[94:95], s88\n\t"
        "v_writelane_b32 %[aK], s85, m0\n\t"
        "s_cbranch_scc0 560f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "s_addc_u32 s90, %[kRP], s84\n\t"
        "s_sub_i32 %[kRP], s90, 2\n\t"
        "350:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 600f\n\t"
        "s_branch 120b\n\t"
        "560:\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_sub_i32 s90, %[kRP], 2\n\t"
        "s_max_i32 %[kRP], s90, 0\n\t"
        "570:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 603f\n\t"
        "s_branch 123b\n\t"
        "561:\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_sub_i32 s90, %[kRP], 2\n\t"
        "s_max_i32 %[kRP], s90, 0\n\t"
        "571:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 604f\n\t"
        "s_branch 104b\n\t"
        "562:\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_sub_i32 s90, %[kRP], 2\n\t"
        "s_max_i32 %[kRP], s90, 0\n\t"
        "572:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 605f\n\t"
        "s_branch 105b\n\t"
        "563:\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_sub_i32 s90, %[kRP], 2\n\t"
        "s_max_i32 %[kRP], s90, 0\n\t"
        "573:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 606f\n\t"
        "s_branch 106b\n\t"
        "684:\n\t"
        "s_sub_u32 m0, m0, s90\n\t"
        "s_cmp_eq_u32 s91, 0\n\t"
        "s_cbranch_scc1 754f\n\t"
        "734:\n\t"
        "s_bitcmp1_b64 s[94:95], 62\n\t"
        "s_cbranch_scc1 704f\n\t"
        "s_bfe_u64 s[90:91], s[94:95], 0x1003d\n\t"
        "s_add_u32 s90, s90, 2\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_cbranch_scc1 774f\n\t"
        "s_lshl_b64 s[92:93], s[94:95], 3\n\t"
        "s_lshr_b32 s89, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[92:93]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s86, s84, 3\n\t"
        "s_add_u32 s86, s86, s89\n\t"
        "s_add_u32 %[nn], %[nn], s86\n\t"
        "v_writelane_b32 %[aT], s93, m0\n\t"
        "s_cbranch_scc1 1134f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "934:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 605f\n\t"
        "s_branch 105b\n\t"
        "685:\n\t"
        "s_sub_u32 m0, m0, s90\n\t"
        "s_cmp_eq_u32 s91, 0\n\t"
        "s_cbranch_scc1 755f\n\t"
        "735:\n\t"
        "s_bitcmp1_b64 s[94:95], 62\n\t"
        "s_cbranch_scc1 705f\n\t"
        "s_bfe_u64 s[90:91], s[94:95], 0x1003d\n\t"
        "s_add_u32 s90, s90, 2\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_cbranch_scc1 775f\n\t"
        "s_lshl_b64 s[92:93], s[94:95], 3\n\t"
        "s_lshr_b32 s89, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[92:93]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s86, s84, 3\n\t"
        "s_add_u32 s86, s86, s89\n\t"
        "s_add_u32 %[nn], %[nn], s86\n\t"
        "v_writelane_b32 %[aT], s93, m0\n\t"
        "s_cbranch_scc1 1135f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "935:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 606f\n\t"
        "s_branch 106b\n\t"
        "686:\n\t"
        "s_sub_u32 m0, m0, s90\n\t"
        "s_cmp_eq_u32 s91, 0\n\t"
        "s_cbranch_scc1 756f\n\t"
        "736:\n\t"
        "s_bitcmp1_b64 s[94:95], 62\n\t"
        "s_cbranch_scc1 706f\n\t"
        "s_bfe_u64 s[90:91], s[94:95], 0x2003c\n\t"
        "s_add_u32 s90, s90, 2\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_cbranch_scc1 776f\n\t"
        "s_lshl_b64 s[92:93], s[94:95], 4\n\t"
        "s_lshr_b32 s89, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[92:93]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s86, s84, 4\n\t"
        "s_add_u32 s86, s86, s89\n\t"
        "s_add_u32 %[nn], %[nn], s86\n\t"
        "v_writelane_b32 %[aT], s93, m0\n\t"
        "s_cbranch_scc1 1136f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "936:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 607f\n\t"
        "s_branch 107b\n\t"
        "687:\n\t"
        "s_sub_u32 m0, m0, s90\n\t"
        "s_cmp_eq_u32 s91, 0\n\t"
        "s_cbranch_scc1 757f\n\t"
        "737:\n\t"
        "s_bitcmp1_b64 s[94:95], 62\n\t"
        "s_cbranch_scc1 707f\n\t"
        "s_bfe_u64 s[90:91], s[94:95], 0x2003c\n\t"
        "s_add_u32 s90, s90, 2\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_cbranch_scc1 777f\n\t"
        "s_lshl_b64 s[92:93], s[94:95], 4\n\t"
        "s_lshr_b32 s89, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[92:93]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s86, s84, 4\n\t"
        "s_add_u32 s86, s86, s89\n\t"
        "s_add_u32 %[nn], %[nn], s86\n\t"
        "v_writelane_b32 %[aT], s93, m0\n\t"
        "s_cbranch_scc1 1137f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "937:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 608f\n\t"
        "s_branch 108b\n\t"
        "688:\n\t"
        "s_sub_u32 m0, m0, s90\n\t"
        "s_cmp_eq_u32 s91, 0\n\t"
        "s_cbranch_scc1 758f\n\t"
        "738:\n\t"
        "s_bitcmp1_b64 s[94:95], 62\n\t"
        "s_cbranch_scc1 708f\n\t"
        "s_bfe_u64 s[90:91], s[94:95], 0x2003c\n\t"
        "s_add_u32 s90, s90, 4\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_cbranch_scc1 778f\n\t"
        "s_lshl_b64 s[92:93], s[94:95], 4\n\t"
        "s_lshr_b32 s89, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[92:93]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s86, s84, 4\n\t"
        "s_add_u32 s86, s86, s89\n\t"
        "s_add_u32 %[nn], %[nn], s86\n\t"
        "v_writelane_b32 %[aT], s93, m0\n\t"
        "s_cbranch_scc1 1138f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "938:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 609f\n\t"
        "s_branch 109b\n\t"
        "689:\n\t"
        "s_sub_u32 m0, m0, s90\n\t"
        "s_cmp_eq_u32 s91, 0\n\t"
        "s_cbranch_scc1 759f\n\t"
        "739:\n\t"
        "s_bitcmp1_b64 s[94:95], 62\n\t"
        "s_cbranch_scc1 709f\n\t"
        "s_bfe_u64 s[90:91], s[94:95], 0x2003c\n\t"
        "s_add_u32 s90, s90, 4\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_cbranch_scc1 779f\n\t"
        "s_lshl_b64 s[92:93], s[94:95], 4\n\t"
        "s_lshr_b32 s89, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[92:93]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s86, s84, 4\n\t"
        "s_add_u32 s86, s86, s89\n\t"
        "s_add_u32 %[nn], %[nn], s86\n\t"
        "v_writelane_b32 %[aT], s93, m0\n\t"
        "s_cbranch_scc1 1139f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "939:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 610f\n\t"
        "s_branch 110b\n\t"
        "690:\n\t"
        "s_sub_u32 m0, m0, s90\n\t"
        "s_cmp_eq_u32 s91, 0\n\t"
        "s_cbranch_scc1 760f\n\t"
        "740:\n\t"
        "s_bitcmp1_b64 s[94:95], 62\n\t"
        "s_cbranch_scc1 710f\n\t"
        "s_bfe_u64 s[90:91], s[94:95], 0x3003b\n\t"
        "s_add_u32 s90, s90, 4\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_cbranch_scc1 780f\n\t"
        "s_lshl_b64 s[92:93], s[94:95], 5\n\t"
        "s_lshr_b32 s89, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[92:93]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s86, s84, 5\n\t"
        "s_add_u32 s86, s86, s89\n\t"
        "s_add_u32 %[nn], %[nn], s86\n\t"
        "v_writelane_b32 %[aT], s93, m0\n\t"
        "s_cbranch_scc1 1140f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "940:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 611f\n\t"
        "s_branch 111b\n\t"
        "691:\n\t"
        "s_sub_u32 m0, m0, s90\n\t"
        "s_cmp_eq_u32 s91, 0\n\t"
        "s_cbranch_scc1 761f\n\t"
        "741:\n\t"
        "s_bitcmp1_b64 s[94:95], 62\n\t"
        "s_cbranch_scc1 711f\n\t"
        "s_bfe_u64 s[90:91], s[94:95], 0x3003b\n\t"
        "s_add_u32 s90, s90, 4\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_cbranch_scc1 781f\n\t"
        "s_lshl_b64 s[92:93], s[94:95], 5\n\t"
        "s_lshr_b32 s89, %[kRP], 2\n\t"
        "s_orn2_b64 s[80:81], 1, s[92:93]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_addc_u32 s86, s84, 5\n\t"
        "s_add_u32 s86, s86, s89\n\t"
        "s_add_u32 %[nn], %[nn], s86\n\t"
        "v_writelane_b32 %[aT], s93, m0\n\t"
        "s_cbranch_scc1 1141f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "941:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 612f\n\t"
        "s_mov_b32 %[kP], 12\n\t"
        "s_branch 100b\n\t"
        "523:\n\t"
        "s_cmp_gt_u32 s87, 32\n\t"
        "s_cbranch_scc1 643f\n\t"
        "s_sub_u32 s88, 63, s85\n\t"
        "s_lshr_b64 s[82:83], s[94:95], s88\n\t"
        "v_writelane_b32 %[aK], s85, m0\n\t"
        "s_cbranch_scc0 583f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "s_addc_u32 s90, %[kRP], s84\n\t"
        "s_sub_i32 %[kRP], s90, 2\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 622f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 323b\n\t"
        "522:\n\t"
        "s_cmp_gt_u32 s87, 32\n\t"
        "s_cbranch_scc1 642f\n\t"
        "s_sub_u32 s88, 63, s85\n\t"
        "s_lshr_b64 s[82:83], s[94:95], s88\n\t"
        "v_writelane_b32 %[aK], s85, m0\n\t"
        "s_cbranch_scc0 582f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "s_addc_u32 s90, %[kRP], s84\n\t"
        "s_sub_i32 %[kRP], s90, 2\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 621f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 322b\n\t"
        "521:\n\t"
        "s_cmp_gt_u32 s87, 32\n\t"
        "s_cbranch_scc1 641f\n\t"
        "s_sub_u32 s88, 63, s85\n\t"
        "s_lshr_b64 s[82:83], s[94:95], s88\n\t"
        "v_writelane_b32 %[aK], s85, m0\n\t"
        "s_cbranch_scc0 581f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "s_addc_u32 s90, %[kRP], s84\n\t"
        "s_sub_i32 %[kRP], s90, 2\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 620f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 321b\n\t"
        "520:\n\t"
        "s_cmp_gt_u32 s87, 32\n\t"
        "s_cbranch_scc1 640f\n\t"
        "s_sub_u32 s88, 63, s85\n\t"
        "s_lshr_b64 s[82:83], s[94:95], s88\n\t"
        "v_writelane_b32 %[aK], s85, m0\n\t"
        "s_cbranch_scc0 580f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "s_addc_u32 s90, %[kRP], s84\n\t"
        "s_sub_i32 %[kRP], s90, 2\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 620f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 320b\n\t"
        "530:\n\t"
        "s_cmp_gt_u32 s87, 32\n\t"
        "s_cbranch_scc1 640f\n\t"
        "s_sub_u32 s88, 63, s85\n\t"
        "s_lshr_b64 s[82:83], s[94:95], s88\n\t"
        "v_writelane_b32 %[aK], s85, m0\n\t"
        "s_cbranch_scc0 580f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "s_addc_u32 s90, %[kRP], s84\n\t"
        "s_sub_i32 %[kRP], s90, 2\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 620f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 330b\n\t"
        "540:\n\t"
        "s_cmp_gt_u32 s87, 32\n\t"
        "s_cbranch_scc1 640f\n\t"
        "s_sub_u32 s88, 63, s85\n\t"
        "s_lshr_b64 s[82:83], s[94:95], s88\n\t"
        "v_writelane_b32 %[aK], s85, m0\n\t"
        "s_cbranch_scc0 580f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "s_addc_u32 s90, %[kRP], s84\n\t"
        "s_sub_i32 %[kRP], s90, 2\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 620f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 340b\n\t"
        "550:\n\t"
        "s_cmp_gt_u32 s87, 32\n\t"
        "s_cbranch_scc1 640f\n\t"
        "s_sub_u32 s88, 63, s85\n\t"
        "s_lshr_b64 s[82:83], s[94:95], s88\n\t"
        "v_writelane_b32 %[aK], s85, m0\n\t"
        "s_cbranch_scc0 580f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "s_addc_u32 s90, %[kRP], s84\n\t"
        "s_sub_i32 %[kRP], s90, 2\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 620f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 350b\n\t"
        "580:\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_sub_i32 s90, %[kRP], 2\n\t"
        "s_max_i32 %[kRP], s90, 0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 623f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 570b\n\t"
        "581:\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_sub_i32 s90, %[kRP], 2\n\t"
        "s_max_i32 %[kRP], s90, 0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 624f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 571b\n\t"
        "582:\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_sub_i32 s90, %[kRP], 2\n\t"
        "s_max_i32 %[kRP], s90, 0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 625f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 572b\n\t"
        "583:\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s87\n\t"
        "s_sub_i32 s90, %[kRP], 2\n\t"
        "s_max_i32 %[kRP], s90, 0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 626f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 573b\n\t"
        "511:\n\t"
        "s_cmp_gt_u32 s86, 32\n\t"
        "s_cbranch_scc1 671f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 630f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 311b\n\t"
        "510:\n\t"
        "s_cmp_gt_u32 s86, 32\n\t"
        "s_cbranch_scc1 670f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 629f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 310b\n\t"
        "509:\n\t"
        "s_cmp_gt_u32 s86, 32\n\t"
        "s_cbranch_scc1 669f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 628f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 309b\n\t"
        "508:\n\t"
        "s_cmp_gt_u32 s86, 32\n\t"
        "s_cbranch_scc1 668f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 627f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 308b\n\t"
        "507:\n\t"
        "s_cmp_gt_u32 s86, 32\n\t"
        "s_cbranch_scc1 667f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 626f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 307b\n\t"
        "506:\n\t"
        "s_cmp_gt_u32 s86, 32\n\t"
        "s_cbranch_scc1 666f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 625f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 306b\n\t"
        "505:\n\t"
        "s_cmp_gt_u32 s86, 32\n\t"
        "s_cbranch_scc1 665f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 624f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 305b\n\t"
        "504:\n\t"
        "s_cmp_gt_u32 s86, 32\n\t"
        "s_cbranch_scc1 664f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 623f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 304b\n\t"
        "1134:\n\t"
        "s_cmp_gt_u32 s86, 32\n\t"
        "s_cbranch_scc1 664f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 625f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 934b\n\t"
        "1135:\n\t"
        "s_cmp_gt_u32 s86, 32\n\t"
        "s_cbranch_scc1 665f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 626f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 935b\n\t"
        "1136:\n\t"
        "s_cmp_gt_u32 s86, 32\n\t"
        "s_cbranch_scc1 666f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 627f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 936b\n\t"
        "1137:\n\t"
        "s_cmp_gt_u32 s86, 32\n\t"
        "s_cbranch_scc1 667f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 628f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 937b\n\t"
        "1138:\n\t"
        "s_cmp_gt_u32 s86, 32\n\t"
        "s_cbranch_scc1 668f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 629f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 938b\n\t"
        "1139:\n\t"
        "s_cmp_gt_u32 s86, 32\n\t"
        "s_cbranch_scc1 669f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 630f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 939b\n\t"
        "1140:\n\t"
        "s_cmp_gt_u32 s86, 32\n\t"
        "s_cbranch_scc1 670f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 631f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 940b\n\t"
        "1141:\n\t"
        "s_cmp_gt_u32 s86, 32\n\t"
        "s_cbranch_scc1 671f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 632f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 941b\n\t"
        "500:\n\t"
        "s_cmp_gt_u32 s86, 32\n\t"
        "s_cbranch_scc1 97f\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 721f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 300b\n\t"
        "600:\n\t"
        "s_mov_b32 %[kP], 0\n\t"
        "s_branch 9f\n\t"
        "601:\n\t"
        "s_mov_b32 %[kP], 1\n\t"
        "s_branch 9f\n\t"
        "602:\n\t"
        "s_mov_b32 %[kP], 2\n\t"
        "s_branch 9f\n\t"
        "603:\n\t"
        "s_mov_b32 %[kP], 3\n\t"
        "s_branch 9f\n\t"
        "604:\n\t"
        "s_mov_b32 %[kP], 4\n\t"
        "s_branch 9f\n\t"
        "605:\n\t"
        "s_mov_b32 %[kP], 5\n\t"
        "s_branch 9f\n\t"
        "606:\n\t"
        "s_mov_b32 %[kP], 6\n\t"
        "s_branch 9f\n\t"
        "607:\n\t"
        "s_mov_b32 %[kP], 7\n\t"
        "s_branch 9f\n\t"
        "608:\n\t"
        "s_mov_b32 %[kP], 8\n\t"
        "s_branch 9f\n\t"
        "609:\n\t"
        "s_mov_b32 %[kP], 9\n\t"
        "s_branch 9f\n\t"
        "610:\n\t"
        "s_mov_b32 %[kP], 10\n\t"
        "s_branch 9f\n\t"
        "611:\n\t"
        "s_mov_b32 %[kP], 11\n\t"
        "s_branch 9f\n\t"
        "612:\n\t"
        "s_mov_b32 %[kP], 12\n\t"
        "s_branch 9f\n\t"
        "620:\n\t"
        "s_mov_b32 %[kP], 0\n\t"
        "s_branch 7f\n\t"
        "621:\n\t"
        "s_mov_b32 %[kP], 1\n\t"
        "s_branch 7f\n\t"
        "622:\n\t"
        "s_mov_b32 %[kP], 2\n\t"
        "s_branch 7f\n\t"
        "623:\n\t"
        "s_mov_b32 %[kP], 3\n\t"
        "s_branch 7f\n\t"
        "624:\n\t"
        "s_mov_b32 %[kP], 4\n\t"
        "s_branch 7f\n\t"
        "625:\n\t"
        "s_mov_b32 %[kP], 5\n\t"
        "s_branch 7f\n\t"
        "626:\n\t"
        "s_mov_b32 %[kP], 6\n\t"
        "s_branch 7f\n\t"
        "627:\n\t"
        "s_mov_b32 %[kP], 7\n\t"
        "s_branch 7f\n\t"
        "628:\n\t"
        "s_mov_b32 %[kP], 8\n\t"
        "s_branch 7f\n\t"
        "629:\n\t"
        "s_mov_b32 %[kP], 9\n\t"
        "s_branch 7f\n\t"
        "630:\n\t"
        "s_mov_b32 %[kP], 10\n\t"
        "s_branch 7f\n\t"
        "631:\n\t"
        "s_mov_b32 %[kP], 11\n\t"
        "s_branch 7f\n\t"
        "632:\n\t"
        "s_mov_b32 %[kP], 12\n\t"
        "s_branch 7f\n\t"
        "640:\n\t"
        "s_mov_b32 %[kP], 0\n\t"
        "s_branch 19f\n\t"
        "641:\n\t"
        "s_mov_b32 %[kP], 1\n\t"
        "s_branch 19f\n\t"
        "642:\n\t"
        "s_mov_b32 %[kP], 2\n\t"
        "s_branch 19f\n\t"
        "643:\n\t"
        "s_mov_b32 %[kP], 3\n\t"
        "s_branch 19f\n\t"
        "664:\n\t"
        "s_mov_b32 %[kP], 4\n\t"
        "s_branch 97f\n\t"
        "665:\n\t"
        "s_mov_b32 %[kP], 5\n\t"
        "s_branch 97f\n\t"
        "666:\n\t"
        "s_mov_b32 %[kP], 6\n\t"
        "s_branch 97f\n\t"
        "667:\n\t"
        "s_mov_b32 %[kP], 7\n\t"
        "s_branch 97f\n\t"
        "668:\n\t"
        "s_mov_b32 %[kP], 8\n\t"
        "s_branch 97f\n\t"
        "669:\n\t"
        "s_mov_b32 %[kP], 9\n\t"
        "s_branch 97f\n\t"
        "670:\n\t"
        "s_mov_b32 %[kP], 10\n\t"
        "s_branch 97f\n\t"
        "671:\n\t"
        "s_mov_b32 %[kP], 11\n\t"
        "s_branch 97f\n\t"
        "754:\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_or_b32 m0, m0, 0x40000000\n\t"
        "s_lshl_b64 s[94:95], s[94:95], 2\n\t"
        "s_add_u32 %[nn], %[nn], 2\n\t"
        "s_mov_b32 %[kP], 4\n\t"
        "s_branch 9f\n\t"
        "755:\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_or_b32 m0, m0, 0x40000000\n\t"
        "s_lshl_b64 s[94:95], s[94:95], 2\n\t"
        "s_add_u32 %[nn], %[nn], 2\n\t"
        "s_mov_b32 %[kP], 5\n\t"
        "s_branch 9f\n\t"
        "756:\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_or_b32 m0, m0, 0x40000000\n\t"
        "s_lshl_b64 s[94:95], s[94:95], 2\n\t"
        "s_add_u32 %[nn], %[nn], 2\n\t"
        "s_mov_b32 %[kP], 6\n\t"
        "s_branch 9f\n\t"
        "757:\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_or_b32 m0, m0, 0x40000000\n\t"
        "s_lshl_b64 s[94:95], s[94:95], 2\n\t"
        "s_add_u32 %[nn], %[nn], 2\n\t"
        "s_mov_b32 %[kP], 7\n\t"
        "s_branch 9f\n\t"
        "758:\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_or_b32 m0, m0, 0x40000000\n\t"
        "s_lshl_b64 s[94:95], s[94:95], 3\n\t"
        "s_add_u32 %[nn], %[nn], 3\n\t"
        "s_mov_b32 %[kP], 8\n\t"
        "s_branch 9f\n\t"
        "759:\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_or_b32 m0, m0, 0x40000000\n\t"
        "s_lshl_b64 s[94:95], s[94:95], 3\n\t"
        "s_add_u32 %[nn], %[nn], 3\n\t"
        "s_mov_b32 %[kP], 9\n\t"
        "s_branch 9f\n\t"
        "760:\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_or_b32 m0, m0, 0x40000000\n\t"
        "s_lshl_b64 s[94:95], s[94:95], 3\n\t"
        "s_add_u32 %[nn], %[nn], 3\n\t"
        "s_mov_b32 %[kP], 10\n\t"
        "s_branch 9f\n\t"
        "761:\n\t"
        "s_add_u32 m0, m0, s90\n\t"
        "s_or_b32 m0, m0, 0x40000000\n\t"
        "s_lshl_b64 s[94:95], s[94:95], 3\n\t"
        "s_add_u32 %[nn], %[nn], 3\n\t"
        "s_mov_b32 %[kP], 11\n\t"
        "s_branch 9f\n\t"
        "774:\n\t"
        "s_or_b32 m0, m0, 0x40000000\n\t"
        "s_lshl_b64 s[94:95], s[94:95], 3\n\t"
        "s_add_u32 %[nn], %[nn], 3\n\t"
        "s_mov_b32 %[kP], 6\n\t"
        "s_branch 9f\n\t"
        "775:\n\t"
        "s_or_b32 m0, m0, 0x40000000\n\t"
        "s_lshl_b64 s[94:95], s[94:95], 3\n\t"
        "s_add_u32 %[nn], %[nn], 3\n\t"
        "s_mov_b32 %[kP], 7\n\t"
        "s_branch 9f\n\t"
        "776:\n\t"
        "s_or_b32 m0, m0, 0x40000000\n\t"
        "s_lshl_b64 s[94:95], s[94:95], 4\n\t"
        "s_add_u32 %[nn], %[nn], 4\n\t"
        "s_mov_b32 %[kP], 8\n\t"
        "s_branch 9f\n\t"
        "777:\n\t"
        "s_or_b32 m0, m0, 0x40000000\n\t"
        "s_lshl_b64 s[94:95], s[94:95], 4\n\t"
        "s_add_u32 %[nn], %[nn], 4\n\t"
        "s_mov_b32 %[kP], 9\n\t"
        "s_branch 9f\n\t"
        "778:\n\t"
        "s_or_b32 m0, m0, 0x40000000\n\t"
        "s_lshl_b64 s[94:95], s[94:95], 4\n\t"
        "s_add_u32 %[nn], %[nn], 4\n\t"
        "s_mov_b32 %[kP], 10\n\t"
        "s_branch 9f\n\t"
        "779:\n\t"
        "s_or_b32 m0, m0, 0x40000000\n\t"
        "s_lshl_b64 s[94:95], s[94:95], 4\n\t"
        "s_add_u32 %[nn], %[nn], 4\n\t"
        "s_mov_b32 %[kP], 11\n\t"
        "s_branch 9f\n\t"
        "780:\n\t"
        "s_or_b32 m0, m0, 0x40000000\n\t"
        "s_lshl_b64 s[94:95], s[94:95], 5\n\t"
        "s_add_u32 %[nn], %[nn], 5\n\t"
        "s_mov_b32 %[kP], 12\n\t"
        "s_branch 9f\n\t"
        "781:\n\t"
        "s_or_b32 m0, m0, 0x40000000\n\t"
        "s_lshl_b64 s[94:95], s[94:95], 5\n\t"
        "s_add_u32 %[nn], %[nn], 5\n\t"
        "s_mov_b32 %[kP], 13\n\t"
        "s_branch 9f\n\t"
        "704:\n\t"
        "s_mov_b32 %[kP], 4\n\t"
        "s_mov_b32 s87, 4\n\t"
        "s_branch 36f\n\t"
        "705:\n\t"
        "s_mov_b32 %[kP], 5\n\t"
        "s_mov_b32 s87, 5\n\t"
        "s_branch 36f\n\t"
        "706:\n\t"
        "s_mov_b32 %[kP], 6\n\t"
        "s_mov_b32 s87, 6\n\t"
        "s_branch 36f\n\t"
        "707:\n\t"
        "s_mov_b32 %[kP], 7\n\t"
        "s_mov_b32 s87, 7\n\t"
        "s_branch 36f\n\t"
        "708:\n\t"
        "s_mov_b32 %[kP], 8\n\t"
        "s_mov_b32 s87, 8\n\t"
        "s_branch 36f\n\t"
        "709:\n\t"
        "s_mov_b32 %[kP], 9\n\t"
        "s_mov_b32 s87, 9\n\t"
        "s_branch 36f\n\t"
        "710:\n\t"
        "s_mov_b32 %[kP], 10\n\t"
        "s_mov_b32 s87, 10\n\t"
        "s_branch 36f\n\t"
        "711:\n\t"
        "s_mov_b32 %[kP], 11\n\t"
        "s_mov_b32 s87, 11\n\t"
        "s_branch 36f\n\t"
        "720:\n\t"
        "s_sub_u32 %[kP], %[kP], 1\n\t"
        "s_branch 9f\n\t"
        "721:\n\t"
        "s_sub_u32 %[kP], %[kP], 1\n\t"
        "s_branch 7f\n\t"
        "722:\n\t"
        "s_sub_u32 m0, m0, s90\n\t"
        "s_branch 9f\n\t"
        "723:\n\t"
        "s_mov_b32 s87, %[kP]\n\t"
        "36:\n\t"
        /* ---- run prefix with one-bits (s87 = k_P) ---- */
        "s_not_b64 s[80:81], s[94:95]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_mov_b32 s82, 0\n\t"
        "31:\n\t"
        "s_cmp_gt_u32 s84, 6\n\t"
        "s_cbranch_scc1 9f\n\t"
        "s_sub_u32 s83, 0, s84\n\t"
        "35:\n\t"
        "s_lshr_b32 s85, s87, 2\n\t"
        "s_lshl_b32 s89, 1, s85\n\t"
        "s_add_u32 s82, s82, s89\n\t"
        "s_add_u32 s87, s87, 2\n\t"
        "s_add_u32 s83, s83, 1\n\t"
        "s_cbranch_scc0 35b\n\t"
        "s_lshr_b32 s85, s87, 2\n\t"
        "s_cmp_gt_u32 s85, 24\n\t"
        "s_cbranch_scc1 9f\n\t"
        "32:\n\t"
        "s_add_u32 s86, s84, 1\n\t"
        "s_lshl_b64 s[80:81], s[94:95], s86\n\t"
        "s_sub_u32 s88, 64, s85\n\t"
        "s_lshr_b64 s[90:91], s[80:81], s88\n\t"
        "s_add_u32 s82, s82, s90\n\t"
        "s_add_u32 s83, m0, s82\n\t"
        "s_cbranch_scc1 9f\n\t"
        "s_lshl_b64 s[92:93], s[80:81], s85\n\t"
        "s_add_u32 s86, s86, s85\n\t"
        "s_not_b64 s[80:81], s[92:93]\n\t"
        "s_flbit_i32_b64 s84, s[80:81]\n\t"
        "s_lshr_b32 s89, %[kRP], 2\n\t"
        "s_add_u32 s86, s86, s84\n\t"
        "s_add_u32 s86, s86, s89\n\t"
        "s_add_u32 s86, s86, 1\n\t"
        "s_add_u32 %[nn], %[nn], s86\n\t"
        "s_cbranch_scc1 38f\n\t"
        "s_mov_b32 m0, s83\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_u32 %[kP], s87, 1\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "v_writelane_b32 %[aT], s93, m0\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "39:\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "s_cbranch_scc1 9f\n\t"
        "s_branch 91b\n\t"
        "38:\n\t"
        "s_cmp_gt_u32 s86, 32\n\t"
        "s_cbranch_scc1 98f\n\t"
        "s_mov_b32 m0, s83\n\t"
        "s_lshl_b64 s[94:95], s[94:95], s86\n\t"
        "s_cmp_lg_u32 s84, 0\n\t"
        "s_addc_u32 s91, %[kRP], s84\n\t"
        "s_sub_u32 %[kP], s87, 1\n\t"
        "s_sub_i32 s91, s91, 2\n\t"
        "s_max_i32 %[kRP], s91, 0\n\t"
        "v_writelane_b32 %[aT], s93, m0\n\t"
        "v_writelane_b32 %[aK], s89, m0\n\t"
        "v_writelane_b32 %[aF], 1, m0\n\t"
        "s_add_u32 %[wpos], %[wpos], 1\n\t"
        "s_and_b32 s88, %[wpos], 63\n\t"
        "s_cbranch_scc0 7f\n\t"
        "s_lshl_b64 s[90:91], s[96:97], %[nn]\n\t"
        "s_or_b64 s[94:95], s[94:95], s[90:91]\n\t"
        "s_sub_u32 %[nn], %[nn], 32\n\t"
        "v_readlane_b32 s96, %[cur], s88\n\t"
        "s_branch 39b\n\t"
        "19:\n\t"
        "v_writelane_b32 %[aT], 0, m0\n\t"
        "s_sub_u32 %[nn], %[nn], s87\n\t"
        "s_branch 9f\n\t"
        "97:\n\t"
        "v_writelane_b32 %[aT], 0, m0\n\t"
        "s_sub_u32 %[nn], %[nn], s86\n\t"
        "s_sub_u32 m0, m0, s90\n\t"
        "s_branch 9f\n\t"
        "98:\n\t"
        "s_sub_u32 %[nn], %[nn], s86\n\t"
        "s_branch 9f\n\t"
        "7:\n\t"
        "s_sub_u32 %[wpos], %[wpos], 1\n\t"
        "s_add_u32 m0, m0, 1\n\t"
        "9:\n\t"
        "s_mov_b32 %[neg], m0\n\t"
        "s_mov_b64 %[win], s[94:95]\n\t"

        : [win] "+s"(win), [nn] "+s"(nn), [wpos] "+s"(wpos), [kRP] "+s"(kRP), [kP] "+s"(kP), [neg] "+s"(neg),
          [aT] "+v"(aT), [aK] "+v"(aK), [aF] "+v"(aF)
        : [cur] "v"(cur), [pre] "s"(pre)
        : "scc", "m0", "s80", "s81", "s82", "s83", "s84", "s85", "s86", "s87", "s88", "s89", "s90", "s91", "s92",
          "s93", "s94", "s95", "s96", "s97");
}

/* The symbol a lane of the scalar loop's accumulators stands for.  K < 32:
 * T's leading bits are a Golomb-Rice code with parameter K (q one-bits, a
 * zero, K remainder bits; the scalar loop writes a lane only for codes that
 * fit in 32 bits), so u = (q << K) | r, plus F (1 for the symbol after a zero
 * run, stored as u + 1).  K >= 32: T is the value itself.  T = 0 (a lane that
 * only a zero run passed over) gives 0 for any K < 32. */
__device__ __forceinline__ uint32_t lane_symbol(uint32_t T, uint32_t K, uint32_t F) {
    const uint32_t q = static_cast<uint32_t>(__clz(static_cast<int>(~T)));
    const uint32_t k = K & 31u;
    const uint32_t r = __builtin_amdgcn_ubfe(T, (31u - q - k) & 31u, k);
    return (K >= 32u ? T : ((q << k) | r)) + F;
}

__device__ __forceinline__ int rlgr_decode_wave(const uint8_t* __restrict__ src, uint32_t nbytes, uint32_t total,
                                                 uint32_t qp, float* coef) {
    BitReader br;
    br_init(br, src, nbytes);
    br_refill(br);
    br_refill(br);
    const uint32_t lane = lane_id();
    /* Symbol acc_base + i sits in lane (i - lim) mod 64, where lim is the
     * chunk's symbol count (64 for all but a slice's last chunk): the lane that
     * the scalar loop's v_writelane selects with the low six bits of its
     * counter nrel - lim.  Full chunks are therefore in lane order.  A lane
     * holds its symbol in the form lane_symbol() decodes: (aT, aK, aF) =
     * (the 32 bits that start with the symbol's Golomb-Rice code, k_R, 1 after
     * a zero run) from the scalar loop, or (raw value, 32, 0) from this
     * function's own decoding.  Lanes that only a zero run passed over keep
     * aT = aF = 0, which decodes to 0. */
    uint32_t aT = 0, aK = 0, aF = 0;
    uint32_t acc_base = 0, nrel = 0;
    uint32_t kP = 0, kRP = 8;
    int st = PVVC_OK;

    auto chunk_lim = [&](uint32_t base) {
        const uint32_t r = total - base;
        return r < 64u ? r : 64u;
    };
    auto store = [&](uint32_t base, uint32_t v) {
        const uint32_t i = base + ((lane + chunk_lim(base)) & 63u);
        const int32_t s = static_cast<int32_t>(unsign2sign(v));
        /* (float)(int64 s * qp) with one rounding: the FP64 product is exact
         * (|s * qp| < 2^39), the FP64 -> FP32 conversion rounds to nearest even
         * like the integer one */
        if (i < total) coef[i] = static_cast<float>(static_cast<double>(s) * static_cast<double>(qp));
    };
    auto store_chunk = [&](uint32_t base) {
        store(base, lane_symbol(aT, aK, aF));
        aT = 0;
        aF = 0;
    };
    auto emit = [&](uint32_t u) {
        const bool mine = lane == ((nrel - chunk_lim(acc_base)) & 63u);
        aT = mine ? u : aT;
        aK = mine ? 32u : aK;
        aF = mine ? 0u : aF;
        ++nrel;
    };

    while (acc_base < total) {
        const uint32_t rest = total - acc_base;
        const uint32_t lim = rest < 64u ? rest : 64u;
        if (nrel < lim && br.nb > 32) {
            kRP = __builtin_amdgcn_readfirstlane(kRP);
            kP = __builtin_amdgcn_readfirstlane(kP);
            uint32_t neg = __builtin_amdgcn_readfirstlane(nrel - lim);
            int32_t nn = __builtin_amdgcn_readfirstlane(32 - br.nb);
            for (;;) {
                /* the loop takes the next dword (the window's next refill)
                 * in an SGPR and reads the one after it during each refill */
                br_sync(br);
                const uint32_t pre = static_cast<uint32_t>(
                    __builtin_amdgcn_readlane(static_cast<int>(br.cur), static_cast<int>(br.wpos & 63u)));
                rlgr_fast_run(br.win, nn, br.wpos, kRP, kP, neg, br.cur, pre, aT, aK, aF);
                /* full chunks back to back: the chunk is complete (neg = 0),
                 * the window holds 33+ bits (nn < 0) and the next chunk is a
                 * full one too, so every lane stores and lim stays 64 */
                if (neg != 0u || nn >= 0 || total - acc_base < 128u) break;
                const int32_t s = unsign2sign(lane_symbol(aT, aK, aF));
                coef[acc_base + lane] = static_cast<float>(static_cast<double>(s) * static_cast<double>(qp));
                aT = 0;
                aF = 0;
                acc_base += 64u;
                neg = 0u - 64u;
            }
            br.nb = 32 - nn;
            if ((neg >> 30) == 1u) {
                /* a zero run reached the chunk end: the loop consumed its
                 * prefix and length and left in neg (bit 30 set) how many of
                 * its zeros lie past this chunk; the run's symbol follows them
                 * (the tail of the general run path below) */
                const uint32_t n2 = acc_base + lim + (neg & 0x3fffffffu);
                store_chunk(acc_base);
                if (n2 >= total) {
                    acc_base += 64u;
                    nrel = 0;
                    break;
                }
                const uint32_t nb2 = n2 & ~63u;
                for (uint32_t b = acc_base + 64u; b < nb2; b += 64u) store(b, 0);
                acc_base = nb2;
                nrel = n2 - acc_base;
                const uint32_t kR = kRP >> 2;
                const uint64_t u = br_gr(br, kR);
                if (u + 1 >= (1ull << 32)) {
                    st = PVVC_ERR_SYMBOL_RANGE;
                    break;
                }
                emit(static_cast<uint32_t>(u + 1));
                kRP = adapt_krp(kRP, u >> kR);
                kP -= 1; /* run mode: k_P >= 4 */
                continue;
            }
            nrel = neg + lim;
        }
        if (nrel >= lim) { /* chunk of 64 symbols complete */
            store_chunk(acc_base);
            acc_base += 64u;
            nrel = 0;
            continue;
        }
        if (br.nb <= 32) {
            br_refill(br);
            continue;
        }
        const uint32_t kR = kRP >> 2;
        if (kP < 4) {
            const uint64_t u = br_gr(br, kR);
            if (u >= (1ull << 32)) {
                st = PVVC_ERR_SYMBOL_RANGE;
                break;
            }
            emit(static_cast<uint32_t>(u));
            kRP = adapt_krp(kRP, u >> kR);
            kP = u ? (kP ? kP - 1 : 0u) : kP + 3;
            continue;
        }
        uint32_t k = kP >> 2;
        uint64_t m = 0;
        bool bad = false;
        for (;;) {
            br_refill(br);
            const uint32_t t = leading_ones(br.win);
            for (uint32_t i = 0; i < t; ++i) {
                m += 1ull << k;
                kP += 2;
                k = kP >> 2;
                if (k > 30) {
                    bad = true;
                    break;
                }
            }
            if (bad) break;
            if (static_cast<int32_t>(t) < br.nb) {
                br_skip(br, t + 1);
                break;
            }
            br_skip(br, t);
        }
        if (bad) {
            st = PVVC_ERR_BITSTREAM;
            break;
        }
        m += br_get(br, k);
        const uint64_t n = static_cast<uint64_t>(acc_base) + nrel;
        if (m >= total - n) {
            store_chunk(acc_base);
            acc_base += 64u;
            nrel = 0;
            break;
        }
        const uint32_t n2 = static_cast<uint32_t>(n + m);
        if (n2 - acc_base >= 64u) {
            store_chunk(acc_base);
            const uint32_t nb2 = n2 & ~63u;
            for (uint32_t b = acc_base + 64u; b < nb2; b += 64u) store(b, 0);
            acc_base = nb2;
        }
        nrel = n2 - acc_base;
        const uint64_t u = br_gr(br, kR);
        if (u + 1 >= (1ull << 32)) {
            st = PVVC_ERR_SYMBOL_RANGE;
            break;
        }
        emit(static_cast<uint32_t>(u + 1));
        kRP = adapt_krp(kRP, u >> kR);
        kP = kP ? kP - 1 : 0u;
    }
    if (acc_base < total) store_chunk(acc_base);
    for (uint32_t b = acc_base + 64u; b < total; b += 64u) store(b, 0);
    const int64_t consumed = static_cast<int64_t>(br.wpos) * 32 - br.nb;
    if (st == PVVC_OK && consumed > 8ll * nbytes) st = PVVC_ERR_BITSTREAM;
    return st;
}

/* ======================================================================== */
/* Device: octree structure (invert_raht_octree.cc:122-254)                  */
/* ======================================================================== */

__device__ __forceinline__ uint32_t prefix_lt(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

/* Exclusive wave scan of v in [0, 15] by bit planes; *total gets the sum. */
__device__ __forceinline__ uint32_t wave_scan_small(uint32_t v, uint32_t* total) {
    uint32_t r = 0, t = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const uint64_t m = __ballot((v >> b) & 1u);
        r += prefix_lt(m) << b;
        t += static_cast<uint32_t>(__popcll(m)) << b;
    }
    *total = t;
    return r;
}

/* Pass 1, one wavefront: level sizes and first-child indices. */
__device__ __forceinline__ int tree_structure(const uint8_t* __restrict__ geo, uint32_t geo_size, uint32_t n, uint32_t* fc,
                              LevelMeta& meta) {
    const uint32_t lane = lane_id();
    const uint32_t h = geo[24];
    if (h == 0 || h > static_cast<uint32_t>(kMaxHeight)) return PVVC_ERR_BAD_SLICE;
    uint32_t base = 0, cnt = 1;
    if (lane == 0) {
        meta.cnt[0] = 1;
        meta.base[0] = 0;
    }
    for (uint32_t l = 0; l + 1 < h; ++l) {
        if (25ull + base + cnt > geo_size) return PVVC_ERR_BAD_SLICE;
        uint32_t next = 0;
        for (uint32_t j0 = 0; j0 < cnt; j0 += 64) {
            const uint32_t j = j0 + lane;
            const uint32_t byte = j < cnt ? geo[25 + base + j] : 1u;
            /* an internal node without children is never written by the
             * encoder (raht_octree.cc:154-167); rejecting it keeps every
             * level within n nodes and the AC positions closed-form */
            if (__ballot(byte == 0u)) return PVVC_ERR_BAD_SLICE;
            uint32_t tot;
            const uint32_t ex = wave_scan_small(j < cnt ? __popc(byte) : 0u, &tot);
            if (j < cnt) fc[base + j] = next + ex;
            next += tot;
        }
        base += cnt;
        cnt = next;
        if (lane == 0) {
            meta.cnt[l + 1] = cnt;
            meta.base[l + 1] = base;
        }
    }
    if (lane == 0) {
        meta.height = h;
        meta.internal = base;
    }
    if (cnt != n) return PVVC_ERR_UNMATCHED_SIZE; /* invert_raht_octree.cc:188-190 */
    return PVVC_OK;
}

/* Pass 2, one wavefront: subtree leaf counts (the node weights). */
__device__ __forceinline__ void tree_weights(const uint8_t* __restrict__ geo, const uint32_t* fc, uint32_t* w, const LevelMeta& meta) {
    const uint32_t lane = lane_id();
    const uint32_t h = meta.height;
    for (int l = static_cast<int>(h) - 2; l >= 0; --l) {
        const uint32_t cnt = meta.cnt[l], base = meta.base[l], cbase = meta.base[l + 1];
        const bool last = (l == static_cast<int>(h) - 2);
        for (uint32_t j = lane; j < cnt; j += 64) {
            const uint32_t pc = __popc(static_cast<uint32_t>(geo[25 + base + j]));
            uint32_t s = pc;
            if (!last) {
                const uint32_t f = fc[base + j];
                s = 0;
                for (uint32_t r = 0; r < pc; ++r) s += w[cbase + f + r];
            }
            w[base + j] = s;
        }
        __threadfence_block();
    }
}

/* ======================================================================== */
/* Device: inverse RAHT + leaf output                                        */
/* ======================================================================== */

__device__ __forceinline__ uint8_t to_u8(float f) {
    /* ConvertRGB (common.h:154-158) with saturation; NaN -> 0 */
    float q = roundf(f);
    if (!(q == q)) return 0;
    q = q < 0.0f ? 0.0f : (q > 255.0f ? 255.0f : q);
    return static_cast<uint8_t>(q);
}

/* Reference Points colour: unrounded RGB / 255 (CUDA2OpenGL.cu:49-51); the
 * white-small-patch artefact (:54-58) only on request. */
__device__ __forceinline__ void write_points_rgb(Points* pt, float r, float g, float b, uint32_t n, int white_small) {
    if (white_small && n < 10) {
        pt->r = pt->g = pt->b = 1.0f;
    } else {
        pt->r = r / 255.0f;
        pt->g = g / 255.0f;
        pt->b = b / 255.0f;
    }
}

/* Leaf output with the reference colour (rY, rU, rV) already loaded by the
 * caller for a predictive slice that is not deferred (ignored otherwise). */
__device__ __forceinline__ void emit_leaf_ref(uint32_t c, float cx, float cy, float cz, float Y, float U, float V,
                                              float rY, float rU, float rV, const SliceArgs& a, float* ref,
                                              uint32_t ref_stride, bool predict, const Sink& sink) {
    const bool deferred = predict && a.stage != nullptr;
    if (deferred) { /* reference added after the launch (finish_kernel) */
        float* sp = a.stage + 3 * (static_cast<uint64_t>(a.stage_off) + c);
        sp[0] = Y;
        sp[1] = U;
        sp[2] = V;
    } else if (predict) { /* invert_raht_octree.cc:304-308 */
        Y = Y + rY;
        U = U + rU;
        V = V + rV;
    } else { /* :309-311 */
        ref[c] = Y;
        ref[ref_stride + c] = U;
        ref[2 * ref_stride + c] = V;
    }
    const float* m = a.mv;
    const float x = cx * m[0] + cy * m[1] + cz * m[2] + m[3];
    const float y = cx * m[4] + cy * m[5] + cz * m[6] + m[7];
    const float z = cx * m[8] + cy * m[9] + cz * m[10] + m[11];
    const float r = Y + 1.4020f * (V - 128.0f);
    const float g = Y - 0.3441f * (U - 128.0f) - 0.7141f * (V - 128.0f);
    const float b = Y + 1.7720f * (U - 128.0f);
    const uint64_t o = a.out_off + c;
    if (sink.mode == PVVC_OUT_NATIVE) {
        float* px = sink.xyz + 3 * o;
        px[0] = x;
        px[1] = y;
        px[2] = z;
        if (deferred) return;
        uint8_t* pc = sink.rgb + 3 * o;
        pc[0] = to_u8(r);
        pc[1] = to_u8(g);
        pc[2] = to_u8(b);
    } else {
        Points* pt = sink.pts + o;
        pt->x = x;
        pt->y = y;
        pt->z = z;
        if (deferred) return;
        write_points_rgb(pt, r, g, b, a.n, sink.white_small);
    }
}

__device__ __forceinline__ void emit_leaf(uint32_t c, float cx, float cy, float cz, float Y, float U, float V,
                                          const SliceArgs& a, float* ref, uint32_t ref_stride, bool predict,
                                          const Sink& sink) {
    const bool load_ref = predict && a.stage == nullptr;
    const float rY = load_ref ? ref[c] : 0.0f;
    const float rU = load_ref ? ref[ref_stride + c] : 0.0f;
    const float rV = load_ref ? ref[2 * ref_stride + c] : 0.0f;
    emit_leaf_ref(c, cx, cy, cz, Y, U, V, rY, rU, rV, a, ref, ref_stride, predict, sink);
}

/* One inverse Haar butterfly pair (octree_base.cc:186-206) on channel values;
 * only called with w0, w1 > 0. */
__device__ __forceinline__ void inv_haar(uint32_t w0, uint32_t w1, float g, float h, float& o0, float& o1) {
    const float base = sqrtf(static_cast<float>(w0 + w1));
    const float ta = sqrtf(static_cast<float>(w0)) / base;
    const float tc = sqrtf(static_cast<float>(w1)) / base;
    o0 = g * ta + h * (-tc);
    o1 = g * tc + h * ta;
}

/* ======================================================================== */
/* Device: one slice, one workgroup                                          */
/* ======================================================================== */

/* Optional per-phase time stamps (built with -DPVVC_PHASE_PROFILE): for
 * item i, stamps [8i, 8i+8) = start, after RLGR (one wavefront per slice;
 * unset otherwise), after the tree (likewise), after RLGR and tree, after
 * the top-down pass, end, and the hardware id; device-wide 100 MHz clock
 * (s_memrealtime). */
#ifdef PVVC_PHASE_PROFILE
__device__ uint64_t* g_prof = nullptr;
#define PVVC_STAMP(item, k)                                              \
    do {                                                                 \
        if (g_prof && threadIdx.x == 0) g_prof[8ull * (item) + (k)] = wall_clock64(); \
    } while (0)
#else
#define PVVC_STAMP(item, k) \
    do {                    \
    } while (0)
#endif

/* Inverse hierarchical transform of one node (octree_base.cc:84-100,
 * 186-206): child weights cw[i] (0 = empty child), the node's DC, and its
 * ACs at reversed stream positions r, r+1, ... (coef is planar Y|U|V of n
 * coefficients in stream order).  Writes the 8 child DCs (meaningful for
 * occupied children).  FP32 in the reference's operation order; where one
 * side of a pair is empty the transform is the identity on the other side. */
/* Haar butterfly factors for small weights: entry (w0 << 6) | w1 holds
 * (sqrtf(w0) / sqrtf(w0 + w1), sqrtf(w1) / sqrtf(w0 + w1)) for 1 <= w0, w1 < 64,
 * written once by raht_table_kernel with the very operations node_inverse
 * uses for larger weights, so a lookup equals the computation bit for bit. */
constexpr uint32_t kRahtTab = 64;
__device__ float2 g_raht_tab[kRahtTab * kRahtTab];

__device__ __forceinline__ float2 haar_factors(uint32_t w0, uint32_t w1) {
    const float base = sqrtf(static_cast<float>(w0 + w1));
    return make_float2(sqrtf(static_cast<float>(w0)) / base, sqrtf(static_cast<float>(w1)) / base);
}

__global__ void raht_table_kernel() {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kRahtTab * kRahtTab) return;
    const uint32_t w0 = i / kRahtTab, w1 = i % kRahtTab;
    g_raht_tab[i] = (w0 && w1) ? haar_factors(w0, w1) : make_float2(0.0f, 0.0f);
}

/* Haar factors of weight pairs below kSmallTab, copied per workgroup into LDS:
 * every pair of a node whose children are leaves has weights of at most 4. */
constexpr uint32_t kSmallTab = 8;

__device__ __forceinline__ void node_inverse(const float2* small_tab, const uint32_t cw[8], float dY, float dU, float dV,
                                             const float* coef,
                                             uint32_t n, uint32_t r, float oY[8], float oU[8], float oV[8]) {
    uint32_t w[16];
#pragma unroll
    for (int i = 0; i < 8; ++i) w[8 + i] = cw[i];
    w[4] = w[8] + w[12];
    w[5] = w[9] + w[13];
    w[6] = w[10] + w[14];
    w[7] = w[11] + w[15];
    w[2] = w[4] + w[6];
    w[3] = w[5] + w[7];
    float cY[16], cU[16], cV[16];
    cY[1] = dY;
    cU[1] = dU;
    cV[1] = dV;
    constexpr int NW[8][2] = {{0, 0}, {2, 3}, {4, 6}, {5, 7}, {8, 12}, {9, 13}, {10, 14}, {11, 15}};
#pragma unroll
    for (int idx = 1; idx < 8; ++idx) {
        const int ia = NW[idx][0], ib = NW[idx][1];
        const uint32_t w0 = w[ia], w1 = w[ib];
        if (w0 != 0 && w1 != 0) {
            const float hY = coef[n - 1 - r], hU = coef[2 * n - 1 - r], hV = coef[3 * n - 1 - r];
            ++r;
            const float2 t = (w0 < kSmallTab && w1 < kSmallTab) ? small_tab[w0 * kSmallTab + w1]
                             : (w0 < kRahtTab && w1 < kRahtTab)   ? g_raht_tab[w0 * kRahtTab + w1]
                                                                  : haar_factors(w0, w1);
            const float ta = t.x, tc = t.y;
            cY[ia] = cY[idx] * ta + hY * (-tc);
            cY[ib] = cY[idx] * tc + hY * ta;
            cU[ia] = cU[idx] * ta + hU * (-tc);
            cU[ib] = cU[idx] * tc + hU * ta;
            cV[ia] = cV[idx] * ta + hV * (-tc);
            cV[ib] = cV[idx] * tc + hV * ta;
        } else {
            cY[ia] = cY[idx];
            cU[ia] = cU[idx];
            cV[ia] = cV[idx];
            cY[ib] = cY[idx];
            cU[ib] = cU[idx];
            cV[ib] = cV[idx];
        }
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        oY[i] = cY[8 + i];
        oU[i] = cU[8 + i];
        oV[i] = cV[8 + i];
    }
}

struct Arena {
    uint8_t* lds;
    uint32_t lds_bytes;
    uint32_t lds_off;
    uint8_t* gws;
    uint64_t gws_bytes;
    uint64_t g_off;
    __device__ void* take(uint64_t bytes) {
        bytes = (bytes + 15u) & ~uint64_t(15);
        if (lds_off + bytes <= lds_bytes) {
            void* p = lds + lds_off;
            lds_off += static_cast<uint32_t>(bytes);
            return p;
        }
        if (g_off + bytes <= gws_bytes) {
            void* p = gws + g_off;
            g_off += bytes;
            return p;
        }
        return nullptr;
    }
    __device__ void* take_global(uint64_t bytes) {
        bytes = (bytes + 255u) & ~uint64_t(255);
        if (g_off + bytes > gws_bytes) return nullptr;
        void* p = gws + g_off;
        g_off += bytes;
        return p;
    }
};

/* pre_coef (optional): coefficients of the launch's items already decoded by
 * lane_rlgr_kernel, at 3 * stage_off (planar, stream order), with their RLGR
 * statuses in pre_status[item]; the slice then skips its own RLGR. */
template <int WAVES>
__device__ __forceinline__ int decode_slice(SliceArgs a, SlotDev* slot, Arena ar, const Sink& sink, LevelMeta& meta,
                            uint32_t launch_id, bool allow_wait, uint32_t item, const float* pre_coef = nullptr,
                            const int* pre_status = nullptr) {
    PVVC_STAMP(item, 0);
    constexpr uint32_t NT = 64u * WAVES;
    const uint32_t tid = threadIdx.x;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6); /* uniform per wavefront */
    const bool predict = (a.type & PVVC_TYPE_PREDICT) != 0;
    const bool wait = predict && allow_wait && a.wait; /* deferred reference add */
    if (!wait) a.stage = nullptr;
    int st = PVVC_OK;

    /* ---- validation (SetSlice, invert_raht_octree.cc:25-40) ---- */
    if (!(a.type & PVVC_TYPE_VALID)) {
        st = PVVC_ERR_BAD_SLICE;
    } else if (!predict) {
        if (!a.geo || a.geo_size < 25) st = PVVC_ERR_BAD_SLICE;
        else if (a.n > slot->n_cap || a.geo_size > slot->geo_cap) st = PVVC_ERR_CAPACITY;
    } else if (!wait) {
        /* a predictive slice decodes on its patch's intra tree: the reference
         * builds the tree only for intra slices (invert_raht_octree.cc:71-78)
         * and ignores any geometry bytes a predictive slice carries */
        const uint32_t sn = slot->n;
        if (sn == 0) st = PVVC_ERR_EMPTY_REFERENCE;
        else if (sn != a.n) st = PVVC_ERR_UNMATCHED_SIZE;
        a.geo = slot->geo;
        a.geo_size = slot->geo_size;
    } else if (a.n > slot->n_cap || !a.geo || a.geo_size < 25) {
        /* deferred: the intra slice decodes in this launch; validity against
         * it is settled by the finish pass */
        st = PVVC_ERR_CAPACITY;
    }
    if (st == PVVC_OK && a.n == 0) st = PVVC_ERR_UNMATCHED_SIZE;

    float* coef = nullptr;
    uint32_t* fc = nullptr;
    uint32_t* wt = nullptr;
    if (st == PVVC_OK) {
        const uint64_t imax = a.geo_size - 25u;
        /* LDS first, then the global workspace.  The first-child and weight
         * arrays (8 B per internal node) come first: the top-down pass reads
         * them on its chain of dependent loads (occupancy -> first child ->
         * child weights -> Haar factors), while coefficients are independent
         * loads at the end of that chain. */
        fc = static_cast<uint32_t*>(ar.take(4ull * imax + 4));
        wt = static_cast<uint32_t*>(ar.take(4ull * imax + 4));
        coef = pre_coef ? const_cast<float*>(pre_coef) + 3ull * a.stage_off : static_cast<float*>(ar.take(12ull * a.n));
        if (!coef || !fc || !wt) st = PVVC_ERR_CAPACITY;
    }

    /* ---- R (RLGR) and T (tree) ---- */
    if (st == PVVC_OK) {
        if (WAVES == 1) {
            /* RLGR first: a launch's first slices then all start on the
             * scalar unit, instead of all building their trees (vector and
             * memory work) while it idles; a tree that fails wins the status */
            const int rs = pre_coef ? pre_status[item] : rlgr_decode_wave(a.col, a.col_size, 3u * a.n, a.qp, coef);
            PVVC_STAMP(item, 1);
            const int ts = tree_structure(a.geo, a.geo_size, a.n, fc, meta);
            __syncthreads();
            if (ts == PVVC_OK) tree_weights(a.geo, fc, wt, meta);
            PVVC_STAMP(item, 2);
            st = ts != PVVC_OK ? ts : rs;
        } else {
            if (wave == 0) {
                int rs = pre_coef ? pre_status[item] : rlgr_decode_wave(a.col, a.col_size, 3u * a.n, a.qp, coef);
                if (lane_id() == 0) meta.rlgr_status = rs;
            } else if (wave == 1) {
                int ts = tree_structure(a.geo, a.geo_size, a.n, fc, meta);
                __threadfence_block();
                if (ts == PVVC_OK) tree_weights(a.geo, fc, wt, meta);
                if (lane_id() == 0) meta.tree_status = ts;
            }
            __syncthreads();
            st = meta.tree_status != PVVC_OK ? meta.tree_status : meta.rlgr_status;
        }
    }
    __syncthreads();

    PVVC_STAMP(item, 3);
    __shared__ float2 s_tab[kSmallTab * kSmallTab];
    for (uint32_t i = tid; i < kSmallTab * kSmallTab; i += NT)
        s_tab[i] = g_raht_tab[(i / kSmallTab) * kRahtTab + i % kSmallTab];
    __syncthreads();
    /* ---- H + M: top-down inverse RAHT, centres, leaf output ----
     * One pass per internal level: a thread per node inverts its transform
     * and writes each child's YUV and centre into the next level's buffer;
     * the leaves, the last level, are then emitted one per thread, in order. */
    const uint32_t h = (st == PVVC_OK) ? meta.height : 0u;
    float* pp = nullptr;
    /* every level holds at most n nodes: internal nodes have children and the
     * leaf count equals n (both checked by tree_structure) */
    const uint32_t wl = a.n;
    if (st == PVVC_OK && h > 1) {
        pp = static_cast<float*>(ar.take(48ull * wl));
        if (!pp) st = PVVC_ERR_CAPACITY;
    }
    if (st == PVVC_OK) {
        const uint32_t n = a.n;
        float* ref = slot->ref;
        const uint32_t rs = slot->n_cap;
        /* root centre / range: LoadTreeCore (octree_base.cc:33-48) */
        float hdr[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            const uint32_t wv = (static_cast<uint32_t>(a.geo[4 * i]) << 24) | (static_cast<uint32_t>(a.geo[4 * i + 1]) << 16) |
                                (static_cast<uint32_t>(a.geo[4 * i + 2]) << 8) | static_cast<uint32_t>(a.geo[4 * i + 3]);
            hdr[i] = __uint_as_float(wv);
        }
        if (h == 1) {
            if (tid == 0) emit_leaf(0, hdr[0], hdr[1], hdr[2], coef[n - 1], coef[2 * n - 1], coef[3 * n - 1], a, ref, rs,
                                    predict, sink);
        } else {
            float* bufA = pp;
            float* bufB = pp + 6 * wl;
            float rx = hdr[3], ry = hdr[4], rz = hdr[5];
            for (uint32_t l = 0; l + 1 < h; ++l) {
                const uint32_t cnt = meta.cnt[l], base = meta.base[l];
                const uint32_t cbase = meta.base[l + 1];
                const bool leaf_level = (l + 2 == h); /* children are leaves (weight 1) */
                /* SubSpaceCenter: child range = range / 2, offset = child range / 2 */
                const float sx = rx / 2.0f, sy = ry / 2.0f, sz = rz / 2.0f;
                const float ox = sx / 2.0f, oy = sy / 2.0f, oz = sz / 2.0f;
                float* cur = (l & 1) ? bufB : bufA;
                float* nxt = (l & 1) ? bufA : bufB;
                for (uint32_t j = tid; j < cnt; j += NT) {
                    const uint32_t occ = a.geo[25 + base + j];
                    const uint32_t f = fc[base + j];
                    float dY, dU, dV, cx, cy, cz;
                    if (l == 0) {
                        dY = coef[n - 1];
                        dU = coef[2 * n - 1];
                        dV = coef[3 * n - 1];
                        cx = hdr[0];
                        cy = hdr[1];
                        cz = hdr[2];
                    } else {
                        dY = cur[j];
                        dU = cur[wl + j];
                        dV = cur[2 * wl + j];
                        cx = cur[3 * wl + j];
                        cy = cur[4 * wl + j];
                        cz = cur[5 * wl + j];
                    }
                    uint32_t cw[8];
                    uint32_t rank = 0;
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        const bool has = (occ >> (7 - i)) & 1u;
                        cw[i] = has ? (leaf_level ? 1u : wt[cbase + f + rank]) : 0u;
                        rank += has ? 1u : 0u;
                    }
                    float oY[8], oU[8], oV[8];
                    node_inverse(s_tab, cw, dY, dU, dV, coef, n, cnt + f - j, oY, oU, oV);
                    uint32_t k = 0;
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        if (!((occ >> (7 - i)) & 1u)) continue;
                        const uint32_t c = f + k;
                        ++k;
                        nxt[c] = oY[i];
                        nxt[wl + c] = oU[i];
                        nxt[2 * wl + c] = oV[i];
                        nxt[3 * wl + c] = (i & 4) ? cx - ox : cx + ox;
                        nxt[4 * wl + c] = (i & 2) ? cy - oy : cy + oy;
                        nxt[5 * wl + c] = (i & 1) ? cz - oz : cz + oz;
                    }
                }
                __syncthreads();
                rx = sx;
                ry = sy;
                rz = sz;
            }
            /* the leaves were written by the pass of level h - 2; each thread
             * issues the loads of four leaves (and their reference colours)
             * before any of their stores */
            const float* lv = ((h - 2) & 1) ? bufA : bufB;
            const bool load_ref = predict && a.stage == nullptr;
            for (uint32_t c0 = tid; c0 < n; c0 += 4 * NT) {
                float v[4][6], r[4][3];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint32_t c = c0 + q * NT;
                    if (c < n) {
#pragma unroll
                        for (int e = 0; e < 6; ++e) v[q][e] = lv[e * wl + c];
#pragma unroll
                        for (int e = 0; e < 3; ++e) r[q][e] = load_ref ? ref[e * rs + c] : 0.0f;
                    }
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint32_t c = c0 + q * NT;
                    if (c < n)
                        emit_leaf_ref(c, v[q][3], v[q][4], v[q][5], v[q][0], v[q][1], v[q][2], r[q][0], r[q][1], r[q][2],
                                      a, ref, rs, predict, sink);
                }
            }
        }
    }

    PVVC_STAMP(item, 4);
    /* ---- intra: keep geometry for the GoP, publish ---- */
    __syncthreads();
    if (!predict && (a.type & PVVC_TYPE_VALID)) {
        if (st == PVVC_OK) {
            for (uint32_t i = tid; i < a.geo_size; i += NT) slot->geo[i] = a.geo[i];
        }
        /* read by later launches only (kernel boundaries order them) */
        __syncthreads();
        if (tid == 0) {
            slot->n = st == PVVC_OK ? a.n : 0u;
            slot->geo_size = st == PVVC_OK ? a.geo_size : 0u;
            slot->status = st;
        }
    } else if (tid == 0) {
        slot->status = st;
    }
    PVVC_STAMP(item, 5);
#ifdef PVVC_PHASE_PROFILE
    /* where the item ran: HW_ID (wave, SIMD, CU, SH, SE) and the XCC id */
    if (g_prof && threadIdx.x == 0)
        g_prof[8ull * item + 6] = static_cast<uint64_t>(__builtin_amdgcn_s_getreg((31 << 11) | 4)) |
                                  (static_cast<uint64_t>(__builtin_amdgcn_s_getreg((31 << 11) | 20)) << 32);
#endif
    return st;
}

/* ======================================================================== */
/* Kernels                                                                   */
/* ======================================================================== */

extern __shared__ __attribute__((aligned(16))) uint8_t g_smem[];

/* Drop-in: one workgroup per slice of the reference's SoA device arrays
 * (CUDA2OpenGL.cu:37-66 processes the same arrays, one thread per slice). */
template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void dropin_kernel(Points* out, const int* __restrict__ inner_offset,
                                                          const int* __restrict__ index, const uint8_t* __restrict__ type,
                                                          float* const* __restrict__ mv, const uint32_t* __restrict__ size,
                                                          const uint8_t* __restrict__ qp,
                                                          uint8_t* const* __restrict__ geometry,
                                                          const uint32_t* __restrict__ geometry_size,
                                                          uint8_t* const* __restrict__ color,
                                                          const uint32_t* __restrict__ color_size, SlotDev* slots,
                                                          int n_slots, int patch_size, uint32_t lds_bytes,
                                                          int white_small, int* err_flag) {
    __shared__ LevelMeta meta;
    const int s = blockIdx.x;
    if (s >= patch_size) return;
    const int pidx = index[s];
    /* err_flag keeps the first failure's status until the host reads it */
    if (pidx < 0 || pidx >= n_slots) {
        if (threadIdx.x == 0) atomicCAS(err_flag, 0, static_cast<int>(PVVC_ERR_INVALID_ARG));
        return;
    }
    SlotDev* slot = slots + pidx;
    SliceArgs a;
    a.geo = geometry[s];
    a.geo_size = geometry_size[s];
    a.col = color[s];
    a.col_size = color_size[s];
    a.n = size[s];
    a.type = type[s];
    a.qp = qp[s];
    a.wait = 0;
    a.stage = nullptr;
    a.out_off = static_cast<uint64_t>(static_cast<int64_t>(inner_offset[s]));
    a.stage_off = 0;
    const float* m = mv[s];
#pragma unroll
    for (int i = 0; i < 12; ++i) a.mv[i] = m[i];
    Sink sink{nullptr, nullptr, out, PVVC_OUT_POINTS, white_small};
    Arena ar{g_smem, lds_bytes, 0, slot->ws, slot->ws_cap, 0};
    int st = decode_slice<WAVES>(a, slot, ar, sink, meta, 0u, false, static_cast<uint32_t>(s));
    if (st != PVVC_OK && threadIdx.x == 0) atomicCAS(err_flag, 0, st);
}

__device__ __forceinline__ SliceArgs item_args(const ItemDev* __restrict__ it) {
    SliceArgs a;
    a.geo = it->geo;
    a.geo_size = it->geo_size;
    a.col = it->col;
    a.col_size = it->col_size;
    a.n = it->n;
    a.type = it->type;
    a.qp = it->qp;
    a.wait = it->wait;
    a.out_off = it->out_off;
    a.stage_off = it->stage_off;
    a.stage = nullptr;
#pragma unroll
    for (int i = 0; i < 12; ++i) a.mv[i] = it->mv[i];
    return a;
}

/* Native per-frame: one workgroup per item of one frame, per-slot workspace. */
template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void frame_kernel(const ItemDev* __restrict__ items, uint32_t first,
                                                         uint32_t count, SlotDev* slots, Sink sink,
                                                         uint32_t lds_bytes, int* __restrict__ status) {
    __shared__ LevelMeta meta;
    const uint32_t i = first + blockIdx.x;
    if (blockIdx.x >= count) return;
    const ItemDev* it = items + i;
    SlotDev* slot = slots + it->patch;
    Arena ar{g_smem, lds_bytes, 0, slot->ws, slot->ws_cap, 0};
    int st = decode_slice<WAVES>(item_args(it), slot, ar, sink, meta, 0u, false, i);
    if (threadIdx.x == 0) status[i] = st;
}

/* Batched persistent: workgroups pop items from a length-sorted queue. */
template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void persistent_kernel(const ItemDev* __restrict__ items,
                                                              const uint32_t* __restrict__ order, uint32_t count,
                                                              uint32_t* queue, SlotDev* slots, Sink sink,
                                                              uint32_t lds_bytes, uint8_t* slab, uint64_t slab_bytes,
                                                              uint32_t launch_id, int* __restrict__ status,
                                                              float* stage, uint32_t chunk, uint32_t prio_items,
                                                              const float* pre_coef, const int* pre_status) {
    __shared__ LevelMeta meta;
    __shared__ uint32_t next;
    uint8_t* ws = slab + static_cast<uint64_t>(blockIdx.x) * slab_bytes;
    /* a workgroup's first entries are its own, so every workgroup starts
     * decoding without an atomic round trip; the shared counter hands out the
     * entries after the first grid * chunk, one atomic per `chunk` consecutive
     * entries.  Workgroup b takes first-wave group b, except that the first
     * quarter of the grid takes its quarter of the (longest-first) queue in
     * reverse.  The dispatcher deals a grid of 16 workgroups per CU out in four
     * rounds of one wavefront per SIMD, and mostly puts workgroups b, b + G/4,
     * b + G/2 and b + 3G/4 (G = grid) on one SIMD; the first quarter holds the
     * longest slices and the steepest fall in length, so reversing it pairs a
     * SIMD's longest first slice with short ones from the later rounds (c1
     * GoP: a SIMD's first four slices at most 5 923 points against 6 470, the
     * launch 0.567 -> 0.558 ms) */
    const uint32_t first_end = gridDim.x * chunk;
    bool first = true;
    for (;;) {
        if (first) {
            const uint32_t quarter = gridDim.x >> 2;
            const uint32_t b = (gridDim.x & 3u) == 0u && blockIdx.x < quarter ? quarter - 1u - blockIdx.x : blockIdx.x;
            next = b * chunk;
            first = false;
        } else {
            if (threadIdx.x == 0) next = first_end + atomicAdd(queue, chunk);
            __syncthreads();
        }
        const uint32_t q0 = next;
        __syncthreads();
        if (q0 >= count) break;
        const uint32_t qe = q0 + chunk < count ? q0 + chunk : count;
        for (uint32_t q = q0; q < qe; ++q) {
            const uint32_t i = order[q];
            const ItemDev* it = items + i;
            SlotDev* slot = slots + it->patch;
            Arena ar{g_smem, lds_bytes, 0, ws, slab_bytes, 0};
            SliceArgs a = item_args(it);
            a.stage = stage;
            /* wave priority falls with queue rank: the first prio_items
             * (longest) slices at 3, the rest of the queue in thirds at 2, 1
             * and 0, so that issue arbitration on a CU favours the slice with
             * more work left over the oldest wavefront */
            const uint32_t rest = count > prio_items ? count - prio_items : 0u;
            const int lvl = prio_items == 0u                 ? 0
                            : q < prio_items                 ? 3
                            : q < prio_items + rest / 3      ? 2
                            : q < prio_items + 2 * (rest / 3) ? 1
                                                             : 0;
            if (lvl == 3) __builtin_amdgcn_s_setprio(3);
            else if (lvl == 2) __builtin_amdgcn_s_setprio(2);
            else if (lvl == 1) __builtin_amdgcn_s_setprio(1);
            int st = decode_slice<WAVES>(a, slot, ar, sink, meta, launch_id, true, i, pre_coef, pre_status);
            if (lvl) __builtin_amdgcn_s_setprio(0);
            if (threadIdx.x == 0) status[i] = st;
            __syncthreads();
        }
    }
    /* queue[1] counts workgroups out; the last one re-arms both counters for
     * the next launch: every other workgroup has made its final pop by then
     * (its exit increment follows the pop that returned past `count`) */
    if (threadIdx.x == 0 && atomicAdd(queue + 1, 1u) == gridDim.x - 1) {
        atomicExch(queue, 0u);
        atomicExch(queue + 1, 0u);
    }
}

/* RLGR of short slices, one lane per slice (rlgr_decode_lane_from); the
 * rest of their decode runs in persistent_kernel with these coefficients.
 * Wavefront w takes list[64 w, 64 w + 64): it zeroes their coefficient
 * ranges together (coalesced), then each lane decodes its own slice into
 * coef_all at 3 * stage_off, reading its colour bytes through an LDS ring
 * (LaneRing), and leaves the status in lane_status[item]. */
__global__ __launch_bounds__(64) void lane_rlgr_kernel(const ItemDev* __restrict__ items,
                                                       const uint32_t* __restrict__ list, uint32_t count,
                                                       float* __restrict__ coef_all, int* __restrict__ lane_status) {
    __shared__ uint32_t ring[(kLaneRing + 1u) * 64u];
    const uint32_t base = blockIdx.x * 64u;
    if (base >= count) return;
    const uint32_t lane = threadIdx.x;
    const uint32_t cnt = count - base < 64u ? count - base : 64u;
    for (uint32_t j = 0; j < cnt; ++j) {
        const ItemDev* it = items + list[base + j];
        float* c = coef_all + 3ull * it->stage_off;
        const uint32_t n3 = 3u * it->n;
        for (uint32_t x = lane; x < n3; x += 64u) c[x] = 0.0f;
    }
    __threadfence_block(); /* the zeroes are written before any lane's coefficients */
    if (lane >= cnt) return;
    const uint32_t i = list[base + lane];
    const ItemDev* it = items + i;
    /* a slice without colour bytes points the reader at its own descriptor,
     * which is readable */
    const uint8_t* src = it->col_size ? it->col : reinterpret_cast<const uint8_t*>(it);
    const uintptr_t a = reinterpret_cast<uintptr_t>(src);
    const uint32_t skip = 8u * static_cast<uint32_t>(a & 3u);
    const uint32_t end = skip + 8u * it->col_size;
    const LaneRing r{reinterpret_cast<const uint32_t*>(a & ~uintptr_t(3)), end, end ? (end - 1u) / 32u : 0u, 0u,
                     ring + lane};
    lane_status[i] = rlgr_decode_lane_from(r, skip, it->col_size, 3u * it->n, it->qp, coef_all + 3ull * it->stage_off);
}

/* Zeroes a queue's counters once, after allocation. */
__global__ void reset_counter(uint32_t* q) {
    q[0] = 0;
    q[1] = 0;
}

/* Finish pass of deferred predictive items: the intra slice of the same
 * patch decoded in the launch before this one; add its colours
 * (invert_raht_octree.cc:304-308), convert (common.h:154-158) and write. */
__global__ __launch_bounds__(256) void finish_kernel(const ItemDev* __restrict__ items,
                                                     const uint32_t* __restrict__ list, uint32_t count,
                                                     const SlotDev* __restrict__ slots, Sink sink,
                                                     const float* __restrict__ stage, int* __restrict__ status) {
    if (blockIdx.x >= count) return;
    const uint32_t i = list[blockIdx.x];
    const ItemDev* it = items + i;
    const SlotDev* slot = slots + it->patch;
    const uint32_t n = it->n;
    const uint32_t sn = slot->n;
    int st = status[i];
    if (sn == 0) st = PVVC_ERR_EMPTY_REFERENCE; /* the intra slice failed */
    else if (sn != n) st = PVVC_ERR_UNMATCHED_SIZE;
    if (st != PVVC_OK) {
        if (threadIdx.x == 0) status[i] = st;
        return;
    }
    const float* ref = slot->ref;
    const uint32_t rs = slot->n_cap;
    for (uint32_t c = threadIdx.x; c < n; c += blockDim.x) {
        const uint64_t o = it->out_off + c;
        const float* sp = stage + 3 * (static_cast<uint64_t>(it->stage_off) + c);
        const float Y = sp[0] + ref[c];
        const float U = sp[1] + ref[rs + c];
        const float V = sp[2] + ref[2 * rs + c];
        const float r = Y + 1.4020f * (V - 128.0f);
        const float g = Y - 0.3441f * (U - 128.0f) - 0.7141f * (V - 128.0f);
        const float b = Y + 1.7720f * (U - 128.0f);
        if (sink.mode == PVVC_OUT_NATIVE) {
            uint8_t* pc = sink.rgb + 3 * o;
            pc[0] = to_u8(r);
            pc[1] = to_u8(g);
            pc[2] = to_u8(b);
        } else {
            write_points_rgb(sink.pts + o, r, g, b, n, sink.white_small);
        }
    }
}

/* ======================================================================== */
/* Forward path: RAHT (raht_octree.cc:189-247) + quantisation               */
/* (GoP_encoding.cc:172-205) + RLGR encoder (entropy_codec.cc:65-159)       */
/* ======================================================================== */

/* One colour stream to encode: the patch's octree bytes, its n leaf colours
 * (YUV, leaf order = the decoder's output order) and the quantisation step. */
struct EncJobDev {
    const uint8_t* geo;
    const float* yuv;      /* n x 3, AoS */
    uint8_t* out;          /* out_cap bytes, 4-byte aligned */
    uint8_t* ws;           /* workspace, ws_cap bytes */
    uint64_t ws_cap;
    uint32_t geo_size;
    uint32_t n;
    uint32_t out_cap;
    uint32_t qp;
    uint32_t idx;          /* the job's index in the caller's table */
};

/* HierarchicalTransform (octree_base.cc:65-82) of one node: child DCs in
 * iY/iU/iV (0 where absent), child weights cw; returns the DC and the coded
 * ACs in NodeWeight order, FP32 in the reference's operation order
 * (HaarTransform, octree_base.cc:164-184). */
__device__ __forceinline__ uint32_t node_forward(const uint32_t cw[8], const float iY[8], const float iU[8],
                                                 const float iV[8], float& dY, float& dU, float& dV, float hY[8],
                                                 float hU[8], float hV[8]) {
    uint32_t w[16];
#pragma unroll
    for (int i = 0; i < 8; ++i) w[8 + i] = cw[i];
    w[4] = w[8] + w[12];
    w[5] = w[9] + w[13];
    w[6] = w[10] + w[14];
    w[7] = w[11] + w[15];
    w[2] = w[4] + w[6];
    w[3] = w[5] + w[7];
    float rY[16], rU[16], rV[16];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        rY[8 + i] = iY[i];
        rU[8 + i] = iU[i];
        rV[8 + i] = iV[i];
    }
    constexpr int NW[8][2] = {{0, 0}, {2, 3}, {4, 6}, {5, 7}, {8, 12}, {9, 13}, {10, 14}, {11, 15}};
#pragma unroll
    for (int i = 7; i >= 1; --i) {
        const int a = NW[i][0], b = NW[i][1];
        const uint32_t w0 = w[a], w1 = w[b];
        /* a pair with no points gets factors (0, 0): its inputs are +0, so the
         * products below give +0, as the reference's skip does, without a
         * branch (which would keep these arrays in scratch memory) */
        float2 t = make_float2(0.0f, 0.0f);
        if (w0 != 0 || w1 != 0)
            t = (w0 && w1 && w0 < kRahtTab && w1 < kRahtTab) ? g_raht_tab[w0 * kRahtTab + w1] : haar_factors(w0, w1);
        const float t00 = t.x, t01 = t.y, t10 = -t.y, t11 = t.x;
        rY[i] = rY[a] * t00 + rY[b] * t01;
        rU[i] = rU[a] * t00 + rU[b] * t01;
        rV[i] = rV[a] * t00 + rV[b] * t01;
        hY[i] = rY[a] * t10 + rY[b] * t11;
        hU[i] = rU[a] * t10 + rU[b] * t11;
        hV[i] = rV[a] * t10 + rV[b] * t11;
    }
    dY = rY[1];
    dU = rU[1];
    dV = rV[1];
    /* bit idx: pair idx (1..7) is coded, its AC is hY/hU/hV[idx]; the ACs go
     * to the stream in idx order (the caller keeps register indices static) */
    uint32_t coded = 0;
#pragma unroll
    for (int idx = 1; idx < 8; ++idx)
        if (w[NW[idx][0]] != 0 && w[NW[idx][1]] != 0) coded |= 1u << idx;
    return coded;
}

/* round(c / qp) (GoP_encoding.cc:174-181); false outside int32. */
__device__ __forceinline__ bool quantise(float c, float qp, int32_t& q) {
    const float r = roundf(c / qp);
    if (!(r >= -2147483648.0f && r < 2147483648.0f)) return false;
    q = static_cast<int32_t>(r);
    return true;
}

/* RLGR encoder (entropy_codec.cc:22-159), one wavefront per stream, 64
 * symbols per step:
 *  1. lane j loads symbol j of the step and maps it to u (Sign2Unsign);
 *  2. a wave-uniform pass walks the step with the adaptive state (k_P, k_RP,
 *     m) and leaves in each lane what its symbol writes: a Golomb-Rice code
 *     with parameter k_R (k = 0, or the end of a zero run, which first writes
 *     '0' and m in k bits), or one '1' bit (the zero at which m reaches 2^k),
 *     or nothing (any other zero of a run).  Zero runs are crossed by whole
 *     runs, found in the ballot of nonzero symbols, one step per '1' bit;
 *  3. the lanes' code lengths are scanned into bit offsets, and each lane ORs
 *     its bits into an LDS window of the output (MSB first);
 *  4. the window's whole words are stored (byte-swapped, so the stream is
 *     big-endian as the reference writes it) and its open word carried.
 * Byte-identical to the reference writer, including its end rule (a stream
 * whose last symbol leaves k > 0 and u = 0 gets '0' and m in k_P / L bits),
 * the 32-ones escape and its k > 32 failure. */
constexpr uint32_t kEncWindow = 256; /* words: one step writes at most 64 x 97 bits, plus the open word */

/* ORs `len` <= 33 bits of `v` (v < 2^len) at bit `pos` of win (MSB first). */
__device__ __forceinline__ void enc_or_bits(uint32_t* win, uint32_t pos, uint64_t v, uint32_t len) {
    if (len == 0 || v == 0) return;
    const uint64_t x = (v << (64u - len)) >> (pos & 31u);
    const uint32_t w = pos >> 5;
    const uint32_t hi = static_cast<uint32_t>(x >> 32), lo = static_cast<uint32_t>(x);
    if (hi) atomicOr(win + w, hi);
    if (lo) atomicOr(win + w + 1, lo);
}

/* Exclusive wave scan of v < 128 by bit planes; *total gets the sum. */
__device__ __forceinline__ uint32_t wave_scan_128(uint32_t v, uint32_t* total) {
    uint32_t r = 0, t = 0;
#pragma unroll
    for (int b = 0; b < 7; ++b) {
        const uint64_t mb = __ballot((v >> b) & 1u);
        r += prefix_lt(mb) << b;
        t += static_cast<uint32_t>(__popcll(mb)) << b;
    }
    *total = t;
    return r;
}

/* Stores window words [0, count) at output word w0 (those below cap_words). */
__device__ __forceinline__ void enc_store(const uint32_t* win, uint32_t count, uint8_t* out, uint32_t w0,
                                          uint32_t cap_words) {
    uint32_t* o = reinterpret_cast<uint32_t*>(out);
    for (uint32_t i = lane_id(); i < count; i += 64u)
        if (w0 + i < cap_words) o[w0 + i] = __builtin_bswap32(win[i]);
}

/* k_RP after a Golomb-Rice code of x with parameter kR (entropy_codec.cc:
 * 92-100): p = x >> kR; p ? min(k_RP + p - 1, 128) : max(k_RP - 2, 0).  32-bit
 * scalar arithmetic (p is capped at 129, which saturates the same way). */
__device__ __forceinline__ uint32_t enc_adapt_krp(uint32_t kRP, uint32_t x, uint32_t kR) {
    const uint32_t p = static_cast<uint32_t>(static_cast<uint64_t>(x) >> kR);
    const int32_t up = static_cast<int32_t>(kRP + (p < 129u ? p : 129u)) - 1;
    const int32_t down = static_cast<int32_t>(kRP) - 2;
    return p ? static_cast<uint32_t>(up < 128 ? up : 128) : static_cast<uint32_t>(down > 0 ? down : 0);
}

/* Lane descriptor bits of one step. */
constexpr uint32_t kEncGr = 1u << 31;    /* a Golomb-Rice code of u (minus 1 at a run end), k_R in bits 0-5 */
constexpr uint32_t kEncRunEnd = 1u << 30; /* '0' and m in k bits first, k in bits 8-13 */
constexpr uint32_t kEncOne = 1u << 29;   /* a single '1' */

/* win: kEncWindow words of LDS, owned by this wavefront.  Returns the byte
 * length, or -1 past out_cap or on k > 32. */
__device__ __forceinline__ int64_t rlgr_encode_wave(const int32_t* __restrict__ sym, uint32_t total, uint8_t* out,
                                                    uint32_t out_cap, uint32_t* win) {
    const uint32_t lane = lane_id();
    const uint32_t cap_words = out_cap / 4u;
    for (uint32_t i = lane; i < kEncWindow; i += 64u) win[i] = 0u;
    __syncthreads();
    uint32_t kP = 0, kRP = 8, m = 0;
    bool last_run = false, last_u0 = false; /* the end rule's k != 0 and u == 0 */
    uint64_t bitpos = 0;                      /* bits written before this step */
    for (uint32_t base = 0; base < total; base += 64u) {
        const uint32_t lim = total - base < 64u ? total - base : 64u;
        const int64_t d = base + lane < total ? sym[base + lane] : 0;
        const uint32_t u = static_cast<uint32_t>(d < 0 ? (static_cast<uint64_t>(-d) << 1) - 1u
                                                       : static_cast<uint64_t>(d) << 1);
        const uint64_t nz = __ballot(u != 0u);
        uint32_t desc = 0, mm = 0;
        uint32_t j = 0;
        while (j < lim) {
            uint32_t k = kP >> 2;
            if (k > 32u) return -1;
            if (k == 0u) {
                /* k = 0: every symbol is a Golomb-Rice code */
                uint32_t uj = __builtin_amdgcn_readlane(u, j);
                do {
                    const uint32_t un = __builtin_amdgcn_readlane(u, j + 1u < lim ? j + 1u : j);
                    const uint32_t kR = kRP >> 2;
                    desc = lane == j ? (kEncGr | kR) : desc;
                    kRP = enc_adapt_krp(kRP, uj, kR);
                    kP = uj ? (kP > 1u ? kP : 1u) - 1u : kP + 3u;
                    uj = un;
                    ++j;
                } while (j < lim && kP < 4u);
                last_run = false;
            } else if ((nz >> j) & 1u) {
                const uint32_t uj = __builtin_amdgcn_readlane(u, j) - 1u;
                const uint32_t kR = kRP >> 2;
                desc = lane == j ? (kEncGr | kEncRunEnd | kR | (k << 8)) : desc;
                mm = lane == j ? m : mm;
                kRP = enc_adapt_krp(kRP, uj, kR);
                kP = (kP > 1u ? kP : 1u) - 1u;
                m = 0;
                last_run = true;
                last_u0 = uj == 0u;
                ++j;
            } else {
                /* zeros [j, t): m counts them; each time it reaches 2^k the
                 * zero there writes '1', k_P += 2 and m restarts */
                const uint64_t rest = nz >> j;
                const uint32_t t = rest ? j + static_cast<uint32_t>(__builtin_ctzll(rest)) : lim;
                const uint32_t end = t < lim ? t : lim;
                uint32_t z = end - j;
                last_run = true;
                last_u0 = true;
                while (z) {
                    if (k > 32u) return -1;
                    const uint64_t need = (uint64_t{1} << k) - m;
                    if (z < need) {
                        m += z;
                        j += z;
                        z = 0;
                    } else {
                        const uint32_t at = j + static_cast<uint32_t>(need) - 1u;
                        desc = lane == at ? kEncOne : desc;
                        j += static_cast<uint32_t>(need);
                        z -= static_cast<uint32_t>(need);
                        m = 0;
                        kP += 2u;
                        k = kP >> 2;
                    }
                }
            }
        }
        /* this lane's bits: [run-end prefix][GR prefix][GR suffix] or a '1' */
        uint32_t l1 = 0, l2 = 0, l3 = 0;
        uint64_t v2 = 0, v3 = 0;
        if (desc & kEncOne) {
            l2 = 1;
            v2 = 1;
        } else if (desc & kEncGr) {
            const uint32_t kR = desc & 63u;
            const uint64_t x = (desc & kEncRunEnd) ? u - 1u : u;
            if (desc & kEncRunEnd) l1 = 1u + ((desc >> 8) & 63u);
            const uint64_t q = x >> kR;
            if (q < 32u) {
                l2 = static_cast<uint32_t>(q) + 1u;
                v2 = (uint64_t{1} << l2) - 2u;
                l3 = kR;
                v3 = kR >= 64u ? x : (x & ((uint64_t{1} << kR) - 1u));
            } else {
                l2 = 32;
                v2 = 0xffffffffull;
                l3 = 32;
                v3 = x & 0xffffffffull;
            }
        }
        uint32_t step_bits;
        const uint32_t off = wave_scan_128(l1 + l2 + l3, &step_bits);
        const uint32_t w0 = static_cast<uint32_t>(bitpos >> 5);
        const uint32_t rel = static_cast<uint32_t>(bitpos & 31u) + off;
        enc_or_bits(win, rel, mm, l1);
        enc_or_bits(win, rel + l1, v2, l2);
        enc_or_bits(win, rel + l1 + l2, v3, l3);
        __syncthreads(); /* one wavefront: the ORs have landed */
        bitpos += step_bits;
        const uint32_t done = static_cast<uint32_t>(bitpos >> 5) - w0; /* whole words of the window */
        if (done) {
            enc_store(win, done, out, w0, cap_words);
            const uint32_t open = win[done];
            __syncthreads();
            for (uint32_t i = lane; i <= done; i += 64u) win[i] = i == 0 ? open : 0u;
            __syncthreads();
        }
    }
    /* the window's word 0 is output word bitpos >> 5 */
    const uint32_t w_open = static_cast<uint32_t>(bitpos >> 5);
    if (total > 0 && last_run && last_u0) {
        const uint32_t kb = kP >> 2;
        if (lane == 0) enc_or_bits(win, static_cast<uint32_t>(bitpos & 31u) + 1u, m, kb); /* m < 2^kb */
        bitpos += 1u + kb;
    }
    bitpos = (bitpos + 7u) & ~uint64_t(7);
    __syncthreads();
    const uint64_t words = (bitpos + 31u) >> 5;
    enc_store(win, static_cast<uint32_t>(words - w_open), out, w_open, cap_words);
    return words > cap_words ? -1 : static_cast<int64_t>(bitpos >> 3);
}

/* Encodes one colour stream with one wavefront.  Returns the byte length or
 * -(status). */
__device__ __forceinline__ int64_t encode_job(const EncJobDev& jb, Arena ar, LevelMeta& meta, uint32_t* win) {
    const uint32_t tid = threadIdx.x;
    const uint32_t n = jb.n;
    if (!jb.geo || jb.geo_size < 25 || n == 0 || !jb.yuv || !jb.out) return -PVVC_ERR_BAD_SLICE;
    const uint64_t imax = jb.geo_size - 25u;
    int32_t* sym = static_cast<int32_t*>(ar.take(12ull * n));
    uint32_t* fc = static_cast<uint32_t*>(ar.take(4ull * imax + 4));
    uint32_t* wt = static_cast<uint32_t*>(ar.take(4ull * imax + 4));
    if (!sym || !fc || !wt) return -PVVC_ERR_CAPACITY;
    int ts = tree_structure(jb.geo, jb.geo_size, n, fc, meta);
    __syncthreads();
    if (ts != PVVC_OK) return -ts;
    tree_weights(jb.geo, fc, wt, meta);
    __syncthreads();
    const uint32_t h = meta.height;
    const float qp = static_cast<float>(jb.qp);
    bool range_ok = true;
    if (h == 1) { /* a single leaf: its colour is the DC */
        if (tid == 0) {
            int32_t q0, q1, q2;
            range_ok = quantise(jb.yuv[0], qp, q0) && quantise(jb.yuv[1], qp, q1) && quantise(jb.yuv[2], qp, q2);
            sym[0] = q0;
            sym[1] = q1;
            sym[2] = q2;
        }
    } else {
        uint32_t wl = 1;
        for (uint32_t l = 0; l + 1 < h; ++l) wl = meta.cnt[l] > wl ? meta.cnt[l] : wl;
        float* bufA = static_cast<float*>(ar.take(12ull * wl));
        float* bufB = static_cast<float*>(ar.take(12ull * wl));
        if (!bufA || !bufB) return -PVVC_ERR_CAPACITY;
        for (int l = static_cast<int>(h) - 2; l >= 0; --l) {
            const uint32_t cnt = meta.cnt[l], base = meta.base[l], cbase = meta.base[l + 1];
            const bool last = l == static_cast<int>(h) - 2;
            const float* cin = ((h - 2 - l) & 1) ? bufA : bufB; /* children's DCs (level l+1) */
            float* dout = ((h - 2 - l) & 1) ? bufB : bufA;
            for (uint32_t j = tid; j < cnt; j += 64u) {
                const uint32_t occ = jb.geo[25 + base + j];
                const uint32_t f = fc[base + j];
                uint32_t cw[8];
                float iY[8], iU[8], iV[8];
                uint32_t kk = 0;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    iY[i] = iU[i] = iV[i] = 0.0f;
                    cw[i] = 0;
                    if ((occ >> (7 - i)) & 1u) {
                        const uint32_t c = f + kk;
                        ++kk;
                        if (last) {
                            cw[i] = 1;
                            iY[i] = jb.yuv[3ull * c];
                            iU[i] = jb.yuv[3ull * c + 1];
                            iV[i] = jb.yuv[3ull * c + 2];
                        } else {
                            cw[i] = wt[cbase + c];
                            iY[i] = cin[3 * c];
                            iU[i] = cin[3 * c + 1];
                            iV[i] = cin[3 * c + 2];
                        }
                    }
                }
                float dY, dU, dV, hY[8], hU[8], hV[8];
                const uint32_t coded = node_forward(cw, iY, iU, iV, dY, dU, dV, hY, hU, hV);
                dout[3 * j] = dY;
                dout[3 * j + 1] = dU;
                dout[3 * j + 2] = dV;
                /* the decoder reads node j's ACs at reversed positions cnt + f - j + t */
                uint32_t rr = cnt + f - j;
#pragma unroll
                for (int idx = 1; idx < 8; ++idx) {
                    if (!((coded >> idx) & 1u)) continue;
                    if (rr >= n) { /* cannot happen for a tree that passed tree_structure; */
                        range_ok = false; /* never write outside */
                        continue;
                    }
                    const uint32_t pos = n - 1u - rr;
                    ++rr;
                    int32_t q0, q1, q2;
                    range_ok &= quantise(hY[idx], qp, q0) && quantise(hU[idx], qp, q1) && quantise(hV[idx], qp, q2);
                    sym[pos] = q0;
                    sym[n + pos] = q1;
                    sym[2 * n + pos] = q2;
                }
                if (l == 0) { /* the root's DC is the last coefficient */
                    int32_t q0, q1, q2;
                    range_ok &= quantise(dY, qp, q0) && quantise(dU, qp, q1) && quantise(dV, qp, q2);
                    sym[n - 1] = q0;
                    sym[2 * n - 1] = q1;
                    sym[3 * n - 1] = q2;
                }
            }
            __syncthreads();
        }
    }
    __syncthreads();
    if (__ballot(!range_ok)) return -PVVC_ERR_SYMBOL_RANGE;
    const int64_t bytes = rlgr_encode_wave(sym, 3u * n, jb.out, jb.out_cap, win);
    return bytes < 0 ? -PVVC_ERR_CAPACITY : bytes;
}

/* Packs each job's bytes at its offset of a contiguous buffer (size and at
 * are indexed by the caller's job index, jobs by rank). */
__global__ __launch_bounds__(256) void gather_kernel(const EncJobDev* __restrict__ jobs, const int64_t* __restrict__ size,
                                                     const uint64_t* __restrict__ at, uint32_t count, uint8_t* packed) {
    if (blockIdx.x >= count) return;
    const EncJobDev& jb = jobs[blockIdx.x];
    const uint32_t i = jb.idx;
    const int64_t n = size[i];
    if (n <= 0) return;
    const uint8_t* src = jb.out;
    uint8_t* dst = packed + at[i];
    for (int64_t b = threadIdx.x; b < n; b += blockDim.x) dst[b] = src[b];
}

/* One wavefront per colour stream.  Jobs come longest first; workgroups
 * start in index order.  Wave priority falls with rank as in the persistent
 * decoder: the first `wave1` jobs (one per resident workgroup) at 3, the rest
 * in thirds at 2, 1 and 0. */
__global__ __launch_bounds__(64) void encode_kernel(const EncJobDev* __restrict__ jobs, uint32_t count,
                                                    uint32_t lds_bytes, int64_t* __restrict__ result,
                                                    uint32_t wave1) {
    __shared__ LevelMeta meta;
    __shared__ uint32_t win[kEncWindow];
    const uint32_t q = blockIdx.x;
    if (q >= count) return;
    const uint32_t rest = count > wave1 ? count - wave1 : 0u;
    if (q < wave1) __builtin_amdgcn_s_setprio(3);
    else if (q < wave1 + rest / 3) __builtin_amdgcn_s_setprio(2);
    else if (q < wave1 + 2 * (rest / 3)) __builtin_amdgcn_s_setprio(1);
    const EncJobDev jb = jobs[q];
    Arena ar{g_smem, lds_bytes, 0, jb.ws, jb.ws_cap, 0};
    const int64_t r = encode_job(jb, ar, meta, win);
    if (threadIdx.x == 0) result[jb.idx] = r;
}

}  // namespace pvvc

/* Device scratch of pvvc_decode_batch_noalloc: the queue counters, the stage
 * of deferred predictive items and the per-workgroup workspaces of a
 * persistent decode, for callers that cannot allocate.  Each GPU the library
 * runs on holds its own copy until the process ends.  Host code never passes
 * this name to a kernel (on the host it is only a stand-in for the symbol):
 * it uses the device address hipGetSymbolAddress returns for the current
 * device (scratch_base()). */
constexpr uint64_t kScratchBytes = 256ull << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

/* ======================================================================== */
/* Host                                                                      */
/* ======================================================================== */
using namespace pvvc;

namespace {

/* The last error is per host thread: two threads driving their own states
 * or decoders never see each other's failures. */
thread_local int t_last_error = 0;
/* Process-wide defaults, read by every object that has not set its own
 * (pvvc_state_set_*, pvvc_decoder_set_*). */
std::atomic<void*> g_default_stream{nullptr};
std::atomic<int> g_white_small{0};
std::mutex g_mu;

void set_error(int e) {
    if (e) t_last_error = e;
}

hipStream_t default_stream() { return static_cast<hipStream_t>(g_default_stream.load()); }

bool hip_ok(hipError_t e) {
    if (e != hipSuccess) {
        set_error(PVVC_ERR_HIP);
        std::fprintf(stderr, "pvvc: HIP error %s\n", hipGetErrorString(e));
        return false;
    }
    return true;
}

struct DeviceInfo {
    int cu = 0, lds_max = 0, warp = 64;
    char name[256] = {0};
    bool ok = false;
};
DeviceInfo& devinfo() {
    static DeviceInfo d;
    static std::once_flag once;
    std::call_once(once, [] {
        int dev = 0;
        hipDeviceProp_t p;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess) {
            d.cu = p.multiProcessorCount;
            d.lds_max = static_cast<int>(p.sharedMemPerBlock);
            d.warp = p.warpSize;
            std::snprintf(d.name, sizeof(d.name), "%s", p.gcnArchName);
            d.ok = true;
        }
    });
    return d;
}

/* Device address of `scratch` on each device, looked up once per device with
 * that device current (pvvc_decoder_create, else the first
 * pvvc_decode_batch_noalloc on the device). */
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_at[kMaxDevices] = {};

uint8_t* scratch_base() {
    int dev = 0;
    if (!hip_ok(hipGetDevice(&dev)) || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_at[dev]) {
        void* p = nullptr;
        if (!hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) || !p) return nullptr;
        g_scratch_at[dev] = static_cast<uint8_t*>(p);
    }
    return g_scratch_at[dev];
}

/* Launch-time records of the last decode call: event pairs per main launch. */
/* Kernel timing is opt-in per host thread (pvvc_set_timing): each recorded
 * event is a marker on the launch stream that costs several microseconds
 * between kernels, ~15 us per persistent GoP decode (its four events). */
thread_local bool g_timing_on = false;
struct Timing {
    std::vector<hipEvent_t> ev;
    size_t used = 0;
    void reset() { used = 0; }
    /* a pair of events to record around a launch, or two null events (and
     * nothing counted) while timing is off */
    std::pair<hipEvent_t, hipEvent_t> next() {
        if (!g_timing_on) return {nullptr, nullptr};
        while (ev.size() < used + 2) {
            hipEvent_t e;
            hipEventCreate(&e);
            ev.push_back(e);
        }
        auto r = std::make_pair(ev[used], ev[used + 1]);
        used += 2;
        return r;
    }
};
/* Records a timing event (null while timing is off: nothing recorded). */
inline void trec(hipEvent_t e, hipStream_t s) {
    if (e) hipEventRecord(e, s);
}
/* Per host thread: a decode call and its timing queries run on the same thread. */
thread_local Timing g_timing;   /* whole decode call: every launch of it, in order */
thread_local Timing g_ktiming;  /* the decode kernel launches alone (frame / persistent / drop-in) */

/* Drop-in launch shape (waves per slice, LDS bytes per slice), packed in one
 * word so that readers never see half an update.  Default 4 waves: the second
 * builds the octree while the first decodes RLGR, and all four share the
 * top-down passes and the leaf output; on a 10-bit 816k-point frame of 820
 * slices this measured 0.521 ms against 0.533 ms for 2 waves and 0.618 ms
 * for 1 (16-32 KiB of LDS alike) */
constexpr uint64_t pack_dropin(int waves, int lds) {
    return (static_cast<uint64_t>(static_cast<uint32_t>(waves)) << 32) | static_cast<uint32_t>(lds);
}
std::atomic<uint64_t> g_dropin_tuning{pack_dropin(4, 32 * 1024)};

int clamp_lds(int want) {
    int mx = devinfo().lds_max > 0 ? devinfo().lds_max : 65536;
    int reserve = static_cast<int>(sizeof(LevelMeta) + sizeof(float2) * kSmallTab * kSmallTab) + 64;
    want = std::max(0, std::min(want, mx - reserve));
    return want & ~15;
}

template <typename K>
void allow_lds(K kernel, int bytes) {
    hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

/* ---- state slots (pvvc_state) ---- */
struct StateInfo {
    SlotDev* slots = nullptr;  /* device */
    uint8_t* arena = nullptr;  /* device */
    int n = 0;
    int max_points = 0;
    uint32_t geo_cap = 0;
    uint64_t ws_cap = 0;
    uint64_t slot_stride = 0;
    int* err = nullptr;        /* device word: status of the first failed slice since the last read */
    /* per-state settings; unset ones follow the process-wide defaults */
    bool has_stream = false;
    hipStream_t stream = nullptr;
    uint64_t dropin = 0;       /* pack_dropin(waves, lds), 0 = default */
    int white_small = -1;      /* -1 = default */
};
std::map<const void*, StateInfo> g_states;

hipStream_t state_stream(const StateInfo& si) { return si.has_stream ? si.stream : default_stream(); }

/* Reads and clears a state's device error word on stream s (synchronising s). */
int take_state_error(const StateInfo& si, hipStream_t s) {
    int v = 0;
    if (!hip_ok(hipMemcpyAsync(&v, si.err, sizeof(int), hipMemcpyDeviceToHost, s)) ||
        !hip_ok(hipStreamSynchronize(s)))
        return PVVC_ERR_HIP;
    if (v && !(hip_ok(hipMemsetAsync(si.err, 0, sizeof(int), s)) && hip_ok(hipStreamSynchronize(s))))
        return PVVC_ERR_HIP;
    return v;
}

bool find_state(const void* state, StateInfo& out) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_states.find(state);
    if (it == g_states.end()) return false;
    out = it->second;
    return true;
}

uint64_t align256(uint64_t x) { return (x + 255) & ~uint64_t(255); }

/* geometry capacity: up to ~8 internal nodes per point + header; workspace: coef
 * 12n + first-child / weight arrays over the geometry + two level buffers. */
uint32_t geo_cap_for(int max_points) { return static_cast<uint32_t>(64 + 8ull * max_points + 25); }
uint64_t ws_cap_for(int max_points) {
    uint64_t n = static_cast<uint64_t>(max_points);
    return align256(12 * n + 2 * (4 * (geo_cap_for(max_points) + 1ull) + 16) + 48 * n + 512);
}

bool state_create(int n_patches, int max_points, StateInfo& si) {
    si.n = n_patches;
    si.max_points = max_points;
    si.geo_cap = geo_cap_for(max_points);
    si.ws_cap = ws_cap_for(max_points);
    si.slot_stride = align256(si.geo_cap) + align256(12ull * max_points) + si.ws_cap;
    if (!hip_ok(hipMalloc(&si.slots, sizeof(SlotDev) * std::max(1, n_patches)))) return false;
    /* on failure the caller's state_destroy frees what was allocated */
    if (!hip_ok(hipMalloc(&si.arena, si.slot_stride * std::max(1, n_patches)))) return false;
    if (!hip_ok(hipMalloc(&si.err, sizeof(int))) || !hip_ok(hipMemset(si.err, 0, sizeof(int)))) return false;
    std::vector<SlotDev> h(std::max(1, n_patches));
    for (int i = 0; i < n_patches; ++i) {
        SlotDev& s = h[i];
        std::memset(&s, 0, sizeof(s));
        uint8_t* b = si.arena + si.slot_stride * i;
        s.geo = b;
        s.ref = reinterpret_cast<float*>(b + align256(si.geo_cap));
        s.ws = b + align256(si.geo_cap) + align256(12ull * max_points);
        s.ws_cap = si.ws_cap;
        s.geo_cap = si.geo_cap;
        s.n_cap = static_cast<uint32_t>(max_points);
    }
    if (!hip_ok(hipMemcpy(si.slots, h.data(), sizeof(SlotDev) * n_patches, hipMemcpyHostToDevice))) return false;
    return true;
}

void state_destroy(StateInfo& si) {
    if (si.slots) hipFree(si.slots);
    if (si.arena) hipFree(si.arena);
    if (si.err) hipFree(si.err);
    si = StateInfo{};
}

bool state_reset(StateInfo& si) {
    std::vector<SlotDev> h(si.n);
    if (!hip_ok(hipMemcpy(h.data(), si.slots, sizeof(SlotDev) * si.n, hipMemcpyDeviceToHost))) return false;
    for (auto& s : h) {
        s.n = 0;
        s.geo_size = 0;
        s.status = 0;
        s.ready = 0;
    }
    return hip_ok(hipMemcpy(si.slots, h.data(), sizeof(SlotDev) * si.n, hipMemcpyHostToDevice));
}

}  // namespace

namespace pvvc {

/* Plans the batch of frames [0, n_frames) in decode order.
 *  - a predictive item whose patch's intra item is earlier in the same
 *    segment expands that item's geometry and is deferred (its reference is
 *    added by the finish pass); otherwise it reads the state slot;
 *  - a patch's new intra slice after earlier items of that patch starts a new
 *    segment, so no launch rewrites state an earlier item still reads;
 *  - queue order per segment: point count descending, an intra item keyed by
 *    the largest size of its patch, intra first on ties.
 * intra_geo (per slot, may be null) tracks the latest intra geometry size for
 * the state-read accounting. */
int plan_batch(const pvvc_frame* const* frames, int n_frames, const int* slot_base, const int* slot_count,
               const uint32_t* geo_cap, uint32_t* intra_geo, BatchPlan& p) {
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    uint32_t n_items = 0;
    size_t data_bytes = 0;
    for (int f = 0; f < n_frames; ++f) {
        if (!frames[f]) return PVVC_ERR_INVALID_ARG;
        n_items += frames[f]->slice_cnt;
        data_bytes += al(frames[f]->data.size());
    }
    p = BatchPlan{};
    p.n_items = n_items;
    p.o_items = 0;
    p.o_order = al(sizeof(ItemDev) * n_items);
    p.o_status = al(p.o_order + 4ull * n_items);
    p.o_defer = al(p.o_status + 4ull * n_items);
    p.o_data = al(p.o_defer + 4ull * n_items);
    p.total = p.o_data + data_bytes;
    p.items.assign(n_items, ItemDev{});
    p.geo_at.assign(n_items, UINT64_MAX);
    p.col_at.assign(n_items, UINT64_MAX);
    std::vector<uint32_t> gcap_of(n_items, 0);
    /* per-slot tables, valid where their stamp equals the current segment's */
    int max_slot = 0;
    for (int f = 0; f < n_frames; ++f) max_slot = std::max(max_slot, slot_base[f] + slot_count[f]);
    thread_local std::vector<uint32_t> intra_at, intra_stamp, seen_stamp, pmax, pmax_stamp;
    thread_local uint32_t epoch = 0;
    if (intra_at.size() < static_cast<size_t>(max_slot)) {
        intra_at.assign(max_slot, 0);
        intra_stamp.assign(max_slot, 0);
        seen_stamp.assign(max_slot, 0);
        pmax.assign(max_slot, 0);
        pmax_stamp.assign(max_slot, 0);
        epoch = 0;
    }
    uint32_t seg_epoch = ++epoch;   /* slot -> intra item / "has items" in this segment */
    std::vector<uint32_t> seg_starts{0};
    size_t data_at = p.o_data;
    int64_t pts_acc = 0;
    uint32_t it = 0;
    for (int f = 0; f < n_frames; ++f) {
        const pvvc_frame* fr = frames[f];
        p.frame_first.push_back(it);
        p.frame_count.push_back(fr->slice_cnt);
        p.frame_off.push_back(pts_acc);
        p.frame_pts.push_back(static_cast<int64_t>(fr->points));
        p.frame_data_at.push_back(data_at);
        bool new_seg = false;
        for (uint32_t s = 0; s < fr->slice_cnt; ++s) {
            const int pi = fr->index[s];
            if (pi < 0 || pi >= slot_count[f]) return PVVC_ERR_INVALID_ARG;
            if (!(fr->type[s] & PVVC_TYPE_PREDICT) && seen_stamp[slot_base[f] + pi] == seg_epoch) new_seg = true;
        }
        if (new_seg && it != 0) {
            seg_starts.push_back(it);
            seg_epoch = ++epoch;
        }
        int64_t inner = 0;
        for (uint32_t s = 0; s < fr->slice_cnt; ++s, ++it) {
            ItemDev& x = p.items[it];
            std::memset(&x, 0, sizeof(x));
            const int pi = fr->index[s];
            if (pi < 0 || pi >= slot_count[f]) return PVVC_ERR_INVALID_ARG;
            const int slot = slot_base[f] + pi;
            if (fr->geo_size[s]) p.geo_at[it] = data_at + fr->geo_off[s];
            if (fr->col_size[s]) p.col_at[it] = data_at + fr->col_off[s];
            x.geo_size = fr->geo_size[s];
            x.col_size = fr->col_size[s];
            x.n = fr->size[s];
            x.patch = slot;
            x.type = fr->type[s];
            x.qp = fr->qp[s];
            x.out_off = static_cast<uint64_t>(pts_acc + inner);
            x.stage_off = static_cast<uint32_t>(pts_acc + inner);
            for (int k = 0; k < 12; ++k) x.mv[k] = fr->mv[16 * s + k];
            gcap_of[it] = geo_cap[f];
            if (x.type & PVVC_TYPE_PREDICT) {
                if (intra_geo) p.state_bytes += static_cast<int64_t>(intra_geo[slot]) + 12ll * x.n;
                if (intra_stamp[slot] == seg_epoch) {
                    const uint32_t ii = intra_at[slot];
                    p.geo_at[it] = p.geo_at[ii];
                    x.geo_size = p.items[ii].geo_size;
                    x.wait = 1;
                } else {
                    p.geo_at[it] = UINT64_MAX;
                    x.geo_size = 0;
                }
            } else {
                if (intra_geo) intra_geo[slot] = x.geo_size;
                intra_at[slot] = it;
                intra_stamp[slot] = seg_epoch;
            }
            seen_stamp[slot] = seg_epoch;
            inner += fr->size[s];
        }
        pts_acc += inner;
        data_at += al(fr->data.size());
        p.input_bytes += 82ll * fr->slice_cnt;
        for (uint32_t s = 0; s < fr->slice_cnt; ++s) p.input_bytes += fr->geo_size[s] + fr->col_size[s];
    }
    p.total_points = pts_acc;
    if (pts_acc >= (int64_t{1} << 32)) return PVVC_ERR_CAPACITY; /* stage offsets are 32-bit */
    p.order.assign(n_items, 0);
    seg_starts.push_back(n_items);
    for (size_t si = 0; si + 1 < seg_starts.size(); ++si) {
        const uint32_t lo = seg_starts[si], hi = seg_starts[si + 1];
        const uint32_t pm_epoch = ++epoch;
        for (uint32_t i = lo; i < hi; ++i) {
            const int sl = p.items[i].patch;
            if (pmax_stamp[sl] != pm_epoch) {
                pmax_stamp[sl] = pm_epoch;
                pmax[sl] = 0;
            }
            pmax[sl] = std::max(pmax[sl], p.items[i].n);
        }
        /* key: larger first; on ties a patch's slices together (they read
         * the same intra geometry), its intra slice first, then item order */
        struct Key {
            uint64_t hi;  /* size (descending) | patch */
            uint32_t lo;  /* predictive flag | item */
            bool operator<(const Key& o) const { return hi != o.hi ? hi < o.hi : lo < o.lo; }
        };
        std::vector<Key> keys(hi - lo);
        for (uint32_t i = lo; i < hi; ++i) {
            const ItemDev& A = p.items[i];
            const bool pa = A.type & PVVC_TYPE_PREDICT;
            const uint32_t ka = pa ? A.n : pmax[A.patch];
            keys[i - lo] = Key{(static_cast<uint64_t>(0xffffffffu - ka) << 32) | static_cast<uint32_t>(A.patch),
                               (static_cast<uint32_t>(pa) << 31) | i};
        }
        std::sort(keys.begin(), keys.end());
        std::vector<uint32_t> ids(hi - lo);
        for (uint32_t j = 0; j < hi - lo; ++j) ids[j] = keys[j].lo & 0x7fffffffu;
        uint64_t need = 0;
        for (uint32_t i : ids) {
            const ItemDev& x = p.items[i];
            need = std::max(need, item_ws_bound(x.n, x.geo_size ? x.geo_size : gcap_of[i]));
        }
        std::copy(ids.begin(), ids.end(), p.order.begin() + lo);
        const uint32_t dlo = static_cast<uint32_t>(p.defer.size());
        for (uint32_t i = lo; i < hi; ++i)
            if (p.items[i].wait) p.defer.push_back(i);
        p.segs.push_back(Segment{lo, hi - lo, dlo, static_cast<uint32_t>(p.defer.size()) - dlo,
                                 (need + 255) & ~uint64_t(255)});
    }
    return PVVC_OK;
}

/* Writes the planned image for a device copy at dev_base into host (p.total bytes). */
void fill_batch_image(const BatchPlan& p, const pvvc_frame* const* frames, uint8_t* dev_base, uint8_t* host) {
    std::vector<ItemDev> items = p.items;
    for (uint32_t i = 0; i < p.n_items; ++i) {
        items[i].geo = p.geo_at[i] == UINT64_MAX ? nullptr : dev_base + p.geo_at[i];
        items[i].col = p.col_at[i] == UINT64_MAX ? nullptr : dev_base + p.col_at[i];
    }
    std::memcpy(host + p.o_items, items.data(), sizeof(ItemDev) * p.n_items);
    std::memcpy(host + p.o_order, p.order.data(), 4ull * p.n_items);
    std::memset(host + p.o_status, 0, 4ull * p.n_items);
    if (!p.defer.empty()) std::memcpy(host + p.o_defer, p.defer.data(), 4ull * p.defer.size());
    for (size_t f = 0; f < p.frame_data_at.size(); ++f)
        std::memcpy(host + p.frame_data_at[f], frames[f]->data.data(), frames[f]->data.size());
}

}  // namespace pvvc

/* ---- native decoder / batch ---- */
struct pvvc_decoder {
    StateInfo st;
    pvvc_tuning tuning{1, 8 * 1024, 16, 1, 0, 0};
    bool auto_small = true;      /* small segments pick their own shape until a tuning is set */
    uint32_t launch_id = 0;
    uint8_t* slab = nullptr;
    uint64_t slab_bytes = 0, slab_total = 0;
    uint32_t* queue = nullptr;
    float* stage = nullptr;      /* unreferenced colours of deferred predictive items */
    uint64_t stage_bytes = 0;
    std::vector<uint32_t> intra_geo; /* geometry bytes of each patch's latest intra slice
                                        seen by pvvc_batch_create (state-read accounting) */
    bool has_stream = false;     /* pvvc_decoder_set_stream; else the process default */
    hipStream_t stream = nullptr;
    int white_small = -1;        /* pvvc_decoder_set_white_small_patches; -1 = process default */
    /* lane decoding of short slices (tuning.lane_max_points): a second stream
     * with its own queue and workspaces, and the lanes' coefficients */
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    uint32_t* queue2 = nullptr;
    uint8_t* slab2 = nullptr;
    uint64_t slab2_total = 0;
    float* lane_coef = nullptr;
    uint64_t lane_coef_bytes = 0;
    int* lane_status = nullptr;
    uint64_t lane_status_bytes = 0;
};

namespace {
hipStream_t dec_stream(const pvvc_decoder* d) { return d->has_stream ? d->stream : default_stream(); }
int dec_white_small(const pvvc_decoder* d) { return d->white_small >= 0 ? d->white_small : g_white_small.load(); }
}  // namespace

struct pvvc_batch {
    BatchPlan plan;
    pvvc_decoder* dec = nullptr;
    int n_frames = 0;
    std::vector<int64_t> frame_off, frame_pts;
    std::vector<uint32_t> frame_first, frame_count; /* items in slice order */
    std::vector<Segment> segs;
    uint32_t n_items = 0;
    int64_t total_points = 0;
    int64_t input_bytes = 0;
    int64_t state_bytes = 0;     /* resident GoP state read by predictive items */
    uint8_t* dev = nullptr;
    ItemDev* items = nullptr;
    uint32_t* order = nullptr;
    uint32_t* defer = nullptr;
    int* status = nullptr;
    hipStream_t last_stream = nullptr; /* stream of the last decode of this batch */
    /* captured per-frame launches, valid for exactly this key: output mode and
     * buffers, waves per slice, LDS bytes and the white-small setting */
    hipGraphExec_t graph = nullptr;
    int graph_output = -1;
    void* graph_ptrs[3] = {nullptr, nullptr, nullptr};
    int graph_waves = 0;
    uint32_t graph_lds = 0;
    int graph_white = -1;
};

namespace {

template <int W>
void launch_frame(const pvvc_batch* b, int f, const Sink& sink, uint32_t lds, hipStream_t s) {
    uint32_t cnt = b->frame_count[f];
    if (!cnt) return;
    hipLaunchKernelGGL(frame_kernel<W>, dim3(cnt), dim3(64 * W), lds, s, b->items, b->frame_first[f], cnt,
                       b->dec->st.slots, sink, lds, b->status);
}

template <int W>
void launch_persistent_w(const BatchDev& b, const Segment& sg, const Sink& sink, uint32_t lds, uint32_t grid,
                         uint32_t chunk, uint32_t prio, const LaunchRes& r, uint32_t launch_id, hipStream_t s) {
    const bool timed = r.timed;
    std::pair<hipEvent_t, hipEvent_t> ev{};
    if (timed) {
        ev = g_ktiming.next();
        trec(ev.first, s);
    }
    hipLaunchKernelGGL(persistent_kernel<W>, dim3(grid), dim3(64 * W), lds, s, b.items, b.order + sg.begin, sg.count,
                       r.queue, r.slots, sink, lds, r.slab, r.slab_bytes, launch_id, b.status, r.stage, chunk, prio,
                       r.pre_coef, r.pre_status);
    if (timed) trec(ev.second, s);
    if (sg.dcount)
        hipLaunchKernelGGL(finish_kernel, dim3(sg.dcount), dim3(256), 0, s, b.items, b.defer + sg.dbegin, sg.dcount,
                           r.slots, sink, r.stage, b.status);
}

template <int W>
void set_lds_attrs(int bytes) {
    allow_lds(frame_kernel<W>, bytes);
    allow_lds(persistent_kernel<W>, bytes);
    allow_lds(dropin_kernel<W>, bytes);
}

void set_all_lds_attrs(int bytes) {
    set_lds_attrs<1>(bytes);
    set_lds_attrs<2>(bytes);
    set_lds_attrs<4>(bytes);
}

}  // namespace

namespace pvvc {

bool rt_hip_ok(hipError_t e) { return hip_ok(e); }
void rt_timing_reset() {
    g_timing.reset();
    g_ktiming.reset();
}
void rt_set_error(int e) { set_error(e); }
hipStream_t rt_stream() { return default_stream(); }

void rt_init_kernels() {
    static std::once_flag once;
    std::call_once(once, [] {
        set_all_lds_attrs(clamp_lds(1 << 30));
        /* the Haar factor table, complete before any decode kernel runs */
        hipLaunchKernelGGL(raht_table_kernel, dim3((kRahtTab * kRahtTab + 255) / 256), dim3(256), 0, nullptr);
        hip_ok(hipGetLastError());
        hip_ok(hipDeviceSynchronize());
    });
}

bool rt_warm_kernels(const pvvc_tuning& t, uint32_t* queue, hipStream_t s) {
    rt_init_kernels();
    const uint32_t lds = static_cast<uint32_t>(clamp_lds(t.lds_bytes));
    const Sink sink{nullptr, nullptr, nullptr, PVVC_OUT_NATIVE, 0};
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, s, queue);
    if (t.waves_per_patch == 2)
        hipLaunchKernelGGL(persistent_kernel<2>, dim3(1), dim3(128), lds, s, nullptr, nullptr, 0u, queue, nullptr, sink,
                           lds, nullptr, 0ull, 0u, nullptr, nullptr, 1u, 0u, nullptr, nullptr);
    else if (t.waves_per_patch == 4)
        hipLaunchKernelGGL(persistent_kernel<4>, dim3(1), dim3(256), lds, s, nullptr, nullptr, 0u, queue, nullptr, sink,
                           lds, nullptr, 0ull, 0u, nullptr, nullptr, 1u, 0u, nullptr, nullptr);
    else
        hipLaunchKernelGGL(persistent_kernel<1>, dim3(1), dim3(64), lds, s, nullptr, nullptr, 0u, queue, nullptr, sink,
                           lds, nullptr, 0ull, 0u, nullptr, nullptr, 1u, 0u, nullptr, nullptr);
    return hip_ok(hipGetLastError()) && hip_ok(hipStreamSynchronize(s));
}

uint32_t persistent_grid(const pvvc_tuning& t, uint32_t count) {
    const uint32_t cap = static_cast<uint32_t>(std::max(1, devinfo().cu) * std::max(1, t.wgs_per_cu));
    return std::max(1u, std::min(count, cap));
}

/* A segment of at most 4 slices per CU cannot fill the CUs with one
 * wavefront per slice.  Unless the caller chose a tuning, it runs 4 wavefronts
 * per slice (32 KiB LDS, 4 workgroups per CU), which share the top-down
 * passes: a c0 GoP (~700 slices) in 0.271 ms against 0.311 ms. */
pvvc_tuning segment_tuning(const pvvc_tuning& t, bool auto_small, uint32_t count) {
    pvvc_tuning e = t;
    if (auto_small && count <= static_cast<uint32_t>(std::max(1, devinfo().cu)) * 4u) {
        e.waves_per_patch = 4;
        e.lds_bytes = 32 * 1024;
        e.wgs_per_cu = 4;
    }
    return e;
}

int launch_segment(const BatchDev& b, const Segment& sg, const Sink& sink, const pvvc_tuning& t, const LaunchRes& r,
                   uint32_t launch_id, hipStream_t s) {
    if (!sg.count) return PVVC_OK;
    const uint32_t lds = static_cast<uint32_t>(clamp_lds(t.lds_bytes));
    uint32_t grid = persistent_grid(t, sg.count);
    if (r.grid_cap) grid = std::min(grid, r.grid_cap);
    /* workgroup g uses slab bytes [g * stride, g * stride + stride) */
    if (sg.ws_need > r.slab_bytes || static_cast<uint64_t>(grid) * r.slab_bytes > r.slab_total) {
        set_error(PVVC_ERR_CAPACITY);
        return PVVC_ERR_CAPACITY;
    }
    const uint32_t chunk = static_cast<uint32_t>(std::max(1, t.queue_chunk));
    /* default: the first wave of the queue, one item per workgroup, at the
     * top priority.  Those items start together; the scalar-issue arbiter
     * otherwise serves the oldest wavefronts first, so a workgroup that pops
     * a second item keeps starving a younger one still on its first (larger)
     * item, and that item ends the launch (c1 GoP: 0.950 -> 0.908 ms; with
     * the graded levels of the rest of the queue 0.899 ms) */
    const uint32_t prio = t.priority_items < 0 ? 0u
                          : t.priority_items == 0 ? grid
                                                  : static_cast<uint32_t>(t.priority_items);
    if (t.waves_per_patch == 2) launch_persistent_w<2>(b, sg, sink, lds, grid, chunk, prio, r, launch_id, s);
    else if (t.waves_per_patch == 4) launch_persistent_w<4>(b, sg, sink, lds, grid, chunk, prio, r, launch_id, s);
    else launch_persistent_w<1>(b, sg, sink, lds, grid, chunk, prio, r, launch_id, s);
    return PVVC_OK;
}

}  // namespace pvvc

extern "C" {

int pvvc_last_error(void) {
    const int e = t_last_error;
    t_last_error = 0;
    return e;
}

#ifdef PVVC_PHASE_PROFILE
/* Profiling build only: installs the stamp buffer (device, 8 u64 per item). */
int pvvc_phase_profile_set(void* dev_buf) {
    return hip_ok(hipMemcpyToSymbol(HIP_SYMBOL(g_prof), &dev_buf, sizeof(void*))) ? PVVC_OK : PVVC_ERR_HIP;
}
#endif

int pvvc_set_stream(void* stream) {
    g_default_stream.store(stream);
    return PVVC_OK;
}

int pvvc_set_white_small_patches(int enable) {
    g_white_small.store(enable ? 1 : 0);
    return PVVC_OK;
}

int pvvc_device_info(int* cu_count, int* lds_per_block, int* warp_size, char* name, int name_len) {
    DeviceInfo& d = devinfo();
    if (!d.ok) return PVVC_ERR_HIP;
    if (cu_count) *cu_count = d.cu;
    if (lds_per_block) *lds_per_block = d.lds_max;
    if (warp_size) *warp_size = d.warp;
    if (name && name_len > 0) std::snprintf(name, name_len, "%s", d.name);
    return PVVC_OK;
}

void* pvvc_dev_alloc(size_t bytes) {
    void* p = nullptr;
    if (!hip_ok(hipMalloc(&p, bytes ? bytes : 1))) return nullptr;
    return p;
}
void pvvc_dev_free(void* p) {
    if (p) hipFree(p);
}
int pvvc_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    return hip_ok(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)) ? PVVC_OK : PVVC_ERR_HIP;
}
int pvvc_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    return hip_ok(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)) ? PVVC_OK : PVVC_ERR_HIP;
}
int pvvc_memcpy_d2d(void* dst, const void* src, size_t bytes) {
    return hip_ok(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, default_stream())) ? PVVC_OK
                                                                                              : PVVC_ERR_HIP;
}
int pvvc_memset(void* dst, int value, size_t bytes) {
    return hip_ok(hipMemset(dst, value, bytes)) ? PVVC_OK : PVVC_ERR_HIP;
}
/* Synchronises the default stream, then folds the device error words of the
 * states that run on it into the calling thread's last error. */
int pvvc_synchronize(void) {
    const hipStream_t s = default_stream();
    if (!hip_ok(hipStreamSynchronize(s))) return PVVC_ERR_HIP;
    std::vector<StateInfo> on_default;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        for (auto& kv : g_states)
            if (!kv.second.has_stream) on_default.push_back(kv.second);
    }
    int first = PVVC_OK;
    for (const StateInfo& si : on_default) {
        const int e = take_state_error(si, s);
        if (e && first == PVVC_OK) first = e;
    }
    set_error(first);
    return PVVC_OK;
}

int pvvc_stream_create(void** stream) {
    if (!stream) return PVVC_ERR_INVALID_ARG;
    hipStream_t s = nullptr;
    if (!hip_ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking))) return PVVC_ERR_HIP;
    *stream = s;
    return PVVC_OK;
}
int pvvc_stream_destroy(void* stream) {
    return hip_ok(hipStreamDestroy(static_cast<hipStream_t>(stream))) ? PVVC_OK : PVVC_ERR_HIP;
}
int pvvc_stream_synchronize(void* stream) {
    return hip_ok(hipStreamSynchronize(static_cast<hipStream_t>(stream))) ? PVVC_OK : PVVC_ERR_HIP;
}

/* Milliseconds of the main kernel launches of the last decode call. */
int pvvc_set_timing(int on) {
    const int prev = g_timing_on ? 1 : 0;
    g_timing_on = on != 0;
    return prev;
}
double pvvc_last_kernel_ms(void) {
    double tot = 0;
    for (size_t i = 0; i + 1 < g_timing.used; i += 2) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, g_timing.ev[i], g_timing.ev[i + 1]) == hipSuccess) tot += ms;
    }
    return tot;
}
int pvvc_last_decode_kernel_times(double* ms, int max) {
    int k = 0;
    for (size_t i = 0; i + 1 < g_ktiming.used && k < max; i += 2, ++k) {
        float t = 0;
        hipEventElapsedTime(&t, g_ktiming.ev[i], g_ktiming.ev[i + 1]);
        if (ms) ms[k] = t;
    }
    return static_cast<int>(g_ktiming.used / 2);
}
int pvvc_last_launch_times(double* ms, int max) {
    int k = 0;
    for (size_t i = 0; i + 1 < g_timing.used && k < max; i += 2, ++k) {
        float t = 0;
        hipEventElapsedTime(&t, g_timing.ev[i], g_timing.ev[i + 1]);
        if (ms) ms[k] = t;
    }
    return static_cast<int>(g_timing.used / 2);
}

/* ---- pvvc_state (drop-in) ---- */
pvvc_state* pvvc_state_alloc(int n_patches, int max_points) {
    if (n_patches <= 0 || max_points <= 0) {
        set_error(PVVC_ERR_INVALID_ARG);
        return nullptr;
    }
    StateInfo si;
    if (!state_create(n_patches, max_points, si)) {
        state_destroy(si);
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_states[si.slots] = si;
    return reinterpret_cast<pvvc_state*>(si.slots);
}

void pvvc_state_free(pvvc_state* state) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_states.find(state);
    if (it == g_states.end()) return;
    state_destroy(it->second);
    g_states.erase(it);
}

int pvvc_state_patch_status(const pvvc_state* state, int patch) {
    StateInfo si;
    if (!find_state(state, si)) return PVVC_ERR_INVALID_ARG;
    if (patch < 0 || patch >= si.n) return PVVC_ERR_INVALID_ARG;
    const hipStream_t s = state_stream(si);
    set_error(take_state_error(si, s)); /* synchronises the state's stream */
    SlotDev sd;
    if (!hip_ok(hipMemcpy(&sd, si.slots + patch, sizeof(sd), hipMemcpyDeviceToHost))) return PVVC_ERR_HIP;
    return sd.status;
}

int pvvc_state_set_stream(pvvc_state* state, void* stream) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_states.find(state);
    if (it == g_states.end()) return PVVC_ERR_INVALID_ARG;
    it->second.has_stream = true;
    it->second.stream = static_cast<hipStream_t>(stream);
    return PVVC_OK;
}

int pvvc_state_set_dropin_tuning(pvvc_state* state, int waves, int lds_bytes) {
    if (waves != 0 && waves != 1 && waves != 2 && waves != 4) return PVVC_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_states.find(state);
    if (it == g_states.end()) return PVVC_ERR_INVALID_ARG;
    it->second.dropin = waves ? pack_dropin(waves, lds_bytes) : 0;
    return PVVC_OK;
}

int pvvc_state_set_white_small_patches(pvvc_state* state, int enable) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_states.find(state);
    if (it == g_states.end()) return PVVC_ERR_INVALID_ARG;
    it->second.white_small = enable < 0 ? -1 : (enable ? 1 : 0);
    return PVVC_OK;
}

int pvvc_state_synchronize(pvvc_state* state) {
    StateInfo si;
    if (!find_state(state, si)) return PVVC_ERR_INVALID_ARG;
    const int e = take_state_error(si, state_stream(si));
    set_error(e);
    return e;
}

int pvvc_set_dropin_tuning(int waves, int lds_bytes) {
    if (waves != 1 && waves != 2 && waves != 4) return PVVC_ERR_INVALID_ARG;
    g_dropin_tuning.store(pack_dropin(waves, lds_bytes));
    return PVVC_OK;
}

void launch_cudaProcess(int grid, int block, Points* cudaData, int timestamp, int* inner_offset, int* index,
                        uint8_t* type, float** mv, uint32_t* size, uint8_t* qp, uint8_t** geometry,
                        uint32_t* geometry_size, uint8_t** color, uint32_t* color_size, pvvc_state* state,
                        int patch_size) {
    (void)grid;
    (void)block;
    (void)timestamp;
    if (patch_size <= 0) return;
    StateInfo si;
    if (!find_state(state, si)) {
        set_error(PVVC_ERR_INVALID_ARG);
        return;
    }
    const uint64_t shape = si.dropin ? si.dropin : g_dropin_tuning.load();
    const int waves = static_cast<int>(shape >> 32);
    const int lds = clamp_lds(static_cast<int>(static_cast<uint32_t>(shape)));
    const int white = si.white_small >= 0 ? si.white_small : g_white_small.load();
    const hipStream_t s = state_stream(si);
    rt_init_kernels();
    g_timing.reset();
    g_ktiming.reset();
    auto ev = g_timing.next();
    auto kev = g_ktiming.next();
    trec(ev.first, s);
    trec(kev.first, s);
    switch (waves) {
        case 2:
            hipLaunchKernelGGL(dropin_kernel<2>, dim3(patch_size), dim3(128), lds, s, cudaData, inner_offset, index,
                               type, mv, size, qp, geometry, geometry_size, color, color_size, si.slots, si.n,
                               patch_size, static_cast<uint32_t>(lds), white, si.err);
            break;
        case 4:
            hipLaunchKernelGGL(dropin_kernel<4>, dim3(patch_size), dim3(256), lds, s, cudaData, inner_offset, index,
                               type, mv, size, qp, geometry, geometry_size, color, color_size, si.slots, si.n,
                               patch_size, static_cast<uint32_t>(lds), white, si.err);
            break;
        default:
            hipLaunchKernelGGL(dropin_kernel<1>, dim3(patch_size), dim3(64), lds, s, cudaData, inner_offset, index,
                               type, mv, size, qp, geometry, geometry_size, color, color_size, si.slots, si.n,
                               patch_size, static_cast<uint32_t>(lds), white, si.err);
    }
    hip_ok(hipGetLastError());
    trec(kev.second, s);
    trec(ev.second, s);
}

/* ---- drop-in staging (test_decode_only.cc:97-123 in one copy) ---- */
}  // extern "C"

namespace {

/* Packed image of one drop-in call: [ptr arrays][u32/int arrays][u8 arrays][mv][blobs]. */
struct StageLayout {
    size_t o_mvp, o_geop, o_colp, o_off, o_idx, o_size, o_gsz, o_csz, o_type, o_qp, o_mv, o_data, total;
    explicit StageLayout(const pvvc_frame* f) {
        const size_t S = f->slice_cnt;
        auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
        o_mvp = 0;
        o_geop = al(o_mvp + 8 * S);
        o_colp = al(o_geop + 8 * S);
        o_off = al(o_colp + 8 * S);
        o_idx = al(o_off + 4 * S);
        o_size = al(o_idx + 4 * S);
        o_gsz = al(o_size + 4 * S);
        o_csz = al(o_gsz + 4 * S);
        o_type = al(o_csz + 4 * S);
        o_qp = al(o_type + S);
        o_mv = al(o_qp + S);
        o_data = al(o_mv + 64 * S);
        total = o_data + f->data.size();
    }
};

/* Fills the host image for a device copy at `dev` and the call's device arrays. */
void stage_fill(const pvvc_frame* f, const StageLayout& L, uint8_t* dev, uint8_t* h, pvvc_dropin_frame* d) {
    const uint32_t S = f->slice_cnt;
    int64_t acc = 0;
    int* off = reinterpret_cast<int*>(h + L.o_off);
    for (uint32_t i = 0; i < S; ++i) {
        off[i] = static_cast<int>(acc);
        acc += f->size[i];
        const uint64_t mvp = reinterpret_cast<uint64_t>(dev + L.o_mv + 64 * i);
        const uint64_t gp = f->geo_size[i] ? reinterpret_cast<uint64_t>(dev + L.o_data + f->geo_off[i]) : 0;
        const uint64_t cp = f->col_size[i] ? reinterpret_cast<uint64_t>(dev + L.o_data + f->col_off[i]) : 0;
        std::memcpy(h + L.o_mvp + 8 * i, &mvp, 8);
        std::memcpy(h + L.o_geop + 8 * i, &gp, 8);
        std::memcpy(h + L.o_colp + 8 * i, &cp, 8);
    }
    std::memcpy(h + L.o_idx, f->index.data(), 4ull * S);
    std::memcpy(h + L.o_size, f->size.data(), 4ull * S);
    std::memcpy(h + L.o_gsz, f->geo_size.data(), 4ull * S);
    std::memcpy(h + L.o_csz, f->col_size.data(), 4ull * S);
    std::memcpy(h + L.o_type, f->type.data(), S);
    std::memcpy(h + L.o_qp, f->qp.data(), S);
    std::memcpy(h + L.o_mv, f->mv.data(), 64ull * S);
    std::memcpy(h + L.o_data, f->data.data(), f->data.size());
    d->mv = reinterpret_cast<float**>(dev + L.o_mvp);
    d->geometry = reinterpret_cast<uint8_t**>(dev + L.o_geop);
    d->color = reinterpret_cast<uint8_t**>(dev + L.o_colp);
    d->inner_offset = reinterpret_cast<int*>(dev + L.o_off);
    d->index = reinterpret_cast<int*>(dev + L.o_idx);
    d->size = reinterpret_cast<uint32_t*>(dev + L.o_size);
    d->geometry_size = reinterpret_cast<uint32_t*>(dev + L.o_gsz);
    d->color_size = reinterpret_cast<uint32_t*>(dev + L.o_csz);
    d->type = dev + L.o_type;
    d->qp = dev + L.o_qp;
    d->slice_cnt = static_cast<int>(S);
    d->point_cnt = acc;
}

}  // namespace

extern "C" {

int pvvc_dropin_stage(const pvvc_frame* f, pvvc_dropin_frame* d) {
    if (!f || !d) return PVVC_ERR_INVALID_ARG;
    std::memset(d, 0, sizeof(*d));
    const StageLayout L(f);
    uint8_t* dev = static_cast<uint8_t*>(pvvc_dev_alloc(L.total));
    if (!dev) return PVVC_ERR_HIP;
    std::vector<uint8_t> h(L.total, 0);
    stage_fill(f, L, dev, h.data(), d);
    if (pvvc_memcpy_h2d(dev, h.data(), L.total) != PVVC_OK) {
        pvvc_dev_free(dev);
        std::memset(d, 0, sizeof(*d));
        return PVVC_ERR_HIP;
    }
    d->blob = dev;
    d->blob_bytes = L.total;
    return PVVC_OK;
}

size_t pvvc_dropin_stage_bytes(const pvvc_frame* f) { return f ? StageLayout(f).total : 0; }

int pvvc_dropin_stage_to(const pvvc_frame* f, void* dev, size_t dev_bytes, void* host, size_t host_bytes,
                         pvvc_dropin_frame* d) {
    if (!f || !d || !dev || !host) return PVVC_ERR_INVALID_ARG;
    std::memset(d, 0, sizeof(*d));
    const StageLayout L(f);
    if (L.total > dev_bytes || L.total > host_bytes) return PVVC_ERR_CAPACITY;
    stage_fill(f, L, static_cast<uint8_t*>(dev), static_cast<uint8_t*>(host), d);
    if (!hip_ok(hipMemcpyAsync(dev, host, L.total, hipMemcpyHostToDevice, default_stream()))) {
        std::memset(d, 0, sizeof(*d));
        return PVVC_ERR_HIP;
    }
    d->blob = nullptr; /* caller-owned */
    d->blob_bytes = L.total;
    return PVVC_OK;
}

void* pvvc_host_alloc(size_t bytes) {
    void* p = nullptr;
    return hip_ok(hipHostMalloc(&p, std::max<size_t>(bytes, 1), hipHostMallocDefault)) ? p : nullptr;
}
void pvvc_host_free(void* p) {
    if (p) hipHostFree(p);
}

void pvvc_dropin_release(pvvc_dropin_frame* d) {
    if (d && d->blob) pvvc_dev_free(d->blob);
    if (d) std::memset(d, 0, sizeof(*d));
}

/* ---- native decoder ---- */
int pvvc_decoder_create(pvvc_decoder** out, int max_patches, int max_points_per_patch) {
    if (!out || max_patches <= 0 || max_points_per_patch <= 0) return PVVC_ERR_INVALID_ARG;
    auto* d = new pvvc_decoder();
    if (!state_create(max_patches, max_points_per_patch, d->st)) {
        state_destroy(d->st);
        delete d;
        return PVVC_ERR_HIP;
    }
    if (!hip_ok(hipMalloc(&d->queue, 64)) || !hip_ok(hipMemset(d->queue, 0, 64))) {
        if (d->queue) hipFree(d->queue);
        state_destroy(d->st);
        delete d;
        return PVVC_ERR_HIP;
    }
    d->intra_geo.assign(static_cast<size_t>(max_patches), 0u);
    /* set-up of the allocation-free entry point: kernel attributes, the Haar
     * table and this device's scratch address */
    rt_init_kernels();
    if (!scratch_base()) {
        state_destroy(d->st);
        hipFree(d->queue);
        delete d;
        return PVVC_ERR_HIP;
    }
    *out = d;
    return PVVC_OK;
}

int pvvc_decoder_destroy(pvvc_decoder* d) {
    if (!d) return PVVC_ERR_INVALID_ARG;
    hipStreamSynchronize(dec_stream(d));
    if (d->aux) hipStreamSynchronize(d->aux);
    state_destroy(d->st);
    if (d->slab) hipFree(d->slab);
    if (d->queue) hipFree(d->queue);
    if (d->stage) hipFree(d->stage);
    if (d->queue2) hipFree(d->queue2);
    if (d->slab2) hipFree(d->slab2);
    if (d->lane_coef) hipFree(d->lane_coef);
    if (d->lane_status) hipFree(d->lane_status);
    if (d->ev_fork) hipEventDestroy(d->ev_fork);
    if (d->ev_join) hipEventDestroy(d->ev_join);
    if (d->aux) hipStreamDestroy(d->aux);
    delete d;
    return PVVC_OK;
}

int pvvc_decoder_set_tuning(pvvc_decoder* d, const pvvc_tuning* t) {
    if (!d || !t) return PVVC_ERR_INVALID_ARG;
    pvvc_tuning n = d->tuning;
    if (t->waves_per_patch) {
        if (t->waves_per_patch != 1 && t->waves_per_patch != 2 && t->waves_per_patch != 4) return PVVC_ERR_INVALID_ARG;
        n.waves_per_patch = t->waves_per_patch;
    }
    if (t->lds_bytes) n.lds_bytes = t->lds_bytes;
    if (t->wgs_per_cu) n.wgs_per_cu = t->wgs_per_cu;
    if (t->queue_chunk) n.queue_chunk = t->queue_chunk;
    if (t->priority_items) n.priority_items = t->priority_items;
    if (t->lane_max_points) n.lane_max_points = t->lane_max_points < 0 ? 0 : t->lane_max_points;
    d->tuning = n;
    d->auto_small = false;
    return PVVC_OK;
}

int pvvc_decoder_get_tuning(const pvvc_decoder* d, pvvc_tuning* t) {
    if (!d || !t) return PVVC_ERR_INVALID_ARG;
    *t = d->tuning;
    return PVVC_OK;
}

int pvvc_decoder_set_stream(pvvc_decoder* d, void* stream) {
    if (!d) return PVVC_ERR_INVALID_ARG;
    d->has_stream = true;
    d->stream = static_cast<hipStream_t>(stream);
    return PVVC_OK;
}

int pvvc_decoder_set_white_small_patches(pvvc_decoder* d, int enable) {
    if (!d) return PVVC_ERR_INVALID_ARG;
    d->white_small = enable < 0 ? -1 : (enable ? 1 : 0);
    return PVVC_OK;
}

int pvvc_decoder_reset(pvvc_decoder* d) {
    if (!d) return PVVC_ERR_INVALID_ARG;
    if (!hip_ok(hipStreamSynchronize(dec_stream(d)))) return PVVC_ERR_HIP;
    return state_reset(d->st) ? PVVC_OK : PVVC_ERR_HIP;
}

int pvvc_batch_create(pvvc_decoder* d, const pvvc_frame* const* frames, int n_frames, pvvc_batch** out) {
    if (!d || !frames || n_frames <= 0 || !out) return PVVC_ERR_INVALID_ARG;
    std::vector<int> base(n_frames, 0), count(n_frames, d->st.n);
    std::vector<uint32_t> gcap(n_frames, d->st.geo_cap);
    auto* b = new pvvc_batch();
    BatchPlan& pl = b->plan;
    int rc = plan_batch(frames, n_frames, base.data(), count.data(), gcap.data(), d->intra_geo.data(), pl);
    if (rc != PVVC_OK) {
        delete b;
        return rc;
    }
    if (!hip_ok(hipMalloc(&b->dev, std::max<size_t>(pl.total, 256)))) {
        delete b;
        return PVVC_ERR_HIP;
    }
    std::vector<uint8_t> h(pl.total, 0);
    fill_batch_image(pl, frames, b->dev, h.data());
    if (!hip_ok(hipMemcpy(b->dev, h.data(), pl.total, hipMemcpyHostToDevice))) {
        hipFree(b->dev);
        delete b;
        return PVVC_ERR_HIP;
    }
    b->dec = d;
    b->last_stream = dec_stream(d);
    b->n_frames = n_frames;
    b->n_items = pl.n_items;
    b->total_points = pl.total_points;
    b->input_bytes = pl.input_bytes;
    b->state_bytes = pl.state_bytes;
    b->frame_off = pl.frame_off;
    b->frame_pts = pl.frame_pts;
    b->frame_first = pl.frame_first;
    b->frame_count = pl.frame_count;
    b->segs = pl.segs;
    b->items = reinterpret_cast<ItemDev*>(b->dev + pl.o_items);
    b->order = reinterpret_cast<uint32_t*>(b->dev + pl.o_order);
    b->defer = reinterpret_cast<uint32_t*>(b->dev + pl.o_defer);
    b->status = reinterpret_cast<int*>(b->dev + pl.o_status);
    *out = b;
    return PVVC_OK;
}

void pvvc_batch_free(pvvc_batch* b) {
    if (!b) return;
    hipStreamSynchronize(b->last_stream);
    if (b->graph) hipGraphExecDestroy(b->graph);
    if (b->dev) hipFree(b->dev);
    delete b;
}

int pvvc_batch_frames(const pvvc_batch* b) { return b ? b->n_frames : 0; }
int64_t pvvc_batch_frame_points(const pvvc_batch* b, int f) {
    return (b && f >= 0 && f < b->n_frames) ? b->frame_pts[f] : -1;
}
int64_t pvvc_batch_frame_offset(const pvvc_batch* b, int f) {
    return (b && f >= 0 && f < b->n_frames) ? b->frame_off[f] : -1;
}
int64_t pvvc_batch_input_bytes(const pvvc_batch* b) { return b ? b->input_bytes : -1; }
int64_t pvvc_batch_state_bytes(const pvvc_batch* b) { return b ? b->state_bytes : -1; }
int pvvc_batch_items(const pvvc_batch* b) { return b ? static_cast<int>(b->n_items) : 0; }

int pvvc_batch_status(pvvc_batch* b, int* status, int n) {
    if (!b || !status) return PVVC_ERR_INVALID_ARG;
    if (!hip_ok(hipStreamSynchronize(b->last_stream))) return PVVC_ERR_HIP;
    int k = std::min<int>(n, static_cast<int>(b->n_items));
    return hip_ok(hipMemcpy(status, b->status, 4ull * k, hipMemcpyDeviceToHost)) ? PVVC_OK : PVVC_ERR_HIP;
}

}  // extern "C"

namespace {

template <typename T>
bool grow_dev(T*& p, uint64_t& cap, uint64_t need, hipStream_t s) {
    if (need <= cap) return true;
    if (p) {
        hipStreamSynchronize(s);
        hipFree(p);
        p = nullptr;
        cap = 0;
    }
    if (!hip_ok(hipMalloc(reinterpret_cast<void**>(&p), need))) return false;
    cap = need;
    return true;
}

/* A segment whose queue tail [ns, count) holds slices of at most
 * lane_max_points points.  On a second stream, forked from s: lane_rlgr_kernel
 * decodes the tail's RLGR one slice per lane, then a persistent launch with
 * those coefficients does the rest of their decode (own queue counter and
 * workspaces).  On s, meanwhile, the scalar persistent launch decodes the
 * head [0, ns).  The lane kernel is enqueued first, but the command
 * processor may dispatch the scalar launch first, whose workgroups then fill
 * every CU (VGPRs) and hold the lane kernel's 33 KiB workgroups off until they
 * drain: the span varies from launch to launch (c2: 4.7 to 5.7 ms).  The
 * streams join before the finish pass of the deferred predictive slices. */
int launch_segment_lanes(pvvc_decoder* d, pvvc_batch* b, const BatchDev& bd, const Segment& sg, uint32_t ns,
                         const Sink& sink, const pvvc_tuning& t, const LaunchRes& res, uint32_t launch_id,
                         hipStream_t s) {
    const uint32_t nl = sg.count - ns;
    if (!d->aux) {
        if (!hip_ok(hipStreamCreateWithFlags(&d->aux, hipStreamNonBlocking)) ||
            !hip_ok(hipEventCreateWithFlags(&d->ev_fork, hipEventDisableTiming)) ||
            !hip_ok(hipEventCreateWithFlags(&d->ev_join, hipEventDisableTiming)) ||
            !hip_ok(hipMalloc(&d->queue2, 64)) || !hip_ok(hipMemset(d->queue2, 0, 64)))
            return PVVC_ERR_HIP;
    }
    /* the tail's workspace bound and the second launch's slab */
    uint64_t need2 = 256;
    for (uint32_t q = sg.begin + ns; q < sg.begin + sg.count; ++q) {
        const ItemDev& x = b->plan.items[b->plan.order[q]];
        need2 = std::max(need2, item_ws_bound(x.n, x.geo_size ? x.geo_size : d->st.geo_cap));
    }
    need2 = (need2 + 255) & ~uint64_t(255);
    pvvc_tuning t2 = t;
    t2.waves_per_patch = 1;
    const uint32_t grid2 = persistent_grid(t2, nl);
    if (!grow_dev(d->slab2, d->slab2_total, need2 * grid2, s) ||
        !grow_dev(d->lane_coef, d->lane_coef_bytes, 12ull * std::max<int64_t>(b->total_points, 1), s) ||
        !grow_dev(d->lane_status, d->lane_status_bytes, 4ull * std::max<uint32_t>(b->n_items, 1), s))
        return PVVC_ERR_HIP;
    auto kev = g_ktiming.next();
    trec(kev.first, s);
    hipEventRecord(d->ev_fork, s);
    hipStreamWaitEvent(d->aux, d->ev_fork, 0);
    hipLaunchKernelGGL(lane_rlgr_kernel, dim3((nl + 63) / 64), dim3(64), 0, d->aux, bd.items, bd.order + sg.begin + ns,
                       nl, d->lane_coef, d->lane_status);
    int rc = PVVC_OK;
    if (ns) {
        Segment head = sg;
        head.count = ns;
        head.dcount = 0;
        LaunchRes r1 = res;
        r1.timed = false;
        rc = launch_segment(bd, head, sink, t, r1, launch_id, s);
    }
    Segment tail{sg.begin + ns, nl, 0, 0, need2};
    LaunchRes r2{res.slots, d->slab2, need2, d->slab2_total, res.stage, d->queue2, 0u};
    r2.pre_coef = d->lane_coef;
    r2.pre_status = d->lane_status;
    r2.timed = false;
    if (rc == PVVC_OK) rc = launch_segment(bd, tail, sink, t2, r2, launch_id, d->aux);
    hipEventRecord(d->ev_join, d->aux);
    hipStreamWaitEvent(s, d->ev_join, 0);
    trec(kev.second, s);
    if (rc == PVVC_OK && sg.dcount)
        hipLaunchKernelGGL(finish_kernel, dim3(sg.dcount), dim3(256), 0, s, bd.items, bd.defer + sg.dbegin, sg.dcount,
                           res.slots, sink, res.stage, bd.status);
    return rc;
}

}  // namespace

extern "C" {

int pvvc_decode_batch(pvvc_decoder* d, pvvc_batch* b, int mode, int output, float* xyz, uint8_t* rgb,
                      Points* points) {
    if (!d || !b || b->dec != d) return PVVC_ERR_INVALID_ARG;
    if (output == PVVC_OUT_NATIVE && (!xyz || !rgb)) return PVVC_ERR_INVALID_ARG;
    if (output == PVVC_OUT_POINTS && !points) return PVVC_ERR_INVALID_ARG;
    rt_init_kernels();
    const hipStream_t s = dec_stream(d);
    b->last_stream = s;
    const int white = dec_white_small(d);
    Sink sink{xyz, rgb, points, output, white};
    const int W = d->tuning.waves_per_patch;
    const uint32_t lds = static_cast<uint32_t>(clamp_lds(d->tuning.lds_bytes));
    g_timing.reset();
    g_ktiming.reset();
    if (mode == PVVC_MODE_PER_FRAME || mode == PVVC_MODE_GRAPH) {
        if (mode == PVVC_MODE_GRAPH) {
            void* ptrs[3] = {xyz, rgb, points};
            const bool same = b->graph && b->graph_output == output && std::equal(ptrs, ptrs + 3, b->graph_ptrs) &&
                              b->graph_waves == W && b->graph_lds == lds && b->graph_white == white;
            if (!same) {
                if (b->graph) hipGraphExecDestroy(b->graph);
                b->graph = nullptr;
                hipStream_t cs;
                if (!hip_ok(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking))) return PVVC_ERR_HIP;
                hipGraph_t g = nullptr;
                bool good = hip_ok(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
                if (good) {
                    for (int f = 0; f < b->n_frames; ++f) {
                        if (W == 2) launch_frame<2>(b, f, sink, lds, cs);
                        else if (W == 4) launch_frame<4>(b, f, sink, lds, cs);
                        else launch_frame<1>(b, f, sink, lds, cs);
                    }
                    good = hip_ok(hipStreamEndCapture(cs, &g));
                }
                if (good) good = hip_ok(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0));
                if (g) hipGraphDestroy(g);
                hipStreamDestroy(cs);
                if (!good) {
                    b->graph = nullptr;
                    return PVVC_ERR_HIP;
                }
                b->graph_output = output;
                std::copy(ptrs, ptrs + 3, b->graph_ptrs);
                b->graph_waves = W;
                b->graph_lds = lds;
                b->graph_white = white;
            }
            auto ev = g_timing.next();
            trec(ev.first, s);
            const bool ok_launch = hip_ok(hipGraphLaunch(b->graph, s));
            trec(ev.second, s);
            return ok_launch ? PVVC_OK : PVVC_ERR_HIP;
        }
        for (int f = 0; f < b->n_frames; ++f) {
            if (!b->frame_count[f]) continue;
            auto ev = g_timing.next();
            auto kev = g_ktiming.next();
            trec(ev.first, s);
            trec(kev.first, s);
            if (W == 2) launch_frame<2>(b, f, sink, lds, s);
            else if (W == 4) launch_frame<4>(b, f, sink, lds, s);
            else launch_frame<1>(b, f, sink, lds, s);
            trec(kev.second, s);
            trec(ev.second, s);
        }
        return hip_ok(hipGetLastError()) ? PVVC_OK : PVVC_ERR_HIP;
    }
    if (mode != PVVC_MODE_PERSISTENT) return PVVC_ERR_INVALID_ARG;
    /* this call's workspace stride and slab: every workgroup of every
     * segment gets `need` bytes at blockIdx * need */
    uint64_t need = 256;
    uint32_t grid_max = 1;
    for (auto& sg : b->segs) {
        need = std::max(need, sg.ws_need);
        grid_max = std::max(grid_max, persistent_grid(segment_tuning(d->tuning, d->auto_small, sg.count), sg.count));
    }
    const uint64_t slab_need = need * grid_max;
    if (slab_need > d->slab_total) {
        if (d->slab) {
            hipStreamSynchronize(s);
            hipFree(d->slab);
            d->slab = nullptr;
            d->slab_total = 0;
        }
        if (!hip_ok(hipMalloc(&d->slab, slab_need))) return PVVC_ERR_HIP;
        d->slab_total = slab_need;
    }
    d->slab_bytes = need;
    const uint64_t stage_need = 12ull * static_cast<uint64_t>(std::max<int64_t>(b->total_points, 1));
    if (stage_need > d->stage_bytes) {
        if (d->stage) {
            hipStreamSynchronize(s);
            hipFree(d->stage);
            d->stage = nullptr;
        }
        if (!hip_ok(hipMalloc(&d->stage, stage_need))) {
            d->stage_bytes = 0;
            return PVVC_ERR_HIP;
        }
        d->stage_bytes = stage_need;
    }
    const BatchDev bd{b->items, b->order, b->defer, b->status};
    const LaunchRes res{d->st.slots, d->slab, need, d->slab_total, d->stage, d->queue, 0u};
    for (auto& sg : b->segs) {
        if (!sg.count) continue;
        const uint32_t id = ++d->launch_id == 0 ? ++d->launch_id : d->launch_id;
        const pvvc_tuning tt = segment_tuning(d->tuning, d->auto_small, sg.count);
        /* slices of at most lane_max_points points: the tail of the
         * length-sorted queue */
        uint32_t ns = sg.count;
        if (tt.lane_max_points > 0)
            while (ns > 0 && b->plan.items[b->plan.order[sg.begin + ns - 1]].n <= static_cast<uint32_t>(tt.lane_max_points))
                --ns;
        auto ev = g_timing.next();
        trec(ev.first, s);
        const int rc = ns == sg.count ? launch_segment(bd, sg, sink, tt, res, id, s)
                                      : launch_segment_lanes(d, b, bd, sg, ns, sink, tt, res, id, s);
        trec(ev.second, s);
        if (rc != PVVC_OK) return rc;
    }
    return hip_ok(hipGetLastError()) ? PVVC_OK : PVVC_ERR_HIP;
}

/* Synchronous one-call forms of the batch API: pack, decode, wait, free.
 * The return value is the call's own error, else the first non-OK slice
 * status in item order (frames in order, each frame's slices in file order),
 * else PVVC_OK; `status` (optional) receives every slice's status. */
namespace {
int decode_once(pvvc_decoder* d, const pvvc_frame* const* frames, int n_frames, int mode, int output, float* xyz,
                uint8_t* rgb, Points* points, int* status, int status_cap) {
    if (!d || !frames || n_frames <= 0 || (status && status_cap < 0)) return PVVC_ERR_INVALID_ARG;
    pvvc_batch* b = nullptr;
    int rc = pvvc_batch_create(d, frames, n_frames, &b);
    if (rc != PVVC_OK) return rc;
    rc = pvvc_decode_batch(d, b, mode, output, xyz, rgb, points);
    std::vector<int> st(b->n_items, PVVC_OK);
    if (rc == PVVC_OK) rc = pvvc_batch_status(b, st.data(), static_cast<int>(st.size()));
    else hipStreamSynchronize(b->last_stream);
    if (rc == PVVC_OK) {
        if (status) std::copy(st.begin(), st.begin() + std::min<size_t>(st.size(), status_cap), status);
        for (int s : st)
            if (s != PVVC_OK) {
                rc = s;
                break;
            }
    }
    pvvc_batch_free(b);
    return rc;
}
}  // namespace

int pvvc_decode_frame(pvvc_decoder* dec, const pvvc_frame* frame, int output, float* xyz, uint8_t* rgb,
                      Points* points, int* status, int status_cap) {
    return decode_once(dec, &frame, 1, PVVC_MODE_PER_FRAME, output, xyz, rgb, points, status, status_cap);
}

int pvvc_decode_gop(pvvc_decoder* dec, const pvvc_frame* const* frames, int n_frames, int output, float* xyz,
                    uint8_t* rgb, Points* points, int* status, int status_cap) {
    return decode_once(dec, frames, n_frames, PVVC_MODE_PERSISTENT, output, xyz, rgb, points, status, status_cap);
}

/* The persistent decode of pvvc_decode_batch with every device buffer it
 * would allocate taken from `scratch`: [queue counters 256 B][stage: 12 B per
 * point of the batch][workspaces: one per workgroup at this call's stride].
 * The grid is cut to the workspaces that fit.  Enqueues one counter-reset
 * kernel and the segments' launches; no allocation, no memset, no events. */
int pvvc_decode_batch_noalloc(pvvc_decoder* d, pvvc_batch* b, int output, float* xyz, uint8_t* rgb, Points* points) {
    if (!d || !b || b->dec != d) return PVVC_ERR_INVALID_ARG;
    if (output == PVVC_OUT_NATIVE && (!xyz || !rgb)) return PVVC_ERR_INVALID_ARG;
    if (output == PVVC_OUT_POINTS && !points) return PVVC_ERR_INVALID_ARG;
    if (output != PVVC_OUT_NATIVE && output != PVVC_OUT_POINTS) return PVVC_ERR_INVALID_ARG;
    uint8_t* base = scratch_base();
    if (!base) return PVVC_ERR_HIP;
    auto al = [](uint64_t x) { return (x + 255) & ~uint64_t(255); };
    const uint64_t o_stage = 256;
    const uint64_t stage_bytes = al(12ull * static_cast<uint64_t>(std::max<int64_t>(b->total_points, 1)));
    uint64_t need = 256;
    for (auto& sg : b->segs) need = std::max(need, sg.ws_need);
    if (o_stage + stage_bytes + need > kScratchBytes) {
        set_error(PVVC_ERR_CAPACITY);
        return PVVC_ERR_CAPACITY; /* the batch's stage or one workspace exceeds the scratch */
    }
    const uint64_t o_slab = o_stage + stage_bytes;
    const uint64_t slab_total = kScratchBytes - o_slab;
    const uint32_t grid_cap = static_cast<uint32_t>(std::min<uint64_t>(slab_total / need, 0xffffffffu));
    const hipStream_t s = dec_stream(d);
    b->last_stream = s;
    const Sink sink{xyz, rgb, points, output, dec_white_small(d)};
    uint32_t* queue = reinterpret_cast<uint32_t*>(base);
    /* the counters start at zero: a kernel, not a memset, so that the call
     * can be captured into a graph */
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, s, queue);
    if (!hip_ok(hipGetLastError())) return PVVC_ERR_HIP;
    const BatchDev bd{b->items, b->order, b->defer, b->status};
    LaunchRes res{d->st.slots, base + o_slab, need, slab_total, reinterpret_cast<float*>(base + o_stage), queue,
                  grid_cap};
    res.timed = false;
    for (auto& sg : b->segs) {
        if (!sg.count) continue;
        const uint32_t id = ++d->launch_id == 0 ? ++d->launch_id : d->launch_id;
        const int rc = launch_segment(bd, sg, sink, segment_tuning(d->tuning, d->auto_small, sg.count), res, id, s);
        if (rc != PVVC_OK) return rc;
    }
    return hip_ok(hipGetLastError()) ? PVVC_OK : PVVC_ERR_HIP;
}

}  // extern "C"

/* ---- forward path (colour encoder) ---- */
namespace {

/* Buffers of the colour encoder, kept between calls and only ever grown: the
 * device image, a pinned host copy of its input part (same offsets) and a
 * pinned landing area for the packed output bytes.  Calls are serialised by
 * g_enc_mu. */
struct EncBuffers {
    uint8_t* dev = nullptr;
    uint64_t dev_cap = 0;
    uint8_t* pin = nullptr;
    uint64_t pin_cap = 0;
    uint8_t* pin_out = nullptr;
    uint64_t pin_out_cap = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
};
EncBuffers g_enc;
std::mutex g_enc_mu;

uint64_t grow_size(uint64_t need) { return ((need + need / 4) + 0xFFFFF) & ~uint64_t(0xFFFFF); }

bool enc_grow_dev(uint64_t need, hipStream_t es) {
    if (need <= g_enc.dev_cap) return true;
    if (g_enc.dev) {
        hipStreamSynchronize(es);
        hipFree(g_enc.dev);
        g_enc.dev = nullptr;
        g_enc.dev_cap = 0;
    }
    const uint64_t cap = grow_size(need);
    if (!hip_ok(hipMalloc(&g_enc.dev, cap))) return false;
    g_enc.dev_cap = cap;
    return true;
}

bool enc_grow_pinned(uint8_t*& p, uint64_t& cap_now, uint64_t need, hipStream_t es) {
    if (need <= cap_now) return true;
    if (p) {
        hipStreamSynchronize(es);
        hipHostFree(p);
        p = nullptr;
        cap_now = 0;
    }
    const uint64_t cap = grow_size(need);
    if (!hip_ok(hipHostMalloc(reinterpret_cast<void**>(&p), cap, hipHostMallocDefault))) return false;
    cap_now = cap;
    return true;
}

int host_threads(int tasks) {
    const unsigned hw = std::thread::hardware_concurrency();
    return std::max(1, std::min({tasks, 8, hw ? static_cast<int>(hw) : 1}));
}

}  // namespace

extern "C" {

/* Steps of one call:
 *  1. layout: device image [EncJobDev x n, longest first][results][offsets][inputs: geo |
 *     yuv per job][outputs][workspaces]; the pinned buffer holds the first
 *     four parts at the same offsets;
 *  2. host threads copy the jobs' inputs into the pinned buffer, in chunks of
 *     consecutive jobs; each finished chunk goes to the device by an
 *     asynchronous copy while the next ones are filled;
 *  3. one encode launch (one wavefront per job), the byte counts back;
 *  4. a gather launch packs the variable-length outputs into the (now free)
 *     input region, one copy brings them into pinned memory, and host threads
 *     scatter them into the jobs' buffers. */
int pvvc_encode_colors(pvvc_color_job* jobs, int n_jobs, double* kernel_ms) {
    if (!jobs || n_jobs <= 0) return PVVC_ERR_INVALID_ARG;
    rt_init_kernels();
    std::lock_guard<std::mutex> lk(g_enc_mu);
    const hipStream_t es = default_stream(); /* one stream for the whole call */
    auto al = [](uint64_t x) { return (x + 255) & ~uint64_t(255); };
    const uint64_t o_jobs = 0, o_res = al(sizeof(EncJobDev) * n_jobs), o_at = al(o_res + 8ull * n_jobs);
    const uint64_t in_begin = al(o_at + 8ull * n_jobs);
    uint64_t at = in_begin;
    std::vector<uint64_t> o_geo(n_jobs), o_yuv(n_jobs), o_out(n_jobs), o_ws(n_jobs), ws_cap(n_jobs), out_cap(n_jobs);
    /* inputs (geometry, leaf YUV) of all jobs first, in job order */
    for (int i = 0; i < n_jobs; ++i) {
        const pvvc_color_job& j = jobs[i];
        if (!j.geometry || j.geometry_size < 25 || !j.yuv || j.n == 0 || !j.out || j.qp == 0)
            return PVVC_ERR_INVALID_ARG;
        o_geo[i] = at;
        at = al(at + j.geometry_size);
        o_yuv[i] = at;
        at = al(at + 12ull * j.n);
    }
    const uint64_t in_end = at;
    for (int i = 0; i < n_jobs; ++i) {
        const uint64_t n = jobs[i].n, imax = jobs[i].geometry_size - 25u;
        out_cap[i] = (std::min<uint64_t>(jobs[i].out_cap, 24 * n + 64) + 3) & ~uint64_t(3);
        o_out[i] = at;
        at = al(at + out_cap[i] + 4);
        ws_cap[i] = al(12 * n + 2 * (4 * (imax + 1) + 16) + 24 * n + 512);
        o_ws[i] = at;
        at += ws_cap[i];
    }
    const uint64_t total = at;
    if (!enc_grow_dev(total, es) || !enc_grow_pinned(g_enc.pin, g_enc.pin_cap, in_end, es)) return PVVC_ERR_HIP;
    if (!g_enc.e0 && !hip_ok(hipEventCreate(&g_enc.e0))) return PVVC_ERR_HIP;
    if (!g_enc.e1 && !hip_ok(hipEventCreate(&g_enc.e1))) return PVVC_ERR_HIP;
    uint8_t* const dev = g_enc.dev;
    uint8_t* const h = g_enc.pin;
    {
        /* longest jobs first, so that the last workgroups to start are short */
        std::vector<int> order(n_jobs);
        for (int i = 0; i < n_jobs; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return jobs[x].n > jobs[y].n; });
        EncJobDev* jd = reinterpret_cast<EncJobDev*>(h + o_jobs);
        for (int r = 0; r < n_jobs; ++r) {
            const int i = order[r];
            EncJobDev d;
            d.geo = dev + o_geo[i];
            d.yuv = reinterpret_cast<const float*>(dev + o_yuv[i]);
            d.out = dev + o_out[i];
            d.ws = dev + o_ws[i];
            d.ws_cap = ws_cap[i];
            d.geo_size = jobs[i].geometry_size;
            d.n = jobs[i].n;
            d.out_cap = static_cast<uint32_t>(out_cap[i]);
            d.qp = jobs[i].qp;
            d.idx = static_cast<uint32_t>(i);
            jd[r] = d;
        }
    }
    bool good = hip_ok(hipMemcpyAsync(dev + o_jobs, h + o_jobs, o_res, hipMemcpyHostToDevice, es));
    /* chunks of consecutive jobs of about 4 MiB of input each (at most 64) */
    const uint64_t in_bytes = in_end - in_begin;
    const int n_chunks = static_cast<int>(std::max<uint64_t>(1, std::min<uint64_t>({64, uint64_t(n_jobs),
                                                                                     in_bytes >> 22})));
    std::vector<int> first(n_chunks + 1, n_jobs);
    for (int c = 0, i = 0; c < n_chunks; ++c) {
        first[c] = i;
        const uint64_t stop = in_begin + in_bytes * (c + 1) / n_chunks;
        while (i < n_jobs && (o_geo[i] < stop || i == first[c])) ++i;
        if (c == n_chunks - 1) i = n_jobs;
    }
    std::unique_ptr<std::atomic<int>[]> done(new std::atomic<int>[n_chunks]);
    for (int c = 0; c < n_chunks; ++c) done[c].store(0);
    auto fill = [&](int c) {
        for (int i = first[c]; i < first[c + 1]; ++i) {
            std::memcpy(h + o_geo[i], jobs[i].geometry, jobs[i].geometry_size);
            std::memcpy(h + o_yuv[i], jobs[i].yuv, 12ull * jobs[i].n);
        }
        done[c].store(1, std::memory_order_release);
    };
    const int T = host_threads(n_chunks);
    std::vector<std::thread> pool;
    for (int t = 0; t < T; ++t)
        pool.emplace_back([&, t] {
            for (int c = t; c < n_chunks; c += T) fill(c);
        });
    /* chunk c's bytes are [start(c), start(c + 1)); a chunk may be empty when
     * the ones before it took every job */
    auto start = [&](int c) { return c == 0 ? in_begin : first[c] < n_jobs ? o_geo[first[c]] : in_end; };
    for (int c = 0; c < n_chunks; ++c) {
        while (!done[c].load(std::memory_order_acquire)) std::this_thread::yield();
        const uint64_t b = start(c), e = start(c + 1);
        if (good && e > b) good = hip_ok(hipMemcpyAsync(dev + b, h + b, e - b, hipMemcpyHostToDevice, es));
    }
    for (auto& th : pool) th.join();
    /* 8 KiB per stream lets 16 encode workgroups share a CU, the most its
     * registers allow; a c1 GoP (6 560 streams) encoded in 3.80 ms against
     * 4.65 ms with 16 KiB (10 per CU) */
    const uint32_t lds = static_cast<uint32_t>(clamp_lds(8 * 1024));
    std::vector<int64_t> res(n_jobs, 0);
    if (good) {
        hipEventRecord(g_enc.e0, es);
        const uint32_t wave1 = static_cast<uint32_t>(std::max(1, devinfo().cu) * 16);
        hipLaunchKernelGGL(encode_kernel, dim3(n_jobs), dim3(64), lds, es,
                           reinterpret_cast<const EncJobDev*>(dev + o_jobs), static_cast<uint32_t>(n_jobs), lds,
                           reinterpret_cast<int64_t*>(dev + o_res), wave1);
        hipEventRecord(g_enc.e1, es);
        good = hip_ok(hipGetLastError()) &&
               hip_ok(hipMemcpyAsync(res.data(), dev + o_res, 8ull * n_jobs, hipMemcpyDeviceToHost, es)) &&
               hip_ok(hipStreamSynchronize(es));
    }
    if (good && kernel_ms) {
        float ms = 0;
        hipEventElapsedTime(&ms, g_enc.e0, g_enc.e1);
        *kernel_ms = ms;
    }
    /* compact the variable-length outputs, one copy back */
    std::vector<uint64_t> pos(n_jobs, 0);
    uint64_t packed_bytes = 0;
    for (int i = 0; i < n_jobs; ++i) {
        pos[i] = packed_bytes;
        if (res[i] > 0) packed_bytes += static_cast<uint64_t>(res[i]);
    }
    uint8_t* spill = nullptr; /* only when the outputs outgrow the input region */
    if (good && packed_bytes) {
        uint8_t* packed = dev + in_begin;
        if (packed_bytes > in_bytes) {
            good = hip_ok(hipMalloc(&spill, packed_bytes));
            packed = spill;
        }
        good = good && enc_grow_pinned(g_enc.pin_out, g_enc.pin_out_cap, packed_bytes, es) &&
               hip_ok(hipMemcpyAsync(dev + o_at, pos.data(), 8ull * n_jobs, hipMemcpyHostToDevice, es));
        if (good) {
            hipLaunchKernelGGL(gather_kernel, dim3(n_jobs), dim3(256), 0, es,
                               reinterpret_cast<const EncJobDev*>(dev + o_jobs),
                               reinterpret_cast<const int64_t*>(dev + o_res),
                               reinterpret_cast<const uint64_t*>(dev + o_at), static_cast<uint32_t>(n_jobs), packed);
            good = hip_ok(hipGetLastError()) &&
                   hip_ok(hipMemcpyAsync(g_enc.pin_out, packed, packed_bytes, hipMemcpyDeviceToHost, es)) &&
                   hip_ok(hipStreamSynchronize(es));
        }
        if (good) {
            const uint8_t* hp = g_enc.pin_out;
            auto scatter = [&](int t, int T2) {
                for (int i = static_cast<int>(int64_t(n_jobs) * t / T2); i < int64_t(n_jobs) * (t + 1) / T2; ++i)
                    if (res[i] > 0) {
                        if (static_cast<uint64_t>(res[i]) > jobs[i].out_cap) res[i] = -PVVC_ERR_CAPACITY;
                        else std::memcpy(jobs[i].out, hp + pos[i], static_cast<size_t>(res[i]));
                    }
            };
            const int T2 = host_threads(static_cast<int>(std::min<uint64_t>(packed_bytes >> 20, 64) + 1));
            std::vector<std::thread> sp;
            for (int t = 1; t < T2; ++t) sp.emplace_back(scatter, t, T2);
            scatter(0, T2);
            for (auto& th : sp) th.join();
        }
    }
    for (int i = 0; i < n_jobs; ++i) {
        jobs[i].out_size = res[i] > 0 ? static_cast<uint32_t>(res[i]) : 0u;
        jobs[i].status = !good ? PVVC_ERR_HIP : (res[i] > 0 ? PVVC_OK : static_cast<int>(-res[i]));
    }
    if (spill) {
        hipStreamSynchronize(es);
        hipFree(spill);
    }
    return good ? PVVC_OK : PVVC_ERR_HIP;
}

/* Frees the colour encoder's kept buffers (they are made again on the next call). */
void pvvc_encoder_release(void) {
    std::lock_guard<std::mutex> lk(g_enc_mu);
    hipStreamSynchronize(default_stream());
    if (g_enc.dev) hipFree(g_enc.dev);
    if (g_enc.pin) hipHostFree(g_enc.pin);
    if (g_enc.pin_out) hipHostFree(g_enc.pin_out);
    if (g_enc.e0) hipEventDestroy(g_enc.e0);
    if (g_enc.e1) hipEventDestroy(g_enc.e1);
    g_enc = EncBuffers{};
}

/* Autotuner: coordinate descent over the launch knobs on a real batch.
 * Each candidate decodes the batch once untimed, then `reps` times; its
 * cost is the median of the summed event times of all launches of a call.
 * Decoding a batch again is idempotent (intra slices rewrite the state they
 * read), so the output equals a single decode's. */
int pvvc_decoder_autotune(pvvc_decoder* d, pvvc_batch* b, int mode, int output, float* xyz, uint8_t* rgb,
                          Points* points, int reps, pvvc_tuning* best, double* best_ms) {
    if (!d || !b || b->dec != d) return PVVC_ERR_INVALID_ARG;
    if (mode != PVVC_MODE_PER_FRAME && mode != PVVC_MODE_PERSISTENT && mode != PVVC_MODE_GRAPH)
        return PVVC_ERR_INVALID_ARG;
    reps = std::max(1, std::min(reps, 50));
    d->auto_small = false; /* every candidate runs as given */
    /* the candidates are timed by events whatever the caller's setting */
    struct TimingOn {
        bool prev = g_timing_on;
        TimingOn() { g_timing_on = true; }
        ~TimingOn() { g_timing_on = prev; }
    } timing_on;
    static const int kWaves[] = {1, 2, 4};
    static const int kLds[] = {4096, 8192, 16384, 24576, 32768, 49152, 65536};
    static const int kWgs[] = {2, 4, 8, 12, 16, 20, 24, 32};
    static const int kChunk[] = {1, 2, 4, 8};
    static const int kLane[] = {0, 256, 1000, 2000, 4000}; /* lane_max_points: 0 = the scalar decoder only */
    const bool persistent = mode == PVVC_MODE_PERSISTENT;
    std::map<std::tuple<int, int, int, int, int>, double> seen;
    auto measure = [&](const pvvc_tuning& t, double& ms) -> int {
        auto key = std::make_tuple(t.waves_per_patch, t.lds_bytes, t.wgs_per_cu, t.queue_chunk, t.lane_max_points);
        auto f = seen.find(key);
        if (f != seen.end()) {
            ms = f->second;
            return PVVC_OK;
        }
        d->tuning = t;
        int rc = pvvc_decode_batch(d, b, mode, output, xyz, rgb, points);
        if (rc == PVVC_OK && !hip_ok(hipStreamSynchronize(dec_stream(d)))) rc = PVVC_ERR_HIP;
        std::vector<double> v;
        for (int r = 0; r < reps && rc == PVVC_OK; ++r) {
            rc = pvvc_decode_batch(d, b, mode, output, xyz, rgb, points);
            if (rc == PVVC_OK && !hip_ok(hipStreamSynchronize(dec_stream(d)))) rc = PVVC_ERR_HIP;
            v.push_back(pvvc_last_kernel_ms());
        }
        if (rc != PVVC_OK) return rc;
        std::sort(v.begin(), v.end());
        ms = v[v.size() / 2];
        seen[key] = ms;
        return PVVC_OK;
    };
    pvvc_tuning cur = d->tuning;
    double cur_ms = 0;
    int rc = measure(cur, cur_ms);
    if (rc != PVVC_OK) return rc;
    for (int pass = 0; pass < 3; ++pass) {
        const pvvc_tuning start = cur;
        for (int dim = 0; dim < 5; ++dim) {
            if (!persistent && dim >= 2) break; /* grid shape is the slice count there */
            const int* vals = dim == 0 ? kWaves : dim == 1 ? kLds : dim == 2 ? kWgs : dim == 3 ? kChunk : kLane;
            const int nv = dim == 0 ? 3 : dim == 1 ? 7 : dim == 2 ? 8 : dim == 3 ? 4 : 5;
            for (int k = 0; k < nv; ++k) {
                pvvc_tuning t = cur;
                int* field = dim == 0   ? &t.waves_per_patch
                             : dim == 1 ? &t.lds_bytes
                             : dim == 2 ? &t.wgs_per_cu
                             : dim == 3 ? &t.queue_chunk
                                        : &t.lane_max_points;
                if (*field == vals[k]) continue;
                *field = vals[k];
                double ms = 0;
                rc = measure(t, ms);
                if (rc != PVVC_OK) {
                    d->tuning = cur;
                    return rc;
                }
                if (ms < cur_ms * 0.99) {
                    cur = t;
                    cur_ms = ms;
                }
            }
        }
        if (std::memcmp(&start, &cur, sizeof(cur)) == 0) break;
    }
    d->tuning = cur;
    if (best) *best = cur;
    if (best_ms) *best_ms = cur_ms;
    /* leave the outputs as the chosen configuration writes them */
    rc = pvvc_decode_batch(d, b, mode, output, xyz, rgb, points);
    if (rc == PVVC_OK && !hip_ok(hipStreamSynchronize(dec_stream(d)))) rc = PVVC_ERR_HIP;
    return rc;
}

}  // extern "C"

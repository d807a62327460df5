"""The decoder's hand-written scalar RLGR loop, checked on the CPU.

`rlgr_fast_run` (patchvvc-plus_amd/csrc/pvvc_decode.hip) is GCN scalar
assembly.  _rlgr_asm.py executes its text instruction by instruction and
states the RLGR decoder (entropy_codec.cc:207-283) independently.  Entry
states are cut from real streams (the oracle's encoder over the crafted
symbol kinds of the GPU stress test) at event boundaries, with random window
fill, refill lane, chunk limit and symbol index; a third of them take a random
k_P and k_RP in place of the stream's (so that every k_P block of the loop is
entered) and a quarter start a few dwords before a chunk rotation, so that
every exit (chunk complete, chunk rotation, long codes, run prefixes, runs
crossing the chunk) is reached and every instruction runs.  Each exit state
must equal the decoder's state after a whole number of events: bits consumed,
window contents, k_P, k_RP, the symbol index and every accumulator lane; a run
whose zeros reach the chunk end may instead stop after its prefix and length
field (the caller decodes its symbol), which is checked against the decoder's
state at that point of the event."""
import numpy as np
import pytest

import _rlgr_asm as R
from test_gpu_rlgr_stress import KINDS, _symbols


def _cases(oracle, seed, streams_per_kind, states_per_stream):
    rng = np.random.default_rng(seed)
    for kind in KINDS:
        for _ in range(streams_per_kind):
            L = int(rng.integers(300, 5000))
            sym = _symbols(kind, L, rng)
            bits = R.to_bits(oracle.rlgr_encode(sym))
            bnd = R.decode_boundaries(bits, L)
            for _ in range(states_per_stream):
                pos, kP, kRP, _n = bnd[int(rng.integers(0, len(bnd)))]
                if rng.random() < 1 / 3:
                    # the decoder's state at this bit replaced by any other:
                    # the bits after it are then read as that state reads
                    # them, which enters every k_P state of the loop
                    kP, kRP = int(rng.integers(0, 16)), int(rng.integers(0, 101))
                nb = int(rng.integers(R.NB_MIN, R.NB_MAX + 1))
                wpos = int(rng.integers(0, 1 << 20))
                if rng.random() < 1 / 4:
                    # a refill within the next few dwords rotates the chunk
                    wpos = (wpos & ~63) | int(rng.integers(60, 64))
                lim = int(rng.integers(1, 65))
                nrel = int(rng.integers(0, lim))
                yield kind, bits, pos, R.make_state(bits, pos, kP, kRP, rng, nb, wpos, lim, nrel)


def _check(oracle, lines, seed=5, streams_per_kind=3, states_per_stream=320):
    loop = R.Loop(lines)
    visited = set()
    fails = []
    for kind, bits, pos, st0 in _cases(oracle, seed, streams_per_kind, states_per_stream):
        st1, _ = loop.run(st0, visited=visited)
        err = R.check_against_decoder(bits, pos, st0, st1)
        if err:
            fails.append((kind, err))
    return fails, len(visited), len(loop.ins)


def test_decoder_statement_matches_oracle(oracle):
    """The Python statement of the decoder reproduces the symbols (and so the
    oracle's decode) of every crafted stream."""
    rng = np.random.default_rng(1)
    for kind in KINDS:
        L = int(rng.integers(300, 4000))
        sym = _symbols(kind, L, rng)
        data = oracle.rlgr_encode(sym)
        got, n = [], 0
        for syms, adv, _, _ in R.events(R.Bits(R.to_bits(data)), 0, 8):
            (m, raw), = syms
            if adv > 1 or m:  # run: m zeros, then the symbol stored as u + 1
                got += [0] * m
            got.append((raw >> 1) ^ -(raw & 1))
            n += adv
            if n >= L:
                break
        np.testing.assert_array_equal(np.array(got[:L]), sym, err_msg=kind)
        ref, st = oracle.rlgr_decode(data, L)
        assert st == 0
        np.testing.assert_array_equal(np.asarray(ref), sym, err_msg=kind)


def test_scalar_loop_matches_decoder(oracle):
    fails, covered, n_ins = _check(oracle, R.extract_loop())
    assert not fails, fails[:5]
    assert covered == n_ins, f"{n_ins - covered} of {n_ins} loop instructions never executed"


@pytest.mark.parametrize("old,new", [
    ("s_branch 104b", "s_branch 105b"),                              # zero symbol at k_P 1: k_P 4 next
    ("s_cbranch_scc1 604f", "s_cbranch_scc1 605f"),                  # chunk end: k_P after the symbol
    ("s_addc_u32 s86, s84, 2", "s_addc_u32 s86, s84, 1"),            # run code: k + 1 bits before it
    ("s_sub_u32 s88, 63, s85", "s_sub_u32 s88, 62, s85"),            # zero test over k_R + 1 bits
    ("v_writelane_b32 %[aF], 1, m0", "v_writelane_b32 %[aF], 0, m0"),  # run symbol stored as u, not u + 1
    ("v_writelane_b32 %[aK], s85, m0", "v_writelane_b32 %[aK], s84, m0"),  # remainder width k_R
    ("v_writelane_b32 %[aT], s95, m0", "v_writelane_b32 %[aT], s94, m0"),  # code bits: the window's top
    ("s_cmp_lg_u32 s84, 0", "s_cmp_eq_u32 s84, 0"),                  # k_RP step: p != 0
    ("s_max_i32 %[kRP], s91, 0", "s_mov_b32 %[kRP], s91"),           # run path: k_RP floor dropped
    ("s_ashr_i64 s[90:91], s[94:95], 62", "s_ashr_i64 s[90:91], s[94:95], 61"),  # k = 1: one length bit
    ("s_sub_u32 m0, m0, s90", "s_sub_u32 m0, m0, 1"),                # run length added before the prefix test
    ("s_lshr_b32 s85, %[kP], 2", "s_lshr_b32 s85, %[kP], 1"),        # k_P >= 12: k = k_P / 4
    ("s_add_u32 s90, s90, 2", "s_add_u32 s90, s90, 1"),              # one prefix one-bit: 2^k zeros
    ("s_bitcmp1_b64 s[94:95], 62", "s_bitcmp1_b64 s[94:95], 61"),    # a second one-bit: general path
    ("s_lshl_b64 s[94:95], s[94:95], 2", "s_lshl_b64 s[94:95], s[94:95], 1"),  # run past the chunk: k + 1 bits
])
def test_planted_loop_mistakes_are_caught(oracle, old, new):
    """A mistake planted in the loop text (on the CPU) fails the check."""
    lines = R.extract_loop()
    i = next(j for j, s in enumerate(lines) if s.strip() == old)
    lines = lines[:i] + [new] + lines[i + 1:]
    fails, _, _ = _check(oracle, lines, streams_per_kind=3, states_per_stream=60)
    assert fails, f"planted '{new}' went unnoticed"


def test_loop_wait_states():
    """Every v_readlane / v_writelane lane select in the loop is written at
    least 4 wait states earlier on every control-flow path, and the loop holds
    no memory instruction (static check of the source text)."""
    lines = R.extract_loop()
    assert R.hazard_report(lines) == []
    d = R.lane_select_distances(lines)
    assert d and min(d.values()) >= R.LANE_SELECT_WAIT
    print("least wait states before each lane-select use:", sorted(set(d.values())))


def test_wait_state_check_catches_planted_violations():
    """The checker fails loops whose lane select is written too close to its
    use: s88 recomputed two instructions before a refill's v_readlane, the
    v_readlane moved up to right after the s_and that computes its lane, or an
    m0 update right before a v_writelane."""
    lines = R.extract_loop()
    i = lines.index("v_readlane_b32 s96, %[cur], s88")
    close = lines[:i - 1] + ["s_and_b32 s88, s88, 63"] + lines[i - 1:]
    assert any("lane select s88" in b for b in R.hazard_report(close))
    # the refill computes its lane (s_and) four instructions before the read
    j = max(k for k in range(i) if lines[k] == "s_and_b32 s88, %[wpos], 63")
    early = lines[:j + 1] + [lines[i]] + lines[j + 1:i] + lines[i + 1:]
    assert any("lane select s88" in b for b in R.hazard_report(early))
    # an m0 write right before a v_writelane
    j = lines.index("v_writelane_b32 %[aK], s85, m0")
    moved = lines[:j] + ["s_add_u32 m0, m0, 0"] + lines[j:]
    assert any("lane select m0" in b for b in R.hazard_report(moved))

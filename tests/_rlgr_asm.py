"""CPU emulation of the decoder's hand-written scalar RLGR loop.

The loop (`rlgr_fast_run` in patchvvc-plus_amd/csrc/pvvc_decode.hip) is GCN
scalar assembly inside an inline-asm string.  This module reads that string
from the source, executes it instruction by instruction with the ISA's
semantics for the subset it uses (SCC rules included), and states the RLGR
decoder independently (entropy_codec.cc:207-283: adaptive Golomb-Rice with
run mode, L = 4, k_P from 0, k_RP from 8), so that a test can check every
state the loop leaves against the decoder's definition.

Register model: the operands %[win] (64-bit), %[nn], %[wpos], %[kRP], %[kP],
%[neg] (32-bit scalars); %[aT], %[aK], %[aF] (per-lane accumulators,
read/write), %[cur] (per-lane input), %[pre] (its dword wpos); s80-s97 and
m0 scratch.  Symbol i of a
chunk of lim symbols sits in accumulator lane (i - lim) mod 64, in the form
lane_symbol() decodes.
"""
from __future__ import annotations

import re
from pathlib import Path

M32 = (1 << 32) - 1
M64 = (1 << 64) - 1
# The loop keeps nn = NN_BIAS - nb (mod 2^32) for a window of nb valid bits
# and is entered with NB_MIN <= nb <= NB_MAX.
NN_BIAS, NB_MIN, NB_MAX = 32, 33, 64
HIP_SOURCE = Path(__file__).resolve().parents[1] / "patchvvc-plus_amd" / "csrc" / "pvvc_decode.hip"


def extract_loop(path: Path = HIP_SOURCE) -> list[str]:
    """Instruction strings of the asm block in rlgr_fast_run, labels included."""
    text = Path(path).read_text()
    f = text.index("void rlgr_fast_run(")
    a = text.index("asm volatile(", f)
    b = text.index('\n        : [win]', a)
    return re.findall(r'"([^"\n]*)\\n\\t"', text[a:b])


class Loop:
    def __init__(self, lines: list[str]):
        self.ins = []
        self.labels: dict[str, list[int]] = {}
        for s in lines:
            m = re.fullmatch(r"(\d+):", s.strip())
            if m:
                self.labels.setdefault(m.group(1), []).append(len(self.ins))
                continue
            op, _, rest = s.strip().partition(" ")
            args = [x.strip() for x in rest.split(",")] if rest else []
            self.ins.append((op, args))

    def target(self, pc: int, ref: str) -> int:
        n, d = ref[:-1], ref[-1]
        pos = self.labels[n]
        if d == "f":
            return min(p for p in pos if p > pc)
        return max(p for p in pos if p <= pc)

    # ------------------------------------------------------------------
    def run(self, st: dict, max_steps: int = 200000, visited: set | None = None) -> tuple[dict, int]:
        s = {k: (list(v) if isinstance(v, list) else v) for k, v in st.items()}
        sg = {}  # scratch SGPRs
        scc = 0
        vcc = 0

        def rd(x):
            if x.startswith("%["):
                return s[x[2:-1]]
            if x == "vcc":
                return vcc
            if x == "m0":
                return sg.get("m0", 0)
            m = re.fullmatch(r"s\[(\d+):(\d+)\]", x)
            if m:
                lo, hi = int(m.group(1)), int(m.group(2))
                assert hi == lo + 1
                return (sg.get(lo, 0) & M32) | ((sg.get(hi, 0) & M32) << 32)
            m = re.fullmatch(r"s(\d+)", x)
            if m:
                return sg.get(int(m.group(1)), 0)
            return int(x, 0) & M32 if not x.startswith("-") else (int(x) & M32)

        def wr(x, v, bits=32):
            if x == "m0":
                sg["m0"] = v & M32
                return
            if x.startswith("%["):
                name = x[2:-1]
                s[name] = v & (M64 if name == "win" else M32)
                return
            m = re.fullmatch(r"s\[(\d+):(\d+)\]", x)
            if m:
                lo = int(m.group(1))
                sg[lo] = v & M32
                sg[lo + 1] = (v >> 32) & M32
                return
            m = re.fullmatch(r"s(\d+)", x)
            assert m, x
            sg[int(m.group(1))] = v & M32

        def sx(v):
            v &= M32
            return v - (1 << 32) if v >> 31 else v

        pc, steps = 0, 0
        while pc < len(self.ins):
            steps += 1
            if steps > max_steps:
                raise RuntimeError("no exit within the step bound")
            op, a = self.ins[pc]
            if visited is not None:
                visited.add(pc)
            nxt = pc + 1
            if op == "s_branch":
                nxt = self.target(pc, a[0])
            elif op in ("s_cbranch_scc0", "s_cbranch_scc1"):
                if scc == (1 if op.endswith("1") else 0):
                    nxt = self.target(pc, a[0])
            elif op == "s_nop":
                pass
            elif op in ("s_mov_b32", "s_mov_b64"):
                wr(a[0], rd(a[1]))
            elif op == "s_not_b64":
                v = ~rd(a[1]) & M64
                wr(a[0], v)
                scc = int(v != 0)
            elif op == "s_flbit_i32_b64":
                v = rd(a[1]) & M64
                wr(a[0], (64 - v.bit_length()) if v else M32)
            elif op == "s_add_u32":
                v = rd(a[1]) + rd(a[2])
                wr(a[0], v)
                scc = int(v > M32)
            elif op == "s_addc_u32":
                v = rd(a[1]) + rd(a[2]) + scc
                wr(a[0], v)
                scc = int(v > M32)
            elif op == "s_sub_u32":
                x, y = rd(a[1]), rd(a[2])
                wr(a[0], x - y)
                scc = int(x < y)
            elif op in ("s_add_i32", "s_sub_i32"):
                x, y = sx(rd(a[1])), sx(rd(a[2]))
                v = x + y if op == "s_add_i32" else x - y
                wr(a[0], v)
                scc = int(not (-(1 << 31) <= v < (1 << 31)))
            elif op == "s_max_i32":
                x, y = sx(rd(a[1])), sx(rd(a[2]))
                wr(a[0], x if x > y else y)
                scc = int(x > y)
            elif op == "s_min_i32":
                x, y = sx(rd(a[1])), sx(rd(a[2]))
                wr(a[0], x if x < y else y)
                scc = int(x < y)
            elif op in ("s_lshl_b32", "s_lshr_b32"):
                x, y = rd(a[1]) & M32, rd(a[2]) & 31
                v = ((x << y) if op == "s_lshl_b32" else (x >> y)) & M32
                wr(a[0], v)
                scc = int(v != 0)
            elif op in ("s_lshl_b64", "s_lshr_b64"):
                x, y = rd(a[1]) & M64, rd(a[2]) & 63
                v = ((x << y) if op == "s_lshl_b64" else (x >> y)) & M64
                wr(a[0], v)
                scc = int(v != 0)
            elif op in ("s_or_b32", "s_and_b32"):
                x, y = rd(a[1]) & M32, rd(a[2]) & M32
                v = (x | y) if op == "s_or_b32" else (x & y)
                wr(a[0], v)
                scc = int(v != 0)
            elif op == "s_or_b64":
                v = (rd(a[1]) | rd(a[2])) & M64
                wr(a[0], v)
                scc = int(v != 0)
            elif op == "s_orn2_b64":
                v = (rd(a[1]) | ~rd(a[2])) & M64
                wr(a[0], v)
                scc = int(v != 0)
            elif op == "s_ashr_i64":
                x, y = rd(a[1]) & M64, rd(a[2]) & 63
                v = ((x - (1 << 64) if x >> 63 else x) >> y) & M64
                wr(a[0], v)
                scc = int(v != 0)
            elif op == "s_bfe_u64":
                # field at bit S1[5:0], S1[22:16] bits wide
                f = rd(a[2])
                off, width = f & 63, (f >> 16) & 127
                v = (rd(a[1]) >> off) & ((1 << width) - 1) & M64
                wr(a[0], v)
                scc = int(v != 0)
            elif op == "s_bitcmp1_b64":
                scc = (rd(a[0]) >> (rd(a[1]) & 63)) & 1
            elif op.startswith("s_cmp_"):
                x, y = rd(a[0]) & M32, rd(a[1]) & M32
                scc = int({"s_cmp_gt_u32": x > y, "s_cmp_eq_u32": x == y, "s_cmp_lg_u32": x != y,
                           "s_cmp_lt_u32": x < y, "s_cmp_ge_u32": x >= y,
                           "s_cmp_ge_i32": sx(x) >= sx(y)}[op])
            elif op == "v_mov_b32":
                v = rd(a[1]) & M32
                s[a[0][2:-1]] = [v] * 64
            elif op == "v_cmp_eq_u32":
                assert a[0] == "vcc"
                x = rd(a[1]) & M32
                vec = s[a[2][2:-1]]
                vcc = sum(1 << l for l in range(64) if (vec[l] & M32) == x)
            elif op == "v_cndmask_b32":
                d, a0, a1 = s[a[1][2:-1]], s[a[2][2:-1]], a[3]
                assert a1 == "vcc"
                s[a[0][2:-1]] = [a0[l] if (vcc >> l) & 1 else d[l] for l in range(64)]
            elif op == "v_readlane_b32":
                wr(a[0], s[a[1][2:-1]][rd(a[2]) & 63])
            elif op == "v_writelane_b32":
                s[a[0][2:-1]][rd(a[2]) & 63] = rd(a[1]) & M32
            else:
                raise NotImplementedError(op)
            pc = nxt
        return s, steps


# ----------------------------------------------------------------------
# RLGR decoder (entropy_codec.cc:207-283), event by event over a bit list
# ----------------------------------------------------------------------
class Bits:
    def __init__(self, bits: list[int], pos: int = 0):
        self.b, self.pos = bits, pos

    def bit(self) -> int:
        v = self.b[self.pos]
        self.pos += 1
        return v

    def read(self, n: int) -> int:
        v = 0
        for _ in range(n):
            v = (v << 1) | self.bit()
        return v

    def gr(self, k: int) -> int:
        p = 0
        while p < 32 and self.b[self.pos] == 1:
            self.pos += 1
            p += 1
        if p == 32:
            return self.read(32)
        self.pos += 1  # the terminating 0
        return (p << k) + self.read(k)


def adapt_krp(krp: int, p: int) -> int:
    return min(krp + min(p, 200) - 1, 128) if p else max(krp - 2, 0)


def events(bits: Bits, kP: int, kRP: int):
    """Yields, per event, (symbols as (offset, raw value) list, symbols advanced,
    kP, kRP) after it.  Raw values are what the loop stores: u for k = 0
    symbols, u + 1 after a zero run; zeros of a run are not stored."""
    while True:
        k = kP >> 2
        kR = kRP >> 2
        if k == 0:
            u = bits.gr(kR)
            kRP = adapt_krp(kRP, u >> kR)
            kP = max(kP - 1, 0) if u else kP + 3
            yield [(0, u)], 1, kP, kRP
        else:
            m = 0
            while bits.bit() == 1:
                m += 1 << k
                kP += 2
                k = kP >> 2
            m += bits.read(k)
            u = bits.gr(kR)
            kRP = adapt_krp(kRP, u >> kR)
            kP = max(kP - 1, 0)
            yield [(m, u + 1)], m + 1, kP, kRP


def to_bits(data: bytes, pad_dwords: int = 80) -> list[int]:
    out = []
    for byte in bytes(data) + bytes(4 * pad_dwords):
        out.extend((byte >> (7 - i)) & 1 for i in range(8))
    return out


def decode_boundaries(bits: list[int], nsym: int):
    """(bit position, kP, kRP, symbol index) at every event boundary."""
    b = Bits(bits)
    kP, kRP, n = 0, 8, 0
    out = [(0, 0, 8, 0)]
    for _, adv, kP, kRP in events(b, kP, kRP):
        n += adv
        out.append((b.pos, kP, kRP, n))
        if n >= nsym:
            break
    return out


def make_state(bits: list[int], pos: int, kP: int, kRP: int, rng, nb: int, wpos: int, lim: int, nrel: int):
    """Loop entry state at bit `pos`: nb window bits, then the stream's dwords
    in lanes wpos & 63 .. 63 of `cur` (the chunk that holds dword wpos), the
    first of them also in `pre`."""
    win = 0
    for i in range(nb):
        win = (win << 1) | bits[pos + i]
    win <<= 64 - nb
    l0 = wpos & 63
    cur = [int(x) for x in rng.integers(0, 1 << 32, 64, dtype="uint64")]
    q = pos + nb
    for lane in range(l0, 64):
        v = 0
        for i in range(32):
            v = (v << 1) | bits[q + i]
        cur[lane] = v
        q += 32
    # lanes of the chunk's earlier symbols hold them in the direct form
    # (value, 32, 0); the others are clear (T = F = 0) with any K
    acc = [int(x) for x in rng.integers(0, 1 << 20, 64)]
    aT, aF = list(acc), [0] * 64
    aK = [32] * 64
    for i in range(nrel, 64):
        lane = (i - lim) & 63
        acc[lane] = aT[lane] = 0
        aK[lane] = (7 * lane + nrel) % 40
    return {
        "win": win, "nn": (NN_BIAS - nb) & M32, "wpos": wpos & M32, "kRP": kRP, "kP": kP,
        "neg": (nrel - lim) & M32, "aT": aT, "aK": aK, "aF": aF, "acc": acc, "cur": cur, "pre": cur[l0],
        "lim": lim,
    }


def lane_symbol(T: int, K: int, F: int) -> int:
    """The symbol one accumulator lane stands for (lane_symbol in
    pvvc_decode.hip): K < 32: T starts with a Golomb-Rice code of parameter K,
    u = (q << K) | r, plus F; K >= 32: T itself, plus F."""
    T &= M32
    inv = ~T & M32
    q = 32 - inv.bit_length() if inv else 32
    k = K & 31
    r = (T >> ((31 - q - k) & 31)) & ((1 << k) - 1)
    u = ((q << k) | r) & M32
    return ((T if (K & M32) >= 32 else u) + F) & M32


def lanes(st: dict) -> list[int]:
    return [lane_symbol(st["aT"][l], st["aK"][l], st["aF"][l]) for l in range(64)]


def run_prefix(b: Bits, kP: int) -> tuple[int, int]:
    """A run event's prefix and length field read at b (k_P >= 4): the zero
    count and k_P after the prefix's one-bits (entropy_codec.cc:219-236)."""
    k, m = kP >> 2, 0
    while b.bit() == 1:
        m += 1 << k
        kP += 2
        k = kP >> 2
    return m + b.read(k), kP


def check_pending_run(bits: list[int], pos: int, st0: dict, st1: dict, consumed: int) -> str | None:
    """An exit with bit 30 of neg set stops inside a run event whose zeros
    reach the chunk end: after whole events from st0, that run's prefix and
    length field are consumed, neg's low bits count its zeros past the chunk
    end, k_P is the value after the prefix, k_RP is unchanged and no lane is
    written for it (the caller decodes the run's symbol)."""
    lim = st0["lim"]
    kP, kRP, n = st0["kP"], st0["kRP"], (st0["neg"] + lim) & M32
    acc = list(st0["acc"])
    got = lanes(st1)
    at = pos
    while at - pos <= consumed:
        if kP >= 4:
            b2 = Bits(bits, at)
            m, kPa = run_prefix(b2, kP)
            if b2.pos - pos == consumed:
                nrel1 = lim + (st1["neg"] & 0x3FFFFFFF)
                if n + m < lim:
                    return f"pending run at index {n} + {m} zeros does not reach the chunk end {lim}"
                if nrel1 != n + m:
                    return f"pending run: symbol index {nrel1} != decoder {n + m}"
                if st1["kP"] != kPa or st1["kRP"] != kRP:
                    return f"pending run: kP/kRP {st1['kP']}/{st1['kRP']} != decoder {kPa}/{kRP}"
                if got != acc:
                    return "pending run: accumulator lanes differ"
                return None
        b3 = Bits(bits, at)
        syms, adv, kP, kRP = next(events(b3, kP, kRP))
        for off, val in syms:
            if n + off < 64:
                acc[(n + off - lim) & 63] = val & M32
        n += adv
        at = b3.pos
    return f"pend exit after {consumed} bits is not inside a run event's prefix"


def check_against_decoder(bits: list[int], pos: int, st0: dict, st1: dict) -> str | None:
    """None when the loop's exit state st1 equals the decoder's state after a
    whole number of events from st0 (or, with bit 30 of neg set, inside a run event
    whose zeros reach the chunk end: check_pending_run); otherwise a
    description."""
    def s32(v):
        return v - (1 << 32) if v >> 31 else v
    nb0, nb1 = NN_BIAS - s32(st0["nn"]), NN_BIAS - s32(st1["nn"])
    dw = (st1["wpos"] - st0["wpos"]) & M32
    consumed = nb0 + 32 * dw - nb1
    lim = st0["lim"]
    pend = (st1["neg"] >> 30) == 1
    nrel0 = (st0["neg"] + lim) & M32
    nrel1 = lim + (st1["neg"] & 0x3FFFFFFF) if pend else (st1["neg"] + lim) & M32
    if not (0 <= nrel1 <= lim or (pend and nrel1 < lim + (1 << 16))) or nrel1 < nrel0 or consumed < 0:
        return f"bad exit counters nrel {nrel0}->{nrel1} lim {lim} consumed {consumed}"
    # window contents: nb1 bits from pos + consumed, left aligned, zero below
    want = 0
    for i in range(nb1):
        want = (want << 1) | bits[pos + consumed + i]
    want = (want << (64 - nb1)) & M64 if nb1 > 0 else 0
    if st1["win"] != want:
        return f"window {st1['win']:016x} != {want:016x}"
    if pend:
        return check_pending_run(bits, pos, st0, st1, consumed)
    b = Bits(bits, pos)
    kP, kRP, n = st0["kP"], st0["kRP"], nrel0
    acc = list(st0["acc"])
    got = lanes(st1)
    if consumed == 0:
        ok = (nrel1 == nrel0 and st1["kP"] == kP and st1["kRP"] == kRP and got == acc)
        return None if ok else "state changed without consuming bits"
    for syms, adv, kP, kRP in events(b, kP, kRP):
        for off, val in syms:
            if n + off < 64:
                acc[(n + off - lim) & 63] = val & M32
        n += adv
        used = b.pos - pos
        if used == consumed:
            if n != nrel1:
                return f"symbol index {nrel1} != decoder {n}"
            if st1["kP"] != kP or st1["kRP"] != kRP:
                return f"kP/kRP {st1['kP']}/{st1['kRP']} != decoder {kP}/{kRP}"
            if got != acc:
                bad = [l for l in range(64) if got[l] != acc[l]]
                return f"accumulator lanes differ: {bad[:8]}"
            return None
        if used > consumed:
            return f"loop consumed {consumed} bits, not at an event boundary (decoder at {used})"


# ----------------------------------------------------------------------
# Static wait-state check of the loop text
# ----------------------------------------------------------------------
# Rules the loop relies on (AMD Instinct MI300 / CDNA3 ISA, "manually
# inserted wait states"; gfx950 keeps them):
#   * a VALU instruction that writes an SGPR (v_readlane) followed by
#     v_readlane / v_writelane using that SGPR as the lane select: 4 wait
#     states;
#   * the loop applies the same 4 wait states when the lane select was written
#     by a scalar instruction (m0 for v_writelane, s88 for v_readlane);
#   * a VALU-written SGPR read by a vector memory instruction: 5 (the loop has
#     no memory instructions, which the check also asserts).
# A scalar instruction reading an SGPR that v_readlane wrote is interlocked by
# the hardware and needs none.  One instruction is one wait state; s_nop N is
# N + 1.
LANE_SELECT_WAIT = 4


def _regs(x: str) -> set[str]:
    m = re.fullmatch(r"s\[(\d+):(\d+)\]", x)
    if m:
        return {f"s{i}" for i in range(int(m.group(1)), int(m.group(2)) + 1)}
    if re.fullmatch(r"s\d+", x) or x == "m0" or x.startswith("%["):
        return {x}
    return set()


def _writes(op: str, a: list[str]) -> tuple[set[str], str]:
    """SGPRs an instruction writes and the writer kind ("salu" / "valu")."""
    if op == "v_readlane_b32":
        return _regs(a[0]), "valu"
    if op.startswith("s_") and a and not op.startswith(("s_cmp", "s_bitcmp", "s_cbranch", "s_branch", "s_nop")):
        return _regs(a[0]), "salu"
    return set(), ""


def _lane_select(op: str, a: list[str]) -> set[str]:
    return _regs(a[2]) if op in ("v_readlane_b32", "v_writelane_b32") else set()


def _wait_states(op: str, a: list[str]) -> int:
    return int(a[0], 0) + 1 if op == "s_nop" else 1


def _predecessors(loop: "Loop") -> list[list[int]]:
    pred = [[] for _ in loop.ins]
    for i, (op, a) in enumerate(loop.ins):
        if i + 1 < len(loop.ins) and op != "s_branch":
            pred[i + 1].append(i)
        if op in ("s_branch", "s_cbranch_scc0", "s_cbranch_scc1"):
            t = loop.target(i, a[0])
            if t < len(loop.ins):
                pred[t].append(i)
    return pred


def hazard_report(lines: list[str]) -> list[str]:
    """Every lane-select use whose last writer, on some path through the
    loop's control flow, lies fewer than LANE_SELECT_WAIT wait states before
    it; plus any memory instruction.  Empty when the loop keeps the rules."""
    loop = Loop(lines)
    pred = _predecessors(loop)
    bad = []
    for u, (op, a) in enumerate(loop.ins):
        if op.startswith(("buffer_", "global_", "flat_", "s_load", "s_store", "s_buffer", "ds_", "s_atomic")):
            bad.append(f"{u}: memory instruction {op}")
        for r in _lane_select(op, a):
            # backward walk: (index of the instruction before the use, wait states so far)
            stack, seen = [(p, 0) for p in pred[u]], set()
            if not pred[u]:
                bad.append(f"{u}: {op} lane select {r} reaches the loop entry unwritten")
            while stack:
                i, d = stack.pop()
                if (i, d) in seen:
                    continue
                seen.add((i, d))
                w, kind = _writes(*loop.ins[i])
                if r in w:
                    if d < LANE_SELECT_WAIT:
                        bad.append(f"{u}: {op} reads lane select {r} {d} wait states after the {kind} write at {i}")
                    continue
                d2 = d + _wait_states(*loop.ins[i])
                if d2 >= LANE_SELECT_WAIT:
                    continue
                if not pred[i]:
                    bad.append(f"{u}: {op} lane select {r} may come from before the loop ({d2} wait states)")
                stack.extend((p, d2) for p in pred[i])
    return bad


def lane_select_distances(lines: list[str]) -> dict[int, int]:
    """Least wait states, over all paths, between each lane-select use and its
    writer (for the record in the test output)."""
    loop = Loop(lines)
    pred = _predecessors(loop)
    out = {}
    for u, (op, a) in enumerate(loop.ins):
        for r in _lane_select(op, a):
            best, stack, seen = 1 << 30, [(p, 0) for p in pred[u]], set()
            while stack:
                i, d = stack.pop()
                if (i, d) in seen or d >= best:
                    continue
                seen.add((i, d))
                if r in _writes(*loop.ins[i])[0]:
                    best = min(best, d)
                    continue
                stack.extend((p, d + _wait_states(*loop.ins[i])) for p in pred[i])
            out[u] = best
    return out

"""Writes wnlqg_eigd_lds32.h: the LDS word of every element of the n = 32 slot grid in k_eigd_sweeps.

The sweep kernel keeps each point's H (upper triangle, slot positions x <= y) in one LDS buffer.  Every round,
lane l of a point reads the elements of its 2x2 block groups in 16 instructions (B[i][e]) and the slot cs = l
diagonal block in 3 more (lanes 0..15), then writes each updated element to the word of its position after the
round, (sigma(x), sigma(y)).  Lanes 28..31 hold two blocks; in the instructions of the other two they read
and write spare words 560..575, and lanes 16..31 repeat the diagonal-block reads of lanes 0..15.  On gfx950 a
16-byte LDS access of a wave is served

  * for reads, 16 lanes per cycle: the lanes of one half wave with equal parity of lane bits 2, 3 and 4,
    from 64 banks (256 bytes): conflict-free when their words differ mod 16;
  * for writes, 8 consecutive lanes per cycle from 32 banks (128 bytes): conflict-free when their words differ
    mod 8

(measured with SQ_LDS_BANK_CONFLICT on single-lane-pair access patterns).  The row-major group layout meets the
read rule but not the write rule (126 instead of 64 write cycles per point and round); this script assigns
each element a bank class c (word = c + 16 k) meeting both rules for every instruction of a round, by a seeded
local search over swaps of classes inside each read group, and places the classes' elements at words
c, c + 16, c + 32, ... below 560 (the buffer size of the row-major layout).  The spare word a lane of 28..31
uses in such an instruction is 560 + c with c a class no other lane of its read or write cycle uses there
(a read cycle's 12 held blocks leave 4 of the 16 classes, a write cycle's 4 leave 4 of the 8).  The diagonal
reads of a and d
load 8 bytes (ds_read_b64); all diagonal reads (a, d and h of the 16 slots) are required to differ mod 16
as one group, which meets both read rules whichever lanes of the two halves are served together.

Run: python3 gen_eigd_lds32.py  (deterministic; rewrites wnlqg_eigd_lds32.h beside this file)
"""
import collections
import math
import os
import random

N = 32
HALF = N // 2
G = HALF // 2
NGF = G * (G - 1) // 2
WORDS = 560   # words per point (16 bytes each)
CAP = WORDS // 16


def sigma(x):
    """Slot position after a round of the round-robin order (WnSlots<N>::sigma)."""
    if x == 0:
        return 0
    if x & 1:
        return x + 2 if (x >> 1) < HALF - 1 else x - 1
    return 1 if (x >> 1) == 1 else x - 2


def lane_blocks(l):
    """Row slots, column slots and active blocks of lane l (k_eigd_sweeps' assignment)."""
    rsl, csl, bon = [0, 0], [0, 0], [False] * 4
    if l < NGF:
        a, rem = 0, l
        while rem >= G - 1 - a:
            rem -= G - 1 - a
            a += 1
        b = a + 1 + rem
        rsl, csl, bon = [2 * a, 2 * a + 1], [2 * b, 2 * b + 1], [True] * 4
    else:
        q0 = 2 * (l - NGF)
        q1 = q0 + 1
        if q0 < G:
            rsl[0], csl[0], bon[0] = 2 * q0, 2 * q0 + 1, True
        if q1 < G:
            rsl[1], csl[1], bon[3] = 2 * q1, 2 * q1 + 1, True
    return rsl, csl, bon


def read_group(l):
    return ((l >> 2) ^ (l >> 3) ^ (l >> 4)) & 1


def norm(x, y):
    return (x, y) if x <= y else (y, x)


def access_sets():
    """(read groups, write groups): instruction x lane group -> elements, for one point's 32 lanes (blocks
    the lanes hold; the spare words are chosen afterwards)."""
    reads, writes = collections.defaultdict(list), collections.defaultdict(list)
    for l in range(N):
        rsl, csl, bon = lane_blocks(l)
        for i in range(4):
            if not bon[i]:
                continue
            for e in range(4):
                x, y = 2 * rsl[i >> 1] + (e >> 1), 2 * csl[i & 1] + (e & 1)
                reads[(("B", i, e), read_group(l))].append((x, y))
                writes[(("B", i, e), l >> 3)].append(norm(sigma(x), sigma(y)))
    for cs in range(HALF):
        for nm, el in (("a", (2 * cs, 2 * cs)), ("d", (2 * cs + 1, 2 * cs + 1)), ("h", (2 * cs, 2 * cs + 1))):
            reads[(("D", nm), 0)].append(el)
        writes[(("D", "a"), cs >> 3)].append(norm(sigma(2 * cs), sigma(2 * cs)))
        writes[(("D", "d"), cs >> 3)].append(norm(sigma(2 * cs + 1), sigma(2 * cs + 1)))
        writes[(("D", "h"), cs >> 3)].append(norm(sigma(2 * cs), sigma(2 * cs + 1)))
    return reads, writes


def solve(seed):
    reads, writes = access_sets()
    wg_of = {e: [k] for k, v in writes.items() for e in v}
    rnd = random.Random(seed)
    cls = {}
    for k in sorted(reads):
        v = reads[k]
        for e, c in zip(v, rnd.sample(range(16), len(v))):
            cls[e] = c
    count = collections.Counter(cls.values())

    def wcost(k):
        m = collections.Counter(cls[e] % 8 for e in writes[k])
        return sum(c * (c - 1) // 2 for c in m.values())

    def ccost():
        return sum(max(0, c - CAP) for c in count.values())

    cost = sum(wcost(k) for k in writes) + 4 * ccost()
    rkeys = sorted(reads)
    temp = 1.0
    for _ in range(3000000):
        if cost == 0:
            break
        v = reads[rnd.choice(rkeys)]
        e1 = rnd.choice(v)
        used = {cls[e] for e in v}
        free = [c for c in range(16) if c not in used]
        if free and rnd.random() < 0.3:
            old, new = cls[e1], rnd.choice(free)
            before = sum(wcost(g) for g in wg_of[e1]) + 4 * ccost()
            cls[e1] = new
            count[old] -= 1
            count[new] += 1
            d = sum(wcost(g) for g in wg_of[e1]) + 4 * ccost() - before
            if d <= 0 or rnd.random() < math.exp(-d / temp):
                cost += d
            else:
                cls[e1] = old
                count[old] += 1
                count[new] -= 1
        else:
            e2 = rnd.choice(v)
            if e2 == e1:
                continue
            gs = set(wg_of[e1]) | set(wg_of[e2])
            before = sum(wcost(g) for g in gs)
            cls[e1], cls[e2] = cls[e2], cls[e1]
            d = sum(wcost(g) for g in gs) - before
            if d <= 0 or rnd.random() < math.exp(-d / temp):
                cost += d
            else:
                cls[e1], cls[e2] = cls[e2], cls[e1]
        temp = max(0.05, temp * 0.99999)
    return cls if cost == 0 else None


def words_of(cls):
    word, nxt = {}, collections.Counter()
    for e in sorted(cls):
        c = cls[e]
        word[e] = c + 16 * nxt[c]
        nxt[c] += 1
    assert max(word.values()) < WORDS
    return word


def spare_words(word):
    """[lane 28..31][block 1, 2][e] -> (read word, write word) of the spare area for the blocks lanes 28..31
    do not hold: classes no other lane of the read cycle (mod 16) or write cycle (mod 8) uses."""
    reads, writes = access_sets()
    out = [[[None] * 4 for _ in range(2)] for _ in range(4)]
    for i in (1, 2):
        for e in range(4):
            used_r = {word[x] % 16 for x in reads[(("B", i, e), read_group(28))]}
            used_w = {word[x] % 8 for x in writes[(("B", i, e), 3)]}
            free_r = [c for c in range(16) if c not in used_r]
            free_w = [c for c in range(8) if c not in used_w]
            assert len(free_r) >= 4 and len(free_w) >= 4
            for m in range(4):
                out[m][i - 1][e] = (WORDS + free_r[m], WORDS + free_w[m])
    return out


def main():
    for seed in range(1, 50):
        cls = solve(seed)
        if cls is not None:
            break
    else:
        raise SystemExit("no conflict-free layout found")
    word = words_of(cls)
    packed = []
    for x in range(N):
        for y in range(x, N):
            packed.append(word[(x, y)])
    lines = [
        "// Generated by gen_eigd_lds32.py (seed %d): LDS word (16 bytes each) of element (x, y), x <= y, of the" % seed,
        "// n = 32 slot grid in k_eigd_sweeps, at index x * 32 - x * (x - 1) / 2 + (y - x).  Every 16-byte read and",
        "// write of a sweep round is free of LDS bank conflicts on gfx950 with this placement.",
        "#pragma once",
        "static __constant__ unsigned short wn_eigd32_word[%d] = {" % len(packed),
    ]
    for i in range(0, len(packed), 16):
        lines.append("    " + ", ".join("%d" % w for w in packed[i:i + 16]) + ",")
    lines.append("};")
    spare = spare_words(word)
    lines.append("// spare words (560 + class) read and written by lane 28 + m in the instructions of blocks 1 and 2, which")
    lines.append("// it does not hold: [m][block - 1][element] = read word | write word << 16")
    lines.append("static __constant__ unsigned int wn_eigd32_spare[4][2][4] = {")
    for m in range(4):
        rows = []
        for i in range(2):
            rows.append("{" + ", ".join("%du" % (r | (w << 16)) for r, w in spare[m][i]) + "}")
        lines.append("    {" + ", ".join(rows) + "},")
    lines.append("};")
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "wnlqg_eigd_lds32.h")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("wrote", path, "seed", seed)


if __name__ == "__main__":
    main()

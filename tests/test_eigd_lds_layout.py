"""LDS placement of the n = 32 sweep kernel's H (wnlqg_eigd_lds32.h, written by gen_eigd_lds32.py).

The table must place the 528 elements of the slot grid's upper triangle on distinct words of the 560-word
buffer, and every read and write instruction of a round must be free of bank conflicts under gfx950's rules
for 16-byte LDS accesses (reads: 16 lanes per cycle, words distinct mod 16; writes: 8 consecutive lanes per
cycle, words distinct mod 8).  The committed header must be what the generator writes.
"""
import collections
import importlib.util
import os
import re
import subprocess
import sys

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "wanniernlqg_amd", "csrc")


def _gen():
    spec = importlib.util.spec_from_file_location("gen_eigd_lds32", os.path.join(CSRC, "gen_eigd_lds32.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _table():
    text = open(os.path.join(CSRC, "wnlqg_eigd_lds32.h")).read()
    body = text[text.index("{") + 1:text.index("}")]
    return [int(v) for v in re.findall(r"\d+", body)]


def _spare():
    """[lane 28..31][block 1, 2][element] -> (read word, write word)"""
    text = open(os.path.join(CSRC, "wnlqg_eigd_lds32.h")).read()
    body = text[text.index("wn_eigd32_spare"):]
    vals = [int(v) for v in re.findall(r"(\d+)u", body)]
    assert len(vals) == 32
    return [[[(vals[m * 8 + i * 4 + e] & 0xffff, vals[m * 8 + i * 4 + e] >> 16) for e in range(4)] for i in range(2)]
            for m in range(4)]


def test_table_is_a_placement_of_the_upper_triangle():
    words = _table()
    assert len(words) == 32 * 33 // 2
    assert len(set(words)) == len(words)
    assert min(words) >= 0 and max(words) < 560


def test_rounds_are_free_of_bank_conflicts():
    gen = _gen()
    words = _table()

    def word(x, y):
        return words[x * 32 - x * (x - 1) // 2 + (y - x)]

    reads, writes = gen.access_sets()
    # every element is read once and written once per round
    rc = collections.Counter(e for v in reads.values() for e in v)
    wc = collections.Counter(e for v in writes.values() for e in v)
    assert set(rc) == set(wc) and len(rc) == 528
    assert all(c == 1 for c in rc.values()) and all(c == 1 for c in wc.values())
    rwords = {k: [word(*e) for e in v] for k, v in reads.items()}
    wwords = {k: [word(*e) for e in v] for k, v in writes.items()}
    # lanes 28..31 (read cycle of parity 1, write cycle 3) use spare words in the instructions of blocks 1, 2
    spare = _spare()
    assert gen.read_group(28) == 1 and all(gen.read_group(l) == 1 for l in range(28, 32))
    for m in range(4):
        for i in (1, 2):
            for e in range(4):
                rw, ww = spare[m][i - 1][e]
                assert 560 <= rw < 576 and 560 <= ww < 576
                rwords[(("B", i, e), 1)].append(rw)
                wwords[(("B", i, e), 3)].append(ww)
    for key, ws in rwords.items():
        assert len({w % 16 for w in ws}) == len(ws), key
    for key, ws in wwords.items():
        assert len({w % 8 for w in ws}) == len(ws), key
    assert len(rwords[(("B", 1, 0), 1)]) == 16 and len(wwords[(("B", 2, 3), 3)]) == 8


def test_round_robin_permutation_of_the_generator():
    """sigma is the round-robin order's move (WnSlots::sigma): a permutation fixing slot 0 whose powers visit
    every pair of bands once in N - 1 rounds."""
    gen = _gen()
    n = gen.N
    perm = [gen.sigma(x) for x in range(n)]
    assert sorted(perm) == list(range(n)) and perm[0] == 0
    pos0 = [2 * b if b < n // 2 else 2 * (n - 1 - b) + 1 for b in range(n)]
    band_at = [0] * n
    for b in range(n):
        band_at[pos0[b]] = b
    seen = set()
    for _ in range(n - 1):
        for t in range(n // 2):
            seen.add(frozenset((band_at[2 * t], band_at[2 * t + 1])))
        nxt = [0] * n
        for x in range(n):
            nxt[gen.sigma(x)] = band_at[x]
        band_at = nxt
    assert len(seen) == n * (n - 1) // 2


def test_header_is_the_generator_output(tmp_path):
    # the generator writes beside itself: run a copy in a scratch folder and compare
    src = open(os.path.join(CSRC, "gen_eigd_lds32.py")).read()
    (tmp_path / "gen_eigd_lds32.py").write_text(src)
    subprocess.run([sys.executable, str(tmp_path / "gen_eigd_lds32.py")], check=True, capture_output=True,
                   timeout=1200)
    assert (tmp_path / "wnlqg_eigd_lds32.h").read_text() == open(os.path.join(CSRC, "wnlqg_eigd_lds32.h")).read()

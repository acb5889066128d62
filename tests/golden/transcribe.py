"""Transcribe the reference's golden BFS layer-count vectors into
tests/golden/reference_vectors.json, recording file:line for every vector.

Run in a tree where the reference is readable (default /root/reference):

    python tests/golden/transcribe.py [REFERENCE_ROOT]

Only numbers are copied (the layer counts and the start states the reference
tests pass); no reference source text is stored.  The output file is committed;
the tests read only the JSON.
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

# (id, file, the test name whose EXPECT line holds the vector or a variable
#  name, puzzle description, start state)
SPECS = [
    # SlidingPuzzleTests/TestFrontierSearch.cpp: TestSearch<w,h>(radius, "counts")
    dict(id="sliding_2x2", file="SlidingPuzzleTests/TestFrontierSearch.cpp", test="Search2x2",
         puzzle=dict(kind="sliding", width=2, height=2), start="0 1 2 3"),
    dict(id="sliding_3x2", file="SlidingPuzzleTests/TestFrontierSearch.cpp", test="Search3x2",
         puzzle=dict(kind="sliding", width=3, height=2), start="0 1 2 3 4 5"),
    dict(id="sliding_3x2_edge", file="SlidingPuzzleTests/TestFrontierSearch.cpp", test="Search3x2_edge",
         puzzle=dict(kind="sliding", width=3, height=2), start="1 0 2 3 4 5"),
    dict(id="sliding_4x2", file="SlidingPuzzleTests/TestFrontierSearch.cpp", test="Search4x2",
         puzzle=dict(kind="sliding", width=4, height=2), start="0 1 2 3 4 5 6 7"),
    dict(id="sliding_4x2_edge", file="SlidingPuzzleTests/TestFrontierSearch.cpp", test="Search4x2_edge",
         puzzle=dict(kind="sliding", width=4, height=2), start="1 0 2 3 4 5 6 7"),
    dict(id="sliding_3x3", file="SlidingPuzzleTests/TestFrontierSearch.cpp", test="Search3x3",
         puzzle=dict(kind="sliding", width=3, height=3), start="0 1 2 3 4 5 6 7 8"),
    dict(id="sliding_3x3_edge", file="SlidingPuzzleTests/TestFrontierSearch.cpp", test="Search3x3_edge",
         puzzle=dict(kind="sliding", width=3, height=3), start="1 0 2 3 4 5 6 7 8"),
    dict(id="sliding_3x3_center", file="SlidingPuzzleTests/TestFrontierSearch.cpp", test="Search3x3_center",
         puzzle=dict(kind="sliding", width=3, height=3), start="1 2 3 4 0 5 6 7 8"),
    dict(id="sliding_5x2", file="SlidingPuzzleTests/TestFrontierSearch.cpp", test="Search5x2",
         puzzle=dict(kind="sliding", width=5, height=2), start="0 1 2 3 4 5 6 7 8 9"),
    dict(id="sliding_6x2", file="SlidingPuzzleTests/TestFrontierSearch.cpp", test="Search6x2",
         puzzle=dict(kind="sliding", width=6, height=2), start="0 1 2 3 4 5 6 7 8 9 10 11"),
    dict(id="sliding_4x3", file="SlidingPuzzleTests/TestFrontierSearch.cpp", test="Search4x3",
         puzzle=dict(kind="sliding", width=4, height=3), start="0 1 2 3 4 5 6 7 8 9 10 11"),
    dict(id="sliding_4x4_prefix", file="SlidingPuzzleTests/TestFrontierSearch.cpp", test="Search4x4",
         puzzle=dict(kind="sliding", width=4, height=4), start="0 1 2 3 4 5 6 7 8 9 10 11 12 13 14 15",
         partial=True),
    # PuzzlesTests/PuzzleSearchTests.cpp
    dict(id="pancake_5", file="PuzzlesTests/PuzzleSearchTests.cpp", test="Opt3BitBFS_5",
         puzzle=dict(kind="pancake", n=5), start="0 1 2 3 4"),
    dict(id="pancake_9", file="PuzzlesTests/PuzzleSearchTests.cpp", test="Opt3BitBFS_9",
         puzzle=dict(kind="pancake", n=9), start="0 1 2 3 4 5 6 7 8"),
    dict(id="pancake_13_prefix", file="PuzzlesTests/PuzzleSearchTests.cpp", test="Opt3BitBFS_13_mt",
         puzzle=dict(kind="pancake", n=13), start="0 1 2 3 4 5 6 7 8 9 10 11 12", partial=True),
    dict(id="hanoi_10", file="PuzzlesTests/PuzzleSearchTests.cpp", test="ClassicBFS_10",
         puzzle=dict(kind="hanoi", disks=10, symmetry=False), start="index:0"),
    dict(id="hanoi_10_sym", file="PuzzlesTests/PuzzleSearchTests.cpp", test="ClassicBFS_10_Sym",
         puzzle=dict(kind="hanoi", disks=10, symmetry=True), start="index:0"),
    dict(id="hanoi_21_prefix", file="PuzzlesTests/PuzzleSearchTests.cpp", test="Opt3BitBFS_21",
         puzzle=dict(kind="hanoi", disks=21, symmetry=False), start="index:0", partial=True),
    dict(id="hanoi_21_sym_prefix", file="PuzzlesTests/PuzzleSearchTests.cpp", test="Opt3BitBFS_21_Sym",
         puzzle=dict(kind="hanoi", disks=21, symmetry=True), start="index:0", partial=True),
    # HanoiTowersTests/HanoiSearchTests.cpp: the first 121 layers (maxSteps 120)
    # of the 4-peg search with 17..21 disks
    dict(id="hanoi_17_120", file="HanoiTowersTests/HanoiSearchTests.cpp", test="Search17",
         puzzle=dict(kind="hanoi", disks=17, symmetry=False), start="index:0", partial=True),
    dict(id="hanoi_18_120", file="HanoiTowersTests/HanoiSearchTests.cpp", test="Search18",
         puzzle=dict(kind="hanoi", disks=18, symmetry=False), start="index:0", partial=True),
    dict(id="hanoi_19_120", file="HanoiTowersTests/HanoiSearchTests.cpp", test="Search19",
         puzzle=dict(kind="hanoi", disks=19, symmetry=False), start="index:0", partial=True),
    dict(id="hanoi_20_120", file="HanoiTowersTests/HanoiSearchTests.cpp", test="Search20",
         puzzle=dict(kind="hanoi", disks=20, symmetry=False), start="index:0", partial=True),
    dict(id="hanoi_21_120", file="HanoiTowersTests/HanoiSearchTests.cpp", test="Search21",
         puzzle=dict(kind="hanoi", disks=21, symmetry=False), start="index:0", partial=True),
    dict(id="hanoi_22_120", file="HanoiTowersTests/HanoiSearchTests.cpp", test="Search22",
         puzzle=dict(kind="hanoi", disks=22, symmetry=False), start="index:0", partial=True),
    dict(id="hanoi_23_120", file="HanoiTowersTests/HanoiSearchTests.cpp", test="Search23",
         puzzle=dict(kind="hanoi", disks=23, symmetry=False), start="index:0", partial=True),
    dict(id="hanoi_24_120", file="HanoiTowersTests/HanoiSearchTests.cpp", test="Search24",
         puzzle=dict(kind="hanoi", disks=24, symmetry=False), start="index:0", partial=True),
    dict(id="sliding_5x3_prefix", file="PuzzlesTests/PuzzleSearchTests.cpp", test="OptFrontierSearch_5x3_1th",
         puzzle=dict(kind="sliding", width=5, height=3), start="0 1 2 3 4 5 6 7 8 9 10 11 12 13 14",
         partial=True),
]

# A BFS layer-count vector starts with the single start state ("1 ...");
# start-state strings in the same tests start with a tile or disk number.
NUMS = re.compile(r'"(1 (?:\d+ )*\d+)"')


def find_vector(path, test):
    """Return (line_no, counts) of the first quoted number list at or after
    the TEST(..., test) / TestSearch line for `test`."""
    with open(path, encoding="utf-8", errors="replace") as f:
        lines = f.readlines()
    start = None
    pat = re.compile(r"\b" + re.escape(test) + r"\)\s*\{")
    for i, line in enumerate(lines):
        if pat.search(line):
            start = i
            break
    if start is None:
        raise SystemExit(f"test {test} not found in {path}")
    for j in range(start, min(start + 20, len(lines))):
        m = NUMS.search(lines[j])
        if m:
            return j + 1, [int(x) for x in m.group(1).split()]
        m = re.search(r"EXPECT_EQ\(ToString\(result\),\s*(\w+)\)", lines[j])
        if m:
            var = m.group(1)
            for k, line in enumerate(lines):
                mm = re.match(r"static std::string " + var + r' = "([\d ]+)";', line)
                if mm:
                    return k + 1, [int(x) for x in mm.group(1).split()]
    raise SystemExit(f"no vector for {test} in {path}")


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    out = []
    for spec in SPECS:
        line, counts = find_vector(os.path.join(ref, spec["file"]), spec["test"])
        rec = dict(spec)
        rec["source"] = f"{spec['file']}:{line}"
        rec["counts"] = counts
        rec.setdefault("partial", False)
        out.append(rec)
    dst = os.path.join(HERE, "reference_vectors.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {len(out)} vectors to {dst}")


if __name__ == "__main__":
    main()

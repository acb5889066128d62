"""Known-answer vectors from the reference's unit tests (state strings, child
strings and operator labels), with the file:line each comes from.  Data only:
shared by the CPU oracle tests and the GPU parity tests.

Each expand KAT: (source, puzzle spec, parent state, [(child state or None, op)]).
`None` marks an empty slot of the GPU layout (INVALID child, op = ops mask).
"""

PANCAKE15_CHILDREN = [
    ("0 1 2 3 4 5 6 7 8 9 10 11 12 14 13", 0), ("0 1 2 3 4 5 6 7 8 9 10 11 14 13 12", 1),
    ("0 1 2 3 4 5 6 7 8 9 10 14 13 12 11", 2), ("0 1 2 3 4 5 6 7 8 9 14 13 12 11 10", 3),
    ("0 1 2 3 4 5 6 7 8 14 13 12 11 10 9", 4), ("0 1 2 3 4 5 6 7 14 13 12 11 10 9 8", 5),
    ("0 1 2 3 4 5 6 14 13 12 11 10 9 8 7", 6), ("0 1 2 3 4 5 14 13 12 11 10 9 8 7 6", 7),
    ("0 1 2 3 4 14 13 12 11 10 9 8 7 6 5", 8), ("0 1 2 3 14 13 12 11 10 9 8 7 6 5 4", 9),
    ("0 1 2 14 13 12 11 10 9 8 7 6 5 4 3", 10), ("0 1 14 13 12 11 10 9 8 7 6 5 4 3 2", 11),
    ("0 14 13 12 11 10 9 8 7 6 5 4 3 2 1", 12), ("14 13 12 11 10 9 8 7 6 5 4 3 2 1 0", 13),
]
# PuzzlesTests/PancakeTest.cpp:101-115: segments (index >> 29) of the 15-pancake children
PANCAKE15_SEGMENTS = [0] * 11 + [12, 181, 2729]

EXPAND_KATS = [
    ("PuzzlesTests/PancakeTest.cpp:12-33", {"kind": "pancake", "n": 5}, "0 1 2 3 4",
     [("0 1 2 4 3", 0), ("0 1 4 3 2", 1), ("0 4 3 2 1", 2), ("4 3 2 1 0", 3)]),
    ("PuzzlesTests/PancakeTest.cpp:199-223", {"kind": "pancake", "n": 7}, "5 1 3 6 0 2 4",
     [("5 1 3 6 0 4 2", 0), ("5 1 3 6 4 2 0", 1), ("5 1 3 4 2 0 6", 2), ("5 1 4 2 0 6 3", 3),
      ("5 4 2 0 6 3 1", 4), ("4 2 0 6 3 1 5", 5)]),
    ("PuzzlesTests/PancakeTest.cpp:58-116", {"kind": "pancake", "n": 15}, "0 1 2 3 4 5 6 7 8 9 10 11 12 13 14",
     PANCAKE15_CHILDREN),
    ("PuzzlesTests/PancakeTest.cpp:118-174", {"kind": "pancake", "n": 15, "inv_index": True},
     "0 1 2 3 4 5 6 7 8 9 10 11 12 13 14", PANCAKE15_CHILDREN),
    # GPU layout keeps empty slots (PuzzlesTests/SlidingPuzzleTest.cpp:43-61)
    ("PuzzlesTests/SlidingPuzzleTest.cpp:43-61", {"kind": "sliding", "width": 3, "height": 2}, "0 1 2 3 4 5",
     [(None, 15), ("3 1 2 0 4 5", None), (None, 15), ("1 0 2 3 4 5", None)]),
    ("PuzzlesTests/SlidingPuzzleTest.cpp:63-78", {"kind": "sliding", "width": 4, "height": 4},
     "0 1 2 3 4 5 6 7 8 9 10 11 12 13 14 15",
     [(None, 15), ("4 1 2 3 0 5 6 7 8 9 10 11 12 13 14 15", None), (None, 15),
      ("1 0 2 3 4 5 6 7 8 9 10 11 12 13 14 15", None)]),
    ("PuzzlesTests/SlidingPuzzleTest.cpp:80-98", {"kind": "sliding_packed", "width": 3, "height": 2}, "0 1 2 3 4 5",
     [(None, 15), ("3 1 2 0 4 5", None), (None, 15), ("1 0 2 3 4 5", None)]),
    ("PuzzlesTests/SlidingPuzzleTest.cpp:100-115", {"kind": "sliding_packed", "width": 4, "height": 4},
     "0 1 2 3 4 5 6 7 8 9 10 11 12 13 14 15",
     [(None, 15), ("4 1 2 3 0 5 6 7 8 9 10 11 12 13 14 15", None), (None, 15),
      ("1 0 2 3 4 5 6 7 8 9 10 11 12 13 14 15", None)]),
    # 4-peg Hanoi, GPU layout (PuzzlesTests/FourPegHanoiTest.cpp:245-377)
    ("PuzzlesTests/FourPegHanoiTest.cpp:245-265", {"kind": "hanoi", "disks": 5, "symmetry": False}, "1 2 3 4 5 0 0 0",
     [("2 3 4 5 0 1 0 0", 1), ("2 3 4 5 0 0 1 0", 2), ("2 3 4 5 0 0 0 1", 3), (None, 15), (None, 15), (None, 15)]),
    ("PuzzlesTests/FourPegHanoiTest.cpp:267-292", {"kind": "hanoi", "disks": 7, "symmetry": False},
     "1 5 6 0 2 0 0 3 4 7",
     [("5 6 0 1 2 0 0 3 4 7", 1), ("5 6 0 2 0 1 0 3 4 7", 2), ("5 6 0 2 0 0 1 3 4 7", 3),
      ("1 5 6 0 0 2 0 3 4 7", 2), ("1 5 6 0 0 0 2 3 4 7", 3), ("1 5 6 0 2 0 3 0 4 7", 2)]),
    ("PuzzlesTests/FourPegHanoiTest.cpp:294-314", {"kind": "hanoi", "disks": 5, "symmetry": True}, "1 2 3 4 5 0 0 0",
     [("2 3 4 5 0 1 0 0", 1), ("2 3 4 5 0 1 0 0", 1), ("2 3 4 5 0 1 0 0", 1), (None, 15), (None, 15), (None, 15)]),
    ("PuzzlesTests/FourPegHanoiTest.cpp:316-337", {"kind": "hanoi", "disks": 7, "symmetry": True},
     "3 4 6 0 1 5 7 0 2 0",
     [("1 3 4 6 0 5 7 0 2 0", 0), ("2 3 4 6 0 1 5 7 0 0", 0), ("4 6 0 1 5 7 0 3 0 2", 2),
      ("3 4 6 0 5 7 0 1 2 0", 2), ("3 4 6 0 5 7 0 2 0 1", 3), ("3 4 6 0 1 5 7 0 2 0", 2)]),
    ("PuzzlesTests/FourPegHanoiTest.cpp:339-353", {"kind": "hanoi", "disks": 19, "symmetry": True},
     "1 2 3 4 5 6 7 8 9 10 11 12 13 14 15 16 17 18 19 0 0 0",
     [("2 3 4 5 6 7 8 9 10 11 12 13 14 15 16 17 18 19 0 1 0 0", 1),
      ("2 3 4 5 6 7 8 9 10 11 12 13 14 15 16 17 18 19 0 1 0 0", 1)]),
    ("PuzzlesTests/FourPegHanoiTest.cpp:355-377", {"kind": "hanoi", "disks": 19, "symmetry": True},
     "18 19 0 13 14 15 16 17 0 1 2 3 4 5 6 7 8 9 10 11 12 0",
     [("13 18 19 0 14 15 16 17 0 1 2 3 4 5 6 7 8 9 10 11 12 0", 0),
      ("1 18 19 0 13 14 15 16 17 0 2 3 4 5 6 7 8 9 10 11 12 0", 0),
      ("19 0 18 0 13 14 15 16 17 0 1 2 3 4 5 6 7 8 9 10 11 12", 1),
      ("18 19 0 1 13 14 15 16 17 0 2 3 4 5 6 7 8 9 10 11 12 0", 1),
      ("18 19 0 14 15 16 17 0 13 0 1 2 3 4 5 6 7 8 9 10 11 12", 2),
      ("18 19 0 13 14 15 16 17 0 2 3 4 5 6 7 8 9 10 11 12 0 1", 3)]),
]

# CPU-twin order of the sliding-tile simple puzzle (SlidingPuzzleTest.cpp:12-41):
# only the existing children, first two listed.
SLIDING_CPU_KATS = [
    ("PuzzlesTests/SlidingPuzzleTest.cpp:12-26", 3, 2, "0 1 2 3 4 5", ["3 1 2 0 4 5", "1 0 2 3 4 5"]),
    ("PuzzlesTests/SlidingPuzzleTest.cpp:28-41", 4, 4, "0 1 2 3 4 5 6 7 8 9 10 11 12 13 14 15",
     ["4 1 2 3 0 5 6 7 8 9 10 11 12 13 14 15", "1 0 2 3 4 5 6 7 8 9 10 11 12 13 14 15"]),
]

# PuzzlesTests/FourPegHanoiTest.cpp:13-32,34-59: indexes of parsed states
HANOI_INDEX_KATS = [("PuzzlesTests/FourPegHanoiTest.cpp:15-16", 5, "1 2 3 4 5 0 0 0", 0),
                    ("PuzzlesTests/FourPegHanoiTest.cpp:36-37", 7, "1 5 6 0 2 0 0 3 4 7", 12532)]

# PuzzlesTests/PermutationMapTest.cpp:74-94
COMPACT_KATS = [("0 1 2 3 4 5 6 7 8 9 10 11 12 13 14 15", "0 0 0 0 0 0 0 0 0 0 0 0 0 0 0 0"),
                ("15 14 13 12 11 10 9 8 7 6 5 4 3 2 1 0", "15 14 13 12 11 10 9 8 7 6 5 4 3 2 1 0"),
                ("8 9 10 11 0 1 2 3 4 5 6 7", "8 8 8 8 0 0 0 0 0 0 0 0"),
                ("5 1 2 4 0 3", "5 1 1 2 0 0")]
RANK_KATS = [("0 1 2", 0), ("0 2 1", 1), ("1 0 2", 2), ("1 2 0", 3), ("2 0 1", 4), ("2 1 0", 5),
             ("0 1 2 3 4 5 6 7 8 9 10 11 12 13 14 15", 0),
             ("15 14 13 12 11 10 9 8 7 6 5 4 3 2 1 0", 20922789888000 - 1),
             ("8 9 10 11 0 1 2 3 4 5 6 7", 351590400), ("5 1 2 4 0 3", 634)]

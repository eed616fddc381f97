"""Writes sift_query_head1000.txt: the first 1000 rows of the reference's
data/sift_query.txt (real SIFT descriptors, integers 0..255, 128 columns),
in the same text-matrix format ("rows cols" header, one row per line).

Run once where the reference checkout is present:
    python3 tests/golden/make_sift_head.py /path/to/reference
"""
import os
import sys

ROWS = 1000


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "data", "sift_query.txt")
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sift_query_head1000.txt")
    with open(src) as f:
        r, c = f.readline().split()
        lines = [f.readline().strip() for _ in range(ROWS)]
    assert int(c) == 128 and all(len(l.split()) == 128 for l in lines)
    with open(dst, "w") as f:
        f.write(f"{ROWS} {c}\n")
        for l in lines:
            f.write(l + "\n")


if __name__ == "__main__":
    main()

"""Writes tests/golden/convert/: a small text tuple file and what the
reference converter writes for it. The converter is compiled from its own
source into oracle/_ref/text_to_bin by `make -C oracle ref`; this script runs
it in the build container. The fixtures are committed; this script is not
needed at test time.

- tuples.txt: 2 header lines, ids from 3 upward, spaces and tabs, repeated
  tuples, no self-loops, trailing newline.
- ref_rev0.npz / ref_rev1.npz: the converter's five output files (int64) for
  reverse = 0 and 1 with 2 lines skipped: beg_pos, csr, deg, head, weight.
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "convert")
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "text_to_bin")


def tuples_text(seed=7, n_tuples=60, n=17, base=3):
    rng = np.random.default_rng(seed)
    lines = ["# tuples for the converter fixture", "% second header line"]
    seen = []
    while len(seen) < n_tuples:
        u, v = (int(x) for x in rng.integers(0, n, 2))
        if u == v:
            continue
        seen.append((u, v))
        if rng.random() < 0.15:
            seen.append((u, v))  # a repeated tuple
    for k, (u, v) in enumerate(seen):
        sep = "\t" if k % 5 == 0 else " " * (1 + k % 3)
        lines.append(f"{u + base}{sep}{v + base}")
    return "\n".join(lines) + "\n"


def run_reference(text, reverse, skip, binary=REF_BIN):
    """The five int64 arrays the reference converter writes for `text`."""
    tmp = tempfile.mkdtemp()
    try:
        path = os.path.join(tmp, "t.txt")
        with open(path, "w") as f:
            f.write(text)
        subprocess.run([binary, path, str(reverse), str(skip)], check=True, capture_output=True, cwd=tmp)
        out = {}
        for name in ("beg_pos", "csr", "deg", "head", "weight"):
            out[name] = np.fromfile(f"{path}_{name}.bin", dtype=np.int64)
        return out
    finally:
        shutil.rmtree(tmp)


def main():
    if not os.path.exists(REF_BIN):
        sys.exit(f"{REF_BIN} missing: run make -C oracle ref")
    os.makedirs(OUT, exist_ok=True)
    text = tuples_text()
    with open(os.path.join(OUT, "tuples.txt"), "w") as f:
        f.write(text)
    for rev in (0, 1):
        np.savez(os.path.join(OUT, f"ref_rev{rev}.npz"), **run_reference(text, rev, 2))
    print("wrote", OUT)


if __name__ == "__main__":
    main()

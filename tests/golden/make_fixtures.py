"""Copies the reference's own golden files for the protein and DNA guide-tree
checks into tests/golden/ (gzip-compressed, unchanged content).

Source: test_data/correct/ of coldfunction/CUDA-clustalW, the files its
run_cpu.sh writes with the CPU clustalw-2.0.11 (`-ALIGN -QUIET`).  Run once
where the reference tree is mounted; the GPU box uses the copies.
"""
import gzip
import os
import shutil

SRC = "/root/reference/test_data/correct"
DST = os.path.dirname(os.path.abspath(__file__))
SETS = ["100_97", "100_498", "1000_97", "1000_498", "100_1002", "100_1523"]

if __name__ == "__main__":
    for s in SETS:
        for ext in ("fa", "dnd", "aln"):
            src = os.path.join(SRC, "%s.%s" % (s, ext))
            with open(src, "rb") as f, gzip.open(os.path.join(DST, "%s.%s.gz" % (s, ext)), "wb", 9) as g:
                shutil.copyfileobj(f, g)
    print("ok")

"""Access to the reference's golden files copied under tests/golden/."""
import gzip
import os

import numpy as np

from cwhip import fasta

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def text(name: str) -> str:
    with gzip.open(os.path.join(GOLDEN, name + ".gz"), "rt", encoding="latin-1") as f:
        return f.read()


def load(setname: str):
    """(names, codes, offs, lens, dna, dnd_text) of one reference data set."""
    names, seqs, dna = fasta.parse(text(setname + ".fa"))
    codes, offs, lens = fasta.pack(seqs)
    return names, codes, offs, lens, dna, text(setname + ".dnd")


def parse_aln(txt: str):
    """(names in file order, {name: gapped row}) of a CLUSTAL .aln file.
    Sequence lines are "name<blanks>residues"; the conservation line starts
    with blanks and is skipped, as are the header and blank lines."""
    order, rows = [], {}
    for line in txt.split("\n")[1:]:
        if not line.strip() or line[0] == " ":
            continue
        parts = line.split()
        if len(parts) < 2:
            continue
        name, res = parts[0], parts[1]
        if name not in rows:
            order.append(name)
            rows[name] = []
        rows[name].append(res)
    return order, {k: "".join(v) for k, v in rows.items()}


def row_text(codes) -> str:
    """Codes of an aligned row as the .aln prints them (gap codes -> '-')."""
    alpha = fasta.ALPHABET
    return "".join("-" if c >= 30 else alpha[c] for c in codes)

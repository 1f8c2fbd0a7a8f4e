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

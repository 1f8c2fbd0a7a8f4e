"""FASTA input as ClustalW 2.0.11 reads it.

Follows fileInput/PearsonFileParser.cpp:51-183 (name: text after '>' up to 150
characters, leading blanks dropped, cut at the first blank, trailing whitespace
trimmed), fileInput/FileParser.cpp:40-46 (only characters of the alphabet
"ABCDEFGHIKLMNOPQRSTUVWXYZ-" survive, lower case folded to upper),
alignment/Sequence.cpp:63-83 (codes: index in that alphabet, '-' -> 31) and
alignment/Sequence.cpp:189-229 with fileInput/FileReader.cpp:228-236 (DNA when
>= 85% of the non-gap characters of the SECOND sequence are A, C, G, T, U or N).
Paths are relative to CPU/clustalw-2.0.11/src.
"""
from __future__ import annotations

import numpy as np

ALPHABET = "ABCDEFGHIKLMNOPQRSTUVWXYZ-"
MAXNAMES = 150
GAP_CODE = 31

_CHARTAB = np.zeros(256, dtype=np.uint8)
for _c in ALPHABET:
    _CHARTAB[ord(_c)] = ord(_c)
    _CHARTAB[ord(_c.lower())] = ord(_c)
_CODE = np.full(256, 255, dtype=np.uint8)
for _i, _c in enumerate(ALPHABET[:-1]):
    _CODE[ord(_c)] = _i
_CODE[ord("-")] = GAP_CODE


def _name(header: str) -> str:
    name = header[1:1 + MAXNAMES]
    while name[:1] == " ":
        name = name[1:1 + MAXNAMES]
    if " " in name:
        name = name[:name.index(" ")]
    name = name.rstrip()
    return name.replace(" ", "_")


def is_dna(chars: bytes) -> bool:
    res = sum(1 for c in chars if c != ord("-"))
    bases = sum(1 for c in chars if c != ord("-") and chr(c) in "ACGTUN")
    if res == 0 or bases == 0:
        return False
    return np.float32(bases) / np.float32(res) >= np.float32(0.85)


def parse(text: str):
    """(names, [uint8 code arrays], dna_flag) for FASTA text."""
    names, raw = [], []
    cur = None
    for line in text.replace("\r", "\n").split("\n"):
        if line.startswith(">"):
            names.append(_name(line))
            cur = bytearray()
            raw.append(cur)
        elif cur is not None:
            b = np.frombuffer(line.encode("latin-1"), dtype=np.uint8)
            kept = _CHARTAB[b]
            cur.extend(kept[kept != 0].tobytes())
    seqs = [_CODE[np.frombuffer(bytes(r), dtype=np.uint8)].astype(np.uint8) for r in raw]
    dna = is_dna(bytes(raw[1])) if len(raw) > 1 else False
    return names, seqs, dna


def read(path: str):
    with open(path, "r", encoding="latin-1") as f:
        return parse(f.read())


def pack(seqs):
    lens = np.array([len(s) for s in seqs], dtype=np.int32)
    offs = np.zeros(len(seqs), dtype=np.int64)
    if len(seqs) > 1:
        offs[1:] = np.cumsum(lens[:-1], dtype=np.int64)
    codes = np.concatenate(seqs).astype(np.uint8) if len(seqs) else np.zeros(0, np.uint8)
    return codes, offs, lens

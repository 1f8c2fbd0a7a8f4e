"""cwhip: MI355X-native ClustalW pairwise DP (host-side Python front end).

The compute lives in libcwhip.so (gfx950 HIP kernels behind a C ABI, see
include/cwhip.h); this package marshals inputs, generates the seeded synthetic
workloads and reads FASTA.
"""
from . import seqgen  # noqa: F401
from .lib import (KERNEL_BATCHED, KERNEL_PLAIN, Context, CwhipError,  # noqa: F401
                  FastParams, MsaParams, TuneTable, clustal_text, distance, fast_params,
                  guide_from_tree, load,
                  JOB_ALIGNMENT, JOB_DISTANCES, MAX_LEN, Server, msa_params, nj_tree,
                  pair_params)

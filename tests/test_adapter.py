"""The C++ drop-in class compiles against the reference's own headers.

Needs the reference tree (present where the library is developed); the GPU box
does not have it, so the test skips there.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/CPU/clustalw-2.0.11/src"


@pytest.mark.skipif(not os.path.isdir(REF_SRC) or shutil.which("g++") is None,
                    reason="reference headers not present")
def test_adapter_compiles_against_reference_headers(tmp_path):
    src = os.path.join(ROOT, "cuda-clustalw_amd", "host", "HipFullPairwiseAlign.cpp")
    r = subprocess.run(["g++", "-std=c++11", "-fsyntax-only", "-w",
                        "-I" + os.path.join(REF_SRC, "pairwise"), "-I" + REF_SRC,
                        "-I" + os.path.join(ROOT, "include"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


@pytest.mark.skipif(not os.path.isdir(REF_SRC) or shutil.which("g++") is None,
                    reason="reference headers not present")
def test_profile_adapter_compiles_against_reference_headers(tmp_path):
    src = os.path.join(ROOT, "cuda-clustalw_amd", "host", "HipMyersMillerProfileAlign.cpp")
    r = subprocess.run(["g++", "-std=c++11", "-fsyntax-only", "-w",
                        "-I" + os.path.join(REF_SRC, "multipleAlign"), "-I" + REF_SRC,
                        "-I" + os.path.join(ROOT, "include"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]

"""The drop-in build (oracle/_ref/libspyker_dropin.so) on the CPU: it loads,
and the HIP seam defines every Core::cuda_* function the reference's dispatch
declares, so no Kind::CUDA call of the reference's code is left unresolved."""

import os
import re
import subprocess

import pytest

import dropin

REF_IMPL = "/root/reference/src/spyker/impl.h"

pytestmark = pytest.mark.skipif(not dropin.available(), reason="oracle/_ref/libspyker_dropin.so is not built")


def test_dropin_library_loads():
    lib = dropin.DropIn()
    assert lib.device_count() >= 0


@pytest.mark.skipif(not os.path.exists(REF_IMPL), reason="reference sources not present")
def test_seam_defines_every_declared_cuda_function():
    text = open(REF_IMPL).read()
    declared = set(re.findall(r"\bcuda_(\w+)\s*\(", text))
    declared |= set(re.findall(r"Declare\([^,]+,\s*(\w+)\s*,", text))
    declared.discard("name")  # the Declare macro's own parameter
    nm = subprocess.run(["nm", "-C", "-D", "--defined-only", dropin.PATH], capture_output=True, text=True,
                        check=True).stdout
    defined = set(re.findall(r"Spyker::Core::cuda_(\w+)\(", nm))
    assert len(declared) > 40
    assert declared <= defined, sorted(declared - defined)

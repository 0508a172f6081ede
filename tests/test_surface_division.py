"""The surface field kernel's division by the fixed radius^2 (mg_surface.hip sf_kernel) against IEEE
division on the CPU: q0 = r2 y, q = q0 + (r2 - radius2 q0) y with fused operations and y = RN(1 /
radius2) must equal RN(r2 / radius2) for every dividend the kernel sends that way (at least
2^-100). Checked over a stride of all such dividends for splat radii of the realtime, 2000-disc and
large drums and for divisors whose significand is all ones or a single bit."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _hex(f):
    return "%08x" % struct.unpack("<I", struct.pack("<f", np.float32(f)))[0]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_reciprocal_fma_quotient_is_ieee(tmp_path):
    exe = str(tmp_path / "division_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "division_check.c"), "-lm"],
                   check=True)
    radii = [np.float32(1.5) * np.float32(h) for h in (0.0013969999, 0.0125, 0.025, 0.1, 0.002)]
    divisors = [_hex(np.float32(r) * np.float32(r)) for r in radii]
    divisors += ["3f7fffff", "37ffffff", "38000000", "3a800001", "3b2aaaab"]
    out = subprocess.run([exe, "1009"] + divisors, check=True, capture_output=True, text=True).stdout.split()
    total, bad = int(out[0]), int(out[1])
    assert total > 5_000_000 and bad == 0

"""div_const (xnyzip_amd/csrc/exact_div.hpp), the quantizers' division by a
per-cloud constant, against IEEE division on the CPU: the divisors of the
truncated-octahedron and cube quantizers at many L2 bounds
(truncated_octahedron_quantizer.hpp:29, cube_quantizer.hpp:20: float(double(2b)
/ sqrt(5 or 3))), the recover divisors float(sqrt(5)) and float(sqrt(3))
(truncated_octahedron_quantizer.hpp:72, cube_quantizer.hpp:34), and divisors
whose significands are all ones or 1 + 2^-23.  The same function runs in the
kernels (v_fma_f32 is the correctly rounded fused multiply-add)."""
import math
import os
import struct
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bits(f: float) -> str:
    return "%08x" % struct.unpack("<I", struct.pack("<f", f))[0]


def _divisors():
    out = []
    for b in (0.001, 0.01, 0.1, 0.05, 0.5, 1.0, 1e-5, 250.0):
        two_b = float(np.float32(2) * np.float32(b))
        for root in (math.sqrt(5.0), math.sqrt(3.0)):
            out.append(float(np.float32(two_b / root)))
    out += [float(np.float32(math.sqrt(5.0))), float(np.float32(math.sqrt(3.0)))]
    out += [struct.unpack("<f", struct.pack("<I", u))[0]
            for u in (0x3FFFFFFF, 0x3F800001, 0x4B7FFFFF, 0x2EFFFFFF)]
    return out


def test_div_const_matches_ieee_division():
    d = tempfile.mkdtemp(prefix="exact_div_")
    exe = os.path.join(d, "exact_div_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", exe,
                    os.path.join(ROOT, "tests", "native", "exact_div_check.cpp")], check=True)
    r = subprocess.run([exe, *[_bits(x) for x in _divisors()]], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert ", 0 mismatches" in r.stdout


def test_div_const_rejects_planted_error():
    """The check fails the plain product by the reciprocal (a planted mistake:
    both corrections dropped), which is one ulp off for many dividends."""
    d = tempfile.mkdtemp(prefix="exact_div_")
    src = os.path.join(d, "one_step.cpp")
    text = open(os.path.join(ROOT, "tests", "native", "exact_div_check.cpp")).read()
    hdr = open(os.path.join(ROOT, "xnyzip_amd", "csrc", "exact_div.hpp")).read()
    hdr = hdr.replace("return std::fma(std::fma(-q1, d, a), r, q1);", "return q0;")
    assert "return q0;" in hdr
    open(os.path.join(d, "exact_div.hpp"), "w").write(hdr)
    text = text.replace('#include "../../xnyzip_amd/csrc/exact_div.hpp"', '#include "exact_div.hpp"')
    open(src, "w").write(text)
    exe = os.path.join(d, "one_step")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", exe, src], check=True)
    r = subprocess.run([exe, _bits(_divisors()[2])], capture_output=True, text=True)
    assert r.returncode == 1 and "mismatch" in r.stdout

"""round_half_away (xnyzip_amd/csrc/fround.hpp), the quantizer's roundf
(truncated_octahedron_quantizer.hpp:31-36 rounds with std::round), against
std::round on the CPU for all 2^32 float bit patterns.  The same expression
runs in the kernels (an addition and a truncation, both exact IEEE
operations on the device)."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _build(src: str, d: str) -> str:
    exe = os.path.join(d, "fround_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", exe, src], check=True)
    return exe


def test_round_half_away_matches_std_round_everywhere():
    d = tempfile.mkdtemp(prefix="fround_")
    exe = _build(os.path.join(ROOT, "tests", "native", "fround_check.cpp"), d)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert ", 0 mismatches" in r.stdout


def test_round_half_away_rejects_planted_error():
    """The check fails the form that adds exactly one half (a planted
    mistake: 0.49999997 then rounds up to 1)."""
    d = tempfile.mkdtemp(prefix="fround_")
    hdr = open(os.path.join(ROOT, "xnyzip_amd", "csrc", "fround.hpp")).read()
    planted = hdr.replace("0x1.fffffep-2f", "0.5f")
    assert planted != hdr
    open(os.path.join(d, "fround.hpp"), "w").write(planted)
    text = open(os.path.join(ROOT, "tests", "native", "fround_check.cpp")).read()
    text = text.replace('#include "../../xnyzip_amd/csrc/fround.hpp"', '#include "fround.hpp"')
    src = os.path.join(d, "planted.cpp")
    open(src, "w").write(text)
    exe = _build(src, d)
    r = subprocess.run([exe, "first"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 1 and "mismatch" in r.stdout

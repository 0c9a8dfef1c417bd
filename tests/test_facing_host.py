"""The renderer's facing test (swegl_amd/csrc/core.h:cross_z_sign), which
skips the normalisation's square root and division where they cannot change
the sign, against the reference's normalize(cross(l, r)).z compared with 0
(points.cpp:32-37, points.hpp:80-90) on 10 M random, tiny, huge, zero,
infinite and NaN inputs.  The host build of the same header computes what
gfx950 computes (IEEE single precision, no contraction)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")


def test_facing_shortcut_matches_normalised_cross():
    exe = os.path.join(CPP, "facing_check")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", CPP, "facing_check"], check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 mismatches" in r.stdout

"""The renderer's pow32 (swegl_amd/csrc/shade.h), the float result of
pow((double)s, 32) at pixel_shaders.cpp:193, against the C library's pow
for every float s in [1/16, 1.2] (all inputs whose s^32 is a normal float)
plus sampled tiny inputs.  pow32 is IEEE double arithmetic only, so the
host build of the same header computes what gfx950 computes."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")


def test_pow32_matches_libm_exhaustively():
    exe = os.path.join(CPP, "pow32_check")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", CPP, "pow32_check"], check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 mismatches" in r.stdout

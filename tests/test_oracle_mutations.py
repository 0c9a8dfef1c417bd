"""Planted mistakes: each is one plausible slip in a restatement of the
reference, planted in a copy of oracle/refcpu.cpp; the CPU pin tests must
reject every one of them.

For each mistake the test compiles the mutated copy with the oracle's own
flags, points the oracle binding at it (SWEGL_ORACLE_LIB) and runs the pin
tests in a child process; the run must fail.  The golden digests
(tests/test_golden.py) are left out on purpose: they are the oracle's own
output, so they pin drift, not the reference.  Everything here runs on the
CPU.

The reference statement each mistake departs from:

  closed_form_x            renderer.cpp:553-554   x += ratio per row
  closed_form_interp       interpolator.hpp:96-100 ta += ts; ba += bs per step
  restart_long_side        renderer.cpp:399-409,439-459 long side carried over
  one_normalisation        vertex_shaders.hpp:63, points.hpp:139-150 (three)
  two_normalisations       the same
  hierarchy_order          vertex_shaders.hpp:18  parent * local
  displace_from_y0         renderer.cpp:401-409   first row clipped to vy
  span_end_floor           renderer.cpp:470       x2 = ceil(xr)
  lr_flipped_upper         renderer.cpp:425       ratio_long > ratio_short
  near_reject_float_literal renderer.cpp:489      z <= 0.001 (double) rejected
  depth_test_strict        renderer.cpp:491       z >= zb rejected (first wins)
  frustum_all_vertices     renderer.cpp:58-70     any vertex meets each side
  near_clip_at_zero        renderer.cpp:290-293   cut at z' = 0.001
  clip_normal_raw          renderer.cpp:294-296   clip normal re-normalised
  bilinear_untransposed    pixel_shaders.cpp:353-354 row index from t.x
  phong_cutoff             pixel_shaders.cpp:178  diffuse < 0.05 skipped
  specular_power           pixel_shaders.cpp:189  pow(s, 32)
  combiner_one_sqrt        pixel_shaders.hpp:170-171 fourth root of light
  ambient_dropped          pixel_shaders.cpp:205  ambient in the light sum
  dof_radius_focus_test    post_shaders.hpp:95    tap counts when bf != 0
  dof_box_inclusive        post_shaders.hpp:88-92 [y - r, y + r)
  strip_winding            renderer.cpp:147,204   odd strip triangles swap
"""
import concurrent.futures as cf
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "refcpu.cpp")
FLAGS = ["-O3", "-msse4", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-std=c++17", "-shared",
         "-w", "-I", os.path.join(ROOT, "oracle"), "-I", os.path.join(ROOT, "include")]
PIN_TESTS = sorted(os.path.join("tests", f) for f in os.listdir(os.path.join(ROOT, "tests"))
                   if f.startswith("test_oracle_") and f.endswith(".py") and f != "test_oracle_mutations.py")
PIN_TESTS += [os.path.join("tests", "test_meshes.py"), os.path.join("tests", "test_assets.py")]

MUTATIONS = {
    "closed_form_x": (
        "    for (; y < y_end; y++) {\n        int x1 = std::max(cvtt(std::ceil(L.x)), vx);",
        "    const float lx0 = L.x, rx0 = R.x;\n    const int ys0 = y;\n"
        "    for (; y < y_end; y++) {\n"
        "        if (y != ys0) { L.x = lx0 + L.ratio * (float)(y - ys0); R.x = rx0 + R.ratio * (float)(y - ys0); }\n"
        "        int x1 = std::max(cvtt(std::ceil(L.x)), vx);"),
    "closed_form_interp": (
        "    void step()\n    {\n        topalpha += topstep;\n        bottomalpha += bottomstep;\n"
        "        ualpha = topalpha / bottomalpha;\n    }",
        "    float ta0 = 0, ba0 = 0;\n    int k = 0;\n"
        "    void step()\n    {\n        if (k == 0) { ta0 = topalpha; ba0 = bottomalpha; }\n        k++;\n"
        "        topalpha = ta0 + topstep * (float)k;\n        bottomalpha = ba0 + bottomstep * (float)k;\n"
        "        ualpha = topalpha / bottomalpha;\n    }"),
    "restart_long_side": (
        "    if (y1 < vy + vh) {\n        side_short.ratio = (v2->x - v1->x) / (v2->y - v1->y);",
        "    if (y1 < vy + vh) {\n"
        "        { int yl = std::max(y1, vy); side_long.it.init_self(v2->y - v0->y, v0->z, v2->z);\n"
        "          side_long.it.displace((float)yl - v0->y);\n"
        "          side_long.x = v0->x + side_long.ratio * ((float)yl - v0->y); }\n"
        "        side_short.ratio = (v2->x - v1->x) / (v2->y - v1->y);"),
    "one_normalisation": (
        "mv.normal_world = normalize(normalize(rotate_n(mv.normal, nd.rs)));",
        "mv.normal_world = rotate_n(mv.normal, nd.rs);"),
    "two_normalisations": (
        "mv.normal_world = normalize(normalize(rotate_n(mv.normal, nd.rs)));",
        "mv.normal_world = normalize(rotate_n(mv.normal, nd.rs));"),
    "hierarchy_order": (
        "    matmul(parent, n.local, n.o2w);",
        "    matmul(n.local, parent, n.o2w);"),
    "displace_from_y0": (
        "        side_long.it.displace((float)vy - v0->y);\n        side_long.x = v0->x + side_long.ratio * ((float)vy - v0->y);",
        "        side_long.it.displace((float)y0 - v0->y);\n        side_long.x = v0->x + side_long.ratio * ((float)y0 - v0->y);"),
    "span_end_floor": (
        "        int x2 = std::min(cvtt(std::ceil(R.x)), vx + vw);",
        "        int x2 = std::min(cvtt(std::floor(R.x)) + 1, vx + vw);"),
    "lr_flipped_upper": (
        "        bool lr = side_long.ratio > side_short.ratio;",
        "        bool lr = side_long.ratio < side_short.ratio;"),
    "near_reject_float_literal": (
        "            if (z <= 0.001)",
        "            if (z <= 0.001f)"),
    "depth_test_strict": (
        "            if (z >= vp->zb[off])",
        "            if (z > vp->zb[off])"),
    "frustum_all_vertices": (
        "    return (p.x >= -1 || q.x >= -1 || r.x >= -1) && (p.y >= -1 || q.y >= -1 || r.y >= -1) &&",
        "    return (p.x >= -1 && q.x >= -1 && r.x >= -1) && (p.y >= -1 || q.y >= -1 || r.y >= -1) &&"),
    "near_clip_at_zero": (
        "        float cut_1 = (v0.z - 0.001f) / (v0.z - v1.z);",
        "        float cut_1 = (v0.z - 0.0f) / (v0.z - v1.z);"),
    "clip_normal_raw": (
        "        nv.normal = normalize(a.normal + (b.normal - a.normal) * cut);",
        "        nv.normal = a.normal + (b.normal - a.normal) * cut;"),
    "bilinear_untransposed": (
        "        float v = t.x;   /* row coordinate comes from t.x */\n        float u = t.y;",
        "        float v = t.y;\n        float u = t.x;"),
    "phong_cutoff": (
        "            float diffuse = psl.intensity / light_distance_squared;\n            if (diffuse < 0.05)",
        "            float diffuse = psl.intensity / light_distance_squared;\n            if (diffuse < 0.005)"),
    "specular_power": (
        "                specular = (float)std::pow((double)specular, (double)p);",
        "                specular = (float)std::pow((double)specular, (double)(p / 2));"),
    "combiner_one_sqrt": (
        "            light = std::sqrt(light);\n            light = std::sqrt(light);",
        "            light = std::sqrt(light);"),
    "ambient_dropped": (
        "        return cvtt(65536 * (sc.ambient + face_sun_intensity + dynamic_lights_intensity));",
        "        return cvtt(65536 * (face_sun_intensity + dynamic_lights_intensity));"),
    "dof_radius_focus_test": (
        "                if (bf[(size_t)j * w + i] != 0) {",
        "                if (cvtt(bf[(size_t)j * w + i]) != 0) {"),
    "dof_box_inclusive": (
        "                for (int j = std::max(0, y - radius); j < std::min(h, y + radius); j++)",
        "                for (int j = std::max(0, y - radius); j < std::min(h, y + radius + 1); j++)"),
    "strip_winding": (
        "                    fill_triangle(I[i - 2], I[i - 1 + (i & 1)], I[i - (i & 1)], n, p, dc, tri_index++);",
        "                    fill_triangle(I[i - 2], I[i - 1], I[i], n, p, dc, tri_index++);"),
}


def _source():
    with open(SRC) as f:
        return f.read()


def test_each_mutation_site_is_unique():
    src = _source()
    for name, (old, new) in MUTATIONS.items():
        assert src.count(old) == 1, f"{name}: the mutated statement occurs {src.count(old)} times"
        assert old != new


def _run_one(name, workdir):
    old, new = MUTATIONS[name]
    src = _source().replace(old, new, 1)
    cpp = os.path.join(workdir, f"{name}.cpp")
    so = os.path.join(workdir, f"liboracle_{name}.so")
    with open(cpp, "w") as f:
        f.write(src)
    cc = subprocess.run(["g++", *FLAGS, "-o", so, cpp], capture_output=True, text=True)
    if cc.returncode != 0:
        return ("compile", cc.stderr[-2000:])
    env = dict(os.environ, SWEGL_ORACLE_LIB=so, PYTHONDONTWRITEBYTECODE="1")
    run = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                          "-m", "not gpu", *PIN_TESTS], cwd=ROOT, env=env,
                         capture_output=True, text=True, timeout=600)
    return (run.returncode, run.stdout[-3000:])


@pytest.fixture(scope="module")
def outcomes(tmp_path_factory):
    workdir = str(tmp_path_factory.mktemp("mutants"))
    workers = max(1, min(6, (os.cpu_count() or 2) - 1))
    with cf.ThreadPoolExecutor(workers) as ex:
        futs = {name: ex.submit(_run_one, name, workdir) for name in MUTATIONS}
        return {name: f.result() for name, f in futs.items()}


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_planted_mistake_is_rejected(outcomes, name):
    rc, out = outcomes[name]
    assert rc != "compile", f"{name}: mutated oracle did not compile:\n{out}"
    # pytest exits 1 when tests ran and some failed; anything else (0: all
    # passed, 2+: interrupted or a collection error) is not a rejection
    assert rc == 1, f"{name}: the pin tests did not reject this mistake (exit {rc}):\n{out}"

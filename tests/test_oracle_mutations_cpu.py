"""The oracle's pins catch plausible mistakes.

Each case plants one mistake in a copy of oracle/oracle.cpp, builds it with
g++ into a temporary directory and runs the pin that should notice it
against that build (through TFG_ORACLE_LIB, in a child pytest); the pin
must fail.  Nothing here touches the GPU or the HIP library.
"""
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "oracle.cpp")
RNG = "tests/test_oracle_rng_camera.py::"
SHD = "tests/test_oracle_shading_cpu.py::"

# (name, text in oracle.cpp, planted replacement, pin that must fail)
MUTATIONS = [
    ("pcg_shift", "((s >> 28u) + 4u)", "((s >> 27u) + 4u)", RNG + "test_pcg32_matches_unbounded_integers"),
    ("seed_constant", "(uint32_t)frame * 719393u", "(uint32_t)frame * 719391u",
     RNG + "test_stream_draws_and_float_conversion"),
    ("camera_left_sign", "out->izquierda = {l[0], l[1], l[2]};", "out->izquierda = {-l[0], -l[1], -l[2]};",
     RNG + "test_camera_basis_for_any_orientation"),
    ("tmax_exclusive", "if (!(tt > tmin) || tt > tmax) return 0;", "if (!(tt > tmin) || tt >= tmax) return 0;",
     RNG + "test_tmin_tmax_ends"),
    ("cosine_map", "std::acos(std::sqrt(1.0 - r2))", "std::acos(1.0 - r2)", SHD + "test_cosine_draw"),
    ("ggx_D_square", "return a2 / (kPiD * d * d);", "return a2 / (kPiD * d);", SHD + "test_ggx_D_normalised"),
    ("ggx_G1", "std::sqrt(a2 + (1.0 - a2) * nv * nv)", "std::sqrt(a2 + (1.0 - a2) * nv)",
     SHD + "test_ggx_G1_tan_form"),
    ("vndf_blend", "const double s = 0.5 * (1.0 + vh.z);", "const double s = 0.5 * (1.0 - vh.z);",
     SHD + "test_ggx_visible_normals_and_sampling"),
    ("sphere_root", "if (t1 > 1e-6f && t1 < t2) return t1;", "if (t1 > 1e-6f && t1 < t2) return t2;",
     SHD + "test_analytic_sphere_and_plane"),
    ("eta_inverted", "const double eta = ior / etat;", "const double eta = etat / ior;",
     SHD + "test_integrator_closed_form_scenes[glass_lens]"),
    ("diffuse_weight", "camino = cdiv(camino * m.kd, (float)pd);", "camino = camino * m.kd;",
     SHD + "test_integrator_closed_form_scenes[mixed_lobes]"),
    ("mirror_weight", "camino = cdiv(camino * m.ks, (float)ps);", "camino = camino * m.ks;",
     SHD + "test_integrator_closed_form_scenes[mixed_lobes]"),
    ("lambert_pi", "Col f = cdiv(m.kd, (float)kPiD);", "Col f = m.kd;",
     SHD + "test_integrator_closed_form_scenes[point_light]"),
    ("distance_clamp", "dist = std::max(dist, 1.0f);", "dist = std::max(dist, 0.5f);",
     SHD + "test_integrator_closed_form_scenes[point_light]"),
    ("shadow_tmax", "if (o_occluded(sc, ps + dl * kEpsF, dl, dist)) continue;",
     "if (o_occluded(sc, ps + dl * kEpsF, dl, dist * 0.5f)) continue;",
     SHD + "test_integrator_closed_form_scenes[point_light]"),
    ("footprint_rule", "const bool foot_ok = acur >= a0 * 0.01f;", "const bool foot_ok = true;",
     SHD + "test_cache_capture_between_two_planes"),
    ("tail_reflectance", "const Col ts = camino * rt;", "const Col ts = camino;",
     SHD + "test_cache_capture_between_two_planes"),
]


def _run(case, tmp):
    name, old, new, pin = case
    src = open(SRC).read()
    assert old in src, f"{name}: anchor not found in oracle.cpp"
    cpp = os.path.join(tmp, f"{name}.cpp")
    lib = os.path.join(tmp, f"lib_{name}.so")
    with open(cpp, "w") as f:
        f.write(src.replace(old, new))
    b = subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I", os.path.join(ROOT, "oracle"),
                        "-shared", "-o", lib, cpp, "-lpthread"], capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-2000:]
    env = dict(os.environ, TFG_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", pin], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=600)
    return name, r.returncode, r.stdout[-400:]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ to build the planted copies")
def test_planted_mistakes_fail_their_pins():
    with tempfile.TemporaryDirectory() as tmp:
        with ThreadPoolExecutor(max_workers=4) as ex:
            results = list(ex.map(lambda c: _run(c, tmp), MUTATIONS))
    missed = [(n, out) for n, rc, out in results if rc != 1]
    assert not missed, missed

"""Each independent check must catch the mistake it exists for.

Builds variants of the oracle with one planted mistake each (a textual
change to a copy of oracle/softrender_oracle.c, compiled with the oracle's
own flags into a temporary folder) and runs tests/oracle_checks.py on
them: every variant must fail at least one check.  The mistakes are the
plausible ones for the reference's triangle path: swapped or misassigned
barycentric weights (src/numeric/interpolate.rs:26-28), a reversed cull
(src/pipeline/stages/rasterization/triangle.rs:54-61), transposed
matrices in the vertex programs, a dropped or misdirected term in the
shading programs, a wrong OVER alpha, a strict depth test (triangle.rs:126),
a pixel corner instead of its centre (triangle.rs:104-105), a reordered
w = 1 - u - v (triangle.rs:110), and a viewport without the y flip
(src/geometry/clipvertex.rs:107-112).
"""
import os
import subprocess

import pytest

import oracle
import oracle_checks

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(os.path.dirname(HERE), "oracle", "softrender_oracle.c")

MUTANTS = {
    "bary_uv_swapped": ("return ((a * u) + (b * v)) + (c * w);", "return ((a * v) + (b * u)) + (c * w);"),
    "bary_wrong_vertex": ("return ((a * u) + (b * v)) + (c * w);", "return ((a * u) + (b * w)) + (c * v);"),
    "cull_reversed": ("int wind = signbit(A) ? SR_CULL_CW : SR_CULL_CCW;",
                      "int wind = signbit(A) ? SR_CULL_CCW : SR_CULL_CW;"),
    "mvp_transposed": ("acc = acc + U[r * 4 + k] * pin[k];", "acc = acc + U[k * 4 + r] * pin[k];"),
    "normal_matrix_transposed": ("acc = acc + U[16 + r * 3 + k] * nin[k];",
                                 "acc = acc + U[16 + k * 3 + r] * nin[k];"),
    "specular_dropped": ("float spec = U[11] * s;", "float spec = 0.0f * s;"),
    "half_vector_is_light": ("float h[3] = {U[0] + v[0], U[1] + v[1], U[2] + v[2]};",
                             "float h[3] = {U[0], U[1], U[2]};"),
    "lambert_term_dropped": ("float lit = U[3] + U[4] * diff;", "float lit = U[3] + U[4];"),
    "over_alpha_ignores_dst": ("uint32_t oa = quant8(a + (unorm8(dst >> 24) * ia));", "uint32_t oa = quant8(a);"),
    "depth_test_strict": ("if (z < 0.0f && z >= fb->depth[index]) {\n                    float k[OR_MAXK];",
                          "if (z < 0.0f && z > fb->depth[index]) {\n                    float k[OR_MAXK];"),
    "pixel_corner": ("float x = (float)px + 0.5f, y = (float)py + 0.5f;", "float x = (float)px, y = (float)py;"),
    "w_reordered": ("float w = 1.0f - u - v;", "float w = 1.0f - v - u;"),
    "viewport_no_y_flip": ("{0.0f, (top - bottom) / -2.0f, 0.0f, (top + bottom) / 2.0f},",
                           "{0.0f, (top - bottom) / 2.0f, 0.0f, (top + bottom) / 2.0f},"),
}


@pytest.fixture(scope="module")
def source():
    with open(SRC) as f:
        return f.read()


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_mistake_is_caught(name, source, tmp_path):
    old, new = MUTANTS[name]
    assert source.count(old) == 1, f"{name}: the line to change is not in the oracle exactly once"
    src = tmp_path / "mutant.c"
    src.write_text(source.replace(old, new))
    lib = tmp_path / f"liboracle_{name}.so"
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-shared",
                    f"-I{os.path.dirname(SRC)}", "-o", str(lib), str(src), "-lm", "-lpthread"], check=True)
    L = oracle.load(str(lib))
    for check in oracle_checks.ALL:  # cheapest first; the random scenes last
        try:
            check(L)
        except AssertionError:
            return
    pytest.fail(f"no check catches the planted mistake {name}")

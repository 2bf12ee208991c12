"""The oracle's triangle path against independent pins (tests/oracle_checks.py):
exact barycentric values at the vertices and float64 inside
(src/pipeline/stages/rasterization/triangle.rs:108-134), the cull
orientation of a known triangle (triangle.rs:54-61), the vertex programs
against float64, closed forms of the shading programs and the OVER blend,
and a numpy-float32 restatement of the whole triangle draw compared bit
for bit on random tiny scenes."""
import pytest

import oracle
import oracle_checks


@pytest.mark.parametrize("check", oracle_checks.ALL, ids=lambda f: f.__name__)
def test_oracle_independent(check):
    check(oracle.lib())

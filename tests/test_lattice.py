"""Lattice declaration -> generated constants (the DSL preamble step of
LatticeDescriptor, src/parsing/registry/LatticeDescriptor.cpp:9-80)."""
from fractions import Fraction

import numpy as np

from fluidloom_amd import lattice as Lt


def test_reference_fixture_and_validation():
    d = Lt.d3q19()
    # the seven velocities and weights of tests/unit/parsing/test_parsing.cpp:23-30
    assert d.stencil_vectors[:7] == [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    assert [float(w) for w in d.weights[:7]] == [1 / 3] + [1 / 18] * 6
    assert d.validate()
    bad = Lt.LatticeDescriptor("X", d.stencil_vectors, [Fraction(1, 19)] * 18 + [Fraction(1, 10)], Fraction(1, 3))
    assert not bad.validate()
    assert not Lt.LatticeDescriptor("E", [], [], Fraction(1, 3)).validate()


def test_isotropy_and_opposites():
    d = Lt.d3q19()
    c = np.array(d.stencil_vectors, float)
    w = np.array([float(x) for x in d.weights])
    assert abs(w.sum() - 1) < 1e-15
    assert np.allclose((w[:, None] * c).sum(0), 0)
    assert np.allclose(np.einsum("q,qa,qb->ab", w, c, c), np.eye(3) / 3)
    # fourth moment of D3Q19: sum w c_a^2 c_b^2 = cs^4 (a != b)
    assert abs((w * c[:, 0] ** 2 * c[:, 1] ** 2).sum() - 1 / 9) < 1e-15
    opp = d.compute_opposites()
    assert all((c[opp[i]] == -c[i]).all() for i in range(19)) and d.has_opposites
    assert opp == [0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17]
    # a stencil with an unpaired velocity has no opposite table
    one = Lt.LatticeDescriptor("U", [(0, 0, 0), (1, 0, 0)], [Fraction(1, 2)] * 2, Fraction(1, 3))
    one.compute_opposites()
    assert not one.has_opposites


def test_generated_header_is_current(tmp_path):
    p = Lt.generate(tmp_path / "lattice_gen.hpp")
    assert p.read_text() == (Lt.CSRC / "lattice_gen.hpp").read_text()
    txt = p.read_text()
    assert "constexpr int D3Q19_Q = 19;" in txt and "constexpr double D3Q19_CS2 = 1.0 / 3.0;" in txt


def test_opencl_preamble_form():
    d = Lt.d3q19()
    pre = d.generate_opencl()
    assert "#define D3Q19_Q 19" in pre and "#define D3Q19_CS2 0.333333f" in pre
    assert "__constant float D3Q19_weights[19] = {0.333333f, 0.0555556f" in pre
    assert "__constant uchar D3Q19_opposite[19] = {0, 2, 1, 4, 3" in pre

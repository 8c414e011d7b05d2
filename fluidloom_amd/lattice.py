"""Lattice declarations and the code generator that turns them into the
HIP constants the lattice kernels are compiled with.

The reference declares lattices in its DSL (lattices.fl) and turns each into
an OpenCL preamble with LatticeDescriptor::generateOpenCLCode
(src/parsing/registry/LatticeDescriptor.cpp:51-80): NAME_Q, NAME_CS2, a
weight table and an opposite table, after validate() (weights sum to 1 +-
1e-6, :9-28) and computeOpposites() (:30-49).  Here the same descriptor
emits a C++/HIP header of constexpr tables (velocities, weights, opposites,
cs^2); `generate()` writes fluidloom_amd/csrc/lattice_gen.hpp, which the
build regenerates before compiling, so the D3Q19 arithmetic in the kernels
comes from this declaration.

    python -m fluidloom_amd.lattice   # rewrite lattice_gen.hpp
"""
from __future__ import annotations

from dataclasses import dataclass, field
from fractions import Fraction
from pathlib import Path

CSRC = Path(__file__).resolve().parent / "csrc"


@dataclass
class LatticeDescriptor:
    name: str
    stencil_vectors: list
    weights: list          # exact fractions
    cs2: Fraction
    opposite: list = field(default_factory=list)
    has_opposites: bool = False

    def validate(self) -> bool:
        """Same rule as the reference: non-empty, matching sizes, sum(w) = 1 +- 1e-6."""
        if not self.stencil_vectors or not self.weights or len(self.stencil_vectors) != len(self.weights):
            return False
        return abs(float(sum(self.weights)) - 1.0) <= 1e-6

    def compute_opposites(self):
        """opposite[i] = first j with c_j = -c_i (the reference's search order)."""
        self.opposite = []
        self.has_opposites = True
        for c in self.stencil_vectors:
            j = next((j for j, d in enumerate(self.stencil_vectors) if tuple(d) == tuple(-v for v in c)), None)
            if j is None:
                self.has_opposites = False
                j = 0
            self.opposite.append(j)
        return self.opposite

    def generate_opencl(self) -> str:
        """The reference's preamble text (for comparison with its generator)."""
        q = len(self.stencil_vectors)
        out = f"// Lattice: {self.name}\n#define {self.name}_Q {q}\n#define {self.name}_CS2 {float(self.cs2):g}f\n\n"
        out += f"__constant float {self.name}_weights[{q}] = {{" + ", ".join(f"{float(w):g}f" for w in self.weights)
        out += "};\n"
        if self.has_opposites:
            out += f"__constant uchar {self.name}_opposite[{q}] = {{" + ", ".join(str(o) for o in self.opposite)
            out += "};\n"
        return out

    def generate_hip(self) -> str:
        """constexpr tables for the HIP kernels (exact weights as ratios)."""
        if not self.validate():
            raise ValueError(f"lattice {self.name} fails validation")
        if not self.has_opposites:
            self.compute_opposites()
        if not self.has_opposites:
            raise ValueError(f"lattice {self.name} has a velocity without an opposite")
        q = len(self.stencil_vectors)
        n = self.name
        lines = [f"// Lattice {n}: Q = {q}, cs^2 = {self.cs2}",
                 f"constexpr int {n}_Q = {q};",
                 f"constexpr double {n}_CS2 = {self.cs2.numerator}.0 / {self.cs2.denominator}.0;",
                 f"constexpr int {n}_C[{q}][3] = {{" + ", ".join(
                     "{" + ", ".join(str(int(v)) for v in c) + "}" for c in self.stencil_vectors) + "};",
                 f"constexpr double {n}_W[{q}] = {{" + ", ".join(
                     f"{w.numerator}.0 / {w.denominator}.0" for w in self.weights) + "};",
                 f"constexpr int {n}_OPP[{q}] = {{" + ", ".join(str(o) for o in self.opposite) + "};"]
        return "\n".join(lines) + "\n"


def d3q19() -> LatticeDescriptor:
    """D3Q19: rest, the six faces in the order of the reference's lattice
    fixture (tests/unit/parsing/test_parsing.cpp:23-30), then the twelve
    edges in opposite pairs."""
    c = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
         (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1), (-1, 0, 1),
         (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1)]
    w = [Fraction(1, 3)] + [Fraction(1, 18)] * 6 + [Fraction(1, 36)] * 12
    d = LatticeDescriptor("D3Q19", c, w, Fraction(1, 3))
    d.compute_opposites()
    return d


HEADER = """// Generated by fluidloom_amd/lattice.py from the lattice declarations there;
// do not edit.  The lattice kernels take their velocities, weights and
// opposites from these tables.
#pragma once

namespace fl {
namespace lattice {

"""


def generate(path: Path | None = None) -> Path:
    path = path or CSRC / "lattice_gen.hpp"
    text = HEADER + d3q19().generate_hip() + "\n}  // namespace lattice\n}  // namespace fl\n"
    if not path.exists() or path.read_text() != text:
        path.write_text(text)
    return path


if __name__ == "__main__":
    print(generate())

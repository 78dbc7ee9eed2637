// The degree-2, 4-point cubature rule of the P1 tetrahedron (points with
// barycentric coordinates a = (5 - sqrt 5)/20 and b = (5 + 3 sqrt 5)/20, weight
// 1/24 of det J each: the rule behind ES::ES's weighted basis values,
// ES.cpp:193-195, basis.hpp:161-163) and the P1 basis values at its points.
// Shared by the host setup, the library's factorisation check and the FP64
// push, so that all of them compute the same doubles.
#pragma once

namespace mpic {

constexpr double kCubA = 0.1381966011250105, kCubB = 0.5854101966249685, kCubW = 1.0 / 24.0;
constexpr double kCub[4][3] = {{kCubA, kCubA, kCubA}, {kCubB, kCubA, kCubA}, {kCubA, kCubB, kCubA}, {kCubA, kCubA, kCubB}};

// Gradients of the P1 basis on the reference tet (basis.hpp): the physical
// gradient is G[i][d] = sum_j J^-1[j][d] * kGradRef[i][j], summed in order j = 0, 1, 2
// from 0 (PIC.cpp:203-242 through Intrepid's HGRAD transform).
constexpr double kGradRef[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};

// N_i at cubature point c, as the setup computes it: 1 - p0 - p1 - p2, p0, p1, p2.
constexpr double cub_basis(int i, int c) {
  return i == 0 ? 1 - kCub[c][0] - kCub[c][1] - kCub[c][2] : kCub[c][i - 1];
}

}  // namespace mpic

// Coefficients of the 4-stage Rosenbrock method ROK4E with its embedded error estimate (Define.h:28-43), in
// the scaled form both chemistry integrators use: stage vectors solve (I - h G J) k = f, the G_ij are divided
// by G, and E_i are the error weights.
#pragma once

namespace corefl {
namespace rok4e {

constexpr double G = 0.572816062482135;
constexpr double A21 = 0.432364435748567, A31 = -0.51421131687617, A32 = 1.38227114461736;
constexpr double G21 = -0.602765307997356 / G, G31 = -1.389195789724843 / G, G32 = 1.072950969011413 / G;
constexpr double G41 = 0.992356412977094 / G, G42 = -1.390032613873701 / G, G43 = -0.440875890223325 / G;
constexpr double B1 = 0.194335256262729, B2 = 0.483167813989227, B4 = 0.322496929748044;
constexpr double E1 = -0.217819895945721 - B1, E2 = 1.03130847478467 - B2, E3 = 0.186511421161047, E4 = -B4;

// A step whose size h satisfies kStretch h >= (time left in the interval) is taken to the interval's end.
constexpr double kStretch = 1.05;

} // namespace rok4e
} // namespace corefl

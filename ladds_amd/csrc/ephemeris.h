// ephemeris.h -- host-side per-sub-step environment.
#pragma once

#include "engine.h"

namespace ladds {

void sun_position(double t, double out[3]);
void moon_position(double t, double out[3]);
double earth_rotation_angle(double t);
void substep_env(double t, SubstepEnv<double> &e64, SubstepEnv<float> &e32);
void yoshida_coefficients(double c[4], double d[3]);

}  // namespace ladds

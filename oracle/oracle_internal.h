/*
 * oracle_internal.h -- constants and the per-sub-step environment shared by the oracle's
 * translation units (oracle.c, gradient.c).  TEST INFRASTRUCTURE ONLY.
 */
#ifndef LADDS_ORACLE_INTERNAL_H
#define LADDS_ORACLE_INTERNAL_H

#include "oracle.h"

/* ------------------------------------------------------------------------- */
/* constants                                                                  */
/* ------------------------------------------------------------------------- */
/* Earth gravity field: EGM96 (GM, reference radius, unnormalised C20/C22/S22
 * obtained from the normalised coefficients C20=-4.84165371736e-4,
 * C22=2.43914352398e-6, S22=-1.40016683654e-6).  GM = 398600.4418 km^3/s^2 is
 * also the value the reference's 1-s Kepler vector (SimulationTest.cpp:39-57)
 * reproduces. */
#define MU_EARTH 398600.4418
#define R_EARTH 6378.1363
#define J2_EARTH 1.0826266835531513e-3
#define C22_EARTH 1.574460374564035e-6
#define S22_EARTH -9.03803806638557e-7
/* third bodies (DE405 GM values) and the astronomical unit (IAU 2012) */
#define GM_SUN 1.32712440018e11
#define GM_MOON 4902.800066
#define AU_KM 149597870.7
/* solar radiation pressure at 1 AU [N/m^2], Montenbruck & Gill eq. 3.69 */
#define P_SUN 4.56e-6
/* Earth rotation rate [rad/s] and GMST at J2000.0 (280.46061837504 deg) */
#define OMEGA_EARTH 7.2921158553e-5
#define THETA_J2000 4.894961212823756
/* obliquity of the ecliptic at J2000 (IAU 1976: 84381.448 arcsec) */
#define OBLIQUITY 0.40909280422232897
#define SEC_PER_CENTURY 3155760000.0
#define DEG (3.14159265358979323846 / 180.0)
#define ARCSEC (DEG / 3600.0)
#define TWO_PI 6.28318530717958647692

/* Vallado (2013) Table 8-4: base altitude [km], nominal density [kg/m^3], scale height [km] */
static const double ATM_H0[28] = {0,   25,  30,  40,  50,  60,  70,  80,  90,  100, 110, 120, 130, 140,
                                  150, 180, 200, 250, 300, 350, 400, 450, 500, 600, 700, 800, 900, 1000};
static const double ATM_RHO[28] = {1.225,     3.899e-2,  1.774e-2,  3.972e-3,  1.057e-3,  3.206e-4,  8.770e-5,
                                   1.905e-5,  3.396e-6,  5.297e-7,  9.661e-8,  2.438e-8,  8.484e-9,  3.845e-9,
                                   2.070e-9,  5.464e-10, 2.789e-10, 7.248e-11, 2.418e-11, 9.518e-12, 3.725e-12,
                                   1.585e-12, 6.967e-13, 1.454e-13, 3.614e-14, 1.170e-14, 5.245e-15, 3.019e-15};
static const double ATM_H[28] = {7.249,  6.349,  6.682,  7.554,  8.382,  7.714,  6.549,  5.799,  5.382,  5.877,
                                 7.263,  9.473,  12.636, 16.149, 22.523, 29.740, 37.105, 45.546, 53.628, 53.298,
                                 58.515, 60.828, 63.822, 71.835, 88.667, 124.64, 181.05, 268.00};

/* uniform environment of one sub-step */
typedef struct {
  double ct, st;                /* cos/sin Earth rotation angle */
  double sun[3], moon[3];       /* positions [km] */
} or_env;

void or_env_at(double t, or_env *e);

#endif

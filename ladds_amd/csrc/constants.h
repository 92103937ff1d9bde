// constants.h -- physical constants of the force model (host and device).
//
// Earth gravity field: EGM96 GM and reference radius; J2, C22, S22 are the
// unnormalised values of the normalised EGM96 coefficients
// C20 = -4.84165371736e-4, C22 = 2.43914352398e-6, S22 = -1.40016683654e-6.
// GM = 398600.4418 km^3/s^2 is also what the 1-s Kepler vector of
// tests/testladds/SimulationTest.cpp:39-57 reproduces.
// Third-body GM: DE405.  Solar pressure at 1 AU: Montenbruck & Gill (2000) eq. 3.69.
// Earth rotation rate and GMST at J2000.0: Montenbruck & Gill (2000) Sec. 5.
#pragma once

namespace ladds {

constexpr double kMuEarth = 398600.4418;        // km^3/s^2
constexpr double kREarth = 6378.1363;           // km
constexpr double kJ2 = 1.0826266835531513e-3;
constexpr double kC22 = 1.574460374564035e-6;
constexpr double kS22 = -9.03803806638557e-7;
constexpr double kGmSun = 1.32712440018e11;     // km^3/s^2
constexpr double kGmMoon = 4902.800066;         // km^3/s^2
constexpr double kAu = 149597870.7;             // km
constexpr double kPSun = 4.56e-6;               // N/m^2 at 1 AU
constexpr double kOmegaEarth = 7.2921158553e-5; // rad/s
constexpr double kThetaJ2000 = 4.894961212823756;  // rad (280.46061837504 deg)
constexpr double kObliquity = 0.40909280422232897; // rad (IAU 1976: 84381.448 arcsec)
constexpr double kSecPerCentury = 3155760000.0;
constexpr double kPi = 3.14159265358979323846;

// Derived coefficients used by the kernels.
constexpr double kJ2Coef = 1.5 * kJ2 * kMuEarth * kREarth * kREarth;   // J2 gradient prefactor
constexpr double kC22Coef = 3.0 * kMuEarth * kREarth * kREarth * kC22; // U22 = k (x^2 - y^2) / r^5
constexpr double kS22Coef = 6.0 * kMuEarth * kREarth * kREarth * kS22; // U22 = k x y / r^5
constexpr double kSrpCoef = kPSun * 1.0e3 * kAu * kAu;                 // km/s^2 * km^2 per (km^2/kg)
constexpr double kDragCoef = -0.5e3;                                   // unit conversion of 1/2 bc rho v^2

// Exponential atmosphere, Vallado (2013) Table 8-4:
// base altitude [km], nominal density [kg/m^3], scale height [km].
constexpr int kAtmRows = 28;
constexpr double kAtmH0[kAtmRows] = {0,   25,  30,  40,  50,  60,  70,  80,  90,  100, 110, 120, 130, 140,
                                     150, 180, 200, 250, 300, 350, 400, 450, 500, 600, 700, 800, 900, 1000};
constexpr double kAtmRho[kAtmRows] = {
    1.225,     3.899e-2,  1.774e-2,  3.972e-3,  1.057e-3,  3.206e-4,  8.770e-5,  1.905e-5,  3.396e-6,  5.297e-7,
    9.661e-8,  2.438e-8,  8.484e-9,  3.845e-9,  2.070e-9,  5.464e-10, 2.789e-10, 7.248e-11, 2.418e-11, 9.518e-12,
    3.725e-12, 1.585e-12, 6.967e-13, 1.454e-13, 3.614e-14, 1.170e-14, 5.245e-15, 3.019e-15};
constexpr double kAtmScale[kAtmRows] = {7.249,  6.349,  6.682,  7.554,  8.382,  7.714,  6.549,
                                        5.799,  5.382,  5.877,  7.263,  9.473,  12.636, 16.149,
                                        22.523, 29.740, 37.105, 45.546, 53.628, 53.298, 58.515,
                                        60.828, 63.822, 71.835, 88.667, 124.64, 181.05, 268.00};

}  // namespace ladds

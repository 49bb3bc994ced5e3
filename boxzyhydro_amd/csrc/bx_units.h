// bx_units.h -- the code-unit scales of units.f90:10-17 (get_units), with the
// constants of parameters.f90:7-11 (gcgs, rgasCGS, pi), in the reference's
// arithmetic order so every scale is bit-identical to the reference's.
#pragma once

namespace bx {

struct Units {
    // declaration order of type units (derived_types.f90:13-16)
    double length, time, mass, density, rgas, eps, kelvin, vel;
};

inline Units get_units()
{
    const double gcgs = 6.67e-8, rgas_cgs = 8.254e7, pi = 3.1415926535897931;
    Units s;
    s.length = 1.496e13;
    s.time = (3600.0 * 365.25 * 24.0) / (2.0 * pi);
    s.density = 1.0 / (gcgs * (s.time * s.time));
    s.mass = s.density * ((s.length * s.length) * s.length);
    s.vel = s.length / s.time;
    s.eps = ((s.vel * s.vel) * s.mass) / ((s.length * s.length) * s.length);
    s.kelvin = 1.0;
    s.rgas = rgas_cgs / ((s.vel * s.vel) / s.kelvin);
    return s;
}

}  // namespace bx

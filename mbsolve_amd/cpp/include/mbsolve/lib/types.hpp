/*
 * Basic type and physical constants of the mbsolve API
 * (reference: mbsolve-lib/include/mbsolve/lib/types.hpp:36-73).  The
 * constants are the values the reference uses; FP64 parity depends on them.
 */
#ifndef MBSOLVE_LIB_TYPES_H
#define MBSOLVE_LIB_TYPES_H

#include <cmath>
#include <complex>
#include <string>

namespace mbsolve {

/** Floating-point type of the API (double precision). */
typedef double real;

/** Reduced Planck constant [J s]. */
static const real HBAR = 1.05457266e-34;

/** pi, computed as 4 atan(1). */
static const real PI = std::atan(1.0) * 4.0;

/** Vacuum permeability [H/m]. */
static const real MU0 = PI * 4e-7;

/** Vacuum permittivity [F/m]. */
static const real EPS0 = 8.854187817e-12;

/** Elementary charge [C]. */
static const real E0 = 1.60217733e-19;

/** Boltzmann constant [J/K]. */
static const real KB = 1.38064852e-23;
}

#endif

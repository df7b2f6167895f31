/* Test instrument for the reverse pass: the frozen-level flux of one electron
 * through a 1-D z grid map, written once over a floating type and built in
 * double (fdref_flux_d) and in long double (fdref_flux_ld).
 *
 * The steps are those of oracle.c (orc_trajectory_rk4 with the acceleration
 * stored one step late, orc_spline with end slopes equal to the end secants,
 * orc_level_samples at memory spacing, orc_flux summing levels 0..L without a
 * level test) for ctstart == T0.  The double build is checked against the
 * oracle; the long-double build gives central differences without the FP64
 * rounding noise of the 1e11-rad phase (64-bit instead of 53-bit significands),
 * which is what the adjoint's gradients are compared with at 1e-4.
 * Only tests import it.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define REAL double
#define FN(x) x##_d
#define SIN sin
#define COS cos
#define SQRT sqrt
#include "fdref_body.h"
#undef REAL
#undef FN
#undef SIN
#undef COS
#undef SQRT

#define REAL long double
#define FN(x) x##_ld
#define SIN sinl
#define COS cosl
#define SQRT sqrtl
#include "fdref_body.h"

// Harness-side definition of the two gravity globals that the reference's
// physics launcher reads (declared extern in code/src/physics/constants.h:20-21).
// The reference defines them, zero-initialised, in code/src/ic/init_cond.cc:12-13,
// a file outside the right-hand-side path that cannot be compiled here (it
// includes <gsl/gsl_integration.h>).  Every configuration exercised here has
// zero gravity, which is the value the reference gets when no initial
// condition sets them.  TEST INFRASTRUCTURE ONLY.
namespace constants {
double GLOBAL_GX = 0.0;
double GLOBAL_GY = 0.0;
}  // namespace constants

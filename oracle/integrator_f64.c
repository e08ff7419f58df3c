/* integrator_f64.c -- oracle (test infrastructure only): the FP64 build of integrator.c. */
#define REAL double
#define SUFFIX(name) name##_f64
#define RSQRT sqrt
#define RATAN atan
#define RCOS cos
#define RSIN sin
#define RTAN tan
#define RFABS fabs
#define RFMAX fmax
#define RFMIN fmin
#define RFLOOR floor
#define RACOS acos
#define RATAN2 atan2
#define RNEXTAFTER nextafter
#include "integrator.c"

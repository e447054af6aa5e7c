/*
 * CPU oracle (test infrastructure only): instantiates eispack_impl.h for float
 * (suffix _s) and double (suffix _d).  Built into oracle/liboracle.so by
 * oracle/Makefile; loaded by tests/, bench.py's CPU-baseline leg and smoke().
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define REAL float
#define SFX _s
#define MSFX f
#include "eispack_impl.h"
#undef REAL
#undef SFX
#undef MSFX

#define REAL double
#define SFX _d
#define MSFX
#include "eispack_impl.h"
#undef REAL
#undef SFX
#undef MSFX

int oracle_max_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Oracle CPU FFT: float and double instantiations of fft_cpu_tmpl.h.
 * Test infrastructure only (see fft_cpu.h). */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "fft_cpu.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define T double
#define SUF _d
#include "fft_cpu_tmpl.h"
#undef T
#undef SUF

#define T float
#define SUF _f
#include "fft_cpu_tmpl.h"
#undef T
#undef SUF

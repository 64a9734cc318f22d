/*
 * CPU ORACLE — test infrastructure only.
 *
 * The CPU-baseline build: the same ref32 restatement (ausm_oracle.c) in the
 * reference's unfused 15-pass structure, built with -O3 and OpenMP over rows
 * so that bench.py can time it on the host cores as a reported baseline.
 */
#include "ausm_oracle.c"
#include <omp.h>

int base32_num_threads(void) { return omp_get_max_threads(); }
void base32_set_threads(int n) { omp_set_num_threads(n); }

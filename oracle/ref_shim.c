/*
 * ref_shim.c -- linked into oracle/_ref/libszp_ref.so beside the reference
 * sources so that a test can pin the reference's OpenMP team size T (the
 * stream format depends on it) in the same OpenMP runtime the reference uses.
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 */
#include <omp.h>

void szp_ref_set_threads(int n) { omp_set_num_threads(n); }
int szp_ref_max_threads(void) { return omp_get_max_threads(); }

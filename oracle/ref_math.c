/*
 * ref_math.c -- exports the reference's own header-only math functions
 * (math/spherical.h YlmR_l1..l6, math/legpoly.h legpoly) from a shared
 * library so that tests can pin the oracle's restatement against them.
 *
 * TEST INFRASTRUCTURE ONLY.  This file contains no reference code: it is a
 * harness that #includes the reference headers where they lie, compiled by
 * oracle/Makefile (`make ref`) into oracle/_ref/.  These two headers are the
 * only parts of the reference that build without FFTW3, which is absent here.
 */
#include <math.h>

#include "spherical.h"
#include "legpoly.h"

/* The headers use C99 `inline` definitions; these declarations emit the
 * external definitions in this translation unit. */
extern inline double YlmR_l1(const int, const double, const double,
    const double, const double);
extern inline double YlmR_l2(const int, const double, const double,
    const double, const double);
extern inline double YlmR_l3(const int, const double, const double,
    const double, const double);
extern inline double YlmR_l4(const int, const double, const double,
    const double, const double);
extern inline double YlmR_l5(const int, const double, const double,
    const double, const double);
extern inline double YlmR_l6(const int, const double, const double,
    const double, const double);
extern inline double legpoly(const int, const double);

double ref_ylm(int ell, int m, double ct, double st, double cp, double sp) {
  switch (ell) {
    case 1: return YlmR_l1(m, ct, st, cp, sp);
    case 2: return YlmR_l2(m, ct, st, cp, sp);
    case 3: return YlmR_l3(m, ct, st, cp, sp);
    case 4: return YlmR_l4(m, ct, st, cp, sp);
    case 5: return YlmR_l5(m, ct, st, cp, sp);
    case 6: return YlmR_l6(m, ct, st, cp, sp);
    default: return NAN;
  }
}

double ref_legpoly(int ell, double x) { return legpoly(ell, x); }

/*
 * CPU oracle for the L-BFGS-B hot path -- TEST INFRASTRUCTURE ONLY.
 * Instantiates oracle_body.h for double (_f64) and float (_f32).
 * Build with -O2 -ffp-contract=off (see Makefile) so that its rounding is
 * reproducible and matches the reference CPU path's plain C++ arithmetic.
 */
#include "lbfgsb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define R double
#define S(x) x##_f64
#define R_IS_F32 0
#include "oracle_body.h"
#undef R
#undef S
#undef R_IS_F32

#define R float
#define S(x) x##_f32
#define R_IS_F32 1
#include "oracle_body.h"
#undef R
#undef S
#undef R_IS_F32

// dogdefs.h -- the common include of FINESS app sources.
#ifndef FINESS_COMPAT_DOGDEFS_H
#define FINESS_COMPAT_DOGDEFS_H
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include "constants.h"
#include "dog_math.h"
#include "tensors.h"
#endif

// dog_math.h -- Max / Min of the FINESS app interface (lib/dog_math.h:18-22).
#ifndef FINESS_COMPAT_DOG_MATH_H
#define FINESS_COMPAT_DOG_MATH_H
inline double Max(double a, double b) { return a > b ? a : b; }
inline double Min(double a, double b) { return a < b ? a : b; }
inline int Max(int a, int b) { return a > b ? a : b; }
inline int Min(int a, int b) { return a < b ? a : b; }
#endif

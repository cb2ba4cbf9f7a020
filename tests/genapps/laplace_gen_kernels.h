// User kernels of the 2D Laplace Jacobi iteration, in the OPS dev API
// (the tutorial's step 7: apps/c/laplace2d_tutorial/step7/laplace_kernels.h).
#ifndef LAPLACE_GEN_KERNELS_H
#define LAPLACE_GEN_KERNELS_H

void apply_stencil(const ACC<double> &A, ACC<double> &Anew, double *error) {
  Anew(0, 0) = 0.25 * (A(1, 0) + A(-1, 0) + A(0, -1) + A(0, 1));
  *error = fmax(*error, fabs(Anew(0, 0) - A(0, 0)));
}

void copy(ACC<double> &A, const ACC<double> &Anew) {
  A(0, 0) = Anew(0, 0);
}

#endif

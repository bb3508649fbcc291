// ref_shim.cpp -- C entry points over the reference's own sources that
// compile without CUDA, OpenCV, Eigen or MPI, so that tests can check the
// oracle's restatements against the reference itself.
//
// TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Built by
// oracle/build_ref.py together with, from the reference tree (paths under
// src/backend/ of RandomWalk_NMR):
//   RNG/xorshift.cpp                 xorshift64 (xorshift.cpp:17-24)
//   Math/LeastSquareAdjust.cpp       the Stejskal-Tanner fit (LeastSquareAdjust.cpp:77-170,
//                                    used at NMR_pfgse.cpp:680-686)
//   Laplace/src/vector_funcs.cpp     accumulated linspace / logspace (vector_funcs.cpp:85-107)
// into oracle/_ref/libref.so.  Nothing here is copied: it calls those
// functions through their own headers.
#include <stdint.h>

#include <cmath>
#include <vector>

#include "LeastSquareAdjust.h"
#include "vector_funcs.h"
#include "xorshift.h"

extern "C" {

// n successive xorshift64 states from s (the reference's generator).
void ref_xorshift64_sequence(uint64_t s, int n, uint64_t *out)
{
    xorshift64_state st{s};
    for (int i = 0; i < n; ++i)
        out[i] = xorshift64(&st);
}

// The reference's linspace_double_vector / logspace_vector into out[size].
void ref_linspace(double start, double end, int size, int include_final, double *out)
{
    const std::vector<double> v = linspace_double_vector(start, end, size, include_final != 0);
    for (int i = 0; i < size; ++i)
        out[i] = v[i];
}

void ref_logspace(double start, double end, int size, int include_final, double *out)
{
    const std::vector<double> v = logspace_vector(start, end, size, include_final != 0);
    for (int i = 0; i < size; ++i)
        out[i] = v[i];
}

// The reference's D(Delta) fit as NMR_PFGSE::recoverDsatWithoutSampling
// runs it (NMR_pfgse.cpp:672-686): LeastSquareAdjust over the first `points`
// of rhs / lhs, setPoints(points), solve; D = B and its standard error
// sqrt(MSE points / (points - 1)).  Also returns A.
void ref_stejskal_tanner(const double *rhs, const double *lhs, int points, int intercept, double *D, double *D_err,
                         double *A)
{
    std::vector<double> x(rhs, rhs + points), y(lhs, lhs + points);
    LeastSquareAdjust lsa(x, y, intercept != 0);
    lsa.setPoints(points);
    lsa.solve();
    *D = lsa.getB();
    *D_err = sqrt(lsa.getMSE() * ((double)points / ((double)points - 1.0)));
    *A = lsa.getA();
}

}  // extern "C"

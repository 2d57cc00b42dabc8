// krylov.hpp -- FP64 flexible GMRES on device vectors
#pragma once
#include <functional>
#include <vector>

#include "mg.hpp"

namespace i2e {

// pinned host buffer (hipHostMalloc, never registered after the fact)
template <class T>
struct HBuf {
  T *p = nullptr;
  size_t n = 0;
  int alloc(size_t count, unsigned flags = hipHostMallocDefault) {
    if (count <= n && p) return I2E_OK;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
    HIP_OK(hipHostMalloc((void **)&p, count * sizeof(T), flags));
    n = count;
    return I2E_OK;
  }
};

struct Fgmres {
  using Op = std::function<int(const double *, double *)>;
  int n = 0, m = 0;
  size_t ld = 0;
  DBuf<double> V, Z, w, part, hd;
  HBuf<double> hh;
  HBuf<double> hc;       // coherent: the projection pass's last kernel stores its coefficients here
  // Optional split of each pending vector that a projection pass leaves
  // (the next preconditioner input): unknown k = 3 m + 1 into sx[m], 3 m + 2
  // into sy[m], as float (split_f32) or double.  split_of is the vector whose
  // split sx/sy hold; the preconditioner clears it when it takes the split.
  void *sx = nullptr, *sy = nullptr;
  bool split_f32 = false;
  const double *split_of = nullptr;
  KTimer *kt = nullptr;  // optional per-launch timing of the Gram-Schmidt kernels
  int reorth = 0;        // extra Gram-Schmidt passes in the last solve
  int orth = 0;          // ORTH_LAGGED or ORTH_DGKS
  enum { ORTH_LAGGED = 0, ORTH_DGKS = 1 };
  int init(int n, int m);
  int gs_pass(int nv, double *w, int off, hipStream_t s);
  int lagged_pass(int j, const double *w, double *wout, hipStream_t s);
  int norm(const double *x, double *out, hipStream_t s);
  // solves A x = b to ||b - A x|| <= rtol ||b||; x holds the initial guess
  int solve(const Op &A, const Op &M, const double *b, double *x, double rtol, int maxit, hipStream_t s, int *iters,
            double *relres);
  int solve_lagged(const Op &A, const Op &M, const double *b, double *x, double rtol, int maxit, hipStream_t s,
                   int *iters, double *relres);
  int solve_dgks(const Op &A, const Op &M, const double *b, double *x, double rtol, int maxit, hipStream_t s,
                 int *iters, double *relres);
  void release();
};

}  // namespace i2e

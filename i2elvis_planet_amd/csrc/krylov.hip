// krylov.hip -- right-preconditioned flexible GMRES in FP64 on the device.
//
// Replaces the direct PARDISO solves of gausmat4 (gausmart.c:174-318).  Each
// restart recomputes the true FP64 residual b - A x, which makes the restart
// loop an iterative refinement around whatever precision the preconditioner
// runs in.  Orthogonalisation is classical Gram-Schmidt applied twice (two
// fused multi-dot / multi-axpy passes over the basis); all reductions go
// through fixed-order block partials, so results are bitwise reproducible.
#include <cmath>
#include <cstdlib>

#include "krylov.hpp"

namespace i2e {

static constexpr int RB = 256;   // threads per block
static constexpr int RG = 480;   // reduction blocks

// part[i*RG + blk] = sum_k V[i][k] w[k] over this block's slice, i < nv
__global__ void k_multidot(int n, int nv, const double *__restrict__ V, size_t ld, const double *__restrict__ w,
                           double *part) {
  __shared__ double sh[RB];
  for (int i0 = 0; i0 < nv; i0 += 4) {
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    const int nb = nv - i0 < 4 ? nv - i0 : 4;
    for (int k = blockIdx.x * RB + threadIdx.x; k < n; k += RG * RB) {
      double wk = w[k];
      a0 += V[(size_t)i0 * ld + k] * wk;
      if (nb > 1) a1 += V[(size_t)(i0 + 1) * ld + k] * wk;
      if (nb > 2) a2 += V[(size_t)(i0 + 2) * ld + k] * wk;
      if (nb > 3) a3 += V[(size_t)(i0 + 3) * ld + k] * wk;
    }
    double acc[4] = {a0, a1, a2, a3};
    for (int q = 0; q < nb; q++) {
      sh[threadIdx.x] = acc[q];
      __syncthreads();
      for (int o = RB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
      }
      if (threadIdx.x == 0) part[(size_t)(i0 + q) * RG + blockIdx.x] = sh[0];
      __syncthreads();
    }
  }
}

// out[i] = sum_blk part[i*RG+blk] (fixed order); if sq, out[i] = sqrt(out[i])
__global__ void k_reduce_parts(int nv, const double *part, double *out, int sq) {
  __shared__ double sh[RB];
  const int i = blockIdx.x;
  if (i >= nv) return;
  double s = 0;
  for (int b = threadIdx.x; b < RG; b += RB) s += part[(size_t)i * RG + b];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = RB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[i] = sq ? sqrt(sh[0]) : sh[0];
}

// w += sgn * sum_i h[i] V[i]
__global__ void k_multiaxpy(int n, int nv, const double *__restrict__ V, size_t ld, const double *__restrict__ h,
                            double sgn, double *__restrict__ w) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    double s = 0;
    for (int i = 0; i < nv; i++) s += h[i] * V[(size_t)i * ld + k];
    w[k] += sgn * s;
  }
}

// v = w / *nrm
__global__ void k_scale_by(int n, const double *__restrict__ w, const double *nrm, double *__restrict__ v) {
  const double inv = 1.0 / nrm[0];
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) v[k] = w[k] * inv;
}

// r = b - r
__global__ void k_bminus(int n, const double *__restrict__ b, double *__restrict__ r) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) r[k] = b[k] - r[k];
}

int Fgmres::init(int n_, int m_) {
  if (n == n_ && m == m_ && V.p) return I2E_OK;
  n = n_;
  m = m_;
  ld = ((size_t)n + 63) & ~(size_t)63;
  ST_OK(V.alloc(ld * (m + 1)));
  ST_OK(Z.alloc(ld * m));
  ST_OK(w.alloc(ld));
  ST_OK(part.alloc((size_t)RG * (m + 2)));
  ST_OK(hd.alloc(4 * (m + 2)));
  ST_OK(hh.alloc(4 * (m + 2)));
  return I2E_OK;
}

int Fgmres::norm(const double *x, double *out_dev, hipStream_t s) {
  hipLaunchKernelGGL(k_multidot, dim3(RG), dim3(RB), 0, s, n, 1, x, ld, x, part.p);
  hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(RB), 0, s, 1, part.p, out_dev, 1);
  return I2E_OK;
}

int Fgmres::solve(const Op &A, const Op &M, const double *b, double *x, double rtol, int maxit, hipStream_t s,
                  int *iters, double *relres) {
  const unsigned gb = 1024;
  double bn;
  ST_OK(norm(b, hd.p, s));
  HIP_OK(hipMemcpyAsync(hh.p, hd.p, sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  bn = hh.p[0];
  *iters = 0;
  *relres = 0;
  if (bn == 0) {
    HIP_OK(hipMemsetAsync(x, 0, sizeof(double) * n, s));
    return I2E_OK;
  }
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
  const bool verbose = std::getenv("I2E_VERBOSE") != nullptr;
  int its = 0;
  double rr = 1;
  while (true) {
    // true residual
    ST_OK(A(x, w.p));
    hipLaunchKernelGGL(k_bminus, dim3(gb), dim3(256), 0, s, n, b, w.p);
    ST_OK(norm(w.p, hd.p, s));
    HIP_OK(hipMemcpyAsync(hh.p, hd.p, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    double beta = hh.p[0];
    rr = beta / bn;
    if (verbose) fprintf(stderr, "[fgmres n=%d] its %d true relres %.3e (||b|| %.3e)\n", n, its, rr, bn);
    if (rr <= rtol || its >= maxit || !(beta > 0)) break;
    hipLaunchKernelGGL(k_scale_by, dim3(gb), dim3(256), 0, s, n, w.p, hd.p, V.p);
    std::fill(H.begin(), H.end(), 0.0);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k = 0;
    for (int j = 0; j < m && its < maxit; j++) {
      double *vj = V.p + (size_t)j * ld, *zj = Z.p + (size_t)j * ld;
      ST_OK(M(vj, zj));
      ST_OK(A(zj, w.p));
      // CGS2: two passes, h accumulated on the host from one packed copy
      for (int pass = 0; pass < 2; pass++) {
        double *hp = hd.p + (size_t)pass * (m + 2);
        hipLaunchKernelGGL(k_multidot, dim3(RG), dim3(RB), 0, s, n, j + 1, V.p, ld, w.p, part.p);
        hipLaunchKernelGGL(k_reduce_parts, dim3(j + 1), dim3(RB), 0, s, j + 1, part.p, hp, 0);
        hipLaunchKernelGGL(k_multiaxpy, dim3(gb), dim3(256), 0, s, n, j + 1, V.p, ld, hp, -1.0, w.p);
      }
      double *nrm = hd.p + 2 * (m + 2);
      ST_OK(norm(w.p, nrm, s));
      hipLaunchKernelGGL(k_scale_by, dim3(gb), dim3(256), 0, s, n, w.p, nrm, V.p + (size_t)(j + 1) * ld);
      HIP_OK(hipMemcpyAsync(hh.p, hd.p, sizeof(double) * 3 * (m + 2), hipMemcpyDeviceToHost, s));
      HIP_OK(hipStreamSynchronize(s));
      for (int i = 0; i <= j; i++) H[(size_t)i * m + j] = hh.p[i] + hh.p[(m + 2) + i];
      double hn = hh.p[2 * (m + 2)];
      H[(size_t)(j + 1) * m + j] = hn;
      for (int i = 0; i < j; i++) {
        double a = H[(size_t)i * m + j], c = H[(size_t)(i + 1) * m + j];
        H[(size_t)i * m + j] = cs[i] * a + sn[i] * c;
        H[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * c;
      }
      double a = H[(size_t)j * m + j], c = H[(size_t)(j + 1) * m + j];
      double den = std::hypot(a, c);
      cs[j] = den > 0 ? a / den : 1.0;
      sn[j] = den > 0 ? c / den : 0.0;
      H[(size_t)j * m + j] = den;
      H[(size_t)(j + 1) * m + j] = 0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      its++;
      k = j + 1;
      if (verbose && (j % 10 == 9)) fprintf(stderr, "[fgmres n=%d]   it %d est %.3e\n", n, its, std::fabs(g[j + 1]) / bn);
      if (std::fabs(g[j + 1]) / bn <= rtol || !(hn > 0)) break;
    }
    // back substitution and x += Z y
    for (int i = k - 1; i >= 0; i--) {
      double t = g[i];
      for (int q = i + 1; q < k; q++) t -= H[(size_t)i * m + q] * y[q];
      y[i] = H[(size_t)i * m + i] != 0 ? t / H[(size_t)i * m + i] : 0.0;
    }
    for (int i = 0; i < k; i++) hh.p[i] = y[i];
    HIP_OK(hipMemcpyAsync(hd.p, hh.p, sizeof(double) * k, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_multiaxpy, dim3(gb), dim3(256), 0, s, n, k, Z.p, ld, hd.p, 1.0, x);
    HIP_OK(hipStreamSynchronize(s));
  }
  *iters = its;
  *relres = rr;
  return (rr <= rtol) ? I2E_OK : I2E_ERR_NOCONV;
}

void Fgmres::release() {
  V.release();
  Z.release();
  w.release();
  part.release();
  hd.release();
  if (hh.p) hipHostFree(hh.p);
  hh.p = nullptr;
  hh.n = 0;
}

}  // namespace i2e

// krylov.hip -- right-preconditioned flexible GMRES in FP64 on the device.
//
// Replaces the direct PARDISO solves of gausmat4 (gausmart.c:174-318).  Each
// restart recomputes the true FP64 residual b - A x, which makes the restart
// loop an iterative refinement around whatever precision the preconditioner
// runs in.  All reductions go through fixed-order block partials, so results are bitwise reproducible.
// Orthogonalisation is classical Gram-Schmidt (one fused projection pass and
// one fused update pass, each a single read of the basis), repeated when the
// first pass cancelled most of the vector.
#include <cmath>
#include <cstdlib>

#include "krylov.hpp"

namespace i2e {

static constexpr int RB = 256;        // threads per block
static constexpr int EPT = 4;        // vector elements per thread in the Gram-Schmidt kernels
static constexpr int CH = RB * EPT;  // vector elements per block

static inline int gs_blocks(int n) { return (n + CH - 1) / CH; }

// Classical Gram-Schmidt projection coefficients in one pass over the basis:
// part[i*G + blk] = <V_i, w> over this block's chunk (i < nv) and
// part[nv*G + blk] = <w, w>.  w stays in registers; each wavefront reduces
// with cross-lane butterflies, the block's four wavefronts meet in LDS once.
__global__ __launch_bounds__(RB) void k_gsdot(int n, int nv, const double *__restrict__ V, size_t ld,
                                              const double *__restrict__ w, double *__restrict__ part) {
  extern __shared__ double sh[];  // [4][nv + 1]
  const int G = gridDim.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6, S = nv + 1;
  const size_t base = (size_t)blockIdx.x * CH + threadIdx.x;
  double we[EPT];
#pragma unroll
  for (int e = 0; e < EPT; e++) {
    const size_t k = base + (size_t)e * RB;
    we[e] = k < (size_t)n ? w[k] : 0.0;
  }
  // four basis vectors per round keep 16 loads in flight per lane
  int i = 0;
  for (; i + 4 <= nv; i += 4) {
    double a[4] = {0, 0, 0, 0};
#pragma unroll
    for (int e = 0; e < EPT; e++) {
      const size_t k = base + (size_t)e * RB;
      if (k < (size_t)n) {
#pragma unroll
        for (int q = 0; q < 4; q++) a[q] += V[(size_t)(i + q) * ld + k] * we[e];
      }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const double t = wave_sum(a[q]);
      if (lane == 0) sh[wv * S + i + q] = t;
    }
  }
  for (; i < nv; i++) {
    const double *v = V + (size_t)i * ld;
    double a = 0;
#pragma unroll
    for (int e = 0; e < EPT; e++) {
      const size_t k = base + (size_t)e * RB;
      if (k < (size_t)n) a += v[k] * we[e];
    }
    a = wave_sum(a);
    if (lane == 0) sh[wv * S + i] = a;
  }
  double a = 0;
#pragma unroll
  for (int e = 0; e < EPT; e++) a += we[e] * we[e];
  a = wave_sum(a);
  if (lane == 0) sh[wv * S + nv] = a;
  __syncthreads();
  for (int i = threadIdx.x; i < S; i += RB)
    part[(size_t)i * G + blockIdx.x] = ((sh[i] + sh[S + i]) + sh[2 * S + i]) + sh[3 * S + i];
}

// w -= sum_i h[i] V_i (i < nv), and part[blk] = <w, w> of the updated w over
// this block's chunk
__global__ __launch_bounds__(RB) void k_gsaxpy(int n, int nv, const double *__restrict__ V, size_t ld,
                                               const double *__restrict__ h, double *__restrict__ w,
                                               double *__restrict__ part) {
  __shared__ double sh[RB / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const size_t base = (size_t)blockIdx.x * CH + threadIdx.x;
  double we[EPT];
#pragma unroll
  for (int e = 0; e < EPT; e++) {
    const size_t k = base + (size_t)e * RB;
    we[e] = k < (size_t)n ? w[k] : 0.0;
  }
  int i = 0;
  for (; i + 4 <= nv; i += 4) {
    const double h0 = h[i], h1 = h[i + 1], h2 = h[i + 2], h3 = h[i + 3];
#pragma unroll
    for (int e = 0; e < EPT; e++) {
      const size_t k = base + (size_t)e * RB;
      if (k < (size_t)n) {
        const double v0 = V[(size_t)i * ld + k], v1 = V[(size_t)(i + 1) * ld + k], v2 = V[(size_t)(i + 2) * ld + k],
                     v3 = V[(size_t)(i + 3) * ld + k];
        we[e] = (((we[e] - h0 * v0) - h1 * v1) - h2 * v2) - h3 * v3;
      }
    }
  }
  for (; i < nv; i++) {
    const double hi = h[i];
    const double *v = V + (size_t)i * ld;
#pragma unroll
    for (int e = 0; e < EPT; e++) {
      const size_t k = base + (size_t)e * RB;
      if (k < (size_t)n) we[e] -= hi * v[k];
    }
  }
  double a = 0;
#pragma unroll
  for (int e = 0; e < EPT; e++) {
    const size_t k = base + (size_t)e * RB;
    if (k < (size_t)n) {
      w[k] = we[e];
      a += we[e] * we[e];
    }
  }
  a = wave_sum(a);
  if (lane == 0) sh[wv] = a;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// out[i] = sum_blk part[i*G + blk] in a fixed order, i < nv (one block per i)
__global__ void k_reduce_parts(int nv, int G, const double *part, double *out) {
  __shared__ double sh[RB];
  const int i = blockIdx.x;
  if (i >= nv) return;
  double s = 0;
  for (int b = threadIdx.x; b < G; b += RB) s += part[(size_t)i * G + b];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = RB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[i] = sh[0];
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

// v = w * f
__global__ void k_scale(int n, const double *__restrict__ w, double f, double *__restrict__ v) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) v[k] = w[k] * f;
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
  ST_OK(part.alloc((size_t)gs_blocks(n) * (m + 2)));
  ST_OK(hd.alloc(4 * (m + 2)));
  ST_OK(hh.alloc(4 * (m + 2)));
  return I2E_OK;
}

// one Gram-Schmidt pass of w against V[0..nv): h = V^T w and <w,w> into
// hd[off .. off+nv], w -= V h, then <w,w> of the result into hd[off+nv+1]
int Fgmres::gs_pass(int nv, double *w_, int off, hipStream_t s) {
  const int G = gs_blocks(n);
  const double bytes_dot = 8.0 * ((double)(nv + 1) * n + (double)(nv + 1) * G);
  const double bytes_axpy = 8.0 * ((double)nv * n + 2.0 * n + G);
  if (kt) kt->begin(K_GSDOT, s);
  hipLaunchKernelGGL(k_gsdot, dim3(G), dim3(RB), sizeof(double) * 4 * (nv + 1), s, n, nv, V.p, ld, w_, part.p);
  if (kt) kt->end(K_GSDOT, s, bytes_dot);
  hipLaunchKernelGGL(k_reduce_parts, dim3(nv + 1), dim3(RB), 0, s, nv + 1, G, part.p, hd.p + off);
  if (nv > 0) {
    if (kt) kt->begin(K_GSAXPY, s);
    hipLaunchKernelGGL(k_gsaxpy, dim3(G), dim3(RB), 0, s, n, nv, V.p, ld, hd.p + off, w_, part.p);
    if (kt) kt->end(K_GSAXPY, s, bytes_axpy);
    hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(RB), 0, s, 1, G, part.p, hd.p + off + nv + 1);
  }
  return I2E_OK;
}

int Fgmres::norm(const double *x, double *out, hipStream_t s) {
  ST_OK(gs_pass(0, const_cast<double *>(x), 0, s));
  HIP_OK(hipMemcpyAsync(hh.p, hd.p, sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  *out = std::sqrt(hh.p[0]);
  return I2E_OK;
}

int Fgmres::solve(const Op &A, const Op &M, const double *b, double *x, double rtol, int maxit, hipStream_t s,
                  int *iters, double *relres) {
  const unsigned gb = 1024;
  double bn;
  ST_OK(norm(b, &bn, s));
  *iters = 0;
  *relres = 0;
  reorth = 0;
  if (bn == 0) {
    HIP_OK(hipMemsetAsync(x, 0, sizeof(double) * n, s));
    return I2E_OK;
  }
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
  const bool verbose = std::getenv("I2E_VERBOSE") != nullptr;
  const size_t P2 = (size_t)(m + 2);  // offset of the second pass's coefficients in hd
  int its = 0;
  double rr = 1;
  while (true) {
    // true residual
    ST_OK(A(x, w.p));
    hipLaunchKernelGGL(k_bminus, dim3(gb), dim3(256), 0, s, n, b, w.p);
    double beta;
    ST_OK(norm(w.p, &beta, s));
    rr = beta / bn;
    if (verbose) fprintf(stderr, "[fgmres n=%d] its %d true relres %.3e (||b|| %.3e)\n", n, its, rr, bn);
    if (rr <= rtol || its >= maxit || !(beta > 0)) break;
    hipLaunchKernelGGL(k_scale, dim3(gb), dim3(256), 0, s, n, w.p, 1.0 / beta, V.p);
    std::fill(H.begin(), H.end(), 0.0);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k = 0;
    for (int j = 0; j < m && its < maxit; j++) {
      double *vj = V.p + (size_t)j * ld, *zj = Z.p + (size_t)j * ld;
      ST_OK(M(vj, zj));
      ST_OK(A(zj, w.p));
      // classical Gram-Schmidt, repeated only when the projection removed
      // most of w (||w'|| < ||w|| / sqrt 2: Daniel-Gragg-Kaufman-Stewart)
      ST_OK(gs_pass(j + 1, w.p, 0, s));
      HIP_OK(hipMemcpyAsync(hh.p, hd.p, sizeof(double) * (j + 3), hipMemcpyDeviceToHost, s));
      HIP_OK(hipStreamSynchronize(s));
      for (int i = 0; i <= j; i++) H[(size_t)i * m + j] = hh.p[i];
      double ww = hh.p[j + 1], nn = hh.p[j + 2];
      if (!(nn >= 0.5 * ww)) {
        ST_OK(gs_pass(j + 1, w.p, (int)P2, s));
        HIP_OK(hipMemcpyAsync(hh.p + P2, hd.p + P2, sizeof(double) * (j + 3), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        for (int i = 0; i <= j; i++) H[(size_t)i * m + j] += hh.p[P2 + i];
        nn = hh.p[P2 + j + 2];
        reorth++;
      }
      const double hn = std::sqrt(nn > 0 ? nn : 0.0);
      if (hn > 0) hipLaunchKernelGGL(k_scale, dim3(gb), dim3(256), 0, s, n, w.p, 1.0 / hn, V.p + (size_t)(j + 1) * ld);
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
      if (verbose && (j % 10 == 9))
        fprintf(stderr, "[fgmres n=%d]   it %d est %.3e reorth %d\n", n, its, std::fabs(g[j + 1]) / bn, reorth);
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

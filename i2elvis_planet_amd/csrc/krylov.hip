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

static constexpr int RB = 128;        // threads per block (2 wavefronts: measured best for the basis passes)
static constexpr int NW = RB / 64;    // wavefronts per block
static constexpr int EPT = 4;        // vector elements per thread in the Gram-Schmidt kernels
static constexpr int CH = RB * EPT;  // vector elements per block

static inline int gs_blocks(int n) { return (n + CH - 1) / CH; }

// Classical Gram-Schmidt projection coefficients in one pass over the basis:
// part[i*G + blk] = <V_i, w> over this block's chunk (i < nv) and
// part[nv*G + blk] = <w, w>.  w stays in registers; each wavefront reduces
// with cross-lane butterflies, the block's four wavefronts meet in LDS once.
__global__ __launch_bounds__(RB) void k_gsdot(int n, int nv, const double *__restrict__ V, size_t ld,
                                              const double *__restrict__ w, double *__restrict__ part) {
  extern __shared__ double sh[];  // [NW][nv + 1]
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
  for (int i = threadIdx.x; i < S; i += RB) {
    double t = 0;
#pragma unroll
    for (int k = 0; k < NW; k++) t += sh[k * S + i];
    part[(size_t)i * G + blockIdx.x] = t;
  }
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
  if (threadIdx.x == 0) {
    double t = 0;
#pragma unroll
    for (int k = 0; k < NW; k++) t += sh[k];
    part[blockIdx.x] = t;
  }
}

// Lagged CGS2 (the second projection of each basis vector is fused into the
// next iteration's pass, normalisation by the Pythagorean identity; after
// Swirydowicz, Langou, Ananthan, Yang, Thomas, "Low synchronization
// Gram-Schmidt and GMRES algorithms", 2020).  One pass forms, per block,
//   a_i = <V_i, u>, b_i = <V_i, w> (i < nv), <u,u>, <u,w>, <w,w>
// into part[q*G + blk], q = 0..2nv+2 (a, then b, then the three scalars).
__global__ __launch_bounds__(RB) void k_gsdot2(int n, int nv, const double *__restrict__ V, size_t ld,
                                               const double *__restrict__ u, const double *__restrict__ w,
                                               double *__restrict__ part) {
  extern __shared__ double sh[];  // [NW][2nv + 3]
  const int G = gridDim.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6, S = 2 * nv + 3;
  const size_t base = (size_t)blockIdx.x * CH + threadIdx.x;
  double ue[EPT], we[EPT];
#pragma unroll
  for (int e = 0; e < EPT; e++) {
    const size_t k = base + (size_t)e * RB;
    ue[e] = k < (size_t)n ? u[k] : 0.0;
    we[e] = k < (size_t)n ? w[k] : 0.0;
  }
  // four basis vectors per round: 16 loads in flight per lane
  int i = 0;
  for (; i + 4 <= nv; i += 4) {
    double a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
#pragma unroll
    for (int e = 0; e < EPT; e++) {
      const size_t k = base + (size_t)e * RB;
      if (k < (size_t)n) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const double v = V[(size_t)(i + q) * ld + k];
          a[q] += v * ue[e];
          b[q] += v * we[e];
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const double ta = wave_sum(a[q]), tb = wave_sum(b[q]);
      if (lane == 0) {
        sh[wv * S + i + q] = ta;
        sh[wv * S + nv + i + q] = tb;
      }
    }
  }
  for (; i < nv; i++) {
    double a0 = 0, b0 = 0;
#pragma unroll
    for (int e = 0; e < EPT; e++) {
      const size_t k = base + (size_t)e * RB;
      if (k < (size_t)n) {
        const double v0 = V[(size_t)i * ld + k];
        a0 += v0 * ue[e];
        b0 += v0 * we[e];
      }
    }
    a0 = wave_sum(a0);
    b0 = wave_sum(b0);
    if (lane == 0) {
      sh[wv * S + i] = a0;
      sh[wv * S + nv + i] = b0;
    }
  }
  double uu = 0, uw = 0, ww = 0;
#pragma unroll
  for (int e = 0; e < EPT; e++) {
    uu += ue[e] * ue[e];
    uw += ue[e] * we[e];
    ww += we[e] * we[e];
  }
  uu = wave_sum(uu);
  uw = wave_sum(uw);
  ww = wave_sum(ww);
  if (lane == 0) {
    sh[wv * S + 2 * nv] = uu;
    sh[wv * S + 2 * nv + 1] = uw;
    sh[wv * S + 2 * nv + 2] = ww;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < S; q += RB) {
    double t = 0;
#pragma unroll
    for (int k = 0; k < NW; k++) t += sh[k * S + q];
    part[(size_t)q * G + blockIdx.x] = t;
  }
}

// Fixed-order sums of the projection pass's block partials, one block per
// quantity: r[q] = sum_blk part[q*G + blk], q < 2nv+3, r = [a(nv), b(nv), uu,
// uw, ww].  The update kernel forms its coefficients from r (no grid-wide
// completion count here: a device-scope fence per block writes back the L2).
__global__ __launch_bounds__(RB) void k_reduce_coef(int G, const double *__restrict__ part, double *__restrict__ r) {
  __shared__ double sa[RB];
  const int q = blockIdx.x;
  double sum = 0;
  for (int b = threadIdx.x; b < G; b += RB) sum += part[(size_t)q * G + b];
  sa[threadIdx.x] = sum;
  __syncthreads();
  for (int o = RB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sa[threadIdx.x] += sa[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) r[q] = sa[0];
}

// Fused update.  Every block first forms the coefficients from the summed
// quantities r = [a(nv), b(nv), uu, uw, ww] (fixed order, so all blocks agree):
//   beta = sqrt(uu - a.a)       (norm of u after its second projection)
//   h    = (uw - a.b) / beta    (<v, w> for the finished v = (u - V a) / beta)
//   c    = b - a h / beta       (so that w - V c - (h/beta) u = w - V b - h v)
// and block 0 stores r[2nv+3] = beta, r[2nv+4] = h (beta = 0 flags a
// breakdown: u in the span of V).  Then vout = (u - V a) / beta (the finished
// basis vector), wout = w - V c - (h/beta) u (w projected on the finished
// basis, once), and part[blk] = <wout, wout> over this block's chunk.
// vout may alias u; wout may alias w.
__global__ __launch_bounds__(RB) void k_gsaxpy2(int n, int nv, const double *__restrict__ V, size_t ld, double *r,
                                                const double *u, const double *w, double *vout, double *wout,
                                                double *__restrict__ part) {
  __shared__ double sa[RB], sb[RB];
  {
    double aa = 0, ab = 0;
    for (int i = threadIdx.x; i < nv; i += RB) {
      aa += r[i] * r[i];
      ab += r[i] * r[nv + i];
    }
    sa[threadIdx.x] = aa;
    sb[threadIdx.x] = ab;
    __syncthreads();
    for (int o = RB / 2; o > 0; o >>= 1) {
      if (threadIdx.x < o) {
        sa[threadIdx.x] += sa[threadIdx.x + o];
        sb[threadIdx.x] += sb[threadIdx.x + o];
      }
      __syncthreads();
    }
  }
  const double b2 = r[2 * nv] - sa[0];
  const double beta = b2 > 0 ? sqrt(b2) : 0.0;
  const double hh = beta > 0 ? (r[2 * nv + 1] - sb[0]) / beta : 0.0;
  const double t = beta > 0 ? hh / beta : 0.0, ib = beta > 0 ? 1.0 / beta : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    r[2 * nv + 3] = beta;
    r[2 * nv + 4] = hh;
  }
  __shared__ double sh[RB / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const size_t base = (size_t)blockIdx.x * CH + threadIdx.x;
  double s1[EPT], s2[EPT];
#pragma unroll
  for (int e = 0; e < EPT; e++) s1[e] = s2[e] = 0.0;
  int i = 0;
  for (; i + 4 <= nv; i += 4) {
    double a[4], c[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      a[q] = r[i + q];
      c[q] = r[nv + i + q] - r[i + q] * t;
    }
#pragma unroll
    for (int e = 0; e < EPT; e++) {
      const size_t k = base + (size_t)e * RB;
      if (k < (size_t)n) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const double v = V[(size_t)(i + q) * ld + k];
          s1[e] += a[q] * v;
          s2[e] += c[q] * v;
        }
      }
    }
  }
  for (; i < nv; i++) {
    const double a0 = r[i], c0 = r[nv + i] - r[i] * t;
#pragma unroll
    for (int e = 0; e < EPT; e++) {
      const size_t k = base + (size_t)e * RB;
      if (k < (size_t)n) {
        const double v0 = V[(size_t)i * ld + k];
        s1[e] += a0 * v0;
        s2[e] += c0 * v0;
      }
    }
  }
  double acc = 0;
#pragma unroll
  for (int e = 0; e < EPT; e++) {
    const size_t k = base + (size_t)e * RB;
    if (k < (size_t)n) {
      const double uk = u[k], wk = w[k];
      const double wn = (wk - s2[e]) - t * uk;
      vout[k] = (uk - s1[e]) * ib;
      wout[k] = wn;
      acc += wn * wn;
    }
  }
  acc = wave_sum(acc);
  if (lane == 0) sh[wv] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
#pragma unroll
    for (int k = 0; k < NW; k++) t += sh[k];
    part[blockIdx.x] = t;
  }
}

// <w',w'> from the update pass's G block partials, summed by every block in
// the same fixed order (so all blocks agree), then v *= 1/sqrt(<w',w'>) over
// this block's chunk; block 0 stores the sum in *nrm2 and, with hout, the
// pass's nc coefficients coef[0..nc) and the sum into hout[0..nc] (host
// memory, read after the stream's synchronisation: no separate copy).  The
// scaling keeps the pending vector's size from compounding from step to
// step.  With sx, the scaled velocity unknowns 3m+1 and 3m+2 also go to
// sx[m] and sy[m] (the next preconditioner's right-hand side).
template <class S>
__global__ __launch_bounds__(RB) void k_norm_scale(int n, int G, const double *__restrict__ part,
                                                   double *__restrict__ nrm2, double *__restrict__ v, int nc,
                                                   const double *__restrict__ coef, double *__restrict__ hout,
                                                   S *__restrict__ sx, S *__restrict__ sy) {
  __shared__ double sh[RB];
  double sum = 0;
  for (int b = threadIdx.x; b < G; b += RB) sum += part[b];
  sh[threadIdx.x] = sum;
  __syncthreads();
  for (int o = RB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  const double q = sh[0];
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) *nrm2 = q;
    if (hout) {
      for (int t = threadIdx.x; t < nc; t += RB) hout[t] = coef[t];
      if (threadIdx.x == 0) hout[nc] = q;
    }
  }
  const double f = q > 0 ? 1.0 / sqrt(q) : 1.0;
  const size_t base = (size_t)blockIdx.x * CH + threadIdx.x;
#pragma unroll
  for (int e = 0; e < EPT; e++) {
    const size_t k = base + (size_t)e * RB;
    if (k < (size_t)n) {
      const double x = v[k] * f;
      v[k] = x;
      if (sx) {
        const size_t mm = k / 3, cmp = k - 3 * mm;
        if (cmp == 1) sx[mm] = S(x);
        if (cmp == 2) sy[mm] = S(x);
      }
    }
  }
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

// out = w / sqrt(*q) (w unchanged when *q is not positive)
__global__ void k_scale_by(int n, const double *__restrict__ q, const double *__restrict__ w,
                           double *__restrict__ out) {
  const double qq = *q;
  const double f = qq > 0 ? 1.0 / sqrt(qq) : 1.0;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) out[k] = w[k] * f;
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
  ST_OK(part.alloc((size_t)gs_blocks(n) * (2 * m + 4)));
  ST_OK(hd.alloc(8 * (m + 4)));
  ST_OK(hh.alloc(8 * (m + 4)));
  ST_OK(hc.alloc(8 * (m + 4), hipHostMallocCoherent));
  return I2E_OK;
}

// one Gram-Schmidt pass of w against V[0..nv): h = V^T w and <w,w> into
// hd[off .. off+nv], w -= V h, then <w,w> of the result into hd[off+nv+1]
int Fgmres::gs_pass(int nv, double *w_, int off, hipStream_t s) {
  const int G = gs_blocks(n);
  const double bytes_dot = 8.0 * ((double)(nv + 1) * n + (double)(nv + 1) * G);
  const double bytes_axpy = 8.0 * ((double)nv * n + 2.0 * n + G);
  KT_LAUNCH(kt, K_GSDOT, bytes_dot, bytes_dot, k_gsdot, dim3(G), dim3(RB), sizeof(double) * NW * (nv + 1), s, n,
            nv, V.p, ld, w_, part.p);
  hipLaunchKernelGGL(k_reduce_parts, dim3(nv + 1), dim3(RB), 0, s, nv + 1, G, part.p, hd.p + off);
  if (nv > 0) {
    KT_LAUNCH(kt, K_GSAXPY, bytes_axpy, bytes_axpy, k_gsaxpy, dim3(G), dim3(RB), 0, s, n, nv, V.p, ld,
              hd.p + off, w_, part.p);
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
  return orth == ORTH_DGKS ? solve_dgks(A, M, b, x, rtol, maxit, s, iters, relres)
                           : solve_lagged(A, M, b, x, rtol, maxit, s, iters, relres);
}

// One fused lagged-CGS2 step on the pending vector u (slot j of V, projected
// once, unit norm) and w = A z_j: finishes u into v_j against V[0..j), and
// projects w once on V[0..j]; w' / ||w'|| goes to wout.  After the caller's
// synchronisation the host holds hc = [a(j), b(j), uu, uw, ww, beta, h, <w',w'>].
int Fgmres::lagged_pass(int j, const double *w_, double *wout, hipStream_t s) {
  const int G = gs_blocks(n);
  double *u = V.p + (size_t)j * ld;
  const size_t R = 0;  // offset in hd
  const double bytes_dot = 8.0 * ((double)j * n + 2.0 * n + (2.0 * j + 3) * G);
  const double bytes_axpy = 8.0 * ((double)j * n + 4.0 * n + G);
  KT_LAUNCH(kt, K_GSDOT, bytes_dot, bytes_dot, k_gsdot2, dim3(G), dim3(RB), sizeof(double) * NW * (2 * j + 3), s,
            n, j, V.p, ld, u, w_, part.p);
  hipLaunchKernelGGL(k_reduce_coef, dim3(2 * j + 3), dim3(RB), 0, s, G, part.p, hd.p + R);
  KT_LAUNCH(kt, K_GSAXPY, bytes_axpy, bytes_axpy, k_gsaxpy2, dim3(G), dim3(RB), 0, s, n, j, V.p, ld,
            hd.p + R, u, w_, u, wout, part.p);
  if (split_f32 || !sx)
    hipLaunchKernelGGL(k_norm_scale<float>, dim3(G), dim3(RB), 0, s, n, G, part.p, hd.p + R + 2 * j + 5, wout,
                       2 * j + 5, hd.p + R, hc.p, (float *)sx, (float *)sy);
  else
    hipLaunchKernelGGL(k_norm_scale<double>, dim3(G), dim3(RB), 0, s, n, G, part.p, hd.p + R + 2 * j + 5, wout,
                       2 * j + 5, hd.p + R, hc.p, (double *)sx, (double *)sy);
  split_of = sx ? wout : nullptr;
  return I2E_OK;
}

// Right-preconditioned FGMRES with lagged CGS2.  Step j preconditions the
// pending vector u_j (w' of the previous step, projected once and scaled by
// 1/sigma_j = 1/||w'||), then one pass finishes u_j into v_j and projects
// w = A z_j once; column j-1 of H becomes final here (A z_{j-1} = V b +
// h v_{j-1} + sigma_j (beta v_j + V a)), column j one step later.  A step whose estimated residual meets the tolerance
// finishes its pending vector with one extra projection pass instead of a
// further preconditioner application.
int Fgmres::solve_lagged(const Op &A, const Op &M, const double *b, double *x, double rtol, int maxit,
                         hipStream_t s, int *iters, double *relres) {
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
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m), col(m + 1);
  const bool verbose = std::getenv("I2E_VERBOSE") != nullptr;
  int its = 0;
  double rr = 1;
  auto rotate_col = [&](double *c, int j) {  // previous rotations on column j (entries 0..j)
    for (int i = 0; i < j; i++) {
      const double a = c[i], d = c[i + 1];
      c[i] = cs[i] * a + sn[i] * d;
      c[i + 1] = -sn[i] * a + cs[i] * d;
    }
  };
  // finish the pending vector u (slot j+1) against V[0..j] without a new
  // step: a' and beta' for column j (into colj, entries 0..j+1)
  auto finish = [&](int j, double sigma, double *colj) -> int {
    const int nv = j + 1;
    const int G = gs_blocks(n);
    double *u = V.p + (size_t)(j + 1) * ld;
    hipLaunchKernelGGL(k_gsdot, dim3(G), dim3(RB), sizeof(double) * NW * (nv + 1), s, n, nv, V.p, ld, u, part.p);
    hipLaunchKernelGGL(k_reduce_parts, dim3(nv + 1), dim3(RB), 0, s, nv + 1, G, part.p, hd.p);
    HIP_OK(hipMemcpyAsync(hh.p, hd.p, sizeof(double) * (nv + 1), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    double aa = 0;
    for (int i = 0; i < nv; i++) {
      colj[i] += sigma * hh.p[i];
      aa += hh.p[i] * hh.p[i];
    }
    const double b2 = hh.p[nv] - aa;
    colj[j + 1] = sigma * (b2 > 0 ? std::sqrt(b2) : 0.0);
    return I2E_OK;
  };
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
    int k = 0;          // finished columns
    double sigma = 0;   // norm of the pending vector before its scaling
    for (int j = 0; j < m && its < maxit; j++) {
      double *vj = V.p + (size_t)j * ld, *zj = Z.p + (size_t)j * ld, *un = V.p + (size_t)(j + 1) * ld;
      ST_OK(M(vj, zj));
      ST_OK(A(zj, w.p));
      its++;
      if (j == 0) {
        // v_0 is finished: one projection of w, the result is pending u_1
        ST_OK(gs_pass(1, w.p, 0, s));
        const double *ww0 = hd.p + 2;
        hipLaunchKernelGGL(k_scale_by, dim3(gb), dim3(256), 0, s, n, ww0, w.p, un);
        HIP_OK(hipMemcpyAsync(hh.p, hd.p, sizeof(double) * 3, hipMemcpyDeviceToHost, s));
      } else {
        ST_OK(lagged_pass(j, w.p, un, s));
      }
      HIP_OK(hipStreamSynchronize(s));
      if (j == 0) {
        H[0] = hh.p[0];
        sigma = std::sqrt(hh.p[2] > 0 ? hh.p[2] : 0.0);
      } else {
        const double *a = hc.p, *bb = hc.p + j;
        const double bet = hc.p[2 * j + 3], h = hc.p[2 * j + 4], ww = hc.p[2 * j + 5];
        // column j-1 is final: A z_{j-1} = ... + sigma (beta v_j + V a)
        for (int i = 0; i < j; i++) H[(size_t)i * m + (j - 1)] += sigma * a[i];
        H[(size_t)j * m + (j - 1)] = sigma * bet;
        for (int i = 0; i <= j; i++) col[i] = H[(size_t)i * m + (j - 1)];
        rotate_col(col.data(), j - 1);
        {
          const double p = col[j - 1], q = col[j], den = std::hypot(p, q);
          cs[j - 1] = den > 0 ? p / den : 1.0;
          sn[j - 1] = den > 0 ? q / den : 0.0;
          col[j - 1] = den;
          col[j] = 0;
          g[j] = -sn[j - 1] * g[j - 1];
          g[j - 1] = cs[j - 1] * g[j - 1];
        }
        for (int i = 0; i <= j; i++) H[(size_t)i * m + (j - 1)] = col[i];
        k = j;
        // column j: V b + h v_j, pending part follows
        for (int i = 0; i < j; i++) H[(size_t)i * m + j] = bb[i];
        H[(size_t)j * m + j] = h;
        sigma = std::sqrt(ww > 0 ? ww : 0.0);
        if (!(bet > 0)) break;  // u_j was in the span of V: breakdown
        if (std::fabs(g[j]) / bn <= rtol) break;
      }
      // estimated residual with column j finished by the pending norm alone
      for (int i = 0; i <= j; i++) col[i] = H[(size_t)i * m + j];
      col[j + 1] = sigma;
      rotate_col(col.data(), j);
      const double est = std::fabs(col[j + 1] / std::hypot(col[j], col[j + 1]) * g[j]);
      if (verbose && (j % 10 == 9)) fprintf(stderr, "[fgmres n=%d]   it %d est %.3e\n", n, its, est / bn);
      const bool last = (j + 1 == m) || its >= maxit || !(sigma > 0);
      if (est / bn <= rtol || last) {
        // finish u_{j+1} with one projection pass and close column j exactly
        for (int i = 0; i <= j + 1; i++) col[i] = i <= j ? H[(size_t)i * m + j] : 0.0;
        if (sigma > 0) ST_OK(finish(j, sigma, col.data()));
        rotate_col(col.data(), j);
        const double p = col[j], q = col[j + 1], den = std::hypot(p, q);
        const double c0 = den > 0 ? p / den : 1.0, s0 = den > 0 ? q / den : 0.0;
        const double gj1 = -s0 * g[j];
        if (std::fabs(gj1) / bn <= rtol || last) {
          for (int i = 0; i <= j + 1; i++) H[(size_t)i * m + j] = col[i];
          H[(size_t)j * m + j] = den;
          H[(size_t)(j + 1) * m + j] = 0;
          cs[j] = c0;
          sn[j] = s0;
          g[j + 1] = gj1;
          g[j] = c0 * g[j];
          k = j + 1;
          break;
        }
        reorth++;  // a finishing pass that did not end the cycle
      }
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

int Fgmres::solve_dgks(const Op &A, const Op &M, const double *b, double *x, double rtol, int maxit,
                       hipStream_t s, int *iters, double *relres) {
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
  if (hc.p) hipHostFree(hc.p);
  hc.p = nullptr;
  hc.n = 0;
  split_of = nullptr;
}

}  // namespace i2e

// sort.hip -- cell sort of the markers (a counting sort on the cell key) and
// the inverse permutation for downloads.
//
// The reference never reorders its markers, and neither does the boundary:
// i2e_download returns them in the caller's order.  Inside a context the
// markers are kept grouped by cell, cells grouped in kTile x kTile tiles
// (tile_key of m1serch, m2serch, markmart.c:1051-1088; markers outside the
// box last), so that a wavefront's markers share a few cells -- coalesced
// marker reads, node gathers that hit the same lines, long same-cell runs in
// the scatter's segmented sums -- and a block's markers share one tile,
// whose node sums it can stage in LDS.  Markers drift by at most maxxystep
// cells per step, so the order decays slowly and a sort every few steps
// restores it.
#include "kernels.hpp"

namespace i2e {

// sort key of every marker and the per-key counts (cnt zeroed by the caller)
__global__ void k_mkey(Grid g, Markers mk, int64_t M, int C, int *key, int *cnt) {
  const int ntj = tiles_per_col(g.Y);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
    const float fx = mk.x[i], fy = mk.y[i];
    int k = C;
    if (fx >= 0 && fy >= 0 && (double)fx <= g.xsize && (double)fy <= g.ysize)
      k = tile_key(serch(g.gx, g.X, (double)fx), serch(g.gy, g.Y, (double)fy), ntj);
    key[i] = k;
    atomicAdd(&cnt[k], 1);
  }
}

// exclusive scan of n ints in three passes: SB-element blocks scanned in
// LDS with their totals kept, the totals scanned by one block, the block
// offsets added back

__global__ __launch_bounds__(256) void k_scan_blocks(int n, const int *in, int *out, int *bsum) {
  __shared__ int s[SB];
  const int base = blockIdx.x * SB;
  for (int k = threadIdx.x; k < SB; k += 256) s[k] = base + k < n ? in[base + k] : 0;
  __syncthreads();
  // each thread scans its 4 consecutive elements, then a Hillis-Steele scan
  // of the 256 thread totals
  int a[4], t = 0;
  for (int q = 0; q < 4; q++) {
    a[q] = t;
    t += s[4 * threadIdx.x + q];
  }
  __shared__ int ts[256];
  ts[threadIdx.x] = t;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const int v = threadIdx.x >= (unsigned)o ? ts[threadIdx.x - o] : 0;
    __syncthreads();
    ts[threadIdx.x] += v;
    __syncthreads();
  }
  const int excl = ts[threadIdx.x] - t;
  for (int q = 0; q < 4; q++) {
    const int k = base + 4 * threadIdx.x + q;
    if (k < n) out[k] = excl + a[q];
  }
  if (threadIdx.x == 255) bsum[blockIdx.x] = ts[255];
}

// one block: exclusive scan of the nb block totals, in chunks of 256
__global__ __launch_bounds__(256) void k_scan_sums(int nb, int *bsum) {
  __shared__ int ts[256];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int c0 = 0; c0 < nb; c0 += 256) {
    const int k = c0 + threadIdx.x;
    const int v = k < nb ? bsum[k] : 0;
    ts[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
      const int u = threadIdx.x >= (unsigned)o ? ts[threadIdx.x - o] : 0;
      __syncthreads();
      ts[threadIdx.x] += u;
      __syncthreads();
    }
    if (k < nb) bsum[k] = carry + ts[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 0) carry += ts[255];
    __syncthreads();
  }
}

__global__ void k_scan_add(int n, const int *bsum, int *out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] += bsum[k / SB];
}

// move every marker to its slot in key order (off: the exclusive scan of the
// counts, advanced atomically; the order inside one cell is the order the
// atomics happen to take).  orig2[slot] = the caller's index of the marker.
// slot0 receives the slot of the caller's marker 0.
__global__ void k_mscatter(Markers src, Markers dst, int64_t M, const int *key, int *off, const int *orig,
                           int *orig2, int *slot0) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
    const int d = atomicAdd(&off[key[i]], 1);
    dst.x[d] = src.x[i];
    dst.y[d] = src.y[i];
    dst.k[d] = src.k[i];
    dst.v[d] = src.v[i];
    dst.p[d] = src.p[i];
    dst.por[d] = src.por[i];
    dst.tmax[d] = src.tmax[i];
    dst.t[d] = src.t[i];
    dst.mg[d] = src.mg[i];
    dst.mgt[d] = src.mgt[i];
    const int o = orig ? orig[i] : (int)i;
    orig2[d] = o;
    if (o == 0) *slot0 = d;
  }
}

// back to the caller's order: dst[orig[i]] = src[i]
__global__ void k_munperm(Markers src, Markers dst, int64_t M, const int *orig) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
    const int j = orig[i];
    dst.x[j] = src.x[i];
    dst.y[j] = src.y[i];
    dst.k[j] = src.k[i];
    dst.v[j] = src.v[i];
    dst.p[j] = src.p[i];
    dst.por[j] = src.por[i];
    dst.tmax[j] = src.tmax[i];
    dst.t[j] = src.t[i];
    dst.mg[j] = src.mg[i];
    dst.mgt[j] = src.mgt[i];
  }
}

}  // namespace i2e

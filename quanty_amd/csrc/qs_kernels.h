// Device building blocks shared by the apply and Lanczos translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#ifndef QS_BLOCK
#define QS_BLOCK 256
#endif

// Term-value tables up to this size are staged in LDS by the assembled kernels
// (one complex value per term; the entry's sign bit picks +V or -V).
#define QS_COEF_LDS_BYTES (16 * 1024)

// One row of the assembled operator for KC interleaved vectors
// (entry j of vector v at x[j*ldx + v]). Lane `lane` of the slice's wavefront
// owns row r; the slice stores its entries column-major (entry e of the lane's
// row at sbase + 64*e + lane), so every load of the index stream is coalesced.
// Entry id t = 2*term + sign: with HALF the table holds V_term only and the
// sign is applied here, otherwise coef[t] = (-1)^sign V_term.
template <int KC, bool HALF = false>
__device__ __forceinline__ void spmv_row(int64_t r, bool active, uint64_t sbase, uint32_t slen,
                                         const uint16_t *__restrict__ rowcnt,
                                         const uint32_t *__restrict__ colv,
                                         const uint16_t *__restrict__ tidv,
                                         const double2 *__restrict__ coef,
                                         const double2 *__restrict__ diag,
                                         const double2 *__restrict__ x, int64_t ldx, int lane,
                                         double *ar, double *ai) {
#pragma unroll
  for (int v = 0; v < KC; ++v) {
    ar[v] = 0.0;
    ai[v] = 0.0;
  }
  uint32_t cnt = 0;
  if (active) {
    cnt = rowcnt[r];
    if (diag) {
      const double2 dg = diag[r];
#pragma unroll
      for (int v = 0; v < KC; ++v) {
        const double2 xv = x[r * ldx + v];
        ar[v] = dg.x * xv.x - dg.y * xv.y;
        ai[v] = dg.x * xv.y + dg.y * xv.x;
      }
    }
  }
  const uint32_t *cp = colv + sbase + lane;
  const uint16_t *tp = tidv + sbase + lane;
  // the index stream is read once per apply: non-temporal loads keep it from
  // evicting the gathered vector from L2 / the Infinity Cache
#pragma unroll 4
  for (uint32_t j = 0; j < slen; ++j) {
    if (j < cnt) {
      const uint32_t c = __builtin_nontemporal_load(cp + (size_t)j * 64);
      const uint16_t t = __builtin_nontemporal_load(tp + (size_t)j * 64);
      double2 cv;
      if (HALF) {
        cv = coef[t >> 1];
        if (t & 1) {
          cv.x = -cv.x;
          cv.y = -cv.y;
        }
      } else {
        cv = coef[t];
      }
#pragma unroll
      for (int v = 0; v < KC; ++v) {
        const double2 xv = x[(int64_t)c * ldx + v];
        ar[v] = fma(cv.x, xv.x, fma(-cv.y, xv.y, ar[v]));
        ai[v] = fma(cv.x, xv.y, fma(cv.y, xv.x, ai[v]));
      }
    }
  }
}

// XCD-aware split of n wave-sized work items. Workgroups are dispatched
// round-robin over the 8 XCDs (workgroup b runs on XCD b % 8), so the items are
// cut into 8 contiguous ranges and the waves of XCD x stride over range x:
// slices that are neighbours in the basis then share one XCD's L2. Grids of
// fewer than 8 workgroups fall back to a plain grid stride.
__device__ __forceinline__ void xcd_items(int64_t n, int64_t &first, int64_t &end, int64_t &stride) {
  constexpr int NX = 8;
  const int wpb = blockDim.x >> 6;
  const int64_t blocks = gridDim.x;
  if (blocks < NX) {
    first = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    end = n;
    stride = (blocks * blockDim.x) >> 6;
    return;
  }
  const int xcd = blockIdx.x % NX;
  const int64_t local = blockIdx.x / NX;
  const int64_t in_xcd = (blocks - xcd + NX - 1) / NX;
  const int64_t b = n * xcd / NX;
  end = n * (xcd + 1) / NX;
  first = b + local * wpb + (threadIdx.x >> 6);
  stride = in_xcd * wpb;
}

// Sum of `np` partials in a fixed order, identical in every block.
__device__ __forceinline__ double block_sum_fixed(const double *__restrict__ part, int np, double *red) {
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x >> 1; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// Block-wide sum of one value per thread, fixed order; result valid in thread 0.
__device__ __forceinline__ double block_sum_values(double v, double *red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int o = blockDim.x >> 1; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

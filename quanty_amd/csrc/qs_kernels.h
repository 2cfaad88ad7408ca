// Device building blocks shared by the apply and Lanczos translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#ifndef QS_BLOCK
#define QS_BLOCK 256
#endif

// One row of the assembled operator for KC interleaved vectors
// (entry j of vector v at x[j*ldx + v]). Lane `lane` of the slice's wavefront
// owns row r; the slice stores its entries column-major (entry e of the lane's
// row at sbase + 64*e + lane), so every load of the index stream is coalesced.
template <int KC>
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
#pragma unroll 4
  for (uint32_t j = 0; j < slen; ++j) {
    if (j < cnt) {
      const uint32_t c = cp[(size_t)j * 64];
      const uint16_t t = tp[(size_t)j * 64];
      const double2 cv = coef[t];
#pragma unroll
      for (int v = 0; v < KC; ++v) {
        const double2 xv = x[(int64_t)c * ldx + v];
        ar[v] = fma(cv.x, xv.x, fma(-cv.y, xv.y, ar[v]));
        ai[v] = fma(cv.x, xv.y, fma(cv.y, xv.x, ai[v]));
      }
    }
  }
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

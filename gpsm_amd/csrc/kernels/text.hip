// The reference's embedding writer (IO::output, io/IO.cpp:207-213: one line
// "(j, k) " per query vertex j = 0..q-1, then "\n", per embedding) on the
// device, as byte work.
//
// Rows are taken R at a time per workgroup (R rows of at most maxlen bytes
// fit the LDS stage of TEXT_STAGE_BYTES).  Pass 1 sums each workgroup's line lengths; an
// exclusive scan of those sums places every workgroup in the output.  Pass 2
// recomputes the lengths, scans them inside the workgroup, writes each row's
// characters into the LDS stage at its offset (one thread per row) and
// copies the stage to HBM with consecutive lanes on consecutive bytes.
#include <algorithm>

#include "launch.h"
#include "wave.h"

namespace gpsm {

namespace {

__device__ __forceinline__ uint32_t decimal_digits(uint32_t x) {
  uint32_t d = 1;
#pragma unroll
  for (uint32_t p = 10; p <= 1000000000u; p *= 10) d += x >= p;
  return d;
}

// Length of one line: the fixed characters plus the digits of every vertex.
__device__ __forceinline__ uint32_t line_length(const uint32_t* __restrict__ t, uint32_t q, uint32_t fixed) {
  uint32_t len = fixed;
  for (uint32_t j = 0; j < q; ++j) len += decimal_digits(t[j]);
  return len;
}

__device__ __forceinline__ uint32_t put_decimal(char* s, uint32_t x) {
  const uint32_t d = decimal_digits(x);
  for (uint32_t i = d; i-- > 0;) {
    s[i] = (char)('0' + x % 10);
    x /= 10;
  }
  return d;
}

__global__ void __launch_bounds__(256) text_len_kernel(const uint32_t* __restrict__ tuples, uint64_t rows,
                                                       uint32_t q, const uint8_t* __restrict__ keep, uint32_t R,
                                                       uint32_t fixed, uint32_t* __restrict__ wg_len) {
  __shared__ uint32_t part[4];
  const uint64_t r = (uint64_t)blockIdx.x * R + threadIdx.x;
  uint32_t len = 0;
  if (threadIdx.x < R && r < rows && (!keep || keep[r])) len = line_length(tuples + r * q, q, fixed);
  len = wave_sum(len);
  if (lane_id() == 0) part[threadIdx.x / WAVE] = len;
  __syncthreads();
  if (threadIdx.x == 0) wg_len[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

__global__ void __launch_bounds__(256) text_write_kernel(const uint32_t* __restrict__ tuples, uint64_t rows,
                                                         uint32_t q, const uint8_t* __restrict__ keep, uint32_t R,
                                                         uint32_t fixed, const uint32_t* __restrict__ wg_off,
                                                         char* __restrict__ out) {
  __shared__ char stage[TEXT_STAGE_BYTES];
  __shared__ uint32_t tmp[8];
  const uint64_t r = (uint64_t)blockIdx.x * R + threadIdx.x;
  const bool mine = threadIdx.x < R && r < rows && (!keep || keep[r]);
  const uint32_t* t = tuples + r * q;
  const uint32_t len = mine ? line_length(t, q, fixed) : 0u;
  uint32_t total;
  const uint32_t at = block_excl_scan(len, tmp, &total);
  if (mine) {
    char* s = stage + at;
    for (uint32_t j = 0; j < q; ++j) {
      *s++ = '(';
      s += put_decimal(s, j);
      *s++ = ',';
      *s++ = ' ';
      s += put_decimal(s, t[j]);
      *s++ = ')';
      *s++ = ' ';
    }
    *s = '\n';
  }
  __syncthreads();
  char* dst = out + wg_off[blockIdx.x];
  for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) dst[i] = stage[i];
}

}  // namespace

uint32_t text_fixed_length(uint32_t q) {
  uint32_t f = 1;  // "\n"
  for (uint32_t j = 0; j < q; ++j) f += 5 + (1 + (j >= 10) + (j >= 100));  // "(" j ", " ") ": k not counted
  return f;
}

uint32_t text_rows_per_wg(uint32_t q) {
  const uint32_t maxlen = text_fixed_length(q) + 10 * q;
  return std::max<uint32_t>(1, std::min<uint32_t>(256, TEXT_STAGE_BYTES / maxlen));
}

void launch_text_len(const uint32_t* tuples, uint64_t rows, uint32_t q, const uint8_t* keep, uint32_t* wg_len,
                     hipStream_t s) {
  const uint32_t R = text_rows_per_wg(q);
  const uint64_t nwg = (rows + R - 1) / R;
  if (!nwg) return;
  hipLaunchKernelGGL(text_len_kernel, dim3((uint32_t)nwg), dim3(256), 0, s, tuples, rows, q, keep, R,
                     text_fixed_length(q), wg_len);
}

void launch_text_write(const uint32_t* tuples, uint64_t rows, uint32_t q, const uint8_t* keep,
                       const uint32_t* wg_off, char* out, hipStream_t s) {
  const uint32_t R = text_rows_per_wg(q);
  const uint64_t nwg = (rows + R - 1) / R;
  if (!nwg) return;
  hipLaunchKernelGGL(text_write_kernel, dim3((uint32_t)nwg), dim3(256), 0, s, tuples, rows, q, keep, R,
                     text_fixed_length(q), wg_off, out);
}

}  // namespace gpsm

// Streaming ceilings for the roofline: plain copies that move a fixed number
// of bytes per workgroup, so a transform's HBM rate can be set beside what a
// copy reaches at the same granularity on the same device.  A workgroup of T
// threads moves U * T * 16 bytes: U non-temporal 16-byte loads per lane (lane
// t of the workgroup reads element t + T * u, so each load instruction of a
// wave covers 1 KiB contiguously), then U non-temporal stores.  Grids are
// one-shot: workgroup b moves bytes [b * U * T * 16, (b + 1) * U * T * 16).
//
// Built into libfftoc_copy.so, which the benchmark loads; nothing in
// libfftoc.so calls it.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

typedef double d2 __attribute__((ext_vector_type(2)));

template <int T, int U>
__global__ __launch_bounds__(T) void copy_blocks(const d2 *__restrict__ in, d2 *__restrict__ out) {
  const size_t base = (size_t)blockIdx.x * (T * U) + threadIdx.x;
  d2 v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(in + base + (size_t)T * u);
#pragma unroll
  for (int u = 0; u < U; ++u) __builtin_nontemporal_store(v[u], out + base + (size_t)T * u);
}

template <int T, int U>
int launch(const void *in, void *out, int64_t bytes, hipStream_t s) {
  constexpr int64_t per_block = (int64_t)T * U * 16;
  if (bytes % per_block != 0) return 2;
  const int64_t blocks = bytes / per_block;
  if (blocks > 0x7fffffffLL) return 2;
  hipLaunchKernelGGL((copy_blocks<T, U>), dim3((unsigned)blocks), dim3(T), 0, s,
                     static_cast<const d2 *>(in), static_cast<d2 *>(out));
  return hipGetLastError() == hipSuccess ? 0 : 7;
}

}  // namespace

extern "C" {

// Copies `bytes` from `in` to `out` (device pointers, 16-byte aligned) on
// `stream`, each workgroup moving `block_kib` KiB: 1 (64 threads, one load
// per lane), 4 (256 threads, one load), 16 (256 threads, four loads) or 64
// (512 threads, eight loads: the shape of one N = 4096 complex128 signal per
// workgroup).  Returns 0, 2 when `bytes` is not a positive multiple of the
// block, 3 for a null pointer, 4 for another block size, 7 when the launch
// fails.
int fftoc_copy_ceiling(const void *in, void *out, int64_t bytes, int block_kib, void *stream) {
  if (!in || !out) return 3;
  if (bytes <= 0) return 2;
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (block_kib) {
    case 1: return launch<64, 1>(in, out, bytes, s);
    case 4: return launch<256, 1>(in, out, bytes, s);
    case 16: return launch<256, 4>(in, out, bytes, s);
    case 64: return launch<512, 8>(in, out, bytes, s);
    default: return 4;
  }
}

}  // extern "C"

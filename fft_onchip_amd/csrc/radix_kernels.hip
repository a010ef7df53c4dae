// Stockham autosort FFT kernels on the vector ALUs (FP64), one template for
// every radix split.  These are the gfx950 counterparts of the reference's
// register-radix kernels legacy_fft<CT,N> (radix 8, includes/legacy8_fft.cuh:
// 57-118) and legacy16_fft<CT,N> (radix 16, includes/legacy16_fft.cuh:
// 99-172), which use the same pass structure: pass p with radix R reads
// a[j + m*N/R], multiplies by W_{R*L}^{m*k} (k = j mod L, L = product of the
// earlier radices), runs a radix-R butterfly and writes
// a[(j - k)*R + k + q*L].
//
// MI355X mapping:
//  * T = N / Rmax lanes own one signal, each holding Rmax complex values in
//    VGPRs; a workgroup processes S signals (a contiguous tile of S*N*16 B).
//  * Pass 0 reads straight from HBM: for fixed m the T lanes of a signal
//    read T consecutive complex values (16 B per lane, dwordx4), so each
//    load instruction covers whole 64..256 B runs.  The last pass writes
//    out[j + q*L] with j < L, again T consecutive values per signal.  For
//    T < 4 (N = 8, 16) the tile is instead staged through LDS with
//    fully contiguous 1 KiB-per-instruction copies.
//  * Between passes the data crosses lanes through LDS.  When a signal lives
//    inside one wave (T <= 64) no workgroup barrier is needed: the LDS
//    operations of a wave are executed in order.
//  * Twiddles W_4096^j come from a 64 KiB FP64 table built on the host in
//    long double (L2/L1 resident); fixed roots inside the butterflies are
//    exact-rounded constants.
#include "fftoc_device.hpp"
#include "fftoc_internal.hpp"

#include <utility>

namespace fftoc {
namespace {

template <int... Rs>
struct Split {
  static constexpr int P = sizeof...(Rs);
  static constexpr int R[P] = {Rs...};
  static constexpr int N = (Rs * ...);
  static constexpr int RM = [] {
    int m = 1;
    for (int r : {Rs...}) m = r > m ? r : m;
    return m;
  }();
  static constexpr int L(int p) {
    int l = 1;
    for (int i = 0; i < p; ++i) l *= R[i];
    return l;
  }
};

// LDS element index of tile element E = s*N + e: one pad slot after every 16
// elements keeps strided butterfly accesses on distinct banks.
__device__ __forceinline__ int lds_index(int E) { return E + (E >> 4); }
constexpr int lds_elems(int n) { return n + (n >> 4); }

template <class SP, int S, bool INV, bool STAGED>
struct StockhamKernel {
  static constexpr int N = SP::N;
  static constexpr int RM = SP::RM;
  static constexpr int T = N / RM;
  static constexpr int THREADS = T * S;
  static constexpr bool WAVE_LOCAL = (T <= 64) && (64 % T == 0);

  __device__ static void sync() {
    if constexpr (WAVE_LOCAL) {
      wave_lds_sync();
    } else {
      __syncthreads();
    }
  }

  template <int p>
  __device__ static void pass(c2 (&reg)[RM], c2 *lds, const c2 *gin, c2 *gout, int t,
                              int sbase, bool valid, const c2 *__restrict__ tw, double scale) {
    constexpr int R = SP::R[p];
    constexpr int L = SP::L(p);
    constexpr int NB = RM / R;
    constexpr bool FIRST = (p == 0);
    constexpr bool LAST = (p == SP::P - 1);

    // ---- gather inputs of this pass
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int j = t + i * T;
#pragma unroll
      for (int m = 0; m < R; ++m) {
        const int e = j + m * (N / R);
        c2 v;
        if constexpr (FIRST && !STAGED) {
          v = valid ? gin[e] : make_double2(0.0, 0.0);
          if constexpr (INV) v = cswap(v);
        } else {
          v = lds[lds_index(sbase + e)];
        }
        reg[i * R + m] = v;
      }
    }
    if constexpr (!FIRST || STAGED) {
      // all lanes have read this pass's inputs before anyone overwrites them
      sync();
    }

    // ---- twiddle W_{R L}^{m k}, k = j mod L, and butterfly
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int j = t + i * T;
      if constexpr (L > 1) {
        const int k = j & (L - 1);
        constexpr int step = kTwN / (R * L);
#pragma unroll
        for (int m = 1; m < R; ++m) {
          reg[i * R + m] = cmul(reg[i * R + m], tw[m * k * step]);
        }
      }
      dft<R>(&reg[i * R]);
    }

    // ---- scatter outputs
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int j = t + i * T;
      const int k = j & (L - 1);
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int d = (j - k) * R + k + q * L;
        c2 v = reg[i * R + q];
        if constexpr (LAST && !STAGED) {
          if (scale != 1.0) v = cscale(v, scale);
          if constexpr (INV) v = cswap(v);
          if (valid) gout[d] = v;
        } else {
          lds[lds_index(sbase + d)] = v;
        }
      }
    }
    if constexpr (!LAST || STAGED) sync();
  }

  template <int... ps>
  __device__ static void run_passes(std::integer_sequence<int, ps...>, c2 (&reg)[RM], c2 *lds,
                                    const c2 *gin, c2 *gout, int t, int sbase, bool valid,
                                    const c2 *__restrict__ tw, double scale) {
    (pass<ps>(reg, lds, gin, gout, t, sbase, valid, tw, scale), ...);
  }

  __device__ static void run(const c2 *in, c2 *out, int64_t batch, const c2 *__restrict__ tw,
                             double scale) {
    __shared__ c2 lds[lds_elems(S * N)];
    const int tid = threadIdx.x;
    const int t = tid % T;
    const int sl = tid / T;
    const int64_t tile0 = (int64_t)blockIdx.x * S;
    const int64_t sig = tile0 + sl;
    const bool valid = sig < batch;

    if constexpr (STAGED) {
      // contiguous tile copy HBM -> LDS (16 B per lane per instruction)
      const int64_t nvalid = (batch - tile0) < S ? (batch - tile0) : S;
      const c2 *src = in + tile0 * N;
#pragma unroll
      for (int c = 0; c < (S * N) / THREADS; ++c) {
        const int E = tid + c * THREADS;
        c2 v = (E < nvalid * N) ? src[E] : make_double2(0.0, 0.0);
        if constexpr (INV) v = cswap(v);
        lds[lds_index(E)] = v;
      }
      __syncthreads();
    }

    c2 reg[RM];
    run_passes(std::make_integer_sequence<int, SP::P>{}, reg, lds, in + sig * N, out + sig * N, t,
               sl * N, valid, tw, scale);

    if constexpr (STAGED) {
      __syncthreads();
      const int64_t nvalid = (batch - tile0) < S ? (batch - tile0) : S;
      c2 *dst = out + tile0 * N;
#pragma unroll
      for (int c = 0; c < (S * N) / THREADS; ++c) {
        const int E = tid + c * THREADS;
        c2 v = lds[lds_index(E)];
        if (scale != 1.0) v = cscale(v, scale);
        if constexpr (INV) v = cswap(v);
        if (E < nvalid * N) dst[E] = v;
      }
    }
  }
};

template <class SP, int S>
constexpr int kThreads = (SP::N / SP::RM) * S;

template <class SP, int S, bool INV, bool STAGED>
__global__ __launch_bounds__((kThreads<SP, S>)) void stockham_kernel(
    const c2 *in, c2 *out, int64_t batch, const c2 *__restrict__ tw, double scale) {
  StockhamKernel<SP, S, INV, STAGED>::run(in, out, batch, tw, scale);
}

template <class SP, int S, bool STAGED>
int launch_split(const LaunchArgs &a) {
  using KF = StockhamKernel<SP, S, false, STAGED>;
  const int64_t grid = (a.batch + S - 1) / S;
  if (grid > 0x7fffffffLL) return FFTOC_ERROR_INVALID_BATCH;
  const c2 *in = static_cast<const c2 *>(a.in);
  c2 *out = static_cast<c2 *>(a.out);
  const c2 *tw = static_cast<const c2 *>(a.twiddles);
  if (a.inverse) {
    hipLaunchKernelGGL((stockham_kernel<SP, S, true, STAGED>), dim3((unsigned)grid),
                       dim3(KF::THREADS), 0, a.stream, in, out, a.batch, tw, a.scale);
  } else {
    hipLaunchKernelGGL((stockham_kernel<SP, S, false, STAGED>), dim3((unsigned)grid),
                       dim3(KF::THREADS), 0, a.stream, in, out, a.batch, tw, a.scale);
  }
  return hipGetLastError() == hipSuccess ? FFTOC_SUCCESS : FFTOC_ERROR_LAUNCH;
}

}  // namespace

// Radix-8 family (legacy_fft): N = 8^d.
int launch_radix8(int n, const LaunchArgs &a) {
  switch (n) {
    case 8: return launch_split<Split<8>, 128, true>(a);
    case 64: return launch_split<Split<8, 8>, 32, false>(a);
    case 512: return launch_split<Split<8, 8, 8>, 4, false>(a);
    case 4096: return launch_split<Split<8, 8, 8, 8>, 1, false>(a);
    default: return FFTOC_ERROR_INVALID_SIZE;
  }
}

// Radix-16 family (legacy16_fft): N = 16^d.
int launch_radix16(int n, const LaunchArgs &a) {
  switch (n) {
    case 16: return launch_split<Split<16>, 64, true>(a);
    case 256: return launch_split<Split<16, 16>, 8, false>(a);
    case 4096: return launch_split<Split<16, 16, 16>, 1, false>(a);
    default: return FFTOC_ERROR_INVALID_SIZE;
  }
}

// Mixed radix for every power of two 8..4096 (largest radices first).
int launch_radix_mixed(int n, const LaunchArgs &a) {
  switch (n) {
    case 8: return launch_split<Split<8>, 128, true>(a);
    case 16: return launch_split<Split<16>, 64, true>(a);
    case 32: return launch_split<Split<8, 4>, 32, true>(a);
    case 64: return launch_split<Split<8, 8>, 32, false>(a);
    case 128: return launch_split<Split<16, 8>, 16, false>(a);
    case 256: return launch_split<Split<16, 16>, 8, false>(a);
    case 512: return launch_split<Split<8, 8, 8>, 4, false>(a);
    case 1024: return launch_split<Split<16, 16, 4>, 2, false>(a);
    case 2048: return launch_split<Split<16, 16, 8>, 1, false>(a);
    case 4096: return launch_split<Split<16, 16, 16>, 1, false>(a);
    default: return FFTOC_ERROR_INVALID_SIZE;
  }
}

}  // namespace fftoc

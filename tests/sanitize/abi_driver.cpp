// Drives every argument-checking path of the C ABI (include/fftoc.h) that
// needs no device, for the AddressSanitizer / UndefinedBehaviorSanitizer
// build of the library's host code.  Exits non-zero on the first unexpected
// status; the sanitizers abort on the first memory or UB error.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fftoc.h"

namespace {

int g_checks = 0;

void expect(int got, int want, const char *what, int line) {
  ++g_checks;
  if (got != want) {
    std::fprintf(stderr, "abi_asan: line %d: %s returned %d, expected %d\n", line, what, got,
                 want);
    std::exit(1);
  }
}
#define EXPECT(call, want) expect((call), (want), #call, __LINE__)

// Aligned, never dereferenced: every call below fails before any device work.
void *const kIn = reinterpret_cast<void *>(uintptr_t(1) << 20);
void *const kOut = reinterpret_cast<void *>(uintptr_t(1) << 32);

using Fn = int (*)(const void *, void *, int64_t, int, uint32_t, void *);

void transform_arguments() {
  const Fn fns[4] = {fftoc_fft, fftoc_tensor_fft, fftoc_legacy8_fft, fftoc_legacy16_fft};
  for (Fn f : fns) {
    for (int n : {-8, 0, 1, 2, 4, 12, 100, 8192, 1 << 30}) EXPECT(f(kIn, kOut, 1, n, 0, nullptr), 1);
    EXPECT(f(kIn, kOut, -1, 4096, 0, nullptr), 2);
    EXPECT(f(nullptr, kOut, 1, 4096, 0, nullptr), 3);
    EXPECT(f(kIn, nullptr, 1, 4096, 0, nullptr), 3);
    EXPECT(f(kIn, kOut, 1, 4096, 0x8u, nullptr), 4);      // unknown bit
    EXPECT(f(kIn, kOut, 1, 4096, 0x30u, nullptr), 4);     // unknown precision
    EXPECT(f(kIn, kOut, 1, 4096, 1u << 31, nullptr), 4);  // unknown high bit
    EXPECT(f(kIn, static_cast<char *>(kIn) + 16, 4, 4096, 0, nullptr), 5);  // partial overlap
    EXPECT(f(static_cast<char *>(kIn) + 8, kOut, 1, 4096, 0, nullptr), 10);
    EXPECT(f(static_cast<char *>(kIn) + 4, kOut, 1, 4096, FFTOC_DATA_C64, nullptr), 10);
    EXPECT(f(nullptr, nullptr, 0, 4096, 0, nullptr), 0);  // empty batch: no-op
    // batch * n * 16 overflows 64 bits
    EXPECT(f(kIn, kOut, INT64_MAX / 1024, 4096, 0, nullptr), 2);
  }
  EXPECT(fftoc_legacy8_fft(kIn, kOut, 1, 4096, FFTOC_PREC_FP32, nullptr), 6);
  EXPECT(fftoc_legacy16_fft(kIn, kOut, 1, 4096, FFTOC_PREC_BF16X3, nullptr), 6);
  EXPECT(fftoc_fft(kIn, kOut, 1, 64, FFTOC_PREC_FP32, nullptr), 6);
  EXPECT(fftoc_tensor_fft(kIn, kOut, 1, 512, FFTOC_PREC_FP32, nullptr), 6);
  EXPECT(fftoc_tensor_fft(kIn, kOut, 1, 64, FFTOC_PREC_FP32 | FFTOC_DATA_C64, nullptr), 6);
  EXPECT(fftoc_tensor_fft(kIn, kOut, 1, 1024, 0, nullptr), 1);
  EXPECT(fftoc_legacy8_fft(kIn, kOut, 1, 16, 0, nullptr), 1);
  EXPECT(fftoc_legacy16_fft(kIn, kOut, 1, 64, 0, nullptr), 1);
  EXPECT(fftoc_fft(kIn, kOut, 4, 256, FFTOC_TILE(7), nullptr), 6);  // no such tile
  using Alias = int (*)(const void *, void *, int64_t, uint32_t, void *);
  const Alias aliases[7] = {fftoc_tensor_fft_8,   fftoc_tensor_fft_16,  fftoc_tensor_fft_64,
                            fftoc_tensor_fft_128, fftoc_tensor_fft_256, fftoc_tensor_fft_512,
                            fftoc_tensor_fft_4096};
  for (Alias a : aliases) {
    EXPECT(a(nullptr, nullptr, 0, 0, nullptr), 0);
    EXPECT(a(kIn, kOut, -3, 0, nullptr), 2);
    EXPECT(a(nullptr, kOut, 5, 0, nullptr), 3);
  }
}

void variant_tables(const char *dir) {
  for (uint32_t dt : {0u, (uint32_t)FFTOC_DATA_C64}) {
    for (int n = -1; n <= 8200; ++n) {
      const int count = fftoc_variant_count(n, dt);
      const bool valid = n >= 8 && n <= 4096 && (n & (n - 1)) == 0;
      if (!valid) {
        EXPECT(count, 0);
        EXPECT(fftoc_get_variant(n, dt, 64), -1);
        EXPECT(fftoc_set_variant(n, dt, 64, 0), 1);
        EXPECT(fftoc_variant_tile(n, dt, 0), -1);
        if (fftoc_variant_name(n, dt, 0) != nullptr) expect(1, 0, "variant_name(invalid n)", __LINE__);
        continue;
      }
      if (count < 1) expect(count, 1, "variant_count", __LINE__);
      for (int v = -2; v <= count + 1; ++v) {
        const char *name = fftoc_variant_name(n, dt, v);
        const bool in = v >= 0 && v < count;
        if ((name != nullptr) != in) expect(0, 1, "variant_name range", __LINE__);
        if (in && std::strlen(name) == 0) expect(0, 1, "variant_name empty", __LINE__);
        if (in != (fftoc_variant_tile(n, dt, v) >= 1)) expect(0, 1, "variant_tile", __LINE__);
        // nothing has been timed in this process
        if (fftoc_variant_time_us(n, dt, v) >= 0) expect(0, 1, "variant_time_us", __LINE__);
        if (!in) EXPECT(fftoc_fft_variant(kIn, kOut, 1, n, dt, v, nullptr), 13);
      }
      for (int64_t b : {int64_t(-1), int64_t(0), int64_t(1), int64_t(3), int64_t(1) << 40,
                        INT64_MAX}) {
        const int v = fftoc_get_variant(n, dt, b);
        if (v < 0 || v >= count) expect(v, 0, "get_variant range", __LINE__);
      }
      EXPECT(fftoc_set_variant(n, dt, 64, count), 13);
      EXPECT(fftoc_set_variant(n, dt, 64, -1), 13);
      EXPECT(fftoc_autotune(n, dt, 0, 3, nullptr), 2);
      EXPECT(fftoc_autotune(n, dt, 64, 0, nullptr), 2);
    }
  }
  EXPECT(fftoc_autotune(100, 0, 64, 3, nullptr), 1);
  // persisted table: save, reload, then malformed files
  const std::string path = std::string(dir) + "/abi_asan_tuning.txt";
  EXPECT(fftoc_tuning_save(path.c_str()), 0);
  EXPECT(fftoc_tuning_load(path.c_str()), 0);
  EXPECT(fftoc_tuning_save(nullptr), 3);
  EXPECT(fftoc_tuning_load(nullptr), 3);
  EXPECT(fftoc_tuning_load((std::string(dir) + "/no/such/file").c_str()), 13);
  const std::string bad = std::string(dir) + "/abi_asan_bad.txt";
  FILE *f = std::fopen(bad.c_str(), "w");
  if (!f) expect(0, 1, "fopen", __LINE__);
  std::fprintf(f, "fft 64 c128 99 radix:8x8/tile/S16\n");          // bucket out of range
  std::fprintf(f, "fft 64 c256 3 radix:8x8/tile/S16\n");           // unknown data type
  std::fprintf(f, "nope 64 c128 3 radix:8x8/tile/S16\n");          // unknown family
  std::fprintf(f, "fft 100 c128 3 radix:8x8/tile/S16\n");          // bad length
  std::fprintf(f, "tensor 64 c128 3 radix:8x8/tile/S16\n");        // variant of another family
  std::fprintf(f, "fft 64 c128 3\n");                              // no name
  std::fprintf(f, "fft 64 c128 -3 x\n");
  std::fprintf(f, "%s\n", std::string(2000, 'x').c_str());         // longer than a line buffer
  std::fprintf(f, "fft 64 3 radix:8x8/tile/S16\n");                // old format
  std::fclose(f);
  EXPECT(fftoc_tuning_load(bad.c_str()), 13);
  EXPECT(fftoc_tuning_load(path.c_str()), 0);  // the saved table still loads after that
  std::remove(path.c_str());
  std::remove(bad.c_str());
}

void plans_and_servers() {
  fftoc_plan p = nullptr;
  fftoc_job j{kIn, kOut, 4, 64, 0};
  EXPECT(fftoc_plan_create(&j, 1, nullptr), 3);
  EXPECT(fftoc_plan_create(nullptr, 1, &p), 3);
  EXPECT(fftoc_plan_create(&j, -1, &p), 3);
  fftoc_job bad = j;
  bad.n = 100;
  EXPECT(fftoc_plan_create(&bad, 1, &p), 1);
  bad = j;
  bad.flags = FFTOC_PREC_FP32;
  EXPECT(fftoc_plan_create(&bad, 1, &p), 6);
  bad.flags = 0x8;
  EXPECT(fftoc_plan_create(&bad, 1, &p), 4);
  bad = j;
  bad.batch = -2;
  EXPECT(fftoc_plan_create(&bad, 1, &p), 2);
  bad = j;
  bad.in = nullptr;
  EXPECT(fftoc_plan_create(&bad, 1, &p), 3);
  bad = j;
  bad.out = static_cast<char *>(kOut) + 8;
  EXPECT(fftoc_plan_create(&bad, 1, &p), 10);
  bad = j;
  bad.out = static_cast<char *>(kIn) + 16;
  EXPECT(fftoc_plan_create(&bad, 1, &p), 5);
  // many jobs, one pair writing into each other: the overlap sweep finds it
  std::vector<fftoc_job> jobs;
  for (int i = 0; i < 300; ++i) {
    char *base = static_cast<char *>(kOut) + (size_t)i * (1 << 16);
    jobs.push_back({base, base, 16, 256, (uint32_t)(i % 2 ? FFTOC_DATA_C64 : 0)});
  }
  jobs[123].out = static_cast<char *>(jobs[200].out) + 256;  // in place inside job 200
  jobs[123].in = jobs[123].out;
  EXPECT(fftoc_plan_create(jobs.data(), (int)jobs.size(), &p), 5);
  // the allocation-free batch: the same checks, before any device work
  EXPECT(fftoc_batch_fft_noalloc(nullptr, 1, nullptr), 3);
  EXPECT(fftoc_batch_fft_noalloc(&j, -1, nullptr), 3);
  EXPECT(fftoc_batch_fft_noalloc(nullptr, 0, nullptr), 0);
  EXPECT(fftoc_batch_fft_noalloc(jobs.data(), FFTOC_NOALLOC_MAX_JOBS + 1, nullptr), 11);
  EXPECT(fftoc_batch_fft_noalloc(&bad, 1, nullptr), 5);
  std::vector<fftoc_job> few(jobs.begin() + 120, jobs.begin() + 120 + FFTOC_NOALLOC_MAX_JOBS);
  few[3].out = static_cast<char *>(few[20].out) + 256;  // writes into job 20's range
  few[3].in = few[3].out;
  EXPECT(fftoc_batch_fft_noalloc(few.data(), (int)few.size(), nullptr), 5);
  few[3].batch = 0;  // empty jobs are not checked for overlap
  few[7].n = 48;
  EXPECT(fftoc_batch_fft_noalloc(few.data(), (int)few.size(), nullptr), 1);
  EXPECT(fftoc_plan_execute(nullptr, nullptr), 13);
  EXPECT(fftoc_plan_execute_graph(nullptr, nullptr), 13);
  EXPECT(fftoc_plan_autotune(nullptr, 2, nullptr), 13);
  const char *name = nullptr;
  EXPECT(fftoc_plan_routine(nullptr, 64, 0, &name), 13);
  EXPECT(fftoc_plan_set_routine(nullptr, 64, 0, "8x8/tile/S32"), 13);
  EXPECT(fftoc_plan_destroy(nullptr), 13);
  for (int n = 0; n <= 8192; n = n ? n * 2 : 1) {
    for (int i = -1; i < 8; ++i) {
      const int rc = fftoc_batch_routine_name(n, i, &name);
      const bool valid_n = n >= 8 && n <= 4096;
      if (!valid_n) {
        expect(rc, 1, "batch_routine_name(bad n)", __LINE__);
      } else if (rc == 0) {
        if (i < 0 || !name || !*name) expect(0, 1, "batch_routine_name", __LINE__);
      } else {
        expect(rc, 13, "batch_routine_name(past the end)", __LINE__);
      }
    }
    EXPECT(fftoc_batch_routine_name(n >= 8 && n <= 4096 ? n : 64, 0, nullptr), 3);
  }
  fftoc_server s = nullptr;
  EXPECT(fftoc_server_create(1 << 20, 4, nullptr), 3);
  EXPECT(fftoc_server_create(100, 4, &s), 2);
  EXPECT(fftoc_server_create(1 << 20, 0, &s), 2);
  EXPECT(fftoc_server_create(1 << 20, 1 << 20, &s), 2);
  uint64_t id = 0;
  int done = 0;
  void *out = nullptr;
  size_t used = 0;
  int inflight = 0;
  double lat = 0, span = 0, arr = 0;
  EXPECT(fftoc_server_submit(nullptr, kIn, 1, 64, 0, &id), 13);
  EXPECT(fftoc_server_poll(nullptr, 0, &done, &out), 13);
  EXPECT(fftoc_server_release(nullptr, 0), 13);
  uint64_t ticks = 0;
  EXPECT(fftoc_server_gpu_time(nullptr, 0, &ticks), 3);
  EXPECT(fftoc_server_replay(nullptr, &j, 1, &arr, &lat, &span), 3);
  EXPECT(fftoc_server_stats(nullptr, &used, &inflight), 13);
  EXPECT(fftoc_server_set_idle_timeout(nullptr, 1.0), 13);
  EXPECT(fftoc_server_destroy(nullptr), 13);
}

void library() {
  for (int st = -5; st < 40; ++st) {
    const char *m = fftoc_strerror(st);
    if (!m || !*m) expect(0, 1, "strerror", __LINE__);
  }
  if (fftoc_version() < 100) expect(fftoc_version(), 100, "version", __LINE__);
  EXPECT(fftoc_twiddle_table(nullptr, nullptr), 3);
  EXPECT(fftoc_shutdown(), 0);
}

}  // namespace

int main(int argc, char **argv) {
  const char *dir = argc > 1 ? argv[1] : "/tmp";
  library();
  transform_arguments();
  variant_tables(dir);
  plans_and_servers();
  std::printf("abi_asan: %d checks passed\n", g_checks);
  return 0;
}

// generator.cpp -- the group-structured temporal contact generator of
// generate_networks/graph_generator.c:39-157, as a library function.
//
// For a seed it writes the file the reference generator writes on a glibc
// system: the same draws from the C library's rand() stream in the same
// order (one per node pair and step, two more per contact), the same
// timestamps and the same text.  The rand() stream is produced here by the
// C library's additive feedback generator (degree 31, separation 3), seeded
// as srand() seeds it, so the output does not depend on the C library the
// program runs on and no global generator state is touched.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "replay_hip.h"

namespace {

// rand() after srand(seed) on glibc: state r[0] = seed (1 for 0), then
// r[i] = 16807 r[i-1] mod (2^31 - 1) for i < 31 (Schrage's method, as the
// library computes it), then r[i] = r[i-31] + r[i-3] mod 2^32; the first
// 310 values of that recurrence are discarded and each later one yields
// r >> 1.
class LibcRand {
 public:
  explicit LibcRand(uint32_t seed) {
    int32_t word = static_cast<int32_t>(seed == 0 ? 1 : seed);
    r_[0] = static_cast<uint32_t>(word);
    for (int i = 1; i < 31; ++i) {
      const int32_t hi = word / 127773, lo = word % 127773;
      word = 16807 * lo - 2836 * hi;
      if (word < 0) word += 2147483647;
      r_[i] = static_cast<uint32_t>(word);
    }
    f_ = 3;
    b_ = 0;
    for (int k = 0; k < 310; ++k) next();
  }
  int32_t next() {
    const uint32_t v = r_[f_] += r_[b_];
    f_ = f_ + 1 == 31 ? 0 : f_ + 1;
    b_ = b_ + 1 == 31 ? 0 : b_ + 1;
    return static_cast<int32_t>(v >> 1);
  }

 private:
  uint32_t r_[31];
  int f_, b_;
};

constexpr double kRandMax = 2147483647.0;

// Civil date of a day count since 1970-01-01 (proleptic Gregorian), by
// 400-year eras.
void civil_from_days(int64_t z, int64_t &y, unsigned &m, unsigned &d) {
  z += 719468;
  const int64_t era = (z >= 0 ? z : z - 146096) / 146097;
  const unsigned doe = static_cast<unsigned>(z - era * 146097);
  const unsigned yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
  y = static_cast<int64_t>(yoe) + era * 400;
  const unsigned doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
  const unsigned mp = (5 * doy + 2) / 153;
  d = doy - (153 * mp + 2) / 5 + 1;
  m = mp < 10 ? mp + 3 : mp - 9;
  if (m <= 2) ++y;
}

// "%Y-%m-%dT%H:%M:%SZ" of UTC seconds t (gmtime + strftime), years 0..9999.
void iso_utc(int64_t t, char out[21]) {
  int64_t days = t / 86400, sod = t % 86400;
  if (sod < 0) {
    sod += 86400;
    --days;
  }
  int64_t y;
  unsigned m, d;
  civil_from_days(days, y, m, d);
  std::snprintf(out, 21, "%04lld-%02u-%02uT%02u:%02u:%02uZ", static_cast<long long>(y), m, d,
                static_cast<unsigned>(sod / 3600), static_cast<unsigned>(sod / 60 % 60),
                static_cast<unsigned>(sod % 60));
}

}  // namespace

extern "C" int32_t replay_generate_groups(int64_t base_time, int32_t steps, int32_t step_size,
                                          int32_t population, int32_t groups,
                                          double prob_inner, double prob_inter, int32_t seed,
                                          const char *path, int64_t *n_contacts, int32_t verbose) {
  if (!path || steps < 0 || step_size <= 0 || population < 0 || groups <= 0 || seed < 0)
    return REPLAY_E_ARG;
  FILE *f = std::fopen(path, "w");
  if (!f) return REPLAY_E_IO;
  std::vector<char> buf;
  buf.reserve(1 << 20);
  auto flush = [&] {
    if (!buf.empty()) std::fwrite(buf.data(), 1, buf.size(), f);
    buf.clear();
  };
  // n: what snprintf returned for a buffer of cap bytes (clamped to what it
  // wrote; nothing on an encoding error)
  auto put = [&](const char *s, int n, size_t cap) {
    if (n <= 0) return;
    buf.insert(buf.end(), s, s + std::min(static_cast<size_t>(n), cap - 1));
    if (buf.size() >= (1u << 20)) flush();
  };
  char tmp[96];
  put("NODE_LIST", 9, 10);
  for (int32_t i = 1; i <= population; ++i)
    put(tmp, std::snprintf(tmp, sizeof tmp, " %d", i), sizeof tmp);
  put("\n", 1, 2);
  LibcRand rng(static_cast<uint32_t>(seed));
  int64_t count = 0;
  char a[21], b[21];
  for (int32_t step = 0; step < steps; ++step) {
    if (verbose) std::printf("Generating step %d/%d...\n", step + 1, steps);
    const int64_t t0 = base_time + static_cast<int64_t>(step) * step_size;
    for (int32_t i = 0; i < population; ++i) {
      for (int32_t j = i + 1; j < population; ++j) {
        const double p = (i % groups == j % groups) ? prob_inner : prob_inter;
        if (static_cast<double>(rng.next()) / kRandMax < p) {
          iso_utc(t0 + rng.next() % step_size, a);
          iso_utc(t0 + rng.next() % step_size, b);
          const bool swap = std::strcmp(a, b) > 0;
          put(tmp,
              std::snprintf(tmp, sizeof tmp, "%d, %d, %s, %s;\n", i + 1, j + 1, swap ? b : a,
                            swap ? a : b),
              sizeof tmp);
          ++count;
        }
      }
    }
  }
  flush();
  const bool ok = std::fclose(f) == 0;
  if (n_contacts) *n_contacts = count;
  if (verbose) std::printf("Temporal graph file generated: %s\n", path);
  return ok ? REPLAY_OK : REPLAY_E_IO;
}

// replay_gen -- command line of the group-structured generator
// (generate_networks/main.c:39-65): the same nine arguments, messages and
// output file, written by replay_generate_groups.  The start time is read
// as UTC; the reference's main() reads it as local time (mktime), which is
// the same on a machine whose time zone is UTC.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "replay_hip.h"

namespace {

// Days from 1970-01-01 to a proleptic Gregorian date, by 400-year eras;
// month and day may lie outside their ranges (normalised as mktime does).
int64_t days_from_civil(int64_t y, int64_t m, int64_t d) {
  y += (m - 1) >= 0 ? (m - 1) / 12 : -((12 - m) / 12);
  m = ((m - 1) % 12 + 12) % 12 + 1;
  y -= m <= 2;
  const int64_t era = (y >= 0 ? y : y - 399) / 400;
  const int64_t yoe = y - era * 400;
  const int64_t doy = (153 * (m + (m > 2 ? -3 : 9)) + 2) / 5 + d - 1;
  const int64_t doe = yoe * 365 + yoe / 4 - yoe / 100 + doy;
  return era * 146097 + doe - 719468;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 10) {
    std::printf("Usage: %s <start_datetime> <steps> <step_size> <population_size> "
                "<num_groups> <prob_inner> <prob_inter> <output_file> <seed>\n",
                argv[0]);
    return 1;
  }
  int year, month, day, hour, minute, second;
  if (std::sscanf(argv[1], "%d-%d-%dT%d:%d:%d", &year, &month, &day, &hour, &minute,
                  &second) != 6) {
    std::fprintf(stderr, "Error parsing datetime string: %s\n", argv[1]);
    return EXIT_FAILURE;
  }
  const int64_t base = days_from_civil(year, month, day) * 86400 + int64_t(hour) * 3600 +
                       int64_t(minute) * 60 + second;
  const int seed = std::atoi(argv[9]);
  if (seed < 0) {
    std::fprintf(stderr, "Seed must be a non-negative integer.\n");
    return 0;
  }
  int64_t n = 0;
  const int rc = replay_generate_groups(base, std::atoi(argv[2]), std::atoi(argv[3]),
                                        std::atoi(argv[4]), std::atoi(argv[5]),
                                        std::atof(argv[6]), std::atof(argv[7]), seed, argv[8],
                                        &n, 1);
  if (rc == REPLAY_E_IO) {
    std::fprintf(stderr, "Error: could not write '%s'\n", argv[8]);
    return 0;
  }
  if (rc != REPLAY_OK) {
    std::fprintf(stderr, "[replay_gen] %s\n", replay_error_string(rc));
    return 1;
  }
  return 0;
}

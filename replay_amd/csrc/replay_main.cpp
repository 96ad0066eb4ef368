// replay_main.cpp -- the `replay` executable on top of libreplay_hip.
//
// Keeps the command-line and output contract of the reference executable
// (tool/gpu_cpu_temporal_sim.cu:570-1185):
//   - positional graph path (last argument not starting with '-', cu:635-638);
//   - the same flags and defaults (cu:25-27, 48-56, 639-699); a value flag
//     without a value or any unknown flag prints "Unrecognized option: %s"
//     and exits 1;
//   - -q / --disable-debug-prints silences stdout (cu:621-631);
//   - output files are removed and re-created while parsing flags
//     (cu:653-666), headers as cu:707-713, rows as cu:553-560, with the
//     initial block at --start-time and one block per iteration at
//     start + llround(iter * step) (cu:829-831, 876, 1093-1095);
//   - the stdout marker line "Starting temporal Monte Carlo simulation for
//     %d iterations..." (cu:835-836).
// Differences: the library has no CPU engine, so --cpu-only (cu:1109-1144,
// the CPU twin) is an error that says so (exit 1) rather than a silent
// GPU run; --cpu-threads, which only sizes that twin's OpenMP pool, is
// accepted and has no effect; --seed <u64> and --device <n> are added.
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "replay_hip.h"

namespace {

const char *kStateNames[4] = {"susceptible", "incubating", "infectious", "resistant"};

int classify(int32_t v, int32_t upper, int32_t Ed, int32_t Id, int32_t Rd) {
  const int32_t elapsed = (v <= 0 || v > upper) ? 0 : (upper - v);
  if (v == 0) return 0;
  if (Ed > 0 && elapsed < Ed) return 1;
  if (Id > 0 && elapsed < Ed + Id) return 2;
  if (Rd > 0 && elapsed < Ed + Id + Rd) return 3;
  return 0;
}

// Decimal digits of v, as printf's %d / %lld write them.
inline char *put_int(char *p, long long v) {
  unsigned long long u = static_cast<unsigned long long>(v);
  if (v < 0) {
    *p++ = '-';
    u = 0ULL - u;
  }
  char tmp[20];
  int n = 0;
  do {
    tmp[n++] = static_cast<char>('0' + u % 10);
    u /= 10;
  } while (u);
  while (n) *p++ = tmp[--n];
  return p;
}

// Buffered writer for the CSV rows: the bytes of "%lld,%d,%d,%d,%d,%d\n"
// (cu:553-560) and "%lld,%d,%d,%s\n", formatted directly into the buffer.
class Out {
 public:
  explicit Out(FILE *f) : f_(f) { buf_.reserve(kChunk + 256); }
  void row_summary(long long t, int sim, const int32_t *c4) {
    char *p = room(128);
    p = put_int(p, t);
    *p++ = ',';
    p = put_int(p, sim);
    for (int c = 0; c < 4; ++c) {
      *p++ = ',';
      p = put_int(p, c4[c]);
    }
    *p++ = '\n';
    commit(p);
  }
  void row_node(long long t, int sim, int node, const char *state) {
    const size_t ls = std::strlen(state);
    char *p = room(64 + ls);
    p = put_int(p, t);
    *p++ = ',';
    p = put_int(p, sim);
    *p++ = ',';
    p = put_int(p, node);
    *p++ = ',';
    std::memcpy(p, state, ls);
    p += ls;
    *p++ = '\n';
    commit(p);
  }
  void flush() {
    if (used_) std::fwrite(buf_.data(), 1, used_, f_);
    used_ = 0;
    std::fflush(f_);
  }

 private:
  static constexpr size_t kChunk = 1u << 20;
  char *room(size_t n) {
    if (used_ + n > buf_.size()) {
      if (used_ >= kChunk) {
        std::fwrite(buf_.data(), 1, used_, f_);
        used_ = 0;
      }
      if (used_ + n > buf_.size()) buf_.resize(used_ + n + kChunk);
    }
    return &buf_[used_];
  }
  void commit(char *p) { used_ = static_cast<size_t>(p - buf_.data()); }
  FILE *f_;
  std::vector<char> buf_;
  size_t used_ = 0;
};

void usage(const char *argv0) {
  std::fprintf(stderr, "Usage: %s <graph file> [options]\n", argv0);
  std::fprintf(stderr, "Options:\n");
  std::fprintf(stderr, "  --cpu-only             Not supported: this build runs on the GPU only.\n");
  std::fprintf(stderr, "  --cpu-threads <int>    Accepted; sizes nothing here (no CPU engine).\n");
  std::fprintf(stderr, "  --M <int>              Number of parallel simulations.\n");
  std::fprintf(stderr, "  --step-size <float>    Time step size.\n");
  std::fprintf(stderr, "  --iterations <int>     Number of main simulation steps.\n");
  std::fprintf(stderr, "  --summary-out <file>   Output file for summary state counts.\n");
  std::fprintf(stderr, "  --node-state-out <file> Output file for per-node states.\n");
  std::fprintf(stderr, "  --initial-infected <float>     Initial infection probability [0-1]\n");
  std::fprintf(stderr, "  --infect-prob <float>          Infection prob per hour of exposure [0-1]\n");
  std::fprintf(stderr, "  --exposed-duration <int>       Seconds before an infected person becomes infectious\n");
  std::fprintf(stderr, "  --infectious-duration <int>    Seconds a person remains infectious\n");
  std::fprintf(stderr, "  --resistant-duration <int>     Seconds a recovered person remains resistant\n");
  std::fprintf(stderr, "  --start-time <epoch>           Unix start time of the simulation.\n");
  std::fprintf(stderr, "  --static-network-duration <seconds>  Duration of each static contact network window.\n");
  std::fprintf(stderr, "  [deprecated] --time-step <seconds>   Alias for --static-network-duration.\n");
  std::fprintf(stderr, "  --seed <u64>           Philox key of the random draws (default 0).\n");
  std::fprintf(stderr, "  --device <int>         HIP device ordinal (default 0).\n");
}

}  // namespace

int main(int argc, char **argv) {
  std::printf("=== replay (HIP, gfx950) temporal epidemic simulation ===\n");
  if (argc < 2) {
    usage(argv[0]);
    return 1;
  }
  bool quiet = false;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--disable-debug-prints") || !std::strcmp(argv[i], "-q")) {
      quiet = true;
      if (!std::freopen("/dev/null", "w", stdout)) return 1;
      break;
    }
  }
  replay_params p;
  replay_params_init(&p);
  const char *filename = nullptr;
  FILE *summary_fp = nullptr, *node_fp = nullptr;
  for (int i = 1; i < argc; ++i) {
    const char *a = argv[i];
    const bool has = i + 1 < argc;
    if (a[0] != '-') {
      filename = a;
      continue;
    }
    if (!std::strcmp(a, "--cpu-only")) {
      std::fprintf(stderr,
                   "Error: --cpu-only is not supported: this build has no CPU engine and runs "
                   "the simulation on the GPU. Run without --cpu-only.\n");
      return 1;
    } else if (!std::strcmp(a, "--disable-debug-prints") || !std::strcmp(a, "-q")) {
      quiet = true;
    } else if (!std::strcmp(a, "--cpu-threads") && has) {
      ++i;
    } else if (!std::strcmp(a, "--M") && has) {
      p.M = std::atoi(argv[++i]);
    } else if (!std::strcmp(a, "--step-size") && has) {
      p.step_size = static_cast<float>(std::atof(argv[++i]));
    } else if (!std::strcmp(a, "--iterations") && has) {
      p.iterations = std::atoi(argv[++i]);
    } else if (!std::strcmp(a, "--summary-out") && has) {
      std::remove(argv[i + 1]);
      summary_fp = std::fopen(argv[++i], "w");
      if (!summary_fp) {
        std::fprintf(stderr, "Error: Could not open summary output file.\n");
        return 1;
      }
    } else if (!std::strcmp(a, "--node-state-out") && has) {
      std::remove(argv[i + 1]);
      node_fp = std::fopen(argv[++i], "w");
      if (!node_fp) {
        std::fprintf(stderr, "Error: Could not open node state output file.\n");
        return 1;
      }
    } else if (!std::strcmp(a, "--initial-infected") && has) {
      p.initial_infected = static_cast<float>(std::atof(argv[++i]));
    } else if (!std::strcmp(a, "--infect-prob") && has) {
      p.infect_prob = static_cast<float>(std::atof(argv[++i]));
    } else if (!std::strcmp(a, "--exposed-duration") && has) {
      p.exposed_duration = std::atoi(argv[++i]);
    } else if (!std::strcmp(a, "--infectious-duration") && has) {
      p.infectious_duration = std::atoi(argv[++i]);
    } else if (!std::strcmp(a, "--resistant-duration") && has) {
      p.resistant_duration = std::atoi(argv[++i]);
    } else if (!std::strcmp(a, "--start-time") && has) {
      p.start_time = static_cast<int64_t>(std::atoll(argv[++i]));
    } else if ((!std::strcmp(a, "--static-network-duration") || !std::strcmp(a, "--time-step")) &&
               has) {
      p.static_network_duration = static_cast<int64_t>(std::atoll(argv[++i]));
    } else if (!std::strcmp(a, "--seed") && has) {
      p.seed = static_cast<uint64_t>(std::strtoull(argv[++i], nullptr, 10));
    } else if (!std::strcmp(a, "--device") && has) {
      p.device = std::atoi(argv[++i]);
    } else {
      std::fprintf(stderr, "Unrecognized option: %s\n", a);
      return 1;
    }
  }
  if (!filename) {
    std::fprintf(stderr, "Missing contact network file argument.\n");
    return 1;
  }
  if (summary_fp) std::fprintf(summary_fp, "time,simulation_id,exposed,infectious,resistant,susceptible\n");
  if (node_fp) std::fprintf(node_fp, "time,simulation_id,node_id,state\n");

  replay_ctx *ctx = nullptr;
  int rc = replay_open(filename, &p, &ctx);
  if (rc != REPLAY_OK) {
    const char *msg = replay_last_error(nullptr);
    if (rc == REPLAY_E_ARG && std::strstr(msg, "static network"))
      std::fprintf(stderr, "Error: static network duration must be positive.\n");
    else if (rc == REPLAY_E_ARG && std::strstr(msg, "step size"))
      std::fprintf(stderr, "Error: step size must be positive.\n");
    else
      std::fprintf(stderr, "No intervals read or parse error.\n");
    std::fprintf(stderr, "[replay_hip] %s (%s)\n", msg, replay_error_string(rc));
    return 1;
  }
  replay_info info;
  replay_get_info(ctx, &info);
  std::printf("[INFO] Node count set from NODE_LIST: %d\n", info.n_nodes);
  std::fprintf(stderr, "Parsed %lld edges (intervals) from graph.\n", static_cast<long long>(info.n_contacts));
  std::fprintf(stderr, "NCLS built. n=%lld, nlists=%lld.\n", static_cast<long long>(info.n_contacts),
               static_cast<long long>(info.nlists));

  const int32_t N = info.n_nodes, M = info.M, T = p.iterations;
  const int32_t upper = p.exposed_duration + p.infectious_duration + p.resistant_duration;
  std::vector<int32_t> state;
  Out sum_out(summary_fp ? summary_fp : stdout), node_out(node_fp ? node_fp : stdout);
  auto write_block = [&](long long t, const int32_t *counts, const int32_t *st) {
    for (int32_t sim = 0; sim < M; ++sim) {
      if (node_fp && st)
        for (int32_t node = 0; node < N; ++node)
          node_out.row_node(t, sim, node,
                            kStateNames[classify(st[static_cast<size_t>(sim) * N + node], upper,
                                                 p.exposed_duration, p.infectious_duration,
                                                 p.resistant_duration)]);
      if (summary_fp) sum_out.row_summary(t, sim, counts + static_cast<size_t>(sim) * 4);
    }
    if (summary_fp) sum_out.flush();
    if (node_fp) node_out.flush();
  };

  std::vector<int32_t> counts(static_cast<size_t>(T + 1) * M * 4);
  if (summary_fp || node_fp) {
    if (node_fp) {
      state.resize(static_cast<size_t>(N) * M);
      replay_get_state(ctx, state.data(), static_cast<int64_t>(state.size()));
    }
    replay_get_counts(ctx, counts.data(), static_cast<int64_t>(M) * 4);
    write_block(p.start_time, counts.data(), node_fp ? state.data() : nullptr);
  }
  std::printf("\n");
  std::printf("Starting temporal Monte Carlo simulation for %d iterations...\n", T);
  std::printf("[INFO] Simulation start time: %lld, time step: %.3f seconds, static network window: %lld seconds\n",
              static_cast<long long>(p.start_time), static_cast<double>(p.step_size),
              static_cast<long long>(p.static_network_duration));
  std::fflush(stdout);

  if (node_fp) {
    // Per-iteration state is needed, so iterations run one at a time.  Each
    // state goes to a writer thread through two host buffers: iteration k+1
    // runs on the device while iteration k's rows are formatted.  The writer
    // classifies every (replica, node) once, for the node row and for the
    // replica's summary counts (cu:505-568 write both from the same state),
    // so no count block is fetched per iteration.
    std::vector<long long> times(static_cast<size_t>(T));
    for (int32_t k = 0; k < T; ++k) times[static_cast<size_t>(k)] = replay_iteration_time(ctx, k);
    std::vector<int32_t> buf[2];
    buf[0].resize(static_cast<size_t>(N) * M);
    buf[1].resize(static_cast<size_t>(N) * M);
    std::mutex mu;
    std::condition_variable cv;
    int32_t queued = -1, written = -1;  // last iteration handed over / written
    bool stop = false;
    std::thread writer([&] {
      for (int32_t k = 0;; ++k) {
        {
          std::unique_lock<std::mutex> lk(mu);
          cv.wait(lk, [&] { return queued >= k || stop; });
          if (queued < k) return;
        }
        const int32_t *st = buf[k & 1].data();
        int32_t *c4 = counts.data() + static_cast<size_t>(k + 1) * M * 4;
        for (int32_t sim = 0; sim < M; ++sim) {
          int32_t n4[4] = {0, 0, 0, 0};  // susceptible, incubating, infectious, resistant
          for (int32_t node = 0; node < N; ++node) {
            const int cls = classify(st[static_cast<size_t>(sim) * N + node], upper,
                                     p.exposed_duration, p.infectious_duration,
                                     p.resistant_duration);
            ++n4[cls];
            node_out.row_node(times[static_cast<size_t>(k)], sim, node, kStateNames[cls]);
          }
          int32_t *c = c4 + static_cast<size_t>(sim) * 4;
          c[0] = n4[1];
          c[1] = n4[2];
          c[2] = n4[3];
          c[3] = n4[0];
          if (summary_fp) sum_out.row_summary(times[static_cast<size_t>(k)], sim, c);
        }
        if (summary_fp) sum_out.flush();
        node_out.flush();
        {
          std::lock_guard<std::mutex> lk(mu);
          written = k;
        }
        cv.notify_all();
      }
    });
    for (int32_t k = 0; k < T && rc == REPLAY_OK; ++k) {
      rc = replay_run(ctx, 1);
      {
        // buffer k & 1 is free once iteration k - 2 is written
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return written >= k - 2; });
      }
      if (rc == REPLAY_OK)
        rc = replay_get_state(ctx, buf[k & 1].data(), static_cast<int64_t>(buf[k & 1].size()));
      if (rc == REPLAY_OK) {
        std::lock_guard<std::mutex> lk(mu);
        queued = k;
      }
      cv.notify_all();
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv.notify_all();
    writer.join();
  } else {
    rc = replay_run(ctx, T);
    if (rc == REPLAY_OK) rc = replay_get_counts(ctx, counts.data(), static_cast<int64_t>(counts.size()));
    if (rc == REPLAY_OK && summary_fp)
      for (int32_t k = 0; k < T; ++k)
        write_block(replay_iteration_time(ctx, k), counts.data() + static_cast<size_t>(k + 1) * M * 4,
                    nullptr);
  }
  if (rc != REPLAY_OK) {
    std::fprintf(stderr, "[replay_hip] %s (%s)\n", replay_last_error(ctx), replay_error_string(rc));
    replay_close(ctx);
    return 1;
  }
  if (!quiet) {
    const int32_t show = M < 5 ? M : 5;
    for (int32_t k = 0; k < T; ++k) {
      std::printf("\n--- Iteration %d (sim time %lld) ---\n", k + 1,
                  static_cast<long long>(replay_iteration_time(ctx, k)));
      for (int32_t sim = 0; sim < show; ++sim) {
        const int32_t *c = counts.data() + (static_cast<size_t>(k + 1) * M + sim) * 4;
        std::printf("Row %d: Exposed=%d, Infectious=%d, Resistant=%d, Susceptible=%d\n", sim, c[0],
                    c[1], c[2], c[3]);
      }
    }
  }
  std::printf("Simulation done.\n");
  replay_close(ctx);
  if (summary_fp) std::fclose(summary_fp);
  if (node_fp) std::fclose(node_fp);
  return 0;
}

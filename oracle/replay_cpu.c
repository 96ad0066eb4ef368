/*
 * replay_cpu.c -- command-line front end of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  Accepts the reference `replay` argv
 * (tool/gpu_cpu_temporal_sim.cu:621-699) and always runs the CPU restatement,
 * so the CSV files it writes can be compared byte for byte with the HIP
 * product's `replay`.  Extra flags: --seed <u64> (Philox key), --libm-exp
 * (glibc expf instead of or_expf, for the reference-faithful timing leg).
 * Prints the reference's stdout marker and "loop_seconds=<s>" at the end.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

int main(int argc, char **argv) {
  or_params p;
  memset(&p, 0, sizeof p);
  p.M = 1000;                          /* cu:25 */
  p.step_size = 3600.0f;               /* cu:26 */
  p.iterations = 42;                   /* cu:27 */
  p.start_time = 946713600;            /* cu:55 */
  p.static_network_duration = 3600;    /* cu:56 */
  p.threads = 1;
  const char *graph = NULL, *summary = NULL, *node = NULL;
  for (int i = 1; i < argc; i++) {
    const char *a = argv[i];
    int has = i + 1 < argc;
    if (a[0] != '-') {
      graph = a;
    } else if (!strcmp(a, "--cpu-only") || !strcmp(a, "-q") ||
               !strcmp(a, "--disable-debug-prints")) {
    } else if (!strcmp(a, "--libm-exp")) {
      p.use_libm_exp = 1;
    } else if (!strcmp(a, "--cpu-threads") && has) {
      p.threads = atoi(argv[++i]);
    } else if (!strcmp(a, "--M") && has) {
      p.M = atoi(argv[++i]);
    } else if (!strcmp(a, "--step-size") && has) {
      p.step_size = (float)atof(argv[++i]);
    } else if (!strcmp(a, "--iterations") && has) {
      p.iterations = atoi(argv[++i]);
    } else if (!strcmp(a, "--summary-out") && has) {
      summary = argv[++i];
    } else if (!strcmp(a, "--node-state-out") && has) {
      node = argv[++i];
    } else if (!strcmp(a, "--initial-infected") && has) {
      p.initial_infected = (float)atof(argv[++i]);
    } else if (!strcmp(a, "--infect-prob") && has) {
      p.infect_prob = (float)atof(argv[++i]);
    } else if (!strcmp(a, "--exposed-duration") && has) {
      p.exposed_duration = atoi(argv[++i]);
    } else if (!strcmp(a, "--infectious-duration") && has) {
      p.infectious_duration = atoi(argv[++i]);
    } else if (!strcmp(a, "--resistant-duration") && has) {
      p.resistant_duration = atoi(argv[++i]);
    } else if (!strcmp(a, "--start-time") && has) {
      p.start_time = (int64_t)atoll(argv[++i]);
    } else if ((!strcmp(a, "--static-network-duration") || !strcmp(a, "--time-step")) && has) {
      p.static_network_duration = (int64_t)atoll(argv[++i]);
    } else if (!strcmp(a, "--seed") && has) {
      p.seed = (uint64_t)strtoull(argv[++i], NULL, 10);
    } else {
      fprintf(stderr, "Unrecognized option: %s\n", a);
      return 1;
    }
  }
  if (!graph) {
    fprintf(stderr, "Missing contact network file argument.\n");
    return 1;
  }
  or_graph *g = NULL;
  int rc = or_parse_file(graph, &g);
  if (rc) {
    fprintf(stderr, "No intervals read or parse error (%d).\n", rc);
    return 1;
  }
  rc = or_build_ncls(g);
  if (rc) {
    fprintf(stderr, "Error: Failed to build interval DB.\n");
    return 1;
  }
  printf("Starting temporal Monte Carlo simulation for %d iterations...\n", p.iterations);
  fflush(stdout);
  double loop = 0.0;
  rc = or_simulate_csv(g, &p, summary, node, NULL, &loop);
  or_graph_free(g);
  if (rc) {
    fprintf(stderr, "Error: simulation failed (%d).\n", rc);
    return 1;
  }
  printf("Simulation done.\nloop_seconds=%.6f\n", loop);
  return 0;
}

#!/bin/bash
# Writes the generator golden files from the reference's group-structured
# generator (generate_networks/graph_generator.c + main.c), built into
# oracle/_ref/graph_gen by `make -C oracle ref`.  TZ=UTC: the reference's
# main() reads the start time as local time.
set -e
cd "$(dirname "$0")"
G=../../../oracle/_ref/graph_gen
TZ=UTC $G "2000-01-01T00:00:00" 5 60 40 4 0.3 0.05 groups_40n_4g_5s_seed12.txt 12 > /dev/null
TZ=UTC $G "2021-06-30T23:59:00" 3 7 25 1 0.5 0.5 groups_25n_1g_3s_seed0.txt 0 > /dev/null

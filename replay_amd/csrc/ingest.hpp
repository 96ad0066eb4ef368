// ingest.hpp -- host-side ingestion for libreplay_hip.
//
// Parses the Replay temporal contact format (NODE_LIST header + "u, v, ISO,
// ISO;" lines), assigns node indices in order of first appearance, ranks the
// contacts the way the reference's nested containment list enumerates them,
// plans the per-iteration network windows, and packs the contacts into
// epoch slices that the HIP kernels turn into per-window weights on the fly.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "replay_hip.h"

namespace rp {

struct Contact {
  int64_t start;
  int64_t end;
  int32_t a;  // smaller node index of the pair
  int32_t b;  // larger node index of the pair
};

struct Network {
  int32_t n_nodes = 0;                 // NODE_LIST count
  int32_t n_indexed = 0;               // distinct endpoint indices
  std::vector<Contact> contacts;       // rank order after finalize()
  std::vector<int64_t> index_to_id;    // node index -> id in the file
  int64_t nlists = 0;                  // hit cap of every window query
  int64_t n_top = 0;                   // top-level list length
  bool well_formed = true;             // every contact has start <= end
};

struct Window {
  int64_t ws;
  int64_t we;
  float scale;          // 1 / iterations sharing the window
  int32_t steps;        // iterations sharing the window
  int64_t hits;         // intervals the reference query returns (after cap)
  uint32_t rank_limit;  // contacts with rank > limit are cut by the cap
  int32_t epoch;
  int32_t int_mode;     // 1: weights are integers and row sums < 2^24
};

struct Plan {
  std::vector<int64_t> iter_time;    // current_time_ts of each iteration
  std::vector<int32_t> iter_window;  // window index of each iteration
  std::vector<Window> windows;       // distinct windows in time order
};

// One packed entry: contact seen from `row`, neighbour `col`, its interval
// clipped to the epoch and made relative to the epoch base time, and its rank.
// The empty constructor leaves a new entry unwritten, so sizing an array of
// hundreds of millions of them does not first zero it: every entry is
// written by the packing pass.
struct Entry {
  int32_t col;
  int32_t ts;
  int32_t te;
  uint32_t rank;
  Entry() {}
};

// Rows with more entries than this in an epoch are "long": the step kernels
// skip them and a long-row kernel gives each of their 64-replica words its
// own wavefront, so a hub's contact list does not serialise on one wavefront.
constexpr int32_t kLongRow = 2048;

struct Slices {
  int32_t windows_per_epoch = 1;
  std::vector<int64_t> epoch_t0;       // base time of each epoch
  std::vector<int64_t> epoch_entry0;   // first entry of each epoch
  std::vector<int32_t> rowptr;         // (n_epochs) x (N+1), epoch-relative
  std::vector<Entry> entries;
  int64_t max_row_entries = 0;         // longest row that is not long
  // long rows of each epoch: rows long_row[long_off[e] .. long_off[e+1]),
  // long_int = 1 where the row's weight bound allows exact integer sums
  std::vector<int64_t> long_off;
  std::vector<int32_t> long_row;
  std::vector<uint8_t> long_int;
};

int parse_file(const char *path, Network &net, std::string &err);
// One contact line by the fast canonical-form parser (fast != 0; -1 when the
// line is not in that form) or by the reference's sscanf patterns; returns
// whether the line yields a contact, and sets *counted as tcm:131-135 would.
int parse_line(const char *line, int64_t len, int fast, int32_t *src, int32_t *tgt,
               int64_t *start, int64_t *end, int *counted);
int from_contacts(int32_t n_list, const int64_t *node_ids, int64_t n,
                  const int32_t *src, const int32_t *tgt, const int64_t *start,
                  const int64_t *end, Network &net, std::string &err);
// Sorts into rank order and computes the nested-list hit cap.
int finalize(Network &net, std::string &err);
int plan_windows(const Network &net, const replay_params &p, Plan &plan,
                 std::string &err);
// windows_per_epoch <= 0 picks a size that keeps the packed bytes under
// max_bytes (or the smallest per-step footprint when everything fits).
int build_slices(const Network &net, Plan &plan, int32_t windows_per_epoch,
                 int64_t max_bytes, Slices &out, std::string &err);
// Packed bytes of the slices for each candidate windows-per-epoch value.
std::vector<int64_t> slice_bytes(const Network &net, const Plan &plan,
                                 const std::vector<int64_t> &candidates);

}  // namespace rp

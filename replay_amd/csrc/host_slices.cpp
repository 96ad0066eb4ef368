// host_slices.cpp -- host evaluation of the packed slices, for CPU tests.
//
// Not part of the public ABI (include/replay_hip.h): these exports let the
// test suite check ingestion, window planning and slice packing on a machine
// without a GPU.  rp_host_slice evaluates one iteration's CSR from the packed
// entries with the same rules the device slice kernel applies (overlap with
// the window, rank limit of the hit cap, rank-order merge of one column).
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>

#include "ingest.hpp"

struct rp_host_plan {
  rp::Network net;
  rp::Plan plan;
  rp::Slices slices;
  std::string err;
};

extern "C" {

int32_t rp_host_plan_open(int32_t n_node_list, int64_t n, const int32_t *src, const int32_t *tgt,
                          const int64_t *start, const int64_t *end, const replay_params *p,
                          rp_host_plan **out) {
  auto *h = new rp_host_plan();
  int rc = rp::from_contacts(n_node_list, nullptr, n, src, tgt, start, end, h->net, h->err);
  if (!rc) rc = rp::finalize(h->net, h->err);
  if (!rc) rc = rp::plan_windows(h->net, *p, h->plan, h->err);
  if (!rc)
    rc = rp::build_slices(h->net, h->plan, p->windows_per_epoch,
                          p->max_slice_bytes > 0 ? p->max_slice_bytes : (int64_t(1) << 32),
                          h->slices, h->err);
  if (rc) {
    delete h;
    return rc;
  }
  *out = h;
  return 0;
}

int32_t rp_host_parse_line(const char *line, int64_t len, int32_t fast, int32_t *src,
                           int32_t *tgt, int64_t *start, int64_t *end, int32_t *counted) {
  int c = 0;
  const int r = rp::parse_line(line, len, fast, src, tgt, start, end, &c);
  *counted = c;
  return r;
}

// Parses a graph file and ranks its contacts: out[k] = {start, end, a, b}
// in rank order; returns the count, or a negative error code.
int64_t rp_host_parse_file(const char *path, int64_t *out, int64_t cap) {
  rp::Network net;
  std::string err;
  int rc = rp::parse_file(path, net, err);
  if (!rc) rc = rp::finalize(net, err);
  if (rc) return rc;
  const int64_t n = static_cast<int64_t>(net.contacts.size());
  for (int64_t k = 0; k < n && k < cap; ++k) {
    out[4 * k] = net.contacts[k].start;
    out[4 * k + 1] = net.contacts[k].end;
    out[4 * k + 2] = net.contacts[k].a;
    out[4 * k + 3] = net.contacts[k].b;
  }
  return n;
}

int64_t rp_host_plan_nlists(rp_host_plan *h) { return h->net.nlists; }
int32_t rp_host_plan_epochs(rp_host_plan *h) { return static_cast<int32_t>(h->slices.epoch_t0.size()); }
int32_t rp_host_plan_int_mode(rp_host_plan *h, int32_t k) {
  return h->plan.windows[h->plan.iter_window[k]].int_mode;
}

int32_t rp_host_slice(rp_host_plan *h, int32_t k, int32_t *rowptr, int32_t *colind, float *val,
                      int64_t cap, int64_t *nnz) {
  const rp::Window &w = h->plan.windows[h->plan.iter_window[k]];
  const int64_t N = h->net.n_nodes;
  const int32_t *rp = &h->slices.rowptr[static_cast<size_t>(w.epoch) * (N + 1)];
  const rp::Entry *en = h->slices.entries.data() + h->slices.epoch_entry0[w.epoch];
  const int64_t t0 = h->slices.epoch_t0[w.epoch];
  const int32_t ws = static_cast<int32_t>(w.ws - t0), we = static_cast<int32_t>(w.we - t0);
  int64_t out = 0;
  rowptr[0] = 0;
  for (int64_t i = 0; i < N; ++i) {
    int32_t e = rp[i];
    while (e < rp[i + 1]) {
      const int32_t col = en[e].col;
      float wt = 0.0f;
      bool any = false;
      for (; e < rp[i + 1] && en[e].col == col; ++e) {
        const int32_t ov = std::min(en[e].te, we) - std::max(en[e].ts, ws);
        if (ov > 0 && en[e].rank <= w.rank_limit) {
          const float wc = static_cast<float>(ov) * w.scale;
          wt = any ? wt + wc : wc;
          any = true;
        }
      }
      if (any) {
        if (out >= cap) return -1;
        colind[out] = col;
        val[out] = wt;
        ++out;
      }
    }
    rowptr[i + 1] = static_cast<int32_t>(out);
  }
  *nnz = out;
  return 0;
}

// Packed entries of epoch e as {col, ts, te, rank} rows of 4 int64, rowptr
// (N+1 entries, epoch-relative) into rowptr_out; returns the entry count, or
// -1 when cap is too small.
int64_t rp_host_plan_epoch(rp_host_plan *h, int32_t e, int32_t *rowptr_out, int64_t *out,
                           int64_t cap) {
  const int64_t N = h->net.n_nodes;
  const int64_t b0 = h->slices.epoch_entry0[e], b1 = h->slices.epoch_entry0[e + 1];
  if (b1 - b0 > cap) return -1;
  std::copy(h->slices.rowptr.begin() + static_cast<size_t>(e) * (N + 1),
            h->slices.rowptr.begin() + static_cast<size_t>(e + 1) * (N + 1), rowptr_out);
  for (int64_t q = b0; q < b1; ++q) {
    const rp::Entry &en = h->slices.entries[static_cast<size_t>(q)];
    int64_t *o = out + 4 * (q - b0);
    o[0] = en.col;
    o[1] = en.ts;
    o[2] = en.te;
    o[3] = en.rank;
  }
  return b1 - b0;
}

// Long rows of epoch e (rows with more than the long-row threshold of
// entries) and their integer-sum flags; returns their count.
int64_t rp_host_plan_long_rows(rp_host_plan *h, int32_t e, int32_t *rows, uint8_t *int_ok,
                               int64_t cap) {
  const auto &sl = h->slices;
  const int64_t q0 = sl.long_off[e], q1 = sl.long_off[e + 1];
  for (int64_t q = q0; q < q1 && q - q0 < cap; ++q) {
    rows[q - q0] = sl.long_row[static_cast<size_t>(q)];
    int_ok[q - q0] = sl.long_int[static_cast<size_t>(q)];
  }
  return q1 - q0;
}

int32_t rp_host_long_row_threshold() { return rp::kLongRow; }

void rp_host_plan_close(rp_host_plan *h) { delete h; }

}  // extern "C"

// ingest.cpp -- host-side ingestion for libreplay_hip (see ingest.hpp).
//
// Reference behaviour kept (paths relative to the reference repository):
//   tool/temporal_contact_matrix.c:148-321  header, line patterns, skip rules,
//       first-appearance node indices (tcm:294-295, 349), pair packing (80-85);
//   tool/ncls/ncls/src/intervaldb.c:30-46, 227-326  contact order and the
//       modified sublist count used as the per-query hit cap (258-261);
//   tool/gpu_cpu_temporal_sim.cu:860-959  window arithmetic and divisor;
//   tool/gpu_cpu_temporal_sim.cu:973-989  hits capped at nlists per query.
//
// For contacts with start <= end the nested containment list's depth-first
// enumeration of the intervals overlapping [ws, we) is exactly the sorted
// order (start ascending, end descending, ties in file order) restricted to
// those intervals: every list is sorted by start with non-decreasing ends,
// children are contained in their parent and follow it.  So a window's hits
// are "contacts overlapping the window, in rank order, first nlists of them",
// which the packed slices evaluate per window with a rank limit.
#include "ingest.hpp"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <thread>

namespace rp {

namespace {

constexpr int kMaxLineLen = 100000;  // tcm:20, the fgets buffer

// open-addressing map: node id -> index in first-appearance order
class IdMap {
 public:
  explicit IdMap(size_t hint) {
    size_t cap = 64;
    while (cap < hint * 2) cap <<= 1;
    keys_.assign(cap, 0);
    vals_.assign(cap, -1);
  }
  // returns {index, created}
  std::pair<int32_t, bool> get_or_create(int64_t key) {
    if ((count_ + 1) * 2 > keys_.size()) grow();
    size_t mask = keys_.size() - 1;
    size_t h = hash(key) & mask;
    while (vals_[h] >= 0) {
      if (keys_[h] == key) return {vals_[h], false};
      h = (h + 1) & mask;
    }
    keys_[h] = key;
    vals_[h] = static_cast<int32_t>(count_);
    ++count_;
    return {vals_[h], true};
  }
  size_t size() const { return count_; }

 private:
  static uint64_t hash(int64_t k) {
    uint64_t x = static_cast<uint64_t>(k);
    x ^= x >> 31;
    x *= 0x9e3779b97f4a7c15ULL;
    x ^= x >> 29;
    return x;
  }
  void grow() {
    std::vector<int64_t> ok;
    std::vector<int32_t> ov;
    ok.swap(keys_);
    ov.swap(vals_);
    keys_.assign(ok.size() * 2, 0);
    vals_.assign(ok.size() * 2, -1);
    size_t mask = keys_.size() - 1;
    for (size_t i = 0; i < ok.size(); ++i) {
      if (ov[i] < 0) continue;
      size_t h = hash(ok[i]) & mask;
      while (vals_[h] >= 0) h = (h + 1) & mask;
      keys_[h] = ok[i];
      vals_[h] = ov[i];
    }
  }
  std::vector<int64_t> keys_;
  std::vector<int32_t> vals_;
  size_t count_ = 0;
};

// tcm:95-120: six integers, then UTC conversion with field normalisation.
int64_t parse_time(const char *s) {
  int year, month, day, hour, min, sec;
  char extra[10];
  if (std::sscanf(s, "%d-%d-%dT%d:%d:%d%9s", &year, &month, &day, &hour, &min,
                  &sec, extra) < 6)
    return -1;
  struct tm t;
  std::memset(&t, 0, sizeof t);
  t.tm_year = year - 1900;
  t.tm_mon = month - 1;
  t.tm_mday = day;
  t.tm_hour = hour;
  t.tm_min = min;
  t.tm_sec = sec;
  return static_cast<int64_t>(timegm(&t));
}

// Days from 1970-01-01 to year-month-01 of the proleptic Gregorian
// calendar (H. Hinnant's days_from_civil), month 1..12.
int64_t days_from_civil(int64_t y, int64_t m) {
  y -= m <= 2;
  const int64_t era = (y >= 0 ? y : y - 399) / 400;
  const int64_t yoe = y - era * 400;
  const int64_t doy = (153 * (m + (m > 2 ? -3 : 9)) + 2) / 5;
  const int64_t doe = yoe * 365 + yoe / 4 - yoe / 100 + doy;
  return era * 146097 + doe - 719468;
}

// Fixed-width digits at p (n of them); -1 when one is not a digit.
inline int digits(const char *p, int n) {
  int v = 0;
  for (int i = 0; i < n; ++i) {
    const unsigned d = static_cast<unsigned>(p[i] - '0');
    if (d > 9) return -1;
    v = v * 10 + static_cast<int>(d);
  }
  return v;
}

// "YYYY-MM-DDTHH:MM:SS" then a suffix of non-space characters other than
// ',' and ';' up to `stop`.  Returns the end of the suffix, or nullptr when
// the text is not in this form or the month is outside 1..12.  Within the
// form, parse_time's sscanf reads the same six fields and timegm's
// normalisation of day, hour, minute and second overflows is linear, so
// the result equals parse_time's.
const char *fast_time(const char *p, const char *end, char stop, int64_t *out) {
  if (end - p < 19) return nullptr;
  if (p[4] != '-' || p[7] != '-' || p[10] != 'T' || p[13] != ':' || p[16] != ':') return nullptr;
  const int y = digits(p, 4), mo = digits(p + 5, 2), d = digits(p + 8, 2);
  const int h = digits(p + 11, 2), mi = digits(p + 14, 2), se = digits(p + 17, 2);
  if (y < 0 || mo < 1 || mo > 12 || d < 0 || h < 0 || mi < 0 || se < 0) return nullptr;
  const char *q = p + 19;
  if (q < end && static_cast<unsigned>(*q - '0') <= 9) return nullptr;  // a third seconds digit
  while (q < end && *q != stop) {
    if (*q == ' ' || *q == '\t' || *q == '\r' || *q == '\n' || *q == ',' || *q == ';') return nullptr;
    ++q;
  }
  if (q >= end || q - p > 63) return nullptr;
  *out = (days_from_civil(y, mo) + (d - 1)) * 86400 + h * 3600 + mi * 60 + se;
  return q;
}

// A non-negative id of at most 9 digits immediately followed by ','.
const char *fast_id(const char *p, const char *end, int32_t *out) {
  int v = 0, n = 0;
  while (p < end && static_cast<unsigned>(*p - '0') <= 9 && n < 10) {
    v = v * 10 + (*p - '0');
    ++p;
    ++n;
  }
  if (n == 0 || n > 9 || p >= end || *p != ',') return nullptr;
  *out = v;
  return p + 1;
}

inline const char *skip_blank(const char *p, const char *end) {
  while (p < end && (*p == ' ' || *p == '\t')) ++p;
  return p;
}

// The canonical line "u, v, TIME, TIME;" (blanks allowed after the line
// start and after each comma, none before a comma or the ';').  When it
// matches, both sscanf patterns of scan_lines match with the same values.
bool fast_line(const char *p, const char *end, int32_t *src, int32_t *tgt, int64_t *ts,
               int64_t *te) {
  p = skip_blank(p, end);
  if (!(p = fast_id(p, end, src))) return false;
  p = skip_blank(p, end);
  if (!(p = fast_id(p, end, tgt))) return false;
  p = skip_blank(p, end);
  if (!(p = fast_time(p, end, ',', ts))) return false;
  p = skip_blank(p + 1, end);
  if (!(p = fast_time(p, end, ';', te))) return false;
  return *ts != -1 && *te != -1;  // -1 is parse_time's failure value
}

struct LineResult {
  int64_t start;
  int64_t end;
  int32_t src;
  int32_t tgt;
  uint8_t counted;  // matches the counting pattern of tcm:135
  uint8_t valid;    // parsed (tcm:278) and both times convert (tcm:285-288)
};

// tcm:131-135 (counting pattern) and tcm:272-295 (parse, then both times)
// on one NUL-terminated line.
void slow_line(const char *line, LineResult &r) {
  r.counted = 0;
  r.valid = 0;
  if (line[0] == '#') return;
  int d1, d2;
  char s[64], e[64];
  if (std::sscanf(line, " %d, %d, %63[^,], %63[^;];", &d1, &d2, s, e) == 4) r.counted = 1;
  int src, tgt;
  if (std::sscanf(line, " %d , %d , %63[^,] , %63[^;\r\n] ;", &src, &tgt, s, e) != 4) return;
  int64_t ts = parse_time(s);
  int64_t te = parse_time(e);
  if (ts == -1 || te == -1) return;
  r.src = src;
  r.tgt = tgt;
  r.start = ts;
  r.end = te;
  r.valid = 1;
}

void scan_lines(const char *data, const std::vector<size_t> &line_start,
                const std::vector<uint32_t> &line_len, size_t lo, size_t hi,
                LineResult *out) {
  std::vector<char> buf(kMaxLineLen + 1);
  for (size_t k = lo; k < hi; ++k) {
    LineResult &r = out[k];
    r.counted = 0;
    r.valid = 0;
    const uint32_t len = line_len[k];
    if (len > 0 && data[line_start[k]] != '#' &&
        fast_line(data + line_start[k], data + line_start[k] + len, &r.src, &r.tgt, &r.start,
                  &r.end)) {
      r.counted = 1;
      r.valid = 1;
      continue;
    }
    std::memcpy(buf.data(), data + line_start[k], len);
    buf[len] = '\0';
    slow_line(buf.data(), r);
  }
}

int set_err(std::string &err, int code, const std::string &msg) {
  err = msg;
  return code;
}

// Appends a contact in first-appearance index order (tcm:294-300).
int add_contact(Network &net, IdMap &ids, int32_t src, int32_t tgt,
                int64_t start, int64_t end, std::string &err) {
  auto s = ids.get_or_create(src);
  if (s.second) net.index_to_id.push_back(src);
  auto t = ids.get_or_create(tgt);
  if (t.second) net.index_to_id.push_back(tgt);
  if (static_cast<int64_t>(ids.size()) > net.n_nodes)
    return set_err(err, REPLAY_E_NODE_RANGE,
                   "contact endpoints name more distinct nodes than NODE_LIST holds");
  Contact c;
  c.start = start;
  c.end = end;
  c.a = std::min(s.first, t.first);
  c.b = std::max(s.first, t.first);
  net.contacts.push_back(c);
  return REPLAY_OK;
}

unsigned worker_count() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 1;
  return std::min(n, 16u);
}

// Runs f(t, lo, hi) over contiguous chunks [lo, hi) of [0, n), one per
// worker thread t; inline when n has fewer than two chunks of min_chunk.
template <typename F>
void par_chunks(int64_t n, int64_t min_chunk, F f) {
  const int64_t T = std::max<int64_t>(
      1, std::min<int64_t>(worker_count(), n / std::max<int64_t>(min_chunk, 1)));
  if (T <= 1) {
    f(0, int64_t(0), n);
    return;
  }
  std::vector<std::thread> th;
  for (int64_t t = 0; t < T; ++t)
    th.emplace_back([&f, t, n, T] { f(static_cast<int>(t), n * t / T, n * (t + 1) / T); });
  for (auto &x : th) x.join();
}

// Stable sort on the worker threads: chunks are stable-sorted, then merged
// in pairs (std::merge takes the left run's element on ties), which yields
// the one order a stable sort defines.
template <typename T, typename C>
void par_stable_sort(std::vector<T> &v, C cmp) {
  const int64_t n = static_cast<int64_t>(v.size());
  const int64_t parts = std::min<int64_t>(worker_count(), n / 16384);
  if (parts <= 1) {
    std::stable_sort(v.begin(), v.end(), cmp);
    return;
  }
  std::vector<int64_t> bound(static_cast<size_t>(parts) + 1);
  for (int64_t t = 0; t <= parts; ++t) bound[static_cast<size_t>(t)] = n * t / parts;
  {
    std::vector<std::thread> th;
    for (int64_t t = 0; t < parts; ++t)
      th.emplace_back([&, t] {
        std::stable_sort(v.begin() + bound[static_cast<size_t>(t)],
                         v.begin() + bound[static_cast<size_t>(t) + 1], cmp);
      });
    for (auto &x : th) x.join();
  }
  std::vector<T> buf(v.size());
  while (bound.size() > 2) {
    const size_t runs = bound.size() - 1;
    std::vector<std::thread> th;
    for (size_t r = 0; r < runs; r += 2) {
      th.emplace_back([&, r] {
        const auto a = v.begin() + bound[r], m = v.begin() + bound[r + 1];
        if (r + 1 < runs)
          std::merge(a, m, m, v.begin() + bound[r + 2], buf.begin() + bound[r], cmp);
        else
          std::copy(a, m, buf.begin() + bound[r]);
      });
    }
    for (auto &x : th) x.join();
    std::vector<int64_t> next;
    for (size_t r = 0; r < runs; r += 2) next.push_back(bound[r]);
    next.push_back(bound.back());
    bound.swap(next);
    v.swap(buf);
  }
}

}  // namespace

int parse_line(const char *line, int64_t len, int fast, int32_t *src, int32_t *tgt,
               int64_t *start, int64_t *end, int *counted) {
  LineResult r{};
  if (fast) {
    if (len <= 0 || line[0] == '#' || !fast_line(line, line + len, &r.src, &r.tgt, &r.start, &r.end))
      return -1;
    r.counted = r.valid = 1;
  } else {
    std::string buf(line, static_cast<size_t>(len));
    slow_line(buf.c_str(), r);
  }
  *src = r.src;
  *tgt = r.tgt;
  *start = r.start;
  *end = r.end;
  *counted = r.counted;
  return r.valid;
}

int parse_file(const char *path, Network &net, std::string &err) {
  FILE *fp = std::fopen(path, "rb");
  if (!fp) return set_err(err, REPLAY_E_IO, std::string("could not open file '") + path + "'");
  std::string data;
  {
    std::fseek(fp, 0, SEEK_END);
    long sz = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    if (sz < 0) {
      std::fclose(fp);
      return set_err(err, REPLAY_E_IO, "cannot size file");
    }
    data.resize(static_cast<size_t>(sz));
    if (sz > 0 && std::fread(&data[0], 1, data.size(), fp) != data.size()) {
      std::fclose(fp);
      return set_err(err, REPLAY_E_IO, "short read");
    }
    std::fclose(fp);
  }
  // Header: the first line, read whole (tcm:156-177).
  size_t nl = data.find('\n');
  size_t hlen = (nl == std::string::npos) ? data.size() : nl + 1;
  if (hlen == 0) return set_err(err, REPLAY_E_FORMAT, "unable to read first line from file");
  std::string header = data.substr(0, hlen);
  if (header.compare(0, 9, "NODE_LIST") != 0 || header.size() < 10 ||
      (header[9] != ' ' && header[9] != '\t'))
    return set_err(err, REPLAY_E_FORMAT, "first line must begin with 'NODE_LIST'");
  // Ids: strtok over " \t\r\n", strtol base 10, >= 0, no duplicates
  // (tcm:180-232).  The reference compares the stored int with the long.
  {
    std::vector<char> hb(header.begin() + 9, header.end());
    hb.push_back('\0');
    char *save = nullptr;
    char *tok = strtok_r(hb.data(), " \t\r\n", &save);
    IdMap seen(1024);
    int32_t count = 0;
    while (tok) {
      errno = 0;
      char *endp = nullptr;
      long id = std::strtol(tok, &endp, 10);
      if (errno != 0 || *endp != '\0' || id < 0)
        return set_err(err, REPLAY_E_NODE_ID, std::string("invalid node ID in NODE_LIST: '") + tok + "'");
      // stored values are ints; a long id only equals a stored int if it fits
      const int stored = static_cast<int>(id);
      const bool fits = static_cast<long>(stored) == id;
      const bool created = seen.get_or_create(stored).second;
      if (fits && !created)
        return set_err(err, REPLAY_E_DUP_ID, "duplicate node ID detected in NODE_LIST: " + std::to_string(id));
      ++count;
      tok = strtok_r(nullptr, " \t\r\n", &save);
    }
    if (count <= 0) return set_err(err, REPLAY_E_NODE_ID, "no valid node IDs found in NODE_LIST");
    net.n_nodes = count;
  }
  // fgets chunking: a "line" ends after '\n' or after kMaxLineLen-1 chars.
  std::vector<size_t> starts;
  std::vector<uint32_t> lens;
  {
    size_t pos = 0;
    const size_t n = data.size();
    starts.reserve(n / 48 + 16);
    lens.reserve(n / 48 + 16);
    while (pos < n) {
      size_t lim = std::min(n, pos + static_cast<size_t>(kMaxLineLen - 1));
      const void *p = std::memchr(data.data() + pos, '\n', lim - pos);
      size_t end = p ? static_cast<size_t>(static_cast<const char *>(p) - data.data()) + 1 : lim;
      starts.push_back(pos);
      lens.push_back(static_cast<uint32_t>(end - pos));
      pos = end;
    }
  }
  const size_t nlines = starts.size();
  std::vector<LineResult> res(nlines);
  {
    unsigned nw = worker_count();
    if (nlines < 4096) nw = 1;
    std::vector<std::thread> th;
    size_t per = (nlines + nw - 1) / nw;
    for (unsigned w = 0; w < nw; ++w) {
      size_t lo = w * per, hi = std::min(nlines, lo + per);
      if (lo >= hi) break;
      th.emplace_back(scan_lines, data.data(), std::cref(starts), std::cref(lens), lo, hi, res.data());
    }
    for (auto &t : th) t.join();
  }
  int64_t edge_count = 0;
  for (size_t k = 0; k < nlines; ++k) edge_count += res[k].counted;
  if (edge_count <= 0) return set_err(err, REPLAY_E_NO_EDGES, std::string("no valid edge lines found in ") + path);
  IdMap ids(static_cast<size_t>(net.n_nodes));
  net.contacts.clear();
  net.contacts.reserve(static_cast<size_t>(edge_count));
  for (size_t k = 0; k < nlines; ++k) {
    const LineResult &r = res[k];
    if (!r.valid) continue;
    int rc = add_contact(net, ids, r.src, r.tgt, r.start, r.end, err);
    if (rc != REPLAY_OK) return rc;
    if (static_cast<int64_t>(net.contacts.size()) > edge_count)
      return set_err(err, REPLAY_E_COUNT, "more edges parsed than expected");
  }
  net.n_indexed = static_cast<int32_t>(ids.size());
  if (net.contacts.empty()) return set_err(err, REPLAY_E_NO_EDGES, "no intervals read");
  return REPLAY_OK;
}

int from_contacts(int32_t n_list, const int64_t *node_ids, int64_t n,
                  const int32_t *src, const int32_t *tgt, const int64_t *start,
                  const int64_t *end, Network &net, std::string &err) {
  if (n_list <= 0) return set_err(err, REPLAY_E_NODE_ID, "empty node list");
  if (node_ids) {
    IdMap seen(static_cast<size_t>(n_list));
    for (int32_t i = 0; i < n_list; ++i) {
      if (node_ids[i] < 0) return set_err(err, REPLAY_E_NODE_ID, "negative node id");
      if (!seen.get_or_create(node_ids[i]).second)
        return set_err(err, REPLAY_E_DUP_ID, "duplicate node id " + std::to_string(node_ids[i]));
    }
  }
  net.n_nodes = n_list;
  if (n <= 0 || !src || !tgt || !start || !end) return set_err(err, REPLAY_E_NO_EDGES, "no contacts");
  IdMap ids(static_cast<size_t>(n_list));
  net.contacts.clear();
  net.contacts.reserve(static_cast<size_t>(n));
  for (int64_t k = 0; k < n; ++k) {
    if (start[k] == -1 || end[k] == -1) continue;
    int rc = add_contact(net, ids, src[k], tgt[k], start[k], end[k], err);
    if (rc != REPLAY_OK) return rc;
  }
  net.n_indexed = static_cast<int32_t>(ids.size());
  if (net.contacts.empty()) return set_err(err, REPLAY_E_NO_EDGES, "no intervals read");
  return REPLAY_OK;
}

int finalize(Network &net, std::string &err) {
  auto &cs = net.contacts;
  const int64_t n = static_cast<int64_t>(cs.size());
  for (const auto &c : cs) {
    if (c.start < 0 || c.end < 0)
      return set_err(err, REPLAY_E_UNSUPPORTED,
                     "contact times before 1970 take the reference index's mirrored-"
                     "orientation path (intervaldb.c:412-424) and are not supported");
    if (c.end < c.start) {
      net.well_formed = false;
      return set_err(err, REPLAY_E_UNSUPPORTED,
                     "a contact ends before it starts; such intervals reorder the "
                     "reference's interval index and are not supported");
    }
  }
  // Rank order: start ascending, end descending, file order on ties
  // (idb:30-46 under glibc 2.35's stable merge-sort qsort).
  par_stable_sort(cs, [](const Contact &x, const Contact &y) {
    if (x.start != y.start) return x.start < y.start;
    return x.end > y.end;
  });
  // Hit cap: idb:243-262 counts a closed top-level chain per new top-level
  // interval; idb:276-281 then adds one per interval that has children.
  std::vector<int32_t> parent(static_cast<size_t>(n), -1);
  std::vector<uint8_t> has_child(static_cast<size_t>(n), 0);
  int64_t closures = 0, nsub = 0;
  int64_t i = 0;
  auto not_contained = [&](int64_t x, int64_t p) {
    return cs[x].end > cs[p].end || (cs[x].end == cs[p].end && cs[x].start == cs[p].start);
  };
  while (i < n) {
    int64_t p = i;
    i = p + 1;
    while (i < n && p >= 0) {
      if (not_contained(i, p)) {
        p = parent[p];
      } else {
        parent[i] = static_cast<int32_t>(p);
        has_child[p] = 1;
        ++nsub;
        p = i;
        ++i;
      }
    }
    if (p < 0) ++closures;
  }
  int64_t parents = 0;
  for (int64_t k = 0; k < n; ++k) parents += has_child[k];
  net.nlists = closures + parents;
  net.n_top = n - nsub;
  if (n > static_cast<int64_t>(UINT32_MAX) - 1)
    return set_err(err, REPLAY_E_UNSUPPORTED, "too many contacts");
  return REPLAY_OK;
}

namespace {

// cu:860-902 for one iteration: offset, network window start.
int window_start(const replay_params &p, int32_t iter, int64_t *offset_out,
                 int64_t *ws_out) {
  double raw = static_cast<double>(iter) * static_cast<double>(p.step_size);
  if (raw > static_cast<double>(INT64_MAX) || raw < static_cast<double>(INT64_MIN))
    return REPLAY_E_OVERFLOW;
  int64_t offset = static_cast<int64_t>(std::llround(raw));
  if ((offset > 0 && p.start_time > INT64_MAX - offset) ||
      (offset < 0 && p.start_time < INT64_MIN - offset))
    return REPLAY_E_OVERFLOW;
  const int64_t D = p.static_network_duration;
  int64_t net = (offset / D) * D;
  if (offset < 0 && (offset % D)) net -= D;
  if ((net > 0 && p.start_time > INT64_MAX - net) ||
      (net < 0 && p.start_time < INT64_MIN - net))
    return REPLAY_E_OVERFLOW;
  int64_t ws = p.start_time + net;
  if (ws > INT64_MAX - D) return REPLAY_E_OVERFLOW;
  *offset_out = offset;
  *ws_out = ws;
  return REPLAY_OK;
}

}  // namespace

int plan_windows(const Network &net, const replay_params &p, Plan &plan,
                 std::string &err) {
  if (p.static_network_duration <= 0)
    return set_err(err, REPLAY_E_ARG, "static network duration must be positive");
  if (!(p.step_size > 0.0f)) return set_err(err, REPLAY_E_ARG, "step size must be positive");
  if (p.iterations < 0) return set_err(err, REPLAY_E_ARG, "negative iteration count");
  const int32_t T = p.iterations;
  plan.iter_time.assign(static_cast<size_t>(T), 0);
  plan.iter_window.assign(static_cast<size_t>(T), 0);
  plan.windows.clear();
  for (int32_t it = 0; it < T; ++it) {
    int64_t offset, ws;
    int rc = window_start(p, it, &offset, &ws);
    if (rc != REPLAY_OK) return set_err(err, rc, "overflow in window arithmetic");
    if (ws < 0)
      return set_err(err, REPLAY_E_UNSUPPORTED,
                     "windows starting before 1970 take the reference index's mirrored "
                     "query path (intervaldb.c:437-444) and are not supported");
    plan.iter_time[it] = p.start_time + offset;
    if (plan.windows.empty() || plan.windows.back().ws != ws) {
      Window w;
      std::memset(&w, 0, sizeof w);
      w.ws = ws;
      w.we = ws + p.static_network_duration;
      w.steps = 0;
      plan.windows.push_back(w);
    }
    plan.windows.back().steps += 1;
    plan.iter_window[it] = static_cast<int32_t>(plan.windows.size() - 1);
  }
  // Hits per window and the cap (cu:973-989 with idb:465-467: the buffer
  // check follows each store, so a cap of 0 still returns one hit).
  const auto &cs = net.contacts;
  const int64_t E = static_cast<int64_t>(cs.size());
  std::vector<int64_t> ends(static_cast<size_t>(E));
  for (int64_t k = 0; k < E; ++k) ends[k] = cs[k].end;
  par_stable_sort(ends, [](int64_t x, int64_t y) { return x < y; });
  const int64_t cap = std::max<int64_t>(net.nlists, 1);
  for (auto &w : plan.windows) {
    w.scale = 1.0f / static_cast<float>(w.steps);
    int64_t n_start_before = std::lower_bound(cs.begin(), cs.end(), w.we,
                                              [](const Contact &c, int64_t v) { return c.start < v; }) -
                             cs.begin();
    int64_t n_end_le = std::upper_bound(ends.begin(), ends.end(), w.ws) - ends.begin();
    int64_t hits = n_start_before - n_end_le;
    w.rank_limit = UINT32_MAX;
    if (hits > cap) {
      int64_t seen = 0;
      for (int64_t k = 0; k < n_start_before; ++k) {
        if (cs[k].end > w.ws) {
          if (++seen == cap) {
            w.rank_limit = static_cast<uint32_t>(k);
            break;
          }
        }
      }
      hits = cap;
    }
    w.hits = hits;
  }
  return REPLAY_OK;
}

namespace {

struct Span {
  int32_t a;  // first window overlapped
  int32_t b;  // last window overlapped
};

}  // namespace

std::vector<int64_t> slice_bytes(const Network &net, const Plan &plan,
                                 const std::vector<int64_t> &candidates) {
  const auto &cs = net.contacts;
  const int32_t W = static_cast<int32_t>(plan.windows.size());
  const int64_t N = net.n_nodes;
  std::vector<int64_t> out(candidates.size(), 0);
  if (W == 0) return out;
  std::vector<int64_t> ws(W), we(W);
  for (int32_t k = 0; k < W; ++k) {
    ws[k] = plan.windows[k].ws;
    we[k] = plan.windows[k].we;
  }
  std::vector<int64_t> ne(candidates.size(), 0);
  for (const Contact &c : cs) {
    if (c.end <= c.start) continue;
    const int64_t a = std::upper_bound(we.begin(), we.end(), c.start) - we.begin();
    const int64_t b = (std::lower_bound(ws.begin(), ws.end(), c.end) - ws.begin()) - 1;
    if (a > b) continue;
    for (size_t j = 0; j < candidates.size(); ++j)
      ne[j] += 2 * (b / candidates[j] - a / candidates[j] + 1);
  }
  for (size_t j = 0; j < candidates.size(); ++j) {
    const int64_t S = candidates[j];
    out[j] = ne[j] * static_cast<int64_t>(sizeof(Entry)) + ((W + S - 1) / S) * (N + 1) * 4;
  }
  return out;
}

int build_slices(const Network &net, Plan &plan, int32_t windows_per_epoch,
                 int64_t max_bytes, Slices &out, std::string &err) {
  const auto &cs = net.contacts;
  const int64_t E = static_cast<int64_t>(cs.size());
  const int32_t W = static_cast<int32_t>(plan.windows.size());
  const int64_t N = net.n_nodes;
  out = Slices();
  if (W == 0) {
    out.windows_per_epoch = 1;
    return REPLAY_OK;
  }
  std::vector<int64_t> ws(W), we(W);
  for (int32_t k = 0; k < W; ++k) {
    ws[k] = plan.windows[k].ws;
    we[k] = plan.windows[k].we;
  }
  // Window span of every contact with positive length (zero-length contacts
  // count as hits but never carry weight).
  std::vector<Span> span(static_cast<size_t>(E));
  par_chunks(E, 1 << 14, [&](int, int64_t lo, int64_t hi) {
    for (int64_t k = lo; k < hi; ++k) {
      const Contact &c = cs[k];
      Span s{1, 0};
      if (c.end > c.start) {
        s.a = static_cast<int32_t>(std::upper_bound(we.begin(), we.end(), c.start) - we.begin());
        s.b = static_cast<int32_t>(std::lower_bound(ws.begin(), ws.end(), c.end) - ws.begin()) - 1;
      }
      span[k] = s;
    }
  });
  auto bytes_for = [&](int64_t S) {
    std::vector<int64_t> part(worker_count(), 0);
    par_chunks(E, 1 << 14, [&](int t, int64_t lo, int64_t hi) {
      int64_t ne = 0;
      for (int64_t k = lo; k < hi; ++k)
        if (span[k].a <= span[k].b) ne += 2 * (span[k].b / S - span[k].a / S + 1);
      part[static_cast<size_t>(t)] = ne;
    });
    int64_t ne = 0;
    for (int64_t x : part) ne += x;
    int64_t ep = (W + S - 1) / S;
    return ne * static_cast<int64_t>(sizeof(Entry)) + ep * (N + 1) * 4;
  };
  if (max_bytes <= 0) max_bytes = int64_t(32) << 30;
  int64_t S = windows_per_epoch;
  if (S <= 0) {
    S = 1;
    while (S < W && bytes_for(S) > max_bytes) S *= 2;
  }
  if (S > W) S = W;
  // Relative times must fit int32 within an epoch.
  const int32_t n_epochs = static_cast<int32_t>((W + S - 1) / S);
  for (int32_t e = 0; e < n_epochs; ++e) {
    int32_t k0 = static_cast<int32_t>(e * S);
    int32_t k1 = static_cast<int32_t>(std::min<int64_t>(W, (e + 1) * S)) - 1;
    if (we[k1] - ws[k0] >= INT32_MAX)
      return set_err(err, REPLAY_E_UNSUPPORTED, "epoch spans more than 2^31 seconds");
  }
  if (bytes_for(S) > max_bytes)
    return set_err(err, REPLAY_E_NOMEM, "packed slices exceed the device budget");
  out.windows_per_epoch = static_cast<int32_t>(S);
  out.epoch_t0.resize(n_epochs);
  out.epoch_entry0.resize(n_epochs + 1);
  for (int32_t e = 0; e < n_epochs; ++e) out.epoch_t0[e] = ws[static_cast<size_t>(e * S)];
  for (int32_t k = 0; k < W; ++k) plan.windows[k].epoch = static_cast<int32_t>(k / S);
  // Counting sort by (epoch, row); entries of one row stay in rank order.
  std::vector<int64_t> cnt(static_cast<size_t>(n_epochs) * static_cast<size_t>(N) + 1, 0);
  for (int64_t k = 0; k < E; ++k) {
    const Span s = span[k];
    if (s.a > s.b) continue;
    for (int64_t e = s.a / S; e <= s.b / S; ++e) {
      cnt[static_cast<size_t>(e * N + cs[k].a)]++;
      cnt[static_cast<size_t>(e * N + cs[k].b)]++;
    }
  }
  int64_t total = 0;
  for (size_t g = 0; g + 1 < cnt.size(); ++g) {
    int64_t c = cnt[g];
    cnt[g] = total;
    total += c;
  }
  cnt.back() = total;
  out.entries.resize(static_cast<size_t>(total));
  out.rowptr.assign(static_cast<size_t>(n_epochs) * static_cast<size_t>(N + 1), 0);
  for (int32_t e = 0; e < n_epochs; ++e) {
    int64_t base = cnt[static_cast<size_t>(e) * N];
    out.epoch_entry0[e] = base;
    for (int64_t r = 0; r <= N; ++r) {
      int64_t g = cnt[static_cast<size_t>(e * N + r)];
      if (r == N) g = (e + 1 < n_epochs) ? cnt[static_cast<size_t>((e + 1) * N)] : total;
      int64_t rel = g - base;
      if (rel > INT32_MAX) return set_err(err, REPLAY_E_UNSUPPORTED, "epoch holds more than 2^31 entries");
      out.rowptr[static_cast<size_t>(e) * (N + 1) + r] = static_cast<int32_t>(rel);
    }
  }
  out.epoch_entry0[n_epochs] = total;
  // Keys g = e * N + row split into contiguous ranges of about equal entry
  // counts, one per worker.  Each worker walks every contact in rank order
  // and writes only its own keys' entries, so a row's entries land in rank
  // order exactly as in one sequential pass.
  const int64_t K = static_cast<int64_t>(n_epochs) * N;
  const int64_t parts = std::max<int64_t>(1, std::min<int64_t>(worker_count(), total / 65536));
  std::vector<int64_t> kcut(static_cast<size_t>(parts) + 1, K);
  kcut[0] = 0;
  for (int64_t t = 1; t < parts; ++t)
    kcut[static_cast<size_t>(t)] =
        std::upper_bound(cnt.begin(), cnt.begin() + K, total * t / parts) - cnt.begin() - 1;
  for (int64_t t = 1; t <= parts; ++t)
    kcut[static_cast<size_t>(t)] = std::max(kcut[static_cast<size_t>(t)], kcut[static_cast<size_t>(t) - 1]);
  std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
  par_chunks(parts, 1, [&](int, int64_t t_lo, int64_t t_hi) {
    for (int64_t t = t_lo; t < t_hi; ++t) {
      const int64_t g_lo = kcut[static_cast<size_t>(t)], g_hi = kcut[static_cast<size_t>(t) + 1];
      if (g_lo >= g_hi) continue;
      for (int64_t k = 0; k < E; ++k) {
        const Span s = span[k];
        if (s.a > s.b) continue;
        const Contact &c = cs[k];
        for (int64_t e = s.a / S; e <= s.b / S; ++e) {
          const int64_t ga = e * N + c.a, gb = e * N + c.b;
          const bool in_a = ga >= g_lo && ga < g_hi, in_b = gb >= g_lo && gb < g_hi;
          if (!in_a && !in_b) continue;
          const int64_t t0 = out.epoch_t0[e];
          const int32_t klast = static_cast<int32_t>(std::min<int64_t>(W, (e + 1) * S)) - 1;
          const int64_t t1 = we[klast];
          Entry en;
          en.ts = static_cast<int32_t>(std::max(c.start, t0) - t0);
          en.te = static_cast<int32_t>(std::min(c.end, t1) - t0);
          en.rank = static_cast<uint32_t>(k);
          if (in_a) {
            en.col = c.b;
            out.entries[static_cast<size_t>(fill[static_cast<size_t>(ga)]++)] = en;
          }
          if (in_b) {
            en.col = c.a;
            out.entries[static_cast<size_t>(fill[static_cast<size_t>(gb)]++)] = en;
          }
        }
      }
    }
  });
  // Columns ascending within a row, rank order kept among equal columns.
  // Integer sums are exact while a row's weights add up to less than 2^24:
  // the bound of short rows decides each window's mode, long rows carry
  // their own flag.
  // Rows sorted and bounded on the same key ranges; each worker keeps its
  // own maxima and long-row list, merged afterwards in key order.
  const int64_t D = we[0] - ws[0];
  struct RowStats {
    int64_t max_row = 0;
    std::vector<int64_t> row_bound;
    std::vector<int64_t> long_key;
    std::vector<uint8_t> long_int;
  };
  std::vector<RowStats> stats(static_cast<size_t>(parts));
  par_chunks(parts, 1, [&](int, int64_t t_lo, int64_t t_hi) {
    for (int64_t t = t_lo; t < t_hi; ++t) {
      RowStats &st = stats[static_cast<size_t>(t)];
      st.row_bound.assign(static_cast<size_t>(n_epochs), 0);
      for (int64_t g = kcut[static_cast<size_t>(t)]; g < kcut[static_cast<size_t>(t) + 1]; ++g) {
        const int64_t e = g / N, r = g - e * N;
        const int32_t *rp = &out.rowptr[static_cast<size_t>(e) * (N + 1)];
        Entry *base = out.entries.data() + out.epoch_entry0[e];
        Entry *b = base + rp[r], *x = base + rp[r + 1];
        const int64_t len = x - b;
        if (len > 1)
          std::stable_sort(b, x, [](const Entry &u, const Entry &v) { return u.col < v.col; });
        int64_t bound = 0;
        for (Entry *q = b; q < x; ++q) bound += std::min<int64_t>(q->te - q->ts, D);
        if (len > kLongRow) {
          st.long_key.push_back(g);
          st.long_int.push_back(bound < (int64_t(1) << 24) ? 1 : 0);
        } else {
          st.max_row = std::max(st.max_row, len);
          st.row_bound[static_cast<size_t>(e)] = std::max(st.row_bound[static_cast<size_t>(e)], bound);
        }
      }
    }
  });
  int64_t max_row = 0;
  std::vector<int64_t> row_bound(static_cast<size_t>(n_epochs), 0);
  out.long_off.assign(static_cast<size_t>(n_epochs) + 1, 0);
  for (const RowStats &st : stats) {
    max_row = std::max(max_row, st.max_row);
    for (int32_t e = 0; e < n_epochs; ++e)
      row_bound[static_cast<size_t>(e)] =
          std::max(row_bound[static_cast<size_t>(e)], st.row_bound[static_cast<size_t>(e)]);
    for (size_t q = 0; q < st.long_key.size(); ++q) {
      const int64_t e = st.long_key[q] / N;
      out.long_row.push_back(static_cast<int32_t>(st.long_key[q] - e * N));
      out.long_int.push_back(st.long_int[q]);
      out.long_off[static_cast<size_t>(e) + 1]++;
    }
  }
  for (int32_t e = 0; e < n_epochs; ++e)
    out.long_off[static_cast<size_t>(e) + 1] += out.long_off[static_cast<size_t>(e)];
  out.max_row_entries = max_row;
  for (auto &w : plan.windows)
    w.int_mode = (w.scale == 1.0f && row_bound[w.epoch] < (int64_t(1) << 24)) ? 1 : 0;
  return REPLAY_OK;
}

}  // namespace rp

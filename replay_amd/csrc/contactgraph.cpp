// contactgraph.cpp -- the ADDNODE / ADDEDGE command language of
// contactgraphgen/contactgraphgen.py, as a native library function.
//
// A command file declares nodes (ADDNODE id) and recurring cliques
// (ADDEDGE [ids] START DURATION DELTA [OCCURRENCES n | UNTIL t]); the
// generator writes the contact graph with every pair of each clique at each
// start, duplicates removed, in (start, u, v, end) order.  Messages for
// rejected lines are the reference's, on stderr, in file order, and the
// graph file is byte-identical to the reference's.
//
// The expansion is held as 24-byte records {start, u, v, end} with u and v
// the nodes' order of declaration, remapped to id order once all nodes are
// known and then sorted; times past the last second the reference can print
// (9999-12-31T23:59:59Z) are kept as one sentinel, which sorts where the
// reference's write stops with an exception.
//
// Lexical rules follow Python's str methods on the file's UTF-8 text:
// whitespace is str.isspace (29 code points), lines end at \n, \r\n or \r.
// Digits are ASCII only (str.isdigit also accepts other scripts' digits),
// and integers of more than 38 significant digits are rejected with a
// message of their own (the reference's integers have no bound).
#include <algorithm>
#include <cerrno>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <string_view>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "replay_hip.h"

namespace {

using u128 = unsigned __int128;
using sv = std::string_view;

constexpr int64_t kLastTime = 253402300799;  // 9999-12-31T23:59:59Z
constexpr int64_t kPastEnd = kLastTime + 1;
constexpr uint64_t kMaxRecords = uint64_t(1) << 30;

// Byte length of the str.isspace code point at p (0 when there is none).
size_t space_at(const char *p, const char *e) {
  const auto c = static_cast<unsigned char>(p[0]);
  if ((c >= 0x09 && c <= 0x0d) || (c >= 0x1c && c <= 0x20)) return 1;
  const auto b = [&](ptrdiff_t i) {
    return p + i < e ? static_cast<unsigned char>(p[i]) : 0u;
  };
  if (c == 0xc2 && (b(1) == 0x85 || b(1) == 0xa0)) return 2;
  if (c == 0xe1 && b(1) == 0x9a && b(2) == 0x80) return 3;
  if (c == 0xe2 && b(1) == 0x80 &&
      ((b(2) >= 0x80 && b(2) <= 0x8a) || b(2) == 0xa8 || b(2) == 0xa9 || b(2) == 0xaf))
    return 3;
  if (c == 0xe2 && b(1) == 0x81 && b(2) == 0x9f) return 3;
  if (c == 0xe3 && b(1) == 0x80 && b(2) == 0x80) return 3;
  return 0;
}

sv lstrip(sv s) {
  const char *p = s.data(), *e = p + s.size();
  while (p < e) {
    const size_t n = space_at(p, e);
    if (!n) break;
    p += n;
  }
  return sv(p, e - p);
}

sv strip(sv s) {
  s = lstrip(s);
  const char *b = s.data(), *e = b + s.size();
  while (e > b) {
    const char *q = e - 1;  // start of the last code point
    while (q > b && (static_cast<unsigned char>(*q) & 0xc0) == 0x80) --q;
    if (space_at(q, e) != static_cast<size_t>(e - q)) break;
    e = q;
  }
  return sv(b, e - b);
}

std::vector<sv> split_ws(sv s) {
  std::vector<sv> out;
  const char *p = s.data(), *e = p + s.size();
  while (p < e) {
    size_t n;
    while (p < e && (n = space_at(p, e)) != 0) p += n;
    if (p == e) break;
    const char *t = p;
    while (p < e && space_at(p, e) == 0) ++p;
    out.emplace_back(t, p - t);
  }
  return out;
}

bool digits(sv s) {
  if (s.empty()) return false;
  for (char c : s)
    if (c < '0' || c > '9') return false;
  return true;
}

// Value of an all-digit token; false when it has more than 38 significant
// digits.
bool to_u128(sv s, u128 &v) {
  size_t i = 0;
  while (i < s.size() && s[i] == '0') ++i;
  if (s.size() - i > 38) return false;
  v = 0;
  for (; i < s.size(); ++i) v = v * 10 + static_cast<unsigned>(s[i] - '0');
  return true;
}

std::string str(u128 v) {
  char b[48];
  int n = 0;
  do {
    b[n++] = static_cast<char>('0' + static_cast<int>(v % 10));
    v /= 10;
  } while (v);
  std::string s(n, '0');
  for (int i = 0; i < n; ++i) s[i] = b[n - 1 - i];
  return s;
}

std::string num(int64_t v) { return std::to_string(v); }

// Strict UTF-8, as Python's decoder reads it.
bool valid_utf8(const std::string &s) {
  const auto *p = reinterpret_cast<const unsigned char *>(s.data());
  const auto *e = p + s.size();
  while (p < e) {
    const unsigned c = *p;
    if (c < 0x80) {
      ++p;
      continue;
    }
    int n;
    unsigned lo = 0x80, hi = 0xbf;
    if (c >= 0xc2 && c <= 0xdf) {
      n = 1;
    } else if (c >= 0xe0 && c <= 0xef) {
      n = 2;
      if (c == 0xe0) lo = 0xa0;
      if (c == 0xed) hi = 0x9f;
    } else if (c >= 0xf0 && c <= 0xf4) {
      n = 3;
      if (c == 0xf0) lo = 0x90;
      if (c == 0xf4) hi = 0x8f;
    } else {
      return false;
    }
    if (e - p <= n) return false;
    if (p[1] < lo || p[1] > hi) return false;
    for (int k = 2; k <= n; ++k)
      if ((p[k] & 0xc0) != 0x80) return false;
    p += n + 1;
  }
  return true;
}

int64_t days_from_civil(int64_t y, unsigned m, unsigned d) {
  y -= m <= 2;
  const int64_t era = (y >= 0 ? y : y - 399) / 400;
  const unsigned yoe = static_cast<unsigned>(y - era * 400);
  const unsigned doy = (153 * (m + (m > 2 ? -3 : 9)) + 2) / 5 + d - 1;
  const unsigned doe = yoe * 365 + yoe / 4 - yoe / 100 + doy;
  return era * 146097 + static_cast<int64_t>(doe) - 719468;
}

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

// One field of datetime.strptime's "%Y-%m-%dT%H:%M:%SZ": 1 or 2 ASCII
// digits whose value lies in [lo2, hi2] when there are two and in
// [lo1, 9] when there is one (the patterns of Python's _strptime for %m %d
// %H %M %S).
bool field(sv s, unsigned lo1, unsigned lo2, unsigned hi2, unsigned &v) {
  if (s.empty() || s.size() > 2 || !digits(s)) return false;
  v = static_cast<unsigned>(s[0] - '0');
  if (s.size() == 1) return v >= lo1;
  v = v * 10 + static_cast<unsigned>(s[1] - '0');
  return v >= lo2 && v <= hi2;
}

// UTC seconds of a token datetime.strptime(tok, "%Y-%m-%dT%H:%M:%SZ")
// accepts (letters T and Z in either case; the date and a second below 60
// must then be valid); false where it raises.
bool parse_iso(sv s, int64_t &out) {
  // split at the separators -, -, T, :, :, Z in order
  const char seps[] = {'-', '-', 'T', ':', ':', 'Z'};
  sv part[6];
  size_t pos = 0;
  for (int k = 0; k < 6; ++k) {
    size_t q = pos;
    const char want = seps[k];
    while (q < s.size() && s[q] != want && !(want >= 'A' && s[q] == want + 32)) ++q;
    if (q == s.size()) return false;
    part[k] = s.substr(pos, q - pos);
    pos = q + 1;
  }
  if (pos != s.size()) return false;
  if (part[0].size() != 4 || !digits(part[0])) return false;
  const int64_t y = (part[0][0] - '0') * 1000 + (part[0][1] - '0') * 100 +
                    (part[0][2] - '0') * 10 + (part[0][3] - '0');
  unsigned mo, d, h, mi, se;
  if (!field(part[1], 1, 1, 12, mo) || !field(part[2], 1, 1, 31, d) ||
      !field(part[3], 0, 0, 23, h) || !field(part[4], 0, 0, 59, mi) ||
      !field(part[5], 0, 0, 61, se))
    return false;
  if (y == 0 || se > 59) return false;
  static const unsigned mdays[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
  const bool leap = (y % 4 == 0 && y % 100 != 0) || y % 400 == 0;
  if (d > mdays[mo - 1] + (mo == 2 && leap)) return false;
  out = days_from_civil(y, mo, d) * 86400 + h * 3600 + mi * 60 + se;
  return true;
}

// "YYYY-MM-DDTHH:MM:SSZ" of 0 <= t <= kLastTime, into 20 bytes.
void iso20(int64_t t, char *o) {
  int64_t y;
  unsigned m, d;
  civil_from_days(t / 86400, y, m, d);
  const unsigned s = static_cast<unsigned>(t % 86400);
  const auto two = [](char *p, unsigned v) {
    p[0] = static_cast<char>('0' + v / 10);
    p[1] = static_cast<char>('0' + v % 10);
  };
  const unsigned yy = static_cast<unsigned>(y);
  two(o, yy / 100);
  two(o + 2, yy % 100);
  o[4] = '-';
  two(o + 5, m);
  o[7] = '-';
  two(o + 8, d);
  o[10] = 'T';
  two(o + 11, s / 3600);
  o[13] = ':';
  two(o + 14, s / 60 % 60);
  o[16] = ':';
  two(o + 17, s % 60);
  o[19] = 'Z';
}

struct Record {
  int64_t start;
  uint32_t u, v;  // node order of declaration, then of id
  int64_t end;
};

struct U128Hash {
  size_t operator()(u128 v) const {
    const uint64_t lo = static_cast<uint64_t>(v), hi = static_cast<uint64_t>(v >> 64);
    return std::hash<uint64_t>()(lo ^ (hi * 0x9e3779b97f4a7c15ull));
  }
};

class Expander {
 public:
  explicit Expander(bool quiet) : quiet_(quiet) {}

  std::string log;  // every message, newline-terminated

  void say(const std::string &m) {
    log += m;
    log += '\n';
    if (!quiet_) {
      std::fputs(m.c_str(), stderr);
      std::fputc('\n', stderr);
    }
  }

  // false: the expansion is past kMaxRecords
  bool line(sv text, int64_t n);

  // Sorted, duplicate-free records in id order, and the ids as text.
  void finish(std::vector<std::string> &ids);

  std::vector<Record> rec;

 private:
  bool uint_tok(sv tok, const char *what, int64_t n, u128 &v);
  void addnode(sv text, int64_t n);
  bool addedge(sv text, int64_t n);

  bool quiet_;
  std::unordered_map<u128, uint32_t, U128Hash> index_;
  std::vector<u128> ids_;
};

bool Expander::uint_tok(sv tok, const char *what, int64_t n, u128 &v) {
  if (!digits(tok)) {
    say("ERROR: Invalid unsigned integer '" + std::string(tok) + "' for " + what + " on line " +
        num(n) + ". Skipping line " + num(n) + ".");
    return false;
  }
  if (!to_u128(tok, v)) {
    say("ERROR: Integer '" + std::string(tok) + "' for " + what + " on line " + num(n) +
        " has more than 38 digits. Skipping line " + num(n) + ".");
    return false;
  }
  return true;
}

void Expander::addnode(sv text, int64_t n) {
  const sv rest = strip(text.substr(7));
  const std::string L = num(n);
  if (rest.empty()) {
    say("ERROR: Missing ID for ADDNODE on line " + L + ". Skipping line " + L + ".");
    return;
  }
  const auto parts = split_ws(rest);
  if (parts.size() != 1) {
    say("ERROR: ADDNODE expects exactly one ID on line " + L + ". Skipping line " + L + ".");
    return;
  }
  u128 v;
  if (!uint_tok(parts[0], "node ID", n, v)) {
    say("Skipping ADDNODE on line " + L + ".");
    return;
  }
  if (v == 0) {
    say("ERROR: Node IDs must be positive integers. Found '0' on line " + L + ". Skipping line " +
        L + ".");
    return;
  }
  if (index_.emplace(v, static_cast<uint32_t>(ids_.size())).second) ids_.push_back(v);
}

bool Expander::addedge(sv text, int64_t n) {
  const std::string L = num(n);
  const std::string skip = "Skipping ADDEDGE on line " + L + ".";
  const sv rem = strip(text.substr(7));
  if (rem.empty() || rem[0] != '[') {
    say("ERROR: Expected ID list in brackets after ADDEDGE on line " + L + ". " + skip);
    return true;
  }
  size_t close = sv::npos;
  int64_t depth = 0;
  for (size_t i = 0; i < rem.size(); ++i) {
    if (rem[i] == '[') {
      ++depth;
    } else if (rem[i] == ']' && --depth == 0) {
      close = i;
      break;
    }
  }
  if (close == sv::npos) {
    say("ERROR: Unterminated ID list '[' on line " + L + ". Skipping line " + L + ".");
    return true;
  }
  std::vector<u128> list;
  const sv body = rem.substr(1, close - 1);
  if (!strip(body).empty()) {
    size_t p = 0;
    for (;;) {
      const size_t c = body.find(',', p);
      const sv item = strip(body.substr(p, c == sv::npos ? sv::npos : c - p));
      if (item.empty()) {
        say("ERROR: Empty entry in ID list on line " + L + ". Skipping line " + L + ".");
        return true;
      }
      if (!digits(item)) {
        say("ERROR: Non-integer ID '" + std::string(item) + "' on line " + L +
            ". Skipping line " + L + ".");
        return true;
      }
      u128 v;
      if (!to_u128(item, v)) {
        say("ERROR: Integer '" + std::string(item) + "' for node ID on line " + L +
            " has more than 38 digits. Skipping line " + L + ".");
        return true;
      }
      if (v == 0) {
        say("ERROR: Node IDs must be positive integers. Found '" + std::string(item) +
            "' on line " + L + ". Skipping line " + L + ".");
        return true;
      }
      list.push_back(v);
      if (c == sv::npos) break;
      p = c + 1;
    }
  }
  std::sort(list.begin(), list.end());
  list.erase(std::unique(list.begin(), list.end()), list.end());
  if (list.size() < 2) {
    say("WARNING: ADDEDGE on line " + L + " has fewer than two unique IDs; skipping.");
    return true;
  }
  const sv tail = strip(rem.substr(close + 1));
  if (tail.empty()) {
    say("ERROR: Missing timing parameters after ID list on line " + L + ". " + skip);
    return true;
  }
  const auto tok = split_ws(tail);
  if (tok.size() < 3) {
    say("ERROR: Expected <UNIX_START> <DURATION> <RECURRENCE_DELTA> on line " + L + ". " + skip);
    return true;
  }
  u128 start = 0, dur = 0, delta = 0;
  const bool ok0 = uint_tok(tok[0], "UNIX_START", n, start);
  const bool ok1 = uint_tok(tok[1], "DURATION", n, dur);
  const bool ok2 = uint_tok(tok[2], "RECURRENCE_DELTA", n, delta);
  if (!(ok0 && ok1 && ok2)) {
    say(skip);
    return true;
  }
  enum { kOnce, kCount, kUntil } horizon = kOnce;
  u128 count = 0;
  bool until_neg = false;
  u128 until = 0;
  if (tok.size() > 3) {
    const auto has = [&](sv w) { return std::find(tok.begin() + 3, tok.end(), w) != tok.end(); };
    if (has("OCCURRENCES") && has("UNTIL")) {
      say("ERROR: Both OCCURRENCES and UNTIL specified on line " + L + ". " + skip);
      return true;
    }
    if (tok[3] == "OCCURRENCES") {
      if (tok.size() != 5) {
        say("ERROR: 'OCCURRENCES' must be followed by a single unsigned integer on line " + L +
            ". " + skip);
        return true;
      }
      if (!uint_tok(tok[4], "OCCURRENCES", n, count)) {
        say(skip);
        return true;
      }
      horizon = kCount;
    } else if (tok[3] == "UNTIL") {
      if (tok.size() != 5) {
        say("ERROR: 'UNTIL' must be followed by exactly one time token on line " + L + ". " +
            skip);
        return true;
      }
      if (digits(tok[4])) {
        if (!to_u128(tok[4], until)) {
          say("ERROR: Integer '" + std::string(tok[4]) + "' for UNTIL on line " + L +
              " has more than 38 digits. Skipping line " + L + ".");
          say(skip);
          return true;
        }
      } else {
        int64_t t;
        if (!parse_iso(tok[4], t)) {
          say("ERROR: Invalid ISO8601Z time '" + std::string(tok[4]) + "' on line " + L +
              ". Expected 'YYYY-MM-DDTHH:MM:SSZ'.");
          say(skip);
          return true;
        }
        until_neg = t < 0;
        until = until_neg ? 0 : static_cast<u128>(t);
      }
      horizon = kUntil;
    } else {
      std::string joined;
      for (size_t k = 3; k < tok.size(); ++k) {
        if (k > 3) joined += ' ';
        joined += tok[k];
      }
      say("ERROR: Unexpected tokens after timing parameters on line " + L + ": '" + joined +
          "'. " + skip);
      return true;
    }
  }
  std::vector<uint32_t> node(list.size());
  bool missing = false;
  for (size_t k = 0; k < list.size(); ++k) {
    const auto it = index_.find(list[k]);
    if (it == index_.end()) {
      const std::string x = str(list[k]);
      say("ERROR: ID " + x + " is undefined. Add it with `ADDNODE " + x + "`.");
      missing = true;
    } else {
      node[k] = it->second;
    }
  }
  if (missing) {
    say(skip);
    return true;
  }
  // Starts in increasing order; after the first one past kLastTime every
  // later one is past it too and sorts after it, so expansion stops there.
  const uint64_t m = list.size(), pairs = m * (m - 1) / 2;
  const auto emit = [&](u128 t) {
    if (rec.size() + pairs > kMaxRecords) return false;
    const int64_t s = t > static_cast<u128>(kLastTime) ? kPastEnd : static_cast<int64_t>(t);
    const u128 te = t + dur;
    const int64_t e = te > static_cast<u128>(kLastTime) ? kPastEnd : static_cast<int64_t>(te);
    for (size_t a = 0; a + 1 < m; ++a)
      for (size_t b = a + 1; b < m; ++b) rec.push_back(Record{s, node[a], node[b], e});
    return true;
  };
  if (delta == 0 || horizon == kOnce) return emit(start);
  u128 t = start;
  if (horizon == kCount) {
    for (u128 k = 0; k < count; ++k, t += delta) {
      if (!emit(t)) return false;
      if (t > static_cast<u128>(kLastTime)) break;
    }
  } else if (!until_neg) {
    for (; t < until; t += delta) {
      if (!emit(t)) return false;
      if (t > static_cast<u128>(kLastTime)) break;
    }
  }
  return true;
}

bool Expander::line(sv raw, int64_t n) {
  const sv text = strip(raw);
  if (text.empty() || text[0] == '#') return true;
  if (text.substr(0, 7) == "ADDNODE") {
    addnode(text, n);
    return true;
  }
  if (text.substr(0, 7) == "ADDEDGE") return addedge(text, n);
  const std::string L = num(n);
  say("ERROR: Unknown command on line " + L + ": '" + std::string(text) + "'. Skipping line " + L +
      ".");
  return true;
}

void Expander::finish(std::vector<std::string> &ids) {
  std::vector<uint32_t> order(ids_.size());
  for (uint32_t k = 0; k < order.size(); ++k) order[k] = k;
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return ids_[a] < ids_[b]; });
  std::vector<uint32_t> rank(ids_.size());
  ids.resize(ids_.size());
  for (uint32_t r = 0; r < order.size(); ++r) {
    rank[order[r]] = r;
    ids[r] = str(ids_[order[r]]);
  }
  for (auto &x : rec) {
    x.u = rank[x.u];
    x.v = rank[x.v];
  }
  const auto key = [](const Record &x) {
    return std::make_tuple(x.start, x.u, x.v, x.end);
  };
  std::sort(rec.begin(), rec.end(), [&](const Record &a, const Record &b) { return key(a) < key(b); });
  rec.erase(std::unique(rec.begin(), rec.end(),
                        [&](const Record &a, const Record &b) { return key(a) == key(b); }),
            rec.end());
}

// Python's str of an OSError from open(): "[Errno N] text: 'path'".
std::string os_error(int err, const char *path) {
  return "[Errno " + std::to_string(err) + "] " + std::strerror(err) + ": '" + path + "'";
}

}  // namespace

extern "C" int32_t replay_contactgraph(const char *commands_path, const char *out_path,
                                       int64_t *n_contacts, int32_t quiet) {
  if (n_contacts) *n_contacts = 0;
  if (!commands_path || !out_path) return REPLAY_E_ARG;
  Expander x(quiet != 0);
  std::string data;
  {
    FILE *f = std::fopen(commands_path, "rb");
    if (!f) {
      const int err = errno;
      if (err == ENOENT)
        x.say(std::string("ERROR: Command file not found: ") + commands_path);
      else if (err == EACCES || err == EPERM)
        x.say(std::string("ERROR: Permission denied when reading: ") + commands_path);
      else
        x.say(std::string("ERROR: Cannot read command file: ") + os_error(err, commands_path));
      return REPLAY_E_IO;
    }
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) data.append(buf, got);
    const bool bad = std::ferror(f) != 0;
    const int err = errno;
    std::fclose(f);
    if (bad) {
      x.say(std::string("ERROR: Cannot read command file: ") + os_error(err, commands_path));
      return REPLAY_E_IO;
    }
  }
  if (!valid_utf8(data)) {
    x.say(std::string("ERROR: Command file is not valid UTF-8: ") + commands_path);
    return REPLAY_E_FORMAT;
  }
  // lines end at \n, \r\n or \r
  int64_t n = 0;
  size_t p = 0;
  while (p < data.size()) {
    size_t q = p;
    while (q < data.size() && data[q] != '\n' && data[q] != '\r') ++q;
    if (!x.line(sv(data.data() + p, q - p), ++n)) {
      x.say("ERROR: the command file expands to more than " + std::to_string(kMaxRecords) +
            " contact occurrences. No graph file written.");
      return REPLAY_E_NOMEM;
    }
    if (q < data.size() && data[q] == '\r' && q + 1 < data.size() && data[q + 1] == '\n') ++q;
    p = q + 1;
  }
  std::string().swap(data);
  std::vector<std::string> ids;
  x.finish(ids);

  FILE *f = std::fopen(out_path, "w");
  if (!f) {
    const int err = errno;
    if (err == EACCES || err == EPERM)
      x.say(std::string("ERROR: Permission denied when writing: ") + out_path);
    else
      x.say(std::string("ERROR: Failed to write output '") + out_path + "': " +
            os_error(err, out_path));
    return REPLAY_E_IO;
  }
  std::string out;
  out.reserve(1 << 22);
  bool ok = true;
  const auto flush = [&] {
    if (!out.empty() && std::fwrite(out.data(), 1, out.size(), f) != out.size()) ok = false;
    out.clear();
  };
  out += "NODE_LIST";
  for (const auto &s : ids) {
    out += ' ';
    out += s;
  }
  out += '\n';
  int32_t rc = REPLAY_OK;
  int64_t written = 0;
  int64_t last_start = -1, last_end = -1;
  char st[20], en[20];
  for (const Record &r : x.rec) {
    if (r.start == kPastEnd || r.end == kPastEnd) {
      x.say("ERROR: contact line " + std::to_string(written + 1) +
            " has a time past 9999-12-31T23:59:59Z; the graph file ends before it.");
      rc = REPLAY_E_OVERFLOW;
      break;
    }
    if (r.start != last_start) iso20(last_start = r.start, st);
    if (r.end != last_end) iso20(last_end = r.end, en);
    out += ids[r.u];
    out += ", ";
    out += ids[r.v];
    out += ", ";
    out.append(st, 20);
    out += ", ";
    out.append(en, 20);
    out += "; \n";
    ++written;
    if (out.size() >= (1u << 22)) flush();
  }
  flush();
  if (std::fclose(f) != 0) ok = false;
  if (n_contacts) *n_contacts = written;
  if (!ok) {
    x.say(std::string("ERROR: Failed to write output '") + out_path + "'");
    return REPLAY_E_IO;
  }
  return rc;
}

// replay_cgg -- command line of the command-file generator
// (contactgraphgen/contactgraphgen.py:385-418): `replay_cgg <commands> --out
// <graph>`, the same messages on stderr and the same graph file, written by
// replay_contactgraph.  Exit status 0 on success, 1 when the command file
// cannot be read, the graph cannot be written or a time passes year 9999,
// 2 for a usage error (argparse's status).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "replay_hip.h"

namespace {

const char kUsage[] = "usage: replay_cgg [-h] --out OUT commands\n";

int usage_error(const std::string &msg) {
  std::fputs(kUsage, stderr);
  std::fprintf(stderr, "replay_cgg: error: %s\n", msg.c_str());
  return 2;
}

// "--o", "--ou" and "--out" name the one long option, as argparse's
// unambiguous prefixes do.
bool is_out(const std::string &a) { return a.size() >= 3 && std::string("--out").rfind(a, 0) == 0; }

}  // namespace

int main(int argc, char **argv) {
  const char *out = nullptr;
  std::vector<const char *> pos;
  bool options = true;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (options && a == "--") {
      options = false;
    } else if (options && (a == "-h" || a == "--help")) {
      std::fputs(kUsage, stdout);
      std::fputs("\nGenerate a contact-graph file from a command file.\n\n"
                 "positional arguments:\n  commands    Path to the command file (text).\n\n"
                 "options:\n  -h, --help  show this help message and exit\n"
                 "  --out OUT   Path to write the contact graph (text).\n",
                 stdout);
      return 0;
    } else if (options && a.rfind("--", 0) == 0 && a.size() > 2) {
      const size_t eq = a.find('=');
      const std::string name = a.substr(0, eq);
      if (!is_out(name)) return usage_error("unrecognized arguments: " + a);
      if (eq != std::string::npos) {
        out = argv[i] + eq + 1;
      } else if (i + 1 < argc && argv[i + 1][0] != '-') {
        out = argv[++i];
      } else {
        return usage_error("argument --out: expected one argument");
      }
    } else if (options && a.size() > 1 && a[0] == '-') {
      return usage_error("unrecognized arguments: " + a);
    } else {
      pos.push_back(argv[i]);
    }
  }
  if (pos.size() > 1) {
    std::string extra;
    for (size_t k = 1; k < pos.size(); ++k) extra += (k > 1 ? " " : "") + std::string(pos[k]);
    return usage_error("unrecognized arguments: " + extra);
  }
  if (pos.empty() || !out) {
    std::string missing = pos.empty() ? "commands" : "";
    if (!out) missing += std::string(missing.empty() ? "" : ", ") + "--out";
    return usage_error("the following arguments are required: " + missing);
  }
  int64_t n = 0;
  return replay_contactgraph(pos[0], out, &n, 0) == REPLAY_OK ? 0 : 1;
}

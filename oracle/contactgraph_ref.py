"""CPU oracle for the command-file generator -- TEST INFRASTRUCTURE ONLY.

A plain-Python restatement of contactgraphgen/contactgraphgen.py, the
ADDNODE / ADDEDGE command language that writes a contact graph.  Only the
tests import it, as the checker of the native generator
(`replay_contactgraph` / `replay_cgg`); nothing in the product calls it.

expand(data) -> Expansion(messages, graph, contacts, failed)

  messages  the lines the reference prints on stderr, in order
  graph     the text of the graph file (None when the reference writes none)
  contacts  number of contact lines written
  failed    True when the reference ends with a non-zero status

Rules restated (reference line numbers):
  - lines: universal newlines, each stripped; blank and '#' lines are
    skipped (py:327-332); commands are prefixes ADDNODE / ADDEDGE,
    anything else is an unknown command (py:334-357);
  - ADDNODE takes exactly one unsigned integer > 0 (py:335-351);
  - ADDEDGE [ids] START DURATION DELTA [OCCURRENCES N | UNTIL T]: the
    bracket list is matched by depth (py:117-128), split on commas, each
    entry stripped and a positive integer (py:131-164); fewer than two
    distinct ids is a warning (py:209-213); the three timing integers are
    all parsed before one "Skipping" line (py:215-231); the horizon clause
    (py:233-275); ids not yet added by ADDNODE reject the line
    (py:277-284); starts are START + k DELTA for k < N, or while < T, one
    start when DELTA is 0 or without a clause (py:286-306); every pair
    u < v of the sorted distinct ids gets [t, t + DURATION] (py:308-313);
  - occurrences form a set, written in (start, u, v, end) order with the
    node ids in ascending order (py:359-381); each line ends "; " before
    the newline (py:380);
  - a time past 9999-12-31T23:59:59Z ends the write with an exception: the
    file keeps the lines before it and the status is non-zero (py:65-68
    and 365-381 under main's handlers py:395-418).

Python's own str.strip / str.split / str.isdigit / datetime.strptime do
the lexical work here, as in the reference.
"""
from __future__ import annotations

import datetime as _dt
import io
from collections import namedtuple

Expansion = namedtuple("Expansion", "messages graph contacts failed")

_ISO = "%Y-%m-%dT%H:%M:%SZ"


def _iso(ts: int) -> str:
    return _dt.datetime.fromtimestamp(ts, tz=_dt.timezone.utc).strftime(_ISO)


class _Line:
    """Message sink for one command line."""

    def __init__(self, n: int, out: list):
        self.n, self.out = n, out

    def say(self, msg: str) -> None:
        self.out.append(msg)

    def uint(self, tok: str, what: str):
        if not tok.isdigit():
            self.say(f"ERROR: Invalid unsigned integer '{tok}' for {what} on line {self.n}. "
                     f"Skipping line {self.n}.")
            return None
        try:
            return int(tok, 10)
        except ValueError:
            self.say(f"ERROR: Could not parse integer '{tok}' for {what} on line {self.n}. "
                     f"Skipping line {self.n}.")
            return None

    def until(self, tok: str):
        if tok.isdigit():
            return int(tok, 10)
        try:
            d = _dt.datetime.strptime(tok, _ISO).replace(tzinfo=_dt.timezone.utc)
            return int(d.timestamp())
        except ValueError:
            self.say(f"ERROR: Invalid ISO8601Z time '{tok}' on line {self.n}. "
                     f"Expected 'YYYY-MM-DDTHH:MM:SSZ'.")
            return None


def _ids(ln: _Line, rem: str):
    """(ids, rest) of a bracket list at rem[0] == '[', or None."""
    depth, close = 0, -1
    for i, c in enumerate(rem):
        depth += (c == "[") - (c == "]")
        if c == "]" and depth == 0:
            close = i
            break
    if close < 0:
        ln.say(f"ERROR: Unterminated ID list '[' on line {ln.n}. Skipping line {ln.n}.")
        return None
    body = rem[1:close]
    out = []
    if body.strip():
        for item in (x.strip() for x in body.split(",")):
            if not item:
                ln.say(f"ERROR: Empty entry in ID list on line {ln.n}. Skipping line {ln.n}.")
                return None
            if not item.isdigit():
                ln.say(f"ERROR: Non-integer ID '{item}' on line {ln.n}. Skipping line {ln.n}.")
                return None
            if int(item) <= 0:
                ln.say(f"ERROR: Node IDs must be positive integers. Found '{item}' on line "
                       f"{ln.n}. Skipping line {ln.n}.")
                return None
            out.append(int(item))
    return out, rem[close + 1:]


def _addedge(ln: _Line, text: str, nodes: set, occ: set) -> None:
    n = ln.n
    skip = f"Skipping ADDEDGE on line {n}."
    rem = text[len("ADDEDGE"):].strip()
    if not rem.startswith("["):
        ln.say(f"ERROR: Expected ID list in brackets after ADDEDGE on line {n}. {skip}")
        return
    got = _ids(ln, rem)
    if got is None:
        return
    ids = sorted(set(got[0]))
    if len(ids) < 2:
        ln.say(f"WARNING: ADDEDGE on line {n} has fewer than two unique IDs; skipping.")
        return
    tail = got[1].strip()
    if not tail:
        ln.say(f"ERROR: Missing timing parameters after ID list on line {n}. {skip}")
        return
    tok = tail.split()
    if len(tok) < 3:
        ln.say(f"ERROR: Expected <UNIX_START> <DURATION> <RECURRENCE_DELTA> on line {n}. {skip}")
        return
    timing = [ln.uint(tok[0], "UNIX_START"), ln.uint(tok[1], "DURATION"),
              ln.uint(tok[2], "RECURRENCE_DELTA")]
    if None in timing:
        ln.say(skip)
        return
    start, dur, delta = timing
    count = until = None
    extra = tok[3:]
    if extra:
        if "OCCURRENCES" in extra and "UNTIL" in extra:
            ln.say(f"ERROR: Both OCCURRENCES and UNTIL specified on line {n}. {skip}")
            return
        if extra[0] == "OCCURRENCES":
            if len(extra) != 2:
                ln.say(f"ERROR: 'OCCURRENCES' must be followed by a single unsigned integer on "
                       f"line {n}. {skip}")
                return
            count = ln.uint(extra[1], "OCCURRENCES")
            if count is None:
                ln.say(skip)
                return
        elif extra[0] == "UNTIL":
            if len(extra) != 2:
                ln.say(f"ERROR: 'UNTIL' must be followed by exactly one time token on line {n}. "
                       f"{skip}")
                return
            until = ln.until(extra[1])
            if until is None:
                ln.say(skip)
                return
        else:
            ln.say(f"ERROR: Unexpected tokens after timing parameters on line {n}: "
                   f"'{' '.join(extra)}'. {skip}")
            return
    missing = [x for x in ids if x not in nodes]
    if missing:
        for x in missing:
            ln.say(f"ERROR: ID {x} is undefined. Add it with `ADDNODE {x}`.")
        ln.say(skip)
        return
    if delta == 0 or (count is None and until is None):
        starts = [start]
    elif count is not None:
        starts = [start + k * delta for k in range(count)]
    else:
        starts = list(range(start, until, delta)) if until > start else []
    for t in starts:
        for a in range(len(ids)):
            for b in range(a + 1, len(ids)):
                occ.add((t, ids[a], ids[b], t + dur))


def expand(data: bytes) -> Expansion:
    """The reference's stderr lines, graph text and status for a command
    file's bytes."""
    msgs: list = []
    nodes: set = set()
    occ: set = set()
    try:
        lines = list(io.TextIOWrapper(io.BytesIO(data), encoding="utf-8"))
    except UnicodeDecodeError:
        return Expansion(None, None, 0, True)
    for n, raw in enumerate(lines, start=1):
        text = raw.strip()
        if not text or text.startswith("#"):
            continue
        ln = _Line(n, msgs)
        if text.startswith("ADDNODE"):
            rest = text[len("ADDNODE"):].strip()
            if not rest:
                ln.say(f"ERROR: Missing ID for ADDNODE on line {n}. Skipping line {n}.")
                continue
            parts = rest.split()
            if len(parts) != 1:
                ln.say(f"ERROR: ADDNODE expects exactly one ID on line {n}. Skipping line {n}.")
                continue
            v = ln.uint(parts[0], "node ID")
            if v is None:
                ln.say(f"Skipping ADDNODE on line {n}.")
                continue
            if v <= 0:
                ln.say(f"ERROR: Node IDs must be positive integers. Found '{v}' on line {n}. "
                       f"Skipping line {n}.")
                continue
            nodes.add(v)
        elif text.startswith("ADDEDGE"):
            _addedge(ln, text, nodes, occ)
        else:
            ln.say(f"ERROR: Unknown command on line {n}: '{text}'. Skipping line {n}.")
    head = "NODE_LIST" + "".join(f" {v}" for v in sorted(nodes)) + "\n"
    body = [head]
    written = 0
    for t0, u, v, t1 in sorted(occ):
        try:
            body.append(f"{u}, {v}, {_iso(t0)}, {_iso(t1)}; \n")
        except (ValueError, OverflowError, OSError):
            return Expansion(msgs, "".join(body), written, True)
        written += 1
    return Expansion(msgs, "".join(body), written, False)

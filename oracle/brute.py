"""ORACLE -- test infrastructure only.

Brute-force evaluator over the raw triple list, for tiny graphs. It never
touches a gstore image: neighbour lists come straight from the triples, using
the reference's storage rules (core/gstore.hpp:546-560 encoding comment,
665-870 insert/index rules):
  (v, p, OUT)  -> objects of (v, p, *), ascending
  (v, p, IN)   -> subjects of (*, p, v), ascending; rdf:type triples have no
                  IN side (core/gstore.hpp:661-664, 689-692)
  (v, 1, OUT)  -> the types of v
  (0, p, IN)   -> subjects of predicate p; (0, p, OUT) -> objects of p
  (0, t, IN)   -> members of type t
Step semantics follow core/engine.hpp:96-255 and the dispatch of
core/engine.hpp:500-570 with request state from core/query.hpp:120-165.
"""
from collections import defaultdict

import numpy as np

VID_BASE = 1 << 17


def _is_tpid(x):
    return 1 < x < VID_BASE


class BruteGraph:
    def __init__(self, s, p, o):
        s = [int(x) for x in s]
        p = [int(x) for x in p]
        o = [int(x) for x in o]
        out, inn = defaultdict(set), defaultdict(set)
        subj, obj, members = defaultdict(set), defaultdict(set), defaultdict(set)
        for a, b, c in zip(s, p, o):
            if b == 1:  # rdf:type: (s,1,OUT) -> types; type index members
                out[(a, 1)].add(c)
                members[c].add(a)
                continue
            out[(a, b)].add(c)
            inn[(c, b)].add(a)
            subj[b].add(a)
            obj[b].add(c)
        self.out = {k: sorted(v) for k, v in out.items()}
        self.inn = {k: sorted(v) for k, v in inn.items()}
        self.subj = {k: sorted(v) for k, v in subj.items()}
        self.obj = {k: sorted(v) for k, v in obj.items()}
        self.members = {k: sorted(v) for k, v in members.items()}

    def edges(self, vid, pid, d):
        if vid < 0:
            return []
        if vid == 0:
            if d == 0:
                if pid in self.subj:
                    return self.subj[pid]
                return self.members.get(pid, [])
            return self.obj.get(pid, [])
        if d == 1:
            return self.out.get((vid, pid), [])
        if pid == 1:
            return []
        return self.inn.get((vid, pid), [])

    def run(self, cmd):
        """Evaluates cmd_chains: returns (rows as a 2-D int32 array, expanded)."""
        cmd = [int(x) for x in cmd]
        table, col_num, expanded = [], 0, 0

        def vtype(v):
            if v >= 0:
                return "const"
            return "unknown" if -v > col_num else "known"

        for step in range(len(cmd) // 4):
            start, pid, d, end = cmd[4 * step:4 * step + 4]
            if step == 0 and _is_tpid(start):
                assert pid in (0, 1) and end == -1
                lst = self.edges(0, start, d)
                table = [[x] for x in lst]
                expanded += len(lst)
                col_num = 1
                continue
            ts, te = vtype(start), vtype(end)
            new = []
            if ts == "const" and te == "unknown":
                assert col_num == 0 and end == -1
                lst = self.edges(start, pid, d)
                new = [[x] for x in lst]
                expanded += len(lst)
                col_num = 1
            elif ts == "known" and te == "unknown":
                assert col_num == -end - 1
                for row in table:
                    lst = self.edges(row[-start - 1], pid, d)
                    expanded += len(lst)
                    for x in lst:
                        new.append(row + [x])
                col_num += 1
            elif ts == "known" and te in ("known", "const"):
                for row in table:
                    lst = self.edges(row[-start - 1], pid, d)
                    target = row[-end - 1] if te == "known" else end
                    if target in lst:
                        new.append(row)
                        expanded += lst.index(target) + 1
                    else:
                        expanded += len(lst)
            else:
                raise ValueError("unsupported pattern")
            table = new
        arr = np.array(table, dtype=np.int32).reshape(len(table), col_num)
        return arr, expanded

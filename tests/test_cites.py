"""Every reference citation in the sources resolves.

A citation is `path:line` or `path:first-last` with path under src/, app/ or
tools/ of the reference tree.  Each must name an existing file and a line
range inside it.  Where a C++ name (`Class::member` or `function(`) stands
just before the citation and occurs in the cited file, it must occur within
the cited lines (two lines of slack either side), so that a citation cannot
drift onto unrelated code.  Needs the reference tree; skipped without it.
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"

CITE = re.compile(r"((?:src|app|tools)/[\w./-]+?\.(?:h|cu|cc|cpp|sh|py)):(\d+)(?:-(\d+))?")
NAME = re.compile(r"([A-Za-z_]\w*)::([A-Za-z_]\w*)|\b([A-Za-z_]\w{3,})\(")
SCANNED = ("include", "lighttraffic_amd", "oracle", "tests", "bench.py", "__graft_entry__.py")

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")), reason="reference tree absent")


def _sources():
    for top in SCANNED:
        path = os.path.join(ROOT, top)
        if os.path.isfile(path):
            yield path
            continue
        for dirpath, _, files in os.walk(path):
            if "__pycache__" in dirpath or os.sep + "build" in dirpath:
                continue
            for f in files:
                if f.endswith((".h", ".hip", ".cpp", ".c", ".py")):
                    yield os.path.join(dirpath, f)


def _citations():
    for path in _sources():
        text = open(path, errors="replace").read()
        for m in CITE.finditer(text):
            yield path, text, m


def test_citations_resolve():
    lines_of = {}
    bad, seen = [], 0
    for path, text, m in _citations():
        seen += 1
        rel, first = m.group(1), int(m.group(2))
        last = int(m.group(3)) if m.group(3) else first
        target = os.path.join(REF, rel)
        if not os.path.isfile(target):
            bad.append(f"{os.path.relpath(path, ROOT)}: {m.group(0)} (no such file)")
            continue
        if target not in lines_of:
            lines_of[target] = open(target, errors="replace").read().split("\n")
        lines = lines_of[target]
        if first < 1 or last < first or last > len(lines):
            bad.append(f"{os.path.relpath(path, ROOT)}: {m.group(0)} (file has {len(lines)} lines)")
            continue
        before = text[max(0, m.start() - 50):m.start()]
        names = [a[1] or a[2] for a in NAME.findall(before)]
        names = [n for n in names if n and any(n in line for line in lines)]
        if names and names[-1] not in "\n".join(lines[max(0, first - 3):last + 2]):
            bad.append(f"{os.path.relpath(path, ROOT)}: {names[-1]} is not at {m.group(0)}")
    assert seen > 50, "citation pattern found almost nothing"
    assert not bad, "\n".join(bad)

"""Every reference citation `file.h:a-b` in the sources stays inside the file.

The headers, kernels, host code, oracle and tests cite the reference's
source files by line range.  This test resolves each cited file name in the
reference tree (a path suffix when the citation gives one) and checks that
every cited line number is at most that file's line count.  It needs the
reference checkout and is skipped where it is absent.
"""
import collections
import os
import re

import pytest

REF = "/root/reference"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCAN_DIRS = ["include", "cfdv_amd", "oracle", "tests"]
SCAN_TOP = ["bench.py", "__graft_entry__.py", "DESIGN.md", "INTEGRATION.md", "README.md"]
EXT = (".h", ".hip", ".cpp", ".hpp", ".c", ".py", ".md")
CITE = re.compile(r"([A-Za-z_][A-Za-z0-9_/.]*\.(?:h|cpp|hpp))\s*:\s*(\d+)(?:\s*[-–]\s*(\d+))?"
                  r"((?:\s*,\s*\d+(?:\s*[-–]\s*\d+)?)*)")


def _reference_lengths():
    out = collections.defaultdict(list)
    for dp, _, files in os.walk(REF):
        if "/.git" in dp:
            continue
        for f in files:
            if f.endswith((".h", ".cpp", ".hpp")):
                p = os.path.join(dp, f)
                with open(p, "rb") as fh:
                    out[f].append((os.path.relpath(p, REF), fh.read().count(b"\n") + 1))
    return out


def _sources():
    for d in SCAN_DIRS:
        for dp, dirs, files in os.walk(os.path.join(ROOT, d)):
            dirs[:] = [x for x in dirs if x not in ("__pycache__", "_build", "golden")]
            for f in files:
                if f.endswith(EXT):
                    yield os.path.join(dp, f)
    for f in SCAN_TOP:
        p = os.path.join(ROOT, f)
        if os.path.exists(p):
            yield p


def _out_of_range():
    lengths = _reference_lengths()
    bad, n = [], 0
    for p in _sources():
        with open(p, errors="replace") as fh:
            for ln, text in enumerate(fh, 1):
                for m in CITE.finditer(text):
                    name = os.path.basename(m.group(1))
                    if name not in lengths:
                        continue
                    nums = [int(m.group(2))] + ([int(m.group(3))] if m.group(3) else [])
                    nums += [int(x) for x in re.findall(r"\d+", m.group(4) or "")]
                    cands = lengths[name]
                    if "/" in m.group(1):
                        cands = [c for c in cands if c[0].endswith(m.group(1))] or cands
                    n += 1
                    if max(nums) > max(c[1] for c in cands):
                        bad.append(f"{os.path.relpath(p, ROOT)}:{ln}: {m.group(0)} "
                                   f"(file has {max(c[1] for c in cands)} lines)")
    return bad, n


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")), reason="reference checkout absent")
def test_citations_inside_reference_files():
    bad, n = _out_of_range()
    assert n > 100, f"only {n} citations found: the pattern no longer matches the sources"
    assert not bad, "citations past the end of the cited file:\n" + "\n".join(bad)

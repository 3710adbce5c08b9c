"""Every reference citation `path/name.ext:N[-M][, K[-L]]` in the committed
sources and documents names a file of the reference tree and lines that exist
in it (a citation past the end of its file cites nothing).  Runs where the
reference tree is mounted; skipped elsewhere."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
SCANNED_DIRS = ["include", "oracle", "cuot_amd", "tests"]
SCANNED_FILES = ["DESIGN.md", "INTEGRATION.md", "README.md", "bench.py", "__graft_entry__.py"]
EXTS = ("cu", "h", "hpp", "cpp", "c", "sh", "cuh", "dockerfile", "txt")
CITE = re.compile(r"([A-Za-z0-9_./-]+\.(?:" + "|".join(EXTS) + r"))\s*:\s*(\d+)(?:\s*-\s*(\d+))?"
                  r"((?:\s*,\s*\d+(?:\s*-\s*\d+)?)*)")


def _reference_files():
    files = {}
    for dp, dn, fn in os.walk(REF):
        dn[:] = [d for d in dn if d != ".git"]
        for f in fn:
            files.setdefault(f, []).append(os.path.join(dp, f))
    return files


def _scanned():
    out = []
    for sub in SCANNED_DIRS:
        for dp, dn, fn in os.walk(os.path.join(ROOT, sub)):
            dn[:] = [d for d in dn if d not in ("_build", "_ref", "__pycache__", "lib")]
            out += [os.path.join(dp, f) for f in fn if f.endswith((".h", ".hip", ".c", ".cpp", ".py", ".md"))]
    out += [os.path.join(ROOT, f) for f in SCANNED_FILES if os.path.exists(os.path.join(ROOT, f))]
    return out


def _line_count(path, cache={}):
    if path not in cache:
        with open(path, errors="replace") as f:
            cache[path] = sum(1 for _ in f)
    return cache[path]


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not mounted")
def test_reference_citations_are_in_range():
    files = _reference_files()
    own = {os.path.basename(p) for p in _scanned()}
    bad, count = [], 0
    for path in _scanned():
        if os.path.abspath(path) == os.path.abspath(__file__):
            continue
        for ln, line in enumerate(open(path, errors="replace"), 1):
            for m in CITE.finditer(line):
                name = m.group(1).lstrip("./")
                base = os.path.basename(name)
                cands = [p for p in files.get(base, []) if p.endswith("/" + name)]
                if not cands:
                    if base in own:  # a citation of this repository's own files
                        continue
                    bad.append(f"{os.path.relpath(path, ROOT)}:{ln}: {m.group(0)} names no reference file")
                    continue
                ranges = [(int(m.group(2)), int(m.group(3) or m.group(2)))]
                for part in re.findall(r"\d+(?:\s*-\s*\d+)?", m.group(4) or ""):
                    xs = re.split(r"\s*-\s*", part)
                    ranges.append((int(xs[0]), int(xs[-1])))
                count += 1
                if not any(all(1 <= lo <= hi <= _line_count(p) for lo, hi in ranges) for p in cands):
                    bad.append(f"{os.path.relpath(path, ROOT)}:{ln}: {m.group(0)} is past the end of "
                               f"{[os.path.relpath(p, REF) for p in cands]} ({[_line_count(p) for p in cands]} lines)")
    assert count > 100
    assert not bad, "\n".join(bad)

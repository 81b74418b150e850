"""Every reference citation `path:N` or `path:N-M` in the repository's sources
and documents names lines that exist in the cited reference file.

Runs where the reference tree is present (the build container); elsewhere it
is skipped.  Citations are relative to the reference's root, e.g.
core/weno.f90:106-153.
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("NYL_REFERENCE_TREE", "/root/reference")

CITE = re.compile(r"((?:core|experiments|tools|doc|docs)/[A-Za-z0-9_./-]+\.(?:py|f90|json|txt|md|png|cfg))"
                  r":(\d+)(?:-(\d+))?")
SCAN_EXT = (".py", ".hip", ".h", ".c", ".cpp", ".md", ".json")
SCAN_DIRS = ("nyles_amd", "include", "oracle", "tests", "profiles")
SKIP_DIRS = {"__pycache__", "_build", ".pytest_cache"}
DOCS = ("DESIGN.md", "INTEGRATION.md")   # the repository's own documents


def _files():
    for d in SCAN_DIRS:
        for dp, dns, fns in os.walk(os.path.join(ROOT, d)):
            dns[:] = [x for x in dns if x not in SKIP_DIRS]
            for fn in fns:
                if fn.endswith(SCAN_EXT):
                    yield os.path.join(dp, fn)
    for fn in sorted(os.listdir(ROOT)):
        if fn.endswith(".py") or fn in DOCS:
            yield os.path.join(ROOT, fn)


def _nlines(path, cache={}):
    if path not in cache:
        with open(path, "rb") as f:
            cache[path] = f.read().count(b"\n") + 1
    return cache[path]


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "core")), reason="reference tree absent")
def test_citations_within_cited_files():
    bad = []
    n = 0
    for fp in _files():
        with open(fp, encoding="utf-8", errors="replace") as f:
            text = f.read()
        for m in CITE.finditer(text):
            rel, a, b = m.group(1).rstrip("."), int(m.group(2)), m.group(3)
            b = int(b) if b else a
            n += 1
            target = os.path.join(REF, rel)
            where = "%s: %s" % (os.path.relpath(fp, ROOT), m.group(0))
            if not os.path.isfile(target):
                bad.append(where + " (no such file)")
            elif not (1 <= a <= b <= _nlines(target)):
                bad.append(where + " (file has %d lines)" % _nlines(target))
    assert n > 100, "citation scan found too few citations (%d)" % n
    assert not bad, "citations past the end of the cited file:\n" + "\n".join(bad)

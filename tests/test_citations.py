"""Reference citations in the product and oracle sources point at real lines.

Every `File.ext:a-b` citation of a reference source file in include/, the
package, oracle/ and bench.py must lie inside that file, and where a
function or method name stands right before the citation
(`Class::method (File.cu:a-b)`, `some_kernel (File.cu:a-b)`), that name
must occur in the cited lines (or the three lines before them, where a
signature starts).  Runs only where the reference checkout is present.
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/GPU_MPM"
SOURCES = ["include", "explicit-gpumpm-ductile-fracture_amd", "oracle", "bench.py", "__graft_entry__.py"]

pytestmark = pytest.mark.skipif(not os.path.isdir(REF), reason="reference checkout not present")

CITE = re.compile(r"\b([A-Za-z_]+\.(?:cu|cpp|h|cuh|hpp)):(\d+(?:-\d+)?(?:,\s*\d+(?:-\d+)?)*)")
NAMED = re.compile(r"([A-Za-z_][A-Za-z_0-9]*(?:::[A-Za-z_][A-Za-z_0-9]*)?)\s*\(([A-Za-z_]+\.(?:cu|cpp|h|cuh|hpp)):"
                   r"(\d+)(?:-(\d+))?")


def _files():
    for s in SOURCES:
        p = os.path.join(ROOT, s)
        if os.path.isfile(p):
            yield p
            continue
        for dp, dn, fn in os.walk(p):
            dn[:] = [d for d in dn if d not in ("_ref", "lib", "__pycache__")]
            for f in fn:
                if f.endswith((".h", ".hip", ".cpp", ".c", ".py")):
                    yield os.path.join(dp, f)


def _ref_lines():
    return {f: open(os.path.join(REF, f), errors="replace").read().split("\n") for f in os.listdir(REF)}


def _is_code_name(name):
    # a method (A::b), a snake_case or camelCase identifier; plain words are prose
    return "::" in name or "_" in name or re.search(r"[a-z][A-Z]", name) is not None


def test_citations_lie_inside_the_cited_files():
    ref = _ref_lines()
    bad, n = [], 0
    for p in _files():
        for ln, line in enumerate(open(p, errors="replace"), 1):
            for m in CITE.finditer(line):
                if m.group(1) not in ref:
                    continue
                length = len(ref[m.group(1)])
                for part in re.split(r",\s*", m.group(2)):
                    a, _, b = part.partition("-")
                    a, b = int(a), int(b) if b else int(a)
                    n += 1
                    if not (1 <= a <= b <= length):
                        bad.append("%s:%d %s:%s (%d lines)" % (os.path.relpath(p, ROOT), ln, m.group(1), part, length))
    assert n > 100
    assert not bad, "\n".join(bad)


def test_named_citations_point_at_the_name():
    ref = _ref_lines()
    bad, n = [], 0
    for p in _files():
        for ln, line in enumerate(open(p, errors="replace"), 1):
            for m in NAMED.finditer(line):
                name, f, a, b = m.group(1), m.group(2), int(m.group(3)), m.group(4)
                if f not in ref or not _is_code_name(name):
                    continue
                b = int(b) if b else a
                n += 1
                seg = "\n".join(ref[f][max(0, a - 4):b])
                if name.split("::")[-1] not in seg:
                    bad.append("%s:%d %s not in %s:%d-%d" % (os.path.relpath(p, ROOT), ln, name, f, a, b))
    assert n > 20
    assert not bad, "\n".join(bad)

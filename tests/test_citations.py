"""The reference citations of the C ABI and the C++ concept layer point at
real lines.

Every `path:line` or `path:first-last` cited in include/*.h and
include/yannq_amd/*.hpp must lie inside the cited file of the yannq source
tree, and the citations of the ABI's entry points must cover the member they
name (a word of its signature inside the cited range).  Skipped when the
reference tree is not mounted (it never is on the GPU machines).
"""
from __future__ import annotations

import glob
import os
import re

import pytest

REF = "/root/reference"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")) +
                 glob.glob(os.path.join(ROOT, "include", "yannq_amd", "*.hpp")))
CITE = re.compile(r"([A-Za-z0-9_/]+\.(?:hpp|cpp|h))((?::\d+(?:-\d+)?)(?:,\s*:?\d+(?:-\d+)?)*)")

pytestmark = pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not mounted")


def _ref_files():
    out = {}
    for p in glob.glob(os.path.join(REF, "**", "*"), recursive=True):
        if os.path.isfile(p) and p.endswith((".hpp", ".cpp", ".h")):
            out.setdefault(os.path.basename(p), []).append(p)
    return out


def _resolve(files, name):
    return [p for p in files.get(os.path.basename(name), []) if p.endswith("/" + name)
            or os.path.basename(p) == name]


def _citations(path):
    for ln, line in enumerate(open(path).read().splitlines(), 1):
        for m in CITE.finditer(line):
            for a, b in re.findall(r"(\d+)(?:-(\d+))?", m.group(2)):
                yield ln, m.group(1), int(a), int(b) if b else int(a)


def test_every_cited_range_is_inside_its_file():
    files = _ref_files()
    bad, seen = [], 0
    for h in HEADERS:
        for ln, name, a, b in _citations(h):
            cands = _resolve(files, name)
            if not cands:
                bad.append(f"{os.path.relpath(h, ROOT)}:{ln}: {name} not in the reference")
                continue
            n = sum(1 for _ in open(cands[0], errors="replace"))
            seen += 1
            if not 1 <= a <= b <= n:
                bad.append(f"{os.path.relpath(h, ROOT)}:{ln}: {name}:{a}-{b} outside 1..{n}")
    assert seen > 50
    assert not bad, "\n".join(bad)


# (cited file and range as the ABI header writes it, a word the range holds)
ANCHORS = [
    ("SamplerMT.hpp:44-60", "SamplerMT(const Machine& qs"),
    ("SwapSweeper.hpp:15", "SwapSweeper(uint32_t n"),
    ("LocalSweeper.hpp:19", "LocalSweeper(uint32_t n"),
    ("SamplerMT.hpp:81-98", "randomizeSigma(int nup)"),
    ("SamplerMT.hpp:103-110", "void randomize("),
    ("SamplerMT.hpp:180-204", "auto sample(uint32_t nSweeps"),
    ("SamplerMT.hpp:120-153", "void mixChains()"),
    ("Energy.hpp:53-58", "void finIter()"),
    ("RBM.hpp:50", "RBM(uint32_t n, uint32_t m"),
    ("RBM.hpp:271-291", "getParams()"),
    ("RBM.hpp:261-269", "updateParams("),
    ("RBM.hpp:301-322", "initializeRandom("),
    ("RBM.hpp:293-296", "hasNaN()"),
    ("SRMat.hpp:89-103", "constructFromSamples"),
    ("SRMat.hpp:170-189", "apply("),
    ("SRMat.hpp:197-223", "solveCG("),
    ("SRMat.hpp:231-237", "solveExact("),
    ("NGDExact.hpp:177-189", "corrMatAmp()"),
    ("FeedForward.hpp:76-82", "addLayer"),
    ("FeedForward.hpp:152-168", "forward("),
    ("FeedForward.hpp:186-205", "backward("),
    ("TFIsing.hpp:28-39", "operator()"),
    ("XXZ.hpp:37-49", "operator()"),
    ("J1J2.hpp:35-55", "operator()"),
    ("ExactSampler.hpp:49-77", "sample("),
    ("SRMatExact.hpp:97-149", "constructExact"),
]


@pytest.mark.parametrize("cite,word", ANCHORS)
def test_abi_citation_covers_its_member(cite, word):
    abi = open(os.path.join(ROOT, "include", "yannq_amd.h")).read()
    assert cite in abi, f"{cite} is no longer cited by the ABI header"
    name, rng = cite.split(":")
    a, _, b = rng.partition("-")
    a, b = int(a), int(b or a)
    cands = _resolve(_ref_files(), name)
    assert cands, name
    lines = open(cands[0], errors="replace").read().splitlines()
    assert word in "\n".join(lines[a - 1:b]), f"{cite} does not hold {word!r}"

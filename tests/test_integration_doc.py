"""The C++ snippets the integration guide shows a reference maintainer are
the code of tests/cpp/test_integration_snippets.cpp (compiled by the library
build and run against the fused frames under -m gpu): every line of a block
marked "compiled and run" must be a line of that program."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROGRAM = os.path.join(ROOT, "tests", "cpp", "test_integration_snippets.cpp")
GUIDE = os.path.join(ROOT, "INTEGRATION.md")
MARK = "<!-- compiled and run: tests/cpp/test_integration_snippets.cpp -->"


def _blocks():
    text = open(GUIDE).read()
    return re.findall(re.escape(MARK) + r"\n```cpp\n(.*?)```", text, flags=re.S)


def test_marked_snippets_are_program_code():
    program = {ln.strip() for ln in open(PROGRAM).read().splitlines()}
    blocks = _blocks()
    assert len(blocks) == 3
    for b in blocks:
        missing = [ln for ln in b.splitlines() if ln.strip() and ln.strip() not in program]
        assert not missing, missing


def test_snippets_use_existing_names():
    """The names the unmarked blocks call exist in the shipped headers."""
    text = open(GUIDE).read()
    headers = ""
    for h in ("include/libatssim_hip.h", "libatssim_amd/csrc/host/launcher.h", "libatssim_amd/csrc/host/vbd_slot.h"):
        headers += open(os.path.join(ROOT, h)).read()
    for name in set(re.findall(r"\blats_[a-z0-9_]+(?=\()", text)):
        assert re.search(r"\b" + name + r"\b", headers), name
    for name in set(re.findall(r"\b(?:s|scheduler)\.([a-z_]+)\(", text)):
        assert re.search(r"\b" + name + r"\(", headers), name

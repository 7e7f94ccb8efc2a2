"""The C-ABI library loads and exports every symbol include/jaqsi_amd.h declares."""

import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_match_header():
    from jaqsi_amd import _native as nat

    header = open(os.path.join(ROOT, "include", "jaqsi_amd.h")).read()
    declared = set(re.findall(r"^(?:int|const char\*)\s+(jq_\w+)\s*\(", header, re.M))
    assert declared, "no declarations found"
    lib = nat.lib()
    for name in declared:
        assert hasattr(lib, name), name
    assert set(nat.EXPORTS) >= declared


def test_version_and_error_text():
    from jaqsi_amd import _native as nat

    assert b"gfx950" in nat.lib().jq_version()
    assert isinstance(nat.lib().jq_last_error(), bytes)


def test_header_constants_match_bindings():
    """Every #define JQ_* integer constant of the header is mirrored in _native."""
    from jaqsi_amd import _native as nat

    header = open(os.path.join(ROOT, "include", "jaqsi_amd.h")).read()
    for name, val in re.findall(r"^#define (JQ_(?:OBS|ENV)_\w+)\s+(\d+)", header, re.M):
        assert getattr(nat, name) == int(val), name

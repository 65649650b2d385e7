"""Array identifiers and shape classes of the C-ABI (include/efdc_hip.h).

The enum values are read from the header itself so that the Python host
side, the library and the tests cannot drift apart.
"""
from __future__ import annotations

import os
import re
from typing import Dict

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
HEADER = os.path.join(ROOT, "include", "efdc_hip.h")


def _parse_enum(name: str) -> Dict[str, int]:
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    m = re.search(r"typedef enum \{([^{}]*)\}\s*" + name + r"\s*;", text)
    if not m:
        raise RuntimeError(f"enum {name} not found in {HEADER}")
    body = m.group(1)
    out: Dict[str, int] = {}
    val = -1
    for tok in body.split(","):
        tok = tok.strip()
        if not tok:
            continue
        if "=" in tok:
            k, v = (s.strip() for s in tok.split("="))
            val = int(v, 0)
        else:
            k = tok
            val += 1
        out[k] = val
    return out


ARRAY_IDS: Dict[str, int] = {
    k[len("EFDC_"):]: v for k, v in _parse_enum("efdc_array_id").items()
    if not k.startswith("EFDC_NUM")}
STATUS: Dict[str, int] = _parse_enum("efdc_status")
STATUS_NAMES = {v: k for k, v in STATUS.items()}


def shape_class(name: str) -> str:
    """'h' for (LCM), 'k' for (LCM,KCM), 'w' for (LCM,0:KCM)."""
    v = ARRAY_IDS[name]
    if v < 64 or 128 <= v < 160:
        return "h"
    if v < 96 or 160 <= v < 192:
        return "k"
    return "w"

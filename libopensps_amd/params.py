"""Per-model parameter slots of the batch, read from the C header that defines them."""
from __future__ import annotations

import os
import re

_HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "opensps_params.h")


def _parse():
    txt = open(_HDR).read()
    slots = {}
    for name, val in re.findall(r"#define\s+OPS_P_(\w+)\s+(\d+)", txt):
        slots[name.lower()] = int(val)
    npar = int(re.search(r"#define\s+OPS_NPAR\s+(\d+)", txt).group(1))
    body = re.search(r"OPS_PARAM_DEFAULTS\s*\{(.*?)\}", txt, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S).replace("\\", " ")
    defaults = [float(v) for v in body.replace("\n", " ").split(",") if v.strip()]
    nbody = re.search(r"OPS_PARAM_NAMES\s*\{(.*?)\}", txt, re.S).group(1)
    names = re.findall(r'"(\w+)"', nbody)
    return slots, npar, defaults, names


_SLOTS, NPAR, PARAM_DEFAULTS, PARAM_NAMES = _parse()
# name -> slot, with the public names (imf_upper_limit, imf_lower_limit) as in PARAM_NAMES
SLOT = {n: i for i, n in enumerate(PARAM_NAMES)}
assert len(PARAM_NAMES) == NPAR == len(PARAM_DEFAULTS)

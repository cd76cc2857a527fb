"""The per-size kernel configuration table (dnsbox_amd/csrc/tuned_cfg.hpp): parse and emit.

One entry per (N, precision):
  zw, zp   tile width (columns) and radices of the z passes (zp symmetric)
  yw, yp   the same for the y passes
  rb, ntx  x-lines per block and threads of the x pass
  rp       radices of the length-N/2 row transforms of the x pass
  xpd, xpad  x-pass LDS row layout: element j at j + j // xpd (xpd = 0: at j), xpad extra
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, replace

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "csrc")
TABLE_PATH = os.path.join(CSRC, "tuned_cfg.hpp")
SIZES = (16, 32, 48, 64, 128, 256, 512)
GROUP_SIZES = (16, 32, 48, 64, 128)   # boxes of a group share 256-thread blocks in every pass


@dataclass(frozen=True)
class Cfg:
    zw: int
    yw: int
    rb: int
    ntx: int
    zp: tuple
    yp: tuple
    rp: tuple
    xpd: int = 0
    xpad: int = 0

    def cxx(self) -> str:
        plan = lambda p: "Plan<" + ", ".join(map(str, p)) + ">"
        return (f"CfgBase<{self.zw}, {self.yw}, {self.rb}, {self.ntx}, {plan(self.zp)}, {plan(self.yp)}, "
                f"{plan(self.rp)}, {self.xpd}, {self.xpad}>")

    def with_(self, **kw) -> "Cfg":
        return replace(self, **kw)


_LINE = re.compile(r"template <> struct CfgDefault<(\d+), (float|double)>\s*:\s*CfgBase<(.*)>\s*\{\};")


def _split_args(s):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    out.append(cur.strip())
    return out


def _plan(s):
    m = re.fullmatch(r"Plan<([\d,\s]+)>", s)
    return tuple(int(x) for x in m.group(1).split(","))


def parse(path: str = TABLE_PATH) -> dict:
    table = {}
    with open(path) as f:
        for line in f:
            m = _LINE.search(line)
            if not m:
                continue
            a = _split_args(m.group(3))
            xpd = int(a[7]) if len(a) > 7 else 0
            xpad = int(a[8]) if len(a) > 8 else 0
            table[(int(m.group(1)), m.group(2))] = Cfg(int(a[0]), int(a[1]), int(a[2]), int(a[3]), _plan(a[4]),
                                                       _plan(a[5]), _plan(a[6]), xpd, xpad)
    return table


def emit(table: dict, notes: list, path: str = TABLE_PATH) -> None:
    lines = [
        "// Kernel configuration per box size and precision (CfgDefault<N, T>, see CfgBase in",
        "// box.hpp).  Written by `python -m dnsbox_amd.tune.autotune`; the notes record the",
        "// measurements behind each entry.",
        "#pragma once",
        "",
        "namespace dnsbox {",
        "",
    ]
    for note in notes:
        lines.append("// " + note)
    if notes:
        lines.append("")
    for prec in ("float", "double"):
        for n in SIZES:
            c = table[(n, prec)]
            lines.append(f"template <> struct CfgDefault<{n}, {prec}> : {c.cxx()} {{}};")
    lines += ["", "}  // namespace dnsbox", ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))

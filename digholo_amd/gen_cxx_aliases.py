"""Generates the C++-linkage entry points of libdigholo.so from include/digHolo.h.

The reference header declares its functions with an empty EXT_C on Linux
(digHolo.h:148-156), so a C++ caller compiled against it references mangled
names (_Z13digHoloCreatev, ...).  For every EXT_C declaration in
include/digHolo.h this writes a C++ function of the same signature that calls
the C entry point (declared under another identifier with an asm label naming
the C symbol):

    extern "C" int digHoloCreate__c(void) __asm__("digHoloCreate");
    int digHoloCreate(void) { return digHoloCreate__c(); }

The declarations keep the reference's parameter types, so the mangled names are
the ones a caller of the reference header uses.

    python -m digholo_amd.gen_cxx_aliases OUT.cpp
"""
from __future__ import annotations

import os
import re
import sys

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "digHolo.h")


def declarations(text: str):
    """(return type, name, parameter list) of every EXT_C function declaration."""
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    out = []
    for m in re.finditer(r"EXT_C\s+([^;{}()]*?)\b(digHolo\w+)\s*\(([^;{}]*?)\)\s*;", text, flags=re.S):
        ret = " ".join(m.group(1).split())
        params = " ".join(m.group(3).split())
        out.append((ret, m.group(2), params))
    return out


def argument_names(params: str):
    if params in ("", "void"):
        return []
    names = []
    for p in params.split(","):
        m = re.search(r"(\w+)\s*(\[[^\]]*\])*\s*$", p.strip())
        if not m or m.group(1) in ("int", "float", "double", "char", "void", "short", "long", "unsigned"):
            raise ValueError(f"unnamed parameter in '{params}'")
        names.append(m.group(1))
    return names


def generate(header: str = HEADER) -> str:
    with open(header, encoding="utf-8") as f:
        decls = declarations(f.read())
    lines = [
        "// Generated by digholo_amd/gen_cxx_aliases.py from include/digHolo.h; do not edit.",
        "// C++-linkage entry points (the names a caller of the reference header links",
        "// against), each calling the C entry point of the same name.",
        "#define DIGHOLO_CXX_LINKAGE",
        '#include "digHolo.h"',
        "",
    ]
    for ret, name, params in decls:
        args = ", ".join(argument_names(params))
        plist = params if params else "void"
        lines.append(f'extern "C" {ret} {name}__c({plist}) __asm__("{name}");')
        call = f"{name}__c({args})"
        body = f"{call};" if ret == "void" else f"return {call};"
        lines.append(f"{ret} {name}({plist}) {{ {body} }}")
    return "\n".join(lines) + "\n"


if __name__ == "__main__":
    out = sys.argv[1] if len(sys.argv) > 1 else "dh_cxx_aliases.cpp"
    with open(out, "w") as f:
        f.write(generate())

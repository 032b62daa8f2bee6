#!/usr/bin/env python3
"""Writes host_calls.json: the functions and globals that the reference's
host C files use and that the reference defines in the files this library
replaces (the CUDA sources and src/mpi_comm.c).  A program built from the
reference's host objects links against libbluebottle.so only if the library
exports every one of them.

Host files (kept, compiled as the reference ships them): bluebottle.c,
domain.c, particle.c, recorder.c, scalar.c, rng.c, vtk.c.  cgns.c is left
out (it needs libcgns).  Replaced files: every *.cu and mpi_comm.c.

Run in a tree that holds the reference sources (reads them as text only):
    python3 tests/golden/extract_host_calls.py /path/to/reference/src
"""
import json
import os
import re
import sys

HOST = ["bluebottle.c", "domain.c", "particle.c", "recorder.c", "scalar.c",
        "rng.c", "vtk.c"]


def strip_comments(text):
    """Comments and string literals blanked, line structure kept."""
    def blank(m):
        return re.sub(r"[^\n]", " ", m.group(0))
    return re.sub(r'/\*.*?\*/|//[^\n]*|"(?:\\.|[^"\\\n])*"', blank, text,
                  flags=re.S)


def defined_functions(text):
    """Host-callable function definitions at file scope (no __global__ or
    __device__ kernels)."""
    out = set()
    pat = re.compile(r'^(?:extern\s+"C"\s+)?((?:[A-Za-z_]\w*[\s\*]+)+?)'
                     r'([A-Za-z_]\w*)\s*\(([^;{}]*)\)\s*\{', re.M)
    for m in pat.finditer(text):
        quals = m.group(1)
        if "__global__" in quals or "__device__" in quals:
            continue
        if m.group(2) in ("if", "for", "while", "switch", "return"):
            continue
        out.add(m.group(2))
    return out


def defined_globals(text):
    """File-scope variables (not __constant__ or __device__)."""
    out = set()
    for m in re.finditer(r'^([A-Za-z_][\w ]*?[\s\*]+)([A-Za-z_]\w*)\s*'
                         r'(?:\[[^\]]*\])?\s*(?:=[^;]*)?;', text, re.M):
        quals = m.group(1)
        if any(q in quals for q in ("__constant__", "__device__", "return",
                                    "typedef", "extern", "static")):
            continue
        out.add(m.group(2))
    return out


def site_key(site):
    f, ln = site.rsplit(":", 1)
    return (f, int(ln))


def main(src):
    provider_files = sorted(f for f in os.listdir(src)
                            if f.endswith(".cu")) + ["mpi_comm.c"]
    funcs, globs = {}, {}
    for f in provider_files:
        text = strip_comments(open(os.path.join(src, f)).read())
        for name in defined_functions(text):
            funcs.setdefault(name, f)
        for name in defined_globals(text):
            globs.setdefault(name, f)
    host_defs = set()
    host_text = {}
    for f in HOST:
        t = strip_comments(open(os.path.join(src, f)).read())
        host_text[f] = t
        host_defs |= defined_functions(t)
    calls, uses, test_only = {}, {}, {}
    for f, t in host_text.items():
        lines = t.split("\n")
        stack = []          # one entry per open #if: inside a TEST branch?
        for ln, line in enumerate(lines, 1):
            d = re.match(r"\s*#\s*(ifdef|ifndef|if|else|elif|endif)\b\s*(\w*)",
                         line)
            if d:
                kw, arg = d.group(1), d.group(2)
                if kw in ("ifdef", "if"):
                    stack.append(arg == "TEST")
                elif kw == "ifndef":
                    stack.append(False)
                elif kw in ("else", "elif") and stack:
                    stack[-1] = False
                elif kw == "endif" and stack:
                    stack.pop()
                continue
            in_test = any(stack)
            for m in re.finditer(r"\b([A-Za-z_]\w*)\s*\(", line):
                name = m.group(1)
                if name in funcs and name not in host_defs:
                    calls.setdefault(name, []).append(f"src/{f}:{ln}")
                    test_only[name] = test_only.get(name, True) and in_test
            for m in re.finditer(r"\b([A-Za-z_]\w*)\b", line):
                name = m.group(1)
                if name in globs and name not in funcs:
                    uses.setdefault(name, []).append(f"src/{f}:{ln}")
    out = {
        "host_files": [f"src/{f}" for f in HOST],
        "replaced_files": [f"src/{f}" for f in provider_files],
        "functions": {k: {"defined_in": f"src/{funcs[k]}",
                          "called_at": sorted(set(v)),
                          "test_build_only": test_only[k]}
                      for k, v in sorted(calls.items())},
        # a global is listed with its number of uses and the first three
        "globals": {k: {"defined_in": f"src/{globs[k]}",
                        "uses": len(set(v)),
                        "used_at": sorted(set(v), key=site_key)[:3]}
                    for k, v in sorted(uses.items())},
    }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                       "host_calls.json")
    with open(dst, "w") as fh:
        json.dump(out, fh, indent=1)
        fh.write("\n")
    print(f"{len(out['functions'])} functions, {len(out['globals'])} "
          f"globals -> {dst}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src")

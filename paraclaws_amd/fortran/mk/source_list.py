#!/usr/bin/env python3
"""Ordered Fortran source list with USER_DIR overrides.

Rule (the reference's USER_DIR mechanism, mk/generate_source_list.py:53-134):
a user .f90 that declares a module or submodule also declared by a base
source replaces that base source at the base source's position; any other
user .f90 is appended after the base sources (sorted by path).  It is an
error for one user file to replace two base files, or for two different
user files to replace the same base file.  Files under USER_DIR/build are
ignored, and so are base files when USER_DIR contains the base tree.

Prints one path per line (or writes a make fragment with --make VAR).
"""
from __future__ import annotations

import argparse
import re
import sys
from pathlib import Path

UNIT_RES = (
    re.compile(r"^\s*module\s+([a-z]\w*)\s*$", re.IGNORECASE),
    re.compile(r"^\s*submodule\s*\([^)]*\)\s*([a-z]\w*)", re.IGNORECASE),
)


def units(path: Path) -> set[str]:
    found = set()
    for line in path.read_text(errors="replace").splitlines():
        code = line.split("!", 1)[0]
        for rx in UNIT_RES:
            m = rx.match(code)
            if m:
                found.add(m.group(1).lower())
    return found


def resolve(base: list[Path], user_dir: Path | None, base_root: Path) -> list[Path]:
    base = [p.resolve() for p in base]
    owner: dict[str, Path] = {}
    for p in base:
        if not p.is_file():
            raise ValueError(f"base source missing: {p}")
        for u in units(p):
            if u in owner and owner[u] != p:
                raise ValueError(f"unit {u!r} declared by {owner[u]} and {p}")
            owner[u] = p
    if user_dir is None:
        return base
    user_dir = user_dir.resolve()
    base_root = base_root.resolve()
    build = user_dir / "build"
    replace: dict[Path, Path] = {}
    extra: list[Path] = []
    for p in sorted(user_dir.rglob("*")):
        if not p.is_file() or p.suffix.lower() != ".f90":
            continue
        r = p.resolve()
        if build == r or build in r.parents:
            continue
        if base_root in r.parents and (user_dir == base_root or user_dir in base_root.parents):
            continue
        targets = {owner[u] for u in units(r) if u in owner and owner[u] != r}
        if len(targets) > 1:
            raise ValueError(f"{p} declares units of several base sources: {sorted(map(str, targets))}")
        if targets:
            t = targets.pop()
            prev = replace.get(t)
            if prev is not None and prev != r:
                if prev.read_bytes() == r.read_bytes():
                    continue
                raise ValueError(f"both {prev} and {p} replace {t}")
            replace[t] = r
        else:
            extra.append(r)
    return [replace.get(p, p) for p in base] + extra


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--base-root", type=Path, required=True)
    ap.add_argument("--user-dir", type=Path)
    ap.add_argument("--make", metavar="VAR")
    ap.add_argument("base", nargs="+", type=Path)
    a = ap.parse_args()
    try:
        srcs = resolve(a.base, a.user_dir, a.base_root)
    except ValueError as e:
        print(f"source_list: {e}", file=sys.stderr)
        return 1
    if a.make:
        print(f"{a.make} := " + " ".join(str(s) for s in srcs))
    else:
        print("\n".join(str(s) for s in srcs))
    return 0


if __name__ == "__main__":
    sys.exit(main())

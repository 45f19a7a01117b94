"""Applies the edits of INTEGRATION.md (its ```cpp edit <file> <mode> blocks) to a
copy of the reference in a temporary directory.

TEST INFRASTRUCTURE ONLY (like the rest of oracle/): it lets the glue a
maintainer adds to the reference be compiled and run exactly as documented.
The copy and the patched files stay outside the repository; only the binary
built from them goes to oracle/_ref/.

    python3 apply_edits.py <reference root> <INTEGRATION.md> <destination>

copies <reference root>/include and src/gravity_backends.cpp to <destination>
and applies every block.  Modes: replace (the anchor occurs once and is
replaced), after / before (the new text goes after / before the anchor, which
occurs once), after-all (after every occurrence, at least one).  Anchors are
whole lines.
"""
import os
import re
import shutil
import sys

BLOCK = re.compile(r"^```cpp edit (\S+) (replace|after|before|after-all)\n(.*?)^```\n",
                   re.M | re.S)


def parse(text):
    edits = []
    for m in BLOCK.finditer(text):
        path, mode, body = m.group(1), m.group(2), m.group(3)
        if "\n---\n" not in body:
            raise ValueError(f"edit of {path}: no '---' line between anchor and new text")
        anchor, new = body.split("\n---\n", 1)
        edits.append((path, mode, anchor, new.rstrip("\n")))
    return edits


def occurrences(text, anchor):
    """Start offsets of `anchor` as whole lines of `text`."""
    out, start = [], 0
    while True:
        k = text.find(anchor, start)
        if k < 0:
            return out
        end = k + len(anchor)
        if (k == 0 or text[k - 1] == "\n") and (end == len(text) or text[end] == "\n"):
            out.append(k)
        start = k + 1


def apply(text, mode, anchor, new, where):
    found = occurrences(text, anchor)
    if mode == "after-all":
        if not found:
            raise ValueError(f"{where}: anchor not found: {anchor!r}")
    elif len(found) != 1:
        raise ValueError(f"{where}: anchor found {len(found)} times (want 1): {anchor!r}")
    for k in reversed(found):
        end = k + len(anchor)
        if mode == "replace":
            text = text[:k] + new + text[end:]
        elif mode == "before":
            text = text[:k] + new + "\n" + text[k:]
        else:
            text = text[:end] + "\n" + new + text[end:]
    return text


def main():
    ref, doc, dest = sys.argv[1:4]
    edits = parse(open(doc).read())
    if not edits:
        raise SystemExit("no edit blocks found")
    shutil.copytree(os.path.join(ref, "include"), os.path.join(dest, "include"))
    os.makedirs(os.path.join(dest, "src"), exist_ok=True)
    shutil.copy(os.path.join(ref, "src", "gravity_backends.cpp"),
                os.path.join(dest, "src", "gravity_backends.cpp"))
    for path, mode, anchor, new in edits:
        full = os.path.join(dest, path)
        with open(full) as f:
            text = f.read()
        text = apply(text, mode, anchor, new, path)
        with open(full, "w") as f:
            f.write(text)
    print(f"applied {len(edits)} edits", file=sys.stderr)


if __name__ == "__main__":
    main()

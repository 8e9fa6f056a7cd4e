"""Hash of the sources libsmilei_amd.so is built from.

The Makefile embeds it in the library (smi_build_hash) and the loader
compares it with the sources beside it, so a library built from other
sources is refused instead of being run.  Run as a script it prints the
hash.
"""
import glob
import hashlib
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def source_files():
    csrc = os.path.join(HERE, "csrc")
    files = []
    for pat in ("*.hip", "*.cpp", "*.h", "Makefile"):
        files += glob.glob(os.path.join(csrc, pat))
    files.append(os.path.join(ROOT, "include", "smilei_amd.h"))
    # names relative to the repository root, so the hash does not depend on
    # where the tree sits
    return sorted(os.path.relpath(f, ROOT) for f in files)


def source_hash():
    h = hashlib.sha256()
    for rel in source_files():
        h.update(rel.encode())
        h.update(b"\0")
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
        h.update(b"\0")
    return h.hexdigest()[:16]


if __name__ == "__main__":
    print(source_hash())

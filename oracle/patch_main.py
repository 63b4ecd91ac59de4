"""Writes the reference main program with the GPU drop-in bound in: a copy
of src/main.F (read from the reference tree, written only into the build
directory given) with four lines added and no call site changed:

  * ``USE FWGPU_DROPIN`` after the program's ``USE GLOBAL`` (src/main.F:136):
    the per-step calls of the time loop (src/main.F:284-535) resolve to the
    module's GPU routines of the same names and argument lists;
  * ``CALL FWGPU_BEGIN`` after ``CALL INITIALIZATION`` (:214);
  * ``CALL FWGPU_SYNC_OUTPUT`` before the loop's ``CALL OUTPUT`` (:288);
  * ``CALL FWGPU_END`` after the loop's ``END DO`` (:535).

Build recipe step (oracle/Makefile target ``dropin``), not product code.
Usage: python patch_main.py <reference main.F> <output .F>
"""
import re
import sys


def patch(src: str) -> str:
    lines = src.splitlines(keepends=True)
    out = []
    state = "head"      # head -> init -> loop -> done
    used = False
    for ln in lines:
        s = ln.strip().upper()
        if state == "head" and not used and re.match(r"USE\s+GLOBAL\s*$", s):
            out.append(ln)
            out.append("     USE FWGPU_DROPIN\n")
            used = True
            continue
        if state == "head" and re.match(r"CALL\s+INITIALIZATION\s*$", s):
            out.append(ln)
            out.append("     CALL FWGPU_BEGIN\n")
            state = "init"
            continue
        if state == "init" and re.match(r"DO\s+WHILE\s*\(\s*TIME\s*<\s*TOTAL_TIME\s*\)", s):
            state = "loop"
            out.append(ln)
            continue
        if state == "loop" and re.match(r"CALL\s+OUTPUT\s*$", s):
            out.append("      CALL FWGPU_SYNC_OUTPUT\n")
            out.append(ln)
            state = "loop_body"
            continue
        if state == "loop_body" and re.match(r"END\s*DO\s*$", s) and ln.startswith("   END"):
            out.append(ln)
            out.append("   CALL FWGPU_END\n")
            state = "done"
            continue
        out.append(ln)
    if not used or state != "done":
        raise SystemExit(f"patch_main.py: anchors not found (USE GLOBAL={used}, state={state})")
    return "".join(out)


if __name__ == "__main__":
    with open(sys.argv[1]) as f:
        text = patch(f.read())
    with open(sys.argv[2], "w") as f:
        f.write(text)

"""Build-time guard over the generated gfx950 assembly: fails if any kernel
contains a scalar-cache store, scalar atomic or scalar-cache write-back.
These are never wanted here (all global writes go through vector memory
instructions).  Runs in the build container only."""
from __future__ import annotations

import re
import subprocess
import sys
from pathlib import Path

FORBIDDEN = re.compile(r"^\s*(s_store_|s_buffer_store|s_scratch_store|s_atomic_|s_buffer_atomic|"
                       r"s_dcache_wb|s_dcache_discard)", re.M)


def check(hipcc: str, flags: list[str], csrc: Path, sources: list[str]) -> None:
    bad = []
    for src in sources:
        r = subprocess.run([hipcc, *flags, "--cuda-device-only", "-S", "-o", "-", str(csrc / src)],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"ISA dump failed for {src}:\n{r.stderr}")
        for m in FORBIDDEN.finditer(r.stdout):
            bad.append(f"{src}: {m.group(0).strip()}")
    if bad:
        raise RuntimeError("scalar-cache stores/atomics in device code:\n" + "\n".join(bad))


if __name__ == "__main__":
    sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
    from cookiedough_amd import build as b
    check(b.hipcc(), b.HIPCC_FLAGS, b.CSRC, b.SOURCES)
    print("isa ok")

"""The persistent SDF tracer's ray set-up replaces two divisions by
cheaper exact forms (cookiedough_amd/csrc/k_sdf.hip, div_rn and the refill's
chunk row):

- u = (2x - W) / H and v = (H - 2y) / H as Markstein's correction of a * rb
  with rb = 1 / H rounded to nearest, fma(fma(-b, a * rb, a), rb, a * rb):
  bit-equal to the IEEE single-precision quotient for every integer a with
  |a| <= 4096 and every 1 <= b <= 2160 (the batch's limits on W and H);
- k / cw as (k * ceil(2^20 / cw)) >> 20 for k < 4096 and cw <= 64, with the
  product below 2^32 (a 24-bit multiply).

Both are checked exhaustively by a small C program (fmaf and the quotient
in IEEE single precision, contraction off)."""
import shutil
import subprocess

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static uint32_t bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
int main(void) {
    long bad = 0, n = 0;
    for (int b = 1; b <= 2160; ++b) {
        const float fb = (float)b, rb = 1.0f / fb;
        for (int a = -4096; a <= 4096; ++a) {
            const float fa = (float)a, q = fa * rb;
            const float got = fmaf(fmaf(-fb, q, fa), rb, q);
            ++n;
            if (bits(got) != bits(fa / fb)) ++bad;
        }
    }
    long bad2 = 0, n2 = 0;
    for (uint32_t cw = 1; cw <= 64; ++cw) {
        const uint32_t M = ((1u << 20) + cw - 1) / cw;
        for (uint32_t k = 0; k < 4096; ++k) {
            const uint64_t p = (uint64_t)k * M;
            ++n2;
            if ((p >> 32) || M >= (1u << 24) || (uint32_t)(p >> 20) != k / cw) ++bad2;
        }
    }
    printf("%ld %ld %ld %ld\n", bad, n, bad2, n2);
    return 0;
}
"""


def test_division_identities(tmp_path):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "arith.c"
    exe = tmp_path / "arith"
    src.write_text(SRC)
    subprocess.run([cc, "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"], check=True)
    bad, n, bad2, n2 = (int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                                        text=True).stdout.split())
    assert n == 2160 * 8193 and bad == 0
    assert n2 == 64 * 4096 and bad2 == 0


def test_planted_mistake_is_caught(tmp_path):
    """the same check fails for a quotient without the correction step"""
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "arith.c"
    exe = tmp_path / "arith"
    src.write_text(SRC.replace("fmaf(fmaf(-fb, q, fa), rb, q)", "q"))
    subprocess.run([cc, "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"], check=True)
    bad = int(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()[0])
    assert bad > 0

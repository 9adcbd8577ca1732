"""Host checks of float identities the kernels rely on for bit-exact results.

The Morton decode divides a raw axis value by the axis scale (2^21 - 1 for
63-bit codes, 2^10 - 1 for 30-bit codes).  The kernels replace that IEEE
division by a reciprocal product and one FMA correction; this test proves the
two agree for every raw value the decode can see, by running all 2^21 of them
through C code with hardware FMA.
"""
import shutil
import subprocess

import pytest

_SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
int main(int argc, char** argv) {
    volatile float d = strtof(argv[1], 0);
    const float r = strtof(argv[2], 0);
    if (r != 1.0f / d) { printf("reciprocal constant is not 1/d\n"); return 2; }
    long bad = 0;
    for (uint32_t x = 0; x < (1u << 21); ++x) {
        const float fx = (float)x;
        const float q0 = fx * r;
        const float q = fmaf(fmaf(-q0, d, fx), r, q0);
        if (q != fx / d) ++bad;
    }
    printf("%ld\n", bad);
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("scale, recip", [("2097151", "0x1.000008p-21"), ("1023", "0x1.00401p-10")])
def test_reciprocal_division_is_exact(tmp_path, scale, recip):
    src = tmp_path / "div.c"
    src.write_text(_SRC)
    exe = tmp_path / "div"
    subprocess.run(["gcc", "-O2", "-mfma", "-ffp-contract=off", str(src), "-o", str(exe), "-lm"], check=True)
    out = subprocess.run([str(exe), scale, recip], check=True, capture_output=True, text=True).stdout.strip()
    assert out == "0", f"{out} raw values differ from the IEEE quotient for scale {scale}"

"""The CLI driver (hsrq3in_amd/driver/main.c, the counterpart of the
reference's main.c:18-116): argument handling, the test-matrix generator and
the validation.  On the CPU the driver is linked against the oracle through a
test shim; on the GPU the in-tree binary linked against libhsrq runs."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAIN = os.path.join(ROOT, "hsrq3in_amd", "driver", "main.c")
BIN = os.path.join(ROOT, "hsrq3in_amd", "driver", "eigvecs")


@pytest.fixture(scope="module")
def oracle_driver(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("drv") / "eigvecs_oracle")
    subprocess.run(["gcc", "-std=gnu99", "-O2", "-fopenmp", "-ffp-contract=off", "-Wall", "-Werror",
                    "-I", os.path.join(ROOT, "include"), MAIN,
                    os.path.join(ROOT, "tests", "driver_oracle_shim.c"),
                    os.path.join(ROOT, "oracle", "hsrq_oracle.c"), "-o", out, "-lm"],
                   check=True)
    return out


def run(cmd, **kw):
    return subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, **kw)


def max_residual(out):
    m = re.search(r"Max residual .* = ([0-9.e+-]+) \(bound 10n = (\d+), (\d+) above\)", out)
    assert m, out
    return float(m.group(1)), float(m.group(2)), int(m.group(3))


def test_usage(oracle_driver):
    r = run([oracle_driver])
    assert r.returncode != 0 and r.stdout.startswith("Usage")
    for arg in ("n:", "tlsz:", "rhs-tlsz:", "select-frac:", "seed:"):
        assert arg in r.stdout


def test_rejects_invalid(oracle_driver):
    r = run([oracle_driver, "100", "64", "16", "1.5", "0"])
    assert r.returncode != 0 and "Invalid" in r.stdout


@pytest.mark.parametrize("n,tl,frac,seed", [(300, 64, 0.2, 3), (257, 30, 0.1, 9), (64, 64, 1.0, 1)])
def test_oracle_driver_end_to_end(oracle_driver, n, tl, frac, seed):
    r = run([oracle_driver, str(n), str(tl), "16", str(frac), str(seed)], timeout=120)
    assert r.returncode == 0, r.stdout
    assert f"Seed = {seed}" in r.stdout           # the seed is honoured
    if tl % 4:
        assert "INFO: tlsz must be divisible by 4" in r.stdout
    worst, bound, above = max_residual(r.stdout)
    assert above == 0 and worst <= bound
    assert "relative error" not in r.stdout.split("Validate...")[1].split("Max")[0]


def test_generator_spectrum(tmp_path):
    """The generated H is upper Hessenberg with eigenvalues 1..n."""
    src = tmp_path / "gen.c"
    txt = open(MAIN).read().replace("int main(int argc", "int driver_main(int argc")
    src.write_text(txt + r'''
#include <stdio.h>
int main(void) {
    int n = 60, ld = 64;
    double *H = calloc((size_t)ld * n, sizeof(double)), lam[60];
    generate_hessenberg(n, H, ld, lam, 5u);
    for (int j = 0; j < n; j++) for (int i = 0; i < n; i++)
        printf("%.17g%c", H[i + (size_t)ld * j], i + 1 == n ? '\n' : ' ');
    return 0;
}
''')
    exe = str(tmp_path / "gen")
    subprocess.run(["gcc", "-std=gnu99", "-O2", "-fopenmp", "-I", os.path.join(ROOT, "include"),
                    str(src), os.path.join(ROOT, "tests", "driver_oracle_shim.c"),
                    os.path.join(ROOT, "oracle", "hsrq_oracle.c"), "-o", exe, "-lm"], check=True)
    cols = np.loadtxt(run([exe]).stdout.splitlines())
    H = cols.T   # printed column by column
    assert np.all(np.tril(H, -2) == 0.0)
    assert np.all(np.diag(H, -1) != 0.0)       # unreduced
    w = np.sort(np.linalg.eigvals(H).real)
    assert np.allclose(w, np.arange(1, 61), rtol=0, atol=1e-8)


@pytest.mark.gpu
def test_gpu_driver():
    assert os.path.exists(BIN), "driver not built"
    r = run([BIN, "2000", "64", "32", "0.1", "4"], timeout=180)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    worst, bound, above = max_residual(r.stdout)
    assert above == 0 and worst <= bound

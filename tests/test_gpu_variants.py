"""Every kernel variant the autotuner chooses between computes the same bits.
Column kernels that split one column's output sequences over 2 or 4 blocks
(variants 6-9) are compared with the unsplit one-column variant of the same
FFT plan, and tiles of 2, 4 and 8 columns (2 also split over 2 or 4 blocks) /
2 and 4 rows with that variant too.  Each
configuration runs in a child process of its own (tests/variant_check.py),
since variants are fixed once per process and grid size."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))

# (reference variants, compared variants) per grid size: column, tangent
# column, row, tangent row; variants 0-2 / 6-7 / 10 / 12-13 use the wide
# plan, 3-5 / 8-9 / 11 / 14-15 the narrow one
CASES = [
    ("32:0,0,0,0;128:0,0,0,0", "32:6,7,1,2;128:7,6,2,1"),
    ("64:3,3,3,3;256:3,3,3,3", "64:9,8,4,5;256:8,9,5,4"),
    # tiles of 8 columns (N >= 128): 10 wide, 11 narrow
    ("128:0,3,0,0;256:3,0,3,3", "128:10,11,0,0;256:11,10,3,3"),
    # tiles of 2 columns split over 2 or 4 blocks: 12, 13 wide, 14, 15 narrow
    ("32:0,3,0,0;64:3,0,3,3;128:3,0,3,3", "32:12,15,0,0;64:14,13,3,3;128:15,12,3,3"),
    # tiles of 8 columns with the middle in registers (16, primal family only)
    # against the same tiles through LDS (11)
    ("128:11,0,3,3;256:11,0,3,3", "128:16,0,3,3;256:16,0,3,3"),
]


def run(variants, path):
    env = dict(os.environ, ECSMHD_VARIANTS=variants)
    res = subprocess.run([sys.executable, os.path.join(HERE, "variant_check.py"), path], env=env,
                         capture_output=True, text=True, timeout=240)
    print(variants, res.stdout, res.stderr[-2000:], flush=True)
    assert res.returncode == 0
    return np.load(path)


@pytest.mark.parametrize("ref,alt", CASES)
def test_variants_bitwise_equal(ref, alt, tmp_path):
    print(f"variants {ref} vs {alt}", flush=True)
    a = run(ref, str(tmp_path / "ref.npz"))
    b = run(alt, str(tmp_path / "alt.npz"))
    assert sorted(a.files) == sorted(b.files)
    for k in a.files:
        assert np.array_equal(a[k], b[k]), k

"""Golden-image regression of the example drivers through the harness
(tests/golden/render_golden.py, the reference's render_golden.py contract):
  * --selftest: two renders of every manifest scene are bit-identical at the
    compare resolution (max_abs 0), a stricter bar than the reference's
    4-LSB / 1 % floor;
  * check: every scene is within the manifest tolerance of its committed
    golden."""
import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]
HARNESS = ROOT / "tests" / "golden" / "render_golden.py"
MANIFEST = ROOT / "tests" / "golden" / "manifest.json"


def run(*mode):
    print("harness:", *mode, flush=True)
    r = subprocess.run([sys.executable, str(HARNESS), str(MANIFEST), *mode], cwd=ROOT, capture_output=True,
                       text=True, timeout=300)
    print(r.stdout, r.stderr[-2000:], flush=True)
    return r.returncode, r.stdout


def test_selftest_renders_are_bit_identical():
    rc, out = run("--selftest")
    assert rc == 0
    lines = [l for l in out.splitlines() if "(selftest)" in l]
    assert len(lines) == 2 and all("max_abs=0 diff_px=0" in l for l in lines)


def test_renders_match_committed_goldens():
    rc, out = run()
    assert rc == 0 and "ALL PASS" in out

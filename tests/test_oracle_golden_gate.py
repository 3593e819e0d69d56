"""The reference's golden gate (tools/check_golden.sh:19-58) with this
repository's oracle in place of the reference evaluator: the reference engine
(LGPEngine, Dataset, main) linked with oracle/ref_engine_shim.cpp must
reproduce golden/hist_*.csv and golden/prog_*.txt byte for byte over 3 targets
x 3 seeds x 1000 generations.  Every fitness the engine sees comes from the
oracle, so one wrong bit anywhere changes the trajectory.

Runs only where the reference sources exist (the shim is built from them by
`make -C oracle ref`); the GPU box has no copy and skips it."""
import os
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

from conftest import REFERENCE, REPO

BIN = os.path.join(REPO, "oracle", "_ref", "lgp_run_oracle")

pytestmark = pytest.mark.skipif(
    not (os.path.isdir(os.path.join(REFERENCE, "golden")) and os.path.exists(BIN)),
    reason="needs the reference sources and `make -C oracle ref`")


def _cell(args):
    t, s, tmp = args
    env = dict(os.environ, OMP_NUM_THREADS="2")
    subprocess.run([BIN, f"datasets/{t}_train.csv", f"datasets/{t}_test.csv", t, str(s),
                    os.path.join(tmp, f"results_{t}_{s}.csv"),
                    os.path.join(tmp, f"prog_{t}_s{s}.txt"),
                    os.path.join(tmp, f"hist_{t}_s{s}.csv")],
                   cwd=REFERENCE, env=env, check=True, stdout=subprocess.DEVNULL)
    out = []
    for kind, ext in (("hist", "csv"), ("prog", "txt")):
        name = f"{kind}_{t}_s{s}.{ext}"
        with open(os.path.join(REFERENCE, "golden", name), "rb") as a, \
                open(os.path.join(tmp, name), "rb") as b:
            out.append((name, a.read() == b.read()))
    return out


def test_golden_gate_with_oracle_evaluator():
    cells = [(t, s) for t in ("nguyen1", "nguyen2", "nguyen3") for s in (0, 1, 2)]
    with tempfile.TemporaryDirectory() as tmp:
        with ThreadPoolExecutor(max_workers=4) as ex:
            results = [r for rs in ex.map(_cell, [(t, s, tmp) for t, s in cells]) for r in rs]
    bad = [name for name, ok in results if not ok]
    assert not bad, f"golden mismatch: {bad}"
    assert len(results) == 18

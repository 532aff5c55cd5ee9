"""The CPU oracle under AddressSanitizer (+ UndefinedBehaviorSanitizer) and
ThreadSanitizer: oracle/Makefile builds the same sources instrumented, and a
child Python process that preloads the sanitizer runtime runs the threaded
drivers (the reference's worker scheme: FixedSource.cpp:22-77) and the
single-event entry points on small problems.  Any sanitizer report fails
the test.  CPU only."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

SCENARIOS = r"""
import numpy as np
from minimc_amd import abi, configs
from oracle.oracle import Oracle, SoA

# k-eigenvalue, multigroup slab with flux and currents, 4 workers
pr = configs.c1_slab(histories=3000)
o = Oracle(pr)
o.set_entropy_mesh(*pr.entropy_mesh)
o.keig_init(pr.histories)
for c in range(3):
    o.run_cycle(active=c > 0, threads=4)
o.get_source()
# fixed source, continuous energy with a sensitivity, 4 workers
pr = configs.c2_slab(n_points=4000, n_res=20)
pr.perturbations.append(dict(name="total", nuclide=0))
pr.estimators.append(dict(pr.estimators[0], name="s", sensitivity=1, parent=0))
o = Oracle(pr)
o.run_fixed_source(3000, threads=4)
# multi-zone k-eigenvalue with thermal scattering (reduced tables), 4 workers
pr = configs.c5_multizone(histories=1500, tables=configs.tsl_tables(reduced=True))
o = Oracle(pr)
o.set_entropy_mesh(*pr.entropy_mesh)
o.keig_init(pr.histories)
for c in range(2):
    o.run_cycle(active=c > 0, threads=4)
# batch thermal-scattering sampling, 4 workers
n = 5000
rng = np.random.default_rng(1)
o.tsl_sample(np.exp(rng.uniform(np.log(1e-11), np.log(3.9e-6), n)),
             rng.uniform(300, 1200, n), rng.integers(1, 2**31 - 1, n), threads=4)
# single events on the delta-tracking linear-field slab
pr = configs.linear_temperature_slab()
o = Oracle(pr)
o.run_fixed_source(3000, threads=4)
s = SoA(2000, E=1e-8, u=1.0, cell=0, rng=rng.integers(1, 2**31 - 1, 2000))
o.eval_event(abi.MMC_EVAL_COLLIDE, s)
print("scenarios done")
"""


def _run(kind):
    lib = os.path.join(ORACLE, "_build", "liboracle_%s.so" % kind)
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitizers"], check=True)
    runtime = subprocess.run(["gcc", "-print-file-name=lib%s.so" % kind], check=True,
                             capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(runtime):
        pytest.skip("no lib%s runtime for gcc" % kind)
    env = dict(os.environ, MMC_ORACLE_LIB=lib, LD_PRELOAD=runtime, PYTHONPATH=ROOT,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", TSAN_OPTIONS="halt_on_error=0")
    p = subprocess.run([sys.executable, "-c", SCENARIOS], env=env, capture_output=True,
                       text=True, timeout=900, cwd=ROOT)
    return p


@pytest.mark.parametrize("kind", ["asan", "tsan"])
def test_oracle_under_sanitizer(kind):
    p = _run(kind)
    report = [l for l in p.stderr.splitlines()
              if "Sanitizer" in l or "runtime error" in l]
    assert p.returncode == 0 and "scenarios done" in p.stdout, p.stderr[-3000:]
    assert not report, "\n".join(report[:20])

"""CPU execution of the library's kernels (tests/emu: HIP compiled as host
C++, work-items as fibers, wave64 collectives and barriers as rendezvous).

The batched persistent path is the one with cross-workgroup waits; this runs
its full query mix (direct, partitioned, 64-bit keys, sampled estimates,
empty and one-row queries) against the oracle, over several graph replays,
in both batch modes.  A stuck kernel is reported by the model with the
source line of every work-item instead of hanging."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "emu")
DRIVER = os.path.join(EMU, "_build", "batch_driver")
PLAIN = os.path.join(EMU, "_build", "plain_driver")


@pytest.fixture(scope="module")
def driver():
    r = subprocess.run(["make", "-s", "-j8"], cwd=EMU, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return DRIVER


@pytest.mark.parametrize("mode,replays", [(0, 3), (1, 1)])
def test_batch_mix_on_cpu_model(driver, mode, replays):
    env = dict(os.environ, EMU_WATCHDOG_S="120")
    r = subprocess.run([driver, str(mode), str(replays)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert r.stdout.strip().endswith("OK")


def test_batch_mix_many_resident_workgroups(driver):
    # the resident set of a full-size device: 256 compute units, 32 running at once
    env = dict(os.environ, EMU_WATCHDOG_S="120", EMU_CUS="256", EMU_RESIDENT="32")
    r = subprocess.run([driver, "0", "1"], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]


def test_plain_operator_kernel_variants_on_cpu_model(driver):
    """Every aggregation-kernel variant (direct / partitioned, 32/64-bit keys,
    one or four value columns, 4- and 8-byte values, wave folding on and off,
    the LDS-full spill path, the empty-marker key, exact fp64 SUM) against
    the oracle."""
    env = dict(os.environ, EMU_WATCHDOG_S="120")
    r = subprocess.run([PLAIN], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert r.stdout.strip().splitlines()[-1].startswith("OK")

"""The scalar modular inversion (field.h inv_mod_plain, Pornin's batched binary GCD),
built for the host from the same source and checked against Python's pow(x, -1, m) for
both BN254 moduli: edge values (1, 2, m - 1, powers of two around the 64-bit
approximation's boundaries, values just below m) and random ones of every length."""
import pathlib
import random
import shutil
import subprocess

import pytest

from halo2_amd.fields import P_MOD, R_MOD

ROOT = pathlib.Path(__file__).resolve().parents[1]
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def inverter(tmp_path_factory):
    if not pathlib.Path(HIPCC).exists():
        pytest.skip("hipcc not available")
    exe = tmp_path_factory.mktemp("inv") / "inverse_host"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-I", str(ROOT / "halo2-gpu_amd" / "csrc"), "-o", str(exe),
                    str(ROOT / "tests" / "inverse_host.hip")], check=True, capture_output=True)
    return exe


def _cases(mod, rng):
    xs = [1, 2, 3, 5, mod - 1, mod - 2, (mod - 1) // 2, 2**31, 2**32, 2**62, 2**63, 2**64 - 1, 2**64, 2**200 + 1,
          2**253]
    xs += [rng.randrange(1, mod) for _ in range(3000)]
    xs += [rng.randrange(1, 2**rng.randrange(1, 254)) for _ in range(3000)]
    xs += [mod - rng.randrange(1, 2**rng.randrange(1, 200)) for _ in range(1000)]
    return [x % mod for x in xs if x % mod]


def test_inverse_matches_pow(inverter):
    rng = random.Random(91)
    cases = [("r", R_MOD, x) for x in _cases(R_MOD, rng)] + [("q", P_MOD, x) for x in _cases(P_MOD, rng)]
    out = subprocess.run([str(inverter)], input="".join(f"{t} {x:064x}\n" for t, _, x in cases),
                         capture_output=True, text=True, check=True).stdout.split()
    assert len(out) == len(cases)
    bad = [(t, hex(x)) for (t, m, x), o in zip(cases, out) if int(o, 16) != pow(x, -1, m)]
    assert not bad, bad[:5]

"""A C host linked with -lp2p_hip (tests/support/dropin_host.c) runs one step
through the reference's call protocol; its accelerations are compared with
the FP64 oracle on the same integer-grid positions, leaves and list."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "support", "dropin_host")


def _read(path):
    raw = open(path, "rb").read()
    n, nleaf, ntask = np.frombuffer(raw, np.int32, 3)
    int2pos = np.frombuffer(raw, np.float64, 1, 12)[0]
    eps, rs, mass = np.frombuffer(raw, np.float32, 3, 20)
    off = 32
    def take(dt, count):
        nonlocal off
        a = np.frombuffer(raw, dt, count, off)
        off += a.nbytes
        return a
    posi = take(np.int32, 3 * n).reshape(n, 3)
    ip, npl = take(np.int32, nleaf), take(np.int32, nleaf)
    tt, ts = take(np.int32, ntask), take(np.int32, ntask)
    acc = take(np.float32, 3 * n).reshape(3, n).T
    return dict(pos=posi.astype(np.float64) * int2pos, ip=ip, np=npl, tt=tt, ts=ts, acc=acc.astype(np.float64),
                eps=float(eps), rs=float(rs), mass=float(mass))


def test_c_host_links_the_dropin_library():
    assert os.path.exists(EXE), "tests/support/dropin_host not built (__graft_entry__.build())"
    out = subprocess.run(["readelf", "-d", EXE], capture_output=True, text=True, check=True).stdout
    assert "libp2p_hip.so" in out


def test_c_host_without_a_device_fails_cleanly(tmp_path):
    """Without a GPU the library reports an error status instead of exiting
    or crashing; with one the step succeeds."""
    r = subprocess.run([EXE, str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=120)
    assert r.returncode >= 0, "killed by a signal"
    assert r.returncode == 0 or "no HIP device" in r.stderr


@pytest.mark.gpu
def test_c_host_step_matches_oracle(tmp_path):
    out = tmp_path / "out.bin"
    print("case: tests/support/dropin_host", flush=True)
    r = subprocess.run([EXE, str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    d = _read(out)
    prm = pyoracle.params(d["eps"], d["rs"], mass=d["mass"], softening=pyoracle.SOFT_REF)
    csr = pyoracle.build_csr(d["tt"], d["ts"], d["ip"].size)
    ref = pyoracle.p2p_f64(prm, d["pos"], d["ip"], d["np"], csr)
    med, p99, mx = pyoracle.rel_err(d["acc"], ref)
    print(f"C host: median {med:.2e} p99 {p99:.2e} max {mx:.2e}")
    assert p99 <= 1e-5 and mx <= 1e-4

"""The reference's own force step, unchanged, linked against the drop-in.

oracle/_ref/link_host is the reference's src/fmm.c (with operator.c,
memory.c, tasklog.c, utility.cc) compiled where it lies in the build
container and linked with -lp2p_hip against photons-on-mt3000_amd/lib/
libp2p_hip.so; its only own code (oracle/link_harness.c) defines the host
program's globals and calls fmm_prepare, fmm_task and the result collection
of src/gravity.c:441-447.  Every call the reference makes into its device
library -- hostmalloc_task, the flushes of task_compute_p2p_gpu from its walk
and the last_time call (src/fmm.c:500-515, :1221-1300, implicitly declared
there), hostfree_task, get_last_kntime, hostmalloc_out,
devMemcpyDH_updateAcc, hostfree_out -- goes to the drop-in.

Without a device the step must run to its end with the library reporting
ERR_NOGPU (no exit, no crash); on the GPU the accelerations must match the
FP64 oracle over the same lists.  The tree and leaves the reference builds
are checked against the host builder, whose lists are pinned to the
reference's."""
import os
import subprocess

import numpy as np
import pytest

import photons_p2p as pp
import pyoracle
from photons_p2p import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "link_host")

needs_exe = pytest.mark.skipif(not os.path.exists(EXE),
                               reason="oracle/_ref/link_host is built only where the reference tree is present")


def _run(tmp_path, w, len_task):
    m = w.meta
    posi = np.ascontiguousarray(m["posi"], np.int32)
    tag = np.ascontiguousarray(m["tag"], np.int32)
    n = posi.shape[0]
    inp = tmp_path / "in.bin"
    out = tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.int64(n).tobytes())
        f.write(np.array([m["maxleaf"], len_task], np.int32).tobytes())
        f.write(np.array([m["open_angle"], m["cutoff_radius"], w.eps, w.rs, w.mass], np.float32).tobytes())
        f.write(np.float64(m["int2pos"]).tobytes())
        f.write(np.asarray(m["bdl"], np.float32).tobytes())
        f.write(np.asarray(m["bdr"], np.float32).tobytes())
        f.write(posi.tobytes())
        f.write(tag.tobytes())
    print(f"case: link_host N={n} LEN_TASK={len_task}", flush=True)
    r = subprocess.run([EXE, str(inp), str(out)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, f"link_host exit {r.returncode}"
    raw = open(out, "rb").read()
    status, nleaf = np.frombuffer(raw, np.int32, 2)
    kt = float(np.frombuffer(raw, np.float64, 1, 8)[0])
    off = 16
    perm = np.frombuffer(raw, np.int32, n, off)
    off += 4 * n
    ip = np.frombuffer(raw, np.int32, nleaf, off)
    off += 4 * nleaf
    npl = np.frombuffer(raw, np.int32, nleaf, off)
    off += 4 * nleaf
    acc = np.frombuffer(raw, np.float32, 3 * n, off).reshape(3, n).T.astype(np.float64)
    # the reference's tree over this input equals the host builder's
    mine = W.build_lists(posi, tag, m["int2pos"], m["maxleaf"], np.float32(m["open_angle"]),
                         np.float32(m["cutoff_radius"]), m["bdl"], m["bdr"])
    np.testing.assert_array_equal(perm, mine["perm"])
    np.testing.assert_array_equal(ip, mine["leaf_ip"])
    np.testing.assert_array_equal(npl, mine["leaf_np"])
    return int(status), kt, acc, posi[perm], mine


@needs_exe
def test_reference_step_links_and_runs_against_the_dropin(tmp_path):
    out = subprocess.run(["readelf", "-d", EXE], capture_output=True, text=True, check=True).stdout
    assert "libp2p_hip.so" in out
    und = subprocess.run(["nm", "-u", EXE], capture_output=True, text=True, check=True).stdout.split()
    for sym in ("task_compute_p2p_gpu", "hostmalloc_task", "hostfree_task", "devMemcpyDH_updateAcc",
                "get_last_kntime"):
        assert sym in und, sym
    w = W.config_plain(12, softening=0)
    status, kt, acc, _, _ = _run(tmp_path, w, len_task=700)
    if pp.device_count() == 0:
        assert status == pp.ERR_NOGPU
        assert np.all(acc == 0.0)
    else:
        assert status == 0


@needs_exe
@pytest.mark.gpu
@pytest.mark.parametrize("len_task", [1500, 1 << 22])
def test_reference_step_matches_oracle(tmp_path, len_task):
    """The reference's walk flushes LEN_TASK tasks at a time into the drop-in
    (many flushes, or one); the collected accelerations equal the FP64 oracle
    on the same list, with the reference's max(r, SoftenScale) softening."""
    w = W.config_plain(20, softening=0)
    status, kt, acc, posi_tree, mine = _run(tmp_path, w, len_task)
    assert status == 0
    assert kt > 0.0
    prm = pyoracle.params(float(np.float32(w.eps)), float(np.float32(w.rs)), mass=float(np.float32(w.mass)),
                          softening=pyoracle.SOFT_REF)
    csr = pyoracle.build_csr(mine["task_t"], mine["task_s"], mine["leaf_ip"].size)
    pos = posi_tree.astype(np.float64) * w.meta["int2pos"]
    ref = pyoracle.p2p_f64(prm, pos, mine["leaf_ip"], mine["leaf_np"], csr)
    med, p99, mx = pyoracle.rel_err(acc, ref)
    print(f"reference fmm_task + drop-in: kernels {kt * 1e3:.3f} ms, median {med:.2e} p99 {p99:.2e} max {mx:.2e}")
    assert p99 <= 1e-5 and mx <= 1e-4

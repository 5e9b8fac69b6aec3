"""The C-ABI libraries load and export every function their headers declare;
host-side logic that needs no device (no compute calls are made here)."""
import ctypes
import os
import re

import numpy as np
import pytest
from scipy.special import erfc

import photons_p2p as pp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared(header):
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while", "for", "return")))


def test_core_library_exports_header():
    lib = pp.load()
    names = declared("photons_p2p.h")
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(pp.CORE_SYMBOLS)


def test_dropin_exports_reference_entry_points():
    host = ctypes.CDLL(os.path.join(ROOT, "tests", "support", "libhostglobals.so"), mode=ctypes.RTLD_GLOBAL)
    lib = ctypes.CDLL(pp.DROPIN_LIB)
    names = declared("p2p_hip.h")
    ref = {"task_compute_p2p_gpu", "p2ppack_compute_gpu", "task_compute_p2p_act_gpu", "p2ppack_compute_act_gpu",
           "hostmalloc_task", "hostrealloc_task", "hostfree_task", "hostmalloc_in", "hostfree_in",
           "hostmalloc_out", "hostfree_out", "dinfofree", "devfree", "updateAcc_fromHostBuff",
           "devMemcpyDH_updateAcc", "sync_with_dev", "get_last_kntime"}
    assert ref <= set(names)
    for n in names:
        assert hasattr(lib, n), n
    del host


def test_no_device_is_an_error_not_a_fallback():
    if pp.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(pp.P2PError) as e:
        pp.Context(0)
    assert e.value.status == pp.ERR_NOGPU


def test_dropin_reports_missing_device_without_exit():
    if pp.device_count() > 0:
        pytest.skip("a device is visible")
    host = ctypes.CDLL(os.path.join(ROOT, "tests", "support", "libhostglobals.so"), mode=ctypes.RTLD_GLOBAL)
    lib = ctypes.CDLL(pp.DROPIN_LIB)
    lib.p2p_hip_status.restype = ctypes.c_int
    lib.task_compute_p2p_gpu.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    ctypes.c_int.in_dll(host, "NPART").value = 0
    lib.task_compute_p2p_gpu(0, None, None, 1)
    assert lib.p2p_hip_status() == pp.ERR_NOGPU


def F_exact(s, ucut):
    u = ucut * np.sqrt(s)
    return (erfc(u) + 2 / np.sqrt(np.pi) * u * np.exp(-u * u)) / s ** 1.5


@pytest.mark.parametrize("softening,eps_div", [(1, 40.0), (0, 1 / 0.03), (1, 50.0)])
def test_device_table_law_accuracy(softening, eps_div):
    d = 400.0 / 256
    law = pp.ForceLaw(eps=d / eps_div, rs=1.25 * d, softening=softening)
    smin = (law.eps / law.rc) ** 2
    s = np.geomspace(smin, 1.0 - 1e-6, 300001).astype(np.float32)
    F, n = pp.law_eval(law, s)
    assert n == 32 * 32
    rel = np.abs(F / F_exact(s.astype(np.float64), law.ucut) - 1)
    assert rel.max() < 5e-7 and np.median(rel) < 5e-8


@pytest.mark.parametrize("softening,eps_div", [(pp.SOFT_PLUMMER, 40.0), (pp.SOFT_REF, 1 / 0.03), (pp.SOFT_PLUMMER, 2.0)])
def test_device_linear_law_accuracy(softening, eps_div):
    """The linear law (a + b s on 512 segments per octave, the line through
    F at each segment's two Chebyshev points) against F(s) from scipy's erfc:
    the interpolation error of a line is about (F''/16) w^2, 0.23 (w/s)^2 for
    the s^-3/2 power law (w/s <= 2^-9: 9e-7) and larger just below the cutoff,
    where g(u) falls steeply."""
    d = 400.0 / 256
    law = pp.ForceLaw(eps=d / eps_div, rs=1.25 * d, softening=softening, law=pp.LAW_LINEAR)
    smin = (law.eps / law.rc) ** 2
    s = np.geomspace(smin, 1.0 - 1e-6, 300001).astype(np.float32)
    s = np.maximum(s, np.float32(smin))
    F, n = pp.law_eval(law, s)
    e_lo = int(np.floor(np.log2(np.float32(smin))))
    assert n == -e_lo * 512 + 1                       # octaves of smin up to 1, and the entry of s = 1
    rel = np.abs(F / F_exact(s.astype(np.float64), law.ucut) - 1)
    low = s < 0.25
    assert rel.max() < 8e-6 and rel[low].max() < 2e-6 and np.median(rel) < 5e-7
    F1, _ = pp.law_eval(law, np.array([1.0, 1.5], np.float32))
    assert F1[0] == 0 and F1[1] == 0                  # at and beyond the cutoff


def test_linear_law_range():
    """The linear law's octaves of s must fit the LDS: eps / rc >= 2^-12.5
    (smin >= 2^-25); below it the law is refused, the table law still works."""
    rs = 1.0
    rc = 2 * rs * 2.8125
    with pytest.raises(pp.P2PError):
        pp.law_eval(pp.ForceLaw(eps=2.0 ** -13 * rc, rs=rs, law=pp.LAW_LINEAR), np.ones(1, np.float32))
    pp.law_eval(pp.ForceLaw(eps=2.0 ** -13 * rc, rs=rs), np.ones(1, np.float32))
    F, _ = pp.law_eval(pp.ForceLaw(eps=2.0 ** -12.4 * rc, rs=rs, law=pp.LAW_LINEAR),
                       np.geomspace(2.0 ** -24.8, 0.999, 1000).astype(np.float32))
    assert np.all(np.isfinite(F)) and np.all(F > 0)


def test_device_table_cutoff_and_continuity():
    law = pp.ForceLaw(eps=0.04, rs=2.0)
    F, _ = pp.law_eval(law, np.array([1.0, 1.25, 2.0, 0.9999999], np.float32))
    assert F[0] == 0 and F[1] == 0 and F[2] == 0
    assert F[3] == pytest.approx(F_exact(0.9999999, law.ucut), rel=1e-6)
    # segment joins: left and right limits agree to FP32 rounding
    s0 = np.float32(2.0 ** -5)
    left = np.nextafter(s0, np.float32(0))
    Fl, _ = pp.law_eval(law, np.array([left, s0], np.float32))
    assert Fl[0] == pytest.approx(Fl[1], rel=2e-6)


def test_law_parameter_validation():
    with pytest.raises(pp.P2PError):
        pp.law_eval(pp.ForceLaw(eps=1e-9, rs=1.0), np.ones(1, np.float32))     # eps / rc outside the table
    with pytest.raises(pp.P2PError):
        pp.law_eval(pp.ForceLaw(eps=10.0, rs=1.0), np.ones(1, np.float32))     # eps beyond the cutoff
    with pytest.raises(pp.P2PError):
        pp.law_eval(pp.ForceLaw(eps=0.1, rs=1.0), np.ones(1, np.float32), segs_per_octave=7)


def test_autotune_table_roundtrip(tmp_path):
    from photons_p2p import autotune as AT
    sig = AT.signature(1000, 5000, 300, 10 ** 9)
    assert sig == AT.signature(1010, 5050, 301, 10 ** 9 + 5)        # same half-octave buckets
    assert sig != AT.signature(4000, 5000, 300, 10 ** 9)
    p = str(tmp_path / "t.json")
    AT.save_table({sig: {"best": {"ipl_max": 3, "split_pairs": 0, "waves_per_cu": 0, "law": 0}, "ms": 1.0}}, p)
    assert AT.lookup(sig, p)["ipl_max"] == 3
    assert AT.lookup("nope", p) is None


def test_table_law_rejects_softening_whose_coefficients_overflow():
    """Below eps / rc ~ 6e-5 the cubic coefficient of the smallest octave
    (about 2.2 s^-4.5) leaves the FP32 range: the table law is refused instead
    of evaluating to NaN; just above it every value is finite and matches F(s)."""
    rs = 3.90625
    rc = 2 * rs * 2.8125
    with pytest.raises(pp.P2PError):
        pp.law_eval(pp.ForceLaw(eps=4.5e-5 * rc, rs=rs, softening=pp.SOFT_PLUMMER), np.ones(1, np.float32))
    eps = 8e-5 * rc
    smin = (eps / rc) ** 2
    s = np.geomspace(smin, 1.0 - 1e-6, 2000).astype(np.float32)
    F, _ = pp.law_eval(pp.ForceLaw(eps=eps, rs=rs, softening=pp.SOFT_PLUMMER), s)
    assert np.all(np.isfinite(F))
    assert np.abs(F / F_exact(s.astype(np.float64), 2.8125) - 1).max() < 5e-7
